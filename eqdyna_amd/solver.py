"""Python front end of the eqdyna_amd C-ABI (include/eqdyna_amd.h).

``Solver`` takes the reference's globalvar arrays (a bundle as produced by the
reference host or by ``eqdyna_amd.mesh``), hands them to libeqd.so, advances
time steps on the GPU and returns the evolved arrays in the same layout.
There is no CPU fallback: without the built library or a GPU this raises.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EQD_LIB", os.path.join(HERE, "libeqd.so"))

EXPORTS = ["eqd_create", "eqd_set", "eqd_commit", "eqd_run", "eqd_run_async", "eqd_wait", "eqd_get", "eqd_clock", "eqd_nan_info",
           "eqd_kernel_time", "eqd_set_timing", "eqd_device_bytes", "eqd_layout_info", "eqd_last_error",
           "eqd_destroy", "eqd_set_guess", "eqd_harvest", "eqd_harvest_get", "eqd_newton_stats", "eqd_mlp_forward",
           "eqd_mlp_grad", "eqd_mlp_train", "eqd_set_launch", "eqd_get_launch", "eqd_tune", "eqd_set_graph",
           "eqd_ensemble_create", "eqd_ensemble_run", "eqd_ensemble_tune", "eqd_ensemble_set_launch",
           "eqd_ensemble_set_graph", "eqd_ensemble_set_timing", "eqd_ensemble_kernel_time",
           "eqd_ensemble_last_error", "eqd_ensemble_destroy", "eqd_precompute", "eqd_record", "eqd_record_get",
           "eqd_mlp_forward_noalloc", "eqd_hypocentre_pairs", "eqd_hypocentre_get"]

# launch kinds (EQD_KIND_*)
KINDS = {"element": 0, "element_pml": 1, "node": 2, "node_pml": 3, "fault": 4}

HYP_VALS = 19
HYP_SLOTS = [78, 79, 80, 23, 73, 71, 72, 74, 75, 76, 77, 1, 2, 3, 9, 10, 20, 81]  # after timeElapsed
MLP_NTRAIN = 8833
MLP_NPARAMS = 8847

FIELDS = {
    "meshCoor": (1, "f8"), "eqNumIndexArr": (2, "i4"), "eqNumStartIndexLoc": (3, "i4"),
    "numOfDofPerNodeArr": (4, "i4"), "nodeElemIdRelation": (5, "i4"), "elemTypeArr": (6, "i4"),
    "mat": (7, "f8"), "eleshp": (8, "f8"), "eledet": (9, "f8"), "elemass": (10, "f8"), "ss": (11, "f8"),
    "phi": (12, "f8"), "stressCompIndexArr": (13, "i4"), "eleporep": (14, "f8"),
    "nodalMassArr": (15, "f8"), "fnms": (16, "f8"), "nsmp": (17, "i4"), "un": (18, "f8"),
    "us": (19, "f8"), "ud": (20, "f8"), "arn": (21, "f8"), "anonfs": (22, "i4"), "idhist": (23, "i4"),
    "stressArr": (24, "f8"), "pstrain": (25, "f8"), "v1": (26, "f8"), "velArr": (27, "f8"),
    "dispArr": (28, "f8"), "nodalForceArr": (29, "f8"), "fric": (30, "f8"), "fnft": (31, "f8"),
    "onFaultQuantHistSCECForm": (32, "f8"), "OffFaultStGramSCEC": (33, "f8"),
}
INPUTS = ["meshCoor", "eqNumIndexArr", "eqNumStartIndexLoc", "numOfDofPerNodeArr", "nodeElemIdRelation",
          "elemTypeArr", "mat", "eleshp", "eledet", "elemass", "ss", "phi", "stressCompIndexArr", "eleporep",
          "nodalMassArr", "fnms", "nsmp", "un", "us", "ud", "arn", "anonfs", "idhist", "stressArr", "pstrain",
          "v1", "velArr", "dispArr", "nodalForceArr", "fric", "fnft"]
OUTPUTS = ["fric", "fnft", "velArr", "dispArr", "v1", "nodalForceArr", "stressArr", "pstrain",
           "onFaultQuantHistSCECForm", "OffFaultStGramSCEC"]

PREC = {"fp64": 0, "fp32": 1}


class Scalars(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dt", "rdampk", "rdampm", "w", "grav", "roumax", "rhow", "gamar",
                 "ccosphi", "sinphi", "tv", "slipRateThres",
                 "xsource", "ysource", "zsource", "nucR", "nucRuptVel", "nucdtau0", "nucT",
                 "kapa_hg", "dx", "R", "vmaxPML")] + \
              [("PMLb", C.c_double * 8), ("pi", C.c_double)] + \
              [(n, C.c_int32) for n in
               ("nstep", "nn", "ne", "neq", "size_eqidx", "size_stress", "nf", "non", "noff", "nPML",
                "C_elastic", "C_Q", "C_hg", "C_nuclea", "friclaw", "TPV", "insertFaultType", "nucfault",
                "precision", "device")] + [("fric_tp_h", C.c_double)]


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64), ("device", C.c_int32)]


class PrecomputeArgs(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("nn", "ne", "neq", "size_eqidx")] + \
        [(n, C.c_void_p) for n in ("meshCoor", "nodeElemIdRelation", "elemTypeArr", "mat", "numOfDofPerNodeArr",
                                   "eqNumStartIndexLoc", "eqNumIndexArr", "eleshp", "eledet", "elemass", "ss",
                                   "phi", "nodalMassArr", "fnms")] + [("device", C.c_int32)]


class EqdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"eqdyna_amd error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libeqd.so; raises if it is not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.eqd_create.argtypes = [C.POINTER(Scalars), C.POINTER(C.c_void_p)]
        L.eqd_set.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]
        L.eqd_commit.argtypes = [C.c_void_p]
        L.eqd_run.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_run_async.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_hypocentre_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.eqd_hypocentre_get.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.eqd_wait.argtypes = [C.c_void_p]
        L.eqd_get.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]
        L.eqd_clock.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.eqd_nan_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.eqd_kernel_time.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.eqd_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_device_bytes.argtypes = [C.c_void_p]
        L.eqd_device_bytes.restype = C.c_int64
        L.eqd_layout_info.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]
        L.eqd_last_error.argtypes = [C.c_void_p]
        L.eqd_last_error.restype = C.c_char_p
        L.eqd_destroy.argtypes = [C.c_void_p]
        fp = C.POINTER(C.c_float)
        L.eqd_set_guess.argtypes = [C.c_void_p, fp, C.c_int64, C.c_int32, C.c_float]
        L.eqd_harvest.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
        L.eqd_harvest_get.argtypes = [C.c_void_p, fp, C.c_int64, C.POINTER(C.c_int64)]
        L.eqd_newton_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.eqd_mlp_forward.argtypes = [fp, fp, C.c_int64, fp, C.c_int32]
        L.eqd_mlp_forward_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p]
        L.eqd_mlp_grad.argtypes = [fp, fp, fp, C.c_int64, fp, fp, C.c_int32]
        L.eqd_mlp_train.argtypes = [fp, fp, C.c_int64, C.POINTER(TrainOpts), fp, fp]
        i32p = C.POINTER(C.c_int32)
        L.eqd_set_launch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.eqd_get_launch.argtypes = [C.c_void_p, C.c_int32, i32p, i32p]
        L.eqd_tune.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64]
        L.eqd_set_graph.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
        L.eqd_ensemble_run.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_ensemble_tune.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64]
        L.eqd_ensemble_set_launch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.eqd_ensemble_set_graph.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_ensemble_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.eqd_ensemble_kernel_time.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.eqd_ensemble_last_error.argtypes = [C.c_void_p]
        L.eqd_ensemble_last_error.restype = C.c_char_p
        L.eqd_ensemble_destroy.argtypes = [C.c_void_p]
        L.eqd_precompute.argtypes = [C.POINTER(PrecomputeArgs), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.eqd_record.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.eqd_record_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def make_scalars(sc: Dict[str, float], precision: str = "fp64", device: int = 0,
                 nstep: int | None = None) -> Scalars:
    s = Scalars()
    for n in ("dt", "rdampk", "rdampm", "w", "grav", "roumax", "rhow", "gamar", "ccosphi", "sinphi", "tv",
              "slipRateThres", "xsource", "ysource", "zsource", "nucR", "nucRuptVel", "nucdtau0", "nucT",
              "kapa_hg", "dx", "R", "vmaxPML", "pi"):
        setattr(s, n, float(sc[n]))
    for i in range(8):
        s.PMLb[i] = float(sc[f"PMLb{i + 1}"])
    s.nstep = int(nstep if nstep is not None else sc["nstep"])
    s.nn = int(sc["totalNumOfNodes"])
    s.ne = int(sc["totalNumOfElements"])
    s.neq = int(sc["totalNumOfEquations"])
    s.size_eqidx = int(sc["sizeOfEqNumIndexArr"])
    s.size_stress = int(sc["sizeOfStressDofIndexArr"])
    s.nf = int(sc["nftnd"])
    s.non = int(sc["numOfOnFaultStCount"])
    s.noff = int(sc["numOfOffFaultStCount"])
    s.nPML = int(sc["nPML"])
    for n in ("C_elastic", "C_Q", "C_hg", "C_nuclea", "friclaw", "TPV", "insertFaultType", "nucfault"):
        setattr(s, n, int(sc[n]))
    s.precision = PREC[precision]
    s.device = int(device)
    # the reference never assigns its module variable fric_tp_h (src/globalvar.f90:42): zero
    s.fric_tp_h = float(sc.get("fric_tp_h", 0.0))
    return s


class Solver:
    """One GPU-resident rupture scenario."""

    def __init__(self, bundle: Dict[str, np.ndarray], sc: Dict[str, float], precision: str = "fp64",
                 device: int = 0, nstep: int | None = None, timing: bool = False):
        self.L = lib()
        self.sc = dict(sc)
        self.s = make_scalars(sc, precision, device, nstep)
        h = C.c_void_p()
        rc = self.L.eqd_create(C.byref(self.s), C.byref(h))
        if rc != 0:
            raise EqdError(rc, "eqd_create rejected the scalars")
        self.h = h
        try:
            for name in INPUTS:
                if name not in bundle:
                    continue
                fid, dt = FIELDS[name]
                a = np.ascontiguousarray(bundle[name], dtype="<" + dt)
                self._check(self.L.eqd_set(self.h, fid, a.ctypes.data, a.size))
            if timing:
                self._check(self.L.eqd_set_timing(self.h, 1))
            self._check(self.L.eqd_commit(self.h))
        except Exception:
            self.close()
            raise

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self.L.eqd_last_error(self.h).decode()
            raise EqdError(rc, msg)

    def _count(self, name: str) -> tuple:
        s = self.s
        non = max(s.non, 1)
        shapes = {
            "fric": (s.nf, 100), "fnft": (s.nf,), "velArr": (s.nn, 3), "dispArr": (s.nn, 3),
            "v1": (s.neq,), "nodalForceArr": (s.neq,), "stressArr": (s.size_stress,), "pstrain": (s.ne,),
            "onFaultQuantHistSCECForm": (non, s.nstep, 12), "OffFaultStGramSCEC": (s.nstep, 6 * s.noff + 1),
        }
        return shapes[name]

    def run(self, nsteps: int) -> None:
        self._check(self.L.eqd_run(self.h, int(nsteps)))

    def run_async(self, nsteps: int) -> None:
        """Enqueue nsteps steps and return at once (wait() finishes the run)."""
        self._check(self.L.eqd_run_async(self.h, int(nsteps)))

    def wait(self) -> None:
        self._check(self.L.eqd_wait(self.h))

    def set_timing(self, on) -> None:
        """Per-kernel event timing: False/0 off, True/1 every step, k every k-th step."""
        self._check(self.L.eqd_set_timing(self.h, int(on)))

    def get(self, name: str) -> np.ndarray:
        fid, dt = FIELDS[name]
        out = np.zeros(self._count(name), dtype="<" + dt)
        self._check(self.L.eqd_get(self.h, fid, out.ctypes.data, out.size))
        return out

    def outputs(self) -> Dict[str, np.ndarray]:
        o = {k: self.get(k) for k in OUTPUTS if not (k == "OffFaultStGramSCEC" and self.s.noff == 0)}
        return o

    def clock(self):
        nt = C.c_int32()
        t = C.c_double()
        self._check(self.L.eqd_clock(self.h, C.byref(nt), C.byref(t)))
        return nt.value, t.value

    def nan_info(self):
        """(node, step) of the first NaN velocity (1-based node id; (0, 0) if none)."""
        node = C.c_int32()
        nt = C.c_int32()
        self._check(self.L.eqd_nan_info(self.h, C.byref(node), C.byref(nt)))
        return node.value, nt.value

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.L.eqd_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def layout_info(self) -> Dict[str, int]:
        out = {}
        for k in ("regular_elements", "pml_elements", "op_classes_regular", "phi_classes_regular",
                  "op_classes_pml", "phi_classes_pml", "phi_compact_regular", "phi_compact_pml",
                  "geom_ops_regular", "geom_ops_pml", "node3_slot_major"):
            v = C.c_int64()
            self._check(self.L.eqd_layout_info(self.h, k.encode(), C.byref(v)))
            out[k] = v.value
        return out

    def set_guess(self, params, k_fallback: int = 6, gate: float = 0.5) -> None:
        """Faster rate-and-state mode: Newton from the network's proposal where
        it predicts |ln(V/V_prev)| >= gate (params: the float32 parameter
        vector), stop at convergence or at the roundoff floor, fall back to the
        reference's solve after k_fallback iterations.  params None with
        k_fallback > 0: the same policy from the plain guess; params None with
        k_fallback 0: the reference's solve."""
        if params is None:
            self._check(self.L.eqd_set_guess(self.h, None, 0, int(k_fallback), 0.0))
            return
        p = _f32(params, MLP_NPARAMS)
        self._check(self.L.eqd_set_guess(self.h, _fptr(p), p.size, int(k_fallback), float(gate)))

    def harvest(self, every: int, capacity: int) -> None:
        self._check(self.L.eqd_harvest(self.h, int(every), int(capacity)))

    def harvest_get(self) -> np.ndarray:
        """Recorded rows [n, 7]: the six network inputs and the converged ln V."""
        n = C.c_int64()
        self._check(self.L.eqd_harvest_get(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 7), dtype=np.float32)
        if n.value:
            self._check(self.L.eqd_harvest_get(self.h, _fptr(out), n.value, C.byref(n)))
        return out

    def newton_stats(self) -> Dict[str, int]:
        v = (C.c_int64 * 5)()
        self._check(self.L.eqd_newton_stats(self.h, v))
        return {"solves": v[0], "iterations": v[1], "fallbacks": v[2], "unconverged": v[3], "roundoff_stops": v[4]}

    def device_bytes(self) -> int:
        return int(self.L.eqd_device_bytes(self.h))

    def hypocentre_pairs(self) -> np.ndarray:
        """1-based pairs at the hypocentre (showSourceDynamics)."""
        n = C.c_int32()
        self._check(self.L.eqd_hypocentre_pairs(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype="<i4")
        self._check(self.L.eqd_hypocentre_pairs(self.h, out.ctypes.data, out.size, C.byref(n)))
        return out[:n.value]

    def hypocentre_get(self, first_step: int, count: int) -> np.ndarray:
        """[count, pairs, 19]: timeElapsed and fric 78, 79, 80, 23, 73, 71, 72, 74,
        75, 76, 77, 1, 2, 3, 9, 10, 20, 81 of each hypocentre pair at each step."""
        npairs = len(self.hypocentre_pairs())
        out = np.zeros((count, npairs, HYP_VALS))
        if npairs and count:
            self._check(self.L.eqd_hypocentre_get(self.h, int(first_step), int(count), out.ctypes.data))
        return out

    def record(self, node_ids, every: int = 10, capacity: int = 64) -> None:
        """Record the velocities of node_ids (1-based) and fric(47) at every
        step nt with (nt - 1) % every == 0 (output_gm / output_src_evol)."""
        ids = np.ascontiguousarray(node_ids, dtype="<i4")
        self._rec_nodes = ids.size
        self._check(self.L.eqd_record(self.h, ids.ctypes.data, ids.size, int(every), int(capacity)))

    def record_get(self, max_records: int = 1024):
        """(vel [count, n_nodes, 3], slip_rate [count, nf]) recorded since the last call."""
        n = getattr(self, "_rec_nodes", 0)
        vel = np.zeros((max_records, max(n, 1), 3))
        sr = np.zeros((max_records, max(self.s.nf, 1)))
        cnt = C.c_int32()
        self._check(self.L.eqd_record_get(self.h, vel.ctypes.data, sr.ctypes.data, int(max_records), C.byref(cnt)))
        return vel[:cnt.value, :n], sr[:cnt.value, :self.s.nf]

    def set_launch(self, kind: str, block: int, remap: int = 0) -> None:
        """Launch shape of one kind (KINDS): block size and XCD remapping."""
        self._check(self.L.eqd_set_launch(self.h, KINDS[kind], int(block), int(remap)))

    def get_launch(self) -> Dict[str, tuple]:
        out = {}
        for k, v in KINDS.items():
            b, r = C.c_int32(), C.c_int32()
            self._check(self.L.eqd_get_launch(self.h, v, C.byref(b), C.byref(r)))
            out[k] = (b.value, r.value)
        return out

    def tune(self, steps: int = 5) -> dict:
        """Autotuner: times every launch shape on this scenario (its state is
        restored afterwards) and keeps the fastest; returns the report."""
        buf = C.create_string_buffer(1 << 16)
        self._check(self.L.eqd_tune(self.h, int(steps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_graph(self, on: bool) -> None:
        """True: capture one step as a HIP graph and replay it in run(); False
        (the default): launch every step."""
        self._check(self.L.eqd_set_graph(self.h, 1 if on else 0))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.eqd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, size=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if size is not None and a.size != size:
        raise ValueError(f"expected {size} float32 values, got {a.size}")
    return a


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_forward(params, X, device: int = 0) -> np.ndarray:
    """The network's r = ln(V / V_prev) for raw inputs X[n, 6] (GPU)."""
    p, X = _f32(params, MLP_NPARAMS), _f32(X)
    out = np.zeros(len(X), dtype=np.float32)
    rc = lib().eqd_mlp_forward(_fptr(p), _fptr(X), len(X), _fptr(out), int(device))
    if rc != 0:
        raise EqdError(rc, "eqd_mlp_forward failed")
    return out


def mlp_forward_noalloc(params_ptr: int, X_ptr: int, n: int, out_ptr: int, device: int = 0, stream: int = 0) -> None:
    """eqd_mlp_forward_noalloc: device arrays given by address (params
    float32[EQD_MLP_NPARAMS], X float32[n][6], out float32[n]), enqueued on the
    HIP stream `stream` (0: the null stream); returns without waiting."""
    rc = lib().eqd_mlp_forward_noalloc(C.c_void_p(params_ptr), C.c_void_p(X_ptr), int(n), C.c_void_p(out_ptr),
                                       int(device), C.c_void_p(stream))
    if rc != 0:
        raise EqdError(rc, "eqd_mlp_forward_noalloc failed")


def mlp_grad(params, X, r, device: int = 0):
    """(loss, gradient[MLP_NTRAIN]) of the mean squared error against targets r over all samples (GPU)."""
    p, X, y = _f32(params, MLP_NPARAMS), _f32(X), _f32(r)
    g = np.zeros(MLP_NTRAIN, dtype=np.float32)
    loss = C.c_float()
    rc = lib().eqd_mlp_grad(_fptr(p), _fptr(X), _fptr(y), len(X), _fptr(g), C.byref(loss), int(device))
    if rc != 0:
        raise EqdError(rc, "eqd_mlp_grad failed")
    return float(loss.value), g


def mlp_train(params, X, r, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, device: int = 0):
    """Train on the GPU against targets r; returns (trained parameters, mean loss per epoch)."""
    p, X, y = _f32(params, MLP_NPARAMS).copy(), _f32(X), _f32(r)
    o = TrainOpts(int(epochs), int(batch), float(lr), float(beta1), float(beta2), float(eps), int(seed), int(device))
    losses = np.zeros(int(epochs), dtype=np.float32)
    rc = lib().eqd_mlp_train(_fptr(X), _fptr(y), len(X), C.byref(o), _fptr(p), _fptr(losses))
    if rc != 0:
        raise EqdError(rc, "eqd_mlp_train failed")
    return p, losses


class Ensemble:
    """Scenarios advanced together: one launch per kind per step over all of
    them, the step captured once as a HIP graph and replayed.  The members stay
    Solver objects for reading results; they must not be run on their own or
    closed while the ensemble exists."""

    def __init__(self, members):
        self.L = lib()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        e = C.c_void_p()
        rc = self.L.eqd_ensemble_create(arr, len(self.members), C.byref(e))
        if rc != 0:
            raise EqdError(rc, "eqd_ensemble_create rejected the members")
        self.e = e

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise EqdError(rc, self.L.eqd_ensemble_last_error(self.e).decode())

    def run(self, nsteps: int) -> None:
        self._check(self.L.eqd_ensemble_run(self.e, int(nsteps)))

    def tune(self, steps: int = 5) -> dict:
        buf = C.create_string_buffer(1 << 16)
        self._check(self.L.eqd_ensemble_tune(self.e, int(steps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def record(self, node_ids, every: int = 10, capacity: int = 64) -> None:
        """Record the velocities of node_ids (1-based) and fric(47) at every
        step nt with (nt - 1) % every == 0 (output_gm / output_src_evol)."""
        ids = np.ascontiguousarray(node_ids, dtype="<i4")
        self._rec_nodes = ids.size
        self._check(self.L.eqd_record(self.h, ids.ctypes.data, ids.size, int(every), int(capacity)))

    def record_get(self, max_records: int = 1024):
        """(vel [count, n_nodes, 3], slip_rate [count, nf]) recorded since the last call."""
        n = getattr(self, "_rec_nodes", 0)
        vel = np.zeros((max_records, max(n, 1), 3))
        sr = np.zeros((max_records, max(self.s.nf, 1)))
        cnt = C.c_int32()
        self._check(self.L.eqd_record_get(self.h, vel.ctypes.data, sr.ctypes.data, int(max_records), C.byref(cnt)))
        return vel[:cnt.value, :n], sr[:cnt.value, :self.s.nf]

    def set_launch(self, kind: str, block: int, remap: int = 0) -> None:
        self._check(self.L.eqd_ensemble_set_launch(self.e, KINDS[kind], int(block), int(remap)))

    def set_graph(self, on: bool) -> None:
        self._check(self.L.eqd_ensemble_set_graph(self.e, 1 if on else 0))

    def set_timing(self, on) -> None:
        self._check(self.L.eqd_ensemble_set_timing(self.e, int(on)))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.L.eqd_ensemble_kernel_time(self.e, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if getattr(self, "e", None):
            self.L.eqd_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def precompute(bundle: Dict[str, np.ndarray], device: int = 0):
    """Device-side element precompute and nodal masses (eqd_precompute) from a
    bundle's mesh arrays; returns (arrays, kernel milliseconds).  nodalMassArr
    and fnms start from zero, as in the reference host."""
    L = lib()
    coor = np.ascontiguousarray(bundle["meshCoor"], dtype="<f8")
    ien = np.ascontiguousarray(bundle["nodeElemIdRelation"], dtype="<i4")
    et = np.ascontiguousarray(bundle["elemTypeArr"], dtype="<i4")
    mat = np.ascontiguousarray(bundle["mat"], dtype="<f8")
    ndof = np.ascontiguousarray(bundle["numOfDofPerNodeArr"], dtype="<i4")
    esl = np.ascontiguousarray(bundle["eqNumStartIndexLoc"], dtype="<i4")
    eqi = np.ascontiguousarray(bundle["eqNumIndexArr"], dtype="<i4")
    nn, ne = ndof.size, et.size
    neq = int(np.asarray(bundle["nodalMassArr"]).size)
    out = {"eleshp": np.zeros(24 * ne), "eledet": np.zeros(ne), "elemass": np.zeros(24 * ne), "ss": np.zeros(6 * ne),
           "phi": np.zeros(32 * ne), "nodalMassArr": np.zeros(neq), "fnms": np.zeros(nn)}
    a = PrecomputeArgs(nn, ne, neq, eqi.size, coor.ctypes.data, ien.ctypes.data, et.ctypes.data, mat.ctypes.data,
                       ndof.ctypes.data, esl.ctypes.data, eqi.ctypes.data, out["eleshp"].ctypes.data,
                       out["eledet"].ctypes.data, out["elemass"].ctypes.data, out["ss"].ctypes.data,
                       out["phi"].ctypes.data, out["nodalMassArr"].ctypes.data, out["fnms"].ctypes.data, device)
    bad = C.c_int32()
    ms = C.c_double()
    rc = L.eqd_precompute(C.byref(a), C.byref(bad), C.byref(ms))
    if rc != 0:
        raise EqdError(rc, f"eqd_precompute failed (element {bad.value})")
    return out, ms.value
