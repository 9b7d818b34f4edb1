"""Test infrastructure: ctypes front end of the C restatement (eqd_oracle.c).

``run_oracle(bundle, scalars, nsteps)`` advances the reference-layout state of
a bundle (as dumped by oracle/_ref/eqd_ref or produced by the case builder)
and returns the same named outputs the reference harness dumps.  Only tests and
the CPU-baseline leg of bench.py use it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _lib.eqdo_run.restype = C.c_int
        _lib.eqdo_run.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        _lib.eqdo_set_grouping.argtypes = [C.c_int32]
        _lib.eqdo_comdampv.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
        _lib.eqdo_qconstant.argtypes = [C.c_double, C.c_int32, C.c_double, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for f in (_lib.eqdo_slip_law, _lib.eqdo_ageing_law):
            f.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return _lib


class Scalars(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dt", "rdampk", "rdampm", "w", "grav", "roumax", "rhow", "gamar",
                 "ccosphi", "sinphi", "tv", "slipRateThres",
                 "xsource", "ysource", "zsource", "nucR", "nucRuptVel", "nucdtau0", "nucT",
                 "kapa_hg", "dx", "R", "vmaxPML")] + \
              [("PMLb", C.c_double * 8), ("pi", C.c_double)] + \
              [(n, C.c_int32) for n in
               ("nstep", "nn", "ne", "neq", "nf", "non", "noff", "nPML",
                "C_elastic", "C_Q", "C_hg", "C_nuclea", "friclaw", "TPV", "insertFaultType",
                "nucfault")] + [("fric_tp_h", C.c_double)]


_PTR_FIELDS = ["meshCoor", "eqNumIndexArr", "eqNumStartIndexLoc", "numOfDofPerNodeArr", "ien",
               "elemTypeArr", "mat", "eleshp", "eledet", "elemass", "ss", "phi",
               "stressCompIndexArr", "eleporep", "nodalMassArr", "fnms", "nsmp", "un", "us", "ud",
               "arn", "anonfs", "idhist", "stressArr", "pstrain", "v1", "velArr", "dispArr",
               "nodalForceArr", "fric", "fnft", "onFault", "offFault"]


class State(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in _PTR_FIELDS] + [
        ("nt", C.c_int32), ("timeElapsed", C.c_double), ("err_node", C.c_int32), ("err_nt", C.c_int32),
        ("tpHist", C.c_void_p)]


def make_scalars(sc: Dict[str, float], nsteps_total: int | None = None) -> Scalars:
    s = Scalars()
    for n in ("dt", "rdampk", "rdampm", "w", "grav", "roumax", "rhow", "gamar", "ccosphi", "sinphi",
              "tv", "slipRateThres", "xsource", "ysource", "zsource", "nucR", "nucRuptVel",
              "nucdtau0", "nucT", "kapa_hg", "dx", "R", "vmaxPML", "pi"):
        setattr(s, n, float(sc[n]))
    for i in range(8):
        s.PMLb[i] = float(sc[f"PMLb{i + 1}"])
    s.nstep = int(nsteps_total if nsteps_total is not None else sc["nstep"])
    s.nn = int(sc["totalNumOfNodes"])
    s.ne = int(sc["totalNumOfElements"])
    s.neq = int(sc["totalNumOfEquations"])
    s.nf = int(sc["nftnd"])
    s.non = int(sc["numOfOnFaultStCount"])
    s.noff = int(sc["numOfOffFaultStCount"])
    s.nPML = int(sc["nPML"])
    for n in ("C_elastic", "C_Q", "C_hg", "C_nuclea", "friclaw", "TPV", "insertFaultType", "nucfault"):
        setattr(s, n, int(sc[n]))
    # the reference never assigns its module variable fric_tp_h (src/globalvar.f90:42), so it is zero
    s.fric_tp_h = float(sc.get("fric_tp_h", 0.0))
    return s


def run_oracle(bundle: Dict[str, np.ndarray], sc: Dict[str, float], nsteps: int | None = None,
               grouping: int = 0) -> Dict[str, np.ndarray]:
    """Run ``nsteps`` steps (default: the case's nstep) from the bundle's
    initial state; returns copies of the evolved state and station series.
    ``grouping`` 1 combines each element's stiffness and hourglass forces
    before the node sums, as the GPU kernels do (0: the reference's order)."""
    L = lib()
    L.eqdo_set_grouping(int(grouping))
    nstep = int(sc["nstep"]) if nsteps is None else int(nsteps)
    s = make_scalars(sc, nstep)
    keep = {}

    def arr(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep[id(a)] = a
        return a

    b = {k: v for k, v in bundle.items()}
    st = State()
    ro = {
        "meshCoor": (b["meshCoor"], np.float64), "eqNumIndexArr": (b["eqNumIndexArr"], np.int32),
        "eqNumStartIndexLoc": (b["eqNumStartIndexLoc"], np.int32),
        "numOfDofPerNodeArr": (b["numOfDofPerNodeArr"], np.int32),
        "ien": (b["nodeElemIdRelation"], np.int32), "elemTypeArr": (b["elemTypeArr"], np.int32),
        "mat": (b["mat"], np.float64), "eleshp": (b["eleshp"], np.float64),
        "eledet": (b["eledet"], np.float64), "elemass": (b["elemass"], np.float64),
        "ss": (b["ss"], np.float64), "phi": (b["phi"], np.float64),
        "stressCompIndexArr": (b["stressCompIndexArr"], np.int32),
        "eleporep": (b["eleporep"], np.float64), "nodalMassArr": (b["nodalMassArr"], np.float64),
        "fnms": (b["fnms"], np.float64), "nsmp": (b["nsmp"], np.int32), "un": (b["un"], np.float64),
        "us": (b["us"], np.float64), "ud": (b["ud"], np.float64), "arn": (b["arn"], np.float64),
    }
    for k, (a, dt) in ro.items():
        setattr(st, k, arr(a, dt).ctypes.data)
    non = max(s.non, 1)
    anonfs = b.get("anonfs", np.zeros((non, 3), np.int32))
    st.anonfs = arr(anonfs, np.int32).ctypes.data
    idhist = b.get("idhist")
    st.idhist = arr(idhist, np.int32).ctypes.data if idhist is not None else None
    nn, ne, neq = s.nn, s.ne, s.neq
    out = {
        "stressArr": np.array(b["stressArr"], dtype=np.float64),
        "pstrain": np.array(b["pstrain"], dtype=np.float64),
        "v1": np.array(b["v1"], dtype=np.float64),
        "velArr": np.array(b.get("velArr", np.zeros((nn, 3))), dtype=np.float64),
        "dispArr": np.array(b.get("dispArr", np.zeros((nn, 3))), dtype=np.float64),
        "nodalForceArr": np.array(b.get("nodalForceArr", np.zeros(neq)), dtype=np.float64),
        "fric": np.array(b["fric"], dtype=np.float64),
        "fnft": np.array(b["fnft"], dtype=np.float64),
        "onFaultQuantHistSCECForm": np.zeros((non, nstep, 12)),
    }
    if s.noff > 0:
        out["OffFaultStGramSCEC"] = np.zeros((nstep, 6 * s.noff + 1))
    for k in ("stressArr", "pstrain", "v1", "velArr", "dispArr", "nodalForceArr", "fric", "fnft"):
        setattr(st, k, out[k].ctypes.data)
    st.onFault = out["onFaultQuantHistSCECForm"].ctypes.data
    st.offFault = out["OffFaultStGramSCEC"].ctypes.data if s.noff > 0 else None
    if s.friclaw == 5:
        out["tpHist"] = np.zeros((nstep, max(s.nf, 1), 2))
        st.tpHist = out["tpHist"].ctypes.data
    st.nt = 0
    st.timeElapsed = 0.0
    rc = L.eqdo_run(C.byref(s), C.byref(st), nstep)
    if rc != 0:
        raise FloatingPointError(f"oracle: velocity NaN at node {st.err_node} step {st.err_nt}")
    out["nt"] = st.nt
    out["timeElapsed"] = st.timeElapsed
    return out
