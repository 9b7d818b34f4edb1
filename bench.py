#!/usr/bin/env python3
"""Throughput of the EQdyna explicit time step on one MI355X.

Default workload (c4): the rough-fault rate-and-state box c3 (240 x 120 x 120 =
3,456,000 hexahedra at 100 m with a 10-element PML, a seeded fractal rough
fault, slip-law rate-and-state friction with strong velocity weakening,
dt = 0.005 s, 3000 steps) plus the learned initial guess of the
rate-and-state Newton solve -- run twice on the GPU: the plain FP64 mode
(the reference's arithmetic and Newton solve) and the FP32-bulk mode with
the learned guess (eqdyna_amd/rsf_guess_c3.npz).  `value` is the FP32+learned
throughput; the FP64 throughput, their ratio and the FP32+learned run's
rupture-time and peak-slip-rate errors against the FP64 run are in `config`.
Other workloads (c1, c2, c3) time one run in the --precision given; c5 is
the batched ensemble (64 seeded scenarios of uneven size, FP32 bulk / FP64
fault, advanced together as one graph-captured step per time step).
Before the timed steps the launch-shape tuner (eqd_tune) picks block sizes
and XCD remapping per kernel on the workload itself (--tune 0 skips it).

A "step" is one explicit time step (src/driver.f90:9-34) over the whole mesh,
with the mesh resident in HBM.  Prints one JSON line:
  value       Mdof-steps/s = 3 * N_nodes * steps / t / 1e6 (N_nodes = grid nodes +
              split master nodes; PML split components are not counted)
  roofline    the run's dominant bulk kernel (largest total time among the
              HBM-bound element and node kernels; the fault kernel is
              latency-bound, runs beside the node kernels and has no byte
              model): its algorithmic bytes per launch (kernel_bytes: every
              array it needs, in the layout the library chose for the mesh,
              moved once) divided by its average launch time, measured with
              HIP events its own dispatches record on every
              --timing-every-th timed step; peak = 8 TB/s, beside a measured
              device copy; and the whole step's algorithmic bytes over the
              time per step.
  cpu_baseline the reference Fortran (oracle/_ref/eqd_ref) on MPI ranks on the
              host's cores, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mdof-steps/s: plain FP64 vs FP32+learned RSF guess at equal rupture-time error"
HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"element": "elem_regular_kernel", "element_pml": "elem_pml_kernel", "node": "node3_kernel",
                "node_pml": "node_pml_kernel"}
KERNEL_KEY = {v: k for k, v in KERNEL_NAMES.items()}
GUESS = {"k_fallback": 6, "gate": 0.5}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str, steps_total: int):
    from eqdyna_amd import case, mesh
    if name == "c2":
        c = case.c2(steps=steps_total)
        desc = ("c2: 480x240x240 = 27,648,000 hexahedra at 50 m incl. 10-element PML, "
                "seeded 5-layer medium (Vs 2000->3464 m/s), Q (reference law), slip-weakening split-node fault")
    elif name == "c1":
        c = case.c1(steps=steps_total)
        desc = "c1: 120x60x60 = 432,000 hexahedra at 200 m incl. 10-element PML, homogeneous, SW fault"
    elif name in ("c3", "c4"):
        c = case.c3(steps=steps_total)
        desc = ("c3: 240x120x120 = 3,456,000 hexahedra at 100 m incl. 10-element PML, seeded fractal "
                "rough fault (rms/length 0.005, H=1), slip-law RSF with strong velocity weakening, seeded a field")
        if name == "c4":
            desc = ("c4: c3 plus the learned RSF Newton guess (6-64-64-64-1 GELU, bf16 MFMA), "
                    "FP64 reference mode vs FP32 bulk + learned guess; " + desc[len("c3: "):])
    else:
        raise SystemExit(f"unknown workload {name}")
    t = time.perf_counter()
    bundle, sc = mesh.build_case(c, nthreads=min(16, os.cpu_count() or 1))
    log(f"built {name}: {int(sc['totalNumOfElements'])} elements, {int(sc['totalNumOfNodes'])} nodes in "
        f"{time.perf_counter() - t:.1f} s")
    return c, bundle, sc, desc


def load_traffic(workload: str, precision: str, kernel: str):
    """Counted bytes per launch of a kernel from profiles/kernel_traffic.json
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes of the same workload) and
    where they come from, or (None, None)."""
    p = os.path.join(ROOT, "profiles", "kernel_traffic.json")
    if not os.path.exists(p):
        return None, None
    try:
        d = json.load(open(p))
        e = d.get(f"{workload}_{precision}_{kernel}")
        if not e:
            return None, None
        return float(e["bytes_per_launch"]), f"profiles/kernel_traffic.json ({d.get('_collected', 'undated')})"
    except Exception:  # noqa: BLE001
        return None, None


def run_one(bundle, sc, precision: str, warmup: int, steps: int, guess=None, tune: int = 0, timing_every: int = 10):
    """One timed run: warm-up steps, the tuner (state restored), then `steps`
    timed, with per-kernel events on every `timing_every`-th of them."""
    from eqdyna_amd.solver import Solver
    t = time.perf_counter()
    s = Solver(bundle, sc, precision, device=0, nstep=warmup + steps)
    if guess is not None:
        s.set_guess(guess, GUESS["k_fallback"], GUESS["gate"])
    layout = s.layout_info()
    log(f"{precision}{'+learned' if guess is not None else ''}: uploaded ({s.device_bytes() / 1e9:.1f} GB on the "
        f"device) in {time.perf_counter() - t:.1f} s; layout {layout}")
    s.run(max(warmup - 1, 0))
    tuned = None
    if tune > 0 and steps > tune + 1:
        tuned = s.tune(tune)
        log(f"  tuner: " + ", ".join(f"{k} block {v['block']} remap {v['remap']} ({v['ms']:.4f} ms)"
                                     for k, v in tuned.items()))
    s.set_timing(timing_every)
    s.run(min(warmup, 1))  # the launch tables for the chosen shapes are built here, not in the timed steps
    t0 = time.perf_counter()
    s.run(steps)  # returns after the stream has drained
    wall = time.perf_counter() - t0
    ktimes = {k: s.kernel_time(k) for k in ("element", "element_pml", "fault", "node", "node_pml", "station")}
    r = {"wall": wall, "ktimes": ktimes, "layout": layout, "nt": s.clock()[0], "newton": s.newton_stats(),
         "fnft": s.get("fnft"), "fric": s.get("fric"), "launch": s.get_launch()}
    s.close()
    for k, (ms, n) in ktimes.items():
        log(f"  kernel {k}: {ms:.4f} ms avg over {n} launches")
    return r


def run_graph(bundle, sc, precision: str, warmup: int, steps: int, guess=None, launch=None) -> float:
    """The same run with every step replayed from one captured HIP graph (no
    per-kernel events); returns its wall time."""
    from eqdyna_amd.solver import Solver
    s = Solver(bundle, sc, precision, device=0, nstep=warmup + steps)
    if guess is not None:
        s.set_guess(guess, GUESS["k_fallback"], GUESS["gate"])
    for k, (b, r) in (launch or {}).items():
        s.set_launch(k, b, r)
    s.set_graph(True)
    s.run(warmup)  # captures and instantiates the step graph outside the timed region
    t0 = time.perf_counter()
    s.run(steps)
    wall = time.perf_counter() - t0
    s.close()
    return wall


def kernel_bytes(bundle, sc, layout: dict, precision: str) -> dict:
    """Algorithmic bytes per launch of each bulk kernel: every array the
    kernel needs, in the layout the library chose for this mesh
    (layout_info), read or written once -- each element's operators (class
    indices when elements share class tables held in LDS; stored rows with
    the full or the compact hourglass operators; or the geometry rows the
    FP32 mode rebuilds them from), connectivity, stresses and forces once per
    element, each node's state once per kernel that touches it, each
    element-to-node force once.  Counts come from the mesh itself."""
    import numpy as np
    f32 = precision == "fp32"
    r = 4 if f32 else 8               # element operator / stress / force values
    nv = 4 if f32 else 8              # node velocity values
    nu_r, nu_w = (4, 6) if f32 else (8, 8)  # displacement: gathered (FP32: hi), read and written (hi + bf16 lo)
    mb = 4 if f32 else 8              # nodal mass as the node kernels read it
    ien = np.asarray(bundle["nodeElemIdRelation"]).reshape(-1, 8) - 1
    et = np.asarray(bundle["elemTypeArr"])
    ndof = np.asarray(bundle["numOfDofPerNodeArr"])
    pml = et == 2
    nr, npl = int((~pml).sum()), int(pml.sum())
    split = np.zeros(len(ndof), dtype=bool)
    split[np.asarray(bundle["nsmp"]).ravel() - 1] = True
    q, dp, hg = int(sc["C_Q"]) == 1, int(sc["C_elastic"]) == 0, int(sc["C_hg"])
    mass_term = float(sc.get("rdampm", 0.0)) != 0.0 or dp

    def ops(kind: str) -> int:
        if layout.get(f"op_classes_{kind}", 0) > 0:
            return 8  # two 32-bit class indices; the tables sit in LDS
        if layout.get(f"geom_ops_{kind}", 0):
            return 23 * r  # 21 relative coordinates, lambda, mu
        hg_vals = (16 if layout.get(f"phi_compact_{kind}", 0) else 32) if hg == 1 else 2
        return (24 + 3 + (6 if hg == 1 else 0) + hg_vals) * r

    occ = np.bincount(ien.ravel(), minlength=len(ndof))        # element-node entries per node
    occ_reg = np.bincount(ien[~pml].ravel(), minlength=len(ndof))
    occ_pml = np.bincount(ien[pml].ravel(), minlength=len(ndof))
    node3 = (ndof == 3) & ~split
    nodepml = (ndof == 12) & (occ_pml > 0)
    out = {}
    ns = 12 if q else 6
    out["element"] = (nr * (32 + ops("regular") + 2 * ns * r + 24 * r + (1 if q else 0) + (24 if dp else 0)
                            + (24 * r if mass_term else 0))
                      + int((occ_reg > 0).sum()) * 3 * (nv + nu_r))
    # PML element: a PML node takes 6 values of its slot, a 3-component node 3
    pml_slots = int(np.where(ndof[ien[pml]] == 12, 6, 3).sum()) if npl else 0
    out["element_pml"] = (npl * (32 + ops("pml") + 2 * 21 * r + 3 * r + 1 + 6 * r + (24 * r if mass_term else 0))
                          + pml_slots * r + int((occ_pml > 0).sum()) * 3 * (nv + nu_r))
    n3, e3 = int(node3.sum()), int(occ[node3].sum())
    if layout.get("node3_slot_major", 0):
        # node id, mass, state; all 8 entry slots of the slot-major table; the forces
        out["node"] = n3 * (4 + mb + 3 * 2 * nv + 3 * 2 * nu_w + 8 * 4) + e3 * 3 * r
    else:
        out["node"] = n3 * (4 + 4 + mb + 3 * 2 * nv + 3 * 2 * nu_w) + e3 * (4 + 3 * r)
    npn, ep = int(nodepml.sum()), int(occ[nodepml].sum())
    out["node_pml"] = (npn * (4 + 4 + mb + 3 * mb + 12 * 2 * nv + 3 * 2 * nv + 3 * 2 * nu_w)
                       + ep * (4 + 6 * r) + npl * 6 * r)
    return {k: int(v) for k, v in out.items()}


def measure_copy_gbs(nbytes: int = 1 << 32, reps: int = 5) -> float:
    """Device-to-device copy bandwidth of this GPU, (read + write) bytes over
    the best of `reps` timed copy kernels of `nbytes` (torch's copy kernel,
    timed with HIP events), in a separate process so that torch's HIP runtime
    never shares a process with libeqd.so's."""
    import subprocess
    code = ("import torch,json;n=%d//4;a=torch.ones(n,device='cuda');b=torch.empty_like(a);b.copy_(a);"
            "e0=torch.cuda.Event(enable_timing=True);e1=torch.cuda.Event(enable_timing=True);best=0.0\n"
            "for _ in range(%d):\n e0.record();b.copy_(a);e1.record();e1.synchronize();"
            "best=max(best,2*n*4/(e0.elapsed_time(e1)*1e-3)/1e9)\nprint(json.dumps(best))" % (nbytes, reps))
    try:
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
        return float(p.stdout.strip().splitlines()[-1]) if p.returncode == 0 else 0.0
    except Exception:  # noqa: BLE001
        return 0.0


def roofline(run, kbytes: dict, precision: str, workload: str, ms_per_step: float, copy_gbs: float):
    """Roofline of the run's dominant bulk kernel (largest total time among
    the kernels with a byte model: element, PML element, node, PML node): its
    algorithmic bytes per launch (kernel_bytes) over its average launch time
    from the HIP events its own dispatches record; beside it the whole step's
    algorithmic bytes over the measured time per step, and the measured copy
    bandwidth."""
    tot = {k: ms * n for k, (ms, n) in run["ktimes"].items() if k in kbytes}
    k = max(tot, key=tot.get)
    ms = run["ktimes"][k][0]
    alg = kbytes[k]
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic, source = load_traffic(workload, precision, k)
    step_bytes = sum(kbytes.values())
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "kernel": KERNEL_NAMES[k], "algorithmic_bytes_per_launch": int(alg), "avg_launch_ms": round(ms, 5),
           "copy_gbs_measured": round(copy_gbs, 1),
           "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs > 0 else None,
           "step": {"algorithmic_bytes": int(step_bytes), "ms": round(ms_per_step, 5),
                    "gbs": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1) if ms_per_step > 0 else None,
                    "frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms_per_step > 0 else None,
                    "per_kernel_bytes": kbytes}}
    if traffic and ms > 0:
        # bytes the L2 fetched from and wrote to memory for the same kernel
        # (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a device copy), over
        # the same launch time; an upper bound on DRAM bytes (the Infinity
        # Cache serves some of the fetches)
        out["traffic_source"] = source
        out["traffic_gbs"] = round(traffic / (ms * 1e-3) / 1e9, 1)
        out["traffic_over_algorithmic"] = round(traffic / alg, 3)
    return out


def errors_vs(run, ref, dt):
    import numpy as np
    a, b = run["fnft"], ref["fnft"]
    both = (a < 5000) & (b < 5000)
    d = np.abs(a[both] - b[both]) / dt if both.any() else np.zeros(1)
    pk_a, pk_b = run["fric"][:, 75], ref["fric"][:, 75]
    return {"rupture_time_mean_dt": float(d.mean()), "rupture_time_max_dt": float(d.max()),
            "ruptured_nodes": int((b < 5000).sum()), "rupture_flips": int(((a < 5000) != (b < 5000)).sum()),
            "peak_slip_rate_rel": float(np.max(np.abs(pk_a - pk_b)) / max(float(np.max(pk_b)), 1e-300))}


def cpu_baseline_reference(workload: str, ranks: int, steps: int):
    """Reference Fortran on `ranks` MPI ranks on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from eqdyna_amd import case, mesh
    binary = refrun.REF_BIN_O3 if os.path.exists(refrun.REF_BIN_O3) else refrun.REF_BIN
    build = "amdflang -O3" if binary == refrun.REF_BIN_O3 else "amdflang -O2 (parity build; the -O3 build is absent)"
    if not os.path.exists(binary) or not os.path.exists(refrun.MPIEXEC):
        return None
    if workload in ("c3", "c4"):
        c = case.c3(steps=steps)
        sample = (f"c3 at full size (3,456,000 hexahedra, rough fault, slip-law RSF), first {steps} of 3000 "
                  f"steps (through the onset of rupture), plain Newton guess")
    elif workload == "c1":
        c = case.c1(steps=steps)
        sample = f"c1 at full size (432,000 hexahedra at 200 m, SW fault), first {steps} steps"
    elif workload == "c5":
        c = case.c5(steps=steps)[0]
        sample = f"scenario 0 of the c5 ensemble ({c.name}, {'slip-law RSF' if c.friclaw == 4 else 'slip weakening'}), {steps} steps"
    else:
        c = case.c2_sample(steps=steps)
        sample = (f"C2 physics (50 m, 5-layer medium, Q, SW fault) on a 240x120x120 = 3,456,000-hexahedron "
                  f"sub-box, {steps} steps")
    nn = mesh.mesh_sizes(c).nn  # one-rank node count: grid nodes + split masters
    fac = {1: (1, 1, 1), 2: (2, 1, 1), 4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 32: (4, 4, 2)}
    if ranks not in fac:
        ranks = 16
    c.nx, c.ny, c.nz = fac[ranks]
    with tempfile.TemporaryDirectory() as d:
        t = time.perf_counter()
        _, _, _, t_loop = refrun.run_reference(c, d, ranks, dump=False, timeout=900, binary=binary)
        log(f"reference on {ranks} ranks: t_loop {t_loop:.2f} s ({time.perf_counter() - t:.1f} s with meshing); "
            f"start-up phases {refrun.LAST_PHASES}")
    return {"value": 3.0 * nn * steps / t_loop / 1e6, "unit": "Mdof-steps/s", "cores": ranks, "kind": "reference",
            "sample": sample + f", reference Fortran ({build}) on {ranks} MPI ranks, time loop only",
            "startup_s": dict(refrun.LAST_PHASES)}


def cpu_baseline_port(steps: int):
    """C restatement (single core) on the C1 mesh, when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from eqdyna_amd import case, mesh
    c = case.c1(steps=steps)
    b, sc = mesh.build_case(c)
    t = time.perf_counter()
    oracle.run_oracle(b, sc, nsteps=steps)
    dt = time.perf_counter() - t
    return {"value": 3.0 * sc["totalNumOfNodes"] * steps / dt / 1e6, "unit": "Mdof-steps/s", "cores": 1,
            "kind": "port", "sample": f"C1 mesh (432,000 hexahedra), {steps} steps, C restatement on one core"}


def calibrate_copy(nbytes: int) -> int:
    """Kernels of known traffic for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes: a plain copy of `nbytes` (read nbytes, write nbytes),
    then a gather of the same array's 128-byte rows in a random order
    (index_select: read nbytes in whole lines plus 8 bytes of index per row,
    write nbytes), the access pattern of the node and element gathers."""
    import torch
    n = nbytes // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    rows = a.view(-1, 32)
    perm = torch.randperm(rows.shape[0], device="cuda")
    for _ in range(3):
        torch.index_select(rows, 0, perm, out=b.view(-1, 32))
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "gather_read_bytes": n * 4 + 8 * rows.shape[0],
                      "gather_write_bytes": n * 4, "launches": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (c4: 2990, others: 200)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--timing-every", type=int, default=10,
                    help="per-kernel HIP events on every k-th timed step (1: every step; the events cost device "
                         "time at the kernel boundaries they mark)")
    ap.add_argument("--workload", default="c4", choices=["c4", "c3", "c2", "c1", "c5"])
    ap.add_argument("--tune", type=int, default=4, help="steps per candidate of the launch-shape tuner (0: off)")
    ap.add_argument("--scenarios", type=int, default=64, help="c5 only: ensemble size")
    ap.add_argument("--c5-guess", default="off", choices=["on", "off"],
                    help="c5 only: the rate-and-state scenarios start Newton from the learned guess")
    ap.add_argument("--host-startup", type=int, default=0, metavar="RANKS",
                    help="only time the reference host's start-up (meshgen, assembleGlobalMass) on the workload's "
                         "full mesh on RANKS MPI ranks, and the device precompute of the same mesh, then exit")
    ap.add_argument("--graph", type=int, default=1,
                    help="also time the run (c4: both modes) replayed from a captured step")
    ap.add_argument("--precompute", type=int, default=1,
                    help="also time the device-side element precompute (eqd_precompute) on the workload's mesh")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"], help="c1-c3 only")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--cpu-ranks", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=60, help="reference steps timed (c3/c4: at least this, through "
                    "the onset of rupture unless --cpu-steps-auto 0)")
    ap.add_argument("--cpu-steps-auto", type=int, default=1)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (for scaling HBM counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Several GPUs would each take a slab of the mesh with a halo exchange; that path is not built.", file=sys.stderr)
        return 2
    if args.host_startup:
        return host_startup(args.workload, args.host_startup)
    if args.workload == "c5":
        return bench_c5(args)
    steps = args.steps if args.steps is not None else (2990 if args.workload == "c4" else 200)
    total = args.warmup + steps
    c, bundle, sc, desc = build_workload(args.workload, total)
    et = bundle["elemTypeArr"]
    n_regular = int((et != 2).sum())
    nn = int(sc["totalNumOfNodes"])
    sc_dt = float(sc["dt"])
    extra = {}
    if args.precompute:
        extra["device_precompute"] = bench_precompute(bundle)
    if args.workload == "c4":
        from eqdyna_amd import learned
        params, meta = learned.load()
        ref = run_one(bundle, sc, "fp64", args.warmup, steps, tune=args.tune, timing_every=args.timing_every)
        run = run_one(bundle, sc, "fp32", args.warmup, steps, guess=params, tune=args.tune,
                      timing_every=args.timing_every)
        precision = "fp32"
        v64 = 3.0 * nn * steps / ref["wall"] / 1e6
        n64, n32 = ref["newton"], run["newton"]
        extra.update({
            "fp64_value": round(v64, 2), "fp64_ms_per_step": round(ref["wall"] / steps * 1e3, 4),
            "ratio_fp32_learned_vs_fp64": None,
            "errors_vs_fp64": errors_vs(run, ref, sc["dt"]),
            "newton_fp64": {**n64, "iterations_per_solve": round(n64["iterations"] / max(n64["solves"], 1), 3)},
            "newton_fp32_learned": {**n32, "iterations_per_solve": round(n32["iterations"] / max(n32["solves"], 1), 3),
                                    "fallback_rate": round(n32["fallbacks"] / max(n32["solves"], 1), 5)},
            "guess": {**GUESS, "weights": "eqdyna_amd/rsf_guess_c3.npz", "training_rows": meta.get("rows")},
        })
    else:
        precision = args.precision
        run = run_one(bundle, sc, precision, args.warmup, steps, tune=args.tune, timing_every=args.timing_every)
    value = 3.0 * nn * steps / run["wall"] / 1e6
    if "fp64_value" in extra:
        extra["ratio_fp32_learned_vs_fp64"] = round(value / extra["fp64_value"], 4)
    if args.graph and args.workload == "c4":
        w64 = run_graph(bundle, sc, "fp64", args.warmup, steps, launch=ref["launch"])
        w32 = run_graph(bundle, sc, "fp32", args.warmup, steps, guess=params, launch=run["launch"])
        g64, g32 = 3.0 * nn * steps / w64 / 1e6, 3.0 * nn * steps / w32 / 1e6
        log(f"graph replay: fp64 {g64:.1f}, fp32+learned {g32:.1f} Mdof-steps/s")
        extra["graph_replay"] = {"fp64_value": round(g64, 2), "fp32_learned_value": round(g32, 2),
                                 "ratio": round(g32 / g64, 4),
                                 "note": "same steps replayed from one captured step; no per-kernel events"}
    elif args.graph:
        wg = run_graph(bundle, sc, precision, args.warmup, steps, launch=run["launch"])
        g = 3.0 * nn * steps / wg / 1e6
        log(f"graph replay: {precision} {g:.1f} Mdof-steps/s")
        extra["graph_replay"] = {"value": round(g, 2), "ms_per_step": round(wg / steps * 1e3, 4),
                                 "note": "same steps replayed from one captured step; no per-kernel events"}
    kbytes = kernel_bytes(bundle, sc, run["layout"], precision)
    copy_gbs = measure_copy_gbs()
    log(f"device copy: {copy_gbs:.1f} GB/s (read + write)")
    roof = roofline(run, kbytes, precision, args.workload, run["wall"] / steps * 1e3, copy_gbs)
    every = "every timed step" if args.timing_every == 1 else f"one timed step in {args.timing_every}"
    roof["kernel_times_from"] = (f"HIP events of the kernels' own dispatches on {every} "
                                 f"({run['ktimes'][KERNEL_KEY[roof['kernel']]][1]} launches)")
    del bundle

    cpu = None
    cpu_steps = args.cpu_steps
    if args.workload in ("c3", "c4") and args.cpu_steps_auto:
        # through the onset of rupture: 40 steps past the first ruptured node of this run
        ft = run["fnft"][run["fnft"] < 5000]
        if ft.size:
            cpu_steps = int(min(400, max(args.cpu_steps, round(float(ft.min()) / sc_dt) + 40)))
    if args.cpu_baseline in ("auto", "reference"):
        try:
            cpu = cpu_baseline_reference(args.workload, args.cpu_ranks, cpu_steps)
        except Exception as e:  # noqa: BLE001
            log(f"reference CPU baseline failed: {e}")
            cpu = None
    if cpu is None and args.cpu_baseline in ("auto", "port"):
        cpu = cpu_baseline_port(20)

    if precision == "fp64":
        dtype = "f64"
    else:
        dtype = "f32 bulk / f64 fault" + (" / bf16 guess" if args.workload == "c4" else "")
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mdof-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(run["wall"] / steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded case generator; no external data)",
        "config": {"workload": desc, "elements": int(sc["totalNumOfElements"]), "nodes": nn,
                   "regular_elements": n_regular, "fault_pairs": int(sc["nftnd"]), "precision": precision,
                   "steps_run": run["nt"],
                   "layout": run["layout"],
                   "launch_shapes": {k: {"block": b, "remap": r} for k, (b, r) in run["launch"].items()}, **extra},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def host_startup(workload: str, ranks: int) -> int:
    """The reference host's start-up phases on the workload's full mesh (one
    step), beside the device precompute of the same mesh."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from eqdyna_amd import case, mesh
    c = {"c1": case.c1, "c2": case.c2, "c3": case.c3, "c4": case.c3}[workload](steps=1)
    fac = {1: (1, 1, 1), 2: (2, 1, 1), 4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 32: (4, 4, 2)}
    c.nx, c.ny, c.nz = fac[ranks]
    with tempfile.TemporaryDirectory() as d:
        t = time.perf_counter()
        refrun.run_reference(c, d, ranks, dump=False, timeout=1800)
        wall = time.perf_counter() - t
    phases = dict(refrun.LAST_PHASES)
    log(f"reference start-up on {ranks} ranks: {phases} ({wall:.1f} s for the whole run)")
    c.nx = c.ny = c.nz = 1
    t = time.perf_counter()
    bundle, sc = mesh.build_case(c, nthreads=min(16, os.cpu_count() or 1))
    build = time.perf_counter() - t
    dev = bench_precompute(bundle)
    print(json.dumps({"workload": workload, "elements": int(sc["totalNumOfElements"]), "reference_ranks": ranks,
                      "reference_startup_s": phases, "reference_run_wall_s": round(wall, 2),
                      "host_builder_s": round(build, 2), "device_precompute": dev}), flush=True)
    return 0


def bench_precompute(bundle) -> dict:
    """Device-side element precompute and nodal masses of the workload's
    mesh (the reference host's assembleGlobalMass element loop), checked bit
    for bit against the host restatement that built the bundle."""
    import numpy as np
    from eqdyna_amd import solver
    t = time.perf_counter()
    out, ms = solver.precompute(bundle)
    wall = time.perf_counter() - t
    same = all(np.array_equal(out[k].ravel(), np.asarray(bundle[k], dtype=np.float64).ravel())
               for k in ("eleshp", "eledet", "elemass", "ss", "phi", "nodalMassArr", "fnms"))
    log(f"device precompute: kernels {ms:.2f} ms, {wall:.2f} s with copies and checks; bitwise equal: {same}")
    return {"kernel_ms": round(ms, 3), "wall_s_with_copies": round(wall, 3), "bitwise_equal_to_host": bool(same)}


def bench_c5(args) -> int:
    """The batched ensemble: every scenario built and uploaded, advanced
    together (one launch per kind per step over all scenarios, the step
    captured as a HIP graph and replayed); value = sum over scenarios of
    3 * N_nodes * steps / wall time.  Per-kernel times for the roofline come
    from an uncaptured pass of --kt steps after the timed ones (events cannot
    sit inside a captured step)."""
    from eqdyna_amd import case, mesh
    from eqdyna_amd.solver import Ensemble, Solver
    steps = args.steps if args.steps is not None else 200
    kt = 20
    kl = 40  # an untimed launched (uncaptured) pass, for comparison with the replayed steps
    total = args.warmup + steps + kl + kt
    cases = case.c5(n=args.scenarios, steps=total)
    guess = None
    if args.c5_guess == "on":
        from eqdyna_amd import learned
        guess, _ = learned.load()
    members = []
    kbytes = {"element": 0, "element_pml": 0, "node": 0, "node_pml": 0}
    nn = ne = nf = 0
    sizes = []
    t = time.perf_counter()
    for c in cases:
        bundle, sc = mesh.build_case(c, nthreads=min(16, os.cpu_count() or 1))
        nn += int(sc["totalNumOfNodes"])
        ne += int(sc["totalNumOfElements"])
        nf += int(sc["nftnd"])
        sizes.append(int(sc["totalNumOfElements"]))
        members.append(Solver(bundle, sc, "fp32", device=0, nstep=total))
        if guess is not None and c.friclaw >= 3:
            members[-1].set_guess(guess, GUESS["k_fallback"], GUESS["gate"])
        for k, v in kernel_bytes(bundle, sc, members[-1].layout_info(), "fp32").items():
            kbytes[k] += v
        del bundle
    dev_gb = sum(m.device_bytes() for m in members) / 1e9
    log(f"c5: {len(members)} scenarios, {ne} elements, {nn} nodes, {nf} fault pairs, {dev_gb:.1f} GB on the device, "
        f"built and uploaded in {time.perf_counter() - t:.1f} s")
    ens = Ensemble(members)
    ens.run(args.warmup)
    tuned = None
    if args.tune > 0:
        t = time.perf_counter()
        tuned = ens.tune(args.tune)
        log(f"  tuner ({time.perf_counter() - t:.1f} s): " +
            ", ".join(f"{k} block {v['block']} remap {v['remap']} ({v['ms']:.4f} ms)" for k, v in tuned.items()))
    t0 = time.perf_counter()
    ens.run(steps)  # graph replay; returns after the stream has drained
    wall = time.perf_counter() - t0
    ens.set_graph(False)
    t0 = time.perf_counter()
    ens.run(kl)
    wall_l = time.perf_counter() - t0
    log(f"  launched without timers: {wall_l / kl * 1e3:.4f} ms per step (replayed {wall / steps * 1e3:.4f})")
    ens.set_timing(True)
    ens.run(kt)
    ktimes = {k: ens.kernel_time(k) for k in ("element", "element_pml", "fault", "node", "node_pml", "station")}
    for k, (ms, n) in ktimes.items():
        log(f"  kernel {k}: {ms:.4f} ms avg over {n} launches")
    ens.close()
    fric = [m.s.friclaw for m in members]
    for m in members:
        m.close()
    value = 3.0 * nn * steps / wall / 1e6
    copy_gbs = measure_copy_gbs()
    log(f"device copy: {copy_gbs:.1f} GB/s (read + write)")
    roof = roofline({"ktimes": ktimes}, kbytes, "fp32", "c5", wall / steps * 1e3, copy_gbs)
    roof["kernel_times_from"] = f"uncaptured pass of {kt} steps after the timed steps"
    cpu = None
    if args.cpu_baseline in ("auto", "reference"):
        try:
            cpu = cpu_baseline_reference("c5", args.cpu_ranks, args.cpu_steps)
        except Exception as e:  # noqa: BLE001
            log(f"reference CPU baseline failed: {e}")
    out = {
        "metric": "Mdof-steps/s (aggregate over the ensemble)", "value": round(value, 2), "unit": "Mdof-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 bulk / f64 fault",
        "data": "synthetic (seeded case generator; no external data)",
        "config": {"workload": f"c5: batched ensemble of {len(members)} seeded scenarios at 100 m (8-24 x 6-12 x 6-12 km "
                               "incl. a 10-element PML, 5-layer media, von Karman tau_0, SW or slip-law RSF), "
                               "one HIP-graph-captured step over all scenarios",
                   "scenarios": len(members), "elements": ne, "nodes": nn, "fault_pairs": nf,
                   "elements_min": min(sizes), "elements_max": max(sizes),
                   "slip_weakening": sum(1 for f in fric if f == 1), "rate_and_state": sum(1 for f in fric if f >= 3),
                   "rate_and_state_guess": ({**GUESS, "weights": "eqdyna_amd/rsf_guess_c3.npz"} if guess is not None
                                            else "plain (the reference's Newton start)"),
                   "device_gb": round(dev_gb, 1), "tuner": tuned,
                   "launched_untimed": {"steps": kl, "value": round(3.0 * nn * kl / wall_l / 1e6, 2),
                                        "ms_per_step": round(wall_l / kl * 1e3, 4),
                                        "note": "the same ensemble step launched, not replayed, no per-kernel events"}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
