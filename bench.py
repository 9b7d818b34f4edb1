#!/usr/bin/env python3
"""Throughput of the EQdyna explicit time step on one MI355X.

Workload (default): BASELINE config 2 at full size -- 480 x 240 x 240 =
27,648,000 hexahedra at 50 m with a 10-element PML, a seeded five-layer medium,
the reference's Q model and a slip-weakening split-node fault -- built by
eqdyna_amd.mesh and stepped by libeqd.so.  A "step" is one explicit time step
(src/driver.f90:9-34) over the whole mesh, with the mesh resident in HBM.

Prints one JSON line:
  value       Mdof-steps/s = 3 * N_nodes * steps / t / 1e6 (N_nodes = grid nodes +
              split master nodes; PML split components are not counted)
  roofline    the element-force kernel: algorithmic bytes per launch (stored-
              operator model, per regular hexahedron: IEN 32 + dN/dx 192 +
              det 8 + lambda,mu 16 + stress R+W 96 + SS 48 + Phi 256 = 648 B in
              FP64, 340 B in FP32; + 96/48 B of Q memory variables) divided by
              its average launch time, measured with HIP events on the library's
              stream over the timed steps; peak = 8 TB/s HBM3E.
  cpu_baseline the reference Fortran (oracle/_ref/eqd_ref) on MPI ranks on the
              host's cores, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mdof-steps/s: plain FP64 vs FP32+learned RSF guess at equal rupture-time error"
HBM_PEAK_GBS = 8000.0
ELEM_BYTES = {"fp64": 648, "fp32": 340}
ELEM_BYTES_Q = {"fp64": 96, "fp32": 48}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str, steps_total: int):
    from eqdyna_amd import case, mesh
    if name == "c2":
        c = case.c2(steps=steps_total)
        desc = ("BASELINE config 2: 480x240x240 = 27,648,000 hexahedra at 50 m incl. 10-element PML, "
                "seeded 5-layer medium (Vs 2000->3464 m/s), Q (reference law), slip-weakening split-node fault")
    elif name == "c1":
        c = case.c1(steps=steps_total)
        desc = "BASELINE config 1: 120x60x60 = 432,000 hexahedra at 200 m incl. 10-element PML, homogeneous, SW fault"
    elif name == "c3":
        c = case.c3(steps=steps_total)
        desc = ("BASELINE config 3: 240x120x120 = 3,456,000 hexahedra at 100 m incl. 10-element PML, seeded fractal "
                "rough fault (rms/length 0.005, H=1), slip-law RSF with strong velocity weakening, seeded a field")
    else:
        raise SystemExit(f"unknown workload {name}")
    t = time.perf_counter()
    bundle, sc = mesh.build_case(c, nthreads=min(16, os.cpu_count() or 1))
    log(f"built {name}: {int(sc['totalNumOfElements'])} elements, {int(sc['totalNumOfNodes'])} nodes in "
        f"{time.perf_counter() - t:.1f} s")
    return c, bundle, sc, desc


def load_traffic(workload: str, precision: str):
    p = os.path.join(ROOT, "profiles", "element_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        e = d.get(f"{workload}_{precision}")
        if not e:
            return None
        return float(e["bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline_reference(ranks: int, steps: int):
    """Reference Fortran on `ranks` MPI ranks, C2 physics on a bounded sub-box."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from eqdyna_amd import case, mesh
    if not refrun.available() or not os.path.exists(refrun.MPIEXEC):
        return None
    c = case.c2_sample(steps=steps)
    nn = mesh.mesh_sizes(c).nn  # one-rank node count: grid nodes + split masters
    fac = {1: (1, 1, 1), 2: (2, 1, 1), 4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 32: (4, 4, 2)}
    if ranks not in fac:
        ranks = 16
    c.nx, c.ny, c.nz = fac[ranks]
    with tempfile.TemporaryDirectory() as d:
        t = time.perf_counter()
        _, _, _, t_loop = refrun.run_reference(c, d, ranks, dump=False, timeout=900)
        log(f"reference on {ranks} ranks: t_loop {t_loop:.2f} s ({time.perf_counter() - t:.1f} s with meshing)")
    return {
        "value": 3.0 * nn * steps / t_loop / 1e6, "unit": "Mdof-steps/s", "cores": ranks, "kind": "reference",
        "sample": f"C2 physics (50 m, 5-layer medium, Q, SW fault) on a 240x120x120 = 3,456,000-hexahedron "
                  f"sub-box, {steps} steps, reference Fortran (amdflang -O2) on {ranks} MPI ranks, time loop only",
    }


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
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give the
    scale factor between counter readings and bytes moved."""
    import torch
    n = nbytes // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "launches": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c3"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--cpu-ranks", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (for scaling HBM counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "The halo-exchange design for several GPUs is described in DESIGN.md.", file=sys.stderr)
        return 2

    from eqdyna_amd.solver import Solver
    total = args.warmup + args.steps
    c, bundle, sc, desc = build_workload(args.workload, total)
    n_reg = int((bundle["elemTypeArr"] != 2).sum())
    nn = int(sc["totalNumOfNodes"])
    t = time.perf_counter()
    s = Solver(bundle, sc, args.precision, device=0, nstep=total)
    layout = s.layout_info()
    log(f"uploaded ({s.device_bytes() / 1e9:.1f} GB on the device) in {time.perf_counter() - t:.1f} s; layout {layout}")
    del bundle
    s.run(args.warmup)
    s.set_timing(True)
    t0 = time.perf_counter()
    s.run(args.steps)  # returns after the stream has drained
    t1 = time.perf_counter()
    wall = t1 - t0
    ktimes = {k: s.kernel_time(k) for k in ("element", "element_pml", "fault", "node", "node_pml", "station")}
    nt, _ = s.clock()
    s.close()
    for k, (ms, n) in ktimes.items():
        log(f"kernel {k}: {ms:.4f} ms avg over {n} launches")
    value = 3.0 * nn * args.steps / wall / 1e6
    elem_ms = ktimes["element"][0]
    per_elem = ELEM_BYTES[args.precision] + (ELEM_BYTES_Q[args.precision] if c.C_Q == 1 else 0)
    alg_bytes = per_elem * n_reg
    achieved = alg_bytes / (elem_ms * 1e-3) / 1e9 if elem_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(args.workload, args.precision),
                "kernel": "elem_regular_kernel", "algorithmic_bytes_per_launch": alg_bytes,
                "avg_launch_ms": round(elem_ms, 5)}

    cpu = None
    if args.cpu_baseline in ("auto", "reference"):
        try:
            cpu = cpu_baseline_reference(args.cpu_ranks, args.cpu_steps)
        except Exception as e:  # noqa: BLE001
            log(f"reference CPU baseline failed: {e}")
            cpu = None
    if cpu is None and args.cpu_baseline in ("auto", "port"):
        cpu = cpu_baseline_port(20)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mdof-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32 bulk / f64 fault",
        "data": "synthetic (seeded case generator; no external data)",
        "config": {"workload": desc, "elements": int(sc["totalNumOfElements"]), "nodes": nn,
                   "regular_elements": n_reg, "fault_pairs": int(sc["nftnd"]), "precision": args.precision,
                   "steps_run": nt, "operator_classes": {k: v for k, v in layout.items() if "classes" in k}},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
