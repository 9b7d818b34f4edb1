"""Learned initial guess of the rate-and-state Newton solve: harvesting training
rows from the solver's own runs, training on the GPU, and an A/B evaluation.

The reference solves each fault node's slip rate by Newton's method started
from the previous slip rate (src/faulting.f90:459-516).  Here a small network
(6 -> 64 -> 64 -> 64 -> 1, GELU; kernels in eqdyna_amd/csrc/eqd_mlp.hip)
predicts r = ln(V / V_prev) from the node's trial shear traction, effective
normal stress, state variable, a, b and ln V_prev.  Where it predicts a change
|r| >= gate, Newton starts from V_prev * exp(r) and gets k_fallback
iterations before it restarts from the plain guess; Newton's convergence
test decides the slip rate either way.

Training data: every `every` steps of seeded variants of the rough-fault
rate-and-state workload (case.c3) the solver records, per node, the
six inputs and the converged ln V (eqd_harvest).

    python -m eqdyna_amd.learned --out eqdyna_amd/rsf_guess_c3.npz
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from typing import Dict, List, Tuple

import numpy as np

from . import case as cases
from . import mesh
from .solver import MLP_NPARAMS, EqdError, Solver, mlp_train

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "rsf_guess_c3.npz")
NIN, NH = 6, 64
SHAPES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH,), (1,)]


def log(msg: str) -> None:
    print(f"[learned] {msg}", file=sys.stderr, flush=True)


def init_params(X: np.ndarray, r: np.ndarray, seed: int) -> np.ndarray:
    """Initial parameter vector: Glorot-uniform weights, zero biases, input
    normalisation (mean, 1/std) and output normalisation (mean, std) of r."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, shape in enumerate(SHAPES):
        if k % 2 == 1:
            parts.append(np.zeros(shape))
        else:
            fan_in = shape[1] if len(shape) == 2 else shape[0]
            fan_out = shape[0] if len(shape) == 2 else 1
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, size=shape))
    X = np.asarray(X, dtype=np.float64)
    sd = X.std(axis=0)
    sd = np.where(sd > 0, sd, 1.0)
    r = np.asarray(r, dtype=np.float64)
    parts += [X.mean(axis=0), 1.0 / sd, np.array([r.mean()]), np.array([max(r.std(), 1e-6)])]
    p = np.concatenate([np.ravel(x) for x in parts]).astype(np.float32)
    assert p.size == MLP_NPARAMS
    return p


def training_case(seed: int, steps: int) -> cases.Case:
    """A seeded variant of case.c3: its own fractal surface and a field, a
    uniform initial shear in [9.5, 10.5] MPa and a hypocentre along strike in
    [-3, 3] km."""
    c = cases.c3(steps=steps, seed=seed)
    rng = np.random.default_rng(10_000 + seed)
    tau0 = float(rng.uniform(9.5e6, 10.5e6))
    c.xsource = float(np.round(rng.uniform(-3.0e3, 3.0e3), -2))
    o = c.ofv
    for ix in range(c.nfx):
        for iz in range(c.nfz):
            o[8, iz, ix] = tau0
            o[20, iz, ix] = cases.state_steady_state(o[9, iz, ix], o[10, iz, ix], o[11, iz, ix], o[12, iz, ix],
                                                     o[13, iz, ix], tau0, o[7, iz, ix], o[46, iz, ix], 4)
    c.name = f"c3_seed{seed}"
    return c


def harvest(case_list: List[cases.Case], every: int, precision: str = "fp32",
            want: int | None = None) -> Tuple[np.ndarray, np.ndarray]:
    """Run each case with the plain guess and collect (inputs, converged ln V).
    A case whose run ends in a NaN velocity (the reference's explicit
    normal-stress state can diverge at high slip rate, see case.c3) is
    dropped; with want, cases are used until that many have run through."""
    Xs, ys = [], []
    used = 0
    for c in case_list:
        if want is not None and used >= want:
            break
        t = time.perf_counter()
        b, sc = mesh.build_case(c)
        s = Solver(b, sc, precision)
        nstep = int(sc["nstep"])
        s.harvest(every, int(sc["nftnd"]) * (nstep // every))
        try:
            s.run(nstep)
        except EqdError as e:
            log(f"{c.name}: dropped ({e})")
            s.close()
            continue
        rows = s.harvest_get()
        s.close()
        used += 1
        ok = np.all(np.isfinite(rows), axis=1)
        Xs.append(rows[ok, :NIN])
        ys.append(rows[ok, NIN])
        log(f"{c.name}: {ok.sum()} rows in {time.perf_counter() - t:.1f} s")
    return np.concatenate(Xs), np.concatenate(ys)


def train(X: np.ndarray, lnv: np.ndarray, epochs: int, batch: int, lr: float, seed: int):
    """Targets r = ln V - ln V_prev; returns (parameters, per-epoch loss)."""
    r = lnv - X[:, 5]
    p0 = init_params(X, r, seed)
    return mlp_train(p0, X, r, epochs=epochs, batch=batch, lr=lr, seed=seed)


def save(path: str, params: np.ndarray, meta: Dict) -> None:
    np.savez(path, params=np.asarray(params, dtype=np.float32), meta=np.array(json.dumps(meta)))


def load(path: str = DEFAULT_WEIGHTS) -> Tuple[np.ndarray, Dict]:
    z = np.load(path, allow_pickle=False)
    p = z["params"]
    if p.size != MLP_NPARAMS:
        raise ValueError(f"{path}: {p.size} parameters, expected {MLP_NPARAMS}")
    return p, json.loads(str(z["meta"]))


def run_case(b, sc, precision: str, params=None, k_fallback: int = 0, gate: float = 0.5, timing: bool = False):
    """One run; returns (fnft/fric outputs, Newton counters, wall seconds, fault kernel ms).
    k_fallback 0 with params None is the reference's solve; k_fallback > 0 the
    faster mode's iteration policy from the plain guess (params None) or from
    the network's proposal."""
    s = Solver(b, sc, precision, timing=timing)
    if params is not None or k_fallback > 0:
        s.set_guess(params, k_fallback, gate)
    nstep = int(sc["nstep"])
    s.run(10)
    t = time.perf_counter()
    s.run(nstep - 10)
    wall = time.perf_counter() - t
    fault_ms = s.kernel_time("fault")[0] if timing else float("nan")
    out = {"fnft": s.get("fnft"), "fric": s.get("fric")}
    stats = s.newton_stats()
    s.close()
    return out, stats, wall, fault_ms


def errors(out, ref, dt) -> Dict[str, float]:
    """Rupture-time error (in dt) over nodes ruptured in both, peak and final slip errors."""
    a, b = out["fnft"], ref["fnft"]
    both = (a < 5000) & (b < 5000)
    d = np.abs(a[both] - b[both]) / dt if both.any() else np.zeros(1)
    pk = float(np.max(np.abs(out["fric"][:, 75] - ref["fric"][:, 75])) / max(np.max(ref["fric"][:, 75]), 1e-300))
    sl_a = np.hypot(out["fric"][:, 70], out["fric"][:, 71])
    sl_b = np.hypot(ref["fric"][:, 70], ref["fric"][:, 71])
    sl = float(np.max(np.abs(sl_a - sl_b)) / max(np.max(sl_b), 1e-300))
    return {"rupture_time_mean_dt": float(d.mean()), "rupture_time_max_dt": float(d.max()),
            "ruptured": int((b < 5000).sum()), "rupture_flips": int(((a < 5000) != (b < 5000)).sum()),
            "peak_slip_rate_rel": pk, "final_slip_rel": sl}


def ab_test(c: cases.Case, params, k_fallback: int = 6, gate: float = 0.5) -> Dict:
    """Four arms on one case: FP64 with the reference's solve (the oracle run
    the errors are measured against), and FP32 bulk with the reference's
    solve, with the iteration policy from the plain guess, and with the
    policy from the network's proposal."""
    b, sc = mesh.build_case(c)
    nn = int(sc["totalNumOfNodes"])
    steps = int(sc["nstep"]) - 10
    res = {}
    ref, st64, w64, f64 = run_case(b, sc, "fp64", timing=True)
    res["fp64_reference"] = {"newton": st64, "wall_s": w64, "fault_ms": f64, "mdof": 3 * nn * steps / w64 / 1e6}
    for tag, p, k in (("fp32_reference", None, 0), ("fp32_policy_plain", None, k_fallback),
                      ("fp32_policy_learned", params, k_fallback)):
        out, st, w, f = run_case(b, sc, "fp32", p, k, gate, timing=True)
        res[tag] = {"newton": st, "wall_s": w, "fault_ms": f, "mdof": 3 * nn * steps / w / 1e6,
                    "errors_vs_fp64": errors(out, ref, sc["dt"])}
    for v in res.values():
        n = v["newton"]
        v["iterations_per_solve"] = n["iterations"] / max(n["solves"], 1)
        v["fallback_rate"] = n["fallbacks"] / max(n["solves"], 1)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seeds", type=int, default=32, help="training variants (seeds 100, 101, ...)")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--every", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-steps", type=int, default=3000)
    ap.add_argument("--gate", type=float, default=0.5)
    ap.add_argument("--k-fallback", type=int, default=6)
    a = ap.parse_args(argv)
    t = time.perf_counter()
    X, lnv = harvest([training_case(100 + k, a.steps) for k in range(3 * a.seeds)], a.every, want=a.seeds)
    log(f"harvested {len(X)} rows in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    p, losses = train(X, lnv, a.epochs, a.batch, a.lr, seed=7)
    log(f"trained {a.epochs} epochs in {time.perf_counter() - t:.1f} s; loss per epoch {np.round(losses, 5).tolist()}")
    meta = {"rows": int(len(X)), "cases": a.seeds, "first_seed": 100, "steps": a.steps, "every": a.every,
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "loss": [float(x) for x in losses]}
    ab = ab_test(cases.c3(steps=a.eval_steps), p, a.k_fallback, a.gate)
    meta["ab_test_c3_seed3"] = ab
    save(a.out, p, meta)
    print(json.dumps(meta["ab_test_c3_seed3"], indent=1))
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
