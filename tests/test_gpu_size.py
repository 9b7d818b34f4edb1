"""The HIP path at a benchmark configuration's full size against the
reference's own one-rank FP64 run (tests/golden/size_*.npz, written by
tests/golden/make_size_fixtures.py from oracle/_ref/eqd_ref).

c1: 120 x 60 x 60 = 432,000 hexahedra at 200 m (232,000 of them PML), 4,050
split-node pairs, slip weakening, 600 steps.  The inputs are rebuilt here by
the project's case builder and must hash to the reference's own inputs, bit
for bit, before any output is compared.  FP64 bar: every channel within
1e-9 of its quantity's largest value over the whole series, rupture times
identical.  FP32 bulk: rupture time within 0.5 dt on average, peak slip rate
within 1 %.

c3: 240 x 120 x 120 = 3,456,000 hexahedra at 100 m, the seeded rough fault
with 15,939 pairs, slip law with strong weakening, 3000 steps; run in FP64
and in the FP32 bulk mode with the learned Newton guess (the c4 mode)."""
import hashlib
import os

import numpy as np
import pytest

from fixtures import GOLDEN

TOL_FP64 = 1e-9
HIST = ["time", "sliprate_s", "sliprate_d", "state", "slip_s", "slip_d", "slip_n", "tau_s", "tau_d", "sigma_n",
        "pore_pressure", "temperature"]


def digest(a) -> str:
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.dtype.str.encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def load_size(name):
    z = np.load(os.path.join(GOLDEN, f"size_{name}.npz"))
    ref = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    sha = dict(zip([str(x) for x in z["in_names"]], [str(x) for x in z["in_sha256"]]))
    return ref, sha


def build_inputs(name, sha):
    from eqdyna_amd import case, mesh
    bundle, sc = mesh.build_case(getattr(case, name)(), nthreads=min(16, os.cpu_count() or 1))
    for k, h in sha.items():
        assert k in bundle, f"case builder does not produce {k}"
        assert digest(bundle[k]) == h, f"{name}: input {k} differs from the reference's"
    return bundle, sc


def series(out, ref):
    """Channels of a run in the fixture's form: on-fault histories, station
    series, the stored fric slots, rupture times, sampled node state."""
    slots = [int(k) for k in ref["fric_slots"]]
    stride = int(ref["node_stride"])
    ch = {f"fric{k}": out["fric"][:, k - 1] for k in slots}
    ch["fnft"] = np.where(out["fnft"] > 5000.0, 0.0, out["fnft"])
    for j, nm in enumerate(HIST):
        ch["onfault_" + nm] = out["onFaultQuantHistSCECForm"][:, :, j]
    ch["station_series"] = out["OffFaultStGramSCEC"]
    ch["velArr"] = out["velArr"][::stride]
    ch["dispArr"] = out["dispArr"][::stride]
    return ch


def ref_series(ref):
    ch = {f"fric{int(k)}": ref["fric"][:, i] for i, k in enumerate(ref["fric_slots"])}
    ch["fnft"] = np.where(ref["fnft"] > 5000.0, 0.0, ref["fnft"])
    for j, nm in enumerate(HIST):
        ch["onfault_" + nm] = ref["onFaultQuantHistSCECForm"][:, :, j]
    ch["station_series"] = ref["OffFaultStGramSCEC"]
    ch["velArr"] = ref["velArr"]
    ch["dispArr"] = ref["dispArr"]
    return ch


def errors(ch, ch_ref):
    from fixtures import channel_errors
    return channel_errors(ch, ch_ref)


def rupture_errors(out, ref, dt):
    """Rupture-time error in dt (mean, max over pairs ruptured in both runs),
    pairs ruptured in one run only, and the peak-slip-rate error (fric(76))
    relative to its largest value."""
    rt_g, rt_r = out["fnft"], ref["fnft"]
    ruptured = rt_r < 5000
    both = ruptured & (rt_g < 5000)
    drt = np.abs(rt_g[both] - rt_r[both]) / dt
    flips = int((ruptured != (rt_g < 5000)).sum())
    pk_i = [int(k) for k in ref["fric_slots"]].index(76)
    pk = float(np.max(np.abs(out["fric"][:, 75] - ref["fric"][:, pk_i])) / np.max(np.abs(ref["fric"][:, pk_i])))
    return int(ruptured.sum()), float(drt.mean()), float(drt.max()), flips, pk


def run_case(bundle, sc, precision, guess=False):
    from eqdyna_amd.solver import Solver
    s = Solver(bundle, sc, precision)
    if guess:
        from eqdyna_amd import learned
        params, _ = learned.load()
        s.set_guess(params, k_fallback=6, gate=0.5)
    s.run(int(sc["nstep"]))
    out = s.outputs()
    s.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c1_full_size_vs_reference(precision):
    print(f"c1 full size, {precision}", flush=True)
    ref, sha = load_size("c1")
    bundle, sc = build_inputs("c1", sha)
    out = run_case(bundle, sc, precision)
    ch, ch_ref = series(out, ref), ref_series(ref)
    err = errors(ch, ch_ref)
    worst = max(err, key=err.get)
    n_rup, drt_mean, drt_max, flips, pk = rupture_errors(out, ref, sc["dt"])
    print(f"[c1 {precision}] {n_rup} of {len(ref['fnft'])} pairs ruptured; worst channel {worst} "
          f"{err[worst]:.2e}; rupture time mean {drt_mean:.3f} dt max {drt_max:.3f} dt, flips {flips}; "
          f"peak slip rate {pk:.2e}", flush=True)
    assert n_rup > 1000  # the rupture crosses the fault within the 600 steps
    if precision == "fp64":
        assert err[worst] <= TOL_FP64, (worst, err[worst])
        assert np.array_equal(out["fnft"], ref["fnft"])
    else:
        assert flips == 0 and drt_mean <= 0.5
        assert pk <= 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "fp32_learned"])
def test_c3_full_size_vs_reference(mode):
    # c3 (and c4, the same case in FP32 bulk with the shipped network, k = 6,
    # gate 0.5) at full size: 3,456,000 hexahedra, 15,939 pairs on the rough
    # fault, slip law with strong weakening, 3000 steps, against the
    # reference's one-rank FP64 run of the same inputs.
    #
    # Rupture times and peak slip rates are the well-conditioned outputs.  The
    # final state (state variable, tractions) 3000 steps into the slip-law
    # rupture is not: moving ONE input by one ulp moves it by percents, so the
    # FP64 bar for those channels is ten times that measured sensitivity
    # (measured here with the FP64 kernels; measured 2.3e-2, against 1.5e-2
    # between this and the reference).
    print(f"c3 full size, {mode}", flush=True)
    ref, sha = load_size("c3")
    bundle, sc = build_inputs("c3", sha)
    dt = sc["dt"]
    out = run_case(bundle, sc, "fp64" if mode == "fp64" else "fp32", guess=mode != "fp64")
    ch_ref = ref_series(ref)
    err = errors(series(out, ref), ch_ref)
    top = sorted(err.items(), key=lambda kv: -kv[1])[:6]
    n_rup, drt_mean, drt_max, flips, pk = rupture_errors(out, ref, dt)
    print(f"[c3 {mode}] {n_rup} of {len(ref['fnft'])} pairs ruptured; rupture time mean {drt_mean:.3f} dt "
          f"max {drt_max:.3f} dt, flips {flips}; peak slip rate {pk:.2e}; largest channel errors "
          + " ".join(f"{k}={v:.2e}" for k, v in top), flush=True)
    assert n_rup == len(ref["fnft"])  # the whole fault ruptures within the 3000 steps
    if mode == "fp64":
        assert np.array_equal(out["fnft"], ref["fnft"]), "rupture times differ"
        assert pk <= 1e-4
        b = dict(bundle)
        fr = b["fric"].copy()
        j = int(np.random.default_rng(0).integers(0, fr.shape[0]))
        fr[j, 19] = np.nextafter(fr[j, 19], np.inf)
        b["fric"] = fr
        ch = series(out, ref)
        sens = errors(series(run_case(b, sc, "fp64"), ref), ch)
        bar = {k: max(TOL_FP64, 10.0 * sens[k]) for k in err}
        print(f"[c3 fp64] one-ulp sensitivity: " + " ".join(f"{k}={v:.2e}" for k, v in
                                                             sorted(sens.items(), key=lambda kv: -kv[1])[:6]),
              flush=True)
        bad = {k: (err[k], bar[k]) for k in err if err[k] > bar[k]}
        assert not bad, bad
    else:
        assert flips == 0 and drt_mean <= 0.5
        assert pk <= 0.01
