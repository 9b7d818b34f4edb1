"""The HIP path at a benchmark configuration's full size against the
reference's own one-rank FP64 run (tests/golden/size_*.npz, written by
tests/golden/make_size_fixtures.py from oracle/_ref/eqd_ref).

c1: 120 x 60 x 60 = 432,000 hexahedra at 200 m (232,000 of them PML), 4,050
split-node pairs, slip weakening, 600 steps.  The inputs are rebuilt here by
the project's case builder and must hash to the reference's own inputs, bit
for bit, before any output is compared.  FP64 bar: every channel within
1e-9 of its quantity's largest value over the whole series, rupture times
identical.  FP32 bulk: rupture time within 0.5 dt on average, peak slip rate
within 1 %."""
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


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c1_full_size_vs_reference(precision):
    from eqdyna_amd.solver import Solver
    print(f"c1 full size, {precision}", flush=True)
    ref, sha = load_size("c1")
    bundle, sc = build_inputs("c1", sha)
    nstep = int(sc["nstep"])
    s = Solver(bundle, sc, precision)
    s.run(nstep)
    out = s.outputs()
    s.close()
    ch, ch_ref = series(out, ref), ref_series(ref)
    err = errors(ch, ch_ref)
    worst = max(err, key=err.get)
    rt_g, rt_r = out["fnft"], ref["fnft"]
    ruptured = rt_r < 5000
    both = ruptured & (rt_g < 5000)
    drt = np.abs(rt_g[both] - rt_r[both]) / sc["dt"]
    flips = int((ruptured != (rt_g < 5000)).sum())
    pk_i = [int(k) for k in ref["fric_slots"]].index(76)
    pk = float(np.max(np.abs(out["fric"][:, 75] - ref["fric"][:, pk_i])) / np.max(np.abs(ref["fric"][:, pk_i])))
    print(f"[c1 {precision}] {int(ruptured.sum())} of {len(rt_r)} pairs ruptured; worst channel {worst} "
          f"{err[worst]:.2e}; rupture time mean {drt.mean():.3f} dt max {drt.max():.3f} dt, flips {flips}; "
          f"peak slip rate {pk:.2e}", flush=True)
    assert ruptured.sum() > 1000  # the rupture crosses the fault within the 600 steps
    if precision == "fp64":
        assert err[worst] <= TOL_FP64, (worst, err[worst])
        assert np.array_equal(rt_g, rt_r)
    else:
        assert flips == 0 and drt.mean() <= 0.5
        assert pk <= 0.01
