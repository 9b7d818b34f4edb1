"""HIP path against the reference's own FP64 run and the C oracle.

FP64 bar (north_star): every channel within 1e-9 relative (infinity norm over
the series, scaled by the reference channel's maximum) and identical rupture
times.  The element forces are regrouped on the GPU (stiffness + hourglass per
element before the node sum), so agreement is to rounding, not bitwise."""
import numpy as np
import pytest

import oracle
from eqdyna_amd.solver import Solver
from fixtures import CASES, channel_errors, channels, load_golden

TOL_FP64 = 1e-9


def report(name, ch_gpu, ch_ref):
    err = channel_errors(ch_gpu, ch_ref)
    worst = max(err.values())
    print(f"\n[{name}] worst {worst:.3e}: " + " ".join(f"{k}={e:.2e}" for k, e in err.items()), flush=True)
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_fp64_matches_reference(name):
    print(f"case {name}", flush=True)
    bundle, sc, ref = load_golden(name)
    s = Solver(bundle, sc, "fp64")
    s.run(int(sc["nstep"]))
    out = s.outputs()
    s.close()
    cg, cr = channels(out), channels(ref)
    worst = report(name, cg, cr)
    assert np.array_equal(out["fnft"], ref["fnft"]), "rupture times differ"
    assert worst <= TOL_FP64


@pytest.mark.gpu
def test_split_runs_equal_one_run():
    # eqd_run(a) + eqd_run(b) continues the clock exactly like eqd_run(a+b)
    bundle, sc, _ = load_golden("box_rsf4")
    n = int(sc["nstep"])
    s1 = Solver(bundle, sc)
    s1.run(n)
    o1 = s1.outputs()
    s1.close()
    s2 = Solver(bundle, sc)
    s2.run(n // 3)
    s2.run(n - n // 3)
    o2 = s2.outputs()
    assert s2.clock()[0] == n
    s2.close()
    for k in o1:
        assert np.array_equal(o1[k], o2[k]), k


@pytest.mark.gpu
def test_gpu_matches_oracle_midway():
    # compare with the oracle at a step count the fixtures do not store
    bundle, sc, _ = load_golden("box_layered")
    k = 17
    s = Solver(bundle, sc)
    s.run(k)
    o = s.outputs()
    s.close()
    r = oracle.run_oracle(bundle, sc, nsteps=k)
    r = {key: v for key, v in r.items() if key in o}
    o["onFaultQuantHistSCECForm"] = o["onFaultQuantHistSCECForm"][:, :k]
    o["OffFaultStGramSCEC"] = o["OffFaultStGramSCEC"][:k]
    err = channel_errors(channels(o), channels(r))
    assert max(err.values()) <= TOL_FP64, err


@pytest.mark.gpu
def test_nan_is_reported():
    from eqdyna_amd.solver import EqdError
    bundle, sc, _ = load_golden("box_sw")
    b = dict(bundle)
    v1 = b["v1"].copy()
    v1[100] = np.nan
    b["v1"] = v1
    s = Solver(b, sc)
    with pytest.raises(EqdError) as e:
        s.run(2)
    assert e.value.code == -2
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_operator_classes_are_lossless(precision, monkeypatch):
    # elements with bit-identical operators share a table row; the result must be
    # bit-identical to per-element operator arrays
    bundle, sc, _ = load_golden("box_layered")
    n = int(sc["nstep"])
    s1 = Solver(bundle, sc, precision)
    info = s1.layout_info()
    s1.run(n)
    o1 = s1.outputs()
    s1.close()
    monkeypatch.setenv("EQD_NO_OPERATOR_CLASSES", "1")
    s2 = Solver(bundle, sc, precision)
    info2 = s2.layout_info()
    s2.run(n)
    o2 = s2.outputs()
    s2.close()
    print(f"\n[{precision}] classes {info} / forced off {info2}", flush=True)
    assert info["op_classes_regular"] > 0 and info2["op_classes_regular"] == 0
    for k in o1:
        assert np.array_equal(o1[k], o2[k]), k
