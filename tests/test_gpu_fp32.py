"""FP32 bulk mode (element force + hourglass in FP32; fault traction, Newton
solve, state and nodal update in FP64) against the reference's FP64 run.

Error bars of the FP32 mode: rupture time within 0.5*dt on average over the
ruptured nodes, peak slip rate within 1 % of the FP64 value."""
import numpy as np
import pytest

from eqdyna_amd.solver import Solver
from fixtures import channel_errors, channels, load_golden, undefined_slots

ELASTIC = ["box_sw", "box_tw", "box_layered", "box_rsf3", "box_rsf4", "box_q", "box_hg2", "box_dip", "box_rough", "box_c5sw", "box_c5rsf",
           "box_tp"]


def fp32_errors(out, ref, dt):
    rt_g, rt_r = out["fnft"], ref["fnft"]
    both = (rt_g < 5000) & (rt_r < 5000)
    flips = int(((rt_g < 5000) != (rt_r < 5000)).sum())
    drt = np.abs(rt_g[both] - rt_r[both]) / dt if both.any() else np.zeros(1)
    pk_g, pk_r = out["fric"][:, 75], ref["fric"][:, 75]
    peak_rel = float(np.max(np.abs(pk_g - pk_r)) / max(np.max(np.abs(pk_r)), 1e-300))
    return {"rupture_time_mean_dt": float(drt.mean()), "rupture_time_max_dt": float(drt.max()),
            "rupture_flips": flips, "peak_slip_rate_rel": peak_rel}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ELASTIC)
def test_fp32_error_vs_fp64_reference(name):
    print(f"case {name}", flush=True)
    bundle, sc, ref = load_golden(name)
    s = Solver(bundle, sc, "fp32")
    s.run(int(sc["nstep"]))
    out = s.outputs()
    s.close()
    e = fp32_errors(out, ref, sc["dt"])
    skip = undefined_slots(sc)
    ch = channel_errors(channels(out, skip), channels(ref, skip))
    worst = max(ch, key=ch.get)
    print(f"[{name}] fp32 vs fp64 reference: {e}; worst channel {worst} {ch[worst]:.2e}", flush=True)
    assert e["rupture_time_mean_dt"] <= 0.5
    assert e["peak_slip_rate_rel"] <= 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_rough", "box_rsf4"])
def test_fp32_learned_mode_vs_fp64_reference(name):
    # the headline fast mode exactly as the benchmark runs it: FP32 bulk plus
    # the shipped network (eqdyna_amd/rsf_guess_c3.npz) with the faster
    # iteration policy, k = 6 iterations before the plain-guess fallback,
    # gate 0.5; judged against the reference's FP64 run of the same inputs
    from eqdyna_amd import learned
    params, _ = learned.load()
    print(f"case {name}", flush=True)
    bundle, sc, ref = load_golden(name)
    s = Solver(bundle, sc, "fp32")
    s.set_guess(params, k_fallback=6, gate=0.5)
    s.run(int(sc["nstep"]))
    out = s.outputs()
    st = s.newton_stats()
    s.close()
    e = fp32_errors(out, ref, sc["dt"])
    ch = channel_errors(channels(out), channels(ref))
    worst = max(ch, key=ch.get)
    print(f"[{name}] fp32+learned vs fp64 reference: {e}; worst channel {worst} {ch[worst]:.2e}; "
          f"newton {st}", flush=True)
    assert st["solves"] == int(sc["nftnd"]) * int(sc["nstep"])
    assert e["rupture_flips"] == 0
    assert e["rupture_time_mean_dt"] <= 0.5
    assert e["peak_slip_rate_rel"] <= 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_rough", "box_rsf4"])
def test_geometry_operators_vs_stored_operators(name, monkeypatch):
    # FP32 bulk mode on a mesh whose elements all differ rebuilds each
    # element's operators from its corner coordinates in the kernels; with
    # EQD_GEOM_OPS=0 it reads the stored (FP64-computed, FP32-rounded) rows.
    # Both are judged against the reference's FP64 run with the same bars.
    print(f"case {name}", flush=True)
    bundle, sc, ref = load_golden(name)
    if name == "box_rsf4":
        monkeypatch.setenv("EQD_NO_OPERATOR_CLASSES", "1")  # a planar mesh: force per-element operators
    res = {}
    for geo in ("1", "0"):
        monkeypatch.setenv("EQD_GEOM_OPS", geo)
        s = Solver(bundle, sc, "fp32")
        info = s.layout_info()
        s.run(int(sc["nstep"]))
        out = s.outputs()
        s.close()
        assert info["geom_ops_regular"] == int(geo) and info["geom_ops_pml"] == int(geo), info
        e = fp32_errors(out, ref, sc["dt"])
        ch = channel_errors(channels(out), channels(ref))
        worst = max(ch, key=ch.get)
        print(f"[{name}] geometry rows {geo}: {e}; worst channel {worst} {ch[worst]:.2e}", flush=True)
        assert e["rupture_flips"] == 0 and e["rupture_time_mean_dt"] <= 0.5 and e["peak_slip_rate_rel"] <= 0.01
        res[geo] = out
    d = channel_errors(channels(res["1"]), channels(res["0"]))
    w = max(d, key=d.get)
    print(f"[{name}] geometry rows against stored rows: worst channel {w} {d[w]:.2e}", flush=True)
