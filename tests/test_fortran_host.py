"""The reference's own Fortran host (meshgen, assembleGlobalMass, output
writers) linked with the GPU time step through the iso_c_binding shim
(eqdyna_amd/fortran/eqd_driver.f90), against the reference build on the same
TPV8 case (case_input/test.tpv8), one rank each.  frt.txt is the reference's
E18.7 text output (src/library_output.f90:160-205)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import refrun
from eqdyna_amd import case
from fixtures import hypocentre_blocks, printed_unit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_HOST = os.path.join(ROOT, "eqdyna_amd", "fortran", "_build", "eqdyna_gpu")


@pytest.mark.gpu
def test_reference_host_with_gpu_driver_matches_reference():
    if not os.path.exists(GPU_HOST) or not refrun.available():
        pytest.skip("the Fortran host build needs the reference sources (built where they are present)")
    c = case.tpv8()
    c.nx = c.ny = c.nz = 1
    with tempfile.TemporaryDirectory() as dg, tempfile.TemporaryDirectory() as dr:
        c.write_inputs(dg)
        print("running the Fortran host with the GPU driver on TPV8", flush=True)
        p = subprocess.run([GPU_HOST], cwd=dg, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0 and "EQD_GPU" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
        refrun.run_reference(c, dr, 1, dump=False, timeout=600)
        g = np.loadtxt(os.path.join(dg, "frt.txt0"))
        r = np.loadtxt(os.path.join(dr, "frt.txt0"))
        gs = sorted(f for f in os.listdir(dg) if f.startswith("faultst"))
    # the 'Hypocenter dynamics' blocks (showSourceDynamics,
    # src/faulting.f90:343-365): one per step, between the same progress
    # lines; each number within one unit of its 7th printed digit or 2e-7 of
    # the largest value of its quantity: coordinates, time, stresses (n, s, d
    # traction, state_normal, nucleation), slips (n, s, d, cumulative; the
    # normal slip is rounding noise next to the others), slip rates, slip
    # weakening parameters, rsf a and b, state
    hg, pg = hypocentre_blocks(p.stdout)
    hr, pr = hypocentre_blocks(refrun.LAST_STDOUT[0])
    assert hr.shape[0] == c.nstep and hg.shape == hr.shape and pg == pr
    assert np.array_equal(hg[:, :3], hr[:, :3])
    hscale = np.zeros(hr.shape[1])
    for grp in [[0, 1, 2], [3], [4, 5, 6, 7, 21], [8, 9, 10, 14], [11, 12, 13], [15, 16, 17], [18, 19], [20]]:
        hscale[grp] = max(np.abs(hr[:, grp]).max(), 1e-300)
    herr = np.abs(hg - hr)
    assert np.all(herr <= np.maximum(printed_unit(hr) * 1.0000001, 2e-7 * hscale)), \
        np.array2string((herr / hscale).max(axis=0), precision=2)
    assert g.shape == r.shape
    # rupture times (column 4) identical; every other column within 2e-6 (one
    # unit of the 7th printed digit) of the largest value of its quantity:
    # coordinates 1-3, slips 5-7, slip rates 8-11, tractions 12-14, velocities
    # 15-20, state 21, normal-stress state 22 (fault opening, column 7, is
    # rounding noise next to the tangential slip)
    assert np.array_equal(g[:, 3], r[:, 3])
    groups = [[0, 1, 2], [3], [4, 5, 6], [7, 8, 9, 10], [11, 12, 13], [14, 15, 16, 17, 18, 19], [20], [21]]
    scale = np.zeros(r.shape[1])
    for grp in groups:
        scale[grp] = max(np.abs(r[:, grp]).max(), 1e-300)
    err = (np.abs(g - r) / scale).max(axis=0)
    print("per-column error", np.array2string(err, precision=2), flush=True)
    assert err.max() <= 2e-6
    # the reference's station writer ran on the GPU results (one file per station)
    assert len(gs) == len(c.st_on_fault)


@pytest.mark.gpu
def test_ground_motion_output_matches_library_records():
    # outputGroundMotion = 1: the library records the surface velocities and
    # fric(47) every 10 steps on the device and the host writes them in the
    # format of output_gm / output_src_evol (src/library_output.f90:267-312;
    # the reference's own routines cannot run under this compiler: they
    # re-open a connected unit).  The files must hold, bit for bit, what the
    # same GPU step records through the Python front end on the same mesh.
    if not os.path.exists(GPU_HOST):
        pytest.skip("the Fortran host build needs the reference sources (built where they are present)")
    from eqdyna_amd import mesh
    from eqdyna_amd.solver import Solver
    c = case.tpv8()
    c.nx = c.ny = c.nz = 1
    c.outputGroundMotion = 1
    with tempfile.TemporaryDirectory() as dg:
        c.write_inputs(dg)
        print("running the Fortran host with the GPU driver on TPV8 with ground-motion output", flush=True)
        p = subprocess.run([GPU_HOST], cwd=dg, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0 and "EQD_GPU" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
        gm = np.fromfile(os.path.join(dg, "gm0"), dtype="<f8")
        se = np.fromfile(os.path.join(dg, "src_evol0"), dtype="<f8")
    bundle, sc = mesh.build_case(c)
    xyz = np.asarray(bundle["meshCoor"]).reshape(-1, 3)
    # surface nodes within 20 km of the fault's extent (find_surfaceNodeIdArr, src/library_output.f90:247-264)
    sel = ((xyz[:, 0] < c.fxmax + 20.0e3) & (xyz[:, 0] > c.fxmin - 20.0e3) & (xyz[:, 1] < c.fymax + 20.0e3) &
           (xyz[:, 1] > c.fymin - 20.0e3) & (np.abs(xyz[:, 2]) < c.dx / 1000))
    ids = (np.nonzero(sel)[0] + 1).astype(np.int32)
    nstep = int(sc["nstep"])
    s = Solver(bundle, sc)
    s.record(ids, every=10, capacity=nstep // 10 + 1)
    s.run(nstep)
    v, r = s.record_get(nstep // 10 + 1)
    s.close()
    print(f"{len(v)} records of {len(ids)} surface nodes and {r.shape[1]} pairs", flush=True)
    assert gm.size == v.size and se.size == r.size
    assert np.array_equal(gm, v.ravel()) and np.array_equal(se, r.ravel())


def _frt_rows(b, o):
    """The 22 columns of output_frt (src/library_output.f90:160-205) from the
    reference's FP64 state."""
    xyz = np.asarray(b["meshCoor"]).reshape(-1, 3)
    nsmp = np.asarray(b["nsmp"]).reshape(-1, 2)
    fr = np.asarray(o["fric"]).reshape(-1, 100)
    f = lambda k: fr[:, k - 1:k]  # noqa: E731
    cols = [xyz[nsmp[:, 0] - 1], np.asarray(o["fnft"]).reshape(-1, 1)] + [f(k) for k in range(71, 77)] + \
        [f(47), f(78), f(79), f(80)] + [f(k) for k in range(31, 37)] + [f(20), f(23)]
    return np.hstack(cols)


def _onfault_cols(h, friclaw):
    # columns of output_onfault_st (src/library_output.f90:16-96) from one
    # station's [nstep][12] history
    c = [h[:, 0], h[:, 4], h[:, 1], h[:, 7] / 1e6, -h[:, 5], -h[:, 2], -h[:, 8] / 1e6, -h[:, 9] / 1e6]
    if friclaw >= 3:
        c += [h[:, 3], h[:, 11], h[:, 10] / 1e6]
    return np.stack(c, axis=1)


@pytest.mark.gpu
def test_full_precision_outputs_match_reference_state():
    # EQD_FULL_PRECISION_OUTPUT=1: frt and the station files also in 17
    # significant digits; checked against the reference build's FP64 state on
    # the same TPV8 case (its binary dump, not its 7-digit text) at the FP64
    # bar: 1e-9 of each quantity's largest value, identical rupture times
    if not os.path.exists(GPU_HOST) or not refrun.available():
        pytest.skip("the Fortran host build needs the reference sources (built where they are present)")
    from fixtures import undefined_slots
    c = case.tpv8()
    c.nx = c.ny = c.nz = 1
    with tempfile.TemporaryDirectory() as dg, tempfile.TemporaryDirectory() as dr:
        c.write_inputs(dg)
        print("running the Fortran host with the GPU driver on TPV8, full-precision output", flush=True)
        env = dict(os.environ, EQD_FULL_PRECISION_OUTPUT="1")
        p = subprocess.run([GPU_HOST], cwd=dg, env=env, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0 and "EQD_GPU" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
        b, sc, o, _ = refrun.run_reference(c, dr, 1, dump=True, timeout=600)
        g = np.loadtxt(os.path.join(dg, "frt.full.txt0"))
        fulls = sorted(f for f in os.listdir(dg) if f.startswith("faultst") and f.endswith(".full.txt"))
        # the reference's writers ran in the same directory on the same state
        texts = sorted(f for f in os.listdir(dg) if f.startswith("faultst") and f.endswith(".txt")
                       and not f.endswith(".full.txt"))
        st = {f: np.loadtxt(os.path.join(dg, f)) for f in fulls}
        bodies = sorted(f for f in os.listdir(dg) if f.startswith("body") and f.endswith(".full.txt"))
        body_texts = sorted(f for f in os.listdir(dg) if f.startswith("body") and f.endswith(".txt")
                            and not f.endswith(".full.txt"))
        bd = {f: np.loadtxt(os.path.join(dg, f)) for f in bodies}
    # frt: the reference's columns, read back to the same doubles
    r = _frt_rows(b, o)
    assert g.shape == r.shape
    assert np.array_equal(g[:, 3], r[:, 3]), "rupture times differ"
    keep = [k for k in range(22) if not (k == 21 and 23 in undefined_slots(sc))]
    groups = [[0, 1, 2], [4, 5, 6], [7, 8, 9, 10], [11, 12, 13], [14, 15, 16, 17, 18, 19], [20], [21]]
    err = 0.0
    for grp in groups:
        grp = [k for k in grp if k in keep]
        if grp:
            err = max(err, np.abs(g[:, grp] - r[:, grp]).max() / max(np.abs(r[:, grp]).max(), 1e-300))
    print(f"frt.full vs reference state: {err:.2e}", flush=True)
    assert err <= 1e-9
    # on-fault stations: named as the reference's writer names them (its
    # files beside them), each equal to one of the reference's histories
    assert fulls == [t[:-4] + ".full.txt" for t in texts]
    hist = np.asarray(o["onFaultQuantHistSCECForm"]).reshape(-1, int(sc["nstep"]), 12)
    want = [_onfault_cols(h, int(sc["friclaw"])) for h in hist]
    scale = np.maximum(np.abs(np.stack(want)).max(axis=(0, 1)), 1e-300)
    used = set()
    worst = 0.0
    for f, a in st.items():
        e = [float((np.abs(a - w) / scale).max()) if a.shape == w.shape else np.inf for w in want]
        i = int(np.argmin(e))
        assert i not in used, f
        used.add(i)
        worst = max(worst, e[i])
    print(f"{len(st)} on-fault station files vs reference histories: {worst:.2e}", flush=True)
    assert worst <= 1e-9
    # off-fault stations: the reference's names and its history columns
    assert bodies == [t[:-4] + ".full.txt" for t in body_texts]
    G = np.asarray(o["OffFaultStGramSCEC"]).reshape(int(sc["nstep"]), -1)
    want = [np.stack([G[:, 0], G[:, 6 * i + 1], G[:, 6 * i + 2], -G[:, 6 * i + 5], -G[:, 6 * i + 6], G[:, 6 * i + 3],
                      G[:, 6 * i + 4]], axis=1) for i in range((G.shape[1] - 1) // 6)]
    scale = np.maximum(np.abs(np.stack(want)).max(axis=(0, 1)), 1e-300)
    worst = 0.0
    for f, a in bd.items():
        worst = max(worst, min(float((np.abs(a - w) / scale).max()) for w in want))
    print(f"{len(bd)} off-fault station files vs reference histories: {worst:.2e}", flush=True)
    assert len(bd) == len(want) and worst <= 1e-9
