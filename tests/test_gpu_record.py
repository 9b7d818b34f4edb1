"""Device-side ground-motion / source-evolution records (eqd_record) equal
the state the reference's output_gm and output_src_evol write
(src/library_output.f90:267-312) at every step with mod(nt, 10) == 1: the
records of a run drained in chunks are compared bit for bit with velArr and
fric(47) copied back at exactly those steps, the records drained while the
next chunk runs (eqd_run_async) bit for bit with those, and the records
against the C oracle's velArr and fric(47) after exactly that many steps
(FP64 bar 1e-9 of each channel's largest value)."""
import numpy as np
import pytest

import oracle
from eqdyna_amd.solver import EqdError, Solver
from fixtures import load_golden

TOL_FP64 = 1e-9


def surface_nodes(bundle, count=24):
    xyz = np.asarray(bundle["meshCoor"]).reshape(-1, 3)
    ids = np.nonzero(np.abs(xyz[:, 2]) < 1e-6)[0]
    return (ids[:: max(1, len(ids) // count)] + 1).astype(np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_rsf4", "box_sw"])
def test_records_equal_state_at_record_steps(name):
    bundle, sc, _ = load_golden(name)
    nstep = int(sc["nstep"])
    nodes = surface_nodes(bundle)
    # reference: stop at every record step and copy the arrays back
    s = Solver(bundle, sc)
    want_v, want_s = [], []
    done = 0
    while done < nstep:
        step = done + 1  # the next record step is the next nt with (nt - 1) % 10 == 0
        nxt = step + (10 - (step - 1) % 10) % 10
        if nxt > nstep:
            s.run(nstep - done)
            break
        s.run(nxt - done)
        done = nxt
        want_v.append(s.get("velArr")[nodes - 1])
        want_s.append(s.get("fric")[:, 46].copy())
    s.close()
    # records, drained after chunks of 25 steps
    s = Solver(bundle, sc)
    s.record(nodes, every=10, capacity=4)
    got_v, got_s = [], []
    done = 0
    while done < nstep:
        n = min(25, nstep - done)
        s.run(n)
        done += n
        v, r = s.record_get(8)
        got_v += list(v)
        got_s += list(r)
    s.close()
    assert len(got_v) == len(want_v) == (nstep - 1) // 10 + 1
    assert np.array_equal(np.array(got_v), np.array(want_v))
    assert np.array_equal(np.array(got_s), np.array(want_s))


@pytest.mark.gpu
def test_record_overflow_is_reported():
    bundle, sc, _ = load_golden("box_sw")
    s = Solver(bundle, sc, nstep=60)
    s.record(surface_nodes(bundle), every=10, capacity=2)
    s.run(35)  # record steps 1, 11, 21, 31: two more than the capacity
    with pytest.raises(EqdError) as e:
        s.record_get(8)
    assert "dropped" in str(e.value)
    s.run(10)  # the buffer started over: step 41
    v, r = s.record_get(8)
    assert len(v) == 1
    s.close()


def _sync_records(bundle, sc, nodes, chunk=25):
    s = Solver(bundle, sc)
    s.record(nodes, every=10, capacity=4)
    got_v, got_s = [], []
    done, nstep = 0, int(sc["nstep"])
    while done < nstep:
        n = min(chunk, nstep - done)
        s.run(n)
        done += n
        v, r = s.record_get(8)
        got_v += list(v)
        got_s += list(r)
    s.close()
    return np.array(got_v), np.array(got_s)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_rsf4", "box_sw"])
def test_async_drain_equals_sync_drain(name):
    # the host drains chunk k while chunk k+1 is in flight: every drain hands
    # over exactly the records of the chunks that have ended, in step order
    bundle, sc, _ = load_golden(name)
    nstep = int(sc["nstep"])
    nodes = surface_nodes(bundle)
    want_v, want_s = _sync_records(bundle, sc, nodes)
    s = Solver(bundle, sc)
    s.record(nodes, every=10, capacity=8)  # two 25-step chunks hold at most 6 records
    chunks = [25] * (nstep // 25) + ([nstep % 25] if nstep % 25 else [])
    got_v, got_s = [], []
    done = 0
    s.run_async(chunks[0])
    for k in range(len(chunks)):
        s.wait()
        done += chunks[k]
        assert s.clock()[0] == done
        if k + 1 < len(chunks):
            s.run_async(chunks[k + 1])
        v, r = s.record_get(8)
        # the records of steps 1..done, none of the chunk in flight
        assert len(got_v) + len(v) == (done - 1) // 10 + 1, (k, len(got_v), len(v))
        got_v += list(v)
        got_s += list(r)
    s.close()
    assert np.array_equal(np.array(got_v), want_v)
    assert np.array_equal(np.array(got_s), want_s)


@pytest.mark.gpu
def test_other_calls_wait_for_the_run_in_flight():
    bundle, sc, _ = load_golden("box_rsf4")
    s = Solver(bundle, sc)
    s.run_async(20)
    # eqd_get waits for the run in flight: the state is that after 20 steps
    v_async = s.get("velArr")
    assert s.clock()[0] == 20
    s.close()
    s = Solver(bundle, sc)
    s.run(20)
    assert np.array_equal(s.get("velArr"), v_async)
    s.close()


@pytest.mark.gpu
def test_async_run_reports_nan_at_wait():
    bundle, sc, _ = load_golden("box_sw")
    b = dict(bundle)
    v1 = b["v1"].copy()
    v1[100] = np.nan
    b["v1"] = v1
    s = Solver(b, sc)
    s.run_async(2)  # the NaN is found when the run is waited for
    with pytest.raises(EqdError) as e:
        s.wait()
    assert e.value.code == -2
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_rsf4", "box_sw"])
def test_records_match_oracle_at_record_steps(name):
    # record r is the state after step 10 r + 1: the C oracle run for exactly
    # that many steps gives the reference's velArr and fric(47) there
    bundle, sc, _ = load_golden(name)
    nstep = int(sc["nstep"])
    nodes = surface_nodes(bundle)
    got_v, got_s = _sync_records(bundle, sc, nodes)
    steps = list(range(1, nstep + 1, 10))
    assert len(got_v) == len(steps)
    ref_v, ref_s = [], []
    for k in steps:
        o = oracle.run_oracle(bundle, sc, nsteps=k)
        ref_v.append(np.asarray(o["velArr"]).reshape(-1, 3)[nodes - 1])
        ref_s.append(np.asarray(o["fric"]).reshape(-1, 100)[:, 46])
    ref_v, ref_s = np.array(ref_v), np.array(ref_s)
    ev = np.abs(got_v - ref_v).max() / max(np.abs(ref_v).max(), 1e-300)
    es = np.abs(got_s - ref_s).max() / max(np.abs(ref_s).max(), 1e-300)
    print(f"[{name}] {len(steps)} records vs oracle: velocity {ev:.2e}, slip rate {es:.2e}", flush=True)
    assert ev <= TOL_FP64 and es <= TOL_FP64
