"""Device-side ground-motion / source-evolution records (eqd_record) equal
the state the reference's output_gm and output_src_evol write
(src/library_output.f90:267-312) at every step with mod(nt, 10) == 1: the
records of a run drained in chunks are compared bit for bit with velArr and
fric(47) copied back at exactly those steps."""
import numpy as np
import pytest

from eqdyna_amd.solver import EqdError, Solver
from fixtures import load_golden


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
