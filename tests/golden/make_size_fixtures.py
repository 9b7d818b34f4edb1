"""Write full-size golden fixtures under tests/golden/.

A full-size case (c1: 120 x 60 x 60 = 432,000 hexahedra, 600 steps; c3:
240 x 120 x 120 = 3,456,000 hexahedra) is run by the reference EQdyna time
step (oracle/_ref/eqd_ref, compiled from the reference's own Fortran sources
by oracle/ref.mk) on ONE rank, so its summation order is the single-domain
one.  The fixture keeps what fits in a few MB:

* the outputs the reference writes after the loop: the on-fault station
  histories (onFaultQuantHistSCECForm), the off-fault station series
  (OffFaultStGramSCEC), the fric slots the time step writes, fnft, and every
  ``stride``-th node of velArr and dispArr;
* a SHA-256 of every input array the reference hands to its time loop, so a
  test can rebuild the inputs with the project's case builder and prove that
  they are the reference's, bit for bit, before comparing outputs.

    python tests/golden/make_size_fixtures.py c1 [c3]

(needs /root/reference and a built oracle/_ref/eqd_ref; c1 takes about two
minutes on one core, c3 about an hour).
"""
import hashlib
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from eqdyna_amd import case  # noqa: E402
import refrun  # noqa: E402
from fixtures import FRIC_SLOTS  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NODE_STRIDE = 61


def digest(a) -> str:
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.dtype.str.encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def make(name: str) -> str:
    c = getattr(case, name)()
    t = time.time()
    with tempfile.TemporaryDirectory(dir=os.environ.get("EQD_FIXTURE_TMP")) as d:
        b, s, o, t_loop = refrun.run_reference(c, d, 1, True, timeout=4 * 3600)
    arrays = {
        "out_onFaultQuantHistSCECForm": o["onFaultQuantHistSCECForm"],
        "out_fric_slots": np.array(FRIC_SLOTS, dtype=np.int32),
        "out_fric": np.ascontiguousarray(o["fric"][:, [k - 1 for k in FRIC_SLOTS]]),
        "out_fnft": o["fnft"],
        "out_node_stride": np.array(NODE_STRIDE),
        "out_velArr": np.ascontiguousarray(o["velArr"][::NODE_STRIDE]),
        "out_dispArr": np.ascontiguousarray(o["dispArr"][::NODE_STRIDE]),
    }
    if "OffFaultStGramSCEC" in o:
        arrays["out_OffFaultStGramSCEC"] = o["OffFaultStGramSCEC"]
    keys = sorted(b)
    arrays["in_names"] = np.array(keys)
    arrays["in_sha256"] = np.array([digest(b[k]) for k in keys])
    sk = sorted(s)
    arrays["scalar_names"] = np.array(sk)
    arrays["scalar_values"] = np.array([s[k] for k in sk])
    arrays["t_loop_s"] = np.array(t_loop)
    p = os.path.join(HERE, f"size_{name}.npz")
    np.savez_compressed(p, **arrays)
    print(f"{p}: {os.path.getsize(p)} bytes; reference loop {t_loop:.1f} s, total {time.time() - t:.1f} s", flush=True)
    return p


if __name__ == "__main__":
    for n in sys.argv[1:] or ["c1"]:
        make(n)
