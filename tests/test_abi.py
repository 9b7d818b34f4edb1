"""The C-ABI library loads and exports every symbol include/eqdyna_amd.h
declares (no compute: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from eqdyna_amd import solver

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "eqdyna_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?\s*(eqd_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_api():
    names = declared_functions()
    assert set(names) == set(solver.EXPORTS), names


def test_library_exports_all_symbols():
    if not os.path.exists(solver.LIB_PATH):
        pytest.fail("libeqd.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(solver.LIB_PATH)
    for n in declared_functions():
        assert hasattr(lib, n), n


def test_create_rejects_bad_scalars():
    lib = solver.lib()
    s = solver.Scalars()
    h = ctypes.c_void_p()
    assert lib.eqd_create(ctypes.byref(s), ctypes.byref(h)) != 0  # nn = ne = 0


@pytest.mark.parametrize("kind", ["shared", "same"])
def test_commit_rejects_bad_split_node_pairs(kind):
    # all pairs update their two nodes at once, so a node in two pairs, or a
    # pair of one node, is rejected before anything touches the device
    from fixtures import load_golden
    bundle, sc, _ = load_golden("box_sw")
    b = dict(bundle)
    nsmp = b["nsmp"].copy()  # (nf, 2): slave, master
    if kind == "shared":
        nsmp[3, 0] = nsmp[2, 0]
        msg = "more than one split-node pair"
    else:
        nsmp[3, 1] = nsmp[3, 0]
        msg = "same node"
    b["nsmp"] = nsmp
    with pytest.raises(solver.EqdError) as e:
        solver.Solver(b, sc)
    assert e.value.code == -1 and msg in str(e.value), str(e.value)
