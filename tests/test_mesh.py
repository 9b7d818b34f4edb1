"""The case builder (eqdyna_amd/csrc/eqd_mesh.cpp) against the reference's own
mesh generation: every golden fixture holds the arrays the reference host
hands to its time loop (after src/meshgen.f90 and src/assembleGlobalMass.f90),
and building the same case here must reproduce them bit for bit."""
import os
import sys

import numpy as np
import pytest

from fixtures import CASES, load_golden

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_fixtures  # noqa: E402

from eqdyna_amd import mesh  # noqa: E402

PRESETS = {c.name: c for c in make_fixtures.cases()}

# scalars the time step reads (the rest are reference bookkeeping)
SCALARS = ["dt", "nstep", "totalNumOfNodes", "totalNumOfElements", "totalNumOfEquations", "sizeOfEqNumIndexArr",
           "sizeOfStressDofIndexArr", "nftnd", "C_elastic", "C_Q", "C_hg", "C_nuclea", "friclaw", "TPV",
           "insertFaultType", "nPML", "R", "vmaxPML", "rdampk", "rdampm", "roumax", "rhow", "gamar", "ccosphi",
           "sinphi", "tv", "slipRateThres", "xsource", "ysource", "zsource", "nucR", "nucRuptVel", "nucdtau0",
           "nucT", "kapa_hg", "numOfOnFaultStCount", "numOfOffFaultStCount"] + [f"PMLb{i}" for i in range(1, 9)]


@pytest.mark.parametrize("name", CASES)
def test_builder_bitwise_vs_reference_meshgen(name):
    bundle, sc, _ = load_golden(name)
    b, s = mesh.build_case(PRESETS[name], nthreads=4)
    for k in SCALARS:
        assert s[k] == sc[k], f"{name}: scalar {k}: {s[k]!r} != {sc[k]!r}"
    for k, ref in bundle.items():
        assert k in b, f"{name}: builder does not produce {k}"
        got = np.asarray(b[k])
        assert got.shape == ref.shape, f"{name}: {k} shape {got.shape} != {ref.shape}"
        assert got.dtype == ref.dtype and np.array_equal(got, ref), f"{name}: {k} differs from the reference meshgen"


@pytest.mark.parametrize("name", ["c1", "c3"])
def test_builder_bitwise_vs_reference_full_size(name):
    # c1 (432,000 hexahedra) and c3 (3,456,000, rough fault) at full size:
    # every input array the reference's one-rank run handed to its time
    # loop, by SHA-256
    import test_gpu_size as T
    ref, sha = T.load_size(name)
    T.build_inputs(name, sha)
