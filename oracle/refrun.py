"""Test infrastructure: run the reference EQdyna time step (oracle/_ref/eqd_ref,
built from the reference's own sources by oracle/ref.mk) on a Case and load its
FP64 input bundle and outputs.  Only tests and the CPU-baseline leg of bench.py
use this module."""
from __future__ import annotations

import os
import re
import subprocess

from eqdyna_amd.bundle import load_arrays, load_scalars

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "eqd_ref")
# timing-only build (-O3), the CPU baseline; never used for parity
REF_BIN_O3 = os.path.join(HERE, "_ref", "eqd_ref_o3")
MPIEXEC = "/opt/conda/bin/mpiexec"


def available() -> bool:
    return os.path.exists(REF_BIN)


LAST_PHASES = {}


LAST_STDOUT = [""]  # standard output of the last run (its printed diagnostics)


def run_reference(case, workdir: str, nranks: int = 1, dump: bool = True, timeout: int = 3600, binary: str = REF_BIN):
    """Write the case inputs into ``workdir``, run the reference (``binary``:
    the parity build, or REF_BIN_O3 for timing) on ``nranks`` MPI ranks
    (ranks = case.nx*case.ny*case.nz must hold), and return (bundle_in,
    scalars, ref_out, t_loop_seconds).  Dumps need one rank."""
    if not os.path.exists(binary):
        raise FileNotFoundError(f"{binary} is not built (make -f oracle/ref.mk)")
    if nranks != case.nx * case.ny * case.nz:
        raise ValueError("nranks must equal case.nx*case.ny*case.nz")
    if dump and nranks != 1:
        raise ValueError("bundle dumps need a one-rank run")
    case.write_inputs(workdir, dump=dump)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    if nranks == 1:
        cmd = [binary]
    else:
        cmd = [MPIEXEC, "-n", str(nranks), binary]
    p = subprocess.run(cmd, cwd=workdir, env=env, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0 or "EQD_REF" not in p.stdout:
        raise RuntimeError(f"reference run failed ({p.returncode}):\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
    LAST_STDOUT[0] = p.stdout
    m = re.search(r"t_loop_s\s+(\S+)", p.stdout)
    t_loop = float(m.group(1)) if m else float("nan")
    # start-up phases of the last run (seconds, max over ranks)
    LAST_PHASES.clear()
    for k in ("t_meshgen_s", "t_mass_s"):
        m = re.search(k + r"\s+(\S+)", p.stdout)
        if m:
            LAST_PHASES[k] = float(m.group(1))
    if not dump:
        return None, None, None, t_loop
    b = load_arrays(os.path.join(workdir, "bundle_in"))
    s = load_scalars(os.path.join(workdir, "bundle_in"))
    o = load_arrays(os.path.join(workdir, "ref_out"))
    return b, s, o, t_loop
