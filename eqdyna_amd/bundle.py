"""Array bundles in the reference's data layout.

A bundle is a directory with ``manifest.txt`` (``name dtype ndim d1 d2 d3``,
dimensions in Fortran order), one raw little-endian ``<name>.bin`` per array
and an optional ``scalars.txt`` (``name value``).  Arrays load as C-order numpy
arrays with the Fortran dimensions reversed, so ``meshCoor(3, N)`` becomes an
``(N, 3)`` array whose row ``n`` is node ``n+1`` of the reference
(src/globalvar.f90:78-101 lists the arrays).
"""
from __future__ import annotations

import os
from typing import Dict

import numpy as np


def load_arrays(d: str) -> Dict[str, np.ndarray]:
    out: Dict[str, np.ndarray] = {}
    with open(os.path.join(d, "manifest.txt")) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            name, dt, nd = t[0], t[1], int(t[2])
            dims = [int(x) for x in t[3:3 + nd]]
            a = np.fromfile(os.path.join(d, name + ".bin"), dtype="<" + dt)
            out[name] = a.reshape(dims[::-1])
    return out


def load_scalars(d: str) -> Dict[str, float]:
    p = os.path.join(d, "scalars.txt")
    out: Dict[str, float] = {}
    if os.path.exists(p):
        with open(p) as f:
            for line in f:
                t = line.split()
                if len(t) == 2:
                    out[t[0]] = float(t[1])
    return out


def save_arrays(d: str, arrays: Dict[str, np.ndarray], scalars: Dict[str, float] | None = None) -> None:
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "manifest.txt"), "w") as f:
        for name, a in arrays.items():
            a = np.ascontiguousarray(a)
            code = {"float64": "f8", "int32": "i4"}[str(a.dtype)]
            dims = list(a.shape[::-1])
            f.write(f"{name} {code} {len(dims)} " + " ".join(str(x) for x in dims) + "\n")
            a.astype("<" + code).tofile(os.path.join(d, name + ".bin"))
    if scalars is not None:
        with open(os.path.join(d, "scalars.txt"), "w") as f:
            for k, v in scalars.items():
                f.write(f"{k} {float(v)!r}\n")
