"""Case description and input-file writer for EQdyna-style rupture cases.

A ``Case`` holds the same parameters as the reference's ``par`` object
(scripts/defaultParameters.py:8-256, defaults taken from there) and writes the
fixed-order text inputs the reference's Fortran host reads
(formats: scripts/case.setup:10-76, readers: src/readInputFiles.f90:29-186).

On-fault fields are held as ``ofv[k, iz, ix]`` -- the reference's
``par.on_fault_vars[iz, ix, k]`` with the slot index first -- on the
(dip, strike) grid of the fault.  The 24 fields the reference reads
(src/netcdf_io.f90:41-64) are written as one raw little-endian float64 file
``on_fault_vars_input.bin`` whose Fortran shape is (nfx, nfz, 24).
"""
from __future__ import annotations

import copy
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# on_fault_vars slots in the order of the 24 netCDF variables
# (scripts/case.setup:163-186 write them; src/netcdf_io.f90:41-64 read them).
ON_FAULT_SLOTS = [1, 2, 3, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 40, 41, 42,
                  46, 8, 7, 20, 5, 4, 49]


def _s(v) -> str:
    """Python's str(), which is what scripts/case.setup writes."""
    return str(v)


@dataclass
class Case:
    # switches (scripts/defaultParameters.py:11-12, 74-106)
    mode: int = 1
    C_elastic: int = 1
    C_nuclea: int = 1
    C_degen: float = 0
    insertFaultType: int = 0
    friclaw: int = 5
    ntotft: int = 1
    nucfault: int = 1
    tpv: int = 105
    output_plastic: int = 0
    outputGroundMotion: int = 0
    outputFinalSurfDisp: int = 0
    nx: int = 1  # MPI ranks along x, y, z
    ny: int = 1
    nz: int = 1
    term: float = 5.0
    dt: float = 0.5 * 500.0 / 6.0e3
    nmat: int = 1
    n2mat: int = 3
    roumax: float = 2.67e3
    rhow: float = 1.0e3
    gamar: float = 0.66
    rdampk: float = 0.1
    vmaxPML: float = 6.0e3
    xsource: float = -4.0e3
    ysource: float = 0.0
    zsource: float = -7.5e3
    nucR: float = 1.5e3
    nucRuptVel: float = -9999.0
    nucdtau0: float = 50.0e6
    nucT: float = 1.0
    str1ToFaultAngle: float = 45.0
    devStrToStrVertRatio: float = 0.33
    bulk: float = 0.75
    coheplas: float = 5.0e6
    strike: float = 270.0
    dip: float = 90.0
    slipRateThres: float = 0.001
    # model geometry (scripts/defaultParameters.py:15-41)
    xmin: float = -42.0e3
    xmax: float = 42.0e3
    ymin: float = -20.0e3
    ymax: float = 22.0e3
    zmin: float = -42.0e3
    zmax: float = 0.0
    nuni_y_minus: int = 5
    nuni_y_plus: int = 5
    enlarging_ratio: float = 1.025
    dx: float = 500.0
    dy: float = 500.0
    dz: float = 500.0
    fxmin: float = -22.0e3
    fxmax: float = 22.0e3
    fymin: float = 0.0
    fymax: float = 0.0
    fzmin: float = -22.0e3
    fzmax: float = 0.0
    # material: (vp, vs, rho) or rows of (bottom depth, vp, vs, rho)
    vp: float = 6.0e3
    vs: float = 3.464e3
    rou: float = 2.67e3
    mat: Optional[np.ndarray] = None
    # stations: on-fault (x, z) km, off-fault (x, y, z) km
    st_on_fault: List[List[float]] = field(default_factory=list)
    st_off_fault: List[List[float]] = field(default_factory=list)
    # on-fault fields, ofv[k, iz, ix]
    ofv: Optional[np.ndarray] = None
    # rough fault interface (insertFaultType > 0): grid origin and spacing
    # (fx_min, fz_min, spacing) and rows (y offset, dy/dx, dy/dz), x-major with
    # z fastest, shape (nnx * nnz, 3) -- bFault_Rough_Geometry.txt,
    # src/readInputFiles.f90:268-312, used by src/meshgen.f90:921-962
    rough_nx: int = 0
    rough_nz: int = 0
    rough_origin: tuple = (0.0, 0.0, 0.0)
    rough: Optional[np.ndarray] = None
    # module knobs of the reference (src/globalvar.f90:16-75)
    nPML: int = 6
    R: float = 0.01
    C_Q: int = 0
    C_hg: int = 1
    rdampm: float = 0.0
    kapa_hg: float = 0.1
    name: str = "case"

    # ------------------------------------------------------------- geometry
    @property
    def nfx(self) -> int:
        return int(round((self.fxmax - self.fxmin) / self.dx + 1))

    @property
    def nfz(self) -> int:
        return int(round((self.fzmax - self.fzmin) / self.dz + 1))

    @property
    def fx(self) -> np.ndarray:
        return np.linspace(self.fxmin, self.fxmax, self.nfx)

    @property
    def fz(self) -> np.ndarray:
        return np.linspace(self.fzmin, self.fzmax, self.nfz)

    @property
    def nstep(self) -> int:
        # src/readInputFiles.f90:184: nstep = idnint(totalSimuTime/dt)
        return int(np.rint(self.term / self.dt))

    def new_ofv(self) -> np.ndarray:
        self.ofv = np.zeros((100, self.nfz, self.nfx))
        return self.ofv

    def copy(self) -> "Case":
        return copy.deepcopy(self)

    # ------------------------------------------------------------- writers
    def write_inputs(self, d: str, dump: bool = False) -> None:
        """Write bGlobal/bModelGeometry/bFaultGeometry/bMaterial/bStations.txt,
        on_fault_vars_input.bin and bKnobs.txt into directory ``d``."""
        os.makedirs(d, exist_ok=True)
        p = self
        g = [p.mode, p.C_elastic, p.C_nuclea, p.C_degen, p.insertFaultType, p.friclaw,
             p.ntotft, p.nucfault, p.tpv, p.output_plastic, p.outputGroundMotion,
             p.outputFinalSurfDisp]
        lines = [_s(v) for v in g]
        lines += [" ", f"{p.nx} {p.ny} {p.nz}", " ", _s(p.term), _s(p.dt), " ",
                  f"{p.nmat} {p.n2mat}", f"{_s(p.roumax)} {_s(p.rhow)} {_s(p.gamar)}",
                  f"{_s(p.rdampk)} {_s(p.vmaxPML)}", " ",
                  f"{_s(p.xsource)} {_s(p.ysource)} {_s(p.zsource)}",
                  f"{_s(p.nucR)} {_s(p.nucRuptVel)} {_s(p.nucdtau0)} {_s(p.nucT)}",
                  f"{_s(p.str1ToFaultAngle)} {_s(p.devStrToStrVertRatio)}",
                  f"{_s(p.bulk)} {_s(p.coheplas)}", f"{_s(p.strike)} {_s(p.dip)}",
                  _s(p.slipRateThres)]
        with open(os.path.join(d, "bGlobal.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")
        with open(os.path.join(d, "bModelGeometry.txt"), "w") as f:
            f.write(f"{_s(p.xmin)}  {_s(p.xmax)}\n{_s(p.ymin)}  {_s(p.ymax)}\n"
                    f"{_s(p.zmin)}  {_s(p.zmax)}\n \n{p.nuni_y_minus} {p.nuni_y_plus}\n"
                    f"{_s(p.enlarging_ratio)}\n{_s(p.dx)} {_s(p.dy)} {_s(p.dz)}\n")
        with open(os.path.join(d, "bFaultGeometry.txt"), "w") as f:
            for i in range(p.ntotft):
                f.write(f"For fault No. {i + 1}\n{_s(p.fxmin)}  {_s(p.fxmax)}\n"
                        f"{_s(p.fymin)}  {_s(p.fymax)}\n{_s(p.fzmin)}  {_s(p.fzmax)}\n")
        with open(os.path.join(d, "bMaterial.txt"), "w") as f:
            if p.nmat == 1:
                f.write(f"{_s(p.vp)} {_s(p.vs)} {_s(p.rou)}\n")
            else:
                for i in range(p.nmat):
                    f.write(" ".join(_s(float(v)) for v in p.mat[i, :4]) + "\n")
        with open(os.path.join(d, "bStations.txt"), "w") as f:
            f.write(f"{len(p.st_off_fault)}\n{len(p.st_on_fault)}\n \n")
            for s in p.st_on_fault:
                f.write(f"{_s(s[0])} {_s(s[1])}\n")
            f.write(" \n")
            for s in p.st_off_fault:
                f.write(f"{_s(s[0])} {_s(s[1])} {_s(s[2])}\n")
        if p.ofv is None:
            raise ValueError("case has no on-fault fields (ofv)")
        sel = np.ascontiguousarray(p.ofv[ON_FAULT_SLOTS], dtype="<f8")  # (24, nfz, nfx)
        sel.tofile(os.path.join(d, "on_fault_vars_input.bin"))
        if p.insertFaultType > 0:
            if p.rough is None or p.rough.shape != (p.rough_nx * p.rough_nz, 3):
                raise ValueError("insertFaultType > 0 needs rough rows of shape (rough_nx*rough_nz, 3)")
            fx0, fz0, sp = p.rough_origin
            with open(os.path.join(d, "bFault_Rough_Geometry.txt"), "w") as f:
                f.write(f"{p.rough_nx} {p.rough_nz}\n{_s(float(sp))} {_s(float(fx0))} {_s(float(fz0))}\n")
                for r in p.rough:
                    f.write(f"{_s(float(r[0]))} {_s(float(r[1]))} {_s(float(r[2]))}\n")
        with open(os.path.join(d, "bKnobs.txt"), "w") as f:
            f.write(f"nPML {p.nPML}\nR {_s(float(p.R))}\nC_Q {p.C_Q}\nC_hg {p.C_hg}\n"
                    f"rdampm {_s(float(p.rdampm))}\nkapa_hg {_s(float(p.kapa_hg))}\n"
                    f"dump {1 if dump else 0}\n")


# ---------------------------------------------------------------- helpers
def state_steady_state(a, b, d0, v0, r0, shear, norm, slip_rate, friclaw):
    """Steady-state RSF state variable, the closed form of scripts/lib.py:17-24."""
    if friclaw == 3:
        tmp = a * math.log(2.0 * math.sinh(abs(shear / norm / a))) - r0 - math.log(slip_rate / v0)
        return d0 / v0 * math.exp(tmp / b)
    return a * math.log(2.0 * v0 / slip_rate * math.sinh(abs(shear / norm / a)))


def shear_steady_state(a, b, v0, r0, load_rate, norm, slip_rate):
    """Steady-state shear stress of the ageing law, scripts/lib.py:11-14."""
    return -norm * a * math.asinh(slip_rate / 2.0 / v0 * math.exp((r0 + b * math.log(v0 / load_rate)) / a))


def boxcar_b1(x, ww, w):
    """Smooth boxcar B1 of TPV104/105 (scripts/lib.py:27-34)."""
    ax = abs(x)
    if ax <= ww:
        return 1.0
    if ax < ww + w:
        return 0.5 * (1.0 + math.tanh(w / (ax - ww - w) + w / (ax - ww)))
    return 0.0


def boxcar_b2(y, ww, w):
    """Smooth boxcar B2 of TPV105 in depth y >= 0 (scripts/lib.py:36-48)."""
    if y < w:
        return 0.5 * (1.0 + math.tanh(w / (w - y) - w / (y + 1e-30)))
    if y <= ww:
        return 1.0
    if y < ww + w:
        return 0.5 * (1.0 + math.tanh(w / (y - ww - w) + w / (y - ww)))
    return 0.0


def boxcar_b3(y, ww, w):
    """Smooth boxcar B3 of TPV105 in depth y >= 0 (scripts/lib.py:49-58)."""
    if y <= ww:
        return 1.0
    if y < ww + w:
        return 0.5 * (1.0 + math.tanh(w / (y - ww - w) + w / (y - ww)))
    return 0.0


# ---------------------------------------------------------------- presets
def tpv1053d() -> Case:
    """SCEC TPV105-3D with thermal pressurization as configured by the
    reference's default parameters (scripts/defaultParameters.py, which
    case_input/test.tpv1053d/user_defined_params.py uses with term = 5 s): the
    class defaults plus the depth-dependent normal stress, the smooth-boxcar
    fields of a, Vw and the hydraulic diffusivity, and 2 x 2 x 1 ranks (the
    case behind test.reference.results/test.tpv1053d)."""
    c = Case(name="tpv1053d")
    c.term = 5.0
    c.nx, c.ny, c.nz = 2, 2, 1
    grav = 9.8
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix], o[2, iz, ix], o[3, iz, ix], o[4, iz, ix] = 0.18, 0.12, 0.3, 0.0
            o[7, iz, ix] = -max(min(grav * 1670.0 * abs(z), 45.0e6), grav * 1670.0 * c.dx / 2.0)
            o[8, iz, ix] = -0.41 * o[7, iz, ix]
            t1 = boxcar_b1(x, 15.0e3, 3.0e3)
            t2 = boxcar_b2(-z, 15.0e3, 3.0e3)
            o[9, iz, ix] = 0.01 + (1.0 - t1 * t2) * 0.01
            o[10, iz, ix], o[11, iz, ix], o[12, iz, ix], o[13, iz, ix] = 0.014, 0.4, 1e-6, 0.6
            o[14, iz, ix] = 0.2
            o[15, iz, ix] = 0.1 + 0.9 * (1.0 - t1 * t2)
            t4 = boxcar_b3(-z, 15.0e3, 3.0e3)
            o[16, iz, ix] = 4.0e-4 + 1.0 * (1.0 - t1 * t4)
            o[17, iz, ix], o[18, iz, ix], o[19, iz, ix] = 1.0e-6, 2.7e6, 0.1e6
            o[40, iz, ix], o[41, iz, ix], o[42, iz, ix] = 0.02, 483.15, 80.0e6
            o[46, iz, ix] = 1e-16
            o[20, iz, ix] = state_steady_state(o[9, iz, ix], o[10, iz, ix], o[11, iz, ix], o[12, iz, ix],
                                               o[13, iz, ix], o[8, iz, ix], o[7, iz, ix], o[46, iz, ix], 5)
    c.st_on_fault = [[0.0, -3.0], [0.0, -7.5], [0.0, -12.0], [9.0, -7.5], [12.0, -3.0], [12.0, -12.0], [15.0, -7.5],
                     [18.0, -7.5], [-9.0, -7.5], [-12.0, -3.0], [-12.0, -12.0], [-15.0, -7.5], [-18.0, -7.5]]
    c.st_off_fault = [[0, 9, 0], [0, -9, 0], [12, 6, 0], [12, -6, 0], [-12, 6, 0], [-12, -6, 0]]
    return c


def tpv8() -> Case:
    """SCEC TPV8 as configured by case_input/test.tpv8/user_defined_params.py
    (the case behind test.reference.results/test.tpv8).  roumax and vmaxPML keep
    the class defaults of scripts/defaultParameters.py:46-48 because that file
    evaluates them before the case overrides vp/rou."""
    c = Case(name="tpv8")
    c.xmin, c.xmax = -22.0e3, 22.0e3
    c.ymin, c.ymax = -10.0e3, 12.0e3
    c.zmin, c.zmax = -22.0e3, 0.0e3
    c.fxmin, c.fxmax = -15.0e3, 15.0e3
    c.fymin, c.fymax = 0.0, 0.0
    c.fzmin, c.fzmax = -15.0e3, 0.0e3
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -12.0e3
    c.dx = c.dy = c.dz = 500.0
    c.nmat = 1
    c.vp, c.vs, c.rou = 5716, 3300, 2700
    c.term = 5.0
    c.dt = 0.5 * c.dx / c.vp
    c.friclaw = 1
    c.tpv = 8
    c.nucR = 1.5e3
    c.nx, c.ny, c.nz = 2, 2, 1
    o = c.new_ofv()
    fs, fd, d0, coh = 0.76, 0.448, 0.5, 1.0e6
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = fs
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01:
                o[1, iz, ix] = 1000.0
            o[2, iz, ix] = fd
            o[3, iz, ix] = d0
            o[4, iz, ix] = coh
            o[7, iz, ix] = 7378.0 * z
            o[8, iz, ix] = abs(0.55 * o[7, iz, ix])
            if abs(x - c.xsource) <= 1.5e3 and abs(z - c.zsource) <= 1.5e3:
                o[8, iz, ix] = 1e6 + abs(1.005 * 0.76 * o[7, iz, ix])
    c.st_on_fault = [[0.0, 0.0], [0.0, -4.5], [0.0, -7.5], [0.0, -12.0], [4.5, -7.5],
                     [12.0, -7.5], [4.5, 0.0], [12.0, 0.0]]
    c.st_off_fault = [[0, 1, 0], [0, -1, 0], [0, 2, 0], [0, -2, 0], [0, 3, 0], [0, -3, 0],
                      [12, 6, 0], [12, -3, 0], [-12, 3, 0], [0, -0.5, -0.3], [0, 0.5, -0.3],
                      [0, -1, -0.3], [0, 1, -0.3], [12, -3, -12], [12, 3, -12]]
    return c


TPV1053D_ON = [[0.0, -3.0], [0.0, -7.5], [0.0, -12.0], [9.0, -7.5], [12.0, -3.0], [12.0, -12.0], [15.0, -7.5],
               [18.0, -7.5], [-9.0, -7.5], [-12.0, -3.0], [-12.0, -12.0], [-15.0, -7.5], [-18.0, -7.5]]
TPV1053D_OFF = [[0, 9, 0], [0, -9, 0], [12, 6, 0], [12, -6, 0], [-12, 6, 0], [-12, -6, 0]]


def tpv104() -> Case:
    """SCEC TPV104 as configured by case_input/test.tpv104/user_defined_params.py
    (the case behind test.reference.results/test.tpv104): slip law with strong
    velocity weakening (friclaw 4), smooth-boxcar fields of a and Vw, TPV104
    nucleation, the class defaults of scripts/defaultParameters.py otherwise
    (stations included), 2 x 2 x 1 ranks."""
    c = Case(name="tpv104")
    c.xmin, c.xmax = -38.0e3, 38.0e3
    c.ymin, c.ymax = -20.0e3, 22.0e3
    c.zmin, c.zmax = -38.0e3, 0.0e3
    c.fxmin, c.fxmax = -18.0e3, 18.0e3
    c.fymin, c.fymax = 0.0, 0.0
    c.fzmin, c.fzmax = -18.0e3, 0.0e3
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -7.5e3
    c.term = 5.0
    c.dx = c.dy = c.dz = 500.0
    c.dt = 0.5 * c.dx / c.vp
    c.friclaw = 4
    c.tpv = 104
    c.nucR = 3.0e3
    c.nucdtau0 = 45.0e6
    c.nx, c.ny, c.nz = 2, 2, 1
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.18, 0.12, 0.3
            o[7, iz, ix], o[8, iz, ix] = -120.0e6, 40.0e6
            t1 = boxcar_b1(x, 15.0e3, 3.0e3)
            t2 = boxcar_b1(-z - 7.5e3, 15.0e3 / 2.0, 3.0e3)
            o[9, iz, ix] = 0.01 + (1.0 - t1 * t2) * 0.01
            o[10, iz, ix], o[11, iz, ix], o[12, iz, ix], o[13, iz, ix] = 0.014, 0.4, 1e-6, 0.6
            o[14, iz, ix] = 0.2
            o[15, iz, ix] = 0.1 + 0.9 * (1.0 - t1 * t2)
            o[16, iz, ix] = 4.0e-4 + 1.0 * (1.0 - t1 * t2)
            o[17, iz, ix], o[18, iz, ix], o[19, iz, ix] = 1.0e-6, 2.7e6, 0.1e6
            o[40, iz, ix], o[41, iz, ix], o[42, iz, ix] = 0.02, 483.15, 80.0e6
            o[46, iz, ix] = 1e-16
            o[20, iz, ix] = state_steady_state(o[9, iz, ix], o[10, iz, ix], o[11, iz, ix], o[12, iz, ix],
                                               o[13, iz, ix], o[8, iz, ix], o[7, iz, ix], o[46, iz, ix], 4)
    c.st_on_fault = [list(s) for s in TPV1053D_ON]
    c.st_off_fault = [list(s) for s in TPV1053D_OFF]
    return c


def _meng2023(name: str) -> Case:
    """Shared set-up of the meng2023 cases (case_input/test.meng2023a,
    test.meng2023cb): a five-layer medium, time weakening (friclaw 2) with the
    TPV201-style forced nucleation, 400 m cells, 2 x 2 x 1 ranks.  roumax and
    vmaxPML keep the class defaults because scripts/defaultParameters.py
    evaluates them before the case sets its layers."""
    c = Case(name=name)
    c.xmin, c.xmax = -30.0e3, 30.0e3
    c.ymin, c.ymax = -28.0e3, 32.0e3
    c.zmin, c.zmax = -34.0e3, 0.0e3
    c.fxmin, c.fxmax = -6.0e3, 6.0e3
    c.fymin, c.fymax = 0.0, 0.0
    c.fzmin, c.fzmax = -8.0e3, 0.0e3
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -3.4e3
    c.dx = c.dy = c.dz = 400.0
    c.nuni_y_plus = c.nuni_y_minus = 10
    c.enlarging_ratio = 1.025
    c.nmat, c.n2mat = 5, 4
    c.mat = np.array([[1.19e3, 2.74e3, 1.45e3, 2.1e3], [2.01e3, 5.75e3, 3.06e3, 2.4e3],
                      [4.94e3, 5.72e3, 3.4e3, 2.6e3], [10.94e3, 6.18e3, 3.62e3, 2.8e3], [1.e6, 6.32e3, 3.67e3, 2.8e3]])
    c.term = 5.0
    c.dt = 0.5 * c.dx / 6.32e3
    c.friclaw = 2
    c.tpv = 201
    c.nucR = 25.e3
    c.nucRuptVel = 1.5e3
    c.nucdtau0 = -9999.
    c.nx, c.ny, c.nz = 2, 2, 1
    c.st_on_fault = [list(s) for s in TPV1053D_ON]
    c.st_off_fault = [list(s) for s in TPV1053D_OFF]
    c.new_ofv()
    return c


def meng2023a() -> Case:
    """case_input/test.meng2023a/user_defined_params.py: a rupturable patch
    of 4.5 x 4.5 km with a tapered dynamic friction, locked elsewhere."""
    c = _meng2023("meng2023a")
    o = c.ofv
    fs, d0, t0 = 0.4, 0.3, 0.2
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = o[2, iz, ix] = 1000.
            if abs(x) < 2.25e3 and z < -3.4e3 + 2.25e3 and z > -3.4e3 - 2.25e3:
                t1 = t2 = 1.
                if abs(x) >= 1.25e3:
                    t1 = (2.25e3 - abs(x)) / 1.0e3
                if abs(z - -3.4e3) >= 1.25e3:
                    t2 = (2.25e3 - abs(z - -3.4e3)) / 1.0e3
                o[1, iz, ix] = fs
                o[2, iz, ix] = fs - 0.1 * t1 * t2
            o[3, iz, ix], o[5, iz, ix] = d0, t0
            o[7, iz, ix], o[8, iz, ix] = -100.0e6, 35.0e6
    return c


def meng2023cb() -> Case:
    """case_input/test.meng2023cb/user_defined_params.py: a 6 x 6 km patch
    with four strengthening barriers."""
    c = _meng2023("meng2023cb")
    o = c.ofv
    fs, d0, t0 = 0.4, 0.3, 0.2
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix], o[2, iz, ix] = 0.4, 0.8
            if abs(x) < 3.e3 and z > -6.4e3 and z < -0.4e3:
                t1 = t2 = 1.
                for cx, cz in ((2.0e3, -3.4e3), (-2.0e3, -3.4e3), (0.0, -1.4e3), (0.0, -5.4e3)):
                    if abs(x - cx) < 1.0e3 and abs(z - cz) < 1.0e3:
                        t1, t2 = 1., -4.
                o[1, iz, ix] = fs
                o[2, iz, ix] = fs - 0.1 * t1 * t2
            o[3, iz, ix], o[5, iz, ix] = d0, t0
            o[7, iz, ix], o[8, iz, ix] = -100.0e6, 35.0e6
    return c


def tpv10() -> Case:
    """SCEC TPV10 as configured by case_input/test.tpv10/user_defined_params.py:
    a 60-degree dipping planar fault inserted as a shifted-node interface
    (insertFaultType 1) with the grid and time step of scripts/defaultParameters.py
    mod4dip, slip weakening with cohesion, depth-dependent normal and dip-direction
    shear stress, 2 x 2 x 1 ranks.  bFault_Rough_Geometry.txt follows the
    construction of scripts/generateFaultInterface for a planar fault: y offset
    (nz - 1 - j) dx cos(dip) at fault row j, the slopes by numpy.gradient, every
    value as printed with '%f'."""
    c = Case(name="tpv10")
    dip = 60.0
    c.dip = dip
    c.xmin, c.xmax = -30.0e3, 30.0e3
    c.ymin, c.ymax = -20.0e3, 40.0e3
    c.zmin, c.zmax = -35.0e3, 0.0e3
    c.fxmin, c.fxmax = -15.0e3, 15.0e3
    c.fzmin, c.fzmax = -15.0e3, 0.0e3
    c.xsource = 0.0
    c.dx = c.dy = 500.
    c.nmat = 1
    c.vp, c.vs, c.rou = 5716, 3300, 2700
    c.term = 5.
    c.C_elastic = 1
    c.insertFaultType = 1
    c.friclaw = 1
    c.tpv = 10
    c.C_nuclea = 0
    c.nucR = 1.5e3
    s, co = math.sin(dip / 180. * math.pi), math.cos(dip / 180. * math.pi)
    c.fzmin = -15.0e3 * s
    c.ysource, c.zsource = 12.0e3 * co, -12.0e3 * s
    c.dz = c.dx * s
    c.dt = 0.5 * c.dz / c.vp
    c.nx, c.ny, c.nz = 2, 2, 1
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            dd = abs(z) / math.sin(abs(dip) / 180. * math.pi)
            o[1, iz, ix] = 0.76
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01:
                o[1, iz, ix] = 1000.
            o[2, iz, ix], o[3, iz, ix], o[4, iz, ix] = 0.448, 0.5, 0.2e6
            o[7, iz, ix] = -7378. * dd
            o[8, iz, ix] = 0.
            o[49, iz, ix] = -abs(0.55 * o[7, iz, ix])
            if abs(x - c.xsource) <= 1.5e3 and abs(z - c.zsource) <= 1.5e3 * math.sin(abs(dip) / 180. * math.pi):
                o[49, iz, ix] = -0.2e6 - abs((0.76 + 0.0057) * o[7, iz, ix])
    C = math.sin(dip / 180. * math.pi)
    c.st_on_fault = [[0.0, -1.5 * C], [0.0, -3.0 * C], [0.0, -4.5 * C], [0.0, -7.5 * C], [0.0, 0.0], [4.5, 0.0],
                     [12.0, 0.0], [0.0, -12.0 * C], [4.5, -7.5 * C], [12.0, -7.5 * C]]
    t = math.tan(abs(dip) / 180. * math.pi)
    c.st_off_fault = [[0, 1, 0], [0, -1, 0], [0, 2, 0], [0, -2, 0], [0, 3, 0], [0, -3, 0], [12, 3, 0], [12, -3, 0],
                      [0, 0.5 + 0.3 / t, -0.3], [0, -0.5 + 0.3 / t, -0.3]]
    nx, nz = c.nfx, c.nfz
    surf = np.zeros((nz, nx))
    for i in range(nx):
        for j in range(nz):
            surf[j, i] = surf[j, i] + (nz - 1 - j) * c.dx * co
    ddx = np.gradient(surf / c.dx, axis=1)
    ddz = np.gradient(surf / c.dz, axis=0)
    pr = lambda v: float("%f" % v)  # noqa: E731  (the generator writes '%f')
    c.rough_nx, c.rough_nz = nx, nz
    c.rough_origin = (pr(c.fxmin), pr(c.fzmin), pr(c.dx))
    c.rough = np.array([[pr(surf[j, i]), pr(ddx[j, i]), pr(ddz[j, i])] for i in range(nx) for j in range(nz)])
    return c


def box_sw(nx_half: int = 6, nz: int = 6, ny_half: int = 3, d: float = 200.0,
           npml: int = 3, steps: int = 40, friclaw: int = 1, layered: bool = False,
           name: str = "box_sw") -> Case:
    """Small homogeneous (or layered) box with a vertical split-node fault at
    y=0, slip-weakening (friclaw 1) or time-weakening (friclaw 2) friction, an
    over-stressed nucleation patch, PML on five faces and a free surface."""
    c = Case(name=name)
    c.dx = c.dy = c.dz = d
    fx = nx_half * d
    c.fxmin, c.fxmax = -fx, fx
    c.fymin, c.fymax = 0.0, 0.0
    c.fzmin, c.fzmax = -nz * d, 0.0
    c.xmin, c.xmax = -fx - 0.5 * d, fx + 0.5 * d
    c.ymin, c.ymax = -ny_half * d - 0.5 * d, ny_half * d + 0.5 * d
    c.zmin, c.zmax = -nz * d - 0.5 * d, 0.0
    c.nuni_y_minus = ny_half
    c.nuni_y_plus = ny_half
    c.enlarging_ratio = 1.0
    c.nPML = npml
    c.R = 0.01
    c.vp, c.vs, c.rou = 6000.0, 3464.0, 2670.0
    c.roumax, c.vmaxPML = 2670.0, 6000.0
    if layered:
        c.nmat, c.n2mat = 3, 4
        c.mat = np.array([[0.4 * nz * d, 5000.0, 2887.0, 2500.0],
                          [0.8 * nz * d, 5600.0, 3233.0, 2600.0],
                          [1.0e6, 6000.0, 3464.0, 2670.0]])
    c.dt = 0.25 * d / 6000.0
    c.term = steps * c.dt
    c.friclaw = friclaw
    c.tpv = 8 if friclaw == 1 else 201
    c.C_nuclea = 0 if friclaw == 1 else 1
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -0.5 * nz * d
    c.nucR = 2.5 * d
    c.nucRuptVel = 1500.0
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = 0.60
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01:
                o[1, iz, ix] = 1.0e4
            o[2, iz, ix] = 0.45
            o[3, iz, ix] = 0.02
            o[5, iz, ix] = 0.05
            o[7, iz, ix] = -100.0e6
            o[8, iz, ix] = 55.0e6
            if abs(x - c.xsource) <= 1.5 * d and abs(z - c.zsource) <= 1.5 * d:
                o[8, iz, ix] = 62.0e6
    c.st_on_fault = [[0.0, c.zsource / 1e3], [2 * d / 1e3, c.zsource / 1e3], [0.0, 0.0]]
    c.st_off_fault = [[0.0, 2 * d / 1e3, 0.0], [2 * d / 1e3, -2 * d / 1e3, 0.0]]
    return c


def box_rsf(friclaw: int = 4, nx_half: int = 6, nz: int = 6, ny_half: int = 3, d: float = 200.0,
            npml: int = 3, steps: int = 60, a_lo: float = 0.008, a_hi: float = 0.016,
            seed: int = 1, name: str = "box_rsf") -> Case:
    """Small box with rate-and-state friction (ageing law 3 or slip law with
    strong velocity weakening 4), TPV104-style smooth nucleation, and a seeded
    spatially varying direct-effect parameter a in [a_lo, a_hi]."""
    c = box_sw(nx_half, nz, ny_half, d, npml, steps, 1, False, name)
    c.friclaw = friclaw
    c.tpv = 104
    c.C_nuclea = 1
    c.nucR = 3.0 * d
    c.nucT = 10.0 * c.dt
    c.nucdtau0 = 40.0e6
    rng = np.random.default_rng(seed)
    o = c.new_ofv()
    b, L, f0, v0, fw, vw = 0.012, 0.02, 0.6, 1.0e-6, 0.1, 0.1
    vini = 1.0e-12 if friclaw != 3 else 1.0e-9
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            a = a_lo + (a_hi - a_lo) * rng.random()
            o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.6, 0.45, 0.3
            o[7, iz, ix] = -60.0e6
            o[8, iz, ix] = 24.0e6
            if friclaw == 3:
                o[8, iz, ix] = shear_steady_state(a, b, v0, f0, vini, o[7, iz, ix], vini)
            o[9, iz, ix] = a
            o[10, iz, ix] = b
            o[11, iz, ix] = L
            o[12, iz, ix] = v0
            o[13, iz, ix] = f0
            o[14, iz, ix] = fw
            o[15, iz, ix] = vw
            o[46, iz, ix] = vini
            if friclaw == 3:
                o[20, iz, ix] = L / vini  # ageing-law steady state theta = L/V
            else:
                o[20, iz, ix] = state_steady_state(a, b, L, v0, f0, o[8, iz, ix], o[7, iz, ix], vini, friclaw)
    return c


def box_tp(steps: int = 60, seed: int = 5, name: str = "box_tp") -> Case:
    """box_rsf with thermal pressurization (friclaw 5: slip law with strong
    velocity weakening plus the pore pressure and temperature of
    src/thermop.f90) and TPV105 nucleation and creep floor: hydraulic
    diffusivity seeded in [4e-4, 4e-3] m^2/s, thermal diffusivity 1e-6 m^2/s,
    rho*c 2.7e6 J/(m^3 K), Lambda 0.1 MPa/K, T_ini 483.15 K, p_ini 80 MPa
    (the parameter values of scripts/defaultParameters.py:139-146)."""
    c = box_rsf(friclaw=4, steps=steps, name=name)
    c.friclaw = 5
    c.tpv = 105
    rng = np.random.default_rng(seed)
    o = c.ofv
    for ix in range(c.nfx):
        for iz in range(c.nfz):
            o[16, iz, ix] = 4.0e-4 + 3.6e-3 * rng.random()
            o[17, iz, ix] = 1.0e-6
            o[18, iz, ix] = 2.7e6
            o[19, iz, ix] = 0.1e6
            o[40, iz, ix] = 0.02
            o[41, iz, ix] = 483.15
            o[42, iz, ix] = 80.0e6
    return c


def box_plastic(steps: int = 40, friclaw: int = 1, name: str = "box_plastic") -> Case:
    """Small box with Drucker-Prager viscoplasticity (C_elastic = 0): gravity,
    lithostatic initial stress (src/meshgen.f90:976-996) and no on-fault
    initial tractions, as the reference defines the plastic mode."""
    c = box_sw(steps=steps, friclaw=friclaw, name=name)
    c.C_elastic = 0
    c.coheplas = 1.0e6
    c.bulk = 0.75
    c.str1ToFaultAngle = 45.0
    c.devStrToStrVertRatio = 0.6
    c.gamar = 0.66
    c.rhow = 1.0e3
    c.tpv = 8
    o = c.ofv
    o[1, :, :] = np.where(o[1] > 100.0, o[1], 0.30)
    o[2, :, :] = 0.20
    return c


def box_dip(dip: float = 30.0, nx_half: int = 6, nd: int = 6, pad: int = 3, d: float = 200.0,
            npml: int = 3, steps: int = 60, name: str = "box_dip") -> Case:
    """Small box with a planar fault dipping ``dip`` degrees towards y+ on
    wedge elements (C_degen = dip, src/library_degeneration.f90:3-55): the
    grid spacing is d along strike and d*cos(dip), d*sin(dip) in y and z so
    that the fault plane z = -y*tan(dip) runs through grid nodes and cuts the
    hexahedra along a diagonal (the set-up of case_input/test.tpv36).  Slip
    weakening with the forced expanding nucleation of TPV36
    (src/faulting.f90:418-440), along-strike and along-dip initial shear."""
    c = Case(name=name)
    s, co = math.sin(math.radians(dip)), math.cos(math.radians(dip))
    c.dip = dip
    c.C_degen = dip
    c.insertFaultType = 0
    c.dx, c.dy, c.dz = d, d * co, d * s
    fx = nx_half * d
    c.fxmin, c.fxmax = -fx, fx
    c.fymin, c.fymax = 0.0, nd * c.dy
    c.fzmin, c.fzmax = -nd * c.dz, 0.0
    c.xmin, c.xmax = -fx - 0.5 * d, fx + 0.5 * d
    c.nuni_y_minus, c.nuni_y_plus = pad, nd + pad
    c.ymin, c.ymax = -(pad + 0.5) * c.dy, (nd + pad + 0.5) * c.dy
    c.zmin, c.zmax = c.fzmin - (pad + 0.5) * c.dz, 0.0
    c.enlarging_ratio = 1.0
    c.nPML = npml
    c.R = 0.01
    c.vp, c.vs, c.rou = 6000.0, 3464.0, 2670.0
    c.roumax, c.vmaxPML = 2670.0, 6000.0
    c.dt = 0.5 * c.dz / 6000.0
    c.term = steps * c.dt
    c.friclaw = 1
    c.tpv = 36
    c.C_nuclea = 1
    src = 0.5 * nd * d  # down-dip distance of the hypocentre
    c.xsource, c.ysource, c.zsource = 0.0, src * co, -src * s
    c.nucR = 2.5 * d
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = 0.60
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01:
                o[1, iz, ix] = 1.0e4
            o[2, iz, ix] = 0.45
            o[3, iz, ix] = 0.02
            o[5, iz, ix] = 0.05
            o[7, iz, ix] = -100.0e6
            o[8, iz, ix] = 20.0e6
            o[49, iz, ix] = 50.0e6
    c.st_on_fault = [[0.0, c.zsource / 1e3], [2 * d / 1e3, -2 * c.dz / 1e3], [0.0, 0.0]]
    c.st_off_fault = [[0.0, 2 * c.dy / 1e3, 0.0], [2 * d / 1e3, -2 * c.dy / 1e3, 0.0]]
    return c


def rough_surface(c: Case, amp: float, lx: float, lz: float):
    """Smooth fault-surface offsets y(x, z) = amp*sin(2*pi*x/lx)*cos(pi*z/lz)
    with their exact slopes, sampled on the fault grid of ``c`` (spacing dx
    along strike, dz along dip) as the rows of bFault_Rough_Geometry.txt.
    (The reference's scripts/generateFaultInterface draws a random fractal
    surface; a closed-form one keeps the fixture reproducible.)"""
    xs, zs = c.fx, c.fz
    rows = []
    for x in xs:            # x-major, z fastest: row nnz*(ix-1)+iz
        for z in zs:
            sx, cx = math.sin(2 * math.pi * x / lx), math.cos(2 * math.pi * x / lx)
            sz, cz = math.sin(math.pi * z / lz), math.cos(math.pi * z / lz)
            rows.append([amp * sx * cz, amp * 2 * math.pi / lx * cx * cz, -amp * math.pi / lz * sx * sz])
    c.rough_nx, c.rough_nz = len(xs), len(zs)
    c.rough_origin = (float(xs[0]), float(zs[0]), float(c.dx))
    c.rough = np.array(rows)


def box_rough(friclaw: int = 4, amp: float = 20.0, steps: int = 60, name: str = "box_rough") -> Case:
    """The rate-and-state box on a rough fault (insertFaultType = 1): every
    node column is shifted in y by the surface offset at its (x, z) and the
    split-node normals follow the local slopes (src/meshgen.f90:874-889,
    921-962); elastic, so the normal stress is held in [-40, -10] MPa
    (src/faulting.f90:208-219)."""
    c = box_rsf(friclaw=friclaw, steps=steps, name=name)
    c.insertFaultType = 1
    rough_surface(c, amp, lx=c.fxmax - c.fxmin, lz=2.0 * (c.fzmax - c.fzmin))
    # initial normal stress at the cap, so that every compressive change is
    # capped; shear and state in steady state at that value
    o = c.ofv
    capped = -40.0e6
    for ix in range(c.nfx):
        for iz in range(c.nfz):
            o[7, iz, ix] = capped
            o[8, iz, ix] = 0.4 * -capped
            if friclaw != 3:
                o[20, iz, ix] = state_steady_state(o[9, iz, ix], o[10, iz, ix], o[11, iz, ix], o[12, iz, ix],
                                                   o[13, iz, ix], o[8, iz, ix], capped, o[46, iz, ix], friclaw)
    return c


def gardner_rho(vp: float) -> float:
    """Gardner's relation in SI units, rho = 310 * Vp**0.25 (kg/m^3, m/s)."""
    return 310.0 * vp ** 0.25


def layered_medium(seed: int, n_layers: int = 5, vs_top: float = 2000.0, vs_bottom: float = 3464.0,
                   max_depth: float = 10.0e3) -> np.ndarray:
    """Seeded 1-D medium: n_layers rows of (bottom depth, Vp, Vs, rho) with Vs
    increasing from vs_top to vs_bottom with depth, Vp/Vs = 1.732 and Gardner
    density; the last layer extends to 1e6 m (the bMaterial.txt row format,
    src/readInputFiles.f90:177-179, src/meshgen.f90:172-198)."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.uniform(0.05, 0.95, n_layers - 1)) * max_depth
    vs = np.linspace(vs_top, vs_bottom, n_layers)
    rows = []
    for i in range(n_layers):
        depth = float(np.round(cuts[i], 0)) if i < n_layers - 1 else 1.0e6
        vp = round(1.732 * vs[i], 3)
        rows.append([depth, vp, float(vs[i]), round(gardner_rho(vp), 3)])
    return np.array(rows)


def _strike_slip_box(d: float, npml: int, name: str) -> Case:
    """Shared set-up of the benchmark boxes (c1, c2): a vertical strike-slip split-node
    fault 16 km along strike, rupturable from 2 to 9.8 km depth (the fault
    plane must reach the free surface in the reference, src/meshgen.f90:476-531,
    so the top 2 km are locked with mu_s = 1e4 as TPV8 does), linear slip
    weakening mu_s = 0.677, mu_d = 0.45, Dc = 0.3 m, sigma_n = -100 MPa,
    tau_0 = 65 MPa and a 3 x 3 km nucleation patch at 75 MPa."""
    c = Case(name=name)
    c.dx = c.dy = c.dz = d
    c.fxmin, c.fxmax = -8.0e3, 8.0e3
    c.fymin = c.fymax = 0.0
    c.fzmin, c.fzmax = -9.8e3, 0.0
    c.enlarging_ratio = 1.0
    c.nPML = npml
    c.R = 0.001  # pairs with nPML = 10 (src/globalvar.f90:23-25)
    c.friclaw = 1
    c.tpv = 8
    c.C_nuclea = 0
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -6.0e3
    c.nucR = 1.5e3
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = 0.677
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01 or z > -2.0e3 + 0.01:
                o[1, iz, ix] = 1.0e4
            o[2, iz, ix] = 0.45
            o[3, iz, ix] = 0.3
            o[7, iz, ix] = -100.0e6
            o[8, iz, ix] = 65.0e6
            if abs(x - c.xsource) <= 1.5e3 and abs(z - c.zsource) <= 1.5e3:
                o[8, iz, ix] = 75.0e6
    xs = [-6.0, -2.0, 2.0, 6.0]
    zs = [-3.0, -5.0, -7.0, -9.0]
    c.st_on_fault = [[x, z] for x in xs for z in zs]
    c.st_off_fault = [[x, y, 0.0] for x in (-6.0, -2.0, 2.0, 6.0) for y in (-3.0, -1.0, 1.0, 3.0)]
    return c


def c1(steps: int = 600) -> Case:
    """c1: homogeneous elastic box at 200 m, 120 x 60 x 60 =
    432,000 hexahedra including a 10-element PML on five faces, rho = 2670,
    Vp = 6000, Vs = 3464, dt = 0.01 s."""
    c = _strike_slip_box(200.0, 10, "c1")
    c.xmin, c.xmax = -10.0e3, 10.0e3
    c.ymin, c.ymax = -4.0e3, 4.0e3
    c.zmin, c.zmax = -10.0e3, 0.0
    c.nuni_y_minus = c.nuni_y_plus = 19
    c.vp, c.vs, c.rou = 6000.0, 3464.0, 2670.0
    c.roumax, c.vmaxPML = 2670.0, 6000.0
    c.dt = 0.01
    c.term = steps * c.dt
    return c


def c2(steps: int = 4000, seed: int = 2, q: bool = True) -> Case:
    """c2, the large elastic box at 50 m: 480 x 240 x 240 = 27,648,000 hexahedra
    including a 10-element PML, a seeded five-layer medium (Vs 2000 -> 3464
    m/s with depth, Vp/Vs = 1.732, Gardner density), the reference's
    frequency-independent Q (C_Q = 1, src/calcElemKU.f90:77-132) and
    dt = 0.0025 s."""
    c = _strike_slip_box(50.0, 10, "c2")
    c.xmin, c.xmax = -11.5e3, 11.5e3
    c.ymin, c.ymax = -5.5e3, 5.5e3
    c.zmin, c.zmax = -11.5e3, 0.0
    c.nuni_y_minus = c.nuni_y_plus = 76
    c.nmat, c.n2mat = 5, 4
    c.mat = layered_medium(seed)
    c.vp, c.vs, c.rou = c.mat[-1, 1], c.mat[-1, 2], c.mat[-1, 3]
    c.roumax, c.vmaxPML = float(c.mat[-1, 3]), float(c.mat[-1, 1])
    c.C_Q = 1 if q else 0
    c.dt = 0.0025
    c.term = steps * c.dt
    return c


def c2_sample(steps: int = 100, seed: int = 2, q: bool = True) -> Case:
    """A bounded sample of the C2 workload for timing the reference on the
    host: the same 50 m resolution, seeded layered medium, Q and friction on a
    240 x 120 x 120 = 3,456,000-hexahedron sub-box (fault 8 km x 4.8 km)."""
    c = _strike_slip_box(50.0, 10, "c2_sample")
    c.fxmin, c.fxmax = -4.0e3, 4.0e3
    c.fzmin, c.fzmax = -4.8e3, 0.0
    c.xmin, c.xmax = -5.5e3, 5.5e3
    c.ymin, c.ymax = -2.5e3, 2.5e3
    c.zmin, c.zmax = -5.5e3, 0.0
    c.nuni_y_minus = c.nuni_y_plus = 35
    c.xsource, c.zsource = 0.0, -3.0e3
    o = c.new_ofv()
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix] = 0.677
            if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01 or z > -1.0e3 + 0.01:
                o[1, iz, ix] = 1.0e4
            o[2, iz, ix], o[3, iz, ix] = 0.45, 0.3
            o[7, iz, ix], o[8, iz, ix] = -100.0e6, 65.0e6
            if abs(x - c.xsource) <= 1.5e3 and abs(z - c.zsource) <= 1.5e3:
                o[8, iz, ix] = 75.0e6
    c.st_on_fault = [[x, z] for x in (-2.0, 2.0) for z in (-2.0, -4.0)]
    c.st_off_fault = [[x, y, 0.0] for x in (-2.0, 2.0) for y in (-1.0, 1.0)]
    c.nmat, c.n2mat = 5, 4
    c.mat = layered_medium(seed)
    c.vp, c.vs, c.rou = c.mat[-1, 1], c.mat[-1, 2], c.mat[-1, 3]
    c.roumax, c.vmaxPML = float(c.mat[-1, 3]), float(c.mat[-1, 1])
    c.C_Q = 1 if q else 0
    c.dt = 0.0025
    c.term = steps * c.dt
    return c


def fractal_surface(n: int, hurst: float, seed: int) -> np.ndarray:
    """Seeded self-affine random surface on an n x n grid (unit spacing): a
    complex Gaussian spectrum shaped to power |k|^-2(H+1), low-pass filtered
    with a fourth-order Butterworth roll-off at |k| = n/4, the three longest
    wavelengths (|k|^2 <= 2) removed so the surface carries no overall tilt,
    transformed back and normalised to zero mean and unit standard deviation
    (the construction of Dunham, Belanger, Cong and Kozdon, BSSA 2011; cf.
    scripts/generateFaultInterface:11-52)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, d=1.0 / n)  # integer wavenumbers 0..n/2, -n/2+1..-1
    ksq = k[:, None] ** 2 + k[None, :] ** 2
    with np.errstate(divide="ignore"):
        power = np.where(ksq > 2.0, ksq ** -(hurst + 1.0), 0.0)
    filt = 1.0 / (1.0 + (ksq / (n / 4.0) ** 2) ** 4)
    spec = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5 * power * filt)
    w = np.real(np.fft.ifft2(spec))
    return (w - w.mean()) / w.std()


def rough_fractal(c: Case, alpha: float, hurst: float, seed: int) -> None:
    """A fractal fault surface for ``c``: rms height alpha times the fault
    length along strike, sampled at the fault grid (dx along strike, dz down
    dip), slopes by centred differences (one-sided at the edges)."""
    nx, nz = c.nfx, c.nfz
    n = 1 << int(math.ceil(math.log2(max(nx, nz))))
    w = fractal_surface(n, hurst, seed)[:nz, :nx]
    h = (w - w.mean()) / w.std() * alpha * (c.fxmax - c.fxmin)
    dhdx = np.gradient(h, c.dx, axis=1)
    dhdz = np.gradient(h, c.dz, axis=0)
    rows = np.stack([h.T.ravel(), dhdx.T.ravel(), dhdz.T.ravel()], axis=1)  # x-major, z fastest
    c.rough_nx, c.rough_nz = nx, nz
    c.rough_origin = (float(c.fxmin), float(c.fzmin), float(c.dx))
    c.rough = np.ascontiguousarray(rows)


def reference_fractal_surface(n: int, hurst: float, seed: int) -> np.ndarray:
    """The reference's own fractal surface, value for value: the spectrum of
    scripts/generateFaultInterface:11-52 built from numpy's legacy
    RandomState stream (seeded once; n*n normal deviates for the real parts,
    then n*n for the imaginary parts, both times sqrt(2)/2), amplitude
    |k|^-(H+1) with the Butterworth roll-off 1/(1+(|k|^2/(n/4)^2)^4), the
    wavenumbers with |k|^2 <= 2 removed, and a quarter of the real part of
    the inverse transform.  Unlike fractal_surface it keeps the reference's
    RNG, wavenumber indexing and scaling, so a case the reference generated
    with a seed can be rebuilt."""
    rs = np.random.RandomState(seed)
    re_part = rs.normal(size=(n, n))
    im_part = rs.normal(size=(n, n))
    spec = 0.5 * math.sqrt(2.0) * (re_part + 1j * im_part)
    half = n // 2
    k = np.arange(n)
    k = np.where(k > half, k - n, k).astype(np.float64)
    ksq = k[:, None] ** 2 + k[None, :] ** 2
    kcut_sq = (n / 4) * (n / 4)
    amp = np.zeros((n, n))
    keep = ksq > 2.0
    amp[keep] = np.sqrt(1.0 / ksq[keep] ** (hurst + 1.0) * (1.0 / (1.0 + (ksq[keep] / kcut_sq) ** 4)))
    shaped = amp * spec
    shaped[~keep] = 0.0
    return 0.25 * np.real(np.fft.ifft2(shaped))


def rough_reference_fractal(c: Case, seed: int = 1, hurst: float = 1.0, alpha: float = 0.005) -> None:
    """bFault_Rough_Geometry.txt as the reference's generator writes it for
    ``c`` (scripts/generateFaultInterface:54-115): the surface on an
    nfx x nfx grid scaled so that its standard deviation is alpha times
    nfx * dx, its first nfz rows, the dip term (nz-1-j) dx cos(dip), slopes
    by numpy's gradient, every value rounded to six decimals (the file's
    '%f' format)."""
    nx, nz = c.nfx, c.nfz
    w = reference_fractal_surface(nx, hurst, seed)
    surf = (w * alpha / (np.std(w) / nx / c.dx))[:nz, :nx]
    surf = surf + (nz - 1 - np.arange(nz))[:, None] * c.dx * math.cos(c.dip / 180.0 * math.pi)
    dhdx = np.gradient(surf / c.dx, axis=1)
    dhdz = np.gradient(surf / c.dz, axis=0)
    rows = np.stack([surf.T.ravel(), dhdx.T.ravel(), dhdz.T.ravel()], axis=1)  # x-major, z fastest
    c.rough_nx, c.rough_nz = nx, nz
    c.rough_origin = (float(c.fxmin), float(c.fzmin), float(c.dx))
    c.rough = np.vectorize(lambda v: float("%f" % v))(rows)


def _taper(x: float, z: float, wwx: float, wx: float, wwz1: float, wz1: float, wwz2: float, wz2: float):
    """Along-strike and down-dip scale factors of case_input/test.drv.a6
    (getScaleCoeff): 1 inside |x| <= wwx and wwz1 <= |z| <= wwz2, falling
    linearly to 0 over wx along strike, wz1 above and wz2 below."""
    ax, az = abs(x), abs(z)
    if ax <= wwx:
        sx = 1.0
    elif ax < wwx + wx:
        sx = 1.0 - (ax - wwx) / wx
    else:
        sx = 0.0
    sz = 1.0
    if az <= wwz1 - wz1 or az >= wwz2 + wz2:
        sz = 0.0
    elif wwz1 - wz1 <= az <= wwz1:
        sz = 1.0 - (wwz1 - az) / wz1
    elif wwz2 <= az <= wwz2 + wz2:
        sz = 1.0 - (az - wwz2) / wz2
    if wwz1 <= az <= wwz2:
        sz = 1.0
    return sx, sz


def drv_a6() -> Case:
    """case_input/test.drv.a6/user_defined_params.py: a Drucker-Prager medium
    (C_elastic 0) on a 17-layer velocity profile, the reference's seed-1
    fractal rough fault (insertFaultType 2), slip-law rate and state with
    strong velocity weakening under tapered a and Vw, TPV2802 nucleation,
    500 m cells, 2 x 2 x 1 ranks.  dt, roumax and vmaxPML keep the class
    defaults (6000 m/s, 2670 kg/m3) because scripts/defaultParameters.py
    evaluates them before the case sets its layers."""
    c = Case(name="drv.a6")
    c.xmin, c.xmax = -40.0e3, 40.0e3
    c.ymin, c.ymax = -31.0e3, 30.0e3
    c.zmin, c.zmax = -40.0e3, 0.0e3
    c.fxmin, c.fxmax = -25.0e3, 25.0e3
    c.fymin, c.fymax = 0.0, 0.0
    c.fzmin, c.fzmax = -25.0e3, 0.0e3
    c.xsource, c.ysource, c.zsource = -12.0e3, 0.0, -12.0e3
    c.C_elastic = 0
    c.str1ToFaultAngle, c.devStrToStrVertRatio = 45.0, 0.33
    c.C_nuclea = 1
    c.friclaw = 4
    c.insertFaultType = 2
    c.enlarging_ratio = 1.0
    c.outputGroundMotion = 0
    c.output_plastic = 1
    c.term = 5.0
    c.dx = c.dy = c.dz = 500.0
    c.tpv = 2802
    # layers: at each 500 m cell centre down to 8 km, Vs from the power laws
    # of the case file, Vp = max(1.4 + 1.14 Vs, 1.68 Vs), rho = 2.4405 + 0.10271 Vs
    rows = []
    nz_max = round((c.zmax - c.zmin) / c.dz + 1)
    for i in range(nz_max):
        zk = (i + 0.5) * c.dz / 1.0e3
        if zk < 0.03:
            vs = 2.206 * zk ** 0.272
        elif zk < 0.19:
            vs = 3.542 * zk ** 0.407
        elif zk < 4.0:
            vs = 2.505 * zk ** 0.199
        elif zk < 8.0:
            vs = 2.927 * zk ** 0.086
        else:
            vs = 2.927 * 8.0 ** 0.086
        vp = max(1.4 + 1.14 * vs, 1.68 * vs)
        rho = 2.4405 + 0.10271 * vs
        rows.append([zk * 1e3 + 0.5 * c.dz, vp * 1e3, vs * 1e3, rho * 1e3])
        if zk > 8.0:
            break
    rows[-1][0] = 1.e6
    c.nmat, c.n2mat = len(rows), 4
    c.mat = np.array(rows)
    c.dt = 0.5 * c.dx / 6.0e3
    c.nucR = 4.0e3
    c.nucdtau0 = 50.0e6
    c.nx, c.ny, c.nz = 2, 2, 1
    c.st_on_fault = [list(s) for s in TPV1053D_ON]
    c.st_off_fault = [list(s) for s in TPV1053D_OFF]
    rough_reference_fractal(c, seed=1)
    o = c.new_ofv()
    a0, da, b, dc, r0, v0, fw, vw, dvw = 0.01, 0.01, 0.014, 0.4, 0.6, 1e-6, 0.2, 0.1, 0.9
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.18, 0.12, 0.3
            o[7, iz, ix] = -120.0e6
            o[8, iz, ix] = 40.0e6
            t1, t2 = _taper(x, z, 20.e3, 2.e3, 2.e3, 2.e3, 14.e3, 1.e3)
            dai = 2.0 * da if abs(z) > 15.e3 else da
            o[9, iz, ix] = a0 + (1. - t1 * t2) * dai
            o[10, iz, ix] = b
            o[11, iz, ix] = dc
            o[12, iz, ix] = v0
            o[13, iz, ix] = r0
            o[14, iz, ix] = fw
            o[15, iz, ix] = vw + dvw * (1. - t1 * t2)
            o[16, iz, ix] = 4.0e-4 + 1.0 * (1. - t1 * t2)
            o[17, iz, ix] = 1.0e-6
            o[18, iz, ix] = 2.7e6
            o[19, iz, ix] = 0.1e6
            o[40, iz, ix] = 0.02
            o[41, iz, ix] = 483.15
            o[42, iz, ix] = 80.0e6
            o[46, iz, ix] = 1e-16
            o[20, iz, ix] = state_steady_state(o[9, iz, ix], b, dc, v0, r0, o[8, iz, ix], o[7, iz, ix],
                                               o[46, iz, ix], c.friclaw)
    return c


def c3(steps: int = 3000, seed: int = 3, tau0: float = 10.5e6, vs_top: float = 2000.0, sn: float = -30.0e6,
       nucdtau0: float = 18.0e6) -> Case:
    """c3, the rough-fault rate-and-state box at 100 m: 240 x 120 x 120 = 3,456,000 hexahedra
    including a 10-element PML; homogeneous rock; a seeded fractal rough
    fault (rms height / length 0.005, H = 1) 16 km along strike and 9.8 km
    down dip; slip law with strong velocity weakening (friclaw 4): a seeded
    in [0.008, 0.016] (0.02 in a velocity-strengthening rim: 1 km on the
    sides and bottom, the top vs_top metres), b = 0.012, L = 0.02 m,
    f0 = 0.6, V0 = 1e-6 m/s, fw = 0.1, Vw = 0.1 m/s; sigma_n = -30 MPa
    (inside the rough-fault cap of [-40, -10] MPa), tau_0 = 10.5 MPa with
    the state that makes the initial slip rate 1e-12 m/s, TPV104 smooth
    nucleation (3 km, 1 s, 18 MPa); dt = 0.005 s.

    Stability: the reference updates the normal-stress state theta_pc
    explicitly (src/faulting.f90:35-52), which diverges where V*dt/L > 2
    (V > 8 m/s here) while the normal stress moves, as it does on a rough
    fault.  With tau_0 = 18.75 MPa and no surface layer the reference's own
    run (and this one, identically) ends in NaN at step 369; the stress
    level and the 2 km velocity-strengthening top above keep the whole
    3000 steps finite."""
    c = _strike_slip_box(100.0, 10, "c3")
    c.xmin, c.xmax = -11.0e3, 11.0e3
    c.ymin, c.ymax = -5.0e3, 5.0e3
    c.zmin, c.zmax = -11.0e3, 0.0
    c.nuni_y_minus = c.nuni_y_plus = 49
    c.vp, c.vs, c.rou = 6000.0, 3464.0, 2670.0
    c.roumax, c.vmaxPML = 2670.0, 6000.0
    c.dt = 0.005
    c.term = steps * c.dt
    c.friclaw = 4
    c.tpv = 104
    c.C_nuclea = 1
    c.insertFaultType = 1
    c.xsource, c.ysource, c.zsource = 0.0, 0.0, -6.0e3
    c.nucR, c.nucT, c.nucdtau0 = 3.0e3, 1.0, nucdtau0
    rough_fractal(c, 0.005, 1.0, seed)
    rng = np.random.default_rng(seed + 1000)
    o = c.new_ofv()
    b, L, f0, v0, fw, vw, vini = 0.012, 0.02, 0.6, 1.0e-6, 0.1, 0.1, 1.0e-12
    a_field = 0.008 + 0.008 * rng.random((c.nfz, c.nfx))
    for ix, x in enumerate(c.fx):
        for iz, z in enumerate(c.fz):
            a = a_field[iz, ix]
            if abs(x) > c.fxmax - 1.0e3 - 0.01 or z < c.fzmin + 1.0e3 + 0.01 or z > -vs_top - 0.01:
                a = 0.02
            o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.6, 0.45, 0.3
            o[7, iz, ix] = sn
            o[8, iz, ix] = tau0
            o[9, iz, ix], o[10, iz, ix], o[11, iz, ix] = a, b, L
            o[12, iz, ix], o[13, iz, ix], o[14, iz, ix], o[15, iz, ix] = v0, f0, fw, vw
            o[46, iz, ix] = vini
            o[20, iz, ix] = state_steady_state(a, b, L, v0, f0, tau0, sn, vini, 4)
    return c


def von_karman_field(nx: int, nz: int, spacing: float, corr_len: float, hurst: float, seed: int) -> np.ndarray:
    """Seeded 2-D von Karman random field on an (nz, nx) grid: a complex
    Gaussian spectrum shaped to the von Karman power spectrum
    P(k) ~ (1 + k^2 a^2)^-(H+1) (correlation length a, Hurst exponent H),
    the mean removed, transformed back on a grid padded to twice the size (so
    the field does not wrap around), normalised to zero mean and unit standard
    deviation."""
    rng = np.random.default_rng(seed)
    px = 1 << int(math.ceil(math.log2(2 * nx)))
    pz = 1 << int(math.ceil(math.log2(2 * nz)))
    kx = 2.0 * math.pi * np.fft.fftfreq(px, d=spacing)
    kz = 2.0 * math.pi * np.fft.fftfreq(pz, d=spacing)
    ksq = kz[:, None] ** 2 + kx[None, :] ** 2
    power = (1.0 + ksq * corr_len ** 2) ** -(hurst + 1.0)
    power[0, 0] = 0.0
    spec = (rng.standard_normal((pz, px)) + 1j * rng.standard_normal((pz, px))) * np.sqrt(0.5 * power)
    w = np.real(np.fft.ifft2(spec))[:nz, :nx]
    return (w - w.mean()) / w.std()


def c5_scenario(seed: int, steps: int = 200, index: int = 0) -> Case:
    """One scenario of the c5 ensemble, drawn from `seed` at 100 m with a
    10-element PML: a box 8-24 km along strike, 6-12 km wide and 6-12 km deep
    (80-240 x 60-120 x 60-120 hexahedra including the PML), a seeded
    five-layer medium (Vs 2000 -> 3464 m/s), a vertical strike-slip fault
    reaching the surface with 0.5 km of rock beyond its ends and bottom, and
    either slip weakening (mu_s 0.677, mu_d 0.45, Dc 0.3 m, sigma_n -100 MPa,
    mean tau_0 62 MPa, a 3 x 3 km patch at 75 MPa) or slip-law
    rate-and-state friction (a in [0.008, 0.016] with a 0.5 km
    velocity-strengthening rim, b 0.012, L 0.02 m, sigma_n -30 MPa, mean tau_0 10.5 MPa, TPV104
    nucleation).  tau_0 carries a von Karman field (correlation length 2 km,
    H 0.75) of 10 % standard deviation, capped below the static strength
    outside the nucleation zone; the hypocentre is drawn on the fault;
    dt 0.005 s."""
    rng = np.random.default_rng(seed)
    d, npml = 100.0, 10
    nxe = 2 * int(rng.integers(40, 121))   # 80..240 elements along strike
    nye = 2 * int(rng.integers(30, 61))    # 60..120 across
    nze = int(rng.integers(60, 121))       # 60..120 in depth
    rsf = bool(rng.random() < 0.5)
    c = Case(name=f"c5_{index}")
    c.dx = c.dy = c.dz = d
    c.nPML, c.R = npml, 0.001
    c.enlarging_ratio = 1.0
    c.xmax = (nxe // 2 - npml) * d
    c.xmin = -c.xmax
    half_y = nye // 2 - npml - 1          # uniform cells per side (one growth cell follows)
    c.nuni_y_minus = c.nuni_y_plus = half_y
    c.ymax = (half_y + 1) * d
    c.ymin = -c.ymax
    c.zmin, c.zmax = -(nze - npml) * d, 0.0
    c.fxmax = c.xmax - 0.5e3
    c.fxmin = -c.fxmax
    c.fymin = c.fymax = 0.0
    c.fzmin, c.fzmax = c.zmin + 0.5e3, 0.0
    c.nmat, c.n2mat = 5, 4
    c.mat = layered_medium(seed + 7, max_depth=-c.zmin)
    c.vp, c.vs, c.rou = c.mat[-1, 1], c.mat[-1, 2], c.mat[-1, 3]
    c.roumax, c.vmaxPML = float(c.mat[:, 3].max()), float(c.mat[:, 1].max())
    c.dt = 0.005
    c.term = steps * c.dt
    c.nucR = 1.5e3
    # hypocentre on the fault, the nucleation zone inside the rupturable part
    c.xsource = float(np.round(rng.uniform(c.fxmin + 1.5e3, c.fxmax - 1.5e3), -2))
    zlo = c.fzmin + 1.5e3
    c.zsource = float(np.round(rng.uniform(zlo, max(zlo, -2.5e3)), -2))
    c.ysource = 0.0
    field = von_karman_field(c.nfx, c.nfz, d, 2.0e3, 0.75, seed + 13)
    o = c.new_ofv()
    if rsf:
        c.friclaw, c.tpv, c.C_nuclea = 4, 104, 1
        c.nucR, c.nucT, c.nucdtau0 = 1.5e3, 1.0, 18.0e6
        b, L, f0, v0, fw, vw, vini, sn, tau_mean = 0.012, 0.02, 0.6, 1.0e-6, 0.1, 0.1, 1.0e-12, -30.0e6, 10.5e6
        a_field = 0.008 + 0.008 * rng.random((c.nfz, c.nfx))
        for ix, x in enumerate(c.fx):
            for iz, z in enumerate(c.fz):
                a = a_field[iz, ix]
                if abs(x) > c.fxmax - 0.5e3 - 0.01 or z < c.fzmin + 0.5e3 + 0.01 or z > -0.5e3 - 0.01:
                    a = 0.02
                tau = tau_mean * (1.0 + 0.1 * field[iz, ix])
                o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.6, 0.45, 0.3
                o[7, iz, ix], o[8, iz, ix] = sn, tau
                o[9, iz, ix], o[10, iz, ix], o[11, iz, ix] = a, b, L
                o[12, iz, ix], o[13, iz, ix], o[14, iz, ix], o[15, iz, ix] = v0, f0, fw, vw
                o[46, iz, ix] = vini
                o[20, iz, ix] = state_steady_state(a, b, L, v0, f0, tau, sn, vini, 4)
    else:
        c.friclaw, c.tpv, c.C_nuclea = 1, 8, 0
        mus, mud, sn, tau_mean = 0.677, 0.45, -100.0e6, 62.0e6
        for ix, x in enumerate(c.fx):
            for iz, z in enumerate(c.fz):
                o[1, iz, ix] = mus
                if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01 or z > -0.5e3 + 0.01:
                    o[1, iz, ix] = 1.0e4
                o[2, iz, ix], o[3, iz, ix], o[7, iz, ix] = mud, 0.3, sn
                tau = min(tau_mean * (1.0 + 0.1 * field[iz, ix]), 0.98 * mus * -sn)
                if abs(x - c.xsource) <= 1.5e3 and abs(z - c.zsource) <= 1.5e3:
                    tau = 75.0e6
                o[8, iz, ix] = tau
    fx_km = [round(v / 1.0e3, 1) for v in (c.fxmin / 2.0, c.fxmax / 2.0)]
    fz_km = [round(v / 1.0e3, 1) for v in (c.fzmin / 2.0, c.fzmin / 4.0)]
    c.st_on_fault = [[x, z] for x in fx_km for z in fz_km]
    yk = round(c.ymax / 2.0e3, 1)
    c.st_off_fault = [[x, y, 0.0] for x in fx_km for y in (-yk, yk)]
    return c


def box_c5(seed: int, rsf: bool, steps: int = 60, name: str = "box_c5") -> Case:
    """A c5 scenario drawn at fixture size: the same kinds of draw as
    c5_scenario (box dimensions, a seeded five-layer medium, tau_0 carrying a
    von Karman field of 10 % standard deviation, a hypocentre drawn on the
    fault, slip weakening or slip-law rate-and-state friction with a seeded a
    field and a velocity-strengthening rim) on a box of 16-24 x 12-16 x 9-12
    hexahedra at 100 m with a 3-element PML, correlation length 300 m, and the
    stress levels of box_sw / box_rsf so that the rupture starts within the
    fixture's steps."""
    rng = np.random.default_rng(seed)
    d, npml = 100.0, 3
    nxe = 2 * int(rng.integers(8, 13))    # 16..24 elements along strike
    nye = 2 * int(rng.integers(6, 9))     # 12..16 across
    nze = int(rng.integers(9, 13))        # 9..12 in depth
    c = Case(name=name)
    c.dx = c.dy = c.dz = d
    c.nPML, c.R = npml, 0.01
    c.enlarging_ratio = 1.0
    c.xmax = (nxe // 2 - npml) * d
    c.xmin = -c.xmax
    half_y = nye // 2 - npml - 1
    c.nuni_y_minus = c.nuni_y_plus = half_y
    c.ymax = (half_y + 1) * d
    c.ymin = -c.ymax
    c.zmin, c.zmax = -(nze - npml) * d, 0.0
    c.fxmax = c.xmax - d
    c.fxmin = -c.fxmax
    c.fymin = c.fymax = 0.0
    c.fzmin, c.fzmax = c.zmin + d, 0.0
    c.nmat, c.n2mat = 5, 4
    c.mat = layered_medium(seed + 7, max_depth=-c.zmin)
    c.vp, c.vs, c.rou = c.mat[-1, 1], c.mat[-1, 2], c.mat[-1, 3]
    c.roumax, c.vmaxPML = float(c.mat[:, 3].max()), float(c.mat[:, 1].max())
    c.dt = 0.25 * d / c.vmaxPML
    c.term = steps * c.dt
    c.xsource = float(np.round(rng.uniform(c.fxmin + 2 * d, c.fxmax - 2 * d), -2))
    c.zsource = float(np.round(rng.uniform(c.fzmin + 2 * d, -2 * d), -2))
    c.ysource = 0.0
    field = von_karman_field(c.nfx, c.nfz, d, 300.0, 0.75, seed + 13)
    o = c.new_ofv()
    if rsf:
        c.friclaw, c.tpv, c.C_nuclea = 4, 104, 1
        c.nucR, c.nucT, c.nucdtau0 = 3.0 * d, 10.0 * c.dt, 40.0e6
        b, L, f0, v0, fw, vw, vini, sn, tau_mean = 0.012, 0.02, 0.6, 1.0e-6, 0.1, 0.1, 1.0e-12, -60.0e6, 24.0e6
        a_field = 0.008 + 0.008 * rng.random((c.nfz, c.nfx))
        for ix, x in enumerate(c.fx):
            for iz, z in enumerate(c.fz):
                a = a_field[iz, ix]
                if abs(x) > c.fxmax - 0.01 or z < c.fzmin + 0.01:
                    a = 0.02
                tau = tau_mean * (1.0 + 0.1 * field[iz, ix])
                o[1, iz, ix], o[2, iz, ix], o[3, iz, ix] = 0.6, 0.45, 0.3
                o[7, iz, ix], o[8, iz, ix] = sn, tau
                o[9, iz, ix], o[10, iz, ix], o[11, iz, ix] = a, b, L
                o[12, iz, ix], o[13, iz, ix], o[14, iz, ix], o[15, iz, ix] = v0, f0, fw, vw
                o[46, iz, ix] = vini
                o[20, iz, ix] = state_steady_state(a, b, L, v0, f0, tau, sn, vini, 4)
    else:
        c.friclaw, c.tpv, c.C_nuclea = 1, 8, 0
        mus, mud, sn, tau_mean = 0.60, 0.45, -100.0e6, 55.0e6
        for ix, x in enumerate(c.fx):
            for iz, z in enumerate(c.fz):
                o[1, iz, ix] = mus
                if abs(abs(x) - c.fxmax) < 0.01 or abs(z - c.fzmin) < 0.01:
                    o[1, iz, ix] = 1.0e4
                o[2, iz, ix], o[3, iz, ix], o[7, iz, ix] = mud, 0.02, sn
                tau = min(tau_mean * (1.0 + 0.1 * field[iz, ix]), 0.98 * mus * -sn)
                if abs(x - c.xsource) <= 1.5 * d and abs(z - c.zsource) <= 1.5 * d:
                    tau = 62.0e6
                o[8, iz, ix] = tau
    c.st_on_fault = [[c.xsource / 1e3, c.zsource / 1e3], [0.0, 0.0]]
    c.st_off_fault = [[0.0, 2 * d / 1e3, 0.0], [c.xsource / 1e3, -2 * d / 1e3, 0.0]]
    return c


def c5(seed: int = 5, n: int = 64, steps: int = 200) -> List[Case]:
    """c5, the batched ensemble: n scenarios of uneven size and physics
    (c5_scenario) drawn from one seed."""
    rng = np.random.default_rng(seed)
    return [c5_scenario(int(s), steps, i) for i, s in enumerate(rng.integers(0, 2**31 - 1, size=n))]
