"""Helpers shared by the tests: golden fixtures and error norms."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = ["box_sw", "box_tw", "box_layered", "box_rsf3", "box_rsf4", "box_q", "box_plastic", "box_hg2",
         "box_dip", "box_rough", "box_tp", "box_c5sw", "box_c5rsf", "box_tpv2802", "box_tpv202"]

# fric slots the time step writes (1-based, src/faulting.f90)
FRIC_SLOTS = [20, 23, 24, 31, 32, 33, 34, 35, 36, 47, 48, 71, 72, 73, 74, 75, 76, 77, 78, 79, 80]


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    bundle = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    sc = {str(k): float(v) for k, v in zip(z["scalar_names"], z["scalar_values"])}
    return bundle, sc, out


def rel_inf(a, ref):
    """max|a-ref| / max|ref| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.max(np.abs(a - ref)) if a.size else 0.0
    s = np.max(np.abs(ref)) if ref.size else 0.0
    if s == 0.0:
        return float(d)
    return float(d / s)


# Error scale of each channel: the infinity norm of its physical quantity
# (all slip components share one scale, all tractions another, ...), so a
# component that is zero up to rounding (e.g. fault opening) is measured
# against the size of that quantity, not against its own rounding noise.
GROUPS = {
    "fric20": "state", "fric23": "traction", "fric24": "theta_pc_dot",
    "fric31": "vel", "fric32": "vel", "fric33": "vel", "fric34": "vel", "fric35": "vel", "fric36": "vel",
    "fric47": "rate", "fric48": "traction",
    "fric71": "slip", "fric72": "slip", "fric73": "slip", "fric74": "rate", "fric75": "rate", "fric76": "rate",
    "fric77": "slip", "fric78": "traction", "fric79": "traction", "fric80": "traction",
    "fnft": "fnft",
    "onfault_time": "time", "onfault_sliprate_s": "st_rate", "onfault_sliprate_d": "st_rate",
    "onfault_state": "state", "onfault_slip_s": "st_slip", "onfault_slip_d": "st_slip", "onfault_slip_n": "st_slip",
    "onfault_tau_s": "st_traction", "onfault_tau_d": "st_traction", "onfault_sigma_n": "st_traction",
    "onfault_pore_pressure": "st_pressure", "onfault_temperature": "st_temperature",
    "station_series": "station", "velArr": "velArr", "dispArr": "dispArr", "stressArr": "stressArr",
    "pstrain": "pstrain",
}


# fric slots the reference leaves undefined: with thermal pressurization
# (friclaw 5) its NewtonRaphson returns the local thetaPcTmp, which it assigns
# only for friclaw < 5 (src/faulting.f90:459-516), into fric(23), and fric(24)
# is computed from that value the next step
UNDEFINED_SLOTS = {5: (23, 24)}


def undefined_slots(sc):
    return UNDEFINED_SLOTS.get(int(sc["friclaw"]), ())


def channels(out, skip=()):
    """Named time series / state channels used for parity reports (fric
    slots in ``skip`` left out)."""
    ch = {}
    fr = out["fric"]
    for k in FRIC_SLOTS:
        if k in skip:
            continue
        ch[f"fric{k}"] = fr[:, k - 1]
    ch["fnft"] = np.where(out["fnft"] > 5000.0, 0.0, out["fnft"])
    h = out["onFaultQuantHistSCECForm"]
    names = ["time", "sliprate_s", "sliprate_d", "state", "slip_s", "slip_d", "slip_n", "tau_s", "tau_d", "sigma_n",
             "pore_pressure", "temperature"]
    for j, nm in enumerate(names):
        ch["onfault_" + nm] = h[:, :, j]
    if "OffFaultStGramSCEC" in out:
        ch["station_series"] = out["OffFaultStGramSCEC"]
    ch["velArr"] = out["velArr"]
    ch["dispArr"] = out["dispArr"]
    ch["stressArr"] = out["stressArr"]
    ch["pstrain"] = out["pstrain"]
    return ch


def channel_errors(ch, ch_ref):
    """{channel: max|x - x_ref| / (infinity norm of the channel's quantity in x_ref)}."""
    scale = {}
    for k, v in ch_ref.items():
        g = GROUPS[k]
        m = float(np.max(np.abs(v))) if np.size(v) else 0.0
        scale[g] = max(scale.get(g, 0.0), m)
    err = {}
    for k, v in ch_ref.items():
        d = float(np.max(np.abs(np.asarray(ch[k], dtype=np.float64) - v))) if np.size(v) else 0.0
        sc = scale[GROUPS[k]]
        err[k] = d / sc if sc > 0 else d
    return err
