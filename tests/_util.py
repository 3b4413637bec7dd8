"""Shared helpers of the test suite: building matched oracle / device configurations."""
from __future__ import annotations

import numpy as np

from langmuir_amd.host import SimulationParameters

# oracle field -> SimulationParameters field conversion
_BOOL = ("coulomb_carriers", "source_metropolis", "source_coulomb", "balance_charges")


def params_pair(**kw):
    """(oracle OrcParams, SimulationParameters) with identical values."""
    from oracle import pyoracle as O
    op = O.OrcParams.make(**kw)
    sp = SimulationParameters()
    for name, _ in O.OrcParams._fields_:
        if name in ("reserved0",):
            continue
        v = getattr(op, name)
        if name == "simulation_type":
            sp.simulation_type = "solarcell" if v == 1 else "transistor"
        elif name in _BOOL:
            setattr(sp, name, bool(v))
        else:
            setattr(sp, name, type(getattr(sp, name))(v))
    return op, sp


def oracle_state(w):
    e, h = w.carriers(0), w.carriers(1)
    win, k = w.rng()
    a, s, _ = w.flux()
    return dict(electrons=e["site"], holes=h["site"], e_lifetime=e["lifetime"],
                e_pathlength=e["pathlength"], h_lifetime=h["lifetime"], h_pathlength=h["pathlength"],
                defects=w.defects(), potential=w.potential(), rng_window=win, rng_consumed=k,
                flux_attempts=a, flux_successes=s, current_step=w.current_step)


def gpu_from_oracle(L, w, sp, context=None, de_path="auto", de_chunks=0):
    st = oracle_state(w)
    return L.Simulation(sp, context=context, de_path=de_path, de_chunks=de_chunks, **st)


def resync(sim, w):
    st = oracle_state(w)
    sim.load(st["electrons"], st["holes"], st["defects"], st["potential"], st["rng_window"],
             st["rng_consumed"], st["e_lifetime"], st["e_pathlength"], st["h_lifetime"],
             st["h_pathlength"], st["flux_attempts"], st["flux_successes"], st["current_step"])


def compare_state(gs, w):
    """Field-by-field equality of a device state dict and the oracle world."""
    e, h = w.carriers(0), w.carriers(1)
    win, k = w.rng()
    a, s, _ = w.flux()
    diffs = []
    for name, g, o in (("e_site", gs["e_site"], e["site"]), ("h_site", gs["h_site"], h["site"]),
                       ("e_lifetime", gs["e_lifetime"], e["lifetime"]),
                       ("h_lifetime", gs["h_lifetime"], h["lifetime"]),
                       ("e_pathlength", gs["e_pathlength"], e["pathlength"]),
                       ("h_pathlength", gs["h_pathlength"], h["pathlength"]),
                       ("flux_attempts", gs["flux_attempts"], a),
                       ("flux_successes", gs["flux_successes"], s)):
        if len(g) != len(o) or not np.array_equal(np.asarray(g), np.asarray(o)):
            diffs.append(name)
    if gs["rng_consumed"] != k or not np.array_equal(gs["rng_window"], win):
        diffs.append("rng")
    if gs["current_step"] != w.current_step:
        diffs.append("step")
    return diffs


def de_tolerance(de_ref):
    """Stated dE tolerance against the FP64 oracle: 1e-5 relative + 2e-7 eV."""
    return 1e-5 * np.abs(de_ref) + 2e-7
