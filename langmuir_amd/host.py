"""Host-side mirror of the reference's hot-path interfaces over the C-ABI.

* :class:`Context` -- one device context (``lmg_create``); stateless pair sums
  and the full-lattice map.
* :class:`GpuHelper` -- the surface of ``LangmuirCore::OpenClHelper``
  (src/langmuirCore/include/openclhelper.h:22-101): ``launch_coulomb_kernel2``
  + ``get_output_host(id)`` / ``get_output_host_future(id)``, and
  ``launch_coulomb_kernel1`` for the map.
* :class:`Simulation` -- ``Simulation::performIterations`` on the device
  (src/langmuirCore/simulation.cpp:29-111) for a loaded configuration.
* :class:`SimulationParameters` and helpers that derive, from the reference's
  input keys, the constants and tables the device needs (the parts of
  ``World::initialize`` that involve no random draws).

Everything runs through ``liblangmuir_gpu.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import math
from typing import Dict, Iterable, Optional, Sequence

import numpy as np

from . import _abi
from ._abi import LmgError, LmgParams, LmgSimConfig, LmgSimState

FLUX = {n: i for i, n in enumerate(_abi.FLUX_NAMES)}
N_SPECIAL = 35


# ---------------------------------------------------------------------------
# parameters (src/langmuirCore/include/parameters.h:46-390, keyvalueparser.cpp:15-105)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class SimulationParameters:
    simulation_type: str = "transistor"
    random_seed: int = 0
    grid_x: int = 128
    grid_y: int = 128
    grid_z: int = 1
    coulomb_carriers: bool = False
    coulomb_gaussian_sigma: float = 0.0
    defects_charge: int = 0
    electron_percentage: float = 0.01
    hole_percentage: float = 0.0
    seed_charges: float = 0.0
    defect_percentage: float = 0.0
    trap_percentage: float = 0.0
    trap_potential: float = 0.10
    gaussian_stdev: float = 0.0
    seed_percentage: float = 1.0
    voltage_right: float = 0.0
    voltage_left: float = 0.0
    exciton_binding: float = 0.0
    temperature_kelvin: float = 300.0
    source_rate: float = 0.90
    e_source_l_rate: float = -1.0
    e_source_r_rate: float = -1.0
    h_source_l_rate: float = -1.0
    h_source_r_rate: float = -1.0
    generation_rate: float = 0.001
    balance_charges: bool = False
    drain_rate: float = 0.90
    e_drain_l_rate: float = -1.0
    e_drain_r_rate: float = -1.0
    h_drain_l_rate: float = -1.0
    h_drain_r_rate: float = -1.0
    boltzmann_constant: float = 1.3806504e-23
    dielectric_constant: float = 3.5
    elementary_charge: float = 1.60217646e-19
    permittivity_space: float = 8.854187817e-12
    grid_factor: float = 1e-9
    electrostatic_cutoff: int = 50
    hopping_range: int = 1
    slope_z: float = 0.0
    source_metropolis: bool = False
    source_coulomb: bool = False
    recombination_rate: float = 0.0
    recombination_range: int = 0
    source_scale_area: float = 65536.0

    # input-file key -> field (keyvalueparser.cpp:15-105)
    KEYS = {
        "simulation.type": "simulation_type", "random.seed": "random_seed", "grid.x": "grid_x",
        "grid.y": "grid_y", "grid.z": "grid_z", "coulomb.carriers": "coulomb_carriers",
        "coulomb.gaussian.sigma": "coulomb_gaussian_sigma", "defects.charge": "defects_charge",
        "electron.percentage": "electron_percentage", "hole.percentage": "hole_percentage",
        "seed.charges": "seed_charges", "defect.percentage": "defect_percentage",
        "trap.percentage": "trap_percentage", "trap.potential": "trap_potential",
        "gaussian.stdev": "gaussian_stdev", "seed.percentage": "seed_percentage",
        "voltage.right": "voltage_right", "voltage.left": "voltage_left",
        "exciton.binding": "exciton_binding", "temperature.kelvin": "temperature_kelvin",
        "source.rate": "source_rate", "e.source.l.rate": "e_source_l_rate",
        "e.source.r.rate": "e_source_r_rate", "h.source.l.rate": "h_source_l_rate",
        "h.source.r.rate": "h_source_r_rate", "generation.rate": "generation_rate",
        "balance.charges": "balance_charges", "drain.rate": "drain_rate",
        "e.drain.l.rate": "e_drain_l_rate", "e.drain.r.rate": "e_drain_r_rate",
        "h.drain.l.rate": "h_drain_l_rate", "h.drain.r.rate": "h_drain_r_rate",
        "hopping.range": "hopping_range", "slope.z": "slope_z",
        "source.metropolis": "source_metropolis", "source.coulomb": "source_coulomb",
        "recombination.rate": "recombination_rate", "recombination.range": "recombination_range",
        "source.scale.area": "source_scale_area",
        # read-only ("Constant") in the reference's parser (variable.h:267-276); the
        # programmatic World(SimulationParameters&) path can still set it (world.cpp:49-76)
        "electrostatic.cutoff": "electrostatic_cutoff",
    }

    @classmethod
    def from_keys(cls, keys: Dict[str, object]) -> "SimulationParameters":
        p = cls()
        for k, v in keys.items():
            name = cls.KEYS.get(k.strip().lower())
            if name is None:
                raise KeyError(f"unknown key {k!r}")
            ftype = type(getattr(p, name))
            if ftype is bool and isinstance(v, str):
                v = v.strip().lower() in ("1", "true", "yes", "on")
            setattr(p, name, ftype(v))
        return p

    # parameters.h:386-387
    @property
    def prefactor(self) -> float:
        return self.elementary_charge / (4.0 * math.pi * self.dielectric_constant *
                                         self.permittivity_space * self.grid_factor)

    @property
    def inverse_kt(self) -> float:
        return self.elementary_charge / (self.boltzmann_constant * self.temperature_kelvin)

    @property
    def volume(self) -> int:
        return self.grid_x * self.grid_y * self.grid_z

    @property
    def is_solarcell(self) -> bool:
        return self.simulation_type == "solarcell"

    def max_electrons(self) -> int:          # world.cpp:597-598 (int truncation)
        return int(self.electron_percentage * float(self.volume))

    def max_holes(self) -> int:              # world.cpp:594-595
        return int(self.hole_percentage * float(self.volume))

    def check(self) -> None:
        """The checks of parameters.h:393-646 that concern this path."""
        if self.simulation_type not in ("transistor", "solarcell"):
            raise ValueError("simulation.type must be transistor or solarcell")
        if min(self.grid_x, self.grid_y, self.grid_z) < 1:
            raise ValueError("grid sizes must be >= 1")
        if self.hopping_range not in (1, 2):
            raise ValueError("hopping.range must be 1 or 2")
        if not self.source_metropolis and self.source_coulomb:
            raise ValueError("source.metropolis = false, yet source.coulomb = true")
        if self.defects_charge != 0 and not self.coulomb_carriers:
            raise ValueError("defects.charge != 0 && coulomb.carriers = false")
        if self.recombination_rate > 0 and not self.is_solarcell:
            raise ValueError("recombination.rate > 0, yet simulation.type != solarcell")
        if self.balance_charges and not self.is_solarcell:
            raise ValueError("balance.charges == true, yet simulation.type != solarcell")
        if self.coulomb_gaussian_sigma > 0 and 1 in (self.grid_x, self.grid_y, self.grid_z):
            raise ValueError("coulomb.gaussian.sigma > 0, yet the simulation is not 3D")


def _smart(rate: float, dflt: float) -> float:   # FluxAgent::setRateSmartly (fluxagent.cpp:156-169)
    if rate < 0:
        if dflt >= 0:
            return dflt
        raise ValueError("negative rate")
    return rate


def flux_rates(p: SimulationParameters) -> np.ndarray:
    """m_probability of the 10 flux agents (world.cpp:997-1028, 1124-1145)."""
    r = np.zeros(10)
    if p.simulation_type == "transistor":
        r[FLUX["xS"]] = 0.0
        r[FLUX["eSL"]] = _smart(p.e_source_l_rate, p.source_rate)
        r[FLUX["eSR"]] = _smart(p.e_source_r_rate, 0.0)
        r[FLUX["hSR"]] = _smart(p.h_source_r_rate, 0.0)
        r[FLUX["hSL"]] = _smart(p.h_source_l_rate, 0.0)
        r[FLUX["xD"]] = 0.0
        r[FLUX["eDR"]] = _smart(p.e_drain_r_rate, p.drain_rate)
        r[FLUX["eDL"]] = _smart(p.e_drain_l_rate, 0.0)
        r[FLUX["hDR"]] = _smart(p.h_drain_r_rate, 0.0)
        r[FLUX["hDL"]] = _smart(p.h_drain_l_rate, 0.0)
    else:
        r[FLUX["xS"]] = _smart(p.generation_rate, p.source_rate)
        r[FLUX["eSR"]] = _smart(p.e_source_r_rate, 0.0)
        r[FLUX["eSL"]] = _smart(p.e_source_l_rate, 0.0)
        r[FLUX["hSR"]] = _smart(p.h_source_r_rate, 0.0)
        r[FLUX["hSL"]] = _smart(p.h_source_l_rate, 0.0)
        xy = p.grid_x * p.grid_y
        if p.source_scale_area > 0 and p.source_scale_area != xy:
            r[FLUX["xS"]] = r[FLUX["xS"]] * (xy / float(p.source_scale_area))
        r[FLUX["xD"]] = p.recombination_rate
        for k, v in (("eDR", p.e_drain_r_rate), ("eDL", p.e_drain_l_rate),
                     ("hDR", p.h_drain_r_rate), ("hDL", p.h_drain_l_rate)):
            r[FLUX[k]] = _smart(v, p.drain_rate)
    return r


def drain_slots(p: SimulationParameters):
    """Special-slot ids of the face drains.  World::createSources/createDrains
    register, per grid and in order (world.cpp:961-1122, cubicgrid.cpp:558-589):
    electron grid eSL V+0, eSR V+1, xS V+2, eDL V+3, eDR V+4, xD V+5;
    hole grid hSL V+0, hSR V+1, hDL V+2, hDR V+3."""
    V = p.volume
    return (V + 3, V + 4), (V + 2, V + 3)


def coupling_table(hopping_range: int) -> np.ndarray:
    """Potential::updateCouplingConstants (potential.cpp:326-349), padded to 27."""
    h = hopping_range + 1
    t = np.zeros((h, h, h))
    for dx in range(h):
        for dy in range(h):
            for dz in range(h):
                r = math.sqrt(dx * dx + dy * dy + dz * dz)
                t[dx, dy, dz] = (1.0 / 3.0) * math.pow(3.0, r - 1) * math.pow(1.0 / 27.0, r - 1)
    t[0, 0, 0] = 0.0
    out = np.zeros(27)
    out[: h ** 3] = t.ravel()
    return out


def background_potential(p: SimulationParameters, traps: Sequence[int] = (),
                         trap_potentials: Optional[Sequence[float]] = None) -> np.ndarray:
    """setPotentialZero + setPotentialLinear + setPotentialGate + the trap adds of
    setPotentialTraps for an explicit trap list (potential.cpp:17-72, 224-244)."""
    X, Y, Z = p.grid_x, p.grid_y, p.grid_z
    v = np.zeros((Z, Y, X))
    m = (p.voltage_right - p.voltage_left) / float(X)
    v += (m * (np.arange(X) + 0.5) + p.voltage_left)[None, None, :]
    if not (Z == 1 or p.slope_z == 0):
        v += (p.slope_z * (np.arange(Z) + 0.5))[:, None, None]
    v = v.ravel()
    traps = np.asarray(traps, dtype=np.int64)
    if len(traps):
        tp = (np.full(len(traps), p.trap_potential) if trap_potentials is None or len(trap_potentials) == 0
              else np.asarray(trap_potentials, dtype=np.float64))
        np.add.at(v, traps, tp)
    return v


def mt19937_seed_window(seed: int) -> np.ndarray:
    """boost::mt19937(seed) initial state x[0..623] (rand.cpp:17-19)."""
    w = np.zeros(624, dtype=np.uint64)
    w[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        prev = int(w[i - 1])
        w[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF
    return w.astype(np.uint32)


# ---------------------------------------------------------------------------
# device context
# ---------------------------------------------------------------------------
def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def device_count() -> int:
    return int(_abi.lib().lmg_device_count())


class Context:
    """One device context (lmg_create).  Fails loudly without a GPU."""

    def __init__(self, dims, cutoff=50, hopping_range=1, prefactor=None, inverse_kt=None,
                 gaussian_sigma=0.0, exciton_binding=0.0, defects_charge=0, device_id=0):
        L = _abi.lib()
        X, Y, Z = (int(d) for d in dims)
        self.dims = (X, Y, Z)
        self.V = X * Y * Z
        base = SimulationParameters()
        self.params = LmgParams(
            grid_x=X, grid_y=Y, grid_z=Z, cutoff=int(cutoff), hopping_range=int(hopping_range),
            defects_charge=int(defects_charge), device_id=int(device_id), reserved=0,
            prefactor=base.prefactor if prefactor is None else float(prefactor),
            inverse_kt=base.inverse_kt if inverse_kt is None else float(inverse_kt),
            gaussian_sigma=float(gaussian_sigma), exciton_binding=float(exciton_binding))
        h = C.c_void_p()
        rc = L.lmg_create(C.byref(self.params), C.byref(h))
        if rc != 0:
            raise LmgError(rc, L.lmg_last_error(None).decode())
        self._h = h

    @classmethod
    def from_parameters(cls, p: SimulationParameters, device_id=0) -> "Context":
        return cls((p.grid_x, p.grid_y, p.grid_z), cutoff=p.electrostatic_cutoff,
                   hopping_range=p.hopping_range, prefactor=p.prefactor, inverse_kt=p.inverse_kt,
                   gaussian_sigma=p.coulomb_gaussian_sigma, exciton_binding=p.exciton_binding,
                   defects_charge=p.defects_charge, device_id=device_id)

    def _chk(self, rc):
        if rc != 0:
            raise LmgError(rc, _abi.lib().lmg_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib().lmg_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def sync(self):
        self._chk(_abi.lib().lmg_sync(self._h))

    @property
    def stream(self) -> int:
        return int(_abi.lib().lmg_stream(self._h) or 0)

    def pair_potential(self, charge_site, charge_q, target_site, flags=0) -> np.ndarray:
        cs, cq, ts = _i32(charge_site), _i32(charge_q), _i32(target_site)
        out = np.empty(len(ts), np.float64)
        self._chk(_abi.lib().lmg_pair_potential(self._h, _p(cs), _p(cq), len(cs), _p(ts), len(ts),
                                                int(flags), _p(out)))
        return out

    def potential_map(self, charge_site, charge_q, flags=0) -> np.ndarray:
        cs, cq = _i32(charge_site), _i32(charge_q)
        out = np.empty(self.V, np.float32)
        self._chk(_abi.lib().lmg_potential_map(self._h, _p(cs), _p(cq), len(cs), int(flags), _p(out)))
        return out

    def pair_potential_dev(self, charge_site_ptr, charge_q_ptr, n_charges, target_ptr, n_targets,
                           out_ptr, flags=0):
        self._chk(_abi.lib().lmg_pair_potential_dev(self._h, charge_site_ptr, charge_q_ptr, n_charges,
                                                    target_ptr, n_targets, int(flags), out_ptr))

    def potential_map_dev(self, charge_site_ptr, charge_q_ptr, n_charges, out_ptr, flags=0):
        self._chk(_abi.lib().lmg_potential_map_dev(self._h, charge_site_ptr, charge_q_ptr, n_charges,
                                                   int(flags), out_ptr))


class GpuHelper:
    """Surface of LangmuirCore::OpenClHelper (openclhelper.h:22-101) over a Context.

    The staging contract of launchCoulombKernel2 (openclhelper.cpp:372-425) is kept:
    charges are [electrons, holes, defects (if charged)] at their current sites;
    targets are the same list followed by the future sites of the electrons and the
    holes; ``get_output_host(id)`` / ``get_output_host_future(id)`` index it with
    the carriers' openCLID (electron i -> i, hole i -> Ne + i) and offset
    n = Ne + Nh + Nd (openclhelper.cpp:567-583).
    """

    def __init__(self, ctx: Context, gauss: bool = False):
        self.ctx = ctx
        self.gauss = gauss
        self._o = np.zeros(0)
        self._offset = 0
        self._map = None

    def launch_coulomb_kernel2(self, e_site, e_fsite, h_site, h_fsite, defect_site=(),
                               defects_charge=0):
        e_site, h_site = _i32(e_site), _i32(h_site)
        d = _i32(defect_site) if defects_charge != 0 else np.zeros(0, np.int32)
        cs = np.concatenate([e_site, h_site, d])
        cq = np.concatenate([np.full(len(e_site), -1, np.int32), np.full(len(h_site), 1, np.int32),
                             np.full(len(d), defects_charge, np.int32)])
        targets = np.concatenate([cs, _i32(e_fsite), _i32(h_fsite)])
        self._offset = len(cs)
        self._o = self.ctx.pair_potential(cs, cq, targets, _abi.LMG_GAUSS if self.gauss else 0)

    launch_gauss_kernel2 = launch_coulomb_kernel2

    def get_output_host(self, index: int) -> float:
        return float(self._o[index])

    def get_output_host_future(self, index: int) -> float:
        return float(self._o[index + self._offset])

    def launch_coulomb_kernel1(self, e_site, h_site, defect_site=(), defects_charge=0,
                               no_cutoff=False) -> np.ndarray:
        d = _i32(defect_site) if defects_charge != 0 else np.zeros(0, np.int32)
        cs = np.concatenate([_i32(e_site), _i32(h_site), d])
        cq = np.concatenate([np.full(len(e_site), -1, np.int32), np.full(len(h_site), 1, np.int32),
                             np.full(len(d), defects_charge, np.int32)])
        flags = (_abi.LMG_GAUSS if self.gauss else 0) | (_abi.LMG_NO_CUTOFF if no_cutoff else 0)
        self._map = self.ctx.potential_map(cs, cq, flags)
        return self._map


# ---------------------------------------------------------------------------
# simulation
# ---------------------------------------------------------------------------
def make_config(p: SimulationParameters) -> LmgSimConfig:
    cfg = LmgSimConfig()
    cfg.simulation_type = 1 if p.is_solarcell else 0
    cfg.coulomb_carriers = int(bool(p.coulomb_carriers))
    cfg.source_metropolis = int(bool(p.source_metropolis))
    cfg.source_coulomb = int(bool(p.source_coulomb))
    cfg.balance_charges = int(bool(p.balance_charges))
    cfg.recombination_range = int(p.recombination_range)
    cfg.max_electrons = p.max_electrons()
    cfg.max_holes = p.max_holes()
    (el, er), (hl, hr) = drain_slots(p)
    cfg.e_drain_slot[0], cfg.e_drain_slot[1] = el, er
    cfg.h_drain_slot[0], cfg.h_drain_slot[1] = hl, hr
    cfg.voltage_left = p.voltage_left
    cfg.voltage_right = p.voltage_right
    for i, v in enumerate(flux_rates(p)):
        cfg.rate[i] = float(v)
    for i, v in enumerate(coupling_table(p.hopping_range)):
        cfg.coupling[i] = float(v)
    return cfg


def pack_state(p: SimulationParameters, electrons=(), holes=(), defects=(), potential=None,
               rng_window=None, rng_consumed=0, e_lifetime=None, e_pathlength=None, h_lifetime=None,
               h_pathlength=None, flux_attempts=None, flux_successes=None, current_step=0):
    """An lmg_sim_state for one configuration; returns (state, arrays it points into)."""
    e, h, d = _i32(list(electrons)), _i32(list(holes)), _i32(list(defects))
    pot = np.ascontiguousarray(background_potential(p) if potential is None else potential,
                               dtype=np.float64)
    if len(pot) != p.volume:
        raise ValueError("potential must have V entries")
    win = mt19937_seed_window(p.random_seed) if rng_window is None else np.asarray(rng_window, np.uint32)
    arrays = dict(e=e, h=h, d=d, pot=pot,
                  el=_i32(np.zeros(len(e)) if e_lifetime is None else e_lifetime),
                  ep=_i32(np.zeros(len(e)) if e_pathlength is None else e_pathlength),
                  hl=_i32(np.zeros(len(h)) if h_lifetime is None else h_lifetime),
                  hp=_i32(np.zeros(len(h)) if h_pathlength is None else h_pathlength))
    st = LmgSimState()
    st.n_electrons, st.n_holes, st.n_defects = len(e), len(h), len(d)
    pi = lambda a: a.ctypes.data_as(_abi.PI32)
    st.e_site, st.e_lifetime, st.e_pathlength = pi(arrays["e"]), pi(arrays["el"]), pi(arrays["ep"])
    st.h_site, st.h_lifetime, st.h_pathlength = pi(arrays["h"]), pi(arrays["hl"]), pi(arrays["hp"])
    st.defect_site = pi(arrays["d"])
    st.potential = arrays["pot"].ctypes.data_as(_abi.PF64)
    for i in range(624):
        st.rng_window[i] = int(win[i])
    st.rng_consumed = int(rng_consumed)
    fa = np.zeros(10, np.uint64) if flux_attempts is None else np.asarray(flux_attempts, np.uint64)
    fs = np.zeros(10, np.uint64) if flux_successes is None else np.asarray(flux_successes, np.uint64)
    for i in range(10):
        st.flux_attempts[i] = int(fa[i])
        st.flux_successes[i] = int(fs[i])
    st.current_step = int(current_step)
    return st, arrays


def lmg_params(p: SimulationParameters, device_id: int = 0) -> LmgParams:
    return LmgParams(grid_x=p.grid_x, grid_y=p.grid_y, grid_z=p.grid_z, cutoff=p.electrostatic_cutoff,
                     hopping_range=p.hopping_range, defects_charge=p.defects_charge,
                     device_id=device_id, reserved=0, prefactor=p.prefactor, inverse_kt=p.inverse_kt,
                     gaussian_sigma=p.coulomb_gaussian_sigma, exciton_binding=p.exciton_binding)


def _unpack_state(st: LmgSimState, bufs) -> dict:
    ne, nh = st.n_electrons, st.n_holes
    return dict(
        e_site=bufs["es"][:ne].copy(), e_lifetime=bufs["el"][:ne].copy(),
        e_pathlength=bufs["ep"][:ne].copy(), h_site=bufs["hs"][:nh].copy(),
        h_lifetime=bufs["hl"][:nh].copy(), h_pathlength=bufs["hp"][:nh].copy(),
        rng_window=np.array(st.rng_window[:], dtype=np.uint32), rng_consumed=int(st.rng_consumed),
        flux_attempts=np.array(st.flux_attempts[:], dtype=np.uint64),
        flux_successes=np.array(st.flux_successes[:], dtype=np.uint64),
        current_step=int(st.current_step))


def _store_buffers(cap_e: int, cap_h: int):
    bufs = {k: np.zeros(cap_e if k[0] == "e" else cap_h, np.int32)
            for k in ("es", "el", "ep", "hs", "hl", "hp")}
    st = LmgSimState()
    pi = lambda a: a.ctypes.data_as(_abi.PI32)
    st.e_site, st.e_lifetime, st.e_pathlength = pi(bufs["es"]), pi(bufs["el"]), pi(bufs["ep"])
    st.h_site, st.h_lifetime, st.h_pathlength = pi(bufs["hs"]), pi(bufs["hl"]), pi(bufs["hp"])
    return st, bufs


class Batch:
    """Many independent simulations advanced together (lmg_batch_*).

    ``replicas`` is a list of dicts with a ``params`` (SimulationParameters) entry and
    the keyword arguments of :func:`pack_state` (electrons, holes, potential, ...).
    """

    def __init__(self, replicas, device_id: int = 0, de_path: str = "auto"):
        """``de_path``: "auto" (the autotuner picks the culled path for the replicas that
        suit it, or none), "pairs" (every replica all-pairs) or "culled" (every replica
        that can be culled)."""
        modes = {"auto": 0, "pairs": 1, "culled": 2}
        if de_path not in modes:
            raise ValueError(f"de_path must be one of {sorted(modes)}")
        L = _abi.lib()
        n = len(replicas)
        self._params = (LmgParams * n)()
        self._configs = (LmgSimConfig * n)()
        self._states = (LmgSimState * n)()
        self._keep = []
        self.caps = []
        for r, rep in enumerate(replicas):
            rep = dict(rep)
            p = rep.pop("params")
            p.check()
            self._params[r] = lmg_params(p, device_id)
            self._configs[r] = make_config(p)
            st, arrays = pack_state(p, **rep)
            self._states[r] = st
            self._keep.append(arrays)
            self.caps.append((max(self._configs[r].max_electrons, st.n_electrons, 1),
                              max(self._configs[r].max_holes, st.n_holes, 1)))
        h = C.c_void_p()
        rc = L.lmg_batch_create_with_path(n, C.cast(self._params, C.c_void_p), C.cast(self._configs, C.c_void_p),
                                          C.cast(self._states, C.c_void_p), device_id, modes[de_path],
                                          C.byref(h))
        if rc != 0:
            raise LmgError(rc, L.lmg_batch_last_error(None).decode())
        self._h = h

    def __len__(self):
        return int(_abi.lib().lmg_batch_size(self._h))

    def _chk(self, rc):
        if rc != 0:
            raise LmgError(rc, _abi.lib().lmg_batch_last_error(self._h).decode())

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib().lmg_batch_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def step(self, n: int):
        self._chk(_abi.lib().lmg_batch_step(self._h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = C.c_double()
        self._chk(_abi.lib().lmg_batch_step_timed(self._h, int(n), C.byref(ms)))
        return float(ms.value)

    def delta_e(self, replica: int):
        """(fsite_e, fsite_h, dE_e, dE_h) of one replica's pending proposals."""
        s = self.state(replica)
        ne, nh = len(s["e_site"]), len(s["h_site"])
        fe, fh = np.zeros(max(ne, 1), np.int32), np.zeros(max(nh, 1), np.int32)
        de, dh = np.zeros(max(ne, 1)), np.zeros(max(nh, 1))
        self._chk(_abi.lib().lmg_batch_delta_e(self._h, int(replica), _p(fe), _p(fh), _p(de), _p(dh)))
        return fe[:ne], fh[:nh], de[:ne], dh[:nh]

    def step_phases(self, n: int):
        """n steps kernel by kernel with HIP events around the replica-wide dE kernel;
        returns (dE kernel ms, rest of the step ms, event span ms), summed over the steps."""
        out = (C.c_double * 3)()
        self._chk(_abi.lib().lmg_batch_step_phases(self._h, int(n), out))
        return float(out[0]), float(out[1]), float(out[2])

    def sync(self):
        self._chk(_abi.lib().lmg_batch_sync(self._h))

    def set_graphs(self, on: bool = True):
        """Replay one captured step of every replica as a HIP graph (default) or launch
        kernel by kernel."""
        self._chk(_abi.lib().lmg_batch_set_graphs(self._h, int(on)))

    @property
    def tuning(self) -> dict:
        """What the autotuner chose and every candidate it timed: the all-pairs replicas'
        common charge-chunk count (0 when all are culled), the culled replicas' workgroup
        total, the dE phase's time for all replicas, and [shape, us] lists."""
        s, us = C.c_int32(0), C.c_double(0.0)
        L = _abi.lib()
        self._chk(L.lmg_batch_get_tuning(self._h, C.byref(s), C.byref(us)))
        out = json.loads(L.lmg_batch_tuning_report(self._h).decode())
        out.update(charge_chunks=int(s.value), de_kernel_us=float(us.value))
        return out

    @property
    def de_paths(self) -> list:
        """Per replica the dE path it runs: "pairs" or "culled"."""
        n = len(self)
        paths = (C.c_int32 * n)()
        self._chk(_abi.lib().lmg_batch_get_de_paths(self._h, paths))
        return ["culled" if v == 2 else "pairs" for v in paths]

    def state(self, replica: int) -> dict:
        st, bufs = _store_buffers(*self.caps[replica])
        self._chk(_abi.lib().lmg_batch_store(self._h, int(replica), C.byref(st)))
        return _unpack_state(st, bufs)


class Simulation:
    """Simulation::performIterations on the device for one configuration.

    The configuration is given explicitly (what a checkpoint holds: carrier
    lists in list order, defects, the background potential, the engine state
    and the flux counters); see :func:`background_potential` and
    :func:`mt19937_seed_window`.
    """

    def __init__(self, params: SimulationParameters, electrons: Iterable[int] = (),
                 holes: Iterable[int] = (), defects: Iterable[int] = (),
                 potential: Optional[np.ndarray] = None, rng_window: Optional[np.ndarray] = None,
                 rng_consumed: int = 0, e_lifetime=None, e_pathlength=None, h_lifetime=None,
                 h_pathlength=None, flux_attempts=None, flux_successes=None, current_step: int = 0,
                 device_id: int = 0, context: Optional[Context] = None, de_path: str = "auto",
                 de_chunks: int = 0):
        params.check()
        self.params = params
        self.ctx = context or Context.from_parameters(params, device_id)
        modes = {"auto": 0, "pairs": 1, "culled": 2}
        if de_path not in modes:
            raise ValueError(f"de_path must be one of {sorted(modes)}")
        self._chk(_abi.lib().lmg_sim_set_de_path(self.ctx.handle, modes[de_path]))
        self._chk(_abi.lib().lmg_sim_set_de_chunks(self.ctx.handle, int(de_chunks)))   # 0: autotune
        self.config = make_config(params)
        self.cap_e = max(self.config.max_electrons, 1)
        self.cap_h = max(self.config.max_holes, 1)
        self.load(electrons, holes, defects, potential, rng_window, rng_consumed, e_lifetime,
                  e_pathlength, h_lifetime, h_pathlength, flux_attempts, flux_successes,
                  current_step)

    def load(self, electrons, holes, defects=(), potential=None, rng_window=None, rng_consumed=0,
             e_lifetime=None, e_pathlength=None, h_lifetime=None, h_pathlength=None,
             flux_attempts=None, flux_successes=None, current_step=0):
        st, keep = pack_state(self.params, electrons, holes, defects, potential, rng_window,
                              rng_consumed, e_lifetime, e_pathlength, h_lifetime, h_pathlength,
                              flux_attempts, flux_successes, current_step)
        self.cap_e = max(self.cap_e, st.n_electrons)
        self.cap_h = max(self.cap_h, st.n_holes)
        self._chk(_abi.lib().lmg_sim_load(self.ctx.handle, C.byref(self.config), C.byref(st)))

    def _chk(self, rc):
        self.ctx._chk(rc)

    @property
    def de_path(self) -> str:
        """The dE path in use: "pairs" (every charge) or "culled" (brick lists)."""
        m = C.c_int32(0)
        self._chk(_abi.lib().lmg_sim_get_de_path(self.ctx.handle, C.byref(m)))
        return "culled" if m.value == 2 else "pairs"

    def set_synchronous(self, on: bool = True):
        self._chk(_abi.lib().lmg_sim_set_sync(self.ctx.handle, int(on)))

    def set_graphs(self, on: bool = True):
        """Replay one captured step as a HIP graph (default) or launch kernel by kernel."""
        self._chk(_abi.lib().lmg_sim_set_graphs(self.ctx.handle, int(on)))

    @property
    def tuning(self) -> dict:
        """The dE kernel's shape chosen at load and its measured time (-1 when forced):
        charge chunks (all pairs) or work items (culled; the time covers the list
        building)."""
        shape, us = C.c_int32(0), C.c_double(0.0)
        self._chk(_abi.lib().lmg_sim_get_tuning(self.ctx.handle, C.byref(shape), C.byref(us)))
        if self.de_path == "culled":
            return {"path": "culled", "work_items": int(shape.value), "de_phase_us": float(us.value)}
        return {"path": "pairs", "charge_chunks": int(shape.value), "de_kernel_us": float(us.value)}

    def perform_iterations(self, n: int):
        """Simulation::performIterations(n) (simulation.cpp:29-111), enqueued."""
        self._chk(_abi.lib().lmg_sim_step(self.ctx.handle, int(n)))

    step = perform_iterations

    def step_timed(self, n: int):
        """n steps with HIP events around each phase; returns (dE phase ms,
        commit phase ms, event span ms, pair-sum kernel ms), each summed over
        the steps."""
        out = (C.c_double * 4)()
        self._chk(_abi.lib().lmg_sim_step_timed(self.ctx.handle, int(n), out))
        return float(out[0]), float(out[1]), float(out[2]), float(out[3])

    def device_error(self) -> int:
        code = C.c_int32(0)
        self._chk(_abi.lib().lmg_sim_device_error(self.ctx.handle, C.byref(code)))
        return int(code.value)

    def state(self) -> dict:
        st, bufs = _store_buffers(self.cap_e, self.cap_h)
        self._chk(_abi.lib().lmg_sim_store(self.ctx.handle, C.byref(st)))
        return _unpack_state(st, bufs)

    def delta_e(self):
        """(fsite_e, fsite_h, dE_e, dE_h) of the pending proposals."""
        s = self.state()
        ne, nh = len(s["e_site"]), len(s["h_site"])
        fe, fh = np.zeros(max(ne, 1), np.int32), np.zeros(max(nh, 1), np.int32)
        de, dh = np.zeros(max(ne, 1)), np.zeros(max(nh, 1))
        self._chk(_abi.lib().lmg_sim_delta_e(self.ctx.handle, _p(fe), _p(fh), _p(de), _p(dh)))
        return fe[:ne], fh[:nh], de[:ne], dh[:nh]
