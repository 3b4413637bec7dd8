"""World set-up on the host: the seeded placement of defects, carriers and traps
that World::initialize performs before the first step (world.cpp:501-660).

The placement consumes the one Boost mt19937 stream of the run (rand.cpp:8-124)
in the reference's order -- defects (world.cpp:659-757), electrons, then holes
(world.cpp:759-958 through SourceAgent::tryToSeed, sourceagent.cpp:71-90), then
traps (potential.cpp:74-244) -- so the engine window handed to the device after
set-up is the one the reference's first step draws from.  Explicit lists (what
an input or checkpoint file lists) are placed first; the rest is drawn:

* defects up to ``int(defect.percentage * V)``;
* carriers up to ``int(max * seed.charges)`` of each kind, where max is
  ``int(percentage * V)`` (also on a reload, which tops a drained list up);
* traps up to ``int(trap.percentage * V)``: ``seed.percentage`` of the drawn
  ones as uniformly drawn seed sites, the rest grown from a drawn seed's
  hopping-range-1 neighbour (potential.cpp:170-220).

uniform_int is Boost's (>= 1.47) bucket rejection on 32-bit words (no engine
call for a one-value range).  With ``gaussian.stdev > 0`` every drawn trap's
depth gets a boost::normal_distribution deviation (potential.cpp:211-220), drawn
as Boost 1.44-1.55 draw it (Box-Muller, two words per value); Boost 1.56
replaced that algorithm, so runs under a newer Boost draw other depths.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .host import FLUX, SimulationParameters, background_potential, flux_rates, mt19937_seed_window

_N, _M = 624, 397
_UPPER, _LOWER, _MATRIX = np.uint32(0x80000000), np.uint32(0x7FFFFFFF), np.uint32(0x9908B0DF)


class MT19937:
    """boost::mt19937 on the host, generated 624 words at a time with numpy.

    ``window()`` is the engine state as the device and the [RandomState] text
    hold it: the 624 most recent raw words x[k .. k+623] after k outputs."""

    def __init__(self, seed: int = 5489, window: Optional[Sequence[int]] = None):
        if window is None:
            self._old = mt19937_seed_window(int(seed) & 0xFFFFFFFF).astype(np.uint32)
        else:
            w = np.asarray(window, dtype=np.uint64)
            if w.shape != (_N,) or np.any(w >> 32):
                raise ValueError("an mt19937 window is 624 words of 32 bits")
            self._old = w.astype(np.uint32)
        self._new = None
        self._out: List[int] = []
        self._pos = _N
        self.consumed = 0

    def _twist(self) -> None:
        s = np.empty(2 * _N, dtype=np.uint32)
        s[:_N] = self._old if self._new is None else self._new
        for lo, hi in ((0, _N - _M), (_N - _M, 2 * (_N - _M)), (2 * (_N - _M), _N)):
            i = np.arange(lo, hi)
            y = (s[i] & _UPPER) | (s[i + 1] & _LOWER)
            s[_N + i] = s[i + _M] ^ (y >> np.uint32(1)) ^ np.where(y & np.uint32(1), _MATRIX, np.uint32(0))
        if self._new is not None:
            self._old = self._new
        self._new = s[_N:].copy()
        y = self._new.copy()
        y ^= y >> np.uint32(11)
        y ^= (y << np.uint32(7)) & np.uint32(0x9D2C5680)
        y ^= (y << np.uint32(15)) & np.uint32(0xEFC60000)
        y ^= y >> np.uint32(18)
        self._out = y.tolist()
        self._pos = 0

    def __call__(self) -> int:
        if self._pos == _N:
            self._twist()
        v = self._out[self._pos]
        self._pos += 1
        self.consumed += 1
        return v

    def window(self) -> np.ndarray:
        if self._new is None:
            return self._old.copy()
        return np.concatenate([self._old[self._pos:], self._new[:self._pos]]).astype(np.uint32)

    def integer(self, lo: int, hi: int) -> int:
        """Random::integer (rand.cpp:68-73): boost::uniform_int<int>(lo, hi)."""
        if hi < lo:
            raise ValueError("uniform_int: empty range")
        rng = (hi - lo) & 0xFFFFFFFF
        if rng == 0:
            return lo
        if rng == 0xFFFFFFFF:
            return lo + self()
        bucket = 0xFFFFFFFF // (rng + 1)
        if 0xFFFFFFFF % (rng + 1) == rng:
            bucket += 1
        while True:
            r = self() // bucket
            if r <= rng:
                return lo + r

    def random(self) -> float:
        """Random::random (rand.cpp:49-52): boost::uniform_01<double>, one word times 2^-32."""
        return self() * (1.0 / 4294967296.0)

    def normal(self, mean: float, sigma: float) -> float:
        """Random::normal (rand.cpp:61-66): a fresh boost::normal_distribution per call.  In
        Boost 1.44-1.55, the versions the reference's build files name, that is Box-Muller
        over two uniform_01 draws r1, r2, and a fresh distribution returns the cosine
        branch: sqrt(-2 log(1 - r2)) cos(2 pi r1) sigma + mean."""
        r1 = self.random()
        r2 = self.random()
        return math.sqrt(-2.0 * math.log(1.0 - r2)) * math.cos(2.0 * 3.14159265358979323846 * r1) * sigma + mean


def _neighbors_h1(dims, s: int) -> List[int]:
    """Grid::neighborsSite(s, 1) of the electron grid: -x, +x, -y, +y, -z, +z
    (cubicgrid.cpp:207-245), then the electron drain's slot on an x face
    (cubicgrid.cpp:447-468), given here as -1: a drain is never a trap site, but its
    entry takes part in the draw."""
    X, Y, Z = dims
    x, y, z = s % X, (s // X) % Y, s // (X * Y)
    out = []
    if x > 0:
        out.append(s - 1)
    if x < X - 1:
        out.append(s + 1)
    if y > 0:
        out.append(s - X)
    if y < Y - 1:
        out.append(s + X)
    if z > 0:
        out.append(s - X * Y)
    if z < Z - 1:
        out.append(s + X * Y)
    if x == 0:
        out.append(-1)
    if x == X - 1:
        out.append(-1)
    return out


@dataclass
class WorldSetup:
    """What World::initialize leaves for the first step."""
    params: SimulationParameters
    electrons: List[int]
    holes: List[int]
    defects: List[int]
    traps: List[int]
    trap_potentials: List[float]
    potential: np.ndarray        # background potential [V]: linear + gate + traps
    rng_window: np.ndarray       # engine window after the placement draws
    rng_consumed: int            # outputs drawn during set-up
    rng_seed: int = 0            # Random::m_seed: the [RandomState] seed, else uint32(random.seed)


def initialize(p: SimulationParameters, electrons: Sequence[int] = (), holes: Sequence[int] = (),
               defects: Sequence[int] = (), traps: Sequence[int] = (),
               trap_potentials: Optional[Sequence[float]] = None,
               rng_window: Optional[Sequence[int]] = None, rng_seed: Optional[int] = None) -> WorldSetup:
    """World::initialize's placement for explicit lists plus the seeded rest.

    The engine starts from ``rng_window`` (a [RandomState]) when given, else from
    ``random.seed`` (0 seeds from the clock, as Random(0) does, rand.cpp:8-20; the
    seed used is written back into the returned parameters).  The set-up keeps the
    engine's seed as Random::m_seed holds it: ``rng_seed`` (the seed a [RandomState]
    carried) with a window, else ``random.seed`` cast to 32 bits (rand.cpp:36-47)."""
    if rng_window is None and p.random_seed == 0:
        p = SimulationParameters(**{**p.__dict__, "random_seed": int(time.time()) & 0xFFFFFFFF})
    p.check()
    rng = MT19937(p.random_seed, rng_window)
    X, Y, Z = p.grid_x, p.grid_y, p.grid_z
    V = X * Y * Z
    max_e, max_h = p.max_electrons(), p.max_holes()
    max_d = int(p.defect_percentage * float(V))
    max_t = int(p.trap_percentage * float(V))

    # createSources / createDrains refusals (world.cpp:961-1187)
    r = {k: float(v) for k, v in zip(FLUX, flux_rates(p))}
    if all(r[k] == 0 for k in ("eSL", "eSR", "hSL", "hSR", "xS")):
        raise ValueError("all sources have zero rate")
    if max_e == 0 and (r["eSR"] > 0 or r["eSL"] > 0 or r["xS"] > 0):
        raise ValueError("electron source rate > 0, yet max electrons == 0")
    if max_h == 0 and (r["hSR"] > 0 or r["hSL"] > 0 or r["xS"] > 0):
        raise ValueError("hole source rate > 0, yet max holes == 0")
    if all(r[k] == 0 for k in ("eDL", "eDR", "hDL", "hDR", "xD")):
        raise ValueError("all drains have zero rate")
    if max_e > 0 and r["eDR"] <= 0 and r["eDL"] <= 0 and r["xD"] <= 0:
        raise ValueError("all electron drain rates are zero")
    if max_h > 0 and r["hDR"] <= 0 and r["hDL"] <= 0 and r["xD"] <= 0:
        raise ValueError("all hole drain rates are zero")

    # placeDefects (world.cpp:659-757)
    d_list: List[int] = []
    d_set = set()
    if len(defects) > max_d:
        raise ValueError("invalid number of defects in checkpoint")
    if max_d > 0:
        for s in defects:
            s = int(s)
            if s in d_set:
                raise ValueError(f"defect already exists at site {s}")
            if not 0 <= s < V:
                raise ValueError(f"defect site {s} out of range")
            d_list.append(s)
            d_set.add(s)
        tries, max_tries = 0, 10 * V
        while len(d_list) < max_d:
            tries += 1
            s = rng.integer(0, V - 1)
            if s not in d_set:
                d_list.append(s)
                d_set.add(s)
            if tries > max_tries:
                raise RuntimeError("can not seed defects")

    # placeElectrons / placeHoles (world.cpp:759-958, sourceagent.cpp:71-90, 204-247)
    def place(ids, mx):
        lst: List[int] = []
        occ = set()
        seeded = int(mx * p.seed_charges)
        to_seed = max(seeded - len(ids), 0)
        if len(ids) + to_seed > mx or len(ids) > mx:
            raise ValueError("invalid number of carriers to be placed")

        def valid(s):
            return len(lst) < mx and 0 <= s < V and s not in occ and s not in d_set

        for s in ids:
            s = int(s)
            if not valid(s):
                raise ValueError(f"can not place a carrier at site {s}")
            lst.append(s)
            occ.add(s)
        if to_seed > 0 and p.seed_charges != 0:
            tries, max_tries = 0, 10 * V
            while len(lst) < mx * p.seed_charges:
                s = rng.integer(0, V - 1)
                if valid(s):
                    lst.append(s)
                    occ.add(s)
                tries += 1
                if tries > max_tries:
                    raise RuntimeError("can not seed carriers")
        return lst

    e_list = place(list(electrons), max_e)
    h_list = place(list(holes), max_h)

    # setPotentialTraps (potential.cpp:74-244)
    forced = [int(t) for t in traps]
    randomly = max_t - len(forced)
    seeds = int(randomly * p.seed_percentage)
    grown = randomly - seeds
    if randomly <= 0:
        seeds = grown = 0
    if seeds == 0 and grown > 0:
        seeds, grown = grown, 0
    if len(forced) > max_t or randomly > max_t or seeds < 0 or grown < 0:
        raise ValueError("invalid trap parameters: more traps listed than trap.percentage allows")
    tp_in = [] if trap_potentials is None else [float(v) for v in trap_potentials]
    if forced and tp_in and len(tp_in) != len(forced):
        raise ValueError("[TrapPotentials] and [Traps] differ in length")
    t_list = list(forced)
    t_pot = tp_in if tp_in else [p.trap_potential] * len(forced)
    if randomly > 0:
        forced_set = set(forced)
        rids: List[int] = []
        rset = set()
        while len(rids) < seeds:
            s = rng.integer(0, V - 1)
            if s not in rset and s not in forced_set:
                rids.append(s)
                rset.add(s)
        progress = 0
        while progress < grown:
            si = rng.integer(0, len(rids) - 1)
            nb = _neighbors_h1((X, Y, Z), rids[si])
            ns = nb[rng.integer(0, len(nb) - 1)]
            if ns >= 0 and ns not in rset and ns not in forced_set:
                rids.append(ns)
                rset.add(ns)
                progress += 1
        r_pot = [p.trap_potential] * len(rids)
        # a normal deviation per randomly placed trap, in list order (potential.cpp:211-220)
        if abs(p.gaussian_stdev) > 0:
            r_pot = [v + rng.normal(0.0, p.gaussian_stdev) for v in r_pot]
        t_list += rids
        t_pot += r_pot

    pot = background_potential(p, t_list, t_pot if t_list else None)
    return WorldSetup(params=p, electrons=e_list, holes=h_list, defects=d_list, traps=t_list,
                      trap_potentials=t_pot, potential=pot, rng_window=rng.window(),
                      rng_consumed=rng.consumed,
                      rng_seed=(int(rng_seed) if rng_window is not None and rng_seed is not None
                                else int(p.random_seed) & 0xFFFFFFFF))
