"""Known-answer pins of the oracle's sweep order, derived by hand from the reference's rules.

Each scenario is a tiny world (one to three carriers, at most 18 sites) stepped once.
The expected outcome is worked out in the test, draw by draw, from numpy's legacy
MT19937 stream for the seed -- an engine independent of the oracle -- with Boost's
uniform_int bucket rule and uniform_01 = u32 * 2^-32 (pinned in test_oracle_engine.py).
Every rule below is written out literally from the reference, and every scenario
carries the plausible mistakes next to the rule: for each mistake at least one seed
must give another outcome (so the check can fail), and the oracle must match the
reference rule on every seed.

* Phase order (simulation.cpp:46-108): electron proposals, hole proposals, electron
  decisions, hole decisions, recombination, ticks (electrons, then holes), injection.
* Decisions draw only for an Empty or a Drain target (chargeagent.cpp:90-161): an
  occupied (same-kind carrier) or defect target rejects without a draw.
* Conflicts: the ticks run in list order and a move into a site taken earlier in the
  same tick is aborted (chargeagent.cpp:163-199).
* Injection: sources eL, eR, hL, hR in that order (simulation.cpp:246-249); each counts
  an attempt and draws its site before validity is checked, even at rate 0
  (sourceagent.cpp:92-103, 204-247); the accept draw only follows a valid site.
* Recombination (drainagent.cpp:113-214): electrons in list order; the candidate holes
  are the same site, then the electron's neighbour list; no draw for one candidate; a
  hole already removed ends the try with no attempt and no draw.
* balanceCharges (simulation.cpp:253-317): after the exciton source, electrons are injected
  while they are fewer than the holes; at equal voltages a uniform_01 draw above 0.5 picks
  the left source, else the right.
* Range-2 neighbour order: a literal transcription of cubicgrid.cpp:246-437.
* Placement order (world.cpp:624-643; 659-958; potential.cpp:74-244): defects,
  electrons, holes, then traps (seeds, then growth from a seed's range-1 neighbour
  list, whose x-face drain entry is drawn and rejected).
"""
import numpy as np
import pytest

TWO32 = 2.0 ** 32
ESL, ESR, HSL, HSR, XS, EDL, EDR, HDL, HDR, XD = range(10)
SEEDS = range(1, 41)


class Words:
    """numpy's legacy MT19937 outputs for one seed, handed out in order and counted."""

    def __init__(self, seed):
        self.w = np.random.RandomState(seed).randint(0, 2 ** 32, size=4096, dtype=np.uint64)
        self.k = 0

    def word(self):
        v = int(self.w[self.k])
        self.k += 1
        return v

    def u01(self):
        return self.word() / TWO32

    def integer(self, lo, hi):
        """boost::uniform_int<int>(lo, hi): no draw for one value; bucket rejection."""
        m = hi - lo + 1
        if m == 1:
            return lo
        bucket = 0xFFFFFFFF // m + (1 if 0xFFFFFFFF % m == m - 1 else 0)
        while True:
            r = self.word() // bucket
            if r < m:
                return lo + r


def seed_window(seed):
    key = np.random.RandomState(seed).get_state()[1]
    return key.astype(np.uint32)   # init_genrand's x[0..623], the engine before any output


def make_world(oracle, seed, electrons=(), holes=(), defects=(), **kw):
    base = dict(grid_z=1, coulomb_carriers=0, seed_charges=0.0, random_seed=seed,
                e_source_l_rate=0.0, e_source_r_rate=0.0, h_source_l_rate=0.0, h_source_r_rate=0.0,
                e_drain_l_rate=0.0, e_drain_r_rate=0.0, h_drain_l_rate=0.0, h_drain_r_rate=0.0)
    w = oracle.World(oracle.OrcParams.make(**{**base, **kw}), electrons=electrons, holes=holes, defects=defects)
    w.set_rng(seed_window(seed), 0)
    return w


def observed(w):
    e, h = w.carriers(0), w.carriers(1)
    a, s, _ = w.flux()
    return dict(e=e["site"].tolist(), h=h["site"].tolist(), e_life=e["lifetime"].tolist(),
                h_life=h["lifetime"].tolist(), e_path=e["pathlength"].tolist(),
                h_path=h["pathlength"].tolist(), att=[int(x) for x in a], suc=[int(x) for x in s],
                k=w.rng()[1])


class Carrier:
    def __init__(self, site, nbrs):
        self.site, self.nbrs, self.fsite = site, nbrs, site
        self.life = self.path = 0
        self.removed = False


def result(W, es, hs, att, suc, injected_e=(), injected_h=()):
    el = [c for c in es if not c.removed]
    hl = [c for c in hs if not c.removed]
    return dict(e=[c.site for c in el] + list(injected_e), h=[c.site for c in hl] + list(injected_h),
                e_life=[c.life for c in el] + [0] * len(injected_e),
                h_life=[c.life for c in hl] + [0] * len(injected_h),
                e_path=[c.path for c in el] + [0] * len(injected_e),
                h_path=[c.path for c in hl] + [0] * len(injected_h), att=att, suc=suc, k=W.k)


def check(oracle, scenario, mistakes, seeds=SEEDS):
    for m in mistakes:
        assert any(scenario(s, m) != scenario(s, None) for s in seeds), f"no seed tells mistake {m!r} apart"
    for s in seeds:
        w, expect = scenario(s, None, oracle)
        w.step(1)
        assert observed(w) == expect, f"seed {s}"


# ----------------------------------------------------------------------------- scenario A
# 5x3x1 transistor, Coulomb off, zero bias: one electron at (2,1) = 7 and one hole at
# (1,1) = 6; every decision target is Empty in its own grid, so each decision draws and
# accepts with u < coupling(1,0,0) = 1/3 (metropolisWithCoupling with dE = 0, rand.cpp:91-106).
# Sources: only hR has a rate (0.5); faces x=0 -> [0, 5, 10], x=4 -> [4, 9, 14] (sorted
# sliceIndex, cubicgrid.cpp:473-481).
def scenario_phase_order(seed, mistake, oracle=None):
    W = Words(seed)
    e = Carrier(7, [6, 8, 2, 12])      # W, E, S, N (cubicgrid.cpp:207-245)
    h = Carrier(6, [5, 7, 1, 11])
    att, suc = [0] * 10, [0] * 10

    def propose(c):
        c.fsite = c.nbrs[W.integer(0, 3)]

    def decide(c):
        c.life += 1
        if W.u01() < 1.0 / 3.0:
            c.path += 1
        else:
            c.fsite = c.site

    if mistake == "interleaved":
        propose(e), decide(e), propose(h), decide(h)
    elif mistake == "holes decide first":
        propose(e), propose(h), decide(h), decide(e)
    else:
        propose(e), propose(h), decide(e), decide(h)
    for c in (e, h):
        c.site = c.fsite
    injected = []
    faces = {ESL: [0, 5, 10], ESR: [4, 9, 14], HSL: [0, 5, 10], HSR: [4, 9, 14]}
    order = [ESL, ESR, HSL, HSR] if mistake != "holes inject first" else [HSL, HSR, ESL, ESR]
    for f in order:
        if mistake == "no site draw at rate 0" and f != HSR:
            continue
        att[f] += 1
        site = faces[f][W.integer(0, 2)]
        if f == HSR:      # the only source with a rate; its site is never h's (x = 4)
            if W.u01() < 0.5:
                suc[f] += 1
                injected.append(site)
    out = result(W, [e], [h], att, suc, injected_h=injected)
    if oracle is None:
        return out
    return make_world(oracle, seed, electrons=[7], holes=[6], grid_x=5, grid_y=3, electron_percentage=0.2, hole_percentage=0.2,
                      h_source_r_rate=0.5, e_drain_l_rate=0.5, e_drain_r_rate=0.5, h_drain_l_rate=0.5,
                      h_drain_r_rate=0.5), out


def test_phase_order(oracle):
    check(oracle, scenario_phase_order, ["interleaved", "holes decide first", "holes inject first",
                                         "no site draw at rate 0"])


# ----------------------------------------------------------------------------- scenario B
# 5x3x1 transistor, three holes in list order at (0,1) = 5, (1,1) = 6, (3,1) = 8 and a
# defect at (1,0) = 1.  Targets occupied by a hole or the defect reject without a draw;
# the left drain slot of h0 (appended after its in-lattice neighbours, cubicgrid.cpp:447-457)
# draws chooseYes(h.drain.l.rate = 0.6); h1 and h2 can both win site 7, and h1 moves first.
def scenario_decision_draws(seed, mistake, oracle=None):
    W = Words(seed)
    D = "drain"
    h0, h1, h2 = Carrier(5, [6, 0, 10, D]), Carrier(6, [5, 7, 1, 11]), Carrier(8, [7, 9, 3, 13])
    hs = [h0, h1, h2]
    att, suc = [0] * 10, [0] * 10
    for c in hs:
        c.fsite = c.nbrs[W.integer(0, 3)]
    occupied = {5, 6, 8}
    defects = {1}
    for c in hs:
        c.life += 1
        f = c.fsite
        if f == D:
            if mistake == "no draw for a drain":
                c.fsite = c.site
                continue
            att[HDL] += 1
            if W.u01() < 0.6:
                suc[HDL] += 1
                c.path += 1
            else:
                c.fsite = c.site
        elif f in occupied or (f in defects and mistake != "defect taken as empty"):
            if mistake == "draw for every target":
                W.u01()
            c.fsite = c.site
        else:
            if W.u01() < 1.0 / 3.0:
                c.path += 1
            else:
                c.fsite = c.site
    taken = set(c.site for c in hs)
    order = hs if mistake != "last in the list moves first" else hs[::-1]
    for c in order:
        if c.fsite == c.site:
            continue
        if c.fsite == D:
            c.removed = True
            taken.discard(c.site)
            continue
        if c.fsite in taken and mistake != "no conflict check":
            c.fsite = c.site
            continue
        taken.discard(c.site)
        taken.add(c.fsite)
        c.site = c.fsite
    injected = []
    for f, face in ((ESL, [0, 5, 10]), (ESR, [4, 9, 14]), (HSL, [0, 5, 10]), (HSR, [4, 9, 14])):
        att[f] += 1
        site = face[W.integer(0, 2)]
        if f == HSR and site not in taken and site not in defects and len([c for c in hs if not c.removed]) < 6:
            if W.u01() < 0.3:
                suc[f] += 1
                injected.append(site)
    out = result(W, [], hs, att, suc, injected_h=injected)
    if oracle is None:
        return out
    return make_world(oracle, seed, holes=[5, 6, 8], defects=[1], grid_x=5, grid_y=3, electron_percentage=0.0, hole_percentage=0.4,
                      defect_percentage=0.07, h_source_r_rate=0.3, h_drain_l_rate=0.6, h_drain_r_rate=0.6), out


def test_decisions_draw_only_for_empty_or_drain_targets(oracle):
    check(oracle, scenario_decision_draws, ["draw for every target", "no draw for a drain",
                                            "defect taken as empty", "last in the list moves first",
                                            "no conflict check"],
          seeds=range(1, 601))   # both h1 and h2 win site 7 with probability 1/144


# ----------------------------------------------------------------------------- scenario C
# 3x3x1 transistor with one electron at (0,1) = 3 (neighbours E 4, S 0, N 6, then the left
# electron drain) and all four face sources: eL 0.7, eR 0 (draws its site all the same),
# hL 0.4, hR 0.9.  An injection onto a site the electron holds after its tick is invalid
# and draws no accept.
def scenario_injection(seed, mistake, oracle=None):
    W = Words(seed)
    D = "drain"
    e = Carrier(3, [4, 0, 6, D])
    att, suc = [0] * 10, [0] * 10
    e.fsite = e.nbrs[W.integer(0, 3)]
    e.life += 1
    if e.fsite == D:
        att[EDL] += 1
        if W.u01() < 0.5:
            suc[EDL] += 1
            e.path += 1
        else:
            e.fsite = e.site
    elif W.u01() < 1.0 / 3.0:
        e.path += 1
    else:
        e.fsite = e.site
    if e.fsite == D:
        e.removed = True
    else:
        e.site = e.fsite
    e_sites = set() if e.removed else {e.site}
    h_sites = set()
    new_e, new_h = [], []
    rate = {ESL: 0.7, ESR: 0.0, HSL: 0.4, HSR: 0.9}
    face = {ESL: [0, 3, 6], ESR: [2, 5, 8], HSL: [0, 3, 6], HSR: [2, 5, 8]}
    order = {"electrons and holes alternate": [ESL, HSL, ESR, HSR],
             "right sources first": [ESR, ESL, HSR, HSL]}.get(mistake, [ESL, ESR, HSL, HSR])
    for f in order:
        if mistake == "no site draw at rate 0" and rate[f] == 0:
            continue
        att[f] += 1
        site = face[f][W.integer(0, 2)]
        mine = e_sites if f in (ESL, ESR) else h_sites
        valid = rate[f] > 0 and site not in mine and len(mine) < 3
        if not valid:
            if mistake == "accept drawn before validity":
                W.u01()
            continue
        if W.u01() < rate[f]:
            suc[f] += 1
            mine.add(site)
            (new_e if f in (ESL, ESR) else new_h).append(site)
    out = result(W, [e], [], att, suc, injected_e=new_e, injected_h=new_h)
    if oracle is None:
        return out
    return make_world(oracle, seed, electrons=[3], grid_x=3, grid_y=3, electron_percentage=0.34, hole_percentage=0.34,
                      e_source_l_rate=0.7, e_source_r_rate=0.0, h_source_l_rate=0.4, h_source_r_rate=0.9,
                      e_drain_l_rate=0.5, e_drain_r_rate=0.5, h_drain_l_rate=0.5, h_drain_r_rate=0.5), out


def test_injection_order_and_site_draws(oracle):
    check(oracle, scenario_injection, ["electrons and holes alternate", "right sources first",
                                       "no site draw at rate 0", "accept drawn before validity"])


# ----------------------------------------------------------------------------- scenario D
# 6x3x1 solar cell, Coulomb off, recombination.rate 0.6 at range 1 (= hopping.range, so the
# electron's own neighbour list is searched): electrons at (2,1) = 8 and (4,1) = 10 in list
# order, one hole at (3,1) = 9 between them, the only candidate of both.  Exciton source
# rate 0.5 (source.scale.area = 0: no area scaling), site drawn over the whole lattice.
def scenario_recombination(seed, mistake, oracle=None):
    W = Words(seed)
    e0, e1, h0 = Carrier(8, [7, 9, 2, 14]), Carrier(10, [9, 11, 4, 16]), Carrier(9, [8, 10, 3, 15])
    att, suc = [0] * 10, [0] * 10

    def recombine():
        for e in (e0, e1):
            if 9 not in e.nbrs:
                continue
            if h0.removed and mistake != "removed hole not checked":
                continue
            att[XD] += 1          # one candidate: uniform_int(0, 0) draws nothing
            if W.u01() < 0.6:
                suc[XD] += 1
                e.removed = True
                h0.removed = True

    for c in (e0, e1, h0):
        c.fsite = c.nbrs[W.integer(0, 3)]
    if mistake == "recombination before decisions":
        recombine()
    for c in (e0, e1, h0):
        c.life += 1
        if W.u01() < 1.0 / 3.0:
            c.path += 1
        else:
            c.fsite = c.site
    if mistake != "recombination before decisions":
        recombine()
    taken = set()
    for c in (e0, e1):
        if not c.removed:
            taken.add(c.site)
    for c in (e0, e1):
        if c.removed or c.fsite == c.site:
            continue
        if c.fsite in taken:
            continue
        taken.discard(c.site)
        taken.add(c.fsite)
        c.site = c.fsite
    if not h0.removed:
        h0.site = h0.fsite
    h_taken = set() if h0.removed else {h0.site}
    att[XS] += 1
    site = W.integer(0, 17)
    new = []
    if site not in taken and site not in h_taken:
        if W.u01() < 0.5:
            suc[XS] += 1
            new.append(site)
    out = result(W, [e0, e1], [h0], att, suc, injected_e=new, injected_h=new)
    if oracle is None:
        return out
    return make_world(oracle, seed, electrons=[8, 10], holes=[9], simulation_type=1, grid_x=6, grid_y=3,
                      electron_percentage=0.2, hole_percentage=0.2, recombination_rate=0.6,
                      recombination_range=1, generation_rate=0.5, source_scale_area=0.0,
                      e_drain_l_rate=-1.0, e_drain_r_rate=-1.0, h_drain_l_rate=-1.0, h_drain_r_rate=-1.0), out


def test_recombination_order(oracle):
    check(oracle, scenario_recombination, ["removed hole not checked", "recombination before decisions"])


# ----------------------------------------------------------------------------- scenario E
# 6x3x1 solar cell, Coulomb off, balance.charges on at equal voltages (simulation.cpp:233-317):
# one electron at (2,1) = 8, holes at (3,1) = 9 and (4,2) = 16 (neighbours W, E, S; no N on
# the last row).  Generation rate 0: the exciton source counts an attempt and draws its site,
# which is then invalid.  Balance: while electrons < holes, a uniform_01 draw above 0.5 picks
# the left electron source, else the right; the source counts an attempt, draws a face site
# (x = 0 -> [0, 6, 12], x = 5 -> [5, 11, 17]) and accepts with u < 0.7.
def scenario_balance(seed, mistake, oracle=None):
    W = Words(seed)
    e0, h0, h1 = Carrier(8, [7, 9, 2, 14]), Carrier(9, [8, 10, 3, 15]), Carrier(16, [15, 17, 10])
    att, suc = [0] * 10, [0] * 10
    for c in (e0, h0, h1):
        c.fsite = c.nbrs[W.integer(0, len(c.nbrs) - 1)]
    for c in (e0, h0, h1):        # every target is Empty in its own grid: each decision draws
        c.life += 1
        if W.u01() < 1.0 / 3.0:
            c.path += 1
        else:
            c.fsite = c.site
    e0.site = e0.fsite
    taken = {h0.site, h1.site}
    for c in (h0, h1):            # list order; a site taken earlier in the tick aborts the move
        if c.fsite != c.site and c.fsite not in taken:
            taken.discard(c.site)
            taken.add(c.fsite)
            c.site = c.fsite
    injected = []

    def exciton():
        att[XS] += 1
        W.integer(0, 17)          # rate 0: the site is drawn, then invalid

    def balance():
        tries = 0
        while 1 + len(injected) < 2:
            coin = W.u01()
            if mistake == "always left":
                f = ESL
            elif mistake == "coin inverted":
                f = ESL if coin < 0.5 else ESR
            else:
                f = ESL if coin > 0.5 else ESR
            att[f] += 1
            site = ([0, 6, 12] if f == ESL else [5, 11, 17])[W.integer(0, 2)]
            if site not in injected and W.u01() < 0.7:
                suc[f] += 1
                injected.append(site)
            tries += 1
            assert tries <= 1024

    if mistake == "balance first":
        balance(), exciton()
    elif mistake == "no balance":
        exciton()
    else:
        exciton(), balance()
    out = result(W, [e0], [h0, h1], att, suc, injected_e=injected)
    if oracle is None:
        return out
    return make_world(oracle, seed, electrons=[8], holes=[9, 16], simulation_type=1, grid_x=6, grid_y=3,
                      electron_percentage=0.2, hole_percentage=0.2, generation_rate=0.0, source_scale_area=0.0,
                      balance_charges=1, e_source_l_rate=0.7, e_source_r_rate=0.7,
                      e_drain_l_rate=-1.0, e_drain_r_rate=-1.0, h_drain_l_rate=-1.0, h_drain_r_rate=-1.0), out


def test_balance_charges_order(oracle):
    check(oracle, scenario_balance, ["no balance", "always left", "coin inverted", "balance first"])


# ----------------------------------------------------------------------------- range 2
# cubicgrid.cpp:246-437, transcribed: (dx, dy, dz) in push order; every push there is
# guarded by exactly the bounds its target needs, so the list is this table filtered to
# in-lattice targets.  Layers (z = 1) take the optimised branch (:250-306).
RANGE2_LAYER = [(-1, 0, 0), (-1, -1, 0), (-1, 1, 0), (-2, 0, 0),
                (1, 0, 0), (1, -1, 0), (1, 1, 0), (2, 0, 0),
                (0, -1, 0), (0, -2, 0),
                (0, 1, 0), (0, 2, 0)]
RANGE2_3D = [(-1, 0, 0), (-1, -1, 0), (-1, -1, -1), (-1, -1, 1), (-1, 1, 0), (-1, 1, -1), (-1, 1, 1),
             (-1, 0, -1), (-1, 0, 1), (-2, 0, 0),
             (1, 0, 0), (1, -1, 0), (1, -1, -1), (1, -1, 1), (1, 1, 0), (1, 1, -1), (1, 1, 1),
             (1, 0, -1), (1, 0, 1), (2, 0, 0),
             (0, 0, -1), (0, -1, -1), (0, 1, -1), (0, 0, -2),
             (0, 0, 1), (0, -1, 1), (0, 1, 1), (0, 0, 2),
             (0, -1, 0), (0, -2, 0),
             (0, 1, 0), (0, 2, 0)]


@pytest.mark.parametrize("dims", [(5, 4, 1), (3, 3, 1), (5, 4, 5), (3, 3, 3), (6, 2, 4)])
def test_range2_order_is_the_reference_table(oracle, dims):
    X, Y, Z = dims
    table = RANGE2_LAYER if Z == 1 else RANGE2_3D
    V = X * Y * Z
    for s in range(V):
        x, y, z = s % X, (s // X) % Y, s // (X * Y)
        expect = [(x + dx) + X * ((y + dy) + Y * (z + dz)) for dx, dy, dz in table
                  if 0 <= x + dx < X and 0 <= y + dy < Y and 0 <= z + dz < Z]
        expect += [V] if x == 0 else []
        expect += [V + 1] if x == X - 1 else []
        got = oracle.neighbors(dims, 2, s, V, V + 1).tolist()
        assert got == expect, (dims, s)


def test_range2_table_tells_swapped_blocks_apart():
    """The plausible mistakes (the -x and +x blocks swapped; south and north swapped;
    z blocks before the x blocks) give other orders on an interior site."""
    for table, dims in ((RANGE2_LAYER, (5, 5, 1)), (RANGE2_3D, (5, 5, 5))):
        half = 4 if table is RANGE2_LAYER else 10
        variants = [table[half:2 * half] + table[:half] + table[2 * half:],
                    table[:-4] + table[-2:] + table[-4:-2]]
        if table is RANGE2_3D:
            variants.append(table[20:28] + table[:20] + table[28:])
        for v in variants:
            assert sorted(v) == sorted(table) and v != table


# ----------------------------------------------------------------------------- placement
# 4x4x1 transistor, seed.charges 1: int(0.13*16) = 2 defects, int(0.19*16) = 3 electrons,
# 3 holes, int(0.25*16) = 4 traps (seed.percentage 0.5: 2 seeds, 2 grown), in that order.
# A drawn carrier site is valid when its own grid is empty there and it is no defect
# (sourceagent.cpp:204-247); electrons never block holes.  Growth draws a seed, then an
# entry of its range-1 list on the electron grid, which on x = 0 / x = 3 ends with that
# face's electron drain slot (rejected: a Drain).
def scenario_placement(seed, mistake, oracle=None):
    W = Words(seed)
    X = Y = 4

    def nbrs1(s):
        x, y = s % X, s // X
        out = [s - 1] if x > 0 else []
        out += [s + 1] if x < X - 1 else []
        out += [s - X] if y > 0 else []
        out += [s + X] if y < Y - 1 else []
        out += ["drain"] if x == 0 else []
        out += ["drain"] if x == X - 1 else []
        if mistake == "growth list without the drain slot":
            out = [v for v in out if v != "drain"]
        return out

    defects, electrons, holes = [], [], []

    def place_defects():
        while len(defects) < 2:
            s = W.integer(0, 15)
            if s not in defects:
                defects.append(s)

    def place(lst):
        while len(lst) < 3:
            s = W.integer(0, 15)
            if s not in lst and (s not in defects or mistake == "defects not excluded"):
                lst.append(s)

    steps = [place_defects, lambda: place(electrons), lambda: place(holes)]
    if mistake == "holes before electrons":
        steps = [place_defects, lambda: place(holes), lambda: place(electrons)]
    for f in steps:
        f()
    traps = []
    while len(traps) < 2:
        s = W.integer(0, 15)
        if s not in traps:
            traps.append(s)
    grown = 0
    while grown < 2:
        seed_site = traps[W.integer(0, len(traps) - 1)]
        nb = nbrs1(seed_site)
        t = nb[W.integer(0, len(nb) - 1)]
        if t != "drain" and t not in traps:
            traps.append(t)
            grown += 1
    out = dict(defects=defects, e=electrons, h=holes, traps=sorted(traps), k=W.k)
    if oracle is None:
        return out
    w = oracle.World(oracle.OrcParams.make(grid_x=4, grid_y=4, grid_z=1, random_seed=seed, seed_charges=1.0,
                                           defect_percentage=0.13, electron_percentage=0.19,
                                           hole_percentage=0.19, trap_percentage=0.25, seed_percentage=0.5,
                                           trap_potential=0.5, e_source_l_rate=0.9, h_source_l_rate=0.9,
                                           h_drain_r_rate=0.9))
    return w, out


def test_placement_order(oracle):
    mistakes = ["holes before electrons", "defects not excluded", "growth list without the drain slot"]
    for m in mistakes:
        assert any(scenario_placement(s, m) != scenario_placement(s, None) for s in SEEDS), m
    for s in SEEDS:
        w, expect = scenario_placement(s, None, oracle)
        pot = w.potential()
        got = dict(defects=w.defects().tolist(), e=w.carriers(0)["site"].tolist(),
                   h=w.carriers(1)["site"].tolist(), traps=sorted(np.nonzero(pot > 0.25)[0].tolist()),
                   k=w.rng()[1])
        assert got == expect, f"seed {s}"
