"""Pins of the oracle's pair sums (Potential::coulomb*/gauss*/coulombImage*,
src/langmuirCore/potential.cpp:351-753) and of its FP64 map (kernel.cl coulomb1).

Each check is fixed by mathematics, not by the oracle's own formula:
closed forms for one charge, the vanishing potential on the bisecting plane of a
+/- pair, linearity in the charges, the exact cutoff edge (r^2 = c^2 - 1 in,
r^2 = c^2 out), the mirror-charge construction of the image term, and an
independent numpy brute force on small lattices.
"""
import math

import numpy as np
import pytest

PREF = 0.41141839736460467   # electrostatic.prefactor at the reference defaults


def site(X, Y, x, y, z):
    return X * (y + z * Y) + x


def xyz(X, Y, s):
    z = s // (X * Y)
    return s % X, s // X - z * Y, z


def brute(dims, cutoff, cs, cq, ts, image=False, sigma=0.0):
    X, Y, Z = dims
    out = []
    for t in ts:
        tx, ty, tz = xyz(X, Y, t)
        v = 0.0
        for s, q in zip(cs, cq):
            x, y, z = xyz(X, Y, s)
            dx = abs(tx + x) + 1 if image else tx - x
            r2 = dx * dx + (ty - y) ** 2 + (tz - z) ** 2
            if 0 < r2 < cutoff * cutoff:
                r = math.sqrt(r2)
                f = q / r
                if sigma > 0:
                    f *= math.erf(r / (math.sqrt(2.0) * sigma))
                v += -f if image else f
        out.append(PREF * v)
    return np.array(out)


def test_prefactor_and_beta(oracle):
    assert oracle.prefactor() == pytest.approx(PREF, rel=1e-15)
    assert oracle.inverse_kt() == pytest.approx(38.68168364706952, rel=1e-15)


@pytest.mark.parametrize("q", [-1, 1, 2])
def test_single_charge_closed_form(oracle, q):
    dims = (20, 20, 20)
    c = site(20, 20, 3, 4, 5)
    targets = [site(20, 20, 3 + a, 4 + b, 5 + d) for a, b, d in [(1, 0, 0), (3, 4, 0), (2, 3, 6), (0, 0, 9)]]
    got = oracle.pair_potential(dims, 50, [c], [q], targets)
    want = [PREF * q / d for d in (1.0, 5.0, 7.0, 9.0)]
    np.testing.assert_allclose(got, want, rtol=1e-14)
    # the charge's own site contributes nothing (iR[0][0][0] = 0)
    assert oracle.pair_potential(dims, 50, [c], [q], [c])[0] == 0.0


def test_dipole_bisecting_plane_is_zero(oracle):
    X = Y = Z = 16
    plus, minus = site(X, Y, 4, 8, 8), site(X, Y, 10, 8, 8)
    targets = [site(X, Y, 7, y, z) for y in range(0, 16, 3) for z in range(0, 16, 5)]
    got = oracle.pair_potential((X, Y, Z), 50, [plus, minus], [1, -1], targets)
    np.testing.assert_allclose(got, 0.0, atol=1e-15)


def test_linearity(oracle):
    rng = np.random.default_rng(0)
    dims = (12, 10, 8)
    V = 12 * 10 * 8
    cs = rng.choice(V, 60, replace=False)
    qa, qb = rng.choice([-1, 1], 60), rng.choice([-1, 1], 60)
    ts = rng.choice(V, 30, replace=False)
    a = oracle.pair_potential(dims, 6, cs, qa, ts)
    b = oracle.pair_potential(dims, 6, cs, qb, ts)
    ab = oracle.pair_potential(dims, 6, cs, qa + qb, ts)
    np.testing.assert_allclose(a + b, ab, atol=1e-13)


@pytest.mark.parametrize("cutoff,inside,outside", [
    (5, (4, 2, 2), (3, 4, 0)),        # r2 = 24 in, r2 = 25 = c^2 out
    (7, (4, 4, 4), (6, 3, 2)),        # r2 = 48 in, r2 = 49 = c^2 out
])
def test_cutoff_edge_exact(oracle, cutoff, inside, outside):
    X = Y = Z = 20
    c = site(X, Y, 0, 0, 0)
    tin = site(X, Y, *inside)
    tout = site(X, Y, *outside)
    r2in = sum(v * v for v in inside)
    assert r2in == cutoff * cutoff - 1
    got = oracle.pair_potential((X, Y, Z), cutoff, [c], [1], [tin, tout])
    assert got[0] == pytest.approx(PREF / math.sqrt(r2in), rel=1e-14)
    assert got[1] == 0.0


def test_image_is_mirror_charge(oracle):
    """Left-electrode image: mirror x -> -x-1 (plane x = -1/2) with opposite sign."""
    X, Y, Z = 16, 8, 8
    cs = [site(X, Y, 2, 3, 4), site(X, Y, 9, 1, 1)]
    cq = [1, -1]
    ts = [site(X, Y, 0, 3, 4), site(X, Y, 5, 5, 5)]
    got = oracle.pair_potential((X, Y, Z), 50, cs, cq, ts, flags=oracle.IMAGE)
    want = []
    for t in ts:
        tx, ty, tz = xyz(X, Y, t)
        v = 0.0
        for s, q in zip(cs, cq):
            x, y, z = xyz(X, Y, s)
            # mirror charge -q at (-x-1, y, z)
            r = math.sqrt((tx - (-x - 1)) ** 2 + (ty - y) ** 2 + (tz - z) ** 2)
            v += -q / r
        want.append(PREF * v)
    np.testing.assert_allclose(got, want, rtol=1e-13)


@pytest.mark.parametrize("flags,sigma,cutoff", [(0, 0.0, 5), (0, 0.0, 50), (1, 1.5, 6), (2, 0.0, 7)])
def test_pair_potential_brute_force(oracle, flags, sigma, cutoff):
    rng = np.random.default_rng(flags * 10 + cutoff)
    dims = (9, 7, 6)
    V = 9 * 7 * 6
    cs = rng.choice(V, 50, replace=False)
    cq = rng.choice([-1, 1, 2], 50)
    ts = rng.choice(V, 40, replace=False)
    got = oracle.pair_potential(dims, cutoff, cs, cq, ts, flags=flags, sigma=sigma)
    want = brute(dims, cutoff, cs, cq, ts, image=bool(flags & 2), sigma=sigma)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("cutoff", [0, 4])
def test_map_brute_force(oracle, cutoff):
    rng = np.random.default_rng(cutoff)
    dims = (8, 6, 5)
    V = 8 * 6 * 5
    cs = rng.choice(V, 30, replace=False)
    cq = rng.choice([-1, 1], 30)
    got = oracle.potential_map(dims, cutoff, cs, cq, threads=2)
    want = brute(dims, cutoff if cutoff > 0 else 10 ** 6, cs, cq, np.arange(V))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_map_equals_pair_potential_on_all_sites(oracle):
    rng = np.random.default_rng(5)
    dims = (10, 9, 4)
    V = 360
    cs = rng.choice(V, 40, replace=False)
    cq = rng.choice([-1, 1], 40)
    m = oracle.potential_map(dims, 5, cs, cq, threads=1)
    p = oracle.pair_potential(dims, 5, cs, cq, np.arange(V))
    np.testing.assert_allclose(m, p, rtol=1e-13, atol=1e-15)
