"""Pins of the surrogate oracle (oracle/surrogate.py) on the CPU.

* ideal weights (rfftn of the circular kernel on the padded grid) make the
  network reproduce the C++ oracle's closed-form map (orc_potential_map, the
  kernel.cl coulomb1/gauss1 restatement) to rounding -- with and without a
  cutoff and with Gaussian screening; one padding site less than the rule
  breaks it, so the padding rule is pinned too;
* the analytic gradient equals central finite differences of the loss on both
  the real and imaginary parts of weights on interior, kx = 0 and Nyquist
  planes;
* the design argument for keeping every mode: truncating the ideal weights to
  12 modes per axis (an FNO layer's usual corners) leaves an error far above
  the 1e-3 gate on the 128x128x32 solar lattice.
"""
import numpy as np
import pytest

from oracle import surrogate as S

PREF = 0.41141839736460467


def random_charges(rng, V, n):
    sites = rng.choice(V, n, replace=False).astype(np.int32)
    q = rng.choice([-1, 1], n).astype(np.int32)
    return sites, q


@pytest.mark.parametrize("dims,cutoff,sigma,none", [
    ((12, 10, 6), 5, 0.0, False),
    ((9, 7, 5), 30, 0.0, False),        # cutoff above the lattice: P = 2N - 1 (+1)
    ((10, 8, 6), 4, 0.0, True),         # no cutoff
    ((11, 9, 7), 6, 1.3, False),        # gaussian screening
    ((16, 12, 1), 7, 0.0, False),       # flat (2-D) lattice
])
def test_ideal_weights_reproduce_exact_map(oracle, dims, cutoff, sigma, none):
    rng = np.random.default_rng(3)
    V = int(np.prod(dims))
    sites, q = random_charges(rng, V, max(3, V // 25))
    W = S.ideal_weights(dims, cutoff, PREF, sigma, none)
    assert W.shape == S.spectrum_shape(dims, cutoff, none)
    occ = S.occupancy(dims, sites, q)
    got = S.forward(W, occ, dims, cutoff, none)[0]
    flags = 1 if sigma > 0 else 0
    ref = oracle.potential_map(dims, 0 if none else cutoff, sites, q, flags=flags, pref=PREF, sigma=sigma)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) + 1e-13


def test_padding_rule_is_tight(oracle):
    """A circular convolution of period P reproduces the map exactly for
    P = N + c - 1 along x and wraps in-cutoff pairs around for P = N + c - 2."""
    dims, cutoff = (12, 6, 4), 5
    rng = np.random.default_rng(4)
    V = int(np.prod(dims))
    sites, q = random_charges(rng, V, 20)
    occ = S.occupancy(dims, sites, q).astype(np.float64).reshape(4, 6, 12)
    ref = oracle.potential_map(dims, cutoff, sites, q, pref=PREF).reshape(4, 6, 12)
    px, py, pz = S.padded_shape(dims, cutoff)
    assert px >= 12 + cutoff - 1 and px % 8 == 6 and py % 4 == 0 and pz % 4 == 0
    K = S.circular_kernel(dims, cutoff, PREF)
    for p_x, exact in ((12 + cutoff - 1, True), (12 + cutoff - 2, False)):
        Kx = np.zeros((pz, py, p_x))
        for dx in range(-(dims[0] - 1), dims[0]):
            Kx[:, :, dx % p_x] += K[:, :, dx % px]
        shp = (pz, py, p_x)
        out = np.fft.irfftn(np.fft.rfftn(Kx) * np.fft.rfftn(occ, s=shp, axes=(0, 1, 2)), s=shp, axes=(0, 1, 2))
        err = np.max(np.abs(out[:4, :6, :12] - ref))
        assert (err < 1e-12) if exact else (err > 1e-3)


def test_gradient_matches_finite_differences():
    dims, cutoff = (8, 6, 4), 4
    rng = np.random.default_rng(5)
    shape = S.spectrum_shape(dims, cutoff)
    V = int(np.prod(dims))
    W = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.05
    occ = np.stack([S.occupancy(dims, *random_charges(rng, V, 12)) for _ in range(2)])
    tgt = rng.normal(size=(2, V))
    L, G = S.loss_and_grad(W, occ, tgt, dims, cutoff)
    assert abs(L - S.loss(W, occ, tgt, dims, cutoff)) < 1e-12 * L
    pz, py, kx = shape
    picks = [(0, 0, 0), (1, 2, 1), (pz - 1, py - 1, kx - 1), (0, 3, kx - 1), (2, 0, 0), (pz // 2, 1, 2)]
    h = 1e-3   # the loss is quadratic in W: central differences are exact up to rounding
    for k in picks:
        for unit in (1.0, 1j):
            Wp, Wm = W.copy(), W.copy()
            Wp[k] += h * unit
            Wm[k] -= h * unit
            fd = (S.loss(Wp, occ, tgt, dims, cutoff) - S.loss(Wm, occ, tgt, dims, cutoff)) / (2 * h)
            an = G[k].real if unit == 1.0 else G[k].imag
            assert abs(fd - an) <= 1e-8 * (abs(an) + np.max(np.abs(G))), (k, unit, fd, an)


def test_truncated_modes_cannot_meet_the_gate():
    """Why the surrogate keeps every mode: an FNO-style truncation to 12 modes
    per axis of even the ideal operator misses the 1e-3 gate by orders of
    magnitude on the 128x128x32 solar lattice (cutoff 30)."""
    dims, cutoff = (128, 128, 32), 30
    rng = np.random.default_rng(6)
    V = int(np.prod(dims))
    sites, q = random_charges(rng, V, 10484)
    W = S.ideal_weights(dims, cutoff, PREF)
    occ = S.occupancy(dims, sites, q)
    exact = S.forward(W, occ, dims, cutoff)[0]
    trunc = S.forward(S.truncate_modes(W, 12), occ, dims, cutoff)[0]
    err = np.linalg.norm(trunc - exact) / np.linalg.norm(exact)
    assert err > 0.1
    full_params = 2 * W.size
    assert 1.5e6 < full_params < 2.0e6     # 64 x 160 x 80 complex weights


def test_occupancy_sums_coincident_charges():
    occ = S.occupancy((4, 2, 1), [0, 0, 3, 5, 5, 5], [-1, 1, 1, -1, -1, 1])
    assert occ.tolist() == [0, 0, 0, 1, 0, -1, 0, 0]


@pytest.mark.parametrize("dims,cutoff,sigma,no_cut", [((20, 12, 6), 5, 0.0, False), ((16, 16, 8), 30, 0.0, False),
                                                      ((12, 10, 4), 3, 1.3, False), ((9, 7, 5), 0, 0.0, True)])
def test_product_analytic_spectrum_is_the_ideal_operator(dims, cutoff, sigma, no_cut):
    """langmuir_amd.surrogate.analytic_spectrum (the untrained, exact weights the product
    can load) equals the oracle's ideal weights, which reproduce the FP64 map."""
    from langmuir_amd.surrogate import analytic_spectrum
    pref = 0.41141839736460467
    padded = S.padded_shape(dims, cutoff, no_cut)
    got = analytic_spectrum(dims, padded, cutoff, pref, sigma, no_cut)
    ref = S.ideal_weights(dims, cutoff, pref, sigma, no_cut)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
