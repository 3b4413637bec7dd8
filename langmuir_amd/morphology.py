"""Two-phase morphologies for the solar-cell lattice, as LangmuirPython makes them.

``isotropic`` is surface.Isotropic with surface.GaussianKernel
(LangmuirPython/langmuir/surface.py:673-783): uniform noise in [-0.5, 0.5)
(``rfunc``, :306-313) convolved, mode 'same', with a normalised product of
Gaussian densities sampled on the integer grid from -3 sigma to +3 sigma per axis
(:695-708 with Kernel's mgrid, :623-628).  ``trap_sites`` is
IndexMapper.map_mesh (grid.py:395-403): the sites of the upper phase, listed in
numpy.ndenumerate order over (x, y, z) -- x outermost -- with the site id
X (y + Y z) + x.  Arrays are indexed [x][y][z] as in LangmuirPython.

The noise comes from numpy's legacy MT19937 ``RandomState(seed)``, the generator
behind ``numpy.random.random`` that LangmuirPython calls, so a seed reproduces
what ``numpy.random.seed(seed)`` followed by the reference's generator gives.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np


def gaussian_kernel(sx: float, sy: float, sz: float, spacing: float = 1.0) -> np.ndarray:
    """GaussianKernel.v: product of N(0, s) densities on [-3s, 3s] (step ``spacing``),
    normalised to sum 1; a zero sigma gives a single point on that axis."""
    axes = []
    for s in (sx, sy, sz):
        if s <= 0:
            axes.append(np.ones(1))
            continue
        x = np.arange(-3.0 * s, 3.0 * s + 1e-10, spacing)
        axes.append(np.exp(-0.5 * (x / s) ** 2) / (s * np.sqrt(2.0 * np.pi)))
    v = axes[0][:, None, None] * axes[1][None, :, None] * axes[2][None, None, :]
    return v / np.sum(v)


def noise(shape: Sequence[int], seed: Optional[int] = None) -> np.ndarray:
    """rfunc: uniform numbers in [-0.5, 0.5) of the given [x][y][z] shape."""
    return np.random.RandomState(seed).random_sample(tuple(int(n) for n in shape)) - 0.5


def isotropic(shape: Sequence[int], sigma, seed: Optional[int] = None,
              field: Optional[np.ndarray] = None) -> np.ndarray:
    """Isotropic(grid, GaussianKernel(sigma...)).image: noise convolved with the kernel,
    mode 'same' (zero outside the lattice).  ``sigma`` is one value or (sx, sy, sz);
    ``field`` replaces the drawn noise."""
    from scipy import signal
    s = (sigma, sigma, sigma) if np.isscalar(sigma) else tuple(sigma)
    f = noise(shape, seed) if field is None else np.asarray(field, dtype=np.float64)
    return signal.convolve(f, gaussian_kernel(*s), mode="same")


def linear_mapping(a: np.ndarray, n: float = 0.0, m: float = 1.0) -> np.ndarray:
    """surface.linear_mapping: values mapped linearly onto [n, m]."""
    a = np.asarray(a, dtype=np.float64)
    lo, hi = np.amin(a), np.amax(a)
    return (a - lo) / (hi - lo) * (m - n) + n


def threshold(a: np.ndarray, v: float = 0.0) -> np.ndarray:
    """surface.threshold (surface.py:265-294), in order: values <= v become 0, then
    values > 0 become 1 (for v < 0 the values in (v, 0] are left as they are)."""
    t = np.array(a, dtype=np.float64)
    t[t <= v] = 0
    t[t > 0] = 1
    return t


def map_mesh(mesh: np.ndarray, value: float = 1.0) -> np.ndarray:
    """IndexMapper.map_mesh: site ids X (y + Y z) + x of the entries that linear_mapping
    onto [0, 1] sends to ``value``, in numpy.ndenumerate order (x outermost)."""
    mesh = linear_mapping(mesh, 0, 1)
    X, Y, Z = mesh.shape
    xi, yi, zi = np.nonzero(mesh == value)
    return (X * (yi + Y * zi) + xi).astype(np.int32)


def two_phase_traps(dims: Sequence[int], sigma: float = 4.0, seed: Optional[int] = None,
                    fraction: float = 0.5) -> Tuple[np.ndarray, np.ndarray]:
    """A morphology whose upper phase holds ``fraction`` of the sites (the median for one
    half): the image is shifted by that level and passed through threshold(., 0) and
    map_mesh.  Returns (trap site ids in map_mesh order, the smoothed image)."""
    image = isotropic(dims, sigma, seed)
    level = np.median(image) if fraction == 0.5 else np.quantile(image, 1.0 - fraction)
    return map_mesh(threshold(image - level, 0.0)), image


def to_checkpoint(traps: Sequence[int], dims: Sequence[int], trap_potential: float = 0.5):
    """surface.save_chk: a checkpoint holding the lattice size and the trap list, with
    CheckPoint.fix_traps applied (trap.percentage set so that World keeps exactly these
    traps); write it with langmuir_amd.io.write_checkpoint."""
    from . import io
    kv = io.KeyValues()
    kv["grid.x"], kv["grid.y"], kv["grid.z"] = (int(n) for n in dims)
    kv["trap.potential"] = float(trap_potential)
    ck = io.Checkpoint(parameters=kv, traps=[int(s) for s in traps])
    io.fix_traps(ck)
    return ck
