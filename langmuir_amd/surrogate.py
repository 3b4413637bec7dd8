"""Learned surrogate of the full-lattice potential map (``lmg_surrogate_*``).

The exact map (``Context.potential_map``; the reference's coulomb1/gauss1
kernel, kernel.cl:56-220, driven by OpenClHelper::launchCoulombKernel1,
openclhelper.cpp:220-370) is a convolution of the charge grid with the
truncated Coulomb kernel.  :class:`Surrogate` learns that operator as one
Fourier layer on a zero-padded grid (every mode kept), trains it on the GPU
from (occupancy, exact map) snapshots, and serves maps behind an error gate
whose judge is the exact pair sum, with the exact map as fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._abi import LmgError
from .device import DeviceBuffer
from .host import Context, Simulation


def analytic_spectrum(dims: Sequence[int], padded: Sequence[int], cutoff: int, prefactor: float,
                      sigma: float = 0.0, no_cutoff: bool = False) -> np.ndarray:
    """The weights that make the surrogate the exact map, without training: the
    half-spectrum (numpy's rfftn over z, y, x) of the map's kernel laid out
    circularly on the padded grid.  Index k of a padded axis of length P stands for
    the offset k (k < N) or k - P (k > P - N) of a lattice axis of length N; the
    kernel there is prefactor * f(r) with f = 1/r, times erf(r / (sqrt(2) sigma))
    when sigma > 0, over 0 < r^2 < cutoff^2 (every r > 0 with no_cutoff): the sum
    kernel.cl's coulomb1/gauss1 evaluate (kernel.cl:56-220)."""
    X, Y, Z = (int(n) for n in dims)
    px, py, pz = (int(n) for n in padded)

    def offsets(n, p):
        k = np.arange(p)
        d = np.where(k < n, k, k - p)
        return d, (k < n) | (k > p - n)

    ox, mx = offsets(X, px)
    oy, my = offsets(Y, py)
    oz, mz = offsets(Z, pz)
    r2 = (oz[:, None, None] ** 2 + oy[None, :, None] ** 2 + ox[None, None, :] ** 2).astype(np.int64)
    keep = mz[:, None, None] & my[None, :, None] & mx[None, None, :] & (r2 > 0)
    if not no_cutoff:
        keep &= r2 < int(cutoff) ** 2
    r = np.sqrt(r2[keep].astype(np.float64))
    f = 1.0 / r
    if sigma > 0:
        erf = np.frompyfunc(math.erf, 1, 1)
        f = f * erf(r / (math.sqrt(2.0) * sigma)).astype(np.float64)
    kern = np.zeros((pz, py, px))
    kern[keep] = prefactor * f
    return np.fft.rfftn(kern, axes=(0, 1, 2))


class Surrogate:
    """Spectral surrogate for the map of ``ctx``'s lattice and cutoff.

    ``flags`` selects the operator it approximates (``LMG_GAUSS``,
    ``LMG_NO_CUTOFF``, as for ``Context.potential_map``); ``max_batch`` bounds
    the batch of training and inference calls.
    """

    def __init__(self, ctx: Context, flags: int = 0, max_batch: int = 8):
        L = _abi.lib()
        h = C.c_void_p()
        rc = L.lmg_surrogate_create(ctx.handle, int(flags), int(max_batch), C.byref(h))
        if rc != 0:
            raise LmgError(rc, L.lmg_surrogate_last_error(None).decode())
        self._h = h
        self.ctx = ctx
        self.flags = int(flags)
        self.max_batch = int(max_batch)
        self.dims = ctx.dims
        self.V = ctx.V
        pad = (C.c_int32 * 3)()
        nw = C.c_int64()
        self._chk(L.lmg_surrogate_shape(h, pad, C.byref(nw)))
        self.padded = (int(pad[0]), int(pad[1]), int(pad[2]))
        self.n_weights = int(nw.value)

    # -- plumbing ------------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            raise LmgError(rc, _abi.lib().lmg_surrogate_last_error(self._h).decode())

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib().lmg_surrogate_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def spectrum_shape(self) -> Tuple[int, int, int]:
        """(Pz, Py, Px // 2 + 1): the [kz][ky][kx] layout of the weights."""
        px, py, pz = self.padded
        return (pz, py, px // 2 + 1)

    # -- weights ---------------------------------------------------------------
    def get_weights(self) -> np.ndarray:
        w = np.empty(2 * self.n_weights, np.float32)
        self._chk(_abi.lib().lmg_surrogate_get_weights(self._h, w.ctypes.data_as(C.c_void_p)))
        return (w[0::2] + 1j * w[1::2]).astype(np.complex64).reshape(self.spectrum_shape)

    def set_weights(self, w: Optional[np.ndarray]) -> None:
        """Complex weights of spectrum_shape; None zeroes them."""
        if w is None:
            self._chk(_abi.lib().lmg_surrogate_set_weights(self._h, None))
            return
        w = np.asarray(w).reshape(-1)
        if w.size != self.n_weights:
            raise ValueError(f"expected {self.n_weights} complex weights, got {w.size}")
        flat = np.empty(2 * self.n_weights, np.float32)
        flat[0::2], flat[1::2] = w.real, w.imag
        self._chk(_abi.lib().lmg_surrogate_set_weights(self._h, flat.ctypes.data_as(C.c_void_p)))

    def set_analytic_weights(self) -> None:
        """Load the exact operator's spectrum (:func:`analytic_spectrum`) instead of
        trained weights: the same pipeline, with nothing learned."""
        p = self.ctx.params
        self.set_weights(analytic_spectrum(self.dims, self.padded, p.cutoff, p.prefactor, p.gaussian_sigma,
                                           bool(self.flags & _abi.LMG_NO_CUTOFF)))

    # -- device-pointer entry points ---------------------------------------------
    def occupancy_dev(self, site_ptr: int, q_ptr: int, n: int, occ_ptr: int) -> None:
        self._chk(_abi.lib().lmg_surrogate_occupancy_dev(self._h, site_ptr, q_ptr, int(n), occ_ptr))

    def infer_dev(self, occ_ptr: int, batch: int, map_ptr: int) -> None:
        self._chk(_abi.lib().lmg_surrogate_infer_dev(self._h, occ_ptr, int(batch), map_ptr))

    def gradient_dev(self, occ_ptr: int, target_ptr: int, batch: int, grad_ptr: int) -> float:
        loss = C.c_double()
        self._chk(_abi.lib().lmg_surrogate_gradient_dev(self._h, occ_ptr, target_ptr, int(batch), grad_ptr,
                                                        C.byref(loss)))
        return float(loss.value)

    def train_dev(self, occ_ptr: int, target_ptr: int, n_snapshots: int, batch: int = 8,
                  iterations: int = 300, momentum: float = 0.9, seed: int = 1) -> np.ndarray:
        """Momentum SGD on the snapshots; returns the per-step minibatch losses."""
        hist = np.zeros(max(int(iterations), 1), np.float32)
        self._chk(_abi.lib().lmg_surrogate_train_dev(self._h, occ_ptr, target_ptr, int(n_snapshots), int(batch),
                                                     int(iterations), float(momentum), C.c_uint64(seed),
                                                     hist.ctypes.data_as(C.c_void_p)))
        return hist[:int(iterations)]

    def map_dev(self, site_ptr: int, q_ptr: int, n: int, map_ptr: int, tolerance: float = 1e-3,
                gate_sites: int = 1024, gate_seed: int = 1, wait: bool = True):
        """Gated map; returns (sampled relative RMS error, exact fallback used),
        or None with wait=False (asynchronous; see :meth:`last_gate`)."""
        if not wait:
            self._chk(_abi.lib().lmg_surrogate_map_dev(self._h, site_ptr, q_ptr, int(n), map_ptr,
                                                       float(tolerance), int(gate_sites), C.c_uint64(gate_seed),
                                                       None, None))
            return None
        err = C.c_double()
        used = C.c_int32()
        self._chk(_abi.lib().lmg_surrogate_map_dev(self._h, site_ptr, q_ptr, int(n), map_ptr, float(tolerance),
                                                   int(gate_sites), C.c_uint64(gate_seed), C.byref(err),
                                                   C.byref(used)))
        return float(err.value), bool(used.value)

    def map_batch_dev(self, site_ptr: int, q_ptr: int, counts: Sequence[int], map_ptr: int,
                      tolerance: float = 1e-3, gate_sites: int = 1024, gate_seed: int = 1, wait: bool = True):
        """A batch of gated maps: map b of the charges counts[0] + ... + counts[b-1]
        onwards (counts[b] of them) into map_ptr + 4 b V.  Returns (errors, used
        exact) per map, or None with wait=False (see :meth:`last_gates`)."""
        c = np.ascontiguousarray(np.asarray(counts, np.int32))
        b = len(c)
        pc = c.ctypes.data_as(_abi.PI32)
        if not wait:
            self._chk(_abi.lib().lmg_surrogate_map_batch_dev(self._h, site_ptr, q_ptr, pc, b, map_ptr,
                                                             float(tolerance), int(gate_sites),
                                                             C.c_uint64(gate_seed), None, None))
            return None
        err = np.zeros(b, np.float64)
        used = np.zeros(b, np.int32)
        self._chk(_abi.lib().lmg_surrogate_map_batch_dev(self._h, site_ptr, q_ptr, pc, b, map_ptr, float(tolerance),
                                                         int(gate_sites), C.c_uint64(gate_seed),
                                                         err.ctypes.data_as(_abi.PF64),
                                                         used.ctypes.data_as(_abi.PI32)))
        return err, used.astype(bool)

    def last_gates(self, batch: int):
        err = np.zeros(int(batch), np.float64)
        used = np.zeros(int(batch), np.int32)
        self._chk(_abi.lib().lmg_surrogate_last_gates(self._h, int(batch), err.ctypes.data_as(_abi.PF64),
                                                      used.ctypes.data_as(_abi.PI32)))
        return err, used.astype(bool)

    def set_graphs(self, on: bool) -> None:
        """Replay the gated map as a recorded HIP graph (default) or launch
        every kernel each call."""
        self._chk(_abi.lib().lmg_surrogate_set_graphs(self._h, int(bool(on))))

    def last_gate(self) -> Tuple[float, bool]:
        err = C.c_double()
        used = C.c_int32()
        self._chk(_abi.lib().lmg_surrogate_last_gate(self._h, C.byref(err), C.byref(used)))
        return float(err.value), bool(used.value)

    # -- host conveniences -----------------------------------------------------
    def occupancy(self, sites: Sequence[int], charges: Sequence[int]) -> np.ndarray:
        s = DeviceBuffer.from_array(np.asarray(sites, np.int32))
        q = DeviceBuffer.from_array(np.asarray(charges, np.int32))
        occ = DeviceBuffer(self.V)
        self.occupancy_dev(s.ptr, q.ptr, len(sites), occ.ptr)
        self.ctx.sync()
        return occ.download(np.int8, self.V)

    def infer(self, occ: np.ndarray) -> np.ndarray:
        occ = np.ascontiguousarray(occ, np.int8).reshape(-1, self.V)
        d_occ = DeviceBuffer.from_array(occ)
        d_map = DeviceBuffer(occ.size * 4)
        self.infer_dev(d_occ.ptr, occ.shape[0], d_map.ptr)
        self.ctx.sync()
        return d_map.download(np.float32, occ.size).reshape(occ.shape)

    def gradient(self, occ: np.ndarray, target: np.ndarray) -> Tuple[np.ndarray, float]:
        """(dL/dW as complex spectrum_shape, L) at the current weights."""
        occ = np.ascontiguousarray(occ, np.int8).reshape(-1, self.V)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, self.V)
        d_occ, d_t = DeviceBuffer.from_array(occ), DeviceBuffer.from_array(target)
        d_g = DeviceBuffer(8 * self.n_weights)
        loss = self.gradient_dev(d_occ.ptr, d_t.ptr, occ.shape[0], d_g.ptr)
        g = d_g.download(np.float32, 2 * self.n_weights)
        return (g[0::2] + 1j * g[1::2]).reshape(self.spectrum_shape), loss

    def map(self, sites, charges, tolerance: float = 1e-3, gate_sites: int = 1024,
            gate_seed: int = 1) -> Tuple[np.ndarray, float, bool]:
        s = DeviceBuffer.from_array(np.asarray(sites, np.int32))
        q = DeviceBuffer.from_array(np.asarray(charges, np.int32))
        out = DeviceBuffer(4 * self.V)
        err, used = self.map_dev(s.ptr, q.ptr, len(sites), out.ptr, tolerance, gate_sites, gate_seed)
        return out.download(np.float32, self.V), err, used


class SnapshotSet:
    """Training pairs on the device: int8 occupancy grids and their exact maps
    (labelled by ``Context.potential_map_dev`` with the surrogate's flags)."""

    def __init__(self, surrogate: Surrogate, capacity: int):
        self.sur = surrogate
        self.V = surrogate.V
        self.capacity = int(capacity)
        self.n = 0
        # grids at V-byte strides; the occupancy kernel writes whole words, so a
        # grid whose start is not 4-byte aligned is staged and copied
        self.occ = DeviceBuffer(self.V * self.capacity + 4)
        self.target = DeviceBuffer(4 * self.V * self.capacity)
        self._stage = DeviceBuffer(self.V + 4)

    def add(self, sites, charges) -> None:
        if self.n >= self.capacity:
            raise ValueError("snapshot set is full")
        sites = np.asarray(sites, np.int32)
        charges = np.asarray(charges, np.int32)
        s = DeviceBuffer.from_array(sites)
        q = DeviceBuffer.from_array(charges)
        off = self.n * self.V
        if off % 4 == 0:
            self.sur.occupancy_dev(s.ptr, q.ptr, len(sites), self.occ.at(off))
        else:
            self.sur.occupancy_dev(s.ptr, q.ptr, len(sites), self._stage.ptr)
            self.sur.ctx.sync()
            self.occ.upload(self._stage.download(np.int8, self.V), off)
        self.sur.ctx.potential_map_dev(s.ptr, q.ptr, len(sites), self.target.at(4 * off), self.sur.flags)
        self.sur.ctx.sync()
        self.n += 1

    def add_simulation(self, sim: Simulation, count: int, stride: int, defects: Iterable[int] = (),
                       defects_charge: int = 0) -> None:
        """count snapshots of a running simulation, stride steps apart: electrons
        -1, holes +1, defects defects_charge (coulomb.carriers charges)."""
        defects = np.asarray(list(defects), np.int32)
        for _ in range(int(count)):
            sim.step(int(stride))
            st = sim.state()
            sites = np.concatenate([st["e_site"], st["h_site"], defects]).astype(np.int32)
            q = np.concatenate([np.full(len(st["e_site"]), -1), np.full(len(st["h_site"]), 1),
                                np.full(len(defects), defects_charge)]).astype(np.int32)
            self.add(sites, q)

    def train(self, batch: int = 8, iterations: int = 300, momentum: float = 0.9, seed: int = 1,
              first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """Train on snapshots [first, first + count)."""
        count = self.n - first if count is None else int(count)
        return self.sur.train_dev(self.occ.at(first * self.V), self.target.at(4 * first * self.V), count,
                                  batch, iterations, momentum, seed)

    def targets(self, i: int) -> np.ndarray:
        return self.target.download(np.float32, self.V, 4 * i * self.V)

    def occupancy(self, i: int) -> np.ndarray:
        return self.occ.download(np.int8, self.V, i * self.V)
