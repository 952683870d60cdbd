"""Grid -> particle interpolation on the HIP path (mirror of l3es/integrator.py)."""

import os

import numpy as np
import torch

from . import _lib
from .relax import relax_wrapper
from .utils import rho_computer


def shift_fn(r, dr, box_size=2 * np.pi):
    """Shift in a periodic box (integrator.py:21-23)."""
    return torch.remainder(r + dr, box_size)


class _Interpolator:
    """interpolate(u, r) -> u_r (integrator.py:36-91) with backend "nufft" or "dft".

    ``interpolate(u, r)`` takes the physical n-grid field like the reference;
    ``from_hat(u_hat, r, dt)`` is the fused path used by ``combined``: it
    applies spectral_filtering to the solver's N-grid spectrum, interpolates,
    and (dt != 0) advances r <- (r + dt u_r) mod 2 pi in place.
    """

    def __init__(self, N, fft_axes, backend, precision, tol):
        self.n = N
        self.dim = len(fft_axes)
        self.dtype = torch.float32 if precision in ("float32", torch.float32) else torch.float64
        self.backend = backend
        self.it = _lib.Interp(self.dim, N, self.dtype, backend, tol)

    def __call__(self, u, r):
        u = u.to(self.dtype).contiguous()
        r = r.to(torch.float64).contiguous()
        ur = torch.empty((r.shape[0], self.dim), dtype=self.dtype, device=r.device)
        _lib.check(_lib.load().qlles_interp_field(self.it.handle, _lib.ptr(u), _lib.ptr(r), r.shape[0],
                                                  _lib.ptr(ur), _lib.stream_ptr()))
        return ur

    def from_hat(self, u_hat, r, dt=0.0, u_r=None):
        N = u_hat.shape[1]
        if u_r is None:
            u_r = torch.empty((r.shape[0], self.dim), dtype=self.dtype, device=r.device)
        _lib.check(_lib.load().qlles_interp_hat(self.it.handle, N, _lib.ptr(u_hat), _lib.ptr(r), r.shape[0],
                                                _lib.ptr(u_r), float(dt), _lib.stream_ptr()))
        return u_r


def spectral_interpolator_wrapper(N, fft_axes=(1, 2, 3), splits=128, backend="nufft", precision="float64",
                                  tol=1e-6):
    """Interpolator factory (integrator.py:26-93); backend in {"dft", "nufft"}."""
    assert splits & (splits - 1) == 0, "Splits must be a power of 2"  # integrator.py:33
    assert backend in ["dft", "nufft"]  # integrator.py:34
    return _Interpolator(N, fft_axes, backend, precision, tol)


def pos_init_cartesian(n, dim, L=2 * np.pi):
    """Cell-centred lattice (i + 1/2) dx in 'ij' order, (n^dim, dim) FP64."""
    dx = L / n
    g = np.meshgrid(*([np.arange(n)] * dim), indexing="ij")
    return (np.stack([a.ravel() for a in g], axis=1) + 0.5) * dx


def pos_init_jittered(n, dim, L=2 * np.pi, jitter=0.25, seed=0):
    """Cell-centred lattice with uniform jitter of +- jitter*dx, wrapped into [0, L)."""
    rng = np.random.default_rng(seed)
    r = pos_init_cartesian(n, dim, L)
    r = r + (rng.random(r.shape) * 2 - 1) * jitter * (L / n)
    return r % L


def set_up_integrator(state_0_path, dim, N, dft_splits, u_ref, interp_backend, precision="float64", tol=1e-6,
                      jitter=0.0, seed=0):
    """Particles, density, interpolator and relaxation closures (integrator.py:96-114).

    Relaxed SPH states (.h5 via jax_sph) are not read here; a ``.npy`` file of
    (np, dim) positions may be given instead.  comp_rho = rho_computer(N, dim,
    L) and (relax_fn, sph_fn) = relax_wrapper(N, dim, L, is_physical=True,
    u_ref) as in the reference; they are None when N < 7 (the kernel support
    3 L/N must stay below L/2) or, for the relaxation, when u_ref is None.
    """
    L = 2 * np.pi
    fft_axes = (1, 2, 3) if dim == 3 else (1, 2)
    if state_0_path is not None and os.path.exists(state_0_path) and state_0_path.endswith(".npy"):
        r = np.load(state_0_path)
    elif jitter > 0:
        r = pos_init_jittered(N, dim, L, jitter, seed)
    else:
        r = pos_init_cartesian(N, dim, L)
    r = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).to(torch.device("cuda", torch.cuda.current_device()))
    comp_rho = rho_computer(N, dim=dim, L=L) if N >= 7 else None
    interpolator = spectral_interpolator_wrapper(N, fft_axes, splits=dft_splits or 1, backend=interp_backend,
                                                 precision=precision, tol=tol)
    relax_fn = sph_fn = None
    if N >= 7 and u_ref is not None:
        relax_fn, sph_fn = relax_wrapper(N, dim, L, is_physical=True, u_ref=u_ref)
    return r, comp_rho, interpolator, relax_fn, sph_fn, L, fft_axes
