"""Particle -> grid transfers and the point-cloud energy spectrum on the HIP path
(mirror of l3es/turbulence.py:253-377).

``ur_to_u_mls_wrapper`` is the second-order moving least squares of
turbulence.py:145-295 (quintic kernel, h = 2/3 dx) at the cell-centred nodes
(i + 1/2) dx; ``ur_to_u_dft_wrapper`` the exact type-1 sum of
turbulence.py:298-336 followed by ifftn(...).real on the nodes m dx.  Both
return (N^dim, dim) FP64 device tensors with the nodes in 'ij' order;
``u_and_spectrum_from_ur`` reshapes to (dim, N, N, N) / (dim, N, N, 1) and adds
the shell-binned spectrum (utils.energy_spectrum).
"""

import numpy as np
import torch

from . import _lib
from .utils import energy_spectrum


def _dev_f64(a, dim, n_expect=None):
    dev = torch.device("cuda", torch.cuda.current_device())
    t = torch.as_tensor(a).to(device=dev, dtype=torch.float64).contiguous()
    if t.dim() != 2 or t.shape[1] != dim:
        raise ValueError(f"expected an (np, {dim}) array, got {tuple(t.shape)}")
    if n_expect is not None and t.shape[0] != n_expect:
        raise ValueError(f"expected {n_expect} particles, got {t.shape[0]}")
    return t


def ur_to_u_mls_wrapper(N, L, dim=3):
    """Particle velocities -> grid by 2nd-order MLS (turbulence.py:253-295)."""
    ctx = _lib.Sph(dim, N, L, True, 0.0)
    lib = _lib.load()

    def body(r, u_r):
        r = _dev_f64(r, dim, N ** dim)  # turbulence.py:285
        u_r = _dev_f64(u_r, dim, N ** dim)
        out = torch.empty((dim, N ** dim), dtype=torch.float64, device=r.device)
        _lib.check(lib.qlles_sph_mls_grid(ctx.handle, _lib.ptr(r), _lib.ptr(u_r), r.shape[0], _lib.ptr(out),
                                          _lib.stream_ptr()))
        return out.T

    body.ctx = ctx
    return body


def ur_to_u_dft_wrapper(N, L, dim=3, fft_axes=(1, 2, 3)):
    """Particle velocities -> grid by the exact DFT (turbulence.py:298-336).

    N must be an FFT size the solver supports (the inverse transform is the solver's).
    """
    del fft_axes
    ctx = _lib.Sph(dim, N, L, True, 0.0)
    plan = _lib.Plan(dim, N, torch.float64, 0.0, 1.0)
    lib = _lib.load()

    def body(r, u_r):
        r = _dev_f64(r, dim)
        u_r = _dev_f64(u_r, dim, r.shape[0])
        hshape = (3, N, N, N // 2 + 1) if dim == 3 else (3, N, N // 2 + 1)
        uh = torch.empty(hshape, dtype=torch.complex128, device=r.device)
        _lib.check(lib.qlles_sph_dft_hat(ctx.handle, _lib.ptr(r), _lib.ptr(u_r), r.shape[0], _lib.ptr(uh),
                                         _lib.stream_ptr()))
        u = torch.empty((3, N, N, N) if dim == 3 else (3, N, N, 1), dtype=torch.float64, device=r.device)
        _lib.check(lib.qlles_irfftn(plan.handle, _lib.ptr(uh), _lib.ptr(u), _lib.stream_ptr()))
        return u[:dim].reshape(dim, -1).T

    body.ctx, body.plan = ctx, plan
    return body


def u_and_spectrum_from_ur(ckp_N, L, dim=3, interp_backend="mls", fft_axes=(1, 2, 3)):
    """body(r, u_r) -> (u_grid, E(k)) (turbulence.py:339-377)."""
    if interp_backend == "dft":
        interpolate = ur_to_u_dft_wrapper(ckp_N, L, dim, fft_axes)
    else:
        interpolate = ur_to_u_mls_wrapper(ckp_N, L, dim)

    def body(r, u_r):
        u_grid = interpolate(r, u_r)
        shape = (dim, ckp_N, ckp_N, ckp_N) if dim == 3 else (dim, ckp_N, ckp_N, 1)
        u_grid = u_grid.T.reshape(shape).contiguous()
        ek = energy_spectrum(u_grid, dim=dim)
        return u_grid, ek

    return body
