"""Spectral utilities on the HIP path (mirror of the hot-path parts of l3es/utils.py)."""

import numpy as np
import torch

from . import _lib

EPS = float(np.finfo(float).eps)


def _plan_for(u_hat, dim, nu=0.0):
    N = u_hat.shape[1]
    dtype = torch.float32 if u_hat.dtype == torch.complex64 else torch.float64
    return _lib.Plan(dim, N, dtype, nu, 1.0)


def spectral_filtering_hat(u_hat, ckp_N, plan):
    """spectral_filtering(irfftn(u_hat), ckp_N) (utils.py:100-138) straight from the spectrum.

    Returns (3, n, n, n) in 3-D and (2, n, n) in 2-D, like the reference.
    """
    shape = (3, ckp_N, ckp_N, ckp_N) if plan.dim == 3 else (2, ckp_N, ckp_N)
    out = torch.empty(shape, dtype=plan.dtype, device=u_hat.device)
    _lib.check(_lib.load().qlles_filter(plan.handle, _lib.ptr(u_hat), int(ckp_N), _lib.ptr(out),
                                        _lib.stream_ptr()))
    return out


def spectral_filtering(u, ckp_N):
    """Reference signature: physical (3,N,N,N) or (2,N,N) field -> (dim, n, ...) field."""
    dim = u.shape[0]
    N = u.shape[1]
    dtype = u.dtype
    plan = _lib.Plan(dim, N, dtype, 0.0, 1.0)
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    if dim == 3:
        full = u.contiguous()
        u_hat = torch.empty((3, N, N, N // 2 + 1), dtype=cdt, device=u.device)
    else:
        full = torch.zeros((3, N, N, 1), dtype=dtype, device=u.device)
        full[:2, :, :, 0] = u
        u_hat = torch.empty((3, N, N // 2 + 1), dtype=cdt, device=u.device)
    _lib.check(_lib.load().qlles_rfftn(plan.handle, _lib.ptr(full), _lib.ptr(u_hat), _lib.stream_ptr()))
    return spectral_filtering_hat(u_hat, ckp_N, plan)


def comp_dissipation_rate(u_hat, nu):
    """nu <|curl u|^2> through irfftn of i k x u_hat (utils.py:300-315), FP64 reduction."""
    dim = 3 if u_hat.dim() == 4 else 2
    plan = _plan_for(u_hat, dim, nu)
    out = torch.zeros(5, dtype=torch.float64, device=u_hat.device)
    _lib.check(_lib.load().qlles_dissipation(plan.handle, _lib.ptr(u_hat), _lib.ptr(out), _lib.stream_ptr()))
    return float(out[0])


def kinetic_energy_hat(u_hat, kf=0):
    """FP64 Parseval energy 0.5 <|u|^2> of u_hat and of its 0 < |k| <= kf shell."""
    dim = 3 if u_hat.dim() == 4 else 2
    N = u_hat.shape[1]
    dtype = torch.float32 if u_hat.dtype == torch.complex64 else torch.float64
    plan = _lib.Plan(dim, N, dtype, 0.0, 1.0, kf=kf)
    out = torch.zeros(2, dtype=torch.float64, device=u_hat.device)
    _lib.check(_lib.load().qlles_energy_hat(plan.handle, _lib.ptr(u_hat), _lib.ptr(out), _lib.stream_ptr()))
    return float(out[0]), float(out[1])


def energy_spectrum(vel, mul_fac=1.0, is_scalar_field=False, dim=3):
    """Shell-binned E(k) (utils.py:35-97).

    The transform runs on the GPU; the shell binning (a diagnostic, n bins)
    runs on the host in FP64 with the reference's weights.
    """
    del is_scalar_field
    if dim == 2:
        v = vel[:2, :, :, 0] if vel.dim() == 4 else vel[:2]
        full = torch.zeros((3,) + tuple(v.shape[1:]) + (1,), dtype=v.dtype, device=v.device)
        full[:2, :, :, 0] = v
        n = v.shape[1]
        shape_hat = (3, n, n // 2 + 1)
    else:
        full = vel.contiguous()
        n = vel.shape[1]
        shape_hat = (3, n, n, n // 2 + 1)
    plan = _lib.Plan(dim, n, full.dtype, 0.0, 1.0)
    cdt = torch.complex64 if full.dtype == torch.float32 else torch.complex128
    vh = torch.empty(shape_hat, dtype=cdt, device=full.device)
    _lib.check(_lib.load().qlles_rfftn(plan.handle, _lib.ptr(full), _lib.ptr(vh), _lib.stream_ptr()))
    a = (vh.abs() ** 2).sum(0).double().cpu().numpy() / float(n ** dim) ** 2
    k = np.fft.fftfreq(n, 1.0 / n)
    kh = np.abs(k[: n // 2 + 1])
    if dim == 3:
        K = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kh[None, None, :] ** 2)
        fact = np.broadcast_to(np.where((kh > 0) & (kh < n // 2), 2, 1)[None, None, :], K.shape)
    else:
        K = np.sqrt(k[:, None] ** 2 + kh[None, :] ** 2)
        fact = np.broadcast_to(np.where((kh > 0) & (kh < n // 2), 2, 1)[None, :], K.shape)
    shell = (K + 0.5).astype(int).ravel()
    n_samples = np.bincount(shell, weights=fact.ravel(), minlength=n)[:n]
    ek = np.bincount(shell, weights=(a * fact * mul_fac).ravel(), minlength=n)[:n]
    return ek * 4 * np.pi * np.arange(n) ** 2 / (n_samples + EPS)
