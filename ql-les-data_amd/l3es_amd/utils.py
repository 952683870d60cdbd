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
    """Shell-binned E(k) (utils.py:35-97): transform and FP64 binning on the GPU (qlles_energy_spectrum).

    Returns a float64 NumPy array of n bins like the reference.
    """
    del is_scalar_field
    if dim == 2:
        if vel.dim() == 3:
            vel = vel[..., None]
        full = torch.zeros((3,) + tuple(vel.shape[1:3]) + (1,), dtype=vel.dtype, device=vel.device)
        full[:2] = vel[:2]
    else:
        full = vel.contiguous()
    n = full.shape[1]
    plan = _lib.Plan(dim, n, full.dtype, 0.0, 1.0)
    ek = torch.empty(n, dtype=torch.float64, device=full.device)
    _lib.check(_lib.load().qlles_energy_spectrum(plan.handle, _lib.ptr(full.contiguous()), float(mul_fac),
                                                 _lib.ptr(ek), _lib.stream_ptr()))
    return ek.cpu().numpy()


def rho_computer(N, dim=3, L=2 * np.pi):
    """Density from coordinates (utils.py:169-218): comp_rho(r) -> rho (np,) FP64
    on the device, rho_i = m sum_j W(|r_i - r_j|) with the quintic kernel of
    smoothing length L/N over every particle within 3 L/N, itself included."""
    ctx = _lib.Sph(dim, N, L, True, 0.0)
    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())

    def comp_rho(r):
        r = torch.as_tensor(r).to(device=dev, dtype=torch.float64).contiguous()
        if r.dim() != 2 or r.shape[1] != dim:
            raise ValueError(f"expected an (np, {dim}) array, got {tuple(r.shape)}")
        rho = torch.empty(r.shape[0], dtype=torch.float64, device=dev)
        _lib.check(lib.qlles_sph_density(ctx.handle, _lib.ptr(r), r.shape[0], _lib.ptr(rho), _lib.stream_ptr()))
        return rho

    comp_rho.ctx = ctx
    return comp_rho
