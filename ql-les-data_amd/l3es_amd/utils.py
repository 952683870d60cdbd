"""Spectral utilities on the HIP path (mirror of the hot-path parts of l3es/utils.py)."""

from collections import OrderedDict

import numpy as np
import torch

from . import _lib

EPS = float(np.finfo(float).eps)

# helper plans of the spectral utilities, one per (device, dim, N, dtype, nu,
# kf), most recently used last: a helper plan holds only the reduction
# partials and, once a transform ran on it, the transform scratch, so
# repeated diagnostics allocate nothing after their first call
_HELPERS = OrderedDict()
_HELPERS_MAX = 4


def helper_plan(dim, N, dtype, nu=0.0, kf=0):
    key = (torch.cuda.current_device(), int(dim), int(N), dtype, float(nu), int(kf or 0))
    plan = _HELPERS.pop(key, None)
    if plan is None:
        plan = _lib.Plan(dim, N, dtype, nu, 1.0, kf=kf)
    _HELPERS[key] = plan
    while len(_HELPERS) > _HELPERS_MAX:
        _HELPERS.popitem(last=False)
    return plan


def _real_dtype(t):
    if t.dtype in (torch.float32, torch.complex64):
        return torch.float32
    if t.dtype in (torch.float64, torch.complex128):
        return torch.float64
    raise ValueError(f"unsupported dtype {t.dtype}")


def spectrum_dim(u_hat):
    """dim of a solver spectrum: (3, N, N, N/2+1) -> 3, (3, N, N/2+1) -> 2; else ValueError."""
    if not torch.is_tensor(u_hat) or not u_hat.is_complex():
        raise ValueError("expected a complex torch tensor spectrum")
    N = u_hat.shape[1] if u_hat.dim() >= 2 else 0
    if u_hat.dim() == 4 and tuple(u_hat.shape) == (3, N, N, N // 2 + 1):
        return 3
    if u_hat.dim() == 3 and tuple(u_hat.shape) == (3, N, N // 2 + 1):
        return 2
    raise ValueError(f"expected a (3, N, N, N/2+1) or (3, N, N/2+1) spectrum, got {tuple(u_hat.shape)}")


def physical_field(u, dim=None):
    """(dim, N, full) for a physical velocity field: 3-D (3, N, N, N); 2-D
    (2, N, N), (2, N, N, 1) or the solver's (3, N, N, 1) with component 2 zero.  ``full``
    is the contiguous (3, N, N, N) or (3, N, N, 1) tensor the library reads.
    Any other shape raises ValueError (the library would read past the array)."""
    if not torch.is_tensor(u) or u.is_complex():
        raise ValueError("expected a real torch tensor field")
    shp = tuple(u.shape)
    N = shp[1] if len(shp) >= 2 else 0
    if len(shp) == 4 and shp == (3, N, N, N) and dim in (None, 3) and N > 1:
        return 3, N, u.contiguous()
    if len(shp) == 4 and shp == (3, N, N, 1) and dim in (None, 2):
        return 2, N, u.contiguous()
    if shp in ((2, N, N), (2, N, N, 1)) and dim in (None, 2) and N > 1:
        full = torch.zeros((3, N, N, 1), dtype=u.dtype, device=u.device)
        full[:2, :, :, 0] = u.reshape(2, N, N)
        return 2, N, full
    raise ValueError(f"expected a (3, N, N, N), (3, N, N, 1) or (2, N, N) field, got {shp}")


def spectral_filtering_hat(u_hat, ckp_N, plan):
    """spectral_filtering(irfftn(u_hat), ckp_N) (utils.py:100-138) straight from the spectrum.

    Returns (3, n, n, n) in 3-D and (2, n, n) in 2-D, like the reference.
    """
    if spectrum_dim(u_hat) != plan.dim or u_hat.shape[1] != plan.N:
        raise ValueError(f"spectrum {tuple(u_hat.shape)} does not match the plan (dim {plan.dim}, N {plan.N})")
    shape = (3, ckp_N, ckp_N, ckp_N) if plan.dim == 3 else (2, ckp_N, ckp_N)
    out = torch.empty(shape, dtype=plan.dtype, device=u_hat.device)
    _lib.check(_lib.load().qlles_filter(plan.handle, _lib.ptr(u_hat.contiguous()), int(ckp_N), _lib.ptr(out),
                                        _lib.stream_ptr()))
    return out


def spectral_filtering(u, ckp_N):
    """Reference signature: physical (3,N,N,N) or (2,N,N) field -> (dim, n, ...) field.

    The solver's 2-D (3,N,N,1) field is taken as the (2,N,N) field it pads."""
    dim, N, full = physical_field(u)
    dtype = _real_dtype(full)
    plan = helper_plan(dim, N, dtype)
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    shape_hat = (3, N, N, N // 2 + 1) if dim == 3 else (3, N, N // 2 + 1)
    u_hat = torch.empty(shape_hat, dtype=cdt, device=u.device)
    _lib.check(_lib.load().qlles_rfftn(plan.handle, _lib.ptr(full), _lib.ptr(u_hat), _lib.stream_ptr()))
    return spectral_filtering_hat(u_hat, ckp_N, plan)


def comp_dissipation_rate(u_hat, nu):
    """nu <|curl u|^2> through irfftn of i k x u_hat (utils.py:300-315), FP64 reduction."""
    dim = spectrum_dim(u_hat)
    plan = helper_plan(dim, u_hat.shape[1], _real_dtype(u_hat), nu)
    out = torch.zeros(5, dtype=torch.float64, device=u_hat.device)
    _lib.check(_lib.load().qlles_dissipation(plan.handle, _lib.ptr(u_hat.contiguous()), _lib.ptr(out),
                                             _lib.stream_ptr()))
    return float(out[0])


def kinetic_energy_hat(u_hat, kf=0):
    """FP64 Parseval energy 0.5 <|u|^2> of u_hat and of its 0 < |k| <= kf shell."""
    dim = spectrum_dim(u_hat)
    plan = helper_plan(dim, u_hat.shape[1], _real_dtype(u_hat), 0.0, kf)
    out = torch.zeros(2, dtype=torch.float64, device=u_hat.device)
    _lib.check(_lib.load().qlles_energy_hat(plan.handle, _lib.ptr(u_hat.contiguous()), _lib.ptr(out),
                                            _lib.stream_ptr()))
    return float(out[0]), float(out[1])


def energy_spectrum(vel, mul_fac=1.0, is_scalar_field=False, dim=3):
    """Shell-binned E(k) (utils.py:35-97): transform and FP64 binning on the GPU (qlles_energy_spectrum).

    vel: (3, N, N, N) for dim 3; (3, N, N, 1) or (2, N, N) for dim 2.  Returns
    a float64 NumPy array of N bins like the reference.
    """
    del is_scalar_field
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    _, n, full = physical_field(vel, dim)
    plan = helper_plan(dim, n, _real_dtype(full))
    ek = torch.empty(n, dtype=torch.float64, device=full.device)
    _lib.check(_lib.load().qlles_energy_spectrum(plan.handle, _lib.ptr(full), float(mul_fac), _lib.ptr(ek),
                                                 _lib.stream_ptr()))
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
