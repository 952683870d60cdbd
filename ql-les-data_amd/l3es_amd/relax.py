"""SPH particle relaxation on the HIP path (mirror of l3es/relax.py).

``relax_wrapper`` returns the reference's pair ``(loop_body, sph_body)``:
``loop_body(r)`` is the relaxation acceleration (pressure + half the
background pressure, relax.py:76-120), ``sph_body(r, u)`` the separate
pressure, viscous and transport-velocity terms (relax.py:123-176).  Both run
as one cell-list pass on the GPU (``qlles_sph_*``) and return FP64 device
tensors.  ``relax_advect`` is the relaxation branch of the combined loop and
its particle shift (combined.py:132-142, :199) as one fused call.
"""

import numpy as np
import torch

from . import _lib


def _dev_f64(a, dim):
    dev = torch.device("cuda", torch.cuda.current_device())
    t = torch.as_tensor(a)
    t = t.to(device=dev, dtype=torch.float64).contiguous()
    if t.dim() != 2 or t.shape[1] != dim:
        raise ValueError(f"expected an (np, {dim}) array, got {tuple(t.shape)}")
    return t


def relax_wrapper(N, dim=3, L=2 * np.pi, is_physical=False, u_ref=None):
    """Relaxation closures for N^dim particles in a periodic box of side L (relax.py:16-74).

    Same static parameters as the reference: smoothing length dx = L/N, quintic
    kernel, Tait EoS with p_ref = (10 u_ref / l_ref)^2, l_ref = 1 (is_physical)
    or dx.
    """
    if u_ref is None:
        raise TypeError("relax_wrapper needs u_ref (reference velocity of the equation of state)")
    ctx = _lib.Sph(dim, N, L, is_physical, u_ref)
    lib = _lib.load()

    def loop_body(r):
        r = _dev_f64(r, dim)
        acc = torch.empty_like(r)
        _lib.check(lib.qlles_sph_relax_acc(ctx.handle, _lib.ptr(r), r.shape[0], _lib.ptr(acc), None,
                                           _lib.stream_ptr()))
        return acc

    def sph_body(r, u):
        assert is_physical, "sph is only implemented for physical units."  # relax.py:126
        r = _dev_f64(r, dim)
        u = _dev_f64(u, dim)
        out = {k: torch.empty_like(r) for k in ("acc_p", "acc_visc", "acc_tvf")}
        _lib.check(lib.qlles_sph_forces(ctx.handle, _lib.ptr(r), _lib.ptr(u), r.shape[0], _lib.ptr(out["acc_p"]),
                                        _lib.ptr(out["acc_visc"]), _lib.ptr(out["acc_tvf"]), _lib.stream_ptr()))
        return out

    loop_body.ctx = sph_body.ctx = ctx
    return loop_body, sph_body


def relax_advect(relax_fn, r, u_r, dt, relax_dt_factor, acc=None):
    """combined.py:132-142 then :199, in place on device tensors:

        r_t = shift(r, dt u_r); a = relax_fn(r_t); r_t = shift(r_t, (f dt)^2 a)
        u_r = displacement(r_t, r) / dt;  r = shift(r, dt u_r)

    ``relax_fn`` is the ``loop_body`` of ``relax_wrapper(..., is_physical=True)``;
    ``r`` is (np, dim) FP64, ``u_r`` (np, dim) FP32 or FP64; ``acc`` (optional,
    FP64) receives a.  Returns (r, u_r).
    """
    ctx = relax_fn.ctx
    if r.dtype != torch.float64 or u_r.dtype not in (torch.float32, torch.float64):
        raise ValueError("r must be float64 and u_r float32 or float64")
    if r.shape != u_r.shape or (acc is not None and acc.shape != r.shape):
        raise ValueError("r, u_r (and acc) must have the same (np, dim) shape")
    dt_code = _lib.QLLES_F32 if u_r.dtype == torch.float32 else _lib.QLLES_F64
    _lib.check(_lib.load().qlles_sph_relax_advect(ctx.handle, _lib.ptr(r), r.shape[0], _lib.ptr(u_r), dt_code,
                                                  float(dt), float(relax_dt_factor), _lib.ptr(acc),
                                                  _lib.stream_ptr()))
    return r, u_r
