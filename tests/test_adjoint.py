"""Tangent-linear and adjoint maps (qlles_*_jvp / _vjp) and the checkpointed differentiable integrator.

No reference counterpart (the reference computes no gradient).  Pins, without the reference:
  - dot-product identity <J v, w> = <v, J^T w> (1e-10 relative in FP64, 1e-5 in FP32);
  - the tangent map equals the central difference of the forward step (which runs the fused kernels);
  - the gradient of an MSE loss after K checkpointed steps matches a central finite difference
    (eps = 1e-3, within 1e-4 relative).
"""

import numpy as np
import pytest
import torch


def _inner(a, b):
    return float(torch.sum(a.conj() * b).real)


def _setup(N, prec, seed=3, power=None):
    from l3es_amd import set_up_solver

    if power is None:
        kw = dict(forcing_type="none")
        kf = None
    else:
        kw = dict(forcing_type="power", power=power)
        kf = 2
    _, uh, _, _, step, _, _, _, _ = set_up_solver(N, 0.02, 3, "HIT", 5e-3, seed, 8, kf, precision=prec,
                                                  warmup=False, **kw)
    return uh, step


def _rand_hat(uh, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    re = torch.randn(uh.shape, generator=g, dtype=torch.float64)
    im = torch.randn(uh.shape, generator=g, dtype=torch.float64)
    return torch.complex(re, im).to(uh.dtype).to(uh.device)


def _valid_dir(uh, step, seed):
    """A direction that is the rfftn of a real field, scaled to the norm of u_hat."""
    from l3es_amd import _lib

    N = uh.shape[1]
    g = torch.Generator(device="cpu").manual_seed(seed)
    real = torch.randn((3, N, N, N), generator=g, dtype=torch.float64)
    real = real.to(torch.float32 if uh.dtype == torch.complex64 else torch.float64).cuda()
    v = torch.empty_like(uh)
    _lib.check(_lib.load().qlles_rfftn(step.plan.handle, _lib.ptr(real), _lib.ptr(v), _lib.stream_ptr()))
    return v * (torch.linalg.vector_norm(uh) / torch.linalg.vector_norm(v))


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("float64", 1e-10), ("float32", 1e-5)])
@pytest.mark.parametrize("which", ["rhs", "step"])
@pytest.mark.parametrize("power", [None, 0.3])
def test_dot_product(gpu, prec, tol, which, power):
    from l3es_amd import adjoint as A

    uh, step = _setup(16, prec, power=power)
    v, w = _rand_hat(uh, 1), _rand_hat(uh, 2)
    jvp, vjp = (A.rhs_jvp, A.rhs_vjp) if which == "rhs" else (A.step_jvp, A.step_vjp)
    lhs = _inner(jvp(step, uh, v), w)
    rhs = _inner(v, vjp(step, uh, w))
    scale = float(torch.linalg.vector_norm(jvp(step, uh, v))) * float(torch.linalg.vector_norm(w))
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)


@pytest.mark.gpu
@pytest.mark.parametrize("power", [None, 0.3])
def test_step_jvp_matches_central_difference(gpu, power):
    from l3es_amd import adjoint as A

    uh, step = _setup(16, "float64", power=power)
    v = _valid_dir(uh, step, 5)
    eps = 1e-4

    def fwd(x):
        y = x.clone()
        step.advance(None, y, 1, None, use_u=False)
        return y

    fd = (fwd(uh + eps * v) - fwd(uh - eps * v)) / (2 * eps)
    jv = A.step_jvp(step, uh, v)
    err = float(torch.linalg.vector_norm(fd - jv) / torch.linalg.vector_norm(jv))
    assert err < 1e-6, err


@pytest.mark.gpu
def test_rhs_vjp_rejects_ekin_tot_plans(gpu):
    from l3es_amd import _lib, set_up_solver
    from l3es_amd import adjoint as A

    _, uh, _, _, step, _, _, _, _ = set_up_solver(16, 0.02, 3, "HIT", 5e-3, 3, 8, 2, forcing_type="ekin_tot",
                                                  e_kin_target=0.5, precision="float64", warmup=False)
    with pytest.raises(_lib.QllesError):
        A.rhs_vjp(step, uh, uh.clone())


@pytest.mark.gpu
def test_checkpointed_gradient_matches_central_difference(gpu):
    from l3es_amd import adjoint as A

    uh0, step = _setup(16, "float64", seed=7)
    K, every = 7, 3
    g = torch.Generator(device="cpu").manual_seed(11)
    with torch.no_grad():
        target = A.irfftn(A.integrate(step, uh0, K, every), step.plan)
        target = target + 0.05 * torch.randn(target.shape, generator=g, dtype=torch.float64).cuda()

    def loss(x):
        return torch.mean((A.irfftn(A.integrate(step, x, K, every), step.plan) - target) ** 2)

    x = uh0.clone().requires_grad_(True)
    L = loss(x)
    L.backward()
    grad = x.grad.detach()
    v = _valid_dir(uh0, step, 13)
    eps = 1e-3
    with torch.no_grad():
        fd = (float(loss(uh0 + eps * v)) - float(loss(uh0 - eps * v))) / (2 * eps)
    an = _inner(grad, v)
    assert abs(fd - an) <= 1e-4 * abs(an), (fd, an)


@pytest.mark.gpu
def test_power_gradient_matches_central_difference(gpu):
    """dL/dP of a fixed-power plan (the forcing-amplitude gradient) against plans at P +- eps."""
    from l3es_amd import adjoint as A

    P, eps, K, every = 0.3, 1e-3, 6, 4
    uh0, step = _setup(16, "float64", seed=9, power=P)
    _, step_p = _setup(16, "float64", seed=9, power=P + eps)
    _, step_m = _setup(16, "float64", seed=9, power=P - eps)
    g = torch.Generator(device="cpu").manual_seed(17)
    with torch.no_grad():
        target = A.irfftn(A.integrate(step, uh0, K, every), step.plan)
        target = target + 0.05 * torch.randn(target.shape, generator=g, dtype=torch.float64).cuda()

    def loss(st, p=None):
        return torch.mean((A.irfftn(A.integrate(st, uh0, K, every, power=p), st.plan) - target) ** 2)

    p = torch.tensor(P, dtype=torch.float64, device=uh0.device, requires_grad=True)
    L = loss(step, p)
    L.backward()
    with torch.no_grad():
        fd = (float(loss(step_p)) - float(loss(step_m))) / (2 * eps)
    an = float(p.grad)
    assert abs(fd - an) <= 1e-4 * abs(an) + 1e-12, (fd, an)


@pytest.mark.gpu
def test_irfftn_adjoint_pairs(gpu):
    """<irfftn(a), g> = <a, irfftn^T g> and <rfftn(x), b> = <x, rfftn^T b>."""
    from l3es_amd import _lib

    uh, step = _setup(16, "float64")
    lib = _lib.load()
    a, b = _rand_hat(uh, 21), _rand_hat(uh, 22)
    gen = torch.Generator(device="cpu").manual_seed(23)
    x = torch.randn((3, 16, 16, 16), generator=gen, dtype=torch.float64).cuda()
    g = torch.randn((3, 16, 16, 16), generator=gen, dtype=torch.float64).cuda()
    ia, gt, rx, bt = torch.empty_like(x), torch.empty_like(uh), torch.empty_like(uh), torch.empty_like(x)
    s = _lib.stream_ptr()
    _lib.check(lib.qlles_irfftn(step.plan.handle, _lib.ptr(a), _lib.ptr(ia), s))
    _lib.check(lib.qlles_irfftn_adjoint(step.plan.handle, _lib.ptr(g), _lib.ptr(gt), s))
    _lib.check(lib.qlles_rfftn(step.plan.handle, _lib.ptr(x), _lib.ptr(rx), s))
    _lib.check(lib.qlles_rfftn_adjoint(step.plan.handle, _lib.ptr(b), _lib.ptr(bt), s))
    l1, r1 = float(torch.sum(ia * g)), _inner(a, gt)
    l2, r2 = _inner(rx, b), float(torch.sum(x * bt))
    assert abs(l1 - r1) <= 1e-12 * max(1.0, abs(l1))
    assert abs(l2 - r2) <= 1e-12 * max(1.0, abs(l2)) * 16 ** 3
    assert np.isfinite(l1) and np.isfinite(l2)


def _interp(n, prec, tol=1e-9):
    from l3es_amd.integrator import spectral_interpolator_wrapper

    return spectral_interpolator_wrapper(n, (1, 2, 3), backend="nufft", precision=prec, tol=tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("float64", 1e-10), ("float32", 1e-5)])
@pytest.mark.parametrize("N,n", [(16, 8), (16, 16)])
def test_interp_vjp_dot_product(gpu, prec, tol, N, n):
    """<I v, ub> = <v, I^T ub> for the filter + NUFFT interpolation, as a linear map of the spectrum."""
    from l3es_amd import adjoint as A

    uh, step = _setup(N, prec)
    it = _interp(n, prec, 1e-9 if prec == "float64" else 1e-6)
    g = torch.Generator(device="cpu").manual_seed(31)
    r = (torch.rand((200, 3), generator=g, dtype=torch.float64) * 2 * np.pi).cuda()
    v = _rand_hat(uh, 32)
    ub = torch.randn((200, 3), generator=g, dtype=torch.float64).to(it.dtype).cuda()
    iv = it.from_hat(v, r.clone())
    lam, _ = A.interp_vjp(it, uh, r, ub)
    lhs = float(torch.sum(iv.to(torch.float64) * ub.to(torch.float64)))
    rhs = _inner(v, lam)
    scale = float(torch.linalg.vector_norm(iv)) * float(torch.linalg.vector_norm(ub))
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)


@pytest.mark.gpu
def test_interp_position_gradient_matches_central_difference(gpu):
    from l3es_amd import adjoint as A

    uh, step = _setup(16, "float64")
    it = _interp(8, "float64")
    g = torch.Generator(device="cpu").manual_seed(41)
    r = (torch.rand((64, 3), generator=g, dtype=torch.float64) * 2 * np.pi).cuda()
    ub = torch.randn((64, 3), generator=g, dtype=torch.float64).cuda()
    _, rbar = A.interp_vjp(it, uh, r, ub)
    eps = 1e-5
    for j in range(3):
        e = torch.zeros_like(r)
        e[:, j] = eps
        fp = it.from_hat(uh, (r + e).clone())
        fm = it.from_hat(uh, (r - e).clone())
        fd = torch.sum((fp - fm) * ub, dim=1) / (2 * eps)
        err = float(torch.max(torch.abs(fd - rbar[:, j])) / torch.max(torch.abs(rbar[:, j])))
        assert err < 1e-5, (j, err)


@pytest.mark.gpu
def test_particle_loss_gradients_match_central_difference(gpu):
    """Loss = MSE of final particle positions vs a perturbed target after K combined-mode steps:
    gradients in u_hat0 and in the injection power against central differences (eps = 1e-3)."""
    from l3es_amd import adjoint as A
    from l3es_amd.integrator import pos_init_cartesian

    P, K, every = 0.3, 5, 2
    uh0, step = _setup(16, "float64", seed=5, power=P)
    _, step_p = _setup(16, "float64", seed=5, power=P + 1e-3)
    _, step_m = _setup(16, "float64", seed=5, power=P - 1e-3)
    it = _interp(8, "float64")
    r0 = torch.from_numpy(pos_init_cartesian(8, 3)).cuda()
    g = torch.Generator(device="cpu").manual_seed(43)
    with torch.no_grad():
        _, rK = A.integrate_particles(step, it, uh0, r0, K, every)
        target = rK + 0.01 * torch.randn(rK.shape, generator=g, dtype=torch.float64).cuda()

    def loss(st, x, p=None):
        _, r = A.integrate_particles(st, it, x, r0, K, every, power=p)
        return torch.mean(A.periodic_diff(r, target) ** 2)

    x = uh0.clone().requires_grad_(True)
    p = torch.tensor(P, dtype=torch.float64, device=uh0.device, requires_grad=True)
    L = loss(step, x, p)
    L.backward()
    v = _valid_dir(uh0, step, 47)
    eps = 1e-3
    with torch.no_grad():
        fd_u = (float(loss(step, uh0 + eps * v)) - float(loss(step, uh0 - eps * v))) / (2 * eps)
        fd_p = (float(loss(step_p, uh0)) - float(loss(step_m, uh0))) / (2 * 1e-3)
    an_u = _inner(x.grad, v)
    an_p = float(p.grad)
    assert abs(fd_u - an_u) <= 1e-4 * abs(an_u), (fd_u, an_u)
    assert abs(fd_p - an_p) <= 1e-4 * abs(an_p) + 1e-14, (fd_p, an_p)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("float64", 1e-12), ("float32", 2e-5)])
@pytest.mark.parametrize("power", [None, 0.3])
def test_recorded_step_and_its_adjoint(gpu, prec, tol, power):
    """qlles_step_record is the fused step bit for bit, and qlles_step_vjp_record equals
    qlles_step_vjp (and the power cotangent) at the same u_hat."""
    from l3es_amd import adjoint as A

    uh, step = _setup(16, prec, power=power)
    rec = torch.empty(A.record_bytes(step), dtype=torch.uint8, device=uh.device)
    nxt = A.step_record(step, uh, rec)
    ref = uh.clone()
    step.advance(None, ref, 1, None, use_u=False)
    assert torch.equal(nxt, ref)
    lam = _rand_hat(uh, 11)
    dp_a = torch.zeros(1, dtype=torch.float64, device=uh.device)
    dp_b = torch.zeros(1, dtype=torch.float64, device=uh.device)
    a = A.step_vjp_record(step, uh, rec, lam, dp_a)
    b = A.step_vjp(step, uh, lam, dp_b)
    assert float(torch.linalg.vector_norm(a - b)) < tol * float(torch.linalg.vector_norm(b))
    if power is not None:
        assert abs(float(dp_a) - float(dp_b)) <= tol * abs(float(dp_b))
    # in place (u_hat_out == u_hat) gives the same step
    uh2 = uh.clone()
    A.step_record(step, uh2, rec, out=uh2)
    assert torch.equal(uh2, nxt)


@pytest.mark.gpu
def test_gradient_with_and_without_step_records(gpu, monkeypatch):
    """The gradient is the same whether every step is recorded in the forward pass, or the forward
    checkpoints and the backward replays each segment with recorded steps, or with fused steps and
    recomputed stages."""
    from l3es_amd import adjoint as A
    from l3es_amd.integrator import pos_init_cartesian

    uh0, step = _setup(16, "float64", seed=7, power=0.3)
    it = _interp(8, "float64")
    r0 = torch.from_numpy(pos_init_cartesian(8, 3)).cuda()

    def grads():
        x = uh0.clone().requires_grad_(True)
        p = torch.tensor(0.3, dtype=torch.float64, device=uh0.device, requires_grad=True)
        uh, r = A.integrate_particles(step, it, x, r0, 4, 3, power=p)
        loss = torch.mean(torch.sin(r) ** 2) + 1e-3 * torch.sum(uh.abs() ** 2)
        loss.backward()
        return x.grad.clone(), float(p.grad)

    g1, p1 = grads()  # every step recorded in the forward pass
    real = A._records
    # only a segment's records fit: the forward checkpoints, the backward replays each segment recorded
    monkeypatch.setattr(A, "_records", lambda step, n, like: None if n > 3 else real(step, n, like))
    g3, p3 = grads()
    monkeypatch.setattr(A, "_records", lambda step, every, like: None)
    g2, p2 = grads()  # no records: fused replay, recomputed stages
    for g, p in ((g1, p1), (g3, p3)):
        assert float(torch.linalg.vector_norm(g - g2)) < 1e-10 * float(torch.linalg.vector_norm(g2))
        assert abs(p - p2) <= 1e-10 * abs(p2)
