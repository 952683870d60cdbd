"""The differentiable configuration at its size: forced HIT at 128^3, K = 50 RK4 steps, a checkpoint
every 10 steps, 16^3 particles advected by the NUFFT, loss = MSE of the final particle positions against
a seeded perturbed target; gradients in the initial spectrum and the injection power.

No reference counterpart (the reference computes no gradient).  Pins:
  - FP32 dot-product identities of the step map and of the filter + NUFFT interpolation at 128^3
    (relative residual <= 1e-5), and the recorded-step adjoint equal to the recomputing one;
  - the FP64 twin's gradients against central differences with eps = 1e-3, within 1e-4 relative.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, K, EVERY, NP, NU, DT, P, KF, E0 = 128, 50, 10, 16, 1e-3, 5e-4, 0.1, 2, 0.5


def _inner(a, b):
    return float(torch.sum(a.conj().to(torch.complex128) * b.to(torch.complex128)).real)


def _setup(prec, power=P):
    from l3es_amd import set_up_solver

    _, uh, _, _, step, _, _, _, _ = set_up_solver(N, NU, 3, "HIT", DT, 7, NP, KF, e_kin_target=E0,
                                                  forcing_type="power", power=power, spectrum="k4exp",
                                                  precision=prec, warmup=False)
    return uh, step


def _interp(prec):
    from l3es_amd.integrator import spectral_interpolator_wrapper

    return spectral_interpolator_wrapper(NP, (1, 2, 3), backend="nufft", precision=prec, tol=1e-6)


def _valid_dir(uh, step, seed):
    """The rfftn of a seeded real field, scaled to the norm of u_hat."""
    from l3es_amd import _lib

    g = torch.Generator(device="cpu").manual_seed(seed)
    real = torch.randn((3, N, N, N), generator=g, dtype=torch.float64)
    real = real.to(torch.float32 if uh.dtype == torch.complex64 else torch.float64).cuda()
    v = torch.empty_like(uh)
    _lib.check(_lib.load().qlles_rfftn(step.plan.handle, _lib.ptr(real), _lib.ptr(v), _lib.stream_ptr()))
    return v * (torch.linalg.vector_norm(uh) / torch.linalg.vector_norm(v))


def test_adjoint128_fp32_dot_products_and_records(gpu):
    from l3es_amd import adjoint as A

    uh, step = _setup("float32")
    print(f"[adjoint128] FP32 {N}^3 step maps", flush=True)
    v, w = _valid_dir(uh, step, 1), _valid_dir(uh, step, 2)
    jv = A.step_jvp(step, uh, v)
    dp = torch.zeros(1, dtype=torch.float64, device=uh.device)
    vw = A.step_vjp(step, uh, w, dp)
    res = abs(_inner(jv, w) - _inner(v, vw)) / (float(torch.linalg.vector_norm(jv)) *
                                                 float(torch.linalg.vector_norm(w)))
    assert res <= 1e-5, res
    # the recorded step and its adjoint (the integrators' replay path) at this size
    rec = torch.empty(A.record_bytes(step), dtype=torch.uint8, device=uh.device)
    nxt = A.step_record(step, uh, rec)
    ref = uh.clone()
    step.advance(None, ref, 1, None, use_u=False)
    assert torch.equal(nxt, ref)  # the recorded step is the fused step
    dpr = torch.zeros(1, dtype=torch.float64, device=uh.device)
    vr = A.step_vjp_record(step, uh, rec, w, dpr)
    assert float(torch.linalg.vector_norm(vr - vw)) < 2e-5 * float(torch.linalg.vector_norm(vw))
    assert abs(float(dpr) - float(dp)) <= 2e-5 * abs(float(dp))
    # the filter + NUFFT interpolation of 16^3 particles as a linear map of the 128^3 spectrum
    it = _interp("float32")
    g = torch.Generator(device="cpu").manual_seed(3)
    r = (torch.rand((NP ** 3, 3), generator=g, dtype=torch.float64) * 2 * np.pi).cuda()
    ub = torch.randn((NP ** 3, 3), generator=g, dtype=torch.float64).to(torch.float32).cuda()
    iv = it.from_hat(v, r.clone())
    lam, _ = A.interp_vjp(it, uh, r, ub)
    lhs = float(torch.sum(iv.to(torch.float64) * ub.to(torch.float64)))
    res_i = abs(lhs - _inner(v, lam)) / (float(torch.linalg.vector_norm(iv)) * float(torch.linalg.vector_norm(ub)))
    assert res_i <= 1e-5, res_i
    print(f"[adjoint128] step residual {res:.2e}, interpolation residual {res_i:.2e}", flush=True)


def _particle_problem(prec):
    from l3es_amd import adjoint as A
    from l3es_amd.integrator import pos_init_jittered

    uh0, step = _setup(prec)
    it = _interp(prec)
    r0 = torch.from_numpy(pos_init_jittered(NP, 3, seed=2)).cuda()
    g = torch.Generator(device="cpu").manual_seed(5)
    with torch.no_grad():
        _, rK = A.integrate_particles(step, it, uh0, r0, K, EVERY)
        target = rK + 0.01 * torch.randn(rK.shape, generator=g, dtype=torch.float64).cuda()

    def loss(st, x, p=None):
        _, r = A.integrate_particles(st, it, x, r0, K, EVERY, power=p)
        return torch.mean(A.periodic_diff(r, target) ** 2)

    return uh0, step, loss


def test_adjoint128_fp32_gradients_finite(gpu):
    """The configuration itself in FP32: gradients in u_hat0 and P exist and are finite."""
    uh0, step, loss = _particle_problem("float32")
    print(f"[adjoint128] FP32 K={K} gradient", flush=True)
    x = uh0.clone().requires_grad_(True)
    p = torch.tensor(P, dtype=torch.float64, device=uh0.device, requires_grad=True)
    L = loss(step, x, p)
    L.backward()
    assert torch.isfinite(x.grad).all() and np.isfinite(float(p.grad))
    assert float(torch.linalg.vector_norm(x.grad)) > 0 and float(p.grad) != 0.0


def test_adjoint128_fp64_gradients_match_central_difference(gpu):
    """FP64 twin of the configuration: dL/du_hat0 along a seeded direction and dL/dP against central
    differences (eps = 1e-3), within 1e-4 relative."""
    _, step_p = _setup("float64", P + 1e-3)
    _, step_m = _setup("float64", P - 1e-3)
    uh0, step, loss = _particle_problem("float64")
    print(f"[adjoint128] FP64 K={K} gradient vs central differences", flush=True)
    x = uh0.clone().requires_grad_(True)
    p = torch.tensor(P, dtype=torch.float64, device=uh0.device, requires_grad=True)
    L = loss(step, x, p)
    L.backward()
    v = _valid_dir(uh0, step, 11)
    eps = 1e-3
    with torch.no_grad():
        fd_u = (float(loss(step, uh0 + eps * v)) - float(loss(step, uh0 - eps * v))) / (2 * eps)
        fd_p = (float(loss(step_p, uh0)) - float(loss(step_m, uh0))) / (2 * 1e-3)
    an_u = _inner(x.grad, v)
    an_p = float(p.grad)
    print(f"[adjoint128] d/du: fd {fd_u:.6e} adjoint {an_u:.6e}; d/dP: fd {fd_p:.6e} adjoint {an_p:.6e}",
          flush=True)
    assert abs(fd_u - an_u) <= 1e-4 * abs(an_u), (fd_u, an_u)
    assert abs(fd_p - an_p) <= 1e-4 * abs(an_p), (fd_p, an_p)
