"""HIP path vs the FP64 oracle on the same seeded inputs (GPU).

Tolerances: FP64 runs must match the oracle to round-off (relative 1e-12 on
fields after a few steps); FP32 runs to 2e-5 relative on fields (FP32 FFT
round-off ~ eps*log2(N)) over the short horizons used here.
"""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-5


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t.to(dtype) if dtype is not None else t


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def tol(prec):
    return F64_TOL if prec == "float64" else F32_TOL


def tdt(prec):
    return torch.float64 if prec == "float64" else torch.float32


def ctdt(prec):
    return torch.complex128 if prec == "float64" else torch.complex64


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("dim,N", [(3, 8), (3, 16), (3, 32), (3, 64), (3, 96), (3, 128),
                                   (2, 8), (2, 32), (2, 96), (2, 256), (2, 512)])
def test_rfftn_irfftn(gpu, dim, N, prec):
    from l3es_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(N + dim)
    shape = (3, N, N, N) if dim == 3 else (3, N, N, 1)
    u = rng.standard_normal(shape)
    if dim == 2:
        u[2] = 0
    axes = (1, 2, 3) if dim == 3 else (1, 2)
    ref = np.squeeze(np.fft.rfftn(u, axes=axes), axis=-1) if dim == 2 else np.fft.rfftn(u, axes=axes)
    plan = _lib.Plan(dim, N, tdt(prec), 0.0, 1.0)
    ud = dev(u, tdt(prec))
    uh = torch.empty(ref.shape, dtype=ctdt(prec), device="cuda")
    _lib.check(lib.qlles_rfftn(plan.handle, _lib.ptr(ud), _lib.ptr(uh), _lib.stream_ptr()))
    assert rel(host(uh), ref) < 50 * tol(prec)
    # inverse of a non-Hermitian spectrum follows numpy's irfftn (Im of DC/Nyquist ignored)
    spec = ref + 0.3 * (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)) * np.abs(ref).mean()
    if dim == 2:
        spec[2] = 0
    ref_u = np.fft.irfftn(spec, axes=axes) if dim == 3 else np.fft.irfftn(spec, axes=axes)[..., None]
    out = torch.empty(shape, dtype=tdt(prec), device="cuda")
    _lib.check(lib.qlles_irfftn(plan.handle, _lib.ptr(dev(spec, ctdt(prec))), _lib.ptr(out), _lib.stream_ptr()))
    assert rel(host(out), ref_u) < 50 * tol(prec)


# ---------------------------------------------------------------------------
# RHS and RK4 step (spectral_solver.py:27-120)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("N", [16, 32, 96])
@pytest.mark.parametrize("rk", [0, 2])
def test_rhs_3d(gpu, N, rk, prec):
    from l3es_amd import rhs_wrapper

    nu = 0.01
    u = O.init_u_hit(N, seed=11)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    ref = O.rhs_wrapper(N, nu)(u, uh, rk)
    du = rhs_wrapper(N, nu, precision=prec)(dev(u, tdt(prec)), dev(uh, ctdt(prec)), rk)
    assert rel(host(du), ref) < 5 * tol(prec)


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("N", [16, 64])
def test_rhs_2d(gpu, N, prec):
    from l3es_amd import rhs_wrapper

    nu = 1e-3
    u = O.init_u_kolm_np(N, seed=2)
    uh = np.squeeze(np.fft.rfftn(u, axes=(1, 2)), -1)
    for rk in (0, 1):
        ref = O.rhs_wrapper(N, nu, (1, 2))(u, uh, rk)
        du = rhs_wrapper(N, nu, (1, 2), precision=prec)(dev(u, tdt(prec)), dev(uh, ctdt(prec)), rk)
        assert rel(host(du), ref) < 5 * tol(prec)


def _run_gpu(N, nu, dim, case, dt, nsteps, prec, **kw):
    from l3es_amd import set_up_solver

    u, uh, _, _, step, _, _, e0, _ = set_up_solver(N, nu, dim, case, dt, 42, N, kw.pop("kf", None),
                                                   precision=prec, warmup=False, **kw)
    e_inj = []
    for _ in range(nsteps):
        u, uh, e = step(u, uh)
        e_inj.append(float(e))
    return host(u), host(uh), e0, e_inj


def _run_oracle(N, nu, dim, case, dt, nsteps, **kw):
    u, uh, step, e0, _ = O.set_up_solver(N, nu, dim, case, dt, seed=42, warmup=False, **kw)
    e_inj = []
    for _ in range(nsteps):
        u, uh, e = step(u, uh)
        e_inj.append(float(e))
    return u, uh, e0, e_inj


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("N", [16, 32])
def test_step_hit_unforced(gpu, N, prec):
    kw = dict(forcing_type="none")
    ug, uhg, _, _ = _run_gpu(N, 0.01, 3, "HIT", 0.005, 4, prec, **kw)
    uo, uho, _, _ = _run_oracle(N, 0.01, 3, "HIT", 0.005, 4, **kw)
    assert rel(ug, uo) < 5 * tol(prec)
    assert rel(uhg, uho) < 5 * tol(prec)


def kat(name):
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")) as fh:
        return json.load(fh)[name]


def test_tgv_known_answer_gpu(gpu):
    """KAT (from_spectralDNS/main_numpy.py:136) on the HIP path in FP64."""
    k = kat("tgv3d_numpy")
    ug, _, _, _ = _run_gpu(k["N"], k["nu"], 3, k["case"], k["dt"], k["steps"], "float64", forcing_type="none")
    e = 0.5 * np.mean(np.sum(ug * ug, axis=0))
    assert round(e - k["e_kin"], k["round_digits"]) == 0
    uo, _, _, _ = _run_oracle(k["N"], k["nu"], 3, k["case"], k["dt"], k["steps"])
    assert rel(ug, uo) < 1e-12


def test_tgv_fp32_energy(gpu):
    k = kat("tgv3d_numpy")
    ug, _, _, _ = _run_gpu(k["N"], k["nu"], 3, k["case"], k["dt"], k["steps"], "float32", forcing_type="none")
    e = 0.5 * np.mean(np.sum(ug.astype(np.float64) ** 2, axis=0))
    assert abs(e - k["e_kin"]) / k["e_kin"] < 1e-6


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_simulate_validation_config(gpu, prec):
    """configs/tgv3d_validate.yaml through the simulation driver: the reference asserts the
    kinetic energy after its 12 steps to two decimals (spectral_solver.py:363-367); the FP64
    run also matches the oracle's 12 steps."""
    from l3es_amd import simulate

    k = kat("tgv3d_validate")
    out = simulate(k["case"], N=k["N"], nu=k["nu"], t_final=k["t_final"], dt=k["dt"], log_freq=5,
                   precision=prec, write_checkpoints=False, verbose=False)
    assert round(out["e_kin"] - k["e_kin"], k["round_digits"]) == 0
    if prec == "float64":
        u, uh, step, _, _ = O.set_up_solver(k["N"], k["nu"], 3, k["case"], k["dt"], warmup=False)
        for _ in range(k["steps"]):
            u, uh, _ = step(u, uh)
        assert rel(host(out["u"]), u) < 1e-12


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_step_hit_ekin_tot(gpu, prec):
    kw = dict(forcing_type="ekin_tot", kf=3, e_kin_target=1.25)
    ug, uhg, e0g, eig = _run_gpu(32, 0.002, 3, "HIT", 0.002, 3, prec, **kw)
    uo, uho, e0o, eio = _run_oracle(32, 0.002, 3, "HIT", 0.002, 3, **kw)
    assert abs(e0g - e0o) < 1e-12 if prec == "float64" else abs(e0g - e0o) < 1e-5
    assert rel(ug, uo) < 5 * tol(prec)
    np.testing.assert_allclose(eig, eio, rtol=1e-9 if prec == "float64" else 2e-3, atol=1e-9)
    e = 0.5 * np.mean(np.sum(ug.astype(np.float64) ** 2, axis=0))
    assert abs(e - e0o) < (1e-12 if prec == "float64" else 1e-5)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_step_hit_power_forcing(gpu, prec):
    kw = dict(forcing_type="power", kf=2, power=0.1, spectrum="k4exp", e_kin_target=0.5)
    ug, uhg, _, _ = _run_gpu(32, 1e-3, 3, "HIT", 5e-4, 3, prec, **kw)
    uo, uho, _, _ = _run_oracle(32, 1e-3, 3, "HIT", 5e-4, 3, **kw)
    assert rel(ug, uo) < 5 * tol(prec)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_step_if_viscous(gpu, prec):
    kw = dict(forcing_type="none", viscous_mode="IF")
    ug, _, _, _ = _run_gpu(16, 0.02, 3, "HIT", 0.005, 3, prec, **kw)
    uo, _, _, _ = _run_oracle(16, 0.02, 3, "HIT", 0.005, 3, **kw)
    assert rel(ug, uo) < 5 * tol(prec)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_step_2d_tgv(gpu, prec):
    ug, _, _, _ = _run_gpu(64, 0.06283, 2, "TGV2D", 0.001, 5, prec, forcing_type="none")
    uo, _, _, _ = _run_oracle(64, 0.06283, 2, "TGV2D", 0.001, 5)
    assert ug.shape == uo.shape == (3, 64, 64, 1)
    assert rel(ug, uo) < 5 * tol(prec)


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("N", [64, 256])
def test_step_2d_kolmogorov(gpu, N, prec):
    kw = dict(forcing_type="kolmogorov", drag=0.1, kolm_amp=1.0, kolm_k=4)
    ug, _, _, _ = _run_gpu(N, 1e-3, 2, "Kolm", 1e-3, 5, prec, **kw)
    uo, _, _, _ = _run_oracle(N, 1e-3, 2, "Kolm", 1e-3, 5, **kw)
    assert rel(ug, uo) < 5 * tol(prec)


def test_step_multi_equals_repeated(gpu):
    """qlles_step(nsteps=K) == K calls of nsteps=1 (in-place chaining), bit for bit."""
    from l3es_amd import set_up_solver

    u, uh, _, _, step, _, _, _, _ = set_up_solver(32, 0.01, 3, "HIT", 0.005, 42, 32, None, forcing_type="none",
                                                  precision="float32", warmup=False)
    u1, uh1 = u.clone(), uh.clone()
    for _ in range(3):
        step.advance(u1, uh1, 1, u1)
    u2, uh2 = u.clone(), uh.clone()
    step.advance(u2, uh2, 3, u2, use_u=False)
    u3, uh3 = u.clone(), uh.clone()
    for _ in range(3):
        step.advance(None, uh3, 1, u3, use_u=False)
    assert np.array_equal(host(uh2), host(uh3))
    assert rel(host(uh1), host(uh2)) < 1e-5


# ---------------------------------------------------------------------------
# init, filter, diagnostics
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [16, 32])
def test_rogallo_init_gpu(gpu, N):
    from l3es_amd.init_fields import init_u_hit

    for spectrum in ("k53", "k4exp"):
        assert rel(host(init_u_hit(N, 7, spectrum=spectrum)), O.init_u_hit(N, 7, spectrum=spectrum)) < 1e-13


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("dim,N,n", [(3, 32, 16), (3, 64, 32), (3, 32, 32), (2, 64, 32), (2, 256, 64)])
def test_filter(gpu, dim, N, n, prec):
    from l3es_amd import _lib
    from l3es_amd.utils import spectral_filtering_hat

    if dim == 3:
        u = O.init_u_hit(N, seed=4)
        uh = np.fft.rfftn(u, axes=(1, 2, 3))
        ref = O.spectral_filtering(u, n) if n < N else u
    else:
        u = O.init_u_kolm_np(N, seed=4)
        uh = np.squeeze(np.fft.rfftn(u, axes=(1, 2)), -1)
        ref = O.spectral_filtering(u[:2, :, :, 0], n) if n < N else u[:2, :, :, 0]
    plan = _lib.Plan(dim, N, tdt(prec), 0.0, 1.0)
    out = spectral_filtering_hat(dev(uh, ctdt(prec)), n, plan)
    assert rel(host(out), ref) < 20 * tol(prec)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_diagnostics_and_dissipation(gpu, prec):
    from l3es_amd import _lib
    from l3es_amd.spectral_solver import comp_dt, diagnostics
    from l3es_amd.utils import comp_dissipation_rate, kinetic_energy_hat

    N, nu = 32, 0.01
    u = O.init_u_hit(N, seed=8)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    plan = _lib.Plan(3, N, tdt(prec), nu, 1.0)
    d = host(diagnostics(dev(u, tdt(prec)), plan))
    t = 1e-13 if prec == "float64" else 1e-6
    assert abs(d[0] - O.kinetic_energy(u)) < t * O.kinetic_energy(u)
    assert abs(d[1] - np.abs(u).max()) < t * np.abs(u).max()
    dx = 2 * np.pi / N
    assert abs(comp_dt(dev(u, tdt(prec)), dx, nu) - O.comp_dt(u, dx, nu)) < 1e-5 * O.comp_dt(u, dx, nu)
    eps = comp_dissipation_rate(dev(uh, ctdt(prec)), nu)
    assert abs(eps - O.comp_dissipation_rate(uh, nu)) < 100 * tol(prec) * O.comp_dissipation_rate(uh, nu)
    e, e_low = kinetic_energy_hat(dev(uh, ctdt(prec)), kf=3)
    assert abs(e - O.kinetic_energy_hat(uh, N, 3)) < 10 * tol(prec) * e
    assert abs(e_low - O.kinetic_energy_hat(uh, N, 3, O.init_forcing_mask_hit(N, 3))) < 10 * tol(prec) * e


# ---------------------------------------------------------------------------
# particle interpolation (integrator.py:26-93) and advection (combined.py:199)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("dim,n", [(3, 16), (3, 32), (2, 64)])
def test_interp_dft_exact(gpu, dim, n, prec):
    from l3es_amd.integrator import spectral_interpolator_wrapper

    rng = np.random.default_rng(n)
    u = rng.standard_normal((dim,) + (n,) * dim)
    r = rng.random((333, dim)) * 2 * np.pi
    ref = O.interpolate_dft(u, r)
    it = spectral_interpolator_wrapper(n, tuple(range(1, dim + 1)), backend="dft", precision=prec)
    out = host(it(dev(u, tdt(prec)), dev(r)))
    assert rel(out, ref) < (1e-12 if prec == "float64" else 5e-6)


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("dim,n", [(3, 16), (3, 32), (2, 64)])
def test_interp_nufft_tolerance(gpu, dim, n, prec):
    """NUFFT vs the exact DFT: relative l2 error <= tol = 1e-6 (FP64); FP32 to its floor."""
    from l3es_amd.integrator import spectral_interpolator_wrapper

    rng = np.random.default_rng(n + 1)
    u = rng.standard_normal((dim,) + (n,) * dim)
    r = rng.random((500, dim)) * 2 * np.pi
    ref = O.interpolate_dft(u, r)
    it = spectral_interpolator_wrapper(n, tuple(range(1, dim + 1)), backend="nufft", precision=prec, tol=1e-6)
    out = host(it(dev(u, tdt(prec)), dev(r)))
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err < (1e-6 if prec == "float64" else 2e-5)


@pytest.mark.parametrize("backend", ["nufft", "dft"])
def test_interp_hat_fused_and_advect(gpu, backend):
    """filter(irfftn(u_hat), n) -> interpolate -> shift, fused, vs the oracle chain."""
    from l3es_amd.integrator import spectral_interpolator_wrapper

    N, n, dt = 32, 16, 0.01
    u = O.init_u_hit(N, seed=5)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    rng = np.random.default_rng(3)
    r = rng.random((400, 3)) * 2 * np.pi
    ref = O.interpolate_dft(O.spectral_filtering(u, n), r)
    it = spectral_interpolator_wrapper(n, (1, 2, 3), backend=backend, precision="float64")
    rd = dev(r)
    ur = it.from_hat(dev(uh), rd, dt)
    t = 1e-12 if backend == "dft" else 1e-6
    assert np.linalg.norm(host(ur) - ref) / np.linalg.norm(ref) < t
    r_new = host(rd)
    np.testing.assert_allclose(r_new, O.shift_fn(r, dt * host(ur)), atol=1e-12)
    assert (r_new >= 0).all() and (r_new < 2 * np.pi).all()


@pytest.mark.parametrize("forcing", ["ekin_tot", "power"])
def test_step_multi_forced_equals_repeated(gpu, forcing):
    """Forced K-step call == K one-step calls, bit for bit (u_n kept in the padded buffer between steps)."""
    from l3es_amd import set_up_solver

    kw = dict(forcing_type=forcing, e_kin_target=0.5, power=0.2)
    u, uh, _, _, step, _, _, _, _ = set_up_solver(32, 0.01, 3, "HIT", 0.005, 7, 32, 2, precision="float32",
                                                  warmup=False, **kw)
    e2 = torch.zeros(3, dtype=torch.float64, device=uh.device)
    uh2, u2 = uh.clone(), torch.empty_like(u)
    step.advance(u2, uh2, 3, u2, e_inj=e2, use_u=False)
    uh3, u3 = uh.clone(), torch.empty_like(u)
    e3 = []
    for _ in range(3):
        e = torch.zeros(1, dtype=torch.float64, device=uh.device)
        step.advance(u3, uh3, 1, u3, e_inj=e, use_u=False)
        e3.append(float(e[0]))
    assert np.array_equal(host(uh2), host(uh3))
    assert np.array_equal(host(u2), host(u3))
    assert np.array_equal(host(e2), np.array(e3))


def test_step_advect_multi_equals_repeated(gpu):
    """qlles_step_advect(K) == K x (one step, then filter + interpolate + shift) on the same arrays."""
    from l3es_amd import _lib, set_up_solver
    from l3es_amd.integrator import spectral_interpolator_wrapper

    N, n, dt = 32, 16, 0.005
    u, uh, _, _, step, _, _, _, _ = set_up_solver(N, 0.01, 3, "HIT", dt, 11, n, 2, forcing_type="power",
                                                  precision="float64", warmup=False, power=0.1)
    it = spectral_interpolator_wrapper(n, (1, 2, 3), backend="nufft", precision="float64")
    r0 = np.random.default_rng(9).random((300, 3)) * 2 * np.pi
    lib = _lib.load()
    uh_a, r_a = uh.clone(), dev(r0)
    ur_a = torch.empty((300, 3), dtype=torch.float64, device=uh.device)
    _lib.check(lib.qlles_step_advect(step.plan.handle, it.it.handle, None, _lib.ptr(uh_a), 3, _lib.ptr(r_a), 300,
                                     _lib.ptr(ur_a), None, _lib.stream_ptr()))
    uh_b, r_b = uh.clone(), dev(r0)
    for _ in range(3):
        step.advance(None, uh_b, 1, None, use_u=False)
        ur_b = it.from_hat(uh_b, r_b, dt)
    torch.cuda.synchronize()
    assert np.array_equal(host(uh_a), host(uh_b))
    assert np.array_equal(host(ur_a), host(ur_b))
    assert np.array_equal(host(r_a), host(r_b))


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("dim,N", [(3, 16), (3, 32), (2, 32), (2, 64)])
def test_energy_spectrum_gpu(gpu, dim, N, prec):
    """Shell-binned E(k) on the GPU vs the oracle's restatement of utils.py:35-97 (random field)."""
    from l3es_amd.utils import energy_spectrum

    rng = np.random.default_rng(N + dim)
    shape = (3, N, N, N) if dim == 3 else (3, N, N, 1)
    u = rng.standard_normal(shape)
    if dim == 2:
        u[2] = 0.0
    ref = O.energy_spectrum(u, mul_fac=1.5, dim=dim)
    ek = energy_spectrum(dev(u, tdt(prec)), mul_fac=1.5, dim=dim)
    assert ek.shape == ref.shape
    assert np.abs(ek - ref).max() <= (1e-12 if prec == "float64" else 2e-6) * np.abs(ref).max()


@pytest.mark.parametrize("dim,N,n", [(3, 32, 16), (2, 64, 32)])
def test_write_u_frames(gpu, tmp_path, dim, N, n):
    """u_{ckp_N}_{step:05d}.npy frames (utils.py:141-166): shape and values of spectral_filtering."""
    from l3es_amd import _lib
    from l3es_amd.spectral_solver import write_u

    rng = np.random.default_rng(7)
    shape = (3, N, N, N) if dim == 3 else (3, N, N, 1)
    u = rng.standard_normal(shape)
    if dim == 2:
        u[2] = 0.0
    uh = np.fft.rfftn(u, axes=(1, 2, 3)) if dim == 3 else np.fft.rfftn(u[..., 0], axes=(1, 2))
    plan = _lib.Plan(dim, N, torch.float64, 0.0, 1.0)
    write_u(dev(uh), 7, str(tmp_path), plan, n)
    arr = np.load(tmp_path / f"u_{n}_00007.npy")
    ref = O.spectral_filtering(u if dim == 3 else u[:2, :, :, 0], n)
    assert arr.shape == ref.shape
    assert np.abs(arr - ref).max() < 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# the benchmark configurations that run as parity cases
# ---------------------------------------------------------------------------
def test_config_tgv64_fp64_dt1e3(gpu):
    """TGV 64^3 FP64, nu = 1/1600, dt = 1e-3, 10 RK4 steps: kinetic energy and full fields vs the oracle."""
    ug, uhg, _, _ = _run_gpu(64, 1.0 / 1600, 3, "TGV3D", 1e-3, 10, "float64", forcing_type="none")
    uo, uho, _, _ = _run_oracle(64, 1.0 / 1600, 3, "TGV3D", 1e-3, 10)
    eg = 0.5 * np.mean(np.sum(ug * ug, axis=0))
    eo = 0.5 * np.mean(np.sum(uo * uo, axis=0))
    assert abs(eg - eo) < 1e-13 * eo
    assert rel(ug, uo) < 1e-12 and rel(uhg, uho) < 1e-12


@pytest.mark.parametrize("prec,tf,tr", [("float64", 1e-11, 2e-7), ("float32", 2e-5, 2e-5)])
def test_config_kolmogorov256_particles(gpu, prec, tf, tr):
    """2-D Kolmogorov 256^2 (f_x = sin 4y, drag 0.1, nu = 1e-3, dt = 1e-3) with 64^2 jittered particles
    advected by the NUFFT (tol 1e-6): 20 combined-mode steps against the oracle's step + filter + DFT
    interpolation + shift chain (combined.py:97,127,130,199).  The random-phase initial field and the
    jittered lattice are the package's own (init_u_kolm, pos_init_jittered), checked against the oracle's."""
    from l3es_amd import _lib, set_up_solver
    from l3es_amd.integrator import pos_init_jittered, spectral_interpolator_wrapper

    N, n, K, dt = 256, 64, 20, 1e-3
    kw = dict(forcing_type="kolmogorov", drag=0.1, kolm_amp=1.0, kolm_k=4)
    u0 = O.init_u_kolm_np(N, seed=3)
    ug0, uh, _, _, step, _, _, _, _ = set_up_solver(N, 1e-3, 2, "Kolm", dt, 3, n, None, precision=prec,
                                                    warmup=False, **kw)
    assert ug0.shape == (3, N, N, 1)
    assert rel(host(ug0), u0) < (1e-13 if prec == "float64" else 1e-7)
    it = spectral_interpolator_wrapper(n, (1, 2), backend="nufft", precision=prec, tol=1e-6)
    r0 = pos_init_jittered(n, 2, seed=5)
    np.testing.assert_array_equal(r0, O.pos_init_jittered(n, 2, seed=5))
    r = dev(r0)
    ur = torch.empty((r.shape[0], 2), dtype=tdt(prec), device=r.device)
    lib = _lib.load()
    _lib.check(lib.qlles_step_advect(step.plan.handle, it.it.handle, None, _lib.ptr(uh), K, _lib.ptr(r), r.shape[0],
                                     _lib.ptr(ur), None, _lib.stream_ptr()))
    torch.cuda.synchronize()
    uo, uho, so, _, _ = O.set_up_solver(N, 1e-3, 2, "Kolm", dt, seed=3, warmup=False, u_init=u0, **kw)
    ro = r0.copy()
    for _ in range(K):
        uo, uho, _ = so(uo, uho)
        uro = O.interpolate_dft(O.spectral_filtering(uo[:2, :, :, 0], n), ro)
        ro = O.shift_fn(ro, dt * uro)
    assert rel(host(uh), uho) < tf
    d = np.abs(host(r) - ro)
    d = np.minimum(d, 2 * np.pi - d)  # periodic distance
    assert d.max() < tr


def test_simulate_driver_tgv(gpu, tmp_path):
    """The simulation driver end to end (spectral_solver.py:233-367) on TGV 32^3 FP64: the set-up's
    warm-up step plus t_final/dt + 1 steps against the oracle, the diagnostics CSV, and the spectrally
    filtered .npy checkpoints (utils.py:141-166)."""
    import csv as _csv
    import os

    from l3es_amd import simulate

    out = simulate("TGV3D", N=32, nu=0.000625, t_final=0.1, dt=0.01, ckp_N=16, log_freq=5, ckp_freq=5,
                   dst_path=str(tmp_path), precision="float64", verbose=False)
    u, uh, step, _, _ = O.set_up_solver(32, 0.000625, 3, "TGV3D", 0.01, seed=42, warmup=True)
    for _ in range(11):
        u, uh, _ = step(u, uh)
    assert rel(host(out["u"]), u) < 1e-12
    rows = list(_csv.reader(open(os.path.join(tmp_path, "diagnostics.csv"))))
    assert rows[0] == ["step", "time", "umax", "ekin", "dt_est", "e_inj"]
    assert [int(r[0]) for r in rows[1:]] == [0, 5, 10]
    assert abs(float(rows[-1][3]) - 0.5 * np.mean(np.sum(u * u, axis=0))) < 1e-12
    ck = np.load(os.path.join(tmp_path, "sim", "u_16_00010.npy"))
    assert ck.shape == (3, 16, 16, 16) and ck.dtype == np.float64
    assert rel(ck, O.spectral_filtering(u, 16)) < 1e-12
    assert np.load(os.path.join(tmp_path, "u_32_00000_burnin.npy")).shape == (3, 32, 32, 32)


def test_simulate_full_resolution_checkpoints(gpu, tmp_path):
    """The driver forms the physical u only on the steps that read it. Full-resolution checkpoints
    (ckp_N = N, written from u itself, utils.py:161) every 2 steps, diagnostics every 3 and the
    returned field must each be the state after their own step: TGV 32^3 FP64 against the oracle."""
    import csv as _csv
    import os

    from l3es_amd import simulate

    N, nu, dt = 32, 0.000625, 0.01
    out = simulate("TGV3D", N=N, nu=nu, t_final=0.06, dt=dt, ckp_N=N, log_freq=3, ckp_freq=2,
                   dst_path=str(tmp_path), precision="float64", verbose=False)
    u, uh, step, _, _ = O.set_up_solver(N, nu, 3, "TGV3D", dt, seed=42, warmup=True)
    states = []
    for _ in range(7):  # loop steps 0..6
        u, uh, _ = step(u, uh)
        states.append(u)
    for i in (0, 2, 4, 6):
        ck = np.load(os.path.join(tmp_path, "sim", f"u_{N}_{i:05d}.npy"))
        assert ck.shape == (3, N, N, N) and rel(ck, states[i]) < 1e-12, f"checkpoint of step {i}"
    assert rel(np.load(os.path.join(tmp_path, f"u_{N}_00000_burnin.npy")), states[0]) < 1e-12
    rows = list(_csv.reader(open(os.path.join(tmp_path, "diagnostics.csv"))))
    assert [int(r[0]) for r in rows[1:]] == [0, 3, 6]
    for r in rows[1:]:
        e = 0.5 * np.mean(np.sum(states[int(r[0])] ** 2, axis=0))
        assert abs(float(r[3]) - e) < 1e-12 and abs(float(r[2]) - np.abs(states[int(r[0])]).max()) < 1e-12
    assert rel(host(out["u"]), states[-1]) < 1e-12


def _oracle_hit_burnin(N, nu, dt, kf, e0, forced_steps, total_steps, log_freq):
    """The reference's forced-HIT driver loop (spectral_solver.py:300-320): the set-up's warm-up step,
    forced steps until the burn-in step, then steps without forcing; the injection rates summed, the
    kinetic energy at the logged steps and the state after each step."""
    u, uh, step, e_init, _ = O.set_up_solver(N, nu, 3, "HIT", dt, seed=42, kf=kf, e_kin_target=e0,
                                             forcing_type="ekin_tot", warmup=True)
    free = O.rk4_wrapper(dt, O.rhs_wrapper(N, nu), None, e_init, "none")
    e_acc, e_log, states = 0.0, [], []
    for i in range(total_steps):
        u, uh, ei = (step if i < forced_steps else free)(u, uh)
        e_acc += ei
        if i % log_freq == 0:
            e_log.append(O.kinetic_energy(u))
        states.append(u)
    return e_acc, e_log, states


def test_simulate_hit_statistics_and_burnin(gpu, tmp_path):
    """simulate's forced-HIT branch end to end: the forcing switched off at the burn-in step
    (spectral_solver.py:305-309), the injection rates summed (:316), the full-resolution _burnin snapshot
    after the burn-in step (:331-332) and the end-of-run statistics (:344-358) from them."""
    import os

    from l3es_amd import simulate
    from l3es_amd.spectral_solver import hit_statistics

    N, nu, dt, kf, e0 = 16, 0.01, 0.01, 3, 1.25
    out = simulate("HIT", N=N, nu=nu, t_final=0.05, t_burnin=0.02, dt=dt, ckp_N=8, forcing_type="ekin_tot",
                   e_kin_target=e0, kf=kf, log_freq=2, ckp_freq=1, dst_path=str(tmp_path), precision="float64",
                   verbose=False)
    e_acc, e_log, states = _oracle_hit_burnin(N, nu, dt, kf, e0, 2, 6, 2)
    assert e_acc > 0
    assert abs(out["e_inj_acc"] - e_acc) <= 1e-10 * e_acc
    assert rel(host(out["u"]), states[-1]) < 1e-12
    snap = np.load(os.path.join(tmp_path, "u_16_00002_burnin.npy"))
    assert rel(snap, states[2]) < 1e-12
    want = hit_statistics(e_acc, 5, kf, N, nu, 0.05, e_log[-1])
    for k, v in want.items():
        assert abs(out["hit_stats"][k] - v) <= 1e-10 * abs(v), k


def test_combined_burnin_snapshot(gpu, tmp_path):
    """combined writes the full-resolution field at the burn-in step before taking it
    (combined.py:89-90), then advances without forcing."""
    import os

    from l3es_amd.combined import combined

    N, nu, dt, kf, e0 = 16, 0.01, 0.01, 3, 1.25
    out = combined("HIT", N=N, dim=3, nu=nu, t_final=0.04, t_burnin=0.02, dt=dt, ckp_N=8, interp_backend="nufft",
                   forcing_type="ekin_tot", e_kin_target=e0, kf=kf, dst_path=str(tmp_path), precision="float64",
                   log_freq=100, ckp_freq=100, verbose=False)
    e_acc, _, states = _oracle_hit_burnin(N, nu, dt, kf, e0, 2, 5, 100)
    snap = np.load(os.path.join(tmp_path, "u_16_00002_burnin.npy"))
    assert snap.shape == (3, N, N, N)
    assert rel(snap, states[1]) < 1e-12  # the state before step 2
    assert rel(host(out["u"]), states[-1]) < 1e-12
    assert abs(out["e_inj_acc"] - e_acc) <= 1e-10 * e_acc
