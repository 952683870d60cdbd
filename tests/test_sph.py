"""SPH density and particle relaxation on the GPU (qlles_sph_*) against the
FP64 oracle (oracle/sph.py, restating relax.py:76-176, utils.py:169-218 and
combined.py:132-142,199).

Tolerances: densities 1e-12 relative; accelerations 1e-11 of the largest
component (cancellation in the pair sums); positions 1e-12 absolute.
"""

import numpy as np
import pytest
import torch

from oracle import sph as osph

L = 2 * np.pi
pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def positions(n, dim, jitter, seed, edge=True):
    rng = np.random.default_rng(seed)
    r = osph.lattice(n, dim) + rng.normal(size=(n ** dim, dim)) * jitter * L / n
    r = np.mod(r, L)
    if edge:  # points on the periodic seam
        r[0] = 0.0
        r[1, 0] = np.nextafter(L, 0.0)
    return r


CASES = [(16, 3, 0.25), (8, 3, 0.3), (10, 3, 0.1), (32, 2, 0.3), (7, 2, 0.2), (64, 2, 0.05)]


def dev(a, gpu):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(gpu)


@pytest.mark.parametrize("n,dim,jit", CASES)
def test_density_matches_oracle(gpu, n, dim, jit):
    from l3es_amd import rho_computer

    r = positions(n, dim, jit, n)
    rho = host(rho_computer(n, dim)(dev(r, gpu)))
    ref = osph.density(r, n, dim)
    assert np.max(np.abs(rho / ref - 1)) < 1e-12


@pytest.mark.parametrize("is_physical", [True, False])
@pytest.mark.parametrize("n,dim,jit", CASES)
def test_relax_acc_matches_oracle(gpu, n, dim, jit, is_physical):
    from l3es_amd import relax_wrapper

    r = positions(n, dim, jit, 7 * n)
    loop_body, _ = relax_wrapper(n, dim, L, is_physical=is_physical, u_ref=1.5)
    acc = host(loop_body(dev(r, gpu)))
    ref = osph.relax_acc(r, n, dim, L, is_physical, 1.5)
    assert np.max(np.abs(acc - ref)) < 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,dim,jit", [(12, 3, 0.2), (32, 2, 0.3)])
def test_sph_forces_match_oracle(gpu, n, dim, jit):
    from l3es_amd import relax_wrapper

    r = positions(n, dim, jit, 3)
    u = np.random.default_rng(4).normal(size=r.shape)
    _, sph_body = relax_wrapper(n, dim, L, is_physical=True, u_ref=0.7)
    out = sph_body(dev(r, gpu), dev(u, gpu))
    ref = osph.sph_forces(r, u, n, dim, L, 0.7)
    for k in ("acc_p", "acc_visc", "acc_tvf"):
        assert np.max(np.abs(host(out[k]) - ref[k])) < 1e-11 * np.max(np.abs(ref[k])), k


@pytest.mark.parametrize("u_dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,dim", [(16, 3), (32, 2)])
def test_relax_advect_matches_oracle(gpu, n, dim, u_dtype):
    """combined.py:132-142,199 with relax_dt_factor 2 (hit32_1.yaml:32)."""
    from l3es_amd import relax_advect, relax_wrapper

    r = positions(n, dim, 0.2, 11)
    u = np.random.default_rng(5).normal(size=r.shape) * 2.0
    dt, f = 0.01, 2.0
    loop_body, _ = relax_wrapper(n, dim, L, is_physical=True, u_ref=4.0)
    r_d, u_d = dev(r, gpu), dev(u, gpu).to(u_dtype)
    u_in = host(u_d).astype(np.float64)  # what the kernel sees
    acc = torch.empty_like(r_d)
    relax_advect(loop_body, r_d, u_d, dt, f, acc=acc)
    r_ref, u_ref, a_ref = osph.relax_advect(r, u_in, dt, f, n, dim, L, 4.0)
    assert np.max(np.abs(host(acc) - a_ref)) < 1e-11 * np.max(np.abs(a_ref))
    d = np.mod(host(r_d) - r_ref + L / 2, L) - L / 2
    assert np.max(np.abs(d)) < 1e-12
    tol = 1e-12 if u_dtype == torch.float64 else 1e-6
    assert np.max(np.abs(host(u_d).astype(np.float64) - u_ref)) < tol * np.max(np.abs(u_ref))
    assert host(r_d).min() >= 0.0 and host(r_d).max() <= L


def test_bitwise_reproducible_and_permutation_equivariant(gpu):
    from l3es_amd import relax_wrapper

    n, dim = 16, 3
    r = positions(n, dim, 0.3, 21)
    loop_body, _ = relax_wrapper(n, dim, L, is_physical=True, u_ref=1.0)
    a1 = host(loop_body(dev(r, gpu)))
    a2 = host(loop_body(dev(r, gpu)))
    assert np.array_equal(a1, a2)
    p = np.random.default_rng(0).permutation(r.shape[0])
    a3 = host(loop_body(dev(r[p], gpu)))
    assert np.max(np.abs(a3 - a1[p])) < 1e-11 * np.max(np.abs(a1))


def test_particle_counts_other_than_the_lattice(gpu):
    """np below n^dim, np above it (workspace growth), np == 0, and a cluster
    of particles sharing few cells."""
    from l3es_amd import _lib, relax_wrapper, rho_computer

    n, dim = 12, 3
    comp_rho = rho_computer(n, dim)
    loop_body, _ = relax_wrapper(n, dim, L, is_physical=True, u_ref=1.0)
    rng = np.random.default_rng(8)
    for count in (100, n ** dim + 777):
        r = rng.random((count, dim)) * L
        assert np.max(np.abs(host(comp_rho(dev(r, gpu))) / osph.density(r, n, dim) - 1)) < 1e-12
    cluster = np.mod(1.0 + rng.normal(size=(600, dim)) * 0.3 * L / n, L)
    a = host(loop_body(dev(cluster, gpu)))
    ref = osph.relax_acc(cluster, n, dim, L, True, 1.0)
    assert np.max(np.abs(a - ref)) < 1e-11 * np.max(np.abs(ref))
    empty = torch.empty((0, dim), dtype=torch.float64, device=gpu)
    assert comp_rho(empty).shape == (0,)
    lib = _lib.load()
    assert lib.qlles_sph_relax_advect(loop_body.ctx.handle, None, 4, None, 1, 0.1, 1.0, None, None) \
        == _lib.QLLES_ERR_ARG
    t = torch.zeros((4, dim), dtype=torch.float64, device=gpu)
    assert lib.qlles_sph_relax_advect(loop_body.ctx.handle, _lib.ptr(t), 4, _lib.ptr(t), 1, 0.0, 1.0, None,
                                      None) == _lib.QLLES_ERR_ARG


def test_relaxation_requires_physical_units(gpu):
    from l3es_amd import _lib, relax_wrapper

    _, sph_body = relax_wrapper(12, 2, L, is_physical=False, u_ref=1.0)
    with pytest.raises(AssertionError):
        sph_body(np.zeros((4, 2)), np.zeros((4, 2)))
    with pytest.raises(TypeError):
        relax_wrapper(12, 2, L)


def test_combined_with_relaxation_matches_stepwise_oracle(gpu, tmp_path):
    """combined(relax_dt_factor=2) over a few steps against the same loop run
    by hand: the GPU interpolation at r, then the oracle's relaxation and
    shift on the host (combined.py:97,127-142,199)."""
    from l3es_amd.combined import combined
    from l3es_amd.integrator import spectral_interpolator_wrapper
    from l3es_amd.spectral_solver import set_up_solver

    N, ckp_N, dt, steps, f, u_ref = 32, 16, 0.01, 3, 2.0, 1.0
    out = combined("TGV3D", N=N, dim=3, nu=0.01, t_final=dt * (steps - 1), dt=dt, u_ref=u_ref, ckp_N=ckp_N,
                   interp_backend="nufft", relax_dt_factor=f, seed=0, log_freq=1, ckp_freq=1,
                   dst_path=str(tmp_path), precision="float64", verbose=False)
    assert len(out["rows"]) == steps and all(np.isfinite(row[6]) for row in out["rows"])
    assert (tmp_path / "com" / "step_00000.npz").exists()
    u, u_hat, *_rest = set_up_solver(N, 0.01, 3, "TGV3D", dt, 0, ckp_N, None, None, "none", True, "float64")
    integrate_fn = _rest[2]
    interp = spectral_interpolator_wrapper(ckp_N, (1, 2, 3), backend="nufft", precision="float64")
    r = osph.lattice(ckp_N, 3)
    e = torch.zeros(1, dtype=torch.float64, device=gpu)
    rho0 = None
    frames = []
    for i in range(steps):
        integrate_fn.advance(u, u_hat, 1, u, e)
        u_r = host(interp.from_hat(u_hat, dev(r, gpu), 0.0))
        frames.append((r.copy(), u_r))
        if i == 0:
            rho0 = osph.density(r, ckp_N, 3).max()
        r, _, _ = osph.relax_advect(r, u_r, dt, f, ckp_N, 3, L, u_ref)
    d = np.mod(host(out["r"]) - r + L / 2, L) - L / 2
    assert np.max(np.abs(d)) < 1e-10
    assert abs(out["rows"][0][6] - rho0) < 1e-12
    # every frame holds the positions and the interpolated velocities of its own step (combined.py:167-169)
    for i, (r_i, u_i) in enumerate(frames):
        fr = np.load(tmp_path / "com" / f"step_{i:05d}.npz")
        assert np.max(np.abs(np.mod(fr["r"] - r_i + L / 2, L) - L / 2)) < 1e-10, f"frame {i}: positions"
        assert np.max(np.abs(fr["u"] - u_i)) < 1e-10, f"frame {i}: velocities"


def test_combined_2d_dft_with_relaxation_matches_stepwise_oracle(gpu, tmp_path):
    """The Kolmogorov dataset loop in small: 2-D, DFT interpolation, relax_dt_factor sqrt 2
    (kolm64_1.yaml:33), against the GPU interpolation followed by the oracle's relaxation."""
    from l3es_amd.combined import combined
    from l3es_amd.integrator import spectral_interpolator_wrapper
    from l3es_amd.spectral_solver import set_up_solver

    # log_freq past the run: only step 0 reads the GPU, so the loop queues the other steps (and their
    # frame copies) ahead of the GPU
    N, ckp_N, dt, steps, f, u_ref = 64, 16, 0.005, 6, 1.4142135624, 6.0
    out = combined("Kolm", N=N, dim=2, nu=0.001, t_final=dt * (steps - 1), dt=dt, u_ref=u_ref, ckp_N=ckp_N,
                   interp_backend="dft", dft_splits=2, relax_dt_factor=f, seed=3, log_freq=100, ckp_freq=1,
                   dst_path=str(tmp_path), precision="float64", verbose=False)
    assert out["steps"] == steps and all(np.isfinite(row[6]) for row in out["rows"])
    u, u_hat, *_rest = set_up_solver(N, 0.001, 2, "Kolm", dt, 3, ckp_N, None, None, "none", True, "float64")
    integrate_fn = _rest[2]
    interp = spectral_interpolator_wrapper(ckp_N, (1, 2), backend="dft", precision="float64")
    r = osph.lattice(ckp_N, 2)
    e = torch.zeros(1, dtype=torch.float64, device=gpu)
    frames = []
    for _ in range(steps):
        integrate_fn.advance(u, u_hat, 1, u, e)
        u_r = host(interp.from_hat(u_hat, dev(r, gpu), 0.0))
        frames.append((r.copy(), u_r))
        r, _, _ = osph.relax_advect(r, u_r, dt, f, ckp_N, 2, L, u_ref)
    d = np.mod(host(out["r"]) - r + L / 2, L) - L / 2
    assert np.max(np.abs(d)) < 1e-10
    for i, (r_i, u_i) in enumerate(frames):
        fr = np.load(tmp_path / "com" / f"step_{i:05d}.npz")
        assert np.max(np.abs(np.mod(fr["r"] - r_i + L / 2, L) - L / 2)) < 1e-10, f"frame {i}: positions"
        assert np.max(np.abs(fr["u"] - u_i)) < 1e-10, f"frame {i}: velocities"
