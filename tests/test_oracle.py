"""Pins of the FP64 oracle: the reference's known-answer value, closed forms,
invariants and brute-force twins (CPU only)."""

import numpy as np
import pytest

import oracle as O

TWO_PI = 2 * np.pi


def rk4_poly(z):
    return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24


def kat(name):
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")) as fh:
        return json.load(fh)[name]


def test_tgv_known_answer():
    """from_spectralDNS/main_numpy.py:126-137: TGV 64^3, nu=6.25e-4, dt=0.01, 10 steps."""
    k = kat("tgv3d_numpy")
    u, uh, step, _, _ = O.set_up_solver(k["N"], k["nu"], 3, k["case"], k["dt"], warmup=False)
    for _ in range(k["steps"]):
        u, uh, _ = step(u, uh)
    assert round(O.kinetic_energy(u) - k["e_kin"], k["round_digits"]) == 0


def _node_grid(N, dim=3):
    return np.array(np.meshgrid(*([np.arange(N)] * dim), indexing="ij")) * TWO_PI / N


@pytest.mark.parametrize("viscous_mode", ["explicit", "IF"])
def test_shear_mode_decay(viscous_mode):
    """u = A sin(k y) x: u x omega is a gradient, so only the viscous term acts.
    Explicit RK4 gives R(-nu k^2 dt) per step, the integrating factor exp(-nu k^2 dt)."""
    N, nu, dt, A, k = 16, 0.05, 0.02, 1.3, 3
    xyz = _node_grid(N)
    u0 = np.zeros((3, N, N, N))
    u0[0] = A * np.sin(k * xyz[1])
    u, uh, step, _, _ = O.set_up_solver(N, nu, 3, "TGV3D", dt, warmup=False, u_init=u0,
                                        viscous_mode=viscous_mode)
    for _ in range(5):
        u, uh, _ = step(u, uh)
    z = -nu * k * k * dt
    g = rk4_poly(z) ** 5 if viscous_mode == "explicit" else np.exp(5 * z)
    np.testing.assert_allclose(u[0], g * u0[0], atol=1e-13)
    np.testing.assert_allclose(u[1:], 0, atol=1e-13)


def test_tgv2d_decay():
    """2-D TGV (init_fields.py:19-30) is an exact solution with k^2 = 2."""
    N, nu, dt = 32, 0.06283, 0.001
    u, uh, step, _, _ = O.set_up_solver(N, nu, 2, "TGV2D", dt, warmup=False)
    u0 = u.copy()
    for _ in range(7):
        u, uh, _ = step(u, uh)
    np.testing.assert_allclose(u, rk4_poly(-2 * nu * dt) ** 7 * u0, atol=1e-12)


def test_rogallo_vectorized_matches_loop_twin():
    """init_u_hit (vectorized) == init_u_hit_slow (scalar loop twin, init_fields.py:109-164)."""
    for spectrum in ("k53", "k4exp"):
        a = O.init_u_hit(8, seed=5, spectrum=spectrum)
        b = O.init_u_hit_slow(8, seed=5, spectrum=spectrum)
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_rogallo_divergence_free_and_seeded():
    N = 16
    u = O.init_u_hit(N, seed=3)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    kkk, _, _, _ = O.wavenumbers(N, 3)
    div = np.sum(kkk * uh, axis=0)
    # interior modes carry no divergence; the Nyquist planes of x and y are
    # aliased +-N/2 modes whose Hermitian projection (irfftn) is not solenoidal
    keep = np.ones(N, bool)
    keep[N // 2] = False
    d = div[keep][:, keep][..., 1:N // 2]
    assert np.abs(d).max() < 1e-10 * np.abs(uh).max()
    np.testing.assert_array_equal(u, O.init_u_hit(N, seed=3))
    assert not np.array_equal(u, O.init_u_hit(N, seed=4))


def test_forcing_mask_enumeration():
    N, kf = 8, 2
    m = O.init_forcing_mask_hit(N, kf)
    ref = np.zeros_like(m)
    for i in range(N):
        for j in range(N):
            for k in range(N // 2 + 1):
                kx = i if i < N // 2 else i - N
                ky = j if j < N // 2 else j - N
                kz = k  # rfft axis: Nyquist index carries +N/2
                ref[i, j, k] = 0 < kx * kx + ky * ky + kz * kz <= kf * kf
    np.testing.assert_array_equal(m, ref)


def test_ekin_tot_forcing_keeps_energy():
    """spectral_solver.py:82-112: the rescaled step returns E == E_init."""
    N = 16
    u, uh, step, e0, _ = O.set_up_solver(N, 0.01, 3, "HIT", 0.01, seed=2, kf=3, e_kin_target=1.25,
                                         forcing_type="ekin_tot", warmup=False)
    assert abs(e0 - 1.25) < 1e-12
    for _ in range(3):
        u, uh, e_inj = step(u, uh)
        assert abs(O.kinetic_energy(u) - e0) < 1e-12
        assert e_inj >= 0


def test_power_forcing_injects_power():
    """Fixed-power forcing: <f . u> == P on the forced shell."""
    N, P = 16, 0.1
    u = O.init_u_hit(N, seed=1)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    f = O.make_forcing("power", N, 3, kf=2, power=P)
    fu = np.fft.irfftn(f(uh), axes=(1, 2, 3))
    assert abs(np.mean(np.sum(fu * u, axis=0)) - P) < 1e-12


def test_kolmogorov_laminar_fixed_point():
    """u_x = sin(4y)/(16 nu + drag) is a steady state of the forced, damped 2-D flow."""
    N, nu, drag = 32, 1e-3, 0.1
    y = np.arange(N) * TWO_PI / N
    u0 = np.zeros((3, N, N, 1))
    u0[0, :, :, 0] = np.sin(4 * y)[None, :] / (16 * nu + drag)
    u, uh, step, _, _ = O.set_up_solver(N, nu, 2, "Kolm", 1e-3, warmup=False, u_init=u0,
                                        forcing_type="kolmogorov", drag=drag)
    for _ in range(5):
        u, uh, _ = step(u, uh)
    np.testing.assert_allclose(u, u0, atol=1e-11)


def test_filter_is_exact_on_band_limited_fields():
    """spectral_filtering of a field band-limited below n/2 samples it on the n-grid."""
    N, n = 32, 16
    x = _node_grid(N)
    xs = _node_grid(n)

    def field(g):
        return np.array([np.sin(3 * g[1]) * np.cos(2 * g[2]), np.cos(g[0] + 5 * g[2]), np.sin(7 * g[0] - 4 * g[1])])

    np.testing.assert_allclose(O.spectral_filtering(field(x), n), field(xs), atol=1e-13)


def test_filter_real_projection_of_nyquist_face():
    """The crop keeps -n/2 without +n/2; .real averages the two (utils.py:127)."""
    N, n = 32, 8
    x = _node_grid(N)
    u = np.array([np.cos(4 * x[0] + 1 * x[1]), np.sin(4 * x[0]), x[0] * 0])
    # the kept coefficient of exp(-4i x) is half of cos; .real makes it a cosine of half amplitude
    xs = _node_grid(n)
    expect = np.array([0.5 * np.cos(4 * xs[0] + xs[1]), 0 * xs[0], 0 * xs[0]])
    np.testing.assert_allclose(O.spectral_filtering(u, n), expect, atol=1e-13)


def test_dft_interpolation_exact_and_at_nodes():
    n = 8
    g = _node_grid(n)
    u = np.array([np.sin(g[0] + 2 * g[1]), np.cos(3 * g[2]) * np.sin(g[1]), np.cos(g[0] - g[2])])
    r_nodes = np.stack([a.ravel() for a in g], axis=1)
    np.testing.assert_allclose(O.interpolate_dft(u, r_nodes), u.reshape(3, -1).T, atol=1e-13)
    rng = np.random.default_rng(0)
    r = rng.random((50, 3)) * TWO_PI
    ex = np.stack([np.sin(r[:, 0] + 2 * r[:, 1]), np.cos(3 * r[:, 2]) * np.sin(r[:, 1]),
                   np.cos(r[:, 0] - r[:, 2])], axis=1)
    np.testing.assert_allclose(O.interpolate_dft(u, r), ex, atol=1e-13)


def test_uniform_translation_advection():
    n = 8
    u = np.zeros((3, n, n, n))
    U = np.array([0.3, -1.1, 2.0])
    u += U[:, None, None, None]
    r = np.random.default_rng(1).random((20, 3)) * TWO_PI
    ur = O.interpolate_dft(u, r)
    np.testing.assert_allclose(ur, np.broadcast_to(U, ur.shape), atol=1e-13)
    dt = 0.7
    np.testing.assert_allclose(O.shift_fn(r, dt * ur), (r + dt * U) % TWO_PI, atol=1e-13)


def test_dissipation_single_mode():
    """u = A sin(k y) x: |omega|^2 = A^2 k^2 cos^2, <> = A^2 k^2 / 2."""
    N, nu, A, k = 16, 0.03, 0.7, 2
    x = _node_grid(N)
    u = np.zeros((3, N, N, N))
    u[0] = A * np.sin(k * x[1])
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    assert abs(O.comp_dissipation_rate(uh, nu) - nu * A * A * k * k / 2) < 1e-14


def test_energy_spectrum_single_mode():
    N, k = 16, 3
    x = _node_grid(N)
    u = np.zeros((3, N, N, N))
    u[1] = np.cos(k * x[0])
    ek = O.energy_spectrum(u)
    e = 0.5 * np.mean(np.sum(u * u, axis=0))
    # all energy sits in shell k; the shell-average normalisation 4 pi k^2 / n_samples
    assert np.argmax(ek) == k
    assert abs(np.sum(ek[np.arange(N) != k])) < 1e-14
    nk = sum(1 for i in range(N) for j in range(N) for m in range(N)
             if int(np.sqrt(min(i, N - i) ** 2 + min(j, N - j) ** 2 + min(m, N - m) ** 2) + 0.5) == k)
    # utils.py:85-95 bins |u^|^2 (no factor 1/2): the shell holds <|u|^2> = 2 E
    assert abs(ek[k] - 2 * e * 4 * np.pi * k * k / nk) < 1e-12


def test_kinetic_energy_parseval():
    u = O.init_u_hit(16, seed=9)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    assert abs(O.kinetic_energy_hat(uh, 16, 3) - O.kinetic_energy(u)) < 1e-13
