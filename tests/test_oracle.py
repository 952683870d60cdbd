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


# --------------------------------------------------------------------------
# Dealias edge (spectral_solver.py:42-48): strict |k_i| < 2/3 (N//2 + 1)
# --------------------------------------------------------------------------
# largest kept |k| per axis, from the exact integer form 3|k| < 2 (N//2 + 1)
# evaluated by hand: 2 (N/2+1) / 3 = 10/3, 6, 34/3, 22, 98/3, 130/3, 86
KEPT_MAX = {8: 3, 16: 5, 32: 11, 64: 21, 96: 32, 128: 43, 256: 85}


@pytest.mark.parametrize("N", sorted(KEPT_MAX))
def test_dealias_mask_edge_by_enumeration(N):
    kkk, kkk2, k_over_k2, dealias = O.wavenumbers(N, 3)
    full = [i if i < N // 2 else i - N for i in range(N)]
    half = list(range(N // 2)) + [N // 2]  # rfft axis: the Nyquist entry is +N/2
    assert list(kkk[0][:, 0, 0]) == full and list(kkk[1][0, :, 0]) == full
    assert list(kkk[2][0, 0, :]) == half
    keep_1d = lambda k: 3 * abs(k) < 2 * (N // 2 + 1)  # noqa: E731
    assert max(k for k in range(N) if keep_1d(k)) == KEPT_MAX[N]
    ref = np.array([[[keep_1d(a) and keep_1d(b) and keep_1d(c) for c in half] for b in full] for a in full])
    np.testing.assert_array_equal(dealias, ref)
    # rfft-axis columns kept: 0 .. KEPT_MAX (22 at N=64, 33 at N=96, 86 at N=256)
    assert int(dealias[0, 0].sum()) == KEPT_MAX[N] + 1
    # k^2 and k/k^2 use the +N/2 Nyquist column and k = 0 -> divide by 1
    assert kkk2[0, 0, N // 2] == (N // 2) ** 2
    assert np.all(k_over_k2[:, 0, 0, 0] == 0)
    np.testing.assert_allclose(k_over_k2[2, 0, 0, N // 2], 1.0 / (N // 2), rtol=0)


def test_dealias_mask_2d_pads_zero_kz():
    N = 32
    kkk, kkk2, _, dealias = O.wavenumbers(N, 2)
    assert kkk.shape == (3, N, N // 2 + 1) and np.all(kkk[2] == 0)
    full = [i if i < N // 2 else i - N for i in range(N)]
    keep_1d = lambda k: 3 * abs(k) < 2 * (N // 2 + 1)  # noqa: E731
    ref = np.array([[keep_1d(a) and keep_1d(b) for b in list(range(N // 2)) + [N // 2]] for a in full])
    np.testing.assert_array_equal(dealias, ref)


def _solenoidal_modes(N, modes, seed):
    """Real field sum_m Re(a_m exp(i k_m.x)) with a_m orthogonal to k_m."""
    rng = np.random.default_rng(seed)
    x = _node_grid(N)
    u = np.zeros((3, N, N, N))
    for k in modes:
        k = np.asarray(k, float)
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        a -= k * (k @ a) / (k @ k)  # k . a = 0
        ph = np.exp(1j * (k[0] * x[0] + k[1] * x[1] + k[2] * x[2]))
        u += np.real(a[:, None, None, None] * ph)
    return u


def test_dealias_edge_on_the_nonlinear_term():
    """Mode pairs whose products land on |k_x| = 5 (the last kept index at N=16,
    2/3 * 9 = 6) and on |k_x| = 6 (the first dropped): the inviscid RHS is
    non-zero on the edge and exactly zero past it."""
    N = 16
    u = _solenoidal_modes(N, [(3, 1, 0), (2, 0, 1), (3, -1, 2), (1, 2, 1)], seed=7)
    assert np.abs(np.fft.rfftn(u, axes=(1, 2, 3))).max() > 1
    rhs = O.rhs_wrapper(N, 0.0, (1, 2, 3))
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    du = rhs(u, uh, 0)
    kkk = O.wavenumbers(N, 3)[0]
    kabs = np.max(np.abs(kkk), axis=0)
    mag = np.sqrt(np.sum(np.abs(du) ** 2, axis=0))
    assert mag[kabs >= 6].max() == 0.0
    # products of (3,.,.) with (2,.,.) reach k_x = 5, products of (3,.,.) with (3,.,.) reach 6
    assert mag[(np.abs(kkk[0]) == 5)].max() > 1e-3 * mag.max()
    full = O.rhs_wrapper(N, 0.0, (1, 2, 3))
    full_nl = np.fft.rfftn(np.cross(u, np.fft.irfftn(1j * np.cross(kkk, uh, axis=0), axes=(1, 2, 3)), axis=0),
                           axes=(1, 2, 3))
    # without the mask the k_x = 6 modes would carry a non-gradient remainder
    p = np.sum(full_nl * O.wavenumbers(N, 3)[2], axis=0)
    undealiased = full_nl - p * kkk
    assert np.sqrt(np.sum(np.abs(undealiased) ** 2, axis=0))[np.abs(kkk[0]) == 6].max() > 1e-3 * mag.max()
    del full


# --------------------------------------------------------------------------
# ekin_tot forcing scalars (spectral_solver.py:93-112): e_inj and alpha
# --------------------------------------------------------------------------
def _shear(N, A, k, comp=0, axis=1):
    x = _node_grid(N)
    u = np.zeros((3, N, N, N))
    u[comp] = A * np.sin(k * x[axis])
    return u


@pytest.mark.parametrize("ratio", [2.0, 0.5])
def test_e_inj_closed_form_both_signs(ratio):
    """A shear mode decays exactly by R(-nu k^2 dt) per RK4 step, so
    E_new = E0 R^2 and e_inj = max(0, (E_init - E_new) / dt) (:109): positive
    when E_init = 2 E0, zero (not negative) when E_init = E0 / 2."""
    N, nu, dt, A, k = 16, 0.05, 0.02, 1.1, 2
    u = _shear(N, A, k)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    E0 = A * A / 4
    e_init = ratio * E0
    mask = O.init_forcing_mask_hit(N, 3)
    step = O.rk4_wrapper(dt, O.rhs_wrapper(N, nu), mask, e_init, "ekin_tot")
    _, uh1, e_inj = step(u, uh)
    R = rk4_poly(-nu * k * k * dt)
    expect = max(0.0, (e_init - E0 * R * R) / dt)
    assert abs(e_inj - expect) < 1e-12 * max(1.0, expect)
    if ratio < 1:
        assert e_inj == 0.0
    else:
        assert e_inj > 1.0
    # the shear mode lies inside the kf = 3 shell: E_high = 0, alpha^2 = E_init / E_new
    assert abs(O.kinetic_energy_hat(uh1, N, 3) - e_init) < 1e-12


def test_ekin_tot_alpha_clamps_at_zero():
    """Two modes of one shear flow u_x(y), k = 1 inside the kf = 1 shell and
    k = 4 outside (u x omega stays a gradient): after the step E_low and E_high
    follow R(-nu k^2 dt)^2; alpha = sqrt(max(0, (E_init -
    E_high) / (E_low + eps))) (:104-106) restores E_init when E_init > E_high
    and removes the shell (alpha = 0) when E_init < E_high."""
    N, nu, dt = 16, 0.05, 0.02
    A1, A4 = 0.8, 1.3
    u = _shear(N, A1, 1, comp=0, axis=1) + _shear(N, A4, 4, comp=0, axis=1)
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    mask = O.init_forcing_mask_hit(N, 1)
    e_low = A1 * A1 / 4 * rk4_poly(-nu * dt) ** 2
    e_high = A4 * A4 / 4 * rk4_poly(-16 * nu * dt) ** 2
    for e_init, low_after in ((e_high + 0.3, 0.3), (0.5 * e_high, 0.0)):
        step = O.rk4_wrapper(dt, O.rhs_wrapper(N, nu), mask, e_init, "ekin_tot")
        _, uh1, e_inj = step(u, uh)
        assert abs(O.kinetic_energy_hat(uh1, N, 3, mask) - low_after) < 1e-12
        assert abs(O.kinetic_energy_hat(uh1, N, 3) - (low_after + e_high)) < 1e-12
        assert abs(e_inj - max(0.0, (e_init - e_low - e_high) / dt)) < 1e-10
    del e_low


# --------------------------------------------------------------------------
# comp_dt (spectral_solver.py:123-127): cfl / min_i(max|u_i| / dx + nu / dx^2)
# --------------------------------------------------------------------------
def test_comp_dt_closed_form():
    """Three shear components of amplitudes A, B, C: max|u_i| is the amplitude
    (sin reaches 1 on the node grid), and the minimum over the components of
    max|u_i|/dx + nu/dx^2 belongs to the smallest amplitude."""
    N, nu, cfl = 16, 0.01, 0.7
    dx = TWO_PI / N
    A, B, C = 2.0, 0.5, 1.25
    u = _shear(N, A, 1, 0, 1) + _shear(N, B, 1, 1, 2) + _shear(N, C, 1, 2, 0)
    dt = O.comp_dt(u, dx, nu, cfl)
    assert abs(dt - cfl / (min(A, B, C) / dx + nu / dx ** 2)) < 1e-15
    # one non-zero component: the other two give nu / dx^2 alone
    assert abs(O.comp_dt(_shear(N, A, 1), dx, nu) - dx * dx / nu) < 1e-12


# --------------------------------------------------------------------------
# Project-defined 2-D random-phase init (the Kolmogorov configuration)
# --------------------------------------------------------------------------
def _shell_enstrophy(u, N):
    uh = np.fft.rfftn(u[0, :, :, 0])
    vh = np.fft.rfftn(u[1, :, :, 0])
    kx = np.fft.fftfreq(N, 1.0 / N)
    ky = np.abs(kx[: N // 2 + 1])
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    w = 1j * KX * vh - 1j * KY * uh
    wgt = np.full(N // 2 + 1, 2.0)
    wgt[0] = wgt[-1] = 1.0
    shell = np.rint(np.sqrt(KX ** 2 + KY ** 2)).astype(int)
    z = np.zeros(N)
    np.add.at(z, shell.ravel(), (np.abs(w) ** 2 * wgt).ravel())
    cnt = np.zeros(N)
    np.add.at(cnt, shell.ravel(), np.broadcast_to(wgt, shell.shape).ravel())
    return z, cnt, (KX, KY, uh, vh)


def test_kolm_init_divergence_free_scaled_and_seeded():
    N = 64
    u = O.init_u_kolm_np(N, seed=3)
    assert u.shape == (3, N, N, 1) and np.all(u[2] == 0)
    assert abs(np.abs(u[:2]).max() - 4.2) < 1e-12
    _, _, (KX, KY, uh, vh) = _shell_enstrophy(u, N)
    div = KX * uh + KY * vh
    assert np.abs(div).max() < 1e-12 * np.abs(uh).max() * N
    np.testing.assert_array_equal(u, O.init_u_kolm_np(N, seed=3))
    assert not np.allclose(u, O.init_u_kolm_np(N, seed=4))


def test_kolm_init_vorticity_spectrum_peaks_at_k0():
    """White-noise vorticity shaped so that the shell enstrophy is
    Z(k) ~ k^4 exp(-2 (k/k0)^2), which peaks at k0 = 4: over 24 seeds the shell
    sums peak at shell 4, and their ratio to the sum of |k|^3 exp(-2 (|k|/k0)^2)
    over each shell's modes (the per-mode density of that Z) is flat."""
    N, k0 = 64, 4.0
    zs = []
    for s in range(24):
        z, cnt, (KX, KY, _, _) = _shell_enstrophy(O.init_u_kolm_np(N, seed=s), N)
        zs.append(z / z.sum())
    z = np.mean(zs, axis=0)
    assert int(np.argmax(z)) == 4
    kmag = np.sqrt(KX ** 2 + KY ** 2)
    wgt = np.full(N // 2 + 1, 2.0)
    wgt[0] = wgt[-1] = 1.0
    shell = np.rint(kmag).astype(int)
    tgt = np.zeros(N)
    np.add.at(tgt, shell.ravel(), (wgt * kmag ** 3 * np.exp(-2 * (kmag / k0) ** 2)).ravel())
    k = np.arange(1, 11)
    ratio = z[k] / tgt[k]
    slope = np.polyfit(np.log(k), np.log(ratio), 1)[0]
    assert abs(slope) < 0.1, slope  # a stray factor k would give slope 1
    assert ratio.max() / ratio.min() < 1.5


# --------------------------------------------------------------------------
# Jittered lattice (the Kolmogorov configuration's 64^2 particles)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_jittered_lattice_bounds_and_seed(dim):
    n, jit = 16, 0.25
    L = TWO_PI
    dx = L / n
    r = O.pos_init_jittered(n, dim, L, jit, seed=5)
    assert r.shape == (n ** dim, dim)
    assert np.all(r >= 0) and np.all(r < L)
    idx = np.stack([a.ravel() for a in np.meshgrid(*([np.arange(n)] * dim), indexing="ij")], axis=1)
    centre = (idx + 0.5) * dx
    d = (r - centre + L / 2) % L - L / 2  # periodic offset from the cell centre
    assert np.abs(d).max() <= jit * dx * (1 + 1e-12)
    assert np.abs(d).max() > 0.95 * jit * dx  # the jitter spans its full width
    assert abs(d.mean()) < 0.02 * dx
    assert abs(d.std() - jit * dx / np.sqrt(3)) < 0.02 * dx  # uniform on [-jit, jit] dx
    np.testing.assert_array_equal(r, O.pos_init_jittered(n, dim, L, jit, seed=5))
    assert not np.array_equal(r, O.pos_init_jittered(n, dim, L, jit, seed=6))
    np.testing.assert_array_equal(O.pos_init_jittered(n, dim, L, 0.0), O.pos_init_cartesian(n, dim, L))
    # past half a cell the lattice leaves the box unless wrapped back into [0, L)
    rw = O.pos_init_jittered(n, dim, L, 0.75, seed=5)
    assert np.all(rw >= 0) and np.all(rw < L)
