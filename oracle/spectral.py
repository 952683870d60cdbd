"""NumPy FP64 restatement of the reference spectral solver (test infrastructure).

Only tests, the smoke check and the CPU-baseline leg of the benchmark use this
module.  Every function cites the reference file:line it restates
(paths relative to the tumaer/ql-les-data checkout).
"""

import numpy as np

EPS = np.finfo(float).eps  # spectral_solver.py:23

__all__ = [
    "EPS", "fft_backend", "wavenumbers", "rhs_wrapper", "rk4_wrapper",
    "init_u_tgv", "init_u_tgv2d", "init_u_hit", "init_u_hit_slow",
    "init_forcing_mask_hit", "init_u_kolm_np", "kolmogorov_forcing_hat",
    "set_up_solver", "spectral_filtering", "interpolate_dft", "shift_fn",
    "comp_dt", "comp_dissipation_rate", "energy_spectrum", "kinetic_energy",
    "kinetic_energy_hat", "pos_init_cartesian", "pos_init_jittered",
    "init_u_hit_spectrum", "make_forcing", "init_forcing_mask_2d",
]


class _FFT:
    """FFT provider: numpy's pocketfft by default, scipy.fft with workers for timing."""

    def __init__(self, workers=None):
        self.workers = workers
        if workers is None:
            self.rfftn, self.irfftn = np.fft.rfftn, np.fft.irfftn
            self.fftn, self.ifftn = np.fft.fftn, np.fft.ifftn
        else:
            import scipy.fft as sf

            w = workers
            self.rfftn = lambda a, axes=None: sf.rfftn(a, axes=axes, workers=w)
            self.irfftn = lambda a, axes=None: sf.irfftn(a, axes=axes, workers=w)
            self.fftn = lambda a, axes=None: sf.fftn(a, axes=axes, workers=w)
            self.ifftn = lambda a, axes=None: sf.ifftn(a, axes=axes, workers=w)


_FFT_DEFAULT = _FFT()


def fft_backend(workers=None):
    return _FFT_DEFAULT if workers is None else _FFT(workers)


# --------------------------------------------------------------------------
# Wavenumbers and RHS: spectral_solver.py:27-62
# --------------------------------------------------------------------------
def wavenumbers(N, dim):
    """k grid, k^2, k/k^2 and the 2/3 dealias mask (spectral_solver.py:32-48).

    The rfft axis keeps the Nyquist entry with its sign flipped to +N/2
    (spectral_solver.py:34-36); 2-D pads a zero third component (:38-39).
    """
    kx = np.fft.fftfreq(N, 1.0 / N)
    kz = kx[: N // 2 + 1].copy()
    kz[-1] = -kz[-1]
    if dim == 3:
        kkk = np.array(np.meshgrid(kx, kx, kz, indexing="ij"), dtype=int)
    else:
        kkk = np.array(np.meshgrid(kx, kz, indexing="ij"), dtype=int)
        kkk = np.concatenate([kkk, np.zeros((1, N, N // 2 + 1), dtype=int)], axis=0)
    kkk2 = np.sum(kkk * kkk, 0, dtype=int)
    k_over_k2 = kkk.astype(float) / np.where(kkk2 == 0, 1, kkk2).astype(float)
    kmax = 2.0 / 3.0 * (N // 2 + 1)
    dealias = (abs(kkk[0]) < kmax) * (abs(kkk[1]) < kmax) * (abs(kkk[2]) < kmax)
    return kkk, kkk2, k_over_k2, dealias.astype(bool)


def rhs_wrapper(N, nu, fft_axes=(1, 2, 3), forcing=None, fft=None, viscous=True):
    """Rotational-form RHS (spectral_solver.py:27-62).

    ``forcing`` (modes this library adds, no reference counterpart) is a callable
    ``f(u_hat) -> f_hat`` added after the projection; ``viscous=False`` drops
    the explicit -nu k^2 u_hat term (used by the integrating-factor scheme).
    """
    fft = fft or _FFT_DEFAULT
    dim = len(fft_axes)
    kkk, kkk2, k_over_k2, dealias = wavenumbers(N, dim)

    def rhs_fn(u, u_hat, rk):
        if rk > 0:  # :51-52
            u = fft.irfftn(u_hat, axes=fft_axes)
        curl = fft.irfftn(1j * np.cross(kkk, u_hat, axis=0), axes=fft_axes)  # :54
        du = fft.rfftn(np.cross(np.squeeze(u), curl, axis=0), axes=fft_axes)  # :55
        du *= dealias  # :56
        p_hat = np.sum(du * k_over_k2, axis=0)  # :57
        du -= p_hat * kkk  # :58
        if viscous:
            du -= nu * kkk2 * u_hat  # :59
        if forcing is not None:
            du += forcing(u_hat)
        return du

    rhs_fn.N, rhs_fn.nu, rhs_fn.fft_axes = N, nu, fft_axes
    rhs_fn.kkk, rhs_fn.kkk2 = kkk, kkk2
    return rhs_fn


def _rfft_weights(N, dim):
    """Parseval weights on the rfft half axis: 1 at index 0 and N/2, else 2."""
    w = np.full(N // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return w  # broadcast against the last axis


def kinetic_energy(u):
    """0.5 <|u|^2> (spectral_solver.py:94, :318)."""
    return 0.5 * float(np.mean(np.sum(u * u, axis=0)))


def kinetic_energy_hat(u_hat, N, dim, mask=None):
    """0.5 <|u|^2> from the rfft spectrum by Parseval (Hermitian weights)."""
    w = _rfft_weights(N, dim)
    a = np.abs(u_hat) ** 2
    if mask is not None:
        a = a * mask
    return 0.5 * float(np.sum(a * w)) / float(N ** dim) ** 2


def rk4_wrapper(dt, rhs_fn, forcing_mask=None, e_kin_init=None, forcing_type="none",
                fft_axes=(1, 2, 3), fft=None, viscous_mode="explicit", nl_fn=None):
    """Classical RK4 (spectral_solver.py:65-120) with the ekin_tot rescaling
    forcing (:82-112).

    ``viscous_mode="IF"`` (an addition of this library, no reference counterpart) integrates
    the viscous term exactly with the factor exp(-nu k^2 h); then ``nl_fn`` is
    an RHS without the viscous term.
    """
    fft = fft or _FFT_DEFAULT
    a_rk4 = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]  # :70
    b_rk4 = [0.5, 0.5, 1.0]  # :71
    dim = len(fft_axes)

    def forcing_step(u_hat):
        e_inj = 0.0
        if forcing_mask is not None and forcing_type == "ekin_tot":
            u_new = fft.irfftn(u_hat, axes=fft_axes)  # :93
            e_kin_new = 0.5 * np.mean(np.sum(u_new * u_new, axis=0))  # :94
            u_lower = fft.irfftn(u_hat * forcing_mask, axes=fft_axes)  # :97
            e_kin_lower = 0.5 * np.mean(np.sum(u_lower * u_lower, axis=0))  # :98
            e_kin_high = e_kin_new - e_kin_lower  # :101
            alpha = np.sqrt(max(0.0, (e_kin_init - e_kin_high) / (e_kin_lower + EPS)))  # :104-106
            e_inj = max(0.0, (e_kin_init - e_kin_new) / dt)  # :109
            u_hat = u_hat * (alpha * forcing_mask + (1 - forcing_mask))  # :112
        return u_hat, e_inj

    def finish(u_hat):
        u_hat, e_inj = forcing_step(u_hat)
        u = fft.irfftn(u_hat, axes=fft_axes)  # :114
        if dim == 2:
            u = u[..., None]  # :115-116
        return u, u_hat, e_inj

    if viscous_mode == "explicit":

        def step_fn(u, u_hat):
            u_temp1, u_temp2 = u_hat.copy(), u_hat.copy()  # :74
            for rk in range(4):  # :75
                du = rhs_fn(u, u_hat, rk)
                if rk < 3:
                    u_hat = u_temp1 + b_rk4[rk] * dt * du  # :77-78
                u_temp2 += a_rk4[rk] * dt * du  # :79
            return finish(u_temp2)  # :80

    elif viscous_mode == "IF":
        kkk2 = rhs_fn.kkk2
        nu = rhs_fn.nu
        e_half = np.exp(-nu * kkk2 * (0.5 * dt))
        e_full = np.exp(-nu * kkk2 * dt)
        nl = nl_fn

        def step_fn(u, u_hat):
            un = u_hat
            n0 = nl(u, un, 0)
            acc = e_full * un + (dt / 6.0) * e_full * n0
            u1 = e_half * (un + 0.5 * dt * n0)
            n1 = nl(None, u1, 1)
            acc = acc + (dt / 3.0) * e_half * n1
            u2 = e_half * un + 0.5 * dt * n1
            n2 = nl(None, u2, 2)
            acc = acc + (dt / 3.0) * e_half * n2
            u3 = e_full * un + dt * e_half * n2
            n3 = nl(None, u3, 3)
            return finish(acc + (dt / 6.0) * n3)

    else:
        raise ValueError(viscous_mode)
    return step_fn


# --------------------------------------------------------------------------
# Initial conditions: init_fields.py
# --------------------------------------------------------------------------
def init_u_tgv(x, y, z):
    """3-D Taylor-Green vortex (init_fields.py:12-16)."""
    return np.array([np.sin(x) * np.cos(y) * np.cos(z),
                     -np.cos(x) * np.sin(y) * np.cos(z), x * 0])


def init_u_tgv2d(Nx, target_dim=2, rescale=1.0):
    """2-D TGV on the cell-centred 'xy' meshgrid (init_fields.py:19-30)."""
    L = rescale
    dx = L / Nx
    x_axis = np.linspace(dx / 2, L - dx / 2, Nx)
    x, y = np.meshgrid(x_axis, x_axis)
    u = np.array([-np.cos(x) * np.sin(y), np.sin(x) * np.cos(y)])
    if target_dim == 3:
        u = u[..., None]
        u = np.concatenate([u, np.zeros((1, Nx, Nx, 1))], axis=0)
    return u * rescale


def _spectrum_table(N, spectrum="k53", k0=4.0):
    """Per-mode energy E(k)/(4 pi k^2) on integer shells (init_fields.py:37-41).

    ``spectrum="k4exp"`` is the added E(k) = k^4 exp(-2 (k/k0)^2).
    """
    nmax = round(np.sqrt(3 * N ** 2) / 2)
    E_k = np.zeros(nmax + 1) + EPS
    k = np.arange(1, nmax + 1, dtype=float)
    if spectrum == "k53":
        E_k[1:] = k ** (-5 / 3) / (4 * np.pi * k ** 2)
    elif spectrum == "k4exp":
        E_k[1:] = k ** 4 * np.exp(-2.0 * (k / k0) ** 2) / (4 * np.pi * k ** 2)
    else:
        raise ValueError(spectrum)
    return nmax, E_k


def init_u_hit(N, seed=42, spectrum="k53", k0=4.0):
    """Vectorized Rogallo init (init_fields.py:33-106)."""
    np.random.seed(seed)  # :35
    nmax, E_k = _spectrum_table(N, spectrum, k0)
    k1_loop = np.arange(N // 2 + 1)
    k2_loop = np.arange(-N // 2, N // 2)
    k3_loop = np.arange(-N // 2, N // 2)
    K1, K2, K3 = np.meshgrid(k1_loop, k2_loop, k3_loop, indexing="ij")  # :44-48
    WK = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)
    WK12 = np.sqrt(K1 ** 2 + K2 ** 2)
    NK = np.clip(np.round(WK).astype(int), 0, nmax)  # :52
    AMP = np.sqrt(E_k[NK])
    phis = 2 * np.pi * np.random.random((N // 2 + 1, N, N, 3))  # :56-57
    phi1, phi2, phi3 = phis[..., 0], phis[..., 1], phis[..., 2]
    ai = AMP * np.exp(1j * phi1) * np.cos(phi3)  # :60
    bi = AMP * np.exp(1j * phi2) * np.sin(phi3)  # :61
    WK_safe = np.where(WK == 0, 1.0, WK)
    WK12_safe = np.where(WK12 == 0, 1.0, WK12)
    u_loop = np.zeros((N // 2 + 1, N, N, 3), dtype=np.complex128)
    u_loop[..., 0] = (ai * WK * K2 + bi * K1 * K3) / (WK_safe * WK12_safe)  # :71-83
    u_loop[..., 1] = (bi * K2 * K3 - ai * WK * K1) / (WK_safe * WK12_safe)
    u_loop[..., 2] = -bi * WK12 / WK_safe
    m12 = WK12 == 0
    u_loop[..., 0] = np.where(m12, ai, u_loop[..., 0])
    u_loop[..., 1] = np.where(m12, bi, u_loop[..., 1])
    m0 = WK == 0
    for c in range(3):
        u_loop[..., c] = np.where(m0, 0.0, u_loop[..., c])
    j_idx = (N + k2_loop) % N  # :86-91
    k_idx = (N + k3_loop) % N
    u = u_loop[:, np.argsort(j_idx), :][:, :, np.argsort(k_idx)]
    Nf_ = N // 2  # :94-102
    u[0, -Nf_ + 1:, -Nf_ + 1:, :] = np.flip(np.conj(u[0, 1:Nf_, 1:Nf_, :]), axis=(0, 1))
    u[0, 1:Nf_, -Nf_ + 1:, :] = np.flip(np.conj(u[0, -Nf_ + 1:, 1:Nf_, :]), axis=(0, 1))
    u[0, 0, -Nf_ + 1:, :] = np.flip(np.conj(u[0, 0, 1:Nf_, :]), axis=0)
    u[0, -Nf_ + 1:, 0, :] = np.flip(np.conj(u[0, 1:Nf_, 0, :]), axis=0)
    # :105 irfftn over axes (-1,-2,-3): the halved axis is x, unnormalized inverse
    return np.fft.irfftn(u.transpose(3, 0, 1, 2), axes=(-1, -2, -3), norm="forward")


def init_u_hit_spectrum(N, seed=42, spectrum="k53", k0=4.0):
    """The x-halved spectrum that init_u_hit feeds to irfftn (for tests)."""
    np.random.seed(seed)
    nmax, E_k = _spectrum_table(N, spectrum, k0)
    K1, K2, K3 = np.meshgrid(np.arange(N // 2 + 1), np.arange(-N // 2, N // 2),
                             np.arange(-N // 2, N // 2), indexing="ij")
    return K1, K2, K3, nmax, E_k


def init_u_hit_slow(N, seed=42, spectrum="k53", k0=4.0):
    """Scalar-loop Rogallo twin (init_fields.py:109-164); brute force for tests."""
    np.random.seed(seed)

    def sym2(A):  # :112-119
        n = A.shape[0]
        f = n // 2
        A[-f + 1:, -f + 1:] = np.flip(np.conj(A[1:f, 1:f]), axis=(0, 1))
        A[1:f, -f + 1:] = np.flip(np.conj(A[-f + 1:, 1:f]), axis=(0, 1))
        A[0, -f + 1:] = np.flip(np.conj(A[0, 1:f]))
        A[-f + 1:, 0] = np.flip(np.conj(A[1:f, 0]))
        return A

    u = np.zeros((N // 2 + 1, N, N, 3), dtype=np.complex128)
    nmax, E_k = _spectrum_table(N, spectrum, k0)
    for k1 in range(N // 2 + 1):
        for k2 in range(-N // 2, N // 2):
            j = (N + k2) % N
            for k3 in range(-N // 2, N // 2):
                k = (N + k3) % N
                wk = np.sqrt(k1 ** 2 + k2 ** 2 + k3 ** 2)
                wk12 = np.sqrt(k1 ** 2 + k2 ** 2)
                nk = round(wk)
                amp = np.sqrt(E_k[nk])
                phi1, phi2, phi3 = 2 * np.pi * np.random.random(size=3)
                ai = amp * (np.cos(phi1) + np.sin(phi1) * 1j) * np.cos(phi3)
                bi = amp * (np.cos(phi2) + np.sin(phi2) * 1j) * np.sin(phi3)
                if wk != 0:
                    if wk12 != 0.0:
                        u[k1, j, k, 0] = (ai * wk * k2 + bi * k1 * k3) / (wk * wk12)
                        u[k1, j, k, 1] = (bi * k2 * k3 - ai * wk * k1) / (wk * wk12)
                    else:
                        u[k1, j, k, 0] = ai
                        u[k1, j, k, 1] = bi
                    u[k1, j, k, 2] = -bi * wk12 / wk
    for ii in range(3):
        u[0, :, :, ii] = sym2(u[0, :, :, ii])
    return np.fft.irfftn(u.transpose(3, 0, 1, 2), axes=(-1, -2, -3), norm="forward")


def init_forcing_mask_hit(N, kf=3):
    """0 < |k|^2 <= kf^2 on the rfft half grid, Nyquist flipped (init_fields.py:167-188)."""
    k1 = np.fft.fftfreq(N, 1.0 / N)
    k3 = k1[: N // 2 + 1].copy()
    k3[-1] = -k3[-1]
    kk = np.asarray(np.meshgrid(k1, k1, k3, indexing="ij"))
    k_mag = np.sum(kk ** 2, axis=0)
    return (k_mag <= kf ** 2) & (k_mag > 0)


def init_forcing_mask_2d(N, kf):
    """2-D analogue of the HIT shell mask (helper of an added mode)."""
    k1 = np.fft.fftfreq(N, 1.0 / N)
    k2 = k1[: N // 2 + 1].copy()
    k2[-1] = -k2[-1]
    kk = np.asarray(np.meshgrid(k1, k2, indexing="ij"))
    k_mag = np.sum(kk ** 2, axis=0)
    return (k_mag <= kf ** 2) & (k_mag > 0)


def init_u_kolm_np(N, seed=42, k0=4.0, max_velocity=4.2):
    """Project-defined seeded 2-D random field (the 256^2 Kolmogorov configuration).

    The reference's init_u_kolm (init_fields.py:191-212) uses jax_cfd and the
    JAX PRNG, which cannot be reproduced here; this generator instead draws a
    Gaussian white-noise vorticity and shapes it so that the shell enstrophy
    (the vorticity spectrum) is Z(k) ~ k^4 exp(-2 (k/k0)^2), peaking at k0;
    the energy spectrum Z(k)/k^2 then peaks at k0/sqrt(2).  It inverts the
    stream function and scales max(|u|, |v|) to max_velocity.  Returns
    (3, N, N, 1).
    """
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((N, N))
    xi_hat = np.fft.rfftn(xi)
    kx = np.fft.fftfreq(N, 1.0 / N)
    ky = np.abs(kx[: N // 2 + 1])
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    k2 = KX ** 2 + KY ** 2
    kmag = np.sqrt(k2)
    amp = np.where(k2 > 0, np.sqrt(kmag ** 4 * np.exp(-2 * (kmag / k0) ** 2) / np.maximum(kmag, 1.0)), 0.0)
    w_hat = xi_hat * amp
    psi_hat = w_hat / np.where(k2 == 0, 1.0, k2)
    u_hat = 1j * KY * psi_hat
    v_hat = -1j * KX * psi_hat
    u = np.fft.irfftn(u_hat, s=(N, N), axes=(0, 1))
    v = np.fft.irfftn(v_hat, s=(N, N), axes=(0, 1))
    s = max_velocity / max(np.abs(u).max(), np.abs(v).max())
    out = np.zeros((3, N, N, 1))
    out[0, :, :, 0] = u * s
    out[1, :, :, 0] = v * s
    return out


def kolmogorov_forcing_hat(N, amp=1.0, kf=4):
    """rfft2 of f_x = amp sin(kf y) on the node grid (the 256^2 Kolmogorov configuration)."""
    y = np.arange(N) * 2 * np.pi / N
    f = np.zeros((3, N, N))
    f[0] = amp * np.sin(kf * y)[None, :]
    return np.fft.rfftn(f, axes=(1, 2))


def make_forcing(kind, N, dim, u_hat_shape=None, kf=2, power=0.1, kolm_amp=1.0,
                 kolm_k=4, drag=0.0):
    """Added RHS forcing terms (no reference counterpart).

    ``power``: f_hat = P / (2 E_f) u_hat on 0 < |k| <= kf, with E_f the
    Parseval energy of those modes, so <f.u> = P exactly.
    ``kolmogorov``: f_x = amp sin(kolm_k y) plus linear drag -drag * u.
    """
    if kind == "power":
        mask = init_forcing_mask_hit(N, kf) if dim == 3 else init_forcing_mask_2d(N, kf)

        def f(u_hat):
            e_f = kinetic_energy_hat(u_hat, N, dim, mask)
            return (power / (2.0 * e_f + EPS)) * mask * u_hat

        f.mask = mask
        return f
    if kind == "kolmogorov":
        fh = kolmogorov_forcing_hat(N, kolm_amp, kolm_k)

        def f(u_hat):
            return fh - drag * u_hat

        return f
    raise ValueError(kind)


def set_up_solver(N, nu, dim, case, dt, seed=42, kf=None, e_kin_target=None,
                  forcing_type="none", fft=None, viscous_mode="explicit",
                  spectrum="k53", power=0.1, drag=0.0, kolm_amp=1.0, kolm_k=4,
                  warmup=True, u_init=None):
    """set_up_solver restated (spectral_solver.py:156-230).

    Returns (u, u_hat, step_fn, e_kin_init, e_inj); the reference's warm-up
    step (:226) is taken when ``warmup`` is true.
    """
    fft = fft or _FFT_DEFAULT
    L = 2 * np.pi
    len_z = N if dim == 3 else 1
    fft_axes = (1, 2, 3) if dim == 3 else (1, 2)
    xyz = np.array(np.meshgrid(np.arange(N), np.arange(N), np.arange(len_z), indexing="ij")) * L / N
    forcing_mask = None
    if case == "HIT" and forcing_type == "ekin_tot":
        forcing_mask = init_forcing_mask_hit(N, kf)
    if u_init is not None:
        u = np.array(u_init, dtype=float)
    elif case == "TGV3D":
        u = init_u_tgv(xyz[0], xyz[1], xyz[2])
    elif case == "TGV2D":
        u = init_u_tgv2d(N, target_dim=3, rescale=L)
    elif case == "Kolm":
        u = init_u_kolm_np(N, seed=seed)
    elif case == "HIT":
        u = init_u_hit(N, seed, spectrum=spectrum)
    else:
        raise ValueError(case)
    e_kin_init = 0.5 * np.mean(np.sum(u * u, axis=0))  # :200
    if forcing_type != "none" and e_kin_target is not None:  # :201-203
        u = u * np.sqrt(e_kin_target / (e_kin_init + EPS))
    u_hat = np.squeeze(fft.rfftn(u, axes=fft_axes))  # :206
    e_kin_init = 0.5 * np.mean(np.sum(u * u, axis=0))  # :209
    force = None
    if forcing_type == "power":
        force = make_forcing("power", N, dim, kf=kf, power=power)
    elif forcing_type == "kolmogorov":
        force = make_forcing("kolmogorov", N, dim, drag=drag, kolm_amp=kolm_amp, kolm_k=kolm_k)
    if viscous_mode == "explicit":
        rhs = rhs_wrapper(N, nu, fft_axes, forcing=force, fft=fft)
        step = rk4_wrapper(dt, rhs, forcing_mask, e_kin_init, forcing_type, fft_axes, fft=fft)
    else:
        rhs = rhs_wrapper(N, nu, fft_axes, forcing=force, fft=fft)
        nl = rhs_wrapper(N, nu, fft_axes, forcing=force, fft=fft, viscous=False)
        step = rk4_wrapper(dt, rhs, forcing_mask, e_kin_init, forcing_type, fft_axes,
                           fft=fft, viscous_mode="IF", nl_fn=nl)
    e_inj = 0.0
    if warmup:
        u, u_hat, e_inj = step(u, u_hat)  # :226
    return u, u_hat, step, e_kin_init, e_inj


# --------------------------------------------------------------------------
# Filtering, interpolation, advection, diagnostics
# --------------------------------------------------------------------------
def spectral_filtering(u, ckp_N):
    """fftn -> fftshift -> crop -> ifftshift -> ifftn/(N/n)^d -> .real (utils.py:100-138)."""
    N_u = u.shape[1]
    dim = len(u)
    fft_axes = (1, 2, 3) if dim == 3 else (1, 2)
    y = np.fft.fftn(u, axes=fft_axes)
    sl = slice(N_u // 2 - ckp_N // 2, N_u // 2 + ckp_N // 2)
    y = np.fft.fftshift(y, axes=fft_axes)[(slice(None),) + (sl,) * dim]
    y = np.fft.ifftshift(y, axes=fft_axes)
    return (np.fft.ifftn(y, axes=fft_axes) / (N_u / ckp_N) ** dim).real


def interpolate_dft(u, r, chunk=2048):
    """Exact direct sum Re mean_k(u_hat_k e^{i k.x}) (integrator.py:62-91)."""
    dim = len(u)
    n = u.shape[1]
    fft_axes = tuple(range(1, dim + 1))
    u_hat = np.fft.fftshift(np.fft.fftn(u, axes=fft_axes), axes=fft_axes)
    k = np.fft.fftshift(np.fft.fftfreq(n, 1.0 / n))
    out = np.zeros((len(r), dim))
    for s in range(0, len(r), chunk):
        rr = np.asarray(r[s:s + chunk], dtype=float)
        ex = [np.exp(1j * np.outer(rr[:, d], k)) for d in range(dim)]  # (P, n)
        if dim == 3:
            t = np.einsum("cijk,pk->cpij", u_hat, ex[2])
            t = np.einsum("cpij,pj->cpi", t, ex[1])
            t = np.einsum("cpi,pi->cp", t, ex[0])
        else:
            t = np.einsum("cij,pj->cpi", u_hat, ex[1])
            t = np.einsum("cpi,pi->cp", t, ex[0])
        out[s:s + chunk] = (t.real / n ** dim).T
    return out


def shift_fn(r, dr, box_size=2 * np.pi):
    """Periodic shift (integrator.py:21-23)."""
    return (r + dr) % box_size


def comp_dt(u, dx, nu, cfl=1.0):
    """CFL estimate (spectral_solver.py:123-127)."""
    umax = np.max(np.abs(u), axis=(1, 2, 3))
    return cfl / (umax / dx + nu / dx ** 2).min()


def comp_dissipation_rate(u_hat, nu):
    """nu <|curl u|^2> through irfftn (utils.py:300-315)."""
    n = u_hat.shape[1]
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = kx[: n // 2 + 1].copy()
    kz[-1] = -kz[-1]
    kkk = np.array(np.meshgrid(kx, kx, kz, indexing="ij"), dtype=int)
    curl = np.fft.irfftn(1j * np.cross(kkk, u_hat, axis=0), axes=(1, 2, 3))
    return nu * np.mean(np.sum(curl ** 2, axis=0))


def energy_spectrum(vel, mul_fac=1.0, dim=3):
    """Shell-binned E(k) (utils.py:35-97); the halved axis is x (axes=(3,2,1))."""
    if dim == 2:
        vel = vel[:2, :, :, 0]
    n = vel.shape[1]
    if dim == 2:
        vel_hat = np.fft.rfftn(vel, axes=(2, 1))
    else:
        vel_hat = np.fft.rfftn(vel, axes=(3, 2, 1))
    Nf = n // 2 + 1
    k = np.fft.fftfreq(n, 1.0 / n)
    kx = k[:Nf].copy()
    kx[-1] *= -1
    if dim == 2:
        k_field = np.array(np.meshgrid(kx, k, indexing="ij"), dtype=int)
    else:
        k_field = np.array(np.meshgrid(kx, k, k, indexing="ij"), dtype=int)
    fact = (2 * (k_field[0] > 0) * (k_field[0] < n // 2) + 1 * (k_field[0] == 0)
            + 1 * (k_field[0] == n // 2))
    knorm = np.linalg.norm(k_field, axis=0, ord=2)
    shell = (knorm + 0.5).astype(int).flatten()
    vel_hat = vel_hat / n ** dim
    abs_energy = np.sum(np.abs(vel_hat ** 2), axis=0) * fact * mul_fac
    n_samples = np.zeros(n)
    np.add.at(n_samples, shell, fact.flatten())
    ek = np.zeros(n)
    np.add.at(ek, shell, abs_energy.flatten())
    ek *= 4 * np.pi * np.arange(n) ** 2 / (n_samples + EPS)
    return ek


def pos_init_cartesian(n, dim, L=2 * np.pi):
    """Cell-centred lattice (i + 1/2) dx in 'ij' order (stand-in for jax_sph's
    pos_init_cartesian_*; its exact ordering is parity unpinned)."""
    dx = L / n
    g = np.meshgrid(*([np.arange(n)] * dim), indexing="ij")
    return (np.stack([a.ravel() for a in g], axis=1) + 0.5) * dx


def pos_init_jittered(n, dim, L=2 * np.pi, jitter=0.25, seed=0):
    """Cell-centred lattice plus uniform jitter of +-jitter*dx (the 256^2 Kolmogorov configuration)."""
    rng = np.random.default_rng(seed)
    r = pos_init_cartesian(n, dim, L)
    r = r + (rng.random(r.shape) * 2 - 1) * jitter * (L / n)
    return r % L
