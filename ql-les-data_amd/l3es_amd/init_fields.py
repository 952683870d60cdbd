"""Initial velocity fields and the HIT forcing mask (mirror of l3es/init_fields.py).

The random-phase construction runs in NumPy on the host exactly as the
reference does (same legacy MT19937 stream, same per-mode arithmetic), so the
phases match the reference bit for bit; the inverse transform to physical
space runs on the GPU through the library's own FFT.
"""

import numpy as np
import torch

from . import _lib

EPS = float(np.finfo(float).eps)


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def init_u_tgv(x, y, z):
    """3-D Taylor-Green vortex on the node grid (init_fields.py:12-16)."""
    return np.array([np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), x * 0])


def init_u_tgv2d(Nx, target_dim=2, rescale=1.0):
    """2-D TGV on the cell-centred 'xy' meshgrid, scaled by L (init_fields.py:19-30)."""
    L = rescale
    dx = L / Nx
    x_axis = np.linspace(dx / 2, L - dx / 2, Nx)
    x, y = np.meshgrid(x_axis, x_axis)
    u = np.array([-np.cos(x) * np.sin(y), np.sin(x) * np.cos(y)])
    if target_dim == 3:
        u = u[..., None]
        u = np.concatenate([u, np.zeros((1, Nx, Nx, 1))], axis=0)
    return u * rescale


def spectrum_table(N, spectrum="k53", k0=4.0):
    """Per-mode energy E(k)/(4 pi k^2) on integer shells (init_fields.py:37-41).

    ``spectrum="k4exp"`` selects E(k) = k^4 exp(-2 (k/k0)^2) (forced-HIT runs).
    """
    nmax = round(np.sqrt(3 * N ** 2) / 2)
    E_k = np.zeros(nmax + 1) + EPS
    k = np.arange(1, nmax + 1, dtype=float)
    if spectrum == "k53":
        E_k[1:] = k ** (-5 / 3) / (4 * np.pi * k ** 2)
    elif spectrum == "k4exp":
        E_k[1:] = k ** 4 * np.exp(-2.0 * (k / k0) ** 2) / (4 * np.pi * k ** 2)
    else:
        raise ValueError(f"unknown spectrum {spectrum!r}")
    return nmax, E_k


def rogallo_spectrum(N, seed=42, spectrum="k53", k0=4.0):
    """Random-phase solenoidal spectrum with the x axis halved (init_fields.py:33-102).

    Returns the (N/2+1, N, N, 3) complex array the reference hands to irfftn.
    """
    np.random.seed(seed)
    nmax, E_k = spectrum_table(N, spectrum, k0)
    k1 = np.arange(N // 2 + 1)
    k2 = np.arange(-N // 2, N // 2)
    k3 = np.arange(-N // 2, N // 2)
    K1, K2, K3 = np.meshgrid(k1, k2, k3, indexing="ij")
    WK = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)
    WK12 = np.sqrt(K1 ** 2 + K2 ** 2)
    NK = np.clip(np.round(WK).astype(int), 0, nmax)
    AMP = np.sqrt(E_k[NK])
    phis = 2 * np.pi * np.random.random((N // 2 + 1, N, N, 3))
    ai = AMP * np.exp(1j * phis[..., 0]) * np.cos(phis[..., 2])
    bi = AMP * np.exp(1j * phis[..., 1]) * np.sin(phis[..., 2])
    del phis
    WK_safe = np.where(WK == 0, 1.0, WK)
    WK12_safe = np.where(WK12 == 0, 1.0, WK12)
    u = np.zeros((N // 2 + 1, N, N, 3), dtype=np.complex128)
    u[..., 0] = (ai * WK * K2 + bi * K1 * K3) / (WK_safe * WK12_safe)
    u[..., 1] = (bi * K2 * K3 - ai * WK * K1) / (WK_safe * WK12_safe)
    u[..., 2] = -bi * WK12 / WK_safe
    m12 = WK12 == 0
    u[..., 0] = np.where(m12, ai, u[..., 0])
    u[..., 1] = np.where(m12, bi, u[..., 1])
    m0 = WK == 0
    for c in range(3):
        u[..., c] = np.where(m0, 0.0, u[..., c])
    j_inv = np.argsort((N + k2) % N)
    k_inv = np.argsort((N + k3) % N)
    u = u[:, j_inv, :][:, :, k_inv]
    f = N // 2
    u[0, -f + 1:, -f + 1:, :] = np.flip(np.conj(u[0, 1:f, 1:f, :]), axis=(0, 1))
    u[0, 1:f, -f + 1:, :] = np.flip(np.conj(u[0, -f + 1:, 1:f, :]), axis=(0, 1))
    u[0, 0, -f + 1:, :] = np.flip(np.conj(u[0, 0, 1:f, :]), axis=0)
    u[0, -f + 1:, 0, :] = np.flip(np.conj(u[0, 1:f, 0, :]), axis=0)
    return u


def _neg(a, axis):
    """a[(-i) mod n] along axis."""
    return np.roll(np.flip(a, axis=axis), 1, axis=axis)


def x_halved_to_z_halved(U, N):
    """Re-express an x-halved spectrum (as irfftn(axes=(-1,-2,-3)) reads it) in the
    solver's z-halved rfft layout (3, N, N, N/2+1).

    irfftn's c2r along x ignores the imaginary part of the kx = 0 and kx = N/2
    planes after the y/z transforms, i.e. it uses their (ky, kz) Hermitian
    projection; the full spectrum F built here is exactly Hermitian, so any
    real inverse of it reproduces the reference field.
    """
    Ut = np.ascontiguousarray(U.transpose(3, 0, 1, 2))  # (3, Nh, N, N): [c][kx][ky][kz]
    Nh = N // 2 + 1
    F = np.empty((3, N, N, N), dtype=np.complex128)
    negyz = np.conj(_neg(_neg(Ut, 2), 3))  # conj(U[kx, -ky, -kz])
    F[:, 1:N // 2] = Ut[:, 1:N // 2]
    F[:, 0] = 0.5 * (Ut[:, 0] + negyz[:, 0])
    F[:, N // 2] = 0.5 * (Ut[:, N // 2] + negyz[:, N // 2])
    # kx in (N/2, N): F(kx) = conj(U(N - kx, -ky, -kz))
    F[:, N // 2 + 1:] = negyz[:, 1:N // 2][:, ::-1]
    return np.ascontiguousarray(F[..., :Nh])


def init_u_hit(N, seed=42, spectrum="k53", k0=4.0, dtype=torch.float64):
    """Rogallo HIT field (init_fields.py:33-106) as a (3, N, N, N) device tensor.

    The reference ends with irfftn(..., norm="forward"), the unnormalized
    inverse; the same real field is produced here by the GPU inverse of the
    z-halved spectrum times N^3.
    """
    V = x_halved_to_z_halved(rogallo_spectrum(N, seed, spectrum, k0), N) * float(N) ** 3
    plan = _lib.Plan(3, N, torch.float64, 0.0, 1.0)
    vh = torch.from_numpy(V).to(_device())
    u = torch.empty((3, N, N, N), dtype=torch.float64, device=_device())
    _lib.check(_lib.load().qlles_irfftn(plan.handle, _lib.ptr(vh), _lib.ptr(u), _lib.stream_ptr()))
    return u.to(dtype)


def init_forcing_mask_hit(N, kf=3):
    """0 < |k|^2 <= kf^2 on the rfft half grid, Nyquist flipped (init_fields.py:167-188)."""
    k1 = np.fft.fftfreq(N, 1.0 / N)
    k3 = k1[: N // 2 + 1].copy()
    k3[-1] = -k3[-1]
    kk = np.asarray(np.meshgrid(k1, k1, k3, indexing="ij"))
    k_mag = np.sum(kk ** 2, axis=0)
    return (k_mag <= kf ** 2) & (k_mag > 0)


def kolm_vorticity_spectrum(N, seed=42, k0=4.0):
    """Seeded random velocity spectrum for 2-D Kolmogorov runs.

    The reference's init_u_kolm (init_fields.py:191-212) draws from jax_cfd with
    the JAX PRNG, which is not reproducible here; this project-defined
    generator shapes a Gaussian white-noise vorticity so that its shell
    enstrophy is Z(k) ~ k^4 exp(-2 (k/k0)^2) (peak at k0; the energy
    spectrum Z/k^2 peaks at k0/sqrt(2)) and inverts the stream function.
    Returns (u_hat, v_hat) in rfft2 layout (N, N/2+1).
    """
    rng = np.random.default_rng(seed)
    xi_hat = np.fft.rfftn(rng.standard_normal((N, N)))
    kx = np.fft.fftfreq(N, 1.0 / N)
    ky = np.abs(kx[: N // 2 + 1])
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    k2 = KX ** 2 + KY ** 2
    kmag = np.sqrt(k2)
    amp = np.where(k2 > 0, np.sqrt(kmag ** 4 * np.exp(-2 * (kmag / k0) ** 2) / np.maximum(kmag, 1.0)), 0.0)
    psi_hat = xi_hat * amp / np.where(k2 == 0, 1.0, k2)
    return 1j * KY * psi_hat, -1j * KX * psi_hat


def init_u_kolm(N, L=2 * np.pi, max_velocity=4.2, seed=42, target_dim=3, iter=4, dtype=torch.float64):
    """Seeded divergence-free 2-D field with max |u| = max_velocity, (3, N, N, 1)."""
    del L, iter  # kept for signature parity with the reference
    uh, vh = kolm_vorticity_spectrum(N, seed)
    spec = np.zeros((3, N, N // 2 + 1), dtype=np.complex128)
    spec[0], spec[1] = uh, vh
    plan = _lib.Plan(2, N, torch.float64, 0.0, 1.0)
    sh = torch.from_numpy(spec).to(_device())
    u = torch.empty((3, N, N, 1), dtype=torch.float64, device=_device())
    _lib.check(_lib.load().qlles_irfftn(plan.handle, _lib.ptr(sh), _lib.ptr(u), _lib.stream_ptr()))
    s = max_velocity / float(torch.abs(u[:2]).max())
    u = u * s
    if target_dim != 3:
        u = u[:2]
    return u.to(dtype)
