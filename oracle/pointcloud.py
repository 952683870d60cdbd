"""NumPy FP64 restatement of the reference's particle -> grid transfers and
point-cloud spectrum (test infrastructure; turbulence.py:145-377).

Only tests and the CPU-baseline leg of the benchmark use this module.

- mls_2nd_order (turbulence.py:145-250) with the quintic kernel the wrapper
  uses (h = 2/3 dx, support 3h = 2 dx, Euclidean neighbourhood).  The
  reference's pbc_copy_scalar + KDTree (:45-142, :186-193) find, for every
  target, the particles within the support under the periodic minimum image;
  neighbor_pairs below does the same directly.  The reference assembles
  M = sum_j w c r^T with c = D r (D = diag(1, 1.., 2.., 2..)) and
  v = sum_j w f c; M x = v is D A x = D b with A = sum_j w r r^T and
  b = sum_j w f r, and x[0] is the value at the target.
- ur_to_u_dft_wrapper (turbulence.py:298-336): u_hat[k] = sum_j u_j e^{-i k.r_j}
  over the fftfreq modes, then ifftn(...).real on the nodes m dx.
- ur_to_u_mls_wrapper (turbulence.py:253-295) interpolates at the cell-centred
  lattice (i + 1/2) dx, taken here in 'ij' order.

Parity pins (tests/test_oracle_pointcloud.py): MLS reproduces quadratic fields
exactly; the DFT transfer is the identity for particles on the nodes and
follows a one-node shift of the particles (fixes the sign of the exponent).
"""

import numpy as np

from . import sph as _sph

__all__ = ["mls_rows", "mls_2nd_order", "ur_to_u_mls", "ur_to_u_dft", "u_and_spectrum_from_ur"]


def mls_rows(rji):
    """Monomial rows [1, r_a, r_a^2 / 2, r_a r_b (a < b)] (turbulence.py:210-218)."""
    rji = np.atleast_2d(rji)
    dim = rji.shape[1]
    cols = [np.ones(rji.shape[0])] + [rji[:, a] for a in range(dim)]
    cols += [0.5 * rji[:, a] ** 2 for a in range(dim)]
    cols += [rji[:, a] * rji[:, b] for a in range(dim) for b in range(a + 1, dim)]
    return np.stack(cols, axis=1)


def _target_pairs(r, r_target, cutoff, side):
    """Ordered (target i, particle j) pairs with minimum-image distance < cutoff."""
    ii, jj = [], []
    for a in range(0, r_target.shape[0], 2048):
        d = _sph.periodic_displacement(r[None, :, :], r_target[a:a + 2048, None, :], side)
        dist = np.sqrt(np.sum(d * d, axis=-1))
        pi, pj = np.nonzero(dist < cutoff)
        ii.append(pi + a)
        jj.append(pj)
    return np.concatenate(ii), np.concatenate(jj)


def mls_2nd_order(r, r_target, f, box, dx, dim, h_factor=2.0 / 3.0):
    """turbulence.py:145-250 with kernel_name="Quintic"; f is (np,) or (np, ncomp)."""
    r = np.asarray(r, dtype=np.float64)
    r_target = np.asarray(r_target, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    f2 = f[:, None] if f.ndim == 1 else f
    h = h_factor * dx
    i_s, j_s = _target_pairs(r, r_target, 3.0 * h, box)
    rji = _sph.periodic_displacement(r[j_s], r_target[i_s], box)  # turbulence.py:196
    w = _sph.quintic_w(np.linalg.norm(rji, axis=1), h, dim)      # :200-201
    rows = mls_rows(rji)
    m = rows.shape[1]
    nt = r_target.shape[0]
    A = np.zeros((nt, m, m))
    np.add.at(A, i_s, w[:, None, None] * rows[:, :, None] * rows[:, None, :])
    out = np.empty((nt, f2.shape[1]))
    for c in range(f2.shape[1]):
        b = np.zeros((nt, m))
        np.add.at(b, i_s, (w * f2[j_s, c])[:, None] * rows)
        out[:, c] = np.linalg.solve(A, b[:, :, None])[:, 0, 0]
    return out[:, 0] if f.ndim == 1 else out


def lattice_ij(N, dim, L):
    dx = L / N
    g = np.meshgrid(*([np.arange(N)] * dim), indexing="ij")
    return (np.stack([a.ravel() for a in g], axis=1) + 0.5) * dx


def ur_to_u_mls(N, L, dim, r, u_r):
    """ur_to_u_mls_wrapper(N, L, dim)(r, u_r) -> (N^dim, dim) at (i + 1/2) dx."""
    return mls_2nd_order(r, lattice_ij(N, dim, L), u_r, L, L / N, dim)


def ur_to_u_dft(N, L, dim, r, u_r):
    """ur_to_u_dft_wrapper(N, L, dim)(r, u_r) -> (N^dim, dim) at the nodes m dx."""
    r = np.asarray(r, dtype=np.float64)
    u_r = np.asarray(u_r, dtype=np.float64)
    k = np.fft.fftfreq(N, 1.0 / N)
    kf = np.stack(np.meshgrid(*([k] * dim), indexing="ij"), 0).reshape(dim, -1).T  # (N^dim, dim)
    u_hat = np.empty((kf.shape[0], dim), dtype=np.complex128)
    for a in range(0, kf.shape[0], 256):
        ph = np.exp(-1j * (kf[a:a + 256] @ r.T))  # turbulence.py:323
        u_hat[a:a + 256] = ph @ u_r                # :324
    u_hat = u_hat.T.reshape((dim,) + (N,) * dim)
    u = np.fft.ifftn(u_hat, axes=tuple(range(1, dim + 1))).real  # :330-331
    return u.reshape(dim, -1).T


def u_and_spectrum_from_ur(ckp_N, L, dim, r, u_r, interp_backend="mls"):
    """turbulence.py:339-377 -> (u_grid (dim, N, N, N | 1), E(k))."""
    from . import spectral

    f = ur_to_u_dft if interp_backend == "dft" else ur_to_u_mls
    ug = f(ckp_N, L, dim, r, u_r)
    shape = (dim,) + (ckp_N,) * dim if dim == 3 else (dim, ckp_N, ckp_N, 1)
    ug = ug.T.reshape(shape)
    return ug, spectral.energy_spectrum(ug, dim=dim)
