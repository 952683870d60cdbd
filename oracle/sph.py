"""NumPy FP64 restatement of the reference SPH density and particle relaxation (test infrastructure).

Only tests, the smoke check and the CPU-baseline leg of the benchmark use this
module.  Functions cite the reference file:line they restate (paths relative
to the tumaer/ql-les-data checkout).

The reference builds these terms from jax_sph 0.1.0 (not vendored, not
importable here), so its pieces are restated from the published methods:
  - QuinticKernel: the quintic spline of Morris, Fox & Zhu (J. Comput. Phys.
    136, 1997), support radius 3h,
        W(q) = sigma [(3-q)_+^5 - 6 (2-q)_+^5 + 15 (1-q)_+^5],  q = r/h,
        sigma_2D = 7 / (478 pi h^2),  sigma_3D = 3 / (359 pi h^3),
    grad_w = dW/dr (the reference takes it by autodiff of w);
  - TaitEoS: p = p_ref ((rho/rho_ref)^gamma - 1) + p_background;
  - jax_md space.periodic: displacement mod(Ra - Rb + L/2, L) - L/2, shift
    mod(R + dR, L);
  - neighbor_list(mask_self=False): every ordered pair (i, j), i == j
    included, closer than the cutoff 3h under the minimum image.  W and
    dW/dr vanish at and beyond 3h, so the list only decides which exactly-zero
    terms are skipped; the restatement enumerates pairs by brute force
    (neighbor_pairs) and, independently, with scipy's periodic cKDTree.

Parity pins (tests/test_oracle_sph.py): the kernel integrates to 1 in 2-D and
3-D (fixes both sigmas), grad_w equals a central difference of w, W is C^4 at
q = 1, 2, 3; the two neighbour searches agree; a perfect lattice has density
~1 and zero acceleration; pairwise forces conserve momentum; a displaced
particle is pushed back (the sign); viscous forces oppose relative motion.
"""

import numpy as np

EPS = np.finfo(float).eps  # relax.py:13 (jax_enable_x64 is on, spectral_solver.py:24)

__all__ = [
    "EPS", "SPHParams", "quintic_w", "quintic_grad_w", "tait_p", "periodic_displacement", "periodic_shift",
    "neighbor_pairs", "density", "relax_acc", "sph_forces", "relax_advect", "lattice",
]


def quintic_sigma(h, dim):
    """Normalisation of the Morris quintic spline (jax_sph QuinticKernel)."""
    if dim == 1:
        return 1.0 / 120.0 / h
    if dim == 2:
        return 7.0 / 478.0 / np.pi / h ** 2
    return 3.0 / 359.0 / np.pi / h ** 3


def quintic_w(r, h, dim):
    """kernel_fn.w (relax.py:40,85; utils.py:186,214)."""
    q = np.asarray(r, dtype=np.float64) / h
    q1 = np.maximum(0.0, 1.0 - q)
    q2 = np.maximum(0.0, 2.0 - q)
    q3 = np.maximum(0.0, 3.0 - q)
    return quintic_sigma(h, dim) * (q3 ** 5 - 6.0 * q2 ** 5 + 15.0 * q1 ** 5)


def quintic_grad_w(r, h, dim):
    """kernel_fn.grad_w = dW/dr (relax.py:95)."""
    q = np.asarray(r, dtype=np.float64) / h
    q1 = np.maximum(0.0, 1.0 - q)
    q2 = np.maximum(0.0, 2.0 - q)
    q3 = np.maximum(0.0, 3.0 - q)
    return quintic_sigma(h, dim) / h * (-5.0 * q3 ** 4 + 30.0 * q2 ** 4 - 75.0 * q1 ** 4)


def tait_p(rho, p_ref, rho_ref=1.0, p_background=0.0, gamma=1.0):
    """eos.p_fn (relax.py:66,88)."""
    return p_ref * ((rho / rho_ref) ** gamma - 1.0) + p_background


def periodic_displacement(ra, rb, side):
    """space.periodic displacement_fn: Ra - Rb wrapped into [-L/2, L/2)."""
    d = ra - rb
    return np.mod(d + side * 0.5, side) - 0.5 * side


def periodic_shift(r, dr, side):
    """space.periodic shift_fn (integrator.py:21-23): mod(R + dR, L)."""
    return np.mod(r + dr, side)


class SPHParams:
    """Static parameters of relax_wrapper (relax.py:16-74) / rho_computer (utils.py:169-206).

    is_physical=False works in units of l_ref = dx (relax.py:33-34).
    """

    def __init__(self, N, dim=3, L=2 * np.pi, is_physical=True, u_ref=1.0):
        self.N, self.dim, self.L = N, dim, L
        dx_phys = L / N
        self.l_ref = 1.0 if is_physical else dx_phys
        self.dx = dx_phys / self.l_ref
        self.rho_ref = 1.0
        self.mass = self.dx ** dim * self.rho_ref
        self.box = L / self.l_ref
        self.h = self.dx
        self.cutoff = 3.0 * self.h
        u_eos = (u_ref if u_ref is not None else 0.0) / self.l_ref
        c_eos = 10.0 * u_eos
        self.p_eos = c_eos ** 2 * self.rho_ref


def neighbor_pairs(r, cutoff, side, method="brute", chunk=2048):
    """Neighbour list (relax.py:48-57, :80-81): ordered pairs (i, j), i == j
    included, with minimum-image distance < cutoff, sorted by (i, j)."""
    r = np.asarray(r, dtype=np.float64)
    n = r.shape[0]
    if method == "kdtree":
        from scipy.spatial import cKDTree

        tree = cKDTree(np.mod(r, side), boxsize=side)
        pairs = tree.query_pairs(cutoff, output_type="ndarray")
        i = np.concatenate([pairs[:, 0], pairs[:, 1], np.arange(n)])
        j = np.concatenate([pairs[:, 1], pairs[:, 0], np.arange(n)])
        d = np.linalg.norm(periodic_displacement(r[i], r[j], side), axis=1)
        keep = d < cutoff  # cKDTree includes d == cutoff
        i, j = i[keep], j[keep]
    else:
        ii, jj = [], []
        for a in range(0, n, chunk):
            dr = periodic_displacement(r[a:a + chunk, None, :], r[None, :, :], side)
            d2 = np.sqrt(np.sum(dr * dr, axis=-1))
            pi, pj = np.nonzero(d2 < cutoff)
            ii.append(pi + a)
            jj.append(pj)
        i = np.concatenate(ii) if ii else np.zeros(0, np.int64)
        j = np.concatenate(jj) if jj else np.zeros(0, np.int64)
    o = np.lexsort((j, i))
    return i[o].astype(np.int64), j[o].astype(np.int64)


def _pair_geometry(r, p, method):
    i, j = neighbor_pairs(r, p.cutoff, p.box, method)
    dr = periodic_displacement(r[i], r[j], p.box)
    d = np.sqrt(np.sum(dr * dr, axis=1))
    return i, j, dr, d


def density(r, N, dim=3, L=2 * np.pi, method="brute"):
    """comp_rho of rho_computer (utils.py:169-218): rho_i = m sum_j W(|r_ij|)."""
    p = SPHParams(N, dim, L, is_physical=True)
    r = np.asarray(r, dtype=np.float64)
    i, _, _, d = _pair_geometry(r, p, method)
    return p.mass * np.bincount(i, weights=quintic_w(d, p.h, dim), minlength=r.shape[0])


def _rho_p(r, p, i, d):
    rho = p.mass * np.bincount(i, weights=quintic_w(d, p.h, p.dim), minlength=r.shape[0])  # relax.py:87
    return rho, tait_p(rho, p.p_eos, p.rho_ref)  # relax.py:88


def relax_acc(r, N, dim=3, L=2 * np.pi, is_physical=False, u_ref=1.0, method="brute", return_rho=False):
    """loop_body of relax_wrapper (relax.py:76-120): pressure + 0.5 * background
    (transport-velocity) acceleration per particle, in physical length units."""
    p = SPHParams(N, dim, L, is_physical, u_ref)
    r = np.asarray(r, dtype=np.float64) / p.l_ref  # relax.py:78
    i, j, dr, d = _pair_geometry(r, p, method)
    rho, pr = _rho_p(r, p, i, d)
    e = dr / (d + EPS)[:, None]  # relax.py:92
    kg = quintic_grad_w(d, p.h, dim)[:, None] * e  # relax.py:95-96
    rho_i, rho_j, p_i, p_j = rho[i], rho[j], pr[i], pr[j]
    p_ij = ((rho_j * p_i + rho_i * p_j) / (rho_i + rho_j))[:, None]  # relax.py:99
    pref = (p.mass * ((1 / rho_i) ** 2 + (1 / rho_j) ** 2))[:, None]  # relax.py:102
    acc = -pref * p_ij * kg  # relax.py:103
    acc = acc + 0.5 * pref * (-p.p_eos) * kg  # relax.py:107
    out = np.stack([np.bincount(i, weights=acc[:, c], minlength=r.shape[0]) for c in range(dim)], axis=1)
    out = out * p.l_ref  # relax.py:119
    return (out, rho) if return_rho else out


def sph_forces(r, u, N, dim=3, L=2 * np.pi, u_ref=1.0, method="brute"):
    """sph_body of relax_wrapper (relax.py:123-176): acc_p, acc_visc, acc_tvf
    (physical units only, relax.py:126)."""
    p = SPHParams(N, dim, L, True, u_ref)
    r = np.asarray(r, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    i, j, dr, d = _pair_geometry(r, p, method)
    rho, pr = _rho_p(r, p, i, d)
    e = dr / (d + EPS)[:, None]  # relax.py:138
    der = quintic_grad_w(d, p.h, dim)
    kg = der[:, None] * e  # relax.py:141-142
    rho_i, rho_j, p_i, p_j = rho[i], rho[j], pr[i], pr[j]
    p_ij = ((rho_j * p_i + rho_i * p_j) / (rho_i + rho_j))[:, None]  # relax.py:145
    pref = (p.mass * ((1 / rho_i) ** 2 + (1 / rho_j) ** 2))[:, None]  # relax.py:148
    acc_p = -pref * p_ij * kg  # relax.py:149
    u_ij = u[i] - u[j]  # relax.py:151
    acc_visc = pref * 1.0 * u_ij / (d + EPS)[:, None] * der[:, None]  # relax.py:153-154
    acc_tvf = pref * (-p.p_eos) * kg  # relax.py:158
    seg = lambda a: np.stack([np.bincount(i, weights=a[:, c], minlength=r.shape[0]) for c in range(dim)], axis=1)
    return {"acc_p": seg(acc_p), "acc_visc": seg(acc_visc), "acc_tvf": seg(acc_tvf)}


def relax_advect(r, u_r, dt, factor, N, dim=3, L=2 * np.pi, u_ref=1.0, method="brute"):
    """The relaxation branch of the combined loop plus the final shift
    (combined.py:132-142, :199), with relax_fn = relax_wrapper(N, dim, L,
    is_physical=True, u_ref) (integrator.py:112):
        r_t = shift(r, dt u_r); a = relax_fn(r_t); r_t = shift(r_t, (f dt)^2 a)
        u_r = displacement(r_t, r) / dt;  r = shift(r, dt u_r)
    Returns (r_new, u_r_new, a)."""
    r = np.asarray(r, dtype=np.float64)
    u_r = np.asarray(u_r, dtype=np.float64)
    r_t = periodic_shift(r, dt * u_r, L)
    a = relax_acc(r_t, N, dim, L, True, u_ref, method)
    r_t = periodic_shift(r_t, (factor * dt) ** 2 * a, L)
    u_new = periodic_displacement(r_t, r, L) / dt
    return periodic_shift(r, dt * u_new, L), u_new, a


def lattice(n, dim, L=2 * np.pi):
    """Cell-centred Cartesian lattice (i + 1/2) dx, 'ij' order (pos_init_cartesian_*d)."""
    dx = L / n
    g = np.meshgrid(*([np.arange(n)] * dim), indexing="ij")
    return (np.stack([a.ravel() for a in g], axis=1) + 0.5) * dx
