"""Spectral Navier-Stokes solver API on the MI355X HIP path (mirror of l3es/spectral_solver.py).

Same factories and return values as the reference:
  rhs_wrapper(N, nu, fft_axes) -> rhs_fn(u, u_hat, rk) -> du_hat        (spectral_solver.py:27-62)
  rk4_wrapper(dt, rhs_fn, forcing_mask, e_kin_init, forcing_type, fft_axes)
      -> step_fn(u, u_hat) -> (u, u_hat, e_inj)                         (:65-120)
  comp_dt(u, dx, nu, cfl)                                                (:123-127)
  custom_jit_integrate(path_pref, rejit, integrate_fn, u, u_hat)         (:130-153)
  set_up_solver(...) -> (u, u_hat, xyz_vis, dx, integrate_fn, L, fft_axes, e_kin_init, e_inj)  (:156-230)
  simulate(...)                                                          (:233-367)
Arrays are torch tensors on the current HIP device; the arithmetic runs in
libqlles.so.  Extra keyword arguments select the modes this library adds:
``precision`` ("float64" | "float32"), ``viscous_mode`` ("explicit" | "IF"),
``forcing_type`` "power" / "kolmogorov", ``spectrum`` ("k53" | "k4exp").
"""

import csv
import os
from time import time

import numpy as np
import torch

from . import _lib
from .init_fields import (init_forcing_mask_hit, init_u_hit, init_u_kolm, init_u_tgv,
                          init_u_tgv2d)

EPS = float(np.finfo(float).eps)  # spectral_solver.py:23

_PREC = {"float64": torch.float64, "float32": torch.float32, torch.float64: torch.float64,
         torch.float32: torch.float32}


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _dtype(precision):
    return _PREC[precision]


def _kf_from_mask(mask):
    """Forcing radius of a init_forcing_mask_hit mask (max |k|^2 on it is kf^2)."""
    m = np.asarray(mask.cpu() if torch.is_tensor(mask) else mask, dtype=bool)
    N = m.shape[0]
    k1 = np.fft.fftfreq(N, 1.0 / N)
    k3 = k1[: N // 2 + 1].copy()
    k3[-1] = -k3[-1]
    kk = np.asarray(np.meshgrid(k1, k1, k3, indexing="ij"))
    k2 = np.sum(kk ** 2, axis=0)
    kf = int(round(np.sqrt(k2[m].max())))
    if not np.array_equal(m, (k2 <= kf ** 2) & (k2 > 0)):
        raise ValueError("forcing_mask is not a shell mask 0 < |k| <= kf")
    return kf


class _RHS:
    """rhs_fn(u, u_hat, rk) of rhs_wrapper: one fused RHS evaluation on the GPU."""

    def __init__(self, N, nu, fft_axes, precision, forcing_type, kf, power, kolm_k, kolm_amp, drag):
        self.N, self.nu, self.fft_axes = N, nu, tuple(fft_axes)
        self.dim = len(self.fft_axes)
        self.dtype = _dtype(precision)
        self.forcing = dict(forcing_type=forcing_type, kf=kf, power=power, kolm_k=kolm_k,
                            kolm_amp=kolm_amp, drag=drag)
        f = forcing_type if forcing_type in ("power", "kolmogorov") else "none"
        self.plan = _lib.Plan(self.dim, N, self.dtype, nu, 1.0, forcing=f, kf=kf, power=power,
                              kolm_k=kolm_k, kolm_amp=kolm_amp, drag=drag)

    def __call__(self, u, u_hat, rk):
        du = torch.empty_like(u_hat)
        u_in = u if (rk == 0 and u is not None) else None
        _lib.check(_lib.load().qlles_rhs(self.plan.handle, _lib.ptr(u_in), _lib.ptr(u_hat), int(rk),
                                         _lib.ptr(du), _lib.stream_ptr()))
        return du


def rhs_wrapper(N, nu, fft_axes=(1, 2, 3), precision="float64", forcing_type="none", kf=0, power=0.1,
                kolm_k=4, kolm_amp=1.0, drag=0.0):
    """Spectral NS right hand side (spectral_solver.py:27-62)."""
    return _RHS(N, nu, fft_axes, precision, forcing_type, kf, power, kolm_k, kolm_amp, drag)


class _Step:
    """step_fn(u, u_hat) -> (u, u_hat, e_inj) of rk4_wrapper (spectral_solver.py:73-118)."""

    def __init__(self, dt, rhs_fn, forcing_mask, e_kin_init, forcing_type, fft_axes, viscous_mode):
        self.dt = dt
        self.rhs = rhs_fn
        self.N, self.dim, self.dtype = rhs_fn.N, rhs_fn.dim, rhs_fn.dtype
        fz = dict(rhs_fn.forcing)
        kind = "none"
        kf = fz["kf"] or 0
        if forcing_mask is not None and forcing_type == "ekin_tot":  # spectral_solver.py:83
            kind = "ekin_tot"
            kf = _kf_from_mask(forcing_mask)
        elif fz["forcing_type"] in ("power", "kolmogorov"):
            kind = fz["forcing_type"]
        self.forcing_type = kind
        self.plan = _lib.Plan(self.dim, self.N, self.dtype, rhs_fn.nu, dt, forcing=kind, kf=kf,
                              e_kin_init=float(e_kin_init or 0.0), power=fz["power"],
                              kolm_k=fz["kolm_k"], kolm_amp=fz["kolm_amp"], drag=fz["drag"],
                              viscous=viscous_mode)

    def advance(self, u, u_hat, nsteps=1, u_out=None, e_inj=None, use_u=True):
        """nsteps steps in place on u_hat; writes irfftn(u_hat) into u_out if given."""
        _lib.check(_lib.load().qlles_step(self.plan.handle, _lib.ptr(u_out), _lib.ptr(u_hat),
                                          _lib.ptr(u) if (use_u and u is not None) else None,
                                          int(nsteps), _lib.ptr(e_inj), _lib.stream_ptr()))

    def __call__(self, u, u_hat):
        u_hat = u_hat.clone()  # step_fn is functional (spectral_solver.py:74)
        u_new = torch.empty(u.shape, dtype=u.dtype, device=u.device)
        e_inj = torch.zeros(1, dtype=torch.float64, device=u.device)
        self.advance(u.contiguous(), u_hat, 1, u_new, e_inj)
        return u_new, u_hat, e_inj[0]


def rk4_wrapper(dt, rhs_fn, forcing_mask, e_kin_init, forcing_type, fft_axes=(1, 2, 3),
                viscous_mode="explicit"):
    """Runge-Kutta 4 step with optional ekin_tot forcing (spectral_solver.py:65-120)."""
    return _Step(dt, rhs_fn, forcing_mask, e_kin_init, forcing_type, fft_axes, viscous_mode)


def diagnostics(u, plan):
    """FP64 (e_kin, umax, umax_x, umax_y, umax_z) of a physical field."""
    out = torch.zeros(5, dtype=torch.float64, device=u.device)
    _lib.check(_lib.load().qlles_diagnostics(plan.handle, _lib.ptr(u.contiguous()), _lib.ptr(out),
                                             _lib.stream_ptr()))
    return out


def comp_dt(u, dx, nu, cfl=1.0, plan=None):
    """CFL estimate (spectral_solver.py:123-127), FP64 reduction on the GPU."""
    if plan is None:
        from .utils import helper_plan, physical_field

        dim, N, u = physical_field(u)
        plan = helper_plan(dim, N, u.dtype)
    d = diagnostics(u, plan).cpu().numpy()
    umax = d[2:5]
    return float(cfl / (umax / dx + nu / dx ** 2).min())


def custom_jit_integrate(path_pref, rejit, integrate_fn, u, u_hat):
    """Compile/cache hook (spectral_solver.py:130-153).

    The HIP kernels are compiled ahead of time into libqlles.so, so the part of
    the reference's compile step that remains is the choice of kernel variants
    (x/y tile width, z radix plan).  rejit or no ``{path_pref}.json``: time the
    variants on a scratch copy of u_hat (qlles_plan_tune) and store the choice;
    otherwise load it.  path_pref None: tune without storing.
    """
    del u
    import json

    plan = integrate_fn.plan
    path = None if path_pref is None else f"{path_pref}.json"
    if rejit or path is None or not os.path.exists(path):
        us = plan.tune(u_hat.contiguous(), trial_steps=4)
        xv, zv, nv = plan.variant()
        if path is not None:
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            with open(path, "w") as f:
                json.dump(dict(xv=xv, zv=zv, variants=nv, us_per_step=us, N=plan.N, dim=plan.dim,
                               dtype=str(plan.dtype)), f)
    else:
        with open(path) as f:
            cfg = json.load(f)
        plan.set_variant(int(cfg["xv"]), int(cfg["zv"]))
    return integrate_fn


def set_up_solver(N, nu, dim, case, dt, seed, ckp_N, kf, e_kin_target=None, forcing_type="ekin_tot",
                  rejit=True, precision="float64", viscous_mode="explicit", spectrum="k53", power=0.1,
                  drag=0.0, kolm_amp=1.0, kolm_k=4, u_init=None, warmup=True, cache_dir=False):
    """Solver setup (spectral_solver.py:156-230); returns the reference's 9-tuple.

    The kernel variants are tuned here (rejit) or loaded from the ``.cache/`` file
    when ``cache_dir`` is true and ``rejit`` false (the reference's compile cache)."""
    dtype = _dtype(precision)
    dev = _dev()
    L = 2 * np.pi
    dx = L / N
    len_z = N if dim == 3 else 1
    len_z_vis = ckp_N if dim == 3 else 1
    fft_axes = (1, 2, 3) if dim == 3 else (1, 2)
    xyz = np.array(np.meshgrid(np.arange(N), np.arange(N), np.arange(len_z), indexing="ij")) * L / N
    x = np.arange(ckp_N)
    xyz_vis = np.array(np.meshgrid(x, x, np.arange(len_z_vis), indexing="ij")) * L / ckp_N
    if dim == 2:
        xyz_vis = (xyz_vis.T + np.array([0, 0, L - dx])).T

    forcing_mask = None
    if case == "HIT" and forcing_type == "ekin_tot":  # spectral_solver.py:185-188
        forcing_mask = init_forcing_mask_hit(N, kf)

    if u_init is not None:
        u = torch.as_tensor(np.asarray(u_init), dtype=torch.float64).to(dev)
    elif case == "TGV3D":
        u = torch.from_numpy(init_u_tgv(xyz[0], xyz[1], xyz[2])).to(dev)
    elif case == "TGV2D":
        u = torch.from_numpy(init_u_tgv2d(N, target_dim=3, rescale=L)).to(dev)
    elif case == "Kolm":
        u = init_u_kolm(N, seed=seed, target_dim=3)
    elif case == "HIT":
        u = init_u_hit(N, seed, spectrum=spectrum)
    else:
        raise ValueError(f"unknown case {case!r}")
    u = u.to(dtype).contiguous()
    from .utils import helper_plan

    plan0 = helper_plan(dim, N, dtype)  # rfftn and diagnostics only: no step workspace
    e_kin_init = float(diagnostics(u, plan0)[0])  # spectral_solver.py:200
    if forcing_type != "none" and e_kin_target is not None:  # :201-203
        u = (u * np.sqrt(e_kin_target / (e_kin_init + EPS))).contiguous()
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    shape_hat = (3, N, N, N // 2 + 1) if dim == 3 else (3, N, N // 2 + 1)
    u_hat = torch.empty(shape_hat, dtype=cdt, device=dev)
    _lib.check(_lib.load().qlles_rfftn(plan0.handle, _lib.ptr(u), _lib.ptr(u_hat), _lib.stream_ptr()))  # :206
    e_kin_init = float(diagnostics(u, plan0)[0])  # :209

    rhs_kind = forcing_type if forcing_type in ("power", "kolmogorov") else "none"
    integrate_fn = rk4_wrapper(dt, rhs_wrapper(N, nu, fft_axes, precision, rhs_kind, kf or 0, power,
                                               kolm_k, kolm_amp, drag),
                               forcing_mask, e_kin_init, forcing_type, fft_axes, viscous_mode)
    # cache name as spectral_solver.py:218-221, plus the precision
    device_type = "gpu"
    if forcing_type != "none":
        path_pref = f".cache/spectral_{device_type}_{N}_{dim}_{kf}_{forcing_type}"
    else:
        path_pref = f".cache/spectral_{device_type}_{N}_{dim}"
    if dtype == torch.float32:
        path_pref += "_f32"
    integrate_fn = custom_jit_integrate(path_pref if cache_dir else None, rejit, integrate_fn, u, u_hat)
    e_inj = torch.zeros((), dtype=torch.float64, device=dev)
    if warmup:  # the reference advances one step during setup (:226)
        u, u_hat, e_inj = integrate_fn(u, u_hat)
    return u, u_hat, xyz_vis, dx, integrate_fn, L, fft_axes, e_kin_init, e_inj


def write_u(u_hat, i, dst_path, plan, ckp_N, suffix="", u=None):
    """Spectrally filtered velocity checkpoint u_{ckp_N}_{i:05d}{suffix}.npy (utils.py:141-166):
    (3, n, n, n) in 3-D, (2, n, n) in 2-D, float64 like the reference's x64 arrays.  With
    ckp_N == N the reference writes the physical field itself (:161); so does this when ``u``
    (the physical field of u_hat) is given."""
    from .utils import spectral_filtering_hat

    if ckp_N > plan.N or ckp_N % 2 or plan.N % 2:
        raise ValueError("ckp_N must be even and <= N (utils.py:157-158)")
    os.makedirs(dst_path, exist_ok=True)
    if u is not None and ckp_N == plan.N:
        arr = u.double().cpu().numpy()
        arr = arr if plan.dim == 3 else arr[:2, :, :, 0]
    else:
        arr = spectral_filtering_hat(u_hat, ckp_N, plan).double().cpu().numpy()
    if plan.dim == 2:
        arr = arr[:2]
    np.save(os.path.join(dst_path, f"u_{ckp_N}_{i:05d}{suffix}.npy"), arr)


def simulate(case="TGV3D", N=64, dim=3, nu=0.000625, t_final=0.1, t_burnin=0.0, dt=0.01, u_ref=1.0,
             ckp_N=32, seed=42, rejit=True, forcing_type="none", e_kin_target=None, kf=None,
             log_freq=20, vis_freq=10 ** 8, ckp_freq=1, dst_path=None, precision="float64",
             viscous_mode="explicit", spectrum="k53", power=0.1, drag=0.0, kolm_amp=1.0, kolm_k=4,
             write_checkpoints=True, verbose=True):
    """Simulation driver (spectral_solver.py:233-367); visualisation is not produced."""
    del u_ref, vis_freq
    u, u_hat, _, dx, integrate_fn, L, _, e_kin_init, e_inj = set_up_solver(
        N, nu, dim, case, dt, seed, ckp_N, kf, e_kin_target, forcing_type, rejit, precision,
        viscous_mode, spectrum, power, drag, kolm_amp, kolm_k)
    dst_ckp = os.path.join(dst_path, "sim") if dst_path else None
    diagnostics_path = os.path.join(dst_path, "diagnostics.csv") if dst_path else None
    if diagnostics_path:
        os.makedirs(dst_path, exist_ok=True)
        with open(diagnostics_path, "w", newline="") as fh:
            csv.writer(fh).writerow(["step", "time", "umax", "ekin", "dt_est", "e_inj"])
    rows = []
    t0 = time()
    tstep_max = round(t_final / dt)
    burnin_steps = round(t_burnin / dt) if t_burnin > 0 else 0
    e_buf = torch.zeros(1, dtype=torch.float64, device=u.device)
    # the loop reads the GPU only on logged steps and checkpoints, so the host queues steps ahead:
    # the injection rates are summed on the device and CUDA events time the steps
    e_sum = torch.zeros((), dtype=torch.float64, device=u.device)
    marks = []
    t_sim = 0.0

    def fold(keep):
        # add the step times of all but the last `keep` steps (waiting for them) and drop their events
        nonlocal t_sim
        while len(marks) > keep:
            a, b = marks.pop(0)
            b.synchronize()
            t_sim += a.elapsed_time(b) / 1e3

    u_fresh = True  # u holds irfftn(u_hat)
    for i in range(tstep_max + 1):
        if i == burnin_steps and case == "HIT":
            # switch to a solver without forcing (spectral_solver.py:305-309)
            integrate_fn = rk4_wrapper(dt, rhs_wrapper(N, nu, (1, 2, 3) if dim == 3 else (1, 2), precision),
                                       None, e_kin_init, "none", viscous_mode=viscous_mode)
        # the step's physical output u is formed only where it is read: the logged diagnostics, the
        # full-resolution snapshots and the returned state (filtered checkpoints read the spectrum).
        # Stage 0 takes u when it is current (spectral_solver.py:51) and transforms u_hat otherwise.
        writes = write_checkpoints and dst_path is not None
        want_u = (i % log_freq == 0 or i == tstep_max or (writes and i == burnin_steps)
                  or (writes and ckp_N == N and i % ckp_freq == 0 and i >= burnin_steps))
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
        integrate_fn.advance(u if u_fresh else None, u_hat, 1, u if want_u else None, e_buf)
        ev[1].record()
        marks.append(ev)
        fold(64)  # the host queues at most 64 steps ahead
        u_fresh = want_u
        e_sum += e_buf[0]
        if i % log_freq == 0:
            e_inj = float(e_buf[0])
            d = diagnostics(u, integrate_fn.plan).cpu().numpy()
            e_kin, umax = float(d[0]), float(d[1])
            dt_est = float(1.0 / (d[2:5] / dx + nu / dx ** 2).min())
            rows.append([i, float(i * dt), umax, e_kin, dt_est, e_inj])
            if diagnostics_path:
                with open(diagnostics_path, "a", newline="") as fh:
                    csv.writer(fh).writerow(rows[-1])
            if verbose:
                print(f"Step {i}/{tstep_max}, u_max = {umax:.3f}, E_kin = {e_kin:.3f}, "
                      f"dt_est = {dt_est:.5f}, E_inj = {e_inj:.3f}")
        if write_checkpoints and dst_path:
            if i == burnin_steps:
                write_u(u_hat, i, dst_path, integrate_fn.plan, N, suffix="_burnin", u=u)
            if i % ckp_freq == 0 and i >= burnin_steps:
                write_u(u_hat, i, dst_ckp, integrate_fn.plan, ckp_N, u=u)
    torch.cuda.synchronize()
    e_inj_acc = float(e_sum)
    fold(0)
    stats = None
    if case == "HIT" and forcing_type != "none":  # spectral_solver.py:344-358
        stats = hit_statistics(e_inj_acc, tstep_max, kf, N, nu, t_final, rows[-1][3])
        if verbose:
            print(f"Eddy turnover time =            {stats['eddy_turnover_time']:.3f}")
            print(f"Number of eddy turnover times = {stats['eddy_turnover_times']:.3f}")
            print(f"Average energy injected =       {stats['e_inj_mean']:.3f}")
            print(f"eta=                            {stats['eta']:.3f}")
            print(f"kmax*eta =                      {stats['kmax_eta']:.3f}")
            print(f"Re_lambda =                     {stats['re_lambda']:.3f}")
    t_tot = time() - t0
    if verbose:
        print(f"t_tot = {t_tot:.3f}, t_sim = {t_sim:.3f}")
    e_kin = float(diagnostics(u, integrate_fn.plan)[0])
    if t_final == 0.1 and case == "TGV3D":  # spectral_solver.py:363-367
        assert round(e_kin - 0.124953117517, 2) == 0, f"e_kin = {e_kin}"
    return dict(u=u, u_hat=u_hat, e_kin=e_kin, rows=rows, e_inj_acc=e_inj_acc, t_sim=t_sim, hit_stats=stats)


def hit_statistics(e_inj_acc, steps, kf, N, nu, t_final, e_kin):
    """End-of-run statistics of forced HIT (spectral_solver.py:344-358), host arithmetic.

    ``e_inj_acc`` is the injection rate summed over the loop, ``steps`` the last loop index
    (the reference divides by it), ``e_kin`` the last logged kinetic energy.  Returns the mean
    injection rate eps, the eddy turnover time 1 / (eps kf^2)^(1/3), the number of turnovers
    in t_final, the Kolmogorov scale eta = (nu^3 / eps)^(1/4), kmax eta with kmax = sqrt(2) N / 3,
    and Re_lambda = sqrt(20 E^2 / (3 nu eps)).
    """
    eps = e_inj_acc / steps
    t_e = 1.0 / (eps * kf ** 2) ** (1.0 / 3.0)
    eta = (nu ** 3 / eps) ** 0.25
    return dict(e_inj_mean=eps, eddy_turnover_time=t_e, eddy_turnover_times=t_final / t_e, eta=eta,
                kmax_eta=np.sqrt(2) * N / 3 * eta, re_lambda=float(np.sqrt(20 * e_kin ** 2 / (3 * nu * eps))))
