"""Combined driver: DNS step + filtered-field particle advection (mirror of l3es/combined.py:18-219).

Per step (combined.py:97,127,130,199): advance the spectral field, evaluate
the ckp_N-filtered velocity at the particles, then r <- (r + dt u_r) mod 2 pi.
Without relaxation the filter, interpolation and shift are one fused GPU
call.  With relax_dt_factor > 0 the velocity is interpolated at r, then one
fused call relaxes it and shifts r (combined.py:132-142, :199).  The density
column rho_max comes from comp_rho(r) (combined.py:145-149).  At the burn-in step the
full-resolution field is written as u_{N}_{step:05d}_burnin.npy (combined.py:89-90).  Plotting is not
part of this library; particle frames are written as .npz (no h5py).
"""

import csv
import os
from concurrent.futures import ThreadPoolExecutor
from time import time

import numpy as np
import torch

from .integrator import set_up_integrator
from .relax import relax_advect
from .spectral_solver import diagnostics, rk4_wrapper, rhs_wrapper, set_up_solver, write_u


def _to_host(t, dtype=None):
    """Queue a copy of device tensor t into pinned host memory on the current stream."""
    h = torch.empty(t.shape, dtype=dtype or t.dtype, pin_memory=True)
    h.copy_(t if dtype is None else t.to(dtype), non_blocking=True)
    return h


def _when_copied(done, fn, *args):
    """fn(*args) once the copies queued before event `done` have landed (runs on the writer thread)."""
    done.synchronize()
    fn(*args)


def _save_frame(path, r, u_r):
    """One particle frame from host copies of the positions and velocities taken at that step."""
    np.savez(path, r=r.numpy(), u=u_r.numpy())


def combined(case, N=256, dim=3, nu=0.000625, t_final=0.1, t_burnin=0, dt=0.01, u_ref=4.0, ckp_N=64,
             state_0_path=None, interp_backend="dft", relax_dt_factor=0.0, dft_splits=8, seed=42, rejit=True,
             forcing_type="none", e_kin_target=None, kf=None, log_freq=100, vis_freq=100, ckp_freq=100,
             dst_path=None, precision="float64", viscous_mode="explicit", spectrum="k53", power=0.1, drag=0.0,
             kolm_amp=1.0, kolm_k=4, tol=1e-6, jitter=0.0, verbose=True, write_frames=True):
    del vis_freq
    int_path = os.path.join(dst_path, "com") if dst_path else None
    if int_path:
        os.makedirs(int_path, exist_ok=True)
    u, u_hat, _, dx, integrate_fn, L, fft_axes, e_kin_init, e_inj = set_up_solver(
        N, nu, dim, case, dt, seed, ckp_N, kf, e_kin_target, forcing_type, rejit, precision, viscous_mode,
        spectrum, power, drag, kolm_amp, kolm_k)
    r, comp_rho, interpolator, relax_fn, _, _, _ = set_up_integrator(state_0_path, dim, ckp_N, dft_splits, u_ref,
                                                                      interp_backend, precision, tol, jitter, seed)
    if relax_dt_factor > 0.0 and relax_fn is None:
        raise ValueError("relax_dt_factor > 0 needs u_ref and ckp_N >= 7")
    diag_path = os.path.join(dst_path, "diagnostics.csv") if dst_path else None
    if diag_path:
        with open(diag_path, "w", newline="") as fh:
            csv.writer(fh).writerow(["step", "time", "umax", "ekin", "dt_est", "e_inj", "rho_max"])
    tstep_max = round(t_final / dt)
    burnin_steps = round(t_burnin / dt) if t_burnin > 0 else 0
    e_buf = torch.zeros(1, dtype=torch.float64, device=u.device)
    e_sum = torch.zeros((), dtype=torch.float64, device=u.device)  # sum of the injection rates, kept on the GPU
    marks = []  # per step: events before the DNS step, after it, and after the particles
    t_sim = t_int = 0.0
    lead = 64  # steps the host may queue ahead; also the frame writes in flight

    def fold(keep):
        # add the phase times of all but the last `keep` steps (waiting for them) and drop their events
        nonlocal t_sim, t_int
        while len(marks) > keep:
            a, b, c = marks.pop(0)
            c.synchronize()
            t_sim += a.elapsed_time(b) / 1e3
            t_int += b.elapsed_time(c) / 1e3
    u_r = torch.empty((r.shape[0], dim), dtype=interpolator.dtype, device=r.device)
    t0 = time()
    rows = []
    u_fresh = True  # u holds irfftn(u_hat)
    # Frames and the burn-in snapshot are copied to pinned host memory in stream order at their step
    # and written by one thread while the loop steps on; the writer makes no library call, and the
    # loop waits for every write (and re-raises its errors) before it stops the clock.  The loop
    # itself reads the GPU only on logged steps (and, when verbose, on checkpoints), so the host
    # queues steps ahead of the GPU; CUDA events time the DNS and particle phases.
    pending = []
    with ThreadPoolExecutor(max_workers=1) as writer:
        for i in range(tstep_max + 1):
            if i == burnin_steps and dst_path:  # combined.py:89-90: the full-resolution field before step i
                u_h = _to_host(u, torch.float64)
                done = torch.cuda.Event()
                done.record()
                pending.append(writer.submit(_when_copied, done, write_u, None, i, dst_path, integrate_fn.plan, N,
                                             "_burnin", u_h))
            if i == burnin_steps and case == "HIT":
                integrate_fn = rk4_wrapper(dt, rhs_wrapper(N, nu, fft_axes, precision), None, e_kin_init, "none",
                                           viscous_mode=viscous_mode)
            # combined.py:97.  The step's physical output u is formed only where it is read: the logged
            # diagnostics, the burn-in snapshot before the next step and the returned state (the particles
            # read the spectrum).  Stage 0 takes u when it is current (spectral_solver.py:51) and
            # transforms u_hat itself otherwise.
            want_u = i % log_freq == 0 or i == tstep_max or (dst_path is not None and i + 1 == burnin_steps)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            ev[0].record()
            integrate_fn.advance(u if u_fresh else None, u_hat, 1, u if want_u else None, e_buf)
            u_fresh = want_u
            ev[1].record()
            e_sum += e_buf[0]
            is_ckp = (i % ckp_freq == 0) and (i >= burnin_steps) and int_path is not None and write_frames
            r0 = _to_host(r) if is_ckp else None
            # combined.py:145-149 forms the density on logged steps and checkpoints; on checkpoints it
            # only prints its maximum, so a silent run forms it on logged steps alone
            need_rho = comp_rho is not None and (i % log_freq == 0 or (is_ckp and verbose))
            rho = None
            if relax_dt_factor > 0.0:
                interpolator.from_hat(u_hat, r, 0.0, u_r)  # combined.py:127,130
                u_r0 = _to_host(u_r) if is_ckp else None  # combined.py:131
                rho = comp_rho(r) if need_rho else None  # combined.py:147
                relax_advect(relax_fn, r, u_r, dt, relax_dt_factor)  # combined.py:132-142, :199
            else:
                rho = comp_rho(r) if need_rho else None
                interpolator.from_hat(u_hat, r, dt, u_r)  # combined.py:127,130,199
                u_r0 = _to_host(u_r) if is_ckp else None
            ev[2].record()
            marks.append(ev)
            fold(lead)
            rho_max = float(rho.max()) if rho is not None else None
            if i % log_freq == 0:
                d = diagnostics(u, integrate_fn.plan).cpu().numpy()
                dt_est = float(1.0 / (d[2:5] / dx + nu / dx ** 2).min())
                rows.append([i, float(i * dt), float(d[1]), float(d[0]), dt_est, float(e_buf[0]), rho_max])
                if diag_path:
                    with open(diag_path, "a", newline="") as fh:
                        csv.writer(fh).writerow(rows[-1])
                if verbose:
                    rho_txt = f", rho_max = {rho_max:.3f}" if rho_max is not None else ""
                    print(f"Step {i}/{tstep_max}, u_max = {d[1]:.3f}, E_kin = {d[0]:.3f}, dt_est = {dt_est:.5f}{rho_txt}")
            if is_ckp:
                pending.append(writer.submit(_when_copied, ev[2], _save_frame,
                                             os.path.join(int_path, f"step_{i:05d}.npz"), r0, u_r0))
                while len(pending) > lead:  # bounds the pinned frames held (and raises a failed write)
                    pending.pop(0).result()
                if verbose and rho_max is not None:
                    print(f"Step {i}/{tstep_max}, rho_max={rho_max:.3f}.")
        for f in pending:
            f.result()
        torch.cuda.synchronize()
    t_tot = time() - t0
    fold(0)
    e_acc = float(e_sum)
    if verbose:
        print(f"t_tot = {t_tot:.3f}, t_sim = {t_sim}, t_int = {t_int:.3f}")
    return dict(u=u, u_hat=u_hat, r=r, u_r=u_r, rows=rows, t_sim=t_sim, t_int=t_int, t_tot=t_tot, e_inj_acc=e_acc,
                steps=tstep_max + 1)
