"""Tangent-linear / adjoint maps of the RK4 step and a differentiable, checkpointed integrator.

The reference computes no gradients (its solver is a plain JAX forward map, spectral_solver.py:73-118).
The maps here are the exact tangent-linear model and adjoint of the classical RK4 step
(spectral_solver.py:70-80) with the rotational-form RHS (:50-60): 3-D, explicit viscous term, forcing
none or fixed power (the power term is differentiated through its energy normalisation, and the
adjoint also returns the cotangent of the injection power).  They run in libqlles.so (qlles_{rhs,step}_{jvp,vjp}).  Inner products: Re sum_k conj(a_k) b_k
over the stored half-spectrum, which is torch's convention for gradients of complex tensors, and
sum_x a.b for physical fields.

``integrate(step, u_hat0, nsteps, every)`` is a torch.autograd.Function.  When the records of all nsteps
steps fit in a quarter of the free device memory (qlles_step_record: a step taken stage by stage that
keeps its stage states and their physical u, omega; 278 MB per step at 128^3 FP32, so 50 steps take
14 GB of the 288 GB), its forward records every step and its backward applies the recorded step
adjoints (qlles_step_vjp_record) in reverse, with no recomputation.  Otherwise the forward takes nsteps
fused RK4 steps and keeps a checkpoint every ``every`` steps, and the backward recomputes each segment
from its checkpoint (with recorded steps when a segment's records fit) before applying its adjoints.
The recorded step equals the fused one to rounding.  ``irfftn(u_hat, plan)`` is differentiable too.
"""

import torch

from . import _lib


def _run(name, plan, a, b, out, dpow=None):
    lib = getattr(_lib.load(), name)
    if name.endswith("_vjp"):
        _lib.check(lib(plan.handle, _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), _lib.ptr(dpow), _lib.stream_ptr()))
    else:
        _lib.check(lib(plan.handle, _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), _lib.stream_ptr()))
    return out


def rhs_jvp(step, u_hat, v):
    """R'(u_hat) v for rhs_fn(u, u_hat, rk > 0)."""
    return _run("qlles_rhs_jvp", step.plan, u_hat.contiguous(), v.contiguous(), torch.empty_like(u_hat))


def rhs_vjp(step, u_hat, lam, dpower=None):
    """R'(u_hat)^T lam; ``dpower`` (float64 device tensor of one element) accumulates <lam, dR/dP>."""
    return _run("qlles_rhs_vjp", step.plan, u_hat.contiguous(), lam.contiguous(), torch.empty_like(u_hat), dpower)


def step_jvp(step, u_hat, v):
    """Tangent of one RK4 step taken from u_hat, applied to v."""
    return _run("qlles_step_jvp", step.plan, u_hat.contiguous(), v.contiguous(), torch.empty_like(u_hat))


def step_vjp(step, u_hat, lam, dpower=None):
    """Adjoint of one RK4 step taken from u_hat, applied to lam (``dpower`` as in rhs_vjp)."""
    return _run("qlles_step_vjp", step.plan, u_hat.contiguous(), lam.contiguous(), torch.empty_like(u_hat), dpower)


def _advance(step, uh):
    step.advance(None, uh, 1, None, use_u=False)


def record_bytes(step):
    """Bytes of one step record (0 where recorded steps are not available)."""
    return int(_lib.load().qlles_step_record_bytes(step.plan.handle))


def step_record(step, u_hat, record, out=None):
    """One RK4 step of u_hat taken stage by stage, its stage states kept in ``record`` (a uint8 device
    tensor of record_bytes(step)); returns the stepped spectrum (``out``, may be u_hat)."""
    out = torch.empty_like(u_hat) if out is None else out
    _lib.check(_lib.load().qlles_step_record(step.plan.handle, _lib.ptr(u_hat), _lib.ptr(out), _lib.ptr(record),
                                             _lib.stream_ptr()))
    return out


def step_vjp_record(step, u_hat, record, lam, dpower=None):
    """step_vjp at the u_hat ``record`` was made from, reading the record."""
    out = torch.empty_like(u_hat)
    _lib.check(_lib.load().qlles_step_vjp_record(step.plan.handle, _lib.ptr(u_hat), _lib.ptr(record),
                                                 _lib.ptr(lam.contiguous()), _lib.ptr(out), _lib.ptr(dpower),
                                                 _lib.stream_ptr()))
    return out


def _records(step, every, like):
    """``every`` step records when they fit in a quarter of the free device memory, else None."""
    rb = record_bytes(step)
    if rb == 0:
        return None
    free, _ = torch.cuda.mem_get_info(like.device)
    if every * rb > free // 4:
        return None
    return [torch.empty(rb, dtype=torch.uint8, device=like.device) for _ in range(every)]


def _replay(step, ckpt, nsteps, records):
    """The states y_0 .. y_nsteps of a segment from its checkpoint y_0; with records, step i is recorded
    into records[i]."""
    states = [ckpt]
    uh = ckpt.clone()
    for i in range(nsteps):
        if records is not None:
            uh = step_record(step, uh, records[i])
        else:
            _advance(step, uh)
        states.append(uh.clone() if records is None else uh)
    return states


def _step_adjoint(step, states, records, i, lam, dpow):
    if records is not None:
        return step_vjp_record(step, states[i], records[i], lam, dpow)
    return step_vjp(step, states[i], lam, dpow)


def _recorded_forward(step, uh, nsteps, records, after=None):
    """nsteps recorded steps from uh: the step inputs u_hat_0 .. u_hat_nsteps (new tensors); after(n, u_hat)
    runs after step n (the particle advection)."""
    states = [uh]
    for n in range(nsteps):
        uh = step_record(step, uh, records[n])
        states.append(uh)
        if after is not None:
            after(n, uh)
    return states


class _Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat0, power, step, nsteps, every, grads):
        uh = u_hat0.detach().contiguous().clone()
        ctx.step, ctx.nsteps, ctx.every = step, nsteps, every
        ctx.records = _records(step, nsteps, uh) if grads else None  # every step recorded when all fit
        if ctx.records is not None:
            ctx.states = _recorded_forward(step, uh, nsteps, ctx.records)
            ctx.ckpts = None
            return ctx.states[-1].clone()
        ckpts = []
        for n in range(nsteps):
            if n % every == 0:
                ckpts.append(uh.clone())
            _advance(step, uh)
        ctx.ckpts = ckpts
        return uh

    @staticmethod
    def backward(ctx, grad):
        step, nsteps, every = ctx.step, ctx.nsteps, ctx.every
        lam = grad.contiguous()
        dpow = torch.zeros(1, dtype=torch.float64, device=grad.device) if ctx.needs_input_grad[1] else None
        if ctx.records is not None:
            for i in reversed(range(nsteps)):
                lam = step_vjp_record(step, ctx.states[i], ctx.records[i], lam, dpow)
            ctx.states = ctx.records = None
            gp = None if dpow is None else dpow.reshape(()).to(torch.float64)
            return lam, gp, None, None, None, None
        records = _records(step, every, lam)
        for seg in reversed(range(len(ctx.ckpts))):
            start = seg * every
            stop = min(start + every, nsteps)
            # the segment's step inputs (the last step is taken only when it is recorded)
            states = _replay(step, ctx.ckpts[seg], stop - start - (records is None), records)
            for i in reversed(range(stop - start)):
                lam = _step_adjoint(step, states, records, i, lam, dpow)
        ctx.ckpts = None
        gp = None if dpow is None else dpow.reshape(()).to(torch.float64)
        return lam, gp, None, None, None, None


def integrate(step, u_hat0, nsteps, every=10, power=None, record=True):
    """nsteps RK4 steps of u_hat0, differentiable: every step recorded when all records fit (and
    ``record``), else checkpoints every ``every`` steps.

    ``power``: optional 0-d float64 tensor equal to the plan's injection power; when it requires grad,
    backward returns dL/dP for a fixed-power plan (the forward uses the plan's own value)."""
    if nsteps < 1 or every < 1:
        raise ValueError("need nsteps >= 1 and every >= 1")
    if power is None:
        power = torch.zeros((), dtype=torch.float64, device=u_hat0.device)
    elif abs(float(power.detach()) - float(step.plan.desc.power)) > 1e-12 * max(1.0, abs(float(power.detach()))):
        raise ValueError("power must equal the plan's injection power (create a plan per value)")
    grads = record and torch.is_grad_enabled() and (u_hat0.requires_grad or power.requires_grad)
    return _Integrate.apply(u_hat0, power, step, int(nsteps), int(every), grads)


class _Irfftn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat, plan, shape):
        ctx.plan = plan
        u = torch.empty(shape, dtype=torch.float32 if u_hat.dtype == torch.complex64 else torch.float64,
                        device=u_hat.device)
        _lib.check(_lib.load().qlles_irfftn(plan.handle, _lib.ptr(u_hat.detach().contiguous()), _lib.ptr(u),
                                            _lib.stream_ptr()))
        ctx.hat_shape = u_hat.shape
        return u

    @staticmethod
    def backward(ctx, g):
        out = torch.empty(ctx.hat_shape, dtype=torch.complex64 if g.dtype == torch.float32 else torch.complex128,
                          device=g.device)
        _lib.check(_lib.load().qlles_irfftn_adjoint(ctx.plan.handle, _lib.ptr(g.contiguous()), _lib.ptr(out),
                                                    _lib.stream_ptr()))
        return out, None, None


def irfftn(u_hat, plan):
    """irfftn(u_hat) over the plan's axes (spectral_solver.py:114), differentiable."""
    N = u_hat.shape[1]
    shape = (3, N, N, N) if u_hat.dim() == 4 else (3, N, N)
    return _Irfftn.apply(u_hat, plan, shape)


# ---------------------------------------------------------------------------
# particles: combined-mode loop (combined.py:97,127,130,199)
#   u_hat_{n+1} = step(u_hat_n);  r_{n+1} = (r_n + dt interp(filter(u_hat_{n+1}), r_n)) mod 2 pi
# ---------------------------------------------------------------------------
def interp_vjp(interp, u_hat, r, u_bar, lam=None, r_bar=None):
    """Adjoint of u_r = interpolate(filter(irfftn(u_hat)), r): adds (du_r/du_hat)^T u_bar to ``lam`` and
    (du_r/dr)^T u_bar to ``r_bar`` (both created when None); returns (lam, r_bar)."""
    if lam is None:
        lam = torch.zeros_like(u_hat)
    if r_bar is None:
        r_bar = torch.zeros_like(r, dtype=torch.float64)
    _lib.check(_lib.load().qlles_interp_hat_vjp(interp.it.handle, u_hat.shape[1], _lib.ptr(u_hat.contiguous()),
                                                _lib.ptr(r.contiguous()), r.shape[0],
                                                _lib.ptr(u_bar.to(interp.dtype).contiguous()), _lib.ptr(lam),
                                                _lib.ptr(r_bar), _lib.stream_ptr()))
    return lam, r_bar


class _IntegrateParticles(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat0, r0, power, step, interp, nsteps, every, grads):
        dt = float(step.plan.desc.dt)
        uh = u_hat0.detach().contiguous().clone()
        r = r0.detach().to(torch.float64).contiguous().clone()
        rs = [r.clone()]
        ctx.step, ctx.interp, ctx.nsteps, ctx.every, ctx.dt = step, interp, nsteps, every, dt
        ctx.records = _records(step, nsteps, uh) if grads else None  # every step recorded when all fit

        def advect(n, u_hat):
            interp.from_hat(u_hat, r, dt)  # r <- (r + dt u_r) mod 2 pi, in place
            rs.append(r.clone())

        if ctx.records is not None:
            ctx.states = _recorded_forward(step, uh, nsteps, ctx.records, advect)
            ctx.ckpts, ctx.rs = None, rs
            return ctx.states[-1].clone(), r
        ckpts = []
        for n in range(nsteps):
            if n % every == 0:
                ckpts.append(uh.clone())
            _advance(step, uh)
            advect(n, uh)
        ctx.ckpts, ctx.rs = ckpts, rs
        return uh, r

    @staticmethod
    def backward(ctx, g_uh, g_r):
        step, interp, nsteps, every, dt = ctx.step, ctx.interp, ctx.nsteps, ctx.every, ctx.dt
        like = ctx.states[0] if ctx.records is not None else ctx.ckpts[0]
        lam = torch.zeros_like(like) if g_uh is None else g_uh.contiguous().clone()
        lr = torch.zeros_like(ctx.rs[0]) if g_r is None else g_r.to(torch.float64).contiguous().clone()
        dpow = torch.zeros(1, dtype=torch.float64, device=lam.device) if ctx.needs_input_grad[2] else None
        if ctx.records is not None:
            for n in reversed(range(nsteps)):
                lam, lr = interp_vjp(interp, ctx.states[n + 1], ctx.rs[n], dt * lr, lam, lr.clone())
                lam = step_vjp_record(step, ctx.states[n], ctx.records[n], lam, dpow)
            ctx.states = ctx.records = ctx.rs = None
            gp = None if dpow is None else dpow.reshape(()).to(torch.float64)
            return lam, lr, gp, None, None, None, None, None
        records = _records(step, every, lam)
        for seg in reversed(range(len(ctx.ckpts))):
            start = seg * every
            stop = min(start + every, nsteps)
            states = _replay(step, ctx.ckpts[seg], stop - start, records)  # u_hat_start .. u_hat_stop
            for n in reversed(range(start, stop)):
                # r_{n+1} = r_n + dt u_r(u_hat_{n+1}, r_n)
                lam, lr = interp_vjp(interp, states[n - start + 1], ctx.rs[n], dt * lr, lam, lr.clone())
                lam = _step_adjoint(step, states, records, n - start, lam, dpow)
        ctx.ckpts = ctx.rs = None
        gp = None if dpow is None else dpow.reshape(()).to(torch.float64)
        return lam, lr, gp, None, None, None, None, None


def integrate_particles(step, interp, u_hat0, r0, nsteps, every=10, power=None, record=True):
    """The combined-mode loop for nsteps: returns (u_hat_K, r_K), differentiable with respect to u_hat0,
    r0 and (fixed-power plans) the injection power; every step recorded when all records fit (and
    ``record``), else checkpoints every ``every`` steps."""
    if nsteps < 1 or every < 1:
        raise ValueError("need nsteps >= 1 and every >= 1")
    if power is None:
        power = torch.zeros((), dtype=torch.float64, device=u_hat0.device)
    elif abs(float(power.detach()) - float(step.plan.desc.power)) > 1e-12 * max(1.0, abs(float(power.detach()))):
        raise ValueError("power must equal the plan's injection power (create a plan per value)")
    grads = record and torch.is_grad_enabled() and (u_hat0.requires_grad or r0.requires_grad or power.requires_grad)
    return _IntegrateParticles.apply(u_hat0, r0, power, step, interp, int(nsteps), int(every), grads)


def periodic_diff(a, b):
    """a - b wrapped to [-pi, pi) (positions live on the 2 pi torus)."""
    two_pi = 2 * torch.pi
    return torch.remainder(a - b + torch.pi, two_pi) - torch.pi
