"""Batched trajectories (qlles_batch_*): every member bit-identical to its solo run.

No reference counterpart: the reference runs one trajectory per process (slurm_*.sh:5).  The pin is
exact equality with the same steps taken alone through qlles_step_advect / qlles_step.
"""

import ctypes

import numpy as np
import pytest
import torch



def has_lib():
    from l3es_amd import _lib

    import os
    return os.path.exists(_lib.LIB_PATH)


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.skipif(not has_lib(), reason="libqlles.so not built")
def test_batch_validation_without_device():
    from l3es_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.qlles_batch_create(ctypes.byref(h)) == 0
    try:
        assert lib.qlles_batch_add(h, None, None, None, None, 0, None, None) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_configure(h, 0, 1, 0) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_configure(h, 2, 0, 0) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_configure(h, 2, 4, 3) == 0
        assert lib.qlles_batch_configure(h, 2, 4, 4) == 0
        assert lib.qlles_batch_configure(h, 2, 4, 5) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_configure(h, 2, 4, -1) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_configure(h, 2, 4, 1) == 0
        lanes, chunk, single, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        assert lib.qlles_batch_get_config(h, ctypes.byref(lanes), ctypes.byref(chunk), ctypes.byref(single),
                                          ctypes.byref(us)) == 0
        assert (lanes.value, chunk.value, single.value) == (2, 4, 1)
        assert lib.qlles_batch_run(h, None, -3, None) == _lib.QLLES_ERR_ARG
        assert lib.qlles_batch_run(h, None, 5, None) == 0  # empty batch: nothing to do
    finally:
        assert lib.qlles_batch_destroy(h) == 0


def _members():
    from l3es_amd import set_up_solver
    from l3es_amd.integrator import pos_init_cartesian, spectral_interpolator_wrapper

    out = []
    specs = [
        # (N, dim, case, precision, kwargs, particle backend, n)
        (32, 3, "HIT", "float32", dict(forcing_type="power", power=0.1, kf=2), "nufft", 16),
        (16, 3, "TGV3D", "float64", dict(forcing_type="none", kf=None), None, 0),
        (64, 2, "Kolm", "float32", dict(forcing_type="kolmogorov", drag=0.1, kolm_amp=1.0, kolm_k=4, kf=None),
         "nufft", 32),
        (32, 2, "TGV2D", "float64", dict(forcing_type="none", kf=None), "dft", 16),
    ]
    for i, (N, dim, case, prec, kw, be, n) in enumerate(specs):
        kf = kw.pop("kf")
        _, uh, _, _, step, _, _, _, _ = set_up_solver(N, 2e-3, dim, case, 2e-3, 10 + i, max(n, 8), kf,
                                                      precision=prec, warmup=False, **kw)
        it = r = None
        if be is not None:
            it = spectral_interpolator_wrapper(n, tuple(range(1, dim + 1)), backend=be, precision=prec)
            r = torch.from_numpy(pos_init_cartesian(n, dim)).cuda()
        out.append((step, uh, it, r))
    return out


def _solo(step, uh, it, r, nsteps):
    from l3es_amd import _lib

    lib = _lib.load()
    if it is None:
        step.advance(None, uh, nsteps, None, use_u=False)
        return None
    ur = torch.empty((r.shape[0], it.dim), dtype=it.dtype, device=r.device)
    _lib.check(lib.qlles_step_advect(step.plan.handle, it.it.handle, None, _lib.ptr(uh), int(nsteps), _lib.ptr(r),
                                     r.shape[0], _lib.ptr(ur), None, _lib.stream_ptr()))
    return ur


@pytest.mark.gpu
def test_batch_bit_identical_to_solo(gpu):
    from l3es_amd import Batch

    mem = _members()
    solo = [(uh.clone(), None if r is None else r.clone()) for (_, uh, _, r) in mem]
    bat = [(uh.clone(), None if r is None else r.clone()) for (_, uh, _, r) in mem]
    b = Batch()
    urs = []
    for (step, _, it, _), (uh, r) in zip(mem, bat):
        ur = None if it is None else torch.empty((r.shape[0], it.dim), dtype=it.dtype, device=r.device)
        urs.append(ur)
        b.add(step, uh, it, r, ur)
    lengths = [7, 5, 9, 3]  # uneven trajectories; chunk 2 leaves eager remainders
    b.configure(lanes=2, chunk=2)
    b.run(lengths)
    b.configure(lanes=3, chunk=2, single_graph=True)  # one joint graph for the common chunk
    b.run([4, 4, 6, 4])
    b.configure(lanes=2, chunk=1, single_graph="whole")  # one graph of every member's whole run
    b.run([3, 1, 2, 5])
    b.run([3, 1, 2, 5])  # replayed
    b.run([1, 2, 0, 1])  # captured again
    cfg = b.tune(trial_steps=4)  # runs on scratch copies: states must not move
    # lane counts the tuner tries: 1, 2, 4, 8, 16, and 3, 6 for the grouped modes
    assert cfg["lanes"] in (1, 2, 3, 4, 6, 8, 16) and cfg["chunk"] in (1, 4, 16, 64) and cfg["us_per_traj_step"] > 0
    assert cfg["lanes"] not in (3, 6) or cfg["single_graph"] in ("grouped", "grouped-whole")
    assert cfg["single_graph"] in (False, True, "grouped", "grouped-whole")
    b.run(5)
    torch.cuda.synchronize()
    for k, ((step, _, it, _), (uh_s, r_s), (uh_b, r_b), ur_b) in enumerate(zip(mem, solo, bat, urs)):
        ur1 = _solo(step, uh_s, it, r_s, lengths[k])
        ur1 = _solo(step, uh_s, it, r_s, [4, 4, 6, 4][k])
        for n in ([3, 1, 2, 5][k], [3, 1, 2, 5][k], [1, 2, 0, 1][k]):
            if n:
                ur1 = _solo(step, uh_s, it, r_s, n)
        ur2 = _solo(step, uh_s, it, r_s, 5)
        torch.cuda.synchronize()
        assert np.array_equal(host(uh_s), host(uh_b)), f"member {k}: spectrum differs from solo run"
        if it is not None:
            assert np.array_equal(host(r_s), host(r_b)), f"member {k}: positions differ"
            assert np.array_equal(host(ur2), host(ur_b)), f"member {k}: particle velocities differ"
        del ur1


@pytest.mark.gpu
@pytest.mark.parametrize("dim,N,case,kw", [
    (3, 32, "HIT", dict(forcing_type="power", power=0.2, kf=2)),
    (2, 64, "Kolm", dict(forcing_type="kolmogorov", drag=0.1, kolm_amp=1.0, kolm_k=4, kf=None)),
])
def test_plan_variants_match_oracle(gpu, dim, N, case, kw):
    """Every kernel variant (tile width x radix plan) takes the same FP64 steps as variant (0, 0), which the
    parity tests pin to the oracle; then the per-plan tuner picks one of them."""
    from l3es_amd import set_up_solver

    kw = dict(kw)
    kf = kw.pop("kf")
    _, uh, _, _, step, _, _, _, _ = set_up_solver(N, 2e-3, dim, case, 2e-3, 21, 8, kf, precision="float64",
                                                  warmup=False, **kw)
    _, _, nv = step.plan.variant()  # set_up_solver has already tuned the plan
    assert nv == 2
    u0 = host(uh)
    ref = None
    for xv, zv in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        if True:
            step.plan.set_variant(xv, zv)
            y = torch.from_numpy(u0).cuda()
            step.advance(None, y, 3, None, use_u=False)
            out = host(y)
            if ref is None:
                ref = out
            err = np.abs(out - ref).max() / np.abs(ref).max()
            assert err < 1e-12, (xv, zv, err)
    best = step.plan.tune(torch.from_numpy(u0).cuda(), trial_steps=2)
    assert best > 0 and step.plan.variant()[:2] in [(a, b) for a in range(nv) for b in range(nv)]


@pytest.mark.gpu
def test_variant_cache_roundtrip(gpu, tmp_path):
    """custom_jit_integrate (spectral_solver.py:130-153): rejit tunes and stores the kernel variants,
    rejit=False loads them into a fresh plan."""
    import json

    from l3es_amd.spectral_solver import custom_jit_integrate, set_up_solver

    _, uh, _, _, step, _, _, _, _ = set_up_solver(32, 2e-3, 3, "HIT", 2e-3, 4, 8, None, forcing_type="none",
                                                  precision="float32", warmup=False)
    pref = str(tmp_path / "spectral_gpu_32_3_f32")
    custom_jit_integrate(pref, True, step, None, uh)
    cfg = json.load(open(pref + ".json"))
    assert (cfg["xv"], cfg["zv"]) == step.plan.variant()[:2] and cfg["us_per_step"] > 0
    _, uh2, _, _, step2, _, _, _, _ = set_up_solver(32, 2e-3, 3, "HIT", 2e-3, 4, 8, None, forcing_type="none",
                                                    precision="float32", warmup=False)
    json.dump(dict(cfg, xv=1 - cfg["xv"]), open(pref + ".json", "w"))
    custom_jit_integrate(pref, False, step2, None, uh2)
    assert step2.plan.variant()[0] == 1 - cfg["xv"]


def _grouped_members():
    """Members of a few kinds, several of each: 3-D FP32 power-forced HIT at 32^3 with NUFFT particles
    (different seeds, viscosities, time steps and powers), 2-D FP32 Kolmogorov at 64^2 with NUFFT
    particles, 3-D FP64 Taylor-Green at 16^3 without particles, 2-D FP64 Taylor-Green at 32^2 with DFT
    particles (advected member by member), a 3-D member held at another kernel variant (a group of its
    own) and an ekin_tot-forced member (runs alone)."""
    from l3es_amd import set_up_solver
    from l3es_amd.integrator import pos_init_jittered, spectral_interpolator_wrapper

    specs = [
        # (N, dim, case, precision, nu, dt, kwargs, particle backend, n, variant)
        (32, 3, "HIT", "float32", 2e-3, 2e-3, dict(forcing_type="power", power=0.1, kf=2), "nufft", 16, (0, 0)),
        (64, 2, "Kolm", "float32", 1e-3, 1e-3,
         dict(forcing_type="kolmogorov", drag=0.1, kolm_amp=1.0, kolm_k=4, kf=None), "nufft", 32, (0, 0)),
        (32, 3, "HIT", "float32", 1e-3, 1e-3, dict(forcing_type="power", power=0.3, kf=2), "nufft", 16, (0, 0)),
        (16, 3, "TGV3D", "float64", 2e-3, 2e-3, dict(forcing_type="none", kf=None), None, 0, (0, 0)),
        (32, 3, "HIT", "float32", 3e-3, 2e-3, dict(forcing_type="power", power=0.2, kf=1), "nufft", 16, (0, 0)),
        (64, 2, "Kolm", "float32", 2e-3, 1e-3,
         dict(forcing_type="kolmogorov", drag=0.05, kolm_amp=0.5, kolm_k=2, kf=None), "nufft", 32, (0, 0)),
        (16, 3, "TGV3D", "float64", 1e-3, 1e-3, dict(forcing_type="none", kf=None), None, 0, (0, 0)),
        (32, 2, "TGV2D", "float64", 2e-3, 2e-3, dict(forcing_type="none", kf=None), "dft", 16, (0, 0)),
        (32, 2, "TGV2D", "float64", 1e-3, 2e-3, dict(forcing_type="none", kf=None), "dft", 16, (0, 0)),
        (32, 3, "HIT", "float32", 2e-3, 2e-3, dict(forcing_type="power", power=0.1, kf=2), "nufft", 16, (1, 1)),
        (16, 3, "HIT", "float64", 2e-3, 2e-3, dict(forcing_type="ekin_tot", e_kin_target=0.5, kf=2), "nufft", 8,
         (0, 0)),
    ]
    out = []
    for i, (N, dim, case, prec, nu, dt, kw, be, n, var) in enumerate(specs):
        kw = dict(kw)
        kf = kw.pop("kf")
        _, uh, _, _, step, _, _, _, _ = set_up_solver(N, nu, dim, case, dt, 30 + i, max(n, 8), kf, precision=prec,
                                                      warmup=False, **kw)
        step.plan.set_variant(*var)
        it = r = None
        if be is not None:
            it = spectral_interpolator_wrapper(n, tuple(range(1, dim + 1)), backend=be, precision=prec)
            r = torch.from_numpy(pos_init_jittered(n, dim, seed=i)).cuda()
        out.append((step, uh, it, r))
    return out


@pytest.mark.gpu
def test_batch_grouped_bit_identical_to_solo(gpu):
    """Grouped modes: multi-member launches per group of one kind (eager, and the whole run as one graph),
    with uneven lengths (members finishing at different steps, zero-step members), group tables remade when
    the order of the step counts changes, and a tuner run in between; every member equals its solo run."""
    from l3es_amd import Batch

    mem = _grouped_members()
    M = len(mem)
    solo = [(uh.clone(), None if r is None else r.clone()) for (_, uh, _, r) in mem]
    bat = [(uh.clone(), None if r is None else r.clone()) for (_, uh, _, r) in mem]
    b = Batch()
    urs = []
    for (step, _, it, _), (uh, r) in zip(mem, bat):
        ur = None if it is None else torch.empty((r.shape[0], it.dim), dtype=it.dtype, device=r.device)
        urs.append(ur)
        b.add(step, uh, it, r, ur)
    # kinds: 3-D FP32 HIT 32^3 (0, 2, 4), 2-D FP32 Kolmogorov 64^2 (1, 5), 3-D FP64 TGV 16^3 (3, 6), 2-D FP64
    # TGV 32^2 with DFT particles (7, 8), the member at another kernel variant (9), the ekin_tot member (10)
    kind_of, us = b.kinds()
    assert kind_of == [0, 1, 0, 2, 0, 1, 2, 3, 3, 4, 5] and us == [0.0] * 6
    runs = []

    def run(cfg, lengths, repeat=1):
        if cfg is not None:
            b.configure(**cfg)
        for _ in range(repeat):
            b.run(lengths)
            runs.append(lengths)

    run(dict(lanes=2, chunk=2, single_graph="grouped"), [5, 3, 7, 4, 2, 6, 0, 3, 5, 4, 3])
    run(None, [2, 4, 1, 3, 5, 2, 2, 1, 0, 2, 2])  # another order within the groups: tables remade
    run(dict(lanes=3, chunk=1, single_graph="grouped-whole"), [4, 2, 3, 1, 4, 5, 2, 3, 3, 1, 2], repeat=2)
    run(None, [1, 3, 0, 2, 2, 1, 3, 0, 1, 2, 1])  # captured again
    run(dict(lanes=16, chunk=4, single_graph="grouped"), [6] * M)
    cfg = b.tune(trial_steps=4)  # scratch copies: the states do not move
    assert cfg["us_per_traj_step"] > 0
    kind_of, us = b.kinds()  # the tuner may have moved members to other kernel variants, hence other kinds
    assert len(kind_of) == M and len(us) == max(kind_of) + 1 and all(v > 0 for v in us)
    b.configure(lanes=4, chunk=1, single_graph="grouped-whole")
    run(None, [3, 1, 2, 2, 1, 3, 1, 2, 2, 1, 1])
    torch.cuda.synchronize()
    for k, ((step, _, it, _), (uh_s, r_s), (uh_b, r_b), ur_b) in enumerate(zip(mem, solo, bat, urs)):
        ur = None
        for lengths in runs:
            if lengths[k]:
                ur = _solo(step, uh_s, it, r_s, lengths[k])
        torch.cuda.synchronize()
        assert np.array_equal(host(uh_s), host(uh_b)), f"member {k}: spectrum differs from solo run"
        if it is not None:
            assert np.array_equal(host(r_s), host(r_b)), f"member {k}: positions differ"
            assert np.array_equal(host(ur), host(ur_b)), f"member {k}: particle velocities differ"
