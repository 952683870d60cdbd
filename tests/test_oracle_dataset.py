"""Pins of the dataset packing / statistics oracle (oracle/dataset.py) and the
host-side packing of l3es_amd.dataset, on the CPU.

The reference holds no fixture for this path (gen_dataset.py runs on the
user's frames), so the pins are the reference's formulas evaluated by hand on
chosen counts, brute force, and closed forms of trajectories whose statistics
are known exactly."""

import ctypes
import json
import os

import numpy as np
import pytest

from oracle import dataset as ods

L = 2 * np.pi


@pytest.mark.parametrize("n,split,bounds", [
    (4, "2_1_1", [0, 2, 3, 4]),      # ceil(1/4 * 4) = 1
    (7, "2_1_1", [0, 3, 5, 7]),      # ceil(7/4) = 2
    (10, "80_10_10", [0, 8, 9, 10]),  # ceil(0.1 * 10) = 1
    (5, "8_1_1", [0, 3, 4, 5]),       # ceil(0.5) = 1
    (12, "1_1_5", [0, 8, 10, 12]),    # the third ratio is not used: ceil(12/7) = 2
])
def test_split_bounds(n, split, bounds):
    got, n_train, n_eval = ods.split_bounds(n, split)
    assert got == bounds
    assert n_train == bounds[1] and n_eval == bounds[2] - bounds[1] == bounds[3] - bounds[2]


@pytest.mark.parametrize("files,skip,every,expect", [(101, 0, 1, 101), (101, 1, 2, 50), (10, 3, 4, 2), (8, 0, 3, 3)])
def test_sequence_length(files, skip, every, expect):
    assert ods.sequence_length(files, skip, every) == expect


@pytest.mark.parametrize("n,expect", [(8, 1.2), (16, 0.59), (32, 0.29), (64, 0.15)])
def test_connectivity_radius(n, expect):
    """1.5 * 2 pi / n to two significant digits: 1.178, 0.589, 0.2945, 0.147."""
    assert ods.connectivity_radius(2 * np.pi / n) == expect


def test_periodic_displacement_is_the_minimum_image():
    rng = np.random.default_rng(0)
    ra = rng.random((2000, 3)) * L
    rb = np.mod(ra + rng.normal(size=ra.shape) * 0.3, L)
    ra[:5] = [[0.01, L - 0.02, 3.0]] * 5  # on the seam
    rb[:5] = [[L - 0.01, 0.02, 3.0]] * 5
    d = ods.periodic_displacement_f32(ra, rb, L).astype(np.float64)
    ra32, rb32 = ra.astype(np.float32).astype(np.float64), rb.astype(np.float32).astype(np.float64)
    shifts = np.array([-1.0, 0.0, 1.0]) * np.float64(np.float32(L))
    cand = (ra32 - rb32)[..., None] + shifts
    brute = np.take_along_axis(cand, np.argmin(np.abs(cand), -1)[..., None], -1)[..., 0]
    assert np.max(np.abs(d - brute)) < 4 * np.spacing(np.float32(L))
    assert np.allclose(d[:5], [[0.02, -0.04, 0.0]] * 5, atol=1e-6)


def test_uniform_translation_across_the_seam():
    """r_t = r0 + t U dt (mod L), u_t = U: vel = U dt, acc = 0, au = 0.  The
    std of vel is its root-mean-square about zero (the reference's centring),
    the stds of u and au are 0 and floored to 1."""
    rng = np.random.default_rng(1)
    T, n, U, dt = 40, 300, np.array([0.7, -1.3, 0.4]), 0.05
    r0 = rng.random((n, 3)) * L
    r = np.stack([np.mod(r0 + t * U * dt, L) for t in range(T)]).astype(np.float32)
    u = np.broadcast_to(U, (T, n, 3)).astype(np.float32)
    s = ods.statistics([(r, u)], L)
    assert np.allclose(s["vel_mean"], U * dt, atol=2e-6)
    assert np.allclose(s["vel_std"], np.abs(U * dt), atol=2e-6)  # RMS about zero, not ~0
    assert np.max(np.abs(s["acc_mean"])) < 2e-6
    assert np.allclose(s["u_mean"], U.astype(np.float32))
    assert np.array_equal(s["u_std"], np.ones(3)) and np.array_equal(s["au_std"], np.ones(3))
    assert np.array_equal(s["au_mean"], np.zeros(3))


def test_constant_acceleration_track_and_centring():
    """r_t = r0 + v0 t + a t^2 / 2 with u_t = c + t: acc = a, au = 1; u's std
    is about its own mean (t uniform on 0..T-1 has std sqrt((T^2-1)/12))."""
    rng = np.random.default_rng(2)
    T, n = 21, 200
    v0, a = np.array([0.03, -0.02]), np.array([2e-3, -1e-3])
    r0 = rng.random((n, 2)) * L
    t = np.arange(T)[:, None, None]
    r = np.mod(r0 + v0 * t + 0.5 * a * t * t, L).astype(np.float32)
    u = (5.0 + t + np.zeros((T, n, 2))).astype(np.float32)
    s = ods.statistics([(r, u)], L)
    assert np.allclose(s["acc_mean"], a, atol=2e-5)
    assert np.allclose(s["vel_mean"], v0 + a * (T - 1) / 2, atol=2e-5)  # v0 + a (t + 1/2), t = 0..T-2
    assert np.allclose(s["u_mean"], 5.0 + (T - 1) / 2)
    assert np.allclose(s["u_std"], np.sqrt((T * T - 1) / 12.0))
    assert np.allclose(s["au_mean"], 1.0) and np.array_equal(s["au_std"], np.ones(2))


def test_moments_and_accumulation():
    """trajectory_moments / counts are the per-trajectory means of statistics;
    the reference's FP32 accumulation agrees with FP64 to FP32 rounding."""
    rng = np.random.default_rng(3)
    trajs = []
    for T in (9, 14):
        r = np.mod(rng.random((1, 64, 3)) * L + np.cumsum(rng.normal(size=(T, 64, 3)) * 0.1, 0), L)
        trajs.append((r.astype(np.float32), rng.normal(size=(T, 64, 3)).astype(np.float32)))
    s64 = ods.statistics(trajs, L)
    s32 = ods.statistics(trajs, L, accumulate=np.float32)
    for k in s64:
        assert np.allclose(s64[k], s32[k], rtol=1e-5, atol=1e-6), k
    r, u = trajs[0]
    m = ods.trajectory_moments(r, u, L).reshape(8, 3)
    counts = np.array([8, 7, 9, 8])[:, None] * 64
    vel, acc, uu, au = ods.trajectory_quantities(r, u, L)
    assert np.allclose(m[:4] / counts, [q.astype(np.float64).mean((0, 1)) for q in (vel, acc, uu, au)], rtol=1e-14)
    c = rng.normal(size=(4, 3))
    m2 = ods.trajectory_moments(r, u, L, c).reshape(8, 3)
    assert np.allclose(m2[:4], m[:4])
    for i, q in enumerate((vel, acc, uu, au)):
        assert np.allclose(m2[4 + i], ((q.astype(np.float64) - c[i]) ** 2).sum((0, 1)), rtol=1e-13)


def write_frames(root, seeds, T, n, dim, cfg):
    rng = np.random.default_rng(7)
    truth = {}
    for s in seeds:
        com = os.path.join(root, f"traj_{s}", "com")
        os.makedirs(com)
        r = rng.random((T, n, dim)) * L
        u = rng.normal(size=(T, n, dim))
        for t in range(T):
            np.savez(os.path.join(com, f"step_{t:05d}.npz"), r=r[t], u=u[t])
        with open(os.path.join(root, f"traj_{s}", "config.yaml"), "w") as fh:
            fh.write(cfg)
        truth[s] = (r, u)
    return truth


CFG = """sim:
  case: hit
  dim: 3
  ckp_N: 4
  dt: 0.002
  t_final: 1.0
  nu: 0.002
  u_ref: 4.0
com:
  ckp_freq: 1
"""


def test_packing_round_trip(tmp_path):
    """single_frames_to_dataset against the reference's layout: seed order
    (2 < 10 < 11 < 30, not lexical), split sizes, frame slicing, FP32 arrays,
    zero particle types and the metadata keys."""
    from l3es_amd.dataset import single_frames_to_dataset

    src, dst = tmp_path / "raw", tmp_path / "ds"
    truth = write_frames(str(src), [10, 2, 30, 11], T=7, n=64, dim=3, cfg=CFG)
    meta = single_frames_to_dataset(str(src), str(dst), split="2_1_1", skip_first_n_frames=1,
                                    slice_every_nth_frame=2)
    order = [2, 10, 11, 30]
    with np.load(dst / "train.npz") as z:
        assert sorted(z.files) == ["00000/particle_type", "00000/position", "00000/u",
                                   "00001/particle_type", "00001/position", "00001/u"]
        for j, s in enumerate(order[:2]):
            pos = z[f"{j:05d}/position"]
            assert pos.dtype == np.float32 and pos.shape == (3, 64, 3)
            assert np.array_equal(pos, truth[s][0][1::2].astype(np.float32))
            assert np.array_equal(z[f"{j:05d}/u"], truth[s][1][1::2].astype(np.float32))
            assert np.array_equal(z[f"{j:05d}/particle_type"], np.zeros(64))
    for name, s in (("valid", order[2]), ("test", order[3])):
        with np.load(dst / f"{name}.npz") as z:
            assert np.array_equal(z["00000/position"], truth[s][0][1::2].astype(np.float32))
    with open(dst / "metadata.json") as fh:
        disk = json.load(fh)
    assert disk == json.loads(json.dumps(meta))
    assert meta["case"] == "HIT" and meta["dim"] == 3 and meta["num_particles_max"] == 64
    assert meta["sequence_length_train"] == 3 and meta["num_trajs_train"] == 2 and meta["num_trajs_test"] == 1
    assert meta["write_every"] == 2 and meta["default_connectivity_radius"] == 2.4  # 1.5 * pi / 2 = 2.356
    assert meta["periodic_boundary_conditions"] == [True] * 3 and len(meta["bounds"]) == 3


def test_stats_abi_rejects_bad_arguments():
    from l3es_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqlles.so not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.qlles_stats_create(4, L, ctypes.byref(h)) == _lib.QLLES_ERR_SHAPE and not h.value
    assert lib.qlles_stats_create(3, -1.0, ctypes.byref(h)) == _lib.QLLES_ERR_ARG and not h.value
    assert lib.qlles_stats_create(3, L, ctypes.byref(h)) == _lib.QLLES_OK and h.value
    try:
        assert lib.qlles_stats_moments(h, None, None, 5, 10, None, None, None) == _lib.QLLES_ERR_ARG
        assert lib.qlles_stats_moments(h, None, None, -1, 10, None, ctypes.c_void_p(8), None) == _lib.QLLES_ERR_ARG
        assert lib.qlles_stats_moments(h, None, None, 5, 10, None, ctypes.c_void_p(8), None) == _lib.QLLES_ERR_ARG
        assert lib.qlles_last_error().decode()
    finally:
        assert lib.qlles_stats_destroy(h) == _lib.QLLES_OK


def test_moment_sums_are_exactly_rounded_enough():
    """The oracle's FP64 sums against math.fsum (exactly rounded): a plain
    numpy sum over the time/particle axes is accurate only to ~n eps, which
    would hide or fake differences of the GPU's sums."""
    import math

    rng = np.random.default_rng(11)
    r = np.mod(rng.random((1, 1000, 3)) * L + np.cumsum(rng.normal(size=(70, 1000, 3)) * 0.15, 0), L)
    r, u = r.astype(np.float32), (rng.normal(size=(70, 1000, 3)) + 0.3).astype(np.float32)
    m = ods.trajectory_moments(r, u, L).reshape(8, 3)
    for i, q in enumerate(ods.trajectory_quantities(r, u, L)):
        x = q.astype(np.float64).reshape(-1, 3)
        for c in range(3):
            assert abs(m[i, c] - math.fsum(x[:, c])) <= 1e-14 * math.fsum(np.abs(x[:, c]))
            ex = math.fsum(x[:, c] ** 2)
            assert abs(m[4 + i, c] - ex) <= 1e-14 * ex
