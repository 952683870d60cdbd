"""Dataset statistics on the GPU (qlles_stats_moments, l3es_amd.dataset) against
the oracle (oracle/dataset.py, restating gen_dataset.py:143-238).

The per-element quantities (periodic displacement, frame differences) are FP32
operations in both, so they agree bit for bit; the FP64 sums differ only in
order.  Tolerance: 5e-14 of the sum of the magnitudes of the summed terms (a
summation-error bound; one term out of place moves a sum by ~1e-1 here, a
sum of magnitudes is at most ~1e6)."""

import json
import os

import numpy as np
import pytest
import torch

from oracle import dataset as ods

L = 2 * np.pi
pytestmark = pytest.mark.gpu


def trajectory(T, n, dim, seed, step=0.15):
    rng = np.random.default_rng(seed)
    r = np.mod(rng.random((1, n, dim)) * L + np.cumsum(rng.normal(size=(T, n, dim)) * step, 0), L)
    if n > 3:
        r[:, 0, 0] = np.linspace(L - 0.05, L + 0.3, T) % L  # crosses the seam
    u = rng.normal(size=(T, n, dim)) + 0.3
    return r.astype(np.float32), u.astype(np.float32)


def magnitudes(r, u, dim, center=None):
    """Per sum of the moment vector, the sum of |term| (the scale of its rounding error)."""
    qs = ods.trajectory_quantities(r, u, L)
    c = np.zeros((4, dim)) if center is None else np.asarray(center, np.float64).reshape(4, dim)
    a = [np.abs(q.astype(np.float64)).reshape(-1, dim).sum(0) for q in qs]
    b = [((q.astype(np.float64).reshape(-1, dim) - c[i]) ** 2).sum(0) for i, q in enumerate(qs)]
    return np.concatenate(a + b)


def close(got, ref, mag):
    return np.all(np.abs(got - ref) <= 5e-14 * mag + 1e-300)


def gpu_moments(r, u, dim, gpu, center=None):
    from l3es_amd import _lib

    lib = _lib.load()
    st = _lib.Stats(dim, L)
    rd, ud = torch.from_numpy(r).to(gpu), torch.from_numpy(u).to(gpu)
    out = torch.full((8 * dim,), float("nan"), dtype=torch.float64, device=gpu)
    c = torch.from_numpy(np.asarray(center, np.float64).reshape(-1)).to(gpu) if center is not None else None
    _lib.check(lib.qlles_stats_moments(st.handle, _lib.ptr(rd), _lib.ptr(ud), r.shape[0], r.shape[1], _lib.ptr(c),
                                       _lib.ptr(out), _lib.stream_ptr()))
    res = out.cpu().numpy()
    st.close()
    return res


@pytest.mark.parametrize("T,n,dim", [(1, 5, 3), (2, 100, 3), (3, 300, 2), (70, 1000, 3), (129, 257, 2),
                                     (200, 4096, 3)])
@pytest.mark.parametrize("centred", [False, True])
def test_moments_match_oracle(gpu, T, n, dim, centred):
    print(f"moments T={T} n={n} dim={dim} centred={centred}", flush=True)
    r, u = trajectory(T, n, dim, T * n + dim)
    center = np.random.default_rng(1).normal(size=(4, dim)) if centred else None
    got = gpu_moments(r, u, dim, gpu, center)
    ref = ods.trajectory_moments(r, u, L, center)
    mag = magnitudes(r, u, dim, center)
    assert close(got, ref, mag), np.max(np.abs(got - ref) / np.maximum(mag, 1e-300))


def test_moments_of_unwrapped_jumps(gpu):
    """Positions outside the box and frame-to-frame jumps of several box lengths: the
    displacement's general path (fmod), not only the in-box shortcut."""
    rng = np.random.default_rng(5)
    T, n = 40, 500
    r = (rng.random((1, n, 3)) * L + np.cumsum(rng.normal(size=(T, n, 3)) * 6.0, 0)).astype(np.float32)
    u = rng.normal(size=(T, n, 3)).astype(np.float32)
    got = gpu_moments(r, u, 3, gpu)
    ref = ods.trajectory_moments(r, u, L)
    assert close(got, ref, magnitudes(r, u, 3))


def test_moments_empty_and_reproducible(gpu):
    r, u = trajectory(50, 700, 3, 9)
    a = gpu_moments(r, u, 3, gpu)
    b = gpu_moments(r, u, 3, gpu)
    assert np.array_equal(a, b)
    z = gpu_moments(np.zeros((0, 10, 3), np.float32), np.zeros((0, 10, 3), np.float32), 3, gpu)
    assert np.array_equal(z, np.zeros(24))
    z = gpu_moments(np.zeros((6, 0, 2), np.float32), np.zeros((6, 0, 2), np.float32), 2, gpu)
    assert np.array_equal(z, np.zeros(16))


def test_context_reuse_grows_and_shrinks(gpu):
    """One context over a small, a large and a small trajectory (the partials
    buffer grows once and is reused)."""
    from l3es_amd import _lib

    lib = _lib.load()
    st = _lib.Stats(3, L)
    out = torch.empty(24, dtype=torch.float64, device=gpu)
    for T, n in ((5, 64), (300, 3000), (7, 100)):
        r, u = trajectory(T, n, 3, T + n)
        rd, ud = torch.from_numpy(r).to(gpu), torch.from_numpy(u).to(gpu)
        _lib.check(lib.qlles_stats_moments(st.handle, _lib.ptr(rd), _lib.ptr(ud), T, n, None, _lib.ptr(out),
                                           _lib.stream_ptr()))
        ref = ods.trajectory_moments(r, u, L)
        assert close(out.cpu().numpy(), ref, magnitudes(r, u, 3))
    st.close()


CFG = """sim:
  case: kolm
  dim: {dim}
  ckp_N: {n}
  dt: 0.005
  t_final: 1.0
  nu: 0.001
  u_ref: 6.0
com:
  ckp_freq: 1
"""


@pytest.mark.parametrize("dim,n,every", [(3, 8, 1), (2, 16, 2)])
@pytest.mark.parametrize("keep_resident", [True, False])
def test_compute_statistics_matches_oracle(gpu, tmp_path, dim, n, every, keep_resident):
    """Frames -> single_frames_to_dataset -> compute_statistics against the
    oracle's statistics over the same train trajectories."""
    from l3es_amd.dataset import compute_statistics, single_frames_to_dataset

    src, dst = tmp_path / "raw", tmp_path / "ds"
    trajs = {}
    for s in (4, 1, 9, 3, 5, 12, 7, 2):
        T = 23
        r, u = trajectory(T, n ** dim, dim, 100 + s)
        com = src / f"traj_{s}" / "com"
        os.makedirs(com)
        for t in range(T):
            np.savez(com / f"step_{t:05d}.npz", r=r[t].astype(np.float64), u=u[t].astype(np.float64))
        (src / f"traj_{s}" / "config.yaml").write_text(CFG.format(dim=dim, n=n))
        trajs[s] = (r, u)
    single_frames_to_dataset(str(src), str(dst), split="2_1_1")
    meta = compute_statistics(str(dst), stats_every_nth=every, device=gpu, keep_resident=keep_resident)
    train = [trajs[s] for s in (1, 2, 3, 4)]  # seeds in order; valid and test take ceil(8/4) = 2 each
    ref = ods.statistics([(r[::every], u[::every]) for r, u in train], L)
    for k, v in ref.items():
        assert np.allclose(meta[k], v, rtol=1e-12, atol=1e-14), k
    with open(dst / f"metadata_every{every}.json") as fh:
        disk = json.load(fh)
    assert disk["write_every"] == every and disk["num_trajs_train"] == 4
    for k in ref:
        assert disk[k] == meta[k]
