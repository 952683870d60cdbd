"""qlles_stats_moments_noalloc: the trajectory moments without a context and without allocating (no
counterpart in the reference), against qlles_stats_moments on the same inputs.  Both entry points run
the same kernels in the same launch shape with fixed-order sums, so the results are bitwise equal.
Two calls in a row, the second on a smaller trajectory, check that the shared scratch array's earlier
contents leak into nothing; a trajectory too large for the scratch array is refused on the host."""

import numpy as np
import pytest
import torch

from oracle import dataset as ods

L = 2 * np.pi
pytestmark = pytest.mark.gpu


def trajectory(T, n, dim, seed, step=0.15):
    rng = np.random.default_rng(seed)
    r = np.mod(rng.random((1, n, dim)) * L + np.cumsum(rng.normal(size=(T, n, dim)) * step, 0), L)
    u = rng.normal(size=(T, n, dim)) + 0.3
    return r.astype(np.float32), u.astype(np.float32)


def both(lib, _lib, r, u, dim, dev, center=None):
    rd, ud = torch.from_numpy(r).to(dev), torch.from_numpy(u).to(dev)
    c = None if center is None else torch.from_numpy(np.asarray(center, np.float64).reshape(-1)).to(dev)
    st = _lib.Stats(dim, L)
    ref = torch.full((8 * dim,), float("nan"), dtype=torch.float64, device=dev)
    _lib.check(lib.qlles_stats_moments(st.handle, _lib.ptr(rd), _lib.ptr(ud), r.shape[0], r.shape[1], _lib.ptr(c),
                                       _lib.ptr(ref), _lib.stream_ptr()))
    got = torch.full((8 * dim,), float("nan"), dtype=torch.float64, device=dev)
    _lib.check(lib.qlles_stats_moments_noalloc(dim, L, _lib.ptr(rd), _lib.ptr(ud), r.shape[0], r.shape[1],
                                               _lib.ptr(c), _lib.ptr(got), _lib.stream_ptr()))
    torch.cuda.synchronize()
    st.close()
    return got.cpu().numpy(), ref.cpu().numpy()


@pytest.mark.parametrize("dim,sizes", [(3, [(300, 4096), (40, 700)]), (2, [(129, 2000), (3, 257)])])
def test_noalloc_moments_equal_context_moments(gpu, dim, sizes):
    from l3es_amd import _lib

    lib = _lib.load()
    for k, (T, n) in enumerate(sizes):  # the second call is smaller than the first
        print(f"[no-alloc] dim {dim}, call {k}: T={T}, np={n}", flush=True)
        r, u = trajectory(T, n, dim, seed=10 * dim + k)
        center = np.linspace(-0.2, 0.4, 4 * dim) if k else None
        got, ref = both(lib, _lib, r, u, dim, gpu, center)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(got, ref)
        # and the oracle's sums (FP64, another order) to summation error
        want = ods.trajectory_moments(r, u, L, center)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-9)


def test_noalloc_moments_empty_and_refused(gpu):
    from l3es_amd import _lib

    lib = _lib.load()
    out = torch.full((24,), float("nan"), dtype=torch.float64, device=gpu)
    _lib.check(lib.qlles_stats_moments_noalloc(3, L, None, None, 0, 5, None, _lib.ptr(out), _lib.stream_ptr()))
    torch.cuda.synchronize()
    assert torch.all(out == 0)
    # 2^20 doubles of partials: 24 per workgroup in 3-D, so 1e6 frames x 32^3 particles is refused before any
    # launch (the pointers are never read)
    r = torch.zeros(3, dtype=torch.float32, device=gpu)
    rc = lib.qlles_stats_moments_noalloc(3, L, _lib.ptr(r), _lib.ptr(r), 10 ** 6, 32 ** 3, None, _lib.ptr(out),
                                         _lib.stream_ptr())
    assert rc == _lib.QLLES_ERR_ARG
    assert b"scratch" in lib.qlles_last_error()
    assert lib.qlles_stats_moments_noalloc(4, L, None, None, 0, 0, None, _lib.ptr(out), None) == _lib.QLLES_ERR_SHAPE
    assert lib.qlles_stats_moments_noalloc(3, -1.0, None, None, 0, 0, None, _lib.ptr(out), None) == _lib.QLLES_ERR_ARG
