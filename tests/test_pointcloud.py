"""Particle -> grid transfers on the GPU (qlles_sph_mls_grid, qlles_sph_dft_hat)
against the FP64 oracle (oracle/pointcloud.py, restating turbulence.py:145-377).

Tolerances: MLS 1e-9 of the largest value (Cholesky vs the reference's LU on
a 10x10 system), DFT 1e-10, spectra 1e-9 relative per shell.
"""

import numpy as np
import pytest
import torch

from oracle import pointcloud as pc
from oracle import sph as osph

L = 2 * np.pi
pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def cloud(n, dim, seed, jitter=0.3):
    rng = np.random.default_rng(seed)
    r = np.mod(osph.lattice(n, dim) + (rng.random((n ** dim, dim)) * 2 - 1) * jitter * L / n, L)
    x = r
    u = np.stack([np.sin(x[:, 0]) * np.cos(x[:, 1 % dim]), np.cos(2 * x[:, 0]) + 0.3 * np.sin(x[:, -1]),
                  np.sin(x[:, 1 % dim] + x[:, -1])][:dim], axis=1) + 0.05 * rng.normal(size=(n ** dim, dim))
    return r, u


@pytest.mark.parametrize("n,dim", [(16, 3), (8, 3), (32, 2), (12, 2)])
def test_mls_matches_oracle(gpu, n, dim):
    from l3es_amd import ur_to_u_mls_wrapper

    r, u = cloud(n, dim, n)
    got = host(ur_to_u_mls_wrapper(n, L, dim)(r, u))
    ref = pc.ur_to_u_mls(n, L, dim, r, u)
    assert np.max(np.abs(got - ref)) < 1e-9 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,dim,count", [(8, 3, None), (16, 3, None), (32, 2, None), (16, 2, 300)])
def test_dft_matches_oracle(gpu, n, dim, count):
    from l3es_amd import ur_to_u_dft_wrapper

    r, u = cloud(n, dim, 3 * n)
    if count is not None:  # any particle count
        r, u = r[:count], u[:count]
    got = host(ur_to_u_dft_wrapper(n, L, dim)(r, u))
    ref = pc.ur_to_u_dft(n, L, dim, r, u)
    assert np.max(np.abs(got - ref)) < 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("backend", ["mls", "dft"])
@pytest.mark.parametrize("n,dim", [(16, 3), (32, 2)])
def test_point_cloud_spectrum_matches_oracle(gpu, backend, n, dim):
    from l3es_amd import u_and_spectrum_from_ur

    r, u = cloud(n, dim, 7)
    ug, ek = u_and_spectrum_from_ur(n, L, dim, interp_backend=backend)(r, u)
    ug_ref, ek_ref = pc.u_and_spectrum_from_ur(n, L, dim, r, u, interp_backend=backend)
    assert tuple(ug.shape) == ug_ref.shape
    assert np.max(np.abs(host(ug) - ug_ref)) < 1e-9 * np.max(np.abs(ug_ref))
    big = ek_ref > 1e-12 * ek_ref.max()
    assert np.max(np.abs(ek[big] / ek_ref[big] - 1)) < 1e-9


def test_mls_requires_the_lattice_count(gpu):
    from l3es_amd import ur_to_u_mls_wrapper

    body = ur_to_u_mls_wrapper(8, L, 3)
    with pytest.raises(ValueError):
        body(np.zeros((10, 3)), np.zeros((10, 3)))
