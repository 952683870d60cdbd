"""The package's host-side initial conditions against the oracle (CPU only).

The 2-D random-phase spectrum and the jittered particle lattice are NumPy on
the host in the package as in the oracle; the oracle's versions are pinned by
invariants in test_oracle.py (divergence, spectral peak, seeding, bounds).
"""

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("N,seed", [(32, 0), (64, 3), (256, 7)])
def test_kolm_spectrum_matches_oracle(N, seed):
    from l3es_amd.init_fields import kolm_vorticity_spectrum

    uh, vh = kolm_vorticity_spectrum(N, seed)
    assert uh.shape == (N, N // 2 + 1)
    u = np.fft.irfftn(uh, s=(N, N), axes=(0, 1))
    v = np.fft.irfftn(vh, s=(N, N), axes=(0, 1))
    s = 4.2 / max(np.abs(u).max(), np.abs(v).max())
    ref = O.init_u_kolm_np(N, seed=seed)
    np.testing.assert_allclose(u * s, ref[0, :, :, 0], rtol=0, atol=1e-13)
    np.testing.assert_allclose(v * s, ref[1, :, :, 0], rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,dim,jitter,seed", [(64, 2, 0.25, 5), (16, 3, 0.25, 1), (8, 2, 0.75, 2)])
def test_pos_init_jittered_matches_oracle(n, dim, jitter, seed):
    from l3es_amd.integrator import pos_init_cartesian, pos_init_jittered

    L = 2 * np.pi
    np.testing.assert_array_equal(pos_init_jittered(n, dim, L, jitter, seed), O.pos_init_jittered(n, dim, L, jitter, seed))
    np.testing.assert_array_equal(pos_init_cartesian(n, dim, L), O.pos_init_cartesian(n, dim, L))
