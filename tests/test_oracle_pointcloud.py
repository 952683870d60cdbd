"""Pins of the particle -> grid oracle (oracle/pointcloud.py) by mathematics:
MLS of 2nd order reproduces quadratics; the DFT transfer is the identity on
the nodes and shift-equivariant."""

import numpy as np
import pytest

from oracle import pointcloud as pc
from oracle import sph

L = 2 * np.pi


@pytest.mark.parametrize("dim,N", [(2, 16), (3, 8)])
def test_mls_reproduces_quadratics(dim, N):
    rng = np.random.default_rng(dim)
    dx = L / N
    r = np.mod(sph.lattice(N, dim) + (rng.random((N ** dim, dim)) * 2 - 1) * 0.3 * dx, L)
    targets = pc.lattice_ij(N, dim, L)
    # a quadratic in the coordinates; only targets whose support (2 dx) stays off the seam
    c = rng.normal(size=(dim, dim))
    q = lambda x: 0.7 + 0.3 * x[:, 0] - 0.2 * x[:, -1] + np.einsum("ia,ab,ib->i", x, c, x)
    f = np.stack([q(r), 2.0 * q(r) + 1.0], axis=1)
    got = pc.mls_2nd_order(r, targets, f, L, dx, dim)
    inner = np.all((targets > 2.5 * dx) & (targets < L - 2.5 * dx), axis=1)
    exact = np.stack([q(targets), 2.0 * q(targets) + 1.0], axis=1)
    assert inner.sum() > 0
    assert np.max(np.abs(got[inner] - exact[inner])) < 1e-9 * np.max(np.abs(exact))

def test_mls_quadratic_check_catches_a_dropped_basis_term(monkeypatch):
    """The reproduction test above fails when the cross terms are left out of the basis."""
    dim, N = 2, 16
    rng = np.random.default_rng(9)
    dx = L / N
    r = np.mod(sph.lattice(N, dim) + (rng.random((N ** dim, dim)) * 2 - 1) * 0.3 * dx, L)
    targets = pc.lattice_ij(N, dim, L)
    f = 1.0 + r[:, 0] * r[:, 1]
    full = pc.mls_rows
    monkeypatch.setattr(pc, "mls_rows", lambda rji: full(rji)[:, :-1])
    got = pc.mls_2nd_order(r, targets, f, L, dx, dim)
    inner = np.all((targets > 2.5 * dx) & (targets < L - 2.5 * dx), axis=1)
    assert np.max(np.abs(got[inner] - (1.0 + targets[inner, 0] * targets[inner, 1]))) > 1e-3


@pytest.mark.parametrize("dim,N", [(2, 12), (3, 6)])
def test_dft_identity_on_nodes_and_shift(dim, N):
    rng = np.random.default_rng(5)
    dx = L / N
    g = np.meshgrid(*([np.arange(N)] * dim), indexing="ij")
    nodes = np.stack([a.ravel() for a in g], axis=1) * dx
    u = rng.normal(size=(N ** dim, dim))
    got = pc.ur_to_u_dft(N, L, dim, nodes, u)
    assert np.max(np.abs(got - u)) < 1e-12
    # particles moved one node along x: the field moves one node along x
    got2 = pc.ur_to_u_dft(N, L, dim, np.mod(nodes + np.eye(dim)[0] * dx, L), u)
    a = got.T.reshape((dim,) + (N,) * dim)
    b = got2.T.reshape((dim,) + (N,) * dim)
    assert np.max(np.abs(np.roll(a, 1, axis=1) - b)) < 1e-11
