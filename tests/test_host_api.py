"""Argument checks of the host API that need no GPU: shapes the library
cannot read are refused with ValueError before any library call."""

import pytest
import torch


def test_physical_field_shapes():
    from l3es_amd.utils import physical_field

    assert physical_field(torch.zeros(3, 8, 8, 8))[:2] == (3, 8)
    assert physical_field(torch.zeros(3, 8, 8, 1))[:2] == (2, 8)
    dim, N, full = physical_field(torch.ones(2, 8, 8))
    assert (dim, N, tuple(full.shape)) == (2, 8, (3, 8, 8, 1))
    assert torch.all(full[2] == 0) and torch.all(full[:2] == 1)
    assert physical_field(torch.ones(2, 8, 8, 1), dim=2)[0] == 2
    for bad in [torch.zeros(3, 8, 8), torch.zeros(3, 8, 8, 4), torch.zeros(2, 8, 8, 8), torch.zeros(3, 8, 4, 8),
                torch.zeros(8, 8), torch.zeros(3, 8, 8, 8, 1), torch.zeros(3, 8, 8, 8, dtype=torch.complex64)]:
        with pytest.raises(ValueError):
            physical_field(bad)
    with pytest.raises(ValueError):
        physical_field(torch.zeros(3, 8, 8, 8), dim=2)
    with pytest.raises(ValueError):
        physical_field(torch.zeros(3, 8, 8, 1), dim=3)


def test_spectrum_dim_shapes():
    from l3es_amd.utils import spectrum_dim

    assert spectrum_dim(torch.zeros(3, 8, 8, 5, dtype=torch.complex128)) == 3
    assert spectrum_dim(torch.zeros(3, 8, 5, dtype=torch.complex64)) == 2
    for bad in [torch.zeros(3, 8, 8, 8, dtype=torch.complex128), torch.zeros(3, 8, 8, 5),
                torch.zeros(2, 8, 5, dtype=torch.complex64), torch.zeros(3, 8, 8, 1, dtype=torch.complex64)]:
        with pytest.raises(ValueError):
            spectrum_dim(bad)


@pytest.mark.parametrize("fn", ["spectral_filtering", "energy_spectrum", "comp_dt"])
def test_helpers_refuse_bad_fields(fn):
    import l3es_amd
    from l3es_amd import utils

    bad = torch.zeros(3, 16, 16)  # neither 3-D nor the solver's 2-D layout
    with pytest.raises(ValueError):
        if fn == "spectral_filtering":
            utils.spectral_filtering(bad, 8)
        elif fn == "energy_spectrum":
            utils.energy_spectrum(bad, dim=2)
        else:
            l3es_amd.comp_dt(bad, 0.1, 0.01)


@pytest.mark.parametrize("fn", ["comp_dissipation_rate", "kinetic_energy_hat"])
def test_helpers_refuse_bad_spectra(fn):
    from l3es_amd import utils

    bad = torch.zeros(3, 16, 16, 16, dtype=torch.complex64)  # a full, not a half spectrum
    with pytest.raises(ValueError):
        if fn == "comp_dissipation_rate":
            utils.comp_dissipation_rate(bad, 0.01)
        else:
            utils.kinetic_energy_hat(bad)


def test_hit_statistics_closed_form():
    """spectral_solver.py:344-358 with numbers chosen so every quantity is exact: mean injection
    eps = 0.25 and kf = 2 give eps kf^2 = 1, a turnover time of 1; nu = 0.25 gives eta =
    (nu^3 / eps)^(1/4) = 0.5; N = 48 gives kmax = 16 sqrt(2); E = 0.75 gives Re_lambda = sqrt(60)."""
    import numpy as np

    from l3es_amd.spectral_solver import hit_statistics

    s = hit_statistics(e_inj_acc=2.5, steps=10, kf=2, N=48, nu=0.25, t_final=3.0, e_kin=0.75)
    assert abs(s["e_inj_mean"] - 0.25) < 1e-15
    assert abs(s["eddy_turnover_time"] - 1.0) < 1e-15
    assert abs(s["eddy_turnover_times"] - 3.0) < 1e-14
    assert abs(s["eta"] - 0.5) < 1e-15
    assert abs(s["kmax_eta"] - 8 * np.sqrt(2)) < 1e-13
    assert abs(s["re_lambda"] - np.sqrt(60)) < 1e-13
