"""The forced-HIT configuration at its full size: 256^3, Rogallo k^4 exp(-2 (k/4)^2)
initial field (E = 0.5), fixed injection power P = 0.1 on 0 < |k| <= 2, nu = 1e-3,
dt = 5e-4.

- FP64: one RK4 step on the GPU equals one step of the FP64 oracle
  (spectral_solver.py:50-118 restated, scipy.fft on the host cores) to 1e-12.
- FP32: ten steps stay within 2e-5 (fields) and 1e-6 (kinetic energy) of the
  FP64 GPU run, and the FP32 spectrum stays divergence-free (k . u_hat = 0 to
  FP32 rounding), properties that do not depend on the grid size.
"""

import os

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

N, NU, DT, P, KF, E0 = 256, 1e-3, 5e-4, 0.1, 2, 0.5


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def gpu_run(prec, steps):
    from l3es_amd import set_up_solver

    u, uh, _, _, step, _, _, _, _ = set_up_solver(N, NU, 3, "HIT", DT, 42, 32, KF, e_kin_target=E0,
                                                  forcing_type="power", power=P, spectrum="k4exp",
                                                  precision=prec, warmup=False)
    e = []
    for _ in range(steps):
        u, uh, ei = step(u, uh)
        e.append(float(ei))
    return u, uh, e


def test_hit256_fp64_step_matches_oracle(gpu):
    u, uh, e = gpu_run("float64", 1)
    ug, uhg = host(u), host(uh)
    del u, uh
    torch.cuda.empty_cache()
    fft = O.fft_backend(workers=min(16, len(os.sched_getaffinity(0))))
    uo, uho, step, _, _ = O.set_up_solver(N, NU, 3, "HIT", DT, seed=42, kf=KF, e_kin_target=E0,
                                          forcing_type="power", power=P, spectrum="k4exp", fft=fft,
                                          warmup=False)
    uo, uho, eo = step(uo, uho)
    assert rel(ug, uo) < 1e-12
    assert rel(uhg, uho) < 1e-12
    assert abs(e[0] - float(eo)) <= 1e-12 * abs(float(eo))


def test_hit256_fp32_against_fp64_and_divergence(gpu):
    u64, uh64, _ = gpu_run("float64", 10)
    a64, h64 = host(u64), host(uh64)
    del u64, uh64
    torch.cuda.empty_cache()
    u32, uh32, _ = gpu_run("float32", 10)
    a32, h32 = host(u32).astype(np.float64), host(uh32)
    assert rel(a32, a64) < 2e-5
    e32 = 0.5 * np.mean(np.sum(a32 * a32, axis=0))
    e64 = 0.5 * np.mean(np.sum(a64 * a64, axis=0))
    assert abs(e32 / e64 - 1) < 1e-6
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = k[: N // 2 + 1].copy()
    kz[-1] *= -1  # the rfft axis keeps +N/2 (spectral_solver.py:34-36)
    div = (k[:, None, None] * h32[0] + k[None, :, None] * h32[1] + kz[None, None, :] * h32[2])
    kmax = np.sqrt(3.0) * N / 2
    assert np.abs(div).max() < 1e-6 * kmax * np.abs(h32).max()
