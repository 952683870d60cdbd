"""FP64 CPU oracle for the spectral Navier-Stokes step and particle interpolation.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for ``cpu_baseline``, as the timed CPU restatement).  The product
path (``ql-les-data_amd/l3es_amd``) never imports it.

The oracle is a NumPy restatement of the reference algorithm
(tumaer/ql-les-data, package ``l3es``), written from the mathematics and the
reference's documented behaviour; each function cites the reference file:line
it follows.  Functions that exist only for modes this library adds (fixed-power HIT
forcing, Kolmogorov forcing + drag, integrating-factor viscous term, the
``k^4 exp(-2(k/k0)^2)`` Rogallo spectrum) have no reference counterpart and are
pinned by closed forms and invariants in ``tests/test_oracle.py``.

Parity pins (see ``tests/test_oracle.py``):
  * RHS + RK4: known-answer constant E_kin = 0.124953117517 +- 5e-8 for TGV 64^3,
    nu = 6.25e-4, dt = 0.01, 10 steps (from_spectralDNS/main_numpy.py:126-137);
    exact discrete decay of a single shear mode and of 2-D TGV.
  * Dealias mask (spectral_solver.py:42-48): enumeration against the exact
    integer rule 3|k| < 2 (N//2 + 1) with the +N/2 rfft Nyquist column at
    N = 8 ... 256, and an inviscid RHS that is non-zero on the last kept
    index and exactly zero past it.  (The known answer above does not see
    the mask.)
  * ekin_tot forcing (:93-112): e_inj and alpha in closed form on shear modes,
    both signs of each max(0, .).
  * comp_dt (:123-127): closed form on three shear components.
  * Rogallo init: vectorized restatement == scalar-loop twin bit-for-bit.
  * Filter, DFT interpolation, dissipation, spectrum: closed forms / Parseval.
  * ``init_u_kolm_np``: project-defined (the reference uses jax_cfd's JAX PRNG,
    so parity against the reference is unpinned); pinned by k.u_hat = 0,
    max|u| = 4.2, seed determinism, and an ensemble shell enstrophy
    ~ k^4 exp(-2 (k/k0)^2) peaking at k0 = 4.
  * ``pos_init_jittered``: inside [0, 2 pi), within +-0.25 dx of each cell
    centre and spanning it uniformly, seeded.
"""

from .spectral import *  # noqa: F401,F403
