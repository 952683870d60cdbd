"""NumPy restatement of the reference's dataset packing and statistics (test infrastructure).

Only tests and the CPU-baseline leg of the benchmark use this module.  Functions
cite the lines of gen_dataset/gen_dataset.py they restate (paths relative to the
tumaer/ql-les-data checkout).

The reference packs the per-step particle frames of several trajectories into
train / valid / test files (single_h5_files_to_h5_dataset, :24-140) and then
computes normalisation statistics over the train split (compute_statistics_h5,
:143-238).  Its third-party pieces are restated from their published
definitions:
  - jax_md space.periodic (jax-md, pulled in by jax_sph 0.1.0): displacement
    mod(Ra - Rb + side/2, side) - side/2, with mod taking the divisor's sign;
  - h5py: only stores arrays, so the packed arrays are compared directly;
  - OmegaConf: only supplies the configuration values of the metadata.

Precision: the frames are FP32 (:91-104) and JAX runs without x64 in that
script, so the displacement and the frame differences are FP32 operations
(restated here in FP32, element by element).  The reference then averages in
FP32 (numpy's pairwise mean of an FP32 array); this restatement accumulates in
FP64 (trajectory_moments, statistics) and keeps an FP32-accumulating twin
(statistics(..., accumulate=np.float32)) to show the difference is FP32
rounding.

One behaviour is kept on purpose: the std loop centres u and au at their means
but vel and acc at zero, because the reference's mean loop assigns vel_mean and
acc_mean (:200-201) while the std loop subtracts v_mean and av_mean, which stay
0.0 from :163.  vel_std and acc_std are therefore root-mean-squares.

Parity pins (tests/test_oracle_dataset.py): split sizes against the
reference's formula evaluated by hand for several counts and ratios; the
connectivity-radius string; the displacement against a brute-force minimum
image on the seam; closed-form statistics of a uniform translation across the
periodic seam and of a constant-acceleration track; the centring quirk on a
field with a non-zero mean; FP64 against FP32 accumulation.
"""

import numpy as np

__all__ = [
    "split_bounds", "sequence_length", "connectivity_radius", "periodic_displacement_f32",
    "trajectory_quantities", "trajectory_moments", "statistics", "STD_FLOOR",
]

STD_FLOOR = 1e-7  # gen_dataset.py:215-216,222-223


def split_bounds(n_trajs, split="2_1_1"):
    """gen_dataset.py:45-52: trajectory index bounds of train / valid / test.
    Valid and test both take ceil(split[1] / sum(split) * n) trajectories (the
    third ratio is not used); train takes the rest."""
    s = np.array([int(x) for x in split.split("_")])
    num_eval = int(np.ceil(s[1] / s.sum() * n_trajs))
    bounds = np.cumsum([0, n_trajs - 2 * num_eval, num_eval, num_eval])
    return [int(b) for b in bounds], n_trajs - 2 * num_eval, num_eval


def sequence_length(n_files, skip_first_n_frames=0, slice_every_nth_frame=1):
    """gen_dataset.py:55-59: frames per trajectory after skipping and slicing
    (n_files counts every entry of the first trajectory's frame folder)."""
    return int(np.ceil((n_files - skip_first_n_frames) / slice_every_nth_frame))


def connectivity_radius(dx):
    """gen_dataset.py:131-135: 1.5 dx rounded to two significant digits."""
    x = np.format_float_positional(1.5 * dx, precision=2, unique=False, fractional=False, trim="k")
    return float(x)


def periodic_displacement_f32(ra, rb, box):
    """jax_md space.periodic displacement in FP32 (gen_dataset.py:150-154,180)."""
    box = np.float32(box)
    half = box * np.float32(0.5)
    d = np.asarray(ra, np.float32) - np.asarray(rb, np.float32)
    return np.mod(d + half, box) - half


def trajectory_quantities(r, u, box):
    """gen_dataset.py:172-186 for one trajectory r, u (T, np, dim) FP32:
    vel (T-1), acc (T-2), u (T), au (T-1) frames, all FP32."""
    r = np.asarray(r, np.float32)
    u = np.asarray(u, np.float32)
    vel = periodic_displacement_f32(r[1:], r[:-1], box)
    return vel, vel[1:] - vel[:-1], u, u[1:] - u[:-1]


def _colsum(x):
    """Per-component FP64 sums of (..., dim) values.  numpy sums a contiguous
    axis pairwise (error ~ log2(n) eps) but other axes one row after another
    (~ n eps), so the components are made the slow axis first."""
    x = np.asarray(x, np.float64)
    return np.ascontiguousarray(x.reshape(-1, x.shape[-1]).T).sum(1)


def trajectory_moments(r, u, box, center=None):
    """What the GPU moment kernel returns for one trajectory: for q in (vel,
    acc, u, au), the FP64 sums of the values, then the FP64 sums of their
    squared deviations from center (4, dim; zeros if None)."""
    qs = trajectory_quantities(r, u, box)
    dim = qs[2].shape[-1]
    c = np.zeros((4, dim)) if center is None else np.asarray(center, np.float64).reshape(4, dim)
    sums = [_colsum(q) for q in qs]
    sq = [_colsum((q.astype(np.float64) - c[i]) ** 2) for i, q in enumerate(qs)]
    return np.concatenate(sums + sq)


def statistics(trajs, box, accumulate=np.float64):
    """compute_statistics_h5 (gen_dataset.py:143-223) over trajectories
    [(r, u), ...] (already sliced by stats_every_nth): per-trajectory means,
    averaged over trajectories without weights (:181-186, :199-204), then the
    same for the squared deviations (:187-196, :206-211) with the reference's
    centring (vel, acc at zero; u, au at their means), square roots, and stds
    below 1e-7 set to 1.  accumulate = np.float32 follows the reference's FP32
    means."""
    if accumulate == np.float64:
        def mean(q):
            return _colsum(q) / (q.size // q.shape[-1])
    else:  # the reference's numpy mean over (time, particle) of an FP32 array
        def mean(q):
            return q.mean((0, 1), dtype=accumulate)
    f = (lambda a: a.astype(np.float64)) if accumulate == np.float64 else (lambda a: a)
    vs, avs, us, aus = [], [], [], []
    for r, u in trajs:
        vel, acc, uu, au = trajectory_quantities(r, u, box)
        vs.append(mean(vel))
        avs.append(mean(acc))
        us.append(mean(uu))
        aus.append(mean(au))
    vel_mean, acc_mean = np.stack(vs).mean(0), np.stack(avs).mean(0)
    u_mean, au_mean = np.stack(us).mean(0), np.stack(aus).mean(0)
    v_sq, av_sq, u_sq, au_sq = [], [], [], []
    for r, u in trajs:
        vel, acc, uu, au = trajectory_quantities(r, u, box)
        v_sq.append(mean(np.square(f(vel) - 0.0)))
        av_sq.append(mean(np.square(f(acc) - 0.0)))
        u_sq.append(mean(np.square(f(uu) - u_mean.astype(accumulate))))
        au_sq.append(mean(np.square(f(au) - au_mean.astype(accumulate))))
    out = {"vel_mean": vel_mean, "acc_mean": acc_mean, "u_mean": u_mean, "au_mean": au_mean}
    for name, sq in (("vel_std", v_sq), ("acc_std", av_sq), ("u_std", u_sq), ("au_std", au_sq)):
        std = np.stack(sq).mean(0) ** 0.5
        out[name] = np.where(std < STD_FLOOR, 1, std)
    return out
