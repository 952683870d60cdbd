"""Dataset packing and statistics (mirror of gen_dataset/gen_dataset.py).

single_frames_to_dataset  <- single_h5_files_to_h5_dataset (gen_dataset.py:24-140)
compute_statistics        <- compute_statistics_h5 (gen_dataset.py:143-238)
main                      <- the script's command line (gen_dataset.py:241-261)

h5py is not available, so the files are NumPy archives with the reference's
layout: the frames are the `com/step_XXXXX.npz` files (`r`, `u`) that
`combined` writes, and each split file `train.npz` / `valid.npz` / `test.npz`
holds the keys of the reference's h5 groups, `"{traj:05d}/position"`,
`"{traj:05d}/u"` (FP32, (T, np, dim)) and `"{traj:05d}/particle_type"`.  The
metadata files are the reference's JSON.

The statistics pass reads every trajectory of the train split once per loop
on the GPU (qlles_stats_moments: periodic displacements, frame differences
and their FP64 moments in one kernel); the split's arrays stay resident in
device memory between the two loops when they fit in a quarter of the free
memory.  The per-trajectory means and their unweighted average follow the
reference, including its centring: vel and acc deviations are taken about
zero, u and au about their means (gen_dataset.py:163,200-204).
"""

import argparse
import json
import os

import numpy as np
import torch

from . import _lib

__all__ = ["single_frames_to_dataset", "compute_statistics", "read_config", "main"]


def _frame_index(name):
    # "step_00012.npz" -> 12 (gen_dataset.py:69 takes the digits after "_")
    return int(os.path.splitext(name.split("_")[1])[0])


def read_config(traj_dir):
    """The run configuration a trajectory folder carries: `config.yaml` (what
    the reference's main.py:73 saves; plain YAML, read with SafeLoader) or
    `config.json`, as nested dicts with `sim` and `com` sections."""
    y = os.path.join(traj_dir, "config.yaml")
    if os.path.exists(y):
        import yaml

        with open(y) as fh:
            return yaml.safe_load(fh)
    with open(os.path.join(traj_dir, "config.json")) as fh:
        return json.load(fh)


def single_frames_to_dataset(src_dir, dst_dir, split="2_1_1", skip_first_n_frames=0, slice_every_nth_frame=1):
    """gen_dataset.py:24-140.  src_dir holds `traj_<seed>/com/step_XXXXX.npz`
    frames; trajectories are ordered by seed, valid and test take
    ceil(split[1] / sum(split) * n) each and train the rest (:45-52); frames
    are skipped and sliced per trajectory (:70).  Writes dst_dir/{train, valid,
    test}.npz and dst_dir/metadata.json and returns the metadata."""
    os.makedirs(dst_dir, exist_ok=True)
    dirs = [d for d in os.listdir(src_dir) if os.path.isdir(os.path.join(src_dir, d)) and d.startswith("traj_")]
    dirs = sorted(dirs, key=lambda x: int(x.split("_")[-1]))
    if not dirs:
        raise ValueError(f"no traj_* folders in {src_dir}")
    ratios = np.array([int(s) for s in split.split("_")])
    num_eval = int(np.ceil(ratios[1] / ratios.sum() * len(dirs)))
    bounds = np.cumsum([0, len(dirs) - 2 * num_eval, num_eval, num_eval])
    # every entry of the first trajectory's frame folder counts (:55-59)
    n_files = len(os.listdir(os.path.join(src_dir, dirs[0], "com")))
    seq_len = int(np.ceil((n_files - skip_first_n_frames) / slice_every_nth_frame))
    for i, name in enumerate(("train", "valid", "test")):
        arrays = {}
        for j, d in enumerate(dirs[bounds[i]:bounds[i + 1]]):
            com = os.path.join(src_dir, d, "com")
            files = sorted((f for f in os.listdir(com) if f.endswith(".npz") and f.startswith("step_")),
                           key=_frame_index)
            files = files[skip_first_n_frames::slice_every_nth_frame]
            pos, vel = [], []
            for f in files:
                with np.load(os.path.join(com, f)) as z:
                    pos.append(np.asarray(z["r"]))
                    vel.append(np.asarray(z["u"]))
            pos, vel = np.stack(pos), np.stack(vel)  # (T, np, dim)
            key = str(j).zfill(5)
            arrays[f"{key}/particle_type"] = np.zeros(pos.shape[1])
            arrays[f"{key}/u"] = vel.astype(np.float32)
            arrays[f"{key}/position"] = pos.astype(np.float32)
        np.savez(os.path.join(dst_dir, f"{name}.npz"), **arrays)
    cfg = read_config(os.path.join(src_dir, dirs[-1]))  # the last trajectory's configuration (:112)
    sim, com = cfg["sim"], cfg["com"]
    metadata = {  # :114-130
        "case": str(sim["case"]).upper(),
        "dim": sim["dim"],
        "dx": 2 * np.pi / sim["ckp_N"],
        "dt": sim["dt"],
        "t_end": sim["t_final"],
        "viscosity": sim["nu"],
        "u_ref": sim["u_ref"],
        "write_every": com["ckp_freq"] * slice_every_nth_frame,
        "sequence_length_train": seq_len,
        "num_trajs_train": int(len(dirs) - 2 * num_eval),
        "sequence_length_test": seq_len,
        "num_trajs_test": int(num_eval),
        "num_particles_max": sim["ckp_N"] ** sim["dim"],
        "periodic_boundary_conditions": [True] * 3,
        "bounds": [[0, 2 * np.pi]] * sim["dim"],
    }
    x = np.format_float_positional(1.5 * metadata["dx"], precision=2, unique=False, fractional=False, trim="k")
    metadata["default_connectivity_radius"] = float(x)  # :131-135
    with open(os.path.join(dst_dir, "metadata.json"), "w") as fh:
        json.dump(metadata, fh)
    return metadata


def _train_trajectories(dst_dir, every):
    with np.load(os.path.join(dst_dir, "train.npz")) as z:
        keys = sorted({k.split("/")[0] for k in z.files})
        for key in keys:
            fluid = np.asarray(z[f"{key}/particle_type"]) == 0  # :173-174
            r = np.ascontiguousarray(np.asarray(z[f"{key}/position"], np.float32)[::every][:, fluid])
            u = np.ascontiguousarray(np.asarray(z[f"{key}/u"], np.float32)[::every][:, fluid])
            yield r, u


def compute_statistics(dst_dir, stats_every_nth=1, device=None, keep_resident=True):
    """gen_dataset.py:143-238: normalisation statistics of the train split,
    written to dst_dir/metadata_every{n}.json (and returned).  The moments of
    each trajectory come from the GPU; the two loops of the reference (means,
    then deviations) are two passes over the split."""
    with open(os.path.join(dst_dir, "metadata.json")) as fh:
        metadata = json.load(fh)
    if not np.array(metadata["periodic_boundary_conditions"]).any():
        raise NotImplementedError("only periodic datasets are produced by this pipeline")
    bounds = np.array(metadata["bounds"], dtype=np.float64)
    box = bounds[:, 1] - bounds[:, 0]
    if not np.allclose(box, box[0]):
        raise NotImplementedError("the moment kernel takes one periodic side for all axes")
    dim = int(metadata["dim"])
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    lib = _lib.load()
    st = _lib.Stats(dim, float(box[0]))
    try:
        means, std = _two_loops(lib, st, dst_dir, stats_every_nth, dim, dev, keep_resident)
    finally:
        st.close()
    for i, name in enumerate(("vel", "acc", "u", "au")):
        metadata[f"{name}_mean"] = means[i].tolist()
        metadata[f"{name}_std"] = std[i].tolist()
    if not (metadata["write_every"] == 1 or stats_every_nth == 1):  # :229-232
        raise AssertionError("Either keep the dataset without slicing and have sliced stats files, "
                             "or slice the dataset and have unsliced stats.")
    metadata["write_every"] = metadata["write_every"] * stats_every_nth
    with open(os.path.join(dst_dir, f"metadata_every{stats_every_nth}.json"), "w") as fh:
        json.dump(metadata, fh)
    return metadata


def _two_loops(lib, st, dst_dir, every, dim, dev, keep_resident):
    """The mean and std loops of gen_dataset.py:168-211 over the train split;
    returns (means, stds), each (4, dim) for vel, acc, u, au."""
    out = torch.empty(8 * dim, dtype=torch.float64, device=dev)
    free = torch.cuda.mem_get_info(dev)[0]
    resident, held = [], 0

    def trajectories():
        if resident:
            yield from resident
            return
        for r, u in _train_trajectories(dst_dir, every):
            pair = (torch.from_numpy(r).to(dev), torch.from_numpy(u).to(dev))
            yield pair

    def moments(r, u, center):
        T, n = r.shape[0], r.shape[1]
        if r.shape != u.shape or r.shape[-1] != dim:
            raise ValueError(f"position {tuple(r.shape)} and u {tuple(u.shape)} must both be (T, np, {dim})")
        _lib.check(lib.qlles_stats_moments(st.handle, _lib.ptr(r), _lib.ptr(u), T, n, _lib.ptr(center),
                                           _lib.ptr(out), _lib.stream_ptr(torch.cuda.current_stream(dev))))
        m = out.cpu().numpy().reshape(8, dim)
        counts = np.array([(T - 1) * n, (T - 2) * n, T * n, (T - 1) * n], dtype=np.float64)[:, None]
        return m[:4] / counts, m[4:] / counts

    # loop "mean" (:168-186, :199-204)
    sums = []
    for r, u in trajectories():
        if keep_resident and held + 2 * r.numel() * 4 <= free // 4:
            resident.append((r, u))
            held += 2 * r.numel() * 4
        mean, _ = moments(r, u, None)
        sums.append(mean)
    if not sums:
        raise ValueError("the train split holds no trajectory")
    if len(resident) < len(sums):
        resident.clear()  # not all fit: the second loop reads the file again
    means = np.stack(sums).mean(0)  # (4, dim): vel, acc, u, au
    # loop "std" (:187-196, :206-211): vel and acc about zero, u and au about their means
    center = np.zeros((4, dim))
    center[2:] = means[2:]
    c_dev = torch.from_numpy(center.reshape(-1)).to(dev)
    sq = np.stack([moments(r, u, c_dev)[1] for r, u in trajectories()]).mean(0) ** 0.5
    return means, np.where(sq < 1e-7, 1, sq)  # :215-223


def main(argv=None):
    """The command line of gen_dataset.py:241-261."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--src_dir", type=str)
    ap.add_argument("--dst_dir", type=str)
    ap.add_argument("--split", default="2_1_1", type=str, help="Relative ratio.")
    ap.add_argument("--skip_first_n_frames", type=int, default=0)
    ap.add_argument("--slice_every_nth_frame", type=int, default=1)
    ap.add_argument("--stats_every_nth", type=int, default=1)
    ap.add_argument("--only_stats", action="store_true")
    args = ap.parse_args(argv)
    if not (args.slice_every_nth_frame == 1 or args.stats_every_nth == 1):
        raise AssertionError("Either keep the dataset without slicing and have sliced stats files, "
                             "or slice the dataset and have unsliced stats.")
    if not args.only_stats:
        single_frames_to_dataset(args.src_dir, args.dst_dir, args.split, args.skip_first_n_frames,
                                 args.slice_every_nth_frame)
    return compute_statistics(args.dst_dir, args.stats_every_nth)


if __name__ == "__main__":
    main()
