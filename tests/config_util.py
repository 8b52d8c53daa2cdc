"""Shared set-up of the full-size GPU tests (test_gpu_config_*): the measured
workloads of synth.config_job with their eigenvector blocks built on the GPU,
and the host copies of only the blocks a sampled oracle check reads."""
import time

import numpy as np

from gpu_intensities_amd import binding, synth


def say(*a):
    print("[config]", *a, flush=True)


def load(name, **kw):
    """(job, device blocks, engine) of workload `name` ("c2", "c3", "c4"),
    with the split-mode digit planes built."""
    import torch
    t = time.time()
    job, seed = synth.config_job(name, **kw)
    say(f"{name}: host job {time.time() - t:.1f}s, {job.nlevels} levels")
    dev = synth.device_blocks(torch, job, seed, log=say)
    eng = binding.Engine(0)
    eng.load(job, device_vectors=dev)
    eng.prepare_split()
    torch.cuda.synchronize()
    say(f"{name}: context ready {time.time() - t:.1f}s")
    return job, dev, eng


def release(job, dev, eng):
    import torch
    eng.close()
    job.vectors.clear()
    dev.clear()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def lower_ok(job):
    e = job.levels["energy"] - job.zpe
    return np.nonzero((e >= job.erange_lower[0]) & (e <= job.erange_lower[1]))[0]


def blocks_read_by(job, lower_ids):
    """(jind, symmetry) blocks the oracle's chain of pair runs reads for these
    lower levels: each level's own block and the partner-symmetry blocks of
    every J' with |J' - J| <= 1."""
    lv = job.levels
    nj = len(job.jvals)
    need = set()
    for L in np.asarray(lower_ids):
        j, g = int(lv["jind"][L]), int(lv["igamma"][L])
        gF = int(job.igamma_pair[g])
        need.add((j, g))
        for jf in (j - 1, j, j + 1):
            if 0 <= jf < nj:
                need.add((jf, gF))
    return need


def to_host(job, dev, keys):
    for k in keys:
        if k not in job.vectors:
            job.vectors[k] = dev[k].cpu().numpy()


def sorted_lines(r):
    k = np.lexsort((r["ilevelF"], r["ilevelI"]))
    return {key: (v[k] if v.ndim == 1 else v[k, :]) for key, v in r.items()}


def same_bits(a, b):
    assert len(a["S"]) == len(b["S"])
    for k in ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp"):
        assert np.array_equal(a[k], b[k]), k
