"""Eigenvector blocks beyond a device budget (gi_set_vector_budget).

Blocks that do not fit the budget stay in pinned host memory; every mode
copies each chunk's lower records and the upper records of its window (in
split mode their digit rows, built at gi_prepare_split and kept in pinned
host memory) to the device while the previous chunk computes (the reference
reads upper vectors in batches sized to free device memory,
cuda_host.cu:1011-1027, 1235-1372).
The kernels sum the same terms in the same order, so every line must be the
same bits as with resident blocks, also when records are stored out of energy
order (staged in several runs), for a J list, and for requests.
"""
import dataclasses

import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

MODES = (binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT)
KEYS = ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp")


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


def _same(a, b, what):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (what, k)


def _nonempty(job):
    return sum(1 for b in job.bsets for n in b.nsize if n > 0)


def _permuted(job, seed):
    """The same job with every block's records stored in a random order
    (irec0 permuted with the rows of the block)."""
    rng = np.random.default_rng(seed)
    lv = {k: v.copy() for k, v in job.levels.items()}
    vecs = {}
    for (jind, g), v in job.vectors.items():
        perm = rng.permutation(v.shape[0])
        w = np.empty_like(v)
        w[perm] = v
        vecs[(jind, g)] = w
        m = (lv["jind"] == jind) & (lv["igamma"] == g)
        lv["irec0"][m] = perm[lv["irec0"][m]]
    return dataclasses.replace(job, levels=lv, vectors=vecs)


@pytest.mark.parametrize("sym", ["C2V", "C3V"])
def test_streamed_blocks_same_bits(engine, sym):
    if sym == "C2V":
        job = synth.make_job("C2V", nvib=120, jvals=(2, 3), seed=71, e_span=12000.0)
    else:
        job = synth.make_job("C3V", nvib=48, jvals=(3, 4), seed=72, e_span=12000.0)
    ids = np.arange(0, job.nlevels, 3).astype(np.int32)
    engine.load(job)
    assert engine.streamed_blocks() == 0
    ref = {m: _copy(engine.compute(mode=m)) for m in MODES}
    sub = {m: _copy(engine.compute(ids, mode=m)) for m in MODES}
    total = sum(v.nbytes for v in job.vectors.values())
    for budget, nhost in ((1, _nonempty(job)), (total // 2, None)):
        print(f"[vec_stream] {sym} budget {budget}", flush=True)
        engine.load(job, vector_budget=budget)
        n = engine.streamed_blocks()
        assert (n == nhost) if nhost else (0 < n < _nonempty(job))
        for m in MODES:
            _same(engine.compute(mode=m), ref[m], (sym, budget, m))
            _same(engine.compute(ids, mode=m), sub[m], (sym, budget, m, "subset"))
    # records out of energy order: the staged copies come in several runs
    engine.load(_permuted(job, 5), vector_budget=1)
    for m in MODES:
        _same(engine.compute(mode=m), ref[m], (sym, "permuted", m))
    compare(ref[binding.GI_MODE_FP64_BATCHED], oracle_py.OracleJob(job).run())
    engine.load(job)


def test_streamed_blocks_j_list_and_requests(engine):
    job = synth.make_job("C2V", nvib=64, jvals=(1, 2, 3, 4), seed=73, e_span=12000.0)
    engine.load(job)
    ref = {m: _copy(engine.compute(mode=m)) for m in MODES}
    req = dict(jindI=2, igammaI=1, jindF=3, temperature=900.0, q_stat=0.0, nu=(100.0, 9000.0))
    engine.serve_start(1 << 30)
    try:
        rref = {m: _copy(engine.serve_fetch(engine.serve_submit(mode=m, **req))) for m in MODES}
    finally:
        engine.serve_stop()
    engine.load(job, vector_budget=1)
    assert engine.streamed_blocks() == _nonempty(job)
    for m in MODES:
        _same(engine.compute(mode=m), ref[m], ("jlist", m))
    engine.serve_start(1 << 30)
    try:
        for m in MODES:
            t = [engine.serve_submit(mode=m, **req) for _ in range(3)]
            for x in t:
                _same(engine.serve_fetch(x), rref[m], ("request", m))
    finally:
        engine.serve_stop()
    engine.load(job)


def test_streamed_blocks_rules(engine):
    job = synth.make_job("C2V", nvib=40, jvals=(0, 1), seed=74)
    engine.load(job, vector_budget=1)
    try:
        with pytest.raises(binding.GIError) as e:
            engine.sweep_create(binding.GI_MODE_FP64_BATCHED)
        assert e.value.status == 4
        assert binding.lib().gi_set_vector_budget(engine.h, -1) == 1  # GI_EINVAL
    finally:
        engine.load(job)
    assert engine.streamed_blocks() == 0
