"""Persistent rounds (gi_set_persist) in GI_MODE_FP64_BATCHED.

With persistence on, each pipeline round runs its half linestrength, fold,
correlation and intensity steps as one launch whose workgroups take tasks in
ticket order and wait per chunk for the previous stage. Every task computes
what the per-stage launches compute, in the same order, so the lines must be
bitwise those of the per-stage launches: for every tile setting, for many
rounds and chunks (uneven blocks of a J list), split-K correlations, captured
sweeps replayed several times, and requests batched by the server. The
FP64 oracle bar of test_gpu_parity holds as well.
"""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

B = binding.GI_MODE_FP64_BATCHED
KEYS = ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp")


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


def _same(a, b):
    assert len(a["S"]) == len(b["S"])
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def _run(eng, persist, **tuning):
    eng.set_persist(persist)
    try:
        if tuning:
            eng.set_tuning(**tuning)
        return _copy(eng.compute(mode=B))
    finally:
        eng.set_persist(0)


@pytest.fixture(scope="module", params=["C2V", "C3V"])
def loaded(request):
    if request.param == "C2V":
        job = synth.make_job("C2V", nvib=160, jvals=(2, 3), seed=41)
    else:
        job = synth.make_job("C3V", nvib=96, jvals=(3, 4), seed=43)
    eng = binding.Engine(0).load(job)
    ref = oracle_py.OracleJob(job).run()
    yield job, eng, ref
    eng.close()


def test_persistent_equals_stages(loaded):
    job, eng, ref = loaded
    print(f"[persist] {job.sym['name']}: persistent vs per-stage launches", flush=True)
    base = _run(eng, 0)
    got = _run(eng, 1)
    _same(got, base)
    compare(got, ref)
    assert eng.get_persist() == 0


@pytest.mark.parametrize("halfls_cfg", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("corr_cfg", [0, 2, 1])
def test_every_tile_setting(loaded, halfls_cfg, corr_cfg):
    job, eng, ref = loaded
    print(f"[persist] {job.sym['name']}: tiles {halfls_cfg}/{corr_cfg}", flush=True)
    eng.set_tuning(halfls_cfg=4, corr_cfg=0, pipe_rounds=4)
    base = _run(eng, 0)
    # corr setting 1 (64 x 128) has no persistent kernel: 64 x 64 is used
    got = _run(eng, 1, halfls_cfg=halfls_cfg, corr_cfg=corr_cfg, pipe_rounds=4)
    eng.set_tuning(halfls_cfg=4, corr_cfg=0, pipe_rounds=4)
    _same(got, base)


def test_split_k_correlation(loaded):
    job, eng, ref = loaded
    eng.set_path_limits(kpart_len=64)
    try:
        base = _run(eng, 0)
        got = _run(eng, 1)
    finally:
        eng.set_path_limits()
    _same(got, base)
    compare(got, ref)


def test_j_list_many_rounds(engine, monkeypatch):
    # uneven blocks (J 0..6) and a small round budget: many rounds of many chunks
    job = synth.make_job("C2V", nvib=60, jvals=(0, 1, 2, 3, 4, 5, 6), seed=45)
    engine.load(job)
    ref = oracle_py.run_chain(job)
    for budget in (None, 1 << 20):
        if budget:
            monkeypatch.setenv("GI_BATCH_BYTES", str(budget))
        print(f"[persist] J 0..6, round budget {budget}", flush=True)
        base = _run(engine, 0)
        got = _run(engine, 1)
        _same(got, base)
        compare(got, ref)


def test_sweep_replays(engine):
    job = synth.make_job("C2V", nvib=40, jvals=(2, 3), seed=61, e_span=15000.0)
    engine.load(job)
    eager = _run(engine, 0)
    engine.set_persist(1)
    sw = engine.sweep_create(B)
    try:
        for i in range(3):
            print(f"[persist] sweep replay {i}", flush=True)
            _same(engine.sweep_run(sw), eager)
            engine.compute(mode=binding.GI_MODE_FP64)  # eager work in between
    finally:
        engine.sweep_destroy(sw)
        engine.set_persist(0)


def test_requests_alone_and_mixed(engine):
    job = synth.make_job("C2V", nvib=120, jvals=(1, 2), seed=21, e_span=9000.0)
    engine.load(job)
    reqs = [dict(jindI=0, igammaI=0, jindF=1, temperature=296.0, q_stat=174.5),
            dict(jindI=1, igammaI=2, jindF=0, temperature=1500.0, q_stat=0.0),
            dict(jindI=1, igammaI=3, jindF=1, temperature=700.0, q_stat=400.0)]
    out = {}
    for persist in (0, 1):
        engine.set_persist(persist)
        engine.serve_start(1 << 30)
        try:
            alone = [_copy(engine.serve_fetch(engine.serve_submit(mode=B, **r)))
                     for r in reqs]
            tickets = [engine.serve_submit(mode=B, **r) for r in reqs]
            mixed = [_copy(engine.serve_fetch(t)) for t in tickets]
        finally:
            engine.serve_stop()
            engine.set_persist(0)
        for a, m in zip(alone, mixed):
            _same(a, m)
        out[persist] = alone
    for a, b in zip(out[0], out[1]):
        _same(a, b)


def test_persist_rules(engine):
    with pytest.raises(binding.GIError):
        engine.set_persist(2)
    assert engine.get_persist() == 0
