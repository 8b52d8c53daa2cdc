"""Captured sweeps (gi_sweep_*): one HIP graph per pipeline round, replayed.

Every replay recomputes the products on the device; at the job's own
temperature a replay must equal gi_compute bit for bit, each of several
replays (with eager work and synchronisations between them, so a graph that
depends on state left by an earlier replay would show). At another
temperature, Q and threshold the lines must match the FP64 oracle run with
those values; S and A do not depend on T and stay bitwise the same.
"""
import dataclasses

import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare
from test_gpu_split import split_error, TOL_S

pytestmark = pytest.mark.gpu

MODES = [binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT]


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


def _same(a, b, keys=("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp")):
    assert len(a["S"]) == len(b["S"])
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def job():
    return synth.make_job("C2V", nvib=40, jvals=(2, 3), seed=61, e_span=15000.0)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("rounds", ["one", "many"])
def test_replays_equal_compute(engine, job, mode, rounds, monkeypatch):
    if rounds == "many":
        monkeypatch.setenv("GI_BATCH_BYTES", str(1 << 20))  # the plan cuts many rounds
    engine.load(job)
    eager = _copy(engine.compute(mode=mode))
    assert len(eager["S"]) > 1000
    sw = engine.sweep_create(mode)
    try:
        for i in range(3):
            got = engine.sweep_run(sw)
            _same(got, eager)
            other = engine.compute(mode=binding.GI_MODE_FP64)  # eager work in between
            assert len(other["S"]) == len(eager["S"])
    finally:
        engine.sweep_destroy(sw)


@pytest.mark.parametrize("mode", MODES)
def test_other_temperature_matches_oracle(engine, job, mode):
    engine.load(job)
    sw = engine.sweep_create(mode)
    try:
        base = _copy(engine.sweep_run(sw))
        T = 1500.0
        Q = engine.partition_function(T)
        hot = _copy(engine.sweep_run(sw, temperature=T, q_stat=0.0))
        again = engine.sweep_run(sw)
        _same(again, base)  # the graphs do not keep the last temperature
    finally:
        engine.sweep_destroy(sw)
    _same(hot, base, keys=("ilevelI", "ilevelF", "nu", "S", "A_einst", "amp"))
    ref = oracle_py.OracleJob(dataclasses.replace(job, temperature=T, q_stat=Q)).run()
    if mode == binding.GI_MODE_SPLIT:
        assert split_error(hot, ref)["max"] <= TOL_S
    else:
        compare(hot, ref)


def test_thresholds_at_replay(engine, job):
    engine.load(job)
    mode = binding.GI_MODE_FP64_BATCHED
    sw = engine.sweep_create(mode)
    try:
        full = _copy(engine.sweep_run(sw))
        thI = float(np.median(full["I"]))
        thS = float(np.percentile(full["S"], 20))
        got = _copy(engine.sweep_run(sw, thresh_linestrength=thS, thresh_intensity=thI))
    finally:
        engine.sweep_destroy(sw)
    keep = (full["S"] >= thS) & (full["I"] >= thI)
    assert 0 < keep.sum() < len(keep)
    _same(got, {k: v[keep] for k, v in full.items()})
    ref = oracle_py.OracleJob(dataclasses.replace(job, thresh_linestrength=thS,
                                                  thresh_intensity=thI)).run()
    compare(got, ref)


def test_sweep_rules(engine, job):
    engine.load(job)
    with pytest.raises(binding.GIError):
        engine.sweep_create(binding.GI_MODE_FP64)
    sw = engine.sweep_create(binding.GI_MODE_FP64_BATCHED, lower_ids=np.arange(0, 40))
    try:
        n = engine.sweep_candidates(sw)
        assert n > 0
        small = engine.alloc_lines(n - 1)
        with pytest.raises(binding.GIError) as e:
            engine.sweep_run(sw, bufs=small)
        assert e.value.status == 5
        with pytest.raises(binding.GIError) as e:
            engine.load(job)  # gi_set_params while a sweep exists
        assert e.value.status == 4
        got = engine.sweep_run(sw)
        ref = engine.compute(np.arange(0, 40), mode=binding.GI_MODE_FP64_BATCHED)
        _same(got, ref)
    finally:
        engine.sweep_destroy(sw)
    engine.load(job)  # allowed again


def test_destroy_context_with_live_sweep(job):
    eng = binding.Engine(0)
    eng.load(job)
    eng.sweep_create(binding.GI_MODE_SPLIT)
    eng.close()  # gi_destroy frees the sweep with the context
