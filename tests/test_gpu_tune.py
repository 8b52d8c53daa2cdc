"""Autotuner and launch configuration (gi_autotune, gi_set_tuning).

The autotuner times the tile shapes of the FP64 matrix-core kernels and the
pipeline round count on the caller's workload and applies the fastest; every
setting must give bitwise-identical lines, so tuning never changes results.
"""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

DEFAULT = dict(halfls_cfg=4, corr_cfg=0, pipe_rounds=4, split_cfg=4, pipe_share=50)


@pytest.fixture
def tuned_engine(engine):
    yield engine
    engine.set_tuning(**DEFAULT)


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


@pytest.mark.parametrize("mode", [binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT])
def test_autotune_picks_a_setting_and_keeps_the_bits(tuned_engine, mode):
    eng = tuned_engine
    job = synth.make_job("C2V", nvib=48, jvals=(3, 4), seed=51, e_span=12000.0)
    eng.load(job)
    assert eng.get_tuning() == DEFAULT
    before = _copy(eng.compute(mode=mode))
    best, table = eng.autotune(mode, reps=2)
    assert eng.get_tuning() == best
    lines = table.strip().splitlines()
    assert lines[0].startswith("# gi_autotune") and lines[-1].startswith("best ")
    kinds = [ln.split()[0] for ln in lines[1:-1]]
    want = (["halfls_tile"] * 6 + ["corr_tile"] * 4 if mode == binding.GI_MODE_FP64_BATCHED
            else ["split_halfls_tile"] * 5) + ["pipe_rounds"] * 21  # 1 round, 5 x 4 pairs
    if mode == binding.GI_MODE_FP64_BATCHED:
        want += ["halfls_tile"] * 6  # the half-ls tile again at the chosen rounds
    assert kinds == want
    assert lines[-1] == ("best halfls_tile=%d corr_tile=%d pipe_rounds=%d split_tile=%d "
                         "pipe_share=%d" % (best["halfls_cfg"], best["corr_cfg"],
                                            best["pipe_rounds"], best["split_cfg"],
                                            best["pipe_share"]))
    after = _copy(eng.compute(mode=mode))
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    if mode == binding.GI_MODE_FP64_BATCHED:
        compare(after, oracle_py.OracleJob(job).run())


def test_every_round_count_gives_the_same_bits(tuned_engine, monkeypatch):
    """Round counts with the pipelined output: forced to several rounds by a
    small round budget, and at each pipe_rounds setting, the lines are the
    same bits as one round."""
    eng = tuned_engine
    job = synth.make_job("C2V", nvib=40, jvals=(2, 3), seed=52, e_span=15000.0)
    eng.load(job)
    for mode in (binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT):
        eng.set_tuning(3, 0, 1)
        ref = _copy(eng.compute(mode=mode))
        for rounds in (2, 3, 8, 64):
            for share in (50, 70, 10, 90):
                eng.set_tuning(3, 0, rounds, 0, share)
                got = eng.compute(mode=mode)
                for k in ref:
                    assert np.array_equal(ref[k], got[k]), (mode, rounds, share, k)
        eng.set_tuning(3, 0, 1)
        monkeypatch.setenv("GI_BATCH_BYTES", str(1 << 20))
        got = eng.compute(mode=mode)
        monkeypatch.delenv("GI_BATCH_BYTES")
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (mode, "small rounds", k)
    # the split half-ls tile shapes: exact integer level sums, so the same bits
    eng.set_tuning(3, 0, 2, 0)
    ref = _copy(eng.compute(mode=binding.GI_MODE_SPLIT))
    for cfg in (1, 2, 3, 4):
        eng.set_tuning(3, 0, 2, cfg)
        got = eng.compute(mode=binding.GI_MODE_SPLIT)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), ("split tile", cfg, k)


def test_autotune_plain_mode_has_nothing_to_tune(tuned_engine):
    eng = tuned_engine
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=53)
    eng.load(job)
    best, table = eng.autotune(binding.GI_MODE_FP64, reps=1)
    assert best == DEFAULT
    assert len(table.strip().splitlines()) == 2


def test_set_tuning_rules(tuned_engine):
    eng = tuned_engine
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=54)
    eng.load(job)
    for bad in ((6, 0, 2, 0), (0, -1, 2, 0), (0, 0, 0, 0), (0, 0, 65, 0), (0, 0, 2, 5),
                (0, 0, 2, 0, 9), (0, 0, 2, 0, 91)):
        with pytest.raises(binding.GIError):
            eng.set_tuning(*bad)
    assert eng.get_tuning() == DEFAULT
    eng.set_tuning(1, 2, 5, 1, 65)
    assert eng.get_tuning() == dict(halfls_cfg=1, corr_cfg=2, pipe_rounds=5, split_cfg=1,
                                    pipe_share=65)
    assert eng.get_tiles() == (1, 2)
    eng.serve_start(1 << 28)
    try:
        with pytest.raises(binding.GIError) as e:
            eng.set_tuning(**DEFAULT)
        assert e.value.status == 4
        with pytest.raises(binding.GIError):
            eng.autotune(binding.GI_MODE_FP64_BATCHED)
    finally:
        eng.serve_stop()
