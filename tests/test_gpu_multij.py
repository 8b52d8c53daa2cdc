"""One context holding many J (gi_set_jvals) on the GPU.

The lines of a context with the J list jvals[0..n-1] must be the lines of the
reference's chain of runs (jvals[i], jvals[i+1]) (the oracle's run_chain):
the same transitions, within the FP64 tolerance of test_gpu_parity for the
FP64 modes and the split-mode bar of test_gpu_split. The lines of the pair
(i, i+1) must be bitwise those of a two-J context loaded with that pair, in
every mode. Q(T) sums over every resident level of every J; requests address
any block pair of the list.
"""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare
from test_gpu_split import split_error, TOL_S

pytestmark = pytest.mark.gpu

PLANCK, VELLGT, BOLTZ = 6.62606896e-27, 2.99792458e10, 1.380658e-16
ALL_MODES = [binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT]


@pytest.fixture(scope="module")
def c2v_list():
    return synth.make_job("C2V", nvib=40, jvals=(0, 1, 2, 3, 4), seed=41)


@pytest.fixture(scope="module")
def c2v_chain(c2v_list):
    return oracle_py.run_chain(c2v_list)


def _check(got, ref, mode):
    if mode == binding.GI_MODE_SPLIT:
        rep = split_error(got, ref)
        assert rep["max"] <= TOL_S, rep
        return rep
    return compare(got, ref)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_list_equals_chain_of_runs(engine, c2v_list, c2v_chain, mode):
    engine.load(c2v_list)
    got = engine.compute(mode=mode)
    st = _check(got, c2v_chain, mode)
    print("C2v J 0..4", mode, st)
    assert len(c2v_chain["S"]) > 10000


@pytest.mark.parametrize("reduced", [0, 1])
def test_c3v_list(engine, reduced):
    job = synth.make_job("C3V", nvib=24, jvals=(1, 2, 3, 4), seed=42, reduced=reduced)
    ref = oracle_py.run_chain(job)
    engine.load(job)
    for mode in ALL_MODES:
        _check(engine.compute(mode=mode), ref, mode)


def test_gapped_list(engine):
    """A list with a gap (3 -> 5): the run (3, 5) only contributes 5 -> 5."""
    job = synth.make_job("C2V", nvib=20, jvals=(2, 3, 5, 6), seed=43)
    ref = oracle_py.run_chain(job)
    engine.load(job)
    for mode in ALL_MODES:
        _check(engine.compute(mode=mode), ref, mode)


def _run_classes(job, i):
    """(jindI, jindF) classes that the run (jvals[i], jvals[i+1]) owns."""
    own = {(i, i + 1), (i + 1, i), (i + 1, i + 1)}
    if i == 0:
        own.add((0, 0))
    return own


def _sorted(r):
    k = np.lexsort((r["ilevelF"], r["ilevelI"]))
    return {key: v[k] for key, v in r.items()}


@pytest.mark.parametrize("mode", ALL_MODES)
def test_pair_lines_bitwise_equal_two_j_context(engine, c2v_list, mode):
    engine.load(c2v_list)
    full = {k: v.copy() for k, v in engine.compute(mode=mode).items()}
    jind = c2v_list.levels["jind"]
    cls = list(zip(jind[full["ilevelI"]].tolist(), jind[full["ilevelF"]].tolist()))
    two = binding.Engine(0)
    try:
        for i in range(len(c2v_list.jvals) - 1):
            sub, keep = oracle_py.pair_job(c2v_list, i)
            two.load(sub)
            r = {k: v.copy() for k, v in two.compute(mode=mode).items()}
            r["ilevelI"] = keep[r["ilevelI"]].astype(np.int32)
            r["ilevelF"] = keep[r["ilevelF"]].astype(np.int32)
            own = _run_classes(c2v_list, i)
            m = np.array([c in own for c in cls])
            a = _sorted({k: v[m] for k, v in full.items()})
            b = _sorted(r)
            assert len(a["S"]) == len(b["S"]) > 0, i
            for k in ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp"):
                assert np.array_equal(a[k], b[k]), (i, k)
    finally:
        two.close()


def test_partition_function_over_every_j(engine, c2v_list):
    engine.load(c2v_list)
    lv = c2v_list.levels
    assert len(np.unique(lv["jval"])) == 5
    w = np.asarray(c2v_list.gns)[lv["igamma"]] * lv["ndeg"] * (2.0 * lv["jval"] + 1.0)
    for T in (296.0, 1500.0):
        beta = PLANCK * VELLGT / (BOLTZ * T)
        ref = float(np.sum(w * np.exp(-(lv["energy"] - c2v_list.zpe) * beta)))
        assert abs(engine.partition_function(T) - ref) <= 1e-13 * ref


def test_requests_across_the_list(engine, c2v_list):
    """A request for any (J, symmetry) -> J' of the list is gi_compute over
    that block's lower levels restricted to J', bit for bit; |J' - J| > 1
    is refused."""
    engine.load(c2v_list)
    lv = c2v_list.levels
    cases = [(0, 0, 1), (2, 1, 1), (2, 3, 3), (3, 2, 2), (4, 0, 3), (4, 2, 4)]
    want = {}
    engine.prepare_split()
    for mode in ALL_MODES:
        for jI, g, jF in cases:
            ids = np.nonzero((lv["jind"] == jI) & (lv["igamma"] == g))[0]
            r = engine.compute(ids, mode=mode)
            keep = lv["jind"][r["ilevelF"]] == jF
            want[mode, jI, g, jF] = {k: v[keep].copy() for k, v in r.items()}
    engine.serve_start(1 << 30)
    try:
        for (mode, jI, g, jF), ref in want.items():
            t = engine.serve_submit(jindI=jI, igammaI=g, jindF=jF, mode=mode)
            got = engine.serve_fetch(t)
            assert len(got["S"]) == len(ref["S"]) > 0
            for k in ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp"):
                assert np.array_equal(got[k], ref[k]), k
        with pytest.raises(binding.GIError):
            engine.serve_submit(jindI=1, igammaI=0, jindF=3)
        with pytest.raises(binding.GIError):
            engine.serve_submit(jindI=0, igammaI=0, jindF=5)
    finally:
        engine.serve_stop()


def test_set_jvals_rules(engine):
    """gi_set_jvals: ascending J >= 0 only; the bases and blocks of the old
    list are dropped, so gi_prepare needs the new ones."""
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1, 2), seed=44)
    engine.load(job)
    L = binding.lib()
    for bad in ([2, 1], [0, 0], [-1, 0]):
        jv = np.asarray(bad, np.int32)
        assert L.gi_set_jvals(engine.h, len(jv), binding._p(jv, binding.P_i)) == 1
    jv = np.asarray([0, 1, 2, 3], np.int32)
    assert L.gi_set_jvals(engine.h, 4, binding._p(jv, binding.P_i)) == 0
    assert L.gi_prepare(engine.h) == 4  # GI_ESTATE: no bases for the new list
