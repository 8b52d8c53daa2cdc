"""Request server (gi_serve_*) and partition function on the GPU.

A request is (lower J, lower symmetry, upper J, lower-energy range,
frequency window, temperature, Q, thresholds). Its lines must equal the FP64
oracle's lines for the same job with those parameters, restricted to the
request's lower levels and upper J (tolerance of test_gpu_parity), and must
be bitwise identical whether the request runs alone or batched with others.
"""
import dataclasses

import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

PLANCK, VELLGT, BOLTZ = 6.62606896e-27, 2.99792458e10, 1.380658e-16


@pytest.fixture(scope="module")
def job():
    return synth.make_job("C2V", nvib=120, jvals=(1, 2), seed=21, e_span=9000.0)


def _block_lowers(job, jindI, g, elow=(-1e300, 1e300)):
    lv = job.levels
    e = lv["energy"] - job.zpe
    m = (lv["jind"] == jindI) & (lv["igamma"] == g) & (e >= elow[0]) & (e <= elow[1])
    return np.nonzero(m)[0]


def _upper_j(res, job, jindF):
    keep = job.levels["jind"][res["ilevelF"]] == jindF
    return {k: v[keep] for k, v in res.items()}


def _oracle_request(job, jindI, g, jindF, T, Q, nu, elow=(-1e300, 1e300), thS=0.0, thI=0.0):
    j2 = dataclasses.replace(job, temperature=T, q_stat=Q, freq_window=tuple(nu),
                             thresh_linestrength=thS, thresh_intensity=thI)
    ref = oracle_py.OracleJob(j2).run(_block_lowers(job, jindI, g, elow))
    return _upper_j(ref, job, jindF)


def _same(a, b):
    assert len(a["S"]) == len(b["S"])
    for k in ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp"):
        assert np.array_equal(a[k], b[k]), k


def _alone(engine, **req):
    t = engine.serve_submit(**req)
    return engine.serve_fetch(t)


def test_partition_function(engine, job):
    engine.load(job)
    lv = job.levels
    w = np.asarray(job.gns)[lv["igamma"]] * lv["ndeg"] * (2.0 * lv["jval"] + 1.0)
    for T in (100.0, 296.0, 1500.0):
        beta = PLANCK * VELLGT / (BOLTZ * T)
        ref = float(np.sum(w * np.exp(-(lv["energy"] - job.zpe) * beta)))
        got = engine.partition_function(T)
        assert abs(got - ref) <= 1e-13 * ref, (T, got, ref)
        assert got == engine.partition_function(T)  # fixed reduction order


@pytest.mark.parametrize("mode", [binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED])
def test_request_matches_oracle(engine, job, mode):
    engine.load(job)
    engine.serve_start(1 << 30)
    try:
        cases = [(0, 0, 1, 296.0, 174.5, (0.0, 8000.0), (-1e300, 1e300)),
                 (1, 2, 0, 1500.0, 900.0, (100.0, 5000.0), (0.0, 4000.0)),
                 (1, 3, 1, 700.0, 400.0, (0.0, 9000.0), (2000.0, 9000.0)),
                 (1, 1, 0, 296.0, 174.5, (50.0, 3000.0), (-1e300, 1e300))]
        for jindI, g, jindF, T, Q, nu, elow in cases:
            got = _alone(engine, jindI=jindI, igammaI=g, jindF=jindF, mode=mode,
                         temperature=T, q_stat=Q, nu=nu, elow=elow)
            ref = _oracle_request(job, jindI, g, jindF, T, Q, nu, elow)
            assert len(ref["S"]) > 0
            compare(got, ref)
    finally:
        engine.serve_stop()


@pytest.mark.parametrize("reduced", [0, 1])
def test_request_c3v_matches_oracle(engine, reduced):
    """Degenerate symmetry (E) with and without SYMMETRY REDUCED."""
    job = synth.make_job("C3V", nvib=30, jvals=(2, 3), seed=5, reduced=reduced)
    engine.load(job)
    engine.serve_start(1 << 30)
    try:
        n = 0
        for mode in (binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED):
            for g in range(job.sym["nrep"]):
                for jindI, jindF in ((0, 1), (1, 1)):
                    got = _alone(engine, jindI=jindI, igammaI=g, jindF=jindF, mode=mode,
                                 temperature=900.0, q_stat=321.0, nu=(0.0, 6000.0))
                    ref = _oracle_request(job, jindI, g, jindF, 900.0, 321.0, (0.0, 6000.0))
                    compare(got, ref)
                    n += len(ref["S"])
        assert n > 0
    finally:
        engine.serve_stop()


def test_request_equals_compute_subset(engine, job):
    """With the job's own T, Q and window, a request is gi_compute over its
    block's lower levels restricted to the upper J, bit for bit."""
    engine.load(job)
    want = {}
    for mode in (binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED):
        for jindI, g, jindF in ((0, 0, 1), (1, 2, 1), (1, 1, 0)):
            r = engine.compute(_block_lowers(job, jindI, g), mode=mode)
            want[mode, jindI, g, jindF] = {k: v.copy() for k, v in _upper_j(r, job, jindF).items()}
    engine.serve_start(1 << 30)
    try:
        for (mode, jindI, g, jindF), ref in want.items():
            got = _alone(engine, jindI=jindI, igammaI=g, jindF=jindF, mode=mode)
            _same(got, ref)
    finally:
        engine.serve_stop()


@pytest.mark.parametrize("mode", [binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED,
                                  binding.GI_MODE_SPLIT])
def test_auto_partition_function(engine, job, mode):
    """q_stat <= 0: the request uses gi_partition_function(T), reduced on the
    device on the batch's stream (the same additions in the same order, so
    the same bits), also for several temperatures in one batch."""
    engine.load(job)
    if mode == binding.GI_MODE_SPLIT:
        engine.prepare_split()
    Ts = (1200.0, 296.0, 2000.0)
    Qs = [engine.partition_function(T) for T in Ts]
    engine.serve_start(1 << 30)
    try:
        a = [_alone(engine, jindI=1, igammaI=0, jindF=1, temperature=T, q_stat=0.0, mode=mode)
             for T in Ts]
        b = [_alone(engine, jindI=1, igammaI=0, jindF=1, temperature=T, q_stat=Q, mode=mode)
             for T, Q in zip(Ts, Qs)]
        tickets = [engine.serve_submit(jindI=1, igammaI=0, jindF=1, temperature=T, q_stat=-1.0,
                                       mode=mode) for T in Ts]
        c = [engine.serve_fetch(t) for t in tickets]
    finally:
        engine.serve_stop()
    for x, y, z in zip(a, b, c):
        assert len(x["S"]) > 0
        _same(x, y)
        _same(x, z)
    assert not np.array_equal(a[0]["I"], a[1]["I"])


def test_mix_equals_alone(engine, job):
    """Requests submitted together (different T, windows, thresholds, modes)
    come back bitwise equal to the same requests served one at a time."""
    engine.load(job)
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(24):
        jindI = int(rng.integers(0, 2))
        reqs.append(dict(jindI=jindI, igammaI=int(rng.integers(0, 4)),
                         jindF=int(rng.integers(0, 2)),
                         mode=int(rng.integers(0, 2)),
                         temperature=float(rng.choice([296.0, 800.0, 1500.0])),
                         q_stat=float(rng.choice([0.0, 200.0])),
                         nu=(float(rng.uniform(0, 500)), float(rng.uniform(2000, 8000))),
                         elow=(float(rng.uniform(-1, 1000)), 1e300),
                         thresh_intensity=float(rng.choice([0.0, 1e-30]))))
    engine.serve_start(1 << 30)
    try:
        alone = [_alone(engine, **r) for r in reqs]
        tickets = [engine.serve_submit(**r) for r in reqs]
        mixed = [engine.serve_fetch(t) for t in tickets]
        st = engine.serve_stats()
    finally:
        engine.serve_stop()
    assert st["max_batch"] > 1, st
    assert sum(len(a["S"]) for a in alone) > 0
    for a, m in zip(alone, mixed):
        _same(a, m)


def test_budget_admission(engine, job):
    engine.load(job)
    engine.serve_start(4096)
    try:
        with pytest.raises(binding.GIError) as e:
            engine.serve_submit(jindI=1, igammaI=0, jindF=1)
        assert e.value.status == binding.GI_EBUDGET
    finally:
        engine.serve_stop()
    # smallest power-of-two budget that admits the request: it holds one copy
    # of it at a time, so six copies run one per batch and all complete
    budget, one = 8192, None
    while one is None:
        engine.serve_start(budget)
        try:
            one = _alone(engine, jindI=1, igammaI=0, jindF=1)
        except binding.GIError as e:
            assert e.status == binding.GI_EBUDGET
            budget *= 2
        finally:
            engine.serve_stop()
        assert budget < (1 << 34)
    engine.serve_start(budget)
    try:
        ts = [engine.serve_submit(jindI=1, igammaI=0, jindF=1) for _ in range(6)]
        outs = [engine.serve_fetch(t) for t in ts]
        st = engine.serve_stats()
    finally:
        engine.serve_stop()
    assert st["max_batch"] == 1 and st["served"] == 6, st
    for o in outs:
        _same(o, one)


def test_state_rules(engine, job):
    engine.load(job)
    engine.serve_start(1 << 30)
    try:
        with pytest.raises(binding.GIError) as e:
            engine.compute(mode=binding.GI_MODE_FP64)
        assert e.value.status == 4  # GI_ESTATE
        with pytest.raises(binding.GIError) as e:
            engine.serve_submit(jindI=0, igammaI=9, jindF=1)
        assert e.value.status == 1  # GI_EINVAL
        t = engine.serve_submit(jindI=0, igammaI=0, jindF=1)
        n, lat = engine.serve_wait(t, 60.0)
        assert lat > 0
        small = engine.alloc_lines(1)
        if n > 1:
            with pytest.raises(binding.GIError):
                engine.serve_fetch(t, bufs=small)
        engine.serve_fetch(t)
    finally:
        engine.serve_stop()
    r = engine.compute(mode=binding.GI_MODE_FP64)
    assert len(r["S"]) > 0
