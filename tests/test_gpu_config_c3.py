"""The J = 0..20 workload at full size, and a request burst against it.

One context holds C2v(M) J = 0..20 at Nvib = 2000: 84 symmetry blocks of 500
to 20,500 functions (99 GB of FP64 eigenvectors and 62 GB of split digit
planes resident), every |dJ| <= 1 pair in the same launches. Lower levels of
J = 20 have records longer than the 19,968 doubles the G builds stage in LDS,
so their G builds gather the rest of the record from global memory; the upper blocks run split-K
correlations. The sample (J = 1 every 4th lower level, four lower levels each
of J = 16 and J = 20, symmetry A1) runs in every mode against the oracle's
chain of pair runs (the reference runs J lists one (J, J+1) pair at a time);
a captured sweep of the sample replays bitwise. Then a Poisson burst of
requests (random pair, symmetry, T, window and threshold, both matrix-core
modes) runs against the resident blocks: three are checked against the oracle,
and five come out bitwise the same alone as in the burst.
"""
import dataclasses
import time

import numpy as np
import pytest

from gpu_intensities_amd import binding
import oracle_py
import config_util as cu
from test_gpu_parity import compare
from test_gpu_split import split_error

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

B, SP = binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT


@pytest.fixture(scope="module")
def c3():
    job, dev, eng = cu.load("c3")
    low = cu.lower_ok(job)
    lv = job.levels
    J, g = lv["jval"][low], lv["igamma"][low]
    parts = [low[J == 1][::4]]
    for jj in (16, 20):
        top = low[(J == jj) & (g == 0)]
        parts.append(top[:: max(1, len(top) // 4)][:4])
    sample = np.sort(np.concatenate(parts)).astype(np.int32)
    cu.to_host(job, dev, cu.blocks_read_by(job, sample))
    t = time.time()
    ref = oracle_py.run_chain(job, sample)
    cu.say(f"c3: oracle chain on {len(sample)} lower levels, {len(ref['S'])} lines, "
           f"{time.time() - t:.1f}s")
    yield job, eng, sample, ref, dev
    cu.release(job, dev, eng)


def test_c3_sizes(c3):
    job, eng, sample, ref, dev = c3
    lv = job.levels
    assert len(job.jvals) == 21
    nsz = {(j, g): int(job.bsets[j].nsize[g]) for j in range(21) for g in range(4)}
    # the sampled J = 20 lower records exceed the LDS record of the G builds
    for L in sample:
        if lv["jval"][L] == 20:
            assert nsz[(int(lv["jind"][L]), int(lv["igamma"][L]))] > 19968
    assert (lv["jval"][sample] >= 16).sum() == 8


@pytest.mark.parametrize("mode", ["fp64", "fp64_batched"])
def test_c3_fp64_modes_vs_oracle(c3, mode):
    job, eng, sample, ref, dev = c3
    got = eng.compute(sample, mode=binding.MODES[mode])
    rep = compare(got, ref)
    cu.say(f"c3 {mode}: {rep}")


def test_c3_split_vs_oracle(c3):
    job, eng, sample, ref, dev = c3
    rep = split_error(eng.compute(sample, mode=SP), ref)
    cu.say(f"c3 split: {rep}")
    assert rep["max"] <= 1e-6


@pytest.mark.parametrize("mode", [B, SP])
def test_c3_sweep_replays_bitwise(c3, mode):
    job, eng, sample, ref, dev = c3
    eager = {k: v.copy() for k, v in eng.compute(sample, mode=mode).items()}
    sw = eng.sweep_create(mode, sample)
    try:
        for _ in range(2):
            cu.same_bits(eng.sweep_run(sw), eager)
    finally:
        eng.sweep_destroy(sw)


def _oracle_request(job, r, Q):
    """The oracle's lines of one request: the chain of pair runs over the
    request's lower levels with its T, Q, window and thresholds, kept to its
    upper J."""
    lv = job.levels
    low = cu.lower_ok(job)
    ids = low[(lv["jind"][low] == r["jindI"]) & (lv["igamma"][low] == r["igammaI"])]
    j2 = dataclasses.replace(job, temperature=r["temperature"], q_stat=Q,
                             freq_window=tuple(r["nu"]), thresh_linestrength=0.0,
                             thresh_intensity=r["thresh_intensity"])
    ref = oracle_py.run_chain(j2, ids.astype(np.int32))
    keep = lv["jind"][ref["ilevelF"]] == r["jindF"]
    return {k: v[keep] for k, v in ref.items()}


def test_c5_request_burst(c3):
    job, eng, sample, ref, dev = c3
    rng = np.random.default_rng(2024)
    jv = job.jvals
    pairs = [(a, b) for a in range(21) for b in range(21)
             if abs(a - b) <= 1 and a + b >= 1 and not (a == b == 0)]

    def req(a, b, g, mode):
        nu0 = float(rng.uniform(0.0, 9000.0))
        return dict(jindI=a, igammaI=g, jindF=b, mode=mode,
                    temperature=float(rng.uniform(200.0, 2000.0)), q_stat=0.0,
                    nu=(nu0, nu0 + 1000.0), thresh_linestrength=0.0,
                    thresh_intensity=float(10.0 ** rng.uniform(-32, -26)))

    reqs = []
    for i in range(17):
        a, b = pairs[int(rng.integers(len(pairs)))]
        reqs.append(req(a, b, int(rng.integers(4)), B if i % 2 else SP))
    # three small-J requests (their blocks are cheap for the oracle)
    checked = [req(1, 2, 0, B), req(2, 1, 2, B), req(3, 3, 1, B)]
    for k, r in zip((2, 9, 15), checked):
        reqs.insert(k, r)
    gaps = rng.exponential(1.0 / 200.0, size=len(reqs))
    eng.serve_start(8 << 30)
    try:
        tickets = []
        t_next = time.perf_counter()
        for gap, r in zip(gaps, reqs):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(eng.serve_submit(**r))
        mixed, lat = [], []
        for t in tickets:
            lat.append(eng.serve_wait(t)[1])
            mixed.append({k: v.copy() for k, v in eng.serve_fetch(t).items()})
        st = eng.serve_stats()
        cu.say(f"c5 burst: {len(reqs)} requests, batches {st['batches']}, "
               f"largest {st['max_batch']}, latency p50 {np.percentile(lat, 50):.2f} ms "
               f"max {max(lat):.2f} ms, lines {sum(len(m['S']) for m in mixed)}")
        # alone vs in the burst: the three checked ones and two large-J ones
        big = [i for i, r in enumerate(reqs) if jv[r["jindI"]] >= 15][:2]
        for i in [2, 9, 15] + big:
            alone = eng.serve_fetch(eng.serve_submit(**reqs[i]))
            cu.same_bits(alone, mixed[i])
    finally:
        eng.serve_stop()
    # the checked requests read the blocks of J <= 4 only
    cu.to_host(job, dev, {(j, g) for j in range(5) for g in range(4)})
    nlines = 0
    for i in (2, 9, 15):
        r = reqs[i]
        Q = eng.partition_function(r["temperature"])
        rr = _oracle_request(job, r, Q)
        rep = compare(mixed[i], rr)
        nlines += rep["n"]
        cu.say(f"c5 request J {jv[r['jindI']]}->{jv[r['jindF']]} g{r['igammaI']} "
               f"T={r['temperature']:.0f}: {rep}")
    assert nlines > 0
