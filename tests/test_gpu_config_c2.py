"""The J 10 <-> 11 workload at full size, against the FP64 oracle.

C2v(M), Nvib = 2000, J 10 <-> 11: symmetry blocks of 10,500 and 11,500
functions, 88,000 levels, 2,968 lower levels with E <= 10,000 cm^-1, window
0-10,000 cm^-1, T = 296 K. At this size the correlation of both matrix-core
modes runs split K (upper blocks of 10,500 functions and more take 4 parts,
the part sums added in order by the intensity step), which no small job
reaches with the default limits. Every mode runs on a sample of the lower
levels (every 100th by energy) at full block size: FP64 modes within 1e-10 of
the oracle's condition floor, split mode within 1e-6 on lines above 1e-6 of
the strongest, identical transition sets. The split mode is also checked at
1500 K through a request with Q(T) from the device.
"""
import dataclasses

import numpy as np
import pytest

from gpu_intensities_amd import binding
import oracle_py
import config_util as cu
from test_gpu_parity import compare
from test_gpu_split import split_error

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module")
def c2():
    job, dev, eng = cu.load("c2")
    low = cu.lower_ok(job)
    sample = low[::100].astype(np.int32)
    cu.to_host(job, dev, cu.blocks_read_by(job, sample))
    ref = oracle_py.OracleJob(job).run(sample)
    cu.say(f"c2: oracle on {len(sample)} lower levels, {len(ref['S'])} lines")
    yield job, eng, sample, ref
    cu.release(job, dev, eng)


def test_c2_sizes(c2):
    job, eng, sample, ref = c2
    nsz = [int(n) for b in job.bsets for n in b.nsize]
    assert min(nsz) >= 4 * 2048            # split K: 4 parts in both matrix-core modes
    assert max(nsz) <= 19968               # records staged in LDS by the G builds
    assert len(cu.lower_ok(job)) > 2500 and len(ref["S"]) > 10000


@pytest.mark.parametrize("mode", ["fp64", "fp64_batched"])
def test_c2_fp64_modes_vs_oracle(c2, mode):
    job, eng, sample, ref = c2
    got = eng.compute(sample, mode=binding.MODES[mode])
    rep = compare(got, ref)
    cu.say(f"c2 {mode}: {rep}")


def test_c2_split_vs_oracle(c2):
    job, eng, sample, ref = c2
    got = eng.compute(sample, mode=binding.GI_MODE_SPLIT)
    rep = split_error(got, ref)
    cu.say(f"c2 split 296 K: {rep}")
    assert rep["max"] <= 1e-6


def test_c2_split_1500K_request(c2):
    """One request (lower block J = 10, A1, E - ZPE <= 1000 cm^-1, upper
    J = 11) at 1500 K with Q(T) from the device, against the oracle run with
    that T, Q and those lower levels."""
    job, eng, sample, ref = c2
    T = 1500.0
    Q = eng.partition_function(T)
    eng.serve_start(8 << 30)
    try:
        got = eng.serve_fetch(eng.serve_submit(jindI=0, igammaI=0, jindF=1,
                                               mode=binding.GI_MODE_SPLIT, temperature=T,
                                               q_stat=0.0, elow=(0.0, 1000.0)))
    finally:
        eng.serve_stop()
    lv = job.levels
    e = lv["energy"] - job.zpe
    ids = np.nonzero((lv["jind"] == 0) & (lv["igamma"] == 0) & (e >= 0) & (e <= 1000.0))[0]
    assert len(ids) > 10
    r = oracle_py.OracleJob(dataclasses.replace(job, temperature=T, q_stat=Q)).run(ids)
    keep = lv["jind"][r["ilevelF"]] == 1
    r = {k: v[keep] for k, v in r.items()}
    rep = split_error(got, r)
    cu.say(f"c2 split 1500 K request: {rep}")
    assert rep["max"] <= 1e-6


def test_c2_full_pass_modes_agree(c2):
    """The full pass (every lower level): both matrix-core modes emit the same
    transitions as each other with every candidate passing (thresholds 0),
    the same nu bits, and S within the split bar of the FP64 result."""
    job, eng, sample, ref = c2
    cap = eng.count_candidates()
    bufs = eng.alloc_lines(cap, pinned=True)
    a = cu.sorted_lines({k: v.copy() for k, v in
                         eng.compute(mode=binding.GI_MODE_FP64_BATCHED, bufs=bufs).items()})
    b = cu.sorted_lines(eng.compute(mode=binding.GI_MODE_SPLIT, bufs=bufs))
    assert len(a["S"]) == len(b["S"]) == cap
    assert np.array_equal(a["ilevelI"], b["ilevelI"])
    assert np.array_equal(a["ilevelF"], b["ilevelF"])
    assert np.array_equal(a["nu"], b["nu"])
    strong = a["S"] >= 1e-6 * a["S"].max()
    rel = np.abs(b["S"] - a["S"])[strong] / a["S"][strong]
    cu.say(f"c2 full pass: {cap} lines, split vs FP64 batched max rel S {rel.max():.3g}")
    assert rel.max() <= 1e-6
