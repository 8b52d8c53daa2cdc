"""J lists on the CPU: the reference covers many J with one run per
consecutive pair (jvals[i], jvals[i+1]) (a run holds two J: fields.h:61,
cuda_host.cu:1005), and its Q-branch rule (no J = J' = jvals[0] within a
run, trove_functions.cpp:1457) makes the runs disjoint. The engine's J list
claims the union of those runs; this pins the claim on the oracle by brute
force over the full level table: the union of the oracle's pair runs is
exactly the set of (lower, upper) pairs that the list rule selects, with no
pair twice.
"""
import numpy as np
import pytest

from gpu_intensities_amd import synth
import oracle_py


def rule_pairs(job):
    """(lower, upper) pairs of the J-list rule, by brute force: gns > 0,
    selection partner, |dJ| <= 1, J + J' >= 1, not J = J' = jvals[0], the
    energy and quanta filters of the lower and upper level, nu in the window
    (trove_functions.cpp:1432-1464)."""
    lv = job.levels
    e = lv["energy"] - job.zpe
    q = lv["quanta"][:, 1:]
    nm = job.nmodes
    lo_q = np.all((q >= job.quanta_lower[:nm]) & (q <= job.quanta_lower[nm:]), axis=1)
    up_q = np.all((q >= job.quanta_upper[:nm]) & (q <= job.quanta_upper[nm:]), axis=1)
    low = lo_q & (e >= job.erange_lower[0]) & (e <= job.erange_lower[1])
    upp = up_q & (e >= job.erange_upper[0]) & (e <= job.erange_upper[1])
    J, g = lv["jval"], lv["igamma"]
    out = set()
    for i in np.nonzero(low)[0]:
        if not job.gns[g[i]] > 0:
            continue
        gF = job.igamma_pair[g[i]]
        if job.isym_pairs[g[i]] != job.isym_pairs[gF]:
            continue
        nu = lv["energy"] - lv["energy"][i]
        m = (upp & (g == gF) & (np.abs(J - J[i]) <= 1) & (J + J[i] >= 1)
             & ~((J == job.jvals[0]) & (J[i] == job.jvals[0]))
             & (nu >= job.freq_window[0]) & (nu <= job.freq_window[1]))
        out.update((int(i), int(f)) for f in np.nonzero(m)[0])
    return out


@pytest.mark.parametrize("sym,kw", [
    ("C2V", dict(nvib=16, jvals=(0, 1, 2, 3), seed=31)),
    ("C3V", dict(nvib=10, jvals=(1, 2, 3), seed=32)),
    ("C2V", dict(nvib=12, jvals=(2, 3, 5, 6), seed=33)),  # a gap: 3 -> 5 has no pair
])
def test_chain_of_pair_runs_is_the_list_rule(sym, kw):
    job = synth.make_job(sym, freq_window=(-1e4, 1e4), **kw)
    ref = oracle_py.run_chain(job)
    got = list(zip(ref["ilevelI"].tolist(), ref["ilevelF"].tolist()))
    assert len(got) == len(set(got)), "a pair came out of two runs"
    want = rule_pairs(job)
    assert set(got) == want
    # every (J, J') class of the list is there, except J = J' = jvals[0]
    J = job.levels["jval"]
    classes = {(int(J[i]), int(J[f])) for i, f in got}
    jv = job.jvals
    expect = {(a, b) for a in jv for b in jv
              if abs(a - b) <= 1 and a + b >= 1 and not (a == b == jv[0])}
    assert classes == expect


def test_pair_job_of_two_j_job_is_the_job():
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=34)
    sub, keep = oracle_py.pair_job(job, 0)
    assert np.array_equal(keep, np.arange(job.nlevels))
    a = oracle_py.OracleJob(job).run()
    b = oracle_py.run_chain(job)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_make_job_list_shares_blocks_with_pair_jobs():
    """The bases and eigenvector blocks of J are the same in every job that
    holds J (seeded per J), so a pair job cut from a list job equals the
    two-J job's inputs apart from the level energies."""
    a = synth.make_job("C2V", nvib=16, jvals=(1, 2, 3), seed=35)
    b = synth.make_job("C2V", nvib=16, jvals=(2, 3), seed=35)
    sub, _ = oracle_py.pair_job(a, 1)
    assert sub.jvals == b.jvals
    for x, y in zip(sub.bsets, b.bsets):
        assert np.array_equal(x.icontr2icase, y.icontr2icase)
        assert np.array_equal(x.ktau, y.ktau)
    for key in b.vectors:
        assert np.array_equal(sub.vectors[key], b.vectors[key])
    with pytest.raises(ValueError):
        synth.make_job("C2V", nvib=16, jvals=(1, 3, 2), seed=35)


def test_chain_with_lower_ids_is_the_subset():
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1, 2, 3), seed=36)
    full = oracle_py.run_chain(job)
    ids = np.arange(3, job.nlevels, 5)
    sub = oracle_py.run_chain(job, lower_ids=ids)
    m = np.isin(full["ilevelI"], ids)
    a = set(zip(full["ilevelI"][m].tolist(), full["ilevelF"][m].tolist()))
    b = list(zip(sub["ilevelI"].tolist(), sub["ilevelF"].tolist()))
    assert len(b) == len(set(b)) and set(b) == a and len(a) > 0
    none = oracle_py.run_chain(job, lower_ids=np.zeros(0, np.int64))
    assert len(none["S"]) == 0

