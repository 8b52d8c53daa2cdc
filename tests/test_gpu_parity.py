"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerance (FP64 path): every transition amplitude A satisfies
|A_gpu - A_ref| <= 1e-10 * max(|A_ref|, sum_r |c_F[r] h[r]|) — the second
term is the condition floor for cancelling sums (the oracle reports it) — and
the emitted transition sets are identical.
"""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _sorted(r):
    k = np.lexsort((r["ilevelF"], r["ilevelI"]))
    return {key: (v[k] if v.ndim == 1 else v[k, :]) for key, v in r.items()}


def compare(got, ref, tol=TOL):
    g, r = _sorted(got), _sorted(ref)
    assert len(g["S"]) == len(r["S"]), (len(g["S"]), len(r["S"]))
    assert np.array_equal(g["ilevelI"], r["ilevelI"])
    assert np.array_equal(g["ilevelF"], r["ilevelF"])
    np.testing.assert_allclose(g["nu"], r["nu"], rtol=0, atol=0)
    floor = np.maximum(np.abs(r["amp"]), r["amp_abs"])
    err = np.abs(g["amp"] - r["amp"])
    bad = err > tol * floor + 1e-300
    assert not bad.any(), f"{bad.sum()} amplitudes out of tolerance, worst {err.max()}"
    rel = np.abs(g["S"] - r["S"]) / np.maximum(r["S"], 1e-300)
    via_floor = int((rel > tol).sum())
    # S is a sum of squares: a relative S error beyond tol can only come from a
    # cancelling amplitude, which the amplitude check above already bounds.
    for f in ("A_einst", "I"):
        relf = np.abs(g[f] - r[f]) / np.maximum(np.abs(r[f]), 1e-300)
        assert (relf[rel <= tol] <= 4 * tol).all(), f
    return dict(n=len(g["S"]), via_floor=via_floor, max_rel_S=float(rel.max()) if len(rel) else 0)


@pytest.fixture(scope="module")
def c1_job():
    return synth.make_job("C2V", nvib=200, jvals=(0, 1), seed=7)


def test_threej_table_matches_oracle():
    for jmax in (0, 1, 5, 11, 20):
        a = binding.threej_table(jmax)
        b = oracle_py.threej_table(jmax)
        np.testing.assert_allclose(a, b, rtol=4e-16, atol=0)


def test_correlate_parity(engine, c1_job):
    engine.load(c1_job)
    oj = oracle_py.OracleJob(c1_job)
    rng = np.random.default_rng(3)
    for jind in (0, 1):
        for g in range(4):
            n = int(c1_job.bsets[jind].nsize[g])
            u = rng.standard_normal(n)
            got = engine.correlate(jind, g, 0, u)
            ref = oj.correlate(jind, 0, g, u)
            np.testing.assert_array_equal(got, ref)


def test_half_ls_parity(engine, c1_job):
    engine.load(c1_job)
    oj = oracle_py.OracleJob(c1_job)
    rng = np.random.default_rng(4)
    for jindI, jindF in ((0, 1), (1, 0), (1, 1)):
        for g in range(4):
            n = int(c1_job.bsets[jindI].nsize[g])
            u = rng.standard_normal(n)
            c = oj.correlate(jindI, 0, g, u)
            ref = oj.half_ls(jindI, jindF, c)
            got = engine.half_ls(jindI, jindF, g, 0, u)
            scale = np.abs(ref).max()
            assert np.abs(got - ref).max() <= 1e-12 * scale, (jindI, jindF, g)


MODES = [binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED]


@pytest.mark.parametrize("mode", MODES)
def test_c1_full_parity(engine, c1_job, mode):
    engine.load(c1_job)
    got = engine.compute(mode=mode)
    ref = oracle_py.OracleJob(c1_job).run()
    st = compare(got, ref)
    assert st["n"] == len(ref["S"]) > 1000
    print("C1", st)


@pytest.mark.parametrize("mode", MODES)
def test_lower_subset_and_order(engine, c1_job, mode):
    engine.load(c1_job)
    ids = np.arange(0, c1_job.nlevels, 7)[::-1].copy()
    got = engine.compute(ids, mode=mode)
    ref = oracle_py.OracleJob(c1_job).run(ids)
    compare(got, ref)
    # rows of one (lower level, upper J) are contiguous, by upper energy
    jF = c1_job.levels["jind"][got["ilevelF"]]
    key = got["ilevelI"].astype(np.int64) * 2 + jF
    change = np.nonzero(np.diff(key))[0] + 1
    starts = np.concatenate([[0], change])
    assert len(set(key[starts].tolist())) == len(starts)
    eF = c1_job.levels["energy"][got["ilevelF"]]
    same = key[1:] == key[:-1]
    assert (eF[1:][same] >= eF[:-1][same]).all()
    if mode == binding.GI_MODE_FP64:
        # lower levels of one (lower block, upper J) in request order
        pos = {int(L): i for i, L in enumerate(ids)}
        blk = (c1_job.levels["jind"][got["ilevelI"]] * 8 + c1_job.levels["igamma"][got["ilevelI"]]) * 2 + jF
        for b_ in np.unique(blk):
            Ls = got["ilevelI"][blk == b_]
            firsts = [int(L) for i, L in enumerate(Ls) if i == 0 or Ls[i - 1] != L]
            assert [pos[L] for L in firsts] == sorted(pos[L] for L in firsts)


@pytest.mark.parametrize("mode", MODES)
def test_lower_ids_in_random_order(engine, mode):
    """Lower levels in random order, with repeats and out-of-range ids: the
    upper-level windows are searched afresh whenever the lower energy drops,
    so every window equals that of a lone search (the oracle's)."""
    job = synth.make_job("C2V", nvib=40, jvals=(3, 4), seed=23, e_span=15000.0,
                         freq_window=(-2000.0, 6000.0))
    engine.load(job)
    rng = np.random.default_rng(5)
    ids = rng.permutation(job.nlevels)[: job.nlevels // 2]
    ids = np.concatenate([ids, ids[:40], [-1, job.nlevels]]).astype(np.int32)
    got = engine.compute(ids, mode=mode)
    keep = ids[(ids >= 0) & (ids < job.nlevels)]
    ref = oracle_py.OracleJob(job).run(keep)
    compare(got, ref)
    assert len(ref["S"]) > 1000


@pytest.mark.parametrize("mode", MODES)
def test_nonsymmetric_mu_shuffled_cases(engine, mode):
    job = synth.make_job("C2V", nvib=48, jvals=(1, 2), seed=21, symmetric_mu=False,
                         shuffle_cases=True)
    engine.load(job)
    compare(engine.compute(mode=mode), oracle_py.OracleJob(job).run())


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("reduced", [0, 1])
def test_c3v_degenerate_parity(engine, reduced, mode):
    job = synth.make_job("C3V", nvib=30, jvals=(2, 3), seed=5, reduced=reduced)
    engine.load(job)
    got = engine.compute(mode=mode)
    ref = oracle_py.OracleJob(job).run()
    st = compare(got, ref)
    assert st["n"] > 100


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("reduced", [0, 1])
def test_d3h_six_irreps_parity(engine, reduced, mode):
    """D3H(M): six irreps, two of them degenerate, dipole partners A1'-A2",
    A2'-A1", E'-E" (fields.cpp:70-105); the synthetic basis is the direct sum
    of two C3v-like bases (the ' and the " vibrational functions)."""
    job = synth.make_job("D3H", nvib=24, jvals=(2, 3), seed=61, reduced=reduced)
    engine.load(job)
    got = engine.compute(mode=mode)
    ref = oracle_py.OracleJob(job).run()
    st = compare(got, ref)
    assert st["n"] > 100
    assert set(np.unique(job.levels["igamma"][got["ilevelI"]])) <= set(range(6))


@pytest.mark.parametrize("mode", MODES)
def test_windows_quanta_thresholds(engine, mode):
    job = synth.make_job("C2V", nvib=40, jvals=(3, 4), seed=9, e_span=20000.0,
                         freq_window=(-3000.0, 5000.0), erange_lower=(1000.0, 12000.0),
                         erange_upper=(500.0, 15000.0), thresh_intensity=1e-3,
                         thresh_linestrength=1e-6)
    job.quanta_lower[:] = [0, 1, 0, 5, 4, 5]
    job.quanta_upper[:] = [1, 0, 0, 5, 5, 3]
    engine.load(job)
    got = engine.compute(mode=mode)
    ref = oracle_py.OracleJob(job).run()
    compare(got, ref)
    assert len(ref["S"]) > 0


def test_empty_inputs(engine, c1_job):
    engine.load(c1_job)
    got = engine.compute(np.zeros(0, np.int32))
    assert len(got["S"]) == 0
    # window that admits nothing
    job = synth.make_job("C2V", nvib=8, jvals=(0, 1), seed=2, freq_window=(1e6, 2e6))
    engine.load(job)
    assert len(engine.compute()["S"]) == 0


def test_capacity_error(engine, c1_job):
    engine.load(c1_job)
    with pytest.raises(binding.GIError) as e:
        _compute_with_capacity(engine, 3)
    assert e.value.status == 5


def _compute_with_capacity(engine, cap):
    import ctypes as C
    L = binding.lib()
    b = engine.alloc_lines(cap)
    s = binding.GiLines(capacity=cap, n=0,
                        ilevelI=b["ilevelI"].ctypes.data_as(binding.P_i),
                        ilevelF=b["ilevelF"].ctypes.data_as(binding.P_i),
                        nu=b["nu"].ctypes.data_as(binding.P_d), S=b["S"].ctypes.data_as(binding.P_d),
                        A_einst=b["A_einst"].ctypes.data_as(binding.P_d),
                        I=b["I"].ctypes.data_as(binding.P_d), amp=b["amp"].ctypes.data_as(binding.P_d))
    rc = L.gi_compute(engine.h, 0, 0, None, C.byref(s))
    engine._chk(rc, "gi_compute")


@pytest.mark.parametrize("mode", MODES)
def test_larger_job_many_chunks(engine, mode):
    """More than one 256-item chunk per group, J 4 <-> 5, wide windows."""
    job = synth.make_job("C2V", nvib=64, jvals=(4, 5), seed=31, e_span=30000.0,
                         freq_window=(-4000.0, 9000.0))
    engine.load(job)
    ids = np.arange(0, job.nlevels, 3)
    got = engine.compute(ids, mode=mode)
    ref = oracle_py.OracleJob(job).run(ids)
    st = compare(got, ref)
    assert st["n"] > 20000


def test_modes_agree_on_transition_set(engine):
    job = synth.make_job("C3V", nvib=24, jvals=(3, 4), seed=13, e_span=12000.0)
    engine.load(job)
    a = engine.compute(mode=binding.GI_MODE_FP64)
    b = engine.compute(mode=binding.GI_MODE_FP64_BATCHED)
    compare(b, dict(a, amp_abs=np.abs(a["amp"]) * 0 + np.abs(a["amp"]).max()))


def test_batched_many_rounds(engine, monkeypatch):
    """A tiny round budget splits the batched mode into many rounds."""
    job = synth.make_job("C2V", nvib=32, jvals=(2, 3), seed=41, e_span=15000.0)
    engine.load(job)
    ref = oracle_py.OracleJob(job).run()
    monkeypatch.setenv("GI_BATCH_BYTES", str(1 << 20))
    got = engine.compute(mode=binding.GI_MODE_FP64_BATCHED)
    compare(got, ref)


def test_plan_threads_same_lines(engine, monkeypatch):
    """Chunks of a large job are built per item group on several host threads
    and joined in group order; forced to 1, 3 and 8 threads on a small job,
    every mode gives the same bits, also for a J list and for requests."""
    job = synth.make_job("C2V", nvib=40, jvals=(1, 2, 3, 4), seed=61, e_span=12000.0)
    engine.load(job)
    modes = (binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT)
    req = dict(jindI=2, igammaI=0, jindF=3, temperature=700.0, q_stat=0.0)
    out = {}
    for n in ("1", "3", "8"):
        monkeypatch.setenv("GI_PLAN_THREADS", n)
        for m in modes:
            out[n, m] = {k: v.copy() for k, v in engine.compute(mode=m).items()}
        engine.prepare_split()
        engine.serve_start(1 << 28)
        try:
            t = [engine.serve_submit(mode=m, **req) for m in modes]
            for m, x in zip(modes, t):
                out[n, "req", m] = {k: v.copy() for k, v in engine.serve_fetch(x).items()}
        finally:
            engine.serve_stop()
    monkeypatch.delenv("GI_PLAN_THREADS")
    for key in [k for k in out if k[0] == "1"]:
        assert len(out[key]["S"]) > 0, key
        for n in ("3", "8"):
            for f, v in out[key].items():
                assert np.array_equal(v, out[(n,) + key[1:]][f]), (key, n, f)


def test_tile_shapes_bitwise_equal(engine):
    """Every tile shape of the batched kernels gives the same bits (each
    product sums K in the same order whatever the tile), and matches the
    oracle; half-ls setting 4 picks the shape per pipeline round, 5 runs
    128 x 128 tiles on eight wavefronts."""
    job = synth.make_job("C3V", nvib=40, jvals=(3, 4), seed=17, e_span=15000.0,
                         freq_window=(-3000.0, 9000.0))
    engine.load(job)
    ids = np.arange(0, job.nlevels, 2)
    ref = oracle_py.OracleJob(job).run(ids)
    first = None
    try:
        for hl in range(6):
            for cr in range(4):
                engine.set_tiles(hl, cr)
                assert engine.get_tiles() == (hl, cr)
                got = engine.compute(ids, mode=binding.GI_MODE_FP64_BATCHED)
                got = {k: v.copy() for k, v in got.items()}
                if first is None:
                    first = got
                    compare(got, ref)
                else:
                    for k in first:
                        assert np.array_equal(first[k], got[k]), (hl, cr, k)
    finally:
        engine.set_tiles(4, 0)
