"""Streamed mu (gi_set_mu_budget), the reference's split dipole.

The reference splits mu into parts of vF rows when it exceeds max_dipole_size
(cuda_host.cu:10; read_dipole_flipped_blocks, trove_functions.cpp:207-318) and
computes each part's half-linestrength rows from it (do_1st_half_ls_blocks,
cuda_host.cu:783-817). Here a mu larger than the device budget stays in pinned
host memory and every mode copies it in parts of vF rows: GI_MODE_FP64 and
GI_MODE_FP64_BATCHED the permuted FP64 rows, GI_MODE_SPLIT their digit rows
(built part by part, each row's digits depending on that row only). Every
output row is a whole sum over the same terms in the same order, so the lines
are the bits of resident mu, and they match the FP64 oracle to its tolerance.
"""
import numpy as np
import pytest
import torch

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare
from test_gpu_split import split_error

pytestmark = pytest.mark.gpu

MODES = (binding.GI_MODE_FP64, binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT)


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.fixture
def engine_resident(engine):
    yield engine
    engine.set_tuning(4, 0, 4, 4, 50)


@pytest.mark.parametrize("sym,kw", [
    ("C2V", dict(nvib=64, jvals=(2, 3), seed=41, e_span=12000.0)),
    ("C3V", dict(nvib=30, jvals=(2, 3), seed=5, reduced=1)),
])
def test_streamed_mu_same_bits(engine_resident, sym, kw):
    eng = engine_resident
    job = synth.make_job(sym, **kw)
    eng.load(job)
    assert eng.mu_parts() == 1
    ref = {m: _copy(eng.compute(mode=m)) for m in MODES}
    oref = oracle_py.OracleJob(job).run()
    compare(ref[binding.GI_MODE_FP64], oref)
    compare(ref[binding.GI_MODE_FP64_BATCHED], oref)
    row = 24 * job.nvib  # bytes of one vF row of mu (3 components)
    for rows in (job.nvib // 3 + 1, 7, 1):
        print(f"[mu_stream] {sym} parts of {rows} rows", flush=True)
        eng.load(job, mu_budget=rows * row + row // 2)
        assert eng.mu_parts() == -(-job.nvib // rows), rows
        for m in MODES:
            _same(ref[m], eng.compute(mode=m), (rows, m))
            ks = eng.kernel_stats()
            assert ks["mu_copy"]["launches"] >= eng.mu_parts(), (rows, m)
    # mu handed over on the device is copied to the host when streamed
    dev = torch.from_numpy(np.ascontiguousarray(job.dipole_flipped)).cuda()
    eng.load(job, device_dipole=dev, mu_budget=5 * row)
    assert eng.mu_parts() == -(-job.nvib // 5)
    for m in MODES:
        _same(ref[m], eng.compute(mode=m), ("device mu", m))
    eng.load(job)  # resident again
    assert eng.mu_parts() == 1
    for m in MODES:
        _same(ref[m], eng.compute(mode=m), ("resident again", m))


def test_streamed_mu_every_tile_and_route(engine_resident, monkeypatch):
    """The part loop under every half-ls tile shape of both matrix-core modes,
    in rounds cut by a small device budget (streamed mu is otherwise one round
    per call), and with the eigenvector blocks streamed as well (FP64
    batched): the bits of resident mu."""
    eng = engine_resident
    job = synth.make_job("C2V", nvib=120, jvals=(2, 3), seed=45, e_span=12000.0)
    eng.load(job)
    B, SP = binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT
    ref = {m: _copy(eng.compute(mode=m)) for m in (B, SP)}
    assert split_error(ref[SP], oracle_py.OracleJob(job).run())["max"] <= 1e-6
    eng.load(job, mu_budget=17 * 24 * job.nvib)
    assert eng.mu_parts() == -(-job.nvib // 17)
    for hl in range(6):
        print(f"[mu_stream] half-ls tile {hl}", flush=True)
        eng.set_tuning(hl, 0, 3, 2, 50)
        _same(ref[B], eng.compute(mode=B), ("tile", hl))
    for sc in range(5):
        print(f"[mu_stream] split half-ls tile {sc}", flush=True)
        eng.set_tuning(4, 0, 3, sc, 50)
        _same(ref[SP], eng.compute(mode=SP), ("split tile", sc))
    eng.set_tuning(4, 0, 4, 4, 50)
    monkeypatch.setenv("GI_BATCH_BYTES", str(1 << 20))
    for m in (B, SP):
        _same(ref[m], eng.compute(mode=m), ("small rounds", m))
        assert eng.kernel_stats()["mu_copy"]["launches"] > eng.mu_parts()
    monkeypatch.delenv("GI_BATCH_BYTES")
    eng.load(job, mu_budget=17 * 24 * job.nvib, vector_budget=1)
    assert eng.streamed_blocks() > 0
    _same(ref[B], eng.compute(mode=B), "blocks and mu streamed")
    _same(ref[SP], eng.compute(mode=SP), "blocks and mu digits streamed")
    eng.load(job)


def test_streamed_mu_rules(engine_resident):
    eng = engine_resident
    job = synth.make_job("C2V", nvib=40, jvals=(1, 2), seed=43)
    row = 24 * job.nvib
    with pytest.raises(binding.GIError) as e:
        eng.load(job, mu_budget=row - 1)  # below one row
    assert e.value.status == 1
    eng.load(job, mu_budget=4 * row)
    assert eng.mu_parts() == 10
    with pytest.raises(binding.GIError) as e:  # a captured sweep does not stream parts
        eng.sweep_create(binding.GI_MODE_FP64_BATCHED)
    assert e.value.status == 4
    ref = oracle_py.OracleJob(job).run()
    compare(eng.compute(mode=binding.GI_MODE_FP64), ref)
    compare(eng.compute(mode=binding.GI_MODE_FP64_BATCHED), ref)
    assert split_error(eng.compute(mode=binding.GI_MODE_SPLIT), ref)["max"] <= 1e-6
    eng.load(job)
    assert eng.mu_parts() == 1


def test_streamed_mu_requests(engine_resident):
    """Requests on the request server stream mu the same way in every mode:
    the lines of a request equal those with resident mu."""
    eng = engine_resident
    job = synth.make_job("C2V", nvib=48, jvals=(1, 2), seed=44, e_span=9000.0)
    req = dict(jindI=1, igammaI=0, jindF=0, temperature=800.0, q_stat=300.0)
    out = {}
    for budget in (0, 6 * 24 * job.nvib):
        eng.load(job, mu_budget=budget)
        eng.prepare_split()
        eng.serve_start(1 << 28)
        try:
            for m in MODES:
                out[budget, m] = _copy(eng.serve_fetch(eng.serve_submit(mode=m, **req)))
        finally:
            eng.serve_stop()
    for m in MODES:
        assert len(out[0, m]["S"]) > 0
        _same(out[0, m], out[6 * 24 * job.nvib, m], ("request", m))
    eng.load(job)
