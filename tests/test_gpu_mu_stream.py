"""Streamed mu (gi_set_mu_budget), the reference's split dipole.

The reference splits mu into parts of vF rows when it exceeds max_dipole_size
(cuda_host.cu:10; read_dipole_flipped_blocks, trove_functions.cpp:207-318) and
computes each part's half-linestrength rows from it (do_1st_half_ls_blocks,
cuda_host.cu:783-817). Here a mu larger than the device budget stays in pinned
host memory and GI_MODE_FP64 copies it in parts of vF rows. Every output row
is a whole sum over the same terms in the same order, so the lines are the
bits of resident mu, and they match the FP64 oracle to its tolerance.
"""
import numpy as np
import pytest
import torch

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def _copy(r):
    return {k: v.copy() for k, v in r.items()}


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.fixture
def engine_resident(engine):
    yield engine
    engine.set_tuning(4, 0, 4, 2, 50)


@pytest.mark.parametrize("sym,kw", [
    ("C2V", dict(nvib=64, jvals=(2, 3), seed=41, e_span=12000.0)),
    ("C3V", dict(nvib=30, jvals=(2, 3), seed=5, reduced=1)),
])
def test_streamed_mu_same_bits(engine_resident, sym, kw):
    eng = engine_resident
    job = synth.make_job(sym, **kw)
    eng.load(job)
    assert eng.mu_parts() == 1
    ref = _copy(eng.compute(mode=binding.GI_MODE_FP64))
    compare(ref, oracle_py.OracleJob(job).run())
    row = 24 * job.nvib  # bytes of one vF row of mu (3 components)
    for rows in (job.nvib // 3 + 1, 7, 1):
        eng.load(job, mu_budget=rows * row + row // 2)
        assert eng.mu_parts() == -(-job.nvib // rows), rows
        got = eng.compute(mode=binding.GI_MODE_FP64)
        _same(ref, got, rows)
        ks = eng.kernel_stats()
        assert ks["mu_copy"]["launches"] >= 1
    # mu handed over on the device is copied to the host when streamed
    dev = torch.from_numpy(np.ascontiguousarray(job.dipole_flipped)).cuda()
    eng.load(job, device_dipole=dev, mu_budget=5 * row)
    assert eng.mu_parts() == -(-job.nvib // 5)
    _same(ref, eng.compute(mode=binding.GI_MODE_FP64), "device mu")
    eng.load(job)  # resident again
    assert eng.mu_parts() == 1
    _same(ref, eng.compute(mode=binding.GI_MODE_FP64), "resident again")


def test_streamed_mu_rules(engine_resident):
    eng = engine_resident
    job = synth.make_job("C2V", nvib=40, jvals=(1, 2), seed=43)
    row = 24 * job.nvib
    with pytest.raises(binding.GIError) as e:
        eng.load(job, mu_budget=row - 1)  # below one row
    assert e.value.status == 1
    eng.load(job, mu_budget=4 * row)
    assert eng.mu_parts() == 10
    for mode in (binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT):
        with pytest.raises(binding.GIError) as e:
            eng.compute(mode=mode)
        assert e.value.status == 4
    with pytest.raises(binding.GIError):
        eng.sweep_create(binding.GI_MODE_FP64_BATCHED)
    ref = oracle_py.OracleJob(job).run()
    compare(eng.compute(mode=binding.GI_MODE_FP64), ref)
    eng.load(job)
    assert eng.mu_parts() == 1


def test_streamed_mu_requests(engine_resident):
    """FP64 requests on the request server stream mu the same way: the lines
    of a request equal those with resident mu; matrix-core requests are
    refused while mu is streamed."""
    eng = engine_resident
    job = synth.make_job("C2V", nvib=48, jvals=(1, 2), seed=44, e_span=9000.0)
    req = dict(jindI=1, igammaI=0, jindF=0, temperature=800.0, q_stat=300.0)
    out = {}
    for budget in (0, 6 * 24 * job.nvib):
        eng.load(job, mu_budget=budget)
        eng.serve_start(1 << 28)
        try:
            out[budget] = _copy(eng.serve_fetch(eng.serve_submit(mode=binding.GI_MODE_FP64,
                                                                 **req)))
            if budget:
                with pytest.raises(binding.GIError):
                    eng.serve_submit(mode=binding.GI_MODE_FP64_BATCHED, **req)
        finally:
            eng.serve_stop()
    assert len(out[0]["S"]) > 0
    _same(out[0], out[6 * 24 * job.nvib], "request")
    eng.load(job)
