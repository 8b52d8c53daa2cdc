"""Size-dependent kernel routes reached at small sizes (gi_set_path_limits).

At real sizes the batched modes switch kernels on the size of a product:
split-K correlation (upper blocks of 2048 functions and more), G builds with
only part of a record in LDS (lower records over 19,968 doubles; the
global-gather builds serve maps wider than 4), the one-step split half-ls
(classes over 512 K steps or 64 segments) and the two-kernel fold and digit
split (folded columns over 32,768). Lowering those thresholds sends small jobs
down the same code. Each route is checked against the FP64 oracle, and the
routes that keep every product's summation order (G builds, split half-ls,
fold) bitwise against the default route; the split-K route changes the
correlation's order and is held to the oracle tolerance of its mode.
"""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth
import oracle_py
from test_gpu_parity import compare
from test_gpu_split import split_error

pytestmark = pytest.mark.gpu

B, SP = binding.GI_MODE_FP64_BATCHED, binding.GI_MODE_SPLIT
KEYS = ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I", "amp")


def _run(eng, mode, **lim):
    eng.set_path_limits(**lim)
    try:
        r = eng.compute(mode=mode)
        return {k: v.copy() for k, v in r.items()}
    finally:
        eng.set_path_limits()


def _bitwise(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module", params=["C2V", "C3V"])
def routed(request):
    # blocks of a few hundred functions: split-K at kpart_len 64 cuts them
    # into the full 4 parts; C3v brings degenerate E blocks (2-term maps)
    if request.param == "C2V":
        job = synth.make_job("C2V", nvib=160, jvals=(2, 3), seed=41)
    else:
        job = synth.make_job("C3V", nvib=96, jvals=(3, 4), seed=43)
    eng = binding.Engine(0).load(job)
    ref = oracle_py.OracleJob(job).run()
    yield job, eng, ref
    eng.close()


@pytest.mark.parametrize("mode", [B, SP])
@pytest.mark.parametrize("rec_lds", [-1, 1, 37])
def test_global_gather_g_build_same_bits(routed, mode, rec_lds):
    # -1: the global-gather builds; 1 and 37: records longer than their part
    # staged in LDS, the rest gathered from global memory
    job, eng, ref = routed
    print(f"[paths] {job.sym['name']} mode {mode}: G builds, rec_lds {rec_lds}", flush=True)
    base = _run(eng, mode)
    glob = _run(eng, mode, rec_lds=rec_lds)
    _bitwise(glob, base)
    if mode == B:
        compare(glob, ref)
    else:
        assert split_error(glob, ref)["max"] <= 1e-6


@pytest.mark.parametrize("split_cfg", [2, 3])
def test_one_step_split_halfls_same_bits(routed, split_cfg):
    # over the step table, 64 x 64 with two steps in flight falls back to
    # 64 x 64 and 128 x 64 with two steps to 128 x 64
    job, eng, ref = routed
    print(f"[paths] {job.sym['name']}: one-step split half-ls from tile {split_cfg}",
          flush=True)
    t = eng.get_tuning()
    eng.set_tuning(**dict(t, split_cfg=split_cfg))
    try:
        base = _run(eng, SP)
        for lim in (dict(split2_steps=1), dict(split2_segs=1)):
            _bitwise(_run(eng, SP, **lim), base)
    finally:
        eng.set_tuning(**t)


def test_two_kernel_fold_split_same_bits(routed):
    job, eng, ref = routed
    print(f"[paths] {job.sym['name']}: two-kernel fold and digit split", flush=True)
    _bitwise(_run(eng, SP, fold_split_rows=1), _run(eng, SP))


@pytest.mark.parametrize("mode", [B, SP])
def test_split_k_correlation_vs_oracle(routed, mode):
    job, eng, ref = routed
    nsz = [int(n) for b in job.bsets for n in b.nsize]
    assert max(nsz) >= 4 * 64  # at least one block takes the full four parts
    print(f"[paths] {job.sym['name']} mode {mode}: split-K correlation, blocks {nsz}",
          flush=True)
    got = _run(eng, mode, kpart_len=64)
    if mode == B:
        rep = compare(got, ref)
        assert rep["n"] > 0
    else:
        assert split_error(got, ref)["max"] <= 1e-6
    # every route lowered at once
    got = _run(eng, mode, kpart_len=64, rec_lds=-1, split2_steps=1, fold_split_rows=1)
    if mode == B:
        compare(got, ref)
    else:
        assert split_error(got, ref)["max"] <= 1e-6


def test_path_limits_rules(routed):
    job, eng, ref = routed
    for bad in (dict(kpart_len=4096), dict(kpart_len=32), dict(rec_lds=30000),
                dict(split2_steps=513), dict(split2_segs=65), dict(fold_split_rows=40000),
                dict(rec_lds=-2)):
        with pytest.raises(binding.GIError):
            eng.set_path_limits(**bad)
    eng.set_path_limits()
