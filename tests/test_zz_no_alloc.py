"""gi_half_ls_noalloc (device buffers in the library's static scratch array)
against gi_half_ls (allocating): bitwise equal, through the C ABI, on two
calls in a row with the second smaller (the array keeps the first call's
contents), and on a second context of the same device."""
import numpy as np
import pytest

from gpu_intensities_amd import binding, synth

pytestmark = pytest.mark.gpu


def _case(eng, job, jindI, jindF, g, ideg, seed):
    rng = np.random.default_rng(seed)
    n = int(job.bsets[jindI].nsize[g])
    vec = rng.standard_normal(n)
    ref = eng.half_ls(jindI, jindF, g, ideg, vec)
    got = eng.half_ls_noalloc(jindI, jindF, g, ideg, vec)
    print(f"[no_alloc] J {job.jvals[jindI]}->{job.jvals[jindF]} g={g} ideg={ideg} "
          f"nsize={n} dimenF={len(ref)}", flush=True)
    assert len(got) == len(ref) > 0
    assert np.array_equal(got, ref)
    assert np.abs(ref).max() > 0
    return len(ref)


def test_half_ls_noalloc_equals_half_ls():
    # C2v, J 2..4: the first call (J 4 -> 4, the largest blocks) fills more of
    # the array than the second (J 3 -> 2)
    job = synth.make_job("C2V", nvib=60, jvals=(2, 3, 4), seed=31)
    eng = binding.Engine(0).load(job)
    try:
        big = _case(eng, job, 2, 2, 1, 0, 1)
        small = _case(eng, job, 1, 0, 3, 0, 2)
        assert small < big
        # and back to a large one after the small one
        _case(eng, job, 2, 1, 0, 0, 3)
    finally:
        eng.close()
    # C3v (a degenerate E block, both components) on a fresh context: the
    # array is shared per device, so it still holds the C2v call's values
    job = synth.make_job("C3V", nvib=24, jvals=(3, 4), seed=5)
    eng = binding.Engine(0).load(job)
    try:
        e = int(np.nonzero(np.asarray(job.sym["degen"]) == 2)[0][0])
        a = _case(eng, job, 1, 1, e, 1, 4)
        b = _case(eng, job, 0, 1, e, 0, 5)
        c = _case(eng, job, 1, 0, 0, 0, 6)
        assert c < a and b == a
    finally:
        eng.close()

