"""TROVE files, the stdin deck and the line list (include/gi_trove.h), on the CPU.

* Round trip: a synthetic job written as TROVE files (trove_io.py) is read
  back by the library's readers into exactly the job's arrays.
* Against the reference: the same files read by the reference's own input
  stage (read_fields, bset_contr_factory, correlate_index, compute_ijterms,
  read_dipole_flipped, read_eigenvalues, built into oracle/_ref) give the
  same deck fields, bases, dipole and level table. The one known difference
  is the reference's qsort, which leaves its last level out of the sort
  (trove_functions.cpp:1414); levels are compared after a stable energy sort.
* The line layout of cuda_host.cu:1444-1473, field by field.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpu_intensities_amd import binding, synth, trove_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtrove_ref.so")

CASES = [
    dict(sym="C2V", nvib=16, jvals=(0, 1), seed=3),
    dict(sym="C3V", nvib=12, jvals=(2, 3), seed=4, reduced=1),
    dict(sym="D3H", nvib=14, jvals=(1, 2), seed=6),
    dict(sym="C2V", nvib=24, jvals=(3, 4), seed=5, erange_lower=(100.0, 6000.0),
         erange_upper=(50.0, 7000.0), shuffle_cases=True),
]


def _write(tmp_path, **kw):
    job = synth.make_job(**kw)
    d = str(tmp_path)
    deck = trove_io.write_trove(job, d)
    return job, d, deck


@pytest.mark.parametrize("case", range(len(CASES)))
def test_roundtrip_matches_job(tmp_path, case):
    job, d, deck = _write(tmp_path, **CASES[case])
    t = binding.Trove(d, deck)
    p = t.params()
    assert p["jvals"] == job.jvals and p["sym_type"] == job.sym["sym_type"]
    assert np.array_equal(p["gns"], job.gns) and np.array_equal(p["isym_pairs"], job.isym_pairs)
    assert p["reduced"] == job.reduced and p["temperature"] == job.temperature
    for jind in range(2):
        b, sb = t.bset(jind, job.sym["degen"]), job.bsets[jind]
        for k in ("icontr2icase", "iroot_correlat_j0", "k", "ktau", "N", "Ntotal", "ijterms",
                  "nsize"):
            assert np.array_equal(b[k], getattr(sb, k)), (jind, k)
        for g in range(job.sym["nrep"]):
            assert np.array_equal(b["repres"][g], sb.repres[g])
    mu, n = t.dipole()
    assert n == job.nvib and np.array_equal(mu, job.dipole_flipped)
    # the reader keeps the levels inside the energy window of the deck
    ref = synth.drop_level_filter(job).levels
    lv = t.levels(job.nmodes)
    for k in ("energy", "jind", "igamma", "ndeg", "irec0", "ilevel", "krot", "taurot"):
        assert np.array_equal(lv[k], ref[k]), k
    assert np.array_equal(lv["quanta"], ref["quanta"])
    for jind in range(2):
        for g in range(job.sym["nrep"]):
            assert np.array_equal(t.vectors(jind, g), job.vectors[(jind, g)])


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libtrove_ref.so not built (needs /root/reference)")
    L = C.CDLL(REF_SO)
    L.ref_read_job.restype = C.c_void_p
    L.ref_read_job.argtypes = [C.c_char_p, C.c_char_p]
    for f in ("ref_job_sizes", "ref_job_params", "ref_job_bset_sizes", "ref_job_bset",
              "ref_job_dipole", "ref_job_levels", "ref_job_ntotal"):
        getattr(L, f).argtypes = None
    return L


def _P(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class RefJob:
    """The reference's own input stage run on the files in d."""

    def __init__(self, d, deck):
        L = self.L = ref_lib()
        path = os.path.join(d, "deck.inp")
        with open(path, "w") as f:
            f.write(deck)
        self.h = C.c_void_p(L.ref_read_job(d.encode(), path.encode()))
        assert self.h.value
        s = np.zeros(5, np.int32)
        L.ref_job_sizes(self.h, _P(s, C.c_int))
        self.nrep, self.nmodes, self.nclasses, self.nlevels, self.N = (int(x) for x in s)

    def params(self):
        nrep, nm = self.nrep, self.nmodes
        dd = np.zeros(10 + nrep)
        ii = np.zeros(4 + 3 * nrep + 4 * nm, np.int32)
        self.L.ref_job_params(self.h, _P(dd, C.c_double), _P(ii, C.c_int))
        q = ii[4 + 3 * nrep:]
        return dict(zpe=dd[0], temperature=dd[1], q_stat=dd[2], thresh=dd[3],
                    freq_window=tuple(dd[4:6]), erange_lower=tuple(dd[6:8]),
                    erange_upper=tuple(dd[8:10]), gns=dd[10:], jvals=tuple(ii[:2]),
                    reduced=int(ii[2]), sym_type=int(ii[3]), isym_pairs=ii[4:4 + nrep],
                    isym_do=ii[4 + nrep:4 + 2 * nrep], igamma_pair=ii[4 + 2 * nrep:4 + 3 * nrep],
                    quanta_lower=np.concatenate([q[:nm], q[nm:2 * nm]]),
                    quanta_upper=np.concatenate([q[2 * nm:3 * nm], q[3 * nm:]]))

    def bset(self, which, degen):
        s = np.zeros(4, np.int32)
        self.L.ref_job_bset_sizes(self.h, which, _P(s, C.c_int))
        J, nc, d, ms = (int(x) for x in s)
        nrep = self.nrep
        a = dict(icontr2icase=np.zeros(2 * d, np.int32), iroot_correlat_j0=np.zeros(d, np.int32),
                 k=np.zeros(d, np.int32), ktau=np.zeros(d, np.int32),
                 N=np.zeros(nrep * nc, np.int32), Ntotal=np.zeros(nrep, np.int32),
                 ijterms=np.zeros(nrep * nc, np.int32), nsize=np.zeros(nrep, np.int32))
        nt = np.zeros(nrep, np.int32)
        self.L.ref_job_ntotal(self.h, which, _P(nt, C.c_int))
        big = [np.zeros(max(int(nt[g]) * degen[g] * ms, 1)) for g in range(nrep)]
        reps = (C.POINTER(C.c_double) * nrep)(*[_P(x, C.c_double) for x in big])
        self.L.ref_job_bset(self.h, which, *[_P(a[k], C.c_int) for k in (
            "icontr2icase", "iroot_correlat_j0", "k", "ktau", "N", "Ntotal", "ijterms", "nsize")],
            reps)
        a["repres"] = [big[g][:int(a["Ntotal"][g]) * degen[g] * ms].copy() for g in range(nrep)]
        a.update(jval=J, Maxsymcoeffs=nc, Maxcontracts=d, mat_size=ms)
        return a

    def dipole(self):
        out = np.zeros(3 * self.N * self.N)
        self.L.ref_job_dipole(self.h, _P(out, C.c_double))
        return out

    def levels(self):
        n, nm = self.nlevels, self.nmodes
        a = dict(energy=np.zeros(n), jind=np.zeros(n, np.int32), igamma=np.zeros(n, np.int32),
                 ndeg=np.zeros(n, np.int32), irec0=np.zeros(n, np.int32),
                 ilevel=np.zeros(n, np.int32), quanta=np.zeros(n * (nm + 1), np.int32),
                 krot=np.zeros(n, np.int32), taurot=np.zeros(n, np.int32))
        self.L.ref_job_levels(self.h, _P(a["energy"], C.c_double), *[_P(a[k], C.c_int) for k in (
            "jind", "igamma", "ndeg", "irec0", "ilevel", "quanta", "krot", "taurot")])
        a["quanta"] = a["quanta"].reshape(n, nm + 1)
        return a


@pytest.mark.parametrize("case", range(len(CASES)))
def test_readers_agree_with_reference(tmp_path, case):
    job, d, deck = _write(tmp_path, **CASES[case])
    ref = RefJob(d, deck)
    t = binding.Trove(d, deck)
    p, rp = t.params(), ref.params()
    assert p["jvals"] == rp["jvals"] and p["sym_type"] == rp["sym_type"]
    assert p["reduced"] == rp["reduced"] and p["zpe"] == rp["zpe"]
    for k in ("temperature", "q_stat", "freq_window", "erange_lower", "erange_upper"):
        assert np.array_equal(np.asarray(p[k]), np.asarray(rp[k])), k
    assert np.array_equal(p["gns"], rp["gns"]) and np.array_equal(p["isym_pairs"], rp["isym_pairs"])
    assert np.array_equal(p["thresh_linestrength"], rp["thresh"])
    degen = job.sym["degen"]
    for jind in range(2):
        b, rb = t.bset(jind, degen), ref.bset(jind + 1, degen)
        for k in ("icontr2icase", "iroot_correlat_j0", "k", "ktau", "N", "Ntotal", "ijterms"):
            assert np.array_equal(b[k], rb[k]), (jind, k)
        for g in range(job.sym["nrep"]):
            assert np.array_equal(b["repres"][g], rb["repres"][g]), (jind, g)
            if rp["isym_do"][g]:
                assert b["nsize"][g] == rb["nsize"][g]
    mu, n = t.dipole()
    assert n == ref.N and np.array_equal(mu, ref.dipole())
    lv, rl = t.levels(job.nmodes), ref.levels()
    assert len(lv["energy"]) == len(rl["energy"])
    order = np.argsort(rl["energy"], kind="stable")
    for k in ("energy", "jind", "igamma", "ndeg", "irec0", "ilevel", "krot", "taurot"):
        assert np.array_equal(lv[k], rl[k][order]), k
    assert np.array_equal(lv["quanta"], rl["quanta"][order])


def test_zpe_from_ground_level(tmp_path):
    """ZPE < 0 in the deck: the first J=0, symmetry-1 level becomes the ZPE
    (trove_functions.cpp:1201), in both readers."""
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=8)
    d = str(tmp_path)
    deck = trove_io.write_trove(job, d).replace("ZPE 0.0", "ZPE -1")
    t = binding.Trove(d, deck)
    e0 = job.levels["energy"][(job.levels["jind"] == 0) & (job.levels["igamma"] == 0)].min()
    assert t.params()["zpe"] == e0
    if os.path.exists(REF_SO):
        assert RefJob(d, deck).params()["zpe"] == e0


DECK_VARIANTS = [
    # lower case keywords, commas, comments, tabs in comments
    ("lower", lambda s: s.replace("TEMPERATURE", "temperature").replace("QSTAT", "qstat")),
    ("comments", lambda s: s.replace("END", "(last line)\nEND").replace(
        "GNS", "GNS (nuclear spin, per symmetry)")),
    ("thresh_v_reduced", lambda s: s.replace(
        "END", "  THRESH_LINE 1.5e-8\n  V 2 LOWER 0 3 UPPER 1 4\n  SYMMETRY REDUCED\nEND")),
    ("energy_order", lambda s: re.sub(r"ENERGY LOW (\S+) (\S+) UPPER (\S+) (\S+)",
                                      r"ENERGY UPPER \3 \4 LOW \1 \2", s)),
]


@pytest.mark.parametrize("name", [v[0] for v in DECK_VARIANTS])
def test_deck_variants_agree_with_reference(tmp_path, name):
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=9)
    d = str(tmp_path)
    deck = dict(DECK_VARIANTS)[name](trove_io.write_trove(job, d))
    p = binding.Trove(d, deck).params()
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built")
    rp = RefJob(d, deck).params()
    for k in ("temperature", "q_stat", "freq_window", "erange_lower", "erange_upper", "gns",
              "isym_pairs"):
        assert np.array_equal(np.asarray(p[k]), np.asarray(rp[k])), k
    assert p["reduced"] == rp["reduced"]
    assert np.array_equal(p["thresh_linestrength"], rp["thresh"])
    if name == "thresh_v_reduced":
        assert p["quanta_lower"][1] == rp["quanta_lower"][1] == 0
        assert p["quanta_lower"][job.nmodes + 1] == rp["quanta_lower"][job.nmodes + 1] == 3
        assert p["quanta_upper"][1] == 1 and p["quanta_upper"][job.nmodes + 1] == 4


def test_disabled_symmetry_needs_no_files(tmp_path):
    """GNS <= 1e-8 switches a symmetry off (fields.cpp:193-195): its descr and
    vector files are not read."""
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=10)
    d = str(tmp_path)
    deck = trove_io.write_trove(job, d).replace("GNS 1.0 1.0 3.0 3.0", "GNS 1.0 0.0 3.0 3.0")
    for J in (0, 1):
        os.remove(os.path.join(d, f"j0eigen_descr{J}_2.chk"))
        os.remove(os.path.join(d, f"j0eigen_vectors{J}_2.chk"))
    t = binding.Trove(d, deck)
    lv = t.levels(job.nmodes)
    assert not (lv["igamma"] == 1).any()


@pytest.mark.parametrize("what", ["deck", "quanta", "dipole", "descr", "vectors"])
def test_errors_are_statuses(tmp_path, what):
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=11)
    d = str(tmp_path)
    deck = trove_io.write_trove(job, d)
    if what == "deck":
        deck = deck.replace("SELECTION 1 1 2 2", "")
    elif what == "quanta":
        with open(os.path.join(d, "j0eigen_quanta1.chk"), "r+") as f:
            f.write("Bogus header")
    elif what == "dipole":
        with open(os.path.join(d, "j0_extfield.chk"), "r+b") as f:
            f.seek(4)
            f.write(b"Stop")
    elif what == "descr":
        os.remove(os.path.join(d, "j0eigen_descr1_3.chk"))
    else:
        with open(os.path.join(d, "j0eigen_vectors1_1.chk"), "r+b") as f:
            f.truncate(8)
    with pytest.raises(binding.GIError) as e:
        binding.Trove(d, deck)
    assert e.value.status == 1 and str(e.value)


def _printf_line(t_lv, labels, nm, ncl, zpe, res, r, cg, normal):
    """The reference's printf (cuda_host.cu:1444-1473) applied with Python's
    %-formatting, field by field."""
    iI, iF = int(res["ilevelI"][r]), int(res["ilevelF"][r])
    J = lambda i: int(t_lv["jval"][i])
    br = {0: "Q", 1: "R", -1: "P"}[J(iF) - J(iI)]
    s = "%4i %4s   <-%4i %4s   %1s  %11.4f <- %11.4f %11.4f  ( %3s %3i ) ( " % (
        J(iF), labels[t_lv["igamma"][iF]], J(iI), labels[t_lv["igamma"][iI]], br,
        t_lv["energy"][iF] - zpe, t_lv["energy"][iI] - zpe, abs(res["nu"][r]),
        labels[cg[iF][0]], t_lv["krot"][iF])
    s += "".join(" %3s" % labels[cg[iF][i]] for i in range(1, ncl + 1)) + " "
    s += "".join(" %3i" % t_lv["quanta"][iF][i] for i in range(1, nm + 1))
    s += " ) <- ( %3s %3i ) ( " % (labels[cg[iI][0]], t_lv["krot"][iI])
    s += "".join(" %3s" % labels[cg[iI][i]] for i in range(1, ncl + 1)) + " "
    s += "".join(" %3i" % t_lv["quanta"][iI][i] for i in range(1, nm + 1))
    s += ")  %16.8e %16.8e %16.8e   %6i <- %6i %8i %8i ( " % (
        res["S"][r], res["A_einst"][r], res["I"][r], t_lv["ilevel"][iF] + 1,
        t_lv["ilevel"][iI] + 1, 0, 0)
    s += "".join(" %3i" % normal[iF][i] for i in range(1, nm + 1)) + " ) <-  ( "
    s += "".join(" %3i" % normal[iI][i] for i in range(1, nm + 1)) + " ) "
    m = int(round(np.sqrt(res["amp"].shape[1])))  # amplitudes idegI + m*idegF
    for dF in range(t_lv["ndeg"][iF]):
        for dI in range(t_lv["ndeg"][iI]):
            s += " %16.9e" % res["amp"][r][dI + m * dF]
    return s + "\n"


@pytest.mark.parametrize("sym", ["C2V", "C3V", "D3H"])
def test_line_layout(tmp_path, sym):
    job = synth.make_job(sym, nvib=16 if sym == "C2V" else 12, jvals=(1, 2), seed=12)
    d = str(tmp_path)
    deck = trove_io.write_trove(job, d)
    t = binding.Trove(d, deck)
    lv = t.levels(job.nmodes)
    lv["jval"] = np.asarray(job.jvals)[lv["jind"]]
    rng = np.random.default_rng(1)
    n = 40
    res = dict(ilevelI=rng.integers(0, len(lv["energy"]), n).astype(np.int32),
               ilevelF=rng.integers(0, len(lv["energy"]), n).astype(np.int32),
               nu=rng.normal(0, 3000, n), S=10 ** rng.uniform(-40, 2, n),
               A_einst=10 ** rng.uniform(-20, 3, n), I=10 ** rng.uniform(-45, -18, n),
               amp=rng.normal(0, 1e-3, (n, max(job.sym["degen"]) ** 2)))
    # ΔJ within ±1 so the branch letter exists
    jI, jF = lv["jval"][res["ilevelI"]], lv["jval"][res["ilevelF"]]
    assert (np.abs(jF - jI) <= 1).all()
    text = t.format(res)
    # cgamma and normal come from the descr files: match levels by (jind, igamma, ilevel)
    key = {(int(a), int(b), int(c)): i for i, (a, b, c) in
           enumerate(zip(job.levels["jind"], job.levels["igamma"], job.levels["ilevel"]))}
    cg, normal = [], []
    for i in range(len(lv["energy"])):
        k = key[(int(lv["jind"][i]), int(lv["igamma"][i]), int(lv["ilevel"][i]))]
        cg.append(job.levels["cgamma"][k])
        normal.append(job.levels["normal"][k])
    want = "".join(_printf_line(lv, job.sym["labels"], job.nmodes, job.nclasses,
                                t.params()["zpe"], res, r, cg, normal) for r in range(n))
    assert text == want
    assert t.header().startswith("Linestrength S(f<-i) [Debye**2]")
    assert t.header().endswith("[cm/mol]\n\n\n")
