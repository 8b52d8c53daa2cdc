"""Pinning of the FP64 oracle (CPU only).

Each function of oracle/oracle.c is checked against something other than its
own formula: exact sympy values, the reference's own CPU code (built from
/root/reference into oracle/_ref/ when present), closed forms, sum rules,
reciprocity, Parseval, and brute-force truth tables.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from gpu_intensities_amd import synth
import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtrove_ref.so")


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libtrove_ref.so not built (needs /root/reference)")
    L = C.CDLL(REF_SO)
    L.ref_fakt.restype = C.c_double
    L.ref_fakt.argtypes = [C.c_double]
    L.ref_three_j.restype = C.c_double
    L.ref_three_j.argtypes = [C.c_int] * 6
    L.ref_precompute_threej.argtypes = [C.c_int, O.P_d, O.P_i]
    L.ref_filter_job.restype = C.c_void_p
    L.ref_filter_job.argtypes = [C.c_int, O.P_i, C.c_int, C.c_int, C.c_double, O.P_d, O.P_d,
                                 O.P_d, O.P_i, O.P_d, O.P_i, O.P_i, C.c_int, O.P_i, O.P_i]
    L.ref_intensity_filter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_int, C.c_int, O.P_i, O.P_i]
    L.ref_energy_filter_lower.argtypes = [C.c_void_p, C.c_int, C.c_double, O.P_i]
    L.ref_energy_filter_upper.argtypes = [C.c_void_p, C.c_int, C.c_double, O.P_i]
    L.ref_degeneracy_filter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    return L


# ---------------------------------------------------------------- 3j symbols

def fixture():
    with open(os.path.join(ROOT, "tests", "golden", "threej_sympy.json")) as f:
        return json.load(f)


def test_three_j_against_sympy_fixture():
    fx = fixture()
    worst = 0.0
    for jI, jF, kI, kF, v in fx["rows"]:
        got = O.three_j(jI, 1, jF, kI, kF - kI, -kF)
        if v == 0.0:
            assert abs(got) < 1e-15
        else:
            worst = max(worst, abs(got - v) / abs(v))
    assert worst < 1e-13, worst


def test_threej_table_layout_against_sympy():
    fx = fixture()
    jmax = fx["jmax"]
    T = O.threej_table(jmax)
    s = jmax + 1
    seen = np.zeros(T.size, bool)
    for jI, jF, kI, kF, v in fx["rows"]:
        i = jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * s * s * 3
        seen[i] = True
        assert T[i] == pytest.approx(v, rel=1e-13, abs=1e-15)
    assert np.all(T[~seen] == 0.0)  # uninitialised in the reference, zero here


def test_three_j_closed_forms():
    # (J 1 J; K 0 -K) = (-1)^(J-K+1) K / sqrt(J (J+1) (2J+1))
    for J in range(1, 21):
        for K in range(J + 1):
            want = (-1) ** (J - K + 1) * K / math.sqrt(J * (J + 1) * (2 * J + 1))
            assert O.three_j(J, 1, J, K, 0, -K) == pytest.approx(want, rel=1e-13, abs=1e-16)
    assert O.three_j(2, 3, 1, 1, -1, 0) == pytest.approx(2 * math.sqrt(210) / 105, rel=1e-14)
    assert O.three_j(0, 1, 1, 0, 0, 0) == pytest.approx(-1 / math.sqrt(3), rel=1e-14)
    assert O.three_j(10, 1, 11, 3, 1, -4) == pytest.approx(-math.sqrt(1265) / 253, rel=1e-13)
    assert O.three_j(11, 1, 10, 5, -1, -4) == pytest.approx(2 * math.sqrt(17710) / 1771, rel=1e-13)
    # selection rules give exact zeros
    assert O.three_j(3, 1, 5, 0, 0, 0) == 0.0
    assert O.three_j(3, 1, 3, 1, 1, -1) == 0.0


def test_fakt_scaled_factorial():
    for a in range(0, 40):
        want = math.factorial(a) / 10.0 ** a
        assert O.fakt(a) == pytest.approx(want, rel=1e-13)


def test_three_j_bitwise_against_reference_code():
    L = ref_lib()
    for j1 in range(0, 12):
        for j3 in range(max(j1 - 1, 0), j1 + 2):
            for k1 in range(-j1, j1 + 1):
                for k2 in (-1, 0, 1):
                    k3 = -(k1 + k2)
                    assert O.three_j(j1, 1, j3, k1, k2, k3) == L.ref_three_j(j1, 1, j3, k1, k2, k3)
    for a in range(30):
        assert O.fakt(a) == L.ref_fakt(float(a))
    for jmax in (1, 7, 20):
        n = (jmax + 1) ** 2 * 9
        t = np.zeros(n)
        v = np.zeros(n, np.int32)
        L.ref_precompute_threej(jmax, t.ctypes.data_as(O.P_d), v.ctypes.data_as(O.P_i))
        mine = O.threej_table(jmax)
        assert np.array_equal(mine[v == 1], t[v == 1])


# ----------------------------------------------------------- filters (truth tables)

def _filter_case(rng, nrep=4, degen=(1, 1, 1, 1), sym_type=1, reduced=0):
    nmodes = 3
    qlo = rng.integers(0, 2, size=nmodes)
    qhi = qlo + rng.integers(1, 4, size=nmodes)
    return dict(nrep=nrep, degen=np.array(degen, np.int32), sym_type=sym_type, nmodes=nmodes,
                zpe=float(rng.uniform(0, 100)),
                fw=np.array(sorted(rng.uniform(-2000, 6000, 2))),
                el=np.array(sorted(rng.uniform(-100, 6000, 2))),
                eu=np.array(sorted(rng.uniform(-100, 9000, 2))),
                jvals=np.array(sorted(rng.integers(0, 4, 2)), np.int32),
                gns=np.array(rng.choice([0.0, 1.0, 3.0], size=nrep)),
                pairs=np.array(rng.integers(1, 3, size=nrep), np.int32),
                reduced=reduced,
                ql=np.concatenate([qlo, qhi]).astype(np.int32),
                qu=np.concatenate([qlo, qhi + 1]).astype(np.int32))


def test_filters_match_reference_truth_tables():
    L = ref_lib()
    rng = np.random.default_rng(17)
    nfail = 0
    for case in range(40):
        sym = [(4, (1, 1, 1, 1), 1), (3, (1, 1, 2), 2), (6, (1, 1, 2, 1, 1, 2), 3)][case % 3]
        f = _filter_case(rng, nrep=sym[0], degen=sym[1], sym_type=sym[2], reduced=case % 2)
        igp = np.arange(f["nrep"], dtype=np.int32)
        for g in range(f["nrep"]):
            for h in range(f["nrep"]):
                if g != h and f["pairs"][g] == f["pairs"][h]:
                    igp[g] = h
        keep = []
        h = L.ref_filter_job(f["nrep"], f["degen"].ctypes.data_as(O.P_i), f["sym_type"],
                             f["nmodes"], f["zpe"], f["fw"].ctypes.data_as(O.P_d),
                             f["el"].ctypes.data_as(O.P_d), f["eu"].ctypes.data_as(O.P_d),
                             f["jvals"].ctypes.data_as(O.P_i), f["gns"].ctypes.data_as(O.P_d),
                             f["pairs"].ctypes.data_as(O.P_i), igp.ctypes.data_as(O.P_i),
                             f["reduced"], f["ql"].ctypes.data_as(O.P_i),
                             f["qu"].ctypes.data_as(O.P_i))
        job = O.OrJob(nrepres=f["nrep"], degen=f["degen"].ctypes.data_as(O.P_i),
                      sym_type=f["sym_type"], nmodes=f["nmodes"], zpe=f["zpe"],
                      freq_window=(C.c_double * 2)(*f["fw"]),
                      erange_lower=(C.c_double * 2)(*f["el"]),
                      erange_upper=(C.c_double * 2)(*f["eu"]),
                      jvals=(C.c_int * 2)(*f["jvals"]), gns=f["gns"].ctypes.data_as(O.P_d),
                      isym_pairs=f["pairs"].ctypes.data_as(O.P_i),
                      igamma_pair=igp.ctypes.data_as(O.P_i), reduced=f["reduced"],
                      quanta_lower=f["ql"].ctypes.data_as(O.P_i),
                      quanta_upper=f["qu"].ctypes.data_as(O.P_i))
        keep.append(job)
        lib = O.lib()
        for _ in range(300):
            jI, jF = (int(x) for x in rng.integers(0, 5, 2))
            eI, eF = (float(x) for x in rng.uniform(-200, 9000, 2))
            gI, gF = (int(x) for x in rng.integers(0, f["nrep"], 2))
            qI = rng.integers(0, 6, size=4).astype(np.int32)
            qF = rng.integers(0, 6, size=4).astype(np.int32)
            pI, pF = qI.ctypes.data_as(O.P_i), qF.ctypes.data_as(O.P_i)
            a = lib.or_intensity_filter(C.byref(job), jI, jF, eI, eF, gI, gF, pI, pF)
            b = L.ref_intensity_filter(h, jI, jF, eI, eF, gI, gF, pI, pF)
            nfail += a != b
            for idegI in range(2):
                for idegF in range(2):
                    a = lib.or_degeneracy_filter(C.byref(job), gI, gF, idegI, idegF)
                    b = L.ref_degeneracy_filter(h, gI, gF, idegI, idegF)
                    nfail += a != b
    assert nfail == 0


def test_find_igamma_pair_matches_reference_semantics():
    igp = synth.find_igamma_pair([1, 1, 2, 2])
    assert list(igp) == [1, 0, 3, 2]
    assert list(synth.find_igamma_pair([1, 1, 2])) == [1, 0, 2]
    with pytest.raises(ValueError):
        synth.find_igamma_pair([1, 1, 1, 2])


# ------------------------------------------------- half linestrength invariants

def _rot_bset(J, nvib):
    rng = np.random.default_rng(J)
    gvib = np.arange(nvib) % 4
    return synth.build_bset_c2v(J, gvib, rng)


def _half_ls_matrix(JI, JF, mu, nvib, jmax):
    """Dense M[rF, rI] of the oracle's half-ls, column by column."""
    keep = O._Keep()
    bI = O.make_bset(_rot_bset(JI, nvib), keep)
    bF = O.make_bset(_rot_bset(JF, nvib), keep)
    T = O.threej_table(jmax)
    mu = np.ascontiguousarray(mu, np.float64)
    dI, dF = bI.Maxcontracts, bF.Maxcontracts
    M = np.zeros((dF, dI))
    h = np.zeros(dF)
    for i in range(dI):
        e = np.zeros(dI)
        e[i] = 1.0
        O.lib().or_half_ls(C.byref(bI), C.byref(bF), mu.ctypes.data_as(O.P_d), nvib,
                           T.ctypes.data_as(O.P_d), jmax, e.ctypes.data_as(O.P_d),
                           h.ctypes.data_as(O.P_d), 0)
        M[:, i] = h
    return M, np.asarray(bI_arr(JI, nvib)), np.asarray(bI_arr(JF, nvib))


def bI_arr(J, nvib):
    b = _rot_bset(J, nvib)
    return b.iroot_correlat_j0


def test_rotational_sum_rule_scalar_mu():
    """sum_J' (2J'+1) M^T M = |mu|^2 * 1 for a rigid rotor with a constant
    dipole (completeness of the direction cosines). Fixes the 3j indexing,
    the 1/sqrt(2) factor and the K mask; the mixed-component case fixes the
    relative signs of the contributions to one rF."""
    for mu3 in ([1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0.3, -0.7, 1.1]):
        mu = np.array(mu3, np.float64)  # nvib = 1: mu_a is 1x1
        norm2 = float(np.sum(mu ** 2))
        for JI in range(0, 7):
            tot = 0.0
            for JF in (JI - 1, JI, JI + 1):
                if JF < 0:
                    continue
                M, _, _ = _half_ls_matrix(JI, JF, mu, 1, max(JI, JF))
                tot = tot + (2 * JF + 1) * M.T @ M
            np.testing.assert_allclose(tot, norm2 * np.eye(2 * JI + 1), atol=1e-14)


def test_sum_rule_nonsymmetric_mu_pins_transposition():
    """With nvib = 3 and a non-symmetric mu the sum rule reads
    1_rot (x) sum_a mu_a^T mu_a with mu_a[vF, vI]; a transposed operand gives
    mu_a mu_a^T instead."""
    rng = np.random.default_rng(5)
    nvib = 3
    mu = rng.standard_normal((3, nvib, nvib))  # C-order [a, vI, vF] = flipped layout
    mu_flat = mu.reshape(-1)
    mats = [mu[a].T for a in range(3)]  # element [vF, vI]
    want_v = sum(m.T @ m for m in mats)
    wrong_v = sum(m @ m.T for m in mats)
    assert np.abs(want_v - wrong_v).max() > 0.1
    for JI in range(0, 5):
        tot = 0.0
        vidx = None
        for JF in (JI - 1, JI, JI + 1):
            if JF < 0:
                continue
            M, vI, _ = _half_ls_matrix(JI, JF, mu_flat, nvib, max(JI, JF))
            vidx = vI
            tot = tot + (2 * JF + 1) * M.T @ M
        # expected[rI, rI'] = delta(slot) * want_v[v(rI), v(rI')]
        b = _rot_bset(JI, nvib)
        same_rot = (b.k[:, None] == b.k[None, :]) & (b.ktau[:, None] == b.ktau[None, :])
        exp = np.where(same_rot, want_v[vidx[:, None], vidx[None, :]], 0.0)
        np.testing.assert_allclose(tot, exp, atol=1e-13)


def test_reciprocity_fixes_both_sign_factors():
    """For a symmetric mu, M(J'<-J) = -M(J<-J')^T (the dipole operator is
    anti-symmetric in this real Wang basis); dropping either (-1)^sigma
    factor breaks it."""
    rng = np.random.default_rng(6)
    nvib = 2
    m = rng.standard_normal((3, nvib, nvib))
    m = m + np.transpose(m, (0, 2, 1))
    for JI in range(0, 6):
        for JF in (JI, JI + 1):
            jmax = max(JI, JF)
            A, _, _ = _half_ls_matrix(JI, JF, m.reshape(-1), nvib, jmax)
            B, _, _ = _half_ls_matrix(JF, JI, m.reshape(-1), nvib, jmax)
            np.testing.assert_allclose(B, -A.T, atol=1e-14)


def test_honl_london_parallel_band():
    """Nvib = 1, mu_z only: S(J', K, tau' <- J, K, tau) equals the textbook
    symmetric-top Honl-London factor mu_z^2 * {Q: (2J+1)K^2/(J(J+1)),
    R: ((J+1)^2 - K^2)/(J+1), P: (J^2 - K^2)/J} (summed over M), with
    S = (2J+1)(2J'+1) h^2 as in cuda_host.cu:1406."""
    mu = np.array([0.0, 0.0, 1.3])
    for JI in range(1, 8):
        for JF in (JI - 1, JI, JI + 1):
            M, _, _ = _half_ls_matrix(JI, JF, mu, 1, max(JI, JF))
            bI, bF = _rot_bset(JI, 1), _rot_bset(JF, 1)
            for i in range(len(bI.k)):
                K, tI = int(bI.k[i]), int(bI.ktau[i] & 1)
                for f in range(len(bF.k)):
                    if int(bF.k[f]) != K:
                        continue
                    tF = int(bF.ktau[f] & 1)
                    S = (2 * JI + 1) * (2 * JF + 1) * M[f, i] ** 2
                    if tF == tI:
                        assert S == pytest.approx(0.0, abs=1e-14)
                        continue
                    if JF == JI:
                        hl = (2 * JI + 1) * K * K / (JI * (JI + 1))
                    elif JF == JI + 1:
                        hl = ((JI + 1) ** 2 - K * K) / (JI + 1)
                    else:
                        hl = (JI * JI - K * K) / JI
                    assert S == pytest.approx(hl * mu[2] ** 2, rel=1e-13, abs=1e-14)


def test_noskip_equals_masked_bitwise():
    job = synth.make_job("C2V", nvib=16, jvals=(2, 3), seed=3)
    oj = O.OracleJob(job)
    rng = np.random.default_rng(1)
    for jI, jF in ((0, 1), (1, 0), (1, 1)):
        c = rng.standard_normal(job.bsets[jI].Maxcontracts)
        assert np.array_equal(oj.half_ls(jI, jF, c, noskip=False), oj.half_ls(jI, jF, c, noskip=True))


# ----------------------------------------------------- correlate and correlation

def test_correlate_is_orthogonal_signed_permutation_c2v():
    job = synth.make_job("C2V", nvib=24, jvals=(2, 3), seed=8)
    oj = O.OracleJob(job)
    for jind in (0, 1):
        b = job.bsets[jind]
        Rs = []
        for g in range(4):
            n = int(b.nsize[g])
            R = np.stack([oj.correlate(jind, 0, g, np.eye(n)[i]) for i in range(n)], axis=1)
            np.testing.assert_allclose(R.T @ R, np.eye(n), atol=0)
            assert set(np.unique(np.abs(R))) <= {0.0, 1.0}
            Rs.append(R)
        full = np.concatenate(Rs, axis=1)  # dimen x dimen, all symmetries
        np.testing.assert_allclose(full @ full.T, np.eye(b.Maxcontracts), atol=0)


def test_correlate_multi_term_c3v():
    job = synth.make_job("C3V", nvib=20, jvals=(2, 3), seed=9)
    oj = O.OracleJob(job)
    for jind in (0, 1):
        b = job.bsets[jind]
        cols = []
        for g, deg in enumerate(job.sym["degen"]):
            n = int(b.nsize[g])
            for ideg in range(deg):
                R = np.stack([oj.correlate(jind, ideg, g, np.eye(n)[i]) for i in range(n)], 1)
                cols.append(R)
        full = np.concatenate(cols, axis=1)
        assert full.shape == (b.Maxcontracts, b.Maxcontracts)
        np.testing.assert_allclose(full.T @ full, np.eye(b.Maxcontracts), atol=1e-14)
        assert (np.count_nonzero(full, axis=1) > 1).any()  # genuinely multi-term rows


def test_parseval_of_correlation():
    """With orthogonal eigenvector blocks and a signed-permutation R, the sum
    of A_F^2 over every upper level of the block equals ||R_F^T h||^2."""
    job = synth.make_job("C2V", nvib=16, jvals=(1, 2), seed=12, freq_window=(-1e9, 1e9))
    oj = O.OracleJob(job)
    lv = job.levels
    L = int(np.nonzero(lv["jval"] == 2)[0][3])  # J=2: both J'=1 and J'=2 allowed
    res = oj.run([L])
    g = int(lv["igamma"][L])
    u = job.vectors[(1, g)][lv["irec0"][L]]
    c = oj.correlate(1, 0, g, u)
    for jindF in (0, 1):
        h = oj.half_ls(1, jindF, c)
        gF = int(job.igamma_pair[g])
        mask = oj.correlate(jindF, 0, gF, np.ones(int(job.bsets[jindF].nsize[gF]))) != 0
        sel = (lv["jind"][res["ilevelF"]] == jindF)
        got = float(np.sum(res["amp"][sel, 0] ** 2))
        assert got == pytest.approx(float(np.sum(h[mask] ** 2)), rel=1e-12)


# ------------------------------------------------------------ constants and A4-A5

def test_constants_closed_forms():
    k = O.constants()
    pi = math.pi
    assert k["pi"] == pi
    assert k["A_coef_s_1"] == pytest.approx(64e-36 * pi ** 4 / (3 * 6.62606896e-27), rel=1e-15)
    assert k["intens_cm_mol"] == pytest.approx(
        8e-36 * pi ** 3 * 6.0221415e23 / (3 * 6.62606896e-27 * 2.99792458e10), rel=1e-15)
    assert k["boltz"] == pytest.approx(1.380658e-16, rel=1e-16)


def test_intensity_step_closed_form():
    """A4-A5 for one line recomputed from the emitted amplitude."""
    job = synth.make_job("C2V", nvib=16, jvals=(0, 1), seed=4)
    oj = O.OracleJob(job)
    r = oj.run()
    k = O.constants()
    beta = k["planck"] * k["vellgt"] / (k["boltz"] * job.temperature)
    lv = job.levels
    for i in range(0, len(r["S"]), max(1, len(r["S"]) // 50)):
        I, F = r["ilevelI"][i], r["ilevelF"][i]
        jI, jF = lv["jval"][I], lv["jval"][F]
        ls = r["amp"][i, 0] ** 2
        nu = lv["energy"][F] - lv["energy"][I]
        S = ls * job.gns[lv["igamma"][I]] * (2 * jI + 1) * (2 * jF + 1)
        A = k["A_coef_s_1"] * (2 * jI + 1) * ls * abs(nu) ** 3
        Int = S * k["intens_cm_mol"] * abs(nu) * math.exp(-(lv["energy"][I] - job.zpe) * beta) * \
            (1 - math.exp(-abs(nu) * beta)) / job.q_stat
        assert r["S"][i] == pytest.approx(S, rel=1e-14)
        assert r["A_einst"][i] == pytest.approx(A, rel=1e-13)
        assert r["I"][i] == pytest.approx(Int, rel=1e-13)


def _hitran_residual(job, r, A, I):
    """Relative residual of the Einstein-A / line-intensity relation of
    Simeckova et al. (JQSRT 98, 2006, the relation HITRAN uses):
        I / N_A = g' A / (8 pi c nu^2) * exp(-c2 E''/T) * (1 - exp(-c2 nu/T)) / Q
    with g' = gns (2J' + 1) the upper level's weight, E'' the lower energy above
    the ZPE, c2 = h c / k, and N_A, c, h, k of cuda_host.cu:12-18 (the
    reference's I is per mole). Lines with nu > 0 (absorption from I to F)."""
    k = O.constants()
    lv = job.levels
    iI, iF = r["ilevelI"], r["ilevelF"]
    nu = lv["energy"][iF] - lv["energy"][iI]
    m = nu > 0
    c2 = k["planck"] * k["vellgt"] / k["boltz"]
    T = job.temperature
    gup = np.asarray(job.gns)[lv["igamma"][iF]] * (2 * lv["jval"][iF] + 1)
    rhs = (gup[m] * A[m] / (8 * math.pi * k["vellgt"] * nu[m] ** 2) *
           np.exp(-c2 * (lv["energy"][iI][m] - job.zpe) / T) *
           (1 - np.exp(-c2 * nu[m] / T)) / job.q_stat)
    lhs = I[m] / k["avogno"]
    return np.abs(lhs - rhs) / np.abs(rhs), m


@pytest.mark.parametrize("sym,jvals", [("C2V", (1, 2)), ("C3V", (2, 3))])
@pytest.mark.parametrize("T", [296.0, 1000.0, 2500.0])
def test_einstein_a_against_line_intensity(sym, jvals, T):
    """The emitted A_einst and I satisfy the HITRAN Einstein-A / intensity
    relation line by line (an independent statement of A4-A5: it fixes the
    (2J+1) factor of A to the lower J, the nu^3 of A against the nu of I, the
    Boltzmann and stimulated-emission factors and the ratio of the two
    constants). Mistakes planted in the oracle's A or I fail it."""
    job = synth.make_job(sym, nvib=24, jvals=jvals, seed=9, temperature=T, e_span=6000.0)
    r = O.OracleJob(job).run()
    res, m = _hitran_residual(job, r, r["A_einst"], r["I"])
    assert m.sum() > 50
    assert res.max() <= 1e-12, res.max()
    lv = job.levels
    jI, jF = lv["jval"][r["ilevelI"]], lv["jval"][r["ilevelF"]]
    nu = np.abs(lv["energy"][r["ilevelF"]] - lv["energy"][r["ilevelI"]])
    dj = (jI != jF)[m]
    assert dj.any()
    # (2J'+1) in A instead of (2J+1): fails on every line with J' != J
    bad, _ = _hitran_residual(job, r, r["A_einst"] * (2 * jF + 1) / (2 * jI + 1), r["I"])
    assert (bad[dj] > 1e-3).all()
    # nu^4 (or nu^2) in A: fails wherever nu != 1
    bad, _ = _hitran_residual(job, r, r["A_einst"] * nu, r["I"])
    assert (bad[np.abs(nu[m] - 1) > 1e-3] > 1e-3).all()
    bad, _ = _hitran_residual(job, r, r["A_einst"] / nu, r["I"])
    assert (bad[np.abs(nu[m] - 1) > 1e-3] > 1e-3).all()
    # I without the stimulated-emission factor: visible on the lines where that
    # factor differs from 1 (low nu against kT)
    k = O.constants()
    beta = k["planck"] * k["vellgt"] / (k["boltz"] * T)
    stim = 1 - np.exp(-nu * beta)
    bad, _ = _hitran_residual(job, r, r["A_einst"], r["I"] / stim)
    sel = stim[m] < 1 - 1e-6
    assert sel.sum() > 10
    assert (bad[sel] > 1e-7).all()


def test_b8_q_branch_of_first_j_is_excluded():
    """trove_functions.cpp:1457: no line with J = J' = jvals[0]."""
    job = synth.make_job("C2V", nvib=16, jvals=(1, 2), seed=4)
    r = O.OracleJob(job).run()
    jI = job.levels["jval"][r["ilevelI"]]
    jF = job.levels["jval"][r["ilevelF"]]
    assert not np.any((jI == 1) & (jF == 1))
    assert np.any((jI == 2) & (jF == 2))
