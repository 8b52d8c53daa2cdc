"""Seeded synthetic TROVE-like inputs for the line-intensity path.

The reference ships no input data: it reads TROVE checkpoints (J=0 contracted
basis, vibrational dipole matrix, per-(J, symmetry) energies and eigenvectors;
trove_functions.cpp:24-28). This module builds the same objects in memory, in
the layouts the reference's readers produce:

* basis per J: the TO_bset_contrT fields after bset_contr_factory
  (trove_functions.cpp:401-679), correlate_index (:680-858) and
  compute_ijterms (:927-966);
* the flipped dipole, mu_a(vF, vI) at vF + vI*N + a*N^2
  (read_dipole_flipped, trove_functions.cpp:113-205);
* the energy-sorted level table of read_eigenvalues (:1074-1429);
* eigenvector blocks, record irec (0-based here) holding nsize doubles
  (cuda_host.cu:1196, 1262).

Symmetry groups: C2v(M) (A1, A2, B1, B2) and C3v(M) (A1, A2, E), with the
labels and degeneracies of fields.cpp:14-68.

C2v construction: the vibrational function v has symmetry g_v, the rotational
Wang function (K, tau) has symmetry rot(K, tau), and the primitive (v, K, tau)
has g_v x rot(K, tau) (the C2v product table is XOR of 2-bit codes). Every
symmetrised function is one primitive times a random sign, so R is a signed
permutation.

C3v construction: a case groups one vibrational function (A1, A2, or an E
pair) with one K (tau = 0, 1 for K > 0). Its lambdas (vibrational component x
tau) are mapped to symmetrised functions by a random orthogonal matrix whose
rows are assigned to (symmetry, component) slots according to the product
decomposition, so R has multi-term rows and R^T R = 1 on each case.

Eigenvectors: QR of a seeded N(0,1) block with diag(R) > 0. Energies: sorted
uniform per block. Dipole: magnitude 10^(-|v-v'|/3) * LogNormal(0,1), random
sign, symmetric unless asked otherwise.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np

C2V = dict(name="C2V(M)", nrep=4, degen=[1, 1, 1, 1], labels=["A1", "A2", "B1", "B2"],
           sym_type=1)
C3V = dict(name="C3V(M)", nrep=3, degen=[1, 1, 2], labels=["A1", "A2", "E"], sym_type=2)
D3H = dict(name="D3H(M)", nrep=6, degen=[1, 1, 2, 1, 1, 2],
           labels=["A1'", "A2'", "E'", 'A1"', 'A2"', 'E"'], sym_type=3)


@dataclass
class Bset:
    jval: int
    Maxsymcoeffs: int
    Maxcontracts: int
    mat_size: int
    icontr2icase: np.ndarray      # int32[2*dimen]
    iroot_correlat_j0: np.ndarray  # int32[dimen]
    k: np.ndarray                 # int32[dimen]
    ktau: np.ndarray              # int32[dimen]
    N: np.ndarray                 # int32[nrep*ncases], N[g*ncases + c]
    Ntotal: np.ndarray            # int32[nrep]
    ijterms: np.ndarray           # int32[nrep*ncases], ijterms[c + g*ncases]
    repres: list                  # per g: float64[Ntotal*degen*mat_size]
    nsize: np.ndarray             # int32[nrep]


@dataclass
class Job:
    sym: dict
    nmodes: int
    nclasses: int
    zpe: float
    temperature: float
    q_stat: float
    freq_window: tuple
    erange_lower: tuple
    erange_upper: tuple
    jvals: tuple
    gns: np.ndarray
    isym_pairs: np.ndarray
    igamma_pair: np.ndarray
    reduced: int
    quanta_lower: np.ndarray      # int32[2*nmodes]
    quanta_upper: np.ndarray
    thresh_linestrength: float
    thresh_intensity: float
    bsets: list                   # [Bset for jvals[0], Bset for jvals[1]]
    nvib: int
    dipole_flipped: np.ndarray    # float64[3*N*N]
    jmax: int
    levels: dict                  # column arrays, energy sorted
    vectors: dict                 # (jind, g) -> float64[nrec, nsize] (or a device tensor)
    meta: dict = field(default_factory=dict)

    @property
    def nlevels(self) -> int:
        return int(self.levels["energy"].shape[0])


def find_igamma_pair(isym_pairs):
    """trove_functions.cpp:1041-1072: partner symmetry with the same
    selection value; raises where the reference exits."""
    n = len(isym_pairs)
    out = np.arange(n, dtype=np.int32)
    for gI in range(n):
        m = 0
        for gF in range(n):
            if gI != gF and isym_pairs[gI] == isym_pairs[gF]:
                out[gI] = gF
                m += 1
        if m > 1:
            raise ValueError("selection rules do not come in pairs")
    return out


def _rot_list(J):
    """(K, tau) of the Wang functions of J: K=0 has tau = J mod 2
    (trove_functions.cpp:1286-1302)."""
    out = [(0, J % 2)]
    for K in range(1, J + 1):
        out += [(K, 0), (K, 1)]
    return out


def _c2v_rot_sym(K, tau):
    return ((0, 1), (3, 2))[K % 2][tau]


def _c3v_rot_sym(K, tau):
    """A1/A2 for K = 0 mod 3 (tau picks), E (code 2) otherwise."""
    if K % 3 == 0:
        return tau
    return 2


def _rng(seed, *tags):
    ss = np.random.SeedSequence([int(seed)] + [int(t) for t in tags])
    return np.random.Generator(np.random.PCG64(ss))


def build_bset_c2v(J, gvib, rng, shuffle_cases=False):
    """Basis of J for C2v(M): cases (v, K), lambdas tau."""
    nvib = len(gvib)
    cases = [(v, K) for v in range(nvib) for K in range(J + 1)]
    if shuffle_cases:
        perm = rng.permutation(len(cases))
        cases = [cases[i] for i in perm]
    ncases = len(cases)
    nrep = 4
    ic2 = []
    v_of, k_of, kt_of, sym_of = [], [], [], []
    lam = []
    for ic, (v, K) in enumerate(cases):
        taus = [J % 2] if K == 0 else [0, 1]
        for il, tau in enumerate(taus):
            ic2.append((ic, il))
            v_of.append(v)
            k_of.append(K)
            kt_of.append(2 * K + tau)
            sym_of.append(gvib[v] ^ _c2v_rot_sym(K, tau))
        lam.append(len(taus))
    dimen = len(ic2)
    icontr2icase = np.empty(2 * dimen, np.int32)
    icontr2icase[:dimen] = [a for a, _ in ic2]
    icontr2icase[dimen:] = [b for _, b in ic2]
    mat_size = 2 if J > 0 else 1
    Nmat = np.zeros((nrep, ncases), np.int32)
    r = 0
    case_syms = []
    for ic in range(ncases):
        ss = sym_of[r:r + lam[ic]]
        case_syms.append(ss)
        for g in ss:
            Nmat[g, ic] += 1
        r += lam[ic]
    ijterms = np.zeros((nrep, ncases), np.int32)
    for g in range(nrep):
        ijterms[g, 1:] = np.cumsum(Nmat[g, :-1])
    Ntotal = Nmat.sum(axis=1).astype(np.int32)
    repres = []
    signs = rng.choice(np.array([-1.0, 1.0]), size=dimen)
    for g in range(nrep):
        repres.append(np.zeros(int(Ntotal[g]) * 1 * mat_size, np.float64))
    r = 0
    for ic in range(ncases):
        seen = [0] * nrep
        for ib, g in enumerate(case_syms[ic]):
            t = ijterms[g, ic] + seen[g]
            seen[g] += 1
            repres[g][t + ib * 1 * Ntotal[g]] = signs[r]
            r += 1
    return Bset(jval=J, Maxsymcoeffs=ncases, Maxcontracts=dimen, mat_size=mat_size,
                icontr2icase=icontr2icase,
                iroot_correlat_j0=np.asarray(v_of, np.int32), k=np.asarray(k_of, np.int32),
                ktau=np.asarray(kt_of, np.int32), N=Nmat.reshape(-1).copy(),
                Ntotal=Ntotal, ijterms=ijterms.reshape(-1).copy(), repres=repres,
                nsize=Ntotal.copy())


def _c3v_product_slots(gv, gr):
    """Symmetrised slots (symmetry, component) of the product of a vibrational
    function of symmetry gv with a rotational pair/single of symmetry gr."""
    if gv < 2 and gr < 2:
        return [(gv ^ gr, 0)]
    if gv < 2 or gr < 2:
        return [(2, 0), (2, 1)]
    return [(0, 0), (1, 0), (2, 0), (2, 1)]


def build_bset_c3v(J, vib_units, rng):
    """Basis of J for C3v(M). vib_units: list of (symmetry, [v indices]) with
    one index for A1/A2 and two (the E pair) for E."""
    nrep = 3
    degen = [1, 1, 2]
    cases = []  # (unit, K, lambdas[(v, tau)], slots)
    for u, (gv, vs) in enumerate(vib_units):
        for K in range(J + 1):
            if K == 0:
                taus = [J % 2]
                gr = _c3v_rot_sym(0, J % 2)
            elif K % 3 == 0:
                taus = [0, 1]
                gr = None
            else:
                taus = [0, 1]
                gr = 2
            lams = [(v, tau) for v in vs for tau in taus]
            if gr is None:
                # K = 0 mod 3, K > 0: tau = 0 -> A1, tau = 1 -> A2 rotational singles
                slots = []
                for tau in taus:
                    slots += _c3v_product_slots(gv, _c3v_rot_sym(K, tau))
            else:
                slots = _c3v_product_slots(gv, gr)
            assert len(slots) == len(lams) * 1 or len(slots) == len(lams), (slots, lams)
            cases.append((u, K, lams, slots))
    ncases = len(cases)
    mat_size = max(len(c[2]) for c in cases)
    ic2, v_of, k_of, kt_of = [], [], [], []
    for ic, (u, K, lams, slots) in enumerate(cases):
        for il, (v, tau) in enumerate(lams):
            ic2.append((ic, il))
            v_of.append(v)
            k_of.append(K)
            kt_of.append(2 * K + tau)
    dimen = len(ic2)
    icontr2icase = np.empty(2 * dimen, np.int32)
    icontr2icase[:dimen] = [a for a, _ in ic2]
    icontr2icase[dimen:] = [b for _, b in ic2]
    # slot -> (symmetry, function index within the case, component)
    Nmat = np.zeros((nrep, ncases), np.int32)
    assigned = []
    for ic, (_, _, _, slots) in enumerate(cases):
        fc = [0, 0, 0]
        asg = []
        for g, comp in slots:
            if g < 2:
                asg.append((g, fc[g], 0))
                fc[g] += 1
            elif comp == 0:
                asg.append((2, fc[2], 0))
            else:
                asg.append((2, fc[2], 1))
                fc[2] += 1
        assigned.append(asg)
        Nmat[:, ic] = fc
    ijterms = np.zeros((nrep, ncases), np.int32)
    for g in range(nrep):
        ijterms[g, 1:] = np.cumsum(Nmat[g, :-1])
    Ntotal = Nmat.sum(axis=1).astype(np.int32)
    repres = [np.zeros(int(Ntotal[g]) * degen[g] * mat_size, np.float64) for g in range(nrep)]
    for ic, (_, _, lams, slots) in enumerate(cases):
        m = len(lams)
        q, rr = np.linalg.qr(rng.standard_normal((m, m)))
        q = q * np.sign(np.diag(rr))[None, :]
        # column i of q is the symmetrised slot i in terms of the lambdas
        for i, (g, fi, comp) in enumerate(assigned[ic]):
            t = ijterms[g, ic] + fi
            for ib in range(m):
                repres[g][t + comp * Ntotal[g] + ib * degen[g] * Ntotal[g]] = q[ib, i]
    return Bset(jval=J, Maxsymcoeffs=ncases, Maxcontracts=dimen, mat_size=mat_size,
                icontr2icase=icontr2icase,
                iroot_correlat_j0=np.asarray(v_of, np.int32), k=np.asarray(k_of, np.int32),
                ktau=np.asarray(kt_of, np.int32), N=Nmat.reshape(-1).copy(),
                Ntotal=Ntotal, ijterms=ijterms.reshape(-1).copy(), repres=repres,
                nsize=Ntotal.copy())


def direct_sum_bsets(bA, bB, nrepA):
    """One basis whose irreps 0..nrepA-1 are those of bA and the next nrepA
    those of bB (cases of bA, then of bB): a structurally valid basis of a
    group with twice the irreps, used for the D3H(M) generator (the ' and "
    functions of two C3v-like bases). Every field keeps the layout of the
    reference's bset_contr (N[g*ncases + c], repres[t + comp*Ntotal + ib*degen*Ntotal])."""
    nA = bA.Maxsymcoeffs
    nB = bB.Maxsymcoeffs
    nc = nA + nB
    dA, dB = bA.Maxcontracts, bB.Maxcontracts
    icA, icB = bA.icontr2icase, bB.icontr2icase
    ic = np.concatenate([icA[:dA], icB[:dB] + nA, icA[dA:], icB[dB:]]).astype(np.int32)
    ms = max(bA.mat_size, bB.mat_size)
    N = np.zeros((2 * nrepA, nc), np.int32)
    N[:nrepA, :nA] = bA.N.reshape(nrepA, nA)
    N[nrepA:, nA:] = bB.N.reshape(nrepA, nB)
    ijterms = np.zeros((2 * nrepA, nc), np.int32)
    for g in range(2 * nrepA):
        ijterms[g, 1:] = np.cumsum(N[g, :-1])
    Ntotal = N.sum(axis=1).astype(np.int32)
    degen = [1, 1, 2] * 2
    repres = []
    for g in range(2 * nrepA):
        src, sg = (bA, g) if g < nrepA else (bB, g - nrepA)
        out = np.zeros(int(Ntotal[g]) * degen[g] * ms, np.float64)
        old = src.repres[sg]
        out[:old.size] = old  # ib * degen * Ntotal: the same offsets for ib < the old mat_size
        repres.append(out)
    return Bset(jval=bA.jval, Maxsymcoeffs=nc, Maxcontracts=dA + dB, mat_size=ms,
                icontr2icase=ic,
                iroot_correlat_j0=np.concatenate([bA.iroot_correlat_j0, bB.iroot_correlat_j0]),
                k=np.concatenate([bA.k, bB.k]), ktau=np.concatenate([bA.ktau, bB.ktau]),
                N=N.reshape(-1).copy(), Ntotal=Ntotal, ijterms=ijterms.reshape(-1).copy(),
                repres=repres, nsize=Ntotal.copy())


def make_dipole(nvib, rng, symmetric=True):
    """Flipped dipole mu_flip[vF + vI*N + a*N^2] = file_a[vI, vF]."""
    v = np.arange(nvib)
    dv = np.abs(v[:, None] - v[None, :]).astype(np.float64)
    out = np.empty((3, nvib, nvib), np.float64)
    for a in range(3):
        mag = 10.0 ** (-dv / 3.0) * np.exp(rng.standard_normal((nvib, nvib)))
        sgn = rng.choice(np.array([-1.0, 1.0]), size=(nvib, nvib))
        m = mag * sgn
        if symmetric:
            m = np.triu(m) + np.triu(m, 1).T
        out[a] = m  # file_a[vI, vF] as a C array m[vI, vF]
    # flipped: element (vF, vI) at vF + vI*N, i.e. C-order array [a, vI, vF]
    return np.ascontiguousarray(out).reshape(-1)


def orthogonal_block(n, rng):
    """Householder QR of an N(0,1) block, diag(R) > 0: columns orthonormal."""
    if n == 0:
        return np.zeros((0, 0))
    a = rng.standard_normal((n, n))
    q, r = np.linalg.qr(a)
    q *= np.sign(np.diag(r))[None, :]
    return q


def _levels_for_blocks(job_syms, jvals, bsets, nmodes, nclasses, e_lo, e_span, rng,
                       ground_zero=True):
    cols = {k: [] for k in ("energy", "jind", "jval", "igamma", "ndeg", "ilevel", "irec0")}
    quanta, cg, normal = [], [], []
    nrep = job_syms["nrep"]
    for jind, J in enumerate(jvals):
        for g in range(nrep):
            n = int(bsets[jind].nsize[g])
            e = np.sort(rng.uniform(e_lo, e_lo + e_span, size=n))
            if ground_zero and J == 0 and g == 0 and n > 0:
                e[0] = 0.0
            cols["energy"].append(e)
            cols["jind"].append(np.full(n, jind))
            cols["jval"].append(np.full(n, J))
            cols["igamma"].append(np.full(n, g))
            cols["ndeg"].append(np.full(n, job_syms["degen"][g]))
            cols["ilevel"].append(np.arange(n))
            cols["irec0"].append(np.arange(n))
            q = np.zeros((n, nmodes + 1), np.int32)
            q[:, 0] = rng.integers(0, 2 * J + 1, size=n)
            q[:, 1:] = rng.integers(0, 6, size=(n, nmodes))
            quanta.append(q)
            cg.append(rng.integers(0, nrep, size=(n, nclasses + 1)).astype(np.int32))
            normal.append(rng.integers(0, 6, size=(n, nmodes + 1)).astype(np.int32))
    lv = {k: np.concatenate(v) for k, v in cols.items()}
    lv["quanta"] = np.concatenate(quanta)
    lv["cgamma"] = np.concatenate(cg)
    lv["normal"] = np.concatenate(normal)
    # energy sort; ties (probability zero) broken by (jind, igamma, ilevel)
    order = np.lexsort((lv["ilevel"], lv["igamma"], lv["jind"], lv["energy"]))
    out = {}
    for k, v in lv.items():
        v = v[order]
        if k == "energy":
            out[k] = np.ascontiguousarray(v, np.float64)
        else:
            out[k] = np.ascontiguousarray(v, np.int32)
    # krot/taurot from quanta[0] (trove_functions.cpp:1387-1388)
    krot = np.where(out["quanta"][:, 0] == 0, 0, (out["quanta"][:, 0] + 1) // 2)
    taurot = np.where(out["quanta"][:, 0] == 0, out["jval"] % 2, (out["quanta"][:, 0] + 1) % 2)
    out["krot"] = krot.astype(np.int32)
    out["taurot"] = taurot.astype(np.int32)
    return out


def make_job(sym="C2V", nvib=200, jvals=(0, 1), seed=1234, *, e_span=8000.0,
             temperature=296.0, q_stat=174.5813, freq_window=(0.0, 8000.0),
             erange_lower=(0.0, 1e9), erange_upper=(0.0, 1e9), gns=None,
             isym_pairs=None, reduced=0, thresh_linestrength=0.0, thresh_intensity=0.0,
             symmetric_mu=True, shuffle_cases=False, vectors="numpy", nmodes=3):
    """Build a synthetic job. vectors='numpy' builds eigenvectors on the host;
    vectors='none' leaves them for the caller (e.g. built on the GPU)."""
    su = sym.upper()
    sy = dict(C2V if su.startswith("C2V") else D3H if su.startswith("D3H") else C3V)
    nrep = sy["nrep"]
    jv = tuple(int(j) for j in jvals)
    if len(jv) < 2 or (len(jv) > 2 and any(b <= a for a, b in zip(jv, jv[1:]))):
        raise ValueError("jvals: the deck's two J values, or an ascending list")
    j0, j1 = jv[0], jv[1]
    rng_b = _rng(seed, 1)
    if sy["sym_type"] == 1:
        if nvib % 4:
            raise ValueError("C2v generator wants nvib divisible by 4")
        gvib = np.repeat(np.arange(4), nvib // 4)
        gvib = gvib[rng_b.permutation(nvib)]
        bsets = [build_bset_c2v(J, gvib, _rng(seed, 2, J), shuffle_cases) for J in jv]
        units = [(int(gvib[v]), [v]) for v in range(nvib)]
        gns_d = [1.0, 1.0, 3.0, 3.0]
        pairs_d = [1, 1, 2, 2]
        nclasses = 1
    elif sy["sym_type"] == 3:
        # D3h: C3v-like units (A1, A2 singles, E pairs), each ' or "; the basis
        # is the direct sum of the C3v bases of the ' and of the " units, and
        # the ' units take the first vibrational indices so that the J = 0
        # functions of that basis come in v order
        rk, rp = _rng(seed, 3), _rng(seed, 7)
        specs, n = [], 0
        while n < nvib:
            kd, p = int(rk.integers(0, 3)), int(rp.integers(0, 2))
            if kd == 2 and n + 2 <= nvib:
                specs.append((2, 2, p))
                n += 2
            else:
                specs.append((kd if kd < 2 else 0, 1, p))
                n += 1
        units, v = [], 0
        for want in (0, 1):
            for kd, sz, p in specs:
                if p == want:
                    units.append((kd, list(range(v, v + sz)), p))
                    v += sz
        bsets = []
        for J in jv:
            b1 = build_bset_c3v(J, [(g, vs) for g, vs, p in units if p == 0], _rng(seed, 2, J))
            b2 = build_bset_c3v(J, [(g, vs) for g, vs, p in units if p == 1],
                                _rng(seed, 8, J))
            bsets.append(direct_sum_bsets(b1, b2, 3))
        units = [(g + 3 * p, vs) for g, vs, p in units]
        gns_d = [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]
        pairs_d = [1, 2, 3, 2, 1, 3]  # A1' - A2", A2' - A1", E' - E"
        nclasses = 1
    else:
        # C3v units: A1, A2 singles and E pairs; nvib counts functions
        units, v = [], 0
        kinds = _rng(seed, 3).integers(0, 3, size=nvib)
        while v < nvib:
            kd = int(kinds[v])
            if kd == 2 and v + 1 < nvib:
                units.append((2, [v, v + 1]))
                v += 2
            else:
                units.append((kd if kd < 2 else 0, [v]))
                v += 1
        bsets = [build_bset_c3v(J, units, _rng(seed, 2, J)) for J in jv]
        gns_d = [1.0, 1.0, 1.0]
        pairs_d = [1, 1, 2]
        nclasses = 1
    gns = np.asarray(gns if gns is not None else gns_d, np.float64)
    pairs = np.asarray(isym_pairs if isym_pairs is not None else pairs_d, np.int32)
    igp = find_igamma_pair(pairs)
    mu = make_dipole(nvib, _rng(seed, 4), symmetric=symmetric_mu)
    levels = _levels_for_blocks(sy, jv, bsets, nmodes, nclasses, 0.0, e_span,
                                _rng(seed, 5))
    vecs = {}
    if vectors == "numpy":
        for jind, J in enumerate(jv):
            for g in range(nrep):
                n = int(bsets[jind].nsize[g])
                q = orthogonal_block(n, _rng(seed, 6, J, g))
                # record irec0 = column irec0 of q; stored record-major
                vecs[(jind, g)] = np.ascontiguousarray(q.T)
    big = 1 << 30
    ql = np.array([0] * nmodes + [big] * nmodes, np.int32)
    return Job(sym=sy, nmodes=nmodes, nclasses=nclasses, zpe=0.0, temperature=temperature,
               q_stat=q_stat, freq_window=tuple(freq_window), erange_lower=tuple(erange_lower),
               erange_upper=tuple(erange_upper), jvals=jv, gns=gns, isym_pairs=pairs,
               igamma_pair=igp, reduced=int(reduced), quanta_lower=ql.copy(),
               quanta_upper=ql.copy(), thresh_linestrength=float(thresh_linestrength),
               thresh_intensity=float(thresh_intensity), bsets=bsets, nvib=nvib,
               dipole_flipped=mu, jmax=max(jv), levels=levels, vectors=vecs,
               meta=dict(seed=seed, e_span=e_span, vib_units=units,
                         shuffle_cases=bool(shuffle_cases)))


def drop_level_filter(job: Job):
    """Apply the read-time level filter of read_eigenvalues
    (trove_functions.cpp:1025-1028, 1202): isym_do and E - ZPE within
    erange = [min(lower, upper) lo, max(lower, upper) hi] (fields.cpp:276-277).
    Returns a job whose level table holds only the kept levels."""
    lo = min(job.erange_lower[0], job.erange_upper[0])
    hi = max(job.erange_lower[1], job.erange_upper[1])
    e = job.levels["energy"] - job.zpe
    isym_do = job.gns > 1e-8
    keep = isym_do[job.levels["igamma"]] & (e >= lo) & (e <= hi)
    job.levels = {k: np.ascontiguousarray(v[keep]) for k, v in job.levels.items()}
    return job



# -- eigenvector blocks built on the GPU (config-size workloads) --------------

def _right_solve(torch, r, q, b=2048):
    """X with X R = Q, R upper triangular: the library triangular solve, or
    (if it fails) column blocks of GEMMs with b x b diagonal-block inverses."""
    try:
        return torch.linalg.solve_triangular(r, q, upper=True, left=False)
    except RuntimeError:
        pass
    x = torch.empty_like(q)
    for j in range(0, r.shape[0], b):
        e = min(j + b, r.shape[0])
        rhs = q[:, j:e] - x[:, :j] @ r[:j, j:e] if j else q[:, j:e]
        x[:, j:e] = rhs @ torch.linalg.inv(r[j:e, j:e])
    return x


def gpu_orthogonal(torch, n, seed):
    """Q of A = QR for a seeded N(0,1) block A on the GPU, diag(R) > 0 (the
    unique such factor), by CholeskyQR2: two passes of R = chol(Q^T Q),
    Q <- Q R^-1 (DGEMM, Cholesky, triangular solve), as orthogonal as
    Householder QR (|Q^T Q - 1| ~ 1e-14 at n = 20480) in a third of the time.
    Householder QR if a Cholesky factorisation fails (A too ill-conditioned)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    a = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    q = a
    for _ in range(2):
        r, info = torch.linalg.cholesky_ex(q.t() @ q, upper=True)
        if int(info) != 0:
            break
        q = _right_solve(torch, r, q)
        del r
    else:
        return q
    q, r = torch.linalg.qr(a)
    q *= torch.sign(torch.diagonal(r))[None, :]
    return q


def device_blocks(torch, job, seed, tag=None, cache_dir=None, log=None):
    """Eigenvector block of every (J, symmetry) of the job on the GPU, as a
    dict {(jind, g): tensor [nrec, nsize]}: record irec0 is column irec0 of a
    Haar-random orthogonal matrix (gpu_orthogonal). With cache_dir (and a
    tag naming the workload) the blocks are saved and reused across
    processes."""
    import os
    import time
    t = time.time()
    dev = {}
    for jind in range(len(job.jvals)):
        for g in range(job.sym["nrep"]):
            n = int(job.bsets[jind].nsize[g])
            path = (os.path.join(cache_dir, f"u_{tag}_{jind}_{g}.npy")
                    if cache_dir and tag else None)
            if path and os.path.exists(path):
                dev[(jind, g)] = torch.from_numpy(np.load(path)).cuda()
                continue
            q = gpu_orthogonal(torch, n, seed * 1000 + jind * 10 + g)
            dev[(jind, g)] = q.t().contiguous()
            del q
            if path:
                os.makedirs(cache_dir, exist_ok=True)
                np.save(path + ".tmp.npy", dev[(jind, g)].cpu().numpy())
                os.replace(path + ".tmp.npy", path)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if log:
        gb = sum(v.numel() for v in dev.values()) * 8 / 1e9
        log(f"eigenvector blocks ready {time.time() - t:.1f}s, {gb:.1f} GB "
            f"(CholeskyQR2 on the GPU, or the cache)")
    return dev


def config_job(name, nvib=None, jmax=None, seed=None):
    """The synthetic jobs of the measured workloads, without eigenvectors
    (pair them with device_blocks):
      "c2":  C2v(M), Nvib = 2000, J 10 <-> 11, level density such that about
             3000 levels lie below 10,000 cm^-1, lower levels E <= 10,000,
             upper unrestricted, window 0-10,000 cm^-1, T = 296 K;
      "c3":  C2v(M) J = 0..jmax (20) in one context, the same level density
             per block as "c2";
      "c4":  C3v(M) (A1, A2, E), Nvib = 3000, J = 0..jmax (15), same density.
    Returns (job, seed)."""
    lower_max, window = 10000.0, (0.0, 10000.0)
    density_span = lower_max * 88000 / 3000   # e_span of the J 10 <-> 11 blocks
    if name == "c2":
        nvib = nvib or 2000
        seed = 20240 if seed is None else seed
        jv = (10, 11)
        nlev = nvib * (2 * jv[0] + 1) + nvib * (2 * jv[1] + 1)
        job = make_job("C2V", nvib=nvib, jvals=jv, seed=seed,
                       e_span=lower_max * nlev / 3000, freq_window=window,
                       erange_lower=(0.0, lower_max), erange_upper=(0.0, 1e9),
                       vectors="none")
    elif name in ("c3", "c4"):
        sym = "C2V" if name == "c3" else "C3V"
        nvib = nvib or (2000 if name == "c3" else 3000)
        jmax = (20 if name == "c3" else 15) if jmax is None else jmax
        seed = 20250 if seed is None else seed
        job = make_job(sym, nvib=nvib, jvals=tuple(range(jmax + 1)), seed=seed,
                       e_span=density_span, freq_window=window,
                       erange_lower=(0.0, lower_max), erange_upper=(0.0, 1e9),
                       vectors="none")
    else:
        raise ValueError(name)
    return job, seed
