"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The product code never imports this module.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ci = C.c_int
cd = C.c_double
P_i = C.POINTER(C.c_int)
P_d = C.POINTER(C.c_double)


def amp_columns(m):
    """Columns of the oracle's 4-wide amplitude rows (entry idegI + 2*idegF)
    in the order of a line's amplitudes for largest degeneracy m: entry
    idegI + m*idegF, the reference's line_str layout per final state with
    sym_maxdegen = m (cuda_host.cu:1467-1471)."""
    return [dI + 2 * dF for dF in range(m) for dI in range(m)]


class OrBset(C.Structure):
    _fields_ = [("jval", ci), ("Maxsymcoeffs", ci), ("Maxcontracts", ci), ("mat_size", ci),
                ("icontr2icase", P_i), ("iroot_correlat_j0", P_i), ("k", P_i), ("ktau", P_i),
                ("N", P_i), ("Ntotal", P_i), ("ijterms", P_i), ("repres", C.POINTER(P_d)),
                ("nsize", P_i)]


class OrLevel(C.Structure):
    _fields_ = [("energy", cd), ("jind", ci), ("jval", ci), ("igamma", ci), ("ndeg", ci),
                ("ilevel", ci), ("irec0", ci), ("quanta", P_i)]


class OrJob(C.Structure):
    _fields_ = [("nrepres", ci), ("degen", P_i), ("sym_type", ci), ("nmodes", ci),
                ("zpe", cd), ("temperature", cd), ("q_stat", cd),
                ("freq_window", cd * 2), ("erange_lower", cd * 2), ("erange_upper", cd * 2),
                ("jvals", ci * 2), ("gns", P_d), ("isym_pairs", P_i), ("igamma_pair", P_i),
                ("reduced", ci), ("quanta_lower", P_i), ("quanta_upper", P_i),
                ("thresh_linestrength", cd), ("thresh_intensity", cd),
                ("bset", C.POINTER(OrBset) * 2), ("nvib", ci), ("dipole_flipped", P_d),
                ("jmax", ci), ("threej", P_d), ("nlevels", ci), ("levels", C.POINTER(OrLevel)),
                ("vectors", C.POINTER(P_d))]


class OrLines(C.Structure):
    _fields_ = [("n", C.c_longlong), ("ilevelI", P_i), ("ilevelF", P_i), ("nu", P_d),
                ("S", P_d), ("A_einst", P_d), ("I", P_d), ("amp", P_d), ("amp_abs", P_d)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.or_fakt.restype = cd
        L.or_fakt.argtypes = [cd]
        L.or_three_j.restype = cd
        L.or_three_j.argtypes = [ci] * 6
        L.or_precompute_threej.argtypes = [ci, P_d]
        L.or_correlate.argtypes = [C.POINTER(OrBset), P_i, ci, ci, ci, P_d, P_d]
        L.or_half_ls.argtypes = [C.POINTER(OrBset), C.POINTER(OrBset), P_d, ci, P_d, ci, P_d,
                                 P_d, ci]
        L.or_constants.argtypes = [P_d]
        L.or_run.argtypes = [C.POINTER(OrJob), ci, P_i, ci, ci, C.POINTER(OrLines)]
        L.or_run.restype = ci
        L.or_free_lines.argtypes = [C.POINTER(OrLines)]
        L.or_max_threads.restype = ci
        L.or_intensity_filter.argtypes = [C.POINTER(OrJob), ci, ci, cd, cd, ci, ci, P_i, P_i]
        L.or_intensity_filter.restype = ci
        L.or_degeneracy_filter.argtypes = [C.POINTER(OrJob), ci, ci, ci, ci]
        L.or_degeneracy_filter.restype = ci
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


class _Keep:
    """Holds the numpy arrays and ctypes objects a struct points into."""

    def __init__(self):
        self.items = []

    def i32(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.items.append(a)
        return _p(a, P_i)

    def f64(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.items.append(a)
        return _p(a, P_d)


def make_bset(b, keep: _Keep) -> OrBset:
    reps = (P_d * len(b.repres))(*[keep.f64(r) for r in b.repres])
    keep.items.append(reps)
    s = OrBset(jval=b.jval, Maxsymcoeffs=b.Maxsymcoeffs, Maxcontracts=b.Maxcontracts,
               mat_size=b.mat_size, icontr2icase=keep.i32(b.icontr2icase),
               iroot_correlat_j0=keep.i32(b.iroot_correlat_j0), k=keep.i32(b.k),
               ktau=keep.i32(b.ktau), N=keep.i32(b.N), Ntotal=keep.i32(b.Ntotal),
               ijterms=keep.i32(b.ijterms), repres=C.cast(reps, C.POINTER(P_d)),
               nsize=keep.i32(b.nsize))
    keep.items.append(s)
    return s


def threej_table(jmax):
    t = np.zeros((jmax + 1) ** 2 * 9, np.float64)
    lib().or_precompute_threej(jmax, _p(t, P_d))
    return t


class OracleJob:
    """An or_job built from a synth.Job (host vectors required)."""

    def __init__(self, job):
        self.keep = _Keep()
        k = self.keep
        self.job = job
        nrep = job.sym["nrep"]
        self.bs = [make_bset(b, k) for b in job.bsets]
        self.threej = threej_table(job.jmax)
        lv = job.levels
        n = job.nlevels
        q = np.ascontiguousarray(lv["quanta"], np.int32)
        k.items.append(q)
        levels = (OrLevel * max(n, 1))()
        base = q.ctypes.data
        stride = q.strides[0]
        for i in range(n):
            levels[i].energy = float(lv["energy"][i])
            levels[i].jind = int(lv["jind"][i])
            levels[i].jval = int(lv["jval"][i])
            levels[i].igamma = int(lv["igamma"][i])
            levels[i].ndeg = int(lv["ndeg"][i])
            levels[i].ilevel = int(lv["ilevel"][i])
            levels[i].irec0 = int(lv["irec0"][i])
            levels[i].quanta = C.cast(base + i * stride, P_i)
        k.items.append(levels)
        vecs = (P_d * (2 * nrep))()
        for jind in range(2):
            for g in range(nrep):
                v = job.vectors.get((jind, g))
                if v is None:
                    v = np.zeros(1)
                vecs[jind * nrep + g] = k.f64(np.asarray(v))
        k.items.append(vecs)
        bsp = (C.POINTER(OrBset) * 2)(C.pointer(self.bs[0]), C.pointer(self.bs[1]))
        self.s = OrJob(
            nrepres=nrep, degen=k.i32(job.sym["degen"]), sym_type=job.sym["sym_type"],
            nmodes=job.nmodes, zpe=job.zpe, temperature=job.temperature, q_stat=job.q_stat,
            freq_window=(cd * 2)(*job.freq_window), erange_lower=(cd * 2)(*job.erange_lower),
            erange_upper=(cd * 2)(*job.erange_upper), jvals=(ci * 2)(*job.jvals),
            gns=k.f64(job.gns), isym_pairs=k.i32(job.isym_pairs),
            igamma_pair=k.i32(job.igamma_pair), reduced=job.reduced,
            quanta_lower=k.i32(job.quanta_lower), quanta_upper=k.i32(job.quanta_upper),
            thresh_linestrength=job.thresh_linestrength, thresh_intensity=job.thresh_intensity,
            bset=bsp, nvib=job.nvib, dipole_flipped=k.f64(job.dipole_flipped), jmax=job.jmax,
            threej=_p(self.threej, P_d), nlevels=n, levels=levels,
            vectors=C.cast(vecs, C.POINTER(P_d)))

    def run(self, lower_ids=None, noskip=False, nthreads=0):
        L = lib()
        out = OrLines()
        if lower_ids is None:
            rc = L.or_run(C.byref(self.s), 0, None, int(noskip), int(nthreads), C.byref(out))
        else:
            ids = np.ascontiguousarray(lower_ids, np.int32)
            rc = L.or_run(C.byref(self.s), len(ids), _p(ids, P_i), int(noskip), int(nthreads),
                          C.byref(out))
        if rc != 0:
            raise RuntimeError("or_run failed")
        n = out.n
        res = {}
        if n > 0:
            res["ilevelI"] = np.ctypeslib.as_array(out.ilevelI, (n,)).copy()
            res["ilevelF"] = np.ctypeslib.as_array(out.ilevelF, (n,)).copy()
            for f in ("nu", "S", "A_einst", "I"):
                res[f] = np.ctypeslib.as_array(getattr(out, f), (n,)).copy()
            cols = amp_columns(max(self.job.sym["degen"]))
            for f in ("amp", "amp_abs"):
                a = np.ctypeslib.as_array(getattr(out, f), (n * 4,)).reshape(n, 4)
                res[f] = a[:, cols].copy()
        else:
            res = OracleJob.empty(len(amp_columns(max(self.job.sym["degen"]))))
        L.or_free_lines(C.byref(out))
        return res

    @staticmethod
    def empty(width=4):
        res = {"ilevelI": np.zeros(0, np.int32), "ilevelF": np.zeros(0, np.int32)}
        for f in ("nu", "S", "A_einst", "I"):
            res[f] = np.zeros(0)
        res["amp"] = np.zeros((0, width))
        res["amp_abs"] = np.zeros((0, width))
        return res

    def correlate(self, jind, ideg, g, vec):
        b = self.bs[jind]
        out = np.zeros(b.Maxcontracts)
        v = np.ascontiguousarray(vec, np.float64)
        lib().or_correlate(C.byref(b), self.s.degen, self.job.sym["nrep"], ideg, g,
                           _p(v, P_d), _p(out, P_d))
        return out

    def half_ls(self, jindI, jindF, vec, noskip=False):
        bI, bF = self.bs[jindI], self.bs[jindF]
        h = np.zeros(bF.Maxcontracts)
        v = np.ascontiguousarray(vec, np.float64)
        lib().or_half_ls(C.byref(bI), C.byref(bF), self.s.dipole_flipped, self.job.nvib,
                         _p(self.threej, P_d), self.job.jmax, _p(v, P_d), _p(h, P_d),
                         int(noskip))
        return h


def pair_job(job, i):
    """The two-J job of the run (jvals[i], jvals[i+1]) cut from a job with a J
    list: the same bases, dipole, eigenvector blocks and parameters, and the
    level table restricted to those two J (jind renumbered 0, 1). Returns the
    job and the indices of its levels in the full table."""
    keep = np.nonzero((job.levels["jind"] == i) | (job.levels["jind"] == i + 1))[0]
    lv = {k: np.ascontiguousarray(v[keep]) for k, v in job.levels.items()}
    lv["jind"] = np.ascontiguousarray(lv["jind"] - i, np.int32)
    vecs = {(jind - i, g): v for (jind, g), v in job.vectors.items() if jind in (i, i + 1)}
    sub = dataclasses.replace(job, jvals=(job.jvals[i], job.jvals[i + 1]),
                              bsets=[job.bsets[i], job.bsets[i + 1]], levels=lv,
                              vectors=vecs, jmax=max(job.jvals[i], job.jvals[i + 1]),
                              meta=dict(job.meta))
    return sub, keep


def run_chain(job, lower_ids=None, nthreads=0):
    """Lines of a job with a J list the way the reference produces them: one
    run per consecutive pair (jvals[i], jvals[i+1]) — a run holds two J
    (fields.h:61, cuda_host.cu:1005) — with level ids mapped back to the full
    table. The Q-branch rule of each run (trove_functions.cpp:1457) keeps
    every (J, J') pair in exactly one run. lower_ids (full-table ids) limits
    every run to those of its lower levels; runs with none are skipped."""
    parts = []
    for i in range(len(job.jvals) - 1):
        sub, keep = pair_job(job, i)
        if lower_ids is None:
            r = OracleJob(sub).run(nthreads=nthreads)
        else:
            ids = np.asarray(lower_ids)
            pos = np.searchsorted(keep, ids)
            ok = (pos < len(keep)) & (keep[np.minimum(pos, len(keep) - 1)] == ids)
            if not ok.any():
                continue
            r = OracleJob(sub).run(pos[ok].astype(np.int32), nthreads=nthreads)
        r["ilevelI"] = keep[r["ilevelI"]].astype(np.int32)
        r["ilevelF"] = keep[r["ilevelF"]].astype(np.int32)
        parts.append(r)
    if not parts:
        return OracleJob.empty(len(amp_columns(max(job.sym["degen"]))))
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def three_j(*a):
    return lib().or_three_j(*[int(x) for x in a])


def fakt(a):
    return lib().or_fakt(float(a))


def constants():
    o = np.zeros(7)
    lib().or_constants(_p(o, P_d))
    return dict(zip(["pi", "A_coef_s_1", "planck", "avogno", "vellgt", "intens_cm_mol",
                     "boltz"], o))


def max_threads():
    return lib().or_max_threads()
