"""ctypes binding of libgi_hip.so (the C ABI in include/gpu_intensities.h).

The library is the product: there is no CPU fallback. Loading fails loudly
when libgi_hip.so has not been built, and every call raises GIError on a
non-zero status with the library's own message.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgi_hip.so")
_LIB = None

ci, cd, cll = C.c_int, C.c_double, C.c_longlong
P_i, P_d = C.POINTER(C.c_int), C.POINTER(C.c_double)

GI_MODE_FP64 = 0
GI_MODE_FP64_BATCHED = 1
GI_MODE_SPLIT = 2
MODES = {"fp64": GI_MODE_FP64, "fp64_batched": GI_MODE_FP64_BATCHED, "split": GI_MODE_SPLIT}
STATUS = {0: "GI_OK", 1: "GI_EINVAL", 2: "GI_EHIP", 3: "GI_ENOMEM", 4: "GI_ESTATE",
          5: "GI_ECAPACITY", 6: "GI_EBUDGET", 7: "GI_ETIMEOUT"}
GI_EBUDGET, GI_ETIMEOUT = 6, 7

# Every symbol declared in include/gpu_intensities.h.
EXPORTS = ["gi_create", "gi_destroy", "gi_last_error", "gi_set_params", "gi_upload_bset",
           "gi_upload_dipole", "gi_set_levels", "gi_upload_vectors", "gi_prepare", "gi_compute",
           "gi_count_candidates", "gi_half_ls", "gi_correlate", "gi_threej_table",
           "gi_kernel_stats", "gi_work_stats", "gi_partition_function", "gi_serve_start",
           "gi_serve_submit", "gi_serve_wait", "gi_serve_fetch", "gi_serve_stats",
           "gi_serve_stop", "gi_set_tiles", "gi_get_tiles", "gi_set_split_digits",
           "gi_prepare_split", "gi_set_jvals", "gi_set_tuning", "gi_get_tuning",
           "gi_autotune", "gi_set_split_digits_each", "gi_sweep_create", "gi_sweep_run", "gi_sweep_candidates",
           "gi_sweep_destroy", "gi_amp_width", "gi_set_split_format", "gi_get_split_format",
           "gi_set_mu_budget", "gi_mu_parts", "gi_half_ls_noalloc",
           "gi_set_path_limits", "gi_set_vector_budget", "gi_streamed_blocks",
           "gi_set_persist", "gi_get_persist"]
# (TROVE_EXPORTS below: the symbols of include/gi_trove.h)


class GIError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class GiBset(C.Structure):
    _fields_ = [("jval", ci), ("Maxsymcoeffs", ci), ("Maxcontracts", ci), ("mat_size", ci),
                ("icontr2icase", P_i), ("iroot_correlat_j0", P_i), ("k", P_i), ("ktau", P_i),
                ("N", P_i), ("Ntotal", P_i), ("ijterms", P_i), ("repres", C.POINTER(P_d)),
                ("nsize", P_i)]


class GiParams(C.Structure):
    _fields_ = [("nrepres", ci), ("degen", P_i), ("sym_type", ci), ("nmodes", ci),
                ("zpe", cd), ("temperature", cd), ("q_stat", cd), ("freq_window", cd * 2),
                ("erange_lower", cd * 2), ("erange_upper", cd * 2), ("jvals", ci * 2),
                ("gns", P_d), ("isym_pairs", P_i), ("reduced", ci), ("quanta_lower", P_i),
                ("quanta_upper", P_i), ("thresh_linestrength", cd), ("thresh_intensity", cd)]


class GiRequest(C.Structure):
    _fields_ = [("jindI", ci), ("igammaI", ci), ("jindF", ci), ("mode", ci),
                ("temperature", cd), ("q_stat", cd), ("elow_lo", cd), ("elow_hi", cd),
                ("nu_lo", cd), ("nu_hi", cd), ("thresh_linestrength", cd),
                ("thresh_intensity", cd)]


class GiTuning(C.Structure):
    _fields_ = [("halfls_cfg", ci), ("corr_cfg", ci), ("pipe_rounds", ci), ("split_cfg", ci),
                ("pipe_share", ci)]


class GiLines(C.Structure):
    _fields_ = [("capacity", cll), ("n", cll), ("ilevelI", P_i), ("ilevelF", P_i),
                ("nu", P_d), ("S", P_d), ("A_einst", P_d), ("I", P_d), ("amp", P_d)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gi_create.argtypes = [ci, C.POINTER(vp)]
        L.gi_destroy.argtypes = [vp]
        L.gi_destroy.restype = None
        L.gi_last_error.argtypes = [vp]
        L.gi_last_error.restype = C.c_char_p
        L.gi_set_params.argtypes = [vp, C.POINTER(GiParams)]
        L.gi_set_jvals.argtypes = [vp, ci, P_i]
        L.gi_upload_bset.argtypes = [vp, ci, C.POINTER(GiBset)]
        L.gi_upload_dipole.argtypes = [vp, ci, vp, ci]
        L.gi_set_mu_budget.argtypes = [vp, cll]
        L.gi_mu_parts.argtypes = [vp]
        L.gi_set_levels.argtypes = [vp, ci, P_d, P_i, P_i, P_i, P_i, P_i]
        L.gi_upload_vectors.argtypes = [vp, ci, ci, ci, ci, vp, ci]
        L.gi_prepare.argtypes = [vp]
        L.gi_compute.argtypes = [vp, ci, ci, P_i, C.POINTER(GiLines)]
        L.gi_count_candidates.argtypes = [vp, ci, P_i, C.POINTER(cll)]
        L.gi_amp_width.argtypes = [vp]
        L.gi_half_ls.argtypes = [vp, ci, ci, ci, ci, P_d, P_d]
        L.gi_half_ls_noalloc.argtypes = [vp, ci, ci, ci, ci, P_d, P_d]
        L.gi_correlate.argtypes = [vp, ci, ci, ci, P_d, P_d]
        L.gi_threej_table.argtypes = [ci, P_d]
        L.gi_kernel_stats.argtypes = [vp, ci, C.POINTER(C.c_char_p), P_d, P_i, P_i]
        L.gi_work_stats.argtypes = [vp, P_d, ci]
        L.gi_partition_function.argtypes = [vp, cd, P_d]
        L.gi_serve_start.argtypes = [vp, cll]
        L.gi_serve_submit.argtypes = [vp, C.POINTER(GiRequest), C.POINTER(cll)]
        L.gi_serve_wait.argtypes = [vp, cll, cd, C.POINTER(cll), P_d]
        L.gi_serve_fetch.argtypes = [vp, cll, C.POINTER(GiLines)]
        L.gi_serve_stats.argtypes = [vp, C.POINTER(cll), C.POINTER(cll), C.POINTER(cll)]
        L.gi_serve_stop.argtypes = [vp]
        L.gi_set_tiles.argtypes = [vp, ci, ci]
        L.gi_set_split_digits.argtypes = [vp, ci]
        L.gi_set_split_digits_each.argtypes = [vp, ci, ci]
        L.gi_set_split_format.argtypes = [vp, ci, ci, ci]
        L.gi_get_split_format.argtypes = [vp, P_i, P_i, P_i]
        L.gi_prepare_split.argtypes = [vp]
        L.gi_get_tiles.argtypes = [vp, P_i, P_i]
        L.gi_sweep_create.argtypes = [vp, ci, ci, P_i, C.POINTER(vp)]
        L.gi_sweep_run.argtypes = [vp, cd, cd, cd, cd, C.POINTER(GiLines)]
        L.gi_sweep_candidates.argtypes = [vp]
        L.gi_sweep_candidates.restype = cll
        L.gi_sweep_destroy.argtypes = [vp]
        L.gi_set_tuning.argtypes = [vp, C.POINTER(GiTuning)]
        L.gi_get_tuning.argtypes = [vp, C.POINTER(GiTuning)]
        L.gi_set_path_limits.argtypes = [vp, ci, ci, ci, ci, ci]
        L.gi_set_persist.argtypes = [vp, ci]
        L.gi_get_persist.argtypes = [vp, P_i]
        L.gi_set_vector_budget.argtypes = [vp, cll]
        L.gi_streamed_blocks.argtypes = [vp]
        L.gi_autotune.argtypes = [vp, ci, ci, P_i, C.POINTER(GiLines), ci, C.POINTER(GiTuning),
                                  C.c_char_p, cll, C.POINTER(cll)]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def threej_table(jmax):
    t = np.zeros((jmax + 1) ** 2 * 9)
    rc = lib().gi_threej_table(jmax, _p(t, P_d))
    if rc:
        raise GIError(rc, "gi_threej_table")
    return t


class Engine:
    """One gi_ctx on a device, loaded with a synth.Job."""

    def __init__(self, device=0):
        self._keep = []
        h = C.c_void_p()
        rc = lib().gi_create(device, C.byref(h))
        if rc:
            raise GIError(rc, "gi_create failed (no HIP device?)")
        self.h = h
        self.job = None

    def close(self):
        if self.h:
            lib().gi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            msg = lib().gi_last_error(self.h)
            raise GIError(rc, f"{what}: {msg.decode() if msg else ''}")

    def _i32(self, a):
        a = np.ascontiguousarray(a, np.int32)
        self._keep.append(a)
        return _p(a, P_i)

    def _f64(self, a):
        a = np.ascontiguousarray(a, np.float64)
        self._keep.append(a)
        return _p(a, P_d)

    def load(self, job, device_vectors=None, device_dipole=None, mu_budget=0, vector_budget=0):
        """Set parameters, bases, dipole, levels and eigenvector blocks.
        device_vectors: optional {(jind, g): torch tensor on the device};
        mu_budget: device bytes for mu (gi_set_mu_budget; 0: no limit);
        vector_budget: device bytes for the blocks uploaded from host memory
        (gi_set_vector_budget; 0: no limit)."""
        L = lib()
        self.job = job
        nrep = job.sym["nrep"]
        p = GiParams(nrepres=nrep, degen=self._i32(job.sym["degen"]),
                     sym_type=job.sym["sym_type"], nmodes=job.nmodes, zpe=job.zpe,
                     temperature=job.temperature, q_stat=job.q_stat,
                     freq_window=(cd * 2)(*job.freq_window),
                     erange_lower=(cd * 2)(*job.erange_lower),
                     erange_upper=(cd * 2)(*job.erange_upper), jvals=(ci * 2)(*job.jvals[:2]),
                     gns=self._f64(job.gns), isym_pairs=self._i32(job.isym_pairs),
                     reduced=job.reduced, quanta_lower=self._i32(job.quanta_lower),
                     quanta_upper=self._i32(job.quanta_upper),
                     thresh_linestrength=job.thresh_linestrength,
                     thresh_intensity=job.thresh_intensity)
        self._chk(L.gi_set_params(self.h, C.byref(p)), "gi_set_params")
        if len(job.jvals) != 2:
            jv = np.ascontiguousarray(job.jvals, np.int32)
            self._chk(L.gi_set_jvals(self.h, len(jv), _p(jv, P_i)), "gi_set_jvals")
        for jind, b in enumerate(job.bsets):
            reps = (P_d * nrep)(*[self._f64(r) for r in b.repres])
            self._keep.append(reps)
            s = GiBset(jval=b.jval, Maxsymcoeffs=b.Maxsymcoeffs, Maxcontracts=b.Maxcontracts,
                       mat_size=b.mat_size, icontr2icase=self._i32(b.icontr2icase),
                       iroot_correlat_j0=self._i32(b.iroot_correlat_j0), k=self._i32(b.k),
                       ktau=self._i32(b.ktau), N=self._i32(b.N), Ntotal=self._i32(b.Ntotal),
                       ijterms=self._i32(b.ijterms), repres=C.cast(reps, C.POINTER(P_d)),
                       nsize=self._i32(b.nsize))
            self._chk(L.gi_upload_bset(self.h, jind, C.byref(s)), "gi_upload_bset")
        self._chk(L.gi_set_mu_budget(self.h, int(mu_budget)), "gi_set_mu_budget")
        self._chk(L.gi_set_vector_budget(self.h, int(vector_budget)), "gi_set_vector_budget")
        if device_dipole is not None:
            self._chk(L.gi_upload_dipole(self.h, job.nvib, C.c_void_p(device_dipole.data_ptr()), 1),
                      "gi_upload_dipole")
        else:
            mu = np.ascontiguousarray(job.dipole_flipped, np.float64)
            self._chk(L.gi_upload_dipole(self.h, job.nvib, C.c_void_p(mu.ctypes.data), 0),
                      "gi_upload_dipole")
        lv = job.levels
        self._chk(L.gi_set_levels(self.h, job.nlevels, self._f64(lv["energy"]),
                                  self._i32(lv["jind"]), self._i32(lv["igamma"]),
                                  self._i32(lv["ndeg"]), self._i32(lv["irec0"]),
                                  self._i32(lv["quanta"].reshape(-1))), "gi_set_levels")
        for jind in range(len(job.jvals)):
            for g in range(nrep):
                nsize = int(job.bsets[jind].nsize[g])
                if device_vectors is not None and (jind, g) in device_vectors:
                    t = device_vectors[(jind, g)]
                    nrec = int(t.shape[0]) if nsize else 0
                    self._keep.append(t)
                    self._chk(L.gi_upload_vectors(self.h, jind, g, nsize, nrec,
                                                  C.c_void_p(t.data_ptr()), 1),
                              "gi_upload_vectors")
                else:
                    v = job.vectors.get((jind, g))
                    if v is None:
                        v = np.zeros((0, nsize))
                    v = np.ascontiguousarray(v, np.float64)
                    nrec = int(v.shape[0])
                    self._chk(L.gi_upload_vectors(self.h, jind, g, nsize, nrec,
                                                  C.c_void_p(v.ctypes.data if v.size else 0), 0),
                              "gi_upload_vectors")
        self._chk(L.gi_prepare(self.h), "gi_prepare")
        return self

    def count_candidates(self, lower_ids=None):
        n = cll()
        if lower_ids is None:
            self._chk(lib().gi_count_candidates(self.h, 0, None, C.byref(n)), "count")
        else:
            ids = np.ascontiguousarray(lower_ids, np.int32)
            self._chk(lib().gi_count_candidates(self.h, len(ids), _p(ids, P_i), C.byref(n)),
                      "count")
        return n.value

    def streamed_blocks(self):
        """Eigenvector blocks kept in host memory (gi_streamed_blocks)."""
        return lib().gi_streamed_blocks(self.h)

    def mu_parts(self):
        """Parts mu is streamed in (gi_mu_parts; 1: resident)."""
        return lib().gi_mu_parts(self.h)

    def amp_width(self):
        """Amplitudes per line (gi_amp_width): m*m, m the largest degeneracy."""
        return lib().gi_amp_width(self.h)

    def alloc_lines(self, cap, pinned=False):
        cap = max(int(cap), 1)
        w = max(self.amp_width(), 1)
        if pinned:
            import torch
            mk = lambda n, dt: torch.empty(n, dtype=dt, pin_memory=True).numpy()
            import torch as T
            out = dict(ilevelI=mk(cap, T.int32), ilevelF=mk(cap, T.int32), nu=mk(cap, T.float64),
                       S=mk(cap, T.float64), A_einst=mk(cap, T.float64), I=mk(cap, T.float64),
                       amp=mk(cap * w, T.float64))
        else:
            out = dict(ilevelI=np.empty(cap, np.int32), ilevelF=np.empty(cap, np.int32),
                       nu=np.empty(cap), S=np.empty(cap), A_einst=np.empty(cap), I=np.empty(cap),
                       amp=np.empty(cap * w))
        return out

    def _lines(self, bufs):
        cap = min(len(bufs["S"]), len(bufs["amp"]) // max(self.amp_width(), 1))
        return GiLines(capacity=cap, n=0, ilevelI=_p(bufs["ilevelI"], P_i),
                       ilevelF=_p(bufs["ilevelF"], P_i), nu=_p(bufs["nu"], P_d),
                       S=_p(bufs["S"], P_d), A_einst=_p(bufs["A_einst"], P_d),
                       I=_p(bufs["I"], P_d), amp=_p(bufs["amp"], P_d))

    def _result(self, bufs, n):
        """Rows [0, n) of bufs; amp as an (n, m*m) array, column idegI + m*idegF."""
        w = self.amp_width()
        res = {k: bufs[k][:n] for k in ("ilevelI", "ilevelF", "nu", "S", "A_einst", "I")}
        res["amp"] = bufs["amp"][:w * n].reshape(n, w)
        return res

    def compute(self, lower_ids=None, mode=GI_MODE_FP64, bufs=None):
        """gi_compute into bufs (allocated for every candidate when None; a
        buffer that turns out too small is replaced and the call repeated)."""
        if bufs is None:
            bufs = self.alloc_lines(self.count_candidates(lower_ids))
        ids = None if lower_ids is None else np.ascontiguousarray(lower_ids, np.int32)
        for _ in range(2):
            s = self._lines(bufs)
            if ids is None:
                rc = lib().gi_compute(self.h, mode, 0, None, C.byref(s))
            else:
                rc = lib().gi_compute(self.h, mode, len(ids), _p(ids, P_i), C.byref(s))
            if rc != 5:  # GI_ECAPACITY: s.n holds the rows needed
                break
            bufs = self.alloc_lines(max(s.n, self.count_candidates(lower_ids)))
        self._chk(rc, "gi_compute")
        return self._result(bufs, s.n)

    def set_tiles(self, halfls=-1, corr=-1):
        """Tile shapes of the batched kernels: 0..3 = 64x64, 64x128, 128x64, 128x128."""
        self._chk(lib().gi_set_tiles(self.h, halfls, corr), "gi_set_tiles")

    def set_split_digits(self, digits):
        self._chk(lib().gi_set_split_digits(self.h, digits), "gi_set_split_digits")

    def set_split_format(self, radix=0, halfls_digits=0, corr_digits=0):
        """gi_set_split_format: radix 0 (automatic), 128 or 254; digits 0 (the
        radix's default) or 3..6 per product."""
        self._chk(lib().gi_set_split_format(self.h, radix, halfls_digits, corr_digits),
                  "gi_set_split_format")

    def split_format(self):
        """(radix, half-ls digits, correlation digits) of the resident planes."""
        r, h, c = ci(), ci(), ci()
        self._chk(lib().gi_get_split_format(self.h, C.byref(r), C.byref(h), C.byref(c)),
                  "gi_get_split_format")
        return r.value, h.value, c.value

    def set_split_digits_each(self, halfls_digits, corr_digits):
        self._chk(lib().gi_set_split_digits_each(self.h, halfls_digits, corr_digits),
                  "gi_set_split_digits_each")

    def prepare_split(self):
        self._chk(lib().gi_prepare_split(self.h), "gi_prepare_split")

    def get_tiles(self):
        a, b = ci(), ci()
        self._chk(lib().gi_get_tiles(self.h, C.byref(a), C.byref(b)), "gi_get_tiles")
        return a.value, b.value

    def set_tuning(self, halfls_cfg, corr_cfg, pipe_rounds, split_cfg=0, pipe_share=50):
        t = GiTuning(halfls_cfg=halfls_cfg, corr_cfg=corr_cfg, pipe_rounds=pipe_rounds,
                     split_cfg=split_cfg, pipe_share=pipe_share)
        self._chk(lib().gi_set_tuning(self.h, C.byref(t)), "gi_set_tuning")

    def set_path_limits(self, kpart_len=0, rec_lds=0, split2_steps=0, split2_segs=0,
                        fold_split_rows=0):
        """gi_set_path_limits: lower the routing thresholds of the batched
        kernels (0 restores a default)."""
        self._chk(lib().gi_set_path_limits(self.h, kpart_len, rec_lds, split2_steps,
                                           split2_segs, fold_split_rows), "gi_set_path_limits")

    def set_persist(self, on):
        """gi_set_persist: FP64 batched rounds as one persistent launch (1) or
        one launch per stage (0)."""
        self._chk(lib().gi_set_persist(self.h, int(on)), "gi_set_persist")

    def get_persist(self):
        v = np.zeros(1, np.int32)
        self._chk(lib().gi_get_persist(self.h, _p(v, P_i)), "gi_get_persist")
        return int(v[0])

    def get_tuning(self):
        t = GiTuning()
        self._chk(lib().gi_get_tuning(self.h, C.byref(t)), "gi_get_tuning")
        return dict(halfls_cfg=t.halfls_cfg, corr_cfg=t.corr_cfg, pipe_rounds=t.pipe_rounds,
                    split_cfg=t.split_cfg, pipe_share=t.pipe_share)

    def autotune(self, mode, lower_ids=None, bufs=None, reps=2):
        """gi_autotune on these lower levels; returns (chosen setting, table text)."""
        cap = self.count_candidates(lower_ids)
        if bufs is None or len(bufs["S"]) < max(cap, 1):
            bufs = self.alloc_lines(cap)
        s = self._lines(bufs)
        best = GiTuning()
        buf = C.create_string_buffer(1 << 16)
        n = cll()
        if lower_ids is None:
            ids, nids = None, 0
        else:
            ids = np.ascontiguousarray(lower_ids, np.int32)
            nids = len(ids)
        rc = lib().gi_autotune(self.h, mode, nids, None if ids is None else _p(ids, P_i),
                               C.byref(s), reps, C.byref(best), buf, len(buf), C.byref(n))
        self._chk(rc, "gi_autotune")
        return (dict(halfls_cfg=best.halfls_cfg, corr_cfg=best.corr_cfg,
                     pipe_rounds=best.pipe_rounds, split_cfg=best.split_cfg,
                     pipe_share=best.pipe_share),
                buf.value.decode())

    # -- captured sweeps ---------------------------------------------------
    def sweep_create(self, mode, lower_ids=None):
        """gi_sweep_create; returns an opaque handle (destroy with sweep_destroy)."""
        h = C.c_void_p()
        if lower_ids is None:
            rc = lib().gi_sweep_create(self.h, mode, 0, None, C.byref(h))
        else:
            ids = np.ascontiguousarray(lower_ids, np.int32)
            rc = lib().gi_sweep_create(self.h, mode, len(ids), _p(ids, P_i), C.byref(h))
        self._chk(rc, "gi_sweep_create")
        return h

    def sweep_candidates(self, sw):
        return int(lib().gi_sweep_candidates(sw))

    def sweep_run(self, sw, temperature=None, q_stat=None, thresh_linestrength=None,
                  thresh_intensity=None, bufs=None):
        """gi_sweep_run; unset fields take the loaded job's values (q_stat <= 0: Q(T))."""
        j = self.job
        if bufs is None:
            bufs = self.alloc_lines(self.sweep_candidates(sw))
        s = self._lines(bufs)
        rc = lib().gi_sweep_run(
            sw, j.temperature if temperature is None else temperature,
            j.q_stat if q_stat is None else q_stat,
            j.thresh_linestrength if thresh_linestrength is None else thresh_linestrength,
            j.thresh_intensity if thresh_intensity is None else thresh_intensity, C.byref(s))
        self._chk(rc, "gi_sweep_run")
        return self._result(bufs, s.n)

    def sweep_destroy(self, sw):
        self._chk(lib().gi_sweep_destroy(sw), "gi_sweep_destroy")

    def partition_function(self, temperature):
        q = cd()
        self._chk(lib().gi_partition_function(self.h, temperature, C.byref(q)),
                  "gi_partition_function")
        return q.value

    # -- request server ----------------------------------------------------
    def serve_start(self, budget_bytes):
        self._chk(lib().gi_serve_start(self.h, int(budget_bytes)), "gi_serve_start")

    def serve_submit(self, jindI, igammaI, jindF, mode=GI_MODE_FP64, temperature=None,
                     q_stat=None, elow=(-1e300, 1e300), nu=None, thresh_linestrength=None,
                     thresh_intensity=None):
        """Queue one request; unset fields take the loaded job's values
        (q_stat <= 0 asks the library for Q(T)). Returns the ticket."""
        j = self.job
        r = GiRequest(jindI=jindI, igammaI=igammaI, jindF=jindF, mode=mode,
                      temperature=j.temperature if temperature is None else temperature,
                      q_stat=j.q_stat if q_stat is None else q_stat,
                      elow_lo=elow[0], elow_hi=elow[1],
                      nu_lo=(j.freq_window if nu is None else nu)[0],
                      nu_hi=(j.freq_window if nu is None else nu)[1],
                      thresh_linestrength=(j.thresh_linestrength if thresh_linestrength is None
                                           else thresh_linestrength),
                      thresh_intensity=(j.thresh_intensity if thresh_intensity is None
                                        else thresh_intensity))
        t = cll()
        self._chk(lib().gi_serve_submit(self.h, C.byref(r), C.byref(t)), "gi_serve_submit")
        return t.value

    def serve_wait(self, ticket, timeout_s=0.0):
        """(nlines, latency_ms) of a finished request; GIError with status
        GI_ETIMEOUT if it is still running after timeout_s."""
        n, lat = cll(), cd()
        self._chk(lib().gi_serve_wait(self.h, ticket, timeout_s, C.byref(n), C.byref(lat)),
                  "gi_serve_wait")
        return n.value, lat.value

    def serve_fetch(self, ticket, bufs=None):
        n, _ = self.serve_wait(ticket)
        if bufs is None:
            bufs = self.alloc_lines(n)
        s = self._lines(bufs)
        self._chk(lib().gi_serve_fetch(self.h, ticket, C.byref(s)), "gi_serve_fetch")
        return self._result(bufs, s.n)

    def serve_stats(self):
        b, s, m = cll(), cll(), cll()
        self._chk(lib().gi_serve_stats(self.h, C.byref(b), C.byref(s), C.byref(m)),
                  "gi_serve_stats")
        return dict(batches=b.value, served=s.value, max_batch=m.value)

    def serve_stop(self):
        self._chk(lib().gi_serve_stop(self.h), "gi_serve_stop")

    def half_ls(self, jindI, jindF, g, ideg, vec):
        v = np.ascontiguousarray(vec, np.float64)
        h = np.zeros(self.job.bsets[jindF].Maxcontracts)
        self._chk(lib().gi_half_ls(self.h, jindI, jindF, g, ideg, _p(v, P_d), _p(h, P_d)),
                  "gi_half_ls")
        return h

    def half_ls_noalloc(self, jindI, jindF, g, ideg, vec):
        """gi_half_ls_noalloc: gi_half_ls with its device buffers in the
        library's static scratch array (no allocation)."""
        v = np.ascontiguousarray(vec, np.float64)
        h = np.zeros(self.job.bsets[jindF].Maxcontracts)
        self._chk(lib().gi_half_ls_noalloc(self.h, jindI, jindF, g, ideg, _p(v, P_d),
                                           _p(h, P_d)), "gi_half_ls_noalloc")
        return h

    def correlate(self, jind, g, ideg, vec):
        v = np.ascontiguousarray(vec, np.float64)
        o = np.zeros(self.job.bsets[jind].Maxcontracts)
        self._chk(lib().gi_correlate(self.h, jind, g, ideg, _p(v, P_d), _p(o, P_d)),
                  "gi_correlate")
        return o

    def kernel_stats(self):
        n = ci()
        names = (C.c_char_p * 32)()
        ms = np.zeros(32)
        la = np.zeros(32, np.int32)
        self._chk(lib().gi_kernel_stats(self.h, 32, names, _p(ms, P_d), _p(la, P_i), C.byref(n)),
                  "gi_kernel_stats")
        return {names[i].decode(): dict(total_ms=float(ms[i]), launches=int(la[i]))
                for i in range(min(n.value, 32))}

    def work_stats(self):
        w = np.zeros(17)
        self._chk(lib().gi_work_stats(self.h, _p(w, P_d), 17), "gi_work_stats")
        keys = ["units", "candidates", "lines", "half_ls_dense_flops", "half_ls_flops",
                "corr_bytes", "mu_bytes", "mu_dense_bytes", "corr_flops", "host_plan_ms",
                "host_enqueue_ms", "host_wait_copy_ms", "out_bytes", "call_ms",
                "half_ls_tiles", "corr_tiles", "corr_useful_flops"]
        return dict(zip(keys, w.tolist()))


# -- TROVE files, deck and line list (include/gi_trove.h) ---------------------
TROVE_EXPORTS = ["gi_trove_open", "gi_trove_error", "gi_trove_close", "gi_trove_load",
                 "gi_trove_apply_thresh_line",
                 "gi_trove_params", "gi_trove_bset", "gi_trove_dipole", "gi_trove_levels",
                 "gi_trove_vectors", "gi_trove_header", "gi_trove_format"]
_TROVE_READY = False


def _trove_lib():
    global _TROVE_READY
    L = lib()
    if not _TROVE_READY:
        vp = C.c_void_p
        PP_i, PP_d = C.POINTER(P_i), C.POINTER(P_d)
        L.gi_trove_open.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
        L.gi_trove_error.argtypes = [vp]
        L.gi_trove_error.restype = C.c_char_p
        L.gi_trove_close.argtypes = [vp]
        L.gi_trove_close.restype = None
        L.gi_trove_load.argtypes = [vp, vp]
        L.gi_trove_apply_thresh_line.argtypes = [vp, ci]
        L.gi_trove_params.argtypes = [vp, C.POINTER(GiParams)]
        L.gi_trove_bset.argtypes = [vp, ci, C.POINTER(GiBset)]
        L.gi_trove_dipole.argtypes = [vp, P_i, PP_d]
        L.gi_trove_levels.argtypes = [vp, P_i, PP_d, PP_i, PP_i, PP_i, PP_i, PP_i, PP_i, PP_i,
                                      PP_i]
        L.gi_trove_vectors.argtypes = [vp, ci, ci, P_i, P_i, PP_d]
        L.gi_trove_header.argtypes = [vp, C.c_char_p, cll, C.POINTER(cll)]
        L.gi_trove_format.argtypes = [vp, C.POINTER(GiLines), cll, cll, C.c_char_p, cll,
                                      C.POINTER(cll)]
        _TROVE_READY = True
    return L


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, (n,)).astype(dtype, copy=True)


class Trove:
    """TROVE input files of one job, read by the library (gi_trove_open)."""

    def __init__(self, directory, deck):
        L = _trove_lib()
        h = C.c_void_p()
        rc = L.gi_trove_open(directory.encode(), deck.encode(), C.byref(h))
        self.h = h
        if rc:
            msg = L.gi_trove_error(h)
            L.gi_trove_close(h)
            self.h = None
            raise GIError(rc, f"gi_trove_open: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            _trove_lib().gi_trove_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = GiParams()
        _trove_lib().gi_trove_params(self.h, C.byref(p))
        nrep, nm = p.nrepres, p.nmodes
        return dict(nrepres=nrep, degen=_arr(p.degen, nrep, np.int32), sym_type=p.sym_type,
                    nmodes=nm, zpe=p.zpe, temperature=p.temperature, q_stat=p.q_stat,
                    freq_window=tuple(p.freq_window), erange_lower=tuple(p.erange_lower),
                    erange_upper=tuple(p.erange_upper), jvals=tuple(p.jvals),
                    gns=_arr(p.gns, nrep, np.float64), isym_pairs=_arr(p.isym_pairs, nrep, np.int32),
                    reduced=p.reduced, quanta_lower=_arr(p.quanta_lower, 2 * nm, np.int64),
                    quanta_upper=_arr(p.quanta_upper, 2 * nm, np.int64),
                    thresh_linestrength=p.thresh_linestrength)

    def bset(self, jind, degen):
        b = GiBset()
        _trove_lib().gi_trove_bset(self.h, jind, C.byref(b))
        d, nc, nrep = b.Maxcontracts, b.Maxsymcoeffs, len(degen)
        Ntot = _arr(b.Ntotal, nrep, np.int32)
        return dict(jval=b.jval, Maxsymcoeffs=nc, Maxcontracts=d, mat_size=b.mat_size,
                    icontr2icase=_arr(b.icontr2icase, 2 * d, np.int32),
                    iroot_correlat_j0=_arr(b.iroot_correlat_j0, d, np.int32),
                    k=_arr(b.k, d, np.int32), ktau=_arr(b.ktau, d, np.int32),
                    N=_arr(b.N, nrep * nc, np.int32), Ntotal=Ntot,
                    ijterms=_arr(b.ijterms, nrep * nc, np.int32),
                    repres=[_arr(b.repres[g], int(Ntot[g]) * int(degen[g]) * b.mat_size,
                                 np.float64) for g in range(nrep)],
                    nsize=_arr(b.nsize, nrep, np.int32))

    def dipole(self):
        n = ci()
        p = P_d()
        _trove_lib().gi_trove_dipole(self.h, C.byref(n), C.byref(p))
        return _arr(p, 3 * n.value * n.value, np.float64), n.value

    def levels(self, nmodes):
        n = ci()
        ptrs = [P_d()] + [P_i() for _ in range(9)]
        _trove_lib().gi_trove_levels(self.h, C.byref(n), *[C.byref(x) for x in ptrs])
        m = n.value
        keys = ["energy", "jind", "igamma", "ndeg", "irec0", "ilevel", "quanta", "krot", "taurot"]
        out = {"energy": _arr(ptrs[0], m, np.float64)}
        for k, p in zip(keys[1:], ptrs[1:]):
            out[k] = _arr(p, m * (nmodes + 1) if k == "quanta" else m, np.int32)
        out["quanta"] = out["quanta"].reshape(m, nmodes + 1)
        return out

    def vectors(self, jind, g):
        ns, nr, p = ci(), ci(), P_d()
        _trove_lib().gi_trove_vectors(self.h, jind, g, C.byref(ns), C.byref(nr), C.byref(p))
        return _arr(p, ns.value * nr.value, np.float64).reshape(nr.value, ns.value)

    def load(self, engine, apply_thresh_line=False):
        """gi_trove_load; THRESH_LINE is applied only when asked (the
        reference parses it and never applies it)."""
        _trove_lib().gi_trove_apply_thresh_line(self.h, int(apply_thresh_line))
        rc = _trove_lib().gi_trove_load(self.h, engine.h)
        engine._chk(rc, "gi_trove_load")

    def header(self):
        n = cll()
        L = _trove_lib()
        L.gi_trove_header(self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        L.gi_trove_header(self.h, buf, n.value + 1, C.byref(n))
        return buf.value.decode()

    def format(self, res):
        """Line-list text of a compute() result (dict of arrays)."""
        nrow = len(res["S"])
        keep = {k: np.ascontiguousarray(res[k]) for k in ("ilevelI", "ilevelF", "nu", "S",
                                                          "A_einst", "I")}
        keep["amp"] = np.ascontiguousarray(res["amp"]).reshape(-1)
        s = GiLines(capacity=nrow, n=nrow, ilevelI=_p(keep["ilevelI"].astype(np.int32), P_i),
                    ilevelF=_p(keep["ilevelF"].astype(np.int32), P_i),
                    nu=_p(keep["nu"], P_d), S=_p(keep["S"], P_d), A_einst=_p(keep["A_einst"], P_d),
                    I=_p(keep["I"], P_d), amp=_p(keep["amp"], P_d))
        iI, iF = keep["ilevelI"].astype(np.int32), keep["ilevelF"].astype(np.int32)
        s.ilevelI, s.ilevelF = _p(iI, P_i), _p(iF, P_i)
        n = cll()
        L = _trove_lib()
        rc = L.gi_trove_format(self.h, C.byref(s), 0, nrow, None, 0, C.byref(n))
        if rc not in (0, 5):
            raise GIError(rc, "gi_trove_format")
        buf = C.create_string_buffer(n.value + 1)
        rc = L.gi_trove_format(self.h, C.byref(s), 0, nrow, buf, n.value + 1, C.byref(n))
        if rc:
            raise GIError(rc, "gi_trove_format")
        return buf.value.decode()
