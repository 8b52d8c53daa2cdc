"""Write a synth.Job as the TROVE files and stdin deck the reference reads.

Files (names trove_functions.cpp:24-28), written so that the reference's
readers rebuild the Job's arrays exactly:

* j0eigen_quanta<J>.chk (text, bset_contr_factory, trove_functions.cpp:401-679):
  a case is (vibrational unit, K); its contractive space is (K level, unit)
  and a lambda's coefficients are (tau index, component in the unit), so that
  correlate_index (:680-858) maps every function to its J=0 contracted
  index (the vibrational index v) and reads (K, tau) from the rotational table;
* j0_extfield.chk (Fortran sequential unformatted records, read_dipole_flipped
  :113-205, ReadFortranRecord Util.cpp:67-75);
* j0eigen_descr<J>_<G>.chk (text, read_eigenvalues :1074-1429);
* j0eigen_vectors<J>_<G>.chk (raw doubles, record irec at (irec-1)*nsize).

The deck follows read_fields (fields.cpp:145-282).
"""
from __future__ import annotations

import os
import struct

import numpy as np

from . import synth


def _units(job):
    units = job.meta.get("vib_units")
    if units is None:
        raise ValueError("job has no vibrational unit table (made by an older synth)")
    unit_of = np.empty(job.nvib, np.int64)
    comp_of = np.empty(job.nvib, np.int64)
    for u, (_, vs) in enumerate(units):
        for c, v in enumerate(vs):
            unit_of[v] = u
            comp_of[v] = c
    return units, unit_of, comp_of


def _j0_basis(job):
    """The J=0 basis: one case per vibrational unit, functions in v order."""
    units, _, _ = _units(job)
    if job.jvals[0] == 0:
        return job.bsets[0]
    if job.sym["sym_type"] == 1:
        gvib = np.array([g for g, _ in units])
        return synth.build_bset_c2v(0, gvib, synth._rng(0, 7))
    if job.sym["sym_type"] == 3:  # D3h: the ' units (irreps 0-2), then the " units
        b1 = synth.build_bset_c3v(0, [(g, vs) for g, vs in units if g < 3], synth._rng(0, 7))
        b2 = synth.build_bset_c3v(0, [(g - 3, vs) for g, vs in units if g >= 3],
                                  synth._rng(0, 8))
        return synth.direct_sum_bsets(b1, b2, 3)
    return synth.build_bset_c3v(0, units, synth._rng(0, 7))


def _fmt(x):
    return repr(float(x))


def quanta_text(job, b):
    """j0eigen_quanta<J>.chk for basis b."""
    _, unit_of, comp_of = _units(job)
    J = b.jval
    nc = b.Maxcontracts
    icase = b.icontr2icase[:nc]
    ilam = b.icontr2icase[nc:]
    ncases = b.Maxsymcoeffs
    cases = [None] * ncases
    lams = [[] for _ in range(ncases)]
    for r in range(nc):
        v = int(b.iroot_correlat_j0[r])
        K = int(b.k[r])
        tau = int(b.ktau[r]) - 2 * K
        tidx = 0 if K == 0 else tau
        cs = (K, int(unit_of[v]))
        c = int(icase[r])
        if cases[c] is None:
            cases[c] = cs
        elif cases[c] != cs:
            raise ValueError("a case mixes K values or vibrational units")
        if int(ilam[r]) != len(lams[c]):
            raise ValueError("functions of a case are not in lambda order")
        lams[c].append((r, tidx, int(comp_of[v])))
    nlmax = max(len(x) for x in lams)
    out = ["Start Primitive basis set", f"{ncases:8d} {nlmax:8d} {nc:8d} {1:8d}"]
    r_expected = 0
    for c in range(ncases):
        K, u = cases[c]
        out.append(f"{len(lams[c]):8d} {K + 1:6d} {u + 1:6d}")
        for r, tidx, comp in lams[c]:
            if r != r_expected:
                raise ValueError("functions are not numbered case by case")
            r_expected += 1
            out.append(f"{r + 1:8d} {tidx + 1:6d} {comp + 1:6d}")
    # rotational table: level K (1-based K+1), degenerate index tau (1-based)
    out.append(f"{J + 1:8d} {2:8d}")
    out.append(f"{1:8d} {1:8d} {J:8d} {0:8d} {J % 2:8d}")
    for K in range(1, J + 1):
        for tau in (0, 1):
            out.append(f"{K + 1:8d} {tau + 1:8d} {J:8d} {K:8d} {tau:8d}")
    out.append("End Primitive basis set")
    out.append("Start irreducible transformation")
    out.append(f"{b.mat_size:8d}")
    nrep = job.sym["nrep"]
    out.append(" ".join(f"{int(b.Ntotal[g]):8d}" for g in range(nrep)))
    degen = job.sym["degen"]
    for g in range(nrep):
        for c in range(ncases):
            out.append(f"{int(b.N[g * ncases + c]):8d}")
        nt, dg = int(b.Ntotal[g]), degen[g]
        rep = np.asarray(b.repres[g])
        for ico in range(nt):
            for d in range(dg):
                vals = [rep[ico + d * nt + i * nt * dg] for i in range(b.mat_size)]
                out.append(" ".join(_fmt(x) for x in vals))
    out.append("End irreducible transformation")
    return "\n".join(out) + "\n"


def _record(f, payload: bytes):
    f.write(struct.pack("<I", len(payload)))
    f.write(payload)
    f.write(struct.pack("<I", len(payload)))


def write_dipole(job, path):
    """j0_extfield.chk: the file blocks are the flipped ones transposed back."""
    N = job.nvib
    flip = np.asarray(job.dipole_flipped, np.float64).reshape(3, N, N)  # [a][vI][vF]
    with open(path, "wb") as f:
        _record(f, b"Start external field")
        _record(f, struct.pack("<i", N))
        for a in range(3):
            _record(f, struct.pack("<i", a + 1))
            # element (i, f) of the file block at i + f*N holds flipped (vF=f, vI=i)
            _record(f, np.ascontiguousarray(flip[a].T).tobytes())
        _record(f, b"End external field")


def descr_text(job, jind, g):
    lv = job.levels
    m = np.nonzero((lv["jind"] == jind) & (lv["igamma"] == g))[0]
    m = m[np.argsort(lv["ilevel"][m], kind="stable")]
    nsize = int(job.bsets[jind].nsize[g])
    degen = job.sym["degen"][g]
    out = ["Synthetic eigenvalues (written by trove_io.py)",
           "Start Quantum numbers and energies", "", "", f"{len(m):8d} {nsize:8d}"]
    for i in m:
        q = " ".join(str(int(x)) for x in lv["quanta"][i])
        cg = " ".join(str(int(x) + 1) for x in lv["cgamma"][i])
        nrm = " ".join(str(int(x)) for x in lv["normal"][i])
        for d in range(1, degen + 1):
            out.append(f"{int(lv['irec0'][i]) + 1:8d} {g + 1:4d} {int(lv['ilevel'][i]) + 1:8d} "
                       f"{d:3d} {_fmt(lv['energy'][i])} {q} 1 {cg} {nrm}")
    out.append("End Quantum numbers and energies")
    return "\n".join(out) + "\n"


def deck_text(job, comments=True):
    nm = job.nmodes
    big = 1 << 30
    lines = [f"NMODES {nm}", f"SYMGROUP {job.sym['name']}", "INTENSITY",
             f"  TEMPERATURE {_fmt(job.temperature)}" + ("  (Kelvin)" if comments else ""),
             f"  QSTAT {_fmt(job.q_stat)}",
             f"  ZPE {_fmt(job.zpe)}",
             f"  J {job.jvals[0]},{job.jvals[1]}",
             f"  FREQ-WINDOW {_fmt(job.freq_window[0])} {_fmt(job.freq_window[1])}",
             f"  ENERGY LOW {_fmt(job.erange_lower[0])} {_fmt(job.erange_lower[1])} "
             f"UPPER {_fmt(job.erange_upper[0])} {_fmt(job.erange_upper[1])}",
             "  GNS " + " ".join(_fmt(x) for x in job.gns),
             "  SELECTION " + " ".join(str(int(x)) for x in job.isym_pairs)]
    if job.thresh_linestrength > 0:
        lines.append(f"  THRESH_LINE {_fmt(job.thresh_linestrength)}")
    for i in range(nm):
        lo, hi = int(job.quanta_lower[i]), int(job.quanta_lower[nm + i])
        ulo, uhi = int(job.quanta_upper[i]), int(job.quanta_upper[nm + i])
        if (lo, hi, ulo, uhi) != (0, big, 0, big):
            lines.append(f"  V {i + 1} LOWER {lo} {hi} UPPER {ulo} {uhi}")
    if job.reduced:
        lines.append("  SYMMETRY REDUCED")
    lines.append("END")
    return "\n".join(lines) + "\n"


def write_trove(job, d):
    """Write every file of `job` into directory d; returns the deck text."""
    if job.meta.get("shuffle_cases") and job.jvals[0] == 0:
        raise ValueError("a shuffled J=0 basis has no TROVE numbering (v is the J=0 order)")
    os.makedirs(d, exist_ok=True)
    b0 = _j0_basis(job)
    with open(os.path.join(d, "j0eigen_quanta0.chk"), "w") as f:
        f.write(quanta_text(job, b0))
    for b in job.bsets:
        if b.jval != 0:
            with open(os.path.join(d, f"j0eigen_quanta{b.jval}.chk"), "w") as f:
                f.write(quanta_text(job, b))
    write_dipole(job, os.path.join(d, "j0_extfield.chk"))
    for jind, J in enumerate(job.jvals):
        if jind == 1 and J == job.jvals[0]:
            continue
        for g in range(job.sym["nrep"]):
            with open(os.path.join(d, f"j0eigen_descr{J}_{g + 1}.chk"), "w") as f:
                f.write(descr_text(job, jind, g))
            v = job.vectors.get((jind, g))
            if v is None:
                v = np.zeros((0, int(job.bsets[jind].nsize[g])))
            np.ascontiguousarray(v, np.float64).tofile(
                os.path.join(d, f"j0eigen_vectors{J}_{g + 1}.chk"))
    deck = deck_text(job)
    with open(os.path.join(d, "deck.inp"), "w") as f:
        f.write(deck)
    return deck
