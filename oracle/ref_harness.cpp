// C entry points over the reference's own CPU helpers (TEST INFRASTRUCTURE).
//
// Built by oracle/Makefile (target `ref`) from this file plus the reference's
// trove_functions.cpp, fields.cpp and Util.cpp as they lie under
// /root/reference, into oracle/_ref/libtrove_ref.so. Used only by tests to
// pin the oracle's 3j and filter functions against the reference's own code.
// Nothing here is shipped; the reference's CUDA files are not built. The
// job reader below runs the reference's input stage on TROVE files the tests
// wrote, so that the build's own readers can be compared with it.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <iostream>
#include <unistd.h>

#include "fields.h"
#include "trove_functions.h"

int G_VERBOSE = 0;  // declared by the reference (common.h:9), never defined there

extern "C" {

double ref_fakt(double a) { return fakt(a); }

double ref_three_j(int j1, int j2, int j3, int k1, int k2, int k3)
{
    return three_j(j1, j2, j3, k1, k2, k3);
}

// precompute_threej leaves entries outside the valid range uninitialised;
// `valid` marks the entries it wrote.
void ref_precompute_threej(int jmax, double *out, int *valid)
{
    double *t = nullptr;
    precompute_threej(&t, jmax);
    const int s = jmax + 1;
    const int n = s * s * 9;
    std::memset(valid, 0, sizeof(int) * n);
    for (int jI = 0; jI <= jmax; jI++)
        for (int jF = (jI - 1 > 0 ? jI - 1 : 0); jF <= (jI + 1 < jmax ? jI + 1 : jmax); jF++)
            for (int kI = 0; kI <= jI; kI++)
                for (int kF = (kI - 1 > 0 ? kI - 1 : 0); kF <= (kI + 1 < jF ? kI + 1 : jF); kF++)
                    valid[jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * s * s * 3] = 1;
    for (int i = 0; i < n; i++) out[i] = valid[i] ? t[i] : 0.0;
    delete[] t;
}

struct RefFilterJob {
    FintensityJob job;
    int *qlo[2], *qhi[2];
};

// Job fields read by the filters (trove_functions.cpp:1432-1510).
void *ref_filter_job(int nrep, const int *degen, int sym_type, int nmodes, double zpe,
                     const double *freq_window, const double *erange_lower,
                     const double *erange_upper, const int *jvals, const double *gns,
                     const int *isym_pairs, const int *igamma_pair, int reduced,
                     const int *quanta_lower, const int *quanta_upper)
{
    RefFilterJob *r = new RefFilterJob();
    FintensityJob &j = r->job;
    j.molec.nmodes = nmodes;
    j.molec.sym_nrepres = nrep;
    j.molec.sym_type = sym_type;
    j.molec.sym_degen = new int[nrep];
    std::memcpy(j.molec.sym_degen, degen, sizeof(int) * nrep);
    j.ZPE = zpe;
    for (int i = 0; i < 2; i++) {
        j.freq_window[i] = freq_window[i];
        j.erange_lower[i] = erange_lower[i];
        j.erange_upper[i] = erange_upper[i];
        j.jvals[i] = jvals[i];
    }
    j.gns = new double[nrep];
    std::memcpy(j.gns, gns, sizeof(double) * nrep);
    j.isym_pairs = new int[nrep];
    std::memcpy(j.isym_pairs, isym_pairs, sizeof(int) * nrep);
    j.igamma_pair = new int[nrep];
    std::memcpy(j.igamma_pair, igamma_pair, sizeof(int) * nrep);
    j.reduced = reduced != 0;
    j.quanta_lower = new int *[2];
    j.quanta_upper = new int *[2];
    for (int k = 0; k < 2; k++) {
        j.quanta_lower[k] = new int[nmodes > 0 ? nmodes : 1];
        j.quanta_upper[k] = new int[nmodes > 0 ? nmodes : 1];
        for (int i = 0; i < nmodes; i++) {
            j.quanta_lower[k][i] = quanta_lower[k * nmodes + i];
            j.quanta_upper[k][i] = quanta_upper[k * nmodes + i];
        }
    }
    return r;
}

int ref_intensity_filter(void *h, int jI, int jF, double eI, double eF, int gI, int gF,
                         int *qI, int *qF)
{
    return intensity_filter(((RefFilterJob *)h)->job, jI, jF, eI, eF, gI, gF, qI, qF) ? 1 : 0;
}

int ref_energy_filter_lower(void *h, int J, double e, int *q)
{
    return energy_filter_lower(((RefFilterJob *)h)->job, J, e, q) ? 1 : 0;
}

int ref_energy_filter_upper(void *h, int J, double e, int *q)
{
    return energy_filter_upper(((RefFilterJob *)h)->job, J, e, q) ? 1 : 0;
}

int ref_degeneracy_filter(void *h, int gI, int gF, int idegI, int idegF)
{
    return degeneracy_filter(((RefFilterJob *)h)->job, gI, gF, idegI, idegF) ? 1 : 0;
}

void ref_find_igamma_pair(void *h, int *out)
{
    FintensityJob &j = ((RefFilterJob *)h)->job;
    find_igamma_pair(j);
    std::memcpy(out, j.igamma_pair, sizeof(int) * j.molec.sym_nrepres);
}


// The reference's input stage (dipole_initialise_cpu, cuda_host.cu:837-898,
// without its GPU part) in directory `dir`, with the deck file as stdin.
// Its console chatter goes to /dev/null.
struct RefJob {
    FintensityJob job;
};

void *ref_read_job(const char *dir, const char *deck_path)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return nullptr;
    if (chdir(dir) != 0) return nullptr;
    if (!std::freopen(deck_path, "r", stdin)) {
        if (chdir(cwd) != 0) return nullptr;
        return nullptr;
    }
    std::cin.clear();
    std::fflush(stdout);
    const int saved = dup(1), nul = open("/dev/null", O_WRONLY);
    dup2(nul, 1);
    RefJob *r = new RefJob();
    FintensityJob &j = r->job;
    read_fields(&j);
    bset_contr_factory(&j.bset_contr[0], 0, j.molec.sym_degen, j.molec.sym_nrepres);
    j.molec.nclasses = j.bset_contr[0].Nclasses;
    bset_contr_factory(&j.bset_contr[1], j.jvals[0], j.molec.sym_degen, j.molec.sym_nrepres);
    bset_contr_factory(&j.bset_contr[2], j.jvals[1], j.molec.sym_degen, j.molec.sym_nrepres);
    correlate_index(j.bset_contr[0], j.bset_contr[0]);
    correlate_index(j.bset_contr[0], j.bset_contr[1]);
    correlate_index(j.bset_contr[0], j.bset_contr[2]);
    read_dipole_flipped(j.bset_contr[0], &j.dipole_me, j.dip_size);
    compute_ijterms(j.bset_contr[1], &j.bset_contr[1].ijterms, j.molec.sym_nrepres);
    compute_ijterms(j.bset_contr[2], &j.bset_contr[2].ijterms, j.molec.sym_nrepres);
    read_eigenvalues(j);
    find_igamma_pair(j);
    std::fflush(stdout);
    std::cout.flush();
    dup2(saved, 1);
    close(saved);
    close(nul);
    if (chdir(cwd) != 0) return r;
    return r;
}

// Sizes: nrep, nmodes, nclasses, nlevels, N (J=0 contracted functions).
void ref_job_sizes(void *h, int *out)
{
    FintensityJob &j = ((RefJob *)h)->job;
    out[0] = j.molec.sym_nrepres;
    out[1] = j.molec.nmodes;
    out[2] = j.molec.nclasses;
    out[3] = j.Neigenlevels;
    out[4] = j.bset_contr[0].Maxcontracts;
}

// Deck fields as read_fields leaves them (isym_do from GNS).
void ref_job_params(void *h, double *d, int *i)
{
    FintensityJob &j = ((RefJob *)h)->job;
    const int nrep = j.molec.sym_nrepres, nm = j.molec.nmodes;
    d[0] = j.ZPE; d[1] = j.temperature; d[2] = j.q_stat; d[3] = j.thresh_linestrength;
    d[4] = j.freq_window[0]; d[5] = j.freq_window[1];
    d[6] = j.erange_lower[0]; d[7] = j.erange_lower[1];
    d[8] = j.erange_upper[0]; d[9] = j.erange_upper[1];
    for (int g = 0; g < nrep; g++) d[10 + g] = j.gns[g];
    i[0] = j.jvals[0]; i[1] = j.jvals[1]; i[2] = j.reduced ? 1 : 0; i[3] = j.molec.sym_type;
    for (int g = 0; g < nrep; g++) {
        i[4 + g] = j.isym_pairs[g];
        i[4 + nrep + g] = j.isym_do[g] ? 1 : 0;
        i[4 + 2 * nrep + g] = j.igamma_pair[g];
    }
    int *q = i + 4 + 3 * nrep;
    for (int m = 0; m < nm; m++) {
        q[m] = j.quanta_lower[0][m]; q[nm + m] = j.quanta_lower[1][m];
        q[2 * nm + m] = j.quanta_upper[0][m]; q[3 * nm + m] = j.quanta_upper[1][m];
    }
}

// Basis `which` (0: J=0, 1: jvals[0], 2: jvals[1]): sizes then arrays in the
// layout of gi_bset (N transposed to [g][case]).
void ref_job_bset_sizes(void *h, int which, int *out)
{
    TO_bset_contrT &b = ((RefJob *)h)->job.bset_contr[which];
    out[0] = b.jval; out[1] = b.Maxsymcoeffs; out[2] = b.Maxcontracts; out[3] = b.mat_size;
}

void ref_job_ntotal(void *h, int which, int *out)
{
    FintensityJob &j = ((RefJob *)h)->job;
    for (int g = 0; g < j.molec.sym_nrepres; g++) out[g] = j.bset_contr[which].Ntotal[g];
}

void ref_job_bset(void *h, int which, int *icontr2icase, int *iroot, int *k, int *ktau, int *N,
                  int *Ntotal, int *ijterms, int *nsize, double **repres)
{
    FintensityJob &j = ((RefJob *)h)->job;
    TO_bset_contrT &b = j.bset_contr[which];
    const int nrep = j.molec.sym_nrepres, d = b.Maxcontracts, nc = b.Maxsymcoeffs;
    std::memcpy(icontr2icase, b.icontr2icase, sizeof(int) * 2 * d);
    std::memcpy(iroot, b.iroot_correlat_j0, sizeof(int) * d);
    std::memcpy(k, b.k, sizeof(int) * d);
    std::memcpy(ktau, b.ktau, sizeof(int) * d);
    for (int g = 0; g < nrep; g++) {
        Ntotal[g] = b.Ntotal[g];
        for (int c = 0; c < nc; c++) N[g * nc + c] = b.irr[g].N[c];
        if (which > 0) {
            for (int c = 0; c < nc; c++) ijterms[c + g * nc] = b.ijterms[c + g * nc];
            nsize[g] = j.isym_do[g] ? b.nsize[g] : -1;
        }
        std::memcpy(repres[g], b.irr[g].repres,
                    sizeof(double) * b.Ntotal[g] * j.molec.sym_degen[g] * b.mat_size);
    }
}

void ref_job_dipole(void *h, double *out)
{
    FintensityJob &j = ((RefJob *)h)->job;
    const size_t n = j.bset_contr[0].Maxcontracts;
    std::memcpy(out, j.dipole_me, sizeof(double) * 3 * n * n);
}

// Level table in the reference's order (its qsort).
void ref_job_levels(void *h, double *energy, int *jind, int *igamma, int *ndeg, int *irec0,
                    int *ilevel, int *quanta, int *krot, int *taurot)
{
    FintensityJob &j = ((RefJob *)h)->job;
    const int nm = j.molec.nmodes;
    for (int i = 0; i < j.Neigenlevels; i++) {
        const TO_PTeigen &e = j.eigen[i];
        energy[i] = e.energy; jind[i] = e.jind; igamma[i] = e.igamma; ndeg[i] = e.ndeg;
        irec0[i] = e.irec[0] - 1; ilevel[i] = e.ilevel; krot[i] = e.krot; taurot[i] = e.taurot;
        for (int q = 0; q <= nm; q++) quanta[i * (nm + 1) + q] = e.quanta[q];
    }
}

}  // extern "C"
