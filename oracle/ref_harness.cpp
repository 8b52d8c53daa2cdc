// C entry points over the reference's own CPU helpers (TEST INFRASTRUCTURE).
//
// Built by oracle/Makefile (target `ref`) from this file plus the reference's
// trove_functions.cpp, fields.cpp and Util.cpp as they lie under
// /root/reference, into oracle/_ref/libtrove_ref.so. Used only by tests to
// pin the oracle's 3j and filter functions against the reference's own code.
// Nothing here is shipped; the reference's CUDA files are not built.
#include <cstdlib>
#include <cstring>

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

}  // extern "C"
