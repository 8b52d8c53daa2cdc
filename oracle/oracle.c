/*
 * FP64 CPU oracle for the line-intensity hot path — TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference's arithmetic, used only as the
 * checker of the HIP path (tests/, __graft_entry__.smoke()) and as the CPU
 * baseline timed by bench.py. The product never links or calls this file.
 *
 * The reference has no CPU implementation of the hot path: the half
 * linestrength, correlation and intensity exist only as CUDA kernels, a
 * cuBLAS ddot and host post-processing. This file restates those steps, term
 * by term and in the reference's summation order, from:
 *   dipole_kernals.cu:38-71     correlate (symmetrised -> primitive)
 *   dipole_kernals.cu:162-227   half linestrength (production kernel)
 *   cuda_host.cu:1174-1407      lower/upper loops, ddot, intensity
 *   cuda_host.cu:12-18          physical constants
 *   trove_functions.cpp:321-399, 899-925  fakt, three_j, 3j table
 *   trove_functions.cpp:1041-1072, 1432-1510  igamma pairs, filters
 *
 * Pinning (which independent check fixes each function; tests/test_oracle.py):
 *   or_fakt, or_three_j, or_precompute_threej: exact sympy values (fixture
 *       tests/golden/threej_sympy.json) and closed forms of (J 1 J'; k q -k').
 *   or_correlate: construction — the generator knows R (signed permutation
 *       and orthogonal multi-term cases); R^T R = 1 checked.
 *   or_half_ls: rotational sum rule sum_J' (2J'+1) M^T M = |mu|^2 1 (fixes 3j
 *       indexing, the 1/sqrt2 factor, the K mask), the reciprocity
 *       M(I<-F) = -M(F<-I)^T (fixes both sign factors), transposition of mu
 *       with a non-symmetric mu, and Honl-London magnitudes at Nvib=1.
 *       Conventions that no physical invariant sees — which of x/y pairs with
 *       tau_F != tau_I and the sign of the y term — are taken as read from
 *       dipole_kernals.cu:206-216 (parity unpinned for those two choices).
 *   correlation (dot): Parseval sum_F A_F^2 = ||P h||^2 with orthogonal U.
 *   or_constants / intensity: closed-form values of the constants.
 *   filters: truth tables from trove_functions.cpp:1432-1510.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* C version of Fortran NINT as used by trove_functions.cpp:20. */
static int nint(double a) { return a >= 0.0 ? (int)(a + 0.5) : (int)(a - 0.5); }

/* a!/10^a for integral a >= 0 (trove_functions.cpp:321-344). */
double or_fakt(double a)
{
    if (fabs(a) < 1e-24) return 1.0;
    if (a < 0.0) return NAN;            /* the reference exits here */
    int ic = nint(a);
    double ax = a / 10.0;
    double f = ax;
    for (int i = 1; i < ic; i++) f *= (ax - (double)i * 0.1);
    return f;
}

/* Wigner 3j symbol via the Racah sum with scaled factorials
 * (trove_functions.cpp:348-399). */
double or_three_j(int j1, int j2, int j3, int k1, int k2, int k3)
{
    double a = j1, b = j2, c = j3, al = k1, be = k2, ga = k3;
    if (c > a + b) return 0.0;
    if (c < fabs(a - b)) return 0.0;
    if (a < 0.0 || b < 0.0 || c < 0.0) return 0.0;
    if (a < fabs(al) || b < fabs(be) || c < fabs(ga)) return 0.0;
    if (-1.0 * ga != al + be) return 0.0;

    double delta = sqrt(or_fakt(a + b - c) * or_fakt(a + c - b) * or_fakt(b + c - a) /
                        or_fakt(a + b + c + 1.0));
    double t1 = or_fakt(a + al) * or_fakt(a - al);
    double t2 = or_fakt(b - be) * or_fakt(b + be);
    double t3 = or_fakt(c + ga) * or_fakt(c - ga);
    double term = sqrt((2.0 * c + 1.0) * t1 * t2 * t3);

    double lo = fmax(fmax(a + be - c, b - c - al), 0.0);
    double hi = fmin(fmin(a - al, b + be), a + b - c);
    int nmin = nint(lo), nmax = nint(hi);
    double summ = 0.0;
    for (int n = nmin; n <= nmax; n++) {
        double dn = (double)n;
        double t4 = or_fakt(a - al - dn) * or_fakt(c - b + al + dn);
        double t5 = or_fakt(b + be - dn) * or_fakt(c - a - be + dn);
        double t6 = or_fakt(dn) * or_fakt(a + b - c - dn);
        summ += pow(-1.0, n) / (t4 * t5 * t6);
    }
    double clebsh = delta * term * summ / sqrt(10.0);
    int iphase = nint(a - b - ga);
    double minus = (iphase % 2 == 0) ? 1.0 : -1.0;
    return minus * clebsh / sqrt(2.0 * c + 1.0);
}

/* T[jI + kI*(jmax+1) + (jF-jI+1)*(jmax+1)^2 + (kF-kI+1)*(jmax+1)^2*3]
 * = (jI 1 jF; kI, kF-kI, -kF) (trove_functions.cpp:899-925). Entries the
 * reference leaves uninitialised are zero here. */
void or_precompute_threej(int jmax, double *T)
{
    int s = jmax + 1;
    memset(T, 0, sizeof(double) * (size_t)s * s * 9);
    for (int jI = 0; jI <= jmax; jI++)
        for (int jF = (jI - 1 > 0 ? jI - 1 : 0); jF <= (jI + 1 < jmax ? jI + 1 : jmax); jF++)
            for (int kI = 0; kI <= jI; kI++)
                for (int kF = (kI - 1 > 0 ? kI - 1 : 0); kF <= (kI + 1 < jF ? kI + 1 : jF); kF++)
                    T[jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * s * s * 3] =
                        or_three_j(jI, 1, jF, kI, kF - kI, -kF);
}

/* out[r] = sum_{e<N} vec[t+e] * repres_g[(t+e) + ideg*Ntot + ib*degen*Ntot]
 * (dipole_kernals.cu:38-71; the host twin at trove_functions.cpp:968-1023 uses
 * degen-1 as the stride, which this follows the device kernel in not doing). */
void or_correlate(const or_bset *b, const int *degen, int nrep, int ideg, int g,
                  const double *vec, double *out)
{
    (void)nrep;
    int dimen = b->Maxcontracts;
    int Ntot = b->Ntotal[g];
    int sdeg = degen[g];
    const double *irr = b->repres[g];
    for (int r = 0; r < dimen; r++) {
        int irow = b->icontr2icase[r];
        int ib = b->icontr2icase[r + dimen];
        int iterm = b->ijterms[irow + g * b->Maxsymcoeffs];
        int nelem = b->N[g * b->Maxsymcoeffs + irow];
        double acc = 0.0;
        for (int i = 0; i < nelem; i++) {
            int is = iterm + i;
            acc += vec[is] * irr[is + ideg * Ntot + ib * sdeg * Ntot];
        }
        out[r] = acc;
    }
}

/* Half linestrength h[rF] for target basis bF from primitive coefficients
 * vec (dimen of bI): dipole_kernals.cu:162-227, with the same factor order
 * per term and, for every rF, the same ascending-rI summation order starting
 * from 0.0, so each h[rF] is the reference's value bit for bit. The loops are
 * interchanged (rI outer) so that mu is read along its contiguous vF axis.
 * With noskip=0 the terms with |kI-kF| > 1, which the reference multiplies
 * by a zero mask, are not visited. */
void or_half_ls(const or_bset *bI, const or_bset *bF, const double *mu, int nvib,
                const double *T, int jmax, const double *vec, double *h, int noskip)
{
    const int dimI = bI->Maxcontracts, dimF = bF->Maxcontracts;
    const int jI = bI->jval, jF = bF->jval;
    const int s = jmax + 1;
    const double sq2 = 1.0 / sqrt(2.0);
    const size_t N = (size_t)nvib, N2 = N * N;

    /* rF lists per kI, ascending rF */
    int kmaxI = 0;
    for (int r = 0; r < dimI; r++) if (bI->k[r] > kmaxI) kmaxI = bI->k[r];
    int *cnt = calloc((size_t)kmaxI + 1, sizeof(int));
    int **lst = calloc((size_t)kmaxI + 1, sizeof(int *));
    for (int kI = 0; kI <= kmaxI; kI++) {
        lst[kI] = malloc(sizeof(int) * (size_t)(dimF > 0 ? dimF : 1));
        for (int r = 0; r < dimF; r++)
            if (noskip || abs(bF->k[r] - kI) <= 1) lst[kI][cnt[kI]++] = r;
    }
    for (int rF = 0; rF < dimF; rF++) h[rF] = 0.0;

    for (int rI = 0; rI < dimI; rI++) {
        const int vI = bI->iroot_correlat_j0[rI];
        const int kI = bI->k[rI];
        const int tauI = bI->ktau[rI] & 1;
        const int sigmaI = (kI % 3) * tauI;
        const int sgnI = 2 * (~(sigmaI + kI) & 1) - 1;
        const double c = vec[rI];
        const int *L = lst[kI];
        for (int q = 0; q < cnt[kI]; q++) {
            int rF = L[q];
            int vF = bF->iroot_correlat_j0[rF];
            int tauF = bF->ktau[rF] & 1;
            int kF = bF->k[rF];
            int diff = abs(kI - kF) <= 1;
            double f3j = T[jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * diff * s * s * 3];
            int keq = (kF == kI), tneq = (tauF != tauI), kzero = ((kI * kF) != 0);
            int a = 2 * keq + (!keq) * (tneq) * 0 + (!keq) * (!tneq) * 1;
            double ls = mu[(size_t)vF + (size_t)vI * N + (size_t)a * N2];
            ls *= (double)(tauF - tauI) * (double)keq + (tauF - tauI) * (kF - kI) * (!keq) * (tneq) +
                  -1.0 * (!keq) * (!tneq);
            h[rF] += (double)sgnI * ls * f3j * c * (double)diff *
                     (1.0 + (sq2 - 1.0) * (double)kzero * (!keq));
        }
    }
    for (int rF = 0; rF < dimF; rF++) {
        int tauF = bF->ktau[rF] & 1;
        int sigmaF = (bF->k[rF] % 3) * tauF;
        h[rF] *= (double)(2 * (~(sigmaF) & 1) - 1);
    }
    for (int kI = 0; kI <= kmaxI; kI++) free(lst[kI]);
    free(lst);
    free(cnt);
}

/* pi, A_coef_s_1, planck, avogno, vellgt, intens_cm_mol, boltz, with the
 * expression shapes of cuda_host.cu:12-18. */
void or_constants(double *o)
{
    double pi = 4.0 * atan2(1.0, 1.0);
    double planck = 6.62606896 * pow(10.0, -27.0);
    double avogno = 6.0221415 * pow(10.0, 23.0);
    double vellgt = 2.99792458 * pow(10.0, 10.0);
    o[0] = pi;
    o[1] = 64.0 * pow(10.0, -36.0) * pow(pi, 4.0) / (3.0 * 6.62606896 * pow(10.0, -27.0));
    o[2] = planck;
    o[3] = avogno;
    o[4] = vellgt;
    o[5] = 8.0 * pow(10.0, -36.0) * pow(pi, 3.0) * avogno / (3.0 * planck * vellgt);
    o[6] = 1.380658 * pow(10.0, -16.0);
}

/* trove_functions.cpp:1432-1441 */
int or_energy_filter_lower(const or_job *j, int J, double e, const int *q)
{
    for (int i = 0; i < j->nmodes; i++) {
        if (q[i + 1] < j->quanta_lower[i]) return 0;
        if (q[i + 1] > j->quanta_lower[j->nmodes + i]) return 0;
    }
    return J >= j->jvals[0] && J <= j->jvals[1] && (e - j->zpe) >= j->erange_lower[0] &&
           (e - j->zpe) <= j->erange_lower[1];
}

/* trove_functions.cpp:1442-1451 */
int or_energy_filter_upper(const or_job *j, int J, double e, const int *q)
{
    for (int i = 0; i < j->nmodes; i++) {
        if (q[i + 1] < j->quanta_upper[i]) return 0;
        if (q[i + 1] > j->quanta_upper[j->nmodes + i]) return 0;
    }
    return J >= j->jvals[0] && J <= j->jvals[1] && (e - j->zpe) >= j->erange_upper[0] &&
           (e - j->zpe) <= j->erange_upper[1];
}

/* trove_functions.cpp:1452-1464, including the exclusion of
 * jI == jF == jvals[0] (:1457). */
int or_intensity_filter(const or_job *j, int jI, int jF, double eI, double eF, int gI, int gF,
                        const int *qI, const int *qF)
{
    double nu = eF - eI;
    int ok = (j->gns[gI] > 0.0) && (nu >= j->freq_window[0]) && (nu <= j->freq_window[1]) &&
             or_energy_filter_lower(j, jI, eI, qI) && or_energy_filter_upper(j, jF, eF, qF);
    if (ok)
        ok = ((jF != j->jvals[0]) || (jI != j->jvals[0])) &&
             (j->isym_pairs[gI] == j->isym_pairs[gF]) && (j->igamma_pair[gI] == gF) &&
             (abs(jF - jI) <= 1) && (jI + jF >= 1);
    return ok;
}

/* trove_functions.cpp:1495-1510 */
int or_degeneracy_filter(const or_job *j, int gI, int gF, int idegI, int idegF)
{
    if (j->degen[gI] == 1 || j->degen[gF] == 1) return 1;
    if (!j->reduced) return 1;
    if (j->sym_type == 2 || j->sym_type == 3) return (idegI == 1 && idegF == 0);
    return 1;
}

/* trove_functions.cpp:1041-1072 (the reference exits when a symmetry has two
 * partners; here the last partner wins and the caller checks). */
void or_find_igamma_pair(int nrep, const int *pairs, int *out)
{
    for (int gI = 0; gI < nrep; gI++) {
        out[gI] = gI;
        for (int gF = 0; gF < nrep; gF++)
            if (gI != gF && pairs[gI] == pairs[gF]) out[gI] = gF;
    }
}

/* trove_functions.cpp:1476-1493 */
static int indF_filter(const or_job *j, int jI, int jF, double eI, int gI, const int *qI)
{
    for (int f = 0; f < j->nlevels; f++) {
        const or_level *L = &j->levels[f];
        if (L->jval != jF) continue;
        if (or_intensity_filter(j, jI, jF, eI, L->energy, gI, L->igamma, qI, L->quanta)) return 1;
    }
    return 0;
}

typedef struct {
    long long n, cap;
    int *iI, *iF;
    double *nu, *S, *A, *I, *amp, *amp_abs;
} linebuf;

static void lb_push(linebuf *b, int iI, int iF, double nu, double S, double A, double I,
                    const double *amp, const double *amp_abs)
{
    if (b->n == b->cap) {
        long long c = b->cap ? b->cap * 2 : 64;
        b->iI = realloc(b->iI, sizeof(int) * c);
        b->iF = realloc(b->iF, sizeof(int) * c);
        b->nu = realloc(b->nu, sizeof(double) * c);
        b->S = realloc(b->S, sizeof(double) * c);
        b->A = realloc(b->A, sizeof(double) * c);
        b->I = realloc(b->I, sizeof(double) * c);
        b->amp = realloc(b->amp, sizeof(double) * 4 * c);
        b->amp_abs = realloc(b->amp_abs, sizeof(double) * 4 * c);
        b->cap = c;
    }
    long long k = b->n++;
    b->iI[k] = iI; b->iF[k] = iF; b->nu[k] = nu; b->S[k] = S; b->A[k] = A; b->I[k] = I;
    memcpy(b->amp + 4 * k, amp, sizeof(double) * 4);
    memcpy(b->amp_abs + 4 * k, amp_abs, sizeof(double) * 4);
}

static void lb_free(linebuf *b)
{
    free(b->iI); free(b->iF); free(b->nu); free(b->S); free(b->A); free(b->I);
    free(b->amp); free(b->amp_abs);
}

/* Body of the lower-level loop, cuda_host.cu:1174-1488: half linestrengths
 * for every allowed (J', ideg) of the lower level (:1204-1228), then for every
 * upper level passing energy_filter_upper and intensity_filter (:1235-1266)
 * the correlation over dimen_F (:1284-1322), the degeneracy sum (:1388-1398)
 * and the intensity (:1400-1407). */
static void run_one(const or_job *j, int ilI, int noskip, double *cons, double beta,
                    double *wI, double *wF, double *const *hls, linebuf *out)
{
    const or_level *LI = &j->levels[ilI];
    const int indI = LI->jind, jI = LI->jval, gI = LI->igamma, ndegI = LI->ndeg;
    const double eI = LI->energy;
    const or_bset *bI = j->bset[indI];
    const int nsizeI = bI->nsize[gI];
    if (!or_energy_filter_lower(j, jI, eI, LI->quanta)) return;
    const double *uI = j->vectors[indI * j->nrepres + gI] + (size_t)LI->irec0 * nsizeI;
    const int gF_t = j->igamma_pair[gI];

    for (int indF = 0; indF < 2; indF++) {
        int jF = j->jvals[indF];
        if (!indF_filter(j, jI, jF, eI, gI, LI->quanta)) continue;
        for (int ideg = 0; ideg < ndegI; ideg++) {
            if (!or_degeneracy_filter(j, gI, gF_t, ideg, 0)) continue;
            or_correlate(bI, j->degen, j->nrepres, ideg, gI, uI, wI);
            or_half_ls(bI, j->bset[indF], j->dipole_flipped, j->nvib, j->threej, j->jmax, wI,
                       hls[indF * 2 + ideg], noskip);
        }
    }

    const double A_coef_s_1 = cons[1], intens_cm_mol = cons[5];
    for (int ilF = 0; ilF < j->nlevels; ilF++) {
        const or_level *LF = &j->levels[ilF];
        const int indF = LF->jind, jF = LF->jval, gF = LF->igamma, ndegF = LF->ndeg;
        const double eF = LF->energy;
        if (!or_energy_filter_upper(j, jF, eF, LF->quanta)) continue;
        if (!or_intensity_filter(j, jI, jF, eI, eF, gI, gF, LI->quanta, LF->quanta)) continue;
        const or_bset *bF = j->bset[indF];
        const int nsizeF = bF->nsize[gF], dimF = bF->Maxcontracts;
        const double *uF = j->vectors[indF * j->nrepres + gF] + (size_t)LF->irec0 * nsizeF;
        double amp[4] = {0, 0, 0, 0}, amp_abs[4] = {0, 0, 0, 0};
        for (int idegF = 0; idegF < ndegF; idegF++) {
            if (j->reduced && idegF != 0) continue;
            for (int idegI = 0; idegI < ndegI; idegI++) {
                if (!or_degeneracy_filter(j, gI, gF, idegI, idegF)) continue;
                or_correlate(bF, j->degen, j->nrepres, idegF, gF, uF, wF);
                const double *h = hls[indF * 2 + idegI];
                double d = 0.0, da = 0.0;
                for (int r = 0; r < dimF; r++) {
                    d += wF[r] * h[r];
                    da += fabs(wF[r] * h[r]);
                }
                amp[idegI + idegF * 2] = d;
                amp_abs[idegI + idegF * 2] = da;
            }
        }
        double ls = 0.0;
        for (int idegF = 0; idegF < ndegF; idegF++)
            for (int idegI = 0; idegI < ndegI; idegI++) {
                double a = amp[idegI + idegF * 2];
                ls += a * a;
            }
        ls /= (double)ndegI;
        if (j->reduced && ndegF != 1 && ndegI != 1) ls *= (double)ndegI;
        double nu = eF - eI;
        double anu = fabs(nu);
        double boltz_fc = anu * exp(-(eI - j->zpe) * beta) * (1.0 - exp(-anu * beta)) / j->q_stat;
        double A_einst = A_coef_s_1 * (double)((2 * jI) + 1) * ls * pow(anu, 3);
        double S = ls * j->gns[gI] * (double)((2 * jI + 1) * (2 * jF + 1));
        double I = S * intens_cm_mol * boltz_fc;
        if (S >= j->thresh_linestrength && I >= j->thresh_intensity)
            lb_push(out, ilI, ilF, nu, S, A_einst, I, amp, amp_abs);
    }
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int or_run(const or_job *j, int nlower, const int *ids, int noskip, int nthreads, or_lines *out)
{
    memset(out, 0, sizeof(*out));
    if (!ids) nlower = j->nlevels;
    double cons[7];
    or_constants(cons);
    const double beta = cons[2] * cons[4] / (cons[6] * j->temperature); /* cuda_host.cu:1170 */
    int dmax = 1;
    for (int i = 0; i < 2; i++)
        if (j->bset[i]->Maxcontracts > dmax) dmax = j->bset[i]->Maxcontracts;
    linebuf *per = calloc((size_t)(nlower > 0 ? nlower : 1), sizeof(linebuf));
    if (!per) return 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *wI = malloc(sizeof(double) * dmax);
        double *wF = malloc(sizeof(double) * dmax);
        double *hbuf = calloc((size_t)dmax * 4, sizeof(double));
        double *hls[4] = {hbuf, hbuf + dmax, hbuf + 2 * (size_t)dmax, hbuf + 3 * (size_t)dmax};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int li = 0; li < nlower; li++)
            run_one(j, ids ? ids[li] : li, noskip, cons, beta, wI, wF, hls, &per[li]);
        free(wI); free(wF); free(hbuf);
    }
    long long n = 0;
    for (int li = 0; li < nlower; li++) n += per[li].n;
    long long c = n > 0 ? n : 1;
    out->n = n;
    out->ilevelI = malloc(sizeof(int) * c);
    out->ilevelF = malloc(sizeof(int) * c);
    out->nu = malloc(sizeof(double) * c);
    out->S = malloc(sizeof(double) * c);
    out->A_einst = malloc(sizeof(double) * c);
    out->I = malloc(sizeof(double) * c);
    out->amp = malloc(sizeof(double) * 4 * c);
    out->amp_abs = malloc(sizeof(double) * 4 * c);
    long long o = 0;
    for (int li = 0; li < nlower; li++) {
        linebuf *b = &per[li];
        if (b->n) {
            memcpy(out->ilevelI + o, b->iI, sizeof(int) * b->n);
            memcpy(out->ilevelF + o, b->iF, sizeof(int) * b->n);
            memcpy(out->nu + o, b->nu, sizeof(double) * b->n);
            memcpy(out->S + o, b->S, sizeof(double) * b->n);
            memcpy(out->A_einst + o, b->A, sizeof(double) * b->n);
            memcpy(out->I + o, b->I, sizeof(double) * b->n);
            memcpy(out->amp + 4 * o, b->amp, sizeof(double) * 4 * b->n);
            memcpy(out->amp_abs + 4 * o, b->amp_abs, sizeof(double) * 4 * b->n);
            o += b->n;
        }
        lb_free(b);
    }
    free(per);
    return 0;
}

void or_free_lines(or_lines *o)
{
    free(o->ilevelI); free(o->ilevelF); free(o->nu); free(o->S); free(o->A_einst); free(o->I);
    free(o->amp); free(o->amp_abs);
    memset(o, 0, sizeof(*o));
}
