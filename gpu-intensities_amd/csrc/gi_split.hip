// Split-precision products (GI_MODE_SPLIT) on the gfx950 int8 matrix cores
// (v_mfma_i32_16x16x64_i8).
//
// Each operand row (A) or column (B) along K is scaled by its own power of
// two and cut into S signed int8 digits of radix R (Ozaki splitting):
//   x = 2^e * sum_{i<S} d_i * R^-(i+1) + r,   |d_i| <= R/2,
//   |r| <= 2^(e-1) * R^-S,   with e such that max|x| <= 2^(e-1).
// R = 254 gives digits in [-127, 127] (8 bits each, the whole signed int8
// range but -128); R = 128 gives 7-bit digits in [-64, 64]. A digit product
// is at most (R/2)^2 in magnitude, so a K sum of them is exact in int32 while
// S * (R/2)^2 * K < 2^31 (R = 254, S = 5: K < 26,600; R = 128, S = 6:
// K < 87,300); the context takes R = 254 when every product allows it. The
// product of a row and a column is then
//   sum_k a_k b_k = 2^(ea+eb) * sum_{l<S} R^-(l+2) * sum_{i+j=l} sum_k da_ik db_jk
// plus the dropped terms (i + j >= S and the remainders), i.e. a relative
// error of about R^-S of the row and column maxima, with S(S+1)/2 digit-pair
// products (15 for R = 254, S = 5 against 21 for R = 128, S = 6 at about the
// same accuracy). The integer sums are kept per level l and combined in FP64
// in the epilogue; everything around the products (G build, fold, intensity)
// stays FP64.
//
// Digit planes: plane q of an operand with R rows of Kp bytes is R*Kp bytes;
// rows are padded with zeros to a multiple of 64 so that a K step never
// needs a bounds check. Tiles are 64 x 64 outputs per 256-thread workgroup
// (4 wavefronts of 32 x 32 as 2 x 2 MFMA blocks), K in steps of 64 staged
// through LDS in unpadded rows of kRowB = 64 bytes (measured on gfx950: both
// the staging stores, 4 lanes per 64-byte row, and the operand reads, 16 rows
// of one 16-byte chunk, then run at the LDS rate of 128 B/clk; an 80-byte
// pitch halves the rate of the stores). Operand layout of the 16x16x64 i8 instruction
// (checked on the device): lane l supplies A[l%16][16*(l/16) .. +15] and
// B[16*(l/16) .. +15][l%16]; accumulator register r holds D[4*(l/16)+r][l%16].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gi_kernels.h"

namespace gi {

constexpr int kRowB = 64;  // LDS row pitch of the digit tiles, bytes

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ inline int digit_exp(double m)
{
    if (!(m > 0.0)) return 0;
    int ex;
    (void)frexp(m, &ex);  // m < 2^ex
    return ex + 1;        // m <= 2^(e-1)
}

// S digits of radix R (128 or 254) of x / 2^e, |x| <= 2^(e-1): each step
// scales the remainder (|r| <= 1/2) by R and rounds, so |d| <= R/2 (64 or 127,
// inside int8) and the next remainder is again within 1/2. With R = 128 every
// step is exact; with R = 254 the product r * R rounds by at most 2^-53 of it.
template <int S>
__device__ inline void to_digits(double x, int e, int (&d)[S], double R)
{
    double r = ldexp(x, -e);
#pragma unroll
    for (int i = 0; i < S; i++) {
        const double t = r * R;
        const double v = rint(t);
        d[i] = (int)v;
        r = t - v;
    }
}

__device__ inline double block_max(double m, double *red)
{
    red[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + k]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Rows of a row-major FP64 matrix (row r at X + r*ld, K valid entries) to
// digit planes D[q*plane + r*Kp + k] and exponents E[r]; optional row maxima.
template <int S>
__device__ inline void split_row(const double *__restrict__ X, long long ld, int K, int Kp,
                                 int8_t *__restrict__ D, long long plane, int *__restrict__ E,
                                 double *__restrict__ Mx, long long r, double *red, double R)
{
    const double *x = X + r * ld;
    double m = 0.0;
    for (int k = threadIdx.x; k < K; k += 256) m = fmax(m, fabs(x[k]));
    m = block_max(m, red);
    const int e = digit_exp(m);
    if (threadIdx.x == 0) {
        E[r] = e;
        if (Mx) Mx[r] = m;
    }
    int8_t *d = D + r * Kp;
    for (int k = threadIdx.x; k < Kp; k += 256) {
        int dg[S];
        to_digits<S>(k < K ? x[k] : 0.0, e, dg, R);
#pragma unroll
        for (int q = 0; q < S; q++) d[q * plane + k] = (int8_t)dg[q];
    }
}

template <int S>
__global__ __launch_bounds__(256) void k_split_rows(const double *__restrict__ X, long long ld,
                                                    int K, int Kp, int8_t *__restrict__ D,
                                                    long long plane, int *__restrict__ E,
                                                    double *__restrict__ Mx, double R)
{
    __shared__ double red[256];
    split_row<S>(X, ld, K, Kp, D, plane, E, Mx, blockIdx.x, red, R);
}

template <int S>
__global__ __launch_bounds__(256) void k_split_rows_jobs(const SplitRowsJob *__restrict__ jobs,
                                                         const int *__restrict__ begin,
                                                         int njobs, double R)
{
    __shared__ double red[256];
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    const SplitRowsJob &J = jobs[lo];
    split_row<S>(J.X, J.ld, J.K, J.Kp, J.D, J.plane, J.E, nullptr, blockIdx.x - begin[lo], red,
                 R);
}

// mu (permuted, mu_a(vF, vI) at vF + vI*ld + a*N*ld: ld = N for the whole
// mu, ld = rows of one part of vF rows) to digit planes with row (a, vF) at
// (a*ld + vF)*kpad_all and column vpad(vI) (classes padded to 64; vinv[p] =
// vI or -1 for padding); one exponent per vF over all a and vI.
template <int S>
__global__ __launch_bounds__(256) void k_split_mu(int N, int ld, int kpad_all,
                                                  const double *__restrict__ mu,
                                                  const int *__restrict__ vinv,
                                                  int8_t *__restrict__ D, long long plane,
                                                  int *__restrict__ E, double R)
{
    __shared__ double red[256];
    const int vF = blockIdx.x;
    const long long NL = (long long)N * ld;
    double m = 0.0;
    for (int i = threadIdx.x; i < 3 * N; i += 256) {
        const int a = i / N, vI = i - a * N;
        m = fmax(m, fabs(mu[vF + (long long)vI * ld + a * NL]));
    }
    m = block_max(m, red);
    const int e = digit_exp(m);
    if (threadIdx.x == 0) E[vF] = e;
    for (int i = threadIdx.x; i < 3 * kpad_all; i += 256) {
        const int a = i / kpad_all, p = i - a * kpad_all;
        const int vI = vinv[p];
        int dg[S];
        to_digits<S>(vI >= 0 ? mu[vF + (long long)vI * ld + a * NL] : 0.0, e, dg, R);
        int8_t *d = D + ((long long)a * ld + vF) * kpad_all + p;
#pragma unroll
        for (int q = 0; q < S; q++) d[q * plane] = (int8_t)dg[q];
    }
}

__device__ inline int find_tile_job(const int *__restrict__ begin, int njobs, int t)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ inline int xcd_tile_s(int w, int grid) { return (w & 7) * (grid >> 3) + (w >> 3); }
static inline int xcd_grid_s(int ntiles) { return (ntiles + 7) & ~7; }

// G of every unit as digits (split layout of the sparse record map): thread
// = one row of one column, up to kGUnits units per workgroup. The column
// exponent comes from the bound max_rows sum|w| * max|u| of the unit's
// record; the thread holding row 0 of a column stores it.
template <int S>
__global__ __launch_bounds__(256) void k_build_g_split(const GSplitJob *__restrict__ jobs,
                                                       const int *__restrict__ begin, int njobs,
                                                       int ntiles, double R)
{
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const GSplitJob &J = jobs[find_tile_job(begin, njobs, tile)];
    const int local = tile - J.tile_begin;
    const int ug = local / J.gtiles, gt = local - ug * J.gtiles;
    const int g = gt * 256 + (int)threadIdx.x;
    if (g >= J.GtotB) return;
    int co = 0;
    while (co + 1 < J.nclasses && J.classes[co + 1].cbaseB <= g) co++;
    const ClassDev &cd = J.classes[co];
    const int j = (g - cd.cbaseB) / cd.kpad;
    const bool first = (g - cd.cbaseB) - j * cd.kpad == 0;
    int ix[kGMapMaxW];
    double wv[kGMapMaxW];
    double bnd = 0.0;
    int cur = -1;
    const int u0 = ug * kGUnits, u1 = min(u0 + kGUnits, J.nunits);
    for (int u = u0; u < u1; u++) {
        const int d = J.uideg[u];
        if (d != cur) {
            const int *ip = J.idx[d] + g;
            const double *wp = J.wt[d] + g;
#pragma unroll
            for (int w = 0; w < kGMapMaxW; w++)
                if (w < J.W) {
                    ix[w] = ip[(size_t)w * J.GtotB];
                    wv[w] = wp[(size_t)w * J.GtotB];
                }
            bnd = J.bnd[d][cd.eoff + j];
            cur = d;
        }
        const double *rec = J.uptr[u];
        const int e = digit_exp(bnd * J.umax[J.urec[u]]);
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kGMapMaxW; w++)
            if (w < J.W) v += wv[w] * rec[ix[w]];
        int dg[S];
        to_digits<S>(v, e, dg, R);
        int8_t *dst = J.D + (size_t)u * J.GtotB + g;
#pragma unroll
        for (int q = 0; q < S; q++) dst[q * J.plane] = (int8_t)dg[q];
        if (first) J.E[(size_t)u * J.EGtot + cd.eoff + j] = e;
    }
}

// The same for record maps of width <= 4, four consecutive rows of one column
// per thread (columns are padded to 64 rows): vector loads of the map and one
// 4-byte store per digit plane instead of four byte stores. The maps of both
// lower components stay in registers (units of a degenerate block alternate
// them), and the gathers of four units are issued before the first digit
// store (byte stores may alias any load, so later loads cannot move above
// them). Every element sums the same products in the same order.
template <int S>
__global__ __launch_bounds__(256) void k_build_g_split4(const GSplitJob *__restrict__ jobs,
                                                        const int *__restrict__ begin, int njobs,
                                                        int ntiles, double R)
{
    constexpr int NB = 4;  // units gathered together
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const GSplitJob &J = jobs[find_tile_job(begin, njobs, tile)];
    const int local = tile - J.tile_begin;
    const int ug = local / J.gtiles, gt = local - ug * J.gtiles;
    const int g = (gt * 256 + (int)threadIdx.x) * 4;
    if (g >= J.GtotB) return;
    int co = 0;
    while (co + 1 < J.nclasses && J.classes[co + 1].cbaseB <= g) co++;
    const ClassDev &cd = J.classes[co];
    const int j = (g - cd.cbaseB) / cd.kpad;
    const bool first = (g - cd.cbaseB) - j * cd.kpad == 0;
    int4 ix[2][4];
    double wv[2][4][4], bnd[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int w = 0; w < 4; w++) {
            ix[d][w] = int4{0, 0, 0, 0};
            wv[d][w][0] = wv[d][w][1] = wv[d][w][2] = wv[d][w][3] = 0.0;
            if (w < J.W) {
                const size_t o = (size_t)w * J.GtotB + g;
                ix[d][w] = *(const int4 *)(J.idx[d] + o);
                const double2 a = *(const double2 *)(J.wt[d] + o);
                const double2 b = *(const double2 *)(J.wt[d] + o + 2);
                wv[d][w][0] = a.x; wv[d][w][1] = a.y; wv[d][w][2] = b.x; wv[d][w][3] = b.y;
            }
        }
        bnd[d] = J.bnd[d][cd.eoff + j];
    }
    const int u0 = ug * kGUnits, u1 = min(u0 + kGUnits, J.nunits);
    for (int ub = u0; ub < u1; ub += NB) {
        double v[NB][4];
        int ex[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            v[q][0] = v[q][1] = v[q][2] = v[q][3] = 0.0;
            ex[q] = 0;
            const int u = ub + q;
            if (u >= u1) continue;
            const bool d1 = J.uideg[u] != 0;
            const double *rec = J.uptr[u];
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (w < J.W) {
                    const int4 x = d1 ? ix[1][w] : ix[0][w];
                    v[q][0] += (d1 ? wv[1][w][0] : wv[0][w][0]) * rec[x.x];
                    v[q][1] += (d1 ? wv[1][w][1] : wv[0][w][1]) * rec[x.y];
                    v[q][2] += (d1 ? wv[1][w][2] : wv[0][w][2]) * rec[x.z];
                    v[q][3] += (d1 ? wv[1][w][3] : wv[0][w][3]) * rec[x.w];
                }
            ex[q] = digit_exp((d1 ? bnd[1] : bnd[0]) * J.umax[J.urec[u]]);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = ub + q;
            if (u >= u1) continue;
            int dg[4][S];
#pragma unroll
            for (int r = 0; r < 4; r++) to_digits<S>(v[q][r], ex[q], dg[r], R);
            int8_t *dst = J.D + (size_t)u * J.GtotB + g;
#pragma unroll
            for (int p = 0; p < S; p++)
                *(int *)(dst + p * J.plane) = (dg[0][p] & 255) | ((dg[1][p] & 255) << 8) |
                                              ((dg[2][p] & 255) << 16) | ((dg[3][p] & 255) << 24);
            if (first) J.E[(size_t)u * J.EGtot + cd.eoff + j] = ex[q];
        }
    }
}

// The G rows of one unit for k_build_g_split_lds: NC chunks of four consecutive
// rows of one column per thread and pass (chunks 4096 rows apart), every map
// load of the pass in flight before the first gather from the record in LDS
// (maps of width <= WMAX).
template <int S, int NC, int WMAX, bool PART>
__device__ inline void g_split_passes(const GSplitJob &J, int d, int u, double um,
                                      const double *rec, gptr<double> src, int nl,
                                      const int (*cls)[kClsLds], double R)
{
    // a record longer than nl keeps its tail in global memory (PART)
    auto at = [&](int i) { return PART ? (i < nl ? rec[i] : src[i]) : rec[i]; };
    const gout<int8_t> D = gmem_w(J.D) + (size_t)u * J.GtotB;
    const int W = J.W, Gt = J.GtotB;
    const gptr<double> bnd = gmem(J.bnd[d]);
    const gptr<int> ixp = gmem(J.idx[d]);
    const gptr<double> wtp = gmem(J.wt[d]);
    const gout<int> E = gmem_w(J.E) + (size_t)u * J.EGtot;
    int co = 0;
    for (int g0 = threadIdx.x * 4; g0 < Gt; g0 += 4096 * NC) {
        int4 ix[NC][WMAX];
        double2 wa[NC][WMAX], wb[NC][WMAX];
        double bv[NC];
        int col[NC];  // column index within the unit's exponents; -1: not its first row
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int g = g0 + 4096 * c;
            if (g >= Gt) continue;
            // class of row g from the table in LDS (cbaseB, kpad, eoff)
            while (co + 1 < J.nclasses && cls[0][co + 1] <= g) co++;
            const int j = (g - cls[0][co]) / cls[1][co];
            bv[c] = bnd[cls[2][co] + j];
            col[c] = (g - cls[0][co]) - j * cls[1][co] == 0 ? cls[2][co] + j : -1;
#pragma unroll
            for (int w = 0; w < WMAX; w++)
                if (w < W) {
                    const size_t o = (size_t)w * Gt + g;
                    const v4i t = *(gptr<v4i>)(ixp + o);
                    const d2v a = *(gptr<d2v>)(wtp + o), b = *(gptr<d2v>)(wtp + o + 2);
                    ix[c][w] = make_int4(t.x, t.y, t.z, t.w);
                    wa[c][w] = make_double2(a.x, a.y);
                    wb[c][w] = make_double2(b.x, b.y);
                }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int g = g0 + 4096 * c;
            if (g >= Gt) break;
            const int e = digit_exp(bv[c] * um);
            double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int w = 0; w < WMAX; w++)
                if (w < W) {
                    v[0] += wa[c][w].x * at(ix[c][w].x);
                    v[1] += wa[c][w].y * at(ix[c][w].y);
                    v[2] += wb[c][w].x * at(ix[c][w].z);
                    v[3] += wb[c][w].y * at(ix[c][w].w);
                }
            int dg[4][S];
#pragma unroll
            for (int r = 0; r < 4; r++) to_digits<S>(v[r], e, dg[r], R);
#pragma unroll
            for (int q = 0; q < S; q++)
                *(gout<int>)(D + q * J.plane + g) = (dg[0][q] & 255) | ((dg[1][q] & 255) << 8) |
                                                    ((dg[2][q] & 255) << 16) |
                                                    ((dg[3][q] & 255) << 24);
            if (col[c] >= 0) E[col[c]] = e;
        }
    }
}

// G digits of one unit per workgroup: the unit's record is read once, with
// coalesced loads, into LDS, and the gathers of the record map read it there
// (they land on scattered cache lines when they read global memory); four
// consecutive rows of one column per thread. PART: records longer than nlds
// keep their tail in global memory, where the gathers beyond nlds read it.
template <int S, bool PART>
__global__ __launch_bounds__(1024) void k_build_g_split_lds(const GSplitJob *__restrict__ jobs,
                                                            const int *__restrict__ ubegin,
                                                            int njobs, double R)
{
    __shared__ double rec[kRecLds];
    __shared__ int cls[3][kClsLds];  // per output class: cbaseB, kpad, eoff
    const int ji = find_tile_job(ubegin, njobs, blockIdx.x);
    const GSplitJob &J = jobs[ji];
    const int u = blockIdx.x - ubegin[ji];
    const gptr<double> src = gmem(J.uptr[u]);
    const int nl = PART ? min(J.nsize, J.nlds) : J.nsize;
    for (int i = threadIdx.x; i < nl; i += 1024) rec[i] = src[i];
    if (threadIdx.x < min(J.nclasses, kClsLds)) {
        const ClassDev &cd = J.classes[threadIdx.x];
        cls[0][threadIdx.x] = cd.cbaseB;
        cls[1][threadIdx.x] = cd.kpad;
        cls[2][threadIdx.x] = cd.eoff;
    }
    __syncthreads();
    const int d = J.uideg[u];
    const double um = J.umax[J.urec[u]];
    if (J.W <= 2) g_split_passes<S, 2, 2, PART>(J, d, u, um, rec, src, nl, cls, R);
    else g_split_passes<S, 1, 4, PART>(J, d, u, um, rec, src, nl, cls, R);
}

// Fold and digit split of one folded column per workgroup: the column's rows
// (PER per thread), its maximum, its digits. Columns of at most 12,288 rows
// (PER 12) fit 64 registers, so two workgroups share a CU and hide each
// other's gather latency.
template <int S, int PER>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(PER * 1024 <= kFoldSplitSmall ? 8 : 1))) void k_fold_split_jobs(const FoldSplitJob *__restrict__ jobs,
                                                          const int *__restrict__ begin,
                                                          int njobs, double R)
{
    __shared__ double red[16];
    const int ji = find_tile_job(begin, njobs, blockIdx.x);
    const FoldSplitJob &J = jobs[ji];
    const FoldJob &F = J.f;
    const int col = blockIdx.x - begin[ji];
    const int ncb = F.ndegI * F.ndegF;
    const int item = col / ncb, cb = col - item * ncb;
    const int idegF = cb / F.ndegI, idegI = cb - idegF * F.ndegI;
    const int unit = F.item_unit[item * GI_MAXDEG + idegI];
    const SymDev &T = idegF ? F.t1 : F.t0;
    const double *Hu = F.H + (size_t)(unit >= 0 ? unit : 0) * F.SFN;
    double v[PER];
    double m = 0.0;
    if constexpr (PER <= 16) {
        // two rows at a time with their loads in flight together (fold_rows;
        // four at a time cost the register room of a second workgroup per CU)
#pragma unroll
        for (int k0 = 0; k0 < PER; k0 += 2) {
            int t[2];
            bool ok[2];
            double acc[2];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                t[r] = (int)threadIdx.x + (k0 + r) * 1024;
                ok[r] = t[r] < F.nsizeF && unit >= 0;
            }
            fold_rows<2>(T, Hu, t, ok, acc);
#pragma unroll
            for (int r = 0; r < 2; r++) {
                v[k0 + r] = acc[r];
                m = fmax(m, fabs(acc[r]));
            }
        }
    } else {
        // one row at a time: the 32 values per thread leave no registers to
        // keep several rows' loads in flight (it spills more that way)
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int t = (int)threadIdx.x + k * 1024;
            double acc = 0.0;
            if (t < F.nsizeF && unit >= 0) {
                int2 cc = make_int2(-2, -2);
                if (T.tfc) {  // fixed pairs, as in fold_rows
                    cc = ((const int2 *)T.tfc)[t];
                    const double2 ww = ((const double2 *)T.tfw)[t];
                    if (cc.x >= 0) acc += ww.x * Hu[cc.x];
                    if (cc.y >= 0) acc += ww.y * Hu[cc.y];
                }
                if (cc.x == -2)
                    for (int e = T.tptr[t]; e < T.tptr[t + 1]; e++) acc += T.tval[e] * Hu[T.tcol[e]];
            }
            v[k] = acc;
            m = fmax(m, fabs(acc));
        }
    }
    // workgroup maximum: wavefront reduction, then the 16 wavefront results
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int w = 1; w < 16; w++) m = fmax(m, red[w]);
    const int e = digit_exp(m);
    if (threadIdx.x == 0) J.E[col] = e;
    int8_t *d = J.D + (size_t)col * J.Kp;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int t = (int)threadIdx.x + k * 1024;
        if (t < J.Kp) {
            int dg[S];
            to_digits<S>(t < F.nsizeF ? v[k] : 0.0, e, dg, R);
#pragma unroll
            for (int q = 0; q < S; q++) d[q * J.plane + t] = (int8_t)dg[q];
        }
    }
}

// One K step of 64 for a 64 x 64 tile: the wave's 2 x 2 blocks, digit pairs
// i + j < S into level accumulators. Every B digit plane of the wave's two
// column blocks stays in registers for the step, so LDS delivers each operand
// once per wavefront; the A fragments of digit i + 1 are read before the
// products of digit i (sched_barrier keeps them there), so their LDS latency
// hides behind those products instead of stalling the next group. Each
// accumulator takes its products in the same order as one digit at a time.
template <int S, int AROWS = 64>
__device__ inline void oz_step(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                               int wm, int wn, int lane, v4i (&acc)[2][2][S])
{
    // nothing of the caller's next-step fetch moves past this point: its scalar
    // loads are then complete, and the LDS waits below can count exactly
    __builtin_amdgcn_sched_barrier(0);
    const int i16 = lane & 15, k16 = (lane >> 4) * 16;
    auto ldA = [&](int i, int x) {
        return *(const v4i *)(As + (i * AROWS + wm + 16 * x + i16) * kRowB + k16);
    };
    v4i a0 = ldA(0, 0), a1 = ldA(0, 1);
    v4i b[S][2];
#pragma unroll
    for (int j = 0; j < S; j++) {
        b[j][0] = *(const v4i *)(Bs + (j * 64 + wn + i16) * kRowB + k16);
        b[j][1] = *(const v4i *)(Bs + (j * 64 + wn + 16 + i16) * kRowB + k16);
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
        v4i n0 = a0, n1 = a1;
        if (i + 1 < S) {
            n0 = ldA(i + 1, 0);
            n1 = ldA(i + 1, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < S - i; j++) {
            acc[0][0][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[j][0], acc[0][0][i + j], 0, 0, 0);
            acc[0][1][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[j][1], acc[0][1][i + j], 0, 0, 0);
            acc[1][0][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[j][0], acc[1][0][i + j], 0, 0, 0);
            acc[1][1][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[j][1], acc[1][1][i + j], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        a0 = n0;
        a1 = n1;
    }
}

// oz_step for the banded correlation: only the wave's 16 x 16 blocks whose bit
// (x*2 + y) is set in `need` (wave-uniform) are multiplied.
template <int S>
__device__ inline void oz_step_mask(const int8_t *__restrict__ As,
                                    const int8_t *__restrict__ Bs, int wm, int wn, int lane,
                                    unsigned need, v4i (&acc)[2][2][S])
{
    // nothing of the caller's next-step fetch moves past this point: its scalar
    // loads are then complete, and the LDS waits below can count exactly
    __builtin_amdgcn_sched_barrier(0);
    const int i16 = lane & 15, k16 = (lane >> 4) * 16;
    auto ldA = [&](int i, int x) {
        return *(const v4i *)(As + (i * 64 + wm + 16 * x + i16) * kRowB + k16);
    };
    v4i a0 = ldA(0, 0), a1 = ldA(0, 1);
    v4i b[S][2];
#pragma unroll
    for (int j = 0; j < S; j++) {
        b[j][0] = *(const v4i *)(Bs + (j * 64 + wn + i16) * kRowB + k16);
        b[j][1] = *(const v4i *)(Bs + (j * 64 + wn + 16 + i16) * kRowB + k16);
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
        v4i n0 = a0, n1 = a1;
        if (i + 1 < S) {
            n0 = ldA(i + 1, 0);
            n1 = ldA(i + 1, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < S - i; j++) {
            if (need & 1u)
                acc[0][0][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[j][0], acc[0][0][i + j], 0, 0, 0);
            if (need & 2u)
                acc[0][1][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b[j][1], acc[0][1][i + j], 0, 0, 0);
            if (need & 4u)
                acc[1][0][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[j][0], acc[1][0][i + j], 0, 0, 0);
            if (need & 8u)
                acc[1][1][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b[j][1], acc[1][1][i + j], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        a0 = n0;
        a1 = n1;
    }
}

// sum over levels l of L_l * R^-(l+2), smallest first (rinv = 1/R; exact
// powers of two for R = 128)
template <int S>
__device__ inline double oz_value(const v4i (&acc)[S], int r, double rinv)
{
    double sc[S];
    sc[0] = rinv * rinv;
#pragma unroll
    for (int l = 1; l < S; l++) sc[l] = sc[l - 1] * rinv;
    double v = 0.0;
#pragma unroll
    for (int l = S - 1; l >= 0; l--) v += (double)acc[l][r] * sc[l];
    return v;
}

// Split half linestrength: H_co[m, (u, j)] = sum over segments of mu digits x
// G digits, per (chunk, class) job; m-tiles fastest.
template <int S>
__global__ __launch_bounds__(256) void k_halfls_split(int kpad_all, MuSrc src,
                                                      const int *__restrict__ mue,
                                                      const SplitHalfJob *__restrict__ jobs,
                                                      const int *__restrict__ tbegin, int njobs,
                                                      int ntiles, double R)
{
    __shared__ __attribute__((aligned(16))) int8_t As[S * 64 * kRowB];
    __shared__ __attribute__((aligned(16))) int8_t Bs[S * 64 * kRowB];
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const SplitHalfJob &J = jobs[find_tile_job(tbegin, njobs, tile)];
    const ClassDev &cd = J.cd;
    const int local = tile - J.tile_begin;
    const int m0 = J.m_lo + (local % J.mt) * 64, n0 = (local / J.mt) * 64;
    const int mhi = J.m_hi;
    const int8_t *__restrict__ mud = (const int8_t *)src.mu;
    const long long muplane = src.plane;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int ncolsTot = J.nunits * cd.ncols;
    // loads: row/column tid/4, 16 bytes at K offset 16*(tid%4), all S planes
    const int lr = tid >> 2, lp = (tid & 3) * 16;
    const bool mok = m0 + lr < mhi;
    const int n = n0 + lr;
    const bool nok = n < ncolsTot;
    long long bcol = 0;
    if (nok) {
        const int u = n / cd.ncols, jj = n - u * cd.ncols;
        bcol = (long long)u * J.GtotB + cd.cbaseB + (long long)jj * cd.kpad;
    }
    v4i acc[2][2][S];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int l = 0; l < S; l++) acc[x][y][l] = v4i{0, 0, 0, 0};
    v4i ra[S], rb[S];
    const gptr<int> seg_a = gmem(cd.seg_a), seg_pad = gmem(cd.seg_pad), seg_n = gmem(cd.seg_n);
    auto fetch = [&](int se, int kk) {
        const long long arow =
            ((long long)seg_a[se] * src.ld + cd.off + m0 + lr - src.v0) * kpad_all +
            seg_pad[2 * se + 1] + kk + lp;
        const long long bofs = bcol + seg_pad[2 * se] + kk + lp;
#pragma unroll
        for (int q = 0; q < S; q++) {
            ra[q] = mok ? *(const v4i *)(mud + q * muplane + arow) : v4i{0, 0, 0, 0};
            rb[q] = nok ? *(gptr<v4i>)(gmem(J.Gd) + q * J.gplane + bofs) : v4i{0, 0, 0, 0};
        }
    };
    int e = 0, kk = 0;
    if (cd.nseg > 0) fetch(0, 0);
    while (e < cd.nseg) {
#pragma unroll
        for (int q = 0; q < S; q++) {
            *(v4i *)(As + (q * 64 + lr) * kRowB + lp) = ra[q];
            *(v4i *)(Bs + (q * 64 + lr) * kRowB + lp) = rb[q];
        }
        __syncthreads();
        kk += 64;
        if (kk >= seg_n[e]) {
            e++;
            kk = 0;
        }
        if (e < cd.nseg) fetch(e, kk);
        oz_step<S>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * x + 4 * (lane >> 4) + r;
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (m >= mhi || nn >= ncolsTot) continue;
                const int u = nn / cd.ncols, jj = nn - u * cd.ncols;
                const int ex = mue[cd.off + m] + J.Ge[(size_t)u * J.EGtot + cd.eoff + jj];
                J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] =
                    ldexp(oz_value<S>(acc[x][y], r, 1.0 / R), ex);
            }
}


// k_halfls_split with the operand loads of two K steps in flight: the K steps
// of the class (64 bytes of one segment each) are first listed in LDS as mu
// and G offsets, then a counted loop alternates two register sets, loading
// step s + 2 while step s multiplies. Same products in the same order as
// k_halfls_split: the same bits. Classes with more than kSplitMaxSteps steps use
// k_halfls_split.
// MW_ wavefronts along the rows: a (32*MW_) x 64 tile on 2*MW_ wavefronts of
// 32 x 32 (MW_ 2: 64 x 64 on 4; MW_ 4: 128 x 64 on 8, which stage each K step's
// 64 G columns once for twice the rows); every thread loads one A row's 16 bytes
// per digit, the first 256 one B row's. The 8-wavefront tile runs one
// workgroup per CU, so it double-buffers its LDS operands (up to 120 KB) and
// waits at one barrier per K step: the step that writes a buffer follows the
// barrier of the step after the one that read it.
template <int S, int MW_ = 2>
__global__ __launch_bounds__(128 * MW_) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_halfls_split2(
    int kpad_all, MuSrc src, const int *__restrict__ mue, const SplitHalfJob *__restrict__ jobs,
    const int *__restrict__ tbegin, int njobs, int ntiles, double R)
{
    constexpr int TMR = 32 * MW_, NT = 128 * MW_;
    constexpr bool DB = MW_ == 4;
    constexpr int ASZ = S * TMR * kRowB, BSZ = S * 64 * kRowB;
    __shared__ __attribute__((aligned(16))) int8_t As[(DB ? 2 : 1) * ASZ];
    __shared__ __attribute__((aligned(16))) int8_t Bs[(DB ? 2 : 1) * BSZ];
    int8_t *const As1 = DB ? As + ASZ : As;
    int8_t *const Bs1 = DB ? Bs + BSZ : Bs;
    __shared__ long long s_a[kSplitMaxSteps];
    __shared__ int s_b[kSplitMaxSteps];
    __shared__ int s_first[kSplitMaxSegs + 1];
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const SplitHalfJob &J = jobs[find_tile_job(tbegin, njobs, tile)];
    const ClassDev &cd = J.cd;
    const int local = tile - J.tile_begin;
    const int m0 = J.m_lo + (local % J.mt) * TMR, n0 = (local / J.mt) * 64;
    const int8_t *__restrict__ mud = (const int8_t *)src.mu;
    const long long muplane = src.plane;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % MW_) * 32, wn = (wave / MW_) * 32;
    const int ncolsTot = J.nunits * cd.ncols;
    // step table: segment e's steps start at s_first[e]
    __shared__ int s_fits;
    if (tid == 0) {
        int n = 0;
        for (int e = 0; e < cd.nseg && e < kSplitMaxSegs; e++) {
            s_first[e] = n;
            n += (cd.seg_n[e] + 63) / 64;
        }
        s_first[min(cd.nseg, kSplitMaxSegs)] = n;
        s_fits = split2_class_fits(cd.seg_n, cd.nseg);
    }
    __syncthreads();
    // the host sends only classes within both limits to this kernel; one that
    // is not gets NaN in its tile rather than a sum with K steps left out
    if (!s_fits) {
        for (int i = tid; i < TMR * 64; i += NT) {
            const int m = m0 + (i % TMR), nn = n0 + (i / TMR);
            if (m >= J.m_hi || nn >= ncolsTot) continue;
            const int u = nn / cd.ncols, jj = nn - u * cd.ncols;
            J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] = __builtin_nan("");
        }
        return;
    }
    const int nsteps = s_first[cd.nseg];
    for (int e = 0; e < cd.nseg; e++)
        for (int st = s_first[e] + tid; st < s_first[e + 1]; st += NT) {
            const int kk = (st - s_first[e]) * 64;
            s_a[st] = ((long long)cd.seg_a[e] * src.ld + cd.off + m0 - src.v0) * kpad_all +
                      cd.seg_pad[2 * e + 1] + kk;
            s_b[st] = cd.seg_pad[2 * e] + kk;
        }
    __syncthreads();
    const int lr = tid >> 2, lp = (tid & 3) * 16;
    const bool mok = m0 + lr < J.m_hi;
    const bool bload = tid < 256;  // the threads that stage a B row (64 rows)
    const int n = n0 + lr;
    const bool nok = bload && n < ncolsTot;
    long long bcol = 0;
    if (nok) {
        const int u = n / cd.ncols, jj = n - u * cd.ncols;
        bcol = (long long)u * J.GtotB + cd.cbaseB + (long long)jj * cd.kpad;
    }
    const long long arow = (long long)lr * kpad_all + lp;
    const long long bofs = bcol + lp;
    v4i acc[2][2][S];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int l = 0; l < S; l++) acc[x][y][l] = v4i{0, 0, 0, 0};
    v4i ra0[S], rb0[S], ra1[S], rb1[S];
#define GI_FETCH(RA, RB, ST)                                                                   \
    do {                                                                                       \
        const long long ao = s_a[ST] + arow;                                                   \
        const long long bo = bofs + s_b[ST];                                                   \
        _Pragma("unroll") for (int q = 0; q < S; q++) {                                        \
            RA[q] = mok ? *(const v4i *)(mud + q * muplane + ao) : v4i{0, 0, 0, 0};           \
            RB[q] = nok ? *(gptr<v4i>)(gmem(J.Gd) + q * J.gplane + bo) : v4i{0, 0, 0, 0};          \
        }                                                                                      \
    } while (0)
#define GI_STAGE(AS, BS, RA, RB)                                                               \
    do {                                                                                       \
        _Pragma("unroll") for (int q = 0; q < S; q++) {                                        \
            *(v4i *)(AS + (q * TMR + lr) * kRowB + lp) = RA[q];                                \
            if (bload) *(v4i *)(BS + (q * 64 + lr) * kRowB + lp) = RB[q];                      \
        }                                                                                      \
    } while (0)
    if (nsteps > 0) GI_FETCH(ra0, rb0, 0);
    if (nsteps > 1) GI_FETCH(ra1, rb1, 1);
    for (int st = 0; st < nsteps; st += 2) {
        GI_STAGE(As, Bs, ra0, rb0);
        __syncthreads();
        if (st + 2 < nsteps) GI_FETCH(ra0, rb0, st + 2);
        oz_step<S, TMR>(As, Bs, wm, wn, lane, acc);
        if (!DB) __syncthreads();
        if (st + 1 >= nsteps) break;
        GI_STAGE(As1, Bs1, ra1, rb1);
        __syncthreads();
        if (st + 3 < nsteps) GI_FETCH(ra1, rb1, st + 3);
        oz_step<S, TMR>(As1, Bs1, wm, wn, lane, acc);
        if (!DB) __syncthreads();
    }
#undef GI_FETCH
#undef GI_STAGE
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * x + 4 * (lane >> 4) + r;
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (m >= J.m_hi || nn >= ncolsTot) continue;
                const int u = nn / cd.ncols, jj = nn - u * cd.ncols;
                const int ex = mue[cd.off + m] + J.Ge[(size_t)u * J.EGtot + cd.eoff + jj];
                J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] =
                    ldexp(oz_value<S>(acc[x][y], r, 1.0 / R), ex);
            }
}

// One K step of 64 for a 128 x 64 tile: the wave's 4 x 2 blocks (64 rows x
// 32 columns), so each LDS operand read feeds twice as many products as in
// oz_step; A rows of a plane are 128 apart in LDS.
template <int S>
__device__ inline void oz_step_wide(const int8_t *__restrict__ As,
                                    const int8_t *__restrict__ Bs, int wm, int wn, int lane,
                                    v4i (&acc)[4][2][S])
{
    const int i16 = lane & 15, k16 = (lane >> 4) * 16;
    v4i b[S][2];
#pragma unroll
    for (int j = 0; j < S; j++) {
        b[j][0] = *(const v4i *)(Bs + (j * 64 + wn + i16) * kRowB + k16);
        b[j][1] = *(const v4i *)(Bs + (j * 64 + wn + 16 + i16) * kRowB + k16);
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
        v4i a[4];
#pragma unroll
        for (int x = 0; x < 4; x++)
            a[x] = *(const v4i *)(As + (i * 128 + wm + 16 * x + i16) * kRowB + k16);
#pragma unroll
        for (int j = 0; j < S - i; j++)
#pragma unroll
            for (int x = 0; x < 4; x++) {
                acc[x][0][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[x], b[j][0], acc[x][0][i + j], 0, 0, 0);
                acc[x][1][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[x], b[j][1], acc[x][1][i + j], 0, 0, 0);
            }
    }
}

// k_halfls_split with 128-row tiles (4 wavefronts of 64 x 32): each G digit
// column is read by half as many tiles. The level sums are exact integers,
// so the result is bitwise that of the 64-row kernel.
template <int S>
__global__ __launch_bounds__(256, 1) void k_halfls_split_wide(
    int kpad_all, MuSrc src, const int *__restrict__ mue, const SplitHalfJob *__restrict__ jobs,
    const int *__restrict__ tbegin, int njobs, int ntiles, double R)
{
    __shared__ __attribute__((aligned(16))) int8_t As[S * 128 * kRowB];
    __shared__ __attribute__((aligned(16))) int8_t Bs[S * 64 * kRowB];
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const SplitHalfJob &J = jobs[find_tile_job(tbegin, njobs, tile)];
    const ClassDev &cd = J.cd;
    const int local = tile - J.tile_begin;
    const int m0 = J.m_lo + (local % J.mt) * 128, n0 = (local / J.mt) * 64;
    const int mhi = J.m_hi;
    const int8_t *__restrict__ mud = (const int8_t *)src.mu;
    const long long muplane = src.plane;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
    const int ncolsTot = J.nunits * cd.ncols;
    const int lr = tid >> 2, lp = (tid & 3) * 16;
    const bool mok0 = m0 + lr < mhi, mok1 = m0 + 64 + lr < mhi;
    const int n = n0 + lr;
    const bool nok = n < ncolsTot;
    long long bcol = 0;
    if (nok) {
        const int u = n / cd.ncols, jj = n - u * cd.ncols;
        bcol = (long long)u * J.GtotB + cd.cbaseB + (long long)jj * cd.kpad;
    }
    v4i acc[4][2][S];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int l = 0; l < S; l++) acc[x][y][l] = v4i{0, 0, 0, 0};
    v4i ra0[S], ra1[S], rb[S];
    const gptr<int> seg_a = gmem(cd.seg_a), seg_pad = gmem(cd.seg_pad), seg_n = gmem(cd.seg_n);
    auto fetch = [&](int se, int kk) {
        const long long arow =
            ((long long)seg_a[se] * src.ld + cd.off + m0 + lr - src.v0) * kpad_all +
            seg_pad[2 * se + 1] + kk + lp;
        const long long bofs = bcol + seg_pad[2 * se] + kk + lp;
#pragma unroll
        for (int q = 0; q < S; q++) {
            ra0[q] = mok0 ? *(const v4i *)(mud + q * muplane + arow) : v4i{0, 0, 0, 0};
            ra1[q] = mok1 ? *(const v4i *)(mud + q * muplane + arow + 64LL * kpad_all)
                          : v4i{0, 0, 0, 0};
            rb[q] = nok ? *(gptr<v4i>)(gmem(J.Gd) + q * J.gplane + bofs) : v4i{0, 0, 0, 0};
        }
    };
    int e = 0, kk = 0;
    if (cd.nseg > 0) fetch(0, 0);
    while (e < cd.nseg) {
#pragma unroll
        for (int q = 0; q < S; q++) {
            *(v4i *)(As + (q * 128 + lr) * kRowB + lp) = ra0[q];
            *(v4i *)(As + (q * 128 + 64 + lr) * kRowB + lp) = ra1[q];
            *(v4i *)(Bs + (q * 64 + lr) * kRowB + lp) = rb[q];
        }
        __syncthreads();
        kk += 64;
        if (kk >= seg_n[e]) {
            e++;
            kk = 0;
        }
        if (e < cd.nseg) fetch(e, kk);
        oz_step_wide<S>(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * x + 4 * (lane >> 4) + r;
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (m >= mhi || nn >= ncolsTot) continue;
                const int u = nn / cd.ncols, jj = nn - u * cd.ncols;
                const int ex = mue[cd.off + m] + J.Ge[(size_t)u * J.EGtot + cd.eoff + jj];
                J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] =
                    ldexp(oz_value<S>(acc[x][y], r, 1.0 / R), ex);
            }
}

// Split banded correlation (the digit form of k_corr_mfma): rows = upper
// records (digits of the resident block), columns = folded vectors (digits).
template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_corr_split(
    const CorrSplitArgs *__restrict__ args,
                                                    const int *__restrict__ cbegin, int nchunks,
                                                    int ntiles, double R)
{
    __shared__ __attribute__((aligned(16))) int8_t As[S * 64 * kRowB];
    __shared__ __attribute__((aligned(16))) int8_t Bs[S * 64 * kRowB];
    __shared__ int s_need;
    const int tile = xcd_tile_s(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const int ci = find_tile_job(cbegin, nchunks, tile);
    const CorrSplitArgs &a = args[ci];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncols = a.nitems * a.ncb;
    const int rt = (a.Hi - a.Lo + 63) / 64, per = rt * ((ncols + 63) / 64);
    const int local = tile - cbegin[ci];
    const int kp = local / per, l2 = local - kp * per;
    const int f0 = a.Lo + (l2 % rt) * 64;
    const int n0 = (l2 / rt) * 64;
    const int kb = kp * a.kchunk, ke = min(a.Kp, kb + a.kchunk);
    double *dst = kp == 0 ? a.amp : a.ampx + (long long)(kp - 1) * a.ampx_stride;
    if (tid == 0) {
        int need = 0;
        const int i0 = n0 / a.ncb, i1 = min(a.nitems, (n0 + 64 + a.ncb - 1) / a.ncb);
        for (int i = i0; i < i1 && !need; i++)
            if (a.item_lo[i] < f0 + 64 && a.item_hi[i] > f0) need = 1;
        s_need = need;
    }
    __syncthreads();
    if (!s_need) return;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    // the wave's 16 x 16 blocks outside every window are not multiplied
    unsigned bneed = 0;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) {
            const int r0 = f0 + wm + 16 * x;
            const int col = n0 + wn + 16 * y + (lane & 15);
            bool hit = false;
            if (col < ncols && r0 < a.Hi) {
                const int it = col / a.ncb;
                hit = a.item_lo[it] < r0 + 16 && a.item_hi[it] > r0;
            }
            if (__ballot(hit) != 0ull) bneed |= 1u << (x * 2 + y);
        }
    bneed = __builtin_amdgcn_readfirstlane(bneed);
    const int lr = tid >> 2, lp = (tid & 3) * 16;
    const int fr = f0 + lr;
    const bool fok = fr < a.Hi;
    const long long arow = fok ? (long long)a.blk_col[fr] * a.Kp : 0;
    const int nr = n0 + lr;
    const bool nok = nr < ncols;
    const long long brow = (long long)nr * a.Kp;
    v4i acc[2][2][S];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int l = 0; l < S; l++) acc[x][y][l] = v4i{0, 0, 0, 0};
    if constexpr (S <= 5) {
    // two register sets: the loads of K step k0 + 128 are issued while step k0
    // multiplies (as in k_halfls_split2; at 6 digits they exceed the registers)
    v4i ra0[S], rb0[S], ra1[S], rb1[S];
#define GI_FETCH(RA, RB, K0)                                                                   \
    do {                                                                                       \
        _Pragma("unroll") for (int q = 0; q < S; q++) {                                        \
            RA[q] = fok ? *(gptr<v4i>)(gmem(a.Ud) + q * a.uplane + arow + (K0) + lp)                 \
                        : v4i{0, 0, 0, 0};                                                     \
            RB[q] = nok ? *(gptr<v4i>)(gmem(a.Hd) + q * a.hplane + brow + (K0) + lp)                 \
                        : v4i{0, 0, 0, 0};                                                     \
        }                                                                                      \
    } while (0)
#define GI_STAGE(RA, RB)                                                                       \
    do {                                                                                       \
        _Pragma("unroll") for (int q = 0; q < S; q++) {                                        \
            *(v4i *)(As + (q * 64 + lr) * kRowB + lp) = RA[q];                                 \
            *(v4i *)(Bs + (q * 64 + lr) * kRowB + lp) = RB[q];                                 \
        }                                                                                      \
    } while (0)
    if (kb < ke) GI_FETCH(ra0, rb0, kb);
    if (kb + 64 < ke) GI_FETCH(ra1, rb1, kb + 64);
    for (int k0 = kb; k0 < ke; k0 += 128) {
        GI_STAGE(ra0, rb0);
        __syncthreads();
        if (k0 + 128 < ke) GI_FETCH(ra0, rb0, k0 + 128);
        // raised issue priority while this wavefront multiplies (as in k_corr_mfma;
        // 0.83 -> 0.80 ms at C2)
        __builtin_amdgcn_s_setprio(1);
        oz_step_mask<S>(As, Bs, wm, wn, lane, bneed, acc);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (k0 + 64 >= ke) break;
        GI_STAGE(ra1, rb1);
        __syncthreads();
        if (k0 + 192 < ke) GI_FETCH(ra1, rb1, k0 + 192);
        __builtin_amdgcn_s_setprio(1);
        oz_step_mask<S>(As, Bs, wm, wn, lane, bneed, acc);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }
#undef GI_FETCH
#undef GI_STAGE
    } else {
    v4i ra[S], rb[S];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < S; q++) {
            ra[q] = fok ? *(gptr<v4i>)(gmem(a.Ud) + q * a.uplane + arow + k0 + lp) : v4i{0, 0, 0, 0};
            rb[q] = nok ? *(gptr<v4i>)(gmem(a.Hd) + q * a.hplane + brow + k0 + lp) : v4i{0, 0, 0, 0};
        }
    };
    if (kb < ke) fetch(kb);  // a part past the end of K (kb >= ke) stays zero
    for (int k0 = kb; k0 < ke; k0 += 64) {
#pragma unroll
        for (int q = 0; q < S; q++) {
            *(v4i *)(As + (q * 64 + lr) * kRowB + lp) = ra[q];
            *(v4i *)(Bs + (q * 64 + lr) * kRowB + lp) = rb[q];
        }
        __syncthreads();
        if (k0 + 64 < ke) fetch(k0 + 64);
        __builtin_amdgcn_s_setprio(1);
        oz_step_mask<S>(As, Bs, wm, wn, lane, bneed, acc);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int f = f0 + wm + 16 * x + 4 * (lane >> 4) + r;
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (f >= a.Hi || nn >= ncols) continue;
                const int i = nn / a.ncb, cb = nn - i * a.ncb;
                if (f < a.item_lo[i] || f >= a.item_hi[i]) continue;
                const long long o = a.item_out_off[i] + (f - a.item_lo[i]);
                const int idegF = cb / a.ndegI, idegI = cb - idegF * a.ndegI;
                const int ex = a.Ue[a.blk_col[f]] + a.He[nn];
                dst[(long long)(a.amp_m * a.amp_m) * o + idegI + a.amp_m * idegF] =
                    ldexp(oz_value<S>(acc[x][y], r, 1.0 / R), ex);
            }
}

// ---------------------------------------------------------------------------

#define GI_SPLIT_DISPATCH(S_, CALL)                                                       \
    switch (S_) {                                                                         \
    case 3: CALL(3); break;                                                               \
    case 4: CALL(4); break;                                                               \
    case 5: CALL(5); break;                                                               \
    default: CALL(6); break;                                                              \
    }

void launch_split_rows(hipStream_t st, int S, double R, const double *X, long long ld, int rows, int K,
                       int Kp, int8_t *D, long long plane, int *E, double *Mx)
{
    if (rows <= 0) return;
#define L_(s) hipLaunchKernelGGL(k_split_rows<s>, dim3(rows), dim3(256), 0, st, X, ld, K, Kp, D, plane, E, Mx, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_split_rows_jobs(hipStream_t st, int S, double R, const SplitRowsJob *d_jobs,
                            const int *d_begin, int njobs, int nrows)
{
    if (njobs <= 0 || nrows <= 0) return;
#define L_(s) hipLaunchKernelGGL(k_split_rows_jobs<s>, dim3(nrows), dim3(256), 0, st, d_jobs, d_begin, njobs, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_fold_split_jobs(hipStream_t st, int S, double R, const FoldSplitJob *d_jobs,
                            const int *d_begin, int njobs, int ncols, int maxrows)
{
    if (njobs <= 0 || ncols <= 0) return;
    if (maxrows <= kFoldSplitSmall) {
#define L_(s) hipLaunchKernelGGL((k_fold_split_jobs<s, kFoldSplitSmall / 1024>), dim3(ncols), dim3(1024), 0, st, d_jobs, d_begin, njobs, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
        return;
    }
    if (maxrows <= 16384) {
#define L_(s) hipLaunchKernelGGL((k_fold_split_jobs<s, 16>), dim3(ncols), dim3(1024), 0, st, d_jobs, d_begin, njobs, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
        return;
    }
#define L_(s) hipLaunchKernelGGL((k_fold_split_jobs<s, kFoldSplitMax / 1024>), dim3(ncols), dim3(1024), 0, st, d_jobs, d_begin, njobs, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_split_mu(hipStream_t st, int S, double R, int N, int ld, int nrows, int kpad_all,
                     const double *mu, const int *vinv, int8_t *D, long long plane, int *E)
{
    if (nrows <= 0) return;
#define L_(s) hipLaunchKernelGGL(k_split_mu<s>, dim3(nrows), dim3(256), 0, st, N, ld, kpad_all, mu, vinv, D, plane, E, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_build_g_split(hipStream_t st, int S, double R, const GSplitJob *d_jobs, const int *d_begin,
                          int njobs, int ntiles, bool four)
{
    if (njobs <= 0 || ntiles <= 0) return;
    if (four) {
#define L_(s) hipLaunchKernelGGL(k_build_g_split4<s>, dim3(xcd_grid_s(ntiles)), dim3(256), 0, st, d_jobs, d_begin, njobs, ntiles, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
        return;
    }
#define L_(s) hipLaunchKernelGGL(k_build_g_split<s>, dim3(xcd_grid_s(ntiles)), dim3(256), 0, st, d_jobs, d_begin, njobs, ntiles, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_build_g_split_lds(hipStream_t st, int S, double R, const GSplitJob *d_jobs,
                              const int *d_begin, int njobs, int nunits, bool part)
{
    if (njobs <= 0 || nunits <= 0) return;
    if (part) {
#define L_(s) hipLaunchKernelGGL((k_build_g_split_lds<s, true>), dim3(nunits), dim3(1024), 0, st, d_jobs, d_begin, njobs, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
    } else {
#define L_(s) hipLaunchKernelGGL((k_build_g_split_lds<s, false>), dim3(nunits), dim3(1024), 0, st, d_jobs, d_begin, njobs, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
    }
}

void launch_halfls_split(hipStream_t st, int S, double R, int kpad_all, const MuSrc &src,
                         const int *mue, const SplitHalfJob *d_jobs, const int *d_tbegin,
                         int njobs, int ntiles, int cfg)
{
    if (njobs <= 0 || ntiles <= 0) return;
    if ((cfg == 2 || cfg == 3) && S <= 5) {  // two register sets exceed the registers at 6 digits
        const int mw = cfg == 3 ? 4 : 2;
#define L_(s, w) hipLaunchKernelGGL((k_halfls_split2<s, w>), dim3(xcd_grid_s(ntiles)), dim3(128 * w), 0, st, kpad_all, src, mue, d_jobs, d_tbegin, njobs, ntiles, R)
        switch (S * 8 + mw) {
        case 3 * 8 + 2: L_(3, 2); break;
        case 4 * 8 + 2: L_(4, 2); break;
        case 5 * 8 + 2: L_(5, 2); break;
        case 3 * 8 + 4: L_(3, 4); break;
        case 4 * 8 + 4: L_(4, 4); break;
        default: L_(5, 4); break;
        }
#undef L_
        return;
    }
    if (cfg == 1 || cfg == 3) {  // 128 rows; setting 3 at 6 digits
#define L_(s) hipLaunchKernelGGL(k_halfls_split_wide<s>, dim3(xcd_grid_s(ntiles)), dim3(256), 0, st, kpad_all, src, mue, d_jobs, d_tbegin, njobs, ntiles, R)
        GI_SPLIT_DISPATCH(S, L_)
#undef L_
        return;
    }
#define L_(s) hipLaunchKernelGGL(k_halfls_split<s>, dim3(xcd_grid_s(ntiles)), dim3(256), 0, st, kpad_all, src, mue, d_jobs, d_tbegin, njobs, ntiles, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

void launch_corr_split(hipStream_t st, int S, double R, const CorrSplitArgs *d_args, const int *d_cbegin,
                       int nchunks, int ntiles)
{
    if (nchunks <= 0 || ntiles <= 0) return;
#define L_(s) hipLaunchKernelGGL(k_corr_split<s>, dim3(xcd_grid_s(ntiles)), dim3(256), 0, st, d_args, d_cbegin, nchunks, ntiles, R)
    GI_SPLIT_DISPATCH(S, L_)
#undef L_
}

}  // namespace gi
