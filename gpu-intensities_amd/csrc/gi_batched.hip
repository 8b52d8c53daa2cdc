// Batched FP64 form of the half linestrength and the correlation on the
// gfx950 matrix cores (v_mfma_f64_16x16x4f64).
//
// K5 (k_halfls_jobs): for one output class co and a chunk of units,
//   H_co[m, (u, j)] = sum_k A[m, k] * B[k, (u, j)],
//   A[m, k] = mu_{a(e)}[off_co + m, off_e + k - row0_e]   (column-major, ld N)
//   B[k, (u, j)] = G[u*Gtot + gbase + k*smax + j]
// so mu blocks are read once per chunk instead of once per lower level.
//
// K6 (k_corr_mfma): for one chunk of items (lower levels of one
// (J, symmetry, J') group, energy sorted), the banded product
//   A[f, n] = sum_t U[col(f)*nsize + t] * Ht[n*nsize + t],
// n = item*ncb + component pair, f in the union of the items' windows; tiles
// outside every window of their items are skipped, entries outside an item's
// own window are not stored.
//
// Both kernels: 256 threads (4 wavefronts) per 64 x 64 output tile, K staged
// through LDS in steps of 16, each wavefront owns a 32 x 32 quarter as 2 x 2
// MFMA 16x16 blocks. Operand layout of the 16x16x4 f64 instruction: lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator register r holds
// D[4*r + l/16][l%16], r = 0..3 (unlike the f32 16x16x4 form, whose register r
// holds row 4*(l/16) + r).
#include <hip/hip_runtime.h>

#include "gi_kernels.h"

namespace gi {

typedef double d4 __attribute__((ext_vector_type(4)));

#define BM 64
#define BN 64
#define BK 16
// LDS row of 80 doubles: the four K rows one MFMA operand read touches fall
// 32 banks apart, so the two half-waves of a 64-bit LDS read are conflict-free.
#define LDP 80

__device__ inline void mfma_tile(const double (*As)[LDP], const double (*Bs)[LDP], int wm, int wn,
                                 int lane, d4 acc[2][2])
{
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        const int k = kk + (lane >> 4);
        const int i = lane & 15;
        double a0 = As[k][wm + i], a1 = As[k][wm + 16 + i];
        double b0 = Bs[k][wn + i], b1 = Bs[k][wn + 16 + i];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// ---------------------------------------------------------------------------
// Flattened launches over many (chunk, output class) jobs: a workgroup finds
// its job by binary search over the jobs' tile prefixes.
__device__ inline int find_job(const int *__restrict__ begin, int njobs, int wg)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= wg) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// K1 over jobs: G rows of every (chunk, class) job (same arithmetic as
// k_build_g_sym). Workgroup = 64 rows of one unit.
__global__ __launch_bounds__(64) void k_build_g_jobs(int N, const HalfJob *__restrict__ jobs,
                                                     const int *__restrict__ gbegin, int njobs)
{
    const int wg = blockIdx.x;
    const HalfJob &J = jobs[find_job(gbegin, njobs, wg)];
    const int local = wg - J.gtile_begin;
    const int unit = local / J.gkt, kt = local - unit * J.gkt;
    const ClassDev &cd = J.cd;
    const int row = kt * 64 + threadIdx.x;
    if (row >= cd.K) return;
    int e = 0;
    while (e + 1 < cd.nseg && cd.seg_row0[e + 1] <= row) e++;
    const int vI = cd.seg_off[e] + (row - cd.seg_row0[e]);
    const double *u = J.uptr[unit];
    const SymDev R = J.uideg[unit] ? J.r1 : J.r0;
    double *Gr = J.G + (size_t)unit * J.Gtot + cd.gbase + (size_t)row * cd.smax;
    for (int j = 0; j < cd.smax; j++) {
        double v = 0.0;
        if (j < cd.ncols) {
            const int q = e * cd.ncols + j;
            for (int p = cd.rptr[q]; p < cd.rptr[q + 1]; p++) {
                const int cell = cd.rsI[p] * N + vI;
                double cv = 0.0;
                for (int x = R.rptr[cell]; x < R.rptr[cell + 1]; x++) cv += u[R.rcol[x]] * R.rval[x];
                v += cd.rcoef[p] * cv;
            }
        }
        Gr[j] = v;
    }
}

// K5 over jobs: H_co[m, (u, j)] = sum_k mu-block[m, k] * G[k, (u, j)] on the
// matrix cores. K runs segment by segment (a tile never straddles two mu
// blocks); the next K tile is fetched into registers while the current one
// is multiplied.
__global__ __launch_bounds__(256) void k_halfls_jobs(int N, const HalfJob *__restrict__ jobs,
                                                     const int *__restrict__ tbegin, int njobs,
                                                     const double *__restrict__ mu)
{
    __shared__ double As[BK][LDP];
    __shared__ double Bs[BK][LDP];
    const int wg = blockIdx.x;
    const HalfJob &J = jobs[find_job(tbegin, njobs, wg)];
    const ClassDev &cd = J.cd;
    const int local = wg - J.tile_begin;
    const int m0 = (local % J.mt) * BM, n0 = (local / J.mt) * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int ncolsTot = J.nunits * cd.smax;
    // this thread's load coordinates: A row m, B column n, K rows kl0 + 4q
    const int lm = tid & 63, kl0 = tid >> 6;
    const bool mok = m0 + lm < cd.n;
    const int n = n0 + lm;
    const bool nok = n < ncolsTot;
    const double *gcol = J.G;
    if (nok) {
        const int u = n / cd.smax, jj = n - u * cd.smax;
        gcol = J.G + (size_t)u * J.Gtot + cd.gbase + jj;
    }
    d4 acc[2][2];
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[4];
    int e = 0, kk0 = 0;
    auto fetch = [&](int se, int sk0) {
        const int len = min(BK, cd.seg_n[se] - sk0);
        const double *mp = mu + (size_t)cd.seg_a[se] * N * N +
                           (size_t)(cd.seg_off[se] + sk0) * N + (cd.off + m0 + lm);
        const double *gp = gcol + (size_t)(cd.seg_row0[se] + sk0) * cd.smax;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kl = kl0 + 4 * q;
            ra[q] = (mok && kl < len) ? mp[(size_t)kl * N] : 0.0;
            rb[q] = (nok && kl < len) ? gp[(size_t)kl * cd.smax] : 0.0;
        }
    };
    if (cd.nseg > 0) fetch(0, 0);
    while (e < cd.nseg) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            As[kl0 + 4 * q][lm] = ra[q];
            Bs[kl0 + 4 * q][lm] = rb[q];
        }
        __syncthreads();
        kk0 += BK;
        if (kk0 >= cd.seg_n[e]) {
            e++;
            kk0 = 0;
        }
        if (e < cd.nseg) fetch(e, kk0);
        mfma_tile(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * x + 4 * r + (lane >> 4);
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (m >= cd.n || nn >= ncolsTot) continue;
                const int u = nn / cd.smax, jj = nn - u * cd.smax;
                if (jj >= cd.ncols) continue;
                J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] = acc[x][y][r];
            }
}

// ---------------------------------------------------------------------------
// K6
__global__ __launch_bounds__(256) void k_corr_mfma(const CorrMMArgs *__restrict__ args)
{
    __shared__ double As[BK][LDP];
    __shared__ double Bs[BK][LDP];
    __shared__ int s_need;
    const CorrMMArgs a = args[blockIdx.z];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int f0 = a.Lo + blockIdx.x * BM;  // block-level index of the first row
    const int n0 = blockIdx.y * BN;
    const int ncols = a.nitems * a.ncb;
    if (f0 >= a.Hi || n0 >= ncols) return;  // this chunk has fewer tiles
    // skip tiles outside every window of the tile's items
    if (tid == 0) {
        int need = 0;
        const int i0 = n0 / a.ncb, i1 = min(a.nitems, (n0 + BN + a.ncb - 1) / a.ncb);
        for (int i = i0; i < i1 && !need; i++)
            if (a.item_lo[i] < f0 + BM && a.item_hi[i] > f0) need = 1;
        s_need = need;
    }
    __syncthreads();
    if (!s_need) return;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    d4 acc[2][2];
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    // per-thread row pointers for the loads: row = tid / 4, 4 consecutive t
    const int lr = tid >> 2, lt = (tid & 3) * 4;
    const int fr = f0 + lr;
    const double *Urow = nullptr;
    if (fr < a.Hi) Urow = a.vec + (size_t)a.blk_col[fr] * a.nsize;
    const int nr = n0 + lr;
    const double *Hrow = nr < ncols ? a.Ht + (size_t)nr * a.nsize : nullptr;
    for (int k0 = 0; k0 < a.nsize; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int t = k0 + lt + q;
            As[lt + q][lr] = (Urow && t < a.nsize) ? Urow[t] : 0.0;
            Bs[lt + q][lr] = (Hrow && t < a.nsize) ? Hrow[t] : 0.0;
        }
        __syncthreads();
        mfma_tile(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int r = 0; r < 4; r++) {
                const int f = f0 + wm + 16 * x + 4 * r + (lane >> 4);
                const int n = n0 + wn + 16 * y + (lane & 15);
                if (f >= a.Hi || n >= ncols) continue;
                const int i = n / a.ncb, cb = n - i * a.ncb;
                if (f < a.item_lo[i] || f >= a.item_hi[i]) continue;
                const long long o = a.item_out_off[i] + (f - a.item_lo[i]);
                const int idegF = cb / a.ndegI, idegI = cb - idegF * a.ndegI;
                a.amp[4 * o + idegI + 2 * idegF] = acc[x][y][r];
            }
}

// ---------------------------------------------------------------------------
// Degeneracy sum, Einstein A, intensity and emit test per candidate, from the
// amplitudes K6 stored (cuda_host.cu:1388-1407).
__global__ __launch_bounds__(256) void k_intensity(IntensArgs a)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.ncand) return;
    int lo = 0, hi = a.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.item_cand_start[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const int item = lo;
    const int j = c - a.item_cand_start[item];
    const int f = a.item_lo[item] + j;
    const long long o = a.item_out_off[item] + j;
    double *am = a.amp + 4 * o;
    double ls = 0.0;
    for (int idegF = 0; idegF < a.ndegF; idegF++)
        for (int idegI = 0; idegI < a.ndegI; idegI++) {
            const int cb = idegF * a.ndegI + idegI;
            double A = 0.0;
            if ((a.cb_mask >> cb) & 1) A = am[idegI + 2 * idegF];
            am[idegI + 2 * idegF] = A;
            ls += A * A;
        }
    for (int q = 0; q < 4; q++) {
        const int idegI = q & 1, idegF = q >> 1;
        if (idegI >= a.ndegI || idegF >= a.ndegF) am[q] = 0.0;
    }
    ls /= (double)a.ndegI;
    if (a.reduced && a.ndegF != 1 && a.ndegI != 1) ls *= (double)a.ndegI;
    const double eI = a.item_energy[item];
    const double eF = a.blk_en[f];
    const double nu = eF - eI;
    const double anu = fabs(nu);
    const double boltz_fc =
        anu * exp(-(eI - a.zpe) * a.beta) * (1.0 - exp(-anu * a.beta)) / a.q_stat;
    const double A_einst = a.A_coef_s_1 * (double)((2 * a.jI) + 1) * ls * pow(anu, 3.0);
    const double S = ls * a.gnsI * (double)((2 * a.jI + 1) * (2 * a.jF + 1));
    const double I = S * a.intens_cm_mol * boltz_fc;
    const bool pass = a.blk_upok[f] && S >= a.thresh_S && I >= a.thresh_I;
    a.flag[o] = pass ? 1 : 0;
    a.out_iI[o] = a.item_level[item];
    a.out_iF[o] = a.blk_lev[f];
    a.out_nu[o] = nu;
    a.out_S[o] = S;
    a.out_A[o] = A_einst;
    a.out_I[o] = I;
}

// ---------------------------------------------------------------------------

void launch_halfls_jobs(hipStream_t st, int N, const HalfJob *d_jobs, const int *d_gbegin,
                        const int *d_tbegin, int njobs, int gtiles, int ttiles, const double *mu)
{
    if (njobs <= 0) return;
    if (gtiles > 0)
        hipLaunchKernelGGL(k_build_g_jobs, dim3(gtiles), dim3(64), 0, st, N, d_jobs, d_gbegin,
                           njobs);
    if (ttiles > 0)
        hipLaunchKernelGGL(k_halfls_jobs, dim3(ttiles), dim3(256), 0, st, N, d_jobs, d_tbegin,
                           njobs, mu);
}

void launch_corr_mfma(hipStream_t st, const CorrMMArgs *d_args, int nchunks, int max_rows,
                      int max_cols)
{
    if (nchunks <= 0 || max_rows <= 0 || max_cols <= 0) return;
    dim3 grid((max_rows + BM - 1) / BM, (max_cols + BN - 1) / BN, nchunks);
    hipLaunchKernelGGL(k_corr_mfma, grid, dim3(256), 0, st, d_args);
}

void launch_intensity(hipStream_t st, const IntensArgs &a)
{
    if (a.ncand <= 0) return;
    hipLaunchKernelGGL(k_intensity, dim3((a.ncand + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace gi
