// Batched FP64 form of the half linestrength and the correlation on the
// gfx950 matrix cores (v_mfma_f64_16x16x4f64).
//
// K5 (k_halfls_mfma): for one output class co and a chunk of units,
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
#define LDP (BM + 2)  // padded LDS row (doubles)

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
// K5
__global__ __launch_bounds__(256) void k_halfls_mfma(int N, int nunits, int Gtot, int Htot,
                                                     ClassDev cd, const double *__restrict__ mu,
                                                     const double *__restrict__ G,
                                                     double *__restrict__ H)
{
    __shared__ double As[BK][LDP];
    __shared__ double Bs[BK][LDP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int ncolsTot = nunits * cd.smax;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    d4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < cd.K; k0 += BK) {
        // A tile: 64 m x 16 k; thread -> m = tid % 64, k = tid / 64 + 4q
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = tid & 63, kl = (tid >> 6) + 4 * q;
            const int k = k0 + kl;
            double v = 0.0;
            if (m0 + m < cd.n && k < cd.K) {
                int e = 0;
                while (e + 1 < cd.nseg && cd.seg_row0[e + 1] <= k) e++;
                const int kk = cd.seg_off[e] + (k - cd.seg_row0[e]);
                v = mu[(size_t)cd.seg_a[e] * N * N + (size_t)kk * N + (cd.off + m0 + m)];
            }
            As[kl][m] = v;
        }
        // B tile: 16 k x 64 n; n = u*smax + j
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            const int nl = e & 63, kl = e >> 6;
            const int n = n0 + nl, k = k0 + kl;
            double v = 0.0;
            if (n < ncolsTot && k < cd.K) {
                const int u = n / cd.smax, j = n - u * cd.smax;
                v = G[(size_t)u * Gtot + cd.gbase + (size_t)k * cd.smax + j];
            }
            Bs[kl][nl] = v;
        }
        __syncthreads();
        mfma_tile(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * a + 4 * r + (lane >> 4);
                const int n = n0 + wn + 16 * b + (lane & 15);
                if (m >= cd.n || n >= ncolsTot) continue;
                const int u = n / cd.smax, j = n - u * cd.smax;
                if (j >= cd.ncols) continue;
                H[(size_t)u * Htot + cd.hbase + (size_t)j * cd.n + m] = acc[a][b][r];
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

void launch_halfls_mfma(hipStream_t st, int nunits, int N, int Gtot, int Htot, const ClassDev &cd,
                        const double *mu, const double *G, double *H)
{
    dim3 grid((cd.n + BM - 1) / BM, (nunits * cd.smax + BN - 1) / BN);
    hipLaunchKernelGGL(k_halfls_mfma, grid, dim3(256), 0, st, N, nunits, Gtot, Htot, cd, mu, G,
                       H);
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
