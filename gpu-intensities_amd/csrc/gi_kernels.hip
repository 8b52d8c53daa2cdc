// gfx950 kernels of the line-intensity engine (FP64 path).
//
// Data layout in HBM (per context):
//   Mu  [3N x N] column-major as an N x 3N matrix: Mu[vF + (a*N + vI)*N]
//       = mu_a(vF, vI), i.e. the flipped dipole of read_dipole_flipped
//       (trove_functions.cpp:193-196) with the three components stacked.
//   slot grid of J: index g = slot*N + v; a slot is one rotational (K, tau).
//   G   per half-ls unit: [3N rows][SMAX] row-major, the rotational
//       coefficients applied to the correlated lower vector.
//   H   per unit: [SF*N] in grid order: the primitive half linestrength.
//   Ht  per unit and component pair: [nsize_F] in symmetrised order.
//   U   eigenvector block: record irec at irec*nsize (the file layout of
//       j0eigen_vectors, cuda_host.cu:1196).
#include <hip/hip_runtime.h>

#include "gi_kernels.h"

namespace gi {

// ---------------------------------------------------------------------------
// Correlate a symmetrised vector into primitive order (the per-r form of
// dipole_kernals.cu:38-71): out[r] = sum_e rval[e] * u[rcol[e]] over the CSR
// row of r's grid cell.
__global__ void k_correlate_prim(int dimen, const int *__restrict__ r_grid,
                                 const int *__restrict__ rptr, const int *__restrict__ rcol,
                                 const double *__restrict__ rval, const double *__restrict__ u,
                                 double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dimen) return;
    const int row = r_grid[r];
    double acc = 0.0;
    for (int e = rptr[row]; e < rptr[row + 1]; e++) acc += u[rcol[e]] * rval[e];
    out[r] = acc;
}

// Gather grid-ordered values into primitive order.
__global__ void k_grid_to_prim(int dimen, const int *__restrict__ r_grid,
                               const double *__restrict__ grid, double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < dimen) out[r] = grid[r_grid[r]];
}

// ---------------------------------------------------------------------------
// K1: correlated lower vector c = R u on the grid, then the rotational
// coefficients: G[a*N + vI][sF] = sum_{(sI, w) in list(sF, a)} w * c[sI*N + vI].
// One 64-lane wave per (unit, 64 vibrational indices); c lives in LDS.
__global__ __launch_bounds__(64) void k_build_g(int N, int SI, int SF, int SMAX,
                                                const double *const *__restrict__ uptr,
                                                const int *__restrict__ unit_ideg,
                                                SymDev r0, SymDev r1,
                                                const int *__restrict__ cptr,
                                                const int *__restrict__ csI,
                                                const double *__restrict__ cval,
                                                double *__restrict__ G)
{
    extern __shared__ double cl[];  // [SI][64]
    const int unit = blockIdx.y;
    const int lane = threadIdx.x;
    const int vI = blockIdx.x * 64 + lane;
    if (vI >= N) return;
    const double *u = uptr[unit];
    const SymDev R = unit_ideg[unit] ? r1 : r0;
    for (int s = 0; s < SI; s++) {
        const int row = s * N + vI;
        double acc = 0.0;
        for (int e = R.rptr[row]; e < R.rptr[row + 1]; e++) acc += u[R.rcol[e]] * R.rval[e];
        cl[s * 64 + lane] = acc;
    }
    double *Gu = G + (size_t)unit * 3 * N * SMAX;
    for (int a = 0; a < 3; a++) {
        double *row = Gu + (size_t)(a * N + vI) * SMAX;
        for (int sF = 0; sF < SMAX; sF++) {
            double v = 0.0;
            if (sF < SF) {
                const int p0 = cptr[sF * 3 + a], p1 = cptr[sF * 3 + a + 1];
                for (int p = p0; p < p1; p++) v += cval[p] * cl[csI[p] * 64 + lane];
            }
            row[sF] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// K2: half linestrength in GEMV form, H = Mu * G per unit. One thread per
// vF row and all SF columns in registers; G rows are wave-uniform (scalar
// loads). Units run fastest in the grid so that concurrent workgroups read
// the same Mu rows (cache reuse across lower levels).
template <int SMAX>
__global__ __launch_bounds__(256) void k_half_ls_gemv(int N, int SF, const double *__restrict__ Mu,
                                                      const double *__restrict__ G,
                                                      double *__restrict__ H)
{
    const int unit = blockIdx.x;
    const int vF = blockIdx.y * 256 + threadIdx.x;
    const int vr = vF < N ? vF : N - 1;
    const double *__restrict__ Gu = G + (size_t)unit * 3 * N * SMAX;
    const double *__restrict__ mp = Mu + vr;
    double acc[SMAX];
#pragma unroll
    for (int s = 0; s < SMAX; s++) acc[s] = 0.0;
    const int K = 3 * N;
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        double m[4];
#pragma unroll
        for (int q = 0; q < 4; q++) m[q] = mp[(size_t)(k + q) * N];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double *gk = Gu + (size_t)(k + q) * SMAX;
#pragma unroll
            for (int s = 0; s < SMAX; s++) acc[s] = __builtin_fma(m[q], gk[s], acc[s]);
        }
    }
    for (; k < K; k++) {
        const double m = mp[(size_t)k * N];
        const double *gk = Gu + (size_t)k * SMAX;
#pragma unroll
        for (int s = 0; s < SMAX; s++) acc[s] = __builtin_fma(m, gk[s], acc[s]);
    }
    if (vF < N) {
        double *Hu = H + (size_t)unit * SF * N;
#pragma unroll
        for (int s = 0; s < SMAX; s++)
            if (s < SF) Hu[(size_t)s * N + vF] = acc[s];
    }
}

// ---------------------------------------------------------------------------
// mu with rows and columns permuted so that vibrational symmetry classes are
// contiguous: out[a][i + j*N] = in[a][perm[i] + perm[j]*N].
__global__ __launch_bounds__(256) void k_permute_mu(int N, const int *__restrict__ perm,
                                                    const double *__restrict__ in,
                                                    double *__restrict__ out)
{
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long NN = (long long)N * N;
    if (idx >= 3 * NN) return;
    const int a = (int)(idx / NN);
    const long long rem = idx - a * NN;
    const int j = (int)(rem / N), i = (int)(rem - (long long)j * N);
    out[idx] = in[a * NN + perm[i] + (long long)perm[j] * N];
}

// ---------------------------------------------------------------------------
// K1 (symmetry-blocked): G rows of one output class. Row (segment e, vI)
// holds, for each upper slot column j, sum_{(sI, w) in recipe(e, j)} w * c[sI, vI]
// with c = R u evaluated on the fly from the CSR of R (one or a few terms).
__global__ __launch_bounds__(64) void k_build_g_sym(int N, int Gtot, ClassDev cd,
                                                    const double *const *__restrict__ uptr,
                                                    const int *__restrict__ unit_ideg,
                                                    SymDev r0, SymDev r1, double *__restrict__ G)
{
    const int unit = blockIdx.y;
    const int row = blockIdx.x * 64 + threadIdx.x;
    if (row >= cd.K) return;
    int e = 0;
    while (e + 1 < cd.nseg && cd.seg_row0[e + 1] <= row) e++;
    const int vI = cd.seg_off[e] + (row - cd.seg_row0[e]);
    const double *u = uptr[unit];
    const SymDev R = unit_ideg[unit] ? r1 : r0;
    double *Gr = G + (size_t)unit * Gtot + cd.gbase + (size_t)row * cd.smax;
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

// K2 (symmetry-blocked, GEMV form): H_co[vF][j] = sum over segments e of
// sum_k mu_{a(e)}[off_co + vF, off_e + k] * G_co[row0_e + k][j], one thread per
// vF row, the class's columns in registers, G rows wave-uniform. Rows
// [rlo, rhi) of the class; mu element (a, vI, vF) at a*N*ld + vI*ld + vF - vbase
// (resident mu: ld = N, vbase = 0; a streamed part: its vF rows from vbase).
template <int SMAX>
__global__ __launch_bounds__(256) void k_gemv_sym(int N, int Gtot, int Htot, ClassDev cd,
                                                  const double *__restrict__ mu, int ld,
                                                  int vbase, int rlo, int rhi,
                                                  const double *__restrict__ G,
                                                  double *__restrict__ H)
{
    const int unit = blockIdx.x;
    const int row = rlo + blockIdx.y * 256 + threadIdx.x;
    const int rr = row < rhi ? row : rhi - 1;
    const double *__restrict__ Gu = G + (size_t)unit * Gtot + cd.gbase;
    double acc[SMAX];
#pragma unroll
    for (int s = 0; s < SMAX; s++) acc[s] = 0.0;
    for (int e = 0; e < cd.nseg; e++) {
        const int a = cd.seg_a[e], off = cd.seg_off[e], cnt = cd.seg_n[e], r0 = cd.seg_row0[e];
        const double *__restrict__ mp =
            mu + (size_t)a * N * ld + (size_t)off * ld + (cd.off + rr - vbase);
        const double *__restrict__ gp = Gu + (size_t)r0 * SMAX;
        int k = 0;
        for (; k + 4 <= cnt; k += 4) {
            double m[4];
#pragma unroll
            for (int q = 0; q < 4; q++) m[q] = mp[(size_t)(k + q) * ld];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double *gk = gp + (size_t)(k + q) * SMAX;
#pragma unroll
                for (int s = 0; s < SMAX; s++) acc[s] = __builtin_fma(m[q], gk[s], acc[s]);
            }
        }
        for (; k < cnt; k++) {
            const double m = mp[(size_t)k * ld];
            const double *gk = gp + (size_t)k * SMAX;
#pragma unroll
            for (int s = 0; s < SMAX; s++) acc[s] = __builtin_fma(m, gk[s], acc[s]);
        }
    }
    if (row < rhi) {
        double *Hc = H + (size_t)unit * Htot + cd.hbase;
#pragma unroll
        for (int s = 0; s < SMAX; s++)
            if (s < cd.ncols) Hc[(size_t)s * cd.n + row] = acc[s];
    }
}

// ---------------------------------------------------------------------------
// K3: fold H into the symmetrised space of the upper block:
// Ht[item][cb][t] = sum_{(g, w) in R_F^T[t]} w * H[unit][g], cb = idegF*ndegI + idegI.
__device__ inline void fold_one(int t, int item, int cb, int nsizeF, int ndegI, int ndegF,
                                int SFN, const int *__restrict__ item_unit, const SymDev &t0,
                                const SymDev &t1, const double *__restrict__ H,
                                double *__restrict__ Ht)
{
    if (t >= nsizeF) return;
    const int idegF = cb / ndegI, idegI = cb % ndegI;
    const int unit = item_unit[item * GI_MAXDEG + idegI];
    double acc = 0.0;
    if (unit >= 0) {
        const SymDev T = idegF ? t1 : t0;
        const double *Hu = H + (size_t)unit * SFN;
        for (int e = T.tptr[t]; e < T.tptr[t + 1]; e++) acc += T.tval[e] * Hu[T.tcol[e]];
    }
    Ht[((size_t)item * ndegI * ndegF + cb) * nsizeF + t] = acc;
}

__global__ __launch_bounds__(256) void k_fold(int nsizeF, int ndegI, int ndegF, int SFN,
                                              const int *__restrict__ item_unit,
                                              SymDev t0, SymDev t1,
                                              const double *__restrict__ H,
                                              double *__restrict__ Ht)
{
    fold_one(blockIdx.x * 256 + threadIdx.x, blockIdx.y, blockIdx.z, nsizeF, ndegI, ndegF, SFN,
             item_unit, t0, t1, H, Ht);
}

// Job owning block b: the last job whose first block is <= b.
__device__ inline int job_of(const int *__restrict__ begin, int njobs, int b)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_fold_jobs(const FoldJob *__restrict__ jobs,
                                                   const int *__restrict__ begin, int njobs)
{
    const int j = job_of(begin, njobs, blockIdx.x);
    fold_block(jobs[j], blockIdx.x - begin[j]);
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// K4: correlation A = U_F[:, col]^T Ht for every candidate upper level of
// every item (one wave per candidate), then the degeneracy sum, Einstein A
// and intensity (cuda_host.cu:1388-1407) and the emit test.
__global__ __launch_bounds__(256) void k_corr_intensity(CorrArgs a)
{
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= a.ncand) return;
    // item owning this candidate: last item with cand_start <= wave
    int lo = 0, hi = a.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.item_cand_start[mid] <= wave) lo = mid; else hi = mid - 1;
    }
    const int item = lo;
    const int j = wave - a.item_cand_start[item];
    const int f = a.item_lo[item] + j;
    const int col = a.blk_col[f];
    const double *__restrict__ U = a.vec + (size_t)col * a.nsizeF;
    const int ncb = a.ndegI * a.ndegF;
    const double *__restrict__ Hi = a.Ht + (size_t)item * ncb * a.nsizeF;
    double acc[GI_MAXDEG * GI_MAXDEG] = {0.0, 0.0, 0.0, 0.0};
    if (ncb == 1) {
        for (int t = lane; t < a.nsizeF; t += 64) acc[0] += U[t] * Hi[t];
    } else {
        for (int t = lane; t < a.nsizeF; t += 64) {
            const double uv = U[t];
            for (int cb = 0; cb < ncb; cb++) acc[cb] += uv * Hi[(size_t)cb * a.nsizeF + t];
        }
    }
    for (int cb = 0; cb < ncb; cb++) acc[cb] = wave_sum(acc[cb]);
    if (lane != 0) return;

    double amp[4] = {0.0, 0.0, 0.0, 0.0};
    double ls = 0.0;
    for (int idegF = 0; idegF < a.ndegF; idegF++)
        for (int idegI = 0; idegI < a.ndegI; idegI++) {
            const int cb = idegF * a.ndegI + idegI;
            const double A = ((a.cb_mask >> cb) & 1) ? acc[cb] : 0.0;
            amp[idegI + idegF * 2] = A;
            ls += A * A;
        }
    ls /= (double)a.ndegI;
    if (a.reduced && a.ndegF != 1 && a.ndegI != 1) ls *= (double)a.ndegI;
    const double eI = a.item_energy[item];
    const double eF = a.blk_en[f];
    const double nu = eF - eI;
    const double anu = fabs(nu);
    double beta = a.beta, q_stat = a.q_stat, thS = a.thresh_S, thI = a.thresh_I;
    if (a.qpar) {  // a request's parameters, Q(T) reduced on the device
        beta = a.qpar[0];
        q_stat = a.qpar[1];
        thS = a.qpar[2];
        thI = a.qpar[3];
    }
    const double boltz_fc =
        anu * exp(-(eI - a.zpe) * beta) * (1.0 - exp(-anu * beta)) / q_stat;
    const double A_einst = a.A_coef_s_1 * (double)((2 * a.jI) + 1) * ls * pow(anu, 3.0);
    const double S = ls * a.gnsI * (double)((2 * a.jI + 1) * (2 * a.jF + 1));
    const double I = S * a.intens_cm_mol * boltz_fc;
    const bool pass = a.blk_upok[f] && S >= thS && I >= thI;
    const long long o = a.item_out_off[item] + j;
    a.flag[o] = pass ? 1 : 0;
    a.out_iI[o] = a.item_level[item];
    a.out_iF[o] = a.blk_lev[f];
    a.out_nu[o] = nu;
    a.out_S[o] = S;
    a.out_A[o] = A_einst;
    a.out_I[o] = I;
    const int m = a.amp_m;
    double *am = a.out_amp + (long long)(m * m) * o;
    for (int idegF = 0; idegF < m; idegF++)
        for (int idegI = 0; idegI < m; idegI++) am[idegI + m * idegF] = amp[idegI + 2 * idegF];
}

// ---------------------------------------------------------------------------
// Stable compaction of the candidate records (flags -> output order).
#define GI_CT 256
#define GI_CPT 8
#define GI_CPB (GI_CT * GI_CPT)

__global__ __launch_bounds__(GI_CT) void k_block_count(long long n, const int *__restrict__ flag,
                                                       int *__restrict__ counts)
{
    __shared__ int red[GI_CT];
    const long long base = (long long)blockIdx.x * GI_CPB + (long long)threadIdx.x * GI_CPT;
    int c = 0;
    for (int i = 0; i < GI_CPT; i++)
        if (base + i < n) c += flag[base + i];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = GI_CT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// exclusive scan of nb block counts by one workgroup; total at offs[nb]
__global__ __launch_bounds__(GI_CT) void k_scan_counts(int nb, const int *__restrict__ counts,
                                                       long long *__restrict__ offs)
{
    __shared__ long long part[GI_CT];
    const int per = (nb + GI_CT - 1) / GI_CT;
    const int b0 = threadIdx.x * per;
    long long s = 0;
    for (int i = 0; i < per; i++)
        if (b0 + i < nb) s += counts[b0 + i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long run = 0;
        for (int i = 0; i < GI_CT; i++) {
            const long long v = part[i];
            part[i] = run;
            run += v;
        }
        offs[nb] = run;
    }
    __syncthreads();
    long long run = part[threadIdx.x];
    for (int i = 0; i < per; i++)
        if (b0 + i < nb) {
            offs[b0 + i] = run;
            run += counts[b0 + i];
        }
}

__global__ __launch_bounds__(GI_CT) void k_scatter(long long n, const int *__restrict__ flag,
                                                   const long long *__restrict__ offs,
                                                   LineSoA in, LineSoA out)
{
    __shared__ int pre[GI_CT];
    if (offs[gridDim.x] == n) return;  // all pass: the input is its own compaction
    const long long base = (long long)blockIdx.x * GI_CPB + (long long)threadIdx.x * GI_CPT;
    int c = 0;
    for (int i = 0; i < GI_CPT; i++)
        if (base + i < n) c += flag[base + i];
    pre[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < GI_CT; i++) {
            const int v = pre[i];
            pre[i] = run;
            run += v;
        }
    }
    __syncthreads();
    long long o = offs[blockIdx.x] + pre[threadIdx.x];
    for (int i = 0; i < GI_CPT; i++) {
        const long long s = base + i;
        if (s < n && flag[s]) {
            out.iI[o] = in.iI[s];
            out.iF[o] = in.iF[s];
            out.nu[o] = in.nu[s];
            out.S[o] = in.S[s];
            out.A[o] = in.A[s];
            out.I[o] = in.I[s];
            for (int q = 0; q < in.amp_w; q++)
                out.amp[(long long)in.amp_w * o + q] = in.amp[(long long)in.amp_w * s + q];
            o++;
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers

void launch_correlate_prim(hipStream_t st, int dimen, const int *r_grid, SymDev R,
                           const double *u, double *out)
{
    hipLaunchKernelGGL(k_correlate_prim, dim3((dimen + 255) / 256), dim3(256), 0, st, dimen,
                       r_grid, R.rptr, R.rcol, R.rval, u, out);
}

void launch_grid_to_prim(hipStream_t st, int dimen, const int *r_grid, const double *grid,
                         double *out)
{
    hipLaunchKernelGGL(k_grid_to_prim, dim3((dimen + 255) / 256), dim3(256), 0, st, dimen,
                       r_grid, grid, out);
}

void launch_build_g(hipStream_t st, int nunits, int N, int SI, int SF, int SMAX,
                    const double *const *uptr, const int *unit_ideg, SymDev r0, SymDev r1,
                    const int *cptr, const int *csI, const double *cval, double *G)
{
    dim3 grid((N + 63) / 64, nunits);
    hipLaunchKernelGGL(k_build_g, grid, dim3(64), sizeof(double) * SI * 64, st, N, SI, SF, SMAX,
                       uptr, unit_ideg, r0, r1, cptr, csI, cval, G);
}

int launch_half_ls_gemv(hipStream_t st, int nunits, int N, int SF, int SMAX, const double *Mu,
                        const double *G, double *H)
{
    dim3 grid(nunits, (N + 255) / 256);
    switch (SMAX) {
    case 8: hipLaunchKernelGGL(k_half_ls_gemv<8>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    case 16: hipLaunchKernelGGL(k_half_ls_gemv<16>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    case 24: hipLaunchKernelGGL(k_half_ls_gemv<24>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    case 32: hipLaunchKernelGGL(k_half_ls_gemv<32>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    case 48: hipLaunchKernelGGL(k_half_ls_gemv<48>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    case 64: hipLaunchKernelGGL(k_half_ls_gemv<64>, grid, dim3(256), 0, st, N, SF, Mu, G, H); break;
    default: return -1;
    }
    return 0;
}

void launch_permute_mu(hipStream_t st, int N, const int *perm, const double *in, double *out)
{
    const long long n = 3LL * N * N;
    hipLaunchKernelGGL(k_permute_mu, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, N,
                       perm, in, out);
}

void launch_build_g_sym(hipStream_t st, int nunits, int N, int Gtot, const ClassDev &cd,
                        const double *const *uptr, const int *unit_ideg, SymDev r0, SymDev r1,
                        double *G)
{
    dim3 grid((cd.K + 63) / 64, nunits);
    hipLaunchKernelGGL(k_build_g_sym, grid, dim3(64), 0, st, N, Gtot, cd, uptr, unit_ideg, r0,
                       r1, G);
}

int launch_gemv_sym(hipStream_t st, int nunits, int N, int Gtot, int Htot, const ClassDev &cd,
                    const double *mu, const double *G, double *H, int ld, int vbase, int rlo,
                    int rhi)
{
    if (ld <= 0) ld = N;
    if (rhi < 0) rhi = cd.n;
    if (rlo >= rhi || nunits <= 0) return 0;
    dim3 grid(nunits, (rhi - rlo + 255) / 256);
#define L_(m) hipLaunchKernelGGL(k_gemv_sym<m>, grid, dim3(256), 0, st, N, Gtot, Htot, cd, mu, ld, vbase, rlo, rhi, G, H)
    switch (cd.smax) {
    case 8: L_(8); break;
    case 16: L_(16); break;
    case 24: L_(24); break;
    case 32: L_(32); break;
    case 48: L_(48); break;
    case 64: L_(64); break;
    default: return -1;
    }
#undef L_
    return 0;
}

void launch_fold(hipStream_t st, int nitems, int nsizeF, int ndegI, int ndegF, int SFN,
                 const int *item_unit, SymDev t0, SymDev t1, const double *H, double *Ht)
{
    dim3 grid((nsizeF + 255) / 256, nitems, ndegI * ndegF);
    hipLaunchKernelGGL(k_fold, grid, dim3(256), 0, st, nsizeF, ndegI, ndegF, SFN, item_unit, t0,
                       t1, H, Ht);
}

void launch_fold_jobs(hipStream_t st, const FoldJob *d_jobs, const int *d_begin, int njobs,
                      int nblocks)
{
    if (njobs <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_fold_jobs, dim3(nblocks), dim3(256), 0, st, d_jobs, d_begin, njobs);
}

void launch_corr_intensity(hipStream_t st, const CorrArgs &a)
{
    const long long threads = (long long)a.ncand * 64;
    const int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(k_corr_intensity, dim3(blocks), dim3(256), 0, st, a);
}

int compact_blocks(long long n) { return (int)((n + GI_CPB - 1) / GI_CPB); }

void launch_compact(hipStream_t st, long long n, const int *flag, int *counts, long long *offs,
                    LineSoA in, LineSoA out)
{
    const int nb = compact_blocks(n);
    if (nb == 0) return;
    // when every candidate passes, k_scatter leaves `out` alone and the
    // caller copies from `in` (same rows, same order)
    hipLaunchKernelGGL(k_block_count, dim3(nb), dim3(GI_CT), 0, st, n, flag, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(GI_CT), 0, st, nb, counts, offs);
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(GI_CT), 0, st, n, flag, offs, in, out);
}

}  // namespace gi
