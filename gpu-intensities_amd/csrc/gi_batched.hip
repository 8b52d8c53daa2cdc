// Batched FP64 form of the half linestrength and the correlation on the
// gfx950 matrix cores (v_mfma_f64_16x16x4f64).
//
// K5 (k_halfls_jobs): for one output class co and a chunk of units,
//   H_co[m, (u, j)] = sum_k A[m, k] * B[k, (u, j)],
//   A[m, k] = mu_{a(e)}[off_co + m, off_e + k - row0_e]   (column-major, ld N,
//             or one streamed part of vF rows: MuRows)
//   B[k, (u, j)] = G[u*Gtot + gbase_p + k*ncols + j]  (packed G)
// so mu blocks are read once per chunk instead of once per lower level.
//
// K6 (k_corr_mfma): for one chunk of items (lower levels of one
// (J, symmetry, J') group, energy sorted), the banded product
//   A[f, n] = sum_t U[col(f)*nsize + t] * Ht[n*nsize + t],
// n = item*ncb + component pair, f in the union of the items' windows; tiles
// outside every window of their items are skipped, entries outside an item's
// own window are not stored.
//
// Operand layout of the 16x16x4 f64 instruction: lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator register r holds
// D[4*r + l/16][l%16], r = 0..3 (unlike the f32 16x16x4 form, whose register r
// holds row 4*(l/16) + r).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gi_kernels.h"

namespace gi {

typedef double d4 __attribute__((ext_vector_type(4)));

#define BK 16

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

// Tile shapes. BM x BN output tile per 256-thread workgroup (4 wavefronts in
// a 2 x 2 layout, each owning a (BM/2) x (BN/2) block of 16 x 16 MFMA
// tiles); K staged through LDS in steps of BK with a register prefetch of
// the next step. LDS rows are padded to BM + 4 doubles (measured on gfx950: the
// correlation's staging stores, 4 lanes on 4 K rows of one record, take 16
// cycles per wave instruction at a pitch of BM + 16 and 6.3 at BM + 4; the
// operand reads and the half-ls stores run at the same rate with either).
template <int BM_, int BN_>
struct Tile {
    static constexpr int PA = BM_ + 4, PB = BN_ + 4;
    static constexpr int MI = BM_ / 32, NI = BN_ / 32;  // 16x16 blocks per wave
    static constexpr int LA = BM_ * BK / 256, LB = BN_ * BK / 256;
};

// Half-ls tiles with NW_ wavefronts in a 2 x (NW_/2) layout (4: 2 x 2 as
// Tile; 8: each wave owns (BM/2) x (BN/4), half the accumulators, so twice
// the waves fit in the register file).
template <int BM_, int BN_, int NW_>
struct TileW {
    static constexpr int NT = 64 * NW_, WR = 2, WC = NW_ / 2;
    static constexpr int PA = BM_ + 4, PB = BN_ + 4;
    static constexpr int MI = BM_ / (16 * WR), NI = BN_ / (16 * WC);
    static constexpr int LA = BM_ * BK / NT, LB = BN_ * BK / NT;
};

template <int MI, int NI, int PA, int PB>
__device__ inline void mfma_step_w(const double *__restrict__ As, const double *__restrict__ Bs,
                                   int wm, int wn, int lane, d4 (&acc)[MI][NI])
{
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        const int k = kk + (lane >> 4);
        const int i = lane & 15;
        double a[MI], b[NI];
#pragma unroll
        for (int x = 0; x < MI; x++) a[x] = As[k * PA + wm + 16 * x + i];
#pragma unroll
        for (int y = 0; y < NI; y++) b[y] = Bs[k * PB + wn + 16 * y + i];
#pragma unroll
        for (int x = 0; x < MI; x++)
#pragma unroll
            for (int y = 0; y < NI; y++)
                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
}

template <int BM_, int BN_>
__device__ inline void mfma_step(const double *__restrict__ As, const double *__restrict__ Bs,
                                 int wm, int wn, int lane,
                                 d4 (&acc)[Tile<BM_, BN_>::MI][Tile<BM_, BN_>::NI])
{
    using T = Tile<BM_, BN_>;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        const int k = kk + (lane >> 4);
        const int i = lane & 15;
        double a[T::MI], b[T::NI];
#pragma unroll
        for (int x = 0; x < T::MI; x++) a[x] = As[k * T::PA + wm + 16 * x + i];
#pragma unroll
        for (int y = 0; y < T::NI; y++) b[y] = Bs[k * T::PB + wn + 16 * y + i];
#pragma unroll
        for (int x = 0; x < T::MI; x++)
#pragma unroll
            for (int y = 0; y < T::NI; y++)
                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
}

// mfma_step for the banded correlation: only the wave's 16 x 16 blocks whose
// bit is set in `need` (wave-uniform) are multiplied; the others stay zero.
template <int BM_, int BN_>
__device__ inline void mfma_step_mask(const double *__restrict__ As,
                                      const double *__restrict__ Bs, int wm, int wn, int lane,
                                      unsigned need,
                                      d4 (&acc)[Tile<BM_, BN_>::MI][Tile<BM_, BN_>::NI])
{
    using T = Tile<BM_, BN_>;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
        const int k = kk + (lane >> 4);
        const int i = lane & 15;
        double a[T::MI], b[T::NI];
#pragma unroll
        for (int x = 0; x < T::MI; x++) a[x] = As[k * T::PA + wm + 16 * x + i];
#pragma unroll
        for (int y = 0; y < T::NI; y++) b[y] = Bs[k * T::PB + wn + 16 * y + i];
#pragma unroll
        for (int x = 0; x < T::MI; x++)
#pragma unroll
            for (int y = 0; y < T::NI; y++)
                if ((need >> (x * T::NI + y)) & 1u)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
}

// Bit x*NI + y set when the wave's 16 x 16 block (rows f0 + wm + 16x, columns
// n0 + wn + 16y) holds an entry inside some column's window [item_lo, item_hi)
// (rows and windows are indices into the upper block). Wave-uniform.
__device__ inline unsigned band_blocks(int MI, int NI, int f0, int n0, int wm, int wn, int lane,
                                       int Hi, int ncols, int ncb, const int *__restrict__ lo,
                                       const int *__restrict__ hi)
{
    unsigned need = 0;
    for (int x = 0; x < MI; x++)
        for (int y = 0; y < NI; y++) {
            const int r0 = f0 + wm + 16 * x;
            const int col = n0 + wn + 16 * y + (lane & 15);
            bool hit = false;
            if (col < ncols && r0 < Hi) {
                const int it = col / ncb;
                hit = lo[it] < r0 + 16 && hi[it] > r0;
            }
            if (__ballot(hit) != 0ull) need |= 1u << (x * NI + y);
        }
    return __builtin_amdgcn_readfirstlane(need);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
// XCD w % 8). Tile t = (w % 8) * (grid / 8) + w / 8 gives each XCD one
// contiguous range of tiles, so tiles that share an operand slab run at the
// same time on the same L2. The grid is a multiple of 8; tiles >= ntiles exit.
__device__ inline int xcd_tile(int w, int grid) { return (w & 7) * (grid >> 3) + (w >> 3); }
static inline int xcd_grid(int ntiles) { return (ntiles + 7) & ~7; }

// K1 over chunks: G of every unit, element by element in G order, so each
// wavefront writes whole cache lines and the tiles of one unit (which read
// the same eigenvector record) run back to back on one XCD.
__global__ __launch_bounds__(256) void k_build_g_units(int N, const GBuildJob *__restrict__ jobs,
                                                       const int *__restrict__ begin, int njobs,
                                                       int ntiles)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const GBuildJob &J = jobs[find_job(begin, njobs, tile)];
    const int local = tile - J.tile_begin;
    const int unit = local / J.tiles_per_unit, t = local - unit * J.tiles_per_unit;
    const double *u = J.uptr[unit];
    const SymDev R = J.uideg[unit] ? J.r1 : J.r0;
    double *Gu = J.G + (size_t)unit * J.Gtot;
    for (int q = 0; q < kGTile / 256; q++) {
        const int g = t * kGTile + q * 256 + (int)threadIdx.x;
        if (g >= J.Gtot) break;
        int co = 0;
        while (co + 1 < J.nclasses && J.classes[co + 1].gbase_p <= g) co++;
        const ClassDev &cd = J.classes[co];
        const int off = g - cd.gbase_p;
        const int row = off / cd.ncols, j = off - row * cd.ncols;
        double v = 0.0;
        {
            int e = 0;
            while (e + 1 < cd.nseg && cd.seg_row0[e + 1] <= row) e++;
            const int vI = cd.seg_off[e] + (row - cd.seg_row0[e]);
            const int qq = e * cd.ncols + j;
            for (int p = cd.rptr[qq]; p < cd.rptr[qq + 1]; p++) {
                const int cell = cd.rsI[p] * N + vI;
                double cv = 0.0;
                for (int x = R.rptr[cell]; x < R.rptr[cell + 1]; x++) cv += u[R.rcol[x]] * R.rval[x];
                v += cd.rcoef[p] * cv;
            }
        }
        Gu[g] = v;
    }
}

// K1 from the sparse record map: tile = (unit group, 256 G elements); each
// thread loads its element's map entries once and applies them to the
// records of up to kGUnits units (writes are whole cache lines per unit).
__global__ __launch_bounds__(256) void k_build_g_map(const GMapJob *__restrict__ jobs,
                                                     const int *__restrict__ begin, int njobs,
                                                     int ntiles)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const GMapJob &J = jobs[find_job(begin, njobs, tile)];
    const int local = tile - J.tile_begin;
    const int ug = local / J.gtiles, gt = local - ug * J.gtiles;
    const int g = gt * 256 + (int)threadIdx.x;
    if (g >= J.Gtot) return;
    int ix[kGMapMaxW];
    double wv[kGMapMaxW];
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
                    ix[w] = ip[(size_t)w * J.Gtot];
                    wv[w] = wp[(size_t)w * J.Gtot];
                }
            cur = d;
        }
        const double *rec = J.uptr[u];
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kGMapMaxW; w++)
            if (w < J.W) v += wv[w] * rec[ix[w]];
        J.G[(size_t)u * J.Gtot + g] = v;
    }
}

// k_build_g_map for maps of width <= 4: four consecutive G elements per thread
// (vector loads of the map), the maps of both lower components kept in
// registers (units of a degenerate block alternate them), and the gathers of
// four units issued before their stores. The same sums in the same order.
__global__ __launch_bounds__(256) void k_build_g_map4(const GMapJob *__restrict__ jobs,
                                                      const int *__restrict__ begin, int njobs,
                                                      int ntiles)
{
    constexpr int NB = 4;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const GMapJob &J = jobs[find_job(begin, njobs, tile)];
    const int local = tile - J.tile_begin;
    const int ug = local / J.gtiles, gt = local - ug * J.gtiles;
    const int g = (gt * 256 + (int)threadIdx.x) * 4;
    if (g >= J.Gtot) return;
    const int ng = min(4, J.Gtot - g);
    int ix[2][4][4];
    double wv[2][4][4];
#pragma unroll
    for (int d = 0; d < 2; d++)
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = w < J.W && r < ng;
                ix[d][w][r] = ok ? J.idx[d][(size_t)w * J.Gtot + g + r] : 0;
                wv[d][w][r] = ok ? J.wt[d][(size_t)w * J.Gtot + g + r] : 0.0;
            }
    const int u0 = ug * kGUnits, u1 = min(u0 + kGUnits, J.nunits);
    for (int ub = u0; ub < u1; ub += NB) {
        double v[NB][4];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            v[q][0] = v[q][1] = v[q][2] = v[q][3] = 0.0;
            const int u = ub + q;
            if (u >= u1) continue;
            const bool d1 = J.uideg[u] != 0;
            const double *rec = J.uptr[u];
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (w < J.W)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        v[q][r] += (d1 ? wv[1][w][r] : wv[0][w][r]) *
                                   rec[d1 ? ix[1][w][r] : ix[0][w][r]];
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int u = ub + q;
            if (u >= u1) continue;
            double *dst = J.G + (size_t)u * J.Gtot + g;
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < ng) dst[r] = v[q][r];
        }
    }
}

// K5 over jobs: H_co[m, (u, j)] = sum_k mu-block[m, k] * G[k, (u, j)] on the
// matrix cores. K runs segment by segment (a tile never straddles two mu
// blocks).
// One output tile (tile `local` of job J). Two LDS buffers: step s multiplies
// from buffer s & 1 while the loads of step s + 1 are in flight, then stores
// them into the other buffer (one barrier per K step). Ends with a barrier
// after the last read of LDS.
template <int BM_, int BN_, int NW_>
__device__ __attribute__((always_inline)) inline void halfls_tile(const MuSrc &src, const HalfJob &J, int local,
                                   double (*__restrict__ As)[BK * TileW<BM_, BN_, NW_>::PA],
                                   double (*__restrict__ Bs)[BK * TileW<BM_, BN_, NW_>::PB])
{
    using T = TileW<BM_, BN_, NW_>;
    const ClassDev &cd = J.cd;
    const int m0 = J.m_lo + (local % J.mt) * BM_, n0 = (local / J.mt) * BN_;
    const int mhi = J.m_hi, ld = src.ld;
    const double *__restrict__ mu = (const double *)src.mu;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * (BM_ / 2), wn = (wave >> 1) * (BN_ / T::WC);
    const int ncolsTot = J.nunits * cd.ncols;
    // loads: A row am, K rows ak0 + q*(NT/BM); B column bn, K rows bk0 + q*(NT/BN)
    const int am = tid % BM_, ak0 = tid / BM_;
    const int bn = tid % BN_, bk0 = tid / BN_;
    const bool mok = m0 + am < mhi;
    const int n = n0 + bn;
    const bool nok = n < ncolsTot;
    gptr<double> gcol = gmem(J.G);
    if (nok) {
        const int u = n / cd.ncols, jj = n - u * cd.ncols;
        gcol = gmem(J.G) + (size_t)u * J.Gtot + cd.gbase_p + jj;
    }
    d4 acc[T::MI][T::NI];
#pragma unroll
    for (int x = 0; x < T::MI; x++)
#pragma unroll
        for (int y = 0; y < T::NI; y++) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[T::LA], rb[T::LB];
    int e = 0, kk0 = 0;
    // the class's segment table, read every K step
    const gptr<int> seg_a = gmem(cd.seg_a), seg_off = gmem(cd.seg_off), seg_n = gmem(cd.seg_n),
                    seg_row0 = gmem(cd.seg_row0);
    auto fetch = [&](int se, int sk0) {
        const int len = min(BK, seg_n[se] - sk0);
        const double *mp = mu + (size_t)seg_a[se] * src.plane +
                           (size_t)(seg_off[se] + sk0) * ld + (cd.off + m0 + am - src.v0);
        gptr<double> gp = gcol + (size_t)(seg_row0[se] + sk0) * cd.ncols;
#pragma unroll
        for (int q = 0; q < T::LA; q++) {
            const int kl = ak0 + q * (T::NT / BM_);
            ra[q] = (mok && kl < len) ? mp[(size_t)kl * ld] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < T::LB; q++) {
            const int kl = bk0 + q * (T::NT / BN_);
            rb[q] = (nok && kl < len) ? gp[(size_t)kl * cd.ncols] : 0.0;
        }
    };
    auto stage = [&](int b) {
#pragma unroll
        for (int q = 0; q < T::LA; q++) As[b][(ak0 + q * (T::NT / BM_)) * T::PA + am] = ra[q];
#pragma unroll
        for (int q = 0; q < T::LB; q++) Bs[b][(bk0 + q * (T::NT / BN_)) * T::PB + bn] = rb[q];
    };
    if (cd.nseg > 0) {
        fetch(0, 0);
        stage(0);
        __syncthreads();
    }
    for (int b = 0; e < cd.nseg; b ^= 1) {
        kk0 += BK;
        if (kk0 >= seg_n[e]) {
            e++;
            kk0 = 0;
        }
        const bool more = e < cd.nseg;
        if (more) fetch(e, kk0);
        mfma_step_w<T::MI, T::NI, T::PA, T::PB>(As[b], Bs[b], wm, wn, lane, acc);
        if (more) stage(b ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < T::MI; x++)
#pragma unroll
        for (int y = 0; y < T::NI; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * x + 4 * r + (lane >> 4);
                const int nn = n0 + wn + 16 * y + (lane & 15);
                if (m >= mhi || nn >= ncolsTot) continue;
                const int u = nn / cd.ncols, jj = nn - u * cd.ncols;
                J.H[(size_t)u * J.Htot + cd.hbase + (size_t)jj * cd.n + m] = acc[x][y][r];
            }
}

template <int BM_, int BN_, int NW_ = 4>
__global__ __launch_bounds__(64 * NW_) __attribute__((amdgpu_waves_per_eu(NW_ == 8 ? 4 : 1)))
void k_halfls_jobs(MuSrc src, const HalfJob *__restrict__ jobs, const int *__restrict__ tbegin,
                   int njobs, int ntiles)
{
    using T = TileW<BM_, BN_, NW_>;
    __shared__ double As[2][BK * T::PA];
    __shared__ double Bs[2][BK * T::PB];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const HalfJob &J = jobs[find_job(tbegin, njobs, tile)];
    halfls_tile<BM_, BN_, NW_>(src, J, tile - J.tile_begin, As, Bs);
}

// ---------------------------------------------------------------------------
// K6 over chunks: tile t of chunk c covers rows f0 = Lo + (local % rt) * BM
// and columns n0 = (local / rt) * BN.
// One correlation tile (tile `local` of chunk a). Returns at once (after one
// barrier) for a tile outside every window; otherwise ends with a barrier
// after the last read of LDS.
template <int BM_, int BN_>
__device__ __attribute__((always_inline)) inline void corr_tile(const CorrMMArgs &a, int local, double *__restrict__ As,
                                 double *__restrict__ Bs, int &s_need)
{
    using T = Tile<BM_, BN_>;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncols = a.nitems * a.ncb;
    const int rt = (a.Hi - a.Lo + BM_ - 1) / BM_, per = rt * ((ncols + BN_ - 1) / BN_);
    const int kp = local / per, l2 = local - kp * per;
    const int f0 = a.Lo + (l2 % rt) * BM_;  // block-level index of the first row
    const int n0 = (l2 / rt) * BN_;
    const int kb = kp * a.kchunk, ke = min(a.nsize, kb + a.kchunk);
    double *dst = kp == 0 ? a.amp : a.ampx + (long long)(kp - 1) * a.ampx_stride;
    // skip tiles outside every window of the tile's items
    if (tid == 0) {
        int need = 0;
        const int i0 = n0 / a.ncb, i1 = min(a.nitems, (n0 + BN_ + a.ncb - 1) / a.ncb);
        for (int i = i0; i < i1 && !need; i++)
            if (a.item_lo[i] < f0 + BM_ && a.item_hi[i] > f0) need = 1;
        s_need = need;
    }
    __syncthreads();
    if (!s_need) return;
    const int wm = (wave & 1) * (BM_ / 2), wn = (wave >> 1) * (BN_ / 2);
    // the wave's blocks outside every window are not multiplied (band edges)
    const unsigned bneed = band_blocks(T::MI, T::NI, f0, n0, wm, wn, lane, a.Hi, ncols, a.ncb,
                                       a.item_lo, a.item_hi);
    d4 acc[T::MI][T::NI];
#pragma unroll
    for (int x = 0; x < T::MI; x++)
#pragma unroll
        for (int y = 0; y < T::NI; y++) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    // loads: K = 2 lt + 8 q + {0, 1} (q < 2) of row lr + 64p (p < BM/64 for A,
    // BN/64 for B): each lane reads 16 contiguous bytes per load, so one load
    // instruction reads 64 contiguous bytes of each of its 16 rows (rows of
    // Ht have an even stride; eigenvector records when vec_al); a pair that
    // would cross the end of the K part, or a record not 16-byte aligned, is
    // read as two doubles
    const int lr = tid >> 2, lt = tid & 3;
    gptr<double> Urow[BM_ / 64];
    gptr<double> Hrow[BN_ / 64];
#pragma unroll
    for (int p = 0; p < BM_ / 64; p++) {
        const int fr = f0 + lr + 64 * p;
        Urow[p] = fr < a.Hi ? gmem(a.vec) + (size_t)gmem(a.blk_col)[fr] * a.nsize : nullptr;
    }
#pragma unroll
    for (int p = 0; p < BN_ / 64; p++) {
        const int nr = n0 + lr + 64 * p;
        Hrow[p] = nr < ncols ? gmem(a.Ht) + (size_t)nr * a.ldh : nullptr;
    }
    typedef double d2a __attribute__((ext_vector_type(2)));
    auto ld2 = [&](gptr<double> row, bool al, int t) -> d2a {
        if (!row) return d2a{0.0, 0.0};
        if (al && t + 1 < ke) return *(gptr<d2a>)(row + t);
        return d2a{t < ke ? row[t] : 0.0, t + 1 < ke ? row[t + 1] : 0.0};
    };
    const bool ual = a.vec_al != 0;
    d2a ra[BM_ / 64][2], rb[BN_ / 64][2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int t = k0 + 2 * lt + 8 * q;
#pragma unroll
            for (int p = 0; p < BM_ / 64; p++) ra[p][q] = ld2(Urow[p], ual, t);
#pragma unroll
            for (int p = 0; p < BN_ / 64; p++) rb[p][q] = ld2(Hrow[p], true, t);
        }
    };
    fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
#pragma unroll
            for (int p = 0; p < BM_ / 64; p++) {
                As[(2 * lt + 8 * q) * T::PA + lr + 64 * p] = ra[p][q].x;
                As[(2 * lt + 8 * q + 1) * T::PA + lr + 64 * p] = ra[p][q].y;
            }
#pragma unroll
            for (int p = 0; p < BN_ / 64; p++) {
                Bs[(2 * lt + 8 * q) * T::PB + lr + 64 * p] = rb[p][q].x;
                Bs[(2 * lt + 8 * q + 1) * T::PB + lr + 64 * p] = rb[p][q].y;
            }
        }
        __syncthreads();
        if (k0 + BK < ke) fetch(k0 + BK);
        // raised issue priority while this wavefront multiplies, so the other
        // workgroup's loads and LDS stores fill the gaps (k_corr_mfma 1.47 -> 1.38 ms
        // at C2; the half-ls kernels gain nothing or lose; two LDS buffers with
        // one barrier per step measured 1.50 ms)
        __builtin_amdgcn_s_setprio(1);
        mfma_step_mask<BM_, BN_>(As, Bs, wm, wn, lane, bneed, acc);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < T::MI; x++)
#pragma unroll
        for (int y = 0; y < T::NI; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int f = f0 + wm + 16 * x + 4 * r + (lane >> 4);
                const int n = n0 + wn + 16 * y + (lane & 15);
                if (f >= a.Hi || n >= ncols) continue;
                const int i = n / a.ncb, cb = n - i * a.ncb;
                if (f < a.item_lo[i] || f >= a.item_hi[i]) continue;
                const long long o = a.item_out_off[i] + (f - a.item_lo[i]);
                const int idegF = cb / a.ndegI, idegI = cb - idegF * a.ndegI;
                dst[(long long)(a.amp_m * a.amp_m) * o + idegI + a.amp_m * idegF] = acc[x][y][r];
            }
}

// The 64 x 64 tile at six wavefronts per SIMD (79 registers, accumulators in
// VGPRs) instead of five: the kernel waits on its operand loads, and the sixth
// wavefront hides more of that (seven spill).
template <int BM_, int BN_>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BM_ * BN_ <= 4096 ? 6 : 1))) void k_corr_mfma(const CorrMMArgs *__restrict__ args,
                                                   const int *__restrict__ cbegin, int nchunks,
                                                   int ntiles)
{
    using T = Tile<BM_, BN_>;
    __shared__ double As[BK * T::PA];
    __shared__ double Bs[BK * T::PB];
    __shared__ int s_need;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const int ci = find_job(cbegin, nchunks, tile);
    corr_tile<BM_, BN_>(args[ci], tile - cbegin[ci], As, Bs, s_need);
}

// ---------------------------------------------------------------------------
// Degeneracy sum, Einstein A, intensity and emit test per candidate, from the
// amplitudes K6 stored (cuda_host.cu:1388-1407).
__device__ __attribute__((always_inline)) inline void intensity_one(const IntensArgs &a, int c)
{
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
    const int m = a.amp_m;
    double *am = a.amp + (long long)(m * m) * o;
    double ls = 0.0;
    for (int idegF = 0; idegF < a.ndegF; idegF++)
        for (int idegI = 0; idegI < a.ndegI; idegI++) {
            const int cb = idegF * a.ndegI + idegI;
            double A = 0.0;
            if ((a.cb_mask >> cb) & 1) {
                A = am[idegI + m * idegF];
                for (int p = 1; p < a.kparts; p++)
                    A += a.ampx[(long long)(p - 1) * a.ampx_stride + (long long)(m * m) * o +
                                idegI + m * idegF];
            }
            am[idegI + m * idegF] = A;
            ls += A * A;
        }
    for (int q = 0; q < m * m; q++) {
        const int idegI = q % m, idegF = q / m;
        if (idegI >= a.ndegI || idegF >= a.ndegF) am[q] = 0.0;
    }
    ls /= (double)a.ndegI;
    if (a.reduced && a.ndegF != 1 && a.ndegI != 1) ls *= (double)a.ndegI;
    const double eI = a.item_energy[item];
    const double eF = a.blk_en[f];
    const double nu = eF - eI;
    const double anu = fabs(nu);
    double beta = a.beta, q_stat = a.q_stat, thS = a.thresh_S, thI = a.thresh_I;
    if (a.qpar) {
        beta = a.qpar[0];
        q_stat = a.qpar[1];
        thS = a.qpar[2];
        thI = a.qpar[3];
    }
    const double boltz_fc = anu * exp(-(eI - a.zpe) * beta) * (1.0 - exp(-anu * beta)) / q_stat;
    const double A_einst = a.A_coef_s_1 * (double)((2 * a.jI) + 1) * ls * pow(anu, 3.0);
    const double S = ls * a.gnsI * (double)((2 * a.jI + 1) * (2 * a.jF + 1));
    const double I = S * a.intens_cm_mol * boltz_fc;
    const bool pass = a.blk_upok[f] && S >= thS && I >= thI;
    a.flag[o] = pass ? 1 : 0;
    a.out_iI[o] = a.item_level[item];
    a.out_iF[o] = a.blk_lev[f];
    a.out_nu[o] = nu;
    a.out_S[o] = S;
    a.out_A[o] = A_einst;
    a.out_I[o] = I;
}

__global__ __launch_bounds__(256) void k_intensity(IntensArgs a)
{
    intensity_one(a, blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_intensity_jobs(const IntensArgs *__restrict__ jobs,
                                                        const int *__restrict__ begin, int njobs)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    intensity_one(jobs[lo], (blockIdx.x - begin[lo]) * 256 + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Persistent round (see PersistRound).

// Every thread's stores of the task are made visible device-wide, then one
// thread counts the task done.
__device__ inline void persist_signal(int *cnt)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// One thread polls until *cnt reaches target, then every thread acquires what
// the signalling tasks stored. A wait that exceeds kPersistSpinMax polls sets
// *err; once *err is set every wait of the launch returns at once, so the
// launch ends after one such wait.
__device__ inline void persist_wait(const int *cnt, int target, int *err)
{
    if (threadIdx.x == 0 && target > 0) {
        long long n = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++n > kPersistSpinMax ||
                __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Three workgroups per CU (two for 128 x 128 half-ls tiles), as the
// half-ls kernel alone keeps.
template <int HM, int HN, int CM, int CN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HM * HN > 8192 ? 2 : 3)))
void k_round_persist(PersistRound p)
{
    using TH = TileW<HM, HN, 4>;
    using TC = Tile<CM, CN>;
    constexpr int kH = 2 * BK * (TH::PA + TH::PB), kC = BK * (TC::PA + TC::PB);
    __shared__ double sm[kH > kC ? kH : kC];
    __shared__ int s_task, s_need;
    int *done = p.sync, *ticket = p.sync + 3 * p.nch;
    const int total = p.nh + p.nf + p.nc + p.ni;
    for (;;) {
        if (threadIdx.x == 0) s_task = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        int t = s_task;
        __syncthreads();
        if (t >= total) return;
        if (t < p.nh) {
            const int j = find_job(p.hbeg, p.nhj, t);
            halfls_tile<HM, HN, 4>(p.src, p.hj[j], t - p.hbeg[j],
                                   (double(*)[BK * TH::PA])sm,
                                   (double(*)[BK * TH::PB])(sm + 2 * BK * TH::PA));
            persist_signal(done + p.hch[j]);
            continue;
        }
        t -= p.nh;
        if (t < p.nf) {
            const int j = find_job(p.fbeg, p.nfj, t);
            const int ch = p.fch[j];
            persist_wait(done + ch, p.need[ch], p.err);
            fold_block(p.fj[j], t - p.fbeg[j]);
            persist_signal(done + p.nch + ch);
            continue;
        }
        t -= p.nf;
        if (t < p.nc) {
            const int ch = find_job(p.cbeg, p.ncj, t);
            persist_wait(done + p.nch + ch, p.need[p.nch + ch], p.err);
            corr_tile<CM, CN>(p.cj[ch], t - p.cbeg[ch], sm, sm + BK * TC::PA, s_need);
            persist_signal(done + 2 * p.nch + ch);
            continue;
        }
        t -= p.nc;
        const int j = find_job(p.ibeg, p.nij, t);
        const int ch = p.ich[j];
        persist_wait(done + 2 * p.nch + ch, p.need[2 * p.nch + ch], p.err);
        intensity_one(p.ij[j], (t - p.ibeg[j]) * 256 + threadIdx.x);
    }
}

__global__ __launch_bounds__(256) void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero_ints(hipStream_t st, int *p, int n)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

template <int HM, int HN, int CM, int CN>
static int persist_launch(hipStream_t st, const PersistRound &p, int ncu)
{
    static int per_cu = 0;  // resident workgroups per CU (the same on every gfx950)
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_round_persist<HM, HN, CM, CN>,
                                                          256, 0) != hipSuccess || n < 1)
            n = 1;
        per_cu = n;
    }
    const int total = p.nh + p.nf + p.nc + p.ni;
    const int grid = std::min(total, per_cu * std::max(ncu, 1));
    if (grid <= 0) return 0;
    hipLaunchKernelGGL((k_round_persist<HM, HN, CM, CN>), dim3(grid), dim3(256), 0, st, p);
    return grid;
}

int launch_round_persist(hipStream_t st, int hcfg, int ccfg, const PersistRound &p, int ncu)
{
    const int hm = tile_rows(hcfg), hn = tile_cols(hcfg);
    if (ccfg == 2) {
        if (hm == 64 && hn == 64) return persist_launch<64, 64, 128, 64>(st, p, ncu);
        if (hm == 64) return persist_launch<64, 128, 128, 64>(st, p, ncu);
        if (hn == 64) return persist_launch<128, 64, 128, 64>(st, p, ncu);
        return persist_launch<128, 128, 128, 64>(st, p, ncu);
    }
    if (hm == 64 && hn == 64) return persist_launch<64, 64, 64, 64>(st, p, ncu);
    if (hm == 64) return persist_launch<64, 128, 64, 64>(st, p, ncu);
    if (hn == 64) return persist_launch<128, 64, 64, 64>(st, p, ncu);
    return persist_launch<128, 128, 64, 64>(st, p, ncu);
}

// ---------------------------------------------------------------------------

int tile_rows(int cfg) { return cfg == kTile128x128w8 ? 128 : (cfg & 2) ? 128 : 64; }
int tile_cols(int cfg) { return cfg == kTile128x128w8 ? 128 : (cfg & 1) ? 128 : 64; }

void launch_build_g_units(hipStream_t st, int N, const GBuildJob *d_jobs, const int *d_begin,
                          int njobs, int ntiles)
{
    if (njobs <= 0 || ntiles <= 0) return;
    hipLaunchKernelGGL(k_build_g_units, dim3(xcd_grid(ntiles)), dim3(256), 0, st, N, d_jobs,
                       d_begin, njobs, ntiles);
}

// G of one unit per workgroup: the unit's record read once into LDS, the map's
// gathers served from there (in global memory they land on scattered lines).
// PART: a record longer than nlds keeps its tail in global memory, and a
// gather beyond nlds reads it there (the same sums in the same order).
template <bool PART>
__global__ __launch_bounds__(1024) void k_build_g_map_lds(const GMapJob *__restrict__ jobs,
                                                          const int *__restrict__ ubegin,
                                                          int njobs)
{
    __shared__ double rec[kRecLds];
    __shared__ double cw[kCmapCodes];  // the compact map's weights
    const int ji = find_job(ubegin, njobs, blockIdx.x);
    const GMapJob &J = jobs[ji];
    const int u = blockIdx.x - ubegin[ji];
    const gptr<double> src = gmem(J.uptr[u]);
    const int nl = PART ? min(J.nsize, J.nlds) : J.nsize;
    const int d = J.uideg[u];
    const bool cm = J.cidx[d] != nullptr;  // uniform over the workgroup
    for (int i = threadIdx.x; i < nl; i += 1024) rec[i] = src[i];
    if (cm && threadIdx.x < kCmapCodes) cw[threadIdx.x] = J.cwt[d][threadIdx.x];
    __syncthreads();
    auto at = [&](int i) { return PART ? (i < nl ? rec[i] : src[i]) : rec[i]; };
    const int Gt = J.Gtot, W = J.W;
    const gout<double> G = gmem_w(J.G) + (size_t)u * Gt;
    if (cm) {
        // the compact map: 3 bytes per entry, the weight from LDS (same doubles)
        const gptr<unsigned short> ip = gmem(J.cidx[d]);
        const gptr<unsigned char> cp = gmem(J.ccode[d]);
        for (int g0 = threadIdx.x; g0 < Gt; g0 += 4096) {
            int ix[4][4], cd[4][4];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    const int g = g0 + 1024 * r;
                    const bool ok = w < W && g < Gt;
                    ix[r][w] = ok ? (int)ip[(size_t)w * Gt + g] : 0;
                    cd[r][w] = ok ? (int)cp[(size_t)w * Gt + g] : 0;
                }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int g = g0 + 1024 * r;
                if (g >= Gt) break;
                double v = 0.0;
#pragma unroll
                for (int w = 0; w < 4; w++)
                    if (w < W) v += cw[cd[r][w]] * at(ix[r][w]);
                G[g] = v;
            }
        }
        return;
    }
    const gptr<int> ip = gmem(J.idx[d]);
    const gptr<double> wp = gmem(J.wt[d]);
    // four elements per thread per pass, 1024 apart: every map load of the pass
    // is in flight before the first gather (one element at a time left the
    // waves waiting on L2 latency most of the time)
    for (int g0 = threadIdx.x; g0 < Gt; g0 += 4096) {
        int ix[4][4];
        double wv[4][4];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const int g = g0 + 1024 * r;
                const bool ok = w < W && g < Gt;
                ix[r][w] = ok ? ip[(size_t)w * Gt + g] : 0;
                wv[r][w] = ok ? wp[(size_t)w * Gt + g] : 0.0;
            }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int g = g0 + 1024 * r;
            if (g >= Gt) break;
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (w < W) v += wv[r][w] * at(ix[r][w]);
            G[g] = v;
        }
    }
}

void launch_build_g_map4(hipStream_t st, const GMapJob *d_jobs, const int *d_begin, int njobs,
                         int ntiles)
{
    if (njobs <= 0 || ntiles <= 0) return;
    hipLaunchKernelGGL(k_build_g_map4, dim3(xcd_grid(ntiles)), dim3(256), 0, st, d_jobs, d_begin,
                       njobs, ntiles);
}

void launch_build_g_map_lds(hipStream_t st, const GMapJob *d_jobs, const int *d_begin,
                            int njobs, int nunits, bool part)
{
    if (njobs <= 0 || nunits <= 0) return;
    if (part)
        hipLaunchKernelGGL(k_build_g_map_lds<true>, dim3(nunits), dim3(1024), 0, st, d_jobs,
                           d_begin, njobs);
    else
        hipLaunchKernelGGL(k_build_g_map_lds<false>, dim3(nunits), dim3(1024), 0, st, d_jobs,
                           d_begin, njobs);
}

void launch_build_g_map(hipStream_t st, const GMapJob *d_jobs, const int *d_begin, int njobs,
                        int ntiles)
{
    if (njobs <= 0 || ntiles <= 0) return;
    hipLaunchKernelGGL(k_build_g_map, dim3(xcd_grid(ntiles)), dim3(256), 0, st, d_jobs, d_begin,
                       njobs, ntiles);
}

void launch_halfls_jobs(hipStream_t st, int cfg, const MuSrc &src, const HalfJob *d_jobs,
                        const int *d_tbegin, int njobs, int ttiles)
{
    if (njobs <= 0 || ttiles <= 0) return;
    const dim3 g(xcd_grid(ttiles)), b(256);
    if (cfg == kTile128x128w8) {
        hipLaunchKernelGGL((k_halfls_jobs<128, 128, 8>), g, dim3(512), 0, st, src, d_jobs,
                           d_tbegin, njobs, ttiles);
        return;
    }
    switch (cfg & 3) {
    case 0: hipLaunchKernelGGL((k_halfls_jobs<64, 64>), g, b, 0, st, src, d_jobs, d_tbegin, njobs, ttiles); break;
    case 1: hipLaunchKernelGGL((k_halfls_jobs<64, 128>), g, b, 0, st, src, d_jobs, d_tbegin, njobs, ttiles); break;
    case 2: hipLaunchKernelGGL((k_halfls_jobs<128, 64>), g, b, 0, st, src, d_jobs, d_tbegin, njobs, ttiles); break;
    default: hipLaunchKernelGGL((k_halfls_jobs<128, 128>), g, b, 0, st, src, d_jobs, d_tbegin, njobs, ttiles); break;
    }
}

void launch_corr_mfma(hipStream_t st, int cfg, const CorrMMArgs *d_args, const int *d_cbegin,
                      int nchunks, int ntiles)
{
    if (nchunks <= 0 || ntiles <= 0) return;
    const dim3 g(xcd_grid(ntiles)), b(256);
    switch (cfg & 3) {
    case 0: hipLaunchKernelGGL((k_corr_mfma<64, 64>), g, b, 0, st, d_args, d_cbegin, nchunks, ntiles); break;
    case 1: hipLaunchKernelGGL((k_corr_mfma<64, 128>), g, b, 0, st, d_args, d_cbegin, nchunks, ntiles); break;
    case 2: hipLaunchKernelGGL((k_corr_mfma<128, 64>), g, b, 0, st, d_args, d_cbegin, nchunks, ntiles); break;
    default: hipLaunchKernelGGL((k_corr_mfma<128, 128>), g, b, 0, st, d_args, d_cbegin, nchunks, ntiles); break;
    }
}

void launch_intensity_jobs(hipStream_t st, const IntensArgs *d_jobs, const int *d_begin,
                           int njobs, int nblocks)
{
    if (njobs <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_intensity_jobs, dim3(nblocks), dim3(256), 0, st, d_jobs, d_begin, njobs);
}

void launch_intensity(hipStream_t st, const IntensArgs &a)
{
    if (a.ncand <= 0) return;
    hipLaunchKernelGGL(k_intensity, dim3((a.ncand + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace gi
