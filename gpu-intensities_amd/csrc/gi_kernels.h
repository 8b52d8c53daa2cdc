// Kernel argument structs and launchers of the gfx950 line-intensity engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#ifndef GI_MAXDEG
#define GI_MAXDEG 2
#endif

namespace gi {

// Device memory seen through a pointer read from a job table. The compiler
// treats such pointers as generic and loads through them as flat loads, which
// also count against the LDS wait counter: an LDS wait in a kernel's multiply
// step then also waits for the next step's operand loads. gmem() marks the
// pointer as device (global) memory, so only the loads' own wait covers them.
template <class T>
using gptr = const __attribute__((address_space(1))) T *;
template <class T>
__device__ inline gptr<T> gmem(const T *p)
{
    return (gptr<T>)p;
}
template <class T>
using gout = __attribute__((address_space(1))) T *;
template <class T>
__device__ inline gout<T> gmem_w(T *p)
{
    return (gout<T>)p;
}

// Device view of one symmetrisation map (R and R^T in CSR form).
struct SymDev {
    const int *rptr, *rcol;
    const double *rval;
    const int *tptr, *tcol;
    const double *tval;
    // the fold's rows as fixed pairs (GroupPlan::tfc) when every row has at
    // most two entries: per row two H offsets (-1: no entry; -2, a longer row
    // read from the CSR, is not produced by the planner but kept by the
    // kernels) and their weights; null: the CSR only
    const int *tfc;
    const double *tfw;
};

// Structure-of-arrays line records; amp_w amplitudes per line.
struct LineSoA {
    int *iI, *iF;
    double *nu, *S, *A, *I, *amp;
    int amp_w = 4;
};

// Device view of one output class of a group plan (see OutClass).
struct ClassDev {
    int off, n, ncols, smax, K, nseg, hbase, gbase, gbase_p;
    int kpad, cbaseB, eoff;          // split mode (see OutClass)
    const int *seg_pad;              // [nseg][2]: row offset in a digit column, mu column offset
    const int *seg_a, *seg_off, *seg_n, *seg_row0, *rptr, *rsI;
    const double *rcoef;
};

struct CorrArgs {
    int ncand, nitems, nsizeF, ndegI, ndegF, cb_mask, reduced, jI, jF;
    const int *item_cand_start, *item_lo, *item_level;
    const long long *item_out_off;
    const double *item_energy;
    const int *blk_col, *blk_lev;
    const double *blk_en;
    const unsigned char *blk_upok;
    const double *vec, *Ht;
    double gnsI, zpe, beta, q_stat, A_coef_s_1, intens_cm_mol, thresh_S, thresh_I;
    const double *qpar;              // non-null: beta, q_stat, thresh_S, thresh_I read here
    int *flag, *out_iI, *out_iF;
    double *out_nu, *out_S, *out_A, *out_I, *out_amp;
    int amp_m;                       // amplitudes per line amp_m^2, entry idegI + amp_m*idegF
};

// One chunk of the banded correlation product (batched mode).
struct CorrMMArgs {
    int Lo, Hi;                      // union of the items' windows (block-level indices)
    int nitems, ncb, ndegI, nsize;
    const int *item_lo, *item_hi;
    const long long *item_out_off;
    const int *blk_col;
    const double *vec, *Ht;
    double *amp;                     // candidate amplitudes, amp_m^2 per candidate
    int amp_m;                       // entry idegI + amp_m*idegF
    int kparts, kchunk;              // split K as in CorrSplitArgs
    double *ampx;
    long long ampx_stride;
    int ldh;                         // row stride of Ht (even: 16-byte aligned rows)
    int vec_al;                      // every record of vec starts 16-byte aligned
};

// Intensity step from stored amplitudes (batched mode).
struct IntensArgs {
    int ncand, nitems, ndegI, ndegF, cb_mask, reduced, jI, jF;
    const int *item_cand_start, *item_lo, *item_level;
    const long long *item_out_off;
    const double *item_energy;
    const int *blk_lev;
    const double *blk_en;
    const unsigned char *blk_upok;
    double gnsI, zpe, beta, q_stat, A_coef_s_1, intens_cm_mol, thresh_S, thresh_I;
    const double *qpar;              // non-null: beta, q_stat, thresh_S, thresh_I read here
    double *amp;
    int amp_m;                       // amp_m^2 amplitudes per candidate, idegI + amp_m*idegF
    int kparts;                      // correlation parts: amp + ampx[0..kparts-2], in order
    const double *ampx;
    long long ampx_stride;
    int *flag, *out_iI, *out_iF;
    double *out_nu, *out_S, *out_A, *out_I;
};

// Where one half-ls launch reads mu: element (a, vI, vF) of the permuted
// FP64 mu at mu + a*plane + vI*ld + (vF - v0), digit row (a, vF) of the split
// mode at (a*ld + vF - v0)*kpad_all in planes `plane` bytes apart. Resident
// mu: ld = N, v0 = 0. Streamed mu: one part of ld vF rows starting at v0,
// whose jobs compute the class rows [m_lo, m_hi) inside it (each row's sum is
// the same either way). A kernel argument, so it stays in scalar registers.
struct MuSrc {
    const void *mu;
    long long plane;
    int ld, v0;
};

// One (chunk, output class) of the batched half linestrength.
struct HalfJob {
    ClassDev cd;
    int m_lo, m_hi;                  // rows of the class computed (see MuSrc)
    int nunits, Gtot, Htot;
    int tile_begin;                  // first tile of the product
    int mt;                          // product tiles along rows
    const double *const *uptr;
    const int *uideg;
    SymDev r0, r1;
    double *G, *H;                   // this chunk's G and H
};

// G build of one chunk in the packed layout: each unit's G is Gtot contiguous
// doubles, its output classes back to back at ClassDev::gbase_p as K rows of
// ncols, built in tiles of 2048.
struct GBuildJob {
    const ClassDev *classes;
    int nclasses, nunits, Gtot, tiles_per_unit, tile_begin;
    const double *const *uptr;
    const int *uideg;
    SymDev r0, r1;
    double *G;
};
constexpr int kGTile = 2048;

// G build from the sparse record map (GroupPlan::gidx/gwt): one thread per G
// element, kGUnits units per workgroup reuse the map entries it loaded.
struct GMapJob {
    const int *idx[2];
    const double *wt[2];
    int W, Gtot, nunits, gtiles, tile_begin;
    const double *const *uptr;
    const int *uideg;
    double *G;
    int nsize;                       // record length of the lower block
    int nlds;                        // leading doubles of a record staged in LDS (<= kRecLds)
    // the compact map (GroupPlan::cmap; null when the group has none): 16-bit
    // record index and 8-bit code into kCmapCodes weights per entry
    const unsigned short *cidx[2];
    const unsigned char *ccode[2];
    const double *cwt[2];
};
// Weights of a compact record map per lower component (code 0 = weight 0).
constexpr int kCmapCodes = 256;
constexpr int kGMapMaxW = 16;
constexpr int kClsLds = 32;  // output classes of a group that k_build_g_split_lds keeps in LDS
constexpr int kGUnits = 16;
// Doubles of a record the one-unit-per-workgroup G builds stage in LDS
// (launch_build_g_map_lds, launch_build_g_split_lds): 156 KB of the 160 KB a
// gfx950 workgroup can hold, leaving room for the compact map's weights and
// the split build's class table. Longer records keep their tail in global
// memory.
constexpr int kRecLds = 19968;
void launch_build_g_map(hipStream_t st, const GMapJob *d_jobs, const int *d_begin, int njobs,
                        int ntiles);
// One workgroup per unit, the first nlds doubles of its record in LDS (maps of
// width <= 4); d_begin = first unit of each job. part: some job's records are
// longer than its nlds, and their gathers beyond it read global memory.
// k_build_g_map for maps of width <= 4 with 1024 G elements per tile (gtiles
// counts 1024-element tiles).
void launch_build_g_map4(hipStream_t st, const GMapJob *d_jobs, const int *d_begin, int njobs,
                         int ntiles);
void launch_build_g_map_lds(hipStream_t st, const GMapJob *d_jobs, const int *d_begin,
                            int njobs, int nunits, bool part);
void launch_build_g_units(hipStream_t st, int N, const GBuildJob *d_jobs, const int *d_begin,
                          int njobs, int ntiles);

// Split-precision mode (gi_split.hip): digit planes of int8, see there.
struct GSplitJob {
    const ClassDev *classes;
    int nclasses;
    const int *idx[2];
    const double *wt[2], *bnd[2];
    int W, GtotB, EGtot, nunits, gtiles, tile_begin;
    const double *const *uptr;
    const int *uideg, *urec;         // per unit: component, record in the lower block
    const double *umax;              // per record of the lower block: max |u|
    int8_t *D;                       // this chunk's G digits (plane 0)
    long long plane;
    int *E;                          // this chunk's column exponents [unit][EGtot]
    int nsize;                       // record length of the lower block
    int nlds;                        // leading doubles of a record staged in LDS (<= kRecLds)
};
struct SplitHalfJob {
    ClassDev cd;
    int m_lo, m_hi;                  // rows of the class computed (see MuSrc)
    int nunits, tile_begin, mt, GtotB, EGtot, Htot;
    const int8_t *Gd;
    long long gplane;
    const int *Ge;
    double *H;
};
struct CorrSplitArgs {
    int Lo, Hi, nitems, ncb, ndegI, Kp;
    const int *item_lo, *item_hi;
    const long long *item_out_off;
    const int *blk_col;
    const int8_t *Ud;                // upper block digits (plane 0), rows of Kp
    long long uplane;
    const int *Ue;                   // per record
    const int8_t *Hd;                // folded vectors' digits (plane 0), rows of Kp
    long long hplane;
    const int *He;                   // per column
    double *amp;
    int amp_m;                       // amp_m^2 amplitudes per candidate, idegI + amp_m*idegF
    // split K: part p of kparts sums K in [p*kchunk, (p+1)*kchunk); part 0 writes
    // amp, part p > 0 writes ampx + (p-1)*ampx_stride (same candidate layout)
    int kparts, kchunk;
    double *ampx;
    long long ampx_stride;
};
void launch_split_rows(hipStream_t st, int S, double R, const double *X, long long ld, int rows, int K,
                       int Kp, int8_t *D, long long plane, int *E, double *Mx);
// Digit rows of nrows vF rows of mu: element (a, vI, vF) read at
// mu[vF + vI*ld + a*N*ld] (ld = N: resident permuted mu; ld = nrows: a part
// of vF rows), digit row (a, vF) written at D + (a*ld + vF)*kpad_all, the
// exponent of row vF at E[vF].
void launch_split_mu(hipStream_t st, int S, double R, int N, int ld, int nrows, int kpad_all,
                     const double *mu, const int *vinv, int8_t *D, long long plane, int *E);
// four: every job's record map has width <= 4 and gtiles counts 1024 rows
// (k_build_g_split4); otherwise 256 rows per tile.
void launch_build_g_split(hipStream_t st, int S, double R, const GSplitJob *d_jobs, const int *d_begin,
                          int njobs, int ntiles, bool four);
// One workgroup per unit: the first nlds doubles of its record in LDS, then
// every G row of the unit (maps of width <= 4); d_begin = first unit of each
// job; part as for launch_build_g_map_lds.
void launch_build_g_split_lds(hipStream_t st, int S, double R, const GSplitJob *d_jobs,
                              const int *d_begin, int njobs, int nunits, bool part);
// cfg 0: 64 x 64 tiles; 1: 128-row tiles (SplitHalfJob.mt counts 128-row
// tiles); 2: 64 x 64 with two K steps of loads in flight, 3: 128 x 64 on eight
// wavefronts with two K steps in flight, both for classes of at most
// kSplitMaxSteps K steps and kSplitMaxSegs segments and at most 5 digits (else
// the one-set kernel of the same rows runs).
constexpr int kSplitMaxSteps = 512, kSplitMaxSegs = 64;
// K steps (64 bytes of one segment each) of a class with segments of seg_n rows.
__host__ __device__ inline int split2_steps(const int *seg_n, int nseg)
{
    int n = 0;
    for (int e = 0; e < nseg; e++) n += (seg_n[e] + 63) / 64;
    return n;
}
// Whether k_halfls_split2's step table holds the class. The host routes only
// such classes to it; a class that does not fit gets NaN results there instead
// of silently dropped K steps.
__host__ __device__ inline bool split2_class_fits(const int *seg_n, int nseg)
{
    return nseg <= kSplitMaxSegs && split2_steps(seg_n, nseg) <= kSplitMaxSteps;
}
void launch_halfls_split(hipStream_t st, int S, double R, int kpad_all, const MuSrc &src,
                         const int *mue, const SplitHalfJob *d_jobs, const int *d_tbegin,
                         int njobs, int ntiles, int cfg);
void launch_corr_split(hipStream_t st, int S, double R, const CorrSplitArgs *d_args, const int *d_cbegin,
                       int nchunks, int ntiles);

// Tile configurations of the two matrix-core kernels: cfg bit 1 selects 128
// rows (else 64), bit 0 selects 128 columns (else 64).
// half-ls tile setting 5: 128 x 128 tiles on 8 wavefronts (each 64 x 32)
constexpr int kTile128x128w8 = 5;
int tile_rows(int cfg);
int tile_cols(int cfg);
void launch_halfls_jobs(hipStream_t st, int cfg, const MuSrc &src, const HalfJob *d_jobs,
                        const int *d_tbegin, int njobs, int ttiles);
void launch_corr_mfma(hipStream_t st, int cfg, const CorrMMArgs *d_args, const int *d_cbegin,
                      int nchunks, int ntiles);
void launch_intensity(hipStream_t st, const IntensArgs &a);

void launch_correlate_prim(hipStream_t st, int dimen, const int *r_grid, SymDev R,
                           const double *u, double *out);
void launch_grid_to_prim(hipStream_t st, int dimen, const int *r_grid, const double *grid,
                         double *out);
void launch_build_g(hipStream_t st, int nunits, int N, int SI, int SF, int SMAX,
                    const double *const *uptr, const int *unit_ideg, SymDev r0, SymDev r1,
                    const int *cptr, const int *csI, const double *cval, double *G);
int launch_half_ls_gemv(hipStream_t st, int nunits, int N, int SF, int SMAX, const double *Mu,
                        const double *G, double *H);
void launch_permute_mu(hipStream_t st, int N, const int *perm, const double *in, double *out);
void launch_build_g_sym(hipStream_t st, int nunits, int N, int Gtot, const ClassDev &cd,
                        const double *const *uptr, const int *unit_ideg, SymDev r0, SymDev r1,
                        double *G);
// rows [rlo, rhi) of the class (rhi < 0: all); mu with leading dimension ld
// (<= 0: N) holding vF rows from vbase (a streamed part of mu)
int launch_gemv_sym(hipStream_t st, int nunits, int N, int Gtot, int Htot, const ClassDev &cd,
                    const double *mu, const double *G, double *H, int ld = 0, int vbase = 0,
                    int rlo = 0, int rhi = -1);
void launch_fold(hipStream_t st, int nitems, int nsizeF, int ndegI, int ndegF, int SFN,
                 const int *item_unit, SymDev t0, SymDev t1, const double *H, double *Ht);

// The folds of every chunk of a round in one launch: job = one chunk's
// launch_fold arguments; blocks of a job = xt * nitems * ndegI * ndegF.
struct FoldJob {
    int nitems, nsizeF, ndegI, ndegF, SFN, xt;  // xt: blocks of kFoldRows rows per column
    const int *item_unit;
    SymDev t0, t1;
    const double *H;
    double *Ht;
    int ldh;                         // row stride of Ht (0: nsizeF)
};
constexpr int kFoldRows = 1024;  // rows of one k_fold_jobs block (4 per thread)

// R rows of one folded column, acc[r] = sum over the CSR row t[r] of R_F^T of
// w * Hu[g] (rows with ok[r] false give 0). The bounds of every row, then the
// first two entries of each row with their H values, are loaded before any is
// used, so R rows' loads are in flight together; longer rows finish in a loop.
// The sum runs in entry order, as one row at a time would.
template <int R>
__device__ inline void fold_rows(const SymDev &T, const double *__restrict__ Hu,
                                 const int (&t)[R], const bool (&ok)[R], double (&acc)[R])
{
    if (T.tfc) {
        // fixed pairs: one load of a row's offsets and weights, then its H values
        int2 cc[R];
        double2 ww[R];
        double h0[R], h1[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            cc[r] = ok[r] ? ((const int2 *)T.tfc)[t[r]] : make_int2(-1, -1);
            ww[r] = ok[r] ? ((const double2 *)T.tfw)[t[r]] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            h0[r] = cc[r].x >= 0 ? Hu[cc[r].x] : 0.0;
            h1[r] = cc[r].y >= 0 ? Hu[cc[r].y] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            double v = 0.0;
            if (cc[r].x >= 0) v += ww[r].x * h0[r];
            if (cc[r].y >= 0) v += ww[r].y * h1[r];
            if (cc[r].x == -2)  // a row of more than two entries, in entry order
                for (int e = T.tptr[t[r]]; e < T.tptr[t[r] + 1]; e++) v += T.tval[e] * Hu[T.tcol[e]];
            acc[r] = v;
        }
        return;
    }
    int p0[R], p1[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        p0[r] = ok[r] ? T.tptr[t[r]] : 0;
        p1[r] = ok[r] ? T.tptr[t[r] + 1] : 0;
    }
    double w0[R], w1[R], h0[R], h1[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool a = p0[r] < p1[r], b = p0[r] + 1 < p1[r];
        const int c0 = a ? T.tcol[p0[r]] : 0, c1 = b ? T.tcol[p0[r] + 1] : 0;
        w0[r] = a ? T.tval[p0[r]] : 0.0;
        w1[r] = b ? T.tval[p0[r] + 1] : 0.0;
        h0[r] = a ? Hu[c0] : 0.0;
        h1[r] = b ? Hu[c1] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        double v = 0.0;
        if (p0[r] < p1[r]) v += w0[r] * h0[r];
        if (p0[r] + 1 < p1[r]) v += w1[r] * h1[r];
        for (int e = p0[r] + 2; e < p1[r]; e++) v += T.tval[e] * Hu[T.tcol[e]];
        acc[r] = v;
    }
}
// Block `local` of fold job J (256 threads, kFoldRows rows of one folded
// column).
__device__ __attribute__((always_inline)) inline void fold_block(const FoldJob &J, int local)
{
    const int xb = local % J.xt, rest = local / J.xt;
    const int item = rest % J.nitems, cb = rest / J.nitems;
    const int idegF = cb / J.ndegI, idegI = cb % J.ndegI;
    const int unit = J.item_unit[item * GI_MAXDEG + idegI];
    const SymDev &T = idegF ? J.t1 : J.t0;
    const double *Hu = J.H + (size_t)(unit >= 0 ? unit : 0) * J.SFN;
    double *out = J.Ht + ((size_t)item * J.ndegI * J.ndegF + cb) * (J.ldh ? J.ldh : J.nsizeF);
    int t[4];
    bool ok[4];
    double acc[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        t[r] = xb * kFoldRows + r * 256 + (int)threadIdx.x;
        ok[r] = t[r] < J.nsizeF && unit >= 0;
    }
    fold_rows<4>(T, Hu, t, ok, acc);
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (t[r] < J.nsizeF) out[t[r]] = acc[r];
}
void launch_fold_jobs(hipStream_t st, const FoldJob *d_jobs, const int *d_begin, int njobs,
                      int nblocks);
// The digit splits of every chunk of a round: one workgroup per row.
struct SplitRowsJob {
    const double *X;
    long long ld;
    int K, Kp;
    int8_t *D;
    long long plane;
    int *E;
};
void launch_split_rows_jobs(hipStream_t st, int S, double R, const SplitRowsJob *d_jobs,
                            const int *d_begin, int njobs, int nrows);
// Fold and digit split in one pass (split mode): one 1024-thread workgroup per
// folded column, the column's values kept in registers, so the FP64 folded
// vectors are never written; columns of up to kFoldSplitMax rows.
constexpr int kFoldSplitMax = 32768;
struct FoldSplitJob {
    FoldJob f;                       // the fold of launch_fold_jobs (xt unused)
    int Kp;
    int8_t *D;                       // digits of column c at D + c * Kp (plane 0)
    long long plane;
    int *E;                          // exponent per column
};
// maxrows: the largest Kp of the jobs (<= kFoldSplitMax); at most
// kFoldSplitSmall rows run at two workgroups per CU.
constexpr int kFoldSplitSmall = 12288;
void launch_fold_split_jobs(hipStream_t st, int S, double R, const FoldSplitJob *d_jobs,
                            const int *d_begin, int njobs, int ncols, int maxrows);
// The intensity steps of every chunk of a round: blocks of a job = ceil(ncand / 256).
void launch_intensity_jobs(hipStream_t st, const IntensArgs *d_jobs, const int *d_begin,
                           int njobs, int nblocks);
void launch_corr_intensity(hipStream_t st, const CorrArgs &a);

// One round of the FP64 batched mode after its G build in one persistent
// launch: workgroups take tasks in ticket order from one counter, stage by
// stage (every half-ls tile of the round, then every fold block, correlation
// tile and intensity block, each stage in chunk order). A task of stage s > 0
// waits until every stage s - 1 task of its chunk has signalled, so a chunk's
// fold starts while later chunks' products still run and no stage waits for
// the whole round's previous stage. Every task depends only on tasks with
// lower tickets, which were taken by running workgroups, so the waits end
// whatever the number of resident workgroups; a wait that exceeds
// kPersistSpinMax polls sets the error word and the round's results are void.
struct PersistRound {
    MuSrc src;
    const HalfJob *hj;
    const int *hbeg, *hch;           // first tile and round-local chunk of each half-ls job
    const FoldJob *fj;
    const int *fbeg, *fch;
    const CorrMMArgs *cj;            // one per chunk of the round
    const int *cbeg;
    const IntensArgs *ij;
    const int *ibeg, *ich;
    const int *need;                 // [3][nch]: half-ls tiles, fold blocks, correlation tiles
    int nhj, nfj, ncj, nij;          // jobs per stage
    int nh, nf, nc, ni;              // tasks per stage
    int nch;
    int *sync;                       // [3][nch] done counts, then the ticket; zeroed per launch
    int *err;                        // set when a wait gave up (never cleared by the kernel)
};
constexpr long long kPersistSpinMax = 1ll << 22;  // polls (a few seconds)
// hcfg: half-ls tile as tile_rows/tile_cols (4 wavefronts; the 8-wavefront
// setting runs 128 x 128 on 4); ccfg: correlation tile (0: 64 x 64, 2: 128 x 64).
// Returns the grid used (workgroups resident on the device, at most the tasks).
int launch_round_persist(hipStream_t st, int hcfg, int ccfg, const PersistRound &p, int ncu);
// n ints set to 0 by a kernel (no memset node in a captured graph)
void launch_zero_ints(hipStream_t st, int *p, int n);
// whether launch_round_persist has a kernel for the correlation tile setting
inline bool persist_corr_cfg_ok(int ccfg) { return ccfg == 0 || ccfg == 2; }
int compact_blocks(long long n);
void launch_compact(hipStream_t st, long long n, const int *flag, int *counts, long long *offs,
                    LineSoA in, LineSoA out);

}  // namespace gi
