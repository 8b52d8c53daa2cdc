// Internal declarations of the gfx950 line-intensity engine.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/gpu_intensities.h"
#include "gi_kernels.h"

#define GI_MAXDEG 2

namespace gi {
// Largest J list of a context (gi_set_jvals).
constexpr int kMaxJ = 128;
// Largest pipeline round count (gi_set_tuning).
constexpr int kMaxPipeRounds = 64;
// Most split-K parts of a correlation product.
constexpr int kMaxKParts = 16;
// tile_halfls setting: per round, 128 x 128 tiles (eight wavefronts) when
// they fill the GPU for at least kAutoTileWaves waves, else 64 x 64 (the same
// bits either way)
constexpr int kTileAuto = 4;
// largest half-ls tile setting (5: 128 x 128 on 8 wavefronts, gi_kernels.h)
constexpr int kTileHalflsMax = 5;
constexpr int kAutoTileWaves = 3;
// tile_split setting: per round, 128 x 64 on eight wavefronts with two K steps
// in flight when the round's 64 x 64 tiles average at least
// kAutoSplitSteps K steps of 64, else 64 x 64 with two steps in flight (the
// same bits either way)
constexpr int kSplitTileAuto = 4;
constexpr int kAutoSplitSteps = 40;
// Counters of gi_work_stats.
constexpr int kWorkStats = 17;
}  // namespace gi

namespace gi {

// Symmetrisation map of one (J, symmetry, component): the sparse matrix R of
// the correlate step (dipole_kernals.cu:38-71) over the slot grid, and its
// transpose for the fold of half linestrengths into symmetrised space.
struct SymMap {
    int nsize = 0;
    std::vector<int> rptr, rcol;     // grid row -> (symmetrised index, value)
    std::vector<double> rval;
    std::vector<int> tptr, tcol;     // symmetrised index -> (grid row, value)
    std::vector<double> tval;
    int *d_rptr = nullptr, *d_rcol = nullptr, *d_tptr = nullptr, *d_tcol = nullptr;
    double *d_rval = nullptr, *d_tval = nullptr;
};

// Slot grid of one J: every primitive r is a (slot, v) cell, where the slot
// carries the rotational labels (K, tau) and v is the J=0 vibrational index.
// Primitives sharing (v, K, tau) get separate slots with the same labels.
struct JPlan {
    int J = 0, dimen = 0, nslots = 0;
    std::vector<int> slot_k, slot_tau, r_slot, r_v;
    std::vector<SymMap> sym;         // [g*GI_MAXDEG + ideg]
    int *d_r_grid = nullptr;         // [dimen]: slot*N + v of primitive r
};

// Rotational coefficients of the half linestrength between the slots of two
// J values: for every (slot_F, component a) the list of (slot_I, coefficient)
// with coefficient = sgnF * sgnI * L * 3j * s (dipole_kernals.cu:179-225).
struct Coef {
    bool valid = false;
    int SI = 0, SF = 0;
    std::vector<int> ptr, sI;        // ptr over sF*3 + a
    std::vector<double> val;
    int *d_ptr = nullptr, *d_sI = nullptr;
    double *d_val = nullptr;
    double P = 0, C = 0;             // useful pair and component counts per Nvib^2
};

// One output class of a group plan: the vibrational indices [off, off+n) of
// the permuted grid (one symmetry class) and the upper slots whose needed
// cells lie in it. Its half linestrength block is H_co = sum over segments e
// of mu_{a(e)}[co, ci(e)] * G_co[e], with G_co[e][vI][j] =
// sum_{(sI, w) in recipe(e, j)} w * c[sI, vI].
struct OutClass {
    int off = 0, n = 0;              // vibrational range (permuted indices)
    std::vector<int> cols;           // upper slots, ascending
    int smax = 0;                    // padded column count
    std::vector<int> seg_a, seg_off, seg_n, seg_row0;  // K segments (component, class range)
    int K = 0;
    std::vector<int> rptr, rsI;      // recipe CSR over (segment, column)
    std::vector<double> rcoef;
    int hbase = 0, gbase = 0;        // offsets (doubles) in a unit's H and G buffers
    int gbase_p = 0;                 // offset in the packed G (rows of ncols, batched mode)
    // split mode: G digits column-major, each segment padded to 64 rows
    int kpad = 0, cbaseB = 0, eoff = 0;  // padded K, offset in a unit's digits, exponent offset
    std::vector<int> seg_pad;        // per segment: (row offset in a column, padded mu column)
    int *d_segpad = nullptr;
    int *d_ints = nullptr;           // packed device copy: seg_a, seg_off, seg_n, seg_row0, rptr, rsI
    double *d_rcoef = nullptr;
};

// Plan of the symmetry-blocked half linestrength for lower block (J, gI) and
// upper J' (the upper symmetry is igamma_pair[gI]).
struct GroupPlan {
    bool valid = false;
    int jindI = 0, gI = 0, jindF = 0, gF = 0;
    std::vector<OutClass> oc;
    int Htot = 0, Gtot = 0, smax_all = 0;
    int Gtot_p = 0;                  // packed G per unit (batched mode)
    double macs = 0;                 // multiply-adds per unit
    std::vector<int> tcol[GI_MAXDEG];    // fold map: symmetrised index -> H offset
    int *d_tcol[GI_MAXDEG] = {nullptr, nullptr};
    // the fold map's rows as fixed pairs (SymDev::tfc, tfw; empty when a row
    // has more than two entries)
    std::vector<int> tfc[GI_MAXDEG];
    std::vector<double> tfw[GI_MAXDEG];
    int *d_tfc[GI_MAXDEG] = {nullptr, nullptr};
    double *d_tfw[GI_MAXDEG] = {nullptr, nullptr};
    ClassDev *d_cd = nullptr;        // device copy of classdev() of every output class
    // packed G as a sparse map of the lower eigenvector record, per lower
    // component d: G[g] = sum_{w < gW} gwt[d][w*Gtot_p + g] * u[gidx[d][w*Gtot_p + g]]
    // (the recipe of g composed with the symmetrisation map; padding has weight 0)
    // the same map in the split-mode layout (columns of kpad rows), and per
    // column the bound max_rows sum_w |w| (times max|u| it bounds |G|)
    int GtotB = 0, EGtot = 0;
    std::vector<int> gidxB[GI_MAXDEG];
    std::vector<double> gwtB[GI_MAXDEG], gbnd[GI_MAXDEG];
    int *d_gidxB[GI_MAXDEG] = {nullptr, nullptr};
    double *d_gwtB[GI_MAXDEG] = {nullptr, nullptr}, *d_gbnd[GI_MAXDEG] = {nullptr, nullptr};
    int gW = 0;
    // compact form of the packed (FP64) layout of the map (build_compact_map):
    // 16-bit record index and 8-bit weight code per entry, kCmapCodes weights
    // per component
    bool cmap = false;
    std::vector<uint16_t> cidx[GI_MAXDEG];
    std::vector<uint8_t> ccode[GI_MAXDEG];
    std::vector<double> cwt[GI_MAXDEG];
    uint16_t *d_cidx[GI_MAXDEG] = {nullptr, nullptr};
    uint8_t *d_ccode[GI_MAXDEG] = {nullptr, nullptr};
    double *d_cwt[GI_MAXDEG] = {nullptr, nullptr};
    std::vector<int> gidx[GI_MAXDEG];
    std::vector<double> gwt[GI_MAXDEG];
    int *d_gidx[GI_MAXDEG] = {nullptr, nullptr};
    double *d_gwt[GI_MAXDEG] = {nullptr, nullptr};
};

// Resident eigenvector block of one (J, symmetry) and its levels in energy order.
struct Block {
    int nsize = 0, nrec = 0;
    const double *d_vec = nullptr;
    bool owned = false;
    // beyond the vector budget (gi_set_vector_budget): the block stays in
    // pinned host memory and the FP64 modes copy the records a chunk needs
    double *h_vec = nullptr;
    std::vector<int> lev;            // level ids, energy order
    std::vector<double> en;
    std::vector<unsigned char> upok; // energy_filter_upper per level
    int *d_lev = nullptr, *d_col = nullptr;
    double *d_en = nullptr;
    unsigned char *d_upok = nullptr;
    // split mode: digit planes of every record (rows of Kp bytes), exponents, max |u|
    int8_t *d_ud = nullptr;
    long long uplane = 0;
    int Kp = 0;
    int *d_ue = nullptr;
    double *d_umax = nullptr;
    // a block in host memory keeps its digit planes and exponents in pinned
    // host memory too (the device keeps d_ue and d_umax)
    int8_t *h_ud = nullptr;
    int *h_ue = nullptr;
};

struct KStat {
    double ms = 0;
    int launches = 0;
};

// Pinned host memory handed out in pieces for the uploads of one launch
// sequence (job descriptors, chunk metadata): the copies are asynchronous and
// can be recorded in a HIP graph. Pieces stay valid until reset(), which the
// owner calls once the sequence has run.
struct HostArena {
    std::vector<std::pair<unsigned char *, size_t>> blocks;
    size_t cur = 0, off = 0;
    unsigned char *alloc(size_t n);  // nullptr if pinned memory runs out
    void reset() { cur = 0; off = 0; }
    void release();
};

// An execution slot: a stream with its own device scratch buffers and pinned
// host staging. The context's main slot runs gi_compute; the request server
// alternates between two more so that one batch drains while the next runs;
// each sweep (gi_sweep_create) owns one.
struct Slot {
    hipStream_t stream = nullptr;
    hipStream_t cstream = nullptr;   // device-to-host copies that overlap `stream`
    hipEvent_t mu_ev[4] = {};        // streamed mu: part ready / part free, two buffers
    hipEvent_t vec_ev[4] = {};       // streamed blocks: records staged / staging free, two buffers
    std::map<std::string, std::pair<void *, size_t>> buf;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    HostArena arena;
};

struct Server;

}  // namespace gi

struct gi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // parameters
    bool have_params = false;
    int nrep = 0, sym_type = 0, nmodes = 0, reduced = 0;
    std::vector<int> degen, isym_pairs, igamma_pair, quanta_lower, quanta_upper;
    // largest degeneracy m: lines carry m*m amplitudes, entry idegI + m*idegF
    int amp_m = 0;
    std::vector<double> gns;
    double zpe = 0, temperature = 0, q_stat = 1;
    double freq_window[2] = {0, 0}, erange_lower[2] = {0, 0}, erange_upper[2] = {0, 0};
    // J values of the context: the deck's two (jvals[0] is the one of the Q-branch
    // exclusion), or an ascending list set by gi_set_jvals; jind indexes it
    std::vector<int> jvals = {0, 0};
    int nJ = 2;
    double thresh_S = 0, thresh_I = 0;

    // inputs
    std::vector<char> have_bset = {0, 0};
    struct HostBset {
        int jval = 0, ncases = 0, dimen = 0, mat_size = 0;
        std::vector<int> icontr2icase, v, k, ktau, N, Ntotal, ijterms, nsize;
        std::vector<std::vector<double>> repres;
    };
    std::vector<HostBset> hb = std::vector<HostBset>(2);
    int nvib = 0;
    double *d_mu = nullptr;
    bool mu_owned = false;
    // streamed mu (gi_set_mu_budget): mu larger than the budget stays on the
    // host; gi_prepare writes its permuted copy to pinned memory and the plain
    // FP64 half linestrength copies it to the device in parts of mu_np rows
    long long mu_budget = 0;         // device bytes for mu; 0: no limit
    bool mu_streamed = false;        // set by gi_upload_dipole
    std::vector<double> h_mu;        // mu as uploaded (streamed mode only)
    double *h_mu_perm = nullptr;     // pinned, permuted (streamed mode only)
    int mu_np = 0;                   // rows (vF) per part
    // eigenvector blocks beyond a device budget (gi_set_vector_budget): device
    // bytes of the blocks uploaded from host memory, the budget (0: none), and
    // an identity map 0..iota_n-1 that addresses staged upper records
    long long vec_budget = 0, vec_dev_bytes = 0;
    int *d_iota = nullptr;
    int iota_n = 0;
    int nlevels = 0;
    std::vector<double> lv_energy;
    std::vector<int> lv_jind, lv_igamma, lv_ndeg, lv_irec0, lv_quanta;
    std::vector<unsigned char> lv_lowok, lv_upok;
    std::vector<int> lv_low;          // ascending ids of the levels with lv_lowok
    std::vector<gi::Block> blocks;   // [jind*nrep + g]

    // plan
    bool prepared = false;
    int jmax = 0;
    std::vector<double> threej;
    std::vector<gi::JPlan> jp;       // [jind]
    std::vector<gi::Coef> coef;      // [jindI*nJ + jindF]
    std::vector<int> vperm, vpos;    // vibrational permutation (new -> old, old -> new)
    std::vector<int> vcls_off, vcls_n;   // symmetry classes: contiguous ranges of new indices
    std::vector<int> vpad_off;       // class offsets with every class padded to 64 (split mode)
    int kpad_all = 0;
    double *d_mu_perm = nullptr;     // mu with rows and columns permuted by vperm
    std::vector<gi::GroupPlan> groups;   // [(jindI*nrep + gI)*nJ + jindF]

    // main execution slot (stream = the context's stream), grown on demand
    gi::Slot main_slot;
    gi::Server *server = nullptr;
    std::vector<gi_sweep *> sweeps;  // live sweeps (they hold pointers into the plan)
    // partition-function inputs on the device: energies and gns*ndeg*(2J+1)
    double *d_lv_energy = nullptr, *d_lv_weight = nullptr;

    // tile configurations of k_halfls_jobs and k_corr_mfma (see tile_rows), and
    // the rounds of a pipelined gi_compute (gi_set_tuning, gi_autotune)
    int tile_halfls = gi::kTileAuto, tile_corr = 0;
    int ncu = 256;                   // compute units of the device (workgroup slots)
    int pipe_rounds = 4;
    int pipe_share = 50;             // percent of the remaining work per round but the last
    // split half-ls: 0 = 64 rows, 1 = 128, 2 = 64 with 2 steps in flight, 3 = 128
    // with 2, 4 = per round 3 or 2 (kSplitTileAuto)
    int tile_split = gi::kSplitTileAuto;
    int persist = 0;                 // FP64 batched rounds in one persistent launch (gi_set_persist)
    // routing thresholds of the batched kernels (gi_set_path_limits; the
    // defaults are the kernels' own limits): split-K part length of the
    // correlation, longest record staged in LDS by the G builds, step and
    // segment limits of k_halfls_split2, longest column of the fused fold-split
    int lim_kpart = 1024, lim_kpart_split = 2048, lim_rec_lds = gi::kRecLds,
        lim_split2_steps = gi::kSplitMaxSteps, lim_split2_segs = gi::kSplitMaxSegs,
        lim_fold_split = gi::kFoldSplitMax;

    // split mode. Requested format (0 = automatic): radix 128 or 254, digits of
    // the correlation (eigenvector blocks, folded vectors) and of the half-ls
    // (mu, G) products, 3..6. Format of the resident digit planes, chosen by
    // prepare_split: split_radix, split_S, split_Sh. split_ready = split_key()
    // of the request the planes were built for (0: none).
    int split_req_radix = 0, split_req_S = 0, split_req_Sh = 0;
    double split_radix = 254.0;
    int split_S = 5, split_Sh = 5, split_ready = 0;
    int split_key() const { return 1 + split_req_radix * 64 + split_req_S * 8 + split_req_Sh; }
    int8_t *d_mud = nullptr;
    int8_t *h_mud = nullptr;         // streamed mu: the digit planes in pinned host memory
    long long muplane = 0;
    int *d_mue = nullptr;

    // statistics of the last gi_compute
    std::map<std::string, gi::KStat> kstats;
    std::vector<std::string> kstat_names;
    double work[gi::kWorkStats] = {};
};

namespace gi {

// Records msg as the context's last error and returns code. On a thread that
// set err_sink (the request server's worker) the message goes there instead,
// so the worker never writes the context's error string.
int fail(gi_ctx *c, int code, const std::string &msg);
extern thread_local std::string *err_sink;

// Plan tables indexed by positions in the context's J list.
inline Coef &coef_of(gi_ctx *c, int jindI, int jindF) { return c->coef[(size_t)jindI * c->nJ + jindF]; }
inline const Coef &coef_of(const gi_ctx *c, int jindI, int jindF)
{
    return c->coef[(size_t)jindI * c->nJ + jindF];
}
inline GroupPlan &group_of(gi_ctx *c, int jindI, int gI, int jindF)
{
    return c->groups[((size_t)jindI * c->nrep + gI) * c->nJ + jindF];
}
inline const GroupPlan &group_of(const gi_ctx *c, int jindI, int gI, int jindF)
{
    return c->groups[((size_t)jindI * c->nrep + gI) * c->nJ + jindF];
}

// Split-K parts of a correlation product of length K (a function of the
// upper block and the mode only, so a candidate's sums never depend on its
// batch): parts of about 1024 (FP64) or 2048 (split digits) K, at most 16.
inline int corr_parts(const gi_ctx *c, int K, bool split)
{
    return std::min(kMaxKParts, std::max(1, K / (split ? c->lim_kpart_split : c->lim_kpart)));
}

// Register-tile width (columns) of the half-ls kernels for `cols` columns.
inline int smax_for(int cols)
{
    const int opts[] = {8, 16, 24, 32, 48, 64};
    for (int o : opts)
        if (cols <= o) return o;
    return -1;
}
void *scratch(gi_ctx *c, const char *name, size_t bytes, int *rc);
void *slot_scratch(gi_ctx *c, Slot &s, const char *name, size_t bytes, int *rc);
void *slot_pinned(gi_ctx *c, Slot &s, size_t bytes, int *rc);
void free_slot(Slot &s);
void server_shutdown(gi_ctx *c);
void sweeps_shutdown(gi_ctx *c);
int partition_function(gi_ctx *c, Slot &S, double T, double *q);
// Looks up the no-allocation scratch array's address on `device` (current).
void noalloc_init(int device);

inline SymDev symdev(const SymMap &m)
{
    return SymDev{m.d_rptr, m.d_rcol, m.d_rval, m.d_tptr, m.d_tcol, m.d_tval};
}

// Device view of output class oc (its packed integer arrays: seg_a, seg_off,
// seg_n, seg_row0[nseg+1], rptr[nseg*ncols+1], rsI).
inline ClassDev classdev(const OutClass &oc)
{
    ClassDev d;
    d.off = oc.off; d.n = oc.n; d.ncols = (int)oc.cols.size(); d.smax = oc.smax; d.K = oc.K;
    d.nseg = (int)oc.seg_a.size(); d.hbase = oc.hbase; d.gbase = oc.gbase;
    d.gbase_p = oc.gbase_p;
    d.kpad = oc.kpad; d.cbaseB = oc.cbaseB; d.eoff = oc.eoff;
    d.seg_pad = oc.d_segpad;
    const int ns = d.nseg;
    d.seg_a = oc.d_ints;
    d.seg_off = oc.d_ints + ns;
    d.seg_n = oc.d_ints + 2 * ns;
    d.seg_row0 = oc.d_ints + 3 * ns;
    d.rptr = oc.d_ints + 4 * ns + 1;
    d.rsI = d.rptr + (ns * d.ncols + 1);
    d.rcoef = oc.d_rcoef;
    return d;
}

// Blocks kept in host memory beyond the vector budget.
inline int host_blocks(const gi_ctx *c)
{
    int n = 0;
    for (const Block &b : c->blocks) n += b.h_vec != nullptr;
    return n;
}

int compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out);
int count_job(gi_ctx *c, int nlower, const int *ids, long long *count);
int autotune(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out, int reps,
             gi_tuning *best, char *table, long long cap, long long *len);
int prepare_split(gi_ctx *c);
void free_split(gi_ctx *c);

// host plan construction (gi_plan.cpp)
int assign_slots(gi_ctx *c, int jind);
int build_vib_classes(gi_ctx *c);
int build_symmaps(gi_ctx *c, int jind);
int build_coef(gi_ctx *c, int jindI, int jindF);
int build_group(gi_ctx *c, int jindI, int gI, int jindF);
int build_gmap(gi_ctx *c, GroupPlan &G);
void build_compact_map(gi_ctx *c, GroupPlan &G);
double three_j(int j1, int j2, int j3, int k1, int k2, int k3);
void threej_table(int jmax, double *out);
bool energy_filter_lower(const gi_ctx *c, int J, double e, const int *q);
bool energy_filter_upper(const gi_ctx *c, int J, double e, const int *q);

}  // namespace gi
