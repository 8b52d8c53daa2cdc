// The hot-path driver: the lower-level loop of cuda_host.cu:1174-1488 as
// batched launches over every requested lower level.
//
// Work is split into items (lower level, J'): the upper levels an item
// correlates with are a contiguous range of the energy-sorted upper block
// (frequency window). Items sharing (J, symmetry, J') form groups, cut into
// chunks of at most kChunk items. Per chunk: the symmetry-blocked half
// linestrength (G build + product per output class), the fold into the upper
// block's symmetrised space, then the correlation and intensity.
//   GI_MODE_FP64: one GEMV-form product per unit (k_gemv_sym) and one
//     wavefront per candidate for the correlation (k_corr_intensity).
//   GI_MODE_FP64_BATCHED: the products of a whole chunk on the matrix cores
//     (k_halfls_mfma), the folded vectors of every chunk kept resident, one
//     banded MFMA product for all correlations (k_corr_mfma), then
//     k_intensity.
// Both end with a stable compaction of the emitted candidates and one copy
// of the lines to the caller's host arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gi_internal.h"
#include "gi_kernels.h"

namespace gi {

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(c, e_ == hipErrorOutOfMemory ? GI_ENOMEM : GI_EHIP,             \
                        std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Timed launches: events recorded on the context's stream around each launch.
struct Timeline {
    gi_ctx *c;
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> recs;
    explicit Timeline(gi_ctx *c_) : c(c_) {}
    int begin(const char *name)
    {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return GI_EHIP;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return GI_EHIP;
        }
        recs.emplace_back(name, a, b);
        return hipEventRecord(a, c->stream) == hipSuccess ? GI_OK : GI_EHIP;
    }
    int end()
    {
        return hipEventRecord(std::get<2>(recs.back()), c->stream) == hipSuccess ? GI_OK : GI_EHIP;
    }
    void collect()
    {
        c->kstats.clear();
        c->kstat_names.clear();
        for (auto &r : recs) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, std::get<1>(r), std::get<2>(r)) != hipSuccess) ms = 0;
            auto &s = c->kstats[std::get<0>(r)];
            if (s.launches == 0) c->kstat_names.push_back(std::get<0>(r));
            s.ms += ms;
            s.launches += 1;
        }
    }
    ~Timeline()
    {
        for (auto &r : recs) {
            (void)hipEventDestroy(std::get<1>(r));
            (void)hipEventDestroy(std::get<2>(r));
        }
    }
};

#define TL_BEGIN(name)                                                         \
    do {                                                                       \
        if (tl.begin(name) != GI_OK) return fail(c, GI_EHIP, "hipEventRecord"); \
    } while (0)
#define TL_END()                                                               \
    do {                                                                       \
        if (tl.end() != GI_OK) return fail(c, GI_EHIP, "hipEventRecord");      \
    } while (0)

// trove_functions.cpp:1495-1510
static bool degeneracy_filter(const gi_ctx *c, int gI, int gF, int idegI, int idegF)
{
    if (c->degen[gI] == 1 || c->degen[gF] == 1) return true;
    if (!c->reduced) return true;
    if (c->sym_type == 2 || c->sym_type == 3) return idegI == 1 && idegF == 0;
    return true;
}

struct Item {
    int level, jindI, gI, jindF, gF, lo, hi;
    long long out_off;
};

// One item per (lower level, J') whose window in the upper block is not
// empty. The lower-level filters (cuda_host.cu:1194) and the J / symmetry
// part of intensity_filter (trove_functions.cpp:1452-1464) are decided here;
// the frequency window (nu = E' - E in [lo, hi]) is a contiguous range of the
// energy-sorted upper block, found with the same comparisons as the filter.
static void make_items(const gi_ctx *c, int nlower, const int *ids, std::vector<Item> &items,
                       long long &ncand)
{
    items.clear();
    ncand = 0;
    const int n = ids ? nlower : c->nlevels;
    for (int q = 0; q < n; q++) {
        const int L = ids ? ids[q] : q;
        if (L < 0 || L >= c->nlevels) continue;
        if (!c->lv_lowok[L]) continue;
        const int gI = c->lv_igamma[L];
        if (!(c->gns[gI] > 0.0)) continue;
        const int jindI = c->lv_jind[L];
        const int jI = c->jvals[jindI];
        const double eI = c->lv_energy[L];
        const int gF = c->igamma_pair[gI];
        if (c->isym_pairs[gI] != c->isym_pairs[gF]) continue;
        for (int jindF = 0; jindF < 2; jindF++) {
            const int jF = c->jvals[jindF];
            if (!((jF != c->jvals[0]) || (jI != c->jvals[0]))) continue;
            if (std::abs(jF - jI) > 1 || jI + jF < 1) continue;
            const auto &en = c->blocks[jindF * c->nrep + gF].en;
            const double w0 = c->freq_window[0], w1 = c->freq_window[1];
            const int a = (int)(std::partition_point(en.begin(), en.end(),
                                                     [&](double e) { return !(e - eI >= w0); }) -
                                en.begin());
            const int b = (int)(std::partition_point(en.begin(), en.end(),
                                                     [&](double e) { return e - eI <= w1; }) -
                                en.begin());
            if (b <= a) continue;
            items.push_back(Item{L, jindI, gI, jindF, gF, a, b, ncand});
            ncand += b - a;
        }
    }
}

struct Chunk {
    int jindI, gI, jindF, gF, nitems, nunits, ncand, cb_mask, Lo, Hi;
    size_t off_uptr, off_uideg, off_iunit, off_cstart, off_lo, off_hi, off_level, off_eI, off_out;
    size_t ht_off;  // doubles, batched mode
};

// Batched FP64 mode. Chunks are grouped into rounds whose G, H and folded
// vectors fit a memory budget; per round one launch builds every G, one
// launch runs every (chunk, class) product on the matrix cores, the folds run
// per chunk, one launch runs every banded correlation, and the intensity step
// runs per chunk.
static int run_batched(gi_ctx *c, const std::vector<Chunk> &chunks, unsigned char *d_meta,
                       Timeline &tl, double beta, double A_coef_s_1, double intens_cm_mol,
                       int *flag, LineSoA cand, double &useful, double &issued,
                       double &corr_bytes, double &mu_bytes, double &mu_dense, double &corr_flops)
{
    const int N = c->nvib;
    int rc = GI_OK;
    // device memory for one round's G, H and folded vectors (GI_BATCH_BYTES overrides)
    size_t kRoundBytes = (size_t)12 << 30;
    if (const char *env = std::getenv("GI_BATCH_BYTES")) {
        const long long v = std::atoll(env);
        if (v > 0) kRoundBytes = (size_t)v;
    }
    struct Need {
        size_t g, h, ht;
        int njobs;
    };
    std::vector<Need> need(chunks.size());
    for (size_t i = 0; i < chunks.size(); i++) {
        const Chunk &ch = chunks[i];
        const GroupPlan &GP = c->groups[((size_t)ch.jindI * c->nrep + ch.gI) * 2 + ch.jindF];
        const int ncb = c->degen[ch.gI] * c->degen[ch.gF];
        const int nsizeF = c->blocks[ch.jindF * c->nrep + ch.gF].nsize;
        need[i] = Need{(size_t)ch.nunits * GP.Gtot, (size_t)ch.nunits * GP.Htot,
                       (size_t)ch.nitems * ncb * nsizeF, ch.nunits > 0 ? (int)GP.oc.size() : 0};
    }
    std::vector<std::pair<size_t, size_t>> rounds;
    size_t start = 0, acc = 0, mg = 1, mh = 1, mht = 1, sg = 0, sh = 0, sht = 0;
    int mjobs = 1, sjobs = 0, mchunks = 1;
    for (size_t i = 0; i <= chunks.size(); i++) {
        const size_t bytes = i < chunks.size() ? 8 * (need[i].g + need[i].h + need[i].ht) : 0;
        if (i == chunks.size() || (i > start && acc + bytes > kRoundBytes)) {
            rounds.push_back({start, i});
            mg = std::max(mg, sg); mh = std::max(mh, sh); mht = std::max(mht, sht);
            mjobs = std::max(mjobs, sjobs);
            mchunks = std::max(mchunks, (int)(i - start));
            start = i;
            acc = sg = sh = sht = 0;
            sjobs = 0;
        }
        if (i == chunks.size()) break;
        acc += bytes;
        sg += need[i].g; sh += need[i].h; sht += need[i].ht;
        sjobs += need[i].njobs;
    }
    // every buffer is sized for the largest round before anything is launched
    double *G = (double *)scratch(c, "G", mg * 8, &rc); if (rc) return rc;
    double *H = (double *)scratch(c, "H", mh * 8, &rc); if (rc) return rc;
    double *Ht = (double *)scratch(c, "Ht", mht * 8, &rc); if (rc) return rc;
    const size_t jobs_bytes = (size_t)mjobs * sizeof(HalfJob);
    unsigned char *d_jobs = (unsigned char *)scratch(c, "jobs", jobs_bytes + 2 * 4 * (size_t)mjobs, &rc);
    if (rc) return rc;
    CorrMMArgs *d_mm = (CorrMMArgs *)scratch(c, "mmargs", (size_t)mchunks * sizeof(CorrMMArgs), &rc);
    if (rc) return rc;
    std::vector<std::vector<unsigned char>> keep;  // host sources of the async copies

    for (auto &rg : rounds) {
        std::vector<HalfJob> jobs;
        std::vector<int> gbegin, tbegin;
        std::vector<CorrMMArgs> mm;
        std::vector<double *> Hc(rg.second - rg.first), Htc(rg.second - rg.first);
        int gt = 0, tt = 0;
        size_t og = 0, oh = 0, oht = 0;
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const JPlan &pI = c->jp[ch.jindI];
            const GroupPlan &GP = c->groups[((size_t)ch.jindI * c->nrep + ch.gI) * 2 + ch.jindF];
            const int ndegI = c->degen[ch.gI];
            SymDev r0 = symdev(pI.sym[ch.gI * GI_MAXDEG + 0]);
            SymDev r1 = ndegI > 1 ? symdev(pI.sym[ch.gI * GI_MAXDEG + 1]) : r0;
            Hc[i - rg.first] = H + oh;
            Htc[i - rg.first] = Ht + oht;
            if (ch.nunits > 0)
                for (const OutClass &oc : GP.oc) {
                    HalfJob j;
                    j.cd = classdev(oc);
                    j.nunits = ch.nunits; j.Gtot = GP.Gtot; j.Htot = GP.Htot;
                    j.tile_begin = tt;
                    j.mt = (oc.n + 63) / 64;
                    tt += j.mt * ((ch.nunits * oc.smax + 63) / 64);
                    j.gtile_begin = gt;
                    j.gkt = (oc.K + 63) / 64;
                    gt += oc.K > 0 ? j.gkt * ch.nunits : 0;
                    j.uptr = (const double *const *)(d_meta + ch.off_uptr);
                    j.uideg = (const int *)(d_meta + ch.off_uideg);
                    j.r0 = r0; j.r1 = r1;
                    j.G = G + og; j.H = H + oh;
                    jobs.push_back(j);
                    gbegin.push_back(j.gtile_begin);
                    tbegin.push_back(j.tile_begin);
                }
            og += need[i].g; oh += need[i].h; oht += need[i].ht;
        }
        const int njobs = (int)jobs.size();
        if (njobs > 0) {
            keep.emplace_back(jobs_bytes + 2 * 4 * (size_t)mjobs);
            unsigned char *hb = keep.back().data();
            std::memcpy(hb, jobs.data(), njobs * sizeof(HalfJob));
            std::memcpy(hb + jobs_bytes, gbegin.data(), njobs * 4);
            std::memcpy(hb + jobs_bytes + 4 * (size_t)mjobs, tbegin.data(), njobs * 4);
            HIPCHK(hipMemcpyAsync(d_jobs, hb, keep.back().size(), hipMemcpyHostToDevice, c->stream));
            const int *d_gb = (const int *)(d_jobs + jobs_bytes);
            const int *d_tb = (const int *)(d_jobs + jobs_bytes + 4 * (size_t)mjobs);
            TL_BEGIN("build_g");
            launch_halfls_jobs(c->stream, N, (const HalfJob *)d_jobs, d_gb, d_tb, njobs, gt, 0,
                               c->d_mu_perm);
            TL_END();
            TL_BEGIN("half_ls_mfma");
            launch_halfls_jobs(c->stream, N, (const HalfJob *)d_jobs, d_gb, d_tb, njobs, 0, tt,
                               c->d_mu_perm);
            TL_END();
        }
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const JPlan &pF = c->jp[ch.jindF];
            const GroupPlan &GP = c->groups[((size_t)ch.jindI * c->nrep + ch.gI) * 2 + ch.jindF];
            const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
            const int ndegI = c->degen[ch.gI], ndegF = c->degen[ch.gF];
            SymDev t0 = symdev(pF.sym[ch.gF * GI_MAXDEG + 0]);
            t0.tcol = GP.d_tcol[0];
            SymDev t1 = t0;
            if (ndegF > 1) {
                t1 = symdev(pF.sym[ch.gF * GI_MAXDEG + 1]);
                t1.tcol = GP.d_tcol[1];
            }
            TL_BEGIN("fold");
            launch_fold(c->stream, ch.nitems, BF.nsize, ndegI, ndegF, GP.Htot,
                        (const int *)(d_meta + ch.off_iunit), t0, t1, Hc[i - rg.first],
                        Htc[i - rg.first]);
            TL_END();
            CorrMMArgs m;
            m.Lo = ch.Lo; m.Hi = ch.Hi; m.nitems = ch.nitems; m.ncb = ndegI * ndegF;
            m.ndegI = ndegI; m.nsize = BF.nsize;
            m.item_lo = (const int *)(d_meta + ch.off_lo);
            m.item_hi = (const int *)(d_meta + ch.off_hi);
            m.item_out_off = (const long long *)(d_meta + ch.off_out);
            m.blk_col = BF.d_col; m.vec = BF.d_vec; m.Ht = Htc[i - rg.first]; m.amp = cand.amp;
            mm.push_back(m);
            corr_flops += 2.0 * (double)(ch.Hi - ch.Lo) * ch.nitems * ndegI * ndegF * BF.nsize;
            double blk_bytes = 0;
            for (const OutClass &oc : GP.oc) blk_bytes += 8.0 * oc.n * oc.K;
            useful += ch.nunits * 2.0 * c->coef[ch.jindI][ch.jindF].C * (double)N * N;
            issued += ch.nunits * 2.0 * GP.macs;
            mu_bytes += blk_bytes;
            mu_dense += ch.nunits * 24.0 * (double)N * N;
            corr_bytes += (double)ch.ncand * 8.0 * BF.nsize;
        }
        int max_rows = 0, max_cols = 0;
        for (auto &m : mm) {
            max_rows = std::max(max_rows, m.Hi - m.Lo);
            max_cols = std::max(max_cols, m.nitems * m.ncb);
        }
        keep.emplace_back(mm.size() * sizeof(CorrMMArgs));
        std::memcpy(keep.back().data(), mm.data(), keep.back().size());
        HIPCHK(hipMemcpyAsync(d_mm, keep.back().data(), keep.back().size(), hipMemcpyHostToDevice,
                              c->stream));
        TL_BEGIN("corr_mfma");
        launch_corr_mfma(c->stream, d_mm, (int)mm.size(), max_rows, max_cols);
        TL_END();
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
            IntensArgs a;
            a.ncand = ch.ncand; a.nitems = ch.nitems;
            a.ndegI = c->degen[ch.gI]; a.ndegF = c->degen[ch.gF]; a.cb_mask = ch.cb_mask;
            a.reduced = c->reduced; a.jI = c->jvals[ch.jindI]; a.jF = c->jvals[ch.jindF];
            a.item_cand_start = (const int *)(d_meta + ch.off_cstart);
            a.item_lo = (const int *)(d_meta + ch.off_lo);
            a.item_level = (const int *)(d_meta + ch.off_level);
            a.item_out_off = (const long long *)(d_meta + ch.off_out);
            a.item_energy = (const double *)(d_meta + ch.off_eI);
            a.blk_lev = BF.d_lev; a.blk_en = BF.d_en; a.blk_upok = BF.d_upok;
            a.gnsI = c->gns[ch.gI]; a.zpe = c->zpe; a.beta = beta; a.q_stat = c->q_stat;
            a.A_coef_s_1 = A_coef_s_1; a.intens_cm_mol = intens_cm_mol;
            a.thresh_S = c->thresh_S; a.thresh_I = c->thresh_I;
            a.amp = cand.amp; a.flag = flag; a.out_iI = cand.iI; a.out_iF = cand.iF;
            a.out_nu = cand.nu; a.out_S = cand.S; a.out_A = cand.A; a.out_I = cand.I;
            TL_BEGIN("intensity");
            launch_intensity(c->stream, a);
            TL_END();
        }
    }
    // the async copies read `keep`; wait before it goes out of scope
    HIPCHK(hipStreamSynchronize(c->stream));
    return GI_OK;
}

int compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out)
{
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_compute before gi_prepare");
    if (mode != GI_MODE_FP64 && mode != GI_MODE_FP64_BATCHED)
        return fail(c, GI_EINVAL, "gi_compute: unknown mode");
    const bool batched = mode == GI_MODE_FP64_BATCHED;
    HIPCHK(hipSetDevice(c->device));
    std::vector<Item> items;
    long long ncand = 0;
    make_items(c, nlower, ids, items, ncand);
    for (double &w : c->work) w = 0;
    c->work[1] = (double)ncand;
    c->kstats.clear();
    c->kstat_names.clear();
    if (ncand == 0) {
        out->n = 0;
        return GI_OK;
    }
    const int N = c->nvib;
    int rc = GI_OK;

    // constants of cuda_host.cu:12-18 and beta of :1170
    const double pi = 4.0 * std::atan2(1.0, 1.0);
    const double planck = 6.62606896 * std::pow(10.0, -27.0);
    const double avogno = 6.0221415 * std::pow(10.0, 23.0);
    const double vellgt = 2.99792458 * std::pow(10.0, 10.0);
    const double A_coef_s_1 =
        64.0 * std::pow(10.0, -36.0) * std::pow(pi, 4.0) / (3.0 * 6.62606896 * std::pow(10.0, -27.0));
    const double intens_cm_mol =
        8.0 * std::pow(10.0, -36.0) * std::pow(pi, 3.0) * avogno / (3.0 * planck * vellgt);
    const double boltz = 1.380658 * std::pow(10.0, -16.0);
    const double beta = planck * vellgt / (boltz * c->temperature);

    // candidate records and compacted output
    const size_t nc = (size_t)ncand;
    int *flag = (int *)scratch(c, "flag", nc * sizeof(int), &rc); if (rc) return rc;
    LineSoA cand, comp;
    cand.iI = (int *)scratch(c, "c_iI", nc * sizeof(int), &rc); if (rc) return rc;
    cand.iF = (int *)scratch(c, "c_iF", nc * sizeof(int), &rc); if (rc) return rc;
    cand.nu = (double *)scratch(c, "c_nu", nc * 8, &rc); if (rc) return rc;
    cand.S = (double *)scratch(c, "c_S", nc * 8, &rc); if (rc) return rc;
    cand.A = (double *)scratch(c, "c_A", nc * 8, &rc); if (rc) return rc;
    cand.I = (double *)scratch(c, "c_I", nc * 8, &rc); if (rc) return rc;
    cand.amp = (double *)scratch(c, "c_amp", nc * 32, &rc); if (rc) return rc;
    comp.iI = (int *)scratch(c, "o_iI", nc * sizeof(int), &rc); if (rc) return rc;
    comp.iF = (int *)scratch(c, "o_iF", nc * sizeof(int), &rc); if (rc) return rc;
    comp.nu = (double *)scratch(c, "o_nu", nc * 8, &rc); if (rc) return rc;
    comp.S = (double *)scratch(c, "o_S", nc * 8, &rc); if (rc) return rc;
    comp.A = (double *)scratch(c, "o_A", nc * 8, &rc); if (rc) return rc;
    comp.I = (double *)scratch(c, "o_I", nc * 8, &rc); if (rc) return rc;
    comp.amp = (double *)scratch(c, "o_amp", nc * 32, &rc); if (rc) return rc;

    // groups and chunks
    const int kChunk = 256;
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (int i = 0; i < (int)items.size(); i++)
        groups[std::make_tuple(items[i].jindI, items[i].gI, items[i].jindF)].push_back(i);
    std::vector<Chunk> chunks;
    std::vector<unsigned char> meta;
    auto put = [&](const void *src, size_t bytes) {
        const size_t o = (meta.size() + 255) & ~size_t(255);
        meta.resize(o + std::max<size_t>(bytes, 8));
        if (bytes) std::memcpy(meta.data() + o, src, bytes);
        return o;
    };
    int max_units = 0, max_items = 0, max_G = 1, max_H = 1, max_ncb = 1, max_nsizeF = 0;
    size_t ht_total = 0;
    for (auto &kv : groups) {
        const int jindI = std::get<0>(kv.first), gI = std::get<1>(kv.first),
                  jindF = std::get<2>(kv.first);
        const int gF = c->igamma_pair[gI];
        const int ndegI = c->degen[gI], ndegF = c->degen[gF];
        const Block &BI = c->blocks[jindI * c->nrep + gI];
        const int nsizeF = c->blocks[jindF * c->nrep + gF].nsize;
        std::vector<int> its = kv.second;
        if (batched)  // energy order keeps every chunk's windows a narrow band
            std::stable_sort(its.begin(), its.end(), [&](int x, int y) {
                return c->lv_energy[items[x].level] < c->lv_energy[items[y].level];
            });
        const GroupPlan &GP = c->groups[((size_t)jindI * c->nrep + gI) * 2 + jindF];
        if (!GP.valid) return fail(c, GI_EINVAL, "internal: no plan for a (J, J') pair");
        for (size_t s0 = 0; s0 < its.size(); s0 += kChunk) {
            const int n_it = (int)std::min<size_t>(kChunk, its.size() - s0);
            std::vector<long long> uptr, outoff(n_it);
            std::vector<int> uideg, iunit(n_it * GI_MAXDEG, -1), cstart(n_it), lo(n_it), hi(n_it),
                lev(n_it);
            std::vector<double> eI(n_it);
            int cnt = 0, Lo = 1 << 30, Hi = 0;
            for (int q = 0; q < n_it; q++) {
                const Item &it = items[its[s0 + q]];
                for (int ideg = 0; ideg < ndegI; ideg++) {
                    if (!degeneracy_filter(c, gI, gF, ideg, 0)) continue;
                    iunit[q * GI_MAXDEG + ideg] = (int)uideg.size();
                    uideg.push_back(ideg);
                    uptr.push_back((long long)c->lv_irec0[it.level] * BI.nsize);
                }
                cstart[q] = cnt;
                lo[q] = it.lo;
                hi[q] = it.hi;
                Lo = std::min(Lo, it.lo);
                Hi = std::max(Hi, it.hi);
                lev[q] = it.level;
                eI[q] = c->lv_energy[it.level];
                outoff[q] = it.out_off;
                cnt += it.hi - it.lo;
            }
            int mask = 0;
            for (int idegF = 0; idegF < ndegF; idegF++)
                for (int idegI = 0; idegI < ndegI; idegI++) {
                    if (c->reduced && idegF != 0) continue;
                    if (!degeneracy_filter(c, gI, gF, idegI, idegF)) continue;
                    if (!degeneracy_filter(c, gI, gF, idegI, 0)) continue;  // unit exists
                    mask |= 1 << (idegF * ndegI + idegI);
                }
            Chunk ch;
            ch.jindI = jindI; ch.gI = gI; ch.jindF = jindF; ch.gF = gF;
            ch.nitems = n_it; ch.nunits = (int)uideg.size(); ch.ncand = cnt; ch.cb_mask = mask;
            ch.Lo = Lo; ch.Hi = Hi;
            ch.off_uptr = put(uptr.data(), uptr.size() * 8);
            ch.off_uideg = put(uideg.data(), uideg.size() * 4);
            ch.off_iunit = put(iunit.data(), iunit.size() * 4);
            ch.off_cstart = put(cstart.data(), cstart.size() * 4);
            ch.off_lo = put(lo.data(), lo.size() * 4);
            ch.off_hi = put(hi.data(), hi.size() * 4);
            ch.off_level = put(lev.data(), lev.size() * 4);
            ch.off_eI = put(eI.data(), eI.size() * 8);
            ch.off_out = put(outoff.data(), outoff.size() * 8);
            ch.ht_off = batched ? ht_total : 0;
            ht_total += (size_t)n_it * ndegI * ndegF * nsizeF;
            chunks.push_back(ch);
            max_units = std::max(max_units, ch.nunits);
            max_items = std::max(max_items, n_it);
            max_G = std::max(max_G, GP.Gtot);
            max_H = std::max(max_H, GP.Htot);
            max_ncb = std::max(max_ncb, ndegI * ndegF);
            max_nsizeF = std::max(max_nsizeF, nsizeF);
        }
    }
    // unit pointers: block offsets -> device addresses
    for (auto &ch : chunks) {
        long long *p = (long long *)(meta.data() + ch.off_uptr);
        const double *base = c->blocks[ch.jindI * c->nrep + ch.gI].d_vec;
        for (int u = 0; u < ch.nunits; u++) p[u] = (long long)(base + p[u]);
    }
    unsigned char *d_meta = (unsigned char *)scratch(c, "meta", meta.size(), &rc); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_meta, meta.data(), meta.size(), hipMemcpyHostToDevice, c->stream));
    (void)ht_total;
    Timeline tl(c);
    double useful = 0, issued = 0, corr_bytes = 0, mu_bytes = 0, mu_dense = 0, corr_flops = 0;
    if (batched) {
        if ((rc = run_batched(c, chunks, d_meta, tl, beta, A_coef_s_1, intens_cm_mol, flag, cand,
                              useful, issued, corr_bytes, mu_bytes, mu_dense, corr_flops)) != GI_OK)
            return rc;
    }
    double *G = nullptr, *H = nullptr, *Ht = nullptr;
    if (!batched) {
        G = (double *)scratch(c, "G", (size_t)max_units * max_G * 8, &rc); if (rc) return rc;
        H = (double *)scratch(c, "H", (size_t)max_units * max_H * 8, &rc); if (rc) return rc;
        Ht = (double *)scratch(c, "Ht", (size_t)max_items * max_ncb * max_nsizeF * 8, &rc);
        if (rc) return rc;
    }
    for (auto &ch : chunks) {
        if (batched) break;
        const JPlan &pI = c->jp[ch.jindI], &pF = c->jp[ch.jindF];
        const GroupPlan &GP = c->groups[((size_t)ch.jindI * c->nrep + ch.gI) * 2 + ch.jindF];
        const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
        const int ndegI = c->degen[ch.gI], ndegF = c->degen[ch.gF];
        SymDev r0 = symdev(pI.sym[ch.gI * GI_MAXDEG + 0]);
        SymDev r1 = ndegI > 1 ? symdev(pI.sym[ch.gI * GI_MAXDEG + 1]) : r0;
        SymDev t0 = symdev(pF.sym[ch.gF * GI_MAXDEG + 0]);
        t0.tcol = GP.d_tcol[0];
        SymDev t1 = t0;
        if (ndegF > 1) {
            t1 = symdev(pF.sym[ch.gF * GI_MAXDEG + 1]);
            t1.tcol = GP.d_tcol[1];
        }
        if (ch.nunits > 0) {
            for (const OutClass &oc : GP.oc) {
                const ClassDev cd = classdev(oc);
                if (cd.K > 0) {
                    TL_BEGIN("build_g");
                    launch_build_g_sym(c->stream, ch.nunits, N, GP.Gtot, cd,
                                       (const double *const *)(d_meta + ch.off_uptr),
                                       (const int *)(d_meta + ch.off_uideg), r0, r1, G);
                    TL_END();
                }
                TL_BEGIN("half_ls_gemv");
                if (launch_gemv_sym(c->stream, ch.nunits, N, GP.Gtot, GP.Htot, cd,
                                    c->d_mu_perm, G, H) != 0)
                    return fail(c, GI_EINVAL, "no half-ls kernel for this column count");
                TL_END();
            }
        }
        double *Htc = Ht;
        TL_BEGIN("fold");
        launch_fold(c->stream, ch.nitems, BF.nsize, ndegI, ndegF, GP.Htot,
                    (const int *)(d_meta + ch.off_iunit), t0, t1, H, Htc);
        TL_END();
        {
            CorrArgs a;
            a.ncand = ch.ncand; a.nitems = ch.nitems; a.nsizeF = BF.nsize;
            a.ndegI = ndegI; a.ndegF = ndegF; a.cb_mask = ch.cb_mask; a.reduced = c->reduced;
            a.jI = c->jvals[ch.jindI]; a.jF = c->jvals[ch.jindF];
            a.item_cand_start = (const int *)(d_meta + ch.off_cstart);
            a.item_lo = (const int *)(d_meta + ch.off_lo);
            a.item_level = (const int *)(d_meta + ch.off_level);
            a.item_out_off = (const long long *)(d_meta + ch.off_out);
            a.item_energy = (const double *)(d_meta + ch.off_eI);
            a.blk_col = BF.d_col; a.blk_lev = BF.d_lev; a.blk_en = BF.d_en; a.blk_upok = BF.d_upok;
            a.vec = BF.d_vec; a.Ht = Htc;
            a.gnsI = c->gns[ch.gI]; a.zpe = c->zpe; a.beta = beta; a.q_stat = c->q_stat;
            a.A_coef_s_1 = A_coef_s_1; a.intens_cm_mol = intens_cm_mol;
            a.thresh_S = c->thresh_S; a.thresh_I = c->thresh_I;
            a.flag = flag; a.out_iI = cand.iI; a.out_iF = cand.iF; a.out_nu = cand.nu;
            a.out_S = cand.S; a.out_A = cand.A; a.out_I = cand.I; a.out_amp = cand.amp;
            TL_BEGIN("corr_intensity");
            launch_corr_intensity(c->stream, a);
            TL_END();
            corr_flops += 2.0 * ch.ncand * ndegI * ndegF * BF.nsize;
        }
        double blk_bytes = 0;
        for (const OutClass &oc : GP.oc) blk_bytes += 8.0 * oc.n * oc.K;
        useful += ch.nunits * 2.0 * c->coef[ch.jindI][ch.jindF].C * (double)N * N;
        issued += ch.nunits * 2.0 * GP.macs;
        mu_bytes += ch.nunits * blk_bytes;
        mu_dense += ch.nunits * 24.0 * (double)N * N;
        corr_bytes += (double)ch.ncand * 8.0 * BF.nsize;
    }
    HIPCHK(hipGetLastError());
    const int nb = compact_blocks(ncand);
    int *counts = (int *)scratch(c, "counts", (size_t)nb * 4, &rc); if (rc) return rc;
    long long *offs = (long long *)scratch(c, "offs", (size_t)(nb + 1) * 8, &rc); if (rc) return rc;
    TL_BEGIN("compact");
    launch_compact(c->stream, ncand, flag, counts, offs, cand, comp);
    TL_END();
    HIPCHK(hipGetLastError());
    long long nout = 0;
    HIPCHK(hipMemcpyAsync(&nout, offs + nb, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    tl.collect();
    c->work[0] = 0;
    for (auto &ch : chunks) c->work[0] += ch.nunits;
    c->work[2] = (double)nout;
    c->work[3] = useful;
    c->work[4] = issued;
    c->work[5] = corr_bytes;
    c->work[6] = mu_bytes;
    c->work[7] = mu_dense;
    c->work[8] = corr_flops;
    out->n = nout;
    if (nout > out->capacity) return fail(c, GI_ECAPACITY, "gi_compute: output capacity too small");
    if (nout > 0) {
        if (!out->ilevelI || !out->ilevelF || !out->nu || !out->S || !out->A_einst || !out->I ||
            !out->amp)
            return fail(c, GI_EINVAL, "gi_compute: output arrays missing");
        HIPCHK(hipMemcpyAsync(out->ilevelI, comp.iI, nout * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->ilevelF, comp.iF, nout * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->nu, comp.nu, nout * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->S, comp.S, nout * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->A_einst, comp.A, nout * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->I, comp.I, nout * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(out->amp, comp.amp, nout * 32, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return GI_OK;
}

}  // namespace gi

extern "C" {

int gi_count_candidates(gi_ctx *c, int nlower, const int *ids, long long *count)
{
    if (!c || !count) return GI_EINVAL;
    if (!c->prepared) return gi::fail(c, GI_ESTATE, "gi_count_candidates before gi_prepare");
    std::vector<gi::Item> items;
    long long n = 0;
    gi::make_items(c, nlower, ids, items, n);
    *count = n;
    return GI_OK;
}

int gi_compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out)
{
    if (!c || !out) return GI_EINVAL;
    return gi::compute(c, mode, nlower, ids, out);
}

}  // extern "C"
