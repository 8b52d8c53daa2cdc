// The hot-path driver (the lower-level loop of cuda_host.cu:1174-1488 as
// batched launches) and the request server built on it.
//
// Work is described by queries. A query carries the temperature, partition
// function and thresholds of one job (gi_compute) or one request
// (gi_serve_*), and owns a contiguous range of candidate slots. Its work is
// split into items (lower level, J'): the upper levels an item correlates
// with are a contiguous range of the energy-sorted upper block (frequency
// window). Items sharing (query, J, symmetry, J') form groups, cut into
// chunks of at most kChunk items, so no chunk ever mixes two queries and a
// query's results do not depend on what it is batched with.
//   GI_MODE_FP64: per chunk, the symmetry-blocked half linestrength in GEMV
//     form (k_gemv_sym), the fold, and one wavefront per candidate for the
//     correlation and intensity (k_corr_intensity).
//   GI_MODE_FP64_BATCHED: per round of chunks, one launch builds every G, one
//     launch runs every (chunk, class) product on the matrix cores
//     (k_halfls_jobs), folds per chunk, one banded matrix-core launch for
//     every correlation (k_corr_mfma), then k_intensity per chunk.
// Each query's candidates are compacted on their own, so its lines come back
// contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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

// Timed launches: events recorded on a stream around each launch (only for
// gi_compute; the server runs untimed).
struct Timeline {
    hipStream_t st;
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> recs;
    explicit Timeline(hipStream_t s) : st(s) {}
    int begin(const char *name, hipStream_t on = nullptr)
    {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return GI_EHIP;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return GI_EHIP;
        }
        recs.emplace_back(name, a, b);
        return hipEventRecord(a, on ? on : st) == hipSuccess ? GI_OK : GI_EHIP;
    }
    int end(hipStream_t on = nullptr)
    {
        return hipEventRecord(std::get<2>(recs.back()), on ? on : st) == hipSuccess ? GI_OK
                                                                                      : GI_EHIP;
    }
    void collect(gi_ctx *c)
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

#define TL_BEGIN(name)                                                                    \
    do {                                                                                  \
        if (tl && tl->begin(name) != GI_OK) return fail(c, GI_EHIP, "hipEventRecord");    \
    } while (0)
#define TL_END()                                                                          \
    do {                                                                                  \
        if (tl && tl->end() != GI_OK) return fail(c, GI_EHIP, "hipEventRecord");          \
    } while (0)
#define TL_BEGIN_ON(name, stream)                                                         \
    do {                                                                                  \
        if (tl && tl->begin(name, stream) != GI_OK)                                       \
            return fail(c, GI_EHIP, "hipEventRecord");                                    \
    } while (0)
#define TL_END_ON(stream)                                                                 \
    do {                                                                                  \
        if (tl && tl->end(stream) != GI_OK) return fail(c, GI_EHIP, "hipEventRecord");    \
    } while (0)

// trove_functions.cpp:1495-1510
static bool degeneracy_filter(const gi_ctx *c, int gI, int gF, int idegI, int idegF)
{
    if (c->degen[gI] == 1 || c->degen[gF] == 1) return true;
    if (!c->reduced) return true;
    if (c->sym_type == 2 || c->sym_type == 3) return idegI == 1 && idegF == 0;
    return true;
}

// Constants of cuda_host.cu:12-18, with the reference's expression shapes.
struct Consts {
    double A_coef_s_1, intens_cm_mol, c2;  // c2 = planck*vellgt/boltz (beta = c2/T)
    Consts()
    {
        const double pi = 4.0 * std::atan2(1.0, 1.0);
        const double planck = 6.62606896 * std::pow(10.0, -27.0);
        const double avogno = 6.0221415 * std::pow(10.0, 23.0);
        const double vellgt = 2.99792458 * std::pow(10.0, 10.0);
        const double boltz = 1.380658 * std::pow(10.0, -16.0);
        A_coef_s_1 = 64.0 * std::pow(10.0, -36.0) * std::pow(pi, 4.0) /
                     (3.0 * 6.62606896 * std::pow(10.0, -27.0));
        intens_cm_mol = 8.0 * std::pow(10.0, -36.0) * std::pow(pi, 3.0) * avogno /
                        (3.0 * planck * vellgt);
        c2 = planck * vellgt / boltz;
    }
    // beta = planck * vellgt / (boltz * T), cuda_host.cu:1170
    static double beta(double T)
    {
        const double planck = 6.62606896 * std::pow(10.0, -27.0);
        const double vellgt = 2.99792458 * std::pow(10.0, 10.0);
        const double boltz = 1.380658 * std::pow(10.0, -16.0);
        return planck * vellgt / (boltz * T);
    }
};

struct Query {
    double beta, q_stat, zpe, thS, thI;
    long long cand_begin = 0, cand_end = 0;
    double T = 0;           // temperature, for Q(T) on the device when q_stat <= 0
};

struct Item {
    int level, jindI, gI, jindF, gF, lo, hi, query;
    long long out_off;
};

struct Chunk {
    int query, jindI, gI, jindF, gF, nitems, nunits, ncand, cb_mask, Lo, Hi;
    // blocks kept in host memory: the lower records (distinct, srec[srec0..
    // srec0+nsrec) of the plan) and the upper records of [Lo, Hi) are staged
    bool lhost = false, uhost = false;
    long long srec0 = 0;
    int nsrec = 0;
    size_t off_urec;       // per unit: record of the lower level in its block
    long long cand_begin;  // first candidate slot (slots are assigned in chunk order)
    size_t off_uptr, off_uideg, off_iunit, off_cstart, off_lo, off_hi, off_level, off_eI, off_out;
};

// Candidate range of one batched round whose compacted lines are copied to
// the caller while later rounds compute (gi_compute, batched mode).
struct RoundOut {
    long long cb = 0, ce = 0;
    int nb = 0;
    size_t cbase = 0, obase = 0;
    hipEvent_t ev = nullptr;
};

// Staging buffers of the blocks kept in host memory (see vec_stage_setup).
struct VecStage {
    double *lbuf[2] = {nullptr, nullptr}, *ubuf[2] = {nullptr, nullptr};
    // split mode: the upper digit rows (planes uplane bytes apart) and exponents
    int8_t *ubd[2] = {nullptr, nullptr};
    int *ube[2] = {nullptr, nullptr};
    long long uplane = 0;
    hipStream_t cs = nullptr;
};

// Everything one launch sequence needs; lives until its work has finished.
struct Plan {
    bool pipe = false;                             // per-round output (see RoundOut)
    std::vector<RoundOut> rounds;
    long long *rcounts = nullptr;                  // pinned, lines per round
    ~Plan()
    {
        for (auto &r : rounds)
            if (r.ev) (void)hipEventDestroy(r.ev);
        for (auto g : graphs) (void)hipGraphExecDestroy(g);
    }
    int mode = GI_MODE_FP64;
    std::vector<Query> q;
    std::vector<Item> items;
    long long ncand = 0;
    std::vector<Chunk> chunks;
    std::vector<unsigned char> meta;
    bool vstream = false;                          // some chunk reads a block in host memory
    std::vector<int> srec;                         // staged lower records of such chunks
    VecStage vs;
    LineSoA comp{};                                // compacted lines (slot buffers)
    LineSoA cand{};                                // candidate lines (their own compaction
                                                   // when every candidate passes)
    double *ampx = nullptr;                        // correlation parts beyond the first
    long long ampx_stride = 0;
    // sweep capture: each round's launches become one graph instead of running;
    // the intensity step then reads (beta, Q, thresholds) from qpar
    bool capture = false;
    std::vector<hipGraphExec_t> graphs;
    const double *qpar = nullptr;
    // requests: (beta, Q, thresholds) of every query in device memory, Q(T)
    // reduced there for the queries with q_stat <= 0 (no host wait)
    bool device_q = false;
    double *qdev = nullptr;
    long long *counts = nullptr;                   // pinned, lines per query
    int *perr = nullptr;                           // error word of the persistent rounds
    double work[kWorkStats] = {};
};

// Window [a, b) of nu = E' - E in [w0, w1] in an energy-sorted block, found
// with the comparisons of intensity_filter (trove_functions.cpp:1453-1454).
static void window(const std::vector<double> &en, double eI, double w0, double w1, int &a, int &b)
{
    a = (int)(std::partition_point(en.begin(), en.end(),
                                   [&](double e) { return !(e - eI >= w0); }) -
              en.begin());
    b = (int)(std::partition_point(en.begin(), en.end(), [&](double e) { return e - eI <= w1; }) -
              en.begin());
}

// First index i >= s of a partitioned range with pred(en[i]) false, given
// pred true on [0, s): exponential steps from s, then a binary search in the
// last step. Equals std::partition_point over the whole range.
template <class Pred>
static int gallop(const std::vector<double> &en, int s, Pred pred)
{
    const int n = (int)en.size();
    int lo = s, hi = s, step = 1;
    while (hi < n && pred(en[hi])) {
        lo = hi + 1;
        hi = lo + step;
        step <<= 1;
    }
    hi = std::min(hi, n);
    return (int)(std::partition_point(en.begin() + lo, en.begin() + hi, pred) - en.begin());
}

// window() for a run of calls on one block: while eI does not decrease, both
// window ends only move up, so each search starts from the previous ends.
struct WindowCursor {
    double eI = 0.0;
    int a = 0, b = 0;
    bool valid = false;
    void find(const std::vector<double> &en, double e, double w0, double w1, int &ra, int &rb)
    {
        if (!(valid && e >= eI)) {
            window(en, e, w0, w1, ra, rb);
        } else {
            ra = gallop(en, a, [&](double x) { return !(x - e >= w0); });
            rb = gallop(en, b, [&](double x) { return x - e <= w1; });
        }
        valid = true;
        eI = e;
        a = ra;
        b = rb;
    }
};

// Host threads for building a large job's chunks (built per group and
// joined in order, so the plan is the same for any count); small jobs plan
// on the calling thread. GI_PLAN_THREADS=n forces n (1..64) at any size.
static int plan_threads(size_t work)
{
    if (const char *env = std::getenv("GI_PLAN_THREADS")) {
        const int n = std::atoi(env);
        if (n >= 1 && n <= 64) return n;
    }
    if (work < 16384) return 1;
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(8u, hw ? hw : 1u));
}

// Runs f(t) for t in [0, n) on n threads (the calling thread takes t = 0).
template <class F>
static void run_threads(int n, F f)
{
    std::vector<std::thread> th;
    for (int t = 1; t < n; t++) th.emplace_back(f, t);
    f(0);
    for (auto &x : th) x.join();
}

// Items of lower levels [k0, k1) of the deck job, candidate offsets from 0.
static void items_job_range(const gi_ctx *c, const int *ids, int k0, int k1, int query,
                            std::vector<Item> &items, long long &ncand)
{
    std::vector<WindowCursor> cur(c->blocks.size());
    items.reserve((size_t)(k1 - k0) * std::min(c->nJ, 3));
    for (int k = k0; k < k1; k++) {
        const int L = ids ? ids[k] : c->lv_low[k];
        if (L < 0 || L >= c->nlevels) continue;
        if (!c->lv_lowok[L]) continue;
        const int gI = c->lv_igamma[L];
        if (!(c->gns[gI] > 0.0)) continue;
        const int jindI = c->lv_jind[L];
        const int jI = c->jvals[jindI];
        const double eI = c->lv_energy[L];
        const int gF = c->igamma_pair[gI];
        if (c->isym_pairs[gI] != c->isym_pairs[gF]) continue;
        // |J' - J| <= 1 keeps J' within one list position of J (a J list is
        // strictly ascending; the deck's two J values are both in range)
        const int jlo = std::max(0, jindI - 1), jhi = std::min(c->nJ - 1, jindI + 1);
        for (int jindF = jlo; jindF <= jhi; jindF++) {
            const int jF = c->jvals[jindF];
            if (!((jF != c->jvals[0]) || (jI != c->jvals[0]))) continue;
            if (std::abs(jF - jI) > 1 || jI + jF < 1) continue;
            int a, b;
            const int bF = jindF * c->nrep + gF;
            cur[bF].find(c->blocks[bF].en, eI, c->freq_window[0], c->freq_window[1], a, b);
            if (b <= a) continue;
            items.push_back(Item{L, jindI, gI, jindF, gF, a, b, query, ncand});
            ncand += b - a;
        }
    }
}

// Items of the deck job: the lower-level filters (cuda_host.cu:1194) and the
// J / symmetry part of intensity_filter (trove_functions.cpp:1452-1464), in
// lower-level order (one thread: at 87k items, ranges built on eight
// threads and joined took 2.5 ms against 1.25 ms).
static void items_job(const gi_ctx *c, int nlower, const int *ids, int query, Plan &P)
{
    // every level: only those that pass the lower-state filters, in id order
    const int n = ids ? nlower : (int)c->lv_low.size();
    std::vector<Item> items;
    long long nc = 0;
    items_job_range(c, ids, 0, n, query, items, nc);
    if (P.items.empty()) {
        P.items.swap(items);
    } else {
        for (Item it : items) {
            it.out_off += P.ncand;
            P.items.push_back(it);
        }
    }
    P.ncand += nc;
}

// Items of one request: every lower level of (jindI, gI) that passes the
// job's lower-level filters and has E - ZPE in [elow_lo, elow_hi], each with
// the window nu in [nu_lo, nu_hi] of the upper block (jindF, igamma_pair[gI]).
static void items_request(const gi_ctx *c, const gi_request &r, int query, Plan &P)
{
    const int gF = c->igamma_pair[r.igammaI];
    const int jI = c->jvals[r.jindI], jF = c->jvals[r.jindF];
    if (!(c->gns[r.igammaI] > 0.0) || c->isym_pairs[r.igammaI] != c->isym_pairs[gF]) return;
    if (!((jF != c->jvals[0]) || (jI != c->jvals[0]))) return;
    if (std::abs(jF - jI) > 1 || jI + jF < 1) return;
    const Block &BI = c->blocks[r.jindI * c->nrep + r.igammaI];
    const auto &enF = c->blocks[r.jindF * c->nrep + gF].en;
    for (size_t i = 0; i < BI.lev.size(); i++) {
        const double eI = BI.en[i];
        if (!c->lv_lowok[BI.lev[i]]) continue;
        if (!((eI - c->zpe) >= r.elow_lo && (eI - c->zpe) <= r.elow_hi)) continue;
        int a, b;
        window(enF, eI, r.nu_lo, r.nu_hi, a, b);
        if (b <= a) continue;
        P.items.push_back(Item{BI.lev[i], r.jindI, r.igammaI, r.jindF, gF, a, b, query, P.ncand});
        P.ncand += b - a;
    }
}

// Chunks of one (query, J, symmetry, J') group of items, with descriptor
// bytes and staged records of their own (offsets relative to the group until
// the join in build_chunks).
struct GroupOut {
    std::vector<Chunk> chunks;
    std::vector<unsigned char> meta;
    std::vector<int> srec;
    bool vstream = false;
    bool bad = false;
};

static void build_group(const gi_ctx *c, const Plan &P, const std::tuple<int, int, int, int> &key,
                        std::vector<int> its, long long run, GroupOut &o)
{
    const bool batched = P.mode != GI_MODE_FP64;  // both matrix-core modes
    const int kChunk = 256;
    auto put = [&](const void *src, size_t bytes) {
        const size_t at = (o.meta.size() + 255) & ~size_t(255);
        o.meta.resize(at + std::max<size_t>(bytes, 8));
        if (bytes) std::memcpy(o.meta.data() + at, src, bytes);
        return at;
    };
    const int query = std::get<0>(key), jindI = std::get<1>(key), gI = std::get<2>(key),
              jindF = std::get<3>(key);
    const int gF = c->igamma_pair[gI];
    const int ndegI = c->degen[gI], ndegF = c->degen[gF];
    const Block &BI = c->blocks[jindI * c->nrep + gI];
    // energy order keeps every chunk's windows a narrow band (level tables
    // usually list a block's levels in energy order already)
    const auto by_energy = [&](int x, int y) {
        return c->lv_energy[P.items[x].level] < c->lv_energy[P.items[y].level];
    };
    if (batched && !std::is_sorted(its.begin(), its.end(), by_energy))
        std::stable_sort(its.begin(), its.end(), by_energy);
    const GroupPlan &GP = group_of(c, jindI, gI, jindF);
    if (!GP.valid) {
        o.bad = true;
        return;
    }
    for (size_t s0 = 0; s0 < its.size(); s0 += kChunk) {
        const int n_it = (int)std::min<size_t>(kChunk, its.size() - s0);
        std::vector<long long> uptr, outoff(n_it);
        std::vector<int> uideg, urec, iunit(n_it * GI_MAXDEG, -1), cstart(n_it), lo(n_it),
            hi(n_it), lev(n_it);
        std::vector<double> eI(n_it);
        int cnt = 0, Lo = 1 << 30, Hi = 0;
        const long long srec0 = (long long)o.srec.size();
        for (int k = 0; k < n_it; k++) {
            const Item &it = P.items[its[s0 + k]];
            const int rec = c->lv_irec0[it.level];
            if (BI.h_vec) o.srec.push_back(rec);  // one staged copy per lower level
            for (int ideg = 0; ideg < ndegI; ideg++) {
                if (!degeneracy_filter(c, gI, gF, ideg, 0)) continue;
                iunit[k * GI_MAXDEG + ideg] = (int)uideg.size();
                uideg.push_back(ideg);
                urec.push_back(rec);
                // host block: the staged slot k, made an address by launch_plan
                uptr.push_back(BI.h_vec ? (long long)k
                                        : (long long)(BI.d_vec + (size_t)rec * BI.nsize));
            }
            cstart[k] = cnt;
            lo[k] = it.lo;
            hi[k] = it.hi;
            Lo = std::min(Lo, it.lo);
            Hi = std::max(Hi, it.hi);
            lev[k] = it.level;
            eI[k] = c->lv_energy[it.level];
            outoff[k] = run + cnt;
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
        ch.query = query;
        ch.jindI = jindI; ch.gI = gI; ch.jindF = jindF; ch.gF = gF;
        ch.lhost = BI.h_vec != nullptr;
        ch.uhost = c->blocks[jindF * c->nrep + gF].h_vec != nullptr;
        ch.srec0 = srec0;
        ch.nsrec = (int)((long long)o.srec.size() - srec0);
        o.vstream = o.vstream || ch.lhost || ch.uhost;
        ch.nitems = n_it; ch.nunits = (int)uideg.size(); ch.ncand = cnt; ch.cb_mask = mask;
        ch.Lo = Lo; ch.Hi = Hi;
        ch.cand_begin = run;
        run += cnt;
        ch.off_uptr = put(uptr.data(), uptr.size() * 8);
        ch.off_uideg = put(uideg.data(), uideg.size() * 4);
        ch.off_urec = put(urec.data(), urec.size() * 4);
        ch.off_iunit = put(iunit.data(), iunit.size() * 4);
        ch.off_cstart = put(cstart.data(), cstart.size() * 4);
        ch.off_lo = put(lo.data(), lo.size() * 4);
        ch.off_hi = put(hi.data(), hi.size() * 4);
        ch.off_level = put(lev.data(), lev.size() * 4);
        ch.off_eI = put(eI.data(), eI.size() * 8);
        ch.off_out = put(outoff.data(), outoff.size() * 8);
        o.chunks.push_back(ch);
    }
}

// Group items and cut chunks; unit pointers become device addresses. Groups
// are built in key order, candidate slots in chunk order; a large job builds
// its groups on several threads (each group's first candidate slot is known
// from the group sizes) and joins them in order, so the plan is the same.
static int build_chunks(gi_ctx *c, Plan &P)
{
    // groups in (query, J, symmetry, J') order through a dense index over
    // those four (at J 0..20 a tree of 240 groups cost 1.8 ms for 87k items)
    int nq = 1;
    for (const Item &it : P.items) nq = std::max(nq, it.query + 1);
    const size_t ng = (size_t)nq * c->nJ * c->nrep * c->nJ;
    auto gkey = [&](const Item &it) {
        return (((size_t)it.query * c->nJ + it.jindI) * c->nrep + it.gI) * c->nJ + it.jindF;
    };
    std::vector<int> gcount(ng, 0);
    for (const Item &it : P.items) gcount[gkey(it)]++;
    std::vector<size_t> gslot(ng, SIZE_MAX);
    std::vector<std::pair<std::tuple<int, int, int, int>, std::vector<int>>> gv;
    for (size_t k = 0; k < ng; k++) {
        if (!gcount[k]) continue;
        size_t r = k;
        const int jindF = (int)(r % c->nJ); r /= c->nJ;
        const int gI = (int)(r % c->nrep); r /= c->nrep;
        const int jindI = (int)(r % c->nJ); r /= c->nJ;
        gslot[k] = gv.size();
        gv.emplace_back(std::make_tuple((int)r, jindI, gI, jindF), std::vector<int>());
        gv.back().second.reserve(gcount[k]);
    }
    for (int i = 0; i < (int)P.items.size(); i++) gv[gslot[gkey(P.items[i])]].second.push_back(i);
    std::vector<long long> run0(gv.size() + 1, 0);
    for (size_t g = 0; g < gv.size(); g++) {
        long long n = 0;
        for (int i : gv[g].second) n += P.items[i].hi - P.items[i].lo;
        run0[g + 1] = run0[g] + n;
    }
    std::vector<GroupOut> out(gv.size());
    const int T = std::min<int>(plan_threads(P.items.size()), (int)gv.size());
    std::atomic<size_t> next{0};
    run_threads(std::max(T, 1), [&](int) {
        for (size_t g; (g = next.fetch_add(1)) < gv.size();)
            build_group(c, P, gv[g].first, std::move(gv[g].second), run0[g], out[g]);
    });
    size_t mtot = P.meta.size();
    for (const GroupOut &o : out) mtot = ((mtot + 255) & ~size_t(255)) + o.meta.size();
    P.meta.reserve(mtot);
    for (GroupOut &o : out) {
        if (o.bad) return fail(c, GI_EINVAL, "internal: no plan for a (J, J') pair");
        const size_t base = (P.meta.size() + 255) & ~size_t(255);
        const long long sbase = (long long)P.srec.size();
        P.meta.resize(base);
        P.meta.insert(P.meta.end(), o.meta.begin(), o.meta.end());
        P.srec.insert(P.srec.end(), o.srec.begin(), o.srec.end());
        P.vstream = P.vstream || o.vstream;
        for (Chunk ch : o.chunks) {
            for (size_t *f : {&ch.off_uptr, &ch.off_uideg, &ch.off_urec, &ch.off_iunit,
                              &ch.off_cstart, &ch.off_lo, &ch.off_hi, &ch.off_level, &ch.off_eI,
                              &ch.off_out})
                *f += base;
            ch.srec0 += sbase;
            P.chunks.push_back(ch);
        }
    }
    // each query's chunks are consecutive, so its candidates are one range
    for (auto &q : P.q) q.cand_begin = q.cand_end = -1;
    for (const Chunk &ch : P.chunks) {
        Query &q = P.q[ch.query];
        if (q.cand_begin < 0) q.cand_begin = ch.cand_begin;
        q.cand_end = ch.cand_begin + ch.ncand;
    }
    long long pos = 0;
    for (auto &q : P.q) {
        if (q.cand_begin < 0) q.cand_begin = q.cand_end = pos;
        pos = q.cand_end;
    }
    return GI_OK;
}

static void fill_corr_args(gi_ctx *c, const Plan &P, const Chunk &ch, unsigned char *d_meta,
                           const Consts &K, CorrArgs &a)
{
    const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
    const Query &q = P.q[ch.query];
    a.ncand = ch.ncand; a.nitems = ch.nitems; a.nsizeF = BF.nsize;
    a.ndegI = c->degen[ch.gI]; a.ndegF = c->degen[ch.gF]; a.cb_mask = ch.cb_mask;
    a.reduced = c->reduced;
    a.jI = c->jvals[ch.jindI]; a.jF = c->jvals[ch.jindF];
    a.item_cand_start = (const int *)(d_meta + ch.off_cstart);
    a.item_lo = (const int *)(d_meta + ch.off_lo);
    a.item_level = (const int *)(d_meta + ch.off_level);
    a.item_out_off = (const long long *)(d_meta + ch.off_out);
    a.item_energy = (const double *)(d_meta + ch.off_eI);
    a.blk_col = BF.d_col; a.blk_lev = BF.d_lev; a.blk_en = BF.d_en; a.blk_upok = BF.d_upok;
    a.vec = BF.d_vec;
    a.gnsI = c->gns[ch.gI]; a.zpe = q.zpe; a.beta = q.beta; a.q_stat = q.q_stat;
    a.A_coef_s_1 = K.A_coef_s_1; a.intens_cm_mol = K.intens_cm_mol;
    a.thresh_S = q.thS; a.thresh_I = q.thI;
    a.qpar = P.qpar ? P.qpar : (P.qdev ? P.qdev + 4 * ch.query : nullptr);
}

static LineSoA shift(LineSoA s, long long o);

// Multiply-add flops the banded correlation kernels execute for a chunk: the
// 16 x 16 output blocks (rows from Lo, columns item*ncb + pair) that hold an
// entry inside some column's window, each over K = nsize_F (the kernels skip
// the others; work[8] against work[16] is the band waste).
static double band_exec_flops(const Plan &P, const Chunk &ch, int ncb, int nsize)
{
    const int *lo = (const int *)(P.meta.data() + ch.off_lo);
    const int *hi = (const int *)(P.meta.data() + ch.off_hi);
    const int ncols = ch.nitems * ncb;
    long long blocks = 0;
    for (int r0 = ch.Lo; r0 < ch.Hi; r0 += 16)
        for (int c0 = 0; c0 < ncols; c0 += 16) {
            bool hit = false;
            for (int c = c0; c < std::min(c0 + 16, ncols) && !hit; c++)
                hit = lo[c / ncb] < r0 + 16 && hi[c / ncb] > r0;
            blocks += hit;
        }
    return 2.0 * 256.0 * (double)blocks * nsize;
}

// Consecutive chunks grouped into rounds of at most kRoundBytes of device
// scratch (GI_BATCH_BYTES overrides). With per-round output the first
// c->pipe_rounds - 1 rounds each take c->pipe_share percent of the work still
// left (at least 1 MB of scratch) and the last takes the rest: each round's
// line copy overlaps the next round's compute while the copy of a round takes
// less than the next round's compute, and the copy left after the last round
// is (1 - share)^(pipe_rounds-1) of the output instead of 1/pipe_rounds.
static std::vector<std::pair<size_t, size_t>> plan_rounds(const gi_ctx *c, const Plan &P,
                                                          const std::vector<size_t> &bytes)
{
    size_t kRoundBytes = (size_t)12 << 30;
    if (const char *env = std::getenv("GI_BATCH_BYTES")) {
        const long long v = std::atoll(env);
        if (v > 0) kRoundBytes = (size_t)v;
    }
    if (P.vstream) {  // staged blocks: a round per chunk, its records staged before it
        std::vector<std::pair<size_t, size_t>> one;
        for (size_t i = 0; i < bytes.size(); i++) one.push_back({i, i + 1});
        return one;
    }
    size_t left = 0;
    for (size_t b : bytes) left += b;
    // streamed mu is copied once per round: its rounds are only those the
    // device budget needs (at J 10<->11, four rounds re-copy 3 x 96 MB of mu
    // to hide 73 MB of line copies)
    auto budget = [&](size_t r) {
        if (!P.pipe || c->mu_streamed || (int)r + 1 >= c->pipe_rounds) return kRoundBytes;
        return std::min(kRoundBytes, std::max(left / 100 * c->pipe_share, (size_t)1 << 20));
    };
    std::vector<std::pair<size_t, size_t>> rounds;
    size_t start = 0, acc = 0, cap = budget(0);
    for (size_t i = 0; i < bytes.size(); i++) {
        if (i > start && acc + bytes[i] > cap) {
            rounds.push_back({start, i});
            left -= acc;
            start = i;
            acc = 0;
            cap = budget(rounds.size());
        }
        acc += bytes[i];
    }
    if (start < bytes.size()) rounds.push_back({start, bytes.size()});
    return rounds;
}

// Events and compaction scratch of the per-round output.
static int setup_pipe(gi_ctx *c, Slot &S, Plan &P,
                      const std::vector<std::pair<size_t, size_t>> &rounds, int **pc_counts,
                      long long **pc_offs)
{
    int rc = GI_OK;
    *pc_counts = nullptr;
    *pc_offs = nullptr;
    if (!P.pipe) return GI_OK;
    size_t ncnt = 0, noff = 0;
    P.rounds.resize(rounds.size());
    for (size_t r = 0; r < rounds.size(); r++) {
        RoundOut &o = P.rounds[r];
        const Chunk &a = P.chunks[rounds[r].first], &b = P.chunks[rounds[r].second - 1];
        o.cb = a.cand_begin;
        o.ce = b.cand_begin + b.ncand;
        o.nb = compact_blocks(o.ce - o.cb);
        o.cbase = ncnt;
        o.obase = noff;
        ncnt += o.nb;
        noff += o.nb + 1;
        if (hipEventCreateWithFlags(&o.ev, hipEventDisableTiming) != hipSuccess)
            return fail(c, GI_EHIP, "hipEventCreate");
    }
    *pc_counts = (int *)slot_scratch(c, S, "rcounts", std::max<size_t>(ncnt, 1) * 4, &rc);
    if (rc) return rc;
    *pc_offs = (long long *)slot_scratch(c, S, "roffs", std::max<size_t>(noff, 1) * 8, &rc);
    if (rc) return rc;
    P.rcounts = (long long *)slot_pinned(c, S, (P.q.size() + rounds.size()) * 8 + 64, &rc);
    if (rc) return rc;
    P.counts = P.rcounts + rounds.size();
    return GI_OK;
}

// Compaction of round r's candidates, its line count to pinned memory and
// the event the host waits on before copying the round's lines.
static int round_compact(gi_ctx *c, Slot &S, Plan &P, size_t r, int *flag, LineSoA cand,
                         int *pc_counts, long long *pc_offs, Timeline *tl)
{
    if (!P.pipe) return GI_OK;
    RoundOut &o = P.rounds[r];
    if (o.nb > 0) {
        TL_BEGIN("compact");
        launch_compact(S.stream, o.ce - o.cb, flag + o.cb, pc_counts + o.cbase, pc_offs + o.obase,
                       shift(cand, o.cb), shift(P.comp, o.cb));
        TL_END();
        HIPCHK(hipMemcpyAsync(P.rcounts + r, pc_offs + o.obase + o.nb, 8, hipMemcpyDeviceToHost,
                              S.stream));
    }
    if (!P.capture) HIPCHK(hipEventRecord(o.ev, S.stream));  // a sweep records it per replay
    return GI_OK;
}

// Split-K parts of a chunk's split-mode correlation: a function of the upper
// block only (never of the batch), so a candidate's sums do not depend on what
// it is batched with. Long products are cut so that a round fills the GPU.
static int corr_kparts(const gi_ctx *c, const Plan &P, const Chunk &ch)
{
    if (P.mode == GI_MODE_FP64) return 1;
    const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
    return P.mode == GI_MODE_SPLIT ? corr_parts(c, BF.Kp, true) : corr_parts(c, BF.nsize, false);
}

// Arguments of one chunk's intensity step from its stored amplitudes.
static IntensArgs intens_args(gi_ctx *c, const Plan &P, const Chunk &ch, unsigned char *d_meta,
                              const Consts &K, int *flag, LineSoA cand)
{
    CorrArgs ca;
    fill_corr_args(c, P, ch, d_meta, K, ca);
    IntensArgs a;
    a.ncand = ca.ncand; a.nitems = ca.nitems; a.ndegI = ca.ndegI; a.ndegF = ca.ndegF;
    a.cb_mask = ca.cb_mask; a.reduced = ca.reduced; a.jI = ca.jI; a.jF = ca.jF;
    a.item_cand_start = ca.item_cand_start; a.item_lo = ca.item_lo;
    a.item_level = ca.item_level; a.item_out_off = ca.item_out_off;
    a.item_energy = ca.item_energy;
    a.blk_lev = ca.blk_lev; a.blk_en = ca.blk_en; a.blk_upok = ca.blk_upok;
    a.gnsI = ca.gnsI; a.zpe = ca.zpe; a.beta = ca.beta; a.q_stat = ca.q_stat;
    a.A_coef_s_1 = ca.A_coef_s_1; a.intens_cm_mol = ca.intens_cm_mol;
    a.thresh_S = ca.thresh_S; a.thresh_I = ca.thresh_I;
    a.qpar = ca.qpar;
    a.kparts = corr_kparts(c, P, ch);
    a.ampx = P.ampx;
    a.ampx_stride = P.ampx_stride;
    a.amp = cand.amp; a.amp_m = c->amp_m;
    a.flag = flag; a.out_iI = cand.iI; a.out_iF = cand.iF;
    a.out_nu = cand.nu; a.out_S = cand.S; a.out_A = cand.A; a.out_I = cand.I;
    return a;
}

static unsigned char *stage(gi_ctx *c, Slot &S, const void *src, size_t n, int *rc);

// Upload a job table (jobs, then the int first-block of each) through the
// slot's pinned arena into dev.
template <typename T>
static int send_jobs(gi_ctx *c, Slot &S, unsigned char *dev, const std::vector<T> &jobs,
                     const std::vector<int> &begin, size_t cap)
{
    int rc = GI_OK;
    unsigned char *hb = stage(c, S, nullptr, cap + 4 * begin.size(), &rc);
    if (rc) return rc;
    std::memcpy(hb, jobs.data(), jobs.size() * sizeof(T));
    std::memcpy(hb + cap, begin.data(), 4 * begin.size());
    HIPCHK(hipMemcpyAsync(dev, hb, cap + 4 * begin.size(), hipMemcpyHostToDevice, S.stream));
    return GI_OK;
}

// The job tables of one round packed into one pinned staging block, each
// table 256-aligned with its int array of first tiles right after it, and
// sent to the device in one copy before the round's first launch.
struct TablePack {
    struct Ent {
        const void *src;
        size_t bytes;
        const int *ints;
        size_t nints, off;
    };
    unsigned char *dev;
    size_t n = 0;
    std::vector<Ent> ents;
    explicit TablePack(unsigned char *d) : dev(d) {}
    // device address of table v; *dbegin gets that of its first tiles
    template <class T>
    const T *add(const std::vector<T> &v, const std::vector<int> &begin, const int **dbegin)
    {
        const size_t o = (n + 255) & ~size_t(255), b = v.size() * sizeof(T);
        ents.push_back(Ent{v.data(), b, begin.data(), begin.size(), o});
        n = o + b + 4 * begin.size();
        *dbegin = (const int *)(dev + o + b);
        return (const T *)(dev + o);
    }
};

static int send_pack(gi_ctx *c, Slot &S, const TablePack &T, size_t cap)
{
    if (T.n == 0) return GI_OK;
    if (T.n > cap) return fail(c, GI_EINVAL, "internal: a round's job tables exceed their buffer");
    int rc = GI_OK;
    unsigned char *hb = stage(c, S, nullptr, T.n, &rc);
    if (rc) return rc;
    for (const auto &e : T.ents) {
        if (e.bytes) std::memcpy(hb + e.off, e.src, e.bytes);
        if (e.nints) std::memcpy(hb + e.off + e.bytes, e.ints, 4 * e.nints);
    }
    HIPCHK(hipMemcpyAsync(T.dev, hb, T.n, hipMemcpyHostToDevice, S.stream));
    return GI_OK;
}

// Bytes a round's tables can take: each table kind at its per-round maximum
// plus the alignment of each.
static size_t pack_cap(std::initializer_list<size_t> tables)
{
    size_t n = 0;
    for (size_t t : tables) n += t + 256;
    return n;
}

// Intensity jobs of chunks [c0, c1): one per chunk with candidates, blocks of
// 256 candidates; returns the block count.
static int round_intensity_jobs(gi_ctx *c, const Plan &P, size_t c0, size_t c1,
                                unsigned char *d_meta, const Consts &K, int *flag, LineSoA cand,
                                std::vector<IntensArgs> &jobs, std::vector<int> &begin)
{
    int nb = 0;
    for (size_t i = c0; i < c1; i++) {
        const Chunk &ch = P.chunks[i];
        if (ch.ncand <= 0) continue;
        jobs.push_back(intens_args(c, P, ch, d_meta, K, flag, cand));
        begin.push_back(nb);
        nb += (int)((ch.ncand + 255) / 256);
    }
    return nb;
}

// Pinned staging of an upload on slot S (valid until the slot's next launch
// sequence resets its arena).
static unsigned char *stage(gi_ctx *c, Slot &S, const void *src, size_t n, int *rc)
{
    unsigned char *p = S.arena.alloc(std::max<size_t>(n, 8));
    if (!p) {
        *rc = fail(c, GI_ENOMEM, "pinned host allocation (upload staging) failed");
        return nullptr;
    }
    if (src && n) std::memcpy(p, src, n);
    *rc = GI_OK;
    return p;
}

// Staging of the blocks kept in host memory (gi_set_vector_budget): two
// device buffers each for lower and upper records, used by alternate chunks;
// chunk i's copies on the copy stream wait until chunk i-2's kernels are done
// with the buffer, and chunk i's kernels wait until its copies are complete.
// Device address of staged upper record f of a chunk whose window starts at
// Lo, for kernels that index records as vec + blk_col[f] * nsize with the
// identity map: the staging buffer shifted back by Lo records.
static const double *staged_rows(const double *buf, int Lo, int nsize)
{
    return (const double *)((uintptr_t)buf - (uintptr_t)Lo * nsize * sizeof(double));
}

static int vec_stage_setup(gi_ctx *c, Slot &S, Plan &P, VecStage &V)
{
    if (!P.vstream) return GI_OK;
    int rc = GI_OK;
    const bool split = P.mode == GI_MODE_SPLIT;
    size_t lmax = 1, umax = 1, urows = 1, ukp = 64;
    for (const Chunk &ch : P.chunks) {
        if (ch.lhost)
            lmax = std::max(lmax, (size_t)ch.nsrec * c->blocks[ch.jindI * c->nrep + ch.gI].nsize);
        if (ch.uhost) {
            const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
            umax = std::max(umax, (size_t)std::max(ch.Hi - ch.Lo, 0) * BF.nsize);
            urows = std::max(urows, (size_t)std::max(ch.Hi - ch.Lo, 0));
            ukp = std::max(ukp, (size_t)BF.Kp);
        }
    }
    // the split mode stages the upper records' digits instead of their FP64 values
    V.uplane = (long long)(urows * ukp);
    for (int b = 0; b < 2; b++) {
        V.lbuf[b] = (double *)slot_scratch(c, S, b ? "vs_l1" : "vs_l0", 8 * lmax, &rc);
        if (rc) return rc;
        if (split) {
            V.ubd[b] = (int8_t *)slot_scratch(c, S, b ? "vs_ud1" : "vs_ud0",
                                              (size_t)c->split_S * V.uplane, &rc);
            if (rc) return rc;
            V.ube[b] = (int *)slot_scratch(c, S, b ? "vs_ue1" : "vs_ue0", 4 * urows, &rc);
        } else {
            V.ubuf[b] = (double *)slot_scratch(c, S, b ? "vs_u1" : "vs_u0", 8 * umax, &rc);
        }
        if (rc) return rc;
    }
    for (auto &e : S.vec_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(c, GI_EHIP, "hipEventCreate");
    V.cs = S.cstream ? S.cstream : S.stream;
    // the staged lower records' addresses into every such chunk's unit table
    for (size_t i = 0; i < P.chunks.size(); i++) {
        const Chunk &ch = P.chunks[i];
        if (!ch.lhost) continue;
        const int nsz = c->blocks[ch.jindI * c->nrep + ch.gI].nsize;
        long long *up = (long long *)(P.meta.data() + ch.off_uptr);
        for (int u = 0; u < ch.nunits; u++)
            up[u] = (long long)(V.lbuf[i & 1] + (size_t)up[u] * nsz);
    }
    return GI_OK;
}

// Copies of chunk i's records in host memory to staging buffer i & 1 (runs
// of consecutive records in one copy each), then the compute stream waits.
static int vec_stage_chunk(gi_ctx *c, Slot &S, const Plan &P, size_t i, const VecStage &V)
{
    const Chunk &ch = P.chunks[i];
    if (!(ch.lhost || ch.uhost)) return GI_OK;
    const int b = (int)(i & 1);
    if (i >= 2) HIPCHK(hipStreamWaitEvent(V.cs, S.vec_ev[2 + b], 0));
    auto runs = [&](const Block &B, double *dst, int n, auto rec_of) -> int {
        int s0 = 0;
        while (s0 < n) {
            int s1 = s0 + 1;
            while (s1 < n && rec_of(s1) == rec_of(s1 - 1) + 1) s1++;
            HIPCHK(hipMemcpyAsync(dst + (size_t)s0 * B.nsize,
                                  B.h_vec + (size_t)rec_of(s0) * B.nsize,
                                  (size_t)(s1 - s0) * B.nsize * sizeof(double),
                                  hipMemcpyHostToDevice, V.cs));
            s0 = s1;
        }
        return GI_OK;
    };
    int rc;
    if (ch.lhost) {
        const Block &BI = c->blocks[ch.jindI * c->nrep + ch.gI];
        if ((rc = runs(BI, V.lbuf[b], ch.nsrec,
                       [&](int s) { return P.srec[(size_t)ch.srec0 + s]; })) != GI_OK)
            return rc;
    }
    if (ch.uhost && ch.Hi > ch.Lo && P.mode == GI_MODE_SPLIT) {
        // digit rows of the window's records (each plane), then their exponents
        const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
        const int n = ch.Hi - ch.Lo;
        auto rec_of = [&](int s) { return c->lv_irec0[BF.lev[ch.Lo + s]]; };
        int s0 = 0;
        while (s0 < n) {
            int s1 = s0 + 1;
            while (s1 < n && rec_of(s1) == rec_of(s1 - 1) + 1) s1++;
            const size_t r0 = (size_t)rec_of(s0);
            for (int q = 0; q < c->split_S; q++)
                HIPCHK(hipMemcpyAsync(V.ubd[b] + q * V.uplane + (size_t)s0 * BF.Kp,
                                      BF.h_ud + q * BF.uplane + r0 * BF.Kp,
                                      (size_t)(s1 - s0) * BF.Kp, hipMemcpyHostToDevice, V.cs));
            HIPCHK(hipMemcpyAsync(V.ube[b] + s0, BF.h_ue + r0, (size_t)(s1 - s0) * 4,
                                  hipMemcpyHostToDevice, V.cs));
            s0 = s1;
        }
    } else if (ch.uhost && ch.Hi > ch.Lo) {
        const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
        if ((rc = runs(BF, V.ubuf[b], ch.Hi - ch.Lo,
                       [&](int s) { return c->lv_irec0[BF.lev[ch.Lo + s]]; })) != GI_OK)
            return rc;
    }
    HIPCHK(hipEventRecord(S.vec_ev[b], V.cs));
    HIPCHK(hipStreamWaitEvent(S.stream, S.vec_ev[b], 0));
    return GI_OK;
}

// After chunk i's last kernel that reads its staged records.
static int vec_stage_done(gi_ctx *c, Slot &S, const Plan &P, size_t i)
{
    const Chunk &ch = P.chunks[i];
    if (!(ch.lhost || ch.uhost)) return GI_OK;
    HIPCHK(hipEventRecord(S.vec_ev[2 + (i & 1)], S.stream));
    return GI_OK;
}

// Streamed mu in the matrix-core modes (gi_set_mu_budget): parts of mu_np vF
// rows of the permuted FP64 mu (batched mode) or of its digit planes (split
// mode) go from pinned host memory into two device buffers on the copy
// stream, and each part's half-ls products compute the class rows inside it,
// every row a whole sum in the order of resident mu (the reference splits mu
// into parts of vF rows the same way, do_1st_half_ls_blocks,
// cuda_host.cu:783-817). Part q fills buffer q & 1, which is refilled only
// after the products of part q - 2 are done.
struct MuParts {
    void *buf[2] = {nullptr, nullptr};
    int nparts = 0;  // parts copied so far in this call
};

static int mu_parts_setup(gi_ctx *c, Slot &S, size_t bytes, MuParts &M)
{
    int rc = GI_OK;
    for (int b = 0; b < 2; b++) {
        M.buf[b] = slot_scratch(c, S, b ? "mupart1" : "mupart0", bytes, &rc);
        if (rc) return rc;
    }
    for (auto &e : S.mu_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(c, GI_EHIP, "hipEventCreate");
    return GI_OK;
}

// Copy of the next part: `height` host rows of `width` bytes, `spitch` bytes
// apart from src, packed into its buffer; the compute stream waits for it.
// *b is the buffer; the caller records mu_ev[2 + *b] after the part's products.
static int mu_part_copy(gi_ctx *c, Slot &S, MuParts &M, const void *src, size_t spitch,
                        size_t width, size_t height, int *b, Timeline *tl)
{
    hipStream_t cs = S.cstream ? S.cstream : S.stream;
    *b = M.nparts++ & 1;
    if (M.nparts > 2) HIPCHK(hipStreamWaitEvent(cs, S.mu_ev[2 + *b], 0));
    TL_BEGIN_ON("mu_copy", cs);
    HIPCHK(hipMemcpy2DAsync(M.buf[*b], width, src, spitch, width, height, hipMemcpyHostToDevice,
                            cs));
    TL_END_ON(cs);
    HIPCHK(hipEventRecord(S.mu_ev[*b], cs));
    HIPCHK(hipStreamWaitEvent(S.stream, S.mu_ev[*b], 0));
    return GI_OK;
}

// The half-ls jobs of one mu part: each job cut to its class rows inside
// vF rows [v0, v0 + np), tiles renumbered from 0 (TM x TN tiles). Returns
// the tile count.
template <class Job>
static int part_jobs(const std::vector<Job> &all, int v0, int np, int TM, int TN,
                     std::vector<Job> &out, std::vector<int> &begin)
{
    int tiles = 0;
    for (const Job &j0 : all) {
        const int lo = std::max(0, v0 - j0.cd.off), hi = std::min(j0.cd.n, v0 + np - j0.cd.off);
        if (lo >= hi) continue;
        Job j = j0;
        j.m_lo = lo;
        j.m_hi = hi;
        j.mt = (hi - lo + TM - 1) / TM;
        j.tile_begin = tiles;
        tiles += j.mt * ((j.nunits * j.cd.ncols + TN - 1) / TN);
        out.push_back(j);
        begin.push_back(j.tile_begin);
    }
    return tiles;
}

// Stream capture of one round into a graph (sweeps). On an error path the
// destructor ends the capture so the stream is usable again.
struct RoundCapture {
    hipStream_t st = nullptr;
    bool on = false;
    int begin(gi_ctx *c, const Plan &P, hipStream_t s)
    {
        if (!P.capture) return GI_OK;
        st = s;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess)
            return fail(c, GI_EHIP, "hipStreamBeginCapture");
        on = true;
        return GI_OK;
    }
    int end(gi_ctx *c, Plan &P)
    {
        if (!on) return GI_OK;
        on = false;
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(st, &g) != hipSuccess || !g)
            return fail(c, GI_EHIP, "hipStreamEndCapture");
        hipGraphExec_t e = nullptr;
        const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (r != hipSuccess) return fail(c, GI_EHIP, "hipGraphInstantiate");
        P.graphs.push_back(e);
        return GI_OK;
    }
    ~RoundCapture()
    {
        if (on) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
        }
    }
};

// Batched FP64 mode over rounds of chunks (see the file comment).
// Row stride of the folded vectors Ht in the FP64 batched mode: even, so
// every row starts 16-byte aligned for the correlation's 16-byte loads.
static int ht_stride(int nsize) { return (nsize + 1) & ~1; }

static int run_batched(gi_ctx *c, Slot &S, Plan &P, unsigned char *d_meta, Timeline *tl,
                       int *flag, LineSoA cand)
{
    const int N = c->nvib;
    const Consts K;
    int rc = GI_OK;
    const auto &chunks = P.chunks;
    struct Need {
        size_t g, h, ht;
        int njobs;
    };
    std::vector<Need> need(chunks.size());
    std::vector<size_t> bytes(chunks.size());
    for (size_t i = 0; i < chunks.size(); i++) {
        const Chunk &ch = chunks[i];
        const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
        const int ncb = c->degen[ch.gI] * c->degen[ch.gF];
        const int nsizeF = c->blocks[ch.jindF * c->nrep + ch.gF].nsize;
        need[i] = Need{(size_t)ch.nunits * GP.Gtot_p, (size_t)ch.nunits * GP.Htot,
                       (size_t)ch.nitems * ncb * ht_stride(nsizeF),
                       ch.nunits > 0 ? (int)GP.oc.size() : 0};
        bytes[i] = 8 * (need[i].g + need[i].h + need[i].ht);
    }
    const auto rounds = plan_rounds(c, P, bytes);
    size_t mg = 1, mh = 1, mht = 1;
    int mjobs = 1, mchunks = 1;
    for (auto &rg : rounds) {
        size_t sg = 0, sh = 0, sht = 0;
        int sjobs = 0;
        for (size_t i = rg.first; i < rg.second; i++) {
            sg += need[i].g; sh += need[i].h; sht += need[i].ht;
            sjobs += need[i].njobs;
        }
        mg = std::max(mg, sg); mh = std::max(mh, sh); mht = std::max(mht, sht);
        mjobs = std::max(mjobs, sjobs);
        mchunks = std::max(mchunks, (int)(rg.second - rg.first));
    }
    // every buffer is sized for the largest round before anything is launched
    double *G = (double *)slot_scratch(c, S, "G", mg * 8, &rc); if (rc) return rc;
    double *H = (double *)slot_scratch(c, S, "H", mh * 8, &rc); if (rc) return rc;
    double *Ht = (double *)slot_scratch(c, S, "Ht", mht * 8, &rc); if (rc) return rc;
    // per-mu-part half-ls tables (streamed mu); the round's other tables go
    // to rtab in one copy
    const size_t jobs_bytes = (size_t)mjobs * sizeof(HalfJob);
    unsigned char *d_jobs = nullptr;
    if (c->mu_streamed) {
        d_jobs = (unsigned char *)slot_scratch(c, S, "jobs", jobs_bytes + 4 * (size_t)mjobs, &rc);
        if (rc) return rc;
    }
    const size_t mc = (size_t)mchunks;
    const size_t rtab_cap = pack_cap({mc * (sizeof(GMapJob) + 4), mc * (sizeof(GMapJob) + 4),
                                      mc * (sizeof(GBuildJob) + 4),
                                      (size_t)mjobs * (sizeof(HalfJob) + 4),
                                      mc * (sizeof(FoldJob) + 4), mc * (sizeof(CorrMMArgs) + 4),
                                      mc * (sizeof(IntensArgs) + 4), (size_t)mjobs * 4, mc * 4,
                                      mc * 4, 3 * mc * 4});
    unsigned char *rtab = (unsigned char *)slot_scratch(c, S, "rtab", rtab_cap, &rc);
    if (rc) return rc;
    // persistent rounds (gi_set_persist): the half-ls, fold, correlation and
    // intensity of a round in one launch after its G build; not with mu or
    // eigenvector blocks in host memory, whose copies sit between the stages
    const bool persist = c->persist && !c->mu_streamed && !P.vstream;
    const int pccfg = persist_corr_cfg_ok(c->tile_corr) ? c->tile_corr : 0;
    int *psync = nullptr;
    if (persist) {
        psync = (int *)slot_scratch(c, S, "psync", (3 * mc + 2) * 4, &rc);
        if (rc) return rc;
        P.perr = psync + 3 * mc + 1;
    }
    // tile shape of the half-ls products of round [c0, c1): the setting, or
    // with kTileAuto 128 x 128 on eight wavefronts while the round has
    // kAutoTileWaves waves of them (two workgroups per CU), else 64 x 64
    // (measured, half-ls per pass: C3v J 0..15 1116 ms against 1189-1258 ms
    // for the 64-row and 64-column tiles; C2v J 10<->11 3.46 ms against 3.39 ms
    // for 64 x 128, which its tuning table picks; 64 x 64 wins rounds under
    // about two waves)
    auto round_halfls_cfg = [&](size_t c0, size_t c1) {
        if (c->tile_halfls != kTileAuto) return c->tile_halfls;
        long long t = 0;
        for (size_t i = c0; i < c1; i++) {
            const Chunk &ch = chunks[i];
            if (ch.nunits <= 0) continue;
            for (const OutClass &oc : group_of(c, ch.jindI, ch.gI, ch.jindF).oc)
                t += (long long)((oc.n + 127) / 128) * ((ch.nunits * (int)oc.cols.size() + 127) / 128);
        }
        return t >= (long long)kAutoTileWaves * 2 * c->ncu ? kTile128x128w8 : 0;
    };

    int maxparts = 1;
    for (const Chunk &ch : chunks) maxparts = std::max(maxparts, corr_kparts(c, P, ch));
    P.ampx_stride = (long long)c->amp_m * c->amp_m * std::max<long long>(P.ncand, 1);
    P.ampx = maxparts > 1 ? (double *)slot_scratch(c, S, "c_ampx",
                                                   8 * (size_t)(maxparts - 1) * P.ampx_stride, &rc)
                          : nullptr;
    if (rc) return rc;
    int *pc_counts = nullptr;
    long long *pc_offs = nullptr;
    if ((rc = setup_pipe(c, S, P, rounds, &pc_counts, &pc_offs)) != GI_OK) return rc;
    MuParts MP;
    if (c->mu_streamed && (rc = mu_parts_setup(c, S, (size_t)3 * N * c->mu_np * 8, MP)) != GI_OK)
        return rc;
    size_t iround = 0;

    for (auto &rg : rounds) {
        RoundCapture cap;
        if ((rc = cap.begin(c, P, S.stream)) != GI_OK) return rc;
        const int hcfg = round_halfls_cfg(rg.first, rg.second);
        const int TM = tile_rows(hcfg), TN = tile_cols(hcfg);
        if (P.vstream && (rc = vec_stage_chunk(c, S, P, rg.first, P.vs)) != GI_OK) return rc;
        std::vector<HalfJob> jobs;
        std::vector<GBuildJob> gjobs;
        // G from the record map: one unit per workgroup with its record in LDS
        // when the record fits and the map is at most 4 wide, else the
        // global-gather build (the same bits)
        std::vector<GMapJob> mjobs_v, mjobs_g;
        std::vector<int> gbegin, tbegin, mbegin, mubeg;
        // persistent round: round-local chunk of each half-ls / fold /
        // intensity job, and per chunk the tasks of each stage
        const int nch = (int)(rg.second - rg.first);
        std::vector<int> hch, fch, ich, ptasks(3 * (size_t)nch, 0), none;
        int munits = 0;
        int mt_tiles = 0;
        // global-gather builds four elements per thread when every such map of
        // the round is at most 4 wide
        bool m4 = true;
        for (size_t i = rg.first; i < rg.second; i++) {
            const GroupPlan &GP = group_of(c, chunks[i].jindI, chunks[i].gI, chunks[i].jindF);
            if (!(GP.gW <= 4 && c->lim_rec_lds > 0)) m4 = m4 && GP.gW <= 4;
        }
        bool mpart = false;  // some LDS-staged record longer than its LDS part
        const int mrows = m4 ? 1024 : 256;
        std::vector<CorrMMArgs> mm;
        std::vector<double *> Hc(rg.second - rg.first), Htc(rg.second - rg.first);
        int gt = 0, tt = 0;
        size_t og = 0, oh = 0, oht = 0;
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const JPlan &pI = c->jp[ch.jindI];
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
            const int ndegI = c->degen[ch.gI];
            SymDev r0 = symdev(pI.sym[ch.gI * GI_MAXDEG + 0]);
            SymDev r1 = ndegI > 1 ? symdev(pI.sym[ch.gI * GI_MAXDEG + 1]) : r0;
            Hc[i - rg.first] = H + oh;
            Htc[i - rg.first] = Ht + oht;
            if (ch.nunits > 0 && GP.Gtot_p > 0 && GP.gW <= kGMapMaxW && GP.d_gidx[0]) {
                GMapJob mj;
                // the compact map at the default routing (another rec_lds limit
                // selects the full map, for comparison)
                const bool cm = GP.cmap && c->lim_rec_lds == kRecLds;
                for (int d = 0; d < 2; d++) {
                    const int dd = d < ndegI ? d : 0;
                    mj.idx[d] = GP.d_gidx[dd];
                    mj.wt[d] = GP.d_gwt[dd];
                    mj.cidx[d] = cm ? GP.d_cidx[dd] : nullptr;
                    mj.ccode[d] = cm ? GP.d_ccode[dd] : nullptr;
                    mj.cwt[d] = cm ? GP.d_cwt[dd] : nullptr;
                }
                mj.W = GP.gW;
                mj.Gtot = GP.Gtot_p;
                mj.nunits = ch.nunits;
                mj.gtiles = (GP.Gtot_p + mrows - 1) / mrows;
                mj.uptr = (const double *const *)(d_meta + ch.off_uptr);
                mj.uideg = (const int *)(d_meta + ch.off_uideg);
                mj.G = G + og;
                mj.nsize = c->blocks[ch.jindI * c->nrep + ch.gI].nsize;
                mj.nlds = std::min(mj.nsize, std::max(c->lim_rec_lds, 0));
                if (GP.gW <= 4 && c->lim_rec_lds > 0) {
                    mpart = mpart || mj.nsize > mj.nlds;
                    mj.tile_begin = 0;
                    mjobs_v.push_back(mj);
                    mubeg.push_back(munits);
                    munits += ch.nunits;
                } else {
                    mj.tile_begin = mt_tiles;
                    mt_tiles += mj.gtiles * ((ch.nunits + kGUnits - 1) / kGUnits);
                    mjobs_g.push_back(mj);
                    mbegin.push_back(mj.tile_begin);
                }
            } else if (ch.nunits > 0 && GP.Gtot_p > 0) {
                GBuildJob gj;
                gj.classes = GP.d_cd;
                gj.nclasses = (int)GP.oc.size();
                gj.nunits = ch.nunits;
                gj.Gtot = GP.Gtot_p;
                gj.tiles_per_unit = (GP.Gtot_p + kGTile - 1) / kGTile;
                gj.tile_begin = gt;
                gt += gj.tiles_per_unit * ch.nunits;
                gj.uptr = (const double *const *)(d_meta + ch.off_uptr);
                gj.uideg = (const int *)(d_meta + ch.off_uideg);
                gj.r0 = r0; gj.r1 = r1;
                gj.G = G + og;
                gjobs.push_back(gj);
                gbegin.push_back(gj.tile_begin);
            }
            if (ch.nunits > 0)
                for (const OutClass &oc : GP.oc) {
                    HalfJob j;
                    j.cd = classdev(oc);
                    j.nunits = ch.nunits; j.Gtot = GP.Gtot_p; j.Htot = GP.Htot;
                    j.tile_begin = tt;
                    j.mt = (oc.n + TM - 1) / TM;
                    const int jt = j.mt * ((ch.nunits * (int)oc.cols.size() + TN - 1) / TN);
                    tt += jt;
                    hch.push_back((int)(i - rg.first));
                    ptasks[i - rg.first] += jt;
                    j.uptr = (const double *const *)(d_meta + ch.off_uptr);
                    j.uideg = (const int *)(d_meta + ch.off_uideg);
                    j.r0 = r0; j.r1 = r1;
                    j.G = G + og; j.H = H + oh;
                    j.m_lo = 0;
                    j.m_hi = oc.n;
                    jobs.push_back(j);
                    tbegin.push_back(j.tile_begin);
                }
            og += need[i].g; oh += need[i].h; oht += need[i].ht;
        }
        const int njobs = (int)jobs.size(), ngj = (int)gjobs.size(), nmj = (int)mjobs_v.size(),
                  nmg = (int)mjobs_g.size();
        // every table of the round is built on the host before its first launch
        // and goes to the device in one copy
        std::vector<FoldJob> fjobs;
        std::vector<int> fbegin;
        int fblocks = 0;
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const JPlan &pF = c->jp[ch.jindF];
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
            const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
            const int ndegI = c->degen[ch.gI], ndegF = c->degen[ch.gF];
            SymDev t0 = symdev(pF.sym[ch.gF * GI_MAXDEG + 0]);
            t0.tcol = GP.d_tcol[0];
            t0.tfc = GP.d_tfc[0];
            t0.tfw = GP.d_tfw[0];
            SymDev t1 = t0;
            if (ndegF > 1) {
                t1 = symdev(pF.sym[ch.gF * GI_MAXDEG + 1]);
                t1.tcol = GP.d_tcol[1];
                t1.tfc = GP.d_tfc[1];
                t1.tfw = GP.d_tfw[1];
            }
            if (ch.nitems > 0 && BF.nsize > 0) {
                FoldJob fj{ch.nitems, BF.nsize, ndegI, ndegF, GP.Htot,
                           (BF.nsize + kFoldRows - 1) / kFoldRows,
                           (const int *)(d_meta + ch.off_iunit), t0, t1, Hc[i - rg.first],
                           Htc[i - rg.first], ht_stride(BF.nsize)};
                fjobs.push_back(fj);
                fbegin.push_back(fblocks);
                fblocks += fj.xt * ch.nitems * ndegI * ndegF;
                fch.push_back((int)(i - rg.first));
                ptasks[nch + (i - rg.first)] += fj.xt * ch.nitems * ndegI * ndegF;
            }
            CorrMMArgs m;
            m.Lo = ch.Lo; m.Hi = ch.Hi; m.nitems = ch.nitems; m.ncb = ndegI * ndegF;
            m.ndegI = ndegI; m.nsize = BF.nsize;
            m.item_lo = (const int *)(d_meta + ch.off_lo);
            m.item_hi = (const int *)(d_meta + ch.off_hi);
            m.item_out_off = (const long long *)(d_meta + ch.off_out);
            m.blk_col = BF.d_col; m.vec = BF.d_vec; m.Ht = Htc[i - rg.first]; m.amp = cand.amp;
            if (ch.uhost) {  // staged upper records of this chunk's window
                m.vec = staged_rows(P.vs.ubuf[i & 1], ch.Lo, BF.nsize);
                m.blk_col = c->d_iota;
            }
            m.ldh = ht_stride(BF.nsize);
            // 16-byte loads of the records: an even record length and an aligned base
            m.vec_al = (BF.nsize % 2 == 0 && ((uintptr_t)m.vec & 15) == 0) ? 1 : 0;
            m.amp_m = c->amp_m;
            m.kparts = corr_kparts(c, P, ch);
            m.kchunk = (BF.nsize + 16 * m.kparts - 1) / (16 * m.kparts) * 16;
            m.ampx = P.ampx;
            m.ampx_stride = P.ampx_stride;
            mm.push_back(m);
            double blk_bytes = 0;
            for (const OutClass &oc : GP.oc) blk_bytes += 8.0 * oc.n * oc.K;
            P.work[3] += ch.nunits * 2.0 * coef_of(c, ch.jindI, ch.jindF).C * (double)N * N;
            P.work[4] += ch.nunits * 2.0 * GP.macs;
            P.work[5] += (double)ch.ncand * 8.0 * BF.nsize;
            P.work[16] += (double)ch.ncand * 2.0 * BF.nsize * __builtin_popcount(ch.cb_mask);
            P.work[6] += blk_bytes;
            P.work[7] += ch.nunits * 24.0 * (double)N * N;
            P.work[8] += band_exec_flops(P, ch, ndegI * ndegF, BF.nsize);
        }
        // corr tiles: per chunk ceil(rows / CM) x ceil(cols / CN), flattened
        const int ccfg = persist ? pccfg : c->tile_corr;
        const int CM = tile_rows(ccfg), CN = tile_cols(ccfg);
        std::vector<int> cbegin(mm.size());
        int ctiles = 0;
        for (size_t i = 0; i < mm.size(); i++) {
            cbegin[i] = ctiles;
            const int rows = mm[i].Hi - mm[i].Lo, cols = mm[i].nitems * mm[i].ncb;
            if (rows > 0 && cols > 0) {
                const int ct = ((rows + CM - 1) / CM) * ((cols + CN - 1) / CN) * mm[i].kparts;
                ctiles += ct;
                ptasks[2 * nch + i] = ct;
            }
        }
        std::vector<IntensArgs> ijobs;
        std::vector<int> ibegin;
        const int iblocks = round_intensity_jobs(c, P, rg.first, rg.second, d_meta, K, flag, cand,
                                                 ijobs, ibegin);
        for (size_t i = rg.first; i < rg.second; i++)
            if (chunks[i].ncand > 0) ich.push_back((int)(i - rg.first));
        TablePack T(rtab);
        const int *b_mv = nullptr, *b_mg = nullptr, *b_gj = nullptr, *b_hl = nullptr,
                  *b_fj = nullptr, *b_mm = nullptr, *b_ij = nullptr;
        const GMapJob *t_mv = T.add(mjobs_v, mubeg, &b_mv), *t_mg = T.add(mjobs_g, mbegin, &b_mg);
        const GBuildJob *t_gj = T.add(gjobs, gbegin, &b_gj);
        const HalfJob *t_hl = c->mu_streamed ? nullptr : T.add(jobs, tbegin, &b_hl);
        const FoldJob *t_fj = T.add(fjobs, fbegin, &b_fj);
        const CorrMMArgs *t_mm = T.add(mm, cbegin, &b_mm);
        const IntensArgs *t_ij = T.add(ijobs, ibegin, &b_ij);
        const int *t_hch = nullptr, *t_fch = nullptr, *t_ich = nullptr, *t_need = nullptr, *nb;
        if (persist) {
            t_hch = T.add(hch, none, &nb);
            t_fch = T.add(fch, none, &nb);
            t_ich = T.add(ich, none, &nb);
            t_need = T.add(ptasks, none, &nb);
        }
        if ((rc = send_pack(c, S, T, rtab_cap)) != GI_OK) return rc;
        if (nmj > 0) {
            TL_BEGIN("build_g");
            launch_build_g_map_lds(S.stream, t_mv, b_mv, nmj, munits, mpart);
            TL_END();
        }
        if (nmg > 0) {
            TL_BEGIN("build_g");
            if (m4)
                launch_build_g_map4(S.stream, t_mg, b_mg, nmg, mt_tiles);
            else
                launch_build_g_map(S.stream, t_mg, b_mg, nmg, mt_tiles);
            TL_END();
        }
        if (ngj > 0) {
            TL_BEGIN("build_g");
            launch_build_g_units(S.stream, N, t_gj, b_gj, ngj, gt);
            TL_END();
        }
        if (persist) {
            PersistRound pr;
            pr.src = MuSrc{c->d_mu_perm, (long long)N * N, N, 0};
            pr.hj = t_hl; pr.hbeg = b_hl; pr.hch = t_hch; pr.nhj = njobs; pr.nh = tt;
            pr.fj = t_fj; pr.fbeg = b_fj; pr.fch = t_fch; pr.nfj = (int)fjobs.size();
            pr.nf = fblocks;
            pr.cj = t_mm; pr.cbeg = b_mm; pr.ncj = (int)mm.size(); pr.nc = ctiles;
            pr.ij = t_ij; pr.ibeg = b_ij; pr.ich = t_ich; pr.nij = (int)ijobs.size();
            pr.ni = iblocks;
            pr.need = t_need;
            pr.nch = nch;
            pr.sync = psync;
            pr.err = P.perr;
            // counters and ticket cleared by a kernel (graph-capturable); the
            // error word once per call, in the first round
            launch_zero_ints(S.stream, psync, 3 * nch + 1);
            if (iround == 0) launch_zero_ints(S.stream, P.perr, 1);
            TL_BEGIN("round_persist");
            P.work[14] += tt;
            P.work[15] += ctiles;
            launch_round_persist(S.stream, hcfg, pccfg, pr, c->ncu);
            TL_END();
        }
        // the half-ls products: one launch over resident mu from the round's
        // tables, or one launch per mu part with its own table
        if (njobs > 0 && !c->mu_streamed && !persist) {
            TL_BEGIN("half_ls_mfma");
            P.work[14] += tt;
            launch_halfls_jobs(S.stream, hcfg, MuSrc{c->d_mu_perm, (long long)N * N, N, 0}, t_hl,
                               b_hl, njobs, tt);
            TL_END();
        }
        for (int v0 = 0; njobs > 0 && c->mu_streamed && v0 < N; v0 += c->mu_np) {
            const int np = std::min(c->mu_np, N - v0);
            std::vector<HalfJob> pj;
            std::vector<int> pb;
            const int pt = part_jobs(jobs, v0, np, TM, TN, pj, pb);
            if (pt == 0) continue;
            int b = 0;
            if ((rc = mu_part_copy(c, S, MP, c->h_mu_perm + v0, (size_t)N * 8, (size_t)np * 8,
                                   (size_t)3 * N, &b, tl)) != GI_OK)
                return rc;
            if ((rc = send_jobs(c, S, d_jobs, pj, pb, jobs_bytes)) != GI_OK) return rc;
            TL_BEGIN("half_ls_mfma");
            P.work[14] += pt;
            launch_halfls_jobs(S.stream, hcfg, MuSrc{MP.buf[b], (long long)N * np, np, v0},
                               (const HalfJob *)d_jobs, (const int *)(d_jobs + jobs_bytes),
                               (int)pj.size(), pt);
            TL_END();
            HIPCHK(hipEventRecord(S.mu_ev[2 + b], S.stream));
        }
        if (!fjobs.empty() && !persist) {
            TL_BEGIN("fold");
            launch_fold_jobs(S.stream, t_fj, b_fj, (int)fjobs.size(), fblocks);
            TL_END();
        }
        if (!persist) {
            TL_BEGIN("corr_mfma");
            P.work[15] += ctiles;
            launch_corr_mfma(S.stream, c->tile_corr, t_mm, b_mm, (int)mm.size(), ctiles);
            TL_END();
        }
        if (P.vstream && (rc = vec_stage_done(c, S, P, rg.first)) != GI_OK) return rc;
        if (!ijobs.empty() && !persist) {
            TL_BEGIN("intensity");
            launch_intensity_jobs(S.stream, t_ij, b_ij, (int)ijobs.size(), iblocks);
            TL_END();
        }
        if ((rc = round_compact(c, S, P, iround, flag, cand, pc_counts, pc_offs, tl)) != GI_OK)
            return rc;
        if ((rc = cap.end(c, P)) != GI_OK) return rc;
        iround++;
    }
    return GI_OK;
}

// Split-precision mode: the batched rounds with both products on the int8
// matrix cores (gi_split.hip). Per round: G digits of every unit, the
// half-ls products per (chunk, class), FP64 folds, digits of the folded
// vectors, the banded correlation, then the FP64 intensity step.
static int run_split(gi_ctx *c, Slot &S, Plan &P, unsigned char *d_meta, Timeline *tl,
                     int *flag, LineSoA cand)
{
    const int N = c->nvib, SD = c->split_S, SH = c->split_Sh;
    const Consts K;
    int rc = GI_OK;
    const auto &chunks = P.chunks;
    struct Need {
        size_t gd, ge, h, ht, hd, hc;
        int njobs;
    };
    std::vector<Need> need(chunks.size());
    std::vector<size_t> bytes(chunks.size());
    for (size_t i = 0; i < chunks.size(); i++) {
        const Chunk &ch = chunks[i];
        const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
        const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
        const size_t cols = (size_t)ch.nitems * c->degen[ch.gI] * c->degen[ch.gF];
        if (GP.gW > kGMapMaxW || !GP.d_gidxB[0])
            return fail(c, GI_EINVAL, "split mode: record map wider than the kernel supports");
        need[i] = Need{(size_t)ch.nunits * GP.GtotB, (size_t)ch.nunits * GP.EGtot,
                       (size_t)ch.nunits * GP.Htot, cols * BF.nsize, cols * BF.Kp, cols,
                       ch.nunits > 0 ? (int)GP.oc.size() : 0};
        bytes[i] = SH * need[i].gd + 4 * need[i].ge + 8 * need[i].h + 8 * need[i].ht +
                   SD * need[i].hd + 4 * need[i].hc;
    }
    const auto rounds = plan_rounds(c, P, bytes);
    size_t mgd = 64, mge = 1, mh = 1, mht = 1, mhd = 64, mhc = 1;
    int mjobs = 1, mchunks = 1;
    for (auto &rg : rounds) {
        Need t{0, 0, 0, 0, 0, 0, 0};
        for (size_t i = rg.first; i < rg.second; i++) {
            t.gd += need[i].gd; t.ge += need[i].ge; t.h += need[i].h; t.ht += need[i].ht;
            t.hd += need[i].hd; t.hc += need[i].hc; t.njobs += need[i].njobs;
        }
        mgd = std::max(mgd, t.gd); mge = std::max(mge, t.ge); mh = std::max(mh, t.h);
        mht = std::max(mht, t.ht); mhd = std::max(mhd, t.hd); mhc = std::max(mhc, t.hc);
        mjobs = std::max(mjobs, t.njobs);
        mchunks = std::max(mchunks, (int)(rg.second - rg.first));
    }
    int8_t *Gd = (int8_t *)slot_scratch(c, S, "Gd", SH * mgd, &rc); if (rc) return rc;
    int *Ge = (int *)slot_scratch(c, S, "Ge", 4 * mge, &rc); if (rc) return rc;
    double *H = (double *)slot_scratch(c, S, "H", 8 * mh, &rc); if (rc) return rc;
    double *Ht = (double *)slot_scratch(c, S, "Ht", 8 * mht, &rc); if (rc) return rc;
    int8_t *Hd = (int8_t *)slot_scratch(c, S, "Hd", SD * mhd, &rc); if (rc) return rc;
    int *He = (int *)slot_scratch(c, S, "He", 4 * mhc, &rc); if (rc) return rc;
    // per-mu-part half-ls tables (streamed mu); the round's other tables go
    // to rtab in one copy
    const size_t hj_bytes = (size_t)mjobs * sizeof(SplitHalfJob);
    unsigned char *d_hj = nullptr;
    if (c->mu_streamed) {
        d_hj = (unsigned char *)slot_scratch(c, S, "sjobs", hj_bytes + 4 * mjobs, &rc);
        if (rc) return rc;
    }
    const size_t mc = (size_t)mchunks;
    // (the fused fold-split jobs come in three tables, by column length, whose
    // jobs together are at most mc: the zeros reserve the others' alignment)
    const size_t rtab_cap = pack_cap(
        {mc * (sizeof(GSplitJob) + 4), mc * (sizeof(GSplitJob) + 4),
         (size_t)mjobs * (sizeof(SplitHalfJob) + 4), mc * (sizeof(FoldSplitJob) + 4), 0, 0,
         mc * (sizeof(FoldJob) + 4), mc * (sizeof(SplitRowsJob) + 4),
         mc * (sizeof(CorrSplitArgs) + 4), mc * (sizeof(IntensArgs) + 4)});
    unsigned char *rtab = (unsigned char *)slot_scratch(c, S, "rtab", rtab_cap, &rc);
    if (rc) return rc;
    // fold and digit split fused when every upper block's column fits one
    // workgroup (GI_FOLD_SPLIT=0 forces the two-kernel path; same bits)
    const char *fenv = std::getenv("GI_FOLD_SPLIT");
    bool fused = !(fenv && fenv[0] == '0');
    int maxkp = 0;
    for (const Chunk &ch : chunks) {
        maxkp = std::max(maxkp, c->blocks[ch.jindF * c->nrep + ch.gF].Kp);
        fused = fused && c->blocks[ch.jindF * c->nrep + ch.gF].Kp <= c->lim_fold_split;
    }
    int maxparts = 1;
    for (const Chunk &ch : chunks) maxparts = std::max(maxparts, corr_kparts(c, P, ch));
    P.ampx_stride = (long long)c->amp_m * c->amp_m * std::max<long long>(P.ncand, 1);
    P.ampx = maxparts > 1 ? (double *)slot_scratch(c, S, "c_ampx",
                                                   8 * (size_t)(maxparts - 1) * P.ampx_stride, &rc)
                          : nullptr;
    if (rc) return rc;
    int *pc_counts = nullptr;
    long long *pc_offs = nullptr;
    if ((rc = setup_pipe(c, S, P, rounds, &pc_counts, &pc_offs)) != GI_OK) return rc;

    // G digits of a chunk: one unit per workgroup with its record in LDS (the
    // first lim_rec_lds doubles of a longer one) when its classes fit the
    // kernel's table and its map is at most 4 wide; else the global-gather
    // build (four rows per thread when every such map of the round is at most 4
    // wide). The builds give the same bits.
    const auto g_lds = [&](const Chunk &ch) {
        const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
        return c->lim_rec_lds > 0 && GP.oc.size() <= (size_t)kClsLds && GP.gW <= 4;
    };
    // split half-ls tile of round [c0, c1): the setting, or with kSplitTileAuto
    // 128 x 64 on eight wavefronts when the round's tiles average at least
    // kAutoSplitSteps K steps, else 64 x 64 (measured, half-ls per pass: C3v J
    // 0..15, 129 steps a tile, 524 ms against 597 ms; C2v J 10<->11 and J
    // 0..20, 23 steps, 1.52 and 34.6 ms against 1.41 and 32.5 ms: one
    // workgroup per CU leaves the short tiles' set-up and stores exposed)
    auto round_split_cfg = [&](size_t c0, size_t c1) {
        if (c->tile_split != kSplitTileAuto) return c->tile_split;
        double tiles = 0, steps = 0;
        for (size_t i = c0; i < c1; i++) {
            const Chunk &ch = chunks[i];
            if (ch.nunits <= 0) continue;
            for (const OutClass &oc : group_of(c, ch.jindI, ch.gI, ch.jindF).oc) {
                const double t = (double)((oc.n + 63) / 64) *
                                 ((ch.nunits * (int)oc.cols.size() + 63) / 64);
                tiles += t;
                steps += t * split2_steps(oc.seg_n.data(), (int)oc.seg_n.size());
            }
        }
        return tiles > 0 && steps >= kAutoSplitSteps * tiles ? 3 : 2;
    };
    MuParts MP;
    if (c->mu_streamed &&
        (rc = mu_parts_setup(c, S, (size_t)SH * 3 * c->mu_np * c->kpad_all, MP)) != GI_OK)
        return rc;
    size_t iround = 0;
    for (auto &rg : rounds) {
        RoundCapture cap;
        if ((rc = cap.begin(c, P, S.stream)) != GI_OK) return rc;
        // blocks in host memory: one chunk per round, its records staged first
        if (P.vstream && (rc = vec_stage_chunk(c, S, P, rg.first, P.vs)) != GI_OK) return rc;
        std::vector<GSplitJob> gj, gjg;  // G builds in LDS, global-gather G builds
        std::vector<SplitHalfJob> hj;
        std::vector<CorrSplitArgs> cj;
        std::vector<int> gbegin, gbegin_g, tbegin, cbegin;
        int gt = 0, tt = 0, ct = 0, ubeg = 0;
        bool g4 = true;  // every global-gather map of the round at most 4 wide
        bool gpart = false;  // some LDS-staged record longer than its LDS part
        for (size_t i = rg.first; i < rg.second; i++)
            if (!g_lds(chunks[i])) g4 = g4 && group_of(c, chunks[i].jindI, chunks[i].gI,
                                                       chunks[i].jindF).gW <= 4;
        const int grows = g4 ? 1024 : 256;
        bool steps_fit = true;  // every class within k_halfls_split2's step table
        // row tile of the split half-ls kernels: k_halfls_split, k_halfls_split2
        // on 4 wavefronts 64; k_halfls_split_wide, k_halfls_split2 on 8 wavefronts 128
        const int rcfg = round_split_cfg(rg.first, rg.second);
        const int srows_tile = rcfg == 1 || rcfg == 3 ? 128 : 64;
        size_t ogd = 0, oge = 0, oh = 0, oht = 0, ohd = 0, ohc = 0;
        std::vector<size_t> Hoff, Htoff, Hdoff, Hcoff;
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
            const Block &BI = c->blocks[ch.jindI * c->nrep + ch.gI];
            const int ndegI = c->degen[ch.gI];
            Hoff.push_back(oh); Htoff.push_back(oht); Hdoff.push_back(ohd); Hcoff.push_back(ohc);
            if (ch.nunits > 0 && GP.GtotB > 0) {
                GSplitJob g;
                g.classes = GP.d_cd;
                g.nclasses = (int)GP.oc.size();
                for (int d = 0; d < 2; d++) {
                    const int dd = d < ndegI ? d : 0;
                    g.idx[d] = GP.d_gidxB[dd];
                    g.wt[d] = GP.d_gwtB[dd];
                    g.bnd[d] = GP.d_gbnd[dd];
                }
                g.W = GP.gW;
                g.GtotB = GP.GtotB;
                g.EGtot = GP.EGtot;
                g.nunits = ch.nunits;
                g.gtiles = (GP.GtotB + grows - 1) / grows;
                g.tile_begin = gt;
                if (!g_lds(ch)) gt += g.gtiles * ((ch.nunits + kGUnits - 1) / kGUnits);
                g.uptr = (const double *const *)(d_meta + ch.off_uptr);
                g.uideg = (const int *)(d_meta + ch.off_uideg);
                g.urec = (const int *)(d_meta + ch.off_urec);
                g.umax = BI.d_umax;
                g.D = Gd + ogd;
                g.plane = (long long)mgd;
                g.E = Ge + oge;
                g.nsize = BI.nsize;
                g.nlds = std::min(BI.nsize, std::max(c->lim_rec_lds, 0));
                if (g_lds(ch)) {
                    gpart = gpart || g.nsize > g.nlds;
                    gj.push_back(g);
                    gbegin.push_back(ubeg);
                    ubeg += ch.nunits;
                } else {
                    gjg.push_back(g);
                    gbegin_g.push_back(g.tile_begin);
                }
                for (const OutClass &oc : GP.oc) {
                    steps_fit = steps_fit &&
                                split2_class_fits(oc.seg_n.data(), (int)oc.seg_n.size()) &&
                                split2_steps(oc.seg_n.data(), (int)oc.seg_n.size()) <=
                                    c->lim_split2_steps &&
                                (int)oc.seg_n.size() <= c->lim_split2_segs;
                    SplitHalfJob h;
                    h.cd = classdev(oc);
                    h.nunits = ch.nunits;
                    h.tile_begin = tt;
                    h.mt = (oc.n + srows_tile - 1) / srows_tile;
                    tt += h.mt * ((ch.nunits * (int)oc.cols.size() + 63) / 64);
                    h.GtotB = GP.GtotB;
                    h.EGtot = GP.EGtot;
                    h.Htot = GP.Htot;
                    h.Gd = Gd + ogd;
                    h.gplane = (long long)mgd;
                    h.Ge = Ge + oge;
                    h.H = H + oh;
                    h.m_lo = 0;
                    h.m_hi = oc.n;
                    hj.push_back(h);
                    tbegin.push_back(h.tile_begin);
                }
            }
            ogd += need[i].gd; oge += need[i].ge; oh += need[i].h;
            oht += need[i].ht; ohd += need[i].hd; ohc += need[i].hc;
        }
        // every table of the round is built on the host before its first launch
        // and goes to the device in one copy
        std::vector<FoldJob> fjobs;
        std::vector<SplitRowsJob> sjobs;
        // fused fold-split jobs by column length (one launch per class): at most
        // kFoldSplitSmall rows, at most 16,384, longer
        std::vector<FoldSplitJob> fsjobs[3];
        std::vector<int> fbegin, sbegin, fsbegin[3];
        int fblocks = 0, srows = 0, fscols[3] = {0, 0, 0};
        for (size_t i = rg.first; i < rg.second; i++) {
            const Chunk &ch = chunks[i];
            const size_t k = i - rg.first;
            const JPlan &pF = c->jp[ch.jindF];
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
            const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
            const int ndegI = c->degen[ch.gI], ndegF = c->degen[ch.gF];
            SymDev t0 = symdev(pF.sym[ch.gF * GI_MAXDEG + 0]);
            t0.tcol = GP.d_tcol[0];
            t0.tfc = GP.d_tfc[0];
            t0.tfw = GP.d_tfw[0];
            SymDev t1 = t0;
            if (ndegF > 1) {
                t1 = symdev(pF.sym[ch.gF * GI_MAXDEG + 1]);
                t1.tcol = GP.d_tcol[1];
                t1.tfc = GP.d_tfc[1];
                t1.tfw = GP.d_tfw[1];
            }
            const int cols = ch.nitems * ndegI * ndegF;
            if (cols > 0 && BF.nsize > 0) {
                FoldJob fj{ch.nitems, BF.nsize, ndegI, ndegF, GP.Htot,
                           (BF.nsize + kFoldRows - 1) / kFoldRows,
                           (const int *)(d_meta + ch.off_iunit), t0, t1, H + Hoff[k],
                           Ht + Htoff[k]};
                fjobs.push_back(fj);
                fbegin.push_back(fblocks);
                fblocks += fj.xt * cols;
                sjobs.push_back(SplitRowsJob{Ht + Htoff[k], (long long)BF.nsize, BF.nsize, BF.Kp,
                                             Hd + Hdoff[k], (long long)mhd, He + Hcoff[k]});
                const int fc = BF.Kp <= kFoldSplitSmall ? 0 : BF.Kp <= 16384 ? 1 : 2;
                fsjobs[fc].push_back(
                    FoldSplitJob{fj, BF.Kp, Hd + Hdoff[k], (long long)mhd, He + Hcoff[k]});
                fsbegin[fc].push_back(fscols[fc]);
                fscols[fc] += cols;
                sbegin.push_back(srows);
                srows += cols;
            }
            CorrSplitArgs a;
            a.Lo = ch.Lo; a.Hi = ch.Hi; a.nitems = ch.nitems; a.ncb = ndegI * ndegF;
            a.ndegI = ndegI; a.Kp = BF.Kp;
            a.item_lo = (const int *)(d_meta + ch.off_lo);
            a.item_hi = (const int *)(d_meta + ch.off_hi);
            a.item_out_off = (const long long *)(d_meta + ch.off_out);
            a.blk_col = BF.d_col;
            a.Ud = BF.d_ud; a.uplane = BF.uplane; a.Ue = BF.d_ue;
            if (ch.uhost) {  // staged digit rows of this chunk's window, identity map
                const int bsel = (int)(i & 1);
                a.Ud = P.vs.ubd[bsel] - (long long)ch.Lo * BF.Kp;
                a.uplane = P.vs.uplane;
                a.Ue = P.vs.ube[bsel] - ch.Lo;
                a.blk_col = c->d_iota;
            }
            a.Hd = Hd + Hdoff[k]; a.hplane = (long long)mhd; a.He = He + Hcoff[k];
            a.amp = cand.amp;
            a.amp_m = c->amp_m;
            a.kparts = corr_kparts(c, P, ch);
            a.kchunk = (BF.Kp + 64 * a.kparts - 1) / (64 * a.kparts) * 64;
            a.ampx = P.ampx;
            a.ampx_stride = P.ampx_stride;
            cj.push_back(a);
            cbegin.push_back(ct);
            const int rows = ch.Hi - ch.Lo;
            if (rows > 0 && cols > 0) ct += ((rows + 63) / 64) * ((cols + 63) / 64) * a.kparts;
            double blk_bytes = 0;
            for (const OutClass &oc : GP.oc) blk_bytes += 8.0 * oc.n * oc.K;
            P.work[3] += ch.nunits * 2.0 * coef_of(c, ch.jindI, ch.jindF).C * (double)N * N;
            P.work[4] += ch.nunits * 2.0 * GP.macs;
            P.work[5] += (double)ch.ncand * 8.0 * BF.nsize;
            P.work[16] += (double)ch.ncand * 2.0 * BF.nsize * __builtin_popcount(ch.cb_mask);
            P.work[6] += blk_bytes;
            P.work[7] += ch.nunits * 24.0 * (double)N * N;
            P.work[8] += band_exec_flops(P, ch, ndegI * ndegF, BF.nsize);
        }
        std::vector<IntensArgs> ijobs;
        std::vector<int> ibegin;
        const int iblocks = round_intensity_jobs(c, P, rg.first, rg.second, d_meta, K, flag, cand,
                                                 ijobs, ibegin);
        TablePack T(rtab);
        const int *b_gj = nullptr, *b_gg = nullptr, *b_hj = nullptr,
                  *b_fj = nullptr, *b_sj = nullptr, *b_cj = nullptr, *b_ij = nullptr;
        const GSplitJob *t_gj = T.add(gj, gbegin, &b_gj), *t_gg = T.add(gjg, gbegin_g, &b_gg);
        const SplitHalfJob *t_hj = c->mu_streamed ? nullptr : T.add(hj, tbegin, &b_hj);
        const FoldSplitJob *t_fs[3] = {nullptr, nullptr, nullptr};
        const int *b_fsc[3] = {nullptr, nullptr, nullptr};
        const FoldJob *t_fj = nullptr;
        const SplitRowsJob *t_sj = nullptr;
        if (fused) {
            for (int fc = 0; fc < 3; fc++) t_fs[fc] = T.add(fsjobs[fc], fsbegin[fc], &b_fsc[fc]);
        } else {
            t_fj = T.add(fjobs, fbegin, &b_fj);
            t_sj = T.add(sjobs, sbegin, &b_sj);
        }
        const CorrSplitArgs *t_cj = T.add(cj, cbegin, &b_cj);
        const IntensArgs *t_ij = T.add(ijobs, ibegin, &b_ij);
        if ((rc = send_pack(c, S, T, rtab_cap)) != GI_OK) return rc;
        if (!gj.empty()) {
            TL_BEGIN("build_g");
            launch_build_g_split_lds(S.stream, SH, c->split_radix, t_gj, b_gj, (int)gj.size(), ubeg,
                                     gpart);
            TL_END();
        }
        if (!gjg.empty()) {
            TL_BEGIN("build_g");
            launch_build_g_split(S.stream, SH, c->split_radix, t_gg, b_gg, (int)gjg.size(), gt, g4);
            TL_END();
        }
        // the half-ls products: one launch over resident mu digits from the
        // round's tables, or one launch per mu part with its own table
        // classes beyond k_halfls_split2's step table: the one-set kernel of the same rows
        const int scfg = rcfg >= 2 && !steps_fit ? rcfg - 2 : rcfg;
        if (!hj.empty() && !c->mu_streamed) {
            TL_BEGIN("half_ls_split");
            P.work[14] += tt;
            launch_halfls_split(S.stream, SH, c->split_radix, c->kpad_all,
                                MuSrc{c->d_mud, c->muplane, N, 0}, c->d_mue, t_hj, b_hj,
                                (int)hj.size(), tt, scfg);
            TL_END();
        }
        for (int v0 = 0; !hj.empty() && c->mu_streamed && v0 < N; v0 += c->mu_np) {
            const int np = std::min(c->mu_np, N - v0);
            const size_t kp = (size_t)c->kpad_all;
            std::vector<SplitHalfJob> pj;
            std::vector<int> pb;
            const int pt = part_jobs(hj, v0, np, srows_tile, 64, pj, pb);
            if (pt == 0) continue;
            int b = 0;
            if ((rc = mu_part_copy(c, S, MP, c->h_mud + v0 * kp, (size_t)N * kp, (size_t)np * kp,
                                   (size_t)3 * SH, &b, tl)) != GI_OK)
                return rc;
            if ((rc = send_jobs(c, S, d_hj, pj, pb, hj_bytes)) != GI_OK) return rc;
            TL_BEGIN("half_ls_split");
            P.work[14] += pt;
            launch_halfls_split(S.stream, SH, c->split_radix, c->kpad_all,
                                MuSrc{MP.buf[b], 3LL * np * (long long)kp, np, v0}, c->d_mue,
                                (const SplitHalfJob *)d_hj, (const int *)(d_hj + hj_bytes),
                                (int)pj.size(), pt, scfg);
            TL_END();
            HIPCHK(hipEventRecord(S.mu_ev[2 + b], S.stream));
        }
        if (!fjobs.empty() && fused) {
            TL_BEGIN("fold_split");
            const int fmax[3] = {kFoldSplitSmall, 16384, maxkp};
            for (int fc = 0; fc < 3; fc++)
                launch_fold_split_jobs(S.stream, SD, c->split_radix, t_fs[fc], b_fsc[fc],
                                       (int)fsjobs[fc].size(), fscols[fc], fmax[fc]);
            TL_END();
        } else if (!fjobs.empty()) {
            TL_BEGIN("fold");
            launch_fold_jobs(S.stream, t_fj, b_fj, (int)fjobs.size(), fblocks);
            TL_END();
            TL_BEGIN("split_ht");
            launch_split_rows_jobs(S.stream, SD, c->split_radix, t_sj, b_sj, (int)sjobs.size(),
                                   srows);
            TL_END();
        }
        if (!cj.empty()) {
            TL_BEGIN("corr_split");
            P.work[15] += ct;
            launch_corr_split(S.stream, SD, c->split_radix, t_cj, b_cj, (int)cj.size(), ct);
            TL_END();
        }
        if (P.vstream && (rc = vec_stage_done(c, S, P, rg.first)) != GI_OK) return rc;
        if (!ijobs.empty()) {
            TL_BEGIN("intensity");
            launch_intensity_jobs(S.stream, t_ij, b_ij, (int)ijobs.size(), iblocks);
            TL_END();
        }
        if ((rc = round_compact(c, S, P, iround, flag, cand, pc_counts, pc_offs, tl)) != GI_OK)
            return rc;
        if ((rc = cap.end(c, P)) != GI_OK) return rc;
        iround++;
    }
    return GI_OK;
}

static LineSoA shift(LineSoA s, long long o)
{
    s.iI += o; s.iF += o; s.nu += o; s.S += o; s.A += o; s.I += o; s.amp += s.amp_w * o;
    return s;
}

static int setup_query_params(gi_ctx *c, Slot &S, Plan &P);

// Enqueue the whole plan on slot S: half linestrengths, correlations,
// intensities, per-query compaction, and the copy of the per-query line
// counts to pinned memory. Nothing here waits for the device.
static int launch_plan(gi_ctx *c, Slot &S, Plan &P, Timeline *tl)
{
    const int N = c->nvib;
    const Consts K;
    int rc = GI_OK;
    HIPCHK(hipSetDevice(c->device));
    S.arena.reset();  // the slot's previous sequence has finished
    const bool batched = P.mode == GI_MODE_FP64_BATCHED;
    const size_t nc = (size_t)std::max<long long>(P.ncand, 1);
    int *flag = (int *)slot_scratch(c, S, "flag", nc * sizeof(int), &rc); if (rc) return rc;
    LineSoA cand, comp;
    cand.iI = (int *)slot_scratch(c, S, "c_iI", nc * 4, &rc); if (rc) return rc;
    cand.iF = (int *)slot_scratch(c, S, "c_iF", nc * 4, &rc); if (rc) return rc;
    cand.nu = (double *)slot_scratch(c, S, "c_nu", nc * 8, &rc); if (rc) return rc;
    cand.S = (double *)slot_scratch(c, S, "c_S", nc * 8, &rc); if (rc) return rc;
    cand.A = (double *)slot_scratch(c, S, "c_A", nc * 8, &rc); if (rc) return rc;
    cand.I = (double *)slot_scratch(c, S, "c_I", nc * 8, &rc); if (rc) return rc;
    const int W = c->amp_m * c->amp_m;
    cand.amp = (double *)slot_scratch(c, S, "c_amp", nc * 8 * W, &rc); if (rc) return rc;
    cand.amp_w = W;
    comp.iI = (int *)slot_scratch(c, S, "o_iI", nc * 4, &rc); if (rc) return rc;
    comp.iF = (int *)slot_scratch(c, S, "o_iF", nc * 4, &rc); if (rc) return rc;
    comp.nu = (double *)slot_scratch(c, S, "o_nu", nc * 8, &rc); if (rc) return rc;
    comp.S = (double *)slot_scratch(c, S, "o_S", nc * 8, &rc); if (rc) return rc;
    comp.A = (double *)slot_scratch(c, S, "o_A", nc * 8, &rc); if (rc) return rc;
    comp.I = (double *)slot_scratch(c, S, "o_I", nc * 8, &rc); if (rc) return rc;
    comp.amp = (double *)slot_scratch(c, S, "o_amp", nc * 8 * W, &rc); if (rc) return rc;
    comp.amp_w = W;
    P.comp = comp;
    P.cand = cand;
    // compaction scratch: per query, block counts and offsets
    std::vector<size_t> cbase(P.q.size()), obase(P.q.size());
    size_t ncnt = 0, noff = 0;
    for (size_t k = 0; k < P.q.size(); k++) {
        const int nb = compact_blocks(P.q[k].cand_end - P.q[k].cand_begin);
        cbase[k] = ncnt;
        obase[k] = noff;
        ncnt += nb;
        noff += nb + 1;
    }
    int *counts = (int *)slot_scratch(c, S, "counts", std::max<size_t>(ncnt, 1) * 4, &rc);
    if (rc) return rc;
    long long *offs = (long long *)slot_scratch(c, S, "offs", std::max<size_t>(noff, 1) * 8, &rc);
    if (rc) return rc;
    if (!P.pipe) {
        P.counts = (long long *)slot_pinned(c, S, P.q.size() * 8 + 64, &rc);
        if (rc) return rc;
    }
    // staging of blocks in host memory (patches the unit tables before the upload)
    if ((rc = vec_stage_setup(c, S, P, P.vs)) != GI_OK) return rc;
    unsigned char *d_meta = nullptr;
    if (!P.meta.empty()) {
        d_meta = (unsigned char *)slot_scratch(c, S, "meta", P.meta.size(), &rc); if (rc) return rc;
        unsigned char *hm = stage(c, S, P.meta.data(), P.meta.size(), &rc);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(d_meta, hm, P.meta.size(), hipMemcpyHostToDevice, S.stream));
    }

    if (P.device_q && (rc = setup_query_params(c, S, P)) != GI_OK) return rc;
    if (P.mode == GI_MODE_SPLIT && !P.chunks.empty()) {
        if ((rc = run_split(c, S, P, d_meta, tl, flag, cand)) != GI_OK) return rc;
    } else if (batched && !P.chunks.empty()) {
        if ((rc = run_batched(c, S, P, d_meta, tl, flag, cand)) != GI_OK) return rc;
    } else if (!P.chunks.empty()) {
        int max_units = 1, max_items = 1, max_G = 1, max_H = 1, max_ncb = 1, max_nsizeF = 1;
        for (auto &ch : P.chunks) {
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
            max_units = std::max(max_units, ch.nunits);
            max_items = std::max(max_items, ch.nitems);
            max_G = std::max(max_G, GP.Gtot);
            max_H = std::max(max_H, GP.Htot);
            max_ncb = std::max(max_ncb, c->degen[ch.gI] * c->degen[ch.gF]);
            max_nsizeF = std::max(max_nsizeF, c->blocks[ch.jindF * c->nrep + ch.gF].nsize);
        }
        double *G = (double *)slot_scratch(c, S, "G", (size_t)max_units * max_G * 8, &rc);
        if (rc) return rc;
        double *H = (double *)slot_scratch(c, S, "H", (size_t)max_units * max_H * 8, &rc);
        if (rc) return rc;
        double *Ht = (double *)slot_scratch(c, S, "Ht", (size_t)max_items * max_ncb * max_nsizeF * 8,
                                            &rc);
        if (rc) return rc;
        // streamed mu: two device parts of mu_np vF rows (every a and vI); the
        // copy stream fills one while the half linestrength reads the other
        double *mupart[2] = {nullptr, nullptr};
        hipStream_t cs = S.cstream ? S.cstream : S.stream;
        if (c->mu_streamed) {
            for (int b = 0; b < 2; b++) {
                mupart[b] = (double *)slot_scratch(c, S, b ? "mupart1" : "mupart0",
                                                   (size_t)3 * N * c->mu_np * 8, &rc);
                if (rc) return rc;
            }
            for (auto &e : S.mu_ev)
                if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                    return fail(c, GI_EHIP, "hipEventCreate");
        }
        int nparts = 0;  // parts copied so far; part q goes to buffer q & 1
        for (size_t ci = 0; ci < P.chunks.size(); ci++) {
            const Chunk &ch = P.chunks[ci];
            if (P.vstream && (rc = vec_stage_chunk(c, S, P, ci, P.vs)) != GI_OK) return rc;
            const JPlan &pI = c->jp[ch.jindI], &pF = c->jp[ch.jindF];
            const GroupPlan &GP = group_of(c, ch.jindI, ch.gI, ch.jindF);
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
            if (ch.nunits > 0)
                for (const OutClass &oc : GP.oc) {
                    const ClassDev cd = classdev(oc);
                    if (cd.K > 0) {
                        TL_BEGIN("build_g");
                        launch_build_g_sym(S.stream, ch.nunits, N, GP.Gtot, cd,
                                           (const double *const *)(d_meta + ch.off_uptr),
                                           (const int *)(d_meta + ch.off_uideg), r0, r1, G);
                        TL_END();
                    }
                    if (c->mu_streamed) continue;  // after every class's G, by parts below
                    TL_BEGIN("half_ls_gemv");
                    if (launch_gemv_sym(S.stream, ch.nunits, N, GP.Gtot, GP.Htot, cd,
                                        c->d_mu_perm, G, H) != 0)
                        return fail(c, GI_EINVAL, "no half-ls kernel for this column count");
                    TL_END();
                }
            // Streamed mu (do_1st_half_ls_blocks, cuda_host.cu:783-817): each
            // part holds vF rows [v0, v0 + np) of mu; the rows of every class
            // inside it are computed from it whole (the same sums in the same
            // order as from resident mu).
            for (int v0 = 0; c->mu_streamed && ch.nunits > 0 && v0 < N; v0 += c->mu_np) {
                const int np = std::min(c->mu_np, N - v0);
                bool used = false;
                for (const OutClass &oc : GP.oc)
                    used = used || (oc.off < v0 + np && oc.off + oc.n > v0);
                if (!used) continue;
                const int b = nparts++ & 1;
                // the buffer is free once the products of the part before last are done
                if (nparts > 2) HIPCHK(hipStreamWaitEvent(cs, S.mu_ev[2 + b], 0));
                TL_BEGIN_ON("mu_copy", cs);
                HIPCHK(hipMemcpy2DAsync(mupart[b], (size_t)np * 8, c->h_mu_perm + v0,
                                        (size_t)N * 8, (size_t)np * 8, (size_t)3 * N,
                                        hipMemcpyHostToDevice, cs));
                TL_END_ON(cs);
                HIPCHK(hipEventRecord(S.mu_ev[b], cs));
                HIPCHK(hipStreamWaitEvent(S.stream, S.mu_ev[b], 0));
                for (const OutClass &oc : GP.oc) {
                    const ClassDev cd = classdev(oc);
                    const int rlo = std::max(0, v0 - cd.off), rhi = std::min(cd.n, v0 + np - cd.off);
                    if (rlo >= rhi) continue;
                    TL_BEGIN("half_ls_gemv");
                    if (launch_gemv_sym(S.stream, ch.nunits, N, GP.Gtot, GP.Htot, cd, mupart[b],
                                        G, H, np, v0, rlo, rhi) != 0)
                        return fail(c, GI_EINVAL, "no half-ls kernel for this column count");
                    TL_END();
                }
                HIPCHK(hipEventRecord(S.mu_ev[2 + b], S.stream));
            }
            TL_BEGIN("fold");
            launch_fold(S.stream, ch.nitems, BF.nsize, ndegI, ndegF, GP.Htot,
                        (const int *)(d_meta + ch.off_iunit), t0, t1, H, Ht);
            TL_END();
            CorrArgs a;
            fill_corr_args(c, P, ch, d_meta, K, a);
            if (ch.uhost) {  // staged upper records of this chunk's window
                a.vec = staged_rows(P.vs.ubuf[ci & 1], ch.Lo, BF.nsize);
                a.blk_col = c->d_iota;
            }
            a.Ht = Ht;
            a.flag = flag; a.out_iI = cand.iI; a.out_iF = cand.iF; a.out_nu = cand.nu;
            a.out_S = cand.S; a.out_A = cand.A; a.out_I = cand.I; a.out_amp = cand.amp;
            a.amp_m = c->amp_m;
            TL_BEGIN("corr_intensity");
            launch_corr_intensity(S.stream, a);
            TL_END();
            if (P.vstream && (rc = vec_stage_done(c, S, P, ci)) != GI_OK) return rc;
            double blk_bytes = 0;
            for (const OutClass &oc : GP.oc) blk_bytes += 8.0 * oc.n * oc.K;
            P.work[3] += ch.nunits * 2.0 * coef_of(c, ch.jindI, ch.jindF).C * (double)N * N;
            P.work[4] += ch.nunits * 2.0 * GP.macs;
            P.work[5] += (double)ch.ncand * 8.0 * BF.nsize;
            P.work[16] += (double)ch.ncand * 2.0 * BF.nsize * __builtin_popcount(ch.cb_mask);
            P.work[6] += ch.nunits * blk_bytes;
            P.work[7] += ch.nunits * 24.0 * (double)N * N;
            P.work[8] += 2.0 * ch.ncand * ndegI * ndegF * BF.nsize;
        }
    }
    HIPCHK(hipGetLastError());
    if (P.pipe) {  // the rounds compacted their own ranges
        for (auto &ch : P.chunks) P.work[0] += ch.nunits;
        P.work[1] = (double)P.ncand;
        return GI_OK;
    }
    TL_BEGIN("compact");
    for (size_t k = 0; k < P.q.size(); k++) {
        const long long b = P.q[k].cand_begin, n = P.q[k].cand_end - b;
        const int nb = compact_blocks(n);
        if (nb > 0)
            launch_compact(S.stream, n, flag + b, counts + cbase[k], offs + obase[k],
                           shift(cand, b), shift(comp, b));
        else
            HIPCHK(hipMemsetAsync(offs + obase[k], 0, 8, S.stream));
        HIPCHK(hipMemcpyAsync(P.counts + k, offs + obase[k] + nb, 8, hipMemcpyDeviceToHost,
                              S.stream));
    }
    TL_END();
    HIPCHK(hipGetLastError());
    for (auto &ch : P.chunks) P.work[0] += ch.nunits;
    P.work[1] = (double)P.ncand;
    return GI_OK;
}

// Host bytes of one emitted line without its amplitudes (ilevelI, ilevelF, nu, S, A, I).
constexpr double kLineBytes = 4 + 4 + 8 * 4;

// Async copy of n compacted lines starting at candidate slot `from` to host
// arrays at row `to`, on stream st.
static int copy_range(gi_ctx *c, hipStream_t st, const LineSoA &comp, long long from, long long n,
                      int *iI, int *iF, double *nu, double *Sv, double *A, double *I,
                      double *amp, long long to)
{
    if (n <= 0) return GI_OK;
    const LineSoA s = shift(comp, from);
    HIPCHK(hipMemcpyAsync(iI + to, s.iI, n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(iF + to, s.iF, n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nu + to, s.nu, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(Sv + to, s.S, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(A + to, s.A, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(I + to, s.I, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(amp + s.amp_w * to, s.amp, n * 8 * s.amp_w, hipMemcpyDeviceToHost,
                          st));
    return GI_OK;
}

// Copy the lines of query k (count already known) to host arrays, async.
static int copy_lines(gi_ctx *c, Slot &S, const Plan &P, size_t k, int *iI, int *iF, double *nu,
                      double *Sv, double *A, double *I, double *amp)
{
    const long long range = P.q[k].cand_end - P.q[k].cand_begin;
    const LineSoA &src = P.counts[k] == range ? P.cand : P.comp;  // all passed: no scatter
    return copy_range(c, S.stream, src, P.q[k].cand_begin, P.counts[k], iI, iF, nu, Sv, A, I,
                      amp, 0);
}

// Pipelined output: as each round's event fires, copy its compacted lines to
// `out` on the copy stream (overlapping the rounds still computing).
static int drain_rounds(gi_ctx *c, Slot &S, const Plan &P, gi_lines *out, long long *nout)
{
    int rc;
    *nout = 0;
    for (size_t r = 0; r < P.rounds.size(); r++) {
        const RoundOut &o = P.rounds[r];
        HIPCHK(hipEventSynchronize(o.ev));
        const long long n = o.nb > 0 ? P.rcounts[r] : 0;
        if (n > 0) {
            HIPCHK(hipStreamWaitEvent(S.cstream, o.ev, 0));
            const LineSoA &src = n == o.ce - o.cb ? P.cand : P.comp;  // all passed: no scatter
            if ((rc = copy_range(c, S.cstream, src, o.cb, n, out->ilevelI, out->ilevelF,
                                 out->nu, out->S, out->A_einst, out->I, out->amp, *nout)) !=
                GI_OK)
                return rc;
        }
        *nout += n;
    }
    HIPCHK(hipStreamSynchronize(S.cstream));
    return GI_OK;
}

// After the plan's work has finished: whether a persistent round gave up a
// dependency wait (its results are then void).
static int persist_check(gi_ctx *c, const Plan &P)
{
    if (!P.perr) return GI_OK;
    int e = 0;
    HIPCHK(hipMemcpy(&e, P.perr, 4, hipMemcpyDeviceToHost));
    if (e) return fail(c, GI_EHIP, "persistent round: a dependency wait gave up");
    return GI_OK;
}

// The deck job's query (temperature, Q, thresholds of gi_set_params).
static Query job_query(const gi_ctx *c)
{
    Query q;
    q.beta = Consts::beta(c->temperature);
    q.q_stat = c->q_stat;
    q.zpe = c->zpe;
    q.thS = c->thresh_S;
    q.thI = c->thresh_I;
    return q;
}

int compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out)
{
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_compute before gi_prepare");
    if (c->server) return fail(c, GI_ESTATE, "gi_compute while the request server runs");
    if (mode != GI_MODE_FP64 && mode != GI_MODE_FP64_BATCHED && mode != GI_MODE_SPLIT)
        return fail(c, GI_EINVAL, "gi_compute: unknown mode");
    int rc;
    if (mode == GI_MODE_SPLIT && (rc = prepare_split(c)) != GI_OK) return rc;
    using clk = std::chrono::steady_clock;
    const auto ms_since = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
    };
    const auto t_start = clk::now();
    Plan P;
    P.mode = mode;
    Query q = job_query(c);
    items_job(c, nlower, ids, 0, P);
    q.cand_begin = 0;
    q.cand_end = P.ncand;
    P.q.push_back(q);
    for (double &w : c->work) w = 0;
    c->kstats.clear();
    c->kstat_names.clear();
    if (P.ncand == 0) {
        out->n = 0;
        return GI_OK;
    }
    const bool arrays = out->ilevelI && out->ilevelF && out->nu && out->S && out->A_einst &&
                        out->I && out->amp;
    // batched mode with room for every candidate: copy each round's lines
    // while the next rounds compute
    P.pipe = mode != GI_MODE_FP64 && out->capacity >= P.ncand && arrays;
    if ((rc = build_chunks(c, P)) != GI_OK) return rc;
    Slot &S = c->main_slot;
    long long nout = 0;
    double t_plan = ms_since(t_start), t_enqueue = 0;
    {
        Timeline tl(S.stream);
        const auto t_launch = clk::now();
        if ((rc = launch_plan(c, S, P, &tl)) != GI_OK) return rc;
        t_enqueue = ms_since(t_launch);
        if (P.pipe && (rc = drain_rounds(c, S, P, out, &nout)) != GI_OK) return rc;
        HIPCHK(hipStreamSynchronize(S.stream));
        tl.collect(c);
    }
    if ((rc = persist_check(c, P)) != GI_OK) return rc;
    for (int i = 0; i < kWorkStats; i++) c->work[i] = P.work[i];
    if (!P.pipe) nout = P.counts[0];
    c->work[2] = (double)nout;
    c->work[9] = t_plan;
    c->work[10] = t_enqueue;
    c->work[12] = (double)nout * (kLineBytes + 8.0 * c->amp_m * c->amp_m);
    out->n = nout;
    if (!P.pipe) {
        if (nout > out->capacity)
            return fail(c, GI_ECAPACITY, "gi_compute: output capacity too small");
        if (nout > 0) {
            if (!arrays) return fail(c, GI_EINVAL, "gi_compute: output arrays missing");
            if ((rc = copy_lines(c, S, P, 0, out->ilevelI, out->ilevelF, out->nu, out->S,
                                 out->A_einst, out->I, out->amp)) != GI_OK)
                return rc;
            HIPCHK(hipStreamSynchronize(S.stream));
        }
    }
    c->work[13] = ms_since(t_start);
    c->work[11] = c->work[13] - t_plan - t_enqueue;
    return GI_OK;
}

// Project autotuner (no reference counterpart): times the launch
// configurations of the batched modes on the caller's workload and keeps the
// fastest. Each setting is scanned in a fixed order from the built-in default;
// a candidate replaces the incumbent only when it is more than 3% faster, so
// timing noise between near-equal settings does not change the pick. Tile
// shapes are judged by their kernel's HIP-event time, the pipeline round count
// by the wall time of the whole call (its effect is the copy overlap). Every
// setting gives bitwise-identical lines (the K order of each product and the
// chunk order do not depend on them).
int autotune(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out, int reps,
             gi_tuning *best, char *table, long long cap, long long *len)
{
    if (mode != GI_MODE_FP64 && mode != GI_MODE_FP64_BATCHED && mode != GI_MODE_SPLIT)
        return fail(c, GI_EINVAL, "gi_autotune: unknown mode");
    if (c->server) return fail(c, GI_ESTATE, "gi_autotune while the request server runs");
    const gi_tuning saved{c->tile_halfls, c->tile_corr, c->pipe_rounds, c->tile_split,
                          c->pipe_share};
    std::string txt;
    char line[256];
    long long ncand = 0;
    int rc = count_job(c, nlower, ids, &ncand);
    if (rc) return rc;
    std::snprintf(line, sizeof line, "# gi_autotune mode=%d lower=%d candidates=%lld reps=%d\n",
                  mode, ids ? nlower : c->nlevels, ncand, reps);
    txt += line;
    // one measurement: min over reps of a kernel's total time in a call, or of the call
    auto measure = [&](const char *kernel, double *ms) -> int {
        *ms = 1e300;
        for (int r = 0; r < reps; r++) {
            const auto t0 = std::chrono::steady_clock::now();
            const int e = compute(c, mode, nlower, ids, out);
            const auto t1 = std::chrono::steady_clock::now();
            if (e) return e;
            double v = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (kernel) {
                auto it = c->kstats.find(kernel);
                v = it == c->kstats.end() ? 0.0 : it->second.ms;
            }
            *ms = std::min(*ms, v);
        }
        return GI_OK;
    };
    auto scan = [&](const char *what, const char *kernel, int *knob, const int *opts, int nopt,
                    const char *const *label) -> int {
        const int start = *knob;
        double tbest = 0;
        int pick = start;
        // the incumbent (built-in default) first, then the others in a fixed order
        std::vector<int> order{start};
        for (int i = 0; i < nopt; i++)
            if (opts[i] != start) order.push_back(opts[i]);
        for (size_t i = 0; i < order.size(); i++) {
            *knob = order[i];
            double ms;
            const int e = measure(kernel, &ms);
            if (e) { *knob = start; return e; }
            int li = 0;
            while (li < nopt && opts[li] != order[i]) li++;
            std::snprintf(line, sizeof line, "%s %d%s%s %.4f ms\n", what, order[i],
                          label ? " " : "", label ? label[li] : "", ms);
            txt += line;
            if (i == 0 || ms < 0.97 * tbest) {
                tbest = ms;
                pick = order[i];
            }
        }
        *knob = pick;
        return GI_OK;
    };
    static const int tiles[6] = {0, 1, 2, 3, kTileAuto, kTile128x128w8};
    static const char *const tname[6] = {"64x64", "64x128", "128x64", "128x128", "auto",
                                         "128x128/8-waves"};
    static const int rounds[6] = {1, 2, 3, 4, 6, 8};
    static const int shares[4] = {50, 60, 70, 80};
    static const int stiles[5] = {0, 1, 2, 3, kSplitTileAuto};
    static const char *const sname[5] = {"64x64", "128x64", "64x64/2-steps",
                                         "128x64/2-steps", "auto"};
    if (mode == GI_MODE_SPLIT &&
        (rc = scan("split_halfls_tile", "half_ls_split", &c->tile_split, stiles, 5, sname))) {
        c->tile_split = saved.split_cfg;
        return rc;
    }
    if (mode == GI_MODE_FP64_BATCHED) {
        if ((rc = scan("halfls_tile", "half_ls_mfma", &c->tile_halfls, tiles, 6, tname)) ||
            (rc = scan("corr_tile", "corr_mfma", &c->tile_corr, tiles, 4, tname))) {
            c->tile_halfls = saved.halfls_cfg; c->tile_corr = saved.corr_cfg;
            return rc;
        }
    }
    if (mode != GI_MODE_FP64 && out->capacity >= ncand) {  // pipelined output only
        // rounds and share together (the best share depends on the round count:
        // at C2 FP64 batched 3 rounds at 70% measured 5.89 ms against 6.20 ms for
        // 4 at 60%, which a scan of one knob after the other did not reach):
        // the incumbent first, then every pair in a fixed order
        std::vector<std::pair<int, int>> order{{c->pipe_rounds, c->pipe_share}};
        for (int r : rounds)
            for (int sh : shares)
                if (!(r == 1 && sh != shares[0]) &&  // one round: the share is unused
                    std::make_pair(r, sh) != order[0])
                    order.push_back({r, sh});
        double tbest = 0;
        std::pair<int, int> pick = order[0];
        for (size_t i = 0; i < order.size(); i++) {
            c->pipe_rounds = order[i].first;
            c->pipe_share = order[i].second;
            double ms;
            if ((rc = measure(nullptr, &ms)) != GI_OK) {
                c->tile_halfls = saved.halfls_cfg; c->tile_corr = saved.corr_cfg;
                c->pipe_rounds = saved.pipe_rounds; c->tile_split = saved.split_cfg;
                c->pipe_share = saved.pipe_share;
                return rc;
            }
            std::snprintf(line, sizeof line, "pipe_rounds %d pipe_share %d %.4f ms\n",
                          order[i].first, order[i].second, ms);
            txt += line;
            if (i == 0 || ms < 0.97 * tbest) {
                tbest = ms;
                pick = order[i];
            }
        }
        c->pipe_rounds = pick.first;
        c->pipe_share = pick.second;
        // the half-ls tile again at the chosen rounds (the best shape depends on the
        // size of the rounds: at C2 one round of 128 x 128 tiles on eight wavefronts
        // measured 2.92 ms against 3.17 ms for 64 x 64, which wins small rounds),
        // from the setting in force at the call again, so that the first scan's
        // pick at other rounds takes no incumbency into this one
        if (mode == GI_MODE_FP64_BATCHED) c->tile_halfls = saved.halfls_cfg;
        if (mode == GI_MODE_FP64_BATCHED &&
            (rc = scan("halfls_tile", "half_ls_mfma", &c->tile_halfls, tiles, 6, tname))) {
            c->tile_halfls = saved.halfls_cfg; c->tile_corr = saved.corr_cfg;
            c->pipe_rounds = saved.pipe_rounds; c->pipe_share = saved.pipe_share;
            return rc;
        }
    }
    std::snprintf(line, sizeof line,
                  "best halfls_tile=%d corr_tile=%d pipe_rounds=%d split_tile=%d pipe_share=%d\n",
                  c->tile_halfls, c->tile_corr, c->pipe_rounds, c->tile_split, c->pipe_share);
    txt += line;
    if (best)
        *best = gi_tuning{c->tile_halfls, c->tile_corr, c->pipe_rounds, c->tile_split,
                          c->pipe_share};
    if (len) *len = (long long)txt.size();
    if (table && cap > 0) {
        const size_t n = std::min<size_t>((size_t)cap - 1, txt.size());
        std::memcpy(table, txt.data(), n);
        table[n] = 0;
    }
    return GI_OK;
}

int count_job(gi_ctx *c, int nlower, const int *ids, long long *count)
{
    Plan P;
    items_job(c, nlower, ids, 0, P);
    *count = P.ncand;
    return GI_OK;
}

// ===========================================================================
// Request server: continuous batching over two execution slots.

struct Request {
    gi_request r;
    long long ticket = 0;
    double est = 0;                         // device bytes while in flight
    std::chrono::steady_clock::time_point t_submit, t_done;
    int status = GI_OK;
    std::string err;
    bool done = false;
    long long n = 0;
    std::vector<int> iI, iF;
    std::vector<double> nu, S, A, I, amp;
};

struct Server {
    gi_ctx *c = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<long long, std::shared_ptr<Request>> all;
    long long next_ticket = 1;
    bool stop = false;
    double budget = 0, in_use = 0;
    Slot slots[2];
    long long batches = 0, served = 0, max_batch = 0;
};

struct Inflight {
    std::unique_ptr<Plan> plan;
    std::vector<std::shared_ptr<Request>> reqs;
    int slot = 0;
    int status = GI_OK;
    std::string err;
};

// Q(T) = sum over resident levels of gns * ndeg * (2J+1) * exp(-(E - ZPE) * beta)
__global__ __launch_bounds__(256) void k_partition(int n, const double *__restrict__ e,
                                                   const double *__restrict__ w, double zpe,
                                                   double beta, double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x)
        s += w[i] * exp(-(e[i] - zpe) * beta);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The host's fixed-order sum of the partial sums of k_partition, on the
// device (the same additions in the same order: the same bits).
__global__ void k_partition_sum(int nb, const double *__restrict__ part, double *__restrict__ q)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += part[i];
    *q = s;
}

constexpr int kPartBlocks = 64;

// The (beta, Q, thresh_S, thresh_I) block of every query of a request batch
// in device memory, read by the intensity steps; Q(T) is reduced on the
// device, on the batch's stream, for the queries without a Q of their own.
static int setup_query_params(gi_ctx *c, Slot &S, Plan &P)
{
    int rc;
    const size_t nq = P.q.size();
    double *qdev = (double *)slot_scratch(c, S, "qdev", 32 * std::max<size_t>(nq, 1), &rc);
    if (rc) return rc;
    size_t ndev = 0;
    for (const Query &q : P.q) ndev += !(q.q_stat > 0.0);
    double *part = nullptr;
    if (ndev) {
        part = (double *)slot_scratch(c, S, "qparts", 8 * kPartBlocks * ndev, &rc);
        if (rc) return rc;
    }
    unsigned char *hb = stage(c, S, nullptr, 32 * std::max<size_t>(nq, 1), &rc);
    if (rc) return rc;
    double *h = (double *)hb;
    for (size_t k = 0; k < nq; k++) {
        h[4 * k] = P.q[k].beta;
        h[4 * k + 1] = P.q[k].q_stat > 0.0 ? P.q[k].q_stat : 0.0;
        h[4 * k + 2] = P.q[k].thS;
        h[4 * k + 3] = P.q[k].thI;
    }
    HIPCHK(hipMemcpyAsync(qdev, hb, 32 * nq, hipMemcpyHostToDevice, S.stream));
    size_t i = 0;
    for (size_t k = 0; k < nq; k++) {
        if (P.q[k].q_stat > 0.0) continue;
        double *pk = part + kPartBlocks * i++;
        hipLaunchKernelGGL(k_partition, dim3(kPartBlocks), dim3(256), 0, S.stream, c->nlevels,
                           c->d_lv_energy, c->d_lv_weight, c->zpe, P.q[k].beta, pk);
        hipLaunchKernelGGL(k_partition_sum, dim3(1), dim3(64), 0, S.stream, kPartBlocks, pk,
                           qdev + 4 * k + 1);
    }
    HIPCHK(hipGetLastError());
    P.qdev = qdev;
    return GI_OK;
}

int partition_function(gi_ctx *c, Slot &S, double T, double *q)
{
    if (!c->prepared) return fail(c, GI_ESTATE, "partition function before gi_prepare");
    if (!(T > 0.0)) return fail(c, GI_EINVAL, "temperature <= 0");
    int rc;
    const int nb = kPartBlocks;
    double *part = (double *)slot_scratch(c, S, "qpart", nb * 8, &rc); if (rc) return rc;
    double *hp = (double *)slot_pinned(c, S, nb * 8, &rc); if (rc) return rc;
    hipLaunchKernelGGL(k_partition, dim3(nb), dim3(256), 0, S.stream, c->nlevels, c->d_lv_energy,
                       c->d_lv_weight, c->zpe, Consts::beta(T), part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hp, part, nb * 8, hipMemcpyDeviceToHost, S.stream));
    HIPCHK(hipStreamSynchronize(S.stream));
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += hp[i];  // fixed order: deterministic
    *q = s;
    return GI_OK;
}

static void finish_batch(Server &sv, Inflight &b)
{
    gi_ctx *c = sv.c;
    Slot &S = sv.slots[b.slot];
    int rc = b.status;
    std::string err = b.err;
    if (rc == GI_OK && hipStreamSynchronize(S.stream) != hipSuccess) {
        rc = GI_EHIP;
        err = "request batch failed on the device";
    }
    if (rc == GI_OK && b.plan->perr) {
        int e = 0;
        if (hipMemcpy(&e, b.plan->perr, 4, hipMemcpyDeviceToHost) != hipSuccess || e) {
            rc = GI_EHIP;
            err = "persistent round: a dependency wait gave up";
        }
    }
    if (rc == GI_OK) {
        for (size_t k = 0; k < b.reqs.size(); k++) {
            Request &R = *b.reqs[k];
            const long long n = b.plan->counts[k];
            R.n = n;
            R.iI.resize(n); R.iF.resize(n); R.nu.resize(n); R.S.resize(n); R.A.resize(n);
            R.I.resize(n); R.amp.resize((size_t)c->amp_m * c->amp_m * n);
            if (n > 0 && copy_lines(c, S, *b.plan, k, R.iI.data(), R.iF.data(), R.nu.data(),
                                    R.S.data(), R.A.data(), R.I.data(), R.amp.data()) != GI_OK)
                rc = GI_EHIP;
        }
        if (rc == GI_OK && hipStreamSynchronize(S.stream) != hipSuccess) rc = GI_EHIP;
        if (rc != GI_OK) err = "copy of request lines failed";
    }
    const auto now = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(sv.mu);
    for (auto &r : b.reqs) {
        r->status = rc;
        if (rc != GI_OK) r->err = err;
        r->t_done = now;
        r->done = true;
        sv.in_use -= r->est;
    }
    sv.served += (long long)b.reqs.size();
    sv.cv_done.notify_all();
}

static void serve_loop(Server &sv)
{
    gi_ctx *c = sv.c;
    std::string werr;
    err_sink = &werr;  // this thread's errors stay with its batches
    (void)hipSetDevice(c->device);
    std::deque<Inflight> inflight;
    int next_slot = 0;
    for (;;) {
        std::vector<std::shared_ptr<Request>> adm;
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            if (inflight.empty())
                sv.cv_work.wait(lk, [&] { return sv.stop || !sv.pending.empty(); });
            if (sv.stop && sv.pending.empty() && inflight.empty()) break;
            if (inflight.size() < 2) {
                int mode = -1;
                while (!sv.pending.empty()) {
                    auto &r = sv.pending.front();
                    if (mode >= 0 && r->r.mode != mode) break;  // one mode per batch
                    if (sv.in_use + r->est > sv.budget) break;  // admission by memory
                    mode = r->r.mode;
                    sv.in_use += r->est;
                    adm.push_back(r);
                    sv.pending.pop_front();
                }
            }
        }
        if (!adm.empty()) {
            Inflight b;
            b.plan.reset(new Plan());
            b.reqs = adm;
            b.slot = next_slot;
            next_slot ^= 1;
            Plan &P = *b.plan;
            P.mode = adm[0]->r.mode;
            P.device_q = true;  // Q(T) of requests without a Q is reduced on the device
            int rc = GI_OK;
            for (size_t k = 0; k < adm.size() && rc == GI_OK; k++) {
                const gi_request &r = adm[k]->r;
                Query q;
                q.beta = Consts::beta(r.temperature);
                q.q_stat = r.q_stat;
                q.T = r.temperature;
                q.zpe = c->zpe;
                q.thS = r.thresh_linestrength;
                q.thI = r.thresh_intensity;
                q.cand_begin = P.ncand;
                items_request(c, r, (int)k, P);
                q.cand_end = P.ncand;
                P.q.push_back(q);
            }
            if (rc == GI_OK) rc = build_chunks(c, P);
            if (rc == GI_OK) rc = launch_plan(c, sv.slots[b.slot], P, nullptr);
            b.status = rc;
            if (rc != GI_OK) b.err = werr;
            sv.batches++;
            sv.max_batch = std::max<long long>(sv.max_batch, (long long)adm.size());
            inflight.push_back(std::move(b));
        }
        // drain the oldest batch when both slots are busy or nothing new came in
        if (!inflight.empty() && (adm.empty() || inflight.size() == 2)) {
            finish_batch(sv, inflight.front());
            inflight.pop_front();
        }
    }
}

void server_shutdown(gi_ctx *c)
{
    Server *sv = c->server;
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    for (auto &s : sv->slots) {
        free_slot(s);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    delete sv;
    c->server = nullptr;
}

// Device bytes a request holds while in flight: candidate and compacted
// records, flags, and the G, H and folded vectors of its units.
static double request_bytes(gi_ctx *c, const gi_request &r, long long *ncand)
{
    Plan P;
    items_request(c, r, 0, P);
    *ncand = P.ncand;
    const int gF = c->igamma_pair[r.igammaI];
    const GroupPlan &GP = group_of(c, r.jindI, r.igammaI, r.jindF);
    const double units = (double)P.items.size() * c->degen[r.igammaI];
    const double ncb = (double)c->degen[r.igammaI] * c->degen[gF];
    const double nsizeF = c->blocks[r.jindF * c->nrep + gF].nsize;
    return (double)P.ncand * (2 * 80.0 + 4.0) + units * 8.0 * (GP.Gtot + GP.Htot) +
           (double)P.items.size() * ncb * nsizeF * 8.0 + 4096.0;
}

}  // namespace gi

// A captured sweep (gi_sweep_create): the plan of one gi_compute with its
// launch sequence recorded as one HIP graph per round, replayed by
// gi_sweep_run with new intensity parameters.
struct gi_sweep {
    gi_ctx *c = nullptr;
    int mode = 0;
    std::unique_ptr<gi::Plan> P;
    gi::Slot slot;
    double *d_qpar = nullptr;        // beta, Q, thresh_S, thresh_I (device, read by k_intensity)
    double *h_qpar = nullptr;        // pinned source of d_qpar
};

namespace gi {

static void sweep_free(gi_sweep *w)
{
    (void)hipSetDevice(w->c->device);
    w->P.reset();  // graphs and events
    free_slot(w->slot);
    if (w->slot.cstream) (void)hipStreamDestroy(w->slot.cstream);
    if (w->slot.stream) (void)hipStreamDestroy(w->slot.stream);
    if (w->h_qpar) (void)hipHostFree(w->h_qpar);
    delete w;
}

void sweeps_shutdown(gi_ctx *c)
{
    for (gi_sweep *w : c->sweeps) sweep_free(w);
    c->sweeps.clear();
}

static int sweep_create(gi_ctx *c, int mode, int nlower, const int *ids, gi_sweep **out)
{
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_sweep_create before gi_prepare");
    if (c->server) return fail(c, GI_ESTATE, "gi_sweep_create while the request server runs");
    if (mode != GI_MODE_FP64_BATCHED && mode != GI_MODE_SPLIT)
        return fail(c, GI_EINVAL, "gi_sweep_create: sweeps run the matrix-core modes");
    if (c->mu_streamed)
        return fail(c, GI_ESTATE, "gi_sweep_create: sweeps need mu resident (a captured "
                                  "sweep does not stream mu parts)");
    if (host_blocks(c))
        return fail(c, GI_ESTATE, "gi_sweep_create: sweeps need every eigenvector block resident");
    int rc;
    if (mode == GI_MODE_SPLIT && (rc = prepare_split(c)) != GI_OK) return rc;
    HIPCHK(hipSetDevice(c->device));
    auto w = std::make_unique<gi_sweep>();
    w->c = c;
    w->mode = mode;
    w->P = std::make_unique<Plan>();
    Plan &P = *w->P;
    P.mode = mode;
    Query q = job_query(c);
    items_job(c, nlower, ids, 0, P);
    q.cand_begin = 0;
    q.cand_end = P.ncand;
    P.q.push_back(q);
    P.pipe = true;
    P.capture = true;
    auto cleanup = [&](int code) {
        gi_sweep *raw = w.release();
        sweep_free(raw);
        return code;
    };
    if ((rc = build_chunks(c, P)) != GI_OK) return cleanup(rc);
    if (hipStreamCreate(&w->slot.stream) != hipSuccess ||
        hipStreamCreate(&w->slot.cstream) != hipSuccess)
        return cleanup(fail(c, GI_EHIP, "hipStreamCreate"));
    if (hipHostMalloc((void **)&w->h_qpar, 64, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(c, GI_ENOMEM, "pinned host allocation failed"));
    w->d_qpar = (double *)slot_scratch(c, w->slot, "qpar", 64, &rc);
    if (rc) return cleanup(rc);
    P.qpar = w->d_qpar;
    if (P.ncand > 0 && (rc = launch_plan(c, w->slot, P, nullptr)) != GI_OK) return cleanup(rc);
    if (hipStreamSynchronize(w->slot.stream) != hipSuccess)
        return cleanup(fail(c, GI_EHIP, "gi_sweep_create: uploads failed"));
    c->sweeps.push_back(w.get());
    *out = w.release();
    return GI_OK;
}

static int sweep_run(gi_sweep *w, double T, double Q, double thS, double thI, gi_lines *out)
{
    gi_ctx *c = w->c;
    Plan &P = *w->P;
    if (!(T > 0.0)) return fail(c, GI_EINVAL, "gi_sweep_run: temperature <= 0");
    if (out->capacity < P.ncand)
        return fail(c, GI_ECAPACITY, "gi_sweep_run: out needs room for every candidate");
    if (!(out->ilevelI && out->ilevelF && out->nu && out->S && out->A_einst && out->I && out->amp))
        return fail(c, GI_EINVAL, "gi_sweep_run: output arrays missing");
    int rc;
    HIPCHK(hipSetDevice(c->device));
    if (!(Q > 0.0) && (rc = partition_function(c, c->main_slot, T, &Q)) != GI_OK) return rc;
    out->n = 0;
    if (P.ncand == 0) return GI_OK;
    Slot &S = w->slot;
    w->h_qpar[0] = Consts::beta(T);
    w->h_qpar[1] = Q;
    w->h_qpar[2] = thS;
    w->h_qpar[3] = thI;
    HIPCHK(hipMemcpyAsync(w->d_qpar, w->h_qpar, 32, hipMemcpyHostToDevice, S.stream));
    for (size_t r = 0; r < P.graphs.size(); r++) {
        HIPCHK(hipGraphLaunch(P.graphs[r], S.stream));
        HIPCHK(hipEventRecord(P.rounds[r].ev, S.stream));
    }
    long long nout = 0;
    if ((rc = drain_rounds(c, S, P, out, &nout)) != GI_OK) return rc;
    HIPCHK(hipStreamSynchronize(S.stream));
    if ((rc = persist_check(c, P)) != GI_OK) return rc;
    out->n = nout;
    return GI_OK;
}

}  // namespace gi

using namespace gi;

extern "C" {

int gi_count_candidates(gi_ctx *c, int nlower, const int *ids, long long *count)
{
    if (!c || !count) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_count_candidates before gi_prepare");
    return count_job(c, nlower, ids, count);
}

int gi_amp_width(const gi_ctx *c) { return c ? c->amp_m * c->amp_m : 0; }

int gi_compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out)
{
    if (!c || !out) return GI_EINVAL;
    return compute(c, mode, nlower, ids, out);
}

int gi_set_tuning(gi_ctx *c, const gi_tuning *t)
{
    if (!c || !t) return GI_EINVAL;
    if (c->server) return fail(c, GI_ESTATE, "gi_set_tuning while the request server runs");
    if (t->halfls_cfg < 0 || t->halfls_cfg > kTileHalflsMax || t->corr_cfg < 0 || t->corr_cfg > 3 ||
        t->pipe_rounds < 1 || t->pipe_rounds > kMaxPipeRounds || t->split_cfg < 0 ||
        t->split_cfg > kSplitTileAuto || t->pipe_share < 10 || t->pipe_share > 90)
        return fail(c, GI_EINVAL, "gi_set_tuning: value out of range");
    c->tile_halfls = t->halfls_cfg;
    c->tile_corr = t->corr_cfg;
    c->pipe_rounds = t->pipe_rounds;
    c->tile_split = t->split_cfg;
    c->pipe_share = t->pipe_share;
    return GI_OK;
}

int gi_set_path_limits(gi_ctx *c, int kpart_len, int rec_lds, int split2_steps,
                       int split2_segs, int fold_split_rows)
{
    if (!c) return GI_EINVAL;
    if (c->server || !c->sweeps.empty())
        return fail(c, GI_ESTATE, "gi_set_path_limits while the server runs or sweeps exist");
    const auto pick = [](int v, int def, int lo, int *out) {
        if (v == 0) v = def;
        if (v < lo || v > def) return false;
        *out = v;
        return true;
    };
    int a, b = -1, d, e, f;
    // rec_lds -1: no record staged in LDS (the global-gather G builds)
    if (!pick(kpart_len, 2048, 64, &a) || (rec_lds != -1 && !pick(rec_lds, kRecLds, 1, &b)) ||
        !pick(split2_steps, kSplitMaxSteps, 1, &d) || !pick(split2_segs, kSplitMaxSegs, 1, &e) ||
        !pick(fold_split_rows, kFoldSplitMax, 1, &f))
        return fail(c, GI_EINVAL, "gi_set_path_limits: a limit outside [minimum, default]");
    // kpart_len 0: the defaults of each mode; else both modes use it
    c->lim_kpart = kpart_len ? a : 1024;
    c->lim_kpart_split = kpart_len ? a : 2048;
    c->lim_rec_lds = b;
    c->lim_split2_steps = d;
    c->lim_split2_segs = e;
    c->lim_fold_split = f;
    return GI_OK;
}

int gi_set_persist(gi_ctx *c, int on)
{
    if (!c) return GI_EINVAL;
    if (c->server || !c->sweeps.empty())
        return fail(c, GI_ESTATE, "gi_set_persist while the server runs or sweeps exist");
    if (on != 0 && on != 1) return fail(c, GI_EINVAL, "gi_set_persist: 0 or 1");
    c->persist = on;
    return GI_OK;
}

int gi_get_persist(gi_ctx *c, int *on)
{
    if (!c || !on) return GI_EINVAL;
    *on = c->persist;
    return GI_OK;
}

int gi_get_tuning(gi_ctx *c, gi_tuning *t)
{
    if (!c || !t) return GI_EINVAL;
    t->halfls_cfg = c->tile_halfls;
    t->corr_cfg = c->tile_corr;
    t->pipe_rounds = c->pipe_rounds;
    t->split_cfg = c->tile_split;
    t->pipe_share = c->pipe_share;
    return GI_OK;
}

int gi_autotune(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out, int reps,
                gi_tuning *best, char *table, long long cap, long long *len)
{
    if (!c || !out || reps < 1 || reps > 100) return GI_EINVAL;
    return autotune(c, mode, nlower, ids, out, reps, best, table, cap, len);
}

int gi_sweep_create(gi_ctx *c, int mode, int nlower, const int *ids, gi_sweep **out)
{
    if (!c || !out) return GI_EINVAL;
    *out = nullptr;
    return sweep_create(c, mode, nlower, ids, out);
}

int gi_sweep_run(gi_sweep *w, double temperature, double q_stat, double thresh_linestrength,
                 double thresh_intensity, gi_lines *out)
{
    if (!w || !out) return GI_EINVAL;
    return sweep_run(w, temperature, q_stat, thresh_linestrength, thresh_intensity, out);
}

long long gi_sweep_candidates(const gi_sweep *w) { return w ? w->P->ncand : -1; }

int gi_sweep_destroy(gi_sweep *w)
{
    if (!w) return GI_EINVAL;
    auto &v = w->c->sweeps;
    v.erase(std::remove(v.begin(), v.end(), w), v.end());
    sweep_free(w);
    return GI_OK;
}

int gi_partition_function(gi_ctx *c, double temperature, double *q)
{
    if (!c || !q) return GI_EINVAL;
    return partition_function(c, c->main_slot, temperature, q);
}

int gi_serve_start(gi_ctx *c, long long budget_bytes)
{
    if (!c || budget_bytes <= 0) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_serve_start before gi_prepare");
    if (c->server) return fail(c, GI_ESTATE, "request server already running");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, GI_EHIP, "hipSetDevice");
    Server *sv = new Server();
    sv->c = c;
    sv->budget = (double)budget_bytes;
    for (auto &s : sv->slots)
        if (hipStreamCreate(&s.stream) != hipSuccess) {
            for (auto &t : sv->slots)
                if (t.stream) (void)hipStreamDestroy(t.stream);
            delete sv;
            return fail(c, GI_EHIP, "hipStreamCreate");
        }
    c->server = sv;
    sv->th = std::thread(serve_loop, std::ref(*sv));
    return GI_OK;
}

int gi_serve_submit(gi_ctx *c, const gi_request *r, long long *ticket)
{
    if (!c || !r || !ticket) return GI_EINVAL;
    Server *sv = c->server;
    if (!sv) return fail(c, GI_ESTATE, "gi_serve_submit without gi_serve_start");
    if (r->jindI < 0 || r->jindI >= c->nJ || r->jindF < 0 || r->jindF >= c->nJ || r->igammaI < 0 ||
        r->igammaI >= c->nrep)
        return GI_EINVAL;
    if (!group_of(c, r->jindI, r->igammaI, r->jindF).valid)
        return GI_EINVAL;  // |J' - J| > 1
    if (r->mode != GI_MODE_FP64 && r->mode != GI_MODE_FP64_BATCHED && r->mode != GI_MODE_SPLIT)
        return GI_EINVAL;
    if (r->mode == GI_MODE_SPLIT && c->split_ready != c->split_key())
        return fail(c, GI_ESTATE, "split-mode request before gi_prepare_split");
    if (!(r->temperature > 0.0)) return GI_EINVAL;
    auto R = std::make_shared<Request>();
    R->r = *r;
    long long nc = 0;
    R->est = request_bytes(c, *r, &nc);
    if (R->est > sv->budget) return GI_EBUDGET;
    R->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        R->ticket = sv->next_ticket++;
        sv->all[R->ticket] = R;
        sv->pending.push_back(R);
    }
    sv->cv_work.notify_one();
    *ticket = R->ticket;
    return GI_OK;
}

int gi_serve_wait(gi_ctx *c, long long ticket, double timeout_s, long long *nlines,
                  double *latency_ms)
{
    if (!c) return GI_EINVAL;
    Server *sv = c->server;
    if (!sv) return GI_ESTATE;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return GI_EINVAL;
    std::shared_ptr<Request> R = it->second;
    const auto until = std::chrono::steady_clock::now() +
                       std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                           std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9));
    if (!sv->cv_done.wait_until(lk, until, [&] { return R->done; })) return GI_ETIMEOUT;
    if (R->status != GI_OK) {
        c->err = R->err;
        return R->status;
    }
    if (nlines) *nlines = R->n;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(R->t_done - R->t_submit).count();
    return GI_OK;
}

int gi_serve_fetch(gi_ctx *c, long long ticket, gi_lines *out)
{
    if (!c || !out) return GI_EINVAL;
    Server *sv = c->server;
    if (!sv) return GI_ESTATE;
    std::shared_ptr<Request> R;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->all.find(ticket);
        if (it == sv->all.end()) return GI_EINVAL;
        R = it->second;
        if (!R->done) return fail(c, GI_ESTATE, "gi_serve_fetch before the request finished");
        if (R->status != GI_OK) return fail(c, R->status, R->err);
    }
    out->n = R->n;
    if (R->n > out->capacity) return fail(c, GI_ECAPACITY, "gi_serve_fetch: capacity too small");
    if (R->n > 0) {
        std::memcpy(out->ilevelI, R->iI.data(), R->n * 4);
        std::memcpy(out->ilevelF, R->iF.data(), R->n * 4);
        std::memcpy(out->nu, R->nu.data(), R->n * 8);
        std::memcpy(out->S, R->S.data(), R->n * 8);
        std::memcpy(out->A_einst, R->A.data(), R->n * 8);
        std::memcpy(out->I, R->I.data(), R->n * 8);
        std::memcpy(out->amp, R->amp.data(), R->amp.size() * 8);
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->all.erase(ticket);
    return GI_OK;
}

int gi_serve_stats(gi_ctx *c, long long *batches, long long *served, long long *max_batch)
{
    if (!c) return GI_EINVAL;
    Server *sv = c->server;
    if (!sv) return GI_ESTATE;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (batches) *batches = sv->batches;
    if (served) *served = sv->served;
    if (max_batch) *max_batch = sv->max_batch;
    return GI_OK;
}

int gi_serve_stop(gi_ctx *c)
{
    if (!c) return GI_EINVAL;
    if (!c->server) return GI_OK;
    server_shutdown(c);
    return GI_OK;
}

}  // extern "C"
