// C ABI of the gfx950 line-intensity engine: context, uploads, plan and the
// hot-path driver (the lower-level loop of cuda_host.cu:1174-1488, batched).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gi_internal.h"
#include "gi_kernels.h"

namespace gi {

thread_local std::string *err_sink = nullptr;

int fail(gi_ctx *c, int code, const std::string &msg)
{
    if (err_sink)
        *err_sink = msg;
    else if (c)
        c->err = msg;
    return code;
}

// Calls that change or read the plan are refused while the request server's
// worker uses it or sweeps hold pointers into it.
static int serving(gi_ctx *c)
{
    if (c->server) return fail(c, GI_ESTATE, "call not allowed while the request server runs");
    if (!c->sweeps.empty())
        return fail(c, GI_ESTATE, "call not allowed while sweeps exist (gi_sweep_destroy first)");
    return 0;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(c, e_ == hipErrorOutOfMemory ? GI_ENOMEM : GI_EHIP,             \
                        std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

void *slot_scratch(gi_ctx *c, Slot &s, const char *name, size_t bytes, int *rc)
{
    *rc = GI_OK;
    auto &e = s.buf[name];
    if (e.second >= bytes && e.first) return e.first;
    if (e.first) {
        // the slot's stream may still read the old buffer
        (void)hipStreamSynchronize(s.stream);
        (void)hipFree(e.first);
    }
    e.first = nullptr;
    e.second = 0;
    size_t want = std::max<size_t>(bytes, 256);
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        *rc = fail(c, GI_ENOMEM, std::string("device allocation of ") + name + " (" +
                                     std::to_string(want) + " bytes) failed");
        return nullptr;
    }
    e.first = p;
    e.second = want;
    return p;
}

void *scratch(gi_ctx *c, const char *name, size_t bytes, int *rc)
{
    return slot_scratch(c, c->main_slot, name, bytes, rc);
}

void *slot_pinned(gi_ctx *c, Slot &s, size_t bytes, int *rc)
{
    *rc = GI_OK;
    if (s.pinned_bytes >= bytes && s.pinned) return s.pinned;
    if (s.pinned) {
        (void)hipStreamSynchronize(s.stream);
        (void)hipHostFree(s.pinned);
    }
    s.pinned = nullptr;
    s.pinned_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    if (hipHostMalloc(&s.pinned, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        s.pinned = nullptr;
        *rc = fail(c, GI_ENOMEM, "pinned host allocation failed");
        return nullptr;
    }
    s.pinned_bytes = want;
    return s.pinned;
}

void free_slot(Slot &s)
{
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.cstream) (void)hipStreamSynchronize(s.cstream);
    for (auto &e : s.mu_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (auto &e : s.vec_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (auto &e : s.buf) (void)hipFree(e.second.first);
    s.buf.clear();
    if (s.pinned) (void)hipHostFree(s.pinned);
    s.pinned = nullptr;
    s.pinned_bytes = 0;
    s.arena.release();
}

unsigned char *HostArena::alloc(size_t n)
{
    n = (n + 255) & ~size_t(255);
    while (cur < blocks.size()) {
        if (off + n <= blocks[cur].second) {
            unsigned char *p = blocks[cur].first + off;
            off += n;
            return p;
        }
        cur++;
        off = 0;
    }
    const size_t last = blocks.empty() ? 0 : blocks.back().second;
    const size_t want = std::max({n, (size_t)4 << 20, 2 * last});
    void *p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    blocks.push_back({(unsigned char *)p, want});
    cur = blocks.size() - 1;
    off = n;
    return (unsigned char *)p;
}

void HostArena::release()
{
    for (auto &b : blocks) (void)hipHostFree(b.first);
    blocks.clear();
    reset();
}

template <typename T>
static int upload(gi_ctx *c, T **dst, const std::vector<T> &src)
{
    if (*dst) {
        (void)hipFree(*dst);
        *dst = nullptr;
    }
    size_t n = std::max<size_t>(src.size(), 1);
    HIPCHK(hipMalloc((void **)dst, n * sizeof(T)));
    if (!src.empty())
        HIPCHK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                              c->stream));
    return GI_OK;
}

static void free_sym(SymMap &m)
{
    (void)hipFree(m.d_rptr); (void)hipFree(m.d_rcol); (void)hipFree(m.d_rval);
    (void)hipFree(m.d_tptr); (void)hipFree(m.d_tcol); (void)hipFree(m.d_tval);
    m.d_rptr = m.d_rcol = m.d_tptr = m.d_tcol = nullptr;
    m.d_rval = m.d_tval = nullptr;
}

void free_split(gi_ctx *c)
{
    (void)hipFree(c->d_mud);
    (void)hipFree(c->d_mue);
    if (c->h_mud) (void)hipHostFree(c->h_mud);
    c->d_mud = nullptr;
    c->h_mud = nullptr;
    c->d_mue = nullptr;
    for (auto &b : c->blocks) {
        (void)hipFree(b.d_ud); (void)hipFree(b.d_ue); (void)hipFree(b.d_umax);
        if (b.h_ud) (void)hipHostFree(b.h_ud);
        if (b.h_ue) (void)hipHostFree(b.h_ue);
        b.d_ud = nullptr;
        b.d_ue = nullptr;
        b.d_umax = nullptr;
        b.h_ud = nullptr;
        b.h_ue = nullptr;
        b.Kp = 0;
        b.uplane = 0;
    }
    c->split_ready = 0;
}

// Digit planes of a streamed mu, built part by part on the device from the
// pinned FP64 parts (the same rows and exponents as from resident mu: each
// vF row's digits depend on that row only) and kept in pinned host memory;
// the split mode's half-ls copies them back in parts.
static int split_mu_streamed(gi_ctx *c, const int *d_vinv)
{
    const int N = c->nvib, np = c->mu_np, SH = c->split_Sh;
    const size_t kp = (size_t)c->kpad_all;
    HIPCHK(hipHostMalloc((void **)&c->h_mud, (size_t)SH * c->muplane, hipHostMallocDefault));
    double *dpart = nullptr;
    int8_t *gpart = nullptr;
    hipError_t e = hipMalloc((void **)&dpart, (size_t)3 * N * np * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&gpart, (size_t)SH * 3 * np * kp);
    for (int v0 = 0; e == hipSuccess && v0 < N; v0 += np) {
        const int n = std::min(np, N - v0);
        // rows (a, vI) of the part's n vF columns, then their digit rows (q, a, vF)
        e = hipMemcpy2DAsync(dpart, (size_t)n * 8, c->h_mu_perm + v0, (size_t)N * 8,
                             (size_t)n * 8, (size_t)3 * N, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) break;
        launch_split_mu(c->stream, SH, c->split_radix, N, n, n, c->kpad_all, dpart, d_vinv, gpart,
                        3LL * n * kp, c->d_mue + v0);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(c->h_mud + v0 * kp, (size_t)N * kp, gpart, (size_t)n * kp,
                                 (size_t)n * kp, (size_t)3 * SH, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // the part buffers are reused
    }
    (void)hipFree(dpart);
    (void)hipFree(gpart);
    if (e != hipSuccess)
        return fail(c, GI_EHIP, std::string("streamed mu digits: ") + hipGetErrorString(e));
    return GI_OK;
}

// Digit planes of mu and of every resident eigenvector block for the split
// mode (built once per digit count; freed with the plan).
// Digit planes of a block kept in host memory (gi_set_vector_budget): its
// records go to the device in batches, are cut into digits there (each
// record's digits and exponent depend on that record only, so they are the
// bytes a resident block would hold) and come back to pinned host memory;
// the exponents and max |u| per record stay on the device as well.
static int split_host_block(gi_ctx *c, Block &b, int SD)
{
    const size_t batch_bytes = (size_t)256 << 20;
    const int rb = (int)std::max<size_t>(1, std::min<size_t>(b.nrec,
                                                             batch_bytes / (8 * (size_t)b.nsize)));
    HIPCHK(hipHostMalloc((void **)&b.h_ud, (size_t)SD * b.uplane, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&b.h_ue, sizeof(int) * b.nrec, hipHostMallocDefault));
    double *dx = nullptr;
    int8_t *dd = nullptr;
    if (hipMalloc((void **)&dx, (size_t)rb * b.nsize * 8) != hipSuccess ||
        hipMalloc((void **)&dd, (size_t)SD * rb * b.Kp) != hipSuccess) {
        (void)hipFree(dx);
        return fail(c, GI_ENOMEM, "split mode: staging of a host block's digits");
    }
    int rc = GI_OK;
    for (int r0 = 0; r0 < b.nrec && rc == GI_OK; r0 += rb) {
        const int n = std::min(rb, b.nrec - r0);
        const long long plane = (long long)n * b.Kp;
        if (hipMemcpyAsync(dx, b.h_vec + (size_t)r0 * b.nsize, (size_t)n * b.nsize * 8,
                           hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            rc = fail(c, GI_EHIP, "split mode: copy of a host block");
            break;
        }
        launch_split_rows(c->stream, SD, c->split_radix, dx, b.nsize, n, b.nsize, b.Kp, dd, plane,
                          b.d_ue + r0, b.d_umax + r0);
        for (int q = 0; q < SD && rc == GI_OK; q++)
            if (hipMemcpyAsync(b.h_ud + q * b.uplane + (size_t)r0 * b.Kp, dd + q * plane,
                               (size_t)plane, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                rc = fail(c, GI_EHIP, "split mode: copy of a host block's digits");
        // the buffers are reused by the next batch
        if (rc == GI_OK && hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, GI_EHIP, "split mode: digits of a host block");
    }
    if (rc == GI_OK &&
        (hipMemcpy(b.h_ue, b.d_ue, sizeof(int) * b.nrec, hipMemcpyDeviceToHost) != hipSuccess))
        rc = fail(c, GI_EHIP, "split mode: exponents of a host block");
    (void)hipFree(dx);
    (void)hipFree(dd);
    return rc;
}

int prepare_split(gi_ctx *c)
{
    if (!c->prepared) return fail(c, GI_ESTATE, "split mode before gi_prepare");
    if (c->split_ready == c->split_key()) return GI_OK;
    HIPCHK(hipSetDevice(c->device));
    free_split(c);
    // The format: an int32 sum of digit products is exact while
    // S * (R/2)^2 * K < 2^31 for its length K: a class of the half linestrength
    // (the sum of its K segments), a split-K part of the correlation. Radix 254
    // (8-bit digits, 5 by default) when both products allow it, else 128
    // (7-bit digits, 6 by default), unless the caller asked for one.
    long long Kh = 0, Kc = 0;
    for (const GroupPlan &G : c->groups)
        for (const OutClass &oc : G.oc) Kh = std::max<long long>(Kh, oc.K);
    for (const Block &b : c->blocks)
        if (b.nrec > 0 && b.nsize > 0) {
            const int Kp = (b.nsize + 63) / 64 * 64;
            const int parts = corr_parts(c, Kp, true);
            Kc = std::max<long long>(Kc, (long long)(Kp + 64 * parts - 1) / (64 * parts) * 64);
        }
    const auto fits = [](double R, int S, long long K) {
        const long long dmax = (long long)(R / 2);
        return (long long)S * dmax * dmax * K < (1LL << 31);
    };
    bool chosen = false;
    for (const double R : {254.0, 128.0}) {
        if (c->split_req_radix && c->split_req_radix != (int)R) continue;
        const int def = R > 200 ? 5 : 6;
        const int sh = c->split_req_Sh ? c->split_req_Sh : def;
        const int sd = c->split_req_S ? c->split_req_S : def;
        if (fits(R, sh, Kh) && fits(R, sd, Kc)) {
            c->split_radix = R;
            c->split_Sh = sh;
            c->split_S = sd;
            chosen = true;
            break;
        }
    }
    if (!chosen)
        return fail(c, GI_EINVAL, "split mode: products too long for exact int32 digit sums");
    const int SD = c->split_S, SH = c->split_Sh, N = c->nvib;
    int rc;
    std::vector<int> vinv(c->kpad_all, -1);
    for (size_t k = 0; k < c->vcls_off.size(); k++)
        for (int i = 0; i < c->vcls_n[k]; i++) vinv[c->vpad_off[k] + i] = c->vcls_off[k] + i;
    int *d_vinv = nullptr;
    if ((rc = upload(c, &d_vinv, vinv)) != GI_OK) return rc;
    c->muplane = 3LL * N * c->kpad_all;
    HIPCHK(hipMalloc((void **)&c->d_mue, sizeof(int) * N));
    if (!c->mu_streamed) {
        HIPCHK(hipMalloc((void **)&c->d_mud, (size_t)SH * c->muplane));
        launch_split_mu(c->stream, SH, c->split_radix, N, N, N, c->kpad_all, c->d_mu_perm, d_vinv,
                        c->d_mud, c->muplane, c->d_mue);
        HIPCHK(hipGetLastError());
    } else if ((rc = split_mu_streamed(c, d_vinv)) != GI_OK) {
        (void)hipFree(d_vinv);
        return rc;
    }
    for (auto &b : c->blocks) {
        if (b.nrec <= 0 || b.nsize <= 0 || !(b.d_vec || b.h_vec)) continue;
        b.Kp = (b.nsize + 63) / 64 * 64;
        b.uplane = (long long)b.nrec * b.Kp;
        HIPCHK(hipMalloc((void **)&b.d_ue, sizeof(int) * b.nrec));
        HIPCHK(hipMalloc((void **)&b.d_umax, sizeof(double) * b.nrec));
        if (b.d_vec) {
            HIPCHK(hipMalloc((void **)&b.d_ud, (size_t)SD * b.uplane));
            launch_split_rows(c->stream, SD, c->split_radix, b.d_vec, b.nsize, b.nrec, b.nsize,
                              b.Kp, b.d_ud, b.uplane, b.d_ue, b.d_umax);
            HIPCHK(hipGetLastError());
        } else if ((rc = split_host_block(c, b, SD)) != GI_OK) {
            (void)hipFree(d_vinv);
            return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(d_vinv);
    c->split_ready = c->split_key();
    return GI_OK;
}

static void free_plan(gi_ctx *c)
{
    free_split(c);
    for (auto &p : c->jp) {
        for (auto &m : p.sym) free_sym(m);
        (void)hipFree(p.d_r_grid);
    }
    c->jp.clear();
    for (auto &cf : c->coef) {
        (void)hipFree(cf.d_ptr); (void)hipFree(cf.d_sI); (void)hipFree(cf.d_val);
    }
    c->coef.clear();
    for (auto &b : c->blocks) {
        (void)hipFree(b.d_lev); (void)hipFree(b.d_col); (void)hipFree(b.d_en);
        (void)hipFree(b.d_upok);
        b.d_lev = b.d_col = nullptr;
        b.d_en = nullptr;
        b.d_upok = nullptr;
    }
    for (auto &G : c->groups) {
        for (auto &oc : G.oc) {
            (void)hipFree(oc.d_ints);
            (void)hipFree(oc.d_rcoef);
            (void)hipFree(oc.d_segpad);
        }
        for (int d = 0; d < GI_MAXDEG; d++) {
            (void)hipFree(G.d_tcol[d]);
            (void)hipFree(G.d_tfc[d]);
            (void)hipFree(G.d_tfw[d]);
        }
        (void)hipFree(G.d_cd);
        for (int d = 0; d < GI_MAXDEG; d++) {
            (void)hipFree(G.d_gidx[d]);
            (void)hipFree(G.d_gwt[d]);
            (void)hipFree(G.d_gidxB[d]);
            (void)hipFree(G.d_gwtB[d]);
            (void)hipFree(G.d_gbnd[d]);
            (void)hipFree(G.d_cidx[d]);
            (void)hipFree(G.d_ccode[d]);
            (void)hipFree(G.d_cwt[d]);
        }
    }
    c->groups.clear();
    (void)hipFree(c->d_mu_perm);
    c->d_mu_perm = nullptr;
    (void)hipHostFree(c->h_mu_perm);
    c->h_mu_perm = nullptr;
    (void)hipFree(c->d_iota);
    c->d_iota = nullptr;
    c->iota_n = 0;
    (void)hipFree(c->d_lv_energy);
    (void)hipFree(c->d_lv_weight);
    c->d_lv_energy = c->d_lv_weight = nullptr;
    c->prepared = false;
}

}  // namespace gi

using namespace gi;

extern "C" {

int gi_create(int device, gi_ctx **out)
{
    if (!out) return GI_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GI_EHIP;
    if (device < 0 || device >= ndev) return GI_EINVAL;
    gi_ctx *c = new gi_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return GI_EHIP;
    }
    c->main_slot.stream = c->stream;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        c->ncu = ncu;
    noalloc_init(device);  // once per device; a failure is reported by gi_half_ls_noalloc
    if (hipStreamCreate(&c->main_slot.cstream) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return GI_EHIP;
    }
    *out = c;
    return GI_OK;
}

void gi_destroy(gi_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    server_shutdown(c);
    sweeps_shutdown(c);
    (void)hipStreamSynchronize(c->stream);
    free_plan(c);
    for (auto &b : c->blocks) {
        if (b.owned) (void)hipFree((void *)b.d_vec);
        if (b.h_vec) (void)hipHostFree(b.h_vec);
    }
    if (c->mu_owned) (void)hipFree(c->d_mu);
    free_slot(c->main_slot);
    (void)hipStreamDestroy(c->main_slot.cstream);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *gi_last_error(const gi_ctx *c) { return c ? c->err.c_str() : "null context"; }

// Drops the bases and eigenvector blocks: they belong to the J list.
static void reset_inputs(gi_ctx *c)
{
    for (auto &b : c->blocks) {
        if (b.owned) (void)hipFree((void *)b.d_vec);
        if (b.h_vec) (void)hipHostFree(b.h_vec);
    }
    c->vec_dev_bytes = 0;
    free_split(c);
    c->blocks.clear();
    c->blocks.resize((size_t)c->nJ * c->nrep);
    c->have_bset.assign(c->nJ, 0);
    c->hb.assign(c->nJ, gi_ctx::HostBset());
}

int gi_set_params(gi_ctx *c, const gi_params *p)
{
    if (!c || !p) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (p->nrepres <= 0 || p->nrepres > 16 || !p->degen || !p->gns || !p->isym_pairs)
        return fail(c, GI_EINVAL, "gi_set_params: bad symmetry description");
    if (p->nmodes < 0 || (p->nmodes > 0 && (!p->quanta_lower || !p->quanta_upper)))
        return fail(c, GI_EINVAL, "gi_set_params: quanta bounds missing");
    if (p->jvals[0] < 0 || p->jvals[1] < p->jvals[0])
        return fail(c, GI_EINVAL, "gi_set_params: need 0 <= jvals[0] <= jvals[1]");
    if (!(p->temperature > 0.0)) return fail(c, GI_EINVAL, "gi_set_params: temperature <= 0");
    free_plan(c);
    c->nrep = p->nrepres;
    c->degen.assign(p->degen, p->degen + p->nrepres);
    for (int d : c->degen)
        if (d < 1 || d > GI_MAXDEG)
            return fail(c, GI_EINVAL, "gi_set_params: degeneracy must be 1 or 2");
    c->amp_m = *std::max_element(c->degen.begin(), c->degen.end());
    c->sym_type = p->sym_type;
    c->nmodes = p->nmodes;
    c->zpe = p->zpe;
    c->temperature = p->temperature;
    c->q_stat = p->q_stat;
    for (int i = 0; i < 2; i++) {
        c->freq_window[i] = p->freq_window[i];
        c->erange_lower[i] = p->erange_lower[i];
        c->erange_upper[i] = p->erange_upper[i];
    }
    c->nJ = 2;
    c->jvals.assign(p->jvals, p->jvals + 2);
    c->gns.assign(p->gns, p->gns + p->nrepres);
    c->isym_pairs.assign(p->isym_pairs, p->isym_pairs + p->nrepres);
    c->reduced = p->reduced;
    c->quanta_lower.assign(p->quanta_lower, p->quanta_lower + 2 * p->nmodes);
    c->quanta_upper.assign(p->quanta_upper, p->quanta_upper + 2 * p->nmodes);
    c->thresh_S = p->thresh_linestrength;
    c->thresh_I = p->thresh_intensity;
    // find_igamma_pair (trove_functions.cpp:1041-1072)
    c->igamma_pair.assign(c->nrep, 0);
    for (int gI = 0; gI < c->nrep; gI++) {
        int m = 0;
        c->igamma_pair[gI] = gI;
        for (int gF = 0; gF < c->nrep; gF++)
            if (gI != gF && c->isym_pairs[gI] == c->isym_pairs[gF]) {
                c->igamma_pair[gI] = gF;
                m++;
            }
        if (m > 1)
            return fail(c, GI_EINVAL, "selection rules do not come in pairs (SELECTION)");
        if (c->gns[gI] != c->gns[c->igamma_pair[gI]])
            return fail(c, GI_EINVAL, "selection rules do not agree with GNS");
    }
    reset_inputs(c);
    c->have_params = true;
    return GI_OK;
}

int gi_set_jvals(gi_ctx *c, int nj, const int *jvals)
{
    if (!c || !jvals || nj < 1 || nj > kMaxJ) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_set_jvals before gi_set_params");
    for (int i = 0; i < nj; i++)
        if (jvals[i] < 0 || (i > 0 && jvals[i] <= jvals[i - 1]))
            return fail(c, GI_EINVAL, "gi_set_jvals: J values must be >= 0 and ascending");
    free_plan(c);
    c->nJ = nj;
    c->jvals.assign(jvals, jvals + nj);
    reset_inputs(c);
    c->nlevels = 0;
    c->lv_energy.clear(); c->lv_jind.clear(); c->lv_igamma.clear();
    c->lv_ndeg.clear(); c->lv_irec0.clear(); c->lv_quanta.clear();
    return GI_OK;
}

int gi_upload_bset(gi_ctx *c, int jind, const gi_bset *b)
{
    if (!c || !b) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_upload_bset before gi_set_params");
    if (jind < 0 || jind >= c->nJ) return fail(c, GI_EINVAL, "gi_upload_bset: jind outside the J list");
    if (b->jval != c->jvals[jind])
        return fail(c, GI_EINVAL, "gi_upload_bset: jval differs from jvals[jind]");
    if (b->Maxcontracts < 0 || b->Maxsymcoeffs < 0 || b->mat_size < 1)
        return fail(c, GI_EINVAL, "gi_upload_bset: bad sizes");
    free_plan(c);
    auto &h = c->hb[jind];
    const int d = b->Maxcontracts, nc = b->Maxsymcoeffs, nr = c->nrep;
    h.jval = b->jval;
    h.ncases = nc;
    h.dimen = d;
    h.mat_size = b->mat_size;
    h.icontr2icase.assign(b->icontr2icase, b->icontr2icase + 2 * d);
    h.v.assign(b->iroot_correlat_j0, b->iroot_correlat_j0 + d);
    h.k.assign(b->k, b->k + d);
    h.ktau.assign(b->ktau, b->ktau + d);
    h.N.assign(b->N, b->N + (size_t)nr * nc);
    h.Ntotal.assign(b->Ntotal, b->Ntotal + nr);
    h.ijterms.assign(b->ijterms, b->ijterms + (size_t)nr * nc);
    h.nsize.assign(b->nsize, b->nsize + nr);
    h.repres.assign(nr, {});
    for (int g = 0; g < nr; g++)
        h.repres[g].assign(b->repres[g],
                           b->repres[g] + (size_t)b->Ntotal[g] * c->degen[g] * b->mat_size);
    c->have_bset[jind] = true;
    return GI_OK;
}

int gi_set_mu_budget(gi_ctx *c, long long bytes)
{
    if (!c || bytes < 0) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    c->mu_budget = bytes;
    return GI_OK;
}

int gi_mu_parts(const gi_ctx *c)
{
    if (!c || c->nvib <= 0) return 0;
    if (!c->mu_streamed) return 1;
    return (c->nvib + c->mu_np - 1) / c->mu_np;
}

int gi_upload_dipole(gi_ctx *c, int nvib, const double *mu, int on_device)
{
    if (!c || nvib <= 0 || !mu) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    if (c->mu_owned) (void)hipFree(c->d_mu);
    c->d_mu = nullptr;
    c->mu_owned = false;
    c->mu_streamed = false;
    c->mu_np = 0;
    std::vector<double>().swap(c->h_mu);
    c->nvib = nvib;
    const size_t bytes = sizeof(double) * 3 * (size_t)nvib * nvib;
    if (c->mu_budget > 0 && (long long)bytes > c->mu_budget) {
        // rows of one part: 3 * nvib doubles each
        const long long row = 3LL * nvib * (long long)sizeof(double);
        if (c->mu_budget < row)
            return fail(c, GI_EINVAL, "gi_upload_dipole: mu budget below one vF row of mu (" +
                                          std::to_string(row) + " bytes)");
        c->mu_np = (int)std::min<long long>(nvib, c->mu_budget / row);
        c->h_mu.resize((size_t)3 * nvib * nvib);
        if (on_device)
            HIPCHK(hipMemcpy(c->h_mu.data(), mu, bytes, hipMemcpyDeviceToHost));
        else
            std::memcpy(c->h_mu.data(), mu, bytes);
        c->mu_streamed = true;
        return GI_OK;
    }
    if (on_device) {
        c->d_mu = const_cast<double *>(mu);
        return GI_OK;
    }
    HIPCHK(hipMalloc((void **)&c->d_mu, bytes));
    c->mu_owned = true;
    HIPCHK(hipMemcpy(c->d_mu, mu, bytes, hipMemcpyHostToDevice));
    return GI_OK;
}

int gi_set_levels(gi_ctx *c, int n, const double *energy, const int *jind, const int *igamma,
                  const int *ndeg, const int *irec0, const int *quanta)
{
    if (!c || n < 0 || (n > 0 && (!energy || !jind || !igamma || !ndeg || !irec0 || !quanta)))
        return GI_EINVAL;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_set_levels before gi_set_params");
    if (serving(c)) return GI_ESTATE;
    free_plan(c);
    for (int i = 0; i < n; i++) {
        if (jind[i] < 0 || jind[i] >= c->nJ || igamma[i] < 0 || igamma[i] >= c->nrep)
            return fail(c, GI_EINVAL, "gi_set_levels: jind/igamma out of range at " + std::to_string(i));
        if (ndeg[i] != c->degen[igamma[i]])
            return fail(c, GI_EINVAL, "gi_set_levels: ndeg differs from the symmetry's degeneracy");
        if (i > 0 && energy[i] < energy[i - 1])
            return fail(c, GI_EINVAL, "gi_set_levels: levels must be sorted by energy");
    }
    c->nlevels = n;
    c->lv_energy.assign(energy, energy + n);
    c->lv_jind.assign(jind, jind + n);
    c->lv_igamma.assign(igamma, igamma + n);
    c->lv_ndeg.assign(ndeg, ndeg + n);
    c->lv_irec0.assign(irec0, irec0 + n);
    c->lv_quanta.assign(quanta, quanta + (size_t)n * (c->nmodes + 1));
    return GI_OK;
}

int gi_upload_vectors(gi_ctx *c, int jind, int igamma, int nsize, int nrec, const double *v,
                      int on_device)
{
    if (!c || nsize < 0 || nrec < 0) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_upload_vectors before gi_set_params");
    if (jind < 0 || jind >= c->nJ) return fail(c, GI_EINVAL, "gi_upload_vectors: jind outside the J list");
    if (igamma < 0 || igamma >= c->nrep) return GI_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    Block &B = c->blocks[jind * c->nrep + igamma];
    if (B.owned) {
        (void)hipFree((void *)B.d_vec);
        c->vec_dev_bytes -= (long long)sizeof(double) * B.nsize * B.nrec;
    }
    if (B.h_vec) (void)hipHostFree(B.h_vec);
    B.h_vec = nullptr;
    B.d_vec = nullptr;
    B.owned = false;
    B.nsize = nsize;
    B.nrec = nrec;
    if ((size_t)nsize * nrec == 0) return GI_OK;
    if (!v) return GI_EINVAL;
    if (on_device) {
        B.d_vec = v;
        return GI_OK;
    }
    const size_t bytes = sizeof(double) * (size_t)nsize * nrec;
    if (c->vec_budget > 0 && c->vec_dev_bytes + (long long)bytes > c->vec_budget) {
        // beyond the budget: pinned host memory, records staged per chunk
        HIPCHK(hipHostMalloc((void **)&B.h_vec, bytes, hipHostMallocDefault));
        std::memcpy(B.h_vec, v, bytes);
        return GI_OK;
    }
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, bytes));
    B.owned = true;
    B.d_vec = d;
    c->vec_dev_bytes += (long long)bytes;
    HIPCHK(hipMemcpy(d, v, bytes, hipMemcpyHostToDevice));
    return GI_OK;
}

int gi_set_vector_budget(gi_ctx *c, long long bytes)
{
    if (!c || bytes < 0) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    c->vec_budget = bytes;
    return GI_OK;
}

int gi_streamed_blocks(const gi_ctx *c) { return c ? host_blocks(c) : 0; }

int gi_prepare(gi_ctx *c)
{
    if (!c) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (!c->have_params || (!c->d_mu && !c->mu_streamed))
        return fail(c, GI_ESTATE, "gi_prepare: parameters, every basis and the dipole are needed");
    for (int j = 0; j < c->nJ; j++)
        if (!c->have_bset[j])
            return fail(c, GI_ESTATE, "gi_prepare: no basis for J=" + std::to_string(c->jvals[j]));
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    const int nJ = c->nJ;
    c->jp.assign(nJ, JPlan());
    c->coef.assign((size_t)nJ * nJ, Coef());
    c->jmax = *std::max_element(c->jvals.begin(), c->jvals.end());
    c->threej.assign((size_t)(c->jmax + 1) * (c->jmax + 1) * 9, 0.0);
    threej_table(c->jmax, c->threej.data());
    int rc;
    for (int j = 0; j < nJ; j++) {
        for (int g = 0; g < c->nrep; g++)
            if (c->hb[j].nsize[g] != c->hb[j].Ntotal[g])
                return fail(c, GI_EINVAL, "nsize differs from Ntotal for J=" +
                                              std::to_string(c->hb[j].jval));
        if ((rc = assign_slots(c, j)) != GI_OK) return rc;
    }
    if ((rc = build_vib_classes(c)) != GI_OK) return rc;
    if (c->mu_streamed) {
        // permuted mu in pinned host memory, element (a, vI, vF) at
        // a*N*N + vI*N + vF as on the device (k_permute_mu)
        const int N = c->nvib;
        const size_t NN = (size_t)N * N;
        HIPCHK(hipHostMalloc((void **)&c->h_mu_perm, sizeof(double) * 3 * NN, hipHostMallocDefault));
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < N; j++) {
                const double *src = c->h_mu.data() + a * NN + (size_t)c->vperm[j] * N;
                double *dst = c->h_mu_perm + a * NN + (size_t)j * N;
                for (int i = 0; i < N; i++) dst[i] = src[c->vperm[i]];
            }
    } else {
        const int N = c->nvib;
        int *d_perm = nullptr;
        if ((rc = upload(c, &d_perm, c->vperm)) != GI_OK) return rc;
        HIPCHK(hipMalloc((void **)&c->d_mu_perm, sizeof(double) * 3 * (size_t)N * N));
        launch_permute_mu(c->stream, N, d_perm, c->d_mu, c->d_mu_perm);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(d_perm);
    }
    for (int j = 0; j < nJ; j++) {
        if ((rc = build_symmaps(c, j)) != GI_OK) return rc;
        JPlan &p = c->jp[j];
        std::vector<int> rg(p.dimen);
        for (int r = 0; r < p.dimen; r++) rg[r] = p.r_slot[r] * c->nvib + p.r_v[r];
        if ((rc = upload(c, &p.d_r_grid, rg)) != GI_OK) return rc;
        for (auto &m : p.sym) {
            if ((rc = upload(c, &m.d_rptr, m.rptr)) != GI_OK) return rc;
            if ((rc = upload(c, &m.d_rcol, m.rcol)) != GI_OK) return rc;
            if ((rc = upload(c, &m.d_rval, m.rval)) != GI_OK) return rc;
            if ((rc = upload(c, &m.d_tptr, m.tptr)) != GI_OK) return rc;
            if ((rc = upload(c, &m.d_tcol, m.tcol)) != GI_OK) return rc;
            if ((rc = upload(c, &m.d_tval, m.tval)) != GI_OK) return rc;
        }
    }
    for (int jI = 0; jI < nJ; jI++)
        for (int jF = 0; jF < nJ; jF++) {
            if (std::abs(c->jvals[jI] - c->jvals[jF]) > 1) continue;
            if ((rc = build_coef(c, jI, jF)) != GI_OK) return rc;
            Coef &cf = coef_of(c, jI, jF);
            if (!cf.valid) continue;
            if ((rc = upload(c, &cf.d_ptr, cf.ptr)) != GI_OK) return rc;
            if ((rc = upload(c, &cf.d_sI, cf.sI)) != GI_OK) return rc;
            if ((rc = upload(c, &cf.d_val, cf.val)) != GI_OK) return rc;
        }
    c->groups.assign((size_t)nJ * c->nrep * nJ, GroupPlan());
    for (int jI = 0; jI < nJ; jI++)
        for (int g = 0; g < c->nrep; g++)
            for (int jF = 0; jF < nJ; jF++) {
                if (!coef_of(c, jI, jF).valid) continue;
                if ((rc = build_group(c, jI, g, jF)) != GI_OK) return rc;
                GroupPlan &G = group_of(c, jI, g, jF);
                if (!G.valid) continue;
                for (auto &oc : G.oc) {
                    std::vector<int> ints;
                    ints.insert(ints.end(), oc.seg_a.begin(), oc.seg_a.end());
                    ints.insert(ints.end(), oc.seg_off.begin(), oc.seg_off.end());
                    ints.insert(ints.end(), oc.seg_n.begin(), oc.seg_n.end());
                    ints.insert(ints.end(), oc.seg_row0.begin(), oc.seg_row0.end());
                    ints.insert(ints.end(), oc.rptr.begin(), oc.rptr.end());
                    ints.insert(ints.end(), oc.rsI.begin(), oc.rsI.end());
                    if ((rc = upload(c, &oc.d_ints, ints)) != GI_OK) return rc;
                    if ((rc = upload(c, &oc.d_rcoef, oc.rcoef)) != GI_OK) return rc;
                }
                for (int d = 0; d < c->degen[G.gF]; d++) {
                    if ((rc = upload(c, &G.d_tcol[d], G.tcol[d])) != GI_OK) return rc;
                    if (G.tfc[d].empty()) {  // the CSR only
                        (void)hipFree(G.d_tfc[d]);
                        (void)hipFree(G.d_tfw[d]);
                        G.d_tfc[d] = nullptr;
                        G.d_tfw[d] = nullptr;
                    } else if ((rc = upload(c, &G.d_tfc[d], G.tfc[d])) != GI_OK ||
                               (rc = upload(c, &G.d_tfw[d], G.tfw[d])) != GI_OK) {
                        return rc;
                    }
                    std::vector<int>().swap(G.tfc[d]);
                    std::vector<double>().swap(G.tfw[d]);
                }
                if ((rc = build_gmap(c, G)) != GI_OK) return rc;
                for (auto &oc : G.oc)
                    if ((rc = upload(c, &oc.d_segpad, oc.seg_pad)) != GI_OK) return rc;
                std::vector<ClassDev> cds;
                for (auto &oc : G.oc) cds.push_back(classdev(oc));
                if ((rc = upload(c, &G.d_cd, cds)) != GI_OK) return rc;
                if (G.gW <= kGMapMaxW)
                    for (int d = 0; d < c->degen[G.gI]; d++) {
                        if ((rc = upload(c, &G.d_gidx[d], G.gidx[d])) != GI_OK) return rc;
                        if ((rc = upload(c, &G.d_gwt[d], G.gwt[d])) != GI_OK) return rc;
                        if ((rc = upload(c, &G.d_gidxB[d], G.gidxB[d])) != GI_OK) return rc;
                        if ((rc = upload(c, &G.d_gwtB[d], G.gwtB[d])) != GI_OK) return rc;
                        if ((rc = upload(c, &G.d_gbnd[d], G.gbnd[d])) != GI_OK) return rc;
                        if (G.cmap &&
                            ((rc = upload(c, &G.d_cidx[d], G.cidx[d])) != GI_OK ||
                             (rc = upload(c, &G.d_ccode[d], G.ccode[d])) != GI_OK ||
                             (rc = upload(c, &G.d_cwt[d], G.cwt[d])) != GI_OK))
                            return rc;
                    }
                for (int d = 0; d < GI_MAXDEG; d++) {  // host copies are not needed again
                    std::vector<int>().swap(G.gidx[d]);
                    std::vector<double>().swap(G.gwt[d]);
                    std::vector<int>().swap(G.gidxB[d]);
                    std::vector<double>().swap(G.gwtB[d]);
                    std::vector<uint16_t>().swap(G.cidx[d]);
                    std::vector<uint8_t>().swap(G.ccode[d]);
                    std::vector<double>().swap(G.cwt[d]);
                }
            }
    // level filters and per-block energy order
    const int nq = c->nmodes + 1;
    c->lv_lowok.assign(c->nlevels, 0);
    c->lv_upok.assign(c->nlevels, 0);
    c->lv_low.clear();
    for (auto &B : c->blocks) {
        B.lev.clear();
        B.en.clear();
        B.upok.clear();
    }
    for (int i = 0; i < c->nlevels; i++) {
        if (c->lv_jind[i] >= nJ)
            return fail(c, GI_EINVAL, "level " + std::to_string(i) + ": jind outside the J list");
        const int J = c->jvals[c->lv_jind[i]];
        const int *q = &c->lv_quanta[(size_t)i * nq];
        c->lv_lowok[i] = energy_filter_lower(c, J, c->lv_energy[i], q);
        if (c->lv_lowok[i]) c->lv_low.push_back(i);
        c->lv_upok[i] = energy_filter_upper(c, J, c->lv_energy[i], q);
        Block &B = c->blocks[c->lv_jind[i] * c->nrep + c->lv_igamma[i]];
        if (c->lv_irec0[i] < 0 || c->lv_irec0[i] >= B.nrec)
            return fail(c, GI_EINVAL, "level " + std::to_string(i) +
                                          ": eigenvector record outside the uploaded block");
        if (B.nsize != c->hb[c->lv_jind[i]].nsize[c->lv_igamma[i]])
            return fail(c, GI_EINVAL, "eigenvector block length differs from the basis nsize");
        B.lev.push_back(i);
        B.en.push_back(c->lv_energy[i]);
        B.upok.push_back(c->lv_upok[i]);
    }
    for (auto &B : c->blocks) {
        std::vector<int> col(B.lev.size());
        for (size_t i = 0; i < B.lev.size(); i++) col[i] = c->lv_irec0[B.lev[i]];
        if ((rc = upload(c, &B.d_lev, B.lev)) != GI_OK) return rc;
        if ((rc = upload(c, &B.d_col, col)) != GI_OK) return rc;
        if ((rc = upload(c, &B.d_en, B.en)) != GI_OK) return rc;
        if ((rc = upload(c, &B.d_upok, B.upok)) != GI_OK) return rc;
    }
    // identity map over block positions: staged upper records of a block kept
    // in host memory are addressed through it
    if (host_blocks(c)) {
        int n = 1;
        for (const Block &B : c->blocks) n = std::max(n, (int)B.lev.size());
        std::vector<int> iota(n);
        for (int i = 0; i < n; i++) iota[i] = i;
        if ((rc = upload(c, &c->d_iota, iota)) != GI_OK) return rc;
        c->iota_n = n;
    }
    // partition-function inputs: level energies and weights gns * ndeg * (2J + 1)
    {
        std::vector<double> w(c->nlevels);
        for (int i = 0; i < c->nlevels; i++)
            w[i] = c->gns[c->lv_igamma[i]] * c->lv_ndeg[i] * (2.0 * c->jvals[c->lv_jind[i]] + 1.0);
        if ((rc = upload(c, &c->d_lv_energy, c->lv_energy)) != GI_OK) return rc;
        if ((rc = upload(c, &c->d_lv_weight, w)) != GI_OK) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prepared = true;
    return GI_OK;
}

int gi_half_ls(gi_ctx *c, int jindI, int jindF, int gI, int ideg, const double *vec, double *h)
{
    if (!c || !vec || !h) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_half_ls before gi_prepare");
    if (jindI < 0 || jindI >= c->nJ || jindF < 0 || jindF >= c->nJ) return GI_EINVAL;
    if (gI < 0 || gI >= c->nrep || ideg < 0 || ideg >= c->degen[gI]) return GI_EINVAL;
    const Coef &cf = coef_of(c, jindI, jindF);
    if (!cf.valid) return fail(c, GI_EINVAL, "gi_half_ls: |J' - J| > 1");
    if (c->mu_streamed) return fail(c, GI_ESTATE, "gi_half_ls: needs mu resident");
    HIPCHK(hipSetDevice(c->device));
    const JPlan &pI = c->jp[jindI], &pF = c->jp[jindF];
    const int N = c->nvib, SI = pI.nslots, SF = pF.nslots, SMAX = smax_for(SF);
    if (SMAX < 0) return fail(c, GI_EINVAL, "too many rotational slots");
    const int nsizeI = c->hb[jindI].nsize[gI];
    int rc;
    double *d_u = (double *)scratch(c, "hl_u", (size_t)std::max(nsizeI, 1) * 8, &rc); if (rc) return rc;
    double *d_G = (double *)scratch(c, "hl_G", (size_t)3 * N * SMAX * 8, &rc); if (rc) return rc;
    double *d_H = (double *)scratch(c, "hl_H", (size_t)SF * N * 8, &rc); if (rc) return rc;
    double *d_h = (double *)scratch(c, "hl_h", (size_t)std::max(pF.dimen, 1) * 8, &rc); if (rc) return rc;
    struct { long long p; int ideg; int pad; } m;
    m.p = (long long)d_u; m.ideg = ideg; m.pad = 0;
    unsigned char *d_m = (unsigned char *)scratch(c, "hl_m", 16, &rc); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_u, vec, (size_t)nsizeI * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_m, &m, 16, hipMemcpyHostToDevice, c->stream));
    SymDev r = symdev(pI.sym[gI * GI_MAXDEG + ideg]);
    // unit_ideg = 0 selects r0; pass the component's map as r0
    launch_build_g(c->stream, 1, N, SI, SF, SMAX, (const double *const *)d_m,
                   (const int *)(d_m + 8) /* ideg field: treated as 0/1 selector */, r, r,
                   cf.d_ptr, cf.d_sI, cf.d_val, d_G);
    if (launch_half_ls_gemv(c->stream, 1, N, SF, SMAX, c->d_mu_perm, d_G, d_H) != 0)
        return fail(c, GI_EINVAL, "no half-ls kernel for this slot count");
    launch_grid_to_prim(c->stream, pF.dimen, pF.d_r_grid, d_H, d_h);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, d_h, (size_t)pF.dimen * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GI_OK;
}

int gi_correlate(gi_ctx *c, int jind, int g, int ideg, const double *vec, double *out)
{
    if (!c || !vec || !out) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_correlate before gi_prepare");
    if (jind < 0 || jind >= c->nJ) return GI_EINVAL;
    if (g < 0 || g >= c->nrep || ideg < 0 || ideg >= c->degen[g]) return GI_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    const JPlan &p = c->jp[jind];
    const int nsize = c->hb[jind].nsize[g];
    int rc;
    double *d_u = (double *)scratch(c, "co_u", (size_t)std::max(nsize, 1) * 8, &rc); if (rc) return rc;
    double *d_o = (double *)scratch(c, "co_o", (size_t)std::max(p.dimen, 1) * 8, &rc); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_u, vec, (size_t)nsize * 8, hipMemcpyHostToDevice, c->stream));
    launch_correlate_prim(c->stream, p.dimen, p.d_r_grid, symdev(p.sym[g * GI_MAXDEG + ideg]), d_u, d_o);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_o, (size_t)p.dimen * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GI_OK;
}

int gi_threej_table(int jmax, double *out)
{
    if (jmax < 0 || !out) return GI_EINVAL;
    threej_table(jmax, out);
    return GI_OK;
}

int gi_kernel_stats(gi_ctx *c, int max, const char **names, double *total_ms, int *launches, int *n)
{
    if (!c || !n) return GI_EINVAL;
    *n = (int)c->kstat_names.size();
    for (int i = 0; i < *n && i < max; i++) {
        const auto &s = c->kstats[c->kstat_names[i]];
        if (names) names[i] = c->kstat_names[i].c_str();
        if (total_ms) total_ms[i] = s.ms;
        if (launches) launches[i] = s.launches;
    }
    return GI_OK;
}

int gi_set_split_format(gi_ctx *c, int radix, int halfls_digits, int corr_digits)
{
    const auto ok_digits = [](int d) { return d == 0 || (d >= 3 && d <= 6); };
    if (!c || (radix != 0 && radix != 128 && radix != 254) || !ok_digits(halfls_digits) ||
        !ok_digits(corr_digits))
        return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (radix != c->split_req_radix || halfls_digits != c->split_req_Sh ||
        corr_digits != c->split_req_S) {
        HIPCHK(hipSetDevice(c->device));
        free_split(c);
        c->split_req_radix = radix;
        c->split_req_Sh = halfls_digits;
        c->split_req_S = corr_digits;
    }
    return GI_OK;
}

int gi_get_split_format(gi_ctx *c, int *radix, int *halfls_digits, int *corr_digits)
{
    if (!c || !radix || !halfls_digits || !corr_digits) return GI_EINVAL;
    *radix = (int)c->split_radix;
    *halfls_digits = c->split_Sh;
    *corr_digits = c->split_S;
    return GI_OK;
}

int gi_set_split_digits_each(gi_ctx *c, int halfls_digits, int corr_digits)
{
    if (!c || halfls_digits < 3 || halfls_digits > 6 || corr_digits < 3 || corr_digits > 6)
        return GI_EINVAL;
    return gi_set_split_format(c, c->split_req_radix, halfls_digits, corr_digits);
}

int gi_set_split_digits(gi_ctx *c, int digits) { return gi_set_split_digits_each(c, digits, digits); }

int gi_prepare_split(gi_ctx *c)
{
    if (!c) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    return prepare_split(c);
}

int gi_set_tiles(gi_ctx *c, int hl, int cr)
{
    if (!c || hl < -1 || hl > kTileHalflsMax || cr < -1 || cr > 3) return GI_EINVAL;
    if (serving(c)) return GI_ESTATE;
    if (hl >= 0) c->tile_halfls = hl;
    if (cr >= 0) c->tile_corr = cr;
    return GI_OK;
}

int gi_get_tiles(gi_ctx *c, int *hl, int *cr)
{
    if (!c) return GI_EINVAL;
    if (hl) *hl = c->tile_halfls;
    if (cr) *cr = c->tile_corr;
    return GI_OK;
}

int gi_work_stats(gi_ctx *c, double *out, int n)
{
    if (!c || !out) return GI_EINVAL;
    for (int i = 0; i < n && i < kWorkStats; i++) out[i] = c->work[i];
    return GI_OK;
}

}  // extern "C"
