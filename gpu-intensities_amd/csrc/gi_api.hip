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

int fail(gi_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(c, e_ == hipErrorOutOfMemory ? GI_ENOMEM : GI_EHIP,             \
                        std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

void *scratch(gi_ctx *c, const char *name, size_t bytes, int *rc)
{
    *rc = GI_OK;
    auto &e = c->scratch[name];
    if (e.second >= bytes && e.first) return e.first;
    if (e.first) (void)hipFree(e.first);
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

static void free_plan(gi_ctx *c)
{
    for (int j = 0; j < 2; j++) {
        for (auto &m : c->jp[j].sym) free_sym(m);
        (void)hipFree(c->jp[j].d_r_grid);
        c->jp[j] = JPlan();
        for (int f = 0; f < 2; f++) {
            Coef &cf = c->coef[j][f];
            (void)hipFree(cf.d_ptr); (void)hipFree(cf.d_sI); (void)hipFree(cf.d_val);
            cf = Coef();
        }
    }
    for (auto &b : c->blocks) {
        (void)hipFree(b.d_lev); (void)hipFree(b.d_col); (void)hipFree(b.d_en);
        (void)hipFree(b.d_upok);
        b.d_lev = b.d_col = nullptr;
        b.d_en = nullptr;
        b.d_upok = nullptr;
    }
    c->prepared = false;
}

static SymDev symdev(const SymMap &m)
{
    return SymDev{m.d_rptr, m.d_rcol, m.d_rval, m.d_tptr, m.d_tcol, m.d_tval};
}

// Timed launches: events recorded on the context's stream around each launch.
struct Timeline {
    gi_ctx *c;
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> recs;
    explicit Timeline(gi_ctx *c_) : c(c_) {}
    int begin(const char *name)
    {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return GI_EHIP;
        if (hipEventCreate(&b) != hipSuccess) return GI_EHIP;
        recs.emplace_back(name, a, b);
        return hipEventRecord(a, c->stream) == hipSuccess ? GI_OK : GI_EHIP;
    }
    int end() { return hipEventRecord(std::get<2>(recs.back()), c->stream) == hipSuccess ? GI_OK : GI_EHIP; }
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

static int smax_for(int SF)
{
    const int opts[] = {8, 16, 24, 32, 48, 64};
    for (int o : opts)
        if (SF <= o) return o;
    return -1;
}

static bool degeneracy_filter(const gi_ctx *c, int gI, int gF, int idegI, int idegF)
{
    // trove_functions.cpp:1495-1510
    if (c->degen[gI] == 1 || c->degen[gF] == 1) return true;
    if (!c->reduced) return true;
    if (c->sym_type == 2 || c->sym_type == 3) return idegI == 1 && idegF == 0;
    return true;
}

struct Item {
    int level, jindI, gI, jindF, gF, lo, hi;
    long long out_off;
};

// Units of work for the given lower levels: one item per (lower level, J')
// whose window in the upper block is not empty (the lower-level filters of
// cuda_host.cu:1194 and the J/symmetry part of intensity_filter,
// trove_functions.cpp:1452-1464, are decided here; the energy window is a
// contiguous range of the energy-sorted block).
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
            const Block &B = c->blocks[jindF * c->nrep + gF];
            const auto &en = B.en;
            // first f with en[f] - eI >= lo; first f with en[f] - eI > hi
            const double w0 = c->freq_window[0], w1 = c->freq_window[1];
            int a = (int)(std::partition_point(en.begin(), en.end(),
                                               [&](double e) { return !(e - eI >= w0); }) -
                          en.begin());
            int b = (int)(std::partition_point(en.begin(), en.end(),
                                               [&](double e) { return e - eI <= w1; }) -
                          en.begin());
            if (b <= a) continue;
            items.push_back(Item{L, jindI, gI, jindF, gF, a, b, ncand});
            ncand += b - a;
        }
    }
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
    *out = c;
    return GI_OK;
}

void gi_destroy(gi_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_plan(c);
    for (auto &b : c->blocks)
        if (b.owned) (void)hipFree((void *)b.d_vec);
    if (c->mu_owned) (void)hipFree(c->d_mu);
    for (auto &s : c->scratch) (void)hipFree(s.second.first);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *gi_last_error(const gi_ctx *c) { return c ? c->err.c_str() : "null context"; }

int gi_set_params(gi_ctx *c, const gi_params *p)
{
    if (!c || !p) return GI_EINVAL;
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
    c->sym_type = p->sym_type;
    c->nmodes = p->nmodes;
    c->zpe = p->zpe;
    c->temperature = p->temperature;
    c->q_stat = p->q_stat;
    for (int i = 0; i < 2; i++) {
        c->freq_window[i] = p->freq_window[i];
        c->erange_lower[i] = p->erange_lower[i];
        c->erange_upper[i] = p->erange_upper[i];
        c->jvals[i] = p->jvals[i];
    }
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
    for (auto &b : c->blocks)
        if (b.owned) (void)hipFree((void *)b.d_vec);
    c->blocks.clear();
    c->blocks.resize(2 * c->nrep);
    c->have_bset[0] = c->have_bset[1] = false;
    c->have_params = true;
    return GI_OK;
}

int gi_upload_bset(gi_ctx *c, int jind, const gi_bset *b)
{
    if (!c || !b || jind < 0 || jind > 1) return GI_EINVAL;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_upload_bset before gi_set_params");
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

int gi_upload_dipole(gi_ctx *c, int nvib, const double *mu, int on_device)
{
    if (!c || nvib <= 0 || !mu) return GI_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    if (c->mu_owned) (void)hipFree(c->d_mu);
    c->d_mu = nullptr;
    c->mu_owned = false;
    c->nvib = nvib;
    if (on_device) {
        c->d_mu = const_cast<double *>(mu);
        return GI_OK;
    }
    const size_t bytes = sizeof(double) * 3 * (size_t)nvib * nvib;
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
    free_plan(c);
    for (int i = 0; i < n; i++) {
        if (jind[i] < 0 || jind[i] > 1 || igamma[i] < 0 || igamma[i] >= c->nrep)
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
    if (!c || jind < 0 || jind > 1 || nsize < 0 || nrec < 0) return GI_EINVAL;
    if (!c->have_params) return fail(c, GI_ESTATE, "gi_upload_vectors before gi_set_params");
    if (igamma < 0 || igamma >= c->nrep) return GI_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    Block &B = c->blocks[jind * c->nrep + igamma];
    if (B.owned) (void)hipFree((void *)B.d_vec);
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
    double *d = nullptr;
    const size_t bytes = sizeof(double) * (size_t)nsize * nrec;
    HIPCHK(hipMalloc((void **)&d, bytes));
    B.owned = true;
    B.d_vec = d;
    HIPCHK(hipMemcpy(d, v, bytes, hipMemcpyHostToDevice));
    return GI_OK;
}

int gi_prepare(gi_ctx *c)
{
    if (!c) return GI_EINVAL;
    if (!c->have_params || !c->have_bset[0] || !c->have_bset[1] || !c->d_mu)
        return fail(c, GI_ESTATE, "gi_prepare: parameters, both bases and the dipole are needed");
    HIPCHK(hipSetDevice(c->device));
    free_plan(c);
    c->jmax = std::max(c->jvals[0], c->jvals[1]);
    c->threej.assign((size_t)(c->jmax + 1) * (c->jmax + 1) * 9, 0.0);
    threej_table(c->jmax, c->threej.data());
    int rc;
    for (int j = 0; j < 2; j++) {
        for (int g = 0; g < c->nrep; g++)
            if (c->hb[j].nsize[g] != c->hb[j].Ntotal[g])
                return fail(c, GI_EINVAL, "nsize differs from Ntotal for J=" +
                                              std::to_string(c->hb[j].jval));
        if ((rc = build_jplan(c, j)) != GI_OK) return rc;
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
    for (int jI = 0; jI < 2; jI++)
        for (int jF = 0; jF < 2; jF++) {
            if ((rc = build_coef(c, jI, jF)) != GI_OK) return rc;
            Coef &cf = c->coef[jI][jF];
            if (!cf.valid) continue;
            if ((rc = upload(c, &cf.d_ptr, cf.ptr)) != GI_OK) return rc;
            if ((rc = upload(c, &cf.d_sI, cf.sI)) != GI_OK) return rc;
            if ((rc = upload(c, &cf.d_val, cf.val)) != GI_OK) return rc;
        }
    // level filters and per-block energy order
    const int nq = c->nmodes + 1;
    c->lv_lowok.assign(c->nlevels, 0);
    c->lv_upok.assign(c->nlevels, 0);
    for (auto &B : c->blocks) {
        B.lev.clear();
        B.en.clear();
        B.upok.clear();
    }
    for (int i = 0; i < c->nlevels; i++) {
        const int J = c->jvals[c->lv_jind[i]];
        const int *q = &c->lv_quanta[(size_t)i * nq];
        c->lv_lowok[i] = energy_filter_lower(c, J, c->lv_energy[i], q);
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
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prepared = true;
    return GI_OK;
}

int gi_count_candidates(gi_ctx *c, int nlower, const int *ids, long long *count)
{
    if (!c || !count) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_count_candidates before gi_prepare");
    std::vector<Item> items;
    long long n = 0;
    make_items(c, nlower, ids, items, n);
    *count = n;
    return GI_OK;
}

int gi_compute(gi_ctx *c, int mode, int nlower, const int *ids, gi_lines *out)
{
    if (!c || !out) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_compute before gi_prepare");
    if (mode != GI_MODE_FP64) return fail(c, GI_EINVAL, "gi_compute: unknown mode");
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

    // group items by (jindI, gI, jindF); chunks of at most kChunk items
    const int kChunk = 256;
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (int i = 0; i < (int)items.size(); i++)
        groups[std::make_tuple(items[i].jindI, items[i].gI, items[i].jindF)].push_back(i);

    struct Chunk {
        int jindI, gI, jindF, gF, nitems, nunits, ncand, cb_mask;
        size_t off_uptr, off_uideg, off_iunit, off_cstart, off_lo, off_level, off_eI, off_out;
    };
    std::vector<Chunk> chunks;
    // packed metadata: 8-byte aligned sections in one host buffer
    std::vector<unsigned char> meta;
    auto put = [&](const void *src, size_t bytes) {
        size_t o = (meta.size() + 255) & ~size_t(255);
        meta.resize(o + std::max<size_t>(bytes, 8));
        if (bytes) std::memcpy(meta.data() + o, src, bytes);
        return o;
    };
    int max_units = 0, max_items = 0, max_SF = 0, max_ncb = 1, max_nsizeF = 0;
    std::vector<size_t> uptr_offsets;  // positions of unit pointers to patch
    for (auto &kv : groups) {
        const int jindI = std::get<0>(kv.first), gI = std::get<1>(kv.first),
                  jindF = std::get<2>(kv.first);
        const int gF = c->igamma_pair[gI];
        const int ndegI = c->degen[gI], ndegF = c->degen[gF];
        const Block &BI = c->blocks[jindI * c->nrep + gI];
        const std::vector<int> &its = kv.second;
        for (size_t s0 = 0; s0 < its.size(); s0 += kChunk) {
            const int n_it = (int)std::min<size_t>(kChunk, its.size() - s0);
            std::vector<long long> uptr;
            std::vector<int> uideg, iunit(n_it * GI_MAXDEG, -1), cstart(n_it), lo(n_it), lev(n_it);
            std::vector<long long> outoff(n_it);
            std::vector<double> eI(n_it);
            int cnt = 0;
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
            // unit pointers are offsets now; patched to device addresses below
            ch.off_uptr = put(uptr.data(), uptr.size() * 8);
            uptr_offsets.push_back(ch.off_uptr);
            ch.off_uideg = put(uideg.data(), uideg.size() * 4);
            ch.off_iunit = put(iunit.data(), iunit.size() * 4);
            ch.off_cstart = put(cstart.data(), cstart.size() * 4);
            ch.off_lo = put(lo.data(), lo.size() * 4);
            ch.off_level = put(lev.data(), lev.size() * 4);
            ch.off_eI = put(eI.data(), eI.size() * 8);
            ch.off_out = put(outoff.data(), outoff.size() * 8);
            chunks.push_back(ch);
            max_units = std::max(max_units, ch.nunits);
            max_items = std::max(max_items, n_it);
            max_SF = std::max(max_SF, c->jp[jindF].nslots);
            max_ncb = std::max(max_ncb, ndegI * ndegF);
            max_nsizeF = std::max(max_nsizeF, c->blocks[jindF * c->nrep + gF].nsize);
        }
    }
    // patch unit pointers (offsets -> device addresses)
    for (auto &ch : chunks) {
        long long *p = (long long *)(meta.data() + ch.off_uptr);
        const double *base = c->blocks[ch.jindI * c->nrep + ch.gI].d_vec;
        for (int u = 0; u < ch.nunits; u++) p[u] = (long long)(base + p[u]);
    }
    unsigned char *d_meta = (unsigned char *)scratch(c, "meta", meta.size(), &rc); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_meta, meta.data(), meta.size(), hipMemcpyHostToDevice, c->stream));
    const int smax_all = smax_for(max_SF);
    if (smax_all < 0) return fail(c, GI_EINVAL, "too many rotational slots (J too large)");
    double *G = (double *)scratch(c, "G", (size_t)max_units * 3 * N * smax_all * 8, &rc); if (rc) return rc;
    double *H = (double *)scratch(c, "H", (size_t)max_units * max_SF * N * 8, &rc); if (rc) return rc;
    double *Ht = (double *)scratch(c, "Ht", (size_t)max_items * max_ncb * max_nsizeF * 8, &rc); if (rc) return rc;

    Timeline tl(c);
    double useful = 0, issued = 0, corr_bytes = 0, mu_bytes = 0;
    for (auto &ch : chunks) {
        const JPlan &pI = c->jp[ch.jindI], &pF = c->jp[ch.jindF];
        const Coef &cf = c->coef[ch.jindI][ch.jindF];
        const Block &BF = c->blocks[ch.jindF * c->nrep + ch.gF];
        const int SI = pI.nslots, SF = pF.nslots, SMAX = smax_for(SF);
        const int ndegI = c->degen[ch.gI], ndegF = c->degen[ch.gF];
        SymDev r0 = symdev(pI.sym[ch.gI * GI_MAXDEG + 0]);
        SymDev r1 = ndegI > 1 ? symdev(pI.sym[ch.gI * GI_MAXDEG + 1]) : r0;
        SymDev t0 = symdev(pF.sym[ch.gF * GI_MAXDEG + 0]);
        SymDev t1 = ndegF > 1 ? symdev(pF.sym[ch.gF * GI_MAXDEG + 1]) : t0;
        if (ch.nunits > 0) {
            if ((rc = tl.begin("build_g")) != GI_OK) return fail(c, rc, "event");
            launch_build_g(c->stream, ch.nunits, N, SI, SF, SMAX,
                           (const double *const *)(d_meta + ch.off_uptr),
                           (const int *)(d_meta + ch.off_uideg), r0, r1, cf.d_ptr, cf.d_sI,
                           cf.d_val, G);
            tl.end();
            if ((rc = tl.begin("half_ls_gemv")) != GI_OK) return fail(c, rc, "event");
            if (launch_half_ls_gemv(c->stream, ch.nunits, N, SF, SMAX, c->d_mu, G, H) != 0)
                return fail(c, GI_EINVAL, "no half-ls kernel for this slot count");
            tl.end();
        }
        if ((rc = tl.begin("fold")) != GI_OK) return fail(c, rc, "event");
        launch_fold(c->stream, ch.nitems, BF.nsize, ndegI, ndegF, SF * N,
                    (const int *)(d_meta + ch.off_iunit), t0, t1, H, Ht);
        tl.end();
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
        a.vec = BF.d_vec; a.Ht = Ht;
        a.gnsI = c->gns[ch.gI]; a.zpe = c->zpe; a.beta = beta; a.q_stat = c->q_stat;
        a.A_coef_s_1 = A_coef_s_1; a.intens_cm_mol = intens_cm_mol;
        a.thresh_S = c->thresh_S; a.thresh_I = c->thresh_I;
        a.flag = flag; a.out_iI = cand.iI; a.out_iF = cand.iF; a.out_nu = cand.nu;
        a.out_S = cand.S; a.out_A = cand.A; a.out_I = cand.I; a.out_amp = cand.amp;
        if ((rc = tl.begin("corr_intensity")) != GI_OK) return fail(c, rc, "event");
        launch_corr_intensity(c->stream, a);
        tl.end();
        useful += ch.nunits * 2.0 * cf.C * (double)N * N;
        issued += ch.nunits * 2.0 * 3.0 * SF * (double)N * N;
        mu_bytes += ch.nunits * 24.0 * (double)N * N;
        corr_bytes += (double)ch.ncand * 8.0 * BF.nsize;
    }
    HIPCHK(hipGetLastError());
    const int nb = compact_blocks(ncand);
    int *counts = (int *)scratch(c, "counts", (size_t)nb * 4, &rc); if (rc) return rc;
    long long *offs = (long long *)scratch(c, "offs", (size_t)(nb + 1) * 8, &rc); if (rc) return rc;
    if ((rc = tl.begin("compact")) != GI_OK) return fail(c, rc, "event");
    launch_compact(c->stream, ncand, flag, counts, offs, cand, comp);
    tl.end();
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

int gi_half_ls(gi_ctx *c, int jindI, int jindF, int gI, int ideg, const double *vec, double *h)
{
    if (!c || !vec || !h || jindI < 0 || jindI > 1 || jindF < 0 || jindF > 1) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_half_ls before gi_prepare");
    if (gI < 0 || gI >= c->nrep || ideg < 0 || ideg >= c->degen[gI]) return GI_EINVAL;
    const Coef &cf = c->coef[jindI][jindF];
    if (!cf.valid) return fail(c, GI_EINVAL, "gi_half_ls: |J' - J| > 1");
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
    if (launch_half_ls_gemv(c->stream, 1, N, SF, SMAX, c->d_mu, d_G, d_H) != 0)
        return fail(c, GI_EINVAL, "no half-ls kernel for this slot count");
    launch_grid_to_prim(c->stream, pF.dimen, pF.d_r_grid, d_H, d_h);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, d_h, (size_t)pF.dimen * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GI_OK;
}

int gi_correlate(gi_ctx *c, int jind, int g, int ideg, const double *vec, double *out)
{
    if (!c || !vec || !out || jind < 0 || jind > 1) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_correlate before gi_prepare");
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

int gi_work_stats(gi_ctx *c, double *out, int n)
{
    if (!c || !out) return GI_EINVAL;
    for (int i = 0; i < n && i < 8; i++) out[i] = c->work[i];
    return GI_OK;
}

}  // extern "C"
