// Host-side plan of the line-intensity engine: Wigner 3j table, slot grids,
// symmetrisation maps and rotational coefficient tables. Everything here runs
// once per job on the host; the per-transition work is on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "gi_internal.h"

namespace gi {

// Fortran NINT, as used by the reference's 3j code (trove_functions.cpp:20).
static int nint_(double a) { return a >= 0.0 ? (int)(a + 0.5) : (int)(a - 0.5); }

// a!/10^a, accumulated as a product of (a/10 - i/10) terms so that large J
// stay in range (TROVE's fakt, trove_functions.cpp:321-344).
static double fakt_(double a)
{
    if (std::fabs(a) < 1e-24) return 1.0;
    if (a < 0.0) return NAN;
    const int ic = nint_(a);
    const double ax = a / 10.0;
    double f = ax;
    for (int i = 1; i < ic; i++) f *= (ax - (double)i * 0.1);
    return f;
}

// Wigner 3j symbol by the Racah formula with the scaled factorials above
// (trove_functions.cpp:348-399).
double three_j(int j1, int j2, int j3, int k1, int k2, int k3)
{
    const double a = j1, b = j2, c = j3, al = k1, be = k2, ga = k3;
    if (c > a + b || c < std::fabs(a - b)) return 0.0;
    if (a < 0.0 || b < 0.0 || c < 0.0) return 0.0;
    if (a < std::fabs(al) || b < std::fabs(be) || c < std::fabs(ga)) return 0.0;
    if (-1.0 * ga != al + be) return 0.0;
    const double delta =
        std::sqrt(fakt_(a + b - c) * fakt_(a + c - b) * fakt_(b + c - a) / fakt_(a + b + c + 1.0));
    const double term = std::sqrt((2.0 * c + 1.0) * (fakt_(a + al) * fakt_(a - al)) *
                                  (fakt_(b - be) * fakt_(b + be)) * (fakt_(c + ga) * fakt_(c - ga)));
    const int n0 = nint_(std::max(std::max(a + be - c, b - c - al), 0.0));
    const int n1 = nint_(std::min(std::min(a - al, b + be), a + b - c));
    double sum = 0.0;
    for (int n = n0; n <= n1; n++) {
        const double x = (double)n;
        const double den = (fakt_(a - al - x) * fakt_(c - b + al + x)) *
                           (fakt_(b + be - x) * fakt_(c - a - be + x)) *
                           (fakt_(x) * fakt_(a + b - c - x));
        sum += ((n & 1) ? -1.0 : 1.0) / den;
    }
    const double cg = delta * term * sum / std::sqrt(10.0);
    const int ph = nint_(a - b - ga);
    return ((ph % 2 == 0) ? 1.0 : -1.0) * cg / std::sqrt(2.0 * c + 1.0);
}

// Table layout of precompute_threej (trove_functions.cpp:899-925); entries
// outside the valid (kI, kF) range are zero.
void threej_table(int jmax, double *T)
{
    const int s = jmax + 1;
    std::memset(T, 0, sizeof(double) * (size_t)s * s * 9);
    for (int jI = 0; jI <= jmax; jI++)
        for (int jF = std::max(jI - 1, 0); jF <= std::min(jI + 1, jmax); jF++)
            for (int kI = 0; kI <= jI; kI++)
                for (int kF = std::max(kI - 1, 0); kF <= std::min(kI + 1, jF); kF++)
                    T[jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * s * s * 3] =
                        three_j(jI, 1, jF, kI, kF - kI, -kF);
}

// trove_functions.cpp:1432-1441
bool energy_filter_lower(const gi_ctx *c, int J, double e, const int *q)
{
    for (int i = 0; i < c->nmodes; i++) {
        if (q[i + 1] < c->quanta_lower[i]) return false;
        if (q[i + 1] > c->quanta_lower[c->nmodes + i]) return false;
    }
    return J >= c->jvals.front() && J <= c->jvals.back() && (e - c->zpe) >= c->erange_lower[0] &&
           (e - c->zpe) <= c->erange_lower[1];
}

// trove_functions.cpp:1442-1451
bool energy_filter_upper(const gi_ctx *c, int J, double e, const int *q)
{
    for (int i = 0; i < c->nmodes; i++) {
        if (q[i + 1] < c->quanta_upper[i]) return false;
        if (q[i + 1] > c->quanta_upper[c->nmodes + i]) return false;
    }
    return J >= c->jvals.front() && J <= c->jvals.back() && (e - c->zpe) >= c->erange_upper[0] &&
           (e - c->zpe) <= c->erange_upper[1];
}

// Slot grid of basis jind: every primitive r becomes a (slot, v) cell. A slot
// carries the rotational labels (K, tau); a repeated (v, K, tau) opens
// another slot with the same labels. Slots are ordered by (K, tau).
int assign_slots(gi_ctx *c, int jind)
{
    const auto &b = c->hb[jind];
    JPlan &p = c->jp[jind];
    p = JPlan();
    p.J = b.jval;
    p.dimen = b.dimen;
    const int N = c->nvib;
    std::map<std::pair<int, int>, std::vector<int>> slots_of;  // (K,tau) -> slot ids
    std::vector<std::vector<char>> used;                       // [slot][v]
    p.r_slot.resize(b.dimen);
    p.r_v.resize(b.dimen);
    for (int r = 0; r < b.dimen; r++) {
        const int v = b.v[r], K = b.k[r], tau = b.ktau[r] & 1;
        if (v < 0 || v >= N)
            return fail(c, GI_EINVAL, "basis J=" + std::to_string(b.jval) +
                                          ": iroot_correlat_j0 out of the dipole's range");
        if (K < 0 || K > b.jval)
            return fail(c, GI_EINVAL, "basis J=" + std::to_string(b.jval) + ": k out of range");
        const int irow = b.icontr2icase[r], ib = b.icontr2icase[r + b.dimen];
        if (irow < 0 || irow >= b.ncases || ib < 0 || ib >= b.mat_size)
            return fail(c, GI_EINVAL, "icontr2icase out of range");
        auto &ids = slots_of[{K, tau}];
        int chosen = -1;
        for (int s : ids)
            if (!used[s][v]) { chosen = s; break; }
        if (chosen < 0) {
            chosen = p.nslots++;
            p.slot_k.push_back(K);
            p.slot_tau.push_back(tau);
            used.emplace_back(N, 0);
            ids.push_back(chosen);
        }
        used[chosen][v] = 1;
        p.r_slot[r] = chosen;
        p.r_v[r] = v;
    }
    std::vector<int> order(p.nslots);
    for (int s = 0; s < p.nslots; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (p.slot_k[x] != p.slot_k[y]) return p.slot_k[x] < p.slot_k[y];
        return p.slot_tau[x] < p.slot_tau[y];
    });
    std::vector<int> newid(p.nslots), sk(p.nslots), st(p.nslots);
    for (int i = 0; i < p.nslots; i++) {
        newid[order[i]] = i;
        sk[i] = p.slot_k[order[i]];
        st[i] = p.slot_tau[order[i]];
    }
    p.slot_k = sk;
    p.slot_tau = st;
    for (int r = 0; r < b.dimen; r++) p.r_slot[r] = newid[p.r_slot[r]];
    return GI_OK;
}

// Symmetry classes of the vibrational index: v and v' are in one class when,
// for every (J, symmetry, component, slot), the cells (slot, v) and
// (slot, v') are both or neither touched by the symmetrisation map R. Classes
// are made contiguous by a permutation applied to the grid and to mu, so the
// half linestrength only multiplies the blocks of mu that connect needed
// cells (for C2v: the four vibrational symmetries).
int build_vib_classes(gi_ctx *c)
{
    const int N = c->nvib;
    std::vector<std::vector<long long>> sig(N);
    for (int jind = 0; jind < c->nJ; jind++) {
        const auto &b = c->hb[jind];
        const JPlan &p = c->jp[jind];
        for (int r = 0; r < b.dimen; r++) {
            const int irow = b.icontr2icase[r], ib = b.icontr2icase[r + b.dimen];
            for (int g = 0; g < c->nrep; g++) {
                const int nelem = b.N[g * b.ncases + irow];
                const int iterm = b.ijterms[irow + g * b.ncases];
                const int Ntot = b.Ntotal[g], sdeg = c->degen[g];
                for (int ideg = 0; ideg < sdeg; ideg++) {
                    bool nz = false;
                    for (int e = 0; e < nelem; e++) {
                        const int is = iterm + e;
                        if (is < 0 || is >= Ntot || is >= b.nsize[g])
                            return fail(c, GI_EINVAL, "ijterms/N exceed Ntotal or nsize");
                        if (b.repres[g][(size_t)is + (size_t)ideg * Ntot +
                                        (size_t)ib * sdeg * Ntot] != 0.0)
                            nz = true;
                    }
                    if (nz)
                        sig[b.v[r]].push_back(
                            (((long long)(jind * c->nrep + g) * GI_MAXDEG + ideg) << 20) +
                            p.r_slot[r]);
                }
            }
        }
    }
    std::map<std::vector<long long>, int> cls_of;
    std::vector<int> vcls(N);
    for (int v = 0; v < N; v++) {
        std::sort(sig[v].begin(), sig[v].end());
        sig[v].erase(std::unique(sig[v].begin(), sig[v].end()), sig[v].end());
        auto it = cls_of.find(sig[v]);
        if (it == cls_of.end()) it = cls_of.emplace(sig[v], (int)cls_of.size()).first;
        vcls[v] = it->second;
    }
    const int ncls = (int)cls_of.size();
    c->vperm.resize(N);
    c->vpos.resize(N);
    for (int v = 0; v < N; v++) c->vperm[v] = v;
    std::stable_sort(c->vperm.begin(), c->vperm.end(),
                     [&](int x, int y) { return vcls[x] < vcls[y]; });
    for (int i = 0; i < N; i++) c->vpos[c->vperm[i]] = i;
    c->vcls_off.assign(ncls, 0);
    c->vcls_n.assign(ncls, 0);
    for (int v = 0; v < N; v++) c->vcls_n[vcls[v]]++;
    for (int k = 1; k < ncls; k++) c->vcls_off[k] = c->vcls_off[k - 1] + c->vcls_n[k - 1];
    c->vpad_off.assign(ncls, 0);
    c->kpad_all = 0;
    for (int k = 0; k < ncls; k++) {
        c->vpad_off[k] = c->kpad_all;
        c->kpad_all += (c->vcls_n[k] + 63) / 64 * 64;
    }
    for (int jind = 0; jind < c->nJ; jind++)
        for (auto &v : c->jp[jind].r_v) v = c->vpos[v];
    return GI_OK;
}

// Symmetrisation maps of basis jind over the (permuted) slot grid.
int build_symmaps(gi_ctx *c, int jind)
{
    const auto &b = c->hb[jind];
    JPlan &p = c->jp[jind];
    const int N = c->nvib;
    const long long G = (long long)p.nslots * N;
    p.sym.assign((size_t)c->nrep * GI_MAXDEG, SymMap());
    for (int g = 0; g < c->nrep; g++) {
        const int Ntot = b.Ntotal[g];
        const int sdeg = c->degen[g];
        for (int ideg = 0; ideg < sdeg && ideg < GI_MAXDEG; ideg++) {
            SymMap &m = p.sym[g * GI_MAXDEG + ideg];
            m.nsize = b.nsize[g];
            std::vector<std::vector<std::pair<int, double>>> rows(G);
            for (int r = 0; r < b.dimen; r++) {
                const int irow = b.icontr2icase[r];
                const int ib = b.icontr2icase[r + b.dimen];
                const int iterm = b.ijterms[irow + g * b.ncases];
                const int nelem = b.N[g * b.ncases + irow];
                const long long gr = (long long)p.r_slot[r] * N + p.r_v[r];
                for (int e = 0; e < nelem; e++) {
                    const int is = iterm + e;
                    const double val =
                        b.repres[g][(size_t)is + (size_t)ideg * Ntot + (size_t)ib * sdeg * Ntot];
                    if (val != 0.0) rows[gr].push_back({is, val});
                }
            }
            m.rptr.assign(G + 1, 0);
            for (long long i = 0; i < G; i++) m.rptr[i + 1] = m.rptr[i] + (int)rows[i].size();
            m.rcol.reserve(m.rptr[G]);
            m.rval.reserve(m.rptr[G]);
            std::vector<std::vector<std::pair<int, double>>> cols(m.nsize);
            for (long long i = 0; i < G; i++)
                for (auto &e : rows[i]) {
                    m.rcol.push_back(e.first);
                    m.rval.push_back(e.second);
                    cols[e.first].push_back({(int)i, e.second});
                }
            m.tptr.assign(m.nsize + 1, 0);
            for (int t = 0; t < m.nsize; t++) m.tptr[t + 1] = m.tptr[t] + (int)cols[t].size();
            for (int t = 0; t < m.nsize; t++)
                for (auto &e : cols[t]) {
                    m.tcol.push_back(e.first);
                    m.tval.push_back(e.second);
                }
        }
    }
    return GI_OK;
}

static bool row_nonempty(const SymMap &m, long long row)
{
    return m.rptr[row + 1] > m.rptr[row];
}

// Symmetry-blocked plan of the half linestrength from (J, gI) to J'.
int build_group(gi_ctx *c, int jindI, int gI, int jindF)
{
    GroupPlan &G = group_of(c, jindI, gI, jindF);
    G = GroupPlan();
    const Coef &cf = coef_of(c, jindI, jindF);
    if (!cf.valid) return GI_OK;
    const int gF = c->igamma_pair[gI];
    const JPlan &pI = c->jp[jindI], &pF = c->jp[jindF];
    const int N = c->nvib, ncls = (int)c->vcls_off.size();
    const int SI = pI.nslots, SF = pF.nslots;
    G.valid = true;
    G.jindI = jindI; G.gI = gI; G.jindF = jindF; G.gF = gF;
    // supports per (class, slot), read off the first member of the class
    std::vector<char> inok((size_t)ncls * SI, 0), outok((size_t)ncls * SF, 0);
    for (int k = 0; k < ncls; k++) {
        const long long v = c->vcls_off[k];
        for (int s = 0; s < SI; s++)
            for (int d = 0; d < c->degen[gI]; d++)
                if (row_nonempty(pI.sym[gI * GI_MAXDEG + d], (long long)s * N + v))
                    inok[(size_t)k * SI + s] = 1;
        for (int s = 0; s < SF; s++)
            for (int d = 0; d < c->degen[gF]; d++)
                if (row_nonempty(pF.sym[gF * GI_MAXDEG + d], (long long)s * N + v))
                    outok[(size_t)k * SF + s] = 1;
    }
    std::vector<int> cls_oc(ncls, -1);
    for (int co = 0; co < ncls; co++) {
        OutClass oc;
        oc.off = c->vcls_off[co];
        oc.n = c->vcls_n[co];
        for (int s = 0; s < SF; s++)
            if (outok[(size_t)co * SF + s]) oc.cols.push_back(s);
        if (oc.cols.empty()) continue;
        const int ncol = (int)oc.cols.size();
        oc.smax = smax_for(ncol);
        if (oc.smax < 0) return fail(c, GI_EINVAL, "too many rotational slots in one class");
        for (int a = 0; a < 3; a++)
            for (int ci = 0; ci < ncls; ci++) {
                bool used = false;
                for (int j = 0; j < ncol && !used; j++) {
                    const int sF = oc.cols[j];
                    for (int p = cf.ptr[sF * 3 + a]; p < cf.ptr[sF * 3 + a + 1]; p++)
                        if (inok[(size_t)ci * SI + cf.sI[p]]) { used = true; break; }
                }
                if (!used) continue;
                oc.seg_a.push_back(a);
                oc.seg_off.push_back(c->vcls_off[ci]);
                oc.seg_n.push_back(c->vcls_n[ci]);
                oc.seg_row0.push_back(oc.K);
                oc.K += c->vcls_n[ci];
                for (int j = 0; j < ncol; j++) {
                    const int sF = oc.cols[j];
                    oc.rptr.push_back((int)oc.rsI.size());
                    for (int p = cf.ptr[sF * 3 + a]; p < cf.ptr[sF * 3 + a + 1]; p++)
                        if (inok[(size_t)ci * SI + cf.sI[p]]) {
                            oc.rsI.push_back(cf.sI[p]);
                            oc.rcoef.push_back(cf.val[p]);
                        }
                }
            }
        oc.seg_row0.push_back(oc.K);
        oc.rptr.push_back((int)oc.rsI.size());
        oc.hbase = G.Htot;
        G.Htot += oc.n * ncol;
        oc.gbase = G.Gtot;
        G.Gtot += oc.K * oc.smax;
        oc.gbase_p = G.Gtot_p;
        G.Gtot_p += oc.K * ncol;
        G.smax_all = std::max(G.smax_all, oc.smax);
        G.macs += (double)oc.n * oc.K * ncol;
        cls_oc[co] = (int)G.oc.size();
        G.oc.push_back(std::move(oc));
    }
    // class of each permuted vibrational index
    std::vector<int> vcls(N);
    for (int k = 0; k < ncls; k++)
        for (int i = 0; i < c->vcls_n[k]; i++) vcls[c->vcls_off[k] + i] = k;
    for (int d = 0; d < c->degen[gF]; d++) {
        const SymMap &m = pF.sym[gF * GI_MAXDEG + d];
        auto &tc = G.tcol[d];
        tc.resize(m.tcol.size());
        for (size_t e = 0; e < m.tcol.size(); e++) {
            const int cell = m.tcol[e];
            const int sF = cell / N, v = cell % N;
            const int o = cls_oc[vcls[v]];
            if (o < 0) return fail(c, GI_EINVAL, "internal: fold cell outside the plan");
            const OutClass &oc = G.oc[o];
            const int j = (int)(std::lower_bound(oc.cols.begin(), oc.cols.end(), sF) -
                                oc.cols.begin());
            if (j >= (int)oc.cols.size() || oc.cols[j] != sF)
                return fail(c, GI_EINVAL, "internal: fold column outside the plan");
            tc[e] = oc.hbase + j * oc.n + (v - oc.off);
        }
        // fixed pairs (-1: no entry), only for maps whose rows all have at most
        // two entries: with longer rows (C3v E blocks) the pair loads come on
        // top of the CSR walk (measured: C3v J 0..15 fold-split 37.6 -> 45.9 ms)
        auto &fc = G.tfc[d];
        auto &fw = G.tfw[d];
        fc.assign(2 * (size_t)m.nsize, -1);
        fw.assign(2 * (size_t)m.nsize, 0.0);
        for (int t = 0; t < m.nsize && !fc.empty(); t++) {
            const int p0 = m.tptr[t], n = m.tptr[t + 1] - p0;
            if (n > 2) {
                fc.clear();
                fw.clear();
                break;
            }
            for (int k = 0; k < n; k++) {
                fc[2 * t + k] = tc[p0 + k];
                fw[2 * t + k] = m.tval[p0 + k];
            }
        }
    }
    return GI_OK;
}

// Sparse map from a lower eigenvector record to the unit's packed G (rows of
// ncols columns, classes back to back at gbase_p): for every G element
// (class, segment row, column) the recipe entries (slot_I, w) and, for each,
// the symmetrisation-map row of cell (slot_I, v_I), composed and merged per
// record index.
int build_gmap(gi_ctx *c, GroupPlan &G)
{
    const int N = c->nvib;
    const JPlan &pI = c->jp[G.jindI];
    std::vector<std::vector<std::pair<int, double>>> ent;
    int W = 0;
    // split-mode layout: per class, ncols columns of kpad rows (segments padded to 64)
    G.GtotB = G.EGtot = 0;
    for (OutClass &oc : G.oc) {
        oc.seg_pad.clear();
        oc.kpad = 0;
        for (size_t e = 0; e < oc.seg_a.size(); e++) {
            int ci = 0;
            while (ci + 1 < (int)c->vcls_off.size() && c->vcls_off[ci + 1] <= oc.seg_off[e]) ci++;
            oc.seg_pad.push_back(oc.kpad);
            oc.seg_pad.push_back(c->vpad_off[ci] + (oc.seg_off[e] - c->vcls_off[ci]));
            oc.kpad += (oc.seg_n[e] + 63) / 64 * 64;
        }
        oc.cbaseB = G.GtotB;
        G.GtotB += (int)oc.cols.size() * oc.kpad;
        oc.eoff = G.EGtot;
        G.EGtot += (int)oc.cols.size();
    }
    for (int d = 0; d < c->degen[G.gI]; d++) {
        const SymMap &m = pI.sym[G.gI * GI_MAXDEG + d];
        ent.assign(G.Gtot_p, {});
        for (const OutClass &oc : G.oc) {
            const int ncol = (int)oc.cols.size();
            for (int e = 0; e < (int)oc.seg_a.size(); e++)
                for (int row = oc.seg_row0[e]; row < oc.seg_row0[e + 1]; row++) {
                    const int vI = oc.seg_off[e] + (row - oc.seg_row0[e]);
                    for (int j = 0; j < ncol; j++) {
                        auto &L = ent[(size_t)oc.gbase_p + (size_t)row * ncol + j];
                        const int q = e * ncol + j;
                        for (int p = oc.rptr[q]; p < oc.rptr[q + 1]; p++) {
                            const long long cell = (long long)oc.rsI[p] * N + vI;
                            for (int x = m.rptr[cell]; x < m.rptr[cell + 1]; x++) {
                                const int idx = m.rcol[x];
                                const double w = oc.rcoef[p] * m.rval[x];
                                auto it = std::find_if(L.begin(), L.end(),
                                                       [&](const std::pair<int, double> &t) {
                                                           return t.first == idx;
                                                       });
                                if (it == L.end()) L.push_back({idx, w});
                                else it->second += w;
                            }
                        }
                        std::sort(L.begin(), L.end());
                        W = std::max(W, (int)L.size());
                    }
                }
        }
        G.gidx[d].clear();
        G.gwt[d].clear();
        G.gidx[d].resize((size_t)std::max(W, 1) * G.Gtot_p, 0);
        G.gwt[d].resize((size_t)std::max(W, 1) * G.Gtot_p, 0.0);
        G.gW = std::max(G.gW, W);
        // split layout: element (row, j) of class co at cbaseB + j*kpad + seg_pad + (row - row0)
        G.gidxB[d].assign((size_t)std::max(W, 1) * G.GtotB, 0);
        G.gwtB[d].assign((size_t)std::max(W, 1) * G.GtotB, 0.0);
        G.gbnd[d].assign(std::max(G.EGtot, 1), 0.0);
        for (const OutClass &oc : G.oc) {
            const int ncol = (int)oc.cols.size();
            for (int e = 0; e < (int)oc.seg_a.size(); e++)
                for (int row = oc.seg_row0[e]; row < oc.seg_row0[e + 1]; row++)
                    for (int j = 0; j < ncol; j++) {
                        const auto &L = ent[(size_t)oc.gbase_p + (size_t)row * ncol + j];
                        const size_t gb = (size_t)oc.cbaseB + (size_t)j * oc.kpad +
                                          oc.seg_pad[2 * e] + (row - oc.seg_row0[e]);
                        double sa = 0.0;
                        for (size_t w = 0; w < L.size(); w++) {
                            G.gidxB[d][w * G.GtotB + gb] = L[w].first;
                            G.gwtB[d][w * G.GtotB + gb] = L[w].second;
                            sa += std::fabs(L[w].second);
                        }
                        double &b = G.gbnd[d][oc.eoff + j];
                        b = std::max(b, sa);
                    }
        }
        // stash this component's entries; widths are fixed below
        for (size_t g = 0; g < ent.size(); g++)
            for (size_t w = 0; w < ent[g].size(); w++) {
                G.gidx[d][w * G.Gtot_p + g] = ent[g][w].first;
                G.gwt[d][w * G.Gtot_p + g] = ent[g][w].second;
            }
    }
    // every component uses the final width gW (layout [w][g])
    for (int d = 0; d < c->degen[G.gI]; d++) {
        const size_t have = G.gidx[d].size() / std::max<size_t>(G.Gtot_p, 1);
        if ((int)have < G.gW) {
            G.gidx[d].resize((size_t)G.gW * G.Gtot_p, 0);
            G.gwt[d].resize((size_t)G.gW * G.Gtot_p, 0.0);
            G.gidxB[d].resize((size_t)G.gW * G.GtotB, 0);
            G.gwtB[d].resize((size_t)G.gW * G.GtotB, 0.0);
        }
    }
    build_compact_map(c, G);
    return GI_OK;
}

// Compact record map for the LDS-staged FP64 G build: the record index of
// each entry in 16 bits and its weight as an 8-bit code into a table of the
// component's distinct weights (code 0 = weight 0, the padding), packed
// layout. A G element then reads 3 bytes per entry instead of 12; its products
// are the same doubles. Only for maps of width <= 4 whose indices fit 16 bits
// and whose components have at most kCmapCodes distinct weights (bit
// patterns). (The split layout keeps the full map: its build loads four rows'
// entries as 16-byte vectors, and the table lookups cost it more than the
// bytes they save.)
void build_compact_map(gi_ctx *c, GroupPlan &G)
{
    G.cmap = false;
    for (int d = 0; d < GI_MAXDEG; d++) {
        G.cidx[d].clear();
        G.ccode[d].clear();
        G.cwt[d].clear();
    }
    if (G.gW > 4 || G.Gtot_p <= 0) return;
    const int nd = c->degen[G.gI];
    for (int d = 0; d < nd; d++) {
        std::map<uint64_t, int> code;
        std::vector<double> tab(1, 0.0);
        code[0] = 0;
        bool ok = true;
        auto enc = [&](const std::vector<int> &ix, const std::vector<double> &wt,
                       std::vector<uint16_t> &cix, std::vector<uint8_t> &ccd) {
            cix.resize(ix.size());
            ccd.resize(wt.size());
            for (size_t i = 0; i < ix.size() && ok; i++) {
                if (ix[i] < 0 || ix[i] > 65535) {
                    ok = false;
                    break;
                }
                cix[i] = (uint16_t)ix[i];
                uint64_t b;
                std::memcpy(&b, &wt[i], 8);
                auto it = code.find(b);
                if (it == code.end()) {
                    if ((int)tab.size() == kCmapCodes) {
                        ok = false;
                        break;
                    }
                    it = code.emplace(b, (int)tab.size()).first;
                    tab.push_back(wt[i]);
                }
                ccd[i] = (uint8_t)it->second;
            }
        };
        enc(G.gidx[d], G.gwt[d], G.cidx[d], G.ccode[d]);
        if (!ok) {
            for (int e = 0; e < GI_MAXDEG; e++) {
                std::vector<uint16_t>().swap(G.cidx[e]);
                std::vector<uint8_t>().swap(G.ccode[e]);
                std::vector<double>().swap(G.cwt[e]);
            }
            return;
        }
        tab.resize(kCmapCodes, 0.0);
        G.cwt[d] = tab;
    }
    G.cmap = true;
}

// Coefficients of the half linestrength jvals[jindI] -> jvals[jindF]: the
// factors of dipole_kernals.cu:194-223 that depend only on the rotational
// labels, gathered per (slot_F, component).
int build_coef(gi_ctx *c, int jindI, int jindF)
{
    Coef &cf = coef_of(c, jindI, jindF);
    cf = Coef();
    const JPlan &pI = c->jp[jindI], &pF = c->jp[jindF];
    const int jI = pI.J, jF = pF.J;
    if (std::abs(jF - jI) > 1) return GI_OK;
    cf.valid = true;
    cf.SI = pI.nslots;
    cf.SF = pF.nslots;
    const int s = c->jmax + 1;
    const double sq2 = 1.0 / std::sqrt(2.0);
    cf.ptr.assign((size_t)cf.SF * 3 + 1, 0);
    std::vector<std::vector<std::pair<int, double>>> lists((size_t)cf.SF * 3);
    for (int sF = 0; sF < cf.SF; sF++) {
        const int kF = pF.slot_k[sF], tauF = pF.slot_tau[sF];
        const int sigmaF = (kF % 3) * tauF;
        const int sgnF = 2 * (~(sigmaF) & 1) - 1;
        bool comp_used[3] = {false, false, false};
        for (int sI = 0; sI < cf.SI; sI++) {
            const int kI = pI.slot_k[sI], tauI = pI.slot_tau[sI];
            if (std::abs(kI - kF) > 1) continue;
            const int sigmaI = (kI % 3) * tauI;
            const int sgnI = 2 * (~(sigmaI + kI) & 1) - 1;
            const double f3j = c->threej[jI + kI * s + (jF - jI + 1) * s * s + (kF - kI + 1) * s * s * 3];
            const int keq = (kF == kI), tneq = (tauF != tauI), kzero = ((kI * kF) != 0);
            const int a = keq ? 2 : (tneq ? 0 : 1);
            const double L = (double)(tauF - tauI) * (double)keq +
                             (double)((tauF - tauI) * (kF - kI) * (!keq) * (tneq)) +
                             -1.0 * (double)((!keq) * (!tneq));
            const double sfac = 1.0 + (sq2 - 1.0) * (double)kzero * (double)(!keq);
            if (L != 0.0) cf.P += 1.0;
            const double v = (double)sgnF * (double)sgnI * L * f3j * sfac;
            if (v == 0.0) continue;
            comp_used[a] = true;
            lists[(size_t)sF * 3 + a].push_back({sI, v});
        }
        for (int a = 0; a < 3; a++) cf.C += comp_used[a] ? 1.0 : 0.0;
    }
    for (size_t i = 0; i < lists.size(); i++) {
        cf.ptr[i + 1] = cf.ptr[i] + (int)lists[i].size();
        for (auto &e : lists[i]) {
            cf.sI.push_back(e.first);
            cf.val.push_back(e.second);
        }
    }
    return GI_OK;
}

}  // namespace gi
