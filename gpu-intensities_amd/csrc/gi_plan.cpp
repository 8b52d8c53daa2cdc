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
    return J >= c->jvals[0] && J <= c->jvals[1] && (e - c->zpe) >= c->erange_lower[0] &&
           (e - c->zpe) <= c->erange_lower[1];
}

// trove_functions.cpp:1442-1451
bool energy_filter_upper(const gi_ctx *c, int J, double e, const int *q)
{
    for (int i = 0; i < c->nmodes; i++) {
        if (q[i + 1] < c->quanta_upper[i]) return false;
        if (q[i + 1] > c->quanta_upper[c->nmodes + i]) return false;
    }
    return J >= c->jvals[0] && J <= c->jvals[1] && (e - c->zpe) >= c->erange_upper[0] &&
           (e - c->zpe) <= c->erange_upper[1];
}

// Slot grid and symmetrisation maps of basis jind.
int build_jplan(gi_ctx *c, int jind)
{
    const auto &b = c->hb[jind];
    JPlan &p = c->jp[jind];
    p = JPlan();
    p.J = b.jval;
    p.dimen = b.dimen;
    const int N = c->nvib;
    // slots: (K, tau, copy); a repeated (v, K, tau) opens a new copy
    std::map<std::pair<int, int>, std::vector<int>> slots_of;  // (K,tau) -> slot ids
    std::map<long long, int> used;                             // (slot, v) taken
    p.r_slot.resize(b.dimen);
    p.r_v.resize(b.dimen);
    for (int r = 0; r < b.dimen; r++) {
        const int v = b.v[r], K = b.k[r], tau = b.ktau[r] & 1;
        if (v < 0 || v >= N)
            return fail(c, GI_EINVAL, "basis J=" + std::to_string(b.jval) +
                                          ": iroot_correlat_j0 out of the dipole's range");
        if (K < 0 || K > b.jval)
            return fail(c, GI_EINVAL, "basis J=" + std::to_string(b.jval) + ": k out of range");
        auto &ids = slots_of[{K, tau}];
        int chosen = -1;
        for (int s : ids)
            if (!used.count((long long)s * N + v)) { chosen = s; break; }
        if (chosen < 0) {
            chosen = p.nslots++;
            p.slot_k.push_back(K);
            p.slot_tau.push_back(tau);
            ids.push_back(chosen);
        }
        used[(long long)chosen * N + v] = r;
        p.r_slot[r] = chosen;
        p.r_v[r] = v;
    }
    // order slots by (K, tau, copy) so the grid is laid out by K
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
                if (irow < 0 || irow >= b.ncases || ib < 0 || ib >= b.mat_size)
                    return fail(c, GI_EINVAL, "icontr2icase out of range");
                const int iterm = b.ijterms[irow + g * b.ncases];
                const int nelem = b.N[g * b.ncases + irow];
                const long long gr = (long long)p.r_slot[r] * N + p.r_v[r];
                for (int e = 0; e < nelem; e++) {
                    const int is = iterm + e;
                    if (is < 0 || is >= Ntot || is >= m.nsize)
                        return fail(c, GI_EINVAL, "ijterms/N exceed Ntotal or nsize");
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

// Coefficients of the half linestrength jvals[jindI] -> jvals[jindF]: the
// factors of dipole_kernals.cu:194-223 that depend only on the rotational
// labels, gathered per (slot_F, component).
int build_coef(gi_ctx *c, int jindI, int jindF)
{
    Coef &cf = c->coef[jindI][jindF];
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
