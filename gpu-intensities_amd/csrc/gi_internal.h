// Internal declarations of the gfx950 line-intensity engine.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/gpu_intensities.h"

#define GI_MAXDEG 2

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

// Resident eigenvector block of one (J, symmetry) and its levels in energy order.
struct Block {
    int nsize = 0, nrec = 0;
    const double *d_vec = nullptr;
    bool owned = false;
    std::vector<int> lev;            // level ids, energy order
    std::vector<double> en;
    std::vector<unsigned char> upok; // energy_filter_upper per level
    int *d_lev = nullptr, *d_col = nullptr;
    double *d_en = nullptr;
    unsigned char *d_upok = nullptr;
};

struct KStat {
    double ms = 0;
    int launches = 0;
};

}  // namespace gi

struct gi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // parameters
    bool have_params = false;
    int nrep = 0, sym_type = 0, nmodes = 0, reduced = 0;
    std::vector<int> degen, isym_pairs, igamma_pair, quanta_lower, quanta_upper;
    std::vector<double> gns;
    double zpe = 0, temperature = 0, q_stat = 1;
    double freq_window[2] = {0, 0}, erange_lower[2] = {0, 0}, erange_upper[2] = {0, 0};
    int jvals[2] = {0, 0};
    double thresh_S = 0, thresh_I = 0;

    // inputs
    bool have_bset[2] = {false, false};
    struct HostBset {
        int jval, ncases, dimen, mat_size;
        std::vector<int> icontr2icase, v, k, ktau, N, Ntotal, ijterms, nsize;
        std::vector<std::vector<double>> repres;
    } hb[2];
    int nvib = 0;
    double *d_mu = nullptr;
    bool mu_owned = false;
    int nlevels = 0;
    std::vector<double> lv_energy;
    std::vector<int> lv_jind, lv_igamma, lv_ndeg, lv_irec0, lv_quanta;
    std::vector<unsigned char> lv_lowok, lv_upok;
    std::vector<gi::Block> blocks;   // [jind*nrep + g]

    // plan
    bool prepared = false;
    int jmax = 0;
    std::vector<double> threej;
    gi::JPlan jp[2];
    gi::Coef coef[2][2];             // [jindI][jindF]

    // scratch (grown on demand)
    std::map<std::string, std::pair<void *, size_t>> scratch;

    // statistics of the last gi_compute
    std::map<std::string, gi::KStat> kstats;
    std::vector<std::string> kstat_names;
    double work[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace gi {

int fail(gi_ctx *c, int code, const std::string &msg);
void *scratch(gi_ctx *c, const char *name, size_t bytes, int *rc);

// host plan construction (gi_plan.cpp)
int build_jplan(gi_ctx *c, int jind);
int build_coef(gi_ctx *c, int jindI, int jindF);
double three_j(int j1, int j2, int j3, int k1, int k2, int k3);
void threej_table(int jmax, double *out);
bool energy_filter_lower(const gi_ctx *c, int J, double e, const int *q);
bool energy_filter_upper(const gi_ctx *c, int J, double e, const int *q);

}  // namespace gi
