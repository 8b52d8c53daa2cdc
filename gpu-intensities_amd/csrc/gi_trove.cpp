// TROVE input files, the stdin deck and the line list (include/gi_trove.h).
//
// Readers restate the reference's (file:line cited per function) with the
// same token rules (strtok on " ,", strtol base 0, strtod) and the same
// checks, returning a status instead of exiting. Two deliberate
// differences, both in the level table: the energy sort is stable and covers
// every level (the reference's qsort leaves the last level out of the sort,
// trove_functions.cpp:1414, because `nlevels` holds the last index there),
// and a quanta[0] outside the rotational table is an error instead of an
// out-of-bounds read.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gi_trove.h"

namespace {

struct Fail {
    std::string msg;
};

[[noreturn]] void fail(const std::string &m) { throw Fail{m}; }

std::string trim(const std::string &s)
{
    const char *ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

// Token stream over one line with the reference's delimiters.
struct Tok {
    std::vector<std::string> t;
    size_t i = 0;
    Tok(const std::string &line, const char *delims)
    {
        size_t p = 0;
        while (p < line.size()) {
            while (p < line.size() && std::strchr(delims, line[p])) p++;
            size_t q = p;
            while (q < line.size() && !std::strchr(delims, line[q])) q++;
            if (q > p) t.push_back(line.substr(p, q - p));
            p = q;
        }
    }
    bool more() const { return i < t.size(); }
    const std::string &next(const char *what)
    {
        if (i >= t.size()) fail(std::string("deck: read error after ") + what);
        return t[i++];
    }
    double d(const char *what) { return std::strtod(next(what).c_str(), nullptr); }
    int n(const char *what) { return (int)std::strtol(next(what).c_str(), nullptr, 0); }
};

// Whitespace-separated numbers of one line, strtol/strtod style (a missing
// or malformed field reads as 0, as with the reference's strtol chains).
struct Nums {
    const char *p;
    explicit Nums(const std::string &s) : p(s.c_str()) {}
    long i()
    {
        char *e;
        const long v = std::strtol(p, &e, 0);
        p = e;
        return v;
    }
    double d()
    {
        char *e;
        const double v = std::strtod(p, &e);
        p = e;
        return v;
    }
};

struct Deck {
    int nmodes = -1, sym_type = 0, nrep = 0;
    std::vector<int> degen;
    std::vector<std::string> labels;
    double temperature = -1, q_stat = -1, zpe = 0, thresh_line = 0;
    double freq_window[2] = {-1e300, 1e300};
    double erange_lower[2] = {-1e300, 1e300}, erange_upper[2] = {-1e300, 1e300};
    double erange[2] = {0, 0};
    int jvals[2] = {-1, -1};
    std::vector<double> gns;
    std::vector<int> isym_pairs, isym_do;
    std::vector<int> quanta_lower, quanta_upper;  // [lo x nmodes, hi x nmodes]
    bool reduced = false, have_gns = false, have_sel = false;
};

// read_symmetry (fields.cpp:11-110): irreps, degeneracies and labels.
void set_symmetry(Deck &d, const std::string &sym)
{
    if (sym == "C2V(M)") {
        d.sym_type = 1;
        d.degen = {1, 1, 1, 1};
        d.labels = {"A1", "A2", "B1", "B2"};
    } else if (sym == "C3V(M)") {
        d.sym_type = 2;
        d.degen = {1, 1, 2};
        d.labels = {"A1", "A2", "E"};
    } else if (sym == "D3H(M)") {
        d.sym_type = 3;
        d.degen = {1, 1, 2, 1, 1, 2};
        d.labels = {"A1'", "A2'", "E'", "A1\"", "A2\"", "E\""};
    } else {
        fail("deck: symmetry not implemented: " + sym);
    }
    d.nrep = (int)d.degen.size();
    d.gns.assign(d.nrep, 0.0);
    d.isym_pairs.assign(d.nrep, 0);
    d.isym_do.assign(d.nrep, 1);
}

std::string upper(std::string s)
{
    for (char &ch : s) ch = (char)std::toupper((unsigned char)ch);
    return s;
}

// read_intensities (fields.cpp:145-234): the INTENSITY ... END block.
void parse_block(Deck &d, std::istream &in)
{
    std::string line;
    while (std::getline(in, line)) {
        const size_t o = line.find('(');
        if (o != std::string::npos) {
            const size_t c = line.find(')');
            if (c != std::string::npos && c > o) line.erase(o, c - o + 1);
            else line.erase(o);
        }
        Tok t(upper(line), " ,\r");
        if (!t.more()) continue;
        const std::string key = t.next("");
        if (key == "END") return;
        if (key == "TEMPERATURE") d.temperature = t.d("TEMPERATURE");
        else if (key == "QSTAT") d.q_stat = t.d("QSTAT");
        else if (key == "ZPE") d.zpe = t.d("ZPE");
        else if (key == "J") {
            d.jvals[0] = t.n("J");
            d.jvals[1] = t.n("J");
        } else if (key == "FREQ-WINDOW") {
            d.freq_window[0] = t.d("FREQ-WINDOW");
            d.freq_window[1] = t.d("FREQ-WINDOW");
        } else if (key == "ENERGY") {
            while (t.more()) {
                const std::string w = t.next("ENERGY");
                if (w == "LOW") {
                    d.erange_lower[0] = t.d("ENERGY LOW");
                    d.erange_lower[1] = t.d("ENERGY LOW");
                } else if (w == "UPPER") {
                    d.erange_upper[0] = t.d("ENERGY UPPER");
                    d.erange_upper[1] = t.d("ENERGY UPPER");
                }
            }
        } else if (key == "GNS") {
            if (!d.nrep) fail("deck: GNS before SYMGROUP");
            for (int i = 0; i < d.nrep; i++) {
                d.gns[i] = t.d("GNS");
                if (d.gns[i] <= 0.00000001) d.isym_do[i] = 0;  // fields.cpp:193-195
            }
            d.have_gns = true;
        } else if (key == "SELECTION") {
            if (!d.nrep) fail("deck: SELECTION before SYMGROUP");
            for (int i = 0; i < d.nrep; i++) d.isym_pairs[i] = t.n("SELECTION");
            d.have_sel = true;
        } else if (key == "THRESH_LINE") {
            d.thresh_line = t.d("THRESH_LINE");
        } else if (key == "V") {
            if (d.nmodes < 0) fail("deck: V before NMODES");
            const int q = t.n("V") - 1;
            if (q < 0 || q >= d.nmodes) fail("deck: V index outside 1..NMODES");
            while (t.more()) {
                const std::string w = t.next("V");
                if (w == "LOWER") {
                    d.quanta_lower[q] = t.n("V LOWER");
                    d.quanta_lower[d.nmodes + q] = t.n("V LOWER");
                } else if (w == "UPPER") {
                    d.quanta_upper[q] = t.n("V UPPER");
                    d.quanta_upper[d.nmodes + q] = t.n("V UPPER");
                }
            }
        } else if (key == "SYMMETRY") {
            if (t.more() && t.next("SYMMETRY") == "REDUCED") d.reduced = true;
        }
    }
}

// read_fields (fields.cpp:236-282).
Deck parse_deck(const std::string &text)
{
    Deck d;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const std::string u = upper(line);
        Tok t(u, " ,\r");
        if (!t.more()) continue;
        const std::string key = t.next("");
        if (key == "NMODES") {
            if (!t.more()) continue;
            d.nmodes = (int)std::strtol(t.next("").c_str(), nullptr, 0);
            if (d.nmodes < 0 || d.nmodes > 64) fail("deck: NMODES out of range");
            d.quanta_lower.assign(2 * d.nmodes, 0);
            d.quanta_upper.assign(2 * d.nmodes, 0);
            for (int i = 0; i < d.nmodes; i++) {
                d.quanta_lower[i] = d.quanta_upper[i] = INT_MIN;
                d.quanta_lower[d.nmodes + i] = d.quanta_upper[d.nmodes + i] = INT_MAX;
            }
        } else if (key == "SYMGROUP") {
            Tok s(u, " \r");  // the symmetry token is split on spaces only
            s.next("");
            set_symmetry(d, s.next("SYMGROUP"));
        } else if (key == "INTENSITY") {
            parse_block(d, in);
        }
    }
    if (d.nmodes < 0) fail("deck: NMODES missing");
    if (!d.nrep) fail("deck: SYMGROUP missing");
    if (d.jvals[0] < 0 || d.jvals[1] < 0) fail("deck: J missing");
    if (!d.have_gns || !d.have_sel) fail("deck: GNS and SELECTION are required");
    if (!(d.temperature > 0)) fail("deck: TEMPERATURE missing or not positive");
    if (!(d.q_stat > 0)) fail("deck: QSTAT missing or not positive");
    d.erange[0] = std::min(d.erange_lower[0], d.erange_upper[0]);
    d.erange[1] = std::max(d.erange_lower[1], d.erange_upper[1]);
    return d;
}

// One J basis as bset_contr_factory (trove_functions.cpp:401-679) reads it,
// plus correlate_index (:680-858) and compute_ijterms (:927-966).
struct Basis {
    int jval = 0, ncases = 0, nlmax = 0, ncontr = 0, nclasses = 0;
    std::vector<int> cspace;              // [(nclasses+1)*ncases], 0-based
    std::vector<std::vector<int>> icoef;  // per case [(nclasses+1)*nlambdas], 0-based
    std::vector<int> icontr2icase;        // [2*ncontr]
    std::vector<int> icase2icontr;        // [ncases*nlmax]
    int rot_ncases = 0, rot_nl = 0;
    std::vector<int> rot_k, rot_tau;      // [rot_ncases*rot_nl]
    std::vector<char> rot_set;
    int mat_size = 0;
    std::vector<int> Ntotal, N;           // N[g*ncases + c]
    std::vector<std::vector<double>> repres;
    std::vector<int> iroot_j0, k, ktau, ijterms, nsize;
};

bool getline_t(std::istream &in, std::string &line)
{
    if (!std::getline(in, line)) return false;
    return true;
}

Basis read_quanta(const std::string &path, int jval, const std::vector<int> &degen)
{
    std::ifstream f(path);
    if (!f) fail("cannot open " + path);
    Basis b;
    b.jval = jval;
    std::string line;
    if (!getline_t(f, line) || trim(line) != "Start Primitive basis set")
        fail(path + ": bad header");
    if (!getline_t(f, line)) fail(path + ": truncated");
    {
        Nums n(line);
        b.ncases = (int)n.i();
        b.nlmax = (int)n.i();
        b.ncontr = (int)n.i();
        b.nclasses = (int)n.i();
    }
    if (b.ncases < 0 || b.nlmax < 0 || b.ncontr < 0 || b.nclasses < 0 || b.nclasses > 64)
        fail(path + ": bad sizes");
    const int nc1 = b.nclasses + 1;
    b.cspace.assign((size_t)nc1 * b.ncases, 0);
    b.icoef.resize(b.ncases);
    b.icontr2icase.assign(2 * (size_t)b.ncontr, 0);
    b.icase2icontr.assign((size_t)b.ncases * b.nlmax, -1);
    int icontr = 0;
    for (int ic = 0; ic < b.ncases; ic++) {
        if (!getline_t(f, line)) fail(path + ": truncated case list");
        Nums n(line);
        const int nl = (int)n.i();
        if (nl < 0 || nl > b.nlmax) fail(path + ": case with more lambdas than declared");
        for (int i = 0; i < nc1; i++) b.cspace[i + (size_t)ic * nc1] = (int)n.i() - 1;
        b.icoef[ic].assign((size_t)nc1 * nl, 0);
        for (int il = 0; il < nl; il++) {
            if (!getline_t(f, line)) fail(path + ": truncated lambda list");
            Nums m(line);
            const int iroot = (int)m.i() - 1;
            for (int i = 0; i < nc1; i++) b.icoef[ic][i + (size_t)il * nc1] = (int)m.i() - 1;
            if (icontr >= b.ncontr) fail(path + ": more functions than declared");
            b.icontr2icase[icontr] = ic;
            b.icontr2icase[icontr + b.ncontr] = il;
            b.icase2icontr[ic + (size_t)il * b.ncases] = icontr;
            if (iroot != icontr) fail(path + ": wrong indexing of contracted functions");
            icontr++;
        }
    }
    if (icontr != b.ncontr) fail(path + ": wrong indexing (function count)");
    if (!getline_t(f, line)) fail(path + ": truncated");
    {
        Nums n(line);
        b.rot_ncases = (int)n.i();
        b.rot_nl = (int)n.i();
    }
    if (b.rot_ncases < 0 || b.rot_nl < 0) fail(path + ": bad rotational index sizes");
    const size_t nrot = (size_t)b.rot_ncases * b.rot_nl;
    b.rot_k.assign(nrot, 0);
    b.rot_tau.assign(nrot, 0);
    b.rot_set.assign(nrot, 0);
    bool ended = false;
    while (getline_t(f, line)) {
        if (trim(line) == "End Primitive basis set") {
            ended = true;
            break;
        }
        Nums n(line);
        const int ic = (int)n.i() - 1, il = (int)n.i() - 1;
        (void)n.i();  // j
        const int kk = (int)n.i(), tau = (int)n.i();
        if (ic < 0 || ic >= b.rot_ncases || il < 0 || il >= b.rot_nl)
            fail(path + ": rotational index outside its table");
        const size_t x = ic + (size_t)il * b.rot_ncases;
        b.rot_k[x] = kk;
        b.rot_tau[x] = tau;
        b.rot_set[x] = 1;
    }
    if (!ended) fail(path + ": missing End Primitive basis set");
    if (!getline_t(f, line) || trim(line) != "Start irreducible transformation")
        fail(path + ": wrong sym-footer");
    if (!getline_t(f, line)) fail(path + ": truncated");
    b.mat_size = (int)Nums(line).i();
    if (b.mat_size < 1) fail(path + ": bad mat_size");
    const int nrep = (int)degen.size();
    if (!getline_t(f, line)) fail(path + ": truncated");
    {
        Nums n(line);
        b.Ntotal.assign(nrep, 0);
        for (int g = 0; g < nrep; g++) b.Ntotal[g] = (int)n.i();
    }
    b.N.assign((size_t)nrep * b.ncases, 0);
    b.repres.resize(nrep);
    for (int g = 0; g < nrep; g++) {
        if (b.Ntotal[g] < 0) fail(path + ": negative Ntotal");
        for (int c = 0; c < b.ncases; c++) {
            if (!getline_t(f, line)) fail(path + ": truncated N list");
            b.N[(size_t)g * b.ncases + c] = (int)Nums(line).i();
        }
        const size_t nt = b.Ntotal[g], dg = degen[g];
        b.repres[g].assign(nt * dg * b.mat_size, 0.0);
        for (size_t ico = 0; ico < nt; ico++)
            for (size_t id = 0; id < dg; id++) {
                if (!getline_t(f, line)) fail(path + ": truncated repres");
                Nums n(line);
                for (int i = 0; i < b.mat_size; i++)
                    b.repres[g][ico + id * nt + (size_t)i * nt * dg] = n.d();
            }
    }
    if (!getline_t(f, line) || trim(line) != "End irreducible transformation")
        fail(path + ": wrong irrep-footer");
    return b;
}

// correlate_index (trove_functions.cpp:680-858): the J=0 contracted index of
// every function of b (matching vibrational class indices and degenerate
// components), and (k, 2k+tau) from the rotational table.
void correlate(const Basis &j0, Basis &b)
{
    if (j0.jval != 0) fail("correlate: first basis is not J=0");
    if (j0.nclasses != b.nclasses) fail("correlate: Nclasses differ between J values");
    const int nc = b.nclasses, nc1 = nc + 1;
    std::vector<int> corr((size_t)b.ncases * b.nlmax, -1);
    for (int ic = 0; ic < b.ncases; ic++) {
        const int *cnu_i = &b.cspace[(size_t)ic * nc1 + 1];
        const int nl = (int)(b.icoef[ic].size() / nc1);
        for (int il = 0; il < nl; il++) {
            const int *co_i = &b.icoef[ic][(size_t)il * nc1 + 1];
            int found = -1;
            for (int jc = 0; jc < j0.ncases && found < 0; jc++) {
                const int *cnu_j = &j0.cspace[(size_t)jc * nc1 + 1];
                if (!std::equal(cnu_i, cnu_i + nc, cnu_j)) continue;
                const int nlj = (int)(j0.icoef[jc].size() / nc1);
                for (int jl = 0; jl < nlj; jl++)
                    if (std::equal(co_i, co_i + nc, &j0.icoef[jc][(size_t)jl * nc1 + 1])) {
                        found = j0.icase2icontr[jc + (size_t)jl * j0.ncases];
                        break;
                    }
            }
            if (found < 0)
                fail("correlate: no J=0 function for J=" + std::to_string(b.jval) + " case " +
                     std::to_string(ic));
            corr[ic + (size_t)il * b.ncases] = found;
        }
    }
    b.iroot_j0.assign(b.ncontr, 0);
    b.k.assign(b.ncontr, 0);
    b.ktau.assign(b.ncontr, 0);
    for (int r = 0; r < b.ncontr; r++) {
        const int ic = b.icontr2icase[r], il = b.icontr2icase[r + b.ncontr];
        const int jc = corr[ic + (size_t)il * b.ncases];
        if (jc >= j0.ncontr) fail("correlate: J=0 index out of range");
        b.iroot_j0[r] = jc;
        const int ilev = b.cspace[(size_t)ic * nc1];
        const int ideg = b.icoef[ic][(size_t)il * nc1];
        if (ilev < 0 || ilev >= b.rot_ncases || ideg < 0 || ideg >= b.rot_nl ||
            !b.rot_set[ilev + (size_t)ideg * b.rot_ncases])
            fail("correlate: rotational index not in the table");
        const int kk = b.rot_k[ilev + (size_t)ideg * b.rot_ncases];
        const int tau = b.rot_tau[ilev + (size_t)ideg * b.rot_ncases];
        b.k[r] = kk;
        b.ktau[r] = 2 * kk + tau;
    }
}

// compute_ijterms (trove_functions.cpp:927-966).
void ijterms(Basis &b, int nrep)
{
    b.ijterms.assign((size_t)b.ncases * nrep, 0);
    for (int g = 0; g < nrep; g++) {
        int s = 0;
        for (int c = 0; c < b.ncases; c++) {
            b.ijterms[c + (size_t)g * b.ncases] = s;
            s += b.N[(size_t)g * b.ncases + c];
        }
    }
}

// ReadFortranRecord (Util.cpp:67-75): 4-byte length, payload, 4-byte length.
std::vector<char> fortran_record(FILE *f, const std::string &path)
{
    unsigned int n = 0, m = 0;
    if (std::fread(&n, 4, 1, f) != 1) fail(path + ": truncated record");
    std::vector<char> buf(n);
    if (n && std::fread(buf.data(), 1, n, f) != n) fail(path + ": truncated record");
    if (std::fread(&m, 4, 1, f) != 1 || m != n) fail(path + ": bad record marker");
    return buf;
}

// read_dipole_flipped (trove_functions.cpp:113-205): three N x N blocks,
// stored flipped as mu_a(vF, vI) at vF + vI*N + a*N^2.
std::vector<double> read_dipole(const std::string &path, int ncontr_j0)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fail("cannot open " + path);
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } cl{f};
    auto h = fortran_record(f, path);
    if (h.size() < 20 || std::memcmp(h.data(), "Start external field", 20) != 0)
        fail(path + ": bogus header");
    auto nr = fortran_record(f, path);
    if (nr.size() != 4) fail(path + ": bad ncontr record");
    int ncontr;
    std::memcpy(&ncontr, nr.data(), 4);
    if (ncontr != ncontr_j0)
        fail(path + ": basis sizes at J=0 do not agree (" + std::to_string(ncontr) + " vs " +
             std::to_string(ncontr_j0) + ")");
    const size_t N = ncontr, N2 = N * N;
    std::vector<double> out(3 * N2);
    for (int a = 0; a < 3; a++) {
        auto ir = fortran_record(f, path);
        int imu;
        if (ir.size() != 4) fail(path + ": bad component record");
        std::memcpy(&imu, ir.data(), 4);
        if (imu != a + 1) fail(path + ": bogus component index");
        auto blk = fortran_record(f, path);
        if (blk.size() != N2 * 8) fail(path + ": bad matrix record size");
        const double *m = (const double *)blk.data();
        for (size_t i = 0; i < N; i++)
            for (size_t fF = 0; fF < N; fF++) out[fF + i * N + a * N2] = m[i + fF * N];
    }
    auto e = fortran_record(f, path);
    if (e.size() < 18 || std::memcmp(e.data(), "End external field", 18) != 0)
        fail(path + ": bogus footer");
    return out;
}

struct Level {
    double energy;
    int jind, jval, igamma, ndeg, ilevel, irec[2];
    std::vector<int> quanta, cgamma, normal;
    int krot, taurot;
    long seq;  // read order (tie-break of the stable sort)
};

}  // namespace

struct gi_trove {
    std::string err, dir;
    Deck deck;
    Basis b0, bj[2];
    std::vector<double> mu;
    std::vector<Level> lv;
    // flat level columns for gi_set_levels and the views
    std::vector<double> energy;
    std::vector<int> ljind, lgamma, lndeg, lirec, lilevel, lquanta, lkrot, ltaurot;
    std::vector<double> vec[2][8];
    int nrec[2][8] = {};
    // views
    std::vector<const double *> rep_ptr[2];
    // gi_trove_load applies the deck's THRESH_LINE as a linestrength cut only
    // when asked (gi_trove_apply_thresh_line): the reference parses it
    // (fields.cpp:204-206) but its cut is commented out (cuda_host.cu:1408)
    bool apply_thresh_line = false;
};

namespace {

char fname_buf[512];

std::string path_of(const std::string &dir, const char *fmt, int a, int b = 0)
{
    std::snprintf(fname_buf, sizeof fname_buf, fmt, a, b);
    return dir.empty() ? std::string(fname_buf) : dir + "/" + fname_buf;
}

// read_eigenvalues (trove_functions.cpp:1074-1429): levels of both J values
// that pass filter() (:1025-1028), with nsize per block from the descr
// header, energy sorted.
void read_levels(gi_trove &t)
{
    Deck &d = t.deck;
    const int nrep = d.nrep, nm = d.nmodes, ncl = t.b0.nclasses;
    long seq = 0;
    for (int jind = 0; jind < 2; jind++) {
        const int J = d.jvals[jind];
        t.bj[jind].nsize.assign(nrep, 0);
        for (int g = 0; g < nrep; g++) t.bj[jind].nsize[g] = t.bj[jind].Ntotal[g];
        for (int g = 0; g < nrep; g++) {
            if (!d.isym_do[g]) continue;
            const std::string path = path_of(t.dir, "j0eigen_descr%i_%i.chk", J, g + 1);
            std::ifstream f(path);
            if (!f) fail("cannot open " + path);
            std::string line;
            bool start = false;
            while (std::getline(f, line))
                if (trim(line) == "Start Quantum numbers and energies") {
                    start = true;
                    break;
                }
            if (!start) fail(path + ": malformed descr file");
            for (int i = 0; i < 3; i++)
                if (!std::getline(f, line)) fail(path + ": truncated header");
            Nums hn(line);
            (void)hn.i();  // nroots
            const int dim_basis = (int)hn.i();
            t.bj[jind].nsize[g] = std::max(dim_basis, 1);
            std::vector<int> lev_of;  // file ilevel -> index in t.lv
            while (std::getline(f, line)) {
                if (trim(line) == "End Quantum numbers and energies") break;
                Nums n(line);
                const int irec = (int)n.i();
                const int igam = (int)n.i();
                if (igam != g + 1) fail(path + ": symmetry index does not match the file");
                const int ilevel = (int)n.i() - 1;
                const int ideg = (int)n.i();
                const double e = n.d();
                if (d.zpe < 0 && igam == 1 && J == 0) d.zpe = e;  // :1201
                const bool pass = d.isym_do[g] && (e - d.zpe) >= d.erange[0] &&
                                  (e - d.zpe) <= d.erange[1];
                if (!pass) continue;
                if (ilevel < 0) fail(path + ": bad level index");
                if ((int)lev_of.size() <= ilevel) lev_of.resize(ilevel + 1, -1);
                if (ideg == 1) {
                    Level L;
                    L.energy = e;
                    L.jind = jind;
                    L.jval = J;
                    L.igamma = g;
                    L.ndeg = d.degen[g];
                    L.ilevel = ilevel;
                    L.irec[0] = irec;
                    L.irec[1] = irec;
                    L.quanta.resize(nm + 1);
                    for (int q = 0; q <= nm; q++) L.quanta[q] = (int)n.i();
                    (void)n.i();  // ilarge_coef
                    L.cgamma.resize(ncl + 1);
                    for (int q = 0; q <= ncl; q++) {
                        const int s = (int)n.i() - 1;
                        if (s < 0 || s >= nrep) fail(path + ": symmetry label out of range");
                        L.cgamma[q] = s;
                    }
                    L.normal.resize(nm + 1);
                    for (int q = 0; q <= nm; q++) L.normal[q] = (int)n.i();
                    const int q0 = L.quanta[0];  // ktau_rot table (:1286-1302)
                    if (q0 < 0 || q0 > 2 * J) fail(path + ": rotational index out of range");
                    L.krot = q0 == 0 ? 0 : (q0 + 1) / 2;
                    L.taurot = q0 == 0 ? J % 2 : (q0 + 1) % 2;
                    L.seq = seq++;
                    lev_of[ilevel] = (int)t.lv.size();
                    t.lv.push_back(std::move(L));
                } else if (ideg == 2) {
                    if (lev_of[ilevel] < 0) fail(path + ": degenerate component before its level");
                    t.lv[lev_of[ilevel]].irec[ideg - 1] = irec;
                } else {
                    fail(path + ": degenerate component index out of range");
                }
            }
        }
    }
    if (t.lv.empty()) fail("no levels pass the energy and symmetry filters");
    std::stable_sort(t.lv.begin(), t.lv.end(),
                     [](const Level &a, const Level &b) { return a.energy < b.energy; });
    const size_t n = t.lv.size();
    t.energy.resize(n);
    t.ljind.resize(n);
    t.lgamma.resize(n);
    t.lndeg.resize(n);
    t.lirec.resize(n);
    t.lilevel.resize(n);
    t.lkrot.resize(n);
    t.ltaurot.resize(n);
    t.lquanta.resize(n * (nm + 1));
    for (size_t i = 0; i < n; i++) {
        const Level &L = t.lv[i];
        t.energy[i] = L.energy;
        t.ljind[i] = L.jind;
        t.lgamma[i] = L.igamma;
        t.lndeg[i] = L.ndeg;
        t.lirec[i] = L.irec[0] - 1;
        t.lilevel[i] = L.ilevel;
        t.lkrot[i] = L.krot;
        t.ltaurot[i] = L.taurot;
        for (int q = 0; q <= nm; q++) t.lquanta[i * (nm + 1) + q] = L.quanta[q];
    }
}

// Eigenvector blocks: raw doubles, record irec (1-based) at (irec-1)*nsize
// (cuda_host.cu:1196).
void read_vectors(gi_trove &t)
{
    const Deck &d = t.deck;
    for (int jind = 0; jind < 2; jind++)
        for (int g = 0; g < d.nrep; g++) {
            if (!d.isym_do[g]) continue;
            const int nsize = t.bj[jind].nsize[g];
            const std::string path =
                path_of(t.dir, "j0eigen_vectors%i_%i.chk", d.jvals[jind], g + 1);
            FILE *f = std::fopen(path.c_str(), "rb");
            if (!f) fail("cannot open " + path);
            std::fseek(f, 0, SEEK_END);
            const long sz = std::ftell(f);
            std::fseek(f, 0, SEEK_SET);
            const long rec = (long)nsize * 8;
            const long nr = rec > 0 ? sz / rec : 0;
            t.vec[jind][g].resize((size_t)nr * nsize);
            const size_t got = nr ? std::fread(t.vec[jind][g].data(), rec, nr, f) : 0;
            std::fclose(f);
            if ((long)got != nr) fail(path + ": short read");
            t.nrec[jind][g] = (int)nr;
        }
    for (size_t i = 0; i < t.lv.size(); i++) {
        const Level &L = t.lv[i];
        if (L.irec[0] < 1 || L.irec[0] > t.nrec[L.jind][L.igamma])
            fail("level record " + std::to_string(L.irec[0]) + " outside j0eigen_vectors" +
                 std::to_string(L.jval) + "_" + std::to_string(L.igamma + 1) + ".chk");
    }
}

const char *branch(int jF, int jI)
{
    switch (jF - jI) {  // trove_functions.cpp:1466-1474
    case 0: return "Q";
    case 1: return "R";
    case -1: return "P";
    }
    return " ";
}

void appendf(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string &s, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n > 0) s.append(buf, std::min<size_t>(n, sizeof buf - 1));
}

int put(const std::string &s, char *buf, long long cap, long long *len)
{
    if (len) *len = (long long)s.size();
    if (!buf || cap < (long long)s.size() + 1) return GI_ECAPACITY;
    std::memcpy(buf, s.data(), s.size());
    buf[s.size()] = 0;
    return GI_OK;
}

}  // namespace

extern "C" {

int gi_trove_open(const char *dir, const char *deck, gi_trove **out)
{
    if (!out) return GI_EINVAL;
    gi_trove *t = new (std::nothrow) gi_trove();
    *out = t;
    if (!t) return GI_ENOMEM;
    if (!deck) {
        t->err = "deck text missing";
        return GI_EINVAL;
    }
    try {
        t->dir = dir ? dir : "";
        t->deck = parse_deck(deck);
        Deck &d = t->deck;
        t->b0 = read_quanta(path_of(t->dir, "j0eigen_quanta%i.chk", 0), 0, d.degen);
        for (int j = 0; j < 2; j++) {
            t->bj[j] = read_quanta(path_of(t->dir, "j0eigen_quanta%i.chk", d.jvals[j]), d.jvals[j],
                                   d.degen);
            correlate(t->b0, t->bj[j]);
            ijterms(t->bj[j], d.nrep);
        }
        t->mu = read_dipole(t->dir.empty() ? "j0_extfield.chk" : t->dir + "/j0_extfield.chk",
                            t->b0.ncontr);
        read_levels(*t);
        read_vectors(*t);
        for (int j = 0; j < 2; j++) {
            t->rep_ptr[j].clear();
            for (int g = 0; g < d.nrep; g++) t->rep_ptr[j].push_back(t->bj[j].repres[g].data());
        }
    } catch (const Fail &e) {
        t->err = e.msg;
        return GI_EINVAL;
    } catch (const std::bad_alloc &) {
        t->err = "out of host memory";
        return GI_ENOMEM;
    }
    return GI_OK;
}

const char *gi_trove_error(const gi_trove *t) { return t ? t->err.c_str() : "null handle"; }

void gi_trove_close(gi_trove *t) { delete t; }

int gi_trove_params(const gi_trove *t, gi_params *p)
{
    if (!t || !p) return GI_EINVAL;
    const Deck &d = t->deck;
    std::memset(p, 0, sizeof *p);
    p->nrepres = d.nrep;
    p->degen = d.degen.data();
    p->sym_type = d.sym_type;
    p->nmodes = d.nmodes;
    p->zpe = d.zpe;
    p->temperature = d.temperature;
    p->q_stat = d.q_stat;
    for (int i = 0; i < 2; i++) {
        p->freq_window[i] = d.freq_window[i];
        p->erange_lower[i] = d.erange_lower[i];
        p->erange_upper[i] = d.erange_upper[i];
        p->jvals[i] = d.jvals[i];
    }
    p->gns = d.gns.data();
    p->isym_pairs = d.isym_pairs.data();
    p->reduced = d.reduced ? 1 : 0;
    p->quanta_lower = d.quanta_lower.data();
    p->quanta_upper = d.quanta_upper.data();
    p->thresh_linestrength = d.thresh_line;  // as read; gi_trove_load applies it only on request
    p->thresh_intensity = 0.0;
    return GI_OK;
}

int gi_trove_bset(const gi_trove *t, int jind, gi_bset *b)
{
    if (!t || !b || jind < 0 || jind > 1) return GI_EINVAL;
    const Basis &x = t->bj[jind];
    b->jval = x.jval;
    b->Maxsymcoeffs = x.ncases;
    b->Maxcontracts = x.ncontr;
    b->mat_size = x.mat_size;
    b->icontr2icase = x.icontr2icase.data();
    b->iroot_correlat_j0 = x.iroot_j0.data();
    b->k = x.k.data();
    b->ktau = x.ktau.data();
    b->N = x.N.data();
    b->Ntotal = x.Ntotal.data();
    b->ijterms = x.ijterms.data();
    b->repres = t->rep_ptr[jind].data();
    b->nsize = x.nsize.data();
    return GI_OK;
}

int gi_trove_dipole(const gi_trove *t, int *nvib, const double **flipped)
{
    if (!t || !nvib || !flipped) return GI_EINVAL;
    *nvib = t->b0.ncontr;
    *flipped = t->mu.data();
    return GI_OK;
}

int gi_trove_levels(const gi_trove *t, int *n, const double **energy, const int **jind,
                    const int **igamma, const int **ndeg, const int **irec0, const int **ilevel,
                    const int **quanta, const int **krot, const int **taurot)
{
    if (!t || !n) return GI_EINVAL;
    *n = (int)t->lv.size();
    if (energy) *energy = t->energy.data();
    if (jind) *jind = t->ljind.data();
    if (igamma) *igamma = t->lgamma.data();
    if (ndeg) *ndeg = t->lndeg.data();
    if (irec0) *irec0 = t->lirec.data();
    if (ilevel) *ilevel = t->lilevel.data();
    if (quanta) *quanta = t->lquanta.data();
    if (krot) *krot = t->lkrot.data();
    if (taurot) *taurot = t->ltaurot.data();
    return GI_OK;
}

int gi_trove_vectors(const gi_trove *t, int jind, int igamma, int *nsize, int *nrec,
                     const double **v)
{
    if (!t || jind < 0 || jind > 1 || igamma < 0 || igamma >= t->deck.nrep || igamma >= 8)
        return GI_EINVAL;
    if (nsize) *nsize = t->bj[jind].nsize[igamma];
    if (nrec) *nrec = t->nrec[jind][igamma];
    if (v) *v = t->vec[jind][igamma].data();
    return GI_OK;
}

int gi_trove_load(gi_trove *t, gi_ctx *ctx)
{
    if (!t || !ctx) return GI_EINVAL;
    if (!t->err.empty()) return GI_ESTATE;
    gi_params p;
    gi_trove_params(t, &p);
    if (!t->apply_thresh_line) p.thresh_linestrength = 0.0;  // the reference prints every line
    int rc;
    if ((rc = gi_set_params(ctx, &p)) != GI_OK) return rc;
    for (int j = 0; j < 2; j++) {
        gi_bset b;
        gi_trove_bset(t, j, &b);
        if ((rc = gi_upload_bset(ctx, j, &b)) != GI_OK) return rc;
    }
    if ((rc = gi_upload_dipole(ctx, t->b0.ncontr, t->mu.data(), 0)) != GI_OK) return rc;
    if ((rc = gi_set_levels(ctx, (int)t->lv.size(), t->energy.data(), t->ljind.data(),
                            t->lgamma.data(), t->lndeg.data(), t->lirec.data(),
                            t->lquanta.data())) != GI_OK)
        return rc;
    for (int j = 0; j < 2; j++)
        for (int g = 0; g < t->deck.nrep; g++) {
            if (!t->deck.isym_do[g]) continue;
            if ((rc = gi_upload_vectors(ctx, j, g, t->bj[j].nsize[g], t->nrec[j][g],
                                        t->vec[j][g].data(), 0)) != GI_OK)
                return rc;
        }
    return gi_prepare(ctx);
}

int gi_trove_apply_thresh_line(gi_trove *t, int on)
{
    if (!t) return GI_EINVAL;
    t->apply_thresh_line = on != 0;
    return GI_OK;
}

int gi_trove_header(const gi_trove *t, char *buf, long long cap, long long *len)
{
    if (!t) return GI_EINVAL;
    // cuda_host.cu:1165
    return put("Linestrength S(f<-i) [Debye**2], Transition moments [Debye],Einstein coefficient "
               "A(if) [1/s],and Intensities [cm/mol]\n\n\n",
               buf, cap, len);
}

// One line per row in the layout of cuda_host.cu:1444-1473.
int gi_trove_format(const gi_trove *t, const gi_lines *L, long long first, long long count,
                    char *buf, long long cap, long long *len)
{
    if (!t || !L || first < 0 || count < 0 || first + count > L->n) return GI_EINVAL;
    const Deck &d = t->deck;
    const int nm = d.nmodes, ncl = t->b0.nclasses;
    // amplitudes per row: m*m, m the largest degeneracy (gi_amp_width)
    const int m = *std::max_element(d.degen.begin(), d.degen.end());
    std::string s;
    s.reserve((size_t)count * 320);
    for (long long r = first; r < first + count; r++) {
        const int iI = L->ilevelI[r], iF = L->ilevelF[r];
        if (iI < 0 || iF < 0 || iI >= (int)t->lv.size() || iF >= (int)t->lv.size())
            return GI_EINVAL;
        const Level &I = t->lv[iI], &F = t->lv[iF];
        appendf(s, "%4i %4s   <-%4i %4s   %1s  %11.4f <- %11.4f %11.4f  ( %3s %3i ) ( ", F.jval,
                d.labels[F.igamma].c_str(), I.jval, d.labels[I.igamma].c_str(),
                branch(F.jval, I.jval), F.energy - d.zpe, I.energy - d.zpe, std::fabs(L->nu[r]),
                d.labels[F.cgamma[0]].c_str(), F.krot);
        for (int i = 1; i <= ncl; i++) appendf(s, " %3s", d.labels[F.cgamma[i]].c_str());
        s += " ";
        for (int i = 1; i <= nm; i++) appendf(s, " %3i", F.quanta[i]);
        appendf(s, " ) <- ( %3s %3i ) ( ", d.labels[I.cgamma[0]].c_str(), I.krot);
        for (int i = 1; i <= ncl; i++) appendf(s, " %3s", d.labels[I.cgamma[i]].c_str());
        s += " ";
        for (int i = 1; i <= nm; i++) appendf(s, " %3i", I.quanta[i]);
        appendf(s, ")  %16.8e %16.8e %16.8e   %6i <- %6i %8i %8i ( ", L->S[r], L->A_einst[r],
                L->I[r], F.ilevel + 1, I.ilevel + 1, 0, 0);
        for (int i = 1; i <= nm; i++) appendf(s, " %3i", F.normal[i]);
        s += " ) <-  ( ";
        for (int i = 1; i <= nm; i++) appendf(s, " %3i", I.normal[i]);
        s += " ) ";
        for (int dF = 0; dF < F.ndeg; dF++)
            for (int dI = 0; dI < I.ndeg; dI++) appendf(s, " %16.9e", L->amp[m * m * r + dI + m * dF]);
        s += "\n";
    }
    return put(s, buf, cap, len);
}

}  // extern "C"
