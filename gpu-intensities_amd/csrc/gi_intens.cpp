// gi_intens: the reference's command-line job (main.cu -> dipole_initialise_cpu
// -> dipole_do_intensities_async_omp) on the MI355X engine. The deck comes on
// stdin, the TROVE files from the working directory (or --dir), the line list
// goes to stdout in the reference's format (cuda_host.cu:1165, 1444-1473).
//
//   gi_intens [--mode fp64|batched|split] [--device N] [--dir D] [--thresh-line] < deck
//
// --thresh-line applies the deck's THRESH_LINE as a linestrength cut; by
// default, as in the reference (cuda_host.cu:1408), it is parsed and ignored.
//
// Exit status 0 on success; errors go to stderr with the library's message.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/gi_trove.h"
#include "../../include/gpu_intensities.h"

static int die(const char *what, const char *msg)
{
    std::fprintf(stderr, "gi_intens: %s: %s\n", what, msg ? msg : "");
    return 1;
}

int main(int argc, char **argv)
{
    int mode = GI_MODE_FP64, device = 0, thresh_line = 0;
    std::string dir;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--mode" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "fp64") mode = GI_MODE_FP64;
            else if (m == "batched") mode = GI_MODE_FP64_BATCHED;
            else if (m == "split") mode = GI_MODE_SPLIT;
            else return die("--mode", "expected fp64, batched or split");
        } else if (a == "--device" && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else if (a == "--dir" && i + 1 < argc) {
            dir = argv[++i];
        } else if (a == "--thresh-line") {
            thresh_line = 1;
        } else {
            std::fprintf(stderr, "usage: gi_intens [--mode fp64|batched|split] [--device N] "
                                 "[--dir D] [--thresh-line] < deck\n");
            return 2;
        }
    }
    const std::string deck((std::istreambuf_iterator<char>(std::cin)),
                           std::istreambuf_iterator<char>());
    gi_trove *t = nullptr;
    if (gi_trove_open(dir.c_str(), deck.c_str(), &t) != GI_OK) {
        const int rc = die("input", gi_trove_error(t));
        gi_trove_close(t);
        return rc;
    }
    gi_ctx *ctx = nullptr;
    if (gi_create(device, &ctx) != GI_OK) {
        gi_trove_close(t);
        return die("device", "no HIP device");
    }
    gi_trove_apply_thresh_line(t, thresh_line);
    int rc = gi_trove_load(t, ctx);
    long long cap = 0;
    if (rc == GI_OK) rc = gi_count_candidates(ctx, 0, nullptr, &cap);
    std::vector<int> iI(cap + 1), iF(cap + 1);
    std::vector<double> nu(cap + 1), S(cap + 1), A(cap + 1), I(cap + 1),
        amp((size_t)gi_amp_width(ctx) * (cap + 1));
    gi_lines out = {cap, 0, iI.data(), iF.data(), nu.data(), S.data(), A.data(), I.data(),
                    amp.data()};
    if (rc == GI_OK) rc = gi_compute(ctx, mode, 0, nullptr, &out);
    if (rc != GI_OK) {
        const int r = die("compute", gi_last_error(ctx));
        gi_destroy(ctx);
        gi_trove_close(t);
        return r;
    }
    std::vector<char> buf(1 << 16);
    long long len = 0;
    gi_trove_header(t, buf.data(), (long long)buf.size(), &len);
    std::fwrite(buf.data(), 1, (size_t)len, stdout);
    const long long step = 4096;
    for (long long r = 0; r < out.n; r += step) {
        const long long n = std::min(step, out.n - r);
        if (gi_trove_format(t, &out, r, n, buf.data(), (long long)buf.size(), &len) ==
            GI_ECAPACITY) {
            buf.resize((size_t)len + 1);
            gi_trove_format(t, &out, r, n, buf.data(), (long long)buf.size(), &len);
        }
        std::fwrite(buf.data(), 1, (size_t)len, stdout);
    }
    std::fflush(stdout);
    gi_destroy(ctx);
    gi_trove_close(t);
    return 0;
}
