// gi_half_ls_noalloc: the half linestrength of gi_half_ls with every device
// buffer taken from one file-scope device array instead of hipMalloc, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation). No reference counterpart.
//
// `scratch` below is a host-side stand-in in host code: its device address
// differs per GPU and is looked up once per device with hipGetSymbolAddress
// (noalloc_init, called by gi_create with the device current). Host code only
// ever passes that looked-up address to kernels and copies.
//
// Every call on a GPU shares that GPU's array; calls that overlap on one GPU
// (two host threads, or two contexts used concurrently) race. Each call fully
// writes every byte it later reads (the lower vector and descriptor by copy,
// G, H and the primitive-order result by the kernels), so nothing has to be
// cleared between calls.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "gi_internal.h"
#include "gi_kernels.h"

// 64 MiB: G (3 * Nvib * SMAX doubles, SMAX <= 64 columns) plus H, the lower
// vector and the result fit for Nvib up to about 25,000. At global scope, and
// always named ::scratch below: inside namespace gi the bare name is the
// allocating scratch() helper.
static constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_mu;
std::map<int, void *> g_addr;        // device -> address of `scratch` on it
std::map<int, std::string> g_err;    // device -> why the lookup failed

size_t round256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

namespace gi {

// Looks up this device's copy of `scratch` (the device must be current).
// Called once per device; a failure is kept and reported by the entry point.
void noalloc_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_addr.count(device) || g_err.count(device)) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_addr[device] = p;
}

static int noalloc_addr(gi_ctx *c, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_addr.find(c->device);
    if (it != g_addr.end()) {
        *out = (unsigned char *)it->second;
        return GI_OK;
    }
    auto er = g_err.find(c->device);
    return fail(c, GI_EHIP, er != g_err.end() ? er->second
                                              : std::string("no-allocation scratch not set up"));
}

}  // namespace gi

using namespace gi;

extern "C" int gi_half_ls_noalloc(gi_ctx *c, int jindI, int jindF, int gI, int ideg,
                                  const double *vec, double *h)
{
    if (!c || !vec || !h) return GI_EINVAL;
    if (!c->prepared) return fail(c, GI_ESTATE, "gi_half_ls_noalloc before gi_prepare");
    if (jindI < 0 || jindI >= c->nJ || jindF < 0 || jindF >= c->nJ) return GI_EINVAL;
    if (gI < 0 || gI >= c->nrep || ideg < 0 || ideg >= c->degen[gI]) return GI_EINVAL;
    const Coef &cf = coef_of(c, jindI, jindF);
    if (!cf.valid) return fail(c, GI_EINVAL, "gi_half_ls_noalloc: |J' - J| > 1");
    if (c->mu_streamed) return fail(c, GI_ESTATE, "gi_half_ls_noalloc: needs mu resident");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, GI_EHIP, "hipSetDevice");
    unsigned char *base = nullptr;
    int rc;
    if ((rc = noalloc_addr(c, &base)) != GI_OK) return rc;
    const JPlan &pI = c->jp[jindI], &pF = c->jp[jindF];
    const int N = c->nvib, SI = pI.nslots, SF = pF.nslots, SMAX = smax_for(SF);
    if (SMAX < 0) return fail(c, GI_EINVAL, "too many rotational slots");
    const int nsizeI = c->hb[jindI].nsize[gI];
    // layout in the array: descriptor, lower vector, G, H, result
    const size_t o_m = 0;
    const size_t o_u = o_m + 256;
    const size_t o_G = o_u + round256((size_t)std::max(nsizeI, 1) * 8);
    const size_t o_H = o_G + round256((size_t)3 * N * SMAX * 8);
    const size_t o_h = o_H + round256((size_t)SF * N * 8);
    const size_t need = o_h + round256((size_t)std::max(pF.dimen, 1) * 8);
    if (need > kNoAllocBytes)
        return fail(c, GI_EINVAL, "gi_half_ls_noalloc: needs " + std::to_string(need) +
                                      " bytes of scratch, the no-allocation array holds " +
                                      std::to_string(kNoAllocBytes));
    double *d_u = (double *)(base + o_u), *d_G = (double *)(base + o_G);
    double *d_H = (double *)(base + o_H), *d_h = (double *)(base + o_h);
    unsigned char *d_m = base + o_m;
    // descriptor read by the G build: the lower vector's address, then the
    // component selector (the component's own map is passed as both maps)
    struct {
        long long p;
        int ideg;
        int pad;
    } m{(long long)d_u, ideg, 0};
    hipError_t e = hipMemcpyAsync(d_u, vec, (size_t)nsizeI * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_m, &m, 16, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return fail(c, GI_EHIP, std::string("upload: ") + hipGetErrorString(e));
    const SymDev r = symdev(pI.sym[gI * GI_MAXDEG + ideg]);
    launch_build_g(c->stream, 1, N, SI, SF, SMAX, (const double *const *)d_m,
                   (const int *)(d_m + 8), r, r, cf.d_ptr, cf.d_sI, cf.d_val, d_G);
    if (launch_half_ls_gemv(c->stream, 1, N, SF, SMAX, c->d_mu_perm, d_G, d_H) != 0)
        return fail(c, GI_EINVAL, "no half-ls kernel for this slot count");
    launch_grid_to_prim(c->stream, pF.dimen, pF.d_r_grid, d_H, d_h);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, d_h, (size_t)pF.dimen * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return fail(c, GI_EHIP, std::string("gi_half_ls_noalloc: ") + hipGetErrorString(e));
    return GI_OK;
}
