// mlp_noalloc.hip -- mcx_mlp_forward_noalloc: the network's forward pass for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  The inputs and outputs of a call are staged through
// `scratch`, a fixed device array of MCX_MLP_NOALLOC_MAXROWS rows; the call
// makes no allocation of any kind.
#include "mlp_internal.h"

#include <mutex>

namespace {

constexpr long MAXROWS = MCX_MLP_NOALLOC_MAXROWS;
constexpr long NIN = MCX_MLP_NIN;

// x [MAXROWS][NIN] followed by y [MAXROWS]: 2.25 MiB.  Each GPU the library
// runs on holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
__device__ __align__(256) float scratch[MAXROWS * (NIN + 1)];

std::mutex scratch_mu;
float *scratch_ptr[256] = {};  // per device

// the address of `scratch` on `device`, looked up once per device with that
// device current
int scratch_on(mcx_mlp *net, float **out) {
    const int d = net->device;
    if (d < 0 || d >= 256) return mfail(net, MCX_EINVAL, "mcx_mlp_forward_noalloc: device index");
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_ptr[d]) {
        MCHK(net, hipSetDevice(d));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mfail(net, MCX_EDEVICE,
                         std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[d] = (float *)p;
    }
    *out = scratch_ptr[d];
    return MCX_OK;
}

}  // namespace

extern "C" int mcx_mlp_forward_noalloc(mcx_mlp *net, int64_t n, const float *x, float *y) {
    if (!net || n < 0 || (n > 0 && (!x || !y))) return MCX_EINVAL;
    if (n > MAXROWS)
        return mfail(net, MCX_EINVAL,
                     "mcx_mlp_forward_noalloc: " + std::to_string((long long)n) + " rows exceed its maximum of " +
                         std::to_string(MAXROWS) + " (mcx_mlp_forward takes any number)");
    if (n == 0) return MCX_OK;
    float *base = nullptr;
    int r = scratch_on(net, &base);
    if (r) return r;
    MCHK(net, hipSetDevice(net->device));
    float *dx = base, *dy = base + MAXROWS * NIN;
    MCHK(net, hipMemcpyAsync(dx, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, net->s));
    if ((r = mlp_forward_launch(net, (long)n, dx, dy))) return r;
    MCHK(net, hipMemcpyAsync(y, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, net->s));
    MCHK(net, hipStreamSynchronize(net->s));
    return MCX_OK;
}
