// mlp_internal.h -- the warm-start network's handle, shared by the sources
// of its entry points (mlp.hip, and the allocation-free forward).
#ifndef MLP_INTERNAL_H
#define MLP_INTERNAL_H
#include <hip/hip_runtime.h>

#include <string>

#include "mcx.h"

typedef __bf16 bf16;

struct mcx_mlp {
    int device = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
    float *P = nullptr, *M = nullptr, *V = nullptr, *G = nullptr, *loss = nullptr;
    bf16 *Wt = nullptr, *Wio = nullptr, *w5 = nullptr;
    long cap = 0;  // rows the work buffers hold
    float *x = nullptr, *tg = nullptr, *y = nullptr, *Z = nullptr, *part = nullptr, *lossp = nullptr;
    long ndata = 0;
    float *X = nullptr, *T = nullptr;  // training set
    long adam_t = 0;
    double fwd_ms = 0.0, step_ms = 0.0;
    long fwd_rows = 0;
};

static inline int mfail(mcx_mlp *net, int code, const std::string &msg) {
    if (net) net->err = msg;
    return code;
}

#define MCHK(net, call)                                                                                        \
    do {                                                                                                        \
        hipError_t e_ = (call);                                                                                 \
        if (e_ != hipSuccess) return mfail(net, MCX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// enqueue the inference forward pass y [n] = network(x [n][nin]) on net->s
// (device pointers)
int mlp_forward_launch(mcx_mlp *net, long n, const float *x, float *y);

#endif
