// mlp.hip -- the warm-start network of the temperature iteration on the
// MI355X matrix cores (gfx950 MFMA, BF16 operands, FP32 accumulation).
//
// The reference starts every model's structure iteration from a fixed
// guess (Init.f:3935) and zeroes T at the start of each Monte Carlo pass
// (MCRadiation.f:276-282); the temperature the first vertical-structure solve
// reads is the first pass's (DiskStructure.f:19, 36-44, 71).  A small network
// that maps a cell's position, its optical depths and the star and disk
// parameters to log T gives that first solve a temperature close to the
// converged one; the Monte Carlo pass that follows judges it.
//
// Network: 8 inputs -> 4 hidden layers of 256 (SiLU) -> 1 output.
//  * layer 1 (K = 8) runs on the VALU in FP32: its inputs are physical
//    coordinates (log r, theta, log tau, ...) that BF16 would quantise to
//    0.4%;
//  * layers 2-4 (256 x 256) run on MFMA v_mfma_f32_32x32x16_bf16 with BF16
//    activations and weights, FP32 accumulation and FP32 biases;
//  * the output layer is a BF16 dot product accumulated in FP32.
// A workgroup of 256 threads (4 waves) owns 64 rows (cells); the rows'
// activations stay in LDS between layers, the weights (128 kB per layer in
// BF16) are read by every workgroup from L2.  Training keeps FP32 master
// weights and Adam state; the backward pass writes one partial gradient per
// workgroup and a second kernel sums them in a fixed order, so a training
// run is bit-reproducible.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mlp_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NIN = MCX_MLP_NIN, WID = MCX_MLP_WIDTH;
constexpr int ROWS = 64;        // rows (cells) per workgroup
constexpr int HS = WID + 8;     // LDS row stride of a row-major tile (bf16; +16 B against bank conflicts)
constexpr int TS = ROWS + 8;    // LDS row stride of a transposed tile (bf16)
constexpr long LW = (long)WID * WID;

// offsets in the flat FP32 parameter vector: W1 [NIN][WID], b1, then per
// hidden layer l = 2..4 W_l [WID in][WID out] and b_l, then w5 [WID], b5
__host__ __device__ constexpr long off_w(int l) { return l == 1 ? 0 : (long)NIN * WID + WID + (long)(l - 2) * (LW + WID); }
__host__ __device__ constexpr long off_b(int l) { return l == 1 ? (long)NIN * WID : off_w(l) + (l == 5 ? WID : LW); }
constexpr long NPAR = off_b(5) + 1;
static_assert(NPAR == MCX_MLP_NPARAMS, "parameter count");

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s + z * s * (1.0f - s);
}

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Row of the 32x32 accumulator element r held by lane `lane` (its column is
// lane % 32): four groups of four consecutive rows.
__device__ __forceinline__ int acc_row(int lane, int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

// Forward pass over rows [blockIdx.x*RT, +RT) of x [n][NIN]: y [n].  TRAIN
// also stores the pre-activations Z [4][n][WID] for the backward pass.  RT
// rows per workgroup: each 256 x 256 layer's BF16 weights (128 kB, from L2)
// serve RT rows, and each wave's B fragment feeds RT/32 MFMA tiles.
template <bool TRAIN, int RT>
__global__ __launch_bounds__(256, 2) void k_mlp_fwd(long n, const float *__restrict__ x, const float *__restrict__ P,
                                                 const bf16 *__restrict__ Wt, const bf16 *__restrict__ w5,
                                                 float *__restrict__ y, float *__restrict__ Z) {
    constexpr int NT = RT / 32;  // row tiles of 32
    __shared__ __attribute__((aligned(16))) bf16 H[RT * HS];
    __shared__ float xs[RT * NIN];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int i32 = lane & 31, kh = (lane >> 5) * 8;
    const long r0 = (long)blockIdx.x * RT;
    for (int k = t; k < RT * NIN; k += 256) xs[k] = (r0 + k / NIN < n) ? x[r0 * NIN + k] : 0.0f;
    __syncthreads();
    // layer 1: thread t owns column t
    {
        float w[NIN];
#pragma unroll
        for (int k = 0; k < NIN; k++) w[k] = P[off_w(1) + k * WID + t];
        const float b = P[off_b(1) + t];
        for (int r = 0; r < RT; r++) {
            float z = b;
#pragma unroll
            for (int k = 0; k < NIN; k++) z = fmaf(xs[r * NIN + k], w[k], z);
            if (TRAIN && r0 + r < n) Z[(r0 + r) * WID + t] = z;
            H[r * HS + t] = (bf16)silu(z);
        }
    }
    // layers 2..4: wave wv computes columns [64 wv, 64 wv + 64) of all RT rows
    for (int l = 2; l <= 4; l++) {
        __syncthreads();
        const bf16 *Wl = Wt + (long)(l - 2) * LW;  // [out][in]
        f32x16 acc[NT][2];
#pragma unroll
        for (int a = 0; a < NT; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = f32x16{};
#pragma unroll 2
        for (int kk = 0; kk < WID; kk += 16) {
            const bf16x8 b0 = *(const bf16x8 *)&Wl[(long)(64 * wv + i32) * WID + kk + kh];
            const bf16x8 b1 = *(const bf16x8 *)&Wl[(long)(64 * wv + 32 + i32) * WID + kk + kh];
#pragma unroll
            for (int ti = 0; ti < NT; ti++) {
                const bf16x8 a = *(const bf16x8 *)&H[(32 * ti + i32) * HS + kk + kh];
                acc[ti][0] = mfma32(a, b0, acc[ti][0]);
                acc[ti][1] = mfma32(a, b1, acc[ti][1]);
            }
        }
        __syncthreads();  // every wave has read H
#pragma unroll
        for (int tj = 0; tj < 2; tj++) {
            const int col = 64 * wv + 32 * tj + i32;
            const float b = P[off_b(l) + col];
#pragma unroll
            for (int ti = 0; ti < NT; ti++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = 32 * ti + acc_row(lane, r);
                    const float z = acc[ti][tj][r] + b;
                    if (TRAIN && r0 + row < n) Z[((long)(l - 1) * n + r0 + row) * WID + col] = z;
                    H[row * HS + col] = (bf16)silu(z);
                }
        }
    }
    __syncthreads();
    // output layer: 256/RT threads per row
    {
        constexpr int TPR = 256 / RT, CPT = WID / TPR;
        const int row = t / TPR, q = t % TPR;
        float s = 0.0f;
        for (int k = q * CPT; k < q * CPT + CPT; k++) s = fmaf((float)H[row * HS + k], (float)w5[k], s);
#pragma unroll
        for (int o = 1; o < TPR; o <<= 1) s += __shfl_xor(s, o);
        if (q == 0 && r0 + row < n) y[r0 + row] = s + P[off_b(5)];
    }
}

// Backward pass of the loss sum_r (y_r - t_r)^2 / ntot over the rows of one
// workgroup: its partial gradient part[blockIdx.x][NPAR] and its sum of
// squared errors lossp[blockIdx.x].
__global__ __launch_bounds__(256) void k_mlp_bwd(long n, float ntot, const float *__restrict__ x,
                                                 const float *__restrict__ tg, const float *__restrict__ y,
                                                 const float *__restrict__ Z, const float *__restrict__ P,
                                                 const bf16 *__restrict__ Wio, const bf16 *__restrict__ w5,
                                                 float *__restrict__ part, float *__restrict__ lossp) {
    // HT: H_{l-1} transposed [in][row]; DZT: dZ_l transposed [out][row];
    // DZ: dZ_l [row][out].  RED aliases HT for the first layer's sums.
    __shared__ __attribute__((aligned(16))) bf16 HT[WID * TS];
    __shared__ __attribute__((aligned(16))) bf16 DZT[WID * TS];
    __shared__ __attribute__((aligned(16))) bf16 DZ[ROWS * HS];
    __shared__ float xs[ROWS * NIN];
    __shared__ float dy[ROWS];
    __shared__ float red[4 * WID];
    float *RED9 = (float *)HT;  // [4][NIN + 1][WID] = WID * TS bf16 exactly
    static_assert(4 * (NIN + 1) * WID * sizeof(float) <= WID * TS * sizeof(bf16), "RED9 fits in HT");
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int i32 = lane & 31, kh = (lane >> 5) * 8;
    const long r0 = (long)blockIdx.x * ROWS;
    float *pb = part + (long)blockIdx.x * NPAR;
    for (int k = t; k < ROWS * NIN; k += 256) xs[k] = (r0 + k / NIN < n) ? x[r0 * NIN + k] : 0.0f;
    if (t < 64) {
        const long r = r0 + t;
        float d = 0.0f, e2 = 0.0f;
        if (r < n) {
            const float e = y[r] - tg[r];
            d = 2.0f * e / ntot;
            e2 = e * e;
        }
        dy[t] = d;
        // fixed-order tree over the wave
        for (int o = 32; o >= 1; o >>= 1) e2 += __shfl_down(e2, o);
        if (t == 0) lossp[blockIdx.x] = e2;
    }
    __syncthreads();
    // output layer and dZ_4: thread t owns column t
    {
        const float w = (float)w5[t];
        float gw = 0.0f, gb = 0.0f;
        for (int r = 0; r < ROWS; r++) {
            const long row = r0 + r;
            const float z = row < n ? Z[(3 * n + row) * WID + t] : 0.0f;
            const float h = (float)(bf16)silu(z);
            gw = fmaf(dy[r], h, gw);
            const float dz = dy[r] * w * dsilu(z);
            gb += dz;
            DZ[r * HS + t] = (bf16)dz;
            DZT[t * TS + r] = (bf16)dz;
        }
        pb[off_w(5) + t] = gw;
        pb[off_b(4) + t] = gb;
        if (t == 0) {
            float s = 0.0f;
            for (int r = 0; r < ROWS; r++) s += dy[r];
            pb[off_b(5)] = s;
        }
    }
    for (int l = 4; l >= 2; l--) {
        // H_{l-1} (the forward pass's BF16 activations) transposed into HT
        for (int r = 0; r < ROWS; r++) {
            const long row = r0 + r;
            const float z = row < n ? Z[((long)(l - 2) * n + row) * WID + t] : 0.0f;
            HT[t * TS + r] = (bf16)silu(z);
        }
        __syncthreads();
        // dW_l partial [in][out] = H_{l-1}^T dZ_l over the 64 rows: 64
        // tiles of 32 x 32, sixteen per wave
        for (int tt = wv * 16; tt < wv * 16 + 16; tt++) {
            const int ti = tt >> 3, tj = tt & 7;
            f32x16 acc = f32x16{};
#pragma unroll
            for (int kk = 0; kk < ROWS; kk += 16) {
                const bf16x8 a = *(const bf16x8 *)&HT[(32 * ti + i32) * TS + kk + kh];
                const bf16x8 b = *(const bf16x8 *)&DZT[(32 * tj + i32) * TS + kk + kh];
                acc = mfma32(a, b, acc);
            }
#pragma unroll
            for (int r = 0; r < 16; r++)
                pb[off_w(l) + (long)(32 * ti + acc_row(lane, r)) * WID + 32 * tj + i32] = acc[r];
        }
        // dH_{l-1} [row][in] = dZ_l W_l^T: wave wv computes columns [64 wv, +64)
        const bf16 *Wl = Wio + (long)(l - 2) * LW;  // [in][out]
        f32x16 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = f32x16{};
#pragma unroll 4
        for (int kk = 0; kk < WID; kk += 16) {
            const bf16x8 a0 = *(const bf16x8 *)&DZ[i32 * HS + kk + kh];
            const bf16x8 a1 = *(const bf16x8 *)&DZ[(32 + i32) * HS + kk + kh];
            const bf16x8 b0 = *(const bf16x8 *)&Wl[(long)(64 * wv + i32) * WID + kk + kh];
            const bf16x8 b1 = *(const bf16x8 *)&Wl[(long)(64 * wv + 32 + i32) * WID + kk + kh];
            acc[0][0] = mfma32(a0, b0, acc[0][0]);
            acc[0][1] = mfma32(a0, b1, acc[0][1]);
            acc[1][0] = mfma32(a1, b0, acc[1][0]);
            acc[1][1] = mfma32(a1, b1, acc[1][1]);
        }
        __syncthreads();  // DZ, DZT and HT have been read
        // dZ_{l-1} = dH_{l-1} silu'(Z_{l-1}); per column the sums over this
        // lane's rows go to RED (first layer: also x_k dZ_1 for dW_1)
        const bool first = (l == 2);
#pragma unroll
        for (int tj = 0; tj < 2; tj++) {
            const int col = 64 * wv + 32 * tj + i32;
#pragma unroll
            for (int ti = 0; ti < 2; ti++) {
                float cs = 0.0f;
                float gx[NIN];
#pragma unroll
                for (int k = 0; k < NIN; k++) gx[k] = 0.0f;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int row = 32 * ti + acc_row(lane, r);
                    const long grow = r0 + row;
                    const float z = grow < n ? Z[((long)(l - 2) * n + grow) * WID + col] : 0.0f;
                    const float dz = acc[ti][tj][r] * dsilu(z);
                    cs += dz;
                    if (first) {
#pragma unroll
                        for (int k = 0; k < NIN; k++) gx[k] = fmaf(xs[row * NIN + k], dz, gx[k]);
                    } else {
                        DZ[row * HS + col] = (bf16)dz;
                        DZT[col * TS + row] = (bf16)dz;
                    }
                }
                const int q = ti * 2 + (lane >> 5);
                if (first) {
                    RED9[(q * (NIN + 1) + NIN) * WID + col] = cs;
#pragma unroll
                    for (int k = 0; k < NIN; k++) RED9[(q * (NIN + 1) + k) * WID + col] = gx[k];
                } else {
                    red[q * WID + col] = cs;
                }
            }
        }
        __syncthreads();
        if (first) {
            for (int k = 0; k <= NIN; k++) {
                float s = 0.0f;
                for (int q = 0; q < 4; q++) s += RED9[(q * (NIN + 1) + k) * WID + t];
                if (k < NIN)
                    pb[off_w(1) + k * WID + t] = s;
                else
                    pb[off_b(1) + t] = s;
            }
        } else {
            pb[off_b(l - 1) + t] = red[t] + red[WID + t] + red[2 * WID + t] + red[3 * WID + t];
        }
        __syncthreads();
    }
}

// grad[p] = sum over workgroups of part[b][p], in workgroup order; the
// loss (mean squared error) from the per-workgroup sums
__global__ void k_mlp_reduce(int nblk, const float *__restrict__ part, float *__restrict__ grad,
                             const float *__restrict__ lossp, float ntot, float *__restrict__ loss) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPAR) {
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += part[(long)b * NPAR + p];
        grad[p] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += lossp[b];
        *loss = (float)(s / ntot);
    }
}

// Adam on the FP32 master weights (update != 0), then the BF16 copies the
// MFMA layers read: W_l as [out][in] (forward) and [in][out] (backward), w5
__global__ void k_mlp_adam(int update, float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                           const float *__restrict__ G, float lr, float b1, float b2, float eps, float bc1, float bc2,
                           bf16 *__restrict__ Wt, bf16 *__restrict__ Wio, bf16 *__restrict__ w5) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    float v = P[p];
    if (update) {
        const float g = G[p];
        const float m1 = b1 * M[p] + (1.0f - b1) * g;
        const float v1 = b2 * V[p] + (1.0f - b2) * g * g;
        M[p] = m1;
        V[p] = v1;
        v -= lr * (m1 / bc1) / (sqrtf(v1 / bc2) + eps);
        P[p] = v;
    }
    for (int l = 2; l <= 4; l++) {
        const long o = p - off_w(l);
        if (o >= 0 && o < LW) {
            const long in = o / WID, out = o % WID;
            Wio[(l - 2) * LW + o] = (bf16)v;
            Wt[(l - 2) * LW + out * WID + in] = (bf16)v;
        }
    }
    const long o5 = p - off_w(5);
    if (o5 >= 0 && o5 < WID) w5[o5] = (bf16)v;
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// a minibatch of B rows drawn with replacement from the n rows of the data
__global__ void k_mlp_gather(int B, long n, const float *__restrict__ X, const float *__restrict__ T,
                             unsigned long long key, float *__restrict__ xb, float *__restrict__ tb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const long r = (long)(splitmix64(key + (unsigned long long)i) % (unsigned long long)n);
#pragma unroll
    for (int k = 0; k < NIN; k++) xb[(long)i * NIN + k] = X[r * NIN + k];
    tb[i] = T[r];
}

}  // namespace

namespace {

template <typename T>
int malloc_dev(mcx_mlp *net, T **p, size_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc((void **)p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return mfail(net, MCX_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    }
    return 0;
}

// work buffers for n rows of a training batch
int ensure_rows(mcx_mlp *net, long n) {
    if (n <= net->cap) return 0;
    const long nblk = (n + ROWS - 1) / ROWS;
    int r;
    if ((r = malloc_dev(net, &net->x, (size_t)n * NIN)) || (r = malloc_dev(net, &net->tg, (size_t)n)) ||
        (r = malloc_dev(net, &net->y, (size_t)n)) || (r = malloc_dev(net, &net->Z, (size_t)4 * n * WID)) ||
        (r = malloc_dev(net, &net->part, (size_t)nblk * NPAR)) || (r = malloc_dev(net, &net->lossp, (size_t)nblk))) {
        net->cap = 0;
        return r;
    }
    net->cap = n;
    return 0;
}

// forward + backward + reduction over the n rows in net->x / net->tg
int launch_grad(mcx_mlp *net, long n) {
    const int nblk = (int)((n + ROWS - 1) / ROWS);
    hipLaunchKernelGGL((k_mlp_fwd<true, ROWS>), dim3(nblk), dim3(256), 0, net->s, n, net->x, net->P, net->Wt,
                       net->w5, net->y, net->Z);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nblk), dim3(256), 0, net->s, n, (float)n, net->x, net->tg, net->y, net->Z,
                       net->P, net->Wio, net->w5, net->part, net->lossp);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)((NPAR + 255) / 256)), dim3(256), 0, net->s, nblk, net->part,
                       net->G, net->lossp, (float)n, net->loss);
    MCHK(net, hipGetLastError());
    return 0;
}

int launch_cast(mcx_mlp *net, int update, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((NPAR + 255) / 256)), dim3(256), 0, net->s, update, net->P, net->M,
                       net->V, net->G, lr, b1, b2, eps, bc1, bc2, net->Wt, net->Wio, net->w5);
    MCHK(net, hipGetLastError());
    return 0;
}

}  // namespace

int mlp_forward_launch(mcx_mlp *net, long n, const float *x, float *y) {
    if (n <= 0) return MCX_OK;
    constexpr int RI = 128;  // rows per workgroup of the inference pass
    hipLaunchKernelGGL((k_mlp_fwd<false, RI>), dim3((unsigned)((n + RI - 1) / RI)), dim3(256), 0, net->s, n, x,
                       net->P, net->Wt, net->w5, y, (float *)nullptr);
    MCHK(net, hipGetLastError());
    return MCX_OK;
}

extern "C" {

int mcx_mlp_create(mcx_mlp **out, int device, int32_t nin, int32_t width, int32_t depth) {
    if (!out) return MCX_EINVAL;
    *out = nullptr;
    if (nin != NIN || width != WID || depth != MCX_MLP_DEPTH) return MCX_EINVAL;
    mcx_mlp *net = new mcx_mlp();
    net->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&net->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&net->ev[0]) != hipSuccess || hipEventCreate(&net->ev[1]) != hipSuccess) {
        delete net;
        return MCX_EDEVICE;
    }
    int r;
    if ((r = malloc_dev(net, &net->P, NPAR)) || (r = malloc_dev(net, &net->M, NPAR)) ||
        (r = malloc_dev(net, &net->V, NPAR)) || (r = malloc_dev(net, &net->G, NPAR)) ||
        (r = malloc_dev(net, &net->loss, 1)) || (r = malloc_dev(net, &net->Wt, 3 * LW)) ||
        (r = malloc_dev(net, &net->Wio, 3 * LW)) || (r = malloc_dev(net, &net->w5, WID))) {
        mcx_mlp_destroy(net);
        return r;
    }
    *out = net;
    return MCX_OK;
}

void mcx_mlp_destroy(mcx_mlp *net) {
    if (!net) return;
    (void)hipSetDevice(net->device);
    if (net->s) (void)hipStreamSynchronize(net->s);
    void *bufs[] = {net->P, net->M, net->V, net->G, net->loss, net->Wt, net->Wio, net->w5, net->x,
                    net->tg, net->y, net->Z, net->part, net->lossp, net->X, net->T};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &e : net->ev)
        if (e) (void)hipEventDestroy(e);
    if (net->s) (void)hipStreamDestroy(net->s);
    delete net;
}

const char *mcx_mlp_last_error(const mcx_mlp *net) { return net ? net->err.c_str() : "no network"; }

int mcx_mlp_set_params(mcx_mlp *net, const float *params) {
    if (!net || !params) return MCX_EINVAL;
    MCHK(net, hipSetDevice(net->device));
    MCHK(net, hipMemcpyAsync(net->P, params, NPAR * sizeof(float), hipMemcpyHostToDevice, net->s));
    MCHK(net, hipMemsetAsync(net->M, 0, NPAR * sizeof(float), net->s));
    MCHK(net, hipMemsetAsync(net->V, 0, NPAR * sizeof(float), net->s));
    net->adam_t = 0;
    int r = launch_cast(net, 0, 0, 0, 0, 0, 1, 1);
    if (r) return r;
    MCHK(net, hipStreamSynchronize(net->s));
    return MCX_OK;
}

int mcx_mlp_get_params(mcx_mlp *net, float *params) {
    if (!net || !params) return MCX_EINVAL;
    MCHK(net, hipSetDevice(net->device));
    MCHK(net, hipMemcpyAsync(params, net->P, NPAR * sizeof(float), hipMemcpyDeviceToHost, net->s));
    MCHK(net, hipStreamSynchronize(net->s));
    return MCX_OK;
}

int mcx_mlp_forward(mcx_mlp *net, int64_t n, const float *x, float *y) {
    if (!net || n < 0 || (n > 0 && (!x || !y))) return MCX_EINVAL;
    if (n == 0) return MCX_OK;
    MCHK(net, hipSetDevice(net->device));
    const long chunk = 1L << 20;
    float *dx = nullptr, *dy = nullptr;
    const long m0 = n < chunk ? (long)n : chunk;
    if (hipMalloc((void **)&dx, m0 * NIN * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&dy, m0 * sizeof(float)) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return mfail(net, MCX_ENOMEM, "mcx_mlp_forward: device buffers");
    }
    int rc = MCX_OK;
    double ms = 0.0;
    for (long a = 0; a < n && rc == MCX_OK; a += chunk) {
        const long m = (n - a) < chunk ? (n - a) : chunk;
        if (hipMemcpyAsync(dx, x + a * NIN, m * NIN * sizeof(float), hipMemcpyHostToDevice, net->s) != hipSuccess) {
            rc = mfail(net, MCX_EDEVICE, "mcx_mlp_forward: upload");
            break;
        }
        (void)hipEventRecord(net->ev[0], net->s);
        if (mlp_forward_launch(net, m, dx, dy) != MCX_OK) {
            rc = MCX_EDEVICE;
            break;
        }
        (void)hipEventRecord(net->ev[1], net->s);
        if (hipMemcpyAsync(y + a, dy, m * sizeof(float), hipMemcpyDeviceToHost, net->s) != hipSuccess ||
            hipStreamSynchronize(net->s) != hipSuccess) {
            rc = mfail(net, MCX_EDEVICE, "mcx_mlp_forward: kernel");
            break;
        }
        float f = 0.0f;
        (void)hipEventElapsedTime(&f, net->ev[0], net->ev[1]);
        ms += f;
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    net->fwd_ms = ms;
    net->fwd_rows = (long)n;
    return rc;
}

int mcx_mlp_gradient(mcx_mlp *net, int64_t n, const float *x, const float *t, double *loss, float *grad) {
    if (!net || n <= 0 || !x || !t) return MCX_EINVAL;
    MCHK(net, hipSetDevice(net->device));
    int r = ensure_rows(net, (long)n);
    if (r) return r;
    MCHK(net, hipMemcpyAsync(net->x, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, net->s));
    MCHK(net, hipMemcpyAsync(net->tg, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->s));
    if ((r = launch_grad(net, (long)n))) return r;
    float l = 0.0f;
    MCHK(net, hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->s));
    if (grad) MCHK(net, hipMemcpyAsync(grad, net->G, NPAR * sizeof(float), hipMemcpyDeviceToHost, net->s));
    MCHK(net, hipStreamSynchronize(net->s));
    if (loss) *loss = l;
    return MCX_OK;
}

int mcx_mlp_set_data(mcx_mlp *net, int64_t n, const float *x, const float *t) {
    if (!net || n <= 0 || !x || !t) return MCX_EINVAL;
    MCHK(net, hipSetDevice(net->device));
    int r;
    if ((r = malloc_dev(net, &net->X, (size_t)n * NIN)) || (r = malloc_dev(net, &net->T, (size_t)n))) {
        net->ndata = 0;
        return r;
    }
    MCHK(net, hipMemcpyAsync(net->X, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, net->s));
    MCHK(net, hipMemcpyAsync(net->T, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->s));
    MCHK(net, hipStreamSynchronize(net->s));
    net->ndata = (long)n;
    return MCX_OK;
}

int mcx_mlp_train(mcx_mlp *net, const mcx_mlp_train_opts *o, float *loss) {
    if (!net || !o || o->batch <= 0 || o->nsteps < 0) return MCX_EINVAL;
    if (net->ndata <= 0) return mfail(net, MCX_ESTATE, "mcx_mlp_train: no data (mcx_mlp_set_data)");
    MCHK(net, hipSetDevice(net->device));
    int r = ensure_rows(net, o->batch);
    if (r) return r;
    float *hist = nullptr;
    if (loss) {
        if ((r = malloc_dev(net, &hist, (size_t)(o->nsteps ? o->nsteps : 1)))) return r;
    }
    (void)hipEventRecord(net->ev[0], net->s);
    for (int k = 0; k < o->nsteps; k++) {
        const unsigned long long key = o->seed * 0x100000001B3ULL + (unsigned long long)net->adam_t * (unsigned long long)o->batch;
        hipLaunchKernelGGL(k_mlp_gather, dim3((o->batch + 255) / 256), dim3(256), 0, net->s, o->batch, net->ndata,
                           net->X, net->T, key, net->x, net->tg);
        if ((r = launch_grad(net, o->batch))) break;
        net->adam_t++;
        const float bc1 = 1.0f - powf(o->beta1, (float)net->adam_t);
        const float bc2 = 1.0f - powf(o->beta2, (float)net->adam_t);
        if ((r = launch_cast(net, 1, o->lr, o->beta1, o->beta2, o->eps, bc1, bc2))) break;
        if (hist && hipMemcpyAsync(hist + k, net->loss, sizeof(float), hipMemcpyDeviceToDevice, net->s) != hipSuccess) {
            r = mfail(net, MCX_EDEVICE, "mcx_mlp_train: loss record");
            break;
        }
    }
    (void)hipEventRecord(net->ev[1], net->s);
    if (!r && hist && o->nsteps > 0)
        r = hipMemcpyAsync(loss, hist, (size_t)o->nsteps * sizeof(float), hipMemcpyDeviceToHost, net->s) == hipSuccess
                ? 0
                : mfail(net, MCX_EDEVICE, "mcx_mlp_train: loss copy");
    if (hipStreamSynchronize(net->s) != hipSuccess && !r) r = mfail(net, MCX_EDEVICE, "mcx_mlp_train: kernels");
    float f = 0.0f;
    (void)hipEventElapsedTime(&f, net->ev[0], net->ev[1]);
    net->step_ms = o->nsteps ? f / o->nsteps : 0.0;
    if (hist) (void)hipFree(hist);
    return r;
}

int mcx_mlp_last_timing(const mcx_mlp *net, double *forward_ms, int64_t *forward_rows, double *step_ms) {
    if (!net) return MCX_EINVAL;
    if (forward_ms) *forward_ms = net->fwd_ms;
    if (forward_rows) *forward_rows = net->fwd_rows;
    if (step_ms) *step_ms = net->step_ms;
    return MCX_OK;
}

}  // extern "C"
