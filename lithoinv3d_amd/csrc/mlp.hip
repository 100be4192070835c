// mlp.hip -- the learned diagonal initial-Hessian model: a 16 -> 64 -> 64 -> 1 GELU MLP
// trained and evaluated on gfx950 with BF16 MFMA (v_mfma_f32_16x16x16_bf16) and FP32
// accumulation.  No reference counterpart: it supplies the diagonal D of H0 = gamma*D for
// the projected L-BFGS (the reference uses H0 = gamma*I, lbfgsb_optimizer.cpp:40-62).
//
// One wave (64 lanes) processes 16 samples per tile.  Products are formed transposed,
// Z^T = W^T X^T, so that a layer's 16x16 MFMA result (lane l: sample l%16, output rows
// 4*(l/16)..+3) is already the B operand of the next layer's MFMA for the matching
// K block -- no data movement between layers in the forward.  The weight gradients need
// the activations with samples as the reduction index; those tiles go through the wave's
// own LDS (2-4 KB) once per tile.
//
// Parameter layout (FP32 master copy): W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lithofwd.h"

namespace {

constexpr int IN = 16, HID = 64;
constexpr int OFF_W1 = 0, OFF_B1 = IN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID, NPARAM = OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float x)
{
    // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float phi = 0.39894228040143268f * __expf(-0.5f * z * z);
    return Phi + z * phi;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ P,
                                                 unsigned short* __restrict__ W1b,
                                                 unsigned short* __restrict__ W2b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < IN * HID) W1b[i] = f2bf(P[OFF_W1 + i]);
    if (i < HID * HID) W2b[i] = f2bf(P[OFF_W2 + i]);
}

// TRAIN = false: Y[s] = MLP(X[s]).  TRAIN = true: also per-wave partial gradients of
// loss = sum_s (y_s - t_s)^2 * inv_n into part[wave][NPARAM] and part_loss[wave].
template <bool TRAIN>
__global__ __launch_bounds__(64) void k_mlp(const float* __restrict__ X, const float* __restrict__ T,
                                            int64_t n, const float* __restrict__ P,
                                            const unsigned short* __restrict__ W1b,
                                            const unsigned short* __restrict__ W2b,
                                            float* __restrict__ Y, float* __restrict__ part,
                                            float* __restrict__ part_loss, float inv_n)
{
    __shared__ unsigned short sH1[16][HID], sdZ2[16][HID], sdZ1[16][HID], sX[16][IN];
    const int l = threadIdx.x;
    const int r = l & 15, g = l >> 4;
    // A operands (lane: row r, k = 4g + j)
    bf16x4 W1T[4], W2T[4][4], W2N[4][4];
    for (int mb = 0; mb < 4; ++mb) {
        for (int j = 0; j < 4; ++j) W1T[mb][j] = (short)W1b[(4 * g + j) * HID + 16 * mb + r];
        for (int kk = 0; kk < 4; ++kk)
            for (int j = 0; j < 4; ++j) {
                W2T[mb][kk][j] = (short)W2b[(16 * kk + 4 * g + j) * HID + 16 * mb + r];
                W2N[mb][kk][j] = (short)W2b[(16 * mb + r) * HID + 16 * kk + 4 * g + j];
            }
    }
    // D-layout rows owned by this lane: 16*mb + 4*g + i
    float b1[4][4], b2[4][4], w3[4][4];
    for (int mb = 0; mb < 4; ++mb)
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mb + 4 * g + i;
            b1[mb][i] = P[OFF_B1 + o];
            b2[mb][i] = P[OFF_B2 + o];
            w3[mb][i] = P[OFF_W3 + o];
        }
    const float b3 = P[OFF_B3];
    f32x4 gW1[4], gW2[4][4];
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.f, lacc = 0.f;
    if (TRAIN) {
        for (int mb = 0; mb < 4; ++mb) {
            gW1[mb] = f32x4{0, 0, 0, 0};
            for (int nb = 0; nb < 4; ++nb) gW2[mb][nb] = f32x4{0, 0, 0, 0};
            for (int i = 0; i < 4; ++i) gb1[mb][i] = gb2[mb][i] = gw3[mb][i] = 0.f;
        }
    }
    const int64_t n_tiles = (n + 15) / 16;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t s = tile * 16 + r;
        const bool valid = s < n;
        float xv[4];
        for (int j = 0; j < 4; ++j) xv[j] = valid ? X[s * IN + 4 * g + j] : 0.f;
        const bf16x4 Bx = pack4(xv[0], xv[1], xv[2], xv[3]);
        float u1[4][4], h1[4][4];
        bf16x4 H1b[4];
        for (int mb = 0; mb < 4; ++mb) {
            const f32x4 z = mfma(W1T[mb], Bx, f32x4{0, 0, 0, 0});
            for (int i = 0; i < 4; ++i) {
                u1[mb][i] = z[i] + b1[mb][i];
                h1[mb][i] = gelu(u1[mb][i]);
            }
            H1b[mb] = pack4(h1[mb][0], h1[mb][1], h1[mb][2], h1[mb][3]);
        }
        float u2[4][4], h2[4][4];
        float p = 0.f;
        for (int mb = 0; mb < 4; ++mb) {
            f32x4 z = f32x4{0, 0, 0, 0};
            for (int kk = 0; kk < 4; ++kk) z = mfma(W2T[mb][kk], H1b[kk], z);
            for (int i = 0; i < 4; ++i) {
                u2[mb][i] = z[i] + b2[mb][i];
                h2[mb][i] = gelu(u2[mb][i]);
                p += w3[mb][i] * h2[mb][i];
            }
        }
        p += __shfl_xor(p, 16);
        p += __shfl_xor(p, 32);
        const float y = p + b3;
        if (!TRAIN) {
            if (g == 0 && valid) Y[s] = y;
            continue;
        }
        const float e = valid ? (y - T[s]) : 0.f;
        const float dy = 2.f * e * inv_n;
        if (g == 0) {
            lacc += e * e * inv_n;
            gb3 += dy;
        }
        float dz2[4][4];
        bf16x4 dZ2b[4];
        for (int mb = 0; mb < 4; ++mb) {
            for (int i = 0; i < 4; ++i) {
                gw3[mb][i] += dy * h2[mb][i];
                dz2[mb][i] = dy * w3[mb][i] * gelu_d(u2[mb][i]);
                gb2[mb][i] += dz2[mb][i];
            }
            dZ2b[mb] = pack4(dz2[mb][0], dz2[mb][1], dz2[mb][2], dz2[mb][3]);
        }
        // dH1 = W2 dZ2 (rows: layer-1 outputs)
        float dz1[4][4];
        for (int mb = 0; mb < 4; ++mb) {
            f32x4 d = f32x4{0, 0, 0, 0};
            for (int kk = 0; kk < 4; ++kk) d = mfma(W2N[mb][kk], dZ2b[kk], d);
            for (int i = 0; i < 4; ++i) {
                dz1[mb][i] = d[i] * gelu_d(u1[mb][i]);
                gb1[mb][i] += dz1[mb][i];
            }
        }
        // tiles with the sample index as the reduction dimension, through LDS
        for (int mb = 0; mb < 4; ++mb)
            for (int i = 0; i < 4; ++i) {
                const int o = 16 * mb + 4 * g + i;
                sH1[r][o] = (unsigned short)H1b[mb][i];
                sdZ2[r][o] = (unsigned short)dZ2b[mb][i];
                sdZ1[r][o] = f2bf(dz1[mb][i]);
            }
        for (int j = 0; j < 4; ++j) sX[r][4 * g + j] = (unsigned short)Bx[j];
        __syncthreads();
        // dW2^T[out][in] += sum_s dZ2^T[out][s] H1[s][in]
        for (int mb = 0; mb < 4; ++mb) {
            bf16x4 A, A1;
            for (int j = 0; j < 4; ++j) {
                A[j] = (short)sdZ2[4 * g + j][16 * mb + r];
                A1[j] = (short)sdZ1[4 * g + j][16 * mb + r];
            }
            for (int nb = 0; nb < 4; ++nb) {
                bf16x4 B;
                for (int j = 0; j < 4; ++j) B[j] = (short)sH1[4 * g + j][16 * nb + r];
                gW2[mb][nb] = mfma(A, B, gW2[mb][nb]);
            }
            // dW1^T[out][in] += sum_s dZ1^T[out][s] X[s][in]
            bf16x4 BX;
            for (int j = 0; j < 4; ++j) BX[j] = (short)sX[4 * g + j][r];
            gW1[mb] = mfma(A1, BX, gW1[mb]);
        }
        __syncthreads();
    }
    if (!TRAIN) return;
    float* out = part + (int64_t)blockIdx.x * NPARAM;
    for (int mb = 0; mb < 4; ++mb)
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mb + 4 * g + i;
            out[OFF_W1 + r * HID + o] = gW1[mb][i];
            for (int nb = 0; nb < 4; ++nb) out[OFF_W2 + (16 * nb + r) * HID + o] = gW2[mb][nb][i];
        }
    // per-row sums over the 16 samples held by lanes r = 0..15 of each group
    for (int mb = 0; mb < 4; ++mb)
        for (int i = 0; i < 4; ++i) {
            float a = gb1[mb][i], b = gb2[mb][i], c = gw3[mb][i];
            for (int m = 1; m < 16; m <<= 1) {
                a += __shfl_xor(a, m);
                b += __shfl_xor(b, m);
                c += __shfl_xor(c, m);
            }
            if (r == 0) {
                const int o = 16 * mb + 4 * g + i;
                out[OFF_B1 + o] = a;
                out[OFF_B2 + o] = b;
                out[OFF_W3 + o] = c;
            }
        }
    for (int m = 1; m < 16; m <<= 1) {
        gb3 += __shfl_xor(gb3, m);
        lacc += __shfl_xor(lacc, m);
    }
    if (l == 0) {
        out[OFF_B3] = gb3;
        part_loss[blockIdx.x] = lacc;
    }
}

__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part,
                                                      const float* __restrict__ part_loss,
                                                      int n_parts, float* __restrict__ grad,
                                                      float* __restrict__ loss_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM) {
        float a = 0.f;
        for (int w = 0; w < n_parts; ++w) a += part[(int64_t)w * NPARAM + k];
        grad[k] = a;
    }
    if (k == 0) {
        float a = 0.f;
        for (int w = 0; w < n_parts; ++w) a += part_loss[w];
        *loss_out = a;
    }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, float* __restrict__ m,
                                              float* __restrict__ v,
                                              const float* __restrict__ grad, float lr,
                                              float b1c, float b2c)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM) return;
    const float g = grad[k];
    const float mk = 0.9f * m[k] + 0.1f * g;
    const float vk = 0.999f * v[k] + 0.001f * g * g;
    m[k] = mk;
    v[k] = vk;
    P[k] -= lr * (mk / b1c) / (sqrtf(vk / b2c) + 1e-8f);
}


}  // namespace

struct lf_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *P = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr,
          *part_loss = nullptr, *losses = nullptr;
    unsigned short *W1b = nullptr, *W2b = nullptr;
    float *X = nullptr, *T = nullptr, *Y = nullptr;
    int64_t cap = 0;
    int losses_cap = 0;  // floats in losses
    int n_parts = 512;
    int64_t step = 0;
    std::string err;
};

namespace {
thread_local std::string g_mlp_err;

int mset(lf_mlp* h, const std::string& s, int code)
{
    if (h) h->err = s;
    else g_mlp_err = s;
    return code;
}

#define MT(h, expr)                                                                   \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return mset(h, std::string(#expr) + ": " + hipGetErrorString(e_), LF_EHIP); \
    } while (0)

int ensure_cap(lf_mlp* h, int64_t n)
{
    if (n <= h->cap) return LF_OK;
    if (h->X) hipFree(h->X);
    if (h->T) hipFree(h->T);
    if (h->Y) hipFree(h->Y);
    h->X = h->T = h->Y = nullptr;
    h->cap = 0;
    MT(h, hipMalloc((void**)&h->X, sizeof(float) * IN * n));
    MT(h, hipMalloc((void**)&h->T, sizeof(float) * n));
    MT(h, hipMalloc((void**)&h->Y, sizeof(float) * n));
    h->cap = n;
    return LF_OK;
}

int upload(lf_mlp* h, int64_t n, const float* x, const float* t)
{
    int rc = ensure_cap(h, std::max<int64_t>(n, 1));
    if (rc) return rc;
    if (n) MT(h, hipMemcpyAsync(h->X, x, sizeof(float) * IN * n, hipMemcpyHostToDevice, h->stream));
    if (t && n) MT(h, hipMemcpyAsync(h->T, t, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    return LF_OK;
}

int grad_pass(lf_mlp* h, int64_t n, float* loss_slot)
{
    hipLaunchKernelGGL(k_to_bf16, dim3((HID * HID + 255) / 256), dim3(256), 0, h->stream, h->P,
                       h->W1b, h->W2b);
    hipLaunchKernelGGL(k_mlp<true>, dim3(h->n_parts), dim3(64), 0, h->stream, h->X, h->T, n, h->P,
                       h->W1b, h->W2b, h->Y, h->part, h->part_loss, n > 0 ? 1.0f / (float)n : 0.f);
    hipLaunchKernelGGL(k_reduce_parts, dim3((NPARAM + 255) / 256), dim3(256), 0, h->stream, h->part,
                       h->part_loss, h->n_parts, h->grad, loss_slot);
    MT(h, hipGetLastError());
    return LF_OK;
}
}  // namespace

extern "C" {

int64_t lf_mlp_param_count(void) { return NPARAM; }

const char* lf_mlp_last_error(const lf_mlp* h) { return h ? h->err.c_str() : g_mlp_err.c_str(); }

void lf_mlp_destroy(lf_mlp* h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    void* b[] = {h->P, h->m, h->v, h->grad, h->part, h->part_loss, h->losses, h->W1b, h->W2b,
                 h->X, h->T, h->Y};
    for (void* p : b)
        if (p) hipFree(p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

int lf_mlp_create(lf_mlp** out, int device, uint64_t seed)
{
    if (!out) return mset(nullptr, "lf_mlp_create: null argument", LF_EINVAL);
    *out = nullptr;
    lf_mlp* h = new lf_mlp();
    h->device = device;
    auto fail = [&](int rc) {
        g_mlp_err = h->err;
        lf_mlp_destroy(h);
        return rc;
    };
#define MC(expr)                                                                       \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) { h->err = std::string(#expr) + ": " + hipGetErrorString(e_); return fail(LF_EHIP); } \
    } while (0)
    MC(hipSetDevice(device));
    MC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    MC(hipMalloc((void**)&h->P, sizeof(float) * NPARAM));
    MC(hipMalloc((void**)&h->m, sizeof(float) * NPARAM));
    MC(hipMalloc((void**)&h->v, sizeof(float) * NPARAM));
    MC(hipMalloc((void**)&h->grad, sizeof(float) * NPARAM));
    MC(hipMalloc((void**)&h->part, sizeof(float) * NPARAM * h->n_parts));
    MC(hipMalloc((void**)&h->part_loss, sizeof(float) * h->n_parts));
    MC(hipMalloc((void**)&h->W1b, sizeof(unsigned short) * IN * HID));
    MC(hipMalloc((void**)&h->W2b, sizeof(unsigned short) * HID * HID));
    // He-normal weights (std sqrt(2/fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> p(NPARAM, 0.f);
    for (int i = 0; i < IN * HID; ++i) p[OFF_W1 + i] = nd(rng) * std::sqrt(2.f / IN);
    for (int i = 0; i < HID * HID; ++i) p[OFF_W2 + i] = nd(rng) * std::sqrt(2.f / HID);
    for (int i = 0; i < HID; ++i) p[OFF_W3 + i] = nd(rng) * std::sqrt(1.f / HID);
    MC(hipMemcpy(h->P, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    std::vector<float> z(NPARAM, 0.f);
    MC(hipMemcpy(h->m, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    MC(hipMemcpy(h->v, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
#undef MC
    *out = h;
    return LF_OK;
}

int lf_mlp_set_params(lf_mlp* h, const float* params)
{
    if (!h || !params) return mset(h, "lf_mlp_set_params: null argument", LF_EINVAL);
    MT(h, hipSetDevice(h->device));
    MT(h, hipMemcpyAsync(h->P, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, h->stream));
    MT(h, hipStreamSynchronize(h->stream));
    return LF_OK;
}

int lf_mlp_get_params(lf_mlp* h, float* params)
{
    if (!h || !params) return mset(h, "lf_mlp_get_params: null argument", LF_EINVAL);
    MT(h, hipSetDevice(h->device));
    MT(h, hipMemcpyAsync(params, h->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream));
    MT(h, hipStreamSynchronize(h->stream));
    return LF_OK;
}

int lf_mlp_forward(lf_mlp* h, int64_t n, const float* x, float* y)
{
    if (!h || n < 0 || (n && (!x || !y))) return mset(h, "lf_mlp_forward: invalid arguments", LF_EINVAL);
    MT(h, hipSetDevice(h->device));
    int rc = upload(h, n, x, nullptr);
    if (rc) return rc;
    if (n) {
        hipLaunchKernelGGL(k_to_bf16, dim3((HID * HID + 255) / 256), dim3(256), 0, h->stream, h->P,
                           h->W1b, h->W2b);
        const int64_t tiles = (n + 15) / 16;
        const unsigned grid = (unsigned)std::min<int64_t>(tiles, 4096);
        hipLaunchKernelGGL(k_mlp<false>, dim3(grid), dim3(64), 0, h->stream, h->X, nullptr, n, h->P,
                           h->W1b, h->W2b, h->Y, nullptr, nullptr, 0.f);
        MT(h, hipGetLastError());
        MT(h, hipMemcpyAsync(y, h->Y, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    }
    MT(h, hipStreamSynchronize(h->stream));
    return LF_OK;
}

int lf_mlp_loss_grad(lf_mlp* h, int64_t n, const float* x, const float* t, float* loss,
                     float* grad)
{
    if (!h || n < 0 || (n && (!x || !t)) || !loss || !grad)
        return mset(h, "lf_mlp_loss_grad: invalid arguments", LF_EINVAL);
    MT(h, hipSetDevice(h->device));
    int rc = upload(h, n, x, t);
    if (rc) return rc;
    float* dloss = nullptr;
    MT(h, hipMalloc((void**)&dloss, sizeof(float)));
    rc = grad_pass(h, n, dloss);
    if (!rc) {
        hipMemcpyAsync(grad, h->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, h->stream);
        hipMemcpyAsync(loss, dloss, sizeof(float), hipMemcpyDeviceToHost, h->stream);
        hipStreamSynchronize(h->stream);
    }
    hipFree(dloss);
    return rc;
}

// Training steps between stream synchronizations: each step enqueues four launches, so at
// most 4 * kTrainSyncSteps launches are outstanding.  Without a bound a long training run
// (3000 steps) queues 12 000 dispatches ahead of the GPU, each of which the HIP runtime,
// and a profiler collecting per-dispatch counters, tracks until it completes; a run of
// that kind under counter collection ended in a host crash inside the enqueue.
constexpr int kTrainSyncSteps = 256;

int lf_mlp_train(lf_mlp* h, int64_t n, const float* x, const float* t, int steps, float lr,
                 float* losses)
{
    if (!h || n <= 0 || !x || !t || steps < 0 || !(lr > 0.f))
        return mset(h, "lf_mlp_train: invalid arguments", LF_EINVAL);
    MT(h, hipSetDevice(h->device));
    int rc = upload(h, n, x, t);
    if (rc) return rc;
    if (steps > h->losses_cap) {
        // the previous buffer may still be read by queued work of an earlier call
        MT(h, hipStreamSynchronize(h->stream));
        if (h->losses) MT(h, hipFree(h->losses));
        h->losses = nullptr;
        h->losses_cap = 0;
        MT(h, hipMalloc((void**)&h->losses, sizeof(float) * steps));
        h->losses_cap = steps;
    }
    for (int k = 0; k < steps; ++k) {
        rc = grad_pass(h, n, h->losses + k);
        if (rc) return rc;
        ++h->step;
        const float b1c = 1.f - std::pow(0.9f, (float)h->step);
        const float b2c = 1.f - std::pow(0.999f, (float)h->step);
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, h->stream, h->P, h->m,
                           h->v, h->grad, lr, b1c, b2c);
        MT(h, hipGetLastError());
        if ((k + 1) % kTrainSyncSteps == 0) MT(h, hipStreamSynchronize(h->stream));
    }
    if (losses && steps)
        MT(h, hipMemcpyAsync(losses, h->losses, sizeof(float) * steps, hipMemcpyDeviceToHost, h->stream));
    MT(h, hipStreamSynchronize(h->stream));
    return LF_OK;
}

}  // extern "C"
