// Spectral neural operator for the warm start (include/pmc_surrogate.h):
// a BF16 matrix-core GEMM with FP32 accumulation, elementwise kernels, and the
// forward / backward / Adam passes composed from them on one HIP stream.
#include "pmc_surrogate.h"
#include "pmc_fno.h"
#include "pmc_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace pmc_nn_dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

struct GemmArgs {
    int m, n, k;
    const float* a;
    int lda;
    long long sa;
    int ta;
    const float* b;
    int ldb;
    long long sb;
    int tb;
    float* c;
    int ldc;
    long long sc;
    float alpha, beta;
    int splits;  // workgroups along k per tile (blockIdx.z = batch * splits + split)
    int kchunk;  // k extent of one split, a multiple of kTK
    float* ws;   // splits > 1: partial products [batch][splits][m][n]
};

constexpr int kTM = 64, kTN = 64, kTK = 64, kPad = 8;

// C = alpha op(A) op(B) + beta C for one 64x64 tile per workgroup.  Four
// waves each own a 32x32 quadrant as 2x2 MFMA tiles of gfx950's
// v_mfma_f32_16x16x32_bf16.  Operands are staged through LDS as BF16 with k
// contiguous (rows padded to 72 elements, so each lane's 8-element fragment is
// one aligned 16-byte read), 64 k per stage (16 loads in flight per operand and
// thread):
//   A fragment of lane l: row l%16, k = 8(l/16)..+7;  B: column l%16, same k;
//   result lane l holds rows 4(l/16)..+3 of column l%16.
// With splits > 1 each workgroup sums one k range and writes alpha * partial
// to the workspace; k_splitk_reduce adds the splits in order.
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[kTM][kTK + kPad];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[kTN][kTK + kPad];
    const long long bz = blockIdx.z / g.splits;
    const int split = static_cast<int>(blockIdx.z % g.splits);
    const int kbeg = split * g.kchunk, kend = min(g.k, kbeg + g.kchunk);
    const float* A = g.a + bz * g.sa;
    const float* B = g.b + bz * g.sb;
    float* C = g.c + bz * g.sc;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += kTK) {
#pragma unroll
        for (int r = 0; r < (kTM * kTK) / 256; ++r) {
            const int e = t + 256 * r;
            int mi, ki;
            if (g.ta) { mi = e % kTM; ki = e / kTM; }   // consecutive threads along m (contiguous for A^T)
            else { ki = e % kTK; mi = e / kTK; }       // consecutive threads along k
            const int gm = m0 + mi, gk = k0 + ki;
            float v = 0.f;
            if (gm < g.m && gk < kend) v = g.ta ? A[static_cast<long long>(gk) * g.lda + gm] : A[static_cast<long long>(gm) * g.lda + gk];
            As[mi][ki] = to_bf16(v);
        }
#pragma unroll
        for (int r = 0; r < (kTN * kTK) / 256; ++r) {
            const int e = t + 256 * r;
            int ni, ki;
            if (g.tb) { ki = e % kTK; ni = e / kTK; }   // B^T: contiguous along k
            else { ni = e % kTN; ki = e / kTN; }       // B: contiguous along n
            const int gn = n0 + ni, gk = k0 + ki;
            float v = 0.f;
            if (gn < g.n && gk < kend) v = g.tb ? B[static_cast<long long>(gn) * g.ldb + gk] : B[static_cast<long long>(gk) * g.ldb + gn];
            Bs[ni][ki] = to_bf16(v);
        }
        __syncthreads();
        const int kq = 8 * (lane >> 4), r16 = lane & 15;
#pragma unroll
        for (int ks = 0; ks < kTK; ks += 32) {
            bf16x8 af[2], bfr[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + r16][ks + kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + r16][ks + kq]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    const int rq = 4 * (lane >> 4), cc = lane & 15;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = m0 + wm + 16 * i + rq + e, col = n0 + wn + 16 * j + cc;
                if (row < g.m && col < g.n) {
                    const float v = g.alpha * acc[i][j][e];
                    if (g.splits > 1) {
                        g.ws[(static_cast<long long>(blockIdx.z) * g.m + row) * g.n + col] = v;
                    } else {
                        float* p = C + static_cast<long long>(row) * g.ldc + col;
                        *p = g.beta == 0.f ? v : v + g.beta * *p;
                    }
                }
            }
}

// C[b] = sum over splits (in order) of the partials + beta C[b].
__global__ void k_splitk_reduce(const float* ws, int splits, int m, int n, long long total, float* c, int ldc,
                                long long sc, float beta) {
    const long long mn = static_cast<long long>(m) * n;
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long b = e / mn, rc = e % mn;
        const float* p = ws + b * splits * mn + rc;
        float v = 0.f;
        for (int s = 0; s < splits; ++s) v += p[s * mn];
        float* q = c + b * sc + (rc / n) * ldc + rc % n;
        *q = beta == 0.f ? v : v + beta * *q;
    }
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
__device__ __forceinline__ float gelu(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * 0.7978845608028654f * (1.f + 3.f * 0.044715f * x * x);
}

// pre[r][c] += bias[c]; out = act ? gelu(pre) : pre.
__global__ void k_bias_act(float* pre, const float* bias, float* out, long long rows, int cols, int act) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const float v = pre[i] + bias[i % cols];
    pre[i] = v;
    out[i] = act ? gelu(v) : v;
}

// d *= gelu'(pre)
__global__ void k_gelu_back(float* d, const float* pre, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] *= gelu_grad(pre[i]);
}

// grad[c] += sum over rows of d[r][c] (cols <= blockDim.x = 256).  A block sums
// a contiguous row range; thread t takes column t % cols of every
// (blockDim.x / cols)-th row, and the row groups meet in LDS.
__global__ void k_colsum(const float* d, long long rows, int cols, float* grad) {
    __shared__ float part[256];
    const int groups = static_cast<int>(blockDim.x) / cols;
    const int c = static_cast<int>(threadIdx.x) % cols, grp = static_cast<int>(threadIdx.x) / cols;
    const long long per = (rows + gridDim.x - 1) / gridDim.x;
    const long long r0 = blockIdx.x * per, r1 = std::min(rows, r0 + per);
    float s = 0.f;
    if (grp < groups)
        for (long long r = r0 + grp; r < r1; r += groups) s += d[r * cols + c];
    part[threadIdx.x] = s;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cols) {
        float t = 0.f;
        for (int k = 0; k < groups; ++k) t += part[k * cols + threadIdx.x];
        atomicAdd(&grad[threadIdx.x], t);
    }
}

// dy = 2 (y - t) / n and the summed squared error.
__global__ void k_mse(const float* y, const float* t, float* dy, long long n, float* sse) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        const float r = y[i] - t[i];
        dy[i] = 2.f * r / static_cast<float>(n);
        e = r * r;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sse, e);
}

// Adam with bias correction.
__global__ void k_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float c1, float c2) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Rows of a device data set by index: dst[j] = src[idx[j]] (row length len).
__global__ void k_gather(const float* src, const int* idx, float* dst, int rows, long long len) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * len) return;
    const long long r = i / len, c = i % len;
    dst[i] = src[static_cast<long long>(idx[r]) * len + c];
}

} // namespace pmc_nn_dev

using namespace pmc_nn_dev;

struct pmc_nn {
    pmc_nn_shape sh{};
    int S = 0, M = 0, W = 0, H = 0, L = 0, Cin = 0;
    int64_t nparam = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    hipError_t err = hipSuccess;
    // parameter offsets
    int64_t oP = 0, obp = 0, oQ1 = 0, ob1 = 0, oQ2 = 0, ob2 = 0;
    std::vector<int64_t> oR, oW, ob;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    float *D = nullptr, *G = nullptr;
    // activations for max_batch
    float *x = nullptr, *y = nullptr, *t = nullptr, *dy = nullptr, *sse = nullptr;
    std::vector<float*> h, pre;          // h[0..L], pre[0..L-1]
    float *Y = nullptr, *Z = nullptr;    // per layer spectral coefficients (kept for backward)
    std::vector<float*> Ys, Zs;
    float *preq = nullptr, *aq = nullptr;
    float *dh = nullptr, *dh2 = nullptr, *dpre = nullptr, *dZ = nullptr, *dY = nullptr, *daq = nullptr;
    // device data set for training
    float *ds_in = nullptr, *ds_t = nullptr;
    int* idx = nullptr;
    int ds_n = 0;
    int64_t adam_t = 0;
    float* ws = nullptr;  // split-k partial products
    size_t ws_cap = 0;
    // GEMM accounting: algorithmic flops (2 m n k per product) of every launch
    // and, with timing on, their HIP-event durations on the operator's stream.
    double gemm_flops = 0.0, gemm_ms = 0.0;
    int64_t gemm_launches = 0;
    bool timing = false;
    static constexpr int kTimed = 64;
    hipEvent_t ev[2 * kTimed] = {};
    int ev_used = 0;

    // Adds the durations of the recorded launches (waits for the last one).
    void harvest() {
        if (ev_used == 0) return;
        note(hipEventSynchronize(ev[2 * ev_used - 1]));
        for (int i = 0; i < ev_used; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) == hipSuccess) gemm_ms += ms;
        }
        ev_used = 0;
    }

    void note(hipError_t e) {
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
            note(hipErrorOutOfMemory);
            return nullptr;
        }
        allocs.push_back(p);
        return static_cast<T*>(p);
    }
};

namespace {

// Launch plan of one product.  Split k when the tile grid leaves most of the
// 256 CUs idle: every stage of a workgroup is one exposed global-memory round
// trip at these sizes, so aim at about 512 workgroups of at most two stages
// (128 k) each.  ws_floats: split-k partial products the plan needs.
struct GemmPlan {
    int splits = 1, kchunk = 1;
    size_t ws_floats = 0;
};

GemmPlan plan_gemm(int m, int n, int k, int batch) {
    GemmPlan p;
    const long long tiles = static_cast<long long>((n + kTN - 1) / kTN) * ((m + kTM - 1) / kTM) * batch;
    p.kchunk = std::max(k, 1);
    if (tiles < 256 && k > 2 * kTK) {
        p.splits = static_cast<int>(std::min<long long>((k + 2 * kTK - 1) / (2 * kTK), (512 + tiles - 1) / tiles));
        if (p.splits > 1) {
            p.kchunk = ((k + p.splits - 1) / p.splits + kTK - 1) / kTK * kTK;
            p.splits = (k + p.kchunk - 1) / p.kchunk;
        }
    }
    if (p.splits > 1) p.ws_floats = static_cast<size_t>(p.splits) * batch * m * n;
    return p;
}

// Enqueues the product on `stream` with the plan's partial products in ws.
void launch_gemm(const GemmPlan& p, int m, int n, int k, int batch, float alpha, const float* a, int lda, long long sa,
                 int ta, const float* b, int ldb, long long sb, int tb, float beta, float* c, int ldc, long long sc,
                 float* ws, hipStream_t stream) {
    GemmArgs g{m, n, k, a, lda, sa, ta, b, ldb, sb, tb, c, ldc, sc, alpha, beta, p.splits, p.kchunk, ws};
    dim3 grid((n + kTN - 1) / kTN, (m + kTM - 1) / kTM, static_cast<unsigned>(batch) * p.splits);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, stream, g);
    if (p.splits > 1) {
        const long long total = static_cast<long long>(batch) * m * n;
        const unsigned nb = static_cast<unsigned>(std::min<long long>((total + 255) / 256, 2048));
        hipLaunchKernelGGL(k_splitk_reduce, dim3(nb), dim3(256), 0, stream, ws, p.splits, m, n, total, c, ldc, sc,
                           beta);
    }
}

void gemm(pmc_nn* nn, int m, int n, int k, int batch, float alpha, const float* a, int lda, long long sa, int ta,
          const float* b, int ldb, long long sb, int tb, float beta, float* c, int ldc, long long sc) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    const GemmPlan plan = plan_gemm(m, n, k, batch);
    float* ws = nullptr;
    if (plan.splits > 1) {
        if (plan.ws_floats > nn->ws_cap) {
            nn->ws = nn->alloc<float>(plan.ws_floats);
            nn->ws_cap = nn->ws ? plan.ws_floats : 0;
        }
        ws = nn->ws;
        if (!ws) return;
    }
    nn->gemm_flops += 2.0 * m * n * static_cast<double>(k) * batch;
    ++nn->gemm_launches;
    const bool timed = nn->timing && nn->ev[0] != nullptr;
    if (timed) {
        if (nn->ev_used == pmc_nn::kTimed) nn->harvest();
        nn->note(hipEventRecord(nn->ev[2 * nn->ev_used], nn->stream));
    }
    launch_gemm(plan, m, n, k, batch, alpha, a, lda, sa, ta, b, ldb, sb, tb, beta, c, ldc, sc, ws, nn->stream);
    if (timed) nn->note(hipEventRecord(nn->ev[2 * nn->ev_used++ + 1], nn->stream));
}

unsigned blocks(long long n, int t = 256) { return static_cast<unsigned>((n + t - 1) / t); }

// Forward pass of `b` samples already in nn->x; leaves the prediction in nn->y.
void forward(pmc_nn* nn, int b) {
    const int S = nn->S, W = nn->W, M2 = 2 * nn->M, H = nn->H, L = nn->L;
    const long long R = static_cast<long long>(b) * S;
    float* P = nn->params;
    gemm(nn, static_cast<int>(R), W, nn->Cin, 1, 1.f, nn->x, nn->Cin, 0, 0, P + nn->oP, W, 0, 0, 0.f, nn->h[0], W, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * W)), dim3(256), 0, nn->stream, nn->h[0], P + nn->obp, nn->h[0], R, W, 0);
    for (int l = 0; l < L; ++l) {
        float* Y = nn->Ys[l];
        float* Z = nn->Zs[l];
        // Y_b = D^T h_b  [2M x W]
        gemm(nn, M2, W, S, b, 1.f, nn->D, M2, 0, 1, nn->h[l], W, static_cast<long long>(S) * W, 0, 0.f, Y, W,
             static_cast<long long>(M2) * W);
        // Z_m = Y_m R_m over the batch rows of each mode (rows b at stride 2M*W)
        gemm(nn, b, 2 * W, 2 * W, nn->M, 1.f, Y, M2 * W, 2 * W, 0, P + nn->oR[l], 2 * W, 4LL * W * W, 0, 0.f, Z,
             M2 * W, 2 * W);
        // pre_b = G Z_b + h_b W_l
        gemm(nn, S, W, M2, b, 1.f, nn->G, M2, 0, 0, Z, W, static_cast<long long>(M2) * W, 0, 0.f, nn->pre[l], W,
             static_cast<long long>(S) * W);
        gemm(nn, static_cast<int>(R), W, W, 1, 1.f, nn->h[l], W, 0, 0, P + nn->oW[l], W, 0, 0, 1.f, nn->pre[l], W, 0);
        hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * W)), dim3(256), 0, nn->stream, nn->pre[l], P + nn->ob[l],
                           nn->h[l + 1], R, W, l < L - 1 ? 1 : 0);
    }
    gemm(nn, static_cast<int>(R), H, W, 1, 1.f, nn->h[L], W, 0, 0, P + nn->oQ1, H, 0, 0, 0.f, nn->preq, H, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * H)), dim3(256), 0, nn->stream, nn->preq, P + nn->ob1, nn->aq, R, H, 1);
    gemm(nn, static_cast<int>(R), 1, H, 1, 1.f, nn->aq, H, 0, 0, P + nn->oQ2, 1, 0, 0, 0.f, nn->y, 1, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R)), dim3(256), 0, nn->stream, nn->y, P + nn->ob2, nn->y, R, 1, 0);
}

// Loss and gradient of `b` samples (x, t on the device): grads overwritten.
void backward(pmc_nn* nn, int b) {
    const int S = nn->S, W = nn->W, M2 = 2 * nn->M, H = nn->H, L = nn->L;
    const long long R = static_cast<long long>(b) * S;
    float* P = nn->params;
    float* Gd = nn->grads;
    nn->note(hipMemsetAsync(Gd, 0, nn->nparam * sizeof(float), nn->stream));
    nn->note(hipMemsetAsync(nn->sse, 0, sizeof(float), nn->stream));
    hipLaunchKernelGGL(k_mse, dim3(blocks(R)), dim3(256), 0, nn->stream, nn->y, nn->t, nn->dy, R, nn->sse);
    // output layer
    gemm(nn, H, 1, static_cast<int>(R), 1, 1.f, nn->aq, H, 0, 1, nn->dy, 1, 0, 0, 0.f, Gd + nn->oQ2, 1, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, nn->stream, nn->dy, R, 1, Gd + nn->ob2);
    gemm(nn, static_cast<int>(R), H, 1, 1, 1.f, nn->dy, 1, 0, 0, P + nn->oQ2, 1, 0, 1, 0.f, nn->daq, H, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3(blocks(R * H)), dim3(256), 0, nn->stream, nn->daq, nn->preq, R * H);
    gemm(nn, W, H, static_cast<int>(R), 1, 1.f, nn->h[L], W, 0, 1, nn->daq, H, 0, 0, 0.f, Gd + nn->oQ1, H, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, nn->stream, nn->daq, R, H, Gd + nn->ob1);
    gemm(nn, static_cast<int>(R), W, H, 1, 1.f, nn->daq, H, 0, 0, P + nn->oQ1, H, 0, 1, 0.f, nn->dh, W, 0);
    for (int l = L - 1; l >= 0; --l) {
        float* dpre = nn->dpre;
        nn->note(hipMemcpyAsync(dpre, nn->dh, R * W * sizeof(float), hipMemcpyDeviceToDevice, nn->stream));
        if (l < L - 1)
            hipLaunchKernelGGL(k_gelu_back, dim3(blocks(R * W)), dim3(256), 0, nn->stream, dpre, nn->pre[l], R * W);
        hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, nn->stream, dpre, R, W, Gd + nn->ob[l]);
        // skip path
        gemm(nn, W, W, static_cast<int>(R), 1, 1.f, nn->h[l], W, 0, 1, dpre, W, 0, 0, 0.f, Gd + nn->oW[l], W, 0);
        gemm(nn, static_cast<int>(R), W, W, 1, 1.f, dpre, W, 0, 0, P + nn->oW[l], W, 0, 1, 0.f, nn->dh2, W, 0);
        // spectral path: dZ_b = G^T dpre_b; dR_m = Y_m^T dZ_m; dY_m = dZ_m R_m^T; dh += D dY_b
        gemm(nn, M2, W, S, b, 1.f, nn->G, M2, 0, 1, dpre, W, static_cast<long long>(S) * W, 0, 0.f, nn->dZ, W,
             static_cast<long long>(M2) * W);
        gemm(nn, 2 * W, 2 * W, b, nn->M, 1.f, nn->Ys[l], M2 * W, 2 * W, 1, nn->dZ, M2 * W, 2 * W, 0, 0.f,
             Gd + nn->oR[l], 2 * W, 4LL * W * W);
        gemm(nn, b, 2 * W, 2 * W, nn->M, 1.f, nn->dZ, M2 * W, 2 * W, 0, P + nn->oR[l], 2 * W, 4LL * W * W, 1, 0.f,
             nn->dY, M2 * W, 2 * W);
        gemm(nn, S, W, M2, b, 1.f, nn->D, M2, 0, 0, nn->dY, W, static_cast<long long>(M2) * W, 0, 1.f, nn->dh2, W,
             static_cast<long long>(S) * W);
        std::swap(nn->dh, nn->dh2);
    }
    gemm(nn, nn->Cin, W, static_cast<int>(R), 1, 1.f, nn->x, nn->Cin, 0, 1, nn->dh, W, 0, 0, 0.f, Gd + nn->oP, W, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, nn->stream, nn->dh, R, W, Gd + nn->obp);
}

int finish(pmc_nn* nn, const char* /*what*/) {
    nn->harvest();
    nn->note(hipStreamSynchronize(nn->stream));
    nn->note(hipGetLastError());
    return nn->err == hipSuccess ? PMC_OK : PMC_ERR_DEVICE;
}

// splitmix64 for parameter initialization and shuffles.
uint64_t mix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return static_cast<double>(mix(s) >> 11) * 0x1.0p-53; }

} // namespace

extern "C" {

int pmc_nn_create(const pmc_nn_shape* sh, pmc_nn** out) {
    if (!sh || !out || sh->nx <= 0 || sh->nz <= 0 || sh->in_channels <= 0 || sh->width <= 0 || sh->width > 128 ||
        sh->layers <= 0 || sh->modes_x <= 0 || sh->modes_z <= 0 || sh->modes_x > sh->nx || sh->modes_z > sh->nz ||
        sh->hidden <= 0 || sh->hidden > 256 || sh->max_batch <= 0)
        return PMC_ERR_ARGUMENT;
    auto* nn = new pmc_nn();
    nn->sh = *sh;
    nn->S = sh->nx * sh->nz;
    nn->M = (2 * sh->modes_x - 1) * sh->modes_z;
    nn->W = sh->width;
    nn->H = sh->hidden;
    nn->L = sh->layers;
    nn->Cin = sh->in_channels;
    if (hipStreamCreateWithFlags(&nn->stream, hipStreamNonBlocking) != hipSuccess) {
        delete nn;
        return PMC_ERR_DEVICE;
    }
    const int W = nn->W, H = nn->H, M2 = 2 * nn->M, S = nn->S;
    int64_t o = 0;
    nn->oP = o; o += static_cast<int64_t>(nn->Cin) * W;
    nn->obp = o; o += W;
    for (int l = 0; l < nn->L; ++l) {
        nn->oR.push_back(o); o += static_cast<int64_t>(nn->M) * 4 * W * W;
        nn->oW.push_back(o); o += static_cast<int64_t>(W) * W;
        nn->ob.push_back(o); o += W;
    }
    nn->oQ1 = o; o += static_cast<int64_t>(W) * H;
    nn->ob1 = o; o += H;
    nn->oQ2 = o; o += H;
    nn->ob2 = o; o += 1;
    nn->nparam = o;
    // Initialization: uniform(-a, a) with a = sqrt(3 / fan_in); zero biases.
    std::vector<float> p(o, 0.f);
    uint64_t seed = sh->seed;
    auto fill = [&](int64_t off, int64_t n, double fan_in) {
        const double a = std::sqrt(3.0 / fan_in);
        for (int64_t i = 0; i < n; ++i) p[off + i] = static_cast<float>((2.0 * uniform(seed) - 1.0) * a);
    };
    fill(nn->oP, static_cast<int64_t>(nn->Cin) * W, nn->Cin);
    for (int l = 0; l < nn->L; ++l) {
        fill(nn->oR[l], static_cast<int64_t>(nn->M) * 4 * W * W, 2.0 * W * nn->M);
        fill(nn->oW[l], static_cast<int64_t>(W) * W, W);
    }
    fill(nn->oQ1, static_cast<int64_t>(W) * H, W);
    fill(nn->oQ2, H, H);
    // Fourier basis of the kept modes.
    std::vector<float> d(static_cast<size_t>(S) * M2), gm(static_cast<size_t>(S) * M2);
    const double two_pi = 6.283185307179586476925286766559;
    for (int ix = 0; ix < sh->nx; ++ix)
        for (int iz = 0; iz < sh->nz; ++iz) {
            const int s = ix * sh->nz + iz;
            int m = 0;
            for (int kx = -(sh->modes_x - 1); kx < sh->modes_x; ++kx)
                for (int kz = 0; kz < sh->modes_z; ++kz, ++m) {
                    const double th = two_pi * (static_cast<double>(kx) * ix / sh->nx + static_cast<double>(kz) * iz / sh->nz);
                    d[static_cast<size_t>(s) * M2 + 2 * m] = static_cast<float>(std::cos(th));
                    d[static_cast<size_t>(s) * M2 + 2 * m + 1] = static_cast<float>(std::sin(th));
                }
        }
    for (size_t i = 0; i < d.size(); ++i) gm[i] = d[i] / static_cast<float>(S);
    const size_t B = sh->max_batch;
    const size_t R = B * S;
    nn->params = nn->alloc<float>(o);
    nn->grads = nn->alloc<float>(o);
    nn->adam_m = nn->alloc<float>(o);
    nn->adam_v = nn->alloc<float>(o);
    nn->D = nn->alloc<float>(d.size());
    nn->G = nn->alloc<float>(d.size());
    nn->x = nn->alloc<float>(R * nn->Cin);
    nn->y = nn->alloc<float>(R);
    nn->t = nn->alloc<float>(R);
    nn->dy = nn->alloc<float>(R);
    nn->sse = nn->alloc<float>(1);
    for (int l = 0; l <= nn->L; ++l) nn->h.push_back(nn->alloc<float>(R * W));
    for (int l = 0; l < nn->L; ++l) {
        nn->pre.push_back(nn->alloc<float>(R * W));
        nn->Ys.push_back(nn->alloc<float>(B * M2 * W));
        nn->Zs.push_back(nn->alloc<float>(B * M2 * W));
    }
    nn->preq = nn->alloc<float>(R * H);
    nn->aq = nn->alloc<float>(R * H);
    nn->daq = nn->alloc<float>(R * H);
    nn->dh = nn->alloc<float>(R * W);
    nn->dh2 = nn->alloc<float>(R * W);
    nn->dpre = nn->alloc<float>(R * W);
    nn->dZ = nn->alloc<float>(B * M2 * W);
    nn->dY = nn->alloc<float>(B * M2 * W);
    nn->idx = nn->alloc<int>(B);
    if (nn->err != hipSuccess) {
        pmc_nn_destroy(nn);
        return PMC_ERR_DEVICE;
    }
    nn->note(hipMemcpy(nn->params, p.data(), o * sizeof(float), hipMemcpyHostToDevice));
    nn->note(hipMemset(nn->adam_m, 0, o * sizeof(float)));
    nn->note(hipMemset(nn->adam_v, 0, o * sizeof(float)));
    nn->note(hipMemcpy(nn->D, d.data(), d.size() * sizeof(float), hipMemcpyHostToDevice));
    nn->note(hipMemcpy(nn->G, gm.data(), gm.size() * sizeof(float), hipMemcpyHostToDevice));
    if (nn->err != hipSuccess) {
        pmc_nn_destroy(nn);
        return PMC_ERR_DEVICE;
    }
    *out = nn;
    return PMC_OK;
}

int64_t pmc_nn_param_count(const pmc_nn* nn) { return nn ? nn->nparam : 0; }

int pmc_nn_get_params(pmc_nn* nn, float* params) {
    if (!nn || !params) return PMC_ERR_ARGUMENT;
    nn->note(hipMemcpyAsync(params, nn->params, nn->nparam * sizeof(float), hipMemcpyDeviceToHost, nn->stream));
    return finish(nn, "get params");
}

int pmc_nn_set_params(pmc_nn* nn, const float* params) {
    if (!nn || !params) return PMC_ERR_ARGUMENT;
    nn->note(hipMemcpyAsync(nn->params, params, nn->nparam * sizeof(float), hipMemcpyHostToDevice, nn->stream));
    return finish(nn, "set params");
}

int pmc_nn_basis(pmc_nn* nn, float* d) {
    if (!nn || !d) return PMC_ERR_ARGUMENT;
    nn->note(hipMemcpyAsync(d, nn->D, static_cast<size_t>(nn->S) * 2 * nn->M * sizeof(float), hipMemcpyDeviceToHost,
                            nn->stream));
    return finish(nn, "basis");
}

int pmc_nn_predict(pmc_nn* nn, int32_t batch, const float* in, float* out) {
    if (!nn || !in || !out || batch <= 0) return PMC_ERR_ARGUMENT;
    const int S = nn->S;
    for (int b0 = 0; b0 < batch; b0 += nn->sh.max_batch) {
        const int b = std::min(nn->sh.max_batch, batch - b0);
        nn->note(hipMemcpyAsync(nn->x, in + static_cast<size_t>(b0) * S * nn->Cin,
                                static_cast<size_t>(b) * S * nn->Cin * sizeof(float), hipMemcpyHostToDevice, nn->stream));
        forward(nn, b);
        nn->note(hipMemcpyAsync(out + static_cast<size_t>(b0) * S, nn->y, static_cast<size_t>(b) * S * sizeof(float),
                                hipMemcpyDeviceToHost, nn->stream));
    }
    return finish(nn, "predict");
}

int pmc_nn_loss_grad(pmc_nn* nn, int32_t batch, const float* in, const float* target, float* loss, float* grad) {
    if (!nn || !in || !target || !loss || batch <= 0 || batch > nn->sh.max_batch) return PMC_ERR_ARGUMENT;
    const size_t R = static_cast<size_t>(batch) * nn->S;
    nn->note(hipMemcpyAsync(nn->x, in, R * nn->Cin * sizeof(float), hipMemcpyHostToDevice, nn->stream));
    nn->note(hipMemcpyAsync(nn->t, target, R * sizeof(float), hipMemcpyHostToDevice, nn->stream));
    forward(nn, batch);
    backward(nn, batch);
    float sse = 0.f;
    nn->note(hipMemcpyAsync(&sse, nn->sse, sizeof(float), hipMemcpyDeviceToHost, nn->stream));
    if (grad) nn->note(hipMemcpyAsync(grad, nn->grads, nn->nparam * sizeof(float), hipMemcpyDeviceToHost, nn->stream));
    const int rc = finish(nn, "loss/grad");
    *loss = sse / static_cast<float>(R);
    return rc;
}

int pmc_nn_train(pmc_nn* nn, int32_t n, const float* in, const float* target, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float* losses) {
    if (!nn || !in || !target || n <= 0 || epochs < 0 || batch <= 0 || batch > nn->sh.max_batch) return PMC_ERR_ARGUMENT;
    const int S = nn->S;
    const size_t in_row = static_cast<size_t>(S) * nn->Cin;
    if (nn->ds_n != n) {
        nn->ds_in = nn->alloc<float>(static_cast<size_t>(n) * in_row);
        nn->ds_t = nn->alloc<float>(static_cast<size_t>(n) * S);
        nn->ds_n = n;
        if (nn->err != hipSuccess) return PMC_ERR_DEVICE;
    }
    nn->note(hipMemcpyAsync(nn->ds_in, in, static_cast<size_t>(n) * in_row * sizeof(float), hipMemcpyHostToDevice,
                            nn->stream));
    nn->note(hipMemcpyAsync(nn->ds_t, target, static_cast<size_t>(n) * S * sizeof(float), hipMemcpyHostToDevice,
                            nn->stream));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int e = 0; e < epochs; ++e) {
        for (int i = n - 1; i > 0; --i) std::swap(order[i], order[mix(s) % static_cast<uint64_t>(i + 1)]);
        double total = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch, ++nb) {
            nn->note(hipMemcpyAsync(nn->idx, order.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, nn->stream));
            hipLaunchKernelGGL(k_gather, dim3(blocks(static_cast<long long>(batch) * in_row)), dim3(256), 0, nn->stream,
                               nn->ds_in, nn->idx, nn->x, batch, static_cast<long long>(in_row));
            hipLaunchKernelGGL(k_gather, dim3(blocks(static_cast<long long>(batch) * S)), dim3(256), 0, nn->stream,
                               nn->ds_t, nn->idx, nn->t, batch, static_cast<long long>(S));
            forward(nn, batch);
            backward(nn, batch);
            ++nn->adam_t;
            const float c1 = 1.f - std::pow(b1, static_cast<float>(nn->adam_t));
            const float c2 = 1.f - std::pow(b2, static_cast<float>(nn->adam_t));
            hipLaunchKernelGGL(k_adam, dim3(blocks(nn->nparam)), dim3(256), 0, nn->stream, nn->params, nn->grads,
                               nn->adam_m, nn->adam_v, static_cast<long long>(nn->nparam), lr, b1, b2, c1, c2);
            float sse = 0.f;
            nn->note(hipMemcpyAsync(&sse, nn->sse, sizeof(float), hipMemcpyDeviceToHost, nn->stream));
            nn->note(hipStreamSynchronize(nn->stream));
            total += sse / (static_cast<double>(batch) * S);
        }
        if (losses) losses[e] = static_cast<float>(total / std::max(1, nb));
    }
    return finish(nn, "train");
}

int pmc_nn_set_timing(pmc_nn* nn, int32_t on) {
    if (!nn) return PMC_ERR_ARGUMENT;
    if (on && nn->ev[0] == nullptr)
        for (auto& e : nn->ev)
            if (hipEventCreate(&e) != hipSuccess) return PMC_ERR_DEVICE;
    nn->harvest();
    nn->timing = on != 0;
    return PMC_OK;
}

int pmc_nn_gemm_stats(pmc_nn* nn, double* flops, int64_t* launches, double* ms, int32_t reset) {
    if (!nn) return PMC_ERR_ARGUMENT;
    nn->harvest();
    nn->note(hipStreamSynchronize(nn->stream));
    if (flops) *flops = nn->gemm_flops;
    if (launches) *launches = nn->gemm_launches;
    if (ms) *ms = nn->gemm_ms;
    if (reset) {
        nn->gemm_flops = 0.0;
        nn->gemm_ms = 0.0;
        nn->gemm_launches = 0;
    }
    return nn->err == hipSuccess ? PMC_OK : PMC_ERR_DEVICE;
}

void pmc_nn_destroy(pmc_nn* nn) {
    if (!nn) return;
    if (nn->stream) (void)hipStreamSynchronize(nn->stream);
    for (auto& e : nn->ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : nn->allocs) (void)hipFree(p);
    if (nn->stream) (void)hipStreamDestroy(nn->stream);
    delete nn;
}

int pmc_nn_gemm(int32_t m, int32_t n, int32_t k, int32_t batch, float alpha, const float* a, int32_t lda,
                int64_t stride_a, int32_t ta, const float* b, int32_t ldb, int64_t stride_b, int32_t tb, float beta,
                float* c, int32_t ldc, int64_t stride_c) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0 || !a || !b || !c) return PMC_ERR_ARGUMENT;
    pmc_nn tmp;
    if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess) return PMC_ERR_DEVICE;
    // Extents of each operand (last element + 1).
    const size_t na = static_cast<size_t>(stride_a) * (batch - 1) + (ta ? static_cast<size_t>(k - 1) * lda + m
                                                                         : static_cast<size_t>(m - 1) * lda + k);
    const size_t nb = static_cast<size_t>(stride_b) * (batch - 1) + (tb ? static_cast<size_t>(n - 1) * ldb + k
                                                                         : static_cast<size_t>(k - 1) * ldb + n);
    const size_t nc = static_cast<size_t>(stride_c) * (batch - 1) + static_cast<size_t>(m - 1) * ldc + n;
    float* da = tmp.alloc<float>(na);
    float* db = tmp.alloc<float>(nb);
    float* dc = tmp.alloc<float>(nc);
    if (tmp.err == hipSuccess) {
        tmp.note(hipMemcpy(da, a, na * sizeof(float), hipMemcpyHostToDevice));
        tmp.note(hipMemcpy(db, b, nb * sizeof(float), hipMemcpyHostToDevice));
        tmp.note(hipMemcpy(dc, c, nc * sizeof(float), hipMemcpyHostToDevice));
        gemm(&tmp, m, n, k, batch, alpha, da, lda, stride_a, ta, db, ldb, stride_b, tb, beta, dc, ldc, stride_c);
        tmp.note(hipStreamSynchronize(tmp.stream));
        tmp.note(hipGetLastError());
        tmp.note(hipMemcpy(c, dc, nc * sizeof(float), hipMemcpyDeviceToHost));
    }
    const int rc = tmp.err == hipSuccess ? PMC_OK : PMC_ERR_DEVICE;
    for (void* p : tmp.allocs) (void)hipFree(p);
    (void)hipStreamDestroy(tmp.stream);
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------- GEMM without allocation
// Split-k partial products of pmc_nn_gemm_noalloc: one copy per GPU, held for
// the life of the process.  Every partial a plan reads is written first by
// k_gemm of the same call, so nothing needs clearing between calls.
namespace {
static __device__ __align__(256) float scratch[PMC_NN_GEMM_SCRATCH_FLOATS];

constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
float* scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch, looked up once per device with that
// device current (the host-side symbol is only a stand-in for it).
hipError_t device_scratch(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(scratch_lock);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}
} // namespace

extern "C" int pmc_nn_gemm_noalloc(int32_t m, int32_t n, int32_t k, int32_t batch, float alpha, const float* a,
                                   int32_t lda, int64_t stride_a, int32_t ta, const float* b, int32_t ldb,
                                   int64_t stride_b, int32_t tb, float beta, float* c, int32_t ldc, int64_t stride_c,
                                   void* stream) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0 || !a || !b || !c) return PMC_ERR_ARGUMENT;
    const GemmPlan plan = plan_gemm(m, n, k, batch);
    if (plan.ws_floats > PMC_NN_GEMM_SCRATCH_FLOATS) return PMC_ERR_ARGUMENT; // partials exceed the fixed buffer
    float* ws = nullptr;
    if (plan.splits > 1 && device_scratch(&ws) != hipSuccess) return PMC_ERR_DEVICE;
    launch_gemm(plan, m, n, k, batch, alpha, a, lda, stride_a, ta, b, ldb, stride_b, tb, beta, c, ldc, stride_c, ws,
                static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? PMC_OK : PMC_ERR_DEVICE;
}

// ---------------------------------------------------------------- 3-D spectral operator
// include/pmc_fno.h.  Every product goes through the same BF16 MFMA GEMM; the
// separable truncated Fourier transform is three axis contractions per
// direction (cos/sin matrices built on the host), the per-mode complex mixing
// one GEMM batched over modes.  Intermediate layouts (channels last, pair index
// p/q/r/s = real/imaginary part):
//   h [B][X][Y][Z][W] -> T1 [B][X][Y][P][Kz][W] -> T2 [B][X][Q][Ky][Kz][W]
//   -> F [B][R][Kx][Ky][Kz][W] = [B][R][M][W] -> Fm [M][B][R][W] (mixing)
//   -> G [B][S][M][W] -> U2 [B][X][Q][Ky][Kz][W] -> U1 [B][X][Y][P][Kz][W] -> [B][X][Y][Z][W].
// Each transform matrix orders its rows so that the pair index lands next to
// the axis the following contraction runs over.
namespace pmc_nn_dev {

// src [b][2][M][W] <-> dst [M][b][2][W] (dir 0 forward, 1 back).
__global__ void k_fno_perm(const float* __restrict__ src, float* __restrict__ dst, int b, int M, int W, int dir) {
    const long long n = static_cast<long long>(b) * 2 * M * W;
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        // i indexes the [b][2][M][W] layout.
        const int c = static_cast<int>(i % W);
        const long long rest = i / W;
        const int m = static_cast<int>(rest % M);
        const long long br = rest / M;
        const int r = static_cast<int>(br % 2);
        const int bb = static_cast<int>(br / 2);
        const long long j = ((static_cast<long long>(m) * b + bb) * 2 + r) * W + c;
        if (dir == 0) dst[j] = src[i];
        else dst[i] = src[j];
    }
}

// Complex weights R [M][2][W][W] -> real block maps B [M][2W][2W] acting on
// (re, im) rows: [[Wr, Wi], [-Wi, Wr]].
__global__ void k_fno_expand(const float* __restrict__ R, float* __restrict__ B, int M, int W) {
    const long long n = static_cast<long long>(M) * 4 * W * W;
    const int W2 = 2 * W;
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int col = static_cast<int>(i % W2);
        const int row = static_cast<int>((i / W2) % W2);
        const long long m = i / (static_cast<long long>(W2) * W2);
        const int r = row / W, c = row % W, s = col / W, cp = col % W;
        const float* Rm = R + m * 2LL * W * W;
        const float wr = Rm[c * W + cp], wi = Rm[W * W + c * W + cp];
        B[i] = (r == s) ? wr : (r == 0 ? wi : -wi);
    }
}

// Gradient of the block maps folded back onto the complex weights.
__global__ void k_fno_fold(const float* __restrict__ gB, float* __restrict__ gR, int M, int W) {
    const long long n = static_cast<long long>(M) * W * W;
    const int W2 = 2 * W;
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int cp = static_cast<int>(i % W), c = static_cast<int>((i / W) % W);
        const long long m = i / (static_cast<long long>(W) * W);
        const float* g = gB + m * W2 * W2;
        const float g00 = g[c * W2 + cp], g11 = g[(W + c) * W2 + W + cp];
        const float g01 = g[c * W2 + W + cp], g10 = g[(W + c) * W2 + cp];
        gR[m * 2LL * W * W + c * W + cp] = g00 + g11;
        gR[m * 2LL * W * W + W * W + c * W + cp] = g01 - g10;
    }
}

// Weighted squared error: dy = 2 w (y - t) / wsum; sse += w (y - t)^2.
__global__ void k_wmse(const float* y, const float* t, const float* w, float* dy, long long n, float wsum, float* sse) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        const float r = y[i] - t[i];
        dy[i] = 2.f * w[i] * r / wsum;
        e = w[i] * r * r;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sse, e);
}

// v[0 .. n) = 0 (a fill kernel rather than a memset node).
__global__ void k_zero(float* v, long long n) {
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        v[i] = 0.f;
}

// Sum of a device array into out[0] (one block).
__global__ void k_sum(const float* v, long long n, float* out) {
    __shared__ float part[256];
    float s = 0.f;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (static_cast<int>(threadIdx.x) < k) part[threadIdx.x] += part[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

} // namespace pmc_nn_dev

struct pmc_fno {
    pmc_nn eng; // stream, allocations, GEMM workspace and accounting
    pmc_fno_shape sh{};
    int X = 0, Y = 0, Z = 0, S = 0, W = 0, H = 0, L = 0, Cin = 0, m = 0, M = 0;
    int64_t nparam = 0;
    int64_t oP = 0, obp = 0, oQ1 = 0, ob1 = 0, oQ2 = 0, ob2 = 0;
    std::vector<int64_t> oR, oW, ob;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    float *Dz = nullptr, *Dy = nullptr, *Dx = nullptr, *Ez = nullptr, *Ey = nullptr, *Ex = nullptr;
    float *x = nullptr, *y = nullptr, *t = nullptr, *w = nullptr, *dy = nullptr, *sse = nullptr, *wsum = nullptr;
    std::vector<float*> h, pre, Fm, Bl;
    float *T1 = nullptr, *T2 = nullptr, *F = nullptr, *Gm = nullptr, *G = nullptr, *dB = nullptr;
    float *preq = nullptr, *aq = nullptr, *daq = nullptr, *dh = nullptr, *dh2 = nullptr, *dpre = nullptr;
    float *ds_in = nullptr, *ds_t = nullptr, *ds_w = nullptr;
    const float* ds_src[3] = {nullptr, nullptr, nullptr};
    int* idx = nullptr;
    int ds_n = 0;
    int64_t adam_t = 0;
};

namespace {
using pmc_nn_dev::k_fno_perm;

unsigned grid_for(long long n) { return static_cast<unsigned>(std::min<long long>((n + 255) / 256, 8192)); }

// K(hin) into out ([b][S][W]); keeps the mode coefficients of the layer in Fm.
void fno_spectral(pmc_fno* f, int b, const float* hin, float* out, float* Fm, const float* Bl) {
    pmc_nn* E = &f->eng;
    const int X = f->X, Y = f->Y, Z = f->Z, W = f->W, m = f->m, M = f->M;
    const int N = 2 * m * m * W; // (ky, kz, c) row length of T2 / F / G / U2 per (x or kx, pair)
    const long long bxy = static_cast<long long>(b) * X * Y;
    gemm(E, 2 * m, W, Z, static_cast<int>(bxy), 1.f, f->Dz, Z, 0, 0, hin, W, static_cast<long long>(Z) * W, 0, 0.f,
         f->T1, W, 2LL * m * W);
    gemm(E, 4 * m, m * W, 2 * Y, b * X, 1.f, f->Dy, 2 * Y, 0, 0, f->T1, m * W, 2LL * Y * m * W, 0, 0.f, f->T2, m * W,
         4LL * m * m * W);
    gemm(E, 4 * m, N, 2 * X, b, 1.f, f->Dx, 2 * X, 0, 0, f->T2, N, 2LL * X * N, 0, 0.f, f->F, N, 4LL * m * N);
    hipLaunchKernelGGL(k_fno_perm, dim3(grid_for(2LL * b * M * W)), dim3(256), 0, E->stream, f->F, Fm, b, M, W, 0);
    gemm(E, b, 2 * W, 2 * W, M, 1.f, Fm, 2 * W, 2LL * b * W, 0, Bl, 2 * W, 4LL * W * W, 0, 0.f, f->Gm, 2 * W,
         2LL * b * W);
    hipLaunchKernelGGL(k_fno_perm, dim3(grid_for(2LL * b * M * W)), dim3(256), 0, E->stream, f->Gm, f->G, b, M, W, 1);
    gemm(E, 2 * X, N, 4 * m, b, 1.f, f->Ex, 4 * m, 0, 0, f->G, N, 4LL * m * N, 0, 0.f, f->T2, N, 2LL * X * N);
    gemm(E, 2 * Y, m * W, 4 * m, b * X, 1.f, f->Ey, 4 * m, 0, 0, f->T2, m * W, 4LL * m * m * W, 0, 0.f, f->T1, m * W,
         2LL * Y * m * W);
    gemm(E, Z, W, 2 * m, static_cast<int>(bxy), 1.f, f->Ez, 2 * m, 0, 0, f->T1, W, 2LL * m * W, 0, 0.f, out, W,
         static_cast<long long>(Z) * W);
}

// Adjoint of fno_spectral for the output gradient dout: dh += K^T dout and the
// gradient of the layer's complex weights into gR (Fm is overwritten).
void fno_spectral_back(pmc_fno* f, int b, const float* dout, float* Fm, const float* Bl, float* gR, float* dh) {
    pmc_nn* E = &f->eng;
    const int X = f->X, Y = f->Y, Z = f->Z, W = f->W, m = f->m, M = f->M;
    const int N = 2 * m * m * W;
    const long long bxy = static_cast<long long>(b) * X * Y;
    gemm(E, 2 * m, W, Z, static_cast<int>(bxy), 1.f, f->Ez, 2 * m, 0, 1, dout, W, static_cast<long long>(Z) * W, 0, 0.f,
         f->T1, W, 2LL * m * W);
    gemm(E, 4 * m, m * W, 2 * Y, b * X, 1.f, f->Ey, 4 * m, 0, 1, f->T1, m * W, 2LL * Y * m * W, 0, 0.f, f->T2, m * W,
         4LL * m * m * W);
    gemm(E, 4 * m, N, 2 * X, b, 1.f, f->Ex, 4 * m, 0, 1, f->T2, N, 2LL * X * N, 0, 0.f, f->G, N, 4LL * m * N);
    hipLaunchKernelGGL(k_fno_perm, dim3(grid_for(2LL * b * M * W)), dim3(256), 0, E->stream, f->G, f->Gm, b, M, W, 0);
    gemm(E, 2 * W, 2 * W, b, M, 1.f, Fm, 2 * W, 2LL * b * W, 1, f->Gm, 2 * W, 2LL * b * W, 0, 0.f, f->dB, 2 * W,
         4LL * W * W);
    hipLaunchKernelGGL(pmc_nn_dev::k_fno_fold, dim3(grid_for(static_cast<long long>(M) * W * W)), dim3(256), 0,
                       E->stream, f->dB, gR, M, W);
    gemm(E, b, 2 * W, 2 * W, M, 1.f, f->Gm, 2 * W, 2LL * b * W, 0, Bl, 2 * W, 4LL * W * W, 1, 0.f, Fm, 2 * W,
         2LL * b * W);
    hipLaunchKernelGGL(k_fno_perm, dim3(grid_for(2LL * b * M * W)), dim3(256), 0, E->stream, Fm, f->F, b, M, W, 1);
    gemm(E, 2 * X, N, 4 * m, b, 1.f, f->Dx, 2 * X, 0, 1, f->F, N, 4LL * m * N, 0, 0.f, f->T2, N, 2LL * X * N);
    gemm(E, 2 * Y, m * W, 4 * m, b * X, 1.f, f->Dy, 2 * Y, 0, 1, f->T2, m * W, 4LL * m * m * W, 0, 0.f, f->T1, m * W,
         2LL * Y * m * W);
    gemm(E, Z, W, 2 * m, static_cast<int>(bxy), 1.f, f->Dz, Z, 0, 1, f->T1, W, 2LL * m * W, 0, 1.f, dh, W,
         static_cast<long long>(Z) * W);
}

void fno_forward(pmc_fno* f, int b) {
    pmc_nn* E = &f->eng;
    const int S = f->S, W = f->W, H = f->H, L = f->L;
    const long long R = static_cast<long long>(b) * S;
    float* P = f->params;
    gemm(E, static_cast<int>(R), W, f->Cin, 1, 1.f, f->x, f->Cin, 0, 0, P + f->oP, W, 0, 0, 0.f, f->h[0], W, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * W)), dim3(256), 0, E->stream, f->h[0], P + f->obp, f->h[0], R, W, 0);
    for (int l = 0; l < L; ++l) {
        hipLaunchKernelGGL(pmc_nn_dev::k_fno_expand, dim3(grid_for(4LL * f->M * W * W)), dim3(256), 0, E->stream,
                           P + f->oR[l], f->Bl[l], f->M, W);
        fno_spectral(f, b, f->h[l], f->pre[l], f->Fm[l], f->Bl[l]);
        gemm(E, static_cast<int>(R), W, W, 1, 1.f, f->h[l], W, 0, 0, P + f->oW[l], W, 0, 0, 1.f, f->pre[l], W, 0);
        hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * W)), dim3(256), 0, E->stream, f->pre[l], P + f->ob[l],
                           f->h[l + 1], R, W, l < L - 1 ? 1 : 0);
    }
    gemm(E, static_cast<int>(R), H, W, 1, 1.f, f->h[L], W, 0, 0, P + f->oQ1, H, 0, 0, 0.f, f->preq, H, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R * H)), dim3(256), 0, E->stream, f->preq, P + f->ob1, f->aq, R, H, 1);
    gemm(E, static_cast<int>(R), 1, H, 1, 1.f, f->aq, H, 0, 0, P + f->oQ2, 1, 0, 0, 0.f, f->y, 1, 0);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(R)), dim3(256), 0, E->stream, f->y, P + f->ob2, f->y, R, 1, 0);
}

// Loss and gradient of `b` samples (x, t, w on the device): grads overwritten.
void fno_backward(pmc_fno* f, int b, float wsum) {
    pmc_nn* E = &f->eng;
    const int S = f->S, W = f->W, H = f->H, L = f->L;
    const long long R = static_cast<long long>(b) * S;
    float* P = f->params;
    float* Gd = f->grads;
    hipLaunchKernelGGL(k_zero, dim3(grid_for(f->nparam)), dim3(256), 0, E->stream, Gd, static_cast<long long>(f->nparam));
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, E->stream, f->sse, 1LL);
    hipLaunchKernelGGL(pmc_nn_dev::k_wmse, dim3(blocks(R)), dim3(256), 0, E->stream, f->y, f->t, f->w, f->dy, R, wsum,
                       f->sse);
    gemm(E, H, 1, static_cast<int>(R), 1, 1.f, f->aq, H, 0, 1, f->dy, 1, 0, 0, 0.f, Gd + f->oQ2, 1, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, E->stream, f->dy, R, 1, Gd + f->ob2);
    gemm(E, static_cast<int>(R), H, 1, 1, 1.f, f->dy, 1, 0, 0, P + f->oQ2, 1, 0, 1, 0.f, f->daq, H, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3(blocks(R * H)), dim3(256), 0, E->stream, f->daq, f->preq, R * H);
    gemm(E, W, H, static_cast<int>(R), 1, 1.f, f->h[L], W, 0, 1, f->daq, H, 0, 0, 0.f, Gd + f->oQ1, H, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, E->stream, f->daq, R, H, Gd + f->ob1);
    gemm(E, static_cast<int>(R), W, H, 1, 1.f, f->daq, H, 0, 0, P + f->oQ1, H, 0, 1, 0.f, f->dh, W, 0);
    for (int l = L - 1; l >= 0; --l) {
        float* dpre = f->dpre;
        E->note(hipMemcpyAsync(dpre, f->dh, R * W * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
        if (l < L - 1)
            hipLaunchKernelGGL(k_gelu_back, dim3(blocks(R * W)), dim3(256), 0, E->stream, dpre, f->pre[l], R * W);
        hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, E->stream, dpre, R, W, Gd + f->ob[l]);
        gemm(E, W, W, static_cast<int>(R), 1, 1.f, f->h[l], W, 0, 1, dpre, W, 0, 0, 0.f, Gd + f->oW[l], W, 0);
        gemm(E, static_cast<int>(R), W, W, 1, 1.f, dpre, W, 0, 0, P + f->oW[l], W, 0, 1, 0.f, f->dh2, W, 0);
        fno_spectral_back(f, b, dpre, f->Fm[l], f->Bl[l], Gd + f->oR[l], f->dh2);
        std::swap(f->dh, f->dh2);
    }
    gemm(E, f->Cin, W, static_cast<int>(R), 1, 1.f, f->x, f->Cin, 0, 1, f->dh, W, 0, 0, 0.f, Gd + f->oP, W, 0);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, E->stream, f->dh, R, W, Gd + f->obp);
}

int fno_finish(pmc_fno* f) { return finish(&f->eng, "fno"); }

// Transform matrices (row-major float, 1/sqrt(n) per axis both ways).
void fno_matrices(int n, int m, std::vector<float>& fwd_pair, std::vector<float>& inv_pair) {
    // Complex axis of length n with kept frequencies k_i = i (i < m) or i - 2m:
    // forward [(q, k) = 2*2m][(y, p) = 2n], inverse [(y, p)][(q, k)].
    const double two_pi = 6.283185307179586476925286766559, sc = 1.0 / std::sqrt(static_cast<double>(n));
    const int K = 2 * m;
    fwd_pair.assign(static_cast<size_t>(2 * K) * 2 * n, 0.f);
    inv_pair.assign(static_cast<size_t>(2 * n) * 2 * K, 0.f);
    for (int ki = 0; ki < K; ++ki) {
        const int k = ki < m ? ki : ki - K;
        for (int y = 0; y < n; ++y) {
            const double ph = two_pi * static_cast<double>(k) * y / n;
            const float c = static_cast<float>(std::cos(ph) * sc), s = static_cast<float>(std::sin(ph) * sc);
            auto F = [&](int q, int p) -> float& { return fwd_pair[static_cast<size_t>(q * K + ki) * 2 * n + 2 * y + p]; };
            auto I = [&](int p, int q) -> float& { return inv_pair[static_cast<size_t>(2 * y + p) * 2 * K + q * K + ki]; };
            // (a + ib) e^{-i ph}: re = a c + b s, im = b c - a s
            F(0, 0) = c;
            F(0, 1) = s;
            F(1, 0) = -s;
            F(1, 1) = c;
            // (a + ib) e^{+i ph}: re = a c - b s, im = a s + b c
            I(0, 0) = c;
            I(0, 1) = -s;
            I(1, 0) = s;
            I(1, 1) = c;
        }
    }
}
} // namespace

extern "C" {

int pmc_fno_create(const pmc_fno_shape* sh, pmc_fno** out) {
    if (!sh || !out || sh->nx <= 0 || sh->ny <= 0 || sh->nz <= 0 || sh->in_channels <= 0 || sh->width <= 0 ||
        sh->width > 128 || sh->layers <= 0 || sh->modes <= 0 || 2 * sh->modes > sh->nx || 2 * sh->modes > sh->ny ||
        sh->modes > sh->nz / 2 || sh->hidden <= 0 || sh->hidden > 256 || sh->max_batch <= 0)
        return PMC_ERR_ARGUMENT;
    auto* f = new pmc_fno();
    f->sh = *sh;
    f->X = sh->nx;
    f->Y = sh->ny;
    f->Z = sh->nz;
    f->S = sh->nx * sh->ny * sh->nz;
    f->W = sh->width;
    f->H = sh->hidden;
    f->L = sh->layers;
    f->Cin = sh->in_channels;
    f->m = sh->modes;
    f->M = 4 * f->m * f->m * f->m;
    pmc_nn* E = &f->eng;
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        delete f;
        return PMC_ERR_DEVICE;
    }
    const int W = f->W, H = f->H, M = f->M, m = f->m, X = f->X, Y = f->Y, Z = f->Z;
    int64_t o = 0;
    f->oP = o; o += static_cast<int64_t>(f->Cin) * W;
    f->obp = o; o += W;
    for (int l = 0; l < f->L; ++l) {
        f->oR.push_back(o); o += static_cast<int64_t>(M) * 2 * W * W;
        f->oW.push_back(o); o += static_cast<int64_t>(W) * W;
        f->ob.push_back(o); o += W;
    }
    f->oQ1 = o; o += static_cast<int64_t>(W) * H;
    f->ob1 = o; o += H;
    f->oQ2 = o; o += H;
    f->ob2 = o; o += 1;
    f->nparam = o;
    // Initialization: uniform(-a, a) with a = sqrt(3 / fan_in) (spectral
    // weights: fan-in of the 2W real inputs of a mode); zero biases.
    std::vector<float> p(o, 0.f);
    uint64_t seed = sh->seed;
    auto fill = [&](int64_t off, int64_t n, double fan_in) {
        const double a = std::sqrt(3.0 / fan_in);
        for (int64_t i = 0; i < n; ++i) p[off + i] = static_cast<float>((2.0 * uniform(seed) - 1.0) * a);
    };
    fill(f->oP, static_cast<int64_t>(f->Cin) * W, f->Cin);
    for (int l = 0; l < f->L; ++l) {
        fill(f->oR[l], static_cast<int64_t>(M) * 2 * W * W, 2.0 * W);
        fill(f->oW[l], static_cast<int64_t>(W) * W, W);
    }
    fill(f->oQ1, static_cast<int64_t>(W) * H, W);
    fill(f->oQ2, H, H);
    // Transform matrices.  z is the real axis: forward [(p, kz)][z], inverse
    // [z][(p, kz)] with the factor 2 of the omitted negative kz (kz > 0).
    std::vector<float> dz(static_cast<size_t>(2 * m) * Z), ez(static_cast<size_t>(Z) * 2 * m);
    {
        const double two_pi = 6.283185307179586476925286766559, sc = 1.0 / std::sqrt(static_cast<double>(Z));
        for (int k = 0; k < m; ++k)
            for (int z = 0; z < Z; ++z) {
                const double ph = two_pi * k * z / Z;
                const double c = std::cos(ph) * sc, s = std::sin(ph) * sc, w2 = k == 0 ? 1.0 : 2.0;
                dz[static_cast<size_t>(k) * Z + z] = static_cast<float>(c);
                dz[static_cast<size_t>(m + k) * Z + z] = static_cast<float>(-s);
                ez[static_cast<size_t>(z) * 2 * m + k] = static_cast<float>(w2 * c);
                ez[static_cast<size_t>(z) * 2 * m + m + k] = static_cast<float>(-w2 * s);
            }
    }
    std::vector<float> dy, ey, dx, ex;
    fno_matrices(Y, m, dy, ey);
    fno_matrices(X, m, dx, ex);
    const size_t B = sh->max_batch;
    const size_t R = B * f->S;
    f->params = E->alloc<float>(o);
    f->grads = E->alloc<float>(o);
    f->adam_m = E->alloc<float>(o);
    f->adam_v = E->alloc<float>(o);
    f->Dz = E->alloc<float>(dz.size());
    f->Ez = E->alloc<float>(ez.size());
    f->Dy = E->alloc<float>(dy.size());
    f->Ey = E->alloc<float>(ey.size());
    f->Dx = E->alloc<float>(dx.size());
    f->Ex = E->alloc<float>(ex.size());
    f->x = E->alloc<float>(R * f->Cin);
    f->y = E->alloc<float>(R);
    f->t = E->alloc<float>(R);
    f->w = E->alloc<float>(R);
    f->dy = E->alloc<float>(R);
    f->sse = E->alloc<float>(1);
    f->wsum = E->alloc<float>(1);
    for (int l = 0; l <= f->L; ++l) f->h.push_back(E->alloc<float>(R * W));
    const size_t modes_rows = B * 2 * static_cast<size_t>(M) * W;
    for (int l = 0; l < f->L; ++l) {
        f->pre.push_back(E->alloc<float>(R * W));
        f->Fm.push_back(E->alloc<float>(modes_rows));
        f->Bl.push_back(E->alloc<float>(static_cast<size_t>(M) * 4 * W * W));
    }
    f->T1 = E->alloc<float>(B * X * Y * 2 * static_cast<size_t>(m) * W);
    f->T2 = E->alloc<float>(B * 2 * X * 2 * static_cast<size_t>(m) * m * W);
    f->F = E->alloc<float>(modes_rows);
    f->Gm = E->alloc<float>(modes_rows);
    f->G = E->alloc<float>(modes_rows);
    f->dB = E->alloc<float>(static_cast<size_t>(M) * 4 * W * W);
    f->preq = E->alloc<float>(R * H);
    f->aq = E->alloc<float>(R * H);
    f->daq = E->alloc<float>(R * H);
    f->dh = E->alloc<float>(R * W);
    f->dh2 = E->alloc<float>(R * W);
    f->dpre = E->alloc<float>(R * W);
    f->idx = E->alloc<int>(B);
    if (E->err != hipSuccess) {
        pmc_fno_destroy(f);
        return PMC_ERR_DEVICE;
    }
    E->note(hipMemcpy(f->params, p.data(), o * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemset(f->adam_m, 0, o * sizeof(float)));
    E->note(hipMemset(f->adam_v, 0, o * sizeof(float)));
    E->note(hipMemcpy(f->Dz, dz.data(), dz.size() * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemcpy(f->Ez, ez.data(), ez.size() * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemcpy(f->Dy, dy.data(), dy.size() * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemcpy(f->Ey, ey.data(), ey.size() * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemcpy(f->Dx, dx.data(), dx.size() * sizeof(float), hipMemcpyHostToDevice));
    E->note(hipMemcpy(f->Ex, ex.data(), ex.size() * sizeof(float), hipMemcpyHostToDevice));
    if (E->err != hipSuccess) {
        pmc_fno_destroy(f);
        return PMC_ERR_DEVICE;
    }
    *out = f;
    return PMC_OK;
}

int64_t pmc_fno_param_count(const pmc_fno* f) { return f ? f->nparam : 0; }

int pmc_fno_get_params(pmc_fno* f, float* params) {
    if (!f || !params) return PMC_ERR_ARGUMENT;
    f->eng.note(hipMemcpyAsync(params, f->params, f->nparam * sizeof(float), hipMemcpyDeviceToHost, f->eng.stream));
    return fno_finish(f);
}

int pmc_fno_set_params(pmc_fno* f, const float* params) {
    if (!f || !params) return PMC_ERR_ARGUMENT;
    f->eng.note(hipMemcpyAsync(f->params, params, f->nparam * sizeof(float), hipMemcpyHostToDevice, f->eng.stream));
    return fno_finish(f);
}

int pmc_fno_predict(pmc_fno* f, int32_t batch, const float* in, float* out) {
    if (!f || !in || !out || batch <= 0) return PMC_ERR_ARGUMENT;
    const int S = f->S;
    pmc_nn* E = &f->eng;
    for (int b0 = 0; b0 < batch; b0 += f->sh.max_batch) {
        const int b = std::min(f->sh.max_batch, batch - b0);
        E->note(hipMemcpyAsync(f->x, in + static_cast<size_t>(b0) * S * f->Cin,
                               static_cast<size_t>(b) * S * f->Cin * sizeof(float), hipMemcpyHostToDevice, E->stream));
        fno_forward(f, b);
        E->note(hipMemcpyAsync(out + static_cast<size_t>(b0) * S, f->y, static_cast<size_t>(b) * S * sizeof(float),
                               hipMemcpyDeviceToHost, E->stream));
    }
    return fno_finish(f);
}

int pmc_fno_loss_grad(pmc_fno* f, int32_t batch, const float* in, const float* target, const float* weight,
                      float* loss, float* grad) {
    if (!f || !in || !target || !weight || !loss || batch <= 0 || batch > f->sh.max_batch) return PMC_ERR_ARGUMENT;
    const size_t R = static_cast<size_t>(batch) * f->S;
    double ws = 0.0;
    for (size_t i = 0; i < R; ++i) ws += weight[i];
    if (!(ws > 0.0)) return PMC_ERR_ARGUMENT;
    pmc_nn* E = &f->eng;
    E->note(hipMemcpyAsync(f->x, in, R * f->Cin * sizeof(float), hipMemcpyHostToDevice, E->stream));
    E->note(hipMemcpyAsync(f->t, target, R * sizeof(float), hipMemcpyHostToDevice, E->stream));
    E->note(hipMemcpyAsync(f->w, weight, R * sizeof(float), hipMemcpyHostToDevice, E->stream));
    fno_forward(f, batch);
    fno_backward(f, batch, static_cast<float>(ws));
    float sse = 0.f;
    E->note(hipMemcpyAsync(&sse, f->sse, sizeof(float), hipMemcpyDeviceToHost, E->stream));
    if (grad) E->note(hipMemcpyAsync(grad, f->grads, f->nparam * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    const int rc = fno_finish(f);
    *loss = static_cast<float>(sse / ws);
    return rc;
}

int pmc_fno_train(pmc_fno* f, int32_t n, const float* in, const float* target, const float* weight, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, float* losses) {
    if (!f || !in || !target || !weight || n <= 0 || epochs < 0 || batch <= 0 || batch > f->sh.max_batch ||
        batch > n)
        return PMC_ERR_ARGUMENT;
    pmc_nn* E = &f->eng;
    const int S = f->S;
    const size_t in_row = static_cast<size_t>(S) * f->Cin;
    const bool upload = f->ds_n != n || f->ds_src[0] != in || f->ds_src[1] != target || f->ds_src[2] != weight;
    if (f->ds_n != n) {
        f->ds_in = E->alloc<float>(static_cast<size_t>(n) * in_row);
        f->ds_t = E->alloc<float>(static_cast<size_t>(n) * S);
        f->ds_w = E->alloc<float>(static_cast<size_t>(n) * S);
        f->ds_n = n;
        if (E->err != hipSuccess) return PMC_ERR_DEVICE;
    }
    if (upload) {
        E->note(hipMemcpyAsync(f->ds_in, in, static_cast<size_t>(n) * in_row * sizeof(float), hipMemcpyHostToDevice,
                               E->stream));
        E->note(hipMemcpyAsync(f->ds_t, target, static_cast<size_t>(n) * S * sizeof(float), hipMemcpyHostToDevice,
                               E->stream));
        E->note(hipMemcpyAsync(f->ds_w, weight, static_cast<size_t>(n) * S * sizeof(float), hipMemcpyHostToDevice,
                               E->stream));
        f->ds_src[0] = in;
        f->ds_src[1] = target;
        f->ds_src[2] = weight;
    }
    // Weight sum of each sample (host), for the batch normalization.
    std::vector<double> wrow(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < S; ++k) wrow[i] += weight[static_cast<size_t>(i) * S + k];
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int e = 0; e < epochs; ++e) {
        for (int i = n - 1; i > 0; --i) std::swap(order[i], order[mix(s) % static_cast<uint64_t>(i + 1)]);
        double total = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch, ++nb) {
            double ws = 0.0;
            for (int j = 0; j < batch; ++j) ws += wrow[order[b0 + j]];
            E->note(hipMemcpyAsync(f->idx, order.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, E->stream));
            hipLaunchKernelGGL(k_gather, dim3(blocks(static_cast<long long>(batch) * in_row)), dim3(256), 0, E->stream,
                               f->ds_in, f->idx, f->x, batch, static_cast<long long>(in_row));
            hipLaunchKernelGGL(k_gather, dim3(blocks(static_cast<long long>(batch) * S)), dim3(256), 0, E->stream,
                               f->ds_t, f->idx, f->t, batch, static_cast<long long>(S));
            hipLaunchKernelGGL(k_gather, dim3(blocks(static_cast<long long>(batch) * S)), dim3(256), 0, E->stream,
                               f->ds_w, f->idx, f->w, batch, static_cast<long long>(S));
            fno_forward(f, batch);
            fno_backward(f, batch, static_cast<float>(ws));
            ++f->adam_t;
            const float c1 = 1.f - std::pow(b1, static_cast<float>(f->adam_t));
            const float c2 = 1.f - std::pow(b2, static_cast<float>(f->adam_t));
            hipLaunchKernelGGL(k_adam, dim3(blocks(f->nparam)), dim3(256), 0, E->stream, f->params, f->grads,
                               f->adam_m, f->adam_v, static_cast<long long>(f->nparam), lr, b1, b2, c1, c2);
            float sse = 0.f;
            E->note(hipMemcpyAsync(&sse, f->sse, sizeof(float), hipMemcpyDeviceToHost, E->stream));
            E->note(hipStreamSynchronize(E->stream));
            total += sse / ws;
        }
        if (losses) losses[e] = static_cast<float>(total / std::max(1, nb));
    }
    return fno_finish(f);
}

int pmc_fno_set_timing(pmc_fno* f, int32_t on) { return f ? pmc_nn_set_timing(&f->eng, on) : PMC_ERR_ARGUMENT; }

int pmc_fno_gemm_stats(pmc_fno* f, double* flops, int64_t* launches, double* ms, int32_t reset) {
    return f ? pmc_nn_gemm_stats(&f->eng, flops, launches, ms, reset) : PMC_ERR_ARGUMENT;
}

void pmc_fno_destroy(pmc_fno* f) {
    if (!f) return;
    pmc_nn* E = &f->eng;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    for (auto& e : E->ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : E->allocs) (void)hipFree(p);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete f;
}

} // extern "C"
