/* pmc_surrogate.h -- spectral neural operator that predicts the steady cell
 * temperature field from source density, solid mask and reservoir temperature,
 * used to seed the initial carrier population (pmc_init.temperature_mode = 2,
 * the per-cell field of src/solver/Initialization.cpp:181-265).
 *
 * No counterpart in the reference: it is the learned warm start of the Monte
 * Carlo step.  All products run in BF16 on CDNA4 matrix cores with FP32
 * accumulation; parameters, gradients and Adam moments are FP32.
 *
 * Network on an nx x nz grid (S = nx*nz points, index x*nz + z):
 *   h0 = x P + bp                                   (in_channels -> width)
 *   h_{l+1} = act(G Mix_l(D^T h_l) + h_l W_l + b_l)   l = 0..layers-1 (no act on the last)
 *   y = act(h_L Q1 + b1) Q2 + b2                     (width -> hidden -> 1)
 * D (S x 2M) holds cos/sin of the kept modes (|kx| < modes_x, 0 <= kz < modes_z),
 * G = D / S; Mix_l multiplies each mode's (cos, sin) coefficient pair of all
 * channels by its own 2width x 2width matrix.  act is the tanh form of GELU.
 * Parameter layout: P, bp, then per layer R_l [M][2w][2w], W_l, b_l, then Q1, b1, Q2, b2. */
#ifndef PMC_SURROGATE_H
#define PMC_SURROGATE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmc_nn pmc_nn;

typedef struct pmc_nn_shape {
    int32_t nx, nz;
    int32_t in_channels;
    int32_t width;
    int32_t layers;
    int32_t modes_x, modes_z;
    int32_t hidden;
    int32_t max_batch;
    uint64_t seed;                 /* parameter initialization */
} pmc_nn_shape;

int pmc_nn_create(const pmc_nn_shape* shape, pmc_nn** out);
int64_t pmc_nn_param_count(const pmc_nn* nn);
int pmc_nn_get_params(pmc_nn* nn, float* params);
int pmc_nn_set_params(pmc_nn* nn, const float* params);
/* Fourier basis D [S][2M] (G = D / S). */
int pmc_nn_basis(pmc_nn* nn, float* d);
/* in: [batch][S][in_channels]; out: [batch][S]. */
int pmc_nn_predict(pmc_nn* nn, int32_t batch, const float* in, float* out);
/* Mean-squared error over batch*S points and its gradient [param_count]. */
int pmc_nn_loss_grad(pmc_nn* nn, int32_t batch, const float* in, const float* target, float* loss, float* grad);
/* Adam (beta 0.9/0.999) over shuffled minibatches of a host data set;
 * losses[epochs] receives the mean training loss of each epoch. */
int pmc_nn_train(pmc_nn* nn, int32_t n, const float* in, const float* target, int32_t epochs, int32_t batch,
                 float lr, uint64_t seed, float* losses);
/* GEMM accounting of the operator's own launches: algorithmic flops
 * (2 m n k per product, BF16 MFMA), launch count and, while timing is on,
 * the summed HIP-event durations of those launches in ms.  reset != 0 zeroes
 * the counters after reading them. */
int pmc_nn_set_timing(pmc_nn* nn, int32_t on);
int pmc_nn_gemm_stats(pmc_nn* nn, double* flops, int64_t* launches, double* ms, int32_t reset);
void pmc_nn_destroy(pmc_nn* nn);

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b] on host arrays through the device
 * GEMM (BF16 products, FP32 accumulation), row-major; op transposes when t != 0. */
int pmc_nn_gemm(int32_t m, int32_t n, int32_t k, int32_t batch, float alpha, const float* a, int32_t lda,
                int64_t stride_a, int32_t ta, const float* b, int32_t ldb, int64_t stride_b, int32_t tb,
                float beta, float* c, int32_t ldc, int64_t stride_c);

/* Fixed capacity (floats) of the split-k buffer of pmc_nn_gemm_noalloc: 16 MiB
 * per GPU.  The launch plan (the one pmc_nn_gemm uses: at most about 512
 * workgroups of two 64-deep stages) needs at most about 3.2 M floats. */
#define PMC_NN_GEMM_SCRATCH_FLOATS (4u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the product of pmc_nn_gemm, same kernels and launch plan,
 * on device arrays a, b and c, enqueued on `stream` (a hipStream_t; NULL is the
 * null stream) without waiting.  It calls no allocator.  Split-k partial
 * products go to a buffer of PMC_NN_GEMM_SCRATCH_FLOATS floats that the
 * library holds on each GPU until the process ends; a product whose plan needs
 * more returns PMC_ERR_ARGUMENT.  Calls that overlap on one GPU (other
 * streams, other threads) race on that buffer. */
int pmc_nn_gemm_noalloc(int32_t m, int32_t n, int32_t k, int32_t batch, float alpha, const float* a, int32_t lda,
                        int64_t stride_a, int32_t ta, const float* b, int32_t ldb, int64_t stride_b, int32_t tb,
                        float beta, float* c, int32_t ldc, int64_t stride_c, void* stream);

#ifdef __cplusplus
}
#endif
#endif
