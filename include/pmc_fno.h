/* pmc_fno.h -- three-dimensional spectral neural operator that predicts the
 * steady temperature rise of a FinFET-like solid from its heat-source density,
 * solid mask and reservoir temperature on a regular nx x ny x nz grid; the
 * prediction seeds the per-cell initial field of the Monte Carlo step
 * (pmc_init.temperature_mode = 2, the role of the uniform and linear fields of
 * src/solver/Initialization.cpp:181-265).
 *
 * No counterpart in the reference: it is the learned warm start.  Every product
 * runs in BF16 on gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32
 * accumulation; parameters, gradients and Adam moments are FP32.
 *
 * Layout: a sample is S = nx*ny*nz points, index (x*ny + y)*nz + z, channels
 * last.  Network (width W, hidden H, L layers):
 *   h0 = in P + bp
 *   h_{l+1} = act(K_l h_l + h_l W_l + b_l)            (no act after the last layer)
 *   y = act(h_L Q1 + b1) Q2 + b2
 * K_l is the spectral convolution: the real 3-D Fourier transform truncated to
 * kx in [0, modes) U [nx - modes, nx), ky likewise, kz in [0, modes), each kept
 * mode's channel vector multiplied by its own complex W x W matrix, then the
 * inverse transform.  The transform is separable: three axis contractions with
 * cos/sin matrices (1/sqrt(n) per axis both ways), each a batched GEMM.
 * act is the tanh form of GELU.
 * Parameters: P [Cin][W], bp [W], per layer R_l [M][2][W][W] (real then imaginary
 * part, [in][out], M = 4 modes^3), W_l [W][W], b_l [W]; Q1 [W][H], b1 [H],
 * Q2 [H], b2. */
#ifndef PMC_FNO_H
#define PMC_FNO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmc_fno pmc_fno;

typedef struct pmc_fno_shape {
    int32_t nx, ny, nz;
    int32_t in_channels;
    int32_t width;
    int32_t layers;
    int32_t modes;                 /* per axis (kx and ky take +-modes, kz 0..modes-1) */
    int32_t hidden;
    int32_t max_batch;
    uint64_t seed;                 /* parameter initialization */
} pmc_fno_shape;

int pmc_fno_create(const pmc_fno_shape* shape, pmc_fno** out);
int64_t pmc_fno_param_count(const pmc_fno* fno);
int pmc_fno_get_params(pmc_fno* fno, float* params);
int pmc_fno_set_params(pmc_fno* fno, const float* params);
/* in: [batch][S][in_channels]; out: [batch][S]. */
int pmc_fno_predict(pmc_fno* fno, int32_t batch, const float* in, float* out);
/* Weighted mean-squared error sum w (y - t)^2 / sum w over batch*S points
 * (w: [batch][S], e.g. the solid mask) and its gradient [param_count]. */
int pmc_fno_loss_grad(pmc_fno* fno, int32_t batch, const float* in, const float* target, const float* weight,
                      float* loss, float* grad);
/* Adam (beta 0.9/0.999) at learning rate lr over shuffled minibatches of a
 * data set held on the device (uploaded on the first call and whenever n or
 * the pointers change); losses[epochs] receives each epoch's mean loss. */
int pmc_fno_train(pmc_fno* fno, int32_t n, const float* in, const float* target, const float* weight, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, float* losses);
/* GEMM accounting of the operator's launches (flops 2 m n k per product,
 * launches, and with timing on the summed HIP-event durations in ms). */
int pmc_fno_set_timing(pmc_fno* fno, int32_t on);
int pmc_fno_gemm_stats(pmc_fno* fno, double* flops, int64_t* launches, double* ms, int32_t reset);
void pmc_fno_destroy(pmc_fno* fno);

#ifdef __cplusplus
}
#endif
#endif
