/* C ABI of libfiphip: the learned subgrid closure that lets the solver run on a grid
 * `factor` times coarser (factor 4: dx 1 m -> 4 m).
 *
 * A small CNN (4 layers of 3x3 convolutions, 32 channels, GELU or hard SiLU) reads six
 * channels of a coarse state (z, qx, qy, bed, subgrid bed std, kSt) and outputs two
 * corrections of the coarse face discharges, c = tanh(o) * 0.1 * |q|. Corrections are
 * added to interior faces only, so water only moves between cells of the domain: mass is
 * conserved by construction. It is trained on the GPU (bf16 matrix-core convolutions, fp32
 * accumulation, Adam) on pairs the solver itself produces: a restricted fine state, one
 * plain coarse step from it, and the restricted fine state one coarse step later.
 *
 * The reference has no counterpart (it solves one grid at one resolution,
 * src/kernels.cu:264-438); the restriction and the closure are pinned by this project's
 * own numpy oracle (closed forms, invariants, finite-difference gradients).
 * Return codes and fip_last_error() as in fip.h.
 */
#ifndef FIP_CLOSURE_H
#define FIP_CLOSURE_H
#include "fip.h"
#include "fip_ensemble.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FIP_CL_LAYERS 4
#define FIP_CL_CH 32
/* padded parameter layout: per layer W[tap 0..8][ci 0..31][co 0..31] then bias[co 0..31];
 * layer 0 reads channels 0..5 only, layer 3 writes channels 0..1 only (others are 0) */
#define FIP_CL_LAYER_PARAMS (9 * 32 * 32 + 32)
#define FIP_CL_PARAMS (FIP_CL_LAYERS * FIP_CL_LAYER_PARAMS)

typedef struct fip_closure fip_closure;

/* Restriction of the fine simulation's current state onto the coarse simulation (both
 * ping-pong sides): depth = block mean, face discharge = mean of the `factor` fine faces
 * of the coarse face (the total discharge is kept); the coarse duration becomes the fine
 * one. The coarse grid must be exactly W/factor x H/factor. */
int fip_restrict(fip_sim* fine, fip_sim* coarse, int factor);

/* The network (He-initialised from seed, last layer scaled by 0.1) and room for
 * `max_samples` training samples of a Wc x Hc coarse grid. */
int fip_closure_create(int device, unsigned long long seed, int Wc, int Hc, int max_samples,
                       fip_closure** out);
int fip_closure_destroy(fip_closure* cl);
/* The activation after layers 0..2. GELU (default): tanh form,
 * x / (1 + 2^(-2 sqrt(2/pi) log2(e) x (1 + 0.044715 x^2))). Hard SiLU:
 * x * clamp(x / 4 + 1 / 2, 0, 1), no transcendental instruction, so the fused inference
 * kernel is faster; trained networks of the two differ in accuracy. Training, inference
 * and gradients all use it; parameters trained with one are meant for that one. Set it
 * before fip_closure_attach*. */
#define FIP_CLOSURE_ACT_GELU 0
#define FIP_CLOSURE_ACT_HSILU 1
int fip_closure_set_activation(fip_closure* cl, int act);
int fip_closure_get_activation(fip_closure* cl, int* act);
int fip_closure_get_params(fip_closure* cl, float* params);
/* Sets the parameters and restarts the optimiser: Adam's moment estimates and step count
 * are reset to zero. */
int fip_closure_set_params(fip_closure* cl, const float* params);

/* One training sample. `coarse` has just taken one plain step from a restricted state: its
 * other side holds that state (the input), its current side the plain result. `fine` has
 * advanced by the same time: its restriction is the target. bed_std: subgrid std of the
 * fine bed per coarse cell (Wc x Hc); zref: reference level of the z and bed features. */
int fip_closure_add_sample(fip_closure* cl, fip_sim* coarse, fip_sim* fine, int factor,
                           const float* bed_std, float zref);
int fip_closure_num_samples(fip_closure* cl, int* n);

/* `steps` Adam steps (lr, beta 0.9/0.999) on batches of `batch` random patch x patch
 * crops from samples [first, first + n); losses[s] = the batch loss (mean squared error
 * of the corrected coarse face discharges against the targets, m^4/s^2) of step s. */
int fip_closure_train(fip_closure* cl, int steps, int batch, int patch, float lr, int first, int n,
                      unsigned long long seed, float* losses);
/* The optimiser step fip_closure_train takes after each backward pass, from a given
 * gradient (FIP_CL_PARAMS floats, padded layout): Adam with beta 0.9 / 0.999, eps 1e-8 on
 * the fp32 master weights, continuing the closure's moment estimates and step count;
 * padded parameters stay 0. For callers that form gradients themselves. */
int fip_closure_adam_step(fip_closure* cl, const float* grad, float lr);
/* Mean squared face-discharge error over samples [first, first + n): of the plain coarse
 * step and of the corrected one. */
int fip_closure_eval(fip_closure* cl, int first, int n, double* mse_plain, double* mse_learned);

/* Inference: after every step of fip_simulate on `coarse`, its face discharges are
 * corrected (faces where |o| > sat keep the plain step), and the mass balance is checked
 * before the next correction; a relative error above mass_limit switches the domain to
 * the plain step for good. */
int fip_closure_attach(fip_closure* cl, fip_sim* coarse, const float* bed_std, float zref,
                       float sat, double mass_limit);
/* The same for every scenario of an ensemble; the scenarios may differ in size. bed_std
 * holds each scenario's W x H values in turn (n x W x H when they have one size), zref n
 * values. */
int fip_closure_attach_ensemble(fip_closure* cl, fip_ens* ens, const float* bed_std,
                                const float* zref, float sat, double mass_limit);
int fip_closure_detach(fip_sim* coarse);
/* State of an attached domain (scenario s of an ensemble, or s = 0 with ens = NULL),
 * the mass check of the last step included. */
int fip_closure_status(fip_closure* cl, fip_sim* coarse, fip_ens* ens, int s, int* fallback,
                       unsigned* saturated_faces, double* max_rel_mass_error);

/* Test hooks on host data. fip_closure_forward: features [N][H][W][6] -> raw outputs
 * [N][H][W][2]. fip_closure_backward: loss and gradient (FIP_CL_PARAMS) of
 * sum m (qp + tanh(o) bound |qp| - qt)^2 / (2 N H W) for the given features, plain fluxes
 * qp, targets qt and face mask ([N][H][W][2] each). */
int fip_closure_forward(fip_closure* cl, const float* feat, int N, int H, int W, float* out);
int fip_closure_backward(fip_closure* cl, const float* feat, int N, int H, int W, const float* qp,
                         const float* qt, const unsigned char* mask, float bound, float* grad,
                         double* loss);
/* fip_closure_forward without allocating: feat and out are DEVICE pointers on the
 * closure's device (feat [N][H][W][6] fp32, ready before the call; out [N][H][W][2] fp32),
 * and the activations live in a static device array of the library (64 MiB per device,
 * held until the process ends) instead of buffers it allocates. N*H*W must not exceed
 * FIP_CL_NOALLOC_MAX_PIXELS (a 512 x 512 coarse grid) and N must not exceed 65535;
 * larger batches are rejected with FIP_E_INVALID. Runs on the closure's stream and
 * returns when out is written. All calls on one device share that array: calls that
 * overlap in time on one device race. No counterpart in the reference: for callers
 * that cannot allocate device memory (tight serving loops, code that must not call
 * hipMalloc). Outputs equal fip_closure_forward's bit for bit. */
#define FIP_CL_NOALLOC_MAX_PIXELS (512 * 512)
int fip_closure_forward_noalloc(fip_closure* cl, const float* feat, int N, int H, int W, float* out);
/* The raw outputs [H][W][2] of the fused inference kernel on the coarse simulation's
 * current state (no correction applied): checks it against fip_closure_forward. */
int fip_closure_fused_outputs(fip_closure* cl, fip_sim* coarse, const float* bed_std, float zref,
                              float* out);
/* Launch form of the fused inference kernel: 2 (default) = row march (strips of 58 columns
 * marched along their rows, each layer's last rows in LDS rings); 1 = persistent, one work-group
 * per CU over runs of 16 x 16 tiles with the weights staged in LDS; 0 = one work-group per
 * 16 x 16 tile. All give bit-identical outputs. It is refused while the closure is
 * attached. */
int fip_closure_set_kernel(fip_closure* cl, int form);
/* D = A B of one bf16 matrix-core instruction (A 16x32, B 32x16 row-major bf16 bits,
 * D 16x16 fp32): checks the operand layout the convolutions assume. */
int fip_mfma_probe(int device, const unsigned short* A, const unsigned short* B, float* D);

#ifdef __cplusplus
}
#endif
#endif
