/*
 * C ABI of the learned surrogate of the D2Q9 time-stepper (libpylbm_hip.so).
 *
 * None of these entry points has a counterpart in pylbm: the surrogate is an
 * addition that advances the moment fields (rho, ux, uy) of a periodic 2D run
 * by K lattice steps per network evaluation.  It reads the moments through the
 * same path as Simulation.m (pylbm/simulation.py:215-224, f2m over the halo
 * domain, pylbm/algorithm/base.py:336-344) and restarts the lattice Boltzmann
 * run with pylbm's "moments -> equilibrium -> m2f" initialisation
 * (pylbm/simulation.py:257-320).
 *
 * Tensors (all device memory, row-major, channels contiguous = NHWC):
 *   state        FP32 [N][Hf][Wf][4]  (rho, ux, uy, 0) on the fine (lattice) grid
 *   coarse       H x W = Hf/4 x Wf/4 pixels; channel (4 py + px) * 4 + ch of coarse
 *                pixel (Y, X) is field ch of fine pixel (4 Y + py, 4 X + px)
 *   input        E [N][H][W][64]      normalised state, field 3 = 0
 *   activations  E [N][H][W][C]       C = 32 or 64 hidden channels
 *   output       FP32 [N][H][W][64]   increment in units of dscale (field 3 unused)
 * Hf is the mid (pylbm x) axis, Wf the inner (pylbm y) axis.  Convolutions need
 * W % 16 == 0, weight gradients W % 32 == 0.  Convolutions are periodic 3x3;
 * taps t = 3 (dy + 1) + (dx + 1) read pixel (y + dy, x + dx).  Master weights:
 * FP32 w[9][cin][cout], b[cout].  Packed operands (E, written by lbm_sur_pack):
 *   forward  A [cout][9 cin],  k = tap * cin + ci
 *   backward A [cin][9 cout],  the input-gradient convolution: weights of tap
 *            8 - t with channel roles swapped, k = tap * cout + co
 * E is the 16-bit element type of every input, activation, gradient and packed
 * operand buffer, chosen per call by `elem`: LBM_SUR_F16 (IEEE half) or
 * LBM_SUR_BF16 (bfloat16); the MFMA products are v_mfma_f32_16x16x32_f16 / _bf16
 * with FP32 accumulation, and FP32 -> E conversions round to nearest even.  All
 * buffers of one network must use the same E.
 * All calls are asynchronous on `stream` unless they return a host value.
 */
#ifndef PYLBM_SURROGATE_H
#define PYLBM_SURROGATE_H

#include <stdint.h>

#include "pylbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_SUR_F16 0  /* elem: IEEE binary16 */
#define LBM_SUR_BF16 1 /* elem: bfloat16 */

/*
 * One periodic 3x3 convolution on the MFMA units (FP32 accumulate).
 * epilogue 0: z = conv + b, out = silu(z) (E, coutp channels), d_zout = z (E) if not NULL
 * epilogue 1: out = conv + b (FP32, coutp channels)
 * epilogue 2: out = conv * silu'(d_zin) (E): the input gradient of a layer
 *             followed by the SiLU of the previous one (d_wpk = backward operand,
 *             d_bias unused)
 * Supported (cinp, coutp, epilogue): (64,64,0) (64,32,0) (32,32,0) (32,64,1) (64,64,1)
 * (64,64,2) (64,32,2) (32,32,2).
 */
int lbm_sur_conv(const void* d_in, int cinp, const void* d_wpk, const float* d_bias, const void* d_zin,
                 void* d_out, void* d_zout, int coutp, int epilogue, int N, int H, int W, int elem, void* stream);

/*
 * The last (64-channel output) convolution of an inference step fused with its use:
 * d_state += (conv + b) * dscale for the 3 fields of every fine pixel, and the next
 * network input (d_state - shift) / scale written to d_xnext (E, coarse layout;
 * must not alias d_in).  dscale, shift, scale: 3 host floats.  cinp = 32 or 64.
 */
int lbm_sur_apply_step(const void* d_in, int cinp, const void* d_wpk, const float* d_bias, float* d_state,
                       void* d_xnext, const float* dscale, const float* shift, const float* scale, int N, int H,
                       int W, int elem, void* stream);

/* the buffers of one network for lbm_sur_run: act[0] = input (E, 64 channels),
 * act[1..3] = hidden activations (E, `hidden` channels; not used by
 * lbm_sur_run), xalt = a second input buffer of act[0]'s size, wpk / bias =
 * forward operands and biases of the 4 layers, elem = E (LBM_SUR_F16 / _BF16) */
typedef struct lbm_sur_net {
  void* act[4];
  void* xalt;
  const void* wpk[4];
  const float* bias[4];
  int32_t hidden;
  int32_t elem;
} lbm_sur_net;

/* nsteps inference steps without returning to the host, one launch per step: a
 * workgroup computes all 4 layers of a 4 x 16 coarse-pixel tile (halo recomputed,
 * activations in LDS), adds the increment to the state and writes the next input.
 * The first step reads d_src (NULL: d_state) and writes d_state; later steps update
 * d_state in place (d_src == d_state is in place throughout; nsteps == 0 copies
 * d_src to d_state).  act[0] must hold the network input of d_src on entry; on
 * return it holds the input of the new d_state: after an odd nsteps the call
 * exchanges act[0] and xalt in *net rather than copying.  The coarse grid must be
 * a multiple of 4 x 16. */
int lbm_sur_run(lbm_sur_net* net, const float* d_src, float* d_state, const float* dscale, const float* shift,
                const float* scale, int64_t nsteps, int N, int H, int W, void* stream);

/* number of workgroups per tap (partial sums) lbm_sur_wgrad uses */
int lbm_sur_wgrad_blocks(void);

/*
 * Weight and bias gradients of one convolution: d_dwt [9][cdz][cinp] = sum over
 * pixels of dZ[p][co] X[p + d_tap][ci], d_db[cdz] = sum dZ.  d_x: E input (cinp
 * channels), d_dz: E gradient of the pre-activation (cdz channels).  d_partial:
 * lbm_sur_wgrad_blocks() * (9 cdz cinp + cdz) floats; the partial sums are reduced
 * in a fixed order (results are deterministic).  cinp, cdz in {32, 64}.
 */
int lbm_sur_wgrad(const void* d_x, int cinp, const void* d_dz, int cdz, int N, int H, int W, float* d_partial,
                  float* d_dwt, float* d_db, int elem, void* stream);

/*
 * Mean-square loss of the FP32 output against the FP32 target over the 48 value
 * channels (field 3 excluded) of npix coarse pixels; d_dy = 2 gscale (y - t) as
 * E [npix][64]: gscale = 1 / (48 npix) is the exact loss gradient, gscale = 1
 * the loss-scaled one the 16-bit backward pass uses (Adam with eps * 48 npix then
 * takes the identical step).  d_partial: 256 doubles.  h_loss (host, may be NULL:
 * then nothing synchronises) = the loss.
 */
int lbm_sur_mse(const float* d_y, const float* d_t, int64_t npix, float gscale, void* d_dy, double* d_partial,
                double* h_loss, int elem, void* stream);

/* pack the FP32 master weights of one layer into the forward (and, if d_abwd is
 * not NULL, backward) E MFMA operands; cin <= cinp <= 64, cout <= coutp <= 64 */
int lbm_sur_pack(const float* d_w, int cin, int cout, int cinp, int coutp, void* d_afwd, void* d_abwd, int elem,
                 void* stream);

/* one Adam step (bias-corrected, step >= 1) on w[9][cin][cout] and b[cout]; d_m / d_v
 * hold 9 cin cout + cout moments each (weights first); gradients from lbm_sur_wgrad
 * (layout [9][coutp][cinp]) */
int lbm_sur_adam(float* d_w, float* d_b, float* d_m, float* d_v, const float* d_dwt, const float* d_db, int cin,
                 int cout, int cinp, int coutp, float lr, float b1, float b2, float eps, int step, void* stream);

/* network input (E coarse [N][Hf/4][Wf/4][64]) = (state - shift) / scale per field
 * (3 host floats each) */
int lbm_sur_make_input(const float* d_state, int N, int Hf, int Wf, const float* shift, const float* scale,
                       void* d_xin, int elem, void* stream);
/* training target (FP32 coarse [N][Hf/4][Wf/4][64]) = (s1 - s0) / dscale */
int lbm_sur_make_target(const float* d_s0, const float* d_s1, int N, int Hf, int Wf, const float* dscale,
                        float* d_t, void* stream);

/* the same distance without draining the stream: d_partial needs 514 doubles; the
 * sums (sum |u_a - u_b|^2, sum |u_b|^2) land in h_out[0..1] (pinned host memory, see
 * lbm_host_malloc) once the stream reaches this point (lbm_event_record / lbm_event_sync) */
int lbm_sur_rel_l2_async(const float* d_a, const float* d_b, int64_t npix, double* d_partial, double* h_out,
                         void* stream);

/* state (rho, qx / rho, qy / rho) straight from plain FP32 populations of a 2D layout:
 * m_i = sum_k mc[i * ns + k] f_k, mc = the rows of M of rho, qx and qy (3 * ns host
 * doubles); the moments Simulation.m reads (pylbm/simulation.py:215-224), conserved
 * rows only */
int lbm_sur_populations_to_state(const float* d_f, const lbm_layout* layout, const double* mc, int H, int W,
                                 float* d_state, void* stream);

/* the read-back of a judge in one pass: the state of plain FP32 D2Q9 populations (as
 * lbm_sur_populations_to_state) into d_state and, if d_ckpt is not NULL, d_ckpt; its
 * network input (as lbm_sur_make_input, E = elem) into d_xin; and the drift sums
 * against d_pred (as lbm_sur_rel_l2_async(d_pred, d_state), same accumulation order)
 * into h_out[0..1] (pinned host memory) once the stream reaches this point.
 * d_partial: 514 doubles.  H, W: multiples of 4. */
int lbm_sur_judge_read(const float* d_f, const lbm_layout* layout, const double* mc, int H, int W,
                       const float* d_pred, float* d_state, float* d_ckpt, const float* shift, const float* scale,
                       void* d_xin, int elem, double* d_partial, double* h_out, void* stream);

/* training noise: state += sigma[f] * N(0, 1) on the fields f = rho, ux, uy of every
 * fine pixel, from a counter-based generator keyed by seed (reproducible) */
int lbm_sur_perturb(float* d_state, int64_t npix, const float* sigma, uint64_t seed, void* stream);

/* FP64 moment planes of a 2D layout (rows irho, iqx, iqy) -> state (rho, qx/rho, qy/rho) */
int lbm_sur_moments_to_state(const double* d_m, const lbm_layout* layout, int irho, int iqx, int iqy,
                             int H, int W, float* d_state, void* stream);
/* state -> conserved FP64 moments (rho, rho ux, rho uy) on the whole halo-inclusive
 * domain (periodic wrap); the caller then runs equilibrium and m2f */
int lbm_sur_state_to_moments(const float* d_state, int H, int W, const lbm_layout* layout, int irho,
                             int iqx, int iqy, double* d_m, void* stream);

/* relative L2 distance of the velocity channels: sqrt(sum |u_a - u_b|^2 / sum |u_b|^2)
 * (host result; synchronises the stream).  d_partial: 512 doubles. */
int lbm_sur_rel_l2(const float* d_a, const float* d_b, int64_t npix, double* d_partial, double* h_out,
                   void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYLBM_SURROGATE_H */
