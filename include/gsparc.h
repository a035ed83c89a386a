/* gsparc.h -- C ABI of the MI355X (gfx950) equirectangular RF rasterizer and
 * fused-SSIM loss.
 *
 * Plain device pointers, sizes and a hipStream_t (passed as void*); every call
 * returns a status code (GSPARC_OK == 0) and never throws.  All tensors are
 * float32, C-contiguous, on the current HIP device.  Receivers are batched:
 * R = 1 reproduces the reference's one-receiver-per-call interface.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the GSpaRC repository root).  The reference binds those through
 * pybind11/torch (cuda_kernel_rasterize/ext.cpp:7-11, fused-ssim/ext.cpp:4-7);
 * the ctypes binding in wireless_rasterizer.py / fused_ssim_cuda.py of this
 * package binds these.
 *
 * Buffers.  The caller owns four byte buffers, sized by the queries below:
 *   geom  per (receiver, Gaussian) records             gsparc_geom_buffer_size
 *   image per pixel / per tile state                   gsparc_image_buffer_size
 *   bin   per duplicated (tile, Gaussian) instance     gsparc_binning_buffer_size
 *   grad  backward accumulation                        gsparc_grad_buffer_size
 * geom, image and bin must survive from the forward to the backward, as the
 * reference's geomBuffer / imgBuffer / binningBuffer do
 * (cuda_rasterizer/__init__.py:48).
 */
#ifndef GSPARC_H
#define GSPARC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSPARC_OK 0
#define GSPARC_ERR_ARG 1        /* bad size / null pointer */
#define GSPARC_ERR_HIP 2        /* a HIP runtime call failed */
#define GSPARC_ERR_CAPACITY 3   /* binning buffer smaller than K */

#define GSPARC_GRAD_EXACT 0     /* true adjoint of the forward */
#define GSPARC_GRAD_REFERENCE 1 /* the reference backward, quirks included */
/* OR-ed into grad_mode of gsparc_rasterize_backward(_scenes): deterministic
 * gradients.  The blend adjoint stores each tile-list instance's sums to a
 * slot of its own (one wave per tile) instead of adding them to the record
 * with float atomics (the reference's atomicAdd, backward.cu:588-613, whose
 * summation order changes from run to run); each record's instances are
 * then summed in a fixed order (its tiles row by row).  Same inputs give
 * bit-identical gradients.  `grad` must hold
 * gsparc_deterministic_grad_size(normal grad bytes, K_capacity) bytes.  No
 * counterpart in the reference, which has no deterministic mode. */
#define GSPARC_GRAD_DETERMINISTIC 16

#define GSPARC_FLAG_KEEP_KEYS 1 /* also store the sorted depth keys */

/* Byte offsets of the arrays inside the caller's buffers (for inspection). */
typedef struct gsparc_layout {
    /* geom: per (receiver, Gaussian) records and the per-receiver depth sort */
    size_t splat, sig, depth, cov3, sort_key0, sort_val0, sort_key1, sort_val1, sort_hist, nvis, plen, rect, srect,
        geom_bytes;
    /* image: per pixel and per tile state, per (chunk, tile) counts */
    size_t final_t, n_contrib, ranges, segcnt, segbase, tile_order, chunk_cnt, total, image_bytes;
    /* bin: the tile lists */
    size_t plist, keys_sorted, bin_bytes;
} gsparc_layout;

const char* gsparc_version(void);
const char* gsparc_last_error(void);

/* Tuning knobs (process-wide; no counterpart in the reference):
 *   "render_fwd_ppl", "render_bwd_ppl"  pixels per lane of the forward and
 *       backward blend kernels: 0 = automatic (1 for few tiles, else 2),
 *       1 = one 8x8 quarter-tile per wave, 2 = half a tile per wave, 4 = one
 *       16x16 tile per wave.  Speed only: the forward result is bit-identical
 *       either way, the backward differs by summation order.
 *   "render_fwd_depth", "render_bwd_depth"  batch depth of the blend
 *       kernels: list entries staged in LDS per batch (the reference stages
 *       BLOCK_SIZE = 256 per batch, forward.cu:193-207, backward.cu:523-541):
 *       0 = automatic (64), 32, 64 or 128.  Speed only, as above.
 *   "render_order"  order in which the blend kernels take the (receiver,
 *       tile) pairs, longest lists first: 0 = automatic (2 from 16
 *       receivers on, else 1), 1 = one list dealt
 *       round robin to the 8 XCDs, 2 = XCD x takes the receivers x mod 8,
 *       3 = as 2, one receiver after the other.  Speed only.
 *   "csi_fwd_ppl"  the multi-frequency forward blend: 0 = the matrix-core
 *       kernel (the blend weights of a batch of entries times their 104
 *       emissions by MFMA), 1 or 2 = the vector-ALU kernel with that many
 *       pixels per lane.  The maps differ only by the order of the sums over
 *       entries; final T and contributor counts are identical.
 * "Automatic" means: what gsparc_autotune chose for a batch of this shape
 * (P, R, W, H, scenes), else the built-in rule above. */
int gsparc_set_option(const char* name, int value);
int gsparc_get_option(const char* name, int* value);

/* Autotuner of the blend kernels' compute mapping: on the buffers of a
 * finished forward (any form, K_capacity and radii as there), times the
 * forward and backward blend under every pixels-per-lane mapping (1, 2, 4),
 * batch depth (32, 64, 128) and work order (1, 2, 3) and keeps the fastest
 * combination for batches of
 * this shape; later calls with the options at 0 use it.  Each candidate is
 * timed `iters` times (median) where it runs in a step: the forward blend
 * after the ordered scatter that writes its lists, the backward blend after a
 * forward blend and the gradient clear.  choice[5] = {forward ppl, backward
 * ppl, order, forward batch depth, backward batch depth}; ms[2] (may be
 * NULL) = their per-launch times; table[54] (may be NULL) = per order, the
 * forward then the backward times for (ppl, depth) = (1, 32), (1, 64),
 * (1, 128), (2, 32), ..., (4, 128).
 * dout_real / dout_imag [R,H,W] drive the timed backward; out_real / out_imag
 * [R,H,W] receive the (unchanged) spectra; `grad` is overwritten, so run the
 * backward after tuning.  Synchronises the stream; not while capturing.  The
 * 16x16 tile binning itself stays fixed: it is part of the numerics
 * (config.h:16-17, auxiliary.h:14-24).  No counterpart in the reference
 * (fixed BLOCK_X x BLOCK_Y thread blocks, forward.cu:151-249). */
int gsparc_autotune(int grad_mode, int P, int R, int W, int H, const float* emission, const int* radii,
                    void* geom, void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout_real,
                    const float* dout_imag, float* out_real, float* out_imag, int iters, int* choice, float* ms,
                    float* table, void* stream);
int gsparc_autotune_clear(void);

size_t gsparc_geom_buffer_size(int P, int R, int W, int H);
size_t gsparc_image_buffer_size(int P, int R, int W, int H);
size_t gsparc_binning_buffer_size(uint64_t K);
size_t gsparc_grad_buffer_size(int P, int R);
/* grad workspace of a deterministic backward: the normal size (P, R form or
 * gsparc_scene_buffer_sizes' sizes[3]) rounded up to 256 bytes, plus 32
 * bytes per tile-list instance the bin buffer can hold */
size_t gsparc_deterministic_grad_size(size_t grad_bytes, uint64_t K_capacity);
int gsparc_layout_query(int P, int R, int W, int H, uint64_t K, gsparc_layout* out);

/* Forward, part 1: per-Gaussian preprocess for R receivers (projection,
 * covariance, conic, radius, tile rect, emission MLP), per-tile instance
 * depth order per receiver and per-tile instance counts and tile ranges.
 * Copies K (the number of duplicated instances,
 * "num_rendered") to *K_out and synchronises the stream, as the reference
 * does before sizing its binning buffer.
 * Replaces: CudaRasterizer::Rasterizer::forward up to the K readback
 * (cuda_wireless_rasterizer/rasterizer_impl.cu:206-278), FORWARD::preprocess
 * (forward.cu:252-327, 359-406) and the inclusive scan (rasterizer_impl.cu:270).
 *   means [P,3], scales [P,3], rotations [P,4] (w,x,y,z), opacities [P],
 *   emission_mlps [P*194] (gaussian.py:554-572; 16-byte aligned) or NULL when
 *   emission is supplied to part 2, viewmatrices [R,16] column-major, tx_pos [3],
 *   rx_pos [R,3]; radii_out [R,P] int32. */
int gsparc_rasterize_forward_geom(int P, int R, int W, int H, const float* means, const float* scales,
                                  const float* rotations, const float* opacities, const float* emission_mlps,
                                  const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                  float scale_modifier, void* geom, void* image, int* radii_out,
                                  uint64_t* K_out, void* stream);

/* Forward, part 2: duplicate instances into tile segments, sort every segment
 * by depth and blend front to back.  emission: NULL (use the MLP output of
 * part 1) or [R,P,2] (real, imag) per receiver and Gaussian.  K_capacity is
 * the instance capacity of `bin`.  Outputs out_real / out_imag [R,H,W].
 * Replaces: duplicateWithKeys, SortPairs, identifyTileRanges and
 * FORWARD::render (rasterizer_impl.cu:102-171, 280-333; forward.cu:151-249). */
int gsparc_rasterize_forward_render(int P, int R, int W, int H, const float* emission, void* geom, void* image,
                                    void* bin, uint64_t K_capacity, const int* radii, float* out_real,
                                    float* out_imag, int flags, void* stream);

/* The whole forward in one call with no host synchronisation: parts 1 and 2
 * back to back, the binning sized by K_capacity (below 2^32 - 1) instead of
 * by K.  K stays on the device; if it exceeds K_capacity the tile lists are
 * not written, the blend writes NaN to out_real / out_imag (and final T = 1,
 * no contributors) and the backward leaves every gradient at zero, which
 * gsparc_instance_count() reveals.  Everything it launches can be
 * captured in a HIP graph.  No counterpart in the reference (which reads K
 * back before sizing its binning buffer, rasterizer_impl.cu:270-278); for
 * callers that cannot synchronise: graph-captured training steps. */
int gsparc_rasterize_forward(int P, int R, int W, int H, const float* means, const float* scales,
                             const float* rotations, const float* opacities, const float* emission_mlps,
                             const float* emission, const float* viewmatrices, const float* tx_pos,
                             const float* rx_pos, float scale_modifier, void* geom, void* image, void* bin,
                             uint64_t K_capacity, int* radii_out, float* out_real, float* out_imag, int flags,
                             void* stream);

/* K of the last forward into `image` (synchronises the stream).  The device
 * sums K in 64 bits and keeps it saturated at 2^32 - 1, which stands for
 * "2^32 - 1 or more": above every admissible capacity, so such a batch reads
 * as an overflow, and gsparc_rasterize_forward_geom fails with
 * GSPARC_ERR_CAPACITY on it. */
int gsparc_instance_count(int P, int R, int W, int H, const void* image, uint64_t* K, void* stream);

/* Backward.  grad_mode: GSPARC_GRAD_EXACT or GSPARC_GRAD_REFERENCE.
 * Inputs as in the forward plus dL/dout_real, dL/dout_imag [R,H,W];
 * K_capacity is the instance capacity `bin` was laid out with.
 * Outputs (all written, none accumulated):
 *   dmeans2D [R,P,3] (screen-space, reference layout; may be NULL),
 *   dmeans3D [P,3], demission_mlps [P*194] (NULL if no MLP; 16-byte aligned),
 *   dopacity [P], dcov3D [P,6] (may be NULL), dscales [P,3], drotations [P,4],
 *   demission [R,P,2] (emission supplied to the forward; may be NULL).
 * Gradients are summed over the R receivers.
 * Replaces: CudaRasterizer::Rasterizer::backward (rasterizer_impl.cu:339-442),
 * BACKWARD::render (backward.cu:442-616, 701-740) and BACKWARD::preprocess
 * (backward.cu:124-437, 618-699) behind RasterizeGaussiansBackwardCUDA
 * (rasterize_points.cu:106-199). */
int gsparc_rasterize_backward(int grad_mode, int P, int R, int W, int H, const float* means, const float* scales,
                              const float* rotations, const float* emission_mlps, const float* emission,
                              const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                              float scale_modifier, const int* radii, void* geom, void* image, void* bin,
                              uint64_t K_capacity, void* grad, const float* dout_real, const float* dout_imag, float* dmeans2D,
                              float* dmeans3D, float* demission_mlps, float* dopacity, float* dcov3D,
                              float* dscales, float* drotations, float* demission, void* stream);

/* Visibility test r^2 > 0.04 (checkFrustum, rasterizer_impl.cu:173-202;
 * markVisible, rasterize_points.cu:203-221).  present [P] uint8. */
int gsparc_mark_visible(int P, const float* means, const float* viewmatrix, uint8_t* present, void* stream);

/* Fused SSIM over `planes` planes of H x W (B*CH in the reference).
 * Forward writes the SSIM map and, when dm_dmu1 != NULL, the three partial
 * derivative maps.  Replaces fusedssim (fused-ssim/ssim.cu:368-404). */
int gsparc_ssim_forward(int planes, int H, int W, float C1, float C2, const float* img1, const float* img2,
                        float* ssim_map, float* dm_dmu1, float* dm_dsigma1_sq, float* dm_dsigma12, void* stream);

/* dL/dimg1.  Replaces fusedssim_backward (fused-ssim/ssim.cu:406-444). */
int gsparc_ssim_backward(int planes, int H, int W, float C1, float C2, const float* img1, const float* img2,
                         const float* dL_dmap, const float* dm_dmu1, const float* dm_dsigma1_sq,
                         const float* dm_dsigma12, float* dL_dimg1, void* stream);

/* ---- multi-frequency (CSI) maps -------------------------------------------
 * 52 complex channels per pixel instead of one.  A (receiver, Gaussian)
 * record with complex amplitude s (the MLP output, or `emission`) and path
 * length L = |tx - mu| + |mu - rx| emits s * exp(-j 2 pi f_k L / c) on
 * subcarrier k = 0..51, f_k = f0_hz + k * df_hz (3.5 GHz centre, 312.5 kHz
 * spacing: f0 = 3.5e9 - 25.5 * 312.5e3, df = 312.5e3); c = 299792458 m/s.
 * out / dout: [R, H, W, 52, 2] float (real, imaginary interleaved).  The
 * geometry, binning and buffers are those of the single-channel path:
 * gsparc_rasterize_forward_geom then gsparc_rasterize_forward_render_csi, or
 * gsparc_rasterize_forward_csi with a capacity.  The backward is the exact
 * adjoint (grad_mode EXACT), including dL/dL to the means.  No counterpart in
 * the reference, which renders one channel (config.h:15) one receiver at a
 * time (renderer.py:39-88). */
int gsparc_rasterize_forward_render_csi(int P, int R, int W, int H, double f0_hz, double df_hz,
                                        const float* emission, void* geom, void* image, void* bin,
                                        uint64_t K_capacity, const int* radii, float* out, void* stream);
int gsparc_rasterize_forward_csi(int P, int R, int W, int H, double f0_hz, double df_hz, const float* means,
                                 const float* scales, const float* rotations, const float* opacities,
                                 const float* emission_mlps, const float* emission, const float* viewmatrices,
                                 const float* tx_pos, const float* rx_pos, float scale_modifier, void* geom,
                                 void* image, void* bin, uint64_t K_capacity, int* radii_out, float* out,
                                 void* stream);
int gsparc_rasterize_backward_csi(int P, int R, int W, int H, double f0_hz, double df_hz, const float* means,
                                  const float* scales, const float* rotations, const float* emission_mlps,
                                  const float* emission, const float* viewmatrices, const float* tx_pos,
                                  const float* rx_pos, float scale_modifier, const int* radii, void* geom,
                                  void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout,
                                  float* dmeans2D, float* dmeans3D, float* demission_mlps, float* dopacity,
                                  float* dcov3D, float* dscales, float* drotations, float* demission, void* stream);

/* Loss of a CSI map against a target map (both [R, H, W, 52, 2]):
 *   (1 - lambda_ssim) * mean |out - target| over all real and imaginary parts
 *   + lambda_ssim * (1 - mean SSIM of the magnitude maps |C_k|),
 * one SSIM plane per (receiver, subcarrier).  loss [1]; dL_dout (may be NULL)
 * like out; work: gsparc_csi_loss_work_size bytes.  Carries the loss
 * composition of train.py:418-427 over to the complex map. */
size_t gsparc_csi_loss_work_size(int R, int H, int W);
int gsparc_csi_loss(int R, int H, int W, float lambda_ssim, float C1, float C2, const float* out, const float* target,
                    void* work, float* loss, float* dL_dout, void* stream);

/* The multi-frequency kernels' phase factors: cos_out[i], sin_out[i] =
 * cos 2 pi x, sin 2 pi x of x = turns[i], computed exactly as the CSI blend
 * kernels compute them (the hardware v_cos_f32 / v_sin_f32, which take the
 * angle in turns), for the accuracy check of that routine over the cycle
 * fractions the kernels pass it ([-1/2, 1/2]).  No counterpart in the
 * reference (one channel, config.h:15). */
int gsparc_csi_phase_factors(size_t n, const float* turns, float* cos_out, float* sin_out, void* stream);

/* The multi-frequency backward's phasors exp(-j 2 pi L f_k / c), k = 0..51,
 * f_k = f0_hz + k df_hz, for n path lengths L (metres): out is n x 52 x 2
 * floats (real, imaginary).  They come from the exact phase of subcarriers 0
 * and 26 and a complex recurrence over the subcarrier spacing, which is what
 * this entry point exposes for its accuracy check.  No counterpart in the
 * reference (one channel, config.h:15). */
int gsparc_csi_phasors(size_t n, const float* path_len, double f0_hz, double df_hz, float* out, void* stream);

/* ---- shared emission MLP on the matrix cores --------------------------------
 * The 68 -> 128 -> 128 -> 1 ReLU MLP of the single-receiver training
 * configuration, for every (receiver r, Gaussian g) pair:
 *   x = [PE(mu_g) (63), d, 1 / max(d^2, 0.04), (tx - mu_g) / |tx - mu_g|],
 *       d = |rx_r - mu_g|; PE = mu, then sin(pi x), cos(pi x) of x = 2^k mu
 *       reduced exactly to [-1, 1], k = 0..9
 *   h1 = relu(W1 x + b1),  h2 = relu(W2 h1 + b2),  a = w3 . h2 + b3
 * means [P,3], rx_pos [R,3], tx_pos [3], W1 [128,68], b1 [128], W2 [128,128]
 * (out, in), b2 [128], w3 [128], b3 [1] (device).  emission [R,P,2] receives
 * (a, 0): the rasterizer's emission input.  The first layer is split: its
 * PE part runs once per Gaussian, the 5 receiver features per pair; the
 * layers run as FP32 MFMA tiles with FP32 accumulation.  work:
 * gsparc_mlp1_work_size bytes, the same buffer for the forward and the
 * backward (the backward reads what the forward left there).  Backward:
 * demission [R,P,2] in (the real part is dL/da); dmeans [P,3] = the MLP's
 * gradient of the means (through PE and the receiver features) and grads =
 * dW1 [128*68] | db1 [128] | dW2 [128*128] | db2 [128] | dw3 [128] | db3 [1]
 * out.  No counterpart in the reference, whose emission MLP is per Gaussian
 * (gaussian.py:47, forward.cu:17-75). */
size_t gsparc_mlp1_work_size(int P, int R);
int gsparc_mlp1_forward(int P, int R, const float* means, const float* rx_pos, const float* tx_pos, const float* W1,
                        const float* b1, const float* W2, const float* b2, const float* w3, const float* b3,
                        void* work, float* emission, void* stream);
int gsparc_mlp1_backward(int P, int R, const float* means, const float* rx_pos, const float* tx_pos, const float* W1,
                         const float* b1, const float* W2, const float* b2, const float* w3, const float* b3,
                         void* work, const float* demission, float* dmeans, float* grads, void* stream);
/* gsparc_mlp1_backward that adds the means gradient to dmeans (for instance
 * the rasterizer's dL/dmeans3D, giving the total) instead of overwriting it;
 * grads are written as there.  No counterpart in the reference. */
int gsparc_mlp1_backward_accumulate(int P, int R, const float* means, const float* rx_pos, const float* tx_pos,
                                    const float* W1, const float* b1, const float* W2, const float* b2,
                                    const float* w3, const float* b3, void* work, const float* demission,
                                    float* dmeans, float* grads, void* stream);

/* Training loss of the reference, value and gradient in one call:
 *   loss = (1 - lambda_ssim) * mean|img - target| + lambda_ssim * (1 - mean SSIM(img, target))
 * over `planes` planes of H x W (one per receiver spectrum), 11x11 Gaussian
 * window (sigma 1.5), zero padding ("same"), constants C1 / C2; loss [1] on
 * the device.  dL_dimg (may be NULL: loss only) receives the gradient, with
 * sign(0) = 0 for the L1 term.  work: gsparc_loss_work_size bytes.  The mean
 * runs over all planes, so a batch of R spectra gives the mean of their
 * per-spectrum losses.  Replaces the loss composition of train.py:418-427
 * (l1_loss, utils.py:40-41; fused_ssim, fused_ssim/__init__.py:34-41;
 * lambda1 = 0.2, optimizer.py:8), which runs as separate torch kernels there.
 * The confidence weighting of train.py:429-436 is left to the caller. */
size_t gsparc_loss_work_size(int planes, int H, int W);
int gsparc_loss_l1_ssim(int planes, int H, int W, float lambda_ssim, float C1, float C2, const float* img,
                        const float* target, void* work, float* loss, float* dL_dimg, void* stream);

/* The reference's confidence-weighted training loss, value and gradients in
 * one call, over `planes` spectra (one receiver each):
 *   L_r = (1 - lambda_ssim) mean|img_r - target_r| + lambda_ssim (1 - mean SSIM(img_r, target_r))
 *   C_r = 1 + exp(fc2 . relu(fc1 rx_r + b1) + b2)      (the confidence network)
 *   loss = mean_r (C_r L_r - alpha log C_r)
 * conf_mlp: the network packed as [fc1.weight hidden*3 | fc1.bias hidden |
 * fc2.weight hidden | fc2.bias 1] (torch Linear layouts), or NULL: C_r = 1,
 * the reference's value without a network.  rx_pos [planes, 3].  Outputs:
 * loss [1]; rec_loss [planes] = L_r and confidence [planes] = C_r (each may
 * be NULL); dL_dimg like img (may be NULL: no image gradient); dconf_mlp
 * like conf_mlp (may be NULL).  Sums run in a fixed order (deterministic).
 * work: gsparc_conf_loss_work_size bytes.  With one plane this is exactly
 * the reference's loss of one training step.  Replaces train.py:418-436
 * (reconstruction loss, DUSt3R-style confidence weighting) and the
 * ConfidenceMLP of gaussian.py:575-609 with its backward. */
size_t gsparc_conf_loss_work_size(int planes, int H, int W);
int gsparc_loss_l1_ssim_conf(int planes, int H, int W, float lambda_ssim, float C1, float C2, float alpha,
                             int hidden, const float* conf_mlp, const float* rx_pos, const float* img,
                             const float* target, void* work, float* loss, float* rec_loss, float* confidence,
                             float* dL_dimg, float* dconf_mlp, void* stream);

/* Evaluation metrics of `planes` rendered spectra img against target [planes,
 * H, W], one SSIM pass: metrics [planes][8] = loss (1 - lambda_ssim) * MAE +
 * lambda_ssim * (1 - SSIM), PSNR 20 log10(1 / sqrt(MSE)), SSIM (mean of the
 * 11 x 11 Gaussian-window map with zero padding, utils.py:81-113), MAE, MSE,
 * RSSI of img 10 log10(mean img^2) - 30, RSSI(img) - RSSI(target), and the
 * confidence 1 + exp(...) of the packed network conf_mlp at rx_pos (1 when
 * conf_mlp is NULL).  Sums in double, in a fixed order.  work:
 * gsparc_eval_work_size bytes.  Replaces the per-sample metric calls of
 * Trainer.evaluate (train.py:76-102, 651-693). */
size_t gsparc_eval_work_size(int planes, int H, int W);
int gsparc_eval_metrics(int planes, int H, int W, float lambda_ssim, float C1, float C2, int hidden,
                        const float* conf_mlp, const float* rx_pos, const float* img, const float* target, void* work,
                        float* metrics, void* stream);

/* gsparc_loss_l1_ssim with no workspace argument, calling no allocator: its
 * scratch (the three SSIM partial-derivative maps and the per-workgroup loss
 * partials) is a library-owned device array of GSPARC_LOSS_SCRATCH_BYTES on
 * every GPU the library runs on, held until the process ends; its address is
 * looked up once per device, on the first call there.  Sizes whose scratch
 * would not fit (gsparc_loss_work_size(planes, H, W) - 256 bytes above it:
 * about 2.79M pixels, e.g. 86 spectra of 90 x 360) fail with GSPARC_ERR_ARG.
 * Every call on one GPU shares that GPU's array and nothing in it needs
 * clearing, but calls that overlap on one GPU (other streams, other threads)
 * race: serialise them.  Results are bit-identical to gsparc_loss_l1_ssim's.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code being captured into a HIP graph, serving loops, harnesses
 * that forbid allocation). */
#define GSPARC_LOSS_SCRATCH_BYTES ((size_t)32 << 20)
int gsparc_loss_l1_ssim_noalloc(int planes, int H, int W, float lambda_ssim, float C1, float C2, const float* img,
                                const float* target, float* loss, float* dL_dimg, void* stream);

/* HIP graphs of this library's launch sequences.  Between begin and end,
 * every call above that does not synchronise (gsparc_rasterize_forward with a
 * capacity, gsparc_rasterize_backward, the SSIM and loss calls, the _scenes
 * variants without _geom) is recorded on `stream` instead of run; end
 * instantiates the graph, launch replays it on any stream with the pointers
 * and sizes given at capture.  The calls that read K back
 * (gsparc_rasterize_forward_geom, gsparc_instance_count) fail while
 * capturing.  Profiling must be off.  Capture is thread-local: the capturing
 * thread may not allocate device memory in between.  No counterpart in the
 * reference (one synchronising launch sequence per call,
 * rasterizer_impl.cu:206-335); for training loops that replay one step. */
typedef struct gsparc_graph gsparc_graph;
int gsparc_graph_begin(void* stream);
int gsparc_graph_end(void* stream, gsparc_graph** graph);
int gsparc_graph_launch(gsparc_graph* graph, void* stream);
int gsparc_graph_destroy(gsparc_graph* graph);

/* Per-kernel timing and counters.  gsparc_profile_enable(mode): 0 off;
 * 2 timing: every kernel launched by the calls above is bracketed by HIP
 * events on its own stream; collect() waits for them, adds the elapsed times
 * to per-kernel totals and returns how many kernels have totals; query()
 * reads entry i (sorted by kernel name); 3 counting: the blend kernels run
 * their counting variants (slower: extra atomics) and count, in the
 * reference's terms whatever this build culls (profile_counts): counts[0]
 * the (pixel, list entry) pairs the reference's forward visits (each pixel's
 * list up to its terminating entry, else all of it; forward.cu:198-236),
 * counts[1] the pairs its backward walks (the sum of the contributor counts;
 * backward.cu:547-563), counts[2] / counts[3] the contributing pairs (alpha
 * >= 1/255, before termination) as the forward / backward found them; 1 both
 * (the timings then include the counting).  No counterpart in the reference (which only times render()
 * on the host, train.py:381-384). */
int gsparc_profile_enable(int on);
int gsparc_profile_collect(int* n_kernels);
int gsparc_profile_query(int i, char* name, int name_len, double* total_ms, long long* launches);
int gsparc_profile_counts(uint64_t counts[4]);

/* ---- batches of uneven scenes --------------------------------------------
 * S independent scenes in one launch sequence: scene s has P_s Gaussians
 * (packed one scene after the other in means / scales / rotations /
 * opacities, and in each section of emission_mlps) and R_s receivers (packed
 * likewise in viewmatrices and rx_pos); tx_pos is [S,3].  Every receiver
 * sees only its own scene.  Per-(receiver, Gaussian) arrays (radii,
 * emission, dmeans2D, demission) hold Nrec = sum_s R_s P_s records, receiver
 * after receiver; maps are [R,H,W] with R = sum_s R_s.
 * gsparc_scene_table_build fills a host table (gsparc_scene_table_words
 * words) and the batch totals; the caller copies the table to the device
 * and passes that copy.  No counterpart in the reference (one receiver of
 * one scene per call, renderer.py:39-88); for callers that train or serve
 * many scenes of different sizes at once. */
typedef struct gsparc_batch {
    int S, R, G, Pmax, Ptot;   /* scenes, receivers, receiver groups, largest and total Gaussian counts */
    uint64_t Nrec, Npart;      /* (receiver, Gaussian) records; backward partial-sum slots */
} gsparc_batch;

size_t gsparc_scene_table_words(int S, int R);
int gsparc_scene_table_build(int S, const int* scene_P, const int* scene_R, uint32_t* table, size_t table_words,
                             gsparc_batch* batch);
/* sizes[0..3] = geom, image, bin (for K_capacity instances) and grad bytes */
int gsparc_scene_buffer_sizes(const gsparc_batch* batch, int W, int H, uint64_t K_capacity, size_t* sizes);
/* As gsparc_rasterize_forward (no host synchronisation; K stays on the device) */
int gsparc_rasterize_forward_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                    const float* means, const float* scales, const float* rotations,
                                    const float* opacities, const float* emission_mlps, const float* emission,
                                    const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                    float scale_modifier, void* geom, void* image, void* bin, uint64_t K_capacity,
                                    int* radii_out, float* out_real, float* out_imag, int flags, void* stream);
/* As gsparc_rasterize_forward_geom / gsparc_rasterize_forward_render (K read back) */
int gsparc_rasterize_forward_geom_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                         const float* means, const float* scales, const float* rotations,
                                         const float* opacities, const float* emission_mlps,
                                         const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                         float scale_modifier, void* geom, void* image, int* radii_out,
                                         uint64_t* K_out, void* stream);
int gsparc_rasterize_forward_render_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                           const float* emission, void* geom, void* image, void* bin,
                                           uint64_t K_capacity, const int* radii, float* out_real,
                                           float* out_imag, int flags, void* stream);
int gsparc_instance_count_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                 const void* image, uint64_t* K, void* stream);
/* As gsparc_rasterize_backward; dmeans2D [Nrec,3], demission [Nrec,2] */
int gsparc_autotune_scenes(int grad_mode, const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                           const float* emission, const int* radii, void* geom, void* image, void* bin,
                           uint64_t K_capacity, void* grad, const float* dout_real, const float* dout_imag,
                           float* out_real, float* out_imag, int iters, int* choice, float* ms, float* table,
                           void* stream);
int gsparc_rasterize_backward_scenes(int grad_mode, const gsparc_batch* batch, const uint32_t* scene_table, int W,
                                     int H, const float* means, const float* scales, const float* rotations,
                                     const float* emission_mlps, const float* emission, const float* viewmatrices,
                                     const float* tx_pos, const float* rx_pos, float scale_modifier,
                                     const int* radii, void* geom, void* image, void* bin, uint64_t K_capacity,
                                     void* grad, const float* dout_real, const float* dout_imag, float* dmeans2D,
                                     float* dmeans3D, float* demission_mlps, float* dopacity, float* dcov3D,
                                     float* dscales, float* drotations, float* demission, void* stream);

/* ---- the training step around the rasterizer ------------------------------
 * The per-Gaussian parameter set of a training run and its optimizer state
 * (gaussian.py:33-38, 122-197), device pointers, P entries each:
 *   centre [P,3], log_scale [P,3], rotation [P,4] (raw quaternion), logit [P],
 *   emission_mlps [P*194] (W1 [P,32,3] | b1 [P,32] | W2 [P,2,32] | b2 [P,2]);
 *   the activations the rasterizer reads: scale [P,3] = exp(log_scale),
 *   rotation_n [P,4] = q / max(|q|, 1e-12), opacity [P] = sigmoid(logit);
 *   Adam moments state [2][11 P] (exp_avg then exp_avg_sq, each centre 3P |
 *   log_scale 3P | rotation 4P | logit P) and state_mlp [2][194 P];
 *   densification statistics accum [P] (centre_gradient_accum), denom [P],
 *   max_radius [P] (max_radii2D).
 * A set laid out for P entries fits in arrays allocated for any capacity
 * >= P (every section starts at a multiple of P). */
typedef struct gsparc_gaussians {
    float *centre, *log_scale, *rotation, *logit, *emission_mlps;
    float *scale, *rotation_n, *opacity;
    float *state, *state_mlp;
    float *accum, *denom, *max_radius;
} gsparc_gaussians;

/* Adam parameter groups (training_setup, gaussian.py:175-197); the four
 * emission MLP groups step together, as do the confidence network's four. */
#define GSPARC_GROUP_CENTRE 1
#define GSPARC_GROUP_SCALING 2
#define GSPARC_GROUP_ROTATION 4
#define GSPARC_GROUP_OPACITY 8
#define GSPARC_GROUP_EMISSION 16
#define GSPARC_GROUP_CONFIDENCE 32
#define GSPARC_GROUPS_ALL 63

/* torch.optim.Adam(eps=1e-15) settings and the centre learning-rate schedule
 * get_expon_lr_func(lr_centre_init, lr_centre_final, lr_centre_delay_steps,
 * lr_centre_delay_mult, lr_centre_max_steps) (utils.py:46-79, gaussian.py:
 * 197-221); clamp_centres with scene_min / scene_max: the clamp after the
 * step (train.py:558-562). */
typedef struct gsparc_adam_config {
    double beta1, beta2, eps;
    double lr_centre_init, lr_centre_final, lr_centre_delay_mult;
    int lr_centre_delay_steps, lr_centre_max_steps;
    double lr_scaling, lr_rotation, lr_opacity, lr_emission, lr_confidence;
    int clamp_centres;
    float scene_min[3], scene_max[3];
} gsparc_adam_config;

/* Device training state of GSPARC_TRAIN_STATE_BYTES: uint32 words [0] the
 * iteration being run (the reference's 1-based `iteration`; the caller sets
 * it, each gsparc_adam_step advances it), [1..6] the steps taken by each
 * group (torch's state["step"]), [7] set to 1 when a step leaves a NaN centre
 * (train.py:565-566); the rest belongs to the library.  Start from zeros. */
#define GSPARC_TRAIN_STATE_BYTES 256

/* Write the activations (scale, rotation_n, opacity) from the raw
 * parameters.  Replaces the get_scaling / get_rotation / get_opacity
 * properties (gaussian.py:86-96). */
int gsparc_gaussians_activate(int P, const gsparc_gaussians* g, void* stream);

/* One optimizer step of the groups in `groups` (the others are left as they
 * are and their step counts stay): gradients with respect to the
 * activations (the rasterizer's dmeans3D, dscales, drotations, dopacity,
 * demission_mlps) are carried through exp / normalize / sigmoid, NaN entries
 * of the centre gradient are zeroed (train.py:441-444), Adam updates the raw
 * parameters, the activations are rewritten and, with clamp_centres, the
 * centres are clamped (also when the centre group does not step).  The
 * confidence network (n_conf floats at conf, its gradient dconf and moments
 * conf_state [2][n_conf]; n_conf may be 0) steps with GSPARC_GROUP_CONFIDENCE.
 * Records no synchronisation: can be captured in a graph.  Replaces
 * update_learning_rate (gaussian.py:214-221), optimizer.step over the groups
 * of training_setup (gaussian.py:175-197; torch.optim.Adam) and the NaN guard
 * and clamp around it (train.py:441-444, 558-566). */
int gsparc_adam_step(int P, int groups, const gsparc_adam_config* cfg, const gsparc_gaussians* g,
                     const float* dcentre, const float* dscale, const float* drotation, const float* dopacity,
                     const float* demission_mlps, int n_conf, float* conf, const float* dconf, float* conf_state,
                     void* train_state, void* stream);

/* Densification statistics of one batch: for every receiver r that sees
 * Gaussian i (radii[r,i] > 0): accum[i] += |dmeans2D[r,i,:2]| * grad_scale,
 * denom[i] += 1, max_radius[i] = max(max_radius[i], radii[r,i]).  dmeans2D
 * [R,P,3] and radii [R,P] as the backward and forward write them;
 * grad_scale turns the gradient of a batch loss into the per-view gradient
 * (R for a loss averaged over R spectra, 1 for one spectrum).  Replaces
 * add_densification_stats (gaussian.py:538-551) and the max-radius update
 * (train.py:510-516). */
int gsparc_densify_stats(int P, int R, float grad_scale, const float* dmeans2D, const int* radii,
                         const gsparc_gaussians* g, void* stream);

/* reset_opacity (gaussian.py:223-231): logit = inverse_sigmoid(min(opacity,
 * 0.01)) (non-finite results mapped as torch.nan_to_num(nan=0, posinf=1,
 * neginf=0)), opacity rewritten, the logit's Adam moments zeroed (its step
 * count stays, replace_tensor_to_optimizer, gaussian.py:233-253). */
int gsparc_opacity_reset(int P, const gsparc_gaussians* g, void* stream);

/* ---- the spectrum dataset on the device ----------------------------------
 * View matrices of N receivers at rx_pos [N,3] under the gateway orientation
 * gateway_quat (x, y, z, w, as in gateway_info.yml; normalised here):
 * viewmatrices [N,16] = [R^T, -R^T C; 0 1] stored transposed (column-major),
 * R the quaternion's rotation computed in double and rounded to float.
 * Replaces Spectrum_dataset.build_view_matrix (preprocess_data.py:177-202),
 * one call for the whole dataset. */
int gsparc_view_matrices(int N, const double* gateway_quat, const float* rx_pos, float* viewmatrices, void* stream);

/* The batch of the iteration stored in train_state (word 0, 1-based, as
 * gsparc_adam_step advances it): sample r = order[((iteration - 1) R + r)
 * mod order_len] of the resident dataset (spectra [N, npix], rx_pos [N,3],
 * viewmatrices [N,16]) is copied to row r of batch_spectra [R, npix],
 * batch_rx [R,3], batch_views [R,16].  Reads the iteration on the device, so
 * a captured training step replays with a new batch each time.  Indices
 * outside [0, N) are clamped.  Replaces the per-iteration
 * wireless_data_train.pop() of train.py:366-370 for R samples at a time. */
int gsparc_gather_batch(int R, int N, int npix, const int* order, int order_len, const void* train_state,
                        const float* spectra, const float* rx_pos, const float* viewmatrices, float* batch_spectra,
                        float* batch_rx, float* batch_views, void* stream);

/* densify_and_prune settings: densify != 0 clones (mean gradient accum/denom
 * >= grad_threshold and max scale <= clone_max_scale, the percent_dense *
 * extent of gaussian.py:495) and splits (same gradient test, max scale
 * above) before pruning; densify == 0 only prunes (the post-densification
 * pruning of train.py:539-555).  Prune: opacity < min_opacity, max scale >
 * prune_max_scale (0.1 * extent when the reference passes max_screen_size,
 * else +inf), with use_bounds a centre outside [scene_min, scene_max].  The
 * thresholds are compared in float, as torch compares a float tensor with a
 * Python number: pass them rounded from double. */
typedef struct gsparc_densify_config {
    int densify;
    float grad_threshold, clone_max_scale, min_opacity, prune_max_scale;
    int use_bounds;
    float scene_min[3], scene_max[3];
} gsparc_densify_config;

size_t gsparc_densify_work_size(int P);

/* Densify and prune src (P Gaussians) into dst (distinct arrays allocated for
 * dst_capacity Gaussians), in the reference's order: the originals that are
 * not split, one clone per clone-selected original, the first child of every
 * split original, then the second, each run in original order; then the
 * pruned ones removed.  A split child of mu, scale s and quaternion q has
 * centre R(q) (s * z) + mu with z row `child` of noise (standard normals
 * [2 P, 3]: the first child of the k-th split original reads row k, the
 * second row n_split + k) and log-scale log(s / 1.2).  Kept originals carry
 * their Adam moments, new Gaussians start with zeros; the statistics of all
 * restart at zero.  Synchronises the stream once (the new count); when
 * P_new exceeds dst_capacity nothing is written to dst and the call returns
 * GSPARC_ERR_CAPACITY with *P_new, *n_clone and *n_split set: call again
 * with room.  Not while capturing.  Replaces densify_and_prune
 * (gaussian.py:510-535) with densify_and_clone (493-508), densify_and_split
 * (450-491), densification_postfix (395-429), prune_points (308-337) and the
 * optimizer-state surgery of cat_tensors_to_optimizer / _prune_optimizer
 * (266-393); with densify == 0, the pruning of train.py:539-555. */
int gsparc_densify_and_prune(int P, const gsparc_densify_config* cfg, const gsparc_gaussians* src,
                             const float* noise, const gsparc_gaussians* dst, int dst_capacity, void* work,
                             int* P_new, int* n_clone, int* n_split, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GSPARC_H */
