/*
 * fpie_hip.h - C ABI of the MI355X (gfx950) Poisson image-editing solver.
 *
 * The entry points replace the pybind11 solver cores that
 * fpie/process.py (Trinkle23897/Fast-Poisson-Image-Editing, fpie 0.3.2)
 * drives through duck typing (process.py:161-176, 187, 270, 275, 300-317,
 * 385, 390):
 *
 *   pybind11 method (fpie/core/gcc/solver.cc:3-15,
 *                    fpie/core/base_solver.h)             C entry point
 *   core_cuda.EquSolver(block_size)   cuda/solver.cc      pie_create(PIE_EQU,...)
 *   core_cuda.GridSolver(gx, gy)      cuda/solver.cc      pie_create(PIE_GRID,...)
 *   EquSolver.partition(mask)         gcc/equ.cc:14-32    pie_partition
 *   EquSolver.reset(N, A, X, B)       base_solver.h:27-59 pie_reset_equ
 *   GridSolver.reset(N, mask, t, g)   base_solver.h:101-140 pie_reset_grid
 *   *.sync()                          base_solver.h:61,142 pie_sync
 *   *.step(iteration) -> (img, err)   gcc/equ.cc:74-86,
 *                                     gcc/grid.cc:57-81   pie_step
 *   (object destruction)              base_solver.h:19-25 pie_destroy
 *   (exceptions -> RuntimeError)      base_solver.h:63-74 status + pie_last_error
 *
 * All pointers are host pointers to plain C arrays; the library copies what
 * it needs (reset deep-copies, as base_solver.h:35-56 does) and writes
 * outputs into caller-owned buffers.  Every call returns PIE_OK (0) or a
 * negative status; pie_last_error() then returns a message for the calling
 * thread.  One solver object must not be used from two threads at once.
 */
#ifndef FPIE_HIP_H_
#define FPIE_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pie_solver pie_solver;

/* solver kinds: the reference's EquSolver (A/X/B system) and GridSolver */
#define PIE_EQU 0
#define PIE_GRID 1

/* update schedules
 *  PIE_JACOBI   every masked pixel from the previous sweep
 *               (fpie/np_solver.py:31-41,79-88; openmp/grid.cc:75-103)
 *  PIE_GS_LEX   lexicographic Gauss-Seidel, the result of the gcc backend's
 *               in-place row-major loop (gcc/equ.cc:74-80, gcc/grid.cc:62-75),
 *               computed by a skewed two-colour wavefront schedule
 *  PIE_REDBLACK (i+j)-odd pixels, then (i+j)-even (openmp/equ.cc:107-118) */
#define PIE_JACOBI 0
#define PIE_GS_LEX 1
#define PIE_REDBLACK 2

/* order in which the four neighbours are added after b (Equ) or grad (Grid)
 *  PIE_ORDER_UDLR  up, down, left, right: every EquSolver (gcc/equ.cc:48-53)
 *                  and the numpy GridSolver (np_solver.py:83-88)
 *  PIE_ORDER_ULRD  up, left, right, down: gcc/openmp/cuda GridSolver
 *                  (gcc/grid.cc:23-35, openmp/grid.cc:35-43) */
#define PIE_ORDER_UDLR 0
#define PIE_ORDER_ULRD 1

/* status codes */
#define PIE_OK 0
#define PIE_EINVAL (-1)       /* bad argument, shape or index            */
#define PIE_EHIP (-2)         /* HIP runtime error (no device, OOM, ...) */
#define PIE_EUNSUPPORTED (-3) /* combination not offered                  */
#define PIE_ESTATE (-4)       /* call out of order (step before reset)    */

/* Create a solver on the current HIP device.  param_a/param_b are the
 * reference constructor arguments (EquSolver: block_size, unused;
 * GridSolver: grid_x, grid_y, unused: the GPU tile is autotuned).
 * Replaces core_cuda.EquSolver(block_size) / GridSolver(grid_x, grid_y)
 * (fpie/core/cuda/solver.cc:4-15, fpie/process.py:174,313). */
int pie_create(int kind, int schedule, int order, int param_a, int param_b,
               pie_solver** out);

/* Release every device and host buffer (base_solver.h:19-25). */
int pie_destroy(pie_solver* s);

/* EquSolver.partition (fpie/core/gcc/equ.cc:14-32): ids[i*w+j] = 1, 2, ...
 * in row-major order over mask > 0, 0 elsewhere.  Records the geometry so
 * that pie_reset_equ can solve the system on the grid layout. */
int pie_partition(pie_solver* s, const int32_t* mask, int h, int w,
                  int32_t* ids);

/* EquSolver.reset(N, A, X, B) (fpie/core/base_solver.h:27-59): A int32[N][4]
 * (ids of the up, down, left, right neighbours, 0 = constant), X and B
 * float[N][3].  Every id in A must lie in [0, N). */
int pie_reset_equ(pie_solver* s, int n, const int32_t* A, const float* X,
                  const float* B);

/* GridSolver.reset(N, mask, tgt, grad) (fpie/core/base_solver.h:101-140):
 * mask int32[h][w] (non-zero = solve), tgt and grad float[h][w][3].  The
 * outermost rows and columns must be unmasked (fpie/process.py:214-217
 * guarantees it; the reference reads out of bounds otherwise). */
int pie_reset_grid(pie_solver* s, int h, int w, const int32_t* mask,
                   const float* tgt, const float* grad);

/* GridSolver.reset built on the device from the uint8 crops: what
 * GridProcessor.reset (fpie/process.py:352-386) computes on the host and
 * then passes to pie_reset_grid.  src and tgt uint8[h][w][3] are the source
 * and target crops, mask int32[h][w] the prepared crop mask; the state is
 * the target as float and the guidance, at each solved pixel, the sum over
 * the up, down, left, right neighbours of mixgrad(src_c - src_n,
 * tgt_c - tgt_n) (process.py:113-122, 362-378), 0 elsewhere.  Integer and
 * half-integer values: identical to the host computation bit for bit. */
#define PIE_GRAD_SRC 0
#define PIE_GRAD_AVG 1
#define PIE_GRAD_MAX 2
int pie_reset_grid_images(pie_solver* s, int h, int w, const int32_t* mask,
                          const uint8_t* src, const uint8_t* tgt,
                          int gradient);

/* EquSolver.partition + EquSolver.reset built on the device from the uint8
 * crops: what EquProcessor.reset (fpie/process.py:180-190, 227-270)
 * computes on the host and passes to pie_partition / pie_reset_equ.  Ids
 * are row-major from 1 over mask > 0 (fpie/core/gcc/equ.cc:14-32),
 * N = count + 1, X[0] = 0, X[id] = tgt, B[id] = the mixed-gradient sum of
 * pie_reset_grid_images plus the target of every unsolved up, left, right,
 * down neighbour.  Integer and half-integer values: identical to the host
 * system bit for bit, for every schedule.  pie_step then returns N x 3
 * values in id order. */
int pie_reset_equ_images(pie_solver* s, int h, int w, const int32_t* mask,
                         const uint8_t* src, const uint8_t* tgt,
                         int gradient);

/* The processors' mask preparation (fpie/process.py:209-221 and 338-349),
 * computed on the device: a pixel of mask uint8[h][w][channels] is solved
 * when the mean of its channels is >= 128, the outermost rows and columns
 * are cleared, and the crop is the bounding box of the solved pixels plus
 * a 1-pixel margin.  box[4] = x0, x1, y0, y1 (the crop window [x0, x1) x
 * [y0, y1) in mask coordinates); crop (may be NULL) receives the int32
 * (x1 - x0) x (y1 - y0) cropped mask (0/1) if crop_capacity (int32 values)
 * holds it.  PIE_EINVAL if no pixel is solved. */
int pie_prepare_mask(const uint8_t* mask, int h, int w, int channels,
                     int* box, int32_t* crop, long long crop_capacity);

/* The output of EquProcessor.step (fpie/process.py:273-280: the solved
 * values clamped and truncated to uint8, fpie/core/gcc/equ.cc:82-85,
 * scattered into the target) as the h x w x 3 crop, assembled on the
 * device: the clamped state at solved pixels, the target crop elsewhere;
 * err as pie_step.  Needs pie_reset_equ_images (which keeps the target
 * crop); on a GridSolver it is pie_read. */
int pie_read_crop(pie_solver* s, uint8_t* crop, float* err);

/* sync() (base_solver.h:61): waits for queued work; no data moves. */
int pie_sync(pie_solver* s);

/* step(iteration) (gcc/equ.cc:74-86, gcc/grid.cc:57-81): run `iterations`
 * sweeps, then write the clamped/truncated image (Equ: uint8[N][3] in id
 * order; Grid: uint8[h][w][3]) and the per-channel sum of
 * |b + sum(neighbours) - 4x| over the solved pixels into err[3].  Either
 * output pointer may be NULL. */
int pie_step(pie_solver* s, int iterations, uint8_t* img, float* err);

/* Message of the last failed call on this thread ("" if none). */
const char* pie_last_error(void);

/* ---- extensions: no counterpart in the reference ---------------------- */

/* Queue `iterations` sweeps on the solver's stream and return at once. */
int pie_iterate(pie_solver* s, int iterations);
/* The outputs of pie_step for the current state, without iterating. */
int pie_read(pie_solver* s, uint8_t* img, float* err);
/* Float32 state: Equ float[N][3] in id order, Grid float[h][w][3]. */
int pie_get_state(pie_solver* s, float* out);
/* Choose the temporal-blocking tile: sweeps per launch, columns per lane
 * and rows per wave task (0 = keep the autotuned value). */
int pie_set_tile(pie_solver* s, int sweeps, int cols_per_lane, int seg_rows);
int pie_get_tile(pie_solver* s, int* sweeps, int* cols_per_lane,
                 int* seg_rows);
/* Colour channels per wave task of the tile: 3 (all), 1 (one channel a
 * task, three times the tasks), or 0 for the default of the tile's column
 * count (3 for one column, 1 for two).  The get call reports the channels
 * the tile uses. */
int pie_set_tile_channels(pie_solver* s, int channels);
int pie_get_tile_channels(pie_solver* s, int* channels);
/* Time every candidate tile on the loaded problem (`launches` launches
 * each, starting at steps spread over a call of 5000 sweeps; state
 * untouched) and keep the fastest.  pie_step / pie_iterate run the same
 * search on their first long call, over that call's own steps. */
int pie_autotune(pie_solver* s, int launches);
/* Event timing of the sweep kernels on the solver's stream. */
int pie_set_timing(pie_solver* s, int enable);
int pie_get_timing(pie_solver* s, double* kernel_ms, long long* launches,
                   double* pixel_updates);
/* Problem size: grid rows/cols, N (Equ ids incl. 0), solved pixels. */
int pie_shape(pie_solver* s, int* h, int* w, int* n, long long* masked);
/* Name and properties of the current device (cuda/utils.cu:5-22). */
int pie_device_info(char* buf, int len);
/* GridSolver: replace the state by x float[h][w][3] (a warm start; solved
 * pixels take x, unsolved pixels should hold the target values). */
int pie_set_state(pie_solver* s, const float* x);
/* Sweep in chunks of check_every until the mean absolute residual
 * err / (3 * solved pixels) is <= tol_mean or max_sweeps sweeps are done;
 * returns the sweeps run and the final mean residual. */
int pie_solve_to_residual(pie_solver* s, double tol_mean, int check_every,
                          long long max_sweeps, long long* sweeps,
                          double* mean_residual);

typedef struct pie_unet pie_unet;

/* ---- coarse-grid correction (GridSolver warm start): no counterpart in
 * the reference.  The coarse grid of `factor` is (h // factor) x
 * (w // factor); block (I, J) covers rows I*factor.., cols J*factor.. */
/* out float[hc][wc][3]: per block, the sum over its solved pixels of the
 * signed residual g + four neighbours - 4 x of the current state, in double
 * (the exact residual of the float state; rounded to float once). */
int pie_restrict(pie_solver* s, int factor, float* out);
/* state += cell-centred bilinear upsampling of corr float[hc][wc][3] at the
 * solved pixels; hc, wc must be the coarse grid of `factor`. */
int pie_prolong(pie_solver* s, int factor, int hc, int wc, const float* corr);
/* Line search for pie_prolong: with d the correction pie_prolong would add
 * (0 off the solved pixels), out[c] = sum r . d and out[3 + c] = sum d . A d
 * over the solved pixels, per channel c, in double (r the exact residual of
 * the state, (A d)_p = 4 d_p - sum of the neighbours' d); scaling corr by
 * out[c] / out[3 + c] minimises the energy norm of the new error. */
int pie_prolong_dots(pie_solver* s, int factor, int hc, int wc,
                     const float* corr, double* out);
/* `sweeps` weighted Jacobi sweeps x <- x + omega * ((g + sum) / 4 - x),
 * 0 < omega <= 1 (omega 1/2: the mean of two consecutive Jacobi iterates). */
int pie_relax(pie_solver* s, int sweeps, float omega);
/* The learned coarse correction on the device.  Restricts the residual
 * (pie_restrict), zeroes it off the coarse mask mc uint8[hc][wc] (outer
 * ring unsolved), then takes `net_steps` (<= 8) flexible conjugate-gradient
 * steps for A_c c = r_c (A_c the 5-point operator on mc) whose directions
 * are the network's answers to the current coarse residual r (input:
 * channels 0..2 = r per channel over its rms on mc, channel 6 = mc),
 * smoothed by `lowpass_sweeps` damped (weight 1/2) homogeneous sweeps,
 * A_c-orthogonalised against the earlier directions and taken with the
 * energy-optimal step; all in double.  net: a batch-1 pie_unet of shape
 * (hc rounded up to 8, wc rounded up to 8).  The float correction stays in
 * the solver for pie_warm_apply; corr_out (may be NULL) receives a copy. */
int pie_warm_net_start(pie_solver* s, pie_unet* net, int factor,
                       const uint8_t* mc, int net_steps, int lowpass_sweeps,
                       float* corr_out);
/* state += alpha_c * (bilinear prolongation of the correction), alpha_c =
 * max(0, out[c] / out[3 + c]) of pie_prolong_dots, with corr float[hc][wc][3]
 * from the host, or NULL for the one pie_warm_net_start left on the device
 * (an error if there is none for this factor).  alpha_out[3] receives
 * alpha, max_corr the largest |alpha_c corr| (as float). */
int pie_warm_apply(pie_solver* s, int factor, const float* corr,
                   double* alpha_out, double* max_corr);

/* ---- batched GridSolver: no counterpart in the reference ---------------
 * Many composites of uneven size and mask, each with its own stop rule,
 * solved together by one persistent Jacobi sweep kernel, R rounds per
 * hipGraph replay: items of few tiles are each run by one work-group,
 * several launches a round, the others share a work queue of tiles, one
 * launch a round.  An item stops after the first launch whose last sweep
 * changed no value by `tol` or more, or after `max_sweeps` sweeps. */
typedef struct pie_batch pie_batch;
/* order: PIE_ORDER_*; tile: sweeps per launch, columns per lane, rows per
 * task (0 = default 8, 1, 64). */
int pie_batch_create(int order, int sweeps_per_launch, int cols_per_lane,
                     int seg_rows, pie_batch** out);
/* Channels per wave task of the batch's sweep kernel (1 or 3; 0 = the
 * tile's default: 3 for one column per lane, 1 for two), as
 * pie_set_tile_channels; results are identical, only the split of the
 * work over waves differs.  Any time before pie_batch_solve. */
int pie_batch_set_tile_channels(pie_batch* b, int channels);
int pie_batch_destroy(pie_batch* b);
/* Time `rounds` rounds of every tile (sweeps, columns per lane, rows per
 * task, channels) on the batch from its initial state and keep the one with
 * the least time per sweep; the next pie_batch_solve starts over. */
int pie_batch_autotune(pie_batch* b, int rounds);
/* The batch's tile: sweeps per launch, columns per lane, rows per task and
 * channels per task (any pointer may be NULL). */
int pie_batch_tile(pie_batch* b, int* sweeps, int* cols_per_lane,
                   int* seg_rows, int* channels);
/* Add one GridSolver problem (arguments as pie_reset_grid); *index gets its
 * position in the batch. */
int pie_batch_add(pie_batch* b, int h, int w, const int32_t* mask,
                  const float* tgt, const float* grad, int max_sweeps,
                  float tol, int* index);
/* Solve every item from its initial state until all have stopped;
 * *seconds = wall time of the solve, *rounds_run = rounds launched. */
int pie_batch_solve(pie_batch* b, int rounds_per_graph, double* seconds,
                    int* rounds_run);
/* Sweeps done, max |update| of its last launch, and the time from the
 * start of the solve until the item stopped (device clock). */
int pie_batch_item(pie_batch* b, int i, int* sweeps, float* last_dmax,
                   double* latency_s);
/* Outputs of item i as pie_step / pie_get_state (any may be NULL). */
int pie_batch_read(pie_batch* b, int i, uint8_t* img, float* err,
                   float* state);
/* Items, wave tasks of one launch over every item (the tile's tasks whose
 * region holds a solved pixel) and solved pixels of the batch. */
int pie_batch_size(pie_batch* b, int* items, int* tasks, long long* masked);

/* ---- allocation-free solve: no counterpart in the reference ------------
 * For callers that cannot allocate device memory (serving loops that must
 * not call hipMalloc, harnesses that forbid it).  GridSolver reset + step +
 * read (pie_reset_grid, pie_step) in one call, on the caller's device
 * buffers: mask int32[h][w], tgt and grad float[h][w][3] in, img
 * uint8[h][w][3] and err float[3] out (all device pointers), `iterations`
 * sweeps of `schedule` in neighbour order `order`, queued on `stream` (a
 * hipStream_t; NULL = the default stream).  Every intermediate lives in a
 * 64 MiB device array of the library, one per GPU, kept until the process
 * ends: crops needing more (about 1000 x 1000 pixels and up) are rejected
 * with PIE_EINVAL.  The call allocates nothing, returns when the outputs
 * are written, and results equal pie_reset_grid + pie_step bit for bit.
 * Calls that overlap on one GPU race on the shared array: serialise them. */
int pie_grid_solve_noalloc(int h, int w, const int32_t* mask,
                           const float* tgt, const float* grad,
                           int iterations, int schedule, int order,
                           uint8_t* img, float* err, void* stream);

/* ---- warm-start U-Net: no counterpart in the reference -----------------
 * 4-level U-Net (base 32 channels, 3x3 convolutions, bf16 weights and
 * activations, fp32 accumulation on MFMA) mapping 8 input channels to 3
 * (the warm start: coarse residual scaled per channel in 0..2, coarse mask
 * in 6, the rest 0 -> a coarse correction direction).  Tensors are NHWC float32 on the host: x [B][H][W][8], target
 * and output [B][H][W][3], mask uint8 [B][H][W].  Parameters: one float32
 * array (pie_unet_param_count), per conv layer w[tap][cin][cout] then
 * b[cout], then the 1x1 head w[32][3], b[3]. */
/* batch >= 1; h, w multiples of 8.  Parameters start at zero. */
int pie_unet_create(int batch, int h, int w, pie_unet** out);
int pie_unet_destroy(pie_unet* u);
int pie_unet_param_count(pie_unet* u, long long* n);
int pie_unet_set_params(pie_unet* u, const float* p);
int pie_unet_get_params(pie_unet* u, float* p);
int pie_unet_forward(pie_unet* u, const float* x, float* out);
/* pie_unet_forward on device memory (no counterpart in the reference): x
 * and out are device pointers of a batch-1 network of shape (h, w); the
 * call returns when out is written. */
int pie_unet_forward_device(pie_unet* u, int h, int w, const float* x,
                            float* out);
/* Masked mean squared error over solved pixels and its parameter gradient
 * (grads may be NULL). */
int pie_unet_loss_grad(pie_unet* u, const float* x, const float* target,
                       const uint8_t* mask, float* loss, float* grads);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch. */
int pie_unet_train_step(pie_unet* u, const float* x, const float* target,
                        const uint8_t* mask, float lr, float* loss);
/* Pipelined training (no counterpart in the reference): two staging slots,
 * each a pinned host batch (x, target, mask in the layouts above) and its
 * device copy.  pie_unet_staging returns slot k's host buffers once its
 * last upload has finished (allocating them on first use); the caller
 * fills them and calls pie_unet_train_async, which enqueues the upload on
 * a copy stream and the Adam step on the network's stream and returns at
 * once.  pie_unet_finish waits for slot k's step and returns its loss; it
 * must be called before the slot's next pie_unet_train_async.  Filling and
 * uploading one slot overlap the other slot's step. */
int pie_unet_staging(pie_unet* u, int k, float** x, float** target,
                     uint8_t** mask);
int pie_unet_train_async(pie_unet* u, int k, float lr);
int pie_unet_finish(pie_unet* u, int k, float* loss);
/* Device time of the forward + backward passes so far and the convolution
 * FLOPs they did (for MFMA utilisation). */
int pie_unet_stats(pie_unet* u, double* ms, double* conv_flops);
/* Convolution tile policy of this process (all networks): layers with 128+
 * input channels use 16 x 16-pixel tiles when that still gives at least
 * min_blocks16 blocks (default 512), 8 x 16 otherwise; 0 = always, -1 =
 * never.  Results do not depend on it. */
int pie_unet_set_conv_tiles(int min_blocks16);

#ifdef __cplusplus
}
#endif

#endif /* FPIE_HIP_H_ */
