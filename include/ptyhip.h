/*
 * ptyhip -- C ABI of the MI355X (gfx950) ERA / difference-map engine.
 *
 * The reference (andyofmelbourne/Ptychography) has no native layer: its hot
 * path is the Python pair ERA / DM (ptychography/era.py:13, ptychography/dm.py:13)
 * and the NumPy helpers they call.  Each entry point below replaces one of
 * those pieces; the Python facade ptychography_amd.ERA / .DM keeps the
 * reference's signature and binds this library through ctypes.
 *
 * Conventions
 *   - Every function returns int status: PTY_OK (0) or a negative code; the
 *     message of the last failure on the calling thread is pty_last_error().
 *   - Pointers named *_dev are device pointers (for example from torch
 *     tensors on a ROCm device); *_host are host pointers.
 *   - Complex data is interleaved float32 (pty_cfloat).  Frame stacks are
 *     [N][M][K] row-major, the object [U][V], the probe [M][K].
 *   - Amplitudes and masks are in the reference's ifftshifted order (DC at
 *     [0][0], ptychography/era.py:497-499).
 *   - A context is not thread-safe; all its work is ordered on one stream.
 */
#ifndef PTYHIP_H
#define PTYHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pty_ctx pty_ctx;
typedef struct { float re, im; } pty_cfloat;

enum {
  PTY_OK = 0,
  PTY_EINVAL = -1,   /* bad argument: shape, pointer, positions outside the object */
  PTY_ENOMEM = -2,   /* device allocation failed */
  PTY_EHIP = -3,     /* a HIP runtime call or kernel launch failed */
  PTY_ESTATE = -4,   /* call order: data / state not bound yet */
  PTY_EUNSUP = -5    /* unsupported frame size */
};

/* Algorithm switches of one run (the keyword arguments of ERA / DM). */
typedef struct pty_opts {
  int32_t algorithm;          /* 0 = ERA (era.py:13), 1 = DM (dm.py:13) */
  int32_t method;             /* 1..6 as resolved by the preamble (era.py:440-456) */
  int32_t op_n, op_every;     /* OP_iters = (op_n, op_every) (era.py:458-459) */
  int64_t pmod_probe_until;   /* Pmod_P while iteration < this; 0 = never (era.py:193) */
  int32_t probe_centering;    /* era.py:172-186, dm.py:65-82 */
  float alpha;                /* division regulariser (era.py:90) */
} pty_opts;

/* 1 if M or K = L is a supported frame side (16, 32, 64, 96, 128, 192, 256). */
int pty_supported_frame(int32_t L);
const char* pty_last_error(void);

/* Geometry and workspace.  R_host holds the reference's R (N x 2, int32) AFTER
 * the preamble's shift (era.py:504-505): R <= 0, frame n sees
 * O[-R[n,0] : M - R[n,0], -R[n,1] : K - R[n,1]].  Every window must lie inside
 * the U x V object.  The context allocates its scratch on the current device. */
int pty_create(pty_ctx** out, int32_t N, int32_t M, int32_t K, int32_t U, int32_t V,
               const int32_t* R_host);
int pty_destroy(pty_ctx* ctx);
/* Run all work of this context on `stream` (a hipStream_t; NULL = default). */
int pty_set_stream(pty_ctx* ctx, void* stream);

/* amp_dev = ifftshift(sqrt(I)) from centred intensities I_dev (float32, N x M x K),
 * and I_norm = sum(mask * I) with the centred mask (era.py:495-498).
 * mask_centred_dev may be NULL (all good).  Synchronous (returns I_norm). */
int pty_prepare_amp(pty_ctx* ctx, const float* I_dev, const uint8_t* mask_centred_dev,
                    float* amp_dev, double* I_norm_out);

/* Bind the modulus data: amplitudes (N x M x K float32, ifftshifted), the
 * ifftshifted mask (M x K, 1 = good; NULL = all good) and I_norm. */
int pty_set_data(pty_ctx* ctx, const float* amp_dev, const uint8_t* mask_dev, double I_norm);

/* Bind the iterate: object (U x V) and probe (M x K), updated in place. */
int pty_set_state(pty_ctx* ctx, pty_cfloat* O_dev, pty_cfloat* P_dev);

/* Background state for methods 4-6: b = sqrt(ifftshift(B)) or the random
 * start (era.py:539-544).  per_frame = 1: N x M x K values, 0: one M x K frame
 * shared by all frames.  Copied into the context. */
int pty_set_background(pty_ctx* ctx, const float* b_dev, int32_t per_frame);

/* Far-field probe modulus |FFT(P)| for Pmod_P (era.py:546-547).  NULL: compute
 * it now from the bound probe. */
int pty_set_pamp(pty_ctx* ctx, const float* pamp_dev);

/* Sub-pixel positions (the Fourier-shift roll of float positions,
 * bagOfns.py:729-764, used on the sample by utility_Ptych.py:42-48; the
 * reference's ERA / DM, era.py:426-433, take integer R only).  The context's
 * R (pty_create) holds each position's integer window; d_host (N x 2 float64,
 * (y, x) per frame) the remainder: frame n's window start is -R[n] + d[n] and
 * it sees the probe shifted by d[n], P_n = IFFT2(FFT2(P) ramp_n) with
 * ramp_n[ky, kx] = exp(-2 pi i (d_n0 f(ky) / M + d_n1 f(kx) / K)), f(k) =
 * ((k + N/2 - 1) mod N) - (N/2 - 1).  The object update uses P_n in place of
 * P (numerator sum conj(P_n) psi_n, heatmap sum |P_n|^2).  Shifts must be
 * finite with |d| <= 1; NULL returns the context to integer positions.  The
 * stack P_n (N x M x K, allocated here on first use) is rebuilt from the bound
 * P by every pty_run with begin = 1 and after pty_set_state.  Runs with
 * shifts set take methods 1 and 4 (object and background updates; the probe
 * stays fixed) without Pmod_probe: anything else is PTY_EUNSUP.  The entry
 * points that take the probe as an argument (pty_make_exits, pty_psup_O,
 * pty_back_projection, pty_hio_modulus, pty_add_exits) ignore the shifts.
 * Synchronous. */
int pty_set_subpixel(pty_ctx* ctx, const double* d_host);

/* Per-frame position terms for refining the positions (the coordinate
 * gradient of the reference's notebooks,
 * ipython_notbooks/1dcoordsUpdate_1dsample_2dprobe.ipynb, for the probe-shift
 * model): with S_n = FFT2(O_win P_n), D_n,a = FFT2(O_win dP_n / dd_n,a) (a = y,
 * x), r = |S_n| - amp_n on the good pixels and J_a = Re(conj(S_n) D_n,a) /
 * |S_n|, terms_host (N x 6 float64) receives per frame f_n = sum r^2 (the
 * frame's share of the unnormalised eMod sum, era.py:560-572), the gradient
 * 2 sum r J_y, 2 sum r J_x of f_n in the frame's shift d_n, and the
 * Gauss-Newton matrix 2 sum J_y J_y, 2 sum J_y J_x, 2 sum J_x J_x.  Uses the
 * bound O, P, data and the shifts of pty_set_subpixel (required; all zeros
 * for integer positions).  Allocates two N x M x K derivative stacks on first
 * use.  Synchronous. */
int pty_position_terms(pty_ctx* ctx, double* terms_host);

/* Run `iters` iterations with global iteration numbers iter0, iter0+1, ...
 * (they drive the OP_iters schedule and the Pmod_probe window).  begin = 1
 * starts a new solve (DM: f = make_exits(O, P), eCon reference = O/P;
 * heatmaps invalidated).  metrics_host receives iters x 3 doubles
 * (eMod, eCon, eSup) per iteration; may be NULL.  Blocks until done. */
int pty_run(pty_ctx* ctx, const pty_opts* opts, int64_t iter0, int32_t iters, int32_t begin,
            double* metrics_host);

/* fftshift(|fftn(exits)|^2) of the final exits into I_dev (N x M x K float32):
 * ERA: make_exits(O, P) (era.py:293-295); DM: make_exits(Os, Ps) (dm.py:296-299). */
int pty_intensity(pty_ctx* ctx, float* I_dev);

/* Copy the cached P heatmap (U x V float64) to H_dev; *valid = 0 if none is
 * cached (the reference's info['heatmap'] is then None). */
int pty_heatmap(pty_ctx* ctx, double* H_dev, int32_t* valid);

/* Current frame-averaged background (M x K float32, ifftshifted). */
int pty_background(pty_ctx* ctx, float* b_dev);

/* Resume from a recorded state: the values the reference's loop carries from
 * one iteration to the next besides O and P (load them after pty_set_state,
 * then pty_run with begin = 0 and iter0 = the next iteration's number).
 *   pty_set_heatmap: the cached P heatmap H (U x V float64, era.py:143,163-189)
 *     and O heatmap G (M x K float64, era.py:144,164); NULL = not cached (the
 *     reference's None: recomputed when next needed).
 *   pty_set_iterate: DM's iterate f (N x M x K, dm.py:45-97; NULL keeps it)
 *     and the eCon reference, the previous iteration's metric object and probe
 *     (dm.py:35-38,158-160; both NULL = the bound O and P).  The per-frame
 *     background state goes through pty_set_background(per_frame = 1).
 *   pty_get_iterate: copies out f (N x M x K, optional) and the per-frame
 *     background state (N x M x K float32, optional). */
int pty_set_heatmap(pty_ctx* ctx, const double* H_dev, const double* G_dev);
int pty_set_iterate(pty_ctx* ctx, const pty_cfloat* f_dev, const pty_cfloat* O_ref_dev,
                    const pty_cfloat* P_ref_dev);
int pty_get_iterate(pty_ctx* ctx, pty_cfloat* f_dev, float* b_dev);

/* Choose between the fused per-frame kernel (one work-group transforms a whole
 * frame through an Infinity-Cache-resident scratch slot; square 128 / 256
 * frames, methods without background) and the staged row / column kernels.
 * threads: 0 = keep, 512 or 1024 per work-group; slots: 0 = keep, else the
 * number of work-groups (and scratch slots).  Default: staged (measured
 * faster at 128 x 128 and 256 x 256; env PTY_FRAME=1 selects the fused
 * kernel), one slot per compute unit. */
int pty_set_path(pty_ctx* ctx, int32_t fused, int32_t threads, int32_t slots);

/* enable != 0: pty_run captures each distinct kind of iteration (its branch
 * decisions and buffer addresses) once as a hipGraph on a private stream and
 * replays it; results are identical to eager launches.  Stage timing is not
 * recorded while graphs are on. */
int pty_set_graphs(pty_ctx* ctx, int32_t enable);
/* Whether graphs are on, and whether a capture or instantiation failed during
 * a run (the run then finished with eager launches and graphs were switched
 * off; nothing of the failed capture ran). */
int pty_get_graphs(pty_ctx* ctx, int32_t* enabled, int32_t* fell_back);

/* ---- launch shapes and the autotuner (no counterpart in the reference:
 * its loop is NumPy on the CPU) ---- */
/* Persistent staged kernels use up to rows_mult * CUs (1..256, default 3)
 * work-groups for the row FFTs and cols_mult * CUs (1..128; 0 = the default per
 * column mode) for the column FFTs; cols_wg threads per column work-group:
 * 256, 512 (32-column blocks, 256^2 frames only) or 0 = automatic (512 for
 * the Pmod_1 column steps at 256^2, else 256; the default). */
int pty_set_launch(pty_ctx* ctx, int32_t rows_mult, int32_t cols_mult, int32_t cols_wg);
/* ERA's staged modulus step over blocks of `frames` frames (0 = the whole
 * stack at once), so the frame spectra stay in the Infinity Cache between its
 * row / column / inverse-row kernels; default 0 (env PTY_CHUNK_FRAMES).
 * Frames per block are rounded up to at least N/16 and to whole column
 * chunks. */
int pty_set_chunk(pty_ctx* ctx, int32_t frames);
/* ERA object-only steps (methods 1 and 4) after the first: fuse the inverse
 * row FFT of psi' into the object update (one work-group per object row
 * segment, the same per-pixel summation order as the unfused update), so psi'
 * never goes back to HBM.  DM with background (methods 4, and 6 on its
 * object-only iterations): the leading modulus step's inverse row FFT writes
 * f and projects it onto the object in the same pass.  Built for K in
 * {64, 96, 128, 192, 256} and V <= 8192; on by default where supported (env
 * PTY_ROW_GATHER=0 turns it off).  enable != 0 on an unsupported geometry
 * leaves it off; pty_get_row_gather reports both. */
int pty_set_row_gather(pty_ctx* ctx, int32_t enable);
int pty_get_row_gather(pty_ctx* ctx, int32_t* supported, int32_t* enabled);
typedef struct pty_tune_result {
  int32_t fused, frame_threads, frame_slots, rows_mult, cols_mult, cols_wg, chunk_frames;  /* chosen */
  int32_t row_gather, fft_plan;                   /* chosen (fft_plan: pty_set_plan's mask) */
  int32_t candidates;                                               /* configurations timed */
  float best_ms, default_ms;                                        /* per iteration */
} pty_tune_result;
/* Times `trial_iters` iterations of opts for every launch configuration
 * (staged: rows_mult {2,3,4} x cols_mult {default,2,4}, and at 256^2 512-thread
 * column work-groups x cols_mult {default,1,3}; the fused inverse-row object
 * update on / off; every other line-FFT plan mask (pty_set_plan); ERA frame
 * blocks of 0 and 1/4..2x the default; fused per-frame path where the frame
 * shape allows it: 512 / 1024 threads x CUs / 2*CUs slots), keeps the fastest
 * and restores O, P and the background; the next pty_run must pass begin = 1. */
int pty_tune(pty_ctx* ctx, const pty_opts* opts, int32_t trial_iters, pty_tune_result* out);

/* Line-FFT plans of the kernel stages (no counterpart in the reference,
 * whose transforms are numpy.fft).  `mask` sets, per stage, the frame side's
 * alternative radix split instead of its default: PTY_PLAN_ROWS (row
 * kernels), PTY_PLAN_COLS (column kernels), PTY_PLAN_ROWG (the fused inverse
 * row FFT + object update), PTY_PLAN_BGERR (the model-error pass with
 * background of DM methods 4-6, a column kernel of its own).  Alternatives: 64: 4x4x4 on 16 threads per line
 * instead of 8x8 on 8; 96: 4x4x6 instead of 6x4x4; 128: 4x8x4 on 16 threads
 * instead of 16x8 on 8; 192: 4x4x12 instead of 12x4x4; 256: 8x4x8 on 32
 * threads instead of 16x16 on 16; sides 16 and 32 keep their plan.  Stages
 * exchange frames in natural order, so plans mix freely; results agree to
 * float32 round-off between plans (the same transform, another order of
 * operations).  The default mask is the library's per-side choice (or the
 * environment's PTY_FFT_PLAN); pty_tune tries every mask.  pty_get_plan
 * reports the mask and whether the frame has alternatives. */
enum { PTY_PLAN_ROWS = 1, PTY_PLAN_COLS = 2, PTY_PLAN_ROWG = 4, PTY_PLAN_BGERR = 8 };
int pty_set_plan(pty_ctx* ctx, int32_t mask);
int pty_get_plan(pty_ctx* ctx, int32_t* mask, int32_t* has_alternative);

/* ---- several scans at once (no counterpart in the reference) ----
 * The contexts (prepared with pty_set_data / pty_set_state / ...) stay owned
 * by the caller and must outlive the batch.  Each scan's kernels run on one
 * of up to 4 lane streams (GPU_MAX_HW_QUEUES, the process's hardware queues,
 * when set; PTY_BATCH_LANES overrides, at most 8) so scans of different
 * sizes overlap; with graphs
 * on, one iteration of every scan is captured as a single hipGraph (fork /
 * join over the lanes) per distinct iteration kind and replayed.  Per scan,
 * results are identical to pty_run on that context alone. */
typedef struct pty_batch pty_batch;
int pty_batch_create(pty_batch** out, int32_t n, pty_ctx* const* ctxs);
int pty_batch_destroy(pty_batch* batch);
int pty_batch_set_graphs(pty_batch* batch, int32_t enable);
int pty_batch_get_graphs(pty_batch* batch, int32_t* enabled, int32_t* fell_back);
/* opts: n entries; metrics_host: n pointers to 3 * S * iters doubles (S = the
 * context's packed scans, else 1; or NULL). */
int pty_batch_run(pty_batch* batch, const pty_opts* opts, int64_t iter0, int32_t iters,
                  int32_t begin, double* const* metrics_host);

/* Packed scans (no counterpart in the reference): S scans of one frame size
 * in one context, so that each stage of an iteration is one launch over all
 * their frames.  The context is created over the concatenated frames (scan s
 * owns frames [frame_start[s], frame_start[s + 1])) and a packed object whose
 * row band [row_start[s], row_start[s + 1]) is scan s's object (bands start
 * on multiples of 16 rows, row_start[S] = U; each frame's window lies in its
 * scan's band).  Call it before pty_prepare_amp / pty_set_data (it unbinds
 * the data); after it, mask arguments are [S][M][K], the probe bound by
 * pty_set_state is [S][M][K] (scan s's probe at s M K), pty_prepare_amp
 * writes S values of I_norm, and pty_run writes 3 * S doubles per iteration
 * (iteration-major, eMod / eCon / eSup of each scan).  Runs take ERA methods
 * 1-3 without background, centring or Pmod_probe, method 3 with OP_iters
 * (1, 1), integer positions, staged kernels; anything else is PTY_EUNSUP.
 * I_norm (S values) normalises each scan's metrics; NULL: pty_prepare_amp
 * computes them.  Per scan the object and probe equal a context of that scan
 * alone bit for bit (the same sums in the same order); the metrics agree to
 * FP64 rounding.  S = 1 returns the context to a single scan (with I_norm[0]
 * when given).  Synchronous. */
int pty_set_scans(pty_ctx* ctx, int32_t S, const int32_t* frame_start_host,
                  const int32_t* row_start_host, const double* I_norm_host);

/* ---- per-stage timing (HIP events on the context stream) ---- */
enum {
  PTY_STAGE_ROWS_FWD = 0,  /* frame formation + forward row FFT */
  PTY_STAGE_COLS = 1,      /* column FFT + modulus step + inverse column FFT */
  PTY_STAGE_ROWS_INV = 2,  /* inverse row FFT (+ DM update) */
  PTY_STAGE_GATHER = 3,    /* object update (overlap gather; fused: with the inverse row FFT) */
  PTY_STAGE_PROBE = 4,     /* probe update (frame reduction) */
  PTY_STAGE_REDUCE = 5,    /* metric / background reductions */
  PTY_STAGE_MISC = 6,      /* copies, centring, far-field probe constraint */
  PTY_NSTAGE = 7
};
/* enable != 0 starts recording an event pair around every stage of pty_run
 * and clears the totals; 0 stops. */
int pty_profile(pty_ctx* ctx, int32_t enable);
/* Totals since pty_profile(ctx, 1): milliseconds and launch counts per stage
 * (arrays of PTY_NSTAGE). */
int pty_profile_read(pty_ctx* ctx, double* ms_host, int64_t* count_host);

/* ---- building blocks (used by the tests; same kernels as pty_run) ---- */

/* In-place batched 2-D FFT of nframes M x K frames: forward unnormalised,
 * inverse scaled by 1/(M K) (numpy.fft.fftn / ifftn over the last two axes). */
int pty_fft2(pty_ctx* ctx, pty_cfloat* data_dev, int32_t nframes, int32_t inverse);

/* exits = make_exits(O, P, R) (era.py:426-433). */
int pty_make_exits(pty_ctx* ctx, const pty_cfloat* O_dev, const pty_cfloat* P_dev,
                   pty_cfloat* exits_dev);

/* exits = pmod_1(amp, exits, mask, alpha, eMod_calc=True) in place (era.py:550-572)
 * with the bound data; *emod_host = the un-normalised eMod sum. */
int pty_pmod1(pty_ctx* ctx, pty_cfloat* exits_dev, float alpha, double* emod_host);

/* Allocation-free variant of pty_pmod1, for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It has no counterpart in the reference.  No
 * context: exits (N x M x K, in place), the ifftshifted amplitudes amp
 * (N x M x K float32) and mask (M x K, 1 = good; NULL = all good) are passed
 * directly, and the un-normalised eMod sum is written to emod_dev (one
 * double, device memory).  Work is queued on `stream` (hipStream_t; NULL =
 * default) and not waited for.  Its scratch (eMod partials, packed mask
 * words, twiddle tables: 284,800 bytes) is one static array per GPU, held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race.  The first call on a device sets the array up with
 * a synchronous copy, so make it outside graph capture.  Limits: M and K in
 * {16, 32, 64, 96, 128, 192, 256}, N x M x K <= 2^34, at most 256 compute
 * units; anything else is rejected before any work is queued.  It calls no
 * hipMalloc / hipFree. */
int pty_pmod1_noalloc(pty_cfloat* exits_dev, const float* amp_dev, const uint8_t* mask_dev,
                      int32_t N, int32_t M, int32_t K, float alpha, double* emod_dev,
                      void* stream);

/* O_out = psup_O(exits, P, R, (U, V), H, alpha) (era.py:336-369).  fresh = 1
 * recomputes the heatmap into H_dev (make_P_heatmap, era.py:402-410); fresh = 0
 * reads it. */
int pty_psup_O(pty_ctx* ctx, const pty_cfloat* exits_dev, const pty_cfloat* P_dev,
               pty_cfloat* O_out_dev, double* H_dev, int32_t fresh, float alpha);

/* P_out = psup_P(exits, O, R, G, alpha) (era.py:371-400); G = O heatmap (M x K
 * float64, make_O_heatmap, era.py:412-417), recomputed when fresh = 1. */
int pty_psup_P(pty_ctx* ctx, const pty_cfloat* exits_dev, const pty_cfloat* O_dev,
               pty_cfloat* P_out_dev, double* G_dev, int32_t fresh, float alpha);

/* Back-projection initial object (back_projection.py:4-88, with the ERA
 * preamble's centring): exits_n = IFFT2(mask * amp_n * FFT2(P) / |FFT2(P)|)
 * from the bound amplitudes / mask, O_out = psup_O(exits, P) and its fresh
 * heatmap H (U x V float64); exits_out (N x M x K) is optional. */
int pty_back_projection(pty_ctx* ctx, const pty_cfloat* P_dev, pty_cfloat* exits_out_dev,
                        pty_cfloat* O_out_dev, double* H_dev, float alpha);

/* One HIO modulus step (utility_Ptych.py:14-17, the update of the legacy
 * Ptychography.HIO method, Ptychography_2dsample_2dprobe_farfield.py:93-127)
 * with the bound data.  The exits are f = F + w * make_exits(O, P, R), F in
 * f_dev (w = 0: f = F; w = beta after an iteration leaves its Psup term
 * pending).  m = pmod_1(f) (era.py:550-572) is formed in m_dev and the update's
 * two linear combinations are applied as it leaves the inverse row FFT:
 * m <- (1 + 1/beta) m - f / beta (Psup's argument), F <- f - beta m.
 * *emod_host = the un-normalised eMod sum of f.  f_dev != m_dev, beta != 0. */
int pty_hio_modulus(pty_ctx* ctx, pty_cfloat* f_dev, pty_cfloat* m_dev, const pty_cfloat* O_dev,
                    const pty_cfloat* P_dev, float w, float beta, float alpha, double* emod_host);

/* f += scale * make_exits(O, P, R): the Psup term of the same update
 * (utility_Ptych.py:16 with makeExits, :24-30). */
int pty_add_exits(pty_ctx* ctx, const pty_cfloat* O_dev, const pty_cfloat* P_dev, float scale,
                  pty_cfloat* f_dev);

/* ---- learned initialiser (no counterpart in the reference; the north-star
 * conv encoder-decoder whose output seeds O) -------------------------------
 * bf16 tensors are uint16 bit patterns; activations NHWC; weights
 * [9*Cin][Cout] (tap-major, tap = 3*dy + dx).  stream: hipStream_t or NULL. */
const char* pty_nn_last_error(void);
/* C (M x N, fp32) = A (M x K) B (K x N), bf16 row-major: MFMA lane-layout probe. */
int pty_nn_gemm_probe(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
                      void* stream);
/* y = act(conv3x3(x, w) + bias), zero padding; y (bf16) and/or yf (fp32).
 * wT: the per-tap transpose [9][Cout][Cin] of w (pty_nn_weight_t), used by the
 * direct-convolution kernel (W in {8,16,32,64,128}, Cin and Cout multiples of
 * 32); NULL selects the generic implicit GEMM. */
int pty_nn_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const uint16_t* wT,
                       const float* bias, uint16_t* y, float* yf, int B, int H, int W, int Cin,
                       int Cout, int relu, void* stream);
/* The first layer (Cin = 1, Cout <= 64) with the network input formed on the
 * fly from centred float32 intensities I (B x H x W): x = bf16(sqrt(max(I, 0))
 * / scale), scale > 0 (the prediction's input normalisation). */
int pty_nn_conv3x3_fwd_intensity(const float* I, float scale, const uint16_t* w, const float* bias,
                                 uint16_t* y, float* yf, int B, int H, int W, int Cout, int relu,
                                 void* stream);
/* dx = conv3x3^T(dy, w); mask (optional, shape of dx): dx = 0 where mask <= 0
 * (the ReLU backward of the layer that produced this conv's input). */
int pty_nn_conv3x3_dgrad(const uint16_t* dy, const uint16_t* w, uint16_t* dx,
                         const uint16_t* mask, int B, int H, int W, int Cin, int Cout,
                         void* stream);
/* wT[t][co][ci] = w[t][ci][co] */
int pty_nn_weight_t(const uint16_t* w, uint16_t* wT, int Cin, int Cout, void* stream);
/* dw (fp32 [9*Cin][Cout]) and db (fp32 [Cout], optional) summed over `splits`
 * fixed partial buffers; ws: splits * (9*Cin*Cout + Cout) floats. */
int pty_nn_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, float* ws,
                         int splits, int B, int H, int W, int Cin, int Cout, void* stream);
/* split count that fills the device for pty_nn_conv3x3_wgrad (>= 1). */
int pty_nn_wgrad_splits(int B, int H, int W, int Cin, int Cout);
int pty_nn_relu_bwd(uint16_t* dy, const uint16_t* y, size_t n, void* stream);
/* 2x2 average pool (backward = 1/4 spread), nearest 2x upsample (backward = sum).
 * These and relu_bwd need 16-byte aligned tensors; pool2 / up2 need C % 8 == 0. */
int pty_nn_pool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, int backward,
                 void* stream);
int pty_nn_up2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, int backward,
               void* stream);
/* loss partials (1024 doubles) of sum w (p-t)^2, w = wgt[i % period] (or 1);
 * dy = gscale * 2 w (p-t) / n (bf16, optional). */
int pty_nn_mse(const float* p, const float* t, size_t n, uint16_t* dy, float gscale, double* part,
               const float* wgt, size_t period, void* stream);
int pty_nn_adam(float* w, uint16_t* wb, const float* g, float* m, float* v, size_t n, float lr,
                float b1, float b2, float eps, int step, float gscale, void* stream);
int pty_nn_to_bf16(const float* x, uint16_t* y, size_t n, void* stream);
/* Training batch: xb[b] = T_b(x[idx[b]]) (bf16, M x M) and tb[b] = T_b(t[idx[b]])
 * (complex as 2 floats), T_b one of the 8 maps from i -> (-i) mod M on rows
 * and columns and a transpose (exact for a probe symmetric under them),
 * code hash(seed, b) & 7 when aug != 0, identity otherwise. */
int pty_nn_batch_gather(const uint16_t* x, const float* t, const int64_t* idx, int B, int M,
                        int aug, uint32_t seed, uint16_t* xb, float* tb, void* stream);
/* Phase synchronisation of predicted patches (each known up to a constant phase).
 * pair_overlap: c[q] (double complex) = sum over the overlap of pair q = (i, j)
 *   of conj(w p_i) (w p_j); off: N (row, col) window offsets >= 0, pairs: npairs
 *   (i, j).  phase_sync: projected power iterations z_i <- s_i / |s_i|,
 *   s_i = sum_j C_ij z_j from z = 1, CSR rowptr (N+1) / entries (j, qs) with
 *   C_ij = c[qs] (qs >= 0) or conj(c[-qs-1]), nent = rowptr[N] entries,
 *   N <= 8192; work: 3 * nent floats (expanded values and columns).
 * phase_exits: out[n] = P * patch_n * z_n, whose overlap projection
 *   (pty_psup_O) is the |P|^2-weighted average of the aligned patches. */
int pty_nn_pair_overlap(const pty_cfloat* patches, const float* w, const int32_t* off,
                        const int32_t* pairs, int npairs, int M, int K, double* c, void* stream);
int pty_nn_phase_sync(const int32_t* rowptr, const int32_t* entries, const double* c, int N,
                      int iters, pty_cfloat* z, pty_cfloat* work, int32_t nent, void* stream);
int pty_nn_phase_exits(const pty_cfloat* patches, const pty_cfloat* z, const pty_cfloat* P, int N,
                       int M, int K, pty_cfloat* out, void* stream);
#ifdef __cplusplus
}
#endif
#endif /* PTYHIP_H */
