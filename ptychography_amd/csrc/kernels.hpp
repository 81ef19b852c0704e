// Device kernels of the ERA / DM iteration (gfx950, wave64).
//
// Frame stacks are [N][M][K] complex64, row-major; the object is [U][V];
// the probe [M][K].  Frame n sees the object window starting at off[n] =
// (y, x) >= 0 (the reference's R negated, ptychography/era.py:356).
//
// A 2-D frame transform is split into a row stage and a column stage; each
// stage keeps whole lines in registers and exchanges once through LDS
// (fft.hpp).  The per-pixel work of each ERA / DM step is fused into those
// stages:
//   rows forward : frame formation  psi = O_win * P  (or 2 O_win P - f, or f)
//   columns      : forward column FFT -> modulus / error / background update
//                  -> inverse column FFT (with FP64 error partials)
//   rows inverse : inverse row FFT, optionally the DM update f += ...; for
//                  ERA object-only steps fused with the object update
//                  (k_rows_gather)
// Reductions are deterministic: every work-group writes its own partial, and
// partials are summed in a fixed order by a single work-group.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fft.hpp"

namespace pty {

constexpr int WG = 256;

struct Geom {
  int N, M, K, U, V;
  const int2* __restrict__ off;  // (y, x) window offset per frame
  // sub-pixel positions: frame n sees its own probe Pn[n] ([N][M][K], the
  // probe Fourier-shifted by the position's fractional part) in place of the
  // shared P; null for integer positions
  const float2* Pn;
  // packed scans: frame n belongs to scan fscan[n] and sees that scan's probe
  // P + fscan[n] M K (and its mask); null for a single scan
  const int* fscan;
};

// probe of frame n: the shared probe, or its scan's probe when scans are packed
__device__ __forceinline__ const float2* frame_probe(const Geom& g, const float2* P, int n) {
  return g.fscan ? P + (size_t)g.fscan[n] * g.M * g.K : P;
}

// the probe row frame row `row` (= n M + i) sees
__device__ __forceinline__ const float2* probe_row(const Geom& g, const float2* P, int row, int i) {
  if (g.Pn) return g.Pn + (size_t)row * g.K;
  if (g.fscan) P = frame_probe(g, P, row / g.M);
  return P + (size_t)i * g.K;
}

// ---------------------------------------------------------------------------
// block reduction of doubles (wave64 shuffles, then LDS across waves)
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red /*[4*NV]*/) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += red[i * NV + k];
      v[k] = s;
    }
  }
}

// Work-groups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// XCD with its own L2.  Re-number them so that consecutive work items land on
// the same XCD: neighbouring column blocks share amplitude cache lines,
// neighbouring frames share object rows, neighbouring tiles share frame rows.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_remap(int b, int grid) {
  if (grid % NXCD) return b;
  return (b % NXCD) * (grid / NXCD) + b / NXCD;
}

// Copy the L-entry twiddle table into LDS (one barrier); line_fft then reads
// its pass-2 twiddles from LDS instead of global memory.
template <int L>
__device__ __forceinline__ void stage_twiddles(float2* s_tw, const float2* __restrict__ tw) {
  for (int k = threadIdx.x; k < L; k += blockDim.x) s_tw[k] = tw[k];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// row stage, forward
// ---------------------------------------------------------------------------
enum RowSrc { SRC_OP = 0, SRC_DM = 1, SRC_F = 2, SRC_HIO = 3 };

// X[n, i, :] = FFT_row( source(n, i, :) ) for rows [0, nrows) of the stack
// (streaming stores: measured 0.62 vs 0.68 ms at 4096 x 256^2, the object
// window and probe rows keep more of L2).
//   SRC_OP: O[y_n + i, x_n + c] * P[i, c]
//   SRC_DM: 2 O[..] P[i, c] - F[n, i, c]
//   SRC_F : F[n, i, c]                 (X may alias F)
//   SRC_HIO: F[n, i, c] + w O[..] P[i, c]  (HIO's iterate with its pending Psup term)
// The sources that also read f (DM, HIO) hold two line groups of inputs in
// registers: at three work-groups per CU they spilled 88 bytes per lane; at
// two they do not (DM with background at 4096 x 256^2 frames: 9.61 vs
// 9.65-9.70 ms per iteration).
template <int PK, int SRC>
__global__ __launch_bounds__(WG, (SRC == SRC_DM || SRC == SRC_HIO) ? 2 : 3)
void k_rows_fwd(Geom g, const float2* O, const float2* P,
                                                 const float2* F, float2* X,
                                                 const float2* __restrict__ tw, int nrows,
                                                 float w) {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T, LPW = WG / T;
  constexpr int W = LineLds<PK>::words;
  constexpr bool HAS_F = (SRC == SRC_DM || SRC == SRC_HIO);
  constexpr int E2 = HAS_F ? E : 1;  // f values (DM, HIO)
  __shared__ float2 lds[LPW * W];
  __shared__ float2 s_tw[L];
  stage_twiddles<L>(s_tw, tw);
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int M = g.M;
  const int ngroups = (nrows + LPW - 1) / LPW;
  // raw inputs of the next line group (f, or the object window and f), loaded
  // one group ahead; the probe row is read from cache when it is needed
  float2 n0[E], n2[E2];
  auto load = [&](int grp) {
    const int row = min(grp * LPW + line, nrows - 1);  // clamped: spare lines re-read a real row
    const size_t rbase = (size_t)row * L;
    if constexpr (SRC == SRC_F) {
#pragma unroll
      for (int e = 0; e < E; ++e) n0[e] = load_stream(&F[rbase + layA<PK>(e, t)]);
    } else {
      const int n = row / M, i = row % M;
      const int2 o = g.off[n];
      const float2* Orow = O + (size_t)(o.x + i) * g.V + o.y;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = layA<PK>(e, t);
        n0[e] = Orow[c];
        if constexpr (HAS_F) n2[e] = load_stream(&F[rbase + c]);
      }
    }
  };
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  if (wg < ngroups) load(wg);
  for (int grp = wg; grp < ngroups; grp += gridDim.x) {
    float2 v[E];
    if constexpr (SRC == SRC_F) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = n0[e];
    } else {
      const int row = min(grp * LPW + line, nrows - 1);
      const float2* Prow = probe_row(g, P, row, row % M);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        float2 psi = cmul(n0[e], Prow[layA<PK>(e, t)]);
        if constexpr (SRC == SRC_DM) psi = make_float2(2.f * psi.x - n2[e].x, 2.f * psi.y - n2[e].y);
        if constexpr (SRC == SRC_HIO) psi = make_float2(n2[e].x + w * psi.x, n2[e].y + w * psi.y);
        v[e] = psi;
      }
    }
    if (grp + (int)gridDim.x < ngroups) load(grp + gridDim.x);
    fft_fwd<PK>(v, lds + line * W, s_tw, t);
    const int row = grp * LPW + line;
    if (row < nrows) {
      const size_t rbase = (size_t)row * L;
#pragma unroll
      for (int e = 0; e < E; ++e) store_stream(&X[rbase + layB<PK>(e, t)], v[e]);
    }
    lds_barrier();
  }
}

// Frame rows listed per object-row segment (k_rows_fwd_orow): item (u, s)
// holds the rows (n, i = u - y_n) whose window starts in columns
// [s segw, (s + 1) segw), ascending n; every frame row is in exactly one list.
struct RowFwdList {
  int segs, segw;                   // column segments per object row and their width
  const int* __restrict__ ptr;      // [U * segs + 1]
  const int2* __restrict__ ent;     // (n*M + i, x_n << 16 | i)
};

// The forward row stage of the sources that read the object (SRC_OP, SRC_DM,
// SRC_HIO) in object-row order: one work-group per (object row u, column
// segment).  The object columns the segment's windows reach ([v0, v0 + segw
// + L - 1), clipped to V) and the segment's list of frame rows (ROWF_ENT
// entries at a time) are staged in LDS; the listed frame rows are then formed
// LPW at a time from LDS, the probe row (L2) and, for DM / HIO, the frame row
// of f (read a batch ahead), and transformed as k_rows_fwd<PK, SRC> does,
// with the same arithmetic, so the spectra are bit-identical (X may alias F:
// a row is read and written by the same batch).  Each object row is read from memory once
// per segment instead of once by every frame whose window covers it (about
// 40 frames at a 40-pixel step), and no global load but the probe row's sits
// between one batch and the next (list entries read from memory a batch
// ahead stalled the batches whenever other traffic had pushed the lists out
// of the Infinity Cache: 0.40 vs 0.56-0.60 ms at 4096 x 256^2).
constexpr int ROWF_ENT = 256;
template <int PK, int SRC>
__global__ __launch_bounds__(WG, (SRC == SRC_DM || SRC == SRC_HIO) ? 2 : 3)
void k_rows_fwd_orow(Geom g, const float2* __restrict__ O, const float2* P, const float2* F,
                     float2* X, RowFwdList a, const float2* __restrict__ tw, float w) {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T, LPW = WG / T;
  constexpr int W = LineLds<PK>::words;
  constexpr bool HAS_F = (SRC == SRC_DM || SRC == SRC_HIO);
  constexpr int E2 = HAS_F ? E : 1;  // f values (DM, HIO)
  static_assert(SRC == SRC_OP || HAS_F, "sources that read the object");
  static_assert(ROWF_ENT % LPW == 0, "whole batches per list chunk");
  __shared__ float2 lds[LPW * W];
  __shared__ float2 s_tw[L];
  __shared__ int2 s_ent[ROWF_ENT];
  extern __shared__ float2 orow[];
  const int item = blockIdx.x;
  const int b0 = a.ptr[item], cnt = a.ptr[item + 1] - b0;
  if (cnt == 0) return;  // uniform: the whole work-group leaves
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int u = item / a.segs, v0 = (item % a.segs) * a.segw;
  const int span = min(g.V, v0 + a.segw + L - 1) - v0;
  const float2* Ou = O + (size_t)u * g.V + v0;
  for (int v = tid; v < span; v += WG) orow[v] = Ou[v];
  float2 pv[E], fv[E2];
  auto load_probe = [&](int2 m) {
    const float2* Prow = probe_row(g, P, m.x, m.y & 0xffff);
#pragma unroll
    for (int e = 0; e < E; ++e) pv[e] = Prow[layA<PK>(e, t)];
    if constexpr (HAS_F) {
      const size_t rbase = (size_t)m.x * L;
#pragma unroll
      for (int e = 0; e < E; ++e) fv[e] = load_stream(&F[rbase + layA<PK>(e, t)]);
    }
  };
  for (int c0 = 0; c0 < cnt; c0 += ROWF_ENT) {
    const int nc = min(ROWF_ENT, cnt - c0);
    for (int k = tid; k < nc; k += WG) s_ent[k] = a.ent[b0 + c0 + k];
    // the first chunk's barrier also publishes the object columns and twiddles
    if (c0 == 0) stage_twiddles<L>(s_tw, tw);
    else __syncthreads();
    const int nbatch = (nc + LPW - 1) / LPW;
    // spare lines of the last batch re-read a real entry and store nothing
    int2 m = s_ent[min(line, nc - 1)];
    load_probe(m);
    for (int bt = 0; bt < nbatch; ++bt) {
      const float2* Ow = orow + ((m.y >> 16) - v0);
      float2 v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        float2 psi = cmul(Ow[layA<PK>(e, t)], pv[e]);
        if constexpr (SRC == SRC_DM) psi = make_float2(2.f * psi.x - fv[e].x, 2.f * psi.y - fv[e].y);
        if constexpr (SRC == SRC_HIO) psi = make_float2(fv[e].x + w * psi.x, fv[e].y + w * psi.y);
        v[e] = psi;
      }
      const int2 mn = s_ent[min((bt + 1) * LPW + line, nc - 1)];
      if (bt + 1 < nbatch) load_probe(mn);
      fft_fwd<PK>(v, lds + line * W, s_tw, t);
      if (bt * LPW + line < nc) {
        const size_t rbase = (size_t)m.x * L;
#pragma unroll
        for (int e = 0; e < E; ++e) store_stream(&X[rbase + layB<PK>(e, t)], v[e]);
      }
      lds_barrier();
      m = mn;
    }
  }
}

// ---------------------------------------------------------------------------
// row stage, inverse
// ---------------------------------------------------------------------------
enum RowDst { DST_STORE = 0, DST_DMUPD = 1, DST_HIO = 2 };

// HIO's linear steps (utility_Ptych.py:16) applied where m = Pmod(f) leaves
// the inverse row FFT, with f = F + w O P (F holds the iterate without its
// pending Psup term): Out <- a m - b f (Psup's argument), F <- f - beta m.
struct HioMix {
  float2* F;
  const float2* O;
  const float2* P;
  float w;            // coefficient of the pending O P term (0: none)
  float a, b, beta;   // 1 + 1/beta, 1/beta, beta
};

// y = scale * IFFT_row(Y[n, i, :]) (unnormalised inverse, times scale), then
//   DST_STORE: Out[n, i, c] = y                         (Out may alias Y)
//   DST_DMUPD: Out[n, i, c] += y - O[..] P[i, c]        (DM: f - Ps f + Pm(..))
//              partial[block] = sum |y - O P|^2          (DM step size)
//   DST_HIO  : f = F + w O[..] P[i, c]; Out = a y - b f, F = f - beta y (HioMix)
template <int PK, int DST>
__global__ __launch_bounds__(WG, 3) void k_rows_inv(Geom g, const float2* Y, float2* Out,
                                                 const float2* O, const float2* P, float scale,
                                                 double* partial, const float2* __restrict__ tw,
                                                 int nrows, HioMix hm) {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T, LPW = WG / T;
  constexpr int W = LineLds<PK>::words;
  __shared__ float2 lds[LPW * W];
  __shared__ float2 s_tw[L];
  __shared__ double red[4];
  stage_twiddles<L>(s_tw, tw);
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int ngroups = (nrows + LPW - 1) / LPW;
  float2 nx[E];
  auto load = [&](int grp) {
    const int row = min(grp * LPW + line, nrows - 1);
    const size_t rbase = (size_t)row * L;
#pragma unroll
    for (int e = 0; e < E; ++e) nx[e] = load_stream(&Y[rbase + layB<PK>(e, t)]);
  };
  double acc[1] = {0.0};
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  if (wg < ngroups) load(wg);
  for (int grp = wg; grp < ngroups; grp += gridDim.x) {
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nx[e];
    if (grp + (int)gridDim.x < ngroups) load(grp + gridDim.x);
    const int row = grp * LPW + line;
    // HIO reads its iterate F at every pixel: issue those loads before the
    // inverse FFT, whose LDS exchanges hide them (DM's update, at its register
    // budget, measured slower the same way)
    float2 fpre[DST == DST_HIO ? E : 1];
    if constexpr (DST == DST_HIO) {
      const size_t rb = (size_t)min(row, nrows - 1) * L;
#pragma unroll
      for (int e = 0; e < E; ++e) fpre[e] = load_stream(&hm.F[rb + layA<PK>(e, t)]);
    }
    fft_inv<PK>(v, lds + line * W, s_tw, t);
    if (row < nrows) {
      const size_t rbase = (size_t)row * L;
      if constexpr (DST == DST_STORE) {
#pragma unroll
        for (int e = 0; e < E; ++e) store_stream(&Out[rbase + layA<PK>(e, t)], v[e] * scale);
      } else if constexpr (DST == DST_HIO) {
        const int M = g.M, n = row / M, i = row % M;
        const int2 o = g.off[n];
        const float2* Orow = hm.O + (size_t)(o.x + i) * g.V + o.y;
        const float2* Prow = hm.P + (size_t)i * L;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int c = layA<PK>(e, t);
          const float2 m = v[e] * scale;
          float2 fv = fpre[e];
          if (hm.w != 0.f) {
            const float2 ex = cmul(Orow[c], Prow[c]);
            fv = make_float2(fv.x + hm.w * ex.x, fv.y + hm.w * ex.y);
          }
          store_stream(&Out[rbase + c], make_float2(hm.a * m.x - hm.b * fv.x, hm.a * m.y - hm.b * fv.y));
          store_stream(&hm.F[rbase + c], make_float2(fv.x - hm.beta * m.x, fv.y - hm.beta * m.y));
        }
      } else {
        const int M = g.M, n = row / M, i = row % M;
        const int2 o = g.off[n];
        const float2* Orow = O + (size_t)(o.x + i) * g.V + o.y;
        const float2* Prow = probe_row(g, P, row, i);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int c = layA<PK>(e, t);
          const float2 ex = cmul(Orow[c], Prow[c]);
          const float2 y = v[e] * scale;
          const float2 d = y - ex;
          acc[0] += (double)norm2(d);
          store_stream(&Out[rbase + c], load_stream(&Out[rbase + c]) + d);
        }
      }
    }
    lds_barrier();
  }
  if constexpr (DST == DST_DMUPD) {
    block_sum<1>(acc, red);
    if (tid == 0) partial[blockIdx.x] = acc[0];
  }
}

// ---------------------------------------------------------------------------
// column stage
// ---------------------------------------------------------------------------
enum ColMode {
  COL_ERA1 = 0,   // Pmod_1 with eMod (era.py:560-572) + sum |psi_hat'|^2
  COL_DM1,        // Pmod_1 without eMod
  COL_ERR1,       // model_error_1, no background (dm.py:337-341); forward only
  COL_ERR7,       // model_error_1 with mean background bbar; forward only
  COL_INTENS,     // fftshift(|psi_hat|^2) -> outI; forward only
  COL_PMODP,      // Pmod_P factor (era.py:597-610) with amp = |FFT P0|
  COL_ERAB,       // Pmod_7 with eMod (era.py:583-595); sums b*fac over frames
  COL_DMB_LEAD,   // Pmod_7 on (f, b_n), b_n updated in place; sums new b_n
  COL_DMB_UPD,    // DM background step: b0 = 2 bbar - b_n, Pmod_7, b_n += b0' - bbar
  COL_FWD,        // plain forward column FFT (2-D FFT building block)
  COL_INV,        // plain inverse column FFT (input in spectrum order), times scale
  // DM with background, iterate carried as spectra F = FFT2(f) (see k_cols):
  COL_DMBF_LEAD,  // F = spectrum of the input; Xo = col-IFFT(Pmod_7(F, b)); sums b'
  COL_DMBF,       // one DM step on (F, b) from the spectra of O P; X = col-IFFT(F_new),
                  // Xo = col-IFFT(Pmod_7(F_new, b_new)); sums b_new and b'
  COL_DM1F,       // one DM step on F (no background) from the spectra S of O P;
                  // X = col-IFFT(F_new); the model error of S (the previous
                  // iteration's metric, O being its metric object) when part is set
};

struct ColArgs {
  float2* X;
  const float* amp;
  long amp_stride;       // elements between frames of amp (0 = one shared frame)
  const uint8_t* mask;   // [M][K], 1 = good, ifftshifted; null = all good
  const uint32_t* maskbits;  // [T][K]: bit e = mask of element e of column-thread t (Pmod_1 modes)
  float alpha;
  float scale;           // applied to the spectrum before the inverse (1/(M K))
  float* bg;             // background b (per frame or shared)
  long bg_stride;
  const float* bbar;     // [M][K] frame-mean background
  float* outI;           // [N][M][K] intensities (COL_INTENS)
  double* part;          // [gridDim.y * gridDim.x] eMod / error partials
  double* part2;         // [gridDim.y * gridDim.x] sum |psi_hat'|^2 partials
  double* partb;         // [gridDim.y][M*K] background sums over the chunk's frames
  double* partb2;        // COL_DMBF: second set of background sums
  float2* F;             // COL_DMBF(_LEAD): the DM iterate's spectra [N][M][K]
  float2* Xo;            // COL_DMBF(_LEAD): second inverse output [N][M][K]
  int nframes, chunk;
  int item_parts;        // 1: one partial per item (at the item's index; the
                         // launch then runs one item per work-group), so the
                         // partials of packed scans are contiguous per scan
  const float2* __restrict__ tw;
};

template <int MODE> struct ColTraits {
  static constexpr bool fwd = MODE != COL_INV;
  static constexpr bool inv = MODE == COL_ERA1 || MODE == COL_DM1 || MODE == COL_PMODP ||
                              MODE == COL_ERAB || MODE == COL_DMB_LEAD ||
                              MODE == COL_DMB_UPD || MODE == COL_INV || MODE == COL_DMBF_LEAD ||
                              MODE == COL_DMBF || MODE == COL_DM1F;
  static constexpr bool store_fwd = MODE == COL_FWD;
  static constexpr bool bsum = MODE == COL_ERAB || MODE == COL_DMB_LEAD || MODE == COL_DMB_UPD ||
                               MODE == COL_DMBF_LEAD || MODE == COL_DMBF;
  static constexpr bool bsum2 = MODE == COL_DMBF;
  static constexpr bool err = MODE == COL_ERA1 || MODE == COL_ERR1 || MODE == COL_ERR7 ||
                              MODE == COL_ERAB || MODE == COL_DMBF || MODE == COL_DM1F;
  // modes that keep amplitude / background values across the FFT need more
  // registers: tuned for 2 waves per SIMD
  static constexpr bool heavy = MODE == COL_ERR1 || MODE == COL_ERR7 || MODE == COL_ERAB ||
                                MODE == COL_DMB_UPD || MODE == COL_DMB_LEAD ||
                                MODE == COL_DMBF_LEAD || MODE == COL_DMBF;
};

// Launch shape of k_cols<PK, MODE>.  Pmod_1 modes with at most 16 points per
// thread load each frame's spectra and amplitudes at the top of the frame step
// (amplitudes in flight across the forward FFT) instead of prefetching the
// next frame's spectra into registers, and take their mask packed one word per
// column (k_pack_mask): that fits 128 VGPRs, 4 waves per SIMD.  Measured at
// 4096 x 256^2: 1.55 ms (prefetch, byte mask, 3 waves) -> 1.47 (hoisted
// loads, 3 waves) -> 1.28 (packed mask, 4 waves) -> 1.18 ms with 512-thread
// work-groups (32-column blocks).  With 24 points per thread (96, 192) the
// hoisted amplitudes would cost 2 waves per SIMD; those keep the prefetch.
template <int PK, int MODE> struct ColShape {
  static constexpr int L = plan_len(PK);
  static constexpr int E = L / Plan<PK>::T;
  static constexpr bool modulus = MODE == COL_ERA1 || MODE == COL_DM1 || MODE == COL_ERR1 ||
                                  MODE == COL_ERR7 || MODE == COL_ERAB ||
                                  MODE == COL_DMB_LEAD || MODE == COL_DMB_UPD ||
                                  MODE == COL_DMBF_LEAD || MODE == COL_DMBF || MODE == COL_DM1F;
  static constexpr bool hoist = modulus && E <= 16;
  // modes that also hold background values across the FFT fit 3 waves
  static constexpr bool bgvals = MODE == COL_ERR7 || MODE == COL_ERAB || MODE == COL_DMB_LEAD ||
                                 MODE == COL_DMB_UPD || MODE == COL_DMBF_LEAD || MODE == COL_DMBF;
  // (plans with at most 8 points per thread hold half the values: one or two
  // waves more; COL_DMBF spills at 4)
  static constexpr int wpe =
      MODE == COL_DMBF ? (E <= 8 ? 3 : 2)
      : (MODE == COL_DM1F && L == 256 && E == 16) ? 3
      : hoist ? (bgvals ? (E <= 8 ? 4 : (MODE == COL_DMB_UPD && L == 128) ? 2 : 3) : 4)
              : ColTraits<MODE>::heavy ? 2 : 3;  // waves per SIMD
};

// Modulus-step arithmetic: the hardware square root and reciprocal (1 ulp;
// the divisor is |psi| + alpha >= alpha > 0) instead of the correctly rounded
// library sequences, which cost ~10 extra VALU instructions per element.
__device__ __forceinline__ float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float mrcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Uniform part of the row index held in register slot e (layoutA / layoutB);
// the per-thread part is t for both layouts.
template <int Ra, int Rb, int T>
__device__ __forceinline__ constexpr int urowA(int e) { return (e / Ra) * T + (e % Ra) * Rb; }
template <int Ra, int Rb, int T>
__device__ __forceinline__ constexpr int urowB(int e) { return (e / Rb) * T + (e % Rb) * Ra; }

// CWG threads per work-group: CPW = CWG / T columns per block, so a wave's
// load reads (64 / T) rows of CPW * 8 contiguous bytes (CWG = 512 at L = 256:
// 256-byte row segments).
// Pack the [M][K] mask for the column stage: word (t, c) holds, in bit e, the
// mask of the pixel that column-thread t keeps in register slot e of column c
// (row t + urowB(e)), so a Pmod_1 column step reads one word per column
// instead of E bytes per frame.
template <int PK>
__global__ __launch_bounds__(WG) void k_pack_mask(const uint8_t* __restrict__ mask, int K,
                                                  uint32_t* __restrict__ bits) {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T;
  static_assert(E <= 32, "one bit per register slot");
  const int idx = blockIdx.x * WG + threadIdx.x;
  if (idx >= T * K) return;
  const int t = idx / K, c = idx % K;
  uint32_t b = 0;
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (mask[(size_t)(t + rowB<PK>(e)) * K + c]) b |= 1u << e;
  bits[idx] = b;
}

template <int PK, int MODE, int CWG = WG>
__global__ __attribute__((amdgpu_flat_work_group_size(1, CWG),
                          amdgpu_waves_per_eu(ColShape<PK, MODE>::wpe))) void k_cols(Geom g, ColArgs a) {
  constexpr int L = plan_len(PK);
  using TR = ColTraits<MODE>;
  constexpr bool HOIST = ColShape<PK, MODE>::hoist;
  // hoisted per-frame (or shared) background values; the frame mean bbar is
  // the same for every frame (an L2 hit) and is read where it is used, which
  // keeps DMB_UPD and ERR7 within their register budget
  constexpr bool HBG = HOIST && (MODE == COL_ERAB || MODE == COL_DMB_LEAD || MODE == COL_DMB_UPD ||
                                 MODE == COL_DMBF_LEAD || MODE == COL_DMBF);
  constexpr bool DMBF = MODE == COL_DMBF_LEAD || MODE == COL_DMBF || MODE == COL_DM1F;
  static_assert(MODE != COL_DMBF || HOIST, "the second step reads the hoisted amplitudes");
  // COL_ERR7 on plans of at most 8 points per thread also loads the frame
  // mean background with the amplitudes, ahead of the FFT (16 points per
  // thread read it where it is used, within their 3-wave register budget)
  constexpr bool HBB = HOIST && MODE == COL_ERR7 && (L / Plan<PK>::T) <= 8;
  constexpr int T = Plan<PK>::T, E = L / T, CPW = CWG / T;
  constexpr int W = LineLds<PK>::words;
  __shared__ float2 lds[CPW * W];
  __shared__ float2 s_tw[L];
  __shared__ double red[2 * (CWG / 64)];
  stage_twiddles<L>(s_tw, a.tw);
  const int tid = threadIdx.x, l = tid % CPW, t = tid / CPW;
  const int K = g.K;
  const int ncb = (K + CPW - 1) / CPW;                 // column blocks per frame
  const int nchunks = (a.nframes + a.chunk - 1) / a.chunk;
  const int nitems = ncb * nchunks;                    // item = (column block, frame chunk)
  float2* mid = lds + l * W;
  double acc[2] = {0.0, 0.0};
  // element (row t + u, column c) of a frame is at [u*K + toff]: u*K is uniform
  float2 nx[E];
  auto load = [&](int item, int n) {
    const int c = min((item % ncb) * CPW + l, K - 1);  // clamped: spare columns re-read a real one
    const uint32_t toff = (uint32_t)(t * K + c);
    const auto rx = buf_rsrc(a.X + (size_t)n * L * K);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int u = TR::fwd ? rowA<PK>(e) : rowB<PK>(e);
      nx[e] = buf_load2(rx, toff * 8u, (uint32_t)(u * K) * 8u);
    }
  };
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  if constexpr (!HOIST)
    if (wg < nitems) load(wg, (wg / ncb) * a.chunk);
  for (int item = wg; item < nitems; item += gridDim.x) {
    const int cb = item % ncb, chunk = item / ncb;
    const int c = cb * CPW + l;
    const bool cv = c < K;
    const uint32_t toff = (uint32_t)(t * K + c);
    const int n0 = chunk * a.chunk;
    const int n1 = min(a.nframes, n0 + a.chunk);
    // packed scans: the mask of the item's scan (an item's frames share one:
    // chunks of several frames only occur in background modes, never packed)
    const size_t mscan = g.fscan ? (size_t)g.fscan[n0] : 0;
    const uint32_t mbits =
        (HOIST && cv && a.maskbits) ? a.maskbits[mscan * T * K + t * K + c] : 0xffffffffu;
    // per-chunk background sums (at most 16 frames): FP32 registers in the
    // hoisted modes (relative rounding <= 16 * 2^-24), written out in FP64
    using BAcc = std::conditional_t<HOIST, float, double>;
    BAcc bacc[TR::bsum ? E : 1], bacc2[TR::bsum2 ? E : 1];
#pragma unroll
    for (int e = 0; e < (TR::bsum ? E : 1); ++e) bacc[e] = 0;
#pragma unroll
    for (int e = 0; e < (TR::bsum2 ? E : 1); ++e) bacc2[e] = 0;
    // the frame-mean background is the same for every frame of the item:
    // loaded once per item (per frame it was re-fetched past L2 for part of
    // the frames)
    float bbh[HBB ? E : 1];
    if constexpr (HBB) {
#pragma unroll
      for (int e = 0; e < E; ++e) bbh[e] = cv ? (a.bbar + (size_t)rowB<PK>(e) * K)[toff] : 0.f;
    }

    for (int n = n0; n < n1; ++n) {
      // opaque, wave-uniform frame index: keeps the frame base in SGPRs
      const int nu = __builtin_amdgcn_readfirstlane(n);
      float2* Xn = a.X + (size_t)nu * L * K;
      float2 v[E];
      float amh[HOIST ? E : 1], bgh[HBG ? E : 1];
      // COL_DMBF keeps b_new for its second step (re-reading it and the
      // amplitudes from cache measured 7.24 vs 6.70 ms per DM iteration at
      // 4096 x 256^2)
      float bnh[MODE == COL_DMBF ? E : 1];
      // ... and F_new where the 2-wave register budget holds it (256 points
      // per line on 16 threads: 255 VGPRs; the 128 and 192 plans spill, and
      // re-read what they just stored)
      constexpr bool KEEPF = MODE == COL_DMBF && (L == 256 || E <= 8);
      float2 fnh[KEEPF ? E : 1];
      // ... and loads the iterate F ahead of the forward transform (its
      // registers then hold F_new)
      constexpr bool HF = (MODE == COL_DMBF || MODE == COL_DM1F) && L == 256;
      float2 fh[HF ? E : 1];
      float2* Fn = DMBF ? a.F + (size_t)nu * L * K : nullptr;
      if constexpr (HOIST) {
        load(item, nu);
        // spare columns (not cv) read a real one; their values are never used
        const auto ra = buf_rsrc(a.amp + (size_t)nu * a.amp_stride);
        const uint32_t toffc = (uint32_t)(t * K + min(c, K - 1));
#pragma unroll
        for (int e = 0; e < E; ++e) amh[e] = buf_load1(ra, toffc * 4u, (uint32_t)(rowB<PK>(e) * K) * 4u);
        if constexpr (HBG) {
          const auto rb = buf_rsrc(a.bg + (size_t)nu * a.bg_stride);
#pragma unroll
          for (int e = 0; e < E; ++e) bgh[e] = buf_load1<0>(rb, toffc * 4u, (uint32_t)(rowB<PK>(e) * K) * 4u);
        }
        if constexpr (HF) {
          const auto rF = buf_rsrc(Fn);
#pragma unroll
          for (int e = 0; e < E; ++e)
            fh[e] = buf_load2<MODE == COL_DM1F ? BUF_NT : 0>(rF, toffc * 8u, (uint32_t)(rowB<PK>(e) * K) * 8u);
        }
      }
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = nx[e];
      if constexpr (!HOIST) {
        // prefetch: next frame of this item, else the first frame of the next item
        if (nu + 1 < n1) {
          load(item, nu + 1);
        } else if (item + (int)gridDim.x < nitems) {
          const int ni = item + gridDim.x;
          load(ni, (ni / ncb) * a.chunk);
        }
      }
      if constexpr (TR::fwd) fft_fwd<PK>(v, mid, s_tw, t);

      // COL_ERA1: the frame's E terms of each error sum in FP32 (relative
      // rounding <= E 2^-24), then added to the FP64 partials
      float fr0 = 0.f, fr1 = 0.f;
      if constexpr (MODE != COL_FWD && MODE != COL_INV) {
        if (cv) {
          const float* ampn = a.amp ? a.amp + (size_t)nu * a.amp_stride : nullptr;
          float* bgn = a.bg ? a.bg + (size_t)nu * a.bg_stride : nullptr;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int u = rowB<PK>(e);
            const size_t uo = (size_t)u * K;
            const float2 z = v[e];
            const float z2 = norm2(z);
            if constexpr (MODE == COL_INTENS) {
              const int r = t + u;
              const int rs = (r + L / 2) % L, cs = (c + K / 2) % K;
              a.outI[(size_t)nu * L * K + (size_t)rs * K + cs] = z2;
            } else {
              float m, am;
              if constexpr (HOIST) {
                m = ((mbits >> e) & 1u) ? 1.f : 0.f;
                am = amh[e];
              } else {
                m = a.mask ? (float)(a.mask + mscan * L * K + uo)[toff] : 1.f;
                am = (ampn + uo)[toff];
              }
              if constexpr (MODE == COL_ERA1 || MODE == COL_DM1 || MODE == COL_PMODP) {
                const float mg = msqrt(z2) + a.alpha;
                if constexpr (MODE == COL_ERA1) {
                  const float d = mg - am;
                  fr0 += m * d * d;
                }
                const float fac = (m != 0.f) ? am * mrcp(mg) : 1.f;
                const float2 zn = z * fac;
                if constexpr (MODE == COL_ERA1) fr1 += norm2(zn);
                v[e] = zn * a.scale;
              } else if constexpr (MODE == COL_ERR1) {
                const float d = msqrt(z2) - am;
                acc[0] += (double)(m * d * d);
              } else if constexpr (MODE == COL_ERR7) {
                float b;
                if constexpr (HBB) b = bbh[e];
                else b = (a.bbar + uo)[toff];
                const float d = msqrt(z2 + b * b) - am;
                acc[0] += (double)(m * d * d);
              } else if constexpr (MODE == COL_ERAB || MODE == COL_DMB_LEAD) {
                float* bp = (bgn + uo) + toff;
                float b;
                if constexpr (HBG) b = bgh[e];
                else b = *bp;
                const float mg = msqrt(z2 + b * b + a.alpha);
                if constexpr (MODE == COL_ERAB) {
                  const float d = mg - am;
                  acc[0] += (double)(m * d * d);
                }
                const float fac = m * am * mrcp(mg);
                const float bn = b * fac;
                bacc[e] += (BAcc)bn;
                if constexpr (MODE == COL_DMB_LEAD) *bp = bn;
                const float2 zn = z * fac;
                if constexpr (MODE == COL_ERAB) acc[1] += (double)norm2(zn);
                v[e] = zn * a.scale;
              } else if constexpr (MODE == COL_DMB_UPD) {
                float* bp = (bgn + uo) + toff;
                float b;
                if constexpr (HBG) b = bgh[e];
                else b = *bp;
                const float bb = (a.bbar + uo)[toff];
                const float b0 = 2.f * bb - b;
                const float mg = msqrt(z2 + b0 * b0 + a.alpha);
                const float fac = m * am * mrcp(mg);
                const float bn = (b - bb) + b0 * fac;
                *bp = bn;
                bacc[e] += (BAcc)bn;
                v[e] = z * (fac * a.scale);
              } else if constexpr (MODE == COL_DMBF_LEAD) {
                // z = F (the iterate's spectrum, stored as the state); the
                // leading Pmod_7 of (f, b) goes to Xo, b itself is not changed
                float b;
                if constexpr (HBG) b = bgh[e];
                else b = ((bgn + uo) + toff)[0];
                (Fn + uo)[toff] = z;
                const float mg = msqrt(z2 + b * b + a.alpha);
                const float fac = m * am * mrcp(mg);
                bacc[e] += (BAcc)(b * fac);
                v[e] = z * (fac * a.scale);
              } else if constexpr (MODE == COL_DMBF) {
                // z = S, the spectrum of O P.  From the state (F, b): the
                // leading Pmod_7 (f1, b1), the DM step y = 2 S - f1,
                // b0 = 2 bbar - b1 (bbar = frame mean of b1), Pmod_7(y, b0),
                // F_new = f1 - S + y', b_new = b1 - bbar + b0'; the step size
                // sum |y' - S|^2 / (M K) = sum |f_new - f1|^2 in real space
                static_assert(HBG, "COL_DMBF reads its hoisted background");
                const auto rF = buf_rsrc(Fn), rb = buf_rsrc(bgn), rbb = buf_rsrc(a.bbar);
                const uint32_t so = (uint32_t)u * (uint32_t)K;
                const float b = bgh[e];
                float2 f;
                if constexpr (HF) f = fh[e];
                else f = buf_load2<0>(rF, toff * 8u, so * 8u);
                const float mg1 = msqrt(norm2(f) + b * b + a.alpha);
                const float fac1 = m * am * mrcp(mg1);
                const float b1 = b * fac1;
                const float2 f1 = f * fac1;
                const float bb = buf_load1<0>(rbb, toff * 4u, so * 4u);
                const float2 y = make_float2(2.f * z.x - f1.x, 2.f * z.y - f1.y);
                const float b0 = 2.f * bb - b1;
                const float mg2 = msqrt(norm2(y) + b0 * b0 + a.alpha);
                const float fac2 = m * am * mrcp(mg2);
                const float2 yp = y * fac2;
                const float2 fn = make_float2(f1.x - z.x + yp.x, f1.y - z.y + yp.y);
                const float bn = (b1 - bb) + b0 * fac2;
                acc[0] += (double)norm2(yp - z);
                buf_store2<0>(fn, rF, toff * 8u, so * 8u);
                buf_store1<0>(bn, rb, toff * 4u, so * 4u);
                bnh[e] = bn;
                if constexpr (KEEPF) fnh[e] = fn;
                bacc[e] += (BAcc)bn;
                v[e] = fn * a.scale;
              } else if constexpr (MODE == COL_DM1F) {
                // z = S: the model error of S (dm.py:337-341) when asked for;
                // y = 2 S - f, y' = Pmod_1(y) (masked pixels keep y),
                // F_new = f - S + y', step size sum |y' - S|^2 / (M K)
                if (a.part) {
                  const float d = msqrt(z2) - am;
                  acc[0] += (double)(m * d * d);
                }
                float2* fp = (Fn + uo) + toff;
                float2 f;
                if constexpr (HF) f = fh[e];
                else f = load_stream(fp);
                const float2 y = make_float2(2.f * z.x - f.x, 2.f * z.y - f.y);
                const float mg = msqrt(norm2(y)) + a.alpha;
                const float fac = (m != 0.f) ? am * mrcp(mg) : 1.f;
                const float2 yp = y * fac;
                const float2 fn = make_float2(f.x - z.x + yp.x, f.y - z.y + yp.y);
                acc[1] += (double)norm2(yp - z);
                store_stream(fp, fn);
                v[e] = fn * a.scale;
              }
            }
          }
        }
      }

      if constexpr (MODE == COL_ERA1) {
        acc[0] += (double)fr0;
        acc[1] += (double)fr1;
      }
      if constexpr (TR::store_fwd) {
        if (cv) {
#pragma unroll
          for (int e = 0; e < E; ++e) (Xn + (size_t)rowB<PK>(e) * K)[toff] = v[e];
        }
      }
      if constexpr (TR::inv) {
        if constexpr (TR::fwd) lds_barrier();
        if constexpr (MODE == COL_INV) {
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = v[e] * a.scale;
        }
        fft_inv<PK>(v, mid, s_tw, t);
        const auto rd = buf_rsrc(MODE == COL_DMBF_LEAD ? a.Xo + (size_t)nu * L * K : Xn);
        if (cv) {
#pragma unroll
          for (int e = 0; e < E; ++e) buf_store2(v[e], rd, toff * 8u, (uint32_t)(rowA<PK>(e) * K) * 8u);
        }
      }
      if constexpr (MODE == COL_DMBF) {
        // the next iteration's leading Pmod_7 of (F_new, b_new) into Xo
        if (cv) {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            float2 fn;
            if constexpr (KEEPF) fn = fnh[e];
            else fn = buf_load2<0>(buf_rsrc(Fn), toff * 8u, (uint32_t)(rowB<PK>(e) * K) * 8u);
            const float m = ((mbits >> e) & 1u) ? 1.f : 0.f;
            const float mg = msqrt(norm2(fn) + bnh[e] * bnh[e] + a.alpha);
            const float fac = m * amh[e] * mrcp(mg);
            bacc2[e] += (BAcc)(bnh[e] * fac);
            v[e] = fn * (fac * a.scale);
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = make_float2(0.f, 0.f);
        }
        lds_barrier();
        fft_inv<PK>(v, mid, s_tw, t);
        const auto rXo = buf_rsrc(a.Xo + (size_t)nu * L * K);
        if (cv) {
#pragma unroll
          for (int e = 0; e < E; ++e) buf_store2(v[e], rXo, toff * 8u, (uint32_t)(rowA<PK>(e) * K) * 8u);
        }
      }
      lds_barrier();
    }

    if constexpr (TR::bsum) {
      if (cv) {
        double* pb = a.partb + (size_t)chunk * L * K;
#pragma unroll
        for (int e = 0; e < E; ++e) (pb + (size_t)rowB<PK>(e) * K)[toff] = (double)bacc[e];
      }
    }
    if constexpr (TR::bsum2) {
      if (cv) {
        double* pb = a.partb2 + (size_t)chunk * L * K;
#pragma unroll
        for (int e = 0; e < E; ++e) (pb + (size_t)rowB<PK>(e) * K)[toff] = (double)bacc2[e];
      }
    }
  }
  if constexpr (TR::err) {
    if constexpr (MODE == COL_DMBF) acc[0] *= (double)a.scale;   // sum |f_new - f1|^2
    if constexpr (MODE == COL_DM1F) acc[1] *= (double)a.scale;   // sum |f_new - f|^2
    block_sum<2>(acc, red);
    if (tid == 0) {
      const int pi = a.item_parts ? wg : (int)blockIdx.x;
      if (a.part) a.part[pi] = acc[0];
      if (a.part2) a.part2[pi] = acc[1];
    }
  }
}

// ---------------------------------------------------------------------------
// deterministic overlap gather: object update (psup_O, era.py:336-369)
// ---------------------------------------------------------------------------
constexpr int TILE_H = 16, TILE_W = 64;

struct GatherArgs {
  const float2* S;         // frames psi' (or f): [N][M][K]
  const float2* P;         // probe [M][K]
  const float2* Oold;      // previous object, for the fused eCon (null = skip)
  float2* Oout;            // new object [U][V] (may alias Oold)
  double* H;               // P heatmap [U][V], FP64
  int fresh;               // 1: accumulate sum |P|^2 here instead of reading H
  int storeH;              // with fresh: write the new heatmap to H
  float alpha;
  const int* __restrict__ tile_ptr;     // CSR over object tiles
  const int* __restrict__ tile_frames;  // ascending frame ids per tile
  int tiles_x;
  int ntiles;
  double* part_econ;       // [2][nblocks]  sum |Oold - O|^2, sum |O|^2
  double* part_esup;       // [2][nblocks]  sum Re(conj(O) num), sum |O|^2 H
};

// Each object pixel sums conj(P) * S_n over the frames whose window covers it,
// in ascending frame order, then divides by (sum |P|^2 + alpha) in FP64.
__global__ __launch_bounds__(WG) void k_gather_O(Geom g, GatherArgs a) {
  constexpr int FCAP = 512;  // frames staged in LDS per batch
  __shared__ double red[16];
  __shared__ int s_n[FCAP];
  __shared__ int2 s_o[FCAP];
  const int tid = threadIdx.x, tx = tid & (TILE_W - 1), ty = tid / TILE_W;
  constexpr int PY = TILE_H / (WG / TILE_W);  // pixels per thread (rows)
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  if (tile >= a.ntiles) return;  // grid padded to a multiple of the XCD count
  const int u0 = (tile / a.tiles_x) * TILE_H, v = (tile % a.tiles_x) * TILE_W + tx;
  const int M = g.M, K = g.K;
  float2 acc[PY];
  double hacc[PY];
#pragma unroll
  for (int k = 0; k < PY; ++k) { acc[k] = make_float2(0.f, 0.f); hacc[k] = 0.0; }
  const int b0 = a.tile_ptr[tile], b1 = a.tile_ptr[tile + 1];
  for (int bs = b0; bs < b1; bs += FCAP) {
    const int cnt = min(FCAP, b1 - bs);
    __syncthreads();
    for (int k = tid; k < cnt; k += WG) {
      const int n = a.tile_frames[bs + k];
      s_n[k] = n;
      s_o[k] = g.off[n];
    }
    __syncthreads();
    // four frames per step, predicated rather than skipped, so their loads
    // are in flight together; the sums still take the frames in order
    constexpr int FB = 4;
    for (int b0f = 0; b0f < cnt; b0f += FB) {
      float2 ps[FB][PY], ss[FB][PY];
      bool ok[FB][PY];
#pragma unroll
      for (int f = 0; f < FB; ++f) {
        const int b = min(b0f + f, cnt - 1);
        const int n = s_n[b];
        const int2 o = s_o[b];
        const int j = v - o.y;
        const bool jin = b0f + f < cnt && (unsigned)j < (unsigned)K;
        const float2* Sn = a.S ? a.S + (size_t)n * M * K + (jin ? j : 0) : nullptr;
        const float2* Pj = (g.Pn ? g.Pn + (size_t)n * M * K : frame_probe(g, a.P, n)) + (jin ? j : 0);
#pragma unroll
        for (int k = 0; k < PY; ++k) {
          const int i = u0 + ty + k * (WG / TILE_W) - o.x;
          ok[f][k] = jin && (unsigned)i < (unsigned)M;
          const size_t ro = ok[f][k] ? (size_t)i * K : 0;
          ps[f][k] = ok[f][k] ? Pj[ro] : make_float2(0.f, 0.f);
          ss[f][k] = (ok[f][k] && Sn) ? load_stream(&Sn[ro]) : make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int f = 0; f < FB; ++f) {
#pragma unroll
        for (int k = 0; k < PY; ++k) {
          if (ok[f][k]) {
            acc[k] = acc[k] + cmulc(ss[f][k], ps[f][k]);
            if (a.fresh) hacc[k] += (double)norm2(ps[f][k]);
          }
        }
      }
    }
  }
  if (!a.S) {
    // heatmap only (make_P_heatmap, era.py:402-410): the sums above, in the
    // same frame order, without the exit waves
#pragma unroll
    for (int k = 0; k < PY; ++k) {
      const int u = u0 + ty + k * (WG / TILE_W);
      if (u < g.U && v < g.V) a.H[(size_t)u * g.V + v] = hacc[k];
    }
    return;
  }
  double r[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < PY; ++k) {
    const int u = u0 + ty + k * (WG / TILE_W);
    if (u < g.U && v < g.V) {
      const size_t pix = (size_t)u * g.V + v;
      const double h = a.fresh ? hacc[k] : a.H[pix];
      if (a.fresh && a.storeH) a.H[pix] = h;
      const double den = h + (double)a.alpha;
      const double ox = (double)acc[k].x / den, oy = (double)acc[k].y / den;
      const float2 on = make_float2((float)ox, (float)oy);
      if (a.Oold) {
        const float2 od = a.Oold[pix];
        const double dx = (double)od.x - (double)on.x, dy = (double)od.y - (double)on.y;
        r[0] += dx * dx + dy * dy;
        r[1] += (double)on.x * on.x + (double)on.y * on.y;
      }
      if (a.part_esup) {
        r[2] += (double)on.x * acc[k].x + (double)on.y * acc[k].y;
        r[3] += ((double)on.x * on.x + (double)on.y * on.y) * h;
      }
      a.Oout[pix] = on;
    }
  }
  if (a.part_econ || a.part_esup) {
    block_sum<4>(r, red);
    if (tid == 0) {
      const int nb = a.ntiles;
      if (a.part_econ) { a.part_econ[tile] = r[0]; a.part_econ[nb + tile] = r[1]; }
      if (a.part_esup) { a.part_esup[tile] = r[2]; a.part_esup[nb + tile] = r[3]; }
    }
  }
}

// ---------------------------------------------------------------------------
// inverse row FFT fused with the object update (ERA object-only steps:
// pmod_1's inverse transform, era.py:560-572, followed by psup_O, :336-369)
// ---------------------------------------------------------------------------
struct RowGatherArgs {
  const float2* Y;         // row spectra after the column stage [N][M][K]
  const float2* P;         // probe [M][K]
  const float2* Oold;      // previous object, for the fused eCon (null = skip)
  float2* Oout;            // new object [U][V] (may alias Oold)
  const double* H;         // heatmap sum |P|^2 [U][V] (valid)
  float alpha;
  int segs, segw;          // column segments per object row and their width
  const int* __restrict__ row_ptr;    // CSR over (object row, segment)
  const int2* __restrict__ row_ent;   // per segment, ascending n: (n*M + i, x_n << 16 | i)
  double* part_econ;       // [2][U*segs]  sum |Oold - O|^2, sum |O|^2
  double* part_esup;       // [2][U*segs]  sum Re(conj(O) num), sum |O|^2 H
  float2* Fout;            // when set: the inverse-transformed frame rows [N][M][K]
                           // (each row written once, by the segment holding its first column)
};

// Largest object width the fused kernel's LDS row accumulator takes (and
// that the packed column offset of a CSR entry holds).
constexpr int ROWG_VMAX = 8192;
static_assert(ROWG_VMAX < (1 << 15), "x_n << 16 must stay a positive int");

// s + the parked values of lines [lf, lf + N) at column vv, added in line
// order (a line that does not cover vv gives the zero word at LPW * W).  All N
// reads are issued before the first add.  oyv holds line l's column offset in
// lane l.
template <int L, int W, int LPW, int N>
__device__ __forceinline__ float2 add_lines(float2 s, const float2* lds, int oyv, int lf, int vv) {
  float2 x[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int l = lf + k;
    const int c = vv - __builtin_amdgcn_readlane(oyv, l);
    x[k] = lds[(unsigned)c < (unsigned)L ? l * W + c : LPW * W];
  }
  // keep the adds after the last read: interleaved, they drain the LDS queue
  // once per group of four reads
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < N; ++k) s = s + x[k];
  return s;
}

// add_lines over a wave-uniform line range [lf, lf + n): n is rounded up to
// the next unrolled group size (1..12 -- with the row value, 13 LDS reads in
// flight stay within the 15-deep LDS counter -- then doubling, capped at LPW)
// and the group shifted to stay within [0, LPW); the extra lines miss the
// wave's columns and add +0.0, which leaves every partial sum unchanged (a
// sum that starts at +0.0 never becomes -0.0).  A 64-column window meets the
// lines whose offsets fall in a 319-column range: 8 or 9 at a 40-pixel step.
template <int L, int W, int LPW, int N = 1>
__device__ __forceinline__ float2 add_line_range(float2 s, const float2* lds, int oyv, int lf, int n,
                                                 int vv) {
  if constexpr (N >= LPW) {
    return add_lines<L, W, LPW, LPW>(s, lds, oyv, 0, vv);
  } else {
    if (n <= N) return add_lines<L, W, LPW, N>(s, lds, oyv, min(lf, LPW - N), vv);
    return add_line_range<L, W, LPW, (N < 12 ? N + 1 : 2 * N)>(s, lds, oyv, lf, n, vv);
  }
}

// One work-group per (object row u, column segment [v0, v0 + segw)).  The
// frames whose window covers the segment's part of row u (ascending n) are
// taken LPW at a time, one per line: frame row i = u - y_n is inverse-
// transformed in registers (the column stage already applied the 1/(M K) of
// the inverse), multiplied by conj(P[i, :]) and parked in the line's own LDS
// exchange area; then each column of the segment adds the parked values of
// the lines covering it, in line order.  Every pixel so sums its frames in
// ascending order, as k_gather_O does, but psi' is never written to or read
// back from HBM.  `row` (dynamic LDS, segw float2) accumulates the segment;
// rows are split into segments so that three work-groups fit a CU's LDS (a
// frame row that straddles two segments is transformed by both).
//
// Global loads are issued in the order their results are needed (CSR entries
// two batches ahead, then this batch's probe row, then the next batch's
// spectra), so no wait on the in-order load counter also waits for the
// prefetch of the next batch.
template <int PK>
__global__ __launch_bounds__(WG, 3) void k_rows_gather(Geom g, RowGatherArgs a,
                                                    const float2* __restrict__ tw) {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T, LPW = WG / T;
  constexpr int W = LineLds<PK>::words;
  constexpr int OFF_NONE = -(1 << 29);  // column offset of a spare line: matches no column
  static_assert(W >= L, "a line's exchange area must hold the parked row");
  static_assert(64 % T == 0, "a line must lie within one wave");
  static_assert((LPW & (LPW - 1)) == 0 && LPW <= 64, "line offsets are spread one per lane");
  __shared__ float2 lds[LPW * W + 1];  // + one zero word read by uncovered slots
  __shared__ float2 s_tw[L];
  __shared__ int s_oy[LPW];
  __shared__ double red[16];
  extern __shared__ float2 row[];
  stage_twiddles<L>(s_tw, tw);
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int item = blockIdx.x, u = item / a.segs, M = g.M, V = g.V;
  const int v0 = (item % a.segs) * a.segw, v1 = min(V, v0 + a.segw);
  for (int v = tid; v < v1 - v0; v += WG) row[v] = make_float2(0.f, 0.f);
  if (tid == 0) lds[LPW * W] = make_float2(0.f, 0.f);
  const int b0 = a.row_ptr[item], cnt = a.row_ptr[item + 1] - b0;
  const int nbatch = (cnt + LPW - 1) / LPW;
  // spare lines of the last batch re-read a real entry and are masked by OFF_NONE
  auto entry = [&](int bt) { return a.row_ent[b0 + min(bt * LPW + line, cnt - 1)]; };
  float2 nx[E];
  auto load = [&](int2 m) {
    const size_t rbase = (size_t)m.x * L;
#pragma unroll
    for (int e = 0; e < E; ++e) nx[e] = load_stream(&a.Y[rbase + layB<PK>(e, t)]);
  };
  int2 m_cur = make_int2(0, 0), m_next = make_int2(0, 0);  // .y = x_n << 16 | i
  if (nbatch > 0) {
    m_cur = entry(0);
    m_next = entry(1);
    load(m_cur);
  }
  for (int bt = 0; bt < nbatch; ++bt) {
    const int i = m_cur.y & 0xffff;
    const float2* Prow = probe_row(g, a.P, m_cur.x, i);
    float2 pv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) pv[e] = Prow[layA<PK>(e, t)];
    const int2 m_after = entry(bt + 2);
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nx[e];
    if (bt + 1 < nbatch) load(m_next);
    fft_inv<PK>(v, lds + line * W, s_tw, t);
    if (a.Fout) {
      const int xn = m_cur.y >> 16;
      if (bt * LPW + line < cnt && xn >= v0 && xn < v0 + a.segw) {
        const size_t rbase = (size_t)m_cur.x * L;
#pragma unroll
        for (int e = 0; e < E; ++e) store_stream(&a.Fout[rbase + layA<PK>(e, t)], v[e]);
      }
    }
    // park conj(P) * psi' in this line's area: every value depends on the
    // pass-2 reads of the same wave, so no barrier is needed before it
    float2* park = lds + line * W;
#pragma unroll
    for (int e = 0; e < E; ++e) park[layA<PK>(e, t)] = cmulc(v[e], pv[e]);
    if (t == 0) s_oy[line] = (bt * LPW + line < cnt) ? (m_cur.y >> 16) : OFF_NONE;
    lds_barrier();
    // lane l < LPW: line l's column offset; other lanes hold OFF_NONE
    const int oyv = ((tid & 63) < LPW) ? s_oy[tid & (LPW - 1)] : OFF_NONE;
    int lo = v1, hi = v0;
#pragma unroll
    for (int l = 0; l < LPW; ++l) {
      const int o = __builtin_amdgcn_readlane(oyv, l);
      if (o != OFF_NONE) {
        lo = min(lo, o);
        hi = max(hi, o + L);
      }
    }
    lo = max(lo, v0);
    hi = min(hi, v1);
    // Each wave reads only the lines that meet its 64 columns: with frames a
    // scan step apart a line covers a few of the batch's windows, and reading
    // every line's parked value or the zero word cost twice the LDS reads.
    // Within a window the range's reads are all in flight before the adds.
    const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
    for (int base = lo; base < hi; base += WG) {
      const int w0 = base + wbase, w1 = min(w0 + 64, hi);
      if (w0 >= hi) continue;
      // line meets [w0, w1) <=> oyv in [w0 - L + 1, w1), one unsigned compare
      // (OFF_NONE, far below, wraps to a large unsigned and never meets)
      const uint64_t meet = __builtin_amdgcn_ballot_w64((unsigned)(oyv - (w0 - L + 1)) <
                                                        (unsigned)(w1 - w0 + L - 1));
      if (meet == 0) continue;
      const int lf = __builtin_ctzll(meet), ll = 63 - __builtin_clzll(meet);
      const int vv = base + tid;
      float2 s = (vv < hi) ? row[vv - v0] : make_float2(0.f, 0.f);
      s = add_line_range<L, W, LPW>(s, lds, oyv, lf, ll - lf + 1, vv);
      if (vv < hi) row[vv - v0] = s;
    }
    lds_barrier();
    m_cur = m_next;
    m_next = m_after;
  }
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v = v0 + tid; v < v1; v += WG) {
    const size_t pix = (size_t)u * V + v;
    const float2 acc = row[v - v0];
    const double h = a.H[pix];
    const double den = h + (double)a.alpha;
    const double ox = (double)acc.x / den, oyv = (double)acc.y / den;
    const float2 on = make_float2((float)ox, (float)oyv);
    if (a.Oold) {
      const float2 od = a.Oold[pix];
      const double dx = (double)od.x - (double)on.x, dy = (double)od.y - (double)on.y;
      r[0] += dx * dx + dy * dy;
      r[1] += (double)on.x * on.x + (double)on.y * on.y;
    }
    if (a.part_esup) {
      r[2] += (double)on.x * acc.x + (double)on.y * acc.y;
      r[3] += ((double)on.x * on.x + (double)on.y * on.y) * h;
    }
    a.Oout[pix] = on;
  }
  if (a.part_econ || a.part_esup) {
    block_sum<4>(r, red);
    if (tid == 0) {
      const int nb = gridDim.x;
      if (a.part_econ) { a.part_econ[item] = r[0]; a.part_econ[nb + item] = r[1]; }
      if (a.part_esup) { a.part_esup[item] = r[2]; a.part_esup[nb + item] = r[3]; }
    }
  }
}

// ---------------------------------------------------------------------------
// probe update (psup_P, era.py:371-400): reduction over frames in two stages
// ---------------------------------------------------------------------------
struct ProbeArgs {
  const float2* S;   // frames [N][M][K]
  const float2* O;   // object [U][V]
  double2* pnum;     // [nchunks][M*K]  sum conj(O_n) S_n
  double* pg;        // [nchunks][M*K]  sum |O_n|^2 (fresh only)
  int fresh;
  int chunk;
  const int* cstart;  // packed scans: chunk y takes frames [cstart[y], cstart[y + 1])
                      // (chunks never straddle scans); null: frames of y * chunk
};

__global__ __launch_bounds__(WG) void k_probe_partial(Geom g, ProbeArgs a) {
  const int MK = g.M * g.K;
  // XCD-aware order: each die takes consecutive probe rows of one frame chunk,
  // whose object-window rows overlap (frame rows a jitter apart)
  const int bl = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int bx = bl % gridDim.x, by = bl / gridDim.x;
  const int p = bx * WG + threadIdx.x;
  if (p >= MK) return;
  const int i = p / g.K, j = p % g.K;
  const int n0 = a.cstart ? a.cstart[by] : by * a.chunk;
  const int n1 = a.cstart ? a.cstart[by + 1] : min(g.N, n0 + a.chunk);
  double nx = 0.0, ny = 0.0, gg = 0.0;
  // eight frames' loads in flight per step; the sums take the frames in order
  constexpr int FB = 8;
  for (int nb = n0; nb < n1; nb += FB) {
    float2 ov[FB], sv[FB];
#pragma unroll
    for (int f = 0; f < FB; ++f) {
      const int n = min(nb + f, n1 - 1);
      const int2 o = g.off[n];
      ov[f] = a.O[(size_t)(o.x + i) * g.V + o.y + j];
      // psi' is read once: streaming loads keep the reused object rows in L2
      sv[f] = load_stream(&a.S[(size_t)n * MK + p]);
    }
#pragma unroll
    for (int f = 0; f < FB; ++f) {
      if (nb + f < n1) {
        const float2 q = cmulc(sv[f], ov[f]);
        nx += (double)q.x;
        ny += (double)q.y;
        if (a.fresh) gg += (double)norm2(ov[f]);
      }
    }
  }
  a.pnum[(size_t)by * MK + p] = make_double2(nx, ny);
  if (a.fresh) a.pg[(size_t)by * MK + p] = gg;
}

struct ProbeFinalArgs {
  const double2* pnum;
  const double* pg;
  int nchunks;
  int fresh, storeG;
  double* G;          // O heatmap [M][K]
  float alpha;
  const float2* Pold; // for the fused eCon (null = skip)
  float2* Pout;
  double* part_econ;  // [2][nblocks]
  double* part_esup;  // [2][nblocks]  sum Re(conj(P) pnum), sum |P|^2 G (null = skip)
  const int* sc;      // packed scans: scan y = blockIdx.y sums chunks [sc[y], sc[y + 1]) into
                      // its own probe (Pold, Pout, G + y M K) and partials (+ y 2 nblocks)
};

// With part_esup, also the two probe-side terms of the eSup identity
// sum |psi' - P O|^2 = sum |psi'|^2 - 2 Re sum conj(P O) psi' + sum |P O|^2:
// pnum and G hold sum_n conj(O_n) psi'_n and sum_n |O_n|^2 of the final O.
__global__ __launch_bounds__(WG) void k_probe_final(Geom g, ProbeFinalArgs a) {
  __shared__ double red[16];
  const int MK = g.M * g.K;
  const int p = blockIdx.x * WG + threadIdx.x;
  const int sy = blockIdx.y;
  const int c0 = a.sc ? a.sc[sy] : 0, c1 = a.sc ? a.sc[sy + 1] : a.nchunks;
  if (a.sc) {
    const size_t so = (size_t)sy * MK;
    if (a.G) a.G += so;
    if (a.Pold) a.Pold += so;
    a.Pout += so;
    if (a.part_econ) a.part_econ += (size_t)sy * 2 * gridDim.x;
    if (a.part_esup) a.part_esup += (size_t)sy * 2 * gridDim.x;
  }
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (p < MK) {
    double nx = 0.0, ny = 0.0, gg = 0.0;
    constexpr int CB = 4;  // chunks' loads in flight per step, summed in order
    for (int cb = c0; cb < c1; cb += CB) {
      double2 q[CB];
      double pgv[CB];
#pragma unroll
      for (int f = 0; f < CB; ++f) {
        const int c = min(cb + f, c1 - 1);
        q[f] = a.pnum[(size_t)c * MK + p];
        pgv[f] = a.fresh ? a.pg[(size_t)c * MK + p] : 0.0;
      }
#pragma unroll
      for (int f = 0; f < CB; ++f) {
        if (cb + f < c1) {
          nx += q[f].x;
          ny += q[f].y;
          if (a.fresh) gg += pgv[f];
        }
      }
    }
    if (!a.fresh) gg = a.G[p];
    else if (a.storeG) a.G[p] = gg;
    const double den = gg + (double)a.alpha;
    const float2 pn = make_float2((float)(nx / den), (float)(ny / den));
    if (a.Pold) {
      const float2 po = a.Pold[p];
      const double dx = (double)po.x - pn.x, dy = (double)po.y - pn.y;
      r[0] = dx * dx + dy * dy;
      r[1] = (double)pn.x * pn.x + (double)pn.y * pn.y;
    }
    if (a.part_esup) {
      r[2] = (double)pn.x * nx + (double)pn.y * ny;
      r[3] = ((double)pn.x * pn.x + (double)pn.y * pn.y) * gg;
    }
    a.Pout[p] = pn;
  }
  if (a.part_econ || a.part_esup) {
    block_sum<4>(r, red);
    if (threadIdx.x == 0) {
      if (a.part_econ) {
        a.part_econ[blockIdx.x] = r[0];
        a.part_econ[gridDim.x + blockIdx.x] = r[1];
      }
      if (a.part_esup) {
        a.part_esup[blockIdx.x] = r[2];
        a.part_esup[gridDim.x + blockIdx.x] = r[3];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// out[n, i, j] = O[y_n + i, x_n + j] * P[i, j]   (make_exits, era.py:426-433)
__global__ __launch_bounds__(WG) void k_make_exits(Geom g, const float2* O, const float2* P,
                                                   float2* out) {
  const size_t MK = (size_t)g.M * g.K, total = MK * g.N;
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * WG) {
    const int n = (int)(idx / MK);
    const int p = (int)(idx % MK), i = p / g.K, j = p % g.K;
    const int2 o = g.off[n];
    out[idx] = cmul(O[(size_t)(o.x + i) * g.V + o.y + j], g.Pn ? g.Pn[idx] : frame_probe(g, P, n)[p]);
  }
}

// f_n += scale * O_win,n * P  (HIO's Psup term, utility_Ptych.py:16)
__global__ __launch_bounds__(WG) void k_add_exits(Geom g, const float2* O, const float2* P,
                                                  float scale, float2* f) {
  const size_t MK = (size_t)g.M * g.K, total = MK * g.N;
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * WG) {
    const int n = (int)(idx / MK);
    const int p = (int)(idx % MK), i = p / g.K, j = p % g.K;
    const int2 o = g.off[n];
    const float2 e = cmul(O[(size_t)(o.x + i) * g.V + o.y + j], P[p]);
    const float2 v = f[idx];
    f[idx] = make_float2(v.x + scale * e.x, v.y + scale * e.y);
  }
}

// partials of sum |psi'_n - O_n P|^2 (support error on an explicit pass)
__global__ __launch_bounds__(WG) void k_esup(Geom g, const float2* S, const float2* O,
                                             const float2* P, double* part) {
  // frame rows (n, i) in a grid-stride loop, blockDim.x / K rows at a time
  // (the host launches a multiple of K threads), threads across the row: the
  // row / column split is one 32-bit division per row
  __shared__ double red[4];
  const int K = g.K, M = g.M;
  const int rpi = (int)blockDim.x / K;
  const int sub = threadIdx.x / K, j = threadIdx.x % K;
  const int nrows = g.N * M;
  double r[1] = {0.0};
  {
    for (int row = blockIdx.x * rpi + sub; row < nrows; row += gridDim.x * rpi) {
      const int n = row / M, i = row - n * M;
      const int2 o = g.off[n];
      const float2 d = S[(size_t)row * K + j] -
                       cmul(O[(size_t)(o.x + i) * g.V + o.y + j], probe_row(g, P, row, i)[j]);
      r[0] += (double)norm2(d);
    }
  }
  block_sum<1>(r, red);
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

// Pn[n, ky, kx] = Phat[ky, kx] exp(-2 pi i (d_n0 f(ky) / M + d_n1 f(kx) / K)):
// the spectrum of the probe shifted by d_n (y[r] = x[r - d]), with the
// frequency of bin k taken as ((k + N/2 - 1) mod N) - (N/2 - 1), i.e. in
// [-(N/2 - 1), N/2] (bagOfns.py:729-764 with phase_ramp, :966-979).
// deriv = 1 / 2: times -2 pi i f(ky) / M / -2 pi i f(kx) / K, the spectrum of
// d Pn / d d_n0 / d d_n1.
__global__ __launch_bounds__(WG) void k_subpixel_ramp(const float2* __restrict__ Phat,
                                                      const double2* __restrict__ d, float2* Pn,
                                                      int N, int M, int K, int deriv) {
  const size_t MK = (size_t)M * K, total = MK * N;
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * WG) {
    const int n = (int)(idx / MK), p = (int)(idx % MK);
    const int ky = p / K, kx = p % K;
    const int fy = (ky + M / 2 - 1) % M - (M / 2 - 1);
    const int fx = (kx + K / 2 - 1) % K - (K / 2 - 1);
    const double2 dn = d[n];
    const double ang = -2.0 * M_PI * (dn.x * fy / M + dn.y * fx / K);
    double sn, cs;
    sincos(ang, &sn, &cs);
    const float2 h = Phat[p];
    double re = h.x * cs - h.y * sn, im = h.x * sn + h.y * cs;
    if (deriv) {
      // times -i w: (re, im) -> (w im, -w re)
      const double w = deriv == 1 ? 2.0 * M_PI * fy / M : 2.0 * M_PI * fx / K;
      const double t = re;
      re = w * im;
      im = -w * t;
    }
    Pn[idx] = make_float2((float)re, (float)im);
  }
}

// Position terms of frame n = blockIdx.x from its spectra S = FFT(O_win P_n)
// and Dy / Dx = FFT(O_win dP_n/dd_n0 / dd_n1), amplitudes amp (ifftshifted)
// and the ifftshifted mask: with r = |S| - amp on good pixels and
// J_a = Re(conj(S) D_a) / |S|, out[n] = (sum r^2, 2 sum r J_y, 2 sum r J_x,
// 2 sum J_y J_y, 2 sum J_y J_x, 2 sum J_x J_x) in FP64: the modulus error of
// the frame, its gradient in the frame's shift and the Gauss-Newton matrix.
__global__ __launch_bounds__(WG) void k_position_terms(const float2* __restrict__ S,
                                                       const float2* __restrict__ Dy,
                                                       const float2* __restrict__ Dx,
                                                       const float* __restrict__ amp,
                                                       const uint8_t* __restrict__ mask, int MK,
                                                       double* __restrict__ out) {
  __shared__ double red[4 * 6];
  const size_t base = (size_t)blockIdx.x * MK;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int p = threadIdx.x; p < MK; p += WG) {
    if (mask && !mask[p]) continue;
    const float2 s = S[base + p], dy = Dy[base + p], dx = Dx[base + p];
    const double m = sqrt((double)s.x * s.x + (double)s.y * s.y);
    const double r = m - (double)amp[base + p];
    double jy = 0.0, jx = 0.0;
    if (m > 0.0) {
      jy = ((double)s.x * dy.x + (double)s.y * dy.y) / m;
      jx = ((double)s.x * dx.x + (double)s.y * dx.y) / m;
    }
    v[0] += r * r;
    v[1] += r * jy;
    v[2] += r * jx;
    v[3] += jy * jy;
    v[4] += jy * jx;
    v[5] += jx * jx;
  }
  block_sum<6>(v, red);
  if (threadIdx.x == 0) {
    out[6 * blockIdx.x] = v[0];
    for (int k = 1; k < 6; ++k) out[6 * blockIdx.x + k] = 2.0 * v[k];
  }
}

// partials of sum |bak - cur|^2 and sum |cur|^2 (convergence error)
__global__ __launch_bounds__(WG) void k_econ(const float2* cur, const float2* bak, size_t count,
                                             double* part) {
  __shared__ double red[8];
  double r[2] = {0.0, 0.0};
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < count;
       idx += (size_t)gridDim.x * WG) {
    const float2 c = cur[idx], b = bak[idx];
    const double dx = (double)b.x - c.x, dy = (double)b.y - c.y;
    r[0] += dx * dx + dy * dy;
    r[1] += (double)c.x * c.x + (double)c.y * c.y;
  }
  block_sum<2>(r, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r[0];
    part[gridDim.x + blockIdx.x] = r[1];
  }
}

// centre of mass of |P|^2 -> cm = rint(com) - shape // 2   (era.py:172-176)
__global__ __launch_bounds__(WG) void k_com(const float2* P, int M, int K, int* cm) {
  __shared__ double red[12];
  double r[3] = {0.0, 0.0, 0.0};
  for (int p = threadIdx.x; p < M * K; p += WG) {
    const double a = (double)norm2(P[p]);
    r[0] += a;
    r[1] += a * (double)(p / K);
    r[2] += a * (double)(p % K);
  }
  block_sum<3>(r, red);
  if (threadIdx.x == 0) {
    if (r[0] > 0.0) {
      cm[0] = (int)rint(r[1] / r[0]) - M / 2;
      cm[1] = (int)rint(r[2] / r[0]) - K / 2;
    } else {
      cm[0] = 0;
      cm[1] = 0;
    }
  }
}

// dst[b, i, j] = src[b, (i + cm0) mod rows, (j + cm1) mod cols]  (roll by -cm)
__global__ __launch_bounds__(WG) void k_roll(const float2* src, float2* dst, int batch, int rows,
                                             int cols, const int* cm) {
  const size_t plane = (size_t)rows * cols, total = plane * batch;
  const int s0 = ((cm[0] % rows) + rows) % rows, s1 = ((cm[1] % cols) + cols) % cols;
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * WG) {
    const size_t b = idx / plane;
    const int p = (int)(idx % plane), i = p / cols, j = p % cols;
    int is = i + s0; if (is >= rows) is -= rows;
    int js = j + s1; if (js >= cols) js -= cols;
    dst[idx] = src[b * plane + (size_t)is * cols + js];
  }
}

// bbar[p] = (sum over chunks of partb[c][p]) / N
// dst[i] = src[i % n_per] for i < n (per-frame copies of one background)
__global__ __launch_bounds__(WG) void k_broadcast(float* __restrict__ dst,
                                                  const float* __restrict__ src, size_t n_per,
                                                  size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i % n_per];
}

// One work-group per BG_PIX pixels; its WG / BG_PIX chunk lanes take chunks
// cl, cl + lanes, ... with eight loads in flight each (eight accumulators
// combined pairwise), then the lanes are added in lane order: a fixed order
// (the one-thread-per-pixel loop of 256 dependent loads ran at ~1.2 TB/s).
constexpr int BG_PIX = 64;
__global__ __launch_bounds__(WG) void k_bg_mean(const double* __restrict__ partb, int nchunks,
                                                int MK, int N, float* bbar) {
  constexpr int CL = WG / BG_PIX;
  __shared__ double red[CL][BG_PIX];
  const int pl = threadIdx.x % BG_PIX, cl = threadIdx.x / BG_PIX;
  const int p = blockIdx.x * BG_PIX + pl;
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (p < MK) {
    int c = cl;
    for (; c + 7 * CL < nchunks; c += 8 * CL) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += partb[(size_t)(c + k * CL) * MK + p];
    }
    for (; c < nchunks; c += CL) a[7] += partb[(size_t)c * MK + p];
  }
  red[cl][pl] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  if (cl == 0 && p < MK) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CL; ++k) s += red[k][pl];
    bbar[p] = (float)(s / (double)N);
  }
}

// amp[n, i, j] = sqrt(I[n, (i + M/2) % M, (j + K/2) % K]) (ifftshift, era.py:497-498)
// and partials of sum mask * I with the unshifted mask (era.py:495).
__global__ __launch_bounds__(WG) void k_prepare_amp(const float* I, int N, int M, int K,
                                                    const uint8_t* mask_centred, float* amp,
                                                    double* part) {
  __shared__ double red[4];
  const size_t MK = (size_t)M * K, total = MK * N;
  double r[1] = {0.0};
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * WG) {
    const size_t n = idx / MK;
    const int p = (int)(idx % MK), i = p / K, j = p % K;
    const int is = (i + M / 2) % M, js = (j + K / 2) % K;
    const float v = I[n * MK + (size_t)is * K + js];
    amp[idx] = sqrtf(v);
    const float m = mask_centred ? (float)mask_centred[(size_t)is * K + js] : 1.f;
    r[0] += (double)m * (double)v;
  }
  block_sum<1>(r, red);
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

// Back-projection spectra (back_projection.py:75-78): X[n][p] = mask[p] amp[n][p]
// Pf[p] / |Pf[p]| with Pf = FFT2(P) (natural order, like amp and mask); a zero
// Pf gives phase 1 (numpy's angle(0) = 0).
__global__ __launch_bounds__(WG) void k_bp_spectra(float2* __restrict__ X,
                                                   const float* __restrict__ amp,
                                                   const uint8_t* __restrict__ mask,
                                                   const float2* __restrict__ Pf, size_t MK,
                                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t p = i % MK;
    const float2 f = Pf[p];
    const float m = sqrtf(norm2(f));
    const float2 ph = m > 0.f ? make_float2(f.x / m, f.y / m) : make_float2(1.f, 0.f);
    const float a = amp[i] * (mask ? (float)mask[p] : 1.f);
    X[i] = make_float2(a * ph.x, a * ph.y);
  }
}

__global__ __launch_bounds__(WG) void k_abs(const float2* X, float* out, size_t count) {
  for (size_t idx = (size_t)blockIdx.x * WG + threadIdx.x; idx < count;
       idx += (size_t)gridDim.x * WG)
    out[idx] = sqrtf(norm2(X[idx]));
}

// ---------------------------------------------------------------------------
// per-iteration metrics: fixed-order sums of the partial regions
// ---------------------------------------------------------------------------
enum Slot {
  S_EMOD = 0,   // sum mask (Mg - amp)^2
  S_ECON_ON,    // sum |Obak - O|^2
  S_ECON_OD,    // sum |O|^2
  S_ECON_PN,    // sum |Pbak - P|^2
  S_ECON_PD,    // sum |P|^2
  S_ESUP_E,     // sum |psi' - P O_n|^2 (explicit, real space) or DM step
  S_ESUP_A,     // sum |psi_hat'|^2 (spectrum after the modulus step)
  S_ESUP_X,     // sum Re(conj(O) num)
  S_ESUP_H,     // sum |O|^2 H
  NSLOT
};

struct Regions {
  const double* p[NSLOT];
  int n[NSLOT];
};

enum EsupMode { ESUP_NONE = 0, ESUP_EXPLICIT = 1, ESUP_IDENTITY = 2 };

// Fixed-order strided sum p[tid] + p[tid + nt] + ... over eight independent
// accumulators (eight loads in flight per thread), combined pairwise.
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int n) {
  const int nt = blockDim.x;
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 7 * nt < n; i += 8 * nt) {
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] += p[i + j * nt];
  }
  for (; i < n; i += nt) a[7] += p[i];
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// metrics[0..2] (or, with a slot counter, metrics[3*slot .. 3*slot+2] followed
// by slot += 1, so one captured graph serves every iteration)
__global__ __launch_bounds__(1024) void k_metrics(Regions rg, double I_norm, double MK,
                                                  int esup_mode, double* metrics, int* slot) {
  __shared__ double red[16 * NSLOT];
  double s[NSLOT];
#pragma unroll
  for (int k = 0; k < NSLOT; ++k) s[k] = rg.p[k] ? strided_sum(rg.p[k], rg.n[k]) : 0.0;
  block_sum<NSLOT>(s, red);
  if (threadIdx.x == 0) {
    if (slot) {
      metrics += 3 * (size_t)(*slot);
      *slot += 1;
    }
    metrics[0] = sqrt(s[S_EMOD] / I_norm);
    const double den = s[S_ECON_OD] + s[S_ECON_PD];
    metrics[1] = sqrt((s[S_ECON_ON] + s[S_ECON_PN]) / den);
    double es = 0.0;
    if (esup_mode == ESUP_EXPLICIT) {
      es = sqrt(MK * s[S_ESUP_E] / I_norm);
    } else if (esup_mode == ESUP_IDENTITY) {
      const double q = s[S_ESUP_A] - MK * (2.0 * s[S_ESUP_X] - s[S_ESUP_H]);
      es = sqrt((q > 0.0 ? q : 0.0) / I_norm);
    } else {
      es = __builtin_nan("");
    }
    metrics[2] = es;
  }
}

// Packed scans: where scan s's partials of one slot lie in p[k].
enum RegionKind {
  RK_COLS = 0,   // one per (frame, column block) item: [fstart[s] ncb, fstart[s + 1] ncb)
  RK_TILES,      // one per 16 x 64 object tile: the tile rows of the scan's object rows
  RK_ROWS,       // one per (object row, column segment): the scan's object rows x segs
  RK_PROBE,      // probe reductions: [2 pb s, 2 pb s + pb) (p[k] already holds + pb for the
                 // second half)
  RK_ESUP,       // support-error pass per scan: [esup_blocks s, esup_blocks (s + 1))
};

struct ScanRegions {
  const double* p[NSLOT];
  int kind[NSLOT];
  const int* fstart;      // [S + 1] first frame of each scan
  const int* rstart;      // [S + 1] first object row of each scan (multiples of TILE_H)
  const double* I_norm;   // [S] sum mask I of each scan
  int S, ncb, tiles_x, segs, pb, esup_blocks;
};

// k_metrics for packed scans: one work-group takes the scans in turn and
// writes metrics[3 (slot S + s) ..] (slot = the iteration's metrics slot).
__global__ __launch_bounds__(1024) void k_metrics_scans(ScanRegions rg, double MK, int esup_mode,
                                                        double* metrics, int* slot) {
  __shared__ double red[16 * NSLOT];
  int base = 0;
  if (slot) base = *slot;   // read by every thread before thread 0 advances it below
  for (int s = 0; s < rg.S; ++s) {
    double v[NSLOT];
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      v[k] = 0.0;
      if (!rg.p[k]) continue;
      int b = 0, n = 0;
      switch (rg.kind[k]) {
        case RK_COLS: b = rg.fstart[s] * rg.ncb; n = rg.fstart[s + 1] * rg.ncb - b; break;
        case RK_TILES:
          // (rounded up: only the last scan's end may not be a tile boundary)
          b = (rg.rstart[s] + TILE_H - 1) / TILE_H * rg.tiles_x;
          n = (rg.rstart[s + 1] + TILE_H - 1) / TILE_H * rg.tiles_x - b;
          break;
        case RK_ROWS: b = rg.rstart[s] * rg.segs; n = rg.rstart[s + 1] * rg.segs - b; break;
        case RK_PROBE: b = 2 * rg.pb * s; n = rg.pb; break;
        default: b = rg.esup_blocks * s; n = rg.esup_blocks; break;
      }
      v[k] = strided_sum(rg.p[k] + b, n);
    }
    block_sum<NSLOT>(v, red);
    if (threadIdx.x == 0) {
      double* m = metrics + 3 * ((size_t)base * rg.S + s);
      const double In = rg.I_norm[s];
      m[0] = sqrt(v[S_EMOD] / In);
      const double den = v[S_ECON_OD] + v[S_ECON_PD];
      m[1] = sqrt((v[S_ECON_ON] + v[S_ECON_PN]) / den);
      double es = __builtin_nan("");
      if (esup_mode == ESUP_EXPLICIT) {
        es = sqrt(MK * v[S_ESUP_E] / In);
      } else if (esup_mode == ESUP_IDENTITY) {
        const double q = v[S_ESUP_A] - MK * (2.0 * v[S_ESUP_X] - v[S_ESUP_H]);
        es = sqrt((q > 0.0 ? q : 0.0) / In);
      }
      m[2] = es;
    }
  }
  __syncthreads();
  if (slot && threadIdx.x == 0) *slot = base + 1;
}

// plain sum of one region into out[0]
__global__ __launch_bounds__(1024) void k_sum(const double* p, int n, double* out) {
  __shared__ double red[16];
  double s[1] = {strided_sum(p, n)};
  block_sum<1>(s, red);
  if (threadIdx.x == 0) out[0] = s[0];
}

}  // namespace pty
