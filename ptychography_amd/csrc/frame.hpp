// Fused per-frame kernel: the whole 2-D transform pipeline of one frame in
// one work-group.
//
//   phase 1  rows:    psi = O_win P (or 2 O_win P - f), forward row FFT -> slot
//   phase 2  columns: forward column FFT, modulus step / error, inverse -> slot
//   phase 3  rows:    inverse row FFT -> psi' (ERA) or the DM update of f
//
// The slot is a per-work-group L x L complex64 scratch area in global memory.
// The grid is sized to the device (one or two work-groups per CU), so all
// slots together are 128-256 MB and are re-used frame after frame: they are
// meant to live in the 256 MB Infinity Cache, leaving HBM with the compulsory
// streams only (amplitudes in, psi' out, and f for DM).  Frames are taken
// n = blockIdx.x, blockIdx.x + gridDim.x, ... and every work-group writes one
// FP64 partial per sum, so all reductions stay deterministic.
#pragma once
#include "kernels.hpp"

namespace pty {

template <int L, int NT>
struct FrameCfg {
  static constexpr int R1 = Plan<L>::R1, R2 = Plan<L>::R2, T = Plan<L>::T, E = L / T;
  static constexpr int LPR = NT / T;                    // lines per round
  static constexpr int ROUNDS = (L + LPR - 1) / LPR;
  static constexpr int W = LineLds<L>::words;
};

enum FrameMode {
  FR_ERA1 = 0,   // O_win P -> Pmod_1 with eMod -> psi'        (ERA methods 1-3)
  FR_DM1,        // 2 O_win P - f -> Pmod_1 -> f += psi'' - O_win P  (DM methods 1-3)
  FR_ERR1,       // O_win P -> model_error_1 (forward only)  (DM metric)
  FR_INTENS,     // O_win P -> fftshift(|psi_hat|^2)         (info['I'])
};

struct FrameArgs {
  const float2* O;
  const float2* P;
  float2* F;            // DM iterate f [N][L][L] (in/out)
  float2* Xout;         // ERA: psi' [N][L][L]
  float2* slots;        // [gridDim.x][L][L]
  const float* amp;     // [N][L][L] ifftshifted
  const uint8_t* mask;  // [L][L] ifftshifted, null = all good
  float alpha;
  float scale;          // 1 / L^2
  float* outI;          // FR_INTENS
  double* part;         // [gridDim.x] eMod / error sums
  double* part2;        // [gridDim.x] sum |psi_hat'|^2 (ERA) or DM step sum
  const float2* __restrict__ tw;
};

template <int L, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_frame(Geom g, FrameArgs a) {
  using C = FrameCfg<L, NT>;
  constexpr int R1 = C::R1, R2 = C::R2, T = C::T, E = C::E, LPR = C::LPR, W = C::W;
  constexpr bool INVERSE = MODE == FR_ERA1 || MODE == FR_DM1;
  __shared__ float2 lds[LPR * W];
  __shared__ float2 s_tw[L];
  __shared__ double red[2 * (NT / 64)];
  stage_twiddles<L>(s_tw, a.tw);
  const int tid = threadIdx.x;
  float2* S = a.slots + (size_t)blockIdx.x * L * L;
  double acc[2] = {0.0, 0.0};

  for (int n0 = xcd_remap(blockIdx.x, gridDim.x); n0 < g.N; n0 += gridDim.x) {
    const int n = __builtin_amdgcn_readfirstlane(n0);
    const int2 o = g.off[n];
    const size_t fbase = (size_t)n * L * L;

    // ---- phase 1: frame formation + forward row FFT -> slot ----
    {
      const int line = tid / T, t = tid % T;
      for (int rd = 0; rd < C::ROUNDS; ++rd) {
        const int i = rd * LPR + line;
        const bool valid = i < L;
        float2 v[E];
        if (valid) {
          const float2* Orow = a.O + (size_t)(o.x + i) * g.V + o.y;
          const float2* Prow = a.P + (size_t)i * L;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int c = idxA<R1, R2, T>(e, t);
            float2 psi = cmul(Orow[c], Prow[c]);
            if constexpr (MODE == FR_DM1) {
              const float2 f = a.F[fbase + (size_t)i * L + c];
              psi = make_float2(2.f * psi.x - f.x, 2.f * psi.y - f.y);
            }
            v[e] = psi;
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = make_float2(0.f, 0.f);
        }
        line_fft<L, R1, R2, T, false>(v, lds + line * W, s_tw, t);
        if (valid) {
#pragma unroll
          for (int e = 0; e < E; ++e) S[(size_t)i * L + idxB<R1, R2, T>(e, t)] = v[e];
        }
        __syncthreads();
      }
    }

    // ---- phase 2: column FFT, modulus / error, inverse column FFT ----
    {
      const int line = tid % LPR, t = tid / LPR;
      for (int rd = 0; rd < C::ROUNDS; ++rd) {
        const int c = rd * LPR + line;
        const bool valid = c < L;
        const uint32_t toff = (uint32_t)(t * L + c);
        float2 v[E];
        if (valid) {
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = (S + (size_t)urowA<R1, R2, T>(e) * L)[toff];
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = make_float2(0.f, 0.f);
        }
        line_fft<L, R1, R2, T, false>(v, lds + line * W, s_tw, t);
        if (valid) {
          const float* ampn = a.amp + fbase;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int u = urowB<R1, R2, T>(e);
            const size_t uo = (size_t)u * L;
            const float2 z = v[e];
            const float z2 = norm2(z);
            if constexpr (MODE == FR_INTENS) {
              const int r = t + u;
              a.outI[fbase + (size_t)((r + L / 2) % L) * L + (c + L / 2) % L] = z2;
            } else {
              const float m = a.mask ? (float)(a.mask + uo)[toff] : 1.f;
              const float am = (ampn + uo)[toff];
              if constexpr (MODE == FR_ERR1) {
                const float d = msqrt(z2) - am;
                acc[0] += (double)(m * d * d);
              } else {
                const float mg = msqrt(z2) + a.alpha;
                if constexpr (MODE == FR_ERA1) {
                  const float d = mg - am;
                  acc[0] += (double)(m * d * d);
                }
                const float fac = (m != 0.f) ? am * mrcp(mg) : 1.f;
                const float2 zn = z * fac;
                if constexpr (MODE == FR_ERA1) acc[1] += (double)norm2(zn);
                v[e] = zn * a.scale;
              }
            }
          }
        }
        if constexpr (INVERSE) {
          __syncthreads();
          line_fft<L, R2, R1, T, true>(v, lds + line * W, s_tw, t);
          if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) (S + (size_t)urowA<R1, R2, T>(e) * L)[toff] = v[e];
          }
        }
        __syncthreads();
      }
    }

    // ---- phase 3: inverse row FFT -> psi' (ERA) or DM update of f ----
    if constexpr (INVERSE) {
      const int line = tid / T, t = tid % T;
      for (int rd = 0; rd < C::ROUNDS; ++rd) {
        const int i = rd * LPR + line;
        const bool valid = i < L;
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
          v[e] = valid ? S[(size_t)i * L + idxA<R2, R1, T>(e, t)] : make_float2(0.f, 0.f);
        line_fft<L, R2, R1, T, true>(v, lds + line * W, s_tw, t);
        if (valid) {
          const size_t rb = fbase + (size_t)i * L;
          if constexpr (MODE == FR_ERA1) {
#pragma unroll
            for (int e = 0; e < E; ++e) a.Xout[rb + idxB<R2, R1, T>(e, t)] = v[e];
          } else {
            const float2* Orow = a.O + (size_t)(o.x + i) * g.V + o.y;
            const float2* Prow = a.P + (size_t)i * L;
#pragma unroll
            for (int e = 0; e < E; ++e) {
              const int c = idxB<R2, R1, T>(e, t);
              const float2 d = v[e] - cmul(Orow[c], Prow[c]);
              acc[1] += (double)norm2(d);
              a.F[rb + c] = a.F[rb + c] + d;
            }
          }
        }
        __syncthreads();
      }
    }
  }

  if constexpr (MODE != FR_INTENS) {
    block_sum<2>(acc, red);
    if (tid == 0) {
      if (a.part) a.part[blockIdx.x] = acc[0];
      if (a.part2) a.part2[blockIdx.x] = acc[1];
    }
  }
}

}  // namespace pty
