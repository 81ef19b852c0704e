// BF16 MFMA kernels of the learned initialiser (a small 3x3-conv
// encoder-decoder that maps a frame's sqrt-intensity to a complex object patch).
//
// Activations are NHWC bf16, weights [9*Cin][Cout] bf16 (tap-major, then input
// channel) plus a per-tap transposed copy [9][Cout][Cin] for the forward pass;
// accumulation is FP32 on v_mfma_f32_32x32x16_bf16.
//
//   k_conv   direct 3x3 convolution for the forward pass and for dgrad (flipped
//            taps): a work-group owns 128 output pixels (whole rows) x BN output
//            channels; per 32-channel chunk the (rows+2) x (W+2) input halo and
//            the 9 x BN x 32 weights are staged in LDS once and all nine taps
//            are read from there with 16-byte fragment loads.
//   k_igemm  generic implicit GEMM (64x64 tile) for weight gradients and for
//            shapes k_conv does not take; wgrad splits the pixel range over
//            fixed partial buffers summed in order (deterministic).
//   k_conv_c1 / k_conv_nsmall / k_dgrad_nsmall   VALU kernels for the first
//            (1 input channel) and last (2 output channels) layers.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "ptyhip.h"

namespace {

thread_local std::string g_nn_err;

int nn_fail(int code, const std::string& m) {
  g_nn_err = m;
  return code;
}

#define PTY_NN_TRY(expr)      \
  do {                        \
    const int r_ = (expr);    \
    if (r_ != PTY_OK) return r_; \
  } while (0)

#define NN_LAUNCHED(name)                                                               \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return nn_fail(PTY_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ f32x16 mfma16(uint4 a, uint4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

__device__ __forceinline__ float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// 8 bf16 <-> 8 floats
__device__ __forceinline__ void unpack8(uint4 v, float* f) {
  const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    f[2 * h] = bf2f((uint16_t)u[h]);
    f[2 * h + 1] = bf2f((uint16_t)(u[h] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float* f) {
  uint32_t u[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) u[h] = (uint32_t)f2bf(f[2 * h]) | ((uint32_t)f2bf(f[2 * h + 1]) << 16);
  return make_uint4(u[0], u[1], u[2], u[3]);
}

constexpr int GT = 64;   // work-group tile (M and N)
constexpr int KC = 32;   // K step staged in LDS
constexpr int NTH = 256;

// ---------------------------------------------------------------------------
// operand loaders: element (row, col) of A (M x K) and B (K x N) as bf16 bits
// ---------------------------------------------------------------------------
struct ConvGeom {
  int B, H, W, Cin, Cout;
};

// forward: A(m, k) = x[b][y+dy][x+dx][ci], m = (b,y,x), k = tap*Cin + ci
struct LoadFwdA {
  const uint16_t* x;
  ConvGeom g;
  __device__ uint16_t operator()(int m, int k) const {
    const int HW = g.H * g.W;
    if (k >= 9 * g.Cin) return 0;
    const int tap = k / g.Cin, ci = k - tap * g.Cin;
    const int b = m / HW, p = m - b * HW, y = p / g.W + tap / 3 - 1, xx = p % g.W + tap % 3 - 1;
    if ((unsigned)y >= (unsigned)g.H || (unsigned)xx >= (unsigned)g.W) return 0;
    return x[(((size_t)b * g.H + y) * g.W + xx) * g.Cin + ci];
  }
};
struct LoadFwdB {  // B(k, n) = w[k][n]
  const uint16_t* w;
  ConvGeom g;
  __device__ uint16_t operator()(int k, int n) const {
    if (k >= 9 * g.Cin) return 0;
    return w[(size_t)k * g.Cout + n];
  }
};
// dgrad: A(m, k) = dy[b][y-dy'][x-dx'][co] with tap (dy', dx'), k = tap*Cout + co
struct LoadDgA {
  const uint16_t* dy;
  ConvGeom g;
  __device__ uint16_t operator()(int m, int k) const {
    const int HW = g.H * g.W;
    if (k >= 9 * g.Cout) return 0;
    const int tap = k / g.Cout, co = k - tap * g.Cout;
    const int b = m / HW, p = m - b * HW, y = p / g.W - (tap / 3 - 1), xx = p % g.W - (tap % 3 - 1);
    if ((unsigned)y >= (unsigned)g.H || (unsigned)xx >= (unsigned)g.W) return 0;
    return dy[(((size_t)b * g.H + y) * g.W + xx) * g.Cout + co];
  }
};
struct LoadDgB {  // B(k, n) = w[tap*Cin + n][co]
  const uint16_t* w;
  ConvGeom g;
  __device__ uint16_t operator()(int k, int n) const {
    if (k >= 9 * g.Cout) return 0;
    const int tap = k / g.Cout, co = k - tap * g.Cout;
    return w[((size_t)tap * g.Cin + n) * g.Cout + co];
  }
};
// wgrad: A(m, k) = x[pixel k shifted by tap(m)][ci(m)], m = tap*Cin + ci
struct LoadWgA {
  const uint16_t* x;
  ConvGeom g;
  __device__ uint16_t operator()(int m, int k) const {
    const int HW = g.H * g.W;
    const int tap = m / g.Cin, ci = m - tap * g.Cin;
    const int pix = k;
    const int b = pix / HW, p = pix - b * HW, y = p / g.W + tap / 3 - 1, xx = p % g.W + tap % 3 - 1;
    if ((unsigned)y >= (unsigned)g.H || (unsigned)xx >= (unsigned)g.W) return 0;
    return x[(((size_t)b * g.H + y) * g.W + xx) * g.Cin + ci];
  }
};
struct LoadWgB {  // B(k, n) = dy[pixel][n]
  const uint16_t* dy;
  ConvGeom g;
  __device__ uint16_t operator()(int k, int n) const { return dy[(size_t)k * g.Cout + n]; }
};

// dense row-major operands (GEMM probe)
struct LoadDenseA {
  const uint16_t* a;
  int K;
  __device__ uint16_t operator()(int m, int k) const { return a[(size_t)m * K + k]; }
};
struct LoadDenseB {
  const uint16_t* b;
  int N;
  __device__ uint16_t operator()(int k, int n) const { return b[(size_t)k * N + n]; }
};

// ---------------------------------------------------------------------------
// implicit GEMM with MFMA 32x32x16 bf16 -> f32
// ---------------------------------------------------------------------------
// MFMA lane layout (wave64, 32x32x16):
//   A fragment: lane l holds A[l % 32][8*(l/32) .. 8*(l/32)+7]
//   B fragment: lane l holds B[8*(l/32) .. +7][l % 32]
//   D: lane l, register j holds D[8*(j/4) + 4*(l/32) + j%4][l % 32]
// Both operands sit in LDS with K contiguous ([m][k] and [n][k], rows padded
// to 80 bytes) so a fragment is one 16-byte read.
// blockIdx.z selects a K range [z*ksplit, min(K, (z+1)*ksplit)); the epilogue
// receives z so split-K partials land in separate buffers.
template <class LA, class LB, class EPI>
__global__ __launch_bounds__(NTH) void k_igemm(int M, int N, int K, int ksplit, LA la, LB lb,
                                               EPI epi) {
  __shared__ __align__(16) uint16_t sA[GT][KC + 8];    // [m][k]
  __shared__ __align__(16) uint16_t sBt[GT][KC + 8];   // [n][k]
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  const int kbeg = blockIdx.z * ksplit, kend = min(K, kbeg + ksplit);
  for (int k0 = kbeg; k0 < kend; k0 += KC) {
    // stage A (64 x 32) and B (32 x 64): 8 elements per thread each
#pragma unroll
    for (int r = 0; r < (GT * KC) / NTH; ++r) {
      const int e = tid + r * NTH;
      const int mi = e / KC, ki = e % KC;
      const int m = m0 + mi, k = k0 + ki;
      sA[mi][ki] = (m < M && k < kend) ? la(m, k) : (uint16_t)0;
      const int kb = e / GT, ni = e % GT;
      const int kk = k0 + kb, n = n0 + ni;
      sBt[ni][kb] = (kk < kend && n < N) ? lb(kk, n) : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC; ks += 16) {
      const int kq = ks + 8 * (lane / 32);
      const uint4 a = *reinterpret_cast<const uint4*>(&sA[wm + lane % 32][kq]);
      const uint4 b = *reinterpret_cast<const uint4*>(&sBt[wn + lane % 32][kq]);
      acc = mfma16(a, b, acc);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int m = m0 + wm + 8 * (j / 4) + 4 * (lane / 32) + (j % 4);
    const int n = n0 + wn + lane % 32;
    if (m < M && n < N) epi(m, n, acc[j], blockIdx.z);
  }
}

// epilogues
struct EpiConv {  // y = act(acc + bias) -> bf16 (and optionally fp32)
  uint16_t* y;
  float* yf;
  const float* bias;
  int N;
  int relu;
  __device__ void operator()(int m, int n, float v, int) const {
    v += bias ? bias[n] : 0.f;
    if (relu) v = fmaxf(v, 0.f);
    if (y) y[(size_t)m * N + n] = f2bf(v);
    if (yf) yf[(size_t)m * N + n] = v;
  }
};
struct EpiBf16 {  // plain bf16 store
  uint16_t* y;
  int N;
  __device__ void operator()(int m, int n, float v, int) const { y[(size_t)m * N + n] = f2bf(v); }
};
struct EpiF32 {  // fp32 store; split z goes to y + z * stride
  float* y;
  int N;
  size_t stride;
  __device__ void operator()(int m, int n, float v, int z) const {
    y[(size_t)z * stride + (size_t)m * N + n] = v;
  }
};

// ---------------------------------------------------------------------------
// direct 3x3 convolution on halo tiles (forward and dgrad)
// ---------------------------------------------------------------------------
// out[g][x][n] = sum_{tap, c} in[g + s*(ky-1)][x + s*(kx-1)][c] * wt[tap][n][c]
// over "global rows" g = b*H + y (NHWC, rows of W pixels), s = +1 forward,
// s = -1 for dgrad (FLIP).  A work-group owns R = 128 / W consecutive global
// rows (they may span images when R > H: taps whose source row leaves the
// image are zeroed per lane) and BN output channels.
constexpr int CK = 32;       // input channels per LDS chunk
constexpr int CS = CK + 8;   // LDS stride per pixel / per weight row (80 B)

struct ConvOut {
  uint16_t* y;             // bf16 out (N channels) or null
  float* yf;               // fp32 out or null
  const float* bias;       // or null
  const uint16_t* mask;    // ReLU backward: zero where mask <= 0 (same shape as y), or null
  int relu;
};

__device__ __forceinline__ void conv_store(const ConvOut& o, size_t pix, int N, int n, float v) {
  if (o.bias) v += o.bias[n];
  if (o.relu) v = fmaxf(v, 0.f);
  const size_t i = pix * N + n;
  if (o.mask && !(bf2f(o.mask[i]) > 0.f)) v = 0.f;
  if (o.y) o.y[i] = f2bf(v);
  if (o.yf) o.yf[i] = v;
}

size_t conv_lds_bytes(int W, int BN) {
  const int R = 128 / W;
  return (size_t)((R + 2) * (W + 2) + 9 * BN) * CS * sizeof(uint16_t);
}

template <int BN, bool FLIP>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ in,
                                              const uint16_t* __restrict__ wt, int rows, int H,
                                              int W, int C, int N, ConvOut o) {
  extern __shared__ uint4 conv_lds[];
  uint16_t* sX = reinterpret_cast<uint16_t*>(conv_lds);
  const int R = 128 / W, W2 = W + 2, HW2 = (R + 2) * W2;
  uint16_t* sW = sX + HW2 * CS;
  constexpr int WN = BN >= 64 ? 2 : 1;   // waves along N
  constexpr int WM = 4 / WN;             // waves along M
  constexpr int MI = 128 / WM / 32;      // 32-row blocks per wave
  constexpr int NJ = BN / WN / 32;       // 32-column blocks per wave
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g0 = blockIdx.x * R, n0 = blockIdx.y * BN;
  const int wm = (wave / WN) * (128 / WM), wn = (wave % WN) * (BN / WN);
  const int kq = 8 * (lane >> 5);
  int pr[MI], pc[MI], py[MI];
#pragma unroll
  for (int i = 0; i < MI; ++i) {
    const int m = wm + 32 * i + (lane & 31);
    pr[i] = m / W;
    pc[i] = m - pr[i] * W;
    py[i] = (g0 + pr[i]) % H;
  }
  f32x16 acc[MI][NJ];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  const uint4 zero = make_uint4(0, 0, 0, 0);
  // staging: every global load of a chunk is issued before the first LDS
  // store, so the chunk pays one load latency (a loop of load -> store paid
  // it once per element); XL covers the halo tile up to W = 128.  (Issuing
  // the next chunk's loads before this chunk's MFMA loop as well took the
  // kernel to 456 VGPRs + AGPRs and one wave per SIMD.)
  constexpr int XL = 8, WL = (9 * BN * (CK / 8) + 255) / 256;
  const int nx = HW2 * (CK / 8);
  for (int c0 = 0; c0 < C; c0 += CK) {
    uint4 xv[XL], wv[WL];
#pragma unroll
    for (int k = 0; k < XL; ++k) {
      const int e = tid + 256 * k;
      const int q = e & (CK / 8 - 1), p = e / (CK / 8);
      const int hr = p / W2, hc = p - hr * W2;
      const int g = g0 - 1 + hr, xx = hc - 1;
      xv[k] = zero;
      if (e < nx && g >= 0 && g < rows && xx >= 0 && xx < W)
        xv[k] = *reinterpret_cast<const uint4*>(in + ((size_t)g * W + xx) * C + c0 + 8 * q);
    }
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const int e = tid + 256 * k;
      const int q = e & (CK / 8 - 1), r = e / (CK / 8);
      const int tap = r / BN, n = r - tap * BN;
      wv[k] = zero;
      if (e < 9 * BN * (CK / 8) && n0 + n < N)
        wv[k] = *reinterpret_cast<const uint4*>(wt + ((size_t)tap * N + n0 + n) * C + c0 + 8 * q);
    }
#pragma unroll
    for (int k = 0; k < XL; ++k) {
      const int e = tid + 256 * k;
      if (e < nx) {
        const int q = e & (CK / 8 - 1), p = e / (CK / 8);
        *reinterpret_cast<uint4*>(sX + p * CS + 8 * q) = xv[k];
      }
    }
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const int e = tid + 256 * k;
      if (e < 9 * BN * (CK / 8)) {
        const int q = e & (CK / 8 - 1), r = e / (CK / 8);
        *reinterpret_cast<uint4*>(sW + r * CS + 8 * q) = wv[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = FLIP ? 2 - tap / 3 : tap / 3, kx = FLIP ? 2 - tap % 3 : tap % 3;
#pragma unroll
      for (int ks = 0; ks < CK; ks += 16) {
        uint4 b[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          b[j] = *reinterpret_cast<const uint4*>(sW + (tap * BN + wn + 32 * j + (lane & 31)) * CS +
                                                 ks + kq);
#pragma unroll
        for (int i = 0; i < MI; ++i) {
          uint4 a = *reinterpret_cast<const uint4*>(sX + ((pr[i] + ky) * W2 + pc[i] + kx) * CS +
                                                    ks + kq);
          if ((unsigned)(py[i] + ky - 1) >= (unsigned)H) a = zero;
#pragma unroll
          for (int j = 0; j < NJ; ++j) acc[i][j] = mfma16(a, b[j], acc[i][j]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = wm + 32 * i + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + 32 * j + (lane & 31);
        if (g0 + m / W < rows && n < N) conv_store(o, (size_t)g0 * W + m, N, n, acc[i][j][r]);
      }
}

// weight gradient on halo tiles:
//   ws[z][t][c][n] = sum over the pixel tiles of split z of x[px + off(t)][c] * dy[px][n]
// grid (Cin/32, Cout/BN, splits); three waves, wave w owns the taps of kernel
// row ky = w.  Both MFMA operands need 8 consecutive pixels per lane, so x and
// dy are staged channel-major in LDS (packed pixel pairs); the kx = 0 / 2 taps
// come from one aligned 16-byte read plus a 4-byte neighbour via alignbyte.
// x rows in LDS: position 8 + col for col in [0, W); 7 and W + 8 hold the
// zero halo; rows are padded so lanes (channels) fall on distinct banks.
// Tiles of split z: t = z, z + splits, ... (fixed order).  With wsb non-null
// the work-groups of blockIdx.x == 0 also write the bias-gradient partials
// wsb[z][n] = sum of dy over those tiles.
constexpr int WG_THREADS = 192;
constexpr int SD = 128 + 8;   // LDS stride of one dy channel (128 pixels + pad)

__host__ __device__ inline int wgrad_x_stride(int W) { return (128 / W + 2) * (W + 16) + 8; }

size_t wgrad_lds_bytes(int W, int BN) {
  return ((size_t)32 * wgrad_x_stride(W) + (size_t)BN * SD) * sizeof(uint16_t);
}

__device__ __forceinline__ uint32_t pack_lo(uint32_t a, uint32_t b) { return (a & 0xffffu) | (b << 16); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b) { return (a >> 16) | (b & 0xffff0000u); }

template <int BN>
__global__ __launch_bounds__(WG_THREADS) void k_wgrad(const uint16_t* __restrict__ x,
                                                      const uint16_t* __restrict__ dy, int rows,
                                                      int H, int W, int Cin, int Cout, int ntiles,
                                                      float* __restrict__ ws,
                                                      float* __restrict__ wsb) {
  extern __shared__ uint4 wg_lds[];
  const int R = 128 / W, RL = W + 16, CSX = wgrad_x_stride(W);
  uint16_t* sX = reinterpret_cast<uint16_t*>(wg_lds);   // [32][R+2][RL] (+pad per channel)
  uint16_t* sD = sX + 32 * CSX;                         // [BN][SD]
  constexpr int NJ = BN / 32;
  const int tid = threadIdx.x, ky = tid >> 6, lane = tid & 63;
  const int c0 = blockIdx.x * 32, n0 = blockIdx.y * BN, z = blockIdx.z;
  const bool do_bias = wsb && blockIdx.x == 0;
  for (int e = tid; e < 32 * (R + 2); e += WG_THREADS) {   // zero halo columns, set once
    const int c = e / (R + 2), hr = e - c * (R + 2);
    sX[c * CSX + hr * RL + 7] = 0;
    sX[c * CSX + hr * RL + W + 8] = 0;
  }
  f32x16 acc[3][NJ];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[k][j][r] = 0.f;
  float bsum = 0.f;
  const uint4 zero = make_uint4(0, 0, 0, 0);
  for (int t = z; t < ntiles; t += gridDim.z) {
    const int g0 = t * R;
    __syncthreads();   // the previous tile's reads are done
    // x halo rows g0-1 .. g0+R, pixel pairs, 8 channels per vector
    for (int e = tid; e < (R + 2) * (W / 2) * 4; e += WG_THREADS) {
      const int q = e & 3, rest = e >> 2;
      const int hr = rest / (W / 2), xp = rest - hr * (W / 2);
      const int g = g0 - 1 + hr;
      uint4 va = zero, vb = zero;
      if (g >= 0 && g < rows) {
        const uint16_t* src = x + ((size_t)g * W + 2 * xp) * Cin + c0 + 8 * q;
        va = *reinterpret_cast<const uint4*>(src);
        vb = *reinterpret_cast<const uint4*>(src + Cin);
      }
      uint32_t* dst = reinterpret_cast<uint32_t*>(sX + (8 * q) * CSX + hr * RL + 8 + 2 * xp);
      const uint32_t a[4] = {va.x, va.y, va.z, va.w}, b[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        dst[(2 * h) * (CSX / 2)] = pack_lo(a[h], b[h]);
        dst[(2 * h + 1) * (CSX / 2)] = pack_hi(a[h], b[h]);
      }
    }
    // dy tile: 128 pixels (R rows), BN channels
    for (int e = tid; e < 64 * (BN / 8); e += WG_THREADS) {
      const int q = e % (BN / 8), pp = e / (BN / 8);
      const int p = 2 * pp;
      uint4 va = zero, vb = zero;
      if (g0 + p / W < rows) {
        const uint16_t* src = dy + ((size_t)g0 * W + p) * Cout + n0 + 8 * q;
        va = *reinterpret_cast<const uint4*>(src);
        vb = *reinterpret_cast<const uint4*>(src + Cout);
      }
      uint32_t* dst = reinterpret_cast<uint32_t*>(sD + (8 * q) * SD + p);
      const uint32_t a[4] = {va.x, va.y, va.z, va.w}, b[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        dst[(2 * h) * (SD / 2)] = pack_lo(a[h], b[h]);
        dst[(2 * h + 1) * (SD / 2)] = pack_hi(a[h], b[h]);
      }
    }
    __syncthreads();
    if (do_bias && tid < BN) {
      const uint4* row = reinterpret_cast<const uint4*>(sD + tid * SD);
      for (int v = 0; v < 16; ++v) {
        const uint4 u = row[v];
        const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int h = 0; h < 4; ++h) bsum += bf2f((uint16_t)w4[h]) + bf2f((uint16_t)(w4[h] >> 16));
      }
    }
#pragma unroll 2
    for (int s = 0; s < 8; ++s) {
      const int p = 16 * s + 8 * (lane >> 5);
      const int pr = p / W, pc = p - pr * W;
      const bool ok = (unsigned)((g0 + pr) % H + ky - 1) < (unsigned)H;
      uint4 b[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        b[j] = *reinterpret_cast<const uint4*>(sD + (32 * j + (lane & 31)) * SD + p);
      const uint16_t* xr = sX + (lane & 31) * CSX + (pr + ky) * RL + pc;
      uint4 a1 = *reinterpret_cast<const uint4*>(xr + 8);
      const uint32_t l0 = *reinterpret_cast<const uint32_t*>(xr + 6);
      const uint32_t r2 = *reinterpret_cast<const uint32_t*>(xr + 16);
      uint4 a0 = make_uint4(__builtin_amdgcn_alignbyte(a1.x, l0, 2),
                            __builtin_amdgcn_alignbyte(a1.y, a1.x, 2),
                            __builtin_amdgcn_alignbyte(a1.z, a1.y, 2),
                            __builtin_amdgcn_alignbyte(a1.w, a1.z, 2));
      uint4 a2 = make_uint4(__builtin_amdgcn_alignbyte(a1.y, a1.x, 2),
                            __builtin_amdgcn_alignbyte(a1.z, a1.y, 2),
                            __builtin_amdgcn_alignbyte(a1.w, a1.z, 2),
                            __builtin_amdgcn_alignbyte(r2, a1.w, 2));
      if (!ok) a0 = a1 = a2 = zero;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        acc[0][j] = mfma16(a0, b[j], acc[0][j]);
        acc[1][j] = mfma16(a1, b[j], acc[1][j]);
        acc[2][j] = mfma16(a2, b[j], acc[2][j]);
      }
    }
  }
  const size_t cnt = (size_t)9 * Cin * Cout;
#pragma unroll
  for (int kx = 0; kx < 3; ++kx)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ci = 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
        const int co = n0 + 32 * j + (lane & 31);
        const int tap = 3 * ky + kx;
        ws[(size_t)z * cnt + ((size_t)tap * Cin + c0 + ci) * Cout + co] = acc[kx][j][r];
      }
  if (do_bias && tid < BN) wsb[(size_t)z * Cout + n0 + tid] = bsum;
}

// The first layer's input value at element i: a bf16 activation, or (FROM_I)
// the network input formed from a centred float32 intensity,
// bf16(sqrt(max(I, 0)) / scale) -- the square root, scaling and rounding of
// the prediction's input done on the fly instead of in a separate pass.
template <bool FROM_I>
__device__ __forceinline__ float c1_input(const void* __restrict__ x, size_t i, float scale) {
  if constexpr (FROM_I) {
    const float I = static_cast<const float*>(x)[i];
    return bf2f(f2bf(sqrtf(fmaxf(I, 0.f)) / scale));
  } else {
    return bf2f(static_cast<const uint16_t*>(x)[i]);
  }
}

// first layer: one input channel, N <= 64 outputs, one pixel per thread (its
// nine inputs formed once).  N % 8 == 0: the channels are stored in 16-byte
// groups of 8, weights and bias read from LDS as float4.
template <bool FROM_I>
__global__ __launch_bounds__(256) void k_conv_c1(const void* __restrict__ x, float scale,
                                                 const uint16_t* __restrict__ w, int B, int H,
                                                 int W, int N, ConvOut o) {
  __shared__ float4 sw4[9 * 64 / 4];
  __shared__ float4 sb4[64 / 4];
  float* sw = reinterpret_cast<float*>(sw4);
  float* sb = reinterpret_cast<float*>(sb4);
  for (int i = threadIdx.x; i < 9 * N; i += blockDim.x) sw[i] = bf2f(w[i]);
  for (int i = threadIdx.x; i < N; i += blockDim.x) sb[i] = o.bias ? o.bias[i] : 0.f;
  __syncthreads();
  const size_t P = (size_t)B * H * W;
  const bool vec = (N % 8) == 0 && o.y && !o.yf && !o.mask && ((uintptr_t)o.y & 15) == 0;
  for (size_t pix = (size_t)blockIdx.x * blockDim.x + threadIdx.x; pix < P;
       pix += (size_t)gridDim.x * blockDim.x) {
    const int xx = pix % W, y = (pix / W) % H;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int yy = y + t / 3 - 1, x2 = xx + t % 3 - 1;
      v[t] = ((unsigned)yy < (unsigned)H && (unsigned)x2 < (unsigned)W)
                 ? c1_input<FROM_I>(x, pix + (long)(t / 3 - 1) * W + (t % 3 - 1), scale)
                 : 0.f;
    }
    if (vec) {
      for (int grp = 0; grp < N / 8; ++grp) {
        float r[8];
        const float4 b0 = sb4[2 * grp], b1 = sb4[2 * grp + 1];
        r[0] = b0.x; r[1] = b0.y; r[2] = b0.z; r[3] = b0.w;
        r[4] = b1.x; r[5] = b1.y; r[6] = b1.z; r[7] = b1.w;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const float4 w0 = sw4[(t * N + 8 * grp) / 4], w1 = sw4[(t * N + 8 * grp) / 4 + 1];
          r[0] += v[t] * w0.x; r[1] += v[t] * w0.y; r[2] += v[t] * w0.z; r[3] += v[t] * w0.w;
          r[4] += v[t] * w1.x; r[5] += v[t] * w1.y; r[6] += v[t] * w1.z; r[7] += v[t] * w1.w;
        }
        if (o.relu) {
#pragma unroll
          for (int h = 0; h < 8; ++h) r[h] = fmaxf(r[h], 0.f);
        }
        *reinterpret_cast<uint4*>(o.y + pix * N + 8 * grp) = pack8(r);
      }
    } else {
      for (int n = 0; n < N; ++n) {
        float acc = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) acc += v[t] * sw[t * N + n];
        conv_store(o, pix, N, n, acc);
      }
    }
  }
}

// weight gradient of thin layers (9*Cin*Cout + Cout <= 768): grid = splits;
// a work-group walks its pixel tiles t = z, z + splits, ... (128 pixels =
// R = 128/W rows each), stages the x halo and the dy tile in LDS as fp32, and
// every thread accumulates fixed (tap, ci, co) entries -- and, after the 9*Cin*Cout
// weight entries, the Cout bias sums -- in registers.
constexpr int THIN_MAX = 768;
size_t thin_lds_bytes(int W, int Cin, int Cout) {
  const int R = 128 / W;
  return ((size_t)(R + 2) * (W + 2) * Cin + (size_t)128 * Cout + 3 * 16) * sizeof(float);
}

__global__ __launch_bounds__(256) void k_wgrad_thin(const uint16_t* __restrict__ x,
                                                    const uint16_t* __restrict__ dy, int rows,
                                                    int H, int W, int Cin, int Cout, int ntiles,
                                                    float* __restrict__ ws,
                                                    float* __restrict__ wsb) {
  extern __shared__ float thin_lds[];
  const int R = 128 / W, W2 = W + 2;
  float* sX = thin_lds;                          // [(R+2)*(W+2)][Cin]
  float* sD = sX + (R + 2) * W2 * Cin;           // [128][Cout]
  float* sOk = sD + 128 * Cout;                  // [R][3]: source row inside the image
  const int tid = threadIdx.x, z = blockIdx.x;
  const int cnt = 9 * Cin * Cout, tot = cnt + Cout;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int t = z; t < ntiles; t += gridDim.x) {
    const int g0 = t * R;
    __syncthreads();
    for (int e = tid; e < (R + 2) * W2 * Cin; e += 256) {
      const int c = e % Cin, p = e / Cin;
      const int hr = p / W2, hc = p - hr * W2;
      const int g = g0 - 1 + hr, xx = hc - 1;
      sX[e] = (g >= 0 && g < rows && xx >= 0 && xx < W) ? bf2f(x[((size_t)g * W + xx) * Cin + c]) : 0.f;
    }
    for (int e = tid; e < 128 * Cout; e += 256) {
      const int p = e / Cout;
      sD[e] = (g0 + p / W < rows) ? bf2f(dy[(size_t)g0 * W * Cout + e]) : 0.f;
    }
    for (int e = tid; e < 3 * R; e += 256) {
      const int pr = e / 3, ky = e % 3;
      sOk[e] = (unsigned)((g0 + pr) % H + ky - 1) < (unsigned)H ? 1.f : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int idx = tid + 256 * k;
      if (idx < cnt) {
        const int tap = idx / (Cin * Cout), rem = idx - tap * Cin * Cout;
        const int ci = rem / Cout, co = rem - ci * Cout;
        const int ky = tap / 3, kx = tap % 3;
        float s = 0.f;
        for (int p = 0; p < 128; ++p) {
          const int pr = p / W, pc = p - pr * W;
          s += sOk[pr * 3 + ky] * sX[((pr + ky) * W2 + pc + kx) * Cin + ci] * sD[p * Cout + co];
        }
        acc[k] += s;
      } else if (idx < tot) {
        const int co = idx - cnt;
        float s = 0.f;
        for (int p = 0; p < 128; ++p) s += sD[p * Cout + co];
        acc[k] += s;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int idx = tid + 256 * k;
    if (idx < cnt) ws[(size_t)z * cnt + idx] = acc[k];
    else if (idx < tot && wsb) wsb[(size_t)z * Cout + idx - cnt] = acc[k];
  }
}

// few output channels (N <= NN <= 8), C % 8 == 0, C <= 64: one pixel per
// thread; NN bounds the unrolled output loop (2 for the network's last layer:
// with 8 the masked iterations cost as many LDS reads as the live ones)
template <int NN>
__global__ __launch_bounds__(256) void k_conv_nsmall(const uint16_t* __restrict__ x,
                                                     const uint16_t* __restrict__ w, int B, int H,
                                                     int W, int C, int N, ConvOut o) {
  __shared__ float sw[9 * 64 * 8];   // [tap][c][n]
  for (int i = threadIdx.x; i < 9 * C * N; i += blockDim.x) sw[i] = bf2f(w[i]);
  __syncthreads();
  const size_t P = (size_t)B * H * W;
  for (size_t pix = (size_t)blockIdx.x * blockDim.x + threadIdx.x; pix < P;
       pix += (size_t)gridDim.x * blockDim.x) {
    const int xx = pix % W, y = (pix / W) % H;
    float s[NN];
#pragma unroll
    for (int n = 0; n < NN; ++n) s[n] = 0.f;
    for (int t = 0; t < 9; ++t) {
      const int yy = y + t / 3 - 1, x2 = xx + t % 3 - 1;
      if ((unsigned)yy >= (unsigned)H || (unsigned)x2 >= (unsigned)W) continue;
      const uint16_t* src = x + (pix + (long)(t / 3 - 1) * W + (t % 3 - 1)) * C;
      for (int c0 = 0; c0 < C; c0 += 8) {
        const uint4 q = *reinterpret_cast<const uint4*>(src + c0);
        const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int h = 0; h < 8; ++h) {
          const float a = bf2f((uint16_t)(u[h / 2] >> (16 * (h & 1))));
          const float* wr = sw + (t * C + c0 + h) * N;
#pragma unroll
          for (int n = 0; n < NN; ++n)
            if (n < N) s[n] += a * wr[n];
        }
      }
    }
#pragma unroll
    for (int n = 0; n < NN; ++n)
      if (n < N) conv_store(o, pix, N, n, s[n]);
  }
}

// dgrad of a few-output-channel layer: dx[p][c] = sum_t sum_n dy[p - off(t)][n] w[t][c][n],
// C % 8 == 0, C <= 64, N <= 8; optional ReLU mask on dx
__global__ __launch_bounds__(256) void k_dgrad_nsmall(const uint16_t* __restrict__ dy,
                                                      const uint16_t* __restrict__ w, int B, int H,
                                                      int W, int C, int N, uint16_t* dx,
                                                      const uint16_t* mask) {
  __shared__ float sw[9 * 64 * 8];
  for (int i = threadIdx.x; i < 9 * C * N; i += blockDim.x) sw[i] = bf2f(w[i]);
  __syncthreads();
  const size_t P = (size_t)B * H * W;
  for (size_t pix = (size_t)blockIdx.x * blockDim.x + threadIdx.x; pix < P;
       pix += (size_t)gridDim.x * blockDim.x) {
    const int xx = pix % W, y = (pix / W) % H;
    float g[9][8];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int yy = y - (t / 3 - 1), x2 = xx - (t % 3 - 1);
      const bool ok = (unsigned)yy < (unsigned)H && (unsigned)x2 < (unsigned)W;
      const size_t q = pix - (long)(t / 3 - 1) * W - (t % 3 - 1);
#pragma unroll
      for (int n = 0; n < 8; ++n) g[t][n] = (ok && n < N) ? bf2f(dy[q * N + n]) : 0.f;
    }
    for (int c0 = 0; c0 < C; c0 += 8) {
      uint32_t packed[4];
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const float* wr = sw + (t * C + c0 + h) * N;
#pragma unroll
          for (int n = 0; n < 8; ++n)
            if (n < N) s += g[t][n] * wr[n];
        }
        if (mask && !(bf2f(mask[pix * C + c0 + h]) > 0.f)) s = 0.f;
        const uint32_t b = f2bf(s);
        if (h & 1) packed[h / 2] |= b << 16; else packed[h / 2] = b;
      }
      *reinterpret_cast<uint4*>(dx + pix * C + c0) =
          make_uint4(packed[0], packed[1], packed[2], packed[3]);
    }
  }
}

// per-tap transpose of the bf16 weights: wT[t][n][c] = w[t][c][n]
__global__ void k_weight_t(const uint16_t* w, uint16_t* wT, int C, int N) {
  const size_t n9 = (size_t)9 * C * N;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n9;
       i += (size_t)gridDim.x * blockDim.x) {
    const int n = i % N;
    const size_t r = i / N;
    const int c = r % C, t = r / C;
    wT[((size_t)t * N + n) * C + c] = w[i];
  }
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// out[i] = sum over splits of part[s][i] (deterministic).  A block covers 64
// consecutive elements (16 float4 columns) with 16 split groups: group g sums
// splits g, g+16, ... in order, then the 16 group sums are added in group
// order through LDS.  count % 4 == 0 (16-byte loads); else one thread per
// element walks the splits in order.
__global__ __launch_bounds__(256) void k_sum_splits(const float* part, int splits, size_t count,
                                                    float* out) {
  if ((count & 3) == 0) {
    __shared__ float4 red[16][16];
    const int col = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const size_t n4 = count / 4;
    for (size_t base = (size_t)blockIdx.x * 16; base < n4; base += (size_t)gridDim.x * 16) {
      const size_t i = base + col;
      float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
      if (i < n4)
        for (int k = grp; k < splits; k += 16) {
          const float4 v = reinterpret_cast<const float4*>(part + (size_t)k * count)[i];
          s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
      red[grp][col] = s;
      __syncthreads();
      if (grp == 0 && i < n4) {
        float4 t = red[0][col];
        for (int g = 1; g < 16; ++g) {
          const float4 v = red[g][col];
          t.x += v.x; t.y += v.y; t.z += v.z; t.w += v.w;
        }
        reinterpret_cast<float4*>(out)[i] = t;
      }
      __syncthreads();
    }
    return;
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * count + i];
    out[i] = s;
  }
}

// bias gradient: db[c] = sum over pixels of dy[p][c] (fixed order per channel)
__global__ void k_bias_grad(const uint16_t* dy, long P, int C, float* db) {
  const int c = blockIdx.x;
  __shared__ float red[256];
  float s = 0.f;
  for (long p = threadIdx.x; p < P; p += blockDim.x) s += bf2f(dy[(size_t)p * C + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

// dy *= (y > 0)   (ReLU backward through the stored post-activation); 8 per thread
__global__ void k_relu_bwd(uint16_t* dy, const uint16_t* y, size_t n) {
  const size_t n8 = n / 8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8;
       i += (size_t)gridDim.x * blockDim.x) {
    uint4 g = reinterpret_cast<uint4*>(dy)[i];
    const uint4 a = reinterpret_cast<const uint4*>(y)[i];
    uint32_t* gg = reinterpret_cast<uint32_t*>(&g);
    const uint32_t aa[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      if (!(bf2f((uint16_t)aa[h]) > 0.f)) gg[h] &= 0xffff0000u;
      if (!(bf2f((uint16_t)(aa[h] >> 16)) > 0.f)) gg[h] &= 0x0000ffffu;
    }
    reinterpret_cast<uint4*>(dy)[i] = g;
  }
  for (size_t i = n8 * 8 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    if (!(bf2f(y[i]) > 0.f)) dy[i] = 0;
}

// 2x2 average pool, NHWC; C % 8 == 0 (8 channels per thread)
__global__ void k_pool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
  const size_t n = (size_t)B * Ho * Wo * C8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = i % C8;
    size_t r = i / C8;
    const int xo = r % Wo; r /= Wo;
    const int yo = r % Ho;
    const int b = r / Ho;
    const uint4* base =
        reinterpret_cast<const uint4*>(x + (((size_t)b * H + 2 * yo) * W + 2 * xo) * C) + c8;
    float a[8], t[8];
    unpack8(base[0], a);
    unpack8(base[C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] += t[h];
    unpack8(base[(size_t)W * C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] += t[h];
    unpack8(base[(size_t)W * C8 + C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] = 0.25f * (a[h] + t[h]);
    reinterpret_cast<uint4*>(y)[i] = pack8(a);
  }
}
__global__ void k_pool2_bwd(const uint16_t* dy, uint16_t* dx, int B, int H, int W, int C) {
  const int Wo = W / 2, Ho = H / 2, C8 = C / 8;
  const size_t n = (size_t)B * H * W * C8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = i % C8;
    size_t r = i / C8;
    const int xx = r % W; r /= W;
    const int y = r % H;
    const int b = r / H;
    float a[8];
    unpack8(reinterpret_cast<const uint4*>(dy + (((size_t)b * Ho + y / 2) * Wo + xx / 2) * C)[c8], a);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] *= 0.25f;
    reinterpret_cast<uint4*>(dx)[i] = pack8(a);
  }
}
// nearest 2x upsample, NHWC; C % 8 == 0
__global__ void k_up2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
  const int C8 = C / 8;
  const size_t n = (size_t)B * 2 * H * 2 * W * C8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = i % C8;
    size_t r = i / C8;
    const int xx = r % (2 * W); r /= (2 * W);
    const int yy = r % (2 * H);
    const int b = r / (2 * H);
    reinterpret_cast<uint4*>(y)[i] =
        reinterpret_cast<const uint4*>(x + (((size_t)b * H + yy / 2) * W + xx / 2) * C)[c8];
  }
}
__global__ void k_up2_bwd(const uint16_t* dy, uint16_t* dx, int B, int H, int W, int C) {
  const int C8 = C / 8;
  const size_t n = (size_t)B * H * W * C8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = i % C8;
    size_t r = i / C8;
    const int xx = r % W; r /= W;
    const int y = r % H;
    const int b = r / H;
    const size_t W2 = 2 * (size_t)W;
    const uint4* base =
        reinterpret_cast<const uint4*>(dy + (((size_t)b * 2 * H + 2 * y) * W2 + 2 * xx) * C) + c8;
    float a[8], t[8];
    unpack8(base[0], a);
    unpack8(base[C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] += t[h];
    unpack8(base[W2 * C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] += t[h];
    unpack8(base[W2 * C8 + C8], t);
#pragma unroll
    for (int h = 0; h < 8; ++h) a[h] += t[h];
    reinterpret_cast<uint4*>(dx)[i] = pack8(a);
  }
}

// mean-squared error between fp32 prediction and fp32 target; dy = 2 (p - t) / n
// as bf16, partial sums of the loss per block
__global__ void k_mse(const float* p, const float* t, size_t n, uint16_t* dy, float gscale,
                      double* part, const float* wgt, size_t period) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float w = wgt ? wgt[i % period] : 1.f;
    const float d = p[i] - t[i];
    s += (double)w * d * d;
    if (dy) dy[i] = f2bf(gscale * 2.f * w * d / (float)n);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on fp32 master weights; refreshes the bf16 copy
__global__ void k_adam(float* w, uint16_t* wb, const float* g, float* m, float* v, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i] / gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
  }
}

__global__ void k_f32_to_bf16(const float* x, uint16_t* y, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// training batches: gather + exact dihedral augmentation
// ---------------------------------------------------------------------------
// xb[b] = T_b(x[idx[b]]), tb[b] = T_b(t[idx[b]]) for M x M frames, where T_b
// is one of the 8 maps built from i -> (-i) mod M on rows (bit 0), on columns
// (bit 1) and a transpose (bit 2).  For a probe symmetric under these maps,
// reflecting the object patch this way reflects the centred diffraction
// pattern the same way (|F f(-r)| = |F f|(-q)), so (input, target) pairs stay
// exact.  code_b = hash(seed, b) & 7 when aug != 0, else 0.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
  return h;
}

__global__ void k_batch_gather(const uint16_t* __restrict__ x, const float2* __restrict__ t,
                               const int64_t* __restrict__ idx, int B, int M, int aug,
                               uint32_t seed, uint16_t* __restrict__ xb, float2* __restrict__ tb) {
  const size_t MM = (size_t)M * M, n = (size_t)B * MM;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int b = e / MM, p = e - (size_t)b * MM;
    const int i = p / M, j = p - i * M;
    const uint32_t code = aug ? (mix32(seed * 0x9e3779b9u + (uint32_t)b) & 7u) : 0u;
    int si = (code & 1) ? (M - i) % M : i, sj = (code & 2) ? (M - j) % M : j;
    if (code & 4) { const int tmp = si; si = sj; sj = tmp; }
    const size_t src = (size_t)idx[b] * MM + (size_t)si * M + sj;
    xb[e] = x[src];
    tb[e] = t[src];
  }
}

// ---------------------------------------------------------------------------
// global phase synchronisation of predicted patches
// ---------------------------------------------------------------------------
// Each patch is known only up to a constant phase.  For every overlapping
// pair q = (i, j), with window offsets off (object row, col),
//   c[q] = sum over the overlap of conj(w p_i) (w p_j)     (double)
// estimates exp(i (theta_i - theta_j)) |c|; the phases z maximise
// Re sum_ij conj(z_i) C_ij z_j over |z_i| = 1 (projected power iterations).
__global__ __launch_bounds__(256) void k_pair_overlap(const float2* __restrict__ patches,
                                                      const float* __restrict__ w,
                                                      const int2* __restrict__ off,
                                                      const int2* __restrict__ pairs, int npairs,
                                                      int M, int K, double2* __restrict__ c) {
  __shared__ double red[2][4];
  for (int q = blockIdx.x; q < npairs; q += gridDim.x) {
    const int i = pairs[q].x, j = pairs[q].y;
    const int dy = off[j].x - off[i].x, dx = off[j].y - off[i].y;
    const int r0 = max(0, dy), r1 = min(M, M + dy), c0 = max(0, dx), c1 = min(K, K + dx);
    const int h = max(0, r1 - r0), wd = max(0, c1 - c0);
    const float2* pi = patches + (size_t)i * M * K;
    const float2* pj = patches + (size_t)j * M * K;
    double sr = 0.0, si = 0.0;
    for (int e = threadIdx.x; e < h * wd; e += 256) {
      const int r = r0 + e / wd, cc = c0 + e % wd;   // pixel in i's window
      const int rj = r - dy, cj = cc - dx;           // the same object pixel in j's window
      const float ww = w ? w[r * K + cc] * w[rj * K + cj] : 1.f;
      const float2 a = pi[r * K + cc], b = pj[rj * K + cj];
      sr += (double)ww * ((double)a.x * b.x + (double)a.y * b.y);
      si += (double)ww * ((double)a.x * b.y - (double)a.y * b.x);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 64);
      si += __shfl_xor(si, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = sr;
      red[1][threadIdx.x >> 6] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      c[q] = make_double2(red[0][0] + red[0][1] + red[0][2] + red[0][3],
                          red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    __syncthreads();
  }
}

// Expand the pair products into CSR order: cv[e] = C_ij as float2, with
// C_ij = c[qs] for qs >= 0 and conj(c[-qs-1]) otherwise; col[e] = j.
__global__ void k_csr_values(const int2* __restrict__ ent, const double2* __restrict__ c, int nent,
                             float2* __restrict__ cv, int* __restrict__ col) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += gridDim.x * blockDim.x) {
    const int2 en = ent[e];
    const double2 v = c[en.y >= 0 ? en.y : -en.y - 1];
    cv[e] = make_float2((float)v.x, (float)(en.y >= 0 ? v.y : -v.y));
    col[e] = en.x;
  }
}

// One work-group; z starts at 1; every iteration sets z_i = s_i / |s_i|,
// s_i = sum_j C_ij z_j over row i in CSR order (kept when s_i = 0).  z lives
// in LDS (two buffers of N <= 8192 entries); the row sum is unrolled by four
// with the partial sums combined in a fixed order.
constexpr int SYNC_MAX = 8192;
__global__ __launch_bounds__(1024) void k_phase_sync(const int* __restrict__ rowptr,
                                                     const int* __restrict__ col,
                                                     const float2* __restrict__ cv, int N,
                                                     int iters, float2* __restrict__ z) {
  extern __shared__ float2 zs[];   // [2][N]
  float2* za = zs;
  float2* zb = zs + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) za[i] = make_float2(1.f, 0.f);
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    const float2* src = (it & 1) ? zb : za;
    float2* dst = (it & 1) ? za : zb;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const int e0 = rowptr[i], e1 = rowptr[i + 1];
      float sr[4] = {0.f, 0.f, 0.f, 0.f}, si[4] = {0.f, 0.f, 0.f, 0.f};
      int e = e0;
      for (; e + 4 <= e1; e += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float2 v = cv[e + u];
          const float2 zj = src[col[e + u]];
          sr[u] += v.x * zj.x - v.y * zj.y;
          si[u] += v.x * zj.y + v.y * zj.x;
        }
      }
      for (; e < e1; ++e) {
        const float2 v = cv[e];
        const float2 zj = src[col[e]];
        sr[0] += v.x * zj.x - v.y * zj.y;
        si[0] += v.x * zj.y + v.y * zj.x;
      }
      const float tr = (sr[0] + sr[1]) + (sr[2] + sr[3]), ti = (si[0] + si[1]) + (si[2] + si[3]);
      const float m = sqrtf(tr * tr + ti * ti);
      dst[i] = m > 0.f ? make_float2(tr / m, ti / m) : src[i];
    }
    __syncthreads();
  }
  const float2* fin = (iters & 1) ? zb : za;
  for (int i = threadIdx.x; i < N; i += blockDim.x) z[i] = fin[i];
}

// exits[n] = P * patch_n * z_n  (so the overlap projection of the exits is the
// |P|^2-weighted average of the phase-aligned patches)
__global__ void k_phase_exits(const float2* __restrict__ patches, const float2* __restrict__ z,
                              const float2* __restrict__ P, int N, int MK, float2* __restrict__ out) {
  const size_t n = (size_t)N * MK;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int f = e / MK, p = e - (size_t)f * MK;
    const float2 a = patches[e], zf = z[f], pp = P[p];
    const float2 az = make_float2(a.x * zf.x - a.y * zf.y, a.x * zf.y + a.y * zf.x);
    out[e] = make_float2(pp.x * az.x - pp.y * az.y, pp.x * az.y + pp.y * az.x);
  }
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

int sum_splits(const float* ws, int splits, size_t cnt, float* out, hipStream_t s) {
  const int blocks = (cnt & 3) == 0 ? (int)std::min<size_t>((cnt / 4 + 15) / 16, 8192) : grid_for(cnt);
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks), dim3(256), 0, s, ws, splits, cnt, out);
  NN_LAUNCHED("k_sum_splits");
  return PTY_OK;
}

template <class LA, class LB, class EPI>
int igemm(int M, int N, int K, LA la, LB lb, EPI epi, hipStream_t s, int splits = 1) {
  const int ksplit = ((K + splits - 1) / splits + KC - 1) / KC * KC;
  dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT, splits);
  hipLaunchKernelGGL((k_igemm<LA, LB, EPI>), grid, dim3(NTH), 0, s, M, N, K, ksplit, la, lb, epi);
  NN_LAUNCHED("k_igemm");
  return PTY_OK;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

bool thin_ok(int W, int Cin, int Cout) {
  return W >= 8 && W <= 128 && 128 % W == 0 && 9 * Cin * Cout + Cout <= THIN_MAX &&
         thin_lds_bytes(W, Cin, Cout) <= 64 * 1024;
}

// k_conv takes whole-row tiles of 128 pixels and 32-channel chunks
bool direct_ok(int W, int C, int N) {
  return W >= 8 && W <= 128 && 128 % W == 0 && C % CK == 0 && N % 32 == 0;
}

template <bool FLIP>
int launch_conv(const uint16_t* in, const uint16_t* wt, int rows, int H, int W, int C, int N,
                ConvOut o, hipStream_t s) {
  const int R = 128 / W;
  const dim3 grid((rows + R - 1) / R, 1, 1);
  if (N % 64 == 0) {
    const size_t lds = conv_lds_bytes(W, 64);
    static bool attr = false;
    if (!attr) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv<64, FLIP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024) != hipSuccess)
        return nn_fail(PTY_EHIP, "k_conv: cannot raise the LDS limit");
      attr = true;
    }
    hipLaunchKernelGGL((k_conv<64, FLIP>), dim3(grid.x, N / 64), dim3(256), lds, s, in, wt, rows,
                       H, W, C, N, o);
  } else {
    hipLaunchKernelGGL((k_conv<32, FLIP>), dim3(grid.x, N / 32), dim3(256), conv_lds_bytes(W, 32),
                       s, in, wt, rows, H, W, C, N, o);
  }
  NN_LAUNCHED("k_conv");
  return PTY_OK;
}

}  // namespace

extern "C" {

const char* pty_nn_last_error(void) { return g_nn_err.c_str(); }

int pty_nn_gemm_probe(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
                      void* stream) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) return nn_fail(PTY_EINVAL, "bad gemm args");
  return igemm(M, N, K, LoadDenseA{A, K}, LoadDenseB{B, N}, EpiF32{C, N, 0}, (hipStream_t)stream);
}

int pty_nn_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const uint16_t* wT,
                       const float* bias, uint16_t* y, float* yf, int B, int H, int W, int Cin,
                       int Cout, int relu, void* stream) {
  if (!x || !w || (!y && !yf) || B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0)
    return nn_fail(PTY_EINVAL, "bad conv args");
  hipStream_t s = (hipStream_t)stream;
  const ConvOut o{y, yf, bias, nullptr, relu};
  const size_t P = (size_t)B * H * W;
  if (Cin == 1 && Cout <= 64) {
    hipLaunchKernelGGL(k_conv_c1<false>, dim3(grid_for(P)), dim3(256), 0, s, (const void*)x, 1.f,
                       w, B, H, W, Cout, o);
    NN_LAUNCHED("k_conv_c1");
    return PTY_OK;
  }
  // the vector kernels read and write 16-byte groups of channels
  const bool al = aligned16(x) && aligned16(wT) && (!y || aligned16(y));
  if (Cout <= 8 && Cin % 8 == 0 && Cin <= 64 && aligned16(x)) {
    if (Cout <= 2)
      hipLaunchKernelGGL(k_conv_nsmall<2>, dim3(grid_for(P)), dim3(256), 0, s, x, w, B, H, W, Cin,
                         Cout, o);
    else
      hipLaunchKernelGGL(k_conv_nsmall<8>, dim3(grid_for(P)), dim3(256), 0, s, x, w, B, H, W, Cin,
                         Cout, o);
    NN_LAUNCHED("k_conv_nsmall");
    return PTY_OK;
  }
  if (wT && al && direct_ok(W, Cin, Cout)) return launch_conv<false>(x, wT, B * H, H, W, Cin, Cout, o, s);
  ConvGeom g{B, H, W, Cin, Cout};
  return igemm(B * H * W, Cout, 9 * Cin, LoadFwdA{x, g}, LoadFwdB{w, g},
               EpiConv{y, yf, bias, Cout, relu}, s);
}

int pty_nn_conv3x3_fwd_intensity(const float* I, float scale, const uint16_t* w, const float* bias,
                                 uint16_t* y, float* yf, int B, int H, int W, int Cout, int relu,
                                 void* stream) {
  if (!I || !w || (!y && !yf) || B <= 0 || H <= 0 || W <= 0 || Cout <= 0 || Cout > 64 ||
      !(scale > 0.f))
    return nn_fail(PTY_EINVAL, "bad first-layer conv args (Cout <= 64, scale > 0)");
  const ConvOut o{y, yf, bias, nullptr, relu};
  const size_t P = (size_t)B * H * W;
  hipLaunchKernelGGL(k_conv_c1<true>, dim3(grid_for(P)), dim3(256), 0, (hipStream_t)stream,
                     (const void*)I, scale, w, B, H, W, Cout, o);
  NN_LAUNCHED("k_conv_c1");
  return PTY_OK;
}

int pty_nn_conv3x3_dgrad(const uint16_t* dy, const uint16_t* w, uint16_t* dx,
                         const uint16_t* mask, int B, int H, int W, int Cin, int Cout,
                         void* stream) {
  if (!dy || !w || !dx || B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0)
    return nn_fail(PTY_EINVAL, "bad dgrad args");
  hipStream_t s = (hipStream_t)stream;
  const size_t P = (size_t)B * H * W;
  if (Cout <= 8 && Cin % 8 == 0 && Cin <= 64 && aligned16(dx)) {
    hipLaunchKernelGGL(k_dgrad_nsmall, dim3(grid_for(P)), dim3(256), 0, s, dy, w, B, H, W, Cin,
                       Cout, dx, mask);
    NN_LAUNCHED("k_dgrad_nsmall");
    return PTY_OK;
  }
  if (direct_ok(W, Cout, Cin) && aligned16(dy) && aligned16(w) && aligned16(dx))
    return launch_conv<true>(dy, w, B * H, H, W, Cout, Cin, ConvOut{dx, nullptr, nullptr, mask, 0},
                             s);
  ConvGeom g{B, H, W, Cin, Cout};
  int r = igemm(B * H * W, Cin, 9 * Cout, LoadDgA{dy, g}, LoadDgB{w, g}, EpiBf16{dx, Cin}, s);
  if (r != PTY_OK || !mask) return r;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(P * Cin)), dim3(256), 0, s, dx, mask, P * Cin);
  NN_LAUNCHED("k_relu_bwd");
  return PTY_OK;
}

int pty_nn_weight_t(const uint16_t* w, uint16_t* wT, int Cin, int Cout, void* stream) {
  if (!w || !wT || Cin <= 0 || Cout <= 0) return nn_fail(PTY_EINVAL, "bad weight_t args");
  const size_t n = (size_t)9 * Cin * Cout;
  hipLaunchKernelGGL(k_weight_t, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, w, wT, Cin,
                     Cout);
  NN_LAUNCHED("k_weight_t");
  return PTY_OK;
}

// Split count for pty_nn_conv3x3_wgrad: enough work-groups to fill the chip,
// at most one pixel tile (direct kernel) or one K step (generic GEMM) per
// split and at most 16 M partial floats.
int pty_nn_wgrad_splits(int B, int H, int W, int Cin, int Cout) {
  if (B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0) return 1;
  const long P = (long)B * H * W;
  const long cnt = 9L * Cin * Cout;
  // partial buffers: at most 16 M floats and at most a quarter of the elements of x + dy
  const long cap = std::max(1L, std::min((16L << 20) / cnt, P * (Cin + Cout) / (4 * cnt)));
  if (direct_ok(W, Cin, Cout) || thin_ok(W, Cin, Cout)) {
    const int R = 128 / W;
    const long tiles = ((long)B * H + R - 1) / R;
    const long base = direct_ok(W, Cin, Cout)
                          ? (long)(Cin / 32) * (Cout / (Cout % 64 == 0 ? 64 : 32))
                          : 1L;
    const long want = (768 + base - 1) / base;
    return (int)std::max(1L, std::min({want, tiles, cap}));
  }
  const long want = std::max(1L, std::min(32L, P / 8192));
  return (int)std::max(1L, std::min({want, cap, (P + KC - 1) / KC}));
}

// dw = sum over `splits` partial buffers (fixed order) -> deterministic.
// ws holds splits * (9*Cin*Cout + Cout) floats.  db (optional) = per-channel sum of dy.
int pty_nn_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, float* ws,
                         int splits, int B, int H, int W, int Cin, int Cout, void* stream) {
  if (!x || !dy || !dw || !ws || splits <= 0 || B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0)
    return nn_fail(PTY_EINVAL, "bad wgrad args");
  hipStream_t s = (hipStream_t)stream;
  ConvGeom g{B, H, W, Cin, Cout};
  const long P = (long)B * H * W;
  const size_t cnt = (size_t)9 * Cin * Cout;
  float* wsb = ws + (size_t)splits * cnt;
  if (direct_ok(W, Cin, Cout) && aligned16(x) && aligned16(dy)) {
    const int R = 128 / W;
    const int ntiles = (B * H + R - 1) / R;
    if (Cout % 64 == 0)
      hipLaunchKernelGGL(k_wgrad<64>, dim3(Cin / 32, Cout / 64, splits), dim3(WG_THREADS),
                         wgrad_lds_bytes(W, 64), s, x, dy, B * H, H, W, Cin, Cout, ntiles, ws,
                         db ? wsb : nullptr);
    else
      hipLaunchKernelGGL(k_wgrad<32>, dim3(Cin / 32, Cout / 32, splits), dim3(WG_THREADS),
                         wgrad_lds_bytes(W, 32), s, x, dy, B * H, H, W, Cin, Cout, ntiles, ws,
                         db ? wsb : nullptr);
    NN_LAUNCHED("k_wgrad");
    PTY_NN_TRY(sum_splits(ws, splits, cnt, dw, s));
    if (db) PTY_NN_TRY(sum_splits(wsb, splits, (size_t)Cout, db, s));
    return PTY_OK;
  }
  if (thin_ok(W, Cin, Cout)) {
    const int R = 128 / W;
    const int ntiles = (B * H + R - 1) / R;
    hipLaunchKernelGGL(k_wgrad_thin, dim3(splits), dim3(256), thin_lds_bytes(W, Cin, Cout), s, x,
                       dy, B * H, H, W, Cin, Cout, ntiles, ws, db ? wsb : nullptr);
    NN_LAUNCHED("k_wgrad_thin");
    PTY_NN_TRY(sum_splits(ws, splits, cnt, dw, s));
    if (db) PTY_NN_TRY(sum_splits(wsb, splits, (size_t)Cout, db, s));
    return PTY_OK;
  }
  // every split gets a non-empty pixel range (ksplit rounds up to the K step)
  const long ksplit = ((P + splits - 1) / splits + KC - 1) / KC * KC;
  if ((long)(splits - 1) * ksplit >= P) return nn_fail(PTY_EINVAL, "too many wgrad splits for the pixels");
  {
    int r = igemm(9 * Cin, Cout, (int)P, LoadWgA{x, g}, LoadWgB{dy, g}, EpiF32{ws, Cout, cnt}, s,
                  splits);
    if (r != PTY_OK) return r;
  }
  PTY_NN_TRY(sum_splits(ws, splits, cnt, dw, s));
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3(Cout), dim3(256), 0, s, dy, P, Cout, db);
    NN_LAUNCHED("k_bias_grad");
  }
  return PTY_OK;
}

int pty_nn_relu_bwd(uint16_t* dy, const uint16_t* y, size_t n, void* stream) {
  if (!dy || !y || !aligned16(dy) || !aligned16(y))
    return nn_fail(PTY_EINVAL, "relu_bwd needs 16-byte aligned pointers");
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(n / 8 + 1)), dim3(256), 0, (hipStream_t)stream, dy,
                     y, n);
  NN_LAUNCHED("k_relu_bwd");
  return PTY_OK;
}

int pty_nn_pool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, int backward,
                 void* stream) {
  if (H % 2 || W % 2) return nn_fail(PTY_EINVAL, "pool2 needs even sides");
  if (C % 8 || !aligned16(x) || !aligned16(y))
    return nn_fail(PTY_EINVAL, "pool2 needs C % 8 == 0 and 16-byte aligned pointers");
  const size_t n = (backward ? (size_t)B * H * W : (size_t)B * (H / 2) * (W / 2)) * (C / 8);
  if (backward)
    hipLaunchKernelGGL(k_pool2_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B,
                       H, W, C);
  else
    hipLaunchKernelGGL(k_pool2, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B, H,
                       W, C);
  NN_LAUNCHED("k_pool2");
  return PTY_OK;
}

// forward: x (B,H,W,C) -> y (B,2H,2W,C); backward: dy (B,2H,2W,C) -> dx (B,H,W,C)
int pty_nn_up2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, int backward,
               void* stream) {
  if (C % 8 || !aligned16(x) || !aligned16(y))
    return nn_fail(PTY_EINVAL, "up2 needs C % 8 == 0 and 16-byte aligned pointers");
  const size_t n = (backward ? (size_t)B * H * W : (size_t)B * 4 * H * W) * (C / 8);
  if (backward)
    hipLaunchKernelGGL(k_up2_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B,
                       H, W, C);
  else
    hipLaunchKernelGGL(k_up2, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B, H,
                       W, C);
  NN_LAUNCHED("k_up2");
  return PTY_OK;
}

// loss = mean (p - t)^2 (sum of the per-block partials / n, done by the caller);
// dy (optional) = gscale * 2 (p - t) / n as bf16.  part: 1024 doubles.
int pty_nn_mse(const float* p, const float* t, size_t n, uint16_t* dy, float gscale, double* part,
               const float* wgt, size_t period, void* stream) {
  if (wgt && period == 0) return nn_fail(PTY_EINVAL, "weight period is 0");
  hipLaunchKernelGGL(k_mse, dim3(1024), dim3(256), 0, (hipStream_t)stream, p, t, n, dy, gscale, part,
                     wgt, period);
  NN_LAUNCHED("k_mse");
  return PTY_OK;
}

int pty_nn_adam(float* w, uint16_t* wb, const float* g, float* m, float* v, size_t n, float lr,
                float b1, float b2, float eps, int step, float gscale, void* stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, w, wb, g, m, v,
                     n, lr, b1, b2, eps, c1, c2, gscale);
  NN_LAUNCHED("k_adam");
  return PTY_OK;
}

int pty_nn_to_bf16(const float* x, uint16_t* y, size_t n, void* stream) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, n);
  NN_LAUNCHED("k_f32_to_bf16");
  return PTY_OK;
}

int pty_nn_batch_gather(const uint16_t* x, const float* t, const int64_t* idx, int B, int M,
                        int aug, uint32_t seed, uint16_t* xb, float* tb, void* stream) {
  if (!x || !t || !idx || !xb || !tb || B <= 0 || M <= 0)
    return nn_fail(PTY_EINVAL, "bad batch_gather args");
  const size_t n = (size_t)B * M * M;
  hipLaunchKernelGGL(k_batch_gather, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x,
                     reinterpret_cast<const float2*>(t), idx, B, M, aug, seed, xb,
                     reinterpret_cast<float2*>(tb));
  NN_LAUNCHED("k_batch_gather");
  return PTY_OK;
}

int pty_nn_pair_overlap(const pty_cfloat* patches, const float* w, const int32_t* off,
                        const int32_t* pairs, int npairs, int M, int K, double* c, void* stream) {
  if (!patches || !off || !pairs || !c || npairs < 0 || M <= 0 || K <= 0)
    return nn_fail(PTY_EINVAL, "bad pair_overlap args");
  if (npairs == 0) return PTY_OK;
  hipLaunchKernelGGL(k_pair_overlap, dim3(std::min(npairs, 65536)), dim3(256), 0,
                     (hipStream_t)stream, reinterpret_cast<const float2*>(patches), w,
                     reinterpret_cast<const int2*>(off), reinterpret_cast<const int2*>(pairs), npairs,
                     M, K, reinterpret_cast<double2*>(c));
  NN_LAUNCHED("k_pair_overlap");
  return PTY_OK;
}

int pty_nn_phase_sync(const int32_t* rowptr, const int32_t* entries, const double* c, int N,
                      int iters, pty_cfloat* z, pty_cfloat* work, int32_t nent, void* stream) {
  if (!rowptr || !entries || !c || !z || !work || N <= 0 || iters < 0 || nent < 0)
    return nn_fail(PTY_EINVAL, "bad phase_sync args");
  if (N > SYNC_MAX) return nn_fail(PTY_EINVAL, "phase_sync: more than 8192 frames");
  hipStream_t s = (hipStream_t)stream;
  float2* cv = reinterpret_cast<float2*>(work);
  int* col = reinterpret_cast<int*>(cv + nent);
  if (nent > 0) {
    hipLaunchKernelGGL(k_csr_values, dim3(grid_for(nent)), dim3(256), 0, s,
                       reinterpret_cast<const int2*>(entries), reinterpret_cast<const double2*>(c),
                       nent, cv, col);
    NN_LAUNCHED("k_csr_values");
  }
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_phase_sync),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            2 * SYNC_MAX * (int)sizeof(float2)) != hipSuccess)
      return nn_fail(PTY_EHIP, "k_phase_sync: cannot raise the LDS limit");
    attr = true;
  }
  hipLaunchKernelGGL(k_phase_sync, dim3(1), dim3(1024), 2 * N * sizeof(float2), s, rowptr, col, cv,
                     N, iters, reinterpret_cast<float2*>(z));
  NN_LAUNCHED("k_phase_sync");
  return PTY_OK;
}

int pty_nn_phase_exits(const pty_cfloat* patches, const pty_cfloat* z, const pty_cfloat* P, int N,
                       int M, int K, pty_cfloat* out, void* stream) {
  if (!patches || !z || !P || !out || N <= 0 || M <= 0 || K <= 0)
    return nn_fail(PTY_EINVAL, "bad phase_exits args");
  const size_t n = (size_t)N * M * K;
  hipLaunchKernelGGL(k_phase_exits, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float2*>(patches), reinterpret_cast<const float2*>(z),
                     reinterpret_cast<const float2*>(P), N, M * K, reinterpret_cast<float2*>(out));
  NN_LAUNCHED("k_phase_exits");
  return PTY_OK;
}

}  // extern "C"
