// BF16 MFMA kernels of the learned initialiser (a small 3x3-conv
// encoder-decoder that maps a frame's sqrt-intensity to a complex object patch).
//
// Activations are NHWC bf16, weights [9*Cin][Cout] bf16 (tap-major, then input
// channel), accumulation FP32.  Every convolution is one implicit GEMM
//     C[m][n] = sum_k A(m,k) B(k,n)
// computed by k_igemm with v_mfma_f32_32x32x8_bf16 (64x64 work-group tile, four
// waves of 32x32, K staged through LDS in steps of 32):
//   forward   m = output pixel, k = (tap, ci), n = co
//   dgrad     m = input pixel,  k = (tap, co), n = ci   (flipped taps)
//   wgrad     m = (tap, ci),    k = pixel,     n = co   (split over pixel
//             ranges into fixed partial buffers, summed in order: deterministic)
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

#define NN_LAUNCHED(name)                                                               \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return nn_fail(PTY_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

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
// implicit GEMM with MFMA 32x32x8 bf16 -> f32
// ---------------------------------------------------------------------------
// MFMA lane layout (wave64, 32x32x8):
//   A fragment: lane l holds A[l % 32][4*(l/32) .. 4*(l/32)+3]
//   B fragment: lane l holds B[4*(l/32) .. +3][l % 32]
//   D: lane l, register j holds D[8*(j/4) + 4*(l/32) + j%4][l % 32]
// blockIdx.z selects a K range [z*ksplit, min(K, (z+1)*ksplit)); the epilogue
// receives z so split-K partials land in separate buffers.
template <class LA, class LB, class EPI>
__global__ __launch_bounds__(NTH) void k_igemm(int M, int N, int K, int ksplit, LA la, LB lb,
                                               EPI epi) {
  __shared__ uint16_t sA[GT][KC + 4];   // [m][k]
  __shared__ uint16_t sB[KC][GT + 4];   // [k][n]
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
      sB[kb][ni] = (kk < kend && n < N) ? lb(kk, n) : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC; ks += 8) {
      bf16x4 a, b;
      const int kq = ks + 4 * (lane / 32);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = (short)sA[wm + lane % 32][kq + q];
        b[q] = (short)sB[kq + q][wn + lane % 32];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
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
// small kernels
// ---------------------------------------------------------------------------
// out[i] = sum over splits of part[s][i], in split order
__global__ void k_sum_splits(const float* part, int splits, size_t count, float* out) {
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

// dy *= (y > 0)   (ReLU backward through the stored post-activation)
__global__ void k_relu_bwd(uint16_t* dy, const uint16_t* y, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    if (!(bf2f(y[i]) > 0.f)) dy[i] = 0;
}

// 2x2 average pool, NHWC
__global__ void k_pool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t n = (size_t)B * Ho * Wo * C;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t r = i / C;
    const int xo = r % Wo; r /= Wo;
    const int yo = r % Ho;
    const int b = r / Ho;
    const size_t base = (((size_t)b * H + 2 * yo) * W + 2 * xo) * C + c;
    const float s = bf2f(x[base]) + bf2f(x[base + C]) + bf2f(x[base + (size_t)W * C]) +
                    bf2f(x[base + (size_t)W * C + C]);
    y[i] = f2bf(0.25f * s);
  }
}
__global__ void k_pool2_bwd(const uint16_t* dy, uint16_t* dx, int B, int H, int W, int C) {
  const size_t n = (size_t)B * H * W * C;
  const int Wo = W / 2, Ho = H / 2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t r = i / C;
    const int xx = r % W; r /= W;
    const int y = r % H;
    const int b = r / H;
    dx[i] = f2bf(0.25f * bf2f(dy[(((size_t)b * Ho + y / 2) * Wo + xx / 2) * C + c]));
  }
}
// nearest 2x upsample, NHWC
__global__ void k_up2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
  const size_t n = (size_t)B * 2 * H * 2 * W * C;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t r = i / C;
    const int xx = r % (2 * W); r /= (2 * W);
    const int yy = r % (2 * H);
    const int b = r / (2 * H);
    y[i] = x[(((size_t)b * H + yy / 2) * W + xx / 2) * C + c];
  }
}
__global__ void k_up2_bwd(const uint16_t* dy, uint16_t* dx, int B, int H, int W, int C) {
  const size_t n = (size_t)B * H * W * C;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t r = i / C;
    const int xx = r % W; r /= W;
    const int y = r % H;
    const int b = r / H;
    const size_t W2 = 2 * (size_t)W;
    const size_t base = (((size_t)b * 2 * H + 2 * y) * W2 + 2 * xx) * C + c;
    const float s = bf2f(dy[base]) + bf2f(dy[base + C]) + bf2f(dy[base + W2 * C]) +
                    bf2f(dy[base + W2 * C + C]);
    dx[i] = f2bf(s);
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

// Stitch complex patches into an object by overlap averaging, rotating each
// patch by the global phase that best matches what is already stitched
// (frames in ascending order).  One work-group walks the frames in order, so
// the result is deterministic.  acc = sum of rotated patches, wsum = count.
__global__ __launch_bounds__(1024) void k_stitch(const float2* patches, const float* pw,
                                                 const int2* off, int N, int M, int K, int V,
                                                 float2* acc, float* wsum) {
  __shared__ double red[2 * 16];
  const int tid = threadIdx.x;
  for (int n = 0; n < N; ++n) {
    const int2 o = off[n];
    const float2* pn = patches + (size_t)n * M * K;
    double cr = 0.0, ci = 0.0;
    for (int p = tid; p < M * K; p += blockDim.x) {
      const float2 a = acc[(size_t)(o.x + p / K) * V + o.y + p % K];
      const float w = pw ? pw[p] : 1.f;
      const float2 b = make_float2(w * pn[p].x, w * pn[p].y);
      cr += (double)a.x * b.x + (double)a.y * b.y;   // a * conj(b)
      ci += (double)a.y * b.x - (double)a.x * b.y;
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      cr += __shfl_xor(cr, s, 64);
      ci += __shfl_xor(ci, s, 64);
    }
    if ((tid & 63) == 0) {
      red[2 * (tid / 64)] = cr;
      red[2 * (tid / 64) + 1] = ci;
    }
    __syncthreads();
    double tr = 0.0, ti = 0.0;
    for (int w = 0; w < (int)blockDim.x / 64; ++w) {
      tr += red[2 * w];
      ti += red[2 * w + 1];
    }
    const double mag = sqrt(tr * tr + ti * ti);
    const float rc = mag > 0.0 ? (float)(tr / mag) : 1.f, rs = mag > 0.0 ? (float)(ti / mag) : 0.f;
    for (int p = tid; p < M * K; p += blockDim.x) {
      const size_t q = (size_t)(o.x + p / K) * V + o.y + p % K;
      const float w = pw ? pw[p] : 1.f;
      const float2 b = make_float2(w * pn[p].x, w * pn[p].y);
      float2 a = acc[q];
      a.x += b.x * rc - b.y * rs;
      a.y += b.x * rs + b.y * rc;
      acc[q] = a;
      wsum[q] += w;
    }
    __syncthreads();
  }
}

__global__ void k_stitch_final(const float2* acc, const float* wsum, size_t n, float2* O,
                               float floor, float fill) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float w = wsum[i];
    O[i] = w > floor ? make_float2(acc[i].x / w, acc[i].y / w) : make_float2(fill, 0.f);
  }
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

template <class LA, class LB, class EPI>
int igemm(int M, int N, int K, LA la, LB lb, EPI epi, hipStream_t s, int splits = 1) {
  const int ksplit = ((K + splits - 1) / splits + KC - 1) / KC * KC;
  dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT, splits);
  hipLaunchKernelGGL((k_igemm<LA, LB, EPI>), grid, dim3(NTH), 0, s, M, N, K, ksplit, la, lb, epi);
  NN_LAUNCHED("k_igemm");
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

int pty_nn_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y,
                       float* yf, int B, int H, int W, int Cin, int Cout, int relu, void* stream) {
  if (!x || !w || (!y && !yf) || B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0)
    return nn_fail(PTY_EINVAL, "bad conv args");
  ConvGeom g{B, H, W, Cin, Cout};
  return igemm(B * H * W, Cout, 9 * Cin, LoadFwdA{x, g}, LoadFwdB{w, g},
               EpiConv{y, yf, bias, Cout, relu}, (hipStream_t)stream);
}

int pty_nn_conv3x3_dgrad(const uint16_t* dy, const uint16_t* w, uint16_t* dx, int B, int H, int W,
                         int Cin, int Cout, void* stream) {
  if (!dy || !w || !dx) return nn_fail(PTY_EINVAL, "bad dgrad args");
  ConvGeom g{B, H, W, Cin, Cout};
  return igemm(B * H * W, Cin, 9 * Cout, LoadDgA{dy, g}, LoadDgB{w, g}, EpiBf16{dx, Cin},
               (hipStream_t)stream);
}

// dw = sum over `splits` pixel ranges (fixed order) of x^T dy; ws holds
// splits * 9*Cin*Cout floats.  db (optional) = per-channel sum of dy.
int pty_nn_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, float* ws,
                         int splits, int B, int H, int W, int Cin, int Cout, void* stream) {
  if (!x || !dy || !dw || !ws || splits <= 0) return nn_fail(PTY_EINVAL, "bad wgrad args");
  hipStream_t s = (hipStream_t)stream;
  ConvGeom g{B, H, W, Cin, Cout};
  const long P = (long)B * H * W;
  const size_t cnt = (size_t)9 * Cin * Cout;
  // every split gets a non-empty pixel range (ksplit rounds up to the K step)
  const long ksplit = ((P + splits - 1) / splits + KC - 1) / KC * KC;
  if ((long)(splits - 1) * ksplit >= P) return nn_fail(PTY_EINVAL, "too many wgrad splits for the pixels");
  {
    int r = igemm(9 * Cin, Cout, (int)P, LoadWgA{x, g}, LoadWgB{dy, g}, EpiF32{ws, Cout, cnt}, s,
                  splits);
    if (r != PTY_OK) return r;
  }
  hipLaunchKernelGGL(k_sum_splits, dim3(grid_for(cnt)), dim3(256), 0, s, (const float*)ws, splits,
                     cnt, dw);
  NN_LAUNCHED("k_sum_splits");
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3(Cout), dim3(256), 0, s, dy, P, Cout, db);
    NN_LAUNCHED("k_bias_grad");
  }
  return PTY_OK;
}

int pty_nn_relu_bwd(uint16_t* dy, const uint16_t* y, size_t n, void* stream) {
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, dy, y, n);
  NN_LAUNCHED("k_relu_bwd");
  return PTY_OK;
}

int pty_nn_pool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, int backward,
                 void* stream) {
  if (H % 2 || W % 2) return nn_fail(PTY_EINVAL, "pool2 needs even sides");
  const size_t n = backward ? (size_t)B * H * W * C : (size_t)B * (H / 2) * (W / 2) * C;
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
  const size_t n = backward ? (size_t)B * H * W * C : (size_t)B * 4 * H * W * C;
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

// O (U x V) = overlap average of N complex patches (M x K) placed at the
// frames' windows (R as for pty_create: R <= 0), each patch rotated by the
// global phase that best matches the patches stitched before it.  acc (U*V
// complex) and wsum (U*V float) are caller workspace, cleared here.
int pty_nn_stitch(const pty_cfloat* patches, const float* pw, const int32_t* R_host, int N, int M,
                  int K, int U, int V, float floor, float fill, pty_cfloat* acc, float* wsum,
                  pty_cfloat* O, void* stream) {
  if (!patches || !R_host || !acc || !wsum || !O) return nn_fail(PTY_EINVAL, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  int2* off = nullptr;
  if (hipMalloc(&off, (size_t)N * sizeof(int2)) != hipSuccess) return nn_fail(PTY_ENOMEM, "offsets");
  int2* h = (int2*)malloc((size_t)N * sizeof(int2));
  for (int n = 0; n < N; ++n) {
    h[n] = make_int2(-R_host[2 * n], -R_host[2 * n + 1]);
    if (h[n].x < 0 || h[n].y < 0 || h[n].x + M > U || h[n].y + K > V) {
      free(h);
      (void)hipFree(off);
      return nn_fail(PTY_EINVAL, "patch window outside the object");
    }
  }
  hipError_t e = hipMemcpy(off, h, (size_t)N * sizeof(int2), hipMemcpyHostToDevice);
  free(h);
  if (e == hipSuccess) e = hipMemsetAsync(acc, 0, (size_t)U * V * sizeof(float2), s);
  if (e == hipSuccess) e = hipMemsetAsync(wsum, 0, (size_t)U * V * sizeof(float), s);
  if (e != hipSuccess) {
    (void)hipFree(off);
    return nn_fail(PTY_EHIP, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_stitch, dim3(1), dim3(1024), 0, s, reinterpret_cast<const float2*>(patches),
                     pw, (const int2*)off, N, M, K, V, reinterpret_cast<float2*>(acc), wsum);
  hipLaunchKernelGGL(k_stitch_final, dim3(grid_for((size_t)U * V)), dim3(256), 0, s,
                     reinterpret_cast<const float2*>(acc), (const float*)wsum, (size_t)U * V,
                     reinterpret_cast<float2*>(O), floor, fill);
  e = hipStreamSynchronize(s);
  (void)hipFree(off);
  if (e != hipSuccess) return nn_fail(PTY_EHIP, hipGetErrorString(e));
  NN_LAUNCHED("k_stitch");
  return PTY_OK;
}

}  // extern "C"
