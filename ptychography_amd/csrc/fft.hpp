// Batched complex-to-complex FFT building blocks for gfx950.
//
// A line of L = R1*R2 points is transformed in two Stockham passes:
//   pass 1: R2 butterflies of radix R1 on x[j + r*R2]              (registers)
//   exchange through LDS (one write, one read, padded rows)
//   pass 2: R1 butterflies of radix R2 on mid[j + r*R1] * W_L^(j*r) (registers)
// giving X[j + r*R1] in natural order.  T threads share a line; every thread
// keeps E = L/T points in registers.  The radix-R1/R2 DFTs are unrolled
// register codelets built from radix 2, 3 and 4 with compile-time twiddles
// (double-precision constants rounded once to float).  Pass-2 twiddles come
// from a table of exp(-2 pi i k / L) computed in double on the host.
//
// Register layouts (t = thread within the line, q = butterfly slot):
//   layoutA<Ra,Rb>: e = q*Ra + r  <->  element (t + q*T) + r*Rb   (q < Rb/T, r < Ra)
//   layoutB<Ra,Rb>: e = q*Rb + r  <->  element (t + q*T) + r*Ra   (q < Ra/T, r < Rb)
// The forward transform uses (Ra,Rb) = (R1,R2) and maps layoutA -> layoutB.
// The inverse uses (R2,R1); its input layout is the forward's output layout,
// so a forward -> pointwise -> inverse chain needs no extra data movement.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace pty {

__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 operator-(float2 a) { return make_float2(-a.x, -a.y); }
__device__ __forceinline__ float2 operator*(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float norm2(float2 a) { return a.x * a.x + a.y * a.y; }

// Streaming loads / store (non-temporal hint) for frame stacks that are read
// or written once per kernel and would only push reused data out of L2.
__device__ __forceinline__ float2 load_stream(const float2* p) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v w = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(w.x, w.y);
}
__device__ __forceinline__ float load_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void store_stream(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream(float2* p, float2 v) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
}

// Raw buffer accesses relative to a wave-uniform base: the buffer resource
// (the base, no range limit) is held in SGPRs, the per-lane byte offset in
// one VGPR and a uniform byte offset in an SGPR, so an access costs no
// per-lane 64-bit address arithmetic.  `nt`: non-temporal (streaming).
// Offsets are bytes and must stay below 2^32.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
constexpr int BUF_NT = 2;  // cache-policy bits of a non-temporal access
template <int AUX = BUF_NT>
__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v w = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
  return make_float2(w.x, w.y);
}
template <int AUX = BUF_NT>
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}
template <int AUX = BUF_NT>
__device__ __forceinline__ void buf_store1(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, AUX);
}
template <int AUX = BUF_NT>
__device__ __forceinline__ void buf_store2(float2 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef unsigned u2v __attribute__((ext_vector_type(2)));
  const u2v w = {__builtin_bit_cast(unsigned, v.x), __builtin_bit_cast(unsigned, v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(w, r, voff, soff, AUX);
}

// multiply by -i (forward sign) or +i (inverse sign)
template <bool INV>
__device__ __forceinline__ float2 mul_negi(float2 a) {
  if constexpr (INV) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}

// Work-group barrier that orders LDS traffic only: global loads issued before
// it stay in flight (a plain __syncthreads also waits for every outstanding
// global load, which would serialise software-pipelined prefetches).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// cos / sin of 2*pi*k/48, k = 0..47 (every codelet radix divides 48)
struct Tw48 {
  static constexpr double c[48] = {
      1.0, 0.991444861373810411145, 0.96592582628906828675, 0.923879532511286756128,
      0.866025403784438646764, 0.79335334029123516458, 0.707106781186547524401,
      0.608761429008720639416, 0.5, 0.382683432365089771728, 0.258819045102520762349,
      0.130526192220051591548, 0.0, -0.130526192220051591548, -0.258819045102520762349,
      -0.382683432365089771728, -0.5, -0.608761429008720639416, -0.707106781186547524401,
      -0.79335334029123516458, -0.866025403784438646764, -0.923879532511286756128,
      -0.96592582628906828675, -0.991444861373810411145, -1.0, -0.991444861373810411145,
      -0.96592582628906828675, -0.923879532511286756128, -0.866025403784438646764,
      -0.79335334029123516458, -0.707106781186547524401, -0.608761429008720639416, -0.5,
      -0.382683432365089771728, -0.258819045102520762349, -0.130526192220051591548, 0.0,
      0.130526192220051591548, 0.258819045102520762349, 0.382683432365089771728, 0.5,
      0.608761429008720639416, 0.707106781186547524401, 0.79335334029123516458,
      0.866025403784438646764, 0.923879532511286756128, 0.96592582628906828675,
      0.991444861373810411145};
  static constexpr double s(int k) { return c[(k + 36) % 48]; }  // sin(x) = cos(x - pi/2)
};

// a * W_R^m, W_R = exp(-2 pi i / R) (forward) or its conjugate (inverse)
template <int R, int m, bool INV>
__device__ __forceinline__ float2 twiddle(float2 a) {
  constexpr int mm = ((m % R) + R) % R;
  if constexpr (mm == 0) {
    return a;
  } else if constexpr (2 * mm == R) {
    return -a;
  } else if constexpr (4 * mm == R) {
    return mul_negi<INV>(a);
  } else if constexpr (4 * mm == 3 * R) {
    return mul_negi<!INV>(a);
  } else {
    constexpr int k = mm * (48 / R);
    constexpr float wc = (float)Tw48::c[k];
    constexpr float ws = INV ? (float)Tw48::s(k) : -(float)Tw48::s(k);
    return make_float2(a.x * wc - a.y * ws, a.x * ws + a.y * wc);
  }
}

template <int R> struct Split;
template <> struct Split<6>  { static constexpr int A = 2, B = 3; };
template <> struct Split<8>  { static constexpr int A = 2, B = 4; };
template <> struct Split<12> { static constexpr int A = 4, B = 3; };
template <> struct Split<16> { static constexpr int A = 4, B = 4; };
template <> struct Split<24> { static constexpr int A = 8, B = 3; };

// In-register DFT of R points v[0..R-1] (natural order in, natural order out).
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 3) {
    constexpr float h = 0.866025403784438646764f;  // sqrt(3)/2
    float2 s = v[1] + v[2], d = v[1] - v[2];
    float2 a = v[0];
    v[0] = a + s;
    float2 m = make_float2(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
    // -i*h*d (forward) / +i*h*d (inverse)
    float2 r = mul_negi<INV>(d) * h;
    v[1] = m + r;
    v[2] = m - r;
  } else if constexpr (R == 4) {
    float2 t0 = v[0] + v[2], t1 = v[0] - v[2];
    float2 t2 = v[1] + v[3], t3 = mul_negi<INV>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  } else {
    // R = A*B: x[B*n1 + n2] -> A-point DFTs over n1 -> twiddle W_R^(n2*k1)
    //          -> B-point DFTs over n2 -> X[k1 + A*k2]
    constexpr int A = Split<R>::A, B = Split<R>::B;
    float2 y[R];
    static_for<0, B>([&](auto n2c) {
      constexpr int n2 = decltype(n2c)::value;
      float2 s[A];
      static_for<0, A>([&](auto n1c) { s[decltype(n1c)::value] = v[B * decltype(n1c)::value + n2]; });
      dft<A, INV>(s);
      static_for<0, A>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        y[n2 * A + k1] = twiddle<R, n2 * k1, INV>(s[k1]);
      });
    });
    static_for<0, A>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      float2 s[B];
      static_for<0, B>([&](auto n2c) { s[decltype(n2c)::value] = y[decltype(n2c)::value * A + k1]; });
      dft<B, INV>(s);
      static_for<0, B>([&](auto k2c) { v[k1 + A * decltype(k2c)::value] = s[decltype(k2c)::value]; });
    });
  }
}

// Line-FFT plans.  A plan key PK names a length L = plan_len(PK) and one
// factorisation of it: L = R1 * R2 (* R3 for three passes), T = threads per
// line, E = L / T points per thread.  Key L itself is the default plan of
// length L; plan_key(L, 1) an alternative the launch-shape tuner may pick
// (pty_set_plan).  96 and 192 take three passes by default so that a thread
// holds 12 points instead of 24 (the two-pass radix-24 / radix-12 x 8 plans
// needed ~250 VGPRs: 2 waves per SIMD).  The alternatives reverse the radix
// order (96, 192) or split the line over more threads with three passes
// (64: 4 x 4 x 4 on 16 threads, 128: 4 x 8 x 4 on 16, 256: 8 x 4 x 8 on 32),
// halving the points a thread holds for one more LDS exchange per transform.
constexpr int PLAN_STRIDE = 1024;
constexpr int plan_len(int PK) { return PK % PLAN_STRIDE; }
constexpr int plan_key(int L, int variant) { return L + PLAN_STRIDE * variant; }
template <int PK> struct Plan;
template <> struct Plan<16>  { static constexpr int R1 = 4,  R2 = 4,  R3 = 1, T = 4; };
template <> struct Plan<32>  { static constexpr int R1 = 8,  R2 = 4,  R3 = 1, T = 4; };
template <> struct Plan<48>  { static constexpr int R1 = 12, R2 = 4,  R3 = 1, T = 4; };
template <> struct Plan<64>  { static constexpr int R1 = 8,  R2 = 8,  R3 = 1, T = 8; };
template <> struct Plan<96>  { static constexpr int R1 = 6,  R2 = 4,  R3 = 4, T = 8; };
template <> struct Plan<128> { static constexpr int R1 = 16, R2 = 8,  R3 = 1, T = 8; };
template <> struct Plan<192> { static constexpr int R1 = 12, R2 = 4,  R3 = 4, T = 16; };
template <> struct Plan<256> { static constexpr int R1 = 16, R2 = 16, R3 = 1, T = 16; };
template <> struct Plan<plan_key(64, 1)>  { static constexpr int R1 = 4, R2 = 4,  R3 = 4,  T = 16; };
template <> struct Plan<plan_key(96, 1)>  { static constexpr int R1 = 4, R2 = 4,  R3 = 6,  T = 8; };
template <> struct Plan<plan_key(128, 1)> { static constexpr int R1 = 4, R2 = 8,  R3 = 4,  T = 16; };
template <> struct Plan<plan_key(192, 1)> { static constexpr int R1 = 4, R2 = 4,  R3 = 12, T = 16; };
template <> struct Plan<plan_key(256, 1)> { static constexpr int R1 = 8, R2 = 4,  R3 = 8,  T = 32; };

// Element index held in register slot e by thread t, for layoutA / layoutB.
template <int Ra, int Rb, int T>
__device__ __forceinline__ int idxA(int e, int t) { return (t + (e / Ra) * T) + (e % Ra) * Rb; }
template <int Ra, int Rb, int T>
__device__ __forceinline__ int idxB(int e, int t) { return (t + (e / Rb) * T) + (e % Rb) * Ra; }

// Bank-conflict-free layout of the three-pass exchange matrices.  A 64-bit
// LDS access serves 16 lanes per cycle over 16 two-bank slots, so it is free
// of conflicts when the 16 lanes' float2 addresses differ mod 16.  Pass 2
// reads mid[k1 S1 + j + Rc m] with k1 = g % Ra fastest and j = g / Ra, and
// writes mid[j S2 + k1 + Ra kb]: with the row strides N2 + 1 / N3 + 1 those
// addresses collide (k1 + j mod 16: the 8 x 4 x 8 plan's row gather lost ~64
// % of its LDS cycles to conflicts).  For Ra in {2, 4, 8}, rows that are a
// multiple of 16 long and lines of at least 16 threads (each 16-lane group
// within one line, starting at a multiple of 16) the matrices are stored
// unpadded with the columns XOR-swizzled: column c of row k1 (pass 1 / 2) at
// c ^ (k1 16 / Ra), column c of row j (pass 2 / 3) at c ^ (j Ra) mod 16.
// Every 16-lane access then covers all 16 slots, and padding-free rows
// keep the exchange area at L words.  Other plans keep odd strides N + 1.
constexpr bool swz3(int Ra, int N2, int N3, int T) {
  return T % 16 == 0 && (Ra == 2 || Ra == 4 || Ra == 8) && N2 % 16 == 0 && N3 % 16 == 0;
}
// words of the two exchange matrices of a three-pass transform (Ra, Rb, Rc)
constexpr int words3(int Ra, int Rb, int Rc, int T) {
  const int N2 = Rb * Rc, N3 = Ra * Rb;
  const bool sw = swz3(Ra, N2, N3, T);
  const int a = Ra * (sw ? N2 : N2 + 1), b = Rc * (sw ? N3 : N3 + 1);
  return a > b ? a : b;
}

// LDS words (float2) one line needs for its exchange buffer (both directions).
template <int PK>
struct LineLds {
  static constexpr int R1 = Plan<PK>::R1, R2 = Plan<PK>::R2, R3 = Plan<PK>::R3;
  // three passes: forward (R1, R2, R3) and inverse (R3, R2, R1) matrices
  static constexpr int f3 = R3 > 1 ? words3(R1, R2, R3, Plan<PK>::T) : 0;
  static constexpr int i3 = R3 > 1 ? words3(R3, R2, R1, Plan<PK>::T) : 0;
  static constexpr int m3 = f3 > i3 ? f3 : i3;
  static constexpr int a = R3 > 1 ? m3 : R2 * (R1 + 1);
  static constexpr int b = R3 > 1 ? m3 : R1 * (R2 + 1);
  static constexpr int words = (a > b ? a : b) | 1;  // odd stride between lines
};

// One two-pass transform of a line held in registers v[E] (layoutA<Ra,Rb> in,
// layoutB<Ra,Rb> out).  `mid` is this line's LDS exchange area, `tw` the
// table exp(-2 pi i k / L).  Contains one LDS barrier (between the passes);
// the caller must put an LDS barrier between two calls that reuse `mid`.
template <int L, int Ra, int Rb, int T, bool INV>
__device__ __forceinline__ void line_fft(float2* v, float2* mid, const float2* __restrict__ tw, int t) {
  static_assert(Ra * Rb == L, "bad factorisation");
  static_assert(Ra % T == 0 && Rb % T == 0, "T must divide both radices");
  constexpr int E = L / T;
  constexpr int NQ1 = Rb / T;  // pass-1 butterflies per thread
  constexpr int NQ2 = Ra / T;  // pass-2 butterflies per thread
  constexpr int PAD = Ra + 1;
  static_for<0, NQ1>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    dft<Ra, INV>(v + q * Ra);
    const int j = t + q * T;  // row of the exchange matrix
#pragma unroll
    for (int r = 0; r < Ra; ++r) mid[j * PAD + r] = v[q * Ra + r];
  });
  lds_barrier();
  static_for<0, NQ2>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int j = t + q * T;
#pragma unroll
    for (int r = 0; r < Rb; ++r) {
      float2 a = mid[r * PAD + j];
      float2 w = tw[j * r];
      if constexpr (INV) w.y = -w.y;
      v[q * Rb + r] = (r == 0) ? a : cmul(a, w);
    }
    dft<Rb, INV>(v + q * Rb);
  });
  static_assert(NQ1 * Ra == E && NQ2 * Rb == E, "layout size mismatch");
}

// Three-pass transform of a line of L = Ra * Rb * Rc points held in registers
// (E = L / T per thread).  With N2 = Rb * Rc and N3 = Ra * Rb:
//   pass 1: radix-Ra DFTs over x[n2 + N2 r] for n2 = t + q T, times W_L^(n2 k1)
//   pass 2: for each (k1, j) = (g % Ra, g / Ra), g = t + q T: radix-Rb DFTs over
//           y[k1][j + Rc m], times W_L^(Ra j kb)          (the N2-point DFT of
//   pass 3: for each b = k1 + Ra kb = t + q T: radix-Rc     row k1, split Rb x Rc)
//           DFTs over j, giving X[b + N3 kc]
// In: slot e = q Ra + r <-> element t + q T + N2 r; out: slot e = q Rc + kc <->
// element t + q T + N3 kc.  The inverse is the same network with the radices
// reversed (input in the forward's output layout, output in its input layout).
// Three LDS barriers; the caller puts one between two calls that reuse `mid`.
template <int L, int Ra, int Rb, int Rc, int T, bool INV>
__device__ __forceinline__ void line_fft3(float2* v, float2* mid, const float2* __restrict__ tw,
                                          int t) {
  static_assert(Ra * Rb * Rc == L, "bad factorisation");
  constexpr int N2 = Rb * Rc, N3 = Ra * Rb;
  static_assert(N2 % T == 0 && (Ra * Rc) % T == 0 && N3 % T == 0, "T must divide every pass");
  constexpr int E = L / T, NQ1 = N2 / T, NQ2 = (Ra * Rc) / T, NQ3 = N3 / T;
  static_assert(NQ1 * Ra == E && NQ2 * Rb == E && NQ3 * Rc == E, "layout size mismatch");
  constexpr bool SW = swz3(Ra, N2, N3, T);
  constexpr int S1 = SW ? N2 : N2 + 1, S2 = SW ? N3 : N3 + 1;
  auto x1 = [](int k1) { return SW ? (k1 * (16 / Ra)) & 15 : 0; };  // row k1 of matrix 1
  auto x2 = [](int j) { return SW ? (j * Ra) & 15 : 0; };           // row j of matrix 2
  auto twid = [&](float2 a, int k) {
    float2 w = tw[k];
    if constexpr (INV) w.y = -w.y;
    return cmul(a, w);
  };
  static_for<0, NQ1>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    dft<Ra, INV>(v + q * Ra);
    const int n2 = t + q * T;
#pragma unroll
    for (int k1 = 0; k1 < Ra; ++k1)
      mid[k1 * S1 + (n2 ^ x1(k1))] = k1 == 0 ? v[q * Ra] : twid(v[q * Ra + k1], n2 * k1);
  });
  lds_barrier();
  static_for<0, NQ2>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int g = t + q * T, k1 = g % Ra, j = g / Ra;
#pragma unroll
    for (int m = 0; m < Rb; ++m) v[q * Rb + m] = mid[k1 * S1 + ((j + Rc * m) ^ x1(k1))];
  });
  lds_barrier();
  static_for<0, NQ2>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int g = t + q * T, k1 = g % Ra, j = g / Ra;
    dft<Rb, INV>(v + q * Rb);
#pragma unroll
    for (int kb = 0; kb < Rb; ++kb)
      mid[j * S2 + ((k1 + Ra * kb) ^ x2(j))] = kb == 0 ? v[q * Rb] : twid(v[q * Rb + kb], Ra * j * kb);
  });
  lds_barrier();
  static_for<0, NQ3>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int b = t + q * T;
#pragma unroll
    for (int j = 0; j < Rc; ++j) v[q * Rc + j] = mid[j * S2 + (b ^ x2(j))];
    dft<Rc, INV>(v + q * Rc);
  });
}

// Plan-level transforms and register layouts used by the kernels.  Layout A
// (forward input, inverse output): slot e of thread t holds element
// t + (e / R1) T + (e % R1) (L / R1); layout B (forward output, inverse input):
// the same with the last radix.  Both are "thread index + per-slot constant".
template <int PK>
__device__ __forceinline__ void fft_fwd(float2* v, float2* mid, const float2* __restrict__ tw, int t) {
  using Pl = Plan<PK>;
  constexpr int L = plan_len(PK);
  if constexpr (Pl::R3 > 1) line_fft3<L, Pl::R1, Pl::R2, Pl::R3, Pl::T, false>(v, mid, tw, t);
  else line_fft<L, Pl::R1, Pl::R2, Pl::T, false>(v, mid, tw, t);
}
template <int PK>
__device__ __forceinline__ void fft_inv(float2* v, float2* mid, const float2* __restrict__ tw, int t) {
  using Pl = Plan<PK>;
  constexpr int L = plan_len(PK);
  if constexpr (Pl::R3 > 1) line_fft3<L, Pl::R3, Pl::R2, Pl::R1, Pl::T, true>(v, mid, tw, t);
  else line_fft<L, Pl::R2, Pl::R1, Pl::T, true>(v, mid, tw, t);
}
template <int PK> struct Lay {
  static constexpr int T = Plan<PK>::T, RA = Plan<PK>::R1;
  static constexpr int RB = Plan<PK>::R3 > 1 ? Plan<PK>::R3 : Plan<PK>::R2;
};
template <int PK>
__host__ __device__ __forceinline__ constexpr int rowA(int e) {
  return (e / Lay<PK>::RA) * Lay<PK>::T + (e % Lay<PK>::RA) * (plan_len(PK) / Lay<PK>::RA);
}
template <int PK>
__host__ __device__ __forceinline__ constexpr int rowB(int e) {
  return (e / Lay<PK>::RB) * Lay<PK>::T + (e % Lay<PK>::RB) * (plan_len(PK) / Lay<PK>::RB);
}
template <int PK>
__host__ __device__ __forceinline__ constexpr int layA(int e, int t) { return t + rowA<PK>(e); }
template <int PK>
__host__ __device__ __forceinline__ constexpr int layB(int e, int t) { return t + rowB<PK>(e); }

// Compile-time checks: for two-pass plans the layouts are the two-pass
// Stockham ones (idxA / idxB of R1, R2), and every layout maps the T x E
// (thread, slot) pairs one-to-one onto 0 .. L-1.
template <int PK>
constexpr bool layouts_ok() {
  constexpr int L = plan_len(PK);
  constexpr int T = Plan<PK>::T, E = L / T, R1 = Plan<PK>::R1, R2 = Plan<PK>::R2;
  static_assert(Plan<PK>::R1 * Plan<PK>::R2 * Plan<PK>::R3 == L, "plan factorisation");
  bool seenA[L] = {}, seenB[L] = {};
  for (int t = 0; t < T; ++t)
    for (int e = 0; e < E; ++e) {
      const int a = layA<PK>(e, t), b = layB<PK>(e, t);
      if (a < 0 || a >= L || b < 0 || b >= L || seenA[a] || seenB[b]) return false;
      seenA[a] = seenB[b] = true;
      if (Plan<PK>::R3 == 1) {
        if (a != (t + (e / R1) * T) + (e % R1) * R2) return false;
        if (b != (t + (e / R2) * T) + (e % R2) * R1) return false;
      }
    }
  return true;
}
static_assert(layouts_ok<16>() && layouts_ok<32>() && layouts_ok<48>() && layouts_ok<64>() &&
                  layouts_ok<96>() && layouts_ok<128>() && layouts_ok<192>() && layouts_ok<256>(),
              "line layouts");
static_assert(layouts_ok<plan_key(64, 1)>() && layouts_ok<plan_key(96, 1)>() &&
                  layouts_ok<plan_key(128, 1)>() && layouts_ok<plan_key(192, 1)>() &&
                  layouts_ok<plan_key(256, 1)>(),
              "alternative line layouts");

}  // namespace pty
