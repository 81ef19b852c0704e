// Host AddressSanitizer run of the C ABI (include/ptyhip.h).
//
// Built by tests/native/Makefile from the library's own sources with
// AddressSanitizer on the host code only (-Xarch_host -fsanitize=address):
// the device code is the ordinary gfx950 build.  It drives every host path
// of the library -- argument checks, the frame / object-row tables built at
// creation (one and several column segments, and an object too wide for the
// row lists), packed scans, batches with and without graphs, the tuner, the
// alternative line-FFT plans, sub-pixel shifts, position terms and the
// building-block entry points -- on small synthetic scans, and checks every
// status code.  Results are not compared here (the GPU tests do that); a
// heap overflow, use after free or double free in host code stops the run
// with ASan's report, and so does any unexpected status.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "ptyhip.h"

static int g_checks = 0;

#define HIPOK(x)                                                                    \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x,                \
                   hipGetErrorString(e_));                                          \
      std::exit(2);                                                                 \
    }                                                                               \
  } while (0)

static void expect(int got, int want, const char* what, int line) {
  ++g_checks;
  if (got != want) {
    std::fprintf(stderr, "line %d: %s returned %d, expected %d (%s)\n", line, what, got, want,
                 pty_last_error());
    std::exit(1);
  }
}
#define OK(x) expect((x), PTY_OK, #x, __LINE__)
#define FAILS(x, code) expect((x), (code), #x, __LINE__)

template <class T>
static T* dev_alloc(size_t n) {
  T* p = nullptr;
  HIPOK(hipMalloc((void**)&p, n * sizeof(T)));
  HIPOK(hipMemset(p, 0, n * sizeof(T)));
  return p;
}

template <class T>
static T* dev_copy(const std::vector<T>& h) {
  T* p = dev_alloc<T>(h.size());
  HIPOK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

// A raster scan of gy x gx frames of side L, `step` pixels apart, in the
// reference's convention after its shift: R <= 0, frame n sees
// O[-R[n,0] : L - R[n,0], -R[n,1] : L - R[n,1]].
struct Scan {
  int N, L, U, V;
  std::vector<int32_t> R;
  std::vector<float> I;
  std::vector<uint8_t> mask;
  std::vector<pty_cfloat> O, P;
};

static Scan make_scan(int L, int gy, int gx, int step, unsigned seed) {
  Scan s;
  s.L = L;
  s.N = gy * gx;
  s.U = L + (gy - 1) * step;
  s.V = L + (gx - 1) * step;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> uni(0.f, 1.f);
  for (int y = 0; y < gy; ++y)
    for (int x = 0; x < gx; ++x) {
      s.R.push_back(-y * step);
      s.R.push_back(-x * step);
    }
  s.I.resize((size_t)s.N * L * L);
  for (auto& v : s.I) v = 1e4f * uni(rng);
  s.mask.assign((size_t)L * L, 1);
  for (int i = 0; i < L * L; i += 17) s.mask[i] = 0;
  s.O.resize((size_t)s.U * s.V);
  for (auto& o : s.O) {
    float ph = 0.3f * uni(rng);
    o = {std::cos(ph), std::sin(ph)};
  }
  s.P.resize((size_t)L * L);
  for (int y = 0; y < L; ++y)
    for (int x = 0; x < L; ++x) {
      float r2 = (y - L / 2.f) * (y - L / 2.f) + (x - L / 2.f) * (x - L / 2.f);
      s.P[(size_t)y * L + x] = {std::exp(-r2 / (2.f * L * L / 16.f)), 0.f};
    }
  return s;
}

static pty_opts opts(int algorithm, int method) {
  pty_opts o;
  std::memset(&o, 0, sizeof o);
  o.algorithm = algorithm;
  o.method = method;
  o.op_n = 1;
  o.op_every = 1;
  o.alpha = 1e-10f;
  return o;
}

// Device buffers of one context's data and state.
struct Bound {
  pty_ctx* c = nullptr;
  float* I = nullptr;
  float* amp = nullptr;
  uint8_t* mask = nullptr;
  pty_cfloat* O = nullptr;
  pty_cfloat* P = nullptr;
  void release() {
    OK(pty_destroy(c));
    HIPOK(hipFree(I));
    HIPOK(hipFree(amp));
    HIPOK(hipFree(mask));
    HIPOK(hipFree(O));
    HIPOK(hipFree(P));
  }
};

static Bound bind(const Scan& s) {
  Bound b;
  OK(pty_create(&b.c, s.N, s.L, s.L, s.U, s.V, s.R.data()));
  b.I = dev_copy(s.I);
  b.mask = dev_copy(s.mask);
  b.amp = dev_alloc<float>(s.I.size());
  b.O = dev_copy(s.O);
  b.P = dev_copy(s.P);
  double I_norm = 0.0;
  OK(pty_prepare_amp(b.c, b.I, b.mask, b.amp, &I_norm));
  OK(pty_set_data(b.c, b.amp, b.mask, I_norm));
  OK(pty_set_state(b.c, b.O, b.P));
  return b;
}

static void argument_checks() {
  Scan s = make_scan(64, 3, 3, 16, 1);
  pty_ctx* c = nullptr;
  FAILS(pty_create(&c, 0, 64, 64, s.U, s.V, s.R.data()), PTY_EINVAL);
  FAILS(pty_create(&c, s.N, 48, 48, s.U, s.V, s.R.data()), PTY_EUNSUP);
  FAILS(pty_create(&c, s.N, 64, 64, s.U, s.V, nullptr), PTY_EINVAL);
  std::vector<int32_t> out = s.R;
  out[2 * s.N - 1] = -(s.V - 63);  // the last window runs past the object's edge
  FAILS(pty_create(&c, s.N, 64, 64, s.U, s.V, out.data()), PTY_EINVAL);
  out = s.R;
  out[0] = 1;  // R > 0
  FAILS(pty_create(&c, s.N, 64, 64, s.U, s.V, out.data()), PTY_EINVAL);
  OK(pty_create(&c, s.N, 64, 64, s.U, s.V, s.R.data()));
  pty_opts o = opts(0, 1);
  FAILS(pty_run(c, &o, 0, 2, 1, nullptr), PTY_ESTATE);  // no data bound
  std::vector<int32_t> fs = {0, 4, s.N}, rs = {0, 7, s.U};  // band start not a multiple of 16
  FAILS(pty_set_scans(c, 2, fs.data(), rs.data(), nullptr), PTY_EINVAL);
  FAILS(pty_set_plan(c, 16), PTY_EINVAL);
  OK(pty_destroy(c));
}

static void solve_paths(const Scan& s, bool extras) {
  Bound b = bind(s);
  std::vector<double> m(3 * 8);
  pty_opts era1 = opts(0, 1), era3 = opts(0, 3), dm1 = opts(1, 1), dm3 = opts(1, 3);
  era3.op_n = 1, era3.op_every = 2;
  dm3.op_n = 1, dm3.op_every = 3;
  OK(pty_run(b.c, &era1, 0, 4, 1, m.data()));
  OK(pty_run(b.c, &era3, 0, 4, 1, m.data()));
  OK(pty_run(b.c, &dm1, 0, 4, 1, m.data()));
  OK(pty_run(b.c, &dm3, 0, 6, 1, m.data()));
  std::vector<float> bg((size_t)s.L * s.L, 0.1f);
  float* bg_d = dev_copy(bg);
  OK(pty_set_background(b.c, bg_d, 0));
  pty_opts dm4 = opts(1, 4), era4 = opts(0, 4);
  OK(pty_run(b.c, &dm4, 0, 4, 1, m.data()));
  OK(pty_run(b.c, &era4, 0, 4, 1, m.data()));
  float* bout = dev_alloc<float>(s.I.size());
  OK(pty_background(b.c, bout));
  float* Iout = dev_alloc<float>(s.I.size());
  OK(pty_intensity(b.c, Iout));
  double* H = dev_alloc<double>((size_t)s.U * s.V);
  int32_t valid = -1;
  OK(pty_heatmap(b.c, H, &valid));
  if (extras) {
    // centring and the far-field probe constraint
    pty_opts e = opts(0, 3);
    e.probe_centering = 1;
    e.pmod_probe_until = 3;
    OK(pty_set_pamp(b.c, nullptr));
    OK(pty_run(b.c, &e, 0, 4, 1, m.data()));
    // graphs, alternative plans, the unfused inverse-row path, profiling
    OK(pty_set_graphs(b.c, 1));
    OK(pty_run(b.c, &era1, 0, 6, 1, m.data()));
    int32_t en = 0, fell = 0;
    OK(pty_get_graphs(b.c, &en, &fell));
    OK(pty_set_graphs(b.c, 0));
    OK(pty_set_plan(b.c, 15));
    OK(pty_run(b.c, &dm1, 0, 3, 1, m.data()));
    OK(pty_set_plan(b.c, 0));
    OK(pty_set_row_gather(b.c, 0));
    OK(pty_run(b.c, &era1, 0, 3, 1, m.data()));
    OK(pty_set_row_gather(b.c, 1));
    OK(pty_profile(b.c, 1));
    OK(pty_run(b.c, &era1, 0, 3, 1, m.data()));
    std::vector<double> ms(PTY_NSTAGE);
    std::vector<int64_t> cnt(PTY_NSTAGE);
    OK(pty_profile_read(b.c, ms.data(), cnt.data()));
    OK(pty_profile(b.c, 0));
    if (s.L == 64) {
      pty_tune_result tr;
      OK(pty_tune(b.c, &era1, 1, &tr));
    }
    OK(pty_run(b.c, &era1, 0, 2, 1, m.data()));
    // sub-pixel shifts and position terms
    std::vector<double> d(2 * s.N);
    for (int n = 0; n < s.N; ++n) d[2 * n] = 0.25 * (n % 3), d[2 * n + 1] = -0.125 * (n % 5);
    OK(pty_set_subpixel(b.c, d.data()));
    OK(pty_run(b.c, &era1, 0, 3, 1, m.data()));
    std::vector<double> terms(6 * s.N);
    OK(pty_position_terms(b.c, terms.data()));
    OK(pty_set_subpixel(b.c, nullptr));
    // building blocks
    size_t nf = s.I.size();
    pty_cfloat* ex = dev_alloc<pty_cfloat>(nf);
    pty_cfloat* f = dev_alloc<pty_cfloat>(nf);
    pty_cfloat* mm = dev_alloc<pty_cfloat>(nf);
    pty_cfloat* O2 = dev_alloc<pty_cfloat>((size_t)s.U * s.V);
    pty_cfloat* P2 = dev_alloc<pty_cfloat>((size_t)s.L * s.L);
    double* G = dev_alloc<double>((size_t)s.L * s.L);
    double emod = 0.0;
    OK(pty_make_exits(b.c, b.O, b.P, ex));
    OK(pty_fft2(b.c, ex, s.N, 0));
    OK(pty_fft2(b.c, ex, s.N, 1));
    OK(pty_pmod1(b.c, ex, 1e-10f, &emod));
    OK(pty_psup_O(b.c, ex, b.P, O2, H, 1, 1e-10f));
    OK(pty_psup_P(b.c, ex, O2, P2, G, 1, 1e-10f));
    OK(pty_back_projection(b.c, b.P, ex, O2, H, 1e-10f));
    OK(pty_hio_modulus(b.c, f, mm, b.O, b.P, 0.f, 0.9f, 1e-10f, &emod));
    OK(pty_add_exits(b.c, b.O, b.P, 0.9f, f));
    OK(pty_hio_modulus(b.c, f, mm, b.O, b.P, 0.9f, 0.9f, 1e-10f, &emod));
    // DM iterate round trip
    OK(pty_run(b.c, &dm1, 0, 2, 1, m.data()));
    float* bi = dev_alloc<float>(nf);
    OK(pty_get_iterate(b.c, f, bi));
    OK(pty_set_iterate(b.c, f, b.O, b.P));
    OK(pty_run(b.c, &dm1, 2, 2, 0, m.data()));
    for (void* p : {(void*)ex, (void*)f, (void*)mm, (void*)O2, (void*)P2, (void*)G, (void*)bi})
      HIPOK(hipFree(p));
  }
  HIPOK(hipFree(bg_d));
  HIPOK(hipFree(bout));
  HIPOK(hipFree(Iout));
  HIPOK(hipFree(H));
  b.release();
}

// Two scans of one frame size packed into one context (row bands on
// multiples of 16 rows), then a batch of that context and a single one.
static void packed_and_batch() {
  Scan a = make_scan(64, 4, 5, 16, 11), b = make_scan(64, 3, 3, 20, 12);
  int U0 = (a.U + 15) / 16 * 16;
  Scan p;
  p.L = 64;
  p.N = a.N + b.N;
  p.U = U0 + b.U;
  p.V = a.V > b.V ? a.V : b.V;
  p.R = a.R;
  for (int n = 0; n < b.N; ++n) p.R.push_back(b.R[2 * n] - U0), p.R.push_back(b.R[2 * n + 1]);
  p.I = a.I;
  p.I.insert(p.I.end(), b.I.begin(), b.I.end());
  p.mask = a.mask;
  p.mask.insert(p.mask.end(), b.mask.begin(), b.mask.end());
  p.O.assign((size_t)p.U * p.V, pty_cfloat{1.f, 0.f});
  p.P = a.P;
  p.P.insert(p.P.end(), b.P.begin(), b.P.end());

  Bound pk;
  OK(pty_create(&pk.c, p.N, 64, 64, p.U, p.V, p.R.data()));
  std::vector<int32_t> fs = {0, a.N, p.N}, rs = {0, U0, p.U};
  OK(pty_set_scans(pk.c, 2, fs.data(), rs.data(), nullptr));
  pk.I = dev_copy(p.I);
  pk.mask = dev_copy(p.mask);
  pk.amp = dev_alloc<float>(p.I.size());
  pk.O = dev_copy(p.O);
  pk.P = dev_copy(p.P);
  double I_norm[2];
  OK(pty_prepare_amp(pk.c, pk.I, pk.mask, pk.amp, I_norm));
  OK(pty_set_data(pk.c, pk.amp, pk.mask, I_norm[0]));
  OK(pty_set_state(pk.c, pk.O, pk.P));
  std::vector<double> m(3 * 2 * 4);
  pty_opts era1 = opts(0, 1), dm1 = opts(1, 1);
  OK(pty_run(pk.c, &era1, 0, 4, 1, m.data()));
  FAILS(pty_run(pk.c, &dm1, 0, 4, 1, m.data()), PTY_EUNSUP);

  Bound single = bind(make_scan(64, 3, 4, 12, 13));
  pty_batch* bt = nullptr;
  pty_ctx* cs[2] = {pk.c, single.c};
  OK(pty_batch_create(&bt, 2, cs));
  pty_opts os[2] = {era1, opts(0, 3)};
  std::vector<double> m0(3 * 2 * 5), m1(3 * 5);
  double* ms[2] = {m0.data(), m1.data()};
  OK(pty_batch_run(bt, os, 0, 5, 1, ms));
  OK(pty_batch_set_graphs(bt, 1));
  OK(pty_batch_run(bt, os, 0, 5, 1, ms));
  int32_t en = 0, fell = 0;
  OK(pty_batch_get_graphs(bt, &en, &fell));
  OK(pty_batch_destroy(bt));
  single.release();
  pk.release();
}

int main() {
  std::printf("argument checks\n");
  std::fflush(stdout);
  argument_checks();
  const int frames[] = {32, 64, 96, 128, 192, 256};
  for (int L : frames) {
    std::printf("solve paths, frame %d\n", L);
    std::fflush(stdout);
    solve_paths(make_scan(L, 3, 3, L / 4, 100 + L), L == 64 || L == 256);
  }
  std::printf("object of several row segments (2 x 32 frames of 256, 130 apart)\n");
  std::fflush(stdout);
  solve_paths(make_scan(256, 2, 32, 130, 7), false);
  std::printf("object wider than the row lists take (2 x 66 frames of 128, 126 apart)\n");
  std::fflush(stdout);
  solve_paths(make_scan(128, 2, 66, 126, 8), false);
  std::printf("packed scans and batches\n");
  std::fflush(stdout);
  packed_and_batch();
  std::printf("asan driver: %d status checks passed\n", g_checks);
  return 0;
}
