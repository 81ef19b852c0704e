// ptyhip: host engine and C ABI of the gfx950 ERA / DM iteration.
//
// One context = one scan geometry (N frames of M x K, object U x V, integer
// window offsets) plus the workspace its kernels need.  pty_run executes the
// reference's loop (ptychography/era.py:152-291, ptychography/dm.py:45-294) as
// a sequence of kernel launches on the context's stream, keeping every array
// on the device; per-iteration metrics are written to device memory and
// copied back once per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "frame.hpp"
#include "kernels.hpp"
#include "ptyhip.h"

using namespace pty;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PTY_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(PTY_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

#define PTY_LAUNCHED(name)                                                             \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess)                                                              \
      return fail(PTY_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PTY_TRY(expr)        \
  do {                       \
    int r_ = (expr);         \
    if (r_ != PTY_OK) return r_; \
  } while (0)

#define PTY_SIZES(X) X(16) X(32) X(64) X(96) X(128) X(192) X(256)
// every line-FFT plan the row / column kernels are built for: the default
// plan of each side and the alternatives (fft.hpp)
// stage bits of a context's plan mask (pty_set_plan)
constexpr int PLAN_ROWS = PTY_PLAN_ROWS, PLAN_COLS = PTY_PLAN_COLS, PLAN_ROWG = PTY_PLAN_ROWG;
constexpr int PLAN_BGERR = PTY_PLAN_BGERR;
constexpr int PLAN_ALL = PLAN_ROWS | PLAN_COLS | PLAN_ROWG | PLAN_BGERR;
#define PTY_PLANS(X) PTY_SIZES(X) X(plan_key(64, 1)) X(plan_key(96, 1)) X(plan_key(128, 1)) \
  X(plan_key(192, 1)) X(plan_key(256, 1))

// threads per line of plan key PK (for a side L: its default plan); 0 = none
int threads_per_line(int PK) {
  switch (PK) {
#define C(Lv) case Lv: return Plan<Lv>::T;
    PTY_PLANS(C)
#undef C
    default: return 0;
  }
}

template <class T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(PTY_ENOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                " bytes: " + hipGetErrorString(e));
  }
  return PTY_OK;
}

template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

constexpr int ECON_BLOCKS = 512;
constexpr int ESUP_BLOCKS = 1024;
// The staged ERA modulus step can run over at most MAX_CHUNKS frame blocks
// (the column kernels' error partials are appended per block); the partial
// buffers hold MAX_CHUNKS x the per-launch capacity.
constexpr int MAX_CHUNKS = 16;

}  // namespace

// host-side state an iteration changes; a graph replay restores what its
// capture left
struct HostIterState {
  bool H_valid, G_valid, era_bg_shared, have_metric_state, os_next, os_next_op;
  float2* F;
  float2* X;
  bool dm1_pending;
  int dm1_par;
  bool f_spec;
};

struct pty_ctx {
  int N = 0, M = 0, K = 0, U = 0, V = 0;
  hipStream_t stream = nullptr;
  Geom g{};
  int2* off = nullptr;
  int tiles_x = 0, tiles_y = 0, ntiles = 0;
  int* tile_ptr = nullptr;
  int* tile_frames = nullptr;
  // object rows -> ascending frame lists (CSR), for the fused inverse-row /
  // object-update kernel of ERA object-only steps
  int* row_ptr = nullptr;
  int2* row_ent = nullptr;   // (frame row n*M + i, column offset x_n << 16 | i)
  bool rowg_ok = false;      // geometry supports the fused kernel
  bool rowg = false;         // currently selected
  int rowg_segs = 1, rowg_segw = 0;  // column segments per object row, their width
  // frame rows listed per object-row segment (by where their window starts),
  // for the forward row stage of O_win P in object-row order
  int* rowf_ptr = nullptr;
  int2* rowf_ent = nullptr;
  bool rowf_ok = false;      // lists built
  bool rowf = false;         // currently selected
  int rowf_segs = 1, rowf_segw = 0;
  float2* twM = nullptr;
  float2* twK = nullptr;
  // column-stage launch shape
  int col_blocks = 0, col_chunk = 1, col_chunks = 0;
  int rows_mult = 3;         // persistent row kernels: up to rows_mult * CUs work-groups
  int cols_mult = 0;         // persistent column kernels: cols_mult * CUs (0 = per-mode default)
  int cols_wg = 0;           // threads per column work-group: 256, 512 (32-column blocks,
                             // 256^2 frames) or 0 = 512 for the Pmod_1 modes at 256^2, else 256
  int chunk_frames = 0;      // ERA modulus step over blocks of this many frames (0 = all at once)
  int cus = 256;             // compute units of the device (persistent grids)
  // probe-reduction launch shape
  int probe_chunk = 1, probe_chunks = 0;

  // bound (caller-owned)
  const float* amp = nullptr;
  const uint8_t* mask = nullptr;
  uint32_t* maskbits = nullptr;  // T * K packed mask words for the column stage (owned)
  uint32_t* maskbits_alt = nullptr;  // the same for the side's alternative plan
  double I_norm = 0.0;
  float2* O = nullptr;
  float2* P = nullptr;

  // owned workspace
  float2* X = nullptr;       // N*M*K frame scratch (ERA: psi')
  float2* F = nullptr;       // N*M*K DM iterate f
  float2* Xp = nullptr;      // M*K probe scratch
  float2* Obak = nullptr;    // U*V
  float2* Pbak = nullptr;    // M*K
  float2* Os = nullptr;      // U*V DM metric object
  float2* Oroll = nullptr;   // U*V scratch for the periodic object roll
  float2* Ps = nullptr;      // M*K DM metric probe / roll scratch
  double* H = nullptr;       // U*V P heatmap
  double* G = nullptr;       // M*K O heatmap
  bool H_valid = false, G_valid = false;
  float* bg = nullptr;       // N*M*K background state
  bool bg_set = false;
  float* bbar = nullptr;     // M*K main-path frame mean
  float* bbar_m = nullptr;   // M*K metric-path frame mean
  bool era_bg_shared = false;  // ERA: the state collapsed to bbar after the first mean
  float* pamp = nullptr;     // M*K |FFT P0|
  bool pamp_set = false;
  double* partb = nullptr;   // col_chunks * M*K background sums
  double2* pnum = nullptr;   // probe_chunks * M*K
  double* pg = nullptr;      // probe_chunks * M*K
  int* cm = nullptr;         // probe centre shift
  double* metrics = nullptr;
  int metrics_cap = 0;
  double* scal = nullptr;    // scalar results
  // partial pool
  double* part_emod = nullptr;   // col WGs
  double* part_esa = nullptr;    // col WGs
  double* part_econO = nullptr;  // 2 * max(ntiles, ECON_BLOCKS)
  double* part_econP = nullptr;  // 2 * max(MK/256, ECON_BLOCKS)
  double* part_esup = nullptr;   // max(ESUP_BLOCKS, row blocks)
  double* part_esxh = nullptr;   // 2 * max(ntiles, U)
  int cap_gather = 0;            // max(ntiles, U): object-update work-groups
  int cap_col = 0, cap_econO = 0, cap_econP = 0, cap_esup = 0;

  // optional per-stage timing with HIP events on the context stream
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;   // pairs (start, stop)
  std::vector<int> prof_stage;       // stage of each used pair
  size_t prof_used = 0;
  int prof_override = -1;            // >= 0: attribute launches to this stage
  double prof_ms[PTY_NSTAGE] = {};
  int64_t prof_cnt[PTY_NSTAGE] = {};

  // fused per-frame path (square frames of a supported side)
  bool frame_ok = false;     // geometry supports it
  bool use_frame = false;    // currently selected
  int frame_nt = 1024;       // threads per work-group
  int frame_slots = 0;       // work-groups (= scratch slots)
  float2* slots = nullptr;   // frame_slots * M*K scratch
  size_t slots_cap = 0;      // slots allocated

  // sub-pixel positions: per-frame probes Pn[n] = P shifted by d_n (the
  // fractional part of frame n's position); Geom::Pn points at them while set
  double2* sub_d = nullptr;  // N shifts d_n (y, x)
  float2* Pn = nullptr;      // N*M*K per-frame probes
  bool sub_set = false;      // shifts given: runs use the per-frame probes
  bool Pn_valid = false;     // Pn holds the shifts of the bound P
  float2* Dy = nullptr;      // N*M*K position-derivative stacks / their spectra
  float2* Dx = nullptr;
  double* pos_terms = nullptr;  // N x 6 per-frame position terms

  // DM with background (method 4) carried in Fourier space: F holds the
  // iterate's spectra FFT2(f), Xo the column-inverted spectra of the next
  // leading Pmod_7 of (f, b), partb2 that step's background sums
  float2* Xo = nullptr;
  double* partb2 = nullptr;
  bool f_spec = false;       // F holds spectra (else f in real space)
  bool dmf_lead = false;     // Xo and partb2 belong to the current (F, b)
  bool dmf_run = false;      // this run takes the Fourier-space DM step
  float dmf_alpha = 0.f;     // alpha of the leading step in Xo
  // DM method 1 on spectra: the metric of iteration i is the model error of
  // O_{i+1} P, formed in iteration i + 1's column pass (or by the run's tail)
  bool dm1_run = false;      // DM method 1 (or 3, see dm1_op) runs on spectra from iteration
                             // pmod_probe_until on
  bool dm1_op = false;       // ... method 3, object + probe every iteration
  bool dm1_pending = false;  // Os awaits its model error; its metrics are not yet written
  int dm1_par = 0;           // which step-size partial buffer the last step wrote
  int dm1_nstep = 0;         // partials per step
  double* dm1_esup[2] = {nullptr, nullptr};

  // hipGraph replay of whole iterations
  struct GraphEntry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    HostIterState st;   // host state after the captured iteration
  };
  bool graphs = false;
  bool graphs_fell_back = false;     // a capture failed; iterations run eagerly since
  hipStream_t gstream = nullptr;     // default-priority stream used for capture / replay
  int* slot = nullptr;               // device metrics slot counter
  bool use_slot = false;             // k_metrics writes through the slot counter
  std::vector<GraphEntry> gcache;

  // DM bookkeeping
  bool dm_ready = false;
  // Os holds the next object-only DM projection of f (the last iteration's
  // metric projection, with the same f, P and heatmap); cleared by every run
  bool os_next = false;
  // Os, Ps hold the next object + probe projection (OP_iters (1, 1), no
  // centring, no Pmod_P: the metric projection with the iterate f the next
  // iteration projects, from the same O, P; its heatmaps were stored)
  bool os_next_op = false;
  bool have_metric_state = false;
  int last_algo = -1;

  // line-FFT plan of each kernel stage, a bit mask: bit PLAN_ROWS (row
  // kernels), PLAN_COLS (column kernels), PLAN_ROWG (fused row gather); a set
  // bit selects the side's alternative plan (sides without one keep the default)
  int plan = 0;

  // packed scans (pty_set_scans): scans of one frame size whose frames form
  // this context's stack in scan order and whose objects occupy consecutive
  // row bands of its object; every stage then runs once over all of them
  int nscan = 1;
  std::vector<int2> h_off;        // window offsets (host copy)
  std::vector<int> h_fstart, h_rstart;
  int* fscan = nullptr;           // [N] scan of each frame (Geom::fscan while packed)
  int* sc_fstart = nullptr;       // [S + 1]
  int* sc_rstart = nullptr;       // [S + 1]
  int* sc_pchunk = nullptr;       // probe chunks: first frame of each, [nchunks + 1]
  int* sc_pscan = nullptr;        // first probe chunk of each scan, [S + 1]
  int pk_chunks = 0;              // probe chunks over all scans
  double* sc_inorm = nullptr;     // [S] I_norm of each scan
  bool sc_norm_set = false;       // sc_inorm given (pty_set_scans) or computed (pty_prepare_amp)
  size_t cap_pchunks = 0;         // probe chunks pnum / pg hold
  size_t cap_G = 0;               // probes G holds (M K each)
  size_t cap_econP_all = 0;       // doubles part_econP holds
  size_t cap_esup_all = 0;        // doubles part_esup holds
  size_t cap_maskbits = 0;        // words maskbits holds
  size_t cap_maskbits_alt = 0;    // words maskbits_alt holds

  size_t MK() const { return (size_t)M * K; }
  size_t NMK() const { return (size_t)N * M * K; }
};

namespace {

// Brackets the launches of one stage with a pair of events when profiling.
struct StageTimer {
  pty_ctx* c;
  long idx = -1;
  StageTimer(pty_ctx* ctx, int stage) : c(ctx) {
    if (!c->prof_on) return;
    if (c->prof_override >= 0) {
      if (c->prof_override != stage && stage != PTY_STAGE_MISC) return;  // outer timer covers it
      stage = c->prof_override;
    }
    if (c->prof_used * 2 + 2 > c->prof_ev.size()) {
      for (int k = 0; k < 2; ++k) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        c->prof_ev.push_back(e);
      }
      c->prof_stage.push_back(stage);
    }
    idx = (long)c->prof_used++;
    c->prof_stage[idx] = stage;
    (void)hipEventRecord(c->prof_ev[2 * idx], c->stream);
  }
  ~StageTimer() {
    if (idx >= 0) (void)hipEventRecord(c->prof_ev[2 * idx + 1], c->stream);
  }
};

// after a stream sync: fold the recorded pairs into the per-stage totals
void prof_collect(pty_ctx* c) {
  for (size_t k = 0; k < c->prof_used; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->prof_ev[2 * k], c->prof_ev[2 * k + 1]) == hipSuccess) {
      c->prof_ms[c->prof_stage[k]] += ms;
      c->prof_cnt[c->prof_stage[k]] += 1;
    }
  }
  c->prof_used = 0;
}

// --------------------------------------------------------------------------
// launch helpers
// --------------------------------------------------------------------------

// plan key of a kernel stage (PLAN_ROWS / PLAN_COLS / PLAN_ROWG) for a frame side L
int pk(const pty_ctx* c, int L, int stage) {
  const int alt = plan_key(L, 1);
  return ((c->plan & stage) && threads_per_line(alt) > 0) ? alt : L;
}

// an iteration's metrics (3 doubles per scan) in the run's device buffer
double* metrics_at(pty_ctx* c, int64_t k) { return c->metrics + 3 * (size_t)c->nscan * k; }

// grid of a persistent launch: all items if few, else `cap` rounded down to a
// multiple of the XCD count (so the XCD-aware renumbering applies)
int persistent_grid(long items, long cap) {
  if (items <= cap) return (int)items;
  return (int)std::max<long>(NXCD, cap / NXCD * NXCD);
}

// static LDS of k_rows_fwd_orow<PK>: line exchange areas, twiddles, list chunk
template <int PK>
constexpr size_t rowf_static_lds() {
  return sizeof(float2) * ((WG / Plan<PK>::T) * LineLds<PK>::words + plan_len(PK)) +
         sizeof(int2) * ROWF_ENT;
}
size_t rowf_static(int PK) {
  switch (PK) {
#define C(Lv) case Lv: return rowf_static_lds<Lv>();
    PTY_PLANS(C)
#undef C
    default: return 0;
  }
}
// The object-row forward stage's dynamic LDS (staged object columns) never
// exceeds this.
constexpr size_t ROWF_DYN_MAX = 64u << 10;

// Column segments of the object-row forward stage for plan key PK: the
// fewest whose staged columns (segment + L - 1) let three work-groups share
// a CU's 160 KB of LDS, each segment at least L wide; 0 = not supported.
int rowf_segments(int PK, int L, int V) {
  const size_t st = rowf_static(PK);
  if (st == 0 || V > ROWG_VMAX) return 0;
  const size_t budget = (size_t)(160u << 10) / 3;
  if (budget < st + 2 * (size_t)L * sizeof(float2)) return 0;
  const size_t cols = std::min((budget - st) / sizeof(float2), ROWF_DYN_MAX / sizeof(float2));
  const size_t segw = cols - (L - 1);
  return (int)((V + segw - 1) / segw);
}

int rowf_attribute() {
  hipError_t e = hipSuccess;
#define A(Lv, SRC)                                                                            \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rows_fwd_orow<Lv, SRC>),         \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)ROWF_DYN_MAX);
#define C(Lv) A(Lv, SRC_OP) A(Lv, SRC_DM) A(Lv, SRC_HIO)
  PTY_PLANS(C)
#undef C
#undef A
  if (e != hipSuccess) return fail(PTY_EHIP, std::string("k_rows_fwd_orow LDS attribute: ") + hipGetErrorString(e));
  return PTY_OK;
}

// The forward row stage of a source that reads the object, over the whole
// stack, in object-row order (k_rows_fwd_orow); false when the lists do not
// apply to the current plan
template <int SRC>
bool rows_fwd_orow(pty_ctx* c, const float2* O, const float2* P, const float2* F, float2* X, float w) {
  const int PK = pk(c, c->K, PLAN_ROWS);
  const size_t dyn = (size_t)std::min(c->V, c->rowf_segw + c->K - 1) * sizeof(float2);
  if (dyn > ROWF_DYN_MAX || rowf_static(PK) + dyn > (size_t)(160u << 10)) return false;
  RowFwdList a{c->rowf_segs, c->rowf_segw, c->rowf_ptr, c->rowf_ent};
  const int grid = c->U * c->rowf_segs;
  switch (PK) {
#define C(Lv)                                                                                \
  case Lv:                                                                                   \
    hipLaunchKernelGGL((k_rows_fwd_orow<Lv, SRC>), dim3(grid), dim3(WG), dyn, c->stream, c->g, O, P, \
                       F, X, a, (const float2*)c->twK, w);                                   \
    break;
    PTY_PLANS(C)
#undef C
    default: return false;
  }
  return true;
}

template <int SRC>
int rows_fwd(pty_ctx* c, const float2* O, const float2* P, const float2* F, float2* X, int nrows,
             float w = 0.f) {
  StageTimer tm(c, PTY_STAGE_ROWS_FWD);
  if constexpr (SRC != SRC_F) {
    if (c->rowf && nrows == c->N * c->M && rows_fwd_orow<SRC>(c, O, P, F, X, w)) {
      PTY_LAUNCHED("k_rows_fwd_orow");
      return PTY_OK;
    }
  }
  const int T = threads_per_line(pk(c, c->K, PLAN_ROWS)), lpw = WG / T;
  dim3 grid(persistent_grid((nrows + lpw - 1) / lpw, (long)c->rows_mult * c->cus));
  switch (pk(c, c->K, PLAN_ROWS)) {
#define C(Lv)                                                                              \
  case Lv:                                                                                 \
    hipLaunchKernelGGL((k_rows_fwd<Lv, SRC>), grid, dim3(WG), 0, c->stream, c->g, O, P, F, X, \
                       (const float2*)c->twK, nrows, w);                                   \
    break;
    PTY_PLANS(C)
#undef C
    default: return fail(PTY_EUNSUP, "unsupported K");
  }
  PTY_LAUNCHED("k_rows_fwd");
  return PTY_OK;
}

template <int DST>
int rows_inv(pty_ctx* c, const float2* Y, float2* Out, const float2* O, const float2* P,
             float scale, double* partial, int nrows, int* nparts = nullptr, HioMix hm = {}) {
  StageTimer tm(c, PTY_STAGE_ROWS_INV);
  const int T = threads_per_line(pk(c, c->K, PLAN_ROWS)), lpw = WG / T;
  dim3 grid(persistent_grid((nrows + lpw - 1) / lpw, (long)c->rows_mult * c->cus));
  if (nparts) *nparts = grid.x;
  switch (pk(c, c->K, PLAN_ROWS)) {
#define C(Lv)                                                                               \
  case Lv:                                                                                  \
    hipLaunchKernelGGL((k_rows_inv<Lv, DST>), grid, dim3(WG), 0, c->stream, c->g, Y, Out, O, P, \
                       scale, partial, (const float2*)c->twK, nrows, hm);                   \
    break;
    PTY_PLANS(C)
#undef C
    default: return fail(PTY_EUNSUP, "unsupported K");
  }
  PTY_LAUNCHED("k_rows_inv");
  return PTY_OK;
}

int row_blocks(pty_ctx* c, int nrows) {
  const int lpw = WG / threads_per_line(pk(c, c->K, PLAN_ROWS));
  return (nrows + lpw - 1) / lpw;
}

ColArgs col_args(pty_ctx* c) {
  ColArgs a{};
  a.mask = c->mask;
  a.maskbits = c->mask ? c->maskbits : nullptr;
  a.scale = 1.0f / (float)c->MK();
  a.tw = c->twM;
  a.nframes = c->N;
  a.chunk = c->col_chunk;
  return a;
}

// persistent launch over (column block, frame chunk) items; returns the
// number of per-work-group partials written (= grid size)
template <int MODE>
int cols(pty_ctx* c, ColArgs a, int* nparts = nullptr) {
  StageTimer tm(c, PTY_STAGE_COLS);
  // frame chunks only matter where per-chunk background sums are accumulated
  if (!ColTraits<MODE>::bsum) a.chunk = 1;
  const int chunks = (a.nframes + a.chunk - 1) / a.chunk;
  int wpe = 3;
  bool hoist = false;
  // the background model-error pass (COL_ERR7) has a plan bit of its own
  const int pkm = pk(c, c->M, MODE == COL_ERR7 ? PLAN_BGERR : PLAN_COLS);
  switch (pkm) {
#define C(Lv)                              \
  case Lv:                                 \
    wpe = ColShape<Lv, MODE>::wpe;         \
    hoist = ColShape<Lv, MODE>::hoist;     \
    break;
    PTY_PLANS(C)
#undef C
    default: return fail(PTY_EUNSUP, "unsupported M");
  }
  // the packed mask in this plan's register layout
  if (a.maskbits) a.maskbits = pkm == c->M ? c->maskbits : c->maskbits_alt;
  const bool wide = c->M == 256 && (c->cols_wg == 512 || (c->cols_wg == 0 && hoist));
  const int cpw = (wide ? 512 : WG) / threads_per_line(pkm);
  const long items = (long)((c->K + cpw - 1) / cpw) * chunks;
  // default grid: Pmod_1 modes without the register prefetch take one item per
  // work-group (up to 128 per CU; measured at 4096 x 256^2: 1.04 ms against
  // 1.17 ms for a persistent grid of 2 per CU, intermediate sizes in between);
  // the others as many work-groups per CU as the register budget keeps resident
  const int mult = c->cols_mult ? c->cols_mult : hoist ? 128 : (wide ? std::max(1, wpe / 2) : wpe);
  int grid;
  if (c->nscan > 1 && ColTraits<MODE>::err) {
    // packed scans: one item per work-group and one partial per item, so each
    // scan's partials are one contiguous range
    if (items > (long)c->cap_col * MAX_CHUNKS) return fail(PTY_EINVAL, "column partials exceed capacity");
    grid = (int)items;
    a.item_parts = 1;
  } else {
    grid = persistent_grid(items, (long)mult * c->cus);
    if (grid > c->cap_col) return fail(PTY_EINVAL, "column partials exceed capacity");
  }
  if (wide) {
    if (pkm == 256)
      hipLaunchKernelGGL((k_cols<256, MODE, 512>), dim3(grid), dim3(512), 0, c->stream, c->g, a);
    else
      hipLaunchKernelGGL((k_cols<plan_key(256, 1), MODE, 512>), dim3(grid), dim3(512), 0, c->stream,
                         c->g, a);
  } else {
    switch (pkm) {
#define C(Lv)                                                                                  \
  case Lv:                                                                                     \
    hipLaunchKernelGGL((k_cols<Lv, MODE>), dim3(grid), dim3(WG), 0, c->stream, c->g, a);       \
    break;
      PTY_PLANS(C)
#undef C
      default: return fail(PTY_EUNSUP, "unsupported M");
    }
  }
  PTY_LAUNCHED("k_cols");
  if (nparts) *nparts = grid;
  return PTY_OK;
}

bool frame_supported(int M, int K) { return M == K && (M == 128 || M == 256); }

int ensure_slots(pty_ctx* c) {
  const size_t need = (size_t)c->frame_slots * c->MK();
  if (c->slots_cap >= need) return PTY_OK;
  dev_free(c->slots);
  c->slots_cap = 0;
  PTY_TRY(dev_alloc(&c->slots, need));
  c->slots_cap = need;
  return PTY_OK;
}

FrameArgs frame_args(pty_ctx* c) {
  FrameArgs a{};
  a.slots = c->slots;
  a.amp = c->amp;
  a.mask = c->mask;
  a.scale = 1.0f / (float)c->MK();
  a.tw = c->twK;
  return a;
}

// one launch of the fused per-frame kernel over all frames; returns the
// number of per-work-group partials it wrote
template <int MODE>
int frame(pty_ctx* c, FrameArgs a, int* nparts = nullptr) {
  StageTimer tm(c, PTY_STAGE_COLS);
  PTY_TRY(ensure_slots(c));
  a.slots = c->slots;
  const int grid = std::min(c->frame_slots, c->N);
  const int key = c->M * 10000 + c->frame_nt;
  switch (key) {
#define C(Lv, NTv)                                                                           \
  case Lv * 10000 + NTv:                                                                     \
    hipLaunchKernelGGL((k_frame<Lv, NTv, MODE>), dim3(grid), dim3(NTv), 0, c->stream, c->g, a); \
    break;
    C(128, 512) C(128, 1024) C(256, 512) C(256, 1024)
#undef C
    default: return fail(PTY_EUNSUP, "fused frame path: unsupported size / threads");
  }
  PTY_LAUNCHED("k_frame");
  if (nparts) *nparts = grid;
  return PTY_OK;
}

int gather(pty_ctx* c, const float2* S, const float2* P, const float2* Oold, float2* Oout,
           double* H, int fresh, int storeH, float alpha, bool econ, bool esup) {
  StageTimer tm(c, PTY_STAGE_GATHER);
  GatherArgs a{};
  a.S = S;
  a.P = P;
  a.Oold = econ ? Oold : nullptr;
  a.Oout = Oout;
  a.H = H;
  a.fresh = fresh;
  a.storeH = storeH;
  a.alpha = alpha;
  a.tile_ptr = c->tile_ptr;
  a.tile_frames = c->tile_frames;
  a.tiles_x = c->tiles_x;
  a.ntiles = c->ntiles;
  a.part_econ = econ ? c->part_econO : nullptr;
  a.part_esup = esup ? c->part_esxh : nullptr;
  const int grid = (c->ntiles + NXCD - 1) / NXCD * NXCD;
  hipLaunchKernelGGL(k_gather_O, dim3(grid), dim3(WG), 0, c->stream, c->g, a);
  PTY_LAUNCHED("k_gather_O");
  return PTY_OK;
}

// frame sides the fused inverse-row / object-update kernel is built for, and
// all their plans
#define PTY_ROWG_SIZES(X) X(64) X(96) X(128) X(192) X(256)
#define PTY_ROWG_PLANS(X) PTY_ROWG_SIZES(X) X(plan_key(64, 1)) X(plan_key(96, 1)) \
  X(plan_key(128, 1)) X(plan_key(192, 1)) X(plan_key(256, 1))

template <int PK>
constexpr size_t rowg_static_lds() {
  constexpr size_t lpw = WG / Plan<PK>::T;
  return sizeof(float2) * (lpw * LineLds<PK>::words + 1 + plan_len(PK)) + 16 * sizeof(double) +
         lpw * sizeof(int);
}
// object rows are cut into segments for the default plan's static LDS; the
// alternatives need no more, so the same segments fit
#define C(Lv) static_assert(rowg_static_lds<plan_key(Lv, 1)>() <= rowg_static_lds<Lv>(), "rowg LDS");
PTY_ROWG_SIZES(C)
#undef C

// Dynamic LDS (the row-segment accumulator) never exceeds this: rowg_segments
// refuses layouts that would need more.
constexpr size_t ROWG_DYN_MAX = 64u << 10;

// Work-groups of the fused kernel of plan key PK that a CU's registers hold
// (a 256-thread work-group is one wave per SIMD; 512 VGPRs per SIMD lane,
// allocated in blocks of 8), at most 8.
int rowg_vgpr_per_cu(int PK) {
  hipFuncAttributes fa{};
  hipError_t e = hipErrorInvalidValue;
  switch (PK) {
#define C(Lv) case Lv: e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_rows_gather<Lv>)); break;
    PTY_ROWG_PLANS(C)
#undef C
    default: break;
  }
  if (e != hipSuccess || fa.numRegs <= 0) return 3;
  return std::max(1, std::min(8, 512 / ((fa.numRegs + 7) / 8 * 8)));
}

// Column segments per object row for the fused kernel of plan key PK: the
// fewest that let as many work-groups as the registers allow (else fewer)
// share a CU's 160 KB of LDS, each segment at least 512 columns wide; 0 =
// geometry not supported.  Measured at 4096 x 256^2 with the alternative
// plan (five work-groups per CU by registers): two segments per object row
// 0.653 ms, one segment (three per CU by LDS) 0.736 ms.  PTY_ROWG_SEGS
// forces a count (1..4).
int rowg_segments(int PK, int V) {
  size_t st = 0;
  switch (PK) {
#define C(Lv) case Lv: st = rowg_static_lds<Lv>(); break;
    PTY_ROWG_PLANS(C)
#undef C
    default: return 0;
  }
  if (V > ROWG_VMAX) return 0;
  if (const char* e_s = getenv("PTY_ROWG_SEGS")) {
    const int f = atoi(e_s);
    if (f >= 1 && f <= 4 && (V + f - 1) / f * sizeof(float2) <= ROWG_DYN_MAX &&
        st + (V + f - 1) / f * sizeof(float2) <= (size_t)(160u << 10))
      return f;
  }
  for (int per_cu = rowg_vgpr_per_cu(PK); per_cu >= 1; --per_cu) {
    const size_t budget = (size_t)(160u << 10) / per_cu;
    if (budget < st + 512 * sizeof(float2)) continue;
    const size_t w = std::min((budget - st) / sizeof(float2), ROWG_DYN_MAX / sizeof(float2));
    const int segs = (int)((V + w - 1) / w);
    if (segs <= 4) return segs;
  }
  return 0;
}

// Set the dynamic-LDS limit of the K-sized instantiation.  The attribute
// belongs to the kernel, not to a context: one constant that covers every
// segment width keeps contexts with different objects from lowering each
// other's limit.
int rowg_attribute(int K) {
  hipError_t e = hipErrorInvalidValue;
  switch (K) {
#define C(Lv)                                                                                 \
  case Lv:                                                                                    \
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rows_gather<Lv>),                 \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)ROWG_DYN_MAX);    \
    if (e == hipSuccess)                                                                      \
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rows_gather<plan_key(Lv, 1)>),  \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)ROWG_DYN_MAX);  \
    break;
    PTY_ROWG_SIZES(C)
#undef C
    default: break;
  }
  if (e != hipSuccess) return fail(PTY_EHIP, std::string("k_rows_gather LDS attribute: ") + hipGetErrorString(e));
  return PTY_OK;
}

// Inverse row FFT of Y fused with the object update O <- sum conj(P) psi' / (H + alpha)
// (H valid); one work-group per object row segment, partials per segment.  Fout, when
// given, also receives the inverse-transformed rows (what rows_inv<DST_STORE> writes).
int rows_gather(pty_ctx* c, const float2* Y, float alpha, bool econ, bool esup, int* nparts,
                float2* Fout = nullptr, float2* Oout = nullptr) {
  StageTimer tm(c, PTY_STAGE_GATHER);
  RowGatherArgs a{};
  a.Y = Y;
  a.P = c->P;
  a.Oout = Oout ? Oout : c->O;
  a.Oold = econ ? a.Oout : nullptr;
  a.H = c->H;
  a.alpha = alpha;
  a.segs = c->rowg_segs;
  a.segw = c->rowg_segw;
  a.row_ptr = c->row_ptr;
  a.row_ent = c->row_ent;
  a.part_econ = econ ? c->part_econO : nullptr;
  a.part_esup = esup ? c->part_esxh : nullptr;
  a.Fout = Fout;
  const size_t dyn = (size_t)c->rowg_segw * sizeof(float2);
  const int grid = c->U * c->rowg_segs;
  switch (pk(c, c->K, PLAN_ROWG)) {
#define C(Lv)                                                                                \
  case Lv:                                                                                   \
    hipLaunchKernelGGL(k_rows_gather<Lv>, dim3(grid), dim3(WG), dyn, c->stream, c->g, a,     \
                       (const float2*)c->twK);                                               \
    break;
    PTY_ROWG_PLANS(C)
#undef C
    default: return fail(PTY_EUNSUP, "fused row gather: unsupported K");
  }
  PTY_LAUNCHED("k_rows_gather");
  if (nparts) *nparts = grid;
  return PTY_OK;
}

int probe_update(pty_ctx* c, const float2* S, const float2* O, const float2* Pold, float2* Pout,
                 double* G, int fresh, int storeG, float alpha, bool econ, bool esup = false) {
  StageTimer tm(c, PTY_STAGE_PROBE);
  const int MK = (int)c->MK();
  ProbeArgs a{};
  a.S = S;
  a.O = O;
  a.pnum = c->pnum;
  a.pg = c->pg;
  a.fresh = fresh;
  a.chunk = c->probe_chunk;
  const bool pk = c->nscan > 1;   // packed scans: per-scan chunk tables, one probe per scan
  if (pk) a.cstart = c->sc_pchunk;
  dim3 g1((MK + WG - 1) / WG, pk ? c->pk_chunks : c->probe_chunks);
  hipLaunchKernelGGL(k_probe_partial, g1, dim3(WG), 0, c->stream, c->g, a);
  PTY_LAUNCHED("k_probe_partial");
  ProbeFinalArgs f{};
  f.pnum = c->pnum;
  f.pg = c->pg;
  f.nchunks = c->probe_chunks;
  f.fresh = fresh;
  f.storeG = storeG;
  f.G = G;
  f.alpha = alpha;
  f.Pold = econ ? Pold : nullptr;
  f.Pout = Pout;
  f.part_econ = econ ? c->part_econP : nullptr;
  f.part_esup = esup ? c->part_esup : nullptr;
  if (pk) f.sc = c->sc_pscan;
  hipLaunchKernelGGL(k_probe_final, dim3((MK + WG - 1) / WG, pk ? c->nscan : 1), dim3(WG), 0,
                     c->stream, c->g, f);
  PTY_LAUNCHED("k_probe_final");
  return PTY_OK;
}

int copy_d2d(pty_ctx* c, void* dst, const void* src, size_t bytes) {
  PTY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return PTY_OK;
}


// P <- Pmod_P(Pamp, P, mask, alpha)   (era.py:597-610)
int pmod_probe(pty_ctx* c, float2* Pbuf, float alpha) {
  StageTimer tm(c, PTY_STAGE_MISC);
  const int saved = c->prof_override;
  c->prof_override = PTY_STAGE_MISC;
  struct Restore { pty_ctx* c; int v; ~Restore() { c->prof_override = v; } } restore{c, saved};
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, Pbuf, c->Xp, c->M));
  ColArgs a = col_args(c);
  a.X = c->Xp;
  a.amp = c->pamp;
  a.amp_stride = 0;
  a.alpha = alpha;
  a.nframes = 1;
  a.chunk = 1;
  PTY_TRY(cols<COL_PMODP>(c, a));
  PTY_TRY(rows_inv<DST_STORE>(c, c->Xp, Pbuf, nullptr, nullptr, 1.0f, nullptr, c->M));
  return PTY_OK;
}

// roll P and O (and, for DM, every frame of f) by -cm, cm from |P|^2
int centre_probe(pty_ctx* c, float2* Pbuf, float2* Obuf, bool roll_f) {
  StageTimer tm(c, PTY_STAGE_MISC);
  hipLaunchKernelGGL(k_com, dim3(1), dim3(WG), 0, c->stream, (const float2*)Pbuf, c->M, c->K, c->cm);
  PTY_LAUNCHED("k_com");
  const int nb = 1024;
  hipLaunchKernelGGL(k_roll, dim3(nb), dim3(WG), 0, c->stream, (const float2*)Pbuf, c->Xp, 1, c->M,
                     c->K, (const int*)c->cm);
  PTY_LAUNCHED("k_roll");
  PTY_TRY(copy_d2d(c, Pbuf, c->Xp, c->MK() * sizeof(float2)));
  hipLaunchKernelGGL(k_roll, dim3(nb), dim3(WG), 0, c->stream, (const float2*)Obuf, c->Oroll, 1,
                     c->U, c->V, (const int*)c->cm);
  PTY_LAUNCHED("k_roll");
  PTY_TRY(copy_d2d(c, Obuf, c->Oroll, (size_t)c->U * c->V * sizeof(float2)));
  if (roll_f) {
    hipLaunchKernelGGL(k_roll, dim3(4096), dim3(WG), 0, c->stream, (const float2*)c->F, c->X, c->N,
                       c->M, c->K, (const int*)c->cm);
    PTY_LAUNCHED("k_roll");
    std::swap(c->F, c->X);
  }
  return PTY_OK;
}

int econ_pass(pty_ctx* c, const float2* cur, const float2* bak, size_t count, double* part) {
  StageTimer tm(c, PTY_STAGE_REDUCE);
  hipLaunchKernelGGL(k_econ, dim3(ECON_BLOCKS), dim3(WG), 0, c->stream, cur, bak, count, part);
  PTY_LAUNCHED("k_econ");
  return PTY_OK;
}

int metrics_launch(pty_ctx* c, const Regions& rg, int esup_mode, double* out) {
  StageTimer tm(c, PTY_STAGE_REDUCE);
  hipLaunchKernelGGL(k_metrics, dim3(1), dim3(1024), 0, c->stream, rg, c->I_norm, (double)c->MK(),
                     esup_mode, out, c->use_slot ? c->slot : nullptr);
  PTY_LAUNCHED("k_metrics");
  return PTY_OK;
}

int sum_region(pty_ctx* c, const double* p, int n, double* host_out) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, c->stream, p, n, c->scal);
  PTY_LAUNCHED("k_sum");
  PTY_HIP(hipMemcpyAsync(host_out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int bg_mean(pty_ctx* c, int nchunks, float* out, const double* partb = nullptr) {
  StageTimer tm(c, PTY_STAGE_REDUCE);
  const int MK = (int)c->MK();
  hipLaunchKernelGGL(k_bg_mean, dim3((MK + BG_PIX - 1) / BG_PIX), dim3(WG), 0, c->stream,
                     partb ? partb : (const double*)c->partb, nchunks, MK, c->N, out);
  PTY_LAUNCHED("k_bg_mean");
  return PTY_OK;
}

int ensure_F(pty_ctx* c) {
  if (!c->F) PTY_TRY(dev_alloc(&c->F, c->NMK()));
  return PTY_OK;
}

int check_bound(pty_ctx* c, bool need_bg) {
  if (!c->amp) return fail(PTY_ESTATE, "pty_set_data has not been called");
  if (!c->O || !c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
  if (need_bg && !c->bg_set) return fail(PTY_ESTATE, "methods 4-6 need pty_set_background");
  return PTY_OK;
}

// --------------------------------------------------------------------------
// one ERA iteration (era.py:152-215 and :223-291)
// --------------------------------------------------------------------------
// The staged modulus step (row FFTs -> column FFTs + modulus -> inverse row
// FFTs) can run over blocks of frames small enough that the frame spectra stay
// in the 256 MB Infinity Cache between its three kernels instead of going
// through HBM.  At most MAX_CHUNKS blocks.

int chunk_frames(const pty_ctx* c) {
  if (c->chunk_frames <= 0 || c->chunk_frames >= c->N) return c->N;
  int G = std::max(c->chunk_frames, (c->N + MAX_CHUNKS - 1) / MAX_CHUNKS);
  G = (G + c->col_chunk - 1) / c->col_chunk * c->col_chunk;   // keeps background sums aligned
  return std::min(G, c->N);
}

// the fused per-frame kernel reads the one shared probe: sub-pixel positions
// take the staged kernels
bool frame_path(const pty_ctx* c) { return c->use_frame && !c->g.Pn; }

// The entry points that take the probe as an argument apply that one probe
// to every frame, whatever sub-pixel shifts the context holds.
struct SharedProbe {
  pty_ctx* c;
  const float2* saved;
  explicit SharedProbe(pty_ctx* c_) : c(c_), saved(c_->g.Pn) { c->g.Pn = nullptr; }
  ~SharedProbe() { c->g.Pn = saved; }
};

// frames [n0, n0 + n) seen as a scan of their own by the row / column kernels
struct ChunkView {
  pty_ctx* c;
  Geom saved;
  ChunkView(pty_ctx* c_, int n0, int n) : c(c_), saved(c_->g) {
    c->g.off = saved.off + n0;
    if (saved.Pn) c->g.Pn = saved.Pn + (size_t)n0 * c->MK();
    if (saved.fscan) c->g.fscan = saved.fscan + n0;
    c->g.N = n;
  }
  ~ChunkView() { c->g = saved; }
};

int era_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out) {
  const bool bgm = o.method >= 4;
  const int upd = (o.method - 1) % 3;  // 0 = O, 1 = P, 2 = OP
  const bool op_step = upd == 2 && (it % o.op_every) == 0;
  const bool pmodp = it < o.pmod_probe_until;
  const int nrows = c->N * c->M;
  const size_t MK = c->MK(), UV = (size_t)c->U * c->V;
  const bool fusedO = upd == 0;
  const bool fusedP = upd == 1 && !pmodp;
  // object + probe step that writes O and P exactly once (one O/P pass, no
  // centring, no far-field probe constraint): the updates read the previous
  // O / P before overwriting them and emit the eCon partials themselves,
  // instead of backup copies and two comparison passes
  const bool fusedOP = op_step && o.op_n == 1 && !o.probe_centering && !pmodp;
  // eSup from the identity: object-only steps take its terms from the object
  // update, object + probe steps whose O and P are final after the probe
  // update (no centring, no far-field constraint) from the probe update
  const bool esup_probe = op_step && !o.probe_centering && !pmodp;
  const bool esup_id = (upd == 0 && !pmodp) || esup_probe;
  const bool staged = !(frame_path(c) && !bgm);
  // object-only step with a valid heatmap: inverse row FFT fused into the update
  const bool rowg = upd == 0 && staged && c->rowg && c->H_valid;
  // single-pass object + probe step: the heatmap of the current probe in a
  // pass of its own, then the inverse row FFT fused into the object update,
  // which also stores psi' (into F) for the probe update
  const bool rowg_op = fusedOP && staged && c->rowg && c->F;

  if (!fusedO && !fusedP && !fusedOP) {
    StageTimer tm(c, PTY_STAGE_MISC);
    if (upd != 1) PTY_TRY(copy_d2d(c, c->Obak, c->O, UV * sizeof(float2)));
    if (upd != 0) PTY_TRY(copy_d2d(c, c->Pbak, c->P, MK * sizeof(float2)));
  }

  // modulus projection: psi' = IFFT(Pmod(FFT(O_n P)))
  int ncol = 0;
  if (!staged) {
    FrameArgs f = frame_args(c);
    f.O = c->O;
    f.P = c->P;
    f.Xout = c->X;
    f.alpha = o.alpha;
    f.part = c->part_emod;
    f.part2 = esup_id ? c->part_esa : nullptr;
    PTY_TRY(frame<FR_ERA1>(c, f, &ncol));
  } else {
    (void)nrows;
    const int G = chunk_frames(c);
    for (int n0 = 0; n0 < c->N; n0 += G) {
      const int nf = std::min(G, c->N - n0);
      ChunkView view(c, n0, nf);
      float2* Xc = c->X + (size_t)n0 * MK;
      PTY_TRY(rows_fwd<SRC_OP>(c, c->O, c->P, nullptr, Xc, nf * c->M));
      ColArgs a = col_args(c);
      a.X = Xc;
      a.nframes = nf;
      a.amp = c->amp + (size_t)n0 * MK;
      a.amp_stride = (long)MK;
      a.alpha = o.alpha;
      a.part = c->part_emod + ncol;
      a.part2 = esup_id ? c->part_esa + ncol : nullptr;
      int nc = 0;
      if (bgm) {
        a.bg = c->era_bg_shared ? c->bbar : c->bg + (size_t)n0 * MK;
        a.bg_stride = c->era_bg_shared ? 0 : (long)MK;
        a.partb = c->partb + (size_t)(n0 / c->col_chunk) * MK;
        PTY_TRY(cols<COL_ERAB>(c, a, &nc));
      } else {
        PTY_TRY(cols<COL_ERA1>(c, a, &nc));
      }
      ncol += nc;
      if (!rowg && !rowg_op)
        PTY_TRY(rows_inv<DST_STORE>(c, Xc, Xc, nullptr, nullptr, 1.0f, nullptr, nf * c->M));
    }
  }

  // consistency projection
  int ngather = c->ntiles;
  if (rowg) {
    PTY_TRY(rows_gather(c, c->X, o.alpha, true, esup_id, &ngather));
  } else if (upd == 0) {
    PTY_TRY(gather(c, c->X, c->P, c->O, c->O, c->H, !c->H_valid, 1, o.alpha, true, esup_id));
    c->H_valid = true;
  } else if (upd == 1) {
    PTY_TRY(probe_update(c, c->X, c->O, c->P, c->P, c->G, !c->G_valid, 1, o.alpha, fusedP));
    c->G_valid = true;
  } else if (rowg_op) {
    c->dm_ready = false;   // F now holds psi', not a DM iterate: a DM run must begin again
    c->f_spec = c->dmf_lead = false;
    PTY_TRY(gather(c, nullptr, c->P, nullptr, nullptr, c->H, 1, 1, o.alpha, false, false));
    c->H_valid = true;
    PTY_TRY(rows_gather(c, c->X, o.alpha, true, false, &ngather, c->F));
    PTY_TRY(probe_update(c, c->F, c->O, c->P, c->P, c->G, 1, 1, o.alpha, true, esup_probe));
    c->G_valid = true;
  } else if (fusedOP) {
    PTY_TRY(gather(c, c->X, c->P, c->O, c->O, c->H, 1, 1, o.alpha, true, false));
    PTY_TRY(probe_update(c, c->X, c->O, c->P, c->P, c->G, 1, 1, o.alpha, true, esup_probe));
    c->H_valid = c->G_valid = true;
  } else if (op_step) {
    for (int r = 0; r < o.op_n; ++r) {
      PTY_TRY(gather(c, c->X, c->P, nullptr, c->O, c->H, 1, 1, o.alpha, false, false));
      PTY_TRY(probe_update(c, c->X, c->O, nullptr, c->P, c->G, 1, 1, o.alpha, false,
                           esup_probe && r == o.op_n - 1));
      c->H_valid = c->G_valid = true;
    }
    if (o.probe_centering) {
      PTY_TRY(centre_probe(c, c->P, c->O, false));
      c->H_valid = c->G_valid = false;
    }
  } else {
    PTY_TRY(gather(c, c->X, c->P, nullptr, c->O, c->H, !c->H_valid, 1, o.alpha, false, false));
    c->H_valid = true;
  }

  if (pmodp) PTY_TRY(pmod_probe(c, c->P, o.alpha));

  if (bgm) {
    PTY_TRY(bg_mean(c, (c->N + c->col_chunk - 1) / c->col_chunk, c->bbar));
    c->era_bg_shared = true;
  }

  // metrics (packed scans: `kind` says how each scan's partials lie in a region)
  Regions rg{};
  int kind[NSLOT] = {};
  const int okind = (rowg || rowg_op) ? RK_ROWS : RK_TILES;   // object-update partials
  rg.p[S_EMOD] = c->part_emod;
  rg.n[S_EMOD] = ncol;
  kind[S_EMOD] = RK_COLS;
  const int pb = (int)((MK + WG - 1) / WG);
  if (fusedO) {
    rg.p[S_ECON_ON] = c->part_econO;
    rg.p[S_ECON_OD] = c->part_econO + ngather;
    rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ngather;
    kind[S_ECON_ON] = kind[S_ECON_OD] = okind;
  } else if (fusedP) {
    rg.p[S_ECON_PN] = c->part_econP;
    rg.p[S_ECON_PD] = c->part_econP + pb;
    rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = pb;
    kind[S_ECON_PN] = kind[S_ECON_PD] = RK_PROBE;
  } else if (fusedOP) {
    rg.p[S_ECON_ON] = c->part_econO;
    rg.p[S_ECON_OD] = c->part_econO + ngather;
    rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ngather;
    kind[S_ECON_ON] = kind[S_ECON_OD] = okind;
    rg.p[S_ECON_PN] = c->part_econP;
    rg.p[S_ECON_PD] = c->part_econP + pb;
    rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = pb;
    kind[S_ECON_PN] = kind[S_ECON_PD] = RK_PROBE;
  } else {
    if (c->nscan > 1) return fail(PTY_EUNSUP, "packed scans: unfused update step");
    if (upd != 1) {
      PTY_TRY(econ_pass(c, c->O, c->Obak, UV, c->part_econO));
      rg.p[S_ECON_ON] = c->part_econO;
      rg.p[S_ECON_OD] = c->part_econO + ECON_BLOCKS;
      rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ECON_BLOCKS;
    }
    if (upd != 0) {
      PTY_TRY(econ_pass(c, c->P, c->Pbak, MK, c->part_econP));
      rg.p[S_ECON_PN] = c->part_econP;
      rg.p[S_ECON_PD] = c->part_econP + ECON_BLOCKS;
      rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = ECON_BLOCKS;
    }
  }
  int esup_mode;
  if (esup_id) {
    esup_mode = ESUP_IDENTITY;
    rg.p[S_ESUP_A] = c->part_esa;
    rg.n[S_ESUP_A] = ncol;
    kind[S_ESUP_A] = RK_COLS;
    if (esup_probe) {
      rg.p[S_ESUP_X] = c->part_esup;
      rg.p[S_ESUP_H] = c->part_esup + pb;
      rg.n[S_ESUP_X] = rg.n[S_ESUP_H] = pb;
      kind[S_ESUP_X] = kind[S_ESUP_H] = RK_PROBE;
    } else {
      rg.p[S_ESUP_X] = c->part_esxh;
      rg.p[S_ESUP_H] = c->part_esxh + ngather;
      rg.n[S_ESUP_X] = rg.n[S_ESUP_H] = ngather;
      kind[S_ESUP_X] = kind[S_ESUP_H] = okind;
    }
  } else {
    esup_mode = ESUP_EXPLICIT;
    StageTimer tm(c, PTY_STAGE_REDUCE);
    // a whole number of frame rows per work-group, every thread busy
    const int esup_threads = (WG / c->K) * c->K;
    if (c->nscan > 1) {
      // one pass per scan, each into its own partials
      for (int sc = 0; sc < c->nscan; ++sc) {
        const int f0 = c->h_fstart[sc], nf = c->h_fstart[sc + 1] - f0;
        ChunkView view(c, f0, nf);
        hipLaunchKernelGGL(k_esup, dim3(ESUP_BLOCKS), dim3(esup_threads), 0, c->stream, c->g,
                           (const float2*)(c->X + (size_t)f0 * MK), (const float2*)c->O,
                           (const float2*)c->P, c->part_esup + (size_t)sc * ESUP_BLOCKS);
        PTY_LAUNCHED("k_esup");
      }
    } else {
      hipLaunchKernelGGL(k_esup, dim3(ESUP_BLOCKS), dim3(esup_threads), 0, c->stream, c->g, (const float2*)c->X,
                         (const float2*)c->O, (const float2*)c->P, c->part_esup);
      PTY_LAUNCHED("k_esup");
    }
    rg.p[S_ESUP_E] = c->part_esup;
    rg.n[S_ESUP_E] = ESUP_BLOCKS;
    kind[S_ESUP_E] = RK_ESUP;
  }
  if (c->nscan > 1) {
    ScanRegions sr{};
    for (int k = 0; k < NSLOT; ++k) {
      sr.p[k] = rg.p[k];
      sr.kind[k] = kind[k];
    }
    sr.fstart = c->sc_fstart;
    sr.rstart = c->sc_rstart;
    sr.I_norm = c->sc_inorm;
    sr.S = c->nscan;
    sr.ncb = ncol / c->N;
    sr.tiles_x = c->tiles_x;
    sr.segs = c->rowg_segs;
    sr.pb = pb;
    sr.esup_blocks = ESUP_BLOCKS;
    StageTimer tm(c, PTY_STAGE_REDUCE);
    hipLaunchKernelGGL(k_metrics_scans, dim3(1), dim3(1024), 0, c->stream, sr, (double)MK, esup_mode,
                       metrics_out, c->use_slot ? c->slot : nullptr);
    PTY_LAUNCHED("k_metrics_scans");
    return PTY_OK;
  }
  return metrics_launch(c, rg, esup_mode, metrics_out);
}

// --------------------------------------------------------------------------
// one DM iteration (dm.py:45-160 and :169-294)
// --------------------------------------------------------------------------
int dm_project(pty_ctx* c, const pty_opts& o, int64_t it, int upd, bool op_step, float2* Ob,
               float2* Pb, bool metric, bool store = false) {
  if (upd == 0) {
    PTY_TRY(gather(c, c->F, Pb, nullptr, Ob, c->H, !c->H_valid, 1, o.alpha, false, false));
    c->H_valid = true;
  } else if (upd == 1) {
    PTY_TRY(probe_update(c, c->F, Ob, nullptr, Pb, c->G, !c->G_valid, 1, o.alpha, false));
    c->G_valid = true;
  } else if (op_step) {
    // (a metric projection that the next iteration reuses stores its heatmaps
    // as that iteration's own projection would)
    const int st = (metric && !store) ? 0 : 1;
    for (int r = 0; r < o.op_n; ++r) {
      PTY_TRY(gather(c, c->F, Pb, nullptr, Ob, c->H, 1, st, o.alpha, false, false));
      PTY_TRY(probe_update(c, c->F, Ob, nullptr, Pb, c->G, 1, st, o.alpha, false));
      if (st) c->H_valid = c->G_valid = true;
    }
    if (o.probe_centering) {
      PTY_TRY(centre_probe(c, Pb, Ob, !metric));
      if (!metric) c->H_valid = c->G_valid = false;
    }
  } else {
    // dm.py:85 / :130: project with the main-iterate probe and the cached heatmap
    PTY_TRY(gather(c, c->F, c->P, nullptr, Ob, c->H, !c->H_valid, 1, o.alpha, false, false));
    c->H_valid = true;
  }
  (void)it;
  return PTY_OK;
}

int dm_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out) {
  const bool bgm = o.method >= 4;
  const int upd = (o.method - 1) % 3;
  const bool op_step = upd == 2 && (it % o.op_every) == 0;
  const bool pmodp = it < o.pmod_probe_until;
  const int nrows = c->N * c->M;
  const size_t MK = c->MK(), UV = (size_t)c->U * c->V;
  const int nchunks = (c->N + c->col_chunk - 1) / c->col_chunk;
  // object-only projections of the freshly modulus-projected f (dm_project's
  // gather of f with P and the cached heatmap) fuse into its inverse row FFT
  const bool fused_proj = bgm && c->rowg && c->H_valid && (upd == 0 || (upd == 2 && !op_step));

  if (bgm) {
    // leading Pmod_7 of (f, b) (dm.py:171)
    PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->F, c->X, nrows));
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.alpha = o.alpha;
    a.bg = c->bg;
    a.bg_stride = (long)MK;
    a.partb = c->partb;
    PTY_TRY(cols<COL_DMB_LEAD>(c, a));
    if (fused_proj) {
      // the inverse row FFT writes f and projects it onto the object in one pass
      PTY_TRY(rows_gather(c, c->X, o.alpha, false, false, nullptr, c->F));
    } else {
      PTY_TRY(rows_inv<DST_STORE>(c, c->X, c->F, nullptr, nullptr, 1.0f, nullptr, nrows));
    }
    PTY_TRY(bg_mean(c, nchunks, c->bbar));
  }

  // object-only projections: upd 0, or upd 2 away from its OP iterations
  const bool plain_proj = upd == 0 || (upd == 2 && !op_step);
  // an object + probe projection that the next iteration would repeat: one
  // pass per iteration, no centring, no Pmod_P on this iteration's metric probe
  const bool op_reuse = !bgm && upd == 2 && o.op_every == 1 && o.op_n == 1 && !o.probe_centering;
  if (!bgm && plain_proj && c->os_next) {
    // dm.py:85 / :130: the projection of this f with P and the heatmap is the
    // one the previous iteration's metric step computed into Os (f, P and the
    // heatmap have not changed since)
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->O, c->Os, UV * sizeof(float2)));
  } else if (op_reuse && op_step && c->os_next_op) {
    // dm.py:65-82: the object + probe projection of this f from this O, P is
    // the previous iteration's metric projection (dm.py:104-130), heatmaps
    // included
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->O, c->Os, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->P, c->Ps, MK * sizeof(float2)));
  } else if (!fused_proj) {
    PTY_TRY(dm_project(c, o, it, upd, op_step, c->O, c->P, false));
  }
  if (pmodp) PTY_TRY(pmod_probe(c, c->P, o.alpha));

  // f <- f - Ps f + Pm(2 Ps f - f)    (dm.py:91-97, :216-224)
  int nstep = 0;
  if (frame_path(c) && !bgm) {
    FrameArgs f = frame_args(c);
    f.O = c->O;
    f.P = c->P;
    f.F = c->F;
    f.alpha = o.alpha;
    f.part2 = c->part_esup;
    PTY_TRY(frame<FR_DM1>(c, f, &nstep));
  } else {
  PTY_TRY(rows_fwd<SRC_DM>(c, c->O, c->P, c->F, c->X, nrows));
  {
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.alpha = o.alpha;
    if (bgm) {
      a.bg = c->bg;
      a.bg_stride = (long)MK;
      a.bbar = c->bbar;
      a.partb = c->partb;
      PTY_TRY(cols<COL_DMB_UPD>(c, a));
    } else {
      PTY_TRY(cols<COL_DM1>(c, a));
    }
  }
  PTY_TRY(rows_inv<DST_DMUPD>(c, c->X, c->F, c->O, c->P, 1.0f, c->part_esup, nrows, &nstep));
  }

  // metrics: Os, Ps = Psup(f)  (dm.py:104-137)
  {
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->Os, c->O, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->Ps, c->P, MK * sizeof(float2)));
  }
  PTY_TRY(dm_project(c, o, it, upd, op_step, c->Os, c->Ps, true, op_reuse && op_step && !pmodp));
  // a plain metric projection used P (Ps == P) and the heatmap; pmod_probe
  // below changes only Ps, and P changes only on OP iterations
  c->os_next = !bgm && plain_proj;
  c->os_next_op = op_reuse && op_step && !pmodp;
  if (pmodp) PTY_TRY(pmod_probe(c, c->Ps, o.alpha));
  if (bgm) PTY_TRY(bg_mean(c, nchunks, c->bbar_m));
  int ncol = 0;
  if (frame_path(c) && !bgm) {
    FrameArgs f = frame_args(c);
    f.O = c->Os;
    f.P = c->Ps;
    f.part = c->part_emod;
    PTY_TRY(frame<FR_ERR1>(c, f, &ncol));
  } else {
    PTY_TRY(rows_fwd<SRC_OP>(c, c->Os, c->Ps, nullptr, c->X, nrows));
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.part = c->part_emod;
    if (bgm) {
      a.bbar = c->bbar_m;
      PTY_TRY(cols<COL_ERR7>(c, a, &ncol));
    } else {
      PTY_TRY(cols<COL_ERR1>(c, a, &ncol));
    }
  }
  Regions rg{};
  rg.p[S_EMOD] = c->part_emod;
  rg.n[S_EMOD] = ncol;
  if (upd != 1) {
    PTY_TRY(econ_pass(c, c->Os, c->Obak, UV, c->part_econO));
    rg.p[S_ECON_ON] = c->part_econO;
    rg.p[S_ECON_OD] = c->part_econO + ECON_BLOCKS;
    rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ECON_BLOCKS;
  }
  if (upd != 0) {
    PTY_TRY(econ_pass(c, c->Ps, c->Pbak, MK, c->part_econP));
    rg.p[S_ECON_PN] = c->part_econP;
    rg.p[S_ECON_PD] = c->part_econP + ECON_BLOCKS;
    rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = ECON_BLOCKS;
  }
  rg.p[S_ESUP_E] = c->part_esup;
  rg.n[S_ESUP_E] = nstep;
  PTY_TRY(metrics_launch(c, rg, ESUP_EXPLICIT, metrics_out));
  {
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->Obak, c->Os, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->Pbak, c->Ps, MK * sizeof(float2)));
  }
  c->have_metric_state = true;
  return PTY_OK;
}

// One DM iteration with background refinement (method 4, fixed probe) on the
// iterate's spectra F (dm.py:169-294).  The reference's steps, with the
// leading Pmod_7 of iteration i + 1 folded into iteration i's column pass:
//   O  = Psup(f1)                 inverse row FFT of Xo fused into the gather
//   S  = FFT2(O P)                row FFT, then (column stage, per pixel)
//   (f1, b1) = Pmod_7(F, b)       recomputed from the state
//   y = 2 S - f1, b0 = 2 bbar - b1, (y', b0') = Pmod_7(y, b0)
//   F <- f1 - S + y', b <- b1 - bbar + b0'   (bbar = frame mean of b1)
//   X  = col-IFFT(F), Xo = col-IFFT(first half of Pmod_7(F, b))
//   Os = Psup(f) from X, eMod of Os P with the frame mean of b.
// Per frame pixel: 8 (Xo) + 8 + 52 (rows forward, column pass) + 8 (X) + 20
// (metric) = 96 bytes, against 156 for the same steps on f in real space.
int dm_bg_fourier_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out) {
  const size_t MK = c->MK(), UV = (size_t)c->U * c->V;
  const int nrows = c->N * c->M;
  const int nchunks = (c->N + c->col_chunk - 1) / c->col_chunk;
  (void)it;
  PTY_TRY(bg_mean(c, nchunks, c->bbar, c->partb2));   // frame mean of b1
  if (!c->H_valid) {
    PTY_TRY(gather(c, nullptr, c->P, nullptr, nullptr, c->H, 1, 1, o.alpha, false, false));
    c->H_valid = true;
  }
  PTY_TRY(rows_gather(c, c->Xo, o.alpha, false, false, nullptr));
  PTY_TRY(rows_fwd<SRC_OP>(c, c->O, c->P, nullptr, c->X, nrows));
  int nstep = 0;
  {
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.alpha = o.alpha;
    a.bg = c->bg;
    a.bg_stride = (long)MK;
    a.bbar = c->bbar;
    a.partb = c->partb;
    a.partb2 = c->partb2;
    a.F = c->F;
    a.Xo = c->Xo;
    a.part = c->part_esa;
    PTY_TRY(cols<COL_DMBF>(c, a, &nstep));
  }
  // metrics: Os = Psup(f), the frame mean of b, eMod of Os P (dm.py:231-271)
  PTY_TRY(bg_mean(c, nchunks, c->bbar_m, c->partb));
  PTY_TRY(rows_gather(c, c->X, o.alpha, false, false, nullptr, nullptr, c->Os));
  PTY_TRY(rows_fwd<SRC_OP>(c, c->Os, c->P, nullptr, c->X, nrows));
  int ncol = 0;
  {
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.part = c->part_emod;
    a.bbar = c->bbar_m;
    PTY_TRY(cols<COL_ERR7>(c, a, &ncol));
  }
  Regions rg{};
  rg.p[S_EMOD] = c->part_emod;
  rg.n[S_EMOD] = ncol;
  PTY_TRY(econ_pass(c, c->Os, c->Obak, UV, c->part_econO));
  rg.p[S_ECON_ON] = c->part_econO;
  rg.p[S_ECON_OD] = c->part_econO + ECON_BLOCKS;
  rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ECON_BLOCKS;
  rg.p[S_ESUP_E] = c->part_esa;
  rg.n[S_ESUP_E] = nstep;
  PTY_TRY(metrics_launch(c, rg, ESUP_EXPLICIT, metrics_out));
  {
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->Obak, c->Os, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->Ps, c->P, MK * sizeof(float2)));
  }
  c->have_metric_state = true;
  return PTY_OK;
}

// One DM method-1 iteration on the iterate's spectra F (dm.py:45-160):
//   O = Os of the previous iteration (= Psup(f), dm.py:85)
//   S = FFT2(O P): row FFT, then (column stage, per pixel) the previous
//     iteration's model error of S when it is pending, y = 2 S - f,
//     y' = Pmod_1(y), F <- f - S + y', step size sum |y' - S|^2 / (M K),
//     X = col-IFFT(F)
//   Os = Psup(f) from X (the metric object, dm.py:104-137), eCon against the
//     previous Os; iteration i's metrics are written once iteration i + 1
//     (or the run's tail, dm1_tail) has formed the model error of Os P.
// Per frame pixel: 8 + 36 (rows forward, column pass) + 8 (gather) = 52
// bytes, against 88 on f in real space.  Method 3 with an object + probe
// step every iteration (dm1_op): O, P = the previous Os, Ps (the metric
// projection is the next projection); the gather also writes f in real
// space (8) for the probe update with Os (8): 68 bytes against 96.
int dm1_fourier_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out) {
  const size_t MK = c->MK(), UV = (size_t)c->U * c->V;
  const int nrows = c->N * c->M;
  (void)it;
  const bool op = c->dm1_op;
  if (c->dm1_pending) {
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->O, c->Os, UV * sizeof(float2)));
    if (op) PTY_TRY(copy_d2d(c, c->P, c->Ps, MK * sizeof(float2)));
  }
  PTY_TRY(rows_fwd<SRC_OP>(c, c->O, c->P, nullptr, c->X, nrows));
  const int par = c->dm1_pending ? 1 - c->dm1_par : c->dm1_par;
  int ncol = 0;
  {
    ColArgs a = col_args(c);
    a.X = c->X;
    a.amp = c->amp;
    a.amp_stride = (long)MK;
    a.alpha = o.alpha;
    a.F = c->F;
    a.part = c->dm1_pending ? c->part_emod : nullptr;
    a.part2 = c->dm1_esup[par];
    PTY_TRY(cols<COL_DM1F>(c, a, &ncol));
  }
  if (c->dm1_pending) {
    // the previous iteration's metrics: eMod of its Os (= this O), its step
    // size and eCon
    Regions rg{};
    rg.p[S_EMOD] = c->part_emod;
    rg.n[S_EMOD] = ncol;
    rg.p[S_ECON_ON] = c->part_econO;
    rg.p[S_ECON_OD] = c->part_econO + ECON_BLOCKS;
    rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ECON_BLOCKS;
    rg.p[S_ESUP_E] = c->dm1_esup[c->dm1_par];
    rg.n[S_ESUP_E] = c->dm1_nstep;
    if (op) {
      rg.p[S_ECON_PN] = c->part_econP;
      rg.p[S_ECON_PD] = c->part_econP + ECON_BLOCKS;
      rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = ECON_BLOCKS;
    }
    PTY_TRY(metrics_launch(c, rg, ESUP_EXPLICIT, c->use_slot ? metrics_out : metrics_out - 3));
  }
  c->dm1_par = par;
  c->dm1_nstep = ncol;
  if (!c->H_valid || op) {
    // (object + probe: the heatmap of this iteration's probe, dm.py:104-130)
    PTY_TRY(gather(c, nullptr, c->P, nullptr, nullptr, c->H, 1, 1, o.alpha, false, false));
    c->H_valid = true;
  }
  // Os = Psup(f); object + probe: the gather also writes f in real space for
  // the probe update with Os, Ps (dm.py:104-130)
  PTY_TRY(rows_gather(c, c->X, o.alpha, false, false, nullptr, op ? c->Xo : nullptr, c->Os));
  if (op) {
    PTY_TRY(probe_update(c, c->Xo, c->Os, nullptr, c->Ps, c->G, 1, 1, o.alpha, false));
    c->G_valid = true;
  }
  PTY_TRY(econ_pass(c, c->Os, c->Obak, UV, c->part_econO));
  if (op) PTY_TRY(econ_pass(c, c->Ps, c->Pbak, MK, c->part_econP));
  {
    StageTimer tm(c, PTY_STAGE_MISC);
    PTY_TRY(copy_d2d(c, c->Obak, c->Os, UV * sizeof(float2)));
    if (op) PTY_TRY(copy_d2d(c, c->Pbak, c->Ps, MK * sizeof(float2)));
  }
  c->dm1_pending = true;
  c->have_metric_state = true;
  return PTY_OK;
}

// The last iteration's metrics: the model error of its Os P (dm.py:337-341).
int dm1_tail(pty_ctx* c, double* metrics_out) {
  if (!c->dm1_pending) return PTY_OK;
  const size_t MK = c->MK();
  PTY_TRY(rows_fwd<SRC_OP>(c, c->Os, c->dm1_op ? c->Ps : c->P, nullptr, c->X, c->N * c->M));
  int ncol = 0;
  ColArgs a = col_args(c);
  a.X = c->X;
  a.amp = c->amp;
  a.amp_stride = (long)MK;
  a.part = c->part_emod;
  PTY_TRY(cols<COL_ERR1>(c, a, &ncol));
  Regions rg{};
  rg.p[S_EMOD] = c->part_emod;
  rg.n[S_EMOD] = ncol;
  rg.p[S_ECON_ON] = c->part_econO;
  rg.p[S_ECON_OD] = c->part_econO + ECON_BLOCKS;
  rg.n[S_ECON_ON] = rg.n[S_ECON_OD] = ECON_BLOCKS;
  rg.p[S_ESUP_E] = c->dm1_esup[c->dm1_par];
  rg.n[S_ESUP_E] = c->dm1_nstep;
  if (c->dm1_op) {
    rg.p[S_ECON_PN] = c->part_econP;
    rg.p[S_ECON_PD] = c->part_econP + ECON_BLOCKS;
    rg.n[S_ECON_PN] = rg.n[S_ECON_PD] = ECON_BLOCKS;
  }
  PTY_TRY(metrics_launch(c, rg, ESUP_EXPLICIT, metrics_out));
  c->dm1_pending = false;
  return PTY_OK;
}

// Host decisions and buffer addresses that shape one iteration's launches.
std::vector<uint64_t> iter_key(pty_ctx* c, const pty_opts& o, int64_t it) {
  const int upd = (o.method - 1) % 3;
  auto bits = [](double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; };
  return {(uint64_t)o.algorithm, (uint64_t)o.method, (uint64_t)o.op_n,
          (uint64_t)(upd == 2 && (it % o.op_every) == 0), (uint64_t)(it < o.pmod_probe_until),
          (uint64_t)o.probe_centering, bits((double)o.alpha), bits(c->I_norm),
          (uint64_t)c->H_valid, (uint64_t)c->G_valid, (uint64_t)c->era_bg_shared,
          (uint64_t)c->os_next, (uint64_t)c->os_next_op,
          (uint64_t)c->use_frame, (uint64_t)c->frame_nt, (uint64_t)c->frame_slots,
          (uint64_t)c->rows_mult, (uint64_t)c->cols_mult, (uint64_t)c->cols_wg,
          (uint64_t)c->chunk_frames, (uint64_t)c->rowg,
          (uint64_t)(uintptr_t)c->O, (uint64_t)(uintptr_t)c->P, (uint64_t)(uintptr_t)c->F,
          (uint64_t)(uintptr_t)c->X, (uint64_t)(uintptr_t)c->amp, (uint64_t)(uintptr_t)c->mask,
          (uint64_t)(uintptr_t)c->bg, (uint64_t)(uintptr_t)c->metrics,
          (uint64_t)(uintptr_t)c->slots, (uint64_t)(uintptr_t)c->pamp,
          (uint64_t)(uintptr_t)c->g.Pn, (uint64_t)c->nscan, (uint64_t)c->plan, (uint64_t)c->dmf_run,
          (uint64_t)c->dm1_run,
          (uint64_t)c->dm1_pending, (uint64_t)c->dm1_par, (uint64_t)c->f_spec,
          (uint64_t)c->dm1_op};
}

HostIterState save_state(const pty_ctx* c) {
  return {c->H_valid, c->G_valid, c->era_bg_shared, c->have_metric_state, c->os_next,
          c->os_next_op, c->F, c->X,
          c->dm1_pending, c->dm1_par, c->f_spec};
}
void load_state(pty_ctx* c, const HostIterState& st) {
  c->H_valid = st.H_valid;
  c->G_valid = st.G_valid;
  c->era_bg_shared = st.era_bg_shared;
  c->have_metric_state = st.have_metric_state;
  c->os_next = st.os_next;
  c->os_next_op = st.os_next_op;
  c->F = st.F;
  c->X = st.X;
  c->dm1_pending = st.dm1_pending;
  c->dm1_par = st.dm1_par;
  c->f_spec = st.f_spec;
}

void drop_graphs(pty_ctx* c) {
  for (auto& e : c->gcache) (void)hipGraphExecDestroy(e.exec);
  c->gcache.clear();
}

// Host-side switch that treats every capture as failed (tests of the eager
// fallback): PTY_FORCE_CAPTURE_FAILURE=1.  Read on every capture.
bool capture_fail_forced() {
  const char* e = getenv("PTY_FORCE_CAPTURE_FAILURE");
  return e && atoi(e) == 1;
}

int run_one(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_dev);

// Iterations as replays of captured graphs, one graph per distinct iteration
// kind (cached with the host-state transitions the iteration implies).
int run_graphs(pty_ctx* c, const pty_opts& o, int64_t iter0, int32_t iters) {
  if (!c->gstream) PTY_HIP(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
  if (!c->slot) PTY_TRY(dev_alloc(&c->slot, 1));
  hipStream_t user = c->stream;
  PTY_HIP(hipStreamSynchronize(user));
  const int zero = 0;
  PTY_HIP(hipMemcpy(c->slot, &zero, sizeof(int), hipMemcpyHostToDevice));
  c->stream = c->gstream;
  c->use_slot = true;
  const bool prof = c->prof_on;
  c->prof_on = false;  // stage events are not recorded inside graphs
  struct Restore {
    pty_ctx* c; hipStream_t s; bool prof;
    ~Restore() { c->stream = s; c->use_slot = false; c->prof_on = prof; }
  } restore{c, user, prof};
  for (int32_t k = 0; k < iters; ++k) {
    const int64_t it = iter0 + k;
    if (!c->graphs) {   // a capture failed earlier in this run: continue eagerly
      PTY_TRY(run_one(c, o, it, metrics_at(c, k)));
      continue;
    }
    std::vector<uint64_t> key = iter_key(c, o, it);
    pty_ctx::GraphEntry* hit = nullptr;
    for (auto& e : c->gcache)
      if (e.key == key) { hit = &e; break; }
    if (!hit) {
      const HostIterState before = save_state(c);
      PTY_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int r = run_one(c, o, it, c->metrics);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (r != PTY_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
      }
      hipGraphExec_t exec = nullptr;
      hipError_t ei = ec;
      if (ec == hipSuccess) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (ei != hipSuccess || capture_fail_forced()) {
        // nothing of the captured iteration ran: undo its host-side effects,
        // drop the graphs and run this and the remaining iterations eagerly
        if (exec) (void)hipGraphExecDestroy(exec);
        load_state(c, before);
        drop_graphs(c);
        c->graphs = false;
        c->graphs_fell_back = true;
        c->use_slot = false;
        PTY_TRY(run_one(c, o, it, metrics_at(c, k)));
        continue;
      }
      c->gcache.push_back({key, exec, save_state(c)});
      hit = &c->gcache.back();
    } else {
      load_state(c, hit->st);
    }
    PTY_HIP(hipGraphLaunch(hit->exec, c->stream));
  }
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

// validation and per-run set-up of pty_run (before any iteration)
// Host-side switch for timing the real-space DM-with-background path:
// PTY_DM_FOURIER=0.  Read at every run.
bool dmf_disabled() {
  const char* e = getenv("PTY_DM_FOURIER");
  return e && atoi(e) == 0;
}

// DM method 1 entering the spectral path from f in real space: the
// iteration's O = Psup(f) (dm.py:85), then F = FFT2(f) in place
int dm1_to_spectra(pty_ctx* c, float alpha) {
  if (c->dm1_op) {
    // dm.py:65-82: object with the current probe, then probe with the new object
    PTY_TRY(gather(c, c->F, c->P, nullptr, c->O, c->H, 1, 1, alpha, false, false));
    PTY_TRY(probe_update(c, c->F, c->O, nullptr, c->P, c->G, 1, 1, alpha, false));
    c->H_valid = c->G_valid = true;
  } else {
    PTY_TRY(gather(c, c->F, c->P, nullptr, c->O, c->H, !c->H_valid, 1, alpha, false, false));
    c->H_valid = true;
  }
  PTY_TRY(copy_d2d(c, c->Ps, c->P, c->MK() * sizeof(float2)));
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->F, c->F, c->N * c->M));
  ColArgs a = col_args(c);
  a.X = c->F;
  a.nframes = c->N;
  a.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
  PTY_TRY(cols<COL_FWD>(c, a));
  c->f_spec = true;
  c->dmf_lead = false;
  return PTY_OK;
}

// F (spectra) -> f in real space, through X
int dm_f_to_real(pty_ctx* c) {
  PTY_TRY(copy_d2d(c, c->X, c->F, c->NMK() * sizeof(float2)));
  ColArgs a = col_args(c);
  a.X = c->X;
  a.nframes = c->N;
  a.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
  PTY_TRY(cols<COL_INV>(c, a));
  PTY_TRY(rows_inv<DST_STORE>(c, c->X, c->F, nullptr, nullptr, 1.0f, nullptr, c->N * c->M));
  c->f_spec = false;
  c->dmf_lead = false;
  return PTY_OK;
}

// f in real space -> F = its spectra, with the leading Pmod_7 of (f, b)
// column-inverted into Xo and its background sums in partb2 (b unchanged)
int dm_lead(pty_ctx* c, float alpha) {
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->F, c->X, c->N * c->M));
  ColArgs a = col_args(c);
  a.X = c->X;
  a.amp = c->amp;
  a.amp_stride = (long)c->MK();
  a.alpha = alpha;
  a.bg = c->bg;
  a.bg_stride = (long)c->MK();
  a.partb = c->partb2;
  a.F = c->F;
  a.Xo = c->Xo;
  PTY_TRY(cols<COL_DMBF_LEAD>(c, a));
  c->f_spec = true;
  c->dmf_lead = true;
  c->dmf_alpha = alpha;
  return PTY_OK;
}

// dst[n] = IFFT(FFT(P) ramp(d_n)) (deriv 0) or its derivative in d_n0 / d_n1
// (deriv 1 / 2), from the spectrum of P in Xp
int probe_stack(pty_ctx* c, float2* dst, int deriv) {
  hipLaunchKernelGGL(k_subpixel_ramp, dim3(2048), dim3(WG), 0, c->stream, (const float2*)c->Xp,
                     (const double2*)c->sub_d, dst, c->N, c->M, c->K, deriv);
  PTY_LAUNCHED("k_subpixel_ramp");
  ColArgs b = col_args(c);
  b.X = dst;
  b.nframes = c->N;
  b.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
  PTY_TRY(cols<COL_INV>(c, b));
  PTY_TRY(rows_inv<DST_STORE>(c, dst, dst, nullptr, nullptr, 1.0f, nullptr, c->N * c->M));
  return PTY_OK;
}

// Pn[n] = IFFT(FFT(P) ramp(d_n)): the forward transform of P into Xp, one
// phase ramp per frame, then the batched inverse transform of the stack
int build_subpixel_probes(pty_ctx* c) {
  if (!c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
  StageTimer tm(c, PTY_STAGE_MISC);
  c->g.Pn = nullptr;
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->P, c->Xp, c->M));
  ColArgs a = col_args(c);
  a.X = c->Xp;
  a.nframes = 1;
  a.chunk = 1;
  PTY_TRY(cols<COL_FWD>(c, a));
  PTY_TRY(probe_stack(c, c->Pn, 0));
  c->g.Pn = c->Pn;
  c->Pn_valid = true;
  c->H_valid = false;   // the heatmap is the sum of the per-frame |Pn|^2
  return PTY_OK;
}

int run_prepare(pty_ctx* c, const pty_opts* o, int64_t iter0, int32_t iters, int32_t begin) {
  if (!c || !o) return fail(PTY_EINVAL, "NULL argument");
  if (o->method < 1 || o->method > 6) return fail(PTY_EINVAL, "method must be 1..6");
  if (o->algorithm != 0 && o->algorithm != 1) return fail(PTY_EINVAL, "algorithm must be 0 or 1");
  if (o->op_n < 0 || o->op_every < 1) return fail(PTY_EINVAL, "bad OP_iters");
  if (iters < 0) return fail(PTY_EINVAL, "iters < 0");
  PTY_TRY(check_bound(c, o->method >= 4));
  if (c->sub_set) {
    if (o->method != 1 && o->method != 4)
      return fail(PTY_EUNSUP, "sub-pixel positions are modelled for the object-only methods 1 and 4");
    if (o->pmod_probe_until > iter0)
      return fail(PTY_EUNSUP, "sub-pixel positions are modelled without the far-field probe constraint");
    if (begin || !c->Pn_valid) PTY_TRY(build_subpixel_probes(c));
  }
  if (c->nscan > 1) {
    // packed scans: the fused ERA steps whose partials are kept per scan
    if (o->algorithm != 0) return fail(PTY_EUNSUP, "packed scans run ERA");
    if (o->method > 3) return fail(PTY_EUNSUP, "packed scans: methods 1-3 (no background)");
    if (o->probe_centering) return fail(PTY_EUNSUP, "packed scans: no probe centring");
    if (o->pmod_probe_until > iter0)
      return fail(PTY_EUNSUP, "packed scans: no far-field probe constraint");
    if (o->method == 3 && (o->op_n != 1 || o->op_every != 1))
      return fail(PTY_EUNSUP, "packed scans: method 3 with OP_iters (1, 1)");
    if (c->use_frame) return fail(PTY_EUNSUP, "packed scans take the staged kernels");
    if (c->chunk_frames > 0 && c->chunk_frames < c->N)
      return fail(PTY_EUNSUP, "packed scans: modulus step over the whole stack");
    if (!c->sc_norm_set) return fail(PTY_ESTATE, "packed scans: no I_norm (pty_prepare_amp)");
  }
  c->os_next = false;   // the caller may have changed O, P or the data since
  c->os_next_op = false;
  if (o->pmod_probe_until > iter0 && !c->pamp_set) PTY_TRY(pty_set_pamp(c, nullptr));
  if ((long)iters * c->nscan > c->metrics_cap) {
    dev_free(c->metrics);
    PTY_TRY(dev_alloc(&c->metrics, 3 * (size_t)iters * c->nscan));
    c->metrics_cap = iters * c->nscan;
  }
  if (begin) {
    c->H_valid = c->G_valid = false;
    c->have_metric_state = false;
    c->era_bg_shared = false;
    const size_t UV = (size_t)c->U * c->V;
    PTY_TRY(copy_d2d(c, c->Obak, c->O, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->Pbak, c->P, c->MK() * sizeof(float2)));
    if (o->algorithm == 1) {
      PTY_TRY(ensure_F(c));
      hipLaunchKernelGGL(k_make_exits, dim3(4096), dim3(WG), 0, c->stream, c->g, (const float2*)c->O,
                         (const float2*)c->P, c->F);
      PTY_LAUNCHED("k_make_exits");
      c->dm_ready = true;
      c->f_spec = false;
      c->dmf_lead = false;
    }
  }
  if (o->algorithm == 1 && !c->dm_ready) return fail(PTY_ESTATE, "DM run needs begin=1 first");
  // DM with background refinement and a fixed probe (method 4, no far-field
  // probe constraint in this run) carries its iterate as spectra when the
  // fused row gather is available; everything else in real space
  c->dmf_run = o->algorithm == 1 && o->method == 4 && o->pmod_probe_until <= iter0 && c->rowg &&
               !dmf_disabled();
  // (iterations inside a Pmod_probe window take the real-space path; the
  // first iteration after it switches, so the split of iterations into runs
  // does not change the result)
  c->dm1_op = o->method == 3 && o->op_every == 1 && o->op_n == 1 && !o->probe_centering;
  c->dm1_run = o->algorithm == 1 && (o->method == 1 || c->dm1_op) && c->rowg && !dmf_disabled() &&
               !c->dmf_run && iters > 0 && o->pmod_probe_until < iter0 + iters;
  c->dm1_pending = false;
  if (c->dm1_run) {
    for (int k = 0; k < 2; ++k)
      if (!c->dm1_esup[k]) PTY_TRY(dev_alloc(&c->dm1_esup[k], (size_t)c->cap_col * MAX_CHUNKS));
    if (c->dm1_op && !c->Xo) PTY_TRY(dev_alloc(&c->Xo, c->NMK()));   // f in real space
  }
  if (o->algorithm == 1) {
    if (c->dm1_run && o->pmod_probe_until <= iter0) {
      PTY_TRY(copy_d2d(c, c->Ps, c->P, c->MK() * sizeof(float2)));
      ColArgs a = col_args(c);
      a.nframes = c->N;
      a.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
      // the first iteration's O = Psup(f) (dm.py:85)
      if (c->f_spec) {
        // from the spectra, through the kernels an iteration's metric
        // projection takes (so split runs equal one run bit for bit)
        PTY_TRY(copy_d2d(c, c->X, c->F, c->NMK() * sizeof(float2)));
        a.X = c->X;
        PTY_TRY(cols<COL_INV>(c, a));
        if (!c->H_valid || c->dm1_op) {
          PTY_TRY(gather(c, nullptr, c->P, nullptr, nullptr, c->H, 1, 1, o->alpha, false, false));
          c->H_valid = true;
        }
        PTY_TRY(rows_gather(c, c->X, o->alpha, false, false, nullptr, c->dm1_op ? c->Xo : nullptr));
        if (c->dm1_op) {
          PTY_TRY(probe_update(c, c->Xo, c->O, nullptr, c->P, c->G, 1, 1, o->alpha, false));
          c->G_valid = true;
        }
      } else {
        PTY_TRY(dm1_to_spectra(c, o->alpha));
      }
    } else if (c->dmf_run) {
      if (!c->Xo) PTY_TRY(dev_alloc(&c->Xo, c->NMK()));
      if (!c->partb2) PTY_TRY(dev_alloc(&c->partb2, (size_t)c->col_chunks * c->MK()));
      if (c->f_spec && (!c->dmf_lead || c->dmf_alpha != o->alpha)) PTY_TRY(dm_f_to_real(c));
      if (!c->f_spec) PTY_TRY(dm_lead(c, o->alpha));
    } else if (c->f_spec) {
      PTY_TRY(dm_f_to_real(c));
    }
  }
  // ERA's object + probe steps keep psi' in F when the fused row gather runs
  // them (allocated here, never inside a capture)
  if (o->algorithm == 0 && (o->method == 3 || o->method == 6) && c->rowg) PTY_TRY(ensure_F(c));
  c->last_algo = o->algorithm;
  if (c->use_frame) PTY_TRY(ensure_slots(c));  // never allocate inside a capture
  return PTY_OK;
}

int dm_bg_fourier_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out);

int dm1_fourier_iter(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_out);
int dm1_tail(pty_ctx* c, double* metrics_out);

int run_one(pty_ctx* c, const pty_opts& o, int64_t it, double* metrics_dev) {
  if (o.algorithm == 0) return era_iter(c, o, it, metrics_dev);
  if (c->dm1_run && it >= o.pmod_probe_until) {
    if (!c->f_spec) PTY_TRY(dm1_to_spectra(c, o.alpha));   // leaving a Pmod_probe window
    return dm1_fourier_iter(c, o, it, metrics_dev);
  }
  return c->dmf_run ? dm_bg_fourier_iter(c, o, it, metrics_dev) : dm_iter(c, o, it, metrics_dev);
}

// after a run's iterations: the metrics a spectral DM method-1 run still owes
int run_tail(pty_ctx* c, int32_t iters) {
  if (!c->dm1_pending || iters <= 0) return PTY_OK;
  return dm1_tail(c, c->use_slot ? c->metrics : metrics_at(c, iters - 1));
}

int run_finish(pty_ctx* c, int32_t iters, double* metrics_host) {
  if (metrics_host && iters > 0)
    PTY_HIP(hipMemcpyAsync(metrics_host, c->metrics, 3 * (size_t)c->nscan * iters * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  PTY_HIP(hipStreamSynchronize(c->stream));
  if (c->prof_on) prof_collect(c);
  return PTY_OK;
}

void twiddles(int L, std::vector<float2>& out) {
  out.resize(L);
  for (int k = 0; k < L; ++k) {
    const double ang = -2.0 * M_PI * (double)k / (double)L;
    out[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
}

}  // namespace

// --------------------------------------------------------------------------
// Scratch of the allocation-free modulus step (pty_pmod1_noalloc): the column
// kernels' eMod partials, the packed mask words and the twiddle tables of every
// supported line length, in one file-scope device array.  Each GPU holds its
// own copy (at its own address) until the process ends; every call on a GPU
// shares it, so calls that overlap on one GPU race.
// --------------------------------------------------------------------------
namespace {
constexpr size_t NA_PARTS = 32768;   // eMod partials: at most 128 work-groups per CU x 256 CUs
constexpr size_t NA_MASKW = 16 * 256;   // packed mask words: T x K, at most 16 x 256
constexpr int NA_SIZES[] = {16, 32, 64, 96, 128, 192, 256};
constexpr int NA_TW = 16 + 32 + 64 + 96 + 128 + 192 + 256;   // twiddle entries, all sides
constexpr size_t NA_OFF_MASK = NA_PARTS * sizeof(double);
constexpr size_t NA_OFF_TW = NA_OFF_MASK + NA_MASKW * sizeof(uint32_t);
constexpr size_t NA_BYTES = NA_OFF_TW + NA_TW * sizeof(float2);   // 284,800 bytes
static_assert(NA_BYTES == 284800, "scratch layout");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NA_MAX_DEV = 64;
struct NoallocDevice {
  unsigned char* base = nullptr;   // this device's address of `scratch`
  int cus = 0;
};
NoallocDevice g_noalloc[NA_MAX_DEV];
std::mutex g_noalloc_mu;

int noalloc_tw_offset(int L) {
  int off = 0;
  for (int s : NA_SIZES) {
    if (s == L) return off;
    off += s;
  }
  return -1;
}

// Once per device, with it current: its address of `scratch` and the twiddle
// tables (a synchronous copy from the host; nothing is allocated).
int noalloc_setup(int dev, NoallocDevice** out) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[dev];
  if (!d.base) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PTY_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    float2 tw[NA_TW];
    int k0 = 0;
    for (int L : NA_SIZES) {
      for (int k = 0; k < L; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)L;
        tw[k0 + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
      k0 += L;
    }
    e = hipMemcpy(static_cast<unsigned char*>(ptr) + NA_OFF_TW, tw, sizeof(tw), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(PTY_EHIP, std::string("twiddle upload: ") + hipGetErrorString(e));
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || cus <= 0)
      return fail(PTY_EHIP, std::string("compute-unit count: ") + hipGetErrorString(e));
    d.cus = cus;
    d.base = static_cast<unsigned char*>(ptr);
  }
  *out = &d;
  return PTY_OK;
}
}  // namespace

// Several scans advanced together: each scan's kernels go to one of up to
// four lane streams (default priority), so small scans fill the device
// side by side; with graphs on, one iteration of every scan is captured as one
// hipGraph (fork / join through events) and replayed.
struct pty_batch {
  std::vector<pty_ctx*> ctx;
  std::vector<hipStream_t> lanes;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  bool graphs = false;
  bool fell_back = false;   // a capture failed; the batch runs eagerly since
  struct Entry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    std::vector<HostIterState> st;
  };
  std::vector<Entry> cache;

  void drop() {
    for (auto& e : cache) (void)hipGraphExecDestroy(e.exec);
    cache.clear();
  }
  ~pty_batch() {
    drop();
    for (auto l : lanes) (void)hipStreamDestroy(l);
    if (origin) (void)hipStreamDestroy(origin);
    if (fork) (void)hipEventDestroy(fork);
    for (auto e : join) (void)hipEventDestroy(e);
  }
};

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char* pty_last_error(void) { return g_err.c_str(); }

int pty_supported_frame(int32_t L) { return threads_per_line(L) > 0 ? 1 : 0; }

int pty_create(pty_ctx** out, int32_t N, int32_t M, int32_t K, int32_t U, int32_t V,
               const int32_t* R_host) {
  if (!out) return fail(PTY_EINVAL, "out is NULL");
  *out = nullptr;
  if (N <= 0) return fail(PTY_EINVAL, "N must be positive");
  if (!pty_supported_frame(M) || !pty_supported_frame(K))
    return fail(PTY_EUNSUP, "frame sides must be one of 16, 32, 64, 96, 128, 192, 256 (got " +
                                std::to_string(M) + " x " + std::to_string(K) + ")");
  if (U < M || V < K) return fail(PTY_EINVAL, "object smaller than a frame");
  if (!R_host) return fail(PTY_EINVAL, "R is NULL");
  if ((size_t)N * M * K > ((size_t)1 << 34)) return fail(PTY_EINVAL, "frame stack too large");
  std::vector<int2> off(N);
  for (int n = 0; n < N; ++n) {
    const int y = -R_host[2 * n], x = -R_host[2 * n + 1];
    if (y < 0 || x < 0 || y + M > U || x + K > V)
      return fail(PTY_EINVAL, "frame " + std::to_string(n) + " window (" + std::to_string(y) +
                                  ", " + std::to_string(x) + ") lies outside the object");
    off[n] = make_int2(y, x);  // .x = row offset, .y = column offset
  }
  pty_ctx* c = new pty_ctx();
  c->N = N; c->M = M; c->K = K; c->U = U; c->V = V;
  c->h_off = off;
  auto cleanup = [&](int code) { pty_destroy(c); return code; };

  // object tiles -> ascending frame lists (CSR)
  c->tiles_y = (U + TILE_H - 1) / TILE_H;
  c->tiles_x = (V + TILE_W - 1) / TILE_W;
  c->ntiles = c->tiles_x * c->tiles_y;
  std::vector<int> cnt(c->ntiles + 1, 0);
  for (int n = 0; n < N; ++n)
    for (int ty = off[n].x / TILE_H; ty <= (off[n].x + M - 1) / TILE_H; ++ty)
      for (int tx = off[n].y / TILE_W; tx <= (off[n].y + K - 1) / TILE_W; ++tx)
        cnt[ty * c->tiles_x + tx + 1]++;
  for (int t = 0; t < c->ntiles; ++t) cnt[t + 1] += cnt[t];
  std::vector<int> frames(cnt[c->ntiles]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int n = 0; n < N; ++n)
    for (int ty = off[n].x / TILE_H; ty <= (off[n].x + M - 1) / TILE_H; ++ty)
      for (int tx = off[n].y / TILE_W; tx <= (off[n].y + K - 1) / TILE_W; ++tx)
        frames[fill[ty * c->tiles_x + tx]++] = n;

  // line-FFT plans of new contexts (pty_set_plan's mask).  Measured per
  // stage on scans of each side (ERA 1 / 3, DM 1 / 4): the alternative
  // plans win only in the fused row gather at 192 and 256 (4-6 %: 32 or 16
  // threads per line halve its parked rows); PTY_FFT_PLAN overrides
  c->plan = (K == 192 || K == 256) ? PLAN_ROWG | PLAN_BGERR : 0;
  if (const char* e_pl = getenv("PTY_FFT_PLAN")) c->plan = std::max(0, std::min(PLAN_ALL, atoi(e_pl)));

  // object rows -> ascending frame lists, when the fused kernel applies; the
  // segments are cut for the row gather's plan at creation (a later plan
  // change keeps them: still correct, possibly fewer work-groups per CU)
  const int segs = rowg_segments(pk(c, K, PLAN_ROWG), V);
  c->rowg_ok = segs > 0 && (size_t)N * M <= ((size_t)1 << 26);
  std::vector<int> rcnt;
  std::vector<int2> rents;
  if (c->rowg_ok) {
    c->rowg_segs = segs;
    c->rowg_segw = (V + segs - 1) / segs;
    const int sw = c->rowg_segw, nitem = U * segs;
    auto seg0 = [&](int n) { return off[n].y / sw; };
    auto seg1 = [&](int n) { return (off[n].y + K - 1) / sw; };
    rcnt.assign(nitem + 1, 0);
    for (int n = 0; n < N; ++n)
      for (int i = 0; i < M; ++i)
        for (int sg = seg0(n); sg <= seg1(n); ++sg) rcnt[(off[n].x + i) * segs + sg + 1]++;
    for (int k = 0; k < nitem; ++k) rcnt[k + 1] += rcnt[k];
    rents.resize(rcnt[nitem]);
    std::vector<int> rfill(rcnt.begin(), rcnt.end() - 1);
    for (int n = 0; n < N; ++n)
      for (int i = 0; i < M; ++i)
        for (int sg = seg0(n); sg <= seg1(n); ++sg)
          rents[rfill[(off[n].x + i) * segs + sg]++] = make_int2(n * M + i, off[n].y << 16 | i);
  }

  // frame rows -> object-row segment lists for the forward row stage, cut for
  // the row plan at creation (a later plan change keeps them if its LDS fits)
  const int fsegs = rowf_segments(pk(c, K, PLAN_ROWS), K, V);
  c->rowf_ok = fsegs > 0 && (size_t)N * M <= ((size_t)1 << 26);
  std::vector<int> fcnt;
  std::vector<int2> fents;
  if (c->rowf_ok) {
    c->rowf_segs = fsegs;
    c->rowf_segw = (V + fsegs - 1) / fsegs;
    const int sw = c->rowf_segw, nitem = U * fsegs;
    fcnt.assign(nitem + 1, 0);
    for (int n = 0; n < N; ++n)
      for (int i = 0; i < M; ++i) fcnt[(off[n].x + i) * fsegs + off[n].y / sw + 1]++;
    for (int k = 0; k < nitem; ++k) fcnt[k + 1] += fcnt[k];
    fents.resize(fcnt[nitem]);
    std::vector<int> ffill(fcnt.begin(), fcnt.end() - 1);
    for (int n = 0; n < N; ++n)
      for (int i = 0; i < M; ++i)
        fents[ffill[(off[n].x + i) * fsegs + off[n].y / sw]++] = make_int2(n * M + i, off[n].y << 16 | i);
  }

  // launch shapes
  c->col_blocks = (K + (WG / threads_per_line(M)) - 1) / (WG / threads_per_line(M));
  c->col_chunk = std::max(1, std::min(16, (int)((long)N * c->col_blocks / 2048)));
  c->col_chunks = (N + c->col_chunk - 1) / c->col_chunk;
  const int MK = M * K;
  const int pblocks = (MK + WG - 1) / WG;
  c->probe_chunks = std::max(1, std::min(N, std::max(1, 4096 / pblocks)));
  c->probe_chunk = (N + c->probe_chunks - 1) / c->probe_chunks;
  c->probe_chunks = (N + c->probe_chunk - 1) / c->probe_chunk;

  c->frame_ok = frame_supported(M, K);
  {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    }
    const char* e_nt = getenv("PTY_FRAME_NT");
    const char* e_sl = getenv("PTY_FRAME_SLOTS");
    const char* e_on = getenv("PTY_FRAME");
    c->frame_nt = (e_nt && atoi(e_nt) == 512) ? 512 : 1024;
    c->frame_slots = e_sl ? std::max(1, atoi(e_sl)) : cus;
    c->cus = cus;
    const char* e_cw = getenv("PTY_COLS_WG");
    if (e_cw && (atoi(e_cw) == 256 || atoi(e_cw) == 512)) c->cols_wg = atoi(e_cw);
    // The modulus step runs over the whole stack at once by default: blocks
    // sized for the 256 MB Infinity Cache (128..512 frames of 256^2) measured
    // 8..18% slower per ERA iteration at 4096 x 256^2, the persistent row and
    // column grids being under-filled by one block.  pty_tune tries them.
    c->chunk_frames = 0;
    const char* e_ch = getenv("PTY_CHUNK_FRAMES");
    if (e_ch) c->chunk_frames = std::max(0, atoi(e_ch));
    // measured: the staged kernels are faster than the fused per-frame one
    // at 256^2 and, with the 4-wave column kernel and the fused object
    // update, at 128^2 too (1024 x 128^2 ERA method 1: 5051 vs 2715 it/s)
    c->use_frame = e_on && atoi(e_on) == 1 && c->frame_ok;
  }
  c->cap_col = std::max(128 * c->cus, c->frame_slots);
  c->cap_gather = std::max(c->ntiles, U * c->rowg_segs);
  c->cap_econO = std::max(c->cap_gather, ECON_BLOCKS);
  c->cap_econP = std::max(pblocks, ECON_BLOCKS);
  c->cap_esup = std::max(std::max(ESUP_BLOCKS, row_blocks(c, N * M)), c->frame_slots);

  int r;
  if ((r = dev_alloc(&c->off, N)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->tile_ptr, c->ntiles + 1)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->tile_frames, frames.size())) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->twM, M)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->twK, K)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->X, c->NMK())) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Xp, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Obak, (size_t)U * V)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Os, (size_t)U * V)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Oroll, (size_t)U * V)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Pbak, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->Ps, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->H, (size_t)U * V)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->G, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->bbar, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->bbar_m, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->pamp, MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->pnum, (size_t)c->probe_chunks * MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->pg, (size_t)c->probe_chunks * MK)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->cm, 2)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->scal, 8)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_emod, (size_t)c->cap_col * MAX_CHUNKS)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_esa, (size_t)c->cap_col * MAX_CHUNKS)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_econO, 2 * (size_t)c->cap_econO)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_econP, 2 * (size_t)c->cap_econP)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_esup, c->cap_esup)) != PTY_OK) return cleanup(r);
  if ((r = dev_alloc(&c->part_esxh, 2 * (size_t)c->cap_gather)) != PTY_OK) return cleanup(r);
  if (c->rowg_ok) {
    if ((r = dev_alloc(&c->row_ptr, rcnt.size())) != PTY_OK) return cleanup(r);
    if ((r = dev_alloc(&c->row_ent, rents.size())) != PTY_OK) return cleanup(r);
    if ((r = rowg_attribute(K)) != PTY_OK) return cleanup(r);
    const char* e_rg = getenv("PTY_ROW_GATHER");
    c->rowg = !(e_rg && atoi(e_rg) == 0);
  }

  if (c->rowf_ok) {
    if ((r = dev_alloc(&c->rowf_ptr, fcnt.size())) != PTY_OK) return cleanup(r);
    if ((r = dev_alloc(&c->rowf_ent, fents.size())) != PTY_OK) return cleanup(r);
    if ((r = rowf_attribute()) != PTY_OK) return cleanup(r);
    const char* e_rf = getenv("PTY_ROW_FWD_OROW");
    c->rowf = !(e_rf && atoi(e_rf) == 0);
  }

  c->cap_pchunks = (size_t)c->probe_chunks;
  c->cap_G = 1;
  c->cap_econP_all = 2 * (size_t)c->cap_econP;
  c->cap_esup_all = (size_t)c->cap_esup;

  std::vector<float2> tw;
  hipError_t e = hipSuccess;
  twiddles(M, tw);
  if (e == hipSuccess) e = hipMemcpy(c->twM, tw.data(), M * sizeof(float2), hipMemcpyHostToDevice);
  twiddles(K, tw);
  if (e == hipSuccess) e = hipMemcpy(c->twK, tw.data(), K * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->off, off.data(), N * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->tile_ptr, cnt.data(), (c->ntiles + 1) * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->tile_frames, frames.data(), frames.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && c->rowg_ok)
    e = hipMemcpy(c->row_ptr, rcnt.data(), rcnt.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && c->rowg_ok)
    e = hipMemcpy(c->row_ent, rents.data(), rents.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess && c->rowf_ok)
    e = hipMemcpy(c->rowf_ptr, fcnt.data(), fcnt.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && c->rowf_ok)
    e = hipMemcpy(c->rowf_ent, fents.data(), fents.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(PTY_EHIP, std::string("upload: ") + hipGetErrorString(e)));
  c->g.N = N; c->g.M = M; c->g.K = K; c->g.U = U; c->g.V = V;
  c->g.off = c->off;
  *out = c;
  return PTY_OK;
}

int pty_destroy(pty_ctx* c) {
  if (!c) return PTY_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  else (void)hipDeviceSynchronize();
  dev_free(c->off); dev_free(c->tile_ptr); dev_free(c->tile_frames);
  dev_free(c->row_ptr); dev_free(c->row_ent);
  dev_free(c->rowf_ptr); dev_free(c->rowf_ent);
  dev_free(c->twM); dev_free(c->twK);
  dev_free(c->X); dev_free(c->F); dev_free(c->Xp);
  dev_free(c->Obak); dev_free(c->Pbak); dev_free(c->Os); dev_free(c->Oroll); dev_free(c->Ps);
  dev_free(c->H); dev_free(c->G); dev_free(c->bg); dev_free(c->bbar); dev_free(c->bbar_m);
  dev_free(c->pamp); dev_free(c->partb); dev_free(c->maskbits); dev_free(c->maskbits_alt); dev_free(c->pnum); dev_free(c->pg);
  dev_free(c->cm); dev_free(c->metrics); dev_free(c->scal); dev_free(c->slots);
  dev_free(c->sub_d); dev_free(c->Pn); dev_free(c->Dy); dev_free(c->Dx); dev_free(c->pos_terms);
  dev_free(c->Xo); dev_free(c->partb2); dev_free(c->dm1_esup[0]); dev_free(c->dm1_esup[1]);
  dev_free(c->part_emod); dev_free(c->part_esa); dev_free(c->part_econO);
  dev_free(c->part_econP); dev_free(c->part_esup); dev_free(c->part_esxh);
  dev_free(c->fscan); dev_free(c->sc_fstart); dev_free(c->sc_rstart); dev_free(c->sc_pchunk);
  dev_free(c->sc_pscan); dev_free(c->sc_inorm);
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->gstream) (void)hipStreamDestroy(c->gstream);
  dev_free(c->slot);
  delete c;
  return PTY_OK;
}

int pty_set_stream(pty_ctx* c, void* stream) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  c->stream = (hipStream_t)stream;
  return PTY_OK;
}

int pty_prepare_amp(pty_ctx* c, const float* I, const uint8_t* mask_centred, float* amp,
                    double* I_norm_out) {
  if (!c || !I || !amp) return fail(PTY_EINVAL, "NULL argument");
  if (c->nscan > 1) {
    // packed scans: each scan's frames with its own mask ([S][M][K]), one I_norm per scan
    const size_t MK = c->MK();
    std::vector<double> norms(c->nscan);
    for (int sc = 0; sc < c->nscan; ++sc) {
      const int f0 = c->h_fstart[sc], nf = c->h_fstart[sc + 1] - f0;
      hipLaunchKernelGGL(k_prepare_amp, dim3(ESUP_BLOCKS), dim3(WG), 0, c->stream,
                         I + (size_t)f0 * MK, nf, c->M, c->K,
                         mask_centred ? mask_centred + (size_t)sc * MK : nullptr,
                         amp + (size_t)f0 * MK, c->part_esup);
      PTY_LAUNCHED("k_prepare_amp");
      double sv = 0.0;
      PTY_TRY(sum_region(c, c->part_esup, ESUP_BLOCKS, &sv));
      norms[sc] = sv;
      if (I_norm_out) I_norm_out[sc] = sv;
    }
    // the scans' metrics normalise by these sums
    PTY_HIP(hipMemcpy(c->sc_inorm, norms.data(), c->nscan * sizeof(double), hipMemcpyHostToDevice));
    c->sc_norm_set = true;
    return PTY_OK;
  }
  hipLaunchKernelGGL(k_prepare_amp, dim3(ESUP_BLOCKS), dim3(WG), 0, c->stream, I, c->N, c->M, c->K,
                     mask_centred, amp, c->part_esup);
  PTY_LAUNCHED("k_prepare_amp");
  double s = 0.0;
  PTY_TRY(sum_region(c, c->part_esup, ESUP_BLOCKS, &s));
  if (I_norm_out) *I_norm_out = s;
  return PTY_OK;
}

// The bound mask in the column kernels' register layouts: one word per
// (column-thread, column), one set per scan, for the side's default plan and
// for its alternative (column modes take either, pty_set_plan).
int pack_mask(pty_ctx* c) {
  for (int v = 0; v < 2; ++v) {
    const int pkm = plan_key(c->M, v);
    const int T = threads_per_line(pkm);
    if (T == 0) continue;   // no alternative plan for this side
    uint32_t*& bits = v ? c->maskbits_alt : c->maskbits;
    size_t& cap = v ? c->cap_maskbits_alt : c->cap_maskbits;
    const size_t words = (size_t)T * c->K * c->nscan;
    if (cap < words) {
      dev_free(bits);
      cap = 0;
      PTY_TRY(dev_alloc(&bits, words));
      cap = words;
    }
    const int grid = (T * c->K + WG - 1) / WG;
    for (int sc = 0; sc < c->nscan; ++sc) {
      const uint8_t* ms = c->mask + (size_t)sc * c->MK();
      uint32_t* mb = bits + (size_t)sc * T * c->K;
      switch (pkm) {
#define C(Lv)                                                                                  \
  case Lv:                                                                                     \
    hipLaunchKernelGGL(k_pack_mask<Lv>, dim3(grid), dim3(WG), 0, c->stream, ms, c->K, mb);     \
    break;
        PTY_PLANS(C)
#undef C
        default: return fail(PTY_EUNSUP, "unsupported M");
      }
      PTY_LAUNCHED("k_pack_mask");
    }
  }
  return PTY_OK;
}

int pty_set_data(pty_ctx* c, const float* amp, const uint8_t* mask, double I_norm) {
  if (!c || !amp) return fail(PTY_EINVAL, "NULL argument");
  c->amp = amp;
  c->mask = mask;
  c->I_norm = I_norm;
  c->dmf_lead = false;   // the leading Pmod_7 in Xo used the previous data
  if (mask) PTY_TRY(pack_mask(c));
  return PTY_OK;
}

int pty_set_scans(pty_ctx* c, int32_t S, const int32_t* frame_start, const int32_t* row_start,
                  const double* I_norm) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (S < 1) return fail(PTY_EINVAL, "S must be >= 1");
  if (S > 1 && (!frame_start || !row_start)) return fail(PTY_EINVAL, "NULL argument");
  if (S > 1 && c->sub_set) return fail(PTY_EUNSUP, "packed scans take integer positions");
  if (S > 1) {
    if (frame_start[0] != 0 || frame_start[S] != c->N)
      return fail(PTY_EINVAL, "frame_start must run from 0 to N");
    if (row_start[0] != 0 || row_start[S] != c->U)
      return fail(PTY_EINVAL, "row_start must run from 0 to U");
    for (int sc = 0; sc < S; ++sc) {
      if (frame_start[sc + 1] <= frame_start[sc]) return fail(PTY_EINVAL, "every scan needs frames");
      if (row_start[sc + 1] <= row_start[sc]) return fail(PTY_EINVAL, "every scan needs object rows");
      if (row_start[sc] % TILE_H)
        return fail(PTY_EINVAL, "scan object bands must start on multiples of 16 rows");
      for (int n = frame_start[sc]; n < frame_start[sc + 1]; ++n) {
        const int y = c->h_off[n].x;
        if (y < row_start[sc] || y + c->M > row_start[sc + 1])
          return fail(PTY_EINVAL, "frame " + std::to_string(n) + " lies outside scan " +
                                      std::to_string(sc) + "'s object rows");
      }
    }
  }
  PTY_HIP(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  // a new packing rebinds data and state
  c->amp = nullptr;
  c->mask = nullptr;
  c->H_valid = c->G_valid = false;
  c->os_next = c->os_next_op = false;
  dev_free(c->fscan); dev_free(c->sc_fstart); dev_free(c->sc_rstart);
  dev_free(c->sc_pchunk); dev_free(c->sc_pscan); dev_free(c->sc_inorm);
  c->h_fstart.clear();
  c->h_rstart.clear();
  c->nscan = 1;
  c->g.fscan = nullptr;
  c->sc_norm_set = false;
  if (S == 1) {
    if (I_norm) c->I_norm = I_norm[0];
    return PTY_OK;
  }

  const int MK = (int)c->MK();
  const int pblocks = (MK + WG - 1) / WG;
  std::vector<int> fscan(c->N), pchunk, pscan(S + 1);
  for (int sc = 0; sc < S; ++sc) {
    const int f0 = frame_start[sc], nf = frame_start[sc + 1] - f0;
    for (int n = f0; n < f0 + nf; ++n) fscan[n] = sc;
    // probe chunks of a scan as a context of its own would cut them
    int chunks = std::max(1, std::min(nf, std::max(1, 4096 / pblocks)));
    const int chunk = (nf + chunks - 1) / chunks;
    chunks = (nf + chunk - 1) / chunk;
    pscan[sc] = (int)pchunk.size();
    for (int k = 0; k < chunks; ++k) pchunk.push_back(f0 + k * chunk);
  }
  pscan[S] = (int)pchunk.size();
  pchunk.push_back(c->N);
  const size_t nch = pchunk.size() - 1;
  // workspace per scan: probe heatmaps, probe-chunk sums, probe / eSup partials
  if (c->cap_pchunks < nch) {
    dev_free(c->pnum); dev_free(c->pg);
    c->cap_pchunks = 0;
    PTY_TRY(dev_alloc(&c->pnum, nch * MK));
    PTY_TRY(dev_alloc(&c->pg, nch * MK));
    c->cap_pchunks = nch;
  }
  if (c->cap_G < (size_t)S) {
    dev_free(c->G);
    c->cap_G = 0;
    PTY_TRY(dev_alloc(&c->G, (size_t)S * MK));
    c->cap_G = S;
  }
  if (c->cap_econP_all < 2 * (size_t)S * pblocks) {
    dev_free(c->part_econP);
    c->cap_econP_all = 0;
    PTY_TRY(dev_alloc(&c->part_econP, 2 * (size_t)S * pblocks));
    c->cap_econP_all = 2 * (size_t)S * pblocks;
  }
  const size_t esup_need = std::max(2 * (size_t)S * pblocks, (size_t)S * ESUP_BLOCKS);
  if (c->cap_esup_all < esup_need) {
    dev_free(c->part_esup);
    c->cap_esup_all = 0;
    PTY_TRY(dev_alloc(&c->part_esup, esup_need));
    c->cap_esup_all = esup_need;
  }
  PTY_TRY(dev_alloc(&c->fscan, (size_t)c->N));
  PTY_TRY(dev_alloc(&c->sc_fstart, (size_t)S + 1));
  PTY_TRY(dev_alloc(&c->sc_rstart, (size_t)S + 1));
  PTY_TRY(dev_alloc(&c->sc_pchunk, pchunk.size()));
  PTY_TRY(dev_alloc(&c->sc_pscan, pscan.size()));
  PTY_TRY(dev_alloc(&c->sc_inorm, (size_t)S));
  PTY_HIP(hipMemcpy(c->fscan, fscan.data(), fscan.size() * sizeof(int), hipMemcpyHostToDevice));
  PTY_HIP(hipMemcpy(c->sc_fstart, frame_start, (S + 1) * sizeof(int), hipMemcpyHostToDevice));
  PTY_HIP(hipMemcpy(c->sc_rstart, row_start, (S + 1) * sizeof(int), hipMemcpyHostToDevice));
  PTY_HIP(hipMemcpy(c->sc_pchunk, pchunk.data(), pchunk.size() * sizeof(int), hipMemcpyHostToDevice));
  PTY_HIP(hipMemcpy(c->sc_pscan, pscan.data(), pscan.size() * sizeof(int), hipMemcpyHostToDevice));
  if (I_norm) PTY_HIP(hipMemcpy(c->sc_inorm, I_norm, S * sizeof(double), hipMemcpyHostToDevice));
  c->sc_norm_set = I_norm != nullptr;
  c->h_fstart.assign(frame_start, frame_start + S + 1);
  c->h_rstart.assign(row_start, row_start + S + 1);
  c->pk_chunks = (int)nch;
  c->nscan = S;
  c->g.fscan = c->fscan;
  return PTY_OK;
}

int pty_set_state(pty_ctx* c, pty_cfloat* O, pty_cfloat* P) {
  if (!c || !O || !P) return fail(PTY_EINVAL, "NULL argument");
  c->O = reinterpret_cast<float2*>(O);
  c->P = reinterpret_cast<float2*>(P);
  c->H_valid = c->G_valid = false;
  c->Pn_valid = false;
  c->g.Pn = nullptr;
  return PTY_OK;
}

int pty_position_terms(pty_ctx* c, double* terms_host) {
  if (!c || !terms_host) return fail(PTY_EINVAL, "NULL argument");
  if (!c->sub_set) return fail(PTY_ESTATE, "pty_set_subpixel has not been called");
  if (!c->amp) return fail(PTY_ESTATE, "pty_set_data has not been called");
  if (!c->O || !c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
  if (!c->Dy) PTY_TRY(dev_alloc(&c->Dy, c->NMK()));
  if (!c->Dx) PTY_TRY(dev_alloc(&c->Dx, c->NMK()));
  if (!c->pos_terms) PTY_TRY(dev_alloc(&c->pos_terms, 6 * (size_t)c->N));
  if (!c->Pn_valid) PTY_TRY(build_subpixel_probes(c));   // leaves FFT(P) in Xp
  else {
    PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->P, c->Xp, c->M));
    ColArgs a = col_args(c);
    a.X = c->Xp;
    a.nframes = 1;
    a.chunk = 1;
    PTY_TRY(cols<COL_FWD>(c, a));
  }
  PTY_TRY(probe_stack(c, c->Dy, 1));
  PTY_TRY(probe_stack(c, c->Dx, 2));
  // spectra of O_win P_n, O_win dP_n/dd_y, O_win dP_n/dd_x: the derivative
  // stacks are transformed in place (each frame row reads its own probe row
  // before the same line writes it)
  ColArgs a = col_args(c);
  a.nframes = c->N;
  a.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
  const int nrows = c->N * c->M;
  PTY_TRY(rows_fwd<SRC_OP>(c, c->O, c->P, nullptr, c->X, nrows));
  a.X = c->X;
  PTY_TRY(cols<COL_FWD>(c, a));
  for (float2* D : {c->Dy, c->Dx}) {
    c->g.Pn = D;
    const int r = rows_fwd<SRC_OP>(c, c->O, c->P, nullptr, D, nrows);
    c->g.Pn = c->Pn;
    PTY_TRY(r);
    a.X = D;
    PTY_TRY(cols<COL_FWD>(c, a));
  }
  hipLaunchKernelGGL(k_position_terms, dim3(c->N), dim3(WG), 0, c->stream, (const float2*)c->X,
                     (const float2*)c->Dy, (const float2*)c->Dx, c->amp, c->mask, (int)c->MK(),
                     c->pos_terms);
  PTY_LAUNCHED("k_position_terms");
  PTY_HIP(hipMemcpyAsync(terms_host, c->pos_terms, 6 * (size_t)c->N * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->os_next = false;   // X no longer holds what a DM projection reuse expects
  c->os_next_op = false;
  return PTY_OK;
}

int pty_set_subpixel(pty_ctx* c, const double* d_host) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->H_valid = false;
  c->Pn_valid = false;
  c->g.Pn = nullptr;
  if (!d_host) {
    c->sub_set = false;
    return PTY_OK;
  }
  if (c->nscan > 1) {
    c->sub_set = false;
    return fail(PTY_EUNSUP, "packed scans take integer positions");
  }
  for (int k = 0; k < 2 * c->N; ++k)
    if (!std::isfinite(d_host[k]) || std::fabs(d_host[k]) > 1.0)
      return fail(PTY_EINVAL, "sub-pixel shifts must be finite and within [-1, 1]");
  if (!c->sub_d) PTY_TRY(dev_alloc(&c->sub_d, (size_t)c->N));
  if (!c->Pn) PTY_TRY(dev_alloc(&c->Pn, c->NMK()));
  PTY_HIP(hipMemcpy(c->sub_d, d_host, (size_t)c->N * sizeof(double2), hipMemcpyHostToDevice));
  c->sub_set = true;
  return PTY_OK;
}

int pty_set_background(pty_ctx* c, const float* b, int32_t per_frame) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_set_background: not available on packed scans");
  if (!c || !b) return fail(PTY_EINVAL, "NULL argument");
  if (!c->bg) PTY_TRY(dev_alloc(&c->bg, c->NMK()));
  if (!c->partb) PTY_TRY(dev_alloc(&c->partb, (size_t)c->col_chunks * c->MK()));
  if (per_frame) {
    PTY_TRY(copy_d2d(c, c->bg, b, c->NMK() * sizeof(float)));
  } else {
    hipLaunchKernelGGL(k_broadcast, dim3(4096), dim3(WG), 0, c->stream, c->bg, b, c->MK(), c->NMK());
    PTY_LAUNCHED("k_broadcast");
  }
  c->bg_set = true;
  c->era_bg_shared = false;
  c->dmf_lead = false;   // the leading Pmod_7 in Xo used the previous background
  return PTY_OK;
}

int pty_set_pamp(pty_ctx* c, const float* pamp) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_set_pamp: not available on packed scans");
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (pamp) {
    PTY_TRY(copy_d2d(c, c->pamp, pamp, c->MK() * sizeof(float)));
  } else {
    if (!c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
    PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, c->P, c->Xp, c->M));
    ColArgs a = col_args(c);
    a.X = c->Xp;
    a.nframes = 1;
    a.chunk = 1;
    PTY_TRY(cols<COL_FWD>(c, a));
    hipLaunchKernelGGL(k_abs, dim3(64), dim3(WG), 0, c->stream, (const float2*)c->Xp, c->pamp, c->MK());
    PTY_LAUNCHED("k_abs");
  }
  c->pamp_set = true;
  return PTY_OK;
}

int pty_run(pty_ctx* c, const pty_opts* o, int64_t iter0, int32_t iters, int32_t begin,
            double* metrics_host) {
  PTY_TRY(run_prepare(c, o, iter0, iters, begin));
  if (c->graphs) {
    PTY_TRY(run_graphs(c, *o, iter0, iters));
  } else {
    for (int32_t k = 0; k < iters; ++k) PTY_TRY(run_one(c, *o, iter0 + k, metrics_at(c, k)));
  }
  PTY_TRY(run_tail(c, iters));
  return run_finish(c, iters, metrics_host);
}

int pty_set_path(pty_ctx* c, int32_t fused, int32_t threads, int32_t slots) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (fused && !c->frame_ok)
    return fail(PTY_EUNSUP, "the fused per-frame path needs square 128 or 256 frames");
  if (threads != 0 && threads != 512 && threads != 1024)
    return fail(PTY_EINVAL, "threads must be 0 (keep), 512 or 1024");
  if (slots < 0) return fail(PTY_EINVAL, "slots < 0");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->use_frame = fused != 0;
  if (threads) c->frame_nt = threads;
  if (slots) c->frame_slots = slots;
  if (c->frame_slots > c->cap_col || c->frame_slots > c->cap_esup)
    return fail(PTY_EINVAL, "slots exceed the partial-sum capacity of this context");
  return PTY_OK;
}

int pty_set_graphs(pty_ctx* c, int32_t enable) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->graphs = enable != 0;
  if (!c->graphs) drop_graphs(c);
  return PTY_OK;
}

int pty_get_graphs(pty_ctx* c, int32_t* enabled, int32_t* fell_back) {
  if (!c || !enabled || !fell_back) return fail(PTY_EINVAL, "NULL argument");
  *enabled = c->graphs ? 1 : 0;
  *fell_back = c->graphs_fell_back ? 1 : 0;
  return PTY_OK;
}

int pty_set_chunk(pty_ctx* c, int32_t frames) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (frames < 0) return fail(PTY_EINVAL, "frames must be >= 0");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->chunk_frames = frames;
  drop_graphs(c);
  return PTY_OK;
}

int pty_set_row_gather(pty_ctx* c, int32_t enable) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->rowg = enable && c->rowg_ok;
  drop_graphs(c);
  return PTY_OK;
}

int pty_get_row_gather(pty_ctx* c, int32_t* supported, int32_t* enabled) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (supported) *supported = c->rowg_ok ? 1 : 0;
  if (enabled) *enabled = c->rowg ? 1 : 0;
  return PTY_OK;
}

int pty_set_launch(pty_ctx* c, int32_t rows_mult, int32_t cols_mult, int32_t cols_wg) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (rows_mult < 1 || rows_mult > 256 || cols_mult < 0 || cols_mult > 128)
    return fail(PTY_EINVAL, "rows_mult must be 1..256 and cols_mult 0..128");
  if (cols_wg != 0 && cols_wg != 256 && cols_wg != 512)
    return fail(PTY_EINVAL, "cols_wg must be 0 (automatic), 256 or 512");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->rows_mult = rows_mult;
  c->cols_mult = cols_mult;
  c->cols_wg = cols_wg;
  drop_graphs(c);
  return PTY_OK;
}

// Time every launch configuration on this scan (`trial_iters` iterations of
// `o` from the bound state, after one warm-up iteration) and keep the fastest.
// O, P, the DM exits and the background are restored afterwards, so the
// caller's next pty_run(begin = 1) starts from the state it set.
int pty_set_plan(pty_ctx* c, int32_t variant) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (variant < 0 || variant > PLAN_ALL) return fail(PTY_EINVAL, "plan mask must be 0..15");
  PTY_HIP(hipStreamSynchronize(c->stream));
  if (c->plan == variant) return PTY_OK;
  c->plan = variant;
  drop_graphs(c);
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_get_plan(pty_ctx* c, int32_t* variant, int32_t* has_alternative) {
  if (!c || !variant || !has_alternative) return fail(PTY_EINVAL, "NULL argument");
  *variant = c->plan;
  *has_alternative = (threads_per_line(plan_key(c->M, 1)) > 0 || threads_per_line(plan_key(c->K, 1)) > 0) ? 1 : 0;
  return PTY_OK;
}

int pty_tune(pty_ctx* c, const pty_opts* o, int32_t trial_iters, pty_tune_result* out) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_tune: not available on packed scans");
  if (!c || !o || !out || trial_iters < 1) return fail(PTY_EINVAL, "bad pty_tune arguments");
  PTY_TRY(check_bound(c, o->method >= 4));
  PTY_HIP(hipStreamSynchronize(c->stream));
  const size_t UV = (size_t)c->U * c->V, MK = c->MK(), NMK = c->NMK();
  struct Snap {
    float2 *O = nullptr, *P = nullptr, *F = nullptr;
    float* B = nullptr;
    ~Snap() { dev_free(O); dev_free(P); dev_free(F); dev_free(B); }
  } snap;
  PTY_TRY(dev_alloc(&snap.O, UV));
  PTY_TRY(dev_alloc(&snap.P, MK));
  PTY_TRY(copy_d2d(c, snap.O, c->O, UV * sizeof(float2)));
  PTY_TRY(copy_d2d(c, snap.P, c->P, MK * sizeof(float2)));
  float2* F0 = c->F;
  if (F0) {
    PTY_TRY(dev_alloc(&snap.F, NMK));
    PTY_TRY(copy_d2d(c, snap.F, F0, NMK * sizeof(float2)));
  }
  if (c->bg && c->bg_set) {
    PTY_TRY(dev_alloc(&snap.B, NMK));
    PTY_TRY(copy_d2d(c, snap.B, c->bg, NMK * sizeof(float)));
  }
  auto restore = [&]() -> int {
    PTY_TRY(copy_d2d(c, c->O, snap.O, UV * sizeof(float2)));
    PTY_TRY(copy_d2d(c, c->P, snap.P, MK * sizeof(float2)));
    if (snap.F && c->F) PTY_TRY(copy_d2d(c, c->F, snap.F, NMK * sizeof(float2)));
    if (snap.B) PTY_TRY(copy_d2d(c, c->bg, snap.B, NMK * sizeof(float)));
    return PTY_OK;
  };
  struct Cfg { int fused, nt, slots, rows, cols, cwg, chunk, rowg, plan; };
  std::vector<Cfg> cands;
  const int ch = c->chunk_frames, rg = c->rowg, pl = c->plan;
  cands.push_back({c->use_frame, c->frame_nt, c->frame_slots, c->rows_mult, c->cols_mult, c->cols_wg, ch, rg, pl});
  if (c->rowg_ok)
    cands.push_back({0, c->frame_nt, c->frame_slots, c->rows_mult, c->cols_mult, c->cols_wg, ch, !rg, pl});
  for (int rows : {2, 3, 4})
    for (int cols : {0, 2, 4}) cands.push_back({0, c->frame_nt, c->frame_slots, rows, cols, 256, ch, rg, pl});
  if (c->M == 256)
    for (int cols : {0, 1, 3}) cands.push_back({0, c->frame_nt, c->frame_slots, 3, cols, 512, ch, rg, pl});
  // the other line-FFT plans (radix split and threads per line) of each
  // stage, with the current launch shapes
  if (threads_per_line(plan_key(c->M, 1)) > 0 || threads_per_line(plan_key(c->K, 1)) > 0)
    for (int m = 0; m <= PLAN_ALL; ++m)
      // (the background model-error bit matters only to DM methods 4-6)
      if (m != pl && ((o->algorithm == 1 && o->method >= 4) || ((m ^ pl) & PLAN_BGERR) == 0))
        cands.push_back({0, c->frame_nt, c->frame_slots, c->rows_mult, c->cols_mult, c->cols_wg, ch, rg, m});
  {
    // frame blocks for the modulus step: whole stack, and 1/4 .. 2x the default block
    const size_t per = (size_t)c->M * c->K * 12, budget = (size_t)192 << 20;
    const int base = std::max<int>(1, (int)(budget / per));
    for (int chunk : {0, base / 4, base / 2, base, 2 * base})
      if (chunk < c->N && chunk != ch) cands.push_back({0, c->frame_nt, c->frame_slots, 3, 0, 256, chunk, rg, pl});
  }
  if (c->frame_ok)
    for (int nt : {512, 1024})
      for (int sl : {c->cus, 2 * c->cus})
        if (sl <= c->cap_col && sl <= c->cap_esup)
          cands.push_back({1, nt, sl, c->rows_mult, c->cols_mult, c->cols_wg, ch, rg, pl});
  auto apply = [&](const Cfg& g) -> int {
    c->use_frame = g.fused != 0;
    c->frame_nt = g.nt;
    c->frame_slots = g.slots;
    c->rows_mult = g.rows;
    c->cols_mult = g.cols;
    c->cols_wg = g.cwg;
    c->chunk_frames = g.chunk;
    c->rowg = g.rowg != 0;
    drop_graphs(c);
    return pty_set_plan(c, g.plan);
  };
  double best = 1e300, first = 0.0;
  int besti = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    PTY_TRY(apply(cands[i]));
    PTY_TRY(restore());
    PTY_TRY(pty_run(c, o, 0, 1, 1, nullptr));
    PTY_TRY(restore());
    PTY_HIP(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    PTY_TRY(pty_run(c, o, 0, trial_iters, 1, nullptr));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
        trial_iters;
    if (i == 0) first = ms;
    if (ms < best) { best = ms; besti = (int)i; }
  }
  PTY_TRY(apply(cands[besti]));
  PTY_TRY(restore());
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->H_valid = c->G_valid = false;
  c->have_metric_state = false;
  const Cfg& g = cands[besti];
  *out = pty_tune_result{g.fused, g.nt, g.slots, g.rows, g.cols, g.cwg, g.chunk, g.rowg, g.plan,
                         (int32_t)cands.size(), (float)best, (float)first};
  return PTY_OK;
}

int pty_batch_create(pty_batch** out, int32_t n, pty_ctx* const* ctxs) {
  if (!out || n < 1 || !ctxs) return fail(PTY_EINVAL, "bad pty_batch_create arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(PTY_EINVAL, "NULL context in the batch");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(PTY_EINVAL, "a context appears twice in the batch");
  }
  pty_batch* b = new pty_batch();
  b->ctx.assign(ctxs, ctxs + n);
  // lanes: as many as the process's hardware queues (GPU_MAX_HW_QUEUES,
  // 4 by default), more would only alias onto them; PTY_BATCH_LANES overrides
  int maxl = 4;
  if (const char* e_q = getenv("GPU_MAX_HW_QUEUES")) maxl = std::max(1, std::min(8, atoi(e_q)));
  if (const char* e_l = getenv("PTY_BATCH_LANES")) maxl = std::max(1, std::min(8, atoi(e_l)));
  const int nl = std::min(n, maxl);
  b->lanes.resize(nl, nullptr);
  b->join.resize(nl, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  for (int l = 0; l < nl && e == hipSuccess; ++l) {
    e = hipStreamCreateWithFlags(&b->lanes[l], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    delete b;
    return fail(PTY_EHIP, std::string("pty_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return PTY_OK;
}

int pty_batch_destroy(pty_batch* b) {
  if (!b) return PTY_OK;
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  for (auto l : b->lanes)
    if (l) (void)hipStreamSynchronize(l);
  delete b;
  return PTY_OK;
}

int pty_batch_get_graphs(pty_batch* b, int32_t* enabled, int32_t* fell_back) {
  if (!b || !enabled || !fell_back) return fail(PTY_EINVAL, "NULL argument");
  *enabled = b->graphs ? 1 : 0;
  *fell_back = b->fell_back ? 1 : 0;
  return PTY_OK;
}

int pty_batch_set_graphs(pty_batch* b, int32_t enable) {
  if (!b) return fail(PTY_EINVAL, "batch is NULL");
  b->graphs = enable != 0;
  if (!b->graphs) b->drop();
  return PTY_OK;
}

// opts: one pty_opts per scan; metrics_host: one pointer per scan (3 * iters
// doubles, or NULL).  Every scan advances `iters` iterations from iter0.
int pty_batch_run(pty_batch* b, const pty_opts* opts, int64_t iter0, int32_t iters, int32_t begin,
                  double* const* metrics_host) {
  if (!b || !opts) return fail(PTY_EINVAL, "NULL argument");
  if (iters < 0) return fail(PTY_EINVAL, "iters < 0");
  const int n = (int)b->ctx.size(), nl = (int)b->lanes.size();
  std::vector<hipStream_t> user(n);
  std::vector<bool> prof(n);
  for (int s = 0; s < n; ++s) {
    user[s] = b->ctx[s]->stream;
    prof[s] = b->ctx[s]->prof_on;
  }
  struct Restore {
    pty_batch* b; std::vector<hipStream_t>& user; std::vector<bool>& prof;
    ~Restore() {
      for (size_t s = 0; s < b->ctx.size(); ++s) {
        b->ctx[s]->stream = user[s];
        b->ctx[s]->use_slot = false;
        b->ctx[s]->prof_on = prof[s];
      }
    }
  } restore{b, user, prof};
  for (int s = 0; s < n; ++s) {
    pty_ctx* c = b->ctx[s];
    PTY_HIP(hipStreamSynchronize(user[s]));
    c->stream = b->lanes[s % nl];
    c->prof_on = false;   // stage events are not recorded in batches
    PTY_TRY(run_prepare(c, &opts[s], iter0, iters, begin));
  }
  for (int l = 0; l < nl; ++l) PTY_HIP(hipStreamSynchronize(b->lanes[l]));
  if (!b->graphs) {
    for (int32_t k = 0; k < iters; ++k)
      for (int s = 0; s < n; ++s)
        PTY_TRY(run_one(b->ctx[s], opts[s], iter0 + k, metrics_at(b->ctx[s], k)));
  } else {
    const int zero = 0;
    for (int s = 0; s < n; ++s) {
      pty_ctx* c = b->ctx[s];
      if (!c->slot) PTY_TRY(dev_alloc(&c->slot, 1));
      PTY_HIP(hipMemcpy(c->slot, &zero, sizeof(int), hipMemcpyHostToDevice));
      c->use_slot = true;
    }
    for (int32_t k = 0; k < iters; ++k) {
      const int64_t it = iter0 + k;
      if (!b->graphs) {   // a capture failed earlier in this run: continue eagerly
        for (int s = 0; s < n; ++s)
          PTY_TRY(run_one(b->ctx[s], opts[s], it, metrics_at(b->ctx[s], k)));
        continue;
      }
      std::vector<uint64_t> key;
      for (int s = 0; s < n; ++s) {
        std::vector<uint64_t> ks = iter_key(b->ctx[s], opts[s], it);
        key.push_back(ks.size());
        key.insert(key.end(), ks.begin(), ks.end());
      }
      pty_batch::Entry* hit = nullptr;
      for (auto& e : b->cache)
        if (e.key == key) { hit = &e; break; }
      if (!hit) {
        std::vector<HostIterState> before;
        for (int s = 0; s < n; ++s) before.push_back(save_state(b->ctx[s]));
        PTY_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
        int r = PTY_OK;
        hipError_t he = hipEventRecord(b->fork, b->origin);
        for (int l = 0; l < nl && he == hipSuccess; ++l) he = hipStreamWaitEvent(b->lanes[l], b->fork, 0);
        for (int s = 0; s < n && r == PTY_OK && he == hipSuccess; ++s)
          r = run_one(b->ctx[s], opts[s], it, b->ctx[s]->metrics);
        for (int l = 0; l < nl && he == hipSuccess; ++l) {
          he = hipEventRecord(b->join[l], b->lanes[l]);
          if (he == hipSuccess) he = hipStreamWaitEvent(b->origin, b->join[l], 0);
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->origin, &graph);
        if (r != PTY_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return r;
        }
        hipGraphExec_t exec = nullptr;
        hipError_t ei = he != hipSuccess ? he : ec;
        if (ei == hipSuccess) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (ei != hipSuccess || capture_fail_forced()) {
          // nothing of the captured iteration ran: undo its host-side effects,
          // drop the graphs and run this and the remaining iterations eagerly
          if (exec) (void)hipGraphExecDestroy(exec);
          PTY_HIP(hipStreamSynchronize(b->origin));
          for (int s = 0; s < n; ++s) {
            load_state(b->ctx[s], before[s]);
            b->ctx[s]->use_slot = false;
          }
          b->drop();
          b->graphs = false;
          b->fell_back = true;
          for (int s = 0; s < n; ++s)
            PTY_TRY(run_one(b->ctx[s], opts[s], it, metrics_at(b->ctx[s], k)));
          continue;
        }
        pty_batch::Entry e{key, exec, {}};
        for (int s = 0; s < n; ++s) e.st.push_back(save_state(b->ctx[s]));
        b->cache.push_back(std::move(e));
        hit = &b->cache.back();
      } else {
        for (int s = 0; s < n; ++s) load_state(b->ctx[s], hit->st[s]);
      }
      PTY_HIP(hipGraphLaunch(hit->exec, b->origin));
    }
    PTY_HIP(hipStreamSynchronize(b->origin));
  }
  for (int s = 0; s < n; ++s) PTY_TRY(run_tail(b->ctx[s], iters));
  for (int s = 0; s < n; ++s) {
    pty_ctx* c = b->ctx[s];
    if (metrics_host && metrics_host[s] && iters > 0)
      PTY_HIP(hipMemcpyAsync(metrics_host[s], c->metrics, 3 * (size_t)c->nscan * iters * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  }
  for (int l = 0; l < nl; ++l) PTY_HIP(hipStreamSynchronize(b->lanes[l]));
  return PTY_OK;
}

int pty_profile(pty_ctx* c, int32_t enable) {
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  PTY_HIP(hipStreamSynchronize(c->stream));
  c->prof_used = 0;
  c->prof_on = enable != 0;
  for (int k = 0; k < PTY_NSTAGE; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_cnt[k] = 0;
  }
  return PTY_OK;
}

int pty_profile_read(pty_ctx* c, double* ms, int64_t* cnt) {
  if (!c || !ms || !cnt) return fail(PTY_EINVAL, "NULL argument");
  for (int k = 0; k < PTY_NSTAGE; ++k) {
    ms[k] = c->prof_ms[k];
    cnt[k] = c->prof_cnt[k];
  }
  return PTY_OK;
}

int pty_intensity(pty_ctx* c, float* I) {
  if (!c || !I) return fail(PTY_EINVAL, "NULL argument");
  if (!c->O || !c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
  const bool dm = c->last_algo == 1 && c->have_metric_state;
  const float2* Ob = dm ? c->Os : c->O;
  const float2* Pb = dm ? c->Ps : c->P;
  if (frame_path(c)) {
    FrameArgs f = frame_args(c);
    f.O = Ob;
    f.P = Pb;
    f.outI = I;
    PTY_TRY(frame<FR_INTENS>(c, f));
  } else {
    PTY_TRY(rows_fwd<SRC_OP>(c, Ob, Pb, nullptr, c->X, c->N * c->M));
    ColArgs a = col_args(c);
    a.X = c->X;
    a.outI = I;
    PTY_TRY(cols<COL_INTENS>(c, a));
  }
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_heatmap(pty_ctx* c, double* H, int32_t* valid) {
  if (!c || !H || !valid) return fail(PTY_EINVAL, "NULL argument");
  *valid = c->H_valid ? 1 : 0;
  if (c->H_valid) PTY_TRY(copy_d2d(c, H, c->H, (size_t)c->U * c->V * sizeof(double)));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_background(pty_ctx* c, float* b) {
  if (!c || !b) return fail(PTY_EINVAL, "NULL argument");
  if (!c->bg_set) return fail(PTY_ESTATE, "no background state");
  const float* src = c->bg;  // first frame of the state before any mean
  if (c->last_algo == 1 && c->have_metric_state) src = c->bbar_m;
  else if (c->last_algo == 0 && c->era_bg_shared) src = c->bbar;
  PTY_TRY(copy_d2d(c, b, src, c->MK() * sizeof(float)));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_set_heatmap(pty_ctx* c, const double* H, const double* G) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_set_heatmap: not available on packed scans");
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (H) PTY_TRY(copy_d2d(c, c->H, H, (size_t)c->U * c->V * sizeof(double)));
  if (G) PTY_TRY(copy_d2d(c, c->G, G, c->MK() * sizeof(double)));
  c->H_valid = H != nullptr;
  c->G_valid = G != nullptr;
  c->os_next = c->os_next_op = false;
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_set_iterate(pty_ctx* c, const pty_cfloat* f, const pty_cfloat* O_ref, const pty_cfloat* P_ref) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_set_iterate: not available on packed scans");
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (!c->O || !c->P) return fail(PTY_ESTATE, "pty_set_state has not been called");
  if ((O_ref == nullptr) != (P_ref == nullptr))
    return fail(PTY_EINVAL, "O_ref and P_ref go together");
  if (f) {
    PTY_TRY(ensure_F(c));
    PTY_TRY(copy_d2d(c, c->F, f, c->NMK() * sizeof(float2)));
    c->dm_ready = true;
    c->f_spec = c->dmf_lead = false;
  }
  const size_t UV = (size_t)c->U * c->V;
  PTY_TRY(copy_d2d(c, c->Obak, O_ref ? (const void*)O_ref : (const void*)c->O, UV * sizeof(float2)));
  PTY_TRY(copy_d2d(c, c->Pbak, P_ref ? (const void*)P_ref : (const void*)c->P, c->MK() * sizeof(float2)));
  c->have_metric_state = false;
  c->os_next = c->os_next_op = false;
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_get_iterate(pty_ctx* c, pty_cfloat* f, float* b) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_get_iterate: not available on packed scans");
  if (!c) return fail(PTY_EINVAL, "ctx is NULL");
  if (f) {
    if (!c->F || !c->dm_ready) return fail(PTY_ESTATE, "no DM iterate (run DM or pty_set_iterate first)");
    float2* fo = reinterpret_cast<float2*>(f);
    PTY_TRY(copy_d2d(c, fo, c->F, c->NMK() * sizeof(float2)));
    if (c->f_spec) {
      // the iterate is held as spectra: inverse 2-D transform in the caller's buffer
      ColArgs a = col_args(c);
      a.X = fo;
      a.nframes = c->N;
      a.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
      PTY_TRY(cols<COL_INV>(c, a));
      PTY_TRY(rows_inv<DST_STORE>(c, fo, fo, nullptr, nullptr, 1.0f, nullptr, c->N * c->M));
    }
  }
  if (b) {
    if (!c->bg_set) return fail(PTY_ESTATE, "no background state");
    if (c->last_algo == 0 && c->era_bg_shared) {
      // ERA keeps the frame mean once it has been taken (era.py:265-270)
      hipLaunchKernelGGL(k_broadcast, dim3(4096), dim3(WG), 0, c->stream, b, (const float*)c->bbar,
                         c->MK(), c->NMK());
      PTY_LAUNCHED("k_broadcast");
    } else {
      PTY_TRY(copy_d2d(c, b, c->bg, c->NMK() * sizeof(float)));
    }
  }
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_fft2(pty_ctx* c, pty_cfloat* data, int32_t nframes, int32_t inverse) {
  if (!c || !data) return fail(PTY_EINVAL, "NULL argument");
  if (nframes < 0) return fail(PTY_EINVAL, "nframes < 0");
  if (nframes == 0) return PTY_OK;
  float2* d = reinterpret_cast<float2*>(data);
  ColArgs a = col_args(c);
  a.X = d;
  a.nframes = nframes;
  a.chunk = std::max(1, std::min(16, (int)((long)nframes * c->col_blocks / 2048)));
  if (!inverse) {
    PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, d, d, nframes * c->M));
    PTY_TRY(cols<COL_FWD>(c, a));
  } else {
    PTY_TRY(cols<COL_INV>(c, a));
    PTY_TRY(rows_inv<DST_STORE>(c, d, d, nullptr, nullptr, 1.0f, nullptr, nframes * c->M));
  }
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_make_exits(pty_ctx* c, const pty_cfloat* O, const pty_cfloat* P, pty_cfloat* ex) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_make_exits: not available on packed scans");
  if (!c || !O || !P || !ex) return fail(PTY_EINVAL, "NULL argument");
  SharedProbe shared(c);
  hipLaunchKernelGGL(k_make_exits, dim3(4096), dim3(WG), 0, c->stream, c->g,
                     reinterpret_cast<const float2*>(O), reinterpret_cast<const float2*>(P),
                     reinterpret_cast<float2*>(ex));
  PTY_LAUNCHED("k_make_exits");
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_hio_modulus(pty_ctx* c, pty_cfloat* f, pty_cfloat* m, const pty_cfloat* O,
                    const pty_cfloat* P, float w, float beta, float alpha, double* emod) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_hio_modulus: not available on packed scans");
  if (!c || !f || !m || !O || !P) return fail(PTY_EINVAL, "NULL argument");
  if (!std::isfinite(w)) return fail(PTY_EINVAL, "w must be finite");
  if (f == m) return fail(PTY_EINVAL, "f and m must be different buffers");
  if (!c->amp) return fail(PTY_ESTATE, "pty_set_data has not been called");
  if (!(beta != 0.0f) || !std::isfinite(beta)) return fail(PTY_EINVAL, "beta must be finite and non-zero");
  SharedProbe shared(c);
  float2* F = reinterpret_cast<float2*>(f);
  float2* Mb = reinterpret_cast<float2*>(m);
  const float2* Ob = reinterpret_cast<const float2*>(O);
  const float2* Pb = reinterpret_cast<const float2*>(P);
  const int nrows = c->N * c->M;
  // FFT rows of f = F + w O P into m; columns with the modulus step and eMod
  // partials; inverse rows with both linear combinations applied on the way out
  PTY_TRY(rows_fwd<SRC_HIO>(c, Ob, Pb, F, Mb, nrows, w));
  ColArgs a = col_args(c);
  a.X = Mb;
  a.amp = c->amp;
  a.amp_stride = (long)c->MK();
  a.alpha = alpha;
  a.part = c->part_emod;
  int ncol = 0;
  PTY_TRY(cols<COL_ERA1>(c, a, &ncol));
  const double binv = 1.0 / (double)beta;
  const HioMix hm{F, Ob, Pb, w, (float)(1.0 + binv), (float)binv, beta};
  PTY_TRY(rows_inv<DST_HIO>(c, Mb, Mb, nullptr, nullptr, 1.0f, nullptr, nrows, nullptr, hm));
  double s = 0.0;
  PTY_TRY(sum_region(c, c->part_emod, ncol, &s));
  if (emod) *emod = s;
  return PTY_OK;
}

int pty_add_exits(pty_ctx* c, const pty_cfloat* O, const pty_cfloat* P, float scale, pty_cfloat* f) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_add_exits: not available on packed scans");
  if (!c || !O || !P || !f) return fail(PTY_EINVAL, "NULL argument");
  SharedProbe shared(c);
  hipLaunchKernelGGL(k_add_exits, dim3(4096), dim3(WG), 0, c->stream, c->g,
                     reinterpret_cast<const float2*>(O), reinterpret_cast<const float2*>(P), scale,
                     reinterpret_cast<float2*>(f));
  PTY_LAUNCHED("k_add_exits");
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_pmod1(pty_ctx* c, pty_cfloat* exits, float alpha, double* emod) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_pmod1: not available on packed scans");
  if (!c || !exits) return fail(PTY_EINVAL, "NULL argument");
  if (!c->amp) return fail(PTY_ESTATE, "pty_set_data has not been called");
  float2* d = reinterpret_cast<float2*>(exits);
  const int nrows = c->N * c->M;
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, d, d, nrows));
  ColArgs a = col_args(c);
  a.X = d;
  a.amp = c->amp;
  a.amp_stride = (long)c->MK();
  a.alpha = alpha;
  a.part = c->part_emod;
  int ncol = 0;
  PTY_TRY(cols<COL_ERA1>(c, a, &ncol));
  PTY_TRY(rows_inv<DST_STORE>(c, d, d, nullptr, nullptr, 1.0f, nullptr, nrows));
  double s = 0.0;
  PTY_TRY(sum_region(c, c->part_emod, ncol, &s));
  if (emod) *emod = s;
  return PTY_OK;
}

int pty_pmod1_noalloc(pty_cfloat* exits, const float* amp, const uint8_t* mask, int32_t N, int32_t M,
                      int32_t K, float alpha, double* emod, void* stream) {
  if (!exits || !amp || !emod) return fail(PTY_EINVAL, "NULL argument");
  if (N <= 0) return fail(PTY_EINVAL, "N must be positive");
  if (!pty_supported_frame(M) || !pty_supported_frame(K))
    return fail(PTY_EUNSUP, "frame sides must be one of 16, 32, 64, 96, 128, 192, 256 (got " +
                                std::to_string(M) + " x " + std::to_string(K) + ")");
  if ((size_t)N * M * K > ((size_t)1 << 34))
    return fail(PTY_EINVAL, "N x M x K above the maximum of 2^34 frame pixels");
  int dev = 0;
  PTY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEV) return fail(PTY_EINVAL, "device index beyond 63");
  NoallocDevice* d = nullptr;
  PTY_TRY(noalloc_setup(dev, &d));
  // the column step takes up to 128 work-groups per compute unit, each with one partial
  if ((size_t)128 * d->cus > NA_PARTS)
    return fail(PTY_EINVAL, "device has more compute units than the scratch's eMod partials cover");

  // a launch description only: nothing of it is allocated
  pty_ctx c;
  c.N = N; c.M = M; c.K = K; c.U = M; c.V = K;
  c.stream = static_cast<hipStream_t>(stream);
  c.g.N = N; c.g.M = M; c.g.K = K; c.g.U = M; c.g.V = K;
  c.g.off = nullptr;   // the plain-frame row kernels read no window offsets
  c.cus = d->cus;
  c.col_chunk = 1;
  c.cap_col = (int)NA_PARTS;
  c.twM = reinterpret_cast<float2*>(d->base + NA_OFF_TW) + noalloc_tw_offset(M);
  c.twK = reinterpret_cast<float2*>(d->base + NA_OFF_TW) + noalloc_tw_offset(K);
  double* parts = reinterpret_cast<double*>(d->base);
  c.mask = mask;
  c.maskbits = mask ? reinterpret_cast<uint32_t*>(d->base + NA_OFF_MASK) : nullptr;
  if (mask) {
    const int T = threads_per_line(M);
    const int grid = (T * K + WG - 1) / WG;
    switch (M) {
#define C(Lv)                                                                                  \
  case Lv:                                                                                     \
    hipLaunchKernelGGL(k_pack_mask<Lv>, dim3(grid), dim3(WG), 0, c.stream, mask, K, c.maskbits); \
    break;
      PTY_SIZES(C)
#undef C
      default: return fail(PTY_EUNSUP, "unsupported M");
    }
    PTY_LAUNCHED("k_pack_mask");
  }
  float2* x = reinterpret_cast<float2*>(exits);
  PTY_TRY(rows_fwd<SRC_F>(&c, nullptr, nullptr, x, x, N * M));
  ColArgs a = col_args(&c);
  a.X = x;
  a.amp = amp;
  a.amp_stride = (long)c.MK();
  a.alpha = alpha;
  a.part = parts;
  int ncol = 0;
  PTY_TRY(cols<COL_ERA1>(&c, a, &ncol));
  PTY_TRY(rows_inv<DST_STORE>(&c, x, x, nullptr, nullptr, 1.0f, nullptr, N * M));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, c.stream, (const double*)parts, ncol, emod);
  PTY_LAUNCHED("k_sum");
  return PTY_OK;
}

int pty_psup_O(pty_ctx* c, const pty_cfloat* ex, const pty_cfloat* P, pty_cfloat* O_out, double* H,
               int32_t fresh, float alpha) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_psup_O: not available on packed scans");
  if (!c || !ex || !P || !O_out || !H) return fail(PTY_EINVAL, "NULL argument");
  SharedProbe shared(c);
  PTY_TRY(gather(c, reinterpret_cast<const float2*>(ex), reinterpret_cast<const float2*>(P), nullptr,
                 reinterpret_cast<float2*>(O_out), H, fresh ? 1 : 0, 1, alpha, false, false));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

int pty_psup_P(pty_ctx* c, const pty_cfloat* ex, const pty_cfloat* O, pty_cfloat* P_out, double* G,
               int32_t fresh, float alpha) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_psup_P: not available on packed scans");
  if (!c || !ex || !O || !P_out || !G) return fail(PTY_EINVAL, "NULL argument");
  PTY_TRY(probe_update(c, reinterpret_cast<const float2*>(ex), reinterpret_cast<const float2*>(O),
                       nullptr, reinterpret_cast<float2*>(P_out), G, fresh ? 1 : 0, 1, alpha, false));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

// Back-projection initial object (back_projection.py:4-88): exits_n =
// IFFT2(mask * amp_n * FFT2(P) / |FFT2(P)|), O = psup_O(exits, P) with a fresh
// heatmap H.  Uses the bound amplitudes / mask (pty_set_data); exits_out
// (N x M x K, optional) receives the exit waves.
int pty_back_projection(pty_ctx* c, const pty_cfloat* P, pty_cfloat* exits_out, pty_cfloat* O_out,
                        double* H, float alpha) {
  if (c && c->nscan > 1) return fail(PTY_EUNSUP, "pty_back_projection: not available on packed scans");
  if (!c || !P || !O_out || !H) return fail(PTY_EINVAL, "NULL argument");
  if (!c->amp) return fail(PTY_ESTATE, "pty_set_data has not been called");
  SharedProbe shared(c);
  const float2* Pd = reinterpret_cast<const float2*>(P);
  PTY_TRY(rows_fwd<SRC_F>(c, nullptr, nullptr, Pd, c->Xp, c->M));
  ColArgs a = col_args(c);
  a.X = c->Xp;
  a.nframes = 1;
  a.chunk = 1;
  PTY_TRY(cols<COL_FWD>(c, a));
  hipLaunchKernelGGL(k_bp_spectra, dim3(4096), dim3(WG), 0, c->stream, c->X, (const float*)c->amp,
                     c->mask, (const float2*)c->Xp, c->MK(), c->NMK());
  PTY_LAUNCHED("k_bp_spectra");
  ColArgs b = col_args(c);
  b.X = c->X;
  b.nframes = c->N;
  b.chunk = std::max(1, std::min(16, (int)((long)c->N * c->col_blocks / 2048)));
  PTY_TRY(cols<COL_INV>(c, b));
  PTY_TRY(rows_inv<DST_STORE>(c, c->X, c->X, nullptr, nullptr, 1.0f, nullptr, c->N * c->M));
  if (exits_out) PTY_TRY(copy_d2d(c, exits_out, c->X, c->NMK() * sizeof(float2)));
  PTY_TRY(gather(c, c->X, Pd, nullptr, reinterpret_cast<float2*>(O_out), H, 1, 1, alpha, false,
                 false));
  PTY_HIP(hipStreamSynchronize(c->stream));
  return PTY_OK;
}

}  // extern "C"
