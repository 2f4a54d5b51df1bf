// Host side of the C ABI declared in include/simple_ts.h: device state of one
// case and the entry points that replace the reference's call sites in
// main() (SIMPLE_TS_MPI.cpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "simple_ts.h"
#include "sts_device.h"
#include "sts_internal.h"
#include "sts_launch.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_OK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(STS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

constexpr int MAX_SWEEPS = 62;
constexpr int NRED = 2 + 2 * MAX_SWEEPS;

}  // namespace

struct sts_ctx {
  int Nx = 0, Ny = 0, Na = 0;
  int device = -1;  // the HIP device current at sts_create
  hipStream_t stream = nullptr;
  KGrid g{};
  double *hx = nullptr, *hy = nullptr;
  uint8_t *m = nullptr;
  uint16_t *nb = nullptr;  // neighbour fluid bits (NB_*), from the mask
  uint32_t *vol = nullptr;
  DevBody *bodies = nullptr;
  double *f[STS_NFIELDS] = {};
  double *p_alt = nullptr, *T_alt = nullptr;  // loop-3 ping-pong partners of p, Temper
  unsigned long long *red = nullptr;          // slotted residual maxima as IEEE bit patterns
  unsigned long long *red_out = nullptr;      // red folded over its slots
  unsigned long long *lred = nullptr;         // slotted maxima of the current outer iteration (loop 2)
  double *probe = nullptr;
  unsigned long long *h_red = nullptr;        // pinned host mirror of red (+ probe)
  double *partial = nullptr;                  // continuity norms, one triple per block
  double *cont_stat = nullptr, *cont = nullptr;
  // device-side loop 2
  LoopCtrl *ctrl = nullptr;
  LoopCtrl *h_ctrl = nullptr;          // pinned mirror
  hipEvent_t ev_ctrl = nullptr;        // after the copy of ctrl to h_ctrl
  unsigned long long *log = nullptr;   // 6 residual maxima per outer iteration
  unsigned long long *h_log = nullptr; // pinned mirror of log
  int log_cap = 0;
  int last_iters = 0;
  // FP32 copy of the state for the mixed-precision time step (precision 32)
  bool mixed = false;
  KGridT<float> g32{};
  float *hx32 = nullptr, *hy32 = nullptr;
  float *f32[STS_NFIELDS] = {};
  float *p_alt32 = nullptr, *T_alt32 = nullptr;
  int last_iters32 = 0, last_iters64m = 0;
  int last_total = -1;  // outer iterations of the previous mixed or warm-started step (-1: none)
  // guarded predicted start: running mean of the FP64 iterations of steps
  // from x^n (-1: none yet), pause left and next pause length, counters
  double plain_mean = -1.0;
  int pred_pause = 0, pred_backoff = 1;
  int pred_degree = -1, pred_wins = 0;  // degree in use (-1: the requested one), successes since a change
  long long pstat[5] = {0, 0, 0, 0, 0};
  long long fluid_cells = 0;
  WallOp *wall_ops = nullptr;  // the wall pass's writes, planned once from the mask (k_wall_ops)
  int n_wall_ops = 0;
  int n_wall_var = 0;  // the first n_wall_var of them depend on the state; the rest write constants
  // drag coefficients: host copy of vol_inf_fb, per-body cell lists, buffers
  std::vector<uint32_t> h_vol;
  std::vector<unsigned *> drag_cells;
  std::vector<int> drag_n;
  double *drag_terms = nullptr, *drag_out = nullptr;
  size_t drag_cap = 0;
  double *p_ws = nullptr, *T_ws = nullptr;  // predicted starting guess (warm start)
  // the corrector (sts_set_corrector): the kept start of loop 2 (u, v, p,
  // T), the network, its normalisation, kept / rejected refinements
  double *xs[4] = {};
  sts_unet *corr_net = nullptr;
  sts_unet_corr_norm corr_norm{};
  long long cstat[3] = {0, 0, 0};  // kept, rejected, skipped (paused)
  // pause after a refinement that did not pay (sts_set_corrector_pause): the
  // next corr_pause predicted steps skip the corrector; corr_backoff doubles
  // up to corr_max_pause on each further one and returns to 1 when one pays.
  // A refinement pays when it is kept and its step then needs at least one
  // outer iteration fewer than corr_ref_mean, the running mean of the
  // predicted steps run without one (< 0: none yet).
  int corr_max_pause = 0, corr_pause = 0, corr_backoff = 1;
  bool corr_kept_now = false;  // this step's start is a kept refinement
  double corr_ref_mean = -1.0;
  // Gamma_* and D_*: whether forming them from sqrt_T, the wall temperatures
  // and the mask gives the values the reference holds (true after a flux and
  // diffusion pass of the library, until those inputs or the arrays are set
  // another way), and whether the arrays lag behind (the loops' passes write
  // only the fluxes and the thermal coefficients, and the assembly forms the
  // rest itself)
  bool diff_inline_ok = false, diff_stale = false;
  ExtrapHist hist{};  // u, v, p, T of the last converged steps (extrapolated starts)
  int n_hist = 0;     // stored levels
  // Anderson acceleration of the FP64 loop 2: ring of FP32 differences
  // (aa_cap slots) and the per-context vectors
  float *aa_dG[MAX_AA] = {}, *aa_dF[MAX_AA] = {};
  int aa_cap = 0;
  AABuf aa{};
  bool fuse_sweeps = true;  // pairs of loop-3 sweeps in one k_sweep2 pass
  // kernel timing: event pairs around the launches of the kernels in the set
  // `timed` (bit per kernel id); a pair launched as outer iteration k of a
  // device-side loop is dropped once that loop has stopped before k (such a
  // launch returns at once)
  unsigned timed = 0;
  struct TimedPair {
    int kernel, loop, k;
  };
  std::vector<hipEvent_t> ev;  // 2 per pending launch
  std::vector<TimedPair> ev_tag;
  int ev_used = 0;
  int loop_seq = 0;
  bool loop_open = false;
  long long timed_launches[2 * STS_K_COUNT] = {};
  double timed_ms[2 * STS_K_COUNT] = {};

  KFields fields() const {
    KFields k;
    for (int q = 0; q < STS_NFIELDS; q++) k.f[q] = f[q];
    return k;
  }
  KFieldsT<float> fields32() const {
    KFieldsT<float> k;
    for (int q = 0; q < STS_NFIELDS; q++) k.f[q] = f32[q];
    return k;
  }
};

extern "C" const char *sts_last_error(void) { return g_err.c_str(); }

namespace {
constexpr int EV_POOL = 4096;

// Folds the durations of completed event pairs into the per-kernel totals.
int drain_events(sts_ctx *c) {
  if (c->ev_used == 0) return STS_OK;
  HIP_OK(hipEventSynchronize(c->ev[2 * (c->ev_used - 1) + 1]));
  for (int k = 0; k < c->ev_used; k++) {
    const int id = c->ev_tag[k].kernel;
    if (id < 0) continue;
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
    c->timed_ms[id] += ms;
    c->timed_launches[id]++;
  }
  c->ev_used = 0;
  return STS_OK;
}

int grow_events(sts_ctx *c, int pairs) {
  const size_t old = c->ev.size();
  c->ev.resize(old + 2 * (size_t)pairs);
  c->ev_tag.resize(c->ev.size() / 2);
  // timing only: no system-scope fence (cache writeback) at each record
  for (size_t q = old; q < c->ev.size(); q++)
    HIP_OK(hipEventCreateWithFlags(&c->ev[q], hipEventDisableSystemFence));
  return STS_OK;
}

// Records the start event when kernel `id` is timed; returns the pair slot or
// -1.  k: the launch's outer iteration inside an open device-side loop, or
// -1 for a launch that always runs.
int timed_begin(sts_ctx *c, int id, int k = -1) {
  if (!(c->timed >> id & 1u)) return -1;
  if (2 * c->ev_used == (int)c->ev.size()) {
    // while a loop is open its pairs are kept until its stop is known
    if (c->loop_open ? grow_events(c, EV_POOL) : drain_events(c)) return -1;
  }
  const int q = c->ev_used++;
  c->ev_tag[q] = sts_ctx::TimedPair{id, k >= 0 ? c->loop_seq : -1, k};
  (void)hipEventRecord(c->ev[2 * q], c->stream);
  return q;
}

void timed_end(sts_ctx *c, int q) {
  if (q >= 0) (void)hipEventRecord(c->ev[2 * q + 1], c->stream);
}

// A device-side loop opens; when it closes at iteration `stop`, the pairs of
// its launches past the stop are dropped.
void timed_loop_open(sts_ctx *c) {
  c->loop_seq++;
  c->loop_open = true;
}
void timed_loop_close(sts_ctx *c, int stop) {
  for (int q = 0; q < c->ev_used; q++) {
    sts_ctx::TimedPair &t = c->ev_tag[q];
    if (t.loop == c->loop_seq && t.k > stop) t.kernel = -1;
  }
  c->loop_open = false;
}
}  // namespace

extern "C" int sts_time_kernel(sts_ctx *c, int kernel) {
  if (!c || kernel < -1 || kernel >= 2 * STS_K_COUNT) return fail(STS_ERR_ARG, "bad kernel id");
  if (c->ev.empty()) {
    int rc = grow_events(c, EV_POOL);
    if (rc) return rc;
  }
  int rc = drain_events(c);
  if (rc) return rc;
  if (kernel < 0) {
    c->timed = 0;
    for (int q = 0; q < 2 * STS_K_COUNT; q++) {
      c->timed_launches[q] = 0;
      c->timed_ms[q] = 0;
    }
    return STS_OK;
  }
  c->timed |= 1u << kernel;
  c->timed_launches[kernel] = 0;
  c->timed_ms[kernel] = 0;
  return STS_OK;
}

extern "C" int sts_kernel_time(sts_ctx *c, int kernel, long long *launches, double *total_ms) {
  if (!c || kernel < 0 || kernel >= 2 * STS_K_COUNT) return fail(STS_ERR_ARG, "bad kernel id");
  int rc = drain_events(c);
  if (rc) return rc;
  if (launches) *launches = c->timed_launches[kernel];
  if (total_ms) *total_ms = c->timed_ms[kernel];
  return STS_OK;
}

static unsigned bc_bits(const sts_grid *g) {
  unsigned b = 0;
  b |= g->dudx_0_xb ? BC_DUDX_XB : 0u;
  b |= g->durhodx_0_xb ? BC_DURHODX_XB : 0u;
  b |= g->drhodt_xb ? BC_DRHODT_XB : 0u;
  b |= g->dudx_0_xe ? BC_DUDX_XE : 0u;
  b |= g->durhodx_0_xe ? BC_DURHODX_XE : 0u;
  b |= g->drhodt_xe ? BC_DRHODT_XE : 0u;
  b |= g->u_Min_Mout_xe ? BC_UMINMOUT_XE : 0u;
  b |= g->dudy_0_yb ? BC_DUDY_YB : 0u;
  b |= g->dudy_0_ye ? BC_DUDY_YE : 0u;
  b |= g->dvdx_0_xb ? BC_DVDX_XB : 0u;
  b |= g->dvdx_0_xe ? BC_DVDX_XE : 0u;
  b |= g->dvdy_0_yb ? BC_DVDY_YB : 0u;
  b |= g->dvdy_0_ye ? BC_DVDY_YE : 0u;
  b |= g->dpdx_0_xb ? BC_DPDX_XB : 0u;
  b |= g->dpdx_0_xe ? BC_DPDX_XE : 0u;
  b |= g->dpdy_0_yb ? BC_DPDY_YB : 0u;
  b |= g->dpdy_0_ye ? BC_DPDY_YE : 0u;
  b |= g->dTdx_0_xb ? BC_DTDX_XB : 0u;
  b |= g->dTdx_0_xe ? BC_DTDX_XE : 0u;
  b |= g->dTdy_0_yb ? BC_DTDY_YB : 0u;
  b |= g->dTdy_0_ye ? BC_DTDY_YE : 0u;
  return b;
}

// Solve boxes must leave the first and last column/row out, as every
// non-periodic box of define_area_for_solving (S:6009-6267) does: the
// stencils then stay inside the grid with the reference's clamped indices.
// Periodic about OX, the x boxes span every column (S:6055-6102).
static int check_box(const char *name, const int *b, int N, bool wrap = false) {
  if (wrap) {
    if (b[0] != 0 || b[1] != N) return fail(STS_ERR_ARG, std::string("periodic solve box ") + name + " must be 0..N");
    return STS_OK;
  }
  if (b[0] < 1 || b[1] > N - 1 || b[0] >= b[1])
    return fail(STS_ERR_ARG, std::string("solve box ") + name + " must satisfy 1 <= lo < hi <= N-1");
  return STS_OK;
}

extern "C" int sts_destroy(sts_ctx *c) {
  if (!c) return STS_OK;
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (int q = 0; q < STS_NFIELDS; q++) (void)hipFree(c->f[q]);
  (void)hipFree(c->p_alt);
  (void)hipFree(c->T_alt);
  for (int q = 0; q < STS_NFIELDS; q++) (void)hipFree(c->f32[q]);
  (void)hipFree(c->p_ws);
  for (double *q : c->xs) (void)hipFree(q);
  for (auto &lv : c->hist.f)
    for (double *q : lv) (void)hipFree(q);
  for (int j = 0; j < MAX_AA; j++) {
    (void)hipFree(c->aa_dG[j]);
    (void)hipFree(c->aa_dF[j]);
  }
  (void)hipFree(c->aa.gp);
  (void)hipFree(c->aa.D);
  (void)hipFree(c->aa.fp);
  (void)hipFree(c->aa.part);
  (void)hipFree(c->aa.gram);
  (void)hipFree(c->aa.gam);
  (void)hipFree(c->wall_ops);
  for (unsigned *q : c->drag_cells) (void)hipFree(q);
  (void)hipFree(c->drag_terms);
  (void)hipFree(c->drag_out);
  (void)hipFree(c->T_ws);
  (void)hipFree(c->p_alt32);
  (void)hipFree(c->T_alt32);
  (void)hipFree(c->hx32);
  (void)hipFree(c->hy32);
  (void)hipFree(c->hx);
  (void)hipFree(c->hy);
  (void)hipFree(c->m);
  (void)hipFree(c->nb);
  (void)hipFree(c->vol);
  (void)hipFree(c->bodies);
  (void)hipFree(c->red);
  (void)hipFree(c->red_out);
  (void)hipFree(c->lred);
  (void)hipFree(c->probe);
  (void)hipFree(c->partial);
  (void)hipFree(c->cont_stat);
  (void)hipFree(c->cont);
  if (c->h_red) (void)hipHostFree(c->h_red);
  if (c->h_ctrl) (void)hipHostFree(c->h_ctrl);
  if (c->h_log) (void)hipHostFree(c->h_log);
  if (c->ev_ctrl) (void)hipEventDestroy(c->ev_ctrl);
  (void)hipFree(c->ctrl);
  (void)hipFree(c->log);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return STS_OK;
}

// The neighbour fluid bits of every cell (KGridT::nb): the mask at the
// neighbour indices of Nbr -- clamped at the grid edge, x wrapping when
// periodic -- and at the v equation's second south row.
static std::vector<uint16_t> neighbour_bits(const std::vector<uint8_t> &mask, int Nx, int Ny, bool px) {
  std::vector<uint16_t> out((size_t)Nx * (size_t)Ny);
  for (int j = 0; j < Ny; j++) {
    const int js = j - (0 < j), jn = j + (j < Ny - 1), jss = j >= 2 ? j - 2 : js;
    for (int i = 0; i < Nx; i++) {
      int iw, iww, ie;
      if (px) {
        iw = i == 0 ? Nx - 1 : i - 1;
        iww = i == 1 ? Nx - 1 : (i == 0 ? Nx - 2 : i - 2);
        ie = i == Nx - 1 ? 0 : i + 1;
      } else {
        iw = i - (0 < i);
        iww = i - (0 < i) - (1 < i);
        ie = i + (i < Nx - 1);
      }
      auto f = [&](int x, int y) { return mask[(size_t)x + (size_t)y * Nx] != 0; };
      unsigned b = 0;
      b |= f(i, j) ? NB_C : 0u;
      b |= f(iw, j) ? NB_W : 0u;
      b |= f(ie, j) ? NB_E : 0u;
      b |= f(iww, j) ? NB_WW : 0u;
      b |= f(i, js) ? NB_S : 0u;
      b |= f(i, jn) ? NB_N : 0u;
      b |= f(iw, js) ? NB_SW : 0u;
      b |= f(ie, js) ? NB_SE : 0u;
      b |= f(iw, jn) ? NB_NW : 0u;
      b |= f(ie, jn) ? NB_NE : 0u;
      b |= f(i, jss) ? NB_SS : 0u;
      out[(size_t)i + (size_t)j * Nx] = (uint16_t)b;
    }
  }
  return out;
}

static int alloc_state(sts_ctx *c, const sts_grid *gi, const std::vector<uint8_t> &mask) {
  const size_t Na = (size_t)c->Na;
  const size_t bytes = Na * sizeof(double);
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int q = 0; q < STS_NFIELDS; q++) {
    HIP_OK(hipMalloc(&c->f[q], bytes));
    HIP_OK(hipMemsetAsync(c->f[q], 0, bytes, c->stream));
  }
  HIP_OK(hipMalloc(&c->p_alt, bytes));
  HIP_OK(hipMalloc(&c->T_alt, bytes));
  HIP_OK(hipMalloc(&c->hx, sizeof(double) * gi->Nx));
  HIP_OK(hipMalloc(&c->hy, sizeof(double) * gi->Ny));
  HIP_OK(hipMalloc(&c->m, Na));
  HIP_OK(hipMalloc(&c->vol, sizeof(uint32_t) * Na));
  const size_t nb = (size_t)(gi->nbodies > 0 ? gi->nbodies : 1);
  HIP_OK(hipMalloc(&c->bodies, sizeof(DevBody) * nb));
  HIP_OK(hipMalloc(&c->red, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE));
  HIP_OK(hipMalloc(&c->red_out, sizeof(unsigned long long) * NRED));
  HIP_OK(hipMalloc(&c->lred, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE));
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  HIP_OK(hipMemsetAsync(c->lred, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  HIP_OK(hipMalloc(&c->probe, sizeof(double)));
  HIP_OK(hipHostMalloc(&c->h_red, sizeof(unsigned long long) * (NRED + 1), hipHostMallocDefault));
  HIP_OK(hipHostMalloc(&c->h_ctrl, sizeof(LoopCtrl), hipHostMallocDefault));
  HIP_OK(hipEventCreateWithFlags(&c->ev_ctrl, hipEventDisableTiming));
  HIP_OK(hipMemcpyAsync(c->hx, gi->hx, sizeof(double) * gi->Nx, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->hy, gi->hy, sizeof(double) * gi->Ny, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->m, mask.data(), Na, hipMemcpyHostToDevice, c->stream));
  const std::vector<uint16_t> nbits = neighbour_bits(mask, gi->Nx, gi->Ny, gi->periodic_x != 0);
  HIP_OK(hipMalloc(&c->nb, sizeof(uint16_t) * Na));
  HIP_OK(hipMemcpyAsync(c->nb, nbits.data(), sizeof(uint16_t) * Na, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->vol, gi->vol_inf_fb, sizeof(uint32_t) * Na, hipMemcpyHostToDevice, c->stream));
  std::vector<DevBody> db(nb);
  for (int k = 0; k < gi->nbodies; k++) {
    const sts_body &b = gi->bodies[k];
    db[k] = DevBody{b.u_body, b.v_body, b.p_body, b.T_body, b.F_slip, b.F_jump, b.dTdn_on_wall_0, b.is_slip, b.is_jump};
  }
  HIP_OK(hipMemcpyAsync(c->bodies, db.data(), sizeof(DevBody) * nb, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return STS_OK;
}

// FP32 grid: the same case with FP32 mesh steps and scalars.
static void narrow_grid(sts_ctx *c) {
  const KGrid &g = c->g;
  KGridT<float> &h = c->g32;
  h.Nx = g.Nx; h.Ny = g.Ny;
  h.hx = c->hx32; h.hy = c->hy32;
  h.m = g.m; h.nb = g.nb; h.vol = g.vol; h.bodies = g.bodies; h.nbodies = g.nbodies;
  h.Nx_u0 = g.Nx_u0; h.Nx_u1 = g.Nx_u1; h.Ny_u0 = g.Ny_u0; h.Ny_u1 = g.Ny_u1;
  h.Nx_v0 = g.Nx_v0; h.Nx_v1 = g.Nx_v1; h.Ny_v0 = g.Ny_v0; h.Ny_v1 = g.Ny_v1;
  h.Nx_o0 = g.Nx_o0; h.Nx_o1 = g.Nx_o1; h.Ny_o0 = g.Ny_o0; h.Ny_o1 = g.Ny_o1;
  h.ht = (float)g.ht; h.gamma1 = (float)g.gamma1; h.Kn = (float)g.Kn; h.cT = (float)g.cT;
  h.cT1 = (float)g.cT1; h.cT2 = (float)g.cT2; h.cPch = (float)g.cPch; h.c_Fr_v = (float)g.c_Fr_v;
  h.rht = 1.0f / h.ht;  // IEEE single division on the host: RN(1/ht) in FP32
  h.c1 = 1.0f - h.gamma1;  // single subtraction in FP32, as the kernels formed it
  h.rht_one = std::fabs(std::fma(-h.rht, h.ht, 1.0f)) <= 0x1p-25f;
  h.given_xb = g.given_xb; h.given_xe = g.given_xe; h.kn_local = g.kn_local;
  h.bc = g.bc;
  h.px = g.px;
}

static int alloc_fp32(sts_ctx *c, const sts_grid *gi) {
  const size_t bytes = sizeof(float) * (size_t)c->Na;
  for (int q = 0; q < STS_NFIELDS; q++) {
    HIP_OK(hipMalloc(&c->f32[q], bytes));
    HIP_OK(hipMemsetAsync(c->f32[q], 0, bytes, c->stream));
  }
  HIP_OK(hipMalloc(&c->p_alt32, bytes));
  HIP_OK(hipMalloc(&c->T_alt32, bytes));
  HIP_OK(hipMalloc(&c->hx32, sizeof(float) * gi->Nx));
  HIP_OK(hipMalloc(&c->hy32, sizeof(float) * gi->Ny));
  std::vector<float> hx(gi->hx, gi->hx + gi->Nx), hy(gi->hy, gi->hy + gi->Ny);
  HIP_OK(hipMemcpyAsync(c->hx32, hx.data(), sizeof(float) * gi->Nx, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->hy32, hy.data(), sizeof(float) * gi->Ny, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  narrow_grid(c);
  return STS_OK;
}

// The wall pass's writes of the listed cells, planned on the device from the
// mask and body table and kept in a compact list (at least one slot, so that
// an empty list still selects the list kernel): first the writes whose value
// depends on the state (slip velocities, wall temperatures of jump or
// zero-gradient bodies), then those that write a body's constant
// temperature.  The writes have distinct destinations, so their order is
// free; a loop applies the constant ones in its first outer iteration only.
static int plan_wall(sts_ctx *c, const std::vector<unsigned> &cells, const sts_body *bodies) {
  const int n = (int)cells.size();
  std::vector<WallOp> all((size_t)4 * n), ops;
  if (n > 0) {
    unsigned *d_cells = nullptr;
    WallOp *d_all = nullptr;
    hipError_t e = hipMalloc(&d_cells, sizeof(unsigned) * n);
    if (e == hipSuccess) e = hipMalloc(&d_all, sizeof(WallOp) * 4 * (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_cells, cells.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_wall_plan(c->g, d_cells, n, d_all, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(all.data(), d_all, sizeof(WallOp) * all.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_cells);
    (void)hipFree(d_all);
    HIP_OK(e);
    std::vector<WallOp> fixed;
    for (const WallOp &o : all) {
      const unsigned kind = o.kb & 7u;
      if (kind == WOP_NONE) continue;
      const sts_body &b = bodies[o.kb >> 3];
      const bool constant = (kind == WOP_TWU || kind == WOP_TWV) && !b.is_jump && b.dTdn_on_wall_0 == 0.0;
      (constant ? fixed : ops).push_back(o);
    }
    c->n_wall_var = (int)ops.size();
    ops.insert(ops.end(), fixed.begin(), fixed.end());
  }
  c->n_wall_ops = (int)ops.size();
  HIP_OK(hipMalloc(&c->wall_ops, sizeof(WallOp) * (ops.empty() ? 1 : ops.size())));
  if (!ops.empty())
    HIP_OK(hipMemcpy(c->wall_ops, ops.data(), sizeof(WallOp) * ops.size(), hipMemcpyHostToDevice));
  return STS_OK;
}

extern "C" int sts_create(const sts_grid *gi, int precision, sts_ctx **out) {
  if (!gi || !out) return fail(STS_ERR_ARG, "null argument");
  *out = nullptr;
  if (precision != 64 && precision != 32) return fail(STS_ERR_UNSUPPORTED, "precision must be 64 or 32");
  if (gi->Nx < 3 || gi->Ny < 3 || !gi->hx || !gi->hy || !gi->vol_inf_fb)
    return fail(STS_ERR_ARG, "grid needs Nx, Ny >= 3, hx, hy and vol_inf_fb");
  if (gi->nbodies < 0 || (gi->nbodies > 0 && !gi->bodies)) return fail(STS_ERR_ARG, "bad body table");
  int rc;
  const bool px = gi->periodic_x != 0;
  if ((rc = check_box("Nx_u", gi->Nx_u, gi->Nx, px)) || (rc = check_box("Ny_u", gi->Ny_u, gi->Ny)) ||
      (rc = check_box("Nx_v", gi->Nx_v, gi->Nx, px)) || (rc = check_box("Ny_v", gi->Ny_v, gi->Ny)) ||
      (rc = check_box("Nx_o", gi->Nx_o, gi->Nx, px)) || (rc = check_box("Ny_o", gi->Ny_o, gi->Ny)))
    return rc;
  if (px && gi->Nx < 4) return fail(STS_ERR_ARG, "periodic grid needs Nx >= 4");
  const long long Na = (long long)gi->Nx * gi->Ny;
  if (Na > (1ll << 31) - 1) return fail(STS_ERR_ARG, "grid too large for 32-bit node indices");
  std::vector<uint8_t> mask((size_t)Na);
  for (long long c = 0; c < Na; c++) {
    uint32_t k = gi->vol_inf_fb[c];
    if ((long long)k > gi->nbodies) return fail(STS_ERR_ARG, "vol_inf_fb names a body beyond nbodies");
    mask[(size_t)c] = (k == 0u);
  }
  // S:6906-6912, S:6977-6983 read rho_v/v one row up; on a fluid cell of the
  // top row that is outside the arrays in the reference, so such a case is
  // refused rather than reproduced.
  const size_t top = (size_t)(gi->Ny - 1) * (size_t)gi->Nx;
  if ((gi->drhodt_xb && mask[top + (size_t)(gi->Nx_u[0] - 1)]) ||
      (gi->drhodt_xe && mask[top + (size_t)(gi->Nx_u[1] - 1)]))
    return fail(STS_ERR_UNSUPPORTED, "drhodt_durhodx_dvrhody edge condition on a fluid cell of the top row");
  sts_ctx *c = new sts_ctx();
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(STS_ERR_HIP, "hipGetDevice failed");
  }
  for (uint8_t b : mask) c->fluid_cells += b;
  c->h_vol.assign(gi->vol_inf_fb, gi->vol_inf_fb + Na);
  c->Nx = gi->Nx;
  c->Ny = gi->Ny;
  c->Na = (int)Na;
  if (const char *e = std::getenv("STS_FUSED_SWEEPS")) c->fuse_sweeps = std::atoi(e) != 0;
  int arc = alloc_state(c, gi, mask);
  if (arc) {
    sts_destroy(c);
    return arc;
  }
  std::vector<unsigned> wl;
  {
    // the wall conditions write only at cells k with k, its west neighbour
    // or k-Nx on a body/fluid interface (a cell with a 4-neighbour of the
    // other kind); periodic about OX, west and east wrap
    const int Nx = gi->Nx, Ny = gi->Ny;
    auto west = [&](int i) { return i > 0 ? i - 1 : (px ? Nx - 1 : -1); };
    auto east = [&](int i) { return i + 1 < Nx ? i + 1 : (px ? 0 : -1); };
    auto iface = [&](int i, int j) {
      const uint8_t a = mask[(size_t)i + (size_t)j * Nx];
      const int w = west(i), e = east(i);
      return (w >= 0 && mask[(size_t)w + (size_t)j * Nx] != a) || (e >= 0 && mask[(size_t)e + (size_t)j * Nx] != a) ||
             (j > 0 && mask[(size_t)i + (size_t)(j - 1) * Nx] != a) ||
             (j + 1 < Ny && mask[(size_t)i + (size_t)(j + 1) * Nx] != a);
    };
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++)
        if (iface(i, j) || (west(i) >= 0 && iface(west(i), j)) || (j > 0 && iface(i, j - 1)))
          wl.push_back((unsigned)i + (unsigned)j * (unsigned)Nx);
  }
  KGrid &g = c->g;
  g.Nx = gi->Nx;
  g.Ny = gi->Ny;
  g.hx = c->hx;
  g.hy = c->hy;
  g.m = c->m;
  g.nb = c->nb;
  g.vol = c->vol;
  g.bodies = c->bodies;
  g.nbodies = gi->nbodies;
  g.Nx_u0 = gi->Nx_u[0]; g.Nx_u1 = gi->Nx_u[1]; g.Ny_u0 = gi->Ny_u[0]; g.Ny_u1 = gi->Ny_u[1];
  g.Nx_v0 = gi->Nx_v[0]; g.Nx_v1 = gi->Nx_v[1]; g.Ny_v0 = gi->Ny_v[0]; g.Ny_v1 = gi->Ny_v[1];
  g.Nx_o0 = gi->Nx_o[0]; g.Nx_o1 = gi->Nx_o[1]; g.Ny_o0 = gi->Ny_o[0]; g.Ny_o1 = gi->Ny_o[1];
  g.ht = gi->ht;
  g.rht = 1.0 / g.ht;  // RN(1/ht); 1 - rht ht is exact in an fma
  g.rht_one = std::fabs(std::fma(-g.rht, g.ht, 1.0)) <= 0x1p-54;
  g.gamma1 = gi->gamma1;
  g.c1 = 1.0 - g.gamma1;
  g.Kn = gi->Kn;
  g.cT = gi->cT;
  g.cT1 = gi->cT1;
  g.cT2 = gi->cT2;
  g.cPch = gi->cPch;
  g.c_Fr_v = gi->c_Fr_v;
  g.given_xb = gi->given_xb;
  g.given_xe = gi->given_xe;
  g.kn_local = gi->kn_local;
  g.bc = bc_bits(gi);
  g.px = px ? 1 : 0;
  if ((arc = plan_wall(c, wl, gi->bodies))) {
    sts_destroy(c);
    return arc;
  }
  if (precision == 32) {
    c->mixed = true;
    if ((arc = alloc_fp32(c, gi))) {
      sts_destroy(c);
      return arc;
    }
  }
  *out = c;
  return STS_OK;
}

// Writes Gamma_* and D_* where the loops left them behind (from the inputs
// as the loops left them); after this the arrays hold the reference's values.
static int refresh_diff(sts_ctx *c, hipStream_t s = nullptr) {
  if (!c->diff_stale) return STS_OK;
  HIP_OK(launch_flux_diff(c->g, c->fields(), s ? s : c->stream, Gate{nullptr, 0}, DIFF_PART));
  c->diff_stale = false;
  return STS_OK;
}

// Before a stage that changes sqrt_T or the wall temperatures: the arrays
// are brought up to date first, and no longer follow from the inputs.
static int diff_inputs_change(sts_ctx *c) {
  int rc = refresh_diff(c);
  c->diff_inline_ok = false;
  return rc;
}

static bool diff_field(int f) {
  return f == STS_GAMMA_XF || f == STS_GAMMA_YF || f == STS_D_UX || f == STS_D_UY || f == STS_D_VX || f == STS_D_VY;
}

extern "C" int sts_set_field(sts_ctx *c, int field, const double *host) {
  if (!c || !host || field < 0 || field >= STS_NFIELDS) return fail(STS_ERR_ARG, "bad field or pointer");
  if (diff_field(field) || field == STS_SQRT_T || field == STS_TW_U || field == STS_TW_V) {
    int rc = diff_inputs_change(c);
    if (rc) return rc;
  }
  c->n_hist = 0;  // the state no longer continues the previous steps
  c->last_total = -1;
  c->plain_mean = -1.0;
  c->pred_pause = 0;
  c->pred_backoff = 1;
  c->pred_degree = -1;
  c->pred_wins = 0;
  HIP_OK(hipMemcpyAsync(c->f[field], host, sizeof(double) * (size_t)c->Na, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return STS_OK;
}

extern "C" int sts_get_field(sts_ctx *c, int field, double *host) {
  if (!c || !host || field < 0 || field >= STS_NFIELDS) return fail(STS_ERR_ARG, "bad field or pointer");
  if (diff_field(field)) {
    int rc = refresh_diff(c);
    if (rc) return rc;
  }
  HIP_OK(hipMemcpyAsync(host, c->f[field], sizeof(double) * (size_t)c->Na, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return STS_OK;
}

extern "C" int sts_set_scalars(sts_ctx *c, double cT, double cPch) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  int rc = diff_inputs_change(c);  // cT enters D_*
  if (rc) return rc;
  c->g.cT = cT;
  c->g.cPch = cPch;
  if (c->mixed) narrow_grid(c);
  return STS_OK;
}

extern "C" int sts_stream_info(sts_ctx *c, void **s) {
  if (!c || !s) return fail(STS_ERR_ARG, "null argument");
  *s = (void *)c->stream;
  return STS_OK;
}

extern "C" int sts_save_previous(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  const double *src[5] = {c->f[STS_P], c->f[STS_U], c->f[STS_V], c->f[STS_T], c->f[STS_RHO]};
  double *dst[5] = {c->f[STS_P_PR], c->f[STS_U_PR], c->f[STS_V_PR], c->f[STS_T_PR], c->f[STS_RHO_PR]};
  HIP_OK(launch_copy5(src, dst, c->Na, c->stream));
  return STS_OK;
}

extern "C" int sts_assemble(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  int rc = refresh_diff(c);
  if (rc) return rc;
  const int t = timed_begin(c, STS_K_ASSEMBLE);
  HIP_OK(launch_assemble(c->g, c->fields(), c->stream));
  timed_end(c, t);
  return STS_OK;
}

static double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

// n sweeps, residuals of sweep s in slots 2+2s, 3+2s; pairs of sweeps run
// fused in one pass.  Each pass leaves its result in the partner buffers.
static int run_sweeps(sts_ctx *c, int n) {
  for (int s = 0; s < n;) {
    KFields kf = c->fields();
    const int t = timed_begin(c, STS_K_SWEEP);
    if (c->fuse_sweeps && s + 1 < n) {
      HIP_OK(launch_sweep2(c->g, kf, c->f[STS_P], c->f[STS_T], c->p_alt, c->T_alt, c->red + 2 + 2 * s,
                           c->red + 3 + 2 * s, c->red + 4 + 2 * s, c->red + 5 + 2 * s, c->stream));
      s += 2;
    } else {
      HIP_OK(launch_sweep(c->g, kf, c->f[STS_P], c->f[STS_T], c->p_alt, c->T_alt, c->red + 2 + 2 * s,
                          c->red + 3 + 2 * s, c->stream));
      s += 1;
    }
    timed_end(c, t);
    std::swap(c->f[STS_P], c->p_alt);
    std::swap(c->f[STS_T], c->T_alt);
  }
  return STS_OK;
}

// Folds words 0..nslots-1 of the slotted residual buffer (clearing them) and
// brings them to h_red.
static int read_red(sts_ctx *c, int nslots, bool with_probe, bool check_finite = true) {
  HIP_OK(launch_fold(c->red, nslots, c->red_out, c->stream));
  HIP_OK(hipMemcpyAsync(c->h_red, c->red_out, sizeof(unsigned long long) * nslots, hipMemcpyDeviceToHost,
                        c->stream));
  if (with_probe)
    HIP_OK(hipMemcpyAsync(c->h_red + NRED, c->probe, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < nslots && check_finite; q++)
    if (!std::isfinite(bits_to_double(c->h_red[q])))
      return fail(STS_ERR_NONFINITE, "non-finite residual (slot " + std::to_string(q) + ")");
  return STS_OK;
}

extern "C" int sts_pt_jacobi(sts_ctx *c, int n, double *res) {
  if (!c || n < 1 || n > MAX_SWEEPS) return fail(STS_ERR_ARG, "sweep count must be 1..62");
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  int rc = run_sweeps(c, n);
  if (rc) return rc;
  if ((rc = read_red(c, 2 + 2 * n, false))) return rc;
  if (res)
    for (int s = 0; s < n; s++) {
      res[2 * s] = bits_to_double(c->h_red[2 + 2 * s]);
      res[2 * s + 1] = bits_to_double(c->h_red[3 + 2 * s]);
    }
  return STS_OK;
}

extern "C" int sts_correct(sts_ctx *c, double *res) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  if (int e = diff_inputs_change(c)) return e;
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  const int t = timed_begin(c, STS_K_CORRECT);
  HIP_OK(launch_correct(c->g, c->fields(), c->red, c->red + 1, 1, c->stream));
  timed_end(c, t);
  int rc = read_red(c, 2, false);
  if (rc) return rc;
  if (res) {
    res[0] = bits_to_double(c->h_red[0]);
    res[1] = bits_to_double(c->h_red[1]);
  }
  return STS_OK;
}

extern "C" int sts_properties(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  if (int e = diff_inputs_change(c)) return e;
  HIP_OK(launch_properties(c->g, c->fields(), c->stream));
  return STS_OK;
}

extern "C" int sts_rho_in_middle_points(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  HIP_OK(launch_rho_mid(c->g, c->fields(), c->stream));
  return STS_OK;
}

static int enqueue_bcs(sts_ctx *c, int parts, bool probe) {
  const int wall = parts & (STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP);
  const int edge = parts & (STS_BC_VELOCITIES | STS_BC_TEMPERATURE | STS_BC_PRESSURE);
  // The jump condition reads only T and rho, which the velocity conditions do
  // not touch, so it runs with the slip conditions ahead of the edge copies;
  // the temperature copies still see the wall values' source unchanged.
  if (wall) {
    const int t = timed_begin(c, STS_K_WALL);
    HIP_OK(launch_wall(c->g, c->fields(), wall, c->stream));
    timed_end(c, t);
  }
  if (edge || probe) {
    const int t = timed_begin(c, STS_K_EDGES);
    HIP_OK(launch_edges(c->g, c->fields(), edge, probe ? c->probe : nullptr, c->stream));
    timed_end(c, t);
  }
  return STS_OK;
}

extern "C" int sts_boundary_conditions(sts_ctx *c, int parts, double *p_probe) {
  if (!c || (parts & ~STS_BC_ALL)) return fail(STS_ERR_ARG, "bad boundary-condition parts");
  if (parts & STS_BC_JUMP)
    if (int e = diff_inputs_change(c)) return e;
  int rc = enqueue_bcs(c, parts, p_probe != nullptr);
  if (rc) return rc;
  if (p_probe) {
    HIP_OK(hipMemcpyAsync(p_probe, c->probe, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return STS_OK;
}

extern "C" int sts_fluxes_diffusion(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  const int t = timed_begin(c, STS_K_FLUX_DIFF);
  HIP_OK(launch_flux_diff(c->g, c->fields(), c->stream));
  timed_end(c, t);
  c->diff_inline_ok = true;
  c->diff_stale = false;
  return STS_OK;
}

extern "C" int sts_apply_bodies(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  HIP_OK(launch_apply_bodies(c->g, c->fields(), c->stream));
  return STS_OK;
}

extern "C" int sts_jump_overwrite(sts_ctx *c) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  if (int e = diff_inputs_change(c)) return e;  // the jump pass rewrites the wall temperatures
  HIP_OK(launch_wall(c->g, c->fields(), STS_BC_JUMP, c->stream));
  HIP_OK(launch_jump_overwrite(c->g, c->fields(), c->stream));
  return STS_OK;
}

extern "C" int sts_set_inflow_u(sts_ctx *c, double u_in) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  HIP_OK(launch_set_inflow(c->g, c->fields(), u_in, c->stream));
  return STS_OK;
}

extern "C" int sts_continuity(sts_ctx *c, double *cont_stat, double *cont, double *norms) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  const int nb = continuity_blocks(c->g);
  if (!c->partial) {
    HIP_OK(hipMalloc(&c->partial, sizeof(double) * 3 * (size_t)nb));
    HIP_OK(hipMalloc(&c->cont_stat, sizeof(double) * (size_t)c->Na));
    HIP_OK(hipMalloc(&c->cont, sizeof(double) * (size_t)c->Na));
    HIP_OK(hipMemsetAsync(c->cont_stat, 0, sizeof(double) * (size_t)c->Na, c->stream));
    HIP_OK(hipMemsetAsync(c->cont, 0, sizeof(double) * (size_t)c->Na, c->stream));
  }
  HIP_OK(launch_continuity(c->g, c->fields(), c->cont_stat, c->cont, c->partial, c->stream));
  std::vector<double> part(3 * (size_t)nb);
  HIP_OK(hipMemcpyAsync(part.data(), c->partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
  if (cont_stat)
    HIP_OK(hipMemcpyAsync(cont_stat, c->cont_stat, sizeof(double) * (size_t)c->Na, hipMemcpyDeviceToHost, c->stream));
  if (cont) HIP_OK(hipMemcpyAsync(cont, c->cont, sizeof(double) * (size_t)c->Na, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (norms) {
    double mx = 0, s2 = 0, s1 = 0;
    for (int b = 0; b < nb; b++) {
      mx = std::fmax(mx, part[3 * b]);
      s2 += part[3 * b + 1];
      s1 += part[3 * b + 2];
    }
    norms[0] = mx;
    norms[1] = std::sqrt(s2);
    norms[2] = s1;
  }
  return STS_OK;
}

extern "C" int sts_outer_iteration(sts_ctx *c, int n_pc, double *res, double *res_p0T0, double *p_probe) {
  if (!c || n_pc < 1 || n_pc > MAX_SWEEPS) return fail(STS_ERR_ARG, "N_I_p_c must be 1..62");
  const int nslots = 2 + 2 * n_pc;
  if (int e = refresh_diff(c)) return e;
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  int t = timed_begin(c, STS_K_ASSEMBLE);
  HIP_OK(launch_assemble(c->g, c->fields(), c->stream));
  timed_end(c, t);
  int rc = run_sweeps(c, n_pc);
  if (rc) return rc;
  t = timed_begin(c, STS_K_CORRECT);
  HIP_OK(launch_correct(c->g, c->fields(), c->red, c->red + 1, 1, c->stream));
  timed_end(c, t);
  if ((rc = enqueue_bcs(c, STS_BC_ALL, true))) return rc;
  t = timed_begin(c, STS_K_FLUX_DIFF);
  HIP_OK(launch_flux_diff(c->g, c->fields(), c->stream));
  timed_end(c, t);
  c->diff_inline_ok = true;
  c->diff_stale = false;
  if ((rc = read_red(c, nslots, true))) return rc;
  if (res) {
    res[0] = bits_to_double(c->h_red[0]);
    res[1] = bits_to_double(c->h_red[1]);
    res[2] = bits_to_double(c->h_red[2 + 2 * (n_pc - 1)]);
    res[3] = bits_to_double(c->h_red[3 + 2 * (n_pc - 1)]);
  }
  if (res_p0T0) {
    res_p0T0[0] = bits_to_double(c->h_red[2]);
    res_p0T0[1] = bits_to_double(c->h_red[3]);
  }
  if (p_probe) std::memcpy(p_probe, &c->h_red[NRED], sizeof(double));
  return STS_OK;
}

// Buffer flips (passes) of one loop 3 of n sweeps.
static int sweep_passes(const sts_ctx *c, int n) { return c->fuse_sweeps ? n / 2 + n % 2 : n; }

// One precision's view of the state for loop 2: grid, fields (p and T as
// they stand at the start of the step) and the p/T ping-pong partners.
template <typename R>
struct Loop2State {
  const KGridT<R> *g;
  KFieldsT<R> f;
  R *p_alt, *T_alt;
  int kid;  // timing id offset: 0 for FP64 kernels, STS_K_COUNT for FP32
  int aa_m = 0, aa_every = 1;  // Anderson acceleration: history depth (0: off), period
  bool inline0 = false;        // the first iteration may form the diffusion coefficients itself
};

// How the store after the K-th block of an Anderson loop of history depth m
// finds the block's start: from D (m > 1), as gp (no mix since the last
// store: K < 2), or as the last mix formed again (m = 1), whose D the mix
// then does not write.
static int aa_start(int K, int m) { return m > 1 ? AA_START_D : K < 2 ? AA_START_GP : AA_START_MIX; }

// Anderson acceleration (history depth m, a step every `every` outer
// iterations): the differences in use after iteration k's store.
static AASel anderson_select(const sts_ctx *c, int k, int m, int every) {
  const int K = (k + 1) / every - 1, n = std::min(K, m);
  AASel sel{};
  sel.n = n;
  sel.keep_dF = m > 1;
  sel.start = aa_start(K, m);
  sel.write_D = m > 1;
  for (int j = 0; j < n; j++) {
    const int sl = (K - n + j) % m;
    sel.slot[j] = sl;
    sel.dG[j] = c->aa_dG[sl];
    sel.dF[j] = c->aa_dF[sl];
  }
  return sel;
}

// The Anderson step after iteration k's edge conditions, gated on iteration
// k + 1 running (the iteration that meets the test is the state the step
// keeps): the iteration is stored; with at least one difference in use the
// mixed start replaces u, v, p, T, with sqrt_T, rho, the face densities,
// the fluxes and the thermal diffusion coefficients.
// The wall and edge conditions are not applied again: every stored iterate
// satisfies them, so the affine combination satisfies each condition that
// is affine in the state (given values, zero-gradient copies, constant-Kn
// slip and jump); the wall temperatures stay the newest iterate's, and the
// rest (local-Kn slip and jump, the rho u copies) the next iteration
// restores, whose own change the test judges.  The assembly of the next
// iteration forms Gamma and D from the mixed sqrt_T.
template <typename R>
static int enqueue_anderson(sts_ctx *c, const Loop2State<R> &st, const KFieldsT<R> &kf, int k, const AASel &sel) {
  const Gate gn{c->ctrl, k + 1};
  const AAFieldsT<R> af{{kf.f[STS_U], kf.f[STS_V], kf.f[STS_P], kf.f[STS_T]}};
  int t = timed_begin(c, st.kid + STS_K_AA_STORE, k + 1);
  HIP_OK(launch_aa_store(af, c->aa, sel, c->Na, c->stream, gn));
  timed_end(c, t);
  if (sel.n == 0) return STS_OK;
  t = timed_begin(c, st.kid + STS_K_AA_MIX, k + 1);
  HIP_OK(launch_aa_mix(*st.g, kf, af, c->aa, sel, c->stream, gn));
  timed_end(c, t);
  return STS_OK;
}

// Buffers of the Anderson step for history depth m.
static int ensure_anderson(sts_ctx *c, int m) {
  if (m < 1 || m > MAX_AA) return fail(STS_ERR_ARG, "Anderson history depth must be 0..8");
  const size_t n4 = 4 * (size_t)c->Na;
  if (!c->aa.gp) {
    c->aa.nb = aa_store_blocks(c->Na);
    int slots = c->aa.nb;  // partial sums: enough for the batched steps' tiles too
    for (int v = 0; v < NVARIANT; v++) slots = std::max(slots, aa_case_tiles(c->Nx, c->Ny, v));
    HIP_OK(hipMalloc(&c->aa.gp, sizeof(double) * n4));
    HIP_OK(hipMalloc(&c->aa.D, sizeof(float) * n4));
    HIP_OK(hipMalloc(&c->aa.fp, sizeof(float) * n4));
    HIP_OK(hipMalloc(&c->aa.part, sizeof(double) * 2 * MAX_AA * (size_t)slots));
    HIP_OK(hipMalloc(&c->aa.gram, sizeof(double) * MAX_AA * MAX_AA));
    HIP_OK(hipMalloc(&c->aa.gam, sizeof(double) * MAX_AA));
  }
  for (; c->aa_cap < m; c->aa_cap++) {
    HIP_OK(hipMalloc(&c->aa_dG[c->aa_cap], sizeof(float) * n4));
    HIP_OK(hipMalloc(&c->aa_dF[c->aa_cap], sizeof(float) * n4));
  }
  return STS_OK;
}

// Enqueues outer iteration k of the current loop 2; the p/T buffer parity at
// its start follows the passes of the iterations before it.  The kernels
// leave their residual maxima in the slotted words lred; the edge kernel folds
// them into log[6k..6k+5] and runs the loop test.
template <typename R>
static int enqueue_outer(sts_ctx *c, const Loop2State<R> &st, int k, int n, LoopTest lt) {
  const KGridT<R> &g = *st.g;
  const Gate gate{c->ctrl, k};
  R *const p0 = st.f.f[STS_P], *const T0 = st.f.f[STS_T], *const p1 = st.p_alt, *const T1 = st.T_alt;
  unsigned long long *lr = c->lred;
  auto fields_at = [&](int parity) {
    KFieldsT<R> kf = st.f;
    kf.f[STS_P] = parity ? p1 : p0;
    kf.f[STS_T] = parity ? T1 : T0;
    return kf;
  };
  int par = (int)(((long long)k * sweep_passes(c, n)) & 1);
  // after the loop's first flux pass the assembly forms the diffusion
  // coefficients from the inputs that pass saw; the pass itself writes only
  // the fluxes and thermal coefficients
  const bool inl = k > 0 || st.inline0;
  int t = timed_begin(c, st.kid + (inl ? STS_K_ASSEMBLE_INL : STS_K_ASSEMBLE), k);
  HIP_OK(launch_assemble(g, fields_at(par), c->stream, gate, inl));
  timed_end(c, t);
  for (int s = 0; s < n;) {
    KFieldsT<R> kf = fields_at(par);
    R *Pi = par ? p1 : p0, *Ti = par ? T1 : T0, *Po = par ? p0 : p1, *To = par ? T0 : T1;
    t = timed_begin(c, st.kid + STS_K_SWEEP, k);
    if (c->fuse_sweeps && s + 1 < n) {
      // sweep s (first?) and s+1 (last?)
      HIP_OK(launch_sweep2(g, kf, Pi, Ti, Po, To, s == 0 ? lr + 4 : nullptr, s == 0 ? lr + 5 : nullptr,
                           s + 1 == n - 1 ? lr + 2 : nullptr, s + 1 == n - 1 ? lr + 3 : nullptr, c->stream, gate));
      s += 2;
    } else {
      HIP_OK(launch_sweep(g, kf, (const R *)Pi, (const R *)Ti, Po, To, s == n - 1 ? lr + 2 : nullptr,
                          s == n - 1 ? lr + 3 : nullptr, c->stream, s == 0 ? lr + 4 : nullptr,
                          s == 0 ? lr + 5 : nullptr, gate));
      s += 1;
    }
    timed_end(c, t);
    par ^= 1;
  }
  const KFieldsT<R> kf = fields_at(par);
  // Anderson acceleration: a step after the edge conditions
  const bool aa = st.aa_m > 0 && (k + 1) % st.aa_every == 0;
  AASel sel{};
  if (aa) sel = anderson_select(c, k, st.aa_m, st.aa_every);
  t = timed_begin(c, st.kid + STS_K_CORRECT, k);
  HIP_OK(launch_correct(g, kf, lr + 0, lr + 1, 1, c->stream, gate));
  timed_end(c, t);
  t = timed_begin(c, st.kid + STS_K_WALL, k);
  HIP_OK(launch_wall(g, kf, STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP, c->stream, gate, c->wall_ops,
                     k == 0 ? c->n_wall_ops : c->n_wall_var));
  timed_end(c, t);
  lt.lred = lr;
  lt.log_k += 6 * (size_t)k;
  t = timed_begin(c, st.kid + STS_K_EDGES, k);
  HIP_OK(launch_edges(g, kf, STS_BC_VELOCITIES | STS_BC_TEMPERATURE | STS_BC_PRESSURE, c->probe, c->stream, gate,
                      lt));
  timed_end(c, t);
  if (aa) {
    int r;
    if ((r = enqueue_anderson(c, st, kf, k, sel))) return r;
  }
  // an Anderson mix forms the fluxes itself, and for the loop's last
  // iteration (where the mix, gated on k + 1, does not run) run_loop2 runs
  // the pass once the loop has stopped
  if (aa && sel.n > 0) return STS_OK;
  t = timed_begin(c, st.kid + STS_K_FLUX_DIFF, k);
  HIP_OK(launch_flux_diff(g, kf, c->stream, gate, FLUX_PART));
  timed_end(c, t);
  return STS_OK;
}

// Runs one loop 2 on the device: outer iterations are enqueued in batches
// (the first of `batch`), each gated on the loop control, and the host checks
// the control once per batch.  lt carries the test (tolerances, cap N_I,
// stagnation) and the log base.  On return *iters holds the iterations run
// and (p, T) of the state are where the last pass left them.
template <typename R>
static int run_loop2(sts_ctx *c, Loop2State<R> &st, R *&p, R *&T, R *&p_alt, R *&T_alt, int n_pc, LoopTest lt,
                     int batch, int *iters) {
  HIP_OK(launch_loop_open(c->ctrl, lt.log_k, 6 * lt.N_I, c->lred, c->stream));
  timed_loop_open(c);
  if (st.aa_m > 0) {
    int r;
    if ((r = ensure_anderson(c, st.aa_m))) return r;
    HIP_OK(launch_aa_begin(AAFieldsT<R>{{st.f.f[STS_U], st.f.f[STS_V], p, T}}, c->aa, c->Na, c->stream));
  }
  // The loop control is read after iteration end - 1 - LOOKAHEAD of each
  // batch (after the last one of a batch that ends at N_I), so that the host
  // decides while the GPU still runs the batch's tail; iterations enqueued
  // past the stop are gated out on the device.
  constexpr int LOOKAHEAD = 1;
  int k = 0, stop = -1, rc;
  while (stop < 0) {
    const int end = std::min(lt.N_I, k + batch);
    const int check = end == lt.N_I ? end - 1 : std::max(k, end - 1 - LOOKAHEAD);
    for (; k < end; k++) {
      if ((rc = enqueue_outer(c, st, k, n_pc, lt))) return rc;
      if (k == check) {
        HIP_OK(hipMemcpyAsync(c->h_ctrl, c->ctrl, sizeof(LoopCtrl), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipEventRecord(c->ev_ctrl, c->stream));
      }
    }
    HIP_OK(hipEventSynchronize(c->ev_ctrl));
    if (c->h_ctrl->stop_at != 0x7fffffff) stop = c->h_ctrl->stop_at;
    else if (check == lt.N_I - 1) return fail(STS_ERR_ARG, "loop 2 did not stop at N_I");  // the test always stops there
    batch = 2;
  }
  timed_loop_close(c, stop);
  *iters = stop + 1;
  if (((long long)*iters * sweep_passes(c, n_pc)) & 1) {  // odd number of passes: result in the partners
    std::swap(p, p_alt);
    std::swap(T, T_alt);
  }
  if (st.aa_m > 0 && *iters % st.aa_every == 0 && (*iters / st.aa_every - 1) > 0) {
    // the last iteration's flux pass (its Anderson mix did not run)
    KFieldsT<R> kf = st.f;
    kf.f[STS_P] = p;
    kf.f[STS_T] = T;
    HIP_OK(launch_flux_diff(*st.g, kf, c->stream, Gate{nullptr, 0}, FLUX_PART));
  }
  if (c->h_ctrl->nonfinite) return fail(STS_ERR_NONFINITE, "non-finite residual in outer iteration " + std::to_string(*iters));
  return STS_OK;
}

static int ensure_loop_buffers(sts_ctx *c, int n_iter) {
  if (n_iter > c->log_cap) {
    (void)hipFree(c->log);
    c->log = nullptr;
    HIP_OK(hipMalloc(&c->log, sizeof(unsigned long long) * 6 * (size_t)n_iter));
    if (c->h_log) (void)hipHostFree(c->h_log);
    c->h_log = nullptr;
    HIP_OK(hipHostMalloc(&c->h_log, sizeof(unsigned long long) * 6 * (size_t)n_iter, hipHostMallocDefault));
    c->log_cap = n_iter;
  }
  if (!c->ctrl) HIP_OK(hipMalloc(&c->ctrl, sizeof(LoopCtrl)));
  return STS_OK;
}

// Copies n_iter iterations of the residual log and the probe to the host.
// With neither asked for the step returns without waiting for its last
// launches (the loop's flux pass, iterations gated out past the stop): the
// next step's launches queue behind them on the context's stream, and every
// call that hands device data back waits for the stream.
static int finish_step(sts_ctx *c, int n_iter, double *res_log, double *p_probe) {
  if (!res_log && !p_probe) return STS_OK;
  const size_t nw = 6 * (size_t)n_iter;  // n_iter <= log_cap
  if (n_iter && res_log)
    HIP_OK(hipMemcpyAsync(c->h_log, c->log, sizeof(unsigned long long) * nw, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&c->h_red[NRED], c->probe, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (res_log)
    for (size_t q = 0; q < nw; q++) res_log[q] = bits_to_double(c->h_log[q]);
  if (p_probe) std::memcpy(p_probe, &c->h_red[NRED], sizeof(double));
  return STS_OK;
}

static LoopTest loop_test(sts_ctx *c, unsigned long long *log, int N_I, double tol_uv, double tol_p, double tol_T,
                          int stall_w, double stall_ratio) {
  return LoopTest{c->ctrl, c->lred, log, N_I, tol_uv, tol_p, tol_T, stall_w, stall_ratio};
}

// Loop 1 body of the plain path: the prologue, then loop 2 on the device
// until the reference's test stops it.
static int plain_step(sts_ctx *c, const sts_loop_control *lc, int *iters) {
  int rc;
  *iters = 0;
  if ((rc = sts_save_previous(c))) return rc;
  Loop2State<double> st{&c->g, c->fields(), c->p_alt, c->T_alt, 0};
  st.inline0 = c->diff_inline_ok;
  const int batch = lc->batch > 0 ? lc->batch : (c->last_iters > 2 ? c->last_iters - 1 : 1);
  rc = run_loop2(c, st, c->f[STS_P], c->f[STS_T], c->p_alt, c->T_alt, lc->N_I_p_c,
                 loop_test(c, c->log, lc->N_I, lc->tol_uv, lc->tol_p, lc->tol_T, 0, 0.0), batch, iters);
  c->diff_inline_ok = c->diff_stale = true;
  c->last_iters = *iters;
  // the step is not stored in the history of predicted starts, so the stored
  // steps no longer precede the state evenly spaced in time
  c->n_hist = 0;
  return rc;
}

static int check_loop_control(const sts_loop_control *lc) {
  if (!lc || lc->N_I < 1 || lc->N_I_p_c < 1 || lc->N_I_p_c > MAX_SWEEPS) return fail(STS_ERR_ARG, "bad loop control");
  return STS_OK;
}

extern "C" int sts_time_step(sts_ctx *c, const sts_loop_control *lc, int *iterations, double *res_log,
                             double *p_probe) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  int rc;
  if ((rc = check_loop_control(lc))) return rc;
  if ((rc = ensure_loop_buffers(c, lc->N_I))) return rc;
  int iters = 0;
  rc = plain_step(c, lc, &iters);
  if (iterations) *iterations = iters;
  int rc2 = finish_step(c, iters, res_log, p_probe);
  return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// The plain time step without allocation (sts_time_step_noalloc): its loop
// control, residual words, probe and residual log live in `scratch`, one
// fixed array per GPU that every context on that GPU shares.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t SCR_CTRL = 0, SCR_PROBE = 256, SCR_LRED = 512;
constexpr size_t SCR_LOG = SCR_LRED + sizeof(unsigned long long) * NSLOT * SLOT_STRIDE;
constexpr size_t SCRATCH_BYTES = SCR_LOG + sizeof(unsigned long long) * 6 * (size_t)STS_NOALLOC_MAX_N_I;
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};  // device address of each GPU's copy
}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// The address of device dev's copy of scratch, looked up once with dev
// current (the array's name is not a device address in host code).
static int scratch_of(int dev, unsigned char **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(STS_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(STS_ERR_HIP, "hipGetSymbolAddress gave a null address for scratch");
    g_scratch[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch[dev];
  return STS_OK;
}

extern "C" int sts_time_step_noalloc(sts_ctx *c, const sts_loop_control *lc, int *iterations, double *res_log,
                                     double *p_probe) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  int rc;
  if ((rc = check_loop_control(lc))) return rc;
  if (lc->N_I > STS_NOALLOC_MAX_N_I)
    return fail(STS_ERR_ARG, "N_I = " + std::to_string(lc->N_I) + " exceeds STS_NOALLOC_MAX_N_I = " +
                                 std::to_string(STS_NOALLOC_MAX_N_I) + " of the allocation-free time step");
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  if (dev != c->device) return fail(STS_ERR_ARG, "the context's device must be the current device");
  unsigned char *scr = nullptr;
  if ((rc = scratch_of(dev, &scr))) return rc;
  // the context's own loop buffers are set aside for the call; timing is off
  // (its event pool may grow)
  LoopCtrl *ctrl0 = c->ctrl;
  double *probe0 = c->probe;
  unsigned long long *lred0 = c->lred, *log0 = c->log;
  const unsigned timed0 = c->timed;
  c->ctrl = reinterpret_cast<LoopCtrl *>(scr + SCR_CTRL);
  c->probe = reinterpret_cast<double *>(scr + SCR_PROBE);
  c->lred = reinterpret_cast<unsigned long long *>(scr + SCR_LRED);
  c->log = reinterpret_cast<unsigned long long *>(scr + SCR_LOG);
  c->timed = 0;
  int iters = 0;
  rc = plain_step(c, lc, &iters);
  // the log words are the residuals' IEEE bit patterns, copied as doubles
  hipError_t e = hipSuccess;
  if (iters > 0 && res_log)
    e = hipMemcpyAsync(res_log, c->log, sizeof(double) * 6 * (size_t)iters, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && p_probe) e = hipMemcpyAsync(p_probe, c->probe, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  c->ctrl = ctrl0;
  c->probe = probe0;
  c->lred = lred0;
  c->log = log0;
  c->timed = timed0;
  if (iterations) *iterations = iters;
  if (rc) return rc;
  if (e != hipSuccess) return fail(STS_ERR_HIP, std::string("sts_time_step_noalloc copies: ") + hipGetErrorString(e));
  return STS_OK;
}

// Fields an outer iteration reads before it writes them: the FP32 phase hands
// these to the FP64 phase (whose assembly then forms Gamma and D itself).
static constexpr unsigned long long bit(int f) { return 1ull << f; }
static constexpr unsigned long long WIDEN_MASK =
    bit(STS_U) | bit(STS_V) | bit(STS_P) | bit(STS_T) | bit(STS_RHO) | bit(STS_SQRT_T) | bit(STS_RHO_U) |
    bit(STS_RHO_V) | bit(STS_FX) | bit(STS_FY) | bit(STS_DTX) | bit(STS_DTY) | bit(STS_TW_U) | bit(STS_TW_V);

// Every field an outer iteration derives from u, v, p, T, rebuilt as the end
// of an outer iteration builds it (after a new start of loop 2).
static int rebuild_derived(sts_ctx *c, hipStream_t s) {
  const KFields kf = c->fields();
  HIP_OK(launch_correct(c->g, kf, nullptr, nullptr, 0, s));
  HIP_OK(launch_wall(c->g, kf, STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP, s, Gate{nullptr, 0}, c->wall_ops,
                     c->n_wall_ops));
  HIP_OK(launch_edges(c->g, kf, STS_BC_VELOCITIES | STS_BC_TEMPERATURE | STS_BC_PRESSURE, nullptr, s));
  HIP_OK(launch_flux_diff(c->g, kf, s, Gate{nullptr, 0}, FLUX_PART));
  c->diff_inline_ok = c->diff_stale = true;
  return STS_OK;
}

// Copies u, v, p, T into the kept start (when the context keeps one), or
// back from it (back = true).
static int copy_start(sts_ctx *c, hipStream_t s, bool back = false) {
  if (!c->xs[0]) return STS_OK;
  const int f[4] = {STS_U, STS_V, STS_P, STS_T};
  const size_t bytes = sizeof(double) * (size_t)c->Na;
  for (int q = 0; q < 4; q++)
    HIP_OK(hipMemcpyAsync(back ? c->f[f[q]] : c->xs[q], back ? c->xs[q] : c->f[f[q]], bytes,
                          hipMemcpyDeviceToDevice, s));
  return STS_OK;
}

// After the prologue: loop 2 starts from the polynomial of degree `order`
// (capped by the stored steps) through x^n and the previous converged steps
// for u, v, p, T, and every field derived from them is rebuilt the way the
// end of an outer iteration builds it.  The history then advances a step.
static int extrapolate_start(sts_ctx *c, int order, hipStream_t s = nullptr, bool store_only = false,
                             int degree_cap = MAX_EXTRAP, int *used = nullptr) {
  if (!s) s = c->stream;
  if (used) *used = 0;
  if (order <= 0) {
    c->n_hist = 0;
    return copy_start(c, s);
  }
  if (order > MAX_EXTRAP) return fail(STS_ERR_ARG, "extrapolation degree must be 0..5");
  const size_t bytes = sizeof(double) * (size_t)c->Na;
  if (!c->hist.f[0][0])
    for (auto &lv : c->hist.f)
      for (double *&q : lv) HIP_OK(hipMalloc(&q, bytes));
  // `order` past steps are kept: x^n goes to a free level, or over the oldest
  // one, and that level becomes level 0 (the rest move down by renaming)
  const int keep = std::min(c->n_hist, order), use = store_only ? 0 : std::min(keep, degree_cap);
  const int slot = keep < order ? keep : order - 1;
  const KFields kf = c->fields();
  HIP_OK(launch_extrapolate(c->g, kf, c->hist, use, slot, s));
  double *top[4];
  for (int q = 0; q < 4; q++) top[q] = c->hist.f[slot][q];
  for (int h = slot; h > 0; h--)
    for (int q = 0; q < 4; q++) c->hist.f[h][q] = c->hist.f[h - 1][q];
  for (int q = 0; q < 4; q++) c->hist.f[0][q] = top[q];
  c->n_hist = std::min(keep + 1, order);
  if (used) *used = use;
  int rc;
  if (use > 0 && (rc = rebuild_derived(c, s))) return rc;
  return copy_start(c, s);
}

// The first sweep's residual max(max|dp|/tol_p, max|dT|/tol_T) from the
// current start (its own assembly; both overwritten by the loop's first
// iteration), through residual slots q, q + 1; *r = +inf if not finite.
static int first_sweep_residual(sts_ctx *c, const sts_loop_control *lc, int q, double *r) {
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  HIP_OK(launch_assemble(c->g, c->fields(), c->stream, Gate{nullptr, 0}, c->diff_inline_ok));
  HIP_OK(launch_sweep(c->g, c->fields(), (const double *)c->f[STS_P], (const double *)c->f[STS_T], c->p_alt,
                      c->T_alt, c->red + q, c->red + q + 1, c->stream));
  HIP_OK(launch_fold(c->red, q + 2, c->red_out, c->stream));
  unsigned long long w[2];
  HIP_OK(hipMemcpyAsync(w, c->red_out + q, sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  double a, b;
  std::memcpy(&a, &w[0], sizeof a);
  std::memcpy(&b, &w[1], sizeof b);
  *r = (std::isfinite(a) && std::isfinite(b)) ? std::max(a / lc->tol_p, b / lc->tol_T) : HUGE_VAL;
  return STS_OK;
}

// The corrector on a predicted start (see sts_set_corrector): the network's
// refinement is kept only if it lowers the first sweep's residual.
static int corrector_refine(sts_ctx *c, const sts_loop_control *lc) {
  int rc;
  double rA, rB;
  c->corr_kept_now = false;
  if (c->corr_pause > 0) {
    c->corr_pause--;
    c->cstat[2]++;
    return STS_OK;
  }
  if ((rc = first_sweep_residual(c, lc, 2, &rA))) return rc;
  StsView v;
  sts_internal_view(c, &v);
  double *const out[4] = {c->f[STS_U], c->f[STS_V], c->f[STS_P], c->f[STS_T]};
  if ((rc = sts_internal_unet_correct(c->corr_net, v, &c->corr_norm, out))) return rc;
  if ((rc = rebuild_derived(c, c->stream)) || (rc = first_sweep_residual(c, lc, 4, &rB))) return rc;
  if (rB < rA) {  // kept; whether it pays is known when the step has run (corrector_outcome)
    c->cstat[0]++;
    c->corr_kept_now = true;
    return STS_OK;
  }
  c->cstat[1]++;
  if (c->corr_max_pause > 0) {
    c->corr_pause = c->corr_backoff;
    c->corr_backoff = std::min(2 * c->corr_backoff, c->corr_max_pause);
  }
  if ((rc = copy_start(c, c->stream, true))) return rc;
  return rebuild_derived(c, c->stream);
}

// After a predicted step of a context with a corrector and a back-off: a
// kept refinement pays when the step needed at least one outer iteration
// (all phases) fewer than the running mean of the predicted steps run
// without one -- a smaller difference is within the step-to-step variation
// of the counts, which fall through a transient anyway; one that does not
// pay backs off like a rejected one.  restarted: the
// predicted attempt was abandoned (the step ran from x^n).
static void corrector_outcome(sts_ctx *c, int total, bool restarted) {
  if (!c->corr_net || c->corr_max_pause <= 0) return;
  if (c->corr_kept_now) {
    c->corr_kept_now = false;
    if (!restarted && (c->corr_ref_mean < 0.0 || total <= c->corr_ref_mean - 1.0)) {
      c->corr_backoff = 1;
    } else {
      c->corr_pause = c->corr_backoff;
      c->corr_backoff = std::min(2 * c->corr_backoff, c->corr_max_pause);
    }
  } else if (!restarted) {
    c->corr_ref_mean = c->corr_ref_mean < 0.0 ? total : 0.5 * (c->corr_ref_mean + total);
  }
}

// Whether this step of context c runs an FP32 phase (mc asks for one and the
// gate, if any, lets it: the previous step needed more than fp32_above).
static bool fp32_phase(const sts_ctx *c, const sts_mixed_control *mc) {
  return mc && mc->N_I32 > 0 && (mc->fp32_above <= 0 || c->last_total < 0 || c->last_total > mc->fp32_above);
}

// Loop 2 after the prologue: the FP32 phase (when mc asks for one), then the
// FP64 phase whose test is the reference's.  Logs go to c->log, FP32 first.
static int run_phases(sts_ctx *c, const sts_loop_control *lc, const sts_mixed_control *mc, int *n32p, int *n64p) {
  int rc, n32 = 0, n64 = 0;
  *n32p = *n64p = 0;
  if (fp32_phase(c, mc)) {
    // the FP32 phase starts from the narrowed FP64 arrays, Gamma and D
    // included (as the batched steps do)
    if ((rc = refresh_diff(c))) return rc;
    const unsigned long long all = (1ull << STS_NFIELDS) - 1ull;
    HIP_OK(launch_convert(c->fields(), c->fields32(), all, c->Na, c->stream));
    Loop2State<float> s32{&c->g32, c->fields32(), c->p_alt32, c->T_alt32, STS_K_COUNT};
    if (mc->anderson32 > 0) {
      s32.aa_m = mc->anderson32;
      s32.aa_every = std::max(1, mc->anderson_every);
    }
    const int b32 = lc->batch > 0 ? lc->batch : (c->last_iters32 > 2 ? c->last_iters32 - 1 : 1);
    rc = run_loop2(c, s32, c->f32[STS_P], c->f32[STS_T], c->p_alt32, c->T_alt32, lc->N_I_p_c,
                   loop_test(c, c->log, mc->N_I32, mc->tol32_uv, mc->tol32_p, mc->tol32_T, mc->stall_w,
                             mc->stall_ratio),
                   b32, &n32);
    c->last_iters32 = n32;
    *n32p = n32;
    if (rc) return rc;
    HIP_OK(launch_convert(c->fields32(), c->fields(), WIDEN_MASK, c->Na, c->stream));
    c->diff_inline_ok = c->diff_stale = true;
  }
  Loop2State<double> s64{&c->g, c->fields(), c->p_alt, c->T_alt, 0};
  s64.inline0 = c->diff_inline_ok;
  if (mc && mc->anderson > 0) {
    s64.aa_m = mc->anderson;
    s64.aa_every = std::max(1, mc->anderson_every);
  }
  int &last = mc ? c->last_iters64m : c->last_iters;
  const int b64 = lc->batch > 0 ? lc->batch : (last > 2 ? last - 1 : 1);
  rc = run_loop2(c, s64, c->f[STS_P], c->f[STS_T], c->p_alt, c->T_alt, lc->N_I_p_c,
                 loop_test(c, c->log + 6 * (size_t)n32, lc->N_I, lc->tol_uv, lc->tol_p, lc->tol_T, 0, 0.0), b64,
                 &n64);
  c->diff_inline_ok = c->diff_stale = true;
  last = n64;
  *n64p = n64;
  c->last_total = n32 + n64;
  return rc;
}

static int check_mixed(const sts_ctx *c, const sts_loop_control *lc, const sts_mixed_control *mc) {
  if (!lc || lc->N_I < 1 || lc->N_I_p_c < 1 || lc->N_I_p_c > MAX_SWEEPS) return fail(STS_ERR_ARG, "bad loop control");
  if (mc && (mc->N_I32 < 0 || mc->stall_w < 0 || mc->anderson < 0 || mc->anderson > MAX_AA ||
             mc->anderson_every < 0 || mc->anderson32 < 0 || mc->anderson32 > MAX_AA))
    return fail(STS_ERR_ARG, "bad mixed control");
  if (mc && mc->N_I32 > 0 && !c->mixed) return fail(STS_ERR_UNSUPPORTED, "the context was not created with precision 32");
  return STS_OK;
}

// Loop 2 from x^n again after an abandoned predicted attempt: u, v, p, T
// back to the step's start values (the *_PR fields) and the derived fields
// rebuilt from them, as the end of the previous step left them.
static int restart_from_previous(sts_ctx *c) {
  const size_t bytes = sizeof(double) * (size_t)c->Na;
  const int src[4] = {STS_U_PR, STS_V_PR, STS_P_PR, STS_T_PR}, dst[4] = {STS_U, STS_V, STS_P, STS_T};
  for (int q = 0; q < 4; q++)
    HIP_OK(hipMemcpyAsync(c->f[dst[q]], c->f[src[q]], bytes, hipMemcpyDeviceToDevice, c->stream));
  return rebuild_derived(c, c->stream);
}

// Whether the last FP64 log entry of a loop 2 of n outer iterations (log
// offset off) meets the reference's test (S:4433-4437).
static int log_converged(sts_ctx *c, size_t off, int n, const sts_loop_control *lc, bool *ok) {
  unsigned long long w[4];
  HIP_OK(hipMemcpyAsync(w, c->log + off + 6 * (size_t)(n - 1), sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  const double r[4] = {bits_to_double(w[0]), bits_to_double(w[1]), bits_to_double(w[2]), bits_to_double(w[3])};
  *ok = !(lc->tol_uv < r[0] || lc->tol_uv < r[1] || lc->tol_p < r[2] || lc->tol_T < r[3]);
  return STS_OK;
}

extern "C" int sts_time_step_mixed(sts_ctx *c, const sts_loop_control *lc, const sts_mixed_control *mc,
                                   int *iters32, int *iters64, double *res_log, double *p_probe) {
  if (!c || !mc) return fail(STS_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_mixed(c, lc, mc))) return rc;
  if ((rc = ensure_loop_buffers(c, lc->N_I + mc->N_I32))) return rc;
  if ((rc = sts_save_previous(c))) return rc;
  int n32 = 0, n64 = 0;
  if (mc->guard <= 0 || mc->extrapolate <= 0) {
    int used = 0;
    if ((rc = extrapolate_start(c, mc->extrapolate, nullptr, false, MAX_EXTRAP, &used))) return rc;
    if (used > 0 && c->corr_net && (rc = corrector_refine(c, lc))) return rc;
    rc = run_phases(c, lc, mc, &n32, &n64);
    if (!rc && used > 0) corrector_outcome(c, n32 + n64, false);
  } else {
    // guarded predicted start: the degree drops by one after an abandoned
    // attempt and climbs back after 4 accepted ones; an abandoned degree-1
    // attempt pauses prediction
    if (c->pred_degree < 0 || c->pred_degree > mc->extrapolate) c->pred_degree = mc->extrapolate;
    const bool predict = c->n_hist > 0 && c->plain_mean > 0.0 && c->pred_pause == 0;
    int used = 0;
    if ((rc = extrapolate_start(c, mc->extrapolate, nullptr, !predict, c->pred_degree, &used))) return rc;
    if (used > 0 && c->corr_net && (rc = corrector_refine(c, lc))) return rc;
    bool plain = !predict;
    if (predict) {
      c->pstat[0]++;
      sts_loop_control lcp = *lc;
      lcp.N_I = std::max(1, std::min(lc->N_I, (int)std::ceil(c->plain_mean)));
      rc = run_phases(c, &lcp, mc, &n32, &n64);
      bool ok = true;
      if (!rc && n64 >= lcp.N_I && lcp.N_I < lc->N_I) rc = log_converged(c, 6 * (size_t)n32, n64, lc, &ok);
      if (rc && rc != STS_ERR_NONFINITE) return rc;
      if (rc == STS_ERR_NONFINITE || !ok) {  // abandon the attempt
        c->pstat[1]++;
        c->pstat[2] += n32;
        c->pstat[3] += n64;
        c->pred_wins = 0;
        if (c->pred_degree > 1) {
          c->pred_degree--;
        } else {
          c->pred_pause = c->pred_backoff;
          c->pred_backoff = std::min(2 * c->pred_backoff, mc->guard);
        }
        corrector_outcome(c, 0, true);
        if ((rc = restart_from_previous(c))) return rc;
        plain = true;
      } else {
        corrector_outcome(c, n32 + n64, false);
        c->pred_backoff = 1;
        if (++c->pred_wins >= 4 && c->pred_degree < mc->extrapolate) {
          c->pred_degree++;
          c->pred_wins = 0;
        }
      }
    } else if (c->pred_pause > 0) {
      c->pred_pause--;
    }
    if (plain) {  // from x^n: FP64 only, with the acceleration if asked for
      c->pstat[4] += !predict;
      n32 = 0;
      sts_mixed_control m64 = *mc;
      m64.N_I32 = 0;
      rc = run_phases(c, lc, &m64, &n32, &n64);
      if (!rc) c->plain_mean = c->plain_mean < 0.0 ? n64 : 0.5 * (c->plain_mean + n64);
    }
  }
  if (iters32) *iters32 = n32;
  if (iters64) *iters64 = n64;
  int rc2 = finish_step(c, n32 + n64, res_log, p_probe);
  return rc ? rc : rc2;
}

extern "C" int sts_predict_stats(sts_ctx *c, long long out[5]) {
  if (!c || !out) return fail(STS_ERR_ARG, "null argument");
  for (int q = 0; q < 5; q++) out[q] = c->pstat[q];
  return STS_OK;
}

// ---------------------------------------------------------------------------
// Warm start (include/sts_unet.h)
// ---------------------------------------------------------------------------
int sts_internal_view(sts_ctx *c, StsView *v) {
  if (!c || !v) return fail(STS_ERR_ARG, "null argument");
  v->Nx = c->Nx;
  v->Ny = c->Ny;
  for (int q = 0; q < STS_NFIELDS; q++) v->f[q] = c->f[q];
  v->mask = c->m;
  v->fluid_cells = c->fluid_cells;
  v->box_o[0] = c->g.Nx_o0;
  v->box_o[1] = c->g.Nx_o1;
  v->box_o[2] = c->g.Ny_o0;
  v->box_o[3] = c->g.Ny_o1;
  const int bu[4] = {c->g.Nx_u0, c->g.Nx_u1, c->g.Ny_u0, c->g.Ny_u1}, bv[4] = {c->g.Nx_v0, c->g.Nx_v1, c->g.Ny_v0,
                                                                              c->g.Ny_v1};
  for (int q = 0; q < 4; q++) {
    v->box_u[q] = bu[q];
    v->box_v[q] = bv[q];
    v->xe[q] = c->xs[q];
  }
  v->px = c->g.px;
  v->stream = c->stream;
  return STS_OK;
}

extern "C" int sts_keep_predicted_start(sts_ctx *c, int on) {
  if (!c) return fail(STS_ERR_ARG, "null context");
  HIP_OK(hipStreamSynchronize(c->stream));
  if (!on) {
    if (c->corr_net) return fail(STS_ERR_ARG, "a corrector is attached (it needs the kept start)");
    for (double *&q : c->xs) {
      (void)hipFree(q);
      q = nullptr;
    }
    return STS_OK;
  }
  for (double *&q : c->xs)
    if (!q) HIP_OK(hipMalloc(&q, sizeof(double) * (size_t)c->Na));
  return copy_start(c, c->stream);
}

extern "C" int sts_set_corrector(sts_ctx *c, sts_unet *net, const sts_unet_corr_norm *nrm) {
  if (!c || (net && !nrm)) return fail(STS_ERR_ARG, "null argument");
  if (net && !(nrm->u_ref > 0.0 && nrm->p_ref > 0.0 && nrm->step > 0.0 && nrm->corr > 0.0))
    return fail(STS_ERR_ARG, "corrector normalisation must be positive");
  c->corr_net = nullptr;
  c->corr_pause = 0;
  c->corr_backoff = 1;
  if (!net) return STS_OK;
  int rc = sts_keep_predicted_start(c, 1);
  if (rc) return rc;
  c->corr_net = net;
  c->corr_norm = *nrm;
  return STS_OK;
}

extern "C" int sts_corrector_stats(sts_ctx *c, long long out[2]) {
  if (!c || !out) return fail(STS_ERR_ARG, "null argument");
  out[0] = c->cstat[0];
  out[1] = c->cstat[1];
  return STS_OK;
}

extern "C" int sts_set_corrector_pause(sts_ctx *c, int max_pause) {
  if (!c || max_pause < 0) return fail(STS_ERR_ARG, "bad corrector pause");
  c->corr_max_pause = max_pause;
  c->corr_pause = 0;
  c->corr_backoff = 1;
  c->corr_kept_now = false;
  c->corr_ref_mean = -1.0;
  return STS_OK;
}

extern "C" int sts_corrector_skipped(sts_ctx *c, long long *skipped) {
  if (!c || !skipped) return fail(STS_ERR_ARG, "null argument");
  *skipped = c->cstat[2];
  return STS_OK;
}

void sts_internal_set_error(const std::string &msg) { g_err = msg; }

extern "C" int sts_time_step_ws(sts_ctx *c, sts_unet *net, const sts_unet_norm *nrm, const sts_loop_control *lc,
                                const sts_mixed_control *mc, int *used, int *iters32, int *iters64, double *res_log,
                                double *p_probe) {
  if (!c || !net || !nrm) return fail(STS_ERR_ARG, "null argument");
  int rc;
  if ((rc = check_mixed(c, lc, mc))) return rc;
  if ((rc = ensure_loop_buffers(c, lc->N_I + (mc ? mc->N_I32 : 0)))) return rc;
  if (!c->p_ws) {
    HIP_OK(hipMalloc(&c->p_ws, sizeof(double) * (size_t)c->Na));
    HIP_OK(hipMalloc(&c->T_ws, sizeof(double) * (size_t)c->Na));
  }
  if ((rc = sts_save_previous(c))) return rc;
  if ((rc = extrapolate_start(c, mc ? mc->extrapolate : 0))) return rc;
  StsView v;
  sts_internal_view(c, &v);
  if ((rc = sts_internal_unet_predict(net, v, nrm, c->p_ws, c->T_ws))) return rc;
  // First sweep of iteration 1 from both guesses; the coefficients do not
  // depend on the guess (assembly reads neither p nor T).
  HIP_OK(hipMemsetAsync(c->red, 0, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE, c->stream));
  HIP_OK(launch_assemble(c->g, c->fields(), c->stream, Gate{nullptr, 0}, c->diff_inline_ok));
  HIP_OK(launch_sweep(c->g, c->fields(), (const double *)c->f[STS_P], (const double *)c->f[STS_T], c->p_alt, c->T_alt,
                      c->red + 2, c->red + 3, c->stream));
  HIP_OK(launch_sweep(c->g, c->fields(), (const double *)c->p_ws, (const double *)c->T_ws, c->p_alt, c->T_alt,
                      c->red + 4, c->red + 5, c->stream));
  if ((rc = read_red(c, 6, false, false))) return rc;
  if (!std::isfinite(bits_to_double(c->h_red[2])) || !std::isfinite(bits_to_double(c->h_red[3])))
    return fail(STS_ERR_NONFINITE, "non-finite residual of the first sweep");
  const double rA = std::max(bits_to_double(c->h_red[2]) / lc->tol_p, bits_to_double(c->h_red[3]) / lc->tol_T);
  const double rB = std::max(bits_to_double(c->h_red[4]) / lc->tol_p, bits_to_double(c->h_red[5]) / lc->tol_T);
  const bool take = rB < rA;
  if (take) {
    HIP_OK(hipMemcpyAsync(c->f[STS_P], c->p_ws, sizeof(double) * (size_t)c->Na, hipMemcpyDeviceToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->f[STS_T], c->T_ws, sizeof(double) * (size_t)c->Na, hipMemcpyDeviceToDevice, c->stream));
  }
  if (used) *used = take ? 1 : 0;
  int n32 = 0, n64 = 0;
  rc = run_phases(c, lc, mc, &n32, &n64);
  if (iters32) *iters32 = n32;
  if (iters64) *iters64 = n64;
  int rc2 = finish_step(c, n32 + n64, res_log, p_probe);
  return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// Batched time steps (include/simple_ts.h, sts_batch_*)
// ---------------------------------------------------------------------------
struct sts_batch {
  std::vector<sts_ctx *> cs;
  hipStream_t stream = nullptr;
  BatchCase<double> *C64 = nullptr;
  BatchCase<float> *C32 = nullptr;
  LoopCtrl *ctrl = nullptr, *h_ctrl = nullptr;
  unsigned long long *lred = nullptr, *log = nullptr;
  int log_cap = 0;
  BatchTile *cell_t[NVARIANT] = {}, *sweep_t[NVARIANT] = {};
  int n_cell[NVARIANT] = {}, n_sweep[NVARIANT] = {};
  BatchTile *wall_t = nullptr;  // chunks of the cases' planned wall writes
  int n_wall = 0;
  // tiles of the cases whose loop is still open (rebuilt as cases stop):
  // device lists for the cell, sweep and wall launches, and their pinned
  // host staging (rewritten only after the stream has passed the last copy)
  BatchTile *act_t[3] = {}, *h_act = nullptr;
  int act_cap[3] = {};
  int cell_v = 0, sweep_v = 0;
  int last_iters32 = 0, last_iters64 = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // autotuner timing
  hipEvent_t ev_ctrl = nullptr;             // after the copy of ctrl to h_ctrl
};

extern "C" int sts_batch_destroy(sts_batch *b) {
  if (!b) return STS_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void *p : {(void *)b->C64, (void *)b->C32, (void *)b->ctrl, (void *)b->lred, (void *)b->log}) (void)hipFree(p);
  if (b->h_ctrl) (void)hipHostFree(b->h_ctrl);
  for (int v = 0; v < NVARIANT; v++) {
    (void)hipFree(b->cell_t[v]);
    (void)hipFree(b->sweep_t[v]);
  }
  (void)hipFree(b->wall_t);
  for (BatchTile *p : b->act_t) (void)hipFree(p);
  if (b->h_act) (void)hipHostFree(b->h_act);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->ev_ctrl) (void)hipEventDestroy(b->ev_ctrl);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return STS_OK;
}

// Tile list of the cases with open[c] (all when open is null) for tiles of
// TX x TY cells, into t.  Workgroup b runs on chiplet b % 8 (round-robin
// dispatch); every chiplet gets one contiguous run of the list (same order
// as xcd_tile).
static void tile_list(const sts_batch *b, int TX, int TY, const char *open, std::vector<BatchTile> &t) {
  std::vector<BatchTile> lt;
  for (int c = 0; c < (int)b->cs.size(); c++) {
    if (open && !open[c]) continue;
    const int nx = (b->cs[c]->Nx + TX - 1) / TX, ny = (b->cs[c]->Ny + TY - 1) / TY;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) lt.push_back(BatchTile{c, x, y});
  }
  const int nt = (int)lt.size(), per = nt / 8, rem = nt % 8;
  t.resize(lt.size());
  for (int bid = 0; bid < nt; bid++) {
    const int x = bid % 8;
    t[bid] = lt[x * per + std::min(x, rem) + bid / 8];
  }
}

// Chunks of WALL_OPS_PER_TILE planned wall writes of the cases with open[c].
static void wall_list(const sts_batch *b, const char *open, std::vector<BatchTile> &t) {
  t.clear();
  for (int c = 0; c < (int)b->cs.size(); c++)
    if (!open || open[c])
      for (int x = 0; x * WALL_OPS_PER_TILE < b->cs[c]->n_wall_ops; x++) t.push_back(BatchTile{c, x, 0});
}

static int build_tiles(sts_batch *b, const int (*shape)[2], BatchTile **out, int *count) {
  for (int v = 0; v < NVARIANT; v++) {
    std::vector<BatchTile> t;
    tile_list(b, shape[v][0], shape[v][1], nullptr, t);
    HIP_OK(hipMalloc(&out[v], sizeof(BatchTile) * (t.empty() ? 1 : t.size())));
    HIP_OK(hipMemcpyAsync(out[v], t.data(), sizeof(BatchTile) * t.size(), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    count[v] = (int)t.size();
  }
  return STS_OK;
}

// Launch lists of one loop 2: the full lists, or those of the open cases.
struct BatchLists {
  const BatchTile *cell, *sweep, *wall;
  int n_cell, n_sweep, n_wall, n_open;
};

// Rebuilds the active lists for the cases with open[c] and enqueues their
// upload on the batch stream (kernels enqueued before it still read the
// previous lists, which held a superset of these tiles).  Called only right
// after the host waited for an event that follows the previous upload, so
// the pinned staging is free.
static int compact_lists(sts_batch *b, const std::vector<char> &open, BatchLists &L) {
  std::vector<BatchTile> t[3];
  tile_list(b, CELL_TILES[b->cell_v][0], CELL_TILES[b->cell_v][1], open.data(), t[0]);
  tile_list(b, SWEEP_TILES[b->sweep_v][0], SWEEP_TILES[b->sweep_v][1], open.data(), t[1]);
  wall_list(b, open.data(), t[2]);
  size_t off = 0;
  for (int q = 0; q < 3; q++) {
    if ((int)t[q].size() > b->act_cap[q]) return fail(STS_ERR_ARG, "active tile list exceeds the full list");
    if (!t[q].empty()) {
      std::memcpy(b->h_act + off, t[q].data(), sizeof(BatchTile) * t[q].size());
      HIP_OK(hipMemcpyAsync(b->act_t[q], b->h_act + off, sizeof(BatchTile) * t[q].size(), hipMemcpyHostToDevice,
                            b->stream));
    }
    off += (size_t)b->act_cap[q];
  }
  L.cell = b->act_t[0];
  L.sweep = b->act_t[1];
  L.wall = b->act_t[2];
  L.n_cell = (int)t[0].size();
  L.n_sweep = (int)t[1].size();
  L.n_wall = (int)t[2].size();
  L.n_open = 0;
  for (char o : open) L.n_open += o ? 1 : 0;
  return STS_OK;
}

extern "C" int sts_batch_create(sts_ctx *const *ctxs, int n, sts_batch **out) {
  if (!ctxs || !out || n < 1) return fail(STS_ERR_ARG, "need at least one context");
  *out = nullptr;
  sts_batch *b = new sts_batch();
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) {
      sts_batch_destroy(b);
      return fail(STS_ERR_ARG, "null context in batch");
    }
    b->cs.push_back(ctxs[i]);
  }
  auto bail = [&](int rc) {
    sts_batch_destroy(b);
    return rc;
  };
#define B_OK(expr)                                                                                            \
  do {                                                                                                        \
    hipError_t e_ = (expr);                                                                                   \
    if (e_ != hipSuccess) return bail(fail(STS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
  } while (0)
  B_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  B_OK(hipMalloc(&b->C64, sizeof(BatchCase<double>) * n));
  B_OK(hipMalloc(&b->C32, sizeof(BatchCase<float>) * n));
  B_OK(hipMalloc(&b->ctrl, sizeof(LoopCtrl) * n));
  B_OK(hipHostMalloc(&b->h_ctrl, sizeof(LoopCtrl) * n, hipHostMallocDefault));
  B_OK(hipMalloc(&b->lred, sizeof(unsigned long long) * NSLOT * SLOT_STRIDE * (size_t)n));
  B_OK(hipEventCreateWithFlags(&b->ev0, hipEventDisableSystemFence));
  B_OK(hipEventCreateWithFlags(&b->ev1, hipEventDisableSystemFence));
  B_OK(hipEventCreateWithFlags(&b->ev_ctrl, hipEventDisableTiming));
#undef B_OK
  int rc;
  if ((rc = build_tiles(b, CELL_TILES, b->cell_t, b->n_cell)) || (rc = build_tiles(b, SWEEP_TILES, b->sweep_t, b->n_sweep)))
    return bail(rc);
  {
    std::vector<BatchTile> wt;
    wall_list(b, nullptr, wt);
    b->n_wall = (int)wt.size();
    if (hipMalloc(&b->wall_t, sizeof(BatchTile) * (wt.empty() ? 1 : wt.size())) != hipSuccess ||
        (!wt.empty() && hipMemcpyAsync(b->wall_t, wt.data(), sizeof(BatchTile) * wt.size(), hipMemcpyHostToDevice,
                                       b->stream) != hipSuccess) ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return bail(fail(STS_ERR_HIP, "wall tiles"));
  }
  b->act_cap[0] = *std::max_element(b->n_cell, b->n_cell + NVARIANT);
  b->act_cap[1] = *std::max_element(b->n_sweep, b->n_sweep + NVARIANT);
  b->act_cap[2] = b->n_wall;
  for (int q = 0; q < 3; q++)
    if (hipMalloc(&b->act_t[q], sizeof(BatchTile) * std::max(1, b->act_cap[q])) != hipSuccess)
      return bail(fail(STS_ERR_HIP, "active tile lists"));
  if (hipHostMalloc(&b->h_act, sizeof(BatchTile) * std::max(1, b->act_cap[0] + b->act_cap[1] + b->act_cap[2]),
                    hipHostMallocDefault) != hipSuccess)
    return bail(fail(STS_ERR_HIP, "active tile staging"));
  if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(fail(STS_ERR_HIP, "stream sync"));
  *out = b;
  return STS_OK;
}

extern "C" int sts_batch_set_tiles(sts_batch *b, int cell_variant, int sweep_variant) {
  if (!b || cell_variant < 0 || cell_variant >= NVARIANT || sweep_variant < 0 || sweep_variant >= NVARIANT)
    return fail(STS_ERR_ARG, "bad tile variant");
  b->cell_v = cell_variant;
  b->sweep_v = sweep_variant;
  return STS_OK;
}

// Case table of one precision at the current p/T pointers of every case.
template <typename R>
static int upload_cases(sts_batch *b, BatchCase<R> *dst, int N_I, double tol_uv, double tol_p, double tol_T,
                        int stall_w, double stall_ratio, size_t log_off, bool gated,
                        const std::vector<char> *skip = nullptr, const std::vector<char> *inline0 = nullptr) {
  std::vector<BatchCase<R>> h(b->cs.size());
  for (size_t i = 0; i < b->cs.size(); i++) {
    sts_ctx *c = b->cs[i];
    BatchCase<R> &x = h[i];
    if constexpr (sizeof(R) == sizeof(double)) {
      x.g = c->g;
      x.f = c->fields();
      x.p_alt = c->p_alt;
      x.T_alt = c->T_alt;
    } else {
      x.g = c->g32;
      x.f = c->fields32();
      x.p_alt = c->p_alt32;
      x.T_alt = c->T_alt32;
    }
    x.ctrl = gated ? b->ctrl + i : nullptr;
    x.lred = b->lred + i * (size_t)NSLOT * SLOT_STRIDE;
    x.log = b->log ? b->log + i * 6 * (size_t)b->log_cap + log_off : nullptr;
    x.probe = c->probe;
    x.N_I = N_I;
    x.stall_w = stall_w;
    x.tol_uv = tol_uv;
    x.tol_p = tol_p;
    x.tol_T = tol_T;
    x.stall_ratio = stall_ratio;
    x.wops = c->wall_ops;
    x.nwops = c->n_wall_ops;
    x.skip = skip ? (*skip)[i] : 0;
    x.inline0 = inline0 ? (*inline0)[i] : 0;
    x.aa = c->aa;
    x.aa.nb = aa_case_tiles(c->Nx, c->Ny, b->cell_v);
    for (int j = 0; j < MAX_AA; j++) {
      x.aa_dG[j] = c->aa_dG[j];
      x.aa_dF[j] = c->aa_dF[j];
    }
  }
  HIP_OK(hipMemcpyAsync(dst, h.data(), sizeof(BatchCase<R>) * h.size(), hipMemcpyHostToDevice, b->stream));
  return STS_OK;
}

// Loop 2 of every case in shared launches; iters[i] gets case i's count.
template <typename R>
static int run_batch_loop2(sts_batch *b, BatchCase<R> *C, int n_pc, int N_I, int batch, std::vector<int> &iters,
                           int aa_m = 0, int aa_every = 1) {
  const int n = (int)b->cs.size(), passes = sweep_passes(b->cs[0], n_pc);
  hipStream_t s = b->stream;
  const int cv = b->cell_v, sv = b->sweep_v;
  // launches cover the tiles of every case until the host has seen cases
  // stop; from then on only those of the open ones (a stopped case's tiles
  // would only read its control word and return)
  BatchLists L{b->cell_t[cv], b->sweep_t[sv], b->wall_t, b->n_cell[cv], b->n_sweep[sv], b->n_wall, n};
  std::vector<char> open_c(n, 1);
  HIP_OK(launch_batch<R>(BK_LOOP_OPEN, 0, C, n, nullptr, 0, 0, 0, 0, 0, s));
  if constexpr (sizeof(R) == sizeof(double))
    if (aa_m > 0) HIP_OK(launch_batch_aa(BKA_BEGIN, cv, C, n, L.cell, L.n_cell, 0, 0, AABatchSel{}, s));
  // the loop control is read one iteration before a batch ends (as in the
  // single-context loop), after the last one of a batch that ends at N_I
  constexpr int LOOKAHEAD = 1;
  int k = 0;
  while (true) {
    const int end = std::min(N_I, k + batch);
    const int check = end == N_I ? end - 1 : std::max(k, end - 1 - LOOKAHEAD);
    for (; k < end; k++) {
      int par = (int)(((long long)k * passes) & 1);
      if (k == 0) HIP_OK(launch_batch<R>(BK_ASSEMBLE_ARR, cv, C, n, L.cell, L.n_cell, k, par, 0, 0, s));
      HIP_OK(launch_batch<R>(BK_ASSEMBLE, cv, C, n, L.cell, L.n_cell, k, par, 0, 0, s));
      for (int q = 0; q < n_pc;) {
        if (b->cs[0]->fuse_sweeps && q + 1 < n_pc) {
          HIP_OK(launch_batch<R>(BK_SWEEP2, sv, C, n, L.sweep, L.n_sweep, k, par, q == 0, q + 1 == n_pc - 1, s));
          q += 2;
        } else {
          HIP_OK(launch_batch<R>(BK_SWEEP, cv, C, n, L.cell, L.n_cell, k, par, q == 0, q == n_pc - 1, s));
          q += 1;
        }
        par ^= 1;
      }
      HIP_OK(launch_batch<R>(BK_CORRECT, cv, C, n, L.cell, L.n_cell, k, par, 0, 0, s));
      HIP_OK(launch_batch<R>(BK_WALL_OPS, 0, C, n, L.wall, L.n_wall, k, par,
                             STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP, 0, s));
      HIP_OK(launch_batch<R>(BK_EDGES, 0, C, n, nullptr, 0, k, par,
                             STS_BC_VELOCITIES | STS_BC_TEMPERATURE | STS_BC_PRESSURE, 0, s));
      if constexpr (sizeof(R) == sizeof(double))
        if (aa_m > 0 && (k + 1) % aa_every == 0) {  // as enqueue_anderson, for every case at once
          const int K = (k + 1) / aa_every - 1;
          AABatchSel bs{};
          bs.n = std::min(K, aa_m);
          bs.keep_dF = aa_m > 1;
          bs.start = aa_start(K, aa_m);
          bs.write_D = aa_m > 1;
          for (int j = 0; j < bs.n; j++) bs.slot[j] = (K - bs.n + j) % aa_m;
          HIP_OK(launch_batch_aa(BKA_STORE, cv, C, n, L.cell, L.n_cell, k, par, bs, s));
          if (bs.n > 0) {
            HIP_OK(launch_batch_aa(BKA_SOLVE, cv, C, n, L.cell, L.n_cell, k, par, bs, s));
            HIP_OK(launch_batch_aa(BKA_MIX, cv, C, n, L.cell, L.n_cell, k, par, bs, s));
          }
        }
      HIP_OK(launch_batch<R>(BK_FLUX_DIFF, cv, C, n, L.cell, L.n_cell, k, par, 0, 0, s));
      if (k == check) {
        HIP_OK(hipMemcpyAsync(b->h_ctrl, b->ctrl, sizeof(LoopCtrl) * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipEventRecord(b->ev_ctrl, s));
      }
    }
    HIP_OK(hipEventSynchronize(b->ev_ctrl));
    int n_open = 0;
    for (int i = 0; i < n; i++) {
      open_c[i] = b->h_ctrl[i].stop_at == 0x7fffffff;
      n_open += open_c[i];
    }
    if (n_open == 0) break;
    if (check == N_I - 1) return fail(STS_ERR_ARG, "loop 2 did not stop at N_I");
    if (n_open < L.n_open) {
      int rc;
      if ((rc = compact_lists(b, open_c, L))) return rc;
    }
    batch = 2;
  }
  iters.assign(n, 0);
  for (int i = 0; i < n; i++) {
    iters[i] = b->h_ctrl[i].stop_at + 1;
    sts_ctx *c = b->cs[i];
    if (((long long)iters[i] * passes) & 1) {
      if constexpr (sizeof(R) == sizeof(double)) {
        std::swap(c->f[STS_P], c->p_alt);
        std::swap(c->f[STS_T], c->T_alt);
      } else {
        std::swap(c->f32[STS_P], c->p_alt32);
        std::swap(c->f32[STS_T], c->T_alt32);
      }
    }
  }
  for (int i = 0; i < n; i++)
    if (b->h_ctrl[i].nonfinite)
      return fail(STS_ERR_NONFINITE, "non-finite residual in case " + std::to_string(i) + ", outer iteration " +
                                         std::to_string(iters[i]));
  return STS_OK;
}

extern "C" int sts_batch_time_step(sts_batch *b, const sts_loop_control *lc, const sts_mixed_control *mc,
                                   int *iters32, int *iters64) {
  if (!b) return fail(STS_ERR_ARG, "null batch");
  int rc;
  const int n = (int)b->cs.size();
  for (sts_ctx *c : b->cs) {
    if ((rc = check_mixed(c, lc, mc))) return rc;
    if (c->fuse_sweeps != b->cs[0]->fuse_sweeps) return fail(STS_ERR_ARG, "cases differ in sweep fusion");
    if (mc && mc->anderson > 0 && (rc = ensure_anderson(c, mc->anderson))) return rc;
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  const int cap = lc->N_I + (mc ? mc->N_I32 : 0);
  if (cap > b->log_cap) {
    (void)hipFree(b->log);
    b->log = nullptr;
    HIP_OK(hipMalloc(&b->log, sizeof(unsigned long long) * 6 * (size_t)cap * n));
    b->log_cap = cap;
  }
  hipStream_t s = b->stream;
  for (sts_ctx *c : b->cs) {  // prologue S:3233-3239
    const double *src[5] = {c->f[STS_P], c->f[STS_U], c->f[STS_V], c->f[STS_T], c->f[STS_RHO]};
    double *dst[5] = {c->f[STS_P_PR], c->f[STS_U_PR], c->f[STS_V_PR], c->f[STS_T_PR], c->f[STS_RHO_PR]};
    HIP_OK(launch_copy5(src, dst, c->Na, s));
  }
  // predicted starts; with a guard, per case: a predicted step that needed
  // more outer iterations than the case's recent steps from x^n pauses its
  // prediction for 1, 2, 4, ... steps (no restart inside shared launches)
  std::vector<char> predicted(n, 0);
  for (int i = 0; i < n; i++) {
    sts_ctx *c = b->cs[i];
    const int q = mc ? mc->extrapolate : 0;
    bool pred = q > 0;
    if (pred && mc->guard > 0) {
      pred = c->n_hist > 0 && c->plain_mean > 0.0 && c->pred_pause == 0;
      if (!pred && c->pred_pause > 0) c->pred_pause--;
    }
    predicted[i] = pred && c->n_hist > 0;
    if ((rc = extrapolate_start(c, q, s, !pred))) return rc;
  }
  std::vector<int> n32(n, 0), n64(n, 0);
  // cases the FP32 gate sends straight to FP64 sit the FP32 phase out
  std::vector<char> skip32(n, 1);
  bool any32 = false;
  for (int i = 0; i < n; i++) {
    skip32[i] = fp32_phase(b->cs[i], mc) ? 0 : 1;
    any32 = any32 || !skip32[i];
  }
  // the first FP64 outer iteration forms Gamma and D itself where the
  // single-case step would (after an FP32 phase, or when forming them gives
  // the values the arrays hold); the FP32 phase starts from the narrowed
  // arrays, brought up to date first
  std::vector<char> inl64(n, 0), inl32(n, 0);
  for (int i = 0; i < n; i++) {
    inl64[i] = (!skip32[i] || b->cs[i]->diff_inline_ok) ? 1 : 0;
    if (!skip32[i] && (rc = refresh_diff(b->cs[i], s))) return rc;
  }
  if (any32) {
    const unsigned long long all = (1ull << STS_NFIELDS) - 1ull;
    for (int i = 0; i < n; i++)
      if (!skip32[i]) HIP_OK(launch_convert(b->cs[i]->fields(), b->cs[i]->fields32(), all, b->cs[i]->Na, s));
    if ((rc = upload_cases<float>(b, b->C32, mc->N_I32, mc->tol32_uv, mc->tol32_p, mc->tol32_T, mc->stall_w,
                                  mc->stall_ratio, 0, true, &skip32, &inl32)))
      return rc;
    const int b32 = lc->batch > 0 ? lc->batch : (b->last_iters32 > 2 ? b->last_iters32 - 1 : 1);
    if ((rc = run_batch_loop2<float>(b, b->C32, lc->N_I_p_c, mc->N_I32, b32, n32))) return rc;
    int lo = mc->N_I32;
    for (int i = 0; i < n; i++)
      if (!skip32[i]) lo = std::min(lo, n32[i]);
    b->last_iters32 = lo;
    for (int i = 0; i < n; i++)
      if (!skip32[i]) HIP_OK(launch_convert(b->cs[i]->fields32(), b->cs[i]->fields(), WIDEN_MASK, b->cs[i]->Na, s));
  }
  // the FP64 logs follow the FP32 ones of each case (offset = the largest FP32 count)
  const int off = mc ? mc->N_I32 : 0;
  if ((rc = upload_cases<double>(b, b->C64, lc->N_I, lc->tol_uv, lc->tol_p, lc->tol_T, 0, 0.0, 6 * (size_t)off, true,
                                 nullptr, &inl64)))
    return rc;
  const int b64 = lc->batch > 0 ? lc->batch : (b->last_iters64 > 2 ? b->last_iters64 - 1 : 1);
  if ((rc = run_batch_loop2<double>(b, b->C64, lc->N_I_p_c, lc->N_I, b64, n64, mc ? mc->anderson : 0,
                                    mc ? std::max(1, mc->anderson_every) : 1)))
    return rc;
  b->last_iters64 = *std::min_element(n64.begin(), n64.end());
  for (int i = 0; i < n; i++) {
    if (iters32) iters32[i] = n32[i];
    if (iters64) iters64[i] = n64[i];
    sts_ctx *c = b->cs[i];
    c->last_iters = n64[i];
    c->last_total = n32[i] + n64[i];
    c->diff_inline_ok = c->diff_stale = true;  // the batched passes write only the fluxes
    if (mc && mc->extrapolate > 0 && mc->guard > 0) {
      if (!predicted[i]) {
        c->plain_mean = c->plain_mean < 0.0 ? n64[i] : 0.5 * (c->plain_mean + n64[i]);
        c->pstat[4]++;
      } else {
        c->pstat[0]++;
        if (n64[i] > c->plain_mean) {
          c->pstat[1]++;
          c->pred_pause = c->pred_backoff;
          c->pred_backoff = std::min(2 * c->pred_backoff, mc->guard);
        } else {
          c->pred_backoff = 1;
        }
      }
    }
  }
  HIP_OK(hipStreamSynchronize(s));
  return STS_OK;
}

// Times the bulk kernels of each tile variant on the current state (the
// timed launches only rewrite what the next outer iteration recomputes before
// reading it: assembly outputs, fluxes from unchanged inputs, and the p/T
// partners), then keeps the fastest.  The assembly is timed in the form the
// loops use after their first iteration (Gamma and D formed in the tile).
extern "C" int sts_batch_autotune(sts_batch *b, int reps, int *cell_variant, int *sweep_variant, double *ms) {
  if (!b || reps < 1) return fail(STS_ERR_ARG, "bad autotune arguments");
  const int n = (int)b->cs.size();
  for (sts_ctx *c : b->cs) HIP_OK(hipStreamSynchronize(c->stream));
  int rc;
  const std::vector<char> inl(n, 1);
  if ((rc = upload_cases<double>(b, b->C64, 1, 0, 0, 0, 0, 0.0, 0, false, nullptr, &inl))) return rc;
  hipStream_t s = b->stream;
  double t_cell[NVARIANT], t_sweep[NVARIANT];
  for (int v = 0; v < NVARIANT; v++) {
    for (int pass = 0; pass < 2; pass++) {  // pass 0 warms up
      HIP_OK(hipEventRecord(b->ev0, s));
      for (int r = 0; r < reps; r++) {
        HIP_OK(launch_batch<double>(BK_ASSEMBLE, v, b->C64, n, b->cell_t[v], b->n_cell[v], 0, 0, 0, 0, s));
        HIP_OK(launch_batch<double>(BK_FLUX_DIFF, v, b->C64, n, b->cell_t[v], b->n_cell[v], 0, 0, 0, 0, s));
      }
      HIP_OK(hipEventRecord(b->ev1, s));
      HIP_OK(hipEventSynchronize(b->ev1));
      float f = 0;
      HIP_OK(hipEventElapsedTime(&f, b->ev0, b->ev1));
      t_cell[v] = f / reps;
      HIP_OK(hipEventRecord(b->ev0, s));
      for (int r = 0; r < reps; r++)
        HIP_OK(launch_batch<double>(BK_SWEEP2, v, b->C64, n, b->sweep_t[v], b->n_sweep[v], 0, 0, 0, 0, s));
      HIP_OK(hipEventRecord(b->ev1, s));
      HIP_OK(hipEventSynchronize(b->ev1));
      HIP_OK(hipEventElapsedTime(&f, b->ev0, b->ev1));
      t_sweep[v] = f / reps;
    }
  }
  int bc = 0, bs = 0;
  for (int v = 1; v < NVARIANT; v++) {
    if (t_cell[v] < t_cell[bc]) bc = v;
    if (t_sweep[v] < t_sweep[bs]) bs = v;
  }
  b->cell_v = bc;
  b->sweep_v = bs;
  if (cell_variant) *cell_variant = bc;
  if (sweep_variant) *sweep_variant = bs;
  if (ms)
    for (int v = 0; v < NVARIANT; v++) {
      ms[v] = t_cell[v];
      ms[NVARIANT + v] = t_sweep[v];
    }
  return STS_OK;
}

// ---------------------------------------------------------------------------
// Drag coefficient sums (WriteDragCoefficient, S:8147-8430)
// ---------------------------------------------------------------------------
extern "C" int sts_drag_sums(sts_ctx *c, int body, double *sums) {
  if (!c || !sums || body < 0 || body >= c->g.nbodies) return fail(STS_ERR_ARG, "bad body index");
  if (c->drag_cells.empty()) {
    c->drag_cells.assign(c->g.nbodies, nullptr);
    c->drag_n.assign(c->g.nbodies, -1);
  }
  if (c->drag_n[body] < 0) {
    // the body's cells of the p/T box in the reference's loop order (i outer, j inner)
    std::vector<unsigned> cells;
    const KGrid &g = c->g;
    for (int i = g.Nx_o0; i < g.Nx_o1; i++)
      for (int j = g.Ny_o0; j < g.Ny_o1; j++) {
        const unsigned node = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
        if ((long long)c->h_vol[node] - 1 == body) cells.push_back(node);
      }
    HIP_OK(hipMalloc(&c->drag_cells[body], sizeof(unsigned) * (cells.empty() ? 1 : cells.size())));
    HIP_OK(hipMemcpy(c->drag_cells[body], cells.data(), sizeof(unsigned) * cells.size(), hipMemcpyHostToDevice));
    c->drag_n[body] = (int)cells.size();
  }
  const size_t need = 12 * (size_t)(c->drag_n[body] > 0 ? c->drag_n[body] : 1);
  if (need > c->drag_cap) {
    (void)hipFree(c->drag_terms);
    c->drag_terms = nullptr;
    HIP_OK(hipMalloc(&c->drag_terms, sizeof(double) * need));
    c->drag_cap = need;
  }
  if (!c->drag_out) HIP_OK(hipMalloc(&c->drag_out, sizeof(double) * 12));
  HIP_OK(launch_drag(c->g, c->fields(), c->drag_cells[body], c->drag_n[body], c->drag_terms, c->drag_out, c->stream));
  HIP_OK(hipMemcpyAsync(sums, c->drag_out, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return STS_OK;
}
