// lq_capi.hip -- extern "C" boundary of libretainwall.so (include/retainwall.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "lq_internal.hpp"
#include "lq_launch.hpp"

using namespace lq;

struct TimedPair {
  int cls;
  hipEvent_t a, b;
};

struct lq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // COLUMNS steps: the corner search beside the interaction
  hipEvent_t ev_fork = nullptr, ev_adami = nullptr, ev_join = nullptr;
  int overlap = 1;             // use the side stream (lq_set_overlap)
  lq_params par{};
  DevParams hp{};          // host image of the static part
  DevParams* dP = nullptr;
  Buffers B{};
  // alternate *_ini set for SortArraysAll
  float4 *pos_ini_alt = nullptr, *vel_ini_alt = nullptr, *str_ini_a_alt = nullptr, *stn_ini_a_alt = nullptr;
  float2 *str_ini_b_alt = nullptr, *stn_ini_b_alt = nullptr;
  std::vector<void*> allocs;
  std::string err;
  bool uploaded = false;
  bool searched = false;  // a neighbour search has left sorted keys in B.part_cel
  int nbits = 0;
  int npm0 = 0;
  int tile_shape = 1, tile_tz = 7;
  int col_wide = 0;  // COLUMNS: register budget of 128 VGPRs instead of 96
  int fuse = 0;      // COLUMNS steps: predictor / corrector fused into the interaction
  int xcd_remap = 1;  // renumber interaction blocks per XCD (lq_set_interaction / lq_autotune)
  long long host_step = 0;
  // instrumentation
  bool timing = false;
  unsigned timing_mask = ~0u;  // kernel classes timed while timing is on
  std::vector<TimedPair> pending;
  std::vector<hipEvent_t> pool;
  double ms[LQ_K_CLASS_COUNT] = {0};
  int64_t launches[LQ_K_CLASS_COUNT] = {0};
  // graph
  hipGraphExec_t graph = nullptr;
  long long layout_version = 0, graph_version = -1;
  // A graph captured with timing on keeps its template (node handles) and,
  // per timed section, the two event-record nodes; each replay points them
  // at fresh events so every launch in a run of replays is timed.
  bool graph_timed = false;
  unsigned graph_mask = ~0u;
  bool capturing = false;  // section events become event-record nodes
  hipGraph_t graph_tmpl = nullptr;
  struct TimedNodes {
    int cls;
    hipGraphNode_t a, b;
  };
  std::vector<TimedNodes> graph_sections;
  // cases (a plain context is a batch of one)
  int n_cases = 1;
  std::vector<lq_params> case_par;  // each case's own parameters
  std::vector<DevCase> cases;       // layout in the global grid, table bases
  std::vector<int> bnd_base, soil_base;  // first wall / soil slot of each case at upload
  std::vector<char> case_uploaded;
  int n_uploaded = 0;
  float (*d_model)[LQ_MATERIAL_SIZE_MAX] = nullptr;
  HypoConst* d_hypo = nullptr;
  DevMotion* d_mv = nullptr;
  DevTable* d_tab = nullptr;
  DevCase* d_cases = nullptr;
  float* d_boxes = nullptr;
  int n_model_rows = 0, n_motion_rows = 0;
  // HypoCS surrogate
  int n_cu = 1;
  int sur_members = 0;
  float* d_sur_norm = nullptr;  // x_mean x_std y_mean y_std
  float* d_sur_lo = nullptr;
  float* d_sur_hi = nullptr;
  unsigned long long* d_sur_stats = nullptr;
  unsigned* d_sur_amax = nullptr;
  int log_capacity = 0;
  // reductions
  unsigned long long* wf_host = nullptr;  // pinned copy of the per-case wall-force accumulators
};

#define HIPCHECK(ctx, call)                                                                  \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
      return LQ_ERR_HIP;                                                                     \
    }                                                                                        \
  } while (0)

static int fail(lq_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

template <class T> static int dalloc(lq_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return LQ_ERR_HIP;
  }
  hipMemsetAsync(q, 0, bytes, c->stream);
  c->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

// ------------------------------------------------------- timing helpers --
static hipEvent_t ev_get(lq_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}
// During stream capture an event record is kept in the graph as a node only
// with the external flag (a plain record just orders the capture).
static void ev_record(lq_ctx* c, hipEvent_t e, hipStream_t s) {
  if (c->capturing) hipEventRecordWithFlags(e, s, hipEventRecordExternal);
  else hipEventRecord(e, s);
}
// A timed section on the library's stream, or on the side stream for work
// that the main stream joins before it reaches its next timed section.
struct Section {
  lq_ctx* c;
  int cls;
  hipStream_t s;
  hipEvent_t a = nullptr;
  Section(lq_ctx* c_, int cls_, hipStream_t s_ = nullptr) : c(c_), cls(cls_), s(s_ ? s_ : c_->stream) {
    if (on()) {
      a = ev_get(c);
      ev_record(c, a, s);
    }
  }
  bool on() const { return c->timing && ((c->timing_mask >> cls) & 1u); }
  ~Section() {
    if (on()) {
      hipEvent_t b = ev_get(c);
      ev_record(c, b, s);
      c->pending.push_back({cls, a, b});
    }
  }
};
// Folds the pending event pairs into the per-class totals (synchronises the
// stream).  lq_step also does it between replays once kTimingPending pairs
// wait, so a long timed run holds a bounded number of events.
constexpr size_t kTimingPending = 8192;
static void resolve_timing(lq_ctx* c) {
  if (c->pending.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    c->ms[p.cls] += ms;
    c->launches[p.cls] += 1;
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

// --------------------------------------------------------------- params --
extern "C" int lq_abi_sizes(size_t* sizes, int n) {
  size_t v[4] = {sizeof(lq_params), sizeof(lq_motion_table), sizeof(lq_rigid_motion), sizeof(lq_host_state)};
  for (int k = 0; k < n && k < 4; k++) sizes[k] = v[k];
  return 0;
}

// InitializeParameters, loquat.cu:130-190.
extern "C" int lq_initialize_parameters(lq_params* p) {
  if (!p || p->dr <= 0.f || p->hdr <= 0.f) return LQ_ERR_ARG;
  p->h = p->dr * p->hdr;
  p->i_h = 1.0f / p->h;
  p->eta = 0.01f * p->h * p->h;
  p->bound_thickness = float(std::ceil(2.0 * double(p->hdr)) * double(p->dr));
  const double hd = double(p->h), pi = 3.14159265358979323846;
  switch (p->dim) {
    case 1:
      p->vol_def = p->dr;
      p->cell_num[0] = 1;
      p->cell_num[1] = 1;
      p->kernel_normalization_par = float(3.0 / (32 * hd));
      p->kernel_gradient_normalization_par = float(-15.0 / (32 * std::pow(hd, 3)));
      break;
    case 2:
      p->vol_def = float(std::pow(double(p->dr), 2));
      p->cell_num[1] = 1;
      p->cell_num[0] = int(0.5f * p->i_h * p->domain_size[0]) + 1;
      p->kernel_normalization_par = float(7.0 / (32.0 * pi * std::pow(hd, 2)));
      p->kernel_gradient_normalization_par = float(-35.0 / (32 * pi * std::pow(hd, 4)));
      break;
    case 3:
      p->vol_def = float(std::pow(double(p->dr), 3));
      p->cell_num[0] = int(0.5f * p->i_h * p->domain_size[0]) + 1;
      p->cell_num[1] = int(0.5f * p->i_h * p->domain_size[1]) + 1;
      p->kernel_normalization_par = float(42.0 / (256.0 * pi * std::pow(hd, 3)));
      p->kernel_gradient_normalization_par = float(-5.0 * 42 / (256.0 * pi * std::pow(hd, 5)));
      break;
    default:
      return LQ_ERR_ARG;
  }
  p->cell_num[2] = int(0.5f * p->i_h * p->domain_size[2]) + 1;
  const int layer = p->cell_num[0] * p->cell_num[1] * p->cell_num[2];
  p->ncell_material = layer;
  p->ncell_delete_m = 2 * layer;
  p->ncell = p->ncell_delete_m + 1;
  p->np_delete_m = 0;
  const float q = p->dr * p->i_h;
  p->kernel_zero = p->kernel_normalization_par * 2.0f * 2.0f * 2.0f * 2.0f * 0.5f;
  p->kernel_dr = p->kernel_normalization_par * (2.0f - q) * (2.0f - q) * (2.0f - q) * (2.0f - q) * (q + 0.5f);
  return LQ_OK;
}

// Largest float r2 with fl(fl(sqrt(r2)) * i_h) < 2: the kernels keep a
// candidate pair when r^2 <= r2_cut, which is the reference's q < 2 test
// (particle_interaction.cu:33-35) bit for bit, because correctly rounded
// sqrt and multiply are monotone.  Non-negative floats order like their bit
// patterns, so the cut is found by bisection over [0, +inf).
extern "C" int lq_interaction_cutoff_sq(float i_h, float* r2_cut) {
  if (!r2_cut || !(i_h > 0.f) || !std::isfinite(i_h)) return LQ_ERR_ARG;
  auto keep = [i_h](uint32_t u) {
    float r2;
    std::memcpy(&r2, &u, sizeof r2);
    return std::sqrt(r2) * i_h < 2.0f;
  };
  uint32_t lo = 0u, hi = 0x7f800000u;  // keep(lo), !keep(hi)
  while (hi - lo > 1u) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (keep(mid)) lo = mid; else hi = mid;
  }
  std::memcpy(r2_cut, &lo, sizeof lo);
  return LQ_OK;
}

// Pair-list capacity per soil slot (a regular lattice at h = 1.3 dr gives
// about 80 pairs inside 2h; overflow falls back to the cell scan).
constexpr int kPairListCap = 160;

static void fill_static(const lq_params& p, DevParams& d) {
  std::memset(&d, 0, sizeof(d));
  d.np = p.np; d.npb = p.npb;
  d.cnx = p.cell_num[0]; d.cny = p.cell_num[1]; d.cnz = p.cell_num[2];
  d.ncell = p.ncell; d.ncell_material = p.ncell_material; d.ncell_delete_m = p.ncell_delete_m;
  d.key_mode = p.key_mode; d.shepard = p.stress_shepard_filter; d.viscosity_type = p.viscosity_type;
  d.n_motion = p.n_motion; d.n_model = p.n_model;
  d.h = p.h; d.i_h = p.i_h; d.dr = p.dr; d.vol_def = p.vol_def; d.bt = p.bound_thickness; d.eta = p.eta;
  d.two_h = 2.0f * p.h;
  lq_interaction_cutoff_sq(p.i_h, &d.r2cut);
  {
    // List radius 1.01 * 2h.  A pair left out at the build is at least
    // 1.01 * 2h apart (up to float rounding, ~1e-6 relative); after moves of
    // at most 0.004 * 2h per particle it is still >= 1.002 * 2h apart, so it
    // fails the r^2 <= r2cut test exactly as the cell scan would.
    const float rl = d.two_h * (1.0f + 0.01f), dl = d.two_h * (0.4f * 0.01f);
    d.r2list = rl * rl;
    d.disp2_lim = dl * dl;
    d.plist_cap = kPairListCap;
    d.plist_stride = std::max(1, p.npm);
  }
  // sub-column windows of the COLUMNS variant: the cut-off radius squared
  // with 1e-3 relative headroom, and a margin of 1% of h for float rounding
  // and for particles the predictor moved past their sub-cell
  d.col_r2 = d.r2cut * 1.001f;
  d.col_margin = 0.01f * p.h;
  d.gx = p.gravity[0]; d.gy = p.gravity[1]; d.gz = p.gravity[2];
  d.damping = p.damping;
  d.knorm = p.kernel_normalization_par; d.kgrad = p.kernel_gradient_normalization_par;
  d.kernel_zero = p.kernel_zero; d.kernel_dr = p.kernel_dr;
  d.art_exp = p.kernel_zero / p.kernel_dr;
  d.alpha_cs = p.artificial_viscosity_alpha_m * p.cs_m;
  d.art_p = p.artificial_pressure_coefficient_material;
  d.cfl = p.cfl; d.cs = p.cs_m;
  d.band = p.dr * 1.0001f;      // particle_interaction.cu:147
  d.half_dr = p.dr / 2;         // particle_interaction.cu:157
  d.mu = 0.58108f;              // particle_interaction.cu:188-189
  d.periodbound = p.periodbound ? 1 : 0;
  d.surcharge_block = -1;
  d.plane_l = make_float4(p.planeleft[0], p.planeleft[1], p.planeleft[2], p.planeleft[3]);
  d.plane_r = make_float4(p.planeright[0], p.planeright[1], p.planeright[2], p.planeright[3]);
  d.npm = p.npm; d.np_delete_m = p.np_delete_m;
  d.time = p.time; d.dt = p.dt; d.time_d = p.time; d.step = 0;
  int maxkey = p.ncell_delete_m, nb = 0;
  while ((1LL << nb) <= maxkey) nb++;
  d.nbits = nb;
}

// constitutive_model.h:226-230, hoisted per material (float arithmetic).
static HypoConst hypo_const(const float* m) {
  HypoConst h{};
  const float Ei = m[0], Vi = m[1], phic = m[2], pref = m[4];
  if (pref != 0.f && phic != 0.f && Vi != 1.f) {
    h.C1 = -Ei / (3.0f * pref * (1.0f + Vi));
    const float a = 3.0f / (sqrtf(2.0f) * sinf(phic));
    h.C2 = h.C1 * ((3.0f * Vi + a * sqrtf(1.0f + Vi * Vi)) / (1.0f - Vi) - a * a / 3.0f);
    h.C3 = a * a * h.C1;
    h.C4 = a * h.C1;
    h.fa = h.C2 / h.C1;
    h.fb = h.C3 / h.C1;
    h.fc = h.C4 / h.C1;
    h.sq2sin = a;
  }
  return h;
}

static DevMotion motion_to_dev(const lq_rigid_motion& m) {
  DevMotion o{};
  o.is_trans = m.is_trans; o.is_rotat = m.is_rotat; o.pattern = m.pattern;
  o.trans_vel = make_float3(m.trans_vel[0], m.trans_vel[1], m.trans_vel[2]);
  o.trans_acc = make_float3(m.trans_acc[0], m.trans_acc[1], m.trans_acc[2]);
  o.pos_rp = make_float3(m.pos_rp[0], m.pos_rp[1], m.pos_rp[2]);
  o.rotat_vel = make_float3(m.rotat_vel[0], m.rotat_vel[1], m.rotat_vel[2]);
  o.rotat_acc = make_float3(m.rotat_acc[0], m.rotat_acc[1], m.rotat_acc[2]);
  return o;
}

// Cases of a batch must share every parameter the step reads globally.
static bool same_physics(const lq_params& a, const lq_params& b, std::string& why) {
  auto chk = [&](bool ok, const char* f) { if (!ok && why.empty()) why = f; return ok; };
  bool ok = true;
  ok &= chk(a.dim == b.dim, "dim");
  ok &= chk(a.dr == b.dr && a.hdr == b.hdr && a.h == b.h && a.i_h == b.i_h, "dr/hdr");
  ok &= chk(a.vol_def == b.vol_def && a.bound_thickness == b.bound_thickness, "vol_def/bound_thickness");
  ok &= chk(a.gravity[0] == b.gravity[0] && a.gravity[1] == b.gravity[1] && a.gravity[2] == b.gravity[2], "gravity");
  ok &= chk(a.damping == b.damping, "damping");
  ok &= chk(a.kernel_normalization_par == b.kernel_normalization_par &&
                a.kernel_gradient_normalization_par == b.kernel_gradient_normalization_par &&
                a.kernel_zero == b.kernel_zero && a.kernel_dr == b.kernel_dr, "kernel constants");
  ok &= chk(a.stress_shepard_filter == b.stress_shepard_filter, "stress_shepard_filter");
  ok &= chk(a.viscosity_type == b.viscosity_type && a.artificial_viscosity_alpha_m == b.artificial_viscosity_alpha_m,
            "viscosity");
  ok &= chk(a.cs_m == b.cs_m && a.eta == b.eta && a.cfl == b.cfl, "cs/eta/cfl");
  ok &= chk(a.artificial_pressure_coefficient_material == b.artificial_pressure_coefficient_material,
            "artificial pressure");
  ok &= chk(a.time == b.time && a.dt == b.dt, "time/dt");
  ok &= chk(a.key_mode == b.key_mode, "key_mode");
  return ok;
}

static int alloc_arrays(lq_ctx* c, Arrays& A, size_t n) {
  int st = 0;
  st |= dalloc(c, &A.codes, n); st |= dalloc(c, &A.cellpos, n); st |= dalloc(c, &A.orig_id, n);
  st |= dalloc(c, &A.case_id, n);
  st |= dalloc(c, &A.pres, n);
  st |= dalloc(c, &A.pos, n); st |= dalloc(c, &A.pos_ori, n); st |= dalloc(c, &A.vel, n);
  st |= dalloc(c, &A.idm, n); st |= dalloc(c, &A.normal, n);
  st |= dalloc(c, &A.str_a, n); st |= dalloc(c, &A.stn_a, n); st |= dalloc(c, &A.str_b, n); st |= dalloc(c, &A.stn_b, n);
  st |= dalloc(c, &A.cvar, n); st |= dalloc(c, &A.cf, n); st |= dalloc(c, &A.cn, n); st |= dalloc(c, &A.cw, n);
  return st;
}

// Places the cases in one global grid: rows of cell blocks along x with one
// empty cell between neighbours in x and y, all starting at z = 0.  A
// single case keeps its own grid unchanged.
static void layout_cases(lq_ctx* c, int gdim[3]) {
  const int n = c->n_cases;
  if (n == 1) {
    const lq_params& p = c->case_par[0];
    c->cases[0] = DevCase{0, 0, 0, p.cell_num[0], p.cell_num[1], p.cell_num[2], 0, 0};
    for (int d = 0; d < 3; d++) gdim[d] = p.cell_num[d];
    return;
  }
  double area = 0;
  for (const lq_params& p : c->case_par) area += double(p.cell_num[0] + 1) * double(p.cell_num[1] + 1);
  const int row_w = std::max(1, int(std::ceil(std::sqrt(area))));
  int x = 0, y = 0, row_h = 0, gx = 0, gz = 0;
  for (int k = 0; k < n; k++) {
    const lq_params& p = c->case_par[k];
    if (x > 0 && x + p.cell_num[0] > row_w) { y += row_h + 1; x = 0; row_h = 0; }
    DevCase& d = c->cases[k];
    d.ox = x; d.oy = y; d.oz = 0;
    d.nx = p.cell_num[0]; d.ny = p.cell_num[1]; d.nz = p.cell_num[2];
    x += p.cell_num[0] + 1;
    row_h = std::max(row_h, p.cell_num[1]);
    gx = std::max(gx, x - 1);
    gz = std::max(gz, p.cell_num[2]);
  }
  gdim[0] = gx; gdim[1] = y + row_h; gdim[2] = gz;
}

// ---- wall pressure profile without allocation ------------------------
// Scratch of lq_wall_pressure_profile_noalloc: the last-slot markers
// (3 x LQ_PROFILE_NOALLOC_MAX_BINS ints) and the staged output rows
// (LQ_PROFILE_NOALLOC_MAX_BINS float4), one copy per GPU for the life of the
// process.  Host code never uses the array's name as an address: the device
// address of each GPU's copy comes from hipGetSymbolAddress, resolved once
// per device (at context creation, with the device current) and kept in
// g_profile_scratch.
constexpr size_t kProfileMarkBytes = size_t(3) * LQ_PROFILE_NOALLOC_MAX_BINS * sizeof(int);
constexpr size_t kProfileScratchBytes = kProfileMarkBytes + size_t(LQ_PROFILE_NOALLOC_MAX_BINS) * sizeof(float4);
static __device__ __align__(256) unsigned char scratch[kProfileScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_profile_scratch_mu;
void* g_profile_scratch[kMaxDevices] = {};

// Device address of the current device's scratch (device must be current).
hipError_t profile_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_profile_scratch_mu);
  if (!g_profile_scratch[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_profile_scratch[device] = ptr;
  }
  *out = g_profile_scratch[device];
  return hipSuccess;
}
}  // namespace

// The COLUMNS variant needs linear keys (a 2h column's cells are
// consecutive keys), at most col_max_cnz() cells along z and no periodic
// image pass.
static bool cols_supported(const lq_params& p) {
  return p.key_mode == LQ_KEY_LINEAR && !p.periodbound && p.cell_num[2] <= col_max_cnz();
}

static int create_impl(const lq_params* cases, int n_cases, int device, lq_ctx** out) {
  if (!cases || !out || n_cases < 1) return LQ_ERR_ARG;
  *out = nullptr;
  long long np_tot = 0;
  for (int k = 0; k < n_cases; k++) {
    const lq_params& p = cases[k];
    if (p.np <= 0 || p.npb < 0 || p.npm < 0 || p.np != p.npb + p.npm) return LQ_ERR_ARG;  // loquat_io.cu:148-151
    if (p.dim != 3) return LQ_ERR_ARG;
    // periodic planes are one case's absolute coordinates: single-case
    // contexts only, with the interaction variants that run the image pass
    if (p.periodbound && (n_cases > 1 || p.interaction_variant == LQ_INTERACTION_TILE ||
                          p.interaction_variant == LQ_INTERACTION_PAIR_LIST ||
                          p.interaction_variant == LQ_INTERACTION_COLUMNS))
      return LQ_ERR_ARG;
    if (p.interaction_variant < LQ_INTERACTION_REFERENCE_ORDER || p.interaction_variant > LQ_INTERACTION_COLUMNS)
      return LQ_ERR_ARG;
    // Morton keys: single-case contexts, with the variants that look cells
    // up one at a time (TILE copies runs of consecutive linear keys, COLUMNS
    // indexes 2h columns of consecutive linear keys)
    if (p.key_mode != LQ_KEY_LINEAR &&
        (p.key_mode != LQ_KEY_MORTON || n_cases > 1 || p.interaction_variant == LQ_INTERACTION_TILE ||
         p.interaction_variant == LQ_INTERACTION_COLUMNS))
      return LQ_ERR_ARG;
    if (p.n_motion < 0 || p.n_motion > LQ_RIGID_MOTION_MAX) return LQ_ERR_ARG;
    np_tot += p.np;
  }
  if (np_tot >= (1 << 26)) return LQ_ERR_ARG;  // neighbour-list and tile entries hold the slot in 26 bits
  lq_ctx* c = new lq_ctx();
  c->device = device;
  c->n_cases = n_cases;
  c->case_par.assign(cases, cases + n_cases);
  auto bail = [&](int st) { *out = c; return st; };
  for (int k = 1; k < n_cases; k++) {
    std::string why;
    if (!same_physics(cases[0], cases[k], why))
      return bail(fail(c, LQ_ERR_ARG, "lq_create_batch: case " + std::to_string(k) + " differs in " + why));
  }
  // combined parameters: case 0's physics, summed counts, the global grid
  c->cases.assign(n_cases, DevCase{});
  int gdim[3];
  layout_cases(c, gdim);
  lq_params par = cases[0];
  par.np = par.npb = par.npm = par.np_delete_m = 0;
  par.n_motion = 0;
  c->bnd_base.resize(n_cases);
  c->soil_base.resize(n_cases);
  for (int k = 0; k < n_cases; k++) {
    c->bnd_base[k] = par.npb;
    c->soil_base[k] = par.npm;
    c->cases[k].model_base = LQ_MODEL_NUMBER_MAX * k;
    c->cases[k].motion_base = par.n_motion;
    par.np += cases[k].np; par.npb += cases[k].npb; par.npm += cases[k].npm;
    par.n_motion += cases[k].n_motion;
  }
  for (int d = 0; d < 3; d++) par.cell_num[d] = gdim[d];
  if (n_cases > 1) {
    const long long layer = (long long)gdim[0] * gdim[1] * gdim[2];
    if (2 * layer + 1 >= (1LL << 31)) return bail(fail(c, LQ_ERR_ARG, "lq_create_batch: grid too large"));
    par.ncell_material = int(layer);
    par.ncell_delete_m = int(2 * layer);
    par.ncell = int(2 * layer + 1);
  }
  c->par = par;
  c->case_uploaded.assign(n_cases, 0);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { c->err = std::string("hipSetDevice: ") + hipGetErrorString(e); return bail(LQ_ERR_HIP); }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { c->err = std::string("hipStreamCreate: ") + hipGetErrorString(e); return bail(LQ_ERR_HIP); }
  e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);  // default priority, as the main stream
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_adami, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) { c->err = std::string("side stream: ") + hipGetErrorString(e); return bail(LQ_ERR_HIP); }
  {
    void* sp = nullptr;  // resolve this device's profile scratch once, device current
    e = profile_scratch(device, &sp);
    if (e != hipSuccess) { c->err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e); return bail(LQ_ERR_HIP); }
  }
  int mbits = 0;
  if (par.key_mode == LQ_KEY_MORTON) {  // a cube of 2^mbits cells per axis holds the grid
    const int nmax = std::max(par.cell_num[0], std::max(par.cell_num[1], par.cell_num[2]));
    while ((1 << mbits) < nmax) mbits++;
    if (3 * mbits > 29) return bail(fail(c, LQ_ERR_ARG, "lq_create: grid too large for Morton keys"));
    par.ncell_material = 1 << (3 * mbits);
    par.ncell_delete_m = 2 * par.ncell_material;
    par.ncell = par.ncell_delete_m + 1;
    c->par = par;
  }
  fill_static(par, c->hp);
  c->hp.mbits = mbits;
  c->hp.n_cases = n_cases;
  c->hp.n_model = LQ_MODEL_NUMBER_MAX * n_cases;
  c->hp.n_motion = par.n_motion;
  c->nbits = c->hp.nbits;
  c->npm0 = par.npm;
  const size_t n = size_t(par.np);
  int st = 0;
  st |= dalloc(c, &c->dP, 1);
  // material, motion and case tables
  c->n_model_rows = LQ_MODEL_NUMBER_MAX * n_cases;
  c->n_motion_rows = std::max(1, par.n_motion);
  st |= dalloc(c, &c->d_model, size_t(c->n_model_rows));
  st |= dalloc(c, &c->d_hypo, size_t(c->n_model_rows));
  st |= dalloc(c, &c->d_mv, size_t(c->n_motion_rows));
  st |= dalloc(c, &c->d_tab, size_t(c->n_motion_rows));
  st |= dalloc(c, &c->d_cases, size_t(n_cases));
  st |= dalloc(c, &c->d_boxes, size_t(3 * n_cases));
  st |= alloc_arrays(c, c->B.cur, n);
  st |= alloc_arrays(c, c->B.alt, n);
  Buffers& B = c->B;
  st |= dalloc(c, &B.pos_ini, n); st |= dalloc(c, &B.vel_ini, n);
  st |= dalloc(c, &B.str_ini_a, n); st |= dalloc(c, &B.stn_ini_a, n);
  st |= dalloc(c, &B.str_ini_b, n); st |= dalloc(c, &B.stn_ini_b, n);
  st |= dalloc(c, &c->pos_ini_alt, n); st |= dalloc(c, &c->vel_ini_alt, n);
  st |= dalloc(c, &c->str_ini_a_alt, n); st |= dalloc(c, &c->stn_ini_a_alt, n);
  st |= dalloc(c, &c->str_ini_b_alt, n); st |= dalloc(c, &c->stn_ini_b_alt, n);
  st |= dalloc(c, &B.accad, n); st |= dalloc(c, &B.spin, n);
  st |= dalloc(c, &B.acestn_a, n); st |= dalloc(c, &B.acestn_b, n);
  st |= dalloc(c, &B.s3, n); st |= dalloc(c, &B.fdef, n); st |= dalloc(c, &B.deferred, n);
  st |= dalloc(c, &B.str_tmp_a, n); st |= dalloc(c, &B.str_tmp_b, n);
  st |= dalloc(c, &B.rec, 4 * n);
  B.rec_rows = int(n);
  if (cols_supported(par)) {
    st |= dalloc(c, &B.aux, n);
    st |= dalloc(c, &B.soil_aux, n);
    st |= dalloc(c, &B.wall_aux, size_t(std::max(1, par.npb)));
    const size_t ncol = size_t(par.cell_num[0]) * size_t(par.cell_num[1]);
    st |= dalloc(c, &B.sub_start, ncol * size_t(8 * par.cell_num[2] + 1));
    st |= dalloc(c, &B.colcnt, ncol);
    st |= dalloc(c, &B.fl_pos, n);
    st |= dalloc(c, &B.fl_cell, n);
    st |= dalloc(c, &B.colinfo, ncol);
    st |= dalloc(c, &B.coltile, ncol / 1024 + 1);
  }
  if (par.np < (1 << 25)) {  // pair-list entries hold the slot in 25 bits
    st |= dalloc(c, &B.plist, size_t(kPairListCap) * size_t(std::max(1, par.npm)));
    st |= dalloc(c, &B.pcount, size_t(std::max(1, par.npm)));
  }
  st |= dalloc(c, &B.keys, n); st |= dalloc(c, &B.vals, n);
  st |= dalloc(c, &B.keys_alt, n); st |= dalloc(c, &B.vals_alt, n);
  st |= dalloc(c, &B.skey[0], n); st |= dalloc(c, &B.skey[1], n); st |= dalloc(c, &B.mflag, n + 1);
  st |= dalloc(c, &B.mkey, n); st |= dalloc(c, &B.mslot, n);
  st |= dalloc(c, &B.mkey_alt, n); st |= dalloc(c, &B.mslot_alt, n);
  st |= dalloc(c, &B.cell_beg, size_t(par.ncell)); st |= dalloc(c, &B.cell_end, size_t(par.ncell));
  st |= dalloc(c, &B.wall_near, size_t(std::max(1, par.ncell_material)));
  st |= dalloc(c, &B.hist, size_t(256) * radix_tiles(par.np) + size_t(radix_tiles(par.np)) + 1);  // + scan tile sums
  st |= dalloc(c, &B.wf_acc, size_t(2 * n_cases));
  st |= dalloc(c, &c->d_sur_norm, size_t(LQ_MLP_NNORM));
  st |= dalloc(c, &c->d_sur_lo, size_t(LQ_MLP_IN));
  st |= dalloc(c, &c->d_sur_hi, size_t(LQ_MLP_IN));
  st |= dalloc(c, &c->d_sur_stats, size_t(8) * sur_slots());
  st |= dalloc(c, &c->d_sur_amax, size_t(16) * sur_slots());
  if (st) return bail(LQ_ERR_HIP);
  hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  c->n_cu = std::max(1, c->n_cu);
  B.part_cel = B.skey[0];
  B.part_idx = B.vals;
  if (hipHostMalloc((void**)&c->wf_host, sizeof(unsigned long long) * 2 * n_cases) != hipSuccess)
    return bail(fail(c, LQ_ERR_HIP, "hipHostMalloc"));
  std::vector<float> mrows(size_t(c->n_model_rows) * LQ_MATERIAL_SIZE_MAX);
  std::vector<HypoConst> hrows(c->n_model_rows);
  std::vector<DevMotion> mvrows(c->n_motion_rows);
  for (int k = 0; k < n_cases; k++) {
    for (int m = 0; m < LQ_MODEL_NUMBER_MAX; m++) {
      const int r = LQ_MODEL_NUMBER_MAX * k + m;
      for (int j = 0; j < LQ_MATERIAL_SIZE_MAX; j++) mrows[size_t(r) * LQ_MATERIAL_SIZE_MAX + j] = cases[k].model_par[m][j];
      hrows[r] = hypo_const(cases[k].model_par[m]);
    }
    for (int m = 0; m < cases[k].n_motion; m++) mvrows[c->cases[k].motion_base + m] = motion_to_dev(cases[k].mv[m]);
  }
  HIPCHECK(c, hipMemcpyAsync(c->d_model, mrows.data(), mrows.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_hypo, hrows.data(), hrows.size() * sizeof(HypoConst), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_mv, mvrows.data(), mvrows.size() * sizeof(DevMotion), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_cases, c->cases.data(), c->cases.size() * sizeof(DevCase), hipMemcpyHostToDevice,
                             c->stream));
  c->hp.sur_norm = c->d_sur_norm;
  c->hp.sur_lo = c->d_sur_lo;
  c->hp.sur_hi = c->d_sur_hi;
  c->hp.sur_stats = c->d_sur_stats;
  c->hp.sur_audit_max = c->d_sur_amax;
  c->hp.model_par = c->d_model;
  c->hp.hypo = c->d_hypo;
  c->hp.cases = c->d_cases;
  c->hp.mv = c->d_mv;
  c->hp.tab = c->d_tab;
  HIPCHECK(c, hipMemcpyAsync(c->dP, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  launch_fill_int(c->stream, B.cell_beg, par.ncell, -1);
  launch_fill_int(c->stream, B.cell_end, par.ncell, -1);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  *out = c;
  return LQ_OK;
}

extern "C" int lq_create(const lq_params* par, int device, lq_ctx** out) { return create_impl(par, 1, device, out); }
extern "C" int lq_create_batch(const lq_params* cases, int n_cases, int device, lq_ctx** out) {
  return create_impl(cases, n_cases, device, out);
}

extern "C" void lq_destroy(lq_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->graph) hipGraphExecDestroy(c->graph);
  if (c->graph_tmpl) hipGraphDestroy(c->graph_tmpl);
  for (auto& p : c->pending) { c->pool.push_back(p.a); c->pool.push_back(p.b); }
  for (auto e : c->pool) hipEventDestroy(e);
  for (void* p : c->allocs) hipFree(p);
  if (c->wf_host) hipHostFree(c->wf_host);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_adami) hipEventDestroy(c->ev_adami);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* lq_last_error(const lq_ctx* c) { return c ? c->err.c_str() : "null context"; }

// ----------------------------------------------------------- host <-> dev --
template <class T> static int h2d(lq_ctx* c, T* dst, const std::vector<T>& v) {
  HIPCHECK(c, hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return 0;
}
template <class T> static int d2h(lq_ctx* c, std::vector<T>& v, const T* src, size_t n) {
  v.resize(n);
  HIPCHECK(c, hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// CopyDataFromHostToDevice (loquat.cu:223-234) + PositionFromAbsolute2Relative
// for case k: its wall particles go to slots bnd_base[k].., its soil to
// npb + soil_base[k].., in .dat order, so after the first stable sort each
// case's slots are in the order a context of that case alone would hold.
// codes.w is moved to the case's rows of the material (soil) and motion
// (wall) tables.
extern "C" int lq_upload_case(lq_ctx* c, int k, const lq_host_state* h) {
  if (!c || !h || !h->codes || !h->pospres || !h->velrhop || !h->idmassidx || !h->str || !h->stn)
    return fail(c, LQ_ERR_ARG, "lq_upload: null buffer");
  if (k < 0 || k >= c->n_cases) return fail(c, LQ_ERR_ARG, "lq_upload_case: case out of range");
  const lq_params& cp = c->case_par[k];
  if (cp.npb > 0 && !h->normal) return fail(c, LQ_ERR_ARG, "lq_upload: normal required when npb > 0");
  hipSetDevice(c->device);
  const int np = cp.np, npb = cp.npb, npm = cp.npm;
  const DevCase& dc = c->cases[k];
  std::vector<int4> codes(np);
  std::vector<int> oid(np), cid(np, k);
  std::vector<float> pres(np);
  std::vector<float4> pos(np), vel(np), idm(np), nrm(np), sa(np), na(np);
  std::vector<float2> sb(np), nb(np);
  for (int i = 0; i < np; i++) {
    const int* cd = h->codes + 4 * i;
    const int base = i < npb ? dc.motion_base : dc.model_base;
    codes[i] = make_int4(cd[0], cd[1], cd[2], cd[3] + base);
    if (i < npb && cd[1] == LQ_NONFIX && (cd[3] < 0 || cd[3] >= cp.n_motion))
      return fail(c, LQ_ERR_ARG, "lq_upload: wall particle with a motion id outside the case's motions");
    if (i >= npb && (cd[3] < 0 || cd[3] >= LQ_MODEL_NUMBER_MAX))
      return fail(c, LQ_ERR_ARG, "lq_upload: soil particle with a material id outside 0..9");
    pos[i] = make_float4(h->pospres[4 * i], h->pospres[4 * i + 1], h->pospres[4 * i + 2], h->pospres[4 * i + 3]);
    vel[i] = make_float4(h->velrhop[4 * i], h->velrhop[4 * i + 1], h->velrhop[4 * i + 2], h->velrhop[4 * i + 3]);
    // idmassidx.z = line index (loquat_io.cu:240,253)
    idm[i] = make_float4(h->idmassidx[4 * i], h->idmassidx[4 * i + 1], float(i), h->idmassidx[4 * i + 3]);
    const float* sv = h->str + 6 * i;
    const float* tv = h->stn + 6 * i;
    sa[i] = make_float4(sv[0], sv[1], sv[2], sv[3]);
    sb[i] = make_float2(sv[4], sv[5]);
    na[i] = make_float4(tv[0], tv[1], tv[2], tv[3]);
    nb[i] = make_float2(tv[4], tv[5]);
    nrm[i] = i < npb ? make_float4(h->normal[4 * i], h->normal[4 * i + 1], h->normal[4 * i + 2], h->normal[4 * i + 3])
                     : make_float4(0.f, 0.f, 0.f, 0.f);
    oid[i] = i;
    pres[i] = h->pospres[4 * i + 3];
  }
  Arrays& A = c->B.cur;
  const size_t b0 = size_t(c->bnd_base[k]), s0 = size_t(c->par.npb) + size_t(c->soil_base[k]);
  // wall rows [0, npb) to b0.., soil rows [npb, np) to s0..
  auto put = [&](auto* dst, const auto& v) -> int {
    using T = typename std::decay<decltype(v[0])>::type;
    if (npb > 0) HIPCHECK(c, hipMemcpyAsync(dst + b0, v.data(), size_t(npb) * sizeof(T), hipMemcpyHostToDevice, c->stream));
    if (npm > 0)
      HIPCHECK(c, hipMemcpyAsync(dst + s0, v.data() + npb, size_t(npm) * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return 0;
  };
  int st = 0;
  st |= put(A.codes, codes); st |= put(A.pos, pos); st |= put(A.pos_ori, pos);
  st |= put(A.vel, vel); st |= put(A.idm, idm); st |= put(A.normal, nrm);
  st |= put(A.str_a, sa); st |= put(A.str_b, sb); st |= put(A.stn_a, na); st |= put(A.stn_b, nb);
  st |= put(A.orig_id, oid); st |= put(A.pres, pres); st |= put(A.case_id, cid);
  if (st) return st;
  const float scale = 0.5f * cp.i_h;
  launch_abs2rel(c->stream, c->dP, A, scale, int(b0), npb);
  launch_abs2rel(c->stream, c->dP, A, scale, int(s0), npm);
  launch_set_massflag_range(c->stream, c->dP, A, int(b0), npb);
  launch_set_massflag_range(c->stream, c->dP, A, int(s0), npm);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (!c->case_uploaded[k]) { c->case_uploaded[k] = 1; c->n_uploaded++; }
  c->uploaded = c->n_uploaded == c->n_cases;
  c->layout_version++;
  return LQ_OK;
}

extern "C" int lq_upload(lq_ctx* c, const lq_host_state* h) {
  if (!c) return LQ_ERR_ARG;
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_upload: batched context, use lq_upload_case");
  return lq_upload_case(c, 0, h);
}

static void tab_to_dev(const lq_motion_table& a, DevTable& m) {
  std::memset(&m, 0, sizeof(m));
  m.is_translation = a.is_translation; m.table_size_trans = a.table_size_trans;
  m.is_rotation = a.is_rotation; m.pattern = a.pattern; m.table_size_rotat = a.table_size_rotat;
  for (int j = 0; j < LQ_TABLE_MAX; j++) {
    m.trans_vel[j] = make_float3(a.trans_vel[j][0], a.trans_vel[j][1], a.trans_vel[j][2]);
    m.trans_acc[j] = make_float3(a.trans_acc[j][0], a.trans_acc[j][1], a.trans_acc[j][2]);
    m.trans_time[j] = a.trans_time[j];
    m.rotat_acc[j] = make_float3(a.rotat_acc[j][0], a.rotat_acc[j][1], a.rotat_acc[j][2]);
    m.rotat_time[j] = a.rotat_time[j];
  }
  m.pos_rp = make_float3(a.pos_rp[0], a.pos_rp[1], a.pos_rp[2]);
  m.rotat_vel = make_float3(a.rotat_vel[0], a.rotat_vel[1], a.rotat_vel[2]);
}
static void tab_from_dev(const DevTable& m, lq_motion_table& a) {
  std::memset(&a, 0, sizeof(a));
  a.is_translation = m.is_translation; a.table_size_trans = m.table_size_trans;
  a.is_rotation = m.is_rotation; a.pattern = m.pattern; a.table_size_rotat = m.table_size_rotat;
  for (int j = 0; j < LQ_TABLE_MAX; j++) {
    a.trans_vel[j][0] = m.trans_vel[j].x; a.trans_vel[j][1] = m.trans_vel[j].y; a.trans_vel[j][2] = m.trans_vel[j].z;
    a.trans_acc[j][0] = m.trans_acc[j].x; a.trans_acc[j][1] = m.trans_acc[j].y; a.trans_acc[j][2] = m.trans_acc[j].z;
    a.trans_time[j] = m.trans_time[j];
    a.rotat_acc[j][0] = m.rotat_acc[j].x; a.rotat_acc[j][1] = m.rotat_acc[j].y; a.rotat_acc[j][2] = m.rotat_acc[j].z;
    a.rotat_time[j] = m.rotat_time[j];
  }
  a.pos_rp[0] = m.pos_rp.x; a.pos_rp[1] = m.pos_rp.y; a.pos_rp[2] = m.pos_rp.z;
  a.rotat_vel[0] = m.rotat_vel.x; a.rotat_vel[1] = m.rotat_vel.y; a.rotat_vel[2] = m.rotat_vel.z;
}

static int pull_params(lq_ctx* c, DevParams& d) {
  HIPCHECK(c, hipMemcpyAsync(&d, c->dP, sizeof(DevParams), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Motion tables of case k (MotionInfoALL, loquat_io.cu:266-303), rows
// motion_base.. of the device table.
extern "C" int lq_set_case_motion_tables(lq_ctx* c, int k, int n, const lq_motion_table* t) {
  if (!c || k < 0 || k >= c->n_cases || n < 0 || n > LQ_RIGID_MOTION_MAX || (n > 0 && !t))
    return fail(c, LQ_ERR_ARG, "lq_set_motion_tables");
  if (n > c->case_par[k].n_motion && n > 0)
    return fail(c, LQ_ERR_ARG, "lq_set_motion_tables: more tables than the case's motions");
  hipSetDevice(c->device);
  std::vector<DevTable> tabs(std::max(n, 1));
  for (int j = 0; j < n; j++) tab_to_dev(t[j], tabs[j]);
  if (n > 0)
    HIPCHECK(c, hipMemcpyAsync(c->d_tab + c->cases[k].motion_base, tabs.data(), sizeof(DevTable) * n,
                               hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return LQ_OK;
}
extern "C" int lq_set_motion_tables(lq_ctx* c, int n, const lq_motion_table* t) {
  if (!c) return LQ_ERR_ARG;
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_set_motion_tables: batched context, use the case form");
  return lq_set_case_motion_tables(c, 0, n, t);
}

extern "C" int lq_get_motion_tables(lq_ctx* c, int n, lq_motion_table* t) {
  if (!c || n < 0 || n > LQ_RIGID_MOTION_MAX || n > c->n_motion_rows || (n > 0 && !t))
    return fail(c, LQ_ERR_ARG, "lq_get_motion_tables");
  hipSetDevice(c->device);
  std::vector<DevTable> tabs(std::max(n, 1));
  if (n > 0) HIPCHECK(c, hipMemcpyAsync(tabs.data(), c->d_tab, sizeof(DevTable) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < n; k++) tab_from_dev(tabs[k], t[k]);
  return LQ_OK;
}

extern "C" int lq_get_params(lq_ctx* c, lq_params* out) {
  if (!c || !out) return fail(c, LQ_ERR_ARG, "lq_get_params");
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  *out = c->par;
  out->npm = d.npm; out->np_delete_m = d.np_delete_m;
  out->time = d.time; out->dt = d.dt;
  const int nm = std::min(LQ_RIGID_MOTION_MAX, c->n_motion_rows);
  std::vector<DevMotion> mvs(nm);
  HIPCHECK(c, hipMemcpyAsync(mvs.data(), c->d_mv, sizeof(DevMotion) * nm, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < nm; k++) {
    const DevMotion& m = mvs[k];
    lq_rigid_motion& o = out->mv[k];
    o.trans_vel[0] = m.trans_vel.x; o.trans_vel[1] = m.trans_vel.y; o.trans_vel[2] = m.trans_vel.z;
    o.trans_acc[0] = m.trans_acc.x; o.trans_acc[1] = m.trans_acc.y; o.trans_acc[2] = m.trans_acc.z;
    o.pos_rp[0] = m.pos_rp.x; o.pos_rp[1] = m.pos_rp.y; o.pos_rp[2] = m.pos_rp.z;
    o.rotat_vel[0] = m.rotat_vel.x; o.rotat_vel[1] = m.rotat_vel.y; o.rotat_vel[2] = m.rotat_vel.z;
    o.rotat_acc[0] = m.rotat_acc.x; o.rotat_acc[1] = m.rotat_acc.y; o.rotat_acc[2] = m.rotat_acc.z;
  }
  return LQ_OK;
}

extern "C" int lq_get_time(lq_ctx* c, double* time, int64_t* step) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  if (time) *time = d.time_d;
  if (step) *step = d.step;
  return LQ_OK;
}

extern "C" int lq_set_time(lq_ctx* c, double time, int64_t step) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHECK(c, hipMemcpyAsync((char*)c->dP + offsetof(DevParams, time_d), &time, sizeof(double), hipMemcpyHostToDevice, c->stream));
  long long s = step;
  HIPCHECK(c, hipMemcpyAsync((char*)c->dP + offsetof(DevParams, step), &s, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->host_step = step;
  return LQ_OK;
}

// Field access: host AoS <-> device SoA.
struct FieldView {
  int width;      // host values per row
  bool is_int;
  size_t rows;
};

static FieldView field_view(const lq_ctx* c, int f) {
  const size_t np = c->par.np, nc = c->par.ncell;
  switch (f) {
    case LQ_F_CODES: case LQ_F_CELLPOS: return {4, true, np};
    case LQ_F_POS: case LQ_F_POS_ORI: case LQ_F_VEL: case LQ_F_IDM: case LQ_F_POS_INI: case LQ_F_VEL_INI:
    case LQ_F_ACCAD: case LQ_F_SPIN: case LQ_F_CVAR: case LQ_F_NORMAL: return {4, false, np};
    case LQ_F_STR: case LQ_F_STN: case LQ_F_STR_INI: case LQ_F_STN_INI: case LQ_F_ACESTN: return {6, false, np};
    case LQ_F_CMAG: return {9, false, np};
    case LQ_F_PART_CEL: case LQ_F_PART_IDX: case LQ_F_ORIG_ID: return {1, true, np};
    case LQ_F_CELL_BEG: case LQ_F_CELL_END: return {1, true, nc};
    default: return {0, false, 0};
  }
}

static float4* f4_ptr(lq_ctx* c, int f) {
  Arrays& A = c->B.cur;
  Buffers& B = c->B;
  switch (f) {
    case LQ_F_POS: return A.pos; case LQ_F_POS_ORI: return A.pos_ori; case LQ_F_VEL: return A.vel;
    case LQ_F_IDM: return A.idm; case LQ_F_POS_INI: return B.pos_ini; case LQ_F_VEL_INI: return B.vel_ini;
    case LQ_F_ACCAD: return B.accad; case LQ_F_SPIN: return B.spin; case LQ_F_CVAR: return A.cvar;
    case LQ_F_NORMAL: return A.normal;
    default: return nullptr;
  }
}
static void sym_ptrs(lq_ctx* c, int f, float4** a, float2** b) {
  Arrays& A = c->B.cur;
  Buffers& B = c->B;
  switch (f) {
    case LQ_F_STR: *a = A.str_a; *b = A.str_b; break;
    case LQ_F_STN: *a = A.stn_a; *b = A.stn_b; break;
    case LQ_F_STR_INI: *a = B.str_ini_a; *b = B.str_ini_b; break;
    case LQ_F_STN_INI: *a = B.stn_ini_a; *b = B.stn_ini_b; break;
    case LQ_F_ACESTN: *a = B.acestn_a; *b = B.acestn_b; break;
    default: *a = nullptr; *b = nullptr;
  }
}
static int* i1_ptr(lq_ctx* c, int f) {
  switch (f) {
    case LQ_F_PART_CEL: return c->B.part_cel; case LQ_F_PART_IDX: return c->B.part_idx;
    case LQ_F_ORIG_ID: return c->B.cur.orig_id; case LQ_F_CELL_BEG: return c->B.cell_beg;
    case LQ_F_CELL_END: return c->B.cell_end;
    default: return nullptr;
  }
}

extern "C" int lq_get_field(lq_ctx* c, int f, void* dst, size_t bytes) {
  if (!c || !dst) return fail(c, LQ_ERR_ARG, "lq_get_field: null");
  FieldView v = field_view(c, f);
  if (v.width == 0) return fail(c, LQ_ERR_ARG, "lq_get_field: unknown field");
  if (bytes != v.rows * v.width * 4) return fail(c, LQ_ERR_ARG, "lq_get_field: size mismatch");
  hipSetDevice(c->device);
  const size_t n = v.rows;
  int st = 0;
  if (f == LQ_F_CODES || f == LQ_F_CELLPOS) {
    std::vector<int4> t;
    st = d2h(c, t, f == LQ_F_CODES ? c->B.cur.codes : c->B.cur.cellpos, n);
    if (st) return st;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::memcpy(dst, t.data(), n * 16);
  } else if (v.width == 1) {
    HIPCHECK(c, hipMemcpyAsync(dst, i1_ptr(c, f), n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  } else if (v.width == 4) {
    HIPCHECK(c, hipMemcpyAsync(dst, f4_ptr(c, f), n * 16, hipMemcpyDeviceToHost, c->stream));
    if (f == LQ_F_POS || f == LQ_F_POS_INI) {  // w on the device is the massflag; the host sees p
      std::vector<float> pr;
      if ((st = d2h(c, pr, c->B.cur.pres, n))) return st;
      HIPCHECK(c, hipStreamSynchronize(c->stream));
      float* o = (float*)dst;
      for (size_t i = 0; i < n; i++) o[4 * i + 3] = pr[i];
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  } else if (v.width == 6) {
    float4* a; float2* b;
    sym_ptrs(c, f, &a, &b);
    std::vector<float4> ta; std::vector<float2> tb;
    if ((st = d2h(c, ta, a, n)) || (st = d2h(c, tb, b, n))) return st;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    float* o = (float*)dst;
    for (size_t i = 0; i < n; i++) {
      o[6 * i] = ta[i].x; o[6 * i + 1] = ta[i].y; o[6 * i + 2] = ta[i].z; o[6 * i + 3] = ta[i].w;
      o[6 * i + 4] = tb[i].x; o[6 * i + 5] = tb[i].y;
    }
  } else {  // CMAG
    std::vector<float4> f0, f1, f2;
    if ((st = d2h(c, f0, c->B.cur.cf, n)) || (st = d2h(c, f1, c->B.cur.cn, n)) || (st = d2h(c, f2, c->B.cur.cw, n)))
      return st;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    float* o = (float*)dst;
    for (size_t i = 0; i < n; i++) {
      o[9 * i] = f0[i].x; o[9 * i + 1] = f0[i].y; o[9 * i + 2] = f0[i].z;
      o[9 * i + 3] = f1[i].x; o[9 * i + 4] = f1[i].y; o[9 * i + 5] = f1[i].z;
      o[9 * i + 6] = f2[i].x; o[9 * i + 7] = f2[i].y; o[9 * i + 8] = f2[i].z;
    }
  }
  return LQ_OK;
}

extern "C" int lq_set_field(lq_ctx* c, int f, const void* src, size_t bytes) {
  if (!c || !src) return fail(c, LQ_ERR_ARG, "lq_set_field: null");
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_set_field: not available on batched contexts");
  FieldView v = field_view(c, f);
  if (v.width == 0) return fail(c, LQ_ERR_ARG, "lq_set_field: unknown field");
  if (bytes != v.rows * v.width * 4) return fail(c, LQ_ERR_ARG, "lq_set_field: size mismatch");
  hipSetDevice(c->device);
  const size_t n = v.rows;
  c->layout_version++;
  if (f == LQ_F_CODES || f == LQ_F_CELLPOS) {
    HIPCHECK(c, hipMemcpyAsync(f == LQ_F_CODES ? c->B.cur.codes : c->B.cur.cellpos, src, n * 16, hipMemcpyHostToDevice, c->stream));
  } else if (v.width == 1) {
    HIPCHECK(c, hipMemcpyAsync(i1_ptr(c, f), src, n * 4, hipMemcpyHostToDevice, c->stream));
  } else if (v.width == 4) {
    HIPCHECK(c, hipMemcpyAsync(f4_ptr(c, f), src, n * 16, hipMemcpyHostToDevice, c->stream));
    if (f == LQ_F_POS) {
      const float* in = (const float*)src;
      std::vector<float> pr(n);
      for (size_t i = 0; i < n; i++) pr[i] = in[4 * i + 3];
      HIPCHECK(c, hipMemcpyAsync(c->B.cur.pres, pr.data(), n * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
  } else if (v.width == 6) {
    float4* a; float2* b;
    sym_ptrs(c, f, &a, &b);
    const float* s = (const float*)src;
    std::vector<float4> ta(n); std::vector<float2> tb(n);
    for (size_t i = 0; i < n; i++) {
      ta[i] = make_float4(s[6 * i], s[6 * i + 1], s[6 * i + 2], s[6 * i + 3]);
      tb[i] = make_float2(s[6 * i + 4], s[6 * i + 5]);
    }
    HIPCHECK(c, hipMemcpyAsync(a, ta.data(), n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(b, tb.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  } else {
    const float* s = (const float*)src;
    std::vector<float4> f0(n), f1(n), f2(n);
    for (size_t i = 0; i < n; i++) {
      f0[i] = make_float4(s[9 * i], s[9 * i + 1], s[9 * i + 2], 0.f);
      f1[i] = make_float4(s[9 * i + 3], s[9 * i + 4], s[9 * i + 5], 0.f);
      f2[i] = make_float4(s[9 * i + 6], s[9 * i + 7], s[9 * i + 8], 0.f);
    }
    HIPCHECK(c, hipMemcpyAsync(c->B.cur.cf, f0.data(), n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->B.cur.cn, f1.data(), n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->B.cur.cw, f2.data(), n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  if (f == LQ_F_POS || f == LQ_F_IDM || f == LQ_F_CODES) launch_set_massflag(c->stream, c->dP, c->par.np, c->B.cur);
  if (f == LQ_F_POS_INI || f == LQ_F_POS || f == LQ_F_IDM || f == LQ_F_CODES)
    launch_copy_w(c->stream, c->dP, c->par.np, c->B.cur.pos, c->B.pos_ini);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return LQ_OK;
}

// ------------------------------------------------------------- batches ----
static int case_rows(lq_ctx* c, int k, std::vector<int>& rows) {
  std::vector<int> cid;
  int st = d2h(c, cid, c->B.cur.case_id, size_t(c->par.np));
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  rows.clear();
  for (int i = 0; i < c->par.np; i++)
    if (cid[i] == k) rows.push_back(i);
  return 0;
}

// One case's rows of a per-slot field, in slot order (the order a context of
// that case alone holds), with codes.w and cellpos in the case's own frame.
extern "C" int lq_get_case_field(lq_ctx* c, int k, int f, void* dst, size_t bytes) {
  if (!c || !dst) return fail(c, LQ_ERR_ARG, "lq_get_case_field: null");
  if (k < 0 || k >= c->n_cases) return fail(c, LQ_ERR_ARG, "lq_get_case_field: case out of range");
  if (f == LQ_F_PART_CEL || f == LQ_F_PART_IDX || f == LQ_F_CELL_BEG || f == LQ_F_CELL_END)
    return fail(c, LQ_ERR_ARG, "lq_get_case_field: grid arrays are global");
  FieldView v = field_view(c, f);
  if (v.width == 0) return fail(c, LQ_ERR_ARG, "lq_get_case_field: unknown field");
  const lq_params& cp = c->case_par[k];
  if (bytes != size_t(cp.np) * v.width * 4) return fail(c, LQ_ERR_ARG, "lq_get_case_field: size mismatch");
  std::vector<uint32_t> all(size_t(c->par.np) * v.width);
  int st = lq_get_field(c, f, all.data(), all.size() * 4);
  if (st) return st;
  std::vector<int> rows;
  if ((st = case_rows(c, k, rows))) return st;
  if (int(rows.size()) != cp.np) return fail(c, LQ_ERR_STATE, "lq_get_case_field: case slot count changed");
  uint32_t* o = static_cast<uint32_t*>(dst);
  const DevCase& dc = c->cases[k];
  const int lnx = dc.nx, lny = dc.ny, lnz = dc.nz, lncm = lnx * lny * lnz;
  for (size_t r = 0; r < rows.size(); r++) {
    const size_t i = size_t(rows[r]);
    std::memcpy(o + r * v.width, all.data() + i * v.width, size_t(v.width) * 4);
    int32_t* q = reinterpret_cast<int32_t*>(o + r * v.width);
    if (f == LQ_F_CODES) {
      q[3] -= int(i) < c->par.npb ? dc.motion_base : dc.model_base;
    } else if (f == LQ_F_CELLPOS) {
      const int key = q[3];
      q[0] -= dc.ox; q[1] -= dc.oy; q[2] -= dc.oz;
      const int lin = q[0] * lny * lnz + q[1] * lnz + q[2];  // loquat_grid.cu:88
      q[3] = key == c->par.ncell_delete_m ? 2 * lncm : (key >= c->par.ncell_material ? lncm + lin : lin);
    }
  }
  return LQ_OK;
}

// Absolute positions (x y z p) of case k in its own frame, slot order.
extern "C" int lq_case_positions_absolute(lq_ctx* c, int k, float* dst, size_t bytes) {
  if (!c || !dst) return fail(c, LQ_ERR_ARG, "lq_case_positions_absolute: null");
  if (k < 0 || k >= c->n_cases) return fail(c, LQ_ERR_ARG, "lq_case_positions_absolute: case out of range");
  if (bytes != size_t(c->case_par[k].np) * 16) return fail(c, LQ_ERR_ARG, "lq_case_positions_absolute: size");
  std::vector<float> all(size_t(c->par.np) * 4);
  int st = lq_positions_absolute(c, all.data(), all.size() * 4);
  if (st) return st;
  std::vector<int> rows;
  if ((st = case_rows(c, k, rows))) return st;
  for (size_t r = 0; r < rows.size(); r++) std::memcpy(dst + 4 * r, all.data() + 4 * size_t(rows[r]), 16);
  return LQ_OK;
}

// Wall, live soil and deleted particle counts of case k (synchronises).
extern "C" int lq_case_counts(lq_ctx* c, int k, int32_t* npb, int32_t* npm, int32_t* ndel) {
  if (!c || k < 0 || k >= c->n_cases) return fail(c, LQ_ERR_ARG, "lq_case_counts");
  int32_t gb = 0, gm = 0, gd = 0;
  int st = lq_get_counts(c, &gb, &gm, &gd);
  if (st) return st;
  std::vector<int> rows;
  if ((st = case_rows(c, k, rows))) return st;
  int32_t b = 0, m = 0, d = 0;
  for (int i : rows) {
    if (i < gb) b++;
    else if (i < gb + gm) m++;
    else d++;
  }
  if (npb) *npb = b;
  if (npm) *npm = m;
  if (ndel) *ndel = d;
  return LQ_OK;
}

extern "C" int lq_case_count(const lq_ctx* c) { return c ? c->n_cases : 0; }

// ------------------------------------------------------------ step pieces --
#define NEED_UPLOAD(c) \
  do { if (!(c)) return LQ_ERR_ARG; if (!(c)->uploaded) return fail((c), LQ_ERR_STATE, "call lq_upload first"); } while (0)

static int enq_step_calc(lq_ctx* c) { launch_step_calc(c->stream, c->dP); return 0; }
// After the first search the sort merges the few slots whose cell changed
// into the previous order (the sorted keys alternate between two buffers).
// flag_cleared: the step's first launch has cleared the sort flag.
static int enq_ns(lq_ctx* c, bool flag_cleared = false) {
  Section s(c, LQ_K_NEIGHBOR_SEARCH);
  Buffers& B = c->B;
  const int* prev = c->searched ? B.part_cel : nullptr;
  int* out = B.part_cel == B.skey[0] ? B.skey[1] : B.skey[0];
  launch_neighbor_search(c->stream, c->dP, c->par.np, c->nbits, c->par.ncell, B.cur, B, prev, out, !flag_cleared);
  c->searched = true;
  return 0;
}
// SortArraysPrimary in place: only the slots the sort moved are rewritten.
static int enq_sort_primary(lq_ctx* c) {
  Section s(c, LQ_K_SORT_ARRAYS);
  launch_gather_moved(c->stream, c->dP, c->par.np, c->B.cur, c->B.alt, c->B.part_idx);
  return 0;
}
static int enq_sort_all(lq_ctx* c) {
  Section s(c, LQ_K_SORT_ARRAYS);
  Buffers& B = c->B;
  launch_gather(c->stream, c->dP, c->par.np, B.cur, B.alt, B.part_idx, B);
  launch_gather_ini(c->stream, c->dP, c->par.np, B.part_idx, B, c->pos_ini_alt, c->vel_ini_alt, c->str_ini_a_alt,
                    c->str_ini_b_alt, c->stn_ini_a_alt, c->stn_ini_b_alt);
  std::swap(B.cur, B.alt);
  std::swap(B.pos_ini, c->pos_ini_alt); std::swap(B.vel_ini, c->vel_ini_alt);
  std::swap(B.str_ini_a, c->str_ini_a_alt); std::swap(B.str_ini_b, c->str_ini_b_alt);
  std::swap(B.stn_ini_a, c->stn_ini_a_alt); std::swap(B.stn_ini_b, c->stn_ini_b_alt);
  c->layout_version++;
  return 0;
}
static int enq_adami(lq_ctx* c) {
  Section s(c, LQ_K_ADAMI);
  if (c->npm0 > 0) launch_adami(c->stream, c->dP, c->par.npb, c->B.cur, c->B.cell_beg, c->B.cell_end);
  return 0;
}
static int enq_corner(lq_ctx* c) {
  Section s(c, LQ_K_CORNER);
  if (c->par.npb > 0)
    launch_corner(c->stream, c->dP, c->npm0, c->B.cur, c->B.cell_beg, c->B.cell_end, c->B.wall_near);
  return 0;
}
static int enq_shepard(lq_ctx* c, bool due_only = false) {
  Section s(c, LQ_K_SHEPARD);
  launch_shepard(c->stream, c->dP, c->npm0, c->B.cur, c->B.cell_beg, c->B.cell_end, c->B.str_tmp_a, c->B.str_tmp_b,
                 true, due_only);
  return 0;
}
static void enq_interaction_kernel(lq_ctx* c) {
  if (c->par.interaction_variant == LQ_INTERACTION_PAIR_LIST) {
    launch_interaction_plist(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end, c->xcd_remap);
    launch_interaction(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end,
                       LQ_INTERACTION_PAIR_LIST, c->xcd_remap, false);  // guarded cell-scan fallback
  } else if (c->par.interaction_variant == LQ_INTERACTION_COLUMNS) {
    launch_interaction_col(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->xcd_remap, c->col_wide);
  } else if (c->par.interaction_variant == LQ_INTERACTION_TILE) {
    if (c->npm0 > 0)
      launch_interaction_tile(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->par.cell_num[2], c->B.cur,
                              c->B, c->B.cell_beg, c->B.cell_end, c->tile_shape, c->tile_tz);
  } else {
    launch_interaction(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end,
                       c->par.interaction_variant, c->xcd_remap, c->par.periodbound != 0);
  }
}
static void enq_plist_build(lq_ctx* c) {
  launch_plist_build(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end, c->xcd_remap);
}
// stage 0: first interaction after a sort (a pair-list variant builds its
// lists); stage 1: the corrector-stage call of the same step (lists reused
// when still exact).
// beside: Adami and the corner search were enqueued on the side stream
// (enq_walls_beside, COLUMNS only); the walls' pair records wait for Adami,
// and the interaction leaves the wall contact of particles near a wall for
// k_contact_after, which waits for the search.
// defer_commit: pass C is left to the predictor / corrector that follows
// (enq_step; a piecewise call publishes it at once).
static int enq_interaction(lq_ctx* c, int stage, bool beside = false, bool defer_commit = false) {
  const bool build = stage == 0 && c->par.interaction_variant == LQ_INTERACTION_PAIR_LIST;
  launch_reset_deferred(c->stream, c->dP, build ? 1 : 0);
  // COLUMNS writes its records in aux order while it indexes the sub-cells
  if (c->par.interaction_variant != LQ_INTERACTION_COLUMNS)
    launch_pair_prep(c->stream, c->dP, c->par.npb + c->npm0, c->B.cur, c->B);
  if (build) {
    Section s(c, LQ_K_PAIR_LIST_BUILD);
    enq_plist_build(c);
  }
  if (beside) {  // the index and soil records beside Adami, the wall records after it
    {
      Section s(c, LQ_K_COLUMN_INDEX);
      launch_col_index(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->B.cur, c->B, 0);
    }
    hipStreamWaitEvent(c->stream, c->ev_adami, 0);
    {
      Section s(c, LQ_K_COLUMN_INDEX);
      launch_col_wall_records(c->stream, c->dP, c->par.npb, c->B.cur, c->B);
    }
  } else if (c->par.interaction_variant == LQ_INTERACTION_COLUMNS) {
    Section s(c, LQ_K_COLUMN_INDEX);
    launch_col_index(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->B.cur, c->B);
  }
  {
    Section s(c, LQ_K_INTERACTION);
    if (beside) launch_interaction_col(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->xcd_remap, c->col_wide, 3);
    else enq_interaction_kernel(c);
  }
  {
    Section s(c, LQ_K_CONTACT);
    if (beside) {
      hipStreamWaitEvent(c->stream, c->ev_join, 0);
      launch_contact_after(c->stream, c->dP, c->npm0, c->B.cur, c->B);
    }
    launch_contact(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end, false, !defer_commit);
  }
  return 0;
}
// The HypoCS update of soil slots after the predictor (stage 0) or corrector
// (stage 1) kinematics: the surrogate when enabled, and tuple logging.
static void enq_constitutive(lq_ctx* c, int stage) {
  if (c->hp.sur_on) launch_surrogate(c->stream, c->dP, c->npm0, c->sur_members, c->n_cu, c->B.cur, c->B, stage);
  if (c->hp.log_every > 0) launch_sur_log(c->stream, c->dP, c->npm0, c->B.cur, c->B, stage);
}
// commit: publish the last interaction's pass C too (its enq_interaction
// deferred it)
static int enq_predictor(lq_ctx* c, bool commit = false) {
  Section s(c, LQ_K_PREDICTOR);
  launch_predictor(c->stream, c->dP, c->par.np, c->B.cur, c->B,
                   c->par.interaction_variant == LQ_INTERACTION_PAIR_LIST ? 1 : 0, commit ? 1 : 0);
  enq_constitutive(c, 0);
  return 0;
}
static int enq_corrector(lq_ctx* c, bool commit = false) {
  Section s(c, LQ_K_CORRECTOR);
  launch_corrector(c->stream, c->dP, c->par.npb, c->npm0, c->par.n_motion, c->B.cur, c->B, commit ? 1 : 0);
  enq_constitutive(c, 1);
  return 0;
}

static int check_launch(lq_ctx* c) {
  HIPCHECK(c, hipGetLastError());
  return 0;
}

extern "C" int lq_step_calculation(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_step_calc(c); return check_launch(c); }
extern "C" int lq_neighbor_search(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_ns(c); return check_launch(c); }
extern "C" int lq_sort_arrays_primary(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_sort_primary(c); return check_launch(c); }
extern "C" int lq_sort_arrays_all(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_sort_all(c); return check_launch(c); }
extern "C" int lq_rigid_body_update(lq_ctx* c) {
  NEED_UPLOAD(c);
  hipSetDevice(c->device);
  if (c->par.n_motion > 0) launch_rigid_body_update(c->stream, c->dP);
  return check_launch(c);
}
extern "C" int lq_adami_boundary(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_adami(c); return check_launch(c); }
extern "C" int lq_find_corner_particle(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_corner(c); return check_launch(c); }
extern "C" int lq_stress_regularization(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_shepard(c); return check_launch(c); }
extern "C" int lq_particle_interaction(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_interaction(c, 0); return check_launch(c); }
extern "C" int lq_predictor(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_predictor(c); return check_launch(c); }
extern "C" int lq_corrector(lq_ctx* c) { NEED_UPLOAD(c); hipSetDevice(c->device); enq_corrector(c); return check_launch(c); }

// A COLUMNS interaction with the predictor (stage 0) or corrector (stage 1)
// fused in: every soil particle whose contact force needs no pass B is
// advanced in the interaction kernel; pass B (reading first-layer positions
// saved before their update) completes the deferred ones, which the finish
// kernels then advance, with the walls' snapshot and the moving walls.
// Bit-identical to enq_interaction + enq_predictor / enq_corrector.
static bool fused(const lq_ctx* c) {
  return c->fuse && c->par.interaction_variant == LQ_INTERACTION_COLUMNS && c->B.fl_pos;
}
static void enq_fused(lq_ctx* c, int stage) {
  launch_reset_deferred(c->stream, c->dP, 0);
  {
    Section s(c, LQ_K_COLUMN_INDEX);
    launch_col_index(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->B.cur, c->B);
  }
  {
    Section s(c, LQ_K_INTERACTION);
    launch_interaction_col(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->xcd_remap, c->col_wide,
                           stage + 1);
  }
  {
    Section s(c, LQ_K_CONTACT);
    launch_contact(c->stream, c->dP, c->npm0, c->B.cur, c->B, c->B.cell_beg, c->B.cell_end, true);
  }
  {
    Section s(c, stage == 0 ? LQ_K_PREDICTOR : LQ_K_CORRECTOR);
    if (stage == 0) launch_predictor_finish(c->stream, c->dP, c->B.cur, c->B);
    else launch_corrector_finish(c->stream, c->dP, c->par.npb, c->par.n_motion, c->B.cur, c->B);
    enq_constitutive(c, stage);
  }
}

// COLUMNS steps: Adami (cuAdamiBoundaryMaterial) and the corner search
// (cuFindCornerParticleMaterial) on the side stream, so that they run beside
// the column build and the interaction (both are latency-bound on few
// waves); enq_interaction(..., true) waits for Adami before it writes the
// walls' pair records and for the search before the wall contact.  Their
// timed sections sit on the side stream; the contact section's wait on
// ev_join orders them before the main stream's next section ends.
static bool corner_beside(const lq_ctx* c) {
  return c->overlap && c->side && c->par.interaction_variant == LQ_INTERACTION_COLUMNS && !(c->fuse && c->B.fl_pos);
}
static void enq_walls_beside(lq_ctx* c) {
  hipEventRecord(c->ev_fork, c->stream);
  hipStreamWaitEvent(c->side, c->ev_fork, 0);
  {
    Section t(c, LQ_K_ADAMI, c->side);
    if (c->npm0 > 0) launch_adami(c->side, c->dP, c->par.npb, c->B.cur, c->B.cell_beg, c->B.cell_end);
  }
  hipEventRecord(c->ev_adami, c->side);
  {
    Section t(c, LQ_K_CORNER, c->side);
    if (c->par.npb > 0)
      launch_corner(c->side, c->dP, c->npm0, c->B.cur, c->B.cell_beg, c->B.cell_end, c->B.wall_near);
  }
  hipEventRecord(c->ev_join, c->side);
}

// One step (loquat.cu:366-436, 548-560).  The corrector-section neighbour
// search and SortArraysAll are skipped: the predictor never re-bins a
// particle and no key changes between the two searches, so the second
// stable sort is the identity permutation (asserted by the oracle tests).
// StepCalculation, the search's flag clear and RigidBodyUpdate go in one
// launch at the start (the bodies' motion tables are read by nothing before
// Adami, which follows the sort).
static int enq_step(lq_ctx* c) {
  const int shep = c->par.stress_shepard_filter;
  launch_step_begin(c->stream, c->dP, c->par.n_motion);
  enq_ns(c, true);
  enq_sort_primary(c);
  const bool beside = corner_beside(c);
  if (beside) {
    enq_walls_beside(c);
  } else {
    enq_adami(c);
    enq_corner(c);
  }
  if (shep > 0) {  // runs when due at the device's step count, after Adami has read the stresses it filters
    if (beside) hipStreamWaitEvent(c->stream, c->ev_adami, 0);
    enq_shepard(c, true);
  }
  if (fused(c)) {
    enq_fused(c, 0);
  } else {
    enq_interaction(c, 0, beside, true);
    enq_predictor(c, true);
  }
  if (beside) {
    enq_walls_beside(c);
  } else {
    enq_adami(c);
    enq_corner(c);
  }
  if (fused(c)) {
    enq_fused(c, 1);
  } else {
    enq_interaction(c, 1, beside, true);
    enq_corrector(c, true);
  }
  launch_advance(c->stream, c->dP);
  c->host_step++;
  return 0;
}

// Captures two steps (the two sorted-key buffers alternate per search, so
// after two steps every pointer is back at its start).  With
// timing on, the timed sections' events become event-record nodes; they are
// matched to their sections here and re-pointed at fresh events per replay.
static int capture_graph(lq_ctx* c) {
  if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
  if (c->graph_tmpl) { hipGraphDestroy(c->graph_tmpl); c->graph_tmpl = nullptr; }
  c->graph_sections.clear();
  const size_t first = c->pending.size();
  hipGraph_t g;
  HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  enq_step(c);
  enq_step(c);
  c->capturing = false;
  c->host_step -= 2;
  HIPCHECK(c, hipStreamEndCapture(c->stream, &g));
  HIPCHECK(c, hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
  c->graph_version = c->layout_version;
  c->graph_timed = c->timing;
  c->graph_mask = c->timing_mask;
  if (!c->timing) {
    hipGraphDestroy(g);
    return 0;
  }
  c->graph_tmpl = g;
  size_t nn = 0;
  HIPCHECK(c, hipGraphGetNodes(g, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  HIPCHECK(c, hipGraphGetNodes(g, nodes.data(), &nn));
  auto node_of = [&](hipEvent_t e) -> hipGraphNode_t {
    for (auto n : nodes) {
      hipGraphNodeType t;
      hipEvent_t ev;
      if (hipGraphNodeGetType(n, &t) == hipSuccess && t == hipGraphNodeTypeEventRecord &&
          hipGraphEventRecordNodeGetEvent(n, &ev) == hipSuccess && ev == e)
        return n;
    }
    return nullptr;
  };
  for (size_t k = first; k < c->pending.size(); k++) {
    const TimedPair p = c->pending[k];
    const hipGraphNode_t na = node_of(p.a), nb = node_of(p.b);
    if (!na || !nb) return fail(c, LQ_ERR_HIP, "lq_step: timed section missing from the captured graph");
    c->graph_sections.push_back({p.cls, na, nb});
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.resize(first);
  return 0;
}

extern "C" int lq_step(lq_ctx* c, int nsteps, int use_graph) {
  NEED_UPLOAD(c);
  if (nsteps < 0) return fail(c, LQ_ERR_ARG, "lq_step: nsteps < 0");
  hipSetDevice(c->device);
  const bool graph_ok = use_graph != 0;
  int done = 0;
  if (graph_ok && !c->searched && nsteps >= 3) {
    enq_step(c);  // the first search sorts in full; the captured steps then merge
    done = 1;
  }
  if (graph_ok && nsteps - done >= 2) {
    if (c->graph == nullptr || c->graph_version != c->layout_version || c->graph_timed != c->timing ||
        (c->timing && c->graph_mask != c->timing_mask)) {
      const int st = capture_graph(c);
      if (st) return st;
    }
    while (nsteps - done >= 2) {
      if (c->pending.size() > kTimingPending) resolve_timing(c);
      if (c->graph_timed) {
        for (const auto& t : c->graph_sections) {
          const hipEvent_t a = ev_get(c), b = ev_get(c);
          HIPCHECK(c, hipGraphExecEventRecordNodeSetEvent(c->graph, t.a, a));
          HIPCHECK(c, hipGraphExecEventRecordNodeSetEvent(c->graph, t.b, b));
          c->pending.push_back({t.cls, a, b});
        }
      }
      HIPCHECK(c, hipGraphLaunch(c->graph, c->stream));
      c->host_step += 2;
      done += 2;
    }
  }
  for (; done < nsteps; done++) {
    if (c->pending.size() > kTimingPending) resolve_timing(c);
    enq_step(c);
  }
  return check_launch(c);
}

extern "C" int lq_synchronize(lq_ctx* c) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  if (d.err == DERR_DELETE) return fail(c, LQ_ERR_DELETE, "ERROR in deleting particles (npb + npm + np_delete_m != np)");
  if (d.err == DERR_DOMAIN) return fail(c, LQ_ERR_DOMAIN, "a boundary particle left the cell grid");
  return LQ_OK;
}

extern "C" int lq_get_counts(lq_ctx* c, int32_t* npb, int32_t* npm, int32_t* ndel) {
  if (!c) return LQ_ERR_ARG;
  int st = lq_synchronize(c);
  if (st) return st;
  DevParams d;
  st = pull_params(c, d);
  if (st) return st;
  if (npb) *npb = d.npb;
  if (npm) *npm = d.npm;
  if (ndel) *ndel = d.np_delete_m;
  return LQ_OK;
}

// Earth-pressure resultant of every case (GetValuesAtPoint,
// utilities.cu:578-640): boxes[3k..3k+2] = case k's probe x_max, z_min,
// z_max in its own frame; F[k] = 0.001 * (exact sum of contact_mag.xx).
extern "C" int lq_wall_force_cases(lq_ctx* c, const float* boxes, double* F) {
  NEED_UPLOAD(c);
  if (!boxes || !F) return fail(c, LQ_ERR_ARG, "lq_wall_force: null");
  hipSetDevice(c->device);
  const int n = c->n_cases;
  HIPCHECK(c, hipMemcpyAsync(c->d_boxes, boxes, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
  launch_wall_force(c->stream, c->dP, c->par.np, n, c->B.cur, c->d_boxes, c->B.wf_acc);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipMemcpyAsync(c->wf_host, c->B.wf_acc, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < n; k++) {
    const __int128 v = (__int128)((unsigned __int128)c->wf_host[2 * k + 1] << 64 | c->wf_host[2 * k]);
    F[k] = 0.001 * std::ldexp(double(v), -64);
  }
  return LQ_OK;
}

extern "C" int lq_wall_force(lq_ctx* c, const float box[3], double* F) {
  if (!c) return LQ_ERR_ARG;
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_wall_force: batched context, use lq_wall_force_cases");
  return lq_wall_force_cases(c, box, F);
}

// Strip surcharge: soil slots whose block id (codes.x) equals block carry a
// downward force pressure * dr^2 (the load on one lattice column); block -1
// removes it.
extern "C" int lq_set_surcharge(lq_ctx* c, int block, float pressure) {
  if (!c) return LQ_ERR_ARG;
  if (block >= 10) return fail(c, LQ_ERR_ARG, "lq_set_surcharge: block ids >= 10 are contact walls");
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  d.surcharge_block = block < 0 ? -1 : block;
  d.surcharge_force = block < 0 ? 0.f : pressure * c->par.dr * c->par.dr;
  HIPCHECK(c, hipMemcpyAsync(c->dP, &d, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->hp.surcharge_block = d.surcharge_block;
  c->hp.surcharge_force = d.surcharge_force;
  return LQ_OK;
}

// GetValuesAlongLine (utilities.cu:224-367) with the case constants passed
// in: out f32[nbins][4] = (depth, band-1, band-2, band-3 pressure).
extern "C" int lq_wall_pressure_profile(lq_ctx* c, const float geom[4], int nbins, float* out, size_t bytes) {
  NEED_UPLOAD(c);
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_wall_pressure_profile: single-case contexts only");
  if (!geom || !out || nbins <= 0 || bytes != size_t(nbins) * 16)
    return fail(c, LQ_ERR_ARG, "lq_wall_pressure_profile: args");
  hipSetDevice(c->device);
  int* last = nullptr;
  float4* dout = nullptr;
  HIPCHECK(c, hipMalloc(&last, size_t(nbins) * 3 * sizeof(int)));
  hipError_t e = hipMalloc(&dout, size_t(nbins) * sizeof(float4));
  if (e == hipSuccess) {
    launch_wall_profile(c->stream, c->dP, c->par.np, c->par.npb, c->B.cur,
                        make_float4(geom[0], geom[1], geom[2], geom[3]), nbins, last, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(last);
  if (dout) hipFree(dout);
  if (e != hipSuccess) return fail(c, LQ_ERR_HIP, std::string("lq_wall_pressure_profile: ") + hipGetErrorString(e));
  return LQ_OK;
}

// lq_wall_pressure_profile with every intermediate in the file-scope
// scratch: no hipMalloc/hipFree.  The clears are fill kernels on the
// context stream; calls that overlap on one GPU share the scratch and race.
extern "C" int lq_wall_pressure_profile_noalloc(lq_ctx* c, const float geom[4], int nbins, float* out,
                                                size_t bytes) {
  NEED_UPLOAD(c);
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_wall_pressure_profile_noalloc: single-case contexts only");
  if (nbins > LQ_PROFILE_NOALLOC_MAX_BINS)
    return fail(c, LQ_ERR_ARG, "lq_wall_pressure_profile_noalloc: nbins above LQ_PROFILE_NOALLOC_MAX_BINS (" +
                                   std::to_string(LQ_PROFILE_NOALLOC_MAX_BINS) + ")");
  if (!geom || !out || nbins <= 0 || bytes != size_t(nbins) * 16)
    return fail(c, LQ_ERR_ARG, "lq_wall_pressure_profile_noalloc: args");
  hipSetDevice(c->device);
  void* base = nullptr;
  hipError_t e = profile_scratch(c->device, &base);
  if (e != hipSuccess)
    return fail(c, LQ_ERR_HIP, std::string("lq_wall_pressure_profile_noalloc: scratch: ") + hipGetErrorString(e));
  int* last = static_cast<int*>(base);
  float4* dout = reinterpret_cast<float4*>(static_cast<unsigned char*>(base) + kProfileMarkBytes);
  launch_wall_profile(c->stream, c->dP, c->par.np, c->par.npb, c->B.cur,
                      make_float4(geom[0], geom[1], geom[2], geom[3]), nbins, last, dout);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess)
    return fail(c, LQ_ERR_HIP, std::string("lq_wall_pressure_profile_noalloc: ") + hipGetErrorString(e));
  return LQ_OK;
}

// Probes at caller-given points (GetSoilStress, GetSurface,
// GetSurfaceAtPoint: utilities.cu:26-132, 371-528, whose points are
// hard-coded there).  Points and results travel through per-call device
// buffers; single-case contexts.
static int probe_points(lq_ctx* c, const char* name, const void* pts, size_t pts_bytes, void* out, size_t out_bytes,
                        const std::function<void(const void*, void*)>& launch) {
  hipSetDevice(c->device);
  void* dpts = nullptr;
  void* dout = nullptr;
  hipError_t e = hipMalloc(&dpts, pts_bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, out_bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(dpts, pts, pts_bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch(dpts, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (dpts) hipFree(dpts);
  if (dout) hipFree(dout);
  if (e != hipSuccess) return fail(c, LQ_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e));
  return LQ_OK;
}

extern "C" int lq_sensor_stress(lq_ctx* c, const float* points, int n, int reference_rows, float* out,
                                size_t bytes) {
  NEED_UPLOAD(c);
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_sensor_stress: single-case contexts only");
  if (!points || !out || n <= 0 || bytes != size_t(n) * 32) return fail(c, LQ_ERR_ARG, "lq_sensor_stress: args");
  std::vector<float4> p(n);
  for (int k = 0; k < n; k++) p[k] = make_float4(points[3 * k], points[3 * k + 1], points[3 * k + 2], 0.f);
  return probe_points(c, "lq_sensor_stress", p.data(), p.size() * sizeof(float4), out, bytes,
                      [&](const void* dp, void* dout) {
                        launch_sensor_stress(c->stream, c->dP, c->B.cur, c->B.cell_beg, c->B.cell_end,
                                             (const float4*)dp, n, reference_rows ? 1 : 0, (float4*)dout);
                      });
}

extern "C" int lq_surface_heights(lq_ctx* c, const float* points, int n, float radius, float* out, size_t bytes) {
  NEED_UPLOAD(c);
  if (c->n_cases != 1) return fail(c, LQ_ERR_ARG, "lq_surface_heights: single-case contexts only");
  if (!points || !out || n <= 0 || bytes != size_t(n) * 4) return fail(c, LQ_ERR_ARG, "lq_surface_heights: args");
  if (!(radius != 0.f && std::fabs(radius) <= c->hp.two_h))
    return fail(c, LQ_ERR_ARG, "lq_surface_heights: 0 < |radius| <= 2h");
  return probe_points(c, "lq_surface_heights", points, size_t(n) * 8, out, bytes, [&](const void* dp, void* dout) {
    launch_surface_heights(c->stream, c->dP, c->B.cur, c->B.cell_beg, c->B.cell_end, (const float2*)dp, n, radius,
                           (float*)dout);
  });
}

extern "C" int lq_positions_absolute(lq_ctx* c, float* dst, size_t bytes) {
  NEED_UPLOAD(c);
  if (!dst || bytes != size_t(c->par.np) * 16) return fail(c, LQ_ERR_ARG, "lq_positions_absolute: size");
  hipSetDevice(c->device);
  float4* tmp = c->B.fdef;  // scratch of np float4
  launch_rel2abs(c->stream, c->dP, c->par.np, c->B.cur, tmp);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return LQ_OK;
}

extern "C" int lq_stress_saving_regularization(lq_ctx* c, float* dst, size_t bytes) {
  NEED_UPLOAD(c);
  const size_t n = size_t(c->par.np);
  if (!dst || bytes != n * 24) return fail(c, LQ_ERR_ARG, "lq_stress_saving_regularization: size");
  hipSetDevice(c->device);
  Buffers& B = c->B;
  HIPCHECK(c, hipMemcpyAsync(B.str_tmp_a, B.cur.str_a, n * 16, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(B.str_tmp_b, B.cur.str_b, n * 8, hipMemcpyDeviceToDevice, c->stream));
  launch_shepard(c->stream, c->dP, c->npm0, B.cur, B.cell_beg, B.cell_end, B.str_tmp_a, B.str_tmp_b, false);
  HIPCHECK(c, hipGetLastError());
  std::vector<float4> ta;
  std::vector<float2> tb;
  int st = 0;
  if ((st = d2h(c, ta, B.str_tmp_a, n)) || (st = d2h(c, tb, B.str_tmp_b, n))) return st;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; i++) {
    dst[6 * i] = ta[i].x; dst[6 * i + 1] = ta[i].y; dst[6 * i + 2] = ta[i].z; dst[6 * i + 3] = ta[i].w;
    dst[6 * i + 4] = tb[i].x; dst[6 * i + 5] = tb[i].y;
  }
  return LQ_OK;
}

extern "C" int lq_stream_copy(lq_ctx* c, size_t bytes, int reps, double* ms) {
  if (!c || bytes < 16 || reps <= 0 || !ms) return fail(c, LQ_ERR_ARG, "lq_stream_copy: args");
  hipSetDevice(c->device);
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  HIPCHECK(c, hipMalloc(&a, n * 16));
  hipError_t e = hipMalloc(&b, n * 16);
  if (e != hipSuccess) { hipFree(a); return fail(c, LQ_ERR_HIP, "hipMalloc"); }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  launch_stream_copy(c->stream, a, b, n);  // warm-up
  hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps; r++) launch_stream_copy(c->stream, a, b, n);
  hipEventRecord(e1, c->stream);
  e = hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(c, LQ_ERR_HIP, "lq_stream_copy");
  *ms = double(t) / reps;
  return LQ_OK;
}

extern "C" int lq_enable_timing(lq_ctx* c, int enable) {
  if (!c) return LQ_ERR_ARG;
  resolve_timing(c);
  c->timing = enable != 0;
  c->timing_mask = ~0u;
  for (int k = 0; k < LQ_K_CLASS_COUNT; k++) { c->ms[k] = 0; c->launches[k] = 0; }
  return LQ_OK;
}

extern "C" int lq_enable_timing_classes(lq_ctx* c, unsigned mask) {
  if (!c) return LQ_ERR_ARG;
  if (mask >> LQ_K_CLASS_COUNT) return fail(c, LQ_ERR_ARG, "lq_enable_timing_classes: unknown class bit");
  const int st = lq_enable_timing(c, mask != 0);
  c->timing_mask = mask;
  return st;
}

extern "C" int lq_kernel_times(lq_ctx* c, double* ms, int64_t* launches, int n) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  resolve_timing(c);
  for (int k = 0; k < n && k < LQ_K_CLASS_COUNT; k++) {
    if (ms) ms[k] = c->ms[k];
    if (launches) launches[k] = c->launches[k];
  }
  return LQ_OK;
}

extern "C" int lq_set_tile(lq_ctx* c, int shape, int tz) {
  if (!c) return LQ_ERR_ARG;
  if (shape < 0 || shape >= interaction_tile_shapes() || tz < 1 || tz > interaction_tile_maxtz(shape))
    return fail(c, LQ_ERR_ARG, "lq_set_tile: shape/tz out of range");
  c->tile_shape = shape;
  c->tile_tz = tz;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

// ------------------------------------------------------------ surrogate --
// Copies the surrogate block of the host DevParams image to the device.
static int push_sur(lq_ctx* c) {
  const size_t off = offsetof(DevParams, sur_on);
  const size_t end = offsetof(DevParams, log_count) + sizeof(unsigned*);
  HIPCHECK(c, hipMemcpyAsync(reinterpret_cast<char*>(c->dP) + off, reinterpret_cast<const char*>(&c->hp) + off,
                             end - off, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

extern "C" int lq_surrogate_set(lq_ctx* c, int members, const float* params, const float* norm, const float* lo,
                                const float* hi, float ens_tol, float audit_frac) {
  if (!c) return LQ_ERR_ARG;
  if (members < 1 || members > 2 || !params || !norm || !lo || !hi || !(ens_tol >= 0.f) ||
      !(audit_frac >= 0.f && audit_frac <= 1.f))
    return fail(c, LQ_ERR_ARG, "lq_surrogate_set: members 1..2, params/norm/lo/hi, ens_tol >= 0, audit in [0,1]");
  for (int k = 0; k < LQ_MLP_NNORM; k++)
    if (!std::isfinite(norm[k])) return fail(c, LQ_ERR_ARG, "lq_surrogate_set: non-finite normalisation");
  for (int k = LQ_MLP_IN; k < 2 * LQ_MLP_IN; k++)
    if (!(norm[k] > 0.f)) return fail(c, LQ_ERR_ARG, "lq_surrogate_set: input std must be > 0");
  hipSetDevice(c->device);
  Buffers& B = c->B;
  const size_t nm = size_t(std::max(1, c->npm0));
  if (!B.sur_x) {
    int st = 0;
    st |= dalloc(c, &B.sur_x, nm * 16);
    st |= dalloc(c, &B.sur_y, 2 * nm * 8);
    st |= dalloc(c, &B.sur_w, size_t(2) * sur_wpack());
    st |= dalloc(c, &B.sur_b, size_t(2) * sur_bpack());
    if (st) return LQ_ERR_HIP;
  }
  std::vector<unsigned short> w(size_t(members) * sur_wpack());
  std::vector<float> b(size_t(members) * sur_bpack());
  for (int m = 0; m < members; m++)
    sur_pack(params + size_t(m) * LQ_MLP_NPARAM, w.data() + size_t(m) * sur_wpack(), b.data() + size_t(m) * sur_bpack());
  HIPCHECK(c, hipMemcpyAsync(B.sur_w, w.data(), w.size() * 2, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(B.sur_b, b.data(), b.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_sur_norm, norm, LQ_MLP_NNORM * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_sur_lo, lo, LQ_MLP_IN * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d_sur_hi, hi, LQ_MLP_IN * 4, hipMemcpyHostToDevice, c->stream));
  c->sur_members = members;
  c->hp.sur_members = members;
  c->hp.sur_ens_tol = ens_tol;
  c->hp.sur_audit_frac = audit_frac;
  return push_sur(c);
}

extern "C" int lq_surrogate_enable(lq_ctx* c, int on) {
  if (!c) return LQ_ERR_ARG;
  if (on && c->sur_members == 0) return fail(c, LQ_ERR_STATE, "lq_surrogate_enable: call lq_surrogate_set first");
  hipSetDevice(c->device);
  c->hp.sur_on = on ? 1 : 0;
  return push_sur(c);
}

extern "C" int lq_surrogate_stats(lq_ctx* c, int64_t counts[5], float* audit_max, int reset) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  const int ns = sur_slots();
  std::vector<unsigned long long> st(size_t(8) * ns);
  std::vector<unsigned> am(size_t(16) * ns);
  HIPCHECK(c, hipMemcpyAsync(st.data(), c->d_sur_stats, st.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipMemcpyAsync(am.data(), c->d_sur_amax, am.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (counts)
    for (int k = 0; k < 5; k++) {
      counts[k] = 0;
      for (int q = 0; q < ns; q++) counts[k] += int64_t(st[size_t(8) * q + k]);
    }
  unsigned mx = 0;
  for (int q = 0; q < ns; q++) mx = std::max(mx, am[size_t(16) * q]);  // non-negative floats order as bits
  if (audit_max) std::memcpy(audit_max, &mx, 4);
  if (reset) {
    HIPCHECK(c, hipMemsetAsync(c->d_sur_stats, 0, st.size() * 8, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_sur_amax, 0, am.size() * 4, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  return LQ_OK;
}

extern "C" int lq_surrogate_log(lq_ctx* c, int every, int capacity) {
  if (!c || every < 0 || capacity < 0) return fail(c, LQ_ERR_ARG, "lq_surrogate_log: every, capacity >= 0");
  hipSetDevice(c->device);
  if (every > 0 && capacity > c->log_capacity) {
    float *lx = nullptr, *ly = nullptr;
    unsigned* lc = nullptr;
    int st = dalloc(c, &lx, size_t(capacity) * 16) | dalloc(c, &ly, size_t(capacity) * 8) | dalloc(c, &lc, 1);
    if (st) return LQ_ERR_HIP;
    c->hp.log_x = lx;
    c->hp.log_y = ly;
    c->hp.log_count = lc;
    c->log_capacity = capacity;
  }
  if (c->hp.log_count) HIPCHECK(c, hipMemsetAsync(c->hp.log_count, 0, sizeof(unsigned), c->stream));
  c->hp.log_every = every;
  c->hp.log_cap = every > 0 ? capacity : 0;
  return push_sur(c);
}

extern "C" int lq_surrogate_log_read(lq_ctx* c, float* x, float* y, int max_rows, int* rows) {
  if (!c || !rows || max_rows < 0 || (max_rows > 0 && (!x || !y))) return fail(c, LQ_ERR_ARG, "lq_surrogate_log_read");
  hipSetDevice(c->device);
  unsigned cnt = 0;
  if (c->hp.log_count)
    HIPCHECK(c, hipMemcpyAsync(&cnt, c->hp.log_count, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const int n = std::min<int>(int(std::min<unsigned>(cnt, unsigned(c->hp.log_cap))), max_rows);
  std::vector<float> tx(size_t(std::max(n, 1)) * 16), ty(size_t(std::max(n, 1)) * 8);
  if (n > 0) {
    HIPCHECK(c, hipMemcpyAsync(tx.data(), c->hp.log_x, size_t(n) * 64, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(ty.data(), c->hp.log_y, size_t(n) * 32, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  for (int r = 0; r < n; r++) {
    std::memcpy(x + size_t(r) * LQ_MLP_IN, &tx[size_t(r) * 16], LQ_MLP_IN * 4);
    std::memcpy(y + size_t(r) * LQ_MLP_OUT, &ty[size_t(r) * 8], LQ_MLP_OUT * 4);
  }
  *rows = n;
  return LQ_OK;
}

// Selects the interaction kernel: variant and, for REFERENCE_ORDER and
// NEIGHBOR_LIST, whether blocks are renumbered per XCD.
extern "C" int lq_set_interaction(lq_ctx* c, int variant, int xcd_remap) {
  if (!c) return LQ_ERR_ARG;
  if (variant < LQ_INTERACTION_REFERENCE_ORDER || variant > LQ_INTERACTION_COLUMNS)
    return fail(c, LQ_ERR_ARG, "lq_set_interaction: variant out of range");
  if (variant == LQ_INTERACTION_PAIR_LIST && !c->B.plist)
    return fail(c, LQ_ERR_ARG, "lq_set_interaction: pair lists need fewer than 2^25 particles");
  if (variant == LQ_INTERACTION_COLUMNS && !c->B.aux)
    return fail(c, LQ_ERR_ARG, "lq_set_interaction: COLUMNS needs linear keys, no periodic planes and at most " +
                                   std::to_string(col_max_cnz()) + " cells along z");
  if (c->par.key_mode == LQ_KEY_MORTON && variant == LQ_INTERACTION_TILE)
    return fail(c, LQ_ERR_ARG, "lq_set_interaction: TILE needs linear keys");
  if (c->par.periodbound && (variant == LQ_INTERACTION_TILE || variant == LQ_INTERACTION_PAIR_LIST))
    return fail(c, LQ_ERR_ARG, "lq_set_interaction: periodic boundaries need REFERENCE_ORDER or NEIGHBOR_LIST");
  c->par.interaction_variant = variant;
  c->xcd_remap = xcd_remap ? 1 : 0;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

// COLUMNS register budget: 0 at most 96 VGPRs, 1 at most 128.
extern "C" int lq_set_columns_build(lq_ctx* c, int build) {
  if (!c) return LQ_ERR_ARG;
  if (build < 0 || build > 1) return fail(c, LQ_ERR_ARG, "lq_set_columns_build: build 0 or 1");
  c->col_wide = build;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

// COLUMNS steps: the corner search beside the interaction on or off (on by default).
extern "C" int lq_set_overlap(lq_ctx* c, int on) {
  if (!c) return LQ_ERR_ARG;
  c->overlap = on ? 1 : 0;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

// Fused COLUMNS steps on or off (off by default).
extern "C" int lq_set_fusion(lq_ctx* c, int on) {
  if (!c) return LQ_ERR_ARG;
  c->fuse = on ? 1 : 0;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}

// Pair-list capacity and skin: the list radius is (1 + skin) * 2h and the
// lists stay in use while no particle moves more than 0.4 * skin * 2h.
extern "C" int lq_set_pair_list(lq_ctx* c, int cap, float skin) {
  if (!c) return LQ_ERR_ARG;
  if (!c->B.plist) return fail(c, LQ_ERR_ARG, "lq_set_pair_list: pair lists need fewer than 2^25 particles");
  if (cap < 1 || cap > kPairListCap || !(skin >= 0.f && skin <= 0.1f))
    return fail(c, LQ_ERR_ARG, "lq_set_pair_list: cap in [1, 160], skin in [0, 0.1]");
  hipSetDevice(c->device);
  const float rl = c->hp.two_h * (1.0f + skin), dl = c->hp.two_h * (0.4f * skin);
  c->hp.r2list = std::max(rl * rl, c->hp.r2cut);
  c->hp.disp2_lim = dl * dl;
  c->hp.plist_cap = cap;
  const size_t off = offsetof(DevParams, r2list);
  const size_t len = offsetof(DevParams, plist_stride) - off;  // r2list, disp2_lim, plist_cap
  HIPCHECK(c, hipMemcpyAsync(reinterpret_cast<char*>(c->dP) + off, reinterpret_cast<const char*>(&c->hp) + off, len,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->layout_version++;
  return LQ_OK;
}

// State of the most recent pair-list build: overflow flag, the largest soil
// displacement since the build and the displacement the skin allows
// (synchronises).
extern "C" int lq_pair_list_status(lq_ctx* c, int* overflow, float* max_disp, float* disp_limit) {
  if (!c) return LQ_ERR_ARG;
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  float d2;
  std::memcpy(&d2, &d.max_disp2_bits, sizeof d2);
  if (overflow) *overflow = d.plist_bad;
  if (max_disp) *max_disp = std::sqrt(d2);
  if (disp_limit) *disp_limit = std::sqrt(d.disp2_lim);
  return LQ_OK;
}

// Times every interaction variant on the current state -- reference order,
// neighbour lists, and each tile shape over a few depths -- and keeps the
// fastest.  The launches mutate the slip flag and contact forces, so
// cvar/cf are saved to the alternate buffers and restored.
extern "C" int lq_autotune(lq_ctx* c, int reps, int* variant_out, int* shape_out, int* tz_out, int* remap_out,
                           double* ms_out) {
  NEED_UPLOAD(c);
  if (reps <= 0) return fail(c, LQ_ERR_ARG, "lq_autotune: reps");
  hipSetDevice(c->device);
  const size_t n = size_t(c->par.np);
  Arrays& A = c->B.cur;
  Arrays& T = c->B.alt;
  HIPCHECK(c, hipMemcpyAsync(T.cvar, A.cvar, n * 16, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(T.cf, A.cf, n * 16, hipMemcpyDeviceToDevice, c->stream));
  launch_pair_prep(c->stream, c->dP, c->par.npb + c->npm0, A, c->B);
  struct Cand { int variant, shape, tz, remap; };
  std::vector<Cand> cands;
  // COLUMNS sums in another order than the others (which agree bit for bit),
  // so it is a candidate only for a context already set to it, and then the
  // only one
  const bool cols = c->par.interaction_variant == LQ_INTERACTION_COLUMNS;
  for (int remap = 0; remap < 2 && cols; remap++)
    for (int wide = 0; wide < 2; wide++) cands.push_back({LQ_INTERACTION_COLUMNS, wide, 0, remap});
  for (int remap = 0; remap < 2 && !cols; remap++) {
    cands.push_back({LQ_INTERACTION_REFERENCE_ORDER, c->tile_shape, c->tile_tz, remap});
    cands.push_back({LQ_INTERACTION_NEIGHBOR_LIST, c->tile_shape, c->tile_tz, remap});
    if (c->B.plist && !c->par.periodbound) cands.push_back({LQ_INTERACTION_PAIR_LIST, c->tile_shape, c->tile_tz, remap});
  }
  const bool no_tile = c->par.periodbound || cols || c->par.key_mode == LQ_KEY_MORTON;
  for (int sh = 0; sh < (no_tile ? 0 : interaction_tile_shapes()); sh++) {
    const int waves = interaction_tile_waves(sh);
    const int base = std::max(1, (waves * 64 * 9) / (10 * 16));  // ~90% of the lanes at ~16 soil per cell
    for (int tz = std::max(1, base - 2); tz <= std::min(interaction_tile_maxtz(sh), base + 2); tz++)
      cands.push_back({LQ_INTERACTION_TILE, sh, tz, c->xcd_remap});
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  double best = 1e30;
  // for COLUMNS, shape is the register budget (1: up to 128 VGPRs)
  auto apply = [&](const Cand& k) {
    c->par.interaction_variant = k.variant;
    if (k.variant == LQ_INTERACTION_COLUMNS) {
      c->col_wide = k.shape;
    } else {
      c->tile_shape = k.shape;
      c->tile_tz = k.tz;
    }
    c->xcd_remap = k.remap;
  };
  Cand win = {c->par.interaction_variant, cols ? c->col_wide : c->tile_shape, c->tile_tz, c->xcd_remap};
  for (const Cand& k : cands) {
    apply(k);
    auto one_call = [&]() {  // COLUMNS rebuilds its sub-cell index and records per call
      if (k.variant == LQ_INTERACTION_COLUMNS)
        launch_col_index(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->B.cur, c->B);
      enq_interaction_kernel(c);
    };
    auto step_calls = [&]() {  // the two interaction calls of one step
      const bool pl = k.variant == LQ_INTERACTION_PAIR_LIST;
      launch_reset_deferred(c->stream, c->dP, pl ? 1 : 0);
      if (pl) enq_plist_build(c);
      one_call();
      launch_reset_deferred(c->stream, c->dP, 0);
      one_call();
    };
    step_calls();  // warm-up
    hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; r++) step_calls();
    hipEventRecord(e1, c->stream);
    hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return fail(c, LQ_ERR_HIP, std::string("lq_autotune: ") + hipGetErrorString(e));
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    const double per = double(ms) / reps;
    if (per < best) { best = per; win = k; }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIPCHECK(c, hipMemcpyAsync(A.cvar, T.cvar, n * 16, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(A.cf, T.cf, n * 16, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  apply(win);
  c->layout_version++;
  if (variant_out) *variant_out = win.variant;
  if (shape_out) *shape_out = win.shape;
  if (tz_out) *tz_out = win.tz;
  if (remap_out) *remap_out = win.remap;
  if (ms_out) *ms_out = best;
  return LQ_OK;
}

extern "C" int lq_neighbor_lists(lq_ctx* c, int max_nb, int32_t* counts, int32_t* ids, size_t ids_bytes) {
  NEED_UPLOAD(c);
  if (max_nb <= 0 || !counts || !ids) return fail(c, LQ_ERR_ARG, "lq_neighbor_lists: args");
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  const size_t n = size_t(d.npm);
  if (ids_bytes < n * max_nb * 4) return fail(c, LQ_ERR_ARG, "lq_neighbor_lists: ids buffer too small");
  int *dc = nullptr, *di = nullptr;
  HIPCHECK(c, hipMalloc(&dc, std::max<size_t>(n, 1) * 4));
  hipError_t e = hipMalloc(&di, std::max<size_t>(n * max_nb, 1) * 4);
  if (e != hipSuccess) { hipFree(dc); return fail(c, LQ_ERR_HIP, "hipMalloc ids"); }
  if (c->par.interaction_variant == LQ_INTERACTION_COLUMNS) {  // the same windows as the interaction
    launch_col_index(c->stream, c->dP, c->par.cell_num[0], c->par.cell_num[1], c->B.cur, c->B);
    launch_neighbor_list_col(c->stream, c->dP, int(n), c->B.cur, c->B, max_nb, dc, di);
  } else {
    launch_neighbor_list(c->stream, c->dP, int(n), c->B.cur, c->B.cell_beg, c->B.cell_end, max_nb, dc, di);
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(counts, dc, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ids, di, n * max_nb * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(dc);
  hipFree(di);
  if (e != hipSuccess) return fail(c, LQ_ERR_HIP, std::string("lq_neighbor_lists: ") + hipGetErrorString(e));
  return LQ_OK;
}

// ------------------------------------------------------------ checkpoint ----
// A checkpoint holds the whole device state a step reads: every per-slot
// array of the current layout, the predictor snapshots and rates, the cell
// ranges, the imposed-motion tables and the device-updated parameters, raw
// (bit patterns), so that a run resumed from it continues bit for bit.  The
// reference restarts from its .dat/.xml output (loquat_io.cu:315-468),
// which drops the contact state, the displacement record and the tables'
// progress.  The next search after a load runs the full sort (its result
// equals the merge path's).
namespace {
struct CkptHeader {
  char magic[8];
  int32_t version, n_cases, np, npb, npm0, ncell, ncm, n_motion_rows;
  int32_t dyn_npm, dyn_np_delete_m, dyn_err, dyn_sort_changed;
  uint32_t dyn_max_disp2_bits;
  int32_t dyn_plist_bad;
  float dyn_time, dyn_dt;
  double dyn_time_d;
  int64_t dyn_step, host_step;
};
constexpr char kCkptMagic[8] = {'L', 'Q', 'C', 'K', 'P', 'T', 0, 1};
struct CkptArray {
  void* p;
  size_t bytes;
};
}  // namespace

static std::vector<CkptArray> ckpt_arrays(lq_ctx* c) {
  const size_t n = size_t(c->par.np), nc = size_t(std::max(1, c->par.ncell));
  const size_t ncm = size_t(std::max(1, c->par.ncell_material));
  Arrays& A = c->B.cur;
  Buffers& B = c->B;
  std::vector<CkptArray> v = {
      {A.codes, n * 16}, {A.cellpos, n * 16}, {A.orig_id, n * 4}, {A.case_id, n * 4}, {A.pres, n * 4},
      {A.pos, n * 16}, {A.pos_ori, n * 16}, {A.vel, n * 16}, {A.idm, n * 16}, {A.normal, n * 16},
      {A.str_a, n * 16}, {A.stn_a, n * 16}, {A.str_b, n * 8}, {A.stn_b, n * 8},
      {A.cvar, n * 16}, {A.cf, n * 16}, {A.cn, n * 16}, {A.cw, n * 16},
      {B.pos_ini, n * 16}, {B.vel_ini, n * 16}, {B.str_ini_a, n * 16}, {B.stn_ini_a, n * 16},
      {B.str_ini_b, n * 8}, {B.stn_ini_b, n * 8}, {B.accad, n * 16}, {B.spin, n * 16},
      {B.acestn_a, n * 16}, {B.acestn_b, n * 8},
      {B.cell_beg, nc * 4}, {B.cell_end, nc * 4}, {B.wall_near, ncm * 4},
      {c->d_mv, size_t(c->n_motion_rows) * sizeof(DevMotion)},
      {c->d_tab, size_t(c->n_motion_rows) * sizeof(DevTable)}};
  return v;
}

extern "C" int lq_checkpoint_write(lq_ctx* c, const char* path) {
  NEED_UPLOAD(c);
  if (!path) return fail(c, LQ_ERR_ARG, "lq_checkpoint_write: null path");
  hipSetDevice(c->device);
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.n_cases = c->n_cases; h.np = c->par.np; h.npb = c->par.npb; h.npm0 = c->npm0;
  h.ncell = c->par.ncell; h.ncm = c->par.ncell_material; h.n_motion_rows = c->n_motion_rows;
  h.dyn_npm = d.npm; h.dyn_np_delete_m = d.np_delete_m; h.dyn_err = d.err; h.dyn_sort_changed = d.sort_changed;
  h.dyn_max_disp2_bits = d.max_disp2_bits; h.dyn_plist_bad = d.plist_bad;
  h.dyn_time = d.time; h.dyn_dt = d.dt; h.dyn_time_d = d.time_d; h.dyn_step = d.step;
  h.host_step = c->host_step;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, LQ_ERR_ARG, std::string("lq_checkpoint_write: cannot open ") + path);
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  std::vector<char> buf;
  for (const CkptArray& a : ckpt_arrays(c)) {
    if (!ok) break;
    buf.resize(a.bytes);
    hipError_t e = hipMemcpyAsync(buf.data(), a.p, a.bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      std::fclose(f);
      return fail(c, LQ_ERR_HIP, std::string("lq_checkpoint_write: ") + hipGetErrorString(e));
    }
    ok = std::fwrite(buf.data(), 1, a.bytes, f) == a.bytes;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(c, LQ_ERR_ARG, std::string("lq_checkpoint_write: short write to ") + path);
  return LQ_OK;
}

extern "C" int lq_checkpoint_read(lq_ctx* c, const char* path) {
  if (!c) return LQ_ERR_ARG;
  if (!path) return fail(c, LQ_ERR_ARG, "lq_checkpoint_read: null path");
  hipSetDevice(c->device);
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, LQ_ERR_ARG, std::string("lq_checkpoint_read: cannot open ") + path);
  CkptHeader h{};
  bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kCkptMagic, 8) == 0 && h.version == 1;
  if (!ok) {
    std::fclose(f);
    return fail(c, LQ_ERR_ARG, "lq_checkpoint_read: not a checkpoint of this format");
  }
  if (h.n_cases != c->n_cases || h.np != c->par.np || h.npb != c->par.npb || h.npm0 != c->npm0 ||
      h.ncell != c->par.ncell || h.ncm != c->par.ncell_material || h.n_motion_rows != c->n_motion_rows) {
    std::fclose(f);
    return fail(c, LQ_ERR_ARG, "lq_checkpoint_read: the checkpoint was written by a context of another shape");
  }
  std::vector<char> buf;
  for (const CkptArray& a : ckpt_arrays(c)) {
    buf.resize(a.bytes);
    if (std::fread(buf.data(), 1, a.bytes, f) != a.bytes) { ok = false; break; }
    hipError_t e = hipMemcpyAsync(a.p, buf.data(), a.bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      std::fclose(f);
      return fail(c, LQ_ERR_HIP, std::string("lq_checkpoint_read: ") + hipGetErrorString(e));
    }
  }
  std::fclose(f);
  if (!ok) return fail(c, LQ_ERR_ARG, "lq_checkpoint_read: truncated file");
  DevParams d;
  int st = pull_params(c, d);
  if (st) return st;
  d.npm = h.dyn_npm; d.np_delete_m = h.dyn_np_delete_m; d.err = h.dyn_err; d.sort_changed = h.dyn_sort_changed;
  d.max_disp2_bits = h.dyn_max_disp2_bits; d.plist_bad = h.dyn_plist_bad;
  d.time = h.dyn_time; d.dt = h.dyn_dt; d.time_d = h.dyn_time_d; d.step = h.dyn_step;
  HIPCHECK(c, hipMemcpyAsync(c->dP, &d, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->host_step = h.host_step;
  c->searched = false;  // no previous sorted keys: the next search sorts in full
  c->uploaded = true;
  c->case_uploaded.assign(c->n_cases, 1);
  c->n_uploaded = c->n_cases;
  c->layout_version++;  // re-capture graphs
  return LQ_OK;
}
