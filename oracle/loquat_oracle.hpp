// loquat_oracle.hpp -- CPU ORACLE (test infrastructure only).
//
// A serial restatement of the reference's 3D SPH retaining-wall step, written
// from reading its source (citations are file:line in the reference's src/).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may load
// it, and only as the checker; the product never links or calls it.
//
// Templated on the arithmetic type T (float mirrors this project's device
// path operation for operation, double is the FP64 reference).  Fused
// multiply-adds are written explicitly (mad) exactly where the device kernels
// use them; the file is compiled with -ffp-contract=off so the compiler adds
// no others.  The reference's own FP32 build evaluates GetKernel /
// GetGradKernel (functions.h:142-143) and the HypoCS constants
// (constitutive_model.h:226-230) in double through double literals; here they
// are evaluated in T, so FP32 bit parity with the reference there is parity
// unpinned (the FP64 instantiation follows the reference's double path).
// Likewise parity unpinned: the rounding of CUDA's norm3df / powf / expf,
// for which std::sqrt of an explicit FMA chain and std::pow / std::exp stand.
//
// Deterministic semantics chosen where the reference races or leaves state
// stale (the same choices are made by the device path):
//  * contact state (contact_var, contact_mag) and the wall normal travel with
//    their particle through every sort;
//  * wall-contact forces are resolved in two passes: pass A gives every
//    active first-layer particle (d <= dr) its own force, pass B lets active
//    particles with d > dr copy a same-row first-layer force, reading only
//    pass-A values; the min-reduced K_i is written back after pass B;
//  * the rotation branch of RigidBodyUpdate reads a zero translation table
//    when the body has none;
//  * a zero rotation increment leaves the wall normal unchanged;
//  * after a deletion, a particle that moves from a slot past npb + npm into
//    the stress/strain rows the sort gathers keeps its own rows (the
//    reference reads an unrefreshed row of its tensor cache there).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include "retainwall.h"

namespace oracle {

template <class T> static inline T mad(T a, T b, T c) { return std::fma(a, b, c); }

template <class T> struct v3 { T x, y, z; };
template <class T> static inline v3<T> V3(T x, T y, T z) { return {x, y, z}; }
template <class T> static inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> static inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> static inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> static inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> static inline v3<T> cross(v3<T> a, v3<T> b) {  // functions.h:100
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> static inline T norm(v3<T> a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

template <class T> struct sym6 { T xx, yy, zz, xy, xz, yz; };
template <class T> static inline sym6<T> operator+(sym6<T> a, sym6<T> b) {
  return {a.xx + b.xx, a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}
template <class T> static inline sym6<T> operator-(sym6<T> a, sym6<T> b) {
  return {a.xx - b.xx, a.yy - b.yy, a.zz - b.zz, a.xy - b.xy, a.xz - b.xz, a.yz - b.yz};
}
template <class T> static inline sym6<T> operator*(sym6<T> a, T s) {
  return {a.xx * s, a.yy * s, a.zz * s, a.xy * s, a.xz * s, a.yz * s};
}
template <class T> static inline sym6<T> operator/(sym6<T> a, T s) {
  return {a.xx / s, a.yy / s, a.zz / s, a.xy / s, a.xz / s, a.yz / s};
}
template <class T> static inline T trace(sym6<T> a) { return a.xx + a.yy + a.zz; }
// functions.h:95 trace(v1, v2) = double contraction with 2x off-diagonals
template <class T> static inline T ddot(sym6<T> a, sym6<T> b) {
  return a.xx * b.xx + a.yy * b.yy + a.zz * b.zz + T(2) * a.xy * b.xy + T(2) * a.xz * b.xz + T(2) * a.yz * b.yz;
}
// functions.h:97 DotProduct(tensor2d, scalar3)
template <class T> static inline v3<T> matvec(sym6<T> s, v3<T> v) {
  return {s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
          s.xz * v.x + s.yz * v.y + s.zz * v.z};
}
template <class T> struct full9 { T xx, xy, xz, yx, yy, yz, zx, zy, zz; };

// ---------------------------------------------------------------- params --
template <class T> struct Motion {  // RigidMotion, types.h:66-76
  int is_trans, is_rotat, pattern;
  v3<T> trans_vel, trans_acc, pos_rp, rotat_vel, rotat_acc;
};
template <class T> struct MotionTab {  // MotionTable, types.h:81-97
  int is_translation, table_size_trans;
  v3<T> trans_vel[LQ_TABLE_MAX];
  T trans_time[LQ_TABLE_MAX];
  v3<T> trans_acc[LQ_TABLE_MAX];
  int is_rotation, pattern, table_size_rotat;
  v3<T> pos_rp, rotat_vel;
  T rotat_time[LQ_TABLE_MAX];
  v3<T> rotat_acc[LQ_TABLE_MAX];
};

template <class T> struct Par {
  int np, npb, npm, np_delete_m;
  int cnx, cny, cnz, ncell, ncell_material, ncell_delete_m;
  T h, i_h, dr, vol_def, bt, eta;
  T two_h;  // 2.0*h as the reference passes it to operator*(int3, scalar)
  v3<T> g;
  T damping;
  T model_par[LQ_MODEL_NUMBER_MAX][LQ_MATERIAL_SIZE_MAX];
  T knorm, kgrad, kernel_zero, kernel_dr;
  int shepard, viscosity_type;
  T alpha_m, cs, cfl, art_p;
  T time, dt;
  int n_motion;
  Motion<T> mv[LQ_RIGID_MOTION_MAX];
  int periodbound;
  T pl[4], pr[4];  // periodic planes a x + b y + c z + d (planeleft, planeright)
  int sur_block;   // soil block id carrying the strip surcharge, -1 for none
  T sur_force;     // its downward force per slot (pressure * dr^2)
};

// Strip surcharge of the large sloping-backfill workload (no reference
// counterpart), set by or_set_surcharge and read by load_par.
inline int g_sur_block = -1;
inline double g_sur_pressure = 0.0;

// Copy the float parameter block (already initialised by the reference rule
// loquat.cu:130-190, see lq_initialize_parameters) into T; for T=double the
// kernel constants are re-derived in double from h.
template <class T> static void load_par(const lq_params* p, Par<T>& P) {
  P.np = p->np; P.npb = p->npb; P.npm = p->npm; P.np_delete_m = p->np_delete_m;
  P.cnx = p->cell_num[0]; P.cny = p->cell_num[1]; P.cnz = p->cell_num[2];
  P.ncell = p->ncell; P.ncell_material = p->ncell_material; P.ncell_delete_m = p->ncell_delete_m;
  P.dr = p->dr;
  P.bt = p->bound_thickness;
  if (sizeof(T) == sizeof(float)) {
    P.h = p->h; P.i_h = p->i_h; P.vol_def = p->vol_def; P.eta = p->eta;
    P.knorm = p->kernel_normalization_par; P.kgrad = p->kernel_gradient_normalization_par;
    P.kernel_zero = p->kernel_zero; P.kernel_dr = p->kernel_dr;
  } else {
    double h = double(p->dr) * double(p->hdr);
    P.h = T(h); P.i_h = T(1.0 / h); P.eta = T(0.01 * h * h);
    P.vol_def = T(double(p->dr) * double(p->dr) * double(p->dr));
    const double pi = 3.14159265358979323846;
    P.knorm = T(42.0 / (256.0 * pi * h * h * h));            // loquat.cu:163
    P.kgrad = T(-5.0 * 42.0 / (256.0 * pi * h * h * h * h * h));  // loquat.cu:165
    double q = double(p->dr) / h;
    P.kernel_zero = T(double(P.knorm) * 16.0 * 0.5);         // loquat.cu:188
    P.kernel_dr = T(double(P.knorm) * std::pow(2.0 - q, 4) * (q + 0.5));  // loquat.cu:189
  }
  P.two_h = T(2) * P.h;
  P.g = V3<T>(p->gravity[0], p->gravity[1], p->gravity[2]);
  P.damping = p->damping;
  for (int k = 0; k < LQ_MODEL_NUMBER_MAX; k++)
    for (int j = 0; j < LQ_MATERIAL_SIZE_MAX; j++) P.model_par[k][j] = p->model_par[k][j];
  P.shepard = p->stress_shepard_filter;
  P.viscosity_type = p->viscosity_type;
  P.alpha_m = p->artificial_viscosity_alpha_m;
  P.cs = p->cs_m; P.cfl = p->cfl;
  P.art_p = p->artificial_pressure_coefficient_material;
  P.time = p->time; P.dt = p->dt;
  P.n_motion = p->n_motion;
  P.periodbound = p->periodbound;
  P.sur_block = g_sur_block;
  P.sur_force = g_sur_block < 0 ? T(0) : T(float(g_sur_pressure) * p->dr * p->dr);
  for (int k = 0; k < 4; k++) { P.pl[k] = p->planeleft[k]; P.pr[k] = p->planeright[k]; }
  for (int k = 0; k < LQ_RIGID_MOTION_MAX; k++) {
    const lq_rigid_motion& m = p->mv[k];
    Motion<T>& o = P.mv[k];
    o.is_trans = m.is_trans; o.is_rotat = m.is_rotat; o.pattern = m.pattern;
    o.trans_vel = V3<T>(m.trans_vel[0], m.trans_vel[1], m.trans_vel[2]);
    o.trans_acc = V3<T>(m.trans_acc[0], m.trans_acc[1], m.trans_acc[2]);
    o.pos_rp = V3<T>(m.pos_rp[0], m.pos_rp[1], m.pos_rp[2]);
    o.rotat_vel = V3<T>(m.rotat_vel[0], m.rotat_vel[1], m.rotat_vel[2]);
    o.rotat_acc = V3<T>(m.rotat_acc[0], m.rotat_acc[1], m.rotat_acc[2]);
  }
}

// ----------------------------------------------------------------- state --
// Per-slot arrays in the host AoS layouts of retainwall.h (caller-owned).
template <class T> struct State {
  T *pos, *pos_ori, *vel, *idm, *pos_ini, *vel_ini, *accad, *spin, *cvar, *normal;  // [np][4]
  T *str, *stn, *str_ini, *stn_ini, *acestn;                                        // [np][6]
  T* cmag;                                                                           // [np][9]
  int32_t *codes, *cellpos;                                                          // [np][4]
  int32_t *part_cel, *part_idx, *orig_id;                                            // [np]
  int32_t *cell_beg, *cell_end;                                                      // [ncell]
  int32_t* deferred;                                                                 // [np]
};

template <class T> static inline v3<T> ld3(const T* a, int i) { return {a[4 * i], a[4 * i + 1], a[4 * i + 2]}; }
template <class T> static inline void st3(T* a, int i, v3<T> v) { a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; }
template <class T> static inline sym6<T> ld6(const T* a, int i) {
  const T* p = a + 6 * i; return {p[0], p[1], p[2], p[3], p[4], p[5]};
}
template <class T> static inline void st6(T* a, int i, sym6<T> s) {
  T* p = a + 6 * i; p[0] = s.xx; p[1] = s.yy; p[2] = s.zz; p[3] = s.xy; p[4] = s.xz; p[5] = s.yz;
}
static inline v3<int> cell_of(const int32_t* cp, int i) { return {cp[4 * i], cp[4 * i + 1], cp[4 * i + 2]}; }

// functions.h:142-143, Wendland C2.  (2-q)^4 (q+1/2) and (2-q)^3.
template <class T> static inline T kernelW(T knorm, T q) { T t = T(2) - q; return knorm * t * t * t * t * (q + T(0.5)); }
template <class T> static inline T kernelG(T kgrad, T q) { T t = T(2) - q; return kgrad * t * t * t; }

// Candidate pair geometry: dr = pos_i - pos_j - off*2h (particle_interaction.cu:31),
// r = |dr| with r^2 accumulated as fma(dz,dz,fma(dy,dy,dx*dx)).
template <class T> static inline T pair_r(v3<T> pi, v3<T> pj, v3<int> off, T two_h, v3<T>& d) {
  d.x = (pi.x - pj.x) - T(off.x) * two_h;
  d.y = (pi.y - pj.y) - T(off.y) * two_h;
  d.z = (pi.z - pj.z) - T(off.z) * two_h;
  return std::sqrt(mad(d.z, d.z, mad(d.y, d.y, d.x * d.x)));
}

template <class T> static inline int linear_key(const Par<T>& P, v3<int> c) {
  return c.x * P.cny * P.cnz + c.y * P.cnz + c.z;  // loquat_grid.cu:88
}
template <class T> static inline bool in_grid(const Par<T>& P, v3<int> c) {
  return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < P.cnx && c.y < P.cny && c.z < P.cnz;
}

// Gravity plus, on slots of the surcharge block, the strip load over the
// slot's mass.
template <class T> static inline v3<T> body_acc(const Par<T>& P, const int32_t* codes, int i, T mi) {
  v3<T> g = P.g;
  if (P.sur_block >= 0 && codes[4 * i] == P.sur_block) g.z = g.z - P.sur_force / mi;
  return g;
}

// ------------------------------------------------ periodic boundaries ----
// Dist2Plane (functions.h:134-136): signed distance of a point to the plane
// a x + b y + c z + d = 0.
template <class T> static inline T plane_dist(v3<T> x, const T* pl) {
  return (x.x * pl[0] + x.y * pl[1] + x.z * pl[2] + pl[3]) / std::sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
}
template <class T> static inline v3<T> abs_pos(const Par<T>& P, v3<int> c, v3<T> p) {
  return V3(p.x + T(c.x) * P.two_h, p.y + T(c.y) * P.two_h, p.z + T(c.z) * P.two_h);
}
// Cell and relative position of an absolute position: int3(x 0.5/h)
// truncated, x - cell 2h.
template <class T> static inline void to_cell(const Par<T>& P, v3<T> a, v3<int>& c, v3<T>& p) {
  const T scale = T(0.5) * P.i_h;
  c = {int(a.x * scale), int(a.y * scale), int(a.z * scale)};
  p = V3(a.x - T(c.x) * P.two_h, a.y - T(c.y) * P.two_h, a.z - T(c.z) * P.two_h);
}
// The periodic image used by the *LoopPeriod helpers (adami_boundary.cu:
// 73-87, field_smoothing.cu:58-72, particle_interaction.cu:243-258): a
// particle closer than 2h to either plane is shifted by the period along
// that plane's normal (left takes precedence) and its 27-cell loop is run
// again from there.  The interaction's shift length is d_l + d_r as
// written; Adami and Shepard take |d_l + d_r|.
template <class T> static bool period_image(const Par<T>& P, v3<int> c, v3<T> p, bool abs_len, v3<int>& cpb,
                                            v3<T>& ppb) {
  const v3<T> a = abs_pos(P, c, p);
  const bool left = std::fabs(plane_dist(a, P.pl)) < P.two_h;
  const bool right = std::fabs(plane_dist(a, P.pr)) < P.two_h;
  if (!left && !right) return false;
  T len = P.pl[3] + P.pr[3];
  if (abs_len) len = std::fabs(len);
  const T* n = left ? P.pl : P.pr;
  to_cell(P, V3(a.x + n[0] * len, a.y + n[1] * len, a.z + n[2] * len), cpb, ppb);
  return true;
}
// cuCheckOutDomainPeriod (predictor_corrector.cu:26-40): a particle behind
// either plane re-enters through the other one.
template <class T> static void wrap_period(const Par<T>& P, v3<int>& c, v3<T>& p) {
  const v3<T> a = abs_pos(P, c, p);
  const bool left = plane_dist(a, P.pl) < T(0);
  const bool right = plane_dist(a, P.pr) < T(0);
  if (!left && !right) return;
  const T len = std::fabs(P.pl[3] + P.pr[3]);
  const T* n = left ? P.pl : P.pr;
  to_cell(P, V3(a.x + n[0] * len, a.y + n[1] * len, a.z + n[2] * len), c, p);
}

// ------------------------------------------------------------- a2: frames --
// cuPositionFromAbsolute2Relative, loquat_grid.cu:253-261.  The cell index is
// computed in float for every T (cell structure is discrete and shared).
template <class T> static void abs2rel(const Par<T>& P, const lq_params* pf, State<T>& S) {
  const float scale = float(0.5 * double(pf->i_h));
  const float two_hf = 2.0f * pf->h;
  for (int i = 0; i < P.np; i++) {
    float xa = float(S.pos[4 * i]), ya = float(S.pos[4 * i + 1]), za = float(S.pos[4 * i + 2]);
    int cx = int(xa * scale), cy = int(ya * scale), cz = int(za * scale);
    S.cellpos[4 * i] = cx; S.cellpos[4 * i + 1] = cy; S.cellpos[4 * i + 2] = cz; S.cellpos[4 * i + 3] = 0;
    if (sizeof(T) == sizeof(float)) {
      S.pos[4 * i] = T(xa - float(cx) * two_hf);
      S.pos[4 * i + 1] = T(ya - float(cy) * two_hf);
      S.pos[4 * i + 2] = T(za - float(cz) * two_hf);
    } else {
      S.pos[4 * i] = S.pos[4 * i] - T(cx) * P.two_h;
      S.pos[4 * i + 1] = S.pos[4 * i + 1] - T(cy) * P.two_h;
      S.pos[4 * i + 2] = S.pos[4 * i + 2] - T(cz) * P.two_h;
    }
  }
}

// --------------------------------------------- a3-a5: neighbour search ---
// cuGetParticleCellIndex (loquat_grid.cu:71-95), stable SortPairs
// (loquat_grid.cu:98-113), cuFindBegEndInCell (116-144) and the deletion
// bookkeeping of NeighborSearch (147-191).  Returns 0, or LQ_ERR_DELETE.
// *identity is set when the sort permutation is the identity.
template <class T> static int neighbor_search(Par<T>& P, State<T>& S, int* identity) {
  const int np = P.np;
  std::vector<int> key(np);
  for (int i = 0; i < np; i++) {
    if (S.cellpos[4 * i + 3] == P.ncell_delete_m) { key[i] = P.ncell_delete_m; continue; }
    int k = linear_key(P, cell_of(S.cellpos, i));
    if (S.idm[4 * i] >= T(100)) k += P.ncell_material;  // MATERIAL_NUMBER, types.h:37
    key[i] = k;
    S.cellpos[4 * i + 3] = k;
  }
  std::vector<int> idx(np);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  int ident = 1;
  for (int i = 0; i < np; i++) {
    S.part_idx[i] = idx[i];
    S.part_cel[i] = key[idx[i]];
    if (idx[i] != i) ident = 0;
  }
  if (identity) *identity = ident;
  for (int c = 0; c < P.ncell; c++) { S.cell_beg[c] = -1; S.cell_end[c] = -1; }
  for (int i = 0; i < np; i++) {
    int c = S.part_cel[i];
    if (i == 0 || c != S.part_cel[i - 1]) {
      S.cell_beg[c] = i;
      if (i > 0) S.cell_end[S.part_cel[i - 1]] = i;
    }
    if (i == np - 1) S.cell_end[c] = i + 1;
  }
  int beg = S.cell_beg[P.ncell_delete_m];
  if (beg != -1) {
    int now = S.cell_end[P.ncell_delete_m] - beg - P.np_delete_m;
    P.npm -= now;
    P.np_delete_m += now;
  }
  if (P.npb + P.npm + P.np_delete_m != P.np) return LQ_ERR_DELETE;
  return 0;
}

// a[i] = a[idx[i]] for the first n rows.  src_rows (>= n, default n) rows are
// readable as sources: the stress and strain gathers cover npb + npm rows,
// and after a deletion the particle that held the last slot can move into
// one of them.  The reference's cuSortArrayTensor2d (loquat_grid.cu:232-251)
// then reads a row of its tensor cache that the memcpy of npb + npm rows did
// not refresh (stale contents of an earlier sort); here the particle keeps
// its own row.
template <class T, class E>
static void gather(std::vector<E>& tmp, E* a, int width, int n, const int32_t* idx, int src_rows = -1) {
  tmp.assign(a, a + size_t(width) * size_t(src_rows < n ? n : src_rows));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int k = 0; k < width; k++) a[size_t(width) * i + k] = tmp[size_t(width) * idx[i] + k];
}

// SortArraysPrimary, loquat.cu:254-266 (+ contact state, orig ids, normal).
template <class T> static void sort_arrays_primary(const Par<T>& P, State<T>& S) {
  std::vector<T> t; std::vector<int32_t> ti;
  const int np = P.np, nm = P.npb + P.npm;
  gather<T>(ti, S.codes, 4, np, S.part_idx);
  gather<T>(ti, S.cellpos, 4, np, S.part_idx);
  gather<T>(ti, S.orig_id, 1, np, S.part_idx);
  gather<T>(t, S.pos, 4, np, S.part_idx);
  gather<T>(t, S.pos_ori, 4, np, S.part_idx);
  gather<T>(t, S.vel, 4, np, S.part_idx);
  gather<T>(t, S.idm, 4, np, S.part_idx);
  gather<T>(t, S.normal, 4, P.npb, S.part_idx);
  gather<T>(t, S.str, 6, nm, S.part_idx, np);
  gather<T>(t, S.stn, 6, nm, S.part_idx, np);
  gather<T>(t, S.cvar, 4, np, S.part_idx);
  gather<T>(t, S.cmag, 9, np, S.part_idx);
}

// SortArraysAll, loquat.cu:269-287 (+ contact state, orig ids, normal).
template <class T> static void sort_arrays_all(const Par<T>& P, State<T>& S) {
  std::vector<T> t; std::vector<int32_t> ti;
  const int np = P.np, nm = P.npb + P.npm;
  gather<T>(ti, S.codes, 4, np, S.part_idx);
  gather<T>(ti, S.cellpos, 4, np, S.part_idx);
  gather<T>(ti, S.orig_id, 1, np, S.part_idx);
  gather<T>(t, S.pos, 4, np, S.part_idx);
  gather<T>(t, S.pos_ori, 4, np, S.part_idx);
  gather<T>(t, S.vel, 4, np, S.part_idx);
  gather<T>(t, S.idm, 4, np, S.part_idx);
  gather<T>(t, S.pos_ini, 4, np, S.part_idx);
  gather<T>(t, S.vel_ini, 4, np, S.part_idx);
  gather<T>(t, S.normal, 4, P.npb, S.part_idx);
  if (P.npm != 0) {
    gather<T>(t, S.str, 6, nm, S.part_idx, np);
    gather<T>(t, S.stn, 6, nm, S.part_idx, np);
    gather<T>(t, S.str_ini, 6, nm, S.part_idx, np);
    gather<T>(t, S.stn_ini, 6, nm, S.part_idx, np);
  }
  gather<T>(t, S.cvar, 4, np, S.part_idx);
  gather<T>(t, S.cmag, 9, np, S.part_idx);
}

// ------------------------------------------------- a8: rigid-body tables --
// LoquatSph::RigidBodyUpdate, adami_boundary.cu:175-203.  time/dt are the
// current step's par.time / par.dt.
template <class T> static void rigid_body_update(Par<T>& P, MotionTab<T>* M) {
  for (int i = 0; i < P.n_motion; i++) {
    MotionTab<T>& m = M[i];
    if (m.is_rotation) {
      for (int j = 0; j < m.table_size_rotat; j++) {
        if (P.time < m.rotat_time[j]) {
          m.rotat_vel = m.rotat_vel + m.rotat_acc[j] * P.dt;
          v3<T> vel = m.trans_vel[j] + m.trans_acc[j] * (T(0.5) * P.dt);
          m.pos_rp = m.pos_rp + vel * P.dt;
          P.mv[i].rotat_acc = m.rotat_acc[j];
          P.mv[i].rotat_vel = m.rotat_vel;
          P.mv[i].pos_rp = m.pos_rp;
          break;
        }
      }
    }
    if (m.is_translation) {
      for (int j = 0; j < m.table_size_trans; j++) {
        if (P.time < m.trans_time[j]) {
          m.trans_vel[j] = m.trans_vel[j] + m.trans_acc[j] * P.dt;
          P.mv[i].trans_acc = m.trans_acc[j];
          P.mv[i].trans_vel = m.trans_vel[j];
          break;
        }
      }
    }
  }
}

// ------------------------------------------------------------ a7: Adami --
// cuAdamiBoundaryMaterial + loops, adami_boundary.cu:26-158.
template <class T> static void adami(const Par<T>& P, State<T>& S) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < P.npb; i++) {
    v3<T> pi = ld3(S.pos, i);
    v3<int> ci = cell_of(S.cellpos, i);
    const bool contact = S.codes[4 * i] >= 10;
    v3<T> n = {0, 0, 0};
    sym6<T> rs = {0, 0, 0, 0, 0, 0};
    v3<T> rpos = {0, 0, 0};
    T rker = 0;
    T rv[4] = {0, 0, 0, 0};
    auto scan = [&](v3<int> ci, v3<T> pi) {
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> off = {cx, cy, cz};
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int cb = linear_key(P, cj);
          for (int pass = 0; pass < 2; pass++) {
            const bool bound = pass == 0;
            int c = bound ? cb : cb + P.ncell_material;
            int b = S.cell_beg[c], e = S.cell_end[c];
            for (int j = b; j < e; j++) {
              v3<T> d;
              T r = pair_r(pi, ld3(S.pos, j), off, P.two_h, d);
              T q = r * P.i_h;
              if (!(q < T(2))) continue;
              T ker = kernelW(P.knorm, q);
              T dwdr = kernelG(P.kgrad, q);
              v3<T> g = d * dwdr;
              if (bound && contact) n = n - g * P.vol_def;
              if (!bound) {
                T volj = S.idm[4 * j + 1] / S.vel[4 * j + 3];
                T wv = ker * volj;
                rs = rs + ld6(S.str, j) * wv;
                rpos = rpos + d * (ker * S.idm[4 * j + 1]);
                rker = rker + wv;
                for (int k = 0; k < 4; k++) rv[k] = rv[k] + S.vel[4 * j + k] * wv;
              }
            }
          }
        }
    };
    scan(ci, pi);
    v3<int> cpb;
    v3<T> ppb;
    if (P.periodbound && period_image(P, ci, pi, true, cpb, ppb)) scan(cpb, ppb);
    T nn = norm(n);
    if (nn > T(0)) {
      S.cmag[9 * i + 3] = n.x / nn; S.cmag[9 * i + 4] = n.y / nn; S.cmag[9 * i + 5] = n.z / nn;
    }
    if (rker > T(0)) {
      T fc = S.normal[4 * i + 3];
      v3<T> ni = ld3(S.normal, i);
      v3<T> vref = V3(rv[0], rv[1], rv[2]) / rker;
      v3<T> vref_n = ni * dot(ni, vref);
      v3<T> vref_t = vref - vref_n;
      v3<T> vrb = {0, 0, 0};
      if (S.codes[4 * i + 1] == LQ_NONFIX) {
        const Motion<T>& m = P.mv[S.codes[4 * i + 3]];
        vrb = m.trans_vel;
        if (m.is_rotat) {
          v3<T> pa = V3(pi.x + T(ci.x) * P.two_h, pi.y + T(ci.y) * P.two_h, pi.z + T(ci.z) * P.two_h);
          vrb = vrb + cross(m.rotat_vel, pa - m.pos_rp);
        }
      }
      v3<T> vrb_n = ni * dot(ni, vrb);
      v3<T> vrb_t = vrb - vrb_n;
      v3<T> vext_n = vrb_n * T(2) - vref_n;
      v3<T> vext_t = fc == T(-1) ? vrb_t * T(2) - vref_t : vref_t;
      v3<T> vext = vext_n + vext_t;
      T rho = rv[3] / rker;
      if (contact) {
        st3(S.vel, i, vrb);
        S.cmag[9 * i + 6] = vext.x; S.cmag[9 * i + 7] = vext.y; S.cmag[9 * i + 8] = vext.z;
      } else {
        st3(S.vel, i, vext);
      }
      S.vel[4 * i + 3] = rho;
      sym6<T> sref = rs / rker;
      v3<T> gr = V3(P.g.x * rpos.x, P.g.y * rpos.y, P.g.z * rpos.z) / rker;
      sref.xx = sref.xx - gr.x; sref.yy = sref.yy - gr.y; sref.zz = sref.zz - gr.z;
      st6(S.str, i, sref);
    } else {
      st6(S.str, i, sym6<T>{0, 0, 0, 0, 0, 0});
    }
  }
}

// ----------------------------------------------------------- a9: corner --
// cuFindCornerParticleMaterial + loop, adami_boundary.cu:209-292.
template <class T> static void find_corner(const Par<T>& P, State<T>& S) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    v3<T> pi = ld3(S.pos, i);
    v3<int> ci = cell_of(S.cellpos, i);
    T* cv = S.cvar + 4 * i;
    T* cm = S.cmag + 9 * i;
    cv[0] = T(0);
    cv[1] = T(999);
    T dij = T(999);
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> off = {cx, cy, cz};
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int c = linear_key(P, cj);
          for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
            v3<T> d;
            T r = pair_r(pi, ld3(S.pos, j), off, P.two_h, d);
            T q = r * P.i_h;
            if (!(q < T(2))) continue;
            T dwdr = kernelG(P.kgrad, q);
            v3<T> g = d * dwdr;
            if (S.codes[4 * j] < 10) continue;
            if (r < dij && r > T(0)) {
              v3<T> na = {S.cmag[9 * j + 3], S.cmag[9 * j + 4], S.cmag[9 * j + 5]};
              v3<T> nm = d / r;
              dij = r;
              cv[1] = std::fabs(dot(d, na));
              cv[3] = T(S.codes[4 * j]);
              cm[6] = S.vel[4 * j]; cm[7] = S.vel[4 * j + 1]; cm[8] = S.vel[4 * j + 2];
              T cosc = std::fabs(dot(nm, na));
              if (cosc < T(0.866f)) { cm[3] = nm.x; cm[4] = nm.y; cm[5] = nm.z; }
              else { cm[3] = na.x; cm[4] = na.y; cm[5] = na.z; }
            }
            cv[0] = cv[0] + std::fabs((cm[3] * g.x + cm[4] * g.y + cm[5] * g.z) * P.vol_def);
          }
        }
  }
}

// ----------------------------------------------------------- a15: Shepard --
// cuStressRegularizationShepard + loop, field_smoothing.cu:26-101.
template <class T> static void shepard(const Par<T>& P, State<T>& S) {
  std::vector<sym6<T>> out(P.npm);
#pragma omp parallel for schedule(dynamic, 256)
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    v3<T> pi = ld3(S.pos, i);
    v3<int> ci = cell_of(S.cellpos, i);
    sym6<T> acc = {0, 0, 0, 0, 0, 0};
    T sk = 0;
    auto scan = [&](v3<int> ci, v3<T> pi) {
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> off = {cx, cy, cz};
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int c = linear_key(P, cj) + P.ncell_material;
          for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
            v3<T> d;
            T r = pair_r(pi, ld3(S.pos, j), off, P.two_h, d);
            T q = r * P.i_h;
            if (!(q < T(2))) continue;
            T ker = kernelW(P.knorm, q);
            T vol = S.idm[4 * j + 1] / S.vel[4 * j + 3];
            T wv = ker * vol;
            sk = sk + wv;
            acc = acc + ld6(S.str, j) * wv;
          }
        }
    };
    scan(ci, pi);
    v3<int> cpb;
    v3<T> ppb;
    if (P.periodbound && period_image(P, ci, pi, true, cpb, ppb)) scan(cpb, ppb);
    if (sk != T(0)) acc = acc / sk;
    out[idx] = acc;
  }
  for (int idx = 0; idx < P.npm; idx++) st6(S.str, P.npb + idx, out[idx]);
}

// ----------------------------------------------- a10/a11: interaction ----
// cuParticleInteractionMaterial, particle_interaction.cu:26-297, with the
// wall contact (LoopContact, :142-241) resolved in two passes.
template <class T> struct Partial { v3<T> s3; v3<T> damp; T accden; T mass; };

template <class T> static void interaction(const Par<T>& P, State<T>& S) {
  const T W0 = P.kernel_zero, Wdr = P.kernel_dr;
  const T art_exp = W0 / Wdr;
  std::vector<Partial<T>> part(P.npm);
  // each iteration writes only slot i's rows; neighbours are read-only here
#pragma omp parallel for schedule(dynamic, 256)
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    S.deferred[i] = 0;
    v3<T> pi = ld3(S.pos, i);
    v3<int> ci = cell_of(S.cellpos, i);
    v3<T> vi = ld3(S.vel, i);
    T rhoi = S.vel[4 * i + 3];
    T rhoi2 = rhoi * rhoi;
    sym6<T> sri = ld6(S.str, i) / rhoi2;
    T tri = trace(sri);
    T artpi = tri > T(0) ? tri * P.art_p : T(0);
    const int shift = int(S.cvar[4 * i + 2]);
    T accden = 0;
    v3<T> as = {0, 0, 0}, av = {0, 0, 0}, aa = {0, 0, 0};
    T gv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // xx xy xz yx yy yz zx zy zz
    auto scan = [&](v3<int> ci, v3<T> pi) {
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> off = {cx, cy, cz};
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int cb = linear_key(P, cj);
          for (int pass = 0; pass < 2; pass++) {
            const bool bound = pass == 0;
            int c = bound ? cb : cb + P.ncell_material;
            for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
              v3<T> d;
              T r = pair_r(pi, ld3(S.pos, j), off, P.two_h, d);
              T q = r * P.i_h;
              if (!(q < T(2))) continue;
              T ker = kernelW(P.knorm, q);
              T dwdr = kernelG(P.kgrad, q);
              v3<T> g = d * dwdr;
              T rhoj = S.vel[4 * j + 3];
              T mj = bound ? P.vol_def * rhoj : S.idm[4 * j + 1];
              v3<T> dv = vi - ld3(S.vel, j);
              const int cj_code = S.codes[4 * j];
              if (cj_code >= 10 && shift == 1) dv = V3<T>(0, 0, 0);
              T volj = mj / rhoj;
              // grad_vel -= (dv (x) g) * vol_j
              v3<T> dvv = dv * volj;
              gv[0] = mad(-dvv.x, g.x, gv[0]); gv[1] = mad(-dvv.x, g.y, gv[1]); gv[2] = mad(-dvv.x, g.z, gv[2]);
              gv[3] = mad(-dvv.y, g.x, gv[3]); gv[4] = mad(-dvv.y, g.y, gv[4]); gv[5] = mad(-dvv.y, g.z, gv[5]);
              gv[6] = mad(-dvv.z, g.x, gv[6]); gv[7] = mad(-dvv.z, g.y, gv[7]); gv[8] = mad(-dvv.z, g.z, gv[8]);
              if (cj_code >= 10) continue;
              T dvg = mad(dv.z, g.z, mad(dv.y, g.y, dv.x * g.x));
              accden = mad(rhoi * dvg, volj, accden);
              T rhoj2 = rhoj * rhoj;
              sym6<T> srj = ld6(S.str, j) / rhoj2;
              sym6<T> st = sri + srj;
              v3<T> f = matvec(st, g);
              as.x = mad(f.x, mj, as.x); as.y = mad(f.y, mj, as.y); as.z = mad(f.z, mj, as.z);
              if (P.viscosity_type == 1) {
                T visc = mad(dv.z, d.z, mad(dv.y, d.y, dv.x * d.x));
                if (visc < T(0)) {
                  T phi = P.h * visc / mad(r, r, P.eta);
                  T artv = P.alpha_m * P.cs * phi / (T(0.5) * (rhoi + rhoj));
                  T s = mj * artv;
                  av.x = mad(g.x, s, av.x); av.y = mad(g.y, s, av.y); av.z = mad(g.z, s, av.z);
                }
              }
              if (P.art_p != T(0)) {
                T trj = trace(srj);
                T artp = artpi + (trj > T(0) ? trj * P.art_p : T(0));
                if (artp > T(0)) {
                  artp = artp * std::pow(ker / Wdr, art_exp);
                  T s = mj * artp;
                  aa.x = mad(-g.x, s, aa.x); aa.y = mad(-g.y, s, aa.y); aa.z = mad(-g.z, s, aa.z);
                }
              }
            }
          }
        }
    };
    scan(ci, pi);
    v3<int> cpb;
    v3<T> ppb;
    if (P.periodbound && period_image(P, ci, pi, false, cpb, ppb)) scan(cpb, ppb);
    v3<T> damp = vi * P.damping;
    v3<T> s3 = (as + av) + aa;
    T mi = S.idm[4 * i + 1];
    const v3<T> gi = body_acc(P, S.codes, i, mi);
    v3<T> acc0 = (s3 - damp) + gi;
    // predicted velocity v + a dt (particle_interaction.cu:287)
    v3<T> vp = V3(mad(acc0.x, P.dt, vi.x), mad(acc0.y, P.dt, vi.y), mad(acc0.z, P.dt, vi.z));
    v3<T> nb = {S.cmag[9 * i + 3], S.cmag[9 * i + 4], S.cmag[9 * i + 5]};
    v3<T> vb = {S.cmag[9 * i + 6], S.cmag[9 * i + 7], S.cmag[9 * i + 8]};
    v3<T> dvc = vb - vp;  // mass_rate = 1
    v3<T> ast = {0, 0, 0};
    T dvn = dot(dvc, nb);
    if (S.cvar[4 * i + 1] <= P.h * T(2) && dvn > T(0)) {
      const T mu = T(0.58108f);  // particle_interaction.cu:188-189
      v3<T> dvn3 = nb * dvn;
      v3<T> fn = dvn3 * mi / P.dt;
      v3<T> ft = (dvc - dvn3) * mi / P.dt;
      T afn = norm(fn), aft = norm(ft);
      if (aft > mu * afn) { ft = ft * (mu * afn / aft); S.cvar[4 * i + 2] = T(1); }
      else S.cvar[4 * i + 2] = T(0);
      v3<T> fr = ft + fn;
      if (S.cvar[4 * i + 1] > P.dr) {
        S.deferred[i] = 1;
      } else {
        S.cmag[9 * i] = fr.x; S.cmag[9 * i + 1] = fr.y; S.cmag[9 * i + 2] = fr.z;
        ast = fr * (T(1) / mi);
      }
    }
    part[idx] = {s3, damp, accden, mi};
    v3<T> acc = ((s3 + ast) - damp) + gi;
    S.accad[4 * i] = acc.x; S.accad[4 * i + 1] = acc.y; S.accad[4 * i + 2] = acc.z; S.accad[4 * i + 3] = accden;
    sym6<T> D = {gv[0], gv[4], gv[8], T(0.5) * (gv[1] + gv[3]), T(0.5) * (gv[2] + gv[6]), T(0.5) * (gv[5] + gv[7])};
    st6(S.acestn, i, D);
    S.spin[4 * i] = T(0.5) * (gv[3] - gv[1]);
    S.spin[4 * i + 1] = T(0.5) * (gv[6] - gv[2]);
    S.spin[4 * i + 2] = T(0.5) * (gv[7] - gv[5]);
    S.spin[4 * i + 3] = T(0);
  }
  // pass B: redistribution for deferred particles (reads pass-A values only)
  struct Out { int i; T K; v3<T> f; };
  std::vector<Out> outs;
  const T band = P.dr * T(1.0001f);
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    if (!S.deferred[i]) continue;
    v3<int> ci = cell_of(S.cellpos, i);
    v3<T> pi = ld3(S.pos, i);
    v3<T> pa = V3(pi.x + T(ci.x) * P.two_h, pi.y + T(ci.y) * P.two_h, pi.z + T(ci.z) * P.two_h);
    v3<T> ni = {S.cmag[9 * i + 3], S.cmag[9 * i + 4], S.cmag[9 * i + 5]};
    T K = S.cvar[4 * i];
    v3<T> fr = {0, 0, 0};
    int count = 0;
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int c = linear_key(P, cj) + P.ncell_material;
          for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
            if (!(S.cvar[4 * j + 1] <= band)) continue;
            v3<int> cjj = cell_of(S.cellpos, j);
            v3<T> pj = ld3(S.pos, j);
            v3<T> pja = V3(pj.x + T(cjj.x) * P.two_h, pj.y + T(cjj.y) * P.two_h, pj.z + T(cjj.z) * P.two_h);
            v3<T> dc = pa - pja;
            T rcol = std::fabs(dc.x * ni.z + dc.y * ni.y + dc.z * ni.x);
            T Kj = S.cvar[4 * j];
            if (rcol <= P.dr / T(2) && Kj != T(0)) {
              if (K > Kj) K = Kj;
              fr = V3(S.cmag[9 * j] * K / Kj, S.cmag[9 * j + 1] * K / Kj, S.cmag[9 * j + 2] * K / Kj);
              count++;
            }
          }
        }
    if (count == 0) fr = fr * T(0);
    outs.push_back({i, K, fr});
  }
  for (const Out& o : outs) {
    const Partial<T>& pt = part[o.i - P.npb];
    S.cvar[4 * o.i] = o.K;
    S.cmag[9 * o.i] = o.f.x; S.cmag[9 * o.i + 1] = o.f.y; S.cmag[9 * o.i + 2] = o.f.z;
    v3<T> ast = o.f * (T(1) / pt.mass);
    v3<T> acc = ((pt.s3 + ast) - pt.damp) + body_acc(P, S.codes, o.i, pt.mass);
    S.accad[4 * o.i] = acc.x; S.accad[4 * o.i + 1] = acc.y; S.accad[4 * o.i + 2] = acc.z;
  }
}

// ------------------------------------------------------- a14: HypoCS ----
// cuMatModHypo (critical-state overload), constitutive_model.h:209-299.
template <class T> static void hypo_cs(T dt, T rho, v3<T> w, sym6<T> D, sym6<T>& sig, const T* mp, T& Mcrt) {
  const T Ei = mp[0], Vi = mp[1], phic = mp[2], coh = mp[3], pref = mp[4], ec0 = mp[5], lam = mp[6];
  const T xi = mp[7], alpha = mp[8], emin = mp[9], emax = mp[10], e0 = mp[11], rho0 = mp[12];
  const T C1 = -Ei / (T(3) * pref * (T(1) + Vi));
  const T a = T(3) / (std::sqrt(T(2)) * std::sin(phic));
  const T C2 = C1 * ((T(3) * Vi + a * std::sqrt(T(1) + Vi * Vi)) / (T(1) - Vi) - a * a / T(3));
  const T C3 = a * a * C1;
  const T C4 = a * C1;
  // Jaumann: W S - S W with W = SkewTensor2FullTensor(w) (functions.h:106-110)
  sym6<T> s = sig;
  if (coh > T(0)) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  // (WS - SW) symmetric part; W = [[0,-wz,wy],[wz,0,-wx],[-wy,wx,0]]
  T Sxx = s.xx, Syy = s.yy, Szz = s.zz, Sxy = s.xy, Sxz = s.xz, Syz = s.yz;
  full9<T> WS = {
      -w.z * Sxy + w.y * Sxz, -w.z * Syy + w.y * Syz, -w.z * Syz + w.y * Szz,
      w.z * Sxx - w.x * Sxz, w.z * Sxy - w.x * Syz, w.z * Sxz - w.x * Szz,
      -w.y * Sxx + w.x * Sxy, -w.y * Sxy + w.x * Syy, -w.y * Sxz + w.x * Syz};
  full9<T> SW = {
      Sxy * w.z - Sxz * w.y, -Sxx * w.z + Sxz * w.x, Sxx * w.y - Sxy * w.x,
      Syy * w.z - Syz * w.y, -Sxy * w.z + Syz * w.x, Sxy * w.y - Syy * w.x,
      Syz * w.z - Szz * w.y, -Sxz * w.z + Szz * w.x, Sxz * w.y - Syz * w.x};
  full9<T> J = {WS.xx - SW.xx, WS.xy - SW.xy, WS.xz - SW.xz, WS.yx - SW.yx, WS.yy - SW.yy,
                WS.yz - SW.yz, WS.zx - SW.zx, WS.zy - SW.zy, WS.zz - SW.zz};
  sym6<T> jau = {J.xx, J.yy, J.zz, T(0.5) * (J.xy + J.yx), T(0.5) * (J.xz + J.zx), T(0.5) * (J.yz + J.zy)};
  T I1 = trace(s);
  if (I1 > T(-4)) { sig = sym6<T>{-1, -1, -2, 0, 0, 0}; return; }
  sym6<T> dev = s; T third = I1 / T(3);
  dev.xx = dev.xx - third; dev.yy = dev.yy - third; dev.zz = dev.zz - third;
  T trs = trace(s), trd = trace(D), trsd = ddot(s, D), nd = std::sqrt(ddot(D, D));
  T p = -I1 / T(3);
  T e = (T(1) + e0) * rho0 / rho - T(1);
  e = std::max(e, emin);
  e = std::min(e, emax);
  T ec = ec0 * std::exp(-lam * std::pow(p / pref, xi));
  T fd = std::pow((e - emin) / (ec - emin), alpha);
  Mcrt = e - ec;
  sym6<T> sd = s + dev;
  sym6<T> rate = jau + D * (C1 * trs) + s * (C2 * trd) + s * (C3 * trsd / trs) + sd * (fd * C4 * nd);
  sym6<T> ns = sig + rate * dt;
  T sph = trace(ns) / T(3);
  if (sph > T(0)) { sig = sym6<T>{T(-1) + coh, T(-1) + coh, T(-2) + coh, 0, 0, 0}; return; }
  sym6<T> nd6 = ns;
  nd6.xx = nd6.xx - sph; nd6.yy = nd6.yy - sph; nd6.zz = nd6.zz - sph;
  T fa = C2 / C1, fb = C3 / C1, fc = C4 / C1;
  T u = T(1) + fa + fb / T(3);
  T f2 = fd * fd * fc * fc;
  T A = f2 / T(3) - u * u;
  T B = T(2) * f2 * (u * u + T(2) * fa + T(3)) - T(4) * fb * u;
  T C = T(4) / T(3) * f2 * fb * fb - T(4) * fb * fb;
  T k2 = (-B - std::sqrt(B * B - T(4) * A * C)) / (T(2) * A);
  T kphi = T(3) / std::sqrt(k2);
  T tau = std::sqrt(T(0.5) * ddot(nd6, nd6));
  if (tau - (-sph * kphi) > T(0)) {
    T ratio = tau < T(1e-10) ? T(1) : (-kphi * sph) / tau;
    nd6 = nd6 * ratio;
  }
  sig = sym6<T>{nd6.xx + sph + coh, nd6.yy + sph + coh, nd6.zz + sph + coh, nd6.xy, nd6.xz, nd6.yz};
}

// ------------------------------- DP, mu(I)-DP, basic hypoplasticity ----
// The Jaumann term sym(W S - S W), W = SkewTensor2FullTensor(w), as hypo_cs
// writes it.
template <class T> static sym6<T> jaumann(v3<T> w, sym6<T> s) {
  const T Sxx = s.xx, Syy = s.yy, Szz = s.zz, Sxy = s.xy, Sxz = s.xz, Syz = s.yz;
  full9<T> WS = {
      -w.z * Sxy + w.y * Sxz, -w.z * Syy + w.y * Syz, -w.z * Syz + w.y * Szz,
      w.z * Sxx - w.x * Sxz, w.z * Sxy - w.x * Syz, w.z * Sxz - w.x * Szz,
      -w.y * Sxx + w.x * Sxy, -w.y * Sxy + w.x * Syy, -w.y * Sxz + w.x * Syz};
  full9<T> SW = {
      Sxy * w.z - Sxz * w.y, -Sxx * w.z + Sxz * w.x, Sxx * w.y - Sxy * w.x,
      Syy * w.z - Syz * w.y, -Sxy * w.z + Syz * w.x, Sxy * w.y - Syy * w.x,
      Syz * w.z - Szz * w.y, -Sxz * w.z + Szz * w.x, Sxz * w.y - Syz * w.x};
  return sym6<T>{WS.xx - SW.xx, WS.yy - SW.yy, WS.zz - SW.zz, T(0.5) * ((WS.xy - SW.xy) + (WS.yx - SW.yx)),
                 T(0.5) * ((WS.xz - SW.xz) + (WS.zx - SW.zx)), T(0.5) * ((WS.yz - SW.yz) + (WS.zy - SW.zy))};
}

// Elastic trial stress shared by both Drucker-Prager models:
// sigma + (Jaumann + 2G D + lambda tr(D) I) dt, lambda = K - 2G/3.
template <class T> static sym6<T> elastic_trial(T dt, v3<T> w, sym6<T> D, sym6<T> sig, T K, T G) {
  const T lame = K - T(2) / T(3) * G;
  sym6<T> rate = jaumann(w, sig) + D * (T(2) * G);
  const T l = lame * trace(D);
  rate.xx = rate.xx + l; rate.yy = rate.yy + l; rate.zz = rate.zz + l;
  return sig + rate * dt;
}

template <class T> static sym6<T> add_iso(sym6<T> s, T p) {
  s.xx = s.xx + p; s.yy = s.yy + p; s.zz = s.zz + p;
  return s;
}

// cuMatModDP, constitutive_model.h:38-84: Drucker-Prager, elastic-perfectly
// plastic with a tension cut-off; mp = {K, G, k_phi, k_psi, k_c, ...}
// (GetMaterial_DP_EPP, GenCase.cu:16-25).  yld is set by a shear return.
template <class T> static void dp_epp(T dt, v3<T> w, sym6<T> D, sym6<T>& sig, const T* mp, bool& yld) {
  const T K = mp[0], G = mp[1], kphi = mp[2], kpsi = mp[3], kc = mp[4];
  const sym6<T> pre = elastic_trial(dt, w, D, sig, K, G);
  T sph = trace(pre) / T(3);
  sym6<T> dev = add_iso(pre, -(trace(pre) / T(3)));
  const T tenlt = kphi == T(0) ? T(0) : kc / kphi;
  if (sph - tenlt >= T(0)) sph = tenlt;
  const T tau = std::sqrt(T(0.5) * ddot(dev, dev));
  const T shr = tau - (-sph * kphi + kc);
  if (shr > T(0)) {
    const T dlam = shr / (G + K * kphi * kpsi);
    sph = sph - K * kpsi * dlam;
    const T ratio = tau < T(1e-10) ? T(1) : (kc - kphi * sph) / tau;
    dev = dev * ratio;
    yld = true;
  }
  sig = add_iso(dev, sph);
}

// cuMatModMuIDP, constitutive_model.h:90-130: Drucker-Prager with the mu(I)
// friction law; mp = {K, G, mu_s, mu_2, I0, rho_g, d, lambda, ...}
// (GetMaterial_MuI_DP_EPP, GenCase.cu:27-40).
template <class T> static void mui_dp(T dt, v3<T> w, sym6<T> D, sym6<T>& sig, const T* mp) {
  const T K = mp[0], G = mp[1], mus = mp[2], dmu = mp[3] - mp[2], I0 = mp[4], rhog = mp[5], dg = mp[6];
  const T lam = mp[7];
  const sym6<T> pre = elastic_trial(dt, w, D, sig, K, G);
  T sph = trace(pre) / T(3);
  sym6<T> dev = add_iso(pre, -(trace(pre) / T(3)));
  const sym6<T> dD = add_iso(D, -(trace(D) / T(3)));
  const T gamma = std::sqrt(T(2) * ddot(dD, dD));
  if (sph > T(0)) sph = T(0);
  const T p = -sph;
  const T mu = mus + dmu * gamma * dg / (I0 * std::sqrt(p / rhog) + dg * std::sqrt(gamma * gamma + lam * lam));
  const T tau = std::sqrt(T(0.5) * ddot(dev, dev));
  if (tau + sph * mu > T(0)) {
    const T ratio = tau < T(1e-10) ? T(1) : (T(0) - mu * sph) / tau;
    dev = dev * ratio;
  }
  sig = add_iso(dev, sph);
}

// cuMatModHypo (basic overload), constitutive_model.h:136-202; mp = {Ei, Vi,
// phi_c, c, p_ref, ...} (GetMaterial_HypoBasic, GenCase.cu:42-50).
template <class T> static void hypo_basic(T dt, v3<T> w, sym6<T> D, sym6<T>& sig, const T* mp) {
  const T Ei = mp[0], Vi = mp[1], phic = mp[2], coh = mp[3], pref = mp[4];
  const T C1 = -Ei / (T(3) * pref * (T(1) + Vi));
  const T a = T(3) / (std::sqrt(T(2)) * std::sin(phic));
  const T C2 = C1 * ((T(3) * Vi + a * std::sqrt(T(1) + Vi * Vi)) / (T(1) - Vi) - a * a / T(3));
  const T C3 = a * a * C1;
  const T C4 = a * C1;
  sym6<T> s = sig;
  if (coh > T(0)) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  const sym6<T> jau = jaumann(w, s);
  const T I1 = trace(s);
  if (I1 > T(-4)) { sig = sym6<T>{-1, -1, -2, 0, 0, 0}; return; }
  const sym6<T> dev = add_iso(s, -(I1 / T(3)));
  const T trs = trace(s), trd = trace(D), trsd = ddot(s, D), nd = std::sqrt(ddot(D, D));
  const sym6<T> rate = jau + D * (C1 * trs) + s * (C2 * trd) + s * (C3 * trsd / trs) + (s + dev) * (C4 * nd);
  const sym6<T> ns = sig + rate * dt;
  const T sph = trace(ns) / T(3);
  if (sph > T(0)) { sig = sym6<T>{T(-1) + coh, T(-1) + coh, T(-2) + coh, 0, 0, 0}; return; }
  sym6<T> nd6 = add_iso(ns, -sph);
  const T kphi = T(3) / (std::sqrt(T(2)) * a);
  const T tau = std::sqrt(T(0.5) * ddot(nd6, nd6));
  if (tau - (-sph * kphi) > T(0)) {
    const T ratio = tau < T(1e-10) ? T(1) : (-kphi * sph) / tau;
    nd6 = nd6 * ratio;
  }
  sig = add_iso(nd6, sph + coh);
}

// The model dispatch of cuPredictorMaterial / cuCorrectorMaterial
// (predictor_corrector.cu:72-75, 169-172); RIGID and unknown types leave
// the stress unchanged.
template <class T> static void constitutive(int model_type, T dt, T rho, v3<T> w, sym6<T> D, sym6<T>& sig, const T* mp,
                                           T& Mcrt, bool& yld) {
  if (model_type == LQ_DP_EPP) dp_epp(dt, w, D, sig, mp, yld);
  else if (model_type == LQ_MUI_DP_EPP) mui_dp(dt, w, D, sig, mp);
  else if (model_type == LQ_HYPO_BASIC) hypo_basic(dt, w, D, sig, mp);
  else if (model_type == LQ_HYPO_CS) hypo_cs(dt, rho, w, D, sig, mp, Mcrt);
}

// ---------------------------------------------------- a12: predictor ----
// LoquatSph::Predictor + cuPredictorMaterial, predictor_corrector.cu:52-103.
template <class T> static void predictor(const Par<T>& P, State<T>& S) {
  std::memcpy(S.pos_ini, S.pos, sizeof(T) * 4 * P.np);
  std::memcpy(S.vel_ini, S.vel, sizeof(T) * 4 * P.np);
  if (P.npm == 0) return;
  std::memcpy(S.str_ini, S.str, sizeof(T) * 6 * (P.npb + P.npm));
  std::memcpy(S.stn_ini, S.stn, sizeof(T) * 6 * (P.npb + P.npm));
  const T hdt = T(0.5) * P.dt;
#pragma omp parallel for schedule(static)
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    if (S.codes[4 * i + 1] != LQ_NONFIX) continue;
    for (int k = 0; k < 4; k++) S.vel[4 * i + k] = mad(S.accad[4 * i + k], hdt, S.vel[4 * i + k]);
    for (int k = 0; k < 3; k++) S.pos[4 * i + k] = mad(S.vel[4 * i + k], hdt, S.pos[4 * i + k]);
    const T* mp = P.model_par[S.codes[4 * i + 3]];
    T rho = std::max(S.vel[4 * i + 3], mp[13]);
    rho = std::min(rho, mp[14]);
    S.vel[4 * i + 3] = rho;
    sym6<T> D = ld6(S.acestn, i);
    st6(S.stn, i, ld6(S.stn, i) + D * hdt);
    sym6<T> sig = ld6(S.str, i);
    T Mcrt = 0;
    bool yld = false;
    constitutive(S.codes[4 * i + 2], hdt, rho, ld3(S.spin, i), D, sig, mp, Mcrt, yld);
    st6(S.str, i, sig);
    if (!yld) S.idm[4 * i] = T(100);  // a Drucker-Prager yield keeps the old value (:78-83)
  }
}

// ---------------------------------------------------- a13: corrector ----
static inline int rebin_d(double p, double two_h) { return int(p >= two_h) - int(p < 0.0); }

// cuCorrectorBound, predictor_corrector.cu:109-137.
template <class T> static void corrector_bound(const Par<T>& P, State<T>& S) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P.npb; i++) {
    if (S.codes[4 * i + 1] != LQ_NONFIX) continue;
    const Motion<T>& m = P.mv[S.codes[4 * i + 3]];
    v3<T> vel = m.trans_vel;
    if (m.is_rotat && m.pattern != LQ_CIRCLE) {
      v3<T> pi = ld3(S.pos, i);
      v3<int> ci = cell_of(S.cellpos, i);
      v3<T> pa = V3(pi.x + T(ci.x) * P.two_h, pi.y + T(ci.y) * P.two_h, pi.z + T(ci.z) * P.two_h);
      vel = vel + cross(m.rotat_vel, pa - m.pos_rp);
      v3<T> ang = m.rotat_vel * P.dt;  // GetRodriguesMatrix, functions.h:126-131
      T th = norm(ang);
      if (th > T(0)) {
        v3<T> k = ang / th;
        T sn = std::sin(th), cc = T(1) - std::cos(th);
        // R = I + K sin + K K (1-cos), K = skew(k)
        T Kxx = -(k.z * k.z + k.y * k.y), Kyy = -(k.z * k.z + k.x * k.x), Kzz = -(k.y * k.y + k.x * k.x);
        full9<T> R = {T(1) + Kxx * cc, -k.z * sn + k.x * k.y * cc, k.y * sn + k.x * k.z * cc,
                      k.z * sn + k.x * k.y * cc, T(1) + Kyy * cc, -k.x * sn + k.y * k.z * cc,
                      -k.y * sn + k.x * k.z * cc, k.x * sn + k.y * k.z * cc, T(1) + Kzz * cc};
        v3<T> n = ld3(S.normal, i);
        v3<T> nr = {R.xx * n.x + R.xy * n.y + R.xz * n.z, R.yx * n.x + R.yy * n.y + R.yz * n.z,
                    R.zx * n.x + R.zy * n.y + R.zz * n.z};
        st3(S.normal, i, nr);
      }
    }
    v3<T> p = ld3(S.pos_ini, i) + vel * P.dt;
    S.pos[4 * i + 3] = S.pos_ini[4 * i + 3];
    int dx = rebin_d(p.x, P.two_h), dy = rebin_d(p.y, P.two_h), dz = rebin_d(p.z, P.two_h);
    p.x = p.x - T(dx) * P.two_h; p.y = p.y - T(dy) * P.two_h; p.z = p.z - T(dz) * P.two_h;
    st3(S.pos, i, p);
    S.cellpos[4 * i] += dx; S.cellpos[4 * i + 1] += dy; S.cellpos[4 * i + 2] += dz;
  }
}

// cuCorrectorMaterial, predictor_corrector.cu:139-199.
template <class T> static int corrector_material(const Par<T>& P, State<T>& S) {
  int left = 0;
#pragma omp parallel for schedule(static) reduction(+ : left)
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx;
    if (S.codes[4 * i + 1] != LQ_NONFIX) continue;
    for (int k = 0; k < 4; k++) S.vel[4 * i + k] = mad(S.accad[4 * i + k], P.dt, S.vel_ini[4 * i + k]);
    const T hdt = T(0.5) * P.dt;
    for (int k = 0; k < 3; k++) S.pos[4 * i + k] = mad(S.vel[4 * i + k] + S.vel_ini[4 * i + k], hdt, S.pos_ini[4 * i + k]);
    S.pos[4 * i + 3] = S.pos_ini[4 * i + 3];
    const T* mp = P.model_par[S.codes[4 * i + 3]];
    T rho = std::max(S.vel[4 * i + 3], mp[13]);
    rho = std::min(rho, mp[14]);
    S.vel[4 * i + 3] = rho;
    int d[3];
    for (int k = 0; k < 3; k++) {
      d[k] = rebin_d(S.pos[4 * i + k], P.two_h);
      S.pos[4 * i + k] = S.pos[4 * i + k] - T(d[k]) * P.two_h;
      S.cellpos[4 * i + k] += d[k];
    }
    if (P.periodbound) {  // :158
      v3<int> c = cell_of(S.cellpos, i);
      v3<T> pr = ld3(S.pos, i);
      wrap_period(P, c, pr);
      st3(S.pos, i, pr);
      S.cellpos[4 * i] = c.x; S.cellpos[4 * i + 1] = c.y; S.cellpos[4 * i + 2] = c.z;
    }
    sym6<T> D = ld6(S.acestn, i);
    st6(S.stn, i, ld6(S.stn_ini, i) + D * P.dt);
    sym6<T> sig = ld6(S.str_ini, i);
    T Mcrt = 0;
    bool yld = false;
    constitutive(S.codes[4 * i + 2], P.dt, rho, ld3(S.spin, i), D, sig, mp, Mcrt, yld);
    st6(S.str, i, sig);
    if (!yld) S.idm[4 * i] = T(100);  // :175-180
    S.idm[4 * i + 3] = Mcrt;
    v3<int> c = cell_of(S.cellpos, i);
    if (!in_grid(P, c)) { S.cellpos[4 * i + 3] = P.ncell_delete_m; S.idm[4 * i] = T(999); left++; }
    if (P.time <= P.dt) for (int k = 0; k < 4; k++) S.pos_ori[4 * i + k] = T(0);
    for (int k = 0; k < 3; k++) S.pos_ori[4 * i + k] = mad(S.vel[4 * i + k], P.dt, S.pos_ori[4 * i + k]);
    for (int k = 0; k < 4; k++) S.accad[4 * i + k] = S.pos_ori[4 * i + k];
  }
  return left;
}

// --------------------------------------------------- a16: wall force ----
// cuGetValuesAtPoint, utilities.cu:578-640 (box bounds passed explicitly).
template <class T> static double wall_force(const Par<T>& P, const State<T>& S, const float box[3]) {
  double F = 0.0;
  for (int i = P.npb; i < P.np; i++) {
    v3<int> c = cell_of(S.cellpos, i);
    T xa = S.pos[4 * i] + T(c.x) * P.two_h;
    T za = S.pos[4 * i + 2] + T(c.z) * P.two_h;
    if (xa <= T(box[0]) && za >= T(box[1]) && za <= T(box[2])) F += 0.001 * double(S.cmag[9 * i]);
  }
  return F;
}

// cuGetValuesAlongLine, utilities.cu:266-320, with the case constants as
// arguments: geom = {x_face, z_lo, z_hi, z_top}; out[nbins][4] = depth
// (set by band-1 writes), band-1..3 contact force x / dr of the last slot
// (serial slot order) in each bin; unwritten entries stay 0.
template <class T> static void wall_profile(const Par<T>& P, const State<T>& S, const float geom[4], int nbins,
                                            double* out) {
  for (int k = 0; k < 4 * nbins; k++) out[k] = 0.0;
  const T dr = P.dr;
  for (int i = P.npb; i < P.np; i++) {
    const v3<T> a = abs_pos(P, cell_of(S.cellpos, i), ld3(S.pos, i));
    if (!(a.z >= T(geom[1]) && a.z <= T(geom[2]))) continue;
    const int bin = int((T(geom[3]) - a.z) / dr);
    if (bin < 0 || bin >= nbins) continue;
    for (int b = 0; b < 3; b++) {
      if (a.x >= T(geom[0]) + dr * T(b) && a.x <= T(geom[0]) + dr * T(b + 1)) {
        if (b == 0) out[4 * bin] = double(T(bin) * dr);
        out[4 * bin + 1 + b] = double(S.cmag[9 * i] / dr);
      }
    }
  }
}

// GetSoilStress + cuGetStressZz, utilities.cu:470-528, at the given
// absolute points: sum_j sigma_j W V_j / sum_j W V_j over soil slots with
// q = |x - x_j| / h < 2 (x_j = relative + cell 2h), summed serially in cell
// order; rows: the reference's cells (own y row, every z level, z outer, x
// inner), else the 27 cells (z, y, x nested).  out[n][8] = (xx yy zz xy xz
// yz, sum W V, slots).
template <class T> static void sensor_stress(const Par<T>& P, const State<T>& S, const float* pts, int n, int rows,
                                             double* out) {
  for (int k = 0; k < n; k++) {
    const T x = pts[3 * k], y = pts[3 * k + 1], z = pts[3 * k + 2];
    const T scale = T(0.5) * P.i_h;
    const int cx0 = int(std::floor(x * scale)), cy0 = int(std::floor(y * scale)), cz0 = int(std::floor(z * scale));
    T sum[6] = {0, 0, 0, 0, 0, 0};
    T wsum = 0;
    int hits = 0;
    const int z0 = rows ? 0 : cz0 - 1, z1 = rows ? P.cnz - 1 : cz0 + 1;
    const int y0 = rows ? cy0 : cy0 - 1, y1 = rows ? cy0 : cy0 + 1;
    for (int cz = z0; cz <= z1; cz++)
      for (int cy = y0; cy <= y1; cy++)
        for (int cx = cx0 - 1; cx <= cx0 + 1; cx++) {
          if (cx < 0 || cy < 0 || cz < 0 || cx >= P.cnx || cy >= P.cny || cz >= P.cnz) continue;
          const int c = linear_key(P, v3<int>{cx, cy, cz}) + P.ncell_material;
          if (S.cell_beg[c] < 0) continue;
          for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
            const v3<T> a = abs_pos(P, cell_of(S.cellpos, j), ld3(S.pos, j));
            const T dx = x - a.x, dy = y - a.y, dz = z - a.z;
            const T q = std::sqrt(mad(dz, dz, mad(dy, dy, dx * dx))) * P.i_h;
            if (!(q < T(2))) continue;
            const T ker = kernelW(P.knorm, q);
            const T vol = S.idm[4 * j + 1] / S.vel[4 * j + 3];
            for (int m = 0; m < 6; m++) sum[m] = sum[m] + S.str[6 * j + m] * ker * vol;
            wsum = wsum + ker * vol;
            hits++;
          }
        }
    if (wsum != T(0))
      for (int m = 0; m < 6; m++) sum[m] = sum[m] / wsum;
    for (int m = 0; m < 6; m++) out[8 * k + m] = double(sum[m]);
    out[8 * k + 6] = double(wsum);
    out[8 * k + 7] = double(hits);
  }
}

// GetSurface (utilities.cu:26-91) / GetSurfaceAtPoint (:371-427) at the
// given (x, y) points, on absolute positions: the largest z of a soil slot
// with horizontal distance < radius over cells x +-1, y +-1 (radius > 0),
// or with |dx| < -radius over the point's own y row, x +-1 (radius < 0);
// every z level; -1e10 when none.
template <class T> static void surface_heights(const Par<T>& P, const State<T>& S, const float* pts, int n,
                                               float radius, double* out) {
  for (int k = 0; k < n; k++) {
    const T x = pts[2 * k], y = pts[2 * k + 1];
    const T scale = T(0.5) * P.i_h;
    const int cx0 = int(std::floor(x * scale)), cy0 = int(std::floor(y * scale));
    const bool strip = radius < 0.f;
    const int y0 = strip ? cy0 : cy0 - 1, y1 = strip ? cy0 : cy0 + 1;
    T top = T(-1.0e10f);
    for (int cz = 0; cz < P.cnz; cz++)
      for (int cy = y0; cy <= y1; cy++)
        for (int cx = cx0 - 1; cx <= cx0 + 1; cx++) {
          if (cx < 0 || cy < 0 || cx >= P.cnx || cy >= P.cny) continue;
          const int c = linear_key(P, v3<int>{cx, cy, cz}) + P.ncell_material;
          if (S.cell_beg[c] < 0) continue;
          for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
            const v3<T> a = abs_pos(P, cell_of(S.cellpos, j), ld3(S.pos, j));
            const T dx = x - a.x, dy = y - a.y;
            const bool near = strip ? std::fabs(dx) < T(-radius) : std::sqrt(mad(dy, dy, dx * dx)) < T(radius);
            if (near && a.z > top) top = a.z;
          }
        }
    out[k] = double(top);
  }
}

}  // namespace oracle
