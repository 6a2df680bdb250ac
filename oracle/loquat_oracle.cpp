// loquat_oracle.cpp -- CPU ORACLE entry points (test infrastructure only).
//
// extern "C" wrappers over loquat_oracle.hpp for T = float (suffix _f32) and
// T = double (suffix _f64), loaded by tests/ through ctypes and run as the
// CPU baseline by bench.py.  The product library never links this file.
#include "loquat_oracle.hpp"

#include <chrono>

#include <omp.h>

using namespace oracle;

extern "C" {

struct or_state {
  void *pos, *pos_ori, *vel, *idm, *pos_ini, *vel_ini, *accad, *spin, *cvar, *normal;
  void *str, *stn, *str_ini, *stn_ini, *acestn;
  void* cmag;
  int32_t *codes, *cellpos;
  int32_t *part_cel, *part_idx, *orig_id;
  int32_t *cell_beg, *cell_end;
  int32_t* deferred;
};

// InitializeParameters, loquat.cu:130-190 (3D branch; 1D/2D kept for the
// kernel constants).  Float arithmetic where the reference uses scalar=float,
// double where it writes double literals.
int or_initialize_parameters(lq_params* p) {
  p->h = p->dr * p->hdr;
  p->i_h = 1.0f / p->h;
  p->eta = 0.01f * p->h * p->h;
  p->bound_thickness = float(std::ceil(2.0 * double(p->hdr)) * double(p->dr));
  const double hd = double(p->h);
  const double pi = 3.14159265358979323846;
  if (p->dim == 1) {
    p->vol_def = p->dr;
    p->cell_num[0] = 1; p->cell_num[1] = 1;
    p->kernel_normalization_par = float(3.0 / (32 * hd));
    p->kernel_gradient_normalization_par = float(-15.0 / (32 * std::pow(hd, 3)));
  } else if (p->dim == 2) {
    p->vol_def = float(std::pow(double(p->dr), 2));
    p->cell_num[1] = 1;
    p->cell_num[0] = int(0.5f * p->i_h * p->domain_size[0]) + 1;
    p->kernel_normalization_par = float(7.0 / (32.0 * pi * std::pow(hd, 2)));
    p->kernel_gradient_normalization_par = float(-35.0 / (32 * pi * std::pow(hd, 4)));
  } else if (p->dim == 3) {
    p->vol_def = float(std::pow(double(p->dr), 3));
    p->cell_num[0] = int(0.5f * p->i_h * p->domain_size[0]) + 1;
    p->cell_num[1] = int(0.5f * p->i_h * p->domain_size[1]) + 1;
    p->kernel_normalization_par = float(42.0 / (256.0 * pi * std::pow(hd, 3)));
    p->kernel_gradient_normalization_par = float(-5.0 * 42 / (256.0 * pi * std::pow(hd, 5)));
  } else {
    return LQ_ERR_ARG;
  }
  p->cell_num[2] = int(0.5f * p->i_h * p->domain_size[2]) + 1;
  int layer = p->cell_num[0] * p->cell_num[1] * p->cell_num[2];
  p->ncell_material = layer;
  p->ncell_delete_m = 2 * layer;
  p->ncell = p->ncell_delete_m + 1;
  p->np_delete_m = 0;
  float q = p->dr * p->i_h;
  p->kernel_zero = p->kernel_normalization_par * 2.0f * 2.0f * 2.0f * 2.0f * 0.5f;
  p->kernel_dr = p->kernel_normalization_par * (2.0f - q) * (2.0f - q) * (2.0f - q) * (2.0f - q) * (q + 0.5f);
  return 0;
}

}  // extern "C"

template <class T> static State<T> view(or_state* s) {
  State<T> S;
  S.pos = (T*)s->pos; S.pos_ori = (T*)s->pos_ori; S.vel = (T*)s->vel; S.idm = (T*)s->idm;
  S.pos_ini = (T*)s->pos_ini; S.vel_ini = (T*)s->vel_ini; S.accad = (T*)s->accad; S.spin = (T*)s->spin;
  S.cvar = (T*)s->cvar; S.normal = (T*)s->normal;
  S.str = (T*)s->str; S.stn = (T*)s->stn; S.str_ini = (T*)s->str_ini; S.stn_ini = (T*)s->stn_ini;
  S.acestn = (T*)s->acestn; S.cmag = (T*)s->cmag;
  S.codes = s->codes; S.cellpos = s->cellpos; S.part_cel = s->part_cel; S.part_idx = s->part_idx;
  S.orig_id = s->orig_id; S.cell_beg = s->cell_beg; S.cell_end = s->cell_end; S.deferred = s->deferred;
  return S;
}

template <class T> static void tab_in(const lq_motion_table& a, MotionTab<T>& m) {
  m.is_translation = a.is_translation; m.table_size_trans = a.table_size_trans;
  for (int j = 0; j < LQ_TABLE_MAX; j++) {
    m.trans_vel[j] = V3<T>(a.trans_vel[j][0], a.trans_vel[j][1], a.trans_vel[j][2]);
    m.trans_acc[j] = V3<T>(a.trans_acc[j][0], a.trans_acc[j][1], a.trans_acc[j][2]);
    m.trans_time[j] = a.trans_time[j];
    m.rotat_acc[j] = V3<T>(a.rotat_acc[j][0], a.rotat_acc[j][1], a.rotat_acc[j][2]);
    m.rotat_time[j] = a.rotat_time[j];
  }
  m.is_rotation = a.is_rotation; m.pattern = a.pattern; m.table_size_rotat = a.table_size_rotat;
  m.pos_rp = V3<T>(a.pos_rp[0], a.pos_rp[1], a.pos_rp[2]);
  m.rotat_vel = V3<T>(a.rotat_vel[0], a.rotat_vel[1], a.rotat_vel[2]);
}
template <class T> static void tab_out(const MotionTab<T>& m, lq_motion_table& a) {
  for (int j = 0; j < LQ_TABLE_MAX; j++) {
    a.trans_vel[j][0] = float(m.trans_vel[j].x); a.trans_vel[j][1] = float(m.trans_vel[j].y);
    a.trans_vel[j][2] = float(m.trans_vel[j].z);
  }
  a.pos_rp[0] = float(m.pos_rp.x); a.pos_rp[1] = float(m.pos_rp.y); a.pos_rp[2] = float(m.pos_rp.z);
  a.rotat_vel[0] = float(m.rotat_vel.x); a.rotat_vel[1] = float(m.rotat_vel.y); a.rotat_vel[2] = float(m.rotat_vel.z);
}
template <class T> static void par_out(const Par<T>& P, lq_params* p) {
  p->npm = P.npm; p->np_delete_m = P.np_delete_m;
  p->time = float(P.time); p->dt = float(P.dt);
  for (int k = 0; k < LQ_RIGID_MOTION_MAX; k++) {
    lq_rigid_motion& m = p->mv[k];
    const Motion<T>& o = P.mv[k];
    m.trans_vel[0] = float(o.trans_vel.x); m.trans_vel[1] = float(o.trans_vel.y); m.trans_vel[2] = float(o.trans_vel.z);
    m.trans_acc[0] = float(o.trans_acc.x); m.trans_acc[1] = float(o.trans_acc.y); m.trans_acc[2] = float(o.trans_acc.z);
    m.pos_rp[0] = float(o.pos_rp.x); m.pos_rp[1] = float(o.pos_rp.y); m.pos_rp[2] = float(o.pos_rp.z);
    m.rotat_vel[0] = float(o.rotat_vel.x); m.rotat_vel[1] = float(o.rotat_vel.y); m.rotat_vel[2] = float(o.rotat_vel.z);
    m.rotat_acc[0] = float(o.rotat_acc.x); m.rotat_acc[1] = float(o.rotat_acc.y); m.rotat_acc[2] = float(o.rotat_acc.z);
  }
}

// One whole step (loquat.cu:366-436, 548-560).  Returns status.
template <class T> static int step_T(Par<T>& P, MotionTab<T>* M, State<T>& S, double& time, int64_t& step,
                                     int* ns2_identity) {
  T dtm = P.h / P.cs;  // StepCalculation, loquat.cu:368-370
  P.dt = P.cfl * dtm;
  P.time = T(time);
  int st = neighbor_search(P, S, nullptr);
  if (st) return st;
  sort_arrays_primary(P, S);
  if (P.n_motion > 0) rigid_body_update(P, M);
  adami(P, S);
  find_corner(P, S);
  if (P.shepard != 0 && step % P.shepard == 0 && step != 0) shepard(P, S);
  interaction(P, S);
  predictor(P, S);
  int ident = 1;
  st = neighbor_search(P, S, &ident);
  if (st) return st;
  if (ns2_identity) *ns2_identity = ident;
  sort_arrays_all(P, S);
  adami(P, S);
  find_corner(P, S);
  interaction(P, S);
  if (P.n_motion > 0) corrector_bound(P, S);
  corrector_material(P, S);
  time += double(P.dt);
  step++;
  return 0;
}

#define OR_DEFINE(T, SFX)                                                                              \
  extern "C" int or_abs2rel_##SFX(const lq_params* p, or_state* s) {                                  \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); abs2rel(P, p, S); return 0;                      \
  }                                                                                                    \
  extern "C" int or_neighbor_search_##SFX(lq_params* p, or_state* s, int* identity) {                 \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s);                                                 \
    int st = neighbor_search(P, S, identity); p->npm = P.npm; p->np_delete_m = P.np_delete_m; return st; \
  }                                                                                                    \
  extern "C" int or_sort_arrays_primary_##SFX(const lq_params* p, or_state* s) {                      \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); sort_arrays_primary(P, S); return 0;             \
  }                                                                                                    \
  extern "C" int or_sort_arrays_all_##SFX(const lq_params* p, or_state* s) {                          \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); sort_arrays_all(P, S); return 0;                 \
  }                                                                                                    \
  extern "C" int or_rigid_body_update_##SFX(lq_params* p, lq_motion_table* tabs) {                    \
    Par<T> P; load_par(p, P); MotionTab<T> M[LQ_RIGID_MOTION_MAX];                                     \
    for (int k = 0; k < p->n_motion; k++) tab_in(tabs[k], M[k]);                                       \
    rigid_body_update(P, M);                                                                           \
    for (int k = 0; k < p->n_motion; k++) tab_out(M[k], tabs[k]);                                      \
    par_out(P, p); return 0;                                                                           \
  }                                                                                                    \
  extern "C" int or_adami_##SFX(const lq_params* p, or_state* s) {                                    \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); adami(P, S); return 0;                           \
  }                                                                                                    \
  extern "C" int or_find_corner_##SFX(const lq_params* p, or_state* s) {                              \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); find_corner(P, S); return 0;                    \
  }                                                                                                    \
  extern "C" int or_shepard_##SFX(const lq_params* p, or_state* s) {                                  \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); shepard(P, S); return 0;                         \
  }                                                                                                    \
  extern "C" int or_interaction_##SFX(const lq_params* p, or_state* s) {                              \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); interaction(P, S); return 0;                     \
  }                                                                                                    \
  extern "C" int or_predictor_##SFX(const lq_params* p, or_state* s) {                                \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); predictor(P, S); return 0;                       \
  }                                                                                                    \
  extern "C" int or_corrector_##SFX(const lq_params* p, or_state* s) {                                \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s);                                                 \
    if (P.n_motion > 0) corrector_bound(P, S);                                                         \
    return corrector_material(P, S);                                                                   \
  }                                                                                                    \
  extern "C" int or_hypo_cs_##SFX(T dt, T rho, const T* w, const T* D, T* sig, const T* mp, T* mcrt) { \
    sym6<T> s = {sig[0], sig[1], sig[2], sig[3], sig[4], sig[5]};                                      \
    sym6<T> d = {D[0], D[1], D[2], D[3], D[4], D[5]};                                                  \
    T m = *mcrt;                                                                                       \
    hypo_cs(dt, rho, V3<T>(w[0], w[1], w[2]), d, s, mp, m);                                            \
    sig[0] = s.xx; sig[1] = s.yy; sig[2] = s.zz; sig[3] = s.xy; sig[4] = s.xz; sig[5] = s.yz;          \
    *mcrt = m; return 0;                                                                               \
  }                                                                                                    \
  extern "C" int or_constitutive_##SFX(int model_type, T dt, T rho, const T* w, const T* D, T* sig,   \
                                       const T* mp, T* mcrt, int* yld) {                               \
    sym6<T> s = {sig[0], sig[1], sig[2], sig[3], sig[4], sig[5]};                                      \
    sym6<T> d = {D[0], D[1], D[2], D[3], D[4], D[5]};                                                  \
    T m = *mcrt; bool y = false;                                                                       \
    constitutive(model_type, dt, rho, V3<T>(w[0], w[1], w[2]), d, s, mp, m, y);                        \
    sig[0] = s.xx; sig[1] = s.yy; sig[2] = s.zz; sig[3] = s.xy; sig[4] = s.xz; sig[5] = s.yz;          \
    *mcrt = m; *yld = y ? 1 : 0; return 0;                                                             \
  }                                                                                                    \
  extern "C" double or_wall_force_##SFX(const lq_params* p, or_state* s, const float* box) {          \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); return wall_force(P, S, box);                   \
  }                                                                                                    \
  extern "C" int or_wall_profile_##SFX(const lq_params* p, or_state* s, const float* geom, int nbins,  \
                                       double* out) {                                                  \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); wall_profile(P, S, geom, nbins, out); return 0; \
  }                                                                                                    \
  extern "C" int or_sensor_stress_##SFX(const lq_params* p, or_state* s, const float* pts, int n, int rows,     \
                                        double* out) {                                                 \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); sensor_stress(P, S, pts, n, rows, out); return 0; \
  }                                                                                                    \
  extern "C" int or_surface_heights_##SFX(const lq_params* p, or_state* s, const float* pts, int n,     \
                                          float radius, double* out) {                                 \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); surface_heights(P, S, pts, n, radius, out);      \
    return 0;                                                                                          \
  }                                                                                                    \
  extern "C" int or_run_##SFX(lq_params* p, lq_motion_table* tabs, or_state* s, int nsteps,          \
                              double* time, int64_t* step, const float* box, int force_every,          \
                              double* forces, int* ns2_identity_all, double* seconds) {                \
    Par<T> P; load_par(p, P); State<T> S = view<T>(s); MotionTab<T> M[LQ_RIGID_MOTION_MAX];            \
    for (int k = 0; k < p->n_motion; k++) tab_in(tabs[k], M[k]);                                       \
    int all_ident = 1; int nf = 0;                                                                     \
    auto t0 = std::chrono::steady_clock::now();                                                        \
    for (int n = 0; n < nsteps; n++) {                                                                 \
      int ident = 1;                                                                                   \
      int st = step_T(P, M, S, *time, *step, &ident);                                                  \
      if (st) { par_out(P, p); return st; }                                                            \
      all_ident &= ident;                                                                              \
      if (forces && force_every > 0 && (n + 1) % force_every == 0) forces[nf++] = wall_force(P, S, box); \
    }                                                                                                  \
    auto t1 = std::chrono::steady_clock::now();                                                        \
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();                             \
    if (ns2_identity_all) *ns2_identity_all = all_ident;                                               \
    for (int k = 0; k < p->n_motion; k++) tab_out(M[k], tabs[k]);                                      \
    par_out(P, p); return 0;                                                                           \
  }

OR_DEFINE(float, f32)
OR_DEFINE(double, f64)

// Neighbour lists in traversal order (particle_interaction.cu:117-140): for
// soil slot npb+s, the orig_id of every j with q < 2.
template <class T> static int neighbor_lists_T(const lq_params* p, or_state* s, int max_nb, int32_t* counts,
                                               int32_t* ids) {
  Par<T> P; load_par(p, P); State<T> S = view<T>(s);
  for (int idx = 0; idx < P.npm; idx++) {
    int i = P.npb + idx, n = 0;
    v3<T> pi = ld3(S.pos, i);
    v3<int> ci = cell_of(S.cellpos, i);
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          v3<int> off = {cx, cy, cz};
          v3<int> cj = {ci.x + cx, ci.y + cy, ci.z + cz};
          if (!in_grid(P, cj)) continue;
          int cb = linear_key(P, cj);
          for (int pass = 0; pass < 2; pass++) {
            int c = pass ? cb + P.ncell_material : cb;
            for (int j = S.cell_beg[c]; j < S.cell_end[c]; j++) {
              v3<T> d;
              T r = pair_r(pi, ld3(S.pos, j), off, P.two_h, d);
              if (!(r * P.i_h < T(2))) continue;
              if (n < max_nb) ids[size_t(idx) * max_nb + n] = S.orig_id[j];
              n++;
            }
          }
        }
    counts[idx] = n;
  }
  return 0;
}
extern "C" int or_neighbor_lists_f32(const lq_params* p, or_state* s, int max_nb, int32_t* counts, int32_t* ids) {
  return neighbor_lists_T<float>(p, s, max_nb, counts, ids);
}
extern "C" int or_neighbor_lists_f64(const lq_params* p, or_state* s, int max_nb, int32_t* counts, int32_t* ids) {
  return neighbor_lists_T<double>(p, s, max_nb, counts, ids);
}

// Threads for the per-particle loops (results do not depend on it: every
// parallel iteration writes only its own particle's rows).
extern "C" void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
extern "C" void or_set_surcharge(int block, double pressure) {
  g_sur_block = block < 0 ? -1 : block;
  g_sur_pressure = pressure;
}
extern "C" int or_get_threads() { return omp_get_max_threads(); }

// Wendland C2 value and gradient factor (functions.h:142-143) for tests.
extern "C" double or_kernel_f64(double knorm, double q) { return kernelW<double>(knorm, q); }
extern "C" double or_grad_f64(double kgrad, double q) { return kernelG<double>(kgrad, q); }
