// lq_internal.hpp -- device-side data model of the gfx950 SPH step.
//
// Layout in HBM (structure of arrays, one row per sorted particle slot):
//   float4 pos      relative x y z (to the 2h cell), massflag (mass of a soil
//                   particle; negative for contact-wall codes)
//   float  pres     the .dat "p" column (carried for output only)
//   float4 vel      vx vy vz rho
//   float4 idm      id mass idx M_crt
//   int4   codes    block id, Fix/NonFix, model, motion id
//   int4   cellpos  cx cy cz key
//   float4 str_a/stn_a (xx yy zz xy) + float2 str_b/stn_b (xz yz)
//   float4 cvar     K d slip wall-id (contact_var)
//   float4 cf/cn/cw contact force / contact normal / wall velocity (contact_mag rows)
// All per-step scalars (time, dt, current wall motion, npm) live in one
// device-resident DevParams block so a whole step runs without host
// round trips and can be captured in a hipGraph.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "retainwall.h"

namespace lq {

struct DevMotion {  // RigidMotion, types.h:66-76
  int is_trans, is_rotat, pattern, pad;
  float3 trans_vel, trans_acc, pos_rp, rotat_vel, rotat_acc;
};

struct DevTable {  // MotionTable, types.h:81-97 (device copy, mutated each step)
  int is_translation, table_size_trans, is_rotation, pattern, table_size_rotat, pad0, pad1, pad2;
  float3 trans_vel[LQ_TABLE_MAX];
  float3 trans_acc[LQ_TABLE_MAX];
  float trans_time[LQ_TABLE_MAX];
  float3 pos_rp, rotat_vel;
  float3 rotat_acc[LQ_TABLE_MAX];
  float rotat_time[LQ_TABLE_MAX];
};

// Per-material HypoCS constants hoisted out of constitutive_model.h:226-230.
struct HypoConst {
  float C1, C2, C3, C4, fa, fb, fc, sq2sin;
};

// One case of a batched context.  Cases sit in one global cell grid as
// cell-aligned blocks separated by at least one empty cell, so no candidate
// range ever reaches another case.  off = the case's first cell, n = its own
// grid (out-of-grid deletion and the domain check use it), and its
// materials and motions start at model_base / motion_base in the tables.
struct DevCase {
  int ox, oy, oz, nx, ny, nz, model_base, motion_base;
};

struct DevParams {
  // ---- static (set at create) ----
  int np, npb, cnx, cny, cnz, ncell, ncell_material, ncell_delete_m;
  int key_mode, shepard, viscosity_type, n_motion, n_model, nbits, periodbound, surcharge_block;
  float h, i_h, dr, vol_def, bt, eta, two_h, r2cut;
  float gx, gy, gz, damping, knorm, kgrad, kernel_zero, kernel_dr;
  float art_exp, alpha_cs, art_p, cfl, cs, band, half_dr, mu;
  // pair lists: squared list radius (cut-off plus skin), largest squared
  // displacement between build and reuse that keeps them exact, capacity
  // (entries per soil slot) and column stride (soil slots)
  float r2list, disp2_lim;
  int plist_cap, plist_stride;
  int n_cases, mbits, pad_c1, pad_c2;  // mbits: Morton bits per axis (LQ_KEY_MORTON)
  float4 plane_l, plane_r;  // periodic planes a x + b y + c z + d = 0 (planeleft, planeright)
  float surcharge_force;  // downward force on each surcharge-block slot (-1 block: none)
  // COLUMNS interaction: squared search radius (the cut-off with a margin)
  // and the distance margin of its sub-column windows
  float col_r2, col_margin, pad_s2;
  const float (*model_par)[LQ_MATERIAL_SIZE_MAX];  // [n_model] material rows (device table)
  const HypoConst* hypo;                           // [n_model]
  const DevCase* cases;                            // [n_cases]
  // ---- dynamic (device-updated) ----
  int npm, np_delete_m, err, n_deferred;
  int sort_changed, pad2, pad3, pad4;
  unsigned max_disp2_bits;  // largest squared soil displacement since the pair-list build (float bits)
  int plist_bad;            // a soil slot overflowed its pair-list capacity
  int pad5, pad6;
  // HypoCS surrogate (MLP) state
  int sur_on, sur_members, log_every, log_cap;
  float sur_ens_tol, sur_audit_frac, pad7, pad8;
  const float* sur_norm;        // x_mean[14] x_std[14] y_mean[7] y_std[7]
  const float* sur_lo;          // [14] guard box of the raw inputs
  const float* sur_hi;          // [14]
  unsigned long long* sur_stats;  // [slots][8]: surrogate, fallback range, fallback ensemble, audited, I1 branch
  unsigned* sur_audit_max;      // [slots][16]: largest audited normalised error (float bits) in [q][0]
  float* log_x;                 // [log_cap][16] logged raw inputs
  float* log_y;                 // [log_cap][8]  logged exact targets
  unsigned* log_count;
  float time, dt;
  double time_d;
  long long step;
  DevMotion* mv;  // [n_motion] current imposed motion of each body (device table)
  DevTable* tab;  // [n_motion] motion tables, advanced on the device
};

// Device error codes written to DevParams::err (first error wins).
enum { DERR_NONE = 0, DERR_DELETE = 1, DERR_DOMAIN = 2 };

struct Arrays {
  int4 *codes, *cellpos;
  int* orig_id;
  int* case_id;  // case of the particle (batched contexts; 0 otherwise)
  float* pres;
  float4 *pos, *pos_ori, *vel, *idm, *normal;
  float4 *str_a, *stn_a;
  float2 *str_b, *stn_b;
  float4 *cvar, *cf, *cn, *cw;
};

struct Buffers {
  Arrays cur;   // current state
  Arrays alt;   // scratch rows of the in-place sort; SortArraysAll gathers into it and swaps
  float4 *pos_ini, *vel_ini, *str_ini_a, *stn_ini_a;
  float2 *str_ini_b, *stn_ini_b;
  float4 *accad, *spin, *acestn_a;
  float2* acestn_b;
  float4* s3;        // per-soil partial acceleration kept for deferred contact
  float4* fdef;      // pass-B contact force scratch
  int* deferred;     // compact list of deferred soil slots
  float4* str_tmp_a; // Shepard output scratch
  float2* str_tmp_b;
  // Per-slot neighbour record written by k_pair_prep, 64 bytes so a pair's
  // four loads share one cache line: pos, vel, then sigma/rho^2 as
  // (xx yy zz xy) and (xz yz, m/rho, trace).  The COLUMNS variant stores
  // the same four float4 per aux position as four planes of rec_rows
  // entries (plane q at rec + q * rec_rows), so that lanes reading nearby
  // aux positions share cache lines.
  float4* rec;
  int rec_rows;
  float* sur_x;      // [npm][16] surrogate inputs (normalised), per soil slot
  float* sur_y;      // [members][npm][8] surrogate outputs (normalised)
  unsigned short* sur_w;  // packed bf16 weights per member (MFMA operand layout)
  float* sur_b;      // biases per member
  int* plist;        // pair lists, entry k of soil slot s at plist[k * stride + s]
  int* pcount;       // pair-list length per soil slot
  // grid / sort
  int *keys, *vals, *keys_alt, *vals_alt, *part_cel, *part_idx;
  // merge search: sorted keys of the last two searches (part_cel is one of
  // them), mover flags scanned in place (np + 1), compacted movers
  int *skey[2], *mflag, *mkey, *mslot, *mkey_alt, *mslot_alt;
  int *cell_beg, *cell_end;
  int* wall_near;    // per cell: a wall particle lies in its 27-cell neighbourhood
  // COLUMNS interaction: sub-cell index rebuilt per call (every particle as
  // (relative position, slot << 2 | cz & 3), grouped by 2h column,
  // sub-column and sub-cell height) and the start of every bucket; in this
  // variant rec is indexed by aux position, not by slot
  float4* aux;
  int* soil_aux;   // aux positions of the soil entries, in aux order
  int* wall_aux;   // aux position of each wall slot (when its record is written after Adami)
  int* sub_start;
  float4* fl_pos;  // fused steps: position and cell of first-layer particles before their update
  int4* fl_cell;
  int2* colcnt;   // per 2h column: wall and soil counts
  int4* colinfo;  // per 2h column: first wall slot, first soil slot, counts
  int2* coltile;  // per 1024 columns: wall and soil totals, then their exclusive scan
  int* hist;         // radix histogram [256][nblocks]
  unsigned long long* wf_acc;  // per case: exact wall-force sum, 128-bit fixed point (lo, hi)
};

// sigma/rho^2 parts of slot j's neighbour record
__device__ __forceinline__ float4 rec_sra(const Buffers& B, int j) { return B.rec[4 * size_t(j) + 2]; }
__device__ __forceinline__ float4 rec_srb(const Buffers& B, int j) { return B.rec[4 * size_t(j) + 3]; }

// ------------------------------------------------------------- helpers ----
__device__ __forceinline__ float kW(float knorm, float q) {  // functions.h:142
  float t = 2.0f - q;
  return knorm * t * t * t * t * (q + 0.5f);
}
__device__ __forceinline__ float kG(float kgrad, float q) {  // functions.h:143
  float t = 2.0f - q;
  return kgrad * t * t * t;
}
// Bits of v (0 <= v < 1024) moved to every third bit position.
__device__ __forceinline__ int spread3(int v) {
  unsigned x = unsigned(v) & 0x3ffu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return int(x);
}
// The inverse of spread3 on every third bit of k.
__device__ __forceinline__ int compact3(int k) {
  unsigned x = unsigned(k) & 0x09249249u;
  x = (x | (x >> 2)) & 0x030C30C3u;
  x = (x | (x >> 4)) & 0x0300F00Fu;
  x = (x | (x >> 8)) & 0x030000FFu;
  x = (x | (x >> 16)) & 0x3ffu;
  return int(x);
}
// Key of cell (cx, cy, cz) of the grid (the soil layer adds ncell_material).
// LQ_KEY_LINEAR: cx * ny * nz + cy * nz + cz (loquat_grid.cu:88);
// LQ_KEY_MORTON: the cell coordinates' bits interleaved (x y z from the top,
// z lowest), on the cube of 2^mbits cells per axis.
__device__ __forceinline__ int grid_key(const DevParams* P, int cx, int cy, int cz) {
  if (P->key_mode == LQ_KEY_MORTON) return (spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz);
  return cx * P->cny * P->cnz + cy * P->cnz + cz;
}
__device__ __forceinline__ bool in_grid(const DevParams* P, int cx, int cy, int cz) {
  return cx >= 0 && cy >= 0 && cz >= 0 && cx < P->cnx && cy < P->cny && cz < P->cnz;
}
__device__ __forceinline__ void set_err(DevParams* P, int code) { atomicCAS(&P->err, 0, code); }

// Gravity plus, on slots of the surcharge block, the strip load's force
// over the slot's mass (a body force like gravity; no reference
// counterpart, the load of the large sloping-backfill workload).
__device__ __forceinline__ float3 body_acc(const DevParams* P, const int4* codes, int i, float mi) {
  float3 g = make_float3(P->gx, P->gy, P->gz);
  if (P->surcharge_block >= 0 && codes[i].x == P->surcharge_block) g.z = g.z - P->surcharge_force / mi;
  return g;
}

// ------------------------------------------------ periodic boundaries ----
// Dist2Plane (functions.h:134-136): signed distance to a x + b y + c z + d = 0.
__device__ __forceinline__ float plane_dist(float3 x, float4 pl) {
  return (x.x * pl.x + x.y * pl.y + x.z * pl.z + pl.w) / sqrtf(pl.x * pl.x + pl.y * pl.y + pl.z * pl.z);
}
// Cell and relative position of an absolute position (int3(x 0.5/h)
// truncated, x - cell 2h), single-case grids.
__device__ __forceinline__ void to_cell(const DevParams* P, float3 a, int4& c, float3& p) {
  const float scale = 0.5f * P->i_h, two_h = P->two_h;
  c.x = int(a.x * scale); c.y = int(a.y * scale); c.z = int(a.z * scale);
  p = make_float3(a.x - float(c.x) * two_h, a.y - float(c.y) * two_h, a.z - float(c.z) * two_h);
}
__device__ __forceinline__ float3 abs_from_cell(const DevParams* P, int4 c, float3 p) {
  const float two_h = P->two_h;
  return make_float3(p.x + float(c.x) * two_h, p.y + float(c.y) * two_h, p.z + float(c.z) * two_h);
}
// The periodic image of the *LoopPeriod helpers (particle_interaction.cu:
// 243-258, adami_boundary.cu:73-87, field_smoothing.cu:58-72): a slot closer
// than 2h to either plane is shifted by the period along that plane's
// normal (left first) and its 27-cell loop runs again from there.  The
// interaction takes d_l + d_r as the period, Adami and Shepard |d_l + d_r|.
__device__ __forceinline__ bool period_image(const DevParams* P, int4 c, float3 p, bool abs_len, int4& cpb,
                                             float3& ppb) {
  const float3 a = abs_from_cell(P, c, p);
  const bool left = fabsf(plane_dist(a, P->plane_l)) < P->two_h;
  const bool right = fabsf(plane_dist(a, P->plane_r)) < P->two_h;
  if (!left && !right) return false;
  float len = P->plane_l.w + P->plane_r.w;
  if (abs_len) len = fabsf(len);
  const float4 n = left ? P->plane_l : P->plane_r;
  to_cell(P, make_float3(a.x + n.x * len, a.y + n.y * len, a.z + n.z * len), cpb, ppb);
  return true;
}
// cuCheckOutDomainPeriod (predictor_corrector.cu:26-40): a slot behind
// either plane re-enters through the other one (cell.w kept).
__device__ __forceinline__ void wrap_period(const DevParams* P, int4& c, float3& p) {
  const float3 a = abs_from_cell(P, c, p);
  const bool left = plane_dist(a, P->plane_l) < 0.f;
  const bool right = plane_dist(a, P->plane_r) < 0.f;
  if (!left && !right) return;
  const float len = fabsf(P->plane_l.w + P->plane_r.w);
  const float4 n = left ? P->plane_l : P->plane_r;
  to_cell(P, make_float3(a.x + n.x * len, a.y + n.y * len, a.z + n.z * len), c, p);
}
// Cell coordinates of a particle in its own case's grid.
__device__ __forceinline__ int3 case_cell(const DevParams* P, const int* case_id, int i, int4 c) {
  const DevCase& k = P->cases[case_id[i]];
  return make_int3(c.x - k.ox, c.y - k.oy, c.z - k.oz);
}
__device__ __forceinline__ bool in_case_grid(const DevParams* P, const int* case_id, int i, int4 c) {
  const DevCase& k = P->cases[case_id[i]];
  const int x = c.x - k.ox, y = c.y - k.oy, z = c.z - k.oz;
  return x >= 0 && y >= 0 && z >= 0 && x < k.nx && y < k.ny && z < k.nz;
}

}  // namespace lq
