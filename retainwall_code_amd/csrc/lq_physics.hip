// lq_physics.hip -- per-particle SPH kernels of the retaining-wall step.
//
// Built with -ffp-contract=off: every fused multiply-add below is written
// out (fmaf) so that each operation sequence is fixed and reproducible.
#include "lq_internal.hpp"
#include "lq_launch.hpp"
#include "lq_device.hpp"

namespace lq {

// ------------------------------------------------------------ frames ----
// cuPositionFromAbsolute2Relative, loquat_grid.cu:253-261, for slots
// [first, first + n): the cell and the relative position are taken in the
// particle's own case frame, then the cell is moved to the case's block.
__global__ void k_abs2rel(const DevParams* __restrict__ P, Arrays A, float scale, int first, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = first + k;
  float4 p = A.pos[i];
  int cx = int(p.x * scale), cy = int(p.y * scale), cz = int(p.z * scale);
  A.pos[i] = make_float4(p.x - float(cx) * P->two_h, p.y - float(cy) * P->two_h, p.z - float(cz) * P->two_h, p.w);
  const DevCase& c = P->cases[A.case_id[i]];
  A.cellpos[i] = make_int4(cx + c.ox, cy + c.oy, cz + c.oz, 0);
}

// pos.w carries the neighbour data the interaction needs with the position:
// the mass of a soil particle (idmassidx.y) and, in the sign, whether the
// particle's code marks a contact wall (codes.x >= 10).
__global__ void k_set_massflag(const DevParams* __restrict__ P, Arrays A, int first, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = first + k;
  float4 p = A.pos[i];
  const float4 id = A.idm[i];
  const float m = id.x >= 100.0f ? id.y : 1.0f;
  p.w = A.codes[i].x >= 10 ? -m : m;
  A.pos[i] = p;
}

__global__ void k_copy_w(const DevParams* __restrict__ P, const float4* __restrict__ src, float4* __restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P->np) return;
  float4 d = dst[i];
  d.w = src[i].w;
  dst[i] = d;
}

// cuPositionFromRelative2Absolute, loquat_grid.cu:273-279.
__global__ void k_rel2abs(const DevParams* __restrict__ P, Arrays A, float4* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P->np) return;
  float4 p = A.pos[i];
  const int3 c = case_cell(P, A.case_id, i, A.cellpos[i]);
  out[i] = make_float4(p.x + float(c.x) * P->two_h, p.y + float(c.y) * P->two_h, p.z + float(c.z) * P->two_h,
                       A.pres[i]);
}

// ------------------------------------------------------ step scalars ----
// StepCalculation, loquat.cu:366-374 (on device).
__global__ void k_step_calc(DevParams* P) {
  float dtm = P->h / P->cs;
  P->dt = P->cfl * dtm;
  P->time = float(P->time_d);
}
// loquat.cu:558-560: time += dt; step++.
__global__ void k_advance(DevParams* P) {
  P->time_d += double(P->dt);
  P->step += 1;
}
// Clears the deferred-contact counter before each interaction; a pair-list
// build also clears the list's overflow flag and the displacement maximum.
__global__ void k_reset_deferred(DevParams* P, int list_build) {
  P->n_deferred = 0;
  if (list_build) {
    P->plist_bad = 0;
    P->max_disp2_bits = 0u;
  }
}


// RigidBodyUpdate, adami_boundary.cu:175-203, for body i at time t.
__device__ void rigid_body_one(DevParams* P, int i, float t, float dt) {
  {
    DevTable& m = P->tab[i];
    DevMotion& mv = P->mv[i];
    if (m.is_rotation) {
      for (int j = 0; j < m.table_size_rotat; j++) {
        if (t < m.rotat_time[j]) {
          m.rotat_vel = f3add(m.rotat_vel, f3mul(m.rotat_acc[j], dt));
          float3 vel = f3add(m.trans_vel[j], f3mul(m.trans_acc[j], 0.5f * dt));
          m.pos_rp = f3add(m.pos_rp, f3mul(vel, dt));
          mv.rotat_acc = m.rotat_acc[j];
          mv.rotat_vel = m.rotat_vel;
          mv.pos_rp = m.pos_rp;
          break;
        }
      }
    }
    if (m.is_translation) {
      for (int j = 0; j < m.table_size_trans; j++) {
        if (t < m.trans_time[j]) {
          m.trans_vel[j] = f3add(m.trans_vel[j], f3mul(m.trans_acc[j], dt));
          mv.trans_acc = m.trans_acc[j];
          mv.trans_vel = m.trans_vel[j];
          break;
        }
      }
    }
  }
}
// one thread per body: each body touches only its own table and motion
__global__ void k_rigid_body_update(DevParams* P) {
  const float t = P->time, dt = P->dt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P->n_motion; i += gridDim.x * blockDim.x)
    rigid_body_one(P, i, t, dt);
}
// k_step_calc, the search's sort_changed clear and k_rigid_body_update in
// one launch: every thread forms dt and time as k_step_calc does, thread 0
// stores them.
__global__ void k_step_begin(DevParams* P) {
  const float dtm = P->h / P->cs;
  const float dt = P->cfl * dtm, t = float(P->time_d);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P->dt = dt;
    P->time = t;
    P->sort_changed = 0;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P->n_motion; i += gridDim.x * blockDim.x)
    rigid_body_one(P, i, t, dt);
}

// ------------------------------------------------------------- Adami ----
// cuAdamiBoundaryMaterial + loops, adami_boundary.cu:26-158.
// Each thread walks the 27 cells in reference order.  The soil extrapolation
// sums and, for contact walls, the wall-normal integral share no inputs or
// outputs.  With split set (few wall slots: about one wave per SIMD, so each
// wave's chain of dependent loads sets the time) they run on two threads,
// [0, npb) the soil terms of slot t and [npb, 2 npb) the normal of slot
// t - npb, halving a contact-wall thread's chain; with many wall slots one
// thread does both and each cell's ranges are read once.  Results are the
// same either way.  The next cell's ranges are fetched while the current
// one is scanned.
__global__ void __launch_bounds__(64) k_adami(const DevParams* __restrict__ P, Arrays A,
                                               const int* __restrict__ cbeg, const int* __restrict__ cend, int remap,
                                               int split) {
  const int t = xcd_block(remap) * blockDim.x + threadIdx.x;
  const int npb = P->npb;
  if (t >= (split ? 2 * npb : npb)) return;
  const int i = t >= npb ? t - npb : t;
  const bool contact = A.codes[i].x >= 10;
  const bool do_soil = !split || t < npb;
  const bool do_normal = contact && (!split || t >= npb);
  if (!do_soil && !do_normal) return;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  const int ncm = P->ncell_material;
  float3 n = make_float3(0.f, 0.f, 0.f);
  Sym rs = {0, 0, 0, 0, 0, 0};
  float3 rpos = make_float3(0.f, 0.f, 0.f);
  float rker = 0.f;
  float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
  // the 27-cell loop around (ci, pi); the periodic image pass
  // (adami_boundary.cu:73-87, 103) runs it again from the shifted position
  auto scan = [&](int4 ci, float3 pi) {
    int bw = 0, ew = 0, bs = 0, es = 0;  // wall and soil ranges of cell k, fetched one cell ahead
    if (in_grid(P, ci.x - 1, ci.y - 1, ci.z - 1)) {
      const int c = grid_key(P, ci.x - 1, ci.y - 1, ci.z - 1);
      if (do_normal) { bw = cbeg[c]; ew = cend[c]; }
      if (do_soil) { bs = cbeg[c + ncm]; es = cend[c + ncm]; }
    }
#pragma unroll 1
    for (int k = 0; k < 27; k++) {
      const int cz = k / 9 - 1, cy = (k / 3) % 3 - 1, cx = k % 3 - 1;
      int bwn = 0, ewn = 0, bsn = 0, esn = 0;
      if (k < 26) {
        const int kn = k + 1;
        const int nx = ci.x + kn % 3 - 1, ny = ci.y + (kn / 3) % 3 - 1, nz = ci.z + kn / 9 - 1;
        if (in_grid(P, nx, ny, nz)) {
          const int c = grid_key(P, nx, ny, nz);
          if (do_normal) { bwn = cbeg[c]; ewn = cend[c]; }
          if (do_soil) { bsn = cbeg[c + ncm]; esn = cend[c + ncm]; }
        }
      }
      for (int j0 = bw; j0 < ew; j0 += 32) {  // normal integral over wall neighbours
        unsigned mask = hit_mask(A.pos, pi, j0, min(ew - j0, 32), cx, cy, cz, two_h, r2cut);
        while (mask) {
          const int j = j0 + __ffs(mask) - 1;
          mask &= mask - 1u;
          float3 d;
          float r = sqrtf(pair_r2(pi, A.pos[j], cx, cy, cz, two_h, d));
          float q = r * i_h;
          float dwdr = kG(P->kgrad, q);
          float3 g = f3mul(d, dwdr);
          n = f3sub(n, f3mul(g, P->vol_def));
        }
      }
      for (int j0 = bs; j0 < es; j0 += 32) {  // extrapolation from soil neighbours
        unsigned mask = hit_mask(A.pos, pi, j0, min(es - j0, 32), cx, cy, cz, two_h, r2cut);
        while (mask) {
          const int j = j0 + __ffs(mask) - 1;
          mask &= mask - 1u;
          float3 d;
          float r = sqrtf(pair_r2(pi, A.pos[j], cx, cy, cz, two_h, d));
          float q = r * i_h;
          float ker = kW(P->knorm, q);
          float4 idmj = A.idm[j];
          float4 vj = A.vel[j];
          float volj = idmj.y / vj.w;
          float wv = ker * volj;
          Sym sj = ldsym(A.str_a, A.str_b, j);
          rs = sadd(rs, smul(sj, wv));
          rpos = f3add(rpos, f3mul(d, ker * idmj.y));
          rker = rker + wv;
          rv = make_float4(rv.x + vj.x * wv, rv.y + vj.y * wv, rv.z + vj.z * wv, rv.w + vj.w * wv);
        }
      }
      bw = bwn; ew = ewn; bs = bsn; es = esn;
    }
  };
  scan(ci, pi);
  int4 cpb;
  float3 ppb;
  if (P->periodbound && period_image(P, ci, pi, true, cpb, ppb)) scan(cpb, ppb);
  if (do_normal) {
    float nn = f3norm(n);
    if (nn > 0.f) A.cn[i] = make_float4(n.x / nn, n.y / nn, n.z / nn, 0.f);
  }
  if (!do_soil) return;
  if (rker > 0.f) {
    float4 nrm = A.normal[i];
    float fc = nrm.w;
    float3 ni = ld3(nrm);
    float3 vref = f3div(make_float3(rv.x, rv.y, rv.z), rker);
    float3 vref_n = f3mul(ni, f3dot(ni, vref));
    float3 vref_t = f3sub(vref, vref_n);
    float3 vrb = make_float3(0.f, 0.f, 0.f);
    int4 cd = A.codes[i];
    if (cd.y == LQ_NONFIX) {
      const DevMotion& m = P->mv[cd.w];
      vrb = m.trans_vel;
      if (m.is_rotat) {
        const int3 lc = case_cell(P, A.case_id, i, ci);
        float3 pa = make_float3(pi.x + float(lc.x) * two_h, pi.y + float(lc.y) * two_h, pi.z + float(lc.z) * two_h);
        vrb = f3add(vrb, f3cross(m.rotat_vel, f3sub(pa, m.pos_rp)));
      }
    }
    float3 vrb_n = f3mul(ni, f3dot(ni, vrb));
    float3 vrb_t = f3sub(vrb, vrb_n);
    float3 vext_n = f3sub(f3mul(vrb_n, 2.0f), vref_n);
    float3 vext_t = fc == -1.0f ? f3sub(f3mul(vrb_t, 2.0f), vref_t) : vref_t;
    float3 vext = f3add(vext_n, vext_t);
    float rho = rv.w / rker;
    if (contact) {
      A.vel[i] = make_float4(vrb.x, vrb.y, vrb.z, rho);
      A.cw[i] = make_float4(vext.x, vext.y, vext.z, 0.f);
    } else {
      A.vel[i] = make_float4(vext.x, vext.y, vext.z, rho);
    }
    Sym sref = sdiv(rs, rker);
    float3 gr = f3div(make_float3(P->gx * rpos.x, P->gy * rpos.y, P->gz * rpos.z), rker);
    sref.xx = sref.xx - gr.x; sref.yy = sref.yy - gr.y; sref.zz = sref.zz - gr.z;
    stsym(A.str_a, A.str_b, i, sref);
  } else {
    stsym(A.str_a, A.str_b, i, Sym{0, 0, 0, 0, 0, 0});
  }
}

// ------------------------------------------------------------ corner ----
// cuFindCornerParticleMaterial + loop, adami_boundary.cu:209-292.
// wall_near[c] != 0 when a wall particle lies in the 27 cells around cell
// c; elsewhere the scan finds nothing and only the reset of cvar.xy remains.
__global__ void __launch_bounds__(256) k_corner(const DevParams* __restrict__ P, Arrays A,
                                                const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                const int* __restrict__ wall_near, int remap) {
  int idx = xcd_block(remap) * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const int4 ci = A.cellpos[i];
  float4 cv = A.cvar[i];
  cv.x = 0.f;
  cv.y = 999.f;
  if (!wall_near[grid_key(P, ci.x, ci.y, ci.z)]) {
    A.cvar[i] = cv;
    return;
  }
  const float3 pi = ld3(A.pos[i]);
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  float4 cn = A.cn[i];
  float4 cw = A.cw[i];
  float dij = 999.f;
  int b = 0, e = 0;  // wall range of cell k, fetched one cell ahead
  if (in_grid(P, ci.x - 1, ci.y - 1, ci.z - 1)) {
    const int c = grid_key(P, ci.x - 1, ci.y - 1, ci.z - 1);
    b = cbeg[c];
    e = cend[c];
  }
#pragma unroll 1
  for (int k = 0; k < 27; k++) {
    const int cz = k / 9 - 1, cy = (k / 3) % 3 - 1, cx = k % 3 - 1;
    int bn = 0, en = 0;
    if (k < 26) {
      const int kn = k + 1;
      const int nx = ci.x + kn % 3 - 1, ny = ci.y + (kn / 3) % 3 - 1, nz = ci.z + kn / 9 - 1;
      if (in_grid(P, nx, ny, nz)) {
        const int c = grid_key(P, nx, ny, nz);
        bn = cbeg[c];
        en = cend[c];
      }
    }
    for (int j0 = b; j0 < e; j0 += 32) {
      unsigned mask = hit_mask(A.pos, pi, j0, min(e - j0, 32), cx, cy, cz, two_h, r2cut);
      while (mask) {
        const int j = j0 + __ffs(mask) - 1;
        mask &= mask - 1u;
        // all four loads of the hit are issued together, ahead of the branches
        const int cj = A.codes[j].x;
        const float4 pj = A.pos[j];
        const float4 cnj = A.cn[j];
        const float4 vj = A.vel[j];
        if (cj < 10) continue;
        float3 d;
        float r = sqrtf(pair_r2(pi, pj, cx, cy, cz, two_h, d));
        float q = r * i_h;
        float dwdr = kG(P->kgrad, q);
        float3 g = f3mul(d, dwdr);
        if (r < dij && r > 0.f) {
          float3 na = ld3(cnj);
          float3 nm = f3div(d, r);
          dij = r;
          cv.y = fabsf(f3dot(d, na));
          cv.w = float(cj);
          cw = make_float4(vj.x, vj.y, vj.z, 0.f);
          float cosc = fabsf(f3dot(nm, na));
          if (cosc < 0.866f) cn = make_float4(nm.x, nm.y, nm.z, 0.f);
          else cn = make_float4(na.x, na.y, na.z, 0.f);
        }
        cv.x = cv.x + fabsf((cn.x * g.x + cn.y * g.y + cn.z * g.z) * P->vol_def);
      }
    }
    b = bn;
    e = en;
  }
  A.cvar[i] = cv;
  A.cn[i] = cn;
  A.cw[i] = cw;
}

// ----------------------------------------------------------- Shepard ----
// The step driver's filter cadence (loquat.cu:548-561: every `shepard`
// steps, not at step 0), decided on the device so that a captured step
// graph carries the filter launches on every replay.
__device__ __forceinline__ bool shepard_due(const DevParams* __restrict__ P) {
  return P->shepard > 0 && P->step != 0 && P->step % P->shepard == 0;
}
// cuStressRegularizationShepard + loop, field_smoothing.cu:26-90.  due_only:
// run only at a step where the filter is due.
__global__ void __launch_bounds__(256) k_shepard(const DevParams* __restrict__ P, Arrays A,
                                                 const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                 float4* __restrict__ oa, float2* __restrict__ ob, int due_only) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm || (due_only && !shepard_due(P))) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float r2cut = P->r2cut;
  Sym acc = {0, 0, 0, 0, 0, 0};
  float sk = 0.f;
  auto scan = [&](int4 ci, float3 pi) {
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
          if (!in_grid(P, jx, jy, jz)) continue;
          int c = grid_key(P, jx, jy, jz) + P->ncell_material;
          for (int j = cbeg[c], e = cend[c]; j < e; j++) {
            float3 d;
            const float r2 = pair_r2(pi, A.pos[j], cx, cy, cz, P->two_h, d);
            if (!(r2 <= r2cut)) continue;
            float r = sqrtf(r2);
            float q = r * P->i_h;
            float ker = kW(P->knorm, q);
            float vol = A.idm[j].y / A.vel[j].w;
            float wv = ker * vol;
            sk = sk + wv;
            acc = sadd(acc, smul(ldsym(A.str_a, A.str_b, j), wv));
          }
        }
  };
  scan(ci, pi);
  int4 cpb;
  float3 ppb;
  if (P->periodbound && period_image(P, ci, pi, true, cpb, ppb)) scan(cpb, ppb);  // field_smoothing.cu:58-72, 85
  if (sk != 0.f) acc = sdiv(acc, sk);
  stsym(oa, ob, i, acc);
}

__global__ void k_copy_soil_str(const DevParams* __restrict__ P, Arrays A, const float4* __restrict__ oa,
                                const float2* __restrict__ ob, int due_only) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm || (due_only && !shepard_due(P))) return;
  int i = P->npb + idx;
  A.str_a[i] = oa[i];
  A.str_b[i] = ob[i];
}

// Per-particle terms of the pair sum that depend on the neighbour alone
// (particle_interaction.cu:52-60): sigma/rho^2, its trace and m/rho, with m
// the Adami volume mass vol_def*rho for wall slots and the particle mass for
// soil.  Computed once per slot here instead of once per pair; the values
// are rounded exactly as the inline expressions.  They go into the slot's
// 64-byte neighbour record next to its position and velocity, so the four
// loads of a pair read one cache line.
__global__ void __launch_bounds__(256) k_pair_prep(const DevParams* __restrict__ P, Arrays A, Buffers B, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float rho = A.vel[i].w;
  const float m = i < P->npb ? P->vol_def * rho : fabsf(A.pos[i].w);
  const Sym sr = sdiv(ldsym(A.str_a, A.str_b, i), rho * rho);
  float4* rc = B.rec + 4 * size_t(i);
  rc[0] = A.pos[i];
  rc[1] = A.vel[i];
  rc[2] = make_float4(sr.xx, sr.yy, sr.zz, sr.xy);
  rc[3] = make_float4(sr.xz, sr.yz, m / rho, strace(sr));
}

// Reference-order variant: each thread walks its 27 cells (cz, cy, cx;
// boundary range then soil range) exactly as particle_interaction.cu:117-140.
// With guard set it is the pair-list variant's fallback and runs only when
// the lists are not valid.
__global__ void __launch_bounds__(256) k_interaction_ref(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                         const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                         int remap, int guard) {
  if (guard && plist_exact(P)) return;
  int idx = xcd_block(remap) * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float4 vri = A.vel[i];
  const float3 vi = ld3(vri);
  const float rhoi = vri.w;
  const float4 sria = rec_sra(B, i), srib = rec_srb(B, i);
  const Sym sri = {sria.x, sria.y, sria.z, sria.w, srib.x, srib.y};
  const float tri = srib.w;
  const float artpi = tri > 0.f ? tri * P->art_p : 0.f;
  const int shift = int(A.cvar[i].z);
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  const PairConst K = pair_const(P);
  IAcc S;
  S.accden = 0.f;
  S.as = S.av = S.aa = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < 9; k++) S.gv[k] = 0.f;
  auto scan = [&](int4 ci, float3 pi) {
    for (int cz = -1; cz <= 1; cz++)
      for (int cy = -1; cy <= 1; cy++)
        for (int cx = -1; cx <= 1; cx++) {
          int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
          if (!in_grid(P, jx, jy, jz)) continue;
          int cb = grid_key(P, jx, jy, jz);
          for (int pass = 0; pass < 2; pass++) {
            const bool bound = pass == 0;
            int c = bound ? cb : cb + P->ncell_material;
            for (int j = cbeg[c], e = cend[c]; j < e; j++) {
              float3 d;
              const float4 pj = A.pos[j];
              const float r2 = pair_r2(pi, pj, cx, cy, cz, two_h, d);
              if (!(r2 <= r2cut)) continue;
              float r = sqrtf(r2);
              float q = r * i_h;
              pair_terms(K, S, d, r, q, vi, rhoi, sri, artpi, shift, bound, A.vel[j], pj.w, rec_sra(B, j),
                         rec_srb(B, j));
            }
          }
        }
  };
  scan(ci, pi);
  int4 cpb;
  float3 ppb;
  if (P->periodbound && period_image(P, ci, pi, false, cpb, ppb)) scan(cpb, ppb);  // particle_interaction.cu:243-258
  interaction_epilogue(P, A, B, i, S, vi);
}

// Neighbour-list variant.  The wave walks the 27 cells x {wall, soil}
// ranges in reference order (loop indices are wave-uniform, so the lanes of
// one home cell read the same candidate at the same time).  Each range is
// tested in chunks of NL_CHUNK candidates into a per-lane hit mask; the hits
// are appended to a per-lane FIFO ring of NL_CAP entries in LDS.  When some
// lane's next hits would not fit, the wave evaluates ring entries -- every
// lane with entries pops its oldest one per iteration -- only until they do,
// and drains the rings at the end.  Lanes thus stay busy on pair terms
// gathered over many ranges, while each particle still sums its pairs in
// reference order (a FIFO keeps the insertion order).
// An entry packs the slot j (bits 6..31, so slots below 2^26, the context
// limit) and the cell offset (bits 0..5, two bits per axis).  Whether j is a
// wall or a soil neighbour needs no bit: wall slots sort before every soil
// slot, so j < npb exactly when j came from a wall range, i.e. when the
// entry is below npb << 6.
constexpr int NL_WAVE = 64;
constexpr int NL_WAVES = 1;
constexpr int NL_CAP = 32;    // ring entries per lane (power of two)
constexpr int NL_CHUNK = 32;  // candidates per hit mask
static_assert((NL_CAP & (NL_CAP - 1)) == 0 && NL_CAP >= NL_CHUNK, "ring must hold a full chunk of hits");

// The neighbour's four 16-byte records (position and mass, velocity and
// density, sigma/rho^2, its last two components with m/rho and the trace).
struct NbrRec {
  float4 p, v, a, b;
};

__device__ __forceinline__ NbrRec nl_load(const Arrays& A, const Buffers& B, int e) {
  const float4* r = B.rec + 4 * size_t(unsigned(e) >> 6);
  return {r[0], r[1], r[2], r[3]};
}

__device__ __forceinline__ void nl_eval(const PairConst& K, IAcc& S, int e, const NbrRec& n, float3 pi, float two_h,
                                        float i_h, float3 vi, float rhoi, Sym sri, float artpi, int shift,
                                        unsigned wall_end) {
  // cell offset o in {-1, 0, 1} per axis, stored as o + 1: (o + 1) * 2h - 2h
  // is exact, so these equal float(o) * 2h of pair_r2
  float3 d;
  d.x = (pi.x - n.p.x) - fmaf(float(e & 3), two_h, -two_h);
  d.y = (pi.y - n.p.y) - fmaf(float((e >> 2) & 3), two_h, -two_h);
  d.z = (pi.z - n.p.z) - fmaf(float((e >> 4) & 3), two_h, -two_h);
  const float r = sqrtf(fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x)));
  const float q = r * i_h;
  pair_terms(K, S, d, r, q, vi, rhoi, sri, artpi, shift, unsigned(e) < wall_end, n.v, n.p.w, n.a, n.b);
}

// Evaluate the lane's `pop` oldest ring entries in order.
__device__ __forceinline__ void nl_drain(const PairConst& K, IAcc& S, const Arrays& A, const Buffers& B,
                                         const int (*lst)[NL_WAVE], int lane, int head, int pop, float3 pi,
                                         float two_h, float i_h, float3 vi, float rhoi, Sym sri, float artpi,
                                         int shift, unsigned wall_end) {
#pragma unroll 1
  for (int k = 0; k < pop; k++) {
    const int en = lst[(head + k) & (NL_CAP - 1)][lane];
    nl_eval(K, S, en, nl_load(A, B, en), pi, two_h, i_h, vi, rhoi, sri, artpi, shift, wall_end);
  }
}

// The wave's scan of the 27 cells x {wall, soil} ranges around (ci, pi)
// into the per-lane rings (see above).  Lanes with active unset have empty
// ranges; the loop indices stay wave-uniform.
__device__ __forceinline__ void nl_scan(const DevParams* __restrict__ P, const PairConst& K, IAcc& S, const Arrays& A,
                                        const Buffers& B, const int* __restrict__ cbeg, const int* __restrict__ cend,
                                        int (*lst)[NL_WAVE], int lane, int& head, int& cnt, int4 ci, float3 pi,
                                        bool active, float3 vi, float rhoi, Sym sri, float artpi, int shift) {
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  const int ncm = P->ncell_material;
  const unsigned wall_end = unsigned(P->npb) << 6;
#pragma unroll 1
  for (int r = 0; r < 54; r++) {
    const int cell = r >> 1, pass = r & 1;
    const int cz = cell / 9 - 1, cy = (cell / 3) % 3 - 1, cx = cell % 3 - 1;
    const int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
    int j = 0, e = 0;
    if (active && in_grid(P, jx, jy, jz)) {
      const int c = grid_key(P, jx, jy, jz) + (pass ? ncm : 0);
      j = cbeg[c];
      e = cend[c];
    }
    const unsigned code = unsigned((cx + 1) | ((cy + 1) << 2) | ((cz + 1) << 4));
    while (j < e) {
      const int m = min(e - j, NL_CHUNK);
      unsigned mask = 0u;
#pragma unroll 4
      for (int t = 0; t < m; t++) {  // unrolled: several candidate loads in flight
        float3 d;
        if (pair_r2(pi, A.pos[j + t], cx, cy, cz, two_h, d) <= r2cut) mask |= 1u << t;
      }
      const int over = cnt + __popc(mask) - NL_CAP;
      if (__ballot(over > 0)) {
        // make room: need = largest overflow over the active lanes (binary
        // search with ballots, 0 < need <= NL_CHUNK); every lane with entries
        // pops up to need of them, the overflowing ones exactly enough
        int need = 0;
#pragma unroll
        for (int b = NL_CHUNK; b; b >>= 1)
          if (__ballot(over >= need + b)) need += b;
        const int pop = min(cnt, need);
        nl_drain(K, S, A, B, lst, lane, head, pop, pi, two_h, i_h, vi, rhoi, sri, artpi, shift, wall_end);
        head = (head + pop) & (NL_CAP - 1);
        cnt -= pop;
      }
      while (mask) {
        const int t = __ffs(mask) - 1;
        mask &= mask - 1u;
        lst[(head + cnt) & (NL_CAP - 1)][lane] = int((unsigned(j + t) << 6) | code);
        cnt++;
      }
      j += m;
    }
  }
}

// At most 96 VGPRs (5 waves per SIMD, 20 per CU, which the 8 KB of ring per
// wave also allows): the kernel is bound by its neighbour gathers and needs
// the waves to keep them in flight.  PERIODIC adds the image pass of
// cuParticleInteractionMaterialLoopPeriod (particle_interaction.cu:243-258)
// after the rings are drained, so each particle still sums its own-position
// pairs first and its image pairs after them, in reference order.
template <bool PERIODIC>
__global__ void __launch_bounds__(NL_WAVE * NL_WAVES, 5) k_interaction_nl(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                                      const int* __restrict__ cbeg,
                                                                      const int* __restrict__ cend, int remap) {
  __shared__ int s_list[NL_WAVES][NL_CAP][NL_WAVE];
  const int lane = threadIdx.x & (NL_WAVE - 1);
  const int idx = xcd_block(remap) * blockDim.x + threadIdx.x;
  const bool active = idx < P->npm;
  const int i = P->npb + (active ? idx : 0);
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float4 vri = A.vel[i];
  const float3 vi = ld3(vri);
  const float rhoi = vri.w;
  const float4 sria = rec_sra(B, i), srib = rec_srb(B, i);
  const Sym sri = {sria.x, sria.y, sria.z, sria.w, srib.x, srib.y};
  const float tri = srib.w;
  const float artpi = tri > 0.f ? tri * P->art_p : 0.f;
  const int shift = int(A.cvar[i].z);
  const float two_h = P->two_h, i_h = P->i_h;
  const PairConst K = pair_const(P);
  int (*lst)[NL_WAVE] = s_list[threadIdx.x / NL_WAVE];
  IAcc S;
  S.accden = 0.f;
  S.as = S.av = S.aa = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < 9; k++) S.gv[k] = 0.f;
  int head = 0, cnt = 0;  // ring: oldest entry, number of entries
  const unsigned wall_end = unsigned(P->npb) << 6;
  nl_scan(P, K, S, A, B, cbeg, cend, lst, lane, head, cnt, ci, pi, active, vi, rhoi, sri, artpi, shift);
  nl_drain(K, S, A, B, lst, lane, head, cnt, pi, two_h, i_h, vi, rhoi, sri, artpi, shift, wall_end);
  if constexpr (PERIODIC) {
    head = 0;
    cnt = 0;
    int4 cpb;
    float3 ppb;
    const bool img = active && period_image(P, ci, pi, false, cpb, ppb);
    if (!img) { cpb = ci; ppb = pi; }
    nl_scan(P, K, S, A, B, cbeg, cend, lst, lane, head, cnt, cpb, ppb, img, vi, rhoi, sri, artpi, shift);
    nl_drain(K, S, A, B, lst, lane, head, cnt, ppb, two_h, i_h, vi, rhoi, sri, artpi, shift, wall_end);
  }
  if (active) interaction_epilogue(P, A, B, i, S, vi);
}

// Pass B: same-row redistribution for deferred particles
// (particle_interaction.cu:142-172, 209-237), reading pass-A values only.
// A group of 32 lanes per deferred particle, lane r < 27 walking the soil
// range of neighbour cell r, four candidates' loads in flight at a time
// (the walk is a chain of dependent loads; splitting it by cell keeps it
// short).  The serial loop keeps K = min(K, Kj) over the qualifying
// neighbours and overwrites fr = fj * K / Kj at each, so its result is the
// minimum K and the force of the last qualifying neighbour in traversal
// order scaled by that minimum: the group reduces exactly those (a min and
// an arg-max of the traversal position, cell r then slot) and evaluates the
// same expression.
// npos / ncell: where the neighbours' positions and cells are read, A.pos
// and A.cellpos, or the copies a fused interaction keeps of its first-layer
// particles before it moves them.
constexpr int RD_GROUP = 32;  // lanes per deferred particle (27 cells)
__global__ void __launch_bounds__(256) k_contact_redistribute(const DevParams* __restrict__ P, Arrays A, Buffers B,
                                                              const int* __restrict__ cbeg,
                                                              const int* __restrict__ cend,
                                                              const float4* __restrict__ npos,
                                                              const int4* __restrict__ ncell) {
  const int nd = P->n_deferred;
  const float two_h = P->two_h, band = P->band, half_dr = P->half_dr;
  const int ncm = P->ncell_material;
  constexpr int G = RD_GROUP;
  const int lane = threadIdx.x & (G - 1);
  const int ngroups = int(gridDim.x * blockDim.x) / G;
  // k is uniform within a group and the shuffles below stay inside it
  for (int k = int(blockIdx.x * blockDim.x + threadIdx.x) / G; k < nd; k += ngroups) {
    const int i = B.deferred[k];
    const int4 ci = A.cellpos[i];
    const float3 pi = ld3(A.pos[i]);
    const DevCase& cs = P->cases[A.case_id[i]];  // neighbours share the case
    const float3 pa = make_float3(pi.x + float(ci.x - cs.ox) * two_h, pi.y + float(ci.y - cs.oy) * two_h,
                                  pi.z + float(ci.z - cs.oz) * two_h);
    const float3 ni = ld3(A.cn[i]);
    float K = A.cvar[i].x;
    int best = -1, bestj = 0;  // traversal position (cell << 20 | offset) and slot of the last qualifying j
    const int r = lane;
    const int jx = ci.x + r % 3 - 1, jy = ci.y + (r / 3) % 3 - 1, jz = ci.z + r / 9 - 1;
    int b = 0, e = 0;
    if (r < 27 && in_grid(P, jx, jy, jz)) {
      const int c = grid_key(P, jx, jy, jz) + ncm;
      b = cbeg[c];
      e = cend[c];
    }
    for (int j0 = b; j0 < e; j0 += 4) {
      float4 cv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) cv[u] = j0 + u < e ? A.cvar[j0 + u] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = j0 + u;
        if (j >= e || !(cv[u].y <= band)) continue;
        const int4 cjj = ncell[j];
        const float4 pj = npos[j];
        const float3 pja = make_float3(pj.x + float(cjj.x - cs.ox) * two_h, pj.y + float(cjj.y - cs.oy) * two_h,
                                       pj.z + float(cjj.z - cs.oz) * two_h);
        const float3 dc = f3sub(pa, pja);
        const float rcol = fabsf(dc.x * ni.z + dc.y * ni.y + dc.z * ni.x);
        const float Kj = cv[u].x;
        if (rcol <= half_dr && Kj != 0.f) {
          if (K > Kj) K = Kj;
          best = (r << 20) | (j - b);  // j ascends in the lane's one cell
          bestj = j;
        }
      }
    }
    for (int o = G / 2; o > 0; o >>= 1) {
      const float Ko = __shfl_xor(K, o);
      if (K > Ko) K = Ko;
      const int bo = __shfl_xor(best, o), jo = __shfl_xor(bestj, o);
      if (bo > best) { best = bo; bestj = jo; }
    }
    if (lane == 0) {
      float3 fr = make_float3(0.f, 0.f, 0.f);
      if (best >= 0) {
        const float4 fj = A.cf[bestj];
        const float Kj = A.cvar[bestj].x;
        fr = make_float3(fj.x * K / Kj, fj.y * K / Kj, fj.z * K / Kj);
      }
      B.fdef[k] = make_float4(fr.x, fr.y, fr.z, K);
      const float4 s3m = B.s3[i];
      const float3 ast = f3mul(fr, 1.0f / s3m.w);
      const float3 vi = ld3(A.vel[i]);
      const float3 damp = f3mul(vi, P->damping);
      const float3 g = body_acc(P, A.codes, i, s3m.w);
      const float3 acc = f3add(f3sub(f3add(make_float3(s3m.x, s3m.y, s3m.z), ast), damp), g);
      const float4 old = B.accad[i];
      B.accad[i] = make_float4(acc.x, acc.y, acc.z, old.w);
    }
  }
}

// Pass A of the wall contact (particle_interaction.cu:174-237) for the
// particles a fuse-3 COLUMNS interaction left for after the corner search
// (their cell has a wall particle within its 27 cells): B.s3 holds their
// pair-force sum and density rate.
__global__ void __launch_bounds__(256) k_contact_after(DevParams* __restrict__ P, Arrays A, Buffers B) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const int4 ci = A.cellpos[i];
  if (!B.wall_near[grid_key(P, ci.x, ci.y, ci.z)]) return;
  const float4 t = B.s3[i];
  interaction_contact(P, A, B, i, make_float3(t.x, t.y, t.z), t.w, ld3(A.vel[i]));
}

// Pass C: publish pass-B forces and the min-reduced K.
__device__ __forceinline__ void contact_commit_one(const DevParams* __restrict__ P, Arrays A, const Buffers& B, int k) {
  const int i = B.deferred[k];
  float4 f = B.fdef[k];
  A.cf[i] = make_float4(f.x, f.y, f.z, 0.f);
  float4 cv = A.cvar[i];
  cv.x = f.w;
  A.cvar[i] = cv;
}
__global__ void k_contact_commit(const DevParams* __restrict__ P, Arrays A, Buffers B) {
  const int nd = P->n_deferred;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nd; k += gridDim.x * blockDim.x) contact_commit_one(P, A, B, k);
}

// With track set, the wave maximum of the squared displacement goes to
// max_disp2_bits for the pair lists (non-negative floats order like their
// bit patterns).  The running maximum only grows, so a wave whose maximum a
// plain read already covers skips the atomic.
// commit: the pass-C publication of the last interaction's deferred forces
// (k_contact_commit) rides along, one entry per thread: it writes cf and
// cvar, which the predictor neither reads nor writes.
__global__ void __launch_bounds__(256) k_predictor(DevParams* __restrict__ P, Arrays A, Buffers B, int track,
                                                   int commit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (commit && i < P->n_deferred) contact_commit_one(P, A, B, i);
  float d2 = i < P->np ? predictor_slot(P, A, B, i) : 0.f;
  if (!track) return;
  for (int o = 32; o > 0; o >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, o));
  if ((threadIdx.x & 63) == 0 && d2 > 0.f) {
    const unsigned bits = __float_as_uint(d2);
    if (bits > __hip_atomic_load(&P->max_disp2_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&P->max_disp2_bits, bits);
  }
}

// --------------------------------------------------------- corrector ----
// cuCorrectorBound, predictor_corrector.cu:109-137.
__device__ __forceinline__ void corrector_bound_slot(DevParams* __restrict__ P, Arrays A, const Buffers& B, int i) {
  int4 cd = A.codes[i];
  if (cd.y != LQ_NONFIX) return;
  const DevMotion& m = P->mv[cd.w];
  const float two_h = P->two_h, dt = P->dt;
  float3 vel = m.trans_vel;
  int4 ci = A.cellpos[i];
  if (m.is_rotat && m.pattern != LQ_CIRCLE) {
    float4 pp = A.pos[i];
    const int3 lc = case_cell(P, A.case_id, i, ci);
    float3 pa = make_float3(pp.x + float(lc.x) * two_h, pp.y + float(lc.y) * two_h, pp.z + float(lc.z) * two_h);
    vel = f3add(vel, f3cross(m.rotat_vel, f3sub(pa, m.pos_rp)));
    float3 ang = f3mul(m.rotat_vel, dt);  // GetRodriguesMatrix, functions.h:126-131
    float th = f3norm(ang);
    if (th > 0.f) {
      float3 k = f3div(ang, th);
      float sn = sinf(th), cc = 1.0f - cosf(th);
      float Kxx = -(k.z * k.z + k.y * k.y), Kyy = -(k.z * k.z + k.x * k.x), Kzz = -(k.y * k.y + k.x * k.x);
      float Rxx = 1.0f + Kxx * cc, Rxy = -k.z * sn + k.x * k.y * cc, Rxz = k.y * sn + k.x * k.z * cc;
      float Ryx = k.z * sn + k.x * k.y * cc, Ryy = 1.0f + Kyy * cc, Ryz = -k.x * sn + k.y * k.z * cc;
      float Rzx = -k.y * sn + k.x * k.z * cc, Rzy = k.x * sn + k.y * k.z * cc, Rzz = 1.0f + Kzz * cc;
      float4 n = A.normal[i];
      A.normal[i] = make_float4(Rxx * n.x + Rxy * n.y + Rxz * n.z, Ryx * n.x + Ryy * n.y + Ryz * n.z,
                                Rzx * n.x + Rzy * n.y + Rzz * n.z, n.w);
    }
  }
  float4 pini = B.pos_ini[i];
  float3 p = f3add(ld3(pini), f3mul(vel, dt));
  int dx = rebin(p.x, two_h), dy = rebin(p.y, two_h), dz = rebin(p.z, two_h);
  p.x = p.x - float(dx) * two_h; p.y = p.y - float(dy) * two_h; p.z = p.z - float(dz) * two_h;
  A.pos[i] = make_float4(p.x, p.y, p.z, pini.w);
  ci.x += dx; ci.y += dy; ci.z += dz;
  A.cellpos[i] = ci;
  if (!in_case_grid(P, A.case_id, i, ci)) set_err(P, DERR_DOMAIN);
}
__global__ void __launch_bounds__(256) k_corrector_bound(DevParams* __restrict__ P, Arrays A, Buffers B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P->npb) corrector_bound_slot(P, A, B, i);
}
// Both corrector kernels in one launch: the first nbb blocks take the wall
// slots, the rest the soil slots (each slot reads and writes only its own
// rows).
__global__ void __launch_bounds__(256) k_corrector_all(DevParams* __restrict__ P, Arrays A, Buffers B, int nbb,
                                                       int commit) {
  // pass C of the last interaction as in k_predictor (the corrector does not
  // touch cf or cvar either)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (commit && g < P->n_deferred) contact_commit_one(P, A, B, g);
  if (int(blockIdx.x) < nbb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P->npb) corrector_bound_slot(P, A, B, i);
    return;
  }
  const int idx = (blockIdx.x - nbb) * blockDim.x + threadIdx.x;
  if (idx < P->npm) corrector_slot(P, A, B, P->npb + idx);
}


// The slots a fused interaction leaves to the predictor (the snapshot of
// walls and of deleted slots, and the deferred soil, whose acceleration pass
// B completes), then those it leaves to the corrector (the deferred soil).
__global__ void __launch_bounds__(256) k_predictor_finish(DevParams* __restrict__ P, Arrays A, Buffers B) {
  const int npb = P->npb, tail = P->npb + P->npm, nd = P->n_deferred;
  const int n = npb + nd + (P->np - tail);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int i = t < npb ? t : t < npb + nd ? B.deferred[t - npb] : tail + (t - npb - nd);
    predictor_slot(P, A, B, i);
  }
}
__global__ void __launch_bounds__(256) k_corrector_finish(DevParams* __restrict__ P, Arrays A, Buffers B) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P->n_deferred; k += gridDim.x * blockDim.x)
    corrector_slot(P, A, B, B.deferred[k]);
}

// -------------------------------------------------------- wall force ----
// cuGetValuesAtPoint, utilities.cu:578-640, per case: the contact force x
// of every soil particle inside the case's probe box is summed exactly, as
// a 128-bit fixed-point integer in units of 2^-64 N (values below 2^-64 N
// truncate toward zero), so the sum does not depend on the order in which
// slots are visited; the host scales it by 0.001 once.
__device__ __forceinline__ void acc128(unsigned long long* a, float v) {
  const unsigned u = __float_as_uint(v);
  const int e = int((u >> 23) & 0xffu);
  if (e == 0 || e == 255) return;  // zero, subnormal (< 2^-126) or non-finite: contributes nothing
  const unsigned long long m = (u & 0x7fffffu) | 0x800000u;
  const int sh = e - 127 - 23 + 64;  // value = m * 2^(e-150) = m * 2^(sh) units of 2^-64
  unsigned __int128 q;
  if (sh >= 0) q = sh >= 104 ? 0 : (unsigned __int128)m << sh;  // |v| >= 2^63 N is out of range
  else q = sh <= -24 ? 0 : (unsigned __int128)(m >> -sh);
  if (u >> 31) q = ~q + 1;  // two's complement
  const unsigned long long lo = (unsigned long long)q, hi = (unsigned long long)(q >> 64);
  const unsigned long long old = atomicAdd(a, lo);
  atomicAdd(a + 1, hi + (old + lo < old ? 1ull : 0ull));
}

__global__ void __launch_bounds__(256) k_wall_force(const DevParams* __restrict__ P, Arrays A,
                                                    const float* __restrict__ boxes,
                                                    unsigned long long* __restrict__ acc) {
  const float two_h = P->two_h;
  // slots npb..np as utilities.cu:592, deleted particles' stale forces included
  for (int i = P->npb + blockIdx.x * blockDim.x + threadIdx.x; i < P->np; i += gridDim.x * blockDim.x) {
    const int k = A.case_id[i];
    const float4 p = A.pos[i];
    const int3 c = case_cell(P, A.case_id, i, A.cellpos[i]);
    const float xa = p.x + float(c.x) * two_h;
    const float za = p.z + float(c.z) * two_h;
    if (xa <= boxes[3 * k] && za >= boxes[3 * k + 1] && za <= boxes[3 * k + 2]) acc128(acc + 2 * k, A.cf[i].x);
  }
}

// -------------------------------------------- wall pressure profile ----
// cuGetValuesAlongLine, utilities.cu:266-320: the soil slots in three
// one-dr bands in front of the wall face x_face (bands closed at both ends)
// and inside [z_lo, z_hi] are binned by depth num_z = int((z_top - z)/dr);
// in each (band, bin) the reference's serial loop over slots keeps the last
// slot's contact force x / dr, and the band-1 write also sets the depth
// num_z dr.  Here every slot records itself in last[band][bin] with
// atomicMax (the largest slot is the serial loop's last writer), and a
// second kernel gathers the values.  geom = {x_face, z_lo, z_hi, z_top}.
// Absolute positions are formed in FP32.
__global__ void __launch_bounds__(256) k_profile_mark(const DevParams* __restrict__ P, Arrays A, float4 geom,
                                                      int nbins, int* __restrict__ last) {
  const int i = P->npb + int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= P->np) return;
  const float3 pa = abs_from_cell(P, A.cellpos[i], ld3(A.pos[i]));
  const float dr = P->dr;
  if (!(pa.z >= geom.y && pa.z <= geom.z)) return;
  const int bin = int((geom.w - pa.z) / dr);
  if (bin < 0 || bin >= nbins) return;
#pragma unroll
  for (int b = 0; b < 3; b++)
    if (pa.x >= geom.x + dr * float(b) && pa.x <= geom.x + dr * float(b + 1)) atomicMax(&last[b * nbins + bin], i);
}

__global__ void __launch_bounds__(256) k_profile_write(const DevParams* __restrict__ P, Arrays A, int nbins,
                                                       const int* __restrict__ last, float4* __restrict__ out) {
  const int bin = blockIdx.x * blockDim.x + threadIdx.x;
  if (bin >= nbins) return;
  const float dr = P->dr;
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);  // depth, band 1, band 2, band 3
  const int j1 = last[bin], j2 = last[nbins + bin], j3 = last[2 * nbins + bin];
  if (j1 >= 0) { r.x = float(bin) * dr; r.y = A.cf[j1].x / dr; }
  if (j2 >= 0) r.z = A.cf[j2].x / dr;
  if (j3 >= 0) r.w = A.cf[j3].x / dr;
  out[bin] = r;
}

// ----------------------------------------------------------- sensors ----
// GetSoilStress + cuGetStressZz (utilities.cu:470-528), at any points: the
// Shepard-weighted stress sum_j sigma_j W V_j / sum_j W V_j over soil slots
// with q = |x - x_j| / h < 2, x_j the absolute position (relative position +
// cell 2h, FP32) and V_j = m_j / rho_j.  rows != 0 takes the reference's
// cells (the point's own y row, every z level, z outer and x inner);
// rows == 0 the 27 cells around the point (z, y, x nested).  One thread per
// point, serial sums.  out[k] = (xx yy zz xy) and (xz yz, sum W V, hits).
__global__ void __launch_bounds__(64) k_sensor_stress(const DevParams* __restrict__ P, Arrays A,
                                                      const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                      const float4* __restrict__ pts, int n, int rows,
                                                      float4* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 x = pts[k];
  const float scale = 0.5f * P->i_h, two_h = P->two_h;
  const int cx0 = int(floorf(x.x * scale)), cy0 = int(floorf(x.y * scale)), cz0 = int(floorf(x.z * scale));
  float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float wsum = 0.f;
  int hits = 0;
  const int z0 = rows ? 0 : cz0 - 1, z1 = rows ? P->cnz - 1 : cz0 + 1;
  const int y0 = rows ? cy0 : cy0 - 1, y1 = rows ? cy0 : cy0 + 1;
  for (int cz = z0; cz <= z1; cz++)
    for (int cy = y0; cy <= y1; cy++)
      for (int cx = cx0 - 1; cx <= cx0 + 1; cx++) {
        if (!in_grid(P, cx, cy, cz)) continue;
        const int c = grid_key(P, cx, cy, cz) + P->ncell_material;
        const int b = cbeg[c], e = cend[c];
        for (int j = b; j < e; j++) {
          const int4 cj = A.cellpos[j];
          const float4 pj = A.pos[j];
          const float ax = pj.x + float(cj.x) * two_h, ay = pj.y + float(cj.y) * two_h,
                      az = pj.z + float(cj.z) * two_h;
          const float dx = x.x - ax, dy = x.y - ay, dz = x.z - az;
          const float q = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) * P->i_h;
          if (!(q < 2.0f)) continue;
          const float ker = kW(P->knorm, q);
          const float vol = A.idm[j].y / A.vel[j].w;
          const Sym sj = ldsym(A.str_a, A.str_b, j);
          s[0] = s[0] + sj.xx * ker * vol; s[1] = s[1] + sj.yy * ker * vol; s[2] = s[2] + sj.zz * ker * vol;
          s[3] = s[3] + sj.xy * ker * vol; s[4] = s[4] + sj.xz * ker * vol; s[5] = s[5] + sj.yz * ker * vol;
          wsum = wsum + ker * vol;
          hits++;
        }
      }
  if (wsum != 0.f)
    for (int q = 0; q < 6; q++) s[q] = s[q] / wsum;
  out[2 * k] = make_float4(s[0], s[1], s[2], s[3]);
  out[2 * k + 1] = make_float4(s[4], s[5], wsum, float(hits));
}

// GetSurface (utilities.cu:26-91) and GetSurfaceAtPoint (:371-427), at any
// points (x, y): the largest absolute z of the soil slots near the vertical
// line through the point, -1e10 when there is none.  radius > 0: horizontal
// distance sqrt(dx^2 + dy^2) < radius over the cells x +-1, y +-1, every z
// (GetSurface's test, with radius 0.5 dr there); radius < 0: |dx| < -radius
// over the point's own y row, x +-1, every z (GetSurfaceAtPoint's, with
// 2h).  Positions are absolute (relative + cell 2h, FP32).
__global__ void __launch_bounds__(64) k_surface_heights(const DevParams* __restrict__ P, Arrays A,
                                                        const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                        const float2* __restrict__ pts, int n, float radius,
                                                        float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float2 x = pts[k];
  const float scale = 0.5f * P->i_h, two_h = P->two_h;
  const int cx0 = int(floorf(x.x * scale)), cy0 = int(floorf(x.y * scale));
  const bool strip = radius < 0.f;
  const int y0 = strip ? cy0 : cy0 - 1, y1 = strip ? cy0 : cy0 + 1;
  float top = -1.0e10f;
  for (int cz = 0; cz < P->cnz; cz++)
    for (int cy = y0; cy <= y1; cy++)
      for (int cx = cx0 - 1; cx <= cx0 + 1; cx++) {
        if (!in_grid(P, cx, cy, cz)) continue;
        const int c = grid_key(P, cx, cy, cz) + P->ncell_material;
        const int b = cbeg[c], e = cend[c];
        for (int j = b; j < e; j++) {
          const int4 cj = A.cellpos[j];
          const float4 pj = A.pos[j];
          const float dx = x.x - (pj.x + float(cj.x) * two_h);
          const float dy = x.y - (pj.y + float(cj.y) * two_h);
          const bool near = strip ? fabsf(dx) < -radius : sqrtf(fmaf(dy, dy, dx * dx)) < radius;
          if (near) top = fmaxf(top, pj.z + float(cj.z) * two_h);
        }
      }
  out[k] = top;
}

__global__ void k_fill_u64(unsigned long long* a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

// ------------------------------------------------- neighbour lists ------
__global__ void k_neighbor_list(const DevParams* __restrict__ P, Arrays A, const int* __restrict__ cbeg,
                                const int* __restrict__ cend, int max_nb, int* __restrict__ counts,
                                int* __restrict__ ids) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  int n = 0;
  for (int cz = -1; cz <= 1; cz++)
    for (int cy = -1; cy <= 1; cy++)
      for (int cx = -1; cx <= 1; cx++) {
        int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
        if (!in_grid(P, jx, jy, jz)) continue;
        int cb = grid_key(P, jx, jy, jz);
        for (int pass = 0; pass < 2; pass++) {
          int c = pass ? cb + P->ncell_material : cb;
          for (int j = cbeg[c], e = cend[c]; j < e; j++) {
            float3 d;
            const float r2 = pair_r2(pi, A.pos[j], cx, cy, cz, P->two_h, d);
            if (!(r2 <= P->r2cut)) continue;
            if (n < max_nb) ids[(size_t)idx * max_nb + n] = A.orig_id[j];
            n++;
          }
        }
      }
  counts[idx] = n;
}

// ------------------------------------------------------- stream copy ----
__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) b[i] = a[i];
}

// --------------------------------------------------------- launchers ----
static inline int blocks(long long n, int b) { return int((n + b - 1) / b); }

void launch_abs2rel(hipStream_t s, const DevParams* P, Arrays A, float scale, int first, int n) {
  if (n > 0) hipLaunchKernelGGL(k_abs2rel, dim3(blocks(n, 256)), dim3(256), 0, s, P, A, scale, first, n);
}
void launch_set_massflag(hipStream_t s, const DevParams* P, int np, Arrays A) {
  launch_set_massflag_range(s, P, A, 0, np);
}
void launch_set_massflag_range(hipStream_t s, const DevParams* P, Arrays A, int first, int n) {
  if (n > 0) hipLaunchKernelGGL(k_set_massflag, dim3(blocks(n, 256)), dim3(256), 0, s, P, A, first, n);
}
void launch_copy_w(hipStream_t s, const DevParams* P, int np, const float4* src, float4* dst) {
  if (np > 0) hipLaunchKernelGGL(k_copy_w, dim3(blocks(np, 256)), dim3(256), 0, s, P, src, dst);
}
void launch_rel2abs(hipStream_t s, const DevParams* P, int np, Arrays A, float4* out) {
  if (np > 0) hipLaunchKernelGGL(k_rel2abs, dim3(blocks(np, 256)), dim3(256), 0, s, P, A, out);
}
void launch_step_calc(hipStream_t s, DevParams* P) { hipLaunchKernelGGL(k_step_calc, dim3(1), dim3(1), 0, s, P); }
void launch_advance(hipStream_t s, DevParams* P) { hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, P); }
void launch_rigid_body_update(hipStream_t s, DevParams* P) {
  hipLaunchKernelGGL(k_rigid_body_update, dim3(4), dim3(64), 0, s, P);
}
void launch_step_begin(hipStream_t s, DevParams* P, int n_motion) {
  hipLaunchKernelGGL(k_step_begin, dim3(n_motion > 0 ? 4 : 1), dim3(64), 0, s, P);
}
void launch_adami(hipStream_t s, const DevParams* P, int npb, Arrays A, const int* cb, const int* ce) {
  // 64-thread blocks spread the wall slots evenly over the CUs; below 16
  // waves per CU (256 CUs) the two sums go to separate threads
  const int split = npb < 256 * 64 * 16 ? 1 : 0;
  if (npb > 0)
    hipLaunchKernelGGL(k_adami, dim3(blocks((split ? 2 : 1) * (long long)npb, 64)), dim3(64), 0, s, P, A, cb, ce, 1,
                       split);
}
void launch_corner(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                   const int* wall_near) {
  if (npm_max > 0)
    hipLaunchKernelGGL(k_corner, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, cb, ce, wall_near, 1);
}
void launch_shepard(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                    float4* oa, float2* ob, bool copy_back, bool due_only) {
  if (npm_max <= 0) return;
  hipLaunchKernelGGL(k_shepard, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, cb, ce, oa, ob, int(due_only));
  if (copy_back) hipLaunchKernelGGL(k_copy_soil_str, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, (const float4*)oa,
                     (const float2*)ob, int(due_only));
}
void launch_reset_deferred(hipStream_t s, DevParams* P, int list_build) {
  hipLaunchKernelGGL(k_reset_deferred, dim3(1), dim3(1), 0, s, P, list_build);
}
void launch_pair_prep(hipStream_t s, const DevParams* P, int n, Arrays A, const Buffers& B) {
  if (n > 0) hipLaunchKernelGGL(k_pair_prep, dim3(blocks(n, 256)), dim3(256), 0, s, P, A, B, n);
}
void launch_interaction(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                        const int* ce, int variant, int xcd_remap, bool periodic) {
  if (npm_max <= 0) return;
  if (variant == LQ_INTERACTION_REFERENCE_ORDER)
    hipLaunchKernelGGL(k_interaction_ref, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, B, cb, ce,
                       xcd_remap, 0);
  else if (variant == LQ_INTERACTION_NEIGHBOR_LIST && periodic)
    hipLaunchKernelGGL(k_interaction_nl<true>, dim3(blocks(npm_max, NL_WAVE * NL_WAVES)), dim3(NL_WAVE * NL_WAVES),
                       0, s, P, A, B, cb, ce, xcd_remap);
  else if (variant == LQ_INTERACTION_NEIGHBOR_LIST)
    hipLaunchKernelGGL(k_interaction_nl<false>, dim3(blocks(npm_max, NL_WAVE * NL_WAVES)), dim3(NL_WAVE * NL_WAVES),
                       0, s, P, A, B, cb, ce, xcd_remap);
  else if (variant == LQ_INTERACTION_PAIR_LIST)  // fallback: runs only if the pair lists are not exact
    hipLaunchKernelGGL(k_interaction_ref, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, B, cb, ce,
                       xcd_remap, 1);
}
void launch_contact(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                    const int* ce, bool fused, bool commit) {
  if (npm_max <= 0) return;
  int g = std::min(blocks(npm_max, 256), 1024);
  // RD_GROUP lanes per deferred particle, up to 8 * 4096 at a time
  hipLaunchKernelGGL(k_contact_redistribute, dim3(std::min(blocks(npm_max, 256 / RD_GROUP), 4096)), dim3(256), 0, s, P,
                     A, B, cb,
                     ce, fused ? (const float4*)B.fl_pos : (const float4*)A.pos,
                     fused ? (const int4*)B.fl_cell : (const int4*)A.cellpos);
  if (commit) hipLaunchKernelGGL(k_contact_commit, dim3(g), dim3(256), 0, s, P, A, B);
}
void launch_contact_after(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B) {
  if (npm_max > 0) hipLaunchKernelGGL(k_contact_after, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, B);
}
void launch_predictor(hipStream_t s, DevParams* P, int np, Arrays A, const Buffers& B, int track_disp, int commit) {
  if (np > 0) hipLaunchKernelGGL(k_predictor, dim3(blocks(np, 256)), dim3(256), 0, s, P, A, B, track_disp, commit);
}
void launch_predictor_finish(hipStream_t s, DevParams* P, Arrays A, const Buffers& B) {
  hipLaunchKernelGGL(k_predictor_finish, dim3(1024), dim3(256), 0, s, P, A, B);
}
void launch_corrector_finish(hipStream_t s, DevParams* P, int npb, int n_motion, Arrays A, const Buffers& B) {
  if (n_motion > 0 && npb > 0)
    hipLaunchKernelGGL(k_corrector_bound, dim3(blocks(npb, 256)), dim3(256), 0, s, P, A, B);
  hipLaunchKernelGGL(k_corrector_finish, dim3(256), dim3(256), 0, s, P, A, B);
}
void launch_corrector(hipStream_t s, DevParams* P, int npb, int npm_max, int n_motion, Arrays A, const Buffers& B,
                      int commit) {
  const int nbb = n_motion > 0 && npb > 0 ? blocks(npb, 256) : 0;
  const int nbm = npm_max > 0 ? blocks(npm_max, 256) : 0;
  if (nbb + nbm > 0) hipLaunchKernelGGL(k_corrector_all, dim3(nbb + nbm), dim3(256), 0, s, P, A, B, nbb, commit);
}
void launch_wall_profile(hipStream_t s, const DevParams* P, int np, int npb, Arrays A, float4 geom, int nbins, int* last,
                         float4* out) {
  launch_fill_int(s, last, 3 * nbins, -1);
  if (np > npb)
    hipLaunchKernelGGL(k_profile_mark, dim3(blocks(np - npb, 256)), dim3(256), 0, s, P, A, geom, nbins, last);
  hipLaunchKernelGGL(k_profile_write, dim3(blocks(nbins, 256)), dim3(256), 0, s, P, A, nbins, (const int*)last, out);
}
void launch_sensor_stress(hipStream_t s, const DevParams* P, Arrays A, const int* cb, const int* ce, const float4* pts,
                          int n, int rows, float4* out) {
  if (n > 0) hipLaunchKernelGGL(k_sensor_stress, dim3(blocks(n, 64)), dim3(64), 0, s, P, A, cb, ce, pts, n, rows, out);
}
void launch_surface_heights(hipStream_t s, const DevParams* P, Arrays A, const int* cb, const int* ce,
                            const float2* pts, int n, float radius, float* out) {
  if (n > 0)
    hipLaunchKernelGGL(k_surface_heights, dim3(blocks(n, 64)), dim3(64), 0, s, P, A, cb, ce, pts, n, radius, out);
}
void launch_wall_force(hipStream_t s, const DevParams* P, int np, int n_cases, Arrays A, const float* boxes,
                       unsigned long long* acc) {
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks(2 * n_cases, 256)), dim3(256), 0, s, acc, 2 * n_cases);
  hipLaunchKernelGGL(k_wall_force, dim3(std::max(1, std::min(blocks(np, 256), 2048))), dim3(256), 0, s, P, A, boxes,
                     acc);
}
void launch_stream_copy(hipStream_t s, const float4* a, float4* b, size_t n) {
  int g = int(std::min<size_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(256), 0, s, a, b, n);
}
void launch_neighbor_list(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const int* cb, const int* ce,
                          int max_nb, int* counts, int* ids) {
  if (npm_max > 0)
    hipLaunchKernelGGL(k_neighbor_list, dim3(blocks(npm_max, 256)), dim3(256), 0, s, P, A, cb, ce, max_nb, counts,
                       ids);
}

}  // namespace lq
