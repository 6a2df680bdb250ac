// lq_device.hpp -- device math shared by the kernel translation units:
// small vector/tensor helpers, the per-pair interaction terms
// (particle_interaction.cu:26-115), the interaction epilogue with pass A of
// the wall contact (particle_interaction.cu:174-237, 284-295) and the HypoCS
// stress update (constitutive_model.h:209-299).  Every fused multiply-add is
// written as fmaf(); the TUs are compiled with -ffp-contract=off.
#pragma once
#include "lq_internal.hpp"

namespace lq {

__device__ __forceinline__ float3 ld3(const float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 f3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 f3mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 f3div(float3 a, float s) { return make_float3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float f3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float f3norm(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 f3cross(float3 a, float3 b) {  // functions.h:100
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

struct Sym { float xx, yy, zz, xy, xz, yz; };
__device__ __forceinline__ Sym ldsym(const float4* a, const float2* b, int i) {
  float4 u = a[i]; float2 v = b[i];
  return {u.x, u.y, u.z, u.w, v.x, v.y};
}
__device__ __forceinline__ void stsym(float4* a, float2* b, int i, Sym s) {
  a[i] = make_float4(s.xx, s.yy, s.zz, s.xy);
  b[i] = make_float2(s.xz, s.yz);
}
__device__ __forceinline__ Sym sadd(Sym a, Sym b) {
  return {a.xx + b.xx, a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}
__device__ __forceinline__ Sym smul(Sym a, float s) {
  return {a.xx * s, a.yy * s, a.zz * s, a.xy * s, a.xz * s, a.yz * s};
}
__device__ __forceinline__ Sym sdiv(Sym a, float s) {
  return {a.xx / s, a.yy / s, a.zz / s, a.xy / s, a.xz / s, a.yz / s};
}
__device__ __forceinline__ float strace(Sym a) { return a.xx + a.yy + a.zz; }
__device__ __forceinline__ float sddot(Sym a, Sym b) {  // functions.h:95
  return a.xx * b.xx + a.yy * b.yy + a.zz * b.zz + 2.0f * a.xy * b.xy + 2.0f * a.xz * b.xz + 2.0f * a.yz * b.yz;
}
__device__ __forceinline__ float3 smatvec(Sym s, float3 v) {  // functions.h:97
  return make_float3(s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
                     s.xz * v.x + s.yz * v.y + s.zz * v.z);
}

// Candidate geometry, particle_interaction.cu:30-33:
// d = pos_i - pos_j - off*2h and r^2 = fma(dz,dz,fma(dy,dy,dx*dx)); r = sqrt(r^2).
// The reference keeps a pair when q = fl(fl(sqrt(r^2)) * (1/h)) < 2.  With
// correctly rounded sqrt and multiply that test is monotone in r^2, so it is
// exactly r^2 <= r2cut for the largest float r2cut that passes
// (lq_interaction_cutoff_sq bisects the float bit patterns for it); the
// square root is then taken only for kept pairs.
__device__ __forceinline__ float pair_r2(float3 pi, float4 pj, int ox, int oy, int oz, float two_h, float3& d) {
  d.x = (pi.x - pj.x) - float(ox) * two_h;
  d.y = (pi.y - pj.y) - float(oy) * two_h;
  d.z = (pi.z - pj.z) - float(oz) * two_h;
  return fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x));
}

// Hit mask of the candidates [j0, j0 + m), m <= 32: bit t set when slot
// j0 + t passes the cut-off.  The loop is unrolled so several candidate
// loads are in flight; callers then walk the set bits in ascending order,
// which keeps the reference's summation order.
__device__ __forceinline__ unsigned hit_mask(const float4* __restrict__ pos, float3 pi, int j0, int m, int ox, int oy,
                                             int oz, float two_h, float r2cut) {
  unsigned mask = 0u;
#pragma unroll 4
  for (int t = 0; t < m; t++) {
    float3 d;
    if (pair_r2(pi, pos[j0 + t], ox, oy, oz, two_h, d) <= r2cut) mask |= 1u << t;
  }
  return mask;
}

// Hardware places workgroup b on XCD b % 8.  Renumber the 1-D grid so each
// XCD runs one contiguous run of logical blocks: consecutive particles are
// cell neighbours, so their candidate data then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int remap) {
  const int nb = int(gridDim.x), b = int(blockIdx.x);
  if (!remap) return b;
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

// True when the pair lists built at the start of the step are still exact:
// no soil slot overflowed its capacity and no particle moved more than the
// skin allows since the build.
__device__ __forceinline__ bool plist_exact(const DevParams* __restrict__ P) {
  return P->plist_bad == 0 && __uint_as_float(P->max_disp2_bits) <= P->disp2_lim;
}

// ------------------------------------------------------- interaction ----
// Constants of the pair sum, loaded once per thread so the inner loops hold
// them in registers.
struct PairConst {
  float knorm, kgrad, vol_def, h, eta, alpha_cs, art_p, kernel_dr, art_exp;
  int viscosity_type;
};
__device__ __forceinline__ PairConst pair_const(const DevParams* __restrict__ P) {
  return {P->knorm, P->kgrad, P->vol_def, P->h, P->eta, P->alpha_cs, P->art_p, P->kernel_dr, P->art_exp,
          P->viscosity_type};
}

// Per-pair terms of cuParticleInteractionMaterialLoopSolution,
// particle_interaction.cu:26-115, accumulated into the per-particle sums.
struct IAcc {
  float accden;
  float3 as, av, aa;
  float gv[9];  // xx xy xz yx yy yz zx zy zz
};

__device__ __forceinline__ void pair_terms(const PairConst& P, IAcc& S, float3 d, float r, float q,
                                           float3 vi, float rhoi, Sym sri, float artpi, int shift, bool bound,
                                           float4 vj, float wj, float4 sra, float4 srb) {
  // wj = pos.w of the neighbour: |wj| is a soil particle's mass, and wj < 0
  // marks a contact-wall code (codes.x >= 10).  sra/srb are the neighbour's
  // per-particle terms from k_pair_prep: sigma/rho^2 (xx yy zz xy | xz yz),
  // m/rho and trace(sigma/rho^2), each rounded exactly as the inline form.
  const bool contactj = wj < 0.f;
  float dwdr = kG(P.kgrad, q);
  float3 g = f3mul(d, dwdr);
  float rhoj = vj.w;
  float mj = bound ? P.vol_def * rhoj : fabsf(wj);
  float3 dv = f3sub(vi, ld3(vj));
  if (contactj && shift == 1) dv = make_float3(0.f, 0.f, 0.f);
  const float volj = srb.z;
  float3 dvv = f3mul(dv, volj);
  S.gv[0] = fmaf(-dvv.x, g.x, S.gv[0]); S.gv[1] = fmaf(-dvv.x, g.y, S.gv[1]); S.gv[2] = fmaf(-dvv.x, g.z, S.gv[2]);
  S.gv[3] = fmaf(-dvv.y, g.x, S.gv[3]); S.gv[4] = fmaf(-dvv.y, g.y, S.gv[4]); S.gv[5] = fmaf(-dvv.y, g.z, S.gv[5]);
  S.gv[6] = fmaf(-dvv.z, g.x, S.gv[6]); S.gv[7] = fmaf(-dvv.z, g.y, S.gv[7]); S.gv[8] = fmaf(-dvv.z, g.z, S.gv[8]);
  // The remaining terms do not apply to contact-wall neighbours
  // (particle_interaction.cu:66).  They are added with a zero factor rather
  // than branched around: fmaf(x, 0, s) == s for every finite x, because
  // these sums start at +0 and a round-to-nearest sum is never -0 unless
  // both addends are, so no sum here ever holds -0.
  float dvg = fmaf(dv.z, g.z, fmaf(dv.y, g.y, dv.x * g.x));
  S.accden = fmaf(contactj ? 0.f : rhoi * dvg, volj, S.accden);
  const Sym srj = {sra.x, sra.y, sra.z, sra.w, srb.x, srb.y};
  Sym st = sadd(sri, srj);
  float3 f = smatvec(st, g);
  const float mf = contactj ? 0.f : mj;
  S.as.x = fmaf(f.x, mf, S.as.x); S.as.y = fmaf(f.y, mf, S.as.y); S.as.z = fmaf(f.z, mf, S.as.z);
  if (P.viscosity_type == 1) {
    // evaluated for every pair (in a wave some lane nearly always needs
    // it); only approaching pairs (visc < 0) add it
    float visc = fmaf(dv.z, d.z, fmaf(dv.y, d.y, dv.x * d.x));
    float phi = P.h * visc / fmaf(r, r, P.eta);
    float artv = P.alpha_cs * phi / (0.5f * (rhoi + rhoj));
    float s = (visc < 0.f && !contactj) ? mj * artv : 0.f;
    S.av.x = fmaf(g.x, s, S.av.x); S.av.y = fmaf(g.y, s, S.av.y); S.av.z = fmaf(g.z, s, S.av.z);
  }
  if (P.art_p != 0.f) {
    float trj = srb.w;
    float artp = artpi + (trj > 0.f ? trj * P.art_p : 0.f);
    if (artp > 0.f && !contactj) {
      float ker = kW(P.knorm, q);
      artp = artp * powf(ker / P.kernel_dr, P.art_exp);
      float s = mj * artp;
      S.aa.x = fmaf(-g.x, s, S.aa.x); S.aa.y = fmaf(-g.y, s, S.aa.y); S.aa.z = fmaf(-g.z, s, S.aa.z);
    }
  }
}

// Epilogue of cuParticleInteractionMaterial (particle_interaction.cu:284-295)
// with pass A of the wall contact (LoopContact, :174-237), in two parts.
// interaction_rates writes the strain rate and spin and returns the sum of
// the pair forces; interaction_contact adds damping, gravity and the pass-A
// contact force of the corner search's nearest wall and writes the
// acceleration (and density rate accden).  It returns whether the
// particle's contact force waits for pass B.
__device__ __forceinline__ float3 interaction_rates(const Buffers& B, int i, const IAcc& S) {
  B.acestn_a[i] = make_float4(S.gv[0], S.gv[4], S.gv[8], 0.5f * (S.gv[1] + S.gv[3]));
  B.acestn_b[i] = make_float2(0.5f * (S.gv[2] + S.gv[6]), 0.5f * (S.gv[5] + S.gv[7]));
  B.spin[i] = make_float4(0.5f * (S.gv[3] - S.gv[1]), 0.5f * (S.gv[6] - S.gv[2]), 0.5f * (S.gv[7] - S.gv[5]), 0.f);
  return f3add(f3add(S.as, S.av), S.aa);
}
// The acceleration of a particle with no contact wall within 2h (what
// interaction_contact gives when the corner search found none).
__device__ __forceinline__ void interaction_free(DevParams* __restrict__ P, Arrays A, const Buffers& B, int i,
                                                 float3 s3, float accden, float3 vi) {
  const float3 damp = f3mul(vi, P->damping);
  const float3 g = body_acc(P, A.codes, i, A.idm[i].y);
  const float3 acc = f3add(f3sub(f3add(s3, make_float3(0.f, 0.f, 0.f)), damp), g);
  B.accad[i] = make_float4(acc.x, acc.y, acc.z, accden);
}
__device__ __forceinline__ bool interaction_contact(DevParams* __restrict__ P, Arrays A, const Buffers& B, int i,
                                                    float3 s3, float accden, float3 vi) {
  float3 damp = f3mul(vi, P->damping);
  float mi = A.idm[i].y;
  const float3 g = body_acc(P, A.codes, i, mi);
  float3 acc0 = f3add(f3sub(s3, damp), g);
  const float dt = P->dt;
  float3 vp = make_float3(fmaf(acc0.x, dt, vi.x), fmaf(acc0.y, dt, vi.y), fmaf(acc0.z, dt, vi.z));
  float4 cvi = A.cvar[i];
  float3 nb = ld3(A.cn[i]);
  float3 vb = ld3(A.cw[i]);
  float3 dvc = f3sub(vb, vp);
  float3 ast = make_float3(0.f, 0.f, 0.f);
  float dvn = f3dot(dvc, nb);
  bool deferred = false;
  if (cvi.y <= P->h * 2.0f && dvn > 0.f) {
    float3 dvn3 = f3mul(nb, dvn);
    float3 fn = f3div(f3mul(dvn3, mi), dt);
    float3 ft = f3div(f3mul(f3sub(dvc, dvn3), mi), dt);
    float afn = f3norm(fn), aft = f3norm(ft);
    if (aft > P->mu * afn) { ft = f3mul(ft, P->mu * afn / aft); cvi.z = 1.f; }
    else cvi.z = 0.f;
    A.cvar[i] = cvi;
    float3 fr = f3add(ft, fn);
    if (cvi.y > P->dr) {
      deferred = true;
    } else {
      A.cf[i] = make_float4(fr.x, fr.y, fr.z, 0.f);
      ast = f3mul(fr, 1.0f / mi);
    }
  }
  if (deferred) {
    int k = atomicAdd(&P->n_deferred, 1);
    B.deferred[k] = i;
    B.s3[i] = make_float4(s3.x, s3.y, s3.z, mi);
  }
  float3 acc = f3add(f3sub(f3add(s3, ast), damp), g);
  B.accad[i] = make_float4(acc.x, acc.y, acc.z, accden);
  return deferred;
}
__device__ __forceinline__ bool interaction_epilogue(DevParams* __restrict__ P, Arrays A, const Buffers& B, int i,
                                                     const IAcc& S, float3 vi) {
  const float3 s3 = interaction_rates(B, i, S);
  return interaction_contact(P, A, B, i, s3, S.accden, vi);
}

// ------------------------------------------------------------ HypoCS ----
// cuMatModHypo (critical state), constitutive_model.h:209-299.
__device__ __forceinline__ void hypo_cs(const DevParams* __restrict__ P, int model, float dt, float rho, float3 w,
                                        Sym D, Sym& sig, float& Mcrt) {
  const float* mp = P->model_par[model];
  const HypoConst& hc = P->hypo[model];
  const float coh = mp[3], pref = mp[4], ec0 = mp[5], lam = mp[6], xi = mp[7], alpha = mp[8];
  const float emin = mp[9], emax = mp[10], e0 = mp[11], rho0 = mp[12];
  Sym s = sig;
  if (coh > 0.f) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  const float Sxx = s.xx, Syy = s.yy, Szz = s.zz, Sxy = s.xy, Sxz = s.xz, Syz = s.yz;
  // W S - S W with W = SkewTensor2FullTensor(w), symmetric part
  float WSxx = -w.z * Sxy + w.y * Sxz, WSxy = -w.z * Syy + w.y * Syz, WSxz = -w.z * Syz + w.y * Szz;
  float WSyx = w.z * Sxx - w.x * Sxz, WSyy = w.z * Sxy - w.x * Syz, WSyz = w.z * Sxz - w.x * Szz;
  float WSzx = -w.y * Sxx + w.x * Sxy, WSzy = -w.y * Sxy + w.x * Syy, WSzz = -w.y * Sxz + w.x * Syz;
  float SWxx = Sxy * w.z - Sxz * w.y, SWxy = -Sxx * w.z + Sxz * w.x, SWxz = Sxx * w.y - Sxy * w.x;
  float SWyx = Syy * w.z - Syz * w.y, SWyy = -Sxy * w.z + Syz * w.x, SWyz = Sxy * w.y - Syy * w.x;
  float SWzx = Syz * w.z - Szz * w.y, SWzy = -Sxz * w.z + Szz * w.x, SWzz = Sxz * w.y - Syz * w.x;
  Sym jau = {WSxx - SWxx, WSyy - SWyy, WSzz - SWzz, 0.5f * ((WSxy - SWxy) + (WSyx - SWyx)),
             0.5f * ((WSxz - SWxz) + (WSzx - SWzx)), 0.5f * ((WSyz - SWyz) + (WSzy - SWzy))};
  float I1 = strace(s);
  if (I1 > -4.0f) { sig = Sym{-1.f, -1.f, -2.f, 0.f, 0.f, 0.f}; return; }
  Sym dev = s;
  float third = I1 / 3.0f;
  dev.xx = dev.xx - third; dev.yy = dev.yy - third; dev.zz = dev.zz - third;
  float trs = strace(s), trd = strace(D), trsd = sddot(s, D), nd = sqrtf(sddot(D, D));
  float p = -I1 / 3.0f;
  float e = (1.0f + e0) * rho0 / rho - 1.0f;
  e = fmaxf(e, emin);
  e = fminf(e, emax);
  float ec = ec0 * expf(-lam * powf(p / pref, xi));
  float fd = powf((e - emin) / (ec - emin), alpha);
  Mcrt = e - ec;
  Sym sd = sadd(s, dev);
  Sym rate = sadd(sadd(sadd(sadd(jau, smul(D, hc.C1 * trs)), smul(s, hc.C2 * trd)), smul(s, hc.C3 * trsd / trs)),
                  smul(sd, fd * hc.C4 * nd));
  Sym ns = sadd(sig, smul(rate, dt));
  float sph = strace(ns) / 3.0f;
  if (sph > 0.f) { sig = Sym{-1.0f + coh, -1.0f + coh, -2.0f + coh, 0.f, 0.f, 0.f}; return; }
  Sym nd6 = ns;
  nd6.xx = nd6.xx - sph; nd6.yy = nd6.yy - sph; nd6.zz = nd6.zz - sph;
  const float fa = hc.fa, fb = hc.fb, fc = hc.fc;
  float u = 1.0f + fa + fb / 3.0f;
  float f2 = fd * fd * fc * fc;
  float A_ = f2 / 3.0f - u * u;
  float B_ = 2.0f * f2 * (u * u + 2.0f * fa + 3.0f) - 4.0f * fb * u;
  float C_ = 4.0f / 3.0f * f2 * fb * fb - 4.0f * fb * fb;
  float k2 = (-B_ - sqrtf(B_ * B_ - 4.0f * A_ * C_)) / (2.0f * A_);
  float kphi = 3.0f / sqrtf(k2);
  float tau = sqrtf(0.5f * sddot(nd6, nd6));
  if (tau - (-sph * kphi) > 0.f) {
    float ratio = tau < 1e-10f ? 1.0f : (-kphi * sph) / tau;
    nd6 = smul(nd6, ratio);
  }
  sig = Sym{nd6.xx + sph + coh, nd6.yy + sph + coh, nd6.zz + sph + coh, nd6.xy, nd6.xz, nd6.yz};
}

// ------------------------------------------------- HypoCS surrogate ----
// The surrogate learns the costly middle of hypo_cs: from
//   x = (s = sigma - coh (xx yy zz xy xz yz), D dt (6), e (clamped), dt)
// it predicts y = ((rate - Jaumann term) dt (6), M_crt).  The Jaumann term,
// the I1 > -4 and tensile branches and the shear cap stay exact
// (hypo_cs_apply).  hypo_cs_parts computes x and, when want_y, the exact y
// with the operations of hypo_cs; it returns false when I1 > -4 (hypo_cs
// then returns a fixed tensor and no surrogate is involved).
constexpr int SUR_IN = 14, SUR_OUT = 7, SUR_XS = 16, SUR_YS = 8;

__device__ __forceinline__ Sym jaumann(float3 w, const Sym& s) {
  const float Sxx = s.xx, Syy = s.yy, Szz = s.zz, Sxy = s.xy, Sxz = s.xz, Syz = s.yz;
  float WSxx = -w.z * Sxy + w.y * Sxz, WSxy = -w.z * Syy + w.y * Syz, WSxz = -w.z * Syz + w.y * Szz;
  float WSyx = w.z * Sxx - w.x * Sxz, WSyy = w.z * Sxy - w.x * Syz, WSyz = w.z * Sxz - w.x * Szz;
  float WSzx = -w.y * Sxx + w.x * Sxy, WSzy = -w.y * Sxy + w.x * Syy, WSzz = -w.y * Sxz + w.x * Syz;
  float SWxx = Sxy * w.z - Sxz * w.y, SWxy = -Sxx * w.z + Sxz * w.x, SWxz = Sxx * w.y - Sxy * w.x;
  float SWyx = Syy * w.z - Syz * w.y, SWyy = -Sxy * w.z + Syz * w.x, SWyz = Sxy * w.y - Syy * w.x;
  float SWzx = Syz * w.z - Szz * w.y, SWzy = -Sxz * w.z + Szz * w.x, SWzz = Sxz * w.y - Syz * w.x;
  return Sym{WSxx - SWxx, WSyy - SWyy, WSzz - SWzz, 0.5f * ((WSxy - SWxy) + (WSyx - SWyx)),
             0.5f * ((WSxz - SWxz) + (WSzx - SWzx)), 0.5f * ((WSyz - SWyz) + (WSzy - SWzy))};
}

__device__ __forceinline__ bool hypo_cs_parts(const DevParams* __restrict__ P, int model, float dt, float rho,
                                              Sym D, const Sym& sig, float* x, float* y, bool want_y) {
  const float* mp = P->model_par[model];
  const HypoConst& hc = P->hypo[model];
  const float coh = mp[3], pref = mp[4], ec0 = mp[5], lam = mp[6], xi = mp[7], alpha = mp[8];
  const float emin = mp[9], emax = mp[10], e0 = mp[11], rho0 = mp[12];
  Sym s = sig;
  if (coh > 0.f) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  const float I1 = strace(s);
  if (I1 > -4.0f) return false;
  float e = (1.0f + e0) * rho0 / rho - 1.0f;
  e = fmaxf(e, emin);
  e = fminf(e, emax);
  x[0] = s.xx; x[1] = s.yy; x[2] = s.zz; x[3] = s.xy; x[4] = s.xz; x[5] = s.yz;
  x[6] = D.xx * dt; x[7] = D.yy * dt; x[8] = D.zz * dt; x[9] = D.xy * dt; x[10] = D.xz * dt; x[11] = D.yz * dt;
  x[12] = e; x[13] = dt;
  if (!want_y) return true;
  Sym dev = s;
  const float third = I1 / 3.0f;
  dev.xx = dev.xx - third; dev.yy = dev.yy - third; dev.zz = dev.zz - third;
  const float trs = strace(s), trd = strace(D), trsd = sddot(s, D), nd = sqrtf(sddot(D, D));
  const float p = -I1 / 3.0f;
  const float ec = ec0 * expf(-lam * powf(p / pref, xi));
  const float fd = powf((e - emin) / (ec - emin), alpha);
  const Sym sd = sadd(s, dev);
  const Sym r = sadd(sadd(sadd(smul(D, hc.C1 * trs), smul(s, hc.C2 * trd)), smul(s, hc.C3 * trsd / trs)),
                     smul(sd, fd * hc.C4 * nd));
  y[0] = r.xx * dt; y[1] = r.yy * dt; y[2] = r.zz * dt; y[3] = r.xy * dt; y[4] = r.xz * dt; y[5] = r.yz * dt;
  y[6] = e - ec;
  return true;
}

// Exact tail of hypo_cs (constitutive_model.h:266-299) applied to a
// predicted increment dsh = (rate - Jaumann) dt and M_crt = e - e_c.
__device__ __forceinline__ void hypo_cs_apply(const DevParams* __restrict__ P, int model, float dt, float3 w,
                                              const float* yv, float e, Sym& sig) {
  const float* mp = P->model_par[model];
  const HypoConst& hc = P->hypo[model];
  const float coh = mp[3], alpha = mp[8], emin = mp[9];
  Sym s = sig;
  if (coh > 0.f) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  const Sym jau = jaumann(w, s);
  const Sym dsh = {yv[0], yv[1], yv[2], yv[3], yv[4], yv[5]};
  const Sym ns = sadd(sig, sadd(dsh, smul(jau, dt)));
  const float sph = strace(ns) / 3.0f;
  if (sph > 0.f) { sig = Sym{-1.0f + coh, -1.0f + coh, -2.0f + coh, 0.f, 0.f, 0.f}; return; }
  Sym nd6 = ns;
  nd6.xx = nd6.xx - sph; nd6.yy = nd6.yy - sph; nd6.zz = nd6.zz - sph;
  const float ec = e - yv[6];
  const float fd = powf(fmaxf((e - emin) / (ec - emin), 0.f), alpha);
  const float fa = hc.fa, fb = hc.fb, fc = hc.fc;
  const float u = 1.0f + fa + fb / 3.0f;
  const float f2 = fd * fd * fc * fc;
  const float A_ = f2 / 3.0f - u * u;
  const float B_ = 2.0f * f2 * (u * u + 2.0f * fa + 3.0f) - 4.0f * fb * u;
  const float C_ = 4.0f / 3.0f * f2 * fb * fb - 4.0f * fb * fb;
  const float k2 = (-B_ - sqrtf(B_ * B_ - 4.0f * A_ * C_)) / (2.0f * A_);
  const float kphi = 3.0f / sqrtf(k2);
  const float tau = sqrtf(0.5f * sddot(nd6, nd6));
  if (tau - (-sph * kphi) > 0.f) {
    const float ratio = tau < 1e-10f ? 1.0f : (-kphi * sph) / tau;
    nd6 = smul(nd6, ratio);
  }
  sig = Sym{nd6.xx + sph + coh, nd6.yy + sph + coh, nd6.zz + sph + coh, nd6.xy, nd6.xz, nd6.yz};
}

// ------------------------------- DP, mu(I)-DP, basic hypoplasticity ----
__device__ __forceinline__ Sym add_iso(Sym s, float p) { return Sym{s.xx + p, s.yy + p, s.zz + p, s.xy, s.xz, s.yz}; }

// Elastic trial stress of both Drucker-Prager models:
// sigma + (Jaumann + 2G D + lambda tr(D) I) dt with lambda = K - 2G/3.
__device__ __forceinline__ Sym elastic_trial(float dt, float3 w, Sym D, const Sym& sig, float K, float G) {
  const float lame = K - 2.0f / 3.0f * G;
  return sadd(sig, smul(add_iso(sadd(jaumann(w, sig), smul(D, 2.0f * G)), lame * strace(D)), dt));
}

// cuMatModDP, constitutive_model.h:38-84: Drucker-Prager, elastic-perfectly
// plastic with a tension cut-off.  mp = {K, G, k_phi, k_psi, k_c, ...}
// (GetMaterial_DP_EPP, GenCase.cu:16-25).  Returns true on a shear return.
__device__ __forceinline__ bool dp_epp(const float* __restrict__ mp, float dt, float3 w, Sym D, Sym& sig) {
  const float K = mp[0], G = mp[1], kphi = mp[2], kpsi = mp[3], kc = mp[4];
  const Sym pre = elastic_trial(dt, w, D, sig, K, G);
  float sph = strace(pre) / 3.0f;
  Sym dev = add_iso(pre, -(strace(pre) / 3.0f));
  const float tenlt = kphi == 0.f ? 0.f : kc / kphi;
  if (sph - tenlt >= 0.f) sph = tenlt;
  const float tau = sqrtf(0.5f * sddot(dev, dev));
  const float shr = tau - (-sph * kphi + kc);
  bool yld = false;
  if (shr > 0.f) {
    const float dlam = shr / (G + K * kphi * kpsi);
    sph = sph - K * kpsi * dlam;
    const float ratio = tau < 1e-10f ? 1.0f : (kc - kphi * sph) / tau;
    dev = smul(dev, ratio);
    yld = true;
  }
  sig = add_iso(dev, sph);
  return yld;
}

// cuMatModMuIDP, constitutive_model.h:90-130: Drucker-Prager with the mu(I)
// friction law.  mp = {K, G, mu_s, mu_2, I0, rho_g, d, lambda, ...}
// (GetMaterial_MuI_DP_EPP, GenCase.cu:27-40).
__device__ __forceinline__ void mui_dp(const float* __restrict__ mp, float dt, float3 w, Sym D, Sym& sig) {
  const float K = mp[0], G = mp[1], mus = mp[2], dmu = mp[3] - mp[2], I0 = mp[4], rhog = mp[5], dg = mp[6];
  const float lam = mp[7];
  const Sym pre = elastic_trial(dt, w, D, sig, K, G);
  float sph = strace(pre) / 3.0f;
  Sym dev = add_iso(pre, -(strace(pre) / 3.0f));
  const Sym dD = add_iso(D, -(strace(D) / 3.0f));
  const float gamma = sqrtf(2.0f * sddot(dD, dD));
  if (sph > 0.f) sph = 0.f;
  const float p = -sph;
  const float mu = mus + dmu * gamma * dg / (I0 * sqrtf(p / rhog) + dg * sqrtf(gamma * gamma + lam * lam));
  const float tau = sqrtf(0.5f * sddot(dev, dev));
  if (tau + sph * mu > 0.f) {
    const float ratio = tau < 1e-10f ? 1.0f : (0.0f - mu * sph) / tau;
    dev = smul(dev, ratio);
  }
  sig = add_iso(dev, sph);
}

// cuMatModHypo (basic overload), constitutive_model.h:136-202.  mp = {Ei,
// Vi, phi_c, c, p_ref, ...} (GetMaterial_HypoBasic, GenCase.cu:42-50); the
// constants C1..C4 are the critical-state model's (same formulas).
__device__ __forceinline__ void hypo_basic(const DevParams* __restrict__ P, int model, float dt, float3 w, Sym D,
                                           Sym& sig) {
  const float coh = P->model_par[model][3];
  const HypoConst& hc = P->hypo[model];
  Sym s = sig;
  if (coh > 0.f) { s.xx = s.xx - coh; s.yy = s.yy - coh; s.zz = s.zz - coh; }
  const Sym jau = jaumann(w, s);
  const float I1 = strace(s);
  if (I1 > -4.0f) { sig = Sym{-1.f, -1.f, -2.f, 0.f, 0.f, 0.f}; return; }
  const Sym dev = add_iso(s, -(I1 / 3.0f));
  const float trs = strace(s), trd = strace(D), trsd = sddot(s, D), nd = sqrtf(sddot(D, D));
  const Sym rate = sadd(sadd(sadd(sadd(jau, smul(D, hc.C1 * trs)), smul(s, hc.C2 * trd)), smul(s, hc.C3 * trsd / trs)),
                        smul(sadd(s, dev), hc.C4 * nd));
  const Sym ns = sadd(sig, smul(rate, dt));
  const float sph = strace(ns) / 3.0f;
  if (sph > 0.f) { sig = Sym{-1.0f + coh, -1.0f + coh, -2.0f + coh, 0.f, 0.f, 0.f}; return; }
  Sym nd6 = add_iso(ns, -sph);
  const float kphi = 3.0f / (sqrtf(2.0f) * hc.sq2sin);
  const float tau = sqrtf(0.5f * sddot(nd6, nd6));
  if (tau - (-sph * kphi) > 0.f) {
    const float ratio = tau < 1e-10f ? 1.0f : (-kphi * sph) / tau;
    nd6 = smul(nd6, ratio);
  }
  sig = add_iso(nd6, sph + coh);
}

// The model dispatch of cuPredictorMaterial / cuCorrectorMaterial
// (predictor_corrector.cu:72-75, 169-172) for every model but the
// critical-state one, which the callers handle (exact or surrogate).
// Returns true when a Drucker-Prager update yielded.
__device__ __forceinline__ bool constitutive_other(const DevParams* __restrict__ P, int4 cd, float dt, float3 w, Sym D,
                                                   Sym& sig) {
  const float* mp = P->model_par[cd.w];
  if (cd.z == LQ_DP_EPP) return dp_epp(mp, dt, w, D, sig);
  if (cd.z == LQ_MUI_DP_EPP) mui_dp(mp, dt, w, D, sig);
  else if (cd.z == LQ_HYPO_BASIC) hypo_basic(P, cd.w, dt, w, D, sig);
  return false;
}

// --------------------------------------------------------- predictor ----
// LoquatSph::Predictor + cuPredictorMaterial, predictor_corrector.cu:52-103:
// the *_ini snapshot and the half-step update in one pass.  Returns the
// squared displacement of slot i (soil that moves; 0 otherwise).
__device__ __forceinline__ float predictor_slot(const DevParams* __restrict__ P, Arrays A, const Buffers& B, int i) {
  const int npm = P->npm, npb = P->npb;
  float4 p = A.pos[i], v = A.vel[i];
  B.pos_ini[i] = p;
  B.vel_ini[i] = v;
  if (npm == 0 || i >= npb + npm) return 0.f;
  float4 sa = A.str_a[i], na = A.stn_a[i];
  float2 sb = A.str_b[i], nb = A.stn_b[i];
  B.str_ini_a[i] = sa; B.str_ini_b[i] = sb;
  B.stn_ini_a[i] = na; B.stn_ini_b[i] = nb;
  if (i < npb) return 0.f;
  int4 cd = A.codes[i];
  if (cd.y != LQ_NONFIX) return 0.f;
  const float hdt = 0.5f * P->dt;
  float4 a = B.accad[i];
  v = make_float4(fmaf(a.x, hdt, v.x), fmaf(a.y, hdt, v.y), fmaf(a.z, hdt, v.z), fmaf(a.w, hdt, v.w));
  const float4 p0 = p;
  p = make_float4(fmaf(v.x, hdt, p.x), fmaf(v.y, hdt, p.y), fmaf(v.z, hdt, p.z), p.w);
  const float* mp = P->model_par[cd.w];
  float rho = fmaxf(v.w, mp[13]);
  rho = fminf(rho, mp[14]);
  v.w = rho;
  Sym D = ldsym(B.acestn_a, B.acestn_b, i);
  Sym stn = {na.x, na.y, na.z, na.w, nb.x, nb.y};
  stsym(A.stn_a, A.stn_b, i, sadd(stn, smul(D, hdt)));
  Sym sig = {sa.x, sa.y, sa.z, sa.w, sb.x, sb.y};
  float Mcrt = 0.f;
  bool yld = false;
  if (cd.z == LQ_HYPO_CS) {
    if (!P->sur_on) hypo_cs(P, cd.w, hdt, rho, ld3(B.spin[i]), D, sig, Mcrt);
  } else {
    yld = constitutive_other(P, cd, hdt, ld3(B.spin[i]), D, sig);
  }
  stsym(A.str_a, A.str_b, i, sig);
  A.pos[i] = p;
  A.vel[i] = v;
  if (!yld) {  // a Drucker-Prager yield keeps the old value (predictor_corrector.cu:78-83)
    float4 id = A.idm[i];
    id.x = 100.f;
    A.idm[i] = id;
  }
  const float dx = p.x - p0.x, dy = p.y - p0.y, dz = p.z - p0.z;
  return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------- corrector (slot) ----
__device__ __forceinline__ int rebin(float p, float two_h) { return int(p >= two_h) - int(p < 0.0f); }

// cuCorrectorMaterial, predictor_corrector.cu:139-199, for soil slot i.
__device__ __forceinline__ void corrector_slot(const DevParams* __restrict__ P, Arrays A, const Buffers& B, int i) {
  int4 cd = A.codes[i];
  if (cd.y != LQ_NONFIX) return;
  const float dt = P->dt, hdt = 0.5f * dt, two_h = P->two_h;
  float4 vini = B.vel_ini[i], a = B.accad[i], pini = B.pos_ini[i];
  float4 v = make_float4(fmaf(a.x, dt, vini.x), fmaf(a.y, dt, vini.y), fmaf(a.z, dt, vini.z), fmaf(a.w, dt, vini.w));
  float3 p = make_float3(fmaf(v.x + vini.x, hdt, pini.x), fmaf(v.y + vini.y, hdt, pini.y), fmaf(v.z + vini.z, hdt, pini.z));
  const float* mp = P->model_par[cd.w];
  float rho = fmaxf(v.w, mp[13]);
  rho = fminf(rho, mp[14]);
  v.w = rho;
  int4 ci = A.cellpos[i];
  int dx = rebin(p.x, two_h), dy = rebin(p.y, two_h), dz = rebin(p.z, two_h);
  p.x = p.x - float(dx) * two_h; p.y = p.y - float(dy) * two_h; p.z = p.z - float(dz) * two_h;
  ci.x += dx; ci.y += dy; ci.z += dz;
  if (P->periodbound) wrap_period(P, ci, p);  // predictor_corrector.cu:158
  Sym D = ldsym(B.acestn_a, B.acestn_b, i);
  Sym stn0 = ldsym(B.stn_ini_a, B.stn_ini_b, i);
  stsym(A.stn_a, A.stn_b, i, sadd(stn0, smul(D, dt)));
  Sym sig = ldsym(B.str_ini_a, B.str_ini_b, i);
  float Mcrt = 0.f;
  bool yld = false;
  if (cd.z == LQ_HYPO_CS) {
    if (!P->sur_on) hypo_cs(P, cd.w, dt, rho, ld3(B.spin[i]), D, sig, Mcrt);
  } else {
    yld = constitutive_other(P, cd, dt, ld3(B.spin[i]), D, sig);
  }
  stsym(A.str_a, A.str_b, i, sig);
  float4 id = A.idm[i];
  if (!yld) id.x = 100.f;  // predictor_corrector.cu:175-180
  id.w = Mcrt;
  if (!in_case_grid(P, A.case_id, i, ci)) { ci.w = P->ncell_delete_m; id.x = 999.f; }
  A.idm[i] = id;
  A.cellpos[i] = ci;
  A.pos[i] = make_float4(p.x, p.y, p.z, pini.w);
  A.vel[i] = v;
  float4 po = A.pos_ori[i];
  if (P->time <= P->dt) po = make_float4(0.f, 0.f, 0.f, 0.f);
  po = make_float4(fmaf(v.x, dt, po.x), fmaf(v.y, dt, po.y), fmaf(v.z, dt, po.z), po.w);
  A.pos_ori[i] = po;
  B.accad[i] = po;
}

}  // namespace lq
