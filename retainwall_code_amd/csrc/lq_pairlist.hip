// lq_pairlist.hip -- pair-list interaction variant.
//
// One scan per time step records, for every soil slot, the candidates that
// lie within the list radius (the pair cut-off 2h plus a 1% skin) in the
// exact traversal order of particle_interaction.cu:117-140 (cz, cy, cx;
// wall range, then soil range; j ascending).  Both interaction calls of the
// step evaluate from the list and re-apply the exact r^2 <= r2_cut test, so
// the pairs, their order and every rounding are those of the cell scan.
//
// The lists are exact while (a) no soil slot overflowed its capacity and
// (b) no particle moved more than disp2_lim since the build: then a pair
// outside the list radius at the build is still outside 2h.  Particles keep
// their cells within a step (the predictor never re-bins), and the build is
// repeated after every sort, so slot indices and ranges are unchanged.  The
// predictor records the largest displacement; when (a) or (b) fails the
// evaluation kernel exits and the reference-order kernel (launched behind it
// with its guard set) recomputes the step's interaction by cell scan.
//
// Layout in HBM: entry k of soil slot s at plist[k * stride + s] (stride =
// soil count at upload), so lane-consecutive slots read and write
// consecutive words.  An entry packs the neighbour slot j (bits 7..31) with
// the cell offset (bits 0..5: cx+1, cy+1, cz+1, two bits each) and the range
// (bit 6: 0 wall, 1 soil).
#include <hip/hip_runtime.h>

#include "lq_device.hpp"
#include "lq_internal.hpp"
#include "lq_launch.hpp"

namespace lq {

__global__ void __launch_bounds__(256) k_plist_build(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                     const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                     int remap) {
  const int idx = xcd_block(remap) * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float two_h = P->two_h, r2list = P->r2list;
  const int cap = P->plist_cap, stride = P->plist_stride, ncm = P->ncell_material;
  int* __restrict__ out = B.plist + idx;
  int n = 0;
  for (int cz = -1; cz <= 1; cz++)
    for (int cy = -1; cy <= 1; cy++)
      for (int cx = -1; cx <= 1; cx++) {
        const int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
        if (!in_grid(P, jx, jy, jz)) continue;
        const int cb = grid_key(P, jx, jy, jz);
        const int off = (cx + 1) | ((cy + 1) << 2) | ((cz + 1) << 4);
        for (int pass = 0; pass < 2; pass++) {
          const int c = pass ? cb + ncm : cb;
          const int code = off | (pass << 6);
          for (int j = cbeg[c], e = cend[c]; j < e; j++) {
            float3 d;
            if (pair_r2(pi, A.pos[j], cx, cy, cz, two_h, d) <= r2list) {
              if (n < cap) out[size_t(n) * stride] = int((unsigned(j) << 7) | unsigned(code));
              n++;
            }
          }
        }
      }
  B.pcount[idx] = n;
  if (n > cap) atomicOr(&P->plist_bad, 1);
}

__global__ void __launch_bounds__(256) k_interaction_plist(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                           int remap) {
  if (!plist_exact(P)) return;  // the guarded reference-order kernel takes over
  const int idx = xcd_block(remap) * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const float4 vri = A.vel[i];
  const float3 vi = ld3(vri);
  const float rhoi = vri.w;
  const float4 sria = rec_sra(B, i), srib = rec_srb(B, i);
  const Sym sri = {sria.x, sria.y, sria.z, sria.w, srib.x, srib.y};
  const float tri = srib.w;
  const float artpi = tri > 0.f ? tri * P->art_p : 0.f;
  const int shift = int(A.cvar[i].z);
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  const int stride = P->plist_stride;
  const PairConst K = pair_const(P);
  IAcc S;
  S.accden = 0.f;
  S.as = S.av = S.aa = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < 9; k++) S.gv[k] = 0.f;
  const int n = B.pcount[idx];
  const int* __restrict__ lst = B.plist + idx;
  int e = n > 0 ? lst[0] : 0;
  for (int k = 0; k < n; k++) {
    const int j = int(unsigned(e) >> 7);
    const int ox = (e & 3) - 1, oy = ((e >> 2) & 3) - 1, oz = ((e >> 4) & 3) - 1;
    const bool bound = (e & 64) == 0;
    const float4* rj = B.rec + 4 * size_t(j);  // the neighbour's 64-byte record
    const float4 pj = rj[0];
    if (k + 1 < n) e = lst[size_t(k + 1) * stride];  // next entry in flight during this pair
    float3 d;
    const float r2 = pair_r2(pi, pj, ox, oy, oz, two_h, d);
    if (r2 <= r2cut) {
      const float r = sqrtf(r2);
      const float q = r * i_h;
      pair_terms(K, S, d, r, q, vi, rhoi, sri, artpi, shift, bound, rj[1], pj.w, rj[2], rj[3]);
    }
  }
  interaction_epilogue(P, A, B, i, S, vi);
}

static int blocks256(int n) { return (n + 255) / 256; }

void launch_plist_build(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                        const int* ce, int xcd_remap) {
  if (npm_max > 0)
    hipLaunchKernelGGL(k_plist_build, dim3(blocks256(npm_max)), dim3(256), 0, s, P, A, B, cb, ce, xcd_remap);
}
void launch_interaction_plist(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, const int* cb,
                              const int* ce, int xcd_remap) {
  (void)cb;
  (void)ce;
  if (npm_max > 0)
    hipLaunchKernelGGL(k_interaction_plist, dim3(blocks256(npm_max)), dim3(256), 0, s, P, A, B, xcd_remap);
}

}  // namespace lq
