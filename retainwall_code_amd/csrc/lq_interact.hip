// lq_interact.hip -- the interaction kernel on LDS-staged cell tiles.
//
// cuParticleInteractionMaterial (particle_interaction.cu:260-297) evaluated
// per tile: a workgroup owns the soil particles of TZ consecutive cells of
// one (x, y) cell column.  With the linear key (z fastest) the neighbour
// region of the tile -- 9 columns x [z0-1, z0+TZ] x {wall, soil} -- is 18
// contiguous slot ranges, which the workgroup copies into LDS with coalesced
// loads.  Each lane then scans its own 27 cells x {wall, soil} from LDS in
// the reference order (cz, cy, cx, wall before soil, slot order), appends
// the hits to a per-lane list in LDS and evaluates the listed pairs with the
// whole wave busy.  Pair arithmetic and summation order are those of
// k_interaction_ref, so the results are bit-identical to it.
//
// Tiles are dealt so that each of the 8 XCDs (workgroups go to XCDs
// round-robin) walks one contiguous run of tiles, keeping its L2 on a
// compact region of the domain.  A tile whose neighbourhood exceeds the LDS
// capacity falls back to reading candidates from global memory.
#include <climits>

#include "lq_device.hpp"
#include "lq_internal.hpp"
#include "lq_launch.hpp"

namespace lq {

// Scan + evaluate for the tile's home particles.  STAGED reads candidate
// positions from the LDS copy; otherwise from global memory (tile overflow).
template <bool STAGED, int WAVES, int NLCAP, int NZW>
__device__ __forceinline__ void tile_home(DevParams* __restrict__ P, const Arrays& A, const Buffers& B,
                                          const int* __restrict__ cbeg, const int* __restrict__ cend,
                                          const float4* __restrict__ s_pos, const int (*s_lo)[NZW],
                                          const int (*s_hi)[NZW], const int* s_rs, const int* s_ro,
                                          unsigned (*s_list)[NLCAP][64], int hb, int he, int zw0, int nzw) {
  constexpr int WG = 64 * WAVES;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int ncm = P->ncell_material;
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut;
  const int H = he - hb;
  for (int base = 0; base < H; base += WG) {
    const int li = base + t;
    const bool active = li < H;
    const int i = hb + (active ? li : 0);
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
    const int zoff = ci.z - zw0;
    const PairConst K = pair_const(P);
    IAcc S;
    S.accden = 0.f;
    S.as = S.av = S.aa = make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 9; k++) S.gv[k] = 0.f;
    int cstep = active ? 0 : 54;
    int j = 0, jend = 0;
    int n = 0;
    for (;;) {
      while (n < NLCAP) {
        if (j >= jend) {
          bool found = false;
          while (cstep < 54) {
            const int cur = cstep++;
            const int cell = cur >> 1, bsel = cur & 1;
            const int cz = cell / 9 - 1, cy = (cell / 3) % 3 - 1, cx = cell % 3 - 1;
            if constexpr (STAGED) {
              const int zi = zoff + cz;
              if (zi < 0 || zi >= nzw) continue;
              const int r = ((cy + 1) * 3 + (cx + 1)) * 2 + bsel;
              const int lo = s_lo[r][zi], hi = s_hi[r][zi];
              if (lo >= hi) continue;
              j = lo;
              jend = hi;
            } else {
              const int jx = ci.x + cx, jy = ci.y + cy, jz = ci.z + cz;
              if (!in_grid(P, jx, jy, jz)) continue;
              const int c = grid_key(P, jx, jy, jz) + (bsel ? ncm : 0);
              const int bb = cbeg[c];
              if (bb < 0) continue;
              j = bb;
              jend = cend[c];
            }
            found = true;
            break;
          }
          if (!found) break;
        }
        const int cur = cstep - 1;
        const int cell = cur >> 1;
        const int cz = cell / 9 - 1, cy = (cell / 3) % 3 - 1, cx = cell % 3 - 1;
        float4 pj;
        if constexpr (STAGED) pj = s_pos[j];
        else pj = A.pos[j];
        float3 d;
        if (pair_r2(pi, pj, cx, cy, cz, two_h, d) <= r2cut) {
          s_list[wv][n][lane] = (unsigned(j) << 6) | unsigned(cur);
          n++;
        }
        j++;
      }
      int nmax = n;
      for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
      for (int k = 0; k < nmax; k++) {
        if (k < n) {
          const unsigned e = s_list[wv][k][lane];
          const int jj = int(e >> 6), cur = int(e & 63u);
          const int cell = cur >> 1, bsel = cur & 1;
          const int cz = cell / 9 - 1, cy = (cell / 3) % 3 - 1, cx = cell % 3 - 1;
          int slot;
          float4 pj;
          if constexpr (STAGED) {
            const int r = ((cy + 1) * 3 + (cx + 1)) * 2 + bsel;
            pj = s_pos[jj];
            slot = s_rs[r] + (jj - s_ro[r]);
          } else {
            pj = A.pos[jj];
            slot = jj;
          }
          float3 d;
          const float r = sqrtf(pair_r2(pi, pj, cx, cy, cz, two_h, d));
          const float q = r * i_h;
          pair_terms(K, S, d, r, q, vi, rhoi, sri, artpi, shift, bsel == 0, A.vel[slot], pj.w, rec_sra(B, slot),
                     rec_srb(B, slot));
        }
      }
      const bool more = (j < jend) || (cstep < 54);
      n = 0;
      if (__ballot(more) == 0ull) break;
    }
    if (active) interaction_epilogue(P, A, B, i, S, vi);
  }
}

template <int WAVES, int CAP, int MAXTZ, int NLCAP>
__global__ void __launch_bounds__(64 * WAVES) k_interaction_tile(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                                 const int* __restrict__ cbeg,
                                                                 const int* __restrict__ cend, int tz, int nzb,
                                                                 int ntile, int per_xcd) {
  constexpr int WG = 64 * WAVES;
  constexpr int NZW = MAXTZ + 2;
  __shared__ float4 s_pos[CAP];
  __shared__ int s_lo[18][NZW];
  __shared__ int s_hi[18][NZW];
  __shared__ int s_rs[18], s_rc[18], s_ro[18];
  __shared__ int s_home[2];
  __shared__ int s_total;
  __shared__ unsigned s_list[WAVES][NLCAP][64];

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntile) return;
  const int cnx = P->cnx, cny = P->cny, cnz = P->cnz, ncm = P->ncell_material;
  const int zb = tile % nzb;
  const int y = (tile / nzb) % cny;
  const int x = tile / (nzb * cny);
  const int z0 = zb * tz;
  const int z1 = min(z0 + tz, cnz);
  if (t == 0) {
    int hb = INT_MAX, he = -1;
    const int kb = (x * cny + y) * cnz + ncm;
    for (int z = z0; z < z1; z++) {
      const int bb = cbeg[kb + z];
      if (bb >= 0) {
        hb = min(hb, bb);
        he = max(he, cend[kb + z]);
      }
    }
    s_home[0] = hb;
    s_home[1] = he;
  }
  __syncthreads();
  const int hb = s_home[0], he = s_home[1];
  if (he < 0) return;
  const int zw0 = max(z0 - 1, 0), zw1 = min(z1, cnz - 1);
  const int nzw = zw1 - zw0 + 1;
  // ranges r = c*2 + type, c = (cy+1)*3 + (cx+1): slot span and per-cell runs
  if (t < 18) {
    const int r = t, c = r >> 1, type = r & 1;
    const int jx = x + c % 3 - 1, jy = y + c / 3 - 1;
    int rs = -1, re = -1;
    for (int zi = 0; zi < NZW; zi++) { s_lo[r][zi] = -1; s_hi[r][zi] = -1; }
    if (jx >= 0 && jy >= 0 && jx < cnx && jy < cny) {
      const int kb = (jx * cny + jy) * cnz + type * ncm;
      for (int z = zw0; z <= zw1; z++) {
        const int bb = cbeg[kb + z];
        if (bb >= 0) {
          const int ee = cend[kb + z];
          if (rs < 0) rs = bb;
          re = ee;
          s_lo[r][z - zw0] = bb;
          s_hi[r][z - zw0] = ee;
        }
      }
    }
    s_rs[r] = rs;
    s_rc[r] = rs >= 0 ? re - rs : 0;
  }
  __syncthreads();
  if (t == 0) {
    int off = 0;
    for (int r = 0; r < 18; r++) {
      s_ro[r] = off;
      off += s_rc[r];
    }
    s_total = off;
  }
  __syncthreads();
  const bool staged = s_total <= CAP;
  if (t < 18) {  // slot runs -> LDS runs (empty cells become [0, 0))
    const int r = t;
    for (int zi = 0; zi < NZW; zi++) {
      const int lo = s_lo[r][zi];
      if (lo >= 0) {
        s_lo[r][zi] = s_ro[r] + (lo - s_rs[r]);
        s_hi[r][zi] = s_ro[r] + (s_hi[r][zi] - s_rs[r]);
      } else {
        s_lo[r][zi] = 0;
        s_hi[r][zi] = 0;
      }
    }
  }
  if (staged) {
    for (int r = 0; r < 18; r++) {
      const int cnt = s_rc[r], so = s_rs[r], lo = s_ro[r];
      for (int k = t; k < cnt; k += WG) s_pos[lo + k] = A.pos[so + k];
    }
  }
  __syncthreads();

  if (staged)
    tile_home<true, WAVES, NLCAP, NZW>(P, A, B, cbeg, cend, s_pos, s_lo, s_hi, s_rs, s_ro, s_list, hb, he, zw0, nzw);
  else
    tile_home<false, WAVES, NLCAP, NZW>(P, A, B, cbeg, cend, s_pos, s_lo, s_hi, s_rs, s_ro, s_list, hb, he, zw0, nzw);
}

// Tile shapes the autotuner may choose from.
struct TileShape {
  int waves, cap, maxtz, nlcap;
};
static const TileShape kShapes[] = {{1, 1024, 6, 32}, {2, 2048, 10, 24}, {4, 3584, 16, 16}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

int interaction_tile_shapes() { return kNumShapes; }
int interaction_tile_maxtz(int shape) { return kShapes[shape].maxtz; }
int interaction_tile_waves(int shape) { return kShapes[shape].waves; }

void launch_interaction_tile(hipStream_t s, DevParams* P, int cnx, int cny, int cnz, Arrays A, const Buffers& B,
                             const int* cb, const int* ce, int shape, int tz) {
  const int nzb = (cnz + tz - 1) / tz;
  const int ntile = cnx * cny * nzb;
  const int per = (ntile + 7) / 8;
  const dim3 grid(8 * per);
  switch (shape) {
    case 0:
      hipLaunchKernelGGL((k_interaction_tile<1, 1024, 6, 32>), grid, dim3(64), 0, s, P, A, B, cb, ce, tz, nzb, ntile,
                         per);
      break;
    case 1:
      hipLaunchKernelGGL((k_interaction_tile<2, 2048, 10, 24>), grid, dim3(128), 0, s, P, A, B, cb, ce, tz, nzb,
                         ntile, per);
      break;
    default:
      hipLaunchKernelGGL((k_interaction_tile<4, 3584, 16, 16>), grid, dim3(256), 0, s, P, A, B, cb, ce, tz, nzb,
                         ntile, per);
      break;
  }
}

}  // namespace lq
