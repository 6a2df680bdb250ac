// lq_columns.hip -- the COLUMNS interaction variant: candidates from
// h-sized sub-cells instead of whole 2h cells.
//
// Every 2h cell splits into 2 x 2 x 2 sub-cells of edge h (sub-cell bit per
// axis: relative position >= h).  Per interaction call an auxiliary index
// lists every particle (wall and soil) grouped by 2h column (cx, cy), then
// by sub-column (sx, sy), then by sub-cell height z' = 2 cz + sz, walls before
// soil and slot order within a sub-cell.  A sub-column's sub-cells are thus
// contiguous along z, and sub_start holds the dense exclusive start of every
// (column, sub-column, z') bucket.  Each aux entry carries the particle's
// relative position and (slot << 2 | cz & 3).
//
// The interaction kernel then visits, per soil particle, the 5 x 5
// sub-columns within two sub-cells of its own; for each it takes only the
// z' window that can hold a particle closer than 2h (from the distance of
// the particle to the sub-column, with a 1% of h margin), one contiguous
// range of the aux index.  The pair predicate is the reference's
// r^2 <= r2_cut on the same FP32 operands (cell offsets from the 2h cells),
// so the neighbour set equals that of the 27-cell scan
// (particle_interaction.cu:117-140); the pairs are summed in column order
// instead of the reference's cell order, and the per-pair terms use the
// hardware reciprocal, square root, exp2 and log2 (one rounding each).  On
// average 180 candidates are tested per particle where the 2h-cell scan
// tests 450.
#include <algorithm>
#include "lq_device.hpp"
#include "lq_launch.hpp"

namespace lq {

__device__ __forceinline__ int sub_bits(float p, float h) { return p >= h ? 1 : 0; }

// Wall and soil counts of every 2h column (cx, cy): cells (cx, cy, 0..cnz-1),
// cpw consecutive columns per wave, its lanes over z (on a large grid of
// mostly empty columns, several columns per wave: fewer, longer waves).
__global__ void __launch_bounds__(256) k_col_count(const DevParams* __restrict__ P, const int* __restrict__ cbeg,
                                                   const int* __restrict__ cend, int2* __restrict__ colcnt, int cpw) {
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ncol = P->cnx * P->cny, cnz = P->cnz, ncm = P->ncell_material;
#pragma unroll 2
  for (int q = 0; q < cpw; q++) {
    const int col = wave * cpw + q;
    if (col >= ncol) break;  // uniform over the wave
    const int k0 = col * cnz;
    int nw = 0, ns = 0;
    for (int z = lane; z < cnz; z += 64) {
      const int bw = cbeg[k0 + z], bs = cbeg[ncm + k0 + z];
      if (bw >= 0) nw += cend[k0 + z] - bw;
      if (bs >= 0) ns += cend[ncm + k0 + z] - bs;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      nw += __shfl_xor(nw, o);
      ns += __shfl_xor(ns, o);
    }
    if (lane == 0) colcnt[col] = make_int2(nw, ns);
  }
}

// Exclusive scan of one int2 per thread over a 1024-thread workgroup; tot
// receives the workgroup's total.
__device__ __forceinline__ int2 col_block_scan(int2 v, int2& tot, int2* wsum) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v.x, y = v.y;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int a = __shfl_up(x, o), b = __shfl_up(y, o);
    if (lane >= o) { x += a; y += b; }
  }
  if (lane == 63) wsum[wv] = make_int2(x, y);
  __syncthreads();
  int px = 0, py = 0, tx = 0, ty = 0;
  for (int w = 0; w < 16; w++) {
    const int2 s = wsum[w];
    if (w < wv) { px += s.x; py += s.y; }
    tx += s.x; ty += s.y;
  }
  __syncthreads();
  tot = make_int2(tx, ty);
  return make_int2(px + x - v.x, py + y - v.y);
}

// Exclusive scans of the column counts: the column's first wall slot w0 and
// first soil slot s0 (columns are runs of consecutive keys, so in slot
// order), with its counts.  One workgroup walks a grid of up to
// COL_SCAN_ONE columns; beyond, per-tile totals (k_col_tile_sums), their
// scan (k_col_tile_scan) and the tiles with their offsets (k_col_apply).
constexpr int COL_SCAN_ONE = 8192;
__global__ void __launch_bounds__(1024) k_col_scan(const DevParams* __restrict__ P, const int2* __restrict__ colcnt,
                                                   int4* __restrict__ colinfo) {
  __shared__ int2 wsum[16];
  const int ncol = P->cnx * P->cny;
  int cw = 0, cs = P->npb;
  for (int c0 = 0; c0 < ncol; c0 += 1024) {
    const int c = c0 + threadIdx.x;
    const int2 v = c < ncol ? colcnt[c] : make_int2(0, 0);
    int2 tot;
    const int2 e = col_block_scan(v, tot, wsum);
    if (c < ncol) colinfo[c] = make_int4(cw + e.x, cs + e.y, v.x, v.y);
    cw += tot.x;
    cs += tot.y;
  }
}
__global__ void __launch_bounds__(1024) k_col_tile_sums(const DevParams* __restrict__ P,
                                                        const int2* __restrict__ colcnt, int2* __restrict__ tsum) {
  __shared__ int2 wsum[16];
  const int c = blockIdx.x * 1024 + threadIdx.x;
  const int2 v = c < P->cnx * P->cny ? colcnt[c] : make_int2(0, 0);
  int2 tot;
  col_block_scan(v, tot, wsum);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(1024) k_col_tile_scan(int2* __restrict__ tsum, int nt) {
  __shared__ int2 wsum[16];
  int cw = 0, cs = 0;
  for (int t0 = 0; t0 < nt; t0 += 1024) {
    const int t = t0 + threadIdx.x;
    const int2 v = t < nt ? tsum[t] : make_int2(0, 0);
    int2 tot;
    const int2 e = col_block_scan(v, tot, wsum);
    if (t < nt) tsum[t] = make_int2(cw + e.x, cs + e.y);
    cw += tot.x;
    cs += tot.y;
  }
}
__global__ void __launch_bounds__(1024) k_col_apply(const DevParams* __restrict__ P, const int2* __restrict__ colcnt,
                                                    const int2* __restrict__ toff, int4* __restrict__ colinfo) {
  __shared__ int2 wsum[16];
  const int ncol = P->cnx * P->cny;
  const int c = blockIdx.x * 1024 + threadIdx.x;
  const int2 v = c < ncol ? colcnt[c] : make_int2(0, 0);
  int2 tot;
  const int2 e = col_block_scan(v, tot, wsum);
  const int2 o = toff[blockIdx.x];
  if (c < ncol) colinfo[c] = make_int4(o.x + e.x, P->npb + o.y + e.y, v.x, v.y);
}

// One workgroup per 2h column: counts per bucket b = q * 2cnz + z'
// (q = 2 sx + sy, z' = 2 cz + sz) for walls and soil, the exclusive scan into
// sub_start, then each particle's place: its bucket's start, plus the
// bucket's wall count for soil, plus the particles of its own cell and kind
// before it in slot order that share its sub-cell (these are the run of
// equal cz just before it, as slots follow the keys).  The same pass writes
// the particle's pair record at its aux position, as four float4 planes of
// rec_rows entries: position with the mass in w (negative for contact
// walls), velocity and density, sigma/rho^2, m/rho and trace(sigma/rho^2)
// (k_pair_prep's quantities).  soil_aux lists the aux positions of the soil
// entries in aux order: the interaction's lanes.
// The pair record of slot j at aux position k (four planes of rec_rows).
__device__ __forceinline__ void col_write_record(const DevParams* __restrict__ P, Arrays A, const Buffers& B, int j,
                                                 int k, bool soil) {
  const float4 p = A.pos[j];
  const float4 v = A.vel[j];
  const float rho = v.w;
  const float m = soil ? fabsf(p.w) : P->vol_def * rho;
  const Sym sr = sdiv(ldsym(A.str_a, A.str_b, j), rho * rho);
  const size_t rows = size_t(B.rec_rows);
  B.rec[k] = make_float4(p.x, p.y, p.z, p.w < 0.f ? -m : m);
  B.rec[rows + k] = v;
  B.rec[2 * rows + k] = make_float4(sr.xx, sr.yy, sr.zz, sr.xy);
  B.rec[3 * rows + k] = make_float4(sr.xz, sr.yz, m / rho, strace(sr));
}

constexpr int COL_MAX_CNZ = 256;
constexpr int COL_MAX_BUCKETS = 8 * COL_MAX_CNZ + 1;
constexpr int COL_LDS_PTS = 8192;  // particles of a column whose codes stay in LDS
constexpr int COL_THREADS = 512;
// wall_rec == 0: wall records are left to k_col_wall_records (which runs
// after Adami has extrapolated the walls' velocity and stress); their aux
// positions go to wall_aux.
__global__ void __launch_bounds__(COL_THREADS) k_col_build(const DevParams* __restrict__ P, Arrays A, Buffers B,
                                                           const int4* __restrict__ colinfo, int wall_rec) {
  __shared__ int cntw[COL_MAX_BUCKETS];
  __shared__ int start[COL_MAX_BUCKETS];
  __shared__ int sstart[COL_MAX_BUCKETS];  // exclusive scan of the soil counts alone
  __shared__ unsigned char oct[COL_LDS_PTS];
  __shared__ unsigned char czs[COL_LDS_PTS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int cnz = P->cnz, npb = P->npb;
  const int col = blockIdx.x;
  const int4 ci = colinfo[col];
  const int w0 = ci.x, s0 = ci.y, nw = ci.z, n = ci.z + ci.w;
  const int nb = 8 * cnz;
  const float h = P->h;
  for (int b = tid; b <= nb; b += COL_THREADS) { cntw[b] = 0; start[b] = 0; sstart[b] = 0; }
  __syncthreads();
  auto slot_of = [&](int t) { return t < nw ? w0 + t : s0 + (t - nw); };
  auto oct_of = [&](float4 p) { return (sub_bits(p.x, h) << 2) | (sub_bits(p.y, h) << 1) | sub_bits(p.z, h); };
  auto bucket = [&](int o, int cz) { return (o >> 1) * (2 * cnz) + 2 * cz + (o & 1); };
  for (int t = tid; t < n; t += COL_THREADS) {
    const int j = slot_of(t);
    const int o = oct_of(A.pos[j]);
    const int cz = A.cellpos[j].z;
    if (t < COL_LDS_PTS) { oct[t] = (unsigned char)o; czs[t] = (unsigned char)cz; }
    atomicAdd(t < nw ? &cntw[bucket(o, cz)] : &start[bucket(o, cz)], 1);
  }
  __syncthreads();
  if (tid < 64) {  // exclusive scans of cntw + cnts and of cnts (start holds cnts) over nb + 1 buckets
    int carry = 0, scarry = 0;
    const size_t sbase = size_t(col) * (nb + 1);
    const int base = w0 + (s0 - npb);
    for (int b0 = 0; b0 <= nb; b0 += 64) {
      const int b = b0 + lane;
      const int vs = b <= nb ? start[b] : 0;
      const int v = b <= nb ? cntw[b] + vs : 0;
      int x = v, xs = vs;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o), ys = __shfl_up(xs, o);
        if (lane >= o) { x += y; xs += ys; }
      }
      if (b <= nb) {
        start[b] = carry + x - v;
        sstart[b] = scarry + xs - vs;
        B.sub_start[sbase + b] = base + carry + x - v;
      }
      carry += __shfl(x, 63);
      scarry += __shfl(xs, 63);
    }
  }
  __syncthreads();
  const int base = w0 + (s0 - npb);
  for (int t = tid; t < n; t += COL_THREADS) {
    const bool soil = t >= nw;
    const int j = slot_of(t);
    const float4 p = A.pos[j];
    int o, cz;
    if (t < COL_LDS_PTS) { o = oct[t]; cz = czs[t]; }
    else { o = oct_of(p); cz = A.cellpos[j].z; }
    const int b = bucket(o, cz);
    int rank = 0;
    for (int u = t - 1; u >= (soil ? nw : 0); u--) {
      int ou, cu;
      if (u < COL_LDS_PTS) { ou = oct[u]; cu = czs[u]; }
      else { const int ju = slot_of(u); ou = oct_of(A.pos[ju]); cu = A.cellpos[ju].z; }
      if (cu != cz) break;
      rank += ou == o;
    }
    const int k = base + start[b] + (soil ? cntw[b] : 0) + rank;
    B.aux[k] = make_float4(p.x, p.y, p.z, __int_as_float((j << 2) | (cz & 3)));
    if (soil) B.soil_aux[(s0 - npb) + sstart[b] + rank] = k;
    if (soil || wall_rec) col_write_record(P, A, B, j, k, soil);
    else B.wall_aux[j] = k;
  }
}

// The wall slots' pair records, at the aux positions k_col_build left.
__global__ void __launch_bounds__(256) k_col_wall_records(const DevParams* __restrict__ P, Arrays A, Buffers B) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P->npb) return;
  col_write_record(P, A, B, j, B.wall_aux[j], false);
}

// ------------------------------------------------------ the window ----
// Per-lane geometry of the sub-column scan.
struct ColHome {
  int xs, ys, zs;     // sub-cell coordinates of the particle (its 2h cell is xs >> 1, ...)
  float fx, fy, fz;   // position inside its sub-cell
};

__device__ __forceinline__ ColHome col_home(const DevParams* __restrict__ P, int4 ci, float3 pi) {
  const float h = P->h;
  ColHome g;
  const int sx = sub_bits(pi.x, h), sy = sub_bits(pi.y, h), sz = sub_bits(pi.z, h);
  g.xs = 2 * ci.x + sx; g.ys = 2 * ci.y + sy; g.zs = 2 * ci.z + sz;
  g.fx = sx ? pi.x - h : pi.x;
  g.fy = sy ? pi.y - h : pi.y;
  g.fz = sz ? pi.z - h : pi.z;
  return g;
}

// Lower bound of the distance along one axis from the particle (offset f in
// its sub-cell) to the sub-cell d sub-cells away, less the margin.
__device__ __forceinline__ float col_gap(int d, float f, float h, float margin) {
  float g = 0.f;
  if (d > 0) g = float(d) * h - f;
  else if (d < 0) g = f - float(d + 1) * h;
  return fmaxf(g - margin, 0.f);
}

// The aux range [k0, k1) of sub-column (dx, dy) around the particle, and the
// cell offsets ox, oy of that sub-column; empty when no particle in it can be
// within 2h.  zb0 and zb1 are the sub-column's aux starts of sub-cell
// heights 2 cz and 2 cz + 2 (cz the particle's own 2h cell level): inside
// the window, candidates before zb0 lie one 2h cell lower, those from zb1 on
// one higher (the window spans at most two sub-cells either side).
__device__ __forceinline__ void col_range(const DevParams* __restrict__ P, const int* __restrict__ sub_start,
                                          const ColHome& g, int dx, int dy, float gx2, float gy2, bool active,
                                          int& k0, int& k1, int& ox, int& oy, int& zb0, int& zb1) {
  k0 = k1 = 0;
  ox = oy = 0;
  zb0 = zb1 = 0;
  const int X = g.xs + dx, Y = g.ys + dy;
  const int cnz = P->cnz;
  if (!active || X < 0 || Y < 0 || X >= 2 * P->cnx || Y >= 2 * P->cny) return;
  const float i_h = P->i_h;
  const float rem = P->col_r2 - gx2 - gy2;
  if (!(rem > 0.f)) return;
  const float dz = __builtin_amdgcn_sqrtf(rem) + P->col_margin;
  const int zlo = max(g.zs + max(-2, int(floorf((g.fz - dz) * i_h))), 0);
  const int zhi = min(g.zs + min(2, int(floorf((g.fz + dz) * i_h))), 2 * cnz - 1);
  if (zlo > zhi) return;
  const int cxj = X >> 1, cyj = Y >> 1;
  ox = cxj - (g.xs >> 1);
  oy = cyj - (g.ys >> 1);
  const size_t base = size_t(cxj * P->cny + cyj) * size_t(8 * cnz + 1) + size_t((2 * (X & 1) + (Y & 1)) * (2 * cnz));
  k0 = sub_start[base + zlo];
  k1 = sub_start[base + zhi + 1];
  const int zc = g.zs & ~1;  // first sub-cell height of the particle's own 2h cell
  zb0 = sub_start[base + zc];
  zb1 = sub_start[base + zc + 2];
}

// z offset of an aux candidate from its stored cz & 3 and the particle's cz
// (|offset| <= 1 inside the window): 0 -> 0, 1 -> +1, 3 -> -1.
__device__ __forceinline__ int col_oz(int code, int cz) {
  const int d = (code - cz) & 3;
  return d == 3 ? -1 : d;
}

// ------------------------------------------------- pair terms (fast) ----
// The terms of particle_interaction.cu:66-111 with one rounding per
// reciprocal, square root and power: (W / W(dr))^(W(0)/W(dr)) as
// exp2(e log2(W / W(dr))).
// wj = the neighbour's mass (the Adami volume mass vol_def rho for walls),
// negative for contact walls.
__device__ __forceinline__ void pair_terms_fast(const PairConst& P, float i_kdr, IAcc& S, float3 d, float r, float q,
                                                float3 vi, float rhoi, Sym sri, float artpi, int shift, float4 vj,
                                                float wj, float4 sra, float4 srb) {
  const bool contactj = wj < 0.f;
  const float t = 2.0f - q;
  const float dwdr = P.kgrad * (t * t * t);
  const float3 g = f3mul(d, dwdr);
  const float rhoj = vj.w;
  const float mj = fabsf(wj);
  float3 dv = f3sub(vi, ld3(vj));
  if (contactj && shift == 1) dv = make_float3(0.f, 0.f, 0.f);
  const float volj = srb.z;
  const float3 dvv = f3mul(dv, volj);
  S.gv[0] = fmaf(-dvv.x, g.x, S.gv[0]); S.gv[1] = fmaf(-dvv.x, g.y, S.gv[1]); S.gv[2] = fmaf(-dvv.x, g.z, S.gv[2]);
  S.gv[3] = fmaf(-dvv.y, g.x, S.gv[3]); S.gv[4] = fmaf(-dvv.y, g.y, S.gv[4]); S.gv[5] = fmaf(-dvv.y, g.z, S.gv[5]);
  S.gv[6] = fmaf(-dvv.z, g.x, S.gv[6]); S.gv[7] = fmaf(-dvv.z, g.y, S.gv[7]); S.gv[8] = fmaf(-dvv.z, g.z, S.gv[8]);
  const float dvg = fmaf(dv.z, g.z, fmaf(dv.y, g.y, dv.x * g.x));
  const float keep = contactj ? 0.f : 1.f;
  S.accden = fmaf(keep * rhoi * dvg, volj, S.accden);
  const float sxx = sri.xx + sra.x, syy = sri.yy + sra.y, szz = sri.zz + sra.z;
  const float sxy = sri.xy + sra.w, sxz = sri.xz + srb.x, syz = sri.yz + srb.y;
  const float fx = fmaf(sxz, g.z, fmaf(sxy, g.y, sxx * g.x));
  const float fy = fmaf(syz, g.z, fmaf(syy, g.y, sxy * g.x));
  const float fz = fmaf(szz, g.z, fmaf(syz, g.y, sxz * g.x));
  const float mf = keep * mj;
  float s = 0.f;
  if (P.viscosity_type == 1) {
    // m_j alpha cs h (dv.d) / ((r^2 + eta) 0.5 (rho_i + rho_j)), approaching pairs only
    const float visc = fmaf(dv.z, d.z, fmaf(dv.y, d.y, dv.x * d.x));
    const float den = fmaf(r, r, P.eta) * (rhoi + rhoj);
    const float sv = 2.0f * P.alpha_cs * P.h * visc * __builtin_amdgcn_rcpf(den);
    s = (visc < 0.f) ? mf * sv : 0.f;
  }
  S.as.x = fmaf(fx, mf, S.as.x); S.as.y = fmaf(fy, mf, S.as.y); S.as.z = fmaf(fz, mf, S.as.z);
  S.av.x = fmaf(g.x, s, S.av.x); S.av.y = fmaf(g.y, s, S.av.y); S.av.z = fmaf(g.z, s, S.av.z);
  if (P.art_p != 0.f) {
    const float trj = srb.w;
    float artp = artpi + (trj > 0.f ? trj * P.art_p : 0.f);
    if (artp > 0.f && !contactj) {
      const float ker = P.knorm * (t * t) * (t * t) * (q + 0.5f);
      artp = artp * __builtin_amdgcn_exp2f(P.art_exp * __builtin_amdgcn_logf(ker * i_kdr));
      const float sa = mj * artp;
      S.aa.x = fmaf(-g.x, sa, S.aa.x); S.aa.y = fmaf(-g.y, sa, S.aa.y); S.aa.z = fmaf(-g.z, sa, S.aa.z);
    }
  }
}

// ----------------------------------------------------- interaction ----
// One lane per soil aux entry (soil_aux, in aux order), so the 64 lanes of
// a wave are consecutive soil particles of one sub-column: for every window
// they read neighbouring stretches of the same target sub-column, and their pair records (stored in
// aux order, one plane per float4, so that 8 lanes share a 128-byte line)
// lie close together.  Hits
// go to per-lane FIFO rings in LDS as (aux index << 2 | z cell-offset code:
// the window fixes the x and y offsets of every entry, as the rings hold
// only the current window's hits) and
// are evaluated in ring order, as in the neighbour-list variant; the rings
// are emptied at the end of every window (and earlier when one would
// overflow), so a window's records are summed while they are still in the
// caches.  Draining earlier or later changes no sum: each lane's hits are
// evaluated in the order they were found.
constexpr int CL_CAP = 16;    // ring entries per lane (power of two): 4 KB of LDS per wave
constexpr int CL_CHUNK = 16;  // candidates per hit mask

// offx, offy: the window's cell offsets along x and y (times 2h)
__device__ __forceinline__ void cl_eval(const PairConst& K, float i_kdr, IAcc& S, const Buffers& B, int e, float3 pi,
                                        float offx, float offy, float two_h, float i_h, float3 vi, float rhoi,
                                        Sym sri, float artpi, int shift) {
  const size_t rows = size_t(B.rec_rows), k = unsigned(e) >> 2;
  const float4 np_ = B.rec[k], nv = B.rec[rows + k], na = B.rec[2 * rows + k], nb = B.rec[3 * rows + k];
  float3 d;
  d.x = (pi.x - np_.x) - offx;
  d.y = (pi.y - np_.y) - offy;
  d.z = (pi.z - np_.z) - fmaf(float(e & 3), two_h, -two_h);
  const float r2 = fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x));
  const float r = __builtin_amdgcn_sqrtf(r2);
  pair_terms_fast(K, i_kdr, S, d, r, r * i_h, vi, rhoi, sri, artpi, shift, nv, np_.w, na, nb);
}

__device__ __forceinline__ void cl_drain(const PairConst& K, float i_kdr, IAcc& S, const Buffers& B,
                                         const int (*lst)[64], int lane, int head, int pop, float3 pi, float offx,
                                         float offy, float two_h, float i_h, float3 vi, float rhoi, Sym sri,
                                         float artpi, int shift) {
#pragma unroll 1
  for (int k = 0; k < pop; k++)
    cl_eval(K, i_kdr, S, B, lst[(head + k) & (CL_CAP - 1)][lane], pi, offx, offy, two_h, i_h, vi, rhoi, sri, artpi,
            shift);
}

// MINW waves per SIMD: 5 holds the kernel to 96 VGPRs, 4 allows 128 (the
// forms compile to 95-100, the steps' fuse-3 form to 96 either way);
// lq_autotune times both.  FUSE 1 / 2: a particle whose
// contact force needs no pass B goes on through cuPredictorMaterial /
// cuCorrectorMaterial (predictor_slot / corrector_slot) right after its
// pair sum.  FUSE 3: the kernel reads nothing the corner search writes
// (the slip flag it reads, cvar.z, is written back unchanged there), so it
// can run beside that search; the wall contact of particles near a wall
// is then finished by k_contact_after.  Nothing in this kernel reads another particle's state arrays
// (neighbours come from aux and rec), so those updates race with nothing;
// a first-layer particle first leaves its position and cell in fl_pos /
// fl_cell for pass B.
template <int MINW, int FUSE>
__global__ void __launch_bounds__(64, MINW) k_interaction_col(DevParams* __restrict__ P, Arrays A, Buffers B,
                                                              int remap) {
  __shared__ int s_list[CL_CAP][64];
  const int lane = threadIdx.x;
  const int l = xcd_block(remap) * 64 + lane;
  const bool active = l < P->npm;
  const int kh = active ? B.soil_aux[l] : 0;
  const int i = __float_as_int(B.aux[kh].w) >> 2;
  const size_t rows = size_t(B.rec_rows);
  const float4 ph = B.rec[kh], vri = B.rec[rows + kh], sria = B.rec[2 * rows + kh], srib = B.rec[3 * rows + kh];
  const float3 pi = ld3(ph);
  const int4 ci = A.cellpos[i];
  const float3 vi = ld3(vri);
  const float rhoi = vri.w;
  const Sym sri = {sria.x, sria.y, sria.z, sria.w, srib.x, srib.y};
  const float artpi = srib.w > 0.f ? srib.w * P->art_p : 0.f;
  const int shift = int(A.cvar[i].z);
  const float two_h = P->two_h, i_h = P->i_h, r2cut = P->r2cut, h = P->h, margin = P->col_margin;
  const float i_kdr = __builtin_amdgcn_rcpf(P->kernel_dr);
  const PairConst K = pair_const(P);
  IAcc S;
  S.accden = 0.f;
  S.as = S.av = S.aa = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int q = 0; q < 9; q++) S.gv[q] = 0.f;
  const ColHome g = col_home(P, ci, pi);
  int head = 0, cnt = 0;
  // window a + 1's range is fetched while window a is scanned
  auto window = [&](int a, int& j, int& e, int& ox, int& oy, int& zb0, int& zb1) {
    const int dx = a / 5 - 2, dy = a % 5 - 2;
    const float gx = col_gap(dx, g.fx, h, margin), gy = col_gap(dy, g.fy, h, margin);
    col_range(P, B.sub_start, g, dx, dy, gx * gx, gy * gy, active && a < 25, j, e, ox, oy, zb0, zb1);
  };
  int nj, ne, nox, noy, nzb0, nzb1;
  window(0, nj, ne, nox, noy, nzb0, nzb1);
#pragma unroll 1
  for (int a = 0; a < 25; a++) {
    int j = nj, e = ne;
    const int ox = nox, oy = noy, zb0 = nzb0, zb1 = nzb1;
    window(a + 1, nj, ne, nox, noy, nzb0, nzb1);
    const float offx = float(ox) * two_h, offy = float(oy) * two_h;
    while (j < e) {
      const int m = min(e - j, CL_CHUNK);
      unsigned mask = 0u;
#pragma unroll 4
      for (int t = 0; t < m; t++) {
        const float4 c = B.aux[j + t];
        // the candidate's cell lies below, at or above the particle's own
        // as its aux position lies before zb0, in [zb0, zb1) or after
        const float offz = j + t < zb0 ? -two_h : (j + t < zb1 ? 0.f : two_h);
        float3 d;
        d.x = (pi.x - c.x) - offx;
        d.y = (pi.y - c.y) - offy;
        d.z = (pi.z - c.z) - offz;
        if (fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x)) <= r2cut) mask |= 1u << t;
      }
      const int over = cnt + __popc(mask) - CL_CAP;
      if (__ballot(over > 0)) {
        int need = 0;
#pragma unroll
        for (int b = CL_CHUNK; b; b >>= 1)
          if (__ballot(over >= need + b)) need += b;
        const int pop = min(cnt, need);
        cl_drain(K, i_kdr, S, B, s_list, lane, head, pop, pi, offx, offy, two_h, i_h, vi, rhoi, sri, artpi, shift);
        head = (head + pop) & (CL_CAP - 1);
        cnt -= pop;
      }
      while (mask) {
        const int t = __ffs(mask) - 1;
        mask &= mask - 1u;
        const unsigned oz = unsigned(j + t >= zb0) + unsigned(j + t >= zb1);  // oz + 1
        s_list[(head + cnt) & (CL_CAP - 1)][lane] = int((unsigned(j + t) << 2) | oz);
        cnt++;
      }
      j += m;
    }
    // drain at the window's end, while its records are still near in the caches
    cl_drain(K, i_kdr, S, B, s_list, lane, head, cnt, pi, offx, offy, two_h, i_h, vi, rhoi, sri, artpi, shift);
    head = (head + cnt) & (CL_CAP - 1);
    cnt = 0;
  }
  if (!active) return;
  if constexpr (FUSE == 3) {
    // contact after the corner search: a particle whose cell has a wall
    // particle within its 27 cells keeps its sums for k_contact_after; any
    // other has no contact wall within 2h (the search will find none)
    const float3 s3 = interaction_rates(B, i, S);
    if (B.wall_near[grid_key(P, ci.x, ci.y, ci.z)]) B.s3[i] = make_float4(s3.x, s3.y, s3.z, S.accden);
    else interaction_free(P, A, B, i, s3, S.accden, vi);
    return;
  }
  const bool deferred = interaction_epilogue(P, A, B, i, S, vi);
  if constexpr (FUSE == 1 || FUSE == 2) {
    if (A.cvar[i].y <= P->band) {
      B.fl_pos[i] = A.pos[i];
      B.fl_cell[i] = ci;
    }
    if (!deferred) {
      if constexpr (FUSE == 1) predictor_slot(P, A, B, i);
      else corrector_slot(P, A, B, i);
    }
  }
}

// Neighbour lists through the same windows (parity and debugging): for soil
// slot npb + s the original ids of its neighbours in column order.
__global__ void k_neighbor_list_col(const DevParams* __restrict__ P, Arrays A, Buffers B, int max_nb,
                                    int* __restrict__ counts, int* __restrict__ ids) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const int i = P->npb + idx;
  const float3 pi = ld3(A.pos[i]);
  const int4 ci = A.cellpos[i];
  const float two_h = P->two_h, h = P->h, margin = P->col_margin;
  const ColHome g = col_home(P, ci, pi);
  int n = 0;
  for (int a = 0; a < 25; a++) {
    const int dx = a / 5 - 2, dy = a % 5 - 2;
    const float gx = col_gap(dx, g.fx, h, margin), gy = col_gap(dy, g.fy, h, margin);
    int j, e, ox, oy, zb0, zb1;
    col_range(P, B.sub_start, g, dx, dy, gx * gx, gy * gy, true, j, e, ox, oy, zb0, zb1);
    for (; j < e; j++) {
      const float4 c = B.aux[j];
      const int w = __float_as_int(c.w);
      float3 d;
      d.x = (pi.x - c.x) - float(ox) * two_h;
      d.y = (pi.y - c.y) - float(oy) * two_h;
      d.z = (pi.z - c.z) - float(col_oz(w, ci.z)) * two_h;
      if (!(fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x)) <= P->r2cut)) continue;
      if (n < max_nb) ids[size_t(idx) * max_nb + n] = A.orig_id[w >> 2];
      n++;
    }
  }
  counts[idx] = n;
}

static inline int cblocks(long long n, int b) { return int((n + b - 1) / b); }

void launch_col_index(hipStream_t s, const DevParams* P, int cnx, int cny, Arrays A, const Buffers& B, int wall_rec) {
  const int ncol = cnx * cny;
  const int cpw = ncol <= COL_SCAN_ONE ? 1 : 8;
  hipLaunchKernelGGL(k_col_count, dim3(cblocks(ncol, 4 * cpw)), dim3(256), 0, s, P, (const int*)B.cell_beg,
                     (const int*)B.cell_end, B.colcnt, cpw);
  if (ncol <= COL_SCAN_ONE) {
    hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(1024), 0, s, P, (const int2*)B.colcnt, B.colinfo);
  } else {
    const int nt = cblocks(ncol, 1024);
    hipLaunchKernelGGL(k_col_tile_sums, dim3(nt), dim3(1024), 0, s, P, (const int2*)B.colcnt, B.coltile);
    hipLaunchKernelGGL(k_col_tile_scan, dim3(1), dim3(1024), 0, s, B.coltile, nt);
    hipLaunchKernelGGL(k_col_apply, dim3(nt), dim3(1024), 0, s, P, (const int2*)B.colcnt, (const int2*)B.coltile,
                       B.colinfo);
  }
  hipLaunchKernelGGL(k_col_build, dim3(ncol), dim3(COL_THREADS), 0, s, P, A, B, (const int4*)B.colinfo, wall_rec);
}
void launch_col_wall_records(hipStream_t s, const DevParams* P, int npb, Arrays A, const Buffers& B) {
  if (npb > 0) hipLaunchKernelGGL(k_col_wall_records, dim3(cblocks(npb, 256)), dim3(256), 0, s, P, A, B);
}
// lanes over the soil aux entries (at most npm_max; the live count is P->npm)
template <int FUSE>
static void launch_col(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, int xcd_remap, int wide) {
  const dim3 grid(cblocks(npm_max, 64));
  if (wide) hipLaunchKernelGGL((k_interaction_col<4, FUSE>), grid, dim3(64), 0, s, P, A, B, xcd_remap);
  else hipLaunchKernelGGL((k_interaction_col<5, FUSE>), grid, dim3(64), 0, s, P, A, B, xcd_remap);
}
void launch_interaction_col(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B,
                            int xcd_remap, int wide, int fuse) {
  if (npm_max <= 0) return;
  if (fuse == 1) launch_col<1>(s, P, npm_max, A, B, xcd_remap, wide);
  else if (fuse == 2) launch_col<2>(s, P, npm_max, A, B, xcd_remap, wide);
  else if (fuse == 3) launch_col<3>(s, P, npm_max, A, B, xcd_remap, wide);
  else launch_col<0>(s, P, npm_max, A, B, xcd_remap, wide);
}
void launch_neighbor_list_col(hipStream_t s, const DevParams* P, int npm_max, Arrays A, const Buffers& B, int max_nb,
                              int* counts, int* ids) {
  if (npm_max > 0)
    hipLaunchKernelGGL(k_neighbor_list_col, dim3(cblocks(npm_max, 256)), dim3(256), 0, s, P, A, B, max_nb, counts,
                       ids);
}
int col_max_cnz() { return COL_MAX_CNZ; }

}  // namespace lq
