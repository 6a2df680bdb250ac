// lq_grid.hip -- cell hash, stable LSD radix sort, cell ranges, reorder.
//
// Replaces LoquatGrid::NeighborSearch (loquat_grid.cu:147-191), whose sort is
// CUB DeviceRadixSort::SortPairs over all 32 key bits.  Here the sort is a
// hand-written stable LSD radix sort over 8-bit digits that runs only over
// the bits the largest key needs (ceil(log2(ncell))), with per-tile digit
// ranking done by 64-lane wave ballots.
#include "lq_internal.hpp"
#include "lq_launch.hpp"

namespace lq {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

__device__ __forceinline__ int cell_key(const DevParams* P, int4 c) { return grid_key(P, c.x, c.y, c.z); }

// cuGetParticleCellIndex, loquat_grid.cu:71-95.  Also flags whether any key
// differs from the previous sorted key at the same slot; with mflag, the
// flags of slots 0..np (slot np: 0) and their per-1024 block sums (the first
// pass of the flag scan).  The same launch empties the cell ranges
// (the memset of loquat_grid.cu:159-160) over ncell keys:
// nothing reads them again before k_beg_end.  1024 threads per block.
constexpr int HASH_THREADS = 1024;
__global__ void __launch_bounds__(HASH_THREADS) k_hash(DevParams* __restrict__ P, Arrays A, int* __restrict__ keys,
                                                       int* __restrict__ vals, const int* __restrict__ prev_sorted,
                                                       int* __restrict__ mflag, int* __restrict__ blocksum,
                                                       int ncell, int* __restrict__ cbeg, int* __restrict__ cend) {
  __shared__ int sh_wave[HASH_THREADS / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = P->np;
  if (i < ncell) {
    cbeg[i] = -1;
    cend[i] = -1;
  }
  int flag = 0;
  if (i < np) {
    int4 c = A.cellpos[i];
    int key;
    if (c.w == P->ncell_delete_m) {
      key = P->ncell_delete_m;
    } else {
      key = cell_key(P, c);
      if (A.idm[i].x >= 100.0f) key += P->ncell_material;  // MATERIAL_NUMBER, types.h:37
      c.w = key;
      A.cellpos[i] = c;
    }
    keys[i] = key;
    vals[i] = i;
    if (prev_sorted != nullptr) {
      const bool moved = key != prev_sorted[i];
      if (moved) P->sort_changed = 1;
      flag = moved ? 1 : 0;
      if (mflag) mflag[i] = flag;
    }
  } else if (i == np && mflag) {
    mflag[np] = 0;
  }
  if (blocksum && blockIdx.x * HASH_THREADS <= np) {  // blocks holding flags 0..np
    int v = flag;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < HASH_THREADS / 64; w++) t += sh_wave[w];
      blocksum[blockIdx.x] = t;
    }
  }
}

__global__ void k_fill_int(int* __restrict__ a, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_clear_flag(DevParams* P) { P->sort_changed = 0; }

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const int* __restrict__ keys, int n, int shift,
                                                           int* __restrict__ hist, int nblocks) {
  __shared__ int cnt[256];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int k = 0; k < RS_ITEMS; k++) {
    int idx = base + k * RS_THREADS + t;
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[t * nblocks + blockIdx.x] = cnt[t];
}

// In-place exclusive scan of the radix histogram (n ints) in three
// launches: per-tile sums of SC_TILE consecutive ints (coalesced int4
// loads), one block scanning the tile sums, and a tile-local scan that adds
// the tile's offset.  sums holds ceil(n / SC_TILE) ints.
constexpr int SC_THREADS = 256;
constexpr int SC_TILE = SC_THREADS * 4;

// Block-wide inclusive scan of one int per thread (SC_THREADS threads).
__device__ __forceinline__ int block_incl_scan(int v, int* sh_wave) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) sh_wave[w] = v;
  __syncthreads();
  int add = 0;
  for (int k = 0; k < w; k++) add += sh_wave[k];
  __syncthreads();
  return v + add;
}

__device__ __forceinline__ int4 load_tile4(const int* __restrict__ data, int n, int base) {
  const int k = base + 4 * int(threadIdx.x);
  if (k + 3 < n) return *reinterpret_cast<const int4*>(data + k);
  return make_int4(k < n ? data[k] : 0, k + 1 < n ? data[k + 1] : 0, k + 2 < n ? data[k + 2] : 0,
                   k + 3 < n ? data[k + 3] : 0);
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tile_sums(const int* __restrict__ data, int n,
                                                               int* __restrict__ sums) {
  __shared__ int sh_wave[SC_THREADS / 64];
  const int4 v = load_tile4(data, n, blockIdx.x * SC_TILE);
  const int tot = block_incl_scan(v.x + v.y + v.z + v.w, sh_wave);
  if (threadIdx.x == SC_THREADS - 1) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_sums(int* __restrict__ sums, int ntiles) {
  __shared__ int sh_wave[SC_THREADS / 64];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < ntiles; base += SC_THREADS) {
    const int k = base + int(threadIdx.x);
    const int v = k < ntiles ? sums[k] : 0;
    const int inc = block_incl_scan(v, sh_wave);
    const int c = carry;
    if (k < ntiles) sums[k] = c + inc - v;
    __syncthreads();
    if (threadIdx.x == SC_THREADS - 1) carry = c + inc;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tiles(int* __restrict__ data, int n,
                                                           const int* __restrict__ sums) {
  __shared__ int sh_wave[SC_THREADS / 64];
  const int base = blockIdx.x * SC_TILE;
  const int4 v = load_tile4(data, n, base);
  const int own = v.x + v.y + v.z + v.w;
  const int excl = block_incl_scan(own, sh_wave) - own + sums[blockIdx.x];
  const int o0 = excl, o1 = o0 + v.x, o2 = o1 + v.y, o3 = o2 + v.z;
  const int k = base + 4 * int(threadIdx.x);
  if (k + 3 < n) {
    *reinterpret_cast<int4*>(data + k) = make_int4(o0, o1, o2, o3);
  } else {
    if (k < n) data[k] = o0;
    if (k + 1 < n) data[k + 1] = o1;
    if (k + 2 < n) data[k + 2] = o2;
  }
}

static int scan_tiles(int n) { return (n + SC_TILE - 1) / SC_TILE; }

// The flag scan's last pass over tiles of HASH_THREADS flags (one per
// thread): the exclusive scan in place (sums: the scanned block sums of
// k_hash) and the movers compacted in slot order into mkey / mslot.
__global__ void __launch_bounds__(HASH_THREADS) k_scan_movers(int* __restrict__ mflag, int n,
                                                              const int* __restrict__ sums,
                                                              const int* __restrict__ keys, int np,
                                                              int* __restrict__ mkey, int* __restrict__ mslot) {
  __shared__ int sh_wave[HASH_THREADS / 64];
  const int k = blockIdx.x * HASH_THREADS + threadIdx.x;
  const int v = k < n ? mflag[k] : 0;
  const int excl = block_incl_scan(v, sh_wave) - v + sums[blockIdx.x];
  if (k < n) mflag[k] = excl;
  if (v && k < np) {
    mkey[excl] = keys[k];
    mslot[excl] = k;
  }
}

// wall_near[c] = 1 when any of the 27 cells around cell c (its key) holds
// a wall particle, else 0 (also for Morton keys outside the grid).
__device__ void delete_count(DevParams* P, const int* __restrict__ cbeg, const int* __restrict__ cend);
// Thread 0 also does the deleted-particle bookkeeping (k_delete_count),
// which reads the same finished cell ranges.
__global__ void __launch_bounds__(256) k_wall_near(DevParams* __restrict__ P, const int* __restrict__ cbeg,
                                                   const int* __restrict__ cend, int* __restrict__ wall_near) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) delete_count(P, cbeg, cend);
  if (c >= P->ncell_material) return;
  int cx, cy, cz;
  if (P->key_mode == LQ_KEY_MORTON) {
    cx = compact3(c >> 2); cy = compact3(c >> 1); cz = compact3(c);
    if (!in_grid(P, cx, cy, cz)) { wall_near[c] = 0; return; }
  } else {
    cz = c % P->cnz; cy = (c / P->cnz) % P->cny; cx = c / (P->cny * P->cnz);
  }
  int any = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++)
        if (in_grid(P, cx + dx, cy + dy, cz + dz) && cbeg[grid_key(P, cx + dx, cy + dy, cz + dz)] >= 0) any = 1;
  wall_near[c] = any;
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(const int* __restrict__ kin, const int* __restrict__ vin,
                                                              int* __restrict__ kout, int* __restrict__ vout, int n,
                                                              int shift, const int* __restrict__ offs, int nblocks) {
  __shared__ int run[256];
  __shared__ int wcnt[RS_THREADS / 64][256];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  run[t] = offs[t * nblocks + blockIdx.x];
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int base = blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; k++) {
    const int idx = base + k * RS_THREADS + t;
    const bool valid = idx < n;
    const int key = valid ? kin[idx] : 0;
    const int d = (key >> shift) & 255;
#pragma unroll
    for (int ww = 0; ww < RS_THREADS / 64; ww++) wcnt[ww][t] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rank;
      for (int ww = 0; ww < w; ww++) pos += wcnt[ww][d];
      kout[pos] = key;
      vout[pos] = vin[idx];
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int ww = 0; ww < RS_THREADS / 64; ww++) add += wcnt[ww][t];
    run[t] += add;
    __syncthreads();
  }
}

// ---- merge search -------------------------------------------------------
// The stable sort by (key, slot) after a previous sort: slot i held key
// prev[i] then (prev is sorted), and between two steps few particles change
// cell.  Slots whose key is unchanged ("stayers") are already in (key, slot)
// order; the others ("movers") are compacted in slot order, sorted stably by
// key, and the two sorted sequences are merged by rank.  The result equals
// the full radix sort for any inputs; it is fast when movers are few.

// Stable LSD radix sort of the M movers by key in one workgroup, over 8-bit
// digits, tiles of 1024 ranked by wave ballots as in k_radix_scatter.  M is
// read on the device; any M is sorted correctly (large M slowly).
constexpr int MS_THREADS = 1024;
__global__ void __launch_bounds__(MS_THREADS) k_mover_sort(const int* __restrict__ mtotal, int* k0, int* v0, int* k1,
                                                           int* v1, int nbits) {
  __shared__ int cnt[256];
  __shared__ int run[256];
  __shared__ int wcnt[MS_THREADS / 64][256];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int M = *mtotal;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  int *kin = k0, *vin = v0, *kout = k1, *vout = v1;
  for (int shift = 0; shift < nbits; shift += 8) {
    if (t < 256) cnt[t] = 0;
    __syncthreads();
    for (int i = t; i < M; i += MS_THREADS) atomicAdd(&cnt[(kin[i] >> shift) & 255], 1);
    __syncthreads();
    if (t == 0) {
      int sum = 0;
      for (int d = 0; d < 256; d++) {
        run[d] = sum;
        sum += cnt[d];
      }
    }
    __syncthreads();
    for (int base = 0; base < M; base += MS_THREADS) {
      for (int k = t; k < (MS_THREADS / 64) * 256; k += MS_THREADS) wcnt[k >> 8][k & 255] = 0;
      __syncthreads();
      const int idx = base + t;
      const bool valid = idx < M;
      const int key = valid ? kin[idx] : 0;
      const int val = valid ? vin[idx] : 0;
      const int d = (key >> shift) & 255;
      unsigned long long peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1;
        const unsigned long long m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const int rank = __popcll(peers & lt);
      if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
      __syncthreads();
      if (valid) {
        int pos = run[d] + rank;
        for (int ww = 0; ww < w; ww++) pos += wcnt[ww][d];
        kout[pos] = key;
        vout[pos] = val;
      }
      __syncthreads();
      if (t < 256) {
        int add = 0;
        for (int ww = 0; ww < MS_THREADS / 64; ww++) add += wcnt[ww][t];
        run[t] += add;
      }
      __syncthreads();
    }
    int* tk = kin; kin = kout; kout = tk;
    int* tv = vin; vin = vout; vout = tv;
  }
}

// Final position of every slot: a stayer i goes after the stayers before it
// and the movers (key, slot) < (key_i, i); the t-th sorted mover goes after
// t movers and the stayers (key, slot) < its own, found by binary search of
// its key in prev (stayers' keys are their prev entries).
__global__ void k_merge_place(const DevParams* __restrict__ P, const int* __restrict__ keys,
                              const int* __restrict__ prev, const int* __restrict__ mscan,
                              const int* __restrict__ smkey, const int* __restrict__ smslot, int* __restrict__ out_key,
                              int* __restrict__ out_idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = P->np;
  if (i >= np) return;
  const int M = mscan[np];
  if (mscan[i + 1] == mscan[i]) {
    const int key = keys[i];
    int lo = 0, hi = M;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int mk = smkey[mid];
      if (mk < key || (mk == key && smslot[mid] < i)) lo = mid + 1;
      else hi = mid;
    }
    const int pos = (i - mscan[i]) + lo;
    out_key[pos] = key;
    out_idx[pos] = i;
  }
  if (i < M) {
    const int key = smkey[i], slot = smslot[i];
    int lo = 0, hi = np;
    while (lo < hi) {  // first prev >= key
      const int mid = (lo + hi) >> 1;
      if (prev[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    int up = lo, uh = np;
    while (up < uh) {  // first prev > key
      const int mid = (up + uh) >> 1;
      if (prev[mid] <= key) up = mid + 1;
      else uh = mid;
    }
    const int p = max(lo, min(slot, up));
    const int pos = i + (p - mscan[p]);
    out_key[pos] = key;
    out_idx[pos] = slot;
  }
}

// cuFindBegEndInCell, loquat_grid.cu:116-144 (half-open runs).
__global__ void k_beg_end(const DevParams* __restrict__ P, const int* __restrict__ sorted, int* __restrict__ cbeg,
                          int* __restrict__ cend) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P->np;
  if (i >= n) return;
  int c = sorted[i];
  if (i == 0 || c != sorted[i - 1]) {
    cbeg[c] = i;
    if (i > 0) cend[sorted[i - 1]] = i;
  }
  if (i == n - 1) cend[c] = n;
}

// Deleted-particle bookkeeping, loquat_grid.cu:165-185, kept on the device.
__device__ void delete_count(DevParams* P, const int* __restrict__ cbeg, const int* __restrict__ cend) {
  const int nd = P->ncell_delete_m;
  int beg = cbeg[nd];
  if (beg != -1) {
    int now = cend[nd] - beg - P->np_delete_m;
    P->npm -= now;
    P->np_delete_m += now;
  }
  if (P->npb + P->npm + P->np_delete_m != P->np) set_err(P, DERR_DELETE);
}
__global__ void k_delete_count(DevParams* P, const int* __restrict__ cbeg, const int* __restrict__ cend) {
  delete_count(P, cbeg, cend);
}

// SortArraysPrimary / SortArraysAll (loquat.cu:254-287) as one gather of
// every per-particle array into the alternate buffer set.  Rows past
// npb+npm keep their stress/strain (the reference gathers only npb+npm rows).
__global__ void __launch_bounds__(256) k_gather(const DevParams* __restrict__ P, Arrays src, Arrays dst,
                                                const int* __restrict__ idx, Buffers B, int with_ini) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P->np) return;
  const int s = idx[i];
  dst.codes[i] = src.codes[s];
  dst.cellpos[i] = src.cellpos[s];
  dst.orig_id[i] = src.orig_id[s];
  dst.case_id[i] = src.case_id[s];
  dst.pres[i] = src.pres[s];
  dst.pos[i] = src.pos[s];
  dst.pos_ori[i] = src.pos_ori[s];
  dst.vel[i] = src.vel[s];
  dst.idm[i] = src.idm[s];
  dst.cvar[i] = src.cvar[s];
  dst.cf[i] = src.cf[s];
  dst.cn[i] = src.cn[s];
  dst.cw[i] = src.cw[s];
  dst.normal[i] = i < P->npb ? src.normal[s] : src.normal[i];
  const bool st = i < P->npb + P->npm;
  const int ss = st ? s : i;
  dst.str_a[i] = src.str_a[ss];
  dst.str_b[i] = src.str_b[ss];
  dst.stn_a[i] = src.stn_a[ss];
  dst.stn_b[i] = src.stn_b[ss];
}

// The same gather applied in place to the slots the sort moved: between two
// steps few particles change cell, so the sorted order is the identity on
// most slots.  Phase 0 writes the gathered rows of slots with idx[i] != i
// into the alternate set (exactly as k_gather would), phase 1 copies them
// back; slots with idx[i] == i already hold their gathered row.
__global__ void __launch_bounds__(256) k_gather_moved(const DevParams* __restrict__ P, Arrays cur, Arrays alt,
                                                      const int* __restrict__ idx, int phase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P->np) return;
  const int s = idx[i];
  if (s == i) return;
  const bool st = i < P->npb + P->npm;
  if (phase == 0) {
    alt.codes[i] = cur.codes[s];
    alt.cellpos[i] = cur.cellpos[s];
    alt.orig_id[i] = cur.orig_id[s];
    alt.case_id[i] = cur.case_id[s];
    alt.pres[i] = cur.pres[s];
    alt.pos[i] = cur.pos[s];
    alt.pos_ori[i] = cur.pos_ori[s];
    alt.vel[i] = cur.vel[s];
    alt.idm[i] = cur.idm[s];
    alt.cvar[i] = cur.cvar[s];
    alt.cf[i] = cur.cf[s];
    alt.cn[i] = cur.cn[s];
    alt.cw[i] = cur.cw[s];
    if (i < P->npb) alt.normal[i] = cur.normal[s];
    if (st) {
      alt.str_a[i] = cur.str_a[s];
      alt.str_b[i] = cur.str_b[s];
      alt.stn_a[i] = cur.stn_a[s];
      alt.stn_b[i] = cur.stn_b[s];
    }
  } else {
    cur.codes[i] = alt.codes[i];
    cur.cellpos[i] = alt.cellpos[i];
    cur.orig_id[i] = alt.orig_id[i];
    cur.case_id[i] = alt.case_id[i];
    cur.pres[i] = alt.pres[i];
    cur.pos[i] = alt.pos[i];
    cur.pos_ori[i] = alt.pos_ori[i];
    cur.vel[i] = alt.vel[i];
    cur.idm[i] = alt.idm[i];
    cur.cvar[i] = alt.cvar[i];
    cur.cf[i] = alt.cf[i];
    cur.cn[i] = alt.cn[i];
    cur.cw[i] = alt.cw[i];
    if (i < P->npb) cur.normal[i] = alt.normal[i];
    if (st) {
      cur.str_a[i] = alt.str_a[i];
      cur.str_b[i] = alt.str_b[i];
      cur.stn_a[i] = alt.stn_a[i];
      cur.stn_b[i] = alt.stn_b[i];
    }
  }
}

// The *_ini arrays of SortArraysAll are permuted in place through scratch.
__global__ void __launch_bounds__(256) k_gather_ini(const DevParams* __restrict__ P, const int* __restrict__ idx,
                                                    const float4* __restrict__ pi, const float4* __restrict__ vi,
                                                    const float4* __restrict__ sia, const float2* __restrict__ sib,
                                                    const float4* __restrict__ nia, const float2* __restrict__ nib,
                                                    float4* __restrict__ po, float4* __restrict__ vo,
                                                    float4* __restrict__ soa, float2* __restrict__ sob,
                                                    float4* __restrict__ noa, float2* __restrict__ nob) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P->np) return;
  const int s = idx[i];
  po[i] = pi[s];
  vo[i] = vi[s];
  const bool st = i < P->npb + P->npm;
  const int ss = st ? s : i;
  soa[i] = sia[ss]; sob[i] = sib[ss];
  noa[i] = nia[ss]; nob[i] = nib[ss];
}

static inline int blocks(long long n, int b) { return int((n + b - 1) / b); }

void launch_fill_int(hipStream_t s, int* a, int n, int v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(std::min(blocks(n, 256), 4096)), dim3(256), 0, s, a, n, v);
}

int radix_tiles(int n) { return blocks(n, RS_TILE); }

// Hash + stable sort by (key, slot) + ranges + deletion bookkeeping.  With
// prev (the sorted keys of the previous search, slot order unchanged since)
// the sort is the merge search; without it a full radix sort.  The sorted
// keys go to out_key (not prev), the permutation to B.vals.
void launch_neighbor_search(hipStream_t s, DevParams* P, int np, int nbits, int ncell, Arrays A, Buffers& B,
                            const int* prev, int* out_key, bool clear_flag) {
  if (clear_flag) hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, P);
  // the merge path scans np + 1 flags in tiles of HASH_THREADS: the hash
  // writes the tile sums, one workgroup scans them, k_scan_movers finishes
  const int n = np + 1, nt = (n + HASH_THREADS - 1) / HASH_THREADS;
  int* sums = B.hist;
  hipLaunchKernelGGL(k_hash, dim3(blocks(std::max(n, ncell), HASH_THREADS)), dim3(HASH_THREADS), 0, s, P, A, B.keys,
                     B.vals, prev, prev ? B.mflag : nullptr, prev ? sums : nullptr, ncell, B.cell_beg, B.cell_end);
  if (prev) {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_THREADS), 0, s, sums, nt);
    hipLaunchKernelGGL(k_scan_movers, dim3(nt), dim3(HASH_THREADS), 0, s, B.mflag, n, (const int*)sums,
                       (const int*)B.keys, np, B.mkey, B.mslot);
    hipLaunchKernelGGL(k_mover_sort, dim3(1), dim3(MS_THREADS), 0, s, (const int*)(B.mflag + np), B.mkey, B.mslot,
                       B.mkey_alt, B.mslot_alt, nbits);
    const bool odd = ((nbits + 7) / 8) & 1;
    hipLaunchKernelGGL(k_merge_place, dim3(blocks(np, 256)), dim3(256), 0, s, (const DevParams*)P,
                       (const int*)B.keys, prev, (const int*)B.mflag, (const int*)(odd ? B.mkey_alt : B.mkey),
                       (const int*)(odd ? B.mslot_alt : B.mslot), out_key, B.vals);
    B.part_idx = B.vals;
  } else {
    int* kin = B.keys; int* vin = B.vals;
    int* kout = B.keys_alt; int* vout = B.vals_alt;
    const int nb = radix_tiles(np);
    for (int shift = 0; shift < nbits; shift += 8) {
      hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(RS_THREADS), 0, s, (const int*)kin, np, shift, B.hist, nb);
      {
        const int nh = 256 * nb, nt = scan_tiles(nh);
        int* sums = B.hist + nh;  // the histogram allocation carries the tile sums
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(nt), dim3(SC_THREADS), 0, s, (const int*)B.hist, nh, sums);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_THREADS), 0, s, sums, nt);
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SC_THREADS), 0, s, B.hist, nh, (const int*)sums);
      }
      hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(RS_THREADS), 0, s, (const int*)kin, (const int*)vin, kout,
                         vout, np, shift, (const int*)B.hist, nb);
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    hipMemcpyAsync(out_key, kin, size_t(np) * sizeof(int), hipMemcpyDeviceToDevice, s);
    B.part_idx = vin;
  }
  B.part_cel = out_key;
  hipLaunchKernelGGL(k_beg_end, dim3(blocks(np, 256)), dim3(256), 0, s, P, (const int*)out_key, B.cell_beg, B.cell_end);
  const int ncm = (ncell - 1) / 2;
  if (ncm > 0)
    hipLaunchKernelGGL(k_wall_near, dim3(blocks(ncm, 256)), dim3(256), 0, s, P, (const int*)B.cell_beg,
                       (const int*)B.cell_end, B.wall_near);
  else
    hipLaunchKernelGGL(k_delete_count, dim3(1), dim3(1), 0, s, P, (const int*)B.cell_beg, (const int*)B.cell_end);
}

void launch_gather(hipStream_t s, const DevParams* P, int np, Arrays src, Arrays dst, const int* idx, const Buffers& B) {
  if (np > 0) hipLaunchKernelGGL(k_gather, dim3(blocks(np, 256)), dim3(256), 0, s, P, src, dst, idx, B, 0);
}

void launch_gather_moved(hipStream_t s, const DevParams* P, int np, Arrays cur, Arrays alt, const int* idx) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_gather_moved, dim3(blocks(np, 256)), dim3(256), 0, s, P, cur, alt, idx, 0);
  hipLaunchKernelGGL(k_gather_moved, dim3(blocks(np, 256)), dim3(256), 0, s, P, cur, alt, idx, 1);
}

void launch_gather_ini(hipStream_t s, const DevParams* P, int np, const int* idx, Buffers& B, float4* t4a,
                       float4* t4b, float4* t4c, float2* t2a, float4* t4d, float2* t2b) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_gather_ini, dim3(blocks(np, 256)), dim3(256), 0, s, P, idx, (const float4*)B.pos_ini,
                     (const float4*)B.vel_ini, (const float4*)B.str_ini_a, (const float2*)B.str_ini_b,
                     (const float4*)B.stn_ini_a, (const float2*)B.stn_ini_b, t4a, t4b, t4c, t2a, t4d, t2b);
}

}  // namespace lq
