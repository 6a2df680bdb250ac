// lq_surrogate.hip -- MLP surrogate of the HypoCS update (constitutive_model.h
// :209-299, called at predictor_corrector.cu:75 and :172).
//
// Network: 14 -> 128 -> 128 -> 128 -> 7, GELU (tanh form) after the three
// hidden layers, inputs and outputs normalised by per-feature mean/std.
// Inputs x = (s = sigma - coh, D dt, e, dt); outputs y = ((rate - Jaumann) dt,
// M_crt).  The step runs, per call site: k_sur_inputs (normalised x per soil
// slot) -> k_mlp_bf16 once per ensemble member (BF16 MFMA, FP32 accumulate)
// -> k_sur_finish (input-range and ensemble guard, 1 % audit, exact tail of
// the update; guarded slots take the exact update itself).
//
// Training (standalone, FP32): a tiled SGEMM, bias/GELU and their
// derivatives, column sums and Adam, driven from the host over shuffled
// mini-batches of logged (x, y) tuples.
//
// Parameter vector (FP32): W1[14][128] b1[128] W2[128][128] b2[128]
// W3[128][128] b3[128] W4[128][7] b4[7]; W[k][n] maps input k to output n.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "lq_device.hpp"
#include "lq_internal.hpp"
#include "lq_launch.hpp"

namespace lq {

constexpr int H = 128;
constexpr int K1P = 40;   // padded row length (bf16) of the transposed W1
constexpr int KP = 136;   // padded row length (bf16) of W2t, W3t, W4t and activations
constexpr int W1T = H * K1P, W2T = H * KP, W4T = 16 * KP;
constexpr int WPACK = W1T + 2 * W2T + W4T;  // bf16 per member
constexpr int BPACK = 3 * H + 16;           // biases per member (FP32, b4 padded)
constexpr int MLP_WAVES = 16;
constexpr int SUR_SLOTS = 64;  // statistics counter lines (8 u64 each; audit max every 16 u32)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float gelu(float x) {  // 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
  const float u = 0.7978845608f * fmaf(0.044715f * x, x * x, x);
  const float t = 1.0f - 2.0f / (__expf(2.0f * u) + 1.0f);
  return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float x2 = x * x;
  const float u = 0.7978845608f * fmaf(0.044715f * x, x2, x);
  const float t = 1.0f - 2.0f / (__expf(2.0f * u) + 1.0f);
  const float du = 0.7978845608f * fmaf(0.134145f, x2, 1.0f);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}
__device__ __forceinline__ unsigned short bf16_bits(float f) {  // v_cvt_pk_bf16_f32, round to nearest even
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f));
}

// ------------------------------------------------------ forward (MFMA) ----
// One 1024-thread block per CU keeps a member's weights in LDS (84 KB); each
// of its 16 waves walks pairs of 16-sample tiles, so every weight fragment
// read from LDS feeds two MFMAs.  Each layer is computed transposed,
// out^T = W x in^T, with v_mfma_f32_16x16x32_bf16: A = 16 output features x
// 32 inputs (W rows, feature = lane % 16), B = 32 inputs x 16 samples
// (sample = lane % 16), 8 inputs per lane (MFMA k = 32 kb + 8 (lane / 16)
// + i).  D holds features 16 nb + 4 (lane / 16) + j of sample lane % 16, so
// over the 8 output blocks nb a lane holds 32 features of its sample: as
// bf16, exactly the B fragments of the next layer once the next layer's
// weight columns are stored in that order (k = 32 kb + 8 g + i reads
// feature 16 (2 kb + i / 4) + 4 g + i % 4; sur_pack).  Activations thus
// stay in registers from layer to layer.  The accumulator starts at the
// bias.
constexpr int MLP_TILES = 2;  // 16-sample tiles per wave and iteration
// The same function as gelu(): 0.5 x (1 + tanh(u)) = x / (1 + exp(-2u)),
// with -2u log2(e) = x (GA + GB x^2), evaluated with the hardware exp2 and
// reciprocal (about 1 ulp each; the result is rounded to bf16 next).
constexpr float GA = -2.0f * 0.7978845608f * 1.4426950409f;
constexpr float GB = GA * 0.044715f;
__device__ __forceinline__ float gelu_folded(float x) {
  const float e = __builtin_amdgcn_exp2f(x * fmaf(x * x, GB, GA));
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ uint2 gelu4_bf16(f32x4 v) {
  return make_uint2(unsigned(bf16_bits(gelu_folded(v[0]))) | (unsigned(bf16_bits(gelu_folded(v[1]))) << 16),
                    unsigned(bf16_bits(gelu_folded(v[2]))) | (unsigned(bf16_bits(gelu_folded(v[3]))) << 16));
}
__device__ __forceinline__ bf16x8 frag(uint2 lo, uint2 hi) {
  return __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
}
__global__ void __launch_bounds__(64 * MLP_WAVES) k_mlp_bf16(const float* __restrict__ X, int n,
                                                             const unsigned short* __restrict__ Wg,
                                                             const float* __restrict__ bg, float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) unsigned short sW[WPACK];
  __shared__ __attribute__((aligned(16))) float sB[BPACK];
  for (int k = threadIdx.x; k < WPACK / 8; k += blockDim.x)
    reinterpret_cast<uint4*>(sW)[k] = reinterpret_cast<const uint4*>(Wg)[k];
  for (int k = threadIdx.x; k < BPACK; k += blockDim.x) sB[k] = bg[k];
  __syncthreads();
  const unsigned short* W1 = sW;
  const unsigned short* W2 = sW + W1T;
  const unsigned short* W3 = W2 + W2T;
  const unsigned short* W4 = W3 + W2T;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, kg = lane >> 4;
  const int ngroups = (n + 16 * MLP_TILES - 1) / (16 * MLP_TILES);
  for (int grp = blockIdx.x * MLP_WAVES + wv; grp < ngroups; grp += gridDim.x * MLP_WAVES) {
    const int row0 = grp * 16 * MLP_TILES;
    // layer 1: 14 (padded to 32) -> 128
    bf16x8 x[MLP_TILES];
    for (int t = 0; t < MLP_TILES; t++) {
      const int row = row0 + 16 * t + r16;
      for (int i = 0; i < 8; i++) {
        const int k = 8 * kg + i;
        x[t][i] = static_cast<__bf16>((row < n && k < SUR_IN) ? X[size_t(row) * SUR_XS + k] : 0.f);
      }
    }
    uint2 h[MLP_TILES][8];  // bf16 activations: features 16 nb + 4 kg + 0..3 of the lane's sample
#pragma unroll
    for (int nb = 0; nb < 8; nb++) {
      const bf16x8 w = *reinterpret_cast<const bf16x8*>(W1 + (nb * 16 + r16) * K1P + 8 * kg);
      const f32x4 bias = *reinterpret_cast<const f32x4*>(sB + nb * 16 + 4 * kg);
#pragma unroll
      for (int t = 0; t < MLP_TILES; t++)
        h[t][nb] = gelu4_bf16(__builtin_amdgcn_mfma_f32_16x16x32_bf16(w, x[t], bias, 0, 0, 0));
    }
    // layers 2 and 3: 128 -> 128
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      const unsigned short* Wl = layer == 0 ? W2 : W3;
      const float* bl = sB + H * (1 + layer);
      uint2 o[MLP_TILES][8];
#pragma unroll
      for (int nb = 0; nb < 8; nb++) {
        const f32x4 bias = *reinterpret_cast<const f32x4*>(bl + nb * 16 + 4 * kg);
        f32x4 acc[MLP_TILES];
#pragma unroll
        for (int t = 0; t < MLP_TILES; t++) acc[t] = bias;
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
          const bf16x8 w = *reinterpret_cast<const bf16x8*>(Wl + (nb * 16 + r16) * KP + kb * 32 + 8 * kg);
#pragma unroll
          for (int t = 0; t < MLP_TILES; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, frag(h[t][2 * kb], h[t][2 * kb + 1]), acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < MLP_TILES; t++) o[t][nb] = gelu4_bf16(acc[t]);
        __builtin_amdgcn_sched_barrier(0);  // keep each block's weight reads next to its MFMAs
      }
#pragma unroll
      for (int t = 0; t < MLP_TILES; t++)
#pragma unroll
        for (int nb = 0; nb < 8; nb++) h[t][nb] = o[t][nb];
    }
    // layer 4: 128 -> 7 (padded to 16), no activation
    const f32x4 b4 = *reinterpret_cast<const f32x4*>(sB + 3 * H + 4 * kg);
#pragma unroll
    for (int t = 0; t < MLP_TILES; t++) {
      f32x4 acc = b4;
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        const bf16x8 w = *reinterpret_cast<const bf16x8*>(W4 + r16 * KP + kb * 32 + 8 * kg);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, frag(h[t][2 * kb], h[t][2 * kb + 1]), acc, 0, 0, 0);
      }
      const int row = row0 + 16 * t + r16;
      for (int i = 0; i < 4; i++)
        if (row < n && 4 * kg + i < SUR_OUT) Y[size_t(row) * SUR_YS + 4 * kg + i] = acc[i];
    }
  }
}

// ------------------------------------------------------- step kernels ----
__device__ __forceinline__ bool sur_slot(const DevParams* __restrict__ P, const Arrays& A, int i) {
  const int4 cd = A.codes[i];
  return cd.y == LQ_NONFIX && cd.z == LQ_HYPO_CS;
}

// Normalised inputs of every soil slot (zeros where the surrogate is not used).
__device__ __forceinline__ float stage_dt(const DevParams* __restrict__ P, int stage) {
  return stage ? P->dt : 0.5f * P->dt;  // predictor_corrector.cu:75 (dt/2) and :172 (dt)
}

__global__ void __launch_bounds__(256) k_sur_inputs(const DevParams* __restrict__ P, Arrays A, Buffers B, int stage) {
  __shared__ float s_mean[SUR_IN], s_inv[SUR_IN];
  if (threadIdx.x < SUR_IN) {
    s_mean[threadIdx.x] = P->sur_norm[threadIdx.x];
    s_inv[threadIdx.x] = 1.0f / P->sur_norm[SUR_IN + threadIdx.x];
  }
  __syncthreads();
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const float dt = stage_dt(P, stage);
  const int i = P->npb + idx;
  float x[SUR_IN];
  bool use = false;
  if (sur_slot(P, A, i)) {
    const Sym sig = ldsym(B.str_ini_a, B.str_ini_b, i);
    const Sym D = ldsym(B.acestn_a, B.acestn_b, i);
    use = hypo_cs_parts(P, A.codes[i].w, dt, A.vel[i].w, D, sig, x, nullptr, false);
  }
  float* o = B.sur_x + size_t(idx) * SUR_XS;
  for (int k = 0; k < SUR_IN; k++) o[k] = use ? (x[k] - s_mean[k]) * s_inv[k] : 0.f;
  o[14] = 0.f;
  o[15] = 0.f;
}

__device__ __forceinline__ unsigned audit_hash(int id, long long step, int stage) {
  unsigned long long z = (unsigned long long)(unsigned)id * 0x9E3779B97F4A7C15ull ^
                         (unsigned long long)(step * 2 + stage) * 0xBF58476D1CE4E5B9ull;
  z ^= z >> 31;
  z *= 0x94D049BB133111EBull;
  z ^= z >> 29;
  return unsigned(z >> 32);
}

// Guard, audit and the update of every soil slot.  stage 0 = predictor
// (dt/2), 1 = corrector (writes M_crt into idmassidx.w).
__global__ void __launch_bounds__(256) k_sur_finish(DevParams* __restrict__ P, Arrays A, Buffers B, int stage,
                                                    int ystride) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const float dt = stage_dt(P, stage);
  const bool live = idx < P->npm;
  const int i = P->npb + (live ? idx : 0);
  bool act = live && sur_slot(P, A, i);
  int cat = 0;  // 1 surrogate, 2 fallback range, 3 fallback ensemble, 4 I1 > -4 branch
  bool audited = false;
  float aerr = 0.f;
  if (act) {
    const int model = A.codes[i].w;
    Sym sig = ldsym(B.str_ini_a, B.str_ini_b, i);
    const Sym D = ldsym(B.acestn_a, B.acestn_b, i);
    const float3 w = ld3(B.spin[i]);
    const float rho = A.vel[i].w;
    float x[SUR_IN];
    float Mcrt = 0.f;
    if (!hypo_cs_parts(P, model, dt, rho, D, sig, x, nullptr, false)) {
      hypo_cs(P, model, dt, rho, w, D, sig, Mcrt);  // fixed-tensor branch
      cat = 4;
    } else {
      const float* nm = P->sur_norm;
      bool in_range = true;
      for (int k = 0; k < SUR_IN; k++) in_range = in_range && x[k] >= P->sur_lo[k] && x[k] <= P->sur_hi[k];
      const float* ya = B.sur_y + size_t(idx) * SUR_YS;
      float yn[SUR_OUT];
      float spread = 0.f;
      for (int k = 0; k < SUR_OUT; k++) yn[k] = ya[k];
      if (P->sur_members > 1) {
        const float* yb = B.sur_y + (size_t(ystride) + idx) * SUR_YS;
        for (int k = 0; k < SUR_OUT; k++) {
          spread = fmaxf(spread, fabsf(ya[k] - yb[k]));
          yn[k] = 0.5f * (ya[k] + yb[k]);
        }
      }
      if (!in_range) cat = 2;
      else if (!(spread <= P->sur_ens_tol)) cat = 3;
      else cat = 1;
      float y[SUR_OUT];
      for (int k = 0; k < SUR_OUT; k++) y[k] = fmaf(yn[k], nm[2 * SUR_IN + SUR_OUT + k], nm[2 * SUR_IN + k]);
      if (cat == 1 && float(audit_hash(A.orig_id[i], P->step, stage)) * 2.3283064e-10f < P->sur_audit_frac) {
        float xe[SUR_IN], ye[SUR_OUT];
        hypo_cs_parts(P, model, dt, rho, D, sig, xe, ye, true);
        for (int k = 0; k < SUR_OUT; k++)
          aerr = fmaxf(aerr, fabsf((ye[k] - nm[2 * SUR_IN + k]) / nm[2 * SUR_IN + SUR_OUT + k] - yn[k]));
        audited = true;
      }
      if (cat == 1) {
        Mcrt = y[6];
        hypo_cs_apply(P, model, dt, w, y, x[12], sig);
      } else {
        hypo_cs(P, model, dt, rho, w, D, sig, Mcrt);  // guard: the exact update
      }
    }
    stsym(A.str_a, A.str_b, i, sig);
    if (stage == 1) {
      float4 id = A.idm[i];
      id.w = Mcrt;
      A.idm[i] = id;
    }
  }
  // statistics: counts summed over the block, then one atomic per category
  // into one of SUR_SLOTS counter lines (so blocks do not all hit one
  // address); the host adds the slots up
  __shared__ unsigned s_cnt[5];
  __shared__ unsigned s_amax;
  if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) s_amax = 0u;
  __syncthreads();
  const unsigned long long m1 = __ballot(cat == 1), m2 = __ballot(cat == 2), m3 = __ballot(cat == 3),
                           m4 = __ballot(cat == 4), ma = __ballot(audited);
  float em = aerr;
  for (int o = 32; o > 0; o >>= 1) em = fmaxf(em, __shfl_xor(em, o));
  if ((threadIdx.x & 63) == 0) {
    if (m1) atomicAdd(&s_cnt[0], unsigned(__popcll(m1)));
    if (m2) atomicAdd(&s_cnt[1], unsigned(__popcll(m2)));
    if (m3) atomicAdd(&s_cnt[2], unsigned(__popcll(m3)));
    if (ma) atomicAdd(&s_cnt[3], unsigned(__popcll(ma)));
    if (m4) atomicAdd(&s_cnt[4], unsigned(__popcll(m4)));
    if (ma) atomicMax(&s_amax, __float_as_uint(em));
  }
  __syncthreads();
  unsigned long long* slot = P->sur_stats + 8 * (blockIdx.x & (SUR_SLOTS - 1));
  if (threadIdx.x < 5 && s_cnt[threadIdx.x]) atomicAdd(&slot[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
  if (threadIdx.x == 5 && s_amax) atomicMax(P->sur_audit_max + (blockIdx.x & (SUR_SLOTS - 1)) * 16, s_amax);
}

// Logs raw (x, exact y) of every log_every-th soil slot (rotating with the
// step and stage) while the exact update runs; stops at log_cap tuples.
__global__ void __launch_bounds__(256) k_sur_log(DevParams* __restrict__ P, Arrays A, Buffers B, int stage) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P->npm) return;
  const float dt = stage_dt(P, stage);
  const int every = P->log_every;
  if (every <= 0 || (idx + int((P->step * 2 + stage) % every)) % every != 0) return;
  const int i = P->npb + idx;
  if (!sur_slot(P, A, i)) return;
  const Sym sig = ldsym(B.str_ini_a, B.str_ini_b, i);
  const Sym D = ldsym(B.acestn_a, B.acestn_b, i);
  float x[SUR_IN], y[SUR_OUT];
  if (!hypo_cs_parts(P, A.codes[i].w, dt, A.vel[i].w, D, sig, x, y, true)) return;
  // one atomic per wave reserves the rows of all its logging lanes
  const unsigned long long m = __ballot(1);
  const int leader = __ffsll((long long)m) - 1;
  const unsigned rank = unsigned(__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull)));
  unsigned base = 0u;
  if (int(threadIdx.x & 63) == leader) base = atomicAdd(P->log_count, unsigned(__popcll(m)));
  base = __shfl(base, leader);
  const unsigned slot = base + rank;
  if (slot >= unsigned(P->log_cap)) return;
  for (int k = 0; k < SUR_IN; k++) P->log_x[size_t(slot) * SUR_XS + k] = x[k];
  for (int k = 0; k < SUR_OUT; k++) P->log_y[size_t(slot) * SUR_YS + k] = y[k];
}

static int blocks_of(long long n, int t) { return int((n + t - 1) / t); }

void launch_surrogate(hipStream_t s, DevParams* P, int npm_max, int members, int n_cu, Arrays A, const Buffers& B,
                      int stage) {
  if (npm_max <= 0) return;
  hipLaunchKernelGGL(k_sur_inputs, dim3(blocks_of(npm_max, 256)), dim3(256), 0, s, P, A, B, stage);
  const int ngroups = (npm_max + 16 * MLP_TILES - 1) / (16 * MLP_TILES);
  const int grid = std::max(1, std::min(n_cu, (ngroups + MLP_WAVES - 1) / MLP_WAVES));
  for (int m = 0; m < members; m++)
    hipLaunchKernelGGL(k_mlp_bf16, dim3(grid), dim3(64 * MLP_WAVES), 0, s, (const float*)B.sur_x, npm_max,
                       (const unsigned short*)(B.sur_w + size_t(m) * WPACK), (const float*)(B.sur_b + m * BPACK),
                       B.sur_y + size_t(m) * npm_max * SUR_YS);
  hipLaunchKernelGGL(k_sur_finish, dim3(blocks_of(npm_max, 256)), dim3(256), 0, s, P, A, B, stage, npm_max);
}
void launch_sur_log(hipStream_t s, DevParams* P, int npm_max, Arrays A, const Buffers& B, int stage) {
  if (npm_max > 0) hipLaunchKernelGGL(k_sur_log, dim3(blocks_of(npm_max, 256)), dim3(256), 0, s, P, A, B, stage);
}
int sur_wpack() { return WPACK; }
int sur_slots() { return SUR_SLOTS; }
int sur_bpack() { return BPACK; }

// Host-side packing of one member's FP32 parameters into the LDS layout.
void sur_pack(const float* p, unsigned short* w, float* b) {
  auto bf = [](float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  };
  const float* W1 = p;
  const float* b1 = W1 + SUR_IN * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * H;
  const float* W4 = b3 + H;
  const float* b4 = W4 + H * SUR_OUT;
  std::fill(w, w + WPACK, (unsigned short)0);
  for (int n = 0; n < H; n++)
    for (int k = 0; k < SUR_IN; k++) w[n * K1P + k] = bf(W1[k * H + n]);
  // layers 2-4 read their inputs in the order k_mlp_bf16's registers hold
  // them: packed column k = 32 kb + 8 g + i is input feature
  // 16 (2 kb + i / 4) + 4 g + i % 4
  auto feat = [](int k) { return 16 * (2 * (k >> 5) + ((k & 7) >> 2)) + 4 * ((k >> 3) & 3) + (k & 3); };
  for (int n = 0; n < H; n++)
    for (int k = 0; k < H; k++) {
      w[W1T + n * KP + k] = bf(W2[feat(k) * H + n]);
      w[W1T + W2T + n * KP + k] = bf(W3[feat(k) * H + n]);
    }
  for (int n = 0; n < SUR_OUT; n++)
    for (int k = 0; k < H; k++) w[W1T + 2 * W2T + n * KP + k] = bf(W4[feat(k) * SUR_OUT + n]);
  std::fill(b, b + BPACK, 0.f);
  for (int n = 0; n < H; n++) { b[n] = b1[n]; b[H + n] = b2[n]; b[2 * H + n] = b3[n]; }
  for (int n = 0; n < SUR_OUT; n++) b[3 * H + n] = b4[n];
}

// ----------------------------------------------------- training (FP32) ----
// C[M][N] = A'[M][K] B'[K][N] (+ C if accumulate); A' = A or A^T, B' = B or
// B^T, all row-major with leading dimensions lda/ldb/ldc.  16 x 16 output
// tile per 256-thread block, K staged through LDS in steps of 16.
__global__ void __launch_bounds__(256) k_sgemm(int M, int N, int K, const float* __restrict__ Am, int lda, int ta,
                                               const float* __restrict__ Bm, int ldb, int tb, float* __restrict__ Cm,
                                               int ldc, int accumulate) {
  __shared__ float sa[16][17], sb[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  float acc = 0.f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int ka = k0 + tx, kb = k0 + ty;
    const int ar = blockIdx.y * 16 + ty;
    sa[ty][tx] = (ar < M && ka < K) ? (ta ? Am[size_t(ka) * lda + ar] : Am[size_t(ar) * lda + ka]) : 0.f;
    const int bc = blockIdx.x * 16 + tx;
    sb[ty][tx] = (kb < K && bc < N) ? (tb ? Bm[size_t(bc) * ldb + kb] : Bm[size_t(kb) * ldb + bc]) : 0.f;
    __syncthreads();
    for (int k = 0; k < 16; k++) acc = fmaf(sa[ty][k], sb[k][tx], acc);
    __syncthreads();
  }
  if (row < M && col < N) {
    float* c = Cm + size_t(row) * ldc + col;
    *c = accumulate ? *c + acc : acc;
  }
}

// Z[M][N] += b; A = gelu(Z) when act (else A = Z).
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ b, float* __restrict__ Ao, int M, int N,
                           int act) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * N) return;
  const float z = Z[t] + b[t % N];
  Z[t] = z;
  Ao[t] = act ? gelu(z) : z;
}

// dY = 2 (Y - T) / (M N); loss partial sums per block.
__global__ void k_out_grad(const float* __restrict__ Yp, const float* __restrict__ T, float* __restrict__ dY, int M,
                           int N, double* __restrict__ loss) {
  __shared__ double sh[256];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (t < (long long)M * N) {
    const float d = Yp[t] - T[t];
    dY[t] = 2.0f * d / float(M * N);
    l = double(d) * double(d);
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, sh[0] / double((long long)M * N));
}

__global__ void k_gelu_back(float* __restrict__ dA, const float* __restrict__ Z, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dA[t] *= gelu_grad(Z[t]);
}

// db[n] = sum over rows of dZ[M][N] (one block per column).
__global__ void k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  __shared__ float sh[256];
  float a = 0.f;
  for (int r = threadIdx.x; r < M; r += blockDim.x) a += dZ[size_t(r) * N + blockIdx.x];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = 0.9f * m[t] + 0.1f * gi;
  const float vi = 0.999f * v[t] + 0.001f * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ void k_gather_rows(const float* __restrict__ src, int ld, const int* __restrict__ idx, int rows, int cols,
                              float* __restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const int r = t / cols, c = t % cols;
  dst[t] = src[size_t(idx[r]) * ld + c];
}

void launch_sgemm(hipStream_t s, int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                  int tb, float* C, int ldc, int accumulate) {
  hipLaunchKernelGGL(k_sgemm, dim3((N + 15) / 16, (M + 15) / 16), dim3(256), 0, s, M, N, K, A, lda, ta, B, ldb, tb,
                     C, ldc, accumulate);
}
void launch_bias_act(hipStream_t s, float* Z, const float* b, float* A, int M, int N, int act) {
  hipLaunchKernelGGL(k_bias_act, dim3(blocks_of((long long)M * N, 256)), dim3(256), 0, s, Z, b, A, M, N, act);
}
void launch_out_grad(hipStream_t s, const float* Y, const float* T, float* dY, int M, int N, double* loss) {
  hipLaunchKernelGGL(k_out_grad, dim3(blocks_of((long long)M * N, 256)), dim3(256), 0, s, Y, T, dY, M, N, loss);
}
void launch_gelu_back(hipStream_t s, float* dA, const float* Z, long long n) {
  hipLaunchKernelGGL(k_gelu_back, dim3(blocks_of(n, 256)), dim3(256), 0, s, dA, Z, n);
}
void launch_colsum(hipStream_t s, const float* dZ, int M, int N, float* db) {
  hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, s, dZ, M, N, db);
}
void launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, int n, float lr, float c1, float c2) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_of(n, 256)), dim3(256), 0, s, p, g, m, v, n, lr, c1, c2);
}
void launch_gather_rows(hipStream_t s, const float* src, int ld, const int* idx, int rows, int cols, float* dst) {
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks_of((long long)rows * cols, 256)), dim3(256), 0, s, src, ld, idx, rows,
                     cols, dst);
}
void launch_mlp_bf16(hipStream_t s, const float* X, int n, const unsigned short* W, const float* b, float* Y,
                     int n_cu) {
  const int ngroups = (n + 16 * MLP_TILES - 1) / (16 * MLP_TILES);
  const int grid = std::max(1, std::min(n_cu, (ngroups + MLP_WAVES - 1) / MLP_WAVES));
  hipLaunchKernelGGL(k_mlp_bf16, dim3(grid), dim3(64 * MLP_WAVES), 0, s, X, n, W, b, Y);
}

}  // namespace lq

// ------------------------------------------------ standalone C entry points --
#include <cmath>
#include <random>
#include <string>
#include <vector>

namespace {

std::string g_mlp_err;

#define MLPCHECK(call)                                                   \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      g_mlp_err = std::string(#call) + ": " + hipGetErrorString(e_);     \
      return LQ_ERR_HIP;                                                 \
    }                                                                    \
  } while (0)

constexpr int NIN = lq::SUR_IN, NOUT = lq::SUR_OUT, NH = lq::H;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH * NOUT;
constexpr int NPARAM = OFF_B4 + NOUT;
static_assert(NPARAM == LQ_MLP_NPARAM, "parameter count");

// Device scratch for one mini-batch of up to cap rows.
struct MlpWork {
  hipStream_t s = nullptr;
  int cap = 0;
  float *P = nullptr, *G = nullptr, *X = nullptr, *T = nullptr, *Y = nullptr, *dY = nullptr;
  float *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Z3 = nullptr, *A3 = nullptr;
  float *D0 = nullptr, *D1 = nullptr;
  double* loss = nullptr;
  std::vector<void*> mem;
  template <class T> int alloc(T** p, size_t n) {
    void* q = nullptr;
    MLPCHECK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    mem.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
  }
  int init(int rows) {
    cap = rows;
    MLPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t r = size_t(rows);
    int st = 0;
    st |= alloc(&P, NPARAM); st |= alloc(&G, NPARAM);
    st |= alloc(&X, r * NIN); st |= alloc(&T, r * NOUT); st |= alloc(&Y, r * NOUT); st |= alloc(&dY, r * NOUT);
    st |= alloc(&Z1, r * NH); st |= alloc(&A1, r * NH); st |= alloc(&Z2, r * NH); st |= alloc(&A2, r * NH);
    st |= alloc(&Z3, r * NH); st |= alloc(&A3, r * NH); st |= alloc(&D0, r * NH); st |= alloc(&D1, r * NH);
    st |= alloc(&loss, 1);
    return st;
  }
  ~MlpWork() {
    if (s) hipStreamSynchronize(s);
    for (void* p : mem) hipFree(p);
    if (s) hipStreamDestroy(s);
  }
  // forward of rows [0, m) of X into Y (normalised units)
  void forward(int m) {
    using namespace lq;
    launch_sgemm(s, m, NH, NIN, X, NIN, 0, P + OFF_W1, NH, 0, Z1, NH, 0);
    launch_bias_act(s, Z1, P + OFF_B1, A1, m, NH, 1);
    launch_sgemm(s, m, NH, NH, A1, NH, 0, P + OFF_W2, NH, 0, Z2, NH, 0);
    launch_bias_act(s, Z2, P + OFF_B2, A2, m, NH, 1);
    launch_sgemm(s, m, NH, NH, A2, NH, 0, P + OFF_W3, NH, 0, Z3, NH, 0);
    launch_bias_act(s, Z3, P + OFF_B3, A3, m, NH, 1);
    launch_sgemm(s, m, NOUT, NH, A3, NH, 0, P + OFF_W4, NOUT, 0, Y, NOUT, 0);
    launch_bias_act(s, Y, P + OFF_B4, Y, m, NOUT, 0);
  }
  // gradient of mean((Y - T)^2) over rows [0, m) into G; loss accumulated
  void backward(int m) {
    using namespace lq;
    launch_out_grad(s, Y, T, dY, m, NOUT, loss);
    launch_sgemm(s, NH, NOUT, m, A3, NH, 1, dY, NOUT, 0, G + OFF_W4, NOUT, 0);
    launch_colsum(s, dY, m, NOUT, G + OFF_B4);
    launch_sgemm(s, m, NH, NOUT, dY, NOUT, 0, P + OFF_W4, NOUT, 1, D0, NH, 0);
    launch_gelu_back(s, D0, Z3, (long long)m * NH);
    launch_sgemm(s, NH, NH, m, A2, NH, 1, D0, NH, 0, G + OFF_W3, NH, 0);
    launch_colsum(s, D0, m, NH, G + OFF_B3);
    launch_sgemm(s, m, NH, NH, D0, NH, 0, P + OFF_W3, NH, 1, D1, NH, 0);
    launch_gelu_back(s, D1, Z2, (long long)m * NH);
    launch_sgemm(s, NH, NH, m, A1, NH, 1, D1, NH, 0, G + OFF_W2, NH, 0);
    launch_colsum(s, D1, m, NH, G + OFF_B2);
    launch_sgemm(s, m, NH, NH, D1, NH, 0, P + OFF_W2, NH, 1, D0, NH, 0);
    launch_gelu_back(s, D0, Z1, (long long)m * NH);
    launch_sgemm(s, NIN, NH, m, X, NIN, 1, D0, NH, 0, G + OFF_W1, NH, 0);
    launch_colsum(s, D0, m, NH, G + OFF_B1);
  }
};

void normalise(const float* src, size_t n, int w, const float* mean, const float* sd, std::vector<float>& out) {
  out.resize(n * size_t(w));
  for (size_t r = 0; r < n; r++)
    for (int k = 0; k < w; k++) out[r * w + k] = (src[r * w + k] - mean[k]) / sd[k];
}

}  // namespace

extern "C" const char* lq_mlp_last_error(void) { return g_mlp_err.c_str(); }

// Per-feature mean and standard deviation (FP64; a constant feature gets 1).
extern "C" int lq_mlp_normalisation(const float* X, const float* Y, size_t n, float* norm) {
  if (!X || !Y || !norm || n == 0) return LQ_ERR_ARG;
  for (int part = 0; part < 2; part++) {
    const int w = part ? NOUT : NIN;
    const float* D = part ? Y : X;
    float* mean = norm + (part ? 2 * NIN : 0);
    float* sd = norm + (part ? 2 * NIN + NOUT : NIN);
    for (int k = 0; k < w; k++) {
      double s = 0, s2 = 0;
      for (size_t r = 0; r < n; r++) s += D[r * w + k];
      const double mu = s / double(n);
      for (size_t r = 0; r < n; r++) { const double d = D[r * w + k] - mu; s2 += d * d; }
      const double v = std::sqrt(s2 / double(n));
      mean[k] = float(mu);
      sd[k] = v > 1e-30 * std::max(1.0, std::fabs(mu)) ? float(v) : 1.0f;
    }
  }
  return LQ_OK;
}

// Forward pass on the device: FP32 (bf16 == 0) or the BF16-MFMA kernel the
// step uses (bf16 != 0).  X raw [n][14] -> Y raw [n][7].
extern "C" int lq_mlp_forward(const float* params, const float* norm, const float* X, size_t n, float* Y, int bf16) {
  if (!params || !norm || !X || !Y || n == 0 || n > (1u << 30)) return LQ_ERR_ARG;
  std::vector<float> xn;
  normalise(X, n, NIN, norm, norm + NIN, xn);
  std::vector<float> yn(n * NOUT);
  if (bf16) {
    hipStream_t s;
    MLPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<float> x16(n * lq::SUR_XS, 0.f);
    for (size_t r = 0; r < n; r++) std::memcpy(&x16[r * lq::SUR_XS], &xn[r * NIN], NIN * 4);
    std::vector<unsigned short> w(lq::WPACK);
    std::vector<float> b(lq::BPACK);
    lq::sur_pack(params, w.data(), b.data());
    float *dx = nullptr, *dy = nullptr, *db = nullptr;
    unsigned short* dw = nullptr;
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipError_t e = hipMalloc(&dx, x16.size() * 4);
    if (e == hipSuccess) e = hipMalloc(&dy, n * lq::SUR_YS * 4);
    if (e == hipSuccess) e = hipMalloc(&dw, w.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&db, b.size() * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x16.data(), x16.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dw, w.data(), w.size() * 2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(db, b.data(), b.size() * 4, hipMemcpyHostToDevice, s);
    std::vector<float> y8(n * lq::SUR_YS);
    if (e == hipSuccess) {
      lq::launch_mlp_bf16(s, dx, int(n), dw, db, dy, std::max(ncu, 1));
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y8.data(), dy, y8.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(dx); hipFree(dy); hipFree(dw); hipFree(db);
    hipStreamDestroy(s);
    if (e != hipSuccess) { g_mlp_err = std::string("lq_mlp_forward: ") + hipGetErrorString(e); return LQ_ERR_HIP; }
    for (size_t r = 0; r < n; r++) std::memcpy(&yn[r * NOUT], &y8[r * lq::SUR_YS], NOUT * 4);
  } else {
    MlpWork W;
    if (W.init(int(n))) return LQ_ERR_HIP;
    MLPCHECK(hipMemcpyAsync(W.P, params, NPARAM * 4, hipMemcpyHostToDevice, W.s));
    MLPCHECK(hipMemcpyAsync(W.X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, W.s));
    W.forward(int(n));
    MLPCHECK(hipGetLastError());
    MLPCHECK(hipMemcpyAsync(yn.data(), W.Y, yn.size() * 4, hipMemcpyDeviceToHost, W.s));
    MLPCHECK(hipStreamSynchronize(W.s));
  }
  for (size_t r = 0; r < n; r++)
    for (int k = 0; k < NOUT; k++) Y[r * NOUT + k] = yn[r * NOUT + k] * norm[2 * NIN + NOUT + k] + norm[2 * NIN + k];
  return LQ_OK;
}

// Loss mean((y_n(x) - y_n)^2) over the n tuples (normalised units) and its
// gradient with respect to the parameter vector (FP32 kernels).
extern "C" int lq_mlp_loss_grad(const float* params, const float* norm, const float* X, const float* Y, size_t n,
                                float* grad, double* loss) {
  if (!params || !norm || !X || !Y || !grad || !loss || n == 0 || n > (1u << 26)) return LQ_ERR_ARG;
  std::vector<float> xn, yn;
  normalise(X, n, NIN, norm, norm + NIN, xn);
  normalise(Y, n, NOUT, norm + 2 * NIN, norm + 2 * NIN + NOUT, yn);
  MlpWork W;
  if (W.init(int(n))) return LQ_ERR_HIP;
  MLPCHECK(hipMemcpyAsync(W.P, params, NPARAM * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemcpyAsync(W.X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemcpyAsync(W.T, yn.data(), yn.size() * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemsetAsync(W.loss, 0, sizeof(double), W.s));
  W.forward(int(n));
  W.backward(int(n));
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipMemcpyAsync(grad, W.G, NPARAM * 4, hipMemcpyDeviceToHost, W.s));
  MLPCHECK(hipMemcpyAsync(loss, W.loss, sizeof(double), hipMemcpyDeviceToHost, W.s));
  MLPCHECK(hipStreamSynchronize(W.s));
  return LQ_OK;
}

// Adam over shuffled mini-batches of the normalised tuples.  init != 0 draws
// Glorot-uniform weights (zero biases) from seed; otherwise training
// continues from params.  loss_hist[e] = mean mini-batch loss of epoch e.
extern "C" int lq_mlp_train(const float* X, const float* Y, size_t n, const float* norm, int epochs, int batch,
                            float lr, uint64_t seed, int init, float* params, float* loss_hist) {
  if (!X || !Y || !norm || !params || n == 0 || epochs < 1 || batch < 1 || !(lr > 0.f)) return LQ_ERR_ARG;
  if (n > (1u << 28)) return LQ_ERR_ARG;
  std::mt19937_64 rng(seed);
  if (init) {
    auto glorot = [&](int off, int fin, int fout) {
      const double a = std::sqrt(6.0 / double(fin + fout));
      std::uniform_real_distribution<double> u(-a, a);
      for (int k = 0; k < fin * fout; k++) params[off + k] = float(u(rng));
    };
    std::fill(params, params + NPARAM, 0.f);
    glorot(OFF_W1, NIN, NH);
    glorot(OFF_W2, NH, NH);
    glorot(OFF_W3, NH, NH);
    glorot(OFF_W4, NH, NOUT);
  }
  std::vector<float> xn, yn;
  normalise(X, n, NIN, norm, norm + NIN, xn);
  normalise(Y, n, NOUT, norm + 2 * NIN, norm + 2 * NIN + NOUT, yn);
  const int bs = int(std::min<size_t>(size_t(batch), n));
  MlpWork W;
  if (W.init(bs)) return LQ_ERR_HIP;
  float *dXall = nullptr, *dYall = nullptr, *mom = nullptr, *vel = nullptr;
  int* dIdx = nullptr;
  if (W.alloc(&dXall, xn.size()) || W.alloc(&dYall, yn.size()) || W.alloc(&mom, NPARAM) || W.alloc(&vel, NPARAM) ||
      W.alloc(&dIdx, n))
    return LQ_ERR_HIP;
  MLPCHECK(hipMemcpyAsync(dXall, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemcpyAsync(dYall, yn.data(), yn.size() * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemcpyAsync(W.P, params, NPARAM * 4, hipMemcpyHostToDevice, W.s));
  MLPCHECK(hipMemsetAsync(mom, 0, NPARAM * 4, W.s));
  MLPCHECK(hipMemsetAsync(vel, 0, NPARAM * 4, W.s));
  std::vector<int> perm(n);
  for (size_t k = 0; k < n; k++) perm[k] = int(k);
  long long t = 0;
  const size_t nb = n / size_t(bs);
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    MLPCHECK(hipMemcpyAsync(dIdx, perm.data(), n * 4, hipMemcpyHostToDevice, W.s));
    MLPCHECK(hipMemsetAsync(W.loss, 0, sizeof(double), W.s));
    for (size_t b = 0; b < nb; b++) {
      lq::launch_gather_rows(W.s, dXall, NIN, dIdx + b * bs, bs, NIN, W.X);
      lq::launch_gather_rows(W.s, dYall, NOUT, dIdx + b * bs, bs, NOUT, W.T);
      W.forward(bs);
      W.backward(bs);
      t++;
      lq::launch_adam(W.s, W.P, W.G, mom, vel, NPARAM, lr, float(1.0 - std::pow(0.9, double(t))),
                      float(1.0 - std::pow(0.999, double(t))));
    }
    MLPCHECK(hipGetLastError());
    double l = 0;
    MLPCHECK(hipMemcpyAsync(&l, W.loss, sizeof(double), hipMemcpyDeviceToHost, W.s));
    MLPCHECK(hipStreamSynchronize(W.s));
    if (loss_hist) loss_hist[ep] = float(l / double(std::max<size_t>(nb, 1)));
  }
  MLPCHECK(hipMemcpyAsync(params, W.P, NPARAM * 4, hipMemcpyDeviceToHost, W.s));
  MLPCHECK(hipStreamSynchronize(W.s));
  return LQ_OK;
}
