// Large-batch HMC for Bayesian logistic regression with the design matrix
// stored in BF16 and the all-chain contraction on MFMA (gfx950
// v_mfma_f32_16x16x32_bf16).
//
// One leapfrog gradient for C chains (`mcmc/hmc.py` target gradient through
// `mcmc/internal/util.py:246-308`, Bernoulli logits `bernoulli.py:119-135`):
//   S = X B^T            [N x C]   (B = chain positions, split B = Bhi + Blo)
//   R = y - sigmoid(S)   [N x C]   (split R = Rhi + Rlo)
//   G = R^T X            [C x D]   (+ prior gradient in the update kernel)
// glm_grad_mfma fuses both GEMMs per tile of 32 rows: the tile of X (and of
// X^T) is staged in LDS once and used for both products, S and R never
// leave registers/LDS (no [N x C] array in HBM), and each workgroup writes
// one partial G for its row range (split-N), reduced in fixed order by the
// per-chain update kernel.  FP32 accumulation throughout; the BF16 hi/lo
// split of B and R keeps ~16 mantissa bits of both operands that are not
// data.
#pragma once

#include "../../include/hipmc.h"
#include "common.h"

namespace hipmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMD = 256;       // state dimension of this path
constexpr int kMRows = 32;     // rows per LDS tile
constexpr int kMChains = 64;   // chain granularity of the path (16 per wave)
constexpr int kRPitch = kMRows + 8;    // bf16 per R row (one chain) in LDS

__device__ inline f32x4 mfma16(const bf16x8& a, const bf16x8& b,
                               const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Direct global -> LDS loads (global_load_lds): the wave's 64 lanes each
// read 16 (or 4) bytes from their own address and the bytes land at the
// wave-uniform LDS base + 16 (4) * lane.  Issued from inline asm so that the
// compiler's wait-count pass does not make every later LDS read wait for
// them; the kernel waits for its own tile loads (vmcnt) before the barrier
// that publishes a tile.  m0 is a reserved register whose clobber the
// compiler does not track (hipcc warns): each asm statement sets m0 and
// issues the load that reads it, and nothing else in glm_grad_mfma uses m0
// (no LDS-DMA builtins, no dynamically indexed register arrays), which the
// gfx950 assembly of the kernel confirms (its only m0 writes are these).
typedef __attribute__((address_space(3))) char lds_byte;
__device__ inline uint32_t lds_addr(void* p) {
  return (uint32_t)(uintptr_t)(lds_byte*)p;
}
__device__ inline void lds_dma16(const void* g, void* lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
               :: "v"(g), "s"(lds_addr(lds_base)) : "memory", "m0");
}
__device__ inline void lds_dma4(const void* g, void* lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off"
               :: "v"(g), "s"(lds_addr(lds_base)) : "memory", "m0");
}

// Chunk swizzle of X^T row d: 16 consecutive d of one chunk column land in
// 16 distinct 16-byte bank groups, and any 8 consecutive aligned d in 8.
__device__ inline int xt_swz(int d) { return ((d >> 1) ^ (d >> 2)) & 3; }

// X tile: 16-byte chunk (row, c) at chunk row * 32 + (c ^ (row & 15));
// X^T tile: chunk (d, c) at d * 4 + (c ^ xt_swz(d)).  The WAVES waves of the
// workgroup share the tile's 1024 + 1024 chunk loads and its 32 labels.
template <int WAVES>
__device__ inline void stage_tile(const __bf16* X, const __bf16* Xt,
                                  const float* y, int n_obs, int n0,
                                  __bf16* xs, __bf16* xts, float* ys, int w,
                                  int l) {
  constexpr int U = 16 / WAVES;      // 1 KiB chunk rows per wave and matrix
  constexpr int YL = kMRows / WAVES; // labels per wave
  w = __builtin_amdgcn_readfirstlane(w);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = (w * U + u) * 64 + l;
    const int row = q >> 5, c = (q & 31) ^ (row & 15);
    lds_dma16(X + (size_t)(n0 + row) * kMD + c * 8,
              (char*)xs + (size_t)(w * U + u) * 1024);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = (w * U + u) * 64 + l;
    const int d = q >> 2, c = (q & 3) ^ xt_swz(d);
    lds_dma16(Xt + (size_t)d * n_obs + n0 + c * 8,
              (char*)xts + (size_t)(w * U + u) * 1024);
  }
  if (l < YL) lds_dma4(y + n0 + w * YL + l, ys + w * YL);
}

// s_waitcnt immediate that waits until at most n vector-memory loads of
// this wave are outstanding (expcnt and lgkmcnt left at their maxima).
template <int N>
constexpr int vmcnt_imm() {
  static_assert(N >= 0 && N < 16, "vmcnt field");
  return 0xF70 | N;
}

// Partial gradient (and optionally partial Bernoulli log-likelihood) of the
// chains [16*WAVES*blockIdx.x, +16*WAVES) over rows
// [blockIdx.y*rows_per_split, +rows), 16 chains per wave.  With 8 waves
// (128 chains) every staged tile serves twice the chains of the 4-wave
// layout, halving the X and X^T bytes staged per gradient; the workgroup
// then takes one CU (84 KiB of LDS) with the same 8 waves per CU.
// Tiles of 32 rows of X and X^T go straight from memory to LDS (no
// registers) into one of two buffers while the previous tile is used, so a
// tile's load latency hides behind the previous tile's MFMAs.  LDS layouts
// (16-byte chunks, XOR-swizzled so that the MFMA fragment reads of 16 rows
// hit distinct banks): X tile chunk (row, c) at row * 32 + (c ^ (row & 15)),
// X^T tile chunk (d, c) at d * 4 + (c ^ xt_swz(d)).  Each lane writes its
// four consecutive residual rows as one 8-byte store (R pitch 40 bf16 keeps
// those stores and the 16-byte fragment reads free of bank conflicts).
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES, 8 / WAVES)
    glm_grad_mfma(const __bf16* __restrict__ X, const __bf16* __restrict__ Xt,
                  const float* __restrict__ y, int n_obs,
                  const __bf16* __restrict__ Bhi,
                  const __bf16* __restrict__ Blo, int n_chains,
                  int rows_per_split, float* __restrict__ out_g,
                  float* __restrict__ out_ll) {
  __shared__ __align__(1024) __bf16 xs0[kMRows * kMD];
  __shared__ __align__(1024) __bf16 xs1[kMRows * kMD];
  __shared__ __align__(1024) __bf16 xts0[kMD * kMRows];
  __shared__ __align__(1024) __bf16 xts1[kMD * kMRows];
  constexpr int CH = 16 * WAVES;  // chains of the workgroup
  __shared__ __align__(16) __bf16 rs[2 * CH * kRPitch];
  __shared__ __align__(16) float ys0[kMRows];
  __shared__ __align__(16) float ys1[kMRows];
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63, m = l & 15, g = l >> 4;
  const int c0 = blockIdx.x * CH + 16 * w;
  bf16x8 bh[kMD / 32], bl[kMD / 32];
#pragma unroll
  for (int ks = 0; ks < kMD / 32; ++ks) {
    const size_t off = (size_t)(c0 + m) * kMD + 32 * ks + 8 * g;
    bh[ks] = *(const bf16x8*)(Bhi + off);
    bl[ks] = *(const bf16x8*)(Blo + off);
  }
  // keep the fragments in registers (not re-read from memory per tile,
  // whose waits would also wait for the tile loads in flight)
#pragma unroll
  for (int ks = 0; ks < kMD / 32; ++ks) asm volatile("" : "+v"(bh[ks]), "+v"(bl[ks]));
  f32x4 acc[kMD / 16];
#pragma unroll
  for (int nt = 0; nt < kMD / 16; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  float lls = 0.f;
  const bool want_ll = out_ll != nullptr;
  const int r0 = blockIdx.y * rows_per_split;
  const int r1 = min(n_obs, r0 + rows_per_split);
  __bf16* rh_w = rs + (16 * w) * kRPitch;
  __bf16* rl_w = rs + CH * kRPitch + (16 * w) * kRPitch;

  auto compute = [&](const __bf16* xs, const __bf16* xts, const float* ys) {
    f32x4 s0 = f32x4{0.f, 0.f, 0.f, 0.f}, s1 = s0;
#pragma unroll
    for (int ks = 0; ks < kMD / 32; ++ks) {
      const int c = 4 * ks + g;
      const bf16x8 a0 = *(const bf16x8*)&xs[(m * 32 + (c ^ m)) * 8];
      const bf16x8 a1 = *(const bf16x8*)&xs[((16 + m) * 32 + (c ^ m)) * 8];
      s0 = mfma16(a0, bh[ks], s0);
      s1 = mfma16(a1, bh[ks], s1);
      s0 = mfma16(a0, bl[ks], s0);
      s1 = mfma16(a1, bl[ks], s1);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      bf16x4 hv, lv;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * mt + 4 * g + i;
        const float sv = mt ? s1[i] : s0[i];
        const float yv = ys[row];
        const float e = __expf(-fabsf(sv));
        const float inv = recip_1p(e);
        const float sig = sv >= 0.f ? inv : e * inv;
        const float r = yv - sig;
        if (want_ll) {
          const float sp_abs = log1p_unit(e);
          const float sp_pos = sp_abs + fmaxf(sv, 0.f);
          const float sp_neg = sp_abs + fmaxf(-sv, 0.f);
          const float w0 = 1.f - yv;
          lls += (w0 == 0.f ? 0.f : -sp_pos * w0) + (yv == 0.f ? 0.f : -sp_neg * yv);
        }
        hv[i] = (__bf16)r;
        lv[i] = (__bf16)(r - (float)hv[i]);
      }
      *(bf16x4*)&rh_w[m * kRPitch + 16 * mt + 4 * g] = hv;
      *(bf16x4*)&rl_w[m * kRPitch + 16 * mt + 4 * g] = lv;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bf16x8 ah = *(const bf16x8*)&rh_w[m * kRPitch + 8 * g];
    const bf16x8 al = *(const bf16x8*)&rl_w[m * kRPitch + 8 * g];
#pragma unroll
    for (int nt0 = 0; nt0 < kMD / 16; nt0 += 4) {
      bf16x8 b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int d = 16 * (nt0 + u) + m;
        b[u] = *(const bf16x8*)&xts[(d * 4 + (g ^ xt_swz(d))) * 8];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc[nt0 + u] = mfma16(ah, b[u], acc[nt0 + u]);
        acc[nt0 + u] = mfma16(al, b[u], acc[nt0 + u]);
      }
    }
  };
  // tiles of this split, two per round (buffers 0 and 1): the next tile's
  // LPW loads per wave are issued, then the wave waits until only those are
  // outstanding (vmcnt(LPW)), i.e. until its loads of the current tile have
  // landed; the barrier makes every wave's part visible
  constexpr int LPW = 2 * (16 / WAVES) + 1;
  if (r0 < r1) stage_tile<WAVES>(X, Xt, y, n_obs, r0, xs0, xts0, ys0, w, l);
  for (int n0 = r0; n0 < r1; n0 += 2 * kMRows) {
    const int n1 = n0 + kMRows;
    if (n1 < r1) {
      stage_tile<WAVES>(X, Xt, y, n_obs, n1, xs1, xts1, ys1, w, l);
      __builtin_amdgcn_s_waitcnt(vmcnt_imm<LPW>());  // tile n0 has landed
    } else {
      __builtin_amdgcn_s_waitcnt(vmcnt_imm<0>());
    }
    __syncthreads();
    compute(xs0, xts0, ys0);
    __syncthreads();
    if (n1 >= r1) break;
    const int n2 = n1 + kMRows;
    if (n2 < r1) {
      stage_tile<WAVES>(X, Xt, y, n_obs, n2, xs0, xts0, ys0, w, l);
      __builtin_amdgcn_s_waitcnt(vmcnt_imm<LPW>());
    } else {
      __builtin_amdgcn_s_waitcnt(vmcnt_imm<0>());
    }
    __syncthreads();
    compute(xs1, xts1, ys1);
    __syncthreads();
  }
  float* og = out_g + ((size_t)blockIdx.y * n_chains + c0) * kMD;
#pragma unroll
  for (int nt = 0; nt < kMD / 16; ++nt)
#pragma unroll
    for (int i = 0; i < 4; ++i) og[(size_t)(4 * g + i) * kMD + 16 * nt + m] = acc[nt][i];
  if (want_ll) {
    lls += __shfl_xor(lls, 16);
    lls += __shfl_xor(lls, 32);
    if (g == 0) out_ll[(size_t)blockIdx.y * n_chains + c0 + m] = lls;
  }
}
// ---------------------------------------------------------------- per chain
// One wave per chain, element e in lane e % 64, slot e / 64 (4 slots).
__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

__device__ inline float mfma_step_size(const hipmc_glm_hmc_args& a, int c) {
  return a.step_size[a.per_chain_step ? c : 0];
}

__device__ inline Seed mfma_iter_seed(const hipmc_glm_hmc_args& a, int it) {
  Seed base{a.seed[0], a.seed[1]};
  return a.fold_seed ? fold_in(base, (uint64_t)(a.iter0 + it)) : base;
}

// Writes Bhi/Blo of the current state (bootstrap gradient input).
__global__ void mfma_prepare(hipmc_glm_hmc_args a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)a.n_chains * kMD) return;
  __bf16 hi, lo;
  split_bf16(a.state[i], hi, lo);
  ((__bf16*)a.ws_b)[i] = hi;
  ((__bf16*)a.ws_b)[(size_t)a.n_chains * kMD + i] = lo;
}

// Sum of the split-N partial gradients of chain c at slot k, plus prior.
__device__ inline float reduce_grad(const hipmc_glm_hmc_args& a, int c, int e,
                                    float q, float inv_ps2) {
  float s = 0.f;
  for (int sp = 0; sp < a.n_split; ++sp)
    s += a.ws_g[((size_t)sp * a.n_chains + c) * kMD + e];
  return s - q * inv_ps2;
}

__device__ inline double reduce_ll(const hipmc_glm_hmc_args& a, int c) {
  double s = 0.0;
  for (int sp = 0; sp < a.n_split; ++sp) s += (double)a.ws_ll[(size_t)sp * a.n_chains + c];
  return s;
}

// mode 0: start of iteration `it` (momentum draw, first half kick, first
// drift); mode 1: after a gradient of a middle leapfrog step (kick, drift);
// mode 2: after the last gradient (kick, un-kick, MH accept); mode 3:
// bootstrap (tlp and grad of the state).
__global__ void __launch_bounds__(256) mfma_chain_step(hipmc_glm_hmc_args a,
                                                        int mode, int it) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= a.n_chains) return;
  const float ps = (float)a.prior_scale;
  const float inv_ps = 1.f / ps;
  const float inv_ps2 = inv_ps * inv_ps;
  const size_t row = (size_t)c * kMD;
  float* q = a.ws_q + row;
  float* p = a.ws_p + row;
  __bf16* bh = (__bf16*)a.ws_b + row;
  __bf16* bl = (__bf16*)a.ws_b + (size_t)a.n_chains * kMD + row;
  const float eps = mfma_step_size(a, c);
  const float half = 0.5f * eps;
  if (mode == 3) {
    double lp = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int e = k * 64 + lane;
      const float qv = a.state[row + e];
      a.grad[row + e] = reduce_grad(a, c, e, qv, inv_ps2);
      const float z = qv * inv_ps;
      lp += (double)(-0.5f * z * z);
    }
    lp = wave_sum(lp);
    if (lane == 0)
      a.tlp[c] = (float)(lp + reduce_ll(a, c) +
                         kMD * (-0.9189385332046727 - log((double)ps)));
    return;
  }
  const Seed sd = mfma_iter_seed(a, it);
  if (mode == 0) {
    const Seed prop = split(sd, 0);
    double ke = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int e = k * 64 + lane;
      const U4 w = philox(U4{(uint32_t)(e >> 2), (uint32_t)c, a.key,
                             kPurposeMomentum},
                          prop.k0, prop.k1);
      const bool hi_pair = (e & 2) != 0;
      double n0, n1;
      box_muller(hi_pair ? w.z : w.x, hi_pair ? w.w : w.y, n0, n1);
      const float pv0 = (float)((e & 1) ? n1 : n0);
      ke += (double)pv0 * (double)pv0;
      float pv = pv0 + half * a.grad[row + e];
      const float qv = a.state[row + e] + eps * pv;
      p[e] = pv;
      q[e] = qv;
      __bf16 hi, lo;
      split_bf16(qv, hi, lo);
      bh[e] = hi;
      bl[e] = lo;
    }
    ke = wave_sum(ke);
    if (lane == 0) a.ws_ke0[c] = 0.5 * ke;
    return;
  }
  if (mode == 1) {
    for (int k = 0; k < 4; ++k) {
      const int e = k * 64 + lane;
      const float gv = reduce_grad(a, c, e, q[e], inv_ps2);
      const float pv = p[e] + eps * gv;
      const float qv = q[e] + eps * pv;
      p[e] = pv;
      q[e] = qv;
      __bf16 hi, lo;
      split_bf16(qv, hi, lo);
      bh[e] = hi;
      bl[e] = lo;
    }
    return;
  }
  // mode 2: last kick, un-kick, kinetic energy, log-prob, MH accept
  float gv[4], qv[4];
  double ke = 0.0, lp = 0.0;
  for (int k = 0; k < 4; ++k) {
    const int e = k * 64 + lane;
    qv[k] = q[e];
    gv[k] = reduce_grad(a, c, e, qv[k], inv_ps2);
    float pv = p[e] + eps * gv[k];
    pv = pv - half * gv[k];
    ke += (double)pv * (double)pv;
    const float z = qv[k] * inv_ps;
    lp += (double)(-0.5f * z * z);
  }
  wave_sum2(ke, lp, ke, lp);
  const double t1 = lp + reduce_ll(a, c) +
                    kMD * (-0.9189385332046727 - log((double)ps));
  const double t0 = (double)a.tlp[c];
  double corr = a.ws_ke0[c] - 0.5 * ke;
  if (!__builtin_isfinite(corr)) corr = -__builtin_inf();
  double lar = (t1 - t0) + corr;
  if (!__builtin_isfinite(lar)) lar = -__builtin_inf();
  const Seed acc = split(sd, 1);
  const U4 uw = philox(U4{0u, (uint32_t)c, a.key, kPurposeMHAccept}, acc.k0,
                       acc.k1);
  const bool accept = log((double)uniform_from<float>(uw.x, uw.y)) < lar;
  if (accept) {
    for (int k = 0; k < 4; ++k) {
      const int e = k * 64 + lane;
      a.state[row + e] = qv[k];
      a.grad[row + e] = gv[k];
    }
  }
  if (lane == 0) {
    if (accept) a.tlp[c] = (float)t1;
    const size_t o = (size_t)it * a.n_chains + c;
    if (a.out_log_accept_ratio) a.out_log_accept_ratio[o] = (float)lar;
    if (a.out_is_accepted) a.out_is_accepted[o] = accept;
  }
  if (a.trace_state) {
    const int64_t gi = a.iter0 + it - a.trace_start;
    const int every = a.trace_every > 0 ? a.trace_every : 1;
    if (gi >= 0 && gi % every == 0 && gi / every < a.n_trace) {
      float* tr = a.trace_state + ((size_t)(gi / every) * a.n_chains + c) * kMD;
      for (int k = 0; k < 4; ++k) {
        const int e = k * 64 + lane;
        tr[e] = a.state[row + e];
      }
    }
  }
}

}  // namespace hipmc
