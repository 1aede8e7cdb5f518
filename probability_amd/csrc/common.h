// Device helpers shared by every hipmc kernel: Philox4x32-10, the
// counter-keyed draw schedule, wave64 reductions and a few numerically
// careful scalar functions.  CDNA4 (gfx950) only: wave64, DPP row ops.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hipmc {

// ---------------------------------------------------------------- Philox
// Salmon et al., SC'11: ten rounds of the 4x32 block function with the Weyl
// key schedule.
struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// Purposes (counter word 3); depth is OR-ed into the low bits.
constexpr uint32_t kPurposeSplit = 0x53504C54u;
constexpr uint32_t kPurposeFold = 0x464F4C44u;
constexpr uint32_t kPurposeMomentum = 0x4D4F4D00u;
constexpr uint32_t kPurposeMHAccept = 0x4D484100u;
constexpr uint32_t kPurposeTree = 0x54520000u;
constexpr uint32_t kPurposeLeaf = 0x4C460000u;

struct Seed {
  uint32_t k0, k1;
};

// Seed of loop step t: one Philox block keyed by the base seed.
__host__ __device__ inline Seed fold_in(Seed s, uint64_t t) {
  U4 o = philox(U4{(uint32_t)t, (uint32_t)(t >> 32), 0u, kPurposeFold}, s.k0,
                s.k1);
  return Seed{o.x, o.y};
}

// Child i of split_seed(s, n).
__host__ __device__ inline Seed split(Seed s, uint32_t i) {
  U4 o = philox(U4{i, 0u, 0u, kPurposeSplit}, s.k0, s.k1);
  return Seed{o.x, o.y};
}

// Uniform [0, 1) from one (float) or two (double) words.
template <typename T>
__device__ inline T uniform_from(uint32_t a, uint32_t b);
template <>
__device__ inline float uniform_from<float>(uint32_t a, uint32_t) {
  return (float)(a >> 8) * 5.9604644775390625e-08f;  // 2^-24
}
template <>
__device__ inline double uniform_from<double>(uint32_t a, uint32_t b) {
  const uint64_t u53 = ((uint64_t)a << 21) | (uint64_t)(b >> 11);
  return (double)u53 * 1.1102230246251565e-16;  // 2^-53
}

// Box-Muller in double on a word pair.
__device__ inline void box_muller(uint32_t a, uint32_t b, double& n0,
                                  double& n1) {
  const double u1 = ((double)a + 1.0) * 2.3283064365386963e-10;  // 2^-32
  const double u2 = (double)b * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586 * u2;
  n0 = r * cos(th);
  n1 = r * sin(th);
}

// ---------------------------------------------------------------- wave64
__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

// DPP control words (GFX9 encoding).
constexpr int kDppQuadXor1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // lane i <-> 7-i within 8
constexpr int kDppMirror = 0x140;      // lane i <-> 15-i within 16

template <int Ctrl>
__device__ inline float dpp_f(float v) {
  return __int_as_float(
      __builtin_amdgcn_mov_dpp(__float_as_int(v), Ctrl, 0xF, 0xF, false));
}
template <int Ctrl>
__device__ inline double dpp_d(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, Ctrl, 0xF, 0xF,
                                          false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), Ctrl, 0xF,
                                          0xF, false);
  return __longlong_as_double(((int64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int Ctrl>
__device__ inline float dpp_any(float v) { return dpp_f<Ctrl>(v); }
template <int Ctrl>
__device__ inline double dpp_any(double v) { return dpp_d<Ctrl>(v); }

__device__ inline float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double readlane_d(double v, int l) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double(((int64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ inline float readlane(float v, int l) { return readlane_f(v, l); }
__device__ inline double readlane(double v, int l) { return readlane_d(v, l); }

// Sum over all 64 lanes; every lane receives the (wave-uniform) result.
// Requires a full EXEC mask.  Rows of 16 are reduced with DPP, the four row
// sums are combined through readlane in a fixed order (deterministic).
__device__ inline float wave_sum(float v) {
  v += dpp_f<kDppQuadXor1>(v);
  v += dpp_f<kDppQuadXor2>(v);
  v += dpp_f<kDppHalfMirror>(v);
  v += dpp_f<kDppMirror>(v);
  return (readlane_f(v, 0) + readlane_f(v, 16)) +
         (readlane_f(v, 32) + readlane_f(v, 48));
}
__device__ inline double wave_sum(double v) {
  v += dpp_d<kDppQuadXor1>(v);
  v += dpp_d<kDppQuadXor2>(v);
  v += dpp_d<kDppHalfMirror>(v);
  v += dpp_d<kDppMirror>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) +
         (readlane_d(v, 32) + readlane_d(v, 48));
}

// Two independent sums reduced together (interleaved for latency).
template <typename A, typename B>
__device__ inline void wave_sum2(A a, B b, A& ra, B& rb) {
  a += dpp_any<kDppQuadXor1>(a);
  b += dpp_any<kDppQuadXor1>(b);
  a += dpp_any<kDppQuadXor2>(a);
  b += dpp_any<kDppQuadXor2>(b);
  a += dpp_any<kDppHalfMirror>(a);
  b += dpp_any<kDppHalfMirror>(b);
  a += dpp_any<kDppMirror>(a);
  b += dpp_any<kDppMirror>(b);
  ra = (readlane(a, 0) + readlane(a, 16)) + (readlane(a, 32) + readlane(a, 48));
  rb = (readlane(b, 0) + readlane(b, 16)) + (readlane(b, 32) + readlane(b, 48));
}

// ---------------------------------------------------------------- scalars
__device__ inline double neg_inf() { return -__builtin_inf(); }

// log(exp(x) + exp(y)) as `math/generic.py:605-611`:
// larger + softplus((x - larger) + (y - larger)).
__device__ inline double log_add_exp(double x, double y) {
  const double larger = fmax(x, y);
  const double z = (x - larger) + (y - larger);
  return larger + (log1p(exp(-fabs(z))) + fmax(z, 0.0));
}

template <typename T>
__device__ inline T softplus(T x) {
  return log1p(exp(-fabs(x))) + fmax(x, T(0));
}
template <typename T>
__device__ inline T sigmoid(T x) {
  const T e = exp(-fabs(x));
  return x >= T(0) ? T(1) / (T(1) + e) : e / (T(1) + e);
}
// Softplus bijector forward (`bijectors/softplus.py:36-50`): log1p(exp(x))
// below -9 (float) / -20 (double), else softplus.
template <typename T>
__device__ inline T softplus_forward(T x);
template <>
__device__ inline float softplus_forward<float>(float x) {
  return x < -9.0f ? log1pf(expf(x)) : softplus(x);
}
template <>
__device__ inline double softplus_forward<double>(double x) {
  return x < -20.0 ? log1p(exp(x)) : softplus(x);
}

}  // namespace hipmc
