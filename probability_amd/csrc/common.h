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
// initial positions of windowed adaptation (uniform on (-2, 2))
constexpr uint32_t kPurposeInit = 0x494E4954u;

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
// A value every lane already holds, marked wave-uniform (a scalar register):
// branches on it become scalar branches and selects on it plain v_cndmask,
// instead of exec-mask regions.
__device__ inline float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline double uniform(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __longlong_as_double(((int64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// v + v[perm(lane)] for a DPP permutation inside each row of 16 lanes.  The
// float form compiles to one v_add_f32_dpp.
template <int Ctrl>
__device__ inline float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(
                 0, __float_as_int(v), Ctrl, 0xF, 0xF, true));
}
template <int Ctrl>
__device__ inline double dpp_add(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, Ctrl, 0xF,
                                             0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), Ctrl,
                                             0xF, 0xF, true);
  return v + __longlong_as_double(((int64_t)(uint32_t)hi << 32) |
                                  (uint32_t)lo);
}
// Row broadcast (row_bcast:15 / row_bcast:31) into the rows of RowMask; the
// other rows add zero.
template <int Ctrl, int RowMask>
__device__ inline float dpp_add_rows(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(
                 0, __float_as_int(v), Ctrl, RowMask, 0xF, false));
}
template <int Ctrl, int RowMask>
__device__ inline double dpp_add_rows(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, Ctrl,
                                             RowMask, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), Ctrl,
                                             RowMask, 0xF, false);
  return v + __longlong_as_double(((int64_t)(uint32_t)hi << 32) |
                                  (uint32_t)lo);
}

constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

// Row sums: after the four butterfly stages every lane of a row of 16 holds
// the same row sum (each add is commutative, so all lanes round alike).
template <typename A>
__device__ inline A row_sum(A v) {
  v = dpp_add<kDppQuadXor1>(v);
  v = dpp_add<kDppQuadXor2>(v);
  v = dpp_add<kDppHalfMirror>(v);
  return dpp_add<kDppMirror>(v);
}
// v of lane i of this lane's quad (DPP quad_perm [i, i, i, i]).  Every lane
// reads a lane of its own quad, so no lane keeps its old value: the move
// carries no old operand, which lets the compiler fold it into the
// instructions that use it.
template <int I>
__device__ inline float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), I * 0x55,
                                                 0xF, 0xF, false));
}
template <int I>
__device__ inline double quad_bcast(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, I * 0x55, 0xF,
                                          0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), I * 0x55,
                                          0xF, 0xF, false);
  return __longlong_as_double(((int64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
// The same for an index that is a constant after unrolling.
template <typename A>
__device__ inline A quad_bcast(A v, int i) {
  switch (i & 3) {
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
  }
}

// Combines row sums r0..r3: rows 1 and 3 take r1 + r0 and r3 + r2, then row 3
// takes (r3 + r2) + (r1 + r0), read from lane 63.  The result equals
// (r0 + r1) + (r2 + r3) bit for bit (commutative adds).
template <typename A>
__device__ inline A rows_total(A v) {
  v = dpp_add_rows<kDppRowBcast15, 0xA>(v);
  v = dpp_add_rows<kDppRowBcast31, 0xC>(v);
  return readlane(v, 63);
}

// Sum over all 64 lanes; every lane receives the (wave-uniform) result.
// Requires a full EXEC mask.  Deterministic: fixed reduction order.
// src read through a DPP pattern into the lanes of RowMask x BankMask (a
// bank is 4 lanes of a row); the other lanes keep `old`.
template <int Ctrl, int BankMask>
__device__ inline float dpp_mov(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      __float_as_int(old), __float_as_int(src), Ctrl, 0xF, BankMask, false));
}
template <int Ctrl, int BankMask>
__device__ inline double dpp_mov(double old, double src) {
  const int64_t o = __double_as_longlong(old), b = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)b,
                                             Ctrl, 0xF, BankMask, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32),
                                             (int)(uint32_t)(b >> 32), Ctrl,
                                             0xF, BankMask, false);
  return __longlong_as_double(((int64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
// x + src[perm(lane)] inside each row of 16 lanes.
template <int Ctrl>
__device__ inline float dpp_add_from(float x, float src) {
  return x + __int_as_float(__builtin_amdgcn_update_dpp(
                 0, __float_as_int(src), Ctrl, 0xF, 0xF, true));
}
template <int Ctrl>
__device__ inline double dpp_add_from(double x, double src) {
  const int64_t b = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, Ctrl, 0xF,
                                             0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), Ctrl,
                                             0xF, 0xF, true);
  return x + __longlong_as_double(((int64_t)(uint32_t)hi << 32) |
                                  (uint32_t)lo);
}
constexpr int kDppRowShl4 = 0x104;   // lane i reads lane i + 4 (in its row)
constexpr int kDppRowShr4 = 0x114;   // lane i reads lane i - 4
constexpr int kDppRowRor8 = 0x128;   // lane i reads lane i ^ 8 (halves swap)
constexpr int kDppRowRor4 = 0x124;   // lane i reads lane i - 4 (mod 16)
constexpr int kDppRowRor12 = 0x12C;  // lane i reads lane i + 4 (mod 16)

// src of the lane a DPP pattern names; every lane reads a lane of its own
// row, so no old operand is needed.
template <int Ctrl>
__device__ inline float dpp_read(float src) {
  return __int_as_float(
      __builtin_amdgcn_mov_dpp(__float_as_int(src), Ctrl, 0xF, 0xF, false));
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// 4x4 transpose inside each quad of lanes: element t of lane q becomes
// element q of lane t.  Two exchange stages: lanes q and q^1 swap the
// elements whose index differs from the lane's in bit 0, then lanes q and
// q^2 those that differ in bit 1.  b0, b1: bits 0 and 1 of the lane.
__device__ inline void quad_transpose(f32x4& v, bool b0, bool b1) {
  const float p0 = dpp_read<kDppQuadXor1>(v[1]);
  const float p1 = dpp_read<kDppQuadXor1>(v[0]);
  const float p2 = dpp_read<kDppQuadXor1>(v[3]);
  const float p3 = dpp_read<kDppQuadXor1>(v[2]);
  const float y0 = b0 ? p0 : v[0], y1 = b0 ? v[1] : p1;
  const float y2 = b0 ? p2 : v[2], y3 = b0 ? v[3] : p3;
  const float s0 = dpp_read<kDppQuadXor2>(y2);
  const float s1 = dpp_read<kDppQuadXor2>(y3);
  const float s2 = dpp_read<kDppQuadXor2>(y0);
  const float s3 = dpp_read<kDppQuadXor2>(y1);
  v[0] = b1 ? s0 : y0;
  v[1] = b1 ? s1 : y1;
  v[2] = b1 ? y2 : s2;
  v[3] = b1 ? y3 : s3;
}

// v_mfma_f32_4x4x1_16b_f32: 16 blocks of 4 lanes; block b adds the outer
// product of its lanes' a (rows) and b (columns) to a 4x4 tile held as
// register m of lane 4b + n (row m, column n).
__device__ inline f32x4 mfma_4x4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

// The row sums of four values a0..a3 (the partial dot products of four rows
// of a batch), delivered so that lane L holds the sum over its 16-lane DPP
// row of a_{L & 3}: bit for bit row_sum(a_{L & 3}) (the same pairs are
// added in the same order: lane pairs L^1, L^2, then the quads 0+1 / 2+3,
// then the halves; float adds are commutative), in 13 instructions instead
// of four row_sum (16) and a pick.  b0, b1: bits 0 and 1 of the lane.
template <typename A>
__device__ inline A row_sum_pick4(A a0, A a1, A a2, A a3, bool b0, bool b1) {
  // pairs L, L^1: keep rows L&1 and 2 + (L&1), send the partner's
  A x = b0 ? a1 : a0, sx = b0 ? a0 : a1;
  A y = b0 ? a3 : a2, sy = b0 ? a2 : a3;
  x = dpp_add_from<kDppQuadXor1>(x, sx);
  y = dpp_add_from<kDppQuadXor1>(y, sy);
  // pairs L, L^2: keep row L & 3
  A k = b1 ? y : x;
  const A s = b1 ? x : y;
  k = dpp_add_from<kDppQuadXor2>(k, s);
  // quads 0+1 and 2+3 (lane L with L ^ 4, the same row)
  A t = dpp_mov<kDppRowShl4, 0x5>(k, k);
  t = dpp_mov<kDppRowShr4, 0xA>(t, k);
  k = k + t;
  // halves (lane L with L ^ 8)
  return dpp_add_from<kDppRowRor8>(k, k);
}

__device__ inline float wave_sum(float v) { return rows_total(row_sum(v)); }
__device__ inline double wave_sum(double v) { return rows_total(row_sum(v)); }

// Two independent sums reduced together (interleaved for latency).
template <typename A, typename B>
__device__ inline void wave_sum2(A a, B b, A& ra, B& rb) {
  a = dpp_add<kDppQuadXor1>(a);
  b = dpp_add<kDppQuadXor1>(b);
  a = dpp_add<kDppQuadXor2>(a);
  b = dpp_add<kDppQuadXor2>(b);
  a = dpp_add<kDppHalfMirror>(a);
  b = dpp_add<kDppHalfMirror>(b);
  a = dpp_add<kDppMirror>(a);
  b = dpp_add<kDppMirror>(b);
  a = dpp_add_rows<kDppRowBcast15, 0xA>(a);
  b = dpp_add_rows<kDppRowBcast15, 0xA>(b);
  a = dpp_add_rows<kDppRowBcast31, 0xC>(a);
  b = dpp_add_rows<kDppRowBcast31, 0xC>(b);
  ra = readlane(a, 63);
  rb = readlane(b, 63);
}

// Four independent wave sums with their DPP stages interleaved (each sum
// in the order of wave_sum).
template <typename A>
__device__ inline void wave_sum4(A a, A b, A c, A d, A& ra, A& rb, A& rc,
                                 A& rd) {
  a = dpp_add<kDppQuadXor1>(a);
  b = dpp_add<kDppQuadXor1>(b);
  c = dpp_add<kDppQuadXor1>(c);
  d = dpp_add<kDppQuadXor1>(d);
  a = dpp_add<kDppQuadXor2>(a);
  b = dpp_add<kDppQuadXor2>(b);
  c = dpp_add<kDppQuadXor2>(c);
  d = dpp_add<kDppQuadXor2>(d);
  a = dpp_add<kDppHalfMirror>(a);
  b = dpp_add<kDppHalfMirror>(b);
  c = dpp_add<kDppHalfMirror>(c);
  d = dpp_add<kDppHalfMirror>(d);
  a = dpp_add<kDppMirror>(a);
  b = dpp_add<kDppMirror>(b);
  c = dpp_add<kDppMirror>(c);
  d = dpp_add<kDppMirror>(d);
  a = dpp_add_rows<kDppRowBcast15, 0xA>(a);
  b = dpp_add_rows<kDppRowBcast15, 0xA>(b);
  c = dpp_add_rows<kDppRowBcast15, 0xA>(c);
  d = dpp_add_rows<kDppRowBcast15, 0xA>(d);
  a = dpp_add_rows<kDppRowBcast31, 0xC>(a);
  b = dpp_add_rows<kDppRowBcast31, 0xC>(b);
  c = dpp_add_rows<kDppRowBcast31, 0xC>(c);
  d = dpp_add_rows<kDppRowBcast31, 0xC>(d);
  ra = readlane(a, 63);
  rb = readlane(b, 63);
  rc = readlane(c, 63);
  rd = readlane(d, 63);
}

// ---------------------------------------------------------------- scalars
// Read-only pointer in the global address space.  Loads through it compile to
// global loads, which count only against vmcnt; loads through a generic
// pointer (one read from a table in memory) compile to flat loads, which
// also count against lgkmcnt, so every LDS wait would wait for them too.
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;
// Pointer in the LDS address space (ds_read / ds_write; through a generic
// pointer the compiler emits flat accesses).
template <typename T>
using lptr = __attribute__((address_space(3))) T*;

__device__ inline double neg_inf() { return -__builtin_inf(); }

// An empty volatile asm: first in a wave-uniform branch whose body is
// rarely taken, it keeps the compiler from if-converting (speculating) the
// body into every execution.
__device__ inline void keep_branch() { asm volatile(""); }

// 1 / d for normal, finite d: hardware reciprocal plus one Newton step
// (float), IEEE division (double).
__device__ inline float recip(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(fmaf(-d, r, 1.0f), r, r);
}
__device__ inline double recip(double d) { return 1.0 / d; }

// log1p(-u) for a uniform u in [0, 1): 1 - u is exact for the 24-bit float
// uniforms, so log(1 - u) (float) is as accurate as the log; libm (double).
__device__ inline float log1m(float u) { return logf(1.0f - u); }
__device__ inline double log1m(double u) { return log1p(-u); }

// 1 / (1 + e) for e in [0, 1]: hardware reciprocal plus one Newton step
// (float), IEEE division (double).
__device__ inline float recip_1p(float e) {
  const float d = 1.0f + e;
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(fmaf(-d, r, 1.0f), r, r);
}
__device__ inline double recip_1p(double e) { return 1.0 / (1.0 + e); }

// log1p(u) for u in [0, 1].  Float: log1p(u) = 2 atanh(s) with
// s = u / (2 + u) in [0, 1/3] (one Newton correction after the hardware
// reciprocal) and the odd series summed to s^15 (truncation below 2^-27
// relative); within 2 ulp of the correctly rounded value on [0, 1] in a
// float32 emulation over 3e5 points.  NaN propagates.  Double: libm.
__device__ inline float log1p_unit(float u) {
  const float d = 2.0f + u;
  const float r = __builtin_amdgcn_rcpf(d);
  float s = u * r;
  s = fmaf(fmaf(-s, d, u), r, s);
  const float s2 = s * s;
  float q = 1.0f / 15.0f;
  q = fmaf(q, s2, 1.0f / 13.0f);
  q = fmaf(q, s2, 1.0f / 11.0f);
  q = fmaf(q, s2, 1.0f / 9.0f);
  q = fmaf(q, s2, 1.0f / 7.0f);
  q = fmaf(q, s2, 1.0f / 5.0f);
  q = fmaf(q, s2, 1.0f / 3.0f);
  const float t = s2 * q;
  const float s2x = 2.0f * s;
  return fmaf(s2x, t, s2x);
}
__device__ inline double log1p_unit(double u) { return log1p(u); }

// log1p(v) for v >= 0: the series above on [0, 1], log(1 + v) beyond
// (where rounding 1 + v costs under one ulp of the result).
__device__ inline float log1p_pos(float v) {
  return v <= 1.0f ? log1p_unit(v) : logf(1.0f + v);
}
__device__ inline double log1p_pos(double v) { return log1p(v); }

// log(exp(x) + exp(y)) as `math/generic.py:605-611`:
// larger + softplus((x - larger) + (y - larger)); the argument is -|x - y|
// (or NaN), so the softplus is log1p(exp(.)) of a value in [0, 1], taken
// at precision P (the chain's state dtype) around double endpoints.
template <typename P>
__device__ inline double log_add_exp(double x, double y) {
  const double larger = fmax(x, y);
  const double z = (x - larger) + (y - larger);
  return larger + (double)log1p_unit(exp((P)z));
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

// softplus(+-x) and sigmoid(+-x) from one exp, one log1p and one
// reciprocal: with e = exp(-|x|) and l = log1p(e),
//   softplus(x) = max(x, 0) + l,      softplus(-x) = max(-x, 0) + l,
//   sigmoid(x) = 1/(1+e) or e/(1+e) (x >= 0 or not), sigmoid(-x) the other.
template <typename T>
struct SoftplusParts {
  T l, sig, sig_neg;
};
template <typename T>
__device__ inline SoftplusParts<T> softplus_parts(T x) {
  const T e = exp(-fabs(x));
  const T r = recip_1p(e);
  const T er = e * r;
  const bool pos = x >= T(0);
  return SoftplusParts<T>{log1p_unit(e), pos ? r : er, pos ? er : r};
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
