// Wave-level analytic log-density + gradient of the target specs.
//
// One wavefront evaluates one chain.  The flat state (parts concatenated) is
// spread over the wave: element e lives in lane e % 64, register slot e / 64.
// Log-density terms follow TFP's distributions (paths relative to
// tensorflow_probability/python/): Normal distributions/normal.py:182-188,
// Bernoulli(logits) bernoulli.py:119-135, HalfCauchy half_cauchy.py:138-145,
// HalfNormal half_normal.py:146-153; the bijector composition follows
// mcmc/transformed_kernel.py:86-128 (tlp(f(x)) + fldj(x)).  Constant terms
// are folded into problem.log_norm on the host.
#pragma once

#include "../../include/hipmc.h"
#include "common.h"

namespace hipmc {

// Target families: one kernel instantiation per family keeps register
// pressure to what that family needs.
enum Family { kFamilyHier = 0, kFamilyGlm = 1, kFamilyTest = 2 };

__host__ __device__ inline int family_of(int kind) {
  switch (kind) {
    case HIPMC_TARGET_HIER_NORMAL: return kFamilyHier;
    case HIPMC_TARGET_LOGISTIC:
    case HIPMC_TARGET_LINEAR: return kFamilyGlm;
    default: return kFamilyTest;
  }
}

// 16-byte vector of T: GLM rows are read 16 bytes per lane (one wave per
// chain is latency bound on X, so bytes in flight per wave set its speed).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
// Gradient blocks of 16 * Vec16<T>::n coefficients covering K*64.
template <typename T, int K>
constexpr int kGlmBlocks = (K * 64 + 16 * Vec16<T>::n - 1) / (16 * Vec16<T>::n);

// Bytes of one half of the buffer where the W waves of one GLM chain add
// their partial gradients [W][K*64], log-likelihoods [W][64] (double) and
// residual sums [W][64].
template <typename T, int K>
__host__ __device__ constexpr size_t glm_red_bytes(int w) {
  return (size_t)w * ((size_t)K * 64 * sizeof(T) + 64 * sizeof(double) +
                      64 * sizeof(T));
}

// Per-wave LDS scratch of a target evaluation: GLMs need beta [K*64], the
// residuals of a chunk [64 V] (narrow designs) and the gradient partials
// [64 NB V]; the other families need [K*64 + 64].
template <typename T, int K, int FAM>
__host__ __device__ constexpr size_t target_scratch_elems() {
  return FAM == kFamilyGlm ? (size_t)K * 64 + 64 * Vec16<T>::n +
                                 64 * kGlmBlocks<T, K> * Vec16<T>::n
                           : (size_t)K * 64 + 64;
}

// Bit-mask blends: v[s] for a wave-uniform slot index s, and
// v[k] = (k == s) ? f : v[k], written with integer masks so that the
// compiler keeps them in registers (a select chain on an index is turned
// into a private-memory array access).
template <typename T>
struct Bits;
template <>
struct Bits<float> {
  typedef uint32_t type;
  static __device__ uint32_t of(float v) { return __float_as_uint(v); }
  static __device__ float to(uint32_t b) { return __uint_as_float(b); }
};
template <>
struct Bits<double> {
  typedef uint64_t type;
  static __device__ uint64_t of(double v) {
    return (uint64_t)__double_as_longlong(v);
  }
  static __device__ double to(uint64_t b) {
    return __longlong_as_double((long long)b);
  }
};
template <typename T, int K>
__device__ inline T pick(const T (&v)[K], int s) {
  using U = typename Bits<T>::type;
  U r = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) r |= Bits<T>::of(v[k]) & (s == k ? ~U(0) : U(0));
  return Bits<T>::to(r);
}
// lane 0 ? a : (lane 1 ? b : v), blended with integer masks so that v, a
// and b are all computed in every lane (no exec-mask region).
template <typename T>
__device__ inline T lanes01(T v, T a, T b, int lane) {
  using U = typename Bits<T>::type;
  const U m0 = lane == 0 ? ~U(0) : U(0);
  const U m1 = lane == 1 ? ~U(0) : U(0);
  return Bits<T>::to((Bits<T>::of(v) & ~(m0 | m1)) | (Bits<T>::of(a) & m0) |
                     (Bits<T>::of(b) & m1));
}
template <typename T, int K>
__device__ inline void put_slot(T (&v)[K], int s, T f) {
  using U = typename Bits<T>::type;
  const U fb = Bits<T>::of(f);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const U m = s == k ? ~U(0) : U(0);
    v[k] = Bits<T>::to((Bits<T>::of(v[k]) & ~m) | (fb & m));
  }
}

// Bijector code paths of Target::eval_partial.
constexpr int kBijAny = 0;            // chosen at run time from the layout
constexpr int kBijSlot0Softplus = 1;  // every bijector lane: Softplus, slot 0

// Shared LDS of a chain group (G chains of one GLM problem whose owner
// waves take turns with all W waves of the block streaming X once for every
// chain): the chains' coefficients beta [G][K*64] and noise scales [G], the
// owners' done flags [G], and per wave and chain the partial gradient
// [W][G][K*64], log-likelihood [W][G] (float64) and residual sum of squares
// [W][G].
template <typename T, int K>
struct GroupLayout {
  static __host__ __device__ constexpr size_t al(size_t n) {
    return (n + 15) & ~(size_t)15;
  }
  static __host__ __device__ constexpr size_t beta(int) { return 0; }
  static __host__ __device__ constexpr size_t sig(int g) {
    return al((size_t)g * K * 64 * sizeof(T));
  }
  static __host__ __device__ constexpr size_t done(int g) {
    return sig(g) + al((size_t)g * sizeof(T));
  }
  static __host__ __device__ constexpr size_t red(int g) {
    return done(g) + al((size_t)g * sizeof(int));
  }
  static __host__ __device__ constexpr size_t lp(int g, int w) {
    return red(g) + al((size_t)w * g * K * 64 * sizeof(T));
  }
  static __host__ __device__ constexpr size_t ssq(int g, int w) {
    return lp(g, w) + al((size_t)w * g * sizeof(double));
  }
  static __host__ __device__ constexpr size_t bytes(int g, int w) {
    return ssq(g, w) + al((size_t)w * g * sizeof(T));
  }
};

// G > 1: the GLM target of a chain group (nuts_group_kernel); every
// evaluation is a round in which all waves of the block pass over X for all
// G chains.
template <typename T, int K, int FAM, int G = 1>
struct Target {
  int kind, dim, n_obs;
  T p0, p1;
  double log_norm;
  gptr<T> x;   // [n_obs, dim_x] row-major (GLM)
  gptr<T> xt;  // [dim_x, n_obs] (GLM)
  gptr<T> y;
  int x_pitch, xt_pitch;  // row pitches of X and X^T (elements)
  // per-lane cached data in the state layout
  T a[K], b[K];
  // bijector code of each register slot of this lane; the masks are
  // wave-uniform: bit k set if some lane of slot k uses that bijector
  int code[K];
  uint32_t sp_slots, ex_slots;
  // when every bijector lane sits in one register slot (the common case:
  // a few constrained scalars), that slot and this lane's code in it
  bool one_slot;
  int bslot, bcode;
  bool sp0;  // the only bijector is Softplus, all of it in slot 0
  T inv_s0sq, inv_s0, inv_s1, s1sq;  // prior constants (hier)
  lptr<T> lds;  // per-wave scratch (target_scratch_elems)
  // GLMs: this wave's index among the waves sharing the chain, their count,
  // and the LDS buffer where their partial sums meet (two halves, used in
  // turn: a wave can only reuse a half after every wave has read it)
  int split_w = 0, split_n = 1, red_half = 0;
  lptr<unsigned char> red = nullptr;
  // chain group: the chain this wave owns (-1: none), this wave and the
  // block's waves, and the group's shared LDS (GroupLayout)
  int grp_j = -1, grp_w = 0, grp_waves = 1;
  lptr<unsigned char> grp = nullptr;

  __device__ void load(const hipmc_problem& pb, T* wave_lds) {
    const int lane = lane_id();
    kind = pb.kind;
    dim = pb.dim;
    n_obs = pb.n_obs;
    p0 = (T)pb.param[0];
    p1 = (T)pb.param[1];
    log_norm = pb.log_norm;
    x = (gptr<T>)pb.x;
    xt = (gptr<T>)pb.xt;
    y = (gptr<T>)pb.y;
    x_pitch = pb.x_pitch;
    xt_pitch = pb.xt_pitch;
    const T* aux = (const T*)pb.aux;
    sp_slots = 0u;
    ex_slots = 0u;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      int c = HIPMC_BIJ_IDENTITY;
#pragma unroll
      for (int i = 0; i < HIPMC_MAX_BIJ_RANGES; ++i)
        if (i < pb.n_bij && e >= pb.bij_off[i] &&
            e < pb.bij_off[i] + pb.bij_len[i])
          c = pb.bij_code[i];
      code[k] = c;
      if (__ballot(c == HIPMC_BIJ_SOFTPLUS)) sp_slots |= 1u << k;
      if (__ballot(c == HIPMC_BIJ_EXP)) ex_slots |= 1u << k;
    }
    const uint32_t any = sp_slots | ex_slots;
    one_slot = any != 0u && (any & (any - 1u)) == 0u;
    sp0 = sp_slots == 1u && ex_slots == 0u;
    bslot = any ? __builtin_ctz(any) : 0;
    bcode = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) bcode |= code[k] & (bslot == k ? ~0 : 0);
    inv_s0 = T(1) / p0;
    inv_s0sq = inv_s0 * inv_s0;
    inv_s1 = T(1) / p1;
    s1sq = p1 * p1;
    lds = (lptr<T>)wave_lds;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      a[k] = T(0);
      b[k] = T(0);
      if (kind == HIPMC_TARGET_HIER_NORMAL) {
        if (e >= 2 && e < dim) {
          a[k] = y[e - 2];
          b[k] = T(1) / aux[e - 2];
        }
      } else if (kind == HIPMC_TARGET_DIAG_NORMAL) {
        if (e < dim) {
          a[k] = y[e];
          b[k] = T(1) / aux[e];
        }
      }
    }
  }

  __device__ void set_split(int w, int n, unsigned char* buf) {
    split_w = w;
    split_n = n;
    red = (lptr<unsigned char>)buf;
  }

  __device__ void set_group(int j, int w, int waves, unsigned char* buf) {
    grp_j = j;
    grp_w = w;
    grp_waves = waves;
    grp = (lptr<unsigned char>)buf;
  }

  // Per-lane partial log-prob (double) and wave-uniform part `lpu`; writes
  // the gradient.  The caller reduces the partial (possibly together with
  // other sums).  Bijector log-det-Jacobians are folded into the partial.
  // BM = kBijSlot0Softplus compiles only the path for bijectors that are all
  // Softplus in register slot 0 (the caller has checked sp0); kBijAny
  // chooses the path at run time.
  template <int BM = kBijAny>
  __device__ double eval_partial(const T (&xs)[K], T (&g)[K], double& lpu) {
    T th[K];
    SoftplusParts<T> sp[K];
    T ex[K];
#pragma unroll
    for (int k = 0; k < K; ++k) th[k] = xs[k];
    // one bijector slot: select it, transform it, put it back (no per-slot
    // branches); otherwise the per-slot loop
    SoftplusParts<T> spb{};
    T exb = T(0);
    if (BM == kBijSlot0Softplus || sp0) {
      spb = softplus_parts(xs[0]);
      th[0] = code[0] == HIPMC_BIJ_SOFTPLUS ? fmax(xs[0], T(0)) + spb.l
                                            : xs[0];
    } else if (BM == kBijSlot0Softplus) {
    } else if (one_slot) {
      const T xb = pick(xs, bslot);
      if (sp_slots) spb = softplus_parts(xb);
      if (ex_slots) {
        keep_branch();
        exb = exp(xb);
      }
      // Softplus forward (bijectors/softplus.py:36-50): below the cutoff TFP
      // takes log1p(exp(x)), which equals max(x, 0) + l there.
      const T f = bcode == HIPMC_BIJ_SOFTPLUS ? fmax(xb, T(0)) + spb.l
                  : (bcode == HIPMC_BIJ_EXP ? exb : xb);
      put_slot(th, bslot, f);
    } else if (sp_slots | ex_slots) {
      keep_branch();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (sp_slots & (1u << k)) {
          sp[k] = softplus_parts(xs[k]);
          const T f = fmax(xs[k], T(0)) + sp[k].l;
          th[k] = code[k] == HIPMC_BIJ_SOFTPLUS ? f : th[k];
        }
        if (ex_slots & (1u << k)) {
          ex[k] = exp(xs[k]);
          th[k] = code[k] == HIPMC_BIJ_EXP ? ex[k] : th[k];
        }
      }
    }
    double lpl;
    if (FAM == kFamilyHier) {
      lpl = hier_normal(th, g, lpu);
    } else if (FAM == kFamilyGlm) {
      lpl = kind == HIPMC_TARGET_LINEAR ? glm<true>(th, g, lpu)
                                        : glm<false>(th, g, lpu);
    } else {
      switch (kind) {
        case HIPMC_TARGET_DIAG_NORMAL: lpl = diag_normal(th, g, lpu); break;
        case HIPMC_TARGET_FUNNEL: lpl = funnel(th, g, lpu); break;
        case HIPMC_TARGET_LOG_GAMMA: lpl = log_gamma(th, g, lpu); break;
        default: lpl = nan_exp(th, g, lpu); break;
      }
    }
    // fldj: Softplus -softplus(-x) with d/dx f = sigmoid(x),
    // d/dx fldj = sigmoid(-x); Exp x with d/dx f = exp(x), d/dx fldj = 1
    if (BM == kBijSlot0Softplus || sp0) {
      const bool on = code[0] == HIPMC_BIJ_SOFTPLUS;
      lpl += on ? (double)(-(fmax(-xs[0], T(0)) + spb.l)) : 0.0;
      g[0] = on ? g[0] * spb.sig + spb.sig_neg : g[0];
    } else if (BM == kBijSlot0Softplus) {
    } else if (one_slot) {
      const T xb = pick(xs, bslot);
      const T gb = pick(g, bslot);
      const bool on_sp = bcode == HIPMC_BIJ_SOFTPLUS;
      const bool on_ex = bcode == HIPMC_BIJ_EXP;
      lpl += on_sp ? (double)(-(fmax(-xb, T(0)) + spb.l))
                   : (on_ex ? (double)xb : 0.0);
      const T gn = on_sp ? gb * spb.sig + spb.sig_neg
                         : (on_ex ? gb * exb + T(1) : gb);
      put_slot(g, bslot, gn);
    } else if (sp_slots | ex_slots) {
      keep_branch();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T xv = xs[k];
        if (sp_slots & (1u << k)) {
          const bool on = code[k] == HIPMC_BIJ_SOFTPLUS;
          lpl += on ? (double)(-(fmax(-xv, T(0)) + sp[k].l)) : 0.0;
          g[k] = on ? g[k] * sp[k].sig + sp[k].sig_neg : g[k];
        }
        if (ex_slots & (1u << k)) {
          const bool on = code[k] == HIPMC_BIJ_EXP;
          lpl += on ? (double)xv : 0.0;
          g[k] = on ? g[k] * ex[k] + T(1) : g[k];
        }
      }
    }
    return lpl;
  }

  // Wave-uniform target log-prob (double); writes the gradient.
  template <int BM = kBijAny>
  __device__ double eval(const T (&xs)[K], T (&g)[K]) {
    double lpu;
    const double lpl = eval_partial<BM>(xs, g, lpu);
    return wave_sum(lpl) + lpu;
  }

  // [mu, tau, z_0..z_{J-1}], y ~ N(mu + tau z, sigma); mu ~ N(0, p0),
  // tau ~ HalfCauchy(0, p1), z ~ N(0, 1).
  __device__ double hier_normal(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    const T mu = readlane(th[0], 0);
    const T tau = readlane(th[0], 1);
    double lpl = 0.0;
    T s1 = T(0), s2 = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool is_z = (e >= 2) && (e < dim);
      const T z = th[k];
      const T t = mu + tau * z;
      const T u = (a[k] - t) * b[k];  // (y - theta) / sigma
      const T r = u * b[k];           // (y - theta) / sigma^2
      lpl += is_z ? (double)(T(-0.5) * z * z - T(0.5) * u * u) : 0.0;
      g[k] = is_z ? (-z + tau * r) : T(0);
      s1 += is_z ? r : T(0);
      s2 += is_z ? r * z : T(0);
    }
    T S1, S2;
    wave_sum2(s1, s2, S1, S2);
    // wave-uniform scalars: computed once, then blended into lanes 0 and 1
    const T gmu = uniform(-mu * inv_s0sq + S1);
    const T gtau = uniform(T(-2) * tau * recip(s1sq + tau * tau) + S2);
    g[0] = lanes01(g[0], gmu, gtau, lane);
    const T zm = mu * inv_s0;
    const T zt = tau * inv_s1;
    lpu = log_norm + (double)(T(-0.5) * zm * zm) -
          (double)log1p_pos(uniform(zt * zt));
    if (tau < T(0)) lpu = neg_inf();
    return lpl;
  }

  __device__ double diag_normal(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    double lpl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool v = e < dim;
      const T u = (th[k] - a[k]) * b[k];
      lpl += v ? (double)(T(-0.5) * u * u) : 0.0;
      g[k] = v ? -u * b[k] : T(0);
    }
    lpu = log_norm;
    return lpl;
  }

  // [v, x_1..x_{dim-1}]: v ~ N(0, 3), x_i ~ N(0, exp(v/2)).
  __device__ double funnel(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    const T v = readlane(th[0], 0);
    const T inv_s2 = exp(-v);
    T sx2 = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool isx = (e >= 1) && (e < dim);
      sx2 += isx ? th[k] * th[k] : T(0);
      g[k] = isx ? -th[k] * inv_s2 : T(0);
    }
    const T S = wave_sum(sx2);
    const T nk = (T)(dim - 1);
    if (lane == 0) g[0] = -v / T(9) + T(0.5) * (S * inv_s2 - nk);
    lpu = log_norm + (double)(T(-0.5) * (v / T(3)) * (v / T(3))) -
          (double)(T(0.5) * nk * v) - (double)(T(0.5) * S * inv_s2);
    return 0.0;
  }

  // x_i = log of a Gamma(concentration p0, rate p1) variable:
  // log p = p0 x - p1 exp(x) (+ constants in log_norm), gradient
  // p0 - p1 exp(x) (the log-gamma target of mcmc/hmc_test.py:146-200).
  __device__ double log_gamma(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    double lpl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool v = e < dim;
      const T ex = exp(th[k]);
      lpl += v ? (double)(p0 * th[k] - p1 * ex) : 0.0;
      g[k] = v ? p0 - p1 * ex : T(0);
    }
    lpu = log_norm;
    return lpl;
  }

  __device__ double nan_exp(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    double lpl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool v = e < dim;
      const T xv = th[k];
      const double term = xv < T(0) ? __builtin_nan("") : (double)(-xv);
      lpl += v ? term : 0.0;
      g[k] = (v && !(xv < T(0))) ? T(-1) : T(0);
    }
    lpu = 0.0;
    return lpl;
  }

  // The row loop of a narrow GLM (one coefficient block, X^T given): two
  // passes per chunk.  This wave handles chunks of 64V rows
  // n0 = split_w * 64V, advancing by split_n * 64V.
  //  s-pass: lane owns rows n0 + V*lane + [0, V) and reads one 16-byte
  //   vector of X^T per coefficient, 8 coefficients per batch; the next
  //   batch is issued before the current one is used, and coefficients past
  //   dx read the last column and meet beta = 0 (exact zero terms), so the
  //   sums equal the plain even/odd split.
  //  gradient pass: lane (phase ph = lane / 16, group grp = lane % 16) adds
  //   rows ph, ph + 4, ... of the chunk, coefficients b*16V + V*grp + [0, V)
  //   of block b (one load instruction covers 4 rows x 16V coefficients);
  //   B rows per phase per batch, the next batch in flight while the current
  //   one is used, and the first batch issued before the residuals are
  //   computed.  Rows past the chunk are clamped and weighted by zero;
  //   blocks past the row pitch are clamped to the row start (their
  //   accumulators are never read).
  template <int NBB, bool LINEAR>
  __device__ void glm_rows(lptr<const T> beta, lptr<T> rbuf, int dx, int xp,
                           int tp,
                           T inv_s2, const int (&cst)[kGlmBlocks<T, K>],
                           typename Vec16<T>::type (&acc)[kGlmBlocks<T, K>],
                           double& lpl, T& ssq) {
    using VT = typename Vec16<T>::type;
    constexpr int V = Vec16<T>::n;
    constexpr int SB = 8;
    constexpr int B = NBB == 1 ? 8 : 4;
    const int lane = lane_id();
    const int ph = lane >> 4;
    const int n = n_obs;
    const int nsb = (dx + SB - 1) / SB;
    const int dlast = dx - 1;
    for (int n0 = split_w * 64 * V; n0 < n; n0 += split_n * 64 * V) {
      const int rr = n0 + V * lane;
      const gptr<T> col = xt + (rr < tp ? rr : tp - V);
      T yv[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int row = rr + v;
        yv[v] = y[row < n ? row : n - 1];
      }
      VT cur[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u)
        cur[u] = *(gptr<VT>)(col + (size_t)(u < dlast ? u : dlast) * tp);
      VT s0 = VT{}, s1 = VT{};
      for (int bi = 0; bi + 1 < nsb; ++bi) {
        const int d = bi * SB;
        VT nxt[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int dd = d + SB + u;
          nxt[u] = *(gptr<VT>)(col + (size_t)(dd < dlast ? dd : dlast) * tp);
        }
#pragma unroll
        for (int u = 0; u < SB; u += 2) {
          s0 += cur[u] * beta[d + u];
          s1 += cur[u + 1] * beta[d + u + 1];
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) cur[u] = nxt[u];
      }
      {
        const int d = (nsb - 1) * SB;
#pragma unroll
        for (int u = 0; u < SB; u += 2) {
          s0 += cur[u] * beta[d + u];
          s1 += cur[u + 1] * beta[d + u + 1];
        }
      }
      const VT sv = s0 + s1;
      const int rows = (n - n0) < 64 * V ? (n - n0) : 64 * V;
      const gptr<T> xb = x + (size_t)n0 * xp;
      VT gv[B][NBB];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int jj = 4 * u + ph;
        const int jl = jj < rows ? jj : rows - 1;
#pragma unroll
        for (int b = 0; b < NBB; ++b)
          gv[u][b] = *(gptr<VT>)(xb + (size_t)jl * xp + cst[b]);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const bool valid = rr + v < n;
        const T s = sv[v];
        const T yy = valid ? yv[v] : T(0);
        T r;
        if (LINEAR) {
          const T res = yy - s;
          r = valid ? res * inv_s2 : T(0);
          lpl += valid ? (double)(T(-0.5) * res * res * inv_s2) : 0.0;
          ssq += valid ? res * res : T(0);
        } else {
          const T w0 = T(1) - yy;
          const SoftplusParts<T> sp = softplus_parts(s);
          const T l0 = w0 == T(0) ? T(0) : -(fmax(s, T(0)) + sp.l) * w0;
          const T l1 = yy == T(0) ? T(0) : -(fmax(-s, T(0)) + sp.l) * yy;
          lpl += valid ? (double)(l0 + l1) : 0.0;
          r = valid ? yy - sp.sig : T(0);
        }
        rbuf[V * lane + v] = r;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      int j0 = 0;
      for (; j0 + 4 * B < rows; j0 += 4 * B) {
        VT nv[B][NBB];
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const int jj = j0 + 4 * B + 4 * u + ph;
          const int jl = jj < rows ? jj : rows - 1;
#pragma unroll
          for (int b = 0; b < NBB; ++b)
            nv[u][b] = *(gptr<VT>)(xb + (size_t)jl * xp + cst[b]);
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const int jj = j0 + 4 * u + ph;
          const T rj = jj < rows ? rbuf[jj] : T(0);
#pragma unroll
          for (int b = 0; b < NBB; ++b) acc[b] += gv[u][b] * rj;
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
          for (int b = 0; b < NBB; ++b) gv[u][b] = nv[u][b];
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int jj = j0 + 4 * u + ph;
        const T rj = jj < rows ? rbuf[jj] : T(0);
#pragma unroll
        for (int b = 0; b < NBB; ++b) acc[b] += gv[u][b] * rj;
      }
      // the next chunk overwrites rbuf only after every lane has read it
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }

  // The row loop of a wider GLM evaluation: one pass over X.  This wave takes
  // chunks of 64V rows n0 = split_w * 64V, advancing by split_n * 64V, in
  // batches of 4B rows.  Lane (phase ph = lane / 16, group grp = lane % 16)
  // holds rows r0 + 4u + ph (u < B) of a batch, coefficients
  // b*16V + V*grp + [0, V) of block b, as one 16-byte vector per row and
  // block (one load instruction covers 4 rows x 16V coefficients); the next
  // batch is loaded while the current one is used.
  //  logits: the lane's partial dot products with beta, summed over the 16
  //   lanes of its DPP row (each of them gets the total);
  //  residuals: lane evaluates rows u = (lane & 3) + 4k (k < B/4; for B = 2,
  //   u = lane & 1) and the quad broadcasts them to its other lanes; only
  //   lanes grp < min(B, 4) add the rows' log-likelihood terms;
  //  gradient: acc[b] += x * r over the lane's coefficients.
  // Rows past n read row n - 1 and get residual zero; coefficients past dx
  // meet beta = 0; blocks past the row pitch are clamped to the row start
  // (their accumulators are never read).
  template <int NBB, bool LINEAR>
  __device__ void glm_pass(lptr<const T> beta, int dx, int xp, T inv_s2,
                           const int (&cst)[kGlmBlocks<T, K>],
                           typename Vec16<T>::type (&acc)[kGlmBlocks<T, K>],
                           double& lpl, T& ssq) {
    using VT = typename Vec16<T>::type;
    constexpr int V = Vec16<T>::n;
    constexpr int CPB = 16 * V;
    constexpr int B = NBB == 1 ? 8 : (NBB == 2 ? 4 : 2);
    constexpr int RPL = B > 4 ? B / 4 : 1;  // rows evaluated per lane
    constexpr int RB = 4 * B;               // rows per batch
    constexpr int CH = 64 * V;              // rows per chunk
    constexpr int NQ = CH / RB;             // batches per chunk
    const int lane = lane_id();
    const int ph = lane >> 4, grp = lane & 15;
    const int n = n_obs;
    const bool counts = grp < (B < 4 ? B : 4);
    VT bs[NBB];
#pragma unroll
    for (int b = 0; b < NBB; ++b)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int c = b * CPB + V * grp + v;
        bs[b][v] = c < dx ? beta[c] : T(0);
      }
    const int q0 = split_w * CH, qs = split_n * CH;
    auto row0 = [&](int q) { return q0 + (q / NQ) * qs + (q % NQ) * RB; };
    auto own = [&](int k) { return B >= 4 ? (lane & 3) + 4 * k : (lane & 1); };
    auto load = [&](int r0, VT (&xv)[B][NBB], T (&yv)[RPL]) {
      if (r0 + RB <= n) {
        // the whole batch lies inside X: rows 4 apart from one base
        const gptr<T> xr = x + (size_t)(r0 + ph) * xp;
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
          for (int b = 0; b < NBB; ++b)
            xv[u][b] = *(gptr<VT>)(xr + (size_t)(4 * u) * xp + cst[b]);
#pragma unroll
        for (int k = 0; k < RPL; ++k) yv[k] = y[r0 + 4 * own(k) + ph];
        return;
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int j = r0 + 4 * u + ph;
        const int jl = j < n ? j : n - 1;
#pragma unroll
        for (int b = 0; b < NBB; ++b)
          xv[u][b] = *(gptr<VT>)(x + (size_t)jl * xp + cst[b]);
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int j = r0 + 4 * own(k) + ph;
        yv[k] = y[j < n ? j : n - 1];
      }
    };
    // one batch: logits, residuals and log-likelihood terms, gradient
    // one batch; `next` issues the loads of the following batch.  Every
    // register of this batch is waited for (the empty asm statements use
    // them) before those loads are issued: across the loop's back edge the
    // wait-count pass does not tell the two buffers' loads apart and would
    // otherwise wait for the newer batch's loads as well, exposing their
    // whole latency every second batch.
    auto process = [&](int r0, const VT (&xv)[B][NBB], const T (&yv)[RPL],
                       auto&& next) {
#pragma unroll
      for (int u = 0; u < B; ++u)
#pragma unroll
        for (int b = 0; b < NBB; ++b) asm volatile("" ::"v"(xv[u][b]));
#pragma unroll
      for (int k = 0; k < RPL; ++k) asm volatile("" ::"v"(yv[k]));
      next();
      T pa[B];  // the lane's partial dot product of each row
#pragma unroll
      for (int u = 0; u < B; ++u) {
        T a = T(0);
#pragma unroll
        for (int b = 0; b < NBB; ++b)
#pragma unroll
          for (int v = 0; v < V; ++v) a += xv[u][b][v] * bs[b][v];
        pa[u] = a;
      }
      T rr[RPL];
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        // the logit of the row this lane evaluates (row 4k + (lane & 3),
        // or lane & 1 for two rows per phase)
        T s;
        if constexpr (B >= 4) {
          s = row_sum_pick4(pa[4 * k], pa[4 * k + 1], pa[4 * k + 2],
                            pa[4 * k + 3], (lane & 1) != 0, (lane & 2) != 0);
        } else {
          T sq[B];
#pragma unroll
          for (int i = 0; i < B; ++i) sq[i] = row_sum(pa[i]);
          s = pick(sq, lane & 1);
        }
        const bool valid = r0 + 4 * own(k) + ph < n;
        const bool add = valid && counts;
        T r;
        if (LINEAR) {
          const T res = yv[k] - s;
          r = res * inv_s2;
          lpl += add ? (double)(T(-0.5) * res * res * inv_s2) : 0.0;
          ssq += add ? res * res : T(0);
        } else {
          const T yy = yv[k];
          const T w0 = T(1) - yy;
          const SoftplusParts<T> sp = softplus_parts(s);
          const T l0 = w0 == T(0) ? T(0) : -(fmax(s, T(0)) + sp.l) * w0;
          const T l1 = yy == T(0) ? T(0) : -(fmax(-s, T(0)) + sp.l) * yy;
          lpl += add ? (double)(l0 + l1) : 0.0;
          r = yy - sp.sig;
        }
        rr[k] = valid ? r : T(0);
      }
      T rj[B];
#pragma unroll
      for (int u = 0; u < B; ++u) rj[u] = quad_bcast(rr[u >> 2], u & 3);
#pragma unroll
      for (int u = 0; u < B; ++u)
#pragma unroll
        for (int b = 0; b < NBB; ++b) acc[b] += xv[u][b] * rj[u];
    };
    int r0 = row0(0);
    if (r0 >= n) return;
    // two register buffers used in turn (the next batch loads into one while
    // the other is processed), unrolled by two so that no copies are needed
    VT xa[B][NBB], xb[B][NBB];
    T ya[RPL], yb[RPL];
    load(r0, xa, ya);
    for (int q = 1;; q += 2) {
      const int r1 = row0(q);
      const bool m1 = r1 < n;
      process(r0, xa, ya, [&] {
        if (m1) load(r1, xb, yb);
      });
      if (!m1) break;
      const int r2 = row0(q + 1);
      const bool m2 = r2 < n;
      process(r1, xb, yb, [&] {
        if (m2) load(r2, xa, ya);
      });
      if (!m2) break;
      r0 = r2;
    }
  }

  // Bayesian GLM.  LINEAR=false: logistic regression over beta (dim = D_x).
  // LINEAR=true: [beta, sigma] with Gaussian likelihood (dim = D_x + 1).
  // The logits X beta and the gradient X^T r come from one pass over the
  // rows of X (glm_pass), or for a design of one coefficient block whose
  // X^T is given, from a pass over X^T and one over X (glm_rows); the four
  // row phases and, for a shared chain, the waves' partial sums are then
  // added in a fixed order.
  template <bool LINEAR>
  __device__ double glm(const T (&th)[K], T (&g)[K], double& lpu) {
    if constexpr (G > 1) {
      return glm_grouped<LINEAR>(th, g, lpu);
    }
    using VT = typename Vec16<T>::type;
    constexpr int V = Vec16<T>::n;
    constexpr int CPB = 16 * V;                 // coefficients per block
    constexpr int NB = kGlmBlocks<T, K>;        // blocks over K*64 coefficients
    const int lane = lane_id();
    const int dx = LINEAR ? dim - 1 : dim;
    const int n = n_obs;
    T sigma = T(1), inv_s2 = T(1);
    if (LINEAR) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == (dx >> 6)) sigma = readlane(th[k], dx & 63);
      inv_s2 = T(1) / (sigma * sigma);
    }
    // X rows are 16-byte aligned (pitch a multiple of 4); any other layout,
    // or an empty X, gives NaN rather than a stray read
    if ((x_pitch & 3) != 0 || x_pitch < dx || dx < 1 || n < 1) {
#pragma unroll
      for (int k = 0; k < K; ++k) g[k] = T(NAN);
      lpu = (double)NAN;
      return 0.0;
    }
    const int xp = x_pitch;
    lptr<T> beta = lds;            // [K*64]
    lptr<T> rbuf = lds + K * 64;   // [64V] residuals of a chunk (narrow)
    lptr<T> part = rbuf + 64 * V;  // [64 lanes][NB][V] gradient partials
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      beta[e] = (e < dx) ? th[k] : T(0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int grp = lane & 15;
    int cst[NB];
    VT acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int c = b * CPB + V * grp;
      cst[b] = c < xp ? c : 0;
      acc[b] = VT{};
    }
    const int nbl = (dx + CPB - 1) / CPB;  // live blocks
    double lpl = 0.0;
    T ssq = T(0);
    // narrow designs with X^T: fewer instructions per row in two passes
    const bool narrow = nbl <= 1 && xt != nullptr && (xt_pitch & 3) == 0 &&
                        xt_pitch >= n;
    if (narrow) {
      glm_rows<1, LINEAR>(beta, rbuf, dx, xp, xt_pitch, inv_s2, cst, acc, lpl,
                          ssq);
    } else if (nbl <= 1 || NB == 1) {
      glm_pass<1, LINEAR>(beta, dx, xp, inv_s2, cst, acc, lpl, ssq);
    } else if (nbl == 2 || NB == 2) {
      glm_pass<(NB >= 2 ? 2 : 1), LINEAR>(beta, dx, xp, inv_s2, cst, acc, lpl,
                                          ssq);
    } else if (nbl == 3 || NB == 3) {
      glm_pass<(NB >= 3 ? 3 : 1), LINEAR>(beta, dx, xp, inv_s2, cst, acc, lpl,
                                          ssq);
    } else {
      glm_pass<NB, LINEAR>(beta, dx, xp, inv_s2, cst, acc, lpl, ssq);
    }
    // sum the four row phases of each coefficient in a fixed order
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int v = 0; v < V; ++v) part[(lane * NB + b) * V + v] = acc[b][v];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    T gl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const int b = e / CPB, gi = (e % CPB) / V, vi = e % V;
      T sum = T(0);
#pragma unroll
      for (int p = 0; p < 4; ++p) sum += part[((p * 16 + gi) * NB + b) * V + vi];
      gl[k] = sum;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (split_n > 1) {
      // every wave adds the waves' partials in the same order, so all of
      // them continue with identical sums
      lptr<unsigned char> half =
          red + (size_t)red_half * glm_red_bytes<T, K>(split_n);
      lptr<T> rg = (lptr<T>)half;
      lptr<double> rl = (lptr<double>)(rg + (size_t)split_n * K * 64);
      lptr<T> rs = (lptr<T>)(rl + (size_t)split_n * 64);
#pragma unroll
      for (int k = 0; k < K; ++k) rg[(split_w * K + k) * 64 + lane] = gl[k];
      rl[split_w * 64 + lane] = lpl;
      rs[split_w * 64 + lane] = ssq;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; ++k) gl[k] = T(0);
      lpl = 0.0;
      ssq = T(0);
      for (int w = 0; w < split_n; ++w) {
#pragma unroll
        for (int k = 0; k < K; ++k) gl[k] += rg[(w * K + k) * 64 + lane];
        lpl += rl[w * 64 + lane];
        ssq += rs[w * 64 + lane];
      }
      red_half ^= 1;
    }
    double lp = log_norm;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool isb = e < dx;
      const T bv = th[k] * inv_s0;
      lpl += isb ? (double)(T(-0.5) * bv * bv) : 0.0;
      g[k] = isb ? gl[k] - th[k] * inv_s0sq : T(0);
    }
    if (LINEAR) {
      const T S = wave_sum(ssq);
      const T ss = p1;
      const T gs = S / (sigma * sigma * sigma) - (T)n / sigma - sigma / (ss * ss);
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == (dx >> 6) && lane == (dx & 63)) g[k] = gs;
      lp += -(double)n * (double)log(sigma) +
            (double)(T(-0.5) * (sigma / ss) * (sigma / ss));
      if (sigma < T(0)) lp = neg_inf();
    }
    lpu = lp;
    return lpl;
  }

  // ---------------------------------------------------------- chain groups
  // Block-wide barrier with LDS ordering; every wave of a group block
  // reaches the same number of them (two per round).
  __device__ static void group_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }

  // The pass over this wave's rows for all G chains (as glm_pass, with B = 2
  // rows per phase per batch): logits, residuals and gradient partials of
  // every chain from each loaded row.
  template <int NBB, bool LINEAR>
  __device__ void group_pass(int dx, int xp, const int (&cst)[kGlmBlocks<T, K>],
                             typename Vec16<T>::type (&acc)[G][kGlmBlocks<T, K>],
                             double (&lpl)[G], T (&ssq)[G]) {
    using VT = typename Vec16<T>::type;
    constexpr int V = Vec16<T>::n;
    constexpr int CPB = 16 * V;
    constexpr int B = 2;
    constexpr int RB = 4 * B;
    constexpr int CH = 64 * V;
    constexpr int NQ = CH / RB;
    const int lane = lane_id();
    const int ph = lane >> 4, grp16 = lane & 15;
    const int n = n_obs;
    const bool counts = grp16 < B;
    lptr<const T> beta = (lptr<const T>)(grp + GroupLayout<T, K>::beta(G));
    lptr<const T> sg = (lptr<const T>)(grp + GroupLayout<T, K>::sig(G));
    VT bs[G][NBB];
    T inv_s2[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int b = 0; b < NBB; ++b)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int c = b * CPB + V * grp16 + v;
          bs[j][b][v] = c < dx ? beta[j * K * 64 + c] : T(0);
        }
      const T s = LINEAR ? sg[j] : T(1);
      inv_s2[j] = T(1) / (s * s);
    }
    const int q0 = grp_w * CH, qs = grp_waves * CH;
    auto row0 = [&](int q) { return q0 + (q / NQ) * qs + (q % NQ) * RB; };
    auto load = [&](int r0, VT (&xv)[B][NBB], T& yv) {
      if (r0 + RB <= n) {
        // the whole batch lies inside X: rows 4 apart from one base
        const gptr<T> xr = x + (size_t)(r0 + ph) * xp;
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
          for (int b = 0; b < NBB; ++b)
            xv[u][b] = *(gptr<VT>)(xr + (size_t)(4 * u) * xp + cst[b]);
        yv = y[r0 + 4 * (lane & 1) + ph];
        return;
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int j = r0 + 4 * u + ph;
        const int jl = j < n ? j : n - 1;
#pragma unroll
        for (int b = 0; b < NBB; ++b)
          xv[u][b] = *(gptr<VT>)(x + (size_t)jl * xp + cst[b]);
      }
      const int jr = r0 + 4 * (lane & 1) + ph;
      yv = y[jr < n ? jr : n - 1];
    };
    // one batch for every chain of the group
    auto process = [&](int r0, const VT (&xv)[B][NBB], T yv) {
      const bool valid = r0 + 4 * (lane & 1) + ph < n;
      const bool add = valid && counts;
#pragma unroll
      for (int j = 0; j < G; ++j) {
        T sr[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
          T a = T(0);
#pragma unroll
          for (int b = 0; b < NBB; ++b)
#pragma unroll
            for (int v = 0; v < V; ++v) a += xv[u][b][v] * bs[j][b][v];
          sr[u] = row_sum(a);
        }
        const T s = (lane & 1) ? sr[1] : sr[0];
        T r;
        if (LINEAR) {
          const T res = yv - s;
          r = res * inv_s2[j];
          lpl[j] += add ? (double)(T(-0.5) * res * res * inv_s2[j]) : 0.0;
          ssq[j] += add ? res * res : T(0);
        } else {
          const T w0 = T(1) - yv;
          const SoftplusParts<T> sp = softplus_parts(s);
          const T l0 = w0 == T(0) ? T(0) : -(fmax(s, T(0)) + sp.l) * w0;
          const T l1 = yv == T(0) ? T(0) : -(fmax(-s, T(0)) + sp.l) * yv;
          lpl[j] += add ? (double)(l0 + l1) : 0.0;
          r = yv - sp.sig;
        }
        r = valid ? r : T(0);
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const T rj = quad_bcast(r, u);
#pragma unroll
          for (int b = 0; b < NBB; ++b) acc[j][b] += xv[u][b] * rj;
        }
      }
    };
    int r0 = row0(0);
    if (r0 >= n) return;
    // two register buffers used in turn, unrolled by two (no copies)
    VT xa[B][NBB], xb[B][NBB];
    T ya, yb;
    load(r0, xa, ya);
    for (int q = 1;; q += 2) {
      const int r1 = row0(q);
      const bool m1 = r1 < n;
      if (m1) load(r1, xb, yb);
      process(r0, xa, ya);
      if (!m1) break;
      const int r2 = row0(q + 1);
      const bool m2 = r2 < n;
      if (m2) load(r2, xa, ya);
      process(r1, xb, yb);
      if (!m2) break;
      r0 = r2;
    }
  }

  // Float32 groups of 4 chains with at most kMfmaSlabs * 16 coefficients
  // take their passes on matrix instructions (group_mfma).
  static constexpr bool kMfmaGroup = sizeof(T) == 4 && G == 4;
  static constexpr int kMfmaSlabs = 4 * K < 8 ? 4 * K : 8;

  // The pass of a float32 group of 4 chains on v_mfma_f32_4x4x1_16b_f32
  // (block b: lanes 4b..4b+3; register m of lane 4b + n holds row m, column
  // n of the block's tile).  Lane L = 16 rg + 4 kp + q; the wave takes
  // sub-chunks of 16 rows r0 = 16 (grp_w + i grp_waves); NS = live slabs of
  // 16 coefficients.
  //  loads: the lane holds row 4 rg + (q + kp) % 4, coefficient quads
  //   4s + kp (s < NS), one 16-byte load each (64 contiguous bytes of a row
  //   over the four lanes with equal rg and row);
  //  logits: MFMA (s, t) with A = x (rows), B = beta of chain n at
  //   coefficient 16s + 4kp + t: block (rg, kp) sums the quads of phase kp
  //   of rows 4 rg + (m + kp) % 4; lane (rg, kp, q) adds the four phases of
  //   row 4 rg + kp, chain q (phase kp + d sits in register (4 - d) % 4 of
  //   the lane 4d further on, a DPP rotation of the 16-lane row);
  //  residual: row 4 rg + kp, chain q, evaluated once per (row, chain);
  //  gradient: each quad's 4x4 tile is transposed across the block, so that
  //   element t of lane m' is coefficient 16s + 4kp + m' of row
  //   4 rg + (t + kp) % 4, and MFMA (s, t) multiplies it by that row's
  //   residual (the residual of the lane 4t further on): register m' of
  //   acc[s] in lane (rg, kp, q) += x[row, 16s + 4kp + m'] r[row, q].
  // Rows past n read row n - 1 and get residual zero; quads past the row
  // pitch read quad 0 and meet beta = 0 (their gradient is never read).
  // At the end the four row groups rg are added (lanes 16 apart) and this
  // wave's partial sums go to its slots, as in group_round.
  template <int NS, bool LINEAR>
  __device__ void group_mfma(int dx, int xp) {
    static_assert(sizeof(T) == 4 && G == 4, "float32 groups of 4 chains");
    static_assert(NS >= 1 && 16 * NS <= K * 64, "slabs within the state");
    const int lane = lane_id();
    const int q = lane & 3, kp = (lane >> 2) & 3, rg = lane >> 4;
    const bool b0 = (q & 1) != 0, b1 = (q & 2) != 0;
    const int n = n_obs;
    lptr<const T> beta = (lptr<const T>)(grp + GroupLayout<T, K>::beta(G));
    lptr<const T> sg = (lptr<const T>)(grp + GroupLayout<T, K>::sig(G));
    f32x4 bq[NS];
    int cs[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int c = 16 * s + 4 * kp;
      bq[s] = *(lptr<const f32x4>)(beta + q * K * 64 + c);
      cs[s] = c < xp ? c : 0;
    }
    T inv_s2 = T(1);
    if (LINEAR) {
      const T sj = sg[q];
      inv_s2 = T(1) / (sj * sj);
    }
    const int ro = 4 * rg + ((q + kp) & 3);  // row of the lane's loads
    const int rr = 4 * rg + kp;              // row of the lane's residual
    double lpl = 0.0;
    T ssq = T(0);
    f32x4 acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto load = [&](int r0, f32x4 (&xv)[NS], T& yv) {
      if (r0 + 16 <= n) {
        const gptr<T> xr = x + (size_t)(r0 + ro) * xp;
#pragma unroll
        for (int s = 0; s < NS; ++s) xv[s] = *(gptr<f32x4>)(xr + cs[s]);
        yv = y[r0 + rr];
        return;
      }
      const int j = r0 + ro;
      const gptr<T> xr = x + (size_t)(j < n ? j : n - 1) * xp;
#pragma unroll
      for (int s = 0; s < NS; ++s) xv[s] = *(gptr<f32x4>)(xr + cs[s]);
      const int jr = r0 + rr;
      yv = y[jr < n ? jr : n - 1];
    };
    auto process = [&](int r0, f32x4 (&xv)[NS], T yv) {
      f32x4 c = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) c = mfma_4x4(xv[s][t], bq[s][t], c);
      // the transposes need only the loaded tiles: they overlap the MFMAs
#pragma unroll
      for (int s = 0; s < NS; ++s) quad_transpose(xv[s], b0, b1);
      T sl = c[0];
      sl = sl + dpp_read<kDppRowRor12>(c[3]);
      sl = sl + dpp_read<kDppRowRor8>(c[2]);
      sl = sl + dpp_read<kDppRowRor4>(c[1]);
      const bool valid = r0 + rr < n;
      T r;
      if (LINEAR) {
        const T res = yv - sl;
        r = res * inv_s2;
        lpl += valid ? (double)(T(-0.5) * res * res * inv_s2) : 0.0;
        ssq += valid ? res * res : T(0);
      } else {
        const T w0 = T(1) - yv;
        const SoftplusParts<T> sp = softplus_parts(sl);
        const T l0 = w0 == T(0) ? T(0) : -(fmax(sl, T(0)) + sp.l) * w0;
        const T l1 = yv == T(0) ? T(0) : -(fmax(-sl, T(0)) + sp.l) * yv;
        lpl += valid ? (double)(l0 + l1) : 0.0;
        r = yv - sp.sig;
      }
      r = valid ? r : T(0);
      const T rt1 = dpp_read<kDppRowRor12>(r);
      const T rt2 = dpp_read<kDppRowRor8>(r);
      const T rt3 = dpp_read<kDppRowRor4>(r);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        acc[s] = mfma_4x4(xv[s][0], r, acc[s]);
        acc[s] = mfma_4x4(xv[s][1], rt1, acc[s]);
        acc[s] = mfma_4x4(xv[s][2], rt2, acc[s]);
        acc[s] = mfma_4x4(xv[s][3], rt3, acc[s]);
      }
    };
    int r0 = 16 * grp_w;
    const int rs = 16 * grp_waves;
    if (r0 < n) {
      // two register buffers used in turn, unrolled by two (no copies)
      f32x4 xa[NS], xb[NS];
      T ya, yb;
      load(r0, xa, ya);
      for (;;) {
        const int r1 = r0 + rs;
        const bool m1 = r1 < n;
        if (m1) load(r1, xb, yb);
        process(r0, xa, ya);
        if (!m1) break;
        const int r2 = r1 + rs;
        const bool m2 = r2 < n;
        if (m2) load(r2, xa, ya);
        process(r1, xb, yb);
        if (!m2) break;
        r0 = r2;
      }
    }
    // add the row groups (lanes 16 and 32 apart) and leave this wave's
    // partial sums in its slots; slabs past NS are zero
    lptr<T> red_g = (lptr<T>)(grp + GroupLayout<T, K>::red(G));
    lptr<double> red_l =
        (lptr<double>)(grp + GroupLayout<T, K>::lp(G, grp_waves));
    lptr<T> red_s = (lptr<T>)(grp + GroupLayout<T, K>::ssq(G, grp_waves));
    lptr<T> dst = red_g + ((size_t)grp_w * G + q) * K * 64;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      f32x4 v = acc[s];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const T s1 = v[e] + __shfl_xor(v[e], 16);
        v[e] = s1 + __shfl_xor(s1, 32);
      }
      if (lane < 16) *(lptr<f32x4>)(dst + 16 * s + 4 * kp) = v;
    }
#pragma unroll
    for (int s = NS; s < 4 * K; ++s)
      if (lane < 16)
        *(lptr<f32x4>)(dst + 16 * s + 4 * kp) = f32x4{0.f, 0.f, 0.f, 0.f};
    // log-likelihood (and residual) sums over the lanes of each chain q
#pragma unroll
    for (int m = 4; m < 64; m <<= 1) {
      lpl += __shfl_xor(lpl, m);
      if (LINEAR) ssq += __shfl_xor(ssq, m);
    }
    if (lane < 4) {
      red_l[grp_w * G + q] = lpl;
      red_s[grp_w * G + q] = LINEAR ? ssq : T(0);
    }
  }

  // One round: every wave passes over its rows for all G chains and leaves
  // its partial sums in LDS.  Returns false (after the first barrier) once
  // every owner has set its done flag.
  template <bool LINEAR>
  __device__ bool group_round() {
    using VT = typename Vec16<T>::type;
    constexpr int V = Vec16<T>::n;
    constexpr int CPB = 16 * V;
    constexpr int NB = kGlmBlocks<T, K>;
    group_barrier();  // 1: coefficients and flags are written
    lptr<const int> done = (lptr<const int>)(grp + GroupLayout<T, K>::done(G));
    bool all = true;
#pragma unroll
    for (int j = 0; j < G; ++j) all = all && done[j] != 0;
    if (all) return false;
    const int lane = lane_id();
    const int dx = LINEAR ? dim - 1 : dim;
    const int xp = x_pitch;
    // a design the passes cannot read (as in glm: rows not 16-byte aligned,
    // a pitch below the coefficients, an empty X) gives NaN partial sums
    if ((xp & 3) != 0 || xp < dx || dx < 1 || n_obs < 1) {
      lptr<T> red_g = (lptr<T>)(grp + GroupLayout<T, K>::red(G));
      lptr<double> red_l =
          (lptr<double>)(grp + GroupLayout<T, K>::lp(G, grp_waves));
      lptr<T> red_s = (lptr<T>)(grp + GroupLayout<T, K>::ssq(G, grp_waves));
#pragma unroll
      for (int j = 0; j < G; ++j) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          red_g[((size_t)grp_w * G + j) * K * 64 + k * 64 + lane] = T(NAN);
        if (lane == 0) {
          red_l[grp_w * G + j] = (double)NAN;
          red_s[grp_w * G + j] = T(NAN);
        }
      }
      group_barrier();  // 2
      return true;
    }
    if constexpr (kMfmaGroup) {
      const int ns = (dx + 15) >> 4;  // live slabs of 16 coefficients
      if (ns <= kMfmaSlabs) {
        switch (ns) {
          case 1: group_mfma<1, LINEAR>(dx, xp); break;
          case 2: group_mfma<(kMfmaSlabs >= 2 ? 2 : 1), LINEAR>(dx, xp); break;
          case 3: group_mfma<(kMfmaSlabs >= 3 ? 3 : 1), LINEAR>(dx, xp); break;
          case 4: group_mfma<(kMfmaSlabs >= 4 ? 4 : 1), LINEAR>(dx, xp); break;
          case 5: group_mfma<(kMfmaSlabs >= 5 ? 5 : 1), LINEAR>(dx, xp); break;
          case 6: group_mfma<(kMfmaSlabs >= 6 ? 6 : 1), LINEAR>(dx, xp); break;
          case 7: group_mfma<(kMfmaSlabs >= 7 ? 7 : 1), LINEAR>(dx, xp); break;
          default: group_mfma<kMfmaSlabs, LINEAR>(dx, xp); break;
        }
        group_barrier();  // 2: every wave's partial sums are in LDS
        return true;
      }
    }
    const int grp16 = lane & 15;
    int cst[NB];
    VT acc[G][NB];
    double lpl[G];
    T ssq[G];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int c = b * CPB + V * grp16;
      cst[b] = c < xp ? c : 0;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[j][b] = VT{};
      lpl[j] = 0.0;
      ssq[j] = T(0);
    }
    const int nbl = (dx + CPB - 1) / CPB;  // live blocks
    if (nbl <= 1 || NB == 1) {
      group_pass<1, LINEAR>(dx, xp, cst, acc, lpl, ssq);
    } else if (nbl == 2 || NB == 2) {
      group_pass<(NB >= 2 ? 2 : 1), LINEAR>(dx, xp, cst, acc, lpl, ssq);
    } else {
      group_pass<NB, LINEAR>(dx, xp, cst, acc, lpl, ssq);
    }
    // sum the four row phases (lanes g, g+16, g+32, g+48) and leave this
    // wave's partial sums in its slots
    lptr<T> red_g = (lptr<T>)(grp + GroupLayout<T, K>::red(G));
    lptr<double> red_l = (lptr<double>)(grp + GroupLayout<T, K>::lp(G, grp_waves));
    lptr<T> red_s = (lptr<T>)(grp + GroupLayout<T, K>::ssq(G, grp_waves));
#pragma unroll
    for (int j = 0; j < G; ++j) {
      lptr<T> dst = red_g + ((size_t)grp_w * G + j) * K * 64;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        VT v = acc[j][b];
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const T s1 = v[e] + __shfl_xor(v[e], 16);
          v[e] = s1 + __shfl_xor(s1, 32);
        }
        if (lane < 16 && b * CPB + V * grp16 < K * 64) {
#pragma unroll
          for (int e = 0; e < V; ++e) dst[b * CPB + V * grp16 + e] = v[e];
        }
      }
      const double lw = wave_sum(lpl[j]);
      const T sw = wave_sum(ssq[j]);
      if (lane == 0) {
        red_l[grp_w * G + j] = lw;
        red_s[grp_w * G + j] = sw;
      }
    }
    group_barrier();  // 2: every wave's partial sums are in LDS
    return true;
  }

  __device__ bool group_round_any() {
    return kind == HIPMC_TARGET_LINEAR ? group_round<true>()
                                       : group_round<false>();
  }

  // A wave with nothing (more) to evaluate joins the rounds until every
  // chain of the group is done; an owner first sets its done flag.
  __device__ void group_finish() {
    if (grp_j >= 0) {
      lptr<int> done = (lptr<int>)(grp + GroupLayout<T, K>::done(G));
      if (lane_id() == 0) done[grp_j] = 1;
    }
    while (group_round_any()) {
    }
  }

  // The owner's evaluation: publish the chain's coefficients (and noise
  // scale), take part in one round, add up the waves' partial sums in a
  // fixed order, then the prior terms as in glm().
  template <bool LINEAR>
  __device__ double glm_grouped(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    const int dx = LINEAR ? dim - 1 : dim;
    const int n = n_obs;
    T sigma = T(1);
    if (LINEAR) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == (dx >> 6)) sigma = readlane(th[k], dx & 63);
    }
    lptr<T> beta = (lptr<T>)(grp + GroupLayout<T, K>::beta(G)) + grp_j * K * 64;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      beta[e] = e < dx ? th[k] : T(0);
    }
    if (lane == 0) ((lptr<T>)(grp + GroupLayout<T, K>::sig(G)))[grp_j] = sigma;
    group_round<LINEAR>();
    lptr<const T> red_g = (lptr<const T>)(grp + GroupLayout<T, K>::red(G));
    lptr<const double> red_l =
        (lptr<const double>)(grp + GroupLayout<T, K>::lp(G, grp_waves));
    lptr<const T> red_s =
        (lptr<const T>)(grp + GroupLayout<T, K>::ssq(G, grp_waves));
    T gl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) gl[k] = T(0);
    double ll = 0.0;
    T S = T(0);
    for (int w = 0; w < grp_waves; ++w) {
      lptr<const T> src = red_g + ((size_t)w * G + grp_j) * K * 64;
#pragma unroll
      for (int k = 0; k < K; ++k) gl[k] += src[k * 64 + lane];
      ll += red_l[w * G + grp_j];
      S += red_s[w * G + grp_j];
    }
    double lpl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool isb = e < dx;
      const T bv = th[k] * inv_s0;
      lpl += isb ? (double)(T(-0.5) * bv * bv) : 0.0;
      g[k] = isb ? gl[k] - th[k] * inv_s0sq : T(0);
    }
    double lp = log_norm + ll;
    if (LINEAR) {
      const T ss = p1;
      const T gs = S / (sigma * sigma * sigma) - (T)n / sigma - sigma / (ss * ss);
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == (dx >> 6) && lane == (dx & 63)) g[k] = gs;
      lp += -(double)n * (double)log(sigma) +
            (double)(T(-0.5) * (sigma / ss) * (sigma / ss));
      if (sigma < T(0)) lp = neg_inf();
    }
    lpu = lp;
    return lpl;
  }
};

}  // namespace hipmc
