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

template <typename T, int K, int FAM>
struct Target {
  int kind, dim, n_obs;
  T p0, p1;
  double log_norm;
  const T* __restrict__ x;   // [n_obs, dim_x] row-major (GLM)
  const T* __restrict__ xt;  // [dim_x, n_obs] (GLM)
  const T* __restrict__ y;
  // per-lane cached data in the state layout
  T a[K], b[K];
  // bijector code of each register slot of this lane; the masks are
  // wave-uniform: bit k set if some lane of slot k uses that bijector
  int code[K];
  uint32_t sp_slots, ex_slots;
  T inv_s0sq, inv_s0, inv_s1, s1sq;  // prior constants (hier)
  T* lds;  // per-wave scratch: GLM beta [K*64] followed by r [64]

  __device__ void load(const hipmc_problem& pb, T* wave_lds) {
    const int lane = lane_id();
    kind = pb.kind;
    dim = pb.dim;
    n_obs = pb.n_obs;
    p0 = (T)pb.param[0];
    p1 = (T)pb.param[1];
    log_norm = pb.log_norm;
    x = (const T*)pb.x;
    xt = (const T*)pb.xt;
    y = (const T*)pb.y;
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
    inv_s0 = T(1) / p0;
    inv_s0sq = inv_s0 * inv_s0;
    inv_s1 = T(1) / p1;
    s1sq = p1 * p1;
    lds = wave_lds;
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

  // Per-lane partial log-prob (double) and wave-uniform part `lpu`; writes
  // the gradient.  The caller reduces the partial (possibly together with
  // other sums).  Bijector log-det-Jacobians are folded into the partial.
  __device__ double eval_partial(const T (&xs)[K], T (&g)[K], double& lpu) {
    T th[K];
    SoftplusParts<T> sp[K];
    T ex[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      th[k] = xs[k];
      if (sp_slots & (1u << k)) {
        keep_branch();
        // Softplus bijector forward (bijectors/softplus.py:36-50): below the
        // cutoff TFP takes log1p(exp(x)), which equals max(x, 0) + l there.
        sp[k] = softplus_parts(xs[k]);
        const T f = fmax(xs[k], T(0)) + sp[k].l;
        th[k] = code[k] == HIPMC_BIJ_SOFTPLUS ? f : th[k];
      }
      if (ex_slots & (1u << k)) {
        keep_branch();
        ex[k] = exp(xs[k]);
        th[k] = code[k] == HIPMC_BIJ_EXP ? ex[k] : th[k];
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
        default: lpl = nan_exp(th, g, lpu); break;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T xv = xs[k];
      if (sp_slots & (1u << k)) {
        keep_branch();
        // fldj = -softplus(-x); d/dx f = sigmoid(x); d/dx fldj = sigmoid(-x)
        const bool on = code[k] == HIPMC_BIJ_SOFTPLUS;
        lpl += on ? (double)(-(fmax(-xv, T(0)) + sp[k].l)) : 0.0;
        g[k] = on ? g[k] * sp[k].sig + sp[k].sig_neg : g[k];
      }
      if (ex_slots & (1u << k)) {
        keep_branch();
        const bool on = code[k] == HIPMC_BIJ_EXP;
        lpl += on ? (double)xv : 0.0;
        g[k] = on ? g[k] * ex[k] + T(1) : g[k];
      }
    }
    return lpl;
  }

  // Wave-uniform target log-prob (double); writes the gradient.
  __device__ double eval(const T (&xs)[K], T (&g)[K]) {
    double lpu;
    const double lpl = eval_partial(xs, g, lpu);
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
    const T gmu = -mu * inv_s0sq + S1;
    const T gtau = T(-2) * tau * recip(s1sq + tau * tau) + S2;
    g[0] = lane == 0 ? gmu : (lane == 1 ? gtau : g[0]);
    const T zm = mu * inv_s0;
    const T zt = tau * inv_s1;
    lpu = log_norm + (double)(T(-0.5) * zm * zm) - (double)log1p_pos(zt * zt);
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

  // Bayesian GLM.  LINEAR=false: logistic regression over beta (dim = D_x).
  // LINEAR=true: [beta, sigma] with Gaussian likelihood (dim = D_x + 1).
  // s = X beta is computed with one row per lane (X^T reads coalesced);
  // X^T r is accumulated with one state element per lane (X reads
  // coalesced), r passing through LDS 64 rows at a time.
  template <bool LINEAR>
  __device__ double glm(const T (&th)[K], T (&g)[K], double& lpu) {
    const int lane = lane_id();
    const int dx = LINEAR ? dim - 1 : dim;
    T sigma = T(1), inv_s2 = T(1);
    if (LINEAR) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == (dx >> 6)) sigma = readlane(th[k], dx & 63);
      inv_s2 = T(1) / (sigma * sigma);
    }
    T* beta = lds;
    T* rbuf = lds + K * 64;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      beta[e] = (e < dx) ? th[k] : T(0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    T gl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) gl[k] = T(0);
    double lpl = 0.0;
    T ssq = T(0);
    const int n = n_obs;
    for (int n0 = 0; n0 < n; n0 += 64) {
      const int row = n0 + lane;
      const bool valid = row < n;
      T s = T(0);
      if (valid) {
        const T* col = xt + row;
        int d = 0;
        for (; d + 4 <= dx; d += 4) {
          s += col[(size_t)d * n] * beta[d];
          s += col[(size_t)(d + 1) * n] * beta[d + 1];
          s += col[(size_t)(d + 2) * n] * beta[d + 2];
          s += col[(size_t)(d + 3) * n] * beta[d + 3];
        }
        for (; d < dx; ++d) s += col[(size_t)d * n] * beta[d];
      }
      const T yv = valid ? y[row] : T(0);
      T r;
      if (LINEAR) {
        const T res = yv - s;
        r = valid ? res * inv_s2 : T(0);
        lpl += valid ? (double)(T(-0.5) * res * res * inv_s2) : 0.0;
        ssq += valid ? res * res : T(0);
      } else {
        const T w0 = T(1) - yv;
        const SoftplusParts<T> sp = softplus_parts(s);
        const T l0 = w0 == T(0) ? T(0) : -(fmax(s, T(0)) + sp.l) * w0;
        const T l1 = yv == T(0) ? T(0) : -(fmax(-s, T(0)) + sp.l) * yv;
        lpl += valid ? (double)(l0 + l1) : 0.0;
        r = valid ? yv - sp.sig : T(0);
      }
      __builtin_amdgcn_wave_barrier();
      rbuf[lane] = r;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int rows = (n - n0) < 64 ? (n - n0) : 64;
      const T* xrow = x + (size_t)n0 * dx;
      for (int j = 0; j < rows; ++j) {
        const T rj = rbuf[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int e = k * 64 + lane;
          if (k * 64 < dx) {
            if (e < dx) gl[k] += xrow[e] * rj;
          }
        }
        xrow += dx;
      }
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
};

}  // namespace hipmc
