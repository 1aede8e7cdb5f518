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
  // bijector code of each register slot of this lane; any_bij is uniform
  int code[K];
  bool any_bij;
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
    any_bij = pb.n_bij > 0;
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
    }
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

  // Returns the wave-uniform target log-prob (double); writes grad.
  __device__ double eval(const T (&xs)[K], T (&g)[K]) {
    T th[K];
#pragma unroll
    for (int k = 0; k < K; ++k) th[k] = xs[k];
    if (any_bij) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (code[k] == HIPMC_BIJ_SOFTPLUS) th[k] = softplus_forward(xs[k]);
        else if (code[k] == HIPMC_BIJ_EXP) th[k] = exp(xs[k]);
      }
    }
    double lp;
    if (FAM == kFamilyHier) {
      lp = hier_normal(th, g);
    } else if (FAM == kFamilyGlm) {
      lp = kind == HIPMC_TARGET_LINEAR ? glm<true>(th, g) : glm<false>(th, g);
    } else {
      switch (kind) {
        case HIPMC_TARGET_DIAG_NORMAL: lp = diag_normal(th, g); break;
        case HIPMC_TARGET_FUNNEL: lp = funnel(th, g); break;
        default: lp = nan_exp(th, g); break;
      }
    }
    if (any_bij) {
      double lj = 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T xv = xs[k];
        if (code[k] == HIPMC_BIJ_SOFTPLUS) {
          // fldj = -softplus(-x); d/dx f = sigmoid(x); d/dx fldj = sigmoid(-x)
          lj += (double)(-softplus(-xv));
          g[k] = g[k] * sigmoid(xv) + sigmoid(-xv);
        } else if (code[k] == HIPMC_BIJ_EXP) {
          lj += (double)xv;
          g[k] = g[k] * exp(xv) + T(1);
        }
      }
      lp += wave_sum(lj);
    }
    return lp;
  }

  // [mu, tau, z_0..z_{J-1}], y ~ N(mu + tau z, sigma); mu ~ N(0, p0),
  // tau ~ HalfCauchy(0, p1), z ~ N(0, 1).
  __device__ double hier_normal(const T (&th)[K], T (&g)[K]) {
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
    const T S1 = wave_sum(s1);
    const T S2 = wave_sum(s2);
    const T ms = p0, ts = p1;
    const T gmu = -mu / (ms * ms) + S1;
    const T gtau = T(-2) * tau / (ts * ts + tau * tau) + S2;
    if (lane == 0) g[0] = gmu;
    if (lane == 1) g[0] = gtau;
    const T zm = mu / ms;
    const T zt = tau / ts;
    double lp = wave_sum(lpl) + log_norm + (double)(T(-0.5) * zm * zm) -
                (double)log1p(zt * zt);
    if (tau < T(0)) lp = neg_inf();
    return lp;
  }

  __device__ double diag_normal(const T (&th)[K], T (&g)[K]) {
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
    return wave_sum(lpl) + log_norm;
  }

  // [v, x_1..x_{dim-1}]: v ~ N(0, 3), x_i ~ N(0, exp(v/2)).
  __device__ double funnel(const T (&th)[K], T (&g)[K]) {
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
    return log_norm + (double)(T(-0.5) * (v / T(3)) * (v / T(3))) -
           (double)(T(0.5) * nk * v) - (double)(T(0.5) * S * inv_s2);
  }

  __device__ double nan_exp(const T (&th)[K], T (&g)[K]) {
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
    return wave_sum(lpl);
  }

  // Bayesian GLM.  LINEAR=false: logistic regression over beta (dim = D_x).
  // LINEAR=true: [beta, sigma] with Gaussian likelihood (dim = D_x + 1).
  // s = X beta is computed with one row per lane (X^T reads coalesced);
  // X^T r is accumulated with one state element per lane (X reads
  // coalesced), r passing through LDS 64 rows at a time.
  template <bool LINEAR>
  __device__ double glm(const T (&th)[K], T (&g)[K]) {
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
        const T l0 = w0 == T(0) ? T(0) : -softplus(s) * w0;
        const T l1 = yv == T(0) ? T(0) : -softplus(-s) * yv;
        lpl += valid ? (double)(l0 + l1) : 0.0;
        r = valid ? yv - sigmoid(s) : T(0);
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
    const T ps = p0;
    double lp = log_norm;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = k * 64 + lane;
      const bool isb = e < dx;
      const T bv = th[k];
      lpl += isb ? (double)(T(-0.5) * (bv / ps) * (bv / ps)) : 0.0;
      g[k] = isb ? gl[k] - bv / (ps * ps) : T(0);
    }
    lp += wave_sum(lpl);
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
    return lp;
  }
};

}  // namespace hipmc
