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

// Loads a GLM evaluation keeps in flight per wave (one wave per chain is
// latency bound on X, so memory-level parallelism sets its speed).
constexpr int kGlmBatch = 16;

// v[s] for a wave-uniform slot index s (selects, no indexed registers).
template <typename T, int K>
__device__ inline T pick(const T (&v)[K], int s) {
  T r = v[0];
#pragma unroll
  for (int k = 1; k < K; ++k) r = s == k ? v[k] : r;
  return r;
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
  // when every bijector lane sits in one register slot (the common case:
  // a few constrained scalars), that slot and this lane's code in it
  bool one_slot;
  int bslot, bcode;
  bool sp0;  // the only bijector is Softplus, all of it in slot 0
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
    const uint32_t any = sp_slots | ex_slots;
    one_slot = any != 0u && (any & (any - 1u)) == 0u;
    sp0 = sp_slots == 1u && ex_slots == 0u;
    bslot = any ? __builtin_ctz(any) : 0;
    bcode = code[0];
#pragma unroll
    for (int k = 1; k < K; ++k) bcode = bslot == k ? code[k] : bcode;
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
    for (int k = 0; k < K; ++k) th[k] = xs[k];
    // one bijector slot: select it, transform it, put it back (no per-slot
    // branches); otherwise the per-slot loop
    SoftplusParts<T> spb{};
    T exb = T(0);
    if (sp0) {
      spb = softplus_parts(xs[0]);
      th[0] = code[0] == HIPMC_BIJ_SOFTPLUS ? fmax(xs[0], T(0)) + spb.l
                                            : xs[0];
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
#pragma unroll
      for (int k = 0; k < K; ++k) th[k] = bslot == k ? f : th[k];
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
        default: lpl = nan_exp(th, g, lpu); break;
      }
    }
    // fldj: Softplus -softplus(-x) with d/dx f = sigmoid(x),
    // d/dx fldj = sigmoid(-x); Exp x with d/dx f = exp(x), d/dx fldj = 1
    if (sp0) {
      const bool on = code[0] == HIPMC_BIJ_SOFTPLUS;
      lpl += on ? (double)(-(fmax(-xs[0], T(0)) + spb.l)) : 0.0;
      g[0] = on ? g[0] * spb.sig + spb.sig_neg : g[0];
    } else if (one_slot) {
      const T xb = pick(xs, bslot);
      const T gb = pick(g, bslot);
      const bool on_sp = bcode == HIPMC_BIJ_SOFTPLUS;
      const bool on_ex = bcode == HIPMC_BIJ_EXP;
      lpl += on_sp ? (double)(-(fmax(-xb, T(0)) + spb.l))
                   : (on_ex ? (double)xb : 0.0);
      const T gn = on_sp ? gb * spb.sig + spb.sig_neg
                         : (on_ex ? gb * exb + T(1) : gb);
#pragma unroll
      for (int k = 0; k < K; ++k) g[k] = bslot == k ? gn : g[k];
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
  // gl[k] += sum_j X[j, k*64 + lane] r[j] over `rows` rows of row-major X
  // (a lane per coefficient).  Unswitched on the number of live slots; the
  // loads are unconditional (lanes past dx read column dx - 1 and are
  // masked) so kGlmBatch (then 4, then 1) rows of loads stay in flight.
  template <int KK>
  __device__ void grad_rows_k(const T* xrow, int dx, int rows, const T* rbuf,
                              T (&gl)[K]) {
    const int lane = lane_id();
    int col[KK];
    bool on[KK];
#pragma unroll
    for (int k = 0; k < KK; ++k) {
      const int e = k * 64 + lane;
      on[k] = e < dx;
      col[k] = on[k] ? e : dx - 1;
    }
    int j = 0;
    for (; j + kGlmBatch <= rows; j += kGlmBatch) {
      T v[kGlmBatch][KK];
#pragma unroll
      for (int u = 0; u < kGlmBatch; ++u)
#pragma unroll
        for (int k = 0; k < KK; ++k) v[u][k] = xrow[(size_t)(j + u) * dx + col[k]];
#pragma unroll
      for (int u = 0; u < kGlmBatch; ++u) {
        const T rj = rbuf[j + u];
#pragma unroll
        for (int k = 0; k < KK; ++k) gl[k] += (on[k] ? v[u][k] : T(0)) * rj;
      }
    }
    for (; j + 4 <= rows; j += 4) {
      T v[4][KK];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < KK; ++k) v[u][k] = xrow[(size_t)(j + u) * dx + col[k]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T rj = rbuf[j + u];
#pragma unroll
        for (int k = 0; k < KK; ++k) gl[k] += (on[k] ? v[u][k] : T(0)) * rj;
      }
    }
    for (; j < rows; ++j) {
      const T rj = rbuf[j];
#pragma unroll
      for (int k = 0; k < KK; ++k)
        gl[k] += (on[k] ? xrow[(size_t)j * dx + col[k]] : T(0)) * rj;
    }
  }

  __device__ void grad_rows(const T* xrow, int dx, int rows, const T* rbuf,
                            T (&gl)[K]) {
    const int live = (dx + 63) >> 6;
    if (K == 1 || live <= 1) {
      grad_rows_k<1>(xrow, dx, rows, rbuf, gl);
    } else if (K == 2 || live == 2) {
      grad_rows_k<(K >= 2 ? 2 : 1)>(xrow, dx, rows, rbuf, gl);
    } else if (K == 3 || live == 3) {
      grad_rows_k<(K >= 3 ? 3 : 1)>(xrow, dx, rows, rbuf, gl);
    } else {
      grad_rows_k<K>(xrow, dx, rows, rbuf, gl);
    }
  }

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
      // s = x_row . beta: a lane per row, coalesced over X^T; rows past the
      // end read row n - 1 and are masked, so the loads carry no branch.
      // Batches of kGlmBatch, then 4, then 1 keep loads in flight.
      const T* col = xt + (valid ? row : n - 1);
      T s0 = T(0), s1 = T(0);
      int d = 0;
      for (; d + kGlmBatch <= dx; d += kGlmBatch) {
        T v[kGlmBatch];
#pragma unroll
        for (int u = 0; u < kGlmBatch; ++u) v[u] = col[(size_t)(d + u) * n];
#pragma unroll
        for (int u = 0; u < kGlmBatch; u += 2) {
          s0 += v[u] * beta[d + u];
          s1 += v[u + 1] * beta[d + u + 1];
        }
      }
      for (; d + 4 <= dx; d += 4) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = col[(size_t)(d + u) * n];
        s0 += v[0] * beta[d] + v[2] * beta[d + 2];
        s1 += v[1] * beta[d + 1] + v[3] * beta[d + 3];
      }
      for (; d < dx; ++d) s0 += col[(size_t)d * n] * beta[d];
      const T s = s0 + s1;
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
      grad_rows(x + (size_t)n0 * dx, dx, rows, rbuf, gl);
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
