// C ABI of libhipmc: argument checks, dispatch to the sampling-kernel
// launchers (built per dtype and target family in inst.hip; the BF16 MFMA
// path in mfma.hip) and the small auxiliary kernels (dual averaging, MH
// accept, bijectors, raw draws, running moments).  No C++ exception crosses
// this boundary; errors are negative return codes plus a thread-local
// message.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <string>

#include "../../include/hipmc.h"
#include "common.h"
#include "runtime.h"
#include "targets.h"

using namespace hipmc;

static_assert(sizeof(hipmc_problem) == 160, "hipmc_problem layout");
static_assert(offsetof(hipmc_problem, x_pitch) == 152, "hipmc_problem layout");
static_assert(sizeof(hipmc_stream) == 48, "hipmc_stream layout");
static_assert(offsetof(hipmc_problem, param) == 56, "hipmc_problem layout");
static_assert(offsetof(hipmc_problem, n_bij) == 96, "hipmc_problem layout");

namespace hipmc {

namespace {
thread_local std::string g_err;
}  // namespace

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return HIPMC_OK;
  return fail(HIPMC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int pick_k(int max_dim, int fam) {
  static const int ks[] = {1, 2, 3, 4, 5, 8, 12, 17};
  for (int k : ks) {
    if (k > max_slots(fam)) break;
    if (k * 64 >= max_dim) return k;
  }
  return -1;
}

int max_state_dim(int fam) { return 64 * max_slots(fam); }

}  // namespace hipmc

namespace {

int check_run(const hipmc_run_args* a, bool nuts) {
  if (!a) return fail(HIPMC_E_INVALID, "args is NULL");
  if (a->dtype != HIPMC_F32 && a->dtype != HIPMC_F64)
    return fail(HIPMC_E_INVALID, "dtype must be HIPMC_F32 or HIPMC_F64");
  if (a->n_chains < 0) return fail(HIPMC_E_INVALID, "n_chains < 0");
  if (a->n_iters < 1) return fail(HIPMC_E_INVALID, "n_iters must be >= 1");
  if (!a->fold_seed && a->n_iters != 1)
    return fail(HIPMC_E_INVALID, "n_iters > 1 requires fold_seed");
  if (a->max_dim < 1 || a->max_dim > a->dim_stride)
    return fail(HIPMC_E_INVALID, "need 1 <= max_dim <= dim_stride");
  {
    const int kind = a->target_kind;
    const int fam = (kind >= HIPMC_TARGET_LOGISTIC && kind <= HIPMC_TARGET_LOG_GAMMA)
                        ? family_of(kind) : kFamilyTest;
    if (pick_k(a->max_dim, fam) < 0)
      return fail(HIPMC_E_UNSUPPORTED,
                  "state dimension above " + std::to_string(max_state_dim(fam)) +
                      " is not built for this target family");
  }
  if (!a->problems || a->n_problems < 1)
    return fail(HIPMC_E_INVALID, "problems is NULL");
  if (!a->step_size || !a->state || !a->tlp || !a->grad)
    return fail(HIPMC_E_INVALID, "step_size/state/tlp/grad must be set");
  if (a->trace_state && (a->trace_every < 1 || a->n_trace < 0))
    return fail(HIPMC_E_INVALID, "trace_every must be >= 1 and n_trace >= 0");
  if (a->out_row0 < 0) return fail(HIPMC_E_INVALID, "out_row0 < 0");
  if (nuts) {
    if (a->max_tree_depth < 1 || a->max_tree_depth > 16)
      return fail(HIPMC_E_INVALID, "max_tree_depth must be in [1, 16]");
    if (a->unrolled_leapfrog_steps < 1)
      return fail(HIPMC_E_INVALID, "unrolled_leapfrog_steps must be >= 1");
  } else if (a->num_leapfrog_steps < 0) {
    return fail(HIPMC_E_INVALID, "num_leapfrog_steps must be >= 0");
  }
  return HIPMC_OK;
}


// Calls L<T, FAM>::run(args...) for the launch's dtype and target family.
template <template <typename, int> class L, typename... A>
int by_family(int dtype, int kind, A&&... args) {
  if (kind < HIPMC_TARGET_LOGISTIC || kind > HIPMC_TARGET_LOG_GAMMA)
    return fail(HIPMC_E_INVALID, "unknown target kind");
  const int fam = family_of(kind);
  if (dtype == HIPMC_F32) {
    if (fam == kFamilyHier) return L<float, kFamilyHier>::run(args...);
    if (fam == kFamilyGlm) return L<float, kFamilyGlm>::run(args...);
    return L<float, kFamilyTest>::run(args...);
  }
  if (dtype == HIPMC_F64) {
    if (fam == kFamilyHier) return L<double, kFamilyHier>::run(args...);
    if (fam == kFamilyGlm) return L<double, kFamilyGlm>::run(args...);
    return L<double, kFamilyTest>::run(args...);
  }
  return fail(HIPMC_E_INVALID, "bad dtype");
}

template <typename T, int FAM>
struct NutsL {
  static int run(const hipmc_run_args& a, hipStream_t s) {
    return launch_nuts_k<T, FAM>(pick_k(a.max_dim, FAM), a, s);
  }
};
template <typename T, int FAM>
struct HmcL {
  static int run(bool calibrated, const hipmc_run_args& a, hipStream_t s) {
    return launch_hmc_k<T, FAM>(pick_k(a.max_dim, FAM), calibrated, a, s);
  }
};
template <typename T, int FAM>
struct EvalL {
  static int run(int k, const hipmc_problem* p, const int32_t* cp, int n,
                 int stride, const void* x, void* tlp, void* grad,
                 hipStream_t s) {
    return launch_eval_k<T, FAM>(k, p, cp, n, stride, x, tlp, grad, s);
  }
};

// ------------------------------------------------------------ DA kernel
// reduce_logmeanexp of hmc_like_log_accept_prob_getter_fn(lar) over chains
// [b, e) (`simple_step_size_adaptation.py:42-48`, `math/generic.py:221-273`),
// float64; every thread of the block returns the same value.
// harmonic = true: reduce_log_harmonic_mean_exp = -logmeanexp(-x)
// (`math/generic.py:387-432`), the reduction windowed adaptation gives its
// dual averaging (`experimental/mcmc/windowed_sampling.py:869-873`).
template <typename T>
__device__ double group_log_mean_accept(const T* lar, int b, int e,
                                        double* red, bool harmonic = false) {
  const int tid = threadIdx.x;
  const double sgn = harmonic ? -1.0 : 1.0;
  auto lap_of = [&](int i) -> double {
    double v = (double)lar[i];
    if (!__builtin_isfinite(v)) v = -__builtin_inf();
    return sgn * fmin(v, 0.0);
  };
  double m = -__builtin_inf();
  for (int i = b + tid; i < e; i += blockDim.x) m = fmax(m, lap_of(i));
  red[tid] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  double mx = red[0];
  if (!__builtin_isfinite(mx)) mx = 0.0;
  __syncthreads();
  double acc = 0.0;
  for (int i = b + tid; i < e; i += blockDim.x) acc += exp(lap_of(i) - mx);
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double total = red[0];
  __syncthreads();
  return sgn * (log(total) + mx - log((double)(e - b)));
}

// One block per step-size group; float64 throughout
// (dual_averaging_step_size_adaptation.py:399-475).
template <typename T>
__global__ void __launch_bounds__(256)
    da_kernel(const int32_t* step_group, const int32_t* state_index,
              const int32_t* group_start,
              const T* lar_base, int n_rows, int64_t row_stride,
              double* error_sum, double* log_avg, const double* log_shrink,
              const double* target_accept, int64_t step_all,
              int64_t num_adapt_all, const int64_t* elem_step,
              const int64_t* elem_num_adapt,
              double gamma, double t0, double kappa, T* step_size,
              double* step_f64, int harmonic) {
  __shared__ double red[256];
  const int elem = blockIdx.x;
  const int64_t step0 = elem_step ? elem_step[elem] : step_all;
  const int64_t num_adapt = elem_num_adapt ? elem_num_adapt[elem] : num_adapt_all;
  const int grp = step_group ? step_group[elem] : elem;
  const int gidx = state_index ? state_index[elem] : elem;
  const int b = group_start[grp], e = group_start[grp + 1];
  const int tid = threadIdx.x;
  double err = error_sum[gidx];
  double la = log_avg[gidx];
  double cur = step_f64 ? step_f64[gidx] : (double)step_size[gidx];
  for (int r = 0; r < n_rows; ++r) {
    const T* lar = lar_base + (size_t)r * row_stride;
    // reduce_logmeanexp, then `_one_step_part`
    const double rlap = group_log_mean_accept(lar, b, e, red, harmonic != 0);
    err = err + target_accept[gidx] - exp(rlap);
    const int64_t step = step0 + r;
    const double t = (double)step + 1.0;
    const double new_log_step =
        log_shrink[gidx] - err * sqrt(t) / ((t0 + t) * gamma);
    const double eta = pow(t, -kappa);
    const double nla = eta * new_log_step + (1.0 - eta) * la;
    const int64_t st = step + 1;
    if (st < num_adapt) cur = exp(new_log_step);
    else if (st == num_adapt) cur = exp(nla);
    if (st <= num_adapt) la = nla;
  }
  if (tid == 0) {
    error_sum[gidx] = err;
    log_avg[gidx] = la;
    step_size[gidx] = (T)cur;
    if (step_f64) step_f64[gidx] = cur;
  }
}

// SimpleStepSizeAdaptation (`simple_step_size_adaptation.py:344-438`): one
// block per step-size element; the element grows by (1 + rate) when its
// group's log-mean accept probability exceeds log(target), else shrinks by
// it, while step < num_adaptation_steps.  The update is in the step-size
// dtype, as the reference multiplies in that dtype.
template <typename T>
__global__ void __launch_bounds__(256)
    simple_step_kernel(const int32_t* step_group, const int32_t* group_start,
                       const T* lar_base, int n_rows, int64_t row_stride,
                       const double* target_accept, double rate,
                       int64_t step0, int64_t num_adapt, T* step_size) {
  __shared__ double red[256];
  const int elem = blockIdx.x;
  const int grp = step_group ? step_group[elem] : elem;
  const int b = group_start[grp], e = group_start[grp + 1];
  const T opr = T(1) + (T)rate;
  const double log_target = (double)log((T)target_accept[elem]);
  T cur = step_size[elem];
  for (int r = 0; r < n_rows; ++r) {
    const double rlap =
        group_log_mean_accept(lar_base + (size_t)r * row_stride, b, e, red);
    if (step0 + r < num_adapt) cur = rlap > log_target ? cur * opr : cur / opr;
  }
  if (threadIdx.x == 0) step_size[elem] = cur;
}

// ------------------------------------------------------------ MH kernels
template <typename T>
__global__ void mh_decide_kernel(int n, Seed sd, const hipmc_problem* problems,
                                 const int32_t* chain_problem,
                                 const int32_t* chain_local, const T* ptlp,
                                 const T* corr, T* tlp, T* out_lar,
                                 uint8_t* accepted) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const uint32_t prob = problems[chain_problem ? chain_problem[c] : 0].key;
  const uint32_t local = chain_local ? chain_local[c] : c;
  double lar = (double)ptlp[c] - (double)tlp[c];
  if (corr) lar += (double)corr[c];
  if (!__builtin_isfinite(lar)) lar = -__builtin_inf();
  const U4 w = philox(U4{0u, local, prob, kPurposeMHAccept}, sd.k0, sd.k1);
  const double u = (double)uniform_from<T>(w.x, w.y);
  const bool acc = log(u) < lar;
  accepted[c] = acc;
  if (out_lar) out_lar[c] = (T)lar;
  if (acc) tlp[c] = ptlp[c];
}

template <typename T>
__global__ void select_rows_kernel(int n, int stride, const uint8_t* accepted,
                                   const T* ps, const T* pg, T* s, T* g) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * stride) return;
  if (accepted[i / stride]) {
    s[i] = ps[i];
    g[i] = pg[i];
  }
}

// ------------------------------------------------------------ bijectors
template <typename T>
__device__ inline T softplus_inverse(T y) {
  // math/generic.py:571-584
  const T eps = sizeof(T) == 4 ? (T)1.1920928955078125e-07
                               : (T)2.220446049250313e-16;
  const T threshold = log(eps) + T(2);
  const bool too_small = y < exp(threshold);
  const bool too_large = y > -threshold;
  const T safe = (too_small || too_large) ? T(1) : y;
  const T mid = safe + log(-expm1(-safe));
  return too_small ? log(y) : (too_large ? y : mid);
}

template <typename T>
__global__ void bijector_kernel(const hipmc_problem* problems,
                                const int32_t* chain_problem, int n,
                                int stride, int period, int direction,
                                const T* x, T* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * stride) return;
  const int c = period > 0 ? (int)((i / stride) % period) : (int)(i / stride);
  const int e = (int)(i % stride);
  const hipmc_problem& pb = problems[chain_problem ? chain_problem[c] : 0];
  T v = x[i];
  int code = HIPMC_BIJ_IDENTITY;
  for (int r = 0; r < pb.n_bij; ++r)
    if (e >= pb.bij_off[r] && e < pb.bij_off[r] + pb.bij_len[r])
      code = pb.bij_code[r];
  if (code == HIPMC_BIJ_SOFTPLUS)
    v = direction == 0 ? softplus_forward(v) : softplus_inverse(v);
  else if (code == HIPMC_BIJ_EXP)
    v = direction == 0 ? exp(v) : log(v);
  out[i] = v;
}

// ------------------------------------------------------------ raw draws
template <typename T>
__global__ void draws_kernel(int kind, Seed sd, uint32_t prob, int n, int dim,
                             int depth, int index, T* out, uint8_t* aux) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kind == 0) {
    if (i >= (size_t)n * dim) return;
    const uint32_t c = (uint32_t)(i / dim);
    const int e = (int)(i % dim);
    const U4 w = philox(U4{(uint32_t)(e >> 2), c, prob, kPurposeMomentum},
                        sd.k0, sd.k1);
    const bool hi = (e & 2) != 0;
    double n0, n1;
    box_muller(hi ? w.z : w.x, hi ? w.w : w.y, n0, n1);
    out[i] = (T)((e & 1) ? n1 : n0);
    return;
  }
  if (kind == 4) {
    // uniform on (-2, 2): element e of chain c from word e % 4 of block
    // (e / 4, c, problem, INIT)
    if (i >= (size_t)n * dim) return;
    const uint32_t c = (uint32_t)(i / dim);
    const int e = (int)(i % dim);
    const U4 w = philox(U4{(uint32_t)(e >> 2), c, prob, kPurposeInit}, sd.k0,
                        sd.k1);
    const uint32_t wd = (e & 3) == 0 ? w.x : (e & 3) == 1 ? w.y
                      : (e & 3) == 2 ? w.z : w.w;
    out[i] = (T)(4.0 * ((double)(wd >> 8) + 0.5) * (1.0 / 16777216.0) - 2.0);
    return;
  }
  if (i >= (size_t)n) return;
  const uint32_t c = (uint32_t)i;
  if (kind == 1) {
    const U4 w = philox(U4{0u, c, prob, kPurposeMHAccept}, sd.k0, sd.k1);
    out[i] = uniform_from<T>(w.x, w.y);
  } else if (kind == 2) {
    const U4 w = philox(U4{0u, c, prob, kPurposeTree | (uint32_t)depth},
                        sd.k0, sd.k1);
    if (aux) aux[i] = (uint8_t)(w.x >> 31);
    out[i] = uniform_from<T>(w.y, w.z);
  } else {
    const U4 w = philox(U4{(uint32_t)index, c, prob,
                           kPurposeLeaf | (uint32_t)depth},
                        sd.k0, sd.k1);
    out[i] = uniform_from<T>(w.x, w.y);
  }
}

int blocks_for(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

// Running moments (Pebay's one-pass update), one thread per element of the
// mean (modes 0 and 1); x is [chunk_n, n_el] in the sample dtype.  Reads the
// sums before the chunk (mean_in, ssr_in) and writes the updated ones
// (mean_out, ssr_out; the same buffers for an in-place update).
template <typename T>
__global__ void moments_elem_kernel(const T* __restrict__ x, int64_t n_el,
                                    int64_t chunk_n, double n_old,
                                    const double* mean_in,
                                    const double* ssr_in, double* mean_out,
                                    double* ssr_out) {
  const double cn = (double)chunk_n, n_new = n_old + cn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_el;
       i += (int64_t)gridDim.x * blockDim.x) {
    double cm = 0.0;
    for (int64_t r = 0; r < chunk_n; ++r) cm += (double)x[r * n_el + i];
    cm /= cn;
    const double m = mean_in[i];
    const double delta = cm - m;
    if (ssr_out) {
      double cs = 0.0;  // exactly 0 for a single sample
      for (int64_t r = 0; r < chunk_n && chunk_n > 1; ++r) {
        const double e = (double)x[r * n_el + i] - cm;
        cs += e * e;
      }
      ssr_out[i] = ssr_in[i] + cs + (n_old * cn / n_new) * delta * delta;
    }
    mean_out[i] = m + cn * delta / n_new;
  }
}

// Cross-product sums (mode 2): blockIdx.y strides over batch members, and
// the threads of the blockIdx.x range stride over the member's ev x ev sums
// (contiguous, so reads and writes coalesce), splitting the flat index with
// 32-bit arithmetic (the host keeps ev * ev below 2^31).  Reads the mean
// before this chunk, so the mean update runs after it on the same stream.
template <typename T>
__global__ void moments_cross_kernel(const T* __restrict__ x, int64_t n_batch,
                                     int64_t ev, int64_t chunk_n,
                                     double n_old,
                                     const double* __restrict__ mean_in,
                                     const double* ssr_in, double* ssr_out,
                                     int paired) {
  const double cn = (double)chunk_n, n_new = n_old + cn;
  const double adj = n_old * cn / n_new;
  const int64_t n_el = n_batch * ev;
  const uint32_t e32 = (uint32_t)ev, ev2 = e32 * e32;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (paired) {
    // even ev and 16-byte aligned sums: each thread takes two neighbours
    // (j, j + 1) of one row, with 16-byte accesses to the sums and the mean
    // (j is even, so j + 1 < ev)
    const uint32_t n_pairs = ev2 >> 1;
    for (int64_t b = blockIdx.y; b < n_batch; b += gridDim.y) {
      const size_t base = (size_t)b * ev2;
      for (uint32_t pi = blockIdx.x * blockDim.x + threadIdx.x; pi < n_pairs;
           pi += stride) {
        const uint32_t ij = pi << 1;
        const uint32_t i = ij / e32, j = ij - i * e32;
        const int64_t ei = b * ev + i, ej = b * ev + j;
        double mi = 0.0, m0 = 0.0, m1 = 0.0;
        for (int64_t r = 0; r < chunk_n; ++r) {
          mi += (double)x[r * n_el + ei];
          m0 += (double)x[r * n_el + ej];
          m1 += (double)x[r * n_el + ej + 1];
        }
        mi /= cn;
        m0 /= cn;
        m1 /= cn;
        double c0 = 0.0, c1 = 0.0;  // exactly 0 for a single sample
        for (int64_t r = 0; r < chunk_n && chunk_n > 1; ++r) {
          const double e = (double)x[r * n_el + ei] - mi;
          c0 += e * ((double)x[r * n_el + ej] - m0);
          c1 += e * ((double)x[r * n_el + ej + 1] - m1);
        }
        const double di = adj * (mi - mean_in[ei]);
        const double2 mj = *(const double2*)(mean_in + ej);
        const double2 old = *(const double2*)(ssr_in + base + ij);
        double2 nv;
        nv.x = old.x + c0 + di * (m0 - mj.x);
        nv.y = old.y + c1 + di * (m1 - mj.y);
        *(double2*)(ssr_out + base + ij) = nv;
      }
    }
    return;
  }
  for (int64_t b = blockIdx.y; b < n_batch; b += gridDim.y) {
    const size_t base = (size_t)b * ev2;
    for (uint32_t ij = blockIdx.x * blockDim.x + threadIdx.x; ij < ev2;
         ij += stride) {
      const uint32_t i = ij / e32, j = ij - i * e32;
      const int64_t ei = b * ev + i, ej = b * ev + j;
      double mi = 0.0, mj = 0.0;
      for (int64_t r = 0; r < chunk_n; ++r) {
        mi += (double)x[r * n_el + ei];
        mj += (double)x[r * n_el + ej];
      }
      mi /= cn;
      mj /= cn;
      double cs = 0.0;  // exactly 0 for a single sample
      for (int64_t r = 0; r < chunk_n && chunk_n > 1; ++r)
        cs += ((double)x[r * n_el + ei] - mi) *
              ((double)x[r * n_el + ej] - mj);
      ssr_out[base + ij] = ssr_in[base + ij] + cs +
                           adj * (mi - mean_in[ei]) * (mj - mean_in[ej]);
    }
  }
}

// Running variance of each group's states (the DiagonalMassMatrixAdaptation
// update, experimental/mcmc/diagonal_mass_matrix_adaptation.py:186-228: the
// chains of a group are one chunk of observations).  One block per group;
// thread e takes element e (strided): the chunk's mean and residual sum in
// float64, then Pebay's merge into (num_samples, mean, ssr).  num_samples
// is read by every thread before thread 0 writes the new count.
template <typename T>
__global__ void __launch_bounds__(256)
    variance_update_kernel(const int32_t* group_start, const int32_t* chain_row,
                           const T* x, int dim_stride, int dim,
                           double* num_samples, double* mean, double* ssr) {
  const int g = blockIdx.x;
  const int b = group_start[g], e1 = group_start[g + 1];
  const double n0 = num_samples[g];
  const double cn = (double)(e1 - b);
  const double nn = n0 + cn;
  for (int e = threadIdx.x; e < dim; e += blockDim.x) {
    double s = 0.0;
    for (int c = b; c < e1; ++c) {
      const int r = chain_row ? chain_row[c] : c;
      s += (double)x[(size_t)r * dim_stride + e];
    }
    const double cm = s / cn;
    double cs = 0.0;
    for (int c = b; c < e1; ++c) {
      const int r = chain_row ? chain_row[c] : c;
      const double d = (double)x[(size_t)r * dim_stride + e] - cm;
      cs += d * d;
    }
    const size_t i = (size_t)g * dim_stride + e;
    const double delta = cm - mean[i];
    ssr[i] = ssr[i] + cs + (n0 * cn / nn) * delta * delta;
    mean[i] = mean[i] + cn * delta / nn;
  }
  __syncthreads();
  if (threadIdx.x == 0) num_samples[g] = nn;
}

static unsigned moments_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return (unsigned)(blocks < 8192 ? (blocks > 0 ? blocks : 1) : 8192);
}

template <typename T>
static int moments_launch(int32_t mode, int64_t n_batch, int64_t ev,
                          int64_t chunk_n, const T* x, double n_old,
                          const double* mean_in, const double* ssr_in,
                          double* mean_out, double* ssr_out, hipStream_t s) {
  const int64_t n_el = n_batch * ev;
  if (mode == 2) {
    if (ev > 46340)  // ev * ev must stay below 2^31
      return fail(HIPMC_E_UNSUPPORTED,
                  "moments_update: cross products of more than 46340 "
                  "event elements are not built");
    const int64_t ev2 = ev * ev;
    const unsigned gx = (unsigned)((ev2 + 255) / 256 < 64 ? (ev2 + 255) / 256 : 64);
    const unsigned gy = (unsigned)(n_batch < 65535 ? n_batch : 65535);
    // the paired path reads mean and sums 16 bytes at a time: it needs an
    // even event size and 16-byte aligned arrays (else the scalar path)
    const auto a16 = [](const void* p) {
      return ((uintptr_t)p & 15u) == 0;
    };
    const int paired = (ev % 2 == 0) && a16(mean_in) && a16(ssr_in) &&
                       a16(ssr_out);
    hipLaunchKernelGGL(moments_cross_kernel<T>, dim3(gx, gy), dim3(256), 0, s,
                       x, n_batch, ev, chunk_n, n_old, mean_in, ssr_in,
                       ssr_out, paired);
    const int rc = hip_check(hipGetLastError(), "moments_cross_kernel launch");
    if (rc) return rc;
  }
  const bool elem_ssr = mode == 1;
  hipLaunchKernelGGL(moments_elem_kernel<T>, dim3(moments_grid(n_el)),
                     dim3(256), 0, s, x, n_el, chunk_n, n_old, mean_in,
                     elem_ssr ? ssr_in : nullptr, mean_out,
                     elem_ssr ? ssr_out : nullptr);
  return hip_check(hipGetLastError(), "moments_elem_kernel launch");
}

extern "C" {

const char* hipmc_last_error(void) { return g_err.c_str(); }
int hipmc_abi_version(void) { return HIPMC_ABI_VERSION; }
int hipmc_sizeof_problem(void) { return (int)sizeof(hipmc_problem); }

int hipmc_seed_split(const uint32_t seed[2], uint32_t i, uint32_t out[2]) {
  if (!seed || !out) return fail(HIPMC_E_INVALID, "seed_split: NULL argument");
  const Seed r = split(Seed{seed[0], seed[1]}, i);
  out[0] = r.k0;
  out[1] = r.k1;
  return HIPMC_OK;
}

int hipmc_seed_fold_in(const uint32_t seed[2], uint64_t t, uint32_t out[2]) {
  if (!seed || !out) return fail(HIPMC_E_INVALID, "seed_fold_in: NULL argument");
  const Seed r = fold_in(Seed{seed[0], seed[1]}, t);
  out[0] = r.k0;
  out[1] = r.k1;
  return HIPMC_OK;
}

int hipmc_seed_salt(const uint32_t seed[2], uint32_t salt_hash,
                    uint32_t out[2]) {
  if (!seed || !out) return fail(HIPMC_E_INVALID, "seed_salt: NULL argument");
  const U4 o = philox(U4{salt_hash, 0u, 0u, 0x53414C54u}, seed[0], seed[1]);
  out[0] = o.x;
  out[1] = o.y;
  return HIPMC_OK;
}

int hipmc_philox_host(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]) {
  if (!ctr || !key || !out) return fail(HIPMC_E_INVALID, "philox_host: NULL argument");
  const U4 o = philox(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = o.x;
  out[1] = o.y;
  out[2] = o.z;
  out[3] = o.w;
  return HIPMC_OK;
}
int hipmc_sizeof_run_args(void) { return (int)sizeof(hipmc_run_args); }
int hipmc_sizeof_stream(void) { return (int)sizeof(hipmc_stream); }
int hipmc_sizeof_glm_hmc_args(void) {
  return (int)sizeof(hipmc_glm_hmc_args);
}

int hipmc_glm_hmc_run(const hipmc_glm_hmc_args* a, uint64_t stream) {
  return launch_glm_hmc(a, as_stream(stream));
}

int hipmc_glm_hmc_run_noalloc(const hipmc_glm_hmc_args* a, uint64_t stream) {
  return launch_glm_hmc_noalloc(a, as_stream(stream));
}

int hipmc_nuts_run(const hipmc_run_args* a, uint64_t stream) {
  int rc = check_run(a, true);
  if (rc) return rc;
  return by_family<NutsL>(a->dtype, a->target_kind, *a, as_stream(stream));
}

int hipmc_hmc_run(const hipmc_run_args* a, uint64_t stream) {
  int rc = check_run(a, false);
  if (rc) return rc;
  return by_family<HmcL>(a->dtype, a->target_kind, true, *a,
                         as_stream(stream));
}

int hipmc_leapfrog_run(const hipmc_run_args* a, uint64_t stream) {
  int rc = check_run(a, false);
  if (rc) return rc;
  if (a->n_iters != 1) return fail(HIPMC_E_INVALID, "leapfrog_run: n_iters must be 1");
  return by_family<HmcL>(a->dtype, a->target_kind, false, *a,
                         as_stream(stream));
}

int hipmc_target_eval(int32_t dtype, int32_t target_kind,
                      const hipmc_problem* problems,
                      const int32_t* chain_problem, int32_t n_chains,
                      int32_t dim_stride, int32_t max_dim, const void* x,
                      void* out_tlp, void* out_grad, uint64_t stream) {
  if (!problems || !x) return fail(HIPMC_E_INVALID, "problems/x is NULL");
  if (max_dim < 1 || max_dim > dim_stride)
    return fail(HIPMC_E_INVALID, "need 1 <= max_dim <= dim_stride");
  if (target_kind < HIPMC_TARGET_LOGISTIC || target_kind > HIPMC_TARGET_LOG_GAMMA)
    return fail(HIPMC_E_INVALID, "unknown target kind");
  const int fam = family_of(target_kind);
  const int k = pick_k(max_dim, fam);
  if (k < 0)
    return fail(HIPMC_E_UNSUPPORTED,
                "state dimension above " + std::to_string(max_state_dim(fam)) +
                    " is not built for this target family");
  return by_family<EvalL>(dtype, target_kind, k, problems, chain_problem,
                         n_chains, dim_stride, x, out_tlp, out_grad,
                         as_stream(stream));
}

int hipmc_da_update(int32_t dtype, int32_t n_groups, const int32_t* step_group,
                    const int32_t* state_index,
                    const int32_t* group_start, const void* lar,
                    int32_t n_rows, int64_t row_stride,
                    double* error_sum, double* log_averaging_step,
                    const double* log_shrinkage_target,
                    const double* target_accept_prob, int64_t step,
                    int64_t num_adaptation_steps, double exploration_shrinkage,
                    double step_count_smoothing, double decay_rate,
                    int32_t reduce_mode, const int64_t* elem_step,
                    const int64_t* elem_num_adapt, void* step_size,
                    double* new_step_size_f64, uint64_t stream) {
  if (reduce_mode != HIPMC_REDUCE_MEAN && reduce_mode != HIPMC_REDUCE_HARMONIC)
    return fail(HIPMC_E_INVALID, "da_update: bad reduce_mode");
  if (n_groups < 1 || !group_start || !lar || !error_sum ||
      !log_averaging_step || !log_shrinkage_target || !target_accept_prob ||
      !step_size)
    return fail(HIPMC_E_INVALID, "da_update: NULL argument or n_groups < 1");
  if (n_rows < 1) return fail(HIPMC_E_INVALID, "da_update: n_rows must be >= 1");
  hipStream_t s = as_stream(stream);
  if (dtype == HIPMC_F32)
    hipLaunchKernelGGL(da_kernel<float>, dim3(n_groups), dim3(256), 0, s,
                       step_group, state_index, group_start,
                       (const float*)lar, n_rows,
                       row_stride, error_sum,
                       log_averaging_step, log_shrinkage_target,
                       target_accept_prob, step, num_adaptation_steps,
                       elem_step, elem_num_adapt,
                       exploration_shrinkage, step_count_smoothing, decay_rate,
                       (float*)step_size, new_step_size_f64, reduce_mode);
  else if (dtype == HIPMC_F64)
    hipLaunchKernelGGL(da_kernel<double>, dim3(n_groups), dim3(256), 0, s,
                       step_group, state_index, group_start,
                       (const double*)lar, n_rows,
                       row_stride, error_sum,
                       log_averaging_step, log_shrinkage_target,
                       target_accept_prob, step, num_adaptation_steps,
                       elem_step, elem_num_adapt,
                       exploration_shrinkage, step_count_smoothing, decay_rate,
                       (double*)step_size, new_step_size_f64, reduce_mode);
  else
    return fail(HIPMC_E_INVALID, "bad dtype");
  return hip_check(hipGetLastError(), "da_kernel launch");
}

int hipmc_simple_step_update(int32_t dtype, int32_t n_elems,
                             const int32_t* step_group,
                             const int32_t* group_start, const void* lar,
                             int32_t n_rows, int64_t row_stride,
                             const double* target_accept_prob,
                             double adaptation_rate, int64_t step,
                             int64_t num_adaptation_steps, void* step_size,
                             uint64_t stream) {
  if (n_elems < 1 || !group_start || !lar || !target_accept_prob || !step_size)
    return fail(HIPMC_E_INVALID,
                "simple_step_update: NULL argument or n_elems < 1");
  if (n_rows < 1)
    return fail(HIPMC_E_INVALID, "simple_step_update: n_rows must be >= 1");
  hipStream_t s = as_stream(stream);
  if (dtype == HIPMC_F32)
    hipLaunchKernelGGL(simple_step_kernel<float>, dim3(n_elems), dim3(256), 0,
                       s, step_group, group_start, (const float*)lar, n_rows,
                       row_stride, target_accept_prob, adaptation_rate, step,
                       num_adaptation_steps, (float*)step_size);
  else if (dtype == HIPMC_F64)
    hipLaunchKernelGGL(simple_step_kernel<double>, dim3(n_elems), dim3(256), 0,
                       s, step_group, group_start, (const double*)lar, n_rows,
                       row_stride, target_accept_prob, adaptation_rate, step,
                       num_adaptation_steps, (double*)step_size);
  else
    return fail(HIPMC_E_INVALID, "bad dtype");
  return hip_check(hipGetLastError(), "simple_step_kernel launch");
}

int hipmc_moments_update(int32_t dtype, int32_t mode, int64_t n_batch,
                         int64_t event_size, int64_t chunk_n,
                         const void* sample, double num_samples,
                         const double* mean_in, const double* ssr_in,
                         double* mean_out, double* ssr_out, uint64_t stream) {
  if (mode < 0 || mode > 2)
    return fail(HIPMC_E_INVALID, "moments_update: mode must be 0, 1 or 2");
  if (n_batch < 1 || event_size < 1 || chunk_n < 1)
    return fail(HIPMC_E_INVALID,
                "moments_update: n_batch, event_size and chunk_n must be >= 1");
  if (!sample || !mean_in || !mean_out || (mode > 0 && (!ssr_in || !ssr_out)))
    return fail(HIPMC_E_INVALID, "moments_update: NULL argument");
  if (!(num_samples >= 0.0))
    return fail(HIPMC_E_INVALID, "moments_update: num_samples must be >= 0");
  hipStream_t s = as_stream(stream);
  if (dtype == HIPMC_F32)
    return moments_launch(mode, n_batch, event_size, chunk_n,
                          (const float*)sample, num_samples, mean_in, ssr_in,
                          mean_out, ssr_out, s);
  if (dtype == HIPMC_F64)
    return moments_launch(mode, n_batch, event_size, chunk_n,
                          (const double*)sample, num_samples, mean_in, ssr_in,
                          mean_out, ssr_out, s);
  return fail(HIPMC_E_INVALID, "bad dtype");
}

int hipmc_variance_update(int32_t dtype, int32_t n_groups,
                          const int32_t* group_start, const int32_t* chain_row,
                          const void* x, int32_t dim_stride, int32_t dim,
                          double* num_samples, double* mean, double* ssr,
                          uint64_t stream) {
  if (n_groups < 0 || !group_start || !x || !num_samples || !mean || !ssr)
    return fail(HIPMC_E_INVALID, "variance_update: NULL argument");
  if (dim < 0 || dim > dim_stride)
    return fail(HIPMC_E_INVALID, "variance_update: need 0 <= dim <= dim_stride");
  if (n_groups == 0) return HIPMC_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == HIPMC_F32)
    hipLaunchKernelGGL(variance_update_kernel<float>, dim3(n_groups),
                       dim3(256), 0, s, group_start, chain_row,
                       (const float*)x, dim_stride, dim, num_samples, mean,
                       ssr);
  else if (dtype == HIPMC_F64)
    hipLaunchKernelGGL(variance_update_kernel<double>, dim3(n_groups),
                       dim3(256), 0, s, group_start, chain_row,
                       (const double*)x, dim_stride, dim, num_samples, mean,
                       ssr);
  else
    return fail(HIPMC_E_INVALID, "bad dtype");
  return hip_check(hipGetLastError(), "variance_update_kernel launch");
}

int hipmc_mh_accept(int32_t dtype, int32_t n, int32_t stride,
                    const uint32_t seed[2], const hipmc_problem* problems,
                    const int32_t* chain_problem,
                    const int32_t* chain_local, const void* ps, const void* pt,
                    const void* pg, const void* corr, void* state, void* tlp,
                    void* grad, void* out_lar, uint8_t* out_acc,
                    uint64_t stream) {
  if (!seed || !problems || !ps || !pt || !pg || !state || !tlp || !grad ||
      !out_acc)
    return fail(HIPMC_E_INVALID, "mh_accept: NULL argument");
  if (n <= 0) return HIPMC_OK;
  hipStream_t s = as_stream(stream);
  const Seed sd{seed[0], seed[1]};
  const size_t ne = (size_t)n * stride;
  if (dtype == HIPMC_F32) {
    hipLaunchKernelGGL(mh_decide_kernel<float>, dim3(blocks_for(n, 256)),
                       dim3(256), 0, s, n, sd, problems, chain_problem,
                       chain_local, (const float*)pt, (const float*)corr, (float*)tlp,
                       (float*)out_lar, out_acc);
    hipLaunchKernelGGL(select_rows_kernel<float>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, n, stride, out_acc, (const float*)ps,
                       (const float*)pg, (float*)state, (float*)grad);
  } else if (dtype == HIPMC_F64) {
    hipLaunchKernelGGL(mh_decide_kernel<double>, dim3(blocks_for(n, 256)),
                       dim3(256), 0, s, n, sd, problems, chain_problem,
                       chain_local, (const double*)pt, (const double*)corr, (double*)tlp,
                       (double*)out_lar, out_acc);
    hipLaunchKernelGGL(select_rows_kernel<double>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, n, stride, out_acc, (const double*)ps,
                       (const double*)pg, (double*)state, (double*)grad);
  } else {
    return fail(HIPMC_E_INVALID, "bad dtype");
  }
  return hip_check(hipGetLastError(), "mh kernels launch");
}

int hipmc_bijector(int32_t dtype, const hipmc_problem* problems,
                   const int32_t* chain_problem, int32_t n, int32_t stride,
                   int32_t chain_period, int32_t direction, const void* x,
                   void* out, uint64_t stream) {
  if (!problems || !x || !out) return fail(HIPMC_E_INVALID, "bijector: NULL argument");
  if (direction != 0 && direction != 1)
    return fail(HIPMC_E_INVALID, "bijector: direction must be 0 or 1");
  const size_t ne = (size_t)n * stride;
  if (ne == 0) return HIPMC_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == HIPMC_F32)
    hipLaunchKernelGGL(bijector_kernel<float>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, problems, chain_problem, n, stride,
                       chain_period, direction, (const float*)x, (float*)out);
  else if (dtype == HIPMC_F64)
    hipLaunchKernelGGL(bijector_kernel<double>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, problems, chain_problem, n, stride,
                       chain_period, direction, (const double*)x,
                       (double*)out);
  else
    return fail(HIPMC_E_INVALID, "bad dtype");
  return hip_check(hipGetLastError(), "bijector_kernel launch");
}

int hipmc_draws(int32_t dtype, int32_t kind, const uint32_t seed[2],
                int32_t problem, int32_t n, int32_t dim, int32_t depth,
                int32_t index, void* out, uint8_t* aux, uint64_t stream) {
  if (!seed || !out) return fail(HIPMC_E_INVALID, "draws: NULL argument");
  if (kind < 0 || kind > 4) return fail(HIPMC_E_INVALID, "draws: bad kind");
  const size_t ne = (kind == 0 || kind == 4) ? (size_t)n * dim : (size_t)n;
  if (ne == 0) return HIPMC_OK;
  hipStream_t s = as_stream(stream);
  const Seed sd{seed[0], seed[1]};
  if (dtype == HIPMC_F32)
    hipLaunchKernelGGL(draws_kernel<float>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, kind, sd, (uint32_t)problem, n, dim,
                       depth, index, (float*)out, aux);
  else if (dtype == HIPMC_F64)
    hipLaunchKernelGGL(draws_kernel<double>, dim3(blocks_for(ne, 256)),
                       dim3(256), 0, s, kind, sd, (uint32_t)problem, n, dim,
                       depth, index, (double*)out, aux);
  else
    return fail(HIPMC_E_INVALID, "bad dtype");
  return hip_check(hipGetLastError(), "draws_kernel launch");
}

}  // extern "C"
