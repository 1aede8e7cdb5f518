// Launch configuration of the sampling kernels (nuts_kernel, hmc_kernel,
// eval_kernel) for one (dtype, family) pair; included by inst.hip only.
#pragma once

#include "common.h"
#include "kernels.h"
#include "runtime.h"
#include "targets.h"

namespace hipmc {

template <typename F>
int allow_lds(F* kernel, size_t bytes) {
  if (bytes > (size_t)kMaxLdsSharedChain)
    return fail(HIPMC_E_UNSUPPORTED, "shared-chain LDS does not fit");
  if (bytes <= (size_t)kMaxLdsPerBlock) return HIPMC_OK;
  return hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bytes),
                   "hipFuncSetAttribute(max dynamic LDS)");
}

inline int waves_for_lds(size_t per_wave_bytes) {
  int w = 4;
  while (w > 1 && (size_t)w * per_wave_bytes > (size_t)kMaxLdsPerBlock) --w;
  return w;
}

// Chain groups of G chains on W waves; HIPMC_E_UNSUPPORTED when the LDS does
// not fit (the caller then tries a smaller group).
template <typename T, int K, int G>
int launch_nuts_group(const hipmc_run_args& a, int W, hipStream_t s) {
  const size_t bytes = nuts_group_lds_bytes<T, K>(a.max_tree_depth, G, W);
  if (bytes > (size_t)kMaxLdsSharedChain)
    return fail(HIPMC_E_UNSUPPORTED, "chain group LDS does not fit");
  const int rc = allow_lds(&nuts_group_kernel<T, K, G>, bytes);
  if (rc != HIPMC_OK) return rc;
  const int blocks = (a.n_chains + G - 1) / G;
  hipLaunchKernelGGL((nuts_group_kernel<T, K, G>), dim3(blocks), dim3(64 * W),
                     bytes, s, a);
  return hip_check(hipGetLastError(), "nuts_group_kernel launch");
}

// Grouped launch of a GLM NUTS run when a.chains_per_block asks for it
// (slot counts up to 3, one leapfrog per leaf): the largest group of at most
// chains_per_block (4 or 2) chains, no more than the waves, whose LDS fits.
// Returns 1 when nothing was launched (the caller runs the per-chain
// kernel).
template <typename T, int K, int FAM>
int try_launch_nuts_group(const hipmc_run_args& a, hipStream_t s) {
  if constexpr (FAM != kFamilyGlm || K > 3) {
    return 1;
  } else {
    if (a.chains_per_block < 2 || a.unrolled_leapfrog_steps != 1) return 1;
    const int W = waves_per_chain<FAM>(a);
    if (a.chains_per_block >= 4 && W >= 4 &&
        nuts_group_lds_bytes<T, K>(a.max_tree_depth, 4, W) <=
            (size_t)kMaxLdsSharedChain)
      return launch_nuts_group<T, K, 4>(a, W, s);
    if (W >= 2 && nuts_group_lds_bytes<T, K>(a.max_tree_depth, 2, W) <=
                      (size_t)kMaxLdsSharedChain)
      return launch_nuts_group<T, K, 2>(a, W, s);
    return 1;
  }
}

template <typename T, int K, int FAM>
int launch_nuts(const hipmc_run_args& a0, hipStream_t s) {
  if (a0.n_chains == 0) return HIPMC_OK;
  {
    const int rc = try_launch_nuts_group<T, K, FAM>(a0, s);
    if (rc != 1) return rc;
  }
  // one chain per block: the chain's LDS (U-turn memory, parked tree
  // vectors), W wave blocks (target scratch, parked leaf vectors) and the
  // two halves of the partial-sum buffer; fewer waves share the chain when
  // that does not fit
  const size_t cl = nuts_chain_lds_elems<T, K, FAM>(a0.max_tree_depth);
  const size_t wl = nuts_wave_lds_elems<T, K, FAM>();
  auto shared_bytes = [&](int w) {
    return (cl + (size_t)w * wl) * sizeof(T) + 2 * glm_red_bytes<T, K>(w);
  };
  hipmc_run_args a = a0;
  int W = waves_per_chain<FAM>(a);
  while (W > 1 && shared_bytes(W) > (size_t)kMaxLdsSharedChain) --W;
  a.waves_per_chain = W;
  if (W > 1) {
    const size_t bytes = shared_bytes(W);
    const int rc = allow_lds(&nuts_kernel<T, K, FAM>, bytes);
    if (rc != HIPMC_OK) return rc;
    hipLaunchKernelGGL((nuts_kernel<T, K, FAM>), dim3(a.n_chains),
                       dim3(64 * W), bytes, s, a);
    return hip_check(hipGetLastError(), "nuts_kernel launch");
  }
  const size_t per_wave = (cl + wl) * sizeof(T);
  if (per_wave > (size_t)kMaxLdsPerBlock) {
    // wide states at deep trees: one chain per block with up to the
    // workgroup's whole LDS
    if (per_wave > (size_t)kMaxLdsSharedChain)
      return fail(HIPMC_E_UNSUPPORTED,
                  "U-turn memory does not fit in LDS (lower max_tree_depth "
                  "or use float32)");
    const int rc = allow_lds(&nuts_kernel<T, K, FAM>, per_wave);
    if (rc != HIPMC_OK) return rc;
    hipLaunchKernelGGL((nuts_kernel<T, K, FAM>), dim3(a.n_chains), dim3(64),
                       per_wave, s, a);
    return hip_check(hipGetLastError(), "nuts_kernel launch");
  }
  const int w = waves_for_lds(per_wave);
  const int blocks = (a.n_chains + w - 1) / w;
  hipLaunchKernelGGL((nuts_kernel<T, K, FAM>), dim3(blocks), dim3(64 * w),
                     per_wave * w, s, a);
  return hip_check(hipGetLastError(), "nuts_kernel launch");
}

template <typename T, int K, int FAM, bool CAL>
int launch_hmc(const hipmc_run_args& a0, hipStream_t s) {
  if (a0.n_chains == 0) return HIPMC_OK;
  const size_t per_wave = hmc_lds_elems<T, K, FAM>() * sizeof(T);
  hipmc_run_args a = a0;
  // hmc_kernel is built for blocks of at most 256 threads (it may use more
  // than 256 registers per lane, one wave per SIMD), so a shared chain
  // spans at most four waves
  int W = waves_per_chain<FAM>(a);
  if (W > 4) W = 4;
  while (W > 1 && per_wave * W + 2 * glm_red_bytes<T, K>(W) >
                      (size_t)kMaxLdsSharedChain)
    --W;
  a.waves_per_chain = W;
  if (W > 1) {
    const size_t bytes = per_wave * W + 2 * glm_red_bytes<T, K>(W);
    const int rc = allow_lds(&hmc_kernel<T, K, FAM, CAL>, bytes);
    if (rc != HIPMC_OK) return rc;
    hipLaunchKernelGGL((hmc_kernel<T, K, FAM, CAL>), dim3(a.n_chains),
                       dim3(64 * W), bytes, s, a);
    return hip_check(hipGetLastError(), "hmc_kernel launch");
  }
  const int w = 4;
  const int blocks = (a.n_chains + w - 1) / w;
  hipLaunchKernelGGL((hmc_kernel<T, K, FAM, CAL>), dim3(blocks), dim3(64 * w),
                     per_wave * w, s, a);
  return hip_check(hipGetLastError(), "hmc_kernel launch");
}

template <typename T, int K, int FAM>
int launch_eval(const hipmc_problem* p, const int32_t* cp, int n, int stride,
                const void* x, void* tlp, void* grad, hipStream_t s) {
  const size_t per_wave = hmc_lds_elems<T, K, FAM>() * sizeof(T);
  const int w = 4;
  const int blocks = (n + w - 1) / w;
  if (blocks == 0) return HIPMC_OK;
  hipLaunchKernelGGL((eval_kernel<T, K, FAM>), dim3(blocks), dim3(64 * w),
                     per_wave * w, s, p, cp, n, stride, (const T*)x,
                     (T*)tlp, (T*)grad);
  return hip_check(hipGetLastError(), "eval_kernel launch");
}

// The slot counts built for every (dtype, family) pair, and the wide ones
// built for the families other than GLMs; pick_k (hipmc.hip) chooses the
// smallest that holds the state.
#define HIPMC_FOR_EACH_K(X) X(1) X(2) X(3) X(4) X(5) X(8)
#define HIPMC_FOR_EACH_WIDE_K(X) X(12) X(17)
static_assert(max_slots(kFamilyGlm) == 8 && max_slots(kFamilyHier) == 17 &&
                  max_slots(kFamilyTest) == 17,
              "slot lists and max_slots disagree");

template <typename T, int FAM>
int launch_nuts_k(int k, const hipmc_run_args& a, hipStream_t s) {
  switch (k) {
#define HIPMC_CASE(KK) \
  case KK: return launch_nuts<T, KK, FAM>(a, s);
    HIPMC_FOR_EACH_K(HIPMC_CASE)
  }
  if constexpr (FAM != kFamilyGlm) {
    switch (k) { HIPMC_FOR_EACH_WIDE_K(HIPMC_CASE) }
  }
#undef HIPMC_CASE
  return fail(HIPMC_E_UNSUPPORTED, "no kernel for this dimension");
}

template <typename T, int FAM>
int launch_hmc_k(int k, bool calibrated, const hipmc_run_args& a,
                 hipStream_t s) {
  switch (k) {
#define HIPMC_CASE(KK)                                        \
  case KK:                                                    \
    return calibrated ? launch_hmc<T, KK, FAM, true>(a, s)    \
                      : launch_hmc<T, KK, FAM, false>(a, s);
    HIPMC_FOR_EACH_K(HIPMC_CASE)
  }
  if constexpr (FAM != kFamilyGlm) {
    switch (k) { HIPMC_FOR_EACH_WIDE_K(HIPMC_CASE) }
  }
#undef HIPMC_CASE
  return fail(HIPMC_E_UNSUPPORTED, "no kernel for this dimension");
}

template <typename T, int FAM>
int launch_eval_k(int k, const hipmc_problem* p, const int32_t* cp, int n,
                  int stride, const void* x, void* tlp, void* grad,
                  hipStream_t s) {
  switch (k) {
#define HIPMC_CASE(KK) \
  case KK: return launch_eval<T, KK, FAM>(p, cp, n, stride, x, tlp, grad, s);
    HIPMC_FOR_EACH_K(HIPMC_CASE)
  }
  if constexpr (FAM != kFamilyGlm) {
    switch (k) { HIPMC_FOR_EACH_WIDE_K(HIPMC_CASE) }
  }
#undef HIPMC_CASE
  return fail(HIPMC_E_UNSUPPORTED, "no kernel for this dimension");
}

}  // namespace hipmc
