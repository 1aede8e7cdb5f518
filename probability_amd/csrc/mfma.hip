// Launch sequence of the BF16 MFMA logistic-regression HMC path
// (hipmc_glm_hmc_run): per leapfrog step one glm_grad_mfma launch (the
// all-chain gradient as fused GEMMs) and one mfma_chain_step launch (split
// reduction, prior, leapfrog / MH arithmetic).
#include <hip/hip_runtime.h>

#include "../../include/hipmc.h"
#include "glm_mfma.h"
#include "runtime.h"

namespace hipmc {

namespace {
int blocks_for(size_t n, int bs) { return (int)((n + bs - 1) / bs); }
}  // namespace

int launch_glm_hmc(const hipmc_glm_hmc_args* a, hipStream_t s) {
  if (!a) return fail(HIPMC_E_INVALID, "glm_hmc: args is NULL");
  if (a->dim != kMD)
    return fail(HIPMC_E_UNSUPPORTED, "glm_hmc: dim must be 256");
  if (a->n_obs <= 0 || a->n_obs % kMRows)
    return fail(HIPMC_E_INVALID, "glm_hmc: n_obs must be a positive multiple of 32");
  if (a->n_chains <= 0 || a->n_chains % kMChains)
    return fail(HIPMC_E_INVALID, "glm_hmc: n_chains must be a positive multiple of 64");
  if (a->n_split < 1) return fail(HIPMC_E_INVALID, "glm_hmc: n_split must be >= 1");
  if (!a->x_bf16 || !a->xt_bf16 || !a->y || !a->step_size || !a->state ||
      !a->tlp || !a->grad || !a->ws_g || !a->ws_ll || !a->ws_q || !a->ws_p ||
      !a->ws_b || !a->ws_ke0)
    return fail(HIPMC_E_INVALID, "glm_hmc: NULL argument");
  if (!a->want_grad_only) {
    if (a->num_leapfrog_steps < 1)
      return fail(HIPMC_E_INVALID, "glm_hmc: num_leapfrog_steps must be >= 1");
    if (a->n_iters < 1) return fail(HIPMC_E_INVALID, "glm_hmc: n_iters must be >= 1");
    if (!a->fold_seed && a->n_iters != 1)
      return fail(HIPMC_E_INVALID, "glm_hmc: n_iters > 1 requires fold_seed");
  }
  // rows per split: a multiple of the tile height
  int rps = (a->n_obs + a->n_split - 1) / a->n_split;
  rps = (rps + kMRows - 1) / kMRows * kMRows;
  const int n_split = (a->n_obs + rps - 1) / rps;
  if (n_split != a->n_split)
    return fail(HIPMC_E_INVALID, "glm_hmc: n_split leaves an empty split; use fewer");
  const dim3 ggrid(a->n_chains / kMChains, n_split);
  const dim3 cgrid((a->n_chains + 3) / 4);
  auto grad = [&](bool ll) {
    hipLaunchKernelGGL(glm_grad_mfma, ggrid, dim3(256), 0, s,
                       (const __bf16*)a->x_bf16, (const __bf16*)a->xt_bf16,
                       a->y, a->n_obs, (const __bf16*)a->ws_b,
                       (const __bf16*)a->ws_b + (size_t)a->n_chains * kMD,
                       a->n_chains, rps, a->ws_g, ll ? a->ws_ll : nullptr);
  };
  if (a->want_grad_only) {
    const size_t ne = (size_t)a->n_chains * kMD;
    hipLaunchKernelGGL(mfma_prepare, dim3(blocks_for(ne, 256)), dim3(256), 0,
                       s, *a);
    grad(true);
    hipLaunchKernelGGL(mfma_chain_step, cgrid, dim3(256), 0, s, *a, 3, 0);
    return hip_check(hipGetLastError(), "glm_hmc bootstrap launch");
  }
  for (int it = 0; it < a->n_iters; ++it) {
    hipLaunchKernelGGL(mfma_chain_step, cgrid, dim3(256), 0, s, *a, 0, it);
    for (int st = 0; st < a->num_leapfrog_steps; ++st) {
      const bool last = st + 1 == a->num_leapfrog_steps;
      grad(last);
      hipLaunchKernelGGL(mfma_chain_step, cgrid, dim3(256), 0, s, *a,
                         last ? 2 : 1, it);
    }
  }
  return hip_check(hipGetLastError(), "glm_hmc launch");
}

}  // namespace hipmc
