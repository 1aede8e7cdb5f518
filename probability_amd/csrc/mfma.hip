// Launch sequence of the BF16 MFMA logistic-regression HMC path
// (hipmc_glm_hmc_run): per leapfrog step one glm_grad_mfma launch (the
// all-chain gradient as fused GEMMs) and one mfma_chain_step launch (split
// reduction, prior, leapfrog / MH arithmetic).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/hipmc.h"
#include "glm_mfma.h"
#include "runtime.h"

// Workspace of hipmc_glm_hmc_run_noalloc: one array per GPU the library runs
// on (each device's copy of the code object holds its own), kept until the
// process ends.  Host code must not use this name as an address: the device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[HIPMC_NOALLOC_SCRATCH_BYTES];

namespace hipmc {

namespace {
int blocks_for(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) on its first call.
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc != HIPMC_OK) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return fail(HIPMC_E_UNSUPPORTED, "glm_hmc_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                   "hipGetSymbolAddress(scratch)");
    if (rc != HIPMC_OK) return rc;
    g_scratch[dev] = p;
  }
  *out = (unsigned char*)g_scratch[dev];
  return HIPMC_OK;
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }
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
  // 128 chains per workgroup when the chain count allows (each staged tile
  // then serves 8 waves), else 64
  const bool wide = a->n_chains % (2 * kMChains) == 0;
  const dim3 ggrid(a->n_chains / (wide ? 2 * kMChains : kMChains), n_split);
  const dim3 cgrid((a->n_chains + 3) / 4);
  auto grad = [&](bool ll) {
    const __bf16* bh = (const __bf16*)a->ws_b;
    const __bf16* bl = bh + (size_t)a->n_chains * kMD;
    float* oll = ll ? a->ws_ll : nullptr;
    if (wide)
      hipLaunchKernelGGL(glm_grad_mfma<8>, ggrid, dim3(512), 0, s,
                         (const __bf16*)a->x_bf16, (const __bf16*)a->xt_bf16,
                         a->y, a->n_obs, bh, bl, a->n_chains, rps, a->ws_g,
                         oll);
    else
      hipLaunchKernelGGL(glm_grad_mfma<4>, ggrid, dim3(256), 0, s,
                         (const __bf16*)a->x_bf16, (const __bf16*)a->xt_bf16,
                         a->y, a->n_obs, bh, bl, a->n_chains, rps, a->ws_g,
                         oll);
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

int launch_glm_hmc_noalloc(const hipmc_glm_hmc_args* a, hipStream_t s) {
  if (!a) return fail(HIPMC_E_INVALID, "glm_hmc_noalloc: args is NULL");
  if (a->n_chains <= 0 || a->n_split < 1 || a->dim != kMD)
    return launch_glm_hmc(a, s);  // the shared argument checks report it
  // workspace layout (each piece 256-byte aligned): ws_g [n_split, C, D]
  // f32, ws_ll [n_split, C] f32, ws_q and ws_p [C, D] f32, ws_b [2, C, D]
  // bf16, ws_ke0 [C] f64
  const size_t c = (size_t)a->n_chains, d = (size_t)kMD;
  const size_t sizes[6] = {(size_t)a->n_split * c * d * 4,
                           (size_t)a->n_split * c * 4, c * d * 4, c * d * 4,
                           2 * c * d * 2, c * 8};
  size_t need = 0;
  for (size_t v : sizes) need += round256(v);
  if (need > (size_t)HIPMC_NOALLOC_SCRATCH_BYTES)
    return fail(HIPMC_E_UNSUPPORTED,
                "glm_hmc_noalloc: n_chains x n_split needs " +
                    std::to_string(need) + " bytes of workspace, more than "
                    "the library's " +
                    std::to_string((size_t)HIPMC_NOALLOC_SCRATCH_BYTES) +
                    "-byte static scratch");
  unsigned char* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc != HIPMC_OK) return rc;
  hipmc_glm_hmc_args b = *a;
  size_t off = 0;
  void* ws[6];
  for (int i = 0; i < 6; ++i) {
    ws[i] = base + off;
    off += round256(sizes[i]);
  }
  b.ws_g = (float*)ws[0];
  b.ws_ll = (float*)ws[1];
  b.ws_q = (float*)ws[2];
  b.ws_p = (float*)ws[3];
  b.ws_b = ws[4];
  b.ws_ke0 = (double*)ws[5];
  return launch_glm_hmc(&b, s);
}

}  // namespace hipmc
