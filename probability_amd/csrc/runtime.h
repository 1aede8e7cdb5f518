// Internal declarations shared by the library's translation units: the
// thread-local error message, HIP error checks, and the launchers of the
// sampling kernels.  Each (dtype, target family) pair of the sampling
// kernels is compiled in its own object (inst.hip), so the launchers are
// declared here and defined by explicit instantiation there.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "../../include/hipmc.h"

namespace hipmc {

// Sets the thread-local message returned by hipmc_last_error; returns code.
int fail(int code, const std::string& msg);
// HIPMC_OK, or HIPMC_E_HIP with the runtime's message.
int hip_check(hipError_t e, const char* what);

inline hipStream_t as_stream(uint64_t s) {
  return reinterpret_cast<hipStream_t>(s);
}

constexpr int kMaxLdsPerBlock = 64 * 1024;
// gfx950 gives one workgroup up to 160 KiB of LDS; a launch that asks for
// more than 64 KiB raises the kernel's dynamic-LDS limit first.
constexpr int kMaxLdsSharedChain = 160 * 1024;

// Register slots per lane (state elements / 64) built for a target family:
// up to 8 for GLMs, whose chains keep a second copy of their vectors parked
// in LDS, and up to 17 (1088 elements) for the others.
constexpr int max_slots(int fam) { return fam == 1 ? 8 : 17; }
// Smallest compiled slot count K of family fam with K*64 >= max_dim, or -1.
int pick_k(int max_dim, int fam);
// The largest state dimension built for family fam.
int max_state_dim(int fam);

// Sampling-kernel launchers of one (dtype T, family FAM) pair; k is the
// register slot count from pick_k.
template <typename T, int FAM>
int launch_nuts_k(int k, const hipmc_run_args& a, hipStream_t s);
template <typename T, int FAM>
int launch_hmc_k(int k, bool calibrated, const hipmc_run_args& a,
                 hipStream_t s);
template <typename T, int FAM>
int launch_eval_k(int k, const hipmc_problem* p, const int32_t* cp, int n,
                  int stride, const void* x, void* tlp, void* grad,
                  hipStream_t s);

// The BF16 MFMA logistic-regression HMC path (mfma.hip), with caller
// workspaces or with the library's static scratch.
int launch_glm_hmc(const hipmc_glm_hmc_args* a, hipStream_t s);
int launch_glm_hmc_noalloc(const hipmc_glm_hmc_args* a, hipStream_t s);

}  // namespace hipmc
