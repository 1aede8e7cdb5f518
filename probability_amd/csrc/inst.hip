// Explicit instantiation of the sampling-kernel launchers for one
// (dtype, target family) pair, so that each pair compiles as its own object
// (the build compiles this file once per pair).  HIPMC_INST_T is float or
// double; HIPMC_INST_FAM a target family; HIPMC_INST_NUTS selects the NUTS
// launcher (1) or the HMC, leapfrog and evaluation launchers (0).
#include "launch.h"

#if !defined(HIPMC_INST_T) || !defined(HIPMC_INST_FAM) || \
    !defined(HIPMC_INST_NUTS)
#error "inst.hip needs HIPMC_INST_T, HIPMC_INST_FAM and HIPMC_INST_NUTS"
#endif

namespace hipmc {
#if HIPMC_INST_NUTS
template int launch_nuts_k<HIPMC_INST_T, HIPMC_INST_FAM>(
    int, const hipmc_run_args&, hipStream_t);
#else
template int launch_hmc_k<HIPMC_INST_T, HIPMC_INST_FAM>(
    int, bool, const hipmc_run_args&, hipStream_t);
template int launch_eval_k<HIPMC_INST_T, HIPMC_INST_FAM>(
    int, const hipmc_problem*, const int32_t*, int, int, const void*, void*,
    void*, hipStream_t);
#endif
}  // namespace hipmc
