"""probability_amd: MI355X-native tfp.mcmc gradient-based sampling.

`probability_amd.mcmc` mirrors `tfp.mcmc` (HamiltonianMonteCarlo,
NoUTurnSampler, TransformedTransitionKernel,
DualAveragingStepSizeAdaptation, sample_chain, diagnostics) and
`probability_amd.distributions` compiles TFP-style JointDistribution models
to the device targets, `probability_amd.experimental` holds the streaming
reducers of
`tfp.experimental.mcmc` / `tfp.experimental.stats`, on top of libhipmc, the
HIP kernels behind the C ABI in include/hipmc.h.
"""

from probability_amd import bijectors
from probability_amd import distributions
from probability_amd import experimental
from probability_amd import mcmc
from probability_amd import random
from probability_amd import targets

__version__ = '0.1.0'
