"""Launch helpers: build `hipmc_run_args` for a target spec and call libhipmc.

Single-problem launches use chain_problem = chain_local = NULL, so chain c's
draws are keyed (problem 0, chain c).
"""

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import tuning

_OUT_FIELDS = ('out_log_accept_ratio', 'out_leapfrogs', 'out_is_accepted',
               'out_reach_max_depth', 'out_has_divergence', 'out_energy',
               'out_tree_depth', 'trace_state', 'out_proposed_state',
               'out_proposed_tlp', 'out_initial_momentum',
               'out_final_momentum')


def step_tensor(step_size, dtype, device):
  """Scalar (or per-chain) step size as a device tensor of `dtype`."""
  if isinstance(step_size, (list, tuple)):
    vals = [torch.as_tensor(s) for s in step_size]
    first = vals[0]
    for v in vals[1:]:
      if not torch.equal(v.detach().cpu().double(), first.detach().cpu().double()):
        raise NotImplementedError(
            'per-part step sizes that differ are not built into the kernels')
    step_size = first
  t = torch.as_tensor(step_size, dtype=dtype)
  return t.to(device).reshape(-1).contiguous()


class Run:
  """Owns the output buffers of one kernel launch."""

  def __init__(self, spec, table, state, tlp, grad, step, seed, n_iters=1,
               fold_seed=False, iter0=0, sampler='nuts'):
    _lib.require_cuda()
    self.spec = spec
    self.table = table
    self.state, self.tlp, self.grad = state, tlp, grad
    c, d = state.shape
    self.n_chains, self.dim = c, d
    self.step = step
    a = _lib.RunArgs()
    a.dtype = _lib.dtype_code(state.dtype)
    a.n_chains = c
    a.dim_stride = d
    a.max_dim = spec.dim
    a.n_iters = n_iters
    a.fold_seed = 1 if fold_seed else 0
    a.iter0 = int(iter0)
    seed = np.asarray(seed, np.uint32).reshape(2)
    a.seed[0], a.seed[1] = int(seed[0]), int(seed[1])
    a.target_kind = spec.kind
    if spec.kind in (_lib.TARGET_LOGISTIC, _lib.TARGET_LINEAR):
      base = getattr(spec, '_base', spec)
      # sampler: 'nuts' or 'hmc' (the HMC and MALA kernels)
      a.waves_per_chain = tuning.glm_waves(spec.kind, base.n_obs, base.dim,
                                           c, sampler, state.dtype)
    a.max_energy_diff = 1000.0
    a.problems = table.data_ptr()
    a.n_problems = 1
    a.step_size = step.data_ptr()
    if step.numel() == c and c > 1:
      self._chain_step = torch.arange(c, dtype=torch.int32, device=state.device)
      a.chain_step = self._chain_step.data_ptr()
    elif step.numel() != 1:
      raise ValueError(f'step size with {step.numel()} elements does not match '
                       f'{c} chains')
    a.state, a.tlp, a.grad = state.data_ptr(), tlp.data_ptr(), grad.data_ptr()
    self.args = a
    self.outputs = {}

  def want(self, name, shape, dtype):
    t = torch.empty(shape, dtype=dtype, device=self.state.device)
    self.outputs[name] = t
    setattr(self.args, name, t.data_ptr())
    return t

  def call(self, fn):
    _lib.check(fn(self.args, _lib.stream_handle(self.state.device)))
    return self.outputs


def flat_inputs(spec, parts, tlp, grads):
  """Packs state/grad parts; returns contiguous working copies."""
  state = spec.pack_state(parts).clone()
  grad = spec.pack_state(grads).clone()
  t = tlp.to(device=spec.device, dtype=spec.dtype).reshape(-1).clone()
  return state, t, grad
