"""Runner for the MFMA large-batch HMC path (hipmc_glm_hmc_run).

Owns the caller-side workspace the library needs (split-N partial
gradients, proposal buffers, BF16 hi/lo positions) for one target and chain
count, and exposes bootstrap (log-prob + gradient) and HMC iterations.
"""

import numpy as np
import torch

from probability_amd import _lib


def default_splits(n_obs, n_chains):
  """Row splits so that the gradient launch has >= 512 workgroups."""
  tiles = n_chains // 64
  want = max(1, -(-512 // max(tiles, 1)))
  rows_tiles = n_obs // 32
  s = min(want, rows_tiles)
  # no empty split: rows per split (in tiles) * splits must cover n_obs
  while s > 1:
    rps = -(-rows_tiles // s)
    if -(-rows_tiles // rps) == s:
      break
    s -= 1
  return max(s, 1)


class GlmHmcRunner:

  def __init__(self, spec, n_chains, n_split=None):
    if not spec.mfma_ok(n_chains):
      raise ValueError('target/chain count not supported by the MFMA path')
    self.spec = spec
    dev = spec.device
    c, d = n_chains, spec.dim
    self.n_chains = c
    self.n_split = n_split or default_splits(spec.n_obs, c)
    f32 = torch.float32
    self.ws_g = torch.empty((self.n_split, c, d), dtype=f32, device=dev)
    self.ws_ll = torch.empty((self.n_split, c), dtype=f32, device=dev)
    self.ws_q = torch.empty((c, d), dtype=f32, device=dev)
    self.ws_p = torch.empty((c, d), dtype=f32, device=dev)
    self.ws_b = torch.empty((2, c, d), dtype=torch.bfloat16, device=dev)
    self.ws_ke0 = torch.empty(c, dtype=torch.float64, device=dev)

  def _args(self, state, tlp, grad, step):
    s = self.spec
    a = _lib.GlmHmcArgs()
    a.n_obs, a.dim, a.n_chains, a.n_split = s.n_obs, s.dim, self.n_chains, \
        self.n_split
    a.prior_scale = s.prior_scale
    a.x_bf16, a.xt_bf16, a.y = (s.x_bf16.data_ptr(), s.xt_bf16.data_ptr(),
                                s.y.data_ptr())
    if step is not None:
      a.step_size = step.data_ptr()
      a.per_chain_step = 1 if step.numel() == self.n_chains and \
          self.n_chains > 1 else 0
    a.state, a.tlp, a.grad = state.data_ptr(), tlp.data_ptr(), grad.data_ptr()
    a.ws_g, a.ws_ll = self.ws_g.data_ptr(), self.ws_ll.data_ptr()
    a.ws_q, a.ws_p = self.ws_q.data_ptr(), self.ws_p.data_ptr()
    a.ws_b, a.ws_ke0 = self.ws_b.data_ptr(), self.ws_ke0.data_ptr()
    return a

  def bootstrap(self, state):
    """(tlp [C], grad [C, D]) at `state` [C, D] float32."""
    state = state.contiguous()
    tlp = torch.empty(self.n_chains, dtype=torch.float32, device=state.device)
    grad = torch.empty_like(state)
    dummy = torch.zeros(1, dtype=torch.float32, device=state.device)
    a = self._args(state, tlp, grad, dummy)
    a.want_grad_only = 1
    _lib.check(_lib.lib().hipmc_glm_hmc_run(
        a, _lib.stream_handle(state.device)))
    return tlp, grad

  def run(self, state, tlp, grad, step, seed, num_leapfrog_steps, n_iters=1,
          fold_seed=False, iter0=0, key=0, out_lar=None, out_acc=None,
          trace=None, trace_start=0, trace_every=1, n_trace=0):
    a = self._args(state, tlp, grad, step)
    a.num_leapfrog_steps = int(num_leapfrog_steps)
    a.n_iters = int(n_iters)
    a.fold_seed = 1 if fold_seed else 0
    a.iter0 = int(iter0)
    a.key = int(key)
    seed = np.asarray(seed, np.uint32).reshape(2)
    a.seed[0], a.seed[1] = int(seed[0]), int(seed[1])
    if out_lar is not None:
      a.out_log_accept_ratio = out_lar.data_ptr()
    if out_acc is not None:
      a.out_is_accepted = out_acc.data_ptr()
    if trace is not None:
      a.trace_state = trace.data_ptr()
      a.trace_start, a.trace_every, a.n_trace = int(trace_start), \
          int(trace_every), int(n_trace)
    _lib.check(_lib.lib().hipmc_glm_hmc_run(
        a, _lib.stream_handle(state.device)))
