"""Many independent posteriors sampled together in one launch.

TFP runs one posterior per graph: independent models need separate
`sample_chain` calls (or a hand-built batch dimension).  `PosteriorBatch`
packs hundreds of target specs of different kinds, data sizes and state
dimensions into one problem table; `sample_posteriors` runs NUTS (or HMC)
with per-problem dual averaging on all of their chains in the same launches
(one launch per adaptation step for all problems, one launch for the
sampling phase).  Every chain's draws are keyed by (seed, problem key,
chain), so a problem gives identical results alone or in any batch.

Per-problem semantics are those of sample_chain(DualAveraging(TTK(NUTS)))
with scalar step size per problem (`mcmc/sample.py`, `mcmc/nuts.py`,
`mcmc/dual_averaging_step_size_adaptation.py`).
"""

import collections

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import bijectors as bij_lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib

__all__ = ['PosteriorBatch', 'sample_posteriors', 'BatchResults']

BatchResults = collections.namedtuple('BatchResults', [
    'draws', 'step_size', 'log_accept_ratio', 'leapfrogs_taken',
    'is_accepted', 'has_divergence', 'final_state'])


def _family(kind):
  if kind == _lib.TARGET_HIER_NORMAL:
    return 0
  if kind in (_lib.TARGET_LOGISTIC, _lib.TARGET_LINEAR):
    return 1
  return 2


class PosteriorBatch:
  """A set of target specs ("problems") laid out for one launch.

  Args:
    specs: list of `probability_amd.targets` specs (one family: GLMs, or
      hierarchical normals, or the test targets); same dtype and device.
    num_chains: chains per problem (int or list).
    bijectors: per-problem list of per-part bijectors (None: identity).
    keys: per-problem Philox key (default: 0..P-1).  A problem with the same
      key, seed and chain count produces the same draws in any batch.
    slot_dim: the kernels hold a chain's state in ceil(dim/64) registers
      per lane, chosen from the largest dim of the launch; passing the same
      slot_dim to every batch a problem may run in makes it run the same
      code (bit-identical results) alone or in any mix.
    waves_per_chain: GLMs only, waves sharing one chain's gradient (1-4;
      None: chosen from the largest design matrix).  Like slot_dim, equal
      values give bit-identical results alone or in any mix.
  """

  def __init__(self, specs, num_chains, bijectors=None, keys=None,
               slot_dim=None, waves_per_chain=None):
    if not specs:
      raise ValueError('need at least one problem')
    self.specs = [targets_lib.as_target(s) for s in specs]
    s0 = self.specs[0]
    self.dtype, self.device = s0.dtype, s0.device
    fams = {_family(s.kind) for s in self.specs}
    if len(fams) != 1:
      raise ValueError('problems of one launch must share a target family')
    for s in self.specs:
      if s.dtype != self.dtype or s.device != self.device:
        raise ValueError('problems must share dtype and device')
    p = len(self.specs)
    self.num_chains = ([int(num_chains)] * p if np.isscalar(num_chains)
                       else [int(c) for c in num_chains])
    if len(self.num_chains) != p:
      raise ValueError('num_chains must be an int or one value per problem')
    self.bijectors = bijectors or [None] * p
    self.keys = list(range(p)) if keys is None else [int(k) for k in keys]
    self.dims = [s.dim for s in self.specs]
    self.max_dim = max(self.dims + [int(slot_dim or 0)])
    self.dim_stride = self.max_dim
    offs = np.concatenate([[0], np.cumsum(self.num_chains)]).astype(np.int64)
    self.chain_offsets = offs
    self.n_chains = int(offs[-1])
    raw = bytearray()
    for i, s in enumerate(self.specs):
      raw += bytes(s.problem(self.bijectors[i], n_chains=self.num_chains[i],
                             chain_offset=int(offs[i]), key=self.keys[i]))
    self.table = torch.frombuffer(raw, dtype=torch.uint8).to(self.device)
    cp = np.repeat(np.arange(p, dtype=np.int32), self.num_chains)
    cl = np.concatenate([np.arange(n, dtype=np.int32) for n in self.num_chains])
    self.chain_problem = torch.as_tensor(cp, device=self.device)
    self.chain_local = torch.as_tensor(cl, device=self.device)
    self.group_start = torch.as_tensor(offs.astype(np.int32), device=self.device)
    self.kind = self.specs[0].kind
    if _family(self.kind) == 2 and len({s.kind for s in self.specs}) > 1:
      self.kind = _lib.TARGET_DIAG_NORMAL
    if _family(self.kind) != 1:
      self.waves_per_chain = 1
    elif waves_per_chain is not None:
      self.waves_per_chain = int(waves_per_chain)
    else:
      self.waves_per_chain = _lib.glm_waves_per_chain(
          max(s.n_obs * s.dim for s in self.specs))

  @property
  def n_problems(self):
    return len(self.specs)

  def eval(self, x):
    """(tlp [C], grad [C, dim_stride]) of every chain at x."""
    tlp = torch.empty(self.n_chains, dtype=self.dtype, device=self.device)
    grad = torch.zeros_like(x)
    _lib.check(_lib.lib().hipmc_target_eval(
        _lib.dtype_code(self.dtype), self.kind, _lib.ptr(self.table),
        _lib.ptr(self.chain_problem), self.n_chains, self.dim_stride,
        self.max_dim, _lib.ptr(x), _lib.ptr(tlp), _lib.ptr(grad),
        _lib.stream_handle(self.device)))
    return tlp, grad

  def bijector(self, x, direction, chain_period=0):
    out = torch.empty_like(x)
    rows = x.numel() // self.dim_stride
    _lib.check(_lib.lib().hipmc_bijector(
        _lib.dtype_code(self.dtype), _lib.ptr(self.table),
        _lib.ptr(self.chain_problem), rows, self.dim_stride, chain_period,
        direction, _lib.ptr(x), _lib.ptr(out),
        _lib.stream_handle(self.device)))
    return out

  def split(self, flat):
    """Per-problem views [..., C_p, D_p] of a [..., C, dim_stride] array."""
    out = []
    for i in range(self.n_problems):
      a, b = int(self.chain_offsets[i]), int(self.chain_offsets[i + 1])
      out.append(flat[..., a:b, :self.dims[i]])
    return out

  def split_chains(self, arr):
    """Per-problem views [..., C_p] of a [..., C] array."""
    return [arr[..., int(self.chain_offsets[i]):int(self.chain_offsets[i + 1])]
            for i in range(self.n_problems)]


def sample_posteriors(batch, num_results, num_burnin_steps=0,
                      num_adaptation_steps=0, init_state=None, step_size=0.1,
                      target_accept_prob=0.75, max_tree_depth=10,
                      max_energy_diff=1000.0, kernel='nuts',
                      num_leapfrog_steps=10, seed=None, trace=True,
                      stats=None, max_iters_per_launch=4096, progress=None):
  """Runs every problem of `batch` for num_burnin_steps + num_results
  transitions (dual averaging over the first num_adaptation_steps) and
  returns per-problem draws in the constrained space.

  init_state: [C, dim_stride] constrained initial states (None: unconstrained
  zeros).  step_size: scalar or one value per problem.  progress, if given,
  is called as progress(transitions_enqueued, total) after each launch.
  """
  if kernel not in ('nuts', 'hmc'):
    raise ValueError("kernel must be 'nuts' or 'hmc'")
  _lib.require_cuda()
  dev, dt = batch.device, batch.dtype
  c, d, p = batch.n_chains, batch.dim_stride, batch.n_problems
  if init_state is None:
    x = torch.zeros((c, d), dtype=dt, device=dev)
  else:
    x = batch.bijector(torch.as_tensor(init_state, dtype=dt, device=dev)
                       .contiguous(), 1)
  tlp, grad = batch.eval(x)
  step = torch.as_tensor(np.broadcast_to(np.asarray(step_size, np.float64),
                                         (p,)).copy(), device=dev).to(dt)
  step64 = step.to(torch.float64).clone()
  err = torch.zeros(p, dtype=torch.float64, device=dev)
  avg = torch.zeros_like(err)
  lst = np.log(10.0) + torch.log(step64)
  tgt = torch.as_tensor(np.broadcast_to(np.asarray(target_accept_prob,
                                                   np.float64), (p,)).copy(),
                        device=dev)
  n_total = num_burnin_steps + num_results
  rows = max(n_total, 1)
  lar = torch.empty((rows, c), dtype=dt, device=dev)
  lf = torch.zeros((rows, c), dtype=torch.int32, device=dev)
  acc = torch.empty((rows, c), dtype=torch.uint8, device=dev)
  div = torch.zeros((rows, c), dtype=torch.uint8, device=dev)
  tr = torch.empty((num_results, c, d), dtype=dt, device=dev) if trace else None
  base_seed = samplers.sanitize_seed(seed, salt='mcmc.sample_chain')
  stream = _lib.stream_handle(dev)
  fn = (_lib.lib().hipmc_nuts_run if kernel == 'nuts'
        else _lib.lib().hipmc_hmc_run)
  events = [] if (stats is not None and stats.get('time_kernels')) else None

  def launch(t0, n):
    a = _lib.RunArgs()
    a.dtype = _lib.dtype_code(dt)
    a.n_chains = c
    a.dim_stride = d
    a.max_dim = batch.max_dim
    a.n_iters = n
    a.fold_seed = 1
    a.iter0 = t0
    a.seed[0], a.seed[1] = int(base_seed[0]), int(base_seed[1])
    a.num_leapfrog_steps = int(num_leapfrog_steps)
    a.max_tree_depth = int(max_tree_depth)
    a.unrolled_leapfrog_steps = 1
    a.target_kind = batch.kind
    a.waves_per_chain = batch.waves_per_chain
    a.max_energy_diff = float(max_energy_diff)
    a.problems = batch.table.data_ptr()
    a.n_problems = p
    a.chain_problem = batch.chain_problem.data_ptr()
    a.chain_local = batch.chain_local.data_ptr()
    a.step_size = step.data_ptr()
    a.chain_step = batch.chain_problem.data_ptr()
    a.state, a.tlp, a.grad = x.data_ptr(), tlp.data_ptr(), grad.data_ptr()
    a.out_row0 = t0
    a.out_log_accept_ratio = lar.data_ptr()
    a.out_is_accepted = acc.data_ptr()
    if kernel == 'nuts':
      a.out_leapfrogs = lf.data_ptr()
      a.out_has_divergence = div.data_ptr()
    if tr is not None:
      a.trace_state = tr.data_ptr()
      a.trace_start = num_burnin_steps
      a.trace_every = 1
      a.n_trace = num_results
    if events is not None:
      e0 = torch.cuda.Event(enable_timing=True)
      e1 = torch.cuda.Event(enable_timing=True)
      e0.record()
    _lib.check(fn(a, stream))
    if events is not None:
      e1.record()
      events.append((e0, e1, t0, n))

  def da(t0, n):
    _lib.check(_lib.lib().hipmc_da_update(
        _lib.dtype_code(dt), p, None, None, _lib.ptr(batch.group_start),
        _lib.ptr(lar[t0]), n, c, _lib.ptr(err), _lib.ptr(avg), _lib.ptr(lst),
        _lib.ptr(tgt), t0, num_adaptation_steps, 0.05, 10.0, 0.75,
        _lib.ptr(step), _lib.ptr(step64), stream))

  t = 0
  while t < n_total:
    n = 1 if t < num_adaptation_steps else min(n_total - t,
                                                max_iters_per_launch)
    launch(t, n)
    if num_adaptation_steps > 0:
      da(t, n)
    t += n
    if progress is not None:
      progress(t, n_total)
  if stats is not None:
    stats['leapfrogs_per_step'] = lf[:n_total].sum(1)
    stats['leapfrogs_total'] = int(lf[:n_total].sum())
    stats['leapfrogs_per_problem'] = torch.zeros(
        p, dtype=torch.float64, device=dev).index_add_(
            0, batch.chain_problem.long(), lf[:n_total].sum(0).double()
        ).cpu().numpy()
    if events is not None:
      torch.cuda.synchronize(dev)
      stats['launches'] = [(e0.elapsed_time(e1), t0, n)
                           for e0, e1, t0, n in events]
  draws = None
  if tr is not None:
    con = batch.bijector(tr, 0, chain_period=c) if tr.numel() else tr
    draws = batch.split(con)
  rec = slice(num_burnin_steps, n_total)
  return BatchResults(
      draws=draws, step_size=step.clone(),
      log_accept_ratio=batch.split_chains(lar[rec]),
      leapfrogs_taken=batch.split_chains(lf[rec]),
      is_accepted=batch.split_chains(acc[rec].bool()),
      has_divergence=batch.split_chains(div[rec].bool()),
      final_state=batch.split(batch.bijector(x, 0)))
