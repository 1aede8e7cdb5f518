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
from probability_amd import tuning
from probability_amd import bijectors as bij_lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib

__all__ = ['PosteriorBatch', 'sample_posteriors', 'sample_ragged',
           'RaggedRun', 'BatchResults', 'WindowState', 'window_sizes']

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
    waves_per_chain: GLMs only, waves sharing one chain's gradient (1-8;
      None: chosen from the largest design matrix).  Like slot_dim, equal
      values give bit-identical results alone or in any mix.
    chains_per_block: GLMs only, chains of one problem that share every
      pass over X (1, 2 or 4; the kernel uses fewer where the LDS or the
      waves do not allow more).  Every problem's chain count must be a
      multiple of it (else 1 is used).  Results depend on it only through
      the order of floating-point sums; equal values give bit-identical
      results alone or in any mix.
    group_adaptation: use the chain groups in the adaptation steps too; by
      default only the sampling phase's launches use them, since a single
      transition of a group lasts as long as its longest trajectory.  The
      choice follows the phase, not a launch's length, so the draws do not
      depend on max_iters_per_launch.
  """

  def __init__(self, specs, num_chains, bijectors=None, keys=None,
               slot_dim=None, waves_per_chain=None, chains_per_block=1,
               group_adaptation=False):
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
    g = int(chains_per_block or 1)
    if g not in (1, 2, 4):
      raise ValueError('chains_per_block must be 1, 2 or 4')
    if _family(self.kind) != 1 or any(n % g for n in self.num_chains):
      g = 1
    self.chains_per_block = g
    self.group_adaptation = bool(group_adaptation)
    if _family(self.kind) != 1:
      self.waves_per_chain = 1
    elif waves_per_chain is not None:
      self.waves_per_chain = int(waves_per_chain)
    else:
      big = max(self.specs, key=lambda s: s.n_obs * s.dim)
      self.waves_per_chain = tuning.glm_waves(
          big.kind, big.n_obs, big.dim, max(self.num_chains), 'nuts',
          big.dtype)

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
  run = _Run(batch, torch.cuda.current_stream(batch.device), num_results,
             num_burnin_steps, num_adaptation_steps, init_state, step_size,
             target_accept_prob, max_tree_depth, max_energy_diff, kernel,
             num_leapfrog_steps, seed, trace, stats, max_iters_per_launch)
  while not run.done():
    run.advance()
    if progress is not None:
      progress(run.t, run.n_total)
  return run.finish()


def sample_ragged(specs, num_chains, num_results, bijectors=None, keys=None,
                  slot_dim=None, progress=None, stats=None, chains_per_block=1,
                  group_adaptation=False, class_splits=None, **kwargs):
  """sample_posteriors for problems of very different sizes: the GLMs are
  split into classes by the waves that share one of their chains (from each
  problem's own N x D), every class is its own batch on its own stream, and
  the classes advance concurrently, each at its own pace (small problems do
  not wait for the largest ones' adaptation steps).  A problem's draws
  depend only on its own size class, key and chain count, so they equal its
  draws alone.  chains_per_block: an int, or a dict from a class's waves per
  chain to the chains per block of its sampling launches (and, with
  group_adaptation, of its adaptation steps).  class_splits: a dict from a
  class's waves per chain to the number of batches (each on its own
  stream) its problems are dealt to, by cost: every adaptation step of a
  batch lasts as long as its longest trajectory, so batches that advance
  independently keep one problem's long trajectories from holding up the
  others.  Returns BatchResults in the order of `specs`.
  """
  run = RaggedRun(specs, num_chains, num_results, bijectors=bijectors,
                  keys=keys, slot_dim=slot_dim, stats=stats,
                  chains_per_block=chains_per_block,
                  group_adaptation=group_adaptation,
                  class_splits=class_splits, **kwargs)
  while not run.done():
    run.advance_launches()
    if progress is not None:
      progress(run.t, run.n_total)
  return run.finish()


class RaggedRun:
  """The size classes of sample_ragged, each a batch on its own stream,
  advanced launch by launch (advance_launches) or up to a transition
  (advance_to)."""

  def __init__(self, specs, num_chains, num_results, bijectors=None,
               keys=None, slot_dim=None, stats=None, chains_per_block=1,
               group_adaptation=False, class_splits=None, **kwargs):
    if not specs:
      raise ValueError('need at least one problem')
    specs = [targets_lib.as_target(s) for s in specs]
    p = len(specs)
    self.specs, self.stats, self.kwargs = specs, stats, kwargs
    # windowed adaptation starts every chain uniform on (-2, 2) from the
    # setup half of the seed, as windowed_adaptive_nuts does
    windowed_init = kwargs.get('adaptation') == 'windowed'
    if windowed_init:
      setup_seed, kwargs['seed'] = samplers.split_seed(
          samplers.sanitize_seed(kwargs.get('seed')), 2)
    chains = ([int(num_chains)] * p if np.isscalar(num_chains)
              else [int(c) for c in num_chains])
    bijs = bijectors or [None] * p
    keys = list(range(p)) if keys is None else [int(k) for k in keys]
    glm = _family(specs[0].kind) == 1
    # a problem's class depends on its own size and chain count only (the
    # tuned wave count for its chains alone), so its draws equal its draws
    # alone
    cls = [tuning.glm_waves(s.kind, s.n_obs, s.dim, chains[i], 'nuts',
                            s.dtype) if glm else 1
           for i, s in enumerate(specs)]
    dev = specs[0].device
    self.caller = torch.cuda.current_stream(dev)
    self.runs = []
    cost = [s.n_obs * s.dim if glm else s.dim for s in specs]
    subs = []
    for w in sorted(set(cls), reverse=True):
      members = [i for i in range(p) if cls[i] == w]
      n_sub = int((class_splits or {}).get(w, 1))
      subs += [(w, idx) for idx in _split_by_cost(members, cost, n_sub)]
    for w, idx in subs:
      if glm:
        idx = _xcd_order(idx, [cost[i] for i in idx])
      g = (chains_per_block.get(w, 1) if isinstance(chains_per_block, dict)
           else chains_per_block)
      batch = PosteriorBatch([specs[i] for i in idx],
                             [chains[i] for i in idx],
                             [bijs[i] for i in idx], [keys[i] for i in idx],
                             slot_dim=slot_dim, waves_per_chain=w,
                             chains_per_block=g,
                             group_adaptation=group_adaptation)
      stream = torch.cuda.Stream(device=dev)
      stream.wait_stream(self.caller)
      sub = {} if stats is not None else None
      if sub is not None and stats.get('time_kernels'):
        sub['time_kernels'] = True
      kw = dict(kwargs)
      # per-problem step sizes follow their problems into the classes
      st = kw.get('step_size')
      if st is not None and np.ndim(st) == 1 and len(st) == p:
        kw['step_size'] = [st[i] for i in idx]
      self.runs.append((idx, sub, _Run(batch, stream, num_results, stats=sub,
                                       **kw)))
    self.n_total = sum(r.n_total for _, _, r in self.runs)
    if windowed_init:
      for _, _, r in self.runs:
        with torch.cuda.stream(r.stream):
          r.x.copy_(_init_uniform(r.batch, setup_seed))
          r.tlp, r.grad = r.batch.eval(r.x)

  @property
  def t(self):
    return sum(r.t for _, _, r in self.runs)

  def done(self):
    return all(r.done() for _, _, r in self.runs)

  def advance_launches(self):
    """One launch (with its updates) for every class not yet done."""
    for _, _, r in self.runs:
      if not r.done():
        r.advance()

  def advance_to(self, t):
    """Enqueues every class up to transition t (launches end there)."""
    while any(r.t < min(t, r.n_total) for _, _, r in self.runs):
      for _, _, r in self.runs:
        if r.t < min(t, r.n_total):
          r.advance(limit=t)

  def finish(self):
    specs, stats, caller = self.specs, self.stats, self.caller
    p = len(specs)
    dev = specs[0].device
    runs = self.runs
    parts = [(idx, sub, r.finish(consumer=caller)) for idx, sub, r in runs]
    for _, _, r in runs:
      caller.wait_stream(r.stream)
    out = {f: [None] * p for f in BatchResults._fields}
    step = torch.empty(p, dtype=specs[0].dtype, device=dev)
    for idx, _, res in parts:
      for f in BatchResults._fields:
        if f == 'step_size':
          step[torch.as_tensor(idx, device=dev)] = res.step_size
        elif getattr(res, f) is not None:
          for j, i in enumerate(idx):
            out[f][i] = getattr(res, f)[j]
    if stats is not None:
      stats['leapfrogs_total'] = sum(sub['leapfrogs_total']
                                     for _, sub, _ in parts)
      lpp = np.zeros(p)
      for idx, sub, _ in parts:
        lpp[idx] = sub['leapfrogs_per_problem']
      stats['leapfrogs_per_problem'] = lpp
      stats['classes'] = [(len(idx), r.batch.waves_per_chain)
                          for idx, _, r in runs]
      stats['class_stats'] = [{k: sub[k] for k in (
          'cost_total', 'cost_adaptation', 'cost_critical',
          'cost_critical_adaptation', 'leapfrogs_total', 'launches')
                               if k in sub} for _, sub, _ in parts]
      if stats.get('time_kernels'):
        stats['launches'] = [l for _, sub, _ in parts for l in sub['launches']]
    out['step_size'] = step
    if not self.kwargs.get('trace', True):
      out['draws'] = None
    return BatchResults(**out)


def _init_uniform(batch, seed):
  """Unconstrained states uniform on (-2, 2), keyed per (problem, chain)
  (hipmc_draws kind 4); zero past each problem's dim."""
  x = torch.zeros((batch.n_chains, batch.dim_stride), dtype=batch.dtype,
                  device=batch.device)
  sd = np.ascontiguousarray(np.asarray(seed, np.uint32).reshape(2))
  for i in range(batch.n_problems):
    a, b = int(batch.chain_offsets[i]), int(batch.chain_offsets[i + 1])
    d = batch.dims[i]
    out = torch.empty((b - a, d), dtype=batch.dtype, device=batch.device)
    _lib.check(_lib.lib().hipmc_draws(
        _lib.dtype_code(batch.dtype), 4, sd.ctypes.data, int(batch.keys[i]),
        b - a, d, 0, 0, _lib.ptr(out), None,
        _lib.stream_handle(batch.device)))
    x[a:b, :d] = out
  return x


def _split_by_cost(idx, cost, n):
  """idx dealt to at most n non-empty lists of similar total cost: most
  expensive first, each to the list with the least cost so far."""
  n = max(1, min(int(n), len(idx)))
  if n == 1:
    return [list(idx)]
  bins, load = [[] for _ in range(n)], [0.0] * n
  for i in sorted(idx, key=lambda i: -cost[i]):
    k = min(range(n), key=lambda j: load[j])
    bins[k].append(i)
    load[k] += cost[i]
  return bins


def _xcd_order(idx, cost, n_xcd=8):
  """Launch order for problems of unequal cost.  The kernels give XCD x the
  x-th contiguous eighth of the launch's blocks and each XCD starts its
  blocks in order, so: deal the problems, most expensive first, to the
  eight ranges in snake order (equal counts, similar total cost, each
  problem's chains on one XCD and its L2), and run each range from its
  most expensive problem down (long trajectories start first and small
  problems fill in around them)."""
  order = sorted(range(len(idx)), key=lambda j: -cost[j])
  bins = [[] for _ in range(n_xcd)]
  for r, j in enumerate(order):
    k = r % (2 * n_xcd)
    bins[k if k < n_xcd else 2 * n_xcd - 1 - k].append(idx[j])
  return [i for b in bins for i in b]


def window_sizes(num_adaptation_steps):
  """First fast window, first slow window and last fast window of windowed
  adaptation (`_get_window_sizes`, experimental/mcmc/windowed_sampling.py:
  367-395): slow windows of s, 2s, 4s, 8s steps with s = A // 20 between a
  3s-step first window and the rest."""
  slow = num_adaptation_steps // 20
  first = 3 * slow
  return first, slow, num_adaptation_steps - 15 * slow - first


class WindowState:
  """Host bookkeeping of WindowedAdaptation(DiagonalMassMatrixAdaptation(
  DualAveraging(...))): the dual averaging's own step count and length, and
  the mass-matrix estimation step and length.  Bootstrapped as the reference
  builds it: both run for the whole adaptation (`windowed_sampling.py:
  323-327, 509-520`) until the first window update."""

  def __init__(self, num_adaptation_steps):
    self.a = int(num_adaptation_steps)
    self.first, self.slow, self.last = window_sizes(self.a)
    self.da_step, self.da_num = 0, self.a
    self.dm_step, self.num_est = 0, self.a

  def option(self, step):
    """The update applied before inner step `step`
    (`windowed_sampling.py:471-494`): 1 first fast window, 2 first slow
    window, 3 a later slow window, 4 the last window; coinciding starts add
    up and the sum is clipped to 4 (`mcmc_util.choose_from`)."""
    f, s = self.first, self.slow
    opt = int(step == 0) + 2 * int(step == f)
    opt += 3 * int(any(step == f + s * (2 ** i - 1) for i in (1, 2, 3)))
    opt += 4 * int(step == f + 15 * s)
    return min(opt, 4)


class _Run:
  """One batch's sampling loop, advanced one launch at a time on `stream`."""

  def __init__(self, batch, stream, num_results, num_burnin_steps=0,
               num_adaptation_steps=0, init_state=None, step_size=0.1,
               target_accept_prob=0.75, max_tree_depth=10,
               max_energy_diff=1000.0, kernel='nuts', num_leapfrog_steps=10,
               seed=None, trace=True, stats=None, max_iters_per_launch=4096,
               adaptation='dual_averaging', init_unconstrained=None,
               history=False, unrolled_leapfrog_steps=1):
    if kernel not in ('nuts', 'hmc'):
      raise ValueError("kernel must be 'nuts' or 'hmc'")
    if adaptation not in ('dual_averaging', 'windowed'):
      raise ValueError("adaptation must be 'dual_averaging' or 'windowed'")
    self.windowed = adaptation == 'windowed'
    self.history = history
    self.unrolled = int(unrolled_leapfrog_steps)
    _lib.require_cuda()
    self.batch, self.stream = batch, stream
    self.kernel = kernel
    self.num_results, self.num_burnin = num_results, num_burnin_steps
    self.num_adapt = num_adaptation_steps
    self.max_tree_depth, self.max_energy_diff = max_tree_depth, max_energy_diff
    self.num_leapfrog_steps = num_leapfrog_steps
    self.stats = stats
    self.max_iters = max_iters_per_launch
    dev, dt = batch.device, batch.dtype
    c, d, p = batch.n_chains, batch.dim_stride, batch.n_problems
    with torch.cuda.stream(stream):
      if init_unconstrained is not None:
        self.x = torch.as_tensor(init_unconstrained, dtype=dt,
                                 device=dev).contiguous().clone()
      elif init_state is None:
        self.x = torch.zeros((c, d), dtype=dt, device=dev)
      else:
        self.x = batch.bijector(torch.as_tensor(init_state, dtype=dt,
                                                device=dev).contiguous(), 1)
      self.tlp, self.grad = batch.eval(self.x)
      self.step = torch.as_tensor(np.broadcast_to(
          np.asarray(step_size, np.float64), (p,)).copy(),
                                  device=dev).to(dt)
      self.step64 = self.step.to(torch.float64).clone()
      self.err = torch.zeros(p, dtype=torch.float64, device=dev)
      self.avg = torch.zeros_like(self.err)
      self.lst = np.log(10.0) + torch.log(self.step64)
      self.tgt = torch.as_tensor(np.broadcast_to(
          np.asarray(target_accept_prob, np.float64), (p,)).copy(), device=dev)
      self.n_total = num_burnin_steps + num_results
      rows = max(self.n_total, 1)
      self.lar = torch.empty((rows, c), dtype=dt, device=dev)
      self.lf = torch.zeros((rows, c), dtype=torch.int32, device=dev)
      self.acc = torch.empty((rows, c), dtype=torch.uint8, device=dev)
      self.div = torch.zeros((rows, c), dtype=torch.uint8, device=dev)
      self.tr = (torch.empty((num_results, c, d), dtype=dt, device=dev)
                 if trace else None)
      # windowed adaptation: diagonal momentum precision per problem, the
      # running variance it is estimated from, and the window bookkeeping
      self.prec = None
      if self.windowed:
        self.prec = torch.ones((p, d), dtype=dt, device=dev)
        self.rv_n = torch.zeros(p, dtype=torch.float64, device=dev)
        self.rv_mean = torch.zeros((p, d), dtype=torch.float64, device=dev)
        self.rv_ssr = torch.zeros((p, d), dtype=torch.float64, device=dev)
        self.win = WindowState(num_adaptation_steps)
        self.prec_changes = [(0, self.prec.clone())]
      # per-step step size and target log-prob (trace fields)
      if history:
        self.step_hist = torch.empty((max(self.n_total, 1), p), dtype=dt,
                                     device=dev)
        self.tlp_hist = torch.empty((max(self.n_total, 1), c), dtype=dt,
                                    device=dev)
        self.corr = torch.empty((max(self.n_total, 1), c), dtype=dt,
                                device=dev)
    self.base_seed = samplers.sanitize_seed(seed, salt='mcmc.sample_chain')
    self.fn = (_lib.lib().hipmc_nuts_run if kernel == 'nuts'
               else _lib.lib().hipmc_hmc_run)
    self.events = ([] if (stats is not None and stats.get('time_kernels'))
                   else None)
    self.t = 0

  def done(self):
    return self.t >= self.n_total

  def advance(self, limit=None):
    """Enqueues the next launch (and its step-size / mass-matrix updates);
    a launch ends at transition `limit` if given."""
    with torch.cuda.stream(self.stream):
      t = self.t
      adapting = t < self.num_adapt
      n = 1 if adapting else min(self.n_total - t, self.max_iters)
      if limit is not None:
        n = max(1, min(n, limit - t))
      if self.windowed and adapting:
        self._window_pre(t)
      if self.history:
        self.step_hist[t:t + n] = self.step
      self._launch(t, n)
      if self.windowed:
        if adapting:
          self._window_post(t)
      elif self.num_adapt > 0:
        self._da(t, n)
      self.t = t + n

  # ---- windowed adaptation (experimental/mcmc/windowed_sampling.py:415-500,
  # diagonal_mass_matrix_adaptation.py:186-244)
  def _da_reset(self):
    """DualAveraging._bootstrap_from_inner_results at the current step size
    (dual_averaging_step_size_adaptation.py:543-610)."""
    self.err.zero_()
    self.avg.zero_()
    self.lst = np.log(10.0) + torch.log(self.step.to(torch.float64))
    self.win.da_step = 0

  def _window_pre(self, t):
    w = self.win
    opt = w.option(t)
    if opt == 1:
      w.da_num = w.first + w.slow
      w.num_est = -1
    elif opt == 2:
      w.dm_step = 0
      w.num_est = w.slow
    elif opt == 3:
      cur = t - w.first + w.slow
      self.rv_n.zero_()
      self.rv_mean.zero_()
      self.rv_ssr.zero_()
      w.dm_step = 0
      w.num_est = cur
      self._da_reset()
      w.da_num = cur
    elif opt == 4:
      self._da_reset()
      w.da_num = w.last

  def _window_post(self, t):
    b, w = self.batch, self.win
    _lib.check(_lib.lib().hipmc_da_update(
        _lib.dtype_code(b.dtype), b.n_problems, None, None,
        _lib.ptr(b.group_start), _lib.ptr(self.lar[t]), 1, b.n_chains,
        _lib.ptr(self.err), _lib.ptr(self.avg), _lib.ptr(self.lst),
        _lib.ptr(self.tgt), w.da_step, w.da_num, 0.05, 10.0, 0.75,
        _lib.REDUCE_HARMONIC, None, None, _lib.ptr(self.step),
        _lib.ptr(self.step64),
        _lib.stream_handle(b.device)))
    w.da_step += 1
    w.dm_step += 1
    if w.dm_step <= w.num_est:
      _lib.check(_lib.lib().hipmc_variance_update(
          _lib.dtype_code(b.dtype), b.n_problems, _lib.ptr(b.group_start),
          None, _lib.ptr(self.x), b.dim_stride, b.dim_stride,
          _lib.ptr(self.rv_n), _lib.ptr(self.rv_mean), _lib.ptr(self.rv_ssr),
          _lib.stream_handle(b.device)))
    if w.dm_step == w.num_est:
      # the momentum precision becomes the population variance
      self.prec.copy_((self.rv_ssr / self.rv_n[:, None]).to(self.prec.dtype))
      # padding elements past a problem's dim are never read; keep them 1
      for i, dd in enumerate(b.dims):
        if dd < b.dim_stride:
          self.prec[i, dd:] = 1
      self.prec_changes.append((t + 1, self.prec.clone()))

  def _launch(self, t0, n):
    b, dt = self.batch, self.batch.dtype
    a = _lib.RunArgs()
    a.dtype = _lib.dtype_code(dt)
    a.n_chains = b.n_chains
    a.dim_stride = b.dim_stride
    a.max_dim = b.max_dim
    a.n_iters = n
    a.fold_seed = 1
    a.iter0 = t0
    a.seed[0], a.seed[1] = int(self.base_seed[0]), int(self.base_seed[1])
    a.num_leapfrog_steps = int(self.num_leapfrog_steps)
    a.max_tree_depth = int(self.max_tree_depth)
    a.unrolled_leapfrog_steps = self.unrolled
    a.target_kind = b.kind
    a.waves_per_chain = b.waves_per_chain
    # chain groups in the sampling phase, and in the adaptation steps (each
    # bound by its groups' slowest chains) on request
    sampling = t0 >= self.num_adapt
    a.chains_per_block = (b.chains_per_block
                          if sampling or b.group_adaptation else 1)
    a.max_energy_diff = float(self.max_energy_diff)
    a.problems = b.table.data_ptr()
    a.n_problems = b.n_problems
    a.chain_problem = b.chain_problem.data_ptr()
    a.chain_local = b.chain_local.data_ptr()
    a.step_size = self.step.data_ptr()
    a.chain_step = b.chain_problem.data_ptr()
    if self.prec is not None:
      a.momentum_precision = self.prec.data_ptr()
    if self.history:
      a.out_target_log_prob = self.tlp_hist.data_ptr()
      if self.kernel == 'hmc':
        a.out_energy = self.corr.data_ptr()
    a.state, a.tlp = self.x.data_ptr(), self.tlp.data_ptr()
    a.grad = self.grad.data_ptr()
    a.out_row0 = t0
    a.out_log_accept_ratio = self.lar.data_ptr()
    a.out_is_accepted = self.acc.data_ptr()
    if self.kernel == 'nuts':
      a.out_leapfrogs = self.lf.data_ptr()
      a.out_has_divergence = self.div.data_ptr()
    if self.tr is not None:
      a.trace_state = self.tr.data_ptr()
      a.trace_start = self.num_burnin
      a.trace_every = 1
      a.n_trace = self.num_results
    if self.events is not None:
      e0 = torch.cuda.Event(enable_timing=True)
      e1 = torch.cuda.Event(enable_timing=True)
      e0.record()
    _lib.check(self.fn(a, _lib.stream_handle(b.device)))
    if self.events is not None:
      e1.record()
      self.events.append((e0, e1, t0, n))

  def _da(self, t0, n):
    b = self.batch
    _lib.check(_lib.lib().hipmc_da_update(
        _lib.dtype_code(b.dtype), b.n_problems, None, None,
        _lib.ptr(b.group_start), _lib.ptr(self.lar[t0]), n, b.n_chains,
        _lib.ptr(self.err), _lib.ptr(self.avg), _lib.ptr(self.lst),
        _lib.ptr(self.tgt), t0, self.num_adapt, 0.05, 10.0, 0.75,
        _lib.REDUCE_MEAN, None, None, _lib.ptr(self.step),
        _lib.ptr(self.step64),
        _lib.stream_handle(b.device)))

  def finish(self, consumer=None):
    """Per-problem results (and stats) once every launch is enqueued.  With
    `consumer`, the result buffers are marked as used on that stream too, so
    the allocator does not hand them out again before its work is done."""
    b, n_total, dev = self.batch, self.n_total, self.batch.device
    with torch.cuda.stream(self.stream):
      lf = self.lf
      if self.stats is not None:
        self.stats['leapfrogs_per_step'] = lf[:n_total].sum(1)
        self.stats['leapfrogs_total'] = int(lf[:n_total].sum())
        self.stats['leapfrogs_per_problem'] = torch.zeros(
            b.n_problems, dtype=torch.float64, device=dev).index_add_(
                0, b.chain_problem.long(), lf[:n_total].sum(0).double()
            ).cpu().numpy()
        # per-chain cost of one leapfrog (design-matrix entries for GLMs,
        # state dimension otherwise): the busiest chain of each adaptation
        # step, and of each sampling launch, bounds that launch's time
        w = torch.as_tensor([float(getattr(s, 'n_obs', 1) * s.dim)
                             for s in b.specs], dtype=torch.float64,
                            device=dev)[b.chain_problem.long()]
        lfw = lf[:n_total].double() * w
        na = min(self.num_adapt, n_total)
        crit = float(lfw[:na].max(1).values.sum()) if na else 0.0
        for t0 in range(na, n_total, self.max_iters):
          crit += float(lfw[t0:t0 + self.max_iters].sum(0).max())
        self.stats['cost_total'] = float(lfw.sum())
        self.stats['cost_adaptation'] = float(lfw[:na].sum())
        self.stats['cost_critical'] = crit
        self.stats['cost_critical_adaptation'] = (
            float(lfw[:na].max(1).values.sum()) if na else 0.0)
        if self.events is not None:
          self.stream.synchronize()
          self.stats['launches'] = [(e0.elapsed_time(e1), t0, n)
                                    for e0, e1, t0, n in self.events]
      draws = con = None
      if self.tr is not None:
        con = (b.bijector(self.tr, 0, chain_period=b.n_chains)
               if self.tr.numel() else self.tr)
        draws = b.split(con)
      rec = slice(self.num_burnin, n_total)
      step = self.step.clone()
      acc, div = self.acc[rec].bool(), self.div[rec].bool()
      final = b.bijector(self.x, 0)
      if consumer is not None:
        for t in (con, step, self.lar, lf, acc, div, final):
          if t is not None:
            t.record_stream(consumer)
      return BatchResults(
          draws=draws, step_size=step,
          log_accept_ratio=b.split_chains(self.lar[rec]),
          leapfrogs_taken=b.split_chains(lf[rec]),
          is_accepted=b.split_chains(acc),
          has_divergence=b.split_chains(div),
          final_state=b.split(final))
