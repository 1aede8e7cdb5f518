"""Continuous-batching posterior sampling server.

Requests (a target spec + chain count + adaptation/draw counts + seed) are
admitted while device memory stays within an HBM budget and chain slots are
free.  Admitted requests keep their chain state, step size and
dual-averaging accumulators resident in device pools; every scheduling tick
launches, per target family, one NUTS launch over all requests still adapting
(one transition, then a per-request dual-averaging update) and one over all
requests in their sampling phase (several transitions).  Each (family,
phase) pair is a lane with its own stream that advances as soon as its own
previous launch has finished, so short requests are not held back by the
slowest chains of another lane.  A request's chains are keyed by (its seed,
its key, chain index) and advance by per-request iteration counters, so its
draws do not depend on which other requests share the launches.

Per-request semantics: sample_chain(DualAveraging(TTK(NUTS))) with
num_burnin_steps = num_adaptation_steps (`mcmc/sample.py`, `mcmc/nuts.py`,
`mcmc/dual_averaging_step_size_adaptation.py`).
"""

import collections
import ctypes
import math
import time

import numpy as np
import torch

from probability_amd import _lib
from probability_amd import bijectors as bij_lib
from probability_amd import random as samplers
from probability_amd import targets as targets_lib
from probability_amd import tuning

__all__ = ['Request', 'SamplerServer']


class Request:
  """One posterior-sampling request."""

  _next = 0

  def __init__(self, spec, num_chains=16, num_adaptation_steps=500,
               num_results=2000, seed=0, bijectors=None, step_size=0.1,
               target_accept_prob=0.8, max_tree_depth=10, key=None,
               arrival=None):
    self.spec = targets_lib.as_target(spec)
    self.num_chains = int(num_chains)
    self.num_adaptation_steps = int(num_adaptation_steps)
    self.num_results = int(num_results)
    self.seed = samplers.sanitize_seed(seed, salt='mcmc.sample_chain')
    self.bijectors = bijectors
    self.step_size = float(step_size)
    self.target_accept_prob = float(target_accept_prob)
    self.max_tree_depth = int(max_tree_depth)
    self.id = Request._next
    Request._next += 1
    self.key = self.id if key is None else int(key)
    self.arrival = arrival
    self.admitted = None
    self.done = None
    self.iter = 0
    self.draws = None
    self.final_step_size = None
    self.rows = None
    self.da_slot = None
    self.in_flight = False
    self.waves = None
    self.pool_key = None

  @property
  def total_steps(self):
    return self.num_adaptation_steps + self.num_results

  @property
  def adapting(self):
    return self.iter < self.num_adaptation_steps

  def nbytes(self, dtype_size=4):
    d = self.spec.dim
    return (self.num_results * self.num_chains * d * dtype_size +
            data_bytes(self.spec))


def data_bytes(spec):
  tot = 0
  for name in ('x', 'xt', 'y', 'sigma', 'loc', 'scale'):
    t = getattr(spec, name, None)
    if isinstance(t, torch.Tensor):
      tot += t.numel() * t.element_size()
  return tot


def _family(spec):
  return 'hier' if spec.kind == _lib.TARGET_HIER_NORMAL else 'glm'


class _Pool:
  """Device-resident chain and dual-averaging state of one target family."""

  def __init__(self, family, slot_dim, max_chains, max_requests, dtype,
               device):
    self.family = family
    self.dim_stride = int(slot_dim)
    self.dtype, self.device = dtype, device
    self.state = torch.zeros((max_chains, slot_dim), dtype=dtype, device=device)
    self.tlp = torch.zeros(max_chains, dtype=dtype, device=device)
    self.grad = torch.zeros((max_chains, slot_dim), dtype=dtype, device=device)
    f64 = torch.float64
    self.step = torch.zeros(max_requests, dtype=dtype, device=device)
    self.step64 = torch.zeros(max_requests, dtype=f64, device=device)
    self.err = torch.zeros(max_requests, dtype=f64, device=device)
    self.avg = torch.zeros(max_requests, dtype=f64, device=device)
    self.lst = torch.zeros(max_requests, dtype=f64, device=device)
    self.tgt = torch.zeros(max_requests, dtype=f64, device=device)
    self.free_rows = list(range(max_chains))
    self.free_da = list(range(max_requests))

  def can_hold(self, req):
    return (len(self.free_rows) >= req.num_chains and self.free_da and
            req.spec.dim <= self.dim_stride)

  def admit(self, req):
    req.rows = [self.free_rows.pop(0) for _ in range(req.num_chains)]
    req.da_slot = self.free_da.pop(0)
    s = req.da_slot
    self.step[s] = req.step_size
    self.step64[s] = float(np.float64(np.asarray(req.step_size,
                                                 np.float32 if self.dtype ==
                                                 torch.float32 else
                                                 np.float64)))
    self.err[s] = 0.0
    self.avg[s] = 0.0
    self.lst[s] = math.log(10.0) + math.log(float(self.step64[s]))
    self.tgt[s] = req.target_accept_prob

  def release(self, req):
    self.free_rows.extend(req.rows)
    self.free_rows.sort()
    self.free_da.append(req.da_slot)
    req.rows = None


class _Lane:
  """The launches of one (slot class, waves per chain, phase): a stream and
  the launch in flight (completion event, its requests, buffers kept alive
  until it ends).  pool_key is the slot class: (family, state slots)."""

  def __init__(self, pool_key, waves, adapting, device):
    self.pool_key = pool_key
    self.family = pool_key[0]
    self.waves = waves
    self.adapting = adapting
    self.stream = torch.cuda.Stream(device=device)
    self.inflight = None
    # device tables of the lane's current set of requests (rebuilt when a
    # request joins or leaves the lane)
    self.tables = None


# hipmc_stream as a NumPy record (one row per request of a launch)
_STREAM_DTYPE = np.dtype([
    ('seed', '<u4', (2,)), ('iter0', '<i8'), ('n_iters', '<i4'),
    ('trace_every', '<i4'), ('trace', '<u8'), ('trace_start', '<i8'),
    ('n_trace', '<i4'), ('pad_', '<i4')])
assert _STREAM_DTYPE.itemsize == ctypes.sizeof(_lib.Stream)


class _LaneTables:
  """Device data of one set of requests in a lane: the problem table, the
  per-chain maps (problem, local index, pool row, step-size slot), the DA
  group table and slots, and the host template of the per-request Philox
  streams.  Uploaded once per set, in one copy."""

  def __init__(self, reqs, device, stream):
    p = len(reqs)
    raw = bytearray()
    cp, cl, rows, cs = [], [], [], []
    offs = [0]
    st = np.zeros(p, _STREAM_DTYPE)
    for i, r in enumerate(reqs):
      prob = _lib.Problem.from_buffer_copy(r.problem_bytes)
      prob.n_chains = r.num_chains
      prob.chain_offset = offs[-1]
      prob.key = r.key
      raw += bytes(prob)
      cp += [i] * r.num_chains
      cl += list(range(r.num_chains))
      rows += r.rows
      cs += [r.da_slot] * r.num_chains
      offs.append(offs[-1] + r.num_chains)
      st['seed'][i] = (int(r.seed[0]), int(r.seed[1]))
      st['trace_every'][i] = 1
      st['trace'][i] = r.draws.data_ptr()
      st['trace_start'][i] = r.num_adaptation_steps
      st['n_trace'][i] = r.num_results
    n = offs[-1]
    ints = np.concatenate([np.asarray(v, np.int32) for v in
                           (cp, cl, rows, cs, offs,
                            [r.da_slot for r in reqs])])
    blob = bytes(raw) + ints.tobytes()
    with torch.cuda.stream(stream):
      self.buf = torch.frombuffer(bytearray(blob), dtype=torch.uint8).to(
          device, non_blocking=True)
    base = self.buf.data_ptr()
    self.p, self.n = p, n
    self.table = base
    o = base + len(raw)
    self.chain_problem, self.chain_local = o, o + 4 * n
    self.chain_row, self.chain_step = o + 8 * n, o + 12 * n
    self.group_start = o + 16 * n
    self.slots = self.group_start + 4 * (p + 1)
    self.streams = st
    self.adapt_len = np.asarray([r.num_adaptation_steps for r in reqs],
                                np.int64)
    self.reqs = reqs


class SamplerServer:
  """Admits requests by HBM budget and advances them in shared launches.

  Args:
    hbm_budget_bytes: device bytes the server may hold for admitted requests
      (their data and draw buffers).
    iters_per_tick: transitions a sampling-phase request makes per launch.
    slot_dims: register-slot classes per family, an int or a list of ints
      ({'glm': 192, 'hier': [512, 1088]}); a request joins the smallest
      class of its family that holds its state, and the requests of a class
      run the same kernel code.  Hierarchical states up to 1088 elements
      (J <= 1086 groups) and GLM states up to 512 are built.
    glm_waves_per_chain: waves sharing each GLM chain's gradient; None (the
      default) takes each request's own from its N x D, a number fixes it
      for every GLM request.  Either way a request's draws do not depend on
      the mix: requests of one wave count share lanes (launches) of their
      own, which run concurrently with the others.
  """

  def __init__(self, dtype=torch.float32, device='cuda',
               hbm_budget_bytes=32 << 30, max_chains=32768, max_requests=2048,
               iters_per_tick=8, slot_dims=None, glm_waves_per_chain=None):
    _lib.require_cuda()
    self.dtype = dtype
    self.device = torch.device(device)
    self.budget = int(hbm_budget_bytes)
    self.used = 0
    self.iters_per_tick = int(iters_per_tick)
    self.glm_waves_per_chain = (None if glm_waves_per_chain is None
                                else int(glm_waves_per_chain))
    slot_dims = slot_dims or {'glm': 192, 'hier': [512, 1088]}
    self.pools = {}
    for f, dims in slot_dims.items():
      for d in sorted([dims] if isinstance(dims, int) else dims):
        self.pools[(f, int(d))] = _Pool(f, int(d), max_chains, max_requests,
                                        dtype, self.device)
    self.queue = collections.deque()
    self.active = []
    self.finished = []
    self.lanes = {}
    self.ticks = 0

  def _waves(self, req):
    if _family(req.spec) != 'glm':
      return 1
    if self.glm_waves_per_chain is not None:
      return self.glm_waves_per_chain
    return tuning.glm_waves(req.spec.kind, req.spec.n_obs, req.spec.dim,
                            req.num_chains, 'nuts', self.dtype)

  def _pool_key(self, req):
    fam = _family(req.spec)
    fits = [k for k in self.pools if k[0] == fam and k[1] >= req.spec.dim]
    if not fits:
      raise ValueError(f'no {fam} slot class holds a state of '
                       f'{req.spec.dim} elements')
    return min(fits, key=lambda k: k[1])

  def _lanes_for(self, req):
    key, w = req.pool_key, req.waves
    for ad in (True, False):
      if (key, w, ad) not in self.lanes:
        self.lanes[(key, w, ad)] = _Lane(key, w, ad, self.device)

  # ------------------------------------------------------------ admission
  def submit(self, req):
    if req.arrival is None:
      req.arrival = time.perf_counter()
    self.queue.append(req)

  def _admit(self):
    admitted = []
    while self.queue:
      req = self.queue[0]
      req.pool_key = self._pool_key(req)
      pool = self.pools[req.pool_key]
      need = req.nbytes(torch.tensor([], dtype=self.dtype).element_size())
      if self.used + need > self.budget or not pool.can_hold(req):
        break
      self.queue.popleft()
      self.used += need
      pool.admit(req)
      req.waves = self._waves(req)
      self._lanes_for(req)
      req.admitted = time.perf_counter()
      req.draws = torch.empty((req.num_results, req.num_chains, req.spec.dim),
                              dtype=self.dtype, device=self.device)
      tspec = (targets_lib.TransformedTarget(req.spec, req.bijectors)
               if req.bijectors else req.spec)
      req.table = tspec.table(tspec.bijectors_for_eval())
      req.problem_bytes = bytes(tspec.problem(tspec.bijectors_for_eval()))
      x = torch.zeros((req.num_chains, req.spec.dim), dtype=self.dtype,
                      device=self.device)
      tlp, grad = tspec.log_prob_and_grad_flat(x)
      rows = torch.as_tensor(req.rows, device=self.device)
      pool.state[rows] = torch.nn.functional.pad(
          x, (0, pool.dim_stride - x.shape[1]))
      pool.grad[rows] = torch.nn.functional.pad(
          grad, (0, pool.dim_stride - grad.shape[1]))
      pool.tlp[rows] = tlp
      self.active.append(req)
      admitted.append(req)
    return admitted

  # ------------------------------------------------------------ launches
  def _launch(self, pool, reqs, adapting, lane):
    if not reqs:
      return None
    dt = self.dtype
    dev = self.device
    stream = lane.stream
    tb = lane.tables
    if tb is None or tb.reqs != reqs:
      tb = lane.tables = _LaneTables(reqs, dev, stream)
    p, n = tb.p, tb.n
    n_iters = 1 if adapting else self.iters_per_tick
    # per launch: each request's iteration offset and count, and (adapting)
    # its DA step count and adaptation length: one upload
    st = tb.streams
    st['iter0'] = [r.iter for r in reqs]
    st['n_iters'] = [1 if adapting else min(self.iters_per_tick,
                                            r.total_steps - r.iter)
                     for r in reqs]
    blob = st.tobytes()
    if adapting:
      blob += np.asarray([r.iter for r in reqs], np.int64).tobytes()
      blob += tb.adapt_len.tobytes()
    with torch.cuda.stream(stream):
      up = torch.frombuffer(bytearray(blob), dtype=torch.uint8).to(
          dev, non_blocking=True)
      lar = torch.empty((n_iters, n), dtype=dt, device=dev)
      a = _lib.RunArgs()
      a.dtype = _lib.dtype_code(dt)
      a.n_chains = n
      a.dim_stride = pool.dim_stride
      a.max_dim = pool.dim_stride
      a.n_iters = n_iters
      a.fold_seed = 1
      a.max_tree_depth = reqs[0].max_tree_depth
      a.unrolled_leapfrog_steps = 1
      a.target_kind = (_lib.TARGET_HIER_NORMAL if pool.family == 'hier'
                       else _lib.TARGET_LOGISTIC)
      a.waves_per_chain = reqs[0].waves
      a.max_energy_diff = 1000.0
      a.problems = tb.table
      a.n_problems = p
      a.chain_problem, a.chain_local = tb.chain_problem, tb.chain_local
      a.chain_row, a.chain_step = tb.chain_row, tb.chain_step
      a.step_size = pool.step.data_ptr()
      a.state, a.tlp, a.grad = (pool.state.data_ptr(), pool.tlp.data_ptr(),
                                pool.grad.data_ptr())
      a.out_log_accept_ratio = lar.data_ptr()
      a.streams = up.data_ptr()
      _lib.check(_lib.lib().hipmc_nuts_run(a, stream.cuda_stream))
      if adapting:
        # every request's DA step in one launch, each at its own step count
        # and adaptation length
        es = up.data_ptr() + st.nbytes
        _lib.check(_lib.lib().hipmc_da_update(
            _lib.dtype_code(dt), p, None, tb.slots, tb.group_start,
            _lib.ptr(lar), 1, n, _lib.ptr(pool.err), _lib.ptr(pool.avg),
            _lib.ptr(pool.lst), _lib.ptr(pool.tgt), 0, 0, 0.05, 10.0, 0.75,
            _lib.REDUCE_MEAN, es, es + 8 * p, _lib.ptr(pool.step),
            _lib.ptr(pool.step64), stream.cuda_stream))
    return (tb, up, lar)

  def tick(self):
    """Admits what fits, then for every lane whose launch has finished
    advances its requests, retires the finished ones and starts the lane's
    next launch; waits for the oldest launch when none has finished.
    Returns the requests finished in this call."""
    if self._admit():
      # the admitted rows were written on the current stream
      ev = torch.cuda.Event()
      ev.record()
      for lane in self.lanes.values():
        lane.stream.wait_event(ev)
    done = []
    while True:
      progressed = False
      for lane in self.lanes.values():
        if lane.inflight is not None:
          if not lane.inflight[0].query():
            continue
          done += self._complete(lane)
          progressed = True
        progressed |= self._start(lane)
      busy = [l for l in self.lanes.values() if l.inflight is not None]
      if progressed or not busy:
        return done
      busy[0].inflight[0].synchronize()

  def _start(self, lane):
    reqs = [r for r in self.active if r.pool_key == lane.pool_key and
            r.waves == lane.waves and r.adapting == lane.adapting and
            not r.in_flight]
    if not reqs:
      return False
    pool = self.pools[lane.pool_key]
    keep = self._launch(pool, reqs, lane.adapting, lane)
    ev = torch.cuda.Event()
    ev.record(lane.stream)
    for r in reqs:
      r.in_flight = True
    lane.inflight = (ev, reqs, keep)
    return True

  def _complete(self, lane):
    _, reqs, _ = lane.inflight
    lane.inflight = None
    self.ticks += 1
    now = time.perf_counter()
    done = []
    for r in reqs:
      r.in_flight = False
      if lane.adapting:
        r.iter += 1
      else:
        r.iter = min(r.total_steps, r.iter + self.iters_per_tick)
      if r.iter >= r.total_steps:
        done.append(r)
    for r in done:
      pool = self.pools[r.pool_key]
      r.final_step_size = float(pool.step[r.da_slot])
      pool.release(r)
      self.active.remove(r)
      self.used -= r.nbytes(torch.tensor([], dtype=self.dtype).element_size())
      r.done = now
      if r.bijectors:
        r.draws = _forward(r, self.dtype)
      self.finished.append(r)
    return done

  def idle(self):
    return not self.queue and not self.active


def _forward(req, dtype):
  """Constrains a finished request's draws (TTK forward map)."""
  d = req.spec.dim
  flat = req.draws.reshape(-1, d)
  out = torch.empty_like(flat)
  _lib.check(_lib.lib().hipmc_bijector(
      _lib.dtype_code(dtype), _lib.ptr(req.table), None, flat.shape[0], d, 0,
      0, _lib.ptr(flat), _lib.ptr(out), _lib.stream_handle(flat.device)))
  return out.reshape(req.draws.shape)


def run_alone(req, dtype=torch.float32, device='cuda', iters_per_tick=8):
  """Runs one request by itself through the same server (reference for the
  alone-vs-mix equality)."""
  srv = SamplerServer(dtype=dtype, device=device, iters_per_tick=iters_per_tick)
  srv.submit(req)
  while not srv.idle():
    srv.tick()
  return req
