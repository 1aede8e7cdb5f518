"""Float64 NumPy restatement of the tfp.mcmc gradient-based kernels.

Paths are relative to `tensorflow_probability/python/`.  Chains are the
leading axis and every kernel is vectorised over chains exactly as TFP's
batched graph is; NUTS keeps TFP's lockstep loop structure ("run until every
chain stops", stopped chains masked) so that the per-chain GPU state machine
can be checked against the batched algorithm it replaces.

Draws come from a provider (`oracle.rng.PhiloxDraws` for the stream the GPU
uses, `oracle.rng.TableDraws` for injected randomness).

Test infrastructure only.
"""

import collections

import numpy as np

from oracle import rng as orng


# --------------------------------------------------------------------------
# small helpers


def safe_sum(xs, alt_value=-np.inf):
  """`mcmc/internal/util.py:205-235`: add, non-finite -> alt_value."""
  x = xs[0]
  for y in xs[1:]:
    x = x + y
  return np.where(np.isfinite(x), x, alt_value)


def log_add_exp(x, y):
  """`math/generic.py:585-611`: larger + softplus((x-larger)+(y-larger))."""
  with np.errstate(invalid='ignore', over='ignore'):
    larger = np.maximum(x, y)
    z = (x - larger) + (y - larger)
    sp = np.log1p(np.exp(-np.abs(z))) + np.maximum(z, 0)
    return larger + sp


def reduce_logmeanexp(x, axis=None, keepdims=False):
  """`math/generic.py:221-273`."""
  x = np.asarray(x, np.float64)
  if axis is None:
    axis = tuple(range(x.ndim))
  axis = tuple(np.atleast_1d(axis).tolist()) if np.ndim(axis) else (int(axis),)
  if len(axis) == 0:
    return x
  m = np.max(x, axis=axis, keepdims=True)
  m = np.where(np.isfinite(m), m, 0.0)
  with np.errstate(divide='ignore'):
    lse = np.log(np.sum(np.exp(x - m), axis=axis, keepdims=True)) + m
  n = 1
  for a in axis:
    n *= x.shape[a]
  out = lse - np.log(n)
  if not keepdims:
    out = np.squeeze(out, axis=axis)
  return out


# --------------------------------------------------------------------------
# leapfrog  (`mcmc/internal/leapfrog_integrator.py:222-362`)


def leapfrog(target, eps, num_steps, p, q, t, g, prec=None):
  """Returns (p', q', t', g') after `num_steps` leapfrog steps.

  Op order as in the reference: p += (0.5*eps)*g; L x [q += eps*v(p);
  (t, g) = value_and_grad(q); p += eps*g]; p -= (0.5*eps)*g, with the
  velocity v(p) = grad of the kinetic energy (`_one_step`, `:319-362`):
  p itself for the identity metric, prec * p for a diagonal momentum
  precision `prec` ([D] or [C, D]; the preconditioned kernels).  `eps` is a
  scalar or per-chain [C] array.
  """
  eps = np.asarray(eps, np.float64)
  if eps.ndim == 1:
    eps = eps[:, None]
  half = 0.5 * eps
  p = p + half * g
  for _ in range(num_steps):
    q = q + eps * (p if prec is None else prec * p)
    with np.errstate(all='ignore'):
      t, g = target.log_prob_and_grad(q)
    p = p + eps * g
  p = p - half * g
  return p, q, t, g


# --------------------------------------------------------------------------
# HMC + Metropolis-Hastings  (`mcmc/hmc.py:661-875`,
#                             `mcmc/metropolis_hastings.py:160-289`)

HMCResults = collections.namedtuple('HMCResults', [
    'is_accepted', 'log_accept_ratio', 'target_log_prob',
    'grads_target_log_prob', 'proposed_state', 'proposed_target_log_prob',
    'log_acceptance_correction', 'initial_momentum', 'final_momentum'])


def hmc_one_step(target, step_size, num_leapfrog_steps, q, t, g, seed,
                 draws=None, prec=None):
  """One calibrated HMC step for all chains.

  Seed use: MH splits `seed` into (proposal, acceptance)
  (`metropolis_hastings.py:183`); the proposal seed keys the momentum
  (`hmc.py:684-695`) and the acceptance seed keys the uniform
  (`metropolis_hastings.py:221-226`).  With a diagonal momentum precision
  `prec` this is PreconditionedHamiltonianMonteCarlo
  (`experimental/mcmc/preconditioned_hmc.py`): momentum z / sqrt(prec),
  velocity prec * p, and the correction is the difference of the kinetic
  energies 0.5 sum(prec p^2) (their normalising terms cancel).
  """
  draws = draws or orng.PhiloxDraws()
  c, d = q.shape
  prop_seed, acc_seed = orng.split_seed(seed, 2)
  p0 = draws.momentum(prop_seed, c, d).astype(np.float64)
  if prec is not None:
    p0 = p0 / np.sqrt(prec)
  p1, q1, t1, g1 = leapfrog(target, step_size, num_leapfrog_steps, p0, q, t, g,
                            prec)
  m = 1.0 if prec is None else prec
  with np.errstate(all='ignore'):
    corr = 0.5 * safe_sum([np.sum(m * p0 ** 2, -1), -np.sum(m * p1 ** 2, -1)])
    lar = safe_sum([t1, -t, corr])
    u = draws.mh_uniform(acc_seed, c).astype(np.float64)
    is_acc = np.log(u) < lar
  sel = is_acc[:, None]
  return np.where(sel, q1, q), HMCResults(
      is_accepted=is_acc, log_accept_ratio=lar,
      target_log_prob=np.where(is_acc, t1, t),
      grads_target_log_prob=np.where(sel, g1, g),
      proposed_state=q1, proposed_target_log_prob=t1,
      log_acceptance_correction=corr, initial_momentum=p0, final_momentum=p1)


# --------------------------------------------------------------------------
# MALA  (`mcmc/langevin.py:427-823`), constant volatility 1

MALAResults = collections.namedtuple('MALAResults', [
    'is_accepted', 'log_accept_ratio', 'target_log_prob',
    'grads_target_log_prob', 'proposed_state', 'proposed_target_log_prob',
    'log_acceptance_correction', 'diffusion_drift'])


def mala_one_step(target, step_size, q, t, g, seed, draws=None):
  """One MetropolisAdjustedLangevinAlgorithm step for all chains, written
  as the reference writes it: drift 0.5 * eps * grad (`_get_drift`,
  `:637-690`, volatility 1 and no volatility gradient), Euler-Maruyama
  proposal state + drift + sqrt(eps) * xi (`_euler_method`, `:579-635`),
  and the normal proposal-density correction
  log q(current | proposed) - log q(proposed | current)
  (`_compute_log_acceptance_correction`, `:692-823`).  The proposal draw xi
  uses the momentum draws of the proposal seed (this library's stream).
  """
  draws = draws or orng.PhiloxDraws()
  c, d = q.shape
  prop_seed, acc_seed = orng.split_seed(seed, 2)
  xi = draws.momentum(prop_seed, c, d).astype(np.float64)
  eps = float(step_size)
  drift = 0.5 * eps * g
  q1 = q + drift + np.sqrt(eps) * xi
  t1, g1 = target.log_prob_and_grad(q1)
  drift1 = 0.5 * eps * g1
  vol = np.sqrt(eps)
  with np.errstate(all='ignore'):
    diff = q1 - q
    e_fwd = (diff - drift) / vol
    e_bwd = (diff + drift1) / vol
    fwd = np.sum(np.log(vol) + 0.5 * e_fwd ** 2, -1)
    bwd = np.sum(np.log(vol) + 0.5 * e_bwd ** 2, -1)
    corr = fwd - bwd
    lar = safe_sum([t1, -t, corr])
    u = draws.mh_uniform(acc_seed, c).astype(np.float64)
    is_acc = np.log(u) < lar
  sel = is_acc[:, None]
  return np.where(sel, q1, q), MALAResults(
      is_accepted=is_acc, log_accept_ratio=lar,
      target_log_prob=np.where(is_acc, t1, t),
      grads_target_log_prob=np.where(sel, g1, g), proposed_state=q1,
      proposed_target_log_prob=t1, log_acceptance_correction=corr,
      diffusion_drift=drift1)


# --------------------------------------------------------------------------
# NUTS  (`mcmc/nuts.py`)


def nuts_instructions(max_depth):
  """U-turn memory write/read tables (role of `nuts.py:1013-1071`).

  Built by simulating the stack of stored left-most leaves of every subtree
  of the depth-`max_depth` tree: leaf r is stored if it is the left-most
  leaf of a subtree of size >= 2 (r even) until that subtree's last leaf;
  at leaf r the subtrees ending at r (sizes 2..2^tz(r+1)) are checked
  against the top tz(r+1) stack slots.  Non-stored leaves write to the
  spare slot `max_depth`.
  """
  n = 1 << max_depth
  write = np.full(n, max_depth, dtype=np.int32)
  read = np.zeros((n, 2), dtype=np.int32)
  stack = []
  for r in range(n):
    while stack and stack[-1] < r:
      stack.pop()
    v = ((r + 1) & -(r + 1)).bit_length() - 1
    if v > 0:
      read[r] = (len(stack) - v, len(stack))
    if r % 2 == 0:
      last = n - 1 if r == 0 else r + (r & -r) - 1
      stack.append(last)
      write[r] = len(stack) - 1
  return write, read


NUTSResults = collections.namedtuple('NUTSResults', [
    'target_log_prob', 'grads_target_log_prob', 'log_accept_ratio',
    'leapfrogs_taken', 'is_accepted', 'reach_max_depth', 'has_divergence',
    'energy', 'tree_depth'])


def compute_hamiltonian(t, p, prec=None):
  """`nuts.py:1085-1102`: returns tlp - 0.5*sum(p^2) (i.e. -H).

  With a diagonal momentum precision (`preconditioned_nuts.py:1091-1099`)
  the kinetic energy is -log N(p; 0, diag(1/prec)) =
  0.5 sum(prec p^2) + 0.5 d log(2 pi) - 0.5 sum(log prec).
  """
  if prec is None:
    return t - 0.5 * np.sum(p ** 2, axis=-1)
  d = p.shape[-1]
  kin = (0.5 * np.sum(prec * p ** 2, axis=-1) + 0.5 * d * np.log(2 * np.pi)
         - 0.5 * np.sum(np.log(prec) * np.ones_like(p), axis=-1))
  return t - kin


def nuts_one_step(target, step_size, q, t, g, seed, draws=None, max_depth=10,
                  max_energy_diff=1000.0, unrolled_leapfrog_steps=1,
                  precision=None):
  """One NUTS transition, batched lockstep restatement of `nuts.py:321-946`.

  `step_size` is a scalar or per-chain [C].  With `precision` ([D] or
  [C, D], the diagonal precision of the momentum distribution) this is the
  PreconditionedNoUTurnSampler (`experimental/mcmc/preconditioned_nuts.py:
  330-1010`): momentum z / sqrt(prec), positions move with the velocity
  prec * p, the U-turn memory holds velocities and the generalized U-turn
  checks take momentum sums against velocities.  Returns (next_state,
  results).
  """
  draws = draws or orng.PhiloxDraws()
  q = np.asarray(q, np.float64)
  c, d = q.shape
  eps0 = np.broadcast_to(np.asarray(step_size, np.float64), (c,))
  write, read = nuts_instructions(max_depth)
  mem = int(write.max()) + 1
  neg_inf = -np.inf
  prec = None if precision is None else np.broadcast_to(
      np.asarray(precision, np.float64), (c, d))
  vel = (lambda p: p) if prec is None else (lambda p: prec * p)

  # _start_trajectory_batched (`nuts.py:512-539`,
  # `preconditioned_nuts.py:525-545`)
  p0 = draws.momentum(seed, c, d).astype(np.float64)
  if prec is not None:
    p0 = p0 / np.sqrt(prec)
  h0 = compute_hamiltonian(t, p0, prec)

  ends = {  # [2, C, ...]: 0 = left, 1 = right  (`nuts.py:342-350`)
      'p': np.stack([p0, p0]), 'q': np.stack([q, q]),
      't': np.stack([t, t]), 'g': np.stack([g, g])}
  cand = {'q': q.copy(), 't': t.copy(), 'g': g.copy(), 'e': h0.copy(),
          'w': np.zeros(c)}
  is_acc = np.zeros(c, bool)
  msum = p0.copy()
  eds = np.zeros(c)
  lf = np.zeros(c, np.int64)
  cont = np.ones(c, bool)
  notdiv = np.ones(c, bool)
  depth_reached = np.zeros(c, np.int64)
  mem_p = np.zeros((mem, c, d))
  mem_c = np.zeros((mem, c, d))

  depth = 0
  while depth < max_depth and cont.any():       # `nuts.py:404-422`
    # _loop_tree_doubling (`nuts.py:541-711`)
    direction, tree_u = draws.tree(seed, c, depth)
    sel = lambda a, b, m=direction: np.where(
        m.reshape((-1,) + (1,) * (a.ndim - 1)), a, b)
    fr = {k: sel(v[1], v[0]) for k, v in ends.items()}
    eps = np.where(direction, eps0, -eps0)

    # _build_sub_tree (`nuts.py:713-791`)
    sc = {'q': fr['q'].copy(), 't': fr['t'].copy(), 'g': fr['g'].copy(),
          'e': fr['t'].copy(), 'w': np.full(c, neg_inf)}
    eds_sub = np.zeros(c)
    cs = np.zeros((c, d))
    lf_sub = np.zeros(c, np.int64)
    c_sub = cont.copy()
    nd_sub = notdiv.copy()
    depth_reached = np.where(cont, depth + 1, depth_reached)
    i = 0
    while i < (1 << depth) and c_sub.any():
      # _loop_build_sub_tree (`nuts.py:793-946`)
      p1, q1, t1, g1 = leapfrog(target, eps, unrolled_leapfrog_steps,
                                fr['p'], fr['q'], fr['t'], fr['g'], prec)
      v1 = vel(p1)
      fr = {'p': p1, 'q': q1, 't': t1, 'g': g1}
      cs_prev = cs
      cs = cs + p1
      lf_sub = np.where(c_sub, lf_sub + 1, lf_sub)
      no_u = np.ones(c, bool)                   # `nuts.py:949-984`
      j = int(read[i, 0])
      while j < read[i, 1] and no_u.any():
        diff = cs - mem_c[j]
        dl = np.sum(diff * mem_p[j], -1)
        dr = np.sum(diff * v1, -1)
        no_u = no_u & (dl >= 0) & (dr >= 0)
        j += 1
      mem_p[write[i]] = v1
      mem_c[write[i]] = cs_prev
      with np.errstate(invalid='ignore'):
        energy = compute_hamiltonian(t1, p1, prec)
      energy = np.where(np.isnan(energy), neg_inf, energy)
      with np.errstate(invalid='ignore'):
        de = energy - h0
        not_div = -de < max_energy_diff
        wsum = log_add_exp(sc['w'], de)
        thresh = de - wsum
        u = np.log1p(-draws.leaf(seed, c, depth, i).astype(np.float64))
        acc = u <= thresh
      for k, v in (('q', q1), ('t', t1), ('g', g1), ('e', energy)):
        sc[k] = sel(v, sc[k], acc)
      sc['w'] = wsum
      cnt = not_div & c_sub
      cnt_next = no_u & cnt
      nd_keep = np.where(c_sub, not_div, True)
      with np.errstate(over='ignore', invalid='ignore'):
        eds_sub = np.where(cnt, eds_sub + np.exp(np.minimum(de, 0.0)), eds_sub)
      nd_sub = nd_sub & nd_keep
      c_sub = cnt_next
      i += 1

    eds = eds + eds_sub
    tree_w = np.where(c_sub, sc['w'], neg_inf)
    with np.errstate(invalid='ignore'):
      wsum = log_add_exp(tree_w, cand['w'])
      thresh = tree_w - cand['w']
    thresh = np.where(np.isnan(thresh), 0.0, thresh)
    with np.errstate(invalid='ignore'):
      u = np.log1p(-tree_u.astype(np.float64))
    choose = (u <= thresh) & c_sub
    for k in ('q', 't', 'g', 'e'):
      cand[k] = sel(sc[k], cand[k], choose)
    cand['w'] = wsum
    other = {k: sel(v[0], v[1]) for k, v in ends.items()}
    ends = {k: np.stack([sel(other[k], fr[k]), sel(fr[k], other[k])])
            for k in ends}
    msum = msum + cs
    no_u_traj = ((np.sum(msum * vel(ends['p'][0]), -1) >= 0)
                 & (np.sum(msum * vel(ends['p'][1]), -1) >= 0))
    is_acc = is_acc | choose
    cont = c_sub & no_u_traj
    notdiv = nd_sub
    lf = lf + lf_sub
    depth += 1

  with np.errstate(divide='ignore', invalid='ignore'):
    lar = np.log(eds / lf.astype(np.float64))
  return cand['q'], NUTSResults(
      target_log_prob=cand['t'], grads_target_log_prob=cand['g'],
      log_accept_ratio=lar, leapfrogs_taken=lf * unrolled_leapfrog_steps,
      is_accepted=is_acc, reach_max_depth=cont, has_divergence=~notdiv,
      energy=cand['e'], tree_depth=depth_reached)


# --------------------------------------------------------------------------
# Dual averaging  (`mcmc/dual_averaging_step_size_adaptation.py:353-610`)

DAState = collections.namedtuple('DAState', [
    'step_size', 'error_sum', 'log_averaging_step', 'log_shrinkage_target',
    'step'])


def da_bootstrap(step_size, shrinkage_target=None):
  step_size = np.asarray(step_size, np.float64)
  if shrinkage_target is None:
    lst = np.log(10.0) + np.log(step_size)
  else:
    lst = np.log(np.asarray(shrinkage_target, np.float64))
  return DAState(step_size=step_size, error_sum=None,
                 log_averaging_step=np.zeros_like(step_size),
                 log_shrinkage_target=lst, step=0)


def log_accept_prob(log_accept_ratio):
  """`simple_step_size_adaptation.py:42-48`."""
  lar = np.asarray(log_accept_ratio, np.float64)
  return np.minimum(np.where(np.isfinite(lar), lar, -np.inf), 0.0)


def _differing_dims(a_shape, b_shape):
  return [i for i, (x, y) in enumerate(zip(a_shape, b_shape[:len(a_shape)]))
          if x != y]


def da_update(state, log_accept_ratio, state_rank, num_adaptation_steps,
              target_accept_prob=0.75, exploration_shrinkage=0.05,
              step_count_smoothing=10.0, decay_rate=0.75,
              reduce_fn=reduce_logmeanexp):
  """One `_one_step_part` update for a single step-size part.

  `log_accept_ratio` has the chain shape; `state_rank` is the rank of the
  corresponding state part (`dual_averaging...py:399-427` reduction rules).
  """
  lap = log_accept_prob(log_accept_ratio)
  step_size = np.asarray(state.step_size, np.float64)
  num_reduce = min(lap.ndim, state_rank - step_size.ndim)
  red = reduce_fn(lap, axis=tuple(range(num_reduce)), keepdims=False)
  idx = _differing_dims(np.shape(red), step_size.shape)
  red = reduce_fn(red, axis=tuple(idx), keepdims=True) if idx else red
  err = 0.0 if state.error_sum is None else state.error_sum
  new_err = err + target_accept_prob - np.exp(red)
  lst = np.asarray(state.log_shrinkage_target)
  pad = max(lst.ndim - np.ndim(new_err), 0)
  new_err_ext = np.reshape(new_err, np.shape(new_err) + (1,) * pad)
  t = float(state.step) + 1.0
  soft_t = step_count_smoothing + t
  new_log_step = lst - new_err_ext * np.sqrt(t) / (soft_t * exploration_shrinkage)
  eta = t ** (-decay_rate)
  new_log_avg = eta * new_log_step + (1.0 - eta) * state.log_averaging_step
  step = state.step + 1
  if step < num_adaptation_steps:
    new_step = np.exp(new_log_step)
  elif step > num_adaptation_steps:
    new_step = step_size
  else:
    new_step = np.exp(new_log_avg)
  if step > num_adaptation_steps:
    new_log_avg = state.log_averaging_step
  return DAState(step_size=new_step, error_sum=new_err,
                 log_averaging_step=new_log_avg,
                 log_shrinkage_target=state.log_shrinkage_target, step=step)


def da_update_grouped(state, log_accept_ratio, groups, n_groups,
                      num_adaptation_steps, target_accept_prob=0.75,
                      exploration_shrinkage=0.05, step_count_smoothing=10.0,
                      decay_rate=0.75, reduce_fn=reduce_logmeanexp):
  """DA for step sizes of shape [n_groups] where chain c belongs to groups[c].

  Same arithmetic as `da_update`; the reduction (log-mean-exp by default)
  runs over each group's chains (one problem's chains in a batched launch).
  """
  lap = log_accept_prob(log_accept_ratio)
  groups = np.asarray(groups)
  red = np.array([reduce_fn(lap[groups == k]) for k in range(n_groups)])
  err = 0.0 if state.error_sum is None else state.error_sum
  new_err = err + np.asarray(target_accept_prob) - np.exp(red)
  t = float(state.step) + 1.0
  new_log_step = (state.log_shrinkage_target - new_err * np.sqrt(t)
                  / ((step_count_smoothing + t) * exploration_shrinkage))
  eta = t ** (-decay_rate)
  new_log_avg = eta * new_log_step + (1.0 - eta) * state.log_averaging_step
  step = state.step + 1
  if step < num_adaptation_steps:
    new_step = np.exp(new_log_step)
  elif step > num_adaptation_steps:
    new_step = np.asarray(state.step_size, np.float64)
  else:
    new_step = np.exp(new_log_avg)
  if step > num_adaptation_steps:
    new_log_avg = state.log_averaging_step
  return DAState(step_size=new_step, error_sum=new_err,
                 log_averaging_step=new_log_avg,
                 log_shrinkage_target=state.log_shrinkage_target, step=step)


# --------------------------------------------------------------------------
# Simple step-size adaptation  (`mcmc/simple_step_size_adaptation.py:344-438`)

def simple_step_update(step_size, log_accept_ratio, state_rank, step,
                       num_adaptation_steps, target_accept_prob=0.75,
                       adaptation_rate=0.01, reduce_fn=reduce_logmeanexp):
  """New value of one step-size part after the inner step number `step`.

  The chains' log accept probabilities are reduced over the leading
  min(rank(lap), state_rank - rank(step)) dimensions, then over the
  dimensions where they and the step size differ (kept); the reduced value
  aligns with the step size from the left (`mcmc_util.choose`).  Above
  log(target) the step grows by (1 + rate), otherwise it shrinks by it; from
  step num_adaptation_steps on it stays (`:410-436`).
  """
  lap = log_accept_prob(log_accept_ratio)
  step_size = np.asarray(step_size, np.float64)
  num_reduce = min(lap.ndim, state_rank - step_size.ndim)
  red = reduce_fn(lap, axis=tuple(range(num_reduce)), keepdims=False)
  idx = _differing_dims(np.shape(red), step_size.shape)
  red = reduce_fn(red, axis=tuple(idx), keepdims=True) if idx else red
  red = np.asarray(red)
  red = np.reshape(red, red.shape + (1,) * max(step_size.ndim - red.ndim, 0))
  up = red > np.log(np.asarray(target_accept_prob, np.float64))
  opr = 1.0 + adaptation_rate
  new = np.where(up, step_size * opr, step_size / opr)
  return new if step < num_adaptation_steps else step_size


# --------------------------------------------------------------------------
# sample_chain drivers  (`mcmc/sample.py:311-383`)


def step_seeds(seed, n_steps):
  """Per-step seeds of a sampling loop: fold_in(sanitize(seed, salt), t)."""
  base = orng.sanitize_seed(seed, salt='mcmc.sample_chain')
  return orng.fold_in_many(base, np.arange(n_steps))


def sample_hmc(target, init, step_size, num_leapfrog_steps, num_results,
               num_burnin_steps=0, seed=0, dtype=np.float64):
  """sample_chain(HamiltonianMonteCarlo): returns (states [R, C, D], is_acc
  [R, C], lar [R, C])."""
  draws = orng.PhiloxDraws(dtype)
  q = np.asarray(init, np.float64)
  t, g = target.log_prob_and_grad(q)
  seeds = step_seeds(seed, num_burnin_steps + num_results)
  out_q, out_a, out_l = [], [], []
  for s in range(num_burnin_steps + num_results):
    q, r = hmc_one_step(target, step_size, num_leapfrog_steps, q, t, g,
                        seeds[s], draws)
    t, g = r.target_log_prob, r.grads_target_log_prob
    if s >= num_burnin_steps:
      out_q.append(q)
      out_a.append(r.is_accepted)
      out_l.append(r.log_accept_ratio)
  return np.stack(out_q), np.stack(out_a), np.stack(out_l)


def sample_nuts_da(target, init_x, step_size, num_adaptation_steps,
                   num_results, num_burnin_steps, seed, target_accept_prob=0.75,
                   max_depth=10, dtype=np.float64, groups=None, n_groups=1,
                   problem=0):
  """sample_chain(DualAveraging(TTK(NUTS))) on an unconstrained target.

  `target` is already the transformed (unconstrained) target; returns the
  unconstrained states [R, C, D] and per-step results.
  """
  draws = orng.PhiloxDraws(dtype, problem=problem)
  x = np.asarray(init_x, np.float64)
  c = x.shape[0]
  groups = np.zeros(c, np.int64) if groups is None else np.asarray(groups)
  t, g = target.log_prob_and_grad(x)
  da = da_bootstrap(np.broadcast_to(np.asarray(step_size, np.float64),
                                    (n_groups,)))
  seeds = step_seeds(seed, num_burnin_steps + num_results)
  out_x, trace = [], []
  for s in range(num_burnin_steps + num_results):
    eps = np.asarray(da.step_size)[groups]
    x, r = nuts_one_step(target, eps, x, t, g, seeds[s], draws,
                         max_depth=max_depth)
    t, g = r.target_log_prob, r.grads_target_log_prob
    da = da_update_grouped(da, r.log_accept_ratio, groups, n_groups,
                           num_adaptation_steps, target_accept_prob)
    if s >= num_burnin_steps:
      out_x.append(x)
      trace.append(r)
  return np.stack(out_x), trace, da


# --------------------------------------------------------------------------
# Windowed adaptation  (`experimental/mcmc/windowed_sampling.py:329-500,
#                       603-961`, `diagonal_mass_matrix_adaptation.py`)


def reduce_log_harmonic_mean_exp(x, axis=None, keepdims=False):
  """`math/generic.py:387-432`: -reduce_logmeanexp(-x)."""
  return -reduce_logmeanexp(-np.asarray(x, np.float64), axis=axis,
                            keepdims=keepdims)


def window_sizes(num_adaptation_steps):
  """`_get_window_sizes` (`windowed_sampling.py:367-395`): first fast
  window, first slow window, last fast window."""
  slow = num_adaptation_steps // 20
  first = 3 * slow
  last = num_adaptation_steps - 15 * slow - first
  return first, slow, last


def window_option(step, num_adaptation_steps):
  """Which update `WindowedAdaptation.one_step` applies before the inner
  step `step` (`windowed_sampling.py:471-494`): 0 none, 1 first fast window
  start, 2 first slow window start, 3 later slow window start, 4 last
  window start; coinciding starts add up and the sum is clipped to 4
  (`mcmc_util.choose_from`, `mcmc/internal/util.py:176-195`)."""
  first, slow, _ = window_sizes(num_adaptation_steps)
  opt = int(step == 0)
  opt += 2 * int(step == first)
  opt += 3 * int(any(step == first + slow * (2 ** i - 1) for i in (1, 2, 3)))
  opt += 4 * int(step == first + (2 ** 4 - 1) * slow)
  return min(opt, 4)


WindowedState = collections.namedtuple('WindowedState', [
    'da', 'da_num_adaptation_steps', 'running', 'dmma_step',
    'num_estimation_steps', 'precision'])


def windowed_bootstrap(step_size, num_adaptation_steps, dim, n_groups=1):
  """Bootstrap of WindowedAdaptation(DiagonalMassMatrixAdaptation(
  DualAveraging(...))) for step sizes [n_groups] and precisions
  [n_groups, dim]: unit running variance with zero samples
  (`windowed_sampling.py:908-913`), DMMA estimating over
  num_adaptation_steps (`:323-327`), identity momentum
  (`_init_momentum`, `:578-585`)."""
  from oracle import reducers
  step = np.broadcast_to(np.asarray(step_size, np.float64), (n_groups,))
  running = reducers.Moments((n_groups, dim), 1, 1)
  return WindowedState(
      da=da_bootstrap(step.copy()),
      da_num_adaptation_steps=num_adaptation_steps, running=running,
      dmma_step=0, num_estimation_steps=num_adaptation_steps,
      precision=np.ones((n_groups, dim)))


def windowed_pre_step(ws, step, num_adaptation_steps, dim, n_groups=1):
  """The window update of `WindowedAdaptation.one_step` before inner step
  `step` (`windowed_sampling.py:415-469`): resets of the dual averaging
  (keeping its current step size, `_bootstrap_from_inner_results`,
  `dual_averaging_step_size_adaptation.py:543-610`) and of the running
  variance, and the window lengths they run for."""
  from oracle import reducers
  first, slow, last = window_sizes(num_adaptation_steps)
  opt = window_option(step, num_adaptation_steps)
  if opt == 1:
    return ws._replace(da_num_adaptation_steps=first + slow,
                       num_estimation_steps=-1)
  if opt == 2:
    return ws._replace(dmma_step=0, num_estimation_steps=slow)
  if opt in (3, 4):
    da = da_bootstrap(np.asarray(ws.da.step_size, np.float64))
    if opt == 3:
      cur = step - first + slow
      return ws._replace(da=da, da_num_adaptation_steps=cur,
                         running=reducers.Moments((n_groups, dim), 1, 1),
                         dmma_step=0, num_estimation_steps=cur)
    return ws._replace(da=da, da_num_adaptation_steps=last)
  return ws


def windowed_post_step(ws, x, log_accept_ratio, groups, n_groups,
                       target_accept_prob):
  """After the inner step: dual averaging with the harmonic-mean reduction
  over each group's chains (`windowed_sampling.py:869-873`), then the
  DiagonalMassMatrixAdaptation bookkeeping
  (`diagonal_mass_matrix_adaptation.py:186-244`): the chains' new states
  enter the group's running variance while step <= num_estimation_steps,
  and at step == num_estimation_steps the momentum precision becomes that
  variance (population variance, ddof 0)."""
  from oracle import reducers
  da = da_update_grouped(ws.da, log_accept_ratio, groups, n_groups,
                         ws.da_num_adaptation_steps, target_accept_prob,
                         reduce_fn=reduce_log_harmonic_mean_exp)
  step = ws.dmma_step + 1
  running, prec = ws.running, ws.precision
  if step <= ws.num_estimation_steps:
    running = _grouped_moments_update(running, x, groups, n_groups)
  if step == ws.num_estimation_steps:
    prec = reducers.covariance(running)
  return ws._replace(da=da, running=running, dmma_step=step, precision=prec)


def _grouped_moments_update(m, x, groups, n_groups):
  """RunningVariance.update of each group by its chains' states
  (`diagonal_mass_matrix_adaptation.py:200-228`: the chain dimensions are
  flattened into one chunk of observations)."""
  from oracle import reducers
  out = reducers.Moments(m.mean.shape, 1, 1)
  nums_in = np.broadcast_to(np.asarray(m.num_samples, np.float64).reshape(-1),
                            (n_groups,))
  means, ssrs, nums = [], [], []
  for k in range(n_groups):
    sub = reducers.Moments((x.shape[1],), 1, 1)
    sub.num_samples = float(nums_in[k])
    sub.mean = m.mean[k].copy()
    sub.ssr = m.ssr[k].copy()
    sub = reducers.update(sub, x[groups == k], axis=0)
    means.append(sub.mean)
    ssrs.append(sub.ssr)
    nums.append(sub.num_samples)
  out.mean = np.stack(means)
  out.ssr = np.stack(ssrs)
  out.num_samples = np.asarray(nums, np.float64)[:, None]
  return out


def sample_windowed(target, init_x, step_size, num_adaptation_steps,
                    num_results, seed, kind='nuts', target_accept_prob=None,
                    max_depth=10, num_leapfrog_steps=10, dtype=np.float64,
                    groups=None, n_groups=1, problem=0):
  """windowed_adaptive_nuts / windowed_adaptive_hmc
  (`windowed_sampling.py:603-961`) on an unconstrained target, discarding
  the tuning draws: sample_chain of WindowedAdaptation(
  DiagonalMassMatrixAdaptation(DualAveraging(Preconditioned{NUTS,HMC})))
  with num_burnin_steps = num_adaptation_steps.  Returns the states
  [R, C, D], the per-step inner results, the final windowed state and a
  dict with every step's state [S, C, D] and the precision [S, G, D] each
  step ran with."""
  if target_accept_prob is None:
    target_accept_prob = 0.85 if kind == 'nuts' else 0.75
  draws = orng.PhiloxDraws(dtype, problem=problem)
  x = np.asarray(init_x, np.float64)
  c, d = x.shape
  groups = np.zeros(c, np.int64) if groups is None else np.asarray(groups)
  t, g = target.log_prob_and_grad(x)
  ws = windowed_bootstrap(step_size, num_adaptation_steps, d, n_groups)
  n_total = num_adaptation_steps + num_results
  seeds = step_seeds(seed, n_total)
  out_x, trace, states, precs = [], [], [], []
  for s in range(n_total):
    ws = windowed_pre_step(ws, s, num_adaptation_steps, d, n_groups)
    precs.append(ws.precision)
    eps = np.asarray(ws.da.step_size)[groups]
    prec = ws.precision[groups]
    if kind == 'nuts':
      x, r = nuts_one_step(target, eps, x, t, g, seeds[s], draws,
                           max_depth=max_depth, precision=prec)
    else:
      x, r = hmc_one_step(target, eps, num_leapfrog_steps, x, t, g, seeds[s],
                          draws, prec=prec)
    t, g = r.target_log_prob, r.grads_target_log_prob
    ws = windowed_post_step(ws, x, r.log_accept_ratio, groups, n_groups,
                            target_accept_prob)
    if s >= num_adaptation_steps:
      out_x.append(x)
    trace.append(r)
    states.append(x)
  info = {'states': np.stack(states), 'precision_used': np.stack(precs)}
  return ((np.stack(out_x) if out_x else np.zeros((0, c, d))), trace, ws,
          info)
