"""Independent recursive NUTS used to pin the oracle's tree termination.

Hoffman & Gelman (2014) Algorithm 3 (recursive BuildTree), with Betancourt
(2017, App. A) generalized U-turn on momentum sums and a max_energy_diff
divergence stop, written as plain recursion over one chain.  Given the same
momentum and the same doubling directions, its leapfrog count, depth,
divergence and reach-max-depth flags must equal the iterative batched
algorithm's (those do not depend on the multinomial uniforms).  The
candidate selection is pinned separately against its exact distribution
(`selection_distribution`).

Test infrastructure only.
"""

import numpy as np


def _leapfrog(target, eps, q, p, g):
  p = p + 0.5 * eps * g
  q = q + eps * p
  with np.errstate(all='ignore'):
    t, g = target.log_prob_and_grad(q[None])
  t, g = t[0], g[0]
  p = p + eps * g
  p = p - 0.5 * eps * g
  return q, p, g, t


def _build(target, q, p, g, eps, depth, h0, max_de, leaves):
  """Returns (ok, q_end, p_end, g_end, rho, p_first, p_last)."""
  if depth == 0:
    q, p, g, t = _leapfrog(target, eps, q, p, g)
    with np.errstate(invalid='ignore'):
      h = t - 0.5 * np.sum(p * p)
    if np.isnan(h):
      h = -np.inf
    leaves.append((q.copy(), h - h0))
    ok = -(h - h0) < max_de
    return ok, q, p, g, p.copy(), p.copy(), p.copy()
  ok1, q1, p1, g1, r1, pf1, pl1 = _build(target, q, p, g, eps, depth - 1, h0,
                                         max_de, leaves)
  if not ok1:
    return False, q1, p1, g1, r1, pf1, pl1
  ok2, q2, p2, g2, r2, pf2, pl2 = _build(target, q1, p1, g1, eps, depth - 1,
                                         h0, max_de, leaves)
  rho = r1 + r2
  if not ok2:
    return False, q2, p2, g2, rho, pf1, pl2
  ok = (np.dot(rho, pf1) >= 0) and (np.dot(rho, pl2) >= 0)
  return ok, q2, p2, g2, rho, pf1, pl2


def termination(target, q, t, g, p0, directions, eps, max_depth=10,
                max_energy_diff=1000.0):
  """One chain: returns dict(leapfrogs, depth, reach_max_depth,
  has_divergence, subtrees) given momentum p0 and directions[depth]."""
  h0 = t - 0.5 * np.sum(p0 * p0)
  ends = {False: (q, p0, g), True: (q, p0, g)}
  rho = p0.copy()
  leapfrogs = 0
  diverged = False
  subtrees = []
  cont = True
  depth = 0
  while depth < max_depth and cont:
    d = bool(directions[depth])
    e = eps if d else -eps
    qs, ps, gs = ends[d]
    leaves = []
    ok, qe, pe, ge, r, _, _ = _build(target, qs, ps, gs, e, depth, h0,
                                     max_energy_diff, leaves)
    leapfrogs += len(leaves)
    if leaves and not (-leaves[-1][1] < max_energy_diff):
      diverged = True
    depth += 1
    subtrees.append((ok, leaves))
    if not ok:
      cont = False
      break
    ends[d] = (qe, pe, ge)
    rho = rho + r
    cont = (np.dot(rho, ends[False][1]) >= 0) and (np.dot(rho, ends[True][1]) >= 0)
  return dict(leapfrogs=leapfrogs, depth=depth, reach_max_depth=cont,
              has_divergence=diverged, subtrees=subtrees)


def selection_distribution(subtrees, q0):
  """Exact distribution of the returned state under multinomial sampling
  with biased progressive acceptance of each completed subtree.

  Returns a list of (state, probability)."""
  states = [np.asarray(q0)]
  probs = np.array([1.0])
  w_old = 0.0
  for ok, leaves in subtrees:
    if not ok:
      break
    de = np.array([l[1] for l in leaves])
    m = de.max()
    w_new = m + np.log(np.sum(np.exp(de - m)))
    q_sub = np.exp(de - w_new)
    a = min(1.0, np.exp(w_new - w_old))
    probs = np.concatenate([(1 - a) * probs, a * q_sub])
    states.extend(l[0] for l in leaves)
    w_old = np.logaddexp(w_old, w_new)
  return list(zip(states, probs))
