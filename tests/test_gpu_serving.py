"""Serving loop: requests admitted at different times share launches; each
request's draws equal the same request run alone, and a request's chains
match the oracle's sample_chain(DA(TTK(NUTS)))."""

import copy

import numpy as np
import pytest
import torch

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import bijectors as pbij
from probability_amd import serving
from probability_amd import targets as ptg

pytestmark = pytest.mark.gpu


def _requests(n=6, seed=0, adapt=6, draws=5):
  r = np.random.default_rng(seed)
  reqs = []
  for k in range(n):
    if k % 3 == 2:
      j = int(r.integers(8, 40))
      y = (4 + 3 * r.normal(size=j) + 10 * r.normal(size=j)).astype(np.float32)
      s = r.uniform(5, 20, j).astype(np.float32)
      spec = ptg.HierarchicalNormal(y, s)
      bij = [pbij.Identity(), pbij.Softplus(), pbij.Identity()]
    else:
      nobs, d = int(r.integers(40, 300)), int(r.integers(3, 60))
      x = r.normal(size=(nobs, d)).astype(np.float32)
      b = r.normal(size=d) * 0.5
      if k % 3 == 0:
        y = (r.uniform(size=nobs) < 1 / (1 + np.exp(-(x @ b)))).astype(
            np.float32)
        spec, bij = ptg.LogisticRegression(x, y), None
      else:
        y = (x @ b + 0.5 * r.normal(size=nobs)).astype(np.float32)
        spec, bij = ptg.LinearRegression(x, y), [pbij.Identity(), pbij.Exp()]
    reqs.append(dict(spec=spec, num_chains=4 + k % 3 * 4,
                     num_adaptation_steps=adapt, num_results=draws,
                     seed=100 + k, bijectors=bij, step_size=0.05, key=k))
  return reqs


def test_requests_same_alone_and_in_the_mix():
  cfgs = _requests()
  srv = serving.SamplerServer(iters_per_tick=3)
  reqs = [serving.Request(**c) for c in cfgs]
  # staggered arrivals: half now, the rest after two ticks
  for r in reqs[:3]:
    srv.submit(r)
  srv.tick()
  srv.tick()
  for r in reqs[3:]:
    srv.submit(r)
  while not srv.idle():
    srv.tick()
  for c, r in zip(cfgs, reqs):
    alone = serving.run_alone(serving.Request(**c), iters_per_tick=5)
    np.testing.assert_array_equal(r.draws.cpu().numpy(),
                                  alone.draws.cpu().numpy())
    assert r.final_step_size == alone.final_step_size


def test_served_request_matches_oracle():
  c = _requests(n=2, seed=3)[1]  # linear regression with Exp on sigma
  r = serving.run_alone(serving.Request(**c), iters_per_tick=4)
  spec = c['spec']
  d = spec.dim
  codes = np.zeros(d, np.int32)
  codes[d - 1] = 2
  ot = otg.Transformed(otg.LinearRegression(
      spec.x.double().cpu().numpy(), spec.y.double().cpu().numpy()), codes)
  ox, otrace, _ = omcmc.sample_nuts_da(
      ot, np.zeros((c['num_chains'], d)), np.float32(0.05),
      c['num_adaptation_steps'], c['num_results'], c['num_adaptation_steps'],
      seed=orng.sanitize_seed(c['seed']), target_accept_prob=0.8,
      dtype=np.float32, problem=c['key'])
  ostates = np.stack([ot.forward(v) for v in ox])
  np.testing.assert_allclose(r.draws.cpu().numpy(), ostates, rtol=2e-4,
                             atol=2e-4)


def _generator_requests(n=40, adapt=3, draws=3, dtype=torch.float32):
  """Requests drawn from the serving workload's own generator: its widest
  hierarchical normal (J up to 1024, a 17-slot state) and its smallest
  logistic and linear GLMs (small enough for the oracle), 16 chains each."""
  from probability_amd import workloads
  _, models = workloads.serving_requests(n)
  hier = max((k for k, m in enumerate(models) if m['kind'] == 'hier'),
             key=lambda k: models[k]['y'].size)
  pick = [hier]
  for kind in ('logistic', 'linear'):
    pick.append(min((k for k, m in enumerate(models) if m['kind'] == kind),
                    key=lambda k: models[k]['x'].size))
  out = []
  for k in pick:
    m = models[k]
    if m['kind'] == 'hier':
      spec = ptg.HierarchicalNormal(m['y'], m['sigma'], dtype=dtype)
      bij = [pbij.Identity(), pbij.Softplus(), pbij.Identity()]
      codes = np.zeros(spec.dim, np.int32)
      codes[1] = 1
      ot = otg.Transformed(otg.HierarchicalNormal(
          m['y'].astype(np.float64), m['sigma'].astype(np.float64)), codes)
    elif m['kind'] == 'logistic':
      spec = ptg.LogisticRegression(m['x'], m['y'], dtype=dtype)
      bij = None
      ot = otg.LogisticRegression(m['x'].astype(np.float64),
                                  m['y'].astype(np.float64))
    else:
      spec = ptg.LinearRegression(m['x'], m['y'], dtype=dtype)
      bij = [pbij.Identity(), pbij.Exp()]
      codes = np.zeros(spec.dim, np.int32)
      codes[-1] = 2
      ot = otg.Transformed(otg.LinearRegression(
          m['x'].astype(np.float64), m['y'].astype(np.float64)), codes)
    out.append((dict(spec=spec, num_chains=16, num_adaptation_steps=adapt,
                     num_results=draws, seed=500 + k, bijectors=bij,
                     step_size=0.05, key=k), ot))
  return out


def _oracle_draws(c, ot, dtype, max_depth):
  d = c['spec'].dim
  ox, _, _ = omcmc.sample_nuts_da(
      ot, np.zeros((16, d)), dtype(0.05), c['num_adaptation_steps'],
      c['num_results'], c['num_adaptation_steps'],
      seed=orng.sanitize_seed(c['seed']), target_accept_prob=0.8,
      dtype=dtype, problem=c['key'], max_depth=max_depth)
  return np.stack([ot.forward(v) for v in ox]) if hasattr(ot, 'forward') \
      else ox


def _normwise(got, want):
  return np.abs(got - want).max() / max(np.abs(want).max(), 1.0)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-9),
                                       (torch.float32, 1e-3)])
def test_generator_requests_in_the_mix_match_alone_and_oracle(dtype, tol):
  # float64: the served trajectories equal the oracle's to rounding (a
  # 17-slot float64 state keeps 8 U-turn levels in LDS, so both run at
  # max_tree_depth 8); float32: six NUTS+DA transitions of f32 rounding
  # drift away from the float64-arithmetic oracle on the same f32 draws
  # (measured ~3e-4 normwise at J = 1021)
  cases = _generator_requests(dtype=dtype)
  assert cases[0][0]['spec'].dim > 512  # the widest request is a wide state
  for c, _ in cases:
    c['max_tree_depth'] = 8
  srv = serving.SamplerServer(dtype=dtype, iters_per_tick=2)
  reqs = [serving.Request(**c) for c, _ in cases]
  for r in reqs:
    srv.submit(r)
  while not srv.idle():
    srv.tick()
  ndt = np.float64 if dtype == torch.float64 else np.float32
  for (c, ot), r in zip(cases, reqs):
    alone = serving.run_alone(serving.Request(**c), dtype=dtype,
                              iters_per_tick=3)
    np.testing.assert_array_equal(r.draws.cpu().numpy(),
                                  alone.draws.cpu().numpy())
    err = _normwise(r.draws.double().cpu().numpy(),
                    _oracle_draws(c, ot, ndt, 8))
    print(f'ERR serving {np.dtype(ndt).name} kind{c["spec"].kind} '
          f'd={c["spec"].dim} draws rel={err:.3e}', flush=True)
    assert err <= tol, err
