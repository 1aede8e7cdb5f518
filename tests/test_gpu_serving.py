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
