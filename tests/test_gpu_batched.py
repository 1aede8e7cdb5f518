"""Batched posteriors: one launch for many problems; each problem's chains
equal the oracle's and equal the same problem run alone."""

import numpy as np
import pytest
import torch

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import bijectors as pbij
from probability_amd import targets as ptg
from probability_amd.mcmc import batched

pytestmark = pytest.mark.gpu


def _problems(dtype, n=4, seed=0):
  r = np.random.default_rng(seed)
  specs, bij, oracles = [], [], []
  for k in range(n):
    nobs = int(r.integers(40, 300))
    d = int(r.integers(3, 90))
    x = r.normal(size=(nobs, d)).astype(np.float32)
    beta = r.normal(size=d) * 0.5
    if k % 2 == 0:
      y = (r.uniform(size=nobs) < 1 / (1 + np.exp(-(x @ beta)))).astype(
          np.float32)
      specs.append(ptg.LogisticRegression(x, y, dtype=dtype))
      bij.append(None)
      oracles.append(otg.LogisticRegression(x.astype(np.float64),
                                            y.astype(np.float64)))
    else:
      y = (x @ beta + 0.7 * r.normal(size=nobs)).astype(np.float32)
      specs.append(ptg.LinearRegression(x, y, dtype=dtype))
      bij.append([pbij.Identity(), pbij.Exp()])
      codes = np.zeros(d + 1, np.int32)
      codes[d] = 2
      oracles.append(otg.Transformed(
          otg.LinearRegression(x.astype(np.float64), y.astype(np.float64)),
          codes))
  return specs, bij, oracles


def _run(batch, seed=3):
  return batched.sample_posteriors(batch, num_results=5, num_burnin_steps=6,
                                   num_adaptation_steps=6, step_size=0.05,
                                   target_accept_prob=0.8, seed=seed)


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_batched_equals_each_problem_alone(dtype):
  specs, bij, _ = _problems(dtype, n=5)
  chains = [8, 5, 8, 3, 6]
  mix = _run(batched.PosteriorBatch(specs, chains, bij, slot_dim=128,
                                    waves_per_chain=2))
  for k in range(len(specs)):
    alone = _run(batched.PosteriorBatch([specs[k]], [chains[k]], [bij[k]],
                                        keys=[k], slot_dim=128,
                                        waves_per_chain=2))
    np.testing.assert_array_equal(mix.draws[k].cpu().numpy(),
                                  alone.draws[0].cpu().numpy())
    np.testing.assert_array_equal(mix.leapfrogs_taken[k].cpu().numpy(),
                                  alone.leapfrogs_taken[0].cpu().numpy())
    assert float(mix.step_size[k]) == float(alone.step_size[0])


@pytest.mark.parametrize('dtype,ndt,tol', [(torch.float32, np.float32, 2e-4),
                                           (torch.float64, np.float64, 1e-9)])
def test_batched_matches_oracle(dtype, ndt, tol):
  specs, bij, oracles = _problems(dtype, n=4, seed=1)
  chains = 6
  res = _run(batched.PosteriorBatch(specs, chains, bij))
  for k, ot in enumerate(oracles):
    d = specs[k].dim
    ox, otrace, _ = omcmc.sample_nuts_da(
        ot, np.zeros((chains, d)), np.asarray(0.05, ndt), 6, 5, 6,
        seed=orng.sanitize_seed(3), target_accept_prob=0.8, dtype=ndt,
        problem=k)
    ostates = np.stack([ot.forward(v) if hasattr(ot, 'forward') else v
                        for v in ox])
    lf = np.stack([r.leapfrogs_taken for r in otrace])
    np.testing.assert_array_equal(res.leapfrogs_taken[k].cpu().numpy(), lf)
    np.testing.assert_allclose(res.draws[k].cpu().numpy(), ostates, rtol=tol,
                               atol=tol)


def _ragged_problems(dtype):
  r = np.random.default_rng(11)
  specs, bij = [], []
  # N x D spans the four wave-count classes (<4096, <16384, <262144,
  # >= 262144).
  for nobs, d in ((50, 7), (600, 40), (120, 12), (300, 20), (40, 3),
                  (2700, 100)):
    x = r.normal(size=(nobs, d)).astype(np.float32)
    beta = r.normal(size=d) * 0.3
    y = (r.uniform(size=nobs) < 1 / (1 + np.exp(-(x @ beta)))).astype(
        np.float32)
    specs.append(ptg.LogisticRegression(x, y, dtype=dtype))
    bij.append(None)
  return specs, bij


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_sample_ragged_equals_each_problem_alone(dtype):
  from probability_amd import tuning
  specs, bij = _ragged_problems(dtype)
  chains = [4, 6, 5, 3, 7, 3]
  # per-problem initial step sizes follow their problems into the classes
  steps = [0.05, 0.04, 0.06, 0.05, 0.03, 0.05]
  kw = dict(num_results=4, num_burnin_steps=5, num_adaptation_steps=5,
            target_accept_prob=0.8, seed=9)
  stats = {}
  mix = batched.sample_ragged(specs, chains, bijectors=bij, slot_dim=64,
                              stats=stats, step_size=steps, **kw)
  # each problem's class: its tuned waves per chain at its own chain count
  waves = [tuning.glm_waves(s.kind, s.n_obs, s.dim, chains[k], 'nuts', dtype)
           for k, s in enumerate(specs)]
  assert sorted(w for _, w in stats['classes']) == sorted(set(waves))
  assert len(set(waves)) > 1
  for k, s in enumerate(specs):
    alone = batched.sample_posteriors(
        batched.PosteriorBatch([s], [chains[k]], [bij[k]], keys=[k],
                               slot_dim=64, waves_per_chain=waves[k]),
        step_size=steps[k], **kw)
    np.testing.assert_array_equal(mix.draws[k].cpu().numpy(),
                                  alone.draws[0].cpu().numpy())
    np.testing.assert_array_equal(mix.leapfrogs_taken[k].cpu().numpy(),
                                  alone.leapfrogs_taken[0].cpu().numpy())
    np.testing.assert_array_equal(mix.final_state[k].cpu().numpy(),
                                  alone.final_state[0].cpu().numpy())
    assert float(mix.step_size[k]) == float(alone.step_size[0])


def test_sample_ragged_class_splits_equal_unsplit():
  # a size class dealt by cost to several batches (class_splits), each on
  # its own stream and advancing at its own pace: every problem's draws,
  # trees and tuned step size equal those of the class's single batch
  specs, bij = _ragged_problems(torch.float32)
  specs, bij = specs * 2, bij * 2
  chains = [4, 6, 5, 3, 7, 3] * 2
  kw = dict(num_results=4, num_burnin_steps=5, num_adaptation_steps=5,
            step_size=0.05, target_accept_prob=0.8, seed=9, slot_dim=64)
  one, split = {}, {}
  base = batched.sample_ragged(specs, chains, bijectors=bij, stats=one, **kw)
  splits = {w: 3 for _, w in one['classes']}
  res = batched.sample_ragged(specs, chains, bijectors=bij, stats=split,
                              class_splits=splits, **kw)
  assert len(split['classes']) > len(one['classes'])
  assert sum(n for n, _ in split['classes']) == len(specs)
  for k in range(len(specs)):
    for field in ('draws', 'leapfrogs_taken', 'final_state'):
      np.testing.assert_array_equal(getattr(res, field)[k].cpu().numpy(),
                                    getattr(base, field)[k].cpu().numpy())
    assert float(res.step_size[k]) == float(base.step_size[k])


@pytest.mark.parametrize('adaptation', ['dual_averaging', 'windowed'])
def test_launch_length_leaves_draws_unchanged(adaptation):
  # the sampling phase in launches of 3 transitions or in one launch, with
  # the chain groups of the sampling launches: the same draws, trees, final
  # states and step sizes (every draw is keyed by its iteration)
  specs, bij = _ragged_problems(torch.float32)
  chains = [4, 6, 5, 3, 7, 3]
  kw = dict(num_results=7, num_burnin_steps=40, num_adaptation_steps=40,
            step_size=0.05, target_accept_prob=0.8, seed=13, slot_dim=64,
            adaptation=adaptation, chains_per_block={8: 2, 4: 2, 2: 2})
  one = batched.sample_ragged(specs, chains, bijectors=bij,
                              max_iters_per_launch=4096, **kw)
  short = batched.sample_ragged(specs, chains, bijectors=bij,
                                max_iters_per_launch=3, **kw)
  for k in range(len(specs)):
    for field in ('draws', 'leapfrogs_taken', 'final_state'):
      np.testing.assert_array_equal(getattr(short, field)[k].cpu().numpy(),
                                    getattr(one, field)[k].cpu().numpy())
    assert float(short.step_size[k]) == float(one.step_size[k])


def _glm_group_problems(dtype):
  r = np.random.default_rng(21)
  specs, bij, oracles = [], [], []
  for k, (nobs, d) in enumerate(((700, 40), (2700, 100), (300, 9),
                                 (1500, 60))):
    x = r.normal(size=(nobs, d)).astype(np.float32) * 0.3
    beta = r.normal(size=d) * 0.3
    if k % 2 == 0:
      y = (r.uniform(size=nobs) < 1 / (1 + np.exp(-(x @ beta)))).astype(
          np.float32)
      specs.append(ptg.LogisticRegression(x, y, dtype=dtype))
      bij.append(None)
      oracles.append(otg.LogisticRegression(x.astype(np.float64),
                                            y.astype(np.float64)))
    else:
      y = (x @ beta + 0.5 * r.normal(size=nobs)).astype(np.float32)
      specs.append(ptg.LinearRegression(x, y, dtype=dtype))
      bij.append([pbij.Identity(), pbij.Exp()])
      codes = np.zeros(d + 1, np.int32)
      codes[d] = 2
      oracles.append(otg.Transformed(
          otg.LinearRegression(x.astype(np.float64), y.astype(np.float64)),
          codes))
  return specs, bij, oracles


@pytest.mark.parametrize('group', [2, 4])
@pytest.mark.parametrize('waves', [4, 8])
@pytest.mark.parametrize('dtype,ndt,tol', [(torch.float32, np.float32, 2e-4),
                                           (torch.float64, np.float64, 1e-9)])
def test_chain_groups_match_oracle(group, waves, dtype, ndt, tol):
  # chains of one problem sharing every pass over X (nuts_group_kernel):
  # each chain still follows the oracle's transitions
  specs, bij, oracles = _glm_group_problems(dtype)
  chains = 8
  batch = batched.PosteriorBatch(specs, chains, bij, slot_dim=128,
                                 waves_per_chain=waves,
                                 chains_per_block=group)
  assert batch.chains_per_block == group
  res = _run(batch, seed=4)
  for k, ot in enumerate(oracles):
    d = specs[k].dim
    ox, otrace, _ = omcmc.sample_nuts_da(
        ot, np.zeros((chains, d)), np.asarray(0.05, ndt), 6, 5, 6,
        seed=orng.sanitize_seed(4), target_accept_prob=0.8, dtype=ndt,
        problem=k)
    ostates = np.stack([ot.forward(v) if hasattr(ot, 'forward') else v
                        for v in ox])
    lf = np.stack([r.leapfrogs_taken for r in otrace])
    np.testing.assert_array_equal(res.leapfrogs_taken[k].cpu().numpy(), lf)
    np.testing.assert_allclose(res.draws[k].cpu().numpy(), ostates, rtol=tol,
                               atol=tol)


@pytest.mark.parametrize('dtype,ndt,tol', [(torch.float32, np.float32, 2e-4),
                                           (torch.float64, np.float64, 1e-9)])
def test_group_adaptation_matches_oracle(dtype, ndt, tol):
  # chain groups of 4 also in the single-transition adaptation launches
  # (group_adaptation): each chain still follows the oracle's transitions
  specs, bij, oracles = _glm_group_problems(dtype)
  chains = 8
  batch = batched.PosteriorBatch(specs, chains, bij, slot_dim=128,
                                 waves_per_chain=8, chains_per_block=4,
                                 group_adaptation=True)
  assert batch.group_adaptation
  res = _run(batch, seed=12)
  for k, ot in enumerate(oracles):
    d = specs[k].dim
    ox, otrace, _ = omcmc.sample_nuts_da(
        ot, np.zeros((chains, d)), np.asarray(0.05, ndt), 6, 5, 6,
        seed=orng.sanitize_seed(12), target_accept_prob=0.8, dtype=ndt,
        problem=k)
    ostates = np.stack([ot.forward(v) if hasattr(ot, 'forward') else v
                        for v in ox])
    lf = np.stack([r.leapfrogs_taken for r in otrace])
    np.testing.assert_array_equal(res.leapfrogs_taken[k].cpu().numpy(), lf)
    np.testing.assert_allclose(res.draws[k].cpu().numpy(), ostates, rtol=tol,
                               atol=tol)


def test_chain_groups_alone_equals_mix():
  specs, bij, _ = _glm_group_problems(torch.float32)
  mk = lambda s, b, keys: batched.PosteriorBatch(
      s, 8, b, keys=keys, slot_dim=128, waves_per_chain=8, chains_per_block=4)
  mix = _run(mk(specs, bij, None), seed=8)
  for k in range(len(specs)):
    alone = _run(mk([specs[k]], [bij[k]], [k]), seed=8)
    np.testing.assert_array_equal(mix.draws[k].cpu().numpy(),
                                  alone.draws[0].cpu().numpy())


def _wide_group_problems():
  # float32 groups of 4 take their passes on matrix instructions for up to
  # 128 coefficients (8 slabs of 16), state widths of 2 and 3 slots; row
  # counts that are not multiples of the 16-row sub-chunks
  r = np.random.default_rng(33)
  specs, bij, oracles = [], [], []
  for k, (nobs, d) in enumerate(((1000, 128), (523, 128), (37, 20),
                                 (611, 113))):
    x = r.normal(size=(nobs, d)).astype(np.float32) * 0.2
    beta = r.normal(size=d) * 0.2
    if k % 2 == 0:
      y = (r.uniform(size=nobs) < 1 / (1 + np.exp(-(x @ beta)))).astype(
          np.float32)
      specs.append(ptg.LogisticRegression(x, y, dtype=torch.float32))
      bij.append(None)
      oracles.append(otg.LogisticRegression(x.astype(np.float64),
                                            y.astype(np.float64)))
    else:
      y = (x @ beta + 0.5 * r.normal(size=nobs)).astype(np.float32)
      specs.append(ptg.LinearRegression(x, y, dtype=torch.float32))
      bij.append([pbij.Identity(), pbij.Exp()])
      codes = np.zeros(d + 1, np.int32)
      codes[d] = 2
      oracles.append(otg.Transformed(
          otg.LinearRegression(x.astype(np.float64), y.astype(np.float64)),
          codes))
  return specs, bij, oracles


@pytest.mark.parametrize('waves', [4, 8])
def test_wide_chain_groups_match_oracle(waves):
  specs, bij, oracles = _wide_group_problems()
  chains = 8
  batch = batched.PosteriorBatch(specs, chains, bij, slot_dim=192,
                                 waves_per_chain=waves, chains_per_block=4)
  assert batch.chains_per_block == 4
  res = _run(batch, seed=6)
  for k, ot in enumerate(oracles):
    d = specs[k].dim
    ox, otrace, _ = omcmc.sample_nuts_da(
        ot, np.zeros((chains, d)), np.asarray(0.05, np.float32), 6, 5, 6,
        seed=orng.sanitize_seed(6), target_accept_prob=0.8, dtype=np.float32,
        problem=k)
    ostates = np.stack([ot.forward(v) if hasattr(ot, 'forward') else v
                        for v in ox])
    lf = np.stack([r.leapfrogs_taken for r in otrace])
    np.testing.assert_array_equal(res.leapfrogs_taken[k].cpu().numpy(), lf)
    got = res.draws[k].double().cpu().numpy()
    err = np.abs(got - ostates).max() / max(np.abs(ostates).max(), 1.0)
    print(f'ERR wide group problem {k} N={specs[k].n_obs} D={d} W={waves} '
          f'draws rel={err:.3e}', flush=True)
    assert err <= 2e-4, (k, err)
