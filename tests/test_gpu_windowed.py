"""GPU checks of windowed adaptation and the preconditioned kernels
(experimental/mcmc/windowed_sampling.py, diagonal_mass_matrix_adaptation.py,
preconditioned_nuts.py, preconditioned_hmc.py) against the oracle.

Tolerances: one preconditioned transition, float32 kernels vs the float64
oracle on float32-rounded inputs <= 2e-4 relative on states (a whole
trajectory of leapfrogs), float64 <= 1e-10; leapfrog counts, accepts and
divergences identical.  Short windowed runs in float64: identical leapfrog
counts and accepts at every step, step sizes and mass matrices <= 1e-8.
"""

import numpy as np
import pytest
import torch

from oracle import mcmc as om
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import _lib
from probability_amd import bijectors as pbij
from probability_amd import distributions as pd
from probability_amd import targets as ptg
from probability_amd.experimental import mcmc as emcmc
from probability_amd.mcmc import _runner
from probability_amd.mcmc import batched

pytestmark = pytest.mark.gpu

DTYPES = [(torch.float32, np.float32, 2e-4), (torch.float64, np.float64, 1e-10)]


def _np(t):
  return t.detach().cpu().numpy()


@pytest.mark.parametrize('tdt,ndt', [(torch.float32, np.float32),
                                     (torch.float64, np.float64)])
def test_init_uniform_draws_match_oracle(tdt, ndt):
  seed = orng.sanitize_seed(77)
  for problem in (0, 3):
    out = torch.empty((9, 13), dtype=tdt, device='cuda')
    _lib.check(_lib.lib().hipmc_draws(_lib.dtype_code(tdt), 4,
                                      seed.ctypes.data, problem, 9, 13, 0, 0,
                                      _lib.ptr(out), None, 0))
    ref = orng.init_uniform(seed, np.arange(9), 13, ndt, problem)
    np.testing.assert_allclose(_np(out), ref, rtol=2 * np.finfo(ndt).eps,
                               atol=4 * np.finfo(ndt).eps)
    assert _np(out).min() > -2 and _np(out).max() < 2


def test_da_harmonic_matches_oracle():
  r = np.random.default_rng(0)
  sizes = [3, 5, 8]
  c = sum(sizes)
  groups = np.repeat(np.arange(3), sizes)
  rows = 4
  lar = r.normal(-0.5, 0.7, size=(rows, c))
  lar[1, 2] = -np.inf
  lar[2, 7] = np.nan
  step0 = np.array([0.3, 0.05, 1.2])
  st = om.da_bootstrap(step0)
  for k in range(rows):
    st = om.da_update_grouped(st, lar[k], groups, 3, 10, 0.85,
                              reduce_fn=om.reduce_log_harmonic_mean_exp)
  dev = 'cuda'
  gstart = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(
      np.int32), device=dev)
  t_lar = torch.as_tensor(lar, device=dev)
  step = torch.as_tensor(step0, device=dev)
  step64 = step.clone()
  err = torch.zeros(3, dtype=torch.float64, device=dev)
  avg = torch.zeros_like(err)
  lst = np.log(10.0) + torch.log(step64)
  tgt = torch.full((3,), 0.85, dtype=torch.float64, device=dev)
  _lib.check(_lib.lib().hipmc_da_update(
      _lib.dtype_code(torch.float64), 3, None, None,
      _lib.ptr(gstart), _lib.ptr(t_lar), rows, c, _lib.ptr(err), _lib.ptr(avg),
      _lib.ptr(lst), _lib.ptr(tgt), 0, 10, 0.05, 10.0, 0.75,
      _lib.REDUCE_HARMONIC, None, None, _lib.ptr(step), _lib.ptr(step64),
      0))
  np.testing.assert_allclose(_np(step), st.step_size, rtol=1e-12)
  np.testing.assert_allclose(_np(err), st.error_sum, rtol=1e-12, atol=1e-14)
  np.testing.assert_allclose(_np(avg), st.log_averaging_step, rtol=1e-12)


def test_da_update_with_per_element_step_counts():
  # elements at different DA steps and adaptation lengths in one launch
  # equal one launch per element with the scalar step and length
  r = np.random.default_rng(4)
  sizes = [4, 6, 5]
  c = sum(sizes)
  lar = np.minimum(0, r.normal(-0.4, 0.5, size=c))
  steps, lens = [0, 7, 3], [10, 8, 3]
  dev = 'cuda'
  gstart = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(
      np.int32), device=dev)
  t_lar = torch.as_tensor(lar, device=dev)

  def fresh():
    step = torch.as_tensor([0.3, 0.05, 1.2], dtype=torch.float64,
                           device=dev)
    err = torch.as_tensor([0.0, 0.2, -0.1], dtype=torch.float64, device=dev)
    avg = torch.as_tensor([0.0, -1.0, 0.3], dtype=torch.float64, device=dev)
    lst = np.log(10.0) + torch.log(step)
    tgt = torch.full((3,), 0.8, dtype=torch.float64, device=dev)
    return step, step.clone(), err, avg, lst, tgt

  step, step64, err, avg, lst, tgt = fresh()
  es = torch.as_tensor(steps, dtype=torch.int64, device=dev)
  el = torch.as_tensor(lens, dtype=torch.int64, device=dev)
  _lib.check(_lib.lib().hipmc_da_update(
      _lib.dtype_code(torch.float64), 3, None, None, _lib.ptr(gstart),
      _lib.ptr(t_lar), 1, c, _lib.ptr(err), _lib.ptr(avg), _lib.ptr(lst),
      _lib.ptr(tgt), 0, 0, 0.05, 10.0, 0.75, _lib.REDUCE_MEAN, _lib.ptr(es),
      _lib.ptr(el), _lib.ptr(step), _lib.ptr(step64), 0))
  for k in range(3):
    s2, s64, e2, a2, l2, t2 = fresh()
    idx = torch.as_tensor([k], dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().hipmc_da_update(
        _lib.dtype_code(torch.float64), 1, _lib.ptr(idx), _lib.ptr(idx),
        _lib.ptr(gstart), _lib.ptr(t_lar), 1, c, _lib.ptr(e2), _lib.ptr(a2),
        _lib.ptr(l2), _lib.ptr(t2), steps[k], lens[k], 0.05, 10.0, 0.75,
        _lib.REDUCE_MEAN, None, None, _lib.ptr(s2), _lib.ptr(s64), 0))
    assert float(step[k]) == float(s2[k])
    assert float(err[k]) == float(e2[k])
    assert float(avg[k]) == float(a2[k])
  # the element past its adaptation (step 3 of 3) keeps its step size
  assert float(step[2]) == 1.2


@pytest.mark.parametrize('tdt', [torch.float32, torch.float64])
def test_variance_update_matches_pooled_variance(tdt):
  r = np.random.default_rng(1)
  sizes = [3, 5, 8]
  c, stride, dim = sum(sizes), 10, 7
  gstart = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
  perm = r.permutation(c).astype(np.int32)   # chain c lives in row perm[c]
  dev = 'cuda'
  n = torch.zeros(3, dtype=torch.float64, device=dev)
  mean = torch.zeros((3, stride), dtype=torch.float64, device=dev)
  ssr = torch.zeros_like(mean)
  seen = [[] for _ in range(3)]
  gs_t = torch.as_tensor(gstart, device=dev)
  perm_t = torch.as_tensor(perm, device=dev)
  for _ in range(3):
    x = r.normal(size=(c, stride)) * r.uniform(0.1, 5, size=stride)
    xt = torch.as_tensor(x, dtype=tdt, device=dev)
    xr = np.asarray(_np(xt), np.float64)
    _lib.check(_lib.lib().hipmc_variance_update(
        _lib.dtype_code(tdt), 3, _lib.ptr(gs_t), _lib.ptr(perm_t),
        _lib.ptr(xt), stride, dim, _lib.ptr(n), _lib.ptr(mean), _lib.ptr(ssr),
        0))
    for g in range(3):
      seen[g].append(xr[perm[gstart[g]:gstart[g + 1]], :dim])
  for g in range(3):
    pooled = np.concatenate(seen[g])
    assert _np(n)[g] == len(pooled)
    np.testing.assert_allclose(_np(mean)[g, :dim], pooled.mean(0), rtol=1e-12,
                               atol=1e-13)
    np.testing.assert_allclose(_np(ssr)[g, :dim] / len(pooled),
                               pooled.var(0), rtol=1e-11)


# ------------------------------------------------- preconditioned kernels
def _case(kind, ndt, seed, c):
  r = np.random.default_rng(seed)
  cast = lambda a: np.asarray(a, ndt).astype(np.float64)
  if kind == 'diag':
    loc, sc = cast(r.normal(size=40)), cast(np.exp(r.uniform(-2, 2, 40)))
    return (otg.DiagNormal(loc, sc),
            lambda dt: ptg.DiagNormal(loc, sc, dtype=dt),
            cast(loc + sc * r.normal(size=(c, 40))), None, sc ** 2)
  if kind == 'hier':
    j = 70
    y = cast(4 + 3 * r.normal(size=j) + r.uniform(5, 20, j) * r.normal(size=j))
    s = cast(r.uniform(5, 20, j))
    q = cast(np.concatenate([r.normal(size=(c, 2)), r.normal(size=(c, j))], 1))
    codes = np.zeros(j + 2, np.int32)
    codes[1] = 1
    prec = np.concatenate([[9.0, 0.5], np.full(j, 0.8)])
    return (otg.HierarchicalNormal(y, s),
            lambda dt: ptg.HierarchicalNormal(y, s, dtype=dt), q, codes, prec)
  if kind == 'linear':
    # noise scale 0.05: posterior sd of beta about 2e-3, of log sigma 0.03;
    # chains start near the posterior
    x = cast(r.normal(size=(600, 20)))
    beta = r.normal(size=20)
    y = cast(x @ beta + 0.05 * r.normal(size=600))
    q = cast(np.concatenate([beta + 0.002 * r.normal(size=(c, 20)),
                             np.log(0.05) + 0.03 * r.normal(size=(c, 1))], 1))
    codes = np.zeros(21, np.int32)
    codes[20] = 2
    prec = np.concatenate([np.full(20, 4e-6), [3e-4]])
    return (otg.LinearRegression(x, y, 1.0, 1.0),
            lambda dt: ptg.LinearRegression(x, y, 1.0, 1.0, dtype=dt), q,
            codes, prec)
  if kind == 'logistic_wide':
    x = cast(r.normal(size=(2700, 100)) * 0.1)
    y = (r.uniform(size=2700) < 0.5).astype(np.float64)
    prec = np.exp(r.uniform(-1, 1, 100)) * 0.05
    return (otg.LogisticRegression(x, y, 1.0),
            lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt),
            cast(r.normal(size=(c, 100)) * 0.03), None, prec)
  raise ValueError(kind)


def _bijs(spec, codes):
  if codes is None:
    return None
  out, off = [], 0
  for size in spec.part_sizes:
    out.append([pbij.Identity(), pbij.Softplus(), pbij.Exp()][int(codes[off])])
    off += size
  return out


def _precond_run(kind, tdt, ndt, seed, c, nuts=True, eps=0.2, steps=6):
  o, mk, q, codes, prec = _case(kind, ndt, seed, c)
  spec = mk(tdt)
  ot = o if codes is None else otg.Transformed(o, codes)
  prec = np.asarray(np.asarray(prec, ndt), np.float64)
  table = spec.table(_bijs(spec, codes))
  dev = 'cuda'
  state = torch.as_tensor(q, dtype=tdt, device=dev).contiguous()
  tlp = torch.empty(c, dtype=tdt, device=dev)
  grad = torch.zeros_like(state)
  _lib.check(_lib.lib().hipmc_target_eval(
      _lib.dtype_code(tdt), spec.kind, _lib.ptr(table), None, c, spec.dim,
      spec.dim, _lib.ptr(state), _lib.ptr(tlp), _lib.ptr(grad), 0))
  t0 = np.asarray(_np(tlp), np.float64)
  g0 = np.asarray(_np(grad), np.float64)
  eps = float(np.asarray(eps, ndt))
  step = torch.as_tensor([eps], dtype=tdt, device=dev)
  sd = orng.sanitize_seed(400 + seed)
  run = _runner.Run(spec, table, state, tlp, grad, step, sd)
  prec_t = torch.as_tensor(prec[None], dtype=tdt, device=dev).contiguous()
  run.args.momentum_precision = prec_t.data_ptr()
  lar = run.want('out_log_accept_ratio', (1, c), tdt)
  acc = run.want('out_is_accepted', (1, c), torch.uint8)
  energy = run.want('out_energy', (1, c), tdt)
  if nuts:
    run.args.max_tree_depth = 10
    run.args.unrolled_leapfrog_steps = 1
    lf = run.want('out_leapfrogs', (1, c), torch.int32)
    div = run.want('out_has_divergence', (1, c), torch.uint8)
    run.call(_lib.lib().hipmc_nuts_run)
    oq, ores = om.nuts_one_step(ot, eps, q, t0, g0, sd,
                                orng.PhiloxDraws(ndt), precision=prec)
    np.testing.assert_array_equal(_np(lf)[0], ores.leapfrogs_taken)
    np.testing.assert_array_equal(_np(div)[0].astype(bool),
                                  ores.has_divergence)
    np.testing.assert_allclose(_np(energy)[0], ores.energy, rtol=1e-5 if
                               ndt == np.float32 else 1e-11)
  else:
    run.args.num_leapfrog_steps = steps
    run.call(_lib.lib().hipmc_hmc_run)
    oq, ores = om.hmc_one_step(ot, eps, steps, q, t0, g0, sd,
                               orng.PhiloxDraws(ndt), prec=prec)
  np.testing.assert_array_equal(_np(acc)[0].astype(bool), ores.is_accepted)
  return _np(state), oq, _np(lar)[0], ores.log_accept_ratio


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('kind,eps', [('diag', 0.6), ('hier', 0.25),
                                      ('linear', 0.3), ('logistic_wide', 0.3)])
def test_preconditioned_nuts_matches_oracle(kind, eps, tdt, ndt, tol):
  for seed in range(2):
    got, want, lar, olar = _precond_run(kind, tdt, ndt, seed, 64, True, eps)
    np.testing.assert_allclose(got, want, rtol=tol, atol=tol)
    fin = np.isfinite(olar)
    np.testing.assert_allclose(lar[fin], olar[fin], rtol=tol * 10,
                               atol=tol * 10)


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('kind,eps', [('diag', 0.5), ('hier', 0.2),
                                      ('linear', 0.2)])
def test_preconditioned_hmc_matches_oracle(kind, eps, tdt, ndt, tol):
  for seed in range(2):
    got, want, lar, olar = _precond_run(kind, tdt, ndt, seed, 64, False, eps)
    err = np.max(np.abs(got - want) / (1 + np.abs(want)))
    print(f'preconditioned hmc {kind} {np.dtype(ndt).name}: max rel err {err:.2e}')
    np.testing.assert_allclose(got, want, rtol=tol, atol=tol)
    fin = np.isfinite(olar)
    np.testing.assert_allclose(lar[fin], olar[fin], rtol=tol * 10,
                               atol=tol * 10)


def test_preconditioned_nuts_equals_rescaled_identity_nuts_on_gpu():
  # the change of variables u = q / sqrt(prec) maps the preconditioned
  # sampler on DiagNormal(loc, scale) to the identity-metric sampler on
  # DiagNormal(loc / s, scale / s)
  r = np.random.default_rng(3)
  d, c = 30, 64
  loc, sc = r.normal(size=d), np.exp(r.uniform(-2, 2, d))
  prec = np.exp(r.uniform(-1.5, 1.5, d))
  s = np.sqrt(prec)
  q = loc + sc * r.normal(size=(c, d))
  dev, dt = 'cuda', torch.float64
  out = []
  for spec, x, pr in ((ptg.DiagNormal(loc, sc, dtype=dt), q, prec),
                      (ptg.DiagNormal(loc / s, sc / s, dtype=dt), q / s, None)):
    table = spec.table()
    state = torch.as_tensor(x, dtype=dt, device=dev).contiguous()
    tlp = torch.empty(c, dtype=dt, device=dev)
    grad = torch.zeros_like(state)
    _lib.check(_lib.lib().hipmc_target_eval(
        _lib.dtype_code(dt), spec.kind, _lib.ptr(table), None, c, d, d,
        _lib.ptr(state), _lib.ptr(tlp), _lib.ptr(grad), 0))
    run = _runner.Run(spec, table, state, tlp, grad,
                      torch.tensor([0.4], dtype=dt, device=dev),
                      orng.sanitize_seed(9))
    if pr is not None:
      pt = torch.as_tensor(pr[None], dtype=dt, device=dev).contiguous()
      run.args.momentum_precision = pt.data_ptr()
    run.args.max_tree_depth = 10
    run.args.unrolled_leapfrog_steps = 1
    lf = run.want('out_leapfrogs', (1, c), torch.int32)
    run.call(_lib.lib().hipmc_nuts_run)
    out.append((_np(state), _np(lf)))
  np.testing.assert_array_equal(out[0][1], out[1][1])
  np.testing.assert_allclose(out[0][0], out[1][0] * s, rtol=1e-10, atol=1e-12)


# ------------------------------------------------- windowed runs
def _diag_batch(dt, sizes=(5, 9), seed=0):
  r = np.random.default_rng(seed)
  specs, otgts = [], []
  for d in sizes:
    loc, sc = r.normal(size=d), np.exp(r.uniform(-2, 2, d))
    specs.append(ptg.DiagNormal(loc, sc, dtype=dt))
    otgts.append(otg.DiagNormal(loc, sc))
  return specs, otgts


def test_windowed_batch_matches_oracle_f64():
  dt = torch.float64
  specs, otgts = _diag_batch(dt)
  c, a, nr = 8, 40, 6
  batch = batched.PosteriorBatch(specs, c, slot_dim=9)
  x0 = torch.as_tensor(np.random.default_rng(5).uniform(-2, 2, (2 * c, 9)),
                       dtype=dt, device='cuda')
  for i, s in enumerate(specs):
    x0[i * c:(i + 1) * c, s.dim:] = 0
  run = batched._Run(batch, torch.cuda.current_stream(), nr,
                     num_burnin_steps=a, num_adaptation_steps=a,
                     step_size=0.3, target_accept_prob=0.85, seed=21,
                     adaptation='windowed', init_unconstrained=x0,
                     history=True)
  while not run.done():
    run.advance()
  res = run.finish()
  torch.cuda.synchronize()
  for i, (s, o) in enumerate(zip(specs, otgts)):
    xi = _np(x0)[i * c:(i + 1) * c, :s.dim]
    dr, trace, ws, _ = om.sample_windowed(o, xi, 0.3, a, nr, seed=21,
                                          problem=i, max_depth=10)
    lf = np.stack([t.leapfrogs_taken for t in trace])
    np.testing.assert_array_equal(_np(run.lf)[:, i * c:(i + 1) * c], lf)
    np.testing.assert_allclose(_np(res.draws[i]), dr, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(_np(run.prec)[i, :s.dim], ws.precision[0],
                               rtol=1e-8)
    np.testing.assert_allclose(float(_np(run.step)[i]),
                               float(np.asarray(ws.da.step_size)[0]),
                               rtol=1e-8)


def test_windowed_adaptive_nuts_recovers_variances():
  # `windowed_sampling_test.py:404-434`: the adapted mass matrix recovers the
  # posterior variances of a diagonal normal
  r = np.random.default_rng(11)
  loc = r.normal(size=6)
  scale = np.exp(r.uniform(-2.5, 2.5, size=6))
  model = pd.JointDistributionNamed(
      {'x': pd.Independent(pd.Normal(loc, scale), 1)})
  draws, trace = emcmc.windowed_adaptive_nuts(
      300, model, n_chains=16, num_adaptation_steps=600, seed=4)
  x = _np(draws['x'])
  assert x.shape == (300, 16, 6)
  var_used = _np(trace['variance_scaling']['x'])[-1, 0]
  np.testing.assert_allclose(1.0 / var_used, scale ** 2, rtol=0.15)
  z = (x.reshape(-1, 6).mean(0) - loc) / (scale / np.sqrt(x.shape[0]))
  assert np.all(np.abs(z) < 5)
  assert _np(trace['step_size']).shape == (300,)
  assert not _np(trace['tune']).any()
  assert _np(trace['n_steps']).shape == (300, 16)


def test_windowed_adaptive_hmc_sequential_structure():
  r = np.random.default_rng(2)
  xd = r.normal(size=(200, 3))
  yv = xd @ np.array([1.0, -0.5, 0.2]) + 0.3 * r.normal(size=200)
  model = pd.JointDistributionSequential([
      pd.Independent(pd.Normal(np.zeros(3), 1.0), 1),
      pd.HalfNormal(1.0),
      lambda sigma, beta: pd.Independent(
          pd.Normal(xd @ beta, sigma), 1)])
  names = [n for n, _ in model._trace()]
  states, trace = emcmc.windowed_adaptive_hmc(
      100, model, num_leapfrog_steps=8, n_chains=8, num_adaptation_steps=200,
      discard_tuning=False, seed=3, **{names[-1]: yv.astype(np.float32)})
  assert isinstance(states, list) and len(states) == 2
  assert tuple(states[0].shape) == (300, 8, 3)
  assert tuple(states[1].shape) == (300, 8)
  tune = _np(trace['tune'])
  assert tune[:200].all() and not tune[200:].any()
  post = _np(states[0])[200:].reshape(-1, 3).mean(0)
  np.testing.assert_allclose(post, [1.0, -0.5, 0.2], atol=0.15)
