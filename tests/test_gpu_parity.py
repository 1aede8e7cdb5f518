"""GPU parity: libhipmc (through the C ABI and the TFP-named Python layer)
against the CPU oracle on the same seeded inputs.

Tolerances: float32 kernels vs the float64 oracle on float32-rounded inputs
<= 1e-5 relative on states/log-probs after one transition (2e-4 for whole
multi-step trajectories), float64 kernels <= 1e-11; accept/reject, tree
depths, leapfrog counts and divergence flags identical.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import targets as otg
from probability_amd import _lib
from probability_amd import bijectors as pbij
from probability_amd import mcmc as pmcmc
from probability_amd import targets as ptg

pytestmark = pytest.mark.gpu

DTYPES = [(torch.float32, np.float32, 1e-5), (torch.float64, np.float64, 1e-11)]


def _np(t):
  return t.detach().cpu().numpy()


def _err(label, got, want):
  """Prints the largest absolute error of got vs want and that error over
  the largest |want| (normwise relative; run with -s to collect them)."""
  got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
  fin = np.isfinite(want)
  if not fin.any():
    return
  d = np.abs(got[fin] - want[fin]).max()
  rel = d / max(np.abs(want[fin]).max(), 1e-300)
  print(f'ERR {label} abs={d:.3e} rel={rel:.3e}', flush=True)


def _check(label, got, want, tol, scale=None):
  """Normwise bound: max |got - want| <= tol * max(scale, 1), scale
  defaulting to the largest finite |want|; prints the measured error."""
  got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
  fin = np.isfinite(want)
  np.testing.assert_array_equal(np.isfinite(got), fin)
  _err(label, got, want)
  if not fin.any():
    return
  if scale is None:
    scale = np.abs(want[fin]).max()
  d = np.abs(got[fin] - want[fin]).max()
  assert d <= tol * max(scale, 1.0), (label, d, tol * max(scale, 1.0))


def _case(kind, dtype_np, seed=0, c=64):
  """(oracle target, product spec ctor, initial state [C, D])."""
  r = np.random.default_rng(seed)
  cast = lambda a: np.asarray(a, dtype_np).astype(np.float64)
  if kind == 'logistic':
    x = cast(r.normal(size=(200, 7)))
    y = (r.uniform(size=200) < 0.5).astype(np.float64)
    o = otg.LogisticRegression(x, y, 1.0)
    mk = lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt)
    q = cast(r.normal(size=(c, 7)) * 0.3)
    return o, mk, q, None
  if kind == 'logistic_split':
    # 2051 x 9: big enough that four waves share each chain's gradient
    x = cast(r.normal(size=(2051, 9)) * 0.3)
    y = (r.uniform(size=2051) < 0.5).astype(np.float64)
    o = otg.LogisticRegression(x, y, 1.0)
    mk = lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt)
    q = cast(r.normal(size=(c, 9)) * 0.1)
    return o, mk, q, None
  if kind == 'logistic_wide':
    # 2700 x 100: eight waves share each chain's gradient
    x = cast(r.normal(size=(2700, 100)) * 0.1)
    y = (r.uniform(size=2700) < 0.5).astype(np.float64)
    o = otg.LogisticRegression(x, y, 1.0)
    mk = lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt)
    q = cast(r.normal(size=(c, 100)) * 0.03)
    return o, mk, q, None
  if kind == 'logistic_tall':
    # 9000 x 8: one coefficient block, X too large to keep X^T (one pass)
    x = cast(r.normal(size=(9000, 8)) * 0.1)
    y = (r.uniform(size=9000) < 0.5).astype(np.float64)
    o = otg.LogisticRegression(x, y, 1.0)
    mk = lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt)
    q = cast(r.normal(size=(c, 8)) * 0.05)
    return o, mk, q, None
  if kind == 'logistic_d150':
    # 150 coefficients: three blocks per row in float32 (batches of 2 rows
    # per phase), five in float64
    x = cast(r.normal(size=(400, 150)) * 0.1)
    y = (r.uniform(size=400) < 0.5).astype(np.float64)
    o = otg.LogisticRegression(x, y, 1.0)
    mk = lambda dt: ptg.LogisticRegression(x, y, 1.0, dtype=dt)
    q = cast(r.normal(size=(c, 150)) * 0.05)
    return o, mk, q, None
  if kind == 'linear_wide':
    # 70 coefficients: one pass over X (no X^T)
    x = cast(r.normal(size=(900, 70)) * 0.3)
    y = cast(x @ r.normal(size=70) * 0.2 + 0.5 * r.normal(size=900))
    o = otg.LinearRegression(x, y, 1.0, 1.0)
    mk = lambda dt: ptg.LinearRegression(x, y, 1.0, 1.0, dtype=dt)
    q = cast(np.concatenate([r.normal(size=(c, 70)) * 0.05,
                             np.full((c, 1), -0.5)], 1))
    codes = np.zeros(71, np.int32)
    codes[70] = 2  # Exp on sigma
    return o, mk, q, codes
  if kind == 'linear_split':
    x = cast(r.normal(size=(1500, 11)))
    y = cast(x @ r.normal(size=11) * 0.3 + 0.5 * r.normal(size=1500))
    o = otg.LinearRegression(x, y, 1.0, 1.0)
    mk = lambda dt: ptg.LinearRegression(x, y, 1.0, 1.0, dtype=dt)
    q = cast(np.concatenate([r.normal(size=(c, 11)) * 0.1,
                             np.full((c, 1), -0.5)], 1))
    codes = np.zeros(12, np.int32)
    codes[11] = 2  # Exp on sigma
    return o, mk, q, codes
  if kind == 'linear':
    x = cast(r.normal(size=(150, 6)))
    y = cast(x @ r.normal(size=6) + 0.5 * r.normal(size=150))
    o = otg.LinearRegression(x, y, 1.0, 1.0)
    mk = lambda dt: ptg.LinearRegression(x, y, 1.0, 1.0, dtype=dt)
    q = cast(np.concatenate([r.normal(size=(c, 6)) * 0.3,
                             np.full((c, 1), 0.2)], 1))
    codes = np.zeros(7, np.int32)
    codes[6] = 2  # Exp on sigma
    return o, mk, q, codes
  if kind in ('hier', 'hier256', 'hier600', 'hier1024'):
    # 256: the headline's groups; 600 and 1024: states of 12 and 17 slots
    j = {'hier': 70, 'hier256': 256, 'hier600': 600, 'hier1024': 1024}[kind]
    y = cast(4 + 3 * r.normal(size=j) + r.uniform(5, 20, j) * r.normal(size=j))
    s = cast(r.uniform(5, 20, j))
    o = otg.HierarchicalNormal(y, s)
    mk = lambda dt: ptg.HierarchicalNormal(y, s, dtype=dt)
    q = cast(np.concatenate([r.normal(size=(c, 1)), r.normal(size=(c, 1)),
                             r.normal(size=(c, j))], 1))
    codes = np.zeros(j + 2, np.int32)
    codes[1] = 1  # Softplus on tau
    return o, mk, q, codes
  if kind == 'diag':
    loc = cast(r.normal(size=130))
    sc = cast(r.uniform(0.5, 2, 130))
    o = otg.DiagNormal(loc, sc)
    mk = lambda dt: ptg.DiagNormal(loc, sc, dtype=dt)
    q = cast(loc + sc * r.normal(size=(c, 130)))
    return o, mk, q, None
  if kind == 'funnel':
    o = otg.NealsFunnel(10)
    mk = lambda dt: ptg.NealsFunnel(10, dtype=dt)
    q = cast(r.normal(size=(c, 10)))
    return o, mk, q, None
  if kind == 'loggamma':
    o = otg.LogGamma(5, 5.0, 10.0)
    mk = lambda dt: ptg.LogGamma(5, 5.0, 10.0, dtype=dt)
    q = cast(-0.8 + 0.4 * r.normal(size=(c, 5)))
    return o, mk, q, None
  raise ValueError(kind)


def _bijs(spec, codes):
  if codes is None:
    return None
  out = []
  off = 0
  for size in spec.part_sizes:
    c = int(codes[off])
    out.append([pbij.Identity(), pbij.Softplus(), pbij.Exp()][c])
    off += size
  return out


def _otarget(o, codes):
  return o if codes is None else otg.Transformed(o, codes)


# ---------------------------------------------------------------- draws
@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_draws_match_oracle(tdt, ndt, tol):
  lib = _lib.lib()
  seed = orng.sanitize_seed(1234)
  sp = np.ascontiguousarray(seed)
  n, d = 300, 37
  for problem in (0, 5):
    out = torch.empty((n, d), dtype=tdt, device='cuda')
    _lib.check(lib.hipmc_draws(_lib.dtype_code(tdt), 0,
                               sp.ctypes.data_as(ctypes.c_void_p), problem, n,
                               d, 0, 0, _lib.ptr(out), None, 0))
    ref = orng.momentum(seed, np.arange(n), d, ndt, problem)
    np.testing.assert_allclose(_np(out), ref, rtol=4 * np.finfo(ndt).eps,
                               atol=4 * np.finfo(ndt).eps)
    out = torch.empty(n, dtype=tdt, device='cuda')
    _lib.check(lib.hipmc_draws(_lib.dtype_code(tdt), 1,
                               sp.ctypes.data_as(ctypes.c_void_p), problem, n,
                               0, 0, 0, _lib.ptr(out), None, 0))
    np.testing.assert_array_equal(_np(out),
                                  orng.mh_uniform(seed, np.arange(n), ndt,
                                                  problem))
    aux = torch.empty(n, dtype=torch.uint8, device='cuda')
    _lib.check(lib.hipmc_draws(_lib.dtype_code(tdt), 2,
                               sp.ctypes.data_as(ctypes.c_void_p), problem, n,
                               0, 3, 0, _lib.ptr(out), _lib.ptr(aux), 0))
    dref, uref = orng.tree_draws(seed, np.arange(n), 3, ndt, problem)
    np.testing.assert_array_equal(_np(aux).astype(bool), dref)
    np.testing.assert_array_equal(_np(out), uref)
    _lib.check(lib.hipmc_draws(_lib.dtype_code(tdt), 3,
                               sp.ctypes.data_as(ctypes.c_void_p), problem, n,
                               0, 4, 9, _lib.ptr(out), None, 0))
    np.testing.assert_array_equal(_np(out), orng.leaf_uniform(
        seed, np.arange(n), 4, 9, ndt, problem))


# ---------------------------------------------------------------- targets
@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('kind', ['logistic', 'linear', 'hier', 'diag',
                                  'funnel', 'loggamma', 'logistic_split',
                                  'linear_split',
                                  'logistic_wide', 'logistic_d150',
                                  'linear_wide', 'logistic_tall', 'hier600',
                                  'hier1024'])
def test_target_eval_matches_oracle(kind, tdt, ndt, tol):
  o, mk, q, codes = _case(kind, ndt)
  spec = mk(tdt)
  ot = _otarget(o, codes)
  x = torch.as_tensor(q, dtype=tdt, device='cuda')
  tlp, grad = spec.log_prob_and_grad_flat(x, _bijs(spec, codes))
  rt, rg = ot.log_prob_and_grad(q)
  lab = f'eval {kind} {np.dtype(ndt).name}'
  _check(lab + ' tlp', _np(tlp), rt, tol)
  _check(lab + ' grad', _np(grad), rg, tol)


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_target_callable_like_tfp(tdt, ndt, tol):
  o, mk, q, _ = _case('hier', ndt, c=8)
  spec = mk(tdt)
  parts = [torch.as_tensor(q[:, 0], dtype=tdt), torch.as_tensor(q[:, 1], dtype=tdt),
           torch.as_tensor(q[:, 2:], dtype=tdt)]
  tlp = spec(*parts)
  assert tlp.shape == (8,)
  np.testing.assert_allclose(_np(tlp), o.log_prob(q), rtol=tol * 5)


# ---------------------------------------------------------------- HMC
@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('kind', ['logistic', 'hier', 'logistic_split',
                                  'logistic_wide', 'loggamma', 'hier1024'])
def test_hmc_one_step_matches_oracle(kind, tdt, ndt, tol):
  o, mk, q, codes = _case(kind, ndt, seed=3)
  spec = mk(tdt)
  ot = _otarget(o, codes)
  eps, nl = {'logistic': (0.05, 10), 'hier': (0.2, 6),
             'logistic_split': (0.05, 8), 'logistic_wide': (0.02, 8),
             'loggamma': (0.3, 4), 'hier1024': (0.1, 5)}[kind]
  kern = pmcmc.HamiltonianMonteCarlo(
      spec if codes is None else ptg.TransformedTarget(spec, _bijs(spec, codes)),
      step_size=eps, num_leapfrog_steps=nl)
  x = torch.as_tensor(q, dtype=tdt, device='cuda')
  parts = spec.unpack_state(x, (q.shape[0],))
  pkr = kern.bootstrap_results(parts)
  seed = orng.sanitize_seed(77)
  nxt, res = kern.one_step(parts, pkr, seed=seed)
  t0, g0 = ot.log_prob_and_grad(q)
  oq, ores = omcmc.hmc_one_step(ot, np.asarray(eps, ndt), nl, q, t0, g0, seed,
                                orng.PhiloxDraws(ndt))
  np.testing.assert_array_equal(_np(res.is_accepted), ores.is_accepted)
  lab = f'hmc {kind} {np.dtype(ndt).name}'
  _check(lab + ' state', _np(spec.pack_state(nxt)), oq, tol)
  _check(lab + ' proposed', _np(spec.pack_state(res.proposed_state)),
         ores.proposed_state, tol)
  # the log accept ratio is a difference of energies: its error is bounded
  # relative to the target log-prob's magnitude
  _check(lab + ' lar', _np(res.log_accept_ratio), ores.log_accept_ratio, tol,
         scale=np.abs(t0).max())


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_mh_over_uncalibrated_equals_fused_hmc(tdt, ndt, tol):
  o, mk, q, _ = _case('logistic', ndt, seed=4, c=32)
  spec = mk(tdt)
  x = torch.as_tensor(q, dtype=tdt, device='cuda')
  parts = spec.unpack_state(x, (32,))
  fused = pmcmc.HamiltonianMonteCarlo(spec, 0.05, 8)
  comp = pmcmc.MetropolisHastings(
      pmcmc.UncalibratedHamiltonianMonteCarlo(spec, 0.05, 8))
  pa, pb = fused.bootstrap_results(parts), comp.bootstrap_results(parts)
  sa, sb = parts, parts
  for s in range(3):
    sa, pa = fused.one_step(sa, pa, seed=orng.fold_in(5, s))
    sb, pb = comp.one_step(sb, pb, seed=orng.fold_in(5, s))
    np.testing.assert_array_equal(_np(pa.is_accepted), _np(pb.is_accepted))
    np.testing.assert_array_equal(_np(sa[0]), _np(sb[0]))


def test_hmc_nan_rejection_on_gpu():
  spec = ptg.NaNExponential(10, dtype=torch.float32)
  x = torch.linspace(0.01, 5, 10, device='cuda')[None]
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=2., num_leapfrog_steps=5)
  nxt, res = kern.one_step(x, kern.bootstrap_results(x), seed=17)
  np.testing.assert_array_equal(_np(nxt), _np(x))
  assert float(torch.exp(torch.clamp(res.log_accept_ratio, max=0.0))[0]) == 0.0


# ---------------------------------------------------------------- NUTS
def _nuts_one(kind, tdt, ndt, seed=0, c=64, eps=None, max_depth=10):
  o, mk, q, codes = _case(kind, ndt, seed=seed, c=c)
  spec = mk(tdt)
  ot = _otarget(o, codes)
  eps = eps or {'logistic': 0.08, 'linear': 0.05, 'hier': 0.3,
                'hier256': 0.3, 'hier600': 0.2, 'hier1024': 0.15, 'diag': 0.5,
                'funnel': 0.9, 'logistic_split': 0.04,
                'linear_split': 0.02, 'logistic_wide': 0.02,
                'logistic_d150': 0.05, 'linear_wide': 0.01,
                'logistic_tall': 0.03, 'loggamma': 0.3}[kind]
  tspec = spec if codes is None else ptg.TransformedTarget(spec,
                                                           _bijs(spec, codes))
  kern = pmcmc.NoUTurnSampler(tspec, step_size=np.asarray(eps, ndt),
                              max_tree_depth=max_depth)
  x = torch.as_tensor(q, dtype=tdt, device='cuda')
  parts = spec.unpack_state(x, (c,))
  pkr = kern.bootstrap_results(parts)
  seed_arr = orng.sanitize_seed(1000 + seed)
  nxt, res = kern.one_step(parts, pkr, seed=seed_arr)
  t0, g0 = ot.log_prob_and_grad(q)
  oq, ores = omcmc.nuts_one_step(ot, float(np.asarray(eps, ndt)), q, t0, g0,
                                 seed_arr, orng.PhiloxDraws(ndt),
                                 max_depth=max_depth)
  return spec, nxt, res, oq, ores


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
@pytest.mark.parametrize('kind', ['hier', 'hier256', 'hier600', 'hier1024',
                                  'logistic', 'linear',
                                  'diag', 'funnel', 'loggamma',
                                  'logistic_split',
                                  'linear_split', 'logistic_wide',
                                  'logistic_d150', 'linear_wide',
                                  'logistic_tall'])
def test_nuts_one_step_matches_oracle(kind, tdt, ndt, tol):
  # float64 states of 17 slots hold the U-turn memory of 8 levels in LDS
  depth = 8 if kind == 'hier1024' and ndt == np.float64 else 10
  for seed in range(3):
    spec, nxt, res, oq, ores = _nuts_one(kind, tdt, ndt, seed=seed,
                                         max_depth=depth)
    np.testing.assert_array_equal(_np(res.leapfrogs_taken),
                                  ores.leapfrogs_taken)
    np.testing.assert_array_equal(_np(res.is_accepted), ores.is_accepted)
    np.testing.assert_array_equal(_np(res.reach_max_depth),
                                  ores.reach_max_depth)
    np.testing.assert_array_equal(_np(res.has_divergence),
                                  ores.has_divergence)
    lab = f'nuts {kind} {np.dtype(ndt).name} seed{seed}'
    _check(lab + ' state', _np(spec.pack_state(nxt)), oq, tol)
    _check(lab + ' tlp', _np(res.target_log_prob), ores.target_log_prob, tol)
    _check(lab + ' energy', _np(res.energy), ores.energy, tol)
    _check(lab + ' lar', _np(res.log_accept_ratio), ores.log_accept_ratio,
           tol, scale=np.abs(ores.energy[np.isfinite(ores.energy)]).max())


def test_nuts_divergence_on_gpu_funnel():
  # `mcmc/nuts_test.py:445-474`
  spec = ptg.NealsFunnel(10, dtype=torch.float32)
  r = np.random.default_rng(0)
  v = 3 * r.normal(size=(5, 1))
  x = np.concatenate([v, np.exp(v / 2) * r.normal(size=(5, 9))], 1)
  parts = spec.unpack_state(torch.as_tensor(x, dtype=torch.float32,
                                            device='cuda'), (5,))
  kern = pmcmc.NoUTurnSampler(lambda_free(spec), step_size=1.0)
  _, div = pmcmc.sample_chain(100, parts, kernel=kern, seed=3,
                              trace_fn=pmcmc.TraceFields('has_divergence'))
  assert int(div.sum()) > 100


def lambda_free(spec):
  return spec


# ---------------------------------------------------------------- DA
@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_da_kernel_matches_oracle(tdt, ndt, tol):
  lar = np.log([[0.70, 0.76, 0.73], [0.76, 0.76, 0.73]]).astype(ndt)
  for step_shape in [(), (3, 1), (2, 3, 1), (2, 1, 1), (1, 1)]:
    step = np.ones(step_shape)

    class FakeInner:
      pass
    # drive the product DA with a fake inner kernel
    from probability_amd.mcmc import kernel as pk
    import collections
    R = collections.namedtuple('R', 'log_accept_ratio, step_size')

    class Fake(pk.TransitionKernel):
      def __init__(self, step_size, store_parameters_in_results=False):
        self._parameters = dict(step_size=step_size,
                                store_parameters_in_results=True)

      def one_step(self, s, pkr, seed=None):
        return s, pkr

      def bootstrap_results(self, s):
        return R(torch.as_tensor(lar, device='cuda'),
                 torch.as_tensor(self._parameters['step_size'], dtype=tdt,
                                 device='cuda'))

      @property
      def is_calibrated(self):
        return True

    k = pmcmc.DualAveragingStepSizeAdaptation(Fake(step), num_adaptation_steps=1)
    state = torch.zeros((2, 3, 4), dtype=tdt, device='cuda')
    pkr = k.bootstrap_results(state)
    for _ in range(2):
      _, pkr = k.one_step(state, pkr)
    ost = omcmc.da_bootstrap(step)
    for _ in range(2):
      ost = omcmc.da_update(ost, lar.astype(np.float64), state_rank=3,
                            num_adaptation_steps=1)
    np.testing.assert_allclose(_np(pkr.inner_results.step_size),
                               np.broadcast_to(ost.step_size, step_shape),
                               rtol=max(tol, 1e-6))


# ---------------------------------------------------------------- drivers
def _hier_kernel(tdt, ndt, j=40, c=64, adapt=10):
  r = np.random.default_rng(11)
  y = (4 + 3 * r.normal(size=j) + r.uniform(5, 20, j) * r.normal(size=j))
  s = r.uniform(5, 20, j)
  spec = ptg.HierarchicalNormal(np.asarray(y, ndt), np.asarray(s, ndt),
                                dtype=tdt)
  nuts = pmcmc.NoUTurnSampler(spec, step_size=np.asarray(0.1, ndt))
  ttk = pmcmc.TransformedTransitionKernel(
      nuts, [pbij.Identity(), pbij.Softplus(), pbij.Identity()])
  da = pmcmc.DualAveragingStepSizeAdaptation(ttk, num_adaptation_steps=adapt,
                                             target_accept_prob=0.8)
  init = [torch.zeros(c, dtype=tdt), torch.full((c,), float(np.log(2.0)),
                                                dtype=tdt),
          torch.zeros((c, j), dtype=tdt)]
  return spec, da, init, y, s


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_fused_sample_chain_equals_step_loop(tdt, ndt, tol):
  spec, da, init, _, _ = _hier_kernel(tdt, ndt)
  tf = pmcmc.TraceFields('log_accept_ratio', 'leapfrogs_taken', 'step_size')
  a = pmcmc.sample_chain(12, init, kernel=da, num_burnin_steps=8,
                         num_steps_between_results=1, trace_fn=tf, seed=5,
                         return_final_kernel_results=True)
  b = pmcmc.sample_chain(12, init, kernel=da, num_burnin_steps=8,
                         num_steps_between_results=1, trace_fn=tf, seed=5,
                         return_final_kernel_results=True,
                         experimental_fused=False)
  for pa, pb in zip(a.all_states, b.all_states):
    np.testing.assert_array_equal(_np(pa), _np(pb))
  np.testing.assert_array_equal(_np(a.trace[0]), _np(b.trace[0]))
  np.testing.assert_array_equal(_np(a.trace[1]), _np(b.trace[1]))
  np.testing.assert_array_equal(_np(a.trace[2]), _np(b.trace[2]))
  assert tuple(a.trace[2].shape) == (12,)  # scalar step size: [num_results]
  np.testing.assert_array_equal(_np(a.final_kernel_results.new_step_size),
                                _np(b.final_kernel_results.new_step_size))


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_resume_in_the_fused_loop_equals_step_loop(tdt, ndt, tol):
  # `mcmc/sample.py:330`: sample_chain(previous_kernel_results=...) resumes
  # the chains and the adaptation; here the first call stops mid-adaptation
  # (6 of 10 steps) and the second finishes it and samples
  spec, da, init, _, _ = _hier_kernel(tdt, ndt)
  tf = pmcmc.TraceFields('log_accept_ratio', 'leapfrogs_taken', 'step_size')
  first = pmcmc.sample_chain(2, init, kernel=da, num_burnin_steps=4,
                             trace_fn=tf, seed=5,
                             return_final_kernel_results=True)
  pkr = first.final_kernel_results
  assert int(pkr.step) == 6
  last = [p[-1] for p in first.all_states]
  kw = dict(num_burnin_steps=3, num_steps_between_results=1, trace_fn=tf,
            seed=8, previous_kernel_results=pkr,
            return_final_kernel_results=True)
  step_before = _np(pkr.new_step_size).copy()
  a = pmcmc.sample_chain(5, last, kernel=da, **kw)
  a2 = pmcmc.sample_chain(5, last, kernel=da, **kw)
  b = pmcmc.sample_chain(5, last, kernel=da, experimental_fused=False, **kw)
  # the caller's kernel results are not changed by a fused resume
  np.testing.assert_array_equal(_np(pkr.new_step_size), step_before)
  for pa, pa2, pb in zip(a.all_states, a2.all_states, b.all_states):
    np.testing.assert_array_equal(_np(pa), _np(pb))
    np.testing.assert_array_equal(_np(pa2), _np(pb))
  for k in range(3):
    np.testing.assert_array_equal(_np(a.trace[k]), _np(b.trace[k]))
  fa, fb = a.final_kernel_results, b.final_kernel_results
  assert int(fa.step) == int(fb.step) == 6 + 3 + 1 + 4 * 2
  for f in ('new_step_size', 'error_sum', 'log_averaging_step'):
    np.testing.assert_array_equal(_np(getattr(fa, f)), _np(getattr(fb, f)))


def test_headline_shape_fused_equals_step_loop():
  """The headline's shape (J = 256 groups, 1024 chains, float32, NUTS + DA
  + TTK): the fused device loop equals the host step loop bit for bit."""
  spec, da, init, _, _ = _hier_kernel(torch.float32, np.float32, j=256,
                                      c=1024, adapt=3)
  tf = pmcmc.TraceFields('log_accept_ratio', 'leapfrogs_taken', 'step_size')
  kw = dict(num_burnin_steps=3, trace_fn=tf, seed=9,
            return_final_kernel_results=True)
  a = pmcmc.sample_chain(3, init, kernel=da, **kw)
  b = pmcmc.sample_chain(3, init, kernel=da, experimental_fused=False, **kw)
  for pa, pb in zip(a.all_states, b.all_states):
    assert tuple(pa.shape[1:2]) == (1024,)
    np.testing.assert_array_equal(_np(pa), _np(pb))
  for k in range(3):
    np.testing.assert_array_equal(_np(a.trace[k]), _np(b.trace[k]))


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_sample_chain_nuts_da_ttk_matches_oracle(tdt, ndt, tol):
  spec, da, init, y, s = _hier_kernel(tdt, ndt, j=12, c=32, adapt=6)
  tf = pmcmc.TraceFields('log_accept_ratio', 'leapfrogs_taken')
  res = pmcmc.sample_chain(4, init, kernel=da, num_burnin_steps=6, trace_fn=tf,
                           seed=21)
  base = otg.HierarchicalNormal(np.asarray(y, ndt).astype(np.float64),
                                np.asarray(s, ndt).astype(np.float64))
  codes = np.zeros(14, np.int32)
  codes[1] = 1
  ot = otg.Transformed(base, codes)
  x0 = np.zeros((32, 14))
  ox, otrace, _ = omcmc.sample_nuts_da(ot, x0, np.asarray(0.1, ndt), 6, 4, 6,
                                       seed=orng.sanitize_seed(21),
                                       target_accept_prob=0.8, dtype=ndt)
  # the oracle returns unconstrained states; compare constrained
  ostates = np.stack([ot.forward(v) for v in ox])
  got = np.concatenate([_np(res.all_states[0])[..., None],
                        _np(res.all_states[1])[..., None],
                        _np(res.all_states[2])], -1)
  lf = np.stack([r.leapfrogs_taken for r in otrace])
  np.testing.assert_array_equal(_np(res.trace[1]), lf)
  np.testing.assert_allclose(got, ostates, rtol=2e-4 if ndt == np.float32
                             else 1e-9, atol=2e-4 if ndt == np.float32 else 1e-9)


@pytest.mark.parametrize('tdt,ndt,tol', DTYPES)
def test_fused_trace_of_target_log_prob_equals_step_loop(tdt, ndt, tol):
  # the kernels write each iteration's target log-prob, so the fused loop
  # traces it in its multi-iteration launches too
  spec, da, init, _, _ = _hier_kernel(tdt, ndt)
  tf = pmcmc.TraceFields('target_log_prob', 'energy')
  kw = dict(num_burnin_steps=8, num_steps_between_results=2, trace_fn=tf,
            seed=6)
  a = pmcmc.sample_chain(10, init, kernel=da, **kw)
  b = pmcmc.sample_chain(10, init, kernel=da, experimental_fused=False, **kw)
  assert tuple(a.trace[0].shape) == (10,) + tuple(init[0].shape)
  np.testing.assert_array_equal(_np(a.trace[0]), _np(b.trace[0]))
  np.testing.assert_array_equal(_np(a.trace[1]), _np(b.trace[1]))


def test_zero_step_sample_chain_returns_bootstrap_results():
  spec, da, init, _, _ = _hier_kernel(torch.float32, np.float32)
  out = pmcmc.sample_chain(0, init, kernel=da, num_burnin_steps=0,
                           trace_fn=None, seed=3,
                           return_final_kernel_results=True)
  boot = da.bootstrap_results(init)
  fk = out.final_kernel_results
  inner, binner = fk.inner_results.inner_results, boot.inner_results.inner_results
  np.testing.assert_array_equal(_np(inner.target_log_prob),
                                _np(binner.target_log_prob))
  np.testing.assert_array_equal(_np(inner.log_accept_ratio),
                                _np(binner.log_accept_ratio))
  np.testing.assert_array_equal(_np(inner.energy), _np(binner.energy))


def test_wide_float64_state_at_depth_10_is_refused():
  # 1026 float64 elements: the U-turn memory of 10 levels (174 KiB) exceeds
  # the workgroup's LDS; the launch is refused on the host with a message
  o, mk, q, codes = _case('hier1024', np.float64, c=4)
  spec = mk(torch.float64)
  kern = pmcmc.NoUTurnSampler(
      ptg.TransformedTarget(spec, _bijs(spec, codes)), step_size=0.1,
      max_tree_depth=10)
  parts = spec.unpack_state(torch.as_tensor(q, device='cuda'), (4,))
  pkr = kern.bootstrap_results(parts)
  with pytest.raises(RuntimeError, match='LDS'):
    kern.one_step(parts, pkr, seed=[1, 2])


def test_state_above_the_built_width_is_refused():
  spec = ptg.DiagNormal(np.zeros(1100), np.ones(1100), dtype=torch.float32)
  kern = pmcmc.HamiltonianMonteCarlo(spec, step_size=0.1,
                                     num_leapfrog_steps=2)
  x = torch.zeros((2, 1100), device='cuda')
  with pytest.raises(RuntimeError, match='above 1088'):
    kern.one_step(x, kern.bootstrap_results(x), seed=[1, 2])
