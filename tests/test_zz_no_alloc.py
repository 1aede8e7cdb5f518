"""hipmc_glm_hmc_run_noalloc (workspaces from the library's static device
scratch) against hipmc_glm_hmc_run (caller workspaces), through the C ABI:
the same kernels on the same inputs, so every output is bit-identical.  Two
configurations in a row, the second smaller, so the second call runs on a
scratch array that holds the first call's contents."""

import ctypes

import numpy as np
import pytest
import torch

from probability_amd import _lib
from probability_amd import targets as ptg
from probability_amd.mcmc import glm_mfma

pytestmark = pytest.mark.gpu


def _spec(n, seed):
  r = np.random.default_rng(seed)
  x = (r.normal(size=(n, 256)) / 16).astype(np.float32)
  beta = r.normal(size=256)
  y = (r.uniform(size=n) < 1 / (1 + np.exp(-(x @ beta)))).astype(np.float32)
  return ptg.LogisticRegressionBF16(x, y)


def _outputs(spec, runner, q, noalloc, n_iters=3, leapfrogs=4):
  dev = 'cuda'
  c = q.shape[0]
  state = torch.as_tensor(q, device=dev).contiguous().clone()
  tlp = torch.empty(c, dtype=torch.float32, device=dev)
  grad = torch.empty_like(state)
  step = torch.full((1,), 0.03, dtype=torch.float32, device=dev)
  lar = torch.empty((n_iters, c), dtype=torch.float32, device=dev)
  acc = torch.empty((n_iters, c), dtype=torch.uint8, device=dev)
  fn = (_lib.lib().hipmc_glm_hmc_run_noalloc if noalloc
        else _lib.lib().hipmc_glm_hmc_run)

  def args():
    a = runner._args(state, tlp, grad, step)
    if noalloc:  # the entry point ignores caller workspaces
      a.ws_g = a.ws_ll = a.ws_q = a.ws_p = a.ws_b = a.ws_ke0 = None
    return a

  a = args()
  a.want_grad_only = 1
  _lib.check(fn(a, 0))
  boot = (tlp.clone(), grad.clone())
  a = args()
  a.num_leapfrog_steps, a.n_iters, a.fold_seed = leapfrogs, n_iters, 1
  a.seed[0], a.seed[1] = 123, 456
  a.out_log_accept_ratio, a.out_is_accepted = lar.data_ptr(), acc.data_ptr()
  _lib.check(fn(a, 0))
  torch.cuda.synchronize()
  return [t.cpu().numpy() for t in (*boot, state, tlp, grad, lar, acc)]


def test_noalloc_matches_caller_workspaces_twice():
  for n, c, seed in ((1024, 128, 0), (256, 64, 1)):
    spec = _spec(n, seed)
    runner = glm_mfma.GlmHmcRunner(spec, c)
    q = (np.random.default_rng(seed + 7).normal(size=(c, 256)) * 0.3).astype(
        np.float32)
    want = _outputs(spec, runner, q, noalloc=False)
    got = _outputs(spec, runner, q, noalloc=True)
    for g, w in zip(got, want):
      np.testing.assert_array_equal(g, w)


def test_noalloc_rejects_workspace_beyond_scratch():
  spec = _spec(256, 2)
  runner = glm_mfma.GlmHmcRunner(spec, 64)
  dev = 'cuda'
  state = torch.zeros((64, 256), device=dev)
  tlp = torch.empty(64, device=dev)
  a = runner._args(state, tlp, state.clone(), torch.ones(1, device=dev))
  a.n_chains = 1 << 20   # would need far more than the static scratch
  a.n_split = 1
  a.want_grad_only = 1
  rc = _lib.lib().hipmc_glm_hmc_run_noalloc(a, 0)
  assert rc == _lib.E_UNSUPPORTED
  assert b'static scratch' in _lib.lib().hipmc_last_error()
