"""Launch order of ragged batches (host logic, no GPU)."""

import numpy as np
import pytest

from probability_amd.mcmc import batched


@pytest.mark.parametrize('n', [1, 7, 8, 20, 391])
def test_xcd_order_is_a_balanced_permutation(n):
  r = np.random.default_rng(n)
  idx = list(range(100, 100 + n))
  cost = list(r.integers(1, 10**6, size=n))
  order = batched._xcd_order(idx, cost)
  assert sorted(order) == idx
  c = dict(zip(idx, cost))
  # deal in snake order: bin k holds ranks k, 15-k, 16+k, 31-k, ...
  ranks = {i: r_ for r_, i in enumerate(sorted(idx, key=lambda i: -c[i]))}
  bins, pos = [], 0
  for k in range(8):
    size = sum(1 for r_ in range(n) if
               (r_ % 16 if r_ % 16 < 8 else 15 - r_ % 16) == k)
    bins.append(order[pos:pos + size])
    pos += size
  sizes = [len(b) for b in bins]
  assert max(sizes) - min(sizes) <= 1
  for k, b in enumerate(bins):
    # each range starts with its most expensive problem and goes down
    assert [c[i] for i in b] == sorted((c[i] for i in b), reverse=True)
    if b:
      assert ranks[b[0]] == k
  if n >= 16:
    totals = [sum(c[i] for i in b) for b in bins]
    assert max(totals) <= 1.5 * min(totals)


def test_sample_ragged_rejects_no_problems():
  with pytest.raises(ValueError):
    batched.sample_ragged([], 4, num_results=1)


@pytest.mark.parametrize('n,k', [(1, 4), (3, 4), (10, 1), (192, 4), (50, 8)])
def test_split_by_cost_partitions_with_balanced_cost(n, k):
  r = np.random.default_rng(n + k)
  idx = list(range(7, 7 + n))
  cost = {i: int(c) for i, c in zip(idx, r.integers(1, 10**5, size=n))}
  full = [0] * (7 + n)
  for i, c in cost.items():
    full[i] = c
  parts = batched._split_by_cost(idx, full, k)
  # every problem once, no empty part, at most k parts
  assert sorted(i for p in parts for i in p) == idx
  assert len(parts) == min(k, n) and all(parts)
  if len(parts) > 1:
    # greedy by decreasing cost: no part exceeds another by more than the
    # largest single cost
    totals = [sum(cost[i] for i in p) for p in parts]
    assert max(totals) - min(totals) <= max(cost.values())
