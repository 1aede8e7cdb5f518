"""Counter-keyed draw schedule shared by the oracle and the HIP kernels.

TFP draws every random number with stateless TF ops and splits seeds along a
call tree (`internal/samplers.py:217-258` split_seed, `:308-325` normal,
`:356-367` uniform; consumption in `mcmc/hmc.py:684-695`,
`mcmc/metropolis_hastings.py:183,221-226`, `mcmc/nuts.py:322-323,512-558,
622-625,793-900`).  Those TF bit streams cannot be imported here, so the
build uses one Philox4x32-10 block per draw, keyed by the step seed and
indexed by (index, chain, problem, purpose|depth).  This makes every chain's
draws independent of the batch it runs in (needed for "same result alone and
in the mix") and lets a per-chain GPU state machine reproduce the batched
algorithm draw for draw.

Layout of one Philox call: key = seed (2 x uint32); counter =
(index, chain, problem, purpose | depth).

  momentum   purpose MOMENTUM, index = element // 4, word = element % 4
             Box-Muller in float64 on word pairs (w0,w1) and (w2,w3).
  MH accept  purpose MH_ACCEPT, index 0.
  NUTS tree  purpose TREE | depth, index 0: word 0 top bit = direction,
             words 1(,2) = uniform for the tree-level accept.
  NUTS leaf  purpose LEAF | depth, index = leapfrog number inside the
             subtree: words 0(,1) = uniform for the in-subtree accept.

float32 uniforms use 24 bits of one word, float64 uniforms 53 bits of two.

Test infrastructure only (see `oracle/__init__.py`).
"""

import zlib

import numpy as np

from oracle.philox import philox4x32_10

PURPOSE_SPLIT = 0x53504C54
PURPOSE_FOLD = 0x464F4C44
PURPOSE_SALT = 0x53414C54
PURPOSE_MOMENTUM = 0x4D4F4D00
PURPOSE_MH_ACCEPT = 0x4D484100
PURPOSE_TREE = 0x54520000
PURPOSE_LEAF = 0x4C460000

_TWO_M24 = 2.0 ** -24
_TWO_M32 = 2.0 ** -32
_TWO_M53 = 2.0 ** -53


def _u32(x):
  return np.asarray(x, dtype=np.uint64).astype(np.uint32)


def sanitize_seed(seed, salt=None):
  """int / [2] seed -> uint32[2]; optional string salt folded in.

  Follows the role of `internal/samplers.py:79-130` (sanitize_seed with salt).
  """
  if isinstance(seed, (int, np.integer)):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    out = np.array([s & 0xFFFFFFFF, s >> 32], dtype=np.uint32)
  else:
    out = _u32(np.asarray(seed).reshape(2))
  if salt is not None:
    h = zlib.crc32(salt.encode('utf-8')) & 0xFFFFFFFF
    out = philox4x32_10(_u32([h, 0, 0, PURPOSE_SALT]), out)[:2]
  return out


def split_seed(seed, n=2):
  """[n, 2] child seeds; restates the role of `internal/samplers.py:217-258`."""
  seed = sanitize_seed(seed)
  idx = np.arange(n, dtype=np.uint64)
  ctr = np.stack([idx, np.zeros_like(idx), np.zeros_like(idx),
                  np.full_like(idx, PURPOSE_SPLIT)], axis=-1)
  return philox4x32_10(_u32(ctr), seed)[:, :2]


def fold_in(seed, t):
  """Seed of step `t` of a sampling loop (one Philox block, no sequential chain)."""
  seed = sanitize_seed(seed)
  t = int(t)
  ctr = _u32([t & 0xFFFFFFFF, t >> 32, 0, PURPOSE_FOLD])
  return philox4x32_10(ctr, seed)[:2]


def fold_in_many(seed, ts):
  seed = sanitize_seed(seed)
  ts = np.asarray(ts, dtype=np.uint64)
  ctr = np.stack([ts & 0xFFFFFFFF, ts >> 32, np.zeros_like(ts),
                  np.full_like(ts, PURPOSE_FOLD)], axis=-1)
  return philox4x32_10(_u32(ctr), seed)[:, :2]


def _block(seed, index, chain, problem, purpose):
  index, chain, problem, purpose = np.broadcast_arrays(
      np.asarray(index, np.uint64), np.asarray(chain, np.uint64),
      np.asarray(problem, np.uint64), np.asarray(purpose, np.uint64))
  ctr = np.stack([index, chain, problem, purpose], axis=-1)
  return philox4x32_10(_u32(ctr), sanitize_seed(seed))


def uniform_from_words(words, dtype):
  """words [..., >=2] uint32 -> uniform in [0, 1) of `dtype`."""
  words = np.asarray(words)
  if np.dtype(dtype) == np.float32:
    return ((words[..., 0] >> np.uint32(8)).astype(np.float64)
            * _TWO_M24).astype(np.float32)
  hi = words[..., 0].astype(np.uint64)
  lo = words[..., 1].astype(np.uint64)
  u53 = (hi << np.uint64(21)) | (lo >> np.uint64(11))
  return u53.astype(np.float64) * _TWO_M53


def box_muller(w_a, w_b):
  """Two standard normals from two uint32 words, computed in float64."""
  u1 = (np.asarray(w_a, np.float64) + 1.0) * _TWO_M32
  u2 = np.asarray(w_b, np.float64) * _TWO_M32
  r = np.sqrt(-2.0 * np.log(u1))
  theta = (2.0 * np.pi) * u2
  return r * np.cos(theta), r * np.sin(theta)


def momentum(seed, chains, dim, dtype=np.float64, problem=0):
  """Standard-normal momentum [len(chains), dim] for the given chain ids."""
  chains = np.asarray(chains, dtype=np.uint64)
  nblk = (dim + 3) // 4
  blk = np.arange(nblk, dtype=np.uint64)
  w = _block(seed, blk[None, :], chains[:, None], problem, PURPOSE_MOMENTUM)
  n0, n1 = box_muller(w[..., 0], w[..., 1])
  n2, n3 = box_muller(w[..., 2], w[..., 3])
  out = np.stack([n0, n1, n2, n3], axis=-1).reshape(len(chains), nblk * 4)
  return out[:, :dim].astype(dtype)


PURPOSE_INIT = 0x494E4954


def init_uniform(seed, chains, dim, dtype=np.float64, problem=0):
  """Initial positions uniform on (-2, 2) [len(chains), dim]: element e
  from word e % 4 of block (e // 4, chain, problem, INIT), 24 bits, centred
  in its bin (4 (w / 2^8 + 0.5) / 2^24 - 2)."""
  chains = np.asarray(chains, dtype=np.uint64)
  nblk = (dim + 3) // 4
  blk = np.arange(nblk, dtype=np.uint64)
  w = _block(seed, blk[None, :], chains[:, None], problem, PURPOSE_INIT)
  w = w.reshape(len(chains), nblk * 4)[:, :dim]
  u = ((w >> np.uint32(8)).astype(np.float64) + 0.5) * _TWO_M24
  return (4.0 * u - 2.0).astype(dtype)


def mh_uniform(seed, chains, dtype=np.float64, problem=0):
  w = _block(seed, 0, np.asarray(chains, np.uint64), problem, PURPOSE_MH_ACCEPT)
  return uniform_from_words(w, dtype)


def tree_draws(seed, chains, depth, dtype=np.float64, problem=0):
  """(direction bool [C], uniform [C]) for tree doubling at `depth`."""
  w = _block(seed, 0, np.asarray(chains, np.uint64), problem,
             PURPOSE_TREE | int(depth))
  direction = (w[..., 0] >> np.uint32(31)).astype(bool)
  return direction, uniform_from_words(w[..., 1:3], dtype)


def leaf_uniform(seed, chains, depth, step, dtype=np.float64, problem=0):
  w = _block(seed, int(step), np.asarray(chains, np.uint64), problem,
             PURPOSE_LEAF | int(depth))
  return uniform_from_words(w, dtype)


class PhiloxDraws:
  """Draw provider used by the oracle kernels (same stream as the GPU)."""

  def __init__(self, dtype=np.float64, problem=0, chains=None):
    self.dtype = np.dtype(dtype)
    self.problem = problem
    self.chains = chains

  def _ch(self, n):
    return np.arange(n) if self.chains is None else np.asarray(self.chains)

  def momentum(self, seed, n_chains, dim):
    return momentum(seed, self._ch(n_chains), dim, self.dtype, self.problem)

  def mh_uniform(self, seed, n_chains):
    return mh_uniform(seed, self._ch(n_chains), self.dtype, self.problem)

  def tree(self, seed, n_chains, depth):
    return tree_draws(seed, self._ch(n_chains), depth, self.dtype, self.problem)

  def leaf(self, seed, n_chains, depth, step):
    return leaf_uniform(seed, self._ch(n_chains), depth, step, self.dtype,
                        self.problem)


class TableDraws:
  """Injected-randomness provider: explicit tables instead of Philox.

  momentum: [C, D]; mh: [C]; direction: [depth, C]; tree_u: [depth, C];
  leaf_u: [depth, 2**depth_max, C].  `seed` arguments are ignored.
  """

  def __init__(self, momentum=None, mh=None, direction=None, tree_u=None,
               leaf_u=None):
    self._m, self._mh = momentum, mh
    self._dir, self._tu, self._lu = direction, tree_u, leaf_u

  def momentum(self, seed, n_chains, dim):
    return np.array(self._m, copy=True)

  def mh_uniform(self, seed, n_chains):
    return np.array(self._mh, copy=True)

  def tree(self, seed, n_chains, depth):
    return np.asarray(self._dir[depth], bool), np.asarray(self._tu[depth])

  def leaf(self, seed, n_chains, depth, step):
    return np.asarray(self._lu[depth][step])
