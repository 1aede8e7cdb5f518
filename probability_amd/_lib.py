"""ctypes binding of libhipmc.so (the C ABI declared in include/hipmc.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing or a call fails, an exception is raised.
"""

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libhipmc.so')

F32, F64 = 0, 1
TARGET_LOGISTIC, TARGET_LINEAR, TARGET_HIER_NORMAL = 1, 2, 3
TARGET_DIAG_NORMAL, TARGET_FUNNEL, TARGET_NAN_EXP = 4, 5, 6
TARGET_LOG_GAMMA = 7
BIJ_IDENTITY, BIJ_SOFTPLUS, BIJ_EXP = 0, 1, 2
MAX_BIJ_RANGES = 4

E_INVALID, E_UNSUPPORTED, E_HIP = -1, -2, -3
REDUCE_MEAN, REDUCE_HARMONIC = 0, 1

EXPORTED_SYMBOLS = (
    'hipmc_last_error', 'hipmc_abi_version', 'hipmc_sizeof_problem',
    'hipmc_sizeof_run_args', 'hipmc_sizeof_stream',
    'hipmc_sizeof_glm_hmc_args', 'hipmc_glm_hmc_run', 'hipmc_seed_split', 'hipmc_seed_fold_in',
    'hipmc_seed_salt', 'hipmc_philox_host', 'hipmc_hmc_run',
    'hipmc_leapfrog_run', 'hipmc_nuts_run', 'hipmc_target_eval',
    'hipmc_da_update', 'hipmc_simple_step_update', 'hipmc_mh_accept',
    'hipmc_bijector', 'hipmc_draws', 'hipmc_moments_update',
    'hipmc_variance_update', 'hipmc_glm_hmc_run_noalloc')

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_f64 = ctypes.c_double


class Problem(ctypes.Structure):
  _fields_ = [
      ('kind', _i32), ('dim', _i32), ('n_obs', _i32), ('n_chains', _i32),
      ('chain_offset', _i64),
      ('x', _vp), ('xt', _vp), ('y', _vp), ('aux', _vp),
      ('param', _f64 * 4), ('log_norm', _f64),
      ('n_bij', _i32),
      ('bij_off', _i32 * MAX_BIJ_RANGES),
      ('bij_len', _i32 * MAX_BIJ_RANGES),
      ('bij_code', _i32 * MAX_BIJ_RANGES),
      ('key', ctypes.c_uint32),
      ('x_pitch', _i32), ('xt_pitch', _i32),
  ]


class Stream(ctypes.Structure):
  _fields_ = [
      ('seed', ctypes.c_uint32 * 2), ('iter0', _i64), ('n_iters', _i32),
      ('trace_every', _i32), ('trace', _vp), ('trace_start', _i64),
      ('n_trace', _i32), ('pad_', _i32),
  ]


class RunArgs(ctypes.Structure):
  _fields_ = [
      ('dtype', _i32), ('n_chains', _i32), ('dim_stride', _i32),
      ('max_dim', _i32), ('n_iters', _i32), ('fold_seed', _i32),
      ('iter0', _i64), ('seed', ctypes.c_uint32 * 2),
      ('num_leapfrog_steps', _i32), ('max_tree_depth', _i32),
      ('unrolled_leapfrog_steps', _i32), ('target_kind', _i32),
      ('max_energy_diff', _f64),
      ('problems', _vp), ('n_problems', _i32), ('waves_per_chain', _i32),
      ('chain_problem', _vp), ('chain_local', _vp),
      ('step_size', _vp), ('chain_step', _vp),
      ('state', _vp), ('tlp', _vp), ('grad', _vp),
      ('out_log_accept_ratio', _vp), ('out_leapfrogs', _vp),
      ('out_is_accepted', _vp), ('out_reach_max_depth', _vp),
      ('out_has_divergence', _vp), ('out_energy', _vp),
      ('out_tree_depth', _vp), ('out_row0', _i64), ('trace_state', _vp),
      ('trace_start', _i64), ('trace_every', _i32), ('n_trace', _i32),
      ('out_proposed_state', _vp), ('out_proposed_tlp', _vp),
      ('out_initial_momentum', _vp), ('out_final_momentum', _vp),
      ('streams', _vp), ('chain_row', _vp),
      ('out_target_log_prob', _vp), ('momentum_precision', _vp),
      ('chains_per_block', _i32), ('pad0_', _i32),
  ]


class GlmHmcArgs(ctypes.Structure):
  _fields_ = [
      ('n_obs', _i32), ('dim', _i32), ('n_chains', _i32), ('n_split', _i32),
      ('num_leapfrog_steps', _i32), ('n_iters', _i32), ('fold_seed', _i32),
      ('key', ctypes.c_uint32), ('iter0', _i64), ('seed', ctypes.c_uint32 * 2),
      ('prior_scale', _f64), ('x_bf16', _vp), ('xt_bf16', _vp), ('y', _vp),
      ('step_size', _vp), ('per_chain_step', _i32), ('want_grad_only', _i32),
      ('state', _vp), ('tlp', _vp), ('grad', _vp), ('ws_g', _vp),
      ('ws_ll', _vp), ('ws_q', _vp), ('ws_p', _vp), ('ws_b', _vp),
      ('ws_ke0', _vp), ('out_log_accept_ratio', _vp),
      ('out_is_accepted', _vp), ('trace_state', _vp), ('trace_start', _i64),
      ('trace_every', _i32), ('n_trace', _i32),
  ]


class HipmcError(RuntimeError):
  pass


_lock = threading.Lock()
_lib = None


def lib():
  """Loads libhipmc.so once; raises if it has not been built."""
  global _lib
  with _lock:
    if _lib is None:
      if not os.path.exists(LIB_PATH):
        raise HipmcError(
            f'{LIB_PATH} not found: build it with `python -m '
            'probability_amd.build` (there is no CPU fallback).')
      l = ctypes.CDLL(LIB_PATH)
      l.hipmc_last_error.restype = ctypes.c_char_p
      for name in ('hipmc_hmc_run', 'hipmc_leapfrog_run', 'hipmc_nuts_run'):
        getattr(l, name).argtypes = [ctypes.POINTER(RunArgs), _u64]
      l.hipmc_glm_hmc_run.argtypes = [ctypes.POINTER(GlmHmcArgs), _u64]
      l.hipmc_glm_hmc_run_noalloc.argtypes = [ctypes.POINTER(GlmHmcArgs),
                                              _u64]
      l.hipmc_target_eval.argtypes = [_i32, _i32, _vp, _vp, _i32, _i32, _i32,
                                      _vp, _vp, _vp, _u64]
      l.hipmc_da_update.argtypes = [_i32, _i32, _vp, _vp, _vp, _vp, _i32,
                                    _i64, _vp, _vp, _vp, _vp, _i64, _i64,
                                    _f64, _f64, _f64, _i32, _vp, _vp, _vp,
                                    _vp, _u64]
      l.hipmc_variance_update.argtypes = [_i32, _i32, _vp, _vp, _vp, _i32,
                                          _i32, _vp, _vp, _vp, _u64]
      l.hipmc_simple_step_update.argtypes = [_i32, _i32, _vp, _vp, _vp, _i32,
                                             _i64, _vp, _f64, _i64, _i64, _vp,
                                             _u64]
      l.hipmc_mh_accept.argtypes = [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _u64]
      l.hipmc_moments_update.argtypes = [_i32, _i32, _i64, _i64, _i64, _vp,
                                         _f64, _vp, _vp, _vp, _vp, _u64]
      l.hipmc_bijector.argtypes = [_i32, _vp, _vp, _i32, _i32, _i32, _i32,
                                   _vp, _vp, _u64]
      l.hipmc_draws.argtypes = [_i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32,
                                _vp, _vp, _u64]
      l.hipmc_seed_split.argtypes = [_vp, ctypes.c_uint32, _vp]
      l.hipmc_seed_fold_in.argtypes = [_vp, _u64, _vp]
      l.hipmc_seed_salt.argtypes = [_vp, ctypes.c_uint32, _vp]
      l.hipmc_philox_host.argtypes = [_vp, _vp, _vp]
      if l.hipmc_sizeof_problem() != ctypes.sizeof(Problem):
        raise HipmcError('hipmc_problem layout mismatch')
      if l.hipmc_sizeof_run_args() != ctypes.sizeof(RunArgs):
        raise HipmcError('hipmc_run_args layout mismatch')
      if l.hipmc_sizeof_stream() != ctypes.sizeof(Stream):
        raise HipmcError('hipmc_stream layout mismatch')
      if l.hipmc_sizeof_glm_hmc_args() != ctypes.sizeof(GlmHmcArgs):
        raise HipmcError('hipmc_glm_hmc_args layout mismatch')
      _lib = l
  return _lib


def check(rc):
  if rc != 0:
    msg = lib().hipmc_last_error().decode()
    if rc == E_INVALID:
      raise ValueError(msg)
    if rc == E_UNSUPPORTED:
      raise NotImplementedError(msg)
    raise HipmcError(f'hipmc error {rc}: {msg}')


def ptr(t):
  """Device (or host) address of a tensor, or None."""
  if t is None:
    return None
  return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
  return torch.cuda.current_stream(device).cuda_stream


def glm_waves_per_chain(max_rows_times_cols):
  """Waves that share one GLM chain's gradient for a launch whose largest
  design matrix has `max_rows_times_cols` entries: a wave's reads of X are
  bounded by the loads it keeps in flight, so big problems split their rows
  over up to eight waves (one chain per CU)."""
  if max_rows_times_cols >= 262144:
    return 8
  if max_rows_times_cols >= 16384:
    return 4
  if max_rows_times_cols >= 4096:
    return 2
  return 1


def dtype_code(dtype):
  if dtype == torch.float32:
    return F32
  if dtype == torch.float64:
    return F64
  raise TypeError(f'unsupported dtype {dtype}; float32 and float64 only')


def require_cuda():
  if not torch.cuda.is_available():
    raise HipmcError('no GPU visible: the hipmc kernels need a ROCm device '
                     '(there is no CPU fallback)')
