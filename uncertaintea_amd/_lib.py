"""ctypes binding of libuncertaintea_hip (the C ABI in include/uncertaintea_hip.h).

The library is loaded from the package directory only.  There is no fallback: if
the shared object is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libuncertaintea_hip.so"

UT_OK = 0
UT_E_UNSUPPORTED = 1
UT_E_DIM = 2
UT_E_PRECISION = 3
UT_E_HIP = 4
UT_E_NONFINITE_INIT = 5
UT_E_ARG = 6
UT_FLOAT32 = 0
UT_FLOAT64 = 1
UT_INIT_PRIOR = 0
UT_INIT_UNIFORM = 1
# general plans (include/uncertaintea_hip.h)
UT_STEP_CHOICE, UT_STEP_DETERMINISTIC, UT_STEP_LOOP = 1, 2, 3
(UT_FAMILY_NORMAL, UT_FAMILY_NONCENTERED_NORMAL, UT_FAMILY_LOGNORMAL, UT_FAMILY_HALFNORMAL, UT_FAMILY_HALFCAUCHY,
 UT_FAMILY_EXPONENTIAL, UT_FAMILY_CAUCHY, UT_FAMILY_LAPLACE, UT_FAMILY_LOGISTIC, UT_FAMILY_STUDENTT, UT_FAMILY_GAMMA,
 UT_FAMILY_BERNOULLI, UT_FAMILY_BERNOULLI_LOGIT, UT_FAMILY_POISSON, UT_FAMILY_DIAG_NORMAL,
 UT_FAMILY_BERNOULLI_LOGIT_GLM, UT_FAMILY_GUMBEL, UT_FAMILY_FRECHET, UT_FAMILY_RAYLEIGH, UT_FAMILY_INVERSEGAUSSIAN,
 UT_FAMILY_BETA, UT_FAMILY_INVERSEGAMMA, UT_FAMILY_WEIBULL, UT_FAMILY_BINOMIAL, UT_FAMILY_GEOMETRIC,
 UT_FAMILY_NEGATIVEBINOMIAL) = range(1, 27)
UT_TRANSFORM_IDENTITY, UT_TRANSFORM_LOG, UT_TRANSFORM_LOGIT = 0, 1, 2
(UT_OP_LITERAL, UT_OP_SLOT, UT_OP_PARAM, UT_OP_OBS, UT_OP_ADD, UT_OP_SUB, UT_OP_MUL, UT_OP_DIV, UT_OP_NEG, UT_OP_EXP,
 UT_OP_LOG, UT_OP_LOG1P, UT_OP_SQRT, UT_OP_POW, UT_OP_ABS, UT_OP_TANH, UT_OP_LOGISTIC) = range(1, 18)
UT_PLAN_GENERIC, UT_PLAN_MATRIX_CORES = 1, 2
(UT_ENGINE_INTERPRETER, UT_ENGINE_LOGISTIC, UT_ENGINE_HIER, UT_ENGINE_FUNNEL, UT_ENGINE_GLM_MATRIX_CORES,
 UT_ENGINE_GP) = range(6)

# names exported by the C ABI (checked by the load test)
EXPORTED = (
    "ut_last_error", "ut_abi_version", "ut_device_available",
    "ut_plan_logistic_glm", "ut_plan_logistic_glm_bf16", "ut_plan_hier_normal_nc", "ut_plan_funnel",
    "ut_plan_gp_poisson", "ut_gp_factor",
    "ut_plan_destroy", "ut_plan_num_params", "ut_plan_validate", "ut_plan_create", "ut_plan_engine",
    "ut_workspace_create", "ut_workspace_destroy", "ut_workspace_set_chain_waves", "ut_workspace_chain_waves",
    "ut_gradient_step_kernel_time",
    "ut_logjoint", "ut_logjoint_gradient", "ut_gradient_kernel_time", "ut_noalloc_scratch_bytes",
    "ut_logjoint_gradient_noalloc",
    "ut_sampler_opts_default", "ut_nuts_run", "ut_nuts_masked_run", "ut_hmc_run", "ut_chees_opts_default", "ut_chees_run",
    "ut_hmc_shared_run", "ut_meads_opts_default", "ut_meads_run", "ut_flow_create", "ut_flow_create_fullrank", "ut_pointwise_loglik_size", "ut_pointwise_loglik", "ut_flow_destroy", "ut_flow_num_weights", "ut_flow_layer_layout",
    "ut_flow_get_weights", "ut_flow_set_weights", "ut_flow_set_base", "ut_flow_train", "ut_flow_forward",
    "ut_flow_inverse", "ut_flow_vjp", "ut_flow_kernel_times", "ut_flow_loss_grad", "ut_workspace_set_flow", "ut_flow_fit_elbo",
    "ut_flow_elbo_grad", "ut_initial_params",
    "ut_philox4x32", "ut_philox_fill", "ut_host_alloc", "ut_host_free", "ut_gemm_bf16",
)


class PlanExpr(ct.Structure):
    _fields_ = [("op", ct.c_int32), ("a", ct.c_int32), ("b", ct.c_int32), ("index", ct.c_int32),
                ("value", ct.c_double)]


class PlanStep(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("family", ct.c_int32), ("value_source", ct.c_int32), ("transform", ct.c_int32),
                ("binding_slot", ct.c_int32), ("args", ct.c_int32 * 3), ("dimension", ct.c_int32),
                ("coef_source", ct.c_int32), ("value_offset", ct.c_int32), ("loop_id", ct.c_int32),
                ("iterator_slot", ct.c_int32), ("body_length", ct.c_int32), ("stride", ct.c_int32)]


class PlanDesc(ct.Structure):
    _fields_ = [("num_params", ct.c_int32), ("num_slots", ct.c_int32), ("num_loops", ct.c_int32),
                ("num_steps", ct.c_int32), ("steps", ct.POINTER(PlanStep)), ("num_exprs", ct.c_int32),
                ("exprs", ct.POINTER(PlanExpr)), ("num_observed", ct.c_int64), ("observed", ct.POINTER(ct.c_double)),
                ("trip_counts", ct.POINTER(ct.c_int32)), ("loop_starts", ct.POINTER(ct.c_int32)),
                ("flags", ct.c_int32)]


class SamplerOpts(ct.Structure):
    _fields_ = [
        ("num_samples", ct.c_int32),
        ("num_warmup", ct.c_int32),
        ("max_tree_depth", ct.c_int32),
        ("num_leapfrog_steps", ct.c_int32),
        ("step_size", ct.c_double),
        ("target_accept", ct.c_double),
        ("divergence_threshold", ct.c_double),
        ("mass_matrix_regularization", ct.c_double),
        ("mass_matrix_min_samples", ct.c_int32),
        ("adapt_step_size", ct.c_int32),
        ("adapt_mass_matrix", ct.c_int32),
        ("find_reasonable_step_size", ct.c_int32),
        ("init_strategy", ct.c_int32),
        ("init_max_retries", ct.c_int32),
        ("initial_params", ct.POINTER(ct.c_double)),
        ("seed", ct.c_uint64),
        ("callback", ct.c_void_p),
        ("callback_user", ct.c_void_p),
        ("callback_every", ct.c_int32),
    ]


# void (*)(void* user, int32_t phase, int32_t iteration, int32_t total, double step_size, int64_t divergences)
PROGRESS_FN = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_double, ct.c_int64)
UT_PHASE_WARMUP = 0
UT_PHASE_SAMPLE = 1


class CheesOpts(ct.Structure):
    _fields_ = [
        ("base", SamplerOpts),
        ("jitter_amount", ct.c_double),
        ("adapt_trajectory_length", ct.c_int32),
        ("max_leapfrog_steps", ct.c_int32),
        ("trajectory_decay_rate", ct.c_double),
        ("trajectory_learning_rate", ct.c_double),
        ("trajectory_adam_beta1", ct.c_double),
        ("trajectory_adam_beta2", ct.c_double),
        ("trajectory_adam_epsilon", ct.c_double),
        ("forward_velocity", ct.c_int32),
    ]


class MeadsOpts(ct.Structure):
    _fields_ = [
        ("num_samples", ct.c_int32),
        ("num_warmup", ct.c_int32),
        ("num_folds", ct.c_int32),
        ("init_strategy", ct.c_int32),
        ("init_max_retries", ct.c_int32),
        ("divergence_threshold", ct.c_double),
        ("initial_params", ct.POINTER(ct.c_double)),
        ("seed", ct.c_uint64),
        ("callback", ct.c_void_p),
        ("callback_user", ct.c_void_p),
        ("callback_every", ct.c_int32),
    ]


class ChainOut(ct.Structure):
    _fields_ = [
        ("samples", ct.c_void_p),
        ("logjoint", ct.POINTER(ct.c_double)),
        ("accept_stat", ct.POINTER(ct.c_double)),
        ("energy", ct.POINTER(ct.c_double)),
        ("energy_error", ct.POINTER(ct.c_double)),
        ("accepted", ct.POINTER(ct.c_uint8)),
        ("divergent", ct.POINTER(ct.c_uint8)),
        ("tree_depth", ct.POINTER(ct.c_int32)),
        ("integration_steps", ct.POINTER(ct.c_int32)),
        ("step_size", ct.POINTER(ct.c_double)),
        ("inverse_mass", ct.POINTER(ct.c_double)),
        ("warmup_seconds", ct.c_double),
        ("sampling_seconds", ct.c_double),
        ("sampling_gradient_evals", ct.c_int64),
        ("sampling_kernel_seconds", ct.c_double),
        ("sampling_launches", ct.c_int32),
    ]


class DeviceError(RuntimeError):
    """A HIP runtime failure reported by libuncertaintea_hip."""


class NonFiniteInitError(ValueError):
    """Initial positions stayed non-finite after the allowed re-draws."""


_lib = None


def lib() -> ct.CDLL:
    """Load (once) and return the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libuncertaintea_hip is not built ({LIB_PATH} missing); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` first")
    L = ct.CDLL(str(LIB_PATH))
    vp, i32, i64, u32, dbl = ct.c_void_p, ct.c_int32, ct.c_int64, ct.c_uint32, ct.c_double
    pp = ct.POINTER(ct.c_void_p)
    sig = {
        "ut_last_error": (ct.c_char_p, []),
        "ut_abi_version": (i32, []),
        "ut_device_available": (i32, []),
        "ut_plan_logistic_glm": (ct.c_int, [vp, vp, i64, i32, dbl, dbl, pp]),
        "ut_plan_logistic_glm_bf16": (ct.c_int, [vp, vp, i64, i32, dbl, dbl, pp]),
        "ut_plan_hier_normal_nc": (ct.c_int, [vp, vp, vp, i32, dbl, dbl, dbl, pp]),
        "ut_plan_funnel": (ct.c_int, [i32, dbl, pp]),
        "ut_plan_gp_poisson": (ct.c_int, [vp, vp, i32, i32, dbl, dbl, dbl, dbl, pp]),
        "ut_gp_factor": (ct.c_int, [vp, vp, i32, vp, vp, vp]),
        "ut_plan_destroy": (ct.c_int, [vp]),
        "ut_plan_num_params": (ct.c_int, [vp, ct.POINTER(i32)]),
        "ut_plan_validate": (ct.c_int, [ct.POINTER(PlanDesc)]),
        "ut_plan_create": (ct.c_int, [ct.POINTER(PlanDesc), pp]),
        "ut_plan_engine": (ct.c_int, [vp, ct.POINTER(i32)]),
        "ut_workspace_create": (ct.c_int, [vp, i32, i32, pp]),
        "ut_workspace_destroy": (ct.c_int, [vp]),
        "ut_workspace_set_chain_waves": (ct.c_int, [vp, i32]),
        "ut_gradient_step_kernel_time": (ct.c_int, [vp, vp, i32, vp]),
        "ut_workspace_chain_waves": (ct.c_int, [vp, vp]),
        "ut_logjoint": (ct.c_int, [vp, vp, vp]),
        "ut_logjoint_gradient": (ct.c_int, [vp, vp, vp, vp]),
        "ut_gradient_kernel_time": (ct.c_int, [vp, vp, i32, i32, vp]),
        "ut_noalloc_scratch_bytes": (i64, [vp, i32]),
        "ut_logjoint_gradient_noalloc": (ct.c_int, [vp, i32, vp, vp, vp, vp]),
        "ut_sampler_opts_default": (None, [ct.POINTER(SamplerOpts)]),
        "ut_nuts_run": (ct.c_int, [vp, ct.POINTER(SamplerOpts), ct.POINTER(ChainOut)]),
        "ut_nuts_masked_run": (ct.c_int, [vp, ct.POINTER(SamplerOpts), ct.POINTER(ChainOut)]),
        "ut_hmc_run": (ct.c_int, [vp, ct.POINTER(SamplerOpts), ct.POINTER(ChainOut)]),
        "ut_chees_opts_default": (None, [ct.POINTER(CheesOpts)]),
        "ut_chees_run": (ct.c_int, [vp, ct.POINTER(CheesOpts), ct.POINTER(ChainOut), vp]),
        "ut_hmc_shared_run": (ct.c_int, [vp, ct.POINTER(SamplerOpts), ct.POINTER(ChainOut)]),
        "ut_meads_opts_default": (None, [ct.POINTER(MeadsOpts)]),
        "ut_meads_run": (ct.c_int, [vp, ct.POINTER(MeadsOpts), ct.POINTER(ChainOut), vp]),
        "ut_flow_create": (ct.c_int, [i32, i32, i32, dbl, ct.c_uint64, pp]),
        "ut_flow_create_fullrank": (ct.c_int, [i32, pp]),
        "ut_pointwise_loglik_size": (ct.c_int, [vp, ct.POINTER(i64)]),
        "ut_pointwise_loglik": (ct.c_int, [vp, vp, i64, vp]),
        "ut_flow_destroy": (ct.c_int, [vp]),
        "ut_flow_num_weights": (ct.c_int, [vp, ct.POINTER(i64)]),
        "ut_flow_layer_layout": (ct.c_int, [vp, i32, ct.POINTER(i32), ct.POINTER(i32), vp]),
        "ut_flow_get_weights": (ct.c_int, [vp, vp]),
        "ut_flow_set_weights": (ct.c_int, [vp, vp]),
        "ut_flow_set_base": (ct.c_int, [vp, vp, vp]),
        "ut_flow_train": (ct.c_int, [vp, vp, i64, i32, i32, dbl, ct.c_uint64, vp]),
        "ut_flow_forward": (ct.c_int, [vp, vp, i64, vp, vp]),
        "ut_flow_inverse": (ct.c_int, [vp, vp, i64, vp, vp]),
        "ut_flow_vjp": (ct.c_int, [vp, vp, i64, vp, vp]),
        "ut_flow_kernel_times": (ct.c_int, [vp, i64, i32, vp, vp]),
        "ut_flow_loss_grad": (ct.c_int, [vp, vp, i64, ct.POINTER(dbl), vp]),
        "ut_workspace_set_flow": (ct.c_int, [vp, vp]),
        "ut_flow_fit_elbo": (ct.c_int, [vp, vp, i32, dbl, dbl, dbl, dbl, dbl, i32, ct.c_uint64, vp, vp, vp, vp,
                                         ct.POINTER(dbl)]),
        "ut_initial_params": (ct.c_int, [vp, i32, ct.c_uint64, i32, vp]),
        "ut_flow_elbo_grad": (ct.c_int, [vp, vp, vp, i32, ct.POINTER(dbl), vp, vp, vp, vp]),
        "ut_philox4x32": (None, [vp, vp, vp]),
        "ut_philox_fill": (ct.c_int, [u32, u32, u32, u32, i64, vp, vp, vp, vp]),
        "ut_host_alloc": (ct.c_int, [i64, pp]),
        "ut_host_free": (ct.c_int, [vp]),
        "ut_gemm_bf16": (ct.c_int, [vp, vp, vp, i32, i32, i32, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    """Map a status code to the reference's exception classes."""
    if rc == UT_OK:
        return
    msg = lib().ut_last_error().decode(errors="replace")
    if rc == UT_E_DIM:
        raise ValueError(f"DimensionMismatch: {msg}")
    if rc == UT_E_HIP:
        raise DeviceError(msg)
    if rc == UT_E_NONFINITE_INIT:
        raise NonFiniteInitError(msg)
    raise ValueError(msg)  # UT_E_ARG / UT_E_UNSUPPORTED / UT_E_PRECISION: ArgumentError


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def dptr(a: np.ndarray | None, ctype):
    return None if a is None else a.ctypes.data_as(ct.POINTER(ctype))


class _PinnedBlock:
    """Owner of one page-locked allocation, exposed through the numpy array
    interface; numpy arrays made from it keep it alive and free it last."""

    def __init__(self, shape, dtype):
        self.ptr = ct.c_void_p()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        check(lib().ut_host_alloc(nbytes, ct.byref(self.ptr)))
        self._free = lib().ut_host_free  # kept: module globals may be gone at exit

    @property
    def __array_interface__(self):
        return {"data": (self.ptr.value, False), "shape": self.shape, "typestr": self.dtype.str, "version": 3}

    def __del__(self):
        if self.ptr and getattr(self, "_free", None) is not None:
            self._free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype) -> np.ndarray:
    """numpy array backed by page-locked host memory (allocated by the library)."""
    if int(np.prod(shape)) == 0:
        return np.empty(shape, dtype=dtype)
    return np.asarray(_PinnedBlock(shape, dtype))


def gemm_bf16(a_bf16: np.ndarray, bt_bf16: np.ndarray, splits: int = 1) -> np.ndarray:
    """C = A @ Bt.T on the matrix cores; A (m x k), Bt (n x k) as uint16 bfloat16 bits."""
    a = np.ascontiguousarray(a_bf16, dtype=np.uint16)
    b = np.ascontiguousarray(bt_bf16, dtype=np.uint16)
    m, k = a.shape
    n = b.shape[0]
    c = np.empty((m, n), dtype=np.float32)
    check(lib().ut_gemm_bf16(ptr(a), ptr(b), ptr(c), m, n, k, splits))
    return c


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even float32 -> bfloat16 bit patterns (uint16)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def device_available() -> bool:
    return bool(lib().ut_device_available())


def philox4x32(key, ctr) -> np.ndarray:
    k = np.ascontiguousarray(key, dtype=np.uint32)
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().ut_philox4x32(ptr(k), ptr(c), ptr(out))
    return out


def philox_fill(chain: int, iteration: int, stream: int, first: int, n: int):
    """Device draws of n consecutive Philox coordinates: (blocks, u32f, normal32, u64f)."""
    blocks = np.zeros((n, 4), dtype=np.uint32)
    uf = np.zeros(n, dtype=np.float32)
    nf = np.zeros(n, dtype=np.float32)
    ud = np.zeros(n, dtype=np.float64)
    check(lib().ut_philox_fill(chain, iteration, stream, first, n, ptr(blocks), ptr(uf), ptr(nf), ptr(ud)))
    return blocks, uf, nf, ud
