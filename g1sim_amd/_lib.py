"""ctypes binding of libg1sim_amd.so (C-ABI in include/g1sim_amd.h).

The library is built in-tree by g1sim_amd.build.  There is no fallback: if
it is missing or cannot be loaded, load() raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg1sim_amd.so")

G1_FLAG_F64_SOLVE = 1
G1_FLAG_F64_BASE = 2
G1_FLAG_F64_CONTACT = 4
G1_FLAG_WAVES4 = 8
G1_FLAG_WAVES2 = 16

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m g1sim_amd.build` "
                           "(the HIP step engine has no CPU fallback)")
    # torch's HIP runtime must be the one in the process when torch tensors and
    # streams are passed; importing torch first makes the loader reuse it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp, i, u, f = C.c_void_p, C.c_int, C.c_uint, C.c_float
    pa = C.POINTER(C.c_void_p)           # host array of device pointers
    lib.g1_cuda_step.restype = None
    lib.g1_cuda_step.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_cuda_reset_done.restype = None
    lib.g1_cuda_reset_done.argtypes = [i, vp, vp, vp, vp, u, f, f, vp]
    lib.g1_hip_step_ex.restype = i
    lib.g1_hip_step_ex.argtypes = [i, i, f, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u, vp]
    lib.g1_hip_step_schedule.restype = i
    lib.g1_hip_step_schedule.argtypes = [i, i, i, i, f, vp, vp, vp, vp, C.c_longlong, vp, vp, vp, vp, vp, u, vp]
    lib.g1_mlp_gemm.restype = i
    lib.g1_mlp_gemm.argtypes = [i, i, i, vp, i, vp, i, i, i, i, i, vp, vp, i, vp, vp, vp]
    lib.g1_mlp_gemm_ex.restype = i
    lib.g1_mlp_gemm_ex.argtypes = [i, i, i, vp, i, vp, i, i, i, i, i, vp, vp, i, vp, vp, vp, vp]
    lib.g1_mlp_backward_layer.restype = i
    lib.g1_mlp_backward_layer.argtypes = [vp, i, vp, i, vp, i, i, vp, vp, i, vp, vp, vp]
    lib.g1_mlp_mse_grad_part.restype = i
    lib.g1_mlp_mse_grad_part.argtypes = [vp, vp, i, i, i, vp, vp, vp, vp]
    lib.g1_mlp_mse_grad.restype = i
    lib.g1_mlp_mse_grad.argtypes = [vp, vp, i, i, i, vp, vp, vp, vp]
    lib.g1_mlp_sum_splits.restype = i
    lib.g1_mlp_sum_splits.argtypes = [vp, i, C.c_longlong, vp, vp]
    lib.g1_mlp_colsum.restype = i
    lib.g1_mlp_colsum.argtypes = [vp, i, i, vp, vp]
    lib.g1_mlp_colsum_det.restype = i
    lib.g1_mlp_colsum_det.argtypes = [vp, i, i, vp, vp, vp]
    lib.g1_mlp_train_schedule.restype = i
    lib.g1_mlp_train_schedule.argtypes = [vp, vp, i, f, f, f, f, vp]
    lib.g1_mlp_sample.restype = i
    lib.g1_mlp_sample.argtypes = [i, i, i, i, u, vp, vp, vp]
    lib.g1_mlp_adam_multi.restype = i
    lib.g1_mlp_adam_multi.argtypes = [i, pa, pa, pa, pa, pa, pa, C.POINTER(C.c_int), pa, C.POINTER(C.c_int),
                                      C.POINTER(C.c_float), vp, f, f, f, vp]
    lib.g1_mlp_adam_multi_sum.restype = i
    lib.g1_mlp_adam_multi_sum.argtypes = [i, pa, pa, pa, pa, pa, pa, C.POINTER(C.c_int), pa, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_float), vp, f,
                                          f, f, vp]
    lib.g1_mlp_adam.restype = i
    lib.g1_mlp_adam.argtypes = [vp, vp, vp, vp, vp, vp, i, f, f, f, f, i, f, vp]
    lib.g1_mlp_fill.restype = i
    lib.g1_mlp_fill.argtypes = [vp, f, C.c_longlong, vp]
    lib.g1_mlp_to_bf16.restype = i
    lib.g1_mlp_to_bf16.argtypes = [vp, vp, C.c_longlong, vp]
    lib.g1_mlp_gather.restype = i
    lib.g1_mlp_gather.argtypes = [vp, vp, vp, i, i, i, i, vp, vp, vp, vp, vp, i, vp, i, vp]
    lib.g1_hip_surrogate_steps.restype = i
    lib.g1_hip_surrogate_steps.argtypes = [i, i, i, f, i, f, vp, vp, vp, vp, vp, vp, vp, pa, pa, vp, vp, vp, vp, vp,
                                           vp, vp, vp, vp, vp, u, vp]
    lib.g1_mlp_gather_rows.restype = i
    lib.g1_mlp_gather_rows.argtypes = [vp, vp, i, i, i, i, vp, vp, vp, vp, vp, i, vp, i, vp]
    lib.g1_mlp_features.restype = i
    lib.g1_mlp_features.argtypes = [i, i, vp, vp, vp, vp, vp, vp, i, vp]
    lib.g1_hip_forces.restype = i
    lib.g1_hip_forces.argtypes = [i, f, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_hip_judge_list.restype = i
    lib.g1_hip_judge_list.argtypes = [i, i, i, vp, vp, vp, vp]
    lib.g1_hip_step_judge.restype = i
    lib.g1_hip_step_judge.argtypes = [i, f, vp, vp, vp, vp, vp, vp, vp, vp, u, vp]
    lib.g1_hip_step_judge_blocks.restype = i
    lib.g1_hip_step_judge_blocks.argtypes = [i, f, vp, vp, vp, vp, vp, vp, i, i, vp, u, vp]
    lib.g1_hip_judge_forces.restype = i
    lib.g1_hip_judge_forces.argtypes = [i, f, vp, vp, vp, vp, vp, vp, i, i, vp, vp, vp, vp, vp, u, vp]
    lib.g1_hip_surrogate_integrate.restype = i
    lib.g1_hip_surrogate_integrate.argtypes = [i, f, i, i, f, vp, vp, vp, i, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_hip_surrogate_forward.restype = i
    lib.g1_hip_surrogate_forward.argtypes = [i, vp, vp, vp, vp, vp, pa, pa, vp, i, vp]
    lib.g1_hip_surrogate_step.restype = i
    lib.g1_hip_surrogate_step.argtypes = [i, f, i, i, f, vp, vp, vp, vp, vp, pa, pa, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_hip_surrogate_step_deferred.restype = i
    lib.g1_hip_surrogate_step_deferred.argtypes = [i, f, i, i, f, vp, vp, vp, vp, vp, pa, pa, vp, vp, vp, vp, vp, vp,
                                                   vp]
    lib.g1_hip_judge_forces_commit.restype = i
    lib.g1_hip_judge_forces_commit.argtypes = [i, f, vp, vp, vp, vp, vp, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, u,
                                               vp]
    lib.g1_hip_surrogate_judge_batch.restype = i
    lib.g1_hip_surrogate_judge_batch.argtypes = [i, f, vp, vp, pa, pa, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_hip_surrogate_steps_deferred.restype = i
    lib.g1_hip_surrogate_steps_deferred.argtypes = [i, i, i, f, i, f, vp, vp, vp, vp, vp, vp, vp, pa, pa, vp, vp, vp,
                                                    vp, vp, vp, vp, vp, vp, vp, u, vp, vp, vp, C.POINTER(i), vp]
    lib.g1_hip_surrogate_steps_schedule.restype = i
    lib.g1_hip_surrogate_steps_schedule.argtypes = [i, i, i, f, i, f, vp, vp, vp, vp, C.c_longlong, i, i, vp, vp, vp,
                                                    pa, pa, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u, vp, vp, vp,
                                                    C.POINTER(i), vp]
    lib.g1_hip_fallback_run.restype = i
    lib.g1_hip_fallback_run.argtypes = [i, i, f, vp, vp, vp, vp, vp, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, u, vp]
    lib.g1_hip_get_obs.restype = i
    lib.g1_hip_get_obs.argtypes = [i, i, vp, vp, vp, vp, vp, vp]
    lib.g1_policy_act.restype = i
    lib.g1_policy_act.argtypes = [i, i, vp, vp, vp, vp, vp, vp, f, pa, pa, vp, u, i, vp, vp, vp, vp, vp]
    lib.g1_rollout_post.restype = i
    lib.g1_rollout_post.argtypes = [i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i, f, f, u, i, vp, vp, vp, vp, vp,
                                    vp]
    lib.g1_policy_obs_norm_update.restype = i
    lib.g1_policy_obs_norm_update.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_policy_obs_norm_update_noalloc.restype = i
    lib.g1_policy_obs_norm_update_noalloc.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.g1_rollout_init.restype = i
    lib.g1_rollout_init.argtypes = [i, f, f, u, vp, vp, vp, vp, vp]
    for name in ("g1_hip_peak_fp32", "g1_hip_peak_bf16"):
        getattr(lib, name).restype = i
        getattr(lib, name).argtypes = [i, i, vp, vp]
    lib.g1_hip_last_error.restype = C.c_char_p
    lib.g1_hip_last_error.argtypes = []
    lib.g1_hip_version.restype = i
    lib.g1_hip_version.argtypes = []
    _lib = lib
    return lib


def exported_symbols():
    return ["g1_cuda_step", "g1_cuda_reset_done", "g1_hip_step_ex", "g1_hip_last_error",
            "g1_hip_version", "g1_mlp_gemm", "g1_mlp_mse_grad", "g1_mlp_colsum", "g1_mlp_sum_splits", "g1_mlp_adam",
            "g1_mlp_fill", "g1_mlp_to_bf16", "g1_mlp_gather", "g1_mlp_features", "g1_hip_forces",
            "g1_hip_judge_list", "g1_hip_step_judge", "g1_hip_step_judge_blocks", "g1_hip_judge_forces",
            "g1_hip_surrogate_integrate", "g1_hip_surrogate_forward", "g1_hip_surrogate_step", "g1_hip_get_obs",
            "g1_policy_act", "g1_rollout_post", "g1_rollout_init", "g1_policy_obs_norm_update", "g1_hip_peak_fp32",
            "g1_hip_peak_bf16", "g1_mlp_colsum_det", "g1_mlp_train_schedule", "g1_mlp_sample", "g1_mlp_adam_multi",
            "g1_policy_obs_norm_update_noalloc", "g1_hip_step_schedule", "g1_mlp_gemm_ex", "g1_mlp_backward_layer", "g1_mlp_mse_grad_part",
            "g1_mlp_adam_multi_sum", "g1_mlp_gather_rows", "g1_hip_surrogate_steps", "g1_hip_surrogate_step_deferred",
            "g1_hip_surrogate_judge_batch", "g1_hip_surrogate_steps_deferred", "g1_hip_judge_forces_commit",
            "g1_hip_fallback_run", "g1_hip_surrogate_steps_schedule"]


def ptr_array(tensors_or_ptrs):
    """Host array of device pointers (the W / b arguments of the fused
    inference entry points)."""
    vals = [None if t is None else t if isinstance(t, int) else t.data_ptr() for t in tensors_or_ptrs]
    return (C.c_void_p * len(vals))(*vals)


def last_error():
    return load().g1_hip_last_error().decode()


def check(rc, what="g1_hip_step_ex"):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
