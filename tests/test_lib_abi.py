"""C-ABI of the product library and the oracle: they load and export every
symbol their headers declare.  No compute call (there is no GPU here); the
argument checks of g1_hip_step_ex return before any HIP call."""
import ctypes as C
import os
import re

import pytest

from g1sim_amd import _lib

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared(header):
    txt = open(os.path.join(ROOT, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)))


def test_product_exports_every_declared_symbol():
    names = declared("include/g1sim_amd.h")
    assert {"g1_cuda_step", "g1_cuda_reset_done", "g1_hip_step_ex"} <= set(names)
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(_lib.exported_symbols())
    assert lib.g1_hip_version() >= 100


def test_oracle_exports_every_declared_symbol():
    from oracle import oracle as O
    names = declared("oracle/g1_oracle.h")
    for lib in (O.lib64(), O.lib32()):
        for n in names:
            assert hasattr(lib, n), n


def test_reference_host_abi_names_kept():
    """The oracle keeps the reference host C-ABI (src/host_lib.cpp:7-53)."""
    names = declared("oracle/g1_oracle.h")
    assert {"g1_sizes", "g1_c_fk", "g1_c_fd", "g1_c_step"} <= set(names)


@pytest.mark.parametrize("args,msg", [
    ((-1, 1, 2e-3), "must be >= 0"),
    ((4, -2, 2e-3), "must be >= 0"),
    ((4, 1, 0.0), "dt must be"),
    ((4, 1, float("nan")), "dt must be"),
])
def test_step_ex_rejects_bad_arguments(args, msg):
    lib = _lib.load()
    nenv, nsub, dt = args
    dummy = C.c_void_p(16)
    rc = lib.g1_hip_step_ex(nenv, nsub, dt, dummy, dummy, dummy, dummy, None, None, None, None, None,
                            None, None, 0, None)
    assert rc == -1
    assert msg in _lib.last_error()


def test_step_ex_requires_state_pointers_and_consistent_flags():
    lib = _lib.load()
    rc = lib.g1_hip_step_ex(4, 1, 2e-3, None, None, None, None, None, None, None, None, None, None, None, 0, None)
    assert rc == -1 and "required" in _lib.last_error()
    d = C.c_void_p(16)
    rc = lib.g1_hip_step_ex(4, 1, 2e-3, d, d, d, d, None, None, None, None, None, None, None,
                            _lib.G1_FLAG_F64_BASE, None)
    assert rc == -1 and "base_pose64" in _lib.last_error()
    rc = lib.g1_hip_step_ex(4, 1, 2e-3, d, d, d, d, None, None, None, None, None, None, None, 64, None)
    assert rc == -1 and "flag" in _lib.last_error()


def test_zero_envs_is_a_no_op():
    lib = _lib.load()
    assert lib.g1_hip_step_ex(0, 5, 2e-3, None, None, None, None, None, None, None, None, None,
                              None, None, 0, None) == 0


def test_fused_inference_entry_points_validate_arguments():
    """Argument checks of the fused inference and rollout entry points run on
    the host, before any launch."""
    lib = _lib.load()
    d = C.c_void_p(16)
    ws = _lib.ptr_array([16] * 5)
    wmiss = _lib.ptr_array([16, 16, 0, 16, 16])
    # ldp below the 64 output columns, a missing weight, a bad period
    assert lib.g1_hip_surrogate_forward(8, d, d, d, d, d, ws, ws, d, 35, None) == -1
    assert lib.g1_hip_surrogate_forward(8, d, d, d, d, d, wmiss, ws, d, 64, None) == -1
    assert lib.g1_hip_surrogate_forward(0, d, d, d, d, d, ws, ws, d, 64, None) == 0
    assert lib.g1_hip_surrogate_step(8, 2e-3, 3, 2, 1e-3, d, d, d, d, d, ws, ws, d, d, d, None, None, None,
                                     None) == -1
    assert lib.g1_hip_surrogate_step(8, 0.0, 0, 16, 1e-3, d, d, d, d, d, ws, ws, d, d, d, None, None, None,
                                     None) == -1
    # observation layout, policy clip and weights
    assert lib.g1_hip_get_obs(8, 2, d, d, d, d, d, None) == -1
    assert lib.g1_hip_get_obs(0, 0, d, d, d, d, d, None) == 0
    w4 = _lib.ptr_array([16] * 4)
    assert lib.g1_policy_act(8, 1, d, d, d, d, d, d, 0.0, w4, w4, None, 0, 0, d, None, None, None, None) == -1
    assert lib.g1_policy_act(8, 3, d, d, d, d, d, d, 5.0, w4, w4, None, 0, 0, d, None, None, None, None) == -1
    assert lib.g1_policy_act(8, 1, d, d, d, d, d, d, 5.0, _lib.ptr_array([16, 0, 16, 16]), w4, None, 0, 0, d,
                             None, None, None, None) == -1
    # rollout bookkeeping needs every buffer except push_interval, and a positive episode length
    assert lib.g1_rollout_post(8, d, d, d, d, d, d, d, d, d, d, None, 0, 1.0, 0.5, 0, 0, d, d, d, None, None,
                               None) == -1
    assert lib.g1_rollout_post(8, d, d, d, d, d, None, d, d, d, d, None, 10, 1.0, 0.5, 0, 0, d, d, d, None, None,
                               None) == -1
    assert lib.g1_rollout_init(8, 1.0, 0.5, 0, None, d, d, None, None) == -1


def test_every_failure_names_its_entry_point():
    """A rejected call leaves its reason in g1_hip_last_error, and the next
    successful call clears it."""
    lib = _lib.load()
    d = C.c_void_p(16)
    ws = _lib.ptr_array([16] * 5)
    assert lib.g1_hip_surrogate_forward(8, d, d, d, d, d, ws, ws, d, 35, None) == -1
    assert "g1_hip_surrogate_forward" in _lib.last_error()
    assert lib.g1_hip_get_obs(0, 0, d, d, d, d, d, None) == 0
    assert _lib.last_error() == ""
    assert lib.g1_mlp_gemm(0, 0, 0, d, 64, d, 64, 100, 64, 64, 1, d, None, 64, d, None, None) == -1
    assert "g1_mlp_gemm" in _lib.last_error() and "tile" in _lib.last_error()
    assert lib.g1_hip_step_judge_blocks(8, 2e-3, d, d, d, d, None, None, 16, 16, d, 0, None) == -1
    assert "phase" in _lib.last_error()
    assert lib.g1_hip_peak_fp32(0, 1, d, None) == -1
    assert "g1_hip_peak_fp32" in _lib.last_error()


@pytest.mark.parametrize("args,msg", [
    ((4, 5, 0, 0, 1e3), "period"),
    ((4, -1, 10, 0, 1e3), "nsteps"),
    ((4, 5, 10, -2, 1e3), "step0"),
    ((4, 5, 10, 0, 100), "block_stride"),
])
def test_step_schedule_rejects_bad_arguments(args, msg):
    lib = _lib.load()
    nenv, nsteps, period, step0, stride = args
    dummy = C.c_void_p(16)
    rc = lib.g1_hip_step_schedule(nenv, nsteps, period, step0, 2e-3, dummy, dummy, dummy, dummy, int(stride),
                                  None, None, None, None, None, 0, None)
    assert rc == -1
    assert msg in _lib.last_error()


def test_training_and_stepper_entry_points_validate_arguments():
    """Host-side checks of the fused-training and multi-step entry points
    (before any launch)."""
    lib = _lib.load()
    d = C.c_void_p(16)
    # column partials belong to the SILU_BWD epilogue
    assert lib.g1_mlp_gemm_ex(0, 0, 0, d, 128, d, 128, 128, 128, 128, 1, d, None, 128, None, None, d, None) == -1
    assert "SILU_BWD" in _lib.last_error()
    # SILU_BWD with partials may leave the fp32 output out, not H
    assert lib.g1_mlp_gemm_ex(0, 0, 5, d, 128, d, 128, 128, 128, 128, 1, None, None, 128, None, d, d, None) == -1
    # loss partials: the batch a multiple of 64, ld a multiple of 32
    assert lib.g1_mlp_mse_grad_part(d, d, 100, 128, 35, d, d, d, None) == -1
    assert lib.g1_mlp_mse_grad_part(d, d, 128, 100, 35, d, d, d, None) == -1
    assert lib.g1_mlp_mse_grad_part(d, d, 128, 128, 129, d, d, d, None) == -1
    # Adam over slices: slices closer than the tensor, no tensors
    n = (C.c_int * 1)(100)
    one = _lib.ptr_array([16])
    assert lib.g1_mlp_adam_multi_sum(1, one, one, one, one, None, None, None, None, n, (C.c_int * 1)(4),
                                     (C.c_longlong * 1)(99), None, d, 0.9, 0.999, 1e-8, None) == -1
    assert lib.g1_mlp_adam_multi_sum(0, one, one, one, one, None, None, None, None, n, None, None, None, d, 0.9,
                                     0.999, 1e-8, None) == -1
    # sample-major gather: the row must hold the features and targets
    assert lib.g1_mlp_gather_rows(d, d, 8, 106, 35, 10, d, d, d, d, d, 64, d, 128, None) == -1
    assert lib.g1_mlp_gather_rows(d, d, 8, 106, 35, 10, d, d, d, d, d, 128, None, 128, None) == -1
    # stepper steps: negative counts, no state
    ws = _lib.ptr_array([16] * 5)
    assert lib.g1_hip_surrogate_steps(8, -1, 0, 2e-3, 16, 1e-3, d, d, d, d, None, d, d, ws, ws, d, d, d, d, d, d,
                                      None, None, None, None, 0, None) == -1
    assert lib.g1_hip_surrogate_steps(8, 2, 0, 2e-3, 16, 1e-3, None, d, d, d, None, d, d, ws, ws, d, d, d, d, d,
                                      d, None, None, None, None, 0, None) == -1
    assert "g1_hip_surrogate_steps" in _lib.last_error()
    # zero steps is a no-op
    assert lib.g1_hip_surrogate_steps(8, 0, 0, 2e-3, 16, 1e-3, d, d, d, d, None, d, d, ws, ws, d, d, d, d, d, d,
                                      None, None, None, None, 0, None) == 0


def test_deferred_judging_entry_points_validate_arguments():
    """Host-side checks of the deferred-judging entry points (before any
    launch): fallback flags, records, statistics and the batch's prediction
    buffer are required."""
    lib = _lib.load()
    d = C.c_void_p(16)
    ws = _lib.ptr_array([16] * 5)
    # judge + forces with commit: no fallback flags, no records, no statistics
    assert lib.g1_hip_judge_forces_commit(8, 2e-3, d, d, d, d, None, None, 0, 16, d, d, None, None, None, None, d, d,
                                          d, 0, None) == -1
    assert lib.g1_hip_judge_forces_commit(8, 2e-3, d, d, d, d, None, d, 0, 16, d, d, None, None, None, None, None, d,
                                          d, 0, None) == -1
    assert lib.g1_hip_judge_forces_commit(8, 2e-3, d, d, d, d, None, d, 0, 16, d, d, None, None, None, None, d, d,
                                          None, 0, None) == -1
    assert "judge_forces_commit" in _lib.last_error()
    # phase outside the period
    assert lib.g1_hip_judge_forces_commit(8, 2e-3, d, d, d, d, None, d, 16, 16, d, d, None, None, None, None, d, d,
                                          d, 0, None) == -1
    # step: no fallback flags or statistics (the committed blocks are found from the flags)
    assert lib.g1_hip_surrogate_step_deferred(8, 2e-3, 0, 16, 1e-3, d, d, d, d, d, ws, ws, d, d, d, None, d, None,
                                              None) == -1
    assert lib.g1_hip_surrogate_step_deferred(8, 2e-3, 0, 16, 1e-3, d, d, d, d, d, ws, ws, d, d, d, d, None, None,
                                              None) == -1
    assert "g1_hip_surrogate_step_deferred" in _lib.last_error()
    # batch: no records, no fallback flags
    assert lib.g1_hip_surrogate_judge_batch(8, 1e-3, d, d, ws, ws, d, d, None, d, d, d, None) == -1
    assert lib.g1_hip_surrogate_judge_batch(8, 1e-3, d, d, ws, ws, d, d, d, d, None, d, None) == -1
    assert "g1_hip_surrogate_judge_batch" in _lib.last_error()
    # steps: missing records; zero steps is a no-op
    assert lib.g1_hip_surrogate_steps_deferred(8, 2, 0, 2e-3, 16, 1e-3, d, d, d, d, None, d, d, ws, ws, d, d, d, d,
                                               d, d, None, None, None, None, 0, None, d, None, None, None) == -1
    flag = C.c_int(1)
    # the all-fallback flag needs the run flags
    assert lib.g1_hip_surrogate_steps_deferred(8, 2, 0, 2e-3, 16, 1e-3, d, d, d, d, None, d, d, ws, ws, d, d, d, d,
                                               d, d, None, None, None, None, 0, d, d, None, C.byref(flag),
                                               None) == -1
    assert lib.g1_hip_surrogate_steps_deferred(8, 0, 0, 2e-3, 16, 1e-3, d, d, d, d, None, d, d, ws, ws, d, d, d, d,
                                               d, d, None, None, None, None, 0, d, d, d, None, None) == 0
    # schedule: no stride, or a stride below one target block
    assert lib.g1_hip_surrogate_steps_schedule(8, 2, 0, 2e-3, 16, 1e-3, d, d, d, d, 0, 10, 0, None, d, d, ws, ws, d,
                                               d, d, d, d, d, None, None, None, None, 0, d, d, d, None, None) == -1
    assert lib.g1_hip_surrogate_steps_schedule(8, 2, 0, 2e-3, 16, 1e-3, d, d, d, d, 100, 10, 0, None, d, d, ws, ws,
                                               d, d, d, d, d, d, None, None, None, None, 0, d, d, d, None,
                                               None) == -1
    assert "g1_hip_surrogate_steps_schedule" in _lib.last_error()
    # fallback run: a run past the period's end, no run flags, no records; zero steps is a no-op
    assert lib.g1_hip_fallback_run(8, 5, 2e-3, d, d, d, d, None, d, 12, 16, d, None, None, None, None, d, d, d, 0,
                                   None) == -1
    assert "fallback_run" in _lib.last_error()
    assert lib.g1_hip_fallback_run(8, 4, 2e-3, d, d, d, d, None, d, 12, 16, None, None, None, None, None, d, d, d, 0,
                                   None) == -1
    assert lib.g1_hip_fallback_run(8, 4, 2e-3, d, d, d, d, None, d, 12, 16, d, None, None, None, None, None, d, d, 0,
                                   None) == -1
    assert lib.g1_hip_fallback_run(8, 0, 2e-3, d, d, d, d, None, d, 12, 16, d, None, None, None, None, d, d, d, 0,
                                   None) == 0
    # both wave splits at once
    both = _lib.G1_FLAG_WAVES4 | _lib.G1_FLAG_WAVES2
    assert lib.g1_hip_judge_forces_commit(8, 2e-3, d, d, d, d, None, d, 0, 16, d, d, None, None, None, None, d, d,
                                          d, both, None) == -1
    assert "flags" in _lib.last_error()
    assert lib.g1_hip_fallback_run(8, 4, 2e-3, d, d, d, d, None, d, 12, 16, d, None, None, None, None, d, d, d, both,
                                   None) == -1
    assert "flags" in _lib.last_error()
    # zero envs is a no-op
    assert lib.g1_hip_surrogate_judge_batch(0, 1e-3, d, d, ws, ws, d, d, d, d, d, d, None) == 0
    assert lib.g1_hip_judge_forces_commit(0, 2e-3, d, d, d, d, None, d, 0, 16, d, d, None, None, None, None, d, d, d,
                                          0, None) == 0
