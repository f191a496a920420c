#!/usr/bin/env python3
"""Benchmark of the fused G1 step on one MI355X.

    python bench.py [--gpus 1] [--steps 250] [--warmup 50] [--nenv 16384]

Workload (synthetic, as the randomized-actuation configuration): 16,384 envs
of the 29-DoF G1, dt = 2e-3, initial base tilt U(+-0.05) rad, base velocity
U(+-0.5) m/s in x/y, joint qd ~ N(0, 0.3), PD targets stand + U(+-0.3) rad
resampled every 10 steps (all target blocks are generated beforehand and are
resident in HBM).  One launch of the step kernel advances all envs by the 10
steps that share a target block (nsub = 10); the timed region is `--steps`
physics steps after `--warmup` untimed ones.

Prints one JSON line: throughput in env-steps/s, the roofline of the step
kernel (algorithmic FP32 FLOP of the reference formulation, 41,491 per
env-step, over the kernel's average launch time from HIP events on the launch
stream), the same FP32 ABA step at 131,072 envs (`aba_131072`), the
surrogate and policy-in-the-loop legs, and the CPU baseline (the reference
physics core compiled from the reference sources, float, OpenMP over the host
cores) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "env-steps/s @16k envs: FP32 ABA vs surrogate; max qacc rel err vs FP64"
FLOP_PER_ENV_STEP = 41491          # reference formulation, stance (8 contacts)
BYTES_PER_ENV_LAUNCH = 812         # SoA loads (qpos, qvel, anchor, ctrl) + stores (qpos, qvel, anchor) of
                                   # one launch without the optional foot/contact outputs (nsub independent)
PEAK_FP32 = 157.3e12               # MI355X FP32 (vector and matrix), dense
PUBLISHED_16K = 12.8e6             # reference README.md:38-42, RTX 3080 Laptop, 16,384 envs
PEAK_BF16 = 2.5e15                 # MI355X dense BF16 MFMA
SURR_FLOP = 1717248                # 106 -> 512 x4 -> 35 MLP forward, per env-step
POLICY_FLOP = 433408               # 96 -> 512 -> 256 -> 128 -> 29 actor forward, per env-step
CTRL_PERIOD = 10


def accuracy_vs_fp64(sample):
    """qacc of one GPU step against the FP64 oracle from the same float32
    state (teacher forcing): per env the max over dofs of |a - b| / max(|b|, 1)
    (tests/test_vs_mujoco.py:74-75).  The reference algorithm evaluated in
    float32 on the CPU is scored the same way: the yardstick of what float32
    arithmetic reaches on these states."""
    from oracle import oracle as O
    before = sample["before"]
    ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], 2e-3,
                           want=("qacc",))["qacc"]
    n = ref.shape[1]
    q32 = np.zeros((35, n))
    args = [np.ascontiguousarray(before[k], np.float32) for k in ("qpos", "qvel", "anchor", "ctrl")]
    O.lib32().g1o_step_batch_f32(n, 2e-3, *(O.fptr(a) for a in args), None, O.dptr(q32), None, None, None, None,
                                 None, None, None, 0)

    def score(q):
        e = (np.abs(q - ref) / np.maximum(np.abs(ref), 1.0)).max(axis=0)
        return {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "median": float(np.median(e))}

    return {"metric": "max over dofs of |qacc - qacc_fp64| / max(|qacc_fp64|, 1), per env",
            "sample": f"{n} envs after the timed region, one step at dt 2e-3 (teacher forced)",
            "gpu_fp32": score(sample["qacc"]), "reference_algorithm_in_float32": score(q32)}


def cpu_baseline(nenv_gpu, acc_sample=None):
    """Reference physics core (src/g1_core.h, float, OpenMP one env per
    iteration) on the host cores, on a bounded sample of the workload; with
    acc_sample, also the FP64 oracle's score of the GPU's qacc on it.
    Returns (baseline, accuracy or None)."""
    from g1sim_amd import workloads as W
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    try:
        fn = O.ref_bench().g1ref_bench_f32
        kind = "reference"
    except (OSError, FileNotFoundError):
        fn = O.lib32().g1o_bench_f32
        kind = "port"
    res = {}
    for threads, nenv, steps in ((cores, 16384, 50), (1, 2048, 25)):
        qp, qv, an, cb = W.randomized_actuation(nenv, seed=1, nblocks=1)
        ct = np.ascontiguousarray(cb[0])
        t = fn(nenv, steps, 2e-3, O.fptr(qp), O.fptr(qv), O.fptr(an), O.fptr(ct), threads)
        res[threads] = (nenv * steps / t, nenv, steps)
    v_all, n_all, s_all = res[cores]
    v_one, n_one, s_one = res[1]
    base = {"value": v_all, "unit": "env-steps/s", "cores": cores, "kind": kind,
            "sample": f"{n_all} envs x {s_all} steps of the same workload (dt 2e-3, fixed targets) "
                      f"on {cores} threads; 1-thread rate from {n_one} envs x {s_one} steps",
            "value_1core": v_one}
    return base, (accuracy_vs_fp64(acc_sample) if acc_sample is not None else None)


def aba_leg(dev, nenv, steps, warmup):
    """The FP32 ABA step on the same randomized-actuation workload at another
    batch size (the 131,072-env target: the 2-wave split), timed as the main
    leg: wall clock between synchronisations for env-steps/s, one HIP event
    pair around the back-to-back launches for the kernel's average duration."""
    import torch
    from g1sim_amd import _lib, workloads as W
    lib = _lib.load()
    nblocks = (warmup + steps + CTRL_PERIOD - 1) // CTRL_PERIOD + 1
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=3, nblocks=nblocks)
    qpos, qvel, anchor = (torch.as_tensor(x, device=dev) for x in (qp, qv, an))
    ctrl = torch.as_tensor(cb, device=dev)
    stream = torch.cuda.current_stream(dev)

    def run(step0, n):
        for s in range(step0, step0 + n, CTRL_PERIOD):
            rc = lib.g1_hip_step_ex(nenv, CTRL_PERIOD, 2e-3, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                                    ctrl[s // CTRL_PERIOD].data_ptr(), None, None, None, None, None, None, None,
                                    0, stream.cuda_stream)
            if rc:
                raise RuntimeError(_lib.last_error())

    run(0, warmup)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    run(warmup, steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    avg_ms = e0.elapsed_time(e1) / (steps // CTRL_PERIOD)
    achieved = FLOP_PER_ENV_STEP * nenv * CTRL_PERIOD / (avg_ms * 1e-3)
    return {"nenv": nenv, "steps": steps, "warmup": warmup, "nsub_per_launch": CTRL_PERIOD,
            "env_steps_per_s": nenv * steps / wall, "avg_launch_ms": avg_ms,
            "kernel": "g1k::step_kernel<false, false, 2, 0>",
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / PEAK_FP32}}


def _event_avg_ms(fn, reps, stream):
    """Average duration of fn()'s launches from HIP events on `stream`."""
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        fn()
        e1.record(stream)
    torch.cuda.synchronize()
    return float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))


def surrogate_leg(dev, nenv, steps, warmup, train_iters):
    """The learned surrogate on the same workload: a 20M-transition dataset
    from the ABA kernel (16,384 envs x 1,221 steps), training by the
    project's MFMA kernels, held-out accuracy, then surrogate + rotating 1/16
    ABA judge + per-env fallback (tolerance 1e-3) stepping `steps` steps.
    Returns (result, trained network)."""
    import torch
    from g1sim_amd import workloads as W
    from g1sim_amd.surrogate import MLP, SurrogateStepper, evaluate, fit_normalization, generate_dataset, train
    t0 = time.perf_counter()
    ds = generate_dataset(dev, nenv=16384, steps=1221, seed=5)
    mlp = MLP(dev, seed=0)
    fit_normalization(mlp, ds)
    hist = train(mlp, ds, iters=train_iters, batch=8192, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    acc = evaluate(mlp, ds, nsamp=65536, tol=1e-3)
    del ds
    torch.cuda.empty_cache()
    nblocks = (warmup + steps + CTRL_PERIOD - 1) // CTRL_PERIOD + 1
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=1, nblocks=nblocks)
    ctrl = torch.as_tensor(cb, device=dev)

    def run(tol):
        qpos = torch.as_tensor(qp, device=dev)
        qvel = torch.as_tensor(qv, device=dev)
        anchor = torch.as_tensor(an, device=dev)
        st = SurrogateStepper(mlp, nenv, dev, period=16, tol=tol)
        for s in range(warmup):
            st.step(qpos, qvel, anchor, ctrl[s // CTRL_PERIOD])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(warmup, warmup + steps):
            st.step(qpos, qvel, anchor, ctrl[s // CTRL_PERIOD])
        torch.cuda.synchronize()
        return time.perf_counter() - t1, st.summary(), (qpos, qvel, st.tau)

    wall, summ, _ = run(1e-3)
    # the same pipeline with every judgement accepted: the cost of the
    # surrogate path itself, independent of the network's accuracy
    wall_acc, _, (qpos, qvel, tau) = run(1e30)
    # looser acceptance tolerances than the configuration's 1e-3, for context
    sweep = []
    for tol in (1e-1, 1.0):
        w_t, s_t, _ = run(tol)
        sweep.append({"tol": tol, "env_steps_per_s": nenv * steps / w_t, "aba_frac": s_t["aba_frac"],
                      "reject_frac": s_t["reject_frac"], "fallback_envs_at_end": s_t["fallback_envs_now"]})
    # the surrogate's own evaluation configuration: 65,536 envs x 1,000 steps,
    # 4,096 judged per step, state drift against plain ABA from the same start
    n4 = 65536
    qp4, qv4, an4, cb4 = W.randomized_actuation(n4, seed=4, nblocks=100)
    ctrl4 = torch.as_tensor(cb4, device=dev)
    sq, sv, sa = (torch.as_tensor(a, device=dev) for a in (qp4, qv4, an4))
    aq, av, aa = (torch.as_tensor(a, device=dev) for a in (qp4, qv4, an4))
    st4 = SurrogateStepper(mlp, n4, dev, period=16, tol=1e-3)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    for s in range(1000):
        st4.step(sq, sv, sa, ctrl4[s // CTRL_PERIOD])
    torch.cuda.synchronize()
    wall4 = time.perf_counter() - t4
    from g1sim_amd import _lib
    lib = _lib.load()
    for blk in range(100):
        rc = lib.g1_hip_step_ex(n4, CTRL_PERIOD, 2e-3, aq.data_ptr(), av.data_ptr(), aa.data_ptr(),
                                ctrl4[blk].data_ptr(), None, None, None, None, None, None, None, 0,
                                torch.cuda.current_stream(dev).cuda_stream)
        if rc:
            raise RuntimeError(_lib.last_error())
    torch.cuda.synchronize()
    s4 = st4.summary()
    dpos = (sq[0:3] - aq[0:3]).norm(dim=0).cpu().numpy()
    eval4 = {"nenv": n4, "steps": 1000, "judged_per_step": n4 // 16, "env_steps_per_s": n4 * 1000 / wall4,
             "aba_frac": s4["aba_frac"], "reject_frac": s4["reject_frac"],
             "fallback_envs_at_end": s4["fallback_envs_now"],
             "base_pos_drift_vs_aba_m": {"median": float(np.median(dpos)), "max": float(dpos.max())}}
    del sq, sv, sa, aq, av, aa, ctrl4, st4
    # the fused network kernel alone (features + 5 layers), HIP events
    pred = torch.empty((nenv, 64), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    fwd = lambda: mlp.forward_state(nenv, qpos, qvel, tau, pred)  # noqa: E731
    fwd()
    k_ms = _event_avg_ms(fwd, 50, stream)
    k_tflops = SURR_FLOP * nenv / (k_ms * 1e-3) / 1e12
    return {"value": nenv * steps / wall, "unit": "env-steps/s", "ms_per_step": wall * 1e3 / steps,
            "judge": "rotating 1/16 of envs per step, per-env fallback at rel. qacc err > 1e-3",
            "heldout_accept_frac": acc["accept_frac"], "heldout_median_rel_err": acc["median_rel_err"],
            "heldout_p90_rel_err": acc["p90_rel_err"], "heldout_accept_frac_at": acc["accept_frac_at"],
            "tolerance_sweep": sweep, "train_iters": train_iters,
            "train_final_mse_z": hist[-1][1], "dataset_and_train_s": t_train,
            "dataset": "20M ABA transitions: 16,384 envs x 1,221 steps",
            "value_all_accepted": nenv * steps / wall_acc, "ms_per_step_all_accepted": wall_acc * 1e3 / steps,
            "aba_frac": summ["aba_frac"], "reject_frac": summ["reject_frac"],
            "fallback_envs_at_end": summ["fallback_envs_now"], "evaluation_65536x1000": eval4,
            "roofline": {"bound": "mfma", "kernel": "g1inf::surrogate_kernel<0> (features + 5 layers)",
                         "achieved": k_tflops, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                         "frac": k_tflops * 1e12 / PEAK_BF16, "avg_launch_ms": k_ms,
                         "flop_per_launch": SURR_FLOP * nenv}}, mlp


def policy_leg(dev, mlp, nenv=65536, steps=24, iters=50, nsub=10):
    """Policy in the loop: a randomly initialised actor (96 -> 512 -> 256 ->
    128 -> 29, ELU, fused on MFMA) drives nenv envs for iters rollouts of
    `steps` policy steps, each nsub physics substeps at dt 2e-3, with
    rewards, terminations, resets, commands (U(+-1) m/s, U(+-1) rad/s) and
    pushes on the device; FP32 ABA physics, then the surrogate with its
    judge."""
    import torch
    from g1sim_amd.policy import PolicyRollout
    out = {"nenv": nenv, "steps": steps, "iterations": iters, "nsub": nsub, "dt": 2e-3,
           "note": "rollouts only: the PPO update between iterations is not part of the loop timed here"}
    stream = torch.cuda.current_stream(dev)
    for physics in ("aba", "surrogate"):
        ro = PolicyRollout(nenv, dev, physics=physics, surrogate=mlp if physics == "surrogate" else None,
                           seed=11, nsub=nsub, cmd_lin=1.0, cmd_yaw=1.0)
        ro.rollout(steps)                                  # warmup iteration
        rs = torch.zeros(steps, dtype=torch.float64, device=dev)
        dc = torch.zeros(steps, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            ro.rollout(steps, rs, dc)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        psteps = nenv * steps * iters
        r = {"env_steps_per_s": psteps * nsub / wall, "policy_steps_per_s": psteps / wall,
             "ms_per_policy_step": wall * 1e3 / (steps * iters),
             "mean_reward": float(rs.sum().item() / psteps), "done_frac": float(dc.sum().item() / psteps)}
        if ro.stepper is not None:
            sm = ro.stepper.summary()
            r.update({"aba_frac": sm["aba_frac"], "reject_frac": sm["reject_frac"],
                      "fallback_envs_at_end": sm["fallback_envs_now"]})
        out[physics] = r
        if physics == "aba":
            pol = ro.policy
            fn = lambda: pol.act(nenv, ro.qpos, ro.qvel, ro.commands, ro.ctrl, ro.ctrl,  # noqa: E731
                                 actions=ro.actions, seed=11, t=0)
            k_ms = _event_avg_ms(fn, 50, stream)
            tf = POLICY_FLOP * nenv / (k_ms * 1e-3) / 1e12
            out["policy_kernel"] = {"bound": "mfma", "kernel": "g1inf::policy_kernel", "avg_launch_ms": k_ms,
                                    "achieved": tf, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                                    "frac": tf * 1e12 / PEAK_BF16}
        del ro
        torch.cuda.empty_cache()
    return out


def kernel_only(kind, nenv, steps, warmup):
    """One fused inference kernel alone, or one surrogate stepper step
    (judge + forces, network + commit; random weights, every judgement
    accepted), on the randomized-actuation state, timed with HIP events: for
    profiling runs."""
    import torch
    from g1sim_amd import workloads as W
    dev = torch.device("cuda:0")
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=1, nblocks=1)
    qpos, qvel = torch.as_tensor(qp, device=dev), torch.as_tensor(qv, device=dev)
    if kind == "stepper":
        from g1sim_amd.surrogate import MLP, SurrogateStepper
        mlp = MLP(dev, seed=0)
        anchor, ctrl = torch.as_tensor(an, device=dev), torch.as_tensor(cb[0], device=dev)
        st = SurrogateStepper(mlp, nenv, dev, period=16, tol=1e30)
        fn = lambda: st.step(qpos, qvel, anchor, ctrl)  # noqa: E731
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ms = _event_avg_ms(fn, steps, torch.cuda.current_stream(dev))
        print(json.dumps({"kernel": "surrogate stepper step", "nenv": nenv, "steps": steps, "avg_step_ms": ms,
                          "env_steps_per_s": nenv / (ms * 1e-3)}), flush=True)
        return
    if kind == "surrogate":
        from g1sim_amd.surrogate import MLP
        mlp = MLP(dev, seed=0)
        tau = torch.randn((35, nenv), device=dev)
        pred = torch.empty((nenv, 64), device=dev)
        fn = lambda: mlp.forward_state(nenv, qpos, qvel, tau, pred)  # noqa: E731
        flop, name = SURR_FLOP, "g1inf::surrogate_kernel<0>"
    else:
        from g1sim_amd.policy import Policy
        pol = Policy(dev, layout=1, seed=0)
        cmd = torch.zeros((3, nenv), device=dev)
        ctrl = torch.zeros((29, nenv), device=dev)
        act = torch.zeros((29, nenv), device=dev)
        fn = lambda: pol.act(nenv, qpos, qvel, cmd, ctrl, ctrl, actions=act, seed=1, t=0)  # noqa: E731
        flop, name = POLICY_FLOP, "g1inf::policy_kernel"
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ms = _event_avg_ms(fn, steps, torch.cuda.current_stream(dev))
    tf = flop * nenv / (ms * 1e-3) / 1e12
    print(json.dumps({"kernel": name, "nenv": nenv, "launches": steps, "avg_launch_ms": ms,
                      "achieved_tflops": tf, "frac_bf16_peak": tf * 1e12 / PEAK_BF16}), flush=True)


def measure_peaks():
    """Measured machine peaks on this device: non-packed FP32 FMA, dense BF16
    MFMA (16x16x32, register operands), and HBM copy bandwidth."""
    import torch
    from g1sim_amd import _lib
    lib = _lib.load()
    dev = torch.device("cuda:0")
    props = torch.cuda.get_device_properties(dev)
    stream = torch.cuda.current_stream(dev)
    cus = props.multi_processor_count
    out = torch.empty(cus * 8 * 256, dtype=torch.float32, device=dev)
    res = {"device": props.name, "compute_units": cus}
    # best over 2 and 8 waves per SIMD (4-wave blocks, 2 or 8 per CU)
    for name, fn, iters, per_wave in (("fp32_fma_tflops", lib.g1_hip_peak_fp32, 20000, 64 * 8 * 2),
                                      ("bf16_mfma_tflops", lib.g1_hip_peak_bf16, 1000, 16 * 16384)):
        best = 0.0
        for blocks in (cus * 2, cus * 8):
            launch = lambda: _lib.check(fn(blocks, iters, out.data_ptr(), stream.cuda_stream), name)  # noqa: E731
            launch()
            ms = _event_avg_ms(launch, 10, stream)
            best = max(best, blocks * 4 * iters * per_wave / (ms * 1e-3) / 1e12)
        res[name] = best
    src = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev)     # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    ms = _event_avg_ms(lambda: dst.copy_(src), 10, stream)
    res["hbm_copy_tbps"] = 2 * src.numel() * 4 / (ms * 1e-3) / 1e12
    res["published"] = {"fp32_tflops": PEAK_FP32 / 1e12, "bf16_mfma_tflops": PEAK_BF16 / 1e12, "hbm_tbps": 8.0}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=250)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--nenv", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--flags", type=int, default=0,
                    help="g1_hip_step_ex flags: 1 float64 base solve, 2 float64 base pose, 4 float64 "
                         "foot-contact geometry")
    ap.add_argument("--nsub", type=int, default=CTRL_PERIOD,
                    help="substeps per launch (must divide the 10-step target period)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate leg")
    ap.add_argument("--train-iters", type=int, default=12000)
    ap.add_argument("--no-policy", action="store_true", help="skip the policy-in-the-loop leg")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 131,072-env FP32 ABA leg of the default run")
    ap.add_argument("--peaks", action="store_true", help="measure FP32 FMA, BF16 MFMA and HBM copy peaks and exit")
    ap.add_argument("--kernel", choices=("surrogate", "policy", "stepper"),
                    help="time one fused inference kernel, or one surrogate stepper step, alone at --nenv "
                         "(for profiling) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy a 256 MiB buffer once (reference kernel of known "
                         "traffic for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step shards envs with no exchange, so N GPUs run N "
              "independent copies of this benchmark; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    if args.kernel:
        kernel_only(args.kernel, args.nenv, args.steps, args.warmup)
        return
    if args.peaks:
        print(json.dumps(measure_peaks()), flush=True)
        return
    import torch
    from g1sim_amd import _lib, workloads as W

    lib = _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nenv = args.nenv
    if CTRL_PERIOD % args.nsub:
        raise SystemExit("--nsub must divide 10")
    base64 = None
    if args.flags & _lib.G1_FLAG_F64_BASE:
        base64 = torch.empty((7, nenv), dtype=torch.float64, device=dev)
    nblocks = (args.warmup + args.steps + CTRL_PERIOD - 1) // CTRL_PERIOD + 1
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=1, nblocks=nblocks)
    qpos = torch.as_tensor(qp, device=dev)
    qvel = torch.as_tensor(qv, device=dev)
    anchor = torch.as_tensor(an, device=dev)
    ctrl = torch.as_tensor(cb, device=dev)                 # all target blocks resident in HBM
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream

    def launch(step0, nsteps):
        """Launches covering physics steps [step0, step0 + nsteps)."""
        s = step0
        while s < step0 + nsteps:
            n = min(args.nsub - s % args.nsub, step0 + nsteps - s)
            rc = lib.g1_hip_step_ex(nenv, n, 2e-3, qpos.data_ptr(), qvel.data_ptr(), anchor.data_ptr(),
                                    ctrl[s // CTRL_PERIOD].data_ptr(), None, None, None, None, None, None,
                                    None if base64 is None else base64.data_ptr(), args.flags, sptr)
            if rc:
                raise RuntimeError(_lib.last_error())
            yield n
            s += n

    if base64 is not None:
        base64.copy_(qpos[:7])
    for _ in launch(0, args.warmup):
        pass
    torch.cuda.synchronize()

    # timed region: wall clock between synchronisations, and one HIP event
    # pair on the launch stream around its back-to-back launches for the
    # kernel's average duration (an event pair around every launch would put
    # ~3 us of event processing between the kernels: 1.18e9 instead of
    # 1.28e9 env-steps/s on the same GPU)
    gen = launch(args.warmup, args.steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    subs = list(gen)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    region_ms = e0.elapsed_time(e1)
    if args.copy_calibration:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    sub_arr = np.array(subs)
    # accuracy sample: one more step of the final state with the qacc output
    acc_sample = None
    if not args.no_cpu_baseline:
        blk = ctrl[(args.warmup + args.steps) // CTRL_PERIOD]
        before = {"qpos": qpos.cpu().numpy(), "qvel": qvel.cpu().numpy(), "anchor": anchor.cpu().numpy(),
                  "ctrl": blk.cpu().numpy()}
        qp2, qv2, an2 = qpos.clone(), qvel.clone(), anchor.clone()
        qacc = torch.zeros((35, nenv), dtype=torch.float32, device=dev)
        rc = lib.g1_hip_step_ex(nenv, 1, 2e-3, qp2.data_ptr(), qv2.data_ptr(), an2.data_ptr(), blk.data_ptr(),
                                None, None, None, qacc.data_ptr(), None, None, None,
                                args.flags & ~_lib.G1_FLAG_F64_BASE, sptr)
        if rc:
            raise RuntimeError(_lib.last_error())
        acc_sample = {"before": before, "qacc": qacc.cpu().numpy().astype(np.float64)}
    h = qpos[2].cpu().numpy()
    ok = bool(np.isfinite(qpos.cpu().numpy()).all() and ((h > 0.2) & (h < 1.2)).mean() > 0.99)

    env_steps = nenv * args.steps
    value = env_steps / wall
    avg_ms = region_ms / len(sub_arr)
    nsub_avg = float(sub_arr.mean())
    flop_per_launch = FLOP_PER_ENV_STEP * nenv * nsub_avg
    achieved = flop_per_launch / (avg_ms * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        if t.get("nenv") == nenv and t.get("nsub") == args.nsub and t.get("flags", 0) == args.flags:
            traffic = t.get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_16K,
        "vs_baseline_ref": "12.8e6 env-steps/s, reference fused CUDA step, 16,384 envs, RTX 3080 Laptop (README.md:38-42)",
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "randomized actuation: 16,384 envs, dt 2e-3, PD targets stand+U(+-0.3) "
                               "resampled every 10 steps, FP32 ABA step",
                   "nenv": nenv, "dt": 2e-3, "nsub_per_launch": args.nsub, "path": "FP32 ABA",
                   "flags": args.flags},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": traffic,
                     "kernel": "g1k::step_kernel<%s, %s, %d, 0>" % (
                         "true" if args.flags & 2 else "false", "true" if args.flags & 4 else "false",
                         4 if args.flags & 8 else 2 if (args.flags & 16 or nenv >= 32768) else 4),
                     "flop_per_launch": flop_per_launch, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_ENV_LAUNCH * nenv},
        "sanity_standing": ok,
    }
    peaks = measure_peaks()
    out["measured_peaks"] = peaks
    out["roofline"]["frac_of_measured_fp32_fma"] = achieved / (peaks["fp32_fma_tflops"] * 1e12)
    if not args.no_large and args.flags == 0 and nenv == 16384 and args.nsub == CTRL_PERIOD:
        out["aba_131072"] = aba_leg(dev, 131072, 200, 20)
    if not args.no_surrogate and args.flags == 0:
        out["surrogate"], mlp = surrogate_leg(dev, nenv, args.steps, args.warmup, args.train_iters)
        if not args.no_policy:
            out["policy_in_the_loop"] = policy_leg(dev, mlp)
    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["max_qacc_rel_err_vs_fp64"] = cpu_baseline(nenv, acc_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
