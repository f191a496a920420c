// g1_launch.h -- host-side launchers shared between the library's sources
// (not part of the C-ABI).
#pragma once
#include <stdint.h>

namespace g1k_host {
// g1_hip_judge_forces_commit with blk_run (nullable): blocks marked there by
// g1_hip_fallback_run were stepped by it and are left alone.
int judge_forces_commit(int nenv, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                        const float* mu, const uint8_t* fallback, int phase, int period, float* qacc_out,
                        float* tau_out, float* contact_forces, float* foot_pos, float* foot_vel, float* qacc_track,
                        float* snap, uint8_t* sflag, unsigned* stats, const uint8_t* blk_run, unsigned flags,
                        void* stream);
// g1_hip_fallback_run with a PD-target schedule (ctrl_stride > 0): substep s
// uses the block ctrl + ((ctrl_phase + s) / ctrl_period) ctrl_stride.
int fallback_run(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                 long long ctrl_stride, int ctrl_period, int ctrl_phase, const float* mu, const uint8_t* fallback,
                 int phase0, int period, uint8_t* blk_run, float* qacc_track, float* contact_forces, float* foot_pos,
                 float* foot_vel, float* snap, uint8_t* sflag, unsigned* stats, unsigned flags, void* stream);
}  // namespace g1k_host
