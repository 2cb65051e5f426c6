/*
 * Physical constants, buffer row layout and track-array layout of the racer
 * simulation and policy (racer/sim.mojo:14-56, racer/ppo.mojo:17-37,
 * racer/track.py:15-18). Shared by the HIP kernels and the CPU oracle; C and
 * HIP compatible. BEAMS, H, ROLLOUT, EPOCHS and MINIBATCHES are run-time
 * parameters here (compile-time constants in the reference).
 */
#ifndef RACER_CONTRACT_H
#define RACER_CONTRACT_H

/* Car (F1TENTH), sim.mojo:16-26 */
#define RACER_WHEELBASE 0.3302f
#define RACER_HALF_DIAG 0.3f
#define RACER_GRIP 1.05f
#define RACER_G 9.81f
#define RACER_STEER_MAX 0.4189f
#define RACER_STEER_RATE 3.2f
#define RACER_ACCEL_MAX 9.51f
#define RACER_SPEED_MAX 5.0f
#define RACER_SUBSTEPS 4
#define RACER_DT ((float)(1.0 / 240))
#define RACER_HORIZON 3000

/* Lidar, sim.mojo:27-31 (BEAMS is a run-time parameter; 108 by default) */
#define RACER_DEFAULT_BEAMS 108
#define RACER_RANGE 10.0f
#define RACER_FOV 4.71238898f
#define RACER_MOUNT_FORWARD 0.27f
#define RACER_MOUNT_UP 0.23f

/* State rows, sim.mojo:33-44 */
#define RACER_X 0
#define RACER_Y 1
#define RACER_Z 2
#define RACER_HEADING 3
#define RACER_SPEED 4
#define RACER_STEER 5
#define RACER_STEPS 6
#define RACER_PROGRESS 7
#define RACER_EPISODE 8
#define RACER_FRICTION 9
#define RACER_RETURN 10
#define RACER_STATE 11

/* Observation rows: steer, speed, slope along, slope across, BEAMS ranges, constant 1
 * (OBS = 4 + BEAMS is the row of the constant, IN = OBS + 1 rows). */
#define RACER_OUT 3
#define RACER_CRASH 1
#define RACER_TIMEOUT 2
#define RACER_HALF_LOG_2PI ((float)0.91893853320467274178)

/* Track array header and grids, sim.mojo:53-56, track.py:15-18 */
#define RACER_HEADER 12
#define RACER_CLEARANCE 0
#define RACER_HEIGHT 1
#define RACER_NEAREST 2

/* PPO, ppo.mojo:17-37 (H, ROLLOUT, EPOCHS, MINIBATCHES are run-time parameters) */
#define RACER_DEFAULT_H 64
#define RACER_DEFAULT_ROLLOUT 32
#define RACER_DEFAULT_EPOCHS 4
#define RACER_DEFAULT_MINIBATCHES 4
#define RACER_GAMMA 0.99f
#define RACER_LAMBDA 0.95f
#define RACER_CLIP 0.2f
#define RACER_LR 3e-4f
#define RACER_MAX_NORM 0.5f
#define RACER_STAT_REWARD 0
#define RACER_STAT_ADV 1
#define RACER_STAT_NORM 3
#define RACER_STAT_KL 4
#define RACER_STATS 5

#endif
