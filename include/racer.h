/*
 * libracer_hip: the C ABI of the MI355X (gfx950) racer hot path.
 *
 * It replaces the reference's Mojo `Device`, `Sim`, `Policy`, `Activations`
 * and `Trainer` (racer/device.mojo, racer/sim.mojo, racer/ppo.mojo) with
 * opaque handles over device memory. Plain pointers and sizes only.
 *
 * Conventions
 *   - Every buffer is float32 device memory, feature-major [rows, cols] as in
 *     the reference (racer/sim.mojo:5-7).
 *   - A racer_view is the reference's `mat[rows](buffer, cols, offset)`
 *     (racer/device.mojo:29-32): `rows` x `cols` floats, row-major, starting
 *     `offset` floats into `buf`. Views are passed by pointer.
 *   - Functions return 0 on success and nonzero on failure; the message is in
 *     racer_last_error() (per thread). The reference raises Error(msg).
 *   - Calls are asynchronous on the device's one in-order stream, except those
 *     documented as synchronous (alloc, read, write, upload, sync, the
 *     scalar-returning rollout/update), as in the reference.
 *   - A handle is not re-entrant: drive each from one host thread.
 */
#ifndef RACER_H
#define RACER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct racer_device racer_device;
typedef struct racer_buf racer_buf;
typedef struct racer_sim racer_sim;
typedef struct racer_policy racer_policy;
typedef struct racer_acts racer_acts;
typedef struct racer_trainer racer_trainer;

typedef struct {
    racer_buf* buf;
    size_t offset; /* floats into buf */
    int64_t rows;
    int64_t cols;
} racer_view;

/* Shape parameters that are compile-time constants in the reference
 * (racer/sim.mojo:28 BEAMS; racer/ppo.mojo:17-26 H, ROLLOUT, EPOCHS, MINIBATCHES). */
typedef struct {
    int beams;          /* lidar beams (reference default 108) */
    int hidden;         /* H, hidden width of both tanh layers (64) */
    int rollout;        /* ROLLOUT steps per iteration (32) */
    int epochs;         /* EPOCHS (4) */
    int minibatches;    /* MINIBATCHES (4) */
    int precision;      /* policy products: 0 = float32, 1 = bf16 MFMA with float32 accumulation,
                           2 = float32 in the reference's decomposition (three products, then a
                           separate sample kernel), kept to time the plain baseline */
    float heading_noise;/* spawn heading noise U(-a, a) in radians; 0 = reference behaviour */
} racer_config;

/* The reference defaults: 108 beams, H 64, ROLLOUT 32, 4 epochs, 4 minibatches, float32. */
void racer_config_default(racer_config* cfg);

/* Version string of the library. */
const char* racer_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* racer_last_error(void);

/* ---------------- Device: racer/device.mojo:35-134 ---------------- */
/* Device(gpu) (device.mojo:39-46). gpu = 0 fails: this library is the GPU path only. */
int racer_device_create(int gpu, racer_device** out);
void racer_device_destroy(racer_device* dev);
/* DeviceContext.name() (main.mojo:176). */
const char* racer_device_name(racer_device* dev);
/* ctx.synchronize() */
int racer_sync(racer_device* dev);
/* alloc(count, fill) (device.mojo:48-54): filled and waited for. */
int racer_alloc(racer_device* dev, size_t count, float fill, racer_buf** out);
void racer_free(racer_buf* buf);
size_t racer_buf_len(const racer_buf* buf);
/* Device address of the buffer's first float (for zero-copy interop). */
void* racer_buf_ptr(racer_buf* buf);
/* read(buffer) / write(buffer, values) / upload(values) (device.mojo:121-134): synchronous. */
int racer_read(racer_buf* buf, float* host);
int racer_write(racer_buf* buf, const float* host);
int racer_read_range(racer_buf* buf, size_t offset, size_t count, float* host);
int racer_write_range(racer_buf* buf, size_t offset, size_t count, const float* host);
int racer_upload(racer_device* dev, const float* host, size_t count, racer_buf** out);
/* ctx.enqueue_copy between sub-buffers (ppo.mojo:190, 210): asynchronous. */
int racer_copy(racer_buf* dst, size_t dst_offset, racer_buf* src, size_t src_offset, size_t count);

/* ---------------- Sim: racer/sim.mojo:269-348 ---------------- */
/* Sim(device, track, n, seed) (sim.mojo:280-289): copies the track array;
 * fails with "not a track array" when len < 12 (sim.mojo:97-98). */
int racer_sim_create(racer_device* dev, const float* track, size_t len, int n, int seed, const racer_config* cfg,
                     racer_sim** out);
void racer_sim_destroy(racer_sim* sim);
/* Sim.reset (sim.mojo:291-305): spawn every car, zero actions, one step. */
int racer_sim_reset(racer_sim* sim);
/* Sim.step (sim.mojo:307-321): physics, rewards, respawns, then every lidar
 * beam. actions [2, n], reward [1, n], done [1, n], obs [beams+5, n]. */
int racer_sim_step(racer_sim* sim, const racer_view* actions, const racer_view* reward, const racer_view* done,
                   const racer_view* obs);
/* Sim.sample (sim.mojo:323-348): out [3, n], log_std [1, 2]. */
int racer_sim_sample(racer_sim* sim, const racer_view* out, const racer_view* log_std, const racer_view* actions,
                     const racer_view* logp, const racer_view* value, int deterministic, int64_t step);
/* `steps` exact physics steps of every car (advance, sim.mojo:184-235) in
 * one launch without the lidar: open-loop steps for planning, step k taking
 * actions rows 2k, 2k+1 of the [2 steps, n] view; the state, reward, done and
 * obs rows 0-3 are those of `steps` Sim.step calls, the obs range rows keep
 * their values (no counterpart in the reference). */
int racer_sim_imagine(racer_sim* sim, const racer_view* actions, int steps);
/* The lidar pass of Sim.step alone (the second launch of sim.mojo:314-321):
 * every beam of every car from the current state and the obs slope rows
 * (2, 3) into the sim's obs range rows. */
int racer_sim_scan(racer_sim* sim);
racer_buf* racer_sim_state(racer_sim* sim);   /* [11, n] */
racer_buf* racer_sim_obs(racer_sim* sim);     /* [beams+5, n] */
racer_buf* racer_sim_actions(racer_sim* sim); /* [2, n] */
racer_buf* racer_sim_reward(racer_sim* sim);  /* [n] */
racer_buf* racer_sim_done(racer_sim* sim);    /* [n] */
int racer_sim_cars(const racer_sim* sim);
int racer_sim_beams(const racer_sim* sim);
/* The viewer's snapshot of every buffer (main.mojo:117-122): state [11, n],
 * obs [beams+5, n], actions [2, n], reward [n], done [n] copied back to back
 * into host (count >= (11 + beams + 5 + 4) n floats); synchronous. */
int racer_sim_snapshot(racer_sim* sim, float* host, size_t count);

/* ---------------- Policy and Activations: racer/ppo.mojo:47-138 ---------------- */
/* Activations(device, cols) (ppo.mojo:56-60): h1, h2 [H+1, cols] filled with 1, o [3, cols]. */
int racer_acts_create(racer_device* dev, int cols, const racer_config* cfg, racer_acts** out);
void racer_acts_destroy(racer_acts* acts);
racer_buf* racer_acts_h1(racer_acts* acts);
racer_buf* racer_acts_h2(racer_acts* acts);
racer_buf* racer_acts_o(racer_acts* acts);
/* Policy(device, seed) (ppo.mojo:68-90): Gaussian fan-in init, then transpose. */
int racer_policy_create(racer_device* dev, int seed, const racer_config* cfg, racer_policy** out);
void racer_policy_destroy(racer_policy* pol);
racer_buf* racer_policy_theta(racer_policy* pol); /* [P] = W0[H,IN] | W1[H,H+1] | W2[3,H+1] | logstd[2] */
size_t racer_policy_params(const racer_policy* pol);
/* Policy.log_std() (ppo.mojo:105-106): writes the [1, 2] view of theta. */
int racer_policy_log_std(racer_policy* pol, racer_view* out);
/* Policy.transpose (ppo.mojo:92-103) */
int racer_policy_transpose(racer_policy* pol);
/* Policy.forward (ppo.mojo:108-111): x [beams+5, acts.cols]. */
int racer_policy_forward(racer_policy* pol, const racer_view* x, racer_acts* acts);
/* Policy.backward (ppo.mojo:113-123): d3 [3, cols] -> d2, d1 [H+1, cols], grad W0/W1/W2 entries. */
int racer_policy_backward(racer_policy* pol, const racer_view* x, racer_acts* acts, const racer_view* d3,
                          racer_buf* d2, racer_buf* d1, racer_buf* grad);

/* racer_policy_backward without allocating (no counterpart in the reference):
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). The weight gradients'
 * split-K partial products go to a fixed scratch array of 16 MiB (4194304
 * floats) that each GPU the library runs on holds until the process ends,
 * instead of the device's grow-only scratch; the kernels and their order are
 * those of racer_policy_backward, so the results are bitwise the same. Shapes
 * whose partials exceed the array fail with an error naming the need (about
 * 61 K columns at 1081 beams and H 256, in float32 and in bf16).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other host threads, other device handles or streams) race. */
int racer_policy_backward_noalloc(racer_policy* pol, const racer_view* x, racer_acts* acts, const racer_view* d3,
                                  racer_buf* d2, racer_buf* d1, racer_buf* grad);

/* ---------------- Trainer: racer/ppo.mojo:141-337 ---------------- */
int racer_trainer_create(racer_device* dev, int n, const racer_config* cfg, racer_trainer** out);
void racer_trainer_destroy(racer_trainer* tr);
/* Trainer.rollout (ppo.mojo:187-232): returns the mean reward per step (synchronous). */
int racer_trainer_rollout(racer_trainer* tr, racer_sim* sim, racer_policy* pol, racer_acts* acts, int iteration,
                          float* mean_reward);
/* Trainer.update (ppo.mojo:234-257): returns the last minibatch's KL estimate (synchronous). */
int racer_trainer_update(racer_trainer* tr, racer_policy* pol, float* kl);
/* Trainer.minibatch (ppo.mojo:259-315): gradient of minibatch m into grad, its squared norm into stats. */
int racer_trainer_minibatch(racer_trainer* tr, racer_policy* pol, int m, float adv_mean, float adv_scale);
/* Trainer.adam (ppo.mojo:317-337): uses and advances the trainer's step count. */
int racer_trainer_adam(racer_trainer* tr, racer_policy* pol);
int racer_trainer_step_count(const racer_trainer* tr);
/* The update's copy of every minibatch's inputs (no reference counterpart;
 * the reference gathers each minibatch, ppo.mojo:276-279) is kept in bf16
 * when the bf16 policy's kernels read them so (bf16 precision, the 64-column
 * tiled forward, minibatches of at least 65536 aligned columns): half the
 * bytes, the same results bit for bit. On by default where it applies; off
 * keeps float32 copies. racer_trainer_bf16_inputs says which is in use. */
int racer_trainer_set_bf16_inputs(racer_trainer* tr, int on);
int racer_trainer_bf16_inputs(const racer_trainer* tr);
racer_buf* racer_trainer_obs(racer_trainer* tr);       /* [(ROLLOUT+1), beams+5, n] */
racer_buf* racer_trainer_actions(racer_trainer* tr);   /* [ROLLOUT, 2, n] */
racer_buf* racer_trainer_logp(racer_trainer* tr);      /* [ROLLOUT * n] */
racer_buf* racer_trainer_value(racer_trainer* tr);     /* [ROLLOUT * n] */
racer_buf* racer_trainer_reward(racer_trainer* tr);    /* [ROLLOUT * n] */
racer_buf* racer_trainer_done(racer_trainer* tr);      /* [ROLLOUT * n] */
racer_buf* racer_trainer_advantage(racer_trainer* tr); /* [ROLLOUT * n] */
racer_buf* racer_trainer_grad(racer_trainer* tr);      /* [P] */
racer_buf* racer_trainer_d3(racer_trainer* tr);        /* [6, batch] */
racer_buf* racer_trainer_stats(racer_trainer* tr);     /* [5] */
racer_acts* racer_trainer_acts(racer_trainer* tr);     /* minibatch activations */

/* ---------------- Checkpoints: racer/ppo.mojo:341-355 ---------------- */
/* save: one float32 .npy of [iteration, theta...]. */
int racer_save(racer_policy* pol, const char* path, int iteration);
/* load: fails with "checkpoint does not match the network size" unless 1+P values. */
int racer_load(racer_policy* pol, const char* path, int* iteration);

/* ---------------- Runner: main.mojo:165-235 ---------------- */
/* Runner.step (main.mojo:183-191): policy forward on sim.obs, sample into
 * sim.actions (logp/value into scratch [2, n]), then Sim.step. */
int racer_runner_step(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch, int deterministic,
                      int64_t t);
/* Runner.benchmark (main.mojo:225-235): 10 untimed steps, then `steps` timed
 * steps; writes milliseconds per step. */
int racer_benchmark(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch, int steps,
                    double* ms_per_step);

/* Launch-shape autotuning (no counterpart in the reference, whose launch
 * shapes are fixed in racer/device.mojo:56-98). Every choice gives bitwise the
 * same results; only the time differs.
 *
 * The forward kernel of racer_policy_act (the Runner step's and the
 * rollout's forward with the sample): RACER_ACT_AUTO picks by size;
 * RACER_ACT_FUSED is the one-kernel forward of 16 columns per workgroup (bf16,
 * or float32 on K = 1 MFMA); RACER_ACT_WIDE its bf16 variant for many
 * workgroups per CU; RACER_ACT_TILED the 64-column tiled bf16 forward (or the
 * float32 products) followed by the sample kernel. A choice that does not
 * apply to a launch falls back to RACER_ACT_AUTO for it. */
#define RACER_ACT_AUTO 0
#define RACER_ACT_FUSED 1
#define RACER_ACT_WIDE 2
#define RACER_ACT_TILED 3
int racer_policy_set_act_kernel(racer_policy* pol, int kernel);
int racer_policy_act_kernel(const racer_policy* pol);
/* Times `steps` Runner steps replayed from a captured graph for every
 * combination of lidar wavefronts per car (1, 2, 4), cooperative or
 * per-wavefront range stores or the fused sim step (where it applies) and
 * forward kernel (those that apply to
 * the policy's precision), from the same saved sim state each time (best of
 * three replays); keeps the fastest in sim and pol and restores the sim's
 * state, obs, actions, reward and done. Writes the choice (any pointer may be
 * null). Not while capturing; the profiler must be off. */
int racer_runner_autotune(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch, int steps,
                          int* waves_per_car, int* fused_step, int* act_kernel);

/* ---------------- Batches (no reference counterpart) ---------------- */
/* The reference steps one Sim per launch (racer/sim.mojo:314-321). A batch
 * groups 1 .. 16 existing sims (any tracks and car counts, one beam count) and
 * steps all of them with one launch per stage; every sim keeps its own
 * buffers and seed, so its results are bitwise those of stepping it alone.
 * The sims must outlive the batch. `cfg` gives the hidden width of the
 * activations the batch keeps for the shared policy. */
typedef struct racer_batch racer_batch;
int racer_batch_create(racer_device* dev, racer_sim** sims, int count, const racer_config* cfg, racer_batch** out);
void racer_batch_destroy(racer_batch* b);
int racer_batch_count(const racer_batch* b);
int64_t racer_batch_cars(const racer_batch* b);
/* Sim.step (racer/sim.mojo:314-321) of every sim, from its own actions buffer. */
int racer_batch_step(racer_batch* b);
/* Runner.step (main.mojo:183-191) of every sim with one shared policy: forward,
 * sample (each sim's seed, Philox offset t), physics, lidar. Log probabilities
 * and values land in the batch scratch: sim k's 2 n_k floats (log
 * probabilities, then values) start at 2 (n_0 + .. + n_(k-1)). */
int racer_batch_runner_step(racer_batch* b, racer_policy* pol, int deterministic, int64_t t);
racer_buf* racer_batch_scratch(racer_batch* b);
/* The batched lidar's wavefronts per car (1, 2, 4), its autotuner (replays
 * of captured graphs of batched sim steps from the same saved states, as
 * racer_sim_autotune), and its march-iteration counter (as the Sim ones). */
int racer_batch_set_waves_per_car(racer_batch* b, int waves);
int racer_batch_waves_per_car(const racer_batch* b);
int racer_batch_autotune(racer_batch* b, int* chosen);
/* The batched step's physics and lidar in one launch (the fused step of
 * racer_sim_set_fused_step over workgroups of 8 cars of one sim; it applies
 * when every sim's track has one or two layers and the lidar runs 1 or 2
 * wavefronts per car). mode 0: two launches; 1: the fused step, one
 * workgroup per 8 cars; 2: the fused step on a persistent grid of the
 * workgroups the GPU holds at once, each running a fixed, XCD-grouped run of
 * the 8-car blocks (no queue). Results are bitwise the same in every mode.
 * racer_batch_autotune times every applicable mode with every wavefront
 * count; racer_batch_fused_step reports the mode that runs (0 when the fused
 * step does not apply). */
int racer_batch_set_fused_step(racer_batch* b, int mode);
int racer_batch_fused_step(const racer_batch* b);
/* racer_runner_autotune for a batch: times `steps` batched Runner steps with
 * `pol` (captured as racer_graph_batch graphs, replayed from the sims' saved
 * states) for every wavefront count, step mode and, when the instances run
 * racer_policy_act, forward kernel; keeps the fastest in b and pol, restores
 * the sims' state, obs, actions, reward and done, and writes the choice (any
 * pointer may be null). Not while capturing; the profiler must be off. */
int racer_batch_runner_autotune(racer_batch* b, racer_policy* pol, int steps, int* waves_per_car, int* fused_step);
int racer_batch_count_iterations(racer_batch* b, int on);
int64_t racer_batch_last_iterations(const racer_batch* b);

/* ---------------- Learned dynamics surrogate (no reference counterpart) ----------------
 * A residual MLP (11 inputs -> 3 x 128 tanh -> 5) that predicts one step's
 * pose change of the RK4 integration (racer/sim.mojo:184-219): inputs x, y, z,
 * yaw, speed, yaw rate, steer, pitch and roll slopes, throttle, steer command;
 * outputs the displacement in the car's frame (forward, left) and the heading,
 * speed and steer changes. Trained on transitions the sim records, by the
 * library's own products (precision 1: bf16 MFMA hidden layers, float32
 * accumulation; precision 0: float32 throughout). `cols` is the largest batch
 * (and car count) the working buffers hold. */
typedef struct racer_surrogate racer_surrogate;
typedef struct racer_dataset racer_dataset;
int racer_surrogate_create(racer_device* dev, int seed, int precision, int64_t cols, racer_surrogate** out);
void racer_surrogate_destroy(racer_surrogate* s);
int64_t racer_surrogate_params(const racer_surrogate* s);
/* Weights: layer l's [w_(l+1), w_l + 1] (bias last) concatenated, l = 0 .. 3.
 * Call racer_surrogate_mark_updated after writing them. */
racer_buf* racer_surrogate_theta(racer_surrogate* s);
int racer_surrogate_mark_updated(racer_surrogate* s);
racer_buf* racer_surrogate_grad(racer_surrogate* s);
/* Normalisation: in_mean[11], in_inv_std[11], out_mean[5], out_std[5], out_inv_std[5]. */
racer_buf* racer_surrogate_stats(racer_surrogate* s);
/* Loss and gradient (into the grad buffer) of one batch: x [12, B] normalised
 * inputs with a last row of ones, y [5, B] normalised targets, mask [1, B];
 * loss = sum_b mask_b sum_o (out - y)^2 / (5 B). Outputs land in
 * racer_surrogate_outputs ([5, B] in its first 5 B floats). */
int racer_surrogate_batch_gradient(racer_surrogate* s, const racer_view* x, const racer_view* y,
                                   const racer_view* mask);
racer_buf* racer_surrogate_outputs(racer_surrogate* s);
/* Transitions: inputs [12, capacity], targets [5, capacity], mask [capacity]
 * (0 for steps that ended in a crash or timeout). */
int racer_dataset_create(racer_device* dev, int64_t capacity, racer_dataset** out);
void racer_dataset_destroy(racer_dataset* d);
int64_t racer_dataset_count(const racer_dataset* d);
racer_buf* racer_dataset_features(racer_dataset* d);
racer_buf* racer_dataset_targets(racer_dataset* d);
racer_buf* racer_dataset_mask(racer_dataset* d);
/* `steps` Runner steps (main.mojo:183-191) with transitions recorded (Philox offsets t0 ..). */
int racer_dataset_record(racer_dataset* d, racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch,
                         int steps, int64_t t0);
int racer_surrogate_fit_normalization(racer_surrogate* s, racer_dataset* d);
/* Adam (lr) over `epochs` passes of interleaved minibatches of `batch`. */
int racer_surrogate_train(racer_surrogate* s, racer_dataset* d, int epochs, int64_t batch, float lr);
/* Mean squared normalised error per usable transition of the dataset. */
int racer_surrogate_loss(racer_surrogate* s, racer_dataset* d, double* loss);
/* Sim.step (racer/sim.mojo:314-321) with the surrogate's pose change in place
 * of the RK4 integration; cars whose prediction lands below min_clearance from
 * the walls or on another layer take the exact step (fallback[i] = 1). The
 * crash test, progress, reward, respawn and lidar are the exact ones. */
int racer_surrogate_step(racer_surrogate* s, racer_sim* sim, float min_clearance, racer_buf* fallback);
/* bf16 surrogates: imagined steps as one fused launch (1, the default) or as
 * separate input, product and step kernels (0); the results are the same. */
int racer_surrogate_set_fused(racer_surrogate* s, int fused);
/* One step of the kinematics (racer/sim.mojo:184-219) in float64 on the GPU
 * (the surrogate's reference): st holds [6, n] doubles (x, y, z, heading, speed,
 * steer) in 12 n floats and is updated in place; friction comes from the sim's
 * state; no crash, progress or respawn. */
int racer_sim_step_f64(racer_sim* sim, racer_buf* st, const racer_view* actions);
/* Open-loop imagined steps (model-based planning): `steps` surrogate steps
 * of every car in one launch, step k taking actions rows 2k, 2k+1 of the
 * [2 steps, n] view, the network's weights loaded once; the lidar is not run
 * (obs range rows keep their values; rows 0-3 and the state, reward, done
 * are those of the last step, RETURN accumulates). fallback[i] = 1 when car i
 * took the exact step on any of them. bf16 fused surrogates only. Bitwise
 * the state of `steps` racer_surrogate_step calls with those actions. */
int racer_surrogate_imagine(racer_surrogate* s, racer_sim* sim, const racer_view* actions, int steps,
                            float min_clearance, racer_buf* fallback);
/* Runner.step (main.mojo:183-191) with racer_surrogate_step. */
int racer_surrogate_runner_step(racer_surrogate* s, racer_sim* sim, racer_policy* pol, racer_acts* acts,
                                racer_buf* scratch, int deterministic, int64_t t, float min_clearance,
                                racer_buf* fallback);

/* ---------------- Graphs (no reference counterpart) ---------------- */
/* The reference enqueues every launch of a step from the host
 * (main.mojo:183-191, racer/ppo.mojo:190-210). These capture the same launch
 * sequences once as a HIP graph and replay them; results are bitwise those of
 * the enqueued calls. Sampling kernels inside a graph take their Philox
 * offset as (device step base + the step's position), the base being written
 * before each replay. A graph holds the buffer addresses of the handles it was
 * captured from: those handles must outlive it. The profiler must be off while
 * capturing, and the iteration counter of the lidar must be off. */
typedef struct racer_graph racer_graph;
/* `steps` Runner steps (forward, sample, physics, lidar); replay i of
 * racer_graph_launch(g, t0) runs the steps t0 .. t0 + steps - 1. */
int racer_graph_runner(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch, int steps,
                       int deterministic, racer_graph** out);
int racer_graph_launch(racer_graph* g, int64_t t0);
/* `steps` batched Runner steps (racer_batch_runner_step) as one graph. */
int racer_graph_batch(racer_batch* b, racer_policy* pol, int steps, int deterministic, racer_graph** out);
/* Trainer.rollout (racer/ppo.mojo:187-232) as one graph: the copy in, ROLLOUT
 * steps, the copy out, the bootstrap forward, GAE and the reward sum. */
int racer_graph_rollout(racer_trainer* tr, racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_graph** out);
/* Replays a rollout graph for `iteration` and writes the mean reward per step
 * (the value racer_trainer_rollout returns). */
int racer_trainer_rollout_graph(racer_trainer* tr, racer_graph* g, int iteration, float* mean_reward);
void racer_graph_destroy(racer_graph* g);

/* ---------------- Instrumentation (no reference counterpart) ---------------- */
/* Kernel timing with HIP events on the device's stream: when enabled, every
 * launch of a named kernel is bracketed by events; totals are read back here. */
int racer_profile_enable(racer_device* dev, int on);
/* Number of kernels with samples, and the i-th: name, launch count, total ms. */
int racer_profile_count(racer_device* dev);
int racer_profile_get(racer_device* dev, int i, const char** name, int64_t* launches, double* total_ms);
int racer_profile_reset(racer_device* dev);
/* Restrict timing to the kernel of this name (NULL or "" times every kernel),
 * so a timed run carries one event pair per step instead of one per launch. */
int racer_profile_only(racer_device* dev, const char* name);
/* Lidar march iterations summed over every beam of the last step (counted by
 * an instrumented variant of the lidar kernel when enabled; 0 otherwise). */
int racer_sim_count_iterations(racer_sim* sim, int on);
int64_t racer_sim_last_iterations(racer_sim* sim);
/* Wavefront march-loop trips of the same counted launch (schedule 0): lane
 * utilisation = iterations / (64 trips). */
int64_t racer_sim_last_trips(racer_sim* sim);
/* The lidar's padded probe takes the floor and conversion of its nearest-cell
 * index as one instruction (v_cvt_flr_i32_f32); this counts the floats x with
 * |x| < 2^23 (every `stride`-th bit pattern, both signs) where that
 * instruction and (int)floorf(x) differ, on the device (0 expected). No
 * reference counterpart. */
int racer_check_floor_index(racer_device* dev, uint32_t stride, uint64_t* mismatches);
/* Lidar launch schedule: 0 (default) = one wavefront per car with beam refill;
 * 1 = one thread per (car, beam), beams fastest, the launch shape of the
 * reference's 2-D Device.run (racer/device.mojo:67-79), for timing the plain
 * baseline. Both give identical results. */
int racer_sim_set_schedule(racer_sim* sim, int schedule);
int racer_sim_schedule(const racer_sim* sim);
/* Wavefronts that share one car's beams in the default schedule (1, 2 or 4;
 * results are identical for every choice). */
int racer_sim_set_waves_per_car(racer_sim* sim, int waves);
int racer_sim_waves_per_car(const racer_sim* sim);
/* Range stores of the default schedule: 1 (default) = cooperative, each
 * workgroup of 8 wavefronts (8 / waves-per-car cars) writes its cars' ranges
 * together as runs of consecutive cars; 0 = each wavefront writes its own.
 * No reference counterpart (a launch-shape choice under racer/sim.mojo:317-321);
 * results are identical either way. */
int racer_sim_set_lidar_coop(racer_sim* sim, int on);
int racer_sim_lidar_coop(const racer_sim* sim);
/* One launch for the whole step (physics, then the lidar, in the same
 * workgroups, with each car's neighbourhood of the track staged in LDS) in
 * place of the two launches of racer/sim.mojo:314-321. It applies with the
 * default lidar schedule, one or two track layers and 1 or 2 wavefronts per
 * car; racer_sim_fused_step returns 1 when the next step will run fused.
 * Results are identical either way. */
int racer_sim_set_fused_step(racer_sim* sim, int on);
int racer_sim_fused_step(const racer_sim* sim);
/* Autotuner: times `racer_sim_step`s replayed from captured graphs with 1, 2
 * and 4 wavefronts per car, each with cooperative and per-wavefront range
 * stores and fused where the fused step applies, from the same saved state
 * each time; keeps the fastest (waves per car written to `chosen`, the other
 * choices read by racer_sim_fused_step and racer_sim_lidar_coop) and restores
 * the sim's state, obs, actions, reward and done. Not while capturing; the
 * profiler must be off. */
int racer_sim_autotune(racer_sim* sim, int* chosen);

#ifdef __cplusplus
}
#endif
#endif
