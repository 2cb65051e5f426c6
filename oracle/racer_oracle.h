/*
 * TEST INFRASTRUCTURE ONLY. The CPU oracle: a plain-C restatement of the
 * reference's simulation and PPO (racer/sim.mojo, racer/ppo.mojo), used by the
 * tests, by smoke() to compare against, and by bench.py's cpu_baseline leg. No
 * product code links or calls it.
 *
 * Layout follows the reference: every buffer is feature-major float32
 * [rows, cols]; a "view" is a base pointer whose rows are `cols` floats apart.
 *
 * What pins each part (tests/test_oracle_*.py), without running the reference
 * (Mojo is not available here):
 *  - Map.at/probe/surface/progress: the reference test's corridor track and its
 *    answers (tests/test_mojo.mojo:27-92), the bridge invariants (:100-107),
 *    brute force on small grids;
 *  - derivative/advance: the closed forms of tests/test_mojo.mojo:48-55, a
 *    float64 restatement of the formula, RK4 against a fine float64 integrator,
 *    the corridor and timeout bookkeeping;
 *  - scan: the corridor ranges (7.1, 0.73), the closed-form ray-circle range on
 *    the flat ring, a brute-force fine march that never passes a wall;
 *  - GAE, the PPO gradient, KL and Adam: the host recurrence, central finite
 *    differences of a float64 loss (tests/test_mojo.mojo:137-214), the textbook
 *    Adam step, the init layout and its absolute-index quirk;
 *  - Philox4x32-10: the Random123 known-answer vectors; uniform and Box-Muller
 *    normals: their range and moment statistics;
 *  - sin, cos, exp, log, tanh: ULP bounds against float64 libm and special
 *    values; the nearest-cell division shortcut: exhaustive checks; the
 *    lidar's quotient from a reciprocal: sampled checks over its range,
 *    rounding midpoints included, on top of its proof (Markstein's theorem).
 * Parity unpinned (no fixture or closed form fixes the bits, and the Mojo
 * standard library that defines them is absent): the mapping of Mojo's
 * Random(seed, subsequence, offset) onto the Philox counter and key and its
 * uint-to-float conversion (this oracle uses counter = (offset, subsequence),
 * key = seed); the association order of SIMD.reduce_add (this oracle uses
 * (v0 + v2) + (v1 + v3)); the last-bit behaviour of Mojo's sin, cos, exp,
 * log and tanh; and the accumulation order of MAX's matmul and sum (this
 * oracle sums products in k order with fmaf, weight gradients in 4096-wide
 * chunks of the batch added in order, and reductions in float64).
 */
#ifndef RACER_ORACLE_H
#define RACER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny;
    float x0, y0, cell;
    int waypoints;
    float spacing;
    int layers;
    float slope;
    long route;
} omap;

typedef struct {
    int layer;
    float clearance;
    float height;
} ospot;

typedef struct {
    float height, sx, sy;
} osurface;

/* Map (sim.mojo:81-152) */
int om_init(omap* m, const float* d, size_t len);
float om_at(const omap* m, const float* d, int layer, int which, float x, float y);
ospot om_probe(const omap* m, const float* d, float x, float y, float z);
osurface om_surface(const omap* m, const float* d, int layer, float x, float y);
float om_progress(const omap* m, const float* d, int layer, float x, float y);
void o_slopes(osurface s, float heading, float* along, float* across);

/* Dynamics (sim.mojo:155-167) */
void o_derivative(const float p[4], float steer, float accel, float grip, float along, float across, float out[4]);

/* Per-car and per-beam functions (sim.mojo:170-266) */
void o_spawn(float* s, long n, const omap* m, const float* d, long i, int seed, float heading_noise);
void o_advance(float* s, long n, const float* act, long act_cols, float* reward, float* done, float* obs, long obs_cols,
               int beams, const omap* m, const float* d, long i, int seed, float heading_noise);
float o_scan(const omap* m, const float* d, const float* s, long n, const float* obs, long obs_cols, long i, int beam,
             int beams, long* iterations);

/* Whole-sim steps (sim.mojo:291-348). Returns the number of lidar march
 * iterations taken (summed over every beam), for the roofline byte count. */
void o_sim_reset(float* s, float* act, float* reward, float* done, float* obs, long n, int beams, const float* d,
                 size_t len, int seed, float heading_noise);
long o_sim_step(float* s, long n, const float* act, long act_cols, float* reward, float* done, float* obs,
                long obs_cols, int beams, const float* d, size_t len, int seed, float heading_noise);
void o_sim_sample(long n, const float* out, long out_cols, const float* log_std, float* act, long act_cols,
                  float* logp, float* value, int deterministic, int64_t step, int seed);

/* Policy (ppo.mojo:63-138). Weights theta = W0[H,IN] | W1[H,HA] | W2[3,HA] | logstd[2]. */
long o_param_count(int beams, int hidden);
void o_policy_init(float* theta, int beams, int hidden, int seed);
void o_forward(const float* theta, int beams, int hidden, const float* x, long x_cols, long cols, float* h1, float* h2,
               float* o);
void o_backward(const float* theta, int beams, int hidden, const float* x, long cols, const float* h1, const float* h2,
                const float* d3, float* d2, float* d1, float* grad);

/* Trainer pieces (ppo.mojo:187-337). */
void o_gae(long n, int rollout, const float* reward, const float* done, const float* value, const float* last_value,
           float* advantage);
void o_moments(long samples, const float* advantage, double* sum, double* sum_sq);
void o_minibatch(const float* theta, int beams, int hidden, long n, int rollout, int minibatches, const float* obs,
                 const float* actions, const float* logp, const float* value, const float* advantage, int m,
                 float adv_mean, float adv_scale, float* x, float* h1, float* h2, float* o, float* d3, float* d2,
                 float* d1, float* grad, float* norm_sq);
void o_adam(float* theta, long P, long logstd_at, const float* grad, float* first, float* second, float norm_sq,
            int step);
float o_bias_correction(float beta, int step);

/* Learned dynamics surrogate: inputs, targets, the float32 MLP and the
 * float64 RK4 step it is measured against (see the .c file). */
void o_surrogate_features(const float* d, size_t len, const float* s, long n, const float* act, long act_cols,
                          float* feat, long ldf);
void o_surrogate_targets(const float* prev, const float* s, const float* done, long n, float* tgt, long ldt,
                         float* mask);
void o_mlp_forward(const float* theta, const int* widths, int layers, const float* x, long cols, float* acts);
double o_mlp_gradient(const float* theta, const int* widths, int layers, const float* x, const float* y,
                      const float* mask, long cols, float* grad);
void o_step_f64(const float* d, size_t len, long n, double* st, const float* friction, const float* act,
                long act_cols);

/* Flat wrappers of the Map lookups and of one lidar beam, for the tests. */
int o_probe(const float* d, size_t len, float x, float y, float z, float out[3]);
int o_surface(const float* d, size_t len, int layer, float x, float y, float out[3]);
float o_progress(const float* d, size_t len, int layer, float x, float y);
float o_scan_one(const float* d, size_t len, const float* s, long n, const float* obs, long obs_cols, long i, int beam,
                 int beams, long* iterations);
/* spawn (sim.mojo:170-181) of car i of the state s [11, n], in place. */
int o_spawn_one(const float* d, size_t len, float* s, long n, long i, int seed, float heading_noise);

/* Exhaustive / sampled checks of the shared header's division shortcuts
 * (each returns the number of mismatches). */
long o_check_nearest_index(float b, float lo, float hi);
long o_check_nearest_index_lim(float b, float lo, float hi);
long o_check_div_by_reciprocal(long count, uint64_t seed);
long o_check_nearest_by_reciprocal(float cell, long stride);
long o_check_nearest_by_reciprocal1(float cell, long stride);
long o_nearest_form_bad(float cell, int nlo, int nhi, int form);
long o_check_tanh_select(long stride);

/* Threads used by the parallel loops (1 = serial). */
void o_set_threads(int threads);
int o_get_threads(void);

/* Elementary functions and RNG, exported for their own tests. */
float o_sin(float x);
float o_cos(float x);
float o_exp(float x);
float o_log(float x);
float o_tanh(float x);
float o_tanh_mlp(float x); /* the bf16 surrogate's tanh with libm exp2f and a division */
void o_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void o_normal4(uint64_t seed, uint64_t subsequence, uint64_t offset, float out[4]);
void o_uniform4(uint64_t seed, uint64_t subsequence, uint64_t offset, float out[4]);

#ifdef __cplusplus
}
#endif
#endif
