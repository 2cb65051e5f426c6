/*
 * TEST INFRASTRUCTURE ONLY: the CPU oracle. A plain-C restatement of the
 * reference's per-step hot path, one function per reference function, each
 * citing the file:line it follows. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, to compare against or as the CPU baseline.
 *
 * It shares exactly two headers with the product: the constant/row layout
 * (warporacer3d_amd/csrc/racer_contract.h, the reference's own constants) and
 * the float32 elementary functions + Philox (racer_math.h), which are pinned on
 * their own against float64 numpy and the Random123 known-answer vectors. All
 * simulation and training logic below is restated independently of the HIP
 * kernels.
 *
 * Arithmetic: float32, evaluated in source order with no contraction
 * (-ffp-contract=off). Matrix products accumulate each output with fmaf in
 * increasing k from 0. SIMD reduce_add over 4 lanes is taken as the pairwise
 * tree (v0 + v2) + (v1 + v3). Where the reference's float32 arithmetic order
 * is fixed by its source, it is followed operation by operation.
 */
#include "racer_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../warporacer3d_amd/csrc/racer_contract.h"
#include "../warporacer3d_amd/csrc/racer_math.h"

static int g_threads = 1;

void o_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }
int o_get_threads(void) { return g_threads; }

/* ---------------- Map: sim.mojo:81-152 ---------------- */

/* sim.mojo:96-102 */
int om_init(omap* m, const float* d, size_t len) {
    if (len < RACER_HEADER) return -1;
    m->nx = (int)d[0];
    m->ny = (int)d[1];
    m->x0 = d[2];
    m->y0 = d[3];
    m->cell = d[4];
    m->waypoints = (int)d[5];
    m->spacing = d[6];
    m->layers = (int)d[7];
    m->slope = d[8];
    m->route = RACER_HEADER + 3L * m->layers * m->nx * m->ny;
    return 0;
}

/* sim.mojo:105-106 */
static long om_grid(const omap* m, int layer, int which) {
    return RACER_HEADER + (long)(3 * layer + which) * m->nx * m->ny;
}

/* sim.mojo:108-114: nearest cell, -1 off the map. */
float om_at(const omap* m, const float* d, int layer, int which, float x, float y) {
    float fx = floorf((x - m->x0) / m->cell + 0.5f);
    float fy = floorf((y - m->y0) / m->cell + 0.5f);
    /* range test on the floats, so a NaN coordinate also reads as off the map */
    if (!(fx >= 0.0f) || !(fy >= 0.0f) || fx >= (float)m->nx || fy >= (float)m->ny) return -1.0f;
    return d[om_grid(m, layer, which) + (long)fy * m->nx + (long)fx];
}

/* sim.mojo:116-124: drivable layer whose height is nearest z; layer 0 with
 * clearance -1 when none is drivable. */
ospot om_probe(const omap* m, const float* d, float x, float y, float z) {
    ospot best;
    best.layer = 0;
    best.clearance = -1.0f;
    best.height = om_at(m, d, 0, RACER_HEIGHT, x, y);
    float gap = 1e30f;
    for (int l = 0; l < m->layers; ++l) {
        float c = om_at(m, d, l, RACER_CLEARANCE, x, y);
        float h = om_at(m, d, l, RACER_HEIGHT, x, y);
        if (c > 0.0f && fabsf(h - z) < gap) {
            best.layer = l;
            best.clearance = c;
            best.height = h;
            gap = fabsf(h - z);
        }
    }
    return best;
}

/* sim.mojo:126-140: bilinear height and its gradient. */
osurface om_surface(const omap* m, const float* d, int layer, float x, float y) {
    float fx = (x - m->x0) / m->cell;
    float fy = (y - m->y0) / m->cell;
    /* clamp(Int(floor(f)), 0, n - 2), clamped on the floats */
    float cx = floorf(fx), cy = floorf(fy);
    long ix = !(cx >= 0.0f) ? 0 : (cx > (float)(m->nx - 2) ? m->nx - 2 : (long)cx);
    long iy = !(cy >= 0.0f) ? 0 : (cy > (float)(m->ny - 2) ? m->ny - 2 : (long)cy);
    float tx = racer_clamp(fx - (float)ix, 0.0f, 1.0f);
    float ty = racer_clamp(fy - (float)iy, 0.0f, 1.0f);
    long k = om_grid(m, layer, RACER_HEIGHT) + iy * m->nx + ix;
    float h0 = d[k], h1 = d[k + 1], h2 = d[k + m->nx], h3 = d[k + m->nx + 1];
    float w0 = (1.0f - tx) * (1.0f - ty);
    float w1 = tx * (1.0f - ty);
    float w2 = (1.0f - tx) * ty;
    float w3 = tx * ty;
    osurface s;
    s.height = (h0 * w0 + h2 * w2) + (h1 * w1 + h3 * w3);
    s.sx = ((h1 - h0) * (1.0f - ty) + (h3 - h2) * ty) / m->cell;
    s.sy = ((h2 - h0) * (1.0f - tx) + (h3 - h1) * tx) / m->cell;
    return s;
}

/* sim.mojo:65-69: rise per metre ahead and to the left. */
void o_slopes(osurface s, float heading, float* along, float* across) {
    float c = racer_cos(heading);
    float sn = racer_sin(heading);
    *along = s.sx * c + s.sy * sn;
    *across = -s.sx * sn + s.sy * c;
}

/* sim.mojo:143-144 */
static float om_waypoint(const omap* m, const float* d, long w, int field) { return d[m->route + 4 * w + field]; }

/* sim.mojo:146-152 */
float om_progress(const omap* m, const float* d, int layer, float x, float y) {
    long w = (long)om_at(m, d, layer, RACER_NEAREST, x, y);
    float heading = om_waypoint(m, d, w, 3);
    float along = (x - om_waypoint(m, d, w, 0)) * racer_cos(heading) + (y - om_waypoint(m, d, w, 1)) * racer_sin(heading);
    return (float)w + along / m->spacing;
}

/* ---------------- Dynamics: sim.mojo:155-167 ---------------- */

void o_derivative(const float p[4], float steer, float accel, float grip, float along, float across, float out[4]) {
    float nz = 1.0f / sqrtf((1.0f + along * along) + across * across);
    float load = (grip * RACER_G) * nz;
    float v = racer_max(fabsf(p[3]), 0.5f) * (p[3] >= 0.0f ? 1.0f : -1.0f);
    float slide = (RACER_G * across) * nz;
    float lateral = racer_clamp((((v * p[3]) * racer_sin(steer)) / racer_cos(steer)) / RACER_WHEELBASE + slide, -load, load);
    float yaw = (lateral - slide) / v;
    float longitudinal = sqrtf(racer_max(load * load - lateral * lateral, 0.0f));
    float dv = racer_clamp(accel, -longitudinal, longitudinal) - (RACER_G * along) * nz;
    out[0] = p[3] * racer_cos(p[2]);
    out[1] = p[3] * racer_sin(p[2]);
    out[2] = yaw;
    out[3] = dv;
}

/* ---------------- spawn: sim.mojo:170-181 ----------------
 * heading_noise is an extension (0 reproduces the reference): the heading is
 * offset by heading_noise * (2 u2 - 1) from the same Philox block. */
void o_spawn(float* s, long n, const omap* m, const float* d, long i, int seed, float heading_noise) {
    float episode = s[RACER_EPISODE * n + i] + 1.0f;
    racer_u32x4 r = racer_random((uint64_t)(int64_t)seed, (uint64_t)i, (uint64_t)episode);
    float u0 = racer_uniform(r.v[0]);
    float u1 = racer_uniform(r.v[1]);
    long w = (long)(u0 * (float)m->waypoints);
    if (w > m->waypoints - 1) w = m->waypoints - 1;
    for (int row = 0; row < RACER_STATE; ++row) s[row * n + i] = 0.0f;
    s[RACER_X * n + i] = om_waypoint(m, d, w, 0);
    s[RACER_Y * n + i] = om_waypoint(m, d, w, 1);
    s[RACER_Z * n + i] = om_waypoint(m, d, w, 2);
    float heading = om_waypoint(m, d, w, 3);
    if (heading_noise != 0.0f) heading = heading + heading_noise * (2.0f * racer_uniform(r.v[2]) - 1.0f);
    s[RACER_HEADING * n + i] = heading;
    s[RACER_PROGRESS * n + i] = (float)w;
    s[RACER_EPISODE * n + i] = episode;
    s[RACER_FRICTION * n + i] = 0.85f + 0.3f * u1;
}

/* ---------------- advance: sim.mojo:184-235 ---------------- */

static void vadd_scaled(const float* p, const float* k, float f, float* out) {
    for (int j = 0; j < 4; ++j) out[j] = p[j] + k[j] * f;
}

void o_advance(float* s, long n, const float* act, long act_cols, float* reward, float* done, float* obs, long obs_cols,
               int beams, const omap* m, const float* d, long i, int seed, float heading_noise) {
    const float DT = RACER_DT;
    float p[4] = {s[RACER_X * n + i], s[RACER_Y * n + i], s[RACER_HEADING * n + i], s[RACER_SPEED * n + i]};
    ospot spot = om_probe(m, d, p[0], p[1], s[RACER_Z * n + i]);
    osurface ground = om_surface(m, d, spot.layer, p[0], p[1]);
    float steer = s[RACER_STEER * n + i];
    float rate = racer_clamp(act[i], -1.0f, 1.0f) * RACER_STEER_RATE;
    float accel = racer_clamp(act[act_cols + i], -1.0f, 1.0f) * RACER_ACCEL_MAX;
    float grip = RACER_GRIP * s[RACER_FRICTION * n + i];
    int crashed = 0;
    for (int sub = 0; sub < RACER_SUBSTEPS; ++sub) {
        float along, across;
        o_slopes(ground, p[2], &along, &across);
        float k1[4], k2[4], k3[4], k4[4], q[4];
        float half_steer = steer + (rate * DT) / 2.0f;
        o_derivative(p, steer, accel, grip, along, across, k1);
        vadd_scaled(p, k1, DT / 2.0f, q);
        o_derivative(q, half_steer, accel, grip, along, across, k2);
        vadd_scaled(p, k2, DT / 2.0f, q);
        o_derivative(q, half_steer, accel, grip, along, across, k3);
        vadd_scaled(p, k3, DT, q);
        o_derivative(q, steer + rate * DT, accel, grip, along, across, k4);
        for (int j = 0; j < 4; ++j) p[j] = p[j] + ((k1[j] + 2.0f * (k2[j] + k3[j])) + k4[j]) * (DT / 6.0f);
        p[3] = racer_clamp(p[3], -RACER_SPEED_MAX, RACER_SPEED_MAX);
        steer = racer_clamp(steer + rate * DT, -RACER_STEER_MAX, RACER_STEER_MAX);
        spot = om_probe(m, d, p[0], p[1], spot.height);
        ground = om_surface(m, d, spot.layer, p[0], p[1]);
        crashed = crashed || spot.clearance < RACER_HALF_DIAG || ground.sx * ground.sx + ground.sy * ground.sy > 1.0f;
    }
    if (p[2] > RACER_PI) p[2] -= RACER_2PI;
    if (p[2] < -RACER_PI) p[2] += RACER_2PI;
    float progress = om_progress(m, d, spot.layer, p[0], p[1]);
    float dp = progress - s[RACER_PROGRESS * n + i];
    float half = (float)m->waypoints / 2.0f;
    dp = dp > half ? dp - 2.0f * half : (dp < -half ? dp + 2.0f * half : dp);
    float score = crashed ? -2.0f : (10.0f * dp) * m->spacing - (0.001f * steer) * steer;
    float steps = s[RACER_STEPS * n + i] + 1.0f;
    float reason = crashed ? (float)RACER_CRASH : (steps >= (float)RACER_HORIZON ? (float)RACER_TIMEOUT : 0.0f);
    s[RACER_X * n + i] = p[0];
    s[RACER_Y * n + i] = p[1];
    s[RACER_Z * n + i] = spot.height;
    s[RACER_HEADING * n + i] = p[2];
    s[RACER_SPEED * n + i] = p[3];
    s[RACER_STEER * n + i] = steer;
    s[RACER_STEPS * n + i] = steps;
    s[RACER_PROGRESS * n + i] = progress;
    s[RACER_RETURN * n + i] += score;
    reward[i] = score;
    done[i] = reason;
    if (reason > 0.0f) {
        o_spawn(s, n, m, d, i, seed, heading_noise);
        float x = s[RACER_X * n + i], y = s[RACER_Y * n + i];
        ground = om_surface(m, d, om_probe(m, d, x, y, s[RACER_Z * n + i]).layer, x, y);
    }
    float along, across;
    o_slopes(ground, s[RACER_HEADING * n + i], &along, &across);
    obs[0 * obs_cols + i] = s[RACER_STEER * n + i] / RACER_STEER_MAX;
    obs[1 * obs_cols + i] = s[RACER_SPEED * n + i] / RACER_SPEED_MAX;
    obs[2 * obs_cols + i] = along;
    obs[3 * obs_cols + i] = across;
    obs[(long)(4 + beams) * obs_cols + i] = 1.0f;
}

/* ---------------- scan: sim.mojo:238-266 ---------------- */

float o_scan(const omap* m, const float* d, const float* s, long n, const float* obs, long obs_cols, long i, int beam,
             int beams, long* iterations) {
    float heading = s[RACER_HEADING * n + i];
    float ch = racer_cos(heading), sh = racer_sin(heading);
    float fwd[3] = {ch, sh, obs[2 * obs_cols + i]};
    float left[3] = {-sh, ch, obs[3 * obs_cols + i]};
    float azimuth = -RACER_FOV / 2.0f + (float)beam * (RACER_FOV / (float)(beams - 1));
    float ca = racer_cos(azimuth), sa = racer_sin(azimuth);
    float dir[3];
    for (int j = 0; j < 3; ++j) dir[j] = ca * fwd[j] + sa * left[j];
    /* reduce_add of (d0^2, d1^2, d2^2, 0) as the tree (d0^2 + d2^2) + (d1^2 + 0) */
    float norm = sqrtf((dir[0] * dir[0] + dir[2] * dir[2]) + (dir[1] * dir[1] + 0.0f));
    for (int j = 0; j < 3; ++j) dir[j] = dir[j] / norm;
    float origin[3];
    origin[0] = (s[RACER_X * n + i] + RACER_MOUNT_FORWARD * fwd[0]) + 0.0f;
    origin[1] = (s[RACER_Y * n + i] + RACER_MOUNT_FORWARD * fwd[1]) + 0.0f;
    origin[2] = (s[RACER_Z * n + i] + RACER_MOUNT_FORWARD * fwd[2]) + RACER_MOUNT_UP;
    float horizontal = sqrtf(dir[0] * dir[0] + dir[1] * dir[1]);
    float descent = (1.5f * m->slope) * horizontal + racer_max(-dir[2], 0.0f);
    float t = 0.0f, last = 0.0f, last_floor = 0.0f;
    long it = 0;
    while (t < RACER_RANGE) {
        ++it;
        float px = origin[0] + t * dir[0];
        float py = origin[1] + t * dir[1];
        float pz = origin[2] + t * dir[2];
        ospot spot = om_probe(m, d, px, py, pz);
        float floor_ = (pz - spot.height) - 0.03f;
        if (floor_ < 0.0f) {
            if (iterations) *iterations += it;
            return racer_min(last + ((t - last) * last_floor) / (last_floor - floor_), RACER_RANGE);
        }
        float step = racer_min(spot.clearance - m->cell, floor_ / descent);
        if (step < 0.5f * m->cell) break;
        last = t;
        last_floor = floor_;
        t += step;
    }
    if (iterations) *iterations += it;
    return racer_min(t, RACER_RANGE);
}

/* ---------------- Sim: sim.mojo:291-348 ---------------- */

long o_sim_step(float* s, long n, const float* act, long act_cols, float* reward, float* done, float* obs,
                long obs_cols, int beams, const float* d, size_t len, int seed, float heading_noise) {
    omap m;
    if (om_init(&m, d, len)) return -1;
    long total = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < n; ++i)
        o_advance(s, n, act, act_cols, reward, done, obs, obs_cols, beams, &m, d, i, seed, heading_noise);
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) reduction(+ : total)
    for (long i = 0; i < n; ++i) {
        long it = 0;
        for (int b = 0; b < beams; ++b)
            obs[(4 + b) * obs_cols + i] = o_scan(&m, d, s, n, obs, obs_cols, i, b, beams, &it) / RACER_RANGE;
        total += it;
    }
    return total;
}

/* sim.mojo:291-305: spawn every car with zero actions, then one step. */
void o_sim_reset(float* s, float* act, float* reward, float* done, float* obs, long n, int beams, const float* d,
                 size_t len, int seed, float heading_noise) {
    omap m;
    if (om_init(&m, d, len)) return;
    for (long i = 0; i < n; ++i) {
        o_spawn(s, n, &m, d, i, seed, heading_noise);
        act[i] = 0.0f;
        act[n + i] = 0.0f;
    }
    o_sim_step(s, n, act, n, reward, done, obs, n, beams, d, len, seed, heading_noise);
}

/* sim.mojo:323-348 */
void o_sim_sample(long n, const float* out, long out_cols, const float* log_std, float* act, long act_cols,
                  float* logp, float* value, int deterministic, int64_t step, int seed) {
    for (long i = 0; i < n; ++i) {
        float noise[4];
        racer_normal4(racer_random((uint64_t)(int64_t)seed + 1, (uint64_t)i, (uint64_t)step), noise);
        float probability = 0.0f;
        for (int a = 0; a < 2; ++a) {
            float z = deterministic ? 0.0f : noise[a];
            act[a * act_cols + i] = out[a * out_cols + i] + racer_exp(log_std[a]) * z;
            probability = probability - (((0.5f * z) * z + log_std[a]) + RACER_HALF_LOG_2PI);
        }
        logp[i] = probability;
        value[i] = out[2 * out_cols + i];
    }
}

/* ---------------- Policy: ppo.mojo:63-138 ---------------- */

long o_param_count(int beams, int hidden) {
    long IN = beams + 5, HA = hidden + 1;
    return (long)hidden * IN + (long)hidden * HA + 3L * HA + 2;
}

/* ppo.mojo:74-89, including the reference's use of the absolute index i in the
 * bias-column tests of W1 and W2. */
void o_policy_init(float* theta, int beams, int hidden, int seed) {
    long IN = beams + 5, OBS = beams + 4, HA = hidden + 1, H = hidden;
    long W1 = H * IN, W2 = W1 + H * HA, LOGSTD = W2 + 3 * HA, P = LOGSTD + 2;
    for (long i = 0; i < P; ++i) {
        float scale = 0.0f;
        if (i < W1) {
            scale = (i % IN < OBS) ? sqrtf(2.0f / (float)OBS) : 0.0f;
        } else if (i < W2) {
            scale = (i % HA < H) ? sqrtf(1.0f / (float)H) : 0.0f;
        } else if (i < LOGSTD) {
            scale = (i % HA < H) ? (((i - W2) / HA < 2) ? 0.00125f : 0.125f) : 0.0f;
        } else {
            theta[i] = -0.5f;
            continue;
        }
        float z[4];
        racer_normal4(racer_random((uint64_t)(int64_t)seed, (uint64_t)i, 0), z);
        theta[i] = scale * z[0];
    }
}

/* c[j, col] = sum_k a[j, k] b[k, col] accumulated with fmaf in increasing k.
 * Columns go 16 at a time so the b reads are contiguous; each output is still
 * its own fmaf chain over k in order. */
#define MM_COLS 16
static void mm(const float* a, long lda, const float* b, long ldb, float* c, long ldc, long rows, long inner, long cols,
               int squash) {
    long blocks = (cols + MM_COLS - 1) / MM_COLS;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long blk = 0; blk < blocks; ++blk) {
        long c0 = blk * MM_COLS, n = cols - c0 < MM_COLS ? cols - c0 : MM_COLS;
        for (long j = 0; j < rows; ++j) {
            float acc[MM_COLS] = {0.0f};
            const float* arow = a + j * lda;
            for (long k = 0; k < inner; ++k) {
                const float w = arow[k];
                const float* brow = b + k * ldb + c0;
                for (long q = 0; q < n; ++q) acc[q] = fmaf(w, brow[q], acc[q]);
            }
            for (long q = 0; q < n; ++q) c[j * ldc + c0 + q] = squash ? racer_tanh(acc[q]) : acc[q];
        }
    }
}

/* ppo.mojo:108-111; h1 and h2 have H+1 rows, the last one the constant 1. */
void o_forward(const float* theta, int beams, int hidden, const float* x, long x_cols, long cols, float* h1, float* h2,
               float* o) {
    long IN = beams + 5, HA = hidden + 1, H = hidden;
    const float* W0 = theta;
    const float* W1 = theta + H * IN;
    const float* W2 = W1 + H * HA;
    mm(W0, IN, x, x_cols, h1, cols, H, IN, cols, 1);
    for (long c = 0; c < cols; ++c) h1[H * cols + c] = 1.0f;
    mm(W1, HA, h1, cols, h2, cols, H, HA, cols, 1);
    for (long c = 0; c < cols; ++c) h2[H * cols + c] = 1.0f;
    mm(W2, HA, h2, cols, o, cols, 3, HA, cols, 0);
}

/* sum_c a[c] b[c] over the batch in the order the weight gradients use (the
 * order of MAX's matmul is not known, see the header): fmaf in increasing c
 * within chunks of 4096, then the chunk sums added in order to 0. */
static float chunked_dot(const float* a, const float* b, long cols) {
    float total = 0.0f;
    for (long c0 = 0; c0 < cols || c0 == 0; c0 += 4096) {
        long c1 = c0 + 4096 < cols ? c0 + 4096 : cols;
        float acc = 0.0f;
        for (long c = c0; c < c1; ++c) acc = fmaf(a[c], b[c], acc);
        total += acc;
        if (c1 >= cols) break;
    }
    return total;
}

/* ppo.mojo:113-138. d3 holds [3, cols] output gradients; d2, d1 get [HA, cols]
 * (bias rows masked to zero); grad receives the W0, W1, W2 entries. */
void o_backward(const float* theta, int beams, int hidden, const float* x, long cols, const float* h1, const float* h2,
                const float* d3, float* d2, float* d1, float* grad) {
    long IN = beams + 5, HA = hidden + 1, H = hidden;
    const float* W1 = theta + H * IN;
    const float* W2 = W1 + H * HA;
    /* through: result[k, c] = (1 - h[k,c]^2) * sum_j wt[k, j] delta[j, c], wt = W^T */
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long c = 0; c < cols; ++c) {
        for (long k = 0; k < HA; ++k) {
            float acc = 0.0f;
            for (long j = 0; j < 3; ++j) acc = fmaf(W2[j * HA + k], d3[j * cols + c], acc);
            float h = h2[k * cols + c];
            d2[k * cols + c] = acc * (1.0f - h * h);
        }
        for (long k = 0; k < HA; ++k) {
            float acc = 0.0f;
            for (long j = 0; j < H; ++j) acc = fmaf(W1[j * HA + k], d2[j * cols + c], acc);
            float h = h1[k * cols + c];
            d1[k * cols + c] = acc * (1.0f - h * h);
        }
    }
    /* weight gradients: grad[j, k] = sum_c delta[j, c] act[k, c] */
    float* g0 = grad;
    float* g1 = grad + H * IN;
    float* g2 = g1 + H * HA;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long j = 0; j < H; ++j) {
        for (long k = 0; k < IN; ++k) g0[j * IN + k] = chunked_dot(d1 + j * cols, x + k * cols, cols);
        for (long k = 0; k < HA; ++k) g1[j * HA + k] = chunked_dot(d2 + j * cols, h1 + k * cols, cols);
    }
    for (long j = 0; j < 3; ++j)
        for (long k = 0; k < HA; ++k) g2[j * HA + k] = chunked_dot(d3 + j * cols, h2 + k * cols, cols);
}

/* ---------------- Trainer: ppo.mojo:187-337 ---------------- */

/* ppo.mojo:218-230. The running advantage is carried in float64 (the return
 * accumulation over long horizons); each stored value is rounded to float32. */
void o_gae(long n, int rollout, const float* reward, const float* done, const float* value, const float* last_value,
           float* advantage) {
    for (long i = 0; i < n; ++i) {
        double next = last_value[i];
        double running = 0.0;
        for (int t = rollout - 1; t >= 0; --t) {
            long k = (long)t * n + i;
            double alive = done[k] == 0.0f ? 1.0 : 0.0;
            double delta = (double)reward[k] + (double)RACER_GAMMA * alive * next - (double)value[k];
            running = delta + (double)RACER_GAMMA * (double)RACER_LAMBDA * alive * running;
            advantage[k] = (float)running;
            next = value[k];
        }
    }
}

/* ppo.mojo:242-250, accumulated in float64. */
void o_moments(long samples, const float* advantage, double* sum, double* sum_sq) {
    double a = 0.0, b = 0.0;
    for (long i = 0; i < samples; ++i) {
        a += advantage[i];
        b += (double)advantage[i] * (double)advantage[i];
    }
    *sum = a;
    *sum_sq = b;
}

/* ppo.mojo:259-315 */
void o_minibatch(const float* theta, int beams, int hidden, long n, int rollout, int minibatches, const float* obs,
                 const float* actions, const float* logp, const float* value, const float* advantage, int m,
                 float adv_mean, float adv_scale, float* x, float* h1, float* h2, float* o, float* d3, float* d2,
                 float* d1, float* grad, float* norm_sq) {
    long IN = beams + 5, H = hidden, HA = hidden + 1;
    long batch = n * rollout / minibatches;
    long LOGSTD = H * IN + H * HA + 3 * HA;
    long P = LOGSTD + 2;
    const float* log_std = theta + LOGSTD;
    for (long b = 0; b < batch; ++b) {
        long s = b * minibatches + m;
        for (long r = 0; r < IN; ++r) x[r * batch + b] = obs[(s / n * IN + r) * n + s % n];
    }
    o_forward(theta, beams, hidden, x, batch, batch, h1, h2, o);
    for (long b = 0; b < batch; ++b) {
        long s = b * minibatches + m;
        long t = s / n, i = s % n;
        float new_logp = 0.0f;
        float error[2], invvar[2];
        for (int a = 0; a < 2; ++a) {
            error[a] = actions[(t * 2 + a) * n + i] - o[a * batch + b];
            invvar[a] = racer_exp(-2.0f * log_std[a]);
            new_logp = new_logp - ((((0.5f * error[a]) * error[a]) * invvar[a] + log_std[a]) + RACER_HALF_LOG_2PI);
        }
        float logratio = new_logp - logp[s];
        float ratio = racer_exp(logratio);
        float a_hat = (advantage[s] - adv_mean) * adv_scale;
        float dlogp = (-a_hat * ratio) / (float)batch;
        if ((a_hat >= 0.0f && ratio > 1.0f + RACER_CLIP) || (a_hat < 0.0f && ratio < 1.0f - RACER_CLIP)) dlogp = 0.0f;
        for (int a = 0; a < 2; ++a) {
            d3[a * batch + b] = (dlogp * error[a]) * invvar[a];
            d3[(3 + a) * batch + b] = dlogp * ((error[a] * error[a]) * invvar[a] - 1.0f);
        }
        d3[2 * batch + b] = (0.5f * racer_clamp((o[2 * batch + b] - value[s]) - advantage[s], -1.0f, 1.0f)) / (float)batch;
        d3[5 * batch + b] = (ratio - 1.0f) - logratio;
    }
    o_backward(theta, beams, hidden, x, batch, h1, h2, d3, d2, d1, grad);
    for (int a = 0; a < 2; ++a) {
        double acc = 0.0;
        for (long b = 0; b < batch; ++b) acc += d3[(3 + a) * batch + b];
        grad[LOGSTD + a] = (float)acc;
    }
    double sq = 0.0;
    for (long i = 0; i < P; ++i) sq += (double)grad[i] * (double)grad[i];
    *norm_sq = (float)sq;
}

/* 1 - beta^step with beta the float32 constant, evaluated in float64 and
 * rounded (the reference's float32 pow, ppo.mojo:331-332). */
float o_bias_correction(float beta, int step) { return (float)(1.0 - pow((double)beta, (double)step)); }

/* ppo.mojo:325-336 */
void o_adam(float* theta, long P, long logstd_at, const float* grad, float* first, float* second, float norm_sq,
            int step) {
    float clip = racer_min(1.0f, RACER_MAX_NORM / (sqrtf(norm_sq) + 1e-8f));
    float c1 = o_bias_correction(0.9f, step);
    float c2 = o_bias_correction(0.999f, step);
    for (long i = 0; i < P; ++i) {
        float g = grad[i] * clip;
        float m = 0.9f * first[i] + 0.1f * g;
        float v = 0.999f * second[i] + (0.001f * g) * g;
        first[i] = m;
        second[i] = v;
        float m_hat = m / c1;
        float v_hat = v / c2;
        float value = theta[i] - (RACER_LR * m_hat) / (sqrtf(v_hat) + 1e-5f);
        theta[i] = i >= logstd_at ? racer_min(racer_max(value, -2.0f), 0.0f) : value;
    }
}

/* ---------------- exported elementary functions ---------------- */
float o_sin(float x) { return racer_sin(x); }
float o_cos(float x) { return racer_cos(x); }
float o_exp(float x) { return racer_exp(x); }
float o_log(float x) { return racer_log(x); }
float o_tanh(float x) { return racer_tanh(x); }
float o_tanh_mlp(float x) { return racer_tanh_mlp(x); }
void o_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    racer_u32x4 r = racer_philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int j = 0; j < 4; ++j) out[j] = r.v[j];
}
void o_normal4(uint64_t seed, uint64_t subsequence, uint64_t offset, float out[4]) {
    racer_normal4(racer_random(seed, subsequence, offset), out);
}
void o_uniform4(uint64_t seed, uint64_t subsequence, uint64_t offset, float out[4]) {
    racer_u32x4 r = racer_random(seed, subsequence, offset);
    for (int j = 0; j < 4; ++j) out[j] = racer_uniform(r.v[j]);
}

/* ---------------- flat wrappers of the Map lookups, for the tests ---------------- */
int o_probe(const float* d, size_t len, float x, float y, float z, float out[3]) {
    omap m;
    if (om_init(&m, d, len)) return -1;
    ospot s = om_probe(&m, d, x, y, z);
    out[0] = (float)s.layer;
    out[1] = s.clearance;
    out[2] = s.height;
    return 0;
}

int o_surface(const float* d, size_t len, int layer, float x, float y, float out[3]) {
    omap m;
    if (om_init(&m, d, len)) return -1;
    osurface s = om_surface(&m, d, layer, x, y);
    out[0] = s.height;
    out[1] = s.sx;
    out[2] = s.sy;
    return 0;
}

float o_progress(const float* d, size_t len, int layer, float x, float y) {
    omap m;
    if (om_init(&m, d, len)) return 0.0f;
    return om_progress(&m, d, layer, x, y);
}

float o_scan_one(const float* d, size_t len, const float* s, long n, const float* obs, long obs_cols, long i, int beam,
                 int beams, long* iterations) {
    omap m;
    if (om_init(&m, d, len)) return 0.0f;
    return o_scan(&m, d, s, n, obs, obs_cols, i, beam, beams, iterations);
}

int o_spawn_one(const float* d, size_t len, float* s, long n, long i, int seed, float heading_noise) {
    omap m;
    if (om_init(&m, d, len)) return -1;
    o_spawn(s, n, &m, d, i, seed, heading_noise);
    return 0;
}

/* ---------------- checks of the shared header's division shortcuts ---------------- */

/* Every float a in [lo, hi] (walking the ordered bit patterns): number of a
 * where racer_nearest_index(a, b, 1/b) differs from floorf(a / b + 0.5f). */
long o_check_nearest_index(float b, float lo, float hi) {
    float inv = 1.0f / b;
    uint32_t ulo = racer_f2u(lo), uhi = racer_f2u(hi);
    int64_t olo = (ulo & 0x80000000u) ? -(int64_t)(ulo & 0x7fffffffu) : (int64_t)ulo;
    int64_t ohi = (uhi & 0x80000000u) ? -(int64_t)(uhi & 0x7fffffffu) : (int64_t)uhi;
    long bad = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (int64_t o = olo; o <= ohi; ++o) {
        uint32_t u = o < 0 ? (uint32_t)(-o) | 0x80000000u : (uint32_t)o;
        float a = racer_u2f(u);
        float want = floorf(a / b + 0.5f);
        float got = racer_nearest_index(a, b, inv);
        if (racer_f2u(want) != racer_f2u(got)) ++bad;
    }
    return bad;
}

/* The same with the caller-supplied limit of a run of calls, Q = max(|lo|,
 * |hi|) / b (with 0.1 % slack) bounding |a / b| over the range. */
long o_check_nearest_index_lim(float b, float lo, float hi) {
    float inv = 1.0f / b;
    float lim = racer_nearest_limit(racer_max(racer_abs(lo), racer_abs(hi)) * inv * 1.001f);
    uint32_t ulo = racer_f2u(lo), uhi = racer_f2u(hi);
    int64_t olo = (ulo & 0x80000000u) ? -(int64_t)(ulo & 0x7fffffffu) : (int64_t)ulo;
    int64_t ohi = (uhi & 0x80000000u) ? -(int64_t)(uhi & 0x7fffffffu) : (int64_t)uhi;
    long bad = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (int64_t o = olo; o <= ohi; ++o) {
        uint32_t u = o < 0 ? (uint32_t)(-o) | 0x80000000u : (uint32_t)o;
        float a = racer_u2f(u);
        float want = floorf(a / b + 0.5f);
        float got = racer_nearest_index_lim(a, b, inv, lim);
        if (racer_f2u(want) != racer_f2u(got)) ++bad;
    }
    return bad;
}

/* `count` seeded random (cap, num, den) triples: mismatches of
 * racer_min_quotient against racer_min(cap, num / den). */
/* A float with the given unbiased exponent (-126 .. 127) and 23-bit fraction. */
static float o_make_float(int exponent, uint32_t fraction, int negative) {
    return racer_u2f(((uint32_t)negative << 31) | ((uint32_t)(exponent + 127) << 23) | (fraction & 0x7fffffu));
}

/* racer_div_by_reciprocal(a, b, RN(1 / b)) against the IEEE quotient a / b,
 * bit for bit, over its stated range. Cases: random a and b (exponents
 * -60 .. 60); a quotient near a rounding midpoint (a = RN(m b) for a midpoint
 * m between two floats, nudged by -2 .. +2 ulps); hard-to-round quotients,
 * within r 2^-25 ulp of a midpoint for r = +-1, +-3 (divisor significand B
 * odd, midpoint significand N = r B^-1 mod 2^25, then N B - r = 2^25 A and
 * a = A scaled; these catch a reciprocal that is one ulp off); divisors with
 * an all-ones fraction or a power of two; a = +0. */
long o_check_div_by_reciprocal(long count, uint64_t seed) {
    long bad = 0;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (long k = 0; k < count; ++k) {
        racer_u32x4 r = racer_random(seed, (uint64_t)k, 0);
        racer_u32x4 r2 = racer_random(seed, (uint64_t)k, 1);
        int kind = (int)(r.v[3] & 7u);
        int eb = (int)(r.v[0] % 121u) - 60;
        float b = o_make_float(eb, r.v[1], 0);
        float a = o_make_float((int)(r2.v[0] % 121u) - 60, r2.v[1], (int)(r2.v[2] & 1u));
        if (kind == 3) {
            /* quotient exponent chosen so that |a| = |m b| stays within [2^-59, 2^59] */
            int eq = (int)(r2.v[0] % 117u) - 58 - eb;
            if (eq < -58) eq = -58;
            if (eq > 58) eq = 58;
            if (eq + eb < -58 || eq + eb > 58) eq = -eb;
            float q = o_make_float(eq, r2.v[1], (int)(r2.v[2] & 1u));
            double m = (double)q + (q > 0 ? 0.5 : -0.5) * ((double)racer_u2f(racer_f2u(fabsf(q)) + 1u) - fabs((double)q));
            a = (float)(m * (double)b);
            int nudge = (int)(r2.v[3] % 5u) - 2;
            a = racer_u2f((uint32_t)((int32_t)racer_f2u(a) + nudge));
        } else if (kind == 4 || kind == 5) {
            uint32_t B = (r.v[1] & 0x7fffffu) | 0x800001u, x = B;
            for (int i = 0; i < 5; ++i) x *= 2u - B * x; /* B^-1 mod 2^32 */
            static const int rs[4] = {1, -1, 3, -3};
            int rr = rs[r2.v[3] & 3u];
            uint32_t N = ((uint32_t)rr * x) & 0x1ffffffu;
            int64_t A = ((int64_t)((uint64_t)N * B) - rr) >> 25;
            if (N < 0x1000000u || A < (1 << 23) || A >= (1 << 24)) continue; /* not a normal 25-bit midpoint */
            int e1 = (int)(r2.v[0] % 61u) - 30, e2 = (int)(r2.v[1] % 61u) - 30;
            b = ldexpf((float)B, e1 - 23);
            a = ldexpf((float)A, e1 + e2 - 22);
            if (r2.v[2] & 1u) a = -a;
        } else if (kind == 6) {
            b = o_make_float(eb, (r.v[3] & 8u) ? 0x7fffffu : 0u, 0);
        } else if (kind == 7) {
            a = 0.0f;
        }
        float y = 1.0f / b;
        float want = a / b;
        float got = racer_div_by_reciprocal(a, b, y);
        if (racer_f2u(want) != racer_f2u(got)) ++bad;
    }
    return bad;
}

/* The lidar's nearest cell through the reciprocal, floorf(RN-quotient + 0.5f)
 * with the quotient from racer_div_by_reciprocal(a, cell, RN(1 / cell)),
 * against floorf(a / cell + 0.5f) for the floats a whose bit patterns are
 * multiples of `stride` (1: all 2^32, NaNs and infinities included): the same
 * index wherever either lies in [0, 2^24), both outside it (or NaN)
 * elsewhere. */
/* racer_tanh_select against racer_tanh on every `stride`-th float bit
 * pattern (stride 1: all 2^32): the number of results that differ. */
long o_check_tanh_select(long stride) {
    long bad = 0;
    if (stride < 1) stride = 1;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (long i = 0; i <= 0xffffffffL; i += stride) {
        float x = racer_u2f((uint32_t)i);
        bad += racer_f2u(racer_tanh(x)) != racer_f2u(racer_tanh_select(x));
    }
    return bad;
}

long o_check_nearest_by_reciprocal(float cell, long stride) {
    const float inv = 1.0f / cell;
    long bad = 0;
    const int64_t count = 0xffffffffll / stride + 1;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (int64_t k = 0; k < count; ++k) {
        float a = racer_u2f((uint32_t)(k * stride));
        float want = floorf(a / cell + 0.5f);
        float got = floorf(racer_div_by_reciprocal(a, cell, inv) + 0.5f);
        int in_want = want >= 0.0f && want < 0x1p24f, in_got = got >= 0.0f && got < 0x1p24f;
        if (in_want != in_got || (in_want && racer_f2u(want) != racer_f2u(got))) ++bad;
    }
    return bad;
}

/* The lidar's one-step nearest cell (racer_nearest_by_reciprocal1) against
 * floorf(a / cell + 0.5f) over every stride-th float pattern: the same index
 * wherever either is a number below 2^22 in magnitude (any grid with its
 * padding). Elsewhere both are off every grid: beyond 2^22 or infinite, or
 * NaN where the quotient overflows (q0 = inf makes the residual NaN); the
 * unpadded probe's on-grid test rejects all of these alike. */
long o_check_nearest_by_reciprocal1(float cell, long stride) {
    const float inv = 1.0f / cell;
    long bad = 0;
    const int64_t count = 0xffffffffll / stride + 1;
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : bad)
    for (int64_t k = 0; k < count; ++k) {
        float a = racer_u2f((uint32_t)(k * stride));
        float want = floorf(a / cell + 0.5f);
        float got = racer_nearest_by_reciprocal1(a, cell, inv);
        const int in_want = racer_abs(want) < 0x1p22f, in_got = racer_abs(got) < 0x1p22f;  /* false for NaN */
        if ((in_want || in_got) && racer_f2u(want) != racer_f2u(got)) ++bad;
    }
    return bad;
}

/* The per-map check the library runs at sim creation (racer_nearest_form_bad). */
long o_nearest_form_bad(float cell, int nlo, int nhi, int form) { return racer_nearest_form_bad(cell, nlo, nhi, form); }

/* ---------------- Learned dynamics surrogate (no reference counterpart) ---------------- */

/* The surrogate's 11 inputs of car i and its constant-1 row (the HIP feature
 * kernel's definition): x, y, z, yaw, speed, yaw rate, steer, along and
 * across slopes, clamped throttle and steer command. */
void o_surrogate_features(const float* d, size_t len, const float* s, long n, const float* act, long act_cols,
                          float* feat, long ldf) {
    omap m;
    if (om_init(&m, d, len)) return;
    for (long i = 0; i < n; ++i) {
        float x = s[RACER_X * n + i], y = s[RACER_Y * n + i], z = s[RACER_Z * n + i];
        float heading = s[RACER_HEADING * n + i], speed = s[RACER_SPEED * n + i], steer = s[RACER_STEER * n + i];
        ospot spot = om_probe(&m, d, x, y, z);
        osurface g = om_surface(&m, d, spot.layer, x, y);
        float along, across;
        o_slopes(g, heading, &along, &across);
        float f[11] = {x, y, z, heading, speed, ((speed * racer_sin(steer)) / racer_cos(steer)) / RACER_WHEELBASE,
                       steer, along, across, racer_clamp(act[act_cols + i], -1.0f, 1.0f),
                       racer_clamp(act[i], -1.0f, 1.0f)};
        for (int r = 0; r < 11; ++r) feat[r * ldf + i] = f[r];
        feat[11 * ldf + i] = 1.0f;
    }
}

/* The one-step targets: displacement in the starting car frame, wrapped
 * heading change, speed and steer changes; mask 0 after a crash or timeout. */
void o_surrogate_targets(const float* prev, const float* s, const float* done, long n, float* tgt, long ldt,
                         float* mask) {
    for (long i = 0; i < n; ++i) {
        float h = prev[RACER_HEADING * n + i];
        float ch = racer_cos(h), sh = racer_sin(h);
        float dx = s[RACER_X * n + i] - prev[RACER_X * n + i];
        float dy = s[RACER_Y * n + i] - prev[RACER_Y * n + i];
        float dh = s[RACER_HEADING * n + i] - h;
        if (dh > RACER_PI) dh -= RACER_2PI;
        if (dh < -RACER_PI) dh += RACER_2PI;
        tgt[0 * ldt + i] = ch * dx + sh * dy;
        tgt[1 * ldt + i] = ch * dy - sh * dx;
        tgt[2 * ldt + i] = dh;
        tgt[3 * ldt + i] = s[RACER_SPEED * n + i] - prev[RACER_SPEED * n + i];
        tgt[4 * ldt + i] = s[RACER_STEER * n + i] - prev[RACER_STEER * n + i];
        mask[i] = done[i] == 0.0f ? 1.0f : 0.0f;
    }
}

/* The surrogate MLP in float32: widths[0 .. layers]; layer l's weights
 * [widths[l+1], widths[l] + 1] (bias last) at consecutive offsets; x holds
 * [widths[0] + 1, cols] with a last row of ones. acts receives each layer's
 * output in turn: hidden [widths[l] + 1, cols] (tanh, last row 1), then the
 * output [widths[layers], cols]. Products sum k in order with fmaf. */
void o_mlp_forward(const float* theta, const int* widths, int layers, const float* x, long cols, float* acts) {
    const float* in = x;
    const float* w = theta;
    float* out = acts;
    for (int l = 0; l < layers; ++l) {
        int rows = widths[l + 1], k = widths[l] + 1;
        int hidden = l < layers - 1;
        mm(w, k, in, cols, out, cols, rows, k, cols, hidden);
        if (hidden)
            for (long c = 0; c < cols; ++c) out[(long)rows * cols + c] = 1.0f;
        w += (long)rows * k;
        in = out;
        out += (long)(rows + hidden) * cols;
    }
}

/* Loss sum_b mask_b sum_o (out - y)^2 / (O cols) and its gradient (the HIP
 * surrogate's orders: output gradient ((e + e) mask) / (O cols), products
 * through the layers in k order times (1 - h^2), weight gradients by
 * chunked_dot). Returns the loss in float64. */
double o_mlp_gradient(const float* theta, const int* widths, int layers, const float* x, const float* y,
                      const float* mask, long cols, float* grad) {
    long total = 0, P = 0;
    for (int l = 0; l < layers; ++l) {
        total += (long)(widths[l + 1] + (l < layers - 1)) * cols;
        P += (long)widths[l + 1] * (widths[l] + 1);
    }
    float* acts = (float*)malloc(sizeof(float) * total);
    o_mlp_forward(theta, widths, layers, x, cols, acts);
    const float* a[16];
    const float* wl[16];
    a[0] = x;
    long ao = 0, wo = 0;
    for (int l = 0; l < layers; ++l) {
        a[l + 1] = acts + ao;
        wl[l] = theta + wo;
        ao += (long)(widths[l + 1] + (l < layers - 1)) * cols;
        wo += (long)widths[l + 1] * (widths[l] + 1);
    }
    int O = widths[layers];
    float* delta = (float*)malloc(sizeof(float) * (size_t)O * cols);
    double loss = 0.0;
    float denom = (float)((long)O * cols);
    for (long c = 0; c < cols; ++c)
        for (int o = 0; o < O; ++o) {
            float e = a[layers][o * cols + c] - y[o * cols + c];
            delta[o * cols + c] = ((e + e) * mask[c]) / denom;
            loss += (double)mask[c] * (double)e * (double)e;
        }
    loss /= (double)denom;
    long go = P;
    for (int l = layers - 1; l >= 0; --l) {
        int rows = widths[l + 1], k = widths[l] + 1;
        go -= (long)rows * k;
        for (int j = 0; j < rows; ++j)
            for (int q = 0; q < k; ++q) grad[go + (long)j * k + q] = chunked_dot(delta + (long)j * cols, a[l] + (long)q * cols, cols);
        if (l == 0) break;
        int below = widths[l];
        float* next = (float*)malloc(sizeof(float) * (size_t)below * cols);
        for (long c = 0; c < cols; ++c)
            for (int q = 0; q < below; ++q) {
                float acc = 0.0f;
                for (int j = 0; j < rows; ++j) acc = fmaf(wl[l][(long)j * k + q], delta[(long)j * cols + c], acc);
                float h = a[l][(long)q * cols + c];
                next[(long)q * cols + c] = acc * (1.0f - h * h);
            }
        free(delta);
        delta = next;
    }
    free(delta);
    free(acts);
    return loss;
}

/* ---------------- FP64 RK4 (the surrogate's float64 reference) ---------------- */

static void f64_probe(const omap* m, const float* d, double x, double y, double z, int* layer, double* clearance,
                      double* height) {
    double fx = floor((x - m->x0) / m->cell + 0.5), fy = floor((y - m->y0) / m->cell + 0.5);
    int on = fx >= 0.0 && fy >= 0.0 && fx < (double)m->nx && fy < (double)m->ny;
    long c = on ? (long)fy * m->nx + (long)fx : 0;
    *layer = 0;
    *clearance = -1.0;
    *height = on ? (double)d[om_grid(m, 0, RACER_HEIGHT) + c] : -1.0;
    double gap = 1e300;
    for (int l = 0; l < m->layers && on; ++l) {
        double cl = d[om_grid(m, l, RACER_CLEARANCE) + c], h = d[om_grid(m, l, RACER_HEIGHT) + c];
        if (cl > 0.0 && fabs(h - z) < gap) {
            *layer = l;
            *clearance = cl;
            *height = h;
            gap = fabs(h - z);
        }
    }
}

static void f64_surface(const omap* m, const float* d, int layer, double x, double y, double* sx, double* sy) {
    double fx = (x - m->x0) / m->cell, fy = (y - m->y0) / m->cell;
    double cx = floor(fx), cy = floor(fy);
    long ix = !(cx >= 0.0) ? 0 : (cx > (double)(m->nx - 2) ? m->nx - 2 : (long)cx);
    long iy = !(cy >= 0.0) ? 0 : (cy > (double)(m->ny - 2) ? m->ny - 2 : (long)cy);
    double tx = fmin(fmax(fx - (double)ix, 0.0), 1.0), ty = fmin(fmax(fy - (double)iy, 0.0), 1.0);
    long k = om_grid(m, layer, RACER_HEIGHT) + iy * m->nx + ix;
    double h0 = d[k], h1 = d[k + 1], h2 = d[k + m->nx], h3 = d[k + m->nx + 1];
    *sx = ((h1 - h0) * (1.0 - ty) + (h3 - h2) * ty) / m->cell;
    *sy = ((h2 - h0) * (1.0 - tx) + (h3 - h1) * tx) / m->cell;
}

static void f64_derivative(const double p[4], double steer, double accel, double grip, double along, double across,
                           double out[4]) {
    const double G = RACER_G, WB = RACER_WHEELBASE;
    double nz = 1.0 / sqrt(1.0 + along * along + across * across);
    double load = grip * G * nz;
    double v = fmax(fabs(p[3]), 0.5) * (p[3] >= 0.0 ? 1.0 : -1.0);
    double slide = G * across * nz;
    double lateral = fmin(fmax(v * p[3] * tan(steer) / WB + slide, -load), load);
    double yaw = (lateral - slide) / v;
    double longitudinal = sqrt(fmax(load * load - lateral * lateral, 0.0));
    out[0] = p[3] * cos(p[2]);
    out[1] = p[3] * sin(p[2]);
    out[2] = yaw;
    out[3] = fmin(fmax(accel, -longitudinal), longitudinal) - G * along * nz;
}

/* One environment step of the kinematics (sim.mojo:184-219) in float64 for n
 * cars: st rows x, y, z, heading, speed, steer (updated in place); friction
 * and actions as the sim's. No crash or respawn: the reference is for windows
 * without them. */
void o_step_f64(const float* d, size_t len, long n, double* st, const float* friction, const float* act,
                long act_cols) {
    omap m;
    if (om_init(&m, d, len)) return;
    const double DT = 1.0 / 240.0;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long i = 0; i < n; ++i) {
        double p[4] = {st[0 * n + i], st[1 * n + i], st[3 * n + i], st[4 * n + i]};
        double steer = st[5 * n + i];
        int layer;
        double clearance, height, sx, sy;
        f64_probe(&m, d, p[0], p[1], st[2 * n + i], &layer, &clearance, &height);
        f64_surface(&m, d, layer, p[0], p[1], &sx, &sy);
        double rate = fmin(fmax((double)act[i], -1.0), 1.0) * RACER_STEER_RATE;
        double accel = fmin(fmax((double)act[act_cols + i], -1.0), 1.0) * RACER_ACCEL_MAX;
        double grip = RACER_GRIP * (double)friction[i];
        for (int sub = 0; sub < RACER_SUBSTEPS; ++sub) {
            double c = cos(p[2]), s = sin(p[2]);
            double along = sx * c + sy * s, across = -sx * s + sy * c;
            double half = steer + rate * DT / 2.0, full = steer + rate * DT;
            double k1[4], k2[4], k3[4], k4[4], q[4];
            f64_derivative(p, steer, accel, grip, along, across, k1);
            for (int j = 0; j < 4; ++j) q[j] = p[j] + k1[j] * DT / 2.0;
            f64_derivative(q, half, accel, grip, along, across, k2);
            for (int j = 0; j < 4; ++j) q[j] = p[j] + k2[j] * DT / 2.0;
            f64_derivative(q, half, accel, grip, along, across, k3);
            for (int j = 0; j < 4; ++j) q[j] = p[j] + k3[j] * DT;
            f64_derivative(q, full, accel, grip, along, across, k4);
            for (int j = 0; j < 4; ++j) p[j] += (k1[j] + 2.0 * (k2[j] + k3[j]) + k4[j]) * DT / 6.0;
            p[3] = fmin(fmax(p[3], -(double)RACER_SPEED_MAX), (double)RACER_SPEED_MAX);
            steer = fmin(fmax(full, -(double)RACER_STEER_MAX), (double)RACER_STEER_MAX);
            f64_probe(&m, d, p[0], p[1], height, &layer, &clearance, &height);
            f64_surface(&m, d, layer, p[0], p[1], &sx, &sy);
        }
        const double PI = 3.14159265358979323846;
        if (p[2] > PI) p[2] -= 2.0 * PI;
        if (p[2] < -PI) p[2] += 2.0 * PI;
        st[0 * n + i] = p[0];
        st[1 * n + i] = p[1];
        st[2 * n + i] = height;
        st[3 * n + i] = p[2];
        st[4 * n + i] = p[3];
        st[5 * n + i] = steer;
    }
}
