// Simulation kernels for gfx950: track lookups, the RK4 slope bicycle with the
// friction circle, crash/respawn/progress, the sphere-traced lidar, action
// sampling (racer/sim.mojo), and the Sim C ABI.
//
// Float32 arithmetic is written in the reference's evaluation order and built
// with -ffp-contract=off, using the shared float32 elementary functions, so
// every result is bit-identical to the CPU restatement.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "racer_internal.h"

namespace {

struct Spot {
    int layer;
    float clearance;
    float height;
};

struct Surf {
    float height, sx, sy;
};

// Nearest cell of (x, y) (racer/sim.mojo:108-114); false off the map. DIV:
// floor((x - x0) / cell + 0.5) with the division, as the reference writes it
// (the reference decomposition's lidar); otherwise racer_nearest_index, the
// same index from the reciprocal with the division only near a half-integer.
template <bool DIV = false>
__device__ __forceinline__ bool cell_of(const racer_dmap& m, float x, float y, int64_t* cell) {
    float fx = DIV ? floorf((x - m.x0) / m.cell + 0.5f) : racer_nearest_index(x - m.x0, m.cell, m.inv_cell);
    float fy = DIV ? floorf((y - m.y0) / m.cell + 0.5f) : racer_nearest_index(y - m.y0, m.cell, m.inv_cell);
    if (!(fx >= 0.0f) || !(fy >= 0.0f) || fx >= (float)m.nx || fy >= (float)m.ny) return false;
    *cell = (int64_t)fy * m.nx + (int64_t)fx;
    return true;
}

__device__ __forceinline__ float grid_at(const racer_dmap& m, int layer, int which, int64_t cell) {
    return m.d[RACER_HEADER + (int64_t)(3 * layer + which) * m.plane + cell];
}

// Map.at (racer/sim.mojo:108-114)
__device__ __forceinline__ float map_at(const racer_dmap& m, int layer, int which, float x, float y) {
    int64_t c;
    return cell_of(m, x, y, &c) ? grid_at(m, layer, which, c) : -1.0f;
}

// Map.probe (racer/sim.mojo:116-124): the drivable layer whose height is nearest z.
template <bool DIV = false>
__device__ Spot map_probe(const racer_dmap& m, float x, float y, float z) {
    int64_t c = 0;
    bool on = cell_of<DIV>(m, x, y, &c);
    Spot best{0, -1.0f, on ? grid_at(m, 0, RACER_HEIGHT, c) : -1.0f};
    float gap = 1e30f;
    for (int l = 0; l < m.layers; ++l) {
        float cl = on ? grid_at(m, l, RACER_CLEARANCE, c) : -1.0f;
        float h = on ? grid_at(m, l, RACER_HEIGHT, c) : -1.0f;
        if (cl > 0.0f && fabsf(h - z) < gap) {
            best.layer = l;
            best.clearance = cl;
            best.height = h;
            gap = fabsf(h - z);
        }
    }
    return best;
}

// Map.surface (racer/sim.mojo:126-140): bilinear height and its gradient.
__device__ Surf map_surface(const racer_dmap& m, int layer, float x, float y) {
    float fx = (x - m.x0) / m.cell;
    float fy = (y - m.y0) / m.cell;
    float cx = floorf(fx), cy = floorf(fy);
    int64_t ix = !(cx >= 0.0f) ? 0 : (cx > (float)(m.nx - 2) ? m.nx - 2 : (int64_t)cx);
    int64_t iy = !(cy >= 0.0f) ? 0 : (cy > (float)(m.ny - 2) ? m.ny - 2 : (int64_t)cy);
    float tx = racer_clamp(fx - (float)ix, 0.0f, 1.0f);
    float ty = racer_clamp(fy - (float)iy, 0.0f, 1.0f);
    const float* h = m.d + RACER_HEADER + (int64_t)(3 * layer + RACER_HEIGHT) * m.plane + iy * m.nx + ix;
    float h0 = h[0], h1 = h[1], h2 = h[m.nx], h3 = h[m.nx + 1];
    float w0 = (1.0f - tx) * (1.0f - ty);
    float w1 = tx * (1.0f - ty);
    float w2 = (1.0f - tx) * ty;
    float w3 = tx * ty;
    Surf s;
    s.height = (h0 * w0 + h2 * w2) + (h1 * w1 + h3 * w3);  // 4-lane tree sum
    s.sx = ((h1 - h0) * (1.0f - ty) + (h3 - h2) * ty) / m.cell;
    s.sy = ((h2 - h0) * (1.0f - tx) + (h3 - h1) * tx) / m.cell;
    return s;
}

// Surface.slopes (racer/sim.mojo:65-69): rise per metre ahead and to the left.
__device__ __forceinline__ void slopes(const Surf& s, float heading, float* along, float* across) {
    float c = racer_cos(heading);
    float sn = racer_sin(heading);
    *along = s.sx * c + s.sy * sn;
    *across = -s.sx * sn + s.sy * c;
}

__device__ __forceinline__ float waypoint(const racer_dmap& m, int64_t w, int field) {
    return m.d[m.route + 4 * w + field];
}

// Map.progress (racer/sim.mojo:146-152) with its first read, Map.at(layer,
// NEAREST, x, y), already done (-1 off the map: then w = -1 reads the floats
// just before the route table, as the reference does).
__device__ float map_progress_at(const racer_dmap& m, float nearest, float x, float y) {
    int64_t w = (int64_t)nearest;
    float heading = waypoint(m, w, 3);
    float along = (x - waypoint(m, w, 0)) * racer_cos(heading) + (y - waypoint(m, w, 1)) * racer_sin(heading);
    return (float)w + along / m.spacing;
}

// derivative (racer/sim.mojo:155-167): the kinematic bicycle on a slope.
// tan(steer) is passed in as (sin, cos) so stages sharing a steer share them.
__device__ __forceinline__ void derivative(const float p[4], float ssin, float scos, float accel, float grip,
                                           float along, float across, float out[4]) {
    float nz = 1.0f / sqrtf((1.0f + along * along) + across * across);
    float load = (grip * RACER_G) * nz;
    float v = racer_max(fabsf(p[3]), 0.5f) * (p[3] >= 0.0f ? 1.0f : -1.0f);
    float slide = (RACER_G * across) * nz;
    float lateral = racer_clamp((((v * p[3]) * ssin) / scos) / RACER_WHEELBASE + slide, -load, load);
    float yaw = (lateral - slide) / v;
    float longitudinal = sqrtf(racer_max(load * load - lateral * lateral, 0.0f));
    float dv = racer_clamp(accel, -longitudinal, longitudinal) - (RACER_G * along) * nz;
    out[0] = p[3] * racer_cos(p[2]);
    out[1] = p[3] * racer_sin(p[2]);
    out[2] = yaw;
    out[3] = dv;
}

// spawn (racer/sim.mojo:170-181); heading_noise != 0 offsets the heading by
// heading_noise * (2 u2 - 1) from the same Philox block.
__device__ void spawn(const racer_dmap& m, float* s, int n, int i, int seed, float heading_noise) {
    float episode = s[RACER_EPISODE * n + i] + 1.0f;
    racer_u32x4 r = racer_random((uint64_t)(int64_t)seed, (uint64_t)i, (uint64_t)episode);
    float u0 = racer_uniform(r.v[0]);
    float u1 = racer_uniform(r.v[1]);
    int64_t w = (int64_t)(u0 * (float)m.waypoints);
    if (w > m.waypoints - 1) w = m.waypoints - 1;
    for (int row = 0; row < RACER_STATE; ++row) s[row * n + i] = 0.0f;
    s[RACER_X * n + i] = waypoint(m, w, 0);
    s[RACER_Y * n + i] = waypoint(m, w, 1);
    s[RACER_Z * n + i] = waypoint(m, w, 2);
    float heading = waypoint(m, w, 3);
    if (heading_noise != 0.0f) heading = heading + heading_noise * (2.0f * racer_uniform(r.v[2]) - 1.0f);
    s[RACER_HEADING * n + i] = heading;
    s[RACER_PROGRESS * n + i] = (float)w;
    s[RACER_EPISODE * n + i] = episode;
    s[RACER_FRICTION * n + i] = 0.85f + 0.3f * u1;
}

// Map.probe at (x, y, z) and Map.surface of the probed layer at (x, y)
// (racer/sim.mojo:116-140). With L = 1 or 2 layers both read the packed
// (clearance, height) records: the probe cell and the four bilinear corners
// are loaded together, every layer's heights at once, and the probed layer's
// are selected afterwards, so the five loads are one round trip. L = 0 reads
// the track array layer by layer.
// nearest (optional): Map.at(layer, NEAREST, x, y) of the probed layer, the
// progress lookup's first read, loaded with the same round of loads.
//
// TILE (L = 1 or 2): a copy of the records and nearest-waypoint cells of a
// PHYS_TILE x PHYS_TILE window around the car sits in LDS; a probe whose five
// cells all lie in the window reads them there, any other reads the track
// array. The values are the same floats either way.
constexpr int PHYS_TILE = 10;
struct PhysTile {
    const float2* rec;  // [PHYS_TILE * PHYS_TILE][L] records, row-major from (cx0, cy0)
    const float* near;  // [L][PHYS_TILE * PHYS_TILE] nearest-waypoint cells
    int cx0, cy0;       // the window's first cell
};

template <int L, bool TILE = false>
__device__ __forceinline__ void probe_surface(const racer_dmap& m, float x, float y, float z, Spot* spot, Surf* g,
                                              float* nearest = nullptr, const PhysTile* tile = nullptr) {
    if constexpr (L == 0) {
        *spot = map_probe(m, x, y, z);
        *g = map_surface(m, spot->layer, x, y);
        if (nearest) *nearest = map_at(m, spot->layer, RACER_NEAREST, x, y);
    } else {
        // One IEEE quotient per axis serves Map.at's nearest cell floor(q +
        // 0.5) (racer/sim.mojo:110-111) and Map.surface's bilinear cell and
        // fraction (:128-133); cells fit 32-bit integers (grids of at most
        // 2^24 cells), and the tests and clamps are selects, not branches (this
        // runs five times in each car's serial chain). fmaxf / fminf clamp a
        // NaN to 0 and +inf to n - 2, as the reference's comparisons do.
        const float2* rec = reinterpret_cast<const float2*>(m.packed);
        const float qx = (x - m.x0) / m.cell, qy = (y - m.y0) / m.cell;
        const float fxn = floorf(qx + 0.5f), fyn = floorf(qy + 0.5f);
        const bool on = (fxn >= 0.0f) & (fyn >= 0.0f) & (fxn < (float)m.nx) & (fyn < (float)m.ny);
        const int c = on ? (int)fyn * m.nx + (int)fxn : 0;
        const int ix = (int)fminf(fmaxf(floorf(qx), 0.0f), (float)(m.nx - 2));
        const int iy = (int)fminf(fmaxf(floorf(qy), 0.0f), (float)(m.ny - 2));
        const float tx = racer_clamp(qx - (float)ix, 0.0f, 1.0f);
        const float ty = racer_clamp(qy - (float)iy, 0.0f, 1.0f);
        const int k = iy * m.nx + ix;
        float2 pr[L], r0[L], r1[L], r2[L], r3[L];
        float nr[L];
        bool inside = false;
        int tc = 0, tk = 0;  // the probe cell and the first corner in the window
        if constexpr (TILE) {
            constexpr unsigned T = PHYS_TILE;
            const int ux = ix - tile->cx0, uy = iy - tile->cy0;
            // clamped before the conversion; only used when on the map
            const int vx = (int)fminf(fmaxf(fxn, -1.0f), (float)m.nx) - tile->cx0;
            const int vy = (int)fminf(fmaxf(fyn, -1.0f), (float)m.ny) - tile->cy0;
            inside = ((unsigned)ux < T - 1) & ((unsigned)uy < T - 1) &
                     (!on | (((unsigned)vx < T) & ((unsigned)vy < T)));
            tc = on ? vy * (int)T + vx : 0;
            tk = uy * (int)T + ux;
        }
        if (inside) {
            constexpr int T = PHYS_TILE;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const float2 v = tile->rec[tc * L + l];
                pr[l] = on ? v : make_float2(-1.0f, -1.0f);
                r0[l] = tile->rec[tk * L + l];
                r1[l] = tile->rec[(tk + 1) * L + l];
                r2[l] = tile->rec[(tk + T) * L + l];
                r3[l] = tile->rec[(tk + T + 1) * L + l];
                if (nearest) {
                    const float w = tile->near[l * T * T + tc];
                    nr[l] = on ? w : -1.0f;
                }
            }
        } else {
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const float2 v = rec[c * L + l];  // c = 0 off the map: a valid record, not used
                pr[l] = on ? v : make_float2(-1.0f, -1.0f);
                r0[l] = rec[k * L + l];
                r1[l] = rec[(k + 1) * L + l];
                r2[l] = rec[(k + m.nx) * L + l];
                r3[l] = rec[(k + m.nx + 1) * L + l];
                if (nearest) {
                    const float w = grid_at(m, l, RACER_NEAREST, c);
                    nr[l] = on ? w : -1.0f;
                }
            }
        }
        Spot best{0, -1.0f, pr[0].y};
        float gap = 1e30f;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const float d = fabsf(pr[l].y - z);
            const bool take = (pr[l].x > 0.0f) & (d < gap);
            best.layer = take ? l : best.layer;
            best.clearance = take ? pr[l].x : best.clearance;
            best.height = take ? pr[l].y : best.height;
            gap = take ? d : gap;
        }
        *spot = best;
        if (nearest) {
            float w = nr[0];
#pragma unroll
            for (int l = 1; l < L; ++l) w = best.layer == l ? nr[l] : w;
            *nearest = w;
        }
        float h0 = r0[0].y, h1 = r1[0].y, h2 = r2[0].y, h3 = r3[0].y;
#pragma unroll
        for (int l = 1; l < L; ++l) {
            const bool sel = best.layer == l;
            h0 = sel ? r0[l].y : h0;
            h1 = sel ? r1[l].y : h1;
            h2 = sel ? r2[l].y : h2;
            h3 = sel ? r3[l].y : h3;
        }
        float w0 = (1.0f - tx) * (1.0f - ty);
        float w1 = tx * (1.0f - ty);
        float w2 = (1.0f - tx) * ty;
        float w3 = tx * ty;
        g->height = (h0 * w0 + h2 * w2) + (h1 * w1 + h3 * w3);
        g->sx = ((h1 - h0) * (1.0f - ty) + (h3 - h2) * ty) / m.cell;
        g->sy = ((h2 - h0) * (1.0f - tx) + (h3 - h1) * tx) / m.cell;
    }
}

template <int L>
__device__ void finish(const racer_dmap& m, float* s, int n, float* reward, float* done, float* obs,
                       int64_t obs_cols, int beams, int i, int seed, float heading_noise, float p[4], float steer,
                       const Spot& spot, Surf ground, bool crashed, float nearest, float* pose = nullptr);

// What a step of car i reads before its first probe: the pose, steer and
// friction rows of the state and the two action rows.
struct CarIn {
    float x, y, z, heading, speed, steer, friction, act_steer, act_throttle;
};

__device__ __forceinline__ CarIn load_car(const float* s, int n, const float* act, int64_t act_cols, int i) {
    return {s[RACER_X * n + i],     s[RACER_Y * n + i],     s[RACER_Z * n + i],
            s[RACER_HEADING * n + i], s[RACER_SPEED * n + i], s[RACER_STEER * n + i],
            s[RACER_FRICTION * n + i], act[i],               act[act_cols + i]};
}

// advance (racer/sim.mojo:184-235): one 60 Hz step of car i from its inputs
// `in`. TILE: the substeps' probes read the LDS window `tile` where it covers
// them, and the new pose goes to `pose` (finish).
template <int L, bool TILE = false>
__device__ void advance(const racer_dmap& m, float* s, int n, const CarIn& in, float* reward, float* done,
                        float* obs, int64_t obs_cols, int beams, int i, int seed, float heading_noise,
                        const PhysTile* tile = nullptr, float* pose = nullptr) {
    const float DT = RACER_DT;
    float p[4] = {in.x, in.y, in.heading, in.speed};
    Spot spot;
    Surf ground;
    probe_surface<L, TILE>(m, p[0], p[1], in.z, &spot, &ground, nullptr, tile);
    float steer = in.steer;
    float rate = racer_clamp(in.act_steer, -1.0f, 1.0f) * RACER_STEER_RATE;
    float accel = racer_clamp(in.act_throttle, -1.0f, 1.0f) * RACER_ACCEL_MAX;
    float grip = RACER_GRIP * in.friction;
    bool crashed = false;
    float nearest = -1.0f;  // the last probe's nearest waypoint (progress)
    for (int sub = 0; sub < RACER_SUBSTEPS; ++sub) {
        float along, across;
        slopes(ground, p[2], &along, &across);
        float half = steer + (rate * DT) / 2.0f;
        float full = steer + rate * DT;
        float s0 = racer_sin(steer), c0 = racer_cos(steer);
        float s1 = racer_sin(half), c1 = racer_cos(half);
        float s2 = racer_sin(full), c2 = racer_cos(full);
        float k1[4], k2[4], k3[4], k4[4], q[4];
        derivative(p, s0, c0, accel, grip, along, across, k1);
        for (int j = 0; j < 4; ++j) q[j] = p[j] + k1[j] * (DT / 2.0f);
        derivative(q, s1, c1, accel, grip, along, across, k2);
        for (int j = 0; j < 4; ++j) q[j] = p[j] + k2[j] * (DT / 2.0f);
        derivative(q, s1, c1, accel, grip, along, across, k3);
        for (int j = 0; j < 4; ++j) q[j] = p[j] + k3[j] * DT;
        derivative(q, s2, c2, accel, grip, along, across, k4);
        for (int j = 0; j < 4; ++j) p[j] = p[j] + ((k1[j] + 2.0f * (k2[j] + k3[j])) + k4[j]) * (DT / 6.0f);
        p[3] = racer_clamp(p[3], -RACER_SPEED_MAX, RACER_SPEED_MAX);
        steer = racer_clamp(full, -RACER_STEER_MAX, RACER_STEER_MAX);
        const float previous = spot.height;  // the re-probe uses the last spot's road height
        probe_surface<L, TILE>(m, p[0], p[1], previous, &spot, &ground,
                               sub == RACER_SUBSTEPS - 1 ? &nearest : nullptr, tile);
        crashed = crashed || spot.clearance < RACER_HALF_DIAG ||
                  ground.sx * ground.sx + ground.sy * ground.sy > 1.0f;
    }
    finish<L>(m, s, n, reward, done, obs, obs_cols, beams, i, seed, heading_noise, p, steer, spot, ground, crashed,
              nearest, pose);
}

// The end of advance (racer/sim.mojo:219-235) from the integrated pose p =
// (x, y, heading, speed), steer, the last probe (with its nearest waypoint at
// (x, y), read in the same round of loads) and the crash flag: heading wrap,
// progress, reward, done code, state and return, respawn, obs rows.
// pose (optional): heading, along, across, x, y, z after the step (and the
// respawn), the lidar's inputs, also written there.
template <int L>
__device__ void finish(const racer_dmap& m, float* s, int n, float* reward, float* done, float* obs,
                       int64_t obs_cols, int beams, int i, int seed, float heading_noise, float p[4], float steer,
                       const Spot& spot, Surf ground, bool crashed, float nearest, float* pose) {
    if (p[2] > RACER_PI) p[2] -= RACER_2PI;
    if (p[2] < -RACER_PI) p[2] += RACER_2PI;
    float progress = map_progress_at(m, nearest, p[0], p[1]);
    float dp = progress - s[RACER_PROGRESS * n + i];
    float halfw = (float)m.waypoints / 2.0f;
    dp = dp > halfw ? dp - 2.0f * halfw : (dp < -halfw ? dp + 2.0f * halfw : dp);
    float score = crashed ? -2.0f : (10.0f * dp) * m.spacing - (0.001f * steer) * steer;
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
        spawn(m, s, n, i, seed, heading_noise);
        float x = s[RACER_X * n + i], y = s[RACER_Y * n + i];
        Spot respawned;
        probe_surface<L>(m, x, y, s[RACER_Z * n + i], &respawned, &ground);
    }
    float along, across;
    slopes(ground, s[RACER_HEADING * n + i], &along, &across);
    obs[0 * obs_cols + i] = s[RACER_STEER * n + i] / RACER_STEER_MAX;
    obs[1 * obs_cols + i] = s[RACER_SPEED * n + i] / RACER_SPEED_MAX;
    obs[2 * obs_cols + i] = along;
    obs[3 * obs_cols + i] = across;
    obs[(int64_t)(4 + beams) * obs_cols + i] = 1.0f;
    if (pose) {
        pose[0] = s[RACER_HEADING * n + i];
        pose[1] = along;
        pose[2] = across;
        pose[3] = s[RACER_X * n + i];
        pose[4] = s[RACER_Y * n + i];
        pose[5] = s[RACER_Z * n + i];
    }
}

template <int L>
__global__ void __launch_bounds__(64) k_physics(racer_dmap m, float* s, int n, const float* act, int64_t act_cols,
                                                float* reward, float* done, float* obs, int64_t obs_cols, int beams,
                                                int seed, float heading_noise) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i < n)
        advance<L>(m, s, n, load_car(s, n, act, act_cols, i), reward, done, obs, obs_cols, beams, i, seed,
                   heading_noise);
}

// ---- learned dynamics surrogate (no reference counterpart) ----
// The state a surrogate step starts from: x, y, z, heading, speed, steer and
// the probe at the car (layer, clearance, height).
struct SurStart {
    float x, y, z, h, speed, steer;
    Spot now;
};

// Inputs of the surrogate for car i: x, y, z, yaw, speed, yaw rate
// v tan(steer) / wheelbase, steer, pitch and roll (the road's rise per metre
// ahead and to the left under the car), throttle and steer command (the
// clamped actions). Also returns the start state for surrogate_apply.
template <int L>
__device__ __forceinline__ void surrogate_inputs(const racer_dmap& m, const float* __restrict__ s, int n,
                                                 const float* __restrict__ act, int64_t act_cols, int i,
                                                 float f[11], SurStart* c) {
    c->x = s[RACER_X * n + i];
    c->y = s[RACER_Y * n + i];
    c->z = s[RACER_Z * n + i];
    c->h = s[RACER_HEADING * n + i];
    c->speed = s[RACER_SPEED * n + i];
    c->steer = s[RACER_STEER * n + i];
    Surf ground;
    probe_surface<L>(m, c->x, c->y, c->z, &c->now, &ground);
    float along, across;
    slopes(ground, c->h, &along, &across);
    f[0] = c->x;
    f[1] = c->y;
    f[2] = c->z;
    f[3] = c->h;
    f[4] = c->speed;
    f[5] = ((c->speed * racer_sin(c->steer)) / racer_cos(c->steer)) / RACER_WHEELBASE;
    f[6] = c->steer;
    f[7] = along;
    f[8] = across;
    f[9] = racer_clamp(act[act_cols + i], -1.0f, 1.0f);
    f[10] = racer_clamp(act[i], -1.0f, 1.0f);
}

// The inputs as rows of feat (ldf apart), normalised with mean / inv_std when
// given; row 11 is the constant 1 that carries the first layer's bias.
template <int L>
__global__ void __launch_bounds__(256) k_surrogate_features(racer_dmap m, const float* __restrict__ s, int n,
                                                            const float* __restrict__ act, int64_t act_cols,
                                                            float* __restrict__ feat, int64_t ldf,
                                                            const float* __restrict__ mean,
                                                            const float* __restrict__ inv_std) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[11];
    SurStart c;
    surrogate_inputs<L>(m, s, n, act, act_cols, i, f, &c);
#pragma unroll
    for (int r = 0; r < 11; ++r) {
        const float v = mean ? (f[r] - mean[r]) * inv_std[r] : f[r];
        feat[(int64_t)r * ldf + i] = v;
    }
    feat[(int64_t)11 * ldf + i] = 1.0f;
}

// One-step targets from the state before (prev) and after (s) a step: the
// displacement in the car's frame at the start (forward, left), the heading
// change wrapped to [-pi, pi], the speed and steer changes; mask 1 unless the
// step ended in a crash or timeout (then the car respawned).
__global__ void __launch_bounds__(256) k_surrogate_targets(const float* __restrict__ prev, const float* __restrict__ s,
                                                           const float* __restrict__ done, int n,
                                                           float* __restrict__ tgt, int64_t ldt,
                                                           float* __restrict__ mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float h = prev[RACER_HEADING * n + i];
    const float ch = racer_cos(h), sh = racer_sin(h);
    const float dx = s[RACER_X * n + i] - prev[RACER_X * n + i];
    const float dy = s[RACER_Y * n + i] - prev[RACER_Y * n + i];
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

// advance from start c with the pose change predicted by the surrogate (pred:
// normalised forward, left, heading, speed, steer changes; out = pred std +
// mean) in place of the RK4 integration. The rest of the step is the exact
// one (finish): crash from the probe at the predicted position, progress,
// reward, respawn, obs rows. A car whose prediction lands below
// min_clearance of the walls or on another layer takes the exact RK4 step
// instead; returns whether it did.
template <int L>
__device__ __forceinline__ bool surrogate_apply(const racer_dmap& m, float* s, int n, const float* act,
                                                int64_t act_cols, float* reward, float* done, float* obs,
                                                int64_t obs_cols, int beams, int seed, float heading_noise,
                                                const float pred[5], const SurStart& c,
                                                const float* __restrict__ out_mean,
                                                const float* __restrict__ out_std, float min_clearance, int i) {
    float d[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) d[o] = pred[o] * out_std[o] + out_mean[o];
    const float ch = racer_cos(c.h), sh = racer_sin(c.h);
    float p[4] = {c.x + (ch * d[0] - sh * d[1]), c.y + (sh * d[0] + ch * d[1]), c.h + d[2],
                  racer_clamp(c.speed + d[3], -RACER_SPEED_MAX, RACER_SPEED_MAX)};
    const float steer = racer_clamp(c.steer + d[4], -RACER_STEER_MAX, RACER_STEER_MAX);
    Spot spot;
    Surf ground;
    float nearest = -1.0f;
    probe_surface<L>(m, p[0], p[1], c.now.height, &spot, &ground, &nearest);
    const bool back = !(spot.clearance >= min_clearance) || spot.layer != c.now.layer;
    if (back) {
        advance<L>(m, s, n, load_car(s, n, act, act_cols, i), reward, done, obs, obs_cols, beams, i, seed,
                   heading_noise);
        return true;
    }
    const bool crashed = spot.clearance < RACER_HALF_DIAG || ground.sx * ground.sx + ground.sy * ground.sy > 1.0f;
    finish<L>(m, s, n, reward, done, obs, obs_cols, beams, i, seed, heading_noise, p, steer, spot, ground, crashed,
              nearest);
    return false;
}

// The surrogate step with the predictions of a separate forward (pred rows ldp apart).
template <int L>
__global__ void __launch_bounds__(64) k_physics_surrogate(racer_dmap m, float* s, int n, const float* act,
                                                          int64_t act_cols, float* reward, float* done, float* obs,
                                                          int64_t obs_cols, int beams, int seed, float heading_noise,
                                                          const float* __restrict__ pred, int64_t ldp,
                                                          const float* __restrict__ out_mean,
                                                          const float* __restrict__ out_std, float min_clearance,
                                                          float* fallback) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float q[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) q[o] = pred[o * ldp + i];
    SurStart c{s[RACER_X * n + i], s[RACER_Y * n + i], s[RACER_Z * n + i], s[RACER_HEADING * n + i],
               s[RACER_SPEED * n + i], s[RACER_STEER * n + i], Spot{}};
    Surf ground;
    probe_surface<L>(m, c.x, c.y, c.z, &c.now, &ground);
    const bool back = surrogate_apply<L>(m, s, n, act, act_cols, reward, done, obs, obs_cols, beams, seed,
                                         heading_noise, q, c, out_mean, out_std, min_clearance, i);
    if (fallback) fallback[i] = back ? 1.0f : 0.0f;
}

// The whole bf16 surrogate step for 64 cars per workgroup in one launch, on
// nine wavefronts. Wavefront 8, the car wavefront (one car per lane),
// computes the inputs while wavefronts 0 .. 7 load the first layer's weight
// fragments and set up LDS, keeps each car's start state in registers through the network,
// and applies the pose update at the end. Wavefronts 0 .. 7 run the three
// 128-wide tanh layers on v_mfma_f32_32x32x16_bf16, wavefront w owning rows
// 32 (w % 4) .. + 31 of each layer for cars 32 (w / 4) .. + 31 (activations
// in LDS as bf16 [car][k]), two per SIMD so that one's tanh epilogue runs
// under the other's products and LDS reads. Wavefronts 0 .. 4 then run the
// float32 output layer, one lane per output and car, k in order, four k per
// LDS read from the car-major h3. The operands, instruction and k order of
// each product are those of the separate bf16 and float32 products, so the
// step is the same as the unfused one: the only difference is the unfused
// products' trailing all-zero k-steps, which change an accumulator only when
// it is exactly -0.
constexpr int SF_CARS = 64, SF_NET_WAVES = 8, SF_THREADS = 64 * (SF_NET_WAVES + 1);
constexpr int SF_NET_THREADS = 64 * SF_NET_WAVES;
constexpr int SF_LDX = 152;  // bf16 row stride: 144 k + 8 (conflict-free 16-byte reads)
constexpr int SF_LDH = 132;  // float row stride of h3 and w3s: 129 k padded to 16 bytes

//
// MULTI: `steps` imagined steps in one launch (step k with act + k act_stride),
// the second and third layers' fragments copied to LDS once (72 KB; kept in
// registers across steps they would spill) and the output layer's weights
// kept in LDS; each car's state is carried through the state rows it owns.
// The one-step kernel (MULTI false) keeps the fragments in registers.
constexpr int SF_FRAGS = 4 * RACER_SF_STEPS * 64;  // bf16x8 fragments of one 128-wide layer
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int L, bool MULTI>
__global__ void __launch_bounds__(SF_THREADS) k_surrogate_fused(racer_dmap m, float* s, int n, const float* act,
                                                                int64_t act_cols, int steps, int64_t act_stride,
                                                                float* reward, float* done, float* obs,
                                                                int64_t obs_cols, int beams, int seed,
                                                                float heading_noise,
                                                                const uint16_t* __restrict__ frags,
                                                                const float* __restrict__ w3,
                                                                const float* __restrict__ st, float min_clearance,
                                                                float* fallback) {
    __shared__ __attribute__((aligned(16))) uint16_t xa[SF_CARS][SF_LDX];
    __shared__ __attribute__((aligned(16))) uint16_t xb[SF_CARS][SF_LDX];
    __shared__ __attribute__((aligned(16))) float h3[SF_CARS][SF_LDH];  // float32 input of the output layer
    __shared__ __attribute__((aligned(16))) float w3s[5][SF_LDH];
    __shared__ float pred[5][SF_CARS];
    __shared__ bf16x8 wfrag[MULTI ? 2 * SF_FRAGS : 1];  // MULTI: layers 1 and 2, [layer][row tile][k][lane]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool net = wave < SF_NET_WAVES;  // wavefront-uniform
    const int rt = wave & 3, car = 32 * (wave >> 2) + (lane & 31);  // a network lane's row tile and car
    const int half = lane >> 5;
    const int i = (int)blockIdx.x * SF_CARS + lane;  // the car wavefront's car
    const bf16x8* fw = reinterpret_cast<const bf16x8*>(frags) + lane;
    const uint16_t one = 0x3f80;  // bf16 1.0
    constexpr int W3R = (5 * 129 + SF_NET_THREADS - 1) / SF_NET_THREADS;
    bf16x8 f0, f1[RACER_SF_STEPS], f2[RACER_SF_STEPS];
    float w3r[W3R];  // this lane's share of the output layer's weights, stored to LDS after the first layer
    SurStart c;
    bool fell = false;  // the car wavefront's car took the exact step at least once
    if (net) {
#pragma unroll
        for (int r = 0; r < W3R; ++r) {
            const int idx = tid + SF_NET_THREADS * r;
            w3r[r] = idx < 5 * 129 ? w3[idx] : 0.0f;
        }
        f0 = fw[(size_t)rt * 64];
        // constant columns: xa k = 11 (layer 0's bias) and k = 128, xb k = 128; zeros past them
        // (four k per 8-byte store from k = 12; k = 128 is the low half of its first word)
        const int r = tid >> 3, part = tid & 7;
        if (part == 0) xa[r][11] = one;
        for (int t = part; t < (SF_LDX - 12) / 4; t += 8) {
            const int k = 12 + 4 * t;
            *reinterpret_cast<uint2*>(&xa[r][k]) = make_uint2(k == 128 ? (uint32_t)one : 0u, 0u);
        }
        for (int t = part; t < (SF_LDX - 128) / 4; t += 8) {
            const int k = 128 + 4 * t;
            *reinterpret_cast<uint2*>(&xb[r][k]) = make_uint2(k == 128 ? (uint32_t)one : 0u, 0u);
        }
    }
    auto tile = [&](const uint16_t (*x)[SF_LDX], const bf16x8* a, int steps_k, int astride_k = 1) {
        f32x16 acc;
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
#pragma unroll
        for (int k = 0; k < RACER_SF_STEPS; ++k) {
            if (k < steps_k) {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&x[car][16 * k + 8 * half]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k * astride_k], bv, acc, 0, 0, 0);
            }
        }
        return acc;
    };
    // accumulator entries 4 g .. 4 g + 3 hold rows 32 rt + 8 g + 4 half + 0 .. 3 of this lane's car
    const int row0 = 32 * rt + 4 * half;
    auto store_bf16 = [&](uint16_t (*y)[SF_LDX], const f32x16& acc) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            uint32_t v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = __builtin_bit_cast(uint16_t, (__bf16)racer_tanh_mlp(acc[4 * g + e]));
            *reinterpret_cast<uint2*>(&y[car][row0 + 8 * g]) = make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16));
        }
    };
    if constexpr (MULTI) {
        if (net) {  // both later layers' fragments into LDS, once
            for (int idx = tid; idx < 2 * SF_FRAGS; idx += SF_NET_THREADS) {
                const int layer = idx / SF_FRAGS, within = idx % SF_FRAGS;
                const int block = (layer == 0 ? RACER_SF_BASE1 : RACER_SF_BASE2) + within / 64;
                wfrag[idx] = reinterpret_cast<const bf16x8*>(frags)[(size_t)block * 64 + within % 64];
            }
        }
    }
    const int nsteps = MULTI ? steps : 1;
    for (int step = 0; step < nsteps; ++step) {
        const float* ak = act + (int64_t)step * act_stride;
        if (!net) {
            float f[11];
            if (i < n) {
                surrogate_inputs<L>(m, s, n, ak, act_cols, i, f, &c);
            } else {
#pragma unroll
                for (int r = 0; r < 11; ++r) f[r] = 0.0f;
            }
#pragma unroll
            for (int r = 0; r < 11; ++r)
                xa[lane][r] = __builtin_bit_cast(uint16_t, (__bf16)((f[r] - st[r]) * st[11 + r]));
            if (step == 0) {
                h3[lane][128] = 1.0f;
            } else {  // the second layer overwrote k 11 .. 15 of the inputs' rows
                xa[lane][11] = one;
                *reinterpret_cast<uint2*>(&xa[lane][12]) = make_uint2(0u, 0u);
            }
        }
        __syncthreads();
        if (net) {
            // the later layers' fragments only now, so that they do not queue ahead of the car wavefront's loads
            if constexpr (!MULTI) {
#pragma unroll
                for (int k = 0; k < RACER_SF_STEPS; ++k)
                    f1[k] = fw[(size_t)(RACER_SF_BASE1 + rt * RACER_SF_STEPS + k) * 64];
            }
            const f32x16 acc = tile(xa, &f0, RACER_SF_STEPS0);
            if constexpr (!MULTI) {
#pragma unroll
                for (int k = 0; k < RACER_SF_STEPS; ++k)
                    f2[k] = fw[(size_t)(RACER_SF_BASE2 + rt * RACER_SF_STEPS + k) * 64];
            }
            store_bf16(xb, acc);
            if (step == 0) {
#pragma unroll
                for (int r = 0; r < W3R; ++r) {
                    const int idx = tid + SF_NET_THREADS * r;
                    if (idx < 5 * 129) w3s[idx / 129][idx % 129] = w3r[r];
                }
            }
        }
        __syncthreads();
        const bf16x8* a1 = MULTI ? wfrag + (size_t)(rt * RACER_SF_STEPS) * 64 + lane : f1;
        const bf16x8* a2 = MULTI ? wfrag + SF_FRAGS + (size_t)(rt * RACER_SF_STEPS) * 64 + lane : f2;
        constexpr int astride = MULTI ? 64 : 1;  // fragments of consecutive k-steps
        if (net) store_bf16(xa, tile(xb, a1, RACER_SF_STEPS, astride));
        __syncthreads();
        if (net) {
            const f32x16 acc = tile(xa, a2, RACER_SF_STEPS, astride);
#pragma unroll
            for (int g = 0; g < 4; ++g)
                *reinterpret_cast<float4*>(&h3[car][row0 + 8 * g]) =
                    make_float4(racer_tanh_mlp(acc[4 * g]), racer_tanh_mlp(acc[4 * g + 1]),
                                racer_tanh_mlp(acc[4 * g + 2]), racer_tanh_mlp(acc[4 * g + 3]));
        }
        __syncthreads();
        if (wave < 5) {
            // output `wave` of car `lane`
            const float4* wv = reinterpret_cast<const float4*>(w3s[wave]);
            const float4* hv = reinterpret_cast<const float4*>(h3[lane]);
            float acc = 0.0f;
#pragma unroll
            for (int kb = 0; kb < 32; ++kb) {
                const float4 a = wv[kb], x = hv[kb];
                acc = fmaf(a.x, x.x, acc);
                acc = fmaf(a.y, x.y, acc);
                acc = fmaf(a.z, x.z, acc);
                acc = fmaf(a.w, x.w, acc);
            }
            pred[wave][lane] = fmaf(w3s[wave][128], h3[lane][128], acc);
        }
        __syncthreads();
        if (!net && i < n) {
            float q[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) q[r] = pred[r][lane];
            fell |= surrogate_apply<L>(m, s, n, ak, act_cols, reward, done, obs, obs_cols, beams, seed,
                                       heading_noise, q, c, st + 22, st + 27, min_clearance, i);
        }
    }
    if (!net && i < n && fallback) fallback[i] = fell ? 1.0f : 0.0f;
}

// Open-loop steps of the physics alone (racer/sim.mojo:184-235 `steps`
// times, the lidar not run): step k takes actions act + k act_stride; each car
// carries its state through the rows it owns.
template <int L>
__global__ void __launch_bounds__(64) k_physics_steps(racer_dmap m, float* s, int n, const float* act,
                                                      int64_t act_cols, int steps, int64_t act_stride,
                                                      float* reward, float* done, float* obs, int64_t obs_cols,
                                                      int beams, int seed, float heading_noise) {
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < steps; ++k) {
        const float* ak = act + (int64_t)k * act_stride;
        advance<L>(m, s, n, load_car(s, n, ak, act_cols, i), reward, done, obs, obs_cols, beams, i, seed,
                   heading_noise);
    }
}

// ---- float64 RK4 (the reference the surrogate is measured against) ----
// The kinematics of advance (racer/sim.mojo:184-219) in float64 with the
// true step 1/240 s: probe and bilinear surface on the float grids read as
// doubles, libm-accurate double sin, cos, tan and sqrt. No crash, progress or
// respawn: it is applied to windows without them.
__device__ void f64_probe(const racer_dmap& m, double x, double y, double z, int* layer, double* clearance,
                          double* height) {
    const double fx = floor((x - (double)m.x0) / (double)m.cell + 0.5);
    const double fy = floor((y - (double)m.y0) / (double)m.cell + 0.5);
    const bool on = fx >= 0.0 && fy >= 0.0 && fx < (double)m.nx && fy < (double)m.ny;
    const int64_t c = on ? (int64_t)fy * m.nx + (int64_t)fx : 0;
    *layer = 0;
    *clearance = -1.0;
    *height = on ? (double)grid_at(m, 0, RACER_HEIGHT, c) : -1.0;
    double gap = 1e300;
    for (int l = 0; l < m.layers && on; ++l) {
        const double cl = grid_at(m, l, RACER_CLEARANCE, c), h = grid_at(m, l, RACER_HEIGHT, c);
        if (cl > 0.0 && fabs(h - z) < gap) {
            *layer = l;
            *clearance = cl;
            *height = h;
            gap = fabs(h - z);
        }
    }
}

__device__ void f64_surface(const racer_dmap& m, int layer, double x, double y, double* sx, double* sy) {
    const double fx = (x - (double)m.x0) / (double)m.cell, fy = (y - (double)m.y0) / (double)m.cell;
    const double cx = floor(fx), cy = floor(fy);
    const int64_t ix = !(cx >= 0.0) ? 0 : (cx > (double)(m.nx - 2) ? m.nx - 2 : (int64_t)cx);
    const int64_t iy = !(cy >= 0.0) ? 0 : (cy > (double)(m.ny - 2) ? m.ny - 2 : (int64_t)cy);
    const double tx = fmin(fmax(fx - (double)ix, 0.0), 1.0), ty = fmin(fmax(fy - (double)iy, 0.0), 1.0);
    const int64_t k = iy * m.nx + ix;
    const double h0 = grid_at(m, layer, RACER_HEIGHT, k), h1 = grid_at(m, layer, RACER_HEIGHT, k + 1);
    const double h2 = grid_at(m, layer, RACER_HEIGHT, k + m.nx), h3 = grid_at(m, layer, RACER_HEIGHT, k + m.nx + 1);
    *sx = ((h1 - h0) * (1.0 - ty) + (h3 - h2) * ty) / (double)m.cell;
    *sy = ((h2 - h0) * (1.0 - tx) + (h3 - h1) * tx) / (double)m.cell;
}

__device__ void f64_derivative(const double p[4], double steer, double accel, double grip, double along,
                               double across, double out[4]) {
    const double G = RACER_G, WB = RACER_WHEELBASE;
    const double nz = 1.0 / sqrt(1.0 + along * along + across * across);
    const double load = grip * G * nz;
    const double v = fmax(fabs(p[3]), 0.5) * (p[3] >= 0.0 ? 1.0 : -1.0);
    const double slide = G * across * nz;
    const double lateral = fmin(fmax(v * p[3] * tan(steer) / WB + slide, -load), load);
    const double longitudinal = sqrt(fmax(load * load - lateral * lateral, 0.0));
    out[0] = p[3] * cos(p[2]);
    out[1] = p[3] * sin(p[2]);
    out[2] = (lateral - slide) / v;
    out[3] = fmin(fmax(accel, -longitudinal), longitudinal) - G * along * nz;
}

// st rows (doubles): x, y, z, heading, speed, steer, updated in place.
__global__ void __launch_bounds__(64) k_step_f64(racer_dmap m, const float* s, int n, double* st, const float* act,
                                                 int64_t act_cols) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double DT = 1.0 / 240.0;
    double p[4] = {st[0 * n + i], st[1 * n + i], st[3 * n + i], st[4 * n + i]};
    double steer = st[5 * n + i];
    int layer;
    double clearance, height, sx, sy;
    f64_probe(m, p[0], p[1], st[2 * n + i], &layer, &clearance, &height);
    f64_surface(m, layer, p[0], p[1], &sx, &sy);
    const double rate = fmin(fmax((double)act[i], -1.0), 1.0) * RACER_STEER_RATE;
    const double accel = fmin(fmax((double)act[act_cols + i], -1.0), 1.0) * RACER_ACCEL_MAX;
    const double grip = RACER_GRIP * (double)s[RACER_FRICTION * n + i];
    for (int sub = 0; sub < RACER_SUBSTEPS; ++sub) {
        const double c = cos(p[2]), sn = sin(p[2]);
        const double along = sx * c + sy * sn, across = -sx * sn + sy * c;
        const double half = steer + rate * DT / 2.0, full = steer + rate * DT;
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
        f64_probe(m, p[0], p[1], height, &layer, &clearance, &height);
        f64_surface(m, layer, p[0], p[1], &sx, &sy);
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

// Sim.reset's spawn kernel (racer/sim.mojo:299-304).
__global__ void __launch_bounds__(256) k_reset(racer_dmap m, float* s, float* act, int n, int seed,
                                               float heading_noise) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    spawn(m, s, n, i, seed, heading_noise);
    act[i] = 0.0f;
    act[n + i] = 0.0f;
}

// Per-cell records for the lidar: (clearance, height) of every layer side by
// side, so one probe is one 8-byte (one layer) or 16-byte (two layers) load.
// A clearance that is not positive (NaN included) is stored as -1: Map.probe
// only takes a layer whose clearance is positive and reports -1 otherwise, so
// every reader's test gives the same answer and a single-layer probe can use
// the stored value as is. Record `plane` (after the grid) is all -1, the
// reference's off-map answer, for probes that load it instead of a cell.
__global__ void k_pack(const float* d, int64_t plane, int layers, float* packed) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > plane) return;
    for (int l = 0; l < layers; ++l) {
        float cl = -1.0f, h = -1.0f;
        if (c < plane) {
            cl = d[RACER_HEADER + (int64_t)(3 * l + RACER_CLEARANCE) * plane + c];
            h = d[RACER_HEADER + (int64_t)(3 * l + RACER_HEIGHT) * plane + c];
        }
        packed[(c * layers + l) * 2 + 0] = cl > 0.0f ? cl : -1.0f;
        packed[(c * layers + l) * 2 + 1] = h;
    }
}

// The same records on the padded grid: cell (x, y) of the map, -pad <= x <
// nx + pad and -pad <= y < ny + pad, at (y + pad) pnx + x + pad, all -1 off the map.
__global__ void k_pack_padded(const float* d, int nx, int ny, int pad, int layers, float* padded) {
    const int pnx = nx + 2 * pad;
    const int64_t cells = (int64_t)pnx * (ny + 2 * pad);
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int x = (int)(c % pnx) - pad, y = (int)(c / pnx) - pad;
    const bool on = x >= 0 && x < nx && y >= 0 && y < ny;
    const int64_t plane = (int64_t)nx * ny, cell = on ? (int64_t)y * nx + x : 0;
    for (int l = 0; l < layers; ++l) {
        float cl = -1.0f, h = -1.0f;
        if (on) {
            cl = d[RACER_HEADER + (int64_t)(3 * l + RACER_CLEARANCE) * plane + cell];
            h = d[RACER_HEADER + (int64_t)(3 * l + RACER_HEIGHT) * plane + cell];
        }
        padded[(c * layers + l) * 2 + 0] = cl > 0.0f ? cl : -1.0f;
        padded[(c * layers + l) * 2 + 1] = h;
    }
}

// Map.probe on the packed records, returning (clearance, height) only. The
// records are read with 32-bit buffer offsets (the descriptor's size makes any
// out-of-range offset read zeros instead of faulting).
template <int L>
__device__ __forceinline__ void probe_packed(const racer_dmap& m, __amdgpu_buffer_rsrc_t rec, float x, float y,
                                             float z, float* clearance, float* height) {
    float fx = racer_nearest_index(x - m.x0, m.cell, m.inv_cell);
    float fy = racer_nearest_index(y - m.y0, m.cell, m.inv_cell);
    bool on = fx >= 0.0f && fy >= 0.0f && fx < (float)m.nx && fy < (float)m.ny;
    int c = on ? (int)fy * m.nx + (int)fx : 0;
    if constexpr (L == 1) {
        auto raw = __builtin_amdgcn_raw_buffer_load_b64(rec, c * 8, 0, 0);
        float2 r = on ? *reinterpret_cast<float2*>(&raw) : make_float2(-1.0f, -1.0f);
        bool take = r.x > 0.0f && fabsf(r.y - z) < 1e30f;
        *clearance = take ? r.x : -1.0f;
        *height = r.y;
    } else if constexpr (L == 2) {
        auto raw = __builtin_amdgcn_raw_buffer_load_b128(rec, c * 16, 0, 0);
        float4 r = on ? *reinterpret_cast<float4*>(&raw) : make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        float cl = -1.0f, h = r.y, gap = 1e30f;
        if (r.x > 0.0f && fabsf(r.y - z) < gap) {
            cl = r.x;
            h = r.y;
            gap = fabsf(r.y - z);
        }
        if (r.z > 0.0f && fabsf(r.w - z) < gap) {
            cl = r.z;
            h = r.w;
        }
        *clearance = cl;
        *height = h;
    } else {
        Spot s = map_probe(m, x, y, z);
        *clearance = s.clearance;
        *height = s.height;
    }
}

// (int)floorf(x) in one instruction (v_cvt_flr_i32_f32) where the compiler
// emits a floor and a conversion; racer_check_floor_index compares the two
// over every float x with |x| < 2^23 on the device. (The instruction that
// rounds x + 0.5 down in one step is no substitute for floorf(x + 0.5f): it
// adds exactly, and the float sum rounds 0.5 - 2^-25 + 0.5 up to 1.)
__device__ __forceinline__ int floor_index(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// sqrtf for x >= 2^-96 (and for +0, +inf and NaN): the hardware
// root s is within one ulp of sqrt(x), and of s - ulp, s and s + ulp the
// correctly rounded root is the one the signs of the residuals x - s' s
// select (the neighbour test of the correctly rounded float square root,
// without the rescaling that keeps it exact for subnormal x).
__device__ __forceinline__ float sqrt_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __uint_as_float(__float_as_uint(s) - 1u), hi = __uint_as_float(__float_as_uint(s) + 1u);
    const float r = fmaf(-lo, s, x) <= 0.0f ? lo : s;
    return fmaf(-hi, s, x) > 0.0f ? hi : r;
}

// The record at byte `off` of a one- or two-layer grid: the clearance and
// height of the layer Map.probe chooses for height z (sim.mojo:116-124).
template <int L>
__device__ __forceinline__ void probe_record(__amdgpu_buffer_rsrc_t rec, uint32_t off, float z, float* clearance,
                                             float* height) {
    if constexpr (L == 1) {
        // the stored clearance is -1 unless positive
        float2 r = racer_ld_f32x2(rec, off);
        // |z - height| = |height - z| exactly; this form shares the
        // subtraction with the caller's floor distance (z - height) - 0.03
        *clearance = fabsf(z - r.y) < 1e30f ? r.x : -1.0f;
        *height = r.y;
    } else {
        // Map.probe's loop over the two layers as selects (a short-circuit
        // test compiles to nested exec-mask branches on every march step)
        float4 r = __builtin_bit_cast(float4, racer_ld_b128(rec, off));
        const float d0 = fabsf(r.y - z), d1 = fabsf(r.w - z);
        const bool take0 = (r.x > 0.0f) & (d0 < 1e30f);
        const float gap = take0 ? d0 : 1e30f;
        const bool take1 = (r.z > 0.0f) & (d1 < gap);
        *clearance = take1 ? r.z : (take0 ? r.x : -1.0f);
        *height = take1 ? r.w : r.y;
    }
}

// Map.probe's (clearance, height) from the packed records with the load
// always issued: off the map it reads an all -1 record.
//
// RECIP: the nearest cell on both axes by racer_nearest_by_reciprocal1, which
// the map's quotient_ok flag has checked against the division over every cell
// boundary of the padded index range. Otherwise racer_nearest_index_lim with
// one branch for the rare division of either axis (a lane whose shortcut holds
// gets the same value from the division, which is the formula the shortcut
// reproduces); `live` = false lanes (their results unused) never take it.
//
// PAD: the probe lies on the padded grid (the caller has checked its origin),
// so its record is read there with no on-map test; else a probe off the map
// reads the all -1 record kept after the packed grid.
template <int L, bool RECIP, bool PAD>
__device__ __forceinline__ void probe_flat(const racer_dmap& m, __amdgpu_buffer_rsrc_t rec, float x, float y, float z,
                                           float limx, float limy, bool live, float* clearance, float* height) {
    const float ax = x - m.x0, ay = y - m.y0;
    uint32_t off;
    if constexpr (PAD) {
        static_assert(RECIP, "the padded grid is read with the reciprocal nearest cell");
        // racer_nearest_by_reciprocal1 with its floor and the conversion as
        // one instruction per axis (floor_index: probes on the padded grid
        // have |q + 0.5| < 2^19); the cell iy pnx + ix relative to cell (0,
        // 0), negative left of or below it, as one 24-bit multiply-add (the
        // padded grid has fewer than 2^24 cells and more than 32 rows, so pnx
        // < 2^19 and |iy pnx| < 2^24); the byte offset wraps modulo 2^32 to
        // the record's
        const float q0x = ax * m.inv_cell, q0y = ay * m.inv_cell;
        const int ix = floor_index(fmaf(fmaf(-q0x, m.cell, ax), m.inv_cell, q0x) + 0.5f);
        const int iy = floor_index(fmaf(fmaf(-q0y, m.cell, ay), m.inv_cell, q0y) + 0.5f);
        off = (uint32_t)(__mul24(iy, m.pnx) + ix) * (uint32_t)(8 * L) + m.pad_base;
    } else {
        float fx, fy;
        if constexpr (RECIP) {
            fx = racer_nearest_by_reciprocal1(ax, m.cell, m.inv_cell);
            fy = racer_nearest_by_reciprocal1(ay, m.cell, m.inv_cell);
        } else {
            // limx, limy: racer_nearest_limit of bounds on |x - x0| / cell and |y - y0| / cell
            const float sx = ax * m.inv_cell + 0.5f, sy = ay * m.inv_cell + 0.5f;
            fx = floorf(sx);
            fy = floorf(sy);
            // non-short-circuit & and |: one exec mask, not a nest of them
            const bool slow_x = !(racer_abs((sx - fx) - 0.5f) < limx), slow_y = !(racer_abs((sy - fy) - 0.5f) < limy);
            if (live & (slow_x | slow_y)) {
                fx = floorf(ax / m.cell + 0.5f);
                fy = floorf(ay / m.cell + 0.5f);
            }
        }
        // 0 <= f < n as one unsigned compare of the bit patterns: f is a floor
        // of a sum with 0.5, never -0.0, so a set sign bit means negative, and
        // NaN and +inf patterns lie above every finite n
        const bool on_x = __float_as_uint(fx) < __float_as_uint((float)m.nx);
        const bool on_y = __float_as_uint(fy) < __float_as_uint((float)m.ny);
        const bool on = on_x & on_y;
        // cell = fy nx + fx, exact in float (both below 2^12, the sum below 2^24)
        off = on ? (uint32_t)fmaf(fy, (float)m.nx, fx) * (uint32_t)(8 * L) : (uint32_t)(m.plane * 8 * L);
    }
    probe_record<L>(rec, off, z, clearance, height);
}

constexpr int LIDAR_RECORD = 5;               // dx, dy, dz, descent, RN(1 / descent)
constexpr int LIDAR_RING = 2 * LIDAR_RECORD * 64;  // two chunks of 64 beam records
constexpr int LIDAR_PIECES = 24;                    // most pieces a wavefront's share is ordered over
// a wavefront's LDS: the ring, each record's beam (16 bits, 128 of them), the pieces' iteration maxima
constexpr int LIDAR_WAVE_LDS = LIDAR_RING + 64 + LIDAR_PIECES;
constexpr int LIDAR_COOP_CARS = 8;                  // cars per workgroup of the fused step
constexpr int LIDAR_COOP_WAVES = 8;                 // wavefronts per workgroup of the cooperative lidar
constexpr size_t LIDAR_COOP_MAX_SHMEM = 80 * 1024;  // two such workgroups per CU (160 KB of LDS)

// The share of a car's beams that wavefront `global_wave` marches: the
// car's wavefronts take the shares in a rotated order, (w + w / 4) mod
// waves_per_car for the workgroup's w-th wavefront, so that the four SIMDs of
// a CU (wavefront w runs on SIMD w mod 4) each get every share equally often.
// The shares differ in work (the beams ahead of a car along a curve march
// longest), and the SIMD with most of it sets a launch's time.
__device__ __forceinline__ int lidar_share(int global_wave, int waves_per_car) {
    return (global_wave + global_wave / 4) % waves_per_car;
}

// scan (racer/sim.mojo:238-266) for every beam of one car per wavefront.
//
// Each lane marches one beam at a time; a lane whose beam ends takes the next
// unclaimed beam of the share (ballot + popcount), so the wavefront stays full
// until the share's last beams. Beam set-up (direction, normalisation, floor
// descent rate and its reciprocal) is done 64 beams at a time by the whole
// wavefront into a two-chunk LDS ring, so a refill is six LDS reads rather
// than a divergent pass of divisions and square roots. Ranges collect in LDS
// and are scaled and stored in one full-lane pass at the end.
//
// Order: a share is cut into pieces, its parts of the track-independent
// 64-beam chunks of the fan, and the pieces are set up longest first by the
// most march iterations any of their beams took in the sim's previous step
// (one byte a piece, `m.lidar_keys`, which this launch rewrites). A
// wavefront's time is set by its longest beams (up to 100+ iterations where
// the mean is 6) when they start late; started first they run beside the
// short ones. Every piece takes a whole set-up chunk: a lane past the piece's
// end gets a spare record that finishes at its first trip (it starts at t =
// NaN) and stores its range into a spare slot. The order changes which lane marches a beam and when,
// never the beam's own arithmetic, so the ranges are those of any order.
//
// lidar_wave marches wavefront `global_wave`'s share (car global_wave /
// waves_per_car); the kernels below map workgroups to wavefronts.
//
// COOP: the workgroup's cars (consecutive, from wg_car0) keep all their
// ranges in LDS ([car][beam] after the wavefronts' rings, total_beams + 1 a
// car) and the workgroup stores them together after one barrier, each obs row
// as a run of consecutive cars: partial 128-byte lines of 32 bytes (8 cars)
// instead of one float per line from each wavefront. Otherwise each wavefront
// stores its own.
//
// POSE: the car's heading, along and across slopes and position come from
// pose[8 * (car - wg_car0) ..] (LDS, written by the same workgroup's physics)
// rather than the state and obs rows.
template <int L, bool COUNT, bool COOP, bool POSE = false>
__device__ __forceinline__ void lidar_wave(const racer_dmap& m, const float* __restrict__ s, int n,
                                           float* __restrict__ obs, int64_t obs_cols, int total_beams, int per_wave,
                                           int waves_per_car, const float2* __restrict__ dirs,
                                           unsigned long long* counter, int global_wave, int wg_car0,
                                           const float* pose = nullptr) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    // wavefront-uniform (the compiler cannot see it through threadIdx.x): the
    // car, its beam share and the refill bookkeeping then live in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    global_wave = __builtin_amdgcn_readfirstlane(global_wave);
    const int car = global_wave / waves_per_car;
    const int share = lidar_share(global_wave, waves_per_car);
    // This wavefront's share of the car's beams: [lo, lo + beams).
    const int lo = share * per_wave;
    const int beams = min(total_beams, lo + per_wave) - lo;
    const bool active = car < n && beams > 0;
    if constexpr (!COOP) {
        if (!active) return;
    }
    const int wg_waves = (int)(blockDim.x >> 6);
    float* ring = COOP ? lds + wave * LIDAR_WAVE_LDS : lds + wave * (LIDAR_WAVE_LDS + per_wave + 1);
    uint16_t* ring_beam = reinterpret_cast<uint16_t*>(ring + LIDAR_RING);
    uint32_t* piece_max = reinterpret_cast<uint32_t*>(ring + LIDAR_RING + 64);
    float* ranges = COOP ? lds + wg_waves * LIDAR_WAVE_LDS + (int64_t)(car - wg_car0) * (total_beams + 1) + lo
                         : ring + LIDAR_WAVE_LDS;
    // the spare range slot of lanes past a piece's end
    const int spare = COOP ? total_beams - lo : per_wave;
    if (active) {
    __amdgpu_buffer_rsrc_t rec = __builtin_amdgcn_make_buffer_rsrc(
        (void*)m.packed, (short)0, (int)((m.plane + 1) * (L > 0 ? L : 1) * 8), 0x00020000);
    __amdgpu_buffer_rsrc_t rec_pad =
        __builtin_amdgcn_make_buffer_rsrc((void*)m.padded, (short)0, (int)m.pad_bytes, 0x00020000);

    // Pieces: chunk k0 + p of the fan cut to [lo, lo + beams), p < pieces; the
    // keys of car `car` hold piece p of share `share` at byte k0 + share + p.
    const int k0 = lo >> 6;
    const int pieces = ((lo + beams - 1) >> 6) - k0 + 1;
    const bool ordered = m.lidar_keys != nullptr && pieces <= LIDAR_PIECES && k0 + share + pieces <= RACER_LIDAR_KEYS;
    uint8_t* keys = ordered ? m.lidar_keys + (int64_t)car * RACER_LIDAR_KEYS + k0 + share : nullptr;
    // rank of piece `lane` in the order: longer first, then by index
    int rank = lane;
    if (ordered) {
        const uint32_t key = lane < pieces ? keys[lane] : 0u;
        if (lane < pieces) {
            rank = 0;
            for (int q = 0; q < pieces; ++q) {
                const uint32_t kq = __builtin_amdgcn_readlane(key, q);
                rank += (int)((kq > key) | ((kq == key) & (q < lane)));
            }
        }
        if (lane < LIDAR_PIECES) piece_max[lane] = 0u;
        // Issue priority by the longest piece's key: a wavefront whose beams
        // march longest sets the launch's end unless the SIMD serves it first,
        // while the short ones fill the issue slots it leaves.
        const int top = __builtin_amdgcn_readfirstlane(
            (int)__builtin_amdgcn_readlane(key, (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(rank == 0))));
        if (top >= 40)
            __builtin_amdgcn_s_setprio(3);
        else if (top >= 20)
            __builtin_amdgcn_s_setprio(2);
        else if (top >= 10)
            __builtin_amdgcn_s_setprio(1);
    }

    const float* cp = POSE ? pose + 8 * (car - wg_car0) : nullptr;
    const float heading = POSE ? cp[0] : s[RACER_HEADING * n + car];
    const float along = POSE ? cp[1] : obs[2 * obs_cols + car];
    const float across = POSE ? cp[2] : obs[3 * obs_cols + car];
    const float ch = racer_cos(heading), sh = racer_sin(heading);
    const float ox = ((POSE ? cp[3] : s[RACER_X * n + car]) + RACER_MOUNT_FORWARD * ch) + 0.0f;
    const float oy = ((POSE ? cp[4] : s[RACER_Y * n + car]) + RACER_MOUNT_FORWARD * sh) + 0.0f;
    const float oz = ((POSE ? cp[5] : s[RACER_Z * n + car]) + RACER_MOUNT_FORWARD * along) + RACER_MOUNT_UP;
    const float slope15 = 1.5f * m.slope;
    const float half_cell = 0.5f * m.cell;
    // Every probe of the march lies within RACER_RANGE of the origin (t < RANGE,
    // |direction| <= 1), so |x - x0| / cell stays below Qx (and likewise y):
    // one nearest-cell limit per axis for the whole wavefront.
    const float qx = (fabsf(ox - m.x0) + (RACER_RANGE + 1.0f)) * m.inv_cell * 1.001f;
    const float qy = (fabsf(oy - m.y0) + (RACER_RANGE + 1.0f)) * m.inv_cell * 1.001f;
    const float limx = racer_nearest_limit(qx), limy = racer_nearest_limit(qy);
    // The march step's quotient fl / descent comes from the beam's RN(1 /
    // descent) (racer_div_by_reciprocal, correctly rounded for a divisor in
    // [2^-60, 2^60] and a dividend of +0 or magnitude in [2^-60, 2^60]), and
    // the probes' nearest cells from RN(1 / cell) (probe_flat), when this
    // car's numbers keep them in range. The map's flag holds the cell and the
    // slope within [2^-20, 2^20] and its heights within 2^40; with |oz| <=
    // 2^40 that bounds |fl| = |(pz - height) - 0.03| by 2^42, and fl is +0 or
    // at least 2^-30 (a float minus 0.03f is a multiple of 2^-30 unless the
    // float is below 2^-7; fl = -0 would need 0.03f = +0); |along| + |across|
    // <= 2^20 keeps the horizontal part of each beam above 2^-20.1, so
    // descent >= 1.5 slope 2^-20.1 >= 2^-41, and descent <= 1.5 slope + 1 <=
    // 2^21. Otherwise the step divides.
    const bool fast_q =
        __builtin_amdgcn_readfirstlane((int)(m.quotient_ok != 0) & (int)(fabsf(oz) <= 0x1p40f) &
                                       (int)(fabsf(along) + fabsf(across) <= 0x1p20f)) != 0;
    // An origin inside the padded grid's box keeps every probe on that grid
    // (the box leaves RACER_RANGE and a margin of cells to its edges).
    const bool pad_q = fast_q && m.padded != nullptr &&
                       __builtin_amdgcn_readfirstlane((int)(ox >= m.pad_x0) & (int)(ox <= m.pad_x1) &
                                                      (int)(oy >= m.pad_y0) & (int)(oy <= m.pad_y1)) != 0;

    // Set-up of the c-th piece in the order, one beam per lane, into ring half
    // c % 2: field f of record r lives at ring[128 f + r % 128], its beam (relative
    // to lo, or the spare slot) at ring_beam[r % 128]. For the reciprocal march
    // (fast_q) the two square roots are sqrt_normal: with |along| + |across| <=
    // 2^20 the squared norm lies in [0.99, 2^40 + 1] ((d0, d1) is the unit fan
    // direction turned by the heading) and the squared horizontal part in
    // [2^-41, 1], no subnormals there.
    auto setup_chunk = [&](int c, auto reciprocal) {
        const int piece = ordered ? (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(rank == c)) : c;
        const int k = k0 + piece;
        const int b = k * 64 + lane - lo;  // this lane's beam, relative to lo
        float* slot = ring + (c & 1) * 64 + lane;
        if ((b >= 0) & (b < beams)) {
            float2 cs = dirs[lo + b];
            float d0 = cs.x * ch + cs.y * (-sh);
            float d1 = cs.x * sh + cs.y * ch;
            float d2 = cs.x * along + cs.y * across;
            const float n2 = (d0 * d0 + d2 * d2) + (d1 * d1 + 0.0f);
            float norm = decltype(reciprocal)::value ? sqrt_normal(n2) : sqrtf(n2);
            float dx = d0 / norm, dy = d1 / norm, dz = d2 / norm;
            const float h2 = dx * dx + dy * dy;
            float horizontal = decltype(reciprocal)::value ? sqrt_normal(h2) : sqrtf(h2);
            float descent = slope15 * horizontal + racer_max(-dz, 0.0f);
            slot[0 * 128] = dx;
            slot[1 * 128] = dy;
            slot[2 * 128] = dz;
            slot[3 * 128] = descent;
            slot[4 * 128] = decltype(reciprocal)::value ? 1.0f / descent : 0.0f;
            ring_beam[(c & 1) * 64 + lane] = (uint16_t)b;
        } else {
            // a spare record: its NaN descent starts the march at t = NaN
            slot[0 * 128] = 0.0f;
            slot[1 * 128] = 0.0f;
            slot[2 * 128] = 0.0f;
            slot[3 * 128] = __builtin_nanf("");
            slot[4 * 128] = 0.0f;
            ring_beam[(c & 1) * 64 + lane] = (uint16_t)spare;
        }
    };
    float dx = 0.0f, dy = 0.0f, dz = 0.0f, descent = 0.0f, inv_descent = 0.0f;
    float t = 0.0f, last = 0.0f, last_floor = 0.0f;
    int beam = spare;   // the beam this lane marches (relative to lo)
    int started = 0;    // the trip it started on
    auto take = [&](int r) {
        const float* slot = ring + (r & 127);
        dx = slot[0 * 128];
        dy = slot[1 * 128];
        dz = slot[2 * 128];
        descent = slot[3 * 128];
        inv_descent = slot[4 * 128];
        beam = ring_beam[r & 127];
        // +0 for a beam (its descent is a number), NaN for a spare record,
        // which !(t < RANGE) then ends at its first trip; one multiply, as
        // many instructions as setting t = 0
        t = fabsf(descent) * 0.0f;
        last = 0.0f;
        last_floor = 0.0f;
    };
    // records: 64 per piece, pieces * 64 in all
    const int records = pieces * 64;
    int b = lane;  // this lane's record
    uint64_t iterations = 0, trips = 0;
    // The march step below runs on every lane with selects in place of
    // branches (lanes past the share's end compute on their last beam's
    // registers and finish nothing): each divergent branch costs the
    // wavefront exec-mask bookkeeping, as much issue time as a VALU
    // instruction, on every trip. The step's quotient is always divided: some
    // live lane needs the exact division on almost every trip (the floor
    // bound is the smaller one on about one iteration in nine), so a
    // shortcut around it only adds its test. The loop is compiled three
    // times, for the reciprocal form of that quotient on the padded grid and
    // on the packed grid, and for the division (fast_q, pad_q and ordered are
    // uniform; an unordered share, one of more pieces than the order covers,
    // takes the dividing loop), so the choice costs nothing per trip.
    auto march = [&](auto reciprocal, auto padded) {
        // declared here rather than captured: the compiler then keeps them in
        // scalar registers in every copy of the loop
        int chunks = pieces > 1 ? 2 : 1;  // records below 64 * chunks are set up in the ring
        int next = 64;                    // the next unclaimed record
        int trip = 0;
        // the claim that needs the next piece set up (none once all are):
        // one scalar compare a trip
        int refill_at = chunks < pieces ? 64 * chunks : 0x7fffffff;
        // the reciprocal loops run for ordered shares only (the caller sends
        // the rest to the dividing loop), so their key updates need no test
        const bool record = decltype(reciprocal)::value || ordered;
        setup_chunk(0, reciprocal);
        if (pieces > 1) setup_chunk(1, reciprocal);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        take(b);
        while (__ballot(b < records) != 0ull) {
            const bool live = b < records;
            if (COUNT) iterations += (live & (beam != spare)) ? 1 : 0;
            const float px = ox + t * dx, py = oy + t * dy, pz = oz + t * dz;
            float clearance, height;
            if constexpr (L == 1 || L == 2)
                probe_flat<L, decltype(reciprocal)::value, decltype(padded)::value>(
                    m, decltype(padded)::value ? rec_pad : rec, px, py, pz, limx, limy, live, &clearance, &height);
            else
                probe_packed<L>(m, rec, px, py, pz, &clearance, &height);
            const float fl = (pz - height) - 0.03f;
            float q;
            if constexpr (decltype(reciprocal)::value)
                q = racer_div_by_reciprocal(fl, descent, inv_descent);
            else
                q = fl / descent;
            const float step = racer_min(clearance - m.cell, q);
            // fl < 0: the floor is hit between last and t; else step < cell / 2:
            // stop at t; else advance, leaving the range at t + step >= RANGE
            const bool hit = fl < 0.0f;
            const bool advance = !hit & !(step < half_cell);
            const float next_t = t + step;
            last = advance ? t : last;
            last_floor = advance ? fl : last_floor;
            t = advance ? next_t : t;
            // an empty asm: the finished lanes' range below then reads these
            // new values rather than the old ones the compiler knows equal
            // them there, which kept both alive and cost three copies a trip
            __asm__ volatile("" : "+v"(t), "+v"(last), "+v"(last_floor));
            const bool finished = live & (!advance | !(t < RACER_RANGE));
            ++trip;
            if (COUNT && lane == 0) ++trips;
            // the builtin on the bool itself: __ballot's int argument costs a
            // select and a compare to turn the lane mask back into one
            const uint64_t done_mask = __builtin_amdgcn_ballot_w64(finished);
            if (done_mask != 0ull) {
                const int count = (int)__popcll(done_mask);
                // The ring slot reused here held chunk `chunks - 2`, whose records
                // are all below `next` and already taken.
                if (next + count > refill_at) {
                    setup_chunk(chunks, reciprocal);
                    ++chunks;
                    refill_at = chunks < pieces ? 64 * chunks : 0x7fffffff;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                // finished lanes below this one
                const int claimed = next + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(done_mask >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)done_mask, 0u));
                // one exec region: the range (interpolated between the last two
                // steps on a floor hit), its store, the piece's iteration
                // maximum, then the next record (a lane whose claim is past the
                // share reads a stale ring slot and stays idle, since b >= records)
                if (finished) {
                    ranges[beam] = hit ? racer_min(last + ((t - last) * last_floor) / (last_floor - fl), RACER_RANGE)
                                       : racer_min(t, RACER_RANGE);
                    if (record) __hip_atomic_fetch_max(piece_max + (b >> 6), (uint32_t)(trip - started),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    b = claimed;
                    started = trip;
                    take(claimed);
                }
                next += count;
            }
        }
    };
    if (pad_q && ordered)
        march(std::integral_constant<bool, true>{}, std::integral_constant<bool, true>{});
    else if (fast_q && ordered)
        march(std::integral_constant<bool, true>{}, std::integral_constant<bool, false>{});
    else
        march(std::integral_constant<bool, false>{}, std::integral_constant<bool, false>{});
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // piece p's key for the next launch: the most iterations of its beams
    // (trips from a beam's take to its end), saturated to a byte
    if (ordered && lane < pieces) keys[lane] = (uint8_t)min(piece_max[rank], 255u);
    if constexpr (!COOP) {
        for (int j = lane; j < beams; j += 64) obs[(int64_t)(4 + lo + j) * obs_cols + car] = ranges[j] / RACER_RANGE;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) iterations += __shfl_xor(iterations, off);
        if (lane == 0) {
            atomicAdd(counter, (unsigned long long)iterations);
            atomicAdd(counter + 1, (unsigned long long)trips);
        }
    }
    }  // active
    if constexpr (COOP) {
        __syncthreads();
        const int cars = wg_waves / waves_per_car;
        const float* all = lds + wg_waves * LIDAR_WAVE_LDS;
        for (int idx = (int)threadIdx.x; idx < cars * total_beams; idx += (int)blockDim.x) {
            const int c = idx % cars, j = idx / cars;
            if (wg_car0 + c < n)
                obs[(int64_t)(4 + j) * obs_cols + wg_car0 + c] = all[c * (total_beams + 1) + j] / RACER_RANGE;
        }
    }
}

template <int L, bool COUNT, bool COOP>
__global__ void __launch_bounds__(COOP ? 1024 : 256) k_lidar(racer_dmap m, const float* __restrict__ s, int n,
                                                             float* __restrict__ obs, int64_t obs_cols,
                                                             int total_beams, int per_wave, int waves_per_car,
                                                             const float2* __restrict__ dirs,
                                                             unsigned long long* counter) {
    const int block = xcd_block();
    const int global_wave = (int)(((int64_t)block * blockDim.x + threadIdx.x) >> 6);
    const int wg_car0 = (int)((int64_t)block * (blockDim.x >> 6) / waves_per_car);
    lidar_wave<L, COUNT, COOP>(m, s, n, obs, obs_cols, total_beams, per_wave, waves_per_car, dirs, counter,
                               global_wave, wg_car0);
}

// One whole step (racer/sim.mojo:314-321, physics then lidar) in one launch,
// on the cooperative lidar's workgroups of LIDAR_COOP_CARS cars.
//
// Each car's first wavefront loads a PHYS_TILE x PHYS_TILE window of records
// and nearest-waypoint cells around the car into LDS (one round of loads;
// cells off the map are left unset and never read) while lanes 0..7 of
// wavefront 0 load the cars' inputs; after a barrier those lanes run the
// physics of the workgroup's cars, whose substep probes read the windows (a
// car moves at most 4 * 5 m/s / 240 Hz = 0.083 m a step, 3.4 cells of
// 0.025 m, inside the window's four cells either side; a probe outside it
// reads the track array). One wavefront rather than one lane in each car's
// wavefront: the serial chain is issue-bound when many wavefronts of a SIMD
// run it at once. The physics hands the new pose to the lidar through LDS;
// after a second barrier every wavefront marches its share of the beams, over
// LDS that the windows no longer need. Results equal k_physics then k_lidar.
//
// LDS: [0, pose_at): the windows (3 * PHYS_TILE^2 * L floats a car), then the
// lidar's rings and ranges; [pose_at, pose_at + 8 * LIDAR_COOP_CARS): poses.
template <int L>
constexpr int phys_tile_floats() {
    return 3 * PHYS_TILE * PHYS_TILE * L;
}

// (eight wavefronts per SIMD, the two workgroups a CU holds: the physics
// alone would take more scalar registers than that allows)
//
// step_block runs workgroup `block` of one sim's grid (cars 8 block ..);
// k_step_fused launches one workgroup per block, k_step_fused_batch the blocks
// of several sims.
template <int L, bool COUNT>
__device__ __forceinline__ void step_block(const racer_dmap& m, float* s, int n, const float* act, float* reward,
                                           float* done, float* obs, int64_t obs_cols, int total_beams, int per_wave,
                                           int waves_per_car, const float2* __restrict__ dirs,
                                           unsigned long long* counter, int seed, float heading_noise, int pose_at,
                                           int block) {
    extern __shared__ float lds[];
    constexpr int T = PHYS_TILE, CELLS = T * T * L;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wg_waves = (int)(blockDim.x >> 6);
    const int global_wave = __builtin_amdgcn_readfirstlane((int)((int64_t)block * wg_waves + wave));
    const int wg_car0 = (int)((int64_t)block * wg_waves / waves_per_car);
    const int car = global_wave / waves_per_car;
    float* pose = lds + pose_at;
    // the window's first cell: four cells before the car's cell on each axis
    auto window = [&](float x, float y, int* cx0, int* cy0) {
        *cx0 = (int)fminf(fmaxf(floorf((x - m.x0) / m.cell), -0x1p20f), 0x1p20f) - 4;
        *cy0 = (int)fminf(fmaxf(floorf((y - m.y0) / m.cell), -0x1p20f), 0x1p20f) - 4;
    };
    // wavefront 0, lane c < the workgroup's cars: car wg_car0 + c's inputs
    const int mine = wg_car0 + lane;
    const bool physics = wave == 0 && lane < LIDAR_COOP_CARS && mine < n;
    CarIn in{};
    if (physics) in = load_car(s, n, act, n, mine);
    if (global_wave % waves_per_car == 0 && car < n) {
        const int slot = car - wg_car0;
        float2* trec = reinterpret_cast<float2*>(lds + slot * phys_tile_floats<L>());
        float* tnear = lds + slot * phys_tile_floats<L>() + 2 * CELLS;
        int cx0, cy0;
        window(s[RACER_X * n + car], s[RACER_Y * n + car], &cx0, &cy0);
        const float2* rec = reinterpret_cast<const float2*>(m.packed);
#pragma unroll
        for (int r = 0; r < (CELLS + 63) / 64; ++r) {
            const int j = r * 64 + lane;
            const int cell = j / L, l = j - cell * L;
            const int cx = cx0 + cell % T, cy = cy0 + cell / T;
            if (j < CELLS && (unsigned)cx < (unsigned)m.nx && (unsigned)cy < (unsigned)m.ny)
                trec[j] = rec[(cy * m.nx + cx) * L + l];
        }
#pragma unroll
        for (int r = 0; r < (CELLS + 63) / 64; ++r) {
            const int j = r * 64 + lane;
            const int l = j / (T * T), cell = j - l * T * T;
            const int cx = cx0 + cell % T, cy = cy0 + cell / T;
            if (j < CELLS && (unsigned)cx < (unsigned)m.nx && (unsigned)cy < (unsigned)m.ny)
                tnear[j] = grid_at(m, l, RACER_NEAREST, cy * m.nx + cx);
        }
    }
    __syncthreads();
    if (physics) {
        PhysTile tile;
        tile.rec = reinterpret_cast<const float2*>(lds + lane * phys_tile_floats<L>());
        tile.near = lds + lane * phys_tile_floats<L>() + 2 * CELLS;
        window(in.x, in.y, &tile.cx0, &tile.cy0);
        advance<L, true>(m, s, n, in, reward, done, obs, obs_cols, total_beams, mine, seed, heading_noise, &tile,
                         pose + 8 * lane);
    }
    __syncthreads();
    lidar_wave<L, COUNT, true, true>(m, s, n, obs, obs_cols, total_beams, per_wave, waves_per_car, dirs, counter,
                                     global_wave, wg_car0, pose);
}

template <int L, bool COUNT>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) k_step_fused(racer_dmap m, float* s, int n, const float* act,
                                                     float* reward, float* done, float* obs, int64_t obs_cols,
                                                     int total_beams, int per_wave, int waves_per_car,
                                                     const float2* __restrict__ dirs, unsigned long long* counter,
                                                     int seed, float heading_noise, int pose_at) {
    step_block<L, COUNT>(m, s, n, act, reward, done, obs, obs_cols, total_beams, per_wave, waves_per_car, dirs,
                         counter, seed, heading_noise, pose_at, xcd_block());
}

// The fused step of several sims (one or two track layers each) in one
// launch. tab.start holds the prefix sums of the instances' workgroups of
// LIDAR_COOP_CARS cars, so a workgroup never spans two sims. With PERSIST the
// grid is what the GPU holds at once and each workgroup runs blocks b, b +
// gridDim.x, ... of a static order in which XCD x takes a contiguous run of
// every gridDim.x blocks (no queue: every workgroup knows its blocks from its
// index); otherwise one workgroup per block in the same XCD-grouped order.
template <bool COUNT, bool PERSIST>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
k_step_fused_batch(racer_sim_table tab, int per_wave, int waves_per_car, unsigned long long* counter, int pose_at) {
    const int64_t total = tab.start[tab.count];
    // the persistent order: round r covers blocks [G r, G r + G) (G = gridDim.x),
    // workgroup w's block in it the (w / 8)-th of XCD w % 8's contiguous share;
    // blocks grow with the round, so a workgroup past the end is done
    const int G = gridDim.x;
    const int64_t slot = PERSIST ? ((G & 7) == 0 ? (int64_t)(blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3)
                                                 : (int64_t)blockIdx.x)
                                 : (int64_t)xcd_block();
    for (int64_t r0 = 0; r0 < total; r0 += PERSIST ? G : total) {
        const int64_t b = r0 + slot;
        if (b >= total) break;
        int k = 0;
        while (k + 1 < tab.count && b >= tab.start[k + 1]) ++k;
        const racer_sim_slot& q = tab.slot[k];
        const int local = (int)(b - tab.start[k]);
        if (PERSIST && r0 > 0) __syncthreads();  // the previous block's range stores read LDS
        if (q.map.layers == 1)
            step_block<1, COUNT>(q.map, q.state, q.n, q.actions, q.reward, q.done, q.obs, q.n, q.beams, per_wave,
                                 waves_per_car, q.dirs, counter, q.seed, q.heading_noise, pose_at, local);
        else
            step_block<2, COUNT>(q.map, q.state, q.n, q.actions, q.reward, q.done, q.obs, q.n, q.beams, per_wave,
                                 waves_per_car, q.dirs, counter, q.seed, q.heading_noise, pose_at, local);
    }
}

// The lidar of several sims in one launch: wavefront g belongs to the
// instance whose range [start[k], start[k + 1]) of wavefronts holds g.
template <bool COUNT>
__global__ void __launch_bounds__(256) k_lidar_batch(racer_sim_table tab, int per_wave, int waves_per_car,
                                                     unsigned long long* counter) {
    // wavefront-uniform, so the instance lookup below is scalar
    const int64_t g = __builtin_amdgcn_readfirstlane((int)(((int64_t)xcd_block() * blockDim.x + threadIdx.x) >> 6));
    int k = 0;
    while (k + 1 < tab.count && g >= tab.start[k + 1]) ++k;
    if (g >= tab.start[tab.count]) return;
    const racer_sim_slot& q = tab.slot[k];
    const int w = (int)(g - tab.start[k]);
    if (q.map.layers == 1)
        lidar_wave<1, COUNT, false>(q.map, q.state, q.n, q.obs, q.n, q.beams, per_wave, waves_per_car, q.dirs, counter,
                                     w, 0);
    else if (q.map.layers == 2)
        lidar_wave<2, COUNT, false>(q.map, q.state, q.n, q.obs, q.n, q.beams, per_wave, waves_per_car, q.dirs, counter,
                                     w, 0);
    else
        lidar_wave<0, COUNT, false>(q.map, q.state, q.n, q.obs, q.n, q.beams, per_wave, waves_per_car, q.dirs, counter,
                                     w, 0);
}

// Physics of several sims in one launch (one thread per car of any instance).
__global__ void __launch_bounds__(64) k_physics_batch(racer_sim_table tab) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = 0;
    while (k + 1 < tab.count && g >= tab.start[k + 1]) ++k;
    if (g >= tab.start[tab.count]) return;
    const racer_sim_slot& q = tab.slot[k];
    const int i = (int)(g - tab.start[k]);
    if (q.map.layers == 1)
        advance<1>(q.map, q.state, q.n, load_car(q.state, q.n, q.actions, q.n, i), q.reward, q.done, q.obs, q.n,
                   q.beams, i, q.seed, q.heading_noise);
    else if (q.map.layers == 2)
        advance<2>(q.map, q.state, q.n, load_car(q.state, q.n, q.actions, q.n, i), q.reward, q.done, q.obs, q.n,
                   q.beams, i, q.seed, q.heading_noise);
    else
        advance<0>(q.map, q.state, q.n, load_car(q.state, q.n, q.actions, q.n, i), q.reward, q.done, q.obs, q.n,
                   q.beams, i, q.seed, q.heading_noise);
}

// The reference's own lidar launch (racer/sim.mojo:317-321 through the 2-D
// Device.run of racer/device.mojo:67-79): one thread per (car, beam), beams
// fastest, every thread recomputing its car's frame and its beam's azimuth,
// and every probe's nearest cell by Map.at's division. The same results as
// k_lidar; kept as the schedule of the plain baseline.
__global__ void __launch_bounds__(256) k_lidar_flat(racer_dmap m, const float* __restrict__ s, int n,
                                                    float* __restrict__ obs, int64_t obs_cols, int beams) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * beams) return;
    const int car = (int)(idx / beams), b = (int)(idx % beams);
    const float heading = s[RACER_HEADING * n + car];
    const float along = obs[2 * obs_cols + car];
    const float across = obs[3 * obs_cols + car];
    const float ch = racer_cos(heading), sh = racer_sin(heading);
    const float az = -RACER_FOV / 2.0f + (float)b * (RACER_FOV / (float)(beams - 1));
    const float ca = racer_cos(az), sa = racer_sin(az);
    float d0 = ca * ch + sa * (-sh);
    float d1 = ca * sh + sa * ch;
    float d2 = ca * along + sa * across;
    float norm = sqrtf((d0 * d0 + d2 * d2) + (d1 * d1 + 0.0f));
    const float dx = d0 / norm, dy = d1 / norm, dz = d2 / norm;
    const float ox = (s[RACER_X * n + car] + RACER_MOUNT_FORWARD * ch) + 0.0f;
    const float oy = (s[RACER_Y * n + car] + RACER_MOUNT_FORWARD * sh) + 0.0f;
    const float oz = (s[RACER_Z * n + car] + RACER_MOUNT_FORWARD * along) + RACER_MOUNT_UP;
    const float horizontal = sqrtf(dx * dx + dy * dy);
    const float descent = (1.5f * m.slope) * horizontal + racer_max(-dz, 0.0f);
    float t = 0.0f, last = 0.0f, last_floor = 0.0f, range = 0.0f;
    bool done = false;
    while (t < RACER_RANGE) {
        float pz = oz + t * dz;
        Spot spot = map_probe<true>(m, ox + t * dx, oy + t * dy, pz);
        float fl = (pz - spot.height) - 0.03f;
        if (fl < 0.0f) {
            range = racer_min(last + ((t - last) * last_floor) / (last_floor - fl), RACER_RANGE);
            done = true;
            break;
        }
        float step = racer_min(spot.clearance - m.cell, fl / descent);
        if (step < 0.5f * m.cell) break;
        last = t;
        last_floor = fl;
        t += step;
    }
    if (!done) range = racer_min(t, RACER_RANGE);
    obs[(int64_t)(4 + b) * obs_cols + car] = range / RACER_RANGE;
}

// Sim.sample's kernel (racer/sim.mojo:337-348).
__global__ void __launch_bounds__(256) k_sample(int n, const float* out, int64_t out_cols, const float* log_std,
                                                float* act, int64_t act_cols, float* logp, float* value,
                                                int deterministic, int64_t step, const int64_t* step_base,
                                                int seed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (step_base) step += *step_base;
    float noise[4];
    racer_normal4(racer_random((uint64_t)(int64_t)seed + 1, (uint64_t)i, (uint64_t)step), noise);
    float probability = 0.0f;
    for (int a = 0; a < 2; ++a) {
        float z = deterministic ? 0.0f : noise[a];
        float ls = log_std[a];
        act[a * act_cols + i] = out[a * out_cols + i] + racer_exp(ls) * z;
        probability = probability - (((0.5f * z) * z + ls) + RACER_HALF_LOG_2PI);
    }
    logp[i] = probability;
    value[i] = out[2 * out_cols + i];
}

}  // namespace

// ---------------------------------------------------------------- launchers
int racer_launch_reset(racer_sim* sim) {
    int n = sim->n;
    {
        racer_timed timer(sim->dev, "reset");
        hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, sim->dev->stream, sim->map,
                           sim->state->ptr, sim->actions->ptr, n, sim->seed, sim->heading_noise);
    }
    return racer_launch_check("reset");
}

template <int L>
static void launch_lidar(racer_sim* sim, float* obs, int64_t obs_cols) {
    int n = sim->n;
    // A car's beams are split over waves_per_car wavefronts; each wavefront
    // holds its beam ring and its share of the ranges in LDS.
    int wpc = sim->waves_per_car;
    int per = (sim->beams + wpc - 1) / wpc;
    int64_t total = (int64_t)n * wpc;
    // cooperative stores from workgroups of 8 wavefronts (8 / wpc cars, runs of
    // 4 (8 / wpc) bytes) when the cars' ranges fit their LDS; else, or when the
    // sim asks for it, each wavefront stores its own. (Workgroups of 16
    // wavefronts measured the same, every workgroup being resident either way.)
    const int coop_waves = LIDAR_COOP_WAVES;
    const int coop_cars = coop_waves / wpc;
    const size_t coop_shmem =
        ((size_t)coop_waves * LIDAR_WAVE_LDS + (size_t)coop_cars * (sim->beams + 1)) * sizeof(float);
    if (sim->lidar_coop && coop_cars >= 1 && coop_shmem <= LIDAR_COOP_MAX_SHMEM) {
        dim3 grid((unsigned)((total + coop_waves - 1) / coop_waves)), block(64 * coop_waves);
        if (sim->count_iterations)
            hipLaunchKernelGGL((k_lidar<L, true, true>), grid, block, coop_shmem, sim->dev->stream, sim->map,
                               sim->state->ptr, n, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter);
        else
            hipLaunchKernelGGL((k_lidar<L, false, true>), grid, block, coop_shmem, sim->dev->stream, sim->map,
                               sim->state->ptr, n, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter);
        return;
    }
    size_t bytes_per_wave = (size_t)(LIDAR_WAVE_LDS + per + 1) * sizeof(float);
    int waves = 4 * bytes_per_wave <= 65536 ? 4 : 1;
    dim3 grid((unsigned)((total + waves - 1) / waves)), block(64 * waves);
    size_t shmem = waves * bytes_per_wave;
    if (sim->count_iterations)
        hipLaunchKernelGGL((k_lidar<L, true, false>), grid, block, shmem, sim->dev->stream, sim->map, sim->state->ptr,
                           n, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter);
    else
        hipLaunchKernelGGL((k_lidar<L, false, false>), grid, block, shmem, sim->dev->stream, sim->map,
                           sim->state->ptr, n, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter);
}

// The cooperative kernels may use more than the default 64 KB of dynamic LDS.
static int lidar_coop_setup() {
    const void* kernels[] = {(const void*)k_lidar<0, false, true>, (const void*)k_lidar<0, true, true>,
                             (const void*)k_lidar<1, false, true>, (const void*)k_lidar<1, true, true>,
                             (const void*)k_lidar<2, false, true>, (const void*)k_lidar<2, true, true>,
                             (const void*)k_step_fused<1, false>,  (const void*)k_step_fused<1, true>,
                             (const void*)k_step_fused<2, false>,  (const void*)k_step_fused<2, true>,
                             (const void*)k_step_fused_batch<false, false>,
                             (const void*)k_step_fused_batch<true, false>,
                             (const void*)k_step_fused_batch<false, true>,
                             (const void*)k_step_fused_batch<true, true>};
    for (const void* k : kernels)
        RACER_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LIDAR_COOP_MAX_SHMEM));
    return 0;
}

static void launch_lidar_any(racer_sim* sim, float* obs, int64_t obs_cols) {
    if (sim->map.layers == 1)
        launch_lidar<1>(sim, obs, obs_cols);
    else if (sim->map.layers == 2)
        launch_lidar<2>(sim, obs, obs_cols);
    else
        launch_lidar<0>(sim, obs, obs_cols);
}

// LDS of the fused step (k_step_fused) for this sim, or 0 when it does not
// apply: the lidar's default schedule with cooperative workgroups, and
// packed one- or two-layer records.
static size_t step_fused_shmem(const racer_sim* sim, int* pose_at) {
    const int layers = sim->map.layers;
    const int coop_waves = LIDAR_COOP_CARS * sim->waves_per_car;
    if (sim->schedule != 0 || (layers != 1 && layers != 2) || coop_waves > 16) return 0;
    const size_t lidar = (size_t)coop_waves * LIDAR_WAVE_LDS + (size_t)LIDAR_COOP_CARS * (sim->beams + 1);
    const size_t tiles = (size_t)LIDAR_COOP_CARS * (layers == 1 ? phys_tile_floats<1>() : phys_tile_floats<2>());
    const size_t at = lidar > tiles ? lidar : tiles;
    const size_t bytes = (at + 8 * LIDAR_COOP_CARS) * sizeof(float);
    if (bytes > LIDAR_COOP_MAX_SHMEM) return 0;
    *pose_at = (int)at;
    return bytes;
}

template <int L>
static void launch_step_fused(racer_sim* sim, float* actions, float* reward, float* done, float* obs,
                              int64_t obs_cols, size_t shmem, int pose_at) {
    const int wpc = sim->waves_per_car, per = (sim->beams + wpc - 1) / wpc;
    const int coop_waves = LIDAR_COOP_CARS * wpc;
    const int64_t total = (int64_t)sim->n * wpc;
    dim3 grid((unsigned)((total + coop_waves - 1) / coop_waves)), block(64 * coop_waves);
    if (sim->count_iterations)
        hipLaunchKernelGGL((k_step_fused<L, true>), grid, block, shmem, sim->dev->stream, sim->map, sim->state->ptr,
                           sim->n, actions, reward, done, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter,
                           sim->seed, sim->heading_noise, pose_at);
    else
        hipLaunchKernelGGL((k_step_fused<L, false>), grid, block, shmem, sim->dev->stream, sim->map, sim->state->ptr,
                           sim->n, actions, reward, done, obs, obs_cols, sim->beams, per, wpc, sim->dirs, sim->counter,
                           sim->seed, sim->heading_noise, pose_at);
}

// k_physics then the lidar (racer/sim.mojo:314-321).
static int launch_pair(racer_sim* sim, float* actions, float* reward, float* done, float* obs, int64_t obs_cols) {
    const int n = sim->n;
    {
        racer_timed timer(sim->dev, "physics");
        auto kernel = sim->map.layers == 1 ? k_physics<1> : (sim->map.layers == 2 ? k_physics<2> : k_physics<0>);
        hipLaunchKernelGGL(kernel, dim3((n + 63) / 64), dim3(64), 0, sim->dev->stream, sim->map, sim->state->ptr, n,
                           actions, (int64_t)n, reward, done, obs, obs_cols, sim->beams, sim->seed,
                           sim->heading_noise);
    }
    if (racer_launch_check("physics")) return 1;
    if (sim->count_iterations)
        RACER_TRY(hipMemsetAsync(sim->counter, 0, 2 * sizeof(unsigned long long), sim->dev->stream));
    if (sim->schedule == 1) {
        int64_t threads = (int64_t)n * sim->beams;
        racer_timed timer(sim->dev, "lidar_flat");
        hipLaunchKernelGGL(k_lidar_flat, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, sim->dev->stream,
                           sim->map, sim->state->ptr, n, obs, obs_cols, sim->beams);
    } else {
        racer_timed timer(sim->dev, "lidar");
        launch_lidar_any(sim, obs, obs_cols);
    }
    return racer_launch_check("lidar");
}

int racer_launch_step(racer_sim* sim, float* actions, float* reward, float* done, float* obs, int64_t obs_cols) {
    int pose_at = 0;
    const size_t fused_shmem = sim->fused_step ? step_fused_shmem(sim, &pose_at) : 0;
    if (fused_shmem) {
        if (sim->count_iterations)
            RACER_TRY(hipMemsetAsync(sim->counter, 0, 2 * sizeof(unsigned long long), sim->dev->stream));
        {
            racer_timed timer(sim->dev, "step_fused");
            if (sim->map.layers == 1)
                launch_step_fused<1>(sim, actions, reward, done, obs, obs_cols, fused_shmem, pose_at);
            else
                launch_step_fused<2>(sim, actions, reward, done, obs, obs_cols, fused_shmem, pose_at);
        }
        if (racer_launch_check("fused step")) return 1;
    } else if (launch_pair(sim, actions, reward, done, obs, obs_cols)) {
        return 1;
    }
    if (sim->count_iterations) {
        unsigned long long count[2] = {0, 0};
        RACER_TRY(hipMemcpyAsync(count, sim->counter, sizeof(count), hipMemcpyDeviceToHost, sim->dev->stream));
        RACER_TRY(hipStreamSynchronize(sim->dev->stream));
        sim->last_iterations = (int64_t)count[0];
        sim->last_trips = (int64_t)count[1];
    }
    return 0;
}

// The instance table of a batch, with work-item prefix sums of `per_car` items per car.
static racer_sim_table sim_table(const racer_batch* b, int per_car) {
    racer_sim_table tab{};
    tab.count = (int)b->sims.size();
    tab.start[0] = 0;
    for (int k = 0; k < tab.count; ++k) {
        const racer_sim* sim = b->sims[k];
        racer_sim_slot& q = tab.slot[k];
        q.map = sim->map;
        q.state = sim->state->ptr;
        q.obs = sim->obs->ptr;
        q.actions = sim->actions->ptr;
        q.reward = sim->reward->ptr;
        q.done = sim->done->ptr;
        q.dirs = sim->dirs;
        q.n = sim->n;
        q.seed = sim->seed;
        q.beams = sim->beams;
        q.heading_noise = sim->heading_noise;
        tab.start[k + 1] = tab.start[k] + (int64_t)sim->n * per_car;
    }
    return tab;
}

size_t racer_batch_fused_shmem(const racer_batch* b, int waves_per_car, int* pose_at) {
    const int coop_waves = LIDAR_COOP_CARS * waves_per_car;
    if (coop_waves > 16) return 0;
    int layers = 1;
    for (const racer_sim* sim : b->sims) {
        if (sim->map.layers != 1 && sim->map.layers != 2) return 0;
        layers = sim->map.layers > layers ? sim->map.layers : layers;
    }
    const size_t lidar = (size_t)coop_waves * LIDAR_WAVE_LDS + (size_t)LIDAR_COOP_CARS * (b->beams + 1);
    const size_t tiles = (size_t)LIDAR_COOP_CARS * (layers == 1 ? phys_tile_floats<1>() : phys_tile_floats<2>());
    const size_t at = lidar > tiles ? lidar : tiles;
    const size_t bytes = (at + 8 * LIDAR_COOP_CARS) * sizeof(float);
    if (bytes > LIDAR_COOP_MAX_SHMEM) return 0;
    *pose_at = (int)at;
    return bytes;
}

// The fused step of every instance: one launch over the instances'
// workgroups of LIDAR_COOP_CARS cars (persistent when b->fused_step == 2).
static int launch_batch_fused(racer_batch* b, size_t shmem, int pose_at) {
    racer_device* dev = b->dev;
    const int wpc = b->waves_per_car, per = (b->beams + wpc - 1) / wpc;
    racer_sim_table tab{};
    tab.count = (int)b->sims.size();
    tab.start[0] = 0;
    for (int k = 0; k < tab.count; ++k) {
        const racer_sim* sim = b->sims[k];
        racer_sim_slot& q = tab.slot[k];
        q.map = sim->map;
        q.state = sim->state->ptr;
        q.obs = sim->obs->ptr;
        q.actions = sim->actions->ptr;
        q.reward = sim->reward->ptr;
        q.done = sim->done->ptr;
        q.dirs = sim->dirs;
        q.n = sim->n;
        q.seed = sim->seed;
        q.beams = sim->beams;
        q.heading_noise = sim->heading_noise;
        tab.start[k + 1] = tab.start[k] + (sim->n + LIDAR_COOP_CARS - 1) / LIDAR_COOP_CARS;
    }
    const int64_t total = tab.start[tab.count];
    const bool persist = b->fused_step == 2;
    int64_t grid = total;
    if (persist) {
        const void* kernel = b->count_iterations ? (const void*)k_step_fused_batch<true, true>
                                                 : (const void*)k_step_fused_batch<false, true>;
        int per_cu = 0;
        RACER_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * LIDAR_COOP_CARS * wpc, shmem));
        const int64_t resident = (int64_t)(per_cu > 0 ? per_cu : 1) * dev->cus;
        grid = total < resident ? total : (resident >= 8 ? resident & ~(int64_t)7 : resident);
    }
    const dim3 block(64 * LIDAR_COOP_CARS * wpc);
    racer_timed timer(dev, persist ? "step_fused_batch_persistent" : "step_fused_batch");
    if (b->count_iterations) {
        if (persist)
            hipLaunchKernelGGL((k_step_fused_batch<true, true>), dim3((unsigned)grid), block, shmem, dev->stream, tab,
                               per, wpc, b->counter, pose_at);
        else
            hipLaunchKernelGGL((k_step_fused_batch<true, false>), dim3((unsigned)grid), block, shmem, dev->stream,
                               tab, per, wpc, b->counter, pose_at);
    } else {
        if (persist)
            hipLaunchKernelGGL((k_step_fused_batch<false, true>), dim3((unsigned)grid), block, shmem, dev->stream,
                               tab, per, wpc, b->counter, pose_at);
        else
            hipLaunchKernelGGL((k_step_fused_batch<false, false>), dim3((unsigned)grid), block, shmem, dev->stream,
                               tab, per, wpc, b->counter, pose_at);
    }
    return 0;
}

int racer_launch_batch_step(racer_batch* b) {
    racer_device* dev = b->dev;
    int pose_at = 0;
    const size_t fused = b->fused_step ? racer_batch_fused_shmem(b, b->waves_per_car, &pose_at) : 0;
    if (fused) {
        if (b->count_iterations)
            RACER_TRY(hipMemsetAsync(b->counter, 0, 2 * sizeof(unsigned long long), dev->stream));
        if (launch_batch_fused(b, fused, pose_at)) return 1;
        if (racer_launch_check("fused batch step")) return 1;
        if (b->count_iterations) {
            unsigned long long count[2] = {0, 0};
            RACER_TRY(hipMemcpyAsync(count, b->counter, sizeof(count), hipMemcpyDeviceToHost, dev->stream));
            RACER_TRY(hipStreamSynchronize(dev->stream));
            b->last_iterations = (int64_t)count[0];
        }
        return 0;
    }
    {
        const racer_sim_table tab = sim_table(b, 1);
        racer_timed timer(dev, "physics");
        hipLaunchKernelGGL(k_physics_batch, dim3((unsigned)((tab.start[tab.count] + 63) / 64)), dim3(64), 0,
                           dev->stream, tab);
    }
    if (racer_launch_check("physics batch")) return 1;
    return racer_launch_batch_lidar(b);
}

int racer_launch_batch_lidar(racer_batch* b) {
    racer_device* dev = b->dev;
    if (b->count_iterations) RACER_TRY(hipMemsetAsync(b->counter, 0, 2 * sizeof(unsigned long long), dev->stream));
    {
        const int wpc = b->waves_per_car;
        const int per = (b->beams + wpc - 1) / wpc;
        const racer_sim_table tab = sim_table(b, wpc);
        const size_t bytes_per_wave = (size_t)(LIDAR_WAVE_LDS + per + 1) * sizeof(float);
        const int waves = 4 * bytes_per_wave <= 65536 ? 4 : 1;
        const int64_t total = tab.start[tab.count];
        dim3 grid((unsigned)((total + waves - 1) / waves)), block(64 * waves);
        racer_timed timer(dev, "lidar");
        if (b->count_iterations)
            hipLaunchKernelGGL(k_lidar_batch<true>, grid, block, waves * bytes_per_wave, dev->stream, tab, per, wpc,
                               b->counter);
        else
            hipLaunchKernelGGL(k_lidar_batch<false>, grid, block, waves * bytes_per_wave, dev->stream, tab, per, wpc,
                               b->counter);
    }
    if (racer_launch_check("lidar batch")) return 1;
    if (b->count_iterations) {
        unsigned long long count[2] = {0, 0};
        RACER_TRY(hipMemcpyAsync(count, b->counter, sizeof(count), hipMemcpyDeviceToHost, dev->stream));
        RACER_TRY(hipStreamSynchronize(dev->stream));
        b->last_iterations = (int64_t)count[0];
    }
    return 0;
}

int racer_launch_surrogate_features(racer_sim* sim, const float* act, int64_t act_cols, float* feat, int64_t ldf,
                                    const float* mean, const float* inv_std) {
    const int n = sim->n;
    auto kernel = sim->map.layers == 1 ? k_surrogate_features<1>
                                       : (sim->map.layers == 2 ? k_surrogate_features<2> : k_surrogate_features<0>);
    {
        racer_timed timer(sim->dev, "surrogate_features");
        hipLaunchKernelGGL(kernel, dim3((n + 255) / 256), dim3(256), 0, sim->dev->stream, sim->map, sim->state->ptr,
                           n, act, act_cols, feat, ldf, mean, inv_std);
    }
    return racer_launch_check("surrogate features");
}

int racer_launch_surrogate_targets(racer_sim* sim, const float* prev, float* tgt, int64_t ldt, float* mask) {
    const int n = sim->n;
    {
        racer_timed timer(sim->dev, "surrogate_targets");
        hipLaunchKernelGGL(k_surrogate_targets, dim3((n + 255) / 256), dim3(256), 0, sim->dev->stream, prev,
                           sim->state->ptr, sim->done->ptr, n, tgt, ldt, mask);
    }
    return racer_launch_check("surrogate targets");
}

int racer_launch_surrogate_step(racer_sim* sim, const float* pred, int64_t ldp, const float* out_mean,
                                const float* out_std, float min_clearance, float* fallback) {
    const int n = sim->n;
    auto kernel = sim->map.layers == 1 ? k_physics_surrogate<1>
                                       : (sim->map.layers == 2 ? k_physics_surrogate<2> : k_physics_surrogate<0>);
    {
        racer_timed timer(sim->dev, "physics_surrogate");
        hipLaunchKernelGGL(kernel, dim3((n + 63) / 64), dim3(64), 0, sim->dev->stream, sim->map, sim->state->ptr, n,
                           sim->actions->ptr, (int64_t)n, sim->reward->ptr, sim->done->ptr, sim->obs->ptr,
                           (int64_t)n, sim->beams, sim->seed, sim->heading_noise, pred, ldp, out_mean, out_std,
                           min_clearance, fallback);
    }
    if (racer_launch_check("physics surrogate")) return 1;
    {
        racer_timed timer(sim->dev, "lidar");
        launch_lidar_any(sim, sim->obs->ptr, (int64_t)n);
    }
    return racer_launch_check("lidar");
}

int racer_launch_surrogate_imagine(racer_sim* sim, const float* actions, int steps, const uint16_t* frags,
                                   const float* w3, const float* stats, float min_clearance, float* fallback) {
    const int n = sim->n;
    auto pick = [&](auto multi) {
        constexpr bool M = decltype(multi)::value;
        return sim->map.layers == 1 ? k_surrogate_fused<1, M>
                                    : (sim->map.layers == 2 ? k_surrogate_fused<2, M> : k_surrogate_fused<0, M>);
    };
    auto kernel = steps == 1 ? pick(std::false_type{}) : pick(std::true_type{});
    {
        racer_timed timer(sim->dev, steps == 1 ? "surrogate_fused" : "surrogate_imagine");
        hipLaunchKernelGGL(kernel, dim3((n + SF_CARS - 1) / SF_CARS), dim3(SF_THREADS), 0, sim->dev->stream, sim->map,
                           sim->state->ptr, n, actions, (int64_t)n, steps, 2 * (int64_t)n, sim->reward->ptr,
                           sim->done->ptr, sim->obs->ptr, (int64_t)n, sim->beams, sim->seed, sim->heading_noise,
                           frags, w3, stats, min_clearance, fallback);
    }
    return racer_launch_check("surrogate fused");
}

int racer_launch_physics_steps(racer_sim* sim, const float* actions, int steps) {
    const int n = sim->n;
    auto kernel = sim->map.layers == 1 ? k_physics_steps<1>
                                       : (sim->map.layers == 2 ? k_physics_steps<2> : k_physics_steps<0>);
    {
        racer_timed timer(sim->dev, "physics_steps");
        hipLaunchKernelGGL(kernel, dim3((n + 63) / 64), dim3(64), 0, sim->dev->stream, sim->map, sim->state->ptr, n,
                           actions, (int64_t)n, steps, 2 * (int64_t)n, sim->reward->ptr, sim->done->ptr,
                           sim->obs->ptr, (int64_t)n, sim->beams, sim->seed, sim->heading_noise);
    }
    return racer_launch_check("physics steps");
}

int racer_launch_surrogate_fused(racer_sim* sim, const uint16_t* frags, const float* w3, const float* stats,
                                 float min_clearance, float* fallback) {
    if (racer_launch_surrogate_imagine(sim, sim->actions->ptr, 1, frags, w3, stats, min_clearance, fallback))
        return 1;
    const int n = sim->n;
    {
        racer_timed timer(sim->dev, "lidar");
        launch_lidar_any(sim, sim->obs->ptr, (int64_t)n);
    }
    return racer_launch_check("lidar");
}

extern "C" int racer_sim_imagine(racer_sim* sim, const racer_view* actions_p, int steps) {
    racer_clear_error();
    RACER_REQUIRE(sim && actions_p, "null argument");
    RACER_REQUIRE(steps >= 1, "imagine at least one step");
    if (racer_check_view(*actions_p, 2 * (int64_t)steps, sim->n, "actions")) return 1;
    return racer_launch_physics_steps(sim, view_ptr(*actions_p), steps);
}

extern "C" int racer_sim_step_f64(racer_sim* sim, racer_buf* st, const racer_view* actions_p) {
    racer_clear_error();
    RACER_REQUIRE(sim && st && actions_p, "null argument");
    const int n = sim->n;
    RACER_REQUIRE(st->len >= (size_t)(12 * n), "the float64 state needs 12 floats (6 doubles) per car");
    RACER_REQUIRE((uintptr_t)st->ptr % 8 == 0, "the float64 state must be 8-byte aligned");
    if (racer_check_view(*actions_p, 2, n, "actions")) return 1;
    {
        racer_timed timer(sim->dev, "step_f64");
        hipLaunchKernelGGL(k_step_f64, dim3((n + 63) / 64), dim3(64), 0, sim->dev->stream, sim->map,
                           sim->state->ptr, n, reinterpret_cast<double*>(st->ptr), view_ptr(*actions_p),
                           actions_p->cols);
    }
    return racer_launch_check("step f64");
}

int racer_launch_sample_raw(racer_device* dev, int n, const float* out, int64_t out_cols, const float* log_std,
                            float* actions, int64_t act_cols, float* logp, float* value, int deterministic,
                            int64_t step, int seed) {
    {
        racer_timed timer(dev, "sample");
        hipLaunchKernelGGL(k_sample, dim3((n + 255) / 256), dim3(256), 0, dev->stream, n, out, out_cols, log_std,
                           actions, act_cols, logp, value, deterministic, step, dev->step_base, seed);
    }
    return racer_launch_check("sample");
}

int racer_launch_sample(racer_sim* sim, const float* out, int64_t out_cols, const float* log_std, float* actions,
                        float* logp, float* value, int deterministic, int64_t step) {
    return racer_launch_sample_raw(sim->dev, sim->n, out, out_cols, log_std, actions, (int64_t)sim->n, logp, value,
                                   deterministic, step, sim->seed);
}

// ---------------------------------------------------------------- Sim C ABI
extern "C" int racer_sim_create(racer_device* dev, const float* track, size_t len, int n, int seed,
                                const racer_config* cfg, racer_sim** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(dev, "null device");
    RACER_REQUIRE(track && len >= RACER_HEADER, "not a track array");
    RACER_REQUIRE(n > 0, "the car count must be positive");
    racer_config def;
    racer_config_default(&def);
    if (!cfg) cfg = &def;
    RACER_REQUIRE(cfg->beams >= 2, "at least two lidar beams are needed");
    RACER_REQUIRE(cfg->beams <= 15000, "at most 15000 lidar beams are supported");
    int nx = (int)track[0], ny = (int)track[1], layers = (int)track[7], wps = (int)track[5];
    RACER_REQUIRE(nx >= 2 && ny >= 2 && layers >= 1 && wps >= 1, "not a track array");
    int64_t plane = (int64_t)nx * ny;
    RACER_REQUIRE(plane <= (1 << 24), "track grids of more than 2^24 cells are not supported");
    int64_t route = RACER_HEADER + 3 * (int64_t)layers * plane;
    RACER_REQUIRE((size_t)(route + 4 * (int64_t)wps) <= len, "not a track array");
    RACER_REQUIRE(track[4] > 0.0f, "not a track array: the cell size must be positive");
    // Route lookups index the waypoint table with the nearest-waypoint grids: keep them in range.
    for (int l = 0; l < layers; ++l) {
        const float* nearest = track + RACER_HEADER + (int64_t)(3 * l + RACER_NEAREST) * plane;
        for (int64_t c = 0; c < plane; ++c)
            RACER_REQUIRE(nearest[c] >= 0.0f && nearest[c] <= (float)(wps - 1),
                          "not a track array: a nearest-waypoint cell is out of range");
    }
    auto* sim = new racer_sim();
    sim->dev = dev;
    racer_device_retain(dev);
    sim->n = n;
    sim->seed = seed;
    sim->beams = cfg->beams;
    sim->heading_noise = cfg->heading_noise;
    int IN = cfg->beams + 5;
    if (racer_upload(dev, track, len, &sim->data) || racer_alloc(dev, (size_t)RACER_STATE * n, 0.0f, &sim->state) ||
        racer_alloc(dev, (size_t)IN * n, 0.0f, &sim->obs) || racer_alloc(dev, 2 * (size_t)n, 0.0f, &sim->actions) ||
        racer_alloc(dev, (size_t)n, 0.0f, &sim->reward) || racer_alloc(dev, (size_t)n, 0.0f, &sim->done)) {
        racer_sim_destroy(sim);
        return 1;
    }
    racer_dmap& m = sim->map;
    m.d = sim->data->ptr;
    m.nx = nx;
    m.ny = ny;
    m.x0 = track[2];
    m.y0 = track[3];
    m.cell = track[4];
    m.waypoints = wps;
    m.spacing = track[6];
    m.layers = layers;
    m.slope = track[8];
    m.inv_cell = 1.0f / m.cell;
    // The lidar's reciprocal quotients need the cell and the slope within
    // [2^-20, 2^20] and every height finite within 2^40 (the bounds are
    // explained where the lidar decides per car).
    m.quotient_ok = m.slope >= 0x1p-20f && m.slope <= 0x1p20f && m.cell >= 0x1p-20f && m.cell <= 0x1p20f;
    for (int l = 0; m.quotient_ok && l < layers; ++l) {
        const float* height = track + RACER_HEADER + (int64_t)(3 * l + RACER_HEIGHT) * plane;
        for (int64_t c = 0; c < plane; ++c)
            if (!(fabsf(height[c]) <= 0x1p40f)) {
                m.quotient_ok = 0;
                break;
            }
    }
    m.route = route;
    m.plane = plane;
    m.packed = nullptr;
    m.padded = nullptr;
    if (layers <= 2) {
        RACER_TRY_OR(hipMalloc(&sim->packed, (size_t)(plane + 1) * layers * 2 * sizeof(float)), racer_sim_destroy(sim));
        hipLaunchKernelGGL(k_pack, dim3((unsigned)((plane + 1 + 255) / 256)), dim3(256), 0, dev->stream, m.d, plane,
                           layers, sim->packed);
        if (racer_launch_check("pack")) {
            racer_sim_destroy(sim);
            return 1;
        }
        m.packed = sim->packed;
        // The padded grid: a lidar probe lies within RACER_RANGE (1 + 2^-20) of
        // its origin, so `pad` cells of RACER_RANGE + 1 m plus 16 leave an
        // origin box reaching 14 cells past the map on every side. Kept only
        // while its cells number fewer than 2^24 (the probe's index is one
        // 24-bit multiply-add).
        const double cell = m.cell;
        const double reach = (double)RACER_RANGE + 1.0;
        const double padc = m.quotient_ok ? std::ceil(reach / cell) + 16.0 : 0.0;
        const double pcells = ((double)nx + 2.0 * padc) * ((double)ny + 2.0 * padc + 1.0);
        if (m.quotient_ok && pcells < 0x1p24) {
            m.pad = (int)padc;
            m.pnx = nx + 2 * m.pad;
            const int64_t cells = (int64_t)m.pnx * (ny + 2 * m.pad);
            m.pad_base = (uint32_t)(((int64_t)m.pad * m.pnx + m.pad) * 8 * layers);
            m.pad_bytes = (uint32_t)(cells * 8 * layers);
            m.pad_x0 = (float)((double)m.x0 - (padc - 2.0) * cell + reach);
            m.pad_x1 = (float)((double)m.x0 + ((double)nx + padc - 3.0) * cell - reach);
            m.pad_y0 = (float)((double)m.y0 - (padc - 2.0) * cell + reach);
            m.pad_y1 = (float)((double)m.y0 + ((double)ny + padc - 3.0) * cell - reach);
            RACER_TRY_OR(hipMalloc(&sim->padded, (size_t)cells * layers * 2 * sizeof(float)), racer_sim_destroy(sim));
            hipLaunchKernelGGL(k_pack_padded, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, dev->stream, m.d, nx,
                               ny, m.pad, layers, sim->padded);
            if (racer_launch_check("pack")) {
                racer_sim_destroy(sim);
                return 1;
            }
            m.padded = sim->padded;
        }
        // The lidar's one-step nearest cell (racer_nearest_by_reciprocal1) must
        // give the division's index at every cell boundary of the padded grid
        // (of the map itself without it).
        if (m.quotient_ok && racer_nearest_form_bad(m.cell, -m.pad - 2, std::max(nx, ny) + m.pad + 2, 1) != 0) {
            m.quotient_ok = 0;
            m.padded = nullptr;
        }
    }
    // The lidar's order keys, all 0 (the natural order) until the first launch.
    RACER_TRY_OR(hipMalloc(&sim->lidar_keys, (size_t)n * RACER_LIDAR_KEYS), racer_sim_destroy(sim));
    RACER_TRY_OR(hipMemsetAsync(sim->lidar_keys, 0, (size_t)n * RACER_LIDAR_KEYS, dev->stream), racer_sim_destroy(sim));
    m.lidar_keys = sim->lidar_keys;
    // Beam directions: azimuth = -FOV/2 + b FOV/(BEAMS-1) (racer/sim.mojo:246).
    std::vector<float2> dirs(cfg->beams);
    for (int b = 0; b < cfg->beams; ++b) {
        float az = -RACER_FOV / 2.0f + (float)b * (RACER_FOV / (float)(cfg->beams - 1));
        dirs[b] = make_float2(racer_cos(az), racer_sin(az));
    }
    RACER_TRY_OR(hipMalloc(&sim->dirs, dirs.size() * sizeof(float2)), racer_sim_destroy(sim));
    RACER_TRY_OR(hipMemcpyAsync(sim->dirs, dirs.data(), dirs.size() * sizeof(float2), hipMemcpyHostToDevice,
                                dev->stream),
                 racer_sim_destroy(sim));
    RACER_TRY_OR(hipMalloc(&sim->counter, 2 * sizeof(unsigned long long)), racer_sim_destroy(sim));
    RACER_TRY_OR(hipStreamSynchronize(dev->stream), racer_sim_destroy(sim));
    if (lidar_coop_setup()) {
        racer_sim_destroy(sim);
        return 1;
    }
    *out = sim;
    return 0;
}

extern "C" void racer_sim_destroy(racer_sim* sim) {
    if (!sim) return;
    if (sim->dev) (void)hipStreamSynchronize(sim->dev->stream);  // destroy reports no errors
    racer_free(sim->data);
    racer_free(sim->state);
    racer_free(sim->obs);
    racer_free(sim->actions);
    racer_free(sim->reward);
    racer_free(sim->done);
    if (sim->packed) (void)hipFree(sim->packed);
    if (sim->padded) (void)hipFree(sim->padded);
    if (sim->lidar_keys) (void)hipFree(sim->lidar_keys);
    if (sim->dirs) (void)hipFree(sim->dirs);
    if (sim->counter) (void)hipFree(sim->counter);
    racer_device* dev = sim->dev;
    delete sim;
    racer_device_release(dev);
}

extern "C" int racer_sim_reset(racer_sim* sim) {
    racer_clear_error();
    if (racer_launch_reset(sim)) return 1;
    return racer_launch_step(sim, sim->actions->ptr, sim->reward->ptr, sim->done->ptr, sim->obs->ptr, sim->n);
}

extern "C" int racer_sim_step(racer_sim* sim, const racer_view* actions_p, const racer_view* reward_p,
                              const racer_view* done_p, const racer_view* obs_p) {
    racer_clear_error();
    RACER_REQUIRE(sim && actions_p && reward_p && done_p && obs_p, "null argument");
    const racer_view &actions = *actions_p, &reward = *reward_p, &done = *done_p, &obs = *obs_p;
    int n = sim->n;
    if (racer_check_view(actions, 2, n, "actions") || racer_check_view(reward, 1, n, "reward") ||
        racer_check_view(done, 1, n, "done") || racer_check_view(obs, sim->beams + 5, n, "obs"))
        return 1;
    return racer_launch_step(sim, view_ptr(actions), view_ptr(reward), view_ptr(done), view_ptr(obs), n);
}

// The second launch of Sim.step (racer/sim.mojo:317-321) alone: every beam of
// every car from the current state and obs slope rows, into the sim's obs.
extern "C" int racer_sim_scan(racer_sim* sim) {
    racer_clear_error();
    RACER_REQUIRE(sim, "null argument");
    if (sim->count_iterations)
        RACER_TRY(hipMemsetAsync(sim->counter, 0, 2 * sizeof(unsigned long long), sim->dev->stream));
    if (sim->schedule == 1) {
        const int64_t threads = (int64_t)sim->n * sim->beams;
        hipLaunchKernelGGL(k_lidar_flat, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, sim->dev->stream,
                           sim->map, sim->state->ptr, sim->n, sim->obs->ptr, (int64_t)sim->n, sim->beams);
    } else {
        launch_lidar_any(sim, sim->obs->ptr, sim->n);
    }
    if (racer_launch_check("lidar")) return 1;
    if (sim->count_iterations) {
        unsigned long long count[2] = {0, 0};
        RACER_TRY(hipMemcpyAsync(count, sim->counter, sizeof(count), hipMemcpyDeviceToHost, sim->dev->stream));
        RACER_TRY(hipStreamSynchronize(sim->dev->stream));
        sim->last_iterations = (int64_t)count[0];
        sim->last_trips = (int64_t)count[1];
    }
    return 0;
}

extern "C" int racer_sim_sample(racer_sim* sim, const racer_view* out_p, const racer_view* log_std_p,
                                const racer_view* actions_p, const racer_view* logp_p, const racer_view* value_p,
                                int deterministic, int64_t step) {
    racer_clear_error();
    RACER_REQUIRE(sim && out_p && log_std_p && actions_p && logp_p && value_p, "null argument");
    const racer_view &out = *out_p, &log_std = *log_std_p, &actions = *actions_p, &logp = *logp_p,
                     &value = *value_p;
    int n = sim->n;
    if (racer_check_view(out, 3, n, "policy output") || racer_check_view(log_std, 1, 2, "log std") ||
        racer_check_view(actions, 2, n, "actions") || racer_check_view(logp, 1, n, "logp") ||
        racer_check_view(value, 1, n, "value"))
        return 1;
    return racer_launch_sample(sim, view_ptr(out), n, view_ptr(log_std), view_ptr(actions), view_ptr(logp),
                               view_ptr(value), deterministic, step);
}

extern "C" racer_buf* racer_sim_state(racer_sim* sim) { return sim->state; }
extern "C" racer_buf* racer_sim_obs(racer_sim* sim) { return sim->obs; }
extern "C" racer_buf* racer_sim_actions(racer_sim* sim) { return sim->actions; }
extern "C" racer_buf* racer_sim_reward(racer_sim* sim) { return sim->reward; }
extern "C" racer_buf* racer_sim_done(racer_sim* sim) { return sim->done; }
extern "C" int racer_sim_cars(const racer_sim* sim) { return sim->n; }

// The viewer's snapshot (main.mojo:117-122): state, obs, actions, reward and
// done, each feature-major, back to back in one host array; waits for the
// step that produced them.
extern "C" int racer_sim_snapshot(racer_sim* sim, float* host, size_t count) {
    racer_clear_error();
    RACER_REQUIRE(sim && host, "null argument");
    const size_t n = (size_t)sim->n, need = (RACER_STATE + (size_t)(sim->beams + 5) + 4) * n;
    RACER_REQUIRE(count >= need, "the snapshot needs (11 + beams + 5 + 4) x cars floats");
    size_t at = 0;
    for (racer_buf* b : {sim->state, sim->obs, sim->actions, sim->reward, sim->done}) {
        RACER_TRY(hipMemcpyAsync(host + at, b->ptr, b->len * sizeof(float), hipMemcpyDeviceToHost, sim->dev->stream));
        at += b->len;
    }
    RACER_TRY(hipStreamSynchronize(sim->dev->stream));
    return 0;
}
extern "C" int racer_sim_beams(const racer_sim* sim) { return sim->beams; }

extern "C" int racer_sim_count_iterations(racer_sim* sim, int on) {
    sim->count_iterations = on != 0;
    return 0;
}
extern "C" int64_t racer_sim_last_iterations(racer_sim* sim) { return sim->last_iterations; }
extern "C" int64_t racer_sim_last_trips(racer_sim* sim) { return sim->last_trips; }

// The floats x with |x| < 2^23 of either sign, every `stride`-th bit pattern
// of that range: count those where floor_index(x) differs from
// (int)floorf(x), the index the padded lidar probe takes it for.
static __global__ void k_check_floor(uint32_t patterns, uint32_t stride, unsigned long long* bad) {
    uint32_t local = 0;
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < patterns; i += step) {
        const uint32_t u = (uint32_t)((uint64_t)i * stride);
#pragma unroll
        for (uint32_t sign = 0; sign < 2; ++sign) {
            const float x = __uint_as_float(u | (sign << 31));
            local += floor_index(x) != (int)floorf(x) ? 1u : 0u;
        }
    }
    if (local != 0) atomicAdd(bad, (unsigned long long)local);
}

extern "C" int racer_check_floor_index(racer_device* dev, uint32_t stride, uint64_t* mismatches) {
    racer_clear_error();
    RACER_REQUIRE(dev && mismatches && stride >= 1, "racer_check_floor_index: device, output and stride >= 1");
    RACER_TRY(hipSetDevice(dev->id));
    // bit patterns below that of 2^23
    const uint32_t patterns = (uint32_t)((0x4B000000ull + stride - 1) / stride);
    unsigned long long* bad = nullptr;
    RACER_TRY(hipMalloc(&bad, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), dev->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_floor, dim3(2048), dim3(256), 0, dev->stream, patterns, stride, bad);
        e = hipGetLastError();
    }
    unsigned long long host = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&host, bad, sizeof(host), hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    (void)hipFree(bad);
    RACER_TRY(e);
    *mismatches = host;
    return 0;
}

extern "C" int racer_sim_set_waves_per_car(racer_sim* sim, int waves) {
    racer_clear_error();
    RACER_REQUIRE(sim && (waves == 1 || waves == 2 || waves == 4), "waves per car must be 1, 2 or 4");
    sim->waves_per_car = waves;
    return 0;
}

extern "C" int racer_sim_waves_per_car(const racer_sim* sim) { return sim->waves_per_car; }

extern "C" int racer_sim_set_lidar_coop(racer_sim* sim, int on) {
    racer_clear_error();
    RACER_REQUIRE(sim, "no sim");
    sim->lidar_coop = on != 0;
    return 0;
}

extern "C" int racer_sim_lidar_coop(const racer_sim* sim) { return sim->lidar_coop ? 1 : 0; }

extern "C" int racer_sim_set_fused_step(racer_sim* sim, int on) {
    racer_clear_error();
    RACER_REQUIRE(sim, "no sim");
    sim->fused_step = on != 0;
    return 0;
}

extern "C" int racer_sim_fused_step(const racer_sim* sim) {
    int pose_at = 0;
    return sim->fused_step && step_fused_shmem(sim, &pose_at) != 0 ? 1 : 0;
}

extern "C" int racer_sim_set_schedule(racer_sim* sim, int schedule) {
    racer_clear_error();
    RACER_REQUIRE(sim && (schedule == 0 || schedule == 1),
                  "schedule must be 0 (wavefront per car) or 1 (thread per beam)");
    sim->schedule = schedule;
    return 0;
}

extern "C" int racer_sim_schedule(const racer_sim* sim) { return sim->schedule; }


