// Internal structures of libracer_hip: device, buffers, sim, policy, trainer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/racer.h"
#include "racer_contract.h"
#include "racer_math.h"

// ---------------------------------------------------------------- errors
void racer_set_error(const std::string& msg);
void racer_clear_error();

#define RACER_TRY(expr)                                                              \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            racer_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            return 1;                                                                \
        }                                                                            \
    } while (0)

// RACER_TRY for a half-built object: `cleanup` (its destroy call) runs before
// the error return, so a failed create leaks nothing.
#define RACER_TRY_OR(expr, cleanup)                                                  \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            racer_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            cleanup;                                                                 \
            return 1;                                                                \
        }                                                                            \
    } while (0)

#define RACER_REQUIRE(cond, msg)   \
    do {                           \
        if (!(cond)) {             \
            racer_set_error(msg);  \
            return 1;              \
        }                          \
    } while (0)

// ---------------------------------------------------------------- device
constexpr int RACER_SUM_PARTIALS = 4096;

struct racer_profile_sample {
    std::string name;
    int64_t launches = 0;
    double total_ms = 0.0;
};

struct racer_pending_event {
    int kernel;
    hipEvent_t start, stop;
};

// Size of the fixed per-GPU scratch array of the allocation-free entry points
// (16 MiB of float32).
constexpr size_t RACER_FIXED_SCRATCH_FLOATS = (size_t)4 << 20;

struct racer_device {
    int id = 0;
    hipStream_t stream = nullptr;
    std::string name;
    int cus = 0;
    bool profiling = false;
    std::vector<racer_profile_sample> samples;
    std::vector<racer_pending_event> pending;
    std::vector<hipEvent_t> event_pool;
    float* scratch = nullptr;  // split-K partial products (grown on demand)
    size_t scratch_len = 0;
    // Device address of this GPU's fixed scratch array (RACER_FIXED_SCRATCH_FLOATS
    // floats, held until the process ends) for the allocation-free entry points.
    float* fixed_scratch = nullptr;
    double* sums = nullptr;    // RACER_SUM_PARTIALS float64 partial sums of the row reductions
    std::string profile_only;  // when set, the profiler times only the kernel of this name
    // Graph capture: while a graph is captured, sampling kernels add the device
    // scalar *step_base to their step argument, so one captured graph serves
    // every replay; graph_base is that scalar, written before each replay.
    bool capturing = false;
    const int64_t* step_base = nullptr;
    int64_t* graph_base = nullptr;
    // Owners: the handle's creator plus every object made on the device. The
    // device is freed when the last of them lets go, so objects may be
    // destroyed in any order (a garbage collector finalises cycles in none).
    int refs = 1;
};

// Workgroups go round-robin over the 8 XCDs; remapping the workgroup index
// gives each XCD a contiguous run of cars (when the grid is a multiple of 8),
// so neighbouring cars' stores (adjacent floats of one obs row) meet in one
// L2, and the physics, lidar and forward kernels of a step keep each car on
// the same XCD.
__device__ __forceinline__ int xcd_block() {
    int block = blockIdx.x;
    if ((gridDim.x & 7) == 0) block = (block & 7) * (gridDim.x >> 3) + (block >> 3);
    return block;
}

void racer_device_retain(racer_device* dev);
void racer_device_release(racer_device* dev);

struct racer_buf {
    racer_device* dev = nullptr;
    float* ptr = nullptr;
    size_t len = 0;
    // Bumped by every write through the C ABI (racer_write*, racer_copy into
    // it), so objects that keep derived copies of a buffer (the policy's packed
    // weights) know to rebuild them.
    uint64_t gen = 0;
};

// Launch bracket for the profiler: records events around a launch when enabled.
// Two timing events, destroyed on every way out of the scope that made them.
struct racer_event_pair {
    hipEvent_t a = nullptr, b = nullptr;
    racer_event_pair() = default;
    racer_event_pair(const racer_event_pair&) = delete;
    racer_event_pair& operator=(const racer_event_pair&) = delete;
    ~racer_event_pair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct racer_timed {
    racer_device* dev;
    int kernel = -1;
    hipEvent_t start = nullptr, stop = nullptr;
    racer_timed(racer_device* d, const char* name);
    ~racer_timed();
};

int racer_check_view(const racer_view& v, int64_t rows, int64_t cols, const char* what);
inline float* view_ptr(const racer_view& v) { return v.buf->ptr + v.offset; }
int racer_ensure_scratch(racer_device* dev, size_t floats);
// The workspace of a split-K product: `ws` when the caller brings one (it must
// hold `floats`), else the device scratch grown to `floats`.
int racer_partials(racer_device* dev, size_t floats, float* ws, size_t ws_len, float** out);
int racer_launch_check(const char* what);

// ---------------------------------------------------------------- map (device side)
// Header fields of the track array (racer/sim.mojo:81-102) plus device pointers.
struct racer_dmap {
    const float* d;        // the track array as uploaded
    const float* packed;   // per-cell records (clearance, height) for every layer, or null
    int nx, ny, layers, waypoints;
    float x0, y0, cell, spacing, slope;
    float inv_cell;        // RN(1 / cell), for the division shortcut of the nearest-cell index
    int quotient_ok;       // heights, slope and cell bounded and the one-step nearest cell exact over the
                           // padded grid's index range: the lidar may divide by reciprocals
    int64_t route;         // offset of the waypoint table
    int64_t plane;         // nx * ny
    // The packed records again on a grid with `pad` cells of all -1 records on
    // every side (row stride pnx = nx + 2 pad), or null: a lidar origin within
    // [pad_x0, pad_x1] x [pad_y0, pad_y1] keeps every probe of its march inside
    // it, so the march needs no on-map test.
    const float* padded;
    int pad, pnx;
    uint32_t pad_base;     // byte offset of cell (0, 0)'s record
    uint32_t pad_bytes;    // bytes of padded records
    float pad_x0, pad_x1, pad_y0, pad_y1;
    // The sim's lidar order keys, RACER_LIDAR_KEYS bytes a car: the most march
    // iterations of any beam of each piece of the fan in the previous step
    // (written by every lidar launch; they order beams, never change a range).
    uint8_t* lidar_keys;
};
constexpr int RACER_LIDAR_KEYS = 32;

struct racer_sim {
    racer_device* dev = nullptr;
    int n = 0, seed = 0, beams = 0;
    float heading_noise = 0.0f;
    racer_buf *data = nullptr, *state = nullptr, *obs = nullptr, *actions = nullptr, *reward = nullptr,
              *done = nullptr;
    float* packed = nullptr;
    float* padded = nullptr;
    uint8_t* lidar_keys = nullptr;
    float2* dirs = nullptr;  // (cos, sin) of every beam's azimuth
    racer_dmap map{};
    int schedule = 0;       // lidar launch: 0 wavefront per car with beam refill, 1 thread per (car, beam)
    int waves_per_car = 2;  // schedule 0: wavefronts sharing one car's beams (1, 2 or 4)
    bool fused_step = false;  // physics and lidar in one launch where that applies
    bool lidar_coop = true;   // schedule 0: cooperative range stores from workgroups of 8 wavefronts
    bool count_iterations = false;
    unsigned long long* counter = nullptr;
    int64_t last_iterations = 0;
    int64_t last_trips = 0;  // wavefront loop trips of the counted launch (schedule 0)
};

// ---------------------------------------------------------------- policy
struct racer_policy {
    racer_device* dev = nullptr;
    int beams = 0, hidden = 0, precision = 0;
    int64_t IN = 0, HA = 0, P = 0, W1 = 0, W2 = 0, LOGSTD = 0;
    racer_buf *theta = nullptr, *w1t = nullptr, *w2t = nullptr;
    uint16_t* theta_bf16 = nullptr;  // bf16 W0 [H, pad32(IN)] then W1 [H, pad32(HA)], zero padded, for MFMA
    int64_t in_pad = 0, ha_pad = 0;
    uint16_t *frag0 = nullptr, *frag1 = nullptr;  // W0, W1 in MFMA fragment order (fused forward, H <= 256)
    uint16_t* w1t_bf16 = nullptr;  // bf16 w1t [HA, pad8(H)] for the backward's through product
    int64_t h_pad = 0;
    bool bf16_stale = true;
    uint64_t bf16_gen = 0;  // theta's write generation the bf16 copies were made from
    // float32 W0, W1 in the K = 1 MFMA operand order (float32 fused forward, H <= 256)
    float *frag0_f32 = nullptr, *frag1_f32 = nullptr;
    bool f32_stale = true;
    uint64_t f32_gen = 0;
    int act_kernel = RACER_ACT_AUTO;  // the forward of racer_policy_act (Runner steps, rollouts)
};

struct racer_acts {
    racer_device* dev = nullptr;
    int cols = 0, hidden = 0;
    racer_buf *h1 = nullptr, *h2 = nullptr, *o = nullptr;
};

struct racer_trainer {
    racer_device* dev = nullptr;
    int n = 0, beams = 0, hidden = 0, rollout = 0, epochs = 0, minibatches = 0, precision = 0;
    int64_t batch = 0, P = 0, IN = 0, HA = 0;
    int step = 0;
    racer_buf *obs = nullptr, *actions = nullptr, *logp = nullptr, *value = nullptr, *reward = nullptr,
              *done = nullptr, *advantage = nullptr, *x = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr,
              *grad = nullptr, *first = nullptr, *second = nullptr, *stats = nullptr;
    // The rollout's observations regrouped by minibatch, [MB][IN][batch], built
    // once per update so that no minibatch gathers strided columns; in bf16
    // (x16, two values a float) when the bf16 update's kernels read them so.
    racer_buf* xall = nullptr;
    bool x16 = false;
    racer_acts* acts = nullptr;
    double* dstats = nullptr;  // float64 accumulators: advantage sum and sum of squares, reward sum
};

// ---------------------------------------------------------------- batched launches
// Up to RACER_BATCH_MAX sims stepped by one launch per stage: each kernel
// finds its instance from the prefix sums `start` of its work items.
constexpr int RACER_BATCH_MAX = 16;

struct racer_sim_slot {
    racer_dmap map;
    float *state, *obs, *actions, *reward, *done;
    const float2* dirs;
    int n, seed, beams;
    float heading_noise;
};

struct racer_sim_table {
    int count;
    int64_t start[RACER_BATCH_MAX + 1];
    racer_sim_slot slot[RACER_BATCH_MAX];
};

struct racer_policy_slot {
    const float* x;
    int64_t ldx, act_cols;
    int64_t cols;
    float *h1, *h2, *o, *act, *logp, *value;
    int seed;
};

struct racer_policy_table {
    int count;
    int64_t start[RACER_BATCH_MAX + 1];  // in 16-column tiles
    racer_policy_slot slot[RACER_BATCH_MAX];
};

struct racer_batch {
    racer_device* dev = nullptr;
    std::vector<racer_sim*> sims;
    int beams = 0, hidden = 0;
    int64_t cars = 0;
    int waves_per_car = 2;
    // 0: physics and lidar as two launches; 1: the fused step, one workgroup
    // per 8 cars; 2: the fused step on a persistent grid (racer_batch_set_fused_step)
    int fused_step = 0;
    std::vector<racer_acts*> acts;  // per instance, written by the batched forward
    racer_buf* scratch = nullptr;   // log probabilities then values: instance k's 2 n_k floats at 2 start_k
    std::vector<int64_t> car_start;
    bool count_iterations = false;
    unsigned long long* counter = nullptr;
    int64_t last_iterations = 0;
};

// Physics, then the lidar, of every instance of the batch (each reads its own
// sim's actions and writes its own sim's reward, done, state and obs).
int racer_launch_batch_step(racer_batch* b);
// LDS bytes of the batched fused step, or 0 when it does not apply to the batch.
size_t racer_batch_fused_shmem(const racer_batch* b, int waves_per_car, int* pose_at);
int racer_launch_batch_lidar(racer_batch* b);
int racer_policy_fused_batch(racer_policy* p, const racer_policy_table& tab, int deterministic, int64_t step);
int racer_batch_runner_enqueue(racer_batch* b, racer_policy* p, int deterministic, int64_t t);

// Graphs of `steps` sim steps (physics and lidar) or batched sim steps, for
// the autotuners; replayed with racer_graph_launch.
int racer_graph_sim_steps(racer_sim* sim, int steps, racer_graph** out);
int racer_graph_batch_steps(racer_batch* b, int steps, racer_graph** out);

// ---------------------------------------------------------------- surrogate (sim side)
// Surrogate inputs (11 rows + a constant-1 row) of a sim's cars for the given
// actions, normalised when mean / inv_std are given.
int racer_launch_surrogate_features(racer_sim* sim, const float* act, int64_t act_cols, float* feat, int64_t ldf,
                                    const float* mean, const float* inv_std);
// One-step targets (5 rows) and the no-crash mask from the state before a step.
int racer_launch_surrogate_targets(racer_sim* sim, const float* prev, float* tgt, int64_t ldt, float* mask);
// Sim.step with the surrogate's predicted pose change (RK4 fallback per car), then the lidar.
int racer_launch_surrogate_step(racer_sim* sim, const float* pred, int64_t ldp, const float* out_mean,
                                const float* out_std, float min_clearance, float* fallback);
// The bf16 surrogate's hidden layers packed for the fused step: 76 blocks of
// 512 bf16 (one v_mfma_f32_32x32x16_bf16 A operand each), layer 0 as 4 row
// tiles x 1 k-step (K = 12 -> 16), layers 1 and 2 as 4 x 9 (K = 129 -> 144).
constexpr int RACER_SF_STEPS0 = 1, RACER_SF_STEPS = 9;
constexpr int RACER_SF_BASE1 = 4, RACER_SF_BASE2 = 40, RACER_SF_BLOCKS = 76;
// One fused launch of the bf16 surrogate step (inputs, 3 MFMA layers, the
// float32 output layer, the pose update with RK4 fallback), then the lidar.
// w3: the output layer [5, 129]; stats: the normalisation (37 floats).
int racer_launch_surrogate_fused(racer_sim* sim, const uint16_t* frags, const float* w3, const float* stats,
                                 float min_clearance, float* fallback);
// `steps` of that step in one launch without the lidar (open-loop imagined
// steps), step k with actions + 2 n k; fallback[i] = 1 if car i took the
// exact step on any of them.
int racer_launch_surrogate_imagine(racer_sim* sim, const float* actions, int steps, const uint16_t* frags,
                                   const float* w3, const float* stats, float min_clearance, float* fallback);
// `steps` exact physics steps (RK4) in one launch without the lidar, step k
// with actions + 2 n k.
int racer_launch_physics_steps(racer_sim* sim, const float* actions, int steps);

// ---------------------------------------------------------------- kernels (launchers)
int racer_launch_reset(racer_sim* sim);
int racer_launch_step(racer_sim* sim, float* actions, float* reward, float* done, float* obs, int64_t obs_cols);
int racer_launch_sample(racer_sim* sim, const float* out, int64_t out_cols, const float* log_std, float* actions,
                        float* logp, float* value, int deterministic, int64_t step);
int racer_launch_sample_raw(racer_device* dev, int n, const float* out, int64_t out_cols, const float* log_std,
                            float* actions, int64_t act_cols, float* logp, float* value, int deterministic,
                            int64_t step, int seed);

// GEMM epilogues
// TANH_MLP: racer_tanh_mlp (the bf16 surrogate's hidden layers), bf16 NN products only
enum { RACER_EPI_NONE = 0, RACER_EPI_TANH = 1, RACER_EPI_MASK = 2, RACER_EPI_TANH_MLP = 3 };
// C[M, N] = epi(A[M, K] B[K, N]); row-major with leading dimensions; mask epilogue
// multiplies by (1 - h^2) with h = H[m, n].
int racer_gemm_nn_f32(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                      int64_t ldc, int M, int N, int K, int epi, const float* H, int64_t ldh);
// C[M, N] = A[M, K] B[N, K]^T (weight gradients); K is the long batch dimension,
// split into chunks whose partial products land in a workspace and are then
// added in chunk order. The workspace is `ws` (ws_len floats, at least
// racer_gemm_nt_*_partials) when given, else the device's grow-only scratch.
// Every partial the sum reads is written first, so `ws` needs no clearing.
size_t racer_gemm_nt_f32_partials(int M, int64_t ldc, int K);
int racer_gemm_nt_f32(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                      int64_t ldc, int M, int N, int K, float* ws = nullptr, size_t ws_len = 0);
// bf16 MFMA variants (operands rounded to bf16, float32 accumulation)
// A is a bf16 copy with rows padded to a multiple of 8 values (lda % 8 == 0, zeros beyond K).
int racer_gemm_nn_bf16(racer_device* dev, const uint16_t* A, int64_t lda, const void* B, int64_t ldb, void* C,
                       int64_t ldc, int M, int N, int K, int epi, const float* H = nullptr, int64_t ldh = 0);
size_t racer_gemm_nt_bf16_partials(int M, int64_t ldc, int K);
int racer_gemm_nt_bf16(racer_device* dev, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                       int64_t ldc, int M, int N, int K, float* ws = nullptr, size_t ws_len = 0);
// C[M, N] = (A[M, 3] B[3, N]) (1 - H^2), each output fmaf over k = 0, 1, 2
// from 0: racer_gemm_nn_f32 with K = 3 and the mask epilogue, bit for bit;
// out16: C stored in bf16 (nearest even).
int racer_mask_k3(racer_device* dev, const float* A, const float* B, int64_t ldb, void* C, bool out16, int64_t ldc,
                  int M, int N, const float* H, int64_t ldh);
// racer_gemm_nt_bf16 with A and / or B already in bf16 (a16, b16; the MFMA
// weight-gradient kernel only: rows and K aligned to 4 elements); the same
// bits as racer_gemm_nt_bf16 on their float32 values.
int racer_gemm_nt_bf16_io(racer_device* dev, const void* A, bool a16, int64_t lda, const void* B, bool b16,
                          int64_t ldb, float* C, int64_t ldc, int M, int N, int K, float* ws, size_t ws_len);
// racer_gemm_nn_bf16 with the mask epilogue, B in bf16 and C stored in bf16
// (rounded to nearest even): the bits its bf16 consumers would load from the
// float32 product.
int racer_gemm_nn_bf16_mask16(racer_device* dev, const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb,
                              uint16_t* C, int64_t ldc, int M, int N, int K, const float* H, int64_t ldh);
// dst[r, c] = bf16(src[r, c]) for c < cols, 0 for cols <= c < ld_dst.
int racer_pack_bf16(racer_device* dev, const float* src, int64_t ld_src, uint16_t* dst, int64_t ld_dst, int rows,
                    int cols);
inline int64_t racer_pad8(int64_t v) { return (v + 7) / 8 * 8; }
inline int64_t racer_pad32(int64_t v) { return (v + 31) / 32 * 32; }
// Fragment-order bf16 weight copies for the fused forward: elements needed for
// `cols` columns, and the packing of a [rows, cols] float32 matrix (rows <= 256).
size_t racer_fragment_elems(int cols);
int racer_pack_fragments(racer_device* dev, const float* src, int64_t ld_src, int rows, int cols, uint16_t* dst);
// The fused bf16 policy forward (and optionally Sim.sample) over `cols` columns of x.
// Its input is read through a buffer descriptor with 32-bit byte offsets: the
// span of IN rows of stride x_cols must stay below 4 GiB.
inline bool racer_fused_reach(int IN, int64_t x_cols, int64_t cols) {
    return x_cols >= 0 && cols >= 0 && ((uint64_t)(IN - 1) * (uint64_t)x_cols + (uint64_t)cols) * 4 < 0xFFFFFF00ull;
}
int racer_policy_fused_setup();
// The float32 forward (and optionally Sim.sample) in one kernel on K = 1
// float32 MFMA, bitwise the float32 products; for H <= 256 within the LDS.
bool racer_policy_f32_fits(int H, int IN);
int racer_policy_f32_setup();
size_t racer_f32_fragment_floats(int K);
int racer_pack_f32_fragments(racer_device* dev, const float* src, int64_t ld_src, int rows, int cols, float* dst);
int racer_policy_f32(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, bool sample,
                     float* act, float* logp, float* value, int deterministic, int64_t step, int seed);
// The forward of the PPO update's minibatches from bf16 inputs (k_policy_tiled
// reading the bf16 minibatch copy): bitwise the forward of their float32
// values. racer_tiled_x16_applies: a bf16 policy of this shape takes the
// 64-column tiled kernel and the shapes are aligned for it.
bool racer_tiled_x16_applies(int precision, int hidden, int64_t IN, int64_t x_cols, int64_t cols);
int racer_policy_tiled_x16(racer_policy* p, const uint16_t* x, int64_t x_cols, racer_acts* a, int cols);
// kernel: RACER_ACT_AUTO picks by the launch's size; RACER_ACT_FUSED the
// 16-column ring kernel, RACER_ACT_WIDE its wide variant, RACER_ACT_TILED the
// 64-column tiled kernel (without the sample only), where they apply.
int racer_policy_fused(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, bool sample,
                       float* act, float* logp, float* value, int deterministic, int64_t step, int seed,
                       int kernel = RACER_ACT_AUTO);

int racer_policy_forward_raw(racer_policy* pol, const float* x, int64_t x_cols, racer_acts* acts, int cols);
int racer_policy_act(racer_policy* p, const float* x, int64_t x_cols, racer_acts* a, int cols, float* act,
                     float* logp, float* value, int deterministic, int64_t step, int seed);
// Re-derive the bf16 weight copies if an update made them stale (no-op otherwise).
int racer_policy_refresh(racer_policy* p);
// Device-to-device copy of count floats on the device's stream (own kernel).
int racer_copy_raw(racer_device* dev, float* dst, const float* src, size_t count);
// The work of Trainer.rollout up to the device-side reward sum (no host read).
int racer_rollout_enqueue(racer_trainer* tr, racer_sim* sim, racer_policy* p, racer_acts* acts, int64_t step0);
// Reads the reward sum of the last rollout (synchronizes) as a mean per step.
int racer_trainer_mean_reward(racer_trainer* tr, float* mean_reward);
int racer_fill(racer_device* dev, float* p, size_t n, float v);
