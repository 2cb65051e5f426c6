// Device, buffers, errors and the launch profiler of libracer_hip
// (the reference's Device, racer/device.mojo:35-134).
#include <cstring>

#include "racer_internal.h"

static thread_local std::string g_error;

void racer_set_error(const std::string& msg) { g_error = msg; }
void racer_clear_error() { g_error.clear(); }

extern "C" const char* racer_last_error(void) { return g_error.c_str(); }
extern "C" const char* racer_version(void) { return "racer_hip 1.0 (gfx950)"; }

extern "C" void racer_config_default(racer_config* cfg) {
    cfg->beams = RACER_DEFAULT_BEAMS;
    cfg->hidden = RACER_DEFAULT_H;
    cfg->rollout = RACER_DEFAULT_ROLLOUT;
    cfg->epochs = RACER_DEFAULT_EPOCHS;
    cfg->minibatches = RACER_DEFAULT_MINIBATCHES;
    cfg->precision = 0;
    cfg->heading_noise = 0.0f;
}

int racer_launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        racer_set_error(std::string(what) + ": " + hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int racer_check_view(const racer_view& v, int64_t rows, int64_t cols, const char* what) {
    if (!v.buf) {
        racer_set_error(std::string(what) + ": null buffer");
        return 1;
    }
    if (rows >= 0 && v.rows != rows) {
        racer_set_error(std::string(what) + ": expected " + std::to_string(rows) + " rows, got " +
                        std::to_string(v.rows));
        return 1;
    }
    if (cols >= 0 && v.cols != cols) {
        racer_set_error(std::string(what) + ": expected " + std::to_string(cols) + " columns, got " +
                        std::to_string(v.cols));
        return 1;
    }
    if (v.rows < 0 || v.cols < 0 || v.offset + (size_t)(v.rows * v.cols) > v.buf->len) {
        racer_set_error(std::string(what) + ": view exceeds its buffer");
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------- device
// The fixed scratch of the allocation-free entry points: one copy per GPU, at
// its own address there, held until the process ends. Host code never passes
// this name to a kernel (on the host it is only a stand-in for the symbol);
// racer_device_create looks up the GPU's address and keeps it in
// racer_device::fixed_scratch.
static __device__ __align__(256) float scratch[RACER_FIXED_SCRATCH_FLOATS];

extern "C" int racer_device_create(int gpu, racer_device** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(gpu, "libracer_hip runs on the GPU only; the CPU device is not provided");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    RACER_REQUIRE(e == hipSuccess && count > 0, "no HIP device is available");
    auto* dev = new racer_device();
    auto drop = [dev] {
        if (dev->sums) (void)hipFree(dev->sums);
        if (dev->stream) (void)hipStreamDestroy(dev->stream);
        delete dev;
    };
    RACER_TRY_OR(hipGetDevice(&dev->id), drop());
    hipDeviceProp_t prop;
    RACER_TRY_OR(hipGetDeviceProperties(&prop, dev->id), drop());
    dev->name = prop.name;
    if (prop.gcnArchName[0]) dev->name += std::string(" (") + prop.gcnArchName + ")";
    dev->cus = prop.multiProcessorCount;
    RACER_TRY_OR(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking), drop());
    RACER_TRY_OR(hipMalloc(&dev->sums, RACER_SUM_PARTIALS * sizeof(double)), drop());
    // with this device current: its copy of the fixed scratch array
    void* fixed = nullptr;
    RACER_TRY_OR(hipGetSymbolAddress(&fixed, HIP_SYMBOL(scratch)), drop());
    if (!fixed) {
        drop();
        racer_set_error("the fixed scratch array has no device address");
        return 1;
    }
    dev->fixed_scratch = static_cast<float*>(fixed);
    *out = dev;
    return 0;
}

void racer_device_retain(racer_device* dev) {
    if (dev) ++dev->refs;
}

void racer_device_release(racer_device* dev) {
    if (!dev || --dev->refs > 0) return;
    (void)hipStreamSynchronize(dev->stream);
    for (auto& p : dev->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : dev->event_pool) (void)hipEventDestroy(e);
    if (dev->scratch) (void)hipFree(dev->scratch);
    if (dev->graph_base) (void)hipFree(dev->graph_base);
    if (dev->sums) (void)hipFree(dev->sums);
    (void)hipStreamDestroy(dev->stream);
    delete dev;
}

extern "C" void racer_device_destroy(racer_device* dev) { racer_device_release(dev); }

extern "C" const char* racer_device_name(racer_device* dev) { return dev ? dev->name.c_str() : ""; }

extern "C" int racer_sync(racer_device* dev) {
    racer_clear_error();
    RACER_TRY(hipStreamSynchronize(dev->stream));
    return 0;
}

int racer_partials(racer_device* dev, size_t floats, float* ws, size_t ws_len, float** out) {
    if (ws) {
        RACER_REQUIRE(floats <= ws_len, "the split-K partial products exceed the workspace");
        *out = ws;
        return 0;
    }
    if (racer_ensure_scratch(dev, floats)) return 1;
    *out = dev->scratch;
    return 0;
}

int racer_ensure_scratch(racer_device* dev, size_t floats) {
    if (dev->scratch_len >= floats) return 0;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    if (dev->scratch) RACER_TRY(hipFree(dev->scratch));
    dev->scratch = nullptr;
    RACER_TRY(hipMalloc(&dev->scratch, floats * sizeof(float)));
    dev->scratch_len = floats;
    return 0;
}

// ---------------------------------------------------------------- buffers
__global__ void k_fill(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_copy4(float4* __restrict__ d, const float4* __restrict__ s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

__global__ void k_copy1(float* __restrict__ d, const float* __restrict__ s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

int racer_fill(racer_device* dev, float* p, size_t n, float v) {
    if (n == 0) return 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, dev->stream, p, n, v);
    return racer_launch_check("fill");
}

extern "C" int racer_alloc(racer_device* dev, size_t count, float fill, racer_buf** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(dev, "null device");
    auto* b = new racer_buf();
    b->dev = dev;
    racer_device_retain(dev);
    b->len = count;
    if (count) {
        hipError_t e = hipMalloc(&b->ptr, count * sizeof(float));
        if (e != hipSuccess) {
            delete b;
            racer_device_release(dev);
            racer_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return 1;
        }
        if (racer_fill(dev, b->ptr, count, fill)) return 1;
        RACER_TRY(hipStreamSynchronize(dev->stream));
    }
    *out = b;
    return 0;
}

extern "C" void racer_free(racer_buf* buf) {
    if (!buf) return;
    if (buf->ptr) {
        (void)hipStreamSynchronize(buf->dev->stream);
        (void)hipFree(buf->ptr);
    }
    racer_device* dev = buf->dev;
    delete buf;
    racer_device_release(dev);
}

extern "C" size_t racer_buf_len(const racer_buf* buf) { return buf ? buf->len : 0; }
extern "C" void* racer_buf_ptr(racer_buf* buf) { return buf ? buf->ptr : nullptr; }

extern "C" int racer_read_range(racer_buf* buf, size_t offset, size_t count, float* host) {
    racer_clear_error();
    RACER_REQUIRE(buf && offset + count <= buf->len, "read exceeds the buffer");
    if (!count) return 0;
    RACER_TRY(hipMemcpyAsync(host, buf->ptr + offset, count * sizeof(float), hipMemcpyDeviceToHost, buf->dev->stream));
    RACER_TRY(hipStreamSynchronize(buf->dev->stream));
    return 0;
}

extern "C" int racer_write_range(racer_buf* buf, size_t offset, size_t count, const float* host) {
    racer_clear_error();
    RACER_REQUIRE(buf && offset + count <= buf->len, "write exceeds the buffer");
    if (!count) return 0;
    ++buf->gen;
    RACER_TRY(hipMemcpyAsync(buf->ptr + offset, host, count * sizeof(float), hipMemcpyHostToDevice, buf->dev->stream));
    RACER_TRY(hipStreamSynchronize(buf->dev->stream));
    return 0;
}

extern "C" int racer_read(racer_buf* buf, float* host) { return racer_read_range(buf, 0, buf ? buf->len : 0, host); }
extern "C" int racer_write(racer_buf* buf, const float* host) {
    return racer_write_range(buf, 0, buf ? buf->len : 0, host);
}

extern "C" int racer_upload(racer_device* dev, const float* host, size_t count, racer_buf** out) {
    if (racer_alloc(dev, count, 0.0f, out)) return 1;
    return racer_write(*out, host);
}

extern "C" int racer_copy(racer_buf* dst, size_t dst_offset, racer_buf* src, size_t src_offset, size_t count) {
    racer_clear_error();
    RACER_REQUIRE(dst && src && dst_offset + count <= dst->len && src_offset + count <= src->len,
                  "copy exceeds a buffer");
    if (!count) return 0;
    ++dst->gen;
    float* d = dst->ptr + dst_offset;
    const float* s = src->ptr + src_offset;
    if (d < s + count && s < d + count) {  // overlapping ranges of one buffer: the runtime's copy
        RACER_TRY(hipMemcpyAsync(d, s, count * sizeof(float), hipMemcpyDeviceToDevice, dst->dev->stream));
        return 0;
    }
    return racer_copy_raw(dst->dev, d, s, count);
}

// Own kernel for disjoint ranges: exactly 4 count bytes read and written (a
// known amount of traffic against which HBM counters can be calibrated), and
// a kernel node rather than a copy node when captured into a graph.
int racer_copy_raw(racer_device* dev, float* d, const float* s, size_t count) {
    if (!count) return 0;
    racer_timed t(dev, "copy");
    bool vec = ((uintptr_t)d % 16 == 0) && ((uintptr_t)s % 16 == 0) && count % 4 == 0;
    size_t items = vec ? count / 4 : count;
    size_t blocks = (items + 255) / 256;
    if (blocks > 32768) blocks = 32768;
    if (vec)
        hipLaunchKernelGGL(k_copy4, dim3((unsigned)blocks), dim3(256), 0, dev->stream, (float4*)d,
                           (const float4*)s, items);
    else
        hipLaunchKernelGGL(k_copy1, dim3((unsigned)blocks), dim3(256), 0, dev->stream, d, s, items);
    return racer_launch_check("copy");
}

// ---------------------------------------------------------------- profiler
static hipEvent_t take_event(racer_device* dev) {
    if (!dev->event_pool.empty()) {
        hipEvent_t e = dev->event_pool.back();
        dev->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static int kernel_index(racer_device* dev, const char* name) {
    for (size_t i = 0; i < dev->samples.size(); ++i)
        if (dev->samples[i].name == name) return (int)i;
    dev->samples.push_back({name, 0, 0.0});
    return (int)dev->samples.size() - 1;
}

racer_timed::racer_timed(racer_device* d, const char* name) : dev(d) {
    if (!dev || !dev->profiling || dev->capturing) return;
    if (!dev->profile_only.empty() && dev->profile_only != name) return;
    start = take_event(dev);
    stop = take_event(dev);
    // a launch whose events cannot be made or recorded is left out of the samples
    if (start && stop && hipEventRecord(start, dev->stream) == hipSuccess) {
        kernel = kernel_index(dev, name);
        return;
    }
    if (start) dev->event_pool.push_back(start);
    if (stop) dev->event_pool.push_back(stop);
}

racer_timed::~racer_timed() {
    if (kernel < 0) return;
    if (hipEventRecord(stop, dev->stream) == hipSuccess) {
        dev->pending.push_back({kernel, start, stop});
        return;
    }
    dev->event_pool.push_back(start);
    dev->event_pool.push_back(stop);
}

static void resolve_pending(racer_device* dev) {
    if (dev->pending.empty()) return;
    (void)hipStreamSynchronize(dev->stream);  // an error shows in the elapsed times below
    for (auto& p : dev->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            dev->samples[p.kernel].launches += 1;
            dev->samples[p.kernel].total_ms += ms;
        }
        dev->event_pool.push_back(p.start);
        dev->event_pool.push_back(p.stop);
    }
    dev->pending.clear();
}

extern "C" int racer_profile_enable(racer_device* dev, int on) {
    resolve_pending(dev);
    dev->profiling = on != 0;
    return 0;
}

extern "C" int racer_profile_count(racer_device* dev) {
    resolve_pending(dev);
    return (int)dev->samples.size();
}

extern "C" int racer_profile_get(racer_device* dev, int i, const char** name, int64_t* launches, double* total_ms) {
    resolve_pending(dev);
    racer_clear_error();
    RACER_REQUIRE(i >= 0 && i < (int)dev->samples.size(), "no such profile entry");
    *name = dev->samples[i].name.c_str();
    *launches = dev->samples[i].launches;
    *total_ms = dev->samples[i].total_ms;
    return 0;
}

extern "C" int racer_profile_reset(racer_device* dev) {
    resolve_pending(dev);
    dev->samples.clear();
    return 0;
}

extern "C" int racer_profile_only(racer_device* dev, const char* name) {
    resolve_pending(dev);
    dev->profile_only = name ? name : "";
    return 0;
}
