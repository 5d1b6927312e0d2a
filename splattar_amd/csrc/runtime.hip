// C-ABI runtime for the MI355X surfel training step (include/gs_hip.h).
//
// Host-side orchestration only: every handle allocates its device buffers
// once at capacity (as crates/gs-kernels/src/rasterizer.rs:1-4: no
// reallocation mid-training), validates shapes before any launch, and enqueues the
// gfx950 kernels of the sibling .hip files on the caller's stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gs_internal.h"

namespace gs {

// ------------------------------------------------------------------ errors

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

gs_status fail(gs_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

gs_status hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? GS_ERR_OOM : GS_ERR_HIP;
}

#define GS_CHECK(expr)                        \
    do {                                      \
        gs_status s_ = (expr);                \
        if (s_ != GS_OK) return s_;           \
    } while (0)

#define GS_REQUIRE(cond, msg)                                    \
    do {                                                         \
        if (!(cond)) return ::gs::fail(GS_ERR_INVALID, (msg));   \
    } while (0)

// --------------------------------------------------------------- allocation

template <typename T>
static gs_status dalloc(T** out, size_t count, const char* what) {
    *out = nullptr;
    if (count == 0) count = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, what);
    // The zero fill runs on the null stream, which is not ordered with the
    // context's non-blocking stream: wait for it, or a kernel enqueued right
    // after the allocation can be overwritten by the late fill.
    e = hipMemset(p, 0, count * sizeof(T));
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, what);
    }
    *out = static_cast<T*>(p);
    return GS_OK;
}

template <typename T>
static void dfree(T*& p) {
    if (p) (void)hipFree(static_cast<void*>(p));
    p = nullptr;
}

static hipStream_t as_stream(void* s, hipStream_t fallback) {
    return s ? static_cast<hipStream_t>(s) : fallback;
}

// Camera-to-world rotation from an unnormalized quaternion (math.rs:14-34),
// evaluated in double and rounded once: the same f32 matrix feeds every
// kernel and the CPU mirror.
static void quat_to_rot(const float q[4], float rot[9]) {
    const double n = std::sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] +
                               (double)q[3] * q[3]);
    const double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
    const double c[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)},
                            {2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)},
                            {2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)}};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) rot[3 * k + j] = (float)c[k][j];
}

static CamParams make_cam(const gs_raster_camera& c, uint32_t w, uint32_t h, uint32_t tile = kTile) {
    CamParams p;
    std::memset(&p, 0, sizeof(p));
    quat_to_rot(c.quat, p.rot);
    p.center[0] = c.center[0];
    p.center[1] = c.center[1];
    p.center[2] = c.center[2];
    p.focal = c.focal;
    p.cx = (float)w * 0.5f;
    p.cy = (float)h * 0.5f;
    p.width = w;
    p.height = h;
    p.tiles_x = div_up(w, tile);
    p.tiles_y = div_up(h, tile);
    p.tile_size = tile;
    p.rtiles_x = div_up(w, kTile);
    p.rtiles_y = div_up(h, kTile);
    p.n_surfels = 0xFFFFFFFFu;
    p.sh_degree = c.sh_degree;
    return p;
}

static gs_status check_camera(const gs_raster_camera* c) {
    GS_REQUIRE(c != nullptr, "camera is null");
    GS_REQUIRE(std::isfinite(c->focal) && c->focal > 0.0f, "camera focal must be finite and > 0");
    GS_REQUIRE(c->sh_degree <= 3, "sh_degree must be <= 3");
    const double qn = (double)c->quat[0] * c->quat[0] + (double)c->quat[1] * c->quat[1] +
                      (double)c->quat[2] * c->quat[2] + (double)c->quat[3] * c->quat[3];
    GS_REQUIRE(qn > 0.0 && std::isfinite(qn), "camera quaternion must be non-zero and finite");
    return GS_OK;
}

}  // namespace gs

using namespace gs;

// =================================================================== context

struct gs_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    // Device address of this device's `scratch` array (the no-allocation
    // scan's partition sums).
    uint32_t* noalloc_scratch = nullptr;
};

namespace gs {

// Scratch of gs_scan_encode_noalloc: the partition sums of a scan of up to
// GS_SCAN_NOALLOC_MAX_ITEMS items, in a module-scope device array so that the
// entry point allocates nothing. Each device holds its own copy until the
// process ends; every call on a device shares it.
constexpr size_t kNoallocScratchWords = (scan_scratch_size(GS_SCAN_NOALLOC_MAX_ITEMS) + 255) / 256 * 64;
static __device__ __align__(256) uint32_t scratch[kNoallocScratchWords];

// The host-side name `scratch` is not a device address: the address is
// looked up once per device (with that device current) and only that
// pointer reaches kernels.
static std::mutex g_noalloc_mu;
static std::map<int, uint32_t*> g_noalloc_ptr;

static gs_status noalloc_scratch_for_current_device(int device, uint32_t** out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    if (it == g_noalloc_ptr.end()) {
        uint32_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        it = g_noalloc_ptr.emplace(device, ptr).first;
    }
    *out = it->second;
    return GS_OK;
}

}  // namespace gs

extern "C" {

const char* gs_last_error(void) { return g_last_error.c_str(); }
int gs_abi_version(void) { return GS_ABI_VERSION; }
const char* gs_build_info(void) { return "splattar_amd gfx950 HIP " __DATE__; }

gs_status gs_ctx_create(int device_index, gs_ctx* out) {
    GS_REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(GS_ERR_NO_DEVICE, "no HIP device available");
    GS_REQUIRE(device_index >= 0 && device_index < count, "device index out of range");
    GS_HIP(hipSetDevice(device_index));
    auto ctx = std::make_unique<gs_context>();
    ctx->device = device_index;
    GS_HIP(hipGetDeviceProperties(&ctx->props, device_index));
    GS_CHECK(noalloc_scratch_for_current_device(device_index, &ctx->noalloc_scratch));
    GS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    *out = ctx.release();
    return GS_OK;
}

void gs_ctx_destroy(gs_ctx ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

void* gs_ctx_stream(gs_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

gs_status gs_ctx_synchronize(gs_ctx ctx) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    GS_HIP(hipStreamSynchronize(ctx->stream));
    GS_HIP(hipDeviceSynchronize());
    return GS_OK;
}

gs_status gs_ctx_device_info(gs_ctx ctx, char* name, size_t name_len, int* compute_units) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", ctx->props.name, ctx->props.gcnArchName);
    }
    if (compute_units) *compute_units = ctx->props.multiProcessorCount;
    return GS_OK;
}

gs_status gs_buffer_alloc(gs_ctx ctx, size_t bytes, void** out) {
    GS_REQUIRE(ctx != nullptr && out != nullptr, "null argument");
    uint8_t* p = nullptr;
    GS_CHECK(dalloc(&p, bytes, "gs_buffer_alloc"));
    *out = p;
    return GS_OK;
}

gs_status gs_buffer_free(gs_ctx ctx, void* ptr) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    if (ptr) GS_HIP(hipFree(ptr));
    return GS_OK;
}

gs_status gs_write_buffer(gs_ctx ctx, void* dst, const void* src, size_t bytes) {
    GS_REQUIRE(ctx != nullptr && dst != nullptr && (src != nullptr || bytes == 0), "null argument");
    if (bytes == 0) return GS_OK;
    // Pageable source: the copy is complete (staged) when the call returns,
    // so the caller may reuse `src`, as with queue.write_buffer.
    GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    GS_HIP(hipStreamSynchronize(ctx->stream));
    return GS_OK;
}

gs_status gs_read_buffer(gs_ctx ctx, void* dst, const void* src, size_t bytes) {
    GS_REQUIRE(ctx != nullptr && dst != nullptr && (src != nullptr || bytes == 0), "null argument");
    if (bytes == 0) return GS_OK;
    GS_HIP(hipStreamSynchronize(ctx->stream));
    GS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GS_OK;
}

gs_status gs_clear_buffer(gs_ctx ctx, void* stream, void* dst, size_t bytes) {
    GS_REQUIRE(ctx != nullptr && dst != nullptr, "null argument");
    GS_REQUIRE(bytes % 4 == 0, "clear size must be a multiple of 4 bytes");
    GS_HIP(launch_fill_u32(as_stream(stream, ctx->stream), static_cast<uint32_t*>(dst), 0u, bytes / 4));
    return GS_OK;
}

}  // extern "C"

// ================================================================ PrefixSum

struct gs_scan_s {
    gs_ctx ctx = nullptr;
    uint32_t capacity = 0;
    uint32_t* data = nullptr;
    uint32_t* total = nullptr;
    uint8_t* scratch = nullptr;
    ~gs_scan_s() {
        dfree(data);
        dfree(total);
        dfree(scratch);
    }
};

extern "C" {

gs_status gs_scan_create(gs_ctx ctx, uint32_t capacity, gs_scan* out) {
    GS_REQUIRE(ctx && out, "null argument");
    *out = nullptr;
    auto s = std::make_unique<gs_scan_s>();
    s->ctx = ctx;
    s->capacity = capacity;
    GS_CHECK(dalloc(&s->data, capacity, "scan data"));
    GS_CHECK(dalloc(&s->total, 1, "scan total"));
    GS_CHECK(dalloc(&s->scratch, scan_scratch_bytes(capacity), "scan scratch"));
    *out = s.release();
    return GS_OK;
}

void gs_scan_destroy(gs_scan s) { delete s; }
uint32_t* gs_scan_data(gs_scan s) { return s ? s->data : nullptr; }
uint32_t* gs_scan_total(gs_scan s) { return s ? s->total : nullptr; }

gs_status gs_scan_encode(gs_scan s, void* stream, uint32_t n) {
    GS_REQUIRE(s != nullptr, "scan is null");
    if (n > s->capacity) return fail(GS_ERR_CAPACITY, "scan: n exceeds capacity");  // prefix_sum.rs:135
    GS_HIP(launch_exclusive_scan(as_stream(stream, s->ctx->stream), s->data, s->data, n, s->total,
                                 s->scratch));
    return GS_OK;
}

gs_status gs_scan_encode_noalloc(gs_ctx ctx, void* stream, uint32_t* data, uint32_t n, uint32_t* total) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    GS_REQUIRE(total != nullptr && (data != nullptr || n == 0), "null argument");
    if (n > GS_SCAN_NOALLOC_MAX_ITEMS)
        return fail(GS_ERR_CAPACITY, "gs_scan_encode_noalloc: n = " + std::to_string(n) +
                                         " exceeds the no-allocation limit of " +
                                         std::to_string(GS_SCAN_NOALLOC_MAX_ITEMS) + " items");
    GS_REQUIRE(ctx->noalloc_scratch != nullptr, "no-allocation scratch not set up for this context");
    // The same kernels as gs_scan_encode; the partition sums live in this
    // device's `scratch` array. The reduce kernel writes every sum the later
    // kernels read, so nothing in it needs clearing between calls.
    GS_HIP(launch_exclusive_scan(as_stream(stream, ctx->stream), data, data, n, total, ctx->noalloc_scratch));
    return GS_OK;
}

}  // extern "C"

// ============================================================== RadixSorter

struct gs_sorter_s {
    gs_ctx ctx = nullptr;
    uint32_t capacity = 0;
    uint32_t key_bits = 64;
    void* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* counts = nullptr;
    uint8_t* scratch = nullptr;
    ~gs_sorter_s() {
        for (int i = 0; i < 2; ++i) {
            if (keys[i]) (void)hipFree(keys[i]);
            dfree(vals[i]);
        }
        dfree(counts);
        dfree(scratch);
    }
};

extern "C" {

gs_status gs_sorter_create(gs_ctx ctx, uint32_t capacity, uint32_t key_bits, gs_sorter* out) {
    GS_REQUIRE(ctx && out, "null argument");
    GS_REQUIRE(key_bits == 32 || key_bits == 64, "key_bits must be 32 or 64");
    GS_REQUIRE(capacity < (1u << 30), "sorter capacity must be < 2^30");
    *out = nullptr;
    auto s = std::make_unique<gs_sorter_s>();
    s->ctx = ctx;
    s->capacity = capacity;
    s->key_bits = key_bits;
    for (int i = 0; i < 2; ++i) {
        uint8_t* k = nullptr;
        GS_CHECK(dalloc(&k, (size_t)(capacity ? capacity : 1) * (key_bits / 8), "sorter keys"));
        s->keys[i] = k;
        GS_CHECK(dalloc(&s->vals[i], capacity, "sorter payloads"));
    }
    GS_CHECK(dalloc(&s->counts, 1, "sorter counts"));
    GS_CHECK(dalloc(&s->scratch, sort_scratch_bytes(capacity), "sorter scratch"));
    *out = s.release();
    return GS_OK;
}

void gs_sorter_destroy(gs_sorter s) { delete s; }
void* gs_sorter_keys(gs_sorter s) { return s ? s->keys[0] : nullptr; }
uint32_t* gs_sorter_payloads(gs_sorter s) { return s ? s->vals[0] : nullptr; }
uint32_t* gs_sorter_counts(gs_sorter s) { return s ? s->counts : nullptr; }

gs_status gs_sorter_encode(gs_sorter s, void* stream, uint32_t begin_bit, uint32_t end_bit,
                           uint32_t digit_bits) {
    GS_REQUIRE(s != nullptr, "sorter is null");
    GS_REQUIRE(digit_bits == 4 || digit_bits == 8, "digit_bits must be 4 or 8");
    GS_REQUIRE(begin_bit < end_bit && end_bit <= s->key_bits, "bad bit range");
    hipStream_t st = as_stream(stream, s->ctx->stream);
    int alt = 0;
    if (s->key_bits == 64)
        GS_HIP(launch_radix_sort_u64(st, static_cast<uint64_t*>(s->keys[0]), s->vals[0],
                                     static_cast<uint64_t*>(s->keys[1]), s->vals[1], s->counts,
                                     s->capacity, begin_bit, end_bit, digit_bits, s->scratch, &alt));
    else
        GS_HIP(launch_radix_sort_u32(st, static_cast<uint32_t*>(s->keys[0]), s->vals[0],
                                     static_cast<uint32_t*>(s->keys[1]), s->vals[1], s->counts,
                                     s->capacity, begin_bit, end_bit, digit_bits, s->scratch, &alt));
    if (alt)  // odd pass count: bring the result back (sort.rs:7: results land in keys())
        GS_HIP(launch_copy_pairs(st, s->keys[1], s->vals[1], s->keys[0], s->vals[0], s->key_bits / 8,
                                 s->counts, s->capacity));
    return GS_OK;
}

}  // extern "C"

// =============================================================== TileBinner

struct gs_binner_s {
    gs_ctx ctx = nullptr;
    uint32_t max_items = 0, max_entries = 0, num_tiles = 0;
    uint4* rects = nullptr;
    uint32_t* depths = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* total = nullptr;
    uint32_t* count = nullptr;
    uint32_t* overflow = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* ranges = nullptr;
    uint8_t* scan_scratch = nullptr;
    uint8_t* sort_scratch = nullptr;
    CamParams* cam = nullptr;
    uint32_t cam_tiles_x = 0;
    ~gs_binner_s() {
        dfree(rects);
        dfree(depths);
        dfree(offsets);
        dfree(total);
        dfree(count);
        dfree(overflow);
        for (int i = 0; i < 2; ++i) {
            dfree(keys[i]);
            dfree(vals[i]);
        }
        dfree(ranges);
        dfree(scan_scratch);
        dfree(sort_scratch);
        dfree(cam);
    }
};

static uint32_t bits_for(uint32_t v) {
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < (uint64_t)v) ++b;
    return b;
}

extern "C" {

gs_status gs_binner_create(gs_ctx ctx, uint32_t max_items, uint32_t max_entries, uint32_t num_tiles,
                           gs_binner* out) {
    GS_REQUIRE(ctx && out, "null argument");
    GS_REQUIRE(max_entries < (1u << 30), "max_entries must be < 2^30");
    GS_REQUIRE(num_tiles > 0, "num_tiles must be > 0");
    *out = nullptr;
    auto b = std::make_unique<gs_binner_s>();
    b->ctx = ctx;
    b->max_items = max_items;
    b->max_entries = max_entries;
    b->num_tiles = num_tiles;
    GS_CHECK(dalloc(&b->rects, max_items, "binner rects"));
    GS_CHECK(dalloc(&b->depths, max_items, "binner depths"));
    GS_CHECK(dalloc(&b->offsets, max_items, "binner offsets"));
    GS_CHECK(dalloc(&b->total, 1, "binner total"));
    GS_CHECK(dalloc(&b->count, 1, "binner count"));
    GS_CHECK(dalloc(&b->overflow, 1, "binner overflow"));
    for (int i = 0; i < 2; ++i) {
        GS_CHECK(dalloc(&b->keys[i], max_entries, "binner keys"));
        GS_CHECK(dalloc(&b->vals[i], max_entries, "binner payloads"));
    }
    GS_CHECK(dalloc(&b->ranges, (size_t)num_tiles * 2, "binner ranges"));
    GS_CHECK(dalloc(&b->scan_scratch, scan_scratch_bytes(max_items), "binner scan scratch"));
    GS_CHECK(dalloc(&b->sort_scratch, sort_scratch_bytes(max_entries), "binner sort scratch"));
    GS_CHECK(dalloc(&b->cam, 1, "binner cam"));
    *out = b.release();
    return GS_OK;
}

void gs_binner_destroy(gs_binner b) { delete b; }
uint32_t* gs_binner_rects(gs_binner b) { return b ? reinterpret_cast<uint32_t*>(b->rects) : nullptr; }
uint32_t* gs_binner_depths(gs_binner b) { return b ? b->depths : nullptr; }
uint32_t* gs_binner_ranges(gs_binner b) { return b ? b->ranges : nullptr; }
uint32_t* gs_binner_sorted_entries(gs_binner b) { return b ? b->vals[0] : nullptr; }
uint64_t* gs_binner_sorted_keys(gs_binner b) { return b ? b->keys[0] : nullptr; }
uint32_t* gs_binner_entry_count(gs_binner b) { return b ? b->count : nullptr; }

gs_status gs_binner_encode(gs_binner b, void* stream, uint32_t num_items, uint32_t tiles_x,
                           uint32_t digit_bits) {
    GS_REQUIRE(b != nullptr, "binner is null");
    if (num_items > b->max_items) return fail(GS_ERR_CAPACITY, "binner: num_items > max_items");
    // binning.rs:181
    GS_REQUIRE(tiles_x > 0 && b->num_tiles % tiles_x == 0, "binner: num_tiles must be a multiple of tiles_x");
    GS_REQUIRE(digit_bits == 4 || digit_bits == 8, "digit_bits must be 4 or 8");
    hipStream_t st = as_stream(stream, b->ctx->stream);
    if (b->cam_tiles_x != tiles_x) {
        CamParams c;
        std::memset(&c, 0, sizeof(c));
        c.tiles_x = c.rtiles_x = tiles_x;
        c.tiles_y = c.rtiles_y = b->num_tiles / tiles_x;
        c.tile_size = kTile;
        c.n_surfels = 0xFFFFFFFFu;
        GS_HIP(hipMemcpyAsync(b->cam, &c, sizeof(c), hipMemcpyHostToDevice, st));
        GS_HIP(hipStreamSynchronize(st));
        b->cam_tiles_x = tiles_x;
    }
    GS_HIP(launch_count_tiles(st, num_items, b->rects, b->offsets));
    GS_HIP(launch_exclusive_scan(st, b->offsets, b->offsets, num_items, b->total, b->scan_scratch));
    GS_HIP(launch_emit_keys(st, num_items, b->rects, b->depths, b->offsets, b->cam,
                            num_items ? num_items : 1, tiles_x, b->keys[0], b->vals[0],
                            b->max_entries, b->total, b->count, b->overflow));
    const uint32_t end_bit = 32 + bits_for(b->num_tiles);
    int alt = 0;
    GS_HIP(launch_radix_sort_u64(st, b->keys[0], b->vals[0], b->keys[1], b->vals[1], b->count,
                                 b->max_entries, 0, end_bit, digit_bits, b->sort_scratch, &alt));
    if (alt)
        GS_HIP(launch_copy_pairs(st, b->keys[1], b->vals[1], b->keys[0], b->vals[0], 8, b->count,
                                 b->max_entries));
    GS_HIP(launch_fill_u32(st, b->ranges, 0u, (size_t)b->num_tiles * 2));  // binning.rs:228
    GS_HIP(launch_tile_ranges(st, b->keys[0], b->count, b->max_entries, b->ranges));
    return GS_OK;
}

gs_status gs_binner_overflowed(gs_binner b, int* overflowed) {
    GS_REQUIRE(b && overflowed, "null argument");
    uint32_t f = 0;
    GS_HIP(hipStreamSynchronize(b->ctx->stream));
    GS_HIP(hipMemcpy(&f, b->overflow, 4, hipMemcpyDeviceToHost));
    *overflowed = f ? 1 : 0;
    const uint32_t zero = 0;
    GS_HIP(hipMemcpy(b->overflow, &zero, 4, hipMemcpyHostToDevice));
    return GS_OK;
}

}  // extern "C"

// ============================================================== raster core

namespace gs {

// Stage timing with HIP events recorded on the step's stream: mark(stage)
// before each stage, mark(-1) at the end; collect() attributes the time
// between consecutive marks to the earlier mark's stage.
struct Profiler {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<size_t, int>> marks;
    double acc[GS_STAGE_COUNT] = {};
    uint32_t steps = 0;

    ~Profiler() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    gs_status mark(hipStream_t st, int stage) {
        if (!enabled) return GS_OK;
        const size_t idx = marks.size();
        if (idx == pool.size()) {
            hipEvent_t e;
            GS_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        GS_HIP(hipEventRecord(pool[idx], st));
        marks.emplace_back(idx, stage);
        return GS_OK;
    }
    gs_status collect() {
        if (marks.empty()) return GS_OK;
        GS_HIP(hipEventSynchronize(pool[marks.back().first]));
        for (size_t i = 0; i + 1 < marks.size(); ++i) {
            const int s = marks[i].second;
            if (s < 0) continue;
            float ms = 0.0f;
            GS_HIP(hipEventElapsedTime(&ms, pool[marks[i].first], pool[marks[i + 1].first]));
            acc[s] += ms;
        }
        marks.clear();
        return GS_OK;
    }
};

// Device state of the projection -> binning -> raster pipeline for up to
// cap_views views of up to cap_n surfels. Used by the single-view Rasterizer
// handle and by the trainer's plain and batched steps.
struct RasterCore {
    gs_ctx ctx = nullptr;
    uint32_t cap_n = 0, cap_views = 0, cap_entries = 0, cap_tiles = 0, cap_pixels = 0;
    CamParams* cams_dev = nullptr;
    CamParams* cams_pinned = nullptr;
    hipEvent_t cams_done = nullptr;
    bool cams_pending = false;
    Record* recs = nullptr;
    uint4* rects = nullptr;
    uint32_t* depth_keys = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* total = nullptr;       // per view slot (plain) or [0] (batched)
    uint32_t* sort_count = nullptr;  // per view slot
    uint32_t* overflow = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* pays[2] = {nullptr, nullptr};
    uint8_t* scan_scratch = nullptr;
    uint8_t* sort_scratch = nullptr;
    uint32_t* ranges = nullptr;
    float4* out_color = nullptr;
    float4* out_aux = nullptr;
    float4* out_normal = nullptr;
    uint32_t* ncontrib = nullptr;
    float* grad_coef = nullptr;
    float* cam_partials = nullptr;
    // Decision trace of the forward (verification aid; lists == nullptr: off).
    FwdTrace trace;
    // Depth-presort path (binning = GS_BINNING_PRESORT or GS_BINNING_TIGHT).
    uint32_t* pre_keys[2] = {nullptr, nullptr};
    uint32_t* pre_vals[2] = {nullptr, nullptr};
    uint32_t* offs_sorted = nullptr;
    uint32_t* item_count = nullptr;
    uint4* emit_info = nullptr;      // presort binnings: per-item rect (16-bit) + tile mask
    // Tile queue: longest-list-first order of the last forward's tiles and
    // the two work cursors (forward, backward) of the persistent raster passes.
    uint32_t* order_keys[2] = {nullptr, nullptr};
    uint32_t* order_vals[2] = {nullptr, nullptr};
    uint32_t* order_count = nullptr;  // one per tile range
    uint32_t* cursors = nullptr;      // per tile range: forward, backward cursor sets
    bool tile_queue = true;
    uint32_t queue_blocks = 0;       // resident workgroups of a persistent raster launch
    // Tile ranges scheduled by the last forward: range s covers global tiles
    // [q_begin[s], q_begin[s] + q_tiles[s]) (q_tiles 0: no queue).
    static constexpr int kRanges = 2;
    uint32_t q_begin[kRanges] = {0, 0}, q_tiles[kRanges] = {0, 0};
    uint32_t* tkeys[2] = {nullptr, nullptr};  // 32-bit tile keys, stored inside keys[1]
    uint32_t binning = GS_BINNING_PRESORT;
    bool cull = true;                // per-block footprint culling in the rasterizers
    RasterShape shape;               // work-tile size and LDS batches (tile != 16 needs TIGHT binning)
    bool aux_outputs = true;         // forward writes depth (aux.y) and normals
    bool lean = false;               // skip rects / depth keys / counts the packed presort path never reads
    bool keys64_valid = false;       // keys[result_alt] holds 64-bit keys of the last forward
    uint32_t digit_bits = 8;
    int result_alt = 0;              // which ping-pong buffer holds the sorted entries
    // Current view set (host mirror of cams_dev).
    std::vector<CamParams> cams;
    uint32_t total_tiles = 0, total_pixels = 0;

    ~RasterCore() {
        if (cams_done) (void)hipEventDestroy(cams_done);
        if (cams_pinned) (void)hipHostFree(cams_pinned);
        dfree(cams_dev);
        dfree(recs);
        dfree(rects);
        dfree(depth_keys);
        dfree(offsets);
        dfree(total);
        dfree(sort_count);
        dfree(overflow);
        for (int i = 0; i < 2; ++i) {
            dfree(keys[i]);
            dfree(pays[i]);
        }
        dfree(scan_scratch);
        dfree(sort_scratch);
        dfree(ranges);
        dfree(out_color);
        dfree(out_aux);
        dfree(out_normal);
        dfree(ncontrib);
        dfree(grad_coef);
        dfree(cam_partials);
        for (int i = 0; i < 2; ++i) {
            dfree(pre_keys[i]);
            dfree(pre_vals[i]);
        }
        dfree(offs_sorted);
        dfree(item_count);
        dfree(emit_info);
        for (int i = 0; i < 2; ++i) {
            dfree(order_keys[i]);
            dfree(order_vals[i]);
        }
        dfree(order_count);
        dfree(cursors);
    }

    gs_status init(gs_ctx c, uint32_t n, uint32_t views, uint32_t entries, uint32_t tiles,
                   uint32_t pixels) {
        ctx = c;
        if (entries >= (1u << 30)) return fail(GS_ERR_CAPACITY, "max_entries must be < 2^30");
        if ((uint64_t)n * views >= (1ull << 32)) return fail(GS_ERR_CAPACITY, "views * surfels must be < 2^32");
        cap_n = n;
        cap_views = views;
        cap_entries = entries;
        cap_tiles = tiles;
        cap_pixels = pixels;
        const size_t nv = (size_t)n * views;
        GS_CHECK(dalloc(&cams_dev, views, "cams"));
        GS_HIP(hipHostMalloc((void**)&cams_pinned, sizeof(CamParams) * views, hipHostMallocDefault));
        GS_HIP(hipEventCreateWithFlags(&cams_done, hipEventDisableTiming));
        GS_CHECK(dalloc(&recs, nv, "records"));
        GS_CHECK(dalloc(&rects, nv, "rects"));
        GS_CHECK(dalloc(&depth_keys, nv, "depth keys"));
        GS_CHECK(dalloc(&offsets, nv, "offsets"));
        GS_CHECK(dalloc(&total, views, "totals"));
        GS_CHECK(dalloc(&sort_count, views, "sort counts"));
        GS_CHECK(dalloc(&overflow, 1, "overflow flag"));
        for (int i = 0; i < 2; ++i) {
            GS_CHECK(dalloc(&keys[i], entries, "keys"));
            GS_CHECK(dalloc(&pays[i], entries, "payloads"));
        }
        // The presorted path's two 32-bit tile-key buffers live in keys[1]
        // (8 B per entry = two u32 arrays); keys[0] stays free for the
        // composed 64-bit keys the debug interface reports.
        tkeys[0] = reinterpret_cast<uint32_t*>(keys[1]);
        tkeys[1] = tkeys[0] + (entries ? entries : 1);
        GS_CHECK(dalloc(&scan_scratch, scan_scratch_bytes((uint32_t)std::min<size_t>(nv, 0xFFFFFFFFu)),
                        "scan scratch"));
        GS_CHECK(dalloc(&sort_scratch,
                        sort_scratch_bytes((uint32_t)std::max<size_t>(
                            std::max<size_t>(entries, tiles), std::min<size_t>(nv, 0xFFFFFFFFu))),
                        "sort scratch"));
        for (int i = 0; i < 2; ++i) {
            GS_CHECK(dalloc(&order_keys[i], tiles ? tiles : 1, "tile order keys"));
            GS_CHECK(dalloc(&order_vals[i], tiles ? tiles : 1, "tile order"));
        }
        GS_CHECK(dalloc(&order_count, kRanges, "tile order count"));
        GS_CHECK(dalloc(&cursors, 2 * kRanges * kTileQueues, "tile cursors"));
        {
            // Persistent raster launches: 8 workgroups per CU (more than can
            // be resident; the rest start as slots free up). GS_TILE_QUEUE=0
            // selects one workgroup per tile in tile order instead.
            const char* env = std::getenv("GS_TILE_QUEUE");
            tile_queue = !(env != nullptr && env[0] == '0');
            queue_blocks = 8u * (uint32_t)std::max(1, ctx->props.multiProcessorCount);
        }
        for (int i = 0; i < 2; ++i) {
            GS_CHECK(dalloc(&pre_keys[i], nv, "presort keys"));
            GS_CHECK(dalloc(&pre_vals[i], nv, "presort order"));
        }
        GS_CHECK(dalloc(&offs_sorted, nv, "presorted offsets"));
        GS_CHECK(dalloc(&emit_info, nv, "emit info"));
        GS_CHECK(dalloc(&item_count, 1, "item count"));
        GS_CHECK(dalloc(&ranges, (size_t)tiles * 2, "ranges"));
        GS_CHECK(dalloc(&out_color, pixels, "out_color"));
        GS_CHECK(dalloc(&out_aux, pixels, "out_aux"));
        GS_CHECK(dalloc(&out_normal, pixels, "out_normal"));
        GS_CHECK(dalloc(&ncontrib, pixels, "ncontrib"));
        GS_CHECK(dalloc(&grad_coef, nv * kGradStride, "grad_coef"));
        uint8_t* cp = nullptr;
        GS_CHECK(dalloc(&cp, geom_cam_partials_bytes(n, views), "camera partials"));
        cam_partials = reinterpret_cast<float*>(cp);
        return GS_OK;
    }

    // Lay out a view set: batched = cumulative tile / pixel / record bases,
    // plain = every view at base 0 (processed one at a time). Uploads the
    // table from pinned memory (the previous upload must have landed before
    // the staging buffer is rewritten).
    gs_status set_views(hipStream_t st, const std::vector<CamParams>& in, bool batched) {
        if (in.size() > cap_views) return fail(GS_ERR_CAPACITY, "too many views for this capacity");
        cams = in;
        uint32_t tb = 0, pb = 0;
        total_tiles = 0;
        total_pixels = 0;
        for (size_t v = 0; v < cams.size(); ++v) {
            CamParams& c = cams[v];
            const uint32_t t = c.tiles_x * c.tiles_y, p = c.width * c.height;
            if (batched) {
                c.tile_base = tb;
                c.pix_base = pb;
                c.rec_base = (uint32_t)v * cap_n;
                tb += t;
                pb += p;
                total_tiles = tb;
                total_pixels = pb;
            } else {
                c.tile_base = 0;
                c.pix_base = 0;
                c.rec_base = 0;
                total_tiles = std::max(total_tiles, t);
                total_pixels = std::max(total_pixels, p);
            }
        }
        if (total_tiles > cap_tiles || total_pixels > cap_pixels)
            return fail(GS_ERR_CAPACITY, "view set exceeds the tile/pixel capacity");
        if (cams_pending) GS_HIP(hipEventSynchronize(cams_done));
        std::memcpy(cams_pinned, cams.data(), sizeof(CamParams) * cams.size());
        GS_HIP(hipMemcpyAsync(cams_dev, cams_pinned, sizeof(CamParams) * cams.size(),
                              hipMemcpyHostToDevice, st));
        GS_HIP(hipEventRecord(cams_done, st));
        cams_pending = true;
        return GS_OK;
    }

    const uint32_t* last_sort_count = nullptr;
    bool keys_lean = false;          // the last forward left depth_keys unwritten

    // 64-bit (tile << 32 | depth) keys of the last forward (composed on
    // demand for the presorted path) and its sorted payloads.
    gs_status sorted_keys(hipStream_t st, const uint64_t** out) {
        if (keys_lean) return fail(GS_ERR_INVALID, "the last forward did not keep its depth keys");
        if (!keys64_valid && last_sort_count) {
            GS_HIP(launch_compose_keys(st, tkeys[result_alt], pays[result_alt], depth_keys, last_sort_count,
                                       cap_entries, keys[0]));
            *out = keys[0];
            return GS_OK;
        }
        *out = keys[result_alt];
        return GS_OK;
    }
    const uint32_t* sorted_pays() const { return pays[result_alt]; }

    // Forward over views [v0, v0 + nv) of the current set (nv > 1 only when
    // the set was laid out batched). Parameters are the activated buffers.
    // With raster = false it stops after the tile ranges (the raster passes
    // then run per tile range: raster_fwd_range / raster_bwd_range).
    gs_status forward(hipStream_t st, uint32_t v0, uint32_t nv, uint32_t n, const float4* pos,
                      const float2* sc, const float4* q, const float* op, const float* sh,
                      Profiler* prof = nullptr, bool raster = true) {
        if (n > cap_n) return fail(GS_ERR_CAPACITY, "num_surfels exceeds capacity");
        if (shape.tile != (uint32_t)kTile && binning != GS_BINNING_TIGHT)
            return fail(GS_ERR_INVALID, "work tiles other than 16 px need GS_BINNING_TIGHT");
        Profiler dummy;
        Profiler& P = prof ? *prof : dummy;
        const CamParams* cd = cams_dev + v0;
        uint32_t tiles = 0;
        for (uint32_t v = v0; v < v0 + nv; ++v) tiles += cams[v].tiles_x * cams[v].tiles_y;
        uint32_t max_tile_id = 0;
        for (uint32_t v = v0; v < v0 + nv; ++v)
            max_tile_id = std::max(max_tile_id, cams[v].tile_base + cams[v].tiles_x * cams[v].tiles_y);
        const uint32_t items = n * nv;
        FwdOut o{out_color, out_aux, out_normal, ncontrib};
        if (binning != GS_BINNING_SORT64) {
            // Stable depth presort of the items, emission in that order, then
            // a stable sort on the tile bits only (binning.hip).
            // Up to 2^24 items the presort payload also carries the tile
            // count (low 8 bits), so the scan needs no gather.
            // 8-px work tiles can give a rect more than 255 tiles: no packing.
            const bool packed = items < (1u << (32 - kPackBits)) && shape.tile >= 16;
            GS_CHECK(P.mark(st, GS_STAGE_PROJECT));
            // Lean (trainer): with packed counts nothing downstream reads the
            // rects, depth keys or counts, and colour-only passes never read
            // the records' normal / depth.
            const bool skip = lean && packed;
            GS_HIP(launch_project(st, cd, nv, n, pos, sc, q, op, sh, recs, skip ? nullptr : rects,
                                  skip ? nullptr : depth_keys, skip ? nullptr : offsets, ranges, 2 * max_tile_id,
                                  pre_keys[0], pre_vals[0], item_count, binning == GS_BINNING_TIGHT, emit_info,
                                  packed, aux_outputs));
            keys_lean = skip;
            GS_CHECK(P.mark(st, GS_STAGE_SORT));
            int alt = 0;
            GS_HIP(launch_radix_sort_u32(st, pre_keys[0], pre_vals[0], pre_keys[1], pre_vals[1],
                                         item_count, items, 0, 32, digit_bits, sort_scratch, &alt));
            const uint32_t* order = pre_vals[alt];
            GS_CHECK(P.mark(st, GS_STAGE_SCAN));
            GS_HIP(launch_exclusive_scan(st, packed ? nullptr : offsets, offs_sorted, items, total + v0,
                                         scan_scratch, order));
            GS_CHECK(P.mark(st, GS_STAGE_EMIT));
            GS_HIP(launch_emit_tile_keys(st, items, order, packed, emit_info, offs_sorted, cd, n ? n : 1,
                                         tkeys[0], pays[0], cap_entries, total + v0, sort_count + v0,
                                         overflow));
            GS_CHECK(P.mark(st, GS_STAGE_SORT));
            GS_HIP(launch_radix_sort_u32(st, tkeys[0], pays[0], tkeys[1], pays[1], sort_count + v0,
                                         cap_entries, 0, bits_for(max_tile_id), digit_bits, sort_scratch,
                                         &result_alt));
            GS_CHECK(P.mark(st, GS_STAGE_RANGES));
            GS_HIP(launch_tile_ranges_u32(st, tkeys[result_alt], sort_count + v0, cap_entries, ranges));
            keys64_valid = false;
            last_sort_count = sort_count + v0;
            if (!raster) return GS_OK;
            GS_CHECK(schedule_range(st, 0, 0, tiles));
            GS_CHECK(P.mark(st, GS_STAGE_RASTER_FWD));
            GS_HIP(launch_raster_fwd(st, cd, cams.data() + v0, nv, tiles, ranges, pays[result_alt], recs, o,
                                     cull, queue(0, 0), aux_outputs, shape, 0, trace));
            return GS_OK;
        }
        // 64-bit (tile || depth) key path: one sort over every key bit that varies.
        GS_CHECK(P.mark(st, GS_STAGE_PROJECT));
        GS_HIP(launch_project(st, cd, nv, n, pos, sc, q, op, sh, recs, rects, depth_keys, offsets,
                              ranges, 2 * max_tile_id));
        GS_CHECK(P.mark(st, GS_STAGE_SCAN));
        GS_HIP(launch_exclusive_scan(st, offsets, offsets, items, total + v0, scan_scratch));
        GS_CHECK(P.mark(st, GS_STAGE_EMIT));
        GS_HIP(launch_emit_keys(st, items, rects, depth_keys, offsets, cd, n ? n : 1, 0, keys[0],
                                pays[0], cap_entries, total + v0, sort_count + v0, overflow));
        const uint32_t end_bit = 32 + bits_for(max_tile_id);
        GS_CHECK(P.mark(st, GS_STAGE_SORT));
        GS_HIP(launch_radix_sort_u64(st, keys[0], pays[0], keys[1], pays[1], sort_count + v0,
                                     cap_entries, 0, end_bit, digit_bits, sort_scratch, &result_alt));
        GS_CHECK(P.mark(st, GS_STAGE_RANGES));
        GS_HIP(launch_tile_ranges(st, keys[result_alt], sort_count + v0, cap_entries, ranges));
        keys64_valid = true;
        keys_lean = false;
        last_sort_count = sort_count + v0;
        if (!raster) return GS_OK;
        GS_CHECK(schedule_range(st, 0, 0, tiles));
        GS_CHECK(P.mark(st, GS_STAGE_RASTER_FWD));
        GS_HIP(launch_raster_fwd(st, cd, cams.data() + v0, nv, tiles, ranges, pays[result_alt], recs, o,
                                 cull, queue(0, 0), aux_outputs, shape, 0, trace));
        return GS_OK;
    }

    // Longest-list-first order of the binned tiles [t0, t0 + nt) (16-bit
    // keys, two stable 8-bit passes) as tile range `slot`, kept in the
    // segment [t0, t0 + nt) of the order buffers, and its two work cursor
    // sets reset to 0.
    gs_status schedule_range(hipStream_t st, int slot, uint32_t t0, uint32_t nt) {
        q_begin[slot] = t0;
        q_tiles[slot] = 0;
        if (!tile_queue || nt == 0) return GS_OK;
        GS_HIP(launch_tile_order_keys(st, ranges, t0, nt, order_keys[0] + t0, order_vals[0] + t0,
                                      order_count + slot, cursors + 2 * slot * kTileQueues));
        int alt = 0;
        GS_HIP(launch_radix_sort_u32(st, order_keys[0] + t0, order_vals[0] + t0, order_keys[1] + t0,
                                     order_vals[1] + t0, order_count + slot, nt, 0, 16, 8, sort_scratch, &alt));
        if (alt != 0) return fail(GS_ERR_INVALID, "tile order sort ended in the alternate buffer");
        q_tiles[slot] = nt;
        return GS_OK;
    }

    // Queue of a raster pass over tile range `slot` (which: 0 = forward,
    // 1 = backward cursor set), or none.
    TileQueue queue(int slot, int which) const {
        if (q_tiles[slot] == 0) return TileQueue{nullptr, nullptr, 0};
        return TileQueue{order_vals[0] + q_begin[slot], cursors + (2 * slot + which) * kTileQueues, queue_blocks};
    }

    // Raster passes over the binned tiles [t0, t0 + nt) of views
    // [v0, v0 + nv) (tile ids are global to the view set), as tile range
    // `slot`; the forward schedules the range's queue first.
    gs_status raster_fwd_range(hipStream_t st, int slot, uint32_t v0, uint32_t nv, uint32_t t0, uint32_t nt) {
        FwdOut o{out_color, out_aux, out_normal, ncontrib};
        GS_CHECK(schedule_range(st, slot, t0, nt));
        GS_HIP(launch_raster_fwd(st, cams_dev + v0, cams.data() + v0, nv, nt, ranges, pays[result_alt], recs, o,
                                 cull, queue(slot, 0), aux_outputs, shape, t0));
        return GS_OK;
    }
    gs_status raster_bwd_range(hipStream_t st, int slot, uint32_t v0, uint32_t nv, uint32_t t0, uint32_t nt,
                               const float4* dl_dcolor) {
        BwdIn in{dl_dcolor, nullptr, out_color, out_aux, ncontrib};
        const TileQueue qb = (q_tiles[slot] == nt && q_begin[slot] == t0) ? queue(slot, 1)
                                                                            : TileQueue{nullptr, nullptr, 0};
        if (qb.cursor != nullptr) GS_HIP(launch_fill_u32(st, qb.cursor, 0u, kTileQueues));
        GS_HIP(launch_raster_bwd(st, cams_dev + v0, cams.data() + v0, nv, nt, ranges, pays[result_alt], recs, in,
                                 grad_coef, false, cull, qb, shape, t0));
        return GS_OK;
    }

    gs_status backward(hipStream_t st, uint32_t v0, uint32_t nv, uint32_t n, const float4* pos,
                       const float2* sc, const float4* q, const float* sh, const float4* dl_dcolor,
                       const float4* dl_dnormal, ParamGrads g, bool accumulate, bool aux,
                       Profiler* prof = nullptr, bool geom = true) {
        Profiler dummy;
        Profiler& P = prof ? *prof : dummy;
        const CamParams* cd = cams_dev + v0;
        uint32_t tiles = 0;
        for (uint32_t v = v0; v < v0 + nv; ++v) tiles += cams[v].tiles_x * cams[v].tiles_y;
        GS_CHECK(P.mark(st, GS_STAGE_GRAD_CLEAR));
        GS_HIP(launch_fill_u32(st, reinterpret_cast<uint32_t*>(grad_coef), 0u,
                               (size_t)n * nv * grad_stride(aux)));
        BwdIn in{dl_dcolor, dl_dnormal, out_color, out_aux, ncontrib};
        GS_CHECK(P.mark(st, GS_STAGE_RASTER_BWD));
        const TileQueue qb = (q_tiles[0] == tiles && q_begin[0] == 0) ? queue(0, 1) : TileQueue{nullptr, nullptr, 0};
        if (qb.cursor != nullptr) GS_HIP(launch_fill_u32(st, qb.cursor, 0u, kTileQueues));
        GS_HIP(launch_raster_bwd(st, cd, cams.data() + v0, nv, tiles, ranges, pays[result_alt], recs, in,
                                 grad_coef, aux, cull, qb, shape));
        if (!geom) return GS_OK;
        GS_CHECK(P.mark(st, GS_STAGE_GEOM_BWD));
        GS_HIP(launch_geom_bwd(st, cd, nv, n, pos, sc, q, sh, grad_coef, aux, g, accumulate,
                               cam_partials));
        return GS_OK;
    }
};

}  // namespace gs

// =============================================================== Rasterizer

struct gs_raster_s {
    gs_ctx ctx = nullptr;
    uint32_t capacity = 0, width = 0, height = 0, max_entries = 0;
    float lambda_dist = 0.0f;
    gs::RasterCore core;
    float4* positions = nullptr;
    float2* scales = nullptr;
    float4* quats = nullptr;
    float* opacities = nullptr;
    float* sh = nullptr;
    float4* dl_dcolor = nullptr;
    float4* dl_dnormal = nullptr;
    float4* grad_pos = nullptr;
    float2* grad_scales = nullptr;
    float4* grad_quat = nullptr;
    float* grad_opacity = nullptr;
    float* grad_sh = nullptr;
    float* grad_cam = nullptr;
    uint32_t last_n = 0;
    bool have_forward = false;
    uint32_t max_views = 1;
    uint32_t last_views = 0;   // views of the last forward (0: none)
    bool last_batch = false;   // the last forward was gs_raster_forward_batch
    uint32_t* trace_lists = nullptr;
    uint32_t* trace_counts = nullptr;
    ~gs_raster_s() {
        dfree(trace_lists);
        dfree(trace_counts);
        dfree(positions);
        dfree(scales);
        dfree(quats);
        dfree(opacities);
        dfree(sh);
        dfree(dl_dcolor);
        dfree(dl_dnormal);
        dfree(grad_pos);
        dfree(grad_scales);
        dfree(grad_quat);
        dfree(grad_opacity);
        dfree(grad_sh);
        dfree(grad_cam);
    }
};

static gs_status upload_scene_into(gs_ctx ctx, const gs_scene_input* s, uint32_t capacity, float4* pos,
                                   float2* sc, float4* q, float* op, float* sh) {
    GS_REQUIRE(s != nullptr, "scene is null");
    if (s->n > capacity) return fail(GS_ERR_CAPACITY, "scene larger than capacity");  // rasterizer.rs:361
    GS_REQUIRE(s->sh_coeffs >= 1 && s->sh_coeffs <= 16, "sh_coeffs must be in 1..16");
    GS_REQUIRE(s->n == 0 || (s->positions && s->scales && s->quats && s->opacities && s->sh),
               "scene arrays must be non-null");
    const size_t n = s->n;
    std::vector<float> p4(n * 4), sh48(n * kShFloats, 0.0f);
    for (size_t i = 0; i < n; ++i) {
        p4[4 * i] = s->positions[3 * i];
        p4[4 * i + 1] = s->positions[3 * i + 1];
        p4[4 * i + 2] = s->positions[3 * i + 2];
        p4[4 * i + 3] = 1.0f;
        std::memcpy(&sh48[i * kShFloats], s->sh + i * s->sh_coeffs * 3, sizeof(float) * s->sh_coeffs * 3);
    }
    hipStream_t st = ctx->stream;
    GS_HIP(hipMemcpyAsync(pos, p4.data(), n * 16, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(sc, s->scales, n * 8, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(q, s->quats, n * 16, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(op, s->opacities, n * 4, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(sh, sh48.data(), n * kShFloats * 4, hipMemcpyHostToDevice, st));
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

extern "C" {

gs_status gs_raster_create_batch(gs_ctx ctx, uint32_t capacity, uint32_t width, uint32_t height,
                                 uint32_t max_entries, uint32_t max_views, gs_raster* out) {
    GS_REQUIRE(ctx && out, "null argument");
    GS_REQUIRE(width > 0 && height > 0, "image size must be > 0");
    GS_REQUIRE(max_views >= 1 && max_views <= 64, "max_views must be in 1..64");
    GS_REQUIRE((uint64_t)width * height * max_views < (1ull << 31), "image too large");
    GS_REQUIRE((uint64_t)max_entries * max_views < (1ull << 30), "max_entries * max_views must be < 2^30");
    *out = nullptr;
    auto r = std::make_unique<gs_raster_s>();
    r->ctx = ctx;
    r->capacity = capacity;
    r->width = width;
    r->height = height;
    r->max_entries = max_entries;
    r->max_views = max_views;
    const uint32_t tiles = div_up(width, kTile) * div_up(height, kTile);
    const uint32_t px = width * height * max_views;
    GS_CHECK(r->core.init(ctx, capacity, max_views, max_entries * max_views, tiles * max_views, px));
    GS_CHECK(dalloc(&r->positions, capacity, "positions"));
    GS_CHECK(dalloc(&r->scales, capacity, "scales"));
    GS_CHECK(dalloc(&r->quats, capacity, "quats"));
    GS_CHECK(dalloc(&r->opacities, capacity, "opacities"));
    GS_CHECK(dalloc(&r->sh, (size_t)capacity * kShFloats, "sh"));
    GS_CHECK(dalloc(&r->dl_dcolor, px, "dl_dcolor"));
    GS_CHECK(dalloc(&r->dl_dnormal, px, "dl_dnormal"));
    GS_CHECK(dalloc(&r->grad_pos, capacity, "grad_pos"));
    GS_CHECK(dalloc(&r->grad_scales, capacity, "grad_scales"));
    GS_CHECK(dalloc(&r->grad_quat, capacity, "grad_quat"));
    GS_CHECK(dalloc(&r->grad_opacity, capacity, "grad_opacity"));
    GS_CHECK(dalloc(&r->grad_sh, (size_t)capacity * kShFloats, "grad_sh"));
    GS_CHECK(dalloc(&r->grad_cam, (size_t)16 * max_views, "grad_cam"));
    *out = r.release();
    return GS_OK;
}

gs_status gs_raster_create(gs_ctx ctx, uint32_t capacity, uint32_t width, uint32_t height,
                           uint32_t max_entries, gs_raster* out) {
    return gs_raster_create_batch(ctx, capacity, width, height, max_entries, 1, out);
}

void gs_raster_destroy(gs_raster r) {
    if (!r) return;
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

gs_status gs_raster_get_buffers(gs_raster r, gs_raster_buffers* o) {
    GS_REQUIRE(r && o, "null argument");
    o->positions = reinterpret_cast<float*>(r->positions);
    o->scales = reinterpret_cast<float*>(r->scales);
    o->quats = reinterpret_cast<float*>(r->quats);
    o->opacities = r->opacities;
    o->sh = r->sh;
    o->out_color = reinterpret_cast<float*>(r->core.out_color);
    o->out_aux = reinterpret_cast<float*>(r->core.out_aux);
    o->out_normal = reinterpret_cast<float*>(r->core.out_normal);
    o->out_ncontrib = r->core.ncontrib;
    o->dl_dcolor = reinterpret_cast<float*>(r->dl_dcolor);
    o->dl_dnormal = reinterpret_cast<float*>(r->dl_dnormal);
    o->grad_pos = reinterpret_cast<float*>(r->grad_pos);
    o->grad_scales = reinterpret_cast<float*>(r->grad_scales);
    o->grad_quat = reinterpret_cast<float*>(r->grad_quat);
    o->grad_opacity = r->grad_opacity;
    o->grad_sh = r->grad_sh;
    o->grad_cam = r->grad_cam;
    return GS_OK;
}

gs_status gs_raster_upload_scene(gs_raster r, const gs_scene_input* scene) {
    GS_REQUIRE(r != nullptr, "raster is null");
    return upload_scene_into(r->ctx, scene, r->capacity, r->positions, r->scales, r->quats,
                             r->opacities, r->sh);
}

gs_status gs_raster_set_lambda_dist(gs_raster r, float lambda_dist) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_REQUIRE(std::isfinite(lambda_dist), "lambda_dist must be finite");
    r->lambda_dist = lambda_dist;
    return GS_OK;
}

gs_status gs_raster_set_digit_bits(gs_raster r, uint32_t digit_bits) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_REQUIRE(digit_bits == 4 || digit_bits == 8, "digit_bits must be 4 or 8");
    r->core.digit_bits = digit_bits;
    return GS_OK;
}

gs_status gs_raster_set_culling(gs_raster r, int enable) {
    GS_REQUIRE(r != nullptr, "raster is null");
    r->core.cull = enable != 0;
    return GS_OK;
}

gs_status gs_raster_set_trace(gs_raster r, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t cap) {
    GS_REQUIRE(r != nullptr, "raster is null");
    dfree(r->trace_lists);
    dfree(r->trace_counts);
    r->core.trace = FwdTrace{};
    if ((uint64_t)w * h * cap == 0) return GS_OK;
    GS_REQUIRE((uint64_t)x0 + w <= r->width && (uint64_t)y0 + h <= r->height,
               "trace window must lie inside the image");
    GS_REQUIRE((uint64_t)w * h * cap <= (1ull << 30), "trace window x cap must be <= 2^30 entries");
    GS_CHECK(dalloc(&r->trace_lists, (size_t)w * h * cap, "trace lists"));
    GS_CHECK(dalloc(&r->trace_counts, (size_t)w * h, "trace counts"));
    FwdTrace t;
    t.lists = r->trace_lists;
    t.counts = r->trace_counts;
    t.x0 = x0;
    t.y0 = y0;
    t.w = w;
    t.h = h;
    t.cap = cap;
    t.view = 0;
    r->core.trace = t;
    return GS_OK;
}

gs_status gs_raster_trace_buffers(gs_raster r, const uint32_t** lists, const uint32_t** counts) {
    GS_REQUIRE(r && lists && counts, "null argument");
    GS_HIP(hipStreamSynchronize(r->ctx->stream));
    *lists = r->trace_lists;
    *counts = r->trace_counts;
    return GS_OK;
}

gs_status gs_raster_set_binning(gs_raster r, uint32_t mode) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_REQUIRE(mode <= GS_BINNING_TIGHT, "unknown binning mode");
    r->core.binning = mode;
    return GS_OK;
}

gs_status gs_raster_forward(gs_raster r, void* stream, const gs_raster_camera* camera,
                            uint32_t num_surfels) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_CHECK(check_camera(camera));
    if (num_surfels > r->capacity) return fail(GS_ERR_CAPACITY, "num_surfels exceeds capacity");  // rasterizer.rs:409
    hipStream_t st = as_stream(stream, r->ctx->stream);
    CamParams c = make_cam(*camera, r->width, r->height);
    c.lambda_dist = r->lambda_dist;
    GS_CHECK(r->core.set_views(st, {c}, true));
    GS_CHECK(r->core.forward(st, 0, 1, num_surfels, r->positions, r->scales, r->quats, r->opacities, r->sh));
    r->last_n = num_surfels;
    r->last_views = 1;
    r->last_batch = false;
    r->have_forward = true;
    return GS_OK;
}

gs_status gs_raster_forward_batch(gs_raster r, void* stream, const gs_view* views, uint32_t n_views) {
    GS_REQUIRE(r && views, "null argument");
    GS_REQUIRE(n_views >= 1 && n_views <= r->max_views, "view count out of range for this rasterizer");
    std::vector<CamParams> cams;
    uint32_t n = 0;
    for (uint32_t j = 0; j < n_views; ++j) {
        const gs_view& v = views[j];
        GS_CHECK(check_camera(&v.camera));
        GS_REQUIRE(v.width > 0 && v.height > 0 && v.width <= r->width && v.height <= r->height,
                   "view size must be within the rasterizer's max width x height");
        if (v.num_surfels > r->capacity) return fail(GS_ERR_CAPACITY, "num_surfels exceeds capacity");
        CamParams c = make_cam(v.camera, v.width, v.height);
        c.lambda_dist = r->lambda_dist;
        c.n_surfels = v.num_surfels;
        cams.push_back(c);
        n = std::max(n, v.num_surfels);
    }
    hipStream_t st = as_stream(stream, r->ctx->stream);
    GS_CHECK(r->core.set_views(st, cams, true));
    GS_CHECK(r->core.forward(st, 0, n_views, n, r->positions, r->scales, r->quats, r->opacities, r->sh));
    r->last_n = n;
    r->last_views = n_views;
    r->last_batch = true;
    r->have_forward = true;
    return GS_OK;
}

gs_status gs_raster_backward_batch(gs_raster r, void* stream, uint32_t flags) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_REQUIRE(r->have_forward && r->last_batch, "backward_batch needs a forward_batch on this rasterizer first");
    hipStream_t st = as_stream(stream, r->ctx->stream);
    const bool accumulate = (flags & GS_BWD_ACCUMULATE) != 0;
    const bool aux = (flags & GS_BWD_COLOR_ONLY) == 0;
    ParamGrads g{r->grad_pos, r->grad_scales, r->grad_quat, r->grad_opacity, r->grad_sh,
                 (flags & GS_BWD_NO_CAMERA) ? nullptr : r->grad_cam};
    GS_CHECK(r->core.backward(st, 0, r->last_views, r->last_n, r->positions, r->scales, r->quats, r->sh,
                              r->dl_dcolor, r->dl_dnormal, g, accumulate, aux));
    return GS_OK;
}

gs_status gs_raster_backward_ex(gs_raster r, void* stream, uint32_t num_surfels, uint32_t flags) {
    GS_REQUIRE(r != nullptr, "raster is null");
    GS_REQUIRE(r->have_forward && !r->last_batch, "backward needs a forward on this rasterizer first");
    GS_REQUIRE(num_surfels == r->last_n, "backward num_surfels must match the forward");
    hipStream_t st = as_stream(stream, r->ctx->stream);
    const bool accumulate = (flags & GS_BWD_ACCUMULATE) != 0;
    const bool aux = (flags & GS_BWD_COLOR_ONLY) == 0;
    ParamGrads g{r->grad_pos, r->grad_scales, r->grad_quat, r->grad_opacity, r->grad_sh,
                 (flags & GS_BWD_NO_CAMERA) ? nullptr : r->grad_cam};
    GS_CHECK(r->core.backward(st, 0, 1, num_surfels, r->positions, r->scales, r->quats, r->sh,
                              r->dl_dcolor, r->dl_dnormal, g, accumulate, aux));
    return GS_OK;
}

gs_status gs_raster_backward(gs_raster r, void* stream, uint32_t num_surfels) {
    return gs_raster_backward_ex(r, stream, num_surfels, 0u);
}

gs_status gs_raster_entry_count(gs_raster r, uint32_t* entries) {
    GS_REQUIRE(r && entries, "null argument");
    GS_HIP(hipStreamSynchronize(r->ctx->stream));
    GS_HIP(hipDeviceSynchronize());
    uint32_t t = 0, of = 0;
    GS_HIP(hipMemcpy(&t, r->core.total, 4, hipMemcpyDeviceToHost));
    GS_HIP(hipMemcpy(&of, r->core.overflow, 4, hipMemcpyDeviceToHost));
    *entries = t;
    if (of) {
        const uint32_t zero = 0;
        GS_HIP(hipMemcpy(r->core.overflow, &zero, 4, hipMemcpyHostToDevice));
        return fail(GS_ERR_OVERFLOW, "tile entries exceeded max_entries (" + std::to_string(t) + " > " +
                                         std::to_string(r->max_entries) + ")");
    }
    return GS_OK;
}

gs_status gs_raster_debug_buffers(gs_raster r, gs_raster_debug* o) {
    GS_REQUIRE(r && o, "null argument");
    GS_REQUIRE(r->have_forward, "no forward has run on this rasterizer");
    GS_CHECK(r->core.sorted_keys(r->ctx->stream, &o->sorted_keys));
    GS_HIP(hipStreamSynchronize(r->ctx->stream));
    o->sorted_payloads = r->core.sorted_pays();
    o->ranges = r->core.ranges;
    o->rects = reinterpret_cast<const uint32_t*>(r->core.rects);
    o->depth_keys = r->core.depth_keys;
    o->offsets = r->core.offsets;
    o->num_tiles = div_up(r->width, kTile) * div_up(r->height, kTile);
    return GS_OK;
}

}  // extern "C"

// ================================================================= SsimLoss

struct gs_loss_s {
    gs_ctx ctx = nullptr;
    uint32_t width = 0, height = 0;
    float lambda = 0.2f, dssim_scale = 0.5f;
    const float4* img = nullptr;
    float4* dl_out = nullptr;
    float4* target = nullptr;
    double* partials = nullptr;
    uint32_t n_partials = 0;
    ~gs_loss_s() {
        dfree(target);
        dfree(partials);
    }
};

extern "C" {

gs_status gs_loss_create(gs_ctx ctx, uint32_t width, uint32_t height, float lambda, const float* img,
                         float* dl_out, gs_loss* out) {
    GS_REQUIRE(ctx && out && img && dl_out, "null argument");
    GS_REQUIRE(width > 0 && height > 0, "image size must be > 0");
    GS_REQUIRE(lambda >= 0.0f && lambda <= 1.0f, "lambda must be in [0, 1]");
    *out = nullptr;
    auto l = std::make_unique<gs_loss_s>();
    l->ctx = ctx;
    l->width = width;
    l->height = height;
    l->lambda = lambda;
    l->img = reinterpret_cast<const float4*>(img);
    l->dl_out = reinterpret_cast<float4*>(dl_out);
    const size_t P = (size_t)width * height;
    GS_CHECK(dalloc(&l->target, P, "loss target"));
    l->n_partials = l1_ssim_partials(width, height);
    GS_CHECK(dalloc(&l->partials, (size_t)l->n_partials * 2, "loss partials"));
    *out = l.release();
    return GS_OK;
}

void gs_loss_destroy(gs_loss l) { delete l; }
float* gs_loss_target(gs_loss l) { return l ? reinterpret_cast<float*>(l->target) : nullptr; }

gs_status gs_loss_set_dssim_scale(gs_loss l, float scale) {
    GS_REQUIRE(l != nullptr, "loss is null");
    GS_REQUIRE(std::isfinite(scale) && scale >= 0.0f, "dssim_scale must be finite and >= 0");
    l->dssim_scale = scale;
    return GS_OK;
}

gs_status gs_loss_encode(gs_loss l, void* stream) {
    GS_REQUIRE(l != nullptr, "loss is null");
    GS_HIP(launch_l1_ssim(as_stream(stream, l->ctx->stream), l->width, l->height, l->lambda,
                          l->dssim_scale, l->img, l->target, l->dl_out, l->partials));
    return GS_OK;
}

static gs_status read_loss_partials(hipStream_t st, const double* partials, uint32_t n, uint32_t w,
                                    uint32_t h, float dssim_scale, double* l1_mean, double* dssim_mean) {
    std::vector<double> p((size_t)n * 2);
    GS_HIP(hipStreamSynchronize(st));
    GS_HIP(hipMemcpy(p.data(), partials, p.size() * sizeof(double), hipMemcpyDeviceToHost));
    double l1 = 0.0, ss = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        l1 += p[2 * i];
        ss += p[2 * i + 1];
    }
    const double n3 = (double)w * h * 3.0;
    if (l1_mean) *l1_mean = l1 / n3;
    if (dssim_mean) *dssim_mean = (double)dssim_scale * (1.0 - ss / n3);
    return GS_OK;
}

gs_status gs_loss_read(gs_loss l, double* l1_mean, double* dssim_mean) {
    GS_REQUIRE(l != nullptr, "loss is null");
    return read_loss_partials(l->ctx->stream, l->partials, l->n_partials, l->width, l->height,
                              l->dssim_scale, l1_mean, dssim_mean);
}

}  // extern "C"

// ================================================================= Tracker

struct gs_tracker_s {
    gs_ctx ctx = nullptr;
    gs_raster raster = nullptr;
    uint32_t width = 0, height = 0;
    float4* target = nullptr;
    float2* partials = nullptr;
    uint32_t n_partials = 0;
    TrackState* state = nullptr;
    TrackState* state_host = nullptr;  // pinned
    float* residuals = nullptr;
    uint32_t residual_cap = 0;
    ~gs_tracker_s() {
        dfree(target);
        dfree(partials);
        dfree(state);
        dfree(residuals);
        if (state_host) (void)hipHostFree(state_host);
    }
};

extern "C" {

gs_status gs_tracker_create(gs_ctx ctx, gs_raster raster, uint32_t width, uint32_t height, gs_tracker* out) {
    GS_REQUIRE(ctx && raster && out, "null argument");
    GS_REQUIRE(width == raster->width && height == raster->height,
               "tracker size must equal the rasterizer's image size");
    *out = nullptr;
    auto t = std::make_unique<gs_tracker_s>();
    t->ctx = ctx;
    t->raster = raster;
    t->width = width;
    t->height = height;
    const uint32_t P = width * height;
    GS_CHECK(dalloc(&t->target, P, "tracker target"));
    t->n_partials = photometric_partials(P);
    GS_CHECK(dalloc(&t->partials, t->n_partials, "photometric partials"));
    GS_CHECK(dalloc(&t->state, 1, "track state"));
    GS_HIP(hipHostMalloc((void**)&t->state_host, sizeof(TrackState), hipHostMallocDefault));
    *out = t.release();
    return GS_OK;
}

void gs_tracker_destroy(gs_tracker t) { delete t; }
float* gs_tracker_target(gs_tracker t) { return t ? reinterpret_cast<float*>(t->target) : nullptr; }

gs_status gs_tracker_set_target(gs_tracker t, const float* frame_rgba) {
    GS_REQUIRE(t && frame_rgba, "null argument");
    GS_HIP(hipMemcpyAsync(t->target, frame_rgba, (size_t)t->width * t->height * sizeof(float4),
                          hipMemcpyHostToDevice, t->ctx->stream));
    GS_HIP(hipStreamSynchronize(t->ctx->stream));
    return GS_OK;
}

gs_status gs_tracker_track(gs_tracker tr, uint32_t surfels, gs_raster_camera* camera, float depth,
                           const gs_track_config* cfg, float* residuals, gs_track_report* report) {
    GS_REQUIRE(tr && camera && cfg && report, "null argument");
    GS_CHECK(check_camera(camera));
    GS_REQUIRE(cfg->iterations >= 1, "iterations must be >= 1");
    GS_REQUIRE(std::isfinite(depth) && std::isfinite(cfg->lr_rot) && std::isfinite(cfg->lr_center),
               "rates and depth must be finite");
    gs_raster r = tr->raster;
    if (surfels > r->capacity) return fail(GS_ERR_CAPACITY, "surfels exceeds the rasterizer's capacity");
    hipStream_t st = tr->ctx->stream;
    if (tr->residual_cap < cfg->iterations) {
        dfree(tr->residuals);
        GS_CHECK(dalloc(&tr->residuals, cfg->iterations, "track residuals"));
        tr->residual_cap = cfg->iterations;
    }
    // Initial state: moments 0, best = +inf, current = best = the camera.
    TrackState s0;
    std::memset(&s0, 0, sizeof(s0));
    s0.best = INFINITY;
    for (int k = 0; k < 4; ++k) s0.cur_q[k] = s0.best_q[k] = camera->quat[k];
    for (int k = 0; k < 3; ++k) s0.cur_c[k] = s0.best_c[k] = camera->center[k];
    *tr->state_host = s0;
    GS_HIP(hipMemcpyAsync(tr->state, tr->state_host, sizeof(TrackState), hipMemcpyHostToDevice, st));
    CamParams c = make_cam(*camera, r->width, r->height);
    c.lambda_dist = 0.0f;
    GS_CHECK(r->core.set_views(st, {c}, true));
    TrackArgs a;
    a.state = tr->state;
    a.partials = tr->partials;
    a.n_partials = tr->n_partials;
    a.npix = r->width * r->height;
    a.grad_cam = r->grad_cam;
    a.cam = r->core.cams_dev;
    a.residuals = tr->residuals;
    a.iterations = cfg->iterations;
    a.patience = cfg->patience;
    a.lr_rot = cfg->lr_rot;
    a.lr_center = cfg->lr_center * depth;  // lib.rs:217
    a.min_improvement = cfg->min_improvement;
    ParamGrads g{r->grad_pos, r->grad_scales, r->grad_quat, r->grad_opacity, r->grad_sh, r->grad_cam};
    // Iterations are enqueued in groups; the stop flag is read between
    // groups, so a stopped call runs at most a group past its last step
    // (those iterations record nothing and leave the pose alone).
    constexpr uint32_t kGroup = 8;
    for (uint32_t t = 1; t <= cfg->iterations; ++t) {
        GS_CHECK(r->core.forward(st, 0, 1, surfels, r->positions, r->scales, r->quats, r->opacities, r->sh));
        GS_HIP(launch_photometric(st, a.npix, 1.0f, 0.0f, r->core.out_color, tr->target, r->dl_dcolor,
                                  tr->partials));
        GS_CHECK(r->core.backward(st, 0, 1, surfels, r->positions, r->scales, r->quats, r->sh, r->dl_dcolor,
                                  nullptr, g, false, false));
        GS_HIP(launch_track_update(st, a, t));
        if (t % kGroup == 0 && t < cfg->iterations) {
            GS_HIP(hipMemcpyAsync(tr->state_host, tr->state, sizeof(TrackState), hipMemcpyDeviceToHost, st));
            GS_HIP(hipStreamSynchronize(st));
            if (tr->state_host->stopped) break;
        }
    }
    GS_HIP(hipMemcpyAsync(tr->state_host, tr->state, sizeof(TrackState), hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    const TrackState& s = *tr->state_host;
    if (residuals && s.iters) GS_HIP(hipMemcpy(residuals, tr->residuals, s.iters * sizeof(float), hipMemcpyDeviceToHost));
    report->iterations = s.iters;
    report->coverage = s.coverage;
    report->best = s.best;
    for (int k = 0; k < 4; ++k) report->last_quat[k] = s.cur_q[k];
    for (int k = 0; k < 3; ++k) report->last_center[k] = s.cur_c[k];
    // lib.rs:219-221: the camera returned is the best pose seen.
    for (int k = 0; k < 4; ++k) camera->quat[k] = s.best_q[k];
    for (int k = 0; k < 3; ++k) camera->center[k] = s.best_c[k];
    r->last_n = surfels;
    r->last_views = 1;
    r->last_batch = false;
    r->have_forward = true;
    return GS_OK;
}

gs_status gs_tracker_photometric(gs_tracker tr, void* stream, float gain, float bias, double* residual,
                                 double* coverage) {
    GS_REQUIRE(tr != nullptr, "tracker is null");
    gs_raster r = tr->raster;
    hipStream_t st = as_stream(stream, tr->ctx->stream);
    const uint32_t P = r->width * r->height;
    GS_HIP(launch_photometric(st, P, gain, bias, r->core.out_color, tr->target, r->dl_dcolor, tr->partials));
    if (residual || coverage) {
        std::vector<float2> p(tr->n_partials);
        GS_HIP(hipStreamSynchronize(st));
        GS_HIP(hipMemcpy(p.data(), tr->partials, p.size() * sizeof(float2), hipMemcpyDeviceToHost));
        double e = 0.0, c = 0.0;
        for (const auto& v : p) {
            e += v.x;
            c += v.y;
        }
        if (residual) *residual = e / (3.0 * P);
        if (coverage) *coverage = c / P;
    }
    return GS_OK;
}

}  // extern "C"

// ======================================================== 3DGS .ply export

namespace gs {

// writer.rs:31-51: the header of the baked export.
static std::string ply_header(uint32_t count) {
    std::string h = "ply\nformat binary_little_endian 1.0\n";
    h += "comment splattar baked export (surfels flattened to 3DGS)\n";
    h += "element vertex " + std::to_string(count) + "\n";
    for (const char* nm : {"x", "y", "z"}) h += std::string("property float ") + nm + "\n";
    for (int i = 0; i < 3; ++i) h += "property float f_dc_" + std::to_string(i) + "\n";
    for (int i = 0; i < 45; ++i) h += "property float f_rest_" + std::to_string(i) + "\n";
    h += "property float opacity\n";
    for (int i = 0; i < 3; ++i) h += "property float scale_" + std::to_string(i) + "\n";
    for (int i = 0; i < 4; ++i) h += "property float rot_" + std::to_string(i) + "\n";
    h += "end_header\n";
    return h;
}

static gs_status write_ply_from_device(gs_ctx ctx, const char* path, uint32_t count, const float4* pos,
                                       const float2* sc, const float4* q, const float* op, const float* sh) {
    // Rows are packed on the device in chunks of up to 2^18 surfels (62 MB),
    // copied to pinned host memory and appended to the file.
    const uint32_t chunk = std::min<uint32_t>(std::max<uint32_t>(count, 1u), 1u << 18);
    const size_t chunk_bytes = (size_t)chunk * kPlyRowFloats * sizeof(float);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(GS_ERR_IO, std::string("cannot open ") + path + " for writing");
    float* rows = nullptr;
    float* host = nullptr;
    gs_status st = GS_OK;
    const std::string header = ply_header(count);
    if (std::fwrite(header.data(), 1, header.size(), f) != header.size())
        st = fail(GS_ERR_IO, std::string("write failed: ") + path);
    hipError_t e = hipSuccess;
    if (st == GS_OK && count > 0) {
        e = hipMalloc((void**)&rows, chunk_bytes);
        if (e == hipSuccess) e = hipHostMalloc((void**)&host, chunk_bytes, hipHostMallocDefault);
        for (uint32_t first = 0; e == hipSuccess && st == GS_OK && first < count; first += chunk) {
            const uint32_t m = std::min(chunk, count - first);
            e = launch_ply_rows(ctx->stream, m, pos + first, sc + first, q + first, op + first,
                                sh + (size_t)first * kShFloats, rows);
            const size_t bytes = (size_t)m * kPlyRowFloats * sizeof(float);
            if (e == hipSuccess) e = hipMemcpyAsync(host, rows, bytes, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e == hipSuccess && std::fwrite(host, 1, bytes, f) != bytes)
                st = fail(GS_ERR_IO, std::string("write failed: ") + path);
        }
    }
    if (e != hipSuccess) st = hip_fail(e, "ply export");
    if (host) (void)hipHostFree(host);
    if (rows) (void)hipFree(rows);
    if (std::fclose(f) != 0 && st == GS_OK) st = fail(GS_ERR_IO, std::string("close failed: ") + path);
    return st;
}

}  // namespace gs

extern "C" {

gs_status gs_mcmc_add_noise(gs_ctx ctx, void* stream, uint32_t n, uint32_t seed, float sigma, float opacity_k,
                           float* pos_raw, const float* scales_act, const float* opac_act) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    GS_REQUIRE(n == 0 || (pos_raw && scales_act && opac_act), "null argument");
    NoiseArgs a;
    a.n = n;
    a.seed = seed;
    a.sigma = sigma;
    a.opacity_k = opacity_k;
    a.pos_raw = reinterpret_cast<float4*>(pos_raw);
    a.scales_act = reinterpret_cast<const float2*>(scales_act);
    a.opac_act = opac_act;
    GS_HIP(launch_add_noise(as_stream(stream, ctx->stream), a));
    return GS_OK;
}

// trainer.rs:1558-1650 (the host part of Trainer::mcmc_relocate).
gs_status gs_mcmc_sample_pairs(const float* opacity, const float* scales, uint32_t n, float dead_threshold,
                               float split_ratio, uint64_t* rng_state, uint32_t* pairs, uint32_t* n_pairs) {
    GS_REQUIRE(opacity && rng_state && pairs && n_pairs, "null argument");
    GS_REQUIRE(split_ratio <= 0.0f || scales != nullptr, "splits need the activated scales");
    *n_pairs = 0;
    std::vector<uint32_t> dead;
    for (uint32_t i = 0; i < n; ++i)
        if (opacity[i] < dead_threshold) dead.push_back(i);
    if (dead.empty()) return GS_OK;
    // Opacity-proportional alive sampling via a cumulative table.
    std::vector<double> cum(n);
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        if (opacity[i] >= dead_threshold) total += (double)opacity[i];
        cum[i] = total;
    }
    if (total <= 0.0) return GS_OK;
    std::vector<uint8_t> used(n, 0), spent(dead.size(), 0);
    uint32_t np = 0;
    uint64_t x = *rng_state;
    for (size_t di = 0; di < dead.size(); ++di) {
        const uint32_t d = dead[di];
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        const double r = (double)(x >> 11) / (double)(1ull << 53) * total;
        size_t a = (size_t)(std::lower_bound(cum.begin(), cum.end(), r) - cum.begin());
        a = std::min<size_t>(a, n - 1);
        if (a == d || used[a]) continue;
        used[a] = 1;
        spent[di] = 1;
        uint32_t* p = pairs + 4 * np++;
        p[0] = d;
        p[1] = (uint32_t)a;
        p[2] = 0;
        p[3] = 0;
    }
    *rng_state = x;
    // Oversized-surfel splits (trainer.rs:1609-1650).
    if (split_ratio > 0.0f) {
        auto size = [&](uint32_t i) { return std::max(scales[2 * (size_t)i], scales[2 * (size_t)i + 1]); };
        // f32 total order (Rust f32::total_cmp).
        auto key = [](float f) {
            int32_t b;
            std::memcpy(&b, &f, 4);
            return b ^ (int32_t)((uint32_t)(b >> 31) >> 1);
        };
        std::vector<float> alive;
        for (uint32_t i = 0; i < n; ++i)
            if (opacity[i] >= dead_threshold) alive.push_back(size(i));
        if (alive.size() >= 16) {
            std::sort(alive.begin(), alive.end(), [&](float a, float b) { return key(a) < key(b); });
            const float med = alive[alive.size() / 2];
            const float cap = split_ratio * med;
            std::vector<uint32_t> big;
            for (uint32_t i = 0; i < n; ++i)
                if (opacity[i] >= dead_threshold && size(i) > cap && !used[i]) big.push_back(i);
            std::stable_sort(big.begin(), big.end(),
                             [&](uint32_t a, uint32_t b) { return key(size(b)) < key(size(a)); });
            size_t li = 0;
            for (uint32_t b : big) {
                while (li < dead.size() && spent[li]) ++li;
                if (li >= dead.size()) break;
                used[b] = 1;
                uint32_t* p = pairs + 4 * np++;
                p[0] = dead[li++];
                p[1] = b;
                p[2] = 1;
                p[3] = 0;
            }
        }
    }
    *n_pairs = np;
    return GS_OK;
}

gs_status gs_ply_pack_rows(gs_ctx ctx, void* stream, uint32_t count, const float* positions, const float* scales,
                           const float* quats, const float* opacities, const float* sh, float* rows) {
    GS_REQUIRE(ctx != nullptr, "ctx is null");
    GS_REQUIRE(count == 0 || (positions && scales && quats && opacities && sh && rows), "null argument");
    GS_HIP(launch_ply_rows(as_stream(stream, ctx->stream), count, reinterpret_cast<const float4*>(positions),
                           reinterpret_cast<const float2*>(scales), reinterpret_cast<const float4*>(quats),
                           opacities, sh, rows));
    return GS_OK;
}

gs_status gs_write_3dgs_ply(gs_ctx ctx, const char* path, uint32_t count, const float* positions,
                            const float* scales, const float* quats, const float* opacities, const float* sh) {
    GS_REQUIRE(ctx != nullptr && path != nullptr, "null argument");
    GS_REQUIRE(count == 0 || (positions && scales && quats && opacities && sh), "null argument");
    GS_HIP(hipStreamSynchronize(ctx->stream));
    return write_ply_from_device(ctx, path, count, reinterpret_cast<const float4*>(positions),
                                 reinterpret_cast<const float2*>(scales), reinterpret_cast<const float4*>(quats),
                                 opacities, sh);
}

}  // extern "C"

// ============================================================== NormalLoss

struct gs_normal_loss_s {
    gs_ctx ctx = nullptr;
    uint32_t width = 0, height = 0;
    float lambda = 0.0f, focal = 1.0f, raw_lambda = 0.0f;
    const float4* color = nullptr;
    const float4* aux = nullptr;
    const float4* normal = nullptr;
    float4* dl_dnormal = nullptr;
    float4* dl_dcolor = nullptr;
    float4* loss_map = nullptr;
    float* partials = nullptr;
    uint32_t n_partials = 0;
    ~gs_normal_loss_s() {
        dfree(loss_map);
        dfree(partials);
    }
};

extern "C" {

gs_status gs_normal_loss_create(gs_ctx ctx, uint32_t width, uint32_t height, const float* out_color,
                                const float* out_aux, const float* out_normal, float* dl_dnormal,
                                float* dl_dcolor, gs_normal_loss* out) {
    GS_REQUIRE(ctx && out && out_color && out_aux && out_normal && dl_dnormal && dl_dcolor, "null argument");
    GS_REQUIRE(width > 0 && height > 0, "image size must be > 0");
    *out = nullptr;
    auto nl = std::make_unique<gs_normal_loss_s>();
    nl->ctx = ctx;
    nl->width = width;
    nl->height = height;
    nl->color = reinterpret_cast<const float4*>(out_color);
    nl->aux = reinterpret_cast<const float4*>(out_aux);
    nl->normal = reinterpret_cast<const float4*>(out_normal);
    nl->dl_dnormal = reinterpret_cast<float4*>(dl_dnormal);
    nl->dl_dcolor = reinterpret_cast<float4*>(dl_dcolor);
    GS_CHECK(dalloc(&nl->loss_map, (size_t)width * height, "normal loss map"));
    nl->n_partials = normal_loss_partials(width, height);
    GS_CHECK(dalloc(&nl->partials, nl->n_partials, "normal loss partials"));
    *out = nl.release();
    return GS_OK;
}

void gs_normal_loss_destroy(gs_normal_loss nl) { delete nl; }
float* gs_normal_loss_loss_map(gs_normal_loss nl) { return nl ? reinterpret_cast<float*>(nl->loss_map) : nullptr; }

gs_status gs_normal_loss_set_lambda(gs_normal_loss nl, float lambda, float focal) {
    GS_REQUIRE(nl != nullptr, "normal loss is null");
    GS_REQUIRE(std::isfinite(lambda) && lambda >= 0.0f, "lambda must be finite and >= 0");
    GS_REQUIRE(std::isfinite(focal) && focal > 0.0f, "focal must be finite and > 0");
    nl->raw_lambda = lambda;
    nl->lambda = lambda / (float)((double)nl->width * nl->height);  // normal_loss.rs:127
    nl->focal = focal;
    return GS_OK;
}

gs_status gs_normal_loss_encode(gs_normal_loss nl, void* stream) {
    GS_REQUIRE(nl != nullptr, "normal loss is null");
    NormalLossArgs a;
    a.width = nl->width;
    a.height = nl->height;
    a.focal = nl->focal;
    a.lambda = nl->lambda;
    a.alpha_min = 0.2f;  // normal_loss.rs:129
    a.color = nl->color;
    a.aux = nl->aux;
    a.normal = nl->normal;
    a.dl_dnormal = nl->dl_dnormal;
    a.dl_dcolor_w = reinterpret_cast<float*>(nl->dl_dcolor) + 3;
    a.loss_map = nl->loss_map;
    a.partials = nl->partials;
    GS_HIP(launch_normal_loss(as_stream(stream, nl->ctx->stream), a));
    return GS_OK;
}

gs_status gs_normal_loss_read(gs_normal_loss nl, double* loss) {
    GS_REQUIRE(nl && loss, "null argument");
    std::vector<float> p(nl->n_partials);
    GS_HIP(hipStreamSynchronize(nl->ctx->stream));
    GS_HIP(hipMemcpy(p.data(), nl->partials, p.size() * sizeof(float), hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (float v : p) sum += v;
    *loss = (double)nl->lambda * sum;
    return GS_OK;
}

}  // extern "C"

// ================================================================ Optimizer

struct OptimClass {
    std::string name;
    uint32_t n = 0;
    int activation = 0;
    float* raw = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    const float* grads = nullptr;
    float* activated = nullptr;
};

struct gs_optim_s {
    gs_ctx ctx = nullptr;
    std::vector<OptimClass> classes;
    AdamState* state = nullptr;
    // Learning rates, schedules and regularisers as set on the host; uploaded
    // into state->hyper on the stream of the next encode that follows a
    // change (the reference writes its uniforms at encode time too,
    // optim.rs:179-197), through a pinned staging copy.
    AdamHyper hyper{};
    bool hyper_dirty = true;
    AdamHyper* hyper_pinned = nullptr;
    hipEvent_t hyper_done = nullptr;
    bool hyper_pending = false;
    ~gs_optim_s() {
        for (auto& c : classes) {
            dfree(c.raw);
            dfree(c.m);
            dfree(c.v);
        }
        dfree(state);
        if (hyper_done) (void)hipEventDestroy(hyper_done);
        if (hyper_pinned) (void)hipHostFree(hyper_pinned);
    }
    AdamTable table(uint32_t* total_blocks) const {
        AdamTable t;
        std::memset(&t, 0, sizeof(t));
        uint32_t b = 0;
        for (size_t k = 0; k < classes.size(); ++k) {
            const OptimClass& c = classes[k];
            t.c[k] = AdamClass{c.raw, c.grads, c.m, c.v, c.activated, c.n, c.activation, b};
            b += adam_blocks_for(c.n);
        }
        t.n_classes = (uint32_t)classes.size();
        *total_blocks = b;
        return t;
    }
    OptimClass* find(const char* name) {
        if (!name) return nullptr;
        for (auto& c : classes)
            if (c.name == name) return &c;
        return nullptr;
    }
    int index_of(const char* name) const {
        if (!name) return -1;
        for (size_t k = 0; k < classes.size(); ++k)
            if (classes[k].name == name) return (int)k;
        return -1;
    }
    // Enqueues the upload of changed hyperparameters on `st` (not while st
    // is being captured: a captured step replays the device copy as it was
    // uploaded last, so callers flush before capturing).
    gs_status sync_hyper(hipStream_t st) {
        if (!hyper_dirty) return GS_OK;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GS_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(GS_ERR_INVALID, "optimizer learning rates changed during stream capture; "
                                        "encode an uncaptured step (or gs_optim_encode_activate) first");
        if (hyper_pending) GS_HIP(hipEventSynchronize(hyper_done));
        *hyper_pinned = hyper;
        GS_HIP(hipMemcpyAsync(&state->hyper, hyper_pinned, sizeof(AdamHyper), hipMemcpyHostToDevice, st));
        GS_HIP(hipEventRecord(hyper_done, st));
        hyper_pending = true;
        hyper_dirty = false;
        return GS_OK;
    }
};

// Serialized optimizer state (gs_optim_export_state): a header, the class
// descriptors, the device AdamState and each class's raw, m and v.
namespace {
constexpr uint32_t kStateMagic = 0x534F5347u;  // "GSOS"
constexpr uint32_t kStateVersion = 1u;
struct StateHeader {
    uint32_t magic, version, n_classes, state_bytes;
};
struct StateClass {
    char name[32];
    uint32_t n;
    int32_t activation;
};
}  // namespace

static uint64_t optim_state_bytes(const gs_optim_s* o) {
    uint64_t b = sizeof(StateHeader) + o->classes.size() * sizeof(StateClass) + sizeof(AdamState);
    for (const auto& c : o->classes) b += 3ull * c.n * sizeof(float);
    return b;
}

extern "C" {

gs_status gs_optim_create(gs_ctx ctx, gs_optim* out) {
    GS_REQUIRE(ctx && out, "null argument");
    *out = nullptr;
    auto o = std::make_unique<gs_optim_s>();
    o->ctx = ctx;
    GS_CHECK(dalloc(&o->state, 1, "adam state"));
    GS_HIP(hipHostMalloc((void**)&o->hyper_pinned, sizeof(AdamHyper), hipHostMallocDefault));
    GS_HIP(hipEventCreateWithFlags(&o->hyper_done, hipEventDisableTiming));
    AdamState s0;
    std::memset(&s0, 0, sizeof(s0));
    s0.p1 = s0.p2 = s0.bc1_inv = s0.bc2_inv = 1.0f;
    GS_HIP(hipMemcpy(o->state, &s0, sizeof(s0), hipMemcpyHostToDevice));
    *out = o.release();
    return GS_OK;
}

void gs_optim_destroy(gs_optim o) {
    if (o && o->hyper_pending) (void)hipEventSynchronize(o->hyper_done);
    delete o;
}

gs_status gs_optim_add_class(gs_optim o, const char* name, uint32_t n, int activation, float lr,
                             const float* grads, float* activated) {
    GS_REQUIRE(o && name && grads && activated, "null argument");
    GS_REQUIRE(activation >= 0 && activation <= 2, "activation must be 0, 1 or 2");
    GS_REQUIRE(o->classes.size() < (size_t)kMaxAdamClasses, "too many optimizer classes");
    GS_REQUIRE(o->find(name) == nullptr, "duplicate optimizer class name");
    GS_REQUIRE(std::strlen(name) < sizeof(StateClass::name), "optimizer class name longer than 31 bytes");
    GS_REQUIRE(std::isfinite(lr), "learning rate must be finite");
    OptimClass c;
    c.name = name;
    c.n = n;
    c.activation = activation;
    c.grads = grads;
    c.activated = activated;
    GS_CHECK(dalloc(&c.raw, n, "adam raw"));
    GS_CHECK(dalloc(&c.m, n, "adam m"));
    GS_CHECK(dalloc(&c.v, n, "adam v"));
    const size_t k = o->classes.size();
    o->hyper.sched[k] = AdamSchedule{lr, 1.0f, 0u, 0u};
    o->hyper.reg[k] = 0.0f;
    o->hyper_dirty = true;
    o->classes.push_back(c);
    return GS_OK;
}

float* gs_optim_class_raw(gs_optim o, const char* name) {
    if (!o) return nullptr;
    OptimClass* c = o->find(name);
    return c ? c->raw : nullptr;
}

gs_status gs_optim_set_lr(gs_optim o, const char* name, float lr) {
    GS_REQUIRE(o != nullptr, "optim is null");
    const int k = o->index_of(name);
    GS_REQUIRE(k >= 0, "unknown optimizer class");
    GS_REQUIRE(std::isfinite(lr), "learning rate must be finite");
    o->hyper.sched[k] = AdamSchedule{lr, 1.0f, 0u, 0u};
    o->hyper_dirty = true;
    return GS_OK;
}

gs_status gs_optim_set_lr_schedule(gs_optim o, const char* name, float base, float final_factor,
                                   uint32_t iters) {
    GS_REQUIRE(o != nullptr, "optim is null");
    const int k = o->index_of(name);
    GS_REQUIRE(k >= 0, "unknown optimizer class");
    GS_REQUIRE(std::isfinite(base) && std::isfinite(final_factor) && final_factor > 0.0f,
               "schedule base must be finite and final_factor finite and > 0");
    o->hyper.sched[k] = AdamSchedule{base, final_factor, iters, 0u};
    o->hyper_dirty = true;
    return GS_OK;
}

gs_status gs_optim_set_reg(gs_optim o, const char* name, float reg) {
    GS_REQUIRE(o != nullptr, "optim is null");
    const int k = o->index_of(name);
    GS_REQUIRE(k >= 0, "unknown optimizer class");
    GS_REQUIRE(std::isfinite(reg), "regulariser must be finite");
    o->hyper.reg[k] = reg;
    o->hyper_dirty = true;
    return GS_OK;
}

gs_status gs_optim_encode_step(gs_optim o, void* stream) {
    GS_REQUIRE(o != nullptr, "optim is null");
    hipStream_t st = as_stream(stream, o->ctx->stream);
    GS_CHECK(o->sync_hyper(st));
    uint32_t blocks = 0;
    AdamTable t = o->table(&blocks);
    GS_HIP(launch_adam_tick(st, o->state));
    GS_HIP(launch_adam(st, t, blocks, o->state, true));
    return GS_OK;
}

gs_status gs_optim_encode_activate(gs_optim o, void* stream) {
    GS_REQUIRE(o != nullptr, "optim is null");
    hipStream_t st = as_stream(stream, o->ctx->stream);
    GS_CHECK(o->sync_hyper(st));
    uint32_t blocks = 0;
    AdamTable t = o->table(&blocks);
    GS_HIP(launch_adam(st, t, blocks, o->state, false));
    return GS_OK;
}

gs_status gs_optim_step_count(gs_optim o, uint32_t* t) {
    GS_REQUIRE(o && t, "null argument");
    AdamState s;
    GS_HIP(hipStreamSynchronize(o->ctx->stream));
    GS_HIP(hipMemcpy(&s, o->state, sizeof(s), hipMemcpyDeviceToHost));
    *t = s.t;
    return GS_OK;
}

gs_status gs_optim_current_lr(gs_optim o, const char* name, float* lr) {
    GS_REQUIRE(o && lr, "null argument");
    const int k = o->index_of(name);
    GS_REQUIRE(k >= 0, "unknown optimizer class");
    AdamState s;
    GS_HIP(hipStreamSynchronize(o->ctx->stream));
    GS_HIP(hipMemcpy(&s, o->state, sizeof(s), hipMemcpyDeviceToHost));
    *lr = s.lr[k];
    return GS_OK;
}

gs_status gs_optim_state_bytes(gs_optim o, uint64_t* bytes) {
    GS_REQUIRE(o && bytes, "null argument");
    *bytes = optim_state_bytes(o);
    return GS_OK;
}

gs_status gs_optim_export_state(gs_optim o, void* dst, uint64_t bytes) {
    GS_REQUIRE(o && dst, "null argument");
    GS_REQUIRE(bytes >= optim_state_bytes(o), "buffer smaller than gs_optim_state_bytes");
    hipStream_t st = o->ctx->stream;
    GS_CHECK(o->sync_hyper(st));
    GS_HIP(hipStreamSynchronize(st));
    char* p = static_cast<char*>(dst);
    StateHeader h{kStateMagic, kStateVersion, (uint32_t)o->classes.size(), (uint32_t)sizeof(AdamState)};
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (const auto& c : o->classes) {
        StateClass sc;
        std::memset(&sc, 0, sizeof(sc));
        std::memcpy(sc.name, c.name.data(), c.name.size());
        sc.n = c.n;
        sc.activation = c.activation;
        std::memcpy(p, &sc, sizeof(sc));
        p += sizeof(sc);
    }
    GS_HIP(hipMemcpy(p, o->state, sizeof(AdamState), hipMemcpyDeviceToHost));
    p += sizeof(AdamState);
    for (const auto& c : o->classes) {
        for (const float* src : {(const float*)c.raw, (const float*)c.m, (const float*)c.v}) {
            GS_HIP(hipMemcpy(p, src, (size_t)c.n * sizeof(float), hipMemcpyDeviceToHost));
            p += (size_t)c.n * sizeof(float);
        }
    }
    return GS_OK;
}

gs_status gs_optim_import_state(gs_optim o, const void* src, uint64_t bytes) {
    GS_REQUIRE(o && src, "null argument");
    GS_REQUIRE(bytes >= sizeof(StateHeader), "state blob too small");
    const char* p = static_cast<const char*>(src);
    StateHeader h;
    std::memcpy(&h, p, sizeof(h));
    GS_REQUIRE(h.magic == kStateMagic, "not an optimizer state blob (bad magic)");
    GS_REQUIRE(h.version == kStateVersion, "unsupported optimizer state version");
    GS_REQUIRE(h.state_bytes == sizeof(AdamState), "optimizer state layout differs from this build");
    GS_REQUIRE(h.n_classes == o->classes.size(), "optimizer state has a different number of classes");
    GS_REQUIRE(bytes >= optim_state_bytes(o), "state blob truncated");
    p += sizeof(h);
    for (const auto& c : o->classes) {
        StateClass sc;
        std::memcpy(&sc, p, sizeof(sc));
        p += sizeof(sc);
        sc.name[sizeof(sc.name) - 1] = 0;
        GS_REQUIRE(c.name == sc.name && c.n == sc.n && c.activation == sc.activation,
                   "optimizer state classes (name, size, activation) differ from this optimizer's");
    }
    AdamState s;
    std::memcpy(&s, p, sizeof(s));
    p += sizeof(s);
    hipStream_t st = o->ctx->stream;
    if (o->hyper_pending) GS_HIP(hipEventSynchronize(o->hyper_done));
    GS_HIP(hipStreamSynchronize(st));
    GS_HIP(hipMemcpy(o->state, &s, sizeof(s), hipMemcpyHostToDevice));
    o->hyper = s.hyper;
    o->hyper_dirty = false;
    for (const auto& c : o->classes) {
        for (float* dst : {c.raw, c.m, c.v}) {
            GS_HIP(hipMemcpy(dst, p, (size_t)c.n * sizeof(float), hipMemcpyHostToDevice));
            p += (size_t)c.n * sizeof(float);
        }
    }
    uint32_t blocks = 0;
    AdamTable t = o->table(&blocks);
    GS_HIP(launch_adam(st, t, blocks, o->state, false));
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

gs_status gs_optim_class_count(gs_optim o, uint32_t* count) {
    GS_REQUIRE(o && count, "null argument");
    *count = (uint32_t)o->classes.size();
    return GS_OK;
}

gs_status gs_optim_class_info(gs_optim o, uint32_t index, const char** name, uint64_t* bytes) {
    GS_REQUIRE(o && name && bytes, "null argument");
    GS_REQUIRE(index < o->classes.size(), "class index out of range");
    *name = o->classes[index].name.c_str();
    *bytes = (uint64_t)o->classes[index].n * sizeof(float);
    return GS_OK;
}

gs_status gs_optim_encode_snapshot(gs_optim o, void* stream, float* const* bufs, uint32_t count) {
    GS_REQUIRE(o && bufs, "null argument");
    GS_REQUIRE(count == o->classes.size(), "one buffer per optimizer class");
    for (uint32_t k = 0; k < count; ++k) GS_REQUIRE(bufs[k] != nullptr, "null snapshot buffer");
    hipStream_t st = as_stream(stream, o->ctx->stream);
    for (uint32_t k = 0; k < count; ++k)
        GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(bufs[k]),
                               reinterpret_cast<const uint32_t*>(o->classes[k].raw), o->classes[k].n));
    return GS_OK;
}

gs_status gs_optim_encode_restore(gs_optim o, void* stream, const float* const* bufs, uint32_t count) {
    GS_REQUIRE(o && bufs, "null argument");
    GS_REQUIRE(count == o->classes.size(), "one buffer per optimizer class");
    for (uint32_t k = 0; k < count; ++k) GS_REQUIRE(bufs[k] != nullptr, "null snapshot buffer");
    hipStream_t st = as_stream(stream, o->ctx->stream);
    for (uint32_t k = 0; k < count; ++k)
        GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(o->classes[k].raw),
                               reinterpret_cast<const uint32_t*>(bufs[k]), o->classes[k].n));
    return GS_OK;
}

}  // extern "C"

// ================================================================== Trainer

struct TrainViewDev {
    uint32_t width, height;
    gs_raster_camera camera;
    size_t target_offset;   // in float4 units
    uint32_t num_surfels;   // surfels [0, num_surfels) are part of this view
};

struct gs_trainer_s {
    gs_ctx ctx = nullptr;
    gs_train_config cfg{};
    uint32_t n = 0;
    std::vector<TrainViewDev> views;
    float4* targets = nullptr;
    // The targets' window mean and second moment per channel (fixed with the
    // targets, so computed once at creation; the step's loss reads them).
    float2* target_stats = nullptr;
    // activated parameters + gradients (layout of gs_raster_buffers)
    float4* positions = nullptr;
    float2* scales = nullptr;
    float4* quats = nullptr;
    float* opacities = nullptr;
    float* sh = nullptr;
    float4* grad_pos = nullptr;
    float2* grad_scales = nullptr;
    float4* grad_quat = nullptr;
    float* grad_opacity = nullptr;
    float* grad_sh = nullptr;
    float4* dl_dcolor = nullptr;       // cap_pixels (batched layout)
    double* loss_partials = nullptr;   // per view slot
    uint32_t partials_per_view = 0;
    gs_optim optim = nullptr;
    gs::RasterCore core;
    uint32_t max_w = 0, max_h = 0;
    std::vector<uint32_t> last_ids;
    bool last_batched = false;
    uint32_t sh_degree = 3;
    // graph mode
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint32_t> graph_shape;   // (w, h) per view of the captured step
    bool tuned = false;
    gs::Profiler prof;
    // Pipelined batched step (opt-in): the loss of one half of the views runs
    // on a second stream while the raster passes of the other half run.
    bool pipeline = false;
    hipStream_t side = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // Batched steps run the geometry backward and Adam as one kernel; the
    // summed parameter gradients reach grad_* only when asked for.
    bool keep_grads = false;
    // Recorded on the stream of every step, so blocking reads and renders
    // on the trainer's own stream are ordered after a step enqueued on a
    // caller's stream.
    hipEvent_t step_done = nullptr;
    bool step_pending = false;

    // Blocks until the last step and everything on the trainer's stream
    // have finished.
    gs_status wait_idle() {
        if (step_pending) GS_HIP(hipEventSynchronize(step_done));
        GS_HIP(hipStreamSynchronize(ctx->stream));
        return GS_OK;
    }
    // Orders later work on the trainer's stream after the last step.
    gs_status order_after_step() {
        if (step_pending) GS_HIP(hipStreamWaitEvent(ctx->stream, step_done, 0));
        return GS_OK;
    }

    ~gs_trainer_s() {
        if (step_pending) (void)hipEventSynchronize(step_done);
        if (step_done) (void)hipEventDestroy(step_done);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        dfree(saved);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (optim) gs_optim_destroy(optim);
        dfree(targets);
        dfree(target_stats);
        dfree(positions);
        dfree(scales);
        dfree(quats);
        dfree(opacities);
        dfree(sh);
        dfree(grad_pos);
        dfree(grad_scales);
        dfree(grad_quat);
        dfree(grad_opacity);
        dfree(grad_sh);
        dfree(dl_dcolor);
        dfree(loss_partials);
        dfree(reloc_pairs);
    }

    gs_status enqueue_step(hipStream_t st, const uint32_t* ids, uint32_t k, bool batched,
                           bool profile = false);
    // MCMC exploration noise after the optimizer step (mcmc.wgsl, trainer.rs:
    // 1256-1272): sigma = mcmc_noise * the step's position rate, seed from
    // the step count, both read on the device.
    float mcmc_noise = 0.0f;
    uint32_t* reloc_pairs = nullptr;  // device (dead, alive, mode, 0) pairs, n of them at most
    gs_status noise_step(hipStream_t st) {
        if (!(mcmc_noise > 0.0f)) return GS_OK;
        NoiseArgs a;
        a.n = n;
        a.opacity_k = 5.0f;  // trainer.rs:1262
        a.state = optim->state;
        a.noise_scale = mcmc_noise;
        a.lr_class = 0;
        a.pos_raw = reinterpret_cast<float4*>(optim->classes[0].raw);
        a.scales_act = scales;
        a.opac_act = opacities;
        GS_HIP(launch_add_noise(st, a));
        return GS_OK;
    }
    gs_status enqueue_pipelined(hipStream_t st, const uint32_t* ids, uint32_t k);
    // Geometry backward of the k views of a batched step fused with the Adam
    // step (bias-correction tick first).
    gs_status geom_adam(hipStream_t st, uint32_t k) {
        ParamGrads keep{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        if (keep_grads) keep = ParamGrads{grad_pos, grad_scales, grad_quat, grad_opacity, grad_sh, nullptr};
        uint32_t blocks = 0;
        const AdamTable tab = optim->table(&blocks);
        GS_HIP(launch_adam_tick(st, optim->state));
        GS_HIP(launch_geom_adam(st, core.cams_dev, k, n, core.grad_coef, tab, optim->state, keep));
        return GS_OK;
    }
    // L1 + D-SSIM of view slot j (the coefficient scratch is shared: the
    // losses of one step run in stream order on one stream).
    gs_status loss_view(hipStream_t s, const uint32_t* ids, uint32_t j) {
        const TrainViewDev& tv = views[ids[j]];
        const CamParams& c = core.cams[j];
        GS_HIP(launch_l1_ssim(s, tv.width, tv.height, cfg.lambda, cfg.dssim_scale, core.out_color + c.pix_base,
                              targets, dl_dcolor + c.pix_base,
                              loss_partials + (size_t)j * partials_per_view * 2, core.cams_dev + j,
                              target_stats));
        return GS_OK;
    }

    // L1 + D-SSIM of view slots [j0, j0 + k) of a batched step, in one
    // launch where the loss kernel supports it.
    gs_status loss_views(hipStream_t s, const uint32_t* ids, uint32_t j0, uint32_t k) {
        if (k == 1) {
            for (uint32_t j = j0; j < j0 + k; ++j) GS_CHECK(loss_view(s, ids, j));
            return GS_OK;
        }
        std::vector<uint32_t> ws(k), hs(k), pb(k);
        for (uint32_t j = 0; j < k; ++j) {
            const TrainViewDev& tv = views[ids[j0 + j]];
            ws[j] = tv.width;
            hs[j] = tv.height;
            pb[j] = core.cams[j0 + j].pix_base;
        }
        GS_HIP(launch_l1_ssim_views(s, j0, k, ws.data(), hs.data(), pb.data(), cfg.lambda, cfg.dssim_scale,
                                    core.out_color, targets, dl_dcolor, loss_partials, partials_per_view,
                                    core.cams_dev, target_stats));
        return GS_OK;
    }

    // Training-state snapshot (raw, m, v of every optimizer class and the
    // Adam step state) in one device buffer; restore also re-activates. The
    // copies are kernels on the trainer's stream.
    float* saved = nullptr;
    size_t saved_floats = 0;
    gs_status save_state() {
        hipStream_t st = ctx->stream;
        GS_CHECK(order_after_step());
        size_t total = 0;
        for (const auto& c : optim->classes) total += 3 * (size_t)c.n;
        if (total != saved_floats) {
            dfree(saved);
            saved = nullptr;
            GS_CHECK(dalloc(&saved, total + sizeof(gs::AdamState) / 4 + 1, "training-state snapshot"));
            saved_floats = total;
        }
        size_t off = 0;
        for (const auto& c : optim->classes) {
            for (const float* src : {c.raw, c.m, c.v}) {
                GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(saved + off),
                                       reinterpret_cast<const uint32_t*>(src), c.n));
                off += c.n;
            }
        }
        GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(saved + off),
                               reinterpret_cast<const uint32_t*>(optim->state), sizeof(gs::AdamState) / 4));
        GS_HIP(hipStreamSynchronize(st));
        return GS_OK;
    }
    gs_status restore_state() {
        if (saved == nullptr) return fail(GS_ERR_INVALID, "no saved training state");
        hipStream_t st = ctx->stream;
        GS_CHECK(order_after_step());
        size_t off = 0;
        for (const auto& c : optim->classes) {
            for (float* dst : {c.raw, c.m, c.v}) {
                GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(dst),
                                       reinterpret_cast<const uint32_t*>(saved + off), c.n));
                off += c.n;
            }
        }
        GS_HIP(launch_copy_u32(st, reinterpret_cast<uint32_t*>(optim->state),
                               reinterpret_cast<const uint32_t*>(saved + off), sizeof(gs::AdamState) / 4));
        GS_CHECK(gs_optim_encode_activate(optim, st));
        GS_HIP(hipStreamSynchronize(st));
        return GS_OK;
    }
};

gs_status gs_trainer_s::enqueue_step(hipStream_t st, const uint32_t* ids, uint32_t k, bool batched,
                                     bool profile) {
    ParamGrads g{grad_pos, grad_scales, grad_quat, grad_opacity, grad_sh, nullptr};
    gs::Profiler* P = profile ? &prof : nullptr;
    if (batched && P == nullptr && pipeline && k >= 2) return enqueue_pipelined(st, ids, k);
    if (batched) {
        GS_CHECK(core.forward(st, 0, k, n, positions, scales, quats, opacities, sh, P));
        if (P) GS_CHECK(P->mark(st, GS_STAGE_LOSS));
        GS_CHECK(loss_views(st, ids, 0, k));
        GS_CHECK(core.backward(st, 0, k, n, positions, scales, quats, sh, dl_dcolor, nullptr, g, false,
                               false, P, false));
        if (P) GS_CHECK(P->mark(st, GS_STAGE_GEOM_BWD));
        GS_CHECK(geom_adam(st, k));
    } else {
        for (uint32_t j = 0; j < k; ++j) {
            const TrainViewDev& tv = views[ids[j]];
            GS_CHECK(core.forward(st, j, 1, n, positions, scales, quats, opacities, sh, P));
            if (P) GS_CHECK(P->mark(st, GS_STAGE_LOSS));
            GS_HIP(launch_l1_ssim(st, tv.width, tv.height, cfg.lambda, cfg.dssim_scale, core.out_color,
                                  targets, dl_dcolor, loss_partials + (size_t)j * partials_per_view * 2,
                                  core.cams_dev + j, target_stats));
            GS_CHECK(core.backward(st, j, 1, n, positions, scales, quats, sh, dl_dcolor, nullptr, g,
                                   j > 0, false, P));
        }
        if (P) GS_CHECK(P->mark(st, GS_STAGE_ADAM));
        GS_CHECK(gs_optim_encode_step(optim, st));
    }
    GS_CHECK(noise_step(st));
    if (P) {
        GS_CHECK(P->mark(st, -1));
        P->steps += 1;
    }
    return GS_OK;
}

// Batched step with the views split in halves A = [0, k/2), B = [k/2, k):
//   stream st:   bin all views | fwd(A) | fwd(B) | .. bwd(A) | .. bwd(B) | geom | Adam
//   side stream:                   loss(A) after fwd(A) | loss(B) after fwd(B)
// so each half's memory-bound loss runs beside a VALU-bound raster pass of
// the other half; bwd(A) waits for loss(A), bwd(B) for loss(B). Events fork
// and join the side stream, so the step captures into one graph as well.
gs_status gs_trainer_s::enqueue_pipelined(hipStream_t st, const uint32_t* ids, uint32_t k) {
    if (side == nullptr) {
        GS_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        for (auto& e : ev) GS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    GS_CHECK(core.forward(st, 0, k, n, positions, scales, quats, opacities, sh, nullptr, false));
    const uint32_t ka = k / 2;
    const uint32_t ta = core.cams[ka].tile_base;
    const uint32_t tt = core.total_tiles;
    GS_HIP(launch_fill_u32(st, reinterpret_cast<uint32_t*>(core.grad_coef), 0u,
                           (size_t)n * k * grad_stride(false)));
    GS_CHECK(core.raster_fwd_range(st, 0, 0, k, 0, ta));
    GS_HIP(hipEventRecord(ev[0], st));
    GS_CHECK(core.raster_fwd_range(st, 1, 0, k, ta, tt - ta));
    GS_HIP(hipEventRecord(ev[1], st));
    GS_HIP(hipStreamWaitEvent(side, ev[0], 0));
    GS_CHECK(loss_views(side, ids, 0, ka));
    GS_HIP(hipEventRecord(ev[2], side));
    GS_HIP(hipStreamWaitEvent(side, ev[1], 0));
    GS_CHECK(loss_views(side, ids, ka, k - ka));
    GS_HIP(hipEventRecord(ev[3], side));
    GS_HIP(hipStreamWaitEvent(st, ev[2], 0));
    GS_CHECK(core.raster_bwd_range(st, 0, 0, k, 0, ta, dl_dcolor));
    GS_HIP(hipStreamWaitEvent(st, ev[3], 0));
    GS_CHECK(core.raster_bwd_range(st, 1, 0, k, ta, tt - ta, dl_dcolor));
    GS_CHECK(geom_adam(st, k));
    return noise_step(st);
}

extern "C" {

gs_status gs_trainer_create(gs_ctx ctx, const gs_train_config* cfg, const gs_scene_input* init,
                            uint32_t n_views, const gs_train_view* views, gs_trainer* out) {
    GS_REQUIRE(ctx && cfg && init && views && out, "null argument");
    GS_REQUIRE(n_views > 0, "need at least one view");
    GS_REQUIRE(cfg->max_views_per_step >= 1 && cfg->max_views_per_step <= 64,
               "max_views_per_step must be in 1..64");
    GS_REQUIRE(cfg->binning <= GS_BINNING_TIGHT, "unknown binning mode");
    GS_REQUIRE(cfg->digit_bits == 0 || cfg->digit_bits == 4 || cfg->digit_bits == 8,
               "digit_bits must be 0, 4 or 8");
    GS_REQUIRE(init->n > 0, "need at least one surfel");
    *out = nullptr;
    auto t = std::make_unique<gs_trainer_s>();
    t->ctx = ctx;
    t->cfg = *cfg;
    t->n = init->n;
    const uint32_t n = init->n;
    size_t toff = 0;
    uint32_t max_tiles = 0, max_px = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        const gs_train_view& tv = views[v];
        GS_REQUIRE(tv.width > 0 && tv.height > 0 && tv.target_rgba != nullptr, "bad view");
        GS_REQUIRE(toff + (size_t)tv.width * tv.height < (1ull << 32), "targets exceed 2^32 pixels");
        GS_CHECK(check_camera(&tv.camera));
        GS_REQUIRE(tv.num_surfels <= init->n, "view num_surfels exceeds the scene");
        t->views.push_back(TrainViewDev{tv.width, tv.height, tv.camera, toff,
                                        tv.num_surfels ? tv.num_surfels : init->n});
        toff += (size_t)tv.width * tv.height;
        // Room for the smallest work tile (8 px).
        max_tiles = std::max(max_tiles, div_up(tv.width, 8) * div_up(tv.height, 8));
        max_px = std::max(max_px, tv.width * tv.height);
        t->max_w = std::max(t->max_w, tv.width);
        t->max_h = std::max(t->max_h, tv.height);
    }
    const uint32_t V = cfg->max_views_per_step;
    const uint64_t per_view_entries = cfg->entries_per_view ? cfg->entries_per_view : (uint64_t)n * 64;
    const uint64_t entries = per_view_entries * V;
    if (entries >= (1ull << 30)) return fail(GS_ERR_CAPACITY, "entries per step must be < 2^30");
    GS_CHECK(t->core.init(ctx, n, V, (uint32_t)entries, max_tiles * V, max_px * V));
    t->core.digit_bits = cfg->digit_bits ? cfg->digit_bits : 8;
    t->core.binning = cfg->binning;
    // The step's losses are colour-only (L1 + D-SSIM): the forward skips the
    // depth and normal images (aux.y = 0; out_normal is left untouched).
    t->core.aux_outputs = false;
    t->core.lean = true;
    {
        const char* env = std::getenv("GS_PIPELINE");
        t->pipeline = env != nullptr && env[0] == '1';
    }
    GS_CHECK(dalloc(&t->targets, toff, "targets"));
    {
        std::vector<float> buf;
        for (uint32_t v = 0; v < n_views; ++v) {
            const size_t P = (size_t)views[v].width * views[v].height;
            GS_HIP(hipMemcpy(t->targets + t->views[v].target_offset, views[v].target_rgba, P * 16,
                             hipMemcpyHostToDevice));
        }
    }
    GS_CHECK(dalloc(&t->target_stats, toff * 3, "target statistics"));
    for (uint32_t v = 0; v < n_views; ++v) {
        const TrainViewDev& tv = t->views[v];
        GS_HIP(launch_target_stats(ctx->stream, tv.width, tv.height, t->targets + tv.target_offset,
                                   t->target_stats + tv.target_offset * 3));
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
    GS_CHECK(dalloc(&t->positions, n, "positions"));
    GS_CHECK(dalloc(&t->scales, n, "scales"));
    GS_CHECK(dalloc(&t->quats, n, "quats"));
    GS_CHECK(dalloc(&t->opacities, n, "opacities"));
    GS_CHECK(dalloc(&t->sh, (size_t)n * kShFloats, "sh"));
    GS_CHECK(dalloc(&t->grad_pos, n, "grad_pos"));
    GS_CHECK(dalloc(&t->grad_scales, n, "grad_scales"));
    GS_CHECK(dalloc(&t->grad_quat, n, "grad_quat"));
    GS_CHECK(dalloc(&t->grad_opacity, n, "grad_opacity"));
    GS_CHECK(dalloc(&t->grad_sh, (size_t)n * kShFloats, "grad_sh"));
    GS_CHECK(dalloc(&t->dl_dcolor, (size_t)max_px * V, "dl_dcolor"));
    t->partials_per_view = l1_ssim_partials(t->max_w, t->max_h);
    GS_CHECK(dalloc(&t->loss_partials, (size_t)t->partials_per_view * 2 * V, "loss partials"));
    GS_CHECK(upload_scene_into(ctx, init, n, t->positions, t->scales, t->quats, t->opacities, t->sh));
    t->sh_degree = 0;
    while ((t->sh_degree + 1) * (t->sh_degree + 1) < init->sh_coeffs) ++t->sh_degree;
    // Optimizer classes (trainer.rs:757-761) with raw = log / logit / identity
    // of the activated init (trainer.rs:763-792).
    GS_CHECK(gs_optim_create(ctx, &t->optim));
    struct Cls {
        const char* name;
        uint32_t count;
        int act;
        float lr;
        const float* grads;
        float* activated;
    } cls[5] = {
        {"pos", n * 4, GS_ACT_IDENTITY, cfg->lr_pos, (const float*)t->grad_pos, (float*)t->positions},
        {"scales", n * 2, GS_ACT_EXP, cfg->lr_scales, (const float*)t->grad_scales, (float*)t->scales},
        {"quat", n * 4, GS_ACT_IDENTITY, cfg->lr_quat, (const float*)t->grad_quat, (float*)t->quats},
        {"opacity", n, GS_ACT_SIGMOID, cfg->lr_opacity, t->grad_opacity, t->opacities},
        {"sh", n * kShFloats, GS_ACT_IDENTITY, cfg->lr_sh, t->grad_sh, t->sh},
    };
    for (auto& c : cls) GS_CHECK(gs_optim_add_class(t->optim, c.name, c.count, c.act, c.lr, c.grads, c.activated));
    // Exponential position-LR decay (trainer.rs:1149-1151), evaluated per
    // step on the device from the step counter.
    if (cfg->iters > 0 && cfg->pos_lr_final_factor > 0.0f)
        GS_CHECK(gs_optim_set_lr_schedule(t->optim, "pos", cfg->lr_pos, cfg->pos_lr_final_factor, cfg->iters));
    GS_HIP(hipEventCreateWithFlags(&t->step_done, hipEventDisableTiming));
    {
        // raw init on the host (log / logit with the reference's clamps).
        std::vector<float> pos4(n * 4), sc(n * 2), q(init->quats, init->quats + n * 4), op(n),
            sh48((size_t)n * kShFloats, 0.0f);
        for (uint32_t i = 0; i < n; ++i) {
            pos4[4 * i] = init->positions[3 * i];
            pos4[4 * i + 1] = init->positions[3 * i + 1];
            pos4[4 * i + 2] = init->positions[3 * i + 2];
            pos4[4 * i + 3] = 1.0f;
            sc[2 * i] = std::log(std::max(init->scales[2 * i], 1e-6f));
            sc[2 * i + 1] = std::log(std::max(init->scales[2 * i + 1], 1e-6f));
            const float o = std::min(std::max(init->opacities[i], 1e-4f), 1.0f - 1e-4f);
            op[i] = std::log(o / (1.0f - o));
            std::memcpy(&sh48[(size_t)i * kShFloats], init->sh + (size_t)i * init->sh_coeffs * 3,
                        sizeof(float) * init->sh_coeffs * 3);
        }
        GS_HIP(hipMemcpy(gs_optim_class_raw(t->optim, "pos"), pos4.data(), pos4.size() * 4, hipMemcpyHostToDevice));
        GS_HIP(hipMemcpy(gs_optim_class_raw(t->optim, "scales"), sc.data(), sc.size() * 4, hipMemcpyHostToDevice));
        GS_HIP(hipMemcpy(gs_optim_class_raw(t->optim, "quat"), q.data(), q.size() * 4, hipMemcpyHostToDevice));
        GS_HIP(hipMemcpy(gs_optim_class_raw(t->optim, "opacity"), op.data(), op.size() * 4, hipMemcpyHostToDevice));
        GS_HIP(hipMemcpy(gs_optim_class_raw(t->optim, "sh"), sh48.data(), sh48.size() * 4, hipMemcpyHostToDevice));
    }
    // Activated buffers are the activation of raw (the "re-run activation"
    // contract of the reference optimizer).
    GS_CHECK(gs_optim_encode_activate(t->optim, ctx->stream));
    GS_HIP(hipStreamSynchronize(ctx->stream));
    *out = t.release();
    return GS_OK;
}

void gs_trainer_destroy(gs_trainer t) {
    if (!t) return;
    (void)hipStreamSynchronize(t->ctx->stream);
    delete t;
}

gs_status gs_trainer_step(gs_trainer t, void* stream, const uint32_t* view_ids, uint32_t k,
                          uint32_t mode) {
    GS_REQUIRE(t && view_ids, "null argument");
    GS_REQUIRE(k >= 1 && k <= t->cfg.max_views_per_step, "view count out of range");
    GS_REQUIRE(mode <= GS_STEP_GRAPH, "unknown step mode");
    for (uint32_t j = 0; j < k; ++j) GS_REQUIRE(view_ids[j] < t->views.size(), "view id out of range");
    hipStream_t st = as_stream(stream, t->ctx->stream);
    const bool batched = mode != GS_STEP_PLAIN;
    std::vector<CamParams> cams;
    for (uint32_t j = 0; j < k; ++j) {
        const TrainViewDev& tv = t->views[view_ids[j]];
        gs_raster_camera c = tv.camera;
        c.sh_degree = t->sh_degree;
        cams.push_back(make_cam(c, tv.width, tv.height, t->core.shape.tile));
        cams.back().target_base = (uint32_t)tv.target_offset;
        cams.back().n_surfels = tv.num_surfels;
    }
    t->last_ids.assign(view_ids, view_ids + k);
    t->last_batched = batched;
    // Learning-rate changes reach the device state before the step (and
    // before a capture: the graph reads them from device memory).
    GS_CHECK(t->optim->sync_hyper(st));
    if (mode == GS_STEP_GRAPH) {
        // The graph bakes grid sizes, so it is valid for one sequence of view
        // sizes; the camera table is refreshed in device memory before each
        // replay (set_views copies it on the stream ahead of the launch).
        std::vector<uint32_t> shape;
        for (uint32_t j = 0; j < k; ++j) {
            shape.push_back(t->views[view_ids[j]].width);
            shape.push_back(t->views[view_ids[j]].height);
        }
        GS_CHECK(t->core.set_views(st, cams, true));
        if (!t->graph_exec || shape != t->graph_shape) {
            if (t->graph_exec) {
                (void)hipGraphExecDestroy(t->graph_exec);
                t->graph_exec = nullptr;
            }
            if (t->graph) {
                (void)hipGraphDestroy(t->graph);
                t->graph = nullptr;
            }
            GS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            gs_status s = t->enqueue_step(st, view_ids, k, true, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (s != GS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            t->graph = g;
            GS_HIP(hipGraphInstantiate(&t->graph_exec, g, nullptr, nullptr, 0));
            t->graph_shape = shape;
        }
        GS_HIP(hipGraphLaunch(t->graph_exec, st));
        GS_HIP(hipEventRecord(t->step_done, st));
        t->step_pending = true;
        return GS_OK;
    }
    GS_CHECK(t->core.set_views(st, cams, batched));
    GS_CHECK(t->enqueue_step(st, view_ids, k, batched, t->prof.enabled));
    GS_HIP(hipEventRecord(t->step_done, st));
    t->step_pending = true;
    // Fold finished events so the event pool stays bounded.
    if (t->prof.enabled) GS_CHECK(t->prof.collect());
    return GS_OK;
}

static void drop_graph(gs_trainer t) {
    // A captured graph holds the previous schedule's launches: recapture.
    if (t->graph_exec) {
        (void)hipGraphExecDestroy(t->graph_exec);
        t->graph_exec = nullptr;
    }
    if (t->graph) {
        (void)hipGraphDestroy(t->graph);
        t->graph = nullptr;
    }
    t->graph_shape.clear();
}

gs_status gs_trainer_set_schedule(gs_trainer t, uint32_t digit_bits, uint32_t blocks_per_cu) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    GS_REQUIRE(digit_bits == 4 || digit_bits == 8, "digit_bits must be 4 or 8");
    GS_REQUIRE(blocks_per_cu <= 64, "blocks_per_cu must be <= 64");
    const uint32_t cus = (uint32_t)std::max(1, t->ctx->props.multiProcessorCount);
    t->core.digit_bits = digit_bits;
    t->core.tile_queue = blocks_per_cu != 0;
    t->core.queue_blocks = std::max(1u, blocks_per_cu) * cus;
    drop_graph(t);
    return GS_OK;
}

gs_status gs_trainer_set_overlap(gs_trainer t, int enable) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    t->pipeline = enable != 0;
    drop_graph(t);
    return GS_OK;
}

gs_status gs_trainer_set_keep_grads(gs_trainer t, int enable) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    t->keep_grads = enable != 0;
    drop_graph(t);
    return GS_OK;
}

static bool tiling_supported(uint32_t tile, uint32_t fb, uint32_t bb) {
    const bool f = (tile == 8 && (fb == 64 || fb == 128)) || (tile == 16 && (fb == 128 || fb == 256)) ||
                   (tile == 32 && (fb == 256 || fb == 512));
    const bool b = (tile == 8 && (bb == 64 || bb == 128)) || (tile == 16 && (bb == 64 || bb == 128)) ||
                   (tile == 32 && (bb == 128 || bb == 256));
    return f && b;
}

gs_status gs_trainer_set_tiling(gs_trainer t, uint32_t tile_size, uint32_t fwd_batch, uint32_t bwd_batch) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    GS_REQUIRE(tiling_supported(tile_size, fwd_batch, bwd_batch),
               "unsupported (tile_size, fwd_batch, bwd_batch); see gs_hip.h");
    GS_REQUIRE(tile_size == (uint32_t)kTile || t->core.binning == GS_BINNING_TIGHT,
               "work tiles other than 16 px need GS_BINNING_TIGHT");
    t->core.shape = gs::RasterShape{tile_size, fwd_batch, bwd_batch};
    drop_graph(t);
    return GS_OK;
}

gs_status gs_trainer_autotune(gs_trainer t, const uint32_t* view_ids, uint32_t n_ids, uint32_t reps,
                              uint32_t* chosen) {
    GS_REQUIRE(t && view_ids, "null argument");
    GS_REQUIRE(reps >= 1 && reps <= 16, "reps must be in 1..16");
    // Candidate schedules; none changes the step's result (the sorts are
    // stable, exact culling makes the per-pixel compositing sequence the same
    // for every work-tile size, and scheduling only reorders float atomics).
    struct Cand {
        uint32_t digit_bits, blocks_per_cu;  // blocks_per_cu 0: one workgroup per tile
        uint32_t tile, fwd_batch, bwd_batch;
    };
    std::vector<Cand> cands;
    const bool tight = t->core.binning == GS_BINNING_TIGHT;
    static const uint32_t tilings[][3] = {{16, 256, 64}, {16, 256, 128}, {16, 128, 64}, {32, 256, 128},
                                          {32, 512, 256}, {8, 128, 128}, {8, 64, 64}};
    static const uint32_t queues[] = {8, 4, 0};
    // Stage 1: work tiling x raster scheduling at 8-bit digits; stage 2: the
    // digit width at the stage-1 winner.
    for (const auto& tl : tilings) {
        if (tl[0] != (uint32_t)kTile && !tight) continue;
        for (uint32_t q : queues) cands.push_back(Cand{8, q, tl[0], tl[1], tl[2]});
    }
    const gs::RasterShape shape0 = t->core.shape;
    const uint32_t digit0 = t->core.digit_bits, queue0 = t->core.tile_queue ? t->core.queue_blocks : 0;
    // The candidates run real steps; the training state (raw parameters,
    // Adam moments and step count) is saved first and restored afterwards,
    // so tuning does not advance training.
    GS_CHECK(t->save_state());
    auto apply = [&](const Cand& c) {
        (void)gs_trainer_set_schedule(t, c.digit_bits, c.blocks_per_cu);
        (void)gs_trainer_set_tiling(t, c.tile, c.fwd_batch, c.bwd_batch);
    };
    hipStream_t st = t->ctx->stream;
    constexpr uint32_t kRounds = 2;
    std::vector<hipEvent_t> ev(reps + 1, nullptr);
    for (auto& e : ev) GS_HIP(hipEventCreate(&e));
    gs_status s = GS_OK;
    // Candidates of a stage interleaved over two rounds (drift hits all
    // alike); each step timed on its own; a candidate's score is its fastest
    // step. A candidate whose binning overflows the entry capacity is out.
    // Scores within kTie of the fastest are ties, won by the earliest
    // candidate in the list (the default schedule first): the steps timed
    // here are the first of a run, and a schedule that wins them by less
    // (such as no tile queue at C3) can lose over a run whose per-pixel
    // walks deepen as the scene trains.
    auto run_stage = [&](const std::vector<Cand>& cs, size_t* best) {
        std::vector<float> best_ms(cs.size(), 3.4e38f);
        std::vector<bool> overflowed(cs.size(), false);
        for (uint32_t round = 0; round < kRounds && s == GS_OK; ++round) {
            for (size_t i = 0; i < cs.size() && s == GS_OK; ++i) {
                apply(cs[i]);
                // One settling step (first use of the schedule), then `reps` timed.
                s = gs_trainer_step(t, st, view_ids, n_ids, GS_STEP_BATCHED);
                if (s == GS_OK && hipEventRecord(ev[0], st) != hipSuccess) s = fail(GS_ERR_HIP, "hipEventRecord");
                for (uint32_t r = 0; r < reps && s == GS_OK; ++r) {
                    s = gs_trainer_step(t, st, view_ids, n_ids, GS_STEP_BATCHED);
                    if (s == GS_OK && hipEventRecord(ev[r + 1], st) != hipSuccess)
                        s = fail(GS_ERR_HIP, "hipEventRecord");
                }
                if (s != GS_OK) break;
                if (hipEventSynchronize(ev[reps]) != hipSuccess) {
                    s = fail(GS_ERR_HIP, "autotune timing");
                    break;
                }
                uint32_t of = 0;
                if (hipMemcpy(&of, t->core.overflow, 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    s = fail(GS_ERR_HIP, "autotune overflow read");
                    break;
                }
                if (of) {
                    overflowed[i] = true;
                    const uint32_t zero = 0;
                    if (hipMemcpy(t->core.overflow, &zero, 4, hipMemcpyHostToDevice) != hipSuccess) {
                        s = fail(GS_ERR_HIP, "autotune overflow reset");
                        break;
                    }
                }
                for (uint32_t r = 0; r < reps; ++r) {
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, ev[r], ev[r + 1]);
                    best_ms[i] = std::min(best_ms[i], ms);
                }
            }
        }
        constexpr float kTie = 0.02f;
        float fastest = 3.4e38f;
        for (size_t i = 0; i < cs.size(); ++i)
            if (!overflowed[i]) fastest = std::min(fastest, best_ms[i]);
        size_t b = cs.size();
        for (size_t i = 0; i < cs.size() && b == cs.size(); ++i)
            if (!overflowed[i] && best_ms[i] <= fastest * (1.0f + kTie)) b = i;
        *best = b;
    };
    size_t b1 = 0, b2 = 0;
    run_stage(cands, &b1);
    std::vector<Cand> stage2;
    if (s == GS_OK && b1 < cands.size()) {
        for (uint32_t d : {8u, 4u}) {
            Cand c = cands[b1];
            c.digit_bits = d;
            stage2.push_back(c);
        }
        run_stage(stage2, &b2);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    const gs_status rs = t->restore_state();
    if (s == GS_OK) s = rs;
    if (s != GS_OK || b1 >= cands.size() || b2 >= stage2.size()) {
        // Failure, or every candidate overflowed: back to the schedule on entry.
        const uint32_t cus = (uint32_t)std::max(1, t->ctx->props.multiProcessorCount);
        (void)gs_trainer_set_schedule(t, digit0, queue0 / cus);
        (void)gs_trainer_set_tiling(t, shape0.tile, shape0.fwd_batch, shape0.bwd_batch);
        GS_CHECK(s);
        return fail(GS_ERR_OVERFLOW, "autotune: every candidate overflowed the entry capacity");
    }
    const Cand& w = stage2[b2];
    apply(w);
    t->tuned = true;
    if (chosen) {
        chosen[0] = w.digit_bits;
        chosen[1] = w.blocks_per_cu;
        chosen[2] = w.tile;
        chosen[3] = w.fwd_batch;
        chosen[4] = w.bwd_batch;
    }
    return GS_OK;
}

gs_status gs_trainer_optim(gs_trainer t, gs_optim* out) {
    GS_REQUIRE(t && out, "null argument");
    *out = t->optim;
    return GS_OK;
}

gs_status gs_trainer_state_bytes(gs_trainer t, uint64_t* bytes) {
    GS_REQUIRE(t && bytes, "null argument");
    return gs_optim_state_bytes(t->optim, bytes);
}

gs_status gs_trainer_export_state(gs_trainer t, void* dst, uint64_t bytes) {
    GS_REQUIRE(t && dst, "null argument");
    GS_CHECK(t->wait_idle());
    return gs_optim_export_state(t->optim, dst, bytes);
}

gs_status gs_trainer_import_state(gs_trainer t, const void* src, uint64_t bytes) {
    GS_REQUIRE(t && src, "null argument");
    GS_CHECK(t->wait_idle());
    return gs_optim_import_state(t->optim, src, bytes);
}

gs_status gs_trainer_set_mcmc_noise(gs_trainer t, float noise) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    GS_REQUIRE(std::isfinite(noise) && noise >= 0.0f, "mcmc noise must be finite and >= 0");
    if ((noise > 0.0f) != (t->mcmc_noise > 0.0f)) drop_graph(t);  // the step gains or loses a launch
    t->mcmc_noise = noise;
    return GS_OK;
}

gs_status gs_trainer_relocate_pairs(gs_trainer t, const uint32_t* pairs, uint32_t n_pairs) {
    GS_REQUIRE(t != nullptr && (pairs != nullptr || n_pairs == 0), "null argument");
    GS_REQUIRE(n_pairs <= t->n, "more pairs than surfels");
    // Pairs must not race: dead and alive indices in range and all distinct.
    std::vector<uint8_t> seen(t->n, 0);
    for (uint32_t i = 0; i < n_pairs; ++i) {
        const uint32_t d = pairs[4 * i], a = pairs[4 * i + 1], mode = pairs[4 * i + 2];
        GS_REQUIRE(d < t->n && a < t->n && mode <= 1u, "pair index or mode out of range");
        GS_REQUIRE(!seen[d] && !seen[a] && d != a, "pairs must use every index at most once");
        seen[d] = seen[a] = 1;
    }
    GS_CHECK(t->wait_idle());
    if (n_pairs == 0) return GS_OK;
    hipStream_t st = t->ctx->stream;
    if (t->reloc_pairs == nullptr) GS_CHECK(dalloc(&t->reloc_pairs, (size_t)t->n * 4, "relocation pairs"));
    GS_HIP(hipMemcpyAsync(t->reloc_pairs, pairs, (size_t)n_pairs * 16, hipMemcpyHostToDevice, st));
    RelocateArgs a;
    a.pairs = reinterpret_cast<const uint4*>(t->reloc_pairs);
    a.n_pairs = n_pairs;
    a.opac_act = t->opacities;
    auto& c = t->optim->classes;
    a.pos_raw = reinterpret_cast<float4*>(c[0].raw);
    a.scales_raw = reinterpret_cast<float2*>(c[1].raw);
    a.quat_raw = reinterpret_cast<float4*>(c[2].raw);
    a.opac_raw = c[3].raw;
    a.sh_raw = reinterpret_cast<float4*>(c[4].raw);
    for (int k = 0; k < 5; ++k) {
        a.m[k] = c[k].m;
        a.v[k] = c[k].v;
        a.class_width[k] = c[k].n / t->n;
    }
    GS_HIP(launch_relocate(st, a));
    GS_CHECK(gs_optim_encode_activate(t->optim, st));  // trainer.rs:1682
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

gs_status gs_trainer_mcmc_relocate(gs_trainer t, float dead_threshold, float split_ratio, uint64_t* rng_state,
                                   uint32_t* n_pairs) {
    GS_REQUIRE(t && rng_state && n_pairs, "null argument");
    GS_CHECK(t->wait_idle());
    std::vector<float> op(t->n), sc((size_t)t->n * 2);
    GS_HIP(hipMemcpy(op.data(), t->opacities, op.size() * 4, hipMemcpyDeviceToHost));
    GS_HIP(hipMemcpy(sc.data(), t->scales, sc.size() * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> pairs((size_t)t->n * 4);
    GS_CHECK(gs_mcmc_sample_pairs(op.data(), sc.data(), t->n, dead_threshold, split_ratio, rng_state, pairs.data(),
                                  n_pairs));
    return gs_trainer_relocate_pairs(t, pairs.data(), *n_pairs);
}

gs_status gs_trainer_write_ply(gs_trainer t, const char* path) {
    GS_REQUIRE(t && path, "null argument");
    GS_CHECK(t->wait_idle());
    return write_ply_from_device(t->ctx, path, t->n, t->positions, t->scales, t->quats, t->opacities, t->sh);
}

gs_status gs_trainer_last_losses(gs_trainer t, double* out, uint32_t n_ids) {
    GS_REQUIRE(t && out, "null argument");
    GS_REQUIRE(n_ids <= t->last_ids.size(), "more views requested than the last step had");
    GS_CHECK(t->order_after_step());
    for (uint32_t j = 0; j < n_ids; ++j) {
        const TrainViewDev& tv = t->views[t->last_ids[j]];
        const uint32_t np = l1_ssim_partials(tv.width, tv.height);
        GS_CHECK(read_loss_partials(t->ctx->stream, t->loss_partials + (size_t)j * t->partials_per_view * 2,
                                    np, tv.width, tv.height, t->cfg.dssim_scale, &out[2 * j],
                                    &out[2 * j + 1]));
    }
    return GS_OK;
}

gs_status gs_trainer_render(gs_trainer t, uint32_t view, float* rgba) {
    GS_REQUIRE(t && rgba, "null argument");
    GS_REQUIRE(view < t->views.size(), "view id out of range");
    const TrainViewDev& tv = t->views[view];
    hipStream_t st = t->ctx->stream;
    gs_raster_camera c = tv.camera;
    c.sh_degree = t->sh_degree;
    CamParams cp = make_cam(c, tv.width, tv.height, t->core.shape.tile);
    cp.n_surfels = tv.num_surfels;
    // The render overwrites the step's camera table, records and images:
    // it must not start before a step enqueued on another stream ends.
    GS_CHECK(t->order_after_step());
    GS_CHECK(t->core.set_views(st, {cp}, true));
    GS_CHECK(t->core.forward(st, 0, 1, t->n, t->positions, t->scales, t->quats, t->opacities, t->sh));
    GS_HIP(hipStreamSynchronize(st));
    GS_HIP(hipMemcpy(rgba, t->core.out_color, (size_t)tv.width * tv.height * 16, hipMemcpyDeviceToHost));
    return GS_OK;
}

gs_status gs_trainer_params(gs_trainer t, gs_raster_buffers* o) {
    GS_REQUIRE(t && o, "null argument");
    std::memset(o, 0, sizeof(*o));
    o->positions = reinterpret_cast<float*>(t->positions);
    o->scales = reinterpret_cast<float*>(t->scales);
    o->quats = reinterpret_cast<float*>(t->quats);
    o->opacities = t->opacities;
    o->sh = t->sh;
    o->out_color = reinterpret_cast<float*>(t->core.out_color);
    o->out_aux = reinterpret_cast<float*>(t->core.out_aux);
    o->out_normal = reinterpret_cast<float*>(t->core.out_normal);
    o->out_ncontrib = t->core.ncontrib;
    o->dl_dcolor = reinterpret_cast<float*>(t->dl_dcolor);
    o->grad_pos = reinterpret_cast<float*>(t->grad_pos);
    o->grad_scales = reinterpret_cast<float*>(t->grad_scales);
    o->grad_quat = reinterpret_cast<float*>(t->grad_quat);
    o->grad_opacity = t->grad_opacity;
    o->grad_sh = t->grad_sh;
    return GS_OK;
}

gs_status gs_trainer_entry_count(gs_trainer t, uint32_t* entries) {
    GS_REQUIRE(t && entries, "null argument");
    GS_CHECK(t->wait_idle());
    std::vector<uint32_t> tot(t->core.cap_views);
    uint32_t of = 0;
    GS_HIP(hipMemcpy(tot.data(), t->core.total, 4 * tot.size(), hipMemcpyDeviceToHost));
    GS_HIP(hipMemcpy(&of, t->core.overflow, 4, hipMemcpyDeviceToHost));
    // A batched step scans all its views at once (one total); a plain step
    // leaves one total per view.
    uint64_t s = 0;
    const size_t nt = t->last_batched ? 1 : std::max<size_t>(1, t->last_ids.size());
    for (size_t v = 0; v < nt && v < tot.size(); ++v) s += tot[v];
    *entries = (uint32_t)std::min<uint64_t>(s, 0xFFFFFFFFu);
    if (of) {
        const uint32_t zero = 0;
        GS_HIP(hipMemcpy(t->core.overflow, &zero, 4, hipMemcpyHostToDevice));
        return fail(GS_ERR_OVERFLOW, "tile entries exceeded the binning capacity");
    }
    return GS_OK;
}

const char* gs_stage_name(uint32_t stage) {
    static const char* names[GS_STAGE_COUNT] = {"project",    "scan",        "emit_keys", "sort",
                                                "tile_ranges", "raster_fwd", "l1_ssim",   "grad_clear",
                                                "raster_bwd", "geom_bwd",    "adam"};
    return stage < GS_STAGE_COUNT ? names[stage] : "";
}

gs_status gs_trainer_set_profiling(gs_trainer t, int enable) {
    GS_REQUIRE(t != nullptr, "trainer is null");
    GS_CHECK(t->prof.collect());
    t->prof.enabled = enable != 0;
    return GS_OK;
}

gs_status gs_trainer_stage_times(gs_trainer t, double* ms, uint32_t* steps) {
    GS_REQUIRE(t && ms, "null argument");
    GS_CHECK(t->prof.collect());
    for (uint32_t s = 0; s < GS_STAGE_COUNT; ++s) {
        ms[s] = t->prof.acc[s];
        t->prof.acc[s] = 0.0;
    }
    if (steps) *steps = t->prof.steps;
    t->prof.steps = 0;
    return GS_OK;
}

gs_status gs_trainer_digit_bits(gs_trainer t, uint32_t* bits) {
    GS_REQUIRE(t && bits, "null argument");
    *bits = t->core.digit_bits;
    return GS_OK;
}

}  // extern "C"
