// pt_noalloc.cpp — pt_trace_noalloc: the trace call without any device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the persistent megakernel (the wavefront backend's path pool is
// gigabytes) with every scratch buffer the kernel needs — work counters, the
// per-XCD tile counters, the seed list, the per-lane medium stacks and the
// traversal-stack spill columns — in one static device array, `scratch`, of
// kScratchBytes per GPU.  The HIP runtime places the array on each device the
// library runs on and keeps it until the process ends.  Its device address
// is fetched once per device by pt_create (noalloc_setup) and kept in
// g_scratch; host code uses that address, never the array's name (which is a
// host-side stand-in).  Every call on one GPU shares the array: calls that
// overlap on one GPU (from contexts on different threads or streams) race.
#include <cstring>
#include <mutex>

#include "pt_ctx.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr uint32_t kMaxPasses = 65536;              // seed-list capacity of one call
// Layout (byte offsets): work counters, 8 tile counters, the seed list, then
// the lanes' medium stacks and stack-spill columns (structure of arrays).
constexpr size_t kOffStats = 0;
constexpr size_t kOffTiles = 512;
constexpr size_t kOffSeeds = 1024;
constexpr size_t kOffLanes = kOffSeeds + 4 * size_t(kMaxPasses);
static_assert(pt::kStCount * 8 <= kOffTiles, "work counters fit");
static_assert(kOffTiles + 8 * 4 <= kOffSeeds, "tile counters fit");
static_assert(kOffLanes % 256 == 0, "lane columns stay 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

}  // namespace

namespace pt {

// Called by pt_create with the context's device current: the device address
// of this device's `scratch`, fetched once per device.
int noalloc_setup(pt_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return c->fail(PT_ERR_ARG, "device %d beyond %d", c->device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[c->device]) return PT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return c->hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    g_scratch[c->device] = p;
    return PT_OK;
}

}  // namespace pt

extern "C" int pt_trace_noalloc(pt_ctx* c, const pt_scene_params* p, uint32_t sw, uint32_t sh, uint32_t n_passes,
                                const uint32_t* seeds, pt_stats* stats) {
    using namespace pt;
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = check_trace_params(c, p);
    if (rc) return rc;
    if (n_passes == 0) {
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return PT_OK;
    }
    if (n_passes > kMaxPasses)
        return c->fail(PT_ERR_CAPACITY, "%u passes: at most %u per allocation-free call", n_passes, kMaxPasses);
    char* base = static_cast<char*>(scratch_of(c->device));
    if (!base) return c->fail(PT_ERR_STATE, "no scratch address for device %d", c->device);
    TraceArgs a{};
    a.P = *p;
    const uint32_t sx = p->section[0], sy = p->section[1];
    if (sx >= p->width || sy >= p->height) return c->fail(PT_ERR_ARG, "section origin outside the frame");
    a.sec_w = sw ? std::min(sw, p->width - sx) : p->width - sx;
    a.sec_h = sh ? std::min(sh, p->height - sy) : p->height - sy;
    a.n_passes = n_passes;
    a.accum = static_cast<float4*>(c->accum.p);
    a.out = static_cast<uchar4*>(c->out.p);
    a.stats = reinterpret_cast<unsigned long long*>(base + kOffStats);
    a.tile_ctr = reinterpret_cast<uint32_t*>(base + kOffTiles);
    a.schedule = c->schedule;
    a.wide = c->traversal;
    a.n_tiles = ((a.sec_w + 7u) / 8u) * ((a.sec_h + 7u) / 8u);
    // the same kernel variant pt_trace's megakernel path picks
    const bool cn = c->count_stats;
    const uint32_t total = (a.wide && !cn) ? c->wstack_cap : c->stack_cap;
    a.stack_lds = c->lds_stack_opt ? std::min<uint32_t>(total, c->lds_stack_opt) : total;
    int bpc = 0;
    hipError_t e = trace_blocks_per_cu(c->dev, a, cn, &bpc);
    if (e != hipSuccess) return c->hip_fail(e, "trace occupancy");
    if (bpc <= 0) return c->fail(PT_ERR_CAPACITY, "trace kernel cannot be resident (stack of %u entries)", total);
    // the persistent grid: resident workgroups, as many as the lane columns fit
    const size_t per_lane = size_t(kLaneMediaFloats) * 4u + size_t(total - a.stack_lds) * 4u;
    const uint64_t fit = (kScratchBytes - kOffLanes) / (per_lane * 64u);
    if (fit == 0)
        return c->fail(PT_ERR_CAPACITY, "a workgroup's scratch (%zu B) exceeds the %zu B array", per_lane * 64u,
                       kScratchBytes - kOffLanes);
    const uint32_t grid = uint32_t(std::min<uint64_t>({a.n_tiles, uint64_t(bpc) * uint64_t(c->n_cu), fit}));
    a.media_stride = grid * 64u;
    a.media = reinterpret_cast<float*>(base + kOffLanes);
    a.stack_spill = reinterpret_cast<uint32_t*>(base + kOffLanes + size_t(a.media_stride) * kLaneMediaFloats * 4u);
    if (seeds) {
        e = hipMemcpyAsync(base + kOffSeeds, seeds, size_t(n_passes) * 4u, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "seeds");
        a.seeds = reinterpret_cast<const uint32_t*>(base + kOffSeeds);
    }
    // counters start at zero: cleared by kernels (the last call left its values)
    e = launch_zero_u64(a.stats, kStCount, c->stream);
    if (e == hipSuccess) e = launch_zero_u64(reinterpret_cast<unsigned long long*>(a.tile_ctr), 4, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) e = launch_trace(c->dev, a, grid, cn, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "trace launch");
    if ((rc = c->sync("trace"))) return rc;
    if (!stats) return PT_OK;
    unsigned long long v[kStCount];
    e = hipMemcpy(v, base + kOffStats, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "stats readback");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    std::memset(stats, 0, sizeof(*stats));
    stats->extension_rays = v[kStExt];
    stats->shadow_rays = v[kStShadow];
    stats->paths = v[kStPaths];
    stats->node_pops = v[kStPops];
    stats->inner_pops = v[kStInner];
    stats->object_visits = v[kStObjs];
    stats->tri_tests = v[kStTris];
    stats->shade_events = v[kStShade];
    stats->kernel_ms = ms;
    stats->launches = 1;
    return PT_OK;
}
