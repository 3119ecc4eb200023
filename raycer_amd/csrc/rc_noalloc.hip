// Allocation-free closest-hit queries on device-resident rays, for callers
// that cannot allocate device memory while they run (code under HIP graph
// capture, tight serving loops).  No counterpart in the reference.
//
// The traversal kernels need scratch beyond the rays and hits: the list of
// rays whose stack outgrew the on-chip depth, the deep-stack kernel's stack
// and the counters.  Here that scratch is one file-scope device array,
// `scratch`, of a fixed size: each GPU the library runs on holds its own copy
// until the process ends.  Its device address is looked up once per device at
// rc_create (with the device current) and kept per device; host code passes
// that address to the kernels, never the array's name (which on the host is
// only a stand-in).  Every call on one GPU shares that GPU's array, so calls
// that overlap on one GPU race; a call clears what must start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rc_internal.hpp"

namespace {

constexpr size_t kNoAllocMaxRays = size_t(1) << 20;  // rays per call
constexpr size_t kCounterBytes = 256;                // DevCounters, padded
constexpr size_t kOvfBytes = kNoAllocMaxRays * sizeof(uint32_t);
constexpr size_t kDeepBytes = size_t(2048) * rcd::kDeepStack * sizeof(uint32_t);
constexpr size_t kScratchBytes = kCounterBytes + kOvfBytes + kDeepBytes;  // 6.25 MiB
static_assert(sizeof(DevCounters) <= kCounterBytes, "counters do not fit their slot");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Looks up this device's copy of the scratch array (the device is current).
int rc_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return rc_host_fail(RC_ERR_INVALID, "rc_create: device index too large");
    if (rc_deep_stack_bytes() > kDeepBytes)
        return rc_host_fail(RC_ERR_INVALID, "rc_create: the deep stack outgrew the allocation-free scratch");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return RC_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return rc_host_fail(RC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return RC_OK;
}

extern "C" int rc_trace_closest_device_noalloc(rc_ctx* c, const rc_scene* s, const rc_ray* d_rays, size_t n,
                                               rc_hit* d_hits) {
    RC_TRY(rc_check_scene(c, s, "rc_trace_closest_device_noalloc"));
    if (n > kNoAllocMaxRays)
        return rc_host_fail(RC_ERR_INVALID, "rc_trace_closest_device_noalloc: " + std::to_string(n) +
                                                " rays exceed the limit of " + std::to_string(kNoAllocMaxRays) +
                                                " per call (split the batch)");
    if (n && (!d_rays || !d_hits)) return rc_host_fail(RC_ERR_INVALID, "rc_trace_closest_device_noalloc: null array");
    if (n == 0) return RC_OK;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = g_scratch[c->device];
    }
    if (!base) return rc_host_fail(RC_ERR_INVALID, "rc_trace_closest_device_noalloc: no scratch on this device");
    uint8_t* p = static_cast<uint8_t*>(base);
    TraceLaunch L;
    L.scene = s;
    L.any = false;
    L.rays = d_rays;
    L.n_bound = uint32_t(n);
    L.hits = d_hits;
    L.ctr = reinterpret_cast<DevCounters*>(p);
    L.ovf_list = reinterpret_cast<uint32_t*>(p + kCounterBytes);
    L.deep_stack = reinterpret_cast<uint32_t*>(p + kCounterBytes + kOvfBytes);
    L.stream = c->stream;
    L.max_depth = c->stack_depth();
    L.max_blocks = uint32_t(c->cu_count) * c->tuning.blocks_per_cu;
    L.counting = false;
    L.dec64 = c->hit_fp64;
    rc_zero_counters(L.ctr, c->stream);  // a kernel: the previous call's counts are stale
    RC_TRY(rc_launch_trace_with(L));
    RC_TRY(rc_launch_deep_with(L));
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}
