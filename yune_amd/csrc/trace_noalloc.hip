// trace_noalloc.hip — traceRay on caller rays without allocating device
// memory: the variant of yn_trace for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops).
//
// Everything the call needs beyond the caller's inputs and outputs (ray
// queues, hit records, counters, the traversal stack's overflow area) lives
// in `scratch`, one static device array per GPU.  Its device address is
// taken once per device with hipGetSymbolAddress at context creation; host
// code only ever uses that address, never the array's name (a host-side
// stand-in).  Every call on one GPU shares the array: calls that overlap on
// one GPU race, and the array keeps the last call's contents (each call
// clears the counters it relies on with a kernel, not a memset).
#include <mutex>

#include "yn_internal_dev.h"
#include "yn_shade.h"

namespace {

using namespace yune;

constexpr int kNaMaxRays = kTraceNoAllocMaxRays;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets inside the scratch array.
constexpr size_t kOffCtr = 0;
constexpr size_t kOffQa = align256(kOffCtr + sizeof(DevCounters));      // closest-hit queue: o_t, d_p
constexpr size_t kOffQb = kOffQa + (size_t)kNaMaxRays * 32;             // any-hit queue: o_t, d_p
constexpr size_t kOffHit = kOffQb + (size_t)kNaMaxRays * 32;            // per ray: t, tri bits, u, v
constexpr size_t kOffOcc = kOffHit + (size_t)kNaMaxRays * 16;           // per ray: any-hit flag
constexpr size_t kOffLight = kOffOcc + (size_t)kNaMaxRays * 4;          // per ray: quad hit (0) or -1
constexpr size_t kOvfEntries = (size_t)kTraceNoAllocGridThreads * kTraceNoAllocOvfDepth;
constexpr size_t kOffOvfRef = align256(kOffLight + (size_t)kNaMaxRays * 4);
constexpr size_t kOffOvfT = align256(kOffOvfRef + kOvfEntries * 4);
constexpr size_t kScratchBytes = align256(kOffOvfT + kOvfEntries * 2);     // 35,663,104 B

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace yune {

namespace {

// Clears the counters (queue counts, trace work counters, flags, stats).
__global__ __launch_bounds__(256) void k_na_clear(DevCounters* ctr)
{
    uint32_t* w = reinterpret_cast<uint32_t*>(ctr);
    for (size_t i = threadIdx.x; i < sizeof(DevCounters) / 4; i += blockDim.x) w[i] = 0u;
}

// The quad test of traceRay (udpt.cl:244-276) and the split into the
// closest-hit and any-hit queues; each ray keeps its number as its result
// slot, so queue order does not matter.
__global__ __launch_bounds__(256) void k_na_prepare(const DevScene* scene, const float* rays, int n, RayQueue qa,
                                                    RayQueue qb, int32_t* light_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const float* r = rays + 8 * (size_t)(live ? i : 0);
    const bool any = live && r[7] != 0.0f;
    const uint32_t ka = wave_append(qa.count, live && !any);
    const uint32_t kb = wave_append(qb.count, any);
    if (!live) return;
    f4 o = mk(r[0], r[1], r[2], 1.0f), d = mk(r[4], r[5], r[6], 0.0f);
    float len = r[3], t;
    int light = -1;
    if (quad_hit(scene->light, o, d, len, t)) { len = t; light = 0; }
    light_out[i] = light;
    const RayQueue& q = any ? qb : qa;
    const uint32_t k = any ? kb : ka;
    q.o_t[k] = make_float4(o.x, o.y, o.z, len);
    q.d_p[k] = make_float4(d.x, d.y, d.z, __int_as_float(i));
}

// Results into the caller's arrays, as yn_trace reports them: a triangle
// nearer than the quad replaces the light hit (udpt.cl:386).
__global__ __launch_bounds__(256) void k_na_finish(int n, const float4* hits, const int32_t* light_in, int32_t* tri,
                                                   int32_t* light, float* t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 h = hits[i];
    const int tr = __float_as_int(h.y);
    tri[i] = tr;
    t[i] = h.x;
    light[i] = tr >= 0 ? -1 : light_in[i];
}

std::mutex g_na_mu;
void* g_na_ptr[64] = {};

}  // namespace

int trace_noalloc_init(int device, void** ptr)
{
    if (device < 0 || device >= 64) return -1;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        // the device is current (the caller set it); each GPU has its own copy
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        g_na_ptr[device] = p;
    }
    *ptr = g_na_ptr[device];
    return 0;
}

int trace_noalloc_enqueue(void* base, const TraceArgs& proto, const DevScene* scene, int width, int lds, int block,
                          const float* rays, int n, int32_t* tri, int32_t* light, float* t, hipStream_t s)
{
    if (!base || n < 0 || n > kNaMaxRays) return -1;
    if (n == 0) return 0;
    uint8_t* b = static_cast<uint8_t*>(base);
    DevCounters* ctr = reinterpret_cast<DevCounters*>(b + kOffCtr);
    RayQueue qa{reinterpret_cast<float4*>(b + kOffQa), reinterpret_cast<float4*>(b + kOffQa) + kNaMaxRays,
                &ctr->q[0][0]};
    RayQueue qb{reinterpret_cast<float4*>(b + kOffQb), reinterpret_cast<float4*>(b + kOffQb) + kNaMaxRays,
                &ctr->q[1][0]};
    float4* hits = reinterpret_cast<float4*>(b + kOffHit);
    int32_t* occ = reinterpret_cast<int32_t*>(b + kOffOcc);
    int32_t* lgt = reinterpret_cast<int32_t*>(b + kOffLight);
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(256), 0, s, ctr);
    hipLaunchKernelGGL(k_na_prepare, dim3((n + 255) / 256), dim3(256), 0, s, scene, rays, n, qa, qb, lgt);
    // one block per trace_block threads of the overflow area's grid
    const int grid = kTraceNoAllocGridThreads / block;
    for (int any = 0; any < 2; ++any) {
        TraceArgs ta = proto;
        ta.q = any ? qb : qa;
        ta.hit_out = hits;
        ta.occ_out = any ? occ : nullptr;
        ta.ctr = ctr;
        ta.work = &ctr->work[any][0][0];
        ta.zero0 = ta.zero1 = nullptr;
        ta.id_shift = 0;
        ta.ovf_ref = reinterpret_cast<uint32_t*>(b + kOffOvfRef);
        ta.ovf_t = reinterpret_cast<uint16_t*>(b + kOffOvfT);
        ta.ovf_depth = kTraceNoAllocOvfDepth;
        // the stack bound passed is what LDS and the overflow area hold; a
        // deeper traversal stops at it and raises the overflow flag
        if (launch_trace(ta, any != 0, width, trace_lds_stack(lds) + kTraceNoAllocOvfDepth, block, grid, false, lds, s))
            return -1;
    }
    hipLaunchKernelGGL(k_na_finish, dim3((n + 255) / 256), dim3(256), 0, s, n, hits, lgt, tri, light, t);
    return 0;
}

DevCounters* trace_noalloc_counters(void* base)
{
    return reinterpret_cast<DevCounters*>(static_cast<uint8_t*>(base) + kOffCtr);
}

}  // namespace yune
