// yn_internal_dev.h — kernel argument blocks and wave helpers shared by the
// .hip translation units and the host runtime.
#pragma once

#include "../../include/yune_hip.h"
#include "yn_device.h"
#include "yn_internal.h"
#include "yn_math.h"

namespace yune {

struct TraceArgs {
    const DevScene* scenes;     // scene table (device)
    const DevJob* jobs;         // job table (device)
    const int32_t* path_job;    // per path: job slot
    int32_t single_scene;       // >= 0: every ray uses this scene
    int32_t chunk;              // rays taken per wave per atomic
    RayQueue q;
    float4* hit_out;            // per path: t, tri bits, u, v
    int32_t* occ_out;           // per path (any-hit)
    uint32_t* work;             // chunk counters, one per queue slice, zero at launch
    int32_t n_slices;           // queue slices (one per XCD, 1..kMaxXcd)
    int32_t refill;             // > 0: lanes take new rays once this many of a wave are idle
    uint32_t* zero0;            // counters this launch clears (next queues)
    uint32_t* zero1;
    DevCounters* ctr;
    uint32_t* ovf_ref;          // traversal stack overflow area: ovf_depth entries
    uint16_t* ovf_t;            //   per grid thread, strided by the grid size
    int32_t ovf_depth;
    int32_t id_shift;           // ray id >> id_shift = path index (result slot = ray id)
    int32_t by_queue;           // 1: a ray's w is its job-table entry and its result goes
                                //    to its queue position (unidirectional paths)
};

// Segment of primary rays: a rectangle of one job's image at one frame.
// Its work items are the rectangle's pixels in tile order (seg_pixel) from
// item `first` on (0 unless a frame's pixels are issued in several parts).
struct GenSeg {
    int32_t job, slot;
    uint32_t rand;              // frame seed (kernel arg 10)
    int32_t x0, y0, w, h;
    int32_t path_base;          // first path index of the segment
    int32_t first;              // item of the rectangle its first path renders
};

// Item q of a w x h rectangle -> pixel (x, y) relative to its corner: bands
// of 8 rows, each cut into tiles 8 pixels wide (the last band and the last
// tile of a band may be smaller), row order inside a tile.  A wave's 64
// camera rays then cover an 8 x 8 tile instead of 64 pixels of one row, and
// fetch the same nodes more often.  A bijection on [0, w h): images do not
// depend on it, since every sample's seed and slot follow from its pixel.
__device__ __forceinline__ void seg_pixel(int q, int w, int h, int& x, int& y)
{
    const int band = q / (8 * w);
    const int l = q - band * 8 * w;
    const int bh = min(8, h - 8 * band);
    const int tc = l / (8 * bh);
    const int li = l - tc * 8 * bh;
    const int tw = min(8, w - 8 * tc);
    x = 8 * tc + li % tw;
    y = 8 * band + li / tw;
}

// Threads per block of the path-tracing shade compiled for `waves` waves
// per SIMD: two blocks fill a CU (4 SIMDs).  Its queue appends are one
// atomic per block; at 4 waves per SIMD, blocks of 512 beat 256 by 1-3%
// and 1024 by 1%.
constexpr int shade_block_pt(int waves) { return waves * 128; }

struct ShadeArgs {
    const DevScene* scenes;
    const DevJob* jobs;
    const float4* hit;          // closest-hit results, per extension-queue entry
    const int32_t* occ;         // any-hit results, per shadow-queue entry
    PathRecs in, out;           // records beside the input and the output queues
    RayQueue ext_in, sh_in;     // results to consume
    RayQueue ext_out, sh_out;   // rays to produce
    uint32_t* zero;             // work counters to clear for the next traces
    int32_t n_zero;
    DevCounters* ctr;
    uint32_t* done_ctr;         // continuous serving: paths ended per job-table entry from
    int32_t done_base;          //   done_base on (done_ctr[entry - done_base]), or null
    const float4* lmats;        // materials staged in LDS (5 float4 each, n_lmats <= kLdsMats):
    int32_t n_lmats;            //   lmode 1: the launch's one scene's, 2: every live scene's
    int32_t lmode;              //   (a scene's at its mat_base); 0: none.  Light quads likewise:
    int32_t lscene;             //   lmode 1: scene lscene's, 2: those of scene slots [0, n_lights)
    int32_t n_lights;
    int32_t ext_primary;        // 1: every extension-queue entry is a camera ray (a batch's
                                //    first iteration), whose record holds only m0, m1, rd
};

// Materials and light quads (scene slots) the shade kernels stage in LDS
// (80 B each).
constexpr int kLdsMats = 256;
constexpr int kLdsLights = 64;

// Shade pass of the bidirectional tracer: advances the paths of `list_in`
// (one entry per path whose rays came back) and lists those that issued
// new rays in `list_out`.
struct BdptShadeArgs {
    const DevScene* scenes;
    const DevJob* jobs;
    PathSoA P;
    BdptSoA B;
    const uint32_t* list_in;
    const uint32_t* list_in_count;
    uint32_t* list_out;
    uint32_t* list_out_count;
    RayQueue ext_out, sh_out;
    uint32_t* zero;             // work counters to clear for the next traces
    int32_t n_zero;
    const float4* lmats;        // materials staged in LDS (5 float4 each, n_lmats <= kLdsMats):
    int32_t n_lmats;            //   lmode 1: the launch's one scene's, 2: every live scene's
    int32_t lmode;              //   (a scene's at its mat_base); 0: none.  Light quads likewise:
    int32_t lscene;             //   lmode 1: scene lscene's, 2: those of scene slots [0, n_lights)
    int32_t n_lights;
};

struct GenArgs {
    const DevScene* scenes;
    const DevJob* jobs;
    PathSoA P;                  // bidirectional mode
    PathRecs rec;               // unidirectional mode: the records beside ext_out
    RayQueue ext_out;
    const GenSeg* segs;
    int32_t n_segs;
    int32_t n_paths;
    BdptSoA B;                  // bidirectional mode only
    uint32_t* list_out;         // bidirectional mode: every path starts listed
    uint32_t* list_count;
    // work items [item0, item0 + n_paths); with queue_base, item i is
    // appended behind the rays already in ext_out (at *queue_base + i, which
    // k_queue_reserve sets just before), else it takes ray position i and
    // the launch sets the queue's count
    int32_t item0;
    uint32_t* queue_base;
};

struct ResolveArgs {
    DevJob* jobs;
    int32_t job;
    int32_t x0, y0, w, h;
    int32_t n_slots;            // frames in this batch, oldest first
    int32_t reset_first;        // first frame of the batch overwrites (kernel arg 9)
    float4* accum;              // the job's RGBA32F image, .w = count
};

// Capture of the first frame for hit-ID parity (null pointers: off).
struct DevCapture {
    int32_t* prim_tri;
    int32_t* prim_light;
    float* prim_t;
    float4* shadow_o;           // xyz origin, w = len
    float4* shadow_d;           // xyz dir, w = 1 when recorded
    int32_t* shadow_occ;
};

#if defined(__HIPCC__)
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Wave-aggregated reservation of n_i consecutive slots per lane (lanes in
// order), one atomic per wave.  Every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint32_t n)
{
    const int lane = (int)__lane_id();
    uint32_t x = n;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    uint32_t total = __shfl(x, 63);
    if (total == 0) return 0;
    uint32_t base = 0;
    if (lane == 63) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + x - n;
}

// Block-aggregated reservation in NQ queues at once: thread t asks for n[q]
// consecutive slots of queue q and gets the first of them in first[q]
// (threads in order).  One atomic per queue and block call, made by threads
// 0..NQ-1 side by side: device-scope atomics on one counter serialise, and
// with one per wave the shade passes spent most of their time waiting for
// them.  Every thread of the block makes the same number of calls (there
// are two barriers inside); `par` alternates between calls, so a call's
// totals stay intact while a slower wave still reads the previous call's.
// A null counter may be passed for a queue whose counts are all zero.
template <int NQ, int BLOCK>
struct BlockReserveLds {
    uint32_t tot[2][BLOCK / 64][NQ];
    uint32_t base[2][NQ];
};

template <int NQ, int BLOCK>
__device__ __forceinline__ void block_reserve(BlockReserveLds<NQ, BLOCK>& sm, int par, uint32_t* const (&ctr)[NQ],
                                              const uint32_t (&n)[NQ], uint32_t (&first)[NQ])
{
    const int lane = (int)__lane_id();
    const int w = (int)(threadIdx.x >> 6);
    uint32_t x[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        x[q] = n[q];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x[q], o);
            if (lane >= o) x[q] += y;
        }
        if (lane == 63) sm.tot[par][w][q] = x[q];
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)NQ) {
        const int q = (int)threadIdx.x;
        uint32_t t = 0;
        for (int v = 0; v < BLOCK / 64; ++v) t += sm.tot[par][v][q];
        uint32_t* c = ctr[0];
#pragma unroll
        for (int k = 1; k < NQ; ++k)
            if (q == k) c = ctr[k];
        sm.base[par][q] = t ? atomicAdd(c, t) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t b = sm.base[par][q];
        for (int v = 0; v < w; ++v) b += sm.tot[par][v][q];
        first[q] = b + x[q] - n[q];
    }
}

// Wave-aggregated append: one atomic per wave, lanes get consecutive slots
// in lane order.  Every active lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return 0;
    const int lane = (int)__lane_id();
    uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + prefix;
}
#endif

// yn_trace_noalloc (trace_noalloc.hip): rays per call at most, threads of
// its trace launches, and traversal stack entries its overflow area holds
// beyond the LDS part.
constexpr int kTraceNoAllocMaxRays = YN_TRACE_NOALLOC_MAX_RAYS;
constexpr int kTraceNoAllocGridThreads = 1 << 16;
constexpr int kTraceNoAllocOvfDepth = 32;
int trace_noalloc_init(int device, void** ptr);
int trace_noalloc_enqueue(void* base, const TraceArgs& proto, const DevScene* scene, int width, int lds, int block,
                          const float* rays, int n, int32_t* tri, int32_t* light, float* t, hipStream_t s);
DevCounters* trace_noalloc_counters(void* base);

int launch_trace(const TraceArgs& a, bool any_hit, int width, int stack, int block, int grid, bool count, int lds,
                 hipStream_t stream);
int trace_lds_stack(int lds);   // traversal stack entries kept in LDS for a configured depth
int trace_resident_waves(int lds);   // waves per SIMD of a trace launch (the default grid fills them)
void launch_generate(const GenArgs& a, const DevCapture& cap, hipStream_t stream);
void launch_shade_pt(const ShadeArgs& a, const DevCapture& cap, int grid, int waves, hipStream_t stream);
void launch_generate_bdpt(const GenArgs& a, const DevCapture& cap, hipStream_t stream);
void launch_shade_bdpt(const BdptShadeArgs& a, const DevCapture& cap, int grid, hipStream_t stream);
void launch_resolve(const ResolveArgs& a, hipStream_t stream);
void launch_fmath(int op, const float* x, const float* y, int n, float* out, hipStream_t stream);
void launch_tonemap(const float4* in, float4* out, int n, int op, hipStream_t stream);
void launch_prepare(const DevScene* scene, const float* rays, int n, RayQueue q, int32_t* light_out,
                    hipStream_t stream);

}  // namespace yune
