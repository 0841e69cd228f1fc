// runtime.cpp — the C ABI (include/yune_hip.h): context, device scenes,
// progressive jobs and the wavefront frame loop.
//
// Stands in for CLManager (src/CLManager.cpp: context, buffers, errors) and
// RendererCore (src/RendererCore.cpp: per-frame seed, reset, tiling,
// ping-pong accumulation, profiling, save).  Instead of one OpenCL launch
// per tile per frame, a batch of frames x pixels becomes one set of paths
// that runs through the generate / extend / shadow / shade kernels until no
// ray is left, after which the frames are folded into the image in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>


#include "../../include/yune_hip.h"
#include "yn_internal_dev.h"

using namespace yune;

namespace {

thread_local std::string g_err;     // errors of calls without a context
}  // namespace

void yune::set_global_error(const std::string& msg) { g_err = msg; }

namespace {

struct SceneRec {
    bool live = false;
    DevScene dev{};
    void* bufs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    yn_scene_info info{};
    // the caller's arrays, kept so the GPU layout can be rebuilt (BVH width)
    std::vector<yn_triangle> h_tris;
    std::vector<yn_material> h_mats;
    std::vector<yn_bvh_node> h_nodes;
    yn_light_quad h_light{};
};

struct JobRec {
    bool live = false;
    int32_t scene = -1;
    yn_cam cam{};
    int32_t W = 0, H = 0;
    yn_render_params p{};
    std::mt19937 engine;
    uint64_t frame = 0;
    float4* accum = nullptr;
    yn_stats st{};
    int counting = 0;
    int tonemap_op = 0;           // 0: Reinhard (the shipped kernel), 1: Hable
    // pending capture (host pointers)
    int32_t* cap_prim_tri = nullptr;
    int32_t* cap_prim_light = nullptr;
    float* cap_prim_t = nullptr;
    float* cap_shadow = nullptr;
    int32_t* cap_shadow_occ = nullptr;
    // serving (yn_submit / yn_serve)
    bool served = false;          // submitted for the scheduler
    bool admitted = true;         // accumulation image allocated
    int64_t target = 0;           // frames requested
    int64_t seq = 0;              // submission order
    int64_t planned = 0;          // frames handed to the continuous loop (pieces)
    size_t bytes = 0;             // accumulation image bytes held
    double t_submit = 0.0, t_start = -1.0, t_done = -1.0;   // ms since context creation
};

constexpr int kMaxScenes = 64;
constexpr int kMaxStack = 256;      // traversal stack entries a scene may need
constexpr int kMaxJobs = 4096;
constexpr int kEvPool = 4096;
// Continuous serving: a piece is a run of consecutive frames of one job with
// its own job-table entry (kMaxJobs + piece index) and samples range.
constexpr int kMaxPieces = 4096;

struct Piece {
    bool live = false;
    bool done = false;            // every path of every item ended
    bool orphan = false;          // its job was released: finish, do not fold
    int32_t job = -1;
    int32_t scene = -1;
    int64_t f0 = 0;               // first frame of the job it renders
    int32_t k = 0;                // frames
    int64_t items = 0;            // k x W x H work items
    int64_t issued = 0;           // items turned into paths so far
    int64_t soff = 0;             // samples buffer offset (entries)
    std::vector<uint32_t> seeds;  // frame seeds, in frame order
};

// The persistent wavefront of continuous serving.  Jobs are cut into pieces
// as they are admitted; the pieces' work items enter the wavefront in order
// whenever enough paths have ended to refill their slots (at the boundary
// of a group of bounce iterations, not at the end of a batch), and a piece
// is folded into its job's image as soon as all its paths have ended and
// the job's earlier pieces are folded.
struct StreamState {
    bool active = false;
    int it = 0;                       // bounce iterations so far (queue parity)
    int64_t cap = 0;                  // paths in flight at most
    std::vector<Piece> pieces;
    std::deque<int> issue;            // pieces with items left to issue, in order
    std::map<int64_t, int64_t> sfree; // free ranges of the samples buffer: offset -> length
    int64_t live = 0;                 // paths in flight at the last sync
    unsigned long long last[6] = {0, 0, 0, 0, 0, 0};   // device statistics at the last sync
    std::vector<GenSeg> segs;         // the last refill's segments (kept until it ran)
};

}  // namespace

struct yn_ctx {
    int device = 0;
    int n_cu = 0;
    std::string name;
    size_t hbm = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;    // any-hit traces when they overlap the closest-hit ones
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    hipGraphExec_t gexec = nullptr;   // captured group of bounce iterations
    std::vector<uint8_t> gkey;        // the launch arguments it was captured with
    std::string err;
    yn_config cfg{};
    std::vector<SceneRec> scenes;
    std::vector<JobRec> jobs;
    DevScene* d_scenes = nullptr;
    yn_material* d_mtab = nullptr;    // every live scene's materials (when at most kLdsMats)
    int32_t n_mtab = 0;
    DevJob* d_jobs = nullptr;
    // wavefront buffers
    int64_t cap_paths = 0;
    int64_t sample_cap = 0;           // samples buffer entries (frames x pixels per batch)
    PathSoA P{};
    RayQueue E[2]{}, Sh[2]{};
    DevCounters* ctr = nullptr;
    DevCounters* h_ctr = nullptr;     // pinned mirror
    GenSeg* d_segs = nullptr;
    int seg_cap = 0;
    float4* samples = nullptr;
    std::vector<void*> allocs;
    hipEvent_t ev[kEvPool];
    int ev_used = 0;
    int pt_job_slot = -1;             // job slot used by yn_pathtracer
    size_t job_bytes = 0;
    // bidirectional mode: vertex arrays, wide shadow queues, path lists
    BdptSoA B{};
    RayQueue ShB[2]{};
    uint32_t* lst[2] = {nullptr, nullptr};
    int64_t cap_bdpt = 0;
    std::vector<void*> ballocs;
    // traversal stack overflow area (entries beyond the LDS part)
    uint32_t* ovf_ref = nullptr;
    uint16_t* ovf_t = nullptr;
    int ovf_depth = 0;
    int64_t ovf_threads = 0;
    // serving
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    int64_t submit_seq = 0;
    JobRec serve_rec;                 // statistics of scheduler rounds
    int64_t batch_paths = 0;          // paths per wavefront batch (0: the whole capacity)
    void* na_scratch = nullptr;       // device address of yn_trace_noalloc's static array on this device
    StreamState strm;                 // continuous serving
    uint32_t* done_ctr = nullptr;     // per piece: paths ended (device)
    uint32_t* h_done = nullptr;       // pinned mirror
};

namespace {

#define HIPCK(ctx, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
            return e_ == hipErrorOutOfMemory ? YN_ERR_NO_MEMORY : YN_ERR_HIP;                   \
        }                                                                                        \
    } while (0)

int fail(yn_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg; else g_err = msg;
    return code;
}

template <class T>
int dalloc(yn_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? YN_ERR_NO_MEMORY : YN_ERR_HIP;
    }
    *p = static_cast<T*>(q);
    return 0;
}

// Overflow area for traversal stacks deeper than the LDS part: `depth`
// entries per thread of a grid of `threads`.
int ensure_overflow(yn_ctx* c, int depth, int64_t threads)
{
    depth = std::max(0, depth);
    if (depth <= c->ovf_depth && threads <= c->ovf_threads && c->ovf_ref) return 0;
    depth = std::max(depth, c->ovf_depth);
    threads = std::max(threads, c->ovf_threads);
    if (c->ovf_ref) { hipFree(c->ovf_ref); hipFree(c->ovf_t); c->ovf_ref = nullptr; c->ovf_t = nullptr; }
    c->ovf_depth = 0;
    c->ovf_threads = 0;
    size_t n = (size_t)std::max(depth, 1) * (size_t)threads;
    int r = dalloc(c, &c->ovf_ref, n);
    if (r) return r;
    r = dalloc(c, &c->ovf_t, n);
    if (r) { hipFree(c->ovf_ref); c->ovf_ref = nullptr; return r; }
    c->ovf_depth = depth;
    c->ovf_threads = threads;
    return 0;
}

int alloc_wavefront(yn_ctx* c, int64_t n)
{
    auto A = [&](auto** p) -> int {
        int r = dalloc(c, p, (size_t)n);
        if (r == 0) c->allocs.push_back(*p);
        return r;
    };
    PathSoA& P = c->P;
    int r = 0;
    r |= A(&P.job); r |= A(&P.hit); r |= A(&P.occ);
    for (int k = 0; k < 2; ++k) {       // path records beside queues 0 and 1
        PathRecs& R = P.R[k];
        int rr = dalloc(c, &R.base, (size_t)PathRecs::alloc_entries(n));
        if (rr == 0) c->allocs.push_back(R.base);
        r |= rr;
        R.cap = n;
    }
    for (int k = 0; k < 2; ++k) {
        r |= A(&c->E[k].o_t); r |= A(&c->E[k].d_p);
        r |= A(&c->Sh[k].o_t); r |= A(&c->Sh[k].d_p);
    }
    {   // a batch may hold more work than paths in flight (regeneration), and
        // the continuous loop keeps a piece's samples until its last path
        // ends: four samples per path of capacity
        int rr = dalloc(c, &c->samples, (size_t)n * 4);
        if (rr == 0) c->allocs.push_back(c->samples);
        r |= rr;
    }
    if (r) return r < 0 ? r : YN_ERR_HIP;
    c->cap_paths = n;
    c->sample_cap = 4 * n;
    return 0;
}

constexpr size_t kBdptBudget = (size_t)16 << 30;  // bytes for bidirectional path state
constexpr int kMaxBdptBounces = 64;

void free_bdpt(yn_ctx* c)
{
    for (void* p : c->ballocs) hipFree(p);
    c->ballocs.clear();
    c->B = BdptSoA{};
    c->ShB[0] = c->ShB[1] = RayQueue{};
    c->lst[0] = c->lst[1] = nullptr;
    c->cap_bdpt = 0;
}

// Buffers of the bidirectional tracer for subpaths of up to `bounces`
// vertices: K vertices per side, 2^shift >= K shadow-ray slots per path and
// round, as many paths as fit the budget (and the unidirectional capacity,
// whose path state and extension queues are shared).
int ensure_bdpt(yn_ctx* c, int bounces)
{
    int K = std::max(bounces, 2);
    if (c->cap_bdpt > 0 && c->B.K >= K) return 0;
    free_bdpt(c);
    int shift = 0;
    while ((1 << shift) < K) ++shift;
    const int64_t R = (int64_t)1 << shift;
    // the per-path records (fixed fields and 2K vertices of 64 bytes), shadow
    // results and queues, path lists
    const int fields = BdptRecs::VTX + 8 * K;
    const int64_t per_path = (int64_t)fields * 16 + R * 4 + 2 * R * 32 + 2 * 4;
    int64_t cap = std::min<int64_t>(c->cap_paths, (int64_t)(kBdptBudget / (size_t)per_path));
    cap = std::min<int64_t>(cap, ((int64_t)1 << (31 - shift)) - 1);
    if (cap < 1) return fail(c, YN_ERR_NO_MEMORY, "no room for bidirectional paths");
    auto A = [&](auto** p, size_t n) -> int {
        int r = dalloc(c, p, n);
        if (r == 0) c->ballocs.push_back(*p);
        return r;
    };
    BdptSoA& B = c->B;
    const size_t n = (size_t)cap, ns = n << shift;
    int r = 0;
    r |= A(&B.R.base, (size_t)BdptRecs::alloc_entries(cap, fields));
    B.R.fields = fields;
    r |= A(&B.occ, ns);
    for (int k = 0; k < 2; ++k) {
        r |= A(&c->ShB[k].o_t, ns); r |= A(&c->ShB[k].d_p, ns);
        r |= A(&c->lst[k], n);
    }
    if (r) { free_bdpt(c); return r < 0 ? r : YN_ERR_HIP; }
    B.K = K;
    B.shift = shift;
    B.cap = (int32_t)cap;
    c->cap_bdpt = cap;
    return 0;
}

void free_all(yn_ctx* c)
{
    free_bdpt(c);
    for (void* p : c->allocs) hipFree(p);
    c->allocs.clear();
    for (auto& s : c->scenes)
        for (void*& b : s.bufs) { if (b) hipFree(b); b = nullptr; }
    for (auto& j : c->jobs) { if (j.accum) hipFree(j.accum); j.accum = nullptr; }
}

DevJob make_devjob(const JobRec& j)
{
    DevJob d{};
    d.cam = j.cam;
    d.p = j.p;
    d.W = j.W;
    d.H = j.H;
    d.scene = j.scene;
    d.samples = nullptr;
    return d;
}

bool valid_params(const yn_render_params* p)
{
    return p && (p->mode == YN_MODE_PT || p->mode == YN_MODE_BDPT) && p->max_bounces >= 0 && p->bdpt_bounces >= 1 &&
           p->bdpt_bounces <= kMaxBdptBounces;
}

bool valid_cam(const yn_cam* c)
{
    return c && c->r4.x == 0.0f && c->r4.y == 0.0f && c->r4.z == 0.0f && c->r4.w == 1.0f;
}

struct Region {
    int32_t job, slot;
    uint32_t rand;
    int32_t x0, y0, w, h;
};

struct Span { hipEvent_t a, b; int kind; };   // kind 0: closest-hit trace, 1: any-hit trace, 2: shade

void add_spans(const std::vector<Span>& spans, yn_stats& st)
{
    for (const Span& sp : spans) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, sp.a, sp.b);
        if (sp.kind == 0) st.ms_trace_ext += ms;
        else if (sp.kind == 1) st.ms_trace_shadow += ms;
        else st.ms_shade += ms;
    }
}

// Launch shapes and buffers shared by every bounce iteration of a batch.
struct LoopCfg {
    bool bdpt = false;
    int single = -1;                  // >= 0: every path renders this scene
    int lds = 16;                     // LDS stack entries of the trace launches (lds_for)
    int width = 2, max_depth = 0;
    int block = 256, grid = 1, shade_grid = 1;
    bool overlap = false, count = false;
    uint32_t* done_ctr = nullptr;     // continuous serving: ended paths per piece (entry - kMaxJobs)
};

// One bounce iteration's launch arguments.  Every argument follows from the
// iteration's parity (queues ping-pong), so a group of 4 iterations is the
// same sequence of launches each time and may be replayed from a graph.
struct IterArgs { TraceArgs te, ts; ShadeArgs sa; BdptShadeArgs ba; };

// LDS traversal-stack entries per lane of a trace launch whose largest
// scene has `node_bytes` of nodes: the configured count, else 12 (7 waves
// per SIMD, registers held to 72) for trees larger than an XCD's 4 MB L2,
// whose node fetches wait on the MALL or HBM and want more waves to hide
// them, and 16 (6 waves, 80 registers) for trees that stay in cache.
// (1M-triangle room, 60 MB of nodes: 12 entries 2.5% faster than 16, and
// launches over several such scenes 1.3% faster; the Cornell box and its
// bidirectional workload: within 1% either way.)
int lds_for(const yn_ctx* c, size_t node_bytes)
{
    if (c->cfg.lds_stack > 0) return c->cfg.lds_stack;
    return node_bytes > ((size_t)4 << 20) ? 12 : 16;
}

int loop_cfg(yn_ctx* c, bool bdpt, const std::vector<int>& scene_ids, bool count, LoopCfg& L)
{
    L = LoopCfg();
    L.bdpt = bdpt;
    L.count = count;
    bool same = !scene_ids.empty();
    for (int sc : scene_ids) same &= sc == scene_ids[0];
    if (same) L.single = scene_ids[0];
    size_t node_bytes = 0;
    for (int sc : scene_ids) {
        const SceneRec& sr = c->scenes[sc];
        L.max_depth = std::max(L.max_depth, sr.info.max_depth);
        L.width = sr.info.width;      // all scenes of a context share the width
        node_bytes = std::max(node_bytes, sr.info.node_bytes);
    }
    L.lds = lds_for(c, node_bytes);
    L.block = c->cfg.trace_block;
    // default: the trace blocks that fit at once (4 SIMDs per CU)
    L.grid = c->cfg.trace_grid > 0 ? c->cfg.trace_grid
                                   : c->n_cu * std::max(1, trace_resident_waves(L.lds) * 256 / L.block);
    // persistent shade grids: the blocks that fit at once (4 SIMDs per CU;
    // the bidirectional shade runs 3 waves per SIMD in blocks of 256)
    L.shade_grid = bdpt ? c->n_cu * 3 : c->n_cu * 2;
    L.overlap = c->cfg.trace_overlap != 0;
    // overlapping traces get disjoint halves of the stack overflow area
    if (int r = ensure_overflow(c, L.max_depth - trace_lds_stack(L.lds),
                                (int64_t)L.grid * L.block * (L.overlap ? 2 : 1)))
        return r;
    if (L.max_depth > trace_lds_stack(L.lds) + c->ovf_depth)
        return fail(c, YN_ERR_INVALID, "BVH too deep for the traversal stack");
    return 0;
}

// What the shade kernels stage in LDS: the materials and light quad of a
// single-scene launch's scene (mode 1), else the table of every live
// scene's materials, when it is small enough, and the light quads of the
// scene slots up to the last live one (mode 2); 0: nothing, each path reads
// its scene's arrays.
template <typename A>
void lds_materials(const yn_ctx* c, const LoopCfg& L, A& a)
{
    static_assert(kMaxScenes <= kLdsLights, "every scene slot's light quad fits the LDS table");
    a.lmats = nullptr;
    a.n_lmats = a.lmode = a.lscene = a.n_lights = 0;
    if (L.single >= 0 && c->scenes[L.single].h_mats.size() <= (size_t)kLdsMats) {
        a.lmats = (const float4*)c->scenes[L.single].bufs[4];
        a.n_lmats = (int32_t)c->scenes[L.single].h_mats.size();
        a.lmode = 1;
        a.lscene = L.single;
        a.n_lights = 1;
    } else if (c->d_mtab && c->n_mtab <= kLdsMats) {
        a.lmats = (const float4*)c->d_mtab;
        a.n_lmats = c->n_mtab;
        a.lmode = 2;
        for (int i = 0; i < kMaxScenes; ++i)
            if (c->scenes[i].live) a.n_lights = i + 1;
    }
}

void make_iter(yn_ctx* c, const LoopCfg& L, int cur, IterArgs& A)
{
    std::memset(&A, 0, sizeof A);
    const int nxt = cur ^ 1;
    auto q = [&](int k) { return &c->ctr->q[k][0]; };
    RayQueue* SQ = L.bdpt ? c->ShB : c->Sh;
    TraceArgs& ta = A.te;
    ta.scenes = c->d_scenes;
    ta.jobs = c->d_jobs;
    ta.path_job = c->P.job;
    ta.single_scene = L.single;
    ta.chunk = c->cfg.trace_chunk;
    ta.ctr = c->ctr;
    ta.ovf_ref = c->ovf_ref;
    ta.ovf_t = c->ovf_t;
    ta.ovf_depth = c->ovf_depth;
    ta.by_queue = L.bdpt ? 0 : 1;
    // extension rays
    ta.q = c->E[cur];
    ta.hit_out = c->P.hit;
    ta.work = &c->ctr->work[0][0][0];
    ta.n_slices = c->cfg.trace_xcd;
    ta.refill = c->cfg.trace_refill > 0 ? c->cfg.trace_refill : 0;
    ta.zero0 = c->E[nxt].count;
    ta.zero1 = SQ[nxt].count;
    // shadow rays
    TraceArgs& ts = A.ts;
    ts = ta;
    ts.q = SQ[cur];
    ts.hit_out = nullptr;
    ts.occ_out = L.bdpt ? c->B.occ : c->P.occ;
    ts.id_shift = L.bdpt ? c->B.shift : 0;
    ts.work = &c->ctr->work[1][0][0];
    ts.zero0 = L.bdpt ? q(8 + nxt) : nullptr;    // the next path list
    ts.zero1 = nullptr;
    if (L.overlap) {        // the second half of the overflow area
        ts.ovf_ref = c->ovf_ref + (size_t)c->ovf_depth * L.grid * L.block;
        ts.ovf_t = c->ovf_t + (size_t)c->ovf_depth * L.grid * L.block;
    }
    if (L.bdpt) {
        BdptShadeArgs& ba = A.ba;
        ba.scenes = c->d_scenes;
        ba.jobs = c->d_jobs;
        ba.P = c->P;
        ba.B = c->B;
        ba.list_in = c->lst[cur];
        ba.list_in_count = q(8 + cur);
        ba.list_out = c->lst[nxt];
        ba.list_out_count = q(8 + nxt);
        ba.ext_out = c->E[nxt];
        ba.sh_out = c->ShB[nxt];
        ba.zero = &c->ctr->work[0][0][0];
        ba.n_zero = 2 * kMaxXcd;
        lds_materials(c, L, ba);
    } else {
        ShadeArgs& sa = A.sa;
        sa.scenes = c->d_scenes;
        sa.jobs = c->d_jobs;
        sa.hit = c->P.hit;
        sa.occ = c->P.occ;
        sa.in = c->P.R[cur];
        sa.out = c->P.R[nxt];
        sa.ext_in = c->E[cur];
        sa.sh_in = c->Sh[cur];
        sa.ext_out = c->E[nxt];
        sa.sh_out = c->Sh[nxt];
        sa.zero = &c->ctr->work[0][0][0];
        sa.n_zero = 2 * kMaxXcd;
        sa.ctr = c->ctr;
        sa.done_ctr = L.done_ctr;
        lds_materials(c, L, sa);
        sa.done_base = kMaxJobs;      // continuous serving: pieces' entries start here
    }
}

// Enqueues one bounce iteration: closest-hit trace, any-hit trace (on the
// second stream when they overlap: it depends only on the previous shade,
// like the closest-hit one), shade.  With `spans`, each launch is bracketed
// by timing events.
int enqueue_iter(yn_ctx* c, const LoopCfg& L, const IterArgs& A, const DevCapture& cap, std::vector<Span>* spans)
{
    hipStream_t s = c->stream;
    hipStream_t s2 = L.overlap ? c->stream2 : s;
    auto ev = [&]() -> hipEvent_t { return c->ev[c->ev_used++ % kEvPool]; };
    if (L.overlap) {
        HIPCK(c, hipEventRecord(c->fork_ev, s));
        HIPCK(c, hipStreamWaitEvent(s2, c->fork_ev, 0));
    }
    hipEvent_t e0 = spans ? ev() : nullptr, e1 = spans ? ev() : nullptr;
    if (spans) hipEventRecord(e0, s);
    launch_trace(A.te, false, L.width, L.max_depth, L.block, L.grid, L.count, L.lds, s);
    if (spans) { hipEventRecord(e1, s); spans->push_back({e0, e1, 0}); }
    hipEvent_t e2 = spans ? ev() : nullptr, e3 = spans ? ev() : nullptr;
    if (spans) hipEventRecord(e2, s2);
    launch_trace(A.ts, true, L.width, L.max_depth, L.block, L.grid, L.count, L.lds, s2);
    if (spans) { hipEventRecord(e3, s2); spans->push_back({e2, e3, 1}); }
    if (L.overlap) {
        HIPCK(c, hipEventRecord(c->join_ev, s2));
        HIPCK(c, hipStreamWaitEvent(s, c->join_ev, 0));
    }
    hipEvent_t e4 = spans ? ev() : nullptr, e5 = spans ? ev() : nullptr;
    if (spans) hipEventRecord(e4, s);
    if (L.bdpt) launch_shade_bdpt(A.ba, cap, L.shade_grid, s);
    else launch_shade_pt(A.sa, cap, L.shade_grid, c->cfg.shade_waves, s);
    if (spans) { hipEventRecord(e5, s); spans->push_back({e4, e5, 2}); }
    return 0;
}

// The graph of one group of `group` bounce iterations (starting at queue
// parity 0), captured once and replayed while the launch arguments stay
// byte-identical (same buffers, scenes, launch shapes).  The iterations
// clear their counters with kernels, so the graph holds no memset node.
int ensure_group_graph(yn_ctx* c, const LoopCfg& L, const IterArgs (&iters)[2], int group)
{
    std::vector<uint8_t> key(sizeof iters);
    std::memcpy(key.data(), iters, sizeof iters);
    const int32_t shape[] = {L.bdpt, L.width, L.max_depth, L.block, L.grid, L.shade_grid, L.count,
                             L.lds, c->cfg.shade_waves, L.overlap, group};
    key.insert(key.end(), (const uint8_t*)shape, (const uint8_t*)shape + sizeof shape);
    if (c->gexec && key == c->gkey) return 0;
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    c->gkey.clear();
    hipStream_t s = c->stream;
    HIPCK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = 0;
    for (int g = 0; g < group && !r; ++g) r = enqueue_iter(c, L, iters[g & 1], DevCapture{}, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (!r && e == hipSuccess) e = hipGraphInstantiateWithFlags(&c->gexec, graph, 0);
    if (graph) hipGraphDestroy(graph);
    if (r) return r;
    if (e != hipSuccess) { c->gexec = nullptr; return fail(c, YN_ERR_HIP, std::string("bounce-loop graph: ") + hipGetErrorString(e)); }
    c->gkey = std::move(key);
    return 0;
}

// Runs one batch of primary-ray regions to completion; samples land in
// samples[slot][pixel] of each region's job (the sample buffer is shared:
// job slots are laid out by the caller so regions do not overlap).
int run_batch(yn_ctx* c, const std::vector<Region>& regs, JobRec& stat_job, const DevCapture& cap)
{
    std::vector<GenSeg> segs;
    int64_t total = 0;
    for (const Region& r : regs) {
        GenSeg g{r.job, r.slot, r.rand, r.x0, r.y0, r.w, r.h, (int32_t)total, 0};
        total += (int64_t)r.w * r.h;
        segs.push_back(g);
    }
    if (total == 0) return 0;
    const bool bdpt = c->jobs[regs[0].job].p.mode == YN_MODE_BDPT;
    // paths in flight: the whole batch when it fits, else the capacity, with
    // ended paths' slots regenerated from the remaining work items
    const int64_t C = bdpt ? c->cap_bdpt : (c->batch_paths > 0 ? std::min(c->batch_paths, c->cap_paths) : c->cap_paths);
    if (bdpt ? total > C : total > c->sample_cap) return fail(c, YN_ERR_INVALID, "batch larger than the path capacity");
    const bool regen = !bdpt && total > C;
    int64_t issued = std::min<int64_t>(total, C);
    if ((int)segs.size() > c->seg_cap) {
        if (c->d_segs) hipFree(c->d_segs);
        c->d_segs = nullptr;
        c->seg_cap = std::max<int>((int)segs.size(), 256);
        int r = dalloc(c, &c->d_segs, (size_t)c->seg_cap);
        if (r) return r;
    }
    hipStream_t s = c->stream;
    HIPCK(c, hipMemcpyAsync(c->d_segs, segs.data(), sizeof(GenSeg) * segs.size(), hipMemcpyHostToDevice, s));
    HIPCK(c, hipMemsetAsync(c->ctr, 0, sizeof(DevCounters), s));

    const DevScene* scenes = c->d_scenes;
    auto q = [&](int k) { return &c->ctr->q[k][0]; };
    // counters (q(k) = q[k][0]): q[0],q[1] extension queue counts, q[2],q[3] shadow queue
    // counts, q[7] stack overflow flag, q[8],q[9] path-list counts
    // (bidirectional mode), q[10] free slots (regeneration), q[11] where a
    // regeneration's rays start in their queue; work[0], work[1]
    // the chunk counters of the extend / shadow traces, per queue slice.
    for (int k = 0; k < 2; ++k) {
        c->E[k].count = q(k);
        c->Sh[k].count = q(2 + k);
        c->ShB[k].count = q(2 + k);
    }
    c->ev_used = 0;

    GenArgs ga{};
    ga.scenes = scenes;
    ga.jobs = c->d_jobs;
    ga.P = c->P;
    ga.rec = c->P.R[0];
    ga.ext_out = c->E[0];
    ga.segs = c->d_segs;
    ga.n_segs = (int32_t)segs.size();
    ga.n_paths = (int32_t)issued;
    if (bdpt) {
        ga.B = c->B;
        ga.list_out = c->lst[0];
        ga.list_count = q(8);
        launch_generate_bdpt(ga, cap, s);
    } else {
        launch_generate(ga, cap, s);
    }

    std::vector<int> scene_ids;
    for (const Region& r : regs) scene_ids.push_back(c->jobs[r.job].scene);
    LoopCfg L;
    if (int r = loop_cfg(c, bdpt, scene_ids, stat_job.counting != 0, L)) return r;
    std::vector<Span> spans;
    IterArgs iters[2];
    make_iter(c, L, 0, iters[0]);
    make_iter(c, L, 1, iters[1]);
    const int group = 4;                 // iterations between host checks (even)
    // Graph replay of a group (not with ray capture for tests)
    const bool any_cap = cap.prim_tri || cap.prim_light || cap.prim_t || cap.shadow_o || cap.shadow_d || cap.shadow_occ;
    const bool use_graph = c->cfg.graph != 0 && !any_cap;
    if (use_graph)
        if (int r = ensure_group_graph(c, L, iters, group)) return r;

    int it = 0;
    for (;;) {
        if (use_graph) {
            HIPCK(c, hipGraphLaunch(c->gexec, s));
            it += group;
        } else {
            for (int g = 0; g < group; ++g, ++it) {
                bool timed = (int)spans.size() + 3 < kEvPool / 2;
                if (it == 0 && !bdpt) {
                    // the first iteration shades camera rays only: their
                    // records hold m0, m1 and the direction
                    IterArgs first = iters[0];
                    first.sa.ext_primary = 1;
                    if (int r = enqueue_iter(c, L, first, cap, timed ? &spans : nullptr)) return r;
                    continue;
                }
                if (int r = enqueue_iter(c, L, iters[it & 1], cap, timed ? &spans : nullptr)) return r;
            }
        }
        HIPCK(c, hipGetLastError());
        HIPCK(c, hipMemcpyAsync(c->h_ctr, c->ctr, kCtrHostBytes, hipMemcpyDeviceToHost, s));
        HIPCK(c, hipStreamSynchronize(s));
        int nxt = it & 1;
        if (c->h_ctr->q[7][0]) return fail(c, YN_ERR_INVALID, "traversal stack overflow");
        if (it > 200000) return fail(c, YN_ERR_HIP, "wavefront loop did not drain");
        bool empty = bdpt ? c->h_ctr->q[8 + nxt][0] == 0 : (c->h_ctr->q[nxt][0] == 0 && c->h_ctr->q[2 + nxt][0] == 0);
        if (regen && issued < total) {
            // new paths behind the rays of the queue the next iteration
            // traces, once there is room for enough of them (or the
            // wavefront has drained): a path's record sits beside its ray,
            // so the room is the capacity less the rays queued
            const int64_t queued = (int64_t)c->h_ctr->q[nxt][0] + c->h_ctr->q[2 + nxt][0];
            const int64_t room = C - queued, need = total - issued;
            if (room > 0 && (room >= std::min<int64_t>(need, std::max<int64_t>(C / 8, 1)) || empty)) {
                GenArgs gr = ga;
                gr.item0 = (int32_t)issued;
                gr.n_paths = (int32_t)std::min<int64_t>(room, need);
                gr.queue_base = q(11);
                gr.ext_out = c->E[nxt];
                gr.rec = c->P.R[nxt];
                launch_generate(gr, cap, s);
                issued += gr.n_paths;
                continue;
            }
        }
        if (empty) break;
    }
    add_spans(spans, stat_job.st);
    stat_job.st.iterations += (uint64_t)it;
    stat_job.st.trace_launches_ext += (uint64_t)it;
    stat_job.st.trace_launches_shadow += (uint64_t)it;
    stat_job.st.rays_ext += c->h_ctr->rays_ext;
    stat_job.st.rays_shadow += c->h_ctr->rays_shadow;
    stat_job.st.nodes_ext += c->h_ctr->nodes_ext;
    stat_job.st.tris_ext += c->h_ctr->tris_ext;
    stat_job.st.nodes_shadow += c->h_ctr->nodes_shadow;
    stat_job.st.tris_shadow += c->h_ctr->tris_shadow;
    stat_job.st.samples += (uint64_t)total;
    return 0;
}

int upload_job_table(yn_ctx* c, int jid, float4* samples)
{
    DevJob d = make_devjob(c->jobs[jid]);
    d.samples = samples;
    HIPCK(c, hipMemcpyAsync(c->d_jobs + jid, &d, sizeof(DevJob), hipMemcpyHostToDevice, c->stream));
    return 0;
}

int stream_drain(yn_ctx* c);

// Render n frames of rectangle [x0,x0+w) x [y0,y0+h) into `accum`, frames
// taking seeds from `seeds`; first frame resets when reset_first.
int render_rect(yn_ctx* c, int jid, float4* accum, int x0, int y0, int w, int h, const std::vector<uint32_t>& seeds,
                bool reset_first, const DevCapture& cap)
{
    JobRec& J = c->jobs[jid];
    int64_t npix = (int64_t)w * h;
    if (npix == 0 || seeds.empty()) return 0;
    // the continuous loop shares the wavefront buffers: it finishes first
    if (c->strm.active)
        if (int r = stream_drain(c)) return r;
    int64_t full = (int64_t)J.W * J.H;
    int64_t pcap = c->cap_paths;
    if (J.p.mode == YN_MODE_BDPT) {
        if (int r = ensure_bdpt(c, J.p.bdpt_bounces)) return r;
        pcap = c->cap_bdpt;
    }
    if (c->batch_paths > 0) pcap = std::min(pcap, c->batch_paths);
    // with regeneration a batch holds more work than paths in flight
    if (J.p.mode == YN_MODE_PT) pcap = std::min<int64_t>(c->sample_cap, pcap * c->cfg.batch_factor);
    // Sample slots are indexed [slot][W*H]; a batch may hold K frames when
    // K full images fit in the path capacity.
    int64_t kcap = pcap / std::max<int64_t>(full, 1);
    size_t f = 0;
    while (f < seeds.size()) {
        std::vector<Region> regs;
        int nslots;
        if (kcap >= 1) {
            nslots = (int)std::min<int64_t>(kcap, (int64_t)(seeds.size() - f));
            for (int k = 0; k < nslots; ++k) regs.push_back({jid, k, seeds[f + k], x0, y0, w, h});
            int r = upload_job_table(c, jid, c->samples);
            if (r) return r;
            DevCapture cp = f == 0 ? cap : DevCapture{};
            r = run_batch(c, regs, J, cp);
            if (r) return r;
            ResolveArgs ra{c->d_jobs, jid, x0, y0, w, h, nslots, (reset_first && f == 0) ? 1 : 0, accum};
            launch_resolve(ra, c->stream);
        } else {
            // image larger than the path capacity: row bands, one frame at a time
            if ((int64_t)J.W > pcap) return fail(c, YN_ERR_INVALID, "image row wider than the path capacity");
            int rows = (int)std::max<int64_t>(1, pcap / J.W);
            nslots = 1;
            for (int yb = y0; yb < y0 + h; yb += rows) {
                int hh = std::min(rows, y0 + h - yb);
                regs.clear();
                regs.push_back({jid, 0, seeds[f], x0, yb, w, hh});
                // sample slot 0 is addressed as [pixel]; offset the buffer so the
                // band's pixels land inside the path-sized sample buffer
                float4* base = c->samples - (int64_t)yb * J.W;
                int r = upload_job_table(c, jid, base);
                if (r) return r;
                DevCapture cp = f == 0 ? cap : DevCapture{};
                r = run_batch(c, regs, J, cp);
                if (r) return r;
                ResolveArgs ra{c->d_jobs, jid, x0, yb, w, hh, 1, (reset_first && f == 0) ? 1 : 0, accum};
                launch_resolve(ra, c->stream);
            }
        }
        f += (size_t)nslots;
    }
    HIPCK(c, hipGetLastError());
    return 0;
}

double ms_since(const yn_ctx* c)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c->t0).count();
}

// Admission: served jobs get their accumulation image, oldest first, while
// the memory budget allows (a job that does not fit waits; younger jobs that
// fit may go ahead of it).
int admit_jobs(yn_ctx* c)
{
    std::vector<int> wait;
    for (int i = 0; i < (int)c->jobs.size(); ++i)
        if (c->jobs[i].live && c->jobs[i].served && !c->jobs[i].admitted) wait.push_back(i);
    std::sort(wait.begin(), wait.end(), [&](int a, int b) { return c->jobs[a].seq < c->jobs[b].seq; });
    for (int jid : wait) {
        JobRec& J = c->jobs[jid];
        size_t bytes = (size_t)J.W * J.H * sizeof(float4);
        if (c->job_bytes + bytes > c->cfg.mem_budget) continue;
        int r = dalloc(c, &J.accum, (size_t)J.W * J.H);
        if (r) return r;
        J.bytes = bytes;
        J.admitted = true;
        c->job_bytes += bytes;
    }
    return 0;
}

// One scheduler round: whole frames of admitted jobs, oldest job first, are
// packed into one wavefront batch (jobs of one mode per round; a job's
// frames stay in order), run, and folded into each job's image.  Returns the
// number of frames rendered (0: nothing to do).
int serve_round(yn_ctx* c, int* frames_out)
{
    *frames_out = 0;
    if (int r = admit_jobs(c)) return r;
    std::vector<int> order;
    int max_kb = 0;
    for (int i = 0; i < (int)c->jobs.size(); ++i) {
        const JobRec& J = c->jobs[i];
        if (J.live && J.served && J.admitted && (int64_t)J.frame < J.target) {
            order.push_back(i);
            if (J.p.mode == YN_MODE_BDPT) max_kb = std::max(max_kb, J.p.bdpt_bounces);
        }
    }
    if (order.empty()) return 0;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return c->jobs[a].seq < c->jobs[b].seq; });
    const int mode = c->jobs[order[0]].p.mode;
    int64_t pcap = c->cap_paths;
    if (mode == YN_MODE_BDPT) {
        if (int r = ensure_bdpt(c, max_kb)) return r;
        pcap = c->cap_bdpt;
    }
    if (c->batch_paths > 0) pcap = std::min(pcap, c->batch_paths);
    if (mode == YN_MODE_PT) pcap = std::min<int64_t>(c->sample_cap, pcap * c->cfg.batch_factor);
    // a job whose frame exceeds the batch runs alone, in row bands
    {
        JobRec& J = c->jobs[order[0]];
        if ((int64_t)J.W * J.H > pcap) {
            if (J.t_start < 0) J.t_start = ms_since(c);
            std::vector<uint32_t> seeds(1, (uint32_t)J.engine());
            int r = render_rect(c, order[0], J.accum, 0, 0, J.W, J.H, seeds, J.frame == 0, DevCapture{});
            if (r) return r;
            HIPCK(c, hipStreamSynchronize(c->stream));
            J.frame += 1;
            J.planned = (int64_t)J.frame;
            J.st.frames += 1;
            J.st.samples += (uint64_t)J.W * J.H;
            if ((int64_t)J.frame >= J.target) J.t_done = ms_since(c);
            *frames_out = 1;
            return 0;
        }
    }
    // Processor sharing: every waiting job of the round's mode first gets an
    // equal share of the path capacity (at least one frame, oldest first),
    // then what is left goes to the oldest jobs.  Small jobs finish within a
    // round or two; large ones keep the batch full.
    std::vector<int> cand;
    for (int jid : order) if (c->jobs[jid].p.mode == mode) cand.push_back(jid);
    std::vector<int64_t> kf(cand.size(), 0);
    int64_t used = 0;
    const int64_t share = pcap / (int64_t)cand.size();
    for (size_t i = 0; i < cand.size(); ++i) {
        const JobRec& J = c->jobs[cand[i]];
        int64_t px = (int64_t)J.W * J.H;
        int64_t want = std::min<int64_t>(J.target - (int64_t)J.frame, std::max<int64_t>(1, share / px));
        int64_t k = std::min<int64_t>(want, (pcap - used) / px);
        if (k > 0) { kf[i] = k; used += k * px; }
    }
    for (size_t i = 0; i < cand.size() && used < pcap; ++i) {
        const JobRec& J = c->jobs[cand[i]];
        int64_t px = (int64_t)J.W * J.H;
        int64_t k = std::min<int64_t>(J.target - (int64_t)J.frame - kf[i], (pcap - used) / px);
        if (k > 0) { kf[i] += k; used += k * px; }
    }
    struct Plan { int jid; int k; int64_t off; bool reset; };
    std::vector<Plan> plan;
    std::vector<Region> regs;
    int64_t off = 0;
    for (size_t i = 0; i < cand.size(); ++i) {
        if (kf[i] == 0) continue;
        int jid = cand[i];
        JobRec& J = c->jobs[jid];
        int64_t px = (int64_t)J.W * J.H;
        for (int64_t f = 0; f < kf[i]; ++f) regs.push_back({jid, (int32_t)f, (uint32_t)J.engine(), 0, 0, J.W, J.H});
        plan.push_back({jid, (int)kf[i], off, J.frame == 0});
        off += kf[i] * px;
        if (J.t_start < 0) J.t_start = ms_since(c);
    }
    if (regs.empty()) return 0;
    for (const Plan& pl : plan)
        if (int r = upload_job_table(c, pl.jid, c->samples + pl.off)) return r;
    if (int r = run_batch(c, regs, c->serve_rec, DevCapture{})) return r;
    for (const Plan& pl : plan) {
        JobRec& J = c->jobs[pl.jid];
        ResolveArgs ra{c->d_jobs, pl.jid, 0, 0, J.W, J.H, pl.k, pl.reset ? 1 : 0, J.accum};
        launch_resolve(ra, c->stream);
    }
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->stream));
    double now = ms_since(c);
    for (const Plan& pl : plan) {
        JobRec& J = c->jobs[pl.jid];
        J.frame += (uint64_t)pl.k;
        J.planned = (int64_t)J.frame;
        J.st.frames += (uint64_t)pl.k;
        J.st.samples += (uint64_t)pl.k * J.W * J.H;
        if ((int64_t)J.frame >= J.target) J.t_done = now;
        *frames_out += pl.k;
    }
    return 0;
}

// ---- continuous serving (path tracing) ----------------------------------

int64_t stream_cap(const yn_ctx* c)
{
    return c->batch_paths > 0 ? std::min(c->batch_paths, c->cap_paths) : c->cap_paths;
}

// First fit in the samples buffer's free ranges; -1 when nothing fits.
int64_t samples_alloc(StreamState& S, int64_t n)
{
    for (auto it = S.sfree.begin(); it != S.sfree.end(); ++it) {
        if (it->second < n) continue;
        const int64_t off = it->first, len = it->second;
        S.sfree.erase(it);
        if (len > n) S.sfree[off + n] = len - n;
        return off;
    }
    return -1;
}

void samples_free(StreamState& S, int64_t off, int64_t n)
{
    auto nx = S.sfree.lower_bound(off);
    if (nx != S.sfree.end() && off + n == nx->first) { n += nx->second; nx = S.sfree.erase(nx); }
    if (nx != S.sfree.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) { pv->second += n; return; }
    }
    S.sfree[off] = n;
}

// Jobs the continuous loop renders: path tracing, a frame that fits the
// samples buffer.  The others are served in rounds.
bool stream_takes(const yn_ctx* c, const JobRec& J)
{
    return J.p.mode == YN_MODE_PT && (int64_t)J.W * J.H <= c->sample_cap;
}

int64_t stream_pending(const yn_ctx* c)
{
    int64_t n = 0;
    for (int i : c->strm.issue) n += c->strm.pieces[i].items - c->strm.pieces[i].issued;
    return n;
}

bool stream_has_pieces(const yn_ctx* c)
{
    for (const Piece& P : c->strm.pieces) if (P.live) return true;
    return false;
}

// An empty wavefront: every slot of the stream's capacity is back on the
// free list and no piece is waiting.
int stream_start(yn_ctx* c)
{
    StreamState& S = c->strm;
    hipStream_t s = c->stream;
    S.active = true;
    S.it = 0;
    S.cap = stream_cap(c);
    S.issue.clear();
    for (Piece& P : S.pieces) P = Piece();
    S.sfree.clear();
    S.sfree[0] = c->sample_cap;
    S.live = 0;
    for (auto& v : S.last) v = 0;
    for (int k = 0; k < 2; ++k) {
        c->E[k].count = &c->ctr->q[k][0];
        c->Sh[k].count = &c->ctr->q[2 + k][0];
    }
    c->ev_used = 0;
    HIPCK(c, hipMemsetAsync(c->ctr, 0, sizeof(DevCounters), s));
    HIPCK(c, hipGetLastError());
    return 0;
}

// Cuts admitted path-tracing jobs (oldest first) into pieces until about
// one capacity of work items waits to enter the wavefront, or the samples
// buffer or the piece table is full.  A piece holds up to a sixteenth of the
// capacity in items (at least one frame).
int stream_plan(yn_ctx* c)
{
    StreamState& S = c->strm;
    int64_t pending = stream_pending(c);
    if (pending >= S.cap) return 0;
    std::vector<int> order;
    for (int i = 0; i < (int)c->jobs.size(); ++i) {
        const JobRec& J = c->jobs[i];
        if (J.live && J.served && J.admitted && stream_takes(c, J) && J.planned < J.target) order.push_back(i);
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) { return c->jobs[a].seq < c->jobs[b].seq; });
    const int64_t piece_items = std::max<int64_t>(S.cap / 16, 1);
    int free_piece = 0;
    for (int jid : order) {
        JobRec& J = c->jobs[jid];
        const int64_t npix = (int64_t)J.W * J.H;
        while (pending < S.cap && J.planned < J.target) {
            const int64_t k = std::min<int64_t>(J.target - J.planned, std::max<int64_t>(1, piece_items / npix));
            const int64_t items = k * npix;
            while (free_piece < kMaxPieces && S.pieces[free_piece].live) ++free_piece;
            if (free_piece >= kMaxPieces) { c->serve_rec.st.plan_stalls++; return 0; }
            const int64_t off = samples_alloc(S, items);
            if (off < 0) { c->serve_rec.st.plan_stalls++; return 0; }
            Piece& P = S.pieces[free_piece];
            P = Piece();
            P.live = true;
            P.job = jid;
            P.scene = J.scene;
            P.f0 = J.planned;
            P.k = (int32_t)k;
            P.items = items;
            P.soff = off;
            for (int64_t f = 0; f < k; ++f) P.seeds.push_back((uint32_t)J.engine());
            J.planned += k;
            DevJob d = make_devjob(J);
            d.samples = c->samples + off;
            HIPCK(c, hipMemcpyAsync(c->d_jobs + kMaxJobs + free_piece, &d, sizeof(DevJob), hipMemcpyHostToDevice,
                                    c->stream));
            HIPCK(c, hipMemsetAsync(c->done_ctr + free_piece, 0, sizeof(uint32_t), c->stream));
            S.issue.push_back(free_piece);
            pending += items;
            if (J.t_start < 0) J.t_start = ms_since(c);
        }
    }
    return 0;
}

// Turns waiting work items into paths: their rays are appended behind the
// rays the next iteration traces, their records beside them.  Once at least
// an eighth of the capacity is free (or all the waiting work fits, or the
// wavefront is empty).
int stream_refill(yn_ctx* c)
{
    StreamState& S = c->strm;
    const int64_t pending = stream_pending(c);
    const int64_t f = S.cap - S.live;
    if (pending == 0 || f <= 0) return 0;
    const bool empty = S.live == 0;
    if (!empty && f < std::min<int64_t>(pending, std::max<int64_t>(S.cap / 8, 1))) return 0;
    const int64_t n = std::min<int64_t>(f, pending);
    hipStream_t s = c->stream;
    // segments: items of the waiting pieces in order, frame by frame
    S.segs.clear();
    int64_t left = n, base = 0;
    while (left > 0 && !S.issue.empty()) {
        Piece& P = S.pieces[S.issue.front()];
        const JobRec& J = c->jobs[P.job];
        const int64_t npix = (int64_t)J.W * J.H;
        const int64_t take = std::min<int64_t>(left, P.items - P.issued);
        const int64_t a = P.issued, b = a + take;
        for (int64_t fi = a / npix; fi * npix < b; ++fi) {
            const int64_t lo = std::max<int64_t>(a, fi * npix), hi = std::min<int64_t>(b, (fi + 1) * npix);
            S.segs.push_back(GenSeg{kMaxJobs + S.issue.front(), (int32_t)fi, P.seeds[fi], 0, 0, J.W, J.H,
                                    (int32_t)base, (int32_t)(lo - fi * npix)});
            base += hi - lo;
        }
        P.issued = b;
        left -= take;
        if (P.issued == P.items) S.issue.pop_front();
    }
    if ((int)S.segs.size() > c->seg_cap) {
        if (c->d_segs) hipFree(c->d_segs);
        c->d_segs = nullptr;
        c->seg_cap = std::max<int>((int)S.segs.size(), 256);
        if (int r = dalloc(c, &c->d_segs, (size_t)c->seg_cap)) return r;
    }
    HIPCK(c, hipMemcpyAsync(c->d_segs, S.segs.data(), sizeof(GenSeg) * S.segs.size(), hipMemcpyHostToDevice, s));
    GenArgs g{};
    g.scenes = c->d_scenes;
    g.jobs = c->d_jobs;
    g.P = c->P;
    g.rec = c->P.R[S.it & 1];
    g.ext_out = c->E[S.it & 1];
    g.segs = c->d_segs;
    g.n_segs = (int32_t)S.segs.size();
    g.n_paths = (int32_t)n;
    g.item0 = 0;
    g.queue_base = &c->ctr->q[11][0];
    launch_generate(g, DevCapture{}, s);
    HIPCK(c, hipGetLastError());
    c->serve_rec.st.refills++;
    S.live += n;
    return 0;
}

// Folds finished pieces into their jobs' images, each job's in frame order,
// and gives their samples ranges and table entries back.
int stream_fold(yn_ctx* c)
{
    StreamState& S = c->strm;
    for (int i = 0; i < kMaxPieces; ++i) {
        Piece& P = S.pieces[i];
        if (P.live && !P.done && P.issued == P.items && c->h_done[i] >= (uint32_t)P.items) P.done = true;
    }
    const double now = ms_since(c);
    for (bool progress = true; progress;) {
        progress = false;
        for (int i = 0; i < kMaxPieces; ++i) {
            Piece& P = S.pieces[i];
            if (!P.live || !P.done) continue;
            if (!P.orphan) {
                JobRec& J = c->jobs[P.job];
                if ((int64_t)J.frame != P.f0) continue;      // an earlier piece of the job is not folded yet
                ResolveArgs ra{c->d_jobs, kMaxJobs + i, 0, 0, J.W, J.H, P.k, P.f0 == 0 ? 1 : 0, J.accum};
                launch_resolve(ra, c->stream);
                J.frame += (uint64_t)P.k;
                J.st.frames += (uint64_t)P.k;
                J.st.samples += (uint64_t)P.items;
                if ((int64_t)J.frame >= J.target) J.t_done = now;
            }
            samples_free(S, P.soff, P.items);
            P = Piece();
            progress = true;
        }
    }
    HIPCK(c, hipGetLastError());
    return 0;
}

// One group of four bounce iterations of the persistent wavefront, with
// admission, planning and refill before it and folding after it.  `plan`
// false: no new pieces (the loop drains).  *ran: whether any path was in
// flight.
int stream_group(yn_ctx* c, bool plan, bool* ran)
{
    StreamState& S = c->strm;
    *ran = false;
    const auto h0 = std::chrono::steady_clock::now();
    if (!S.active)
        if (int r = stream_start(c)) return r;
    if (int r = admit_jobs(c)) return r;
    if (plan)
        if (int r = stream_plan(c)) return r;
    if (int r = stream_refill(c)) return r;
    if (S.live == 0) {
        // nothing in flight: every issued path has ended
        if (int r = stream_fold(c)) return r;
        if (S.issue.empty() && !stream_has_pieces(c)) S.active = false;
        return 0;
    }
    std::vector<int> scene_ids;
    for (const Piece& P : S.pieces)
        if (P.live) scene_ids.push_back(P.scene);
    LoopCfg L;
    if (int r = loop_cfg(c, false, scene_ids, false, L)) return r;
    L.done_ctr = c->done_ctr;
    IterArgs iters[2];
    make_iter(c, L, 0, iters[0]);
    make_iter(c, L, 1, iters[1]);
    std::vector<Span> spans;
    c->ev_used = 0;
    c->serve_rec.st.groups++;
    c->serve_rec.st.paths_in_flight += (uint64_t)S.live;
    c->serve_rec.st.ms_host += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    if (c->cfg.graph) {     // a group starts at queue parity 0 (four iterations per group)
        if (int r = ensure_group_graph(c, L, iters, 4)) return r;
        HIPCK(c, hipGraphLaunch(c->gexec, c->stream));
        S.it += 4;
    } else {
        for (int g = 0; g < 4; ++g, ++S.it)
            if (int r = enqueue_iter(c, L, iters[S.it & 1], DevCapture{}, &spans)) return r;
    }
    hipStream_t s = c->stream;
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipMemcpyAsync(c->h_ctr, c->ctr, kCtrHostBytes, hipMemcpyDeviceToHost, s));
    HIPCK(c, hipMemcpyAsync(c->h_done, c->done_ctr, sizeof(uint32_t) * kMaxPieces, hipMemcpyDeviceToHost, s));
    HIPCK(c, hipStreamSynchronize(s));
    *ran = true;
    if (c->h_ctr->q[7][0]) return fail(c, YN_ERR_INVALID, "traversal stack overflow");
    yn_stats& st = c->serve_rec.st;
    add_spans(spans, st);
    const unsigned long long now[6] = {c->h_ctr->rays_ext, c->h_ctr->rays_shadow, c->h_ctr->nodes_ext,
                                       c->h_ctr->tris_ext, c->h_ctr->nodes_shadow, c->h_ctr->tris_shadow};
    st.rays_ext += now[0] - S.last[0];
    st.rays_shadow += now[1] - S.last[1];
    st.nodes_ext += now[2] - S.last[2];
    st.tris_ext += now[3] - S.last[3];
    st.nodes_shadow += now[4] - S.last[4];
    st.tris_shadow += now[5] - S.last[5];
    for (int k = 0; k < 6; ++k) S.last[k] = now[k];
    st.iterations += 4;
    st.trace_launches_ext += 4;
    st.trace_launches_shadow += 4;
    // every path in flight has one ray in the queues the next iteration reads
    S.live = (int64_t)c->h_ctr->q[S.it & 1][0] + c->h_ctr->q[2 + (S.it & 1)][0];
    return stream_fold(c);
}

// Runs the continuous loop without new pieces until every issued path has
// ended and every piece is folded: other entry points use the same
// wavefront buffers.
int stream_drain(yn_ctx* c)
{
    for (int guard = 0; c->strm.active; ++guard) {
        if (guard > 1000000) return fail(c, YN_ERR_HIP, "continuous loop did not drain");
        // items not yet issued stay with their pieces; finish them too
        bool ran = false;
        if (int r = stream_group(c, false, &ran)) return r;
    }
    return 0;
}

// GPU layout of scene `sid` at `width` from its kept host arrays, uploaded
// in place of the previous one (the scene id stays valid).
// Whether every present child box of the laid-out tree has finite
// coordinates within 2^60, the bound under which the traversal's FMA slab
// form cannot overflow (boxes of absent children are never tested).
static bool fma_slab_ok(const GpuLayout& L)
{
    const int w = L.width, fpn = w == 2 ? 16 : 32;
    for (int n = 0; n < L.n_nodes; ++n) {
        const float* g = &L.nodes[(size_t)n * fpn];
        const int32_t* ref = reinterpret_cast<const int32_t*>(g + (w == 2 ? 12 : 24));
        for (int k = 0; k < w; ++k) {
            if ((uint32_t)ref[k] == kEmptyRef) continue;
            for (int a = 0; a < 3; ++a) {
                const float lo = g[2 * w * a + k], hi = g[2 * w * a + w + k];
                if (!(std::fabs(lo) <= 0x1p60f) || !(std::fabs(hi) <= 0x1p60f)) return false;
            }
        }
    }
    return true;
}

// The table of every live scene's materials, in scene order, and each
// scene's mat_base in it.  A shade launch over several scenes stages the
// table in LDS when it has at most kLdsMats entries (else it is not made).
int material_table(yn_ctx* c)
{
    HIPCK(c, hipStreamSynchronize(c->stream));
    std::vector<yn_material> all;
    for (int i = 0; i < kMaxScenes; ++i) {
        SceneRec& S = c->scenes[i];
        if (!S.live) continue;
        S.dev.mat_base = (int32_t)all.size();
        all.insert(all.end(), S.h_mats.begin(), S.h_mats.end());
        HIPCK(c, hipMemcpy(c->d_scenes + i, &S.dev, sizeof(DevScene), hipMemcpyHostToDevice));
    }
    if (c->d_mtab) hipFree(c->d_mtab);
    c->d_mtab = nullptr;
    c->n_mtab = (int32_t)all.size();
    if (!all.empty() && all.size() <= (size_t)kLdsMats) {
        HIPCK(c, hipMalloc(&c->d_mtab, all.size() * sizeof(yn_material)));
        HIPCK(c, hipMemcpy(c->d_mtab, all.data(), all.size() * sizeof(yn_material), hipMemcpyHostToDevice));
    }
    return 0;
}

int build_scene(yn_ctx* c, int sid, int width)
{
    SceneRec& S = c->scenes[sid];
    GpuLayout L;
    std::string err;
    LayoutOpts lo;
    lo.width = width;
    lo.hoist_frac = c->cfg.hoist_frac;
    lo.leaf_max = c->cfg.leaf_max;
    lo.rebuild = c->cfg.bvh_build == 1;
    int r = make_layout(S.h_tris.data(), (int32_t)S.h_tris.size(), (int32_t)S.h_mats.size(), S.h_nodes.data(),
                        (int32_t)S.h_nodes.size(), lo, L, err);
    if (r == 0 && lo.rebuild && L.max_depth > kMaxStack) {
        // a pathological triangle set can make the own tree deeper than the
        // traversal stack; the caller's tree is laid out instead
        lo.rebuild = false;
        r = make_layout(S.h_tris.data(), (int32_t)S.h_tris.size(), (int32_t)S.h_mats.size(), S.h_nodes.data(),
                        (int32_t)S.h_nodes.size(), lo, L, err);
    }
    if (r) return fail(c, r, err);
    if (L.max_depth > kMaxStack) return fail(c, YN_ERR_INVALID, "BVH deeper than the traversal stack bound");
    HIPCK(c, hipStreamSynchronize(c->stream));
    void* nb[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    auto up = [&](int k, const void* src, size_t bytes) -> int {
        HIPCK(c, hipMalloc(&nb[k], std::max<size_t>(bytes, 16)));
        if (bytes) HIPCK(c, hipMemcpy(nb[k], src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if ((r = up(0, L.nodes.data(), L.nodes.size() * 4)) || (r = up(1, L.tri_isect.data(), L.tri_isect.size() * 4)) ||
        (r = up(2, L.tri_shade.data(), L.tri_shade.size() * 4)) || (r = up(3, L.tri_mat.data(), L.tri_mat.size() * 4)) ||
        (r = up(4, S.h_mats.data(), sizeof(yn_material) * S.h_mats.size())) ||
        (r = up(5, L.hoist_box.data(), L.hoist_box.size() * 4))) {
        for (void*& b : nb) { if (b) hipFree(b); b = nullptr; }
        return r;
    }
    for (int k = 0; k < 6; ++k) {
        if (S.bufs[k]) hipFree(S.bufs[k]);
        S.bufs[k] = nb[k];
    }
    DevScene& D = S.dev;
    D.nodes = (const float4*)S.bufs[0];
    D.tri_isect = (const float4*)S.bufs[1];
    D.tri_shade = (const float4*)S.bufs[2];
    D.tri_mat = (const int32_t*)S.bufs[3];
    D.mats = (const yn_material*)S.bufs[4];
    D.hoist_box = (const float4*)S.bufs[5];
    D.light = S.h_light;
    // the reference constant has these w's (udpt.cl:97-104)
    D.light.pos.w = 1.0f; D.light.normal.w = 0.0f; D.light.ke.w = 0.0f;
    D.light.edge_l.w = 0.0f; D.light.edge_w.w = 0.0f;
    D.root_ref = L.root_ref;
    std::memcpy(D.root_box, L.root_box, sizeof D.root_box);
    D.width = L.width;
    D.brute = L.brute;
    D.n_slots = (int32_t)L.slot_tri.size();
    D.n_hoist = L.n_hoist;
    // YUNE_SLAB_REFERENCE=1 keeps the scene's rays on the reference's slab
    // arithmetic, for comparing the two forms on one scene
    const char* ref_slab = std::getenv("YUNE_SLAB_REFERENCE");
    D.fma_slab = fma_slab_ok(L) && !(ref_slab && ref_slab[0] == '1') ? 1 : 0;
    HIPCK(c, hipMemcpy(c->d_scenes + sid, &D, sizeof(DevScene), hipMemcpyHostToDevice));
    S.info.n_tri = (int32_t)S.h_tris.size();
    S.info.n_ref_nodes = (int32_t)S.h_nodes.size();
    S.info.n_gpu_nodes = L.n_nodes;
    S.info.width = L.width;
    S.info.max_depth = L.max_depth;
    S.info.n_hoist = L.n_hoist;
    S.info.node_bytes = L.nodes.size() * 4;
    S.info.tri_bytes = L.tri_isect.size() * 4;
    S.info.shade_bytes = L.tri_shade.size() * 4 + L.tri_mat.size() * 4;
    return 0;
}

int check_ctx(yn_ctx* c) { return c ? 0 : fail(nullptr, YN_ERR_INVALID, "null context"); }
int check_job(yn_ctx* c, int jid)
{
    if (jid < 0 || jid >= (int)c->jobs.size() || !c->jobs[jid].live) return fail(c, YN_ERR_NOT_FOUND, "unknown job");
    return 0;
}
int check_scene(yn_ctx* c, int sid)
{
    if (sid < 0 || sid >= (int)c->scenes.size() || !c->scenes[sid].live)
        return fail(c, YN_ERR_NOT_FOUND, "unknown scene");
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

void yn_params_reference(yn_render_params* p, int32_t mode)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->mode = mode;
    p->gi = 1;
    p->mis = 0;                                   // udpt.cl:9
    p->max_bounces = 100000;                      // udpt.cl:463
    p->rr_threshold = mode == YN_MODE_BDPT ? 4 : 6;   // bdpt.cl:4, udpt.cl:6
    p->bdpt_bounces = 20;                         // bdpt.cl:7
    p->oren_nayar = 0;
    p->fix_mis = 0;
    p->background[0] = p->background[1] = p->background[2] = 0.4f;   // udpt.cl:109
    p->background[3] = 1.0f;
}

const char* yn_last_error(yn_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int yn_create(const yn_config* cfg, yn_ctx** out)
{
    if (!out) return fail(nullptr, YN_ERR_INVALID, "null output");
    *out = nullptr;
    yn_config k{};
    if (cfg) k = *cfg;
    if (k.max_paths <= 0) k.max_paths = 1 << 26;     // ~27 GB of wavefront state
    if (k.mem_budget == 0) k.mem_budget = (size_t)64 << 30;
    if (k.bvh_width == 0) k.bvh_width = 4;
    if (k.trace_block == 0) k.trace_block = 256;
    if (k.trace_chunk <= 0) k.trace_chunk = 64;
    if (k.hoist_frac == 0.0f) k.hoist_frac = 0.01f;
    if (k.bvh_build == 0) k.bvh_build = 1;
    if (k.serve_mode != 0 && k.serve_mode != 1) return fail(nullptr, YN_ERR_INVALID, "serve_mode must be 0 or 1");
    if (k.bvh_build != 1 && k.bvh_build != 2) return fail(nullptr, YN_ERR_INVALID, "bvh_build must be 1 or 2");
    if (k.leaf_max <= 0) k.leaf_max = k.bvh_build == 1 ? 1 : 2;
    if (k.leaf_max > 10) k.leaf_max = 10;
    if (k.shade_waves == 0) k.shade_waves = 4;
    if (k.lds_stack != 0 && k.lds_stack != 8 && k.lds_stack != 12 && k.lds_stack != 16 && k.lds_stack != 32)
        return fail(nullptr, YN_ERR_INVALID, "lds_stack must be 0, 8, 12, 16 or 32");
    if (k.shade_waves < 2 || k.shade_waves > 4) return fail(nullptr, YN_ERR_INVALID, "shade_waves must be 2, 3 or 4");
    if (k.batch_factor == 0) k.batch_factor = 1;
    if (k.batch_factor != 1 && k.batch_factor != 2) return fail(nullptr, YN_ERR_INVALID, "batch_factor must be 1 or 2");
    if (k.trace_overlap != 0 && k.trace_overlap != 1) return fail(nullptr, YN_ERR_INVALID, "trace_overlap must be 0 or 1");
    if (k.graph != 0 && k.graph != 1) return fail(nullptr, YN_ERR_INVALID, "graph must be 0 or 1");
    if (k.trace_xcd == 0) k.trace_xcd = 8;
    if (k.trace_refill == 0) k.trace_refill = 48;
    if (k.trace_refill < -1 || k.trace_refill > 64) return fail(nullptr, YN_ERR_INVALID, "trace_refill must be -1 or 1..64");
    if (k.trace_xcd < 1 || k.trace_xcd > kMaxXcd) return fail(nullptr, YN_ERR_INVALID, "trace_xcd must be 1..16");
    if (k.bvh_width != 2 && k.bvh_width != 4) return fail(nullptr, YN_ERR_INVALID, "bvh_width must be 2 or 4");
    if (k.trace_block != 64 && k.trace_block != 128 && k.trace_block != 256)
        return fail(nullptr, YN_ERR_INVALID, "trace_block must be 64, 128 or 256");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= k.device || k.device < 0)
        return fail(nullptr, YN_ERR_NO_DEVICE, "no HIP device available");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, k.device) != hipSuccess)
        return fail(nullptr, YN_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, YN_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", the library is built for gfx950");
    yn_ctx* c = new yn_ctx();
    c->cfg = k;
    c->device = k.device;
    c->n_cu = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    c->hbm = prop.totalGlobalMem;
    int r = 0;
    auto bail = [&](int code) { int rr = code; yn_destroy(c); return rr; };
    if (hipSetDevice(k.device) != hipSuccess) { g_err = "hipSetDevice failed"; delete c; return YN_ERR_HIP; }
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess) {
        g_err = "stream setup failed"; delete c; return YN_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipStreamCreate failed"; delete c; return YN_ERR_HIP;
    }
    for (int i = 0; i < kEvPool; ++i) hipEventCreate(&c->ev[i]);
    if (trace_noalloc_init(k.device, &c->na_scratch) != 0) {
        g_err = "hipGetSymbolAddress of the no-allocation scratch array failed"; return bail(YN_ERR_HIP);
    }
    c->scenes.resize(kMaxScenes);
    c->jobs.resize(kMaxJobs);
    if ((r = dalloc(c, &c->d_scenes, kMaxScenes)) || (r = dalloc(c, &c->d_jobs, kMaxJobs + kMaxPieces)) ||
        (r = dalloc(c, &c->ctr, 1)) || (r = dalloc(c, &c->done_ctr, kMaxPieces))) {
        g_err = c->err; return bail(r);
    }
    c->allocs.push_back(c->d_scenes);
    c->allocs.push_back(c->d_jobs);
    c->allocs.push_back(c->ctr);
    c->allocs.push_back(c->done_ctr);
    if (hipHostMalloc((void**)&c->h_ctr, sizeof(DevCounters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->h_done, sizeof(uint32_t) * kMaxPieces, hipHostMallocDefault) != hipSuccess) {
        g_err = "hipHostMalloc failed"; return bail(YN_ERR_HIP);
    }
    c->strm.pieces.resize(kMaxPieces);
    if ((r = alloc_wavefront(c, k.max_paths))) { g_err = c->err; return bail(r); }
    *out = c;
    return 0;
}

void yn_destroy(yn_ctx* c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    free_all(c);
    if (c->d_segs) hipFree(c->d_segs);
    if (c->d_mtab) hipFree(c->d_mtab);
    if (c->ovf_ref) hipFree(c->ovf_ref);
    if (c->ovf_t) hipFree(c->ovf_t);
    if (c->h_ctr) hipHostFree(c->h_ctr);
    if (c->h_done) hipHostFree(c->h_done);
    for (int i = 0; i < kEvPool; ++i) if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->join_ev) hipEventDestroy(c->join_ev);
    delete c;
}

int yn_device_info(yn_ctx* c, char* name, int name_len, int32_t* n_cu, size_t* hbm)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (name && name_len > 0) { std::strncpy(name, c->name.c_str(), name_len - 1); name[name_len - 1] = 0; }
    if (n_cu) *n_cu = c->n_cu;
    if (hbm) *hbm = c->hbm;
    return 0;
}

int yn_scene_create(yn_ctx* c, const yn_triangle* tris, int32_t n_tri, const yn_material* mats, int32_t n_mat,
                    const yn_bvh_node* nodes, int32_t n_node, const yn_light_quad* light, int32_t* scene_id)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (!light || !scene_id || n_mat <= 0 || !mats || n_tri < 0 || n_node < 0 || (n_tri > 0 && !tris) ||
        (n_node > 0 && !nodes))
        return fail(c, YN_ERR_INVALID, "bad scene arguments");
    int sid = -1;
    for (int i = 0; i < kMaxScenes; ++i) if (!c->scenes[i].live) { sid = i; break; }
    if (sid < 0) return fail(c, YN_ERR_NO_MEMORY, "too many scenes");
    SceneRec& S = c->scenes[sid];
    S = SceneRec();
    S.h_tris.assign(tris, tris + n_tri);
    S.h_mats.assign(mats, mats + n_mat);
    S.h_nodes.assign(nodes, nodes + n_node);
    S.h_light = *light;
    if (int r = build_scene(c, sid, c->cfg.bvh_width)) { S = SceneRec(); return r; }
    S.live = true;
    if (int r = material_table(c)) {
        for (void*& bf : S.bufs) { if (bf) hipFree(bf); bf = nullptr; }
        S = SceneRec();
        material_table(c);
        return r;
    }
    *scene_id = sid;
    return 0;
}

int yn_scene_release(yn_ctx* c, int32_t sid)
{
    if (check_ctx(c) || check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    for (auto& j : c->jobs)
        if (j.live && j.scene == sid) return fail(c, YN_ERR_INVALID, "scene still used by a job");
    for (const Piece& P : c->strm.pieces)
        if (P.live && P.scene == sid) {     // paths of a released job still trace it
            if (int r = stream_drain(c)) return r;
            break;
        }
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (void*& b : c->scenes[sid].bufs) { if (b) hipFree(b); b = nullptr; }
    c->scenes[sid] = SceneRec();
    return material_table(c);
}

int yn_scene_info_get(yn_ctx* c, int32_t sid, yn_scene_info* out)
{
    if (check_ctx(c) || check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (!out) return fail(c, YN_ERR_INVALID, "null output");
    *out = c->scenes[sid].info;
    return 0;
}

int yn_job_create(yn_ctx* c, int32_t sid, const yn_cam* cam, int32_t W, int32_t H, const yn_render_params* p,
                  uint32_t base_seed, int32_t* job_id)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (!valid_cam(cam)) return fail(c, YN_ERR_INVALID, "camera row 4 must be (0,0,0,1)");
    if (!valid_params(p)) return fail(c, YN_ERR_INVALID, "bad render parameters");
    if (W <= 0 || H <= 0 || (int64_t)W * H > (1ll << 28) || !job_id) return fail(c, YN_ERR_INVALID, "bad image size");
    size_t bytes = (size_t)W * H * sizeof(float4);
    if (c->job_bytes + bytes > c->cfg.mem_budget) return fail(c, YN_ERR_NO_MEMORY, "job memory budget exceeded");
    int jid = -1;
    for (int i = 0; i < kMaxJobs; ++i) if (!c->jobs[i].live && i != c->pt_job_slot) { jid = i; break; }
    if (jid < 0) return fail(c, YN_ERR_NO_MEMORY, "too many jobs");
    JobRec& J = c->jobs[jid];
    J = JobRec();
    int r = dalloc(c, &J.accum, (size_t)W * H);
    if (r) return r;
    HIPCK(c, hipMemset(J.accum, 0, bytes));
    J.live = true;
    J.scene = sid;
    J.cam = *cam;
    J.W = W;
    J.H = H;
    J.p = *p;
    J.engine.seed(base_seed);
    J.bytes = bytes;
    c->job_bytes += bytes;
    *job_id = jid;
    return upload_job_table(c, jid, c->samples);
}

int yn_job_render(yn_ctx* c, int32_t jid, int32_t n_frames)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!c->jobs[jid].admitted) return fail(c, YN_ERR_INVALID, "job not admitted yet (memory budget)");
    if (n_frames < 0) return fail(c, YN_ERR_INVALID, "negative frame count");
    JobRec& J = c->jobs[jid];
    auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> seeds(n_frames);
    for (auto& s : seeds) s = (uint32_t)J.engine();
    DevCapture cap{};
    std::vector<void*> cap_bufs;
    size_t npix = (size_t)J.W * J.H;
    if (J.cap_prim_tri && n_frames > 0) {
        int r = 0;
        r |= dalloc(c, &cap.prim_tri, npix); cap_bufs.push_back(cap.prim_tri);
        r |= dalloc(c, &cap.prim_light, npix); cap_bufs.push_back(cap.prim_light);
        r |= dalloc(c, &cap.prim_t, npix); cap_bufs.push_back(cap.prim_t);
        r |= dalloc(c, &cap.shadow_o, npix); cap_bufs.push_back(cap.shadow_o);
        r |= dalloc(c, &cap.shadow_d, npix); cap_bufs.push_back(cap.shadow_d);
        r |= dalloc(c, &cap.shadow_occ, npix); cap_bufs.push_back(cap.shadow_occ);
        if (r) { for (void* b : cap_bufs) if (b) hipFree(b); return YN_ERR_NO_MEMORY; }
    }
    int r = render_rect(c, jid, J.accum, 0, 0, J.W, J.H, seeds, J.frame == 0, cap);
    if (r == 0) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) r = fail(c, YN_ERR_HIP, std::string("render: ") + hipGetErrorString(e));
    }
    if (r == 0 && cap.prim_tri) {
        std::vector<float4> so(npix), sd(npix);
        hipMemcpy(J.cap_prim_tri, cap.prim_tri, npix * 4, hipMemcpyDeviceToHost);
        hipMemcpy(J.cap_prim_light, cap.prim_light, npix * 4, hipMemcpyDeviceToHost);
        hipMemcpy(J.cap_prim_t, cap.prim_t, npix * 4, hipMemcpyDeviceToHost);
        hipMemcpy(J.cap_shadow_occ, cap.shadow_occ, npix * 4, hipMemcpyDeviceToHost);
        hipMemcpy(so.data(), cap.shadow_o, npix * 16, hipMemcpyDeviceToHost);
        hipMemcpy(sd.data(), cap.shadow_d, npix * 16, hipMemcpyDeviceToHost);
        for (size_t i = 0; i < npix; ++i) {
            float* q = J.cap_shadow + 8 * i;
            q[0] = so[i].x; q[1] = so[i].y; q[2] = so[i].z; q[3] = so[i].w;
            q[4] = sd[i].x; q[5] = sd[i].y; q[6] = sd[i].z; q[7] = sd[i].w;
        }
        J.cap_prim_tri = nullptr;
    }
    for (void* b : cap_bufs) hipFree(b);
    if (r) return r;
    J.frame += (uint64_t)n_frames;
    J.st.frames += (uint64_t)n_frames;
    J.st.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int yn_job_capture(yn_ctx* c, int32_t jid, int32_t* prim_tri, int32_t* prim_light, float* prim_t, float* shadow_ray,
                   int32_t* shadow_occ)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!prim_tri || !prim_light || !prim_t || !shadow_ray || !shadow_occ) return fail(c, YN_ERR_INVALID, "null capture");
    JobRec& J = c->jobs[jid];
    J.cap_prim_tri = prim_tri; J.cap_prim_light = prim_light; J.cap_prim_t = prim_t;
    J.cap_shadow = shadow_ray; J.cap_shadow_occ = shadow_occ;
    return 0;
}

int yn_job_set_counting(yn_ctx* c, int32_t jid, int32_t on)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    c->jobs[jid].counting = on;
    return 0;
}

int yn_job_stats(yn_ctx* c, int32_t jid, yn_stats* out)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!out) return fail(c, YN_ERR_INVALID, "null output");
    *out = c->jobs[jid].st;
    return 0;
}

int yn_job_read(yn_ctx* c, int32_t jid, float* rgba)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!c->jobs[jid].admitted) return fail(c, YN_ERR_INVALID, "job not admitted yet (memory budget)");
    if (!rgba) return fail(c, YN_ERR_INVALID, "null output");
    JobRec& J = c->jobs[jid];
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(rgba, J.accum, (size_t)J.W * J.H * 16, hipMemcpyDeviceToHost));
    return 0;
}

int yn_job_read_tonemapped(yn_ctx* c, int32_t jid, float* rgba)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!c->jobs[jid].admitted) return fail(c, YN_ERR_INVALID, "job not admitted yet (memory budget)");
    if (!rgba) return fail(c, YN_ERR_INVALID, "null output");
    JobRec& J = c->jobs[jid];
    size_t n = (size_t)J.W * J.H;
    float4* tmp = nullptr;
    int r = dalloc(c, &tmp, n);
    if (r) return r;
    launch_tonemap(J.accum, tmp, (int)n, J.tonemap_op, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(rgba, tmp, n * 16, hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, YN_ERR_HIP, std::string("tonemap: ") + hipGetErrorString(e));
    return 0;
}

int yn_job_save(yn_ctx* c, int32_t jid, const char* path)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!c->jobs[jid].admitted) return fail(c, YN_ERR_INVALID, "job not admitted yet (memory budget)");
    if (!path) return fail(c, YN_ERR_INVALID, "null path");
    JobRec& J = c->jobs[jid];
    std::string p(path);
    size_t n = (size_t)J.W * J.H;
    std::vector<float> img(n * 4);
    bool hdr = p.size() >= 4 && p.compare(p.size() - 4, 4, ".hdr") == 0;
    int r = hdr ? yn_job_read(c, jid, img.data()) : yn_job_read_tonemapped(c, jid, img.data());
    if (r) return r;
    // stbi_flip_vertically_on_write(1) (RendererCore.cpp:615): file row 0 is
    // image row H-1.
    if (hdr) {
        std::vector<float> rgb(n * 3);
        for (int y = 0; y < J.H; ++y)
            for (int x = 0; x < J.W; ++x)
                for (int k = 0; k < 3; ++k)
                    rgb[((size_t)y * J.W + x) * 3 + k] = img[((size_t)(J.H - 1 - y) * J.W + x) * 4 + k];
        if (write_hdr(p, J.W, J.H, rgb.data())) return fail(c, YN_ERR_IO, "cannot write " + p);
    } else {
        std::vector<uint8_t> rgb(n * 3);
        for (int y = 0; y < J.H; ++y)
            for (int x = 0; x < J.W; ++x)
                for (int k = 0; k < 3; ++k) {
                    float v = img[((size_t)(J.H - 1 - y) * J.W + x) * 4 + k];
                    v = std::isnan(v) ? 0.0f : std::min(1.0f, std::max(0.0f, v));
                    rgb[((size_t)y * J.W + x) * 3 + k] = (uint8_t)std::lround(v * 255.0f);
                }
        if (write_png(p, J.W, J.H, rgb.data())) return fail(c, YN_ERR_IO, "cannot write " + p);
    }
    return 0;
}

int yn_job_release(yn_ctx* c, int32_t jid)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    JobRec& J = c->jobs[jid];
    // pieces of the job in the continuous loop: the paths already issued
    // finish (their samples are dropped), the rest is never issued
    StreamState& S = c->strm;
    for (Piece& P : S.pieces)
        if (P.live && P.job == jid && !P.orphan) { P.orphan = true; P.items = P.issued; }
    S.issue.erase(std::remove_if(S.issue.begin(), S.issue.end(),
                                 [&](int i) { return S.pieces[i].issued == S.pieces[i].items; }),
                  S.issue.end());
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (J.accum) hipFree(J.accum);
    c->job_bytes -= J.bytes;
    J = JobRec();
    return 0;
}

int yn_submit(yn_ctx* c, int32_t sid, const yn_cam* cam, int32_t W, int32_t H, const yn_render_params* p,
              uint32_t base_seed, int32_t n_frames, int32_t* job_id)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (!valid_cam(cam)) return fail(c, YN_ERR_INVALID, "camera row 4 must be (0,0,0,1)");
    if (!valid_params(p)) return fail(c, YN_ERR_INVALID, "bad render parameters");
    if (W <= 0 || H <= 0 || (int64_t)W * H > (1ll << 28) || !job_id || n_frames < 0)
        return fail(c, YN_ERR_INVALID, "bad image size or frame count");
    if ((size_t)W * H * sizeof(float4) > c->cfg.mem_budget)
        return fail(c, YN_ERR_NO_MEMORY, "image larger than the job memory budget");
    int jid = -1;
    for (int i = 0; i < kMaxJobs; ++i) if (!c->jobs[i].live && i != c->pt_job_slot) { jid = i; break; }
    if (jid < 0) return fail(c, YN_ERR_NO_MEMORY, "too many jobs");
    JobRec& J = c->jobs[jid];
    J = JobRec();
    J.live = true;
    J.served = true;
    J.admitted = false;
    J.scene = sid;
    J.cam = *cam;
    J.W = W;
    J.H = H;
    J.p = *p;
    J.engine.seed(base_seed);
    J.target = n_frames;
    J.seq = c->submit_seq++;
    J.t_submit = ms_since(c);
    if (n_frames == 0) J.t_done = J.t_submit;
    *job_id = jid;
    return admit_jobs(c);
}

int yn_serve(yn_ctx* c, int32_t max_rounds, int32_t* pending)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    for (int r = 0; r < max_rounds; ++r) {
        if (int e = admit_jobs(c)) return e;
        // the oldest admitted job with frames left decides what runs next
        int oldest = -1;
        for (int i = 0; i < (int)c->jobs.size(); ++i) {
            const JobRec& J = c->jobs[i];
            if (J.live && J.served && J.admitted && (int64_t)J.frame < J.target &&
                (oldest < 0 || J.seq < c->jobs[oldest].seq))
                oldest = i;
        }
        const bool rounds = c->cfg.serve_mode == 1;
        const bool round_job = oldest >= 0 && (rounds || !stream_takes(c, c->jobs[oldest]));
        if (round_job || rounds) {
            // a scheduler round; the continuous loop finishes its pieces first
            if (c->strm.active) {
                bool ran = false;
                if (int e = stream_group(c, false, &ran)) return e;
                continue;
            }
            int frames = 0;
            if (int e = serve_round(c, &frames)) return e;
            if (frames == 0) break;
            continue;
        }
        if (oldest < 0 && !c->strm.active) break;
        bool ran = false;
        if (int e = stream_group(c, true, &ran)) return e;
        if (!ran && oldest < 0 && !c->strm.active) break;
    }
    if (pending) {
        int n = 0;
        for (const JobRec& J : c->jobs) n += (J.live && J.served && (int64_t)J.frame < J.target) ? 1 : 0;
        *pending = n;
    }
    return 0;
}

int yn_job_status_get(yn_ctx* c, int32_t jid, yn_job_status* out)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (!out) return fail(c, YN_ERR_INVALID, "null output");
    const JobRec& J = c->jobs[jid];
    out->frames_target = (int32_t)(J.served ? J.target : (int64_t)J.frame);
    out->frames_done = (int32_t)J.frame;
    out->admitted = J.admitted ? 1 : 0;
    out->done = (J.served && (int64_t)J.frame >= J.target) ? 1 : 0;
    out->t_submit_ms = J.t_submit;
    out->t_start_ms = J.t_start;
    out->t_done_ms = J.t_done;
    return 0;
}

int yn_serve_stats(yn_ctx* c, yn_stats* out)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (!out) return fail(c, YN_ERR_INVALID, "null output");
    *out = c->serve_rec.st;
    return 0;
}

double yn_clock_ms(yn_ctx* c) { return c ? ms_since(c) : 0.0; }

int yn_config_get(yn_ctx* c, yn_config* out, int64_t* batch_paths)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (!out) return fail(c, YN_ERR_INVALID, "null output");
    *out = c->cfg;
    if (batch_paths) *batch_paths = c->batch_paths > 0 ? c->batch_paths : c->cap_paths;
    return 0;
}

// Coordinate descent over the launch and layout knobs on short timed
// renders of the caller's view: each knob in turn takes each of its values
// with the others held, and keeps a value that is more than 1% faster.
int yn_autotune(yn_ctx* c, int32_t sid, const yn_cam* cam, int32_t W, int32_t H, const yn_render_params* p,
                int32_t frames, yn_tune_result* out)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (!out || frames <= 0 || !valid_params(p) || !valid_cam(cam) || W <= 0 || H <= 0)
        return fail(c, YN_ERR_INVALID, "bad autotune arguments");
    for (const JobRec& J : c->jobs)
        if (J.live && J.served && (int64_t)J.frame < J.target)
            return fail(c, YN_ERR_INVALID, "autotune while served jobs are pending");
    if (c->strm.active)
        if (int r = stream_drain(c)) return r;
    int tried = 0;
    int err = 0;
    auto measure = [&]() -> double {
        int32_t jid = -1;
        if ((err = yn_job_create(c, sid, cam, W, H, p, 0x9e3779b9u, &jid))) return -1.0;
        double rate = -1.0;
        if (!(err = yn_job_render(c, jid, 1))) {
            yn_stats a = c->jobs[jid].st;
            auto t0 = std::chrono::steady_clock::now();
            if (!(err = yn_job_render(c, jid, frames))) {
                double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                yn_stats b = c->jobs[jid].st;
                rate = (double)((b.rays_ext - a.rays_ext) + (b.rays_shadow - a.rays_shadow)) / std::max(dt, 1e-9);
            }
        }
        yn_job_release(c, jid);
        ++tried;
        return rate;
    };
    auto set_width = [&](int w) -> int {
        for (int i = 0; i < kMaxScenes; ++i)
            if (c->scenes[i].live)
                if (int r = build_scene(c, i, w)) return r;
        c->cfg.bvh_width = w;
        return 0;
    };
    const int64_t cap = c->cap_paths;
    struct Knob { int kind; std::vector<int64_t> vals; };
    std::vector<Knob> knobs = {
        {0, {4, 2}},                            // BVH width
        {1, {256, 128, 64}},                    // trace block
        {2, {64, 32, 128}},                     // rays per wave per queue fetch
        {3, {0, 16, 12, 8, 32}},                // LDS stack entries (0: per launch)
        {4, {2, 3, 4}},                         // shade occupancy
        {5, {cap, cap / 2, cap / 4}},           // paths per batch
        {6, {1, 2}},                            // work per batch over paths in flight (regeneration)
        {7, {0, 1}},                            // overlapping closest-hit and any-hit traces
        {8, {0, 1}},                            // bounce loop replayed from a graph
        {9, {8, 16, 1}},                        // queue slices of the traces (one per XCD)
        {10, {48, 32, -1}},                     // idle lanes before a trace wave takes new rays
    };
    auto get = [&](int kind) -> int64_t {
        switch (kind) {
        case 0: return c->cfg.bvh_width;
        case 1: return c->cfg.trace_block;
        case 2: return c->cfg.trace_chunk;
        case 3: return c->cfg.lds_stack;
        case 4: return c->cfg.shade_waves;
        case 5: return c->batch_paths > 0 ? c->batch_paths : cap;
        case 6: return c->cfg.batch_factor;
        case 7: return c->cfg.trace_overlap;
        case 8: return c->cfg.graph;
        case 9: return c->cfg.trace_xcd;
        default: return c->cfg.trace_refill;
        }
    };
    auto set = [&](int kind, int64_t v) -> int {
        switch (kind) {
        case 0: return set_width((int)v);
        case 1: c->cfg.trace_block = (int)v; return 0;
        case 2: c->cfg.trace_chunk = (int)v; return 0;
        case 3: c->cfg.lds_stack = (int)v; return 0;
        case 4: c->cfg.shade_waves = (int)v; return 0;
        case 5: c->batch_paths = v; return 0;
        case 6: c->cfg.batch_factor = (int)v; return 0;
        case 7: c->cfg.trace_overlap = (int)v; return 0;
        case 8: c->cfg.graph = (int)v; return 0;
        case 9: c->cfg.trace_xcd = (int)v; return 0;
        default: c->cfg.trace_refill = (int)v; return 0;
        }
    };
    double best = measure();
    if (best < 0) return err ? err : fail(c, YN_ERR_HIP, "autotune: first render failed");
    for (int pass = 0; pass < 2; ++pass) {
        bool changed = false;
        for (const Knob& k : knobs) {
            int64_t keep = get(k.kind);
            for (int64_t v : k.vals) {
                if (v == keep || v <= 0) continue;
                if (int r = set(k.kind, v)) return r;
                double m = measure();
                if (m < 0) return err;
                if (m > best * 1.01) { best = m; keep = v; changed = true; }
            }
            if (get(k.kind) != keep)
                if (int r = set(k.kind, keep)) return r;
        }
        if (!changed) break;
    }
    out->bvh_width = c->cfg.bvh_width;
    out->trace_block = c->cfg.trace_block;
    out->trace_chunk = c->cfg.trace_chunk;
    out->lds_stack = c->cfg.lds_stack;
    out->shade_waves = c->cfg.shade_waves;
    out->batch_paths = c->batch_paths > 0 ? c->batch_paths : cap;
    out->batch_factor = c->cfg.batch_factor;
    out->trace_overlap = c->cfg.trace_overlap;
    out->graph = c->cfg.graph;
    out->trace_xcd = c->cfg.trace_xcd;
    out->trace_refill = c->cfg.trace_refill;
    out->mrays_per_s = best / 1e6;
    out->n_tried = tried;
    return 0;
}

int yn_fmath_eval(yn_ctx* c, int32_t op, const float* x, const float* y, int32_t n, float* out)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (op < 0 || op > 2 || n < 0 || (n > 0 && (!x || !out || (op == 2 && !y))))
        return fail(c, YN_ERR_INVALID, "bad fmath arguments");
    if (n == 0) return 0;
    float* d = nullptr;
    if (int r = dalloc(c, &d, (size_t)n * 3)) return r;
    hipStream_t s = c->stream;
    hipError_t e = hipMemcpyAsync(d, x, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && op == 2) e = hipMemcpyAsync(d + n, y, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        launch_fmath(op, d, d + n, n, d + 2 * (size_t)n, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return fail(c, YN_ERR_HIP, std::string("fmath: ") + hipGetErrorString(e));
    return 0;
}

int yn_pathtracer(yn_ctx* c, float* out_img, const float* in_img, const yn_cam* cam, int32_t sid, int32_t W, int32_t H,
                  const yn_render_params* params, int32_t gi_check, int32_t reset, uint32_t rand, int32_t block,
                  int32_t block_x, int32_t block_y)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (!out_img || (!in_img && !reset)) return fail(c, YN_ERR_INVALID, "null image");
    if (!valid_cam(cam)) return fail(c, YN_ERR_INVALID, "camera row 4 must be (0,0,0,1)");
    if (!valid_params(params)) return fail(c, YN_ERR_INVALID, "bad render parameters");
    if (W <= 0 || H <= 0 || block_x <= 0 || block_y <= 0 || block < 0 || block >= block_x * block_y)
        return fail(c, YN_ERR_INVALID, "bad image or block arguments");
    // tile of `block` (udpt.cl:166-170)
    int tw = (int)std::ceil((float)W / block_x), th = (int)std::ceil((float)H / block_y);
    int x0 = tw * (block % block_x), y0 = th * (block / block_x);
    int w = std::min(tw, W - x0), h = std::min(th, H - y0);
    if (w <= 0 || h <= 0) return 0;
    if (c->pt_job_slot < 0) {
        // a job slot no live job holds, kept for this entry point from now
        // on (yn_job_create and yn_submit skip it)
        for (int i = kMaxJobs - 1; i >= 0 && c->pt_job_slot < 0; --i)
            if (!c->jobs[i].live) c->pt_job_slot = i;
        if (c->pt_job_slot < 0) return fail(c, YN_ERR_NO_MEMORY, "too many jobs: no slot for yn_pathtracer");
    }
    int jid = c->pt_job_slot;
    JobRec& J = c->jobs[jid];
    J = JobRec();
    J.scene = sid; J.cam = *cam; J.W = W; J.H = H; J.p = *params; J.p.gi = gi_check;
    // previous frame -> output for the tile, then fold this frame in place
    if (!reset)
        HIPCK(c, hipMemcpy2DAsync(out_img + 4 * ((size_t)y0 * W + x0), (size_t)W * 16, in_img + 4 * ((size_t)y0 * W + x0),
                                  (size_t)W * 16, (size_t)w * 16, (size_t)h, hipMemcpyDeviceToDevice, c->stream));
    std::vector<uint32_t> seeds(1, rand);
    int r = render_rect(c, jid, (float4*)out_img, x0, y0, w, h, seeds, reset != 0, DevCapture{});
    if (r == 0) HIPCK(c, hipStreamSynchronize(c->stream));
    return r;
}

int yn_tonemap_op(yn_ctx* c, const float* in_img, float* out_img, int32_t W, int32_t H, int32_t op)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (!in_img || !out_img || W <= 0 || H <= 0) return fail(c, YN_ERR_INVALID, "bad tonemap arguments");
    if (op != YN_TONEMAP_REINHARD && op != YN_TONEMAP_HABLE) return fail(c, YN_ERR_INVALID, "unknown tonemap operator");
    launch_tonemap((const float4*)in_img, (float4*)out_img, W * H, op, c->stream);
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int yn_tonemap(yn_ctx* c, const float* in_img, float* out_img, int32_t W, int32_t H)
{
    return yn_tonemap_op(c, in_img, out_img, W, H, YN_TONEMAP_REINHARD);
}

int yn_job_set_tonemap(yn_ctx* c, int32_t jid, int32_t op)
{
    if (check_ctx(c) || check_job(c, jid)) return YN_ERR_NOT_FOUND;
    if (op != YN_TONEMAP_REINHARD && op != YN_TONEMAP_HABLE) return fail(c, YN_ERR_INVALID, "unknown tonemap operator");
    c->jobs[jid].tonemap_op = op;
    return 0;
}

int yn_trace(yn_ctx* c, int32_t sid, const float* rays, int32_t n, int32_t* tri, int32_t* light, float* t)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (n < 0 || (n > 0 && (!rays || !tri || !light || !t))) return fail(c, YN_ERR_INVALID, "bad trace arguments");
    if (n == 0) return 0;
    if (n > c->cap_paths) return fail(c, YN_ERR_INVALID, "more rays than the path capacity");
    if (c->strm.active)
        if (int r = stream_drain(c)) return r;
    hipStream_t s = c->stream;
    std::vector<int> idx[2];
    for (int i = 0; i < n; ++i) idx[rays[8 * (size_t)i + 7] != 0.0f ? 1 : 0].push_back(i);
    float* d_rays = nullptr;
    int32_t* d_light = nullptr;
    int r = dalloc(c, &d_rays, (size_t)n * 8);
    if (r) return r;
    r = dalloc(c, &d_light, (size_t)n);
    if (r) { hipFree(d_rays); return r; }
    std::vector<float> sub;
    std::vector<float4> hits(n);
    std::vector<int32_t> lights(n), occ(n);
    const SceneRec& S = c->scenes[sid];
    for (int any = 0; any < 2; ++any) {
        int m = (int)idx[any].size();
        if (m == 0) continue;
        sub.resize((size_t)m * 8);
        for (int k = 0; k < m; ++k) std::memcpy(&sub[8 * (size_t)k], rays + 8 * (size_t)idx[any][k], 32);
        HIPCK(c, hipMemcpyAsync(d_rays, sub.data(), sub.size() * 4, hipMemcpyHostToDevice, s));
        HIPCK(c, hipMemsetAsync(c->ctr, 0, sizeof(DevCounters), s));
        RayQueue q = c->E[0];
        q.count = &c->ctr->q[0][0];
        launch_prepare(c->d_scenes + sid, d_rays, m, q, d_light, s);
        TraceArgs ta{};
        ta.scenes = c->d_scenes;
        ta.jobs = c->d_jobs;
        ta.path_job = c->P.job;
        ta.single_scene = sid;
        ta.chunk = c->cfg.trace_chunk;
        ta.ctr = c->ctr;
        ta.q = q;
        ta.hit_out = c->P.hit;
        ta.occ_out = any ? c->P.occ : nullptr;
        ta.work = &c->ctr->work[0][0][0];
        ta.n_slices = c->cfg.trace_xcd;
        ta.refill = c->cfg.trace_refill > 0 ? c->cfg.trace_refill : 0;
        const int lds = lds_for(c, S.info.node_bytes);
        int grid = c->cfg.trace_grid > 0 ? c->cfg.trace_grid
                                         : c->n_cu * std::max(1, trace_resident_waves(lds) * 256 / c->cfg.trace_block);
        if (int rr = ensure_overflow(c, S.info.max_depth - trace_lds_stack(lds), (int64_t)grid * c->cfg.trace_block))
            return rr;
        ta.ovf_ref = c->ovf_ref;
        ta.ovf_t = c->ovf_t;
        ta.ovf_depth = c->ovf_depth;
        if (launch_trace(ta, any != 0, S.info.width, S.info.max_depth, c->cfg.trace_block, grid, false, lds, s) != 0)
            return fail(c, YN_ERR_INVALID, "BVH too deep");
        HIPCK(c, hipGetLastError());
        HIPCK(c, hipMemcpyAsync(hits.data(), c->P.hit, (size_t)m * 16, hipMemcpyDeviceToHost, s));
        HIPCK(c, hipMemcpyAsync(lights.data(), d_light, (size_t)m * 4, hipMemcpyDeviceToHost, s));
        HIPCK(c, hipMemcpyAsync(c->h_ctr, c->ctr, kCtrHostBytes, hipMemcpyDeviceToHost, s));
        HIPCK(c, hipStreamSynchronize(s));
        if (c->h_ctr->q[7][0]) return fail(c, YN_ERR_INVALID, "traversal stack overflow");
        for (int k = 0; k < m; ++k) {
            int i = idx[any][k];
            int tr;
            std::memcpy(&tr, &hits[k].y, 4);
            tri[i] = tr;
            t[i] = hits[k].x;
            // a triangle nearer than the quad replaces the light hit (udpt.cl:386)
            light[i] = tr >= 0 ? -1 : lights[k];
        }
    }
    hipFree(d_rays);
    hipFree(d_light);
    return 0;
}

int yn_trace_noalloc(yn_ctx* c, int32_t sid, const float* rays, int32_t n, int32_t* tri, int32_t* light, float* t)
{
    if (check_ctx(c)) return YN_ERR_INVALID;
    if (check_scene(c, sid)) return YN_ERR_NOT_FOUND;
    if (n < 0 || (n > 0 && (!rays || !tri || !light || !t))) return fail(c, YN_ERR_INVALID, "bad trace arguments");
    if (n > YN_TRACE_NOALLOC_MAX_RAYS)
        return fail(c, YN_ERR_INVALID, "yn_trace_noalloc takes at most " + std::to_string(YN_TRACE_NOALLOC_MAX_RAYS) +
                                           " rays per call, got " + std::to_string(n));
    if (n == 0) return 0;
    const SceneRec& S = c->scenes[sid];
    hipStream_t s = c->stream;
    TraceArgs ta{};
    ta.scenes = c->d_scenes;
    ta.jobs = c->d_jobs;
    ta.path_job = c->P.job;
    ta.single_scene = sid;
    ta.chunk = c->cfg.trace_chunk;
    ta.n_slices = c->cfg.trace_xcd;
    ta.refill = c->cfg.trace_refill > 0 ? c->cfg.trace_refill : 0;
    // the no-allocation entry keeps 16 LDS entries unless configured
    const int na_lds = c->cfg.lds_stack > 0 ? c->cfg.lds_stack : 16;
    if (trace_noalloc_enqueue(c->na_scratch, ta, c->d_scenes + sid, S.info.width, na_lds, c->cfg.trace_block,
                              rays, n, tri, light, t, s) != 0)
        return fail(c, YN_ERR_INVALID, "trace launch refused");
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipMemcpyAsync(&c->h_ctr->q[7][0], &trace_noalloc_counters(c->na_scratch)->q[7][0], sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
    HIPCK(c, hipStreamSynchronize(s));
    if (c->h_ctr->q[7][0])
        return fail(c, YN_ERR_INVALID, "traversal deeper than the no-allocation stack (" +
                                           std::to_string(trace_lds_stack(na_lds) + kTraceNoAllocOvfDepth) +
                                           " entries)");
    return 0;
}

}  // extern "C"
