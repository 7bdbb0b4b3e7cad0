// step.cpp -- one launch of the search kernel k_teams, as the plain solve, a batch
// sub-batch or a serving step issues it.
//
// Everything a launch reads besides the search state -- its shape, the per-class query
// counters, the groups' TeamVars, the query descriptors and the results reported for
// queries it does not reach -- forms one I/O block (StepHeader first). The host builds
// the block in pinned memory and it reaches the device in one copy; the kernel's
// arguments (the block and the insert buffers) and its grid stay the same
// from launch to launch. A launch is five operations on the context's stream: copy in,
// event, cooperative kernel, event, copy of the results out. With UA_FLAG_GRAPH they are
// an instantiated hipGraph per (mode, problem set, block count), replayed with only the two copy
// sizes updated; otherwise they are issued call by call. Where the runtime refuses a
// part of the graph, the context notes why and issues the calls from then on.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "internal.hpp"

using namespace ua;
using namespace uart;

namespace uart {

void StepGraph::release() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    *this = StepGraph{};
}

void release_steps(ua_ctx *c) {
    for (auto &g : c->graph) g.release();
    if (c->h_io) (void)hipHostFree(c->h_io);
    c->h_io = nullptr;
    c->h_io_n = 0;
    c->io.release();
    c->segs.release();
}

namespace {

struct Layout {
    size_t tv_off, qd_off, qr_off, end;
};

Layout layout_of(int groups, int nq) {
    Layout l;
    l.tv_off = align256(sizeof(StepHeader));
    l.qd_off = l.tv_off + align256((size_t)groups * sizeof(TeamVars));
    l.qr_off = l.qd_off + align256((size_t)nq * sizeof(QueryDesc));
    l.end = l.qr_off + (size_t)nq * sizeof(QueryResult);
    return l;
}

// Builds and instantiates the graph of one launch for the current addresses. The copy
// sizes set here are placeholders: every replay sets its own.
hipError_t build_graph(ua_ctx *c, StepGraph &sg, int mode, int set, int blocks) {
    sg.release();
    char *io = c->io.as<char>();
    uint64_t *segs = c->segs.as<uint64_t>();
    hipError_t e = hipGraphCreate(&sg.g, 0);
    if (e) return e;
    e = hipGraphAddMemcpyNode1D(&sg.h2d, sg.g, nullptr, 0, io, c->h_io, sizeof(StepHeader), hipMemcpyHostToDevice);
    hipGraphNode_t ev0 = nullptr, ev1 = nullptr;
    if (!e) e = hipGraphAddEventRecordNode(&ev0, sg.g, &sg.h2d, 1, c->ev0);
    void *args[] = {(void *)&io, (void *)&segs};
    hipKernelNodeParams kp{};
    kp.func = const_cast<void *>(teams_fn(mode, set));
    kp.gridDim = dim3(blocks);
    kp.blockDim = dim3(kBlockWaves * 64);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    if (!e) e = hipGraphAddKernelNode(&sg.kern, sg.g, &ev0, 1, &kp);
    if (!e) {
        hipLaunchAttributeValue v{};
        v.cooperative = 1;
        e = hipGraphKernelNodeSetAttribute(sg.kern, hipLaunchAttributeCooperative, &v);
    }
    if (!e) e = hipGraphAddEventRecordNode(&ev1, sg.g, &sg.kern, 1, c->ev1);
    if (!e) e = hipGraphAddMemcpyNode1D(&sg.d2h, sg.g, &ev1, 1, c->h_io, io, sizeof(StepHeader),
                                        hipMemcpyDeviceToHost);
    if (!e) e = hipGraphInstantiate(&sg.x, sg.g, nullptr, nullptr, 0);
    if (e) {
        sg.release();
        return e;
    }
    sg.io = io;
    sg.host = c->h_io;
    sg.segs = segs;
    sg.mode = mode;
    sg.set = set;
    sg.blocks = blocks;
    return hipSuccess;
}

// Points the graph's two copies at this launch's extent of the I/O block.
hipError_t set_copies(ua_ctx *c, StepGraph &sg, const Layout &l) {
    char *io = c->io.as<char>();
    hipError_t e = hipGraphExecMemcpyNodeSetParams1D(sg.x, sg.h2d, io, c->h_io, l.end, hipMemcpyHostToDevice);
    if (!e)
        e = hipGraphExecMemcpyNodeSetParams1D(sg.x, sg.d2h, c->h_io + l.qr_off, io + l.qr_off, l.end - l.qr_off,
                                              hipMemcpyDeviceToHost);
    return e;
}

// Replays the launch's graph (building it first if needed). Returns false, with the
// reason in c->graph_err, if the runtime refuses to build or update it: nothing of the
// launch has run then.
bool try_graph(ua_ctx *c, int mode, int set, int blocks, const Layout &l, hipError_t *launch_err) {
    const int ng = (int)(sizeof c->graph / sizeof c->graph[0]);
    StepGraph *pick = nullptr;
    for (auto &g : c->graph)
        if (g.x && g.mode == mode && g.set == set && g.blocks == blocks) pick = &g;
    if (!pick) {
        pick = &c->graph[c->graph_next];
        c->graph_next = (c->graph_next + 1) % ng;
    }
    StepGraph &sg = *pick;
    const bool same = sg.x && sg.mode == mode && sg.set == set && sg.io == c->io.p && sg.host == c->h_io && sg.segs == c->segs.p &&
                      sg.blocks == blocks;
    hipError_t e = hipSuccess;
    if (!same) e = build_graph(c, sg, mode, set, blocks);
    if (!e) {
        e = set_copies(c, sg, l);
        if (e && same) {  // an update the instantiated graph does not take: build it afresh once
            e = build_graph(c, sg, mode, set, blocks);
            if (!e) e = set_copies(c, sg, l);
        }
    }
    if (e) {
        c->graph_err = std::string("launch graph refused: ") + hipGetErrorString(e);
        sg.release();
        (void)hipGetLastError();
        return false;
    }
    *launch_err = hipGraphLaunch(sg.x, c->st);
    return true;
}

}  // namespace

int run_step(ua_ctx *c, int mode, int set, int blocks, const LaunchShape &shape, const QueryDesc *descs,
             const QueryResult *init, int nq, QueryResult *out) {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    if (mode < 0 || mode > 2 || set < 0 || set >= kSets)
        return fail(c, UA_ERR_INVALID, "bad launch mode %d / set %d", mode, set);
    int per_cu = 0;
    HIPCHK(c, teams_max_blocks(&per_cu));
    const int resident = std::max(1, c->num_cu * std::max(1, per_cu));
    int groups = 0;
    size_t seg_words = 0;
    for (int k = 0; k < kClasses; ++k) {
        groups += shape.groups[k];
        seg_words += (size_t)shape.groups[k] * 2 * (size_t)shape.segcap[k];
    }
    // insert buffers of every block of the device, at most 8 successors per pop: sized
    // once, so that the kernel's arguments never change
    const size_t seg_bytes = (size_t)resident * 2 * kBlockWaves * 64 * 8 * 8;
    if (blocks < shape_blocks(shape) || blocks > resident || seg_words * 8 > seg_bytes || nq < 0)
        return fail(c, UA_ERR_INVALID, "launch of %d blocks does not fit its shape or the device", blocks);
    HIPCHK(c, c->segs.ensure(seg_bytes));

    const Layout l = layout_of(groups, nq);
    if (l.end > UINT32_MAX) return fail(c, UA_ERR_INVALID, "launch of %d queries is too large", nq);
    if (l.end > c->io.n) HIPCHK(c, c->io.ensure(std::max(l.end, 2 * c->io.n)));
    if (l.end > c->h_io_n) {
        size_t n = std::max(l.end, 2 * c->h_io_n);
        if (c->h_io) HIPCHK(c, hipHostFree(c->h_io));
        c->h_io = nullptr;
        c->h_io_n = 0;
        HIPCHK(c, hipHostMalloc((void **)&c->h_io, n, hipHostMallocDefault));
        c->h_io_n = n;
    }
    // the block: header, zero TeamVars, descriptors, initial results
    StepHeader h{};
    h.shape = shape;
    h.tv_off = (uint32_t)l.tv_off;
    h.qd_off = (uint32_t)l.qd_off;
    h.qr_off = (uint32_t)l.qr_off;
    h.nq = (uint32_t)nq;
    std::memset(c->h_io, 0, l.qd_off);
    std::memcpy(c->h_io, &h, sizeof h);
    if (nq) std::memcpy(c->h_io + l.qd_off, descs, (size_t)nq * sizeof(QueryDesc));
    if (init) std::memcpy(c->h_io + l.qr_off, init, (size_t)nq * sizeof(QueryResult));
    else std::memset(c->h_io + l.qr_off, 0, (size_t)nq * sizeof(QueryResult));

    bool replayed = false;
    if (c->graphs) {
        hipError_t le = hipSuccess;
        replayed = try_graph(c, mode, set, blocks, l, &le);
        if (!replayed) c->graphs = false;
        else HIPCHK(c, le);
    }
    if (!replayed) {
        char *io = c->io.as<char>();
        HIPCHK(c, hipMemcpyAsync(io, c->h_io, l.end, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipEventRecord(c->ev0, c->st));
        HIPCHK(c, launch_teams(io, blocks, c->segs.as<uint64_t>(), mode, set, c->st));
        HIPCHK(c, hipEventRecord(c->ev1, c->st));
        HIPCHK(c, hipMemcpyAsync(c->h_io + l.qr_off, io + l.qr_off, l.end - l.qr_off, hipMemcpyDeviceToHost, c->st));
    }
    if (mode != 1) {
        // A serving step ends within its time budget, and most plain solves take a few
        // milliseconds: poll the stream rather than let the thread sleep until the runtime
        // wakes it, which adds latency to every launch. After 50 ms of polling (a serving
        // step never runs that long) the thread waits normally.
        const auto spin_until = Clock::now() + std::chrono::milliseconds(50);
        hipError_t q;
        while ((q = hipStreamQuery(c->st)) == hipErrorNotReady && Clock::now() < spin_until) {
        }
        if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(c, q);
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (nq) std::memcpy(out, c->h_io + l.qr_off, (size_t)nq * sizeof(QueryResult));
    float kms = 0;
    HIPCHK(c, hipEventElapsedTime(&kms, c->ev0, c->ev1));
    c->last_kernel_ms = kms;
    (replayed ? c->graph_launches : c->direct_launches)++;
    c->step_kernel_ms += kms;
    c->step_wall_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    return UA_OK;
}

}  // namespace uart

extern "C" int ua_launch_stats(ua_ctx *c, int64_t *graph_launches, int64_t *direct_launches, double *wall_ms,
                               double *kernel_ms) {
    if (!c) return UA_ERR_INVALID;
    if (graph_launches) *graph_launches = c->graph_launches;
    if (direct_launches) *direct_launches = c->direct_launches;
    if (wall_ms) *wall_ms = c->step_wall_ms;
    if (kernel_ms) *kernel_ms = c->step_kernel_ms;
    return UA_OK;
}

extern "C" const char *ua_launch_graph_error(ua_ctx *c) { return c ? c->graph_err.c_str() : "no context"; }
