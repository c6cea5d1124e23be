// Batches of independent runs on one GPU: many 3-D and P-SV contexts of
// uneven size advanced together, graph_k steps of every member captured
// once into one HIP graph (one branch per member, joined at the end) and
// replayed with one launch per graph_k steps.  Each member keeps its own
// stream, step counter and state, so results equal the members' own runs;
// the graph removes the per-kernel launch cost of the many small grids
// and lets the runtime overlap the branches.  With grouping on (the
// default) the members that qualify are not branches: one group advances
// all FP32 P-SV members (group2_eligible) with one launch per phase (three
// per step), and another all FP32 3-D members (group3_eligible) with seven
// launches per step, the workgroups of every member's kernels of a phase
// end to end.  The P-SV group runs on the batch stream and the 3-D group
// on a stream of its own, beside the branches.
#include "swpc_batch.h"

#include <algorithm>
#include <vector>

namespace swpc {

__global__ void k_batch_set_steps(int *const *__restrict__ dits, int n, int it)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) *dits[q] = it;
}

}  // namespace swpc

struct swpc_batch {
    struct M {
        swpc3d_ctx *c3 = nullptr;
        swpc2d_ctx *c2 = nullptr;
        swpc::BatchMember m;
        hipEvent_t ev = nullptr;
        unsigned gen_captured = 0;
    };
    std::vector<M> mem;
    hipStream_t stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_k = 0;             // steps in the captured graph
    swpc::DBuf dits;             // device array of the members' step counters
    long long bytes = 0;
    int grouping = 3;            // groups: bit 0 P-SV, bit 1 3-D members (swpc_batch_set_grouping)
    swpc::Group2 *group = nullptr;   // the P-SV group of the captured graph
    std::vector<swpc::Group3 *> groups3;   // its 3-D group (none or one)
    hipStream_t stream3 = nullptr;   // the 3-D group's stream
    hipEvent_t ev3 = nullptr;
    long long group_bytes = 0;
};

namespace swpc {

static void batch_drop(swpc_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_k = 0;
    group2_free(b->group);
    b->group = nullptr;
    for (auto *g : b->groups3) group3_free(g);
    b->groups3.clear();
    b->bytes -= b->group_bytes;
    b->group_bytes = 0;
}

static int member_info(swpc_batch::M &x)
{
    return x.c3 ? batch_member3(x.c3, x.m) : batch_member2(x.c2, x.m);
}

static int member_step(swpc_batch::M &x) { return x.c3 ? batch_step3(x.c3) : batch_step2(x.c2); }

// Capture k steps of every member: the batch stream forks to each member's
// stream (an event each), the member enqueues its k steps there, and the
// batch stream joins them all.
static int batch_capture(swpc_batch *b, int k)
{
    batch_drop(b);
    // the group: every qualifying P-SV member, stepped on the batch stream
    std::vector<swpc2d_ctx *> gc;
    std::vector<char> grouped(b->mem.size(), 0);
    if (b->grouping & 1)
        for (size_t q = 0; q < b->mem.size(); q++)
            if (b->mem[q].c2 && group2_eligible(b->mem[q].c2)) {
                gc.push_back(b->mem[q].c2);
                grouped[q] = 1;
            }
    if (gc.size() >= 2) {
        long long gb = 0;
        if (int r = group2_build(gc, &b->group, gb)) return r;
        b->group_bytes += gb;
        b->bytes += gb;
    } else {
        for (size_t q = 0; q < b->mem.size(); q++)
            if (b->mem[q].c2) grouped[q] = 0;
    }
    // the 3-D group: every qualifying 3-D member of at most GROUP3_MAX_CELLS
    // cells (larger ones fill the GPU with their own launches, whose tuned
    // tile shapes the group's fixed ones would give up)
    if (b->grouping & 2) {
        std::vector<swpc3d_ctx *> g3;
        std::vector<size_t> idx;
        for (size_t q = 0; q < b->mem.size(); q++)
            if (b->mem[q].c3 && group3_eligible(b->mem[q].c3) && group3_cells(b->mem[q].c3) <= GROUP3_MAX_CELLS) {
                g3.push_back(b->mem[q].c3);
                idx.push_back(q);
            }
        if (g3.size() >= 2) {
            Group3 *g = nullptr;
            long long gb = 0;
            if (int r = group3_build(g3, &g, gb)) return r;
            b->groups3.push_back(g);
            b->group_bytes += gb;
            b->bytes += gb;
            for (size_t q : idx) grouped[q] = 1;
        }
    }
    SWPC_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int r = SWPC_OK;
    hipError_t e = hipEventRecord(b->ev_fork, b->stream);
    for (size_t q = 0; q < b->mem.size(); q++) {
        auto &x = b->mem[q];
        if (r != SWPC_OK || e != hipSuccess) break;
        if (grouped[q]) continue;
        e = hipStreamWaitEvent(x.m.stream, b->ev_fork, 0);
        for (int s = 0; s < k && r == SWPC_OK && e == hipSuccess; s++) r = member_step(x);
        if (r == SWPC_OK && e == hipSuccess) e = hipEventRecord(x.ev, x.m.stream);
    }
    // the 3-D groups run on their own stream and the P-SV group on the batch
    // stream, beside the branches; all join the batch stream at the end
    if (!b->groups3.empty() && e == hipSuccess) {
        e = hipStreamWaitEvent(b->stream3, b->ev_fork, 0);
        for (int s = 0; s < k && r == SWPC_OK && e == hipSuccess; s++)
            for (auto *g : b->groups3)
                if (r == SWPC_OK) r = group3_step(g, b->stream3);
        if (r == SWPC_OK && e == hipSuccess) e = hipEventRecord(b->ev3, b->stream3);
    }
    if (b->group)
        for (int s = 0; s < k && r == SWPC_OK && e == hipSuccess; s++) r = group2_step(b->group, b->stream);
    if (!b->groups3.empty() && r == SWPC_OK && e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev3, 0);
    for (size_t q = 0; q < b->mem.size() && r == SWPC_OK && e == hipSuccess; q++)
        if (!grouped[q]) e = hipStreamWaitEvent(b->stream, b->mem[q].ev, 0);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->stream, &g);
    if (r != SWPC_OK || e != hipSuccess || ee != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        for (auto &x : b->mem)
            if (x.c3) batch_abort3(x.c3);
        if (r != SWPC_OK) return r;
        return fail(SWPC_EHIP, "batch graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : ee));
    }
    hipGraphExec_t ge = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        for (auto &x : b->mem)
            if (x.c3) batch_abort3(x.c3);
        return fail(SWPC_EHIP, "batch graph instantiation failed: %s", hipGetErrorString(ei));
    }
    b->graph = g;
    b->gexec = ge;
    b->graph_k = k;
    for (auto &x : b->mem) x.gen_captured = x.m.gen;
    return SWPC_OK;
}

}  // namespace swpc

using namespace swpc;

extern "C" {

int swpc_batch_create(swpc_batch **out)
{
    clear_error();
    if (!out) return fail(SWPC_EINVAL, "null argument");
    *out = nullptr;
    auto *b = new swpc_batch();
    if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->stream3) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev3, hipEventDisableTiming) != hipSuccess) {
        swpc_batch_destroy(b);
        return fail(SWPC_EHIP, "batch stream/event creation failed");
    }
    *out = b;
    return SWPC_OK;
}

int swpc_batch_destroy(swpc_batch *b)
{
    if (!b) return SWPC_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop(b);
    for (auto &x : b->mem)
        if (x.ev) (void)hipEventDestroy(x.ev);
    dfree(b->dits);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->ev3) (void)hipEventDestroy(b->ev3);
    if (b->stream3) (void)hipStreamDestroy(b->stream3);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SWPC_OK;
}

static int batch_add(swpc_batch *b, swpc3d_ctx *c3, swpc2d_ctx *c2)
{
    clear_error();
    if (!b || (!c3 && !c2)) return fail(SWPC_EINVAL, "null argument");
    for (auto &x : b->mem)
        if ((c3 && x.c3 == c3) || (c2 && x.c2 == c2)) return fail(SWPC_EINVAL, "context already in the batch");
    swpc_batch::M x;
    x.c3 = c3;
    x.c2 = c2;
    if (int r = member_info(x)) return r;
    // the step-counter table with the new member goes into a fresh buffer
    // first; the batch takes the member only once the table is on the device
    std::vector<int *> d;
    for (auto &y : b->mem) d.push_back(y.m.dit);
    d.push_back(x.m.dit);
    DBuf nd;
    long long nbytes = 0;
    if (int r = dalloc(nd, d.size() * sizeof(int *), nbytes)) {
        dfree(nd);
        return r;
    }
    hipError_t e = hipMemcpy(nd.p, d.data(), d.size() * sizeof(int *), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x.ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        dfree(nd);
        return fail(SWPC_EHIP, "batch_add: %s", hipGetErrorString(e));
    }
    batch_drop(b);
    b->bytes -= (long long)b->dits.bytes;
    dfree(b->dits);
    b->dits = nd;
    nd = DBuf();
    b->bytes += nbytes;
    b->mem.push_back(x);
    return SWPC_OK;
}

int swpc_batch_add3d(swpc_batch *b, swpc3d_ctx *c) { return batch_add(b, c, nullptr); }
int swpc_batch_add2d(swpc_batch *b, swpc2d_ctx *c) { return batch_add(b, nullptr, c); }

int swpc_batch_run(swpc_batch *b, int it0, int nsteps, int graph_k)
{
    clear_error();
    if (!b || nsteps < 0 || graph_k < 0) return fail(SWPC_EINVAL, "bad argument");
    if (b->mem.empty() || nsteps == 0) return SWPC_OK;
    if (!b->dits.p) return fail(SWPC_EINVAL, "batch has no step-counter table (a failed add?)");
    bool stale = false;
    for (auto &x : b->mem) {
        if (int r = member_info(x)) return r;
        if (int r = x.c3 ? batch_check3(x.c3, it0, nsteps) : batch_check2(x.c2, it0, nsteps)) return r;
        stale |= x.m.gen != x.gen_captured;
    }
    // order after every member's earlier work, then set all step counters
    for (auto &x : b->mem) {
        SWPC_HIP(hipEventRecord(x.ev, x.m.stream));
        SWPC_HIP(hipStreamWaitEvent(b->stream, x.ev, 0));
    }
    const int n = (int)b->mem.size();
    hipLaunchKernelGGL(k_batch_set_steps, dim3((n + 63) / 64), dim3(64), 0, b->stream, (int *const *)b->dits.p, n, it0);
    SWPC_CHECK_LAUNCH();
    int done = 0;
    if (graph_k > 0 && nsteps >= graph_k) {
        if (!b->gexec || stale || b->graph_k != graph_k)
            if (int r = batch_capture(b, graph_k)) return r;
        for (; done + graph_k <= nsteps; done += graph_k) SWPC_HIP(hipGraphLaunch(b->gexec, b->stream));
    }
    // the remaining steps go member by member on their own streams
    SWPC_HIP(hipEventRecord(b->ev_join, b->stream));
    for (auto &x : b->mem) {
        SWPC_HIP(hipStreamWaitEvent(x.m.stream, b->ev_join, 0));
        for (int q = done; q < nsteps; q++)
            if (int r = member_step(x)) return r;
        if (x.c3) batch_done3(x.c3, it0, it0 + nsteps - 1);
        else batch_done2(x.c2, it0, it0 + nsteps - 1);
    }
    return SWPC_OK;
}

int swpc_batch_set_grouping(swpc_batch *b, int mode)
{
    clear_error();
    if (!b) return fail(SWPC_EINVAL, "null batch");
    if (mode < 0 || mode > 3) return fail(SWPC_EINVAL, "grouping mode %d: a bit mask of 1 (P-SV) and 2 (3-D)", mode);
    batch_drop(b);
    b->grouping = mode;
    return SWPC_OK;
}

int swpc_batch_sync(swpc_batch *b)
{
    clear_error();
    if (!b) return fail(SWPC_EINVAL, "null batch");
    SWPC_HIP(hipStreamSynchronize(b->stream));
    for (auto &x : b->mem) SWPC_HIP(hipStreamSynchronize(x.m.stream));
    return SWPC_OK;
}

}  // extern "C"
