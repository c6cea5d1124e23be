// Batches of independent runs (3-D and P-SV contexts) advanced together by
// one captured graph: the hooks each context file exposes to
// swpc_batch.hip.
#pragma once
#include "swpc_common.h"

#include <vector>

namespace swpc {

// What a batch needs of one member context.
struct BatchMember {
    hipStream_t stream = nullptr;   // the member's own stream (its kernels launch there)
    int *dit = nullptr;             // the member's device step number
    unsigned gen = 0;               // changes whenever the member's captured steps would change
};

int batch_member3(swpc3d_ctx *c, BatchMember &m);   // fails unless the member can step
int batch_check3(swpc3d_ctx *c, int it0, int n);    // its sources cover steps it0..it0+n-1
int batch_step3(swpc3d_ctx *c);                     // enqueue one step on its stream
void batch_done3(swpc3d_ctx *c, int it0, int it1);  // host bookkeeping after steps it0..it1
void batch_abort3(swpc3d_ctx *c);                   // captured steps that will not run: forget their effects

int batch_member2(swpc2d_ctx *c, BatchMember &m);
int batch_check2(swpc2d_ctx *c, int it0, int n);
int batch_step2(swpc2d_ctx *c);
void batch_done2(swpc2d_ctx *c, int it0, int it1);

// A group of P-SV members stepped together, one launch per phase for all
// of them (FP32, fused steps, no snapshots, sources inside the fused
// kernels); results equal each member's own steps bit for bit.
struct Group2;
bool group2_eligible(const swpc2d_ctx *c);
int group2_build(const std::vector<swpc2d_ctx *> &cs, Group2 **out, long long &bytes);
int group2_step(Group2 *g, hipStream_t s);   // one step of every member, enqueued on s
void group2_free(Group2 *g);

// A group of 3-D members stepped together, one launch per phase for all of
// them (FP32 with a density, the one-pass stress and z-marching velocity
// forms, no snapshots, no body force); results equal each member's own
// steps bit for bit.
struct Group3;
bool group3_eligible(const swpc3d_ctx *c);
long long group3_cells(const swpc3d_ctx *c);   // nx * ny * nz
// 3-D members above this size stay branches (about 128^3 cells)
constexpr long long GROUP3_MAX_CELLS = 1LL << 21;
int group3_build(const std::vector<swpc3d_ctx *> &cs, Group3 **out, long long &bytes);
int group3_step(Group3 *g, hipStream_t s);
void group3_free(Group3 *g);

}  // namespace swpc
