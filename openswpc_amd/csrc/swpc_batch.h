// Batches of independent runs (3-D and P-SV contexts) advanced together by
// one captured graph: the hooks each context file exposes to
// swpc_batch.hip.
#pragma once
#include "swpc_common.h"

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

int batch_member2(swpc2d_ctx *c, BatchMember &m);
int batch_check2(swpc2d_ctx *c, int it0, int n);
int batch_step2(swpc2d_ctx *c);
void batch_done2(swpc2d_ctx *c, int it0, int it1);

}  // namespace swpc
