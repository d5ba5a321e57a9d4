// The two halves of one call of the kernel chain, for callers that keep
// several calls in flight (the serving loop): chain_enqueue queues
// everything on the context's stream and returns; chain_collect, once
// chain_done, reads the results.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "atg.h"

namespace atg {

int chain_enqueue(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device);
bool chain_done(atg_ctx *ctx);
// ATG_OK, 1 if a frame ran out of point capacity, or an error.
int chain_collect(atg_ctx *ctx);
// enqueue + collect
int run_chain(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device);
size_t batch_arena_bytes(const atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device);
int ctx_reserve(atg_ctx *ctx, size_t bytes, int32_t max_frames);
hipStream_t ctx_stream(atg_ctx *ctx);
const std::vector<at_detection> &ctx_dets(const atg_ctx *ctx, int32_t i);
int ctx_frame_status(const atg_ctx *ctx, int32_t i);
int64_t ctx_nquads(const atg_ctx *ctx, int32_t i);
size_t ctx_result_bytes(int32_t max_frames);

}  // namespace atg
