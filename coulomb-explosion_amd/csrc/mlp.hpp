// Internal interface of the closure network (mlp.hip) for the closure pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "coulomb_explosion.h"
#include "hip_check.hpp"

namespace ce {

// Layer widths: mlp_width(m, 0) inputs ... mlp_width(m, mlp_layers(m)) outputs.
int mlp_width(const ce_mlp* m, int layer);
int mlp_layers(const ce_mlp* m);
// y [B][w_out] = net(x [B][w_in]) for device buffers, on stream s (in chunks
// of the network's batch workspace).  Fused kernel only: with rows_dev the rows
// are min(B, *rows_dev), read on the device (a count another kernel left there,
// without a host round trip); with idx, row r of y is row idx[r] of x.
int mlp_forward_device(ce_mlp* m, const float* x, int64_t B, float* y, hipStream_t s,
                       const uint32_t* rows_dev = nullptr, const uint32_t* idx = nullptr);
// Whether inference runs through the fused kernel (and so can take rows_dev).
bool mlp_fused_ready(const ce_mlp* m);

}  // namespace ce
