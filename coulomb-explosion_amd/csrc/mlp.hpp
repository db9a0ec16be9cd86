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

// Layout options of the fused inference kernel (the defaults: x [B][w_in] and
// y [B][w_out], row-major, B rows).
struct FusedIo {
  const uint32_t* rows_dev = nullptr;  // the rows are min(B, *rows_dev), read on the device
  const uint32_t* idx = nullptr;       // row r of y is row idx[r] of x
  int64_t ldx = 0;  // > 0: x column-major, input k of row i at x[k ldx + i]
  int64_t ldy = 0;  // > 0: y column-major, output n of row r at y[n ldy + r]
};

// y [B][w_out] = net(x [B][w_in]) for device buffers, on stream s (in chunks
// of the network's batch workspace).  A count left on the device by another
// kernel, a row index and column-major layouts (FusedIo) need the fused kernel.
int mlp_forward_device(ce_mlp* m, const float* x, int64_t B, float* y, hipStream_t s,
                       const FusedIo& io = FusedIo{});
// Whether inference runs through the fused kernel (and so can take FusedIo).
bool mlp_fused_ready(const ce_mlp* m);

}  // namespace ce
