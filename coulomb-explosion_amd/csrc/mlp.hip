// Closure network: a small MLP (SiLU hidden layers) trained and evaluated on the
// GPU with hand-written BF16 MFMA GEMMs accumulating in FP32 (gfx950
// v_mfma_f32_32x32x16_bf16), FP32 master weights and Adam.
//
// Layouts.  Activations are row-major [batch][features] in bf16.  Layer l has
// W_l [out][in] (bf16 copy of the FP32 master, refreshed after every update) and
// bias b_l [out] (FP32).  Parameters are stored W_0, b_0, W_1, b_1, ... in one
// FP32 vector.
//   forward   Z = X W^T + b,  Y = silu(Z)     (A = X, B(k=i, n=o) = W[o][i])
//   backward  dZ = dY * silu'(Z)
//             dX = dZ W                        (A = dZ, B(k=o, n=i) = W[o][i])
//             dW = dZ^T X,  db = sum_b dZ      (A(m=o, k=b) = dZ[b][o], B(k=b, n=i) = X[b][i])
// The GEMM takes general element strides for A and B and stages 64 x 32 / 32 x 64
// tiles in LDS with k contiguous, so every product above uses the same kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "coulomb_explosion.h"
#include "mlp.hpp"

namespace ce {
int set_error(int code, const std::string& msg);
}

namespace {

// Compile-time loop: f(std::integral_constant<int, i>) for i in [B, E).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even with the gfx950 conversion instruction (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}
// sigmoid with the hardware exp2 and reciprocal (about 1 ulp each; the result is
// rounded to bf16 or feeds a bf16 product)
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

constexpr int kTM = 128, kTN = 128, kTK = 32, kLdsK = kTK + 8;  // LDS row: 80 bytes
enum Epilogue : int { kEpiStoreF32 = 0, kEpiBiasSiluBf16 = 1, kEpiBiasF32 = 2, kEpiAtomicF32 = 3 };

struct GemmArgs {
  const uint16_t* A;  // bf16 bits
  const uint16_t* B;
  int64_t sAm, sAk, sBk, sBn;  // element strides; per operand one of them is 1
  int M, N, K;
  int k_split;                 // K range of one blockIdx.z (multiple of kTK)
  const float* bias;           // [N] or nullptr
  float* C32;                  // fp32 out [M][ldc] (pre-activation / plain / atomic sum)
  uint16_t* C16;               // bf16 out [M][ldc]
  int64_t ldc;
  int epilogue;
};

// Eight bf16 of one operand tile chunk, held in registers between the global
// load and the LDS store.
struct Chunk8 {
  uint4 v;
};

// Operand tile (128 rows of m or n, 32 of k) from global memory into registers:
// thread t handles chunks t and t + 256 of 8 elements.  KC: k is the contiguous
// dimension (chunk = 8 consecutive k of one row); otherwise the row dimension is
// contiguous (chunk = 8 consecutive rows at one k).  Out-of-range elements are 0.
template <bool KC>
__device__ __forceinline__ void load_tile(const uint16_t* X, int64_t s_mn, int64_t s_k, int MN,
                                          int K, int mn0, int k0, bool vec, Chunk8 (&c)[2]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int id = tid + 256 * i;
    int mn, k;
    if constexpr (KC) {
      mn = mn0 + (id >> 2);
      k = k0 + (id & 3) * 8;
    } else {
      k = k0 + (id >> 4);
      mn = mn0 + (id & 15) * 8;
    }
    const bool full = KC ? (mn < MN && k + 8 <= K) : (k < K && mn + 8 <= MN);
    if (vec && full) {
      c[i].v = *reinterpret_cast<const uint4*>(X + mn * s_mn + k * s_k);
    } else {
      uint16_t e[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int mj = KC ? mn : mn + j, kj = KC ? k + j : k;
        e[j] = (mj < MN && kj < K) ? X[mj * s_mn + kj * s_k] : uint16_t(0);
      }
      c[i].v.x = e[0] | (uint32_t(e[1]) << 16);
      c[i].v.y = e[2] | (uint32_t(e[3]) << 16);
      c[i].v.z = e[4] | (uint32_t(e[5]) << 16);
      c[i].v.w = e[6] | (uint32_t(e[7]) << 16);
    }
  }
}

// Registers -> LDS as T[row][k] (row = m or n, k contiguous, padded rows).
template <bool KC>
__device__ __forceinline__ void store_tile(uint16_t (*T)[kLdsK], const Chunk8 (&c)[2]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int id = tid + 256 * i;
    if constexpr (KC) {
      *reinterpret_cast<uint4*>(&T[id >> 2][(id & 3) * 8]) = c[i].v;
    } else {
      const int k = id >> 4, r = (id & 15) * 8;
      const uint32_t w[4] = {c[i].v.x, c[i].v.y, c[i].v.z, c[i].v.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        T[r + 2 * j][k] = static_cast<uint16_t>(w[j] & 0xFFFFu);
        T[r + 2 * j + 1][k] = static_cast<uint16_t>(w[j] >> 16);
      }
    }
  }
}

// C (M x N) = A (M x K) B (K x N) in bf16 with FP32 accumulation.  One
// workgroup = 4 waves = a 128 x 128 tile of C; wave w owns the 64 x 64 quarter
// (w / 2, w % 2) as 2 x 2 blocks of v_mfma_f32_32x32x16_bf16.  Per 32-deep K
// step the A and B tiles are staged in LDS as [m][k] and [n][k]; the next
// step's global loads are issued before the current step's MFMAs.
// blockIdx.z splits K (kEpiAtomicF32 sums the parts into a zeroed C32).
// MFMA operand layout (32x32x16): lane l holds A[l % 32][8 (l / 32) + 0..7] and
// B[8 (l / 32) + 0..7][l % 32]; the result register r of lane l is
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
template <bool AK, bool BK>
__global__ void __launch_bounds__(256) gemm_bf16(const GemmArgs g) {
  __shared__ __align__(16) uint16_t As[kTM][kLdsK];
  __shared__ __align__(16) uint16_t Bs[kTN][kLdsK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int kb = blockIdx.z * g.k_split;
  const int ke = min(g.K, kb + g.k_split);
  // 16-byte loads need the contiguous dimension's partner stride and the base aligned
  const bool va = (reinterpret_cast<uintptr_t>(g.A) & 15) == 0 && ((AK ? g.sAm : g.sAk) % 8) == 0;
  const bool vb = (reinterpret_cast<uintptr_t>(g.B) & 15) == 0 && ((BK ? g.sBn : g.sBk) % 8) == 0;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  Chunk8 ra[2], rb[2];
  if (kb < ke) {
    load_tile<AK>(g.A, g.sAm, g.sAk, g.M, ke, m0, kb, va, ra);
    load_tile<BK>(g.B, g.sBn, g.sBk, g.N, ke, n0, kb, vb, rb);
  }
  for (int k0 = kb; k0 < ke; k0 += kTK) {
    store_tile<AK>(As, ra);
    store_tile<BK>(Bs, rb);
    __syncthreads();
    if (k0 + kTK < ke) {
      load_tile<AK>(g.A, g.sAm, g.sAk, g.M, ke, m0, k0 + kTK, va, ra);
      load_tile<BK>(g.B, g.sBn, g.sBk, g.N, ke, n0, k0 + kTK, vb, rb);
    }
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 16) {
      const int row = lane & 31, kq = kk + 8 * (lane >> 5);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 32 * i + row][kq]);
        b[i] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 32 * i + row][kq]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 32 * j + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row >= g.M || col >= g.N) continue;
        float v = acc[i][j][r];
        const int64_t o = row * g.ldc + col;
        switch (g.epilogue) {
          case kEpiStoreF32:
            g.C32[o] = v;
            break;
          case kEpiBiasF32:
            g.C32[o] = v + g.bias[col];
            break;
          case kEpiBiasSiluBf16:
            v += g.bias[col];
            if (g.C32) g.C32[o] = v;
            g.C16[o] = f32_to_bf16_bits(silu(v));
            break;
          case kEpiAtomicF32:
            atomicAdd(g.C32 + o, v);
            break;
        }
      }
    }
  }
}

// elementwise helpers -------------------------------------------------------
__global__ void to_bf16(const float* x, uint16_t* y, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) y[i] = f32_to_bf16_bits(x[i]);
}

// dZ = dY * silu'(Z) (bf16 out), dY bf16/fp32 in
__global__ void silu_backward(const float* dY, const float* Z, int64_t n, uint16_t* dZ16,
                              float* dZ32) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float d = dY[i] * silu_grad(Z[i]);
  dZ16[i] = f32_to_bf16_bits(d);
  if (dZ32) dZ32[i] = d;
}

// column sums of a [rows][cols] fp32 matrix (bias gradients): one block per column.
__global__ void column_sums(const float* x, int64_t rows, int cols, int64_t ld, float* out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += x[r * ld + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// MSE loss on the output layer: dY = 2 (Y - T) / (rows * cols), loss partials.
__global__ void mse_grad(const float* Y, int64_t ldy, const float* T, int64_t rows, int cols,
                         float* dY, float* loss_partial) {
  __shared__ float red[256];
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  float l = 0.0f;
  if (i < rows * cols) {
    const int64_t r = i / cols;
    const int c = static_cast<int>(i % cols);
    const float d = Y[r * ldy + c] - T[i];
    dY[r * ldy + c] = 2.0f * d / static_cast<float>(rows * cols);
    l = d * d;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss_partial, red[0] / static_cast<float>(rows * cols));
}

// Adam (Kingma & Ba) on fp32 master parameters.
__global__ void adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                     float b2, float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}


// ---------------------------------------------------------- fused forward ----
// Inference of a whole network [w0, H, ..., H, w_out] (H = 256 or 512, w0 + 2 <=
// 64 or 128, w_out <= 32, 2..kFusedMaxLayers weight layers) in one kernel,
// computed transposed: C^T[n][m] = sum_k W[n][k] act[m][k], so the weights are the
// MFMA A operand, streamed through LDS, and the activations the B operand, held in
// registers from the first layer to the last.  A workgroup owns 4 waves x 32
// samples; the accumulator layout of a 32-feature block (lane holds features
// 8q + 4h + t, h = lane / 32) becomes the next layer's operand layout (8
// consecutive features per lane) with one v_permlane32_swap per dword pair.
//
// Weight stream.  All layers' bf16 weights lie back to back ([H][k0], then [H][H]
// per hidden layer, then [32][H]) and are consumed as equal tiles of 32 H
// elements: layer 0 as tiles of 32 H / k0 rows (several output blocks each), every
// later layer as one 32-row output block per tile.  Tile t + 2's global loads
// (t + 1's at H = 512) are issued when tile t starts and land in registers; tile
// t + 1 goes from registers to the other LDS buffer while tile t feeds the MFMAs.  Tiles, layers and blocks
// are unrolled at compile time (the layer count is a template parameter), so every
// address is a constant offset and every wait count exact.
//
// Log2-scaled activations.  Hidden layers carry a = log2(e) SiLU(z) (rounded to
// bf16) instead of SiLU(z), and their biases are log2(e) b (FP32), so the MFMA sums
// are z' = log2(e) z and SiLU costs one exp2, one add, one reciprocal and one
// multiply: a = z' / (1 + 2^-z').  Layer 0 (raw inputs) scales its sums by log2(e)
// in its epilogue; the output layer's bias enters as log2(e) b_out and its sums are
// divided by log2(e) at the end.  The weights are the training path's bf16
// values; results agree with the layer-by-layer path to bf16 rounding of the
// activations, not bit for bit.
//
// Biases.  Hidden and output blocks start their accumulators from the block's
// biases, loaded before the weight prefetch issued with them (so waiting for them
// never waits for weights); layer 0's bias is two extra weight columns,
// the bf16 high and low parts of b, against constant-1 inputs at w0 and w0 + 1.
//
// Scheduling.  A block's SiLU epilogue (VALU) is emitted with the next block's
// MFMA chain, which does not depend on it, so the two interleave.
constexpr int kFusedMaxLayers = 6;     // H = 256
constexpr int kFusedMaxLayers512 = 5;  // H = 512 (code size)
constexpr float kLog2e = 1.4426950408889634f;
// sched_barrier mask: VALU, SALU and transcendental instructions may cross it
// (MFMA, LDS and global memory instructions may not).
constexpr int kSchedAlu = 0x2 | 0x4 | 0x400;

struct FusedArgs {
  const float* x;     // [rows][w0] fp32
  int64_t rows;
  const uint32_t* rows_dev;  // or nullptr: the row count, at most `rows`, read on the device
  const uint32_t* idx;       // or nullptr: sample m is row idx[m] of x (else row m)
  int64_t ldx, ldy;          // > 0: x / y column-major with these leading dimensions
  int w0;             // inputs; columns w0 and w0 + 1 of layer 0 hold its bias
  int w_out;          // outputs (<= 32)
  const uint16_t* W;  // the weight stream
  const float* b;     // log2(e) b of layers 1 .. L - 1: layer l at b + (l - 1) H (output: 32)
  float* y;           // [rows][w_out]
};

// Global-address loads (vmcnt only): through a generic pointer they would be flat
// loads, whose lgkmcnt share makes every LDS wait also wait for the prefetch.
typedef const u32x4 __attribute__((address_space(1))) global_u32x4;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef const f32x4 __attribute__((address_space(1))) global_f32x4;
__device__ __forceinline__ u32x4 load_global16(const void* p) { return *(global_u32x4*)(p); }
__device__ __forceinline__ f32x4 load_global_f32x4(const float* p) {
  return *(global_f32x4*)(p);
}

template <int H, int K0S, int L>
struct Fz {
  static constexpr int HID = H;                       // hidden width
  static constexpr int K0 = 16 * K0S;                 // padded input width
  static constexpr int KS = H / 16;                   // k16 steps of a later layer
  static constexpr int NB = H / 32;                   // output blocks of layer 0 / a hidden layer
  static constexpr int TILE = 32 * H;                 // elements of one streamed tile
  static constexpr int CH = TILE / (8 * 256);         // 16-byte chunks per thread and tile
  static constexpr int R0 = TILE / K0;                // rows of a layer-0 tile
  static constexpr int B0 = R0 / 32;                  // output blocks of a layer-0 tile
  static constexpr int T0 = H / R0;                   // layer-0 tiles
  static constexpr int NT = T0 + (L - 2) * NB + 1;    // tiles in the stream
  static constexpr int NBLK = NB + (L - 2) * NB + 1;  // output blocks
  static constexpr int S0 = K0 + 8, SH = H + 8;       // LDS row strides (elements)
  static constexpr int LDS = (R0 * S0 > 32 * SH) ? R0 * S0 : 32 * SH;
  // Tiles in flight in registers: two ahead at H = 256; one at H = 512, whose
  // activations (two 128-register sets) leave no room for a second.
  static constexpr int PD = H == 256 ? 2 : 1;
  // Weight fragments read ahead of the MFMAs: at H = 512 (one wave per SIMD, no
  // other wave to cover the LDS latency) three; at H = 256 the compiler's one
  // (two waves per SIMD; three measured no faster).
  static constexpr int RA = H == 256 ? 1 : 3;
  static_assert(K0 % 32 == 0 && B0 * K0S == KS && T0 * B0 == NB, "tile geometry");
  static constexpr int tile_of(int g) { return g < NB ? g / B0 : T0 + (g - NB); }
  static constexpr int pos_of(int g) { return g < NB ? g % B0 : 0; }
  static constexpr int first_of(int t) { return t < T0 ? t * B0 : NB + (t - T0); }
  static constexpr int blocks_of(int t) { return t < T0 ? B0 : 1; }
  static constexpr int layer_of(int g) { return g < NB ? 0 : 1 + (g - NB) / NB; }
  static constexpr int nb_of(int g) { return g < NB ? g : (g - NB) % NB; }
};

template <int P, class A>
__device__ __forceinline__ A& pick(A& a0, A& a1) {
  if constexpr (P == 0)
    return a0;
  else
    return a1;
}

// Tile TL's weights into registers (nothing past the last tile).
template <class F, int TL>
__device__ __forceinline__ void fz_fetch(const uint16_t* W, u32x4 (&r)[F::CH]) {
  if constexpr (TL < F::NT) {
    const uint16_t* p = W + static_cast<int64_t>(TL) * F::TILE + 8 * threadIdx.x;
#pragma unroll
    for (int c = 0; c < F::CH; ++c) r[c] = load_global16(p + 2048 * c);
  }
}

// Registers -> LDS in the tile's row layout (layer 0: rows of K0, else of H).
template <class F, int TL>
__device__ __forceinline__ void fz_stage(uint16_t* T, const u32x4 (&r)[F::CH]) {
  if constexpr (TL < F::NT) {
    constexpr int per_row = (TL < F::T0 ? F::K0 : F::HID) / 8;  // 16-byte chunks per row
    constexpr int stride = TL < F::T0 ? F::S0 : F::SH;
#pragma unroll
    for (int c = 0; c < F::CH; ++c) {
      const int e = threadIdx.x + 256 * c;
      *reinterpret_cast<u32x4*>(T + (e / per_row) * stride + 8 * (e % per_row)) = r[c];
    }
  }
}

// The 16 biases of block g (a later layer) this lane's accumulators start from.
template <class F, int G>
__device__ __forceinline__ void fz_bias(const float* b, f32x16& acc) {
  constexpr int l = F::layer_of(G), nb = F::nb_of(G);
  const float* p = b + (l - 1) * F::HID + 32 * nb + 4 * ((threadIdx.x & 63) >> 5);
  f32x4 v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = load_global_f32x4(p + 8 * q);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[4 * q + u] = v[q][u];
}

// a = z' / (1 + 2^-z') of a block (z' = SCALE * acc) -> the next layer's fragments.
template <bool SCALE>
__device__ __forceinline__ void fz_epilogue(const f32x16& acc, bf16x8& lo_frag, bf16x8& hi_frag) {
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const float z = SCALE ? acc[i] * kLog2e : acc[i];
    v[i] = z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z));
  }
  uint32_t P0 = pack_bf16(v[0], v[1]), P1 = pack_bf16(v[2], v[3]);
  uint32_t Q0 = pack_bf16(v[4], v[5]), Q1 = pack_bf16(v[6], v[7]);
  uint32_t R0 = pack_bf16(v[8], v[9]), R1 = pack_bf16(v[10], v[11]);
  uint32_t S0 = pack_bf16(v[12], v[13]), S1 = pack_bf16(v[14], v[15]);
  auto x0 = __builtin_amdgcn_permlane32_swap(P0, Q0, false, false);
  auto x1 = __builtin_amdgcn_permlane32_swap(P1, Q1, false, false);
  auto x2 = __builtin_amdgcn_permlane32_swap(R0, S0, false, false);
  auto x3 = __builtin_amdgcn_permlane32_swap(R1, S1, false, false);
  lo_frag = __builtin_bit_cast(bf16x8, u32x4{x0[0], x1[0], x0[1], x1[1]});
  hi_frag = __builtin_bit_cast(bf16x8, u32x4{x2[0], x3[0], x2[1], x3[1]});
}

template <int H, int K0S, int L>
__global__ void __launch_bounds__(256, H == 256 ? 2 : 1) mlp_fused_forward(const FusedArgs a) {
  using F = Fz<H, K0S, L>;
  __shared__ __align__(16) uint16_t lds[2][F::LDS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, ln = lane & 31, h = lane >> 5;
  const int64_t m = static_cast<int64_t>(blockIdx.x) * 128 + wave * 32 + ln;  // this lane's sample
  int64_t rows = a.rows;
  if (a.rows_dev != nullptr) rows = min(rows, static_cast<int64_t>(*a.rows_dev));
  if (static_cast<int64_t>(blockIdx.x) * 128 >= rows) return;  // the whole workgroup

  u32x4 rg[F::PD][F::CH];
  fz_fetch<F, 0>(a.W, rg[0]);
  // inputs, then the two constant-1 columns of layer 0's bias (branch free: every
  // load is in bounds, rows past the end read the last row and are zeroed)
  bf16x8 xin[K0S];
  const int64_t mx = m < rows ? m : rows - 1;
  const int64_t xr = a.idx != nullptr ? static_cast<int64_t>(a.idx[mx]) : mx;
  // input k of this sample at xs + k * xk (row- or column-major)
  const float* xs = a.x + (a.ldx > 0 ? xr : xr * a.w0);
  const int64_t xk = a.ldx > 0 ? a.ldx : 1;
#pragma unroll
  for (int ks = 0; ks < K0S; ++ks) {
    uint32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float f[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int k = 16 * ks + 8 * h + 2 * j + u;
        const float xv = xs[(k < a.w0 ? k : a.w0 - 1) * xk];
        const float fv = k < a.w0 ? xv : (k < a.w0 + 2 ? 1.0f : 0.0f);
        f[u] = m < rows ? fv : 0.0f;
      }
      d[j] = pack_bf16(f[0], f[1]);
    }
    xin[ks] = __builtin_bit_cast(bf16x8, u32x4{d[0], d[1], d[2], d[3]});
  }
  fz_stage<F, 0>(lds[0], rg[0]);
  if constexpr (F::PD == 2) fz_fetch<F, 1>(a.W, rg[1]);

  bf16x8 act0[F::KS], act1[F::KS];  // layer l's output: act0 for even l, act1 for odd
  f32x16 prev, bnext, bcur;
  static_for<0, F::NT>([&](auto Tc) {
    constexpr int t = decltype(Tc)::value;
    __syncthreads();  // tile t staged in lds[t & 1]; the other buffer is free
    // biases before the weight prefetch, so waiting for them never waits for weights
    if constexpr (F::PD == 2) {
      if constexpr (t + 1 < F::NT && t + 1 >= F::T0) fz_bias<F, F::first_of(t + 1)>(a.b, bnext);
      fz_fetch<F, t + 2>(a.W, rg[t & 1]);
    } else {
      if constexpr (t >= F::T0) fz_bias<F, F::first_of(t)>(a.b, bcur);
      fz_fetch<F, t + 1>(a.W, rg[0]);
    }
    const uint16_t* T = lds[t & 1];
    static_for<0, F::blocks_of(t)>([&](auto Jc) {
      constexpr int g = F::first_of(t) + decltype(Jc)::value;
      constexpr int layer = F::layer_of(g);
      constexpr int ksteps = layer == 0 ? K0S : F::KS;
      constexpr int stride = layer == 0 ? F::S0 : F::SH;
      // The previous block's epilogue.  Inside a layer this block's MFMA chain does
      // not depend on it, so the scheduler interleaves the two; the first block of
      // a layer reads its last two fragments, so there it comes first (program
      // order), and only the last two steps of the chain wait for it.
      auto prev_epilogue = [&]() {
        if constexpr (g > 0) {
          constexpr int pl = F::layer_of(g - 1), pnb = F::nb_of(g - 1);
          auto& out = pick<pl & 1>(act0, act1);
          fz_epilogue<pl == 0>(prev, out[2 * pnb], out[2 * pnb + 1]);
        }
      };
      constexpr bool first_of_layer = g > 0 && F::layer_of(g - 1) != layer;
      if constexpr (first_of_layer) prev_epilogue();
      f32x16 acc;
      if constexpr (layer == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
      } else {
        acc = bcur;
      }
      const uint16_t* arow = T + (32 * F::pos_of(g) + ln) * stride + 8 * h;
      // Weight fragments RA - 1 steps ahead of their MFMA (a ring in registers); the
      // scheduling barrier keeps the LDS reads there (ALU work may still cross).
      constexpr int RA = ksteps < F::RA ? ksteps : F::RA;
      bf16x8 wr[RA];
#pragma unroll
      for (int p = 0; p < RA - 1; ++p) wr[p] = *reinterpret_cast<const bf16x8*>(arow + 16 * p);
#pragma unroll
      for (int ks = 0; ks < ksteps; ++ks) {
        if (ks + RA - 1 < ksteps)
          wr[(ks + RA - 1) % RA] = *reinterpret_cast<const bf16x8*>(arow + 16 * (ks + RA - 1));
        if constexpr (RA > 1) __builtin_amdgcn_sched_barrier(kSchedAlu);
        const bf16x8 w = wr[ks % RA];
        if constexpr (layer == 0)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, xin[ks], acc, 0, 0, 0);
        else
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              w, pick<(layer - 1) & 1>(act0, act1)[ks], acc, 0, 0, 0);
      }
      if constexpr (!first_of_layer) prev_epilogue();
      prev = acc;
    });
    if constexpr (t + 1 < F::NT) fz_stage<F, t + 1>(lds[(t + 1) & 1], rg[(t + 1) % F::PD]);
    if constexpr (F::PD == 2) bcur = bnext;
  });
  // prev holds the output block: log2(e) (W_out a + b_out) with a = log2(e) SiLU
  if (m < rows) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n = 8 * (r >> 2) + 4 * h + (r & 3);
      if (n < a.w_out) a.y[a.ldy > 0 ? n * a.ldy + m : m * a.w_out + n] = prev[r] * (1.0f / kLog2e);
    }
  }
}

// One layer's part of the weight stream: [out_p][in_p] bf16 of an fp32 [out][in]
// matrix, zero padded; with `bias_cols`, columns in and in + 1 hold the bf16
// high and low parts of b (layer 0's bias against constant-1 inputs).
__global__ void fz_pack_weights(const float* w, const float* b, int out, int in, int out_p,
                                int in_p, int bias_cols, uint16_t* dst) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= int64_t(out_p) * in_p) return;
  const int n = static_cast<int>(i / in_p), k = static_cast<int>(i % in_p);
  uint16_t v = 0;
  if (n < out) {
    if (k < in) {
      v = f32_to_bf16_bits(w[int64_t(n) * in + k]);
    } else if (bias_cols && k < in + 2) {
      const uint16_t hi = f32_to_bf16_bits(b[n]);
      const float hf = __builtin_bit_cast(float, static_cast<uint32_t>(hi) << 16);
      v = k == in ? hi : f32_to_bf16_bits(b[n] - hf);
    }
  }
  dst[i] = v;
}

// s * b of one layer, zero padded to n_p entries.
__global__ void fz_scale_bias(const float* b, int n, int n_p, float s, float* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_p) dst[i] = i < n ? s * b[i] : 0.0f;
}

}  // namespace

// ------------------------------------------------------------------ host ----

struct ce_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<int> widths;  // w[0] inputs ... w[L] outputs
  int L = 0;
  std::vector<int64_t> w_off, b_off;  // offsets into the parameter vector
  int64_t n_params = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;  // master, grad, Adam moments
  std::vector<uint16_t*> Wb;  // bf16 copies [out][in]
  int fused_h = 0;            // hidden width of the fused inference kernel, 0 = none
  bool use_fused = true;      // inference (ce_mlp_forward, closure) through it
  bool wp_stale = true;       // the fused kernel's weight stream older than the master
  uint16_t* Wp_all = nullptr;  // the fused kernel's weight stream (all layers back to back)
  std::vector<uint16_t*> Wp;   // layer l's part of it
  float* Bs = nullptr;         // the fused kernel's scaled biases of layers 1 .. L - 1
  float *fx = nullptr, *fy = nullptr;  // ce_mlp_forward's device input / output through it
  int64_t fcap = 0;
  int64_t adam_t = 0;
  // batch workspace
  int64_t cap = 0;
  std::vector<uint16_t*> act;  // act[l]: bf16 input of layer l [cap][w[l]]
  std::vector<float*> Z;       // Z[l]: pre-activation of hidden layer l [cap][w[l+1]]
  float* out = nullptr;        // [cap][w[L]]
  float *dY = nullptr, *dZ32 = nullptr;
  uint16_t* dZ16 = nullptr;
  float* loss = nullptr;
  float* xin = nullptr;  // fp32 staging of inputs [cap][w[0]]
  float* tgt = nullptr;  // fp32 targets [cap][w[L]]
  // device-resident training set
  float *ds_x = nullptr, *ds_t = nullptr;
  int64_t ds_n = 0;
  int64_t* perm = nullptr;
};

namespace {

int fail(int code, const std::string& m) { return ce::set_error(code, m); }

#define MLP_HIP(call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ce::clear_stale_error(); /* reported here: not left for the next check */   \
      return fail(CE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                             \
  } while (0)

template <class T>
int dalloc(T** p, int64_t n) {
  *p = nullptr;
  MLP_HIP(hipMalloc(reinterpret_cast<void**>(p), (n > 0 ? n : 1) * sizeof(T)));
  return CE_OK;
}

unsigned blocks_for(int64_t n, int t = 256) { return static_cast<unsigned>((n + t - 1) / t); }

void free_mlp(ce_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(m->P), f(m->G), f(m->M1), f(m->M2), f(m->out), f(m->dY), f(m->dZ32), f(m->dZ16);
  f(m->loss), f(m->xin), f(m->tgt), f(m->ds_x), f(m->ds_t), f(m->perm);
  for (auto* p : m->Wb) f(p);
  f(m->Wp_all), f(m->Bs), f(m->fx), f(m->fy);
  for (auto* p : m->act) f(p);
  for (auto* p : m->Z) f(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int ensure_batch(ce_mlp* m, int64_t B) {
  if (B <= m->cap) return CE_OK;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (auto* p : m->act) f(p);
  for (auto* p : m->Z) f(p);
  f(m->out), f(m->dY), f(m->dZ32), f(m->dZ16), f(m->xin), f(m->tgt);
  m->act.assign(m->L, nullptr);
  m->Z.assign(m->L, nullptr);
  int maxw = 0;
  for (int w : m->widths) maxw = w > maxw ? w : maxw;
  int rc = CE_OK;
  for (int l = 0; l < m->L && rc == CE_OK; ++l) {
    rc = dalloc(&m->act[l], B * m->widths[l]);
    if (rc == CE_OK && l + 1 < m->L) rc = dalloc(&m->Z[l], B * m->widths[l + 1]);
  }
  if (rc == CE_OK) rc = dalloc(&m->out, B * m->widths[m->L]);
  if (rc == CE_OK) rc = dalloc(&m->dY, B * maxw);
  if (rc == CE_OK) rc = dalloc(&m->dZ32, B * maxw);
  if (rc == CE_OK) rc = dalloc(&m->dZ16, B * maxw);
  if (rc == CE_OK) rc = dalloc(&m->xin, B * m->widths[0]);
  if (rc == CE_OK) rc = dalloc(&m->tgt, B * m->widths[m->L]);
  if (rc == CE_OK) m->cap = B;
  return rc;
}

// Launches C = A B.  Each operand has unit stride along k (row-major A / a
// [n][k] B) or along its m / n dimension.  With `split_k`, kEpiAtomicF32 sums
// K slices into C32, which the caller has zeroed.
int gemm(hipStream_t s, GemmArgs g, bool split_k = false) {
  const bool ak = g.sAk == 1, bk = g.sBk == 1;
  if ((!ak && g.sAm != 1) || (!bk && g.sBn != 1))
    return fail(CE_ERR_INVALID_ARGUMENT, "gemm operands need a unit stride");
  const unsigned gx = blocks_for(g.N, kTN), gy = blocks_for(g.M, kTM);
  unsigned gz = 1;
  g.k_split = ((g.K + kTK - 1) / kTK) * kTK;
  if (split_k) {
    // enough workgroups for the chip: ~512 over the tiles, >= 4 K steps each
    const int steps = (g.K + kTK - 1) / kTK;
    const int want = static_cast<int>(std::max<unsigned>(1u, 512u / (gx * gy)));
    const int per = std::max(4, (steps + want - 1) / want);
    g.k_split = per * kTK;
    gz = static_cast<unsigned>((steps + per - 1) / per);
  }
  const dim3 grid(gx, gy, gz);
  if (ak && bk) hipLaunchKernelGGL((gemm_bf16<true, true>), grid, dim3(256), 0, s, g);
  else if (ak) hipLaunchKernelGGL((gemm_bf16<true, false>), grid, dim3(256), 0, s, g);
  else if (bk) hipLaunchKernelGGL((gemm_bf16<false, true>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_bf16<false, false>), grid, dim3(256), 0, s, g);
  return CE_OK;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

int export_all(ce_mlp* m) {
  ce::clear_stale_error();
  for (int l = 0; l < m->L; ++l) {
    const int in = m->widths[l], out = m->widths[l + 1];
    const int64_t n = static_cast<int64_t>(in) * out;
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(n)), dim3(256), 0, m->stream, m->P + m->w_off[l],
                       m->Wb[l], n);
  }
  m->wp_stale = true;  // the padded copies are rebuilt before the next fused inference
  MLP_HIP(hipGetLastError());
  return CE_OK;
}

// Forward pass of x (fp32 [B][w0], B <= cap) into y (fp32 [B][wL]) on stream s;
// keep = true also stores Z for backward (the bf16 activations always stay).
int forward(ce_mlp* m, const float* x, int64_t B, bool keep, float* y, hipStream_t s) {
  const int w0 = m->widths[0];
  hipLaunchKernelGGL(to_bf16, dim3(blocks_for(B * w0)), dim3(256), 0, s, x, m->act[0], B * w0);
  for (int l = 0; l < m->L; ++l) {
    const int in = m->widths[l], out = m->widths[l + 1];
    GemmArgs g{};
    g.A = m->act[l];
    g.sAm = in;
    g.sAk = 1;
    g.B = m->Wb[l];
    g.sBk = 1;
    g.sBn = in;
    g.M = static_cast<int>(B);
    g.N = out;
    g.K = in;
    g.bias = m->P + m->b_off[l];
    g.ldc = out;
    if (l + 1 < m->L) {
      g.epilogue = kEpiBiasSiluBf16;
      g.C16 = m->act[l + 1];
      g.C32 = keep ? m->Z[l] : nullptr;
    } else {
      g.epilogue = kEpiBiasF32;
      g.C32 = y;
    }
    const int rc = gemm(s, g);
    if (rc != CE_OK) return rc;
  }
  MLP_HIP(hipGetLastError());
  return CE_OK;
}

// Backward of the last forward (keep = true), from dL/dout in m->dY [B][wL];
// writes parameter gradients into m->G.
int backward(ce_mlp* m, int64_t B) {
  float* dY = m->dY;
  for (int l = m->L - 1; l >= 0; --l) {
    const int in = m->widths[l], out = m->widths[l + 1];
    const int64_t n = B * out;
    const float* dz32 = m->dZ32;
    if (l + 1 < m->L) {
      hipLaunchKernelGGL(silu_backward, dim3(blocks_for(n)), dim3(256), 0, m->stream, dY, m->Z[l],
                         n, m->dZ16, m->dZ32);
    } else {  // linear output layer: dZ = dY
      hipLaunchKernelGGL(to_bf16, dim3(blocks_for(n)), dim3(256), 0, m->stream, dY, m->dZ16, n);
      dz32 = dY;
    }
    hipLaunchKernelGGL(column_sums, dim3(out), dim3(256), 0, m->stream, dz32, B, out,
                       static_cast<int64_t>(out), m->G + m->b_off[l]);
    GemmArgs gw{};  // dW[o][i] = sum_b dZ[b][o] X[b][i]
    gw.A = m->dZ16;
    gw.sAm = 1;
    gw.sAk = out;
    gw.B = m->act[l];
    gw.sBk = in;
    gw.sBn = 1;
    gw.M = out;
    gw.N = in;
    gw.K = static_cast<int>(B);
    gw.C32 = m->G + m->w_off[l];
    gw.ldc = in;
    gw.epilogue = kEpiAtomicF32;  // split over the batch, summed into zeroed G
    MLP_HIP(hipMemsetAsync(gw.C32, 0, sizeof(float) * in * out, m->stream));
    int rc = gemm(m->stream, gw, true);
    if (rc != CE_OK) return rc;
    if (l > 0) {  // dX[b][i] = sum_o dZ[b][o] W[o][i]
      GemmArgs gx{};
      gx.A = m->dZ16;
      gx.sAm = out;
      gx.sAk = 1;
      gx.B = m->Wb[l];
      gx.sBk = in;
      gx.sBn = 1;
      gx.M = static_cast<int>(B);
      gx.N = in;
      gx.K = out;
      gx.C32 = m->dY;  // dY of this layer is consumed (dZ16 / dz32) by now
      gx.ldc = in;
      gx.epilogue = kEpiStoreF32;
      rc = gemm(m->stream, gx);
      if (rc != CE_OK) return rc;
      dY = m->dY;
    }
  }
  MLP_HIP(hipGetLastError());
  return CE_OK;
}

// deterministic init on the host: U(-sqrt(6/(in+out)), +) (Glorot), zero biases
void init_params(const ce_mlp* m, uint64_t seed, std::vector<float>& p) {
  p.assign(m->n_params, 0.0f);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto next = [&s]() {  // splitmix64
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int l = 0; l < m->L; ++l) {
    const int in = m->widths[l], out = m->widths[l + 1];
    const float lim = std::sqrt(6.0f / static_cast<float>(in + out));
    for (int64_t i = 0; i < static_cast<int64_t>(in) * out; ++i) {
      const double u = (next() >> 11) * (1.0 / 9007199254740992.0);
      p[m->w_off[l] + i] = static_cast<float>((2.0 * u - 1.0) * lim);
    }
  }
}

}  // namespace

extern "C" {

int ce_mlp_create(const int32_t* widths, int32_t n_widths, uint64_t seed, ce_mlp** out) {
  if (out == nullptr || widths == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (n_widths < 2 || n_widths > 16) return fail(CE_ERR_INVALID_ARGUMENT, "2..16 widths");
  for (int i = 0; i < n_widths; ++i)
    if (widths[i] < 1 || widths[i] > 4096) return fail(CE_ERR_INVALID_ARGUMENT, "width out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CE_ERR_NO_DEVICE, "no HIP device");
  ce_mlp* m = new ce_mlp();
  m->widths.assign(widths, widths + n_widths);
  m->L = n_widths - 1;
  for (int l = 0; l < m->L; ++l) {
    m->w_off.push_back(m->n_params);
    m->n_params += static_cast<int64_t>(widths[l]) * widths[l + 1];
    m->b_off.push_back(m->n_params);
    m->n_params += widths[l + 1];
  }
  int rc = CE_OK;
  if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(CE_ERR_HIP, "stream");
  if (rc == CE_OK) rc = dalloc(&m->P, m->n_params);
  if (rc == CE_OK) rc = dalloc(&m->G, m->n_params);
  if (rc == CE_OK) rc = dalloc(&m->M1, m->n_params);
  if (rc == CE_OK) rc = dalloc(&m->M2, m->n_params);
  if (rc == CE_OK) rc = dalloc(&m->loss, 1);
  m->Wb.assign(m->L, nullptr);
  for (int l = 0; l < m->L && rc == CE_OK; ++l)
    rc = dalloc(&m->Wb[l], static_cast<int64_t>(widths[l]) * widths[l + 1]);
  // the fused inference kernel covers [w0, H, ..., H, w_out <= 32] with H = 256,
  // w0 + 2 <= 64 and up to 6 layers, or H = 512, w0 + 2 <= 128 and up to 5 layers
  {
    const int h = m->L >= 2 ? widths[1] : 0;
    bool ok = ((h == 256 && widths[0] + 2 <= 64 && m->L <= kFusedMaxLayers) ||
               (h == 512 && widths[0] + 2 <= 128 && m->L <= kFusedMaxLayers512)) &&
              widths[m->L] <= 32;
    for (int l = 1; l < m->L && ok; ++l) ok = widths[l] == h;
    m->fused_h = ok ? h : 0;
  }
  if (m->fused_h) {
    // The weight stream: [H][round64(w0 + 2)] (two bias columns), then (L - 2) x
    // [H][H], then [32][H], back to back; the scaled biases of layers 1 .. L - 1.
    m->Wp.assign(m->L, nullptr);
    int64_t total = 0;
    std::vector<int64_t> off(m->L);
    for (int l = 0; l < m->L; ++l) {
      off[l] = total;
      total += static_cast<int64_t>(round_up(widths[l + 1], 32)) *
               (l == 0 ? round_up(widths[0] + 2, 64) : widths[l]);
    }
    if (rc == CE_OK) rc = dalloc(&m->Wp_all, total);
    for (int l = 0; l < m->L && rc == CE_OK; ++l) m->Wp[l] = m->Wp_all + off[l];
    if (rc == CE_OK) rc = dalloc(&m->Bs, static_cast<int64_t>(m->L - 1) * m->fused_h);
  }
  if (rc == CE_OK) {
    std::vector<float> p;
    init_params(m, seed, p);
    if (hipMemcpyAsync(m->P, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice,
                       m->stream) != hipSuccess ||
        hipMemsetAsync(m->M1, 0, m->n_params * sizeof(float), m->stream) != hipSuccess ||
        hipMemsetAsync(m->M2, 0, m->n_params * sizeof(float), m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      rc = fail(CE_ERR_HIP, "parameter upload");
  }
  if (rc == CE_OK) rc = export_all(m);
  if (rc == CE_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = fail(CE_ERR_HIP, "sync");
  if (rc != CE_OK) {
    free_mlp(m);
    return rc;
  }
  *out = m;
  return CE_OK;
}

int ce_mlp_destroy(ce_mlp* m) {
  free_mlp(m);
  return CE_OK;
}

int64_t ce_mlp_param_count(const ce_mlp* m) { return m ? m->n_params : 0; }

int ce_mlp_get_params(ce_mlp* m, float* host) {
  if (!m || !host) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(host, m->P, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return CE_OK;
}

int ce_mlp_set_params(ce_mlp* m, const float* host) {
  if (!m || !host) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(m->P, host, m->n_params * sizeof(float), hipMemcpyHostToDevice, m->stream));
  m->adam_t = 0;
  MLP_HIP(hipMemsetAsync(m->M1, 0, m->n_params * sizeof(float), m->stream));
  MLP_HIP(hipMemsetAsync(m->M2, 0, m->n_params * sizeof(float), m->stream));
  const int rc = export_all(m);
  if (rc != CE_OK) return rc;
  MLP_HIP(hipStreamSynchronize(m->stream));
  return CE_OK;
}

int ce_mlp_forward(ce_mlp* m, const float* x_host, size_t batch, float* y_host) {
  if (!m || (!x_host && batch) || (!y_host && batch)) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  // rows per device pass: the fused kernel needs only input and output buffers
  const bool fused = m->fused_h && m->use_fused;
  const int64_t chunk = fused ? (1 << 20) : 65536;
  for (size_t b0 = 0; b0 < batch; b0 += chunk) {
    const int64_t B = static_cast<int64_t>(std::min<size_t>(chunk, batch - b0));
    int rc = CE_OK;
    if (fused && B > m->fcap) {
      if (m->fx) (void)hipFree(m->fx);
      if (m->fy) (void)hipFree(m->fy);
      m->fx = m->fy = nullptr;
      m->fcap = 0;
      rc = dalloc(&m->fx, B * m->widths[0]);
      if (rc == CE_OK) rc = dalloc(&m->fy, B * m->widths[m->L]);
      if (rc == CE_OK) m->fcap = B;
    } else if (!fused) {
      rc = ensure_batch(m, B);
    }
    if (rc != CE_OK) return rc;
    float* xd = fused ? m->fx : m->xin;
    float* yd = fused ? m->fy : m->out;
    MLP_HIP(hipMemcpyAsync(xd, x_host + b0 * m->widths[0], B * m->widths[0] * sizeof(float),
                           hipMemcpyHostToDevice, m->stream));
    ce::clear_stale_error();
    rc = fused ? ce::mlp_forward_device(m, xd, B, yd, m->stream)
               : forward(m, xd, B, false, yd, m->stream);
    if (rc != CE_OK) return rc;
    MLP_HIP(hipMemcpyAsync(y_host + b0 * m->widths[m->L], yd,
                           B * m->widths[m->L] * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  return CE_OK;
}

// Loss (mean squared error over the batch and outputs) and its gradient with
// respect to every parameter, for the batch given on the host.
int ce_mlp_gradients(ce_mlp* m, const float* x_host, const float* t_host, size_t batch,
                     float* grad_host, float* loss_host) {
  if (!m || !x_host || !t_host || !grad_host || batch == 0)
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument or empty batch");
  MLP_HIP(hipSetDevice(m->device));
  const int64_t B = static_cast<int64_t>(batch);
  int rc = ensure_batch(m, B);
  if (rc != CE_OK) return rc;
  const int wo = m->widths[m->L];
  MLP_HIP(hipMemcpyAsync(m->xin, x_host, B * m->widths[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemcpyAsync(m->tgt, t_host, B * wo * sizeof(float), hipMemcpyHostToDevice, m->stream));
  ce::clear_stale_error();
  rc = forward(m, m->xin, B, true, m->out, m->stream);
  if (rc != CE_OK) return rc;
  MLP_HIP(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
  hipLaunchKernelGGL(mse_grad, dim3(blocks_for(B * wo)), dim3(256), 0, m->stream, m->out,
                     static_cast<int64_t>(wo), m->tgt, B, wo, m->dY, m->loss);
  rc = backward(m, B);
  if (rc != CE_OK) return rc;
  MLP_HIP(hipMemcpyAsync(grad_host, m->G, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss_host) MLP_HIP(hipMemcpyAsync(loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return CE_OK;
}

int ce_mlp_set_dataset(ce_mlp* m, const float* x_host, const float* t_host, size_t n) {
  if (!m || (!x_host && n) || (!t_host && n)) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipStreamSynchronize(m->stream));
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(m->ds_x), f(m->ds_t), f(m->perm);
  m->ds_x = m->ds_t = nullptr;
  m->perm = nullptr;
  m->ds_n = 0;
  int rc = dalloc(&m->ds_x, static_cast<int64_t>(n) * m->widths[0]);
  if (rc == CE_OK) rc = dalloc(&m->ds_t, static_cast<int64_t>(n) * m->widths[m->L]);
  if (rc == CE_OK) rc = dalloc(&m->perm, static_cast<int64_t>(n));
  if (rc != CE_OK) return rc;
  MLP_HIP(hipMemcpyAsync(m->ds_x, x_host, n * m->widths[0] * sizeof(float), hipMemcpyHostToDevice,
                         m->stream));
  MLP_HIP(hipMemcpyAsync(m->ds_t, t_host, n * m->widths[m->L] * sizeof(float),
                         hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->ds_n = static_cast<int64_t>(n);
  return CE_OK;
}

}  // extern "C"

namespace {

__global__ void gather_rows(const float* src, const int64_t* idx, int64_t rows, int cols, float* dst) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows * cols) return;
  const int64_t r = i / cols;
  const int c = static_cast<int>(i % cols);
  dst[i] = src[idx[r] * cols + c];
}

int adam_step(ce_mlp* m, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m->adam_t += 1;
  const float c1 = 1.0f - std::pow(b1, static_cast<float>(m->adam_t));
  const float c2 = 1.0f - std::pow(b2, static_cast<float>(m->adam_t));
  ce::clear_stale_error();
  hipLaunchKernelGGL(adam, dim3(blocks_for(m->n_params)), dim3(256), 0, m->stream, m->P, m->G,
                     m->M1, m->M2, m->n_params, lr, b1, b2, eps, c1, c2);
  MLP_HIP(hipGetLastError());
  return export_all(m);
}

}  // namespace

extern "C" {

// One pass over the device-resident training set in a shuffled order (the
// permutation comes from the host, seeded), Adam after every batch.
int ce_mlp_train_epoch(ce_mlp* m, size_t batch, float lr, uint64_t shuffle_seed, float* mean_loss) {
  if (!m || batch == 0) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument or zero batch");
  if (m->ds_n == 0) return fail(CE_ERR_RUNTIME, "no training set: call ce_mlp_set_dataset");
  MLP_HIP(hipSetDevice(m->device));
  const int64_t n = m->ds_n, B = static_cast<int64_t>(batch);
  int rc = ensure_batch(m, B);
  if (rc != CE_OK) return rc;
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  uint64_t s = shuffle_seed * 0x9E3779B97F4A7C15ull + 1;
  for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates with splitmix64
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int64_t j = static_cast<int64_t>(z % static_cast<uint64_t>(i + 1));
    std::swap(perm[i], perm[j]);
  }
  MLP_HIP(hipMemcpyAsync(m->perm, perm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice,
                         m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));  // perm is a local host buffer
  const int wi = m->widths[0], wo = m->widths[m->L];
  MLP_HIP(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
  int64_t n_batches = 0;
  for (int64_t b0 = 0; b0 + B <= n; b0 += B, ++n_batches) {  // whole batches only
    ce::clear_stale_error();  // the batch's launches are checked together, by backward
    hipLaunchKernelGGL(gather_rows, dim3(blocks_for(B * wi)), dim3(256), 0, m->stream, m->ds_x,
                       m->perm + b0, B, wi, m->xin);
    hipLaunchKernelGGL(gather_rows, dim3(blocks_for(B * wo)), dim3(256), 0, m->stream, m->ds_t,
                       m->perm + b0, B, wo, m->tgt);
    rc = forward(m, m->xin, B, true, m->out, m->stream);
    if (rc != CE_OK) return rc;
    hipLaunchKernelGGL(mse_grad, dim3(blocks_for(B * wo)), dim3(256), 0, m->stream, m->out,
                       static_cast<int64_t>(wo), m->tgt, B, wo, m->dY, m->loss);
    rc = backward(m, B);
    if (rc == CE_OK) rc = adam_step(m, lr);
    if (rc != CE_OK) return rc;
  }
  float total = 0.0f;
  MLP_HIP(hipMemcpyAsync(&total, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (mean_loss) *mean_loss = n_batches ? total / static_cast<float>(n_batches) : 0.0f;
  return CE_OK;
}

// Inference through the fused whole-network kernel when the shape allows it
// (default) or layer by layer; returns 1 when the fused kernel is used.
int ce_mlp_set_fused(ce_mlp* m, int32_t enable) {
  if (!m) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  m->use_fused = enable != 0;
  return (m->fused_h && m->use_fused) ? 1 : 0;
}

// One Adam step on the gradients left by ce_mlp_gradients (tests).
int ce_mlp_adam_step(ce_mlp* m, float lr) {
  if (!m) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  const int rc = adam_step(m, lr);
  if (rc != CE_OK) return rc;
  MLP_HIP(hipStreamSynchronize(m->stream));
  return CE_OK;
}

}  // extern "C"

namespace ce {

int mlp_width(const ce_mlp* m, int layer) {
  return (m != nullptr && layer >= 0 && layer <= m->L) ? m->widths[layer] : 0;
}
int mlp_layers(const ce_mlp* m) { return m != nullptr ? m->L : 0; }

bool mlp_fused_ready(const ce_mlp* m) { return m != nullptr && m->fused_h && m->use_fused; }

int mlp_forward_device(ce_mlp* m, const float* x, int64_t B, float* y, hipStream_t s,
                       const FusedIo& io) {
  ce::clear_stale_error();
  const bool plain_io = io.rows_dev == nullptr && io.idx == nullptr && io.ldx == 0 && io.ldy == 0;
  if (!plain_io && !(m->fused_h && m->use_fused))
    return fail(CE_ERR_INVALID_ARGUMENT,
                "a device row count, a row index or a column-major layout needs the fused "
                "inference kernel");
  if (m->fused_h && m->use_fused && B > 0) {
    const int H = m->fused_h, L = m->L, k0 = round_up(m->widths[0] + 2, 64);
    if (m->wp_stale) {
      for (int l = 0; l < L; ++l) {
        const int in = m->widths[l], out = m->widths[l + 1];
        const int op = round_up(out, 32), ip = l == 0 ? k0 : in;
        hipLaunchKernelGGL(fz_pack_weights, dim3(blocks_for(int64_t(op) * ip)), dim3(256), 0, s,
                           m->P + m->w_off[l], m->P + m->b_off[l], out, in, op, ip, l == 0 ? 1 : 0,
                           m->Wp[l]);
        if (l > 0)
          hipLaunchKernelGGL(fz_scale_bias, dim3(blocks_for(H)), dim3(256), 0, s,
                             m->P + m->b_off[l], out, H, kLog2e, m->Bs + int64_t(l - 1) * H);
      }
      m->wp_stale = false;
    }
    FusedArgs a{};
    a.x = x;
    a.rows = B;
    a.rows_dev = io.rows_dev;
    a.idx = io.idx;
    a.ldx = io.ldx;
    a.ldy = io.ldy;
    a.w0 = m->widths[0];
    a.w_out = m->widths[L];
    a.W = m->Wp_all;
    a.b = m->Bs;
    a.y = y;
    // 32 samples per wave, 4 waves per workgroup
    const dim3 grid(static_cast<unsigned>((B + 127) / 128)), blk(256);
    using Kern = void (*)(const FusedArgs);
    static const Kern k256[] = {mlp_fused_forward<256, 4, 2>, mlp_fused_forward<256, 4, 3>,
                                mlp_fused_forward<256, 4, 4>, mlp_fused_forward<256, 4, 5>,
                                mlp_fused_forward<256, 4, 6>};
    static const Kern k512a[] = {mlp_fused_forward<512, 4, 2>, mlp_fused_forward<512, 4, 3>,
                                 mlp_fused_forward<512, 4, 4>, mlp_fused_forward<512, 4, 5>};
    static const Kern k512b[] = {mlp_fused_forward<512, 8, 2>, mlp_fused_forward<512, 8, 3>,
                                 mlp_fused_forward<512, 8, 4>, mlp_fused_forward<512, 8, 5>};
    const Kern kern = H == 256 ? k256[L - 2] : (k0 == 64 ? k512a[L - 2] : k512b[L - 2]);
    hipLaunchKernelGGL(kern, grid, blk, 0, s, a);
    MLP_HIP(hipGetLastError());
    return CE_OK;
  }
  const int64_t chunk = 262144;
  const int64_t c = B < chunk ? B : chunk;
  int rc = ensure_batch(m, c);
  if (rc != CE_OK) return rc;
  const int w0 = m->widths[0], wl = m->widths[m->L];
  for (int64_t b0 = 0; b0 < B && rc == CE_OK; b0 += chunk) {
    const int64_t n = (B - b0) < chunk ? (B - b0) : chunk;
    rc = forward(m, x + b0 * w0, n, false, y + b0 * wl, s);
  }
  return rc;
}

}  // namespace ce
