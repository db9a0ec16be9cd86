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
// Inference of a whole network [w0, H, ..., H, w_out] (H = 256 with w0 <= 64, or
// H = 512 with w0 <= 128; w_out <= 32) in one kernel, computed transposed:
// C^T[n][m] = sum_k W[n][k] act[m][k], so the weights are the MFMA A operand,
// read from LDS, and the activations are the B operand, held in registers for
// the whole network.  A workgroup owns 4 waves x 32 samples; the weights stream
// through LDS in 32-row output blocks (all K columns), double-buffered and shared
// by the 4 waves, later tiles' global loads in flight while the current one feeds
// the MFMAs.  Between layers each lane turns its accumulators
// (features 8q + 4h + t of a 32-feature block, h = lane / 32) into the next
// layer's B fragments (8 consecutive features per lane) with one
// v_permlane32_swap per dword pair.  Products and their K order are those of
// the layer-by-layer GEMM.
constexpr int kFusedMaxLayers = 8;

struct FusedArgs {
  const float* x;  // [rows][w0] fp32
  int64_t rows;
  int w0;
  int layers;      // weight layers (>= 2): w0->H, (layers - 2) x H->H, H->w_out
  int w_out;
  // Padded bf16 weights of all layers back to back: [H][k0] (k0 = round64(w0)),
  // then [H][H] per hidden layer, then [32][H]; layer l >= 1 starts at
  // W + H k0 + (l - 1) H H (uniform arithmetic, no per-block table lookup).
  const uint16_t* W;
  // Per-layer bias pointers, read only through layer_b (constant indices:
  // indexing a kernel-argument array with a run-time layer would copy the
  // arguments to scratch).
  const float* b[kFusedMaxLayers];     // [out]
  float* y;                            // [rows][w_out]
};

__device__ __forceinline__ const float* layer_b(const FusedArgs& a, int l) {
  switch (l) {
    case 0: return a.b[0];
    case 1: return a.b[1];
    case 2: return a.b[2];
    case 3: return a.b[3];
    case 4: return a.b[4];
    case 5: return a.b[5];
    case 6: return a.b[6];
    default: return a.b[7];
  }
}

// Output-block-major fused inference (H = 256 or 512): the weights stream through
// LDS one 32-row output block (all K columns) at a time, so a block's accumulators
// are final after its tile and its SiLU epilogue (VALU) overlaps the next block's
// MFMAs.  Each wave holds 32 samples: the input fragments (act) and the next
// layer's fragments (nxt) in registers, H / 4 VGPRs each -- two waves per SIMD at
// H = 256 (their MFMA and VALU work also overlap), one at H = 512.
// Loads through pointers read from the device pointer table: cast to the global
// address space, so they are global loads (vmcnt only) rather than flat loads,
// whose lgkmcnt share would make every LDS wait also wait for the weight prefetch.
typedef const u32x4 __attribute__((address_space(1))) global_u32x4;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef const f32x4 __attribute__((address_space(1))) global_f32x4;
__device__ __forceinline__ u32x4 load_global16(const void* p) {
  return *(global_u32x4*)(p);
}
__device__ __forceinline__ f32x4 load_global_f32x4(const float* p) {
  return *(global_f32x4*)(p);
}

template <int H>
struct Nout {
  static constexpr int TS = H + 8;         // LDS tile row stride (elements)
  static constexpr int NB = H / 32;        // output blocks of a hidden layer
  static constexpr int KS = H / 16;        // k16 steps of a hidden layer's input
  static constexpr int LD = 32 * H / 2048; // 16-byte chunks per thread of a full tile
};

// Tile t of the stream: layer 0's NB blocks (k0 columns), the hidden layers'
// blocks, then the output layer's single block.
template <int H>
__device__ __forceinline__ void nout_tile(const FusedArgs& a, int t, int k0, const uint16_t** W,
                                          int* ldw, int* row0) {
  constexpr int NB = Nout<H>::NB;
  if (t < NB) {
    *W = a.W;
    *ldw = k0;
    *row0 = 32 * t;
  } else {
    const int l = 1 + (t - NB) / NB;
    *W = a.W + static_cast<int64_t>(H) * k0 + static_cast<int64_t>(l - 1) * H * H;
    *ldw = H;
    *row0 = l == a.layers - 1 ? 0 : 32 * ((t - NB) % NB);
  }
}

// Tile t's weights into registers; past the last tile the last one is loaded
// again, so every block issues the same loads and the compiler's wait counts stay
// exact (a load issued on only some paths makes later waits conservative).
template <int H>
__device__ __forceinline__ void nout_fetch(const FusedArgs& a, int t, int total, int k0,
                                           u32x4 (&r)[Nout<H>::LD]) {
  const uint16_t* W;
  int ldw, row0;
  nout_tile<H>(a, t < total ? t : total - 1, k0, &W, &ldw, &row0);
  const int per_row = ldw / 8;  // 16-byte chunks per row
#pragma unroll
  for (int i = 0; i < Nout<H>::LD; ++i) {  // every register written (chunks past the tile reload chunk 0)
    const int e = threadIdx.x + 256 * i < 32 * per_row ? threadIdx.x + 256 * i : 0;
    r[i] = load_global16(W + (row0 + e / per_row) * ldw + 8 * (e % per_row));
  }
}

template <int H>
__device__ __forceinline__ void nout_stage(int t, int k0, uint16_t (*T)[Nout<H>::TS],
                                           const u32x4 (&r)[Nout<H>::LD]) {
  const int per_row = t < Nout<H>::NB ? k0 / 8 : H / 8;
#pragma unroll
  for (int i = 0; i < Nout<H>::LD; ++i) {
    const int e = threadIdx.x + 256 * i;
    if (e < 32 * per_row) *reinterpret_cast<u32x4*>(&T[e / per_row][8 * (e % per_row)]) = r[i];
  }
}

// acc = W-tile (32 rows x 16 KS columns) x act[0..KS) for this wave's 32 samples.
template <int H, int KS>
__device__ __forceinline__ void nout_mfma(const uint16_t (*T)[Nout<H>::TS],
                                          const bf16x8 (&act)[Nout<H>::KS], f32x16& acc) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const bf16x8 w = *reinterpret_cast<const bf16x8*>(&T[lane & 31][16 * ks + 8 * (lane >> 5)]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, act[ks], acc, 0, 0, 0);
  }
}

// The 16 biases of output block nb this lane's accumulators need (features
// 8q + 4h + u of the block).
__device__ __forceinline__ void nout_bias(const float* bias, int nb, f32x4 (&bv)[4]) {
  const int h = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int q = 0; q < 4; ++q) bv[q] = load_global_f32x4(bias + 32 * nb + 8 * q + 4 * h);
}

// SiLU(acc + bias) of an output block -> the next layer's fragments 2 nb, 2 nb + 1.
__device__ __forceinline__ void nout_epilogue(const f32x4 (&bv)[4], const f32x16& acc,
                                              bf16x8& lo_frag, bf16x8& hi_frag) {
  float v[16];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int u = 0; u < 4; ++u) v[4 * q + u] = silu(acc[4 * q + u] + bv[q][u]);
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

// One layer of NB output blocks (KS k16-steps of input); t advances by NB (a
// layer starts at an even t, so the parity of t is that of the block).  Tile
// t + 1 is in registers (r1 for odd t + 1, r0 for even) when block t starts: its
// global loads were issued a whole block earlier; block t issues tile t + 2's.
template <int H, int KS>
__device__ __forceinline__ void nout_layer(const FusedArgs& a, int k0, const float* bias,
                                           uint16_t (*tile)[32][Nout<H>::TS], int& t, int total,
                                           const bf16x8 (&act)[Nout<H>::KS],
                                           bf16x8 (&nxt)[Nout<H>::KS], u32x4 (&r0)[Nout<H>::LD],
                                           u32x4 (&r1)[Nout<H>::LD]) {
  constexpr int NB = Nout<H>::NB;
  f32x16 prev;
  static_for<0, NB>([&](auto Bc) {
    constexpr int nb = decltype(Bc)::value;
    auto block = [&](u32x4 (&next)[Nout<H>::LD], u32x4 (&after)[Nout<H>::LD]) {  // t + 1, t + 2
      __syncthreads();  // tile t staged; the other buffer is free
      // The previous block's biases are loaded before tile t + 2's weights, so
      // waiting for them does not wait for the weight prefetch.
      f32x4 bv[4];
      if constexpr (nb > 0) nout_bias(bias, nb - 1, bv);
      nout_fetch<H>(a, t + 2, total, k0, after);
      f32x16 cur;
      nout_mfma<H, KS>(tile[t & 1], act, cur);
      if constexpr (nb > 0) nout_epilogue(bv, prev, nxt[2 * nb - 2], nxt[2 * nb - 1]);
      if (t + 1 < total) nout_stage<H>(t + 1, k0, tile[(t + 1) & 1], next);
      prev = cur;
      ++t;
    };
    if constexpr (nb & 1)
      block(r0, r1);
    else
      block(r1, r0);
  });
  f32x4 bv[4];
  nout_bias(bias, NB - 1, bv);
  nout_epilogue(bv, prev, nxt[2 * NB - 2], nxt[2 * NB - 1]);
}

// K0S = k16 steps of the padded input width (round64(w0) / 16).
template <int H, int K0S>
__global__ void __launch_bounds__(256, H == 256 ? 2 : 1) mlp_fused_forward_nout(const FusedArgs a) {
  using NO = Nout<H>;
  __shared__ __align__(16) uint16_t tile[2][32][NO::TS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, ln = lane & 31, h = lane >> 5;
  const int64_t m = static_cast<int64_t>(blockIdx.x) * 128 + wave * 32 + ln;  // this lane's sample
  constexpr int k0 = 16 * K0S;
  bf16x8 act[NO::KS], nxt[NO::KS];
#pragma unroll
  for (int ks = 0; ks < NO::KS; ++ks) {
    uint32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float f[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int k = 16 * ks + 8 * h + 2 * j + u;
        f[u] = (ks < K0S && k < a.w0 && m < a.rows) ? a.x[m * a.w0 + k] : 0.0f;
      }
      d[j] = pack_bf16(f[0], f[1]);
    }
    act[ks] = __builtin_bit_cast(bf16x8, u32x4{d[0], d[1], d[2], d[3]});
  }
  const int total = NO::NB * (a.layers - 1) + 1;
  int t = 0;
  u32x4 r0[NO::LD], r1[NO::LD];
  nout_fetch<H>(a, 0, total, k0, r0);
  nout_stage<H>(0, k0, tile[0], r0);
  nout_fetch<H>(a, 1, total, k0, r1);  // tile 1 in flight
  nout_layer<H, K0S>(a, k0, layer_b(a, 0), tile, t, total, act, nxt, r0, r1);
  for (int l = 1; l + 1 < a.layers; ++l) {
#pragma unroll
    for (int ks = 0; ks < NO::KS; ++ks) act[ks] = nxt[ks];
    nout_layer<H, NO::KS>(a, k0, layer_b(a, l), tile, t, total, act, nxt, r0, r1);
  }
  __syncthreads();  // output layer tile staged
  f32x16 out;
  nout_mfma<H, NO::KS>(tile[t & 1], nxt, out);
  const float* bias = layer_b(a, a.layers - 1);
  if (m < a.rows) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n = 8 * (r >> 2) + 4 * h + (r & 3);
      if (n < a.w_out) a.y[m * a.w_out + n] = out[r] + bias[n];
    }
  }
}

// Padded bf16 copy [round32(out)][round64(in)] of an fp32 [out][in] matrix.
__global__ void pad_bf16(const float* w, int out, int in, int out_p, int in_p, uint16_t* dst) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= int64_t(out_p) * in_p) return;
  const int n = static_cast<int>(i / in_p), k = static_cast<int>(i % in_p);
  dst[i] = (n < out && k < in) ? f32_to_bf16_bits(w[int64_t(n) * in + k]) : uint16_t(0);
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
  bool wp_stale = true;       // Wp older than the master parameters
  uint16_t* Wp_all = nullptr;  // padded bf16 copies for the fused kernel, one allocation
  std::vector<uint16_t*> Wp;   // layer l's copy inside Wp_all
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
  f(m->Wp_all);
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
  // the fused inference kernel covers [w0, H, ..., H, w_out <= 32] with H = 256 and
  // w0 <= 64, or H = 512 and w0 <= 128
  {
    const int h = m->L >= 2 ? widths[1] : 0;
    bool ok = ((h == 256 && widths[0] <= 64) || (h == 512 && widths[0] <= 128)) &&
              widths[m->L] <= 32 && m->L <= kFusedMaxLayers;
    for (int l = 1; l < m->L && ok; ++l) ok = widths[l] == h;
    m->fused_h = ok ? h : 0;
  }
  if (m->fused_h) {
    // Layers back to back: [H][round64(w0)], then (L - 2) x [H][H], then [32][H], so
    // the kernel finds layer l >= 1 at a uniform stride from the first.
    m->Wp.assign(m->L, nullptr);
    int64_t total = 0;
    std::vector<int64_t> off(m->L);
    for (int l = 0; l < m->L; ++l) {
      off[l] = total;
      total += static_cast<int64_t>(round_up(widths[l + 1], 32)) * round_up(widths[l], 64);
    }
    if (rc == CE_OK) rc = dalloc(&m->Wp_all, total);
    for (int l = 0; l < m->L && rc == CE_OK; ++l) m->Wp[l] = m->Wp_all + off[l];
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
  const int64_t chunk = 65536;
  for (size_t b0 = 0; b0 < batch; b0 += chunk) {
    const int64_t B = static_cast<int64_t>(std::min<size_t>(chunk, batch - b0));
    int rc = ensure_batch(m, B);
    if (rc != CE_OK) return rc;
    MLP_HIP(hipMemcpyAsync(m->xin, x_host + b0 * m->widths[0], B * m->widths[0] * sizeof(float),
                           hipMemcpyHostToDevice, m->stream));
    ce::clear_stale_error();
    rc = (m->fused_h && m->use_fused) ? ce::mlp_forward_device(m, m->xin, B, m->out, m->stream)
                                      : forward(m, m->xin, B, false, m->out, m->stream);
    if (rc != CE_OK) return rc;
    MLP_HIP(hipMemcpyAsync(y_host + b0 * m->widths[m->L], m->out,
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

int mlp_forward_device(ce_mlp* m, const float* x, int64_t B, float* y, hipStream_t s) {
  ce::clear_stale_error();
  if (m->fused_h && m->use_fused && B > 0) {
    if (m->wp_stale) {
      for (int l = 0; l < m->L; ++l) {
        const int in = m->widths[l], out = m->widths[l + 1];
        const int op = round_up(out, 32), ip = round_up(in, 64);
        hipLaunchKernelGGL(pad_bf16, dim3(blocks_for(int64_t(op) * ip)), dim3(256), 0, s,
                           m->P + m->w_off[l], out, in, op, ip, m->Wp[l]);
      }
      m->wp_stale = false;
    }
    FusedArgs a{};
    a.x = x;
    a.rows = B;
    a.w0 = m->widths[0];
    a.layers = m->L;
    a.w_out = m->widths[m->L];
    a.W = m->Wp_all;
    for (int l = 0; l < m->L; ++l) a.b[l] = m->P + m->b_off[l];
    a.y = y;
    // 32 samples per wave, 4 waves per workgroup, output-block-major
    const unsigned grid = static_cast<unsigned>((B + 127) / 128);
    if (m->fused_h == 256)
      hipLaunchKernelGGL((mlp_fused_forward_nout<256, 4>), dim3(grid), dim3(256), 0, s, a);
    else if (round_up(a.w0, 64) == 64)
      hipLaunchKernelGGL((mlp_fused_forward_nout<512, 4>), dim3(grid), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((mlp_fused_forward_nout<512, 8>), dim3(grid), dim3(256), 0, s, a);
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
