// Learned row-size predictor: a 16 -> 64 -> 64 -> 8 MLP on gfx950 MFMA
// (BF16 operands, FP32 accumulation), with hand-written inference,
// training (forward + backward) and Adam kernels.
//
// The network predicts, per row of C = A*B, the class of the row's exact
// output count (thresholds = the numeric hash-table capacities, then
// 65536, then unbounded).  Its per-row capacity feeds the estimation
// workflow in place of the HyperLogLog merge; rows whose count exceeds the
// predicted capacity overflow and are recomputed exactly, so C never
// depends on the prediction.
//
// A 256-thread workgroup takes 64 rows at a time (inference: persistent,
// the weights staged once per workgroup); each wave runs 16 rows through
// the three layers with v_mfma_f32_16x16x32_bf16.  Activations move between
// layers through LDS (the accumulator layout of one MFMA is not the operand
// layout of the next).  The backward pass recomputes the forward pass,
// forms dZ/dH with MFMA and the weight gradients with MFMA over the
// workgroup's 64 rows (K = rows); per-workgroup partial gradients are
// summed by a reduce kernel in a fixed order, so gradients are
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dev_common.hpp"
#include "launch.hpp"

namespace oc {

#define HIP_TRY(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                    \
                hipGetErrorString(e_));                                       \
      return e_ == hipErrorOutOfMemory ? OC_ERR_NOMEM : OC_ERR_HIP;           \
    }                                                                         \
  } while (0)
#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

// ------------------------------------------------------------ parameters
// Logical layout of the 5768 parameters (FP32 master copy, BF16 copy,
// gradients and Adam moments all use it): W1[f][h], b1[h], W2[h1][h2],
// b2[h2], W3[h][c], b3[c].
constexpr int kF = OC_MLP_FEATURES, kH = OC_MLP_HIDDEN, kC = OC_MLP_CLASSES;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kF * kH, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kC,
              kNParam = kOffB3 + kC;
static_assert(kNParam == OC_MLP_NPARAM, "parameter count");
constexpr int kRowsPerBlock = 64;      // 4 waves x 16 rows
constexpr int kMlpBlock = 256;
constexpr uint32_t kMaxBatch = 1u << 16;  // rows per backward call

// LDS strides (u16 elements): K padded to a multiple of 32 plus 8 to move
// consecutive rows to other banks.
constexpr int kS32 = 40, kS64 = 72;

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// ReLU that never yields -0, so "active" is exactly "bf16 bits != 0"
__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }

// MFMA 16x16x32 operand layout: lane l supplies row/column (l % 16) and the
// eight k values (l / 16) * 8 + v of the 32-wide k block.
__device__ inline int kofs(int lane) { return (lane >> 4) * 8; }

// Operand from a row-major LDS tile T[row][k] (row = l%16 + r0): 8
// contiguous k values.
__device__ inline bf8 ld_rowk(const uint16_t *T, int ld, int r0, int k0) {
  const int l = lane_id();
  const us8 v = *reinterpret_cast<const us8 *>(T + (r0 + (l & 15)) * ld + k0 + kofs(l));
  return __builtin_bit_cast(bf8, v);
}
// Operand whose k index runs over rows of an LDS tile T[k][m]: column
// m0 + l%16 of rows k0 + (l/16)*8 + v (used for gradients, K = rows).
__device__ inline bf8 ld_colk(const uint16_t *T, int ld, int m0, int k0) {
  const int l = lane_id();
  const uint16_t *p = T + (k0 + kofs(l)) * ld + m0 + (l & 15);
  us8 v;
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = p[u * ld];
  return __builtin_bit_cast(bf8, v);
}
__device__ inline f4 mfma(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Shared memory of one workgroup.
struct MlpSmem {
  uint16_t W1t[kH * kS32];   // [h][f]   f < 16, zero to 32
  uint16_t W2t[kH * kS64];   // [h2][h1]
  uint16_t W3t[16 * kS64];   // [c][h]   c < 8, zero rows to 16
  uint16_t X[kRowsPerBlock * kS32];
  uint16_t H1[kRowsPerBlock * kS64];
  uint16_t H2[kRowsPerBlock * kS64];
  float b1[kH], b2[kH], b3[16];
};
struct MlpSmemBwd {
  MlpSmem f;
  uint16_t W2n[kH * kS64];   // [h1][h2]
  uint16_t W3h[kH * kS32];   // [h][c]   c < 8, zero to 32
  uint16_t dZ3[kRowsPerBlock * kS32];
  uint16_t dZ2[kRowsPerBlock * kS64];
  uint16_t dZ1[kRowsPerBlock * kS64];
  float lossrow[kRowsPerBlock];
};

__device__ void stage_weights(MlpSmem &s, const uint16_t *__restrict__ wbf,
                              const float *__restrict__ master) {
  for (int i = threadIdx.x; i < kH * 32; i += blockDim.x) {
    const int h = i / 32, f = i % 32;
    s.W1t[h * kS32 + f] = f < kF ? wbf[kOffW1 + f * kH + h] : (uint16_t)0;
  }
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
    const int h2 = i / kH, h1 = i % kH;
    s.W2t[h2 * kS64 + h1] = wbf[kOffW2 + h1 * kH + h2];
  }
  for (int i = threadIdx.x; i < 16 * kH; i += blockDim.x) {
    const int c = i / kH, h = i % kH;
    s.W3t[c * kS64 + h] = c < kC ? wbf[kOffW3 + h * kC + c] : (uint16_t)0;
  }
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    s.b1[i] = master[kOffB1 + i];
    s.b2[i] = master[kOffB2 + i];
  }
  for (int i = threadIdx.x; i < 16; i += blockDim.x) s.b3[i] = i < kC ? master[kOffB3 + i] : 0.f;
}

// Features of rows [row0, row0+64) into s.X (bf16, zero padding); rows
// beyond n are zero.
__device__ void stage_x(MlpSmem &s, const float *__restrict__ feat, uint64_t row0, uint64_t n) {
  for (int i = threadIdx.x; i < kRowsPerBlock * 32; i += blockDim.x) {
    const int r = i / 32, f = i % 32;
    const uint64_t g = row0 + r;
    const float v = (f < kF && g < n) ? feat[g * kF + f] : 0.f;
    s.X[r * kS32 + f] = f2bf(v);
  }
}

// Features of one wave's 16 rows [g0, g0+16) into s.X rows r0..r0+15,
// columns 0..15 (one float4 per lane; rows beyond n are zero).  The padding
// columns 16..31 are zeroed once by the caller.
__device__ void stage_x_wave(MlpSmem &s, const float *__restrict__ feat, uint64_t g0,
                             uint64_t n, int r0) {
  static_assert(kF == 16, "one float4 per lane covers 16 rows of 16 features");
  const int l = lane_id(), r = l >> 2, f = (l & 3) * 4;
  const uint64_t g = g0 + r;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (g < n) v = *reinterpret_cast<const float4 *>(feat + g * kF + f);
  uint16_t *x = s.X + (r0 + r) * kS32 + f;
  x[0] = f2bf(v.x);
  x[1] = f2bf(v.y);
  x[2] = f2bf(v.z);
  x[3] = f2bf(v.w);
}

// Forward pass of the wave's 16 rows (block rows r0..r0+15): H1, H2 into
// LDS; returns the logits accumulator (lane: class l%16, rows
// r0 + (l/16)*4 + r) with b3 added.
__device__ f4 forward_wave(MlpSmem &s, int r0) {
  const int l = lane_id();
  const int col = l & 15, rb = (l >> 4) * 4;
  for (int nt = 0; nt < 4; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_rowk(s.X, kS32, r0, 0), ld_rowk(s.W1t, kS32, nt * 16, 0), acc);
    const float bb = s.b1[nt * 16 + col];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      s.H1[(r0 + rb + r) * kS64 + nt * 16 + col] = f2bf(relu(acc[r] + bb));
  }
  wave_sync();
  for (int nt = 0; nt < 4; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_rowk(s.H1, kS64, r0, 0), ld_rowk(s.W2t, kS64, nt * 16, 0), acc);
    acc = mfma(ld_rowk(s.H1, kS64, r0, 32), ld_rowk(s.W2t, kS64, nt * 16, 32), acc);
    const float bb = s.b2[nt * 16 + col];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      s.H2[(r0 + rb + r) * kS64 + nt * 16 + col] = f2bf(relu(acc[r] + bb));
  }
  wave_sync();
  f4 z = {0.f, 0.f, 0.f, 0.f};
  z = mfma(ld_rowk(s.H2, kS64, r0, 0), ld_rowk(s.W3t, kS64, 0, 0), z);
  z = mfma(ld_rowk(s.H2, kS64, r0, 32), ld_rowk(s.W3t, kS64, 0, 32), z);
#pragma unroll
  for (int r = 0; r < 4; ++r) z[r] += s.b3[col];
  return z;
}

// Softmax over the 8 class lanes of each row group (lanes with l%16 >= 8
// carry padding and compute values nobody reads).
__device__ inline void softmax8(float z, float &p, float &m, float &sum) {
  const bool valid = (lane_id() & 15) < kC;
  float zz = valid ? z : -1e30f;
  m = zz;
  m = fmaxf(m, __shfl_xor(m, 1));
  m = fmaxf(m, __shfl_xor(m, 2));
  m = fmaxf(m, __shfl_xor(m, 4));
  const float e = valid ? expf(zz - m) : 0.f;
  sum = e;
  sum += __shfl_xor(sum, 1);
  sum += __shfl_xor(sum, 2);
  sum += __shfl_xor(sum, 4);
  p = e / sum;
}

// ----------------------------------------------------------- inference
// MODE 0: logits[n][8].  MODE 1: per-row capacity (estimate) for the
// estimation workflow: the smallest class whose cumulative probability
// reaches q, mapped to its threshold, raised to the longest referenced B
// row and clipped to min(products, column span).
struct EstArgs {
  const uint32_t *products, *max_b, *left, *right;
  uint32_t thr[kC];
  float q;
  uint32_t *est;
};

// Persistent: each workgroup stages the weights once and strides over the
// 64-row groups; within a group every wave loads and runs its own 16 rows,
// so the loop needs no workgroup barrier.
template <int MODE>
__global__ __launch_bounds__(kMlpBlock) void k_mlp_forward(
    const uint16_t *__restrict__ wbf, const float *__restrict__ master,
    const float *__restrict__ feat, uint64_t n, float *__restrict__ logits, EstArgs ea) {
  extern __shared__ __align__(16) uint8_t smem_raw[];
  MlpSmem &s = *reinterpret_cast<MlpSmem *>(smem_raw);
  stage_weights(s, wbf, master);
  for (int i = threadIdx.x; i < kRowsPerBlock * 16; i += blockDim.x)
    s.X[(i / 16) * kS32 + 16 + i % 16] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave, l = lane_id();
  const int r0 = w * 16, col = l & 15, rb = (l >> 4) * 4;
  const uint64_t ngroups = (n + kRowsPerBlock - 1) / kRowsPerBlock;
  for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const uint64_t row0 = grp * kRowsPerBlock;
    stage_x_wave(s, feat, row0 + r0, n, r0);
    wave_sync();
    const f4 z = forward_wave(s, r0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint64_t g = row0 + r0 + rb + r;
      if (MODE == 0) {
        if (col < kC && g < n) logits[g * kC + col] = z[r];
      } else {
        float p, m, sum;
        softmax8(z[r], p, m, sum);
        // inclusive prefix of p over the 8 class lanes of the row
        float cum = p;
#pragma unroll
        for (int d = 1; d < 8; d <<= 1) {
          const float o = __shfl_up(cum, d, 8);
          if ((col & 7) >= d) cum += o;
        }
        const uint64_t below = __ballot(col < kC && cum < ea.q);
        const int cls = min(kC - 1, (int)__popcll((below >> ((l >> 4) * 16)) & 0xFFull));
        if (col == 0 && g < n) {
          const uint32_t P = ea.products[g];
          const uint32_t span = ea.right[g] >= ea.left[g] ? ea.right[g] - ea.left[g] + 1 : 0u;
          const uint32_t ub = min(P, span);
          uint32_t e = cls < kC - 1 ? ea.thr[cls] : ub;
          e = max(e, ea.max_b[g]);
          ea.est[g] = min(e, ub);
        }
      }
    }
    wave_sync();  // this wave's X, H1 and H2 rows are reused by its next group
  }
}

// ------------------------------------------------------------ training
// Gradients of the mean softmax cross-entropy over the n rows of a
// minibatch: per-workgroup partials partial[blk][5768] and loss[blk].
__global__ __launch_bounds__(kMlpBlock) void k_mlp_backward(
    const uint16_t *__restrict__ wbf, const float *__restrict__ master,
    const float *__restrict__ feat, const uint8_t *__restrict__ labels, uint64_t n,
    float inv_n, float *__restrict__ partial, float *__restrict__ ploss) {
  extern __shared__ __align__(16) uint8_t smem_raw[];
  MlpSmemBwd &sb = *reinterpret_cast<MlpSmemBwd *>(smem_raw);
  MlpSmem &s = sb.f;
  stage_weights(s, wbf, master);
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
    const int h1 = i / kH, h2 = i % kH;
    sb.W2n[h1 * kS64 + h2] = wbf[kOffW2 + h1 * kH + h2];
  }
  for (int i = threadIdx.x; i < kH * 32; i += blockDim.x) {
    const int h = i / 32, c = i % 32;
    sb.W3h[h * kS32 + c] = c < kC ? wbf[kOffW3 + h * kC + c] : (uint16_t)0;
  }
  const uint64_t row0 = (uint64_t)blockIdx.x * kRowsPerBlock;
  stage_x(s, feat, row0, n);
  __syncthreads();
  const int w = threadIdx.x / kWave, l = lane_id();
  const int r0 = w * 16, col = l & 15, rb = (l >> 4) * 4;
  const f4 z = forward_wave(s, r0);
  // dZ3 = (softmax - onehot) / n, loss per row
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int br = r0 + rb + r;
    const uint64_t g = row0 + br;
    float p, m, sum;
    softmax8(z[r], p, m, sum);
    const bool valid = g < n;
    const int y = valid ? (int)labels[g] : -1;
    float d = 0.f;
    if (valid && col < kC) d = (p - (col == y ? 1.f : 0.f)) * inv_n;
    sb.dZ3[br * kS32 + col] = f2bf(d);
    sb.dZ3[br * kS32 + 16 + col] = 0;  // k padding of the class dimension
    if (col == (y >= 0 ? y : 0)) sb.lossrow[br] = valid ? (m + logf(sum) - z[r]) : 0.f;
  }
  wave_sync();
  // dZ2 = (dZ3 W3^T) * [H2 > 0]
  for (int nt = 0; nt < 4; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_rowk(sb.dZ3, kS32, r0, 0), ld_rowk(sb.W3h, kS32, nt * 16, 0), acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = (r0 + rb + r) * kS64 + nt * 16 + col;
      sb.dZ2[idx] = f2bf(s.H2[idx] != 0 ? acc[r] : 0.f);
    }
  }
  wave_sync();
  // dZ1 = (dZ2 W2^T) * [H1 > 0]
  for (int nt = 0; nt < 4; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_rowk(sb.dZ2, kS64, r0, 0), ld_rowk(sb.W2n, kS64, nt * 16, 0), acc);
    acc = mfma(ld_rowk(sb.dZ2, kS64, r0, 32), ld_rowk(sb.W2n, kS64, nt * 16, 32), acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = (r0 + rb + r) * kS64 + nt * 16 + col;
      sb.dZ1[idx] = f2bf(s.H1[idx] != 0 ? acc[r] : 0.f);
    }
  }
  __syncthreads();
  // weight gradients over the workgroup's 64 rows (K = rows)
  float *out = partial + (uint64_t)blockIdx.x * kNParam;
  {  // dW2[h1][h2]: wave w takes h1 tile w
    const int mt = w;
    for (int nt = 0; nt < 4; ++nt) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(ld_colk(s.H1, kS64, mt * 16, 0), ld_colk(sb.dZ2, kS64, nt * 16, 0), acc);
      acc = mfma(ld_colk(s.H1, kS64, mt * 16, 32), ld_colk(sb.dZ2, kS64, nt * 16, 32), acc);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[kOffW2 + (mt * 16 + rb + r) * kH + nt * 16 + col] = acc[r];
    }
  }
  {  // dW1[f][h]: wave w takes h tile w
    const int nt = w;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_colk(s.X, kS32, 0, 0), ld_colk(sb.dZ1, kS64, nt * 16, 0), acc);
    acc = mfma(ld_colk(s.X, kS32, 0, 32), ld_colk(sb.dZ1, kS64, nt * 16, 32), acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) out[kOffW1 + (rb + r) * kH + nt * 16 + col] = acc[r];
  }
  {  // dW3[h][c]: wave w takes h tile w
    const int mt = w;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_colk(s.H2, kS64, mt * 16, 0), ld_colk(sb.dZ3, kS32, 0, 0), acc);
    acc = mfma(ld_colk(s.H2, kS64, mt * 16, 32), ld_colk(sb.dZ3, kS32, 0, 32), acc);
    if (col < kC) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[kOffW3 + (mt * 16 + rb + r) * kC + col] = acc[r];
    }
  }
  // bias gradients: sums of the bf16 dZ over the rows, in row order
  const int t = threadIdx.x;
  if (t < kH) {
    float a1 = 0.f, a2 = 0.f;
    for (int i = 0; i < kRowsPerBlock; ++i) {
      a1 += bf2f(sb.dZ1[i * kS64 + t]);
      a2 += bf2f(sb.dZ2[i * kS64 + t]);
    }
    out[kOffB1 + t] = a1;
    out[kOffB2 + t] = a2;
  } else if (t < kH + kC) {
    float a3 = 0.f;
    for (int i = 0; i < kRowsPerBlock; ++i) a3 += bf2f(sb.dZ3[i * kS32 + (t - kH)]);
    out[kOffB3 + (t - kH)] = a3;
  } else if (t == kH + kC) {
    float a = 0.f;
    for (int i = 0; i < kRowsPerBlock; ++i) a += sb.lossrow[i];
    ploss[blockIdx.x] = a;
  }
}

// Sum of the per-workgroup partials in workgroup order.
__global__ void k_mlp_reduce(const float *__restrict__ partial, uint32_t nblk,
                             const float *__restrict__ ploss, float inv_n,
                             float *__restrict__ grads, float *__restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNParam) {
    float a = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) a += partial[(uint64_t)b * kNParam + p];
    grads[p] = a;
  }
  if (p == 0) {
    float a = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) a += ploss[b];
    loss[0] = a * inv_n;
  }
}

// Adam on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_mlp_adam(float *__restrict__ master, float *__restrict__ m1,
                           float *__restrict__ m2, const float *__restrict__ grads,
                           uint16_t *__restrict__ wbf, float lr, float beta1, float beta2,
                           float eps, float corr1, float corr2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float g = grads[p];
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float w = master[p] - lr * (a / corr1) / (sqrtf(b / corr2) + eps);
  master[p] = w;
  wbf[p] = f2bf(w);
}

// --------------------------------------------------- features and labels
// Sixteen features per row of C = A*B from the row analysis (all O(1)
// after the 1/16 scaling of the base-2 logarithms).  Every feature is a
// property of the row alone, so a row scores the same in a product of its
// own and inside a batch of many products.
struct FeatArgs {
  const uint32_t *a_rp, *products, *max_b, *avg_b, *left, *right;
  uint64_t rows;
};

__global__ void k_mlp_features(FeatArgs fa, float *__restrict__ feat) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < fa.rows;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const float P = (float)fa.products[r];
    const float nA = (float)(fa.a_rp[r + 1] - fa.a_rp[r]);
    const float mb = (float)fa.max_b[r], ab = (float)fa.avg_b[r];
    const uint32_t lo = fa.left[r], hi = fa.right[r];
    const float span = hi >= lo ? (float)(hi - lo) + 1.f : 0.f;
    const float ub = fminf(P, span);
    const float uni = span > 0.f ? span * (1.f - expf(-P / span)) : 0.f;
    const float lP = log2f(1.f + P), lmb = log2f(1.f + mb);
    const float lub = log2f(1.f + ub), luni = log2f(1.f + uni);
    float *f = feat + r * kF;
    f[0] = lP * 0.0625f;
    f[1] = log2f(1.f + nA) * 0.0625f;
    f[2] = lmb * 0.0625f;
    f[3] = log2f(1.f + ab) * 0.0625f;
    f[4] = log2f(1.f + span) * 0.0625f;
    f[5] = lub * 0.0625f;
    f[6] = luni * 0.0625f;
    f[7] = (lP - lmb) * 0.0625f;
    f[8] = (nA > 0.f && mb > 0.f) ? P / (nA * mb) : 0.f;
    f[9] = span > 0.f ? fminf(P / span, 4.f) * 0.25f : 0.f;  // product density
    f[10] = span > 0.f ? mb / span : 0.f;                     // longest B row / span
    f[11] = (lub - lmb) * 0.0625f;   // room between the bounds mb <= nnz <= ub
    f[12] = (luni - lmb) * 0.0625f;
    f[13] = (lub - luni) * 0.0625f;
    f[14] = 0.f;
    f[15] = 1.f;
  }
}

struct Thr {
  uint32_t t[kC];
};
__global__ void k_mlp_labels(const uint32_t *__restrict__ nnz, uint64_t n, Thr th,
                             uint8_t *__restrict__ labels) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = nnz[i];
    int c = 0;
    while (c < kC - 1 && v > th.t[c]) ++c;
    labels[i] = (uint8_t)c;
  }
}

// ------------------------------------------------------------ host side
static void class_thresholds(uint32_t *t) {
  const Tables &tb = tables();
  for (int c = 0; c < 6; ++c) t[c] = tb.hash_num[c + 1];
  t[6] = 65536;
  t[7] = 0xFFFFFFFFu;
}

template <typename K>
static int set_lds_mlp(K kernel, size_t bytes) {
  if (bytes <= 65536) return OC_OK;
  HIP_TRY(hipFuncSetAttribute((const void *)kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return OC_OK;
}

// Workgroups of the persistent forward pass: every resident one, or one per
// 64-row group when there are fewer groups.
template <typename K>
static uint32_t forward_grid(K kernel, size_t lds, uint64_t n) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMlpBlock, lds) !=
          hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  const uint64_t full = (uint64_t)std::max<uint32_t>(1, tables().num_cu) * (uint32_t)per_cu;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(full, (n + kRowsPerBlock - 1) / kRowsPerBlock));
}

static uint32_t grid_1d(uint64_t n, uint32_t bs) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + bs - 1) / bs, 1u << 16));
}

int launch_mlp_features(const DevCsr &A, const DevCsr &B, const RowInfo &ri, float *feat,
                        hipStream_t s) {
  if (A.rows == 0) return OC_OK;
  FeatArgs fa;
  fa.a_rp = A.rp;
  fa.products = ri.products;
  fa.max_b = ri.max_b;
  fa.avg_b = ri.avg_b;
  fa.left = ri.left;
  fa.right = ri.right;
  fa.rows = A.rows;
  (void)B;
  hipLaunchKernelGGL(k_mlp_features, dim3(grid_1d(A.rows, 256)), dim3(256), 0, s, fa, feat);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_mlp_estimates(const oc_mlp *m, const float *feat, uint64_t n, const RowInfo &ri,
                         double q, uint32_t *est, hipStream_t s);

}  // namespace oc

using namespace oc;

struct oc_mlp {
  float *master = nullptr, *m1 = nullptr, *m2 = nullptr, *grads = nullptr;
  float *partial = nullptr, *ploss = nullptr, *loss = nullptr;
  uint16_t *wbf = nullptr;
  uint64_t steps = 0;
  float last_inv_n = 0.f;
};

namespace oc {
int launch_mlp_estimates(const oc_mlp *m, const float *feat, uint64_t n, const RowInfo &ri,
                         double q, uint32_t *est, hipStream_t s) {
  if (n == 0) return OC_OK;
  OC_REQUIRE((uintptr_t)feat % 16 == 0, "predictor features must be 16-byte aligned");
  EstArgs ea;
  ea.products = ri.products;
  ea.max_b = ri.max_b;
  ea.left = ri.left;
  ea.right = ri.right;
  class_thresholds(ea.thr);
  ea.q = (float)q;
  ea.est = est;
  const size_t lds = sizeof(MlpSmem);
  OC_TRY(set_lds_mlp(k_mlp_forward<1>, lds));
  const uint32_t nblk = forward_grid(k_mlp_forward<1>, lds, n);
  hipLaunchKernelGGL((k_mlp_forward<1>), dim3(nblk), dim3(kMlpBlock), lds, s,
                     m->wbf, m->master, feat, n, nullptr, ea);
  LAUNCH_CHECK();
  return OC_OK;
}
}  // namespace oc

// ------------------------------------------------------------------ C-ABI
static int mlp_upload(oc_mlp *m, const float *host) {
  std::vector<uint16_t> bf(kNParam);
  for (int i = 0; i < kNParam; ++i) bf[i] = f2bf(host[i]);
  HIP_TRY(hipMemcpy(m->master, host, kNParam * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->wbf, bf.data(), kNParam * 2, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->m1, 0, kNParam * 4));
  HIP_TRY(hipMemset(m->m2, 0, kNParam * 4));
  m->steps = 0;
  return OC_OK;
}

extern "C" oc_mlp *oc_mlp_create(uint64_t seed) {
  oc_mlp *m = new oc_mlp();
  const uint32_t maxblk = kMaxBatch / kRowsPerBlock;
  bool ok = hipMalloc((void **)&m->master, kNParam * 4) == hipSuccess &&
            hipMalloc((void **)&m->m1, kNParam * 4) == hipSuccess &&
            hipMalloc((void **)&m->m2, kNParam * 4) == hipSuccess &&
            hipMalloc((void **)&m->grads, kNParam * 4) == hipSuccess &&
            hipMalloc((void **)&m->partial, (size_t)maxblk * kNParam * 4) == hipSuccess &&
            hipMalloc((void **)&m->ploss, maxblk * 4) == hipSuccess &&
            hipMalloc((void **)&m->loss, 4) == hipSuccess &&
            hipMalloc((void **)&m->wbf, kNParam * 2) == hipSuccess;
  if (!ok) {
    set_error("oc_mlp_create: device allocation failed");
    oc_mlp_destroy(m);
    return nullptr;
  }
  // He-normal weights (fan-in), zero biases
  std::vector<float> w(kNParam, 0.f);
  SplitMix64 g(seed);
  auto normal = [&]() {
    double u1 = g.uniform(), u2 = g.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < kF * kH; ++i) w[kOffW1 + i] = (float)(normal() * std::sqrt(2.0 / kF));
  for (int i = 0; i < kH * kH; ++i) w[kOffW2 + i] = (float)(normal() * std::sqrt(2.0 / kH));
  for (int i = 0; i < kH * kC; ++i) w[kOffW3 + i] = (float)(normal() * std::sqrt(1.0 / kH));
  if (mlp_upload(m, w.data()) != OC_OK) {
    oc_mlp_destroy(m);
    return nullptr;
  }
  return m;
}

extern "C" void oc_mlp_destroy(oc_mlp *m) {
  if (!m) return;
  for (void *p : {(void *)m->master, (void *)m->m1, (void *)m->m2, (void *)m->grads,
                  (void *)m->partial, (void *)m->ploss, (void *)m->loss, (void *)m->wbf})
    if (p) (void)hipFree(p);
  delete m;
}

extern "C" int oc_mlp_get_params(const oc_mlp *m, float *host) {
  OC_REQUIRE(m && host, "oc_mlp_get_params: null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host, m->master, kNParam * 4, hipMemcpyDeviceToHost));
  return OC_OK;
}

extern "C" int oc_mlp_set_params(oc_mlp *m, const float *host) {
  OC_REQUIRE(m && host, "oc_mlp_set_params: null argument");
  for (int i = 0; i < kNParam; ++i)
    OC_REQUIRE(std::isfinite(host[i]), "oc_mlp_set_params: non-finite parameter %d", i);
  return mlp_upload(m, host);
}

extern "C" int oc_mlp_get_grads(const oc_mlp *m, float *host, float *loss) {
  OC_REQUIRE(m && host, "oc_mlp_get_grads: null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host, m->grads, kNParam * 4, hipMemcpyDeviceToHost));
  if (loss) HIP_TRY(hipMemcpy(loss, m->loss, 4, hipMemcpyDeviceToHost));
  return OC_OK;
}

static const char kModelMagic[8] = {'O', 'C', 'M', 'L', 'P', '1', 0, 0};

extern "C" int oc_mlp_save(const oc_mlp *m, const char *path) {
  OC_REQUIRE(m && path, "oc_mlp_save: null argument");
  std::vector<float> w(kNParam);
  OC_TRY(oc_mlp_get_params(m, w.data()));
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    set_error("could not open \"%s\" for writing", path);
    return OC_ERR_IO;
  }
  const uint32_t np = kNParam;
  bool ok = std::fwrite(kModelMagic, 8, 1, f) == 1 && std::fwrite(&np, 4, 1, f) == 1 &&
            std::fwrite(w.data(), 4, kNParam, f) == (size_t)kNParam;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_error("write to \"%s\" failed", path);
    return OC_ERR_IO;
  }
  return OC_OK;
}

extern "C" int oc_mlp_load(oc_mlp *m, const char *path) {
  OC_REQUIRE(m && path, "oc_mlp_load: null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    set_error("could not open model \"%s\"", path);
    return OC_ERR_IO;
  }
  char magic[8];
  uint32_t np = 0;
  std::vector<float> w(kNParam);
  bool ok = std::fread(magic, 8, 1, f) == 1 && std::fread(&np, 4, 1, f) == 1;
  ok = ok && std::memcmp(magic, kModelMagic, 8) == 0 && np == (uint32_t)kNParam;
  ok = ok && std::fread(w.data(), 4, kNParam, f) == (size_t)kNParam;
  std::fclose(f);
  if (!ok) {
    set_error("\"%s\" is not a %d-parameter row-size model", path, kNParam);
    return OC_ERR_IO;
  }
  return oc_mlp_set_params(m, w.data());
}

extern "C" int oc_mlp_features(const oc_csr *dA, const oc_csr *dB, float *feat, void *stream) {
  OC_REQUIRE(dA && dB && feat && dA->on_device && dB->on_device && dA->off_bytes == 4 &&
                 dB->off_bytes == 4,
             "oc_mlp_features: need device matrices with 32-bit offsets");
  OC_REQUIRE(dA->cols == dB->rows, "Incompatible matrix dimensions!");
  const uint64_t m = dA->rows;
  if (m == 0) return OC_OK;
  hipStream_t s = (hipStream_t)stream;
  uint32_t *buf = nullptr;
  HIP_TRY(hipMallocAsync((void **)&buf, (5 * m + 64) * 4, s));
  RowInfo ri{buf, buf + m, buf + 2 * m, buf + 3 * m, buf + 4 * m};
  const DevCsr A{dA->rows, dA->cols, dA->nnz, (const uint32_t *)dA->row_offsets,
                 dA->col_ids, dA->values, dA->vtype};
  const DevCsr B{dB->rows, dB->cols, dB->nnz, (const uint32_t *)dB->row_offsets,
                 dB->col_ids, dB->values, dB->vtype};
  int rc = launch_analysis(A, B, ri, buf + 5 * m, buf + 5 * m + 1, nullptr, s);
  if (rc == OC_OK) rc = launch_mlp_features(A, B, ri, feat, s);
  (void)hipFreeAsync(buf, s);
  OC_TRY(rc);
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_mlp_labels(const uint32_t *nnz, uint64_t n, uint8_t *labels, void *stream) {
  OC_REQUIRE((nnz && labels) || n == 0, "oc_mlp_labels: null argument");
  if (n == 0) return OC_OK;
  hipStream_t s = (hipStream_t)stream;
  Thr th;
  class_thresholds(th.t);
  hipLaunchKernelGGL(k_mlp_labels, dim3(grid_1d(n, 256)), dim3(256), 0, s, nnz, n, th, labels);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_mlp_class_thresholds(uint32_t *thr) {
  OC_REQUIRE(thr, "oc_mlp_class_thresholds: null argument");
  OC_TRY(tables_init());
  class_thresholds(thr);
  return OC_OK;
}

extern "C" int oc_mlp_forward(const oc_mlp *m, const float *feat, uint64_t n, float *logits,
                              void *stream) {
  OC_REQUIRE(m && ((feat && logits) || n == 0), "oc_mlp_forward: null argument");
  if (n == 0) return OC_OK;
  OC_REQUIRE((uintptr_t)feat % 16 == 0, "oc_mlp_forward: features must be 16-byte aligned");
  OC_TRY(tables_init());
  hipStream_t s = (hipStream_t)stream;
  const size_t lds = sizeof(MlpSmem);
  OC_TRY(set_lds_mlp(k_mlp_forward<0>, lds));
  const uint32_t nblk = forward_grid(k_mlp_forward<0>, lds, n);
  EstArgs ea;
  std::memset(&ea, 0, sizeof(ea));
  hipLaunchKernelGGL((k_mlp_forward<0>), dim3(nblk), dim3(kMlpBlock), lds, s,
                     m->wbf, m->master, feat, n, logits, ea);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

static int mlp_backward_async(oc_mlp *m, const float *feat, const uint8_t *labels, uint64_t n,
                              hipStream_t s) {
  const uint32_t nblk = (uint32_t)((n + kRowsPerBlock - 1) / kRowsPerBlock);
  const size_t lds = sizeof(MlpSmemBwd);
  OC_TRY(set_lds_mlp(k_mlp_backward, lds));
  const float inv_n = 1.f / (float)n;
  hipLaunchKernelGGL(k_mlp_backward, dim3(nblk), dim3(kMlpBlock), lds, s, m->wbf, m->master,
                     feat, labels, n, inv_n, m->partial, m->ploss);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, s, m->partial,
                     nblk, m->ploss, inv_n, m->grads, m->loss);
  LAUNCH_CHECK();
  return OC_OK;
}

static int mlp_step_async(oc_mlp *m, float lr, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m->steps += 1;
  const float c1 = (float)(1.0 - std::pow((double)b1, (double)m->steps));
  const float c2 = (float)(1.0 - std::pow((double)b2, (double)m->steps));
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, s, m->master,
                     m->m1, m->m2, m->grads, m->wbf, lr, b1, b2, eps, c1, c2);
  LAUNCH_CHECK();
  return OC_OK;
}

extern "C" int oc_mlp_backward(oc_mlp *m, const float *feat, const uint8_t *labels, uint64_t n,
                               void *stream) {
  OC_REQUIRE(m && feat && labels, "oc_mlp_backward: null argument");
  OC_REQUIRE(n > 0 && n <= kMaxBatch, "oc_mlp_backward: batch of %llu rows (1..%u allowed)",
             (unsigned long long)n, kMaxBatch);
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(mlp_backward_async(m, feat, labels, n, s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_mlp_step(oc_mlp *m, float lr, void *stream) {
  OC_REQUIRE(m && lr > 0.f, "oc_mlp_step: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(mlp_step_async(m, lr, s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_mlp_train(oc_mlp *m, const float *feat, const uint8_t *labels, uint64_t n,
                            uint32_t batch, uint32_t epochs, float lr, float *epoch_loss,
                            void *stream) {
  OC_REQUIRE(m && feat && labels && n > 0, "oc_mlp_train: bad arguments");
  OC_REQUIRE(batch > 0 && batch <= kMaxBatch, "oc_mlp_train: batch must be 1..%u", kMaxBatch);
  hipStream_t s = (hipStream_t)stream;
  for (uint32_t e = 0; e < epochs; ++e) {
    double sum = 0.0;
    uint64_t seen = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      const uint64_t nb = std::min<uint64_t>(batch, n - b0);
      OC_TRY(mlp_backward_async(m, feat + b0 * kF, labels + b0, nb, s));
      OC_TRY(mlp_step_async(m, lr, s));
      float l = 0.f;
      HIP_TRY(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      sum += (double)l * (double)nb;
      seen += nb;
    }
    if (epoch_loss) epoch_loss[e] = (float)(sum / (double)seen);
  }
  return OC_OK;
}

extern "C" int oc_mlp_estimates(const oc_mlp *m, const oc_csr *dA, const oc_csr *dB,
                                double quantile, uint32_t *est, void *stream) {
  OC_REQUIRE(m && dA && dB && est && dA->on_device && dB->on_device && dA->off_bytes == 4 &&
                 dB->off_bytes == 4,
             "oc_mlp_estimates: need device matrices with 32-bit offsets");
  OC_REQUIRE(quantile > 0.0 && quantile <= 1.0, "oc_mlp_estimates: quantile in (0, 1]");
  const uint64_t rows = dA->rows;
  if (rows == 0) return OC_OK;
  hipStream_t s = (hipStream_t)stream;
  uint32_t *buf = nullptr;
  float *feat = nullptr;
  HIP_TRY(hipMallocAsync((void **)&buf, (5 * rows + 64) * 4, s));
  HIP_TRY(hipMallocAsync((void **)&feat, rows * kF * 4, s));
  RowInfo ri{buf, buf + rows, buf + 2 * rows, buf + 3 * rows, buf + 4 * rows};
  const DevCsr A{dA->rows, dA->cols, dA->nnz, (const uint32_t *)dA->row_offsets,
                 dA->col_ids, dA->values, dA->vtype};
  const DevCsr B{dB->rows, dB->cols, dB->nnz, (const uint32_t *)dB->row_offsets,
                 dB->col_ids, dB->values, dB->vtype};
  int rc = launch_analysis(A, B, ri, buf + 5 * rows, buf + 5 * rows + 1, nullptr, s);
  if (rc == OC_OK) rc = launch_mlp_features(A, B, ri, feat, s);
  if (rc == OC_OK) rc = launch_mlp_estimates(m, feat, rows, ri, quantile, est, s);
  (void)hipFreeAsync(feat, s);
  (void)hipFreeAsync(buf, s);
  OC_TRY(rc);
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}
