// Learned surrogate of the grid + Born + far-field interaction: a
// 9 -> 256 -> 256 -> 256 -> 256 -> 13 SiLU MLP, forward and backward on
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulate).  No reference counterpart.
//
// Layout: parameters arrive as one FP32 vector in [out][in] order
// (W0 256x9, b0, W1 256x256, b1, W2, b2, W3, b3, W4 13x256, b4).  The first
// layer (9 inputs) runs on the FP32 units with FP32 inputs and weights
// (BF16 inputs would round the pose to 1 part in 256: ~0.2 A at the b-sphere),
// the other layers on MFMA with padded BF16 copies: W1..W3 [256][256], W4p
// [64][256] (outputs padded 13 -> 64); activations are BF16 [rows][256]
// (rows padded to 64), pre-activations FP32.  The backward keeps a BF16
// copy of the inputs, W0p [256][64], for the first layer's weight gradient.
//
// One GEMM kernel serves every product: a 256-thread workgroup owns a 64x64
// output tile, stages 64x32 slices of A and B in LDS as [row][k] with k
// contiguous, and each wave runs a 2x2 block of 16x16x32 MFMAs.  Transposed
// operands are transposed on their way into LDS.  Epilogues fuse bias,
// SiLU, SiLU' and the FP32 weight-gradient accumulation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pst_device.hpp"
#include "pst_kernels.hpp"
#include "pst_mlp.hpp"

namespace pst {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// SiLU z sigmoid(z) with the hardware reciprocal (1 ulp) in place of an
// IEEE division: the activation epilogues, not the MFMAs, bound the fused
// inference, and the division's correction sequence doubled their cost.
__device__ inline float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu_f(float z) { return z * sigmoid_f(z); }
__device__ inline float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}

// First layer on the FP32 units: z = sum_k x_k W0[n][k] (k = 0..8 in order,
// fused multiply-adds) + b0[n].  Every path computes it with this function,
// so inference and training see the same bits.
__device__ __forceinline__ float layer0_z(const float* x, const float* w, float b) {
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) acc = fmaf(x[k], w[k], acc);
  return acc + b;
}

// Training-path first layer: Z (train) and H = bf16(SiLU(Z)) for rows < *m_dev.
template <bool TRAIN>
__global__ void k_mlp_layer0(const float* x0f, const float* w0f, const float* b0, int rows,
                             const int* m_dev, float* Z, __bf16* H) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(t / kMlpWidth), n = (int)(t % kMlpWidth);
  const int live = m_dev ? *m_dev : rows;
  if (row >= rows || row >= ((live + 63) & ~63)) return;
  float x[kMlpIn], w[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    x[k] = x0f[(size_t)row * kMlpInLd + k];
    w[k] = w0f[n * kMlpIn + k];
  }
  const float z = layer0_z(x, w, b0[n]);
  if (TRAIN) Z[(size_t)row * kMlpWidth + n] = z;
  H[(size_t)row * kMlpWidth + n] = (__bf16)silu_f(z);
}

enum { LAY_MK = 0, LAY_KM = 1 };  // A: [m][k] or [k][m];  B: [n][k] (=MK) or [k][n] (=KM)
enum { EPI_HIDDEN = 0, EPI_HIDDEN_INFER = 1, EPI_OUT = 2, EPI_DATA = 3, EPI_WGRAD = 4 };

struct GemmArgs {
  const __bf16* A; int lda;
  const __bf16* B; int ldb;
  int M, N, K;          // M, N multiples of 64; K multiple of 32
  const int* m_dev;     // live rows (rows >= *m_dev are skipped by 64-row tiles), or null
  int k_split;          // K range split across blockIdx.z (weight gradients)
  // epilogue operands
  const float* bias;    // [N]
  float* Z; int ldz;    // pre-activation out (EPI_HIDDEN) / in (EPI_DATA)
  __bf16* H; int ldh;   // bf16 activation / gradient out
  float* Y; int ldy;    // fp32 out (EPI_OUT, EPI_WGRAD accumulate)
};

constexpr int kPad = 8;

template <int AL, int BL, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 As[64][32 + kPad];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[64][32 + kPad];
  // Workgroups go round-robin over the 8 XCDs by linear id.  With the M tiles
  // a multiple of 8 (no K split), launch ids are regrouped so that the N
  // tiles of one M tile (which read the same A rows) run on one XCD and share
  // its L2, while consecutive M tiles still alternate over the XCDs (the
  // rows past *m_dev stay spread evenly).  Each tile's arithmetic is
  // unchanged.
  int bx = blockIdx.x, by = blockIdx.y;
  if (gridDim.z == 1 && (gridDim.y & 7) == 0) {
    const int L = blockIdx.y * gridDim.x + blockIdx.x;
    const int per = 8 * gridDim.x;
    const int grp = L / per, r = L - grp * per;
    by = grp * 8 + (r & 7);
    bx = r >> 3;
  }
  const int tile_n = bx * 64;
  const int tile_m = by * 64;
  if (g.m_dev && tile_m >= *g.m_dev) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  int kbeg = 0, kend = g.K;
  if (g.k_split > 1) {
    const int chunk = ((g.K / 32 + g.k_split - 1) / g.k_split) * 32;
    kbeg = blockIdx.z * chunk;
    kend = min(g.K, kbeg + chunk);
  }
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    if (AL == LAY_MK) {
      const int row = tid >> 2, kc = (tid & 3) * 8;
      *(bf16x8*)&As[row][kc] = *(const bf16x8*)&g.A[(size_t)(tile_m + row) * g.lda + k0 + kc];
    } else {
      const int kr = tid >> 3, mc = (tid & 7) * 8;
      const bf16x8 v = *(const bf16x8*)&g.A[(size_t)(k0 + kr) * g.lda + tile_m + mc];
#pragma unroll
      for (int j = 0; j < 8; ++j) As[mc + j][kr] = v[j];
    }
    if (BL == LAY_MK) {
      const int row = tid >> 2, kc = (tid & 3) * 8;
      *(bf16x8*)&Bs[row][kc] = *(const bf16x8*)&g.B[(size_t)(tile_n + row) * g.ldb + k0 + kc];
    } else {
      const int kr = tid >> 3, nc = (tid & 7) * 8;
      const bf16x8 v = *(const bf16x8*)&g.B[(size_t)(k0 + kr) * g.ldb + tile_n + nc];
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[nc + j][kr] = v[j];
    }
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 a0 = *(const bf16x8*)&As[wm + r][kq];
    bf16x8 a1 = *(const bf16x8*)&As[wm + 16 + r][kq];
    bf16x8 b0 = *(const bf16x8*)&Bs[wn + r][kq];
    bf16x8 b1 = *(const bf16x8*)&Bs[wn + 16 + r][kq];
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    __syncthreads();
  }
  // D layout of the 16x16 MFMA tile: lane holds rows 4*(lane/16)+i, column lane%16
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = tile_m + wm + mt * 16 + (lane >> 4) * 4 + i;
        const int n = tile_n + wn + nt * 16 + (lane & 15);
        const float v = acc[mt][nt][i];
        if (EPI == EPI_HIDDEN) {
          const float z = v + g.bias[n];
          g.Z[(size_t)m * g.ldz + n] = z;
          g.H[(size_t)m * g.ldh + n] = (__bf16)silu_f(z);
        } else if (EPI == EPI_HIDDEN_INFER) {
          g.H[(size_t)m * g.ldh + n] = (__bf16)silu_f(v + g.bias[n]);
        } else if (EPI == EPI_OUT) {
          g.Y[(size_t)m * g.ldy + n] = v + g.bias[n];
        } else if (EPI == EPI_DATA) {
          g.H[(size_t)m * g.ldh + n] = (__bf16)(v * dsilu_f(g.Z[(size_t)m * g.ldz + n]));
        } else {
          atomicAdd(&g.Y[(size_t)m * g.ldy + n], v);
        }
      }
}

template <int AL, int BL, int EPI>
static hipError_t gemm(const GemmArgs& g, hipStream_t st, int splits = 1) {
  GemmArgs a = g;
  a.k_split = splits;
  hipLaunchKernelGGL((k_gemm<AL, BL, EPI>), dim3(g.N / 64, g.M / 64, splits), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---- fused inference forward ---------------------------------------------------------
// All five layers of the inference forward in one kernel: a 256-thread
// workgroup owns kInferRows (128) rows and keeps their activations in LDS as
// BF16 [128][256 + 8] (one buffer: a layer's outputs wait in registers until
// every wave has read its inputs).  Wave w computes neurons 64w..64w+63 of a
// hidden layer for all 128 rows as an 8 x 4 block of 16x16x32 MFMA tiles per
// 32-wide k step, with the weights as the A operand and the activations as
// the B operand (the transposed product W A^T): a lane's four results are
// then four consecutive neurons of one row, stored as one 8-byte LDS write
// instead of four 2-byte ones.  B fragments come from LDS; A fragments
// (weights, [out][in] with k contiguous, shared by every workgroup) straight
// from L2 / L1, issued two k steps ahead of the MFMAs that consume them,
// across layer boundaries too.  128 rows per workgroup read each weight once
// per 128 rows; two workgroups per CU (256 VGPRs) let one's SiLU epilogue
// run beside the other's MFMAs.  Layer 0 runs on the FP32 units (thread t:
// neurons 4 (t % 64) .. + 3 of 32 rows), the output layer as one 16-neuron
// tile (13 outputs), two row tiles per wave.  Each output is the same sum of
// the same products in the same k order as in the layer GEMMs, and every
// activation is rounded to BF16 after bias + SiLU in FP32 by the same
// functions, so outputs are the same bit for bit; rows past *m_dev skip
// their tile.  Row buffers are allocated in multiples of kInferRows, so a
// tile never leaves them.
constexpr int kActLd = kMlpWidth + 8;
constexpr int kInferMT = kInferRows / 16;   // 16-row tiles per workgroup
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Tile `tile` (rows kInferRows tile .. + kInferRows - 1) of one network's inference.
__device__ __forceinline__ void mlp_infer_tile(const MlpInfer& p, int tile) {
  __shared__ __attribute__((aligned(16))) __bf16 act[kInferRows][kActLd];
  __shared__ float xs[kInferRows][kMlpIn];
  const int tile_m = tile * kInferRows;
  const int n_rows = p.m_dev ? *p.m_dev : p.rows;
  if (tile_m >= n_rows) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 8, q4 = (lane >> 4) * 4;
  const int wn = wave * 64;
  const __bf16* ws[3] = {p.w1, p.w2, p.w3};
  const float* bs[3] = {p.b1, p.b2, p.b3};
  // weight fragments of the 24 (layer, k step) pairs of the hidden layers; at
  // most three are live at once (the one in use and two in flight)
  bf16x8 wf[24][4];
  auto load_w = [&](int step) {
    const __bf16* W = ws[step >> 3];
    const int k0 = (step & 7) * 32;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      wf[step][nt] = *(const bf16x8*)&W[(size_t)(wn + nt * 16 + r) * kMlpWidth + k0 + kq];
  };
  load_w(0);
  load_w(1);
  // layer 0 on the FP32 units (features staged once in LDS, read as broadcasts)
  for (int i = threadIdx.x; i < kInferRows * kMlpIn; i += 256)
    xs[i / kMlpIn][i % kMlpIn] = p.x0f[(size_t)(tile_m + i / kMlpIn) * kMlpInLd + i % kMlpIn];
  {
    const int n0 = 4 * lane, rb = wave * (kInferRows / 4);
    float w[4][kMlpIn], bn0[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) w[j][k] = p.w0f[(n0 + j) * kMlpIn + k];
      bn0[j] = p.b0[n0 + j];
    }
    __syncthreads();
#pragma unroll 2
    for (int row = rb; row < rb + kInferRows / 4; ++row) {
      bf16x4 v;
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = (__bf16)silu_f(layer0_z(xs[row], w[j], bn0[j]));
      *(bf16x4*)&act[row][n0] = v;
    }
  }
  __syncthreads();
  // hidden layers 1..3: [256][256] x [128][256]^T
  f32x4 acc[kInferMT][4];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    float bias[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) bias[nt][i] = bs[l][wn + nt * 16 + q4 + i];
#pragma unroll
    for (int mt = 0; mt < kInferMT; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kMlpWidth / 32; ++ks) {
      const int step = l * 8 + ks;
      if (step + 2 < 24) load_w(step + 2);
#pragma unroll
      for (int mt = 0; mt < kInferMT; ++mt) {
        const bf16x8 a = *(const bf16x8*)&act[mt * 16 + r][ks * 32 + kq];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[step][nt], a, acc[mt][nt], 0, 0, 0);
      }
    }
    __syncthreads();   // every wave has read this layer's inputs
#pragma unroll
    for (int mt = 0; mt < kInferMT; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x4 v;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = (__bf16)silu_f(acc[mt][nt][i] + bias[nt][i]);
        *(bf16x4*)&act[mt * 16 + r][wn + nt * 16 + q4] = v;
      }
    __syncthreads();
  }
  // output layer: one 16-neuron tile (13 outputs); wave w takes row tiles
  // 2w and 2w + 1
  constexpr int kOutMT = kInferMT / 4;
  f32x4 o[kOutMT];
#pragma unroll
  for (int j = 0; j < kOutMT; ++j) o[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kMlpWidth / 32; ++ks) {
    const bf16x8 wv = *(const bf16x8*)&p.w4[(size_t)r * kMlpWidth + ks * 32 + kq];
#pragma unroll
    for (int j = 0; j < kOutMT; ++j) {
      const bf16x8 av = *(const bf16x8*)&act[(wave * kOutMT + j) * 16 + r][ks * 32 + kq];
      o[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, av, o[j], 0, 0, 0);
    }
  }
  float b4[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) b4[i] = p.b4[q4 + i];
#pragma unroll
  for (int j = 0; j < kOutMT; ++j) {
    const int row = tile_m + (wave * kOutMT + j) * 16 + r;
    const float4 v = {o[j][0] + b4[0], o[j][1] + b4[1], o[j][2] + b4[2], o[j][3] + b4[3]};
    *(float4*)&p.y[(size_t)row * 64 + q4] = v;
  }
}

__global__ void __launch_bounds__(256, 2) k_mlp_infer(MlpInfer p) { mlp_infer_tile(p, blockIdx.x); }

// Several networks (one per ligand) in one launch: workgroup b runs tile
// b - start[l] of network l for start[l] <= b < start[l + 1].
__global__ void __launch_bounds__(256, 2) k_mlp_infer_group(const MlpInfer* __restrict__ P,
                                                         const int* __restrict__ start, int L) {
  const int b = blockIdx.x;
  int l = 0;
  while (l + 1 < L && b >= start[l + 1]) ++l;
  l = __builtin_amdgcn_readfirstlane(l);
  mlp_infer_tile(P[l], b - __builtin_amdgcn_readfirstlane(start[l]));
}

MlpInfer mlp_infer_args(const MlpBufs& m, int Bp, const int* m_dev, float* y) {
  MlpInfer p;
  p.x0f = m.x0f;
  p.w0f = m.w0f;
  p.w1 = m.w1; p.w2 = m.w2; p.w3 = m.w3; p.w4 = m.w4;
  p.b0 = m.b0; p.b1 = m.b1; p.b2 = m.b2; p.b3 = m.b3; p.b4 = m.b4;
  p.m_dev = m_dev;
  p.rows = Bp;
  p.y = y;
  return p;
}

hipError_t launch_mlp_infer_group(const MlpInfer* P, const int* start, int L, int blocks,
                                  hipStream_t st) {
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_infer_group, dim3(blocks), dim3(256), 0, st, P, start, L);
  return hipGetLastError();
}

// ---- small element-wise kernels --------------------------------------------------
// FP32 parameters -> padded BF16 weights + FP32 biases.
__global__ void k_mlp_cast(const float* p, __bf16* w0, __bf16* w1, __bf16* w2, __bf16* w3,
                           __bf16* w4, float* b0, float* b1, float* b2, float* b3, float* b4,
                           float* w0f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int H = kMlpWidth;
  if (t < H * kMlpIn) w0f[t] = p[kOffW0 + t];
  // W0p [256][64]
  if (t < H * 64) {
    const int o = t / 64, i = t % 64;
    w0[t] = (__bf16)(i < kMlpIn ? p[kOffW0 + o * kMlpIn + i] : 0.0f);
  }
  if (t < H * H) {
    w1[t] = (__bf16)p[kOffW1 + t];
    w2[t] = (__bf16)p[kOffW2 + t];
    w3[t] = (__bf16)p[kOffW3 + t];
  }
  if (t < 64 * H) {
    const int o = t / H, i = t % H;
    w4[t] = (__bf16)(o < kMlpOut ? p[kOffW4 + o * H + i] : 0.0f);
  }
  if (t < H) {
    b0[t] = p[kOffB0 + t];
    b1[t] = p[kOffB1 + t];
    b2[t] = p[kOffB2 + t];
    b3[t] = p[kOffB3 + t];
  }
  if (t < 64) b4[t] = t < kMlpOut ? p[kOffB4 + t] : 0.0f;
}

// (B,9) FP32 features -> [Bp][16] FP32 (the first layer's input) and
// [Bp][64] BF16 (the first layer's weight gradient), zero padded.
__global__ void k_mlp_pack_x(const float* x, int B, int Bp, __bf16* x0, float* x0f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Bp * 64) return;
  const int r = t / 64, c = t % 64;
  const float v = (r < B && c < kMlpIn) ? x[r * kMlpIn + c] : 0.0f;
  x0[t] = (__bf16)v;
  if (c < kMlpInLd) x0f[(size_t)r * kMlpInLd + c] = v;
}

// [Bp][64] FP32 outputs -> (B,13).
__global__ void k_mlp_unpack_y(const float* y, int B, float* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kMlpOut) return;
  out[t] = y[(t / kMlpOut) * 64 + t % kMlpOut];
}

// (B,13) FP32 output gradient -> [Bp][64] BF16, zero padded.
__global__ void k_mlp_pack_gy(const float* gy, int B, int Bp, __bf16* g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Bp * 64) return;
  const int r = t / 64, c = t % 64;
  g[t] = (__bf16)((r < B && c < kMlpOut) ? gy[r * kMlpOut + c] : 0.0f);
}

// Bias gradient: column sums of a BF16 [Bp][ld] gradient into out[n], in two
// passes with a fixed order (no atomics): block b sums rows
// kColRows b .. kColRows b + kColRows - 1 of every column (thread n reads
// column n, so a wave reads whole row segments; a training batch of 16,384
// rows gives 256 blocks, one per CU), then one block adds the blocks'
// partial sums in order.
__global__ void k_colsum_part(const __bf16* g, int rows, int ld, float* part) {
  const int n = threadIdx.x;
  if (n >= ld) return;
  const int r0 = blockIdx.x * kColRows, r1 = min(rows, r0 + kColRows);
  float s = 0.f;
#pragma unroll 16
  for (int r = r0; r < r1; ++r) s += (float)g[(size_t)r * ld + n];
  part[(size_t)blockIdx.x * 256 + n] = s;
}

__global__ void k_colsum_final(const float* part, int n_part, int n_cols, float* out) {
  const int n = threadIdx.x;
  if (n >= n_cols) return;
  float s = 0.f;
#pragma unroll 16
  for (int b = 0; b < n_part; ++b) s += part[(size_t)b * 256 + n];
  out[n] = s;
}

static hipError_t colsum(const __bf16* g, int rows, int ld, int n_cols, float* part, float* out,
                         hipStream_t st) {
  const int n_part = (rows + kColRows - 1) / kColRows;
  hipLaunchKernelGGL(k_colsum_part, dim3(n_part), dim3(256), 0, st, g, rows, ld, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(256), 0, st, part, n_part, n_cols, out);
  return hipGetLastError();
}

// Padded weight gradients -> the flat FP32 [out][in] layout.
__global__ void k_mlp_gather_grads(const float* gw0, const float* gw1, const float* gw2,
                                   const float* gw3, const float* gw4, float* gp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int H = kMlpWidth;
  if (t < H * kMlpIn) gp[kOffW0 + t] = gw0[(t / kMlpIn) * 64 + t % kMlpIn];
  if (t < H * H) {
    gp[kOffW1 + t] = gw1[t];
    gp[kOffW2 + t] = gw2[t];
    gp[kOffW3 + t] = gw3[t];
  }
  if (t < kMlpOut * H) gp[kOffW4 + t] = gw4[t];
}

__global__ void k_zero(float* p, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.f;
}

static inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- host-side orchestration ----------------------------------------------------------
hipError_t mlp_cast(MlpBufs& m, const float* params, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_cast, dim3(nblk(kMlpWidth * kMlpWidth)), dim3(256), 0, st, params,
                     m.w0, m.w1, m.w2, m.w3, m.w4, m.b0, m.b1, m.b2, m.b3, m.b4, m.w0f);
  return hipGetLastError();
}

// Forward on the FP32 features x0f [Bp][16] (training also reads the BF16
// copy x0 in backward); rows >= *m_dev skipped when m_dev is set.
// Inference runs the fused kernel; training the layer GEMMs, which keep the
// pre-activations for the backward pass.
hipError_t mlp_forward_packed(const MlpBufs& m, int Bp, const int* m_dev, bool train, float* y,
                              hipStream_t st) {
  if (!train) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((Bp + kInferRows - 1) / kInferRows), dim3(256), 0, st,
                       mlp_infer_args(m, Bp, m_dev, y));
    return hipGetLastError();
  }
  return mlp_forward_layers(m, Bp, m_dev, train, y, st);
}

// The same forward as five layer GEMMs (training mode keeps Z for backward).
hipError_t mlp_forward_layers(const MlpBufs& m, int Bp, const int* m_dev, bool train, float* y,
                              hipStream_t st) {
  const int H = kMlpWidth;
  __bf16* hs[4] = {m.h1, m.h2, m.h3, m.h4};
  float* zs[4] = {m.z1, m.z2, m.z3, m.z4};
  const __bf16* ws[4] = {m.w0, m.w1, m.w2, m.w3};
  const float* bs[4] = {m.b0, m.b1, m.b2, m.b3};
  if (train)
    hipLaunchKernelGGL((k_mlp_layer0<true>), dim3(nblk((size_t)Bp * H)), dim3(256), 0, st, m.x0f,
                       m.w0f, m.b0, Bp, m_dev, m.z1, m.h1);
  else
    hipLaunchKernelGGL((k_mlp_layer0<false>), dim3(nblk((size_t)Bp * H)), dim3(256), 0, st, m.x0f,
                       m.w0f, m.b0, Bp, m_dev, m.z1, m.h1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const __bf16* in = m.h1;
  int ld_in = H;
  for (int l = 1; l < 4; ++l) {
    GemmArgs g = {};
    g.A = in; g.lda = ld_in; g.B = ws[l]; g.ldb = ld_in;
    g.M = Bp; g.N = H; g.K = ld_in; g.m_dev = m_dev;
    g.bias = bs[l]; g.Z = zs[l]; g.ldz = H; g.H = hs[l]; g.ldh = H;
    e = train ? gemm<LAY_MK, LAY_MK, EPI_HIDDEN>(g, st) : gemm<LAY_MK, LAY_MK, EPI_HIDDEN_INFER>(g, st);
    if (e != hipSuccess) return e;
    in = hs[l];
    ld_in = H;
  }
  GemmArgs g = {};
  g.A = m.h4; g.lda = H; g.B = m.w4; g.ldb = H;
  g.M = Bp; g.N = 64; g.K = H; g.m_dev = m_dev;
  g.bias = m.b4; g.Y = y; g.ldy = 64;
  return gemm<LAY_MK, LAY_MK, EPI_OUT>(g, st);
}

hipError_t mlp_forward(const MlpBufs& m, int B, const float* x, float* y_out, bool train,
                       hipStream_t st) {
  const int Bp = (B + 63) / 64 * 64;
  hipLaunchKernelGGL(k_mlp_pack_x, dim3(nblk((size_t)Bp * 64)), dim3(256), 0, st, x, B, Bp, m.x0,
                     m.x0f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = mlp_forward_packed(m, Bp, nullptr, train, m.y, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_unpack_y, dim3(nblk((size_t)B * kMlpOut)), dim3(256), 0, st, m.y, B,
                     y_out);
  return hipGetLastError();
}

// Backward of the last train-mode forward: gy (B,13) FP32 -> grads (flat FP32).
hipError_t mlp_backward(const MlpBufs& m, int B, const float* gy, float* gparams, hipStream_t st) {
  const int H = kMlpWidth;
  const int Bp = (B + 63) / 64 * 64;
  hipError_t e;
  hipLaunchKernelGGL(k_mlp_pack_gy, dim3(nblk((size_t)Bp * 64)), dim3(256), 0, st, gy, B, Bp, m.gy);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const size_t nw = (size_t)H * 64 * 2 + (size_t)H * H * 3;
  hipLaunchKernelGGL(k_zero, dim3(nblk(nw)), dim3(256), 0, st, m.gw, nw);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  float* gw0 = m.gw;                    // [256][64]
  float* gw1 = gw0 + (size_t)H * 64;    // [256][256]
  float* gw2 = gw1 + (size_t)H * H;
  float* gw3 = gw2 + (size_t)H * H;
  float* gw4 = gw3 + (size_t)H * H;     // [64][256]
  const int splits = max(1, min(64, Bp / 256));
  // layer 4: dW4 = gy^T h4, db4 = sum gy, g_h4 = gy W4, g_z4 = g_h4 * silu'(z4)
  {
    GemmArgs g = {};
    g.A = m.gy; g.lda = 64; g.B = m.h4; g.ldb = H; g.M = 64; g.N = H; g.K = Bp;
    g.Y = gw4; g.ldy = H;
    if ((e = gemm<LAY_KM, LAY_KM, EPI_WGRAD>(g, st, splits)) != hipSuccess) return e;
    if ((e = colsum(m.gy, Bp, 64, kMlpOut, m.gcol, gparams + kOffB4, st)) != hipSuccess) return e;
    GemmArgs d = {};
    d.A = m.gy; d.lda = 64; d.B = m.w4; d.ldb = H; d.M = Bp; d.N = H; d.K = 64;
    d.Z = m.z4; d.ldz = H; d.H = m.g4; d.ldh = H;
    if ((e = gemm<LAY_MK, LAY_KM, EPI_DATA>(d, st)) != hipSuccess) return e;
  }
  __bf16* gz[4] = {m.g1, m.g2, m.g3, m.g4};
  __bf16* hin[4] = {m.x0, m.h1, m.h2, m.h3};
  float* zin[4] = {nullptr, m.z1, m.z2, m.z3};
  const __bf16* ws[4] = {m.w0, m.w1, m.w2, m.w3};
  float* gws[4] = {gw0, gw1, gw2, gw3};
  const int boff[4] = {kOffB0, kOffB1, kOffB2, kOffB3};
  for (int l = 3; l >= 0; --l) {
    const int nin = l == 0 ? 64 : H;
    GemmArgs g = {};
    g.A = gz[l]; g.lda = H; g.B = hin[l]; g.ldb = nin; g.M = H; g.N = nin; g.K = Bp;
    g.Y = gws[l]; g.ldy = nin;
    if ((e = gemm<LAY_KM, LAY_KM, EPI_WGRAD>(g, st, splits)) != hipSuccess) return e;
    if ((e = colsum(gz[l], Bp, H, H, m.gcol, gparams + boff[l], st)) != hipSuccess) return e;
    if (l > 0) {
      GemmArgs d = {};
      d.A = gz[l]; d.lda = H; d.B = ws[l]; d.ldb = H; d.M = Bp; d.N = H; d.K = H;
      d.Z = zin[l]; d.ldz = H; d.H = gz[l - 1]; d.ldh = H;
      if ((e = gemm<LAY_MK, LAY_KM, EPI_DATA>(d, st)) != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL(k_mlp_gather_grads, dim3(nblk((size_t)H * H)), dim3(256), 0, st, gw0, gw1,
                     gw2, gw3, gw4, gparams);
  return hipGetLastError();
}

}  // namespace pst
