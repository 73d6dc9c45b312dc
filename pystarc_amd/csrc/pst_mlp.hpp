// Surrogate MLP: sizes, flat FP32 parameter offsets ([out][in] per layer)
// and device buffers.
#pragma once
#include <hip/hip_runtime.h>

namespace pst {

constexpr int kMlpIn = 9;
constexpr int kMlpWidth = 256;
constexpr int kMlpOut = 13;
constexpr int kOffW0 = 0;
constexpr int kOffB0 = kOffW0 + kMlpWidth * kMlpIn;
constexpr int kOffW1 = kOffB0 + kMlpWidth;
constexpr int kOffB1 = kOffW1 + kMlpWidth * kMlpWidth;
constexpr int kOffW2 = kOffB1 + kMlpWidth;
constexpr int kOffB2 = kOffW2 + kMlpWidth * kMlpWidth;
constexpr int kOffW3 = kOffB2 + kMlpWidth;
constexpr int kOffB3 = kOffW3 + kMlpWidth * kMlpWidth;
constexpr int kOffW4 = kOffB3 + kMlpWidth;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpWidth;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 203,277

constexpr int kMlpInLd = 16;   // FP32 feature row stride (9 used)
// rows of one fused-inference tile; every row buffer (features, outputs) is
// allocated in a multiple of it
constexpr int kInferRows = 128;
// rows per partial column sum of the bias gradients; the partial-sum buffer
// holds 256 floats per kColRows rows
constexpr int kColRows = 64;

struct MlpBufs {
  int max_rows;  // padded to kInferRows
  __bf16 *w0, *w1, *w2, *w3, *w4;
  float* w0f;                 // [256][9] FP32 first-layer weights
  float* x0f;                 // [rows][16] FP32 features (9 used)
  float *b0, *b1, *b2, *b3, *b4;
  __bf16* x0;                 // [rows][64]
  __bf16 *h1, *h2, *h3, *h4;  // [rows][256]
  float *z1, *z2, *z3, *z4;   // [rows][256] (train mode)
  float* y;                   // [rows][64]
  __bf16* gy;                 // [rows][64]
  __bf16 *g1, *g2, *g3, *g4;  // [rows][256]
  float* gw;                  // padded weight gradients
  float* gcol;                // [rows / kColRows][256] partial column sums (bias gradients)
};

// Operands of the fused inference forward of one network.
struct MlpInfer {
  const float* x0f;                 // [rows][16] FP32 features
  const float* w0f;                 // [256][9] FP32 first-layer weights
  const __bf16 *w1, *w2, *w3, *w4;
  const float *b0, *b1, *b2, *b3, *b4;
  const int* m_dev;                 // live rows, or null
  int rows;                         // rows when m_dev is null
  float* y;                         // [rows][64] outputs (columns 0..15 written)
};
MlpInfer mlp_infer_args(const MlpBufs& m, int Bp, const int* m_dev, float* y);
// L networks' inference in one launch: device descriptors P[L], workgroup
// partition start[L + 1] (kInferRows-row tiles), blocks = start[L].
hipError_t launch_mlp_infer_group(const MlpInfer* P, const int* start, int L, int blocks,
                                  hipStream_t st);

hipError_t mlp_cast(MlpBufs& m, const float* params, hipStream_t st);
hipError_t mlp_forward(const MlpBufs& m, int B, const float* x, float* y_out, bool train,
                       hipStream_t st);
hipError_t mlp_forward_packed(const MlpBufs& m, int Bp, const int* m_dev, bool train, float* y,
                              hipStream_t st);
hipError_t mlp_forward_layers(const MlpBufs& m, int Bp, const int* m_dev, bool train, float* y,
                              hipStream_t st);
hipError_t mlp_backward(const MlpBufs& m, int B, const float* gy, float* gparams, hipStream_t st);

}  // namespace pst
