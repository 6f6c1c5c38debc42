// nn_launch.hpp — host-side launchers of the warm-start network kernels
// (nn_kernels.hip).  BF16 tensors are passed as uint16_t storage.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbmnn {

typedef uint16_t u16;

// Input of a 3x3 conv: NHWC tensor in0 (C0 channels, B x H x W), optionally
// concatenated with the nearest-upsampled NHWC tensor in1 (C1 channels,
// B x ceil(H/2) x ceil(W/2)).  C0 and C1 are multiples of 32.
struct ConvSrc {
  const u16* in0 = nullptr;
  int C0 = 0;
  const u16* in1 = nullptr;
  int C1 = 0;
  int B = 0, H = 0, W = 0;
  int dil = 1;  // dilation of the 3x3 stencil
};

// out = act(conv3x3(src, w) + bias); w is [Cout][9][C0 + C1], bias may be null.
hipError_t conv3x3(const ConvSrc& s, const u16* w, const float* bias, u16* out, int Cout,
                   bool relu, hipStream_t st);
// Weight-gradient partial sums, one [Cout][9][Cin] block per chunk of
// grp_per_chunk groups of 64 consecutive pixels (Cout, Cin multiples of 32).
// (co, ci) tile edge of a workgroup: 64 when both counts are multiples of 64
int wgrad_tile(int Cout, int Cin);
int wgrad_groups(const ConvSrc& s);
int wgrad_chunks(const ConvSrc& s, int grp_per_chunk);
hipError_t wgrad(const ConvSrc& s, const u16* dout, int Cout, float* part, int grp_per_chunk,
                 hipStream_t st);
hipError_t colsum_part(const u16* g, int P, int C, int px_per_chunk, float* part, hipStream_t st);
hipError_t reduce_chunks(const float* part, int nchunk, long n, float* out, hipStream_t st);
hipError_t flip_w(const u16* w, u16* wt, int Cout, int Cin, hipStream_t st);
hipError_t to_bf16(const float* a, u16* b, long n, hipStream_t st);
hipError_t relu_bwd(u16* g, const u16* y, long n, hipStream_t st);
// Divergence penalty of the predicted velocity: d (per pixel) and its
// per-256-pixel sums of d^2 into part; dy += the gradient of (g / 2) sum d^2.
hipError_t div_penalty(const float* y, const float* mask, float* d, float* part, float* dy, int B,
                       int H, int W, float su, float sv, float g, hipStream_t st);
hipError_t pool(const u16* in, u16* out, int B, int H, int W, int C, hipStream_t st);
hipError_t skip_grad(const u16* dcat, int Ccat, const u16* dpool, const u16* s, u16* ds, int B,
                     int H, int W, int C, hipStream_t st);
hipError_t up_grad(const u16* dcat, int Ccat, int C0, const u16* y, u16* dy, int B, int H, int W,
                   int C1, hipStream_t st);
hipError_t head_fwd(const u16* x, const float* w, const float* b, float* out, int P, int C, int K,
                    hipStream_t st);
hipError_t mse(const float* y, const float* t, const float* mask, float* dy, float* part, int P,
               int K, float inv_norm, hipStream_t st);
hipError_t head_bwd(const float* dy, const float* w, const u16* x, u16* dx, int P, int C, int K,
                    hipStream_t st);
hipError_t head_wgrad_part(const float* dy, const u16* x, float* part, int P, int C, int K,
                           int px_per_chunk, hipStream_t st);
hipError_t adam(float* p, const float* g, float* m, float* v, u16* pb, long n, float lr, float b1,
                float b2, float eps, float c1, float c2, float wd, hipStream_t st);
hipError_t gather(const float* inputs, const float* targets, const int* idx, u16* x, float* t,
                  float* mask, int B, int HW, int Ci, int Cp, hipStream_t st);

}  // namespace lbmnn
