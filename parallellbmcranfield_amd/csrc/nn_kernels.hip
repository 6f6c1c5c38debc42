// nn_kernels.hip — kernels of the warm-start network (a 3x3-conv U-Net on a
// square grid): BF16 weights and activations, FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), hand-written backward
// passes and Adam.
//
// Tensors are NHWC (channels contiguous); every conv channel count is a
// multiple of 32, so one 32-deep K slice of the implicit GEMM is one tap of
// 32 consecutive channels: a 16-byte load per lane.  A decoder conv reads a
// virtual concatenation [skip (C0 channels, full size), nearest-upsampled
// coarse tensor (C1 channels, size ceil(H/2) x ceil(W/2))] without
// materialising it.
//
// MFMA 16x16x32 (bf16) operand layout, lane l = 16*kb + r (r = 0..15):
//   A (16 x 32, M x K): lane holds A[r][8kb .. 8kb+7]
//   B (32 x 16, K x N): lane holds B[8kb .. 8kb+7][r]
//   D (16 x 16):        lane holds D[4kb + i][r], i = 0..3
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nn_launch.hpp"

namespace lbmnn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Src {  // virtual concat input
  const bf16* in0;
  int C0;
  const bf16* in1;  // coarse, nearest-upsampled (may be null)
  int C1;
  int B, H, W;      // full size
  int dil;          // tap spacing (dilation) of the 3x3 stencil
};

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int e = 0; e < 8; e++) z[e] = (__bf16)0.0f;
  return z;
}

// 8 channels [ci, ci+8) of the virtual input at (b, y, x); zeros outside.
// Offsets in 32-bit arithmetic: every tensor holds fewer than 2^31 elements
// (lbm_net_create checks), and a 32-bit offset from a uniform base is one
// address register per load.
__device__ inline bf16x8 load8(const Src& s, int b, int y, int x, int ci) {
  if (y < 0 || y >= s.H || x < 0 || x >= s.W) return zero8();
  if (ci < s.C0) {
    const unsigned off = ((unsigned)(b * s.H + y) * (unsigned)s.W + (unsigned)x) * (unsigned)s.C0 +
                         (unsigned)ci;
    return *reinterpret_cast<const bf16x8*>(s.in0 + off);
  }
  const int Hc = (s.H + 1) >> 1, Wc = (s.W + 1) >> 1;
  const unsigned off = ((unsigned)(b * Hc + (y >> 1)) * (unsigned)Wc + (unsigned)(x >> 1)) *
                           (unsigned)s.C1 +
                       (unsigned)(ci - s.C0);
  return *reinterpret_cast<const bf16x8*>(s.in1 + off);
}

// (b, y, x) of pixel p in the flattened (b, y, x) order, and the step that
// moves it n pixels on (n <= W or a few row wraps): pixel coordinates advance
// by additions instead of a division per fetch.
struct PixelCursor {
  int x, y, b;
  __device__ void at(int p, int W, int H) {
    const int q = p / W;
    x = p - q * W;
    b = q / H;
    y = q - b * H;
  }
  __device__ void advance(int n, int W, int H) {
    x += n;
    while (x >= W) {
      x -= W;
      if (++y == H) {
        y = 0;
        b++;
      }
    }
  }
};

__device__ inline float load1(const Src& s, int b, int y, int x, int ci) {
  if (y < 0 || y >= s.H || x < 0 || x >= s.W) return 0.0f;
  if (ci < s.C0) return (float)s.in0[(((size_t)b * s.H + y) * s.W + x) * s.C0 + ci];
  const int Hc = (s.H + 1) >> 1, Wc = (s.W + 1) >> 1;
  return (float)s.in1[(((size_t)b * Hc + (y >> 1)) * Wc + (x >> 1)) * s.C1 + (ci - s.C0)];
}

// ------------------------------------------------------------ conv forward
// out[p][co] = act(bias[co] + sum_{t, ci} in(nbr(p, t))[ci] * w[co][t][ci]),
// nbr(p, t) = p + dil * (t / 3 - 1, t % 3 - 1).
// Implicit GEMM, M = pixels, N = output channels, K = 9 taps x Cin in steps
// of 32 channels of one tap.  Both operands are already in MFMA layout in
// memory (a lane's 8 consecutive K values are 8 consecutive channels: one
// 16-byte load).  Each wave computes TM x TN tiles of 16 x 16 (16 TM pixels
// x 16 TN channels); the 4 waves of a workgroup take consecutive pixel
// blocks of the same channel block.  The weight rows of a K step, the same
// for all 4 waves, are loaded once per workgroup (one 16-byte load per
// thread) into a double-buffered LDS tile; each wave loads its own pixels'
// channels from memory.  The next K step's loads are issued before the
// current step's MFMAs; one workgroup barrier per step.  Every wave stays in
// the loop (waves past the last pixel load zeros) so every barrier is
// reached.  Cout % (16 TN) == 0 (the host checks).
constexpr int CV_LD = 40;  // LDS row of a weight tile (bf16): 80-byte rows, 16-byte aligned reads
// The grid is one-dimensional, a multiple of 8 workgroups, dealt to the 8
// XCDs in turn (workgroup b on XCD b % 8); workgroup b takes work item
// (b % 8) q + b / 8 (q = grid / 8), item = pixel block x channel blocks +
// channel block.  Each XCD then holds a contiguous run of pixel blocks, so
// the rows above and below a block, which its stencil reads, are the
// neighbouring blocks' own rows in the same L2, and the channel blocks of a
// pixel block, which read the same pixels, run side by side.  Items past
// nwork return before any barrier.
template <int TM, int TN, bool RELU>
__global__ __launch_bounds__(256) void k_conv3x3(Src s, const bf16* __restrict__ w,
                                                  const float* __restrict__ bias,
                                                  bf16* __restrict__ out, int Cout, int nwork) {
  __shared__ __attribute__((aligned(16))) bf16 sw[2][16 * TN][CV_LD];
  const unsigned item = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (item >= (unsigned)nwork) return;
  const int ncb = Cout / (16 * TN);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kb = lane >> 4;
  const int P = s.B * s.H * s.W;
  const int Cin = s.C0 + s.C1;
  const int pb = ((int)(item / (unsigned)ncb) * 4 + wave) * (16 * TM);
  const int cb = (int)(item % (unsigned)ncb) * (16 * TN);
  int pbb[TM], py[TM], px[TM];
  bool pv[TM];
#pragma unroll
  for (int g = 0; g < TM; g++) {
    const int p = pb + g * 16 + r;
    pv[g] = p < P;
    const int pp = pv[g] ? p : 0;
    px[g] = pp % s.W;
    const int q = pp / s.W;
    py[g] = q % s.H;
    pbb[g] = q / s.H;
  }
  // cooperative weight loader: thread t < 64 TN loads row t / 4 (output
  // channel cb + t / 4), channels 8 (t % 4) .. + 7 of the K step
  const int wl_row = threadIdx.x >> 2, wl_col = (threadIdx.x & 3) * 8;
  const bool wl = threadIdx.x < 64 * TN;
  const bf16* wsrc = w + (size_t)(cb + (wl ? wl_row : 0)) * 9 * Cin + wl_col;
  f32x4 acc[TM][TN];
#pragma unroll
  for (int g = 0; g < TM; g++)
#pragma unroll
    for (int h = 0; h < TN; h++) acc[g][h] = f32x4{0.f, 0.f, 0.f, 0.f};
  // K position of the next load: tap lt, channels lci .. lci + 31
  int lt = 0, lci = 0;
  bf16x8 wreg = zero8();
  // per tap: whether each pixel's neighbour is inside, and the offsets of its
  // channel kb * 8 in the two halves of the virtual input (set up once per
  // tap, at its first K step; 32-bit, as load8)
  bool tok[TM];
  unsigned toff0[TM], toff1[TM];
  const int Hc = (s.H + 1) >> 1, Wc = (s.W + 1) >> 1;
  auto tap_setup = [&]() {
    const int dy = (lt / 3 - 1) * s.dil, dx = (lt % 3 - 1) * s.dil;
#pragma unroll
    for (int g = 0; g < TM; g++) {
      const int y = py[g] + dy, x = px[g] + dx;
      tok[g] = pv[g] && y >= 0 && y < s.H && x >= 0 && x < s.W;
      const int yy = tok[g] ? y : 0, xx = tok[g] ? x : 0;
      toff0[g] = ((unsigned)(pbb[g] * s.H + yy) * (unsigned)s.W + (unsigned)xx) * (unsigned)s.C0 +
                 (unsigned)(kb * 8);
      toff1[g] = ((unsigned)(pbb[g] * Hc + (yy >> 1)) * (unsigned)Wc + (unsigned)(xx >> 1)) *
                     (unsigned)s.C1 +
                 (unsigned)(kb * 8);
    }
  };
  auto load = [&](bf16x8 (&A)[TM]) {
    if (lci == 0) tap_setup();
#pragma unroll
    for (int g = 0; g < TM; g++) {
      if (!tok[g])
        A[g] = zero8();
      else if (lci < s.C0)
        A[g] = *reinterpret_cast<const bf16x8*>(s.in0 + (toff0[g] + (unsigned)lci));
      else
        A[g] = *reinterpret_cast<const bf16x8*>(s.in1 + (toff1[g] + (unsigned)(lci - s.C0)));
    }
    if (wl) wreg = *reinterpret_cast<const bf16x8*>(wsrc + (size_t)lt * Cin + lci);
    lci += 32;
    if (lci == Cin) {
      lci = 0;
      lt++;
    }
  };
  auto stage = [&](int buf) {
    if (wl) *reinterpret_cast<bf16x8*>(&sw[buf][wl_row][wl_col]) = wreg;
  };
  auto mma = [&](const bf16x8 (&A)[TM], int buf) {
    bf16x8 Bm[TN];
#pragma unroll
    for (int h = 0; h < TN; h++)
      Bm[h] = *reinterpret_cast<const bf16x8*>(&sw[buf][h * 16 + r][kb * 8]);
#pragma unroll
    for (int g = 0; g < TM; g++)
#pragma unroll
      for (int h = 0; h < TN; h++)
        acc[g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[g], Bm[h], acc[g][h], 0, 0, 0);
  };
  const int nk = 9 * (Cin / 32);
  bf16x8 a0[TM], a1[TM];
  load(a0);
  stage(0);
  __syncthreads();
  for (int kk = 0;;) {  // two steps per trip: pixel sets and weight buffers alternate
    if (kk + 1 < nk) load(a1);
    mma(a0, 0);
    if (kk + 1 < nk) stage(1);
    __syncthreads();
    if (++kk == nk) break;
    if (kk + 1 < nk) load(a0);
    mma(a1, 1);
    if (kk + 1 < nk) stage(0);
    __syncthreads();
    if (++kk == nk) break;
  }
  if (pb >= P) return;  // after the last barrier
#pragma unroll
  for (int h = 0; h < TN; h++) {
    const int co = cb + h * 16 + r;
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int g = 0; g < TM; g++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int p = pb + g * 16 + kb * 4 + i;
        if (p >= P) continue;
        float v = acc[g][h][i] + bv;
        if (RELU) v = v > 0.0f ? v : 0.0f;
        out[(size_t)p * Cout + co] = (__bf16)v;
      }
  }
}

// Orders a wave's LDS accesses: what every lane of the wave stored before it
// is visible to every lane's loads after it (a wave's LDS instructions run in
// order; the fences stop the compiler moving memory accesses across).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ weight gradient
// part[chunk][co][t][ci] = sum over the chunk's pixels p of
// dout[p][co] * in(nbr(p, t))[ci].  K runs over groups of 64 consecutive
// pixels of the flattened (b, y, x) order; the 4 waves of a workgroup take
// every 4th group of the chunk.  A wave loads its group's 64 x 32 blocks of
// dout and of the (shifted) input with 16-byte loads, four pixels per lane,
// and stores them transposed ([channel][pixel]) in its own LDS slice with
// 8-byte writes (the four pixels of one channel), so every MFMA operand is
// one 16-byte LDS read; the next group is loaded into registers while the
// current one is multiplied.  The slices are wave-private, so the loop syncs
// waves, not the workgroup; a wave past the chunk's end loads zeros.  The 4
// waves' results are summed in a fixed order through LDS (reusing the
// slices) after a workgroup barrier.
// Work item of a weight-gradient workgroup: (chunk, co block, tap, ci block)
// of T x T tiles.  The grid is one-dimensional, a multiple of 8 workgroups:
// workgroups are dealt to the 8 XCDs in turn, so workgroup b runs on XCD
// b % 8, and it takes work item (b % 8) q + b / 8 (q = grid / 8).  Each XCD
// then holds a contiguous range of items, and the items of one chunk (every
// tap and channel block, all reading the same pixel groups) run side by
// side on one XCD, where its L2 serves the pixels to all of them.  Items past
// nwork return before any barrier.
struct WgItem {
  int chunk, co, t, ci;
};
__device__ inline bool wgrad_item(int T, int Cout, int Cin, int nwork, WgItem& w) {
  const unsigned q = gridDim.x >> 3;
  const unsigned id = (blockIdx.x & 7u) * q + (blockIdx.x >> 3);
  if (id >= (unsigned)nwork) return false;
  const int nci = Cin / T, per = (Cout / T) * 9 * nci;
  w.chunk = (int)id / per;
  int rem = (int)id % per;
  w.ci = (rem % nci) * T;
  rem /= nci;
  w.t = rem % 9;
  w.co = (rem / 9) * T;
  return true;
}

constexpr int WG_PX = 64;          // pixels per K step of a wave
constexpr int WG_LD = WG_PX + 8;   // LDS row (bf16): 144-byte rows keep 16-byte reads aligned
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_wgrad(Src s, const bf16* __restrict__ dout, int Cout,
                                               float* __restrict__ part, int grp_per_chunk,
                                               int nwork) {
  // [0]: dout^T [co][px], [1]: input^T [ci][px], per wave (36 864 B); the
  // final reduction reads it as float [4][32][33] (16 896 B)
  __shared__ __attribute__((aligned(16))) bf16 lds[2][4][32][WG_LD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kb = lane >> 4;
  const int Cin = s.C0 + s.C1;
  WgItem wi;
  if (!wgrad_item(32, Cout, Cin, nwork, wi)) return;
  const int cob = wi.co, t = wi.t, cib = wi.ci;
  const int dy = (t / 3 - 1) * s.dil, dx = (t % 3 - 1) * s.dil;
  const int P = s.B * s.H * s.W;
  const int ngrp = (P + WG_PX - 1) / WG_PX;
  const int g0 = wi.chunk * grp_per_chunk;
  const int g1 = min(ngrp, g0 + grp_per_chunk);
  const int nit = (g1 - g0 + 3) / 4;  // the same for every wave
  // loader lanes: pixels 4 lq .. 4 lq + 3 of the group x 8-channel group lc
  const int lq = lane >> 2, lc = (lane & 3) * 8;
  bf16x8 va[4], vb[4];
  PixelCursor cur;  // the lane's first pixel of the next group it fetches
  cur.at((g0 + wave) * WG_PX + lq * 4, s.W, s.H);
  auto fetch = [&](int it) {  // called for it = 0, 1, 2, ... in turn
    const int gi = g0 + it * 4 + wave;
    const bool live = it < nit && gi < g1;
    const int p0 = gi * WG_PX + lq * 4;
    int x = cur.x, y = cur.y, b = cur.b;
    cur.advance(4 * WG_PX, s.W, s.H);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      va[j] = zero8();
      vb[j] = zero8();
      if (live && p0 + j < P) {
        va[j] = *reinterpret_cast<const bf16x8*>(dout + ((unsigned)(p0 + j) * (unsigned)Cout + cob + lc));
        vb[j] = load8(s, b, y + dy, x + dx, cib + lc);
      }
      if (++x == s.W) {  // next pixel in (b, y, x) order
        x = 0;
        if (++y == s.H) {
          y = 0;
          b++;
        }
      }
    }
  };
  f32x4 acc[2][2];
#pragma unroll
  for (int g = 0; g < 2; g++)
#pragma unroll
    for (int h = 0; h < 2; h++) acc[g][h] = f32x4{0.f, 0.f, 0.f, 0.f};
  // a half-wave's store writes rows lc + e (lc = 0, 8, 16, 24) at 8
  // consecutive 8-byte slots; rows 16 apart share banks, so rows 16-31 take
  // their slots XOR 8 (the other half of the row): conflict-free.  Operand
  // reads cover 16 aligned rows (one swizzle per read).
  auto swz = [](int row) { return ((row >> 4) & 1) * 32; };  // in pixels
  fetch(0);
  for (int it = 0; it < nit; it++) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int row = lc + e;
      *reinterpret_cast<bf16x4*>(&lds[0][wave][row][(lq * 4) ^ swz(row)]) =
          bf16x4{va[0][e], va[1][e], va[2][e], va[3][e]};
      *reinterpret_cast<bf16x4*>(&lds[1][wave][row][(lq * 4) ^ swz(row)]) =
          bf16x4{vb[0][e], vb[1][e], vb[2][e], vb[3][e]};
    }
    wave_sync();
    fetch(it + 1);  // the next group's loads are in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < WG_PX / 32; ks++) {
      bf16x8 a[2], bm[2];
#pragma unroll
      for (int g = 0; g < 2; g++) {
        const int row = g * 16 + r;
        a[g] = *reinterpret_cast<const bf16x8*>(&lds[0][wave][row][(ks * 32 + kb * 8) ^ swz(row)]);
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = h * 16 + r;
        bm[h] = *reinterpret_cast<const bf16x8*>(&lds[1][wave][row][(ks * 32 + kb * 8) ^ swz(row)]);
      }
#pragma unroll
      for (int g = 0; g < 2; g++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          acc[g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g], bm[h], acc[g][h], 0, 0, 0);
    }
    wave_sync();
  }
  __syncthreads();  // every wave is done with its slice
  float(*red)[32][33] = reinterpret_cast<float(*)[32][33]>(&lds[0][0][0][0]);
#pragma unroll
  for (int g = 0; g < 2; g++)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int i = 0; i < 4; i++) red[wave][g * 16 + kb * 4 + i][h * 16 + r] = acc[g][h][i];
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int row = e / 32, col = e % 32;
    const int co = cob + row, ci = cib + col;
    const float v = ((red[0][row][col] + red[1][row][col]) + red[2][row][col]) + red[3][row][col];
    part[(((size_t)wi.chunk * Cout + co) * 9 + t) * Cin + ci] = v;
  }
}

// The weight gradient for Cout and Cin multiples of 64: a workgroup takes a
// 64 x 64 (co, ci) tile of one tap and one group of 64 consecutive pixels
// per iteration, loaded by all 256 threads (threads 0-127 dout, 128-255 the
// shifted input; 4 pixels x 8 channels each) and stored transposed in a
// double-buffered LDS tile that the 4 waves share: wave w multiplies the co
// half (w & 1) by the ci half (w >> 1), so each wave owns a 32 x 32 block of
// the result.  The next group's loads are issued before the current group's
// MFMAs; one workgroup barrier per group, reached by every thread (the
// iteration count is the workgroup's).
__global__ __launch_bounds__(256) void k_wgrad64(Src s, const bf16* __restrict__ dout, int Cout,
                                                 float* __restrict__ part, int grp_per_chunk,
                                                 int nwork) {
  // [buffer][dout^T | input^T][channel][pixel] (36 864 B)
  __shared__ __attribute__((aligned(16))) bf16 lds[2][2][64][WG_LD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kb = lane >> 4;
  const int Cin = s.C0 + s.C1;
  WgItem wi;
  if (!wgrad_item(64, Cout, Cin, nwork, wi)) return;
  const int cob = wi.co, t = wi.t, cib = wi.ci;
  const int dy = (t / 3 - 1) * s.dil, dx = (t % 3 - 1) * s.dil;
  const int P = s.B * s.H * s.W;
  const int ngrp = (P + WG_PX - 1) / WG_PX;
  const int g0 = wi.chunk * grp_per_chunk;
  const int nit = min(ngrp, g0 + grp_per_chunk) - g0;  // >= 1 for every launched workgroup
  // loader: operand `which`, pixels 4 lq .. 4 lq + 3 of the group, channels lc .. lc + 7
  const int which = threadIdx.x >> 7, li = threadIdx.x & 127;
  const int lq = li >> 3, lc = (li & 7) * 8;
  bf16x8 v[4];
  PixelCursor cur;  // the lane's first pixel of the next group it fetches
  cur.at(g0 * WG_PX + lq * 4, s.W, s.H);
  auto fetch = [&](int it) {  // called for it = 0, 1, 2, ... in turn
    const bool live = it < nit;
    const int p0 = (g0 + it) * WG_PX + lq * 4;
    int x = cur.x, y = cur.y, b = cur.b;
    cur.advance(WG_PX, s.W, s.H);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      v[j] = zero8();
      if (live && p0 + j < P) {
        if (which == 0)
          v[j] = *reinterpret_cast<const bf16x8*>(dout + ((unsigned)(p0 + j) * (unsigned)Cout + cob + lc));
        else
          v[j] = load8(s, b, y + dy, x + dx, cib + lc);
      }
      if (++x == s.W) {  // next pixel in (b, y, x) order
        x = 0;
        if (++y == s.H) {
          y = 0;
          b++;
        }
      }
    }
  };
  // LDS rows of 64 pixels = 16 slots of 4 pixels (8 bytes).  A half-wave's
  // store writes rows lc + e (lc = 0, 8, .., 56) at 4 consecutive slots; the
  // 144-byte rows put rows 16 apart on the same banks, so a row takes its
  // slots XOR 4 ((row >> 4) & 3): the eight rows then cover all 64 banks.
  // Operand reads cover 16 aligned rows, one swizzle for the whole read,
  // which keeps each 16-byte pair of slots together.
  auto swz = [](int row) { return ((row >> 4) & 3) * 16; };  // in pixels
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int row = lc + e;
      *reinterpret_cast<bf16x4*>(&lds[buf][which][row][(lq * 4) ^ swz(row)]) =
          bf16x4{v[0][e], v[1][e], v[2][e], v[3][e]};
    }
  };
  const int arow = (wave & 1) * 32, brow = (wave >> 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int g = 0; g < 2; g++)
#pragma unroll
    for (int h = 0; h < 2; h++) acc[g][h] = f32x4{0.f, 0.f, 0.f, 0.f};
  fetch(0);
  store(0);
  __syncthreads();
  for (int it = 0; it < nit; it++) {
    const int buf = it & 1;
    fetch(it + 1);
#pragma unroll
    for (int ks = 0; ks < WG_PX / 32; ks++) {
      bf16x8 a[2], bm[2];
#pragma unroll
      for (int g = 0; g < 2; g++) {
        const int row = arow + g * 16 + r;
        a[g] = *reinterpret_cast<const bf16x8*>(&lds[buf][0][row][(ks * 32 + kb * 8) ^ swz(row)]);
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = brow + h * 16 + r;
        bm[h] = *reinterpret_cast<const bf16x8*>(&lds[buf][1][row][(ks * 32 + kb * 8) ^ swz(row)]);
      }
#pragma unroll
      for (int g = 0; g < 2; g++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          acc[g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g], bm[h], acc[g][h], 0, 0, 0);
    }
    if (it + 1 < nit) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int g = 0; g < 2; g++)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int co = cob + arow + g * 16 + kb * 4 + i, ci = cib + brow + h * 16 + r;
        part[(((size_t)wi.chunk * Cout + co) * 9 + t) * Cin + ci] = acc[g][h][i];
      }
}

// part[chunk][c] = sum over the chunk's pixels of g[p][c] (bias gradient).
// Threads are (lane, group) pairs: a group is 8 consecutive channels read as
// one 16-byte load, lane = t / (C/8) over 256 / (C/8) pixel lanes, so a wave
// reads whole pixel rows with wide loads; lanes are summed in a fixed order
// in LDS.  C % 8 == 0 and C <= 2048 (the host checks).
__global__ __launch_bounds__(256) void k_colsum_part(const bf16* __restrict__ g, int P, int C,
                                                     int px_per_chunk, float* __restrict__ part) {
  __shared__ float red[256][9];  // 9: odd row stride, conflict-free column reads
  const int p0 = blockIdx.x * px_per_chunk, p1 = min(P, p0 + px_per_chunk);
  const int G = C / 8, lanes = 256 / G;
  const int t = threadIdx.x, grp = t % G, lane = t / G;
  float s[8];
  for (int e = 0; e < 8; e++) s[e] = 0.0f;
  if (lane < lanes)
    for (int p = p0 + lane; p < p1; p += lanes) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(g + (size_t)p * C + grp * 8);
      for (int e = 0; e < 8; e++) s[e] += (float)v[e];
    }
  for (int e = 0; e < 8; e++) red[t][e] = s[e];
  __syncthreads();
  for (int c = t; c < C; c += 256) {
    float r = 0.0f;
    for (int l = 0; l < lanes; l++) r += red[l * G + c / 8][c % 8];
    part[(size_t)blockIdx.x * C + c] = r;
  }
}

// out[i] = sum_c part[c][i] in a fixed order (deterministic): a workgroup
// takes 32 consecutive outputs; its 8 chunk lanes each sum every 8th chunk
// (c = lane, lane + 8, ...) with two running sums, and the 8 lane sums are
// added in lane order through LDS.
__global__ __launch_bounds__(256) void k_reduce_chunks(const float* __restrict__ part, int nchunk,
                                                       long n, float* __restrict__ out) {
  __shared__ float red[8][33];
  const int o = threadIdx.x & 31, lane = threadIdx.x >> 5;
  const long i = (long)blockIdx.x * 32 + o;
  float a0 = 0.0f, a1 = 0.0f;
  if (i < n) {
    int c = lane;
    for (; c + 8 < nchunk; c += 16) {
      a0 += part[(size_t)c * n + i];
      a1 += part[(size_t)(c + 8) * n + i];
    }
    if (c < nchunk) a0 += part[(size_t)c * n + i];
  }
  red[lane][o] = a0 + a1;
  __syncthreads();
  if (lane == 0 && i < n) {
    float r = 0.0f;
    for (int l = 0; l < 8; l++) r += red[l][o];
    out[i] = r;
  }
}

// ------------------------------------------------------------ elementwise
// w'[ci][t][co] = w[co][8 - t][ci]: the data-gradient conv's weights.
__global__ void k_flip_w(const bf16* __restrict__ w, bf16* __restrict__ wt, int Cout, int Cin) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = (long)Cout * 9 * Cin;
  if (i >= n) return;
  const int ci = (int)(i % Cin);
  const int t = (int)((i / Cin) % 9);
  const int co = (int)(i / (9L * Cin));
  wt[((size_t)ci * 9 + (8 - t)) * Cout + co] = w[i];
}

__global__ void k_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = (__bf16)a[i];
}

// g *= (y > 0) (ReLU backward; y is the stored post-activation output).
// g[i] = 0 where y[i] <= 0; eight elements per thread (n % 8 == 0, 16-byte
// aligned buffers: the host checks)
__global__ void k_relu_bwd(bf16* __restrict__ g, const bf16* __restrict__ y, long n8) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n8) return;
  const bf16x8 yv = reinterpret_cast<const bf16x8*>(y)[i];
  bf16x8 gv = reinterpret_cast<const bf16x8*>(g)[i];
  for (int e = 0; e < 8; e++)
    if (!((float)yv[e] > 0.0f)) gv[e] = (__bf16)0.0f;
  reinterpret_cast<bf16x8*>(g)[i] = gv;
}

// Eight consecutive channels as floats, and back.
__device__ inline void unpack8(const bf16x8 v, float* f) {
#pragma unroll
  for (int e = 0; e < 8; e++) f[e] = (float)v[e];
}
__device__ inline bf16x8 pack8(const float* f) {
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; e++) v[e] = (__bf16)f[e];
  return v;
}
__device__ inline bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ inline void st8(bf16* p, const bf16x8 v) { *reinterpret_cast<bf16x8*>(p) = v; }

// The elementwise kernels below take eight consecutive channels per thread
// (16-byte accesses; every channel count is a multiple of 8 and the element
// counts fit 32-bit indices: the host checks).

// 2x2 average pooling, ceil size: out (ceil(H/2), ceil(W/2)).
__global__ void k_pool(const bf16* __restrict__ in, bf16* __restrict__ out, int B, int H, int W,
                       int C) {
  const int Ho = (H + 1) / 2, Wo = (W + 1) / 2, C8 = C / 8;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B * Ho * Wo * C8) return;
  const int cg = i % C8;
  int q = i / C8;
  const int xo = q % Wo;
  q /= Wo;
  const int yo = q % Ho;
  const int b = q / Ho;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
  int cnt = 0;
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      const int y = 2 * yo + dy, x = 2 * xo + dx;
      if (y < H && x < W) {
        unpack8(ld8(in + ((size_t)(b * H + y) * W + x) * C + cg * 8), f);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += f[e];
        cnt++;
      }
    }
#pragma unroll
  for (int e = 0; e < 8; e++) s[e] /= (float)cnt;
  st8(out + (size_t)i * 8, pack8(s));
}

// Gradient of a skip tensor s (full size, C channels): the skip half of the
// decoder's concat gradient plus the pooling's backward of the gradient of
// pool(s).  Then ReLU backward against s.
__global__ void k_skip_grad(const bf16* __restrict__ dcat, int Ccat, const bf16* __restrict__ dpool,
                            const bf16* __restrict__ s, bf16* __restrict__ ds, int B, int H, int W,
                            int C) {
  const int C8 = C / 8;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B * H * W * C8) return;
  const int cg = i % C8;
  int q = i / C8;
  const int x = q % W;
  q /= W;
  const int y = q % H;
  const int b = q / H;
  const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
  const int yo = y >> 1, xo = x >> 1;
  const float inv = 1.0f / (float)((2 * yo + 1 < H ? 2 : 1) * (2 * xo + 1 < W ? 2 : 1));
  float v[8], d[8], sv[8];
  unpack8(ld8(dcat + ((size_t)(b * H + y) * W + x) * Ccat + cg * 8), v);
  unpack8(ld8(dpool + ((size_t)(b * Ho + yo) * Wo + xo) * C + cg * 8), d);
  unpack8(ld8(s + (size_t)i * 8), sv);
#pragma unroll
  for (int e = 0; e < 8; e++) {
    v[e] += d[e] * inv;
    if (!(sv[e] > 0.0f)) v[e] = 0.0f;
  }
  st8(ds + (size_t)i * 8, pack8(v));
}

// Gradient of the coarse (upsampled) input of a decoder conv: sum of the
// concat gradient's coarse channels over the up-to-4 fine children; then
// ReLU backward against the coarse tensor y.
__global__ void k_up_grad(const bf16* __restrict__ dcat, int Ccat, int C0,
                          const bf16* __restrict__ y, bf16* __restrict__ dy_, int B, int H, int W,
                          int C1) {
  const int Hc = (H + 1) / 2, Wc = (W + 1) / 2, C8 = C1 / 8;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B * Hc * Wc * C8) return;
  const int cg = i % C8;
  int q = i / C8;
  const int xc = q % Wc;
  q /= Wc;
  const int yc = q % Hc;
  const int b = q / Hc;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      const int yy = 2 * yc + dy, xx = 2 * xc + dx;
      if (yy < H && xx < W) {
        unpack8(ld8(dcat + ((size_t)(b * H + yy) * W + xx) * Ccat + C0 + cg * 8), f);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += f[e];
      }
    }
  unpack8(ld8(y + (size_t)i * 8), f);
#pragma unroll
  for (int e = 0; e < 8; e++)
    if (!(f[e] > 0.0f)) s[e] = 0.0f;
  st8(dy_ + (size_t)i * 8, pack8(s));
}

// ------------------------------------------------------------ head, loss
// out[p][k] = b[k] + sum_ci x[p][ci] w[k][ci]   (1x1 conv, FP32 output; one
// thread per pixel, K <= 3)
__global__ void k_head_fwd(const bf16* __restrict__ x, const float* __restrict__ w,
                           const float* __restrict__ b, float* __restrict__ out, int P, int C,
                           int K) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  float s[3] = {0.f, 0.f, 0.f}, f[8];
  for (int c0 = 0; c0 < C; c0 += 8) {
    unpack8(ld8(x + (size_t)p * C + c0), f);
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int e = 0; e < 8; e++) s[k] += f[e] * w[k * C + c0 + e];
  }
  for (int k = 0; k < K; k++) out[(size_t)p * K + k] = b[k] + s[k];
}

// Masked mean-square error: per-workgroup partial sums of mask*(y-t)^2 and
// dy = 2 mask (y - t) * inv_norm.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             const float* __restrict__ mask, float* __restrict__ dy,
                                             float* __restrict__ part, int P, int K, float inv_norm) {
  __shared__ float red[256];
  const int p = blockIdx.x * 256 + threadIdx.x;
  float s = 0.0f;
  if (p < P) {
    const float m = mask[p];
    for (int k = 0; k < K; k++) {
      const float d = m * (y[(size_t)p * K + k] - t[(size_t)p * K + k]);
      s += d * d;
      dy[(size_t)p * K + k] = 2.0f * d * inv_norm;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Divergence penalty of the predicted velocity (channels 0, 1 of y, scaled
// to physical units by su, sv): d[p] = su (y0[x+1] - y0[x-1]) / 2 +
// sv (y1[y+1] - y1[y-1]) / 2 where p and its four neighbours are fluid, else
// 0; part[block] = sum of d^2 over the block's pixels.
__global__ __launch_bounds__(256) void k_div_fwd(const float* __restrict__ y,
                                                 const float* __restrict__ mask,
                                                 float* __restrict__ d, float* __restrict__ part,
                                                 int B, int H, int W, float su, float sv) {
  __shared__ float red[256];
  const int p = blockIdx.x * 256 + threadIdx.x;
  float v = 0.0f;
  if (p < B * H * W) {
    const int x = p % W, yy = (p / W) % H;
    if (x > 0 && x < W - 1 && yy > 0 && yy < H - 1 && mask[p] > 0.0f && mask[p - 1] > 0.0f &&
        mask[p + 1] > 0.0f && mask[p - W] > 0.0f && mask[p + W] > 0.0f)
      v = 0.5f * su * (y[(size_t)(p + 1) * 3] - y[(size_t)(p - 1) * 3]) +
          0.5f * sv * (y[(size_t)(p + W) * 3 + 1] - y[(size_t)(p - W) * 3 + 1]);
    d[p] = v;
  }
  red[threadIdx.x] = v * v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dy += the gradient of g/2 * sum d^2 (the adjoint of k_div_fwd's stencil):
// dy[q][0] += g su/2 (d[q-1] - d[q+1]), dy[q][1] += g sv/2 (d[q-W] - d[q+W]),
// neighbours outside the image contributing 0.
__global__ void k_div_bwd(const float* __restrict__ d, float* __restrict__ dy, int B, int H, int W,
                          float gsu, float gsv) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= B * H * W) return;
  const int x = p % W, yy = (p / W) % H;
  const float dl = x > 0 ? d[p - 1] : 0.0f, dr = x < W - 1 ? d[p + 1] : 0.0f;
  const float du = yy > 0 ? d[p - W] : 0.0f, dd = yy < H - 1 ? d[p + W] : 0.0f;
  dy[(size_t)p * 3] += gsu * (dl - dr);
  dy[(size_t)p * 3 + 1] += gsv * (du - dd);
}

// dx[p][c] = relu'(x) * sum_k dy[p][k] w[k][c]
__global__ void k_head_bwd(const float* __restrict__ dy, const float* __restrict__ w,
                           const bf16* __restrict__ x, bf16* __restrict__ dx, int P, int C, int K) {
  const int C8 = C / 8;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P * C8) return;
  const int cg = i % C8, p = i / C8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
  for (int k = 0; k < K; k++) {
    const float d = dy[(size_t)p * K + k];
#pragma unroll
    for (int e = 0; e < 8; e++) s[e] += d * w[k * C + cg * 8 + e];
  }
  unpack8(ld8(x + (size_t)i * 8), f);
#pragma unroll
  for (int e = 0; e < 8; e++)
    if (!(f[e] > 0.0f)) s[e] = 0.0f;
  st8(dx + (size_t)i * 8, pack8(s));
}

// part[chunk][k*C + c] (weights) and part[chunk][K*C + k] (bias) of the head
// (K <= 3 outputs, C <= 2048 inputs).  Threads are (lane, group) pairs as in
// k_colsum_part; group-0 threads also sum the bias terms.
__global__ __launch_bounds__(256) void k_head_wgrad_part(const float* __restrict__ dy,
                                                         const bf16* __restrict__ x,
                                                         float* __restrict__ part, int P, int C,
                                                         int K, int px_per_chunk) {
  __shared__ float red[256][28];  // 24 weight + 3 bias sums (+1: odd row stride)
  const int p0 = blockIdx.x * px_per_chunk, p1 = min(P, p0 + px_per_chunk);
  const int G = C / 8, lanes = 256 / G;
  const int t = threadIdx.x, grp = t % G, lane = t / G;
  float a[3][8], bsum[3] = {0.f, 0.f, 0.f}, f[8];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int e = 0; e < 8; e++) a[k][e] = 0.0f;
  if (lane < lanes)
    for (int p = p0 + lane; p < p1; p += lanes) {
      unpack8(ld8(x + (size_t)p * C + grp * 8), f);
      for (int k = 0; k < K; k++) {
        const float d = dy[(size_t)p * K + k];
#pragma unroll
        for (int e = 0; e < 8; e++) a[k][e] += d * f[e];
        bsum[k] += d;
      }
    }
#pragma unroll
  for (int k = 0; k < 3; k++) {
#pragma unroll
    for (int e = 0; e < 8; e++) red[t][k * 8 + e] = a[k][e];
    red[t][24 + k] = bsum[k];
  }
  __syncthreads();
  const int nw = K * C + K;
  for (int o = t; o < K * C; o += 256) {
    const int k = o / C, c = o % C;
    float r = 0.0f;
    for (int l = 0; l < lanes; l++) r += red[l * G + c / 8][k * 8 + c % 8];
    part[(size_t)blockIdx.x * nw + o] = r;
  }
  if (t < K) {  // bias: the group-0 threads of every lane saw each pixel once
    float r = 0.0f;
    for (int l = 0; l < lanes; l++) r += red[l * G][24 + t];
    part[(size_t)blockIdx.x * nw + K * C + t] = r;
  }
}

// ------------------------------------------------------------ optimiser
// Adam on the FP32 master copy; bf16 copy of the same element refreshed.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16* __restrict__ pb, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float wd) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  const float np = p[i] - upd - lr * wd * p[i];  // decoupled weight decay (AdamW)
  p[i] = np;
  pb[i] = (__bf16)np;
}

// ------------------------------------------------------------ data
// x[b][y][x][0..Cp) bf16 from samples idx[b] of an FP32 [N][H][W][Ci]
// input set (channels >= Ci zero); t[b] from the [N][H][W][3] target set and
// mask[b] from input channel 0.
__global__ void k_gather(const float* __restrict__ inputs, const float* __restrict__ targets,
                         const int* __restrict__ idx, bf16* __restrict__ x, float* __restrict__ t,
                         float* __restrict__ mask, int B, int HW, int Ci, int Cp) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B * HW) return;
  const int b = i / HW, q = i % HW;
  const size_t src = ((size_t)idx[b] * HW + q) * Ci;
  float f[8];
  for (int c0 = 0; c0 < Cp; c0 += 8) {
#pragma unroll
    for (int e = 0; e < 8; e++) f[e] = c0 + e < Ci ? inputs[src + c0 + e] : 0.0f;
    st8(x + (size_t)i * Cp + c0, pack8(f));
  }
  if (t) {
    const size_t ts = ((size_t)idx[b] * HW + q) * 3;
    for (int c = 0; c < 3; c++) t[(size_t)i * 3 + c] = targets[ts + c];
    mask[i] = inputs[src];
  }
}

// ------------------------------------------------------------ launchers
static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }
static inline const bf16* cb(const u16* p) { return reinterpret_cast<const bf16*>(p); }
static inline bf16* mb(u16* p) { return reinterpret_cast<bf16*>(p); }
static inline Src src_of(const ConvSrc& c) {
  return Src{cb(c.in0), c.C0, cb(c.in1), c.C1, c.B, c.H, c.W, c.dil};
}

template <int TM, int TN>
static void launch_conv(const Src& s, const u16* w, const float* bias, u16* out, int Cout,
                        long P, bool relu, hipStream_t st) {
  const long nwork = (P + 64 * TM - 1) / (64 * TM) * (Cout / (16 * TN));
  const unsigned g = (unsigned)((nwork + 7) / 8 * 8);  // a multiple of 8 (k_conv3x3)
  if (relu) k_conv3x3<TM, TN, true><<<g, 256, 0, st>>>(s, cb(w), bias, mb(out), Cout, (int)nwork);
  else k_conv3x3<TM, TN, false><<<g, 256, 0, st>>>(s, cb(w), bias, mb(out), Cout, (int)nwork);
}

// Tile choice: 64 output channels per wave where Cout allows, 64 pixels per
// wave while that still gives >= 1024 workgroups (4 per CU), else 32.
hipError_t conv3x3(const ConvSrc& cs, const u16* w, const float* bias, u16* out, int Cout,
                   bool relu, hipStream_t st) {
  const Src s = src_of(cs);
  const long P = (long)cs.B * cs.H * cs.W;
  if (Cout % 32 || (cs.C0 + cs.C1) % 32 || P + 256 > 0x7fffffffL) return hipErrorInvalidValue;
  const int TN = Cout % 64 == 0 ? 4 : 2;
  const long wg4 = (P + 255) / 256 * (Cout / (16 * TN));
  if (TN == 4) {
    if (wg4 >= 1024) launch_conv<4, 4>(s, w, bias, out, Cout, P, relu, st);
    else launch_conv<2, 4>(s, w, bias, out, Cout, P, relu, st);
  } else {
    if (wg4 >= 1024) launch_conv<4, 2>(s, w, bias, out, Cout, P, relu, st);
    else launch_conv<2, 2>(s, w, bias, out, Cout, P, relu, st);
  }
  return hipGetLastError();
}

int wgrad_tile(int Cout, int Cin) { return Cout % 64 == 0 && Cin % 64 == 0 ? 64 : 32; }

int wgrad_groups(const ConvSrc& cs) {
  return (int)(((long)cs.B * cs.H * cs.W + WG_PX - 1) / WG_PX);
}

int wgrad_chunks(const ConvSrc& cs, int grp_per_chunk) {
  return (wgrad_groups(cs) + grp_per_chunk - 1) / grp_per_chunk;
}

hipError_t wgrad(const ConvSrc& cs, const u16* dout, int Cout, float* part, int grp_per_chunk,
                 hipStream_t st) {
  const int Cin = cs.C0 + cs.C1;
  if (Cout % 32 || Cin % 32 || grp_per_chunk < 1 ||
      (long)cs.B * cs.H * cs.W + WG_PX > 0x7fffffffL)
    return hipErrorInvalidValue;
  const int T = wgrad_tile(Cout, Cin);
  const long nwork = (long)wgrad_chunks(cs, grp_per_chunk) * (Cout / T) * 9 * (Cin / T);
  if (nwork > 0x7ffffff0L) return hipErrorInvalidValue;
  const unsigned g = (unsigned)((nwork + 7) / 8 * 8);  // a multiple of 8 (wgrad_item)
  if (T == 64) k_wgrad64<<<g, 256, 0, st>>>(src_of(cs), cb(dout), Cout, part, grp_per_chunk, (int)nwork);
  else k_wgrad<<<g, 256, 0, st>>>(src_of(cs), cb(dout), Cout, part, grp_per_chunk, (int)nwork);
  return hipGetLastError();
}

hipError_t colsum_part(const u16* g, int P, int C, int px_per_chunk, float* part, hipStream_t st) {
  if (C % 8 || C > 2048 || (reinterpret_cast<uintptr_t>(g) & 15)) return hipErrorInvalidValue;
  k_colsum_part<<<(unsigned)((P + px_per_chunk - 1) / px_per_chunk), 256, 0, st>>>(
      cb(g), P, C, px_per_chunk, part);
  return hipGetLastError();
}

hipError_t reduce_chunks(const float* part, int nchunk, long n, float* out, hipStream_t st) {
  k_reduce_chunks<<<(unsigned)((n + 31) / 32), 256, 0, st>>>(part, nchunk, n, out);
  return hipGetLastError();
}

hipError_t flip_w(const u16* w, u16* wt, int Cout, int Cin, hipStream_t st) {
  k_flip_w<<<nblk((long)Cout * 9 * Cin), 256, 0, st>>>(cb(w), mb(wt), Cout, Cin);
  return hipGetLastError();
}

hipError_t to_bf16(const float* a, u16* b, long n, hipStream_t st) {
  k_to_bf16<<<nblk(n), 256, 0, st>>>(a, mb(b), n);
  return hipGetLastError();
}

hipError_t relu_bwd(u16* g, const u16* y, long n, hipStream_t st) {
  if (n % 8 || ((reinterpret_cast<uintptr_t>(g) | reinterpret_cast<uintptr_t>(y)) & 15))
    return hipErrorInvalidValue;
  k_relu_bwd<<<nblk(n / 8), 256, 0, st>>>(mb(g), cb(y), n / 8);
  return hipGetLastError();
}

// the eight-channel elementwise kernels: C % 8 == 0 and n / 8 fits an int
static inline bool vec_ok(long n, int C) { return C % 8 == 0 && n / 8 < 0x7fffff00L; }

hipError_t pool(const u16* in, u16* out, int B, int H, int W, int C, hipStream_t st) {
  const long n = (long)B * ((H + 1) / 2) * ((W + 1) / 2) * C;
  if (!vec_ok((long)B * H * W * C, C)) return hipErrorInvalidValue;
  k_pool<<<nblk(n / 8), 256, 0, st>>>(cb(in), mb(out), B, H, W, C);
  return hipGetLastError();
}

hipError_t skip_grad(const u16* dcat, int Ccat, const u16* dpool, const u16* s, u16* ds, int B,
                     int H, int W, int C, hipStream_t st) {
  if (!vec_ok((long)B * H * W * Ccat, C) || Ccat % 8) return hipErrorInvalidValue;
  k_skip_grad<<<nblk((long)B * H * W * C / 8), 256, 0, st>>>(cb(dcat), Ccat, cb(dpool), cb(s),
                                                             mb(ds), B, H, W, C);
  return hipGetLastError();
}

hipError_t up_grad(const u16* dcat, int Ccat, int C0, const u16* y, u16* dy, int B, int H, int W,
                   int C1, hipStream_t st) {
  if (!vec_ok((long)B * H * W * Ccat, C1) || Ccat % 8 || C0 % 8) return hipErrorInvalidValue;
  k_up_grad<<<nblk((long)B * ((H + 1) / 2) * ((W + 1) / 2) * C1 / 8), 256, 0, st>>>(
      cb(dcat), Ccat, C0, cb(y), mb(dy), B, H, W, C1);
  return hipGetLastError();
}

hipError_t head_fwd(const u16* x, const float* w, const float* b, float* out, int P, int C, int K,
                    hipStream_t st) {
  if (!vec_ok((long)P * C, C) || K > 3) return hipErrorInvalidValue;
  k_head_fwd<<<nblk(P), 256, 0, st>>>(cb(x), w, b, out, P, C, K);
  return hipGetLastError();
}

hipError_t mse(const float* y, const float* t, const float* mask, float* dy, float* part, int P,
               int K, float inv_norm, hipStream_t st) {
  k_mse<<<nblk(P), 256, 0, st>>>(y, t, mask, dy, part, P, K, inv_norm);
  return hipGetLastError();
}

hipError_t div_penalty(const float* y, const float* mask, float* d, float* part, float* dy, int B,
                       int H, int W, float su, float sv, float g, hipStream_t st) {
  const long P = (long)B * H * W;
  if (P > 0x7fffff00L) return hipErrorInvalidValue;
  k_div_fwd<<<nblk(P), 256, 0, st>>>(y, mask, d, part, B, H, W, su, sv);
  k_div_bwd<<<nblk(P), 256, 0, st>>>(d, dy, B, H, W, 0.5f * g * su, 0.5f * g * sv);
  return hipGetLastError();
}

hipError_t head_bwd(const float* dy, const float* w, const u16* x, u16* dx, int P, int C, int K,
                    hipStream_t st) {
  if (!vec_ok((long)P * C, C)) return hipErrorInvalidValue;
  k_head_bwd<<<nblk((long)P * C / 8), 256, 0, st>>>(dy, w, cb(x), mb(dx), P, C, K);
  return hipGetLastError();
}

hipError_t head_wgrad_part(const float* dy, const u16* x, float* part, int P, int C, int K,
                           int px_per_chunk, hipStream_t st) {
  if (C % 8 || C > 2048 || K > 3) return hipErrorInvalidValue;
  k_head_wgrad_part<<<(unsigned)((P + px_per_chunk - 1) / px_per_chunk), 256, 0, st>>>(
      dy, cb(x), part, P, C, K, px_per_chunk);
  return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m, float* v, u16* pb, long n, float lr, float b1,
                float b2, float eps, float c1, float c2, float wd, hipStream_t st) {
  k_adam<<<nblk(n), 256, 0, st>>>(p, g, m, v, mb(pb), n, lr, b1, b2, eps, c1, c2, wd);
  return hipGetLastError();
}

hipError_t gather(const float* inputs, const float* targets, const int* idx, u16* x, float* t,
                  float* mask, int B, int HW, int Ci, int Cp, hipStream_t st) {
  if (Cp % 8 || Ci > Cp || !vec_ok((long)B * HW * Cp, Cp)) return hipErrorInvalidValue;
  k_gather<<<nblk((long)B * HW), 256, 0, st>>>(inputs, targets, idx, mb(x), t, mask, B, HW, Ci,
                                              Cp);
  return hipGetLastError();
}

}  // namespace lbmnn
