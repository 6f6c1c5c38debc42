// nn.cpp — the warm-start network: a U-Net of 3x3 convolutions that maps a
// cavity's (fluid mask, log-Re, lid speed) on a square grid to its steady
// (u, v, rho) there.  Device-resident BF16 weights and activations with an
// FP32 master copy, hand-written forward/backward (nn_kernels.hip) and Adam.
//
// Inputs carry `cin` channels (<= 32, zero-padded to 32 for the first conv);
// channel 0 is the fluid mask that weighs the loss.
// Architecture for `levels` L and base width c (C_l = c 2^l):
//   encoder l < L-1:  a_l = relu(conv(x_l)), s_l = relu(conv(a_l)), x_{l+1} = avgpool2(s_l)
//   bottleneck:       a_{L-1} = relu(conv_d(x_{L-1})), s_{L-1} = relu(conv_2d(a_{L-1}))
//                     (3x3 stencils dilated by d and 2d: the receptive field spans the grid)
//   decoder l = L-2..0: da_l = relu(conv([s_l, up2(prev)])), dd_l = relu(conv(da_l)), prev = dd_l
//   head:             y = 1x1 conv(dd_0) -> 3 channels (FP32)
// Loss: mean over fluid cells (input channel 0) and the 3 channels of (y - t)^2.
// Parameters, canonical order: per conv (encoder a, b per level, bottleneck
// a, b, decoder a, b from the coarsest level) W[Cout][9][Cin] then b[Cout];
// head W[3][c] then b[3].  The first conv's Cin is padded from 3 to 32.
//
// No counterpart in the reference.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "lbm_host.hpp"
#include "nn_launch.hpp"

using namespace lbmnn;
using lbmhost::check_hip;
using lbmhost::set_error;

namespace {
constexpr int CIN0 = 32;  // padded input channels
constexpr int KOUT = 3;   // u, v, rho
constexpr int RED_CHUNK = 1024;  // pixels per workgroup of the bias / head-gradient partials
struct Conv {
  int C0, C1, Cout, level, dil;
  size_t w, b;  // offsets of W and bias in the parameter vector
};
}  // namespace

#define NC(x)                                    \
  do {                                           \
    int rc_ = check_hip((x), #x);                \
    if (rc_ != LBM_OK) return rc_;               \
  } while (0)

struct lbm_net {
  int H = 0, W = 0, Bmax = 0, base = 32, L = 4, cin = 3;  // cin: input channels
  float wd = 0.0f;  // decoupled weight decay of the Adam step
  float div_w = 0.0f, div_su = 1.0f, div_sv = 1.0f;  // divergence penalty of the loss
  int dil = 1;  // bottleneck dilation: its convs use dil and 2 dil
  hipStream_t st = nullptr;
  std::vector<Conv> conv;
  size_t nparam = 0, head_w = 0, head_b = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 master, grad, Adam
  u16 *pb = nullptr, *wt = nullptr;  // bf16 weights, flipped weights (same offsets)
  int step = 0;
  std::vector<int> Hl, Wl;
  // activations [level], gradients
  u16* x0 = nullptr;
  std::vector<u16*> a, s, xp, da, dd;  // xp[l] = pool(s[l]) at level l+1
  std::vector<u16*> ga, gs, gxp, gda, gdd, dcat;
  float *y = nullptr, *dy = nullptr, *t = nullptr, *mask = nullptr, *lpart = nullptr;
  float *dv = nullptr, *dvpart = nullptr;  // divergence per pixel, its block sums
  float* part = nullptr;
  size_t part_cap = 0;
  int* idx = nullptr;
  float *stage_in = nullptr, *stage_t = nullptr;  // host-batch staging [Bmax][H][W][3]
  // training set
  float *set_in = nullptr, *set_t = nullptr;
  int nset = 0;
  std::vector<double> set_fluid;
  std::vector<void*> allocs;
};

static int nalloc(lbm_net* n, void** ptr, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc failed for the network (" + std::to_string(bytes) + " bytes)");
    *ptr = nullptr;
    return LBM_ENOMEM;
  }
  n->allocs.push_back(*ptr);
  return LBM_OK;
}

template <class T>
static int nalloc_t(lbm_net* n, T** ptr, size_t count) {
  return nalloc(n, (void**)ptr, sizeof(T) * count);
}

static int C_of(const lbm_net* n, int l) { return n->base << l; }
static long P_of(const lbm_net* n, int l, int B) { return (long)B * n->Hl[l] * n->Wl[l]; }

static ConvSrc src1(const lbm_net* n, const u16* x, int C, int l, int B, int dil = 1) {
  ConvSrc s;
  s.in0 = x;
  s.C0 = C;
  s.B = B;
  s.H = n->Hl[l];
  s.W = n->Wl[l];
  s.dil = dil;
  return s;
}

// chunking of a conv's weight gradient: about 2048 workgroups per launch
// (fewer chunks = less partial-sum traffic for the fixed-order reduction)
static int grp_per_chunk(const lbm_net* n, const Conv& c, const ConvSrc& s) {
  const long ngrp = wgrad_groups(s);
  const int T = wgrad_tile(c.Cout, c.C0 + c.C1);
  const long per = (long)(c.Cout / T) * 9 * ((c.C0 + c.C1) / T);
  long chunks = std::max(1L, 2048 / per);
  (void)n;
  return (int)std::max(1L, (ngrp + chunks - 1) / chunks);
}

// uniform (0,1] from a 64-bit LCG: deterministic initialisation on every host
static double urand(uint64_t& st) {
  st = st * 6364136223846793005ULL + 1442695040888963407ULL;
  return ((st >> 11) + 1) * (1.0 / 9007199254740992.0);
}

int lbm_net_create(int H, int W, int batch, int base, int levels, int in_channels, int dilation,
                   unsigned seed, lbm_net** out) {
  if (!out || H < 2 || W < 2 || batch < 1 || base < 32 || base % 32 || levels < 2 || levels > 6 ||
      ((long)base << (levels - 1)) > 2048 ||
      in_channels < 1 || in_channels > CIN0 || dilation < 1 || dilation > 64)
    return LBM_EINVAL;
  // the kernels index activations with 32-bit offsets: the largest tensor
  // (batch x H x W x max(base, padded input channels)) must stay below 2^31
  if ((long long)batch * H * W * (base > CIN0 ? base : CIN0) >= (1LL << 31)) {
    set_error("network tensors must hold fewer than 2^31 elements (batch x H x W x channels)");
    return LBM_EINVAL;
  }
  *out = nullptr;
  lbm_net* n = new lbm_net();
  n->H = H;
  n->W = W;
  n->Bmax = batch;
  n->base = base;
  n->L = levels;
  n->cin = in_channels;
  n->dil = dilation;
  const int L = levels;
  n->Hl.resize(L);
  n->Wl.resize(L);
  n->Hl[0] = H;
  n->Wl[0] = W;
  for (int l = 1; l < L; l++) {
    n->Hl[l] = (n->Hl[l - 1] + 1) / 2;
    n->Wl[l] = (n->Wl[l - 1] + 1) / 2;
  }
  // layers and parameter offsets
  size_t off = 0;
  auto add = [&](int C0, int C1, int Cout, int l, int dil) {
    Conv c{C0, C1, Cout, l, dil, off, 0};
    off += (size_t)Cout * 9 * (C0 + C1);
    c.b = off;
    off += Cout;
    n->conv.push_back(c);
  };
  for (int l = 0; l < L; l++) {
    const bool bot = l == L - 1;
    add(l == 0 ? CIN0 : C_of(n, l - 1), 0, C_of(n, l), l, bot ? dilation : 1);
    add(C_of(n, l), 0, C_of(n, l), l, bot ? 2 * dilation : 1);
  }
  for (int l = L - 2; l >= 0; l--) {
    add(C_of(n, l), C_of(n, l + 1), C_of(n, l), l, 1);
    add(C_of(n, l), 0, C_of(n, l), l, 1);
  }
  n->head_w = off;
  off += (size_t)KOUT * base;
  n->head_b = off;
  off += KOUT;
  n->nparam = off;

  int rc = check_hip(hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking), "hipStreamCreate");
  if (!rc) rc = nalloc_t(n, &n->p, off);
  if (!rc) rc = nalloc_t(n, &n->g, off);
  if (!rc) rc = nalloc_t(n, &n->m, off);
  if (!rc) rc = nalloc_t(n, &n->v, off);
  if (!rc) rc = nalloc_t(n, &n->pb, off);
  if (!rc) rc = nalloc_t(n, &n->wt, off);
  const int B = batch;
  n->a.assign(L, nullptr);
  n->s.assign(L, nullptr);
  n->xp.assign(L, nullptr);
  n->da.assign(L, nullptr);
  n->dd.assign(L, nullptr);
  n->ga.assign(L, nullptr);
  n->gs.assign(L, nullptr);
  n->gxp.assign(L, nullptr);
  n->gda.assign(L, nullptr);
  n->gdd.assign(L, nullptr);
  n->dcat.assign(L, nullptr);
  if (!rc) rc = nalloc_t(n, &n->x0, (size_t)P_of(n, 0, B) * CIN0);
  for (int l = 0; l < L && !rc; l++) {
    const size_t e = (size_t)P_of(n, l, B) * C_of(n, l);
    rc = nalloc_t(n, &n->a[l], e);
    if (!rc) rc = nalloc_t(n, &n->s[l], e);
    if (!rc) rc = nalloc_t(n, &n->ga[l], e);
    if (!rc) rc = nalloc_t(n, &n->gs[l], e);
    if (l < L - 1) {
      const size_t ep = (size_t)P_of(n, l + 1, B) * C_of(n, l);
      if (!rc) rc = nalloc_t(n, &n->xp[l], ep);
      if (!rc) rc = nalloc_t(n, &n->gxp[l], ep);
      if (!rc) rc = nalloc_t(n, &n->da[l], e);
      if (!rc) rc = nalloc_t(n, &n->dd[l], e);
      if (!rc) rc = nalloc_t(n, &n->gda[l], e);
      if (!rc) rc = nalloc_t(n, &n->gdd[l], e);
      if (!rc) rc = nalloc_t(n, &n->dcat[l], (size_t)P_of(n, l, B) * (C_of(n, l) + C_of(n, l + 1)));
    }
  }
  const size_t P0 = (size_t)P_of(n, 0, B);
  if (!rc) rc = nalloc_t(n, &n->y, P0 * KOUT);
  if (!rc) rc = nalloc_t(n, &n->dy, P0 * KOUT);
  if (!rc) rc = nalloc_t(n, &n->t, P0 * KOUT);
  if (!rc) rc = nalloc_t(n, &n->mask, P0);
  if (!rc) rc = nalloc_t(n, &n->lpart, (P0 + 255) / 256);
  if (!rc) rc = nalloc_t(n, &n->dv, P0);
  if (!rc) rc = nalloc_t(n, &n->dvpart, (P0 + 255) / 256);
  if (!rc) rc = nalloc_t(n, &n->idx, (size_t)B);
  if (!rc) rc = nalloc_t(n, &n->stage_in, P0 * in_channels);
  if (!rc) rc = nalloc_t(n, &n->stage_t, P0 * 3);
  // partial-sum scratch: the largest conv weight-gradient partials, the
  // bias and head partials
  size_t cap = 0;
  for (const Conv& c : n->conv) {
    ConvSrc s = src1(n, nullptr, c.C0, c.level, B);
    s.C1 = c.C1;
    const int ch = wgrad_chunks(s, grp_per_chunk(n, c, s));
    cap = std::max(cap, (size_t)ch * c.Cout * 9 * (c.C0 + c.C1));
    cap = std::max(cap, (size_t)((P_of(n, c.level, B) + RED_CHUNK - 1) / RED_CHUNK) * c.Cout);
  }
  cap = std::max(cap, (size_t)((P0 + RED_CHUNK - 1) / RED_CHUNK) * (KOUT * base + KOUT));
  n->part_cap = cap;
  if (!rc) rc = nalloc_t(n, &n->part, cap);
  if (rc) {
    lbm_net_destroy(n);
    return rc;
  }
  // He initialisation (fan-in of the real input channels), zero biases
  std::vector<float> hp(off, 0.0f);
  uint64_t st = 0x9e3779b97f4a7c15ULL ^ seed;
  for (size_t ci = 0; ci < n->conv.size(); ci++) {
    const Conv& c = n->conv[ci];
    const int Cin = c.C0 + c.C1;
    const int real = ci == 0 ? n->cin : Cin;
    const double sd = std::sqrt(2.0 / (9.0 * real));
    for (int co = 0; co < c.Cout; co++)
      for (int tp = 0; tp < 9; tp++)
        for (int k = 0; k < Cin; k++) {
          const double u1 = urand(st), u2 = urand(st);
          const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
          hp[c.w + ((size_t)co * 9 + tp) * Cin + k] = k < real ? (float)(sd * z) : 0.0f;
        }
  }
  for (size_t k = 0; k < (size_t)KOUT * base; k++) {
    const double u1 = urand(st), u2 = urand(st);
    hp[n->head_w + k] = (float)(0.1 * std::sqrt(1.0 / base) * std::sqrt(-2.0 * std::log(u1)) *
                                std::cos(6.283185307179586 * u2));
  }
  rc = lbm_net_set_params(n, hp.data());
  if (rc) {
    lbm_net_destroy(n);
    return rc;
  }
  *out = n;
  return LBM_OK;
}

void lbm_net_destroy(lbm_net* n) {
  if (!n) return;
  if (n->st) (void)hipStreamSynchronize(n->st);
  for (void* q : n->allocs) (void)hipFree(q);
  if (n->st) (void)hipStreamDestroy(n->st);
  delete n;
}

long lbm_net_num_params(const lbm_net* n) { return n ? (long)n->nparam : -1; }

static int refresh_weights(lbm_net* n) {
  NC(to_bf16(n->p, n->pb, (long)n->nparam, n->st));
  for (const Conv& c : n->conv) NC(flip_w(n->pb + c.w, n->wt + c.w, c.Cout, c.C0 + c.C1, n->st));
  return LBM_OK;
}

int lbm_net_set_params(lbm_net* n, const float* params) {
  if (!n || !params) return LBM_EINVAL;
  NC(hipMemcpyAsync(n->p, params, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->st));
  NC(hipMemsetAsync(n->m, 0, sizeof(float) * n->nparam, n->st));
  NC(hipMemsetAsync(n->v, 0, sizeof(float) * n->nparam, n->st));
  n->step = 0;
  int rc = refresh_weights(n);
  if (rc) return rc;
  NC(hipStreamSynchronize(n->st));
  return LBM_OK;
}

int lbm_net_get_params(lbm_net* n, float* params) {
  if (!n || !params) return LBM_EINVAL;
  NC(hipMemcpyAsync(params, n->p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->st));
  NC(hipStreamSynchronize(n->st));
  return LBM_OK;
}

// forward of B samples already in x0
static int forward(lbm_net* n, int B) {
  const int L = n->L;
  const u16* x = n->x0;
  int Cx = CIN0;
  size_t ci = 0;
  for (int l = 0; l < L; l++) {
    const Conv& ca = n->conv[ci++];
    const Conv& cb = n->conv[ci++];
    NC(conv3x3(src1(n, x, Cx, l, B, ca.dil), n->pb + ca.w, n->p + ca.b, n->a[l], ca.Cout, true,
               n->st));
    NC(conv3x3(src1(n, n->a[l], C_of(n, l), l, B, cb.dil), n->pb + cb.w, n->p + cb.b, n->s[l],
               cb.Cout, true, n->st));
    if (l < L - 1) {
      NC(pool(n->s[l], n->xp[l], B, n->Hl[l], n->Wl[l], C_of(n, l), n->st));
      x = n->xp[l];
      Cx = C_of(n, l);
    }
  }
  const u16* prev = n->s[L - 1];
  for (int l = L - 2; l >= 0; l--) {
    const Conv& ca = n->conv[ci++];
    const Conv& cb = n->conv[ci++];
    ConvSrc s = src1(n, n->s[l], C_of(n, l), l, B);
    s.in1 = prev;
    s.C1 = C_of(n, l + 1);
    NC(conv3x3(s, n->pb + ca.w, n->p + ca.b, n->da[l], ca.Cout, true, n->st));
    NC(conv3x3(src1(n, n->da[l], C_of(n, l), l, B), n->pb + cb.w, n->p + cb.b, n->dd[l], cb.Cout,
               true, n->st));
    prev = n->dd[l];
  }
  NC(head_fwd(n->dd[0], n->p + n->head_w, n->p + n->head_b, n->y, (int)P_of(n, 0, B), n->base,
              KOUT, n->st));
  return LBM_OK;
}

// weight + bias gradient of conv c (input s, output gradient gout) into n->g
static int conv_grads(lbm_net* n, const Conv& c, const ConvSrc& s, const u16* gout) {
  const int spc = grp_per_chunk(n, c, s);
  const int ch = wgrad_chunks(s, spc);
  const long nw = (long)c.Cout * 9 * (c.C0 + c.C1);
  if ((size_t)ch * nw > n->part_cap) {
    set_error("network scratch too small");
    return LBM_EINVAL;
  }
  NC(wgrad(s, gout, c.Cout, n->part, spc, n->st));
  NC(reduce_chunks(n->part, ch, nw, n->g + c.w, n->st));
  const int P = s.B * s.H * s.W;
  NC(colsum_part(gout, P, c.Cout, RED_CHUNK, n->part, n->st));
  NC(reduce_chunks(n->part, (P + RED_CHUNK - 1) / RED_CHUNK, c.Cout, n->g + c.b, n->st));
  return LBM_OK;
}

// data gradient of conv c: gin = conv(gout, flipped W), Cin channels
static int conv_dgrad(lbm_net* n, const Conv& c, const u16* gout, int l, int B, u16* gin) {
  return check_hip(conv3x3(src1(n, gout, c.Cout, l, B, c.dil), n->wt + c.w, nullptr, gin,
                           c.C0 + c.C1, false, n->st),
                   "conv dgrad");
}

// loss + backward of B samples (x0, t, mask filled); gradients into n->g
static int backward(lbm_net* n, int B, double fluid, double* loss) {
  const int L = n->L;
  const int P0 = (int)P_of(n, 0, B);
  const float inv_norm = (float)(1.0 / (KOUT * std::max(1.0, fluid)));
  NC(mse(n->y, n->t, n->mask, n->dy, n->lpart, P0, KOUT, inv_norm, n->st));
  if (n->div_w > 0.0f)
    NC(div_penalty(n->y, n->mask, n->dv, n->dvpart, n->dy, B, n->H, n->W, n->div_su, n->div_sv,
                   2.0f * n->div_w * inv_norm, n->st));
  // head
  NC(head_bwd(n->dy, n->p + n->head_w, n->dd[0], n->gdd[0], P0, n->base, KOUT, n->st));
  const int hch = (P0 + RED_CHUNK - 1) / RED_CHUNK;
  NC(head_wgrad_part(n->dy, n->dd[0], n->part, P0, n->base, KOUT, RED_CHUNK, n->st));
  NC(reduce_chunks(n->part, hch, (long)KOUT * n->base + KOUT, n->g + n->head_w, n->st));
  // decoder, fine to coarse
  const size_t ndec0 = 2 * (size_t)L;  // index of the first decoder conv (coarsest level)
  for (int l = 0; l <= L - 2; l++) {
    const size_t k = ndec0 + 2 * (size_t)(L - 2 - l);
    const Conv& ca = n->conv[k];
    const Conv& cb = n->conv[k + 1];
    int rc = conv_grads(n, cb, src1(n, n->da[l], C_of(n, l), l, B), n->gdd[l]);
    if (rc) return rc;
    rc = conv_dgrad(n, cb, n->gdd[l], l, B, n->gda[l]);
    if (rc) return rc;
    NC(relu_bwd(n->gda[l], n->da[l], P_of(n, l, B) * C_of(n, l), n->st));
    const u16* prev = (l + 1 <= L - 2) ? n->dd[l + 1] : n->s[L - 1];
    u16* gprev = (l + 1 <= L - 2) ? n->gdd[l + 1] : n->gs[L - 1];
    ConvSrc s = src1(n, n->s[l], C_of(n, l), l, B);
    s.in1 = prev;
    s.C1 = C_of(n, l + 1);
    rc = conv_grads(n, ca, s, n->gda[l]);
    if (rc) return rc;
    rc = conv_dgrad(n, ca, n->gda[l], l, B, n->dcat[l]);
    if (rc) return rc;
    NC(up_grad(n->dcat[l], ca.C0 + ca.C1, ca.C0, prev, gprev, B, n->Hl[l], n->Wl[l], ca.C1, n->st));
  }
  // bottleneck and encoder, coarse to fine
  for (int l = L - 1; l >= 0; l--) {
    const Conv& ca = n->conv[2 * (size_t)l];
    const Conv& cb = n->conv[2 * (size_t)l + 1];
    if (l < L - 1)
      NC(skip_grad(n->dcat[l], C_of(n, l) + C_of(n, l + 1), n->gxp[l], n->s[l], n->gs[l], B,
                   n->Hl[l], n->Wl[l], C_of(n, l), n->st));
    int rc = conv_grads(n, cb, src1(n, n->a[l], C_of(n, l), l, B, cb.dil), n->gs[l]);
    if (rc) return rc;
    rc = conv_dgrad(n, cb, n->gs[l], l, B, n->ga[l]);
    if (rc) return rc;
    NC(relu_bwd(n->ga[l], n->a[l], P_of(n, l, B) * C_of(n, l), n->st));
    const u16* xin = l == 0 ? n->x0 : n->xp[l - 1];
    const int Cx = l == 0 ? CIN0 : C_of(n, l - 1);
    rc = conv_grads(n, ca, src1(n, xin, Cx, l, B, ca.dil), n->ga[l]);
    if (rc) return rc;
    if (l > 0) {
      rc = conv_dgrad(n, ca, n->ga[l], l, B, n->gxp[l - 1]);
      if (rc) return rc;
    }
  }
  if (loss) {
    std::vector<float> lp((size_t)(P0 + 255) / 256), dp(lp.size(), 0.0f);
    NC(hipMemcpyAsync(lp.data(), n->lpart, sizeof(float) * lp.size(), hipMemcpyDeviceToHost,
                      n->st));
    if (n->div_w > 0.0f)
      NC(hipMemcpyAsync(dp.data(), n->dvpart, sizeof(float) * dp.size(), hipMemcpyDeviceToHost,
                        n->st));
    NC(hipStreamSynchronize(n->st));
    double sum = 0.0, dsum = 0.0;
    for (float f : lp) sum += f;
    for (float f : dp) dsum += f;
    *loss = (sum + n->div_w * dsum) / (KOUT * std::max(1.0, fluid));
  }
  return LBM_OK;
}

// x0 / t / mask from host arrays (nb <= Bmax samples, [H][W][3] each)
static int upload_batch(lbm_net* n, const float* in, const float* tg, int nb, double* fluid) {
  const size_t HW = (size_t)n->H * n->W;
  std::vector<int> id(nb);
  for (int b = 0; b < nb; b++) id[b] = b;
  NC(hipMemcpyAsync(n->stage_in, in, sizeof(float) * HW * n->cin * nb, hipMemcpyHostToDevice,
                    n->st));
  if (tg)
    NC(hipMemcpyAsync(n->stage_t, tg, sizeof(float) * HW * 3 * nb, hipMemcpyHostToDevice, n->st));
  NC(hipMemcpyAsync(n->idx, id.data(), sizeof(int) * nb, hipMemcpyHostToDevice, n->st));
  NC(gather(n->stage_in, n->stage_t, n->idx, n->x0, tg ? n->t : nullptr, tg ? n->mask : nullptr,
            nb, (int)HW, n->cin, CIN0, n->st));
  NC(hipStreamSynchronize(n->st));  // the host arrays may go after return
  if (fluid) {
    double f = 0.0;
    for (size_t q = 0; q < HW * nb; q++) f += in[q * n->cin];
    *fluid = f;
  }
  return LBM_OK;
}

int lbm_net_predict(lbm_net* n, const float* in, int nb, float* out) {
  if (!n || !in || !out || nb < 0) return LBM_EINVAL;
  const size_t HW = (size_t)n->H * n->W;
  for (int b0 = 0; b0 < nb; b0 += n->Bmax) {
    const int B = std::min(n->Bmax, nb - b0);
    int rc = upload_batch(n, in + HW * n->cin * b0, nullptr, B, nullptr);
    if (rc) return rc;
    rc = forward(n, B);
    if (rc) return rc;
    NC(hipMemcpyAsync(out + HW * 3 * b0, n->y, sizeof(float) * HW * 3 * B, hipMemcpyDeviceToHost,
                      n->st));
    NC(hipStreamSynchronize(n->st));
  }
  return LBM_OK;
}

int lbm_net_gradients(lbm_net* n, const float* in, const float* target, int nb, double* loss,
                      float* grads) {
  if (!n || !in || !target || nb < 1 || nb > n->Bmax) return LBM_EINVAL;
  double fluid = 0.0;
  int rc = upload_batch(n, in, target, nb, &fluid);
  if (!rc) rc = forward(n, nb);
  if (!rc) rc = backward(n, nb, fluid, loss);
  if (rc) return rc;
  if (grads) {
    NC(hipMemcpyAsync(grads, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->st));
    NC(hipStreamSynchronize(n->st));
  }
  return LBM_OK;
}

int lbm_net_set_data(lbm_net* n, const float* in, const float* target, int count) {
  if (!n || !in || !target || count < 1) return LBM_EINVAL;
  const size_t HW = (size_t)n->H * n->W;
  if (n->set_in) {
    (void)hipFree(n->set_in);
    (void)hipFree(n->set_t);
    for (auto& q : n->allocs)
      if (q == n->set_in || q == n->set_t) q = nullptr;
    n->set_in = n->set_t = nullptr;
  }
  NC(hipMalloc((void**)&n->set_in, sizeof(float) * HW * n->cin * count));
  NC(hipMalloc((void**)&n->set_t, sizeof(float) * HW * 3 * count));
  n->allocs.push_back(n->set_in);
  n->allocs.push_back(n->set_t);
  NC(hipMemcpy(n->set_in, in, sizeof(float) * HW * n->cin * count, hipMemcpyHostToDevice));
  NC(hipMemcpy(n->set_t, target, sizeof(float) * HW * 3 * count, hipMemcpyHostToDevice));
  n->nset = count;
  n->set_fluid.assign(count, 0.0);
  for (int k = 0; k < count; k++) {
    double f = 0.0;
    for (size_t q = 0; q < HW; q++) f += in[((size_t)k * HW + q) * n->cin];
    n->set_fluid[k] = f;
  }
  return LBM_OK;
}

int lbm_net_train(lbm_net* n, const int* idx, int nb, float lr, double* loss) {
  if (!n || !idx || nb < 1 || nb > n->Bmax || !n->set_in) return LBM_EINVAL;
  double fluid = 0.0;
  for (int b = 0; b < nb; b++) {
    if (idx[b] < 0 || idx[b] >= n->nset) {
      set_error("training sample index out of range");
      return LBM_EINVAL;
    }
    fluid += n->set_fluid[idx[b]];
  }
  NC(hipMemcpyAsync(n->idx, idx, sizeof(int) * nb, hipMemcpyHostToDevice, n->st));
  NC(gather(n->set_in, n->set_t, n->idx, n->x0, n->t, n->mask, nb, n->H * n->W, n->cin, CIN0,
            n->st));
  int rc = forward(n, nb);
  if (!rc) rc = backward(n, nb, fluid, loss);
  if (rc) return rc;
  n->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
  NC(adam(n->p, n->g, n->m, n->v, n->pb, (long)n->nparam, lr, b1, b2, 1e-8f, c1, c2, n->wd,
          n->st));
  for (const Conv& c : n->conv) NC(flip_w(n->pb + c.w, n->wt + c.w, c.Cout, c.C0 + c.C1, n->st));
  if (!loss) NC(hipStreamSynchronize(n->st));
  return LBM_OK;
}

int lbm_net_set_div_weight(lbm_net* n, float weight, float su, float sv) {
  if (!n || !(weight >= 0.0f) || !(su > 0.0f) || !(sv > 0.0f)) return LBM_EINVAL;
  n->div_w = weight;
  n->div_su = su;
  n->div_sv = sv;
  return LBM_OK;
}

int lbm_net_set_weight_decay(lbm_net* n, float wd) {
  if (!n || !(wd >= 0.0f) || wd > 1.0f) return LBM_EINVAL;
  n->wd = wd;
  return LBM_OK;
}

int lbm_net_save(lbm_net* n, const char* path) {
  if (!n || !path) return LBM_EINVAL;
  std::vector<float> hp(n->nparam);
  int rc = lbm_net_get_params(n, hp.data());
  if (rc) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return LBM_EIO;
  }
  const int hdr[8] = {0x4c4e4554 /* "LNET" */, n->H, n->W, n->base, n->L, n->cin, n->dil,
                      (int)n->nparam};
  bool ok = std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
            std::fwrite(hp.data(), sizeof(float), hp.size(), f) == hp.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_error(std::string("short write to ") + path);
    return LBM_EIO;
  }
  return LBM_OK;
}

int lbm_net_load(lbm_net* n, const char* path) {
  if (!n || !path) return LBM_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return LBM_EIO;
  }
  int hdr[8];
  std::vector<float> hp(n->nparam);
  bool ok = std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == 0x4c4e4554 && hdr[1] == n->H &&
            hdr[2] == n->W && hdr[3] == n->base && hdr[4] == n->L && hdr[5] == n->cin &&
            hdr[6] == n->dil && hdr[7] == (int)n->nparam &&
            std::fread(hp.data(), sizeof(float), hp.size(), f) == hp.size();
  std::fclose(f);
  if (!ok) {
    set_error(std::string(path) + " does not hold weights of this network shape");
    return LBM_EPARSE;
  }
  return lbm_net_set_params(n, hp.data());
}
