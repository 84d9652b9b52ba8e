// unet.hip -- the learned initial guess: a 3-level 3-D U-Net (16/32/64
// channels, 3x3x3 convolutions, ReLU, 2x2x2 average pooling, 2x2x2
// transposed-convolution upsampling, skip concatenation, 1x1x1 head) on BF16
// activations with FP32 accumulation and FP32 master weights, trained with
// Adam by its own forward and backward kernels (unet_kernels.h).
//
// Inputs per voxel (5 channels): divergence, fluid / air / solid masks,
// previous pressure.  Output: the pressure guess handed to the projection's
// PCG as its starting iterate.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mac_internal.h"
#include "unet_kernels.h"

namespace {

#define UCHK(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return mac_internal_set_error(MAC_E_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)
#define URC(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)

inline int pad32(int k) { return (k + 31) / 32 * 32; }
inline int nblk(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

struct Conv {
  int Cin, Cout, fan_in, Kpad, Kpad2;
  size_t w_off, b_off;  // offsets into the parameter buffer
  bf16 *wf = nullptr, *wb = nullptr;
};
struct Up {
  int Cin, Cout;
  size_t w_off, b_off;
  bf16* wb = nullptr;  // BF16 copy of the weights for the forward pass
};

enum { ENC1A, ENC1B, ENC2A, ENC2B, ENC3A, ENC3B, DEC2A, DEC2B, DEC1A, DEC1B, NCONV };

}  // namespace

struct mac_unet {
  Vol v1, v2, v3;
  size_t n1, n2, n3;
  hipStream_t st = nullptr;
  Conv conv[NCONV];
  Up up2, up1;
  size_t head_w, head_b;
  size_t nparams = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  int adam_t = 0;
  // activations (bf16, channels-last)
  bf16 *X0, *E1a, *C1, *P1, *E2a, *C2, *P2, *E3a, *E3, *D2a, *D2, *D1a, *D1;
  float* out = nullptr;
  // gradients
  bf16 *gD1, *gD1a, *gC1, *gD2, *gD2a, *gC2, *gE3, *gE3a, *gP2, *gE2a, *gP1, *gE1a;
  float *gout = nullptr, *tgt = nullptr, *msk = nullptr, *in5 = nullptr;
  float *partial = nullptr, *hpart = nullptr;
  size_t partial_n = 0;
  std::vector<void*> allocs;
  // conv timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evx[2] = {nullptr, nullptr};  // hand-over to and from a projection stream
  double conv_ms = 0.0, conv_flop = 0.0;
  bool time_convs = false;
  bool wgrad_fp32 = false;  // weight gradients on FP32 FMA instead of the matrix cores
};

namespace {

template <typename T>
int ualloc(mac_unet* u, T** p, size_t n) {
  UCHK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
  u->allocs.push_back((void*)*p);
  return 0;
}

int repack(mac_unet* u) {
  for (int l = 0; l < NCONV; ++l) {
    Conv& c = u->conv[l];
    size_t n = (size_t)c.Cout * c.Kpad + (size_t)c.Cin * c.Kpad2;
    k_pack_weights<<<nblk(n), 256, 0, u->st>>>(c.Cout, c.Cin, u->params + c.w_off, c.wf, c.Kpad, c.wb, c.Kpad2);
  }
  for (Up* up : {&u->up2, &u->up1}) {
    const size_t n = (size_t)8 * up->Cout * up->Cin;
    k_to_bf16<<<nblk(n), 256, 0, u->st>>>(n, u->params + up->w_off, up->wb);
  }
  UCHK(hipGetLastError());
  return 0;
}

inline int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// k_conv3 over a volume: NB = C_out / 16 column blocks, ROWS output rows per
// wave (4 when the height allows it).
void launch_conv(mac_unet* u, const bf16* in, int in_cs, int in_co, int cin, Vol vol, const bf16* wt, int kpad, const float* bias,
                 bf16* out, int out_cs, int out_co, int cout, int relu) {
  // the LDS-tiled kernel where its tile fits: H, W multiples of 16, Cin <= 32
  if (vol.H % 16 == 0 && vol.W % 16 == 0 && (cin == 8 || cin == 16 || cin == 32)) {
    const int blocks = vol.D * (vol.H / 16) * (vol.W / 16);
    const int lc = ilog2(cin);
#define CONVL(NB, C) \
  k_conv3_lds<NB, C><<<blocks, 256, 0, u->st>>>(in, in_cs, in_co, lc, vol, wt, kpad, bias, out, out_cs, out_co, relu, 0)
#define CONVC(NB)                          \
  switch (cin) {                           \
    case 8: CONVL(NB, 8); break;           \
    case 16: CONVL(NB, 16); break;         \
    default: CONVL(NB, 32); break;         \
  }
    switch (cout / 16) {
      case 1: CONVC(1); break;
      case 2: CONVC(2); break;
      default: CONVC(4); break;
    }
#undef CONVC
#undef CONVL
    return;
  }
  const int rows = (vol.H % 4 == 0) ? 4 : 1;
  const int waves = vol.D * (vol.H / rows) * (vol.W / 16);
  const int blocks = (waves + 3) / 4;
  const int lc = ilog2(cin);
#define CONV(NB, R) \
  k_conv3<NB, R><<<blocks, 256, 0, u->st>>>(in, in_cs, in_co, lc, vol, wt, kpad, bias, out, out_cs, out_co, relu, 0)
  if (rows == 4) {
    switch (cout / 16) {
      case 1: CONV(1, 4); break;
      case 2: CONV(2, 4); break;
      default: CONV(4, 4); break;
    }
  } else {
    switch (cout / 16) {
      case 1: CONV(1, 1); break;
      case 2: CONV(2, 1); break;
      default: CONV(4, 1); break;
    }
  }
#undef CONV
}

void conv_fwd(mac_unet* u, int l, const bf16* in, int in_cs, int in_co, Vol vol, bf16* out, int out_cs, int out_co, int relu) {
  Conv& c = u->conv[l];
  if (u->time_convs) hipEventRecord(u->ev0, u->st);
  launch_conv(u, in, in_cs, in_co, c.Cin, vol, c.wf, c.Kpad, u->params + c.b_off, out, out_cs, out_co, c.Cout, relu);
  if (u->time_convs) {
    hipEventRecord(u->ev1, u->st);
    hipEventSynchronize(u->ev1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, u->ev0, u->ev1);
    u->conv_ms += ms;
    u->conv_flop += 2.0 * 27.0 * c.Cin * c.Cout * (double)vol.D * vol.H * vol.W;
  }
}

// data gradient of conv layer l: dX = conv(G, mirrored transposed weights)
void conv_bwd_data(mac_unet* u, int l, const bf16* g, int g_cs, int g_co, Vol vol, bf16* dx, int dx_cs, int dx_co) {
  Conv& c = u->conv[l];
  launch_conv(u, g, g_cs, g_co, c.Cout, vol, c.wb, c.Kpad2, nullptr, dx, dx_cs, dx_co, c.Cin, 0);
}

int conv_wgrad(mac_unet* u, int l, const bf16* g, int g_cs, int g_co, const bf16* x, int x_cs, int x_co, Vol vol) {
  Conv& c = u->conv[l];
  const int nrows = vol.D * vol.H;
  const int rpc = std::max(1, nrows / 128);
  const int nchunk = (nrows + rpc - 1) / rpc;
  const size_t need = (size_t)nchunk * 27 * c.Cout * c.Cin;
  if (need > u->partial_n) return mac_internal_set_error(MAC_E_ARG, "weight-gradient scratch too small");
  if (u->wgrad_fp32) {
    k_conv3_wgrad<<<27 * nchunk, 256, 0, u->st>>>(g, g_cs, g_co, c.Cout, x, x_cs, x_co, c.Cin, vol, rpc, u->partial);
  } else {
    if (vol.W > WG_MAXW || c.Cout % 16 || c.Cin % 8 || c.Cin > 64 || c.Cout + ((c.Cin + 15) & ~15) > 128)
      return mac_internal_set_error(MAC_E_ARG, "weight gradient on the matrix cores: volume wider than 128 voxels");
    k_conv3_wgrad_mfma<<<27 * nchunk, 256, 0, u->st>>>(g, g_cs, g_co, c.Cout, x, x_cs, x_co, c.Cin, vol, rpc, u->partial);
  }
  k_wgrad_reduce<<<nblk(27 * c.Cout * c.Cin), 256, 0, u->st>>>(nchunk, c.Cout, c.Cin, u->partial, u->grads + c.w_off);
  k_bias_grad<<<c.Cout, 256, 0, u->st>>>(g, g_cs, g_co, (size_t)vol.D * vol.H * vol.W, u->grads + c.b_off);
  UCHK(hipGetLastError());
  return 0;
}

// weight and bias gradients of a transposed convolution (its output gradient
// g: the first Cout channels of a g_cs-channel fine tensor; input x: the
// coarse tensor, Cin channels)
int up_wgrad(mac_unet* u, const Up& up, const bf16* g, int g_cs, Vol vc, const bf16* x) {
  if (u->wgrad_fp32) {
    k_up_wgrad<<<nblk(8 * up.Cout * up.Cin + up.Cout), 256, 0, u->st>>>(g, g_cs, 0, up.Cout, vc, x, up.Cin, u->grads + up.w_off,
                                                                      u->grads + up.b_off);
  } else {
    if (vc.W > WG_MAXW || up.Cout % 16 || up.Cin % 16 || up.Cout + up.Cin > 128)
      return mac_internal_set_error(MAC_E_ARG, "transposed-convolution weight gradient on the matrix cores: coarse volume wider than 128 voxels");
    const int nrows = vc.D * vc.H;
    const int rpc = std::max(1, nrows / 256);
    const int nchunk = (nrows + rpc - 1) / rpc;
    if ((size_t)nchunk * 8 * up.Cout * up.Cin > u->partial_n) return mac_internal_set_error(MAC_E_ARG, "weight-gradient scratch too small");
    k_up_wgrad_mfma<<<8 * nchunk, 256, 0, u->st>>>(g, g_cs, 0, up.Cout, vc, x, up.Cin, rpc, u->partial);
    k_up_wgrad_reduce<<<nblk(8 * up.Cout * up.Cin), 256, 0, u->st>>>(nchunk, up.Cout, up.Cin, u->partial, u->grads + up.w_off);
    k_bias_grad<<<up.Cout, 256, 0, u->st>>>(g, g_cs, 0, (size_t)vc.D * vc.H * vc.W * 8, u->grads + up.b_off);
  }
  UCHK(hipGetLastError());
  return 0;
}

void relu_mask(mac_unet* u, size_t nvox, int C, bf16* g, int g_cs, int g_co, const bf16* y, int y_cs, int y_co) {
  k_relu_mask<<<nblk(nvox * C), 256, 0, u->st>>>(nvox, C, g, g_cs, g_co, y, y_cs, y_co);
}

// The forward pass over the volume v1 (the network's own, or a smaller one
// stored contiguously in the same buffers: D and H multiples of 16, W of 64).
int forward(mac_unet* u, Vol v1) {
  const Vol v2 = {v1.D / 2, v1.H / 2, v1.W / 2}, v3 = {v1.D / 4, v1.H / 4, v1.W / 4};
  const size_t n1 = (size_t)v1.D * v1.H * v1.W, n2 = n1 / 8, n3 = n1 / 64;
  conv_fwd(u, ENC1A, u->X0, 8, 0, v1, u->E1a, 16, 0, 1);
  conv_fwd(u, ENC1B, u->E1a, 16, 0, v1, u->C1, 32, 16, 1);
  k_pool_fwd<<<nblk(n2 * 16), 256, 0, u->st>>>(u->C1, 32, 16, 16, v1, u->P1);
  conv_fwd(u, ENC2A, u->P1, 16, 0, v2, u->E2a, 32, 0, 1);
  conv_fwd(u, ENC2B, u->E2a, 32, 0, v2, u->C2, 64, 32, 1);
  k_pool_fwd<<<nblk(n3 * 32), 256, 0, u->st>>>(u->C2, 64, 32, 32, v2, u->P2);
  conv_fwd(u, ENC3A, u->P2, 32, 0, v3, u->E3a, 64, 0, 1);
  conv_fwd(u, ENC3B, u->E3a, 64, 0, v3, u->E3, 64, 0, 1);
  k_up_fwd_mfma<2><<<(v3.D * v3.H * (v3.W / 16) + 3) / 4, 256, 0, u->st>>>(u->E3, 64, v3, u->up2.wb, u->params + u->up2.b_off, u->C2,
                                                                           64, 0);
  conv_fwd(u, DEC2A, u->C2, 64, 0, v2, u->D2a, 32, 0, 1);
  conv_fwd(u, DEC2B, u->D2a, 32, 0, v2, u->D2, 32, 0, 1);
  k_up_fwd_mfma<1><<<(v2.D * v2.H * (v2.W / 16) + 3) / 4, 256, 0, u->st>>>(u->D2, 32, v2, u->up1.wb, u->params + u->up1.b_off, u->C1,
                                                                           32, 0);
  conv_fwd(u, DEC1A, u->C1, 32, 0, v1, u->D1a, 16, 0, 1);
  conv_fwd(u, DEC1B, u->D1a, 16, 0, v1, u->D1, 16, 0, 1);
  k_head_fwd<<<nblk(n1), 256, 0, u->st>>>(u->D1, 16, n1, u->params + u->head_w, u->params + u->head_b, u->out);
  UCHK(hipGetLastError());
  return 0;
}

// gradients of all parameters for d loss / d out = u->gout
int backward(mac_unet* u) {
  const Vol v1 = u->v1, v2 = u->v2, v3 = u->v3;
  const size_t n1 = u->n1, n2 = u->n2, n3 = u->n3;
  // head
  {
    const int nb = 256;
    k_head_wgrad<<<nb, 256, 0, u->st>>>(u->gout, u->D1, 16, n1, u->hpart);
    std::vector<float> hp((size_t)nb * 17);
    UCHK(hipMemcpyAsync(hp.data(), u->hpart, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, u->st));
    UCHK(hipStreamSynchronize(u->st));
    float s[17] = {0};
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < 17; ++c) s[c] += hp[(size_t)b * 17 + c];
    UCHK(hipMemcpyAsync(u->grads + u->head_w, s, sizeof(float) * 16, hipMemcpyHostToDevice, u->st));
    UCHK(hipMemcpyAsync(u->grads + u->head_b, s + 16, sizeof(float), hipMemcpyHostToDevice, u->st));
    k_head_bwd_data<<<nblk(n1 * 16), 256, 0, u->st>>>(u->gout, 16, n1, u->params + u->head_w, u->gD1);
  }
  // dec1b, dec1a
  relu_mask(u, n1, 16, u->gD1, 16, 0, u->D1, 16, 0);
  URC(conv_wgrad(u, DEC1B, u->gD1, 16, 0, u->D1a, 16, 0, v1));
  conv_bwd_data(u, DEC1B, u->gD1, 16, 0, v1, u->gD1a, 16, 0);
  relu_mask(u, n1, 16, u->gD1a, 16, 0, u->D1a, 16, 0);
  URC(conv_wgrad(u, DEC1A, u->gD1a, 16, 0, u->C1, 32, 0, v1));
  conv_bwd_data(u, DEC1A, u->gD1a, 16, 0, v1, u->gC1, 32, 0);
  // up1 (U1 = C1[:, 0:16])
  relu_mask(u, n1, 16, u->gC1, 32, 0, u->C1, 32, 0);
  URC(up_wgrad(u, u->up1, u->gC1, 32, v2, u->D2));
  k_up_bwd_data<<<nblk(n2 * 32), 256, 0, u->st>>>(u->gC1, 32, 0, 16, v2, u->params + u->up1.w_off, 32, u->gD2);
  // dec2b, dec2a
  relu_mask(u, n2, 32, u->gD2, 32, 0, u->D2, 32, 0);
  URC(conv_wgrad(u, DEC2B, u->gD2, 32, 0, u->D2a, 32, 0, v2));
  conv_bwd_data(u, DEC2B, u->gD2, 32, 0, v2, u->gD2a, 32, 0);
  relu_mask(u, n2, 32, u->gD2a, 32, 0, u->D2a, 32, 0);
  URC(conv_wgrad(u, DEC2A, u->gD2a, 32, 0, u->C2, 64, 0, v2));
  conv_bwd_data(u, DEC2A, u->gD2a, 32, 0, v2, u->gC2, 64, 0);
  // up2 (U2 = C2[:, 0:32])
  relu_mask(u, n2, 32, u->gC2, 64, 0, u->C2, 64, 0);
  URC(up_wgrad(u, u->up2, u->gC2, 64, v3, u->E3));
  k_up_bwd_data<<<nblk(n3 * 64), 256, 0, u->st>>>(u->gC2, 64, 0, 32, v3, u->params + u->up2.w_off, 64, u->gE3);
  // enc3b, enc3a
  relu_mask(u, n3, 64, u->gE3, 64, 0, u->E3, 64, 0);
  URC(conv_wgrad(u, ENC3B, u->gE3, 64, 0, u->E3a, 64, 0, v3));
  conv_bwd_data(u, ENC3B, u->gE3, 64, 0, v3, u->gE3a, 64, 0);
  relu_mask(u, n3, 64, u->gE3a, 64, 0, u->E3a, 64, 0);
  URC(conv_wgrad(u, ENC3A, u->gE3a, 64, 0, u->P2, 32, 0, v3));
  conv_bwd_data(u, ENC3A, u->gE3a, 64, 0, v3, u->gP2, 32, 0);
  // pool2 into the E2 half of gC2, then enc2b, enc2a
  k_pool_bwd<<<nblk(n2 * 32), 256, 0, u->st>>>(u->gP2, 32, v2, u->gC2, 64, 32);
  relu_mask(u, n2, 32, u->gC2, 64, 32, u->C2, 64, 32);
  URC(conv_wgrad(u, ENC2B, u->gC2, 64, 32, u->E2a, 32, 0, v2));
  conv_bwd_data(u, ENC2B, u->gC2, 64, 32, v2, u->gE2a, 32, 0);
  relu_mask(u, n2, 32, u->gE2a, 32, 0, u->E2a, 32, 0);
  URC(conv_wgrad(u, ENC2A, u->gE2a, 32, 0, u->P1, 16, 0, v2));
  conv_bwd_data(u, ENC2A, u->gE2a, 32, 0, v2, u->gP1, 16, 0);
  // pool1 into the E1 half of gC1, then enc1b, enc1a
  k_pool_bwd<<<nblk(n1 * 16), 256, 0, u->st>>>(u->gP1, 16, v1, u->gC1, 32, 16);
  relu_mask(u, n1, 16, u->gC1, 32, 16, u->C1, 32, 16);
  URC(conv_wgrad(u, ENC1B, u->gC1, 32, 16, u->E1a, 16, 0, v1));
  conv_bwd_data(u, ENC1B, u->gC1, 32, 16, v1, u->gE1a, 16, 0);
  relu_mask(u, n1, 16, u->gE1a, 16, 0, u->E1a, 16, 0);
  URC(conv_wgrad(u, ENC1A, u->gE1a, 16, 0, u->X0, 8, 0, v1));
  UCHK(hipGetLastError());
  return 0;
}

int upload_input(mac_unet* u, const float* in5) {
  UCHK(hipMemcpyAsync(u->in5, in5, sizeof(float) * u->n1 * 5, hipMemcpyHostToDevice, u->st));
  k_pack_input<<<nblk(u->n1 * 8), 256, 0, u->st>>>(u->n1, u->in5, u->X0);
  UCHK(hipGetLastError());
  return 0;
}

// The grid of a projection context as a volume of the network: its own
// volume (whole = true, the layout of training pairs) or the grid's bounding
// box rounded up to what the kernels take (D and H to 16, W to 64; inference
// then costs what the grid spans, not the network's volume); fails when the
// bounding box does not fit.
int guess_grid(mac_unet* u, mac_ctx* c, GuessGrid* g, mac_grid_view* v, bool whole) {
  URC(mac_internal_view(c, v));
  if (8 * v->ext[0] > u->v1.W || 8 * v->ext[1] > u->v1.H || 8 * v->ext[2] > u->v1.D)
    return mac_internal_set_error(MAC_E_ARG, "grid bounding box larger than the network volume");
  auto up = [](int n, int m, int cap) { return std::min(cap, (n + m - 1) / m * m); };
  Vol a = u->v1;
  if (!whole) a = {up(8 * v->ext[2], 16, u->v1.D), up(8 * v->ext[1], 16, u->v1.H), up(8 * v->ext[0], 64, u->v1.W)};
  g->nb = v->nb;
  g->keys = v->dkeys;
  g->flags = v->flags;
  g->code = v->code;
  g->ff = v->ff;
  g->ndom = (double)v->ndom;
  g->ih2m = v->ih2m;
  g->D = a.D;
  g->H = a.H;
  g->W = a.W;
  return 0;
}

// Network input from the context's right-hand side (already computed, on its
// stream) and `prev`, on the network's stream after the context's work.
int guess_input(mac_unet* u, const GuessGrid& g, const mac_grid_view& v, const double* prev) {
  UCHK(hipEventRecord(u->evx[0], v.st));
  UCHK(hipStreamWaitEvent(u->st, u->evx[0], 0));
  const size_t n1 = (size_t)g.D * g.H * g.W;
  k_guess_fill<<<nblk(n1 * 5), 256, 0, u->st>>>(n1, u->in5);
  k_guess_pack<<<nblk((size_t)g.nb * 512), 256, 0, u->st>>>(g, v.f, prev, u->in5);
  k_pack_input<<<nblk(n1 * 8), 256, 0, u->st>>>(n1, u->in5, u->X0);
  UCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

MAC_HIDDEN int mac_internal_unet_guess(mac_unet* u, mac_ctx* c, const double* prev, float out_scale) {
  if (!u || !c) return mac_internal_set_error(MAC_E_ARG, "null argument");
  if (!(out_scale > 0.f)) return mac_internal_set_error(MAC_E_ARG, "out_scale must be positive");
  GuessGrid g;
  mac_grid_view v;
  URC(guess_grid(u, c, &g, &v, false));
  URC(mac_internal_rhs(c));
  URC(guess_input(u, g, v, prev));
  URC(forward(u, Vol{g.D, g.H, g.W}));
  k_guess_unpack<<<nblk((size_t)g.nb * 512), 256, 0, u->st>>>(g, u->out, out_scale, v.x);
  UCHK(hipGetLastError());
  UCHK(hipEventRecord(u->evx[1], u->st));
  UCHK(hipStreamWaitEvent(v.st, u->evx[1], 0));
  return 0;
}

int mac_unet_guess(mac_unet* u, mac_ctx* c, int use_prev, float out_scale) {
  if (!u || !c) return mac_internal_set_error(MAC_E_ARG, "null argument");
  mac_grid_view v;
  URC(mac_internal_view(c, &v));
  URC(mac_internal_unet_guess(u, c, use_prev ? v.x : nullptr, out_scale));
  UCHK(hipStreamSynchronize(v.st));
  return 0;
}

MAC_HIDDEN int mac_internal_unet_sample(mac_unet* u, mac_ctx* c, const double* prev, float out_scale, float* in5, float* target,
                                        float* mask) {
  if (!u || !c || !in5 || !target || !mask) return mac_internal_set_error(MAC_E_ARG, "null argument");
  if (!(out_scale > 0.f)) return mac_internal_set_error(MAC_E_ARG, "out_scale must be positive");
  GuessGrid g;
  mac_grid_view v;
  URC(guess_grid(u, c, &g, &v, true));
  URC(guess_input(u, g, v, prev));
  UCHK(hipMemsetAsync(u->tgt, 0, sizeof(float) * u->n1, u->st));
  UCHK(hipMemsetAsync(u->msk, 0, sizeof(float) * u->n1, u->st));
  k_guess_target<<<nblk((size_t)g.nb * 512), 256, 0, u->st>>>(g, v.x, out_scale, u->tgt, u->msk);
  UCHK(hipGetLastError());
  UCHK(hipMemcpyAsync(in5, u->in5, sizeof(float) * u->n1 * 5, hipMemcpyDeviceToHost, u->st));
  UCHK(hipMemcpyAsync(target, u->tgt, sizeof(float) * u->n1, hipMemcpyDeviceToHost, u->st));
  UCHK(hipMemcpyAsync(mask, u->msk, sizeof(float) * u->n1, hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return 0;
}

int mac_unet_sample(mac_unet* u, mac_ctx* c, const double* prev, float out_scale, float* in5, float* target, float* mask) {
  if (!u || !c) return mac_internal_set_error(MAC_E_ARG, "null argument");
  mac_grid_view v;
  URC(mac_internal_view(c, &v));
  double* dprev = nullptr;
  const size_t n = (size_t)v.nb * 512;
  if (prev) {
    UCHK(hipMalloc(&dprev, n * sizeof(double)));
    if (hipMemcpyAsync(dprev, prev, n * sizeof(double), hipMemcpyHostToDevice, v.st) != hipSuccess) {
      hipFree(dprev);
      return mac_internal_set_error(MAC_E_HIP, "hipMemcpyAsync of prev failed");
    }
  }
  int rc = mac_internal_unet_sample(u, c, dprev, out_scale, in5, target, mask);
  if (dprev) {
    hipStreamSynchronize(v.st);
    hipFree(dprev);
  }
  return rc;
}

int mac_unet_create(mac_unet** out, int D, int H, int W, unsigned long long seed) {
  if (!out) return mac_internal_set_error(MAC_E_ARG, "null out");
  *out = nullptr;
  if (D <= 0 || H <= 0 || W <= 0 || D % 4 || H % 4 || W % 64)
    return mac_internal_set_error(MAC_E_ARG, "U-Net volume: D, H multiples of 4 and W a multiple of 64");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mac_internal_set_error(MAC_E_NODEV, "no HIP device visible");
  mac_unet* u = new mac_unet();
  u->v1 = {D, H, W};
  u->v2 = {D / 2, H / 2, W / 2};
  u->v3 = {D / 4, H / 4, W / 4};
  u->n1 = (size_t)D * H * W;
  u->n2 = u->n1 / 8;
  u->n3 = u->n1 / 64;
  const int cfg[NCONV][3] = {{8, 16, 5},  {16, 16, 16}, {16, 32, 16}, {32, 32, 32}, {32, 64, 32},
                             {64, 64, 64}, {64, 32, 64}, {32, 32, 32}, {32, 16, 32}, {16, 16, 16}};
  size_t off = 0;
  for (int l = 0; l < NCONV; ++l) {
    Conv& c = u->conv[l];
    c.Cin = cfg[l][0];
    c.Cout = cfg[l][1];
    c.fan_in = 27 * cfg[l][2];
    c.Kpad = pad32(27 * c.Cin);
    c.Kpad2 = pad32(27 * c.Cout);
    c.w_off = off;
    off += (size_t)c.Cout * 27 * c.Cin;
    c.b_off = off;
    off += c.Cout;
  }
  u->up2 = {64, 32, off, 0};
  off += 8 * 32 * 64;
  u->up2.b_off = off;
  off += 32;
  u->up1 = {32, 16, off, 0};
  off += 8 * 16 * 32;
  u->up1.b_off = off;
  off += 16;
  u->head_w = off;
  off += 16;
  u->head_b = off;
  off += 1;
  u->nparams = off;
  int rc = 0;
#define A(p, n) \
  if (!rc) rc = ualloc(u, &(p), (n))
  if (hipStreamCreateWithFlags(&u->st, hipStreamNonBlocking) != hipSuccess) rc = MAC_E_HIP;
  A(u->params, off);
  A(u->grads, off);
  A(u->m, off);
  A(u->v, off);
  for (int l = 0; l < NCONV && !rc; ++l) {
    Conv& c = u->conv[l];
    A(c.wf, (size_t)c.Cout * c.Kpad);
    A(c.wb, (size_t)c.Cin * c.Kpad2);
  }
  A(u->up2.wb, (size_t)8 * 32 * 64);
  A(u->up1.wb, (size_t)8 * 16 * 32);
  const size_t n1 = u->n1, n2 = u->n2, n3 = u->n3;
  A(u->X0, n1 * 8);
  A(u->E1a, n1 * 16);
  A(u->C1, n1 * 32);
  A(u->P1, n2 * 16);
  A(u->E2a, n2 * 32);
  A(u->C2, n2 * 64);
  A(u->P2, n3 * 32);
  A(u->E3a, n3 * 64);
  A(u->E3, n3 * 64);
  A(u->D2a, n2 * 32);
  A(u->D2, n2 * 32);
  A(u->D1a, n1 * 16);
  A(u->D1, n1 * 16);
  A(u->out, n1);
  A(u->gD1, n1 * 16);
  A(u->gD1a, n1 * 16);
  A(u->gC1, n1 * 32);
  A(u->gD2, n2 * 32);
  A(u->gD2a, n2 * 32);
  A(u->gC2, n2 * 64);
  A(u->gE3, n3 * 64);
  A(u->gE3a, n3 * 64);
  A(u->gP2, n3 * 32);
  A(u->gE2a, n2 * 32);
  A(u->gP1, n2 * 16);
  A(u->gE1a, n1 * 16);
  A(u->gout, n1);
  A(u->tgt, n1);
  A(u->msk, n1);
  A(u->in5, n1 * 5);
  A(u->hpart, 256 * 17);
  u->partial_n = (size_t)130 * 27 * 64 * 64;
  A(u->partial, u->partial_n);
#undef A
  if (rc) {
    mac_unet_destroy(u);
    return rc;
  }
  hipEventCreate(&u->ev0);
  hipEventCreate(&u->ev1);
  hipEventCreateWithFlags(&u->evx[0], hipEventDisableTiming);
  hipEventCreateWithFlags(&u->evx[1], hipEventDisableTiming);
  // He-normal initial weights (std sqrt(2 / fan_in)), zero biases
  std::vector<float> p(off, 0.f);
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int l = 0; l < NCONV; ++l) {
    Conv& c = u->conv[l];
    double sd = std::sqrt(2.0 / c.fan_in);
    for (size_t i = 0; i < (size_t)c.Cout * 27 * c.Cin; ++i) p[c.w_off + i] = (float)(sd * nd(gen));
  }
  for (Up* up : {&u->up2, &u->up1}) {
    double sd = std::sqrt(2.0 / up->Cin);
    for (size_t i = 0; i < (size_t)8 * up->Cout * up->Cin; ++i) p[up->w_off + i] = (float)(sd * nd(gen));
  }
  for (int i = 0; i < 16; ++i) p[u->head_w + i] = (float)(0.25 * nd(gen));
  rc = mac_unet_set_params(u, p.data());
  if (rc) {
    mac_unet_destroy(u);
    return rc;
  }
  *out = u;
  return 0;
}

void mac_unet_destroy(mac_unet* u) {
  if (!u) return;
  if (u->st) hipStreamSynchronize(u->st);
  for (void* p : u->allocs) hipFree(p);
  if (u->ev0) hipEventDestroy(u->ev0);
  if (u->ev1) hipEventDestroy(u->ev1);
  for (hipEvent_t e : u->evx)
    if (e) hipEventDestroy(e);
  if (u->st) hipStreamDestroy(u->st);
  delete u;
}

long long mac_unet_num_params(mac_unet* u) { return u ? (long long)u->nparams : 0; }

int mac_unet_get_params(mac_unet* u, float* out) {
  if (!u || !out) return mac_internal_set_error(MAC_E_ARG, "null argument");
  UCHK(hipMemcpyAsync(out, u->params, sizeof(float) * u->nparams, hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return 0;
}

int mac_unet_set_params(mac_unet* u, const float* in) {
  if (!u || !in) return mac_internal_set_error(MAC_E_ARG, "null argument");
  UCHK(hipMemcpyAsync(u->params, in, sizeof(float) * u->nparams, hipMemcpyHostToDevice, u->st));
  UCHK(hipMemsetAsync(u->m, 0, sizeof(float) * u->nparams, u->st));
  UCHK(hipMemsetAsync(u->v, 0, sizeof(float) * u->nparams, u->st));
  u->adam_t = 0;
  URC(repack(u));
  UCHK(hipStreamSynchronize(u->st));
  return 0;
}

int mac_unet_forward(mac_unet* u, const float* in5, float* out) {
  if (!u || !in5) return mac_internal_set_error(MAC_E_ARG, "null argument");
  URC(upload_input(u, in5));
  URC(forward(u, u->v1));
  if (out) UCHK(hipMemcpyAsync(out, u->out, sizeof(float) * u->n1, hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return 0;
}

int mac_unet_backward(mac_unet* u, const float* gout, float* grads) {
  if (!u || !gout) return mac_internal_set_error(MAC_E_ARG, "null argument");
  UCHK(hipMemcpyAsync(u->gout, gout, sizeof(float) * u->n1, hipMemcpyHostToDevice, u->st));
  URC(backward(u));
  if (grads) UCHK(hipMemcpyAsync(grads, u->grads, sizeof(float) * u->nparams, hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return 0;
}

int mac_unet_train_step(mac_unet* u, const float* in5, const float* target, const float* mask, float lr, float* loss) {
  if (!u || !in5 || !target || !mask) return mac_internal_set_error(MAC_E_ARG, "null argument");
  URC(upload_input(u, in5));
  UCHK(hipMemcpyAsync(u->tgt, target, sizeof(float) * u->n1, hipMemcpyHostToDevice, u->st));
  UCHK(hipMemcpyAsync(u->msk, mask, sizeof(float) * u->n1, hipMemcpyHostToDevice, u->st));
  double msum = 0;
  for (size_t i = 0; i < u->n1; ++i) msum += mask[i];
  if (msum <= 0) return 0;
  URC(forward(u, u->v1));
  const int nb = 256;
  k_mse<<<nb, 256, 0, u->st>>>(u->n1, u->out, u->tgt, u->msk, (float)(1.0 / msum), u->gout, u->hpart);
  std::vector<float> part(nb);
  UCHK(hipMemcpyAsync(part.data(), u->hpart, sizeof(float) * nb, hipMemcpyDeviceToHost, u->st));
  URC(backward(u));
  u->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)u->adam_t)) / (1.f - std::pow(b1, (float)u->adam_t));
  k_adam<<<nblk(u->nparams), 256, 0, u->st>>>(u->nparams, u->params, u->grads, u->m, u->v, lr_t, b1, b2, 1e-8f);
  URC(repack(u));
  UCHK(hipStreamSynchronize(u->st));
  double s = 0;
  for (float x : part) s += x;
  if (loss) *loss = (float)(s / msum);
  return 0;
}

// Forward-convolution timing (for the matrix-core utilisation figure).
int mac_unet_wgrad_mode(mac_unet* u, int mode) {
  if (!u || (mode != 0 && mode != 1)) return mac_internal_set_error(MAC_E_ARG, "bad arguments");
  u->wgrad_fp32 = mode == 1;
  return 0;
}

int mac_unet_conv_timing(mac_unet* u, int on, double* ms, double* flop) {
  if (!u) return mac_internal_set_error(MAC_E_ARG, "null argument");
  if (ms) *ms = u->conv_ms;
  if (flop) *flop = u->conv_flop;
  u->time_convs = on != 0;
  u->conv_ms = 0.0;
  u->conv_flop = 0.0;
  return 0;
}

}  // extern "C"
