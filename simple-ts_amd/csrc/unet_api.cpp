// Host runtime of the warm-start U-Net (include/sts_unet.h): parameter and
// activation buffers, the forward and backward schedules over the kernels of
// unet_kernels.hip, Adam, the device-resident training set.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sts_internal.h"
#include "sts_unet.h"
#include "unet_launch.h"

namespace {

thread_local std::string u_err;

int ufail(int code, const std::string &m) {
  u_err = m;
  sts_internal_set_error(m);
  return code;
}

#define U_OK(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return ufail(STS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef uint16_t u16;

struct Conv {
  int Cin, Cx, Cout, ks, level;  // Cx: channel stride of the input tensor
  int Kp, Kpb;                   // forward / backward-data k extents
  size_t w_off, b_off;           // offsets in the flat parameter array
  u16 *Wt = nullptr, *Wb = nullptr;
};

int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Conv indices in parameter order.
enum { E1A, E1B, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1A, D1B, HEAD, NCONV };

}  // namespace

struct sts_unet {
  int cin = 0, cout = 0, base = 0, nmax = 0, H = 0, W = 0, Cx0 = 0;
  hipStream_t stream = nullptr;
  std::vector<Conv> L;
  size_t nparams = 0;
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
  int adam_t = 0;
  // activations
  float *xin = nullptr, *y = nullptr, *dy = nullptr, *tg = nullptr;
  uint8_t *mask = nullptr;
  u16 *x = nullptr, *act[NCONV] = {}, *pl[3] = {}, *cat[3] = {};  // cat[0] = c1 (level 0) .. cat[2] = c3
  u16 *dz = nullptr, *gA = nullptr, *gB = nullptr, *gP = nullptr, *dc[3] = {};
  double *loss = nullptr;
  int *slots_dev = nullptr;
  // training set
  int ds_n = 0;
  u16 *ds_x = nullptr;
  float *ds_t = nullptr;
  uint8_t *ds_m = nullptr;
  std::vector<long long> ds_count;

  int h(int l) const { return H >> l; }
  int w(int l) const { return W >> l; }
  size_t npix(int n, int l) const { return (size_t)n * h(l) * w(l); }
};

namespace {

int alloc_u16(u16 **p, size_t n) {
  U_OK(hipMalloc(p, sizeof(u16) * (n ? n : 1)));
  return STS_OK;
}

void free_all(sts_unet *u) {
  for (Conv &c : u->L) {
    (void)hipFree(c.Wt);
    (void)hipFree(c.Wb);
  }
  for (void *p : {(void *)u->P, (void *)u->G, (void *)u->Mo, (void *)u->Vo, (void *)u->xin, (void *)u->y,
                  (void *)u->dy, (void *)u->tg, (void *)u->mask, (void *)u->x, (void *)u->dz, (void *)u->gA,
                  (void *)u->gB, (void *)u->gP, (void *)u->loss, (void *)u->slots_dev, (void *)u->ds_x,
                  (void *)u->ds_t, (void *)u->ds_m})
    (void)hipFree(p);
  for (int k = 0; k < NCONV; k++) (void)hipFree(u->act[k]);
  for (int k = 0; k < 3; k++) {
    (void)hipFree(u->pl[k]);
    (void)hipFree(u->cat[k]);
    (void)hipFree(u->dc[k]);
  }
  if (u->stream) (void)hipStreamDestroy(u->stream);
}

// Re-derives the BF16 weight layouts from the FP32 master parameters.
int pack_weights(sts_unet *u) {
  for (int k = 0; k < NCONV; k++) {
    Conv &c = u->L[k];
    const float *W = u->P + c.w_off;
    if (k == HEAD) {
      U_OK(unet_pack_head(W, c.Cin * c.Cout, c.Wt, u->stream));
      continue;
    }
    U_OK(unet_pack_fwd(W, 9, c.Cin, c.Cx, c.Cout, c.Kp, c.Wt, u->stream));
    if (c.Wb) U_OK(unet_pack_bwd(W, 9, c.Cin, c.Cout, c.Kpb, c.Wb, u->stream));
  }
  return STS_OK;
}

int conv_fwd(sts_unet *u, int k, const u16 *X, int n, u16 *Y, const u16 *skip = nullptr, int Cb = 0) {
  const Conv &c = u->L[k];
  ConvArgs a{X, c.Cx, n, u->h(c.level), u->w(c.level), 3, c.Wt, c.Kp, u->P + c.b_off, c.Cout, Y, 1, skip, Cb};
  U_OK(unet_conv(a, u->stream));
  return STS_OK;
}

// Decoder conv on [nearest-up(half), skip]: materialised (training keeps the
// concatenation for the backward pass) or read in place by the convolution.
int conv_cat(sts_unet *u, int k, const u16 *half, int Cb, const u16 *skip, int Cs, int n, u16 *cat, bool keep,
             u16 *Y) {
  const int l = u->L[k].level;
  if (!keep) return conv_fwd(u, k, half, n, Y, skip, Cb);
  U_OK(unet_upcat(half, Cb, skip, Cs, n, u->h(l), u->w(l), cat, u->stream));
  return conv_fwd(u, k, cat, n, Y);
}

// Backward of conv k + ReLU: g = dL/d(output) (BF16); writes the parameter
// gradients and, if dX, the input gradient (BF16).
int conv_back(sts_unet *u, int k, const u16 *Xin, int n, const u16 *g, u16 *dX) {
  const Conv &c = u->L[k];
  const int Hh = u->h(c.level), Ww = u->w(c.level);
  const size_t np = u->npix(n, c.level);
  U_OK(unet_relu_mask(g, u->act[k], np * c.Cout, u->dz, u->stream));
  WgradArgs wa{Xin, c.Cx, c.Cin, n, Hh, Ww, 3, u->dz, c.Cout, u->G + c.w_off, 2048};
  U_OK(unet_wgrad(wa, u->stream));
  U_OK(unet_bgrad(u->dz, (int)np, c.Cout, u->G + c.b_off, u->stream));
  if (dX) {
    ConvArgs a{u->dz, c.Cout, n, Hh, Ww, 3, c.Wb, c.Kpb, nullptr, c.Cx, dX, 0};
    U_OK(unet_conv(a, u->stream));
  }
  return STS_OK;
}

// Forward of the batch already in u->x; result in u->y.  keep: materialise
// the concatenations the backward pass reads.
int forward_dev(sts_unet *u, int n, bool keep = true) {
  int rc;
  const int B = u->base;
  if ((rc = conv_fwd(u, E1A, u->x, n, u->act[E1A])) || (rc = conv_fwd(u, E1B, u->act[E1A], n, u->act[E1B])))
    return rc;
  U_OK(unet_pool(u->act[E1B], n, u->h(0), u->w(0), B, u->pl[0], u->stream));
  if ((rc = conv_fwd(u, E2A, u->pl[0], n, u->act[E2A])) || (rc = conv_fwd(u, E2B, u->act[E2A], n, u->act[E2B])))
    return rc;
  U_OK(unet_pool(u->act[E2B], n, u->h(1), u->w(1), 2 * B, u->pl[1], u->stream));
  if ((rc = conv_fwd(u, E3A, u->pl[1], n, u->act[E3A])) || (rc = conv_fwd(u, E3B, u->act[E3A], n, u->act[E3B])))
    return rc;
  U_OK(unet_pool(u->act[E3B], n, u->h(2), u->w(2), 4 * B, u->pl[2], u->stream));
  if ((rc = conv_fwd(u, BA, u->pl[2], n, u->act[BA])) || (rc = conv_fwd(u, BB, u->act[BA], n, u->act[BB])))
    return rc;
  if ((rc = conv_cat(u, D3A, u->act[BB], 8 * B, u->act[E3B], 4 * B, n, u->cat[2], keep, u->act[D3A])) ||
      (rc = conv_fwd(u, D3B, u->act[D3A], n, u->act[D3B])))
    return rc;
  if ((rc = conv_cat(u, D2A, u->act[D3B], 4 * B, u->act[E2B], 2 * B, n, u->cat[1], keep, u->act[D2A])) ||
      (rc = conv_fwd(u, D2B, u->act[D2A], n, u->act[D2B])))
    return rc;
  if ((rc = conv_cat(u, D1A, u->act[D2B], 2 * B, u->act[E1B], B, n, u->cat[0], keep, u->act[D1A])) ||
      (rc = conv_fwd(u, D1B, u->act[D1A], n, u->act[D1B])))
    return rc;
  const Conv &hd = u->L[HEAD];
  U_OK(unet_head(u->act[D1B], B, hd.Wt, u->P + hd.b_off, u->cout, (int)u->npix(n, 0), u->y, u->stream));
  return STS_OK;
}

// Backward from u->dy (FP32) after forward_dev of the same batch.
int backward_dev(sts_unet *u, int n) {
  const int B = u->base;
  U_OK(hipMemsetAsync(u->G, 0, sizeof(float) * u->nparams, u->stream));
  const Conv &hd = u->L[HEAD];
  U_OK(unet_head_back(u->dy, u->act[D1B], B, hd.Wt, u->cout, (int)u->npix(n, 0), u->gA, u->G + hd.w_off,
                      u->G + hd.b_off, u->stream));
  int rc;
  // decoder level 0
  if ((rc = conv_back(u, D1B, u->act[D1A], n, u->gA, u->gB)) || (rc = conv_back(u, D1A, u->cat[0], n, u->gB, u->dc[0])))
    return rc;
  U_OK(unet_up_back(u->dc[0], 3 * B, 2 * B, n, u->h(0), u->w(0), u->gA, u->stream));
  // decoder level 1
  if ((rc = conv_back(u, D2B, u->act[D2A], n, u->gA, u->gB)) || (rc = conv_back(u, D2A, u->cat[1], n, u->gB, u->dc[1])))
    return rc;
  U_OK(unet_up_back(u->dc[1], 6 * B, 4 * B, n, u->h(1), u->w(1), u->gA, u->stream));
  // decoder level 2
  if ((rc = conv_back(u, D3B, u->act[D3A], n, u->gA, u->gB)) || (rc = conv_back(u, D3A, u->cat[2], n, u->gB, u->dc[2])))
    return rc;
  U_OK(unet_up_back(u->dc[2], 12 * B, 8 * B, n, u->h(2), u->w(2), u->gA, u->stream));
  // bottleneck (level 3)
  if ((rc = conv_back(u, BB, u->act[BA], n, u->gA, u->gB)) || (rc = conv_back(u, BA, u->pl[2], n, u->gB, u->gP)))
    return rc;
  U_OK(unet_pool_back_add(u->gP, u->dc[2], 12 * B, 8 * B, n, u->h(2), u->w(2), 4 * B, u->gA, u->stream));
  // encoder level 2
  if ((rc = conv_back(u, E3B, u->act[E3A], n, u->gA, u->gB)) || (rc = conv_back(u, E3A, u->pl[1], n, u->gB, u->gP)))
    return rc;
  U_OK(unet_pool_back_add(u->gP, u->dc[1], 6 * B, 4 * B, n, u->h(1), u->w(1), 2 * B, u->gA, u->stream));
  // encoder level 1
  if ((rc = conv_back(u, E2B, u->act[E2A], n, u->gA, u->gB)) || (rc = conv_back(u, E2A, u->pl[0], n, u->gB, u->gP)))
    return rc;
  U_OK(unet_pool_back_add(u->gP, u->dc[0], 3 * B, 2 * B, n, u->h(0), u->w(0), B, u->gA, u->stream));
  // encoder level 0
  if ((rc = conv_back(u, E1B, u->act[E1A], n, u->gA, u->gB)) || (rc = conv_back(u, E1A, u->x, n, u->gB, nullptr)))
    return rc;
  return STS_OK;
}

int check_n(const sts_unet *u, int n) {
  if (n < 1 || n > u->nmax) return ufail(STS_ERR_ARG, "batch size must be 1..max_n");
  return STS_OK;
}

}  // namespace

extern "C" int sts_unet_create(int cin, int cout, int base, int max_n, int H, int W, sts_unet **out) {
  if (!out) return ufail(STS_ERR_ARG, "null argument");
  *out = nullptr;
  if (cin < 1 || cin > 32 || cout < 1 || cout > 8 || base < 8 || base % 8 || max_n < 1 || H < 8 || H % 8 ||
      W < 64 || W % 64)
    return ufail(STS_ERR_ARG, "U-Net needs 1 <= cin <= 32, 1 <= cout <= 8, base a multiple of 8, H % 8 == 0, "
                              "W % 64 == 0");
  sts_unet *u = new sts_unet();
  u->cin = cin;
  u->cout = cout;
  u->base = base;
  u->nmax = max_n;
  u->H = H;
  u->W = W;
  u->Cx0 = round_up(cin, 8);
  const int B = base;
  // name order of include/sts_unet.h; (Cin, Cx, Cout, level)
  const int spec[NCONV][4] = {{cin, u->Cx0, B, 0},     {B, B, B, 0},          {B, B, 2 * B, 1},
                              {2 * B, 2 * B, 2 * B, 1}, {2 * B, 2 * B, 4 * B, 2}, {4 * B, 4 * B, 4 * B, 2},
                              {4 * B, 4 * B, 8 * B, 3}, {8 * B, 8 * B, 8 * B, 3}, {12 * B, 12 * B, 4 * B, 2},
                              {4 * B, 4 * B, 4 * B, 2}, {6 * B, 6 * B, 2 * B, 1}, {2 * B, 2 * B, 2 * B, 1},
                              {3 * B, 3 * B, B, 0},     {B, B, B, 0},          {B, B, cout, 0}};
  size_t off = 0;
  for (int k = 0; k < NCONV; k++) {
    Conv c;
    c.Cin = spec[k][0];
    c.Cx = spec[k][1];
    c.Cout = spec[k][2];
    c.level = spec[k][3];
    c.ks = (k == HEAD) ? 1 : 3;
    c.Kp = round_up(c.ks * c.ks * c.Cx, 32);
    c.Kpb = round_up(c.ks * c.ks * c.Cout, 32);
    c.w_off = off;
    off += (size_t)c.ks * c.ks * c.Cin * c.Cout;
    c.b_off = off;
    off += (size_t)c.Cout;
    u->L.push_back(c);
  }
  u->nparams = off;
  auto fail_out = [&](int rc) {
    free_all(u);
    delete u;
    return rc;
  };
#define A_OK(expr)                            \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return fail_out(ufail(STS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
  } while (0)
  A_OK(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  for (int k = 0; k < NCONV; k++) {
    Conv &c = u->L[k];
    A_OK(hipMalloc(&c.Wt, sizeof(u16) * (k == HEAD ? (size_t)c.Cin * c.Cout : (size_t)c.Cout * c.Kp)));
    if (k != E1A && k != HEAD) A_OK(hipMalloc(&c.Wb, sizeof(u16) * (size_t)c.Cin * c.Kpb));
  }
  for (float **p : {&u->P, &u->G, &u->Mo, &u->Vo}) {
    A_OK(hipMalloc(p, sizeof(float) * u->nparams));
    A_OK(hipMemsetAsync(*p, 0, sizeof(float) * u->nparams, u->stream));
  }
  const size_t n0 = u->npix(max_n, 0);
  A_OK(hipMalloc(&u->xin, sizeof(float) * n0 * cin));
  A_OK(hipMalloc(&u->y, sizeof(float) * n0 * cout));
  A_OK(hipMalloc(&u->dy, sizeof(float) * n0 * cout));
  A_OK(hipMalloc(&u->tg, sizeof(float) * n0 * cout));
  A_OK(hipMalloc(&u->mask, n0));
  A_OK(hipMalloc(&u->loss, sizeof(double)));
  A_OK(hipMalloc(&u->slots_dev, sizeof(int) * max_n));
  A_OK(hipMalloc(&u->x, sizeof(u16) * n0 * u->Cx0));
  size_t big = 0;  // largest per-layer tensor (elements)
  for (int k = 0; k < NCONV; k++) {
    const Conv &c = u->L[k];
    const size_t np = u->npix(max_n, c.level);
    if (k != HEAD) A_OK(hipMalloc(&u->act[k], sizeof(u16) * np * c.Cout));
    big = std::max(big, np * (size_t)std::max(c.Cx, c.Cout));
  }
  for (int l = 0; l < 3; l++) {
    A_OK(hipMalloc(&u->pl[l], sizeof(u16) * u->npix(max_n, l + 1) * (B << l)));
    A_OK(hipMalloc(&u->cat[l], sizeof(u16) * u->npix(max_n, l) * 3 * (B << l)));
    A_OK(hipMalloc(&u->dc[l], sizeof(u16) * u->npix(max_n, l) * 3 * (B << l)));
  }
  A_OK(hipMalloc(&u->dz, sizeof(u16) * big));
  A_OK(hipMalloc(&u->gA, sizeof(u16) * big));
  A_OK(hipMalloc(&u->gB, sizeof(u16) * big));
  A_OK(hipMalloc(&u->gP, sizeof(u16) * big));
#undef A_OK
  int rc = pack_weights(u);
  if (rc) return fail_out(rc);
  if (hipStreamSynchronize(u->stream) != hipSuccess) return fail_out(ufail(STS_ERR_HIP, "stream sync"));
  *out = u;
  return STS_OK;
}

extern "C" int sts_unet_destroy(sts_unet *u) {
  if (!u) return STS_OK;
  (void)hipStreamSynchronize(u->stream);
  free_all(u);
  delete u;
  return STS_OK;
}

extern "C" long long sts_unet_param_count(const sts_unet *u) { return u ? (long long)u->nparams : 0; }

extern "C" int sts_unet_set_params(sts_unet *u, const float *host) {
  if (!u || !host) return ufail(STS_ERR_ARG, "null argument");
  U_OK(hipMemcpyAsync(u->P, host, sizeof(float) * u->nparams, hipMemcpyHostToDevice, u->stream));
  int rc = pack_weights(u);
  if (rc) return rc;
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_get_params(sts_unet *u, float *host) {
  if (!u || !host) return ufail(STS_ERR_ARG, "null argument");
  U_OK(hipMemcpyAsync(host, u->P, sizeof(float) * u->nparams, hipMemcpyDeviceToHost, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_get_grads(sts_unet *u, float *host) {
  if (!u || !host) return ufail(STS_ERR_ARG, "null argument");
  U_OK(hipMemcpyAsync(host, u->G, sizeof(float) * u->nparams, hipMemcpyDeviceToHost, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_forward(sts_unet *u, int n, const float *x, float *y) {
  if (!u || !x || !y) return ufail(STS_ERR_ARG, "null argument");
  int rc = check_n(u, n);
  if (rc) return rc;
  const size_t np = u->npix(n, 0);
  U_OK(hipMemcpyAsync(u->xin, x, sizeof(float) * np * u->cin, hipMemcpyHostToDevice, u->stream));
  U_OK(unet_to_bf16(u->xin, u->cin, u->Cx0, np, u->x, u->stream));
  if ((rc = forward_dev(u, n))) return rc;
  U_OK(hipMemcpyAsync(y, u->y, sizeof(float) * np * u->cout, hipMemcpyDeviceToHost, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_infer(sts_unet *u, int n, const float *x, float *y) {
  if (!u || !x || !y) return ufail(STS_ERR_ARG, "null argument");
  int rc = check_n(u, n);
  if (rc) return rc;
  const size_t np = u->npix(n, 0);
  U_OK(hipMemcpyAsync(u->xin, x, sizeof(float) * np * u->cin, hipMemcpyHostToDevice, u->stream));
  U_OK(unet_to_bf16(u->xin, u->cin, u->Cx0, np, u->x, u->stream));
  if ((rc = forward_dev(u, n, false))) return rc;
  U_OK(hipMemcpyAsync(y, u->y, sizeof(float) * np * u->cout, hipMemcpyDeviceToHost, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_backward(sts_unet *u, int n, const float *dy) {
  if (!u || !dy) return ufail(STS_ERR_ARG, "null argument");
  int rc = check_n(u, n);
  if (rc) return rc;
  U_OK(hipMemcpyAsync(u->dy, dy, sizeof(float) * u->npix(n, 0) * u->cout, hipMemcpyHostToDevice, u->stream));
  if ((rc = backward_dev(u, n))) return rc;
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

// Loss + dy on the device, then backward; count = masked cells.
static int loss_backward_dev(sts_unet *u, int n, long long count, double *loss, bool backward) {
  U_OK(hipMemsetAsync(u->loss, 0, sizeof(double), u->stream));
  const float inv = 1.0f / (float)((double)(count > 0 ? count : 1) * u->cout);
  U_OK(unet_mse(u->y, u->tg, u->mask, u->cout, (int)u->npix(n, 0), inv, u->dy, u->loss, u->stream));
  int rc;
  if (backward && (rc = backward_dev(u, n))) return rc;
  if (loss) U_OK(hipMemcpyAsync(loss, u->loss, sizeof(double), hipMemcpyDeviceToHost, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

extern "C" int sts_unet_backward_mse(sts_unet *u, int n, const float *t, const unsigned char *mask, double *loss) {
  if (!u || !t || !mask) return ufail(STS_ERR_ARG, "null argument");
  int rc = check_n(u, n);
  if (rc) return rc;
  const size_t np = u->npix(n, 0);
  long long count = 0;
  for (size_t q = 0; q < np; q++) count += mask[q] ? 1 : 0;
  U_OK(hipMemcpyAsync(u->tg, t, sizeof(float) * np * u->cout, hipMemcpyHostToDevice, u->stream));
  U_OK(hipMemcpyAsync(u->mask, mask, np, hipMemcpyHostToDevice, u->stream));
  return loss_backward_dev(u, n, count, loss, true);
}

extern "C" int sts_unet_adam(sts_unet *u, float lr, float b1, float b2, float eps) {
  if (!u) return ufail(STS_ERR_ARG, "null argument");
  u->adam_t++;
  const float c1 = 1.0f - (float)std::pow((double)b1, u->adam_t), c2 = 1.0f - (float)std::pow((double)b2, u->adam_t);
  U_OK(unet_adam(u->P, u->G, u->Mo, u->Vo, u->nparams, lr, b1, b2, eps, c1, c2, u->stream));
  int rc = pack_weights(u);
  if (rc) return rc;
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

// ---------------------------------------------------------------------------
// Training set and the solver coupling
// ---------------------------------------------------------------------------
extern "C" int sts_unet_dataset(sts_unet *u, int slots) {
  if (!u || slots < 1) return ufail(STS_ERR_ARG, "bad dataset size");
  (void)hipFree(u->ds_x);
  (void)hipFree(u->ds_t);
  (void)hipFree(u->ds_m);
  u->ds_x = nullptr;
  u->ds_t = nullptr;
  u->ds_m = nullptr;
  const size_t np = (size_t)u->H * u->W;
  U_OK(hipMalloc(&u->ds_x, sizeof(u16) * np * u->Cx0 * slots));
  U_OK(hipMalloc(&u->ds_t, sizeof(float) * np * u->cout * slots));
  U_OK(hipMalloc(&u->ds_m, np * slots));
  u->ds_n = slots;
  u->ds_count.assign(slots, 0);
  return STS_OK;
}

static int check_solver(sts_unet *u, const StsView &v) {
  if (v.Nx != u->W || v.Ny != u->H) return ufail(STS_ERR_ARG, "solver grid differs from the network's H x W");
  if (u->cin != 5 || u->cout != 2) return ufail(STS_ERR_ARG, "the solver coupling needs cin = 5, cout = 2");
  return STS_OK;
}

extern "C" int sts_unet_record(sts_unet *u, int slot, sts_ctx *ctx, const sts_unet_norm *nrm) {
  if (!u || !ctx || !nrm || slot < 0 || slot >= u->ds_n) return ufail(STS_ERR_ARG, "bad record arguments");
  StsView v;
  int rc = sts_internal_view(ctx, &v);
  if (rc) return rc;
  if ((rc = check_solver(u, v))) return rc;
  U_OK(hipStreamSynchronize(v.stream));
  const int np = v.Nx * v.Ny;
  const double *const *f = v.f;
  U_OK(unet_pack_state(f[STS_U_PR], f[STS_V_PR], f[STS_P_PR], f[STS_T_PR], v.mask, np, u->Cx0, 1.0 / nrm->u_ref,
                       1.0 / nrm->p_ref, u->ds_x + (size_t)slot * np * u->Cx0, u->stream));
  U_OK(unet_pack_target(f[STS_P], f[STS_P_PR], f[STS_T], f[STS_T_PR], v.mask, np, 1.0 / (nrm->p_ref * nrm->dp_scale),
                        1.0 / nrm->dT_scale, u->ds_t + (size_t)slot * np * u->cout, u->ds_m + (size_t)slot * np,
                        u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  u->ds_count[slot] = v.fluid_cells;
  return STS_OK;
}

static int check_corrector(sts_unet *u, const StsView &v) {
  if (v.Nx != u->W || v.Ny != u->H) return ufail(STS_ERR_ARG, "solver grid differs from the network's H x W");
  if (u->cin != 9 || u->cout != 4) return ufail(STS_ERR_ARG, "the corrector needs cin = 9, cout = 4");
  if (!v.xe[0]) return ufail(STS_ERR_ARG, "the context does not keep its predicted start (sts_keep_predicted_start)");
  return STS_OK;
}

static CorrFields state_fields(const StsView &v, bool previous) {
  const double *const *f = v.f;
  return previous ? CorrFields{{f[STS_U_PR], f[STS_V_PR], f[STS_P_PR], f[STS_T_PR]}}
                  : CorrFields{{f[STS_U], f[STS_V], f[STS_P], f[STS_T]}};
}

extern "C" int sts_unet_record_correction(sts_unet *u, int slot, sts_ctx *ctx, const sts_unet_corr_norm *nrm) {
  if (!u || !ctx || !nrm || slot < 0 || slot >= u->ds_n) return ufail(STS_ERR_ARG, "bad record arguments");
  StsView v;
  int rc = sts_internal_view(ctx, &v);
  if (rc) return rc;
  if ((rc = check_corrector(u, v))) return rc;
  U_OK(hipStreamSynchronize(v.stream));
  const int np = v.Nx * v.Ny;
  const CorrFields xe{{v.xe[0], v.xe[1], v.xe[2], v.xe[3]}};
  U_OK(unet_pack_corr_in(state_fields(v, true), xe, v.mask, np, u->Cx0, 1.0 / nrm->u_ref, 1.0 / nrm->p_ref,
                         1.0 / nrm->step, u->ds_x + (size_t)slot * np * u->Cx0, u->stream));
  U_OK(unet_pack_corr_target(state_fields(v, false), xe, v.mask, np, 1.0 / nrm->u_ref, 1.0 / nrm->p_ref,
                             1.0 / nrm->corr, u->ds_t + (size_t)slot * np * u->cout, u->ds_m + (size_t)slot * np,
                             u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  u->ds_count[slot] = v.fluid_cells;
  return STS_OK;
}

static int batch_loss(sts_unet *u, const int *slots, int n, double *loss, bool train, float lr) {
  if (!u || !slots) return ufail(STS_ERR_ARG, "null argument");
  int rc = check_n(u, n);
  if (rc) return rc;
  long long count = 0;
  for (int b = 0; b < n; b++) {
    if (slots[b] < 0 || slots[b] >= u->ds_n) return ufail(STS_ERR_ARG, "slot out of range");
    count += u->ds_count[slots[b]];
  }
  const int np = u->H * u->W;
  U_OK(hipMemcpyAsync(u->slots_dev, slots, sizeof(int) * n, hipMemcpyHostToDevice, u->stream));
  U_OK(unet_gather(u->ds_x, u->ds_t, u->ds_m, u->slots_dev, n, np, u->Cx0, u->cout, u->x, u->tg, u->mask, u->stream));
  if ((rc = forward_dev(u, n))) return rc;
  if ((rc = loss_backward_dev(u, n, count, loss, train))) return rc;
  if (train) return sts_unet_adam(u, lr, 0.9f, 0.999f, 1e-8f);
  return STS_OK;
}

extern "C" int sts_unet_train_step(sts_unet *u, const int *slots, int n, float lr, double *loss) {
  return batch_loss(u, slots, n, loss, true, lr);
}

extern "C" int sts_unet_eval(sts_unet *u, const int *slots, int n, double *loss) {
  return batch_loss(u, slots, n, loss, false, 0.f);
}

// Corrected start for the solver (see sts_set_corrector).
int sts_internal_unet_correct(sts_unet *u, const StsView &v, const sts_unet_corr_norm *nrm, double *const out[4]) {
  int rc = check_corrector(u, v);
  if (rc) return rc;
  U_OK(hipStreamSynchronize(v.stream));
  const int np = v.Nx * v.Ny;
  const CorrFields xe{{v.xe[0], v.xe[1], v.xe[2], v.xe[3]}};
  U_OK(unet_pack_corr_in(state_fields(v, true), xe, v.mask, np, u->Cx0, 1.0 / nrm->u_ref, 1.0 / nrm->p_ref,
                         1.0 / nrm->step, u->x, u->stream));
  if ((rc = forward_dev(u, 1, false))) return rc;
  CorrBoxes bx{v.Nx, v.px, {}, {}, {}};
  for (int q = 0; q < 4; q++) {
    bx.u[q] = v.box_u[q];
    bx.v[q] = v.box_v[q];
    bx.o[q] = v.box_o[q];
  }
  U_OK(unet_apply_corr(u->y, xe, v.mask, bx, np, nrm->u_ref * nrm->corr, nrm->p_ref * nrm->corr, nrm->corr,
                       CorrOut{{out[0], out[1], out[2], out[3]}}, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}

// Prediction for the solver: candidate p, T into pw, Tw (device, FP64).
int sts_internal_unet_predict(sts_unet *u, const StsView &v, const sts_unet_norm *nrm, double *pw, double *Tw) {
  int rc = check_solver(u, v);
  if (rc) return rc;
  U_OK(hipStreamSynchronize(v.stream));
  const int np = v.Nx * v.Ny;
  const double *const *f = v.f;
  U_OK(unet_pack_state(f[STS_U], f[STS_V], f[STS_P], f[STS_T], v.mask, np, u->Cx0, 1.0 / nrm->u_ref, 1.0 / nrm->p_ref,
                       u->x, u->stream));
  if ((rc = forward_dev(u, 1, false))) return rc;
  U_OK(unet_apply_pred(u->y, f[STS_P], f[STS_T], v.mask, v.Nx, np, v.box_o, nrm->p_ref * nrm->dp_scale,
                       nrm->dT_scale, pw, Tw, u->stream));
  U_OK(hipStreamSynchronize(u->stream));
  return STS_OK;
}
