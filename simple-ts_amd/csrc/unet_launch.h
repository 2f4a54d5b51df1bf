// Launch wrappers of the U-Net kernels (unet_kernels.hip), host side.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Implicit-GEMM convolution Y[p][co] = act(sum_k A[p][k] Wt[co][k] + bias[co]).
struct ConvArgs {
  const uint16_t *X;  // input [N][H][W][Cx], BF16
  int Cx;             // channel stride of X (multiple of 8)
  int N, H, W;
  int ks;              // 3 or 1
  const uint16_t *Wt;  // [Cout][Kp], k = tap*Cx + ci
  int Kp;              // multiple of 32, >= ks*ks*Cx
  const float *bias;   // [Cout] or null
  int Cout;
  uint16_t *Y;  // [N][H][W][Cout], BF16
  int relu;
  // Optional concatenated input read in place: channels [0, Cb) are X at
  // half resolution, nearest-upsampled; channels [Cb, Cx) are X2 (stride
  // Cx - Cb).  X2 == null: X is one tensor of stride Cx.
  const uint16_t *X2 = nullptr;
  int Cb = 0;
};

// Weight gradient dW[t][ci][co] += sum_p X[p+off(t)][ci] dZ[p][co].
struct WgradArgs {
  const uint16_t *X;  // layer input [N][H][W][Cx]
  int Cx, Cin;
  int N, H, W;  // W a multiple of 8
  int ks;
  const uint16_t *dZ;  // [N][H][W][Cout]
  int Cout;
  float *dW;  // [ks*ks][Cin][Cout]
  int chunk;  // pixels per block (multiple of 32)
};

hipError_t unet_conv(const ConvArgs &a, hipStream_t s);
hipError_t unet_wgrad(const WgradArgs &a, hipStream_t s);
hipError_t unet_bgrad(const uint16_t *dZ, int npix, int Cout, float *db, hipStream_t s);
hipError_t unet_pool(const uint16_t *X, int N, int H, int W, int C, uint16_t *Y, hipStream_t s);
hipError_t unet_upcat(const uint16_t *B, int Cb, const uint16_t *S, int Cs, int N, int H, int W, uint16_t *O,
                      hipStream_t s);
hipError_t unet_relu_mask(const uint16_t *G, const uint16_t *Act, size_t n, uint16_t *dZ, hipStream_t s);
hipError_t unet_up_back(const uint16_t *D, int Cd, int Cu, int N, int H, int W, uint16_t *O, hipStream_t s);
hipError_t unet_pool_back_add(const uint16_t *P, const uint16_t *Skip, int Cs, int off, int N, int H, int W, int C,
                              uint16_t *O, hipStream_t s);
hipError_t unet_head(const uint16_t *X, int C, const uint16_t *Wh, const float *b, int Co, int npix, float *Y,
                     hipStream_t s);
hipError_t unet_head_back(const float *dY, const uint16_t *X, int C, const uint16_t *Wh, int Co, int npix,
                          uint16_t *dX, float *dW, float *db, hipStream_t s);
hipError_t unet_mse(const float *Y, const float *T, const uint8_t *M, int Co, int npix, float inv_count, float *dY,
                    double *loss, hipStream_t s);
hipError_t unet_adam(float *P, const float *G, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                     float c1, float c2, hipStream_t s);
hipError_t unet_pack_fwd(const float *W, int taps, int Cin, int Cx, int Co, int Kp, uint16_t *Wt, hipStream_t s);
hipError_t unet_pack_bwd(const float *W, int taps, int Cin, int Co, int Kpb, uint16_t *Wb, hipStream_t s);
hipError_t unet_pack_head(const float *W, int n, uint16_t *Wh, hipStream_t s);
hipError_t unet_to_bf16(const float *X, int C, int Cx, size_t npix, uint16_t *Y, hipStream_t s);
hipError_t unet_pack_state(const double *u, const double *v, const double *p, const double *T, const uint8_t *m,
                           int npix, int Cx, double inv_u, double inv_p, uint16_t *X, hipStream_t s);
hipError_t unet_pack_target(const double *p, const double *p_pr, const double *T, const double *T_pr,
                            const uint8_t *m, int npix, double sp, double sT, float *Tg, uint8_t *M, hipStream_t s);
hipError_t unet_apply_pred(const float *Y, const double *p, const double *T, const uint8_t *m, int Nx, int npix,
                           const int box[4], double sp, double sT, double *pw, double *Tw, hipStream_t s);
// The corrector's coupling (sts_set_corrector): u, v, p, T of one state
// (device, FP64), and the grid with its solve boxes.
struct CorrFields {
  const double *f[4];
};
struct CorrOut {
  double *f[4];
};
struct CorrBoxes {
  int Nx, px;
  int u[4], v[4], o[4];  // x0, x1, y0, y1 of the u, v and p/T boxes
};
hipError_t unet_pack_corr_in(const CorrFields &xn, const CorrFields &xe, const uint8_t *m, int npix, int Cx,
                            double inv_u, double inv_p, double inv_step, uint16_t *X, hipStream_t s);
hipError_t unet_pack_corr_target(const CorrFields &x1, const CorrFields &xe, const uint8_t *m, int npix, double inv_u,
                                 double inv_p, double inv_corr, float *Tg, uint8_t *M, hipStream_t s);
hipError_t unet_apply_corr(const float *Y, const CorrFields &xe, const uint8_t *m, const CorrBoxes &bx, int npix,
                           double su, double sp, double sT, const CorrOut &out, hipStream_t s);
hipError_t unet_gather(const uint16_t *Xs, const float *Ts, const uint8_t *Ms, const int *slots, int n, int npix,
                       int Cx, int Co, uint16_t *X, float *Tg, uint8_t *M, hipStream_t s);
