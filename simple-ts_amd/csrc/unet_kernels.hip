// Kernels of the warm-start U-Net (gfx950): BF16 activations and weights,
// FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Tensors are NHWC with channels fastest; a pixel index p runs over
// (n, j, i) with i fastest.  Every BF16 value is stored as its 16 raw bits.
//
// 3x3 convolution = implicit GEMM  Y[p][co] = sum_k A[p][k] * Wt[co][k],
// k = tap * Cx + ci (Cx = channel stride of the input, a multiple of 8, so
// the 8 consecutive k of one MFMA lane are 8 channels of one tap: one 16-byte
// load), Kp = k range rounded up to 32 with zero weights.  A 16x16x32 MFMA
// takes from lane l = r + 16 g the A row r (pixel) and the B column r (output
// channel), both at k = 8g..8g+7, and leaves D rows 4g..4g+3 of column r.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "unet_launch.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

struct alignas(16) U16x8 {
  u16 v[8];
};

__device__ __forceinline__ u16 f2bf(float f) {  // round to nearest even
  unsigned b = __float_as_uint(f);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (u16)(b >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8 load8(const u16 *p) {
  const U16x8 q = *reinterpret_cast<const U16x8 *>(p);
  return __builtin_bit_cast(bf16x8, q);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, U16x8{}); }

// ---------------------------------------------------------------------------
// Convolution (forward, and backward-data with flipped weights)
// Block: 4 waves stacked along the pixels; a wave computes TM x TN tiles of
// 16 pixels x 16 channels: pixels [p0, p0 + 16 TM) x channels
// [c0, c0 + 16 TN), p0 = (4 blockIdx.x + w) * 16 TM, c0 = 16 TN blockIdx.y.
// Per 32-wide k chunk a lane loads TM + TN 16-byte fragments for TM * TN
// MFMAs; every output sums its k chunks in the same order for any TM, TN.
// ---------------------------------------------------------------------------
template <int TM, int TN>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int npix = a.N * a.H * a.W;
  const int p0 = (blockIdx.x * 4 + wave) * (16 * TM), c0 = blockIdx.y * (16 * TN);
  if (p0 >= npix) return;
  const int rad = a.ks / 2, taps = a.ks * a.ks;
  int pn[TM], pj[TM], pi[TM];
  bool pv[TM];
#pragma unroll
  for (int s = 0; s < TM; s++) {
    const int p = p0 + 16 * s + r;
    pv[s] = p < npix;
    const int q = pv[s] ? p : 0;
    pi[s] = q % a.W;
    const int t = q / a.W;
    pj[s] = t % a.H;
    pn[s] = t / a.H;
  }
  floatx4 acc[TM][TN];
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < a.Kp; kc += 32) {
    const int k = kc + 8 * g;
    const int t = k / a.Cx, ci = k - t * a.Cx;
    const int dy = t / a.ks - rad, dx = t % a.ks - rad;
    bf16x8 A[TM], B[TN];
#pragma unroll
    for (int s = 0; s < TM; s++) {
      A[s] = zero8();
      if (t < taps && pv[s]) {
        const int jj = pj[s] + dy, ii = pi[s] + dx;
        if (jj >= 0 && jj < a.H && ii >= 0 && ii < a.W) {
          if (!a.X2)
            A[s] = load8(a.X + ((size_t)(pn[s] * a.H + jj) * a.W + ii) * a.Cx + ci);
          else if (ci < a.Cb)  // nearest-upsampled half-size tensor (channels [0, Cb))
            A[s] = load8(a.X + ((size_t)(pn[s] * (a.H / 2) + jj / 2) * (a.W / 2) + ii / 2) * a.Cb + ci);
          else  // skip tensor (channels [Cb, Cx))
            A[s] = load8(a.X2 + ((size_t)(pn[s] * a.H + jj) * a.W + ii) * (a.Cx - a.Cb) + (ci - a.Cb));
        }
      }
    }
#pragma unroll
    for (int s = 0; s < TN; s++) {
      const int co = c0 + 16 * s + r;
      B[s] = (co < a.Cout) ? load8(a.Wt + (size_t)co * a.Kp + k) : zero8();
    }
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int y = 0; y < TN; y++) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[x], B[y], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int y = 0; y < TN; y++) {
    const int co = c0 + 16 * y + r;
    if (co >= a.Cout) continue;
    const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int p = p0 + 16 * x + 4 * g + q;
        if (p >= npix) continue;
        float v = acc[x][y][q] + bias;
        if (a.relu) v = fmaxf(v, 0.f);
        a.Y[(size_t)p * a.Cout + co] = f2bf(v);
      }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution with its operands staged in LDS (input channel strides
// that are multiples of 32).  A block computes a 4 RW x CONV_PW pixel tile
// of one image (wave w: rows RW w .. RW w + RW - 1 of the tile, 4 sub-tiles
// of 16 pixels each) for 16 TN output channels.  Per chunk of 32 input channels it stages the
// input tile with a one-pixel halo (zeros outside the image) and the chunk's
// weights of all 9 taps in LDS once; every wave then takes its A and B
// fragments of the 9 taps from there, so an input value is read from memory
// once per chunk instead of once per tap and the weights once per block
// instead of once per wave.  Sums run over chunks, then taps.
// ---------------------------------------------------------------------------
constexpr int CONV_PW = 64, CONV_CI = 32;

// LDS slot of 16-byte vector g (of 4) of the pixel or output channel at
// position c: a pixel's 4 vectors are rotated by c / 2, so that the 8 lanes
// reading vector g of 8 consecutive positions in one LDS access hit 8
// different groups of 4 banks (unrotated, the 64-byte stride between
// positions puts them on 2 groups: 4-way conflicts).
__device__ __forceinline__ int conv_slot(int c, int g) { return (g + (c >> 1)) & 3; }

template <int TN, int RW>
__global__ void __launch_bounds__(256) k_conv_lds(ConvArgs a) {
  constexpr int CONV_PH = 4 * RW, TM = 4 * RW;
  constexpr int NV = CONV_CI / 8;                               // 16-byte vectors per pixel per chunk
  constexpr int NA = (CONV_PH + 2) * (CONV_PW + 2) * NV;        // input vectors per chunk
  constexpr int NB = 9 * 16 * TN * NV;                          // weight vectors per chunk
  constexpr int RA = (NA + 255) / 256, RB = (NB + 255) / 256;  // per thread
  __shared__ u32x4 As[CONV_PH + 2][CONV_PW + 2][NV];
  __shared__ u32x4 Bs[9][16 * TN][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
  const int r = lane & 15, g = lane >> 4;
  const int tiles_x = (a.W + CONV_PW - 1) / CONV_PW, tiles_y = (a.H + CONV_PH - 1) / CONV_PH;
  const int bx = blockIdx.x, n = bx / (tiles_x * tiles_y), ty = (bx / tiles_x) % tiles_y, tx = bx % tiles_x;
  const int i0 = tx * CONV_PW, j0 = ty * CONV_PH, c0 = blockIdx.y * (16 * TN);
  floatx4 acc[TM][TN];
#pragma unroll
  for (int x = 0; x < TM; x++)
#pragma unroll
    for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.f, 0.f, 0.f, 0.f};
  // the chunk at ci0 into registers: input tile (zeros outside the image) and weights
  u32x4 qa[RA], qb[RB];
  auto fetch = [&](int ci0) {
#pragma unroll
    for (int k = 0; k < RA; k++) {
      const int e = t + 256 * k;
      qa[k] = u32x4{0u, 0u, 0u, 0u};
      if (e < NA) {
        const int v = e % NV, pos = e / NV;
        const int ii = i0 + pos % (CONV_PW + 2) - 1, jj = j0 + pos / (CONV_PW + 2) - 1, ci = ci0 + 8 * v;
        if (ii >= 0 && ii < a.W && jj >= 0 && jj < a.H) {
          const u16 *src;
          if (!a.X2) src = a.X + ((size_t)(n * a.H + jj) * a.W + ii) * a.Cx + ci;
          else if (ci < a.Cb)  // nearest-upsampled half-size tensor
            src = a.X + ((size_t)(n * (a.H / 2) + jj / 2) * (a.W / 2) + ii / 2) * a.Cb + ci;
          else  // skip tensor
            src = a.X2 + ((size_t)(n * a.H + jj) * a.W + ii) * (a.Cx - a.Cb) + (ci - a.Cb);
          qa[k] = *reinterpret_cast<const u32x4 *>(src);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < RB; k++) {
      const int e = t + 256 * k;
      qb[k] = u32x4{0u, 0u, 0u, 0u};
      if (e < NB) {
        const int v = e % NV, col = (e / NV) % (16 * TN), tap = e / (NV * 16 * TN), co = c0 + col;
        if (co < a.Cout) qb[k] = *reinterpret_cast<const u32x4 *>(a.Wt + (size_t)co * a.Kp + tap * a.Cx + ci0 + 8 * v);
      }
    }
  };
  fetch(0);
  for (int ci0 = 0; ci0 < a.Cx; ci0 += CONV_CI) {
#pragma unroll
    for (int k = 0; k < RA; k++) {
      const int e = t + 256 * k;
      if (e < NA) {
        const int col = (e / NV) % (CONV_PW + 2);
        As[(e / NV) / (CONV_PW + 2)][col][conv_slot(col, e % NV)] = qa[k];
      }
    }
#pragma unroll
    for (int k = 0; k < RB; k++) {
      const int e = t + 256 * k;
      if (e < NB) {
        const int co = (e / NV) % (16 * TN);
        Bs[e / (NV * 16 * TN)][co][conv_slot(co, e % NV)] = qb[k];
      }
    }
    __syncthreads();
    if (ci0 + CONV_CI < a.Cx) fetch(ci0 + CONV_CI);  // in flight while this chunk's MFMAs run
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      bf16x8 A[TM], B[TN];
#pragma unroll
      for (int s = 0; s < TM; s++) {
        const int col = 16 * (s % 4) + r + 1 + dx;
        A[s] = __builtin_bit_cast(bf16x8, As[RW * wave + s / 4 + 1 + dy][col][conv_slot(col, g)]);
      }
#pragma unroll
      for (int s = 0; s < TN; s++) B[s] = __builtin_bit_cast(bf16x8, Bs[tap][16 * s + r][conv_slot(16 * s + r, g)]);
#pragma unroll
      for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[x], B[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int y = 0; y < TN; y++) {
    const int co = c0 + 16 * y + r;
    if (co >= a.Cout) continue;
    const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int x = 0; x < TM; x++) {
      const int j = j0 + RW * wave + x / 4;
      if (j >= a.H) continue;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i0 + 16 * (x % 4) + 4 * g + q;
        if (i >= a.W) continue;
        float v = acc[x][y][q] + bias;
        if (a.relu) v = fmaxf(v, 0.f);
        a.Y[((size_t)(n * a.H + j) * a.W + i) * a.Cout + co] = f2bf(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient dW[t][ci][co] += sum_p X[p + off(t)][ci] * dZ[p][co]:
// GEMM with M = co, N = k = t*Cx + ci, reduced over pixels.  Wave w of block
// (bx, by, bz) owns k tile [64 bx + 16 w, +16), co tile [16 by, +16), pixels
// [bz * chunk, +chunk); partial sums go to dW with atomicAdd.  W (row length)
// is a multiple of 8, so the 8 pixels of a lane are one row segment.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int npix = a.N * a.H * a.W;
  const int kk = blockIdx.x * 64 + wave * 16 + r;  // B column of this lane
  const int co = blockIdx.y * 16 + r;              // A row of this lane
  const int rad = a.ks / 2, taps = a.ks * a.ks;
  const int t = kk / a.Cx, ci = kk - t * a.Cx;
  const bool kv = t < taps && ci < a.Cin;
  const int dy = kv ? t / a.ks - rad : 0, dx = kv ? t % a.ks - rad : 0;
  const bool cv = co < a.Cout;
  const int pbeg = blockIdx.z * a.chunk, pend = min(npix, pbeg + a.chunk);
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int pc = pbeg; pc < pend; pc += 32) {
    const int p = pc + 8 * g;  // 8 pixels p..p+7, one row segment
    U16x8 av{}, bv{};
    if (p < pend) {
      const int i0 = p % a.W, tt = p / a.W, j = tt % a.H, n = tt / a.H;
      if (cv) {
#pragma unroll
        for (int q = 0; q < 8; q++) av.v[q] = a.dZ[(size_t)(p + q) * a.Cout + co];
      }
      const int jj = j + dy;
      if (kv && jj >= 0 && jj < a.H) {
        const size_t rowb = (size_t)(n * a.H + jj) * a.W;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int ii = i0 + q + dx;
          bv.v[q] = (ii >= 0 && ii < a.W) ? a.X[(rowb + ii) * a.Cx + ci] : (u16)0;
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc,
                                                  0, 0, 0);
  }
  // D rows = co (4g + q of the tile), column = k tile entry r
  const int col_k = blockIdx.x * 64 + wave * 16 + r;
  const int tc = col_k / a.Cx, cc = col_k - tc * a.Cx;
  if (tc >= taps || cc >= a.Cin) return;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int c = blockIdx.y * 16 + 4 * g + q;
    if (c < a.Cout) atomicAdd(&a.dW[((size_t)tc * a.Cin + cc) * a.Cout + c], acc[q]);
  }
}

// db[co] += sum_p dZ[p][co]; block = 256 threads over a pixel chunk.
__global__ void __launch_bounds__(256) k_bgrad(const u16 *dZ, int npix, int Cout, int chunk, float *db) {
  const int pbeg = blockIdx.x * chunk, pend = min(npix, pbeg + chunk);
  for (int co = threadIdx.x; co < Cout; co += blockDim.x) {
    float s = 0.f;
    for (int p = pbeg; p < pend; p++) s += bf2f(dZ[(size_t)p * Cout + co]);
    atomicAdd(&db[co], s);
  }
}

// ---------------------------------------------------------------------------
// Element-wise and layout kernels
// ---------------------------------------------------------------------------
// 2x2 average pool: X [N][H][W][C] -> Y [N][H/2][W/2][C]
__global__ void k_pool(const u16 *X, int N, int H, int W, int C, u16 *Y) {
  const size_t n_out = (size_t)N * (H / 2) * (W / 2) * C;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_out) return;
  const int c = (int)(e % C);
  size_t q = e / C;
  const int i = (int)(q % (W / 2));
  q /= (W / 2);
  const int j = (int)(q % (H / 2));
  const int n = (int)(q / (H / 2));
  const size_t b = ((size_t)(n * H + 2 * j) * W + 2 * i) * C + c;
  const float s = (bf2f(X[b]) + bf2f(X[b + C])) + (bf2f(X[b + (size_t)W * C]) + bf2f(X[b + (size_t)W * C + C]));
  Y[e] = f2bf(s * 0.25f);
}

// concat [nearest-up(B), S]: B [N][H/2][W/2][Cb], S [N][H][W][Cs] -> O [N][H][W][Cb+Cs]
__global__ void k_upcat(const u16 *B, int Cb, const u16 *S, int Cs, int N, int H, int W, u16 *O) {
  const int C = Cb + Cs;
  const size_t n_out = (size_t)N * H * W * C;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_out) return;
  const int c = (int)(e % C);
  const size_t p = e / C;
  const int i = (int)(p % W);
  const size_t t = p / W;
  const int j = (int)(t % H), n = (int)(t / H);
  O[e] = (c < Cb) ? B[((size_t)(n * (H / 2) + j / 2) * (W / 2) + i / 2) * Cb + c] : S[p * Cs + (c - Cb)];
}

// dZ = G * (Act > 0)
__global__ void k_relu_mask(const u16 *G, const u16 *Act, size_t n, u16 *dZ) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  dZ[e] = (bf2f(Act[e]) > 0.f) ? G[e] : (u16)0;
}

// Backward of nearest-up on channels [0, Cu) of D [N][H][W][Cd]:
// O [N][H/2][W/2][Cu] = round(sum of the 4 children)
__global__ void k_up_back(const u16 *D, int Cd, int Cu, int N, int H, int W, u16 *O) {
  const size_t n_out = (size_t)N * (H / 2) * (W / 2) * Cu;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_out) return;
  const int c = (int)(e % Cu);
  size_t q = e / Cu;
  const int i = (int)(q % (W / 2));
  q /= (W / 2);
  const int j = (int)(q % (H / 2)), n = (int)(q / (H / 2));
  const size_t b = ((size_t)(n * H + 2 * j) * W + 2 * i) * Cd + c;
  const size_t rw = (size_t)W * Cd;
  const float s = (bf2f(D[b]) + bf2f(D[b + Cd])) + (bf2f(D[b + rw]) + bf2f(D[b + rw + Cd]));
  O[e] = f2bf(s);
}

// Backward of the pool plus the skip branch: O [N][H][W][C] =
// round(round(P[parent] / 4) + Skip[p][off + c]) with Skip of stride Cs.
__global__ void k_pool_back_add(const u16 *P, const u16 *Skip, int Cs, int off, int N, int H, int W, int C, u16 *O) {
  const size_t n_out = (size_t)N * H * W * C;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_out) return;
  const int c = (int)(e % C);
  const size_t p = e / C;
  const int i = (int)(p % W);
  const size_t t = p / W;
  const int j = (int)(t % H), n = (int)(t / H);
  const float pb = bf2f(f2bf(bf2f(P[((size_t)(n * (H / 2) + j / 2) * (W / 2) + i / 2) * C + c]) * 0.25f));
  O[e] = f2bf(pb + bf2f(Skip[p * Cs + off + c]));
}

// Head (1x1 conv, linear): Y[p][o] = sum_c X[p][c] * Wb[c][o] + b[o], FP32 out.
// Co <= HEAD_MAX_CO; the sums are kept in registers (loops over a constant
// bound, outputs past Co skipped).
// The weights (C x Co <= HEAD_MAX_W) are read from LDS, one broadcast read
// per weight and wavefront.
constexpr int HEAD_MAX_CO = 8, HEAD_MAX_W = 1024;
__global__ void __launch_bounds__(256) k_head(const u16 *X, int C, const u16 *Wb, const float *b, int Co, int npix,
                                              float *Y) {
  __shared__ float Ws[HEAD_MAX_W];
  for (int e = threadIdx.x; e < C * Co; e += blockDim.x) Ws[e] = bf2f(Wb[e]);
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float s[HEAD_MAX_CO];
#pragma unroll
  for (int o = 0; o < HEAD_MAX_CO; o++) s[o] = 0.f;
  for (int c0 = 0; c0 < C; c0 += 8) {  // C is a multiple of 8: 16-byte loads
    const U16x8 x = *reinterpret_cast<const U16x8 *>(X + (size_t)p * C + c0);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const float xv = bf2f(x.v[q]);
#pragma unroll
      for (int o = 0; o < HEAD_MAX_CO; o++)
        if (o < Co) s[o] += xv * Ws[(c0 + q) * Co + o];
    }
  }
#pragma unroll
  for (int o = 0; o < HEAD_MAX_CO; o++)
    if (o < Co) Y[(size_t)p * Co + o] = s[o] + b[o];
}

// Head backward: dX[p][c] = round(sum_o dY[p][o] Wb[c][o]); dW[c][o] and
// db[o] accumulate per block, then atomically.
__global__ void __launch_bounds__(256) k_head_back(const float *dY, const u16 *X, int C, const u16 *Wb, int Co,
                                                   int npix, u16 *dX, float *dW, float *db) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = p < npix;
  for (int c = 0; c < C; c++) {
    float s = 0.f;
    if (ok)
      for (int o = 0; o < Co; o++) s += dY[(size_t)p * Co + o] * bf2f(Wb[c * Co + o]);
    if (ok) dX[(size_t)p * C + c] = f2bf(s);
  }
  __shared__ float red[256];
  for (int q = 0; q < C * Co; q++) {  // dW[c][o], one block reduction each
    const int c = q / Co, o = q - c * Co;
    red[threadIdx.x] = ok ? dY[(size_t)p * Co + o] * bf2f(X[(size_t)p * C + c]) : 0.f;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&dW[q], red[0]);
    __syncthreads();
  }
  for (int o = 0; o < Co; o++) {
    red[threadIdx.x] = ok ? dY[(size_t)p * Co + o] : 0.f;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[o], red[0]);
    __syncthreads();
  }
}

// Masked MSE: dY = (Y - T) * m / count; loss += sum((Y - T)^2 m) / (2 count)
__global__ void __launch_bounds__(256) k_mse(const float *Y, const float *T, const uint8_t *M, int Co, int npix,
                                             float inv_count, float *dY, double *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (p < npix) {
    const float m = M[p] ? 1.f : 0.f;
    for (int o = 0; o < Co; o++) {
      const float e = (Y[(size_t)p * Co + o] - T[(size_t)p * Co + o]) * m;
      dY[(size_t)p * Co + o] = e * inv_count;
      s += (double)e * (double)e;
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * 0.5 * (double)inv_count);
}

// Adam on FP32 master parameters.
__global__ void k_adam(float *P, const float *G, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float g = G[e];
  const float mm = b1 * m[e] + (1.f - b1) * g;
  const float vv = b2 * v[e] + (1.f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  const float mh = mm / c1, vh = vv / c2;
  P[e] -= lr * mh / (sqrtf(vh) + eps);
}

// Weight layouts from the FP32 master W[t][ci][co]:
//   Wt[co][t*Cx + ci]       forward (zero for ci >= Cin and k >= taps*Cx)
//   Wb[ci][t'*Co + co] = W[taps-1-t'][ci][co]   backward data (flipped taps)
//   Wh[ci][co]              1x1 head
__global__ void k_pack_fwd(const float *W, int taps, int Cin, int Cx, int Co, int Kp, u16 *Wt) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)Co * Kp) return;
  const int co = (int)(e / Kp), k = (int)(e % Kp);
  const int t = k / Cx, ci = k - t * Cx;
  Wt[e] = (t < taps && ci < Cin) ? f2bf(W[((size_t)t * Cin + ci) * Co + co]) : (u16)0;
}
__global__ void k_pack_bwd(const float *W, int taps, int Cin, int Co, int Kpb, u16 *Wb) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)Cin * Kpb) return;
  const int ci = (int)(e / Kpb), k = (int)(e % Kpb);
  const int t = k / Co, co = k - t * Co;
  Wb[e] = (t < taps) ? f2bf(W[((size_t)(taps - 1 - t) * Cin + ci) * Co + co]) : (u16)0;
}
__global__ void k_pack_head(const float *W, int n, u16 *Wh) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) Wh[e] = f2bf(W[e]);
}

// FP32 NHWC [n][C] -> BF16 with channel stride Cx (zero padding channels).
__global__ void k_to_bf16(const float *X, int C, int Cx, size_t npix, u16 *Y) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  for (int c0 = 0; c0 < Cx; c0 += 8) {  // Cx is a multiple of 8: 16-byte stores
    U16x8 o;
#pragma unroll
    for (int q = 0; q < 8; q++) o.v[q] = (c0 + q < C) ? f2bf(X[p * C + c0 + q]) : (u16)0;
    *reinterpret_cast<U16x8 *>(Y + p * Cx + c0) = o;
  }
}

// Solver state -> network input [H][W][Cx] (H = Ny rows j, W = Nx columns i):
// (u/u_ref, v/u_ref, p/p_ref - 1, T - 1, fluid mask), zero padding channels.
__global__ void k_pack_state(const double *u, const double *v, const double *p, const double *T, const uint8_t *m,
                             int npix, int Cx, double inv_u, double inv_p, u16 *X) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  u16 *o = X + (size_t)q * Cx;
  o[0] = f2bf((float)(u[q] * inv_u));
  o[1] = f2bf((float)(v[q] * inv_u));
  o[2] = f2bf((float)(p[q] * inv_p - 1.0));
  o[3] = f2bf((float)(T[q] - 1.0));
  o[4] = m[q] ? f2bf(1.f) : (u16)0;
  for (int c = 5; c < Cx; c++) o[c] = 0;
}

// Training target (dp / (p_ref dp_scale), dT / dT_scale) and mask.
__global__ void k_pack_target(const double *p, const double *p_pr, const double *T, const double *T_pr,
                              const uint8_t *m, int npix, double sp, double sT, float *Tg, uint8_t *M) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  Tg[2 * (size_t)q] = (float)((p[q] - p_pr[q]) * sp);
  Tg[2 * (size_t)q + 1] = (float)((T[q] - T_pr[q]) * sT);
  M[q] = m[q];
}

// Predicted guess: p' = p + y0 * (p_ref dp_scale), T' = T + y1 * dT_scale on
// the cells loop 3 updates (fluid cells of the p/T box [x0,x1) x [y0,y1)),
// copies elsewhere: fixed-value cells keep their boundary values.
__global__ void k_apply_pred(const float *Y, const double *p, const double *T, const uint8_t *m, int Nx, int npix,
                             int x0, int x1, int y0, int y1, double sp, double sT, double *pw, double *Tw) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  const int i = q % Nx, j = q / Nx;
  const bool upd = m[q] && i >= x0 && i < x1 && j >= y0 && j < y1;
  pw[q] = upd ? p[q] + (double)Y[2 * (size_t)q] * sp : p[q];
  Tw[q] = upd ? T[q] + (double)Y[2 * (size_t)q + 1] * sT : T[q];
}

// Corrector input [H][W][Cx]: the state x^n of the step (u/u_ref, v/u_ref,
// p/p_ref - 1, T - 1), the predicted start's change x_e - x^n in the same
// units divided by `step`, the fluid mask; zero padding channels.
__global__ void k_pack_corr_in(CorrFields xn, CorrFields xe, const uint8_t *m, int npix, int Cx, double inv_u,
                               double inv_p, double inv_step, u16 *X) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  u16 *o = X + (size_t)q * Cx;
  const double sc[4] = {inv_u, inv_u, inv_p, 1.0};
  o[0] = f2bf((float)(xn.f[0][q] * inv_u));
  o[1] = f2bf((float)(xn.f[1][q] * inv_u));
  o[2] = f2bf((float)(xn.f[2][q] * inv_p - 1.0));
  o[3] = f2bf((float)(xn.f[3][q] - 1.0));
#pragma unroll
  for (int c = 0; c < 4; c++) o[4 + c] = f2bf((float)((xe.f[c][q] - xn.f[c][q]) * sc[c] * inv_step));
  o[8] = m[q] ? f2bf(1.f) : (u16)0;
  for (int c = 9; c < Cx; c++) o[c] = 0;
}

// Corrector target: the converged step x^(n+1) minus the predicted start
// x_e, normalised as the input and divided by `corr`; mask = fluid cells.
__global__ void k_pack_corr_target(CorrFields x1, CorrFields xe, const uint8_t *m, int npix, double inv_u,
                                   double inv_p, double inv_corr, float *Tg, uint8_t *M) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  const double sc[4] = {inv_u, inv_u, inv_p, 1.0};
#pragma unroll
  for (int c = 0; c < 4; c++) Tg[4 * (size_t)q + c] = (float)((x1.f[c][q] - xe.f[c][q]) * sc[c] * inv_corr);
  M[q] = m[q];
}

// Corrected start x = x_e + y * scale on the cells loop 2 updates (u: fluid
// faces of the u box, v: of the v box, p and T: fluid cells of the p/T
// box); every other cell keeps x_e.  A value that is not finite, or a p or T
// that is not positive, keeps x_e.
__global__ void k_apply_corr(const float *Y, CorrFields xe, const uint8_t *m, CorrBoxes bx, int npix, double su,
                             double sp, double sT, CorrOut out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  const int Nx = bx.Nx, i = q % Nx, j = q / Nx;
  const int iw = i > 0 ? q - 1 : (bx.px ? q + Nx - 1 : q), js = j > 0 ? q - Nx : q;
  const bool c = m[q];
  const bool upd[4] = {c && m[iw] && i >= bx.u[0] && i < bx.u[1] && j >= bx.u[2] && j < bx.u[3],
                       c && m[js] && i >= bx.v[0] && i < bx.v[1] && j >= bx.v[2] && j < bx.v[3],
                       c && i >= bx.o[0] && i < bx.o[1] && j >= bx.o[2] && j < bx.o[3],
                       c && i >= bx.o[0] && i < bx.o[1] && j >= bx.o[2] && j < bx.o[3]};
  const double sc[4] = {su, su, sp, sT};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double e = xe.f[k][q];
    double x = upd[k] ? e + (double)Y[4 * (size_t)q + k] * sc[k] : e;
    if (!isfinite(x) || (k >= 2 && !(x > 0.0))) x = e;
    out.f[k][q] = x;
  }
}

// Gathers dataset samples into a batch.
__global__ void k_gather(const u16 *Xs, const float *Ts, const uint8_t *Ms, const int *slots, int n, int npix, int Cx,
                         int Co, u16 *X, float *Tg, uint8_t *M) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * npix) return;
  const int b = (int)(e / npix), q = (int)(e % npix);
  const size_t src = (size_t)slots[b] * npix + q;
  for (int c = 0; c < Cx; c++) X[e * Cx + c] = Xs[src * Cx + c];
  for (int c = 0; c < Co; c++) Tg[e * Co + c] = Ts[src * Co + c];
  M[e] = Ms[src];
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

hipError_t unet_conv(const ConvArgs &a, hipStream_t s) {
  const size_t npix = (size_t)a.N * a.H * a.W;
  if (a.ks == 3 && a.Cx % CONV_CI == 0 && (!a.X2 || a.Cb % CONV_CI == 0)) {
    // 64 or 32 channels x 4 rows per block (32 channels x 8 rows, two rows per
    // wave, needs 291 registers: one wave per SIMD, 45% slower)
    const unsigned tiles = (unsigned)a.N * nblk(a.H, 4) * nblk(a.W, CONV_PW);
    if (a.Cout >= 64) hipLaunchKernelGGL((k_conv_lds<4, 1>), dim3(tiles, nblk(a.Cout, 64)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_conv_lds<2, 1>), dim3(tiles, nblk(a.Cout, 32)), dim3(256), 0, s, a);
    return hipGetLastError();
  }
  // 4 pixel tiles per wave; 4 channel tiles where there are 64+ channels
  if (a.Cout >= 64)
    hipLaunchKernelGGL((k_conv<4, 4>), dim3(nblk(npix, 256), nblk(a.Cout, 64)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_conv<4, 2>), dim3(nblk(npix, 256), nblk(a.Cout, 32)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t unet_wgrad(const WgradArgs &a, hipStream_t s) {
  const size_t npix = (size_t)a.N * a.H * a.W;
  const int Kp = ((a.ks * a.ks * a.Cx + 31) / 32) * 32;
  hipLaunchKernelGGL(k_wgrad, dim3(nblk(Kp, 64), nblk(a.Cout, 16), nblk(npix, a.chunk)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t unet_bgrad(const uint16_t *dZ, int npix, int Cout, float *db, hipStream_t s) {
  const int chunk = 1024;
  hipLaunchKernelGGL(k_bgrad, dim3(nblk(npix, chunk)), dim3(256), 0, s, dZ, npix, Cout, chunk, db);
  return hipGetLastError();
}
hipError_t unet_pool(const uint16_t *X, int N, int H, int W, int C, uint16_t *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_pool, dim3(nblk((size_t)N * (H / 2) * (W / 2) * C, 256)), dim3(256), 0, s, X, N, H, W, C, Y);
  return hipGetLastError();
}
hipError_t unet_upcat(const uint16_t *B, int Cb, const uint16_t *S, int Cs, int N, int H, int W, uint16_t *O,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_upcat, dim3(nblk((size_t)N * H * W * (Cb + Cs), 256)), dim3(256), 0, s, B, Cb, S, Cs, N, H, W,
                     O);
  return hipGetLastError();
}
hipError_t unet_relu_mask(const uint16_t *G, const uint16_t *Act, size_t n, uint16_t *dZ, hipStream_t s) {
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(n, 256)), dim3(256), 0, s, G, Act, n, dZ);
  return hipGetLastError();
}
hipError_t unet_up_back(const uint16_t *D, int Cd, int Cu, int N, int H, int W, uint16_t *O, hipStream_t s) {
  hipLaunchKernelGGL(k_up_back, dim3(nblk((size_t)N * (H / 2) * (W / 2) * Cu, 256)), dim3(256), 0, s, D, Cd, Cu, N, H,
                     W, O);
  return hipGetLastError();
}
hipError_t unet_pool_back_add(const uint16_t *P, const uint16_t *Skip, int Cs, int off, int N, int H, int W, int C,
                              uint16_t *O, hipStream_t s) {
  hipLaunchKernelGGL(k_pool_back_add, dim3(nblk((size_t)N * H * W * C, 256)), dim3(256), 0, s, P, Skip, Cs, off, N, H,
                     W, C, O);
  return hipGetLastError();
}
hipError_t unet_head(const uint16_t *X, int C, const uint16_t *Wh, const float *b, int Co, int npix, float *Y,
                     hipStream_t s) {
  if (Co > HEAD_MAX_CO || C * Co > HEAD_MAX_W) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_head, dim3(nblk(npix, 256)), dim3(256), 0, s, X, C, Wh, b, Co, npix, Y);
  return hipGetLastError();
}
hipError_t unet_head_back(const float *dY, const uint16_t *X, int C, const uint16_t *Wh, int Co, int npix,
                          uint16_t *dX, float *dW, float *db, hipStream_t s) {
  hipLaunchKernelGGL(k_head_back, dim3(nblk(npix, 256)), dim3(256), 0, s, dY, X, C, Wh, Co, npix, dX, dW, db);
  return hipGetLastError();
}
hipError_t unet_mse(const float *Y, const float *T, const uint8_t *M, int Co, int npix, float inv_count, float *dY,
                    double *loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nblk(npix, 256)), dim3(256), 0, s, Y, T, M, Co, npix, inv_count, dY, loss);
  return hipGetLastError();
}
hipError_t unet_adam(float *P, const float *G, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                     float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, s, P, G, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}
hipError_t unet_pack_fwd(const float *W, int taps, int Cin, int Cx, int Co, int Kp, uint16_t *Wt, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_fwd, dim3(nblk((size_t)Co * Kp, 256)), dim3(256), 0, s, W, taps, Cin, Cx, Co, Kp, Wt);
  return hipGetLastError();
}
hipError_t unet_pack_bwd(const float *W, int taps, int Cin, int Co, int Kpb, uint16_t *Wb, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_bwd, dim3(nblk((size_t)Cin * Kpb, 256)), dim3(256), 0, s, W, taps, Cin, Co, Kpb, Wb);
  return hipGetLastError();
}
hipError_t unet_pack_head(const float *W, int n, uint16_t *Wh, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_head, dim3(nblk(n, 256)), dim3(256), 0, s, W, n, Wh);
  return hipGetLastError();
}
hipError_t unet_to_bf16(const float *X, int C, int Cx, size_t npix, uint16_t *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(npix, 256)), dim3(256), 0, s, X, C, Cx, npix, Y);
  return hipGetLastError();
}
hipError_t unet_pack_state(const double *u, const double *v, const double *p, const double *T, const uint8_t *m,
                           int npix, int Cx, double inv_u, double inv_p, uint16_t *X, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_state, dim3(nblk(npix, 256)), dim3(256), 0, s, u, v, p, T, m, npix, Cx, inv_u, inv_p, X);
  return hipGetLastError();
}
hipError_t unet_pack_target(const double *p, const double *p_pr, const double *T, const double *T_pr,
                            const uint8_t *m, int npix, double sp, double sT, float *Tg, uint8_t *M, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_target, dim3(nblk(npix, 256)), dim3(256), 0, s, p, p_pr, T, T_pr, m, npix, sp, sT, Tg, M);
  return hipGetLastError();
}
hipError_t unet_apply_pred(const float *Y, const double *p, const double *T, const uint8_t *m, int Nx, int npix,
                           const int box[4], double sp, double sT, double *pw, double *Tw, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_pred, dim3(nblk(npix, 256)), dim3(256), 0, s, Y, p, T, m, Nx, npix, box[0], box[1],
                     box[2], box[3], sp, sT, pw, Tw);
  return hipGetLastError();
}
hipError_t unet_pack_corr_in(const CorrFields &xn, const CorrFields &xe, const uint8_t *m, int npix, int Cx,
                            double inv_u, double inv_p, double inv_step, uint16_t *X, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_corr_in, dim3(nblk(npix, 256)), dim3(256), 0, s, xn, xe, m, npix, Cx, inv_u, inv_p,
                     inv_step, X);
  return hipGetLastError();
}
hipError_t unet_pack_corr_target(const CorrFields &x1, const CorrFields &xe, const uint8_t *m, int npix, double inv_u,
                                 double inv_p, double inv_corr, float *Tg, uint8_t *M, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_corr_target, dim3(nblk(npix, 256)), dim3(256), 0, s, x1, xe, m, npix, inv_u, inv_p,
                     inv_corr, Tg, M);
  return hipGetLastError();
}
hipError_t unet_apply_corr(const float *Y, const CorrFields &xe, const uint8_t *m, const CorrBoxes &bx, int npix,
                           double su, double sp, double sT, const CorrOut &out, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_corr, dim3(nblk(npix, 256)), dim3(256), 0, s, Y, xe, m, bx, npix, su, sp, sT, out);
  return hipGetLastError();
}
hipError_t unet_gather(const uint16_t *Xs, const float *Ts, const uint8_t *Ms, const int *slots, int n, int npix,
                       int Cx, int Co, uint16_t *X, float *Tg, uint8_t *M, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)n * npix, 256)), dim3(256), 0, s, Xs, Ts, Ms, slots, n, npix, Cx, Co,
                     X, Tg, M);
  return hipGetLastError();
}
