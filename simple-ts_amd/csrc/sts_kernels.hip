// FP64 kernels of the SIMPLE-TS time step for CDNA4 (gfx950).
//
// One thread per cell (or per face), 64-wide wavefronts along i, so every
// field access of a wavefront is one contiguous 512-byte row segment of the
// reference's i-fastest layout.  The update sets of the reference's rule
// vectors (SIMPLE_TS_MPI.cpp:8581-9061) are evaluated as predicates on the
// body mask; inside a set the reference's "bc" formula is used, which is
// bit-identical to its "nch" formula where all masks are 1.  Built with
// -ffp-contract=off so each a*b+c rounds twice, as the reference's SSE2 build.
#include "sts_device.h"

#define FLD(name) fl.f[STS_##name]

namespace {

constexpr int BX = 64, BY = 4;

// Workgroups are dispatched to the chiplets (XCDs, each with its own L2)
// round-robin by linear id.  Tile t of a launch of nt tiles is given to the
// workgroup ids that make every XCD work through one contiguous run of tiles
// (whole rows of tiles), so tiles sharing halo cache lines meet in one L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_tile(int bid, int nt) {
  const int per = nt / NXCD, rem = nt % NXCD, x = bid % NXCD;
  return x * per + min(x, rem) + bid / NXCD;
}

// Origin (in cells) of this workgroup's TX x TY tile of an Nx-wide grid
// launched as a 1-D grid of tiles.
template <int TX, int TY>
__device__ __forceinline__ void tile_origin(int Nx, int &i0, int &j0) {
  const int nbx = (Nx + TX - 1) / TX;
  const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  i0 = (t % nbx) * TX;
  j0 = (t / nbx) * TY;
}

__device__ __forceinline__ bool in_box(int i, int j, int x0, int x1, int y0, int y1) {
  return i >= x0 && i < x1 && j >= y0 && j < y1;
}

// a / b, correctly rounded, from y = RN(1/b) when RN(a y) is within one ulp
// of a/b -- Markstein's theorem: then r = a - b q is exact in an fma and
// RN(q + r y) = RN(a/b), barring underflow and overflow.  RN(a y) is that
// close for every a when |1 - b y| <= 2^-(p+1) (half the unit roundoff).
// r = 0 (an exact quotient, a = +-0) or NaN (a infinite or NaN) keeps q, so
// signed zeros and non-finite quotients are those of a / b.
template <typename R>
__device__ __forceinline__ R quot1(R a, R b, R y) {
  const R q = a * y;
  const R r = fma(-q, b, a);
  return (fabs(r) > R(0)) ? fma(r, y, q) : q;
}

// a / 3: RN(1/3) = (1/3)(1 - 2^-(p+1)) meets quot1's condition.
template <typename R>
__device__ __forceinline__ R third(R a) {
  return quot1(a, R(3.0), R(1.0) / R(3.0));
}

// Non-negative doubles order like their bit patterns; a NaN (0x7ff8...)
// compares above +inf and so survives the max and is reported.  FP32
// differences are widened to double (exactly) first.
template <typename R>
__device__ __forceinline__ unsigned long long absbits(R r) {
  return (unsigned long long)__double_as_longlong(fabs((double)r));
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}

// One step of wave_max_last: v and its DPP-moved copy (both 32-bit halves
// moved alike); rows outside ROW_MASK keep v.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_max_step(unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
  const unsigned long long w =
      ((unsigned long long)(unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false) << 32) |
      (unsigned long long)(unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  return (w > v) ? w : v;
}

// Max over a full wave of a 64-bit word, complete in lane 63: within quads
// (quad permutations), half rows and rows (mirrors), then row 1 and 3 take
// the maxima of rows 0 and 2 (broadcast of lane 15) and rows 2 and 3 that of
// rows 0-1 (broadcast of lane 31).  DPP moves are VALU operations, with no
// LDS round trip per step.  Every lane of the wave must be active.
__device__ __forceinline__ unsigned long long wave_max_last(unsigned long long v) {
  v = dpp_max_step<0xb1, 0xf>(v);   // quad_perm(1, 0, 3, 2)
  v = dpp_max_step<0x4e, 0xf>(v);   // quad_perm(2, 3, 0, 1)
  v = dpp_max_step<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_max_step<0x140, 0xf>(v);  // row_mirror
  v = dpp_max_step<0x142, 0xa>(v);  // row_bcast:15
  v = dpp_max_step<0x143, 0xc>(v);  // row_bcast:31
  return v;
}

// Destinations of block_max_atomic's NV maxima: value k goes to d[k] and,
// when set, also to e[k] (the first and the last sweep of a single-sweep
// loop 3 coincide); null destinations are skipped.
template <int NV>
struct MaxDst {
  unsigned long long *d[NV];
  unsigned long long *e[NV];
};

// Block-wide max of NV values over a block of NT threads (whole waves), one
// atomicMax per block and destination, into the block's slot copy of each
// destination.  Every thread leaves its values in LDS, one barrier, then the
// first wave alone folds the NT / 64 entries of each lane and reduces across
// its lanes; the other waves are done at the barrier.  Called once per
// kernel by every thread of the block (its LDS words are not reused).
template <int NV, int NT>
__device__ __forceinline__ void block_max_atomic(const unsigned long long (&v)[NV], const MaxDst<NV> &dst) {
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ unsigned long long sm[NV][NT];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
  for (int k = 0; k < NV; k++) sm[k][tid] = v[k];
  __syncthreads();
  if (tid >= 64) return;
  const unsigned slot = ((unsigned)blockIdx.x + (unsigned)blockIdx.y * gridDim.x) % (unsigned)NSLOT;
  const unsigned off = slot * (unsigned)SLOT_STRIDE;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    unsigned long long b = sm[k][tid];
#pragma unroll
    for (int q = 1; q < NT / 64; q++) {
      const unsigned long long w = sm[k][tid + 64 * q];
      b = (w > b) ? w : b;
    }
    b = wave_max_last(b);
    if (tid == 63 && b) {
      if (dst.d[k]) atomicMax(dst.d[k] + off, b);
      if (dst.e[k]) atomicMax(dst.e[k] + off, b);
    }
  }
}

// ---------------------------------------------------------------------------
// A10 + A11 per face: fluxes (S:10440-10454) and diffusion coefficients
// (S:10457-10618).
// ---------------------------------------------------------------------------
// Inputs of one face/cell (load), then the reference's expressions on them
// (one method per array); a thread issues the loads of all its cells before
// computing any.  FLUX: the inputs of Fx, Fy; DIFF: those of Gamma and the
// velocity diffusion coefficients (the thermal ones need neither set alone).
template <typename R>
struct FluxIn {
  bool in, m00, m10, m01, m11;  // fluid: (i-1,j-1), (i,j-1), (i-1,j), (i,j)
  unsigned ij;
  R hxi, hxw, hyj, hys;  // hx[i], hx[i-1], hy[j], hy[j-1] (clamped or wrapped as the reference's)
  R ru, u, rv, v;
  R s00, s10, s01, s11;      // sqrt_T at the four cells
  R twu10, twu11, twv01, twv11;  // Temper_of_fluid_on_the_wall_u[ij_1], [ij], _v[i_1j], [ij]
  template <bool FLUX = true, bool DIFF = true>
  __device__ __forceinline__ void load(const KGridT<R> &g, const KFieldsT<R> &fl, int i, int j) {
    in = i < g.Nx && j < g.Ny;
    if (!in) return;
    const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
    const R *sT = FLD(SQRT_T), *Twu = FLD(TW_U), *Twv = FLD(TW_V);
    ij = n.ij;
    const unsigned jm = (unsigned)j - 1u * (0 < j);
    hxi = g.hx[i];
    hxw = g.hx[n.i_1];
    hyj = g.hy[j];
    hys = g.hy[jm];
    if (FLUX) {
      ru = FLD(RHO_U)[ij];
      u = FLD(U)[ij];
      rv = FLD(RHO_V)[ij];
      v = FLD(V)[ij];
    }
    const unsigned b = g.nb[ij];
    m00 = DIFF ? (b & NB_SW) != 0 : false;
    m10 = (b & NB_S) != 0;
    m01 = (b & NB_W) != 0;
    m11 = (b & NB_C) != 0;
    s00 = DIFF ? sT[n.i_1j_1] : R(0);
    s10 = sT[n.ij_1];
    s01 = sT[n.i_1j];
    s11 = sT[ij];
    twu10 = DIFF ? Twu[n.ij_1] : R(0);
    twu11 = Twu[ij];
    twv01 = DIFF ? Twv[n.i_1j] : R(0);
    twv11 = Twv[ij];
  }
  __device__ __forceinline__ R fx() const { return hyj * ru * u; }
  __device__ __forceinline__ R fy() const { return hxi * rv * v; }
  // Gamma_yf, Gamma_xf; false where all four cells are body and the
  // reference leaves them untouched (S:10477-10543), with Gy = Gx = 0.
  __device__ __forceinline__ bool gamma(R &Gy, R &Gx) const {
    Gy = R(0.0);
    Gx = R(0.0);
    if (m00 && m10 && m01 && m11) {
      Gy = R(0.5) * (Li(s00, s01, hys, hyj) + Li(s10, s11, hys, hyj));
      Gx = R(0.5) * (Li(s00, s10, hxw, hxi) + Li(s01, s11, hxw, hxi));
      return true;
    }
    if (m00 || m10 || m01 || m11) {
      Gy = R(0.5) * (R(m00 && m01) * Li(s00, s01, hys, hyj) + R(m00 != m01) * sqrt(twv01)
                     + R(m10 && m11) * Li(s10, s11, hys, hyj) + R(m10 != m11) * sqrt(twv11)
                     + R((!(m00 && m01)) || (!(m10 && m11))) * Li(sqrt(twu10), sqrt(twu11), hys, hyj));
      Gx = R(0.5) * (R(m00 && m10) * Li(s00, s10, hxw, hxi) + R(m00 != m10) * sqrt(twu10)
                     + R(m01 && m11) * Li(s01, s11, hxw, hxi) + R(m01 != m11) * sqrt(twu11)
                     + R((!(m00 && m10)) || (!(m01 && m11))) * Li(sqrt(twv01), sqrt(twv11), hxw, hxi));
      return true;
    }
    return false;
  }
  __device__ __forceinline__ R d_ux(const KGridT<R> &g) const { return g.cT * s01 * (hyj / hxw); }
  __device__ __forceinline__ R d_uy(const KGridT<R> &g, R Gy) const {
    return ((m11 && m01) || (m10 && m00)) ? g.cT * Gy * ((hxi + hxw) / (R(m11 * m01) * hyj + R(m10 * m00) * hys))
                                          : R(0.0);
  }
  __device__ __forceinline__ R d_vx(const KGridT<R> &g, R Gx) const {
    return ((m11 && m10) || (m01 && m00)) ? g.cT * Gx * ((hyj + hys) / (R(m11 * m10) * hxi + R(m01 * m00) * hxw))
                                          : R(0.0);
  }
  __device__ __forceinline__ R d_vy(const KGridT<R> &g) const { return g.cT * s10 * (hxi / hys); }
  __device__ __forceinline__ R dtx(const KGridT<R> &g) const {
    return (m11 || m01) ? g.cT1 * (hyj / (R(0.5) * (R(m11) * hxi + R(m01) * hxw)))
                              * (R((!m11) || (!m01)) * sqrt(twu11) + R(m11 * m01) * Hi(s01, s11, hxw, hxi))
                        : R(0.0);
  }
  __device__ __forceinline__ R dty(const KGridT<R> &g) const {
    return (m11 || m10) ? g.cT1 * (hxi / (R(0.5) * (R(m11) * hyj + R(m10) * hys)))
                              * (R((!m11) || (!m10)) * sqrt(twv11) + R(m11 * m10) * Hi(s10, s11, hys, hyj))
                        : R(0.0);
  }
  // Gamma and the velocity diffusion coefficients (needs DIFF)
  __device__ __forceinline__ void apply_diff(const KGridT<R> &g, const KFieldsT<R> &fl) const {
    if (!in) return;
    R Gy, Gx;
    if (gamma(Gy, Gx)) {
      FLD(GAMMA_YF)[ij] = Gy;
      FLD(GAMMA_XF)[ij] = Gx;
    }
    FLD(D_UX)[ij] = d_ux(g);
    FLD(D_UY)[ij] = d_uy(g, Gy);
    FLD(D_VX)[ij] = d_vx(g, Gx);
    FLD(D_VY)[ij] = d_vy(g);
  }
  // fluxes and thermal diffusion coefficients (needs FLUX)
  __device__ __forceinline__ void apply_flux(const KGridT<R> &g, const KFieldsT<R> &fl) const {
    if (!in) return;
    FLD(FX)[ij] = fx();
    FLD(FY)[ij] = fy();
    FLD(DTX)[ij] = dtx(g);
    FLD(DTY)[ij] = dty(g);
  }
  __device__ __forceinline__ void apply(const KGridT<R> &g, const KFieldsT<R> &fl) const {
    apply_flux(g, fl);
    apply_diff(g, fl);
  }
};

// ---------------------------------------------------------------------------
// A2 + A3 + A4: coefficient assembly (S:3269-3756)
// ---------------------------------------------------------------------------
// The diffusion coefficients of the faces an assembled cell reads: Gamma_yf
// and D_uy at ij and ij1, Gamma_xf and D_vx at ij and i1j, D_ux at ij and
// i1j, D_vy at ij and ij1.
template <typename R>
struct CellDiff {
  R Gy0, Gy1, Gx0, Gx1, Dux0, Dux1, Duy0, Duy1, Dvx0, Dvx1, Dvy0, Dvy1;
};

// The values of the fields a cell's assembly reads around it -- Fx, Fy, u,
// v, sqrt_T at offsets (di, dj) in {-1, 0, 1}^2, the clamped (x: wrapped when
// periodic) neighbours of Nbr; rho and rho_pr at the cell and its west and
// south neighbours; the mesh steps hx of the cell's column and its clamped
// (wrapped) neighbours, hy of its row and the clamped row below -- either
// from global memory through Nbr's indices or from a workgroup's LDS tile
// of them.  The offsets are compile-time constants wherever the accessors
// are used.
template <typename R>
struct GlobalStencil {
  const R *Fx_, *Fy_, *u_, *v_, *sT_, *rho_, *rho_pr_, *hx_, *hy_;
  const Nbr &n;
  __device__ __forceinline__ GlobalStencil(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &nb)
      : Fx_(FLD(FX)), Fy_(FLD(FY)), u_(FLD(U)), v_(FLD(V)), sT_(FLD(SQRT_T)), rho_(FLD(RHO)), rho_pr_(FLD(RHO_PR)),
        hx_(g.hx), hy_(g.hy), n(nb) {}
  __device__ __forceinline__ unsigned at(int di, int dj) const {
    if (dj == 0) return di < 0 ? n.i_1j : (di > 0 ? n.i1j : n.ij);
    if (dj < 0) return di < 0 ? n.i_1j_1 : (di > 0 ? n.i1j_1 : n.ij_1);
    return di < 0 ? n.i_1j1 : (di > 0 ? n.i1j1 : n.ij1);
  }
  __device__ __forceinline__ R Fx(int di, int dj) const { return Fx_[at(di, dj)]; }
  __device__ __forceinline__ R Fy(int di, int dj) const { return Fy_[at(di, dj)]; }
  __device__ __forceinline__ R u(int di, int dj) const { return u_[at(di, dj)]; }
  __device__ __forceinline__ R v(int di, int dj) const { return v_[at(di, dj)]; }
  __device__ __forceinline__ R sT(int di, int dj) const { return sT_[at(di, dj)]; }
  __device__ __forceinline__ R rho(int di, int dj) const { return rho_[at(di, dj)]; }
  __device__ __forceinline__ R rho_pr(int di, int dj) const { return rho_pr_[at(di, dj)]; }
  __device__ __forceinline__ R hx(int di) const { return hx_[di < 0 ? n.i_1 : (di > 0 ? n.i1 : n.i)]; }
  __device__ __forceinline__ R hy(int dj) const { return hy_[dj < 0 ? n.j - (0u < n.j) : n.j]; }  // dj = -1, 0
  __device__ __forceinline__ R div_ht(R a, const KGridT<R> &g) const { return a / g.ht; }
};

// LDS tile [TY + 2][TX + 2] of each field: entry [y][x] holds the field at
// cell (i0 + x - 1, j0 + y - 1) with the coordinates clamped (x wrapped when
// periodic) as Nbr clamps a neighbour, so that the neighbour (di, dj) of the
// tile's cell (tx, ty) is entry [ty + 1 + dj][tx + 1 + di].  HT1: the grid's
// 1/ht meets quot1's condition (rht_one), quotients by ht are formed from it.
template <typename R, int TX, int TY, bool HT1>
struct TileStencil {
  const R (*Fx_)[TX + 2], (*Fy_)[TX + 2], (*u_)[TX + 2], (*v_)[TX + 2], (*sT_)[TX + 2], (*rho_)[TX + 2],
      (*rho_pr_)[TX + 2];
  const R *hx_, *hy_;  // [TX + 2], [TY + 2]: the tile's columns and rows with their one-cell ring
  int tx, ty;
  __device__ __forceinline__ R Fx(int di, int dj) const { return Fx_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R Fy(int di, int dj) const { return Fy_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R u(int di, int dj) const { return u_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R v(int di, int dj) const { return v_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R sT(int di, int dj) const { return sT_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R rho(int di, int dj) const { return rho_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R rho_pr(int di, int dj) const { return rho_pr_[ty + 1 + dj][tx + 1 + di]; }
  __device__ __forceinline__ R hx(int di) const { return hx_[tx + 1 + di]; }
  __device__ __forceinline__ R hy(int dj) const { return hy_[ty + 1 + dj]; }
  __device__ __forceinline__ R div_ht(R a, const KGridT<R> &g) const {
    if constexpr (HT1) return quot1(a, g.ht, g.rht);
    else return a / g.ht;
  }
};

template <typename R, typename St>
__device__ __forceinline__ void cell_assemble_u(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &n,
                                                const CellDiff<R> &df, unsigned nb, const St &S) {
  const R *u_pr = FLD(U_PR), *rho_u = FLD(RHO_U);
  const unsigned ij = n.ij;
  const R cT = g.cT;
  R F_xu1 = R(0.5) * (S.Fx(-1, 0) + S.Fx(0, 0));
  R F_xu2 = R(0.5) * (S.Fx(0, 0) + S.Fx(1, 0));
  R Scu = cT * (R((nb & NB_E) != 0) * S.u(1, 0) * S.sT(0, 0) * S.hy(0) / (S.hx(0) * R(3.0))
                     + R((nb & NB_WW) != 0) * S.u(-1, 0) * S.sT(-1, 0) * S.hy(0) / (S.hx(-1) * R(3.0))
                     + df.Gy1 * (S.v(0, 1) * R((nb & NB_N) != 0) - S.v(-1, 1) * R((nb & NB_NW) != 0))
                     - df.Gy0 * (S.v(0, 0) * R((nb & NB_S) != 0) - S.v(-1, 0) * R((nb & NB_SW) != 0))
                     - third(S.sT(0, 0) * (S.v(0, 1) * R((nb & NB_N) != 0) - S.v(0, 0) * R((nb & NB_S) != 0)) * R(2.0))
                     + third(S.sT(-1, 0) * (S.v(-1, 1) * R((nb & NB_NW) != 0) - S.v(-1, 0) * R((nb & NB_SW) != 0)) * R(2.0)));
  R Spu = third(-(df.Dux0 + df.Dux1));
  R au1 = maxi0(F_xu1) + df.Dux0;
  R au2 = maxi0(-F_xu2) + df.Dux1;
  R au3 = R(0.5) * (maxi0(S.Fy(-1, 0)) + maxi0(S.Fy(0, 0))) + df.Duy0;
  R au4 = R(0.5) * (maxi0(-S.Fy(-1, 1)) + maxi0(-S.Fy(0, 1))) + df.Duy1;
  R bu3 = S.div_ht((S.rho_pr(-1, 0) * S.hx(-1) + S.rho_pr(0, 0) * S.hx(0)) * u_pr[ij] * S.hy(0) * R(0.5), g) + Scu;
  R au0 = au1 + au2 + au3 + au4 + R(0.5) * (-S.Fy(-1, 0) - S.Fy(0, 0) + S.Fy(-1, 1) + S.Fy(0, 1)) + F_xu2 - F_xu1 - Spu
               + S.div_ht((S.rho(-1, 0) * S.hx(-1) + S.rho(0, 0) * S.hx(0)) * S.hy(0) * R(0.5), g);
  R ups = (au1 * R((nb & NB_E) != 0) * S.u(-1, 0) + au2 * R((nb & NB_WW) != 0) * S.u(1, 0) + au3 * S.u(0, -1) + au4 * S.u(0, 1) + bu3) / au0;
  R dpc = g.cPch * S.hy(0) / au0;
  R h = S.hy(0) * rho_u[ij];
  FLD(U_PSEUDO)[ij] = ups;
  FLD(D_P_C_12)[ij] = dpc;
  FLD(APX)[ij] = dpc * h;
  FLD(BPX)[ij] = R((nb & NB_W) != 0 && (nb & NB_C) != 0) * h * ups;
}

template <typename R, typename St>
__device__ __forceinline__ void cell_assemble_v(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &n,
                                                const CellDiff<R> &df, unsigned nb, const St &S) {
  const R *v_pr = FLD(V_PR), *rho_v = FLD(RHO_V);
  const unsigned ij = n.ij;
  const R cT = g.cT;
  // S:3399 the mask at ij_1 - Nx*(0 <= (int)(j-2)) (NB_SS); S:3399, S:3404 hy[j - 1*(0 <= (int)(j-1))],
  // the clamped row below; S:3419, S:3426 read hy[j - 1] unclamped, the same row inside the v box (j >= 1)
  const R hyjm = S.hy(-1);
  R F_yv3 = R(0.5) * (S.Fy(0, -1) + S.Fy(0, 0));
  R F_yv4 = R(0.5) * (S.Fy(0, 0) + S.Fy(0, 1));
  R Scv = cT * (R((nb & NB_N) != 0) * S.v(0, 1) * S.sT(0, 0) * S.hx(0) / (S.hy(0) * R(3.0))
                     + R((nb & NB_SS) != 0) * S.v(0, -1) * S.sT(0, -1) * S.hx(0) / (hyjm * R(3.0))
                     + df.Gx1 * (S.u(1, 0) * R((nb & NB_E) != 0) - S.u(1, -1) * R((nb & NB_SE) != 0))
                     - df.Gx0 * (S.u(0, 0) * R((nb & NB_W) != 0) - S.u(0, -1) * R((nb & NB_SW) != 0))
                     - third(S.sT(0, 0) * (S.u(1, 0) * R((nb & NB_E) != 0) - S.u(0, 0) * R((nb & NB_W) != 0)) * R(2.0))
                     + third(S.sT(0, -1) * (S.u(1, -1) * R((nb & NB_SE) != 0) - S.u(0, -1) * R((nb & NB_SW) != 0)) * R(2.0)))
               - g.c_Fr_v * (S.rho(0, -1) * S.hx(0) * R(0.5) * hyjm + S.rho(0, 0) * S.hx(0) * R(0.5) * S.hy(0));
  R Spv = third(-(df.Dvy0 + df.Dvy1));
  R av1 = R(0.5) * (maxi0(S.Fx(0, 0)) + maxi0(S.Fx(0, -1))) + df.Dvx0;
  R av2 = R(0.5) * (maxi0(-S.Fx(1, 0)) + maxi0(-S.Fx(1, -1))) + df.Dvx1;
  R av3 = maxi0(F_yv3) + df.Dvy0;
  R av4 = maxi0(-F_yv4) + df.Dvy1;
  R bv3 = S.div_ht((S.rho_pr(0, -1) * S.hy(-1) + S.rho_pr(0, 0) * S.hy(0)) * v_pr[ij] * S.hx(0) * R(0.5), g) + Scv;
  R av0 = av1 + av2 + av3 + av4 + R(0.5) * (S.Fx(1, 0) - S.Fx(0, 0) + S.Fx(1, -1) - S.Fx(0, -1)) + F_yv4 - F_yv3 - Spv
               + S.div_ht((S.rho(0, -1) * S.hy(-1) + S.rho(0, 0) * S.hy(0)) * S.hx(0) * R(0.5), g);
  R vps = (av1 * S.v(-1, 0) + av2 * S.v(1, 0) + av3 * R((nb & NB_SS) != 0) * S.v(0, -1) + av4 * R((nb & NB_N) != 0) * S.v(0, 1) + bv3) / av0;
  R dpc = g.cPch * S.hx(0) / av0;
  R h = S.hx(0) * rho_v[ij];
  FLD(V_PSEUDO)[ij] = vps;
  FLD(D_P_C_34)[ij] = dpc;
  FLD(APY)[ij] = dpc * h;
  FLD(BPY)[ij] = R((nb & NB_S) != 0 && (nb & NB_C) != 0) * h * vps;
}

template <typename R, typename St>
__device__ __forceinline__ void cell_energy_source(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &n,
                                                   unsigned nb, const St &S) {
  const unsigned Nx = (unsigned)g.Nx;
  const R *p_pr = FLD(P_PR);
  const unsigned j = n.j, ij = n.ij;
  // S:3729-3740: the upper neighbour row is guarded by j < Nx-1 (from global
  // memory: that row may lie past the grid's last one); the lower is clamped
  const unsigned jp = j + 1 * (j < (Nx - 1));
  const R hyp = g.hy[jp];
  R dudx_T = (R(((nb & NB_E) != 0)) * S.u(1, 0) - R(((nb & NB_W) != 0)) * S.u(0, 0)) / S.hx(0);
  R dvdy_T = (R(((nb & NB_N) != 0)) * S.v(0, 1) - R(((nb & NB_S) != 0)) * S.v(0, 0)) / S.hy(0);
  R s0 = ((S.u(0, 0) - S.u(0, -1)) / (R(((nb & NB_SW) != 0) && ((nb & NB_S) != 0)) * S.hy(-1) + S.hy(0)))
              + ((S.v(0, 0) - S.v(-1, 0)) / (R(((nb & NB_SW) != 0) && ((nb & NB_W) != 0)) * S.hx(-1) + S.hx(0)));
  R s1 = ((S.u(1, 0) - S.u(1, -1)) / (R(((nb & NB_S) != 0) && ((nb & NB_SE) != 0)) * S.hy(-1) + S.hy(0)))
              + ((S.v(1, 0) - S.v(0, 0)) / (S.hx(0) + R(((nb & NB_SE) != 0) && ((nb & NB_E) != 0)) * S.hx(1)));
  R s2 = ((S.u(0, 1) - S.u(0, 0)) / (S.hy(0) + R(((nb & NB_NW) != 0) && ((nb & NB_N) != 0)) * hyp))
              + ((S.v(0, 1) - S.v(-1, 1)) / (R(((nb & NB_W) != 0) && ((nb & NB_NW) != 0)) * S.hx(-1) + S.hx(0)));
  R s3 = ((S.u(1, 1) - S.u(1, 0)) / (S.hy(0) + R(((nb & NB_N) != 0) && ((nb & NB_NE) != 0)) * hyp))
              + ((S.v(1, 1) - S.v(0, 1)) / (S.hx(0) + R(((nb & NB_E) != 0) && ((nb & NB_NE) != 0)) * S.hx(1)));
  R dsum = dudx_T + dvdy_T;
  FLD(SCT1)[ij] = g.ht * S.sT(0, 0) * g.cT2 * S.hx(0) * S.hy(0)
                      * (R(2.0) * (dudx_T * dudx_T + dvdy_T * dvdy_T) + s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3
                         - third(dsum * dsum * R(2.0)))
                  + p_pr[ij] * S.hx(0) * S.hy(0);
}

// The assembly of cell (i, j) with Gamma and D read from their arrays (stage
// calls, batched steps; the loops use k_assemble_inl).
template <typename R>
__device__ __forceinline__ void assemble_cell(const KGridT<R> &g, const KFieldsT<R> &fl, int i, int j) {
  if (i >= g.Nx || j >= g.Ny) return;
  const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
  const unsigned nb = g.nb[n.ij];
  const bool c = (nb & NB_C) != 0;
  const bool du = in_box(i, j, g.Nx_u0, g.Nx_u1, g.Ny_u0, g.Ny_u1) && (nb & NB_W) && c;
  const bool dv = in_box(i, j, g.Nx_v0, g.Nx_v1, g.Ny_v0, g.Ny_v1) && (nb & NB_S) && c;
  if (du || dv) {
    const R *Gyf = FLD(GAMMA_YF), *Gxf = FLD(GAMMA_XF), *Dux = FLD(D_UX), *Duy = FLD(D_UY), *Dvx = FLD(D_VX),
            *Dvy = FLD(D_VY);
    const CellDiff<R> df{Gyf[n.ij], Gyf[n.ij1], Gxf[n.ij], Gxf[n.i1j], Dux[n.ij], Dux[n.i1j],
                         Duy[n.ij], Duy[n.ij1], Dvx[n.ij], Dvx[n.i1j], Dvy[n.ij], Dvy[n.ij1]};
    const GlobalStencil<R> S(g, fl, n);
    if (du) cell_assemble_u(g, fl, n, df, nb, S);
    if (dv) cell_assemble_v(g, fl, n, df, nb, S);
  }
  if (in_box(i, j, g.Nx_o0, g.Nx_o1, g.Ny_o0, g.Ny_o1) && c)
    cell_energy_source(g, fl, n, nb, GlobalStencil<R>(g, fl, n));
}

template <typename R>
__global__ void __launch_bounds__(BX *BY) k_assemble(KGridT<R> g, KFieldsT<R> fl, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<BX, BY>(g.Nx, i0, j0);
  assemble_cell<R>(g, fl, i0 + (int)threadIdx.x, j0 + (int)threadIdx.y);
}

// Gamma and the velocity diffusion coefficients of one face as the flux and
// diffusion pass writes them (the assembly's faces always have a fluid cell,
// so the untouched case does not arise there).
template <typename R>
struct FaceDiff {
  R Gy, Gx, Dux, Duy, Dvx, Dvy;
};

// Gamma and the velocity diffusion coefficients of the face of cell (fi, fj)
// as face_diff forms them, with sqrt_T and the mesh steps taken from a
// workgroup's LDS tile: (cx, cy) is the face cell's position in the tile
// (entry [cy + 1][cx + 1] of sT holds its sqrt_T, [cy][cx] that of its
// clamped or wrapped south-west neighbour; Shx[cx + 1], Shx[cx] its hx and
// the west one, Shy likewise in y), b its neighbour fluid bits.  The wall
// temperatures are read only where some but not all of the four cells are
// fluid, the one case gamma() uses them.
template <typename R, int TX>
__device__ __forceinline__ FaceDiff<R> face_diff_tile(const KGridT<R> &g, const KFieldsT<R> &fl, unsigned fi,
                                                      unsigned fj, unsigned b, const R (*sT)[TX + 2], const R *Shx,
                                                      const R *Shy, int cx, int cy) {
  FluxIn<R> f;
  f.in = true;
  f.hxi = Shx[cx + 1];
  f.hxw = Shx[cx];
  f.hyj = Shy[cy + 1];
  f.hys = Shy[cy];
  f.m00 = (b & NB_SW) != 0;
  f.m10 = (b & NB_S) != 0;
  f.m01 = (b & NB_W) != 0;
  f.m11 = (b & NB_C) != 0;
  f.s00 = sT[cy][cx];
  f.s10 = sT[cy][cx + 1];
  f.s01 = sT[cy + 1][cx];
  f.s11 = sT[cy + 1][cx + 1];
  f.twu10 = f.twu11 = f.twv01 = f.twv11 = R(0);
  if ((f.m00 || f.m10 || f.m01 || f.m11) && !(f.m00 && f.m10 && f.m01 && f.m11)) {
    const Nbr n(fi, fj, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
    f.twu10 = FLD(TW_U)[n.ij_1];
    f.twu11 = FLD(TW_U)[n.ij];
    f.twv01 = FLD(TW_V)[n.i_1j];
    f.twv11 = FLD(TW_V)[n.ij];
  }
  FaceDiff<R> d;
  f.gamma(d.Gy, d.Gx);
  d.Dux = f.d_ux(g);
  d.Duy = f.d_uy(g, d.Gy);
  d.Dvx = f.d_vx(g, d.Gx);
  d.Dvy = f.d_vy(g);
  return d;
}

// The assembly of the TX x TY tile at (i0, j0) with the diffusion
// coefficients formed in the workgroup: each face of the tile once, plus the
// faces its cells read beyond it (column TX: the i1j face of each row's last
// cell, row TY: the ij1 face of each column's last cell, as the clamped or
// wrapped neighbour indices give them); then every cell assembles from them.
// A cell reads the y-face values (Gamma_y, D_uy, D_vy) of its own face and the
// north one, kept in LDS, and the x-face values (Gamma_x, D_ux, D_vx) of its
// own face and the east one: when a tile row lies within one wave (64 % TX
// == 0) the east cell's come from the next lane and only the ring column's
// from LDS (a quarter less LDS per workgroup, so more workgroups per CU);
// otherwise they are all kept in LDS.  (tx, ty): the thread's cell in the
// tile; the workgroup has TX * TY >= TX + TY + 4 threads, the linear thread
// index tx + TX ty.
template <typename R, int TX, int TY, bool HT1 = false>
__device__ __forceinline__ void assemble_inl_tile(const KGridT<R> &g, const KFieldsT<R> &fl, int i0, int j0, int tx,
                                                  int ty) {
  static_assert(TX * TY >= TX + TY + 4, "the face ring and the mesh steps need TX + TY + 4 threads");
  constexpr bool XL = 64 % TX == 0;      // x-face values of the east cell by lane exchange
  constexpr int LXW = XL ? 1 : TX + 1;  // x-face arrays: the ring column only, or every face
  constexpr int LXR = XL ? 0 : TX;      // their ring column
  __shared__ R Lgy[TY + 1][TX], Luy[TY + 1][TX], Lvy[TY + 1][TX], Lgx[TY][LXW], Lux[TY][LXW], Lvx[TY][LXW];
  __shared__ R SFx[TY + 2][TX + 2], SFy[TY + 2][TX + 2], Su[TY + 2][TX + 2], Sv[TY + 2][TX + 2],
      SsT[TY + 2][TX + 2], Srho[TY + 2][TX + 2], Srho_pr[TY + 2][TX + 2], Shx[TX + 2], Shy[TY + 2];
  const int Nx = g.Nx, Ny = g.Ny, t = tx + ty * TX;
  // column x - 1 and row y - 1 of the tile, clamped (x wrapped when periodic)
  // as Nbr takes a neighbour; x - 1 >= -1, and x - 1 < Nx + TX
  auto col = [&](int x) {
    int gi = i0 + x - 1;
    if (g.px) {
      if (gi < 0) gi += Nx;
      while (gi >= Nx) gi -= Nx;
      return gi;
    }
    return min(max(gi, 0), Nx - 1);
  };
  auto row = [&](int y) { return min(max(j0 + y - 1, 0), Ny - 1); };
  // the tile's neighbourhood of the fields the cells and faces read
  // (TileStencil, face_diff_tile); the entries past the grid's last tile
  // column are never read, only kept in it
  for (int e = t; e < (TX + 2) * (TY + 2); e += TX * TY) {
    const int x = e % (TX + 2), y = e / (TX + 2);
    const unsigned k = (unsigned)col(x) + (unsigned)row(y) * (unsigned)Nx;
    SFx[y][x] = FLD(FX)[k];
    SFy[y][x] = FLD(FY)[k];
    Su[y][x] = FLD(U)[k];
    Sv[y][x] = FLD(V)[k];
    SsT[y][x] = FLD(SQRT_T)[k];
    Srho[y][x] = FLD(RHO)[k];
    Srho_pr[y][x] = FLD(RHO_PR)[k];
  }
  if (t < TX + 2) Shx[t] = g.hx[col(t)];
  else if (t < TX + TY + 4) Shy[t - (TX + 2)] = g.hy[row(t - (TX + 2))];
  const int i = i0 + tx, j = j0 + ty;
  const int il = min(i0 + TX, Nx) - 1, jl = min(j0 + TY, Ny) - 1;  // last cell of the tile in x, y
  const bool in = i < Nx && j < Ny;
  const unsigned ij = (unsigned)i + (unsigned)j * (unsigned)Nx, nb = in ? g.nb[ij] : 0u;
  __syncthreads();
  // the cell's own face
  R gx = R(0), ux = R(0), vx = R(0);
  if (in) {
    const FaceDiff<R> d = face_diff_tile<R, TX>(g, fl, (unsigned)i, (unsigned)j, nb, SsT, Shx, Shy, tx, ty);
    Lgy[ty][tx] = d.Gy;
    Luy[ty][tx] = d.Duy;
    Lvy[ty][tx] = d.Dvy;
    gx = d.Gx;
    ux = d.Dux;
    vx = d.Dvx;
    if constexpr (!XL) {
      Lgx[ty][tx] = gx;
      Lux[ty][tx] = ux;
      Lvx[ty][tx] = vx;
    }
  }
  if (t < TX) {  // north: the ij1 face of column t's last cell (jl itself on the grid's top row); its y-face values
    const int fj = jl + (jl < Ny - 1 ? 1 : 0);
    if (i0 + t < Nx) {
      const FaceDiff<R> d = face_diff_tile<R, TX>(g, fl, (unsigned)(i0 + t), (unsigned)fj,
                                                  g.nb[(unsigned)(i0 + t) + (unsigned)fj * (unsigned)Nx], SsT, Shx,
                                                  Shy, t, fj - j0);
      Lgy[TY][t] = d.Gy;
      Luy[TY][t] = d.Duy;
      Lvy[TY][t] = d.Dvy;
    }
  } else if (t < TX + TY) {  // east: the i1j face of row t - TX's last cell (Nbr's i1 of il); its x-face values
    const int r = t - TX;
    const int fi = g.px ? (il == Nx - 1 ? 0 : il + 1) : il + (il < Nx - 1 ? 1 : 0);
    if (j0 + r < Ny) {
      const FaceDiff<R> d = face_diff_tile<R, TX>(g, fl, (unsigned)fi, (unsigned)(j0 + r),
                                                  g.nb[(unsigned)fi + (unsigned)(j0 + r) * (unsigned)Nx], SsT, Shx,
                                                  Shy, il - i0 + (fi != il ? 1 : 0), r);
      Lgx[r][LXR] = d.Gx;
      Lux[r][LXR] = d.Dux;
      Lvx[r][LXR] = d.Dvx;
    }
  }
  // the next lane's own x-face values (lanes of one wave, every lane active)
  R egx = R(0), eux = R(0), evx = R(0);
  if constexpr (XL) {
    egx = __shfl_down(gx, 1);
    eux = __shfl_down(ux, 1);
    evx = __shfl_down(vx, 1);
  }
  __syncthreads();
  if (!in) return;
  const Nbr n((unsigned)i, (unsigned)j, (unsigned)Nx, (unsigned)Ny, g.px);
  const bool c = (nb & NB_C) != 0;
  const bool du = in_box(i, j, g.Nx_u0, g.Nx_u1, g.Ny_u0, g.Ny_u1) && (nb & NB_W) && c;
  const bool dv = in_box(i, j, g.Nx_v0, g.Nx_v1, g.Ny_v0, g.Ny_v1) && (nb & NB_S) && c;
  const TileStencil<R, TX, TY, HT1> S{SFx, SFy, Su, Sv, SsT, Srho, Srho_pr, Shx, Shy, tx, ty};
  if (du || dv) {
    const int ny = (j == jl) ? TY : ty + 1;
    const bool last = i == il;  // the east face is the ring column's
    const R gx1 = last ? Lgx[ty][LXR] : XL ? egx : Lgx[ty][tx + 1];
    const R ux1 = last ? Lux[ty][LXR] : XL ? eux : Lux[ty][tx + 1];
    const R vx1 = last ? Lvx[ty][LXR] : XL ? evx : Lvx[ty][tx + 1];
    const CellDiff<R> df{Lgy[ty][tx], Lgy[ny][tx], gx, gx1, ux, ux1,
                         Luy[ty][tx], Luy[ny][tx], vx, vx1, Lvy[ty][tx], Lvy[ny][tx]};
    if (du) cell_assemble_u(g, fl, n, df, nb, S);
    if (dv) cell_assemble_v(g, fl, n, df, nb, S);
  }
  if (in_box(i, j, g.Nx_o0, g.Nx_o1, g.Ny_o0, g.Ny_o1) && c) cell_energy_source(g, fl, n, nb, S);
}

// Tile of the loops' assembly: 64 x 8 cells at six waves per SIMD (78
// registers; 51.6 KB of LDS, three workgroups per CU).
constexpr int AX = 64, AY = 8;

// HT1: launched for grids with rht_one.
template <typename R, bool HT1>
__global__ void __launch_bounds__(AX *AY) __attribute__((amdgpu_waves_per_eu(6, 8))) k_assemble_inl(KGridT<R> g, KFieldsT<R> fl, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<AX, AY>(g.Nx, i0, j0);
  assemble_inl_tile<R, AX, AY, HT1>(g, fl, i0, j0, (int)threadIdx.x, (int)threadIdx.y);
}

// ---------------------------------------------------------------------------
// A5: fused T/p Jacobi sweep (S:3805-4027).
// ---------------------------------------------------------------------------
// Old p and T of a cell and of its four (clamped) neighbours.
template <typename R>
struct PTStencil {
  R Pc, Tc, Pw, Tw, Pe, Te, Ps, Ts, Pn, Tn;
};

// The reference's T update (S:3837-3857) then p update (S:3884-3895) of one
// cell of the update set, with the new T in the p denominator (S:3894).  A
// body neighbour contributes the fluid temperature on its wall (S:3838-3848:
// m*T_old + !m*T_wall with T_wall finite, i.e. a select).  The per-cell
// factors that do not depend on the old p, T are formed once (load) with the
// reference's operations and kept, so a cell swept twice reads its
// coefficient arrays once; update() then performs the rest of each
// expression in the reference's order.
template <typename R>
struct PTCoef {
  R aT1, aT2, aT3, aT4;  // T neighbour coefficients (S:3959-3962)
  R X1, X2, X3, X4;      // wall temperature of a body neighbour
  R div, sct, denT;      // velocity divergence term, ScT1, T denominator
  R a1, a2, a3, a4;      // p neighbour coefficients apx_W, apx_E, apy_S, apy_N
  R b1, b2, b3, b4;      // bpx_W, bpx_E, bpy_S, bpy_N
  R gxb, gxe;            // given-velocity flux terms (S:3887-3888)
  R q, Ap, hh;           // p_pr hx hy / T_pr, ht * sum(a), hx hy
  R c1, ht;
  unsigned fluid;  // bit q: neighbour q (W, E, S, N) is fluid
  bool has_gxb, has_gxe;

  __device__ __forceinline__ void load(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &n) {
    const unsigned ij = n.ij, i1j = n.i1j, ij1 = n.ij1;
    const R *Fx = FLD(FX), *Fy = FLD(FY), *DTx = FLD(DTX), *DTy = FLD(DTY), *u = FLD(U), *v = FLD(V);
    const R *apx = FLD(APX), *apy = FLD(APY), *bpx = FLD(BPX), *bpy = FLD(BPY);
    ht = g.ht;
    c1 = g.c1;
    const R hxi = g.hx[n.i], hyj = g.hy[n.j];
    aT1 = maxi0(Fx[ij]) + DTx[ij];
    aT2 = maxi0(-Fx[i1j]) + DTx[i1j];
    aT3 = maxi0(Fy[ij]) + DTy[ij];
    aT4 = maxi0(-Fy[ij1]) + DTy[ij1];
    const unsigned nb = g.nb[ij];
    fluid = ((nb & NB_W) ? 1u : 0u) | ((nb & NB_E) ? 2u : 0u) | ((nb & NB_S) ? 4u : 0u) | ((nb & NB_N) ? 8u : 0u);
    X1 = (fluid & 1u) ? R(0) : FLD(TW_U)[ij];
    X2 = (fluid & 2u) ? R(0) : FLD(TW_U)[i1j];
    X3 = (fluid & 4u) ? R(0) : FLD(TW_V)[ij];
    X4 = (fluid & 8u) ? R(0) : FLD(TW_V)[ij1];
    div = (u[i1j] - u[ij]) * hyj + (v[ij1] - v[ij]) * hxi;
    sct = FLD(SCT1)[ij];
    denT = ht * (aT1 + aT2 + aT3 + aT4 + Fx[i1j] - Fx[ij] + Fy[ij1] - Fy[ij]) + hxi * hyj * FLD(RHO)[ij];
    a1 = apx[ij];
    a2 = apx[i1j];
    a3 = apy[ij];
    a4 = apy[ij1];
    b1 = bpx[ij];
    b2 = bpx[i1j];
    b3 = bpy[ij];
    b4 = bpy[ij1];
    has_gxb = g.given_xb && (int)n.i == g.Nx_o0;
    has_gxe = g.given_xe && (int)n.i == g.Nx_o1 - 1;
    gxb = has_gxb ? Fx[ij] : R(0);
    gxe = has_gxe ? Fx[i1j] : R(0);
    q = FLD(P_PR)[ij] * hxi * hyj / (FLD(T_PR)[ij]);
    Ap = ht * (a1 + a2 + a3 + a4);
    hh = hxi * hyj;
  }

  __device__ __forceinline__ void update(const PTStencil<R> &x, R &pn, R &tn) const {
    const R t1 = (fluid & 1u) ? x.Tw : X1, t2 = (fluid & 2u) ? x.Te : X2;
    const R t3 = (fluid & 4u) ? x.Ts : X3, t4 = (fluid & 8u) ? x.Tn : X4;
    tn = (ht * (aT1 * t1 + aT2 * t2 + aT3 * t3 + aT4 * t4 + c1 * x.Pc * div) + sct) / denT;
    R s = a1 * x.Pw + a2 * x.Pe + a3 * x.Ps + a4 * x.Pn + b1 - b2 + b3 - b4;
    if (has_gxb) s = s + gxb;
    if (has_gxe) s = s - gxe;
    pn = (ht * s + q) / (Ap + hh / (tn));
  }
};

template <typename R>
__device__ __forceinline__ void pt_update(const KGridT<R> &g, const KFieldsT<R> &fl, const Nbr &n, const PTStencil<R> &x,
                                          R &pn, R &tn) {
  PTCoef<R> c;
  c.load(g, fl, n);
  c.update(x, pn, tn);
}

template <typename R>
__device__ __forceinline__ bool in_pt_set(const KGridT<R> &g, int i, int j, unsigned ij) {
  return in_box(i, j, g.Nx_o0, g.Nx_o1, g.Ny_o0, g.Ny_o1) && g.m[ij];
}

template <typename R>
__device__ __forceinline__ PTStencil<R> stencil_global(const R *Po, const R *To, const Nbr &n) {
  return PTStencil<R>{Po[n.ij], To[n.ij], Po[n.i_1j], To[n.i_1j], Po[n.i1j], To[n.i1j],
                   Po[n.ij_1], To[n.ij_1], Po[n.ij1], To[n.ij1]};
}

// The first-sweep value of one cell from (Po, To): the update set gets the
// new values, any other cell keeps its old ones.
template <typename R>
__device__ __forceinline__ void sweep1_cell(const KGridT<R> &g, const KFieldsT<R> &fl, const R *__restrict__ Po,
                                            const R *__restrict__ To, int i, int j, R &pn, R &tn) {
  const unsigned ij = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
  if (in_pt_set(g, i, j, ij)) {
    const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
    pt_update(g, fl, n, stencil_global(Po, To, n), pn, tn);
  } else {
    pn = Po[ij];
    tn = To[ij];
  }
}

// One sweep: reads (Po, To), writes (Pn, Tn) on every cell: the update set
// gets the new values, every other cell is copied through, which restates
// the loop-3 entry copy S:3763-3764 and the pointer swap S:3792-3795.
template <typename R>
__global__ void __launch_bounds__(BX *BY) k_sweep(KGridT<R> g, KFieldsT<R> fl, const R *__restrict__ Po,
                                                  const R *__restrict__ To, R *__restrict__ Pn,
                                                  R *__restrict__ Tn, unsigned long long *red_p,
                                                  unsigned long long *red_T, unsigned long long *red_p2,
                                                  unsigned long long *red_T2, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<BX, BY>(g.Nx, i0, j0);
  const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
  unsigned long long r[2] = {0ull, 0ull};
  if (i < g.Nx && j < g.Ny) {
    const unsigned ij = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
    if (in_pt_set(g, i, j, ij)) {
      const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
      const PTStencil<R> x = stencil_global(Po, To, n);
      R pn, tn;
      pt_update(g, fl, n, x, pn, tn);
      Tn[ij] = tn;
      Pn[ij] = pn;
      r[0] = absbits(pn - x.Pc);
      r[1] = absbits(tn - x.Tc);
    } else {
      Pn[ij] = Po[ij];
      Tn[ij] = To[ij];
    }
  }
  if (red_p || red_T || red_p2 || red_T2) block_max_atomic<2, BX * BY>(r, MaxDst<2>{{red_p, red_T}, {red_p2, red_T2}});
}

// Two sweeps in one pass (temporal blocking): each block first sweeps its
// tile plus a one-cell halo from (Po, To) into LDS, then sweeps the tile from
// LDS into (Pn, Tn).  Same arithmetic as two k_sweep launches; the
// intermediate p and T never go to HBM.  Residual maxima of both sweeps over
// the tile cells go to (r1p, r1T) and (r2p, r2T).
constexpr int S2X = 64, S2Y = 8;

// r: the thread's residuals of sweep 1 (p, T) and sweep 2 (p, T).
template <typename R>
__device__ __forceinline__ void sweep2_tile_body(const KGridT<R> &g, const KFieldsT<R> &fl, const R *__restrict__ Po,
                                                 const R *__restrict__ To, R *__restrict__ Pn, R *__restrict__ Tn,
                                                 int i0, int j0, int tx, int ty, int SX, int SY, R *sP, R *sT,
                                                 unsigned long long (&r)[4]) {
  const int Nx = g.Nx, Ny = g.Ny;
  const int RX = SX + 2, RY = SY + 2;
  const int tid = tx + ty * SX;
  // Sweep 1 on the one-cell halo ring of the tile (its cells belong to the
  // neighbouring tiles, which count their residuals).  Periodic about OX, the
  // ring column left of the first grid column holds the last one, and the
  // column right of the last grid column holds the first one: if the last
  // column lies inside the tile, SY extra entries fill that LDS column.
  const int NR = 2 * RX + 2 * SY;
  const bool east_extra = g.px && i0 <= Nx - 1 && Nx - 1 < i0 + SX - 1;
  const int NRx = NR + (east_extra ? SY : 0);
  for (int q = tid; q < NRx; q += SX * SY) {
    int rx, ry;
    if (q >= NR) {
      rx = Nx - i0 + 1;
      ry = 1 + (q - NR);
    } else if (q < RX) {
      rx = q;
      ry = 0;
    } else if (q < 2 * RX) {
      rx = q - RX;
      ry = RY - 1;
    } else if (q < 2 * RX + SY) {
      rx = 0;
      ry = 1 + (q - 2 * RX);
    } else {
      rx = RX - 1;
      ry = 1 + (q - 2 * RX - SY);
    }
    int hi = i0 - 1 + rx;
    const int hj = j0 - 1 + ry;
    if (g.px) hi = (hi == -1) ? Nx - 1 : (hi == Nx) ? 0 : hi;
    if (hi < 0 || hj < 0 || hi >= Nx || hj >= Ny) continue;
    R pn, tn;
    sweep1_cell(g, fl, Po, To, hi, hj, pn, tn);
    sP[ry * RX + rx] = pn;
    sT[ry * RX + rx] = tn;
  }
  // Sweep 1 on the thread's own cell; its factors stay in registers.
  const int i = i0 + tx, j = j0 + ty;
  const bool in = i < Nx && j < Ny;
  const unsigned ij = in ? (unsigned)i + (unsigned)j * (unsigned)Nx : 0u;
  const bool upd = in && in_pt_set(g, i, j, ij);
  const Nbr n((unsigned)(in ? i : 0), (unsigned)(in ? j : 0), (unsigned)Nx, (unsigned)Ny, g.px);
  const int cx = tx + 1, cy = ty + 1;
  unsigned long long r1[2] = {0ull, 0ull};
  PTCoef<R> c;
  if (upd) {
    c.load(g, fl, n);
    const PTStencil<R> x = stencil_global(Po, To, n);
    R pn, tn;
    c.update(x, pn, tn);
    r1[0] = absbits(pn - x.Pc);
    r1[1] = absbits(tn - x.Tc);
    sP[cy * RX + cx] = pn;
    sT[cy * RX + cx] = tn;
  } else if (in) {
    sP[cy * RX + cx] = Po[ij];
    sT[cy * RX + cx] = To[ij];
  }
  __syncthreads();
  // Sweep 2 on the own cell from LDS.
  unsigned long long r2[2] = {0ull, 0ull};
  if (upd) {
    // clamped neighbours map to the cell itself at the domain edge; periodic
    // about OX, the LDS column beside the seam holds the wrapped neighbour
    const int wx = cx - (g.px ? 1 : (int)(n.i - n.i_1)), ex = cx + (g.px ? 1 : (int)(n.i1 - n.i));
    const int sy = cy - (int)((n.ij - n.ij_1) / (unsigned)Nx), ny = cy + (int)((n.ij1 - n.ij) / (unsigned)Nx);
    const PTStencil<R> x{sP[cy * RX + cx], sT[cy * RX + cx], sP[cy * RX + wx], sT[cy * RX + wx],
                         sP[cy * RX + ex], sT[cy * RX + ex], sP[sy * RX + cx], sT[sy * RX + cx],
                         sP[ny * RX + cx], sT[ny * RX + cx]};
    R pn, tn;
    c.update(x, pn, tn);
    Tn[ij] = tn;
    Pn[ij] = pn;
    r2[0] = absbits(pn - x.Pc);
    r2[1] = absbits(tn - x.Tc);
  } else if (in) {
    Pn[ij] = sP[cy * RX + cx];
    Tn[ij] = sT[cy * RX + cx];
  }
  r[0] = r1[0];
  r[1] = r1[1];
  r[2] = r2[0];
  r[3] = r2[1];
}

// Tile (i0, j0) of SX x SY cells with its one-cell halo in LDS.
template <typename R, int SX, int SY>
__device__ __forceinline__ void sweep2_tile(const KGridT<R> &g, const KFieldsT<R> &fl, const R *__restrict__ Po,
                                            const R *__restrict__ To, R *__restrict__ Pn, R *__restrict__ Tn, int i0,
                                            int j0, int tx, int ty, unsigned long long *r1p, unsigned long long *r1T,
                                            unsigned long long *r2p, unsigned long long *r2T) {
  __shared__ R sP[(SY + 2) * (SX + 2)], sT[(SY + 2) * (SX + 2)];
  unsigned long long r[4];
  sweep2_tile_body(g, fl, Po, To, Pn, Tn, i0, j0, tx, ty, SX, SY, sP, sT, r);
  if (r1p || r1T || r2p || r2T) block_max_atomic<4, SX * SY>(r, MaxDst<4>{{r1p, r1T, r2p, r2T}, {}});
}

// Six waves per SIMD (80 registers, three workgroups per CU): the pass is
// bound by the latency of its two load phases, and the third resident
// workgroup hides more of it.  Uniform factors come from the grid constants
// (scalar registers), so nothing spills.
template <typename R>
__global__ void __launch_bounds__(S2X *S2Y) __attribute__((amdgpu_waves_per_eu(6, 8))) k_sweep2(KGridT<R> g, KFieldsT<R> fl, const R *__restrict__ Po,
                                                     const R *__restrict__ To, R *__restrict__ Pn,
                                                     R *__restrict__ Tn, unsigned long long *r1p,
                                                     unsigned long long *r1T, unsigned long long *r2p,
                                                     unsigned long long *r2T, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<S2X, S2Y>(g.Nx, i0, j0);
  sweep2_tile<R, S2X, S2Y>(g, fl, Po, To, Pn, Tn, i0, j0, (int)threadIdx.x,
                           (int)threadIdx.y, r1p, r1T, r2p, r2T);
}

// ---------------------------------------------------------------------------
// A6 + A7 + A9: velocity correction (S:4170-4305), sqrt_T and rho
// (S:4323-4330), face densities (S:10382-10437).  The face density of face
// ij needs rho of the west/south cell; it is recomputed here from p and T
// with the same operations, so no second pass is needed.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R rho_of(const uint8_t *m, const R *p, const R *T, unsigned c) {
  return m[c] ? p[c] / T[c] : R(0.0);
}

template <typename R>
__device__ __forceinline__ void correct_cell(const KGridT<R> &g, const KFieldsT<R> &fl, int i, int j, int do_correct,
                                             unsigned long long (&r)[2]) {
  if (i < g.Nx && j < g.Ny) {
    const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
    const uint8_t *m = g.m;
    const unsigned ij = n.ij, i_1j = n.i_1j, ij_1 = n.ij_1;
    const R *p = FLD(P), *T = FLD(T);
    R uf = FLD(U)[ij], vf = FLD(V)[ij];
    if (do_correct) {
      if (in_box(i, j, g.Nx_u0, g.Nx_u1, g.Ny_u0, g.Ny_u1) && m[i_1j] && m[ij]) {
        R un = FLD(U_PSEUDO)[ij] - (p[ij] - p[i_1j]) * FLD(D_P_C_12)[ij];
        r[0] = absbits(un - uf);
        uf = un;
        FLD(U)[ij] = un;
      }
      if (in_box(i, j, g.Nx_v0, g.Nx_v1, g.Ny_v0, g.Ny_v1) && m[ij_1] && m[ij]) {
        R vn = FLD(V_PSEUDO)[ij] - (p[ij] - p[ij - (unsigned)g.Nx]) * FLD(D_P_C_34)[ij];
        r[1] = absbits(vn - vf);
        vf = vn;
        FLD(V)[ij] = vn;
      }
    }
    const R rij = rho_of(m, p, T, ij);
    FLD(SQRT_T)[ij] = sqrt(T[ij]);
    FLD(RHO)[ij] = rij;
    R ru, rv;
    if (m[i_1j] && m[ij]) ru = upwind_rho(rho_of(m, p, T, i_1j), rij, uf);
    else if (m[i_1j] != m[ij]) ru = m[i_1j] ? rho_of(m, p, T, i_1j) : rij;
    else ru = R(0.0);
    if (m[ij_1] && m[ij]) rv = upwind_rho(rho_of(m, p, T, ij_1), rij, vf);
    else if (m[ij_1] != m[ij]) rv = m[ij_1] ? rho_of(m, p, T, ij_1) : rij;
    else rv = R(0.0);
    FLD(RHO_U)[ij] = ru;
    FLD(RHO_V)[ij] = rv;
  }
}

// correct_cell in two phases, so that a thread can issue the loads of
// several cells before it computes any of them.
template <typename R>
struct CorrectIn {
  unsigned ij, i_1j, ij_1;
  bool in, cu, cv;                    // cell in grid; u, v corrected
  bool mc, mw, ms;                    // fluid: cell, west, south
  R uf, vf, ups, vps, d12, d34, pc, pw, ps, Tc, Tw, Ts;
  __device__ __forceinline__ void load(const KGridT<R> &g, const KFieldsT<R> &fl, int i, int j, int do_correct) {
    in = i < g.Nx && j < g.Ny;
    if (!in) return;
    const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
    ij = n.ij;
    i_1j = n.i_1j;
    ij_1 = n.ij_1;
    const unsigned nb = g.nb[ij];
    mc = (nb & NB_C) != 0;
    mw = (nb & NB_W) != 0;
    ms = (nb & NB_S) != 0;
    cu = do_correct && in_box(i, j, g.Nx_u0, g.Nx_u1, g.Ny_u0, g.Ny_u1) && mw && mc;
    cv = do_correct && in_box(i, j, g.Nx_v0, g.Nx_v1, g.Ny_v0, g.Ny_v1) && ms && mc;
    const R *p = FLD(P), *T = FLD(T);
    uf = FLD(U)[ij];
    vf = FLD(V)[ij];
    ups = cu ? FLD(U_PSEUDO)[ij] : R(0);
    d12 = cu ? FLD(D_P_C_12)[ij] : R(0);
    vps = cv ? FLD(V_PSEUDO)[ij] : R(0);
    d34 = cv ? FLD(D_P_C_34)[ij] : R(0);
    pc = p[ij];
    pw = p[i_1j];
    ps = p[ij_1];
    Tc = T[ij];
    Tw = T[i_1j];
    Ts = T[ij_1];
  }
  // S:4170-4305 (u, v), S:4323-4330 (sqrt_T, rho), S:10382-10437 (face densities)
  // rr (optional): the face densities rho_u, rho_v written
  __device__ __forceinline__ void apply(const KFieldsT<R> &fl, unsigned long long (&r)[2], R *rr = nullptr) const {
    if (!in) return;
    R u = uf, v = vf;
    if (cu) {
      const R un = ups - (pc - pw) * d12;
      const unsigned long long e = absbits(un - uf);
      r[0] = e > r[0] ? e : r[0];
      u = un;
      FLD(U)[ij] = un;
    }
    if (cv) {  // p[ij - Nx] is p[ij_1] inside the v box (j >= 1)
      const R vn = vps - (pc - ps) * d34;
      const unsigned long long e = absbits(vn - vf);
      r[1] = e > r[1] ? e : r[1];
      v = vn;
      FLD(V)[ij] = vn;
    }
    const R rij = mc ? pc / Tc : R(0.0);
    FLD(SQRT_T)[ij] = sqrt(Tc);
    FLD(RHO)[ij] = rij;
    const R rw = mw ? pw / Tw : R(0.0), rs = ms ? ps / Ts : R(0.0);
    R ru, rv;
    if (mw && mc) ru = upwind_rho(rw, rij, u);
    else if (mw != mc) ru = mw ? rw : rij;
    else ru = R(0.0);
    if (ms && mc) rv = upwind_rho(rs, rij, v);
    else if (ms != mc) rv = ms ? rs : rij;
    else rv = R(0.0);
    FLD(RHO_U)[ij] = ru;
    FLD(RHO_V)[ij] = rv;
    if (rr) {
      rr[0] = ru;
      rr[1] = rv;
    }
  }
};

// CORRECT_ROWS cells per thread, rows BY apart, all loads issued first.
constexpr int CORRECT_ROWS = 2;

template <typename R>
__global__ void __launch_bounds__(BX *BY) k_correct(KGridT<R> g, KFieldsT<R> fl, unsigned long long *red_u,
                                                    unsigned long long *red_v, int do_correct, Gate gate) {
  if (gated_out(gate)) return;
  unsigned long long r[2] = {0ull, 0ull};
  int i0, j0;
  tile_origin<BX, BY * CORRECT_ROWS>(g.Nx, i0, j0);
  CorrectIn<R> c[CORRECT_ROWS];
#pragma unroll
  for (int q = 0; q < CORRECT_ROWS; q++)
    c[q].load(g, fl, i0 + (int)threadIdx.x, j0 + (int)threadIdx.y + q * BY, do_correct);
#pragma unroll
  for (int q = 0; q < CORRECT_ROWS; q++) c[q].apply(fl, r);
  if (do_correct) block_max_atomic<2, BX * BY>(r, MaxDst<2>{{red_u, red_v}, {}});
}

// A7 alone (S:4323-4330)
template <typename R>
__global__ void __launch_bounds__(BX *BY) k_properties(KGridT<R> g, KFieldsT<R> fl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned c = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
  const R T = FLD(T)[c];
  FLD(SQRT_T)[c] = sqrt(T);
  FLD(RHO)[c] = g.m[c] ? FLD(P)[c] / T : R(0.0);
}

// A9 alone (S:10382-10437), from the stored rho
template <typename R>
__global__ void __launch_bounds__(BX *BY) k_rho_mid(KGridT<R> g, KFieldsT<R> fl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
  const uint8_t *m = g.m;
  const R *rho = FLD(RHO);
  const unsigned ij = n.ij, i_1j = n.i_1j, ij_1 = n.ij_1;
  R ru, rv;
  if (m[i_1j] && m[ij]) ru = upwind_rho(rho[i_1j], rho[ij], FLD(U)[ij]);
  else if (m[i_1j] != m[ij]) ru = m[i_1j] ? rho[i_1j] : rho[ij];
  else ru = R(0.0);
  if (m[ij_1] && m[ij]) rv = upwind_rho(rho[ij_1], rho[ij], FLD(V)[ij]);
  else if (m[ij_1] != m[ij]) rv = m[ij_1] ? rho[ij_1] : rho[ij];
  else rv = R(0.0);
  FLD(RHO_U)[ij] = ru;
  FLD(RHO_V)[ij] = rv;
}

// ---------------------------------------------------------------------------
// A12 wall part: velocity slip (S:7098-7331) and the fluid temperature on the
// walls (S:7744-8045), per face.  The reference walks lists of cells in
// increasing node order and each cell writes up to two faces; per face the
// last writer of that walk is evaluated here.  Every value read is a face or
// cell that no writer of this pass modifies, so faces are independent.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ const DevBody *body_of(const KGridT<R> &g, uint32_t k) {
  return (k >= 1u && (int)k <= g.nbodies) ? &g.bodies[k - 1] : nullptr;
}

// Slip cell for u: body cell with fluid below or above, slip on (S:10169-10176).
template <typename R>
__device__ __forceinline__ const DevBody *slip_cell_u(const KGridT<R> &g, const Nbr &c) {
  const uint32_t *vol = g.vol;
  if (!((0u < vol[c.ij]) && ((0u == vol[c.ij_1]) || (0u == vol[c.ij1])))) return nullptr;
  const DevBody *b = body_of(g, vol[c.ij]);
  return (b && b->is_slip) ? b : nullptr;
}
template <typename R>
__device__ __forceinline__ const DevBody *slip_cell_v(const KGridT<R> &g, const Nbr &c) {
  const uint32_t *vol = g.vol;
  if (!((0u < vol[c.ij]) && ((0u == vol[c.i_1j]) || (0u == vol[c.i1j])))) return nullptr;
  const DevBody *b = body_of(g, vol[c.ij]);
  return (b && b->is_slip) ? b : nullptr;
}

// The two writes of slip cell c for u (S:7155-7206): which==0 -> u[c.ij],
// which==1 -> u[c.i1j].  Decided by the mask alone: returns false if the
// write does not happen, else the face it reads (src) and the hy row of the
// half cell (h).
template <typename R>
__device__ __forceinline__ bool slip_u_src(const KGridT<R> &g, const Nbr &c, int which, unsigned &src, unsigned &h) {
  const uint8_t *m = g.m;
  if (0u == g.vol[c.ij_1]) {
    h = c.j - 1 * (0 < c.j);
    if (which == 0) {
      if (!m[c.i_1j_1]) return false;
      src = c.ij_1;
    } else {
      if (!m[c.i1j_1]) return false;
      src = c.i1j_1;
    }
  } else {
    h = c.j + 1 * (c.j < (unsigned)(g.Ny - 1));
    if (which == 0) {
      if (!m[c.i_1j1]) return false;
      src = c.ij1;
    } else {
      if (!m[c.i1j1]) return false;
      src = c.i1j1;
    }
  }
  return true;
}

// S:7273-7324: which==0 -> v[c.ij], which==1 -> v[c.ij1]; h is an hx column.
template <typename R>
__device__ __forceinline__ bool slip_v_src(const KGridT<R> &g, const Nbr &c, int which, unsigned &src, unsigned &h) {
  const uint8_t *m = g.m;
  if (0u == g.vol[c.i_1j]) {
    h = c.i_1;
    if (which == 0) {
      if (!m[c.i_1j_1]) return false;
      src = c.i_1j;
    } else {
      if (!m[c.i_1j1]) return false;
      src = c.i_1j1;
    }
  } else {
    h = c.i1;
    if (which == 0) {
      if (!m[c.i1j_1]) return false;
      src = c.i1j;
    } else {
      if (!m[c.i1j1]) return false;
      src = c.i1j1;
    }
  }
  return true;
}

// Slip velocity (S:7160-7205): (vel a + vb) / (1 + a) with
// a = F Kn / (rho_face h/2) (local Kn) or F Kn / (h/2).
template <typename R>
__device__ __forceinline__ R slip_value(const KGridT<R> &g, const DevBody *b, R vb, R vel, R rho_face, R hh) {
  const R FK = R(b->F_slip) * g.Kn;
  const R a = g.kn_local ? FK / (rho_face * R(0.5) * hh) : FK / (R(0.5) * hh);
  return (vel * a + vb) / (R(1.0) + a);
}

// Wall temperature seen from a fluid cell through a face with body b (S:7762-7797).
template <typename R>
__device__ __forceinline__ R wall_T(const KGridT<R> &g, const DevBody *b, R T_fluid, R rho_fluid, R half_h) {
  if (b->is_jump) {
    R FK = g.kn_local ? (R(b->F_jump) * g.Kn) / rho_fluid : R(b->F_jump) * g.Kn;
    R a = FK / half_h;
    return (a * T_fluid + R(b->T_body)) / (R(1.0) + a);
  } else if (b->dTdn) {
    return T_fluid;
  }
  return R(b->T_body);
}

__device__ __forceinline__ void set_op(WallOp &o, unsigned kind, unsigned dst, unsigned src, unsigned h,
                                       unsigned body) {
  o.dst = dst;
  o.src = src;
  o.h = h;
  o.kb = kind | (body << 3);
}

// The writes of face k = (i, j) for the requested parts: per array the last
// writer of the reference's walk (increasing node order; periodic about OX
// the last column's cell follows the first column's).  op[q] for U, V, TW_U,
// TW_V; kind WOP_NONE where nothing is written.
template <typename R>
__device__ __forceinline__ void plan_wall_cell(const KGridT<R> &g, int i, int j, int parts, WallOp (&op)[4]) {
#pragma unroll
  for (int q = 0; q < 4; q++) op[q].kb = WOP_NONE;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned Nx = (unsigned)g.Nx, Ny = (unsigned)g.Ny;
  const unsigned k = (unsigned)i + (unsigned)j * Nx;
  const uint8_t *m = g.m;
  const bool px = g.px != 0;
  unsigned src = 0u, h = 0u;
  if (parts & STS_BC_SLIP_U) {
    const Nbr c((unsigned)i, (unsigned)j, Nx, Ny, px);
    const DevBody *b = slip_cell_u(g, c);
    const unsigned bi = b ? (unsigned)(b - g.bodies) : 0u;
    if (px && i == 0) {
      // periodic: u[ij] of the first column is written by the cell itself,
      // then by the last column's cell (node ij + Nx - 1, later in the walk)
      if (b && slip_u_src(g, c, 0, src, h)) set_op(op[0], WOP_U, k, src, h, bi);
      const Nbr cw(Nx - 1u, (unsigned)j, Nx, Ny, true);
      const DevBody *bw = slip_cell_u(g, cw);
      if (bw && slip_u_src(g, cw, 1, src, h)) set_op(op[0], WOP_U, k, src, h, (unsigned)(bw - g.bodies));
    } else {
      if (i >= 1) {  // cell (i-1, j) writes u[i1j]
        const Nbr cw((unsigned)i - 1u, (unsigned)j, Nx, Ny, px);
        const DevBody *bw = slip_cell_u(g, cw);
        if (bw && slip_u_src(g, cw, 1, src, h)) set_op(op[0], WOP_U, k, src, h, (unsigned)(bw - g.bodies));
      }
      if (b) {
        if (slip_u_src(g, c, 0, src, h)) set_op(op[0], WOP_U, k, src, h, bi);
        if (!px && (unsigned)i == Nx - 1u && slip_u_src(g, c, 1, src, h))  // clamped i1j == ij
          set_op(op[0], WOP_U, k, src, h, bi);
      }
    }
  }
  if (parts & STS_BC_SLIP_V) {
    if (j >= 1) {  // cell (i, j-1) writes v[ij1]
      const Nbr c((unsigned)i, (unsigned)j - 1u, Nx, Ny, px);
      const DevBody *b = slip_cell_v(g, c);
      if (b && slip_v_src(g, c, 1, src, h)) set_op(op[1], WOP_V, k, src, h, (unsigned)(b - g.bodies));
    }
    const Nbr c((unsigned)i, (unsigned)j, Nx, Ny, px);
    const DevBody *b = slip_cell_v(g, c);
    if (b) {
      const unsigned bi = (unsigned)(b - g.bodies);
      if (slip_v_src(g, c, 0, src, h)) set_op(op[1], WOP_V, k, src, h, bi);
      if ((unsigned)j == Ny - 1u && slip_v_src(g, c, 1, src, h))  // clamped ij1 == ij
        set_op(op[1], WOP_V, k, src, h, bi);
    }
  }
  if (parts & STS_BC_JUMP) {
    if (i >= 1 || px) {  // face between cell iw (west, across the seam if periodic) and cell i
      const unsigned iw = (i >= 1) ? (unsigned)i - 1u : Nx - 1u, w_ = iw + (unsigned)j * Nx;
      if (m[k] && !m[w_]) set_op(op[2], WOP_TWU, k, k, (unsigned)i, g.vol[w_] - 1u);
      else if (m[w_] && !m[k]) set_op(op[2], WOP_TWU, k, w_, iw, g.vol[k] - 1u);
    }
    if (j >= 1) {
      const unsigned s_ = k - Nx;
      if (m[k] && !m[s_]) set_op(op[3], WOP_TWV, k, k, (unsigned)j, g.vol[s_] - 1u);
      else if (m[s_] && !m[k]) set_op(op[3], WOP_TWV, k, s_, (unsigned)j - 1u, g.vol[k] - 1u);
    }
  }
}

// Value of one planned write; every value read is a face or cell that no
// write of the pass modifies, so the writes are independent.
template <typename R>
__device__ __forceinline__ void apply_wall_op(const KGridT<R> &g, const KFieldsT<R> &fl, const WallOp &o) {
  const unsigned kind = o.kb & 7u;
  if (kind == WOP_NONE) return;
  const DevBody *b = &g.bodies[o.kb >> 3];
  if (kind == WOP_U)
    FLD(U)[o.dst] = slip_value(g, b, R(b->u_body), FLD(U)[o.src], FLD(RHO_U)[o.src], g.hy[o.h]);
  else if (kind == WOP_V)
    FLD(V)[o.dst] = slip_value(g, b, R(b->v_body), FLD(V)[o.src], FLD(RHO_V)[o.src], g.hx[o.h]);
  else if (kind == WOP_TWU)
    FLD(TW_U)[o.dst] = wall_T(g, b, FLD(T)[o.src], FLD(RHO)[o.src], R(0.5) * g.hx[o.h]);
  else
    FLD(TW_V)[o.dst] = wall_T(g, b, FLD(T)[o.src], FLD(RHO)[o.src], R(0.5) * g.hy[o.h]);
}

template <typename R>
__device__ __forceinline__ void wall_cell(const KGridT<R> &g, const KFieldsT<R> &fl, int i, int j, int parts) {
  WallOp op[4];
  plan_wall_cell(g, i, j, parts, op);
#pragma unroll
  for (int q = 0; q < 4; q++) apply_wall_op(g, fl, op[q]);
}

template <typename R>
__global__ void __launch_bounds__(BX *BY) k_wall(KGridT<R> g, KFieldsT<R> fl, int parts, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<BX, BY>(g.Nx, i0, j0);
  wall_cell(g, fl, i0 + (int)threadIdx.x, j0 + (int)threadIdx.y, parts);
}

// The planned writes of a case (built once from its mask), one per thread.
template <typename R>
__global__ void __launch_bounds__(256) k_wall_ops(KGridT<R> g, KFieldsT<R> fl, const WallOp *__restrict__ ops, int n,
                                                  int parts, Gate gate) {
  if (gated_out(gate)) return;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= n) return;
  const WallOp o = ops[t];
  const unsigned kind = o.kb & 7u;
  const int need = kind == WOP_U ? STS_BC_SLIP_U : kind == WOP_V ? STS_BC_SLIP_V : STS_BC_JUMP;
  if (parts & need) apply_wall_op(g, fl, o);
}

// Planning pass: the four planned writes of each listed cell (out[4 t + q]).
__global__ void __launch_bounds__(256) k_wall_plan(KGrid g, const unsigned *cells, int n, WallOp *out) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= n) return;
  const unsigned k = cells[t];
  WallOp op[4];
  plan_wall_cell(g, (int)(k % (unsigned)g.Nx), (int)(k / (unsigned)g.Nx), STS_BC_SLIP_U | STS_BC_SLIP_V | STS_BC_JUMP,
                 op);
#pragma unroll
  for (int q = 0; q < 4; q++) out[4 * (size_t)t + q] = op[q];
}

// ---------------------------------------------------------------------------
// A12 edge part: Neumann copies and inflow/outflow mass conditions at the
// solve-box edges, S:6790-7096 (velocities), S:7650-7741 (temperature),
// S:7335-7428 (pressure).  O(Nx+Ny) work whose phases read each other's
// results, so one workgroup runs the phases in the reference's order with a
// barrier between them.
// ---------------------------------------------------------------------------
constexpr int EDGE_THREADS = 1024;

template <typename R>
__device__ __forceinline__ void edge_x(const KGridT<R> &g, R *a, int b0, int b1, bool lo, bool hi) {
  const int Nx = g.Nx, Ny = g.Ny;
  const uint8_t *m = g.m;
  if (lo) {
    for (int t = threadIdx.x; t < Ny * b0; t += EDGE_THREADS) {
      int j = t / b0, i = t - j * b0;
      int node = b0 + j * Nx;
      if (m[node]) a[i + j * Nx] = a[node];
    }
  }
  if (hi) {
    const int w = Nx - b1;
    for (int t = threadIdx.x; t < Ny * w; t += EDGE_THREADS) {
      int j = t / w, i = b1 + (t - j * w);
      int node = b1 - 1 + j * Nx;
      if (m[node]) a[i + j * Nx] = a[node];
    }
  }
}

template <typename R>
__device__ __forceinline__ void edge_y(const KGridT<R> &g, R *a, int b0, int b1, bool lo, bool hi) {
  const int Nx = g.Nx, Ny = g.Ny;
  const uint8_t *m = g.m;
  if (lo) {
    for (int t = threadIdx.x; t < Nx * b0; t += EDGE_THREADS) {
      int j = t / Nx, i = t - j * Nx;
      int node = i + b0 * Nx;
      if (m[node]) a[i + j * Nx] = a[node];
    }
  }
  if (hi) {
    const int h = Ny - b1;
    for (int t = threadIdx.x; t < Nx * h; t += EDGE_THREADS) {
      int jj = t / Nx, i = t - jj * Nx, j = b1 + jj;
      int node = i + (b1 - 1) * Nx;
      if (m[node]) a[i + j * Nx] = a[node];
    }
  }
}

// The loop-2 test of iteration gate.k (S:4433-4437, Iter = k + 1 after the
// increment), one wave: fold the slot copies of the six residual words,
// clear them, log the maxima, and end the loop by setting stop_at = k.  The
// residuals stay in registers; the stagnation window's older log entries are
// fetched alongside the slots.
__device__ __forceinline__ void loop_test_wave(const Gate &gate, const LoopTest &lt) {
  if (!lt.ctrl || threadIdx.x >= NSLOT) return;
  const bool stall = lt.stall_w > 0 && gate.k >= 2 * lt.stall_w;
  unsigned long long old[4] = {0ull, 0ull, 0ull, 0ull};
  if (stall && threadIdx.x == 0) {
    const unsigned long long *lb = lt.log_k - 6 * lt.stall_w;
#pragma unroll
    for (int q = 0; q < 4; q++) old[q] = lb[q];
  }
  unsigned long long *sl = lt.lred + threadIdx.x * SLOT_STRIDE;
  unsigned long long w[6];
#pragma unroll
  for (int q = 0; q < 6; q++) w[q] = sl[q];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    sl[q] = 0ull;
    w[q] = wave_max(w[q]);
  }
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int q = 0; q < 6; q++) lt.log_k[q] = w[q];
  double r[4];
  bool fin = true;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    r[q] = __longlong_as_double((long long)w[q]);
    fin = fin && isfinite(r[q]);
  }
  bool cont = (gate.k + 1 < lt.N_I) && (lt.tol_uv < r[0] || lt.tol_uv < r[1] || lt.tol_p < r[2] || lt.tol_T < r[3]);
  if (cont && stall) {
    // stagnation (FP32 phase): the largest residual relative to its
    // tolerance fell by less than stall_ratio over the last stall_w iterations
    const double tol[4] = {lt.tol_uv, lt.tol_uv, lt.tol_p, lt.tol_T};
    double now = 0.0, then = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      now = fmax(now, r[q] / tol[q]);
      then = fmax(then, __longlong_as_double((long long)old[q]) / tol[q]);
    }
    if (now > lt.stall_ratio * then) cont = false;
  }
  if (!fin) {
    lt.ctrl->nonfinite = 1;
    cont = false;
  }
  if (!cont) lt.ctrl->stop_at = gate.k;
}

constexpr int EDGE_GROUPS = 4;

template <typename R>
__device__ void edges_block(const KGridT<R> &g, const KFieldsT<R> &fl, int parts, double *probe, const Gate &gate,
                            const LoopTest &lt, int group) {
  // The velocity, temperature and pressure conditions touch disjoint arrays
  // (u, v read rho and the face densities, which none of them writes), and
  // the loop test reads only the residual words, so four workgroups run them
  // side by side: group 0 = velocities, 1 = temperature, 2 = pressure and the
  // probe, 3 = loop test (group < 0: everything in one workgroup).
  if (group == 3 || group < 0) loop_test_wave(gate, lt);
  if (group == 3) return;
  if (group == 1) parts &= STS_BC_TEMPERATURE;
  if (group == 2) parts &= STS_BC_PRESSURE;
  if (group == 0) parts &= STS_BC_VELOCITIES;
  // periodic about OX: the box-edge conditions are off (S:6795, S:7340, S:7655)
  if (g.px) parts &= ~(STS_BC_VELOCITIES | STS_BC_TEMPERATURE | STS_BC_PRESSURE);
  if (group >= 0 && group != 2) probe = nullptr;
  const int Nx = g.Nx, Ny = g.Ny;
  const uint8_t *m = g.m;
  const unsigned bc = g.bc;
  R *u = FLD(U), *v = FLD(V);
  const R *rho = FLD(RHO), *rho_pr = FLD(RHO_PR), *rho_u = FLD(RHO_U), *rho_v = FLD(RHO_V);
  const R *hx = g.hx, *hy = g.hy;
  if (parts & STS_BC_VELOCITIES) {
    edge_x(g, v, g.Nx_v0, g.Nx_v1, bc & BC_DVDX_XB, false);
    __syncthreads();
    edge_x(g, v, g.Nx_v0, g.Nx_v1, false, bc & BC_DVDX_XE);
    __syncthreads();
    edge_y(g, v, g.Ny_v0, g.Ny_v1, bc & BC_DVDY_YB, false);
    __syncthreads();
    edge_y(g, v, g.Ny_v0, g.Ny_v1, false, bc & BC_DVDY_YE);
    __syncthreads();
    edge_x(g, u, g.Nx_u0, g.Nx_u1, bc & BC_DUDX_XB, false);
    __syncthreads();
    if (bc & BC_DURHODX_XB)
      for (int j = threadIdx.x; j < Ny; j += EDGE_THREADS) {
        int node = g.Nx_u0 - 1 + j * Nx;
        if (m[node]) u[node] = (rho_u[node + 1] * u[node + 1]) / rho_u[node];
      }
    __syncthreads();
    if (bc & BC_DRHODT_XB) {
      const int i = g.Nx_u0 - 1;
      for (int j = threadIdx.x; j < Ny; j += EDGE_THREADS) {
        int node = i + j * Nx;
        if (m[node] && j + 1 < Ny)
          u[node] = ((rho[node] - rho_pr[node]) * hx[i] * hy[j] / g.ht
                     + (rho_v[node + Nx] * v[node + Nx] - rho_v[node] * v[node]) * hx[i]
                     + rho_u[node + 1] * u[node + 1] * hy[j])
                    / (rho_u[node] * hy[j]);
      }
    }
    __syncthreads();
    edge_x(g, u, g.Nx_u0, g.Nx_u1, false, bc & BC_DUDX_XE);
    __syncthreads();
    if (bc & BC_DURHODX_XE)
      for (int j = threadIdx.x; j < Ny; j += EDGE_THREADS) {
        int node = g.Nx_u1 - 1 + j * Nx;
        if (m[node]) u[node + 1] = (rho_u[node] * u[node]) / rho_u[node + 1];
      }
    __syncthreads();
    if (bc & BC_DRHODT_XE) {
      const int i = g.Nx_u1 - 1;
      for (int j = threadIdx.x; j < Ny; j += EDGE_THREADS) {
        int node = i + j * Nx;
        if (m[node] && j + 1 < Ny)
          u[node + 1] = (-(rho[node] - rho_pr[node]) * hx[i] * hy[j] / g.ht
                         - (rho_v[node + Nx] * v[node + Nx] - rho_v[node] * v[node]) * hx[i]
                         + rho_u[node] * u[node] * hy[j])
                        / (rho_u[node + 1] * hy[j]);
      }
    }
    __syncthreads();
    if (bc & BC_UMINMOUT_XE) {  // S:7031-7059 (serial build): the last row's rho_u
      const R Min = rho_u[g.Nx_u0 + (Ny - 1) * Nx];
      const R Mout = rho_u[g.Nx_u1 + (Ny - 1) * Nx];
      for (int j = threadIdx.x; j < Ny; j += EDGE_THREADS) {
        int node = g.Nx_u1 - 1 + j * Nx;
        u[node + 1] = u[node] * Min / Mout;
      }
    }
    __syncthreads();
    edge_y(g, u, g.Ny_u0, g.Ny_u1, bc & BC_DUDY_YB, false);
    __syncthreads();
    edge_y(g, u, g.Ny_u0, g.Ny_u1, false, bc & BC_DUDY_YE);
    __syncthreads();
  }
  if (parts & STS_BC_TEMPERATURE) {
    R *T = FLD(T);
    edge_x(g, T, g.Nx_o0, g.Nx_o1, bc & BC_DTDX_XB, false);
    __syncthreads();
    edge_x(g, T, g.Nx_o0, g.Nx_o1, false, bc & BC_DTDX_XE);
    __syncthreads();
    edge_y(g, T, g.Ny_o0, g.Ny_o1, bc & BC_DTDY_YB, false);
    __syncthreads();
    edge_y(g, T, g.Ny_o0, g.Ny_o1, false, bc & BC_DTDY_YE);
    __syncthreads();
  }
  if (parts & STS_BC_PRESSURE) {
    R *p = FLD(P);
    edge_x(g, p, g.Nx_o0, g.Nx_o1, bc & BC_DPDX_XB, false);
    __syncthreads();
    edge_x(g, p, g.Nx_o0, g.Nx_o1, false, bc & BC_DPDX_XE);
    __syncthreads();
    edge_y(g, p, g.Ny_o0, g.Ny_o1, bc & BC_DPDY_YB, false);
    __syncthreads();
    edge_y(g, p, g.Ny_o0, g.Ny_o1, false, bc & BC_DPDY_YE);
    __syncthreads();
  }
  if (probe && threadIdx.x == 0) {
    // S:7512-7513: the cell p_BC_xb is read from (not touched by the edge copies)
    int jj = int((g.Ny_o1 - 1.0) - g.Ny_o0);
    probe[0] = (jj >= 0 && jj < Ny) ? (double)FLD(P)[g.Nx_o0 + jj * Nx] : 0.0;
  }
}

template <typename R>
__global__ void __launch_bounds__(EDGE_THREADS) k_edges(KGridT<R> g, KFieldsT<R> fl, int parts, double *probe,
                                                        Gate gate, LoopTest lt) {
  if (gated_out(gate)) return;
  edges_block(g, fl, parts, probe, gate, lt, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// A10 + A11: fluxes (S:10440-10454) and diffusion coefficients
// (S:10457-10618) over all cells.
// ---------------------------------------------------------------------------
// FLUX_ROWS cells per thread, rows BY apart, all loads issued first.
constexpr int FLUX_ROWS = 2;

// PARTS: FLUX_PART (Fx, Fy, DTx, DTy) and/or DIFF_PART (Gamma, D_ux, D_uy,
// D_vx, D_vy).
template <typename R, int PARTS>
__global__ void __launch_bounds__(BX *BY) k_flux_diff(KGridT<R> g, KFieldsT<R> fl, Gate gate) {
  if (gated_out(gate)) return;
  int i0, j0;
  tile_origin<BX, BY * FLUX_ROWS>(g.Nx, i0, j0);
  FluxIn<R> c[FLUX_ROWS];
#pragma unroll
  for (int q = 0; q < FLUX_ROWS; q++)
    c[q].template load<(PARTS & FLUX_PART) != 0, (PARTS & DIFF_PART) != 0>(g, fl, i0 + (int)threadIdx.x,
                                                                          j0 + (int)threadIdx.y + q * BY);
#pragma unroll
  for (int q = 0; q < FLUX_ROWS; q++) {
    if (PARTS & FLUX_PART) c[q].apply_flux(g, fl);
    if (PARTS & DIFF_PART) c[q].apply_diff(g, fl);
  }
}

// ---------------------------------------------------------------------------
// Setup and output helpers
// ---------------------------------------------------------------------------
// ApplyBodiesConditionsToVariables, S:6470-6584, per face: the reference's
// i-major walk lets cell (i,j) overwrite what cell (i-1,j) wrote to u[ij]
// (and cell (i,j) overwrite cell (i,j-1) for v[ij]).
__global__ void __launch_bounds__(BX *BY) k_apply_bodies(KGrid g, KFields fl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned Nx = (unsigned)g.Nx;
  const unsigned k = (unsigned)i + (unsigned)j * Nx;
  const DevBody *b = body_of(g, g.vol[k]);
  const DevBody *bw = (i >= 1) ? body_of(g, g.vol[k - 1u]) : nullptr;
  const DevBody *bs = (j >= 1) ? body_of(g, g.vol[k - Nx]) : nullptr;
  if (b) FLD(U)[k] = b->u_body;
  else if (bw) FLD(U)[k] = bw->u_body;
  if (b) FLD(V)[k] = b->v_body;
  else if (bs) FLD(V)[k] = bs->v_body;
  if (b) {
    FLD(P)[k] = b->p_body;
    FLD(T)[k] = b->T_body;
    FLD(RHO)[k] = g.m[k] ? b->p_body / b->T_body : 0.0;
  }
}

// S:8048-8099 after the wall temperatures are refreshed: each body cell gets
// the wall temperature of the face to its last fluid neighbour in the
// reference's walk order (above, right, left, below; periodic about OX, the
// neighbour across the seam ranks by its node number).
__global__ void __launch_bounds__(BX *BY) k_jump_overwrite(KGrid g, KFields fl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned Nx = (unsigned)g.Nx;
  const unsigned k = (unsigned)i + (unsigned)j * Nx;
  const uint8_t *m = g.m;
  if (m[k]) return;
  double *T = FLD(T);
  const bool seam_w = g.px && i == 0, seam_e = g.px && i == g.Nx - 1;
  if (j + 1 < g.Ny && m[k + Nx]) T[k] = FLD(TW_V)[k + Nx];
  else if (seam_w && m[k + Nx - 1u]) T[k] = FLD(TW_U)[k];  // last column, node k + Nx - 1
  else if (i + 1 < g.Nx && m[k + 1u]) T[k] = FLD(TW_U)[k + 1u];
  else if (i >= 1 && m[k - 1u]) T[k] = FLD(TW_U)[k];
  else if (seam_e && m[k + 1u - Nx]) T[k] = FLD(TW_U)[k + 1u - Nx];  // first column, node k - Nx + 1
  else if (j >= 1 && m[k - Nx]) T[k] = FLD(TW_V)[k];
}

__global__ void __launch_bounds__(BX *BY) k_copy5(const double *a0, const double *a1, const double *a2,
                                                  const double *a3, const double *a4, double *b0, double *b1,
                                                  double *b2, double *b3, double *b4, int n) {
  const int t = (blockIdx.x + blockIdx.y * gridDim.x) * (BX * BY) + threadIdx.x + threadIdx.y * BX;
  if (t >= n) return;
  b0[t] = a0[t];
  b1[t] = a1[t];
  b2[t] = a2[t];
  b3[t] = a3[t];
  b4[t] = a4[t];
}

// Start of loop 2 from a linear extrapolation in time: x = 2 x^n - x^(n-1)
// for u, v, p, T on their update sets (x^n = the *_pr fields after the
// prologue); x^(n-1) (prev) then takes x^n.
// Polynomial extrapolation in time of degree `order` through x^n and the
// stored steps, as Newton's backward-difference sum x0 + D1 + ... + Dq
// (D1 = x0 - x1, D(k+1) = Dk(x0) - Dk(x1)), so that a value constant in time
// (fixed inflow, outflow pressure, body values) is kept exactly; then the
// history shifts (keep = number of levels stored).  Every cell is
// extrapolated: besides the update sets, the cells the boundary conditions
// copy into (e.g. the inflow pressure, which sets the inflow density) carry
// the step's slowest-converging change.
__device__ __forceinline__ double extrap(int order, const double (&x)[MAX_EXTRAP + 1]) {
  double d[MAX_EXTRAP + 1];
#pragma unroll
  for (int h = 0; h <= MAX_EXTRAP; h++) d[h] = x[h];
  double s = x[0];
#pragma unroll
  for (int k = 1; k <= MAX_EXTRAP; k++) {
    if (k > order) break;
#pragma unroll
    for (int h = 0; h + k <= MAX_EXTRAP; h++) d[h] = d[h] - d[h + 1];  // d[h] = D^k x^(n-h)
    s += d[0];
  }
  return s;
}

// Reads x^n and hist[0 .. order-1]; stores x^n into hist[slot] (the host then
// rotates the level pointers, so no level is copied).
__global__ void __launch_bounds__(BX *BY) k_extrapolate(KGrid g, KFields fl, ExtrapHist hist, int order, int slot) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned k = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
  const int src[4] = {STS_U_PR, STS_V_PR, STS_P_PR, STS_T_PR}, dst[4] = {STS_U, STS_V, STS_P, STS_T};
  // the inflow columns i < Nx_u0 hold the given velocity, which no kernel of
  // loop 2 rewrites and which the driver advances in time itself
  // (BoundaryConditions_function_of_time, S:8102-8144): u keeps x^n there
  const bool given_u = i < g.Nx_u0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double x[MAX_EXTRAP + 1];
    x[0] = fl.f[src[q]][k];
#pragma unroll
    for (int h = 0; h < MAX_EXTRAP; h++) x[h + 1] = (h < order) ? hist.f[h][q][k] : 0.0;
    if (order > 0 && !(q == 0 && given_u)) {
      // a guess that is not finite, or a pressure or temperature that is not
      // positive, keeps the previous step's value in that cell
      const double e = extrap(order, x);
      const bool ok = isfinite(e) && (q < 2 || e > 0.0 || x[0] <= 0.0);
      fl.f[dst[q]][k] = ok ? e : x[0];
    }
    hist.f[slot][q][k] = x[0];
  }
}

// ---------------------------------------------------------------------------
// Anderson acceleration of loop 2 (FP64), difference form.  The vectors of
// the acceleration hold the four state values of a cell side by side
// (v[4 k + q], q = u, v, p, T), so a thread moves a cell's entry of each
// vector with one or two 16-byte accesses.
// ---------------------------------------------------------------------------
constexpr int AA_THREADS = 256;
constexpr int AA_SOLVE_THREADS = 1024;

// One step of wave_sum_last: v plus its DPP-moved copy; rows outside
// ROW_MASK add +0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_sum_step(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const unsigned long long m =
      ((unsigned long long)(unsigned)__builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false) << 32) |
      (unsigned long long)(unsigned)__builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
  return v + __longlong_as_double((long long)m);
}

// Sum over a full wave, complete in lane 63, in a fixed order: quads, half
// rows and rows by DPP permutations and mirrors, then rows 1 and 3 add rows 0
// and 2 (broadcast of lane 15) and rows 2 and 3 add rows 0-1 (broadcast of
// lane 31).  Every lane of the wave must be active.
__device__ __forceinline__ double wave_sum_last(double v) {
  v = dpp_sum_step<0xb1, 0xf>(v);   // quad_perm(1, 0, 3, 2)
  v = dpp_sum_step<0x4e, 0xf>(v);   // quad_perm(2, 3, 0, 1)
  v = dpp_sum_step<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_sum_step<0x140, 0xf>(v);  // row_mirror
  v = dpp_sum_step<0x142, 0xa>(v);  // row_bcast:15
  v = dpp_sum_step<0x143, 0xc>(v);  // row_bcast:31
  return v;
}

// Sums acc[0..NV) over a workgroup of NT threads (fixed order: lanes, then
// waves in turn); thread t < NV gets sum t in out[t].  Only the sums j with
// live(j) are formed, the others are 0 (their acc is unused).
template <int NV, int NT, typename Live>
__device__ __forceinline__ void block_sums(const double (&acc)[NV], double *out, Live live) {
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ double sm[NV][NT / 64];
  const int lane = (int)threadIdx.x & 63, wid = (int)threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; j++) {
    const double w = live(j) ? wave_sum_last(acc[j]) : 0.0;
    if (lane == 63) sm[j][wid] = w;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NT / 64; q++) s += sm[threadIdx.x][q];
    out[threadIdx.x] = s;
  }
}

__device__ __forceinline__ void load4(const double *p, double (&v)[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void store4(double *p, const double (&v)[4]) {
  reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void load4(const float *p, double (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4 *>(p);
  v[0] = (double)a.x;
  v[1] = (double)a.y;
  v[2] = (double)a.z;
  v[3] = (double)a.w;
}
__device__ __forceinline__ void store4(float *p, const double (&v)[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}

// A cell's mixed value v of component q (0..3: u, v, p, T) from the newest
// iterate gp: v itself, or gp if v is not finite or is a p or T that is not
// positive where gp's is.
__device__ __forceinline__ double aa_keep(double v, double gp, int q) {
  return (!isfinite(v) || (q >= 2 && !(v > 0.0) && gp > 0.0)) ? gp : v;
}

// Loop start: the start x = gp + D with gp = the state, D = 0.  The state
// may be FP32 (the FP32 phase); gp holds it widened, exactly.
template <typename R>
__global__ void __launch_bounds__(AA_THREADS) k_aa_begin(AAFieldsT<R> s, AABuf ab, int Na) {
  const int k = (int)(blockIdx.x * AA_THREADS + threadIdx.x);
  if (k >= Na) return;
  const double g[4] = {(double)s.x[0][k], (double)s.x[1][k], (double)s.x[2][k], (double)s.x[3][k]},
               z[4] = {0.0, 0.0, 0.0, 0.0};
  store4(ab.gp + 4 * (size_t)k, g);
  store4(ab.D + 4 * (size_t)k, z);
}

// After an outer iteration: g = the state of cell k, x its start (sel.start:
// gp + D, gp, or the last mix formed again exactly as aa_mix_tile formed it,
// rounded to R), f = g - x.  First iteration of the loop (sel.n = 0): the next start is g
// (gp = g, D = 0), fp = f.  Later: the newest differences dF = f - fp,
// dG = g - gp go to sel.dF/dG[n-1] (dF only if a later store reads it,
// sel.keep_dF), fp = f, gp = g, and acc[j] += dF_j . dF_new,
// acc[NM + j] += dF_j . f for every difference in use.
// NM >= sel.n bounds the differences at compile time.
template <int NM, typename R>
__device__ __forceinline__ void aa_store_cell(const AAFieldsT<R> &s, const AABuf &ab, const AASel &sel, int k,
                                              double (&acc)[2 * NM]) {
  const int n = sel.n;
  const size_t e = 4 * (size_t)k;
  double g[4], gp[4], D[4] = {0.0, 0.0, 0.0, 0.0}, fp[4], o[NM][4];
#pragma unroll
  for (int q = 0; q < 4; q++) g[q] = (double)s.x[q][k];
  load4(ab.gp + e, gp);
  if (sel.start == AA_START_D) load4(ab.D + e, D);
  else if (sel.start == AA_START_MIX) load4(sel.dG[0] + e, D);  // the last mix's difference (depth 1)
  if (n > 0) load4(ab.fp + e, fp);
#pragma unroll
  for (int j = 0; j < NM; j++) {
#pragma unroll
    for (int q = 0; q < 4; q++) o[j][q] = 0.0;
    if (j < n - 1) load4(sel.dF[j] + e, o[j]);
  }
  double f[4], dF[4], dG[4];
  if (sel.start == AA_START_MIX) {
    const double gam0 = ab.gam[0];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double v = gp[q];
      v -= gam0 * D[q];
      f[q] = g[q] - (double)(R)aa_keep(v, gp[q], q);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) f[q] = g[q] - (sel.start == AA_START_D ? gp[q] + D[q] : gp[q]);
  }
  if (n == 0) {
    const double z[4] = {0.0, 0.0, 0.0, 0.0};
    if (sel.write_D) store4(ab.D + e, z);
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      dF[q] = (double)(float)(f[q] - fp[q]);
      dG[q] = g[q] - gp[q];
    }
#pragma unroll
    for (int j = 0; j < NM; j++)
      if (j == n - 1) {
        if (sel.keep_dF) store4(sel.dF[j] + e, dF);
        store4(sel.dG[j] + e, dG);
#pragma unroll
        for (int q = 0; q < 4; q++) o[j][q] = dF[q];
      }
#pragma unroll
    for (int j = 0; j < NM; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        acc[j] += o[j][q] * dF[q];
        acc[NM + j] += o[j][q] * f[q];
      }
  }
  store4(ab.gp + e, g);
  store4(ab.fp + e, f);
}

// The workgroup's (NT threads) sums of acc into partial slot idx:
// part[j nb + idx] (a_j) and part[(MAX_AA + j) nb + idx] (b_j).
template <int NM, int NT>
__device__ __forceinline__ void aa_store_sums(const double (&acc)[2 * NM], const AABuf &ab, int idx, int n) {
  __shared__ double sums[2 * NM];
  block_sums<2 * NM, NT>(acc, sums, [n](int j) { return j % NM < n; });
  __syncthreads();
  const int j = (int)threadIdx.x;
  if (j < 2 * NM) ab.part[(size_t)(j < NM ? j : MAX_AA + j - NM) * ab.nb + idx] = sums[j];
}

// AA_STORE_CELLS cells per thread (AA_THREADS apart), so that the partial
// sums the weights are folded from are few.
constexpr int AA_STORE_CELLS = 2;
template <int NM, typename R>
__global__ void __launch_bounds__(AA_THREADS) k_aa_store(AAFieldsT<R> s, AABuf ab, AASel sel, int Na, Gate gate) {
  if (gated_out(gate)) return;
  double acc[2 * NM];
#pragma unroll
  for (int j = 0; j < 2 * NM; j++) acc[j] = 0.0;
#pragma unroll
  for (int r = 0; r < AA_STORE_CELLS; r++) {
    const int k = (int)((blockIdx.x * AA_STORE_CELLS + r) * AA_THREADS + threadIdx.x);
    if (k < Na) aa_store_cell<NM, R>(s, ab, sel, k, acc);
  }
  if (sel.n > 0) aa_store_sums<NM, AA_THREADS>(acc, ab, (int)blockIdx.x, sel.n);
}

// One workgroup of NT threads: folds the store pass's ab.nb partial sums,
// completes the Gram matrix of the differences in use (the newest row from
// the sums; the rest as stored), and solves (Gram + 1e-12 trace I) gamma = b
// by Gaussian elimination with partial pivoting.  A singular or non-finite
// solve gives gamma = 0 (the next start is g).  The weights go to gam_out
// (shared memory of the workgroup, valid after a barrier); a leader
// workgroup also stores the newest Gram row, for the next steps, and the
// weights in ab.gam.  Workgroups that solve the same system concurrently
// read only rows other than the newest, which the leader alone writes.
template <int NM, int NT>
__device__ __forceinline__ void aa_solve_group(const AABuf &ab, const AASel &sel, double *gam_out, bool leader) {
  const int n = sel.n;
  double acc[2 * NM];
#pragma unroll
  for (int j = 0; j < 2 * NM; j++) acc[j] = 0.0;
  // four rows of partial sums per thread and pass, their loads issued together
  for (int b0 = 0; b0 < ab.nb; b0 += 4 * NT) {
    double v[4][2 * NM];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int b = b0 + r * NT + (int)threadIdx.x;
#pragma unroll
      for (int j = 0; j < NM; j++) {
        const bool use = j < n && b < ab.nb;
        v[r][j] = use ? ab.part[(size_t)j * ab.nb + b] : 0.0;
        v[r][NM + j] = use ? ab.part[(size_t)(MAX_AA + j) * ab.nb + b] : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int j = 0; j < 2 * NM; j++) acc[j] += v[r][j];
  }
  __shared__ double dots[2 * NM];
  __shared__ double A[NM][NM + 1];
  double *gam = gam_out;
  if ((int)threadIdx.x < NM * NM) {  // the stored part of the Gram matrix
    const int a = (int)threadIdx.x / NM, b = (int)threadIdx.x % NM;
    if (a < n - 1 && b < n - 1) A[a][b] = ab.gram[sel.slot[a] * MAX_AA + sel.slot[b]];
  }
  block_sums<2 * NM, NT>(acc, dots, [n](int j) { return j % NM < n; });
  __syncthreads();
  if (leader && (int)threadIdx.x < n) {  // the newest row, kept for the next steps
    const int kn = sel.slot[n - 1], sa = sel.slot[threadIdx.x];
    ab.gram[sa * MAX_AA + kn] = dots[threadIdx.x];
    ab.gram[kn * MAX_AA + sa] = dots[threadIdx.x];
  }
  if (threadIdx.x != 0) return;
  double tr = 0.0;
  for (int a = 0; a < n; a++) {
    A[a][n - 1] = dots[a];
    A[n - 1][a] = dots[a];
  }
  for (int a = 0; a < n; a++) {
    A[a][n] = dots[NM + a];
    tr += A[a][a];
  }
  for (int a = 0; a < n; a++) A[a][a] += 1e-12 * tr;
  bool ok = true;
  for (int col = 0; col < n && ok; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (!(fabs(A[piv][col]) > 0.0)) {
      ok = false;
      break;
    }
    if (piv != col)
      for (int q = col; q <= n; q++) {
        const double t = A[col][q];
        A[col][q] = A[piv][q];
        A[piv][q] = t;
      }
    for (int r = col + 1; r < n; r++) {
      const double l = A[r][col] / A[col][col];
      for (int q = col; q <= n; q++) A[r][q] -= l * A[col][q];
    }
  }
  for (int a = n - 1; a >= 0 && ok; a--) {
    double t = A[a][n];
    for (int q = a + 1; q < n; q++) t -= A[a][q] * gam[q];
    gam[a] = t / A[a][a];
    ok = isfinite(gam[a]);
  }
  for (int a = 0; a < n; a++)
    if (!ok) gam[a] = 0.0;
  if (leader)
    for (int a = 0; a < MAX_AA; a++) ab.gam[a] = a < n ? gam[a] : 0.0;
}

// Mixed state of cell k: x = gp - sum_j gamma_j dG_j (gp = the newest
// iterate, also returned); a value that is not finite, or a p or T that is
// not positive where gp's is, falls back to gp.
template <int NM>
__device__ __forceinline__ void aa_mixed(const AABuf &ab, const AASel &sel, const double (&gam)[NM], unsigned k,
                                         double (&x)[4], double (&gp)[4]) {
  const size_t e = 4 * (size_t)k;
  double dG[NM][4];
  load4(ab.gp + e, gp);
#pragma unroll
  for (int j = 0; j < NM; j++)
    if (j < sel.n) load4(sel.dG[j] + e, dG[j]);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double v = gp[q];
#pragma unroll
    for (int j = 0; j < NM; j++)
      if (j < sel.n) v -= gam[j] * dG[j][q];
    x[q] = aa_keep(v, gp[q], q);
  }
}

// The next start x of a TX x TY tile (origin i0, j0; this thread tx, ty)
// into the state arrays (and D = x - gp when sel.write_D), and, from it,
// what the correction pass derives without correcting: sqrt_T, rho and the
// face densities
// (CorrectIn::apply on the mixed p, T of the cell and of its west and south
// neighbours); with FLUX also what the loop's flux pass writes (Fx, Fy, DTx,
// DTy; FluxIn::apply_flux).  The mixed p, T of the tile and of its west
// column and south row (formed by the first threads) meet in LDS (sp, sT:
// [1 + row][1 + column], row 0 and column 0 the halo); with FLUX their
// sqrt(T) too (ss), each formed once.
template <int NM, int TX, int TY, bool FLUX, typename R>
__device__ __forceinline__ void aa_mix_tile(const KGridT<R> &g, const KFieldsT<R> &fl, const AAFieldsT<R> &s,
                                            const AABuf &ab, const AASel &sel, const double *gsrc, int i0, int j0,
                                            int tx, int ty, double (*sp)[TX + 1], double (*sT)[TX + 1],
                                            double (*ss)[TX + 1]) {
  double gam[NM];
#pragma unroll
  for (int j = 0; j < NM; j++) gam[j] = gsrc[j];
  const int Nx = g.Nx, Ny = g.Ny;
  const int i = i0 + tx, j = j0 + ty;
  const bool in = i < Nx && j < Ny;
  const Nbr nb((unsigned)(in ? i : 0), (unsigned)(in ? j : 0), (unsigned)Nx, (unsigned)Ny, g.px);
  // the mixed values as the state holds them (rounded to R), here and in LDS
  double x[4] = {0.0, 0.0, 0.0, 0.0}, gp[4] = {0.0, 0.0, 0.0, 0.0};
  if (in) {
    aa_mixed<NM>(ab, sel, gam, nb.ij, x, gp);
#pragma unroll
    for (int q = 0; q < 4; q++) x[q] = (double)(R)x[q];
    sp[ty + 1][tx + 1] = x[2];
    sT[ty + 1][tx + 1] = x[3];
    if constexpr (FLUX) ss[ty + 1][tx + 1] = (double)sqrt((R)x[3]);
  }
  // halo: thread (t, 0) the cell south of column t, thread (t < TY, 1) the
  // cell west of row t (the clamped or wrapped neighbour of the reference)
  double h[4], hg[4];
  if (ty == 0 && in) {
    aa_mixed<NM>(ab, sel, gam, nb.ij_1, h, hg);
    sp[0][tx + 1] = (double)(R)h[2];
    sT[0][tx + 1] = (double)(R)h[3];
    if constexpr (FLUX) ss[0][tx + 1] = (double)sqrt((R)h[3]);
  }
  if (ty == 1 && tx < TY && i0 < Nx && j0 + tx < Ny) {
    const Nbr w((unsigned)i0, (unsigned)(j0 + tx), (unsigned)Nx, (unsigned)Ny, g.px);
    aa_mixed<NM>(ab, sel, gam, w.i_1j, h, hg);
    sp[tx + 1][0] = (double)(R)h[2];
    sT[tx + 1][0] = (double)(R)h[3];
    if constexpr (FLUX) ss[tx + 1][0] = (double)sqrt((R)h[3]);
  }
  __syncthreads();
  if (!in) return;
  CorrectIn<R> ci;
  ci.in = true;
  ci.cu = ci.cv = false;
  ci.ij = nb.ij;
  ci.i_1j = nb.i_1j;
  ci.ij_1 = nb.ij_1;
  const unsigned bits = g.nb[nb.ij];
  ci.mc = (bits & NB_C) != 0;
  ci.mw = (bits & NB_W) != 0;
  ci.ms = (bits & NB_S) != 0;
  ci.uf = (R)x[0];
  ci.vf = (R)x[1];
  ci.pc = (R)x[2];
  ci.Tc = (R)x[3];
  ci.pw = (R)sp[ty + 1][tx];
  ci.Tw = (R)sT[ty + 1][tx];
  ci.ps = (R)sp[ty][tx + 1];
  ci.Ts = (R)sT[ty][tx + 1];
  double d[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    s.x[q][nb.ij] = (R)x[q];
    d[q] = x[q] - gp[q];
  }
  if (sel.write_D) store4(ab.D + 4 * (size_t)nb.ij, d);
  KFieldsT<R> kf = fl;
  kf.f[STS_P] = s.x[2];
  kf.f[STS_T] = s.x[3];
  unsigned long long rr[2] = {0ull, 0ull};
  R rho_uv[2];
  ci.apply(kf, rr, rho_uv);
  if (!FLUX) return;
  // the loop's flux pass on the mixed state (fluxes, thermal coefficients):
  // its inputs at this cell and the west and south sqrt_T are all here
  FluxIn<R> fi;
  fi.in = true;
  fi.ij = nb.ij;
  fi.hxi = g.hx[i];
  fi.hxw = g.hx[nb.i_1];
  fi.hyj = g.hy[j];
  fi.hys = g.hy[j - 1 * (0 < j)];
  fi.ru = rho_uv[0];
  fi.u = (R)x[0];
  fi.rv = rho_uv[1];
  fi.v = (R)x[1];
  fi.m10 = ci.ms;
  fi.m01 = ci.mw;
  fi.m11 = ci.mc;
  fi.s10 = (R)ss[ty][tx + 1];
  fi.s01 = (R)ss[ty + 1][tx];
  fi.s11 = (R)ss[ty + 1][tx + 1];
  // a wall temperature enters DTx (DTy) only where exactly one of the cell
  // and its west (south) neighbour is fluid: with both fluid its term is
  // 0 * sqrt(Tw), +0 for every wall temperature the state holds (finite,
  // >= 0) as for 0, and with neither DTx is 0; those loads are left out
  fi.twu11 = ci.mc == ci.mw ? R(0) : FLD(TW_U)[nb.ij];
  fi.twv11 = ci.mc == ci.ms ? R(0) : FLD(TW_V)[nb.ij];
  fi.apply_flux(g, kf);
}

// The weights and the mix in one launch: every workgroup folds the store
// pass's partial sums itself (the same sums in the same order, so the same
// weights; workgroup 0 stores the Gram row), then mixes AA_MIX_TILES tiles
// of AA_MIX_TX x AA_MIX_TY cells in turn.
constexpr int AA_MIX_TX = 64, AA_MIX_TY = 8, AA_MIX_TILES = 3;
template <int NM, typename R>
__global__ void __launch_bounds__(AA_MIX_TX *AA_MIX_TY) k_aa_mix(KGridT<R> g, KFieldsT<R> fl, AAFieldsT<R> s,
                                                                 AABuf ab, AASel sel, Gate gate) {
  if (gated_out(gate)) return;
  __shared__ double gam[NM];
  aa_solve_group<NM, AA_MIX_TX * AA_MIX_TY>(ab, sel, gam, blockIdx.x == 0);
  __syncthreads();
  __shared__ double sp[AA_MIX_TY + 1][AA_MIX_TX + 1], sT[AA_MIX_TY + 1][AA_MIX_TX + 1],
      ss[AA_MIX_TY + 1][AA_MIX_TX + 1];
  const int ntx = (g.Nx + AA_MIX_TX - 1) / AA_MIX_TX, ntiles = ntx * ((g.Ny + AA_MIX_TY - 1) / AA_MIX_TY);
  const int tx = (int)threadIdx.x % AA_MIX_TX, ty = (int)threadIdx.x / AA_MIX_TX;
  for (int t = (int)blockIdx.x; t < ntiles; t += (int)gridDim.x) {
    aa_mix_tile<NM, AA_MIX_TX, AA_MIX_TY, true, R>(g, fl, s, ab, sel, gam, (t % ntx) * AA_MIX_TX,
                                                  (t / ntx) * AA_MIX_TY, tx, ty, sp, sT, ss);
    __syncthreads();  // the next tile reuses sp, sT, ss
  }
}

__global__ void __launch_bounds__(BX *BY) k_set_inflow(KGrid g, KFields fl, double u_in) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx_u0 || j >= g.Ny) return;
  const unsigned k = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
  if (g.m[k]) FLD(U)[k] = u_in;
}

// WriteDragCoefficient (S:8203-8262): the friction and pressure terms of
// the four faces of each listed body cell, 12 per cell in the order
// (bottom, front, top, behind) x (friction, pressure, face length); a face
// whose neighbour is not fluid contributes +0 (an exact no-op in the sums).
__global__ void __launch_bounds__(256) k_drag_terms(KGrid g, KFields fl, const unsigned *cells, int n, double *terms) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= n) return;
  const unsigned Nx = (unsigned)g.Nx, Ny = (unsigned)g.Ny, k = cells[t];
  const Nbr c(k % Nx, k / Nx, Nx, Ny, false);  // clamped in every case (S:8214)
  const uint8_t *m = g.m;
  const double *hx = g.hx, *hy = g.hy, *u = FLD(U), *v = FLD(V), *p = FLD(P);
  const unsigned i = c.i, j = c.j;
  const unsigned ij = c.ij, ij_1 = c.ij_1, ij1 = c.ij1, i_1j = c.i_1j, i1j = c.i1j;
  const unsigned i1j_1 = c.i1j_1, i1j1 = c.i1j1, i_1j1 = c.i_1j1;
  double *o = terms + 12 * (size_t)t;
  for (int q = 0; q < 12; q++) o[q] = 0.0;
  if (m[ij_1]) {
    const double hb = hy[j - 1 * (j < (Ny - 1))];
    o[0] = hx[i] * sqrt(FLD(TW_V)[ij]) * ((u[ij_1] - u[ij]) / hb + (u[i1j_1] - u[i1j]) / hb);
    o[1] = p[ij_1] * hx[i];
    o[2] = hx[i];
  }
  if (m[i_1j]) {
    o[3] = hy[j] * sqrt(FLD(TW_U)[ij]) * ((v[i_1j] - v[ij]) / hx[c.i_1] + (v[i_1j1] - v[ij1]) / hx[c.i_1]);
    o[4] = p[i_1j] * hy[j];
    o[5] = hy[j];
  }
  if (m[ij1]) {
    const double ht_ = hy[j + 1 * (0 < j)];
    o[6] = hx[i] * sqrt(FLD(TW_V)[ij1]) * ((u[ij1] - u[ij]) / ht_ + (u[i1j1] - u[i1j]) / ht_);
    o[7] = p[ij1] * hx[i];
    o[8] = hx[i];
  }
  if (m[i1j]) {
    o[9] = hy[j] * sqrt(FLD(TW_U)[i1j]) * ((v[i1j] - v[ij]) / hx[c.i1] + (v[i1j1] - v[ij1]) / hx[c.i1]);
    o[10] = p[i1j] * hy[j];
    o[11] = hy[j];
  }
}

// The sums of k_drag_terms in the reference's cell order (one thread: the
// order of the additions is the reference's, so the sums are bit-identical).
__global__ void k_drag_sum(const double *terms, int n, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[12];
  for (int q = 0; q < 12; q++) s[q] = 0.0;
  for (int t = 0; t < n; t++)
    for (int q = 0; q < 12; q++) s[q] += terms[12 * (size_t)t + q];
  for (int q = 0; q < 12; q++) out[q] = s[q];
}

// Opens the loop-2 control of a time step and clears its residual log and
// the slotted residual words.
__global__ void k_loop_open(LoopCtrl *ctrl, unsigned long long *log, int n, unsigned long long *lred) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    ctrl->stop_at = 0x7fffffff;
    ctrl->nonfinite = 0;
  }
  for (int q = t; q < n; q += gridDim.x * blockDim.x) log[q] = 0ull;
  for (int q = t; q < NSLOT * SLOT_STRIDE; q += gridDim.x * blockDim.x) lred[q] = 0ull;
}

// Element-wise type conversion of the fields selected by mask (bit f = field
// f) between the FP64 state and its FP32 copy (round to nearest / exact).
template <typename S, typename D>
__global__ void __launch_bounds__(256) k_convert(KFieldsT<S> src, KFieldsT<D> dst, unsigned long long mask, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int f = 0; f < STS_NFIELDS; f++)
    if ((mask >> f) & 1ull) dst.f[f][t] = (D)src.f[f][t];
}

// Folds the NSLOT copies of words 0..n-1 of a slotted buffer into out[0..n-1]
// and clears the copies.
__global__ void k_fold(unsigned long long *slots, int n, unsigned long long *out) {
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    unsigned long long b = 0ull;
    for (int s = 0; s < NSLOT; s++) {
      const unsigned long long w = slots[q + s * SLOT_STRIDE];
      b = (w > b) ? w : b;
      slots[q + s * SLOT_STRIDE] = 0ull;
    }
    out[q] = b;
  }
}

// SolveFluxes alone (S:10440-10454), as SolveContinuityEquation calls it.
__global__ void __launch_bounds__(BX *BY) k_fluxes(KGrid g, KFields fl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const unsigned ij = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
  FLD(FX)[ij] = g.hy[j] * FLD(RHO_U)[ij] * FLD(U)[ij];
  FLD(FY)[ij] = g.hx[i] * FLD(RHO_V)[ij] * FLD(V)[ij];
}

// S:10621-10646 plus its norms over fluid cells (one partial per block).
__global__ void __launch_bounds__(BX *BY) k_continuity(KGrid g, KFields fl, double *cs, double *cf,
                                                       double *partial) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  const int Nx = g.Nx;
  double amax = 0.0, s2 = 0.0, s1 = 0.0;
  if (i < g.Nx - 1 && j < g.Ny - 1) {
    const int ij = i + j * Nx;
    const double *rho_u = FLD(RHO_U), *rho_v = FLD(RHO_V), *u = FLD(U), *v = FLD(V);
    double Fx0 = g.hy[j] * rho_u[ij] * u[ij], Fx1 = g.hy[j] * rho_u[ij + 1] * u[ij + 1];
    double Fy0 = g.hx[i] * rho_v[ij] * v[ij], Fy1 = g.hx[i] * rho_v[ij + Nx] * v[ij + Nx];
    double st = Fx1 - Fx0 + Fy1 - Fy0;
    double c = (FLD(RHO)[ij] - FLD(RHO_PR)[ij]) * g.hx[i] * g.hy[j] / g.ht + st;
    cs[ij] = st;
    cf[ij] = c;
    if (g.m[ij]) {
      amax = fabs(c);
      s2 = c * c;
      s1 = fabs(c);
    }
  }
  // wavefront reductions by lane exchange, then one value per wave in LDS
  // folded by the first wave
  constexpr int NW = (BX * BY) / 64;
  __shared__ double sm[3][NW];
  const int t = threadIdx.x + threadIdx.y * BX, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    amax = fmax(amax, __shfl_xor(amax, o, 64));
    s2 += __shfl_xor(s2, o, 64);
    s1 += __shfl_xor(s1, o, 64);
  }
  if (lane == 0) {
    sm[0][w] = amax;
    sm[1][w] = s2;
    sm[2][w] = s1;
  }
  __syncthreads();
  if (t == 0) {
    double a = sm[0][0], b2 = sm[1][0], b1 = sm[2][0];
#pragma unroll
    for (int q = 1; q < NW; q++) {
      a = fmax(a, sm[0][q]);
      b2 += sm[1][q];
      b1 += sm[2][q];
    }
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    partial[3 * b] = a;
    partial[3 * b + 1] = b2;
    partial[3 * b + 2] = b1;
  }
}

// ---------------------------------------------------------------------------
// Batched loop-2 kernels: one launch covers the tiles of every case of a
// batch (BatchTile list); each workgroup takes its case's grid and fields
// from the BatchCase table and obeys that case's loop control.  The per-cell
// work is the same code as the single-case kernels above.  par = 1 selects
// the p/T partners as the current p, T.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ KFieldsT<R> fields_at(const BatchCase<R> &c, int par) {
  KFieldsT<R> f = c.f;
  if (par) {
    f.f[STS_P] = c.p_alt;
    f.f[STS_T] = c.T_alt;
  }
  return f;
}

// The assembly forms Gamma and D itself (assemble_inl_tile), except in the
// first outer iteration of a case whose arrays are to be read (inline0 = 0),
// as the single-case loops do: kb_assemble_arr does those, in a kernel of its
// own so that the two forms do not share one register budget.
template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_assemble(const BatchCase<R> *C, const BatchTile *T, int k, int par) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k}) || (k == 0 && !c.inline0)) return;
  assemble_inl_tile<R, TX, TY>(c.g, fields_at(c, par), t.x * TX, t.y * TY, (int)threadIdx.x % TX,
                               (int)threadIdx.x / TX);
}

template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_assemble_arr(const BatchCase<R> *C, const BatchTile *T, int k, int par) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k}) || k != 0 || c.inline0) return;
  assemble_cell(c.g, fields_at(c, par), t.x * TX + (int)threadIdx.x % TX, t.y * TY + (int)threadIdx.x / TX);
}

// Pass of two fused sweeps from parity par; first/last: whether it holds
// loop 3's first / last sweep (their residuals go to the loop-2 test).  Six
// waves per SIMD, as k_sweep2.
template <typename R, int SX, int SY>
__global__ void __launch_bounds__(SX *SY) __attribute__((amdgpu_waves_per_eu(6, 8))) kb_sweep2(const BatchCase<R> *C, const BatchTile *T, int k, int par,
                                                    int first, int last) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const KFieldsT<R> f = fields_at(c, par);
  const R *Pi = par ? c.p_alt : c.f.f[STS_P], *Ti = par ? c.T_alt : c.f.f[STS_T];
  R *Po = par ? c.f.f[STS_P] : c.p_alt, *To = par ? c.f.f[STS_T] : c.T_alt;
  sweep2_tile<R, SX, SY>(c.g, f, Pi, Ti, Po, To, t.x * SX, t.y * SY, (int)threadIdx.x % SX, (int)threadIdx.x / SX,
                         first ? c.lred + 4 : nullptr, first ? c.lred + 5 : nullptr, last ? c.lred + 2 : nullptr,
                         last ? c.lred + 3 : nullptr);
}

// Single sweep (odd sweep counts).
template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_sweep(const BatchCase<R> *C, const BatchTile *T, int k, int par,
                                                   int first, int last) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const KFieldsT<R> fl = fields_at(c, par);
  const KGridT<R> &g = c.g;
  const R *Po = par ? c.p_alt : c.f.f[STS_P], *To = par ? c.T_alt : c.f.f[STS_T];
  R *Pn = par ? c.f.f[STS_P] : c.p_alt, *Tn = par ? c.f.f[STS_T] : c.T_alt;
  const int i = t.x * TX + (int)threadIdx.x % TX, j = t.y * TY + (int)threadIdx.x / TX;
  unsigned long long r[2] = {0ull, 0ull};
  if (i < g.Nx && j < g.Ny) {
    const unsigned ij = (unsigned)i + (unsigned)j * (unsigned)g.Nx;
    if (in_pt_set(g, i, j, ij)) {
      const Nbr n((unsigned)i, (unsigned)j, (unsigned)g.Nx, (unsigned)g.Ny, g.px);
      const PTStencil<R> x = stencil_global(Po, To, n);
      R pn, tn;
      pt_update(g, fl, n, x, pn, tn);
      Tn[ij] = tn;
      Pn[ij] = pn;
      r[0] = absbits(pn - x.Pc);
      r[1] = absbits(tn - x.Tc);
    } else {
      Pn[ij] = Po[ij];
      Tn[ij] = To[ij];
    }
  }
  block_max_atomic<2, TX * TY>(r, MaxDst<2>{{last ? c.lred + 2 : nullptr, last ? c.lred + 3 : nullptr},
                                         {first ? c.lred + 4 : nullptr, first ? c.lred + 5 : nullptr}});
}

template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_correct(const BatchCase<R> *C, const BatchTile *T, int k, int par) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const KFieldsT<R> f = fields_at(c, par);
  unsigned long long r[2] = {0ull, 0ull};
  correct_cell(c.g, f, t.x * TX + (int)threadIdx.x % TX, t.y * TY + (int)threadIdx.x / TX, 1, r);
  block_max_atomic<2, TX * TY>(r, MaxDst<2>{{c.lred + 0, c.lred + 1}, {}});
}

template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_wall(const BatchCase<R> *C, const BatchTile *T, int k, int par,
                                                  int parts) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const KFieldsT<R> f = fields_at(c, par);
  wall_cell(c.g, f, t.x * TX + (int)threadIdx.x % TX, t.y * TY + (int)threadIdx.x / TX, parts);
}

template <typename R>
__global__ void __launch_bounds__(WALL_OPS_PER_TILE) kb_wall_ops(const BatchCase<R> *C, const BatchTile *T, int k,
                                                                 int par, int parts) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const int q = t.x * WALL_OPS_PER_TILE + (int)threadIdx.x;
  if (q >= c.nwops) return;
  const WallOp o = c.wops[q];
  const unsigned kind = o.kb & 7u;
  const int need = kind == WOP_U ? STS_BC_SLIP_U : kind == WOP_V ? STS_BC_SLIP_V : STS_BC_JUMP;
  if (parts & need) apply_wall_op(c.g, fields_at(c, par), o);
}

// The fluxes and thermal coefficients (FLUX_PART); the assembly forms the
// diffusion coefficients itself.
template <typename R, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_flux_diff(const BatchCase<R> *C, const BatchTile *T, int k, int par) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<R> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k})) return;
  const KFieldsT<R> f = fields_at(c, par);
  FluxIn<R> x;
  x.template load<true, false>(c.g, f, t.x * TX + (int)threadIdx.x % TX, t.y * TY + (int)threadIdx.x / TX);
  x.apply_flux(c.g, f);
}

// Batched Anderson steps (FP64): the single-context kernels' code per tile
// of a case, with the case's vectors and the shared ring slots.
__device__ __forceinline__ AASel case_sel(const BatchCase<double> &c, const AABatchSel &bs) {
  AASel sel;
  sel.n = bs.n;
  sel.keep_dF = bs.keep_dF;
  sel.start = bs.start;
  sel.write_D = bs.write_D;
#pragma unroll
  for (int j = 0; j < MAX_AA; j++) {
    sel.slot[j] = bs.slot[j];
    sel.dG[j] = j < bs.n ? c.aa_dG[bs.slot[j]] : nullptr;
    sel.dF[j] = j < bs.n ? c.aa_dF[bs.slot[j]] : nullptr;
  }
  return sel;
}

__device__ __forceinline__ AAFields case_state(const BatchCase<double> &c, int par) {
  const KFields f = fields_at(c, par);
  return AAFields{{f.f[STS_U], f.f[STS_V], f.f[STS_P], f.f[STS_T]}};
}

template <int TX, int TY>
__global__ void __launch_bounds__(TX *TY) kb_aa_begin(const BatchCase<double> *C, const BatchTile *T) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<double> &c = C[t.c];
  const int i = t.x * TX + (int)threadIdx.x % TX, j = t.y * TY + (int)threadIdx.x / TX;
  if (i >= c.g.Nx || j >= c.g.Ny) return;
  const AAFields s = case_state(c, 0);
  const int k = i + j * c.g.Nx;
  const double g4[4] = {s.x[0][k], s.x[1][k], s.x[2][k], s.x[3][k]}, z[4] = {0.0, 0.0, 0.0, 0.0};
  store4(c.aa.gp + 4 * (size_t)k, g4);
  store4(c.aa.D + 4 * (size_t)k, z);
}

// Partial slot of a tile: its position in the case's row-major tile grid.
template <int TX, int TY, int NM>
__global__ void __launch_bounds__(TX *TY) kb_aa_store(const BatchCase<double> *C, const BatchTile *T, int k, int par,
                                                      AABatchSel bs) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<double> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k + 1})) return;
  const AASel sel = case_sel(c, bs);
  const int i = t.x * TX + (int)threadIdx.x % TX, j = t.y * TY + (int)threadIdx.x / TX;
  double acc[2 * NM];
#pragma unroll
  for (int q = 0; q < 2 * NM; q++) acc[q] = 0.0;
  if (i < c.g.Nx && j < c.g.Ny) aa_store_cell<NM>(case_state(c, par), c.aa, sel, i + j * c.g.Nx, acc);
  if (bs.n > 0) aa_store_sums<NM, TX * TY>(acc, c.aa, t.y * ((c.g.Nx + TX - 1) / TX) + t.x, sel.n);
}

// One workgroup per case.
template <int NM>
__global__ void __launch_bounds__(AA_SOLVE_THREADS) kb_aa_solve(const BatchCase<double> *C, int k, AABatchSel bs) {
  const BatchCase<double> &c = C[blockIdx.x];
  if (gated_out(Gate{c.ctrl, k + 1})) return;
  __shared__ double gam[NM];
  aa_solve_group<NM, AA_SOLVE_THREADS>(c.aa, case_sel(c, bs), gam, true);
}

// The mix without the flux part (the batched flux and diffusion pass follows).
template <int TX, int TY, int NM>
__global__ void __launch_bounds__(TX *TY) kb_aa_mix(const BatchCase<double> *C, const BatchTile *T, int k, int par,
                                                    AABatchSel bs) {
  const BatchTile t = T[blockIdx.x];
  const BatchCase<double> &c = C[t.c];
  if (gated_out(Gate{c.ctrl, k + 1})) return;
  __shared__ double sp[TY + 1][TX + 1], sT[TY + 1][TX + 1];
  aa_mix_tile<NM, TX, TY, false>(c.g, fields_at(c, par), case_state(c, par), c.aa, case_sel(c, bs), c.aa.gam,
                                 t.x * TX, t.y * TY, (int)threadIdx.x % TX, (int)threadIdx.x / TX, sp, sT, nullptr);
}

// Edge conditions and the loop-2 test: one workgroup per case.
template <typename R>
__global__ void __launch_bounds__(EDGE_THREADS) kb_edges(const BatchCase<R> *C, int k, int par, int parts) {
  const BatchCase<R> &c = C[blockIdx.x / EDGE_GROUPS];
  const Gate gate{c.ctrl, k};
  if (gated_out(gate)) return;
  const KFieldsT<R> f = fields_at(c, par);
  const LoopTest lt{c.ctrl, c.lred, c.log + 6 * (size_t)k, c.N_I, c.tol_uv, c.tol_p, c.tol_T, c.stall_w,
                    c.stall_ratio};
  edges_block(c.g, f, parts, c.probe, gate, lt, (int)(blockIdx.x % EDGE_GROUPS));
}

// Opens every case's loop 2: control, log (6 N_I words), slotted words.
template <typename R>
__global__ void __launch_bounds__(256) kb_loop_open(const BatchCase<R> *C) {
  const BatchCase<R> &c = C[blockIdx.x];
  if (threadIdx.x == 0) {
    c.ctrl->stop_at = c.skip ? -1 : 0x7fffffff;
    c.ctrl->nonfinite = 0;
  }
  for (int q = threadIdx.x; q < 6 * c.N_I; q += blockDim.x) c.log[q] = 0ull;
  for (int q = threadIdx.x; q < NSLOT * SLOT_STRIDE; q += blockDim.x) c.lred[q] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch wrappers used by the host API (sts_api.cpp)
// ---------------------------------------------------------------------------
#include "sts_launch.h"

template <typename R>
static inline dim3 grid2d(const KGridT<R> &g) {
  return dim3((g.Nx + BX - 1) / BX, (g.Ny + BY - 1) / BY);
}
// 1-D grid of TX x TY tiles (kernels that place their tiles with tile_origin)
template <int TX, int TY, typename R>
static inline dim3 grid_tiles(const KGridT<R> &g) {
  return dim3(((g.Nx + TX - 1) / TX) * ((g.Ny + TY - 1) / TY));
}

template <typename R>
hipError_t launch_assemble(const KGridT<R> &g, const KFieldsT<R> &f, hipStream_t s, Gate gate, bool inline_diff) {
  if (inline_diff && g.rht_one)
    hipLaunchKernelGGL((k_assemble_inl<R, true>), (grid_tiles<AX, AY>(g)), dim3(AX, AY), 0, s, g, f, gate);
  else if (inline_diff)
    hipLaunchKernelGGL((k_assemble_inl<R, false>), (grid_tiles<AX, AY>(g)), dim3(AX, AY), 0, s, g, f, gate);
  else hipLaunchKernelGGL(k_assemble<R>, (grid_tiles<BX, BY>(g)), dim3(BX, BY), 0, s, g, f, gate);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_sweep(const KGridT<R> &g, const KFieldsT<R> &f, const R *Po, const R *To, R *Pn, R *Tn,
                        unsigned long long *rp, unsigned long long *rT, hipStream_t s, unsigned long long *rp2,
                        unsigned long long *rT2, Gate gate) {
  hipLaunchKernelGGL(k_sweep<R>, (grid_tiles<BX, BY>(g)), dim3(BX, BY), 0, s, g, f, Po, To, Pn, Tn, rp, rT, rp2, rT2, gate);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_sweep2(const KGridT<R> &g, const KFieldsT<R> &f, const R *Po, const R *To, R *Pn, R *Tn,
                         unsigned long long *r1p, unsigned long long *r1T, unsigned long long *r2p,
                         unsigned long long *r2T, hipStream_t s, Gate gate) {
  hipLaunchKernelGGL(k_sweep2<R>, (grid_tiles<S2X, S2Y>(g)), dim3(S2X, S2Y), 0, s, g, f, Po, To, Pn, Tn, r1p, r1T, r2p, r2T, gate);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_correct(const KGridT<R> &g, const KFieldsT<R> &f, unsigned long long *ru, unsigned long long *rv,
                          int do_correct, hipStream_t s, Gate gate) {
  hipLaunchKernelGGL(k_correct<R>, (grid_tiles<BX, BY * CORRECT_ROWS>(g)), dim3(BX, BY), 0, s, g, f, ru, rv, do_correct,
                     gate);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_wall(const KGridT<R> &g, const KFieldsT<R> &f, int parts, hipStream_t s, Gate gate,
                       const WallOp *ops, int nops) {
  if (ops) {
    if (nops > 0)
      hipLaunchKernelGGL(k_wall_ops<R>, dim3((nops + 255) / 256), dim3(256), 0, s, g, f, ops, nops, parts, gate);
  } else {
    hipLaunchKernelGGL(k_wall<R>, (grid_tiles<BX, BY>(g)), dim3(BX, BY), 0, s, g, f, parts, gate);
  }
  return hipGetLastError();
}
template <typename R>
hipError_t launch_edges(const KGridT<R> &g, const KFieldsT<R> &f, int parts, double *probe, hipStream_t s,
                        Gate gate, LoopTest lt) {
  hipLaunchKernelGGL(k_edges<R>, dim3(lt.ctrl ? EDGE_GROUPS : EDGE_GROUPS - 1), dim3(EDGE_THREADS), 0, s, g, f, parts,
                     probe, gate, lt);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_flux_diff(const KGridT<R> &g, const KFieldsT<R> &f, hipStream_t s, Gate gate, int parts) {
  if (parts == FLUX_PART)
    hipLaunchKernelGGL((k_flux_diff<R, FLUX_PART>), (grid_tiles<BX, BY * FLUX_ROWS>(g)), dim3(BX, BY), 0, s, g, f,
                       gate);
  else if (parts == DIFF_PART)
    hipLaunchKernelGGL((k_flux_diff<R, DIFF_PART>), (grid_tiles<BX, BY * FLUX_ROWS>(g)), dim3(BX, BY), 0, s, g, f,
                       gate);
  else
    hipLaunchKernelGGL((k_flux_diff<R, FLUX_PART | DIFF_PART>), (grid_tiles<BX, BY * FLUX_ROWS>(g)), dim3(BX, BY), 0,
                       s, g, f, gate);
  return hipGetLastError();
}

#define STS_INSTANTIATE(R)                                                                                     \
  template hipError_t launch_assemble<R>(const KGridT<R> &, const KFieldsT<R> &, hipStream_t, Gate, bool);     \
  template hipError_t launch_sweep<R>(const KGridT<R> &, const KFieldsT<R> &, const R *, const R *, R *, R *,  \
                                      unsigned long long *, unsigned long long *, hipStream_t,                 \
                                      unsigned long long *, unsigned long long *, Gate);                       \
  template hipError_t launch_sweep2<R>(const KGridT<R> &, const KFieldsT<R> &, const R *, const R *, R *, R *, \
                                       unsigned long long *, unsigned long long *, unsigned long long *,       \
                                       unsigned long long *, hipStream_t, Gate);                               \
  template hipError_t launch_correct<R>(const KGridT<R> &, const KFieldsT<R> &, unsigned long long *,          \
                                        unsigned long long *, int, hipStream_t, Gate);                         \
  template hipError_t launch_wall<R>(const KGridT<R> &, const KFieldsT<R> &, int, hipStream_t, Gate,           \
                                     const WallOp *, int);                                                    \
  template hipError_t launch_edges<R>(const KGridT<R> &, const KFieldsT<R> &, int, double *, hipStream_t, Gate, \
                                      LoopTest);                                                               \
  template hipError_t launch_flux_diff<R>(const KGridT<R> &, const KFieldsT<R> &, hipStream_t, Gate, int);
STS_INSTANTIATE(double)
STS_INSTANTIATE(float)
#undef STS_INSTANTIATE

hipError_t launch_properties(const KGrid &g, const KFields &f, hipStream_t s) {
  hipLaunchKernelGGL(k_properties<double>, grid2d(g), dim3(BX, BY), 0, s, g, f);
  return hipGetLastError();
}
hipError_t launch_rho_mid(const KGrid &g, const KFields &f, hipStream_t s) {
  hipLaunchKernelGGL(k_rho_mid<double>, grid2d(g), dim3(BX, BY), 0, s, g, f);
  return hipGetLastError();
}
hipError_t launch_apply_bodies(const KGrid &g, const KFields &f, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_bodies, grid2d(g), dim3(BX, BY), 0, s, g, f);
  return hipGetLastError();
}
hipError_t launch_jump_overwrite(const KGrid &g, const KFields &f, hipStream_t s) {
  hipLaunchKernelGGL(k_jump_overwrite, grid2d(g), dim3(BX, BY), 0, s, g, f);
  return hipGetLastError();
}
hipError_t launch_copy5(const double *const *src, double *const *dst, int n, hipStream_t s) {
  const int blocks = (n + BX * BY - 1) / (BX * BY);
  dim3 gr(blocks < 65535 ? blocks : 65535, (blocks + 65534) / 65535);
  hipLaunchKernelGGL(k_copy5, gr, dim3(BX, BY), 0, s, src[0], src[1], src[2], src[3], src[4], dst[0], dst[1], dst[2],
                     dst[3], dst[4], n);
  return hipGetLastError();
}
hipError_t launch_extrapolate(const KGrid &g, const KFields &f, const ExtrapHist &hist, int order, int slot,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_extrapolate, grid2d(g), dim3(BX, BY), 0, s, g, f, hist, order, slot);
  return hipGetLastError();
}
int aa_store_blocks(int Na) { return (Na + AA_THREADS * AA_STORE_CELLS - 1) / (AA_THREADS * AA_STORE_CELLS); }
template <typename R>
hipError_t launch_aa_begin(const AAFieldsT<R> &f, const AABuf &ab, int Na, hipStream_t s) {
  hipLaunchKernelGGL(k_aa_begin<R>, dim3((Na + AA_THREADS - 1) / AA_THREADS), dim3(AA_THREADS), 0, s, f, ab, Na);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_aa_store(const AAFieldsT<R> &f, const AABuf &ab, const AASel &sel, int Na, hipStream_t s,
                           Gate gate) {
  if (sel.n < 0 || sel.n > MAX_AA || ab.nb != aa_store_blocks(Na)) return hipErrorInvalidValue;
  // one difference (history depth 1, or the loop's first steps): the
  // one-difference instantiation holds far fewer registers, same sums
  if (sel.n <= 1) hipLaunchKernelGGL((k_aa_store<1, R>), dim3(ab.nb), dim3(AA_THREADS), 0, s, f, ab, sel, Na, gate);
  else if (sel.n <= 2) hipLaunchKernelGGL((k_aa_store<2, R>), dim3(ab.nb), dim3(AA_THREADS), 0, s, f, ab, sel, Na, gate);
  else if (sel.n <= 4) hipLaunchKernelGGL((k_aa_store<4, R>), dim3(ab.nb), dim3(AA_THREADS), 0, s, f, ab, sel, Na, gate);
  else hipLaunchKernelGGL((k_aa_store<MAX_AA, R>), dim3(ab.nb), dim3(AA_THREADS), 0, s, f, ab, sel, Na, gate);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_aa_mix(const KGridT<R> &g, const KFieldsT<R> &fl, const AAFieldsT<R> &f, const AABuf &ab,
                         const AASel &sel, hipStream_t s, Gate gate) {
  const int Na = g.Nx * g.Ny;
  if (sel.n < 1 || sel.n > MAX_AA || ab.nb != aa_store_blocks(Na)) return hipErrorInvalidValue;
  const int ntiles = ((g.Nx + AA_MIX_TX - 1) / AA_MIX_TX) * ((g.Ny + AA_MIX_TY - 1) / AA_MIX_TY);
  const dim3 gm((ntiles + AA_MIX_TILES - 1) / AA_MIX_TILES);
  if (sel.n <= 1)
    hipLaunchKernelGGL((k_aa_mix<1, R>), gm, dim3(AA_MIX_TX * AA_MIX_TY), 0, s, g, fl, f, ab, sel, gate);
  else if (sel.n <= 2)
    hipLaunchKernelGGL((k_aa_mix<2, R>), gm, dim3(AA_MIX_TX * AA_MIX_TY), 0, s, g, fl, f, ab, sel, gate);
  else if (sel.n <= 4)
    hipLaunchKernelGGL((k_aa_mix<4, R>), gm, dim3(AA_MIX_TX * AA_MIX_TY), 0, s, g, fl, f, ab, sel, gate);
  else hipLaunchKernelGGL((k_aa_mix<MAX_AA, R>), gm, dim3(AA_MIX_TX * AA_MIX_TY), 0, s, g, fl, f, ab, sel, gate);
  return hipGetLastError();
}
template hipError_t launch_aa_begin<double>(const AAFields &, const AABuf &, int, hipStream_t);
template hipError_t launch_aa_begin<float>(const AAFieldsT<float> &, const AABuf &, int, hipStream_t);
template hipError_t launch_aa_store<double>(const AAFields &, const AABuf &, const AASel &, int, hipStream_t, Gate);
template hipError_t launch_aa_store<float>(const AAFieldsT<float> &, const AABuf &, const AASel &, int, hipStream_t,
                                           Gate);
template hipError_t launch_aa_mix<double>(const KGrid &, const KFields &, const AAFields &, const AABuf &,
                                          const AASel &, hipStream_t, Gate);
template hipError_t launch_aa_mix<float>(const KGridT<float> &, const KFieldsT<float> &, const AAFieldsT<float> &,
                                         const AABuf &, const AASel &, hipStream_t, Gate);
hipError_t launch_set_inflow(const KGrid &g, const KFields &f, double u_in, hipStream_t s) {
  dim3 gr((g.Nx_u0 + BX - 1) / BX > 0 ? (g.Nx_u0 + BX - 1) / BX : 1, (g.Ny + BY - 1) / BY);
  hipLaunchKernelGGL(k_set_inflow, gr, dim3(BX, BY), 0, s, g, f, u_in);
  return hipGetLastError();
}
hipError_t launch_loop_open(LoopCtrl *ctrl, unsigned long long *log, int n, unsigned long long *lred,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_loop_open, dim3(16), dim3(256), 0, s, ctrl, log, n, lred);
  return hipGetLastError();
}
template <typename S, typename D>
hipError_t launch_convert(const KFieldsT<S> &src, const KFieldsT<D> &dst, unsigned long long mask, int n,
                          hipStream_t s) {
  hipLaunchKernelGGL((k_convert<S, D>), dim3((n + 255) / 256), dim3(256), 0, s, src, dst, mask, n);
  return hipGetLastError();
}
template hipError_t launch_convert<double, float>(const KFieldsT<double> &, const KFieldsT<float> &,
                                                  unsigned long long, int, hipStream_t);
template hipError_t launch_convert<float, double>(const KFieldsT<float> &, const KFieldsT<double> &,
                                                  unsigned long long, int, hipStream_t);
hipError_t launch_drag(const KGrid &g, const KFields &f, const unsigned *cells, int n, double *terms, double *out,
                       hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_drag_terms, dim3((n + 255) / 256), dim3(256), 0, s, g, f, cells, n, terms);
  hipLaunchKernelGGL(k_drag_sum, dim3(1), dim3(64), 0, s, terms, n, out);
  return hipGetLastError();
}
hipError_t launch_wall_plan(const KGrid &g, const unsigned *cells, int n, WallOp *out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_wall_plan, dim3((n + 255) / 256), dim3(256), 0, s, g, cells, n, out);
  return hipGetLastError();
}

hipError_t launch_fold(unsigned long long *slots, int n, unsigned long long *out, hipStream_t s) {
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(128), 0, s, slots, n, out);
  return hipGetLastError();
}
int continuity_blocks(const KGrid &g) {
  dim3 gr = grid2d(g);
  return (int)(gr.x * gr.y);
}
hipError_t launch_continuity(const KGrid &g, const KFields &f, double *cs, double *cf, double *partial,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_fluxes, grid2d(g), dim3(BX, BY), 0, s, g, f);
  hipLaunchKernelGGL(k_continuity, grid2d(g), dim3(BX, BY), 0, s, g, f, cs, cf, partial);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched launches; tile shape variants (the autotuner's choices)
// ---------------------------------------------------------------------------
const int CELL_TILES[NVARIANT][2] = {{64, 4}, {64, 8}, {128, 2}, {32, 8}};
const int SWEEP_TILES[NVARIANT][2] = {{64, 8}, {64, 4}, {128, 4}, {64, 16}};

#define CELL_SWITCH(v, KERNEL, ...)                                                                   \
  switch (v) {                                                                                        \
    case 0: hipLaunchKernelGGL((KERNEL<R, 64, 4>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break;  \
    case 1: hipLaunchKernelGGL((KERNEL<R, 64, 8>), dim3(ntiles), dim3(512), 0, s, __VA_ARGS__); break;  \
    case 2: hipLaunchKernelGGL((KERNEL<R, 128, 2>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL((KERNEL<R, 32, 8>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break; \
  }

#define AA_SWITCH(v, NM, KERNEL, ...)                                                                 \
  switch (v) {                                                                                        \
    case 0: hipLaunchKernelGGL((KERNEL<64, 4, NM>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break;  \
    case 1: hipLaunchKernelGGL((KERNEL<64, 8, NM>), dim3(ntiles), dim3(512), 0, s, __VA_ARGS__); break;  \
    case 2: hipLaunchKernelGGL((KERNEL<128, 2, NM>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL((KERNEL<32, 8, NM>), dim3(ntiles), dim3(256), 0, s, __VA_ARGS__); break; \
  }

int aa_case_tiles(int Nx, int Ny, int variant) {
  const int TX = CELL_TILES[variant][0], TY = CELL_TILES[variant][1];
  return ((Nx + TX - 1) / TX) * ((Ny + TY - 1) / TY);
}

hipError_t launch_batch_aa(int kernel, int variant, const BatchCase<double> *C, int ncases, const BatchTile *T,
                           int ntiles, int k, int par, const AABatchSel &bs, hipStream_t s) {
  if (bs.n < 0 || bs.n > MAX_AA || variant < 0 || variant >= NVARIANT) return hipErrorInvalidValue;
  if (kernel == BKA_SOLVE) {
    // the instantiation for the differences in use: the same sums, fewer registers
    if (bs.n <= 2) hipLaunchKernelGGL(kb_aa_solve<2>, dim3(ncases), dim3(AA_SOLVE_THREADS), 0, s, C, k, bs);
    else if (bs.n <= 4) hipLaunchKernelGGL(kb_aa_solve<4>, dim3(ncases), dim3(AA_SOLVE_THREADS), 0, s, C, k, bs);
    else hipLaunchKernelGGL(kb_aa_solve<MAX_AA>, dim3(ncases), dim3(AA_SOLVE_THREADS), 0, s, C, k, bs);
    return hipGetLastError();
  }
  if (ntiles == 0) return hipSuccess;
  if (kernel == BKA_BEGIN) {
    switch (variant) {
      case 0: hipLaunchKernelGGL((kb_aa_begin<64, 4>), dim3(ntiles), dim3(256), 0, s, C, T); break;
      case 1: hipLaunchKernelGGL((kb_aa_begin<64, 8>), dim3(ntiles), dim3(512), 0, s, C, T); break;
      case 2: hipLaunchKernelGGL((kb_aa_begin<128, 2>), dim3(ntiles), dim3(256), 0, s, C, T); break;
      default: hipLaunchKernelGGL((kb_aa_begin<32, 8>), dim3(ntiles), dim3(256), 0, s, C, T); break;
    }
  } else if (kernel == BKA_STORE) {
    if (bs.n <= 2) { AA_SWITCH(variant, 2, kb_aa_store, C, T, k, par, bs); }
    else if (bs.n <= 4) { AA_SWITCH(variant, 4, kb_aa_store, C, T, k, par, bs); }
    else { AA_SWITCH(variant, MAX_AA, kb_aa_store, C, T, k, par, bs); }
  } else if (kernel == BKA_MIX) {
    if (bs.n <= 2) { AA_SWITCH(variant, 2, kb_aa_mix, C, T, k, par, bs); }
    else if (bs.n <= 4) { AA_SWITCH(variant, 4, kb_aa_mix, C, T, k, par, bs); }
    else { AA_SWITCH(variant, MAX_AA, kb_aa_mix, C, T, k, par, bs); }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_batch(int kernel, int variant, const BatchCase<R> *C, int ncases, const BatchTile *T, int ntiles,
                        int k, int par, int a, int b, hipStream_t s) {
  if (kernel != BK_EDGES && kernel != BK_LOOP_OPEN && ntiles == 0) return hipSuccess;
  switch (kernel) {
    case BK_ASSEMBLE: CELL_SWITCH(variant, kb_assemble, C, T, k, par); break;
    case BK_ASSEMBLE_ARR: CELL_SWITCH(variant, kb_assemble_arr, C, T, k, par); break;
    case BK_SWEEP: CELL_SWITCH(variant, kb_sweep, C, T, k, par, a, b); break;
    case BK_CORRECT: CELL_SWITCH(variant, kb_correct, C, T, k, par); break;
    case BK_WALL: CELL_SWITCH(variant, kb_wall, C, T, k, par, a); break;
    case BK_WALL_OPS:
      if (ntiles > 0)
        hipLaunchKernelGGL(kb_wall_ops<R>, dim3(ntiles), dim3(WALL_OPS_PER_TILE), 0, s, C, T, k, par, a);
      break;
    case BK_FLUX_DIFF: CELL_SWITCH(variant, kb_flux_diff, C, T, k, par); break;
    case BK_SWEEP2:
      switch (variant) {
        case 0: hipLaunchKernelGGL((kb_sweep2<R, 64, 8>), dim3(ntiles), dim3(512), 0, s, C, T, k, par, a, b); break;
        case 1: hipLaunchKernelGGL((kb_sweep2<R, 64, 4>), dim3(ntiles), dim3(256), 0, s, C, T, k, par, a, b); break;
        case 2: hipLaunchKernelGGL((kb_sweep2<R, 128, 4>), dim3(ntiles), dim3(512), 0, s, C, T, k, par, a, b); break;
        default:
          hipLaunchKernelGGL((kb_sweep2<R, 64, 16>), dim3(ntiles), dim3(1024), 0, s, C, T, k, par, a, b);
          break;
      }
      break;
    case BK_EDGES:
      hipLaunchKernelGGL(kb_edges<R>, dim3(EDGE_GROUPS * ncases), dim3(EDGE_THREADS), 0, s, C, k, par, a);
      break;
    case BK_LOOP_OPEN: hipLaunchKernelGGL(kb_loop_open<R>, dim3(ncases), dim3(256), 0, s, C); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#undef CELL_SWITCH
template hipError_t launch_batch<double>(int, int, const BatchCase<double> *, int, const BatchTile *, int, int, int,
                                         int, int, hipStream_t);
template hipError_t launch_batch<float>(int, int, const BatchCase<float> *, int, const BatchTile *, int, int, int,
                                        int, int, hipStream_t);
