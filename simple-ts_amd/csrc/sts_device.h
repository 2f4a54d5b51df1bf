// Device-side data structures and per-cell helpers shared by the SIMPLE-TS
// kernels.  Index semantics follow the reference's ij_function macro
// (SIMPLE_TS_MPI.cpp:140-153): neighbour indices are clamped at the domain
// edge, so a cell on the edge sees itself as its missing neighbour.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "simple_ts.h"

struct DevBody {
  double u_body, v_body, p_body, T_body, F_slip, F_jump, dTdn;
  int is_slip, is_jump;
};

// Edge boundary-condition switches, one bit each (EnteredData.txt:81-109).
enum : unsigned {
  BC_DUDX_XB = 1u << 0, BC_DURHODX_XB = 1u << 1, BC_DRHODT_XB = 1u << 2,
  BC_DUDX_XE = 1u << 3, BC_DURHODX_XE = 1u << 4, BC_DRHODT_XE = 1u << 5,
  BC_UMINMOUT_XE = 1u << 6, BC_DUDY_YB = 1u << 7, BC_DUDY_YE = 1u << 8,
  BC_DVDX_XB = 1u << 9, BC_DVDX_XE = 1u << 10, BC_DVDY_YB = 1u << 11, BC_DVDY_YE = 1u << 12,
  BC_DPDX_XB = 1u << 13, BC_DPDX_XE = 1u << 14, BC_DPDY_YB = 1u << 15, BC_DPDY_YE = 1u << 16,
  BC_DTDX_XB = 1u << 17, BC_DTDX_XE = 1u << 18, BC_DTDY_YB = 1u << 19, BC_DTDY_YE = 1u << 20,
};

// Grid, scalars and switches of one case; R is the arithmetic type of the
// kernels that take it (double, or float for the FP32 phase of the mixed
// precision time step).
// Fluid bits of the cells the stencils read around a cell, one 16-bit word
// per cell formed once from the (static) mask with Nbr's neighbour indices
// (clamped at the grid edge, x wrapping when periodic): the cell itself, west,
// east, west-west (i_2), south, north, the four diagonals, and the v
// equation's second south row (ij_1 - Nx*(0 <= j-2), S:3399).  One load
// replaces the byte loads of every neighbour's mask.
enum : unsigned {
  NB_C = 1u << 0, NB_W = 1u << 1, NB_E = 1u << 2, NB_WW = 1u << 3, NB_S = 1u << 4, NB_N = 1u << 5,
  NB_SW = 1u << 6, NB_SE = 1u << 7, NB_NW = 1u << 8, NB_NE = 1u << 9, NB_SS = 1u << 10,
};

template <typename R>
struct KGridT {
  int Nx, Ny;
  const R *hx, *hy;
  const uint8_t *m;       // vol_inf_fb_bool
  const uint16_t *nb;     // neighbour fluid bits (NB_*)
  const uint32_t *vol;    // vol_inf_fb
  const DevBody *bodies;
  int nbodies;
  int Nx_u0, Nx_u1, Ny_u0, Ny_u1, Nx_v0, Nx_v1, Ny_v0, Ny_v1, Nx_o0, Nx_o1, Ny_o0, Ny_o1;
  R ht, gamma1, Kn, cT, cT1, cT2, cPch, c_Fr_v;
  R rht;      // 1/ht rounded to nearest in R
  R c1;       // 1 - gamma1 in R, formed on the host (a uniform kernel argument, no vector register)
  int rht_one;  // |1 - rht ht| <= 2^-(p+1): a rht is within an ulp of a/ht for every a (quot1)
  int given_xb, given_xe, kn_local;
  unsigned bc;
  int px;  // periodic about OX: x neighbours wrap, box-edge conditions off
};
using KGrid = KGridT<double>;

template <typename R>
struct KFieldsT {
  R *f[STS_NFIELDS];
};
using KFields = KFieldsT<double>;

// Device-side loop-2 control (S:4406-4450 evaluated on the GPU): the outer
// iterations of one time step are enqueued ahead; iteration k runs only while
// k <= stop_at, and the edge kernel of iteration k ends the loop by setting
// stop_at = k when the reference's test says so.
struct LoopCtrl {
  int stop_at;    // last outer iteration (0-based) to run, INT_MAX while open
  int nonfinite;  // a residual was not finite
};

struct Gate {
  const LoopCtrl *ctrl;  // nullptr: always run
  int k;                 // outer-iteration index of this launch
};

__device__ __forceinline__ bool gated_out(const Gate &g) { return g.ctrl && g.k > g.ctrl->stop_at; }

// Residual maxima are accumulated with one atomicMax per workgroup into one
// of NSLOT copies of each destination (copy = workgroup index mod NSLOT,
// copies SLOT_STRIDE words apart), so no address takes more than a few dozen
// atomics per launch; the reader folds the copies.  A destination pointer
// passed to a kernel is the address of copy 0.
constexpr int NSLOT = 64, SLOT_STRIDE = 128;

// Loop-2 test of iteration k (S:4433-4437).  The kernels of iteration k leave
// their maxima in the slotted words lred[0..5]: max|du|, max|dv|, max|dp|,
// max|dT| (last sweep), first-sweep dp, dT.  The test folds them into log_k,
// clears lred for the next iteration and decides.
struct LoopTest {
  LoopCtrl *ctrl;
  unsigned long long *lred;
  unsigned long long *log_k;
  int N_I;
  double tol_uv, tol_p, tol_T;
  int stall_w;         // > 0: also stop when the residuals stagnate (see k_edges)
  double stall_ratio;
};

// Clamped neighbour indices of cell (i, j).
// Neighbour indices of cell (i, j): clamped at the grid edge (ij_function,
// S:140-153), or, with px, wrapping in x (ij_function_PeriodicBC,
// S:201-232; j stays clamped).
struct Nbr {
  unsigned i, j, ij, i_1, i_2, i1, i_1j, i1j, ij_1, ij1, i_1j_1, i_1j1, i1j_1, i1j1, i_2j;
  __device__ __forceinline__ Nbr(unsigned ii, unsigned jj, unsigned Nx, unsigned Ny, bool px) {
    i = ii;
    j = jj;
    ij = ii + jj * Nx;
    if (px) {
      i_1 = (ii == 0u) ? Nx - 1u : ii - 1u;
      i_2 = (ii == 1u) ? Nx - 1u : (ii == 0u) ? Nx - 2u : ii - 2u;
      i1 = (ii == Nx - 1u) ? 0u : ii + 1u;
    } else {
      const unsigned l = (0u < ii);
      i_1 = ii - l;
      i_2 = ii - l - (1u < ii);
      i1 = ii + (ii < Nx - 1u);
    }
    const unsigned rs = (jj - (0u < jj)) * Nx, rn = (jj + (jj < Ny - 1u)) * Nx, rj = jj * Nx;
    i_1j = i_1 + rj;
    i1j = i1 + rj;
    ij_1 = ii + rs;
    ij1 = ii + rn;
    i_1j_1 = i_1 + rs;
    i_1j1 = i_1 + rn;
    i1j_1 = i1 + rs;
    i1j1 = i1 + rn;
    i_2j = i_2 + rj;
  }
};

template <typename R>
__device__ __forceinline__ R maxi0(R a) {  // maximum(0, a), S:121
  return (R(0.0) > a) ? R(0.0) : a;
}
template <typename R>
__device__ __forceinline__ R upwind_rho(R r0, R r1, R v1) {  // S:130
  return (v1 == R(0)) ? (R(0.5) * (r0 + r1)) : ((v1 > R(0)) ? r0 : r1);
}
template <typename R>
__device__ __forceinline__ R Li(R f1, R f2, R h1, R h2) {  // S:133
  return (f1 * h2 + f2 * h1) / (h1 + h2);
}
template <typename R>
__device__ __forceinline__ R Hi(R f1, R f2, R h1, R h2) {  // S:136
  return (h1 + h2) * f1 * f2 / (h1 * f2 + h2 * f1);
}

// One write of the wall pass.  Which faces are written, from which face or
// cell, with which body and mesh step, follows from the mask and the body
// table alone, which are fixed for a case; only the values change.
enum : unsigned { WOP_U = 0u, WOP_V = 1u, WOP_TWU = 2u, WOP_TWV = 3u, WOP_NONE = 7u };
struct WallOp {
  unsigned dst;  // face written: U, V, TW_U or TW_V [dst]
  unsigned src;  // slip: the fluid-side face read; wall temperature: the fluid cell
  unsigned h;    // hy row (U, TW_V) or hx column (V, TW_U) of the half cell
  unsigned kb;   // kind | body index << 3
};
constexpr int WALL_OPS_PER_TILE = 256;

// Parts of the flux and diffusion pass (S:10440-10618): fluxes and thermal
// diffusion coefficients; Gamma and the velocity diffusion coefficients.
enum : int { FLUX_PART = 1, DIFF_PART = 2 };

// Converged u, v, p, T of the last steps: f[h] = x^(n-1-h) (extrapolated starts).
constexpr int MAX_EXTRAP = 5;
struct ExtrapHist {
  double *f[MAX_EXTRAP][4];  // [h][u, v, p, T]
};

// Anderson acceleration of loop 2 (Walker and Ni, SIAM J. Numer. Anal. 49,
// 2011), in the difference form: the state is (u, v, p, T) of every cell, an
// outer iteration the map x -> g(x), f = g - x.  After an iteration the
// differences dF = f - f_prev, dG = g - g_prev of the last MAX_AA or fewer
// iterations give gamma = argmin |f - dF gamma|_2 and the next start
// x = g - dG gamma.  Every vector holds a cell's four values side by side
// (entry 4 k + q); the differences are stored in FP32.
constexpr int MAX_AA = 8;
// The differences in use, oldest first, newest last: ring slots (rows of the
// Gram matrix) and buffers.
enum { AA_START_D = 0, AA_START_GP = 1, AA_START_MIX = 2 };
struct AASel {
  int n;
  bool keep_dF;  // the newest dF is read by later stores (history depth > 1)
  // where the store finds the start x of the iterations it follows:
  // AA_START_D x = gp + D; AA_START_GP x = gp (no mix since the last store);
  // AA_START_MIX (history depth 1) x = the last mix of gp, formed again from
  // gp, gam[0] and dG[0] before the store overwrites dG[0]
  int start;
  bool write_D;  // the mix writes D (AA_START_D)
  int slot[MAX_AA];
  float *dG[MAX_AA];
  float *dF[MAX_AA];
};
// Per-context vectors: gp = the newest iterate g (FP64), D = the current
// start minus gp, fp = the newest f (both FP32); part: partial sums of the
// store pass (2 MAX_AA rows of nb workgroups); gram: dF_a . dF_b by ring slot;
// gam: the weights.
struct AABuf {
  double *gp;
  float *D, *fp;
  double *part, *gram, *gam;
  int nb;
};
// The four state arrays as the outer iteration leaves them (p and T in
// whichever ping-pong buffer holds them), of the loop's precision R (the
// acceleration's own vectors stay as above for both).
template <typename R>
struct AAFieldsT {
  R *x[4];  // u, v, p, T
};
using AAFields = AAFieldsT<double>;

// One case of a batch run in shared launches (device-resident): grid and
// fields with p and T as they stand at the start of the step, the p/T
// ping-pong partners, loop control, residual words and log, loop-2 test.
template <typename R>
struct BatchCase {
  KGridT<R> g;
  KFieldsT<R> f;
  R *p_alt, *T_alt;
  LoopCtrl *ctrl;
  unsigned long long *lred;  // slotted residual words
  unsigned long long *log;   // 6 words per outer iteration
  double *probe;
  int N_I, stall_w;
  double tol_uv, tol_p, tol_T, stall_ratio;
  const WallOp *wops;  // the case's planned wall writes
  int nwops;
  int skip;     // the case sits this loop out (its loop control opens closed)
  int inline0;  // the first outer iteration forms Gamma and D itself (else reads the arrays)
  // Anderson acceleration (FP64 loops): the case's vectors (aa.nb = its
  // tiles of the cell variant in use) and its ring of differences
  AABuf aa;
  float *aa_dG[MAX_AA], *aa_dF[MAX_AA];
};

// The differences in use in every case of a batched Anderson step (the cases
// step together, so their rings rotate alike): count and ring slots.
struct AABatchSel {
  int n;
  bool keep_dF;  // as AASel::keep_dF, start and write_D
  int start;
  bool write_D;
  int slot[MAX_AA];
};

// Work item of a batched launch: tile (x, y) of case c.
struct BatchTile {
  int c, x, y;
};
