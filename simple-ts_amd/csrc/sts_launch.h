// Launch wrappers of the kernels in sts_kernels.hip (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "sts_device.h"

// Loop-2 kernels, instantiated for double and float.
// inline_diff: form the diffusion coefficients from sqrt_T, the wall
// temperatures and the mask instead of reading Gamma_* and D_* (equal while
// those inputs are unchanged since the last flux and diffusion pass).
template <typename R>
hipError_t launch_assemble(const KGridT<R> &g, const KFieldsT<R> &f, hipStream_t s, Gate gate = Gate{nullptr, 0},
                           bool inline_diff = false);
template <typename R>
hipError_t launch_sweep(const KGridT<R> &g, const KFieldsT<R> &f, const R *Po, const R *To, R *Pn, R *Tn,
                        unsigned long long *rp, unsigned long long *rT, hipStream_t s,
                        unsigned long long *rp2 = nullptr, unsigned long long *rT2 = nullptr,
                        Gate gate = Gate{nullptr, 0});
template <typename R>
hipError_t launch_sweep2(const KGridT<R> &g, const KFieldsT<R> &f, const R *Po, const R *To, R *Pn, R *Tn,
                         unsigned long long *r1p, unsigned long long *r1T, unsigned long long *r2p,
                         unsigned long long *r2T, hipStream_t s, Gate gate = Gate{nullptr, 0});
template <typename R>
hipError_t launch_correct(const KGridT<R> &g, const KFieldsT<R> &f, unsigned long long *ru, unsigned long long *rv,
                          int do_correct, hipStream_t s, Gate gate = Gate{nullptr, 0});
// cells: list of the cells next to a body/fluid interface (or null: all cells)
template <typename R>
hipError_t launch_wall(const KGridT<R> &g, const KFieldsT<R> &f, int parts, hipStream_t s,
                       Gate gate = Gate{nullptr, 0}, const WallOp *ops = nullptr, int nops = 0);
// The planned writes of the listed cells: out[4 t + q], q = U, V, TW_U, TW_V.
hipError_t launch_wall_plan(const KGrid &g, const unsigned *cells, int n, WallOp *out, hipStream_t s);
template <typename R>
hipError_t launch_edges(const KGridT<R> &g, const KFieldsT<R> &f, int parts, double *probe, hipStream_t s,
                        Gate gate = Gate{nullptr, 0},
                        LoopTest lt = LoopTest{nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 0});
// parts: FLUX_PART (Fx, Fy, DTx, DTy) and/or DIFF_PART (Gamma_*, D_*).
template <typename R>
hipError_t launch_flux_diff(const KGridT<R> &g, const KFieldsT<R> &f, hipStream_t s, Gate gate = Gate{nullptr, 0},
                            int parts = FLUX_PART | DIFF_PART);

// FP64-only kernels.
hipError_t launch_properties(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_rho_mid(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_apply_bodies(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_jump_overwrite(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_copy5(const double *const *src, double *const *dst, int n, hipStream_t s);
hipError_t launch_set_inflow(const KGrid &g, const KFields &f, double u_in, hipStream_t s);
// u, v, p, T extrapolated from x^n (the *_PR fields) and hist[h] = x^(n-1-h)
// (h < order); x^n is stored into hist[slot]; order 0 only stores.
hipError_t launch_extrapolate(const KGrid &g, const KFields &f, const ExtrapHist &hist, int order, int slot,
                              hipStream_t s);
// Anderson acceleration of loop 2.  launch_aa_begin at the loop start;
// after an outer iteration's edge conditions, launch_aa_store (its
// differences and their partial dot products), then, when sel.n > 0,
// launch_aa_mix (solve for the weights, mix into the state arrays, with
// sqrt_T, rho and the face densities of the mixed state).
// ab.nb = aa_store_blocks(Na); ab.part: 2 MAX_AA * ab.nb doubles; ab.gram:
// MAX_AA^2; ab.gam: MAX_AA; ab.gp: 4 Na doubles; ab.D, ab.fp and every
// sel.dG/dF: 4 Na floats.
int aa_store_blocks(int Na);
// R: the loop's precision (double, or float in the FP32 phase of the mixed
// time step); the acceleration's vectors are the same for both.
template <typename R>
hipError_t launch_aa_begin(const AAFieldsT<R> &f, const AABuf &ab, int Na, hipStream_t s);
template <typename R>
hipError_t launch_aa_store(const AAFieldsT<R> &f, const AABuf &ab, const AASel &sel, int Na, hipStream_t s,
                           Gate gate);
template <typename R>
hipError_t launch_aa_mix(const KGridT<R> &g, const KFieldsT<R> &fl, const AAFieldsT<R> &f, const AABuf &ab,
                         const AASel &sel, hipStream_t s, Gate gate);
hipError_t launch_loop_open(LoopCtrl *ctrl, unsigned long long *log, int n, unsigned long long *lred,
                            hipStream_t s);
template <typename S, typename D>
hipError_t launch_convert(const KFieldsT<S> &src, const KFieldsT<D> &dst, unsigned long long mask, int n,
                          hipStream_t s);
hipError_t launch_fold(unsigned long long *slots, int n, unsigned long long *out, hipStream_t s);
hipError_t launch_drag(const KGrid &g, const KFields &f, const unsigned *cells, int n, double *terms, double *out,
                       hipStream_t s);
int continuity_blocks(const KGrid &g);
hipError_t launch_continuity(const KGrid &g, const KFields &f, double *cs, double *cf, double *partial,
                             hipStream_t s);

// Batched loop-2 launches over the tiles T[0..ntiles) of the cases
// C[0..ncases).  kernel: BK_*; variant: index into CELL_TILES (cell
// kernels) or SWEEP_TILES (BK_SWEEP2).  (a, b) = (first, last) for the
// sweeps, a = parts for BK_WALL and BK_EDGES.
// BK_ASSEMBLE forms Gamma and D in the tile; BK_ASSEMBLE_ARR (first outer
// iteration only) assembles the cases whose arrays are to be read instead.
enum {
  BK_ASSEMBLE,
  BK_SWEEP2,
  BK_SWEEP,
  BK_CORRECT,
  BK_WALL,
  BK_EDGES,
  BK_FLUX_DIFF,
  BK_LOOP_OPEN,
  BK_WALL_OPS,
  BK_ASSEMBLE_ARR
};
// BK_WALL_OPS: tile (c, x) = planned wall writes [x W, x W + W) of case c,
// W = WALL_OPS_PER_TILE.
constexpr int NVARIANT = 4;
extern const int CELL_TILES[NVARIANT][2];   // (TX, TY)
extern const int SWEEP_TILES[NVARIANT][2];  // (SX, SY)
template <typename R>
hipError_t launch_batch(int kernel, int variant, const BatchCase<R> *C, int ncases, const BatchTile *T, int ntiles,
                        int k, int par, int a, int b, hipStream_t s);

// Batched Anderson steps (FP64) over the cell tiles T of variant `variant`:
// BKA_BEGIN at the loop start (parity 0), then after outer iteration k's edge
// conditions BKA_STORE and, when bs.n > 0, BKA_SOLVE (one workgroup per
// case) and BKA_MIX, all gated per case on iteration k + 1.  Each case's
// aa.nb must be aa_case_tiles(Nx, Ny, variant) and its aa.part that many
// partial slots.
enum { BKA_BEGIN, BKA_STORE, BKA_SOLVE, BKA_MIX };
int aa_case_tiles(int Nx, int Ny, int variant);
hipError_t launch_batch_aa(int kernel, int variant, const BatchCase<double> *C, int ncases, const BatchTile *T,
                           int ntiles, int k, int par, const AABatchSel &bs, hipStream_t s);
