// HIP kernels of the 3x3-block Krylov solver for gfx950 (MI355X).
//
// Everything here is HBM-bandwidth bound (0.23 flop/byte in FP64), so:
//  * matrix data is read as contiguous wavefront segments (SELL-64 planes);
//  * each lane issues a whole chunk of block loads (values, columns, gathers)
//    before it accumulates, so one wavefront keeps many HBM requests in flight
//    (the per-colour SSOR launches have few wavefronts);
//  * workgroups are remapped so each XCD sweeps a contiguous run of slices and
//    x/z gathers hit that XCD's L2;
//  * scalars stay on the device (KrylovState); a reduction kernel writes one
//    partial per workgroup and a one-workgroup tail kernel sums them in a fixed
//    order and runs the Krylov scalar logic.  No grid-wide fences or atomics:
//    on gfx950 an agent-scope fence writes the XCD's L2 back, which made the
//    in-kernel "last block" pattern cost ~0.2 ms per reduction.
//
// FP64 row arithmetic is written with contraction off and in the reference's
// operation order, so SpMV, SSOR and DIAG results are bitwise those of the
// reference CPU loops (hecmw_solver_las_33.F90:281-308,
// hecmw_precond_SSOR_33.F90:302-405, hecmw_precond_DIAG_33.F90:115-140).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>

#include "device.h"

namespace hecmw33 {

namespace {

constexpr int32_t kErrNoconv = 3001;
constexpr int32_t kErrDivergeMat = 3002;
constexpr int32_t kErrDivergePc = 3003;
constexpr int32_t kErrDivergeNan = 3004;

__device__ __forceinline__ int xcd_block(int bid, int nblk) {
  // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous
  // run of logical blocks so neighbouring slices share that XCD's L2.
  const int xcd = bid % kNumXcd, local = bid / kNumXcd;
  const int per = nblk / kNumXcd, rem = nblk % kNumXcd;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-reduce K values and store the block's partial sums (fixed tree).
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], double* part, int slot) {
  __shared__ double red[K][kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double w = wave_sum(v[k]);
    if (lane == 0) red[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = red[k][0];
#pragma unroll
      for (int w = 1; w < kWavesPerBlock; ++w) s += red[k][w];
      part[k * kMaxPartials + slot] = s;
    }
  }
}

// Sum nparts partials of K components in a fixed order (one workgroup of
// kTailThreads): thread t adds partials t, t+T, t+2T, ... with eight loads in
// flight, then a fixed tree.
constexpr int kTailThreads = 1024;
template <int K>
__device__ __forceinline__ void sum_partials(const double* part, int nparts, double (&tot)[K]) {
  constexpr int kW = kTailThreads / kWave;
  __shared__ double red[K][kW];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double* p = part + k * kMaxPartials;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i0 = threadIdx.x; i0 < nparts; i0 += 8 * kTailThreads) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * kTailThreads;
        if (i < nparts) acc[u] += p[i];
      }
    }
    double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    s = wave_sum(s);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = red[k][0];
#pragma unroll
    for (int w = 1; w < kW; ++w) t += red[k][w];
    tot[k] = t;
  }
}

// ---- scalar logic of the Krylov loops ------------------------------------------
enum Tail : int {
  kTailStore = 0,      // out[0..K) = totals
  kTailCgRho = 1,      // hecmw_solver_CG.f90:179-207
  kTailCgC1 = 2,       // :221-232
  kTailDnrm2 = 3,      // :248-250, 267 (recurrence residual)
  kTailBiRho = 5,      // hecmw_solver_BiCGSTAB.f90:169-176
  kTailBiC2 = 6,       // :200-202
  kTailBiOmega = 7,    // :227-234
  kTailFcgBeta = 8,    // flexible CG: rho = r.z, beta = -(z.q_prev)/(p.q)_prev
  kTailFcgAlpha = 9,   // alpha = (r.p)/(p.q)
  kTailFcgRes = 10,    // recurrence |r|^2, reliable-update decision
};

template <int TAIL, int K>
__device__ void run_tail(KrylovState* st, const double (&tot)[K], double* out, double* hist) {
  if (TAIL == kTailStore) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = tot[k];
  } else if (TAIL == kTailCgRho) {
    const double rho = tot[0];
    st->rho = rho;
    if (rho == 0.0) {
      st->done = 1;
    } else if (st->iter > 1 && rho * st->rho1 <= 0.0) {
      st->n_indef += 1;
      if (st->n_indef >= 3) {
        st->error = kErrDivergePc;
        st->done = 1;
      }
    } else if (rho != rho) {
      st->error = kErrDivergeNan;
      st->done = 1;
    }
    if (!st->done && st->iter > 1) st->beta = rho / st->rho1;
  } else if (TAIL == kTailCgC1) {
    const double c1 = tot[0];
    st->c1 = c1;
    if (c1 <= 0.0) {
      st->error = kErrDivergeMat;
      st->done = 1;
    } else if (c1 != c1) {
      st->error = kErrDivergeNan;
      st->done = 1;
    } else {
      st->alpha = st->rho / c1;
    }
  } else if (TAIL == kTailDnrm2) {
    // iterations that recompute the true residual produce no recurrence partials
    if ((st->iter % st->nrec) == 0) return;
    st->dnrm2 = tot[0];
    st->resid = sqrt(tot[0] / st->bnrm2);
    if (hist) hist[st->iter - 1] = st->resid;
    if (st->resid <= st->tol) st->recheck = 1;
  } else if (TAIL == kTailBiRho) {
    st->rho = tot[0];
    if (st->iter > 1) st->beta = (st->rho / st->rho1) * (st->alpha / st->omega);
  } else if (TAIL == kTailBiC2) {
    st->c2 = tot[0];
    st->alpha = st->rho / tot[0];
  } else if (TAIL == kTailBiOmega) {
    st->ts = tot[0];
    st->tt = tot[K > 1 ? 1 : 0];
    st->omega = st->ts / st->tt;
  } else if (TAIL == kTailFcgBeta) {
    // Notay's FCG(1): beta keeps p A-orthogonal to the previous direction
    // whatever the (FP32) preconditioner did; r.z only checks positivity
    // with CG's rules.
    const double rho = tot[0];
    st->zq = tot[K > 1 ? 1 : 0];
    st->rho = rho;
    if (rho == 0.0) {
      st->done = 1;
    } else if (rho != rho) {
      st->error = kErrDivergeNan;
      st->done = 1;
    } else if (st->iter > 1 && rho * st->rho1 <= 0.0) {
      st->n_indef += 1;
      if (st->n_indef >= 3) {
        st->error = kErrDivergePc;
        st->done = 1;
      }
    }
    if (!st->done) st->beta = st->iter > 1 ? -st->zq / st->pq : 0.0;
  } else if (TAIL == kTailFcgAlpha) {
    const double pq = tot[0];
    st->pq = pq;
    st->rp = tot[K > 1 ? 1 : 0];
    if (pq <= 0.0) {
      st->error = kErrDivergeMat;
      st->done = 1;
    } else if (pq != pq || st->rp != st->rp) {
      st->error = kErrDivergeNan;
      st->done = 1;
    } else {
      st->alpha = st->rp / pq;
      st->edrop += st->alpha * st->rp;
    }
  } else if (TAIL == kTailFcgRes) {
    st->dnrm2 = tot[0];
    st->resid = sqrt(tot[0] / st->bnrm2);
    if (hist) hist[st->iter - 1] = st->resid;
    // replace the FP32 recurrence residual by b - A x in FP64 when it has
    // dropped by delta^2 in |r|^2, when it claims convergence, and every nrec
    // iterations
    st->reliable = st->resid * st->gap <= st->tol || tot[0] <= st->delta2 * st->rr_ref ||
                   (st->iter % st->nrec) == 0 || tot[0] != tot[0];
  }
}

__device__ __forceinline__ bool true_resid_active(const KrylovState* st) {
  return (st->iter % st->nrec) == 0 || st->recheck;
}

template <int TAIL, int K>
__global__ __launch_bounds__(kTailThreads) void k_tail(KrylovState* st, const double* part, int nparts, double* out,
                                                       double* hist) {
  if (TAIL != kTailStore && st->done) return;
  double tot[K];
  sum_partials<K>(part, nparts, tot);
  if (threadIdx.x == 0) run_tail<TAIL, K>(st, tot, out, hist);
}

// True residual bookkeeping and the end of one iteration
// (hecmw_solver_CG.f90:242-277): convergence after a recomputed residual,
// NOCONV at MAXIT (the Fortran DO variable then reads MAXIT+1), else
// rho1 = rho and the next iteration.
__global__ __launch_bounds__(kTailThreads) void k_tail_true(KrylovState* st, const double* part, int nparts,
                                                            double* hist) {
  if (st->done) return;
  if (true_resid_active(st)) {
    double tot[1];
    sum_partials<1>(part, nparts, tot);
    if (threadIdx.x == 0) {
      st->dnrm2 = tot[0];
      st->resid = sqrt(tot[0] / st->bnrm2);
      if ((st->iter % st->nrec) == 0 && hist) hist[st->iter - 1] = st->resid;
      if (st->resid <= st->tol) st->done = 1;
      st->recheck = 0;
    }
  }
  if (threadIdx.x == 0 && !st->done) {
    if (st->iter >= st->maxit) {
      st->error = kErrNoconv;
      st->done = 1;
      st->iter = st->maxit + 1;
    } else {
      st->rho1 = st->rho;
      st->iter += 1;
    }
  }
}

// Flexible CG: after a reliable update (r = b - A x in FP64) the true
// residual decides convergence and becomes the new reference norm; then the
// end of the iteration as in k_tail_true.
__global__ __launch_bounds__(kTailThreads) void k_tail_fcg_rel(KrylovState* st, const double* part, int nparts,
                                                               double* hist) {
  if (st->done) return;
  if (st->reliable) {
    double tot[1];
    sum_partials<1>(part, nparts, tot);
    if (threadIdx.x == 0) {
      // the learned two-level step is dropped for the rest of the solve once
      // the FP64 residual grows to drop2 (in |r|^2) times the smallest one seen
      // at a replacement (the residual of CG is not monotone, so a transient
      // rise is no failure; a divergence is): the iterations go on as the
      // plain SSOR flexible CG
      if (!st->learned_off && tot[0] > st->drop2 * st->rr_min) {
        st->learned_off = 1;
        st->coarse_off = 1;
        st->fallback_iter = st->iter;
      }
      st->rr_min = fmin(st->rr_min, tot[0]);
      const double rec = st->resid;
      st->dnrm2 = tot[0];
      st->rr_ref = tot[0];
      st->resid = sqrt(tot[0] / st->bnrm2);
      if (rec > 0.0) st->gap = fmin(1e3, fmax(1.0, st->resid / rec));
      if (hist) hist[st->iter - 1] = st->resid;
      st->nreliable += 1;
      st->reliable = 0;
      if (st->resid <= st->tol) {
        st->done = 1;
      } else if (st->resid != st->resid) {
        st->error = kErrDivergeNan;
        st->done = 1;
      }
    }
  }
  if (threadIdx.x == 0 && !st->done) {
    if (st->iter >= st->maxit) {
      st->error = kErrNoconv;
      st->done = 1;
      st->iter = st->maxit + 1;
    } else {
      st->rho1 = st->rho;
      st->iter += 1;
    }
  }
}

// ---- SELL-16 block row product (four lanes per row) -------------------------
// Lane 4r+d of a wavefront works on row r of the slice and output component
// d (the fourth lane of every quad mirrors d = 2).  The three component lanes
// of a row gather the three entries of a neighbour's x from one cache line
// and share them with DPP quad broadcasts, so every component's sum runs
// over the row's blocks in storage order exactly as the reference's loop.
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// DPP quad_perm broadcast of lane J of each group of four lanes
template <int J>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = J | (J << 2) | (J << 4) | (J << 6);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
template <int J>
__device__ __forceinline__ float quad_bcast(float v) {
  constexpr int ctrl = J | (J << 2) | (J << 4) | (J << 6);
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, false));
}

// SpMV rows, SELL-64, one row per lane (kRowsA).  FP64 in the reference's
// order: the D block first, then L, then U in storage order;
// y = ((y + a1*x1) + a2*x2) + a3*x3, no fused multiply-add.  A chunk of U
// blocks is loaded before it is accumulated in order; the D block is slot 0
// of the first chunk, so no load waits for a block of its own.  y starts at
// -0, for which -0 + a is a exactly (also for a = -0), so the D block's sum is
// the reference's (a1*x1 + a2*x2) + a3*x3.  xo = x of the row itself.
// One chunk of U blocks from k0.  FULL: every block is live, so no load or sum
// of the chunk sits under a condition and all its loads are in flight
// together (a per-block condition lets the compiler move a block's loads into
// a branch of their own, one more dependent round trip per chunk).  Not FULL
// (the row's last chunk): slots past the end re-read the last block and are
// masked out of the sum.
template <int U, bool FULL>
__device__ __forceinline__ void sell_chunk(const Sell<double>& a, int base, int lane, const int32_t* __restrict__ colp,
                                           const double* __restrict__ x, int k0, int len, double (&y)[3],
                                           double (&xo)[3]) {
#pragma clang fp contract(off)
  int kk[U], c[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    kk[j] = (FULL || k0 + j < len) ? k0 + j : len - 1;
    c[j] = colp[(int64_t)kk[j] * kRowsA];
  }
  double v[U][9], xv[U][3];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    aval_load(a.val, (int64_t)base + kk[j], lane, v[j]);
    xv[j][0] = x[3 * (int64_t)c[j]];
    xv[j][1] = x[3 * (int64_t)c[j] + 1];
    xv[j][2] = x[3 * (int64_t)c[j] + 2];
  }
  if (k0 == 0) {
    xo[0] = xv[0][0];
    xo[1] = xv[0][1];
    xo[2] = xv[0][2];
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const double t0 = y[0] + v[j][0] * xv[j][0] + v[j][1] * xv[j][1] + v[j][2] * xv[j][2];
    const double t1 = y[1] + v[j][3] * xv[j][0] + v[j][4] * xv[j][1] + v[j][5] * xv[j][2];
    const double t2 = y[2] + v[j][6] * xv[j][0] + v[j][7] * xv[j][1] + v[j][8] * xv[j][2];
    if (FULL) {
      y[0] = t0;
      y[1] = t1;
      y[2] = t2;
    } else {
      const bool live = k0 + j < len;
      y[0] = live ? t0 : y[0];
      y[1] = live ? t1 : y[1];
      y[2] = live ? t2 : y[2];
    }
  }
}

template <int U>
__device__ __forceinline__ void sell_row(const Sell<double>& a, int s, int lane, const double* __restrict__ x,
                                         double (&y)[3], double (&xo)[3]) {
#pragma clang fp contract(off)
  const int base = a.ptr[s];
  const int len = a.len[s * kRowsA + lane];
  y[0] = y[1] = y[2] = 0.0;
  xo[0] = xo[1] = xo[2] = 0.0;
  if (len == 0) return;
  y[0] = y[1] = y[2] = -0.0;
  const int32_t* __restrict__ colp = a.col + (int64_t)base * kRowsA + lane;
  int k0 = 0;
  for (; k0 + U <= len; k0 += U) sell_chunk<U, true>(a, base, lane, colp, x, k0, len, y, xo);
  if (k0 < len) sell_chunk<U, false>(a, base, lane, colp, x, k0, len, y, xo);
}

// FP32 values and vectors, FP32 accumulation with FMA (mixed-precision inner
// solves only).
template <int U, bool FULL>
__device__ __forceinline__ void sell_chunk_f32(const Sell<float>& a, int base, int lane,
                                               const int32_t* __restrict__ colp, const float* __restrict__ x, int k0,
                                               int len, float (&y)[3]) {
  int kk[U], c[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    kk[j] = (FULL || k0 + j < len) ? k0 + j : len - 1;
    c[j] = colp[(int64_t)kk[j] * kRowsA];
  }
  float v[U][9], xv[U][3];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    aval_load(a.val, (int64_t)base + kk[j], lane, v[j]);
    xv[j][0] = x[3 * (int64_t)c[j]];
    xv[j][1] = x[3 * (int64_t)c[j] + 1];
    xv[j][2] = x[3 * (int64_t)c[j] + 2];
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const float t0 = fmaf(v[j][2], xv[j][2], fmaf(v[j][1], xv[j][1], fmaf(v[j][0], xv[j][0], y[0])));
    const float t1 = fmaf(v[j][5], xv[j][2], fmaf(v[j][4], xv[j][1], fmaf(v[j][3], xv[j][0], y[1])));
    const float t2 = fmaf(v[j][8], xv[j][2], fmaf(v[j][7], xv[j][1], fmaf(v[j][6], xv[j][0], y[2])));
    const bool live = FULL || k0 + j < len;
    y[0] = live ? t0 : y[0];
    y[1] = live ? t1 : y[1];
    y[2] = live ? t2 : y[2];
  }
}

template <int U>
__device__ __forceinline__ void sell_row_f32(const Sell<float>& a, int s, int lane, const float* __restrict__ x,
                                             float (&y)[3]) {
  const int base = a.ptr[s];
  const int len = a.len[s * kRowsA + lane];
  y[0] = y[1] = y[2] = 0.0f;
  const int32_t* __restrict__ colp = a.col + (int64_t)base * kRowsA + lane;
  int k0 = 0;
  for (; k0 + U <= len; k0 += U) sell_chunk_f32<U, true>(a, base, lane, colp, x, k0, len, y);
  if (k0 < len) sell_chunk_f32<U, false>(a, base, lane, colp, x, k0, len, y);
}

// Slices handled by this wavefront: a contiguous chunk after XCD remapping.
struct SliceRange {
  int begin, end;
};
__device__ __forceinline__ SliceRange wave_slices(int nslices) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int nwaves = gridDim.x * kWavesPerBlock;
  const int w = blk * kWavesPerBlock + threadIdx.x / kWave;
  const int per = (nslices + nwaves - 1) / nwaves;
  SliceRange r;
  r.begin = w * per;
  r.end = min(nslices, r.begin + per);
  return r;
}

int spmv_grid(int nslices) {
  int g = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
  return g < 16384 ? (g < 1 ? 1 : g) : 16384;
}

int vec_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  return g < 1 ? 1 : (int)g;
}

constexpr int kSpmvChunk = 4;

// ---- SpMV kernels --------------------------------------------------------------
enum SpmvEpi : int {
  kEpiStore = 0,   // y = A x
  kEpiResid = 1,   // y = b - A x
  kEpiDotX = 2,    // y = A x, partial x.y            (CG: q = A p, p.q)
  kEpiDotW = 3,    // y = A x, partial w.y            (BiCGSTAB: v = A pt, rt.v)
  kEpiDot2W = 4,   // y = A x, partials w.y and y.y   (BiCGSTAB: t = A st, t.s, t.t)
  kEpiTrue = 5,    // y = b - A x, partial y.y, only in residual-recompute iterations
  kEpiStoreTrue = 6,  // y = A x, only in residual-recompute iterations (MPC: A T x of a true residual)
};

template <int EPI>
__global__ __launch_bounds__(kBlock) void k_spmv(Sell<double> a, const double* __restrict__ x,
                                                   const double* __restrict__ bw, double* __restrict__ y,
                                                   const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st) {
    if (st->done) return;
    if ((EPI == kEpiTrue || EPI == kEpiStoreTrue) && !true_resid_active(st)) return;
  }
  constexpr int K = EPI == kEpiDot2W ? 2 : 1;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(a.nslices);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int s = sr.begin; s < sr.end; ++s) {
    double yv[3], xo[3];
    sell_row<kSpmvChunk>(a, s, lane, x, yv, xo);
    const int64_t row = (int64_t)s * kRowsA + lane;
    if (row < a.nrows) {
      double o[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = (EPI == kEpiResid || EPI == kEpiTrue) ? bw[3 * row + d] - yv[d] : yv[d];
#pragma unroll
      for (int d = 0; d < 3; ++d) y[3 * row + d] = o[d];
      if (EPI == kEpiDotX) acc[0] += xo[0] * o[0] + xo[1] * o[1] + xo[2] * o[2];
      if (EPI == kEpiDotW || EPI == kEpiDot2W)
        acc[0] += bw[3 * row] * o[0] + bw[3 * row + 1] * o[1] + bw[3 * row + 2] * o[2];
      if (EPI == kEpiDot2W || EPI == kEpiTrue) acc[K - 1] += o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    }
  }
  if (EPI >= kEpiDotX && EPI <= kEpiTrue) block_partials<K>(acc, part, blockIdx.x);
}

// ---- MPC-CG: y = T^T A T x (hecmw_TtmatTvec_33, hecmw_solver_las_33.F90:442-586) ------------
// T x for the internal dofs (hecmw_Tvec_33): x, except a slave dof, which is
// 0 - val_2 x_2 - val_3 x_3 - ... over its constraint's master terms in order.
__global__ __launch_bounds__(kBlock) void k_mpc_tvec(MpcOp m, const double* __restrict__ x, double* __restrict__ y,
                                                       const KrylovState* st, int guard) {
#pragma clang fp contract(off)
  if (st) {
    if (st->done) return;
    if (guard && !true_resid_active(st)) return;
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m.n3; i += (int64_t)gridDim.x * kBlock) {
    const int c = m.slave_row[i];
    double v = x[i];
    if (c >= 0) {
      v = 0.0;
      for (int t = m.mptr[c]; t < m.mptr[c + 1]; ++t) v = v - m.mval[t] * x[m.mdof[t]];
    }
    y[i] = v;
  }
}

// T^T w for the internal rows (hecmw_Ttvec_33), w = A T x in m.w2: a slave entry
// is 0; a master dof subtracts val * w_slave for every constraint it is a
// master of, in constraint order (the reference's single-thread order), then
// the SpMV epilogue of EPI on the node's three entries (x: the product's input).
template <int EPI>
__global__ __launch_bounds__(kBlock) void k_mpc_ttvec(MpcOp m, const double* __restrict__ x,
                                                        const double* __restrict__ bw, double* __restrict__ y,
                                                        const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st) {
    if (st->done) return;
    if (EPI == kEpiTrue && !true_resid_active(st)) return;
  }
  constexpr int K = EPI == kEpiDot2W ? 2 : 1;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  const int64_t n = m.n3 / 3;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < n; row += (int64_t)gridDim.x * kBlock) {
    double o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t k = 3 * row + d;
      double v = m.slave_row[k] >= 0 ? 0.0 : m.w2[k];
      for (int t = m.tptr[k]; t < m.tptr[k + 1]; ++t) v = v - m.tval[t] * m.w2[m.tslave[t]];
      o[d] = (EPI == kEpiResid || EPI == kEpiTrue) ? bw[k] - v : v;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) y[3 * row + d] = o[d];
    if (EPI == kEpiDotX) acc[0] += x[3 * row] * o[0] + x[3 * row + 1] * o[1] + x[3 * row + 2] * o[2];
    if (EPI == kEpiDotW || EPI == kEpiDot2W)
      acc[0] += bw[3 * row] * o[0] + bw[3 * row + 1] * o[1] + bw[3 * row + 2] * o[2];
    if (EPI == kEpiDot2W || EPI == kEpiTrue) acc[K - 1] += o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
  }
  if (EPI >= kEpiDotX && EPI <= kEpiTrue) block_partials<K>(acc, part, blockIdx.x);
}

// FP32 SpMV; with DOTS (flexible CG: q = A p) also the FP64 partials of p.q
// and r.p, r passed in w.
template <bool DOTS>
__global__ __launch_bounds__(kBlock) void k_spmv_f32(Sell<float> a, const float* __restrict__ x,
                                                       const float* __restrict__ w, float* __restrict__ y,
                                                       const KrylovState* st, double* part) {
  if (st && st->done) return;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(a.nslices);
  double acc[2] = {0.0, 0.0};
  for (int s = sr.begin; s < sr.end; ++s) {
    float yv[3];
    sell_row_f32<kSpmvChunk>(a, s, lane, x, yv);
    const int64_t row = (int64_t)s * kRowsA + lane;
    if (row < a.nrows) {
      y[3 * row] = yv[0];
      y[3 * row + 1] = yv[1];
      y[3 * row + 2] = yv[2];
      if (DOTS) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double pv = x[3 * row + d];
          acc[0] += pv * (double)yv[d];
          acc[1] += (double)w[3 * row + d] * pv;
        }
      }
    }
  }
  if (DOTS) block_partials<2>(acc, part, blockIdx.x);
}

// A chunk of the FP64-value, FP32-vector row product (FULL as in sell_chunk)
template <int U, bool FULL>
__device__ __forceinline__ void mixed_chunk(const Sell<double>& a, int base, int lane,
                                            const int32_t* __restrict__ colp, const float* __restrict__ x, int k0,
                                            int len, double (&yv)[3]) {
  int kk[U], c[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    kk[j] = (FULL || k0 + j < len) ? k0 + j : len - 1;
    c[j] = colp[(int64_t)kk[j] * kRowsA];
  }
  double v[U][9], xv[U][3];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    aval_load(a.val, (int64_t)base + kk[j], lane, v[j]);
#pragma unroll
    for (int d = 0; d < 3; ++d) xv[j][d] = (double)x[3 * (int64_t)c[j] + d];
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const bool live = FULL || k0 + j < len;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t = fma(v[j][3 * d + 2], xv[j][2], fma(v[j][3 * d + 1], xv[j][1], fma(v[j][3 * d], xv[j][0], yv[d])));
      yv[d] = live ? t : yv[d];
    }
  }
}

// FP64 values, FP32 vectors, FP64 accumulation: q = A p for the flexible CG
// when the FP32-rounded matrix is too inexact (high material contrast), with
// the FP64 partials of p.q and r.p.
template <bool DOTS>
__global__ __launch_bounds__(kBlock) void k_spmv_mixed(Sell<double> a, const float* __restrict__ x,
                                                         const float* __restrict__ w, float* __restrict__ y,
                                                         const KrylovState* st, double* part) {
  if (st && st->done) return;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(a.nslices);
  double acc[2] = {0.0, 0.0};
  for (int s = sr.begin; s < sr.end; ++s) {
    const int base = a.ptr[s];
    const int len = a.len[s * kRowsA + lane];
    const int32_t* __restrict__ colp = a.col + (int64_t)base * kRowsA + lane;
    double yv[3] = {0.0, 0.0, 0.0};
    int k0 = 0;
    for (; k0 + kSpmvChunk <= len; k0 += kSpmvChunk) mixed_chunk<kSpmvChunk, true>(a, base, lane, colp, x, k0, len, yv);
    if (k0 < len) mixed_chunk<kSpmvChunk, false>(a, base, lane, colp, x, k0, len, yv);
    const int64_t row = (int64_t)s * kRowsA + lane;
    if (row < a.nrows) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        y[3 * row + d] = (float)yv[d];
        if (DOTS) {
          const double pv = x[3 * row + d];
          acc[0] += pv * yv[d];
          acc[1] += (double)w[3 * row + d] * pv;
        }
      }
    }
  }
  if (DOTS) block_partials<2>(acc, part, blockIdx.x);
}

// r32 = (float)(b - A x) with A, x, b in FP64 and the partials of |b - A x|^2;
// with st, only when the flexible CG asked for a reliable update.
__global__ __launch_bounds__(kBlock) void k_spmv_rel(Sell<double> a, const double* __restrict__ x,
                                                       const double* __restrict__ b, float* __restrict__ r,
                                                       const KrylovState* st, double* part) {
  if (st && (st->done || !st->reliable)) return;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(a.nslices);
  double acc[1] = {0.0};
  for (int s = sr.begin; s < sr.end; ++s) {
    double yv[3], xo[3];
    sell_row<kSpmvChunk>(a, s, lane, x, yv, xo);
    const int64_t row = (int64_t)s * kRowsA + lane;
    if (row < a.nrows) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double v = b[3 * row + d] - yv[d];
        r[3 * row + d] = (float)v;
        acc[0] += v * v;
      }
    }
  }
  block_partials<1>(acc, part, blockIdx.x);
}

// ---- SSOR sweeps (one launch per colour) ---------------------------------------
// LU solve with the factors of hecmw_precond_SSOR_33.F90:183-195 (inverted
// pivots on the diagonal, multipliers below): the reference's statements.
template <typename T>
__device__ __forceinline__ void lu3_solve(const T (&a)[9], T& x0, T& x1, T& x2) {
#pragma clang fp contract(off)
  x1 = x1 - a[3] * x0;
  x2 = x2 - a[6] * x0 - a[7] * x1;
  x2 = a[8] * x2;
  x1 = a[4] * (x1 - a[5] * x2);
  x0 = a[0] * (x0 - a[2] * x2 - a[1] * x1);
}

// 9 LU factors of row r of slice s, 16-row planes (all lanes of the quad)
template <typename T>
__device__ __forceinline__ void load_alu(const T* __restrict__ alu, int s, int r, T (&a)[9]) {
  const T* __restrict__ f = alu + (int64_t)s * 9 * kRowsS + r;
#pragma unroll
  for (int e = 0; e < 9; ++e) a[e] = f[e * kRowsS];
}

// an L'/U' value as the sweep's arithmetic type: stored as T, or as BF16 (the
// upper 16 bits of an FP32) for the FP32 sweeps
template <typename T, typename S>
__device__ __forceinline__ T ldv(const S* __restrict__ p) {
  if constexpr (sizeof(S) == 2)
    return __uint_as_float((uint32_t)*p << 16);
  else
    return *p;
}

template <typename T>
__device__ __forceinline__ T pick3(int d, T a, T b, T c) {
  return d == 0 ? a : (d == 1 ? b : c);
}

// Accumulate blocks [k0, k0+CH) of a row (descending when DESC) into the
// lane's component sum; all CH column indices, gathers and values are
// requested before the first one is used.  SIGN -1: sw - a*x (forward),
// +1: sw + a*x (backward), each block as the reference writes it.
template <typename T, typename S, int CH, bool DESC, int SIGN, bool FULL>
__device__ __forceinline__ void ssor_chunk(const int32_t* __restrict__ colp, const S* __restrict__ vp,
                                           const T* __restrict__ z, int k0, int len, int dd, T& sw) {
#pragma clang fp contract(off)
  // FULL (every block live): no load or sum under a condition.  The row's
  // last chunk re-reads a valid block in the slots past the end and masks
  // them out of the sum.
  int kk[CH], c[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const int k = DESC ? k0 - j : k0 + j;
    kk[j] = FULL ? k : (DESC ? (k < 0 ? 0 : k) : (k < len ? k : len - 1));
    c[j] = colp[(int64_t)kk[j] * kRowsS];
  }
  T xd[CH], v[CH][3];
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const S* __restrict__ vk = vp + (int64_t)kk[j] * 9 * kRowsS;
    xd[j] = z[3 * (int64_t)c[j] + dd];
    v[j][0] = ldv<T>(vk);
    v[j][1] = ldv<T>(vk + kRowsS);
    v[j][2] = ldv<T>(vk + 2 * kRowsS);
  }
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const int k = DESC ? k0 - j : k0 + j;
    const bool live = FULL || (DESC ? k >= 0 : k < len);
    const T x0 = quad_bcast<0>(xd[j]), x1 = quad_bcast<1>(xd[j]), x2 = quad_bcast<2>(xd[j]);
    const T t = SIGN < 0 ? sw - v[j][0] * x0 - v[j][1] * x1 - v[j][2] * x2
                         : sw + v[j][0] * x0 + v[j][1] * x1 + v[j][2] * x2;
    sw = live ? t : sw;
  }
}

// The sweeps work on zs, the preconditioned vector in slot order (3 entries
// per slot; L'/U' columns hold slot ids), so the rows of a slice, which are
// consecutive in colour order, gather neighbours that are consecutive too.
// r is staged into zs once before the forward sweep (ZP = R,
// hecmw_precond.f90:114-117) and z leaves in the caller's (natural) node
// order once after the backward sweep, so the colour launches read and write
// slot-order memory only: a colour's rows are spread over the whole mesh, and
// per-colour accesses in node order would touch a separate cache line for
// every row.
//
// Forward, rows of one colour independent:
//   zs[i] = LU^-1 (zs[i] - sum_{j in L'(i), ascending} L'_ij zs[j])
template <typename T, typename S, int CH>
__global__ __launch_bounds__(kBlock) void k_ssor_fwd(Sell<T> L, const S* __restrict__ vals, const int32_t* __restrict__ row_old,
                                                       const T* __restrict__ alu, int s0, int ns,
                                                       T* __restrict__ zs, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = lane_id(), r = lane >> 2, d = lane & 3, dd = d < 3 ? d : 2;
  const SliceRange sr = wave_slices(ns);
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kRowsS + r;
    const int iold = row_old[slot];
    const int len = L.len[slot];
    const int base = L.ptr[s];
    T a[9];
    load_alu<T>(alu, s, r, a);
    if (iold < 0) continue;
    const int32_t* __restrict__ colp = L.col + (int64_t)base * kRowsS + r;
    const S* __restrict__ vp = vals + ((int64_t)base * 9 + 3 * dd) * kRowsS + r;
    T sw = zs[3 * slot + dd];
    {
      int k0 = 0;
      for (; k0 + CH <= len; k0 += CH) ssor_chunk<T, S, CH, false, -1, true>(colp, vp, zs, k0, len, dd, sw);
      if (k0 < len) ssor_chunk<T, S, CH, false, -1, false>(colp, vp, zs, k0, len, dd, sw);
    }
    T x0 = quad_bcast<0>(sw), x1 = quad_bcast<1>(sw), x2 = quad_bcast<2>(sw);
    lu3_solve<T>(a, x0, x1, x2);
    if (d < 3) zs[3 * slot + d] = pick3(d, x0, x1, x2);
  }
}

// Backward, colours in reverse:
//   zs[i] -= LU^-1 sum_{j in U'(i), descending} U'_ij zs[j]
template <typename T, typename S, int CH>
__global__ __launch_bounds__(kBlock) void k_ssor_bwd(Sell<T> U, const S* __restrict__ vals, const int32_t* __restrict__ row_old,
                                                       const T* __restrict__ alu, int s0, int ns,
                                                       T* __restrict__ zs, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = lane_id(), r = lane >> 2, d = lane & 3, dd = d < 3 ? d : 2;
  const SliceRange sr = wave_slices(ns);
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kRowsS + r;
    const int iold = row_old[slot];
    const int len = U.len[slot];
    const int base = U.ptr[s];
    T a[9];
    load_alu<T>(alu, s, r, a);
    if (iold < 0) continue;
    const int32_t* __restrict__ colp = U.col + (int64_t)base * kRowsS + r;
    const S* __restrict__ vp = vals + ((int64_t)base * 9 + 3 * dd) * kRowsS + r;
    const T zo = zs[3 * slot + dd];
    T sw = 0;
    {
      int top = len - 1;
      for (; top - CH + 1 >= 0; top -= CH) ssor_chunk<T, S, CH, true, 1, true>(colp, vp, zs, top, len, dd, sw);
      if (top >= 0) ssor_chunk<T, S, CH, true, 1, false>(colp, vp, zs, top, len, dd, sw);
    }
    T x0 = quad_bcast<0>(sw), x1 = quad_bcast<1>(sw), x2 = quad_bcast<2>(sw);
    lu3_solve<T>(a, x0, x1, x2);
    const T zn = zo - pick3(dd, x0, x1, x2);
    if (d < 3) zs[3 * slot + d] = zn;
  }
}

// FP32 -> BF16, round to nearest even (finite values)
__global__ __launch_bounds__(kBlock) void k_to_bf16(int64_t n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint32_t u = __float_as_uint(a[i]);
    b[i] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
}

// ---- SSOR sweeps, SELL-64 layout: one lane per row ------------------------------
// For large colours (many more slices than resident waves) a lane per row
// needs a quarter of the waves of the four-lane layout for the same dependent
// load chains.  Same arithmetic per row and component as ssor_chunk + the
// quad layout: the reference's statements in the reference's order.
template <typename T, typename S, int CH, bool DESC, int SIGN, bool FULL>
__device__ __forceinline__ void ssor64_chunk(const int32_t* __restrict__ colp, const S* __restrict__ vp,
                                             const T* __restrict__ z, int k0, int len, T (&sw)[3]) {
#pragma clang fp contract(off)
  int kk[CH], c[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const int k = DESC ? k0 - j : k0 + j;
    kk[j] = FULL ? k : (DESC ? (k < 0 ? 0 : k) : (k < len ? k : len - 1));
    c[j] = colp[(int64_t)kk[j] * kRowsA];
  }
  T v[CH][9], x[CH][3];
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const S* __restrict__ vk = vp + (int64_t)kk[j] * 9 * kRowsA;
#pragma unroll
    for (int e = 0; e < 9; ++e) v[j][e] = ldv<T>(vk + e * kRowsA);
#pragma unroll
    for (int d = 0; d < 3; ++d) x[j][d] = z[3 * (int64_t)c[j] + d];
  }
#pragma unroll
  for (int j = 0; j < CH; ++j) {
    const int k = DESC ? k0 - j : k0 + j;
    const bool live = FULL || (DESC ? k >= 0 : k < len);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T t = SIGN < 0 ? sw[d] - v[j][3 * d] * x[j][0] - v[j][3 * d + 1] * x[j][1] - v[j][3 * d + 2] * x[j][2]
                           : sw[d] + v[j][3 * d] * x[j][0] + v[j][3 * d + 1] * x[j][1] + v[j][3 * d + 2] * x[j][2];
      sw[d] = live ? t : sw[d];
    }
  }
}

template <typename T>
__device__ __forceinline__ void load_alu64(const T* __restrict__ alu, int s, int lane, T (&a)[9]) {
  const T* __restrict__ f = alu + (int64_t)s * 9 * kRowsA + lane;
#pragma unroll
  for (int e = 0; e < 9; ++e) a[e] = f[e * kRowsA];
}

template <typename T, typename S, int CH>
__global__ __launch_bounds__(kBlock) void k_ssor64_fwd(Sell<T> L, const S* __restrict__ vals, const int32_t* __restrict__ row_old,
                                                         const T* __restrict__ alu, int s0, int ns,
                                                         T* __restrict__ zs, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(ns);
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kRowsA + lane;
    const int iold = row_old[slot];
    const int len = L.len[slot];
    const int base = L.ptr[s];
    T a[9];
    load_alu64<T>(alu, s, lane, a);
    if (iold < 0) continue;
    const int32_t* __restrict__ colp = L.col + (int64_t)base * kRowsA + lane;
    const S* __restrict__ vp = vals + (int64_t)base * 9 * kRowsA + lane;
    T sw[3] = {zs[3 * slot], zs[3 * slot + 1], zs[3 * slot + 2]};
    {
      int k0 = 0;
      for (; k0 + CH <= len; k0 += CH) ssor64_chunk<T, S, CH, false, -1, true>(colp, vp, zs, k0, len, sw);
      if (k0 < len) ssor64_chunk<T, S, CH, false, -1, false>(colp, vp, zs, k0, len, sw);
    }
    lu3_solve<T>(a, sw[0], sw[1], sw[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) zs[3 * slot + d] = sw[d];
  }
}

template <typename T, typename S, int CH>
__global__ __launch_bounds__(kBlock) void k_ssor64_bwd(Sell<T> U, const S* __restrict__ vals, const int32_t* __restrict__ row_old,
                                                         const T* __restrict__ alu, int s0, int ns,
                                                         T* __restrict__ zs, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = lane_id();
  const SliceRange sr = wave_slices(ns);
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kRowsA + lane;
    const int iold = row_old[slot];
    const int len = U.len[slot];
    const int base = U.ptr[s];
    T a[9];
    load_alu64<T>(alu, s, lane, a);
    if (iold < 0) continue;
    const int32_t* __restrict__ colp = U.col + (int64_t)base * kRowsA + lane;
    const S* __restrict__ vp = vals + (int64_t)base * 9 * kRowsA + lane;
    T sw[3] = {0, 0, 0};
    {
      int top = len - 1;
      for (; top - CH + 1 >= 0; top -= CH) ssor64_chunk<T, S, CH, true, 1, true>(colp, vp, zs, top, len, sw);
      if (top >= 0) ssor64_chunk<T, S, CH, true, 1, false>(colp, vp, zs, top, len, sw);
    }
    lu3_solve<T>(a, sw[0], sw[1], sw[2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) zs[3 * slot + d] = zs[3 * slot + d] - sw[d];
  }
}

// r into slot order: zs[slot(i)] = r[i] (node order in, so the reads are
// contiguous; a line of zs is completed by waves of nearby nodes)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ssor_stage(int32_t n, const int32_t* __restrict__ node_slot,
                                                         const T* __restrict__ r, T* __restrict__ zs,
                                                         const KrylovState* st) {
  if (st && st->done) return;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t t = node_slot[i];
    const T a = r[3 * i], b = r[3 * i + 1], c = r[3 * i + 2];
    zs[3 * t] = a;
    zs[3 * t + 1] = b;
    zs[3 * t + 2] = c;
  }
}

// z = zs back in node order; ND fused dots w0.z (and w1.z), FP64 accumulation
template <typename T, int ND>
__global__ __launch_bounds__(kBlock) void k_ssor_unstage(int32_t n, const int32_t* __restrict__ node_slot,
                                                           const T* __restrict__ zs, T* __restrict__ z,
                                                           const T* __restrict__ w0, const T* __restrict__ w1,
                                                           const KrylovState* st, double* part) {
  if (st && st->done) return;
  constexpr int KD = ND > 0 ? ND : 1;
  double acc[KD] = {};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t t = node_slot[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T v = zs[3 * t + d];
      z[3 * i + d] = v;
      if constexpr (ND > 0) acc[0] += (double)w0[3 * i + d] * (double)v;
      if constexpr (ND > 1) acc[KD - 1] += (double)w1[3 * i + d] * (double)v;
    }
  }
  if constexpr (ND > 0) block_partials<KD>(acc, part, blockIdx.x);
}

// ---- DIAG (block Jacobi), hecmw_precond_DIAG_33.F90:106-142 --------------------
// ND fused dots: r.z (and w1.z), FP64 accumulation.
template <typename T, int ND>
__global__ __launch_bounds__(kBlock) void k_diag(int32_t n, const T* __restrict__ alu, const T* __restrict__ r,
                                                   T* __restrict__ z, const T* __restrict__ w1,
                                                   const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  constexpr int KD = ND > 0 ? ND : 1;
  double d[KD] = {};
  for (int64_t i = (int64_t)blk * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t s = i / kWave, lane = i % kWave;
    const T* f = alu + s * 9 * kWave + lane;
    T a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = f[e * kWave];
    const T r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    T x0 = r0, x1 = r1, x2 = r2;
    lu3_solve<T>(a, x0, x1, x2);
    z[3 * i] = x0;
    z[3 * i + 1] = x1;
    z[3 * i + 2] = x2;
    if constexpr (ND > 0) d[0] += (double)r0 * (double)x0 + (double)r1 * (double)x1 + (double)r2 * (double)x2;
    if constexpr (ND > 1)
      d[KD - 1] += (double)w1[3 * i] * (double)x0 + (double)w1[3 * i + 1] * (double)x1 +
                   (double)w1[3 * i + 2] * (double)x2;
  }
  if constexpr (ND > 0) block_partials<KD>(d, part, blockIdx.x);
}

// ---- vector kernels -----------------------------------------------------------
// z = r with partial r.r (no preconditioning: rho = r.z)
__global__ __launch_bounds__(kBlock) void k_copy_dot(int64_t n, const double* __restrict__ r, double* __restrict__ z,
                                                       const KrylovState* st, double* part, int with_dot) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double v = r[i];
    z[i] = v;
    d[0] += v * v;
  }
  if (with_dot) block_partials<1>(d, part, blockIdx.x);
}

// x = 0 / y += x over n entries (the ITERPREMAX > 1 accumulation of
// hecmw_precond.f90:120-165)
__global__ __launch_bounds__(kBlock) void k_zero(int64_t n, double* __restrict__ x, const KrylovState* st) {
  if (st && st->done) return;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) x[i] = 0.0;
}
__global__ __launch_bounds__(kBlock) void k_add(int64_t n, const double* __restrict__ x, double* __restrict__ y,
                                                  const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = y[i] + x[i];
}

// partials of x.y
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                  const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[0] += x[i] * y[i];
  block_partials<1>(d, part, blockIdx.x);
}

// p = z (first iteration) or p = z + beta p (hecmw_xpay_R: Y = X + a*Y)
__global__ __launch_bounds__(kBlock) void k_cg_xpay(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                                      const KrylovState* st) {
#pragma clang fp contract(off)
  if (st->done) return;
  const bool first = st->iter == 1;
  const double beta = st->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

// x += alpha p; r -= alpha q (hecmw_axpy_R: Y = a*X + Y) with partial r.r,
// except in iterations that recompute the true residual (x only).
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, const double* __restrict__ p,
                                                        const double* __restrict__ q, double* __restrict__ x,
                                                        double* __restrict__ r, const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st->done) return;
  const double alpha = st->alpha;
  const bool periodic = (st->iter % st->nrec) == 0;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] = alpha * p[i] + x[i];
    if (!periodic) {
      const double v = -alpha * q[i] + r[i];
      r[i] = v;
      d[0] += v * v;
    }
  }
  if (!periodic) block_partials<1>(d, part, blockIdx.x);
}

// BiCGSTAB p (hecmw_solver_BiCGSTAB.f90:176-182): p = r on the first
// iteration, else p = -omega v + p (axpy) then p = r + beta p (xpay).
__global__ __launch_bounds__(kBlock) void k_bi_p(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                                   double* __restrict__ p, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st->done) return;
  const bool first = st->iter == 1;
  const double beta = st->beta, omega = st->omega;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    if (first) {
      p[i] = r[i];
    } else {
      const double t = -omega * v[i] + p[i];
      p[i] = r[i] + beta * t;
    }
  }
}

// s = -alpha v + r (hecmw_axpyz_R: Z = a*X + Y)
__global__ __launch_bounds__(kBlock) void k_bi_s(int64_t n, const double* __restrict__ v, const double* __restrict__ r,
                                                   double* __restrict__ s, const KrylovState* st) {
#pragma clang fp contract(off)
  if (st->done) return;
  const double alpha = st->alpha;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s[i] = -alpha * v[i] + r[i];
}

// x = alpha pt + x; x = omega st + x; r = -omega t + s with partial r.r unless
// the iteration recomputes the true residual (hecmw_solver_BiCGSTAB.f90:241-249).
__global__ __launch_bounds__(kBlock) void k_bi_update(int64_t n, const double* __restrict__ pt,
                                                        const double* __restrict__ stv, const double* __restrict__ t,
                                                        const double* __restrict__ s, double* __restrict__ x,
                                                        double* __restrict__ r, const KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st->done) return;
  const double alpha = st->alpha, omega = st->omega;
  const bool periodic = (st->iter % st->nrec) == 0;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double x1 = alpha * pt[i] + x[i];
    x[i] = omega * stv[i] + x1;
    if (!periodic) {
      const double v = -omega * t[i] + s[i];
      r[i] = v;
      d[0] += v * v;
    }
  }
  if (!periodic) block_partials<1>(d, part, blockIdx.x);
}

// ---- flexible CG vector kernels (FP32 vectors, FP64 scalars and x) ---------------
// z = r with the partials of r.r and w1.r (no preconditioning)
__global__ __launch_bounds__(kBlock) void k_copy_dots_f32(int64_t n, const float* __restrict__ r,
                                                            float* __restrict__ z, const float* __restrict__ w1,
                                                            const KrylovState* st, double* part, int ndots) {
  if (st && st->done) return;
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float v = r[i];
    z[i] = v;
    d[0] += (double)v * (double)v;
    if (ndots > 1) d[1] += (double)w1[i] * (double)v;
  }
  if (ndots > 0) block_partials<2>(d, part, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_fcg_p(int64_t n, const float* __restrict__ z, float* __restrict__ p,
                                                    const KrylovState* st) {
  if (st->done) return;
  const bool first = st->iter == 1;
  const double beta = st->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = first ? z[i] : (float)((double)z[i] + beta * (double)p[i]);
}

// x += alpha p (FP64 x); r = r - alpha q rounded to FP32; partials of |r|^2
__global__ __launch_bounds__(kBlock) void k_fcg_update(int64_t n, const float* __restrict__ p,
                                                         const float* __restrict__ q, double* __restrict__ x,
                                                         float* __restrict__ r, const KrylovState* st, double* part) {
  if (st->done) return;
  const double alpha = st->alpha;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] = x[i] + alpha * (double)p[i];
    const float v = (float)((double)r[i] - alpha * (double)q[i]);
    r[i] = v;
    d[0] += (double)v * (double)v;
  }
  block_partials<1>(d, part, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = (float)x[i];
}

// ---- diagonal scaling, hecmw_solver_scaling_33.f90:20-208 -----------------------
// s = 1/sqrt(|D_ii|) per DOF of the internal nodes; external nodes get 1 (the
// reference takes them from a halo exchange, absent on one domain).
__global__ void k_make_scale(int32_t n, int32_t np, const double* __restrict__ D, double* __restrict__ sc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int d = 0; d < 3; ++d) sc[3 * i + d] = i < n ? 1.0 / sqrt(fabs(D[9 * i + 4 * d])) : 1.0;
}

// FW: a = a*s_p*s_q (left to right); BK: a = a/(s_p*s_q), over the caller's
// D/AL/AU rows 0..NP-1 as the reference loops them.
template <bool FW>
__device__ __forceinline__ void scale_block(double* __restrict__ a, const double* __restrict__ sp,
                                            const double* __restrict__ sq) {
#pragma clang fp contract(off)
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) a[3 * p + q] = FW ? a[3 * p + q] * sp[p] * sq[q] : a[3 * p + q] / (sp[p] * sq[q]);
}
template <bool FW>
__global__ void k_scale_raw(int32_t np, const int32_t* __restrict__ indexL, const int32_t* __restrict__ itemL,
                            const int32_t* __restrict__ indexU, const int32_t* __restrict__ itemU,
                            double* __restrict__ D, double* __restrict__ AL, double* __restrict__ AU,
                            const double* __restrict__ sc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const double* sp = sc + 3 * i;
    scale_block<FW>(D + 9 * i, sp, sp);
    for (int64_t k = indexL[i]; k < indexL[i + 1]; ++k) scale_block<FW>(AL + 9 * k, sp, sc + 3 * (int64_t)(itemL[k] - 1));
    for (int64_t k = indexU[i]; k < indexU[i + 1]; ++k) scale_block<FW>(AU + 9 * k, sp, sc + 3 * (int64_t)(itemU[k] - 1));
  }
}

// v = v*s (MUL) or v = v/s over n entries
template <bool MUL>
__global__ void k_scale_vec(int64_t n, double* __restrict__ v, const double* __restrict__ sc) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = MUL ? v[i] * sc[i] : v[i] / sc[i];
}

// ---- numeric setup ---------------------------------------------------------------
template <typename T, bool PAIRED>
__global__ void k_fill_values(Sell<T> a, const int64_t* __restrict__ src, const double* __restrict__ D,
                              const double* __restrict__ AL, const double* __restrict__ AU, int64_t nD, int64_t nAL) {
  const int64_t nslot = a.nslots;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nslot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t C = a.rows;
    const int64_t grp = t / C, r = t % C;  // slot = grp*C + r
    const int64_t b = src[t];
    const double* blk = nullptr;
    if (b >= 0) blk = b < nD ? D + 9 * b : (b < nD + nAL ? AL + 9 * (b - nD) : AU + 9 * (b - nD - nAL));
    if (PAIRED) {
      using V2 = typename Pair2<T>::type;
      T* out = a.val + grp * (9 * kRowsA);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        V2 t;
        t.x = blk ? (T)blk[2 * q] : (T)0;
        t.y = blk ? (T)blk[2 * q + 1] : (T)0;
        *reinterpret_cast<V2*>(out + q * (2 * kRowsA) + 2 * r) = t;
      }
      out[8 * kRowsA + r] = blk ? (T)blk[8] : (T)0;
    } else {
      T* out = a.val + grp * 9 * C + r;
#pragma unroll
      for (int e = 0; e < 9; ++e) out[e * C] = blk ? (T)blk[e] : (T)0;
    }
  }
}

template <typename T>
__global__ void k_block_lu(int32_t nslots, const int32_t* __restrict__ row_src, const double* __restrict__ D,
                           double sigma, T* __restrict__ alu, int rows) {
#pragma clang fp contract(off)
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nslots; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t grp = t / rows, r = t % rows;
    const int32_t i = row_src[t];
    double a[3][3];
    if (i >= 0) {
      const double* d = D + 9 * (int64_t)i;
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) a[p][q] = d[3 * p + q];
      a[0][0] = a[0][0] * sigma;
      a[1][1] = a[1][1] * sigma;
      a[2][2] = a[2][2] * sigma;
      // hecmw_precond_SSOR_33.F90:183-195: inverted pivots on the diagonal,
      // multipliers below, updated U above.
      for (int k = 0; k < 3; ++k) {
        a[k][k] = 1.0 / a[k][k];
        for (int p = k + 1; p < 3; ++p) {
          a[p][k] = a[p][k] * a[k][k];
          double pw[3];
          for (int q = k + 1; q < 3; ++q) pw[q] = a[p][q] - a[p][k] * a[k][q];
          for (int q = k + 1; q < 3; ++q) a[p][q] = pw[q];
        }
      }
    } else {
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) a[p][q] = 0.0;
    }
    T* out = alu + grp * 9 * rows + r;
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) out[(3 * p + q) * rows] = (T)a[p][q];
  }
}

// ---- HBM read ceiling -----------------------------------------------------------------
// Sum of n2 double2 values with a grid-stride loop, one partial per workgroup:
// the plain streaming read that bounds what any kernel here can get from HBM.
__global__ __launch_bounds__(kBlock) void k_stream_read(const double2* __restrict__ a, int64_t n2,
                                                         double* __restrict__ part) {
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n2; i += (int64_t)gridDim.x * kBlock) {
    const double2 v = a[i];
    acc[0] += v.x + v.y;
  }
  block_partials<1>(acc, part, blockIdx.x);
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------
template <int EPI>
static void spmv_launch(int g, hipStream_t s, const Sell<double>& a, const double* x, const double* bw, double* y,
                        const KrylovState* st, double* part) {
  k_spmv<EPI><<<g, kBlock, 0, s>>>(a, x, bw, y, st, part);
}

// y = A x (or T^T A T x with MPC) with the epilogue EPI; returns the number of
// partials written
template <int EPI>
static int spmv_any(hipStream_t s, const Sell<double>& a, const MpcOp* m, const double* x, const double* bw,
                    double* y, const KrylovState* st, double* part) {
  if (!m || m->n3 == 0) {
    const int g = spmv_grid(a.nslices);
    spmv_launch<EPI>(g, s, a, x, bw, y, st, part);
    return g;
  }
  k_mpc_tvec<<<vec_grid(m->n3), kBlock, 0, s>>>(*m, x, m->w1, st, EPI == kEpiTrue ? 1 : 0);
  if (EPI == kEpiTrue)
    spmv_launch<kEpiStoreTrue>(spmv_grid(a.nslices), s, a, m->w1, nullptr, m->w2, st, nullptr);
  else
    spmv_launch<kEpiStore>(spmv_grid(a.nslices), s, a, m->w1, nullptr, m->w2, st, nullptr);
  const int g = vec_grid(m->n3 / 3);
  k_mpc_ttvec<EPI><<<g, kBlock, 0, s>>>(*m, x, bw, y, st, part);
  return g;
}

void launch_stream_read(const double* a, int64_t n, double* part, int grid, hipStream_t s) {
  k_stream_read<<<grid, kBlock, 0, s>>>(reinterpret_cast<const double2*>(a), n / 2, part);
}
void launch_spmv(const Sell<double>& a, const double* x, double* y, hipStream_t s) {
  spmv_launch<kEpiStore>(spmv_grid(a.nslices), s, a, x, nullptr, y, nullptr, nullptr);
}
void launch_spmv_mpc(const Sell<double>& a, const MpcOp& mpc, const double* x, double* y, hipStream_t s) {
  spmv_any<kEpiStore>(s, a, &mpc, x, nullptr, y, nullptr, nullptr);
}
void launch_spmv_f32(const Sell<float>& a, const float* x, float* y, hipStream_t s) {
  k_spmv_f32<false><<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, x, nullptr, y, nullptr, nullptr);
}
void launch_matresid(const Sell<double>& a, const double* x, const double* b, double* r, hipStream_t s,
                     const MpcOp* mpc) {
  spmv_any<kEpiResid>(s, a, mpc, x, b, r, nullptr, nullptr);
}
void launch_dot_store(int64_t n, const double* x, const double* y, double* out, KrylovState* st, double* part,
                      hipStream_t s) {
  const int g = vec_grid(n);
  k_dot<<<g, kBlock, 0, s>>>(n, x, y, nullptr, part);
  k_tail<kTailStore, 1><<<1, kTailThreads, 0, s>>>(st, part, g, out, nullptr);
}
void launch_fill_values(const Sell<double>& a, const int64_t* src, const double* D, const double* AL,
                        const double* AU, int64_t nD, int64_t nAL, hipStream_t s) {
  k_fill_values<double, false><<<4096, 256, 0, s>>>(a, src, D, AL, AU, nD, nAL);
}
void launch_fill_values_a(const Sell<double>& a, const int64_t* src, const double* D, const double* AL,
                          const double* AU, int64_t nD, int64_t nAL, hipStream_t s) {
  k_fill_values<double, true><<<4096, 256, 0, s>>>(a, src, D, AL, AU, nD, nAL);
}
void launch_fill_values_a_f32(const Sell<float>& a, const int64_t* src, const double* D, const double* AL,
                              const double* AU, int64_t nD, int64_t nAL, hipStream_t s) {
  k_fill_values<float, true><<<4096, 256, 0, s>>>(a, src, D, AL, AU, nD, nAL);
}
void launch_block_lu(int32_t nslots, const int32_t* row_src, const double* D, double sigma, double* alu, int rows,
                     hipStream_t s) {
  k_block_lu<double><<<2048, 256, 0, s>>>(nslots, row_src, D, sigma, alu, rows);
}
void launch_block_lu_f32(int32_t nslots, const int32_t* row_src, const double* D, double sigma, float* alu, int rows,
                         hipStream_t s) {
  k_block_lu<float><<<2048, 256, 0, s>>>(nslots, row_src, D, sigma, alu, rows);
}
// Blocks whose loads the four-lane sweeps issue together (HECMW33_SSOR_CHUNK =
// 4, 8 or 16; 4 by default: with the row's last chunk alone masked, 4 runs
// the 100^3 FP64 application in 1.000 ms against 1.094 ms for 8).
static int ssor_chunk_sel() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("HECMW33_SSOR_CHUNK");
    const int c = e ? atoi(e) : 4;
    v = (c == 8 || c == 16) ? c : 4;
  }
  return v;
}

// Blocks in flight per lane of the one-lane-per-row sweeps (HECMW33_SSOR64_CHUNK
// = 4 or 8; 4 by default).
static int ssor64_chunk_sel() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("HECMW33_SSOR64_CHUNK");
    v = (e && atoi(e) == 8) ? 8 : 4;
  }
  return v;
}

template <typename T, bool FWD, int CH>
static void ssor_launch4(const SsorOp<T>& op, int g, int s0, int ns, const KrylovState* st, hipStream_t s) {
  const uint16_t* v16 = FWD ? op.Lv16 : op.Uv16;
  if constexpr (std::is_same<T, float>::value) {
    if (v16) {
      if constexpr (FWD)
        k_ssor_fwd<T, uint16_t, CH><<<g, kBlock, 0, s>>>(op.L, v16, op.row_old, op.alu, s0, ns, op.zs, st);
      else
        k_ssor_bwd<T, uint16_t, CH><<<g, kBlock, 0, s>>>(op.U, v16, op.row_old, op.alu, s0, ns, op.zs, st);
      return;
    }
  }
  if constexpr (FWD)
    k_ssor_fwd<T, T, CH><<<g, kBlock, 0, s>>>(op.L, op.L.val, op.row_old, op.alu, s0, ns, op.zs, st);
  else
    k_ssor_bwd<T, T, CH><<<g, kBlock, 0, s>>>(op.U, op.U.val, op.row_old, op.alu, s0, ns, op.zs, st);
}

template <typename T, bool FWD, int CH>
static void ssor_launch64(const SsorOp<T>& op, int g, int s0, int ns, const KrylovState* st, hipStream_t s) {
  const uint16_t* v16 = FWD ? op.Lv16 : op.Uv16;
  if constexpr (std::is_same<T, float>::value) {
    if (v16) {
      if constexpr (FWD)
        k_ssor64_fwd<T, uint16_t, CH><<<g, kBlock, 0, s>>>(op.L, v16, op.row_old, op.alu, s0, ns, op.zs, st);
      else
        k_ssor64_bwd<T, uint16_t, CH><<<g, kBlock, 0, s>>>(op.U, v16, op.row_old, op.alu, s0, ns, op.zs, st);
      return;
    }
  }
  if constexpr (FWD)
    k_ssor64_fwd<T, T, CH><<<g, kBlock, 0, s>>>(op.L, op.L.val, op.row_old, op.alu, s0, ns, op.zs, st);
  else
    k_ssor64_bwd<T, T, CH><<<g, kBlock, 0, s>>>(op.U, op.U.val, op.row_old, op.alu, s0, ns, op.zs, st);
}

template <typename T, bool FWD>
static void ssor_colour(const SsorOp<T>& op, int g, int s0, int ns, const KrylovState* st, hipStream_t s) {
  if ((FWD ? op.L.rows : op.U.rows) == kRowsA) {
    if (ssor64_chunk_sel() == 8)
      ssor_launch64<T, FWD, 8>(op, g, s0, ns, st, s);
    else
      ssor_launch64<T, FWD, 4>(op, g, s0, ns, st, s);
    return;
  }
  switch (ssor_chunk_sel()) {
    case 4: ssor_launch4<T, FWD, 4>(op, g, s0, ns, st, s); break;
    case 16: ssor_launch4<T, FWD, 16>(op, g, s0, ns, st, s); break;
    default: ssor_launch4<T, FWD, 8>(op, g, s0, ns, st, s); break;
  }
}

template <typename T>
static void ssor_fwd_any(const SsorOp<T>& op, int c, const KrylovState* st, hipStream_t s) {
  const int32_t s0 = op.color_slice[c], ns = op.color_slice[c + 1] - s0;
  if (ns <= 0) return;
  const int g = spmv_grid(ns);
  ssor_colour<T, true>(op, g, s0, ns, st, s);
}

template <typename T>
static void ssor_bwd_any(const SsorOp<T>& op, int c, const KrylovState* st, hipStream_t s) {
  const int32_t s0 = op.color_slice[c], ns = op.color_slice[c + 1] - s0;
  if (ns <= 0) return;
  const int g = spmv_grid(ns);
  ssor_colour<T, false>(op, g, s0, ns, st, s);
}

// stage r, every colour forward, then backward in reverse (the last colour's
// backward subtracts zero when it has no U' entries and is left out), unstage
// z with ND fused dots w0.z (and w1.z); returns the partials written
template <typename T, int ND>
static int ssor_apply_run(const SsorOp<T>& op, const T* r, T* z, const T* w1, const KrylovState* st, double* part,
                          hipStream_t s) {
  const int gv = vec_grid(op.n);
  k_ssor_stage<T><<<gv, kBlock, 0, s>>>(op.n, op.node_slot, r, op.zs, st);
  for (int c = 0; c < op.ncolor; ++c) ssor_fwd_any<T>(op, c, st, s);
  for (int c = op.ncolor - 1; c >= 0; --c) {
    if (c == op.ncolor - 1 && op.last_bwd_empty) continue;
    ssor_bwd_any<T>(op, c, st, s);
  }
  k_ssor_unstage<T, ND><<<gv, kBlock, 0, s>>>(op.n, op.node_slot, op.zs, z, r, w1, st, part);
  return ND > 0 ? gv : 0;
}

template <typename T>
static void ssor_apply_any(const SsorOp<T>& op, const T* r, T* z, const KrylovState* st, hipStream_t s) {
  ssor_apply_run<T, 0>(op, r, z, nullptr, st, nullptr, s);
}

template <typename T, int ND>
static int ssor_apply_dots(const SsorOp<T>& op, const T* r, T* z, const T* w1, const KrylovState* st, double* part,
                           hipStream_t s) {
  return ssor_apply_run<T, ND>(op, r, z, w1, st, part, s);
}

void ssor_apply(const SsorOp<double>& op, const double* r, double* z, const KrylovState* st, hipStream_t s) {
  ssor_apply_any<double>(op, r, z, st, s);
}
void ssor_apply_f32(const SsorOp<float>& op, const float* r, float* z, const KrylovState* st, hipStream_t s) {
  ssor_apply_any<float>(op, r, z, st, s);
}
void launch_diag_apply(int32_t n, const double* alu, const double* r, double* z, const KrylovState* st,
                       hipStream_t s) {
  k_diag<double, 0><<<vec_grid(n), kBlock, 0, s>>>(n, alu, r, z, nullptr, st, nullptr);
}
void launch_copy(int64_t n3, const double* r, double* z, const KrylovState* st, hipStream_t s) {
  k_copy_dot<<<vec_grid(n3), kBlock, 0, s>>>(n3, r, z, st, nullptr, 0);
}
void launch_zero(int64_t n, double* x, const KrylovState* st, hipStream_t s) {
  k_zero<<<vec_grid(n), kBlock, 0, s>>>(n, x, st);
}
void launch_add(int64_t n, const double* x, double* y, const KrylovState* st, hipStream_t s) {
  k_add<<<vec_grid(n), kBlock, 0, s>>>(n, x, y, st);
}
void launch_matresid_guarded(const Sell<double>& a, const double* x, const double* b, double* r,
                             const KrylovState* st, hipStream_t s, const MpcOp* mpc) {
  spmv_any<kEpiResid>(s, a, mpc, x, b, r, st, nullptr);
}

// ---- CG iteration pieces ------------------------------------------------------------
void cg_precond_ssor_rho(const SsorOp<double>& op, const double* r, double* z, KrylovState* st, double* part,
                         hipStream_t s) {
  const int nparts = ssor_apply_dots<double, 1>(op, r, z, nullptr, st, part, s);
  k_tail<kTailCgRho, 1><<<1, kTailThreads, 0, s>>>(st, part, nparts, nullptr, nullptr);
}
void cg_precond_diag_rho(int32_t n, const double* alu, const double* r, double* z, KrylovState* st, double* part,
                         hipStream_t s) {
  const int g = vec_grid(n);
  k_diag<double, 1><<<g, kBlock, 0, s>>>(n, alu, r, z, nullptr, st, part);
  k_tail<kTailCgRho, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void cg_precond_none_rho(int64_t n3, const double* r, double* z, KrylovState* st, double* part, hipStream_t s) {
  const int g = vec_grid(n3);
  k_copy_dot<<<g, kBlock, 0, s>>>(n3, r, z, st, part, 1);
  k_tail<kTailCgRho, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void cg_rho(int64_t n3, const double* r, const double* z, KrylovState* st, double* part, hipStream_t s) {
  const int g = vec_grid(n3);
  k_dot<<<g, kBlock, 0, s>>>(n3, r, z, st, part);
  k_tail<kTailCgRho, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void cg_xpay(int64_t n3, const double* z, double* p, const KrylovState* st, hipStream_t s) {
  k_cg_xpay<<<vec_grid(n3), kBlock, 0, s>>>(n3, z, p, st);
}
void cg_spmv_dot(const Sell<double>& a, const double* p, double* q, KrylovState* st, double* part, hipStream_t s,
                 hipEvent_t after, const MpcOp* mpc) {
  const int g = spmv_any<kEpiDotX>(s, a, mpc, p, nullptr, q, st, part);
  if (after) (void)hipEventRecord(after, s);
  k_tail<kTailCgC1, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void cg_update(int64_t n3, const double* p, const double* q, double* x, double* r, KrylovState* st, double* part,
               double* hist, hipStream_t s) {
  const int g = vec_grid(n3);
  k_cg_update<<<g, kBlock, 0, s>>>(n3, p, q, x, r, st, part);
  k_tail<kTailDnrm2, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, hist);
}
// only the non-periodic iterations produce the dnrm2 partials; the tail is
// guarded the same way
void cg_true_resid(const Sell<double>& a, const double* x, const double* b, double* r, KrylovState* st, double* part,
                   double* hist, hipStream_t s, const MpcOp* mpc) {
  const int g = spmv_any<kEpiTrue>(s, a, mpc, x, b, r, st, part);
  k_tail_true<<<1, kTailThreads, 0, s>>>(st, part, g, hist);
}

// ---- BiCGSTAB iteration pieces (hecmw_solver_BiCGSTAB.f90:164-275) --------------------
void bi_rho(int64_t n3, const double* r, const double* rt, KrylovState* st, double* part, hipStream_t s) {
  const int g = vec_grid(n3);
  k_dot<<<g, kBlock, 0, s>>>(n3, r, rt, st, part);
  k_tail<kTailBiRho, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void bi_p(int64_t n3, const double* r, const double* v, double* p, const KrylovState* st, hipStream_t s) {
  k_bi_p<<<vec_grid(n3), kBlock, 0, s>>>(n3, r, v, p, st);
}
void bi_spmv_c2(const Sell<double>& a, const double* pt, const double* rt, double* v, KrylovState* st, double* part,
                hipStream_t s, hipEvent_t after, const MpcOp* mpc) {
  const int g = spmv_any<kEpiDotW>(s, a, mpc, pt, rt, v, st, part);
  if (after) (void)hipEventRecord(after, s);
  k_tail<kTailBiC2, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void bi_s(int64_t n3, const double* v, const double* r, double* sv, const KrylovState* st, hipStream_t s) {
  k_bi_s<<<vec_grid(n3), kBlock, 0, s>>>(n3, v, r, sv, st);
}
void bi_spmv_omega(const Sell<double>& a, const double* stv, const double* sv, double* t, KrylovState* st,
                   double* part, hipStream_t s, hipEvent_t after, const MpcOp* mpc) {
  const int g = spmv_any<kEpiDot2W>(s, a, mpc, stv, sv, t, st, part);
  if (after) (void)hipEventRecord(after, s);
  k_tail<kTailBiOmega, 2><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void bi_update(int64_t n3, const double* pt, const double* stv, const double* t, const double* sv, double* x,
               double* r, KrylovState* st, double* part, double* hist, hipStream_t s) {
  const int g = vec_grid(n3);
  k_bi_update<<<g, kBlock, 0, s>>>(n3, pt, stv, t, sv, x, r, st, part);
  k_tail<kTailDnrm2, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, hist);
}
// generic preconditioner application z = M^-1 r for BiCGSTAB (no fused dot)
void precond_apply(int32_t kind, const SsorOp<double>& op, const double* alu_diag, int32_t n, const double* r,
                   double* z, const KrylovState* st, hipStream_t s) {
  if (kind == 0) {
    launch_copy(3 * (int64_t)n, r, z, st, s);
  } else if (kind == 3) {
    launch_diag_apply(n, alu_diag, r, z, st, s);
  } else {
    ssor_apply(op, r, z, st, s);
  }
}

// ---- diagonal scaling ------------------------------------------------------------------
void launch_make_scale(int32_t n, int32_t np, const double* D, double* sc, hipStream_t s) {
  k_make_scale<<<vec_grid(np), kBlock, 0, s>>>(n, np, D, sc);
}
void launch_scale_raw(bool fw, int32_t np, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                      const int32_t* itemU, double* D, double* AL, double* AU, const double* sc, hipStream_t s) {
  if (fw)
    k_scale_raw<true><<<vec_grid(np), kBlock, 0, s>>>(np, indexL, itemL, indexU, itemU, D, AL, AU, sc);
  else
    k_scale_raw<false><<<vec_grid(np), kBlock, 0, s>>>(np, indexL, itemL, indexU, itemU, D, AL, AU, sc);
}
void launch_scale_vec(bool mul, int64_t n, double* v, const double* sc, hipStream_t s) {
  if (mul)
    k_scale_vec<true><<<vec_grid(n), kBlock, 0, s>>>(n, v, sc);
  else
    k_scale_vec<false><<<vec_grid(n), kBlock, 0, s>>>(n, v, sc);
}

// ---- flexible CG with FP32 operators ------------------------------------------------
void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t s) {
  int64_t g = (n + kBlock - 1) / kBlock;
  g = g < 1 ? 1 : (g > 8192 ? 8192 : g);
  k_to_f32<<<(int)g, kBlock, 0, s>>>(n, x, y);
}
void precond_apply_f32(int32_t kind, const SsorOp<float>& op, const float* alu_diag, int32_t n, const float* r,
                       float* z, const KrylovState* st, hipStream_t s) {
  if (kind == 0) {
    k_copy_dots_f32<<<vec_grid(3 * (int64_t)n), kBlock, 0, s>>>(3 * (int64_t)n, r, z, nullptr, st, nullptr, 0);
  } else if (kind == 3) {
    k_diag<float, 0><<<vec_grid(n), kBlock, 0, s>>>(n, alu_diag, r, z, nullptr, st, nullptr);
  } else {
    ssor_apply_f32(op, r, z, st, s);
  }
}
void fcg_precond_dots(int32_t kind, const SsorOp<float>& op, const float* alu_diag, int32_t n, const float* r,
                      float* z, const float* q, KrylovState* st, double* part, hipStream_t s) {
  int nparts = 0;
  if (kind == 0) {
    nparts = vec_grid(3 * (int64_t)n);
    k_copy_dots_f32<<<nparts, kBlock, 0, s>>>(3 * (int64_t)n, r, z, q, st, part, 2);
  } else if (kind == 3) {
    nparts = vec_grid(n);
    k_diag<float, 2><<<nparts, kBlock, 0, s>>>(n, alu_diag, r, z, q, st, part);
  } else {
    nparts = ssor_apply_dots<float, 2>(op, r, z, q, st, part, s);
  }
  k_tail<kTailFcgBeta, 2><<<1, kTailThreads, 0, s>>>(st, part, nparts, nullptr, nullptr);
}
void fcg_alpha(KrylovState* st, const double* part, int nparts, hipStream_t s) {
  k_tail<kTailFcgAlpha, 2><<<1, kTailThreads, 0, s>>>(st, part, nparts, nullptr, nullptr);
}
void fcg_spmv_store(const Sell<float>& a32, const Sell<double>& a64, bool fp64_values, const float* x, float* y,
                    const KrylovState* st, hipStream_t s) {
  const int g = spmv_grid(a64.nslices);
  if (fp64_values)
    k_spmv_mixed<false><<<g, kBlock, 0, s>>>(a64, x, nullptr, y, st, nullptr);
  else
    k_spmv_f32<false><<<g, kBlock, 0, s>>>(a32, x, nullptr, y, st, nullptr);
}
void fcg_beta(KrylovState* st, const double* part, int nparts, hipStream_t s) {
  k_tail<kTailFcgBeta, 2><<<1, kTailThreads, 0, s>>>(st, part, nparts, nullptr, nullptr);
}
void fcg_p(int64_t n3, const float* z, float* p, const KrylovState* st, hipStream_t s) {
  k_fcg_p<<<vec_grid(n3), kBlock, 0, s>>>(n3, z, p, st);
}
void fcg_spmv(const Sell<float>& a, const float* p, const float* r, float* q, KrylovState* st, double* part,
              hipStream_t s, hipEvent_t after) {
  const int g = spmv_grid(a.nslices);
  k_spmv_f32<true><<<g, kBlock, 0, s>>>(a, p, r, q, st, part);
  if (after) (void)hipEventRecord(after, s);
  k_tail<kTailFcgAlpha, 2><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void fcg_spmv_mixed(const Sell<double>& a, const float* p, const float* r, float* q, KrylovState* st, double* part,
                    hipStream_t s, hipEvent_t after) {
  const int g = spmv_grid(a.nslices);
  k_spmv_mixed<true><<<g, kBlock, 0, s>>>(a, p, r, q, st, part);
  if (after) (void)hipEventRecord(after, s);
  k_tail<kTailFcgAlpha, 2><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, nullptr);
}
void fcg_update(int64_t n3, const float* p, const float* q, double* x, float* r, KrylovState* st, double* part,
                double* hist, hipStream_t s) {
  const int g = vec_grid(n3);
  k_fcg_update<<<g, kBlock, 0, s>>>(n3, p, q, x, r, st, part);
  k_tail<kTailFcgRes, 1><<<1, kTailThreads, 0, s>>>(st, part, g, nullptr, hist);
}
void fcg_reliable(const Sell<double>& a, const double* x, const double* b, float* r, KrylovState* st, double* part,
                  double* hist, hipStream_t s) {
  const int g = spmv_grid(a.nslices);
  k_spmv_rel<<<g, kBlock, 0, s>>>(a, x, b, r, st, part);
  k_tail_fcg_rel<<<1, kTailThreads, 0, s>>>(st, part, g, hist);
}
// r32 = b - A x (FP64) and |r|^2 stored to *rr (start of a flexible CG solve)
void fcg_init_resid(const Sell<double>& a, const double* x, const double* b, float* r, double* rr, KrylovState* st,
                    double* part, hipStream_t s) {
  const int g = spmv_grid(a.nslices);
  k_spmv_rel<<<g, kBlock, 0, s>>>(a, x, b, r, nullptr, part);
  k_tail<kTailStore, 1><<<1, kTailThreads, 0, s>>>(st, part, g, rr, nullptr);
}

}  // namespace hecmw33

namespace hecmw33 {
void launch_to_bf16(int64_t n, const float* a, uint16_t* b, hipStream_t s) {
  k_to_bf16<<<vec_grid(n), kBlock, 0, s>>>(n, a, b);
}
}  // namespace hecmw33
