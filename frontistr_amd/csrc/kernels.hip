// HIP kernels of the 3x3-block Krylov solver for gfx950 (MI355X).
//
// Everything here is HBM-bandwidth bound (0.23 flop/byte in FP64), so the
// design goals are: every wavefront load of matrix data is one contiguous
// segment (SELL-64 planes), x/z gathers stay inside an XCD's L2 (blocks are
// remapped so each XCD sweeps a contiguous run of slices), no host round trip
// inside an iteration (scalars live in KrylovState), and reductions are
// deterministic (per-block partials summed in a fixed order by the last block).
//
// FP64 row arithmetic is written with contraction off and in the reference's
// operation order, so SpMV, SSOR and DIAG results are bitwise those of the
// reference CPU loops (hecmw_solver_las_33.F90:281-308,
// hecmw_precond_SSOR_33.F90:302-405, hecmw_precond_DIAG_33.F90:115-140).
#include <hip/hip_runtime.h>
#include <cmath>
#include <algorithm>
#include <cstdint>

#include "device.h"

namespace hecmw33 {

namespace {

constexpr int32_t kErrDivergeMat = 3002;
constexpr int32_t kErrDivergePc = 3003;
constexpr int32_t kErrDivergeNan = 3004;
constexpr int32_t kErrNoconv = 3001;

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

// Block-reduce K values, publish one partial per block, and let the last block
// to arrive sum all partials in a fixed order.  Returns true in every thread
// of the last block, with tot[] holding the totals.
template <int K>
__device__ bool reduce_publish(double (&v)[K], double* part, int pstride, int pbase, int nparts_total,
                               bool finish, uint32_t* ticket, int nblk, double (&tot)[K]) {
  __shared__ double red[K][kWavesPerBlock];
  __shared__ int is_last;
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
      part[k * pstride + pbase + blockIdx.x] = s;
    }
    is_last = 0;
    if (finish) {
      __threadfence();
      const uint32_t t = atomicAdd(ticket, 1u);
      is_last = (t == (uint32_t)(nblk - 1));
    }
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts_total; i += blockDim.x) {
      s += __builtin_nontemporal_load(part + k * pstride + i);
    }
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) red[k][wave] = s;
    __syncthreads();
    double t = red[k][0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) t += red[k][w];
    tot[k] = t;
  }
  if (threadIdx.x == 0) *ticket = 0u;
  return true;
}

// ---- scalar logic run by one thread of the last block ------------------------
enum Tail : int {
  kTailStore = 0,      // out[0..K) = totals
  kTailCgRho = 1,      // hecmw_solver_CG.f90:179-207
  kTailCgC1 = 2,       // :221-232
  kTailDnrm2 = 3,      // :248-250, 267 (recurrence residual)
  kTailTrue = 4,       // :242-275 (true residual, convergence, loop end)
  kTailBiRho = 5,      // hecmw_solver_BiCGSTAB.f90:169-176
  kTailBiC2 = 6,       // :200-202
  kTailBiOmega = 7,    // :227-234
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
    st->dnrm2 = tot[0];
    st->resid = sqrt(tot[0] / st->bnrm2);
    if (hist) hist[st->iter - 1] = st->resid;
    if (st->resid <= st->tol) st->recheck = 1;
  } else if (TAIL == kTailTrue) {
    st->dnrm2 = tot[0];
    st->resid = sqrt(tot[0] / st->bnrm2);
    const bool periodic = (st->iter % st->nrec) == 0;
    if (periodic && hist) hist[st->iter - 1] = st->resid;
    if (st->resid <= st->tol) st->done = 1;
    st->recheck = 0;
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
  }
}

// End of one Krylov iteration (hecmw_solver_CG.f90:275-277): NOCONV at MAXIT
// (the Fortran DO variable then reads MAXIT+1), else rho1 = rho, iter += 1.
__device__ void end_iteration(KrylovState* st) {
  if (st->done) return;
  if (st->iter >= st->maxit) {
    st->error = kErrNoconv;
    st->done = 1;
    st->iter = st->maxit + 1;
  } else {
    st->rho1 = st->rho;
    st->iter += 1;
  }
}

// ---- SELL-64 block row product -----------------------------------------------
// FP64: reference order, D block first then L then U in storage order,
// y = ((y + a1*x1) + a2*x2) + a3*x3 per row, no fused multiply-add.
template <typename VT, typename XT>
__device__ __forceinline__ void sell_row(const Sell<VT>& a, int s, int lane, const XT* __restrict__ x, double (&y)[3],
                                         double (&xo)[3]) {
#pragma clang fp contract(off)
  const int base = a.ptr[s];
  const int len = a.len[s * kWave + lane];
  y[0] = y[1] = y[2] = 0.0;
  xo[0] = xo[1] = xo[2] = 0.0;
  if (len == 0) return;
  const int32_t* __restrict__ colp = a.col + (int64_t)base * kWave + lane;
  const VT* __restrict__ vp = a.val + (int64_t)base * 9 * kWave + lane;
  {
    const int c = colp[0];
    const double x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
    xo[0] = x0;
    xo[1] = x1;
    xo[2] = x2;
    double v[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) v[e] = vp[e * kWave];
    y[0] = v[0] * x0 + v[1] * x1 + v[2] * x2;
    y[1] = v[3] * x0 + v[4] * x1 + v[5] * x2;
    y[2] = v[6] * x0 + v[7] * x1 + v[8] * x2;
  }
#pragma unroll 2
  for (int k = 1; k < len; ++k) {
    const int c = colp[(int64_t)k * kWave];
    const double x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
    const VT* __restrict__ vk = vp + (int64_t)k * 9 * kWave;
    double v[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) v[e] = vk[e * kWave];
    y[0] = y[0] + v[0] * x0 + v[1] * x1 + v[2] * x2;
    y[1] = y[1] + v[3] * x0 + v[4] * x1 + v[5] * x2;
    y[2] = y[2] + v[6] * x0 + v[7] * x1 + v[8] * x2;
  }
}

// FP32 values and vectors, FP32 accumulation with FMA (mixed-precision inner
// solves only).
__device__ __forceinline__ void sell_row_f32(const Sell<float>& a, int s, int lane, const float* __restrict__ x,
                                             float (&y)[3]) {
  const int base = a.ptr[s];
  const int len = a.len[s * kWave + lane];
  y[0] = y[1] = y[2] = 0.0f;
  const int32_t* __restrict__ colp = a.col + (int64_t)base * kWave + lane;
  const float* __restrict__ vp = a.val + (int64_t)base * 9 * kWave + lane;
#pragma unroll 2
  for (int k = 0; k < len; ++k) {
    const int c = colp[(int64_t)k * kWave];
    const float x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
    const float* __restrict__ vk = vp + (int64_t)k * 9 * kWave;
    float v[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) v[e] = vk[e * kWave];
    y[0] = fmaf(v[2], x2, fmaf(v[1], x1, fmaf(v[0], x0, y[0])));
    y[1] = fmaf(v[5], x2, fmaf(v[4], x1, fmaf(v[3], x0, y[1])));
    y[2] = fmaf(v[8], x2, fmaf(v[7], x1, fmaf(v[6], x0, y[2])));
  }
}

// Slices handled by this wavefront: contiguous chunk after XCD remapping.
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

// ---- SpMV kernels --------------------------------------------------------------
enum SpmvMode : int {
  kSpmvPlain = 0,   // y = A x
  kSpmvResid = 1,   // y = b - A x
};

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(Sell<double> a, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ y) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(a.nslices);
  for (int s = sr.begin; s < sr.end; ++s) {
    double acc[3], xo[3];
    sell_row(a, s, lane, x, acc, xo);
    const int64_t row = (int64_t)s * kWave + lane;
    if (row < a.nrows) {
      if (MODE == kSpmvResid) {
        y[3 * row] = b[3 * row] - acc[0];
        y[3 * row + 1] = b[3 * row + 1] - acc[1];
        y[3 * row + 2] = b[3 * row + 2] - acc[2];
      } else {
        y[3 * row] = acc[0];
        y[3 * row + 1] = acc[1];
        y[3 * row + 2] = acc[2];
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_spmv_f32(Sell<float> a, const float* __restrict__ x,
                                                       float* __restrict__ y) {
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(a.nslices);
  for (int s = sr.begin; s < sr.end; ++s) {
    float acc[3];
    sell_row_f32(a, s, lane, x, acc);
    const int64_t row = (int64_t)s * kWave + lane;
    if (row < a.nrows) {
      y[3 * row] = acc[0];
      y[3 * row + 1] = acc[1];
      y[3 * row + 2] = acc[2];
    }
  }
}

// CG: q = A p and the partial sums of p.q; tail computes alpha.
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dot(Sell<double> a, const double* __restrict__ p,
                                                          double* __restrict__ q, KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st->done) return;
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(a.nslices);
  double d[1] = {0.0};
  for (int s = sr.begin; s < sr.end; ++s) {
    double acc[3], xo[3];
    sell_row(a, s, lane, p, acc, xo);
    const int64_t row = (int64_t)s * kWave + lane;
    if (row < a.nrows) {
      q[3 * row] = acc[0];
      q[3 * row + 1] = acc[1];
      q[3 * row + 2] = acc[2];
      d[0] += xo[0] * acc[0] + xo[1] * acc[1] + xo[2] * acc[2];
    }
  }
  double tot[1];
  if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[0], gridDim.x, tot) &&
      threadIdx.x == 0)
    run_tail<kTailCgC1, 1>(st, tot, nullptr, nullptr);
}

// True residual r = b - A x with its squared norm; runs only in iterations
// that recompute the residual (every nrec-th, or a convergence recheck).
// Ends the iteration.
__global__ __launch_bounds__(kBlock) void k_true_resid(Sell<double> a, const double* __restrict__ x,
                                                         const double* __restrict__ b, double* __restrict__ r,
                                                         KrylovState* st, double* part, double* hist) {
#pragma clang fp contract(off)
  if (st->done) return;
  const bool active = (st->iter % st->nrec) == 0 || st->recheck;
  if (!active) {
    // Every block has read the state above before the last one to arrive
    // advances the iteration.
    if (threadIdx.x == 0) {
      __threadfence();
      const uint32_t t = atomicAdd(&st->ticket[1], 1u);
      if (t == (uint32_t)(gridDim.x - 1)) {
        __threadfence();
        end_iteration(st);
        st->ticket[1] = 0u;
      }
    }
    return;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(a.nslices);
  double d[1] = {0.0};
  for (int s = sr.begin; s < sr.end; ++s) {
    double acc[3], xo[3];
    sell_row(a, s, lane, x, acc, xo);
    const int64_t row = (int64_t)s * kWave + lane;
    if (row < a.nrows) {
      const double r0 = b[3 * row] - acc[0], r1 = b[3 * row + 1] - acc[1], r2 = b[3 * row + 2] - acc[2];
      r[3 * row] = r0;
      r[3 * row + 1] = r1;
      r[3 * row + 2] = r2;
      d[0] += r0 * r0 + r1 * r1 + r2 * r2;
    }
  }
  double tot[1];
  if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[1], gridDim.x, tot) &&
      threadIdx.x == 0) {
    run_tail<kTailTrue, 1>(st, tot, nullptr, hist);
    end_iteration(st);
  }
}

// ---- SSOR sweeps (one launch per colour) ---------------------------------------
// Forward, colour rows independent: z[i] = LU^-1 (r[i] - sum_{j in L'(i)} L'_ij z[j])
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ssor_fwd(Sell<T> L, const int32_t* __restrict__ row_old,
                                                       const T* __restrict__ alu, int s0, int ns,
                                                       const T* __restrict__ rin, T* __restrict__ z,
                                                       const KrylovState* st) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(ns);
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kWave + lane;
    const int iold = row_old[slot];
    if (iold < 0) continue;
    T sw0 = rin[3 * (int64_t)iold], sw1 = rin[3 * (int64_t)iold + 1], sw2 = rin[3 * (int64_t)iold + 2];
    const int len = L.len[slot];
    const int base = L.ptr[s];
    const int32_t* __restrict__ colp = L.col + (int64_t)base * kWave + lane;
    const T* __restrict__ vp = L.val + (int64_t)base * 9 * kWave + lane;
#pragma unroll 2
    for (int k = 0; k < len; ++k) {
      const int c = colp[(int64_t)k * kWave];
      const T x0 = z[3 * (int64_t)c], x1 = z[3 * (int64_t)c + 1], x2 = z[3 * (int64_t)c + 2];
      const T* __restrict__ vk = vp + (int64_t)k * 9 * kWave;
      T v[9];
#pragma unroll
      for (int e = 0; e < 9; ++e) v[e] = vk[e * kWave];
      sw0 = sw0 - v[0] * x0 - v[1] * x1 - v[2] * x2;
      sw1 = sw1 - v[3] * x0 - v[4] * x1 - v[5] * x2;
      sw2 = sw2 - v[6] * x0 - v[7] * x1 - v[8] * x2;
    }
    const T* __restrict__ f = alu + (int64_t)s * 9 * kWave + lane;
    T a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = f[e * kWave];
    T x0 = sw0, x1 = sw1, x2 = sw2;
    x1 = x1 - a[3] * x0;
    x2 = x2 - a[6] * x0 - a[7] * x1;
    x2 = a[8] * x2;
    x1 = a[4] * (x1 - a[5] * x2);
    x0 = a[0] * (x0 - a[2] * x2 - a[1] * x1);
    z[3 * (int64_t)iold] = x0;
    z[3 * (int64_t)iold + 1] = x1;
    z[3 * (int64_t)iold + 2] = x2;
  }
}

// Backward, colours in reverse: z[i] -= LU^-1 sum_{j in U'(i), descending} U'_ij z[j].
// With RHO, also accumulates r.z over this colour's rows (their z is final
// now) into part[pbase + block]; the launch with `finish` sums every colour's
// partials and runs the CG rho logic.
template <typename T, bool RHO>
__global__ __launch_bounds__(kBlock) void k_ssor_bwd(Sell<T> U, const int32_t* __restrict__ row_old,
                                                       const T* __restrict__ alu, int s0, int ns,
                                                       T* __restrict__ z, const double* __restrict__ r,
                                                       KrylovState* st, double* part, int pbase, int ptotal,
                                                       int finish) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int lane = threadIdx.x & (kWave - 1);
  const SliceRange sr = wave_slices(ns);
  double d[1] = {0.0};
  for (int si = sr.begin; si < sr.end; ++si) {
    const int s = s0 + si;
    const int64_t slot = (int64_t)s * kWave + lane;
    const int iold = row_old[slot];
    if (iold < 0) continue;
    T sw0 = 0, sw1 = 0, sw2 = 0;
    const int len = U.len[slot];
    const int base = U.ptr[s];
    const int32_t* __restrict__ colp = U.col + (int64_t)base * kWave + lane;
    const T* __restrict__ vp = U.val + (int64_t)base * 9 * kWave + lane;
#pragma unroll 2
    for (int k = len - 1; k >= 0; --k) {
      const int c = colp[(int64_t)k * kWave];
      const T x0 = z[3 * (int64_t)c], x1 = z[3 * (int64_t)c + 1], x2 = z[3 * (int64_t)c + 2];
      const T* __restrict__ vk = vp + (int64_t)k * 9 * kWave;
      T v[9];
#pragma unroll
      for (int e = 0; e < 9; ++e) v[e] = vk[e * kWave];
      sw0 = sw0 + v[0] * x0 + v[1] * x1 + v[2] * x2;
      sw1 = sw1 + v[3] * x0 + v[4] * x1 + v[5] * x2;
      sw2 = sw2 + v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    const T* __restrict__ f = alu + (int64_t)s * 9 * kWave + lane;
    T a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = f[e * kWave];
    T x0 = sw0, x1 = sw1, x2 = sw2;
    x1 = x1 - a[3] * x0;
    x2 = x2 - a[6] * x0 - a[7] * x1;
    x2 = a[8] * x2;
    x1 = a[4] * (x1 - a[5] * x2);
    x0 = a[0] * (x0 - a[2] * x2 - a[1] * x1);
    const T z0 = z[3 * (int64_t)iold] - x0, z1 = z[3 * (int64_t)iold + 1] - x1, z2 = z[3 * (int64_t)iold + 2] - x2;
    z[3 * (int64_t)iold] = z0;
    z[3 * (int64_t)iold + 1] = z1;
    z[3 * (int64_t)iold + 2] = z2;
    if (RHO) {
      d[0] += r[3 * (int64_t)iold] * (double)z0 + r[3 * (int64_t)iold + 1] * (double)z1 +
              r[3 * (int64_t)iold + 2] * (double)z2;
    }
  }
  if (RHO) {
    double tot[1];
    if (reduce_publish<1>(d, part, kMaxPartials, pbase, ptotal, finish != 0, &st->ticket[2], gridDim.x, tot) &&
        threadIdx.x == 0)
      run_tail<kTailCgRho, 1>(st, tot, nullptr, nullptr);
  }
}

// ---- DIAG (block Jacobi) --------------------------------------------------------
template <bool RHO>
__global__ __launch_bounds__(kBlock) void k_diag(int32_t n, const double* __restrict__ alu,
                                                   const double* __restrict__ r, double* __restrict__ z,
                                                   KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st && st->done) return;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blk * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t s = i / kWave, lane = i % kWave;
    const double* f = alu + s * 9 * kWave + lane;
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = f[e * kWave];
    const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
    double x0 = r0, x1 = r1, x2 = r2;
    x1 = x1 - a[3] * x0;
    x2 = x2 - a[6] * x0 - a[7] * x1;
    x2 = a[8] * x2;
    x1 = a[4] * (x1 - a[5] * x2);
    x0 = a[0] * (x0 - a[2] * x2 - a[1] * x1);
    z[3 * i] = x0;
    z[3 * i + 1] = x1;
    z[3 * i + 2] = x2;
    if (RHO) d[0] += r0 * x0 + r1 * x1 + r2 * x2;
  }
  if (RHO) {
    double tot[1];
    if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[2], gridDim.x, tot) &&
        threadIdx.x == 0)
      run_tail<kTailCgRho, 1>(st, tot, nullptr, nullptr);
  }
}

// ---- vector kernels -----------------------------------------------------------
int vec_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  return g < 1 ? 1 : (int)g;
}

// z = r (no preconditioning) with rho = r.z
__global__ __launch_bounds__(kBlock) void k_copy_rho(int64_t n, const double* __restrict__ r, double* __restrict__ z,
                                                       KrylovState* st, double* part) {
#pragma clang fp contract(off)
  if (st->done) return;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double v = r[i];
    z[i] = v;
    d[0] += v * v;
  }
  double tot[1];
  if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[2], gridDim.x, tot) &&
      threadIdx.x == 0)
    run_tail<kTailCgRho, 1>(st, tot, nullptr, nullptr);
}

// generic dot with a selectable tail
template <int TAIL>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                  KrylovState* st, double* part, double* out, int guard) {
#pragma clang fp contract(off)
  if (guard && st->done) return;
  double d[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[0] += x[i] * y[i];
  double tot[1];
  if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[3], gridDim.x, tot) &&
      threadIdx.x == 0)
    run_tail<TAIL, 1>(st, tot, out, nullptr);
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

// x += alpha p; r -= alpha q (hecmw_axpy_R: Y = a*X + Y) and r.r, except in
// iterations that recompute the true residual (there only x is updated).
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, const double* __restrict__ p,
                                                        const double* __restrict__ q, double* __restrict__ x,
                                                        double* __restrict__ r, KrylovState* st, double* part,
                                                        double* hist) {
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
  if (periodic) return;
  double tot[1];
  if (reduce_publish<1>(d, part, kMaxPartials, 0, gridDim.x, true, &st->ticket[3], gridDim.x, tot) &&
      threadIdx.x == 0)
    run_tail<kTailDnrm2, 1>(st, tot, nullptr, hist);
}

// ---- numeric setup ---------------------------------------------------------------
template <typename T>
__global__ void k_fill_values(Sell<T> a, const int64_t* __restrict__ src, const double* __restrict__ D,
                              const double* __restrict__ AL, const double* __restrict__ AU, int64_t nD, int64_t nAL) {
  const int64_t nslot = a.nslots;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nslot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t grp = t / kWave, lane = t % kWave;  // slot = grp*64 + lane
    const int64_t b = src[t];
    const double* blk = nullptr;
    if (b >= 0) blk = b < nD ? D + 9 * b : (b < nD + nAL ? AL + 9 * (b - nD) : AU + 9 * (b - nD - nAL));
    T* out = a.val + grp * 9 * kWave + lane;
#pragma unroll
    for (int e = 0; e < 9; ++e) out[e * kWave] = blk ? (T)blk[e] : (T)0;
  }
}

template <typename T>
__global__ void k_block_lu(int32_t nslots, const int32_t* __restrict__ row_src, const double* __restrict__ D,
                           double sigma, T* __restrict__ alu) {
#pragma clang fp contract(off)
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nslots; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t grp = t / kWave, lane = t % kWave;
    const int32_t i = row_src[t];
    double a[3][3];
    if (i >= 0) {
      const double* d = D + 9 * (int64_t)i;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = d[3 * r + c];
      a[0][0] = a[0][0] * sigma;
      a[1][1] = a[1][1] * sigma;
      a[2][2] = a[2][2] * sigma;
      // hecmw_precond_SSOR_33.F90:183-195: inverted pivots on the diagonal,
      // multipliers below, updated U above.
      for (int k = 0; k < 3; ++k) {
        a[k][k] = 1.0 / a[k][k];
        for (int r = k + 1; r < 3; ++r) {
          a[r][k] = a[r][k] * a[k][k];
          double pw[3];
          for (int c = k + 1; c < 3; ++c) pw[c] = a[r][c] - a[r][k] * a[k][c];
          for (int c = k + 1; c < 3; ++c) a[r][c] = pw[c];
        }
      }
    } else {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = 0.0;
    }
    T* out = alu + grp * 9 * kWave + lane;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out[(3 * r + c) * kWave] = (T)a[r][c];
  }
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------
void launch_spmv(const Sell<double>& a, const double* x, double* y, hipStream_t s) {
  k_spmv<kSpmvPlain><<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, x, nullptr, y);
}
void launch_spmv_f32(const Sell<float>& a, const float* x, float* y, hipStream_t s) {
  k_spmv_f32<<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, x, y);
}
void launch_matresid(const Sell<double>& a, const double* x, const double* b, double* r, hipStream_t s) {
  k_spmv<kSpmvResid><<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, x, b, r);
}
void launch_dot_store(int64_t n, const double* x, const double* y, double* out, KrylovState* st, double* part,
                      hipStream_t s) {
  k_dot<kTailStore><<<vec_grid(n), kBlock, 0, s>>>(n, x, y, st, part, out, 0);
}
void launch_fill_values(const Sell<double>& a, const int64_t* src, const double* D, const double* AL,
                        const double* AU, int64_t nD, int64_t nAL, hipStream_t s) {
  k_fill_values<double><<<4096, 256, 0, s>>>(a, src, D, AL, AU, nD, nAL);
}
void launch_fill_values_f32(const Sell<float>& a, const int64_t* src, const double* D, const double* AL,
                            const double* AU, int64_t nD, int64_t nAL, hipStream_t s) {
  k_fill_values<float><<<4096, 256, 0, s>>>(a, src, D, AL, AU, nD, nAL);
}
void launch_block_lu(int32_t nslots, const int32_t* row_src, const double* D, double sigma, double* alu,
                     hipStream_t s) {
  k_block_lu<double><<<2048, 256, 0, s>>>(nslots, row_src, D, sigma, alu);
}
void launch_block_lu_f32(int32_t nslots, const int32_t* row_src, const double* D, double sigma, float* alu,
                         hipStream_t s) {
  k_block_lu<float><<<2048, 256, 0, s>>>(nslots, row_src, D, sigma, alu);
}

void launch_ssor_fwd(const Sell<double>& L, const int32_t* row_old, const double* alu, int32_t s0, int32_t ns,
                     const double* rin, double* z, const KrylovState* st, hipStream_t s) {
  if (ns <= 0) return;
  k_ssor_fwd<double><<<spmv_grid(ns), kBlock, 0, s>>>(L, row_old, alu, s0, ns, rin, z, st);
}
void launch_ssor_bwd(const Sell<double>& U, const int32_t* row_old, const double* alu, int32_t s0, int32_t ns,
                     double* z, const KrylovState* st, hipStream_t s) {
  if (ns <= 0) return;
  k_ssor_bwd<double, false><<<spmv_grid(ns), kBlock, 0, s>>>(U, row_old, alu, s0, ns, z, nullptr,
                                                              const_cast<KrylovState*>(st), nullptr, 0, 0, 0);
}
void launch_ssor_fwd_f32(const Sell<float>& L, const int32_t* row_old, const float* alu, int32_t s0, int32_t ns,
                         const float* rin, float* z, hipStream_t s) {
  if (ns <= 0) return;
  k_ssor_fwd<float><<<spmv_grid(ns), kBlock, 0, s>>>(L, row_old, alu, s0, ns, rin, z, nullptr);
}
void launch_ssor_bwd_f32(const Sell<float>& U, const int32_t* row_old, const float* alu, int32_t s0, int32_t ns,
                         float* z, hipStream_t s) {
  if (ns <= 0) return;
  k_ssor_bwd<float, false><<<spmv_grid(ns), kBlock, 0, s>>>(U, row_old, alu, s0, ns, z, nullptr, nullptr, nullptr,
                                                             0, 0, 0);
}
void launch_diag_apply(int32_t n, const double* alu, const double* r, double* z, const KrylovState* st,
                       hipStream_t s) {
  k_diag<false><<<vec_grid(n), kBlock, 0, s>>>(n, alu, r, z, const_cast<KrylovState*>(st), nullptr);
}

// ---- CG iteration pieces ------------------------------------------------------------
void cg_precond_ssor_rho(const Sell<double>& L, const Sell<double>& U, const int32_t* row_old, const double* alu,
                         const int32_t* color_slice, int32_t ncolor, const double* r, double* z, KrylovState* st,
                         double* part, hipStream_t s) {
  for (int c = 0; c < ncolor; ++c)
    launch_ssor_fwd(L, row_old, alu, color_slice[c], color_slice[c + 1] - color_slice[c], r, z, st, s);
  // partial slots per colour launch, laid out in launch order; the grid of
  // each colour is capped so that all colours fit in the partials buffer
  const int gcap = std::max(1, kMaxPartials / std::max(1, ncolor));
  int ptotal = 0;
  for (int c = ncolor - 1; c >= 0; --c) ptotal += std::min(gcap, spmv_grid(color_slice[c + 1] - color_slice[c]));
  int pbase = 0;
  for (int c = ncolor - 1; c >= 0; --c) {
    const int ns = color_slice[c + 1] - color_slice[c];
    const int g = std::min(gcap, spmv_grid(ns));
    k_ssor_bwd<double, true><<<g, kBlock, 0, s>>>(U, row_old, alu, color_slice[c], ns, z, r, st, part, pbase, ptotal,
                                                  c == 0 ? 1 : 0);
    pbase += g;
  }
}
void cg_precond_diag_rho(int32_t n, const double* alu, const double* r, double* z, KrylovState* st, double* part,
                         hipStream_t s) {
  k_diag<true><<<vec_grid(n), kBlock, 0, s>>>(n, alu, r, z, st, part);
}
void cg_precond_none_rho(int64_t n3, const double* r, double* z, KrylovState* st, double* part, hipStream_t s) {
  k_copy_rho<<<vec_grid(n3), kBlock, 0, s>>>(n3, r, z, st, part);
}
void cg_rho(int64_t n3, const double* r, const double* z, KrylovState* st, double* part, hipStream_t s) {
  k_dot<kTailCgRho><<<vec_grid(n3), kBlock, 0, s>>>(n3, r, z, st, part, nullptr, 1);
}
void cg_xpay(int64_t n3, const double* z, double* p, const KrylovState* st, hipStream_t s) {
  k_cg_xpay<<<vec_grid(n3), kBlock, 0, s>>>(n3, z, p, st);
}
void cg_spmv_dot(const Sell<double>& a, const double* p, double* q, KrylovState* st, double* part, hipStream_t s) {
  k_cg_spmv_dot<<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, p, q, st, part);
}
void cg_update(int64_t n3, const double* p, const double* q, double* x, double* r, KrylovState* st, double* part,
               double* hist, hipStream_t s) {
  k_cg_update<<<vec_grid(n3), kBlock, 0, s>>>(n3, p, q, x, r, st, part, hist);
}
void cg_true_resid(const Sell<double>& a, const double* x, const double* b, double* r, KrylovState* st, double* part,
                   double* hist, hipStream_t s) {
  k_true_resid<<<spmv_grid(a.nslices), kBlock, 0, s>>>(a, x, b, r, st, part, hist);
}

}  // namespace hecmw33
