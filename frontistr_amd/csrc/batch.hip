// Batched multi-RHS solves: many independent 3x3-block systems, each with up
// to 16 right-hand sides, solved together by DIAG-preconditioned CG.
//
// The reference solves one system and one right-hand side per
// hecmw_solve_CG call (hecmw_solver_CG.f90:19-320).  Here every system of the
// batch is one diagonal block of a single block-diagonal operator, so one
// launch of each kernel serves all systems and all right-hand sides:
//
//   rows      : the systems' rows concatenated (system s owns rows
//               [row0[s], row0[s]+N_s)), merged BCSR per row in the
//               reference's SpMV order (D, then L, then U in storage order)
//   vectors   : [row][16 rhs][xyz]; a (system, rhs) pair is one CG instance
//   tiles     : runs of rows inside one system; one workgroup per tile writes
//               that tile's partial inner products for all 16 right-hand
//               sides, and one workgroup per system sums its tiles in a fixed
//               order (deterministic, no atomics)
//   precision : FP32 matrix, SpMM, DIAG and work vectors; FP64 x, inner
//               products and CG scalars; the residuals of a system's
//               instances are replaced by b - A x computed with the FP64
//               matrix whenever one of them has dropped by delta, every 50
//               iterations and before convergence is declared (flexible CG,
//               Notay's FCG(1) recurrences, as the single-system mixed path)
//   launches  : the kernels of kChunk iterations are captured once into a
//               hipGraph and replayed until every instance is done; tiles of
//               finished systems return at once
//   SpMM shape: lanes per row, right-hand sides per lane and blocks in flight
//               per lane are picked by timing the candidates on the loaded
//               matrix (autotuner)
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hecmw33.h"
#include "device.h"

namespace hecmw33 {
namespace {

constexpr int KS = 16;          // right-hand-side stride of every batch vector
constexpr int kBT = 256;        // threads per workgroup
constexpr int kTileRows = 128;  // rows per tile (at least; grown to cap the tile count)
constexpr int kMaxTiles = 65536;
constexpr int kChunk = 8;       // iterations per captured graph

constexpr int32_t kErrNoconv = 3001, kErrDivergeMat = 3002, kErrDivergePc = 3003, kErrDivergeNan = 3004;

// per-instance state, structure of arrays over nsys*16 instances
struct BState {
  double *bnrm2, *rho, *rho1, *zq, *pq, *rp, *alpha, *beta, *rr_ref, *resid;
  // gap: true / recurrence residual seen at the last replacement; the FP32
  // recurrence runs ahead of the FP64 residual near convergence, so it claims
  // convergence when resid * gap <= tol
  double* gap;
  int32_t *iter, *done, *error, *nindef, *rel;
  int32_t *sys_active, *sys_rel, *active_total;
  // solve parameters in device memory, so a captured graph serves any call:
  // prm = {tol, delta^2}, iprm = {maxit, nrec}
  double* prm;
  int32_t* iprm;
};

struct Tiles {
  const int32_t *row0, *nrows, *sys;  // per tile
  const int32_t *sys_tile0, *sys_ntile;
  int32_t ntiles;
};

// ---- persistent kernels over live-tile work queues -----------------------------------
// The per-tile kernels of an iteration run persistently: a grid of resident
// workgroups walks a list of the live tiles (those of the systems still
// running, or of those whose residuals are being replaced), rebuilt on the
// device after every tail, instead of one workgroup per tile of the whole
// batch, so finished systems cost nothing.  The list is dealt without atomics
// (a ticket counter per XCD measured 13% slower on the 48-system workload: its
// atomics serialise across the XCDs): runs of kQRun consecutive list entries
// go round-robin to the XCDs (workgroup b runs on XCD b % kNumXcd, the
// hardware's deal), so an XCD's workgroups share neighbouring rows whose x
// stays in its L2 while every system's tiles spread over all XCDs, and the
// workgroups of one XCD take that XCD's entries in turn.  Which workgroup
// computes a tile does not matter: every tile writes its own partials in a
// fixed order.
enum QList : int { kqActive = 0, kqRel = 1, kqLists = 2 };
constexpr int kQRun = 16;

struct Queue {
  int32_t* tile;  // [kqLists][ntiles] live tile ids, system by system
  int32_t* n;     // [kqLists] live tiles per list
};

// what one launch walks (tile == nullptr: one tile per workgroup over every
// tile, the layout of the set-up launches)
struct QView {
  const int32_t* tile;
  const int32_t* n;
};

__device__ __forceinline__ int spmm_tile();

template <class F>
__device__ __forceinline__ void walk_tiles(const QView& q, F&& body) {
  if (!q.tile) {
    body(spmm_tile());
    return;
  }
  const int n = *q.n, G = gridDim.x, b = blockIdx.x;
  if (G < kNumXcd) {
    for (int i = b; i < n; i += G) body(q.tile[i]);
    return;
  }
  // XCD x holds list entries ((j / kQRun) * kNumXcd + x) * kQRun + j % kQRun,
  // j = 0, 1, ...; its Gx workgroups b = x, x + 8, ... take j = b / 8 + k Gx
  const int x = b % kNumXcd, gx = (G - x + kNumXcd - 1) / kNumXcd;
  for (int j = b / kNumXcd;; j += gx) {
    // entries only grow along a workgroup's walk: the first one past the list ends it
    const int64_t i = ((int64_t)(j / kQRun) * kNumXcd + x) * kQRun + j % kQRun;
    if (i >= n) break;
    body(q.tile[i]);
  }
}

// Rebuild list LIST from the systems' flags (active: sys_active != 0; rel:
// sys_rel != 0) and, for the active list, publish the live instance count the
// host polls.
template <int LIST>
__global__ __launch_bounds__(1024) void kb_queue(Tiles T, int32_t nsys, BState S, Queue Q) {
  int32_t* out = Q.tile + (int64_t)LIST * T.ntiles;
  int32_t off = 0, live_inst = 0;
  for (int s = 0; s < nsys; ++s) {
    const int32_t a = S.sys_active[s];
    const bool sel = LIST == kqActive ? a != 0 : S.sys_rel[s] != 0;
    live_inst += a;
    if (!sel) continue;
    const int32_t t0 = T.sys_tile0[s], nt = T.sys_ntile[s];
    for (int i = threadIdx.x; i < nt; i += blockDim.x) out[off + i] = t0 + i;
    off += nt;
  }
  if (threadIdx.x == 0) {
    Q.n[LIST] = off;
    if (LIST == kqActive) *S.active_total = live_inst;
  }
}

// The set-up launches (one workgroup per tile) keep the hardware's round-robin
// deal over the XCDs: a contiguous per-XCD remap of every tile measured 27%
// slower on the 48-system workload (load imbalance across systems of very
// different sizes).

// Partials of ND inner products per right-hand side: thread tid holds the
// values of rhs k = tid % 16 (rows tid / 16, ...); fixed-order sum over the
// 16 row lanes, one partial per (dot, rhs, tile).
template <int ND>
__device__ __forceinline__ void tile_partials_rowlanes(const double (&acc)[ND], double* part, int pitch, int tile) {
  __shared__ double red[ND][kBT];
#pragma unroll
  for (int d = 0; d < ND; ++d) red[d][threadIdx.x] = acc[d];
  __syncthreads();
  if (threadIdx.x < ND * KS) {
    const int d = threadIdx.x / KS, k = threadIdx.x % KS;
    double s = 0.0;
    for (int r = 0; r < kBT / KS; ++r) s += red[d][r * KS + k];
    part[(int64_t)(d * KS + k) * pitch + tile] = s;
  }
}

// ---- preconditioner + fused dots: z = D^-1 r, r.z, z.q ----------------------------
// z of one (row, rhs) from its three residual entries (the reference's DIAG
// statements, hecmw_precond_DIAG_33.F90)
__device__ __forceinline__ void kb_jacobi(const float* __restrict__ alu, int64_t g, float r0v, float r1v, float r2v,
                                          float& x0, float& x1, float& x2) {
#pragma clang fp contract(off)
  float a[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) a[e] = alu[9 * g + e];
  x0 = r0v;
  x1 = r1v;
  x2 = r2v;
  x1 = x1 - a[3] * x0;
  x2 = x2 - a[6] * x0 - a[7] * x1;
  x2 = a[8] * x2;
  x1 = a[4] * (x1 - a[5] * x2);
  x0 = a[0] * (x0 - a[2] * x2 - a[1] * x1);
}

// guard 0: every active system; 2: only systems whose residuals were just
// replaced (their z is recomputed from the new r)
__global__ __launch_bounds__(kBT) void kb_diag(Tiles T, QView Qv, const float* __restrict__ alu,
                                               const float* __restrict__ r, float* __restrict__ z,
                                               const float* __restrict__ q, BState S, double* part, int guard) {
  walk_tiles(Qv, [&](int t) {
#pragma clang fp contract(off)
    const int s = T.sys[t];
    if (S.sys_active[s] == 0) return;
    if (guard == 2 && S.sys_rel[s] == 0) return;
    const int k = threadIdx.x % KS, rl = threadIdx.x / KS;
    const bool live = !S.done[s * KS + k];
    double acc[2] = {0.0, 0.0};
    const int r0 = T.row0[t], nr = T.nrows[t];
    for (int i = rl; i < nr && live; i += kBT / KS) {
      const int64_t g = r0 + i;
      const int64_t o = (g * KS + k) * 3;
      const float r0v = r[o], r1v = r[o + 1], r2v = r[o + 2];
      float x0, x1, x2;
      kb_jacobi(alu, g, r0v, r1v, r2v, x0, x1, x2);
      z[o] = x0;
      z[o + 1] = x1;
      z[o + 2] = x2;
      acc[0] += (double)r0v * x0 + (double)r1v * x1 + (double)r2v * x2;
      acc[1] += (double)q[o] * x0 + (double)q[o + 1] * x1 + (double)q[o + 2] * x2;
    }
    tile_partials_rowlanes<2>(acc, part, T.ntiles, t);
  });
}

// ---- p = z (first iteration of an instance) or z + beta p --------------------------
__global__ __launch_bounds__(kBT) void kb_p(Tiles T, QView Qv, const float* __restrict__ z, float* __restrict__ p,
                                            BState S) {
  walk_tiles(Qv, [&](int t) {
    const int s = T.sys[t];
    if (S.sys_active[s] == 0) return;
    const int k = threadIdx.x % KS, rl = threadIdx.x / KS;
    const int id = s * KS + k;
    if (S.done[id]) return;
    const bool first = S.iter[id] == 1;
    const double beta = S.beta[id];
    const int r0 = T.row0[t], nr = T.nrows[t];
    for (int i = rl; i < nr; i += kBT / KS) {
      const int64_t o = ((int64_t)(r0 + i) * KS + k) * 3;
#pragma unroll
      for (int d = 0; d < 3; ++d) p[o + d] = first ? z[o + d] : (float)((double)z[o + d] + beta * (double)p[o + d]);
    }
  });
}

// ---- dx += alpha p (FP32 increments of the FP64 x, folded into x before each
//      residual replacement by kb_fold), r = r - alpha q (FP32), r.r; then the
//      next iteration's preconditioner on the new r while it is in registers:
//      z = D^-1 r with r.z and z.q (q of this iteration, the next beta's
//      q_prev), the values and partials kb_diag would give
__global__ __launch_bounds__(kBT) void kb_update_diag(Tiles T, QView Qv, const float* __restrict__ alu,
                                                      const float* __restrict__ p, const float* __restrict__ q,
                                                      float* __restrict__ dx, float* __restrict__ r,
                                                      float* __restrict__ z, BState S, double* part_rr,
                                                      double* part_rz) {
  walk_tiles(Qv, [&](int t) {
    const int s = T.sys[t];
    if (S.sys_active[s] == 0) return;
    const int k = threadIdx.x % KS, rl = threadIdx.x / KS;
    const int id = s * KS + k;
    const bool live = !S.done[id];
    const double alpha = S.alpha[id];
    double acc[1] = {0.0}, acc2[2] = {0.0, 0.0};
    const int r0 = T.row0[t], nr = T.nrows[t];
    for (int i = rl; i < nr && live; i += kBT / KS) {
      const int64_t g = r0 + i;
      const int64_t o = (g * KS + k) * 3;
      float v[3], qv[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        qv[d] = q[o + d];
        dx[o + d] = (float)((double)dx[o + d] + alpha * (double)p[o + d]);
        v[d] = (float)((double)r[o + d] - alpha * (double)qv[d]);
        r[o + d] = v[d];
        acc[0] += (double)v[d] * (double)v[d];
      }
      float x0, x1, x2;
      kb_jacobi(alu, g, v[0], v[1], v[2], x0, x1, x2);
      z[o] = x0;
      z[o + 1] = x1;
      z[o + 2] = x2;
      {
#pragma clang fp contract(off)
        acc2[0] += (double)v[0] * x0 + (double)v[1] * x1 + (double)v[2] * x2;
        acc2[1] += (double)qv[0] * x0 + (double)qv[1] * x1 + (double)qv[2] * x2;
      }
    }
    tile_partials_rowlanes<1>(acc, part_rr, T.ntiles, t);
    tile_partials_rowlanes<2>(acc2, part_rz, T.ntiles, t);
  });
}

// ---- x += dx, dx = 0 for every row and right-hand side of the listed tiles'
//      systems: before a residual replacement (guard 2: the systems being
//      replaced) and at the end of the solve (guard 0: all)
__global__ __launch_bounds__(kBT) void kb_fold(Tiles T, QView Qv, float* __restrict__ dx, double* __restrict__ x,
                                               BState S, int guard) {
  walk_tiles(Qv, [&](int t) {
    const int s = T.sys[t];
    if (guard == 2 && S.sys_rel[s] == 0) return;
    const int64_t o0 = (int64_t)T.row0[t] * KS * 3, n = (int64_t)T.nrows[t] * KS * 3;
    for (int64_t i = threadIdx.x; i < n; i += kBT) {
      const float d = dx[o0 + i];
      if (d != 0.0f) {
        x[o0 + i] += (double)d;
        dx[o0 + i] = 0.0f;
      }
    }
  });
}

// ---- SpMM ----------------------------------------------------------------------------
// Tile of this workgroup for the SpMM: runs of kSpmmRun consecutive tiles are
// dealt round-robin to the 8 XCDs (workgroup b runs on XCD b % 8), so one XCD
// sweeps contiguous rows whose x neighbourhoods stay in its L2, while the
// tiles of any one system still spread over every XCD.  Tiles past the last
// whole round keep the identity map.
constexpr int kSpmmRun = 64;
__device__ __forceinline__ int spmm_tile() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int round = kNumXcd * kSpmmRun;
  const int whole = (nb / round) * round;
  if (b >= whole) return b;
  const int xcd = b % kNumXcd, l = b / kNumXcd;
  return ((l / kSpmmRun) * kNumXcd + xcd) * kSpmmRun + (l % kSpmmRun);
}

// A row is handled by LPR lanes: G = 16/RPL right-hand-side groups (RPL rhs
// per lane) times NPH = LPR/G phases that take every NPH-th block of the row;
// phases are summed with lane shuffles.  U blocks per lane are in flight.
enum BEpi : int {
  kbEpiQ = 0,      // FP32: q = A p, partials p.q, r.p
  kbEpiRes = 1,    // FP64: r32 = (float)(b - A x) for instances with rel set, partial |b - A x|^2
  kbEpiStore = 2,  // y = A x (either precision), no partials
};

template <typename T, int LPR, int RPL, int U, int EPI>
__device__ __forceinline__ void spmm_one_tile(int t, const Tiles& TL, const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ x, const void* __restrict__ w,
                                              void* __restrict__ y, const BState& S, double* part, int guard) {
#pragma clang fp contract(off)
  constexpr int G = KS / RPL, NPH = LPR / G, RPW = 64 / LPR;
  constexpr int ND = EPI == kbEpiQ ? 2 : (EPI == kbEpiRes ? 1 : 0);
  const int s = TL.sys[t];
  if (guard == 1 && S.sys_active[s] == 0) return;
  if (guard == 2 && S.sys_rel[s] == 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / LPR, l = lane % LPR, jg = l % G, ph = l / G;
  const int r0 = TL.row0[t], nr = TL.nrows[t];
  double acc[RPL][ND > 0 ? ND : 1];
#pragma unroll
  for (int j = 0; j < RPL; ++j)
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d) acc[j][d] = 0.0;
  const int npass = (nr + 4 * RPW - 1) / (4 * RPW);
  for (int pass = 0; pass < npass; ++pass) {
    const int i = pass * 4 * RPW + wave * RPW + slot;
    const bool valid = i < nr;
    const int64_t g = r0 + (valid ? i : 0);
    const int b0 = rowptr[g], b1 = valid ? rowptr[g + 1] : b0;
    T yv[RPL][3];
#pragma unroll
    for (int j = 0; j < RPL; ++j) yv[j][0] = yv[j][1] = yv[j][2] = (T)0;
    // a chunk of U blocks from bb; FULL (all U live) keeps every load and sum
    // out of conditions, so the compiler cannot move a block's loads into a
    // branch of their own; the row's last, partial chunk re-reads its last
    // block in the dead slots and masks them out
    auto chunk = [&](int bb, auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      int c[U];
      T a[U][9];
      T xv[U][RPL][3];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = bb + u * NPH;
        const int bc = (FULL || b < b1) ? b : b1 - 1;
        c[u] = col[bc];
#pragma unroll
        for (int e = 0; e < 9; ++e) a[u][e] = val[9 * (int64_t)bc + e];
#pragma unroll
        for (int j = 0; j < RPL; ++j)
#pragma unroll
          for (int d = 0; d < 3; ++d) xv[u][j][d] = x[((int64_t)c[u] * KS + jg * RPL + j) * 3 + d];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool live = FULL || bb + u * NPH < b1;
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          T n0, n1, n2;
          if constexpr (sizeof(T) == 4) {
            n0 = fmaf(a[u][2], xv[u][j][2], fmaf(a[u][1], xv[u][j][1], fmaf(a[u][0], xv[u][j][0], yv[j][0])));
            n1 = fmaf(a[u][5], xv[u][j][2], fmaf(a[u][4], xv[u][j][1], fmaf(a[u][3], xv[u][j][0], yv[j][1])));
            n2 = fmaf(a[u][8], xv[u][j][2], fmaf(a[u][7], xv[u][j][1], fmaf(a[u][6], xv[u][j][0], yv[j][2])));
          } else {
            // the reference's order: ((y + a1 x1) + a2 x2) + a3 x3, blocks in storage order
            n0 = yv[j][0] + a[u][0] * xv[u][j][0] + a[u][1] * xv[u][j][1] + a[u][2] * xv[u][j][2];
            n1 = yv[j][1] + a[u][3] * xv[u][j][0] + a[u][4] * xv[u][j][1] + a[u][5] * xv[u][j][2];
            n2 = yv[j][2] + a[u][6] * xv[u][j][0] + a[u][7] * xv[u][j][1] + a[u][8] * xv[u][j][2];
          }
          yv[j][0] = live ? n0 : yv[j][0];
          yv[j][1] = live ? n1 : yv[j][1];
          yv[j][2] = live ? n2 : yv[j][2];
        }
      }
    };
    int bb = b0 + ph;
    for (; bb + (U - 1) * NPH < b1; bb += NPH * U) chunk(bb, std::true_type{});
    if (bb < b1) chunk(bb, std::false_type{});
    if constexpr (NPH > 1) {
#pragma unroll
      for (int off = G; off < LPR; off <<= 1)
#pragma unroll
        for (int j = 0; j < RPL; ++j)
#pragma unroll
          for (int d = 0; d < 3; ++d) yv[j][d] += __shfl_xor(yv[j][d], off, 64);
    }
    if (valid && ph == 0) {
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        const int k = jg * RPL + j;
        const int64_t o = (g * KS + k) * 3;
        if constexpr (EPI == kbEpiQ) {
          const float* rr = static_cast<const float*>(w);
          float* q = static_cast<float*>(y);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            q[o + d] = (float)yv[j][d];
            const double pv = (double)x[o + d];
            acc[j][0] += pv * (double)yv[j][d];
            acc[j][ND - 1] += (double)rr[o + d] * pv;
          }
        } else if constexpr (EPI == kbEpiRes) {
          const double* b = static_cast<const double*>(w);
          float* r = static_cast<float*>(y);
          if (S.rel[s * KS + k]) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double v = b[o + d] - (double)yv[j][d];
              r[o + d] = (float)v;
              acc[j][0] += v * v;
            }
          }
        } else {
          T* yo = static_cast<T*>(y);
#pragma unroll
          for (int d = 0; d < 3; ++d) yo[o + d] = yv[j][d];
        }
      }
    }
  }
  if constexpr (ND > 0) {
    // thread (wave, slot, jg, ph=0) holds rhs jg*RPL+j; fixed-order sum over
    // the workgroup's rows for every (dot, rhs)
    __shared__ double red[ND][kBT][RPL];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int j = 0; j < RPL; ++j) red[d][threadIdx.x][j] = ph == 0 ? acc[j][d] : 0.0;
    __syncthreads();
    if (threadIdx.x < ND * KS) {
      const int d = threadIdx.x / KS, k = threadIdx.x % KS;
      const int jgk = k / RPL, j = k % RPL;
      double sum = 0.0;
      for (int m = 0; m < kBT / LPR; ++m) sum += red[d][m * LPR + jgk][j];
      part[(int64_t)(d * KS + k) * TL.ntiles + t] = sum;
    }
  }
}

template <typename T, int LPR, int RPL, int U, int EPI>
__global__ __launch_bounds__(kBT) void kb_spmm(Tiles TL, QView Qv, const int32_t* __restrict__ rowptr,
                                               const int32_t* __restrict__ col, const T* __restrict__ val,
                                               const T* __restrict__ x, const void* __restrict__ w,
                                               void* __restrict__ y, BState S, double* part, int guard) {
  walk_tiles(Qv, [&](int t) { spmm_one_tile<T, LPR, RPL, U, EPI>(t, TL, rowptr, col, val, x, w, y, S, part, guard); });
}

// ---- per-system tails: sum a system's tile partials, then the CG scalar logic ---------
enum BTail : int { kbInit = 0, kbBeta = 1, kbAlpha = 2, kbRes = 3, kbRel = 4 };

template <int TAIL, int ND>
__global__ __launch_bounds__(kBT) void kb_tail(Tiles T, BState S, const double* part) {
  const int s = blockIdx.x;
  if (TAIL != kbInit && S.sys_active[s] == 0) return;
  const int k = threadIdx.x % KS;
  const int t0 = T.sys_tile0[s], nt = T.sys_ntile[s];
  __shared__ double red[ND][kBT];
  __shared__ int32_t flag[KS];
  {
    // 16 threads per right-hand side sweep the system's tile partials, which
    // are contiguous per (dot, rhs), with 8 independent loads in flight; a
    // fixed order throughout
    constexpr int kSub = kBT / KS;
    const int kl = threadIdx.x / kSub, sub = threadIdx.x % kSub;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double* __restrict__ pp = part + (int64_t)(d * KS + kl) * T.ntiles + t0;
      double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int i = sub;
      for (; i + 7 * kSub < nt; i += 8 * kSub)
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += pp[i + u * kSub];
      for (; i < nt; i += kSub) a[0] += pp[i];
      red[d][threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
  }
  __syncthreads();
  if (threadIdx.x < KS) {
    double tot[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double a = 0.0;
      for (int m = 0; m < kBT / KS; ++m) a += red[d][k * (kBT / KS) + m];
      tot[d] = a;
    }
    const int id = s * KS + k;
    int32_t fl = 0;
    if (TAIL == kbInit) {
      // tot[0] = b.b, tot[1] = |b - A x0|^2
      S.bnrm2[id] = tot[0];
      S.rr_ref[id] = tot[ND - 1];
      S.rho[id] = S.rho1[id] = S.zq[id] = S.pq[id] = S.rp[id] = S.alpha[id] = S.beta[id] = 0.0;
      S.gap[id] = 1.0;
      S.iter[id] = 1;
      S.nindef[id] = 0;
      S.rel[id] = 0;
      S.error[id] = 0;
      S.resid[id] = tot[0] == 0.0 ? 0.0 : sqrt(tot[ND - 1] / tot[0]);
      // b = 0: x = 0 and nothing to do (hecmw_solver_CG.f90:135-141); the
      // caller zeroes x for such instances
      S.done[id] = (tot[0] == 0.0 || S.iprm[0] < 1) ? 1 : 0;
      fl = !S.done[id];
    } else if (!S.done[id]) {
      if (TAIL == kbBeta) {
        const double rho = tot[0];
        S.rho[id] = rho;
        S.zq[id] = tot[1];
        const int it = S.iter[id];
        if (rho == 0.0) {
          S.done[id] = 1;
        } else if (rho != rho) {
          S.error[id] = kErrDivergeNan;
          S.done[id] = 1;
        } else if (it > 1 && rho * S.rho1[id] <= 0.0) {
          S.nindef[id] += 1;
          if (S.nindef[id] >= 3) {
            S.error[id] = kErrDivergePc;
            S.done[id] = 1;
          }
        }
        if (!S.done[id]) S.beta[id] = it > 1 ? -S.zq[id] / S.pq[id] : 0.0;
      } else if (TAIL == kbAlpha) {
        const double pq = tot[0], rp = tot[1];
        S.pq[id] = pq;
        S.rp[id] = rp;
        if (pq <= 0.0) {
          S.error[id] = kErrDivergeMat;
          S.done[id] = 1;
        } else if (pq != pq || rp != rp) {
          S.error[id] = kErrDivergeNan;
          S.done[id] = 1;
        } else {
          S.alpha[id] = rp / pq;
        }
      } else if (TAIL == kbRes) {
        const double rr = tot[0];
        S.resid[id] = sqrt(rr / S.bnrm2[id]);
        S.rel[id] = S.resid[id] * S.gap[id] <= S.prm[0] || rr <= S.prm[1] * S.rr_ref[id] ||
                    (S.iter[id] % S.iprm[1]) == 0 || rr != rr;
        fl = S.rel[id];
      } else if (TAIL == kbRel) {
        if (S.rel[id]) {
          const double rr = tot[0];
          const double rec = S.resid[id];
          S.rr_ref[id] = rr;
          S.resid[id] = sqrt(rr / S.bnrm2[id]);
          if (rec > 0.0) S.gap[id] = fmin(1e3, fmax(1.0, S.resid[id] / rec));
          S.rel[id] = 0;
          if (S.resid[id] <= S.prm[0]) {
            S.done[id] = 1;
          } else if (rr != rr) {
            S.error[id] = kErrDivergeNan;
            S.done[id] = 1;
          }
        }
        if (!S.done[id]) {
          if (S.iter[id] >= S.iprm[0]) {
            S.error[id] = kErrNoconv;
            S.done[id] = 1;
            S.iter[id] = S.iprm[0] + 1;
          } else {
            S.rho1[id] = S.rho[id];
            S.iter[id] += 1;
          }
        }
      }
    }
    if (TAIL == kbRel || TAIL == kbInit) fl = !S.done[id];
    flag[k] = fl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t n = 0;
    for (int j = 0; j < KS; ++j) n += flag[j];
    // The FP64 residual costs the same for one right-hand side as for all of
    // a system's: when one asks for it, every live one is refreshed, so their
    // reference norms stay aligned and the next replacement comes later.
    if (TAIL == kbRes && n > 0)
      for (int j = 0; j < KS; ++j)
        if (!S.done[s * KS + j]) S.rel[s * KS + j] = 1;
    if (TAIL == kbRes) S.sys_rel[s] = n > 0;
    // a finished system's tails return early from now on: clear its request
    // here, or the FP64 residual SpMM would keep computing its tiles
    if (TAIL == kbRel) S.sys_rel[s] = 0;
    if (TAIL == kbRel || TAIL == kbInit) S.sys_active[s] = n;
    if (TAIL == kbInit) S.sys_rel[s] = n > 0;
  }
}

// b.b partials per tile (initial phase, one workgroup per tile)
__global__ __launch_bounds__(kBT) void kb_bb(Tiles T, const double* __restrict__ b, double* part) {
  const int t = blockIdx.x;
  const int k = threadIdx.x % KS, rl = threadIdx.x / KS;
  double acc[1] = {0.0};
  const int r0 = T.row0[t], nr = T.nrows[t];
  for (int i = rl; i < nr; i += kBT / KS) {
    const int64_t o = ((int64_t)(r0 + i) * KS + k) * 3;
    acc[0] += b[o] * b[o] + b[o + 1] * b[o + 1] + b[o + 2] * b[o + 2];
  }
  tile_partials_rowlanes<1>(acc, part, T.ntiles, t);
}

__global__ void kb_fill0(int64_t n, float* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0f;
}
__global__ void kb_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ---- SpMM variant table ------------------------------------------------------------------
using SpmmFn = void (*)(int grid, hipStream_t, Tiles, QView, const int32_t*, const int32_t*, const void*,
                        const void*, const void*, void*, BState, double*, int);
using OccFn = int (*)();

template <typename T, int LPR, int RPL, int U, int EPI>
void spmm_launch(int grid, hipStream_t st, Tiles tl, QView qv, const int32_t* rp, const int32_t* col, const void* val,
                 const void* x, const void* w, void* y, BState S, double* part, int guard) {
  kb_spmm<T, LPR, RPL, U, EPI><<<grid, kBT, 0, st>>>(tl, qv, rp, col, static_cast<const T*>(val),
                                                     static_cast<const T*>(x), w, y, S, part, guard);
}

// resident workgroups of a kernel per CU (the persistent grid is that times the CU count)
template <class K>
int occupancy_of(K kernel) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBT, 0) != hipSuccess) nb = 1;
  return nb < 1 ? 1 : nb;
}
template <typename T, int LPR, int RPL, int U, int EPI>
int spmm_occupancy() {
  return occupancy_of(kb_spmm<T, LPR, RPL, U, EPI>);
}

struct Variant {
  const char* name;
  int lpr, rpl, u;
  SpmmFn q32, st32;
  OccFn occ;
};
#define HECMW33_VARIANT(L, R, U)                                                                        \
  Variant {                                                                                             \
    "lpr" #L "_rpl" #R "_u" #U, L, R, U, &spmm_launch<float, L, R, U, kbEpiQ>,                          \
        &spmm_launch<float, L, R, U, kbEpiStore>, &spmm_occupancy<float, L, R, U, kbEpiQ>               \
  }
const Variant kVariants[] = {
    HECMW33_VARIANT(16, 1, 1), HECMW33_VARIANT(16, 1, 2), HECMW33_VARIANT(8, 2, 1), HECMW33_VARIANT(8, 2, 2),
    HECMW33_VARIANT(4, 4, 1),  HECMW33_VARIANT(4, 4, 2),  HECMW33_VARIANT(32, 1, 1), HECMW33_VARIANT(32, 1, 2),
    HECMW33_VARIANT(64, 1, 1), HECMW33_VARIANT(16, 2, 1), HECMW33_VARIANT(32, 2, 1), HECMW33_VARIANT(64, 4, 1),
    HECMW33_VARIANT(16, 4, 1), HECMW33_VARIANT(32, 4, 1), HECMW33_VARIANT(8, 4, 1),  HECMW33_VARIANT(16, 8, 1),
    HECMW33_VARIANT(32, 8, 1), HECMW33_VARIANT(8, 8, 1),  HECMW33_VARIANT(4, 8, 1),
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
#undef HECMW33_VARIANT

// FP64 SpMM shapes (residual replacement, products on request): one phase per
// row (LPR = 16 / RPL), so every (row, rhs) sums its blocks in storage order
// and the result is bitwise the reference's SpMV whichever shape runs.
struct Variant64 {
  const char* name;
  SpmmFn res, st;
  OccFn occ;
};
#define HECMW33_VARIANT64(L, R, U)                                                                   \
  Variant64 {                                                                                        \
    "lpr" #L "_rpl" #R "_u" #U, &spmm_launch<double, L, R, U, kbEpiRes>,                             \
        &spmm_launch<double, L, R, U, kbEpiStore>, &spmm_occupancy<double, L, R, U, kbEpiRes>        \
  }
const Variant64 kVariants64[] = {
    HECMW33_VARIANT64(16, 1, 2), HECMW33_VARIANT64(16, 1, 1), HECMW33_VARIANT64(8, 2, 1), HECMW33_VARIANT64(8, 2, 2),
    HECMW33_VARIANT64(4, 4, 1),  HECMW33_VARIANT64(4, 4, 2),  HECMW33_VARIANT64(2, 8, 1),
};
constexpr int kNumVariants64 = sizeof(kVariants64) / sizeof(kVariants64[0]);
#undef HECMW33_VARIANT64

}  // namespace

// ---------------------------------------------------------------------------------------------
// host side
static thread_local std::string g_batch_error;
#define BHIPC(expr)                                                               \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      g_batch_error = std::string(#expr) + ": " + hipGetErrorString(_e);          \
      return HECMW33_ERROR_DEVICE;                                                \
    }                                                                             \
  } while (0)

struct BMem {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) n = 8;
    if (hipMalloc(&p, n) != hipSuccess) {
      g_batch_error = "hipMalloc failed";
      return HECMW33_ERROR_DEVICE;
    }
    bytes = n;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace hecmw33

using namespace hecmw33;

struct hecmw33_batch {
  hipStream_t stream = nullptr;
  int32_t nsys = 0, nrhs = 0;
  int64_t nrows = 0, nnzb = 0;
  std::vector<int32_t> N, NP, row0;     // per system
  std::vector<int32_t> tile_sys_h, tile_row0_h;
  int32_t ntiles = 0;
  BMem rowptr, col, val64, val32, alu32, d64;
  BMem tile_row0, tile_nrows, tile_sys, sys_tile0, sys_ntile;
  BMem b, x, dx, r, z, p, q, part, state;
  BState S{};
  Tiles T{};
  int variant = -1;
  float variant_ms[64] = {};
  int variant64 = 0;
  float variant64_ms[16] = {};
  hipGraphExec_t graph = nullptr;
  int32_t* h_active = nullptr;  // pinned, 2 slots
  double spmm_ms = 0.0;
  // live-tile work queues and the persistent grids that walk them
  BMem qtile, qn;
  int ncu = 0;
  int grid_p = 0, grid_q = 0, grid_upd = 0, grid_res = 0, grid_diag = 0;
  int tune_cached = 0;  // 1: the SpMM shapes came from the autotuner's persisted winners

  ~hecmw33_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph) (void)hipGraphExecDestroy(graph);
    for (BMem* m : {&rowptr, &col, &val64, &val32, &alu32, &d64, &tile_row0, &tile_nrows, &tile_sys, &sys_tile0,
                    &sys_ntile, &b, &x, &dx, &r, &z, &p, &q, &part, &state, &qtile, &qn})
      m->release();
    if (h_active) (void)hipHostFree(h_active);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace hecmw33 {

static int32_t batch_state_layout(hecmw33_batch& B) {
  const int64_t ni = (int64_t)B.nsys * KS;
  const size_t bytes = sizeof(double) * (11 * ni + 2) + sizeof(int32_t) * (5 * ni + 2 * B.nsys + 3) + 256;
  if (B.state.ensure(bytes)) return HECMW33_ERROR_DEVICE;
  char* base = B.state.as<char>();
  double* d = reinterpret_cast<double*>(base);
  BState& S = B.S;
  double** dp[11] = {&S.bnrm2, &S.rho, &S.rho1, &S.zq, &S.pq, &S.rp, &S.alpha, &S.beta, &S.rr_ref, &S.resid, &S.gap};
  for (int i = 0; i < 11; ++i) *dp[i] = d + i * ni;
  S.prm = d + 11 * ni;
  int32_t* ip = reinterpret_cast<int32_t*>(d + 11 * ni + 2);
  int32_t** ipp[5] = {&S.iter, &S.done, &S.error, &S.nindef, &S.rel};
  for (int i = 0; i < 5; ++i) *ipp[i] = ip + i * ni;
  S.sys_active = ip + 5 * ni;
  S.sys_rel = S.sys_active + B.nsys;
  S.active_total = S.sys_rel + B.nsys;
  S.iprm = S.active_total + 1;
  return 0;
}

static int grid_of(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + kBT - 1) / kBT, 1), 8192); }

static Queue queue_of(const hecmw33_batch& B) {
  return Queue{B.qtile.as<int32_t>(), B.qn.as<int32_t>()};
}

// a launch that walks list `list`
static QView qview(const hecmw33_batch& B, int list) {
  return QView{B.qtile.as<int32_t>() + (int64_t)list * B.ntiles, B.qn.as<int32_t>() + list};
}

// one workgroup per tile of the whole batch (set-up launches)
static QView qall() { return QView{nullptr, nullptr}; }

// one FCG iteration for every live instance
static void batch_enqueue_iteration(hecmw33_batch& B) {
  hipStream_t st = B.stream;
  double* part = B.part.as<double>();
  float *r = B.r.as<float>(), *z = B.z.as<float>(), *p = B.p.as<float>(), *q = B.q.as<float>();
  // z = D^-1 r and its dots (r.z, z.q) come from the previous iteration's
  // kb_update_diag (or kb_diag after a residual replacement, or before the
  // first iteration), in the second half of the partials
  double* part_rz = part + (size_t)2 * KS * B.ntiles;
  kb_tail<kbBeta, 2><<<B.nsys, kBT, 0, st>>>(B.T, B.S, part_rz);
  kb_p<<<B.grid_p, kBT, 0, st>>>(B.T, qview(B, kqActive), z, p, B.S);
  kVariants[B.variant].q32(B.grid_q, st, B.T, qview(B, kqActive), B.rowptr.as<int32_t>(), B.col.as<int32_t>(),
                           B.val32.p, p, r, q, B.S, part, 1);
  kb_tail<kbAlpha, 2><<<B.nsys, kBT, 0, st>>>(B.T, B.S, part);
  kb_update_diag<<<B.grid_upd, kBT, 0, st>>>(B.T, qview(B, kqActive), B.alu32.as<float>(), p, q,
                                             B.dx.as<float>(), r, z, B.S, part, part_rz);
  kb_tail<kbRes, 1><<<B.nsys, kBT, 0, st>>>(B.T, B.S, part);
  // the systems whose residuals are replaced now, their x brought up to date
  kb_queue<kqRel><<<1, 1024, 0, st>>>(B.T, B.nsys, B.S, queue_of(B));
  kb_fold<<<B.grid_res, kBT, 0, st>>>(B.T, qview(B, kqRel), B.dx.as<float>(), B.x.as<double>(), B.S, 2);
  kVariants64[B.variant64].res(B.grid_res, st, B.T, qview(B, kqRel), B.rowptr.as<int32_t>(),
                               B.col.as<int32_t>(), B.val64.p, B.x.p, B.b.p, r, B.S, part, 2);
  // systems whose residuals were just replaced: z from the new r
  kb_diag<<<B.grid_diag, kBT, 0, st>>>(B.T, qview(B, kqRel), B.alu32.as<float>(), r, z, q, B.S, part_rz, 2);
  kb_tail<kbRel, 1><<<B.nsys, kBT, 0, st>>>(B.T, B.S, part);
  // the systems still running after this iteration, and their live count for the host
  kb_queue<kqActive><<<1, 1024, 0, st>>>(B.T, B.nsys, B.S, queue_of(B));
}

// ---- the autotuner's winners, persisted per (blocks per row, rows) bucket ------------------
// Key: the device (architecture and CU count), the mean blocks per row rounded, and floor(log2(rows)).
// Value: the names of the FP32 and FP64 SpMM shapes that won on a batch of that bucket.  Kept in a
// text file, one "<arch> <cus> <blocks/row> <log2 rows> <fp32 shape> <fp64 shape>" line per bucket,
// and in this process for that file: $HECMW33_TUNE_CACHE names the file ("off": nothing is kept and
// every load times the shapes), else $XDG_CACHE_HOME/hecmw33/batch_tune.txt, else
// $HOME/.cache/hecmw33/batch_tune.txt.  A line whose shape names this library does not have is
// ignored (and re-timed).
struct TuneKey {
  std::string arch;
  int cus, bpr, lrows;
  std::string str() const {
    return arch + " " + std::to_string(cus) + " " + std::to_string(bpr) + " " + std::to_string(lrows);
  }
};

static std::mutex g_tune_mu;
static std::map<std::string, std::pair<int, int>> g_tune_mem;

static std::string tune_path() {
  const char* e = getenv("HECMW33_TUNE_CACHE");
  if (e && *e) return std::string(e) == "off" ? std::string() : std::string(e);
  const char* x = getenv("XDG_CACHE_HOME");
  if (x && *x) return std::string(x) + "/hecmw33/batch_tune.txt";
  const char* h = getenv("HOME");
  if (h && *h) return std::string(h) + "/.cache/hecmw33/batch_tune.txt";
  return std::string();
}

static int variant_by_name(const std::string& n) {
  for (int v = 0; v < kNumVariants; ++v)
    if (n == kVariants[v].name) return v;
  return -1;
}
static int variant64_by_name(const std::string& n) {
  for (int v = 0; v < kNumVariants64; ++v)
    if (n == kVariants64[v].name) return v;
  return -1;
}

// every valid line of the file, key -> (fp32, fp64)
static std::map<std::string, std::pair<int, int>> tune_read(const std::string& path) {
  std::map<std::string, std::pair<int, int>> m;
  if (path.empty()) return m;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string arch, n32, n64;
    int cus, bpr, lrows;
    if (!(is >> arch >> cus >> bpr >> lrows >> n32 >> n64)) continue;
    const int v32 = variant_by_name(n32), v64 = variant64_by_name(n64);
    if (v32 < 0 || v64 < 0) continue;
    m[TuneKey{arch, cus, bpr, lrows}.str()] = {v32, v64};
  }
  return m;
}

static void tune_write(const std::string& path, const std::map<std::string, std::pair<int, int>>& m) {
  if (path.empty()) return;
  // create the missing directories, then replace the file through a rename
  for (size_t i = 1; i < path.size(); ++i)
    if (path[i] == '/') (void)::mkdir(path.substr(0, i).c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string((long)::getpid());
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return;
    for (const auto& kv : m)
      f << kv.first << " " << kVariants[kv.second.first].name << " " << kVariants64[kv.second.second].name << "\n";
    if (!f) return;
  }
  (void)::rename(tmp.c_str(), path.c_str());
}

static TuneKey tune_key(const hecmw33_batch& B) {
  int dev = 0;
  hipDeviceProp_t prop{};
  std::string arch = "unknown";
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    arch = prop.gcnArchName;
    const size_t c = arch.find(':');
    if (c != std::string::npos) arch = arch.substr(0, c);
  }
  const int bpr = (int)std::lround((double)B.nnzb / (double)std::max<int64_t>(1, B.nrows));
  int lrows = 0;
  while ((int64_t(2) << lrows) <= B.nrows) ++lrows;
  return TuneKey{arch, B.ncu, bpr, lrows};
}

// persistent grids of the iteration's per-tile launches for the chosen shapes
static void batch_grids(hecmw33_batch& B) {
  auto grid = [&](int per_cu) {
    const int64_t g = (int64_t)per_cu * std::max(1, B.ncu);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, B.ntiles));
  };
  B.grid_p = grid(occupancy_of(kb_p));
  B.grid_upd = grid(occupancy_of(kb_update_diag));
  B.grid_diag = grid(occupancy_of(kb_diag));
  B.grid_q = grid(kVariants[B.variant].occ());
  B.grid_res = grid(kVariants64[B.variant64].occ());
}

static int32_t batch_autotune(hecmw33_batch& B) {
  const TuneKey key = tune_key(B);
  const std::string ks = key.str(), path = tune_path(), mk = path + "|" + ks;
  bool cached = false;
  if (!path.empty()) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_mem.find(mk);
    if (it == g_tune_mem.end()) {
      const auto disk = tune_read(path);
      auto jt = disk.find(ks);
      if (jt != disk.end()) it = g_tune_mem.emplace(mk, jt->second).first;
    }
    if (it != g_tune_mem.end()) {
      B.variant = it->second.first;
      B.variant64 = it->second.second;
      cached = true;
    }
  }
  B.tune_cached = cached ? 1 : 0;
  std::fill(std::begin(B.variant_ms), std::end(B.variant_ms), 0.f);
  std::fill(std::begin(B.variant64_ms), std::end(B.variant64_ms), 0.f);
  // time the FP32 shapes' SpMM with its fused dots on the loaded matrix, as
  // the iteration runs it: persistent, over the live-tile list of every system
  // (the FP64 shapes over the list of residual replacements, also every
  // system).  With a persisted winner only the winners are timed, for the
  // figures hecmw33_batch_info reports.
  hipEvent_t e0, e1;
  BHIPC(hipEventCreate(&e0));
  BHIPC(hipEventCreate(&e1));
  std::vector<int32_t> ones(B.nsys, 1);
  BHIPC(hipMemcpyAsync(B.S.sys_active, ones.data(), sizeof(int32_t) * B.nsys, hipMemcpyHostToDevice, B.stream));
  BHIPC(hipMemcpyAsync(B.S.sys_rel, ones.data(), sizeof(int32_t) * B.nsys, hipMemcpyHostToDevice, B.stream));
  // both live-tile lists: every tile
  kb_queue<kqActive><<<1, 1024, 0, B.stream>>>(B.T, B.nsys, B.S, queue_of(B));
  kb_queue<kqRel><<<1, 1024, 0, B.stream>>>(B.T, B.nsys, B.S, queue_of(B));
  auto timed = [&](auto one, int warm, int reps, float* ms) -> int32_t {
    for (int rep = 0; rep < warm; ++rep) one();
    BHIPC(hipEventRecord(e0, B.stream));
    for (int rep = 0; rep < reps; ++rep) one();
    BHIPC(hipEventRecord(e1, B.stream));
    BHIPC(hipEventSynchronize(e1));
    BHIPC(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    return 0;
  };
  auto time32 = [&](int v) -> int32_t {
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)kVariants[v].occ() * B.ncu, B.ntiles));
    return timed(
        [&]() {
          kVariants[v].q32(g, B.stream, B.T, qview(B, kqActive), B.rowptr.as<int32_t>(), B.col.as<int32_t>(),
                           B.val32.p, B.p.p, B.r.p, B.q.p, B.S, B.part.as<double>(), 1);
        },
        2, 5, &B.variant_ms[v]);
  };
  auto time64 = [&](int v) -> int32_t {
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)kVariants64[v].occ() * B.ncu, B.ntiles));
    return timed(
        [&]() {
          kVariants64[v].res(g, B.stream, B.T, qview(B, kqRel), B.rowptr.as<int32_t>(), B.col.as<int32_t>(),
                             B.val64.p, B.x.p, B.b.p, B.r.p, B.S, B.part.as<double>(), 2);
        },
        1, 3, &B.variant64_ms[v]);
  };
  if (cached) {
    if (int32_t e = time32(B.variant)) return e;
  } else {
    B.variant = 0;
    for (int v = 0; v < kNumVariants; ++v) {
      if (int32_t e = time32(v)) return e;
      if (B.variant_ms[v] < B.variant_ms[B.variant]) B.variant = v;
    }
  }
  B.spmm_ms = B.variant_ms[B.variant];
  // the FP64 shapes on the residual replacement of every instance (all
  // flagged), from x = 0; the solve starts from 0 and rewrites r
  const int64_t nv = B.nrows * KS * 3;
  BHIPC(hipMemsetAsync(B.x.p, 0, sizeof(double) * nv, B.stream));
  std::vector<int32_t> rel((size_t)B.nsys * KS, 1);
  BHIPC(hipMemcpyAsync(B.S.rel, rel.data(), sizeof(int32_t) * rel.size(), hipMemcpyHostToDevice, B.stream));
  if (cached) {
    if (int32_t e = time64(B.variant64)) return e;
  } else {
    B.variant64 = 0;
    for (int v = 0; v < kNumVariants64; ++v) {
      if (int32_t e = time64(v)) return e;
      if (B.variant64_ms[v] < B.variant64_ms[B.variant64]) B.variant64 = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  BHIPC(hipGetLastError());
  batch_grids(B);
  if (!cached && !path.empty()) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_mem[mk] = {B.variant, B.variant64};
    auto disk = tune_read(path);
    disk[ks] = {B.variant, B.variant64};
    tune_write(path, disk);
  }
  return 0;
}

static int32_t batch_load(hecmw33_batch& B, int32_t nsys, const hecmw33_sys* sys, int32_t nrhs, const double* const* Bh,
                          const double* rarray) {
  if (nsys < 1 || !sys || nrhs < 1 || nrhs > KS || !Bh) {
    g_batch_error = "batch: need nsys >= 1, 1 <= nrhs <= 16 and right-hand sides";
    return HECMW33_ERROR_ARGS;
  }
  if (!B.stream) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      g_batch_error = "no HIP device visible";
      return HECMW33_ERROR_DEVICE;
    }
    BHIPC(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    BHIPC(hipHostMalloc((void**)&B.h_active, 2 * sizeof(int32_t), hipHostMallocDefault));
  }
  if (B.graph) {
    (void)hipGraphExecDestroy(B.graph);
    B.graph = nullptr;
  }
  B.nsys = nsys;
  B.nrhs = nrhs;
  B.N.assign(nsys, 0);
  B.NP.assign(nsys, 0);
  B.row0.assign(nsys + 1, 0);
  int64_t nnz = 0;
  for (int s = 0; s < nsys; ++s) {
    const hecmw33_sys& A = sys[s];
    if (A.N < 1 || A.NP != A.N || !A.D || !A.indexL || !A.indexU || (A.NPL > 0 && (!A.AL || !A.itemL)) ||
        (A.NPU > 0 && (!A.AU || !A.itemU)) || A.indexL[A.N] != A.NPL || A.indexU[A.N] != A.NPU || !Bh[s]) {
      g_batch_error = "batch: invalid system " + std::to_string(s) + " (one domain, NP == N)";
      return HECMW33_ERROR_ARGS;
    }
    for (int64_t q = 0; q < A.NPL; ++q)
      if (A.itemL[q] < 1 || A.itemL[q] > A.N) {
        g_batch_error = "batch: itemL out of range";
        return HECMW33_ERROR_ARGS;
      }
    for (int64_t q = 0; q < A.NPU; ++q)
      if (A.itemU[q] < 1 || A.itemU[q] > A.N) {
        g_batch_error = "batch: itemU out of range";
        return HECMW33_ERROR_ARGS;
      }
    B.N[s] = A.N;
    B.NP[s] = A.NP;
    B.row0[s + 1] = B.row0[s] + A.N;
    nnz += (int64_t)A.N + A.NPL + A.NPU;
  }
  B.nrows = B.row0[nsys];
  B.nnzb = nnz;
  if (nnz > INT32_MAX) {
    g_batch_error = "batch: more than 2^31 blocks";
    return HECMW33_ERROR_ARGS;
  }
  // merged rows D, L, U (storage order), global columns; values FP64
  std::vector<int32_t> rp(B.nrows + 1), cl(nnz);
  std::vector<double> v64(9 * nnz), d64(9 * (size_t)B.nrows);
  for (int s = 0; s < nsys; ++s) {
    const hecmw33_sys& A = sys[s];
    const int32_t g0 = B.row0[s];
    for (int32_t i = 0; i < A.N; ++i) rp[g0 + i + 1] = 1 + (A.indexL[i + 1] - A.indexL[i]) + (A.indexU[i + 1] - A.indexU[i]);
  }
  rp[0] = 0;
  for (int64_t g = 0; g < B.nrows; ++g) rp[g + 1] += rp[g];
  for (int s = 0; s < nsys; ++s) {
    const hecmw33_sys& A = sys[s];
    const int32_t g0 = B.row0[s];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < A.N; ++i) {
      int64_t o = rp[g0 + i];
      cl[o] = g0 + i;
      std::memcpy(&v64[9 * o], A.D + 9 * (size_t)i, 72);
      ++o;
      for (int32_t q = A.indexL[i]; q < A.indexL[i + 1]; ++q, ++o) {
        cl[o] = g0 + A.itemL[q] - 1;
        std::memcpy(&v64[9 * o], A.AL + 9 * (size_t)q, 72);
      }
      for (int32_t q = A.indexU[i]; q < A.indexU[i + 1]; ++q, ++o) {
        cl[o] = g0 + A.itemU[q] - 1;
        std::memcpy(&v64[9 * o], A.AU + 9 * (size_t)q, 72);
      }
      std::memcpy(&d64[9 * (size_t)(g0 + i)], A.D + 9 * (size_t)i, 72);
    }
  }
  // tiles: runs of rows inside a system, at least kTileRows, count capped
  int tr = kTileRows;
  for (;;) {
    int64_t nt = 0;
    for (int s = 0; s < nsys; ++s) nt += (B.N[s] + tr - 1) / tr;
    if (nt <= kMaxTiles) break;
    tr *= 2;
  }
  std::vector<int32_t> t_row0, t_nrows, t_sys, s_t0(nsys), s_nt(nsys);
  for (int s = 0; s < nsys; ++s) {
    s_t0[s] = (int32_t)t_row0.size();
    for (int32_t i = 0; i < B.N[s]; i += tr) {
      t_row0.push_back(B.row0[s] + i);
      t_nrows.push_back(std::min(tr, B.N[s] - i));
      t_sys.push_back(s);
    }
    s_nt[s] = (int32_t)t_row0.size() - s_t0[s];
  }
  B.ntiles = (int32_t)t_row0.size();
  // device buffers
  const int64_t nv = B.nrows * KS * 3;
  if (B.rowptr.ensure(sizeof(int32_t) * rp.size()) || B.col.ensure(sizeof(int32_t) * nnz) ||
      B.val64.ensure(sizeof(double) * 9 * nnz) || B.val32.ensure(sizeof(float) * 9 * nnz) ||
      B.d64.ensure(sizeof(double) * 9 * B.nrows) || B.alu32.ensure(sizeof(float) * 9 * B.nrows) ||
      B.tile_row0.ensure(sizeof(int32_t) * B.ntiles) || B.tile_nrows.ensure(sizeof(int32_t) * B.ntiles) ||
      B.tile_sys.ensure(sizeof(int32_t) * B.ntiles) || B.sys_tile0.ensure(sizeof(int32_t) * nsys) ||
      B.sys_ntile.ensure(sizeof(int32_t) * nsys) || B.b.ensure(sizeof(double) * nv) ||
      B.x.ensure(sizeof(double) * nv) || B.dx.ensure(sizeof(float) * nv) || B.r.ensure(sizeof(float) * nv) ||
      B.z.ensure(sizeof(float) * nv) ||
      B.p.ensure(sizeof(float) * nv) || B.q.ensure(sizeof(float) * nv) ||
      B.part.ensure(sizeof(double) * 4 * KS * (size_t)B.ntiles) || batch_state_layout(B) ||
      B.qtile.ensure(sizeof(int32_t) * kqLists * (size_t)B.ntiles) || B.qn.ensure(sizeof(int32_t) * kqLists))
    return HECMW33_ERROR_DEVICE;
  {
    int dev = 0;
    BHIPC(hipGetDevice(&dev));
    BHIPC(hipDeviceGetAttribute(&B.ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  hipStream_t st = B.stream;
  BHIPC(hipMemcpyAsync(B.rowptr.p, rp.data(), sizeof(int32_t) * rp.size(), hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.col.p, cl.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.val64.p, v64.data(), sizeof(double) * 9 * nnz, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.d64.p, d64.data(), sizeof(double) * 9 * B.nrows, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.tile_row0.p, t_row0.data(), sizeof(int32_t) * B.ntiles, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.tile_nrows.p, t_nrows.data(), sizeof(int32_t) * B.ntiles, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.tile_sys.p, t_sys.data(), sizeof(int32_t) * B.ntiles, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.sys_tile0.p, s_t0.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.sys_ntile.p, s_nt.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, st));
  // right-hand sides into [row][16][3] (unused columns zero)
  {
    std::vector<double> bh(nv, 0.0);
    for (int s = 0; s < nsys; ++s) {
      const int64_t n3 = 3 * (int64_t)B.NP[s];
      for (int k = 0; k < nrhs; ++k)
        for (int32_t i = 0; i < B.N[s]; ++i)
          for (int d = 0; d < 3; ++d) bh[((int64_t)(B.row0[s] + i) * KS + k) * 3 + d] = Bh[s][n3 * k + 3 * (int64_t)i + d];
    }
    BHIPC(hipMemcpyAsync(B.b.p, bh.data(), sizeof(double) * nv, hipMemcpyHostToDevice, st));
    BHIPC(hipStreamSynchronize(st));
  }
  kb_to_f32<<<grid_of(9 * nnz), kBT, 0, st>>>(9 * nnz, B.val64.as<double>(), B.val32.as<float>());
  // DIAG factors (hecmw_precond_DIAG_33.F90:27-104) in FP64, rounded to FP32
  {
    BMem rows, alu64;
    std::vector<int32_t> id(B.nrows);
    for (int64_t g = 0; g < B.nrows; ++g) id[g] = (int32_t)g;
    if (rows.ensure(sizeof(int32_t) * B.nrows) || alu64.ensure(sizeof(double) * 9 * B.nrows))
      return HECMW33_ERROR_DEVICE;
    BHIPC(hipMemcpyAsync(rows.p, id.data(), sizeof(int32_t) * B.nrows, hipMemcpyHostToDevice, st));
    const double sigma = rarray ? rarray[1] : 1.0;
    launch_block_lu((int32_t)B.nrows, rows.as<int32_t>(), B.d64.as<double>(), sigma > 0.0 ? sigma : 1.0,
                    alu64.as<double>(), 1, st);
    kb_to_f32<<<grid_of(9 * B.nrows), kBT, 0, st>>>(9 * B.nrows, alu64.as<double>(), B.alu32.as<float>());
    BHIPC(hipStreamSynchronize(st));
    rows.release();
    alu64.release();
  }
  B.T.row0 = B.tile_row0.as<int32_t>();
  B.T.nrows = B.tile_nrows.as<int32_t>();
  B.T.sys = B.tile_sys.as<int32_t>();
  B.T.sys_tile0 = B.sys_tile0.as<int32_t>();
  B.T.sys_ntile = B.sys_ntile.as<int32_t>();
  B.T.ntiles = B.ntiles;
  for (BMem* m : {&B.r, &B.z, &B.p, &B.q}) kb_fill0<<<grid_of(nv), kBT, 0, st>>>(nv, m->as<float>());
  BHIPC(hipGetLastError());
  BHIPC(hipStreamSynchronize(st));
  return batch_autotune(B);
}

static int32_t batch_solve(hecmw33_batch& B, const int32_t* iarray, const double* rarray, int32_t* iter,
                           int32_t* error, double* resid, double* times) {
  if (!B.stream || B.nsys < 1) {
    g_batch_error = "batch: nothing loaded";
    return HECMW33_ERROR_ARGS;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = B.stream;
  const int32_t maxit = iarray ? iarray[0] : 100;
  const int32_t pc = iarray ? iarray[2] : 3;
  const int32_t iterpremax = iarray ? iarray[4] : 1;
  if (pc != 3 || iterpremax != 1) {
    g_batch_error = "batch: DIAG (PRECOND=3, ITERPREMAX=1) only";
    return HECMW33_ERROR_UNSUPPORTED;
  }
  const double delta = rarray && rarray[40] > 0.0 && rarray[40] < 1.0 ? rarray[40] : 1.0e-2;
  const double prm[2] = {rarray ? rarray[0] : 1e-8, delta * delta};
  const int32_t iprm[2] = {maxit, 50};
  BHIPC(hipMemcpyAsync(B.S.prm, prm, sizeof(prm), hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.S.iprm, iprm, sizeof(iprm), hipMemcpyHostToDevice, st));
  const int64_t nv = B.nrows * KS * 3;
  double* part = B.part.as<double>();
  // x0 = 0; r = b - A x0 in FP64 -> r32; b.b; instance state
  BHIPC(hipMemsetAsync(B.x.p, 0, sizeof(double) * nv, st));
  for (BMem* m : {&B.dx, &B.z, &B.p, &B.q}) kb_fill0<<<grid_of(nv), kBT, 0, st>>>(nv, m->as<float>());
  kb_bb<<<B.ntiles, kBT, 0, st>>>(B.T, B.b.as<double>(), part);
  std::vector<int32_t> ones((size_t)B.nsys * KS, 1), sys1(B.nsys, 1);
  BHIPC(hipMemcpyAsync(B.S.rel, ones.data(), sizeof(int32_t) * ones.size(), hipMemcpyHostToDevice, st));
  BHIPC(hipMemcpyAsync(B.S.sys_rel, sys1.data(), sizeof(int32_t) * B.nsys, hipMemcpyHostToDevice, st));
  // |b - A x0|^2 into the second dot slot
  kVariants64[B.variant64].res(B.ntiles, st, B.T, qall(), B.rowptr.as<int32_t>(), B.col.as<int32_t>(), B.val64.p,
                               B.x.p, B.b.p, B.r.p, B.S, part + (size_t)KS * B.ntiles, 0);
  kb_tail<kbInit, 2><<<B.nsys, kBT, 0, st>>>(B.T, B.S, part);
  // the first iteration's z = D^-1 r0 and its dots
  kb_diag<<<B.ntiles, kBT, 0, st>>>(B.T, qall(), B.alu32.as<float>(), B.r.as<float>(), B.z.as<float>(),
                                    B.q.as<float>(), B.S, part + (size_t)2 * KS * B.ntiles, 0);
  // the systems the first iteration runs
  kb_queue<kqActive><<<1, 1024, 0, st>>>(B.T, B.nsys, B.S, queue_of(B));
  BHIPC(hipGetLastError());
  BHIPC(hipStreamSynchronize(st));
  const auto t1 = std::chrono::steady_clock::now();
  // capture kChunk iterations once (pointers and sizes are fixed after load)
  if (!B.graph) {
    hipGraph_t g;
    BHIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < kChunk; ++i) batch_enqueue_iteration(B);
    BHIPC(hipStreamEndCapture(st, &g));
    BHIPC(hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
  }
  hipEvent_t ev[2];
  BHIPC(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  BHIPC(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  const int64_t nchunks = (maxit + kChunk - 1) / kChunk;
  int64_t launched = 0;
  bool finished = maxit < 1;
  for (int64_t c = 0; c < nchunks && !finished; ++c) {
    BHIPC(hipGraphLaunch(B.graph, st));
    const int slot = (int)(c & 1);
    BHIPC(hipMemcpyAsync(&B.h_active[slot], B.S.active_total, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    BHIPC(hipEventRecord(ev[slot], st));
    ++launched;
    if (c > 0) {
      BHIPC(hipEventSynchronize(ev[slot ^ 1]));
      if (B.h_active[slot ^ 1] == 0) finished = true;
    }
  }
  // the increments of instances that stopped without a final replacement
  kb_fold<<<B.ntiles, kBT, 0, st>>>(B.T, qall(), B.dx.as<float>(), B.x.as<double>(), B.S, 0);
  BHIPC(hipGetLastError());
  BHIPC(hipStreamSynchronize(st));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  const auto t2 = std::chrono::steady_clock::now();
  const int64_t ni = (int64_t)B.nsys * KS;
  std::vector<int32_t> it(ni), er(ni);
  std::vector<double> rs(ni);
  BHIPC(hipMemcpy(it.data(), B.S.iter, sizeof(int32_t) * ni, hipMemcpyDeviceToHost));
  BHIPC(hipMemcpy(er.data(), B.S.error, sizeof(int32_t) * ni, hipMemcpyDeviceToHost));
  BHIPC(hipMemcpy(rs.data(), B.S.resid, sizeof(double) * ni, hipMemcpyDeviceToHost));
  for (int s = 0; s < B.nsys; ++s)
    for (int k = 0; k < B.nrhs; ++k) {
      const int64_t o = (int64_t)s * B.nrhs + k;
      if (iter) iter[o] = it[s * KS + k];
      if (error) error[o] = er[s * KS + k];
      if (resid) resid[o] = rs[s * KS + k];
    }
  if (times) {
    times[0] = std::chrono::duration<double>(t1 - t0).count();
    times[1] = std::chrono::duration<double>(t2 - t1).count();
    times[2] = B.spmm_ms;
    times[3] = (double)launched * kChunk;
  }
  return 0;
}

}  // namespace hecmw33

extern "C" {

hecmw33_batch* hecmw33_batch_create(void) { return new (std::nothrow) hecmw33_batch(); }

void hecmw33_batch_destroy(hecmw33_batch* b) { delete b; }

const char* hecmw33_batch_last_error(void) { return g_batch_error.c_str(); }

int32_t hecmw33_batch_load(hecmw33_batch* b, int32_t nsys, const hecmw33_sys* sys, int32_t nrhs,
                           const double* const* B, const int32_t iarray[100], const double rarray[100]) {
  (void)iarray;
  if (!b) return HECMW33_ERROR_ARGS;
  return batch_load(*b, nsys, sys, nrhs, B, rarray);
}

int32_t hecmw33_batch_solve(hecmw33_batch* b, const int32_t iarray[100], const double rarray[100], int32_t* iter,
                            int32_t* error, double* resid, double* times) {
  if (!b) return HECMW33_ERROR_ARGS;
  return batch_solve(*b, iarray, rarray, iter, error, resid, times);
}

int32_t hecmw33_batch_get_x(hecmw33_batch* b, int32_t s, double* X) {
  if (!b || s < 0 || s >= b->nsys || !X) return HECMW33_ERROR_ARGS;
  const int64_t n = (int64_t)b->N[s] * KS * 3;
  std::vector<double> h(n);
  BHIPC(hipMemcpy(h.data(), b->x.as<double>() + (int64_t)b->row0[s] * KS * 3, sizeof(double) * n,
                  hipMemcpyDeviceToHost));
  const int64_t n3 = 3 * (int64_t)b->NP[s];
  for (int k = 0; k < b->nrhs; ++k)
    for (int32_t i = 0; i < b->N[s]; ++i)
      for (int d = 0; d < 3; ++d) X[n3 * k + 3 * (int64_t)i + d] = h[((int64_t)i * KS + k) * 3 + d];
  return 0;
}

// Y_s = A_s X_s for every system and right-hand side (host arrays, nrhs
// columns of 3*N_s each).  fp64 != 0: FP64 in the reference's order;
// otherwise the FP32 SpMM of the tuned variant (values rounded to FP32).
int32_t hecmw33_batch_spmm(hecmw33_batch* b, int32_t fp64, const double* const* X, double* const* Y) {
  if (!b || !b->stream || !X || !Y || fp64 < 0 || fp64 >= 2 + kNumVariants64) return HECMW33_ERROR_ARGS;
  const int64_t nv = b->nrows * KS * 3;
  std::vector<double> h(nv, 0.0);
  for (int s = 0; s < b->nsys; ++s) {
    const int64_t n3 = 3 * (int64_t)b->NP[s];
    for (int k = 0; k < b->nrhs; ++k)
      for (int32_t i = 0; i < b->N[s]; ++i)
        for (int d = 0; d < 3; ++d) h[((int64_t)(b->row0[s] + i) * KS + k) * 3 + d] = X[s][n3 * k + 3 * (int64_t)i + d];
  }
  hipStream_t st = b->stream;
  BMem xin, yout;
  if (xin.ensure(sizeof(double) * nv) || yout.ensure(sizeof(double) * nv)) return HECMW33_ERROR_DEVICE;
  BHIPC(hipMemcpy(xin.p, h.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
  if (fp64) {
    const int v = fp64 >= 2 ? fp64 - 2 : b->variant64;
    kVariants64[v].st(b->ntiles, st, b->T, qall(), b->rowptr.as<int32_t>(), b->col.as<int32_t>(), b->val64.p, xin.p,
                      nullptr, yout.p, b->S, nullptr, 0);
  } else {
    float* x32 = b->z.as<float>();  // scratch: the solve re-initialises it
    float* y32 = b->q.as<float>();
    kb_to_f32<<<grid_of(nv), kBT, 0, st>>>(nv, xin.as<double>(), x32);
    kVariants[b->variant].st32(b->ntiles, st, b->T, qall(), b->rowptr.as<int32_t>(), b->col.as<int32_t>(),
                               b->val32.p, x32, nullptr, y32, b->S, nullptr, 0);
    BHIPC(hipStreamSynchronize(st));
    std::vector<float> hf(nv);
    BHIPC(hipMemcpy(hf.data(), y32, sizeof(float) * nv, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nv; ++i) h[i] = hf[i];
  }
  BHIPC(hipGetLastError());
  BHIPC(hipStreamSynchronize(st));
  if (fp64) BHIPC(hipMemcpy(h.data(), yout.p, sizeof(double) * nv, hipMemcpyDeviceToHost));
  for (int s = 0; s < b->nsys; ++s) {
    const int64_t n3 = 3 * (int64_t)b->NP[s];
    for (int k = 0; k < b->nrhs; ++k)
      for (int32_t i = 0; i < b->N[s]; ++i)
        for (int d = 0; d < 3; ++d) Y[s][n3 * k + 3 * (int64_t)i + d] = h[((int64_t)(b->row0[s] + i) * KS + k) * 3 + d];
  }
  return 0;
}

// [0] rows, [1] blocks, [2] tiles, [3] FP32 shape, [4] nv32, [5..5+nv32) FP32 shape times in us (SpMM +
// dots; 0 when the shapes came from the persisted winners), FP64 shape, nv64, FP64 shape times in us,
// then: shapes from the persisted winners (0/1), CU count, the persistent grids of the p, q = A p,
// update, residual and DIAG launches
int32_t hecmw33_batch_info(hecmw33_batch* b, int64_t* info, int32_t n) {
  if (!b || !info) return HECMW33_ERROR_ARGS;
  std::vector<int64_t> v = {b->nrows, b->nnzb, b->ntiles, b->variant, kNumVariants};
  for (int i = 0; i < kNumVariants; ++i) v.push_back((int64_t)std::llround(b->variant_ms[i] * 1000.0));
  v.push_back(b->variant64);
  v.push_back(kNumVariants64);
  for (int i = 0; i < kNumVariants64; ++i) v.push_back((int64_t)std::llround(b->variant64_ms[i] * 1000.0));
  for (int64_t e : {(int64_t)b->tune_cached, (int64_t)b->ncu, (int64_t)b->grid_p, (int64_t)b->grid_q,
                    (int64_t)b->grid_upd, (int64_t)b->grid_res, (int64_t)b->grid_diag})
    v.push_back(e);
  for (int i = 0; i < n && i < (int)v.size(); ++i) info[i] = v[i];
  return (int32_t)v.size();
}

const char* hecmw33_batch_variant_name(int32_t v) {
  return v >= 0 && v < kNumVariants ? kVariants[v].name : "";
}

const char* hecmw33_batch_variant64_name(int32_t v) {
  return v >= 0 && v < kNumVariants64 ? kVariants64[v].name : "";
}

// One-shot: load, solve, copy every solution back.
int32_t hecmw33_solve_batch(int32_t nsys, const hecmw33_sys* sys, int32_t nrhs, const double* const* B,
                            double* const* X, const int32_t iarray[100], const double rarray[100], int32_t* iter,
                            int32_t* error, double* resid) {
  hecmw33_batch b;
  if (int32_t e = batch_load(b, nsys, sys, nrhs, B, rarray)) return e;
  if (int32_t e = batch_solve(b, iarray, rarray, iter, error, resid, nullptr)) return e;
  for (int s = 0; s < nsys; ++s)
    if (int32_t e = hecmw33_batch_get_x(&b, s, X[s])) return e;
  return 0;
}

}  // extern "C"
