// Two-level correction for the flexible CG: a coarse space of piecewise
// rigid-body modes on geometric aggregates of nodes, added to the FP32 SSOR
// output as z = M^-1 r + Z E^-1 Z^T r with E = Z^T A Z.
//
// The SSOR sweeps damp the error components that vary from node to node;
// what is left after many iterations is smooth, close to a rigid motion of
// each region of the body.  Six modes per aggregate (three translations,
// three rotations about the aggregate's centroid) capture that part, and the
// dense coarse inverse removes it in one step.  Z^T A Z + M^-1 is symmetric
// positive definite, so the flexible CG's conjugacy is kept.
//
// Layout:
//  * agg[i]: aggregate of node i (NP entries, -1 for external nodes);
//  * aptr/anodes: the nodes of each aggregate, ascending;
//  * dv[3i..3i+2]: (x_i - centroid of agg(i)) / scale, FP32 -- the same numbers
//    define Z in the assembly, the restriction and the prolongation;
//  * pairs p = (a, b) of aggregates coupled by a matrix block, and per pair the
//    nodes of a whose rows touch b (pn_ptr/pn);
//  * W: E^-1 in FP32, row-major n_c x ldw (symmetric).
//
// Z_i (3x6) = [ I | R(d_i) ] with R's columns the rotations about x, y, z:
// (0, -dz, dy), (dz, 0, -dx), (-dy, dx, 0).
//
// When a learned solve drops its two-level step (KrylovState::coarse_off)
// the per-iteration kernels leave z to the SSOR sweeps: no coarse term, A z = w.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device.h"

namespace hecmw33 {

namespace {

constexpr int kCoarseBlock = 256;
constexpr int kCoarseWaves = kCoarseBlock / kWave;

__device__ __forceinline__ double cwave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Fixed-order sum of K values over the workgroup; thread k < K gets total k.
template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], double* out) {
  __shared__ double red[K][kCoarseWaves];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double w = cwave_sum(v[k]);
    if (lane == 0) red[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = red[threadIdx.x][0];
#pragma unroll
    for (int w = 1; w < kCoarseWaves; ++w) s += red[threadIdx.x][w];
    out[threadIdx.x] = s;
  }
}

// E block of the aggregate pair p = (a, b):
//   E_ab = sum over the nodes i of a that touch b, blocks j of row i with
//   agg(j) = b, of Z_i^T A_ij Z_j.
// One workgroup per pair; each thread sums its nodes' blocks in registers, then
// a fixed-order reduction, so E is the same on every run.
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_E(Sell<double> A, const int32_t* __restrict__ agg,
                                                            const int32_t* __restrict__ pb,
                                                            const int32_t* __restrict__ pn_ptr,
                                                            const int32_t* __restrict__ pn,
                                                            const float* __restrict__ dv, double* __restrict__ Eblk) {
  const int p = blockIdx.x;
  const int b = pb[p];
  double acc[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) acc[k] = 0.0;
  const int t1 = pn_ptr[p + 1];
  for (int t = pn_ptr[p] + threadIdx.x; t < t1; t += kCoarseBlock) {
    const int i = pn[t];
    const double ix = dv[3 * (int64_t)i], iy = dv[3 * (int64_t)i + 1], iz = dv[3 * (int64_t)i + 2];
    const int sl = i / A.rows, l = i % A.rows;
    const int64_t base = (int64_t)A.ptr[sl] * A.rows;
    const int len = A.len[i];
    for (int k = 0; k < len; ++k) {
      const int64_t slot = base + (int64_t)k * A.rows + l;
      const int j = A.col[slot];
      if (agg[j] != b) continue;
      double v[9];
      aval_load(A.val, (int64_t)A.ptr[sl] + k, l, v);
      double m[3][6];
      const double jx = dv[3 * (int64_t)j], jy = dv[3 * (int64_t)j + 1], jz = dv[3 * (int64_t)j + 2];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double a0 = v[3 * r], a1 = v[3 * r + 1], a2 = v[3 * r + 2];
        m[r][0] = a0;
        m[r][1] = a1;
        m[r][2] = a2;
        m[r][3] = -a1 * jz + a2 * jy;
        m[r][4] = a0 * jz - a2 * jx;
        m[r][5] = -a0 * jy + a1 * jx;
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        acc[q] += m[0][q];
        acc[6 + q] += m[1][q];
        acc[12 + q] += m[2][q];
        acc[18 + q] += -iz * m[1][q] + iy * m[2][q];
        acc[24 + q] += iz * m[0][q] - ix * m[2][q];
        acc[30 + q] += -iy * m[0][q] + ix * m[1][q];
      }
    }
  }
  block_reduce<36>(acc, Eblk + (int64_t)p * 36);
}

// dense E, column-major (the solver library's layout): E(6a+r, 6b+q) = Eblk[p][r][q]
__global__ void k_coarse_scatter(int32_t npair, const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                                 const double* __restrict__ Eblk, double* __restrict__ E, int32_t lda) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)npair * 36) return;
  const int p = (int)(t / 36), e = (int)(t % 36);
  const int a = pa[p], b = pb[p];
  const int r = e / 6, q = e % 6;
  E[(int64_t)(6 * b + q) * lda + 6 * a + r] = Eblk[t];
}

// unit pivot for a coarse unknown whose column of Z is zero (E(k,k) exactly 0):
// its restriction is exactly 0, so its value stays 0 and E stays invertible
__global__ void k_coarse_unit_pivots(int32_t nc, double* __restrict__ E, int32_t lda) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nc && E[(int64_t)k * lda + k] == 0.0) E[(int64_t)k * lda + k] = 1.0;
}

// W = E^-1 in FP32, full symmetric from the lower triangle the inverse leaves
__global__ void k_coarse_inv32(int32_t nc, const double* __restrict__ Ei, int32_t lda, float* __restrict__ W,
                               int32_t ldw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nc * ldw) return;
  const int row = (int)(t / ldw), col = (int)(t % ldw);
  float v = 0.f;
  if (col < nc) {
    const int lo = row > col ? row : col, hi = row > col ? col : row;  // (lo, hi) is in the lower triangle
    v = (float)Ei[(int64_t)hi * lda + lo];
  }
  W[t] = v;
}

// W = the lower triangle of Ei mirrored into a full FP64 array: L^-1 below and
// on the diagonal, L^-T above (the twelve-mode step: with the learned modes E
// is too ill-conditioned for FP32)
__global__ void k_coarse_inv64(int32_t nc, const double* __restrict__ Ei, int32_t lda, double* __restrict__ W,
                               int32_t ldw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nc * ldw) return;
  const int row = (int)(t / ldw), col = (int)(t % ldw);
  double v = 0.0;
  if (col < nc) {
    const int lo = row > col ? row : col, hi = row > col ? col : row;
    v = Ei[(int64_t)hi * lda + lo];
  }
  W[t] = v;
}

// y = W c in FP64, one wavefront per row, double2 loads (ldw a multiple of 64).
// W holds L^-1 in its lower triangle and L^-T in its upper one (the mirror):
// PART 1 takes the columns j <= row (y = L^-1 c), PART 2 the columns j >= row
// (y = L^-T c), each half of the row's bytes
template <int PART>
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_gemv64(int32_t nc, int32_t ldw, const double* __restrict__ W,
                                                                 const double* __restrict__ c, double* __restrict__ y,
                                                                 const KrylovState* st) {
  if (st->done || st->coarse_off) return;
  const int row = blockIdx.x * kCoarseWaves + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= nc) return;
  const double2* w = reinterpret_cast<const double2*>(W + (int64_t)row * ldw);
  const double2* c2 = reinterpret_cast<const double2*>(c);
  const int k0 = PART == 2 ? row / 2 : 0, k1 = PART == 1 ? row / 2 + 1 : ldw / 2;
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += kWave) {
    double2 a = w[k];
    const double2 b = c2[k];
    if (PART == 1) {
      if (2 * k > row) a.x = 0.0;
      if (2 * k + 1 > row) a.y = 0.0;
    } else {
      if (2 * k < row) a.x = 0.0;
      if (2 * k + 1 < row) a.y = 0.0;
    }
    acc += a.x * b.x + a.y * b.y;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
  if (lane == 0) y[row] = acc;
}

// c = Z^T r per aggregate (FP64 sums, fixed order)
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_restrict(const int32_t* __restrict__ aptr,
                                                                   const int32_t* __restrict__ anodes,
                                                                   const float* __restrict__ dv,
                                                                   const float* __restrict__ r, float* __restrict__ c,
                                                                   const KrylovState* st) {
  if (st->done || st->coarse_off) return;
  const int a = blockIdx.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int t1 = aptr[a + 1];
  for (int t = aptr[a] + threadIdx.x; t < t1; t += kCoarseBlock) {
    const int64_t i = anodes[t];
    const double rx = r[3 * i], ry = r[3 * i + 1], rz = r[3 * i + 2];
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    acc[0] += rx;
    acc[1] += ry;
    acc[2] += rz;
    acc[3] += -dz * ry + dy * rz;
    acc[4] += dz * rx - dx * rz;
    acc[5] += -dy * rx + dx * ry;
  }
  __shared__ double tot[6];
  block_reduce<6>(acc, tot);
  if (threadIdx.x < 6) c[6 * a + threadIdx.x] = (float)tot[threadIdx.x];
}

// y = W c, one wavefront per row, float4 loads of W
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_gemv(int32_t nc, int32_t ldw, const float* __restrict__ W,
                                                               const float* __restrict__ c, float* __restrict__ y,
                                                               const KrylovState* st) {
  if (st->done || st->coarse_off) return;
  const int row = blockIdx.x * kCoarseWaves + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= nc) return;
  const float4* w = reinterpret_cast<const float4*>(W + (int64_t)row * ldw);
  const float4* c4 = reinterpret_cast<const float4*>(c);
  float acc = 0.f;
  for (int k = lane; k < ldw / 4; k += kWave) {
    const float4 a = w[k], b = c4[k];
    acc += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
  if (lane == 0) y[row] = acc;
}

// z += Z y per node; with DOTS the partials of z.r and z.q for the flexible
// CG's beta (the layout of the other partial-sum kernels)
template <bool DOTS>
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_prolong(int32_t n, const int32_t* __restrict__ agg,
                                                                  const float* __restrict__ dv,
                                                                  const float* __restrict__ y, float* __restrict__ z,
                                                                  const float* __restrict__ r,
                                                                  const float* __restrict__ q, const KrylovState* st,
                                                                  double* part) {
  if (st->done) return;
  const bool off = st->coarse_off;
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    float z0 = z[3 * i], z1 = z[3 * i + 1], z2 = z[3 * i + 2];
    if (!off) {
      const int a = agg[i];
      const float dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
      const float* ya = y + 6 * a;
      const float t0 = ya[0], t1 = ya[1], t2 = ya[2], w0 = ya[3], w1 = ya[4], w2 = ya[5];
      z0 = z0 + (t0 + (w1 * dz - w2 * dy));
      z1 = z1 + (t1 + (w2 * dx - w0 * dz));
      z2 = z2 + (t2 + (w0 * dy - w1 * dx));
    }
    z[3 * i] = z0;
    z[3 * i + 1] = z1;
    z[3 * i + 2] = z2;
    if (DOTS) {
      d[0] += (double)z0 * (double)r[3 * i] + (double)z1 * (double)r[3 * i + 1] + (double)z2 * (double)r[3 * i + 2];
      d[1] += (double)z0 * (double)q[3 * i] + (double)z1 * (double)q[3 * i + 1] + (double)z2 * (double)q[3 * i + 2];
    }
  }
  if (DOTS) {
    __shared__ double tot[2];
    block_reduce<2>(d, tot);
    if (threadIdx.x < 2) part[threadIdx.x * kMaxPartials + blockIdx.x] = tot[threadIdx.x];
  }
}

// ---- symmetric two-level step: z1 = Z E^-1 Z^T r, z2 = z1 + M^-1 (r - A z1),
//      z = z2 + Z E^-1 Z^T (r - A z2), with A Z kept per node and neighbour
//      aggregate (W), so A z1 and the A Z part of A z need no SpMV ----------------
// A Z of row i restricted to aggregate b: sum over the row's blocks j with
// agg(j) = b of A_ij Z_j (3x6, FP64), and the row's scale
// sum_j max|A_ij| (1 + max|d_j|), which bounds every component.
__device__ __forceinline__ void row_AZ(const Sell<double>& A, int64_t i, int b, const int32_t* __restrict__ agg,
                                       const float* __restrict__ dv, double (&m)[18], double& scale) {
  const int sl = (int)(i / A.rows), l = (int)(i % A.rows);
  const int64_t base = (int64_t)A.ptr[sl] * A.rows;
  const int len = A.len[i];
#pragma unroll
  for (int k = 0; k < 18; ++k) m[k] = 0.0;
  scale = 0.0;
  for (int k = 0; k < len; ++k) {
    const int j = A.col[base + (int64_t)k * A.rows + l];
    if (agg[j] != b) continue;
    double v[9];
    aval_load(A.val, (int64_t)A.ptr[sl] + k, l, v);
    const double jx = dv[3 * (int64_t)j], jy = dv[3 * (int64_t)j + 1], jz = dv[3 * (int64_t)j + 2];
    double amax = 0.0;
#pragma unroll
    for (int e = 0; e < 9; ++e) amax = fmax(amax, fabs(v[e]));
    scale += amax * (1.0 + fmax(fabs(jx), fmax(fabs(jy), fabs(jz))));
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double a0 = v[3 * r], a1 = v[3 * r + 1], a2 = v[3 * r + 2];
      m[6 * r] += a0;
      m[6 * r + 1] += a1;
      m[6 * r + 2] += a2;
      m[6 * r + 3] += -a1 * jz + a2 * jy;
      m[6 * r + 4] += a0 * jz - a2 * jx;
      m[6 * r + 5] += -a0 * jy + a1 * jx;
    }
  }
}

// A Z per node and touched aggregate, compacted for the iterations: node i's
// entries are slots sc[i].x .. sc[i].x + sc[i].y - 1 of aggW (the aggregate)
// and W (its 3x6 FP32 block).
//  * A row whose blocks all lie in its own aggregate applies A to a whole
//    rigid motion: zero in exact arithmetic unless a boundary condition
//    changed the row.  What is left is the FP32 rounding of the offsets
//    (about 1e-9 of the row's scale), the level at which the FP32 SpMV of
//    the same iteration already works; such an entry is dropped when every
//    component is below 1e-6 of the scale, so the iterations do not stream
//    it.  Rows that touch other aggregates keep every entry.
//  * Workgroup g compacts its 256 nodes inside their own symbolic slots
//    (from ncnt[256 g] on), so no grid-wide scan is needed; kept[g] counts its
//    entries.
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_W(Sell<double> A, int32_t n, const int32_t* __restrict__ agg,
                                                            const int32_t* __restrict__ ncnt,
                                                            const int32_t* __restrict__ nagg,
                                                            const float* __restrict__ dv, int2* __restrict__ sc,
                                                            int32_t* __restrict__ aggW, float* __restrict__ W,
                                                            int32_t* __restrict__ kept) {
  const int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x;
  const bool valid = i < n;
  const int e0 = valid ? ncnt[i] : 0;
  const int ne = valid ? ncnt[i + 1] - e0 : 0;
  int cnt = ne;
  float w1[18];
  if (ne == 1) {
    double m[18], scale;
    row_AZ(A, i, nagg[e0], agg, dv, m, scale);
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      mx = fmax(mx, fabs(m[k]));
      w1[k] = (float)m[k];
    }
    cnt = mx > 1e-6 * scale ? 1 : 0;
  }
  // exclusive scan of cnt over the workgroup: per wave by shuffles, then the
  // wave totals in order
  __shared__ int wtot[kCoarseWaves];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int inc = cnt;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += t;
  }
  if (lane == kWave - 1) wtot[wave] = inc;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kCoarseWaves; ++w) {
    before += w < wave ? wtot[w] : 0;
    total += wtot[w];
  }
  const int slot = ncnt[(int64_t)blockIdx.x * kCoarseBlock] + before + inc - cnt;
  if (threadIdx.x == 0) kept[blockIdx.x] = total;
  if (!valid) return;
  sc[i] = make_int2(slot, cnt);
  if (ne == 1) {
    if (cnt) {
      aggW[slot] = nagg[e0];
#pragma unroll
      for (int k = 0; k < 18; ++k) W[18 * (int64_t)slot + k] = w1[k];
    }
    return;
  }
  for (int t = 0; t < ne; ++t) {
    double m[18], scale;
    row_AZ(A, i, nagg[e0 + t], agg, dv, m, scale);
    aggW[slot + t] = nagg[e0 + t];
#pragma unroll
    for (int k = 0; k < 18; ++k) W[18 * (int64_t)(slot + t) + k] = (float)m[k];
  }
}

__device__ __forceinline__ void zy(const float* __restrict__ dv, const float* __restrict__ ya, int64_t i, float (&o)[3]) {
  const float dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
  o[0] = ya[0] + (ya[4] * dz - ya[5] * dy);
  o[1] = ya[1] + (ya[5] * dx - ya[3] * dz);
  o[2] = ya[2] + (ya[3] * dy - ya[4] * dx);
}

__device__ __forceinline__ void wy(const int2* __restrict__ sc, const int32_t* __restrict__ aggW,
                                   const float* __restrict__ W, const float* __restrict__ y, int64_t i,
                                   float (&o)[3]) {
  o[0] = o[1] = o[2] = 0.f;
  const int2 r = sc[i];
  for (int e = r.x; e < r.x + r.y; ++e) {
    const float* w = W + 18 * (int64_t)e;
    const float* yb = y + 6 * aggW[e];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 6; ++q) o[r] += w[6 * r + q] * yb[q];
  }
}

// z = Z y, r1 = r - (A Z) y
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_pre(int32_t n, const int32_t* __restrict__ agg,
                                                              const float* __restrict__ dv,
                                                              const int2* __restrict__ sc,
                                                              const int32_t* __restrict__ aggW,
                                                              const float* __restrict__ W, const float* __restrict__ y,
                                                              const float* __restrict__ r, float* __restrict__ z,
                                                              float* __restrict__ r1, const KrylovState* st) {
  if (st->done) return;
  const bool off = st->coarse_off;
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    float zi[3] = {0.f, 0.f, 0.f}, ai[3] = {0.f, 0.f, 0.f};
    if (!off) {
      zy(dv, y + 6 * agg[i], i, zi);
      wy(sc, aggW, W, y, i, ai);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      z[3 * i + d] = zi[d];
      r1[3 * i + d] = r[3 * i + d] - ai[d];
    }
  }
}

// z = z + t (the SSOR output added to the coarse part)
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_add(int64_t n3, const float* __restrict__ t,
                                                              float* __restrict__ z, const KrylovState* st) {
  if (st->done) return;
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kCoarseBlock)
    z[i] = z[i] + t[i];
}

// c = Z^T (r - w) per aggregate
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_restrict2(const int32_t* __restrict__ aptr,
                                                                    const int32_t* __restrict__ anodes,
                                                                    const float* __restrict__ dv,
                                                                    const float* __restrict__ r,
                                                                    const float* __restrict__ w,
                                                                    float* __restrict__ c, const KrylovState* st) {
  if (st->done || st->coarse_off) return;
  const int a = blockIdx.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int t1 = aptr[a + 1];
  for (int t = aptr[a] + threadIdx.x; t < t1; t += kCoarseBlock) {
    const int64_t i = anodes[t];
    const double rx = (double)r[3 * i] - (double)w[3 * i], ry = (double)r[3 * i + 1] - (double)w[3 * i + 1],
                 rz = (double)r[3 * i + 2] - (double)w[3 * i + 2];
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    acc[0] += rx;
    acc[1] += ry;
    acc[2] += rz;
    acc[3] += -dz * ry + dy * rz;
    acc[4] += dz * rx - dx * rz;
    acc[5] += -dy * rx + dx * ry;
  }
  __shared__ double tot[6];
  block_reduce<6>(acc, tot);
  if (threadIdx.x < 6) c[6 * a + threadIdx.x] = (float)tot[threadIdx.x];
}

// z = z + Z y, az = w + (A Z) y, partials r.z and q_prev.z (the flexible CG's beta)
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_post(int32_t n, const int32_t* __restrict__ agg,
                                                               const float* __restrict__ dv,
                                                               const int2* __restrict__ sc,
                                                               const int32_t* __restrict__ aggW,
                                                               const float* __restrict__ W,
                                                               const float* __restrict__ y, float* __restrict__ z,
                                                               const float* __restrict__ w, float* __restrict__ az,
                                                               const float* __restrict__ r,
                                                               const float* __restrict__ q, const KrylovState* st,
                                                               double* part) {
  if (st->done) return;
  const bool off = st->coarse_off;
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    float zi[3] = {0.f, 0.f, 0.f}, ai[3] = {0.f, 0.f, 0.f};
    if (!off) {
      zy(dv, y + 6 * agg[i], i, zi);
      wy(sc, aggW, W, y, i, ai);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float zn = z[3 * i + k] + zi[k];
      z[3 * i + k] = zn;
      az[3 * i + k] = w[3 * i + k] + ai[k];
      d[0] += (double)zn * (double)r[3 * i + k];
      d[1] += (double)zn * (double)q[3 * i + k];
    }
  }
  __shared__ double tot[2];
  block_reduce<2>(d, tot);
  if (threadIdx.x < 2) part[threadIdx.x * kMaxPartials + blockIdx.x] = tot[threadIdx.x];
}

// p = z + beta p, q = az + beta q (first iteration: z, az), partials p.q and r.p
__global__ __launch_bounds__(kCoarseBlock) void k_fcg_pq(int64_t n3, const float* __restrict__ z,
                                                          const float* __restrict__ az, float* __restrict__ p,
                                                          float* __restrict__ q, const float* __restrict__ r,
                                                          const KrylovState* st, double* part) {
  if (st->done) return;
  const bool first = st->iter == 1;
  const double beta = st->beta;
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kCoarseBlock) {
    const float pn = first ? z[i] : (float)((double)z[i] + beta * (double)p[i]);
    const float qn = first ? az[i] : (float)((double)az[i] + beta * (double)q[i]);
    p[i] = pn;
    q[i] = qn;
    d[0] += (double)pn * (double)qn;
    d[1] += (double)r[i] * (double)pn;
  }
  __shared__ double tot[2];
  block_reduce<2>(d, tot);
  if (threadIdx.x < 2) part[threadIdx.x * kMaxPartials + blockIdx.x] = tot[threadIdx.x];
}

// ---- twelve modes per aggregate: the six rigid-body modes and six learned ones -------------
// Z_i (3x12) = [ I | R(d_i) | X_i ], X_i (3x6) = the learned modes' values at node i,
// xm[18 i + 3 j + d] (mode j, component d).  E, A Z and the per-iteration kernels
// follow the six-mode ones; E_ab is assembled as four 6x6 tiles.

// the 3x6 block t of Z_i: t = 0 the rigid-body modes, t = 1 the learned ones
__device__ __forceinline__ void zblock(const float* __restrict__ dv, const float* __restrict__ xm, int64_t i, int t,
                                       double (&z)[3][6]) {
  if (t == 0) {
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    z[0][0] = 1.0, z[0][1] = 0.0, z[0][2] = 0.0, z[0][3] = 0.0, z[0][4] = dz, z[0][5] = -dy;
    z[1][0] = 0.0, z[1][1] = 1.0, z[1][2] = 0.0, z[1][3] = -dz, z[1][4] = 0.0, z[1][5] = dx;
    z[2][0] = 0.0, z[2][1] = 0.0, z[2][2] = 1.0, z[2][3] = dy, z[2][4] = -dx, z[2][5] = 0.0;
  } else {
    const float* x = xm + 18 * i;
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int d = 0; d < 3; ++d) z[d][j] = x[3 * j + d];
  }
}

// tile (tr, tc) = blockIdx.y of E_ab: rows 6 tr.., columns 6 tc.. of Z_a^T A Z_b
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_E12(Sell<double> A, const int32_t* __restrict__ agg,
                                                              const int32_t* __restrict__ pb,
                                                              const int32_t* __restrict__ pn_ptr,
                                                              const int32_t* __restrict__ pn,
                                                              const float* __restrict__ dv,
                                                              const float* __restrict__ xm, double* __restrict__ Eblk) {
  const int p = blockIdx.x, tr = blockIdx.y >> 1, tc = blockIdx.y & 1;
  const int b = pb[p];
  double acc[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) acc[k] = 0.0;
  const int t1 = pn_ptr[p + 1];
  for (int t = pn_ptr[p] + threadIdx.x; t < t1; t += kCoarseBlock) {
    const int i = pn[t];
    double zi[3][6];
    zblock(dv, xm, i, tr, zi);
    const int sl = i / A.rows, l = i % A.rows;
    const int64_t base = (int64_t)A.ptr[sl] * A.rows;
    const int len = A.len[i];
    for (int k = 0; k < len; ++k) {
      const int j = A.col[base + (int64_t)k * A.rows + l];
      if (agg[j] != b) continue;
      double v[9], zj[3][6];
      aval_load(A.val, (int64_t)A.ptr[sl] + k, l, v);
      zblock(dv, xm, j, tc, zj);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          // (A_ij Z_j)[r][q], then Z_i^T of it
          const double m = v[3 * r] * zj[0][q] + v[3 * r + 1] * zj[1][q] + v[3 * r + 2] * zj[2][q];
#pragma unroll
          for (int u = 0; u < 6; ++u) acc[6 * u + q] += zi[r][u] * m;
        }
    }
  }
  block_reduce<36>(acc, Eblk + ((int64_t)p * 4 + blockIdx.y) * 36);
}

// dense E (column-major): E(12a + 6tr + r, 12b + 6tc + q) = tile (tr, tc) of pair p
__global__ void k_coarse_scatter12(int32_t npair, const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                                   const double* __restrict__ Eblk, double* __restrict__ E, int32_t lda) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)npair * 144) return;
  const int p = (int)(t / 144), e = (int)(t % 144);
  const int tile = e / 36, tr = tile >> 1, tc = tile & 1, r = (e % 36) / 6, q = e % 6;
  const int a = pa[p], b = pb[p];
  E[(int64_t)(12 * b + 6 * tc + q) * lda + 12 * a + 6 * tr + r] = Eblk[t];
}

// A Z per node and touched aggregate, 3x12 FP32 per entry, every entry kept
// (the learned modes do not vanish under A inside their aggregate); the
// entries of node i are its symbolic slots ncnt[i].. (sc[i] = (ncnt[i], count))
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_W12(Sell<double> A, int32_t n,
                                                              const int32_t* __restrict__ agg,
                                                              const int32_t* __restrict__ ncnt,
                                                              const int32_t* __restrict__ nagg,
                                                              const float* __restrict__ dv,
                                                              const float* __restrict__ xm, int2* __restrict__ sc,
                                                              int32_t* __restrict__ aggW, float* __restrict__ W,
                                                              int32_t* __restrict__ kept) {
  const int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x;
  if (threadIdx.x == 0) {
    const int64_t i0 = (int64_t)blockIdx.x * kCoarseBlock, i1 = i0 + kCoarseBlock < n ? i0 + kCoarseBlock : n;
    kept[blockIdx.x] = ncnt[i1] - ncnt[i0];
  }
  if (i >= n) return;
  const int e0 = ncnt[i], ne = ncnt[i + 1] - e0;
  sc[i] = make_int2(e0, ne);
  const int sl = (int)(i / A.rows), l = (int)(i % A.rows);
  const int64_t base = (int64_t)A.ptr[sl] * A.rows;
  const int len = A.len[i];
  for (int t = 0; t < ne; ++t) {
    const int b = nagg[e0 + t];
    double m[3][12];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 12; ++q) m[r][q] = 0.0;
    for (int k = 0; k < len; ++k) {
      const int j = A.col[base + (int64_t)k * A.rows + l];
      if (agg[j] != b) continue;
      double v[9];
      aval_load(A.val, (int64_t)A.ptr[sl] + k, l, v);
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) {
        double zj[3][6];
        zblock(dv, xm, j, tb, zj);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int q = 0; q < 6; ++q)
            m[r][6 * tb + q] += v[3 * r] * zj[0][q] + v[3 * r + 1] * zj[1][q] + v[3 * r + 2] * zj[2][q];
      }
    }
    aggW[e0 + t] = b;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 12; ++q) W[36 * (int64_t)(e0 + t) + 12 * r + q] = (float)m[r][q];
  }
}

// c = Z^T (r - w) per aggregate, 12 sums (w == nullptr: Z^T r)
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_restrict12(const int32_t* __restrict__ aptr,
                                                                     const int32_t* __restrict__ anodes,
                                                                     const float* __restrict__ dv,
                                                                     const float* __restrict__ xm,
                                                                     const float* __restrict__ r,
                                                                     const float* __restrict__ w,
                                                                     double* __restrict__ c, const KrylovState* st) {
  if (st->done || st->coarse_off) return;
  const int a = blockIdx.x;
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  const int t1 = aptr[a + 1];
  for (int t = aptr[a] + threadIdx.x; t < t1; t += kCoarseBlock) {
    const int64_t i = anodes[t];
    double rv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) rv[d] = w ? (double)r[3 * i + d] - (double)w[3 * i + d] : (double)r[3 * i + d];
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    acc[0] += rv[0];
    acc[1] += rv[1];
    acc[2] += rv[2];
    acc[3] += -dz * rv[1] + dy * rv[2];
    acc[4] += dz * rv[0] - dx * rv[2];
    acc[5] += -dy * rv[0] + dx * rv[1];
    const float* x = xm + 18 * i;
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[6 + j] += (double)x[3 * j] * rv[0] + (double)x[3 * j + 1] * rv[1] +
                                              (double)x[3 * j + 2] * rv[2];
  }
  __shared__ double tot[12];
  block_reduce<12>(acc, tot);
  if (threadIdx.x < 12) c[12 * a + threadIdx.x] = tot[threadIdx.x];
}

// Z_i y_a and (A Z)_i y with FP64 coarse values, accumulated in FP64
__device__ __forceinline__ void zy12(const float* __restrict__ dv, const float* __restrict__ xm,
                                     const double* __restrict__ ya, int64_t i, float (&o)[3]) {
  const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
  double u0 = ya[0] + (ya[4] * dz - ya[5] * dy);
  double u1 = ya[1] + (ya[5] * dx - ya[3] * dz);
  double u2 = ya[2] + (ya[3] * dy - ya[4] * dx);
  const float* x = xm + 18 * i;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    u0 += (double)x[3 * j] * ya[6 + j];
    u1 += (double)x[3 * j + 1] * ya[6 + j];
    u2 += (double)x[3 * j + 2] * ya[6 + j];
  }
  o[0] = (float)u0;
  o[1] = (float)u1;
  o[2] = (float)u2;
}

__device__ __forceinline__ void wy12(const int2* __restrict__ sc, const int32_t* __restrict__ aggW,
                                     const float* __restrict__ W, const double* __restrict__ y, int64_t i,
                                     float (&o)[3]) {
  double u[3] = {0.0, 0.0, 0.0};
  const int2 r = sc[i];
  for (int e = r.x; e < r.x + r.y; ++e) {
    const float* w = W + 36 * (int64_t)e;
    const double* yb = y + 12 * aggW[e];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int q = 0; q < 12; ++q) u[rr] += (double)w[12 * rr + q] * yb[q];
  }
  o[0] = (float)u[0];
  o[1] = (float)u[1];
  o[2] = (float)u[2];
}

// z = Z y, r1 = r - (A Z) y
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_pre12(int32_t n, const int32_t* __restrict__ agg,
                                                                const float* __restrict__ dv,
                                                                const float* __restrict__ xm,
                                                                const int2* __restrict__ sc,
                                                                const int32_t* __restrict__ aggW,
                                                                const float* __restrict__ W,
                                                                const double* __restrict__ y,
                                                                const float* __restrict__ r, float* __restrict__ z,
                                                                float* __restrict__ r1, const KrylovState* st) {
  if (st->done) return;
  const bool off = st->coarse_off;
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    float zi[3] = {0.f, 0.f, 0.f}, ai[3] = {0.f, 0.f, 0.f};
    if (!off) {
      zy12(dv, xm, y + 12 * agg[i], i, zi);
      wy12(sc, aggW, W, y, i, ai);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      z[3 * i + d] = zi[d];
      r1[3 * i + d] = r[3 * i + d] - ai[d];
    }
  }
}

// z = z + Z y, az = w + (A Z) y, partials r.z and q_prev.z
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_post12(int32_t n, const int32_t* __restrict__ agg,
                                                                 const float* __restrict__ dv,
                                                                 const float* __restrict__ xm,
                                                                 const int2* __restrict__ sc,
                                                                 const int32_t* __restrict__ aggW,
                                                                 const float* __restrict__ W,
                                                                 const double* __restrict__ y, float* __restrict__ z,
                                                                 const float* __restrict__ w, float* __restrict__ az,
                                                                 const float* __restrict__ r,
                                                                 const float* __restrict__ q, const KrylovState* st,
                                                                 double* part) {
  if (st->done) return;
  const bool off = st->coarse_off;
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    float zi[3] = {0.f, 0.f, 0.f}, ai[3] = {0.f, 0.f, 0.f};
    if (!off) {
      zy12(dv, xm, y + 12 * agg[i], i, zi);
      wy12(sc, aggW, W, y, i, ai);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float zn = z[3 * i + k] + zi[k];
      z[3 * i + k] = zn;
      az[3 * i + k] = w[3 * i + k] + ai[k];
      d[0] += (double)zn * (double)r[3 * i + k];
      d[1] += (double)zn * (double)q[3 * i + k];
    }
  }
  __shared__ double tot[2];
  block_reduce<2>(d, tot);
  if (threadIdx.x < 2) part[threadIdx.x * kMaxPartials + blockIdx.x] = tot[threadIdx.x];
}

}  // namespace

void coarse_build_W(const Sell<double>& A, const CoarseOp& op, int32_t n, const int32_t* ncnt, const int32_t* nagg,
                    int2* sc, int32_t* aggW, float* W, int32_t* kept, hipStream_t s) {
  if (n <= 0) return;
  const int g = (n + kCoarseBlock - 1) / kCoarseBlock;
  if (op.nm == 12)
    k_coarse_W12<<<g, kCoarseBlock, 0, s>>>(A, n, op.agg, ncnt, nagg, op.dv, op.xm, sc, aggW, W, kept);
  else
    k_coarse_W<<<g, kCoarseBlock, 0, s>>>(A, n, op.agg, ncnt, nagg, op.dv, sc, aggW, W, kept);
}

static int coarse_vec_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kCoarseBlock - 1) / kCoarseBlock, 4096));
}

void coarse_sym_pre(const CoarseOp& op, int32_t n, const int2* sc, const int32_t* aggW, const float* W,
                    const float* r, float* z, float* r1, const KrylovState* st, hipStream_t s) {
  const int gr = (op.nc + kCoarseWaves - 1) / kCoarseWaves;
  if (op.nm == 12) {
    k_coarse_restrict12<<<op.n_agg, kCoarseBlock, 0, s>>>(op.aptr, op.anodes, op.dv, op.xm, r, nullptr, op.cv64, st);
    k_coarse_gemv64<1><<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W64, op.cv64, op.tv64, st);
    k_coarse_gemv64<2><<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W64, op.tv64, op.yv64, st);
    k_coarse_pre12<<<coarse_vec_grid(n), kCoarseBlock, 0, s>>>(n, op.agg, op.dv, op.xm, sc, aggW, W, op.yv64, r, z,
                                                               r1, st);
  } else {
    k_coarse_restrict<<<op.n_agg, kCoarseBlock, 0, s>>>(op.aptr, op.anodes, op.dv, r, op.cv, st);
    k_coarse_gemv<<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W, op.cv, op.yv, st);
    k_coarse_pre<<<coarse_vec_grid(n), kCoarseBlock, 0, s>>>(n, op.agg, op.dv, sc, aggW, W, op.yv, r, z, r1, st);
  }
}

void coarse_sym_add(int64_t n3, const float* t, float* z, const KrylovState* st, hipStream_t s) {
  k_coarse_add<<<coarse_vec_grid(n3), kCoarseBlock, 0, s>>>(n3, t, z, st);
}

int coarse_sym_post(const CoarseOp& op, int32_t n, const int2* sc, const int32_t* aggW, const float* W,
                    const float* r, const float* w, float* z, float* az, const float* q, const KrylovState* st,
                    double* part, hipStream_t s) {
  const int gr = (op.nc + kCoarseWaves - 1) / kCoarseWaves;
  const int g = coarse_vec_grid(n);
  if (op.nm == 12) {
    k_coarse_restrict12<<<op.n_agg, kCoarseBlock, 0, s>>>(op.aptr, op.anodes, op.dv, op.xm, r, w, op.cv64, st);
    k_coarse_gemv64<1><<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W64, op.cv64, op.tv64, st);
    k_coarse_gemv64<2><<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W64, op.tv64, op.yv64, st);
    k_coarse_post12<<<g, kCoarseBlock, 0, s>>>(n, op.agg, op.dv, op.xm, sc, aggW, W, op.yv64, z, w, az, r, q, st,
                                               part);
  } else {
    k_coarse_restrict2<<<op.n_agg, kCoarseBlock, 0, s>>>(op.aptr, op.anodes, op.dv, r, w, op.cv, st);
    k_coarse_gemv<<<gr, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W, op.cv, op.yv, st);
    k_coarse_post<<<g, kCoarseBlock, 0, s>>>(n, op.agg, op.dv, sc, aggW, W, op.yv, z, w, az, r, q, st, part);
  }
  return g;
}

int coarse_fcg_pq(int64_t n3, const float* z, const float* az, float* p, float* q, const float* r,
                  const KrylovState* st, double* part, hipStream_t s) {
  const int g = coarse_vec_grid(n3);
  k_fcg_pq<<<g, kCoarseBlock, 0, s>>>(n3, z, az, p, q, r, st, part);
  return g;
}

void coarse_assemble(const Sell<double>& A, const CoarseOp& op, int32_t npair, const int32_t* pa, const int32_t* pb,
                     const int32_t* pn_ptr, const int32_t* pn, double* Eblk, double* E, int32_t lda, hipStream_t s) {
  if (npair <= 0) return;
  if (op.nm == 12) {
    k_coarse_E12<<<dim3((unsigned)npair, 4), kCoarseBlock, 0, s>>>(A, op.agg, pb, pn_ptr, pn, op.dv, op.xm, Eblk);
    const int64_t ne = (int64_t)npair * 144;
    k_coarse_scatter12<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(npair, pa, pb, Eblk, E, lda);
  } else {
    k_coarse_E<<<(unsigned)npair, kCoarseBlock, 0, s>>>(A, op.agg, pb, pn_ptr, pn, op.dv, Eblk);
    const int64_t ne = (int64_t)npair * 36;
    k_coarse_scatter<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(npair, pa, pb, Eblk, E, lda);
  }
  k_coarse_unit_pivots<<<(op.nc + 255) / 256, 256, 0, s>>>(op.nc, E, lda);
}

void coarse_inverse_f64(const CoarseOp& op, const double* Ei, int32_t lda, hipStream_t s) {
  const int64_t ne = (int64_t)op.nc * op.ldw;
  k_coarse_inv64<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(op.nc, Ei, lda, op.W64, op.ldw);
}

void coarse_inverse_f32(const CoarseOp& op, const double* Ei, int32_t lda, hipStream_t s) {
  const int64_t ne = (int64_t)op.nc * op.ldw;
  k_coarse_inv32<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(op.nc, Ei, lda, op.W, op.ldw);
}

int coarse_apply(const CoarseOp& op, int32_t n, const float* r, float* z, const float* q, const KrylovState* st,
                 double* part, bool dots, hipStream_t s) {
  k_coarse_restrict<<<op.n_agg, kCoarseBlock, 0, s>>>(op.aptr, op.anodes, op.dv, r, op.cv, st);
  k_coarse_gemv<<<(op.nc + kCoarseWaves - 1) / kCoarseWaves, kCoarseBlock, 0, s>>>(op.nc, op.ldw, op.W, op.cv, op.yv,
                                                                                   st);
  int g = (int)std::min<int64_t>(((int64_t)n + kCoarseBlock - 1) / kCoarseBlock, 4096);
  if (g < 1) g = 1;
  if (dots)
    k_coarse_prolong<true><<<g, kCoarseBlock, 0, s>>>(n, op.agg, op.dv, op.yv, z, r, q, st, part);
  else
    k_coarse_prolong<false><<<g, kCoarseBlock, 0, s>>>(n, op.agg, op.dv, op.yv, z, r, q, st, part);
  return g;
}

}  // namespace hecmw33
