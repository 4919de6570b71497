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
      const double* v = A.val + (base + (int64_t)k * A.rows) * 9 + l;
      double m[3][6];
      const double jx = dv[3 * (int64_t)j], jy = dv[3 * (int64_t)j + 1], jz = dv[3 * (int64_t)j + 2];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double a0 = v[(3 * r) * A.rows], a1 = v[(3 * r + 1) * A.rows], a2 = v[(3 * r + 2) * A.rows];
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

// c = Z^T r per aggregate (FP64 sums, fixed order)
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_restrict(const int32_t* __restrict__ aptr,
                                                                   const int32_t* __restrict__ anodes,
                                                                   const float* __restrict__ dv,
                                                                   const float* __restrict__ r, float* __restrict__ c,
                                                                   const KrylovState* st) {
  if (st->done) return;
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
  if (st->done) return;
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
  double d[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kCoarseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCoarseBlock) {
    const int a = agg[i];
    const float dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    const float* ya = y + 6 * a;
    const float t0 = ya[0], t1 = ya[1], t2 = ya[2], w0 = ya[3], w1 = ya[4], w2 = ya[5];
    const float z0 = z[3 * i] + (t0 + (w1 * dz - w2 * dy));
    const float z1 = z[3 * i + 1] + (t1 + (w2 * dx - w0 * dz));
    const float z2 = z[3 * i + 2] + (t2 + (w0 * dy - w1 * dx));
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

}  // namespace

void coarse_assemble(const Sell<double>& A, const CoarseOp& op, int32_t npair, const int32_t* pa, const int32_t* pb,
                     const int32_t* pn_ptr, const int32_t* pn, double* Eblk, double* E, int32_t lda, hipStream_t s) {
  if (npair <= 0) return;
  k_coarse_E<<<(unsigned)npair, kCoarseBlock, 0, s>>>(A, op.agg, pb, pn_ptr, pn, op.dv, Eblk);
  const int64_t ne = (int64_t)npair * 36;
  k_coarse_scatter<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(npair, pa, pb, Eblk, E, lda);
  k_coarse_unit_pivots<<<(op.nc + 255) / 256, 256, 0, s>>>(op.nc, E, lda);
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
