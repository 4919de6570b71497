// Learned coarse space for the flexible CG: a per-node MLP evaluated and
// trained with BF16 matrix-core (MFMA) kernels accumulating in FP32.
//
// No counterpart in the reference (its preconditioners are SSOR, DIAG and
// the ones hecmw_precond_setup dispatches to, hecmw_precond.f90:30-58).
//
// The two-level step of the flexible CG takes its coarse space from
// aggregates of nodes (six rigid-body modes each).  Geometric cells alone
// miss what dominates the slow error of high-contrast bodies: a stiff
// inclusion moves as one rigid piece inside softer material that deforms
// around it.  The network predicts, per node, the logit y_i of "the slow
// error moves this node rigidly with a stiff piece"; the nodes with y_i > 0
// of every cell are split off into one aggregate per connected piece, whose
// rigid-body modes then enter the coarse space.  The preconditioner stays the
// symmetric positive definite two-level step; only its aggregation is
// learned.
//
// Training targets come from the solver itself: errors e = x* - x_k = A^-1 r_k
// of multicolour-SSOR CG iterates on sample systems.  An edge (i, j) of the
// matrix graph whose axial strain |(e_j - e_i).(x_j - x_i)| / |x_j - x_i|^2
// stays below tau times the sample's median in every sample lies inside a
// rigidly moving region, and its nodes are labelled 1 (solver.hip,
// rigid_labels).  The loss is the logistic loss of y against the labels.
//
// Features g_i (16): m_i = tanh((ln tr D_ii - mean)/2), the mean, maximum and
// minimum of m over the node's stencil neighbours, m_i^2, the local contrast
// max - min, 1 (a bias), the stencil size / 27, zeros.
//
// Network: 16 -> 64 -> ReLU -> 64 -> ReLU -> 3 (output 0 is the logit; 1 and 2
// are unused), no biases beyond the constant feature.  Layouts (row-major
// [k][n], BF16): W1 [16][64], W2 [64][64], W3 [64][16] (columns 3..15 zero).
// A v_mfma_f32_16x16x32_bf16 tile takes, in lane l, A[i = l%16][k = 8(l/16)+e]
// and B[k = 8(l/16)+e][j = l%16], e = 0..7, and returns D[i = 4(l/16)+v][j =
// l%16], v = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.h"

namespace hecmw33 {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kNT = 256;  // threads per workgroup (4 waves)
constexpr int kFt = 16, kHid = 64, kOut = 16;

struct NetW {
  const __bf16 *W1, *W2, *W3;
};

__device__ __forceinline__ floatx4 mfma(const bf16x8& a, const bf16x8& b, const floatx4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// B operand of W [K][N] (row-major), k-block kb, column tile nt; rows >= K are zero
__device__ __forceinline__ bf16x8 load_b(const __bf16* W, int K, int N, int kb, int nt, int lane) {
  bf16x8 v;
  const int n = nt * 16 + (lane & 15), k0 = kb * 32 + 8 * (lane >> 4);
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (k0 + e < K) ? W[(k0 + e) * N + n] : (__bf16)0.0f;
  return v;
}
// B operand of W^T where W is [N][K] row-major: B[k][n] = W[n][k]; k >= K zero
__device__ __forceinline__ bf16x8 load_bt(const __bf16* W, int K, int kb, int nt, int lane) {
  bf16x8 v;
  const int n = nt * 16 + (lane & 15), k0 = kb * 32 + 8 * (lane >> 4);
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (k0 + e < K) ? W[n * K + k0 + e] : (__bf16)0.0f;
  return v;
}

// Forward of one 16-node tile from the layer-1 A operand; hs: this wave's
// [16][64] BF16 staging.  Returns the output tile (lane: rows 4g+v, col l%16).
// If H1/H2 are given, the post-ReLU activations of valid nodes are stored.
__device__ __forceinline__ floatx4 mlp_tile(const bf16x8& a1, const bf16x8 (&w1)[4], const bf16x8 (&w2)[2][4],
                                            const bf16x8 (&w3)[2], __bf16 (*hs)[kHid], int lane, int64_t node0,
                                            int64_t n, __bf16* H1, __bf16* H2) {
  const int g = lane >> 4, j = lane & 15;
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const floatx4 c = mfma(a1, w1[nt], zero);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __bf16 h = (__bf16)fmaxf(c[v], 0.0f);
      hs[4 * g + v][nt * 16 + j] = h;
      const int64_t node = node0 + 4 * g + v;
      if (H1 && node < n) H1[node * kHid + nt * 16 + j] = h;
    }
  }
  __syncthreads();
  bf16x8 a2[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) a2[kb] = *reinterpret_cast<const bf16x8*>(&hs[j][kb * 32 + 8 * g]);
  floatx4 c2[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    c2[nt] = mfma(a2[0], w2[0][nt], zero);
    c2[nt] = mfma(a2[1], w2[1][nt], c2[nt]);
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __bf16 h = (__bf16)fmaxf(c2[nt][v], 0.0f);
      hs[4 * g + v][nt * 16 + j] = h;
      const int64_t node = node0 + 4 * g + v;
      if (H2 && node < n) H2[node * kHid + nt * 16 + j] = h;
    }
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) a2[kb] = *reinterpret_cast<const bf16x8*>(&hs[j][kb * 32 + 8 * g]);
  floatx4 c3 = mfma(a2[0], w3[0], zero);
  c3 = mfma(a2[1], w3[1], c3);
  __syncthreads();
  return c3;
}

__device__ __forceinline__ void load_weights(const NetW& W, int lane, bf16x8 (&w1)[4], bf16x8 (&w2)[2][4],
                                             bf16x8 (&w3)[2]) {
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) w1[nt] = load_b(W.W1, kFt, kHid, 0, nt, lane);
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) w2[kb][nt] = load_b(W.W2, kHid, kHid, kb, nt, lane);
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) w3[kb] = load_b(W.W3, kHid, kOut, kb, 0, lane);
}

// fixed-order block reduction of K doubles, partial to part[k*kMaxPartials + slot]
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* part, int slot) {
  __shared__ double red[K][kNT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double w = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    if (lane == 0) red[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = red[k][0];
      for (int w = 1; w < kNT / 64; ++w) s += red[k][w];
      part[(int64_t)k * kMaxPartials + slot] = s;
    }
}

// ---- static features and the linear combination -------------------------------------
// g_i from the node stiffness m and A's stencil (the SELL-64 structure of A:
// block 0 of a row is its diagonal block)
__global__ __launch_bounds__(kNT) void k_static_features(Sell<double> A, int64_t n, const float* __restrict__ mat,
                                                         __bf16* __restrict__ X) {
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT) {
    const int64_t s = i / kRowsA, lane = i % kRowsA;
    const int base = A.ptr[s];
    const int len = A.len[s * kRowsA + lane];
    const float mi = mat[i];
    float sum = 0.0f, mx = mi, mn = mi;
    for (int k = 1; k < len; ++k) {
      const float mj = mat[A.col[((int64_t)base + k) * kRowsA + lane]];
      sum += mj;
      mx = fmaxf(mx, mj);
      mn = fminf(mn, mj);
    }
    const float mean = len > 1 ? sum / (float)(len - 1) : mi;
    const float f[16] = {mi, mean, mx, mn, mi * mi, mx - mn, 1.0f, (float)len / 27.0f, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 16; ++e) X[i * kFt + e] = (__bf16)f[e];
  }
}

// logistic loss of output 0 against labels t: l_i = softplus(y_i) - t_i y_i;
// dY_i = ((sigmoid(y_i) - t_i) / n, 0, 0); FP64 partials of the loss and of
// the nodes classified as their label (y_i > 0 exactly when t_i >= 1/2)
__global__ __launch_bounds__(kNT) void k_net_bce(int64_t n, const float* __restrict__ Y,
                                                 const float* __restrict__ t, float* __restrict__ dY, double* part) {
  double acc[2] = {0.0, 0.0};
  const float inv = 1.0f / (float)n;
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT) {
    const float y = Y[3 * i], ti = t[i];
    const float sp = y > 20.0f ? y : log1pf(expf(y));
    const float sg = 1.0f / (1.0f + expf(-y));
    acc[0] += (double)(sp - ti * y);
    acc[1] += ((y > 0.0f) == (ti >= 0.5f)) ? 1.0 : 0.0;
    dY[3 * i] = (sg - ti) * inv;
    dY[3 * i + 1] = 0.0f;
    dY[3 * i + 2] = 0.0f;
  }
  block_sum<2>(acc, part, blockIdx.x);
}

// the nodes the network splits off: mask_i = (y_i > 0)
__global__ void k_net_mask(int64_t n, const float* __restrict__ Y, uint8_t* __restrict__ mask) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    mask[i] = Y[3 * i] > 0.0f ? 1 : 0;
}

// ---- training -------------------------------------------------------------------
// forward on stored features; keeps H1, H2; writes z = z0 + N(x) (z0 may be
// null: z = N(x), the raw network output)
__global__ __launch_bounds__(kNT) void k_net_fwd(int64_t n, NetW W, const __bf16* __restrict__ X,
                                                 __bf16* __restrict__ H1, __bf16* __restrict__ H2,
                                                 const float* __restrict__ z0, float* __restrict__ z) {
  __shared__ __attribute__((aligned(16))) __bf16 hs[kNT / 64][16][kHid];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
  bf16x8 w1[4], w2[2][4], w3[2];
  load_weights(W, lane, w1, w2, w3);
  const int64_t ntile = (n + 15) / 16;
  const int64_t per_pass = (int64_t)gridDim.x * (kNT / 64);
  const int64_t npass = (ntile + per_pass - 1) / per_pass;
  for (int64_t ps = 0; ps < npass; ++ps) {
    const int64_t t = ps * per_pass + (int64_t)blockIdx.x * (kNT / 64) + wave;
    const int64_t node0 = t * 16;
    const int64_t mynode = node0 + j;
    const bool valid = t < ntile && mynode < n;
    bf16x8 a1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * g + e;
      a1[e] = (valid && k < kFt) ? X[mynode * kFt + k] : (__bf16)0.0f;
    }
    const floatx4 c3 = mlp_tile(a1, w1, w2, w3, hs[wave], lane, node0, n, H1, H2);
    if (j < 3 && t < ntile)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t node = node0 + 4 * g + v;
        if (node < n) z[3 * node + j] = (z0 ? z0[3 * node + j] : 0.0f) + c3[v];
      }
  }
}

// Backward: from dY (n x 3, FP32) to the weight gradients.  Each wave sums
// its 32-node chunks in MFMA accumulators and stores one partial
// (W1 16x64, W2 64x64, W3 64x16 = 6144 floats) to part[wave].
constexpr int kWsz = kFt * kHid + kHid * kHid + kHid * kOut;
__global__ __launch_bounds__(kNT) void k_net_bwd(int64_t n, NetW W, const __bf16* __restrict__ X,
                                                 const __bf16* __restrict__ H1, const __bf16* __restrict__ H2,
                                                 const float* __restrict__ dY, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 xs[kNT / 64][32][kFt];
  __shared__ __attribute__((aligned(16))) __bf16 h1s[kNT / 64][32][kHid];
  __shared__ __attribute__((aligned(16))) __bf16 h2s[kNT / 64][32][kHid];
  __shared__ __attribute__((aligned(16))) __bf16 dys[kNT / 64][32][kOut];
  __shared__ __attribute__((aligned(16))) __bf16 d1s[kNT / 64][32][kHid];
  __shared__ __attribute__((aligned(16))) __bf16 d2s[kNT / 64][32][kHid];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
  // W3^T (B[k=j][n=h] = W3[h][j], K = 16) and W2^T (B[k=h2][n=h1] = W2[h1][h2])
  bf16x8 w3t[4], w2t[2][4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) w3t[nt] = load_bt(W.W3, kOut, 0, nt, lane);
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) w2t[kb][nt] = load_bt(W.W2, kHid, kb, nt, lane);
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  floatx4 g1[4], g2[4][4], g3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    g1[a] = zero;
    g3[a] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) g2[a][b] = zero;
  }
  const int64_t nchunk = (n + 31) / 32;
  const int64_t per_pass = (int64_t)gridDim.x * (kNT / 64);
  const int64_t npass = (nchunk + per_pass - 1) / per_pass;
  for (int64_t ps = 0; ps < npass; ++ps) {
    const int64_t ch = ps * per_pass + (int64_t)blockIdx.x * (kNT / 64) + wave;
    const int64_t base = ch * 32;
    // stage X, H1, H2, dY of the chunk (zeros past the end)
    for (int idx = lane; idx < 32 * kHid; idx += 64) {
      const int rr = idx / kHid, cc = idx % kHid;
      const int64_t node = base + rr;
      const bool ok = ch < nchunk && node < n;
      h1s[wave][rr][cc] = ok ? H1[node * kHid + cc] : (__bf16)0.0f;
      h2s[wave][rr][cc] = ok ? H2[node * kHid + cc] : (__bf16)0.0f;
    }
    for (int idx = lane; idx < 32 * kFt; idx += 64) {
      const int rr = idx / kFt, cc = idx % kFt;
      const int64_t node = base + rr;
      const bool ok = ch < nchunk && node < n;
      xs[wave][rr][cc] = ok ? X[node * kFt + cc] : (__bf16)0.0f;
      dys[wave][rr][cc] = (ok && cc < 3) ? (__bf16)dY[3 * node + cc] : (__bf16)0.0f;
    }
    __syncthreads();
    // dH2 = (dY W3^T) * relu'(H2), per 16-node sub-tile
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = (8 * g + e < kOut) ? dys[wave][s * 16 + j][8 * g + e] : (__bf16)0.0f;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const floatx4 c = mfma(a, w3t[nt], zero);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = s * 16 + 4 * g + v, cc = nt * 16 + j;
          d2s[wave][rr][cc] = (float)h2s[wave][rr][cc] > 0.0f ? (__bf16)c[v] : (__bf16)0.0f;
        }
      }
    }
    __syncthreads();
    // dH1 = (dH2 W2^T) * relu'(H1)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      bf16x8 a[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) a[kb] = *reinterpret_cast<const bf16x8*>(&d2s[wave][s * 16 + j][kb * 32 + 8 * g]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        floatx4 c = mfma(a[0], w2t[0][nt], zero);
        c = mfma(a[1], w2t[1][nt], c);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = s * 16 + 4 * g + v, cc = nt * 16 + j;
          d1s[wave][rr][cc] = (float)h1s[wave][rr][cc] > 0.0f ? (__bf16)c[v] : (__bf16)0.0f;
        }
      }
    }
    __syncthreads();
    // weight gradients, the 32 nodes of the chunk as the MFMA K dimension
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      bf16x8 ah2, ah1, bdy;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        ah2[e] = h2s[wave][8 * g + e][mt * 16 + j];
        ah1[e] = h1s[wave][8 * g + e][mt * 16 + j];
        bdy[e] = dys[wave][8 * g + e][j];
      }
      g3[mt] = mfma(ah2, bdy, g3[mt]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x8 bd2;
#pragma unroll
        for (int e = 0; e < 8; ++e) bd2[e] = d2s[wave][8 * g + e][nt * 16 + j];
        g2[mt][nt] = mfma(ah1, bd2, g2[mt][nt]);
      }
    }
    {
      bf16x8 ax;
#pragma unroll
      for (int e = 0; e < 8; ++e) ax[e] = xs[wave][8 * g + e][j];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x8 bd1;
#pragma unroll
        for (int e = 0; e < 8; ++e) bd1[e] = d1s[wave][8 * g + e][nt * 16 + j];
        g1[nt] = mfma(ax, bd1, g1[nt]);
      }
    }
    __syncthreads();
  }
  // this wave's partial: D tile (rows 4g+v, col j)
  float* out = part + (int64_t)(blockIdx.x * (kNT / 64) + wave) * kWsz;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rr = 4 * g + v;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) out[rr * kHid + nt * 16 + j] = g1[nt][v];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) out[kFt * kHid + (mt * 16 + rr) * kHid + nt * 16 + j] = g2[mt][nt][v];
      out[kFt * kHid + kHid * kHid + (mt * 16 + rr) * kOut + j] = g3[mt][v];
    }
  }
}

// dW = sum of the wave partials in a fixed order
__global__ void k_net_reduce(int nparts, const float* __restrict__ part, float* __restrict__ dW) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= kWsz) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * kWsz + w];
  dW[w] = s;
}

// Adam on the FP32 master weights; BF16 copies for the MFMA kernels.
// Columns 3..15 of W3 stay zero.
__global__ void k_net_adam(float* __restrict__ Wm, float* __restrict__ M, float* __restrict__ V,
                           const float* __restrict__ dW, __bf16* __restrict__ Wb, float lr, float b1, float b2,
                           float eps, float c1, float c2, float scale) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= kWsz) return;
  const bool frozen = w >= kFt * kHid + kHid * kHid && ((w - kFt * kHid - kHid * kHid) % kOut) >= 3;
  if (!frozen) {
    const float gr = dW[w] * scale;
    const float m = b1 * M[w] + (1.0f - b1) * gr;
    const float v = b2 * V[w] + (1.0f - b2) * gr * gr;
    M[w] = m;
    V[w] = v;
    Wm[w] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  } else {
    Wm[w] = 0.0f;
  }
  Wb[w] = (__bf16)Wm[w];
}

__global__ void k_net_to_bf16(int n, const float* __restrict__ a, __bf16* __restrict__ b) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < n) b[w] = (__bf16)a[w];
}

__global__ void k_sum_parts(int nparts, int K, const double* __restrict__ part, double* __restrict__ out) {
  if ((int)threadIdx.x < K) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(int64_t)threadIdx.x * kMaxPartials + p];
    out[threadIdx.x] = s;
  }
}
// m_i = tanh((ln tr D_i - mean)/2): pass 1 ln tr D and partial sums, pass 2 normalise
__global__ __launch_bounds__(kNT) void k_mat_log(int64_t n, const double* __restrict__ D, float* __restrict__ m,
                                                 double* part) {
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT) {
    const double tr = fabs(D[9 * i]) + fabs(D[9 * i + 4]) + fabs(D[9 * i + 8]);
    const double l = log(tr > 0.0 ? tr : 1e-300);
    m[i] = (float)l;
    acc[0] += l;
  }
  block_sum<1>(acc, part, blockIdx.x);
}
__global__ void k_mat_norm(int64_t n, float* __restrict__ m, const double* __restrict__ sum) {
  const float mean = (float)(sum[0] / (double)n);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m[i] = tanhf(0.5f * (m[i] - mean));
}

int grid_for(int64_t n, int per) {
  int64_t g = (n + per - 1) / per;
  return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

// ---- learned coarse modes: per-aggregate linear autoencoders -----------------------------
// Samples: the solver's own slow errors e_s = A^-1 r_s, stored node-major with
// the sample index inner, S[(3 i + d) * 16 + s] (FP32, up to 16 samples; unused
// slots zero).  Per aggregate a linear autoencoder x -> W^T x -> W W^T x with
// k columns in W is fitted to the samples' part beyond the rigid-body modes;
// its optimum is the span of the top-k eigenvectors of the aggregate's Gram
// matrix G = R^T R (16 x 16), which a BF16 MFMA pass accumulates in FP32.  The
// decoder columns W = R V diag(lambda)^-1/2 become the aggregate's extra
// coarse modes.

// v = a seeded pseudo-random vector, uniform in [-1, 1) (splitmix64 of seed and index)
__global__ void k_modes_random(int64_t n3, uint64_t seed, double* __restrict__ v) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)k * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[k] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// S[:, s] = v - x (FP32)
__global__ void k_modes_sample(int64_t n3, const double* __restrict__ v, const double* __restrict__ x,
                               float* __restrict__ S, int s) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x)
    S[16 * k + s] = (float)(v[k] - x[k]);
}

// per (aggregate a = blockIdx.x, sample s = blockIdx.y): Z_a^T e_s over the six
// rigid-body modes (FP64, fixed order); s == 0 also writes Z_a^T Z_a (36)
__global__ __launch_bounds__(kNT) void k_modes_rigid(const int32_t* __restrict__ aptr,
                                                      const int32_t* __restrict__ anodes,
                                                      const float* __restrict__ dv, const float* __restrict__ S,
                                                      double* __restrict__ rhs, double* __restrict__ ztz) {
  const int a = blockIdx.x, s = blockIdx.y;
  double acc[6 + 21];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.0;
  for (int t = aptr[a] + threadIdx.x; t < aptr[a + 1]; t += kNT) {
    const int64_t i = anodes[t];
    const double ex = S[16 * (3 * i) + s], ey = S[16 * (3 * i + 1) + s], ez = S[16 * (3 * i + 2) + s];
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    // Z_i columns: e_x, e_y, e_z, (0,-dz,dy), (dz,0,-dx), (-dy,dx,0)
    const double z[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, -dz, dy}, {dz, 0, -dx}, {-dy, dx, 0}};
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += z[q][0] * ex + z[q][1] * ey + z[q][2] * ez;
    if (s == 0) {
      int k = 6;
#pragma unroll
      for (int q = 0; q < 6; ++q)
#pragma unroll
        for (int u = q; u < 6; ++u) acc[k++] += z[q][0] * z[u][0] + z[q][1] * z[u][1] + z[q][2] * z[u][2];
    }
  }
  __shared__ double red[27][kNT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kNT / 64; ++w) v += red[threadIdx.x][w];
    if (threadIdx.x < 6) {
      rhs[((int64_t)a * 16 + s) * 6 + threadIdx.x] = v;
    } else if (s == 0) {
      // unpack the upper triangle into the full 6 x 6
      int q = 0, u = 0, k = 6;
      while (k < (int)threadIdx.x) {
        if (++u == 6) u = ++q;
        ++k;
      }
      ztz[(int64_t)a * 36 + 6 * q + u] = v;
      ztz[(int64_t)a * 36 + 6 * u + q] = v;
    }
  }
}

// R = S - Z_rigid c: the samples' part beyond the rigid-body modes of each
// node's aggregate (c[(a * 16 + s) * 6 + q], zero for unused samples)
__global__ void k_modes_resid(int32_t n, const int32_t* __restrict__ agg, const float* __restrict__ dv,
                              const double* __restrict__ c, float* __restrict__ S) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = agg[i];
    const double dx = dv[3 * i], dy = dv[3 * i + 1], dz = dv[3 * i + 2];
    for (int s = 0; s < 16; ++s) {
      const double* cs = c + ((int64_t)a * 16 + s) * 6;
      const double u0 = cs[0] + (cs[4] * dz - cs[5] * dy);
      const double u1 = cs[1] + (cs[5] * dx - cs[3] * dz);
      const double u2 = cs[2] + (cs[3] * dy - cs[4] * dx);
      S[16 * (3 * i) + s] = (float)((double)S[16 * (3 * i) + s] - u0);
      S[16 * (3 * i + 1) + s] = (float)((double)S[16 * (3 * i + 1) + s] - u1);
      S[16 * (3 * i + 2) + s] = (float)((double)S[16 * (3 * i + 2) + s] - u2);
    }
  }
}

// G_a = R_a^T R_a (16 x 16) with v_mfma_f32_16x16x32_bf16: the A operand of a
// 32-dof chunk, A[i = l%16][k = 8(l/16)+e] = R[k][i], equals its B operand
// B[k][j = l%16] = R[k][j] lane for lane, so one load feeds both.  Four waves
// take every fourth chunk; their tiles are summed in a fixed order.
__global__ __launch_bounds__(kNT) void k_modes_gram(const int32_t* __restrict__ aptr,
                                                     const int32_t* __restrict__ anodes, const float* __restrict__ S,
                                                     float* __restrict__ G) {
  const int a = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t0 = aptr[a], nd = 3 * (aptr[a + 1] - t0);
  const int g = lane >> 4, col = lane & 15;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int kb = wave; kb * 32 < nd; kb += kNT / 64) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kb * 32 + 8 * g + e;
      float x = 0.f;
      if (k < nd) {
        const int64_t i = anodes[t0 + k / 3];
        x = S[16 * (3 * i + k % 3) + col];
      }
      v[e] = (__bf16)x;
    }
    acc = mfma(v, v, acc);
  }
  __shared__ float red[kNT / 64][16][16];
#pragma unroll
  for (int vv = 0; vv < 4; ++vv) red[wave][4 * g + vv][col] = acc[vv];
  __syncthreads();
  if (threadIdx.x < 256) {
    const int i = threadIdx.x / 16, j = threadIdx.x % 16;
    float sum = 0.f;
#pragma unroll
    for (int w = 0; w < kNT / 64; ++w) sum += red[w][i][j];
    G[(int64_t)a * 256 + 16 * i + j] = sum;
  }
}

// xm[18 i + 3 j + d] = sum_s R[i][d][s] V[a][s][j]  (the fitted decoder, mode j of node i)
__global__ void k_modes_make(int32_t n, const int32_t* __restrict__ agg, const float* __restrict__ S,
                             const float* __restrict__ V, float* __restrict__ xm) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float* va = V + (int64_t)agg[i] * 96;
    for (int d = 0; d < 3; ++d) {
      float o[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      for (int s = 0; s < 16; ++s) {
        const float r = S[16 * (3 * i + d) + s];
#pragma unroll
        for (int j = 0; j < 6; ++j) o[j] += r * va[6 * s + j];
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) xm[18 * i + 3 * j + d] = o[j];
    }
  }
}

}  // namespace

// ---- launchers ---------------------------------------------------------------------
int net_weight_count() { return kWsz; }

void net_static_features(const Sell<double>& A, int64_t n, const float* mat, __bf16* X, hipStream_t s) {
  k_static_features<<<grid_for(n, kNT), kNT, 0, s>>>(A, n, mat, X);
}

void net_forward(int64_t n, const __bf16* Wb, const __bf16* X, __bf16* H1, __bf16* H2, const float* z0, float* z,
                 hipStream_t s) {
  const NetW W{Wb, Wb + kFt * kHid, Wb + kFt * kHid + kHid * kHid};
  k_net_fwd<<<grid_for((n + 15) / 16, kNT / 64), kNT, 0, s>>>(n, W, X, H1, H2, z0, z);
}

// gradients of the weights for upstream dY (n x 3) into dW (6144 floats);
// part needs net_bwd_parts() * 6144 floats
int net_bwd_parts(int64_t n) { return grid_for((n + 31) / 32, kNT / 64) * (kNT / 64); }
void net_backward(int64_t n, const __bf16* Wb, const __bf16* X, const __bf16* H1, const __bf16* H2, const float* dY,
                  float* part, float* dW, hipStream_t s) {
  const NetW W{Wb, Wb + kFt * kHid, Wb + kFt * kHid + kHid * kHid};
  const int g = grid_for((n + 31) / 32, kNT / 64);
  k_net_bwd<<<g, kNT, 0, s>>>(n, W, X, H1, H2, dY, part);
  k_net_reduce<<<(kWsz + 255) / 256, 256, 0, s>>>(g * (kNT / 64), part, dW);
}

void net_adam(float* Wm, float* M, float* V, const float* dW, __bf16* Wb, float lr, float b1, float b2, float eps,
              int step, float scale, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  k_net_adam<<<(kWsz + 255) / 256, 256, 0, s>>>(Wm, M, V, dW, Wb, lr, b1, b2, eps, c1, c2, scale);
}

void net_to_bf16(const float* Wm, __bf16* Wb, hipStream_t s) {
  k_net_to_bf16<<<(kWsz + 255) / 256, 256, 0, s>>>(kWsz, Wm, Wb);
}

// logistic loss against labels t (n), gradient into dY (3n); out[0] = summed
// loss, out[1] = correctly classified nodes
void net_bce(int64_t n, const float* Y, const float* t, float* dY, double* part, double* out, hipStream_t s) {
  const int g = grid_for(n, kNT);
  k_net_bce<<<g, kNT, 0, s>>>(n, Y, t, dY, part);
  k_sum_parts<<<1, 64, 0, s>>>(g, 2, part, out);
}
void net_mask(int64_t n, const float* Y, uint8_t* mask, hipStream_t s) {
  k_net_mask<<<grid_for(n, 256), 256, 0, s>>>(n, Y, mask);
}

void modes_random(int64_t n3, uint64_t seed, double* v, hipStream_t st) {
  k_modes_random<<<grid_for(n3, 256), 256, 0, st>>>(n3, seed, v);
}
void modes_sample(int64_t n3, const double* v, const double* x, float* S, int s, hipStream_t st) {
  k_modes_sample<<<grid_for(n3, 256), 256, 0, st>>>(n3, v, x, S, s);
}
void modes_rigid(int32_t n_agg, int ns, const int32_t* aptr, const int32_t* anodes, const float* dv, const float* S,
                 double* rhs, double* ztz, hipStream_t st) {
  k_modes_rigid<<<dim3((unsigned)n_agg, (unsigned)ns), kNT, 0, st>>>(aptr, anodes, dv, S, rhs, ztz);
}
void modes_resid(int32_t n, const int32_t* agg, const float* dv, const double* c, float* S, hipStream_t st) {
  k_modes_resid<<<grid_for(n, 256), 256, 0, st>>>(n, agg, dv, c, S);
}
void modes_gram(int32_t n_agg, const int32_t* aptr, const int32_t* anodes, const float* S, float* G, hipStream_t st) {
  k_modes_gram<<<(unsigned)n_agg, kNT, 0, st>>>(aptr, anodes, S, G);
}
void modes_make(int32_t n, const int32_t* agg, const float* S, const float* V, float* xm, hipStream_t st) {
  k_modes_make<<<grid_for(n, 256), 256, 0, st>>>(n, agg, S, V, xm);
}

void net_material(int64_t n, const double* D, float* m, double* part, double* sum, hipStream_t s) {
  const int g = grid_for(n, kNT);
  k_mat_log<<<g, kNT, 0, s>>>(n, D, m, part);
  k_sum_parts<<<1, 64, 0, s>>>(g, 1, part, sum);
  k_mat_norm<<<grid_for(n, 256), 256, 0, s>>>(n, m, sum);
}

}  // namespace hecmw33
