// The relative residual sqrt(|b - A x|^2 / |b|^2) of a system whose arrays are
// all in device memory, for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// hecmw33_rel_resid_L2 (hecmw_rel_resid_L2_33, hecmw_solver_las_33.F90:406-435)
// takes host arrays and keeps a device copy of the system and its reduction
// partials in the library's context.  This variant reads the caller's device
// copies of the hecmwST_matrix arrays directly and keeps its partial sums in a
// fixed device array, `scratch`, so a call allocates nothing:
//   k_resid_parts : one workgroup per kRowsPerPart rows; row i computes
//                   r_i = b_i - A_i x in the reference's order (D, then the L
//                   items, then the U items, hecmw_solver_las_33.F90:262-300)
//                   and the workgroup writes its partials of |r|^2 and |b|^2
//   k_resid_tail  : one workgroup sums the partials in a fixed order and
//                   writes sqrt(rr / bb) (bb = 0 counts as 1, as the
//                   reference) to the caller's device double
// Every partial a call reads is written by that call, so nothing has to start
// at zero; calls that overlap on one GPU share its scratch array and race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/hecmw33.h"
#include "device.h"

namespace hecmw33 {
namespace {

constexpr int kPartThreads = 256;
constexpr int kRowsPerPart = kPartThreads;  // a thread per row
constexpr int kTailThreads = 1024;
constexpr int kMaxDevices = 64;

// the partial sums of |r|^2 (even slots) and |b|^2 (odd slots), one pair per workgroup
static __device__ __align__(256) double scratch[2 * HECMW33_NOALLOC_MAX_PARTS];

// per device: scratch's device address (hipGetSymbolAddress), looked up once
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

__global__ __launch_bounds__(kPartThreads) void k_resid_parts(int32_t N, const double* __restrict__ D,
                                                                const double* __restrict__ AL,
                                                                const double* __restrict__ AU,
                                                                const int32_t* __restrict__ indexL,
                                                                const int32_t* __restrict__ itemL,
                                                                const int32_t* __restrict__ indexU,
                                                                const int32_t* __restrict__ itemU,
                                                                const double* __restrict__ b,
                                                                const double* __restrict__ x, double* part) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kRowsPerPart + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (i < N) {
    const double* d = D + 9 * i;
    double x1 = x[3 * i], x2 = x[3 * i + 1], x3 = x[3 * i + 2];
    double y1 = d[0] * x1 + d[1] * x2 + d[2] * x3;
    double y2 = d[3] * x1 + d[4] * x2 + d[5] * x3;
    double y3 = d[6] * x1 + d[7] * x2 + d[8] * x3;
    for (int32_t j = indexL[i]; j < indexL[i + 1]; ++j) {
      const int64_t c = itemL[j] - 1;
      const double* a = AL + 9 * (int64_t)j;
      x1 = x[3 * c];
      x2 = x[3 * c + 1];
      x3 = x[3 * c + 2];
      y1 = y1 + a[0] * x1 + a[1] * x2 + a[2] * x3;
      y2 = y2 + a[3] * x1 + a[4] * x2 + a[5] * x3;
      y3 = y3 + a[6] * x1 + a[7] * x2 + a[8] * x3;
    }
    for (int32_t j = indexU[i]; j < indexU[i + 1]; ++j) {
      const int64_t c = itemU[j] - 1;
      const double* a = AU + 9 * (int64_t)j;
      x1 = x[3 * c];
      x2 = x[3 * c + 1];
      x3 = x[3 * c + 2];
      y1 = y1 + a[0] * x1 + a[1] * x2 + a[2] * x3;
      y2 = y2 + a[3] * x1 + a[4] * x2 + a[5] * x3;
      y3 = y3 + a[6] * x1 + a[7] * x2 + a[8] * x3;
    }
    const double b1 = b[3 * i], b2 = b[3 * i + 1], b3 = b[3 * i + 2];
    const double r1 = b1 - y1, r2 = b2 - y2, r3 = b3 - y3;
    rr = r1 * r1 + r2 * r2 + r3 * r3;
    bb = b1 * b1 + b2 * b2 + b3 * b3;
  }
  __shared__ double red[2][kPartThreads];
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = bb;
  __syncthreads();
  for (int w = kPartThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ __launch_bounds__(kTailThreads) void k_resid_tail(int32_t nparts, const double* __restrict__ part,
                                                               double* out) {
  double rr = 0.0, bb = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kTailThreads) {
    rr += part[2 * k];
    bb += part[2 * k + 1];
  }
  __shared__ double red[2][kTailThreads];
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = bb;
  __syncthreads();
  for (int w = kTailThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double b2 = red[1][0] == 0.0 ? 1.0 : red[1][0];
    *out = sqrt(red[0][0] / b2);
  }
}

// scratch's address on the current device, looked up on the first call there
int32_t scratch_ptr(double** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_last_error(std::string("hecmw33_rel_resid_L2_device_noalloc: no current HIP device: ") +
                   hipGetErrorString(e));
    return HECMW33_ERROR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_last_error(std::string("hecmw33_rel_resid_L2_device_noalloc: hipGetSymbolAddress(scratch): ") +
                     hipGetErrorString(e));
      return HECMW33_ERROR_DEVICE;
    }
    g_scratch[dev] = ptr;
  }
  *p = static_cast<double*>(g_scratch[dev]);
  return 0;
}

}  // namespace
}  // namespace hecmw33

using namespace hecmw33;

extern "C" int32_t hecmw33_rel_resid_L2_device_noalloc(const hecmw33_sys* A, const double* B, const double* X,
                                                       double* resid, void* stream) {
  if (!A || !B || !X || !resid || A->N < 1 || A->NP < A->N || A->NPL < 0 || A->NPU < 0 || !A->D || !A->indexL ||
      !A->indexU || (A->NPL > 0 && (!A->AL || !A->itemL)) || (A->NPU > 0 && (!A->AU || !A->itemU))) {
    set_last_error("hecmw33_rel_resid_L2_device_noalloc: invalid system or vector pointers");
    return HECMW33_ERROR_ARGS;
  }
  const int64_t nparts = ((int64_t)A->N + kRowsPerPart - 1) / kRowsPerPart;
  if (nparts > HECMW33_NOALLOC_MAX_PARTS) {
    set_last_error("hecmw33_rel_resid_L2_device_noalloc: N = " + std::to_string(A->N) + " exceeds the " +
                   std::to_string((int64_t)HECMW33_NOALLOC_MAX_PARTS * kRowsPerPart) +
                   " rows its fixed scratch array holds partial sums for");
    return HECMW33_ERROR_ARGS;
  }
  double* part = nullptr;
  if (int32_t e = scratch_ptr(&part)) return e;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_resid_parts<<<(int)nparts, kPartThreads, 0, s>>>(A->N, A->D, A->AL, A->AU, A->indexL, A->itemL, A->indexU,
                                                     A->itemU, B, X, part);
  k_resid_tail<<<1, kTailThreads, 0, s>>>((int32_t)nparts, part, resid);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_last_error(std::string("hecmw33_rel_resid_L2_device_noalloc: launch: ") + hipGetErrorString(e));
    return HECMW33_ERROR_DEVICE;
  }
  return 0;
}
