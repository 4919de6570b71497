// Host runtime of libhecmw33: device context, storage build, preconditioner
// setup and the Krylov drivers behind the C ABI of include/hecmw33.h.
//
// Layout of one context in HBM (FP64 path):
//   A      : full matrix, natural row order, SELL-64, D block first per row
//            then L then U in the caller's storage order (the reference CPU
//            SpMV order, hecmw_solver_las_33.F90:281-308)
//   raw    : the caller's D/AL/AU uploaded as-is (source of every value fill)
//   L', U' : SSOR factors in multicolour order, SELL-16 (four lanes per row)
//            or SELL-64 (a lane per row) slices never crossing a colour; the
//            sweeps keep their vector in slot order and the columns hold slot
//            ids, so rows that are consecutive in colour order gather
//            neighbours that are consecutive too (the reference maps items
//            back to old ids, hecmw_precond_SSOR_33.F90:151-155, and sweeps
//            ZP in natural order; the arithmetic is the same)
//   ALU    : 3x3 LU of the diagonal blocks, 9-plane slice layout
//   vectors: b, x, r, z, p, q (+ BiCGSTAB's) of 3*NP doubles
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hecmw33.h"
#include "device.h"
#include "ordering.h"

struct hecmw33_net;

namespace hecmw33 {

static thread_local std::string g_last_error;

static void set_error(const std::string& s) { g_last_error = s; }
void set_last_error(const std::string& s) { g_last_error = s; }

#define HIPC(expr)                                                                          \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                         \
      return HECMW33_ERROR_DEVICE;                                                          \
    }                                                                                       \
  } while (0)

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Iarray / Rarray slots (1-based in the reference, hecmw_matrix_misc.F90:107-150)
enum {
  I_ITER = 1, I_METHOD = 2, I_PRECOND = 3, I_ITERPREMAX = 5, I_SCALING = 7, I_METHOD2 = 8,
  I_ITERLOG = 21, I_TIMELOG = 22, I_NCOLOR_IN = 34, I_MAXRECYCLE = 35, I_SOLVER_OPT = 41,
  I_MPC_METHOD = 13, I_FLAG_CONVERGED = 81, I_FLAG_DIVERGED = 82, I_FLAG_MPCMATVEC = 83, I_NRECYCLE = 96,
  I_FLAG_NUMFACT = 97, I_FLAG_SYMBFACT = 98,
  R_RESID = 1, R_SIGMA_DIAG = 2, R_SOLVER_OPT = 41, R_SOLVER_OPT2 = 42,
};
// SOLVER_OPT(1) (Iarray(41), unused by the reference's CG/BiCGSTAB with
// PRECOND 1-3) = 1 selects the flexible CG with FP32 operators;
// Rarray(41) in (0,1) overrides its reliable-update factor.
constexpr int32_t kOptMixedFcg = 1;
// = 2: the same flexible CG with the two-level step (Iarray 43, 1 when not
// set) on the aggregation an attached network refines: the nodes it marks as
// moving rigidly with a stiff piece leave their cell, one aggregate per
// connected piece (hecmw33_ctx_set_net / hecmw33_set_net)
constexpr int32_t kOptLearnedFcg = 2;
constexpr double kFcgDelta = 1.0e-2;
// a piece the network marks becomes an aggregate from this many nodes on
// (smaller ones stay with their cell); the largest pieces are kept first
// within kCoarseAggMax
constexpr int32_t kPieceMinNodes = 32;
// SOLVER_OPT(2) (Iarray 42) for the flexible CG's q = A p: 1 = FP32-rounded
// values, 2 = FP64 values with FP32 vectors (FP64 accumulation), 0 = FP32
// values with the two-level correction, else chosen from the stiffness
// contrast of the diagonal blocks (FP64 values above kContrastFp64Spmv)
constexpr int I_SOLVER_OPT2 = 42;
constexpr double kContrastFp64Spmv = 1.0e4;
// SOLVER_OPT(3) (Iarray 43) for the flexible CG: > 0 adds the two-level
// correction z += Z E^-1 Z^T r (rigid-body modes on geometric aggregates of
// the nodes, which hecmw33_set_coords supplies); 1 = kCoarseAggDefault
// aggregates, larger values the aggregate count to aim at
constexpr int I_SOLVER_OPT3 = 43;
constexpr int32_t kCoarseAggDefault = 1000;
// SOLVER_OPT(4) (Iarray 44) picks the form of the two-level correction: 0
// (default) the symmetric step z1 = Q r, z2 = z1 + M^-1 (r - A z1),
// z = z2 + Q (r - A z2) with Q = Z E^-1 Z^T (A z1 and the coarse part of A z
// from A Z kept per node; q = A p carried as A z + beta q, so an iteration
// still has one SpMV); 1 the additive z = M^-1 r + Q r; 2 the deflation form
// y = M^-1 r, z = y + Q (r - A y): one coarse solve per iteration instead of
// two (the A-DEF2 preconditioner of Tang, Nabben, Vuik and Erlangga, J. Sci.
// Comput. 39 (2009) 340-370, whose spectrum is the symmetric step's; it is
// not symmetric, which the flexible CG's explicit beta allows).
constexpr int I_SOLVER_OPT4 = 44;
// SOLVER_OPT(7) (Iarray 47): the L'/U' values the flexible CG's FP32 SSOR
// sweeps read: 0 (default) BF16 copies (22 B per block instead of 40, the
// sweeps' dominant stream; D^-1 stays FP32, and since the rounding is
// entrywise M stays symmetric positive definite), 1 FP32
constexpr int I_SOLVER_OPT7 = 47;
// the dense coarse inverse is n_c^2 entries (n_c = 6 per aggregate): at most
// 2048 aggregates (n_c = 12288, 0.6 GB in FP32)
constexpr int32_t kCoarseAggMax = 2048;
// and at least this many nodes per aggregate on average
constexpr int32_t kCoarseMinNodes = 32;
static inline int32_t& IA(int32_t* a, int k) { return a[k - 1]; }
static inline int32_t IA(const int32_t* a, int k) { return a[k - 1]; }
static inline double RA(const double* a, int k) { return a[k - 1]; }

// ---------------------------------------------------------------------------
// device memory helper
struct DMem {
  void* p = nullptr;
  size_t bytes = 0;
  int32_t ensure(size_t n) {
    if (n <= bytes && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) n = 8;
    HIPC(hipMalloc(&p, n));
    bytes = n;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

struct SellHost {
  int32_t nslices = 0, nrows = 0, rows = kRowsA;
  std::vector<int32_t> ptr, len, col;
  std::vector<int64_t> src;
};

struct SellDev {
  // val: FP64 values; val32: their FP32 image; val16: BF16 copies of the FP32
  // image (the L'/U' values the flexible CG's SSOR sweeps read by default)
  DMem ptr, len, col, src, val, val32, val16;
  Sell<double> view;
  Sell<float> view32;
};

// Build SELL-C storage (C = kRows) for rows given slot by slot
// (lane_rows[s*C+r] = row key or -1).  count(key) = blocks in that row;
// fill(key, emit) calls emit(k, col, src) for k = 0..count-1 in storage order.
template <int kRows, class Count, class Fill>
static void build_sell(const std::vector<int32_t>& lane_rows, int32_t nrows, Count count, Fill fill, SellHost& h) {
  h.rows = kRows;
  const int32_t ns = (int32_t)(lane_rows.size() / kRows);
  h.nslices = ns;
  h.nrows = nrows;
  h.len.assign(lane_rows.size(), 0);
  h.ptr.assign(ns + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)lane_rows.size(); ++t) h.len[t] = lane_rows[t] >= 0 ? count(lane_rows[t]) : 0;
  for (int32_t s = 0; s < ns; ++s) {
    int32_t w = 0;
    for (int l = 0; l < kRows; ++l) w = std::max(w, h.len[(int64_t)s * kRows + l]);
    h.ptr[s + 1] = h.ptr[s] + w;
  }
  const int64_t nslots = (int64_t)h.ptr[ns] * kRows;
  h.col.assign(nslots, 0);
  h.src.assign(nslots, -1);
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t s = 0; s < ns; ++s)
    for (int l = 0; l < kRows; ++l) {
      const int32_t key = lane_rows[(int64_t)s * kRows + l];
      if (key < 0) continue;
      const int64_t base = (int64_t)h.ptr[s] * kRows + l;
      fill(key, [&](int32_t k, int32_t c, int64_t sidx) {
        h.col[base + (int64_t)k * kRows] = c;
        h.src[base + (int64_t)k * kRows] = sidx;
      });
    }
}

static int32_t upload_sell(const SellHost& h, SellDev& d, hipStream_t s) {
  const int64_t nslots = (int64_t)h.ptr[h.nslices] * h.rows;
  if (d.ptr.ensure(sizeof(int32_t) * (h.nslices + 1)) || d.len.ensure(sizeof(int32_t) * h.len.size()) ||
      d.col.ensure(sizeof(int32_t) * std::max<int64_t>(nslots, 1)) ||
      d.src.ensure(sizeof(int64_t) * std::max<int64_t>(nslots, 1)) ||
      d.val.ensure(sizeof(double) * 9 * std::max<int64_t>(nslots, 1)))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(d.ptr.p, h.ptr.data(), sizeof(int32_t) * (h.nslices + 1), hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.len.p, h.len.data(), sizeof(int32_t) * h.len.size(), hipMemcpyHostToDevice, s));
  if (nslots > 0) {
    HIPC(hipMemcpyAsync(d.col.p, h.col.data(), sizeof(int32_t) * nslots, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.src.p, h.src.data(), sizeof(int64_t) * nslots, hipMemcpyHostToDevice, s));
  }
  d.view.nslices = h.nslices;
  d.view.nrows = h.nrows;
  d.view.rows = h.rows;
  d.view32.rows = h.rows;
  d.view.nslots = nslots;
  d.view.ptr = d.ptr.as<int32_t>();
  d.view.len = d.len.as<int32_t>();
  d.view.col = d.col.as<int32_t>();
  d.view.val = d.val.as<double>();
  d.view32.nslices = h.nslices;
  d.view32.nrows = h.nrows;
  d.view32.nslots = nslots;
  d.view32.ptr = d.view.ptr;
  d.view32.len = d.view.len;
  d.view32.col = d.view.col;
  d.view32.val = nullptr;
  HIPC(hipStreamSynchronize(s));  // host vectors may be released by the caller
  return 0;
}

// ---------------------------------------------------------------------------
struct Context {
  hipStream_t stream = nullptr;
  // structure key
  int32_t N = -1, NP = -1, NPL = -1, NPU = -1;
  const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
  bool have_struct = false;
  // raw arrays and full matrix
  DMem dD, dAL, dAU;
  SellDev A;
  // host copy of the structure for preconditioner builds
  std::vector<int32_t> h_indexL, h_indexU, h_itemL, h_itemU;
  // SSOR
  bool ssor_init = false;
  bool ssor_sym = false;        // ordering and L'/U' structure built
  int32_t ssor_ncolor_in = -1;  // NCOLOR_IN they were built with
  int64_t ssor_rows = 0;        // row slots of the SSOR slices
  int32_t ssor_R = kRowsS;      // rows per SSOR slice (16 or 64)
  int32_t ncolor = 0;
  std::vector<int32_t> perm, iperm, colorindex, color_slice;
  int32_t ssor_last_bwd_empty = 0;  // the last colour's U' rows are all empty
  SellDev L, U;
  DMem row_old, node_slot, alu_ssor, vzs, vzs32;
  bool ssor16 = true;          // the flexible CG's FP32 sweeps read BF16 copies of the L'/U' values
  bool ssor16_valid = false;   // L.val16 / U.val16 hold the current values
  // DIAG
  bool diag_init = false;
  DMem row_diag, alu_diag;
  // device copies of the index arrays and the scale vector (SCALING=1 only)
  bool dindex_valid = false;
  DMem dIndexL, dIndexU, dItemL, dItemU, vscale;
  // FP32 copies for the flexible CG (rounded from the FP64 storage)
  bool a32_valid = false, ssor32_valid = false, diag32_valid = false;
  DMem alu_ssor32, alu_diag32, vr32, vz32, vp32, vq32;
  // the node classifier network (hecmw33_net_*): attached network and its static node features
  hecmw33_net* net = nullptr;
  bool mat_valid = false;
  DMem vmat, scal, vfeat, vnety, vmask;  // vfeat: BF16 static node features, vnety: network outputs (3 per node)
  // learned coarse modes (Iarray 41 = 2): per-aggregate linear autoencoders fitted to the
  // solver's own slow errors; cg_xm holds the fitted modes ([N][6][3] FP32), valid for the
  // current values and aggregation
  bool modes_valid = false;
  bool x_zero = false;  // the next krylov_core starts from x = 0 (its r0 = b needs no product)
  int32_t modes_kept = 0;          // learned modes with a usable eigenvalue, over all aggregates
  int32_t modes_samples = 0, modes_iters = 0;  // error samples and FCG iterations per sample of the last fit
  double modes_harvest_s = 0.0, modes_fit_s = 0.0;
  DMem cg_xm, ml_S, ml_v, ml_bsave, ml_xsave, ml_rhs, ml_ztz, ml_c, ml_G, ml_V;
  // Krylov vectors and state
  DMem vb, vx, vr, vz, vp, vq, vrt, vs, vt, vv, vpt, vst, vzp, vzt;
  DMem st, part, hist;
  KrylovState* h_st = nullptr;  // pinned, 2 slots for polling
  int32_t last_iter = 0;
  int32_t last_nreliable = 0;  // FP64 residual replacements of the last flexible CG solve
  int32_t last_fallback_iter = 0;  // iteration at which the last learned solve dropped its step (0: kept)
  int32_t last_learned_off = 1;    // the last solve ran without the learned step (not learned, or dropped)
  bool last_spmv64 = false;      // the last flexible CG used the FP64 matrix values in q = A p
  // two-level correction: node coordinates (hecMESH%node), aggregation, coarse inverse
  std::vector<double> h_coord;   // 3*NP
  int64_t coord_gen = 0;         // bumped whenever coordinates are set
  bool coord_fresh = false;      // set since the last structure load (kept by the next one)
  bool coarse_sym = false;       // aggregation built for (structure, coordinates, target)
  bool coarse_num = false;       // E^-1 built for the current values
  bool coarse_ok = false;        // E was positive definite
  int32_t coarse_target = 0;
  int64_t coarse_coord_gen = -1;
  bool coarse_use = false;       // this solve applies it
  int32_t last_coarse = 0;       // aggregates used by the last solve (-1: E not positive definite)
  CoarseOp coarse;
  int32_t coarse_npair = 0;       // coupled aggregate pairs (blocks of E)
  int32_t coarse_band_agg = 0;    // max |a - b| over the coupled pairs: E's bandwidth in aggregates
  bool coarse_sym_use = false;    // this solve takes the symmetric two-level step
  bool coarse_defl = false;       // ... in its deflation form (no pre-smoothing coarse solve)
  bool coarse_w_valid = false;    // A Z (per node and touched aggregate) built for the current values
  int32_t last_modes_nm = 0;      // modes per aggregate of the last two-level solve
  int32_t coarse_nm_num = 0;      // modes per aggregate of the E^-1 built (6 or 12)
  int32_t coarse_w_nm = 0;        // modes per aggregate of the A Z built
  int64_t coarse_nw = 0;          // node-aggregate entries of A Z
  DMem cg_agg, cg_aptr, cg_anodes, cg_dv, cg_pa, cg_pb, cg_pnptr, cg_pn, cg_Eblk, cg_E, cg_W, cg_cv, cg_tv, cg_yv, cg_info;
  DMem cg_ncnt, cg_nagg, cg_AZ, cg_sc, cg_aggW, cg_kept, vr1_32, vt32, vw32, vaz32;
  rocblas_handle blas = nullptr;
  // multipoint constraints (hecMESH%mpc) for the MPC-CG product T^T A T
  // (Iarray 83 set, as hecmw_solve_iterative sets it for MPC_METHOD = 2)
  int32_t mpc_n = 0, mpc_N = 0;  // constraints given, for N internal nodes
  std::vector<int32_t> h_mpc_index, h_mpc_item, h_mpc_dof;
  std::vector<double> h_mpc_val;
  bool mpc_built = false;  // device structure built for the constraints and N
  bool mpc_use = false;    // the products of this solve are T^T A T
  MpcOp mpc;
  DMem mpc_slave, mpc_mptr, mpc_mdof, mpc_mval, mpc_tptr, mpc_tslave, mpc_tval, mpc_w1, mpc_w2;
  const MpcOp* mpc_op() const { return mpc_use ? &mpc : nullptr; }
  double contrast = -1.0;        // max/min trace of the diagonal blocks (< 0: not computed for these values)
  // TIMELOG: event pairs around the preconditioner and the SpMV of the first
  // iteration of every enqueued chunk (4 events, 8 for BiCGSTAB); a record
  // that fails surfaces in the checked elapsed-time query after the solve
  std::vector<hipEvent_t> events;
  hipEvent_t event(size_t i) {
    while (events.size() <= i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      events.push_back(e);
    }
    return events[i];
  }

  SsorOp<double> ssor_op() {
    return SsorOp<double>{L.view,  U.view, row_old.as<int32_t>(), alu_ssor.as<double>(), color_slice.data(), ncolor,
                          vzs.as<double>(), node_slot.as<int32_t>(), N, ssor_last_bwd_empty};
  }
  SsorOp<float> ssor_op32(bool v16) {
    SsorOp<float> op{L.view32, U.view32, row_old.as<int32_t>(), alu_ssor32.as<float>(), color_slice.data(),
                     ncolor, vzs32.as<float>(), node_slot.as<int32_t>(), N, ssor_last_bwd_empty};
    if (v16) {
      op.Lv16 = L.val16.as<uint16_t>();
      op.Uv16 = U.val16.as<uint16_t>();
    }
    return op;
  }

  ~Context() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (DMem* m : {&dIndexL, &dIndexU, &dItemL, &dItemU, &vscale, &vzs, &vzs32, &vmat, &scal, &vfeat, &vnety,
                    &vmask})
      m->release();
    for (DMem* m : {&cg_agg, &cg_aptr, &cg_anodes, &cg_dv, &cg_pa, &cg_pb, &cg_pnptr, &cg_pn, &cg_Eblk, &cg_E, &cg_W,
                    &cg_cv, &cg_tv, &cg_yv, &cg_info, &cg_ncnt, &cg_nagg, &cg_AZ, &cg_sc, &cg_aggW, &cg_kept, &vr1_32, &vt32, &vw32, &vaz32})
      m->release();
    if (blas) rocblas_destroy_handle(blas);
    for (DMem* m : {&mpc_slave, &mpc_mptr, &mpc_mdof, &mpc_mval, &mpc_tptr, &mpc_tslave, &mpc_tval, &mpc_w1, &mpc_w2})
      m->release();
    for (DMem* m : {&dD, &dAL, &dAU, &row_old, &node_slot, &alu_ssor, &row_diag, &alu_diag, &alu_ssor32, &alu_diag32, &vr32,
                    &vz32, &vp32, &vq32, &vb, &vx, &vr, &vz, &vp, &vq, &vrt, &vs, &vt, &vv, &vpt, &vst, &vzp, &vzt, &st,
                    &part, &hist})
      m->release();
    for (SellDev* s : {&A, &L, &U})
      for (DMem* m : {&s->ptr, &s->len, &s->col, &s->src, &s->val, &s->val32, &s->val16}) m->release();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (h_st) (void)hipHostFree(h_st);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace hecmw33

// A learned aggregation network: FP32 master weights, BF16 copies for the
// MFMA kernels, Adam moments, and per context the training labels built from
// the error samples added on it (contexts must outlive their samples).
struct hecmw33_net {
  struct Sample {
    hecmw33::Context* c = nullptr;
    std::vector<float> worst;  // per stored edge: max over the samples of its axial strain / the sample median
    int32_t nsamp = 0;
    double tau = 0.1;
    hecmw33::DMem t;           // labels on the device (built at training)
    bool t_valid = false;
  };
  hecmw33::DMem Wm, Wb, M, V, dW, wpart, H1, H2, y, z, dY, dots, dpart, xb;
  hipStream_t stream = nullptr;
  int32_t step = 0;
  uint64_t gen = 1;  // bumped whenever the weights change
  std::vector<Sample*> samples;
  void clear_samples() {
    for (Sample* p : samples) {
      p->t.release();
      delete p;
    }
    samples.clear();
  }
  ~hecmw33_net() {
    if (stream) (void)hipStreamSynchronize(stream);
    clear_samples();
    for (hecmw33::DMem* m : {&Wm, &Wb, &M, &V, &dW, &wpart, &H1, &H2, &y, &z, &dY, &dots, &dpart, &xb}) m->release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace hecmw33 {

static int32_t ctx_init(Context& c) {
  if (c.stream) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    set_error("no HIP device visible");
    return HECMW33_ERROR_DEVICE;
  }
  HIPC(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  HIPC(hipHostMalloc((void**)&c.h_st, 2 * sizeof(KrylovState), hipHostMallocDefault));
  if (c.st.ensure(sizeof(KrylovState)) || c.part.ensure(sizeof(double) * 4 * kMaxPartials))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemsetAsync(c.st.p, 0, sizeof(KrylovState), c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

static int32_t check_sys(const hecmw33_sys* A) {
  if (!A || A->N < 1 || A->NP < A->N || A->NPL < 0 || A->NPU < 0 || !A->D || !A->indexL || !A->indexU ||
      (A->NPL > 0 && (!A->AL || !A->itemL)) || (A->NPU > 0 && (!A->AU || !A->itemU))) {
    set_error("invalid hecmw33_sys");
    return HECMW33_ERROR_ARGS;
  }
  if (A->indexL[0] != 0 || A->indexU[0] != 0 || A->indexL[A->NP] != A->NPL || A->indexU[A->NP] != A->NPU) {
    set_error("index arrays inconsistent with NPL/NPU");
    return HECMW33_ERROR_ARGS;
  }
  return 0;
}

// Structure of A: rebuilt when the symbolic flag asks for it or when the
// sizes or the contents of index/item differ from the kept host copies (a
// new system may sit at the addresses of a freed one, so addresses alone
// do not identify a structure).
static int32_t load_structure(Context& c, const hecmw33_sys* A, bool force) {
  bool same = c.have_struct && c.N == A->N && c.NP == A->NP && c.NPL == A->NPL && c.NPU == A->NPU;
  if (same) {
    const size_t np1 = (size_t)A->NP + 1;
    same = std::equal(A->indexL, A->indexL + np1, c.h_indexL.data()) &&
           std::equal(A->indexU, A->indexU + np1, c.h_indexU.data()) &&
           (A->NPL == 0 || std::equal(A->itemL, A->itemL + A->NPL, c.h_itemL.data())) &&
           (A->NPU == 0 || std::equal(A->itemU, A->itemU + A->NPU, c.h_itemU.data()));
  }
  if (same && !force) return 0;
  // node coordinates given before a different structure was loaded belong
  // to an earlier system; those given since the last load are for this one
  if (!same && !c.coord_fresh) c.h_coord.clear();
  c.coord_fresh = false;
  const int32_t N = A->N, NP = A->NP;
  c.h_indexL.assign(A->indexL, A->indexL + NP + 1);
  c.h_indexU.assign(A->indexU, A->indexU + NP + 1);
  c.h_itemL.assign(A->itemL, A->itemL + A->NPL);
  c.h_itemU.assign(A->itemU, A->itemU + A->NPU);
  for (int64_t q = 0; q < A->NPL; ++q)
    if (A->itemL[q] < 1 || A->itemL[q] > NP) {
      set_error("itemL out of range");
      return HECMW33_ERROR_ARGS;
    }
  for (int64_t q = 0; q < A->NPU; ++q)
    if (A->itemU[q] < 1 || A->itemU[q] > NP) {
      set_error("itemU out of range");
      return HECMW33_ERROR_ARGS;
    }
  const int32_t ns = (N + kRowsA - 1) / kRowsA;
  std::vector<int32_t> lanes((int64_t)ns * kRowsA, -1);
  for (int32_t i = 0; i < N; ++i) lanes[i] = i;
  const int32_t* iL = A->indexL;
  const int32_t* iU = A->indexU;
  const int32_t* tL = A->itemL;
  const int32_t* tU = A->itemU;
  const int64_t offL = N, offU = (int64_t)N + A->NPL;
  SellHost h;
  build_sell<kRowsA>(
      lanes, N, [&](int32_t i) { return 1 + (iL[i + 1] - iL[i]) + (iU[i + 1] - iU[i]); },
      [&](int32_t i, auto emit) {
        int32_t k = 0;
        emit(k++, i, (int64_t)i);
        for (int32_t q = iL[i]; q < iL[i + 1]; ++q) emit(k++, tL[q] - 1, offL + q);
        for (int32_t q = iU[i]; q < iU[i + 1]; ++q) emit(k++, tU[q] - 1, offU + q);
      },
      h);
  if (int32_t e = upload_sell(h, c.A, c.stream)) return e;
  c.N = N;
  c.NP = NP;
  c.NPL = A->NPL;
  c.NPU = A->NPU;
  c.key[0] = A->indexL;
  c.key[1] = A->indexU;
  c.key[2] = A->itemL;
  c.key[3] = A->itemU;
  c.have_struct = true;
  c.ssor_init = false;
  c.ssor_sym = false;
  c.coarse_sym = false;
  c.mpc_built = false;
  c.diag_init = false;
  c.dindex_valid = false;
  const size_t vb = sizeof(double) * 3 * (size_t)NP;
  for (DMem* m : {&c.vb, &c.vx, &c.vr, &c.vz, &c.vp, &c.vq})
    if (m->ensure(vb)) return HECMW33_ERROR_DEVICE;
  return 0;
}

// Values: upload D/AL/AU and refill the SpMV storage.
static int32_t load_values(Context& c, const hecmw33_sys* A) {
  if (c.dD.ensure(sizeof(double) * 9 * (size_t)A->NP) || c.dAL.ensure(sizeof(double) * 9 * (size_t)A->NPL) ||
      c.dAU.ensure(sizeof(double) * 9 * (size_t)A->NPU))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.dD.p, A->D, sizeof(double) * 9 * (size_t)A->NP, hipMemcpyHostToDevice, c.stream));
  if (A->NPL) HIPC(hipMemcpyAsync(c.dAL.p, A->AL, sizeof(double) * 9 * (size_t)A->NPL, hipMemcpyHostToDevice, c.stream));
  if (A->NPU) HIPC(hipMemcpyAsync(c.dAU.p, A->AU, sizeof(double) * 9 * (size_t)A->NPU, hipMemcpyHostToDevice, c.stream));
  launch_fill_values_a(c.A.view, c.A.src.as<int64_t>(), c.dD.as<double>(), c.dAL.as<double>(), c.dAU.as<double>(),
                     c.N, c.NPL, c.stream);
  HIPC(hipGetLastError());
  c.a32_valid = false;
  c.mat_valid = false;
  c.coarse_num = false;
  c.contrast = -1.0;
  return 0;
}

// Stiffness contrast of the system: max / min over nodes of |tr D_ii|,
// leaving out the unit rows of fixed degrees of freedom (trace 3 after
// hecmw_mat_ass_bc: all three fixed) -- computed on the host copy of D once
// per value upload.
static int32_t diag_contrast(Context& c, double* out) {
  if (c.contrast >= 0.0) {
    *out = c.contrast;
    return 0;
  }
  std::vector<double> d(9 * (size_t)c.N);
  HIPC(hipMemcpyAsync(d.data(), c.dD.p, sizeof(double) * d.size(), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  double mn = 1e308, mx = 0.0;
  for (int32_t i = 0; i < c.N; ++i) {
    const double* b = &d[9 * (size_t)i];
    const bool unit = b[0] == 1.0 && b[4] == 1.0 && b[8] == 1.0 && b[1] == 0.0 && b[2] == 0.0 && b[3] == 0.0 &&
                      b[5] == 0.0 && b[6] == 0.0 && b[7] == 0.0;
    if (unit) continue;
    const double t = std::fabs(b[0]) + std::fabs(b[4]) + std::fabs(b[8]);
    if (t > 0.0) {
      mn = std::min(mn, t);
      mx = std::max(mx, t);
    }
  }
  c.contrast = mx > 0.0 ? mx / mn : 1.0;
  *out = c.contrast;
  return 0;
}

// ---- preconditioners ----------------------------------------------------------
// Symbolic part: ordering, L'/U' structure, lane assignment.  Depends on the
// structure and NCOLOR_IN only.
static int32_t setup_ssor_symbolic(Context& c, int32_t ncolor_in) {
  const int32_t N = c.N;
  const int32_t* iL = c.h_indexL.data();
  const int32_t* iU = c.h_indexU.data();
  const int32_t* tL = c.h_itemL.data();
  const int32_t* tU = c.h_itemU.data();
  std::vector<int32_t> perm_rcm(N), iperm_rcm(N);
  ordering_rcm(N, iL, tL, iU, tU, perm_rcm.data(), iperm_rcm.data());
  c.perm.assign(N, 0);
  c.iperm.assign(N, 0);
  c.colorindex.assign(N + 1, 0);
  c.ncolor = ordering_mc(N, iL, tL, iU, tU, perm_rcm.data(), ncolor_in, c.colorindex.data(), c.perm.data(),
                         c.iperm.data());
  std::vector<int32_t> iLp, tLp, iUp, tUp;
  reorder_profile(N, c.perm.data(), c.iperm.data(), iL, tL, iU, tU, iLp, tLp, iUp, tUp);
  // Rows per slice: 16 (four lanes per row) or 64 (a lane per row).  The
  // lane-per-row sweep needs a quarter of the waves, which pays once a colour
  // holds many more slices than the GPU keeps resident; HECMW33_SSOR_ROWS
  // (16 or 64) overrides.
  int R = (int64_t)N / std::max(1, c.ncolor) > kSsorRows64Min ? kRowsA : kRowsS;
  if (const char* e = getenv("HECMW33_SSOR_ROWS")) R = atoi(e) == 64 ? kRowsA : (atoi(e) == 16 ? kRowsS : R);
  c.ssor_R = R;
  // slices per colour
  c.color_slice.assign(c.ncolor + 1, 0);
  for (int32_t k = 0; k < c.ncolor; ++k) {
    const int32_t cnt = c.colorindex[k + 1] - c.colorindex[k];
    c.color_slice[k + 1] = c.color_slice[k] + (cnt + R - 1) / R;
  }
  const int32_t ns = c.color_slice[c.ncolor];
  std::vector<int32_t> lanes((int64_t)ns * R, -1), row_old((int64_t)ns * R, -1);
  // Rows of one colour are independent, so the lane that sweeps a row is
  // free.  A colour's rows are dealt in natural node order, so neighbouring
  // lanes gather neighbouring nodes (2-6% faster sweeps than the colour order
  // on the 100^3 and 160^3 cubes), and then, inside windows of 4 slices, by
  // falling L' length: an interior row's L' and U' lengths add up to its
  // degree, so the rows of a slice end their forward and their backward
  // walks together and the slices carry little padding (SELL padding 1.19-1.26x
  // -> 1.06-1.09x; 100^3 FP64 application 1.000 -> 0.929 ms, 160^3 FP32
  // 1.802 -> 1.544 ms).  HECMW33_SSOR_SORT=0 keeps the colour order, 1 the
  // natural order alone; HECMW33_SSOR_WIN sets the window in slices.
  const char* sort_env = getenv("HECMW33_SSOR_SORT");
  const int sort_rows = sort_env ? atoi(sort_env) : 2;
  const char* win_env = getenv("HECMW33_SSOR_WIN");
  const int64_t win = (int64_t)R * (win_env ? std::max(1, atoi(win_env)) : 4);
  std::vector<int32_t> order;
  for (int32_t k = 0; k < c.ncolor; ++k) {
    order.resize(c.colorindex[k + 1] - c.colorindex[k]);
    for (int32_t q = 0; q < (int32_t)order.size(); ++q) order[q] = c.colorindex[k] + q;
    if (sort_rows) std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return c.perm[a] < c.perm[b]; });
    if (sort_rows == 2)
      for (int64_t w0 = 0; w0 < (int64_t)order.size(); w0 += win)
        std::stable_sort(order.begin() + w0, order.begin() + std::min<int64_t>(w0 + win, order.size()),
                         [&](int32_t a, int32_t b) { return iLp[a + 1] - iLp[a] > iLp[b + 1] - iLp[b]; });
    for (int32_t q = 0; q < (int32_t)order.size(); ++q) {
      const int32_t inew = order[q];
      const int64_t t = (int64_t)c.color_slice[k] * R + q;
      lanes[t] = inew;                    // 0-based new position
      row_old[t] = c.perm[inew] - 1;      // 0-based old id
    }
  }
  // slot of every row (new order): the sweeps' columns
  std::vector<int32_t> slot_of(N, -1);
  for (int64_t t = 0; t < (int64_t)lanes.size(); ++t)
    if (lanes[t] >= 0) slot_of[lanes[t]] = (int32_t)t;
  const int64_t offL = N, offU = (int64_t)N + c.NPL;
  const std::vector<int32_t>& perm = c.perm;
  // source block of A(iold, kold) inside the caller's AL/AU (row iold)
  auto find_src = [&](int32_t iold1, int32_t kold1) -> int64_t {
    if (kold1 < iold1) {
      for (int32_t q = iL[iold1 - 1]; q < iL[iold1]; ++q)
        if (tL[q] == kold1) return offL + q;
    } else {
      for (int32_t q = iU[iold1 - 1]; q < iU[iold1]; ++q)
        if (tU[q] == kold1) return offU + q;
    }
    return -1;
  };
  SellHost hl, hu;
  auto build_lu = [&](auto rows_tag) {
    constexpr int RR = decltype(rows_tag)::value;
    build_sell<RR>(
        lanes, N, [&](int32_t inew) { return iLp[inew + 1] - iLp[inew]; },
        [&](int32_t inew, auto emit) {
          const int32_t iold1 = perm[inew];
          int32_t k = 0;
          for (int32_t q = iLp[inew]; q < iLp[inew + 1]; ++q) {
            const int32_t kold1 = perm[tLp[q] - 1];
            emit(k++, slot_of[tLp[q] - 1], find_src(iold1, kold1));
          }
        },
        hl);
    build_sell<RR>(
        lanes, N, [&](int32_t inew) { return iUp[inew + 1] - iUp[inew]; },
        [&](int32_t inew, auto emit) {
          const int32_t iold1 = perm[inew];
          int32_t k = 0;
          for (int32_t q = iUp[inew]; q < iUp[inew + 1]; ++q) {
            const int32_t kold1 = perm[tUp[q] - 1];
            emit(k++, slot_of[tUp[q] - 1], find_src(iold1, kold1));
          }
        },
        hu);
  };
  if (R == kRowsA)
    build_lu(std::integral_constant<int, kRowsA>{});
  else
    build_lu(std::integral_constant<int, kRowsS>{});
  if (int32_t e = upload_sell(hl, c.L, c.stream)) return e;
  if (int32_t e = upload_sell(hu, c.U, c.stream)) return e;
  // the slot of every node, for staging r into slot order and z back out
  std::vector<int32_t> node_slot(N, -1);
  for (int64_t t = 0; t < (int64_t)row_old.size(); ++t)
    if (row_old[t] >= 0) node_slot[row_old[t]] = (int32_t)t;
  // the last colour is an independent set with the largest new indices, so
  // its rows have no U' entries and its backward step subtracts LU^-1 0 = 0
  c.ssor_last_bwd_empty = 1;
  for (int32_t inew = c.colorindex[c.ncolor - 1]; inew < N; ++inew)
    if (iUp[inew + 1] != iUp[inew]) c.ssor_last_bwd_empty = 0;
  if (c.row_old.ensure(sizeof(int32_t) * row_old.size()) || c.alu_ssor.ensure(sizeof(double) * 9 * row_old.size()) ||
      c.node_slot.ensure(sizeof(int32_t) * (size_t)N))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.row_old.p, row_old.data(), sizeof(int32_t) * row_old.size(), hipMemcpyHostToDevice,
                      c.stream));
  HIPC(hipMemcpyAsync(c.node_slot.p, node_slot.data(), sizeof(int32_t) * (size_t)N, hipMemcpyHostToDevice, c.stream));
  if (c.vzs.ensure(sizeof(double) * 3 * row_old.size())) return HECMW33_ERROR_DEVICE;
  HIPC(hipMemsetAsync(c.vzs.p, 0, sizeof(double) * 3 * row_old.size(), c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  c.ssor_rows = (int64_t)row_old.size();
  c.ssor_ncolor_in = ncolor_in;
  c.ssor_sym = true;
  return 0;
}

// hecmw_precond_SSOR_33_setup (hecmw_precond_SSOR_33.F90:57-218).  The
// reference redoes the ordering also when only a numeric setup is asked for
// (Iarray 97 = 1, 98 = 0: ":75-79, numerical setup only ... TEMPORARY"); the
// ordering depends on the structure and NCOLOR_IN alone, so it is kept here
// and only the values and factors are refreshed, with the same result.
static int32_t setup_ssor(Context& c, double sigma, int32_t ncolor_in, bool symbolic) {
  if (symbolic || !c.ssor_sym || c.ssor_ncolor_in != ncolor_in)
    if (int32_t e = setup_ssor_symbolic(c, ncolor_in)) return e;
  const int32_t N = c.N;
  launch_fill_values(c.L.view, c.L.src.as<int64_t>(), c.dD.as<double>(), c.dAL.as<double>(), c.dAU.as<double>(), N,
                     c.NPL, c.stream);
  launch_fill_values(c.U.view, c.U.src.as<int64_t>(), c.dD.as<double>(), c.dAL.as<double>(), c.dAU.as<double>(), N,
                     c.NPL, c.stream);
  launch_block_lu((int32_t)c.ssor_rows, c.row_old.as<int32_t>(), c.dD.as<double>(), sigma, c.alu_ssor.as<double>(),
                  c.ssor_R, c.stream);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(c.stream));
  c.ssor_init = true;
  c.ssor32_valid = false;
  c.ssor16_valid = false;
  return 0;
}

static int32_t setup_diag(Context& c, double sigma) {
  const int32_t ns = (c.N + kWave - 1) / kWave;
  std::vector<int32_t> rows((int64_t)ns * kWave, -1);
  for (int32_t i = 0; i < c.N; ++i) rows[i] = i;
  if (c.row_diag.ensure(sizeof(int32_t) * rows.size()) || c.alu_diag.ensure(sizeof(double) * 9 * rows.size()))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.row_diag.p, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, c.stream));
  launch_block_lu((int32_t)rows.size(), c.row_diag.as<int32_t>(), c.dD.as<double>(), sigma, c.alu_diag.as<double>(),
                  kWave, c.stream);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(c.stream));
  c.diag_init = true;
  c.diag32_valid = false;
  return 0;
}

// FP32 images of A and of the preconditioner in use, rounded from the FP64
// storage (so a reused preconditioner keeps the values of its setup).
static int32_t round_sell(SellDev& d, hipStream_t s) {
  const int64_t n = 9 * std::max<int64_t>(d.view.nslots, 1);
  if (d.val32.ensure(sizeof(float) * n)) return HECMW33_ERROR_DEVICE;
  d.view32.val = d.val32.as<float>();
  launch_to_f32(9 * d.view.nslots, d.view.val, d.view32.val, s);
  return 0;
}
// v16: the SSOR sweeps will read BF16 copies of the L'/U' values (rounded to
// nearest even from the FP32 image)
static int32_t setup_f32(Context& c, int32_t kind, bool v16) {
  if (!c.a32_valid) {
    if (int32_t e = round_sell(c.A, c.stream)) return e;
    c.a32_valid = true;
  }
  if (kind == 1 && !c.ssor32_valid) {
    if (round_sell(c.L, c.stream) || round_sell(c.U, c.stream)) return HECMW33_ERROR_DEVICE;
    c.ssor16_valid = false;
    const int64_t nslot = c.ssor_rows;
    if (c.alu_ssor32.ensure(sizeof(float) * 9 * std::max<int64_t>(nslot, 1)) ||
        c.vzs32.ensure(sizeof(float) * 3 * std::max<int64_t>(nslot, 1)))
      return HECMW33_ERROR_DEVICE;
    HIPC(hipMemsetAsync(c.vzs32.p, 0, sizeof(float) * 3 * std::max<int64_t>(nslot, 1), c.stream));
    launch_to_f32(9 * nslot, c.alu_ssor.as<double>(), c.alu_ssor32.as<float>(), c.stream);
    c.ssor32_valid = true;
  }
  if (kind == 1 && v16 && !c.ssor16_valid) {
    for (SellDev* d : {&c.L, &c.U}) {
      const int64_t n = 9 * std::max<int64_t>(d->view.nslots, 1);
      if (d->val16.ensure(sizeof(uint16_t) * n)) return HECMW33_ERROR_DEVICE;
      launch_to_bf16(9 * d->view.nslots, d->view32.val, d->val16.as<uint16_t>(), c.stream);
    }
    c.ssor16_valid = true;
  }
  if (kind == 3 && !c.diag32_valid) {
    const int64_t nslot = (int64_t)((c.N + kWave - 1) / kWave) * kWave;
    if (c.alu_diag32.ensure(sizeof(float) * 9 * nslot)) return HECMW33_ERROR_DEVICE;
    launch_to_f32(9 * nslot, c.alu_diag.as<double>(), c.alu_diag32.as<float>(), c.stream);
    c.diag32_valid = true;
  }
  const size_t vb = sizeof(float) * 3 * (size_t)c.NP;
  for (DMem* m : {&c.vr32, &c.vz32, &c.vp32, &c.vq32})
    if (m->ensure(vb)) return HECMW33_ERROR_DEVICE;
  HIPC(hipGetLastError());
  return 0;
}

// hecmw_precond_setup (hecmw_precond.f90:30-58) with the SSOR/DIAG
// reuse rule (hecmw_precond_SSOR_33.F90:74-82, hecmw_precond_DIAG_33.F90:38-41).
// Static node features of the learned aggregation (node stiffness and its
// stencil statistics), refreshed with the values.
static int32_t features_setup(Context& c) {
  if (c.vmat.ensure(sizeof(float) * (size_t)c.N) || c.scal.ensure(sizeof(double) * 8) ||
      c.vfeat.ensure(2 * 16 * (size_t)c.N) || c.vnety.ensure(sizeof(float) * 3 * (size_t)c.N) ||
      c.vmask.ensure((size_t)c.N))
    return HECMW33_ERROR_DEVICE;
  if (!c.mat_valid) {
    net_material(c.N, c.dD.as<double>(), c.vmat.as<float>(), c.part.as<double>(), c.scal.as<double>(), c.stream);
    net_static_features(c.A.view, c.N, c.vmat.as<float>(), c.vfeat.as<__bf16>(), c.stream);
    c.mat_valid = true;
  }
  HIPC(hipGetLastError());
  return 0;
}

static int32_t precond_setup(Context& c, int32_t* iarray, const double* rarray) {
  if (IA(iarray, I_ITERPREMAX) <= 0) return 0;
  const int32_t pc = IA(iarray, I_PRECOND);
  const double sigma = RA(rarray, R_SIGMA_DIAG);
  const bool want = IA(iarray, I_FLAG_SYMBFACT) == 1 || IA(iarray, I_FLAG_NUMFACT) == 1;
  if (pc == 1 || pc == 2) {
    if (c.ssor_init && !want) return 0;
    if (int32_t e = setup_ssor(c, sigma, IA(iarray, I_NCOLOR_IN), IA(iarray, I_FLAG_SYMBFACT) == 1)) return e;
  } else if (pc == 3) {
    if (c.diag_init && IA(iarray, I_FLAG_SYMBFACT) == 0 && IA(iarray, I_FLAG_NUMFACT) == 0) return 0;
    if (int32_t e = setup_diag(c, sigma)) return e;
  } else {
    set_error("PRECOND not provided by the GPU library (1, 2: SSOR; 3: DIAG)");
    return HECMW33_ERROR_INCONS_PC;
  }
  // the two-level correction's coarse inverse and its learned modes are part of the numeric set-up
  c.coarse_num = false;
  c.modes_valid = false;
  IA(iarray, I_FLAG_SYMBFACT) = 0;
  IA(iarray, I_FLAG_NUMFACT) = 0;
  return 0;
}

// ---- Krylov iterations ----------------------------------------------------------
// Preconditioner kind as the sweeps see it: 0 none, 1 multicolour SSOR, 3 DIAG.
static int32_t pc_kind(int32_t pc, int32_t iterpremax) {
  if (iterpremax <= 0) return 0;
  return pc == 3 ? 3 : 1;
}

// z = M^-1 r with hecmw_precond_apply's ITERPREMAX loop (hecmw_precond.f90:114-168):
// ZP = R; Z = 0; repeat: ZP = M^-1 ZP; Z += ZP; ZP = R - A Z (except after the last).
static void enqueue_precond(Context& c, int32_t kind, int32_t iterpremax, const double* r, double* z) {
  KrylovState* st = c.st.as<KrylovState>();
  const int64_t n3 = 3 * (int64_t)c.N;
  auto apply = [&](const double* in, double* out) {
    precond_apply(kind, c.ssor_op(), c.alu_diag.as<double>(), c.N, in, out, st, c.stream);
  };
  if (kind == 0 || iterpremax <= 1) {
    apply(r, z);
    return;
  }
  double* zp = c.vzp.as<double>();
  double* zt = c.vzt.as<double>();
  launch_zero(3 * (int64_t)c.NP, z, st, c.stream);
  launch_copy(n3, r, zp, st, c.stream);
  for (int it = 1; it <= iterpremax; ++it) {
    apply(zp, zt);
    launch_add(n3, zt, z, st, c.stream);
    if (it == iterpremax) break;
    launch_matresid_guarded(c.A.view, z, r, zp, st, c.stream, c.mpc_op());
  }
}

// One CG iteration (hecmw_solver_CG.f90:164-279).
static void enqueue_cg_iteration(Context& c, int32_t kind, int32_t iterpremax, hipEvent_t* ev) {
  const int64_t n3 = 3 * (int64_t)c.N;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  double* r = c.vr.as<double>();
  double* z = c.vz.as<double>();
  if (ev) (void)hipEventRecord(ev[0], c.stream);
  if (iterpremax > 1 && kind != 0) {
    enqueue_precond(c, kind, iterpremax, r, z);
    cg_rho(n3, r, z, st, part, c.stream);
  } else if (kind == 0) {
    cg_precond_none_rho(n3, r, z, st, part, c.stream);
  } else if (kind == 3) {
    cg_precond_diag_rho(c.N, c.alu_diag.as<double>(), r, z, st, part, c.stream);
  } else {
    cg_precond_ssor_rho(c.ssor_op(), r, z, st, part, c.stream);
  }
  if (ev) (void)hipEventRecord(ev[1], c.stream);
  cg_xpay(n3, z, c.vp.as<double>(), st, c.stream);
  if (ev) (void)hipEventRecord(ev[2], c.stream);
  cg_spmv_dot(c.A.view, c.vp.as<double>(), c.vq.as<double>(), st, part, c.stream, ev ? ev[3] : nullptr,
              c.mpc_op());
  cg_update(n3, c.vp.as<double>(), c.vq.as<double>(), c.vx.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
  cg_true_resid(c.A.view, c.vx.as<double>(), c.vb.as<double>(), r, st, part, c.hist.as<double>(), c.stream,
                c.mpc_op());
}

// One BiCGSTAB iteration (hecmw_solver_BiCGSTAB.f90:164-275).  The reference
// keeps s~ in R's slot (ST aliases R); a separate vector gives the same values.
static void enqueue_bicg_iteration(Context& c, int32_t kind, int32_t iterpremax, hipEvent_t* ev) {
  const int64_t n3 = 3 * (int64_t)c.N;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  double *r = c.vr.as<double>(), *rt = c.vrt.as<double>(), *p = c.vp.as<double>(), *pt = c.vpt.as<double>();
  double *v = c.vv.as<double>(), *sv = c.vs.as<double>(), *stv = c.vst.as<double>(), *t = c.vt.as<double>();
  bi_rho(n3, r, rt, st, part, c.stream);
  bi_p(n3, r, v, p, st, c.stream);
  if (ev) (void)hipEventRecord(ev[0], c.stream);
  enqueue_precond(c, kind, iterpremax, p, pt);
  if (ev) (void)hipEventRecord(ev[1], c.stream);
  if (ev) (void)hipEventRecord(ev[2], c.stream);
  bi_spmv_c2(c.A.view, pt, rt, v, st, part, c.stream, ev ? ev[3] : nullptr, c.mpc_op());
  bi_s(n3, v, r, sv, st, c.stream);
  if (ev) (void)hipEventRecord(ev[4], c.stream);
  enqueue_precond(c, kind, iterpremax, sv, stv);
  if (ev) (void)hipEventRecord(ev[5], c.stream);
  if (ev) (void)hipEventRecord(ev[6], c.stream);
  bi_spmv_omega(c.A.view, stv, sv, t, st, part, c.stream, ev ? ev[7] : nullptr, c.mpc_op());
  bi_update(n3, pt, stv, t, sv, c.vx.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
  cg_true_resid(c.A.view, c.vx.as<double>(), c.vb.as<double>(), r, st, part, c.hist.as<double>(), c.stream,
                c.mpc_op());
}

// One flexible CG iteration with FP32 operators: z = M32^-1 r (fused r.z,
// z.q_prev), p = z + beta p, q = A32 p (fused p.q, r.p), x += alpha p (FP64),
// r -= alpha q, then, when the recurrence asks for it, r = b - A x in FP64.
// ---- two-level correction: aggregation (symbolic) and coarse inverse (numeric) ----------
// Aggregates are the non-empty cells of a uniform grid over the nodes' bounding
// box, sized for about `target` of them (grown while there are more than
// kCoarseAggMax).  E_ab is assembled for every pair of aggregates coupled by a
// matrix block; the pair list and, per pair, the nodes of a whose rows touch b
// come from the matrix graph, so any aggregation works.

static int32_t coarse_symbolic(Context& c, int32_t target) {
  const int32_t N = c.N, NP = c.NP;
  const double* X = c.h_coord.data();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int32_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], X[3 * (int64_t)i + d]);
      hi[d] = std::max(hi[d], X[3 * (int64_t)i + d]);
    }
  double ext[3], vol = 1.0;
  int ndim = 0;
  for (int d = 0; d < 3; ++d) {
    ext[d] = hi[d] - lo[d];
    if (!(ext[d] > 0.0)) ext[d] = 0.0;
  }
  const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
  for (int d = 0; d < 3; ++d)
    if (ext[d] > 1e-9 * emax) {
      vol *= ext[d];
      ++ndim;
    } else {
      ext[d] = 0.0;
    }
  if (ndim == 0) {
    set_error("coarse correction: all nodes at one point");
    return HECMW33_ERROR_ARGS;
  }
  const int32_t* iL = c.h_indexL.data();
  const int32_t* iU = c.h_indexU.data();
  const int32_t* tL = c.h_itemL.data();
  const int32_t* tU = c.h_itemU.data();

  // aggregates: the non-empty cells, numbered in order of first appearance over the nodes
  std::vector<int32_t> agg(NP, -1);
  int32_t na = 0;
  double H = std::pow(vol / std::max(1, target), 1.0 / ndim);
  for (int attempt = 0; attempt < 60; ++attempt, H *= 1.15) {
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = ext[d] > 0.0 ? std::max(1, (int)std::ceil(ext[d] / H - 1e-9)) : 1;
    const int64_t ncells = (int64_t)nc[0] * nc[1] * nc[2];
    if (ncells > 64 * (int64_t)kCoarseAggMax) continue;
    std::vector<int32_t> cell_agg((size_t)ncells, -1);
    na = 0;
    bool over = false;
    for (int32_t i = 0; i < N && !over; ++i) {
      int cd[3];
      for (int d = 0; d < 3; ++d)
        cd[d] = ext[d] > 0.0 ? std::min(nc[d] - 1, (int)((X[3 * (int64_t)i + d] - lo[d]) / ext[d] * nc[d])) : 0;
      int32_t& slot = cell_agg[cd[0] + (int64_t)nc[0] * (cd[1] + (int64_t)nc[1] * cd[2])];
      if (slot < 0) slot = na++;
      agg[i] = slot;
      over = na > kCoarseAggMax;
    }
    if (!over) break;
    na = 0;
  }
  if (na == 0) {
    set_error("coarse correction: no aggregation within the size limit");
    return HECMW33_ERROR_ARGS;
  }
  std::vector<int32_t> cnt(na + 1, 0);
  for (int32_t i = 0; i < N; ++i) ++cnt[agg[i] + 1];
  for (int32_t a = 0; a < na; ++a) cnt[a + 1] += cnt[a];
  std::vector<int32_t> anodes(N), fillp(cnt.begin(), cnt.end() - 1);
  for (int32_t i = 0; i < N; ++i) anodes[fillp[agg[i]]++] = i;

  // centroids; an aggregate whose nodes are collinear (or a single node) has a
  // rotation that moves none of them: its offsets are set to 0, so its three
  // rotation columns of Z vanish, and E gets unit pivots for them
  std::vector<double> cen(3 * (size_t)na, 0.0);
  std::vector<char> degen(na, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t a = 0; a < na; ++a) {
    double sx = 0, sy = 0, sz = 0;
    for (int32_t t = cnt[a]; t < cnt[a + 1]; ++t) {
      const int64_t i = anodes[t];
      sx += X[3 * i];
      sy += X[3 * i + 1];
      sz += X[3 * i + 2];
    }
    const double k = 1.0 / (cnt[a + 1] - cnt[a]);
    cen[3 * (size_t)a] = sx * k;
    cen[3 * (size_t)a + 1] = sy * k;
    cen[3 * (size_t)a + 2] = sz * k;
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int32_t t = cnt[a]; t < cnt[a + 1]; ++t) {
      const int64_t i = anodes[t];
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = X[3 * i + d] - cen[3 * (size_t)a + d];
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) M[p][q] += r[p] * r[q];
    }
    const double trm = M[0][0] + M[1][1] + M[2][2];
    double f2 = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) f2 += M[p][q] * M[p][q];
    // (tr^2 - |M|_F^2) / 2 = l1 l2 + l1 l3 + l2 l3: ~0 when rank(M) <= 1
    degen[a] = !(0.5 * (trm * trm - f2) > 1e-8 * trm * trm);
  }
  const double sc = 2.0 / H;
  std::vector<float> dv(3 * (size_t)N);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < N; ++i) {
    const int32_t a = agg[i];
    for (int d = 0; d < 3; ++d)
      dv[3 * (size_t)i + d] = degen[a] ? 0.f : (float)((X[3 * (int64_t)i + d] - cen[3 * (size_t)a + d]) * sc);
  }

  // per node, the distinct aggregates its row touches (its own first)
  std::vector<int32_t> ncnt(N + 1, 0);
  auto row_aggs = [&](int32_t i, int32_t* out) {
    int n = 0;
    out[n++] = agg[i];
    auto add = [&](int32_t j) {
      if (j >= N) return;
      const int32_t b = agg[j];
      for (int k = 0; k < n; ++k)
        if (out[k] == b) return;
      out[n++] = b;
    };
    for (int32_t q = iL[i]; q < iL[i + 1]; ++q) add(tL[q] - 1);
    for (int32_t q = iU[i]; q < iU[i + 1]; ++q) add(tU[q] - 1);
    return n;
  };
  int32_t maxrow = 1;
  for (int32_t i = 0; i < N; ++i) maxrow = std::max(maxrow, 1 + (iL[i + 1] - iL[i]) + (iU[i + 1] - iU[i]));
#pragma omp parallel
  {
    std::vector<int32_t> buf(maxrow);
#pragma omp for schedule(dynamic, 4096)
    for (int32_t i = 0; i < N; ++i) ncnt[i + 1] = row_aggs(i, buf.data());
  }
  for (int32_t i = 0; i < N; ++i) ncnt[i + 1] += ncnt[i];
  std::vector<int32_t> nagg(ncnt[N]);
#pragma omp parallel
  {
    std::vector<int32_t> buf(maxrow);
#pragma omp for schedule(dynamic, 4096)
    for (int32_t i = 0; i < N; ++i) {
      const int n = row_aggs(i, buf.data());
      std::copy(buf.data(), buf.data() + n, nagg.data() + ncnt[i]);
    }
  }
  // aggregate pairs (a, b), b ascending per a, and the nodes of a touching b
  std::vector<std::vector<int32_t>> adj(na);
  for (int32_t i = 0; i < N; ++i)
    for (int32_t k = ncnt[i]; k < ncnt[i + 1]; ++k) adj[agg[i]].push_back(nagg[k]);
  std::vector<int32_t> pptr(na + 1, 0);
  for (int32_t a = 0; a < na; ++a) {
    std::sort(adj[a].begin(), adj[a].end());
    adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
    pptr[a + 1] = pptr[a] + (int32_t)adj[a].size();
  }
  const int32_t npair = pptr[na];
  std::vector<int32_t> pa(npair), pb(npair), pn_ptr(npair + 1, 0);
  for (int32_t a = 0; a < na; ++a)
    for (size_t k = 0; k < adj[a].size(); ++k) {
      pa[pptr[a] + k] = a;
      pb[pptr[a] + k] = adj[a][k];
    }
  auto pair_of = [&](int32_t a, int32_t b) {
    return pptr[a] + (int32_t)(std::lower_bound(adj[a].begin(), adj[a].end(), b) - adj[a].begin());
  };
  for (int32_t i = 0; i < N; ++i)
    for (int32_t k = ncnt[i]; k < ncnt[i + 1]; ++k) ++pn_ptr[pair_of(agg[i], nagg[k]) + 1];
  for (int32_t p = 0; p < npair; ++p) pn_ptr[p + 1] += pn_ptr[p];
  std::vector<int32_t> pn(pn_ptr[npair]), pfill(pn_ptr.begin(), pn_ptr.end() - 1);
  for (int32_t i = 0; i < N; ++i)
    for (int32_t k = ncnt[i]; k < ncnt[i + 1]; ++k) pn[pfill[pair_of(agg[i], nagg[k])]++] = i;

  CoarseOp& op = c.coarse;
  op.n_agg = na;
  op.nc = 0;
  op.ldw = 0;
  c.coarse_npair = npair;
  c.coarse_band_agg = 0;
  for (int32_t p = 0; p < npair; ++p) c.coarse_band_agg = std::max(c.coarse_band_agg, std::abs(pa[p] - pb[p]));
  c.coarse_nw = ncnt[N];
  if (c.cg_ncnt.ensure(sizeof(int32_t) * (N + 1)) || c.cg_nagg.ensure(sizeof(int32_t) * std::max<int64_t>(ncnt[N], 1)))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.cg_ncnt.p, ncnt.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_nagg.p, nagg.data(), sizeof(int32_t) * ncnt[N], hipMemcpyHostToDevice, c.stream));
  c.coarse_w_valid = false;
  const int64_t npn = pn_ptr[npair];
  if (c.cg_agg.ensure(sizeof(int32_t) * NP) || c.cg_aptr.ensure(sizeof(int32_t) * (na + 1)) ||
      c.cg_anodes.ensure(sizeof(int32_t) * N) || c.cg_dv.ensure(sizeof(float) * 3 * (size_t)N) ||
      c.cg_pa.ensure(sizeof(int32_t) * npair) || c.cg_pb.ensure(sizeof(int32_t) * npair) ||
      c.cg_pnptr.ensure(sizeof(int32_t) * (npair + 1)) || c.cg_pn.ensure(sizeof(int32_t) * std::max<int64_t>(npn, 1)))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.cg_agg.p, agg.data(), sizeof(int32_t) * NP, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_aptr.p, cnt.data(), sizeof(int32_t) * (na + 1), hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_anodes.p, anodes.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_dv.p, dv.data(), sizeof(float) * 3 * (size_t)N, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_pa.p, pa.data(), sizeof(int32_t) * npair, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_pb.p, pb.data(), sizeof(int32_t) * npair, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.cg_pnptr.p, pn_ptr.data(), sizeof(int32_t) * (npair + 1), hipMemcpyHostToDevice, c.stream));
  if (npn) HIPC(hipMemcpyAsync(c.cg_pn.p, pn.data(), sizeof(int32_t) * npn, hipMemcpyHostToDevice, c.stream));
  op.agg = c.cg_agg.as<int32_t>();
  op.aptr = c.cg_aptr.as<int32_t>();
  op.anodes = c.cg_anodes.as<int32_t>();
  op.dv = c.cg_dv.as<float>();
  HIPC(hipStreamSynchronize(c.stream));  // the host vectors above go out of scope
  c.coarse_sym = true;
  c.coarse_num = false;
  c.modes_valid = false;
  c.coarse_target = target;
  c.coarse_coord_gen = c.coord_gen;
  return 0;
}

// E = Z^T A Z on the device with nm modes per aggregate (6 rigid-body; 12 with
// the learned ones in cg_xm), Cholesky factor and inverse (six modes) or
// inverse factor (twelve) in FP64 (chol_band, then rocSOLVER), then the copy the iterations
// read
// Cholesky E = L L^T of the coarse matrix (lower, column-major, n x n), which
// is banded: aggregates couple only when their cells touch, and the cells are
// numbered in order of first appearance over the nodes, so on a mesh numbered
// along its axes every nonzero lies within `band` rows below the diagonal (at
// 160^3 with 1000 aggregates: 1332 of 12,000 rows), and so does every nonzero
// of L.  Right-looking over panels of kCholPanel columns: rocSOLVER factors the
// panel's diagonal block, a triangular solve gives the at most `band` rows
// below it, and a rank-kb product updates the trailing band; the zeros outside
// the band are never touched.  info[p] != 0: panel p's block is not positive
// definite.  Same operations as the dense right-looking factorization, minus
// the ones on zeros.
constexpr int32_t kCholPanel = 256;
static int32_t chol_band(rocblas_handle h, int32_t n, double* A, int32_t band, int32_t* info) {
  const double one = 1.0, mone = -1.0;
  for (int32_t k = 0, p = 0; k < n; k += kCholPanel, ++p) {
    const int32_t kb = std::min(kCholPanel, n - k);
    double* Akk = A + (size_t)k * n + k;
    if (rocsolver_dpotrf(h, rocblas_fill_lower, kb, Akk, n, info + p) != rocblas_status_success) return 1;
    const int32_t m = std::min(n - k - kb, band);
    if (m <= 0) continue;
    if (rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, m, kb, &one, Akk, n, Akk + kb, n) != rocblas_status_success ||
        rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, m, kb, &mone, Akk + kb, n, &one,
                      A + (size_t)(k + kb) * n + k + kb, n) != rocblas_status_success)
      return 1;
  }
  return 0;
}

static int32_t coarse_numeric(Context& c, int32_t nm) {
  CoarseOp& op = c.coarse;
  op.nm = nm;
  op.nc = nm * op.n_agg;
  op.ldw = (op.nc + 63) / 64 * 64;
  op.xm = nm == 12 ? c.cg_xm.as<float>() : nullptr;
  const int32_t nc = op.nc;
  // six modes: E^-1, c and y in FP32; twelve: in FP64
  const size_t we = nm == 12 ? sizeof(double) : sizeof(float);
  const int32_t npanel = (nc + kCholPanel - 1) / kCholPanel;
  if (c.cg_info.ensure(sizeof(int32_t) * (npanel + 1)) ||
      c.cg_Eblk.ensure(sizeof(double) * nm * nm * (size_t)std::max(c.coarse_npair, 1)) ||
      c.cg_E.ensure(sizeof(double) * (size_t)nc * nc) || c.cg_W.ensure(we * (size_t)nc * op.ldw) ||
      c.cg_cv.ensure(we * op.ldw) || c.cg_yv.ensure(we * op.ldw) || c.cg_tv.ensure(we * op.ldw))
    return HECMW33_ERROR_DEVICE;
  op.W = nm == 12 ? nullptr : c.cg_W.as<float>();
  op.cv = nm == 12 ? nullptr : c.cg_cv.as<float>();
  op.yv = nm == 12 ? nullptr : c.cg_yv.as<float>();
  op.W64 = nm == 12 ? c.cg_W.as<double>() : nullptr;
  op.cv64 = nm == 12 ? c.cg_cv.as<double>() : nullptr;
  op.yv64 = nm == 12 ? c.cg_yv.as<double>() : nullptr;
  op.tv64 = nm == 12 ? c.cg_tv.as<double>() : nullptr;
  HIPC(hipMemsetAsync(c.cg_cv.p, 0, we * op.ldw, c.stream));
  HIPC(hipMemsetAsync(c.cg_tv.p, 0, we * op.ldw, c.stream));
  HIPC(hipMemsetAsync(c.cg_E.p, 0, sizeof(double) * (size_t)nc * nc, c.stream));
  coarse_assemble(c.A.view, op, c.coarse_npair, c.cg_pa.as<int32_t>(), c.cg_pb.as<int32_t>(),
                  c.cg_pnptr.as<int32_t>(), c.cg_pn.as<int32_t>(), c.cg_Eblk.as<double>(), c.cg_E.as<double>(), nc,
                  c.stream);
  HIPC(hipGetLastError());
  if (!c.blas && rocblas_create_handle(&c.blas) != rocblas_status_success) {
    c.blas = nullptr;
    set_error("rocblas_create_handle failed");
    return HECMW33_ERROR_DEVICE;
  }
  if (rocblas_set_stream(c.blas, c.stream) != rocblas_status_success) {
    set_error("rocblas_set_stream failed");
    return HECMW33_ERROR_DEVICE;
  }
  // six modes: E^-1 (potri) for one FP32 product per coarse solve; twelve:
  // L^-1 (trtri) for two FP64 triangular products, y = L^-T (L^-1 c), the
  // same bytes read without forming L^-T L^-1 (at 12,000 coarse unknowns
  // potri takes 57 ms, trtri 22 ms)
  int32_t* info = c.cg_info.as<int32_t>();
  const int32_t band = std::max(0, std::min(nc - 1, nm * (c.coarse_band_agg + 1) - 1));
  if (chol_band(c.blas, nc, c.cg_E.as<double>(), band, info) ||
      (nm == 12 ? rocsolver_dtrtri(c.blas, rocblas_fill_lower, rocblas_diagonal_non_unit, nc, c.cg_E.as<double>(), nc,
                                   info + npanel)
                : rocsolver_dpotri(c.blas, rocblas_fill_lower, nc, c.cg_E.as<double>(), nc, info + npanel)) !=
          rocblas_status_success) {
    set_error("coarse factorization (rocSOLVER potrf/potri/trtri, rocBLAS trsm/syrk) failed");
    return HECMW33_ERROR_DEVICE;
  }
  std::vector<int32_t> hinfo(npanel + 1, 0);
  HIPC(hipMemcpyAsync(hinfo.data(), info, sizeof(int32_t) * (npanel + 1), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  c.coarse_ok = std::all_of(hinfo.begin(), hinfo.end(), [](int32_t v) { return v == 0; });
  c.coarse_w_valid = false;
  if (c.coarse_ok) {
    if (nm == 12)
      coarse_inverse_f64(op, c.cg_E.as<double>(), nc, c.stream);
    else
      coarse_inverse_f32(op, c.cg_E.as<double>(), nc, c.stream);
  }
  HIPC(hipGetLastError());
  c.coarse_num = true;
  c.coarse_nm_num = nm;
  return 0;
}

static int32_t coarse_aim(const Context& c, int32_t opt3) {
  const int32_t t = opt3 == 1 ? kCoarseAggDefault : opt3;
  return std::max(1, std::min(std::min(t, kCoarseAggMax), c.N / kCoarseMinNodes));
}

static int32_t coarse_setup(Context& c, int32_t target, int32_t nm) {
  if ((int64_t)c.h_coord.size() != 3 * (int64_t)c.NP) {
    set_error(c.h_coord.empty()
                  ? "SOLVER_OPT(3) (two-level correction) needs the node coordinates of this system (hecmw33_set_coords)"
                  : "SOLVER_OPT(3): the node coordinates given do not have 3*NP entries for this system");
    return HECMW33_ERROR_ARGS;
  }
  target = coarse_aim(c, target);
  if (!c.coarse_sym || c.coarse_target != target || c.coarse_coord_gen != c.coord_gen)
    if (int32_t e = coarse_symbolic(c, target)) return e;
  if (nm == 12 && !c.modes_valid) {
    set_error("the learned coarse modes are not fitted for these values");
    return HECMW33_ERROR_ARGS;
  }
  if (!c.coarse_num || c.coarse_nm_num != nm)
    if (int32_t e = coarse_numeric(c, nm)) return e;
  return 0;
}

static void enqueue_fcg_iteration(Context& c, int32_t kind, bool spmv64, hipEvent_t* ev) {
  const int64_t n3 = 3 * (int64_t)c.N;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  float *r = c.vr32.as<float>(), *z = c.vz32.as<float>(), *p = c.vp32.as<float>(), *q = c.vq32.as<float>();
  if (ev) (void)hipEventRecord(ev[0], c.stream);
  if (c.coarse_use) {
    precond_apply_f32(kind, c.ssor_op32(c.ssor16), c.alu_diag32.as<float>(), c.N, r, z, st, c.stream);
    fcg_beta(st, part, coarse_apply(c.coarse, c.N, r, z, q, st, part, true, c.stream), c.stream);
  } else {
    fcg_precond_dots(kind, c.ssor_op32(c.ssor16), c.alu_diag32.as<float>(), c.N, r, z, q, st, part, c.stream);
  }
  if (ev) (void)hipEventRecord(ev[1], c.stream);
  fcg_p(n3, z, p, st, c.stream);
  if (ev) (void)hipEventRecord(ev[2], c.stream);
  if (spmv64)
    fcg_spmv_mixed(c.A.view, p, r, q, st, part, c.stream, ev ? ev[3] : nullptr);
  else
    fcg_spmv(c.A.view32, p, r, q, st, part, c.stream, ev ? ev[3] : nullptr);
  fcg_update(n3, p, q, c.vx.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
  fcg_reliable(c.A.view, c.vx.as<double>(), c.vb.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
}

// One flexible CG iteration with the symmetric two-level step (FP32 vectors):
//   z1 = Q r, r1 = r - A z1 (from A Z), z2 = z1 + M32^-1 r1, w = A z2 (the one SpMV),
//   z = z2 + Q (r - w), A z = w + A Z E^-1 Z^T (r - w);
// in the deflation form z2 = M32^-1 r (no z1);
// then beta from (r.z, z.q_prev), p = z + beta p and q = A z + beta q (so q = A p
// without a second SpMV), alpha from (p.q, r.p), and the update as the flexible CG.
static void enqueue_fcg_sym_iteration(Context& c, int32_t kind, bool spmv64, hipEvent_t* ev) {
  const int64_t n3 = 3 * (int64_t)c.N;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  float *r = c.vr32.as<float>(), *z = c.vz32.as<float>(), *p = c.vp32.as<float>(), *q = c.vq32.as<float>();
  float *r1 = c.vr1_32.as<float>(), *t = c.vt32.as<float>(), *w = c.vw32.as<float>(), *az = c.vaz32.as<float>();
  const int2* sc = c.cg_sc.as<int2>();
  const int32_t* aggW = c.cg_aggW.as<int32_t>();
  const float* AZ = c.cg_AZ.as<float>();
  if (ev) (void)hipEventRecord(ev[0], c.stream);
  if (c.coarse_defl) {
    precond_apply_f32(kind, c.ssor_op32(c.ssor16), c.alu_diag32.as<float>(), c.N, r, z, st, c.stream);
  } else {
    coarse_sym_pre(c.coarse, c.N, sc, aggW, AZ, r, z, r1, st, c.stream);
    precond_apply_f32(kind, c.ssor_op32(c.ssor16), c.alu_diag32.as<float>(), c.N, r1, t, st, c.stream);
    coarse_sym_add(n3, t, z, st, c.stream);
  }
  if (ev) (void)hipEventRecord(ev[1], c.stream);
  if (ev) (void)hipEventRecord(ev[2], c.stream);
  fcg_spmv_store(c.A.view32, c.A.view, spmv64, z, w, st, c.stream);
  if (ev) (void)hipEventRecord(ev[3], c.stream);
  fcg_beta(st, part, coarse_sym_post(c.coarse, c.N, sc, aggW, AZ, r, w, z, az, q, st, part, c.stream), c.stream);
  fcg_alpha(st, part, coarse_fcg_pq(n3, z, az, p, q, r, st, part, c.stream), c.stream);
  fcg_update(n3, p, q, c.vx.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
  fcg_reliable(c.A.view, c.vx.as<double>(), c.vb.as<double>(), r, st, part, c.hist.as<double>(), c.stream);
}

// Diagonal scaling of the device copies, hecmw_solver_scaling_fw_33 / _bk_33
// (hecmw_solver_scaling_33.f90:20-208): forward scales the uploaded D/AL/AU
// and b in place (x0 is left as given), backward unscales x and restores b
// and the matrix by division, as the reference restores hecMAT.  The SpMV
// storage is refilled after each.
static int32_t scaling_apply(Context& c, bool fw) {
  const int32_t N = c.N, NP = c.NP;
  if (!c.dindex_valid) {
    if (c.dIndexL.ensure(sizeof(int32_t) * (NP + 1)) || c.dIndexU.ensure(sizeof(int32_t) * (NP + 1)) ||
        c.dItemL.ensure(sizeof(int32_t) * std::max(c.NPL, 1)) || c.dItemU.ensure(sizeof(int32_t) * std::max(c.NPU, 1)))
      return HECMW33_ERROR_DEVICE;
    HIPC(hipMemcpyAsync(c.dIndexL.p, c.h_indexL.data(), sizeof(int32_t) * (NP + 1), hipMemcpyHostToDevice, c.stream));
    HIPC(hipMemcpyAsync(c.dIndexU.p, c.h_indexU.data(), sizeof(int32_t) * (NP + 1), hipMemcpyHostToDevice, c.stream));
    if (c.NPL)
      HIPC(hipMemcpyAsync(c.dItemL.p, c.h_itemL.data(), sizeof(int32_t) * c.NPL, hipMemcpyHostToDevice, c.stream));
    if (c.NPU)
      HIPC(hipMemcpyAsync(c.dItemU.p, c.h_itemU.data(), sizeof(int32_t) * c.NPU, hipMemcpyHostToDevice, c.stream));
    c.dindex_valid = true;
  }
  if (c.vscale.ensure(sizeof(double) * 3 * (size_t)NP)) return HECMW33_ERROR_DEVICE;
  double* sc = c.vscale.as<double>();
  const int64_t n3 = 3 * (int64_t)N;
  if (fw) {
    launch_make_scale(N, NP, c.dD.as<double>(), sc, c.stream);
    launch_scale_vec(true, n3, c.vb.as<double>(), sc, c.stream);
  } else {
    launch_scale_vec(true, n3, c.vx.as<double>(), sc, c.stream);
    launch_scale_vec(false, n3, c.vb.as<double>(), sc, c.stream);
  }
  launch_scale_raw(fw, NP, c.dIndexL.as<int32_t>(), c.dItemL.as<int32_t>(), c.dIndexU.as<int32_t>(),
                   c.dItemU.as<int32_t>(), c.dD.as<double>(), c.dAL.as<double>(), c.dAU.as<double>(), sc, c.stream);
  launch_fill_values_a(c.A.view, c.A.src.as<int64_t>(), c.dD.as<double>(), c.dAL.as<double>(), c.dAU.as<double>(),
                     c.N, c.NPL, c.stream);
  c.a32_valid = false;
  c.coarse_num = false;
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

static int32_t krylov_core(Context& c, int32_t* iarray, double* rarray, hecmw33_out* out);

// Krylov solve on the context's vectors vb (b) and vx (x0 in, x out), with
// the reference's optional diagonal scaling around it (Iarray 7).
static int32_t krylov_resident(Context& c, int32_t* iarray, double* rarray, hecmw33_out* out) {
  const bool scaling = IA(iarray, I_SCALING) != 0;
  double tsc = 0.0;
  if (scaling) {
    const double t0 = now_s();
    if (int32_t e = scaling_apply(c, true)) return e;
    tsc = now_s() - t0;
  }
  int32_t e = krylov_core(c, iarray, rarray, out);
  out->Tset += tsc;
  if (scaling) {
    const int32_t e2 = scaling_apply(c, false);
    if (!e) e = e2;
  }
  return e;
}

// ---- multipoint constraints: the MPC-CG product T^T A T ---------------------------------
// Validates hecMESH%mpc-style arrays for N internal nodes (1-based items and
// dofs, mpc_index(0:n) cumulative): constraints with a dof above 3 are skipped,
// as hecmw_Tvec_33 / hecmw_Ttvec_33 skip them; items must be internal nodes; a
// slave dof may be the slave of one constraint only and no constraint's master
// (the reference's parallel loops assume that too).
static int32_t mpc_check(int32_t N, int32_t n, const int32_t* index, const int32_t* item, const int32_t* dof) {
  if (n < 0 || N < 1 || (n > 0 && (!index || !item || !dof))) return HECMW33_ERROR_ARGS;
  if (n == 0) return 0;
  if (index[0] != 0) {
    set_error("mpc_index(0) must be 0");
    return HECMW33_ERROR_ARGS;
  }
  std::vector<char> role(3 * (size_t)N, 0);  // 1 slave, 2 master
  for (int32_t i = 0; i < n; ++i) {
    if (index[i + 1] <= index[i]) {
      set_error("a multipoint constraint without terms");
      return HECMW33_ERROR_ARGS;
    }
    bool skip = false;
    for (int32_t j = index[i]; j < index[i + 1]; ++j) {
      if (dof[j] > 3) skip = true;
      if (dof[j] < 1 || item[j] < 1 || item[j] > N) {
        set_error("a multipoint constraint on an external node or a dof out of range is not supported");
        return HECMW33_ERROR_UNSUPPORTED;
      }
    }
    if (skip) continue;
    for (int32_t j = index[i]; j < index[i + 1]; ++j) {
      const int64_t k = 3 * (int64_t)(item[j] - 1) + dof[j] - 1;
      const char want = j == index[i] ? 1 : 2;
      if (role[k] == 1 || (role[k] == 2 && want == 1)) {
        set_error("chained multipoint constraints (a slave dof that is another constraint's slave or master)");
        return HECMW33_ERROR_UNSUPPORTED;
      }
      role[k] = want;
    }
  }
  return 0;
}

static int32_t mpc_build(Context& c) {
  if (c.mpc_built) return 0;
  if (c.mpc_N != c.N || c.NP != c.N) {
    set_error("the multipoint constraints were given for another system (or NP > N)");
    return HECMW33_ERROR_UNSUPPORTED;
  }
  const int32_t N = c.N, n = c.mpc_n;
  const int32_t* index = c.h_mpc_index.data();
  const int32_t* item = c.h_mpc_item.data();
  const int32_t* dof = c.h_mpc_dof.data();
  const double* val = c.h_mpc_val.data();
  const int64_t n3 = 3 * (int64_t)N;
  std::vector<int32_t> slave_row(n3, -1), mptr(1, 0), mdof, tcnt(n3 + 1, 0);
  std::vector<double> mval;
  std::vector<int32_t> kept;
  for (int32_t i = 0; i < n; ++i) {
    bool skip = false;
    for (int32_t j = index[i]; j < index[i + 1]; ++j) skip = skip || dof[j] > 3;
    if (skip) continue;
    const int32_t r = (int32_t)kept.size();
    kept.push_back(i);
    slave_row[3 * (int64_t)(item[index[i]] - 1) + dof[index[i]] - 1] = r;
    for (int32_t j = index[i] + 1; j < index[i + 1]; ++j) {
      const int32_t k = 3 * (item[j] - 1) + dof[j] - 1;
      mdof.push_back(k);
      mval.push_back(val[j]);
      ++tcnt[k + 1];
    }
    mptr.push_back((int32_t)mdof.size());
  }
  for (int64_t k = 0; k < n3; ++k) tcnt[k + 1] += tcnt[k];
  std::vector<int32_t> tslave(std::max<int32_t>(tcnt[n3], 1)), fill(tcnt.begin(), tcnt.end() - 1);
  std::vector<double> tval(std::max<int32_t>(tcnt[n3], 1));
  for (size_t r = 0; r < kept.size(); ++r) {
    const int32_t i = kept[r];
    const int32_t ks = 3 * (item[index[i]] - 1) + dof[index[i]] - 1;
    for (int32_t j = index[i] + 1; j < index[i + 1]; ++j) {
      const int32_t k = 3 * (item[j] - 1) + dof[j] - 1;
      tslave[fill[k]] = ks;
      tval[fill[k]++] = val[j];
    }
  }
  const size_t nk = kept.size(), nm = std::max<size_t>(mdof.size(), 1), nt = tslave.size();
  if (c.mpc_slave.ensure(sizeof(int32_t) * n3) || c.mpc_mptr.ensure(sizeof(int32_t) * (nk + 1)) ||
      c.mpc_mdof.ensure(sizeof(int32_t) * nm) || c.mpc_mval.ensure(sizeof(double) * nm) ||
      c.mpc_tptr.ensure(sizeof(int32_t) * (n3 + 1)) || c.mpc_tslave.ensure(sizeof(int32_t) * nt) ||
      c.mpc_tval.ensure(sizeof(double) * nt) || c.mpc_w1.ensure(sizeof(double) * 3 * (size_t)c.NP) ||
      c.mpc_w2.ensure(sizeof(double) * 3 * (size_t)c.NP))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.mpc_slave.p, slave_row.data(), sizeof(int32_t) * n3, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.mpc_mptr.p, mptr.data(), sizeof(int32_t) * (nk + 1), hipMemcpyHostToDevice, c.stream));
  if (!mdof.empty()) {
    HIPC(hipMemcpyAsync(c.mpc_mdof.p, mdof.data(), sizeof(int32_t) * mdof.size(), hipMemcpyHostToDevice, c.stream));
    HIPC(hipMemcpyAsync(c.mpc_mval.p, mval.data(), sizeof(double) * mval.size(), hipMemcpyHostToDevice, c.stream));
  }
  HIPC(hipMemcpyAsync(c.mpc_tptr.p, tcnt.data(), sizeof(int32_t) * (n3 + 1), hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.mpc_tslave.p, tslave.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.mpc_tval.p, tval.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  MpcOp& m = c.mpc;
  m.n3 = (int32_t)n3;
  m.slave_row = c.mpc_slave.as<int32_t>();
  m.mptr = c.mpc_mptr.as<int32_t>();
  m.mdof = c.mpc_mdof.as<int32_t>();
  m.mval = c.mpc_mval.as<double>();
  m.tptr = c.mpc_tptr.as<int32_t>();
  m.tslave = c.mpc_tslave.as<int32_t>();
  m.tval = c.mpc_tval.as<double>();
  m.w1 = c.mpc_w1.as<double>();
  m.w2 = c.mpc_w2.as<double>();
  c.mpc_built = true;
  return 0;
}

static int32_t set_mpc(Context& c, int32_t N, int32_t n, const int32_t* index, const int32_t* item,
                       const int32_t* dof, const double* val) {
  if (n > 0 && !val) return HECMW33_ERROR_ARGS;
  if (int32_t e = mpc_check(N, n, index, item, dof)) return e;
  c.mpc_n = n;
  c.mpc_N = N;
  c.mpc_built = false;
  c.mpc_use = false;
  if (n == 0) return 0;
  const int32_t nt = index[n];
  c.h_mpc_index.assign(index, index + n + 1);
  c.h_mpc_item.assign(item, item + nt);
  c.h_mpc_dof.assign(dof, dof + nt);
  c.h_mpc_val.assign(val, val + nt);
  return 0;
}

// Eigen-decomposition of a symmetric n x n matrix (cyclic Jacobi, FP64):
// a is destroyed, lam[k] and the columns v[.][k] are its eigenpairs, unsorted.
static void jacobi_eigen(int n, double* a, double* lam, double* v) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i * n + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += a[i * n + i] * a[i * n + i];
      for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (apq == 0.0) continue;
        const double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = cs * akp - sn * akq;
          a[k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = cs * apk - sn * aqk;
          a[q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = cs * vkp - sn * vkq;
          v[k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) lam[i] = a[i * n + i];
}

// Default sample count and FCG iterations per sample of the learned coarse
// modes (Iarray 45 and 46 override them), deflation form.  8 x 8: on the
// 160^3 lognormal cube 108 iterations in 0.64 s against 107 in 0.68 s with
// 8 x 10 and 100 in 1.01 s with 8 x 20; on the 120^3 inclusion cube 2111
// iterations in 3.69 s against 2085 in 3.67 s with 8 x 10; on the small
// two-phase cubes the gain holds (24^3: 62 against 57, 32^3: 369 against
// 364), while 8 x 6 loses it (105, 452) and fewer samples lose more
// iterations than the harvest saves (6 x 10 on the inclusion cube: 2201).
constexpr int32_t I_SOLVER_OPT5 = 45, I_SOLVER_OPT6 = 46;
constexpr int32_t kModeSamples = 8, kModeIters = 8;
// default growth of the FP64 |r| over its smallest value at a replacement that drops the learned step
constexpr double kLearnedDropGrowth = 1000.0;
constexpr int32_t kModesPerAgg = 6;

// Learned coarse modes: fit, per aggregate, a linear autoencoder with six
// decoder columns to slow errors the solver generates on this system, and
// keep the decoders as six extra coarse modes beside the rigid-body ones.
//  * Samples (the solver's own (residual, A^-1 residual) pairs): for seeded
//    random v_s, b_s = A v_s, and the flexible CG with the rigid-body
//    two-level step run for a few iterations from x = 0; e_s = v_s - x_s =
//    A^-1 r_s is what that solver has not yet removed.
//  * Per aggregate: the part R of the samples beyond the rigid-body modes (a
//    6 x 6 least-squares fit per sample), its Gram matrix R^T R (BF16 MFMA,
//    FP32 accumulation), whose top eigenvectors V give the autoencoder's
//    optimum; the decoder columns R V lambda^-1/2, scaled to the norm of a
//    translation mode, become the modes.  Eigenvalues below 1e-6 of the
//    largest leave their mode at zero (a unit pivot of E).
// The caller's b and x in vb/vx are kept; harvest and fit times go to
// modes_harvest_s / modes_fit_s.
static int32_t learned_modes(Context& c, const int32_t* iarray, const double* rarray, int32_t o3) {
  const int32_t N = c.N, na = c.coarse.n_agg;
  const int ns = std::max(1, std::min(16, IA(iarray, I_SOLVER_OPT5) > 0 ? IA(iarray, I_SOLVER_OPT5) : kModeSamples));
  const int nu = std::max(1, IA(iarray, I_SOLVER_OPT6) > 0 ? IA(iarray, I_SOLVER_OPT6) : kModeIters);
  const int64_t n3 = 3 * (int64_t)N;
  const size_t vbytes = sizeof(double) * 3 * (size_t)c.NP;
  const double t0 = now_s();
  if (c.ml_S.ensure(sizeof(float) * 16 * (size_t)n3) || c.ml_v.ensure(vbytes) || c.ml_bsave.ensure(vbytes) ||
      c.ml_xsave.ensure(vbytes) || c.ml_rhs.ensure(sizeof(double) * 96 * (size_t)na) ||
      c.ml_ztz.ensure(sizeof(double) * 36 * (size_t)na) || c.ml_c.ensure(sizeof(double) * 96 * (size_t)na) ||
      c.ml_G.ensure(sizeof(float) * 256 * (size_t)na) || c.ml_V.ensure(sizeof(float) * 96 * (size_t)na) ||
      c.cg_xm.ensure(sizeof(float) * 18 * (size_t)std::max(N, 1)))
    return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(c.ml_bsave.p, c.vb.p, vbytes, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.ml_xsave.p, c.vx.p, vbytes, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipMemsetAsync(c.ml_S.p, 0, sizeof(float) * 16 * (size_t)n3, c.stream));
  HIPC(hipMemsetAsync(c.ml_v.p, 0, vbytes, c.stream));
  std::vector<int32_t> ia(iarray, iarray + 100);
  std::vector<double> ra(rarray, rarray + 100);
  ia[I_ITER - 1] = nu;
  ia[I_SOLVER_OPT - 1] = kOptMixedFcg;
  ia[I_SOLVER_OPT3 - 1] = o3;
  ia[I_SOLVER_OPT4 - 1] = IA(iarray, I_SOLVER_OPT4) == 2 ? 2 : 0;  // the solve's own form of the step
  ia[I_FLAG_NUMFACT - 1] = 0;
  ia[I_FLAG_SYMBFACT - 1] = 0;
  ia[I_ITERLOG - 1] = 0;
  ia[I_TIMELOG - 1] = 0;
  int32_t e = 0, got = 0;
  for (int s = 0; s < ns; ++s) {
    // v_s (external entries stay zero), b_s = A v_s, x = 0
    modes_random(n3, 0x5EED0000ull + (uint64_t)s, c.ml_v.as<double>(), c.stream);
    launch_spmv(c.A.view, c.ml_v.as<double>(), c.vb.as<double>(), c.stream);
    HIPC(hipMemsetAsync(c.vx.p, 0, vbytes, c.stream));
    hecmw33_out o{};
    std::vector<int32_t> ias(ia);
    c.x_zero = true;
    e = krylov_core(c, ias.data(), ra.data(), &o);
    c.x_zero = false;
    // a failed sample solve leaves the modes unfitted (the caller keeps the rigid-body modes)
    if (e || (o.error != 0 && o.error != HECMW33_ERROR_NOCONV_MAXIT)) break;
    modes_sample(n3, c.ml_v.as<double>(), c.vx.as<double>(), c.ml_S.as<float>(), s, c.stream);
    ++got;
  }
  HIPC(hipMemcpyAsync(c.vb.p, c.ml_bsave.p, vbytes, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.vx.p, c.ml_xsave.p, vbytes, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(c.stream));
  if (e) return e;
  c.modes_harvest_s = now_s() - t0;
  c.modes_samples = got;
  c.modes_iters = nu;
  if (got != ns) return 0;
  const double t1 = now_s();
  const CoarseOp& op = c.coarse;
  // the rigid-body part of every sample, aggregate by aggregate
  modes_rigid(na, ns, op.aptr, op.anodes, op.dv, c.ml_S.as<float>(), c.ml_rhs.as<double>(), c.ml_ztz.as<double>(),
              c.stream);
  std::vector<double> rhs(96 * (size_t)na), ztz(36 * (size_t)na), coef(96 * (size_t)na, 0.0);
  HIPC(hipMemcpyAsync(rhs.data(), c.ml_rhs.p, sizeof(double) * 96 * na, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipMemcpyAsync(ztz.data(), c.ml_ztz.p, sizeof(double) * 36 * na, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
#pragma omp parallel for schedule(static)
  for (int32_t a = 0; a < na; ++a) {
    // least squares on the rigid-body modes through the eigenpairs of Z^T Z (an
    // aggregate whose rotations vanish has zero eigenvalues there: those
    // directions get no coefficient)
    double m[36], lam[6], v[36];
    std::copy(&ztz[36 * (size_t)a], &ztz[36 * (size_t)a] + 36, m);
    jacobi_eigen(6, m, lam, v);
    double lmax = 0.0;
    for (int k = 0; k < 6; ++k) lmax = std::max(lmax, lam[k]);
    for (int s = 0; s < ns; ++s) {
      const double* b = &rhs[((size_t)a * 16 + s) * 6];
      double* x = &coef[((size_t)a * 16 + s) * 6];
      for (int k = 0; k < 6; ++k) {
        if (lam[k] <= 1e-12 * lmax) continue;
        double pk = 0.0;
        for (int q = 0; q < 6; ++q) pk += v[q * 6 + k] * b[q];
        pk /= lam[k];
        for (int q = 0; q < 6; ++q) x[q] += v[q * 6 + k] * pk;
      }
    }
  }
  HIPC(hipMemcpyAsync(c.ml_c.p, coef.data(), sizeof(double) * 96 * na, hipMemcpyHostToDevice, c.stream));
  modes_resid(N, op.agg, op.dv, c.ml_c.as<double>(), c.ml_S.as<float>(), c.stream);
  modes_gram(na, op.aptr, op.anodes, c.ml_S.as<float>(), c.ml_G.as<float>(), c.stream);
  std::vector<float> G(256 * (size_t)na), V(96 * (size_t)na, 0.f);
  std::vector<int32_t> acnt(na + 1);
  HIPC(hipMemcpyAsync(G.data(), c.ml_G.p, sizeof(float) * 256 * na, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipMemcpyAsync(acnt.data(), op.aptr, sizeof(int32_t) * (na + 1), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  int32_t kept = 0;
#pragma omp parallel for schedule(static) reduction(+ : kept)
  for (int32_t a = 0; a < na; ++a) {
    double m[256], lam[16], v[256];
    for (int k = 0; k < 256; ++k) m[k] = G[256 * (size_t)a + k];
    jacobi_eigen(16, m, lam, v);
    int order[16];
    for (int k = 0; k < 16; ++k) order[k] = k;
    std::sort(order, order + 16, [&](int x, int y) { return lam[x] > lam[y]; });
    const double lmax = lam[order[0]];
    const double scale = std::sqrt((double)std::max(1, acnt[a + 1] - acnt[a]));
    for (int j = 0; j < kModesPerAgg; ++j) {
      const int k = order[j];
      if (!(lmax > 0.0) || lam[k] <= 1e-6 * lmax) continue;
      ++kept;
      const double f = scale / std::sqrt(lam[k]);
      for (int s = 0; s < 16; ++s) V[96 * (size_t)a + 6 * s + j] = (float)(v[s * 16 + k] * f);
    }
  }
  HIPC(hipMemcpyAsync(c.ml_V.p, V.data(), sizeof(float) * 96 * na, hipMemcpyHostToDevice, c.stream));
  modes_make(N, op.agg, c.ml_S.as<float>(), c.ml_V.as<float>(), c.cg_xm.as<float>(), c.stream);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(c.stream));
  c.modes_kept = kept;
  c.modes_fit_s = now_s() - t1;
  c.modes_valid = true;
  return 0;
}

static int32_t krylov_core(Context& c, int32_t* iarray, double* rarray, hecmw33_out* out) {
  const double t0 = now_s();
  const int32_t method = IA(iarray, I_METHOD);
  const int32_t pc = IA(iarray, I_PRECOND);
  const int32_t iterpremax = IA(iarray, I_ITERPREMAX);
  const int32_t maxit = IA(iarray, I_ITER);
  const double tol = RA(rarray, R_RESID);
  const bool timelog = IA(iarray, I_TIMELOG) >= 1;
  if (method != 1 && method != 2) {
    set_error("METHOD not provided by the GPU library (1: CG, 2: BiCGSTAB)");
    return HECMW33_ERROR_UNSUPPORTED;
  }
  const int32_t opt = IA(iarray, I_SOLVER_OPT);
  const bool learned = opt == kOptLearnedFcg;
  const bool mixed = opt == kOptMixedFcg || learned;
  // MPC-CG: every product is T^T A T while Iarray 83 is set (hecmw_matvec_33,
  // hecmw_solver_las_33.F90:50-53)
  c.mpc_use = false;
  if (c.mpc_n > 0 && IA(iarray, I_FLAG_MPCMATVEC) != 0) {
    if (mixed) {
      set_error("the MPC-CG product is provided for the FP64 CG/BiCGSTAB only");
      return HECMW33_ERROR_UNSUPPORTED;
    }
    if (int32_t e = mpc_build(c)) return e;
    c.mpc_use = true;
  }
  if (mixed && (method != 1 || iterpremax > 1)) {
    set_error("SOLVER_OPT(1)=1/2 (flexible CG, FP32 operators) needs METHOD=1 and ITERPREMAX<=1");
    return HECMW33_ERROR_UNSUPPORTED;
  }
  if (learned && (iterpremax < 1 || (IA(iarray, I_SOLVER_OPT3) > 0 && IA(iarray, I_SOLVER_OPT4) == 1))) {
    set_error("SOLVER_OPT(1)=2 (learned coarse modes) needs ITERPREMAX >= 1 and the symmetric two-level step "
              "(SOLVER_OPT(4) = 0 or 2)");
    return HECMW33_ERROR_UNSUPPORTED;
  }
  if (int32_t e = precond_setup(c, iarray, rarray)) return e;
  const int32_t kind = pc_kind(pc, iterpremax);
  c.ssor16 = IA(iarray, I_SOLVER_OPT7) != 1;
  if (mixed)
    if (int32_t e = setup_f32(c, kind, c.ssor16)) return e;
  bool spmv64 = false;
  const int32_t o3 = learned && IA(iarray, I_SOLVER_OPT3) <= 0 ? 1 : IA(iarray, I_SOLVER_OPT3);
  if (mixed) {
    const int32_t o2 = IA(iarray, I_SOLVER_OPT2);
    if (o2 == 2) {
      spmv64 = true;
    } else if (o2 != 1 && o3 <= 0) {
      // without the two-level correction the FCG runs long enough on
      // high-contrast systems for FP32-rounded values to lose conjugacy; with
      // it FP32 values win (120^3 inclusions, contrast 1e4: 2409 iterations,
      // 4.16 s against 2241 iterations, 4.57 s with FP64 values)
      if (int32_t e = diag_contrast(c, &c.contrast)) return e;
      spmv64 = c.contrast > kContrastFp64Spmv;
    }
  }
  c.last_spmv64 = spmv64;
  c.coarse_use = false;
  c.last_coarse = 0;
  c.last_modes_nm = 0;
  c.coarse_sym_use = false;
  c.coarse_defl = false;
  if (mixed && o3 > 0) {
    // the learned solve first fits its modes with solves on the rigid-body
    // two-level step (learned_modes), then takes the twelve-mode space; if
    // the fit or its E fails, the rigid-body modes stay
    int32_t nm = 6;
    if (learned) {
      if (int32_t e = coarse_setup(c, o3, 6)) return e;
      if (c.coarse_ok && !c.modes_valid)
        if (int32_t e = learned_modes(c, iarray, rarray, o3)) return e;
      if (c.modes_valid) nm = 12;
    }
    if (int32_t e = coarse_setup(c, o3, nm)) return e;
    if (nm == 12 && !c.coarse_ok) {
      nm = 6;
      if (int32_t e = coarse_setup(c, o3, 6)) return e;
    }
    c.coarse_use = c.coarse_ok;
    c.last_coarse = c.coarse_ok ? c.coarse.n_agg : -1;
    if (c.coarse_use && IA(iarray, I_SOLVER_OPT4) != 1) {
      const size_t vb32 = sizeof(float) * 3 * (size_t)c.NP;
      for (DMem* m : {&c.vr1_32, &c.vt32, &c.vw32, &c.vaz32})
        if (m->ensure(vb32)) return HECMW33_ERROR_DEVICE;
      if (!c.coarse_w_valid || c.coarse_w_nm != nm) {
        const size_t nw = (size_t)std::max<int64_t>(c.coarse_nw, 1);
        const size_t nblk = (size_t)std::max<int32_t>((c.N + 255) / 256, 1);
        if (c.cg_AZ.ensure(sizeof(float) * 3 * nm * nw) || c.cg_aggW.ensure(sizeof(int32_t) * nw) ||
            c.cg_sc.ensure(sizeof(int2) * (size_t)std::max<int32_t>(c.N, 1)) || c.cg_kept.ensure(sizeof(int32_t) * nblk))
          return HECMW33_ERROR_DEVICE;
        coarse_build_W(c.A.view, c.coarse, c.N, c.cg_ncnt.as<int32_t>(), c.cg_nagg.as<int32_t>(), c.cg_sc.as<int2>(),
                       c.cg_aggW.as<int32_t>(), c.cg_AZ.as<float>(), c.cg_kept.as<int32_t>(), c.stream);
        HIPC(hipGetLastError());
        c.coarse_w_valid = true;
        c.coarse_w_nm = nm;
      }
      c.coarse_sym_use = true;
      c.coarse_defl = IA(iarray, I_SOLVER_OPT4) == 2;
    }
    c.last_modes_nm = nm;
  }
  const int64_t n3 = 3 * (int64_t)c.N;
  const size_t vbytes = sizeof(double) * 3 * (size_t)c.NP;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  if (c.hist.ensure(sizeof(double) * std::max(1, maxit))) return HECMW33_ERROR_DEVICE;
  if (iterpremax > 1 && (c.vzp.ensure(vbytes) || c.vzt.ensure(vbytes))) return HECMW33_ERROR_DEVICE;
  if (method == 2) {
    for (DMem* m : {&c.vrt, &c.vs, &c.vt, &c.vv, &c.vpt, &c.vst})
      if (m->ensure(vbytes)) return HECMW33_ERROR_DEVICE;
  }
  // r0 = b - A x0; bnrm2 = b.b  (hecmw_solver_CG.f90:131-141)
  if (mixed && c.x_zero) {
    // x0 = 0 (the learned modes' sample solves): r0 = b exactly, without the FP64 product
    launch_to_f32(n3, c.vb.as<double>(), c.vr32.as<float>(), c.stream);
    launch_dot_store(n3, c.vb.as<double>(), c.vb.as<double>(), &st->rr_ref, st, part, c.stream);
  } else if (mixed) {
    fcg_init_resid(c.A.view, c.vx.as<double>(), c.vb.as<double>(), c.vr32.as<float>(), &st->rr_ref, st, part,
                   c.stream);
  } else {
    launch_matresid(c.A.view, c.vx.as<double>(), c.vb.as<double>(), c.vr.as<double>(), c.stream, c.mpc_op());
  }
  launch_dot_store(n3, c.vb.as<double>(), c.vb.as<double>(), &st->bnrm2, st, part, c.stream);
  // work vectors start at zero (WW = 0), including the external entries the
  // SpMV reads; nothing writes those afterwards
  HIPC(hipMemsetAsync(c.vp.p, 0, vbytes, c.stream));
  if (mixed)
    for (DMem* m : {&c.vz32, &c.vp32, &c.vq32}) HIPC(hipMemsetAsync(m->p, 0, vbytes / 2, c.stream));
  if (method == 2) {
    HIPC(hipMemcpyAsync(c.vrt.p, c.vr.p, vbytes, hipMemcpyDeviceToDevice, c.stream));  // r~ = r0
    for (DMem* m : {&c.vv, &c.vpt, &c.vst}) HIPC(hipMemsetAsync(m->p, 0, vbytes, c.stream));
  }
  HIPC(hipGetLastError());
  KrylovState h{};
  HIPC(hipMemcpyAsync(&c.h_st[0], st, sizeof(KrylovState), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  const double bnrm2 = c.h_st[0].bnrm2;
  out->method = method;
  out->ncolor = kind == 1 ? c.ncolor : 0;
  if (bnrm2 == 0.0 || maxit < 1) {
    // zero-trip DO loop: ITER reads 1 (hecmw_solver_CG.f90:135-141,164)
    if (bnrm2 == 0.0) HIPC(hipMemsetAsync(c.vx.p, 0, vbytes, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    out->iter = 1;
    out->error = 0;
    out->resid = bnrm2 == 0.0 ? 0.0 : tol;
    out->Tset = now_s() - t0;
    out->Tsol = 0.0;
    c.last_iter = 0;
    return 0;
  }
  h = c.h_st[0];
  h.tol = tol;
  h.rho = h.rho1 = h.alpha = h.beta = h.c1 = h.dnrm2 = h.resid = 0.0;
  h.omega = h.c2 = h.ts = h.tt = 0.0;
  h.iter = 1;
  h.maxit = maxit;
  h.done = 0;
  h.error = 0;
  h.n_indef = 0;
  h.recheck = 0;
  h.nrec = method == 1 ? 50 : 100;
  h.zq = h.pq = h.rp = 0.0;
  const double delta = RA(rarray, R_SOLVER_OPT) > 0.0 && RA(rarray, R_SOLVER_OPT) < 1.0 ? RA(rarray, R_SOLVER_OPT)
                                                                                         : kFcgDelta;
  h.delta2 = delta * delta;
  // Rarray 42: the growth of the FP64 |r| over its smallest value at a replacement that drops
  // the learned step
  const double grow = RA(rarray, R_SOLVER_OPT2) >= 1.0 ? RA(rarray, R_SOLVER_OPT2) : kLearnedDropGrowth;
  h.drop2 = grow * grow;
  h.rr_min = h.rr_ref;
  h.reliable = 0;
  h.nreliable = 0;
  // a learned solve whose two-level step was built keeps it unless its FP64
  // residual grows between replacements (k_tail_fcg_rel)
  h.learned_off = learned && c.coarse_use && c.last_modes_nm == 12 ? 0 : 1;
  h.coarse_off = 0;
  h.fallback_iter = 0;
  h.edrop = 0.0;
  h.gap = 1.0;
  c.h_st[1] = h;
  HIPC(hipMemcpyAsync(st, &c.h_st[1], sizeof(KrylovState), hipMemcpyHostToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  const double t1 = now_s();
  out->Tset = t1 - t0;
  // iterations: enqueue in chunks and poll the previous chunk's state while
  // the current one runs; kernels after the exit return at once
  const int chunk = 8;
  const int nev = method == 1 ? 4 : 8;
  hipEvent_t ev[2];
  HIPC(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int32_t enq = 0, nchunk = 0;
  bool finished = false;
  std::vector<hipEvent_t> itev(nev);
  while (enq < maxit && !finished) {
    const int k = std::min(chunk, maxit - enq);
    for (int i = 0; i < k; ++i) {
      // TIMELOG times the first iteration of each chunk: an event between two
      // kernels costs a few microseconds of idle GPU, so one iteration in
      // `chunk` carries them and the totals are scaled to all iterations
      hipEvent_t* e = nullptr;
      if (timelog && i == 0) {
        for (int j = 0; j < nev; ++j) itev[j] = c.event((size_t)nev * nchunk + j);
        e = itev.data();
      }
      if (c.coarse_sym_use)
        enqueue_fcg_sym_iteration(c, kind, spmv64, e);
      else if (mixed)
        enqueue_fcg_iteration(c, kind, spmv64, e);
      else if (method == 1)
        enqueue_cg_iteration(c, kind, iterpremax, e);
      else
        enqueue_bicg_iteration(c, kind, iterpremax, e);
    }
    enq += k;
    const int slot = nchunk & 1;
    HIPC(hipMemcpyAsync(&c.h_st[slot], st, sizeof(KrylovState), hipMemcpyDeviceToHost, c.stream));
    HIPC(hipEventRecord(ev[slot], c.stream));
    if (nchunk > 0) {
      HIPC(hipEventSynchronize(ev[slot ^ 1]));
      if (c.h_st[slot ^ 1].done) finished = true;
    }
    ++nchunk;
  }
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(&c.h_st[0], st, sizeof(KrylovState), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  h = c.h_st[0];
  out->Tsol = now_s() - t1;
  if (timelog) {
    // kernel time of the sampled iterations that ran (later ones returned at
    // once), scaled to the iterations that ran
    const int64_t ran = std::min<int64_t>(std::min(h.iter, maxit), enq);
    double tax = 0.0, tpc = 0.0;
    int64_t sampled = 0;
    for (int64_t ch = 0; ch < nchunk && ch * chunk < ran; ++ch, ++sampled) {
      for (int j = 0; j < nev; j += 4) {
        float a = 0.f, b = 0.f;
        HIPC(hipEventElapsedTime(&a, c.events[nev * ch + j], c.events[nev * ch + j + 1]));
        HIPC(hipEventElapsedTime(&b, c.events[nev * ch + j + 2], c.events[nev * ch + j + 3]));
        tpc += a;
        tax += b;
      }
    }
    const double scale = sampled > 0 ? (double)ran / (double)sampled : 0.0;
    out->time_precond = tpc * scale * 1e-3;
    out->time_Ax = tax * scale * 1e-3;
  }
  out->iter = h.iter;
  out->error = h.error;
  out->resid = h.resid;
  c.last_iter = std::min(h.iter, maxit);
  c.last_nreliable = mixed ? h.nreliable : 0;
  c.last_fallback_iter = learned ? h.fallback_iter : 0;
  c.last_learned_off = learned ? h.learned_off : 1;
  return 0;
}

static std::mutex g_mu;
static Context* g_ctx = nullptr;

static Context* default_ctx() {
  if (!g_ctx) g_ctx = new Context();
  return g_ctx;
}

// upload x0/b (host) into the context, run, copy x back
static int32_t krylov_host(Context& c, const hecmw33_sys* A, const double* B, double* X, int32_t* iarray,
                           double* rarray, hecmw33_out* out) {
  const double t0 = now_s();
  const size_t vb = sizeof(double) * 3 * (size_t)A->NP;
  HIPC(hipMemcpyAsync(c.vb.p, B, vb, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.vx.p, X, vb, hipMemcpyHostToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  out->t_upload += now_s() - t0;
  if (int32_t e = krylov_resident(c, iarray, rarray, out)) return e;
  HIPC(hipMemcpyAsync(X, c.vx.p, sizeof(double) * 3 * (size_t)A->N, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

static int32_t prepare(Context& c, const hecmw33_sys* A, bool force_symbolic, hecmw33_out* out) {
  if (int32_t e = check_sys(A)) return e;
  if (int32_t e = ctx_init(c)) return e;
  const double t0 = now_s();
  if (int32_t e = load_structure(c, A, force_symbolic)) return e;
  if (int32_t e = load_values(c, A)) return e;
  HIPC(hipStreamSynchronize(c.stream));
  if (out) out->t_upload += now_s() - t0;
  return 0;
}

static int32_t rel_resid_resident(Context& c, double* resid) {
  const int64_t n3 = 3 * (int64_t)c.N;
  KrylovState* st = c.st.as<KrylovState>();
  double* part = c.part.as<double>();
  launch_dot_store(n3, c.vb.as<double>(), c.vb.as<double>(), &st->ts, st, part, c.stream);
  launch_matresid(c.A.view, c.vx.as<double>(), c.vb.as<double>(), c.vr.as<double>(), c.stream, c.mpc_op());
  launch_dot_store(n3, c.vr.as<double>(), c.vr.as<double>(), &st->tt, st, part, c.stream);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(&c.h_st[0], st, sizeof(KrylovState), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  double bn = c.h_st[0].ts;
  if (bn == 0.0) bn = 1.0;
  *resid = std::sqrt(c.h_st[0].tt / bn);
  return 0;
}

}  // namespace hecmw33

using namespace hecmw33;

struct hecmw33_ctx {
  Context c;
};

extern "C" {

const char* hecmw33_last_error(void) { return g_last_error.c_str(); }

int32_t hecmw33_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int32_t hecmw33_solve_krylov(const hecmw33_sys* A, const double* B, double* X, int32_t iarray[100],
                             double rarray[100], hecmw33_out* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  hecmw33_out tmp{};
  if (!out) out = &tmp;
  *out = hecmw33_out{};
  Context& c = *default_ctx();
  if (int32_t e = prepare(c, A, IA(iarray, I_FLAG_SYMBFACT) == 1, out)) return e;
  if (int32_t e = krylov_host(c, A, B, X, iarray, rarray, out)) return e;
  return out->error;
}

int32_t hecmw33_solve(const hecmw33_sys* A, const double* B, double* X, int32_t iarray[100], double rarray[100],
                      hecmw33_out* out) {
  hecmw33_out tmp{};
  if (!out) out = &tmp;
  *out = hecmw33_out{};
  if (int32_t e = check_sys(A)) return e;
  const int32_t pc = IA(iarray, I_PRECOND), iterpremax = IA(iarray, I_ITERPREMAX);
  // hecmw_solve_check_zerodiag (hecmw_solver_Iterative.f90:256-283)
  for (int32_t i = 0; i < A->N; ++i)
    for (int j = 0; j < 3; ++j)
      if (std::fabs(A->D[9 * (int64_t)i + 4 * j]) == 0.0 && pc < 10 && iterpremax > 0) {
        set_error("zero component in a diagonal block");
        out->error = HECMW33_ERROR_ZERO_DIAG;
        return HECMW33_ERROR_ZERO_DIAG;
      }
  // hecmw_mat_recycle_precond_setting (hecmw_matrix_misc.F90:858-877)
  if (IA(iarray, I_FLAG_SYMBFACT) >= 1) {
    IA(iarray, I_FLAG_NUMFACT) = 1;
    IA(iarray, I_NRECYCLE) = 0;
  } else if (IA(iarray, I_FLAG_NUMFACT) > 1) {
    IA(iarray, I_NRECYCLE) = 0;
    IA(iarray, I_FLAG_NUMFACT) = 1;
  } else if (IA(iarray, I_FLAG_NUMFACT) == 1) {
    if (IA(iarray, I_NRECYCLE) < IA(iarray, I_MAXRECYCLE)) {
      IA(iarray, I_FLAG_NUMFACT) = 0;
      IA(iarray, I_NRECYCLE) += 1;
    } else {
      IA(iarray, I_NRECYCLE) = 0;
    }
  }
  double sigma = rarray[R_SIGMA_DIAG - 1];
  const bool auto_sigma = sigma < 0.0;
  if (auto_sigma) sigma = 1.0;
  int32_t method = IA(iarray, I_METHOD);
  const int32_t method2 = IA(iarray, I_METHOD2);
  std::lock_guard<std::mutex> lk(g_mu);
  Context& c = *default_ctx();
  if (int32_t e = prepare(c, A, IA(iarray, I_FLAG_SYMBFACT) == 1, out)) return e;
  int32_t error = 0;
  for (;;) {
    IA(iarray, I_FLAG_CONVERGED) = 0;
    IA(iarray, I_FLAG_DIVERGED) = 0;
    if (auto_sigma) rarray[R_SIGMA_DIAG - 1] = std::min(2.0, std::max(1.0, sigma));
    int32_t ia2[100];
    std::memcpy(ia2, iarray, sizeof(ia2));
    IA(ia2, I_METHOD) = method;
    // MPC-CG (hecmw_solver_Iterative.f90:79-84): T^T A T products for MPC_METHOD 2
    if (c.mpc_n > 0 && IA(iarray, I_MPC_METHOD) == 2) IA(ia2, I_FLAG_MPCMATVEC) = 1;
    if (int32_t e = krylov_host(c, A, B, X, ia2, rarray, out)) return e;
    IA(iarray, I_FLAG_SYMBFACT) = IA(ia2, I_FLAG_SYMBFACT);
    IA(iarray, I_FLAG_NUMFACT) = IA(ia2, I_FLAG_NUMFACT);
    error = out->error;
    if (error == HECMW33_ERROR_DIVERGE_PC || error == HECMW33_ERROR_DIVERGE_MAT || error == HECMW33_ERROR_DIVERGE_NAN) {
      IA(iarray, I_FLAG_DIVERGED) = 1;
      if (method == 1 && method2 > 1) {
        if (auto_sigma) sigma = 1.0;
        method = method2;
        continue;
      }
    }
    if (auto_sigma) rarray[R_SIGMA_DIAG - 1] = -1.0;
    break;
  }
  // hecmw_rel_resid_L2 still with the MPC product (the reference clears the flag after it)
  c.mpc_use = c.mpc_n > 0 && IA(iarray, I_MPC_METHOD) == 2 && c.mpc_built;
  double r2 = 0.0;
  const int32_t er = rel_resid_resident(c, &r2);
  c.mpc_use = false;
  if (er) return er;
  out->resid_true = r2;
  if (r2 < RA(rarray, R_RESID)) IA(iarray, I_FLAG_CONVERGED) = 1;
  return error;
}

int32_t hecmw33_matvec(const hecmw33_sys* A, const double* x, double* y) {
  std::lock_guard<std::mutex> lk(g_mu);
  Context& c = *default_ctx();
  if (int32_t e = prepare(c, A, false, nullptr)) return e;
  HIPC(hipMemcpyAsync(c.vx.p, x, sizeof(double) * 3 * (size_t)A->NP, hipMemcpyHostToDevice, c.stream));
  launch_spmv(c.A.view, c.vx.as<double>(), c.vq.as<double>(), c.stream);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(y, c.vq.p, sizeof(double) * 3 * (size_t)A->N, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

int32_t hecmw33_set_mpc(int32_t N, int32_t n_mpc, const int32_t* mpc_index, const int32_t* mpc_item,
                        const int32_t* mpc_dof, const double* mpc_val) {
  std::lock_guard<std::mutex> lk(g_mu);
  return set_mpc(*default_ctx(), N, n_mpc, mpc_index, mpc_item, mpc_dof, mpc_val);
}

int32_t hecmw33_matvec_mpc(const hecmw33_sys* A, const double* x, double* y) {
  std::lock_guard<std::mutex> lk(g_mu);
  Context& c = *default_ctx();
  if (int32_t e = prepare(c, A, false, nullptr)) return e;
  if (c.mpc_n == 0) {
    set_error("no multipoint constraints set (hecmw33_set_mpc)");
    return HECMW33_ERROR_ARGS;
  }
  if (int32_t e = mpc_build(c)) return e;
  HIPC(hipMemcpyAsync(c.vx.p, x, sizeof(double) * 3 * (size_t)A->NP, hipMemcpyHostToDevice, c.stream));
  launch_spmv_mpc(c.A.view, c.mpc, c.vx.as<double>(), c.vq.as<double>(), c.stream);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(y, c.vq.p, sizeof(double) * 3 * (size_t)A->N, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

int32_t hecmw33_precond_apply(const hecmw33_sys* A, const int32_t iarray[100], const double rarray[100],
                              const double* r, double* z) {
  std::lock_guard<std::mutex> lk(g_mu);
  Context& c = *default_ctx();
  if (int32_t e = prepare(c, A, false, nullptr)) return e;
  int32_t ia[100];
  std::memcpy(ia, iarray, sizeof(ia));
  IA(ia, I_FLAG_SYMBFACT) = 1;
  IA(ia, I_FLAG_NUMFACT) = 1;
  if (int32_t e = precond_setup(c, ia, rarray)) return e;
  const size_t vb = sizeof(double) * 3 * (size_t)A->NP;
  HIPC(hipMemcpyAsync(c.vr.p, r, vb, hipMemcpyHostToDevice, c.stream));
  const int32_t pc = IA(ia, I_PRECOND);
  if (IA(ia, I_ITERPREMAX) <= 0) {
    HIPC(hipMemcpyAsync(c.vz.p, c.vr.p, vb, hipMemcpyDeviceToDevice, c.stream));
  } else if (pc == 3) {
    launch_diag_apply(c.N, c.alu_diag.as<double>(), c.vr.as<double>(), c.vz.as<double>(), nullptr, c.stream);
  } else {
    ssor_apply(c.ssor_op(), c.vr.as<double>(), c.vz.as<double>(), nullptr, c.stream);
  }
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(z, c.vz.p, sizeof(double) * 3 * (size_t)A->N, hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return 0;
}

int32_t hecmw33_rel_resid_L2(const hecmw33_sys* A, const double* B, const double* X, double* resid) {
  // no MPC flag here: the plain product, as hecmw_rel_resid_L2 with Iarray 83 = 0
  std::lock_guard<std::mutex> lk(g_mu);
  Context& c = *default_ctx();
  c.mpc_use = false;
  if (int32_t e = prepare(c, A, false, nullptr)) return e;
  const size_t vb = sizeof(double) * 3 * (size_t)A->NP;
  HIPC(hipMemcpyAsync(c.vb.p, B, vb, hipMemcpyHostToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.vx.p, X, vb, hipMemcpyHostToDevice, c.stream));
  return rel_resid_resident(c, resid);
}

int32_t hecmw33_last_history(double* hist, int32_t cap) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx || !hist || cap < 0) return HECMW33_ERROR_ARGS;
  Context& c = *g_ctx;
  const int32_t n = std::min(cap, c.last_iter);
  if (n > 0) {
    HIPC(hipMemcpyAsync(hist, c.hist.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
  }
  return n;
}

int32_t hecmw33_ordering(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                         const int32_t* itemU, int32_t ncolor_in, int32_t* perm_rcm, int32_t* perm, int32_t* iperm,
                         int32_t* ncolor, int32_t* colorindex) {
  if (N < 1 || !indexL || !indexU || !perm_rcm || !perm || !iperm || !ncolor || !colorindex) return HECMW33_ERROR_ARGS;
  std::vector<int32_t> iperm_rcm(N);
  ordering_rcm(N, indexL, itemL, indexU, itemU, perm_rcm, iperm_rcm.data());
  *ncolor = ordering_mc(N, indexL, itemL, indexU, itemU, perm_rcm, ncolor_in, colorindex, perm, iperm);
  return 0;
}

// ---- context API -----------------------------------------------------------
hecmw33_ctx* hecmw33_ctx_create(void) {
  hecmw33_ctx* c = new hecmw33_ctx();
  if (ctx_init(c->c) != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

void hecmw33_ctx_destroy(hecmw33_ctx* ctx) { delete ctx; }

int32_t hecmw33_ctx_load(hecmw33_ctx* ctx, const hecmw33_sys* A, const int32_t iarray[100], const double rarray[100]) {
  if (!ctx) return HECMW33_ERROR_ARGS;
  if (int32_t e = prepare(ctx->c, A, true, nullptr)) return e;
  int32_t ia[100];
  std::memcpy(ia, iarray, sizeof(ia));
  IA(ia, I_FLAG_SYMBFACT) = 1;
  IA(ia, I_FLAG_NUMFACT) = 1;
  if (int32_t e = precond_setup(ctx->c, ia, rarray)) return e;
  // and the two-level correction's set-up when the load asks for it: the
  // aggregation (symbolic) and the first coarse inverse, which also brings
  // up the dense-solver library once
  const int32_t o3 = IA(iarray, I_SOLVER_OPT3);
  if (o3 > 0 && (int64_t)ctx->c.h_coord.size() == 3 * (int64_t)ctx->c.NP)
    if (int32_t e = coarse_setup(ctx->c, o3, 6)) return e;
  return 0;
}

int32_t hecmw33_ctx_solve_device(hecmw33_ctx* ctx, const double* B_dev, double* X_dev, const int32_t iarray[100],
                                 const double rarray[100], hecmw33_out* out) {
  if (!ctx || !ctx->c.have_struct) return HECMW33_ERROR_ARGS;
  hecmw33_out tmp{};
  if (!out) out = &tmp;
  *out = hecmw33_out{};
  Context& c = ctx->c;
  const size_t vb = sizeof(double) * 3 * (size_t)c.NP;
  HIPC(hipMemcpyAsync(c.vb.p, B_dev, vb, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipMemcpyAsync(c.vx.p, X_dev, vb, hipMemcpyDeviceToDevice, c.stream));
  int32_t ia[100];
  double ra[100];
  std::memcpy(ia, iarray, sizeof(ia));
  std::memcpy(ra, rarray, sizeof(ra));
  if (int32_t e = krylov_resident(c, ia, ra, out)) return e;
  HIPC(hipMemcpyAsync(X_dev, c.vx.p, vb, hipMemcpyDeviceToDevice, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  return out->error;
}

int32_t hecmw33_stream_read(const double* buf_dev, int64_t n, double* part_dev, int32_t reps, double* gbs) {
  if (!buf_dev || !part_dev || !gbs || n < 2 || (n & 1) || ((uintptr_t)buf_dev & 15) || reps < 1)
    return HECMW33_ERROR_ARGS;
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  double best = 0.0;
  for (const int grid : {1024, 2048, 4096, 8192, 65536}) {
    launch_stream_read(buf_dev, n, part_dev, grid, nullptr);
    HIPC(hipEventRecord(e0, nullptr));
    for (int i = 0; i < reps; ++i) launch_stream_read(buf_dev, n, part_dev, grid, nullptr);
    HIPC(hipEventRecord(e1, nullptr));
    HIPC(hipEventSynchronize(e1));
    float t = 0.f;
    HIPC(hipEventElapsedTime(&t, e0, e1));
    if (t > 0.f) best = std::max(best, 8.0 * (double)n * reps / ((double)t * 1e6));
  }
  HIPC(hipGetLastError());
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbs = best;
  return 0;
}

int32_t hecmw33_ctx_matvec_device(hecmw33_ctx* ctx, const double* x_dev, double* y_dev, int32_t count, double* ms) {
  if (!ctx || !ctx->c.have_struct || count < 1) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  HIPC(hipEventRecord(e0, c.stream));
  for (int i = 0; i < count; ++i) launch_spmv(c.A.view, x_dev, y_dev, c.stream);
  HIPC(hipEventRecord(e1, c.stream));
  HIPC(hipEventSynchronize(e1));
  float t = 0.f;
  HIPC(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = (double)t / count;
  return 0;
}

int32_t hecmw33_ctx_matvec_device_f32(hecmw33_ctx* ctx, const float* x_dev, float* y_dev, int32_t count, double* ms) {
  if (!ctx || !ctx->c.have_struct || count < 1) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if (!c.A.val32.p) {
    if (c.A.val32.ensure(sizeof(float) * 9 * (size_t)std::max<int64_t>(c.A.view.nslots, 1))) return HECMW33_ERROR_DEVICE;
    c.A.view32.val = c.A.val32.as<float>();
    launch_fill_values_a_f32(c.A.view32, c.A.src.as<int64_t>(), c.dD.as<double>(), c.dAL.as<double>(),
                           c.dAU.as<double>(), c.N, c.NPL, c.stream);
    HIPC(hipGetLastError());
  }
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  HIPC(hipEventRecord(e0, c.stream));
  for (int i = 0; i < count; ++i) launch_spmv_f32(c.A.view32, x_dev, y_dev, c.stream);
  HIPC(hipEventRecord(e1, c.stream));
  HIPC(hipEventSynchronize(e1));
  float t = 0.f;
  HIPC(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = (double)t / count;
  return 0;
}

// `count` back-to-back SSOR applications z = M^-1 r on device vectors with the
// context's multicolour SSOR set-up: FP64 (fp32 = 0), or its FP32 image on
// float vectors with the L'/U' values in BF16 (fp32 = 1, the flexible CG's
// default) or FP32 (fp32 = 2); the mean time per application in *ms (HIP
// events on the context's stream)
int32_t hecmw33_ctx_ssor_device(hecmw33_ctx* ctx, const void* r_dev, void* z_dev, int32_t fp32, int32_t count,
                                double* ms) {
  if (!ctx || !ctx->c.have_struct || !r_dev || !z_dev || count < 1 || fp32 < 0 || fp32 > 2) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if (!c.ssor_init) {
    set_error("hecmw33_ctx_ssor_device: the context has no SSOR set-up (load with PRECOND 1 or 2)");
    return HECMW33_ERROR_ARGS;
  }
  if (fp32)
    if (int32_t e = setup_f32(c, 1, fp32 == 1)) return e;
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  HIPC(hipEventRecord(e0, c.stream));
  for (int i = 0; i < count; ++i) {
    if (fp32)
      ssor_apply_f32(c.ssor_op32(fp32 == 1), static_cast<const float*>(r_dev), static_cast<float*>(z_dev), nullptr, c.stream);
    else
      ssor_apply(c.ssor_op(), static_cast<const double*>(r_dev), static_cast<double*>(z_dev), nullptr, c.stream);
  }
  HIPC(hipEventRecord(e1, c.stream));
  HIPC(hipEventSynchronize(e1));
  float t = 0.f;
  HIPC(hipEventElapsedTime(&t, e0, e1));
  HIPC(hipGetLastError());
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = (double)t / count;
  return 0;
}

int32_t hecmw33_ctx_history(hecmw33_ctx* ctx, double* hist, int32_t cap) {
  if (!ctx || !hist) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  const int32_t n = std::min(cap, c.last_iter);
  if (n > 0) {
    HIPC(hipMemcpyAsync(hist, c.hist.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
  }
  return n;
}

int32_t hecmw33_ctx_info(hecmw33_ctx* ctx, int64_t* info, int32_t n) {
  if (!ctx || !info) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  // A Z entries of the symmetric two-level step: kept after the last build, and symbolic
  int64_t az_kept = 0, az_sym = 0;
  if (c.coarse_w_valid && c.cg_kept.p && c.N > 0) {
    const int32_t nblk = (c.N + 255) / 256;
    std::vector<int32_t> kept(nblk);
    HIPC(hipMemcpyAsync(kept.data(), c.cg_kept.p, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    for (int32_t k : kept) az_kept += k;
    az_sym = c.coarse_nw;
  }
  const int64_t v[21] = {c.A.view.nslots,
                        (int64_t)c.N,
                        (int64_t)c.NPL,
                        (int64_t)c.NPU,
                        (int64_t)c.ncolor,
                        c.L.view.nslots + c.U.view.nslots,
                        (int64_t)c.last_nreliable,
                        (int64_t)c.last_fallback_iter,
                        (int64_t)c.last_learned_off,
                        (int64_t)c.last_spmv64,
                        (int64_t)(c.modes_valid ? c.modes_kept : 0),
                        (int64_t)c.last_coarse,
                        (int64_t)c.coarse_npair,
                        az_kept,
                        az_sym,
                        (int64_t)c.last_modes_nm,
                        (int64_t)c.modes_samples,
                        (int64_t)c.modes_iters,
                        (int64_t)std::llround(c.modes_harvest_s * 1e6),
                        (int64_t)std::llround(c.modes_fit_s * 1e6),
                        (int64_t)c.coarse.nc};
  for (int i = 0; i < n && i < 21; ++i) info[i] = v[i];
  return 0;
}

}  // extern "C"

// ---- learned aggregation ---------------------------------------------------------------------
// Training labels from the solver's own errors e = x* - x_k = A^-1 r_k.  Per
// stored edge k = (i, j) of the profile (the L items of every row, then the U
// items; both nodes internal) the axial strain of e along the edge,
// |(e_j - e_i).(x_j - x_i)| / |x_j - x_i|^2, divided by its median over the
// edges of the sample; worst[k] keeps the largest ratio over the samples.
// An edge whose e is zero at both ends (fixed degrees of freedom) or that
// leaves the internal nodes says nothing: +inf.  In a body whose stiffness
// varies by orders of magnitude the slow error strains the soft material and
// moves the stiff pieces rigidly, so their edges stay orders of magnitude
// below the median in every sample.
static void rigid_edge_ratios(int32_t N, const int32_t* iL, const int32_t* tL, const int32_t* iU, const int32_t* tU,
                              const double* X, const double* e, std::vector<float>& worst) {
  const int64_t nl = iL[N], ne = (int64_t)iL[N] + iU[N];
  const float inf = std::numeric_limits<float>::infinity();
  if ((int64_t)worst.size() != ne) worst.assign(ne, 0.0f);
  std::vector<double> eps(ne, -1.0);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < N; ++i)
    for (int part = 0; part < 2; ++part) {
      const int32_t* ix = part ? iU : iL;
      const int32_t* it = part ? tU : tL;
      for (int32_t q = ix[i]; q < ix[i + 1]; ++q) {
        const int32_t j = it[q] - 1;
        const int64_t k = part ? nl + q : q;
        if (j >= N) continue;
        double de[3], dx[3], num = 0.0, den = 0.0, mag = 0.0;
        for (int d = 0; d < 3; ++d) {
          de[d] = e[3 * (int64_t)j + d] - e[3 * (int64_t)i + d];
          dx[d] = X[3 * (int64_t)j + d] - X[3 * (int64_t)i + d];
          num += de[d] * dx[d];
          den += dx[d] * dx[d];
          mag += std::fabs(e[3 * (int64_t)j + d]) + std::fabs(e[3 * (int64_t)i + d]);
        }
        if (mag > 0.0 && den > 0.0) eps[k] = std::fabs(num) / den;
      }
    }
  std::vector<double> live;
  live.reserve(ne);
  for (double v : eps)
    if (v >= 0.0) live.push_back(v);
  double med = 0.0;
  if (!live.empty()) {
    std::nth_element(live.begin(), live.begin() + live.size() / 2, live.end());
    med = live[live.size() / 2];
  }
  for (int64_t k = 0; k < ne; ++k) {
    float r;
    if (eps[k] < 0.0)
      r = inf;
    else if (med > 0.0)
      r = (float)(eps[k] / med);
    else
      r = eps[k] > 0.0 ? inf : 0.0f;
    worst[k] = std::max(worst[k], r);
  }
}

// label 1 for both nodes of every edge whose worst ratio is below tau
static int32_t rigid_labels_from_ratios(int32_t N, const int32_t* iL, const int32_t* tL, const int32_t* iU,
                                        const int32_t* tU, const std::vector<float>& worst, double tau, float* t) {
  const int64_t nl = iL[N];
  std::fill(t, t + N, 0.0f);
  for (int32_t i = 0; i < N; ++i)
    for (int part = 0; part < 2; ++part) {
      const int32_t* ix = part ? iU : iL;
      const int32_t* it = part ? tU : tL;
      for (int32_t q = ix[i]; q < ix[i + 1]; ++q) {
        const int64_t k = part ? nl + q : q;
        if (k < (int64_t)worst.size() && worst[k] < tau) {
          t[i] = 1.0f;
          t[it[q] - 1] = 1.0f;
        }
      }
    }
  int32_t cnt = 0;
  for (int32_t i = 0; i < N; ++i) cnt += t[i] > 0.5f;
  return cnt;
}

static uint16_t to_bf16_bits(float f) {
  // round to nearest even (what the device's float -> __bf16 conversion does)
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16 | ((u & 0xffff) ? 0x40 : 0));
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

static int32_t net_init(hecmw33_net& N) {
  if (!N.stream) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      set_error("no HIP device visible");
      return HECMW33_ERROR_DEVICE;
    }
    HIPC(hipStreamCreateWithFlags(&N.stream, hipStreamNonBlocking));
  }
  const int w = net_weight_count();
  if (N.Wm.ensure(sizeof(float) * w) || N.Wb.ensure(2 * (size_t)w) || N.M.ensure(sizeof(float) * w) ||
      N.V.ensure(sizeof(float) * w) || N.dW.ensure(sizeof(float) * w) || N.dots.ensure(sizeof(double) * 8) ||
      N.dpart.ensure(sizeof(double) * 4 * kMaxPartials))
    return HECMW33_ERROR_DEVICE;
  return 0;
}

static int32_t net_upload(hecmw33_net& N, const float* w) {
  const int nw = net_weight_count();
  HIPC(hipMemcpyAsync(N.Wm.p, w, sizeof(float) * nw, hipMemcpyHostToDevice, N.stream));
  HIPC(hipMemsetAsync(N.M.p, 0, sizeof(float) * nw, N.stream));
  HIPC(hipMemsetAsync(N.V.p, 0, sizeof(float) * nw, N.stream));
  net_to_bf16(N.Wm.as<float>(), N.Wb.as<__bf16>(), N.stream);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(N.stream));
  N.step = 0;
  return 0;
}

// working buffers for n nodes
static int32_t net_work(hecmw33_net& N, int64_t n) {
  const size_t n3 = 3 * (size_t)n;
  if (N.H1.ensure(2 * 64 * (size_t)n) || N.H2.ensure(2 * 64 * (size_t)n) || N.z.ensure(sizeof(float) * n3) ||
      N.y.ensure(sizeof(float) * n3) || N.dY.ensure(sizeof(float) * n3) ||
      N.wpart.ensure(sizeof(float) * (size_t)net_bwd_parts(n) * net_weight_count()))
    return HECMW33_ERROR_DEVICE;
  return 0;
}

extern "C" {

hecmw33_net* hecmw33_net_create(uint64_t seed) {
  hecmw33_net* N = new (std::nothrow) hecmw33_net();
  if (!N) return nullptr;
  if (net_init(*N)) {
    delete N;
    return nullptr;
  }
  // He initialisation of W1 (12 live inputs) and W2, small W3 (columns 0..2)
  const int nw = net_weight_count();
  std::vector<float> w(nw, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto gauss = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
  for (int k = 0; k < 12; ++k)
    for (int h = 0; h < 64; ++h) w[k * 64 + h] = (float)(gauss() * std::sqrt(2.0 / 12.0));
  for (int k = 0; k < 64; ++k)
    for (int h = 0; h < 64; ++h) w[16 * 64 + k * 64 + h] = (float)(gauss() * std::sqrt(2.0 / 64.0));
  for (int k = 0; k < 64; ++k)
    for (int j = 0; j < 3; ++j) w[16 * 64 + 64 * 64 + k * 16 + j] = (float)(gauss() * 0.01);
  if (net_upload(*N, w.data())) {
    delete N;
    return nullptr;
  }
  return N;
}

void hecmw33_net_destroy(hecmw33_net* net) { delete net; }

int32_t hecmw33_net_weight_count(void) { return net_weight_count(); }

int32_t hecmw33_net_get_weights(hecmw33_net* net, float* w, int32_t n) {
  if (!net || !w || n < net_weight_count()) return HECMW33_ERROR_ARGS;
  HIPC(hipMemcpyAsync(w, net->Wm.p, sizeof(float) * net_weight_count(), hipMemcpyDeviceToHost, net->stream));
  HIPC(hipStreamSynchronize(net->stream));
  return 0;
}

int32_t hecmw33_net_set_weights(hecmw33_net* net, const float* w, int32_t n) {
  if (!net || !w || n < net_weight_count()) return HECMW33_ERROR_ARGS;
  return net_upload(*net, w);
}

int32_t hecmw33_net_eval(hecmw33_net* net, int64_t n, const float* X, const float* dY, float* Y, float* dW) {
  if (!net || n < 1 || !X || !Y) return HECMW33_ERROR_ARGS;
  hecmw33_net& N = *net;
  if (int32_t e = net_work(N, n)) return e;
  std::vector<uint16_t> xb(16 * (size_t)n);
  for (size_t i = 0; i < xb.size(); ++i) xb[i] = to_bf16_bits(X[i]);
  if (N.xb.ensure(2 * xb.size())) return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpyAsync(N.xb.p, xb.data(), 2 * xb.size(), hipMemcpyHostToDevice, N.stream));
  net_forward(n, N.Wb.as<__bf16>(), N.xb.as<__bf16>(), N.H1.as<__bf16>(), N.H2.as<__bf16>(), nullptr, N.z.as<float>(),
              N.stream);
  if (dY && dW) {
    HIPC(hipMemcpyAsync(N.dY.p, dY, sizeof(float) * 3 * n, hipMemcpyHostToDevice, N.stream));
    net_backward(n, N.Wb.as<__bf16>(), N.xb.as<__bf16>(), N.H1.as<__bf16>(), N.H2.as<__bf16>(), N.dY.as<float>(),
                 N.wpart.as<float>(), N.dW.as<float>(), N.stream);
    HIPC(hipMemcpyAsync(dW, N.dW.p, sizeof(float) * net_weight_count(), hipMemcpyDeviceToHost, N.stream));
  }
  HIPC(hipMemcpyAsync(Y, N.z.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, N.stream));
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(N.stream));
  return 0;
}

// Error samples e = A^-1 r (device, FP64, 3*NP) on a context with node
// coordinates: per stored edge of the profile, the largest ratio over the
// samples of its axial strain to the sample's median (rigid_edge_ratios);
// the labels follow at training.  Returns the context's sample count.
int32_t hecmw33_net_add_sample(hecmw33_net* net, hecmw33_ctx* ctx, const double* e_dev, double tau) {
  if (!net || !ctx || !e_dev || !ctx->c.have_struct || !(tau > 0.0)) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if ((int64_t)c.h_coord.size() != 3 * (int64_t)c.NP) {
    set_error("learned aggregation samples need the node coordinates of the system (hecmw33_ctx_set_coords)");
    return HECMW33_ERROR_ARGS;
  }
  hecmw33_net::Sample* S = nullptr;
  for (auto* p : net->samples)
    if (p->c == &c) S = p;
  if (!S) {
    S = new (std::nothrow) hecmw33_net::Sample();
    if (!S) return HECMW33_ERROR_DEVICE;
    S->c = &c;
    net->samples.push_back(S);
  }
  std::vector<double> e(3 * (size_t)c.N);
  HIPC(hipMemcpyAsync(e.data(), e_dev, sizeof(double) * e.size(), hipMemcpyDeviceToHost, c.stream));
  HIPC(hipStreamSynchronize(c.stream));
  rigid_edge_ratios(c.N, c.h_indexL.data(), c.h_itemL.data(), c.h_indexU.data(), c.h_itemU.data(), c.h_coord.data(),
                    e.data(), S->worst);
  S->nsamp += 1;
  S->tau = tau;
  S->t_valid = false;
  return S->nsamp;
}

int32_t hecmw33_net_clear_samples(hecmw33_net* net) {
  if (!net) return HECMW33_ERROR_ARGS;
  net->clear_samples();
  return 0;
}

// labels of a sample's context on the device (once per set of samples)
static int32_t sample_labels(hecmw33_net::Sample& S) {
  if (S.t_valid) return 0;
  Context& c = *S.c;
  std::vector<float> t(c.N);
  rigid_labels_from_ratios(c.N, c.h_indexL.data(), c.h_itemL.data(), c.h_indexU.data(), c.h_itemU.data(), S.worst,
                           S.tau, t.data());
  if (S.t.ensure(sizeof(float) * (size_t)c.N)) return HECMW33_ERROR_DEVICE;
  HIPC(hipMemcpy(S.t.p, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
  S.t_valid = true;
  return 0;
}

// loss and accuracy of the current weights on one sample (forward on the context's features)
static int32_t sample_bce(hecmw33_net& N, hecmw33_net::Sample& S, bool keep, double* out) {
  Context& c = *S.c;
  if (int32_t e = features_setup(c)) return e;
  HIPC(hipStreamSynchronize(c.stream));
  if (int32_t e = net_work(N, c.N)) return e;
  net_forward(c.N, N.Wb.as<__bf16>(), c.vfeat.as<__bf16>(), keep ? N.H1.as<__bf16>() : nullptr,
              keep ? N.H2.as<__bf16>() : nullptr, nullptr, N.y.as<float>(), N.stream);
  net_bce(c.N, N.y.as<float>(), S.t.as<float>(), N.dY.as<float>(), N.dpart.as<double>(), N.dots.as<double>(),
          N.stream);
  HIPC(hipMemcpyAsync(out, N.dots.p, 2 * sizeof(double), hipMemcpyDeviceToHost, N.stream));
  return 0;
}

// Adam on the mean logistic loss of the network's output 0 against the labels,
// one sample (context) per step, in a seeded shuffled order every epoch.
// loss[0] = mean loss of the weights before training, loss[1+e] = mean loss
// over epoch e (before each step's update).
int32_t hecmw33_net_train(hecmw33_net* net, int32_t epochs, double lr, uint64_t seed, double* loss) {
  if (!net || epochs < 0 || net->samples.empty()) return HECMW33_ERROR_ARGS;
  hecmw33_net& N = *net;
  const size_t ns = N.samples.size();
  int64_t nmax = 0;
  for (auto* S : N.samples) {
    if (int32_t e = sample_labels(*S)) return e;
    nmax = std::max<int64_t>(nmax, S->c->N);
  }
  if (int32_t e = net_work(N, nmax)) return e;  // no reallocation inside the loop
  std::vector<double> d(2 * ns);
  double base = 0.0;
  for (size_t k = 0; k < ns; ++k)
    if (int32_t e = sample_bce(N, *N.samples[k], false, &d[2 * k])) return e;
  HIPC(hipStreamSynchronize(N.stream));
  for (size_t k = 0; k < ns; ++k) base += d[2 * k] / (double)N.samples[k]->c->N;
  if (loss) loss[0] = base / (double)ns;
  std::vector<size_t> order(ns);
  for (size_t i = 0; i < ns; ++i) order[i] = i;
  uint64_t st = seed + 0x1234567ull;
  for (int ep = 0; ep < epochs; ++ep) {
    for (size_t i = ns; i > 1; --i) {  // Fisher-Yates, splitmix64
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      std::swap(order[i - 1], order[z % i]);
    }
    for (size_t k = 0; k < ns; ++k) {
      hecmw33_net::Sample& S = *N.samples[order[k]];
      Context& c = *S.c;
      if (int32_t e = sample_bce(N, S, true, &d[2 * k])) return e;
      net_backward(c.N, N.Wb.as<__bf16>(), c.vfeat.as<__bf16>(), N.H1.as<__bf16>(), N.H2.as<__bf16>(),
                   N.dY.as<float>(), N.wpart.as<float>(), N.dW.as<float>(), N.stream);
      net_adam(N.Wm.as<float>(), N.M.as<float>(), N.V.as<float>(), N.dW.as<float>(), N.Wb.as<__bf16>(), (float)lr,
               0.9f, 0.999f, 1e-8f, ++N.step, 1.0f, N.stream);
    }
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(N.stream));
    double acc = 0.0;
    for (size_t k = 0; k < ns; ++k) acc += d[2 * k] / (double)N.samples[order[k]]->c->N;
    if (loss) loss[1 + ep] = acc / (double)ns;
  }
  N.gen += 1;
  return 0;
}

// Mean logistic loss (*mean) and the fraction of nodes classified as their
// label (*accuracy, optional) of the current weights over the samples.
int32_t hecmw33_net_score(hecmw33_net* net, double* mean, double* accuracy) {
  if (!net || !mean || net->samples.empty()) return HECMW33_ERROR_ARGS;
  hecmw33_net& N = *net;
  double l = 0.0, a = 0.0;
  for (auto* Sp : N.samples) {
    if (int32_t e = sample_labels(*Sp)) return e;
    double d[2];
    if (int32_t e = sample_bce(N, *Sp, false, d)) return e;
    HIPC(hipStreamSynchronize(N.stream));
    l += d[0] / (double)Sp->c->N;
    a += d[1] / (double)Sp->c->N;
  }
  *mean = l / (double)N.samples.size();
  if (accuracy) *accuracy = a / (double)N.samples.size();
  return 0;
}

static int32_t set_coords(Context& c, int32_t NP, const double* xyz) {
  if (NP < 0 || (NP > 0 && !xyz)) return HECMW33_ERROR_ARGS;
  c.h_coord.assign(xyz, xyz + 3 * (size_t)NP);
  ++c.coord_gen;
  c.coord_fresh = true;
  return 0;
}

int32_t hecmw33_ctx_set_mpc(hecmw33_ctx* ctx, int32_t N, int32_t n_mpc, const int32_t* mpc_index,
                            const int32_t* mpc_item, const int32_t* mpc_dof, const double* mpc_val) {
  if (!ctx) return HECMW33_ERROR_ARGS;
  return set_mpc(ctx->c, N, n_mpc, mpc_index, mpc_item, mpc_dof, mpc_val);
}

int32_t hecmw33_ctx_set_coords(hecmw33_ctx* ctx, int32_t NP, const double* xyz) {
  if (!ctx) return HECMW33_ERROR_ARGS;
  return set_coords(ctx->c, NP, xyz);
}

int32_t hecmw33_set_coords(int32_t NP, const double* xyz) {
  std::lock_guard<std::mutex> lk(g_mu);
  return set_coords(*default_ctx(), NP, xyz);
}

int32_t hecmw33_ctx_coarse_get(hecmw33_ctx* ctx, int32_t* n_agg, int32_t* n_pair, int32_t* agg, float* dv,
                               int32_t* pair_a, int32_t* pair_b, double* Eblk) {
  if (!ctx || !n_agg || !n_pair) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if (!c.coarse_sym) {
    *n_agg = 0;
    *n_pair = 0;
    return 0;
  }
  const int32_t na = c.coarse.n_agg, np = c.coarse_npair;
  *n_agg = na;
  *n_pair = np;
  HIPC(hipStreamSynchronize(c.stream));
  if (agg) HIPC(hipMemcpy(agg, c.cg_agg.p, sizeof(int32_t) * c.N, hipMemcpyDeviceToHost));
  if (dv) HIPC(hipMemcpy(dv, c.cg_dv.p, sizeof(float) * 3 * (size_t)c.N, hipMemcpyDeviceToHost));
  if (pair_a) HIPC(hipMemcpy(pair_a, c.cg_pa.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
  if (pair_b) HIPC(hipMemcpy(pair_b, c.cg_pb.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
  if (Eblk) {
    if (!c.coarse_num || c.coarse_nm_num != 6) {
      set_error("coarse matrix of the rigid-body modes not assembled for the current values");
      return HECMW33_ERROR_ARGS;
    }
    HIPC(hipMemcpy(Eblk, c.cg_Eblk.p, sizeof(double) * 36 * (size_t)np, hipMemcpyDeviceToHost));
  }
  return 0;
}

int32_t hecmw33_ctx_coarse_inverse_get(hecmw33_ctx* ctx, int32_t* nc, int32_t* ldw, int32_t* fp64, void* W) {
  if (!ctx || !nc || !ldw || !fp64) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if (!c.coarse_sym || !c.coarse_num || !c.coarse_ok) {
    set_error("no coarse operator set up for the current values");
    return HECMW33_ERROR_ARGS;
  }
  const CoarseOp& op = c.coarse;
  *nc = op.nc;
  *ldw = op.ldw;
  *fp64 = op.nm == 12 ? 1 : 0;
  if (W) {
    HIPC(hipStreamSynchronize(c.stream));
    const size_t bytes = (op.nm == 12 ? sizeof(double) : sizeof(float)) * (size_t)op.nc * op.ldw;
    HIPC(hipMemcpy(W, op.nm == 12 ? (const void*)op.W64 : (const void*)op.W, bytes, hipMemcpyDeviceToHost));
  }
  return 0;
}

int32_t hecmw33_ctx_modes_get(hecmw33_ctx* ctx, float* xm, double* Eblk, float* R) {
  if (!ctx) return HECMW33_ERROR_ARGS;
  Context& c = ctx->c;
  if (!c.coarse_sym || !c.modes_valid) {
    set_error("no learned coarse modes fitted for the current values");
    return HECMW33_ERROR_ARGS;
  }
  HIPC(hipStreamSynchronize(c.stream));
  if (xm) HIPC(hipMemcpy(xm, c.cg_xm.p, sizeof(float) * 18 * (size_t)c.N, hipMemcpyDeviceToHost));
  if (R) HIPC(hipMemcpy(R, c.ml_S.p, sizeof(float) * 48 * (size_t)c.N, hipMemcpyDeviceToHost));
  if (Eblk) {
    if (!c.coarse_num || c.coarse_nm_num != 12) {
      set_error("coarse matrix of the twelve modes not assembled for the current values");
      return HECMW33_ERROR_ARGS;
    }
    // the four 6x6 tiles of every pair into row-major 12x12 blocks
    const int32_t np = c.coarse_npair;
    std::vector<double> t(144 * (size_t)np);
    HIPC(hipMemcpy(t.data(), c.cg_Eblk.p, sizeof(double) * 144 * (size_t)np, hipMemcpyDeviceToHost));
    for (int32_t p = 0; p < np; ++p)
      for (int tile = 0; tile < 4; ++tile)
        for (int r = 0; r < 6; ++r)
          for (int q = 0; q < 6; ++q)
            Eblk[144 * (size_t)p + 12 * (6 * (tile >> 1) + r) + 6 * (tile & 1) + q] =
                t[(4 * (size_t)p + tile) * 36 + 6 * r + q];
  }
  return 0;
}

int32_t hecmw33_rigid_labels(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                             const int32_t* itemU, const double* xyz, int32_t nsamp, const double* e, double tau,
                             float* labels) {
  if (N < 1 || !indexL || !indexU || !xyz || nsamp < 1 || !e || !labels || !(tau > 0.0) ||
      (indexL[N] > 0 && !itemL) || (indexU[N] > 0 && !itemU))
    return HECMW33_ERROR_ARGS;
  std::vector<float> worst;
  for (int32_t k = 0; k < nsamp; ++k)
    rigid_edge_ratios(N, indexL, itemL, indexU, itemU, xyz, e + 3 * (int64_t)N * k, worst);
  return rigid_labels_from_ratios(N, indexL, itemL, indexU, itemU, worst, tau, labels);
}

int32_t hecmw33_ctx_set_net(hecmw33_ctx* ctx, hecmw33_net* net) {
  if (!ctx) return HECMW33_ERROR_ARGS;
  ctx->c.net = net;
  return 0;
}

int32_t hecmw33_set_net(hecmw33_net* net) {
  std::lock_guard<std::mutex> lk(g_mu);
  default_ctx()->net = net;
  return 0;
}

}  // extern "C"
