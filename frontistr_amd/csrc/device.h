// Device data layout and kernel launchers of the 3x3-block Krylov solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace hecmw33 {

// the message hecmw33_last_error() returns on this thread
void set_last_error(const std::string& s);

constexpr int kWave = 64;       // CDNA wavefront
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kNumXcd = 8;      // MI355X: 8 XCDs, each with its own L2

// Block rows in SELL-C form, C = rows per slice: slice s holds
// width(s) = ptr[s+1]-ptr[s] block slots per row; slot k of row r (0 <= r < C)
// keeps its column in col[(ptr[s]+k)*C + r] and its 3x3 values in 9 planes
// val[((ptr[s]+k)*9 + e)*C + r], e = 3*i+j for block entry (i,j).
//  * The full matrix for the SpMV uses C = 64, one row per lane: in the
//    natural order consecutive rows have consecutive neighbours, so the x
//    gathers of a wavefront are already nearly contiguous, and one wavefront
//    load of a plane is one 512-byte segment.
//  * The multicolour SSOR factors use C = 16 with four lanes per row (lane 4r+d
//    computes component d of row r, the fourth lane mirrors d = 2): inside a
//    colour the rows are scattered nodes, so the three component lanes of a row
//    gather a neighbour's z from one cache line and share it with DPP quad
//    broadcasts, and a colour launch has four times as many wavefronts.
constexpr int kRowsA = 64;
constexpr int kRowsS = 16;

// The SpMV matrix A (C = 64) pairs its planes instead: in slot group
// g = ptr[s]+k, entries e = 2p and 2p+1 of lane r's block are adjacent at
// val[g*576 + 128p + 2r + (e&1)] (p = 0..3), entry 8 at val[g*576 + 512 + r],
// so a lane fetches a block with four 16-byte loads (8-byte in FP32) and one
// 8-byte load, and each wavefront load is still one contiguous segment.
__host__ __device__ inline int64_t aval_index(int64_t g, int e, int r) {
  return g * (9 * kRowsA) + (e < 8 ? (int64_t)(e >> 1) * (2 * kRowsA) + 2 * r + (e & 1) : 8 * kRowsA + r);
}
template <typename T>
struct Pair2;
template <>
struct Pair2<double> {
  using type = double2;
};
template <>
struct Pair2<float> {
  using type = float2;
};
// the 9 entries of lane r's block in slot group g of A
template <typename T>
__device__ __forceinline__ void aval_load(const T* __restrict__ val, int64_t g, int r, T (&v)[9]) {
  using V2 = typename Pair2<T>::type;
  const T* __restrict__ b = val + g * (9 * kRowsA);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const V2 t = *reinterpret_cast<const V2*>(b + q * (2 * kRowsA) + 2 * r);
    v[2 * q] = t.x;
    v[2 * q + 1] = t.y;
  }
  v[8] = b[8 * kRowsA + r];
}
// colours above this many rows on average sweep with a lane per row (SELL-64):
// 4-5% faster at 245k rows per colour (160^3 cube), slower at 60k (100^3)
constexpr int kSsorRows64Min = 150000;

template <typename T>
struct Sell {
  int32_t nslices = 0;
  int32_t nrows = 0;
  int32_t rows = kRowsA;       // rows per slice (C)
  int64_t nslots = 0;          // sum of widths * C
  int32_t* ptr = nullptr;      // [nslices+1]
  int32_t* len = nullptr;      // [nslices*C] blocks in the row (0 = idle row)
  int32_t* col = nullptr;      // [nslots] 0-based node id of the block column
  T* val = nullptr;            // [9*nslots]
};

// Device-resident scalars of one Krylov solve; kernels read them and the
// one-workgroup tail kernels update them, so the host never waits inside an
// iteration.
struct KrylovState;

// Multicolour SSOR operator (hecmw_precond_SSOR_33) as the sweeps see it.
template <typename T>
struct SsorOp {
  Sell<T> L, U;                // L'/U' per colour slice; columns are slot ids
  const int32_t* row_old;      // node (0-based, natural order) of every slot, -1 = padding
  const T* alu;                // 3x3 LU factors, planes of L.rows rows
  const int32_t* color_slice;  // host: first slice of every colour, ncolor+1 entries
  int32_t ncolor;
  T* zs;                       // sweep vector in slot order, 3 entries per slot
  const int32_t* node_slot;    // slot of every internal node (the inverse of row_old)
  int32_t n;                   // internal nodes
  int32_t last_bwd_empty;      // 1: the last colour's rows have no U' entries (their backward step is z -= 0)
  const uint16_t* Lv16 = nullptr;  // FP32 sweeps: L'/U' values in BF16 instead of L.val/U.val (null: T)
  const uint16_t* Uv16 = nullptr;
};
void launch_to_bf16(int64_t n, const float* a, uint16_t* b, hipStream_t s);
void ssor_apply(const SsorOp<double>& op, const double* r, double* z, const KrylovState* st, hipStream_t s);
void ssor_apply_f32(const SsorOp<float>& op, const float* r, float* z, const KrylovState* st, hipStream_t s);

struct KrylovState {
  double bnrm2, tol;
  double rho, rho1, alpha, beta, c1, dnrm2, resid;
  double omega, c2, ts, tt;
  int32_t iter;      // 1-based iteration being executed
  int32_t maxit;
  int32_t done;      // 1: the loop has exited, every later kernel returns at once
  int32_t error;
  int32_t n_indef;
  int32_t recheck;   // recompute the true residual in this iteration
  int32_t nrec;      // true-residual period (50 CG, 100 BiCGSTAB)
  int32_t pad;
  // flexible CG with FP32 operators (reliable updates)
  double zq, pq, rp;   // z.q_prev, p.q, r.p of the current iteration
  double rr_ref;       // |r|^2 at the last FP64 true-residual update
  double delta2;       // update when |r|^2 <= delta2 * rr_ref
  int32_t reliable;    // recompute r = b - A x in FP64 in this iteration
  int32_t nreliable;   // FP64 true-residual updates so far
  // learned two-level step: learned_off = 1 when the solve is not a learned one or has dropped
  // the step; coarse_off = 1: the coarse kernels leave z to the SSOR sweeps (plain SSOR FCG);
  // fallback_iter: the iteration at which the learned step was dropped (0: kept)
  int32_t learned_off, coarse_off, fallback_iter, pad2;
  double edrop;  // sum of alpha (r.p): the decrease of |e|_A^2 over the iterations so far
  // true / recurrence residual seen at the last replacement: near convergence
  // the FP32 recurrence runs ahead, so it claims convergence at resid * gap <= tol
  double gap;
  // the learned step is dropped when |r|^2 at a replacement exceeds drop2 times the smallest
  // FP64 |r|^2 seen so far (rr_min)
  double drop2, rr_min;
};

constexpr int kMaxPartials = 65536;  // per reduction component

// ---- setup ---------------------------------------------------------------------
// Fill SELL values from the caller's D/AL/AU through a slot->source map
// (src >= 0: block index into [D | AL | AU], -1: padding).
// Plain streaming read of n doubles (n even, a 16-byte aligned), grid
// workgroups, one partial each in part.
void launch_stream_read(const double* a, int64_t n, double* part, int grid, hipStream_t s);
// launch_fill_values: 9 planes (the SSOR factors); the _a variants: A's paired
// planes (aval_index).
void launch_fill_values(const Sell<double>& a, const int64_t* src, const double* D, const double* AL,
                        const double* AU, int64_t nD, int64_t nAL, hipStream_t s);
void launch_fill_values_a(const Sell<double>& a, const int64_t* src, const double* D, const double* AL,
                          const double* AU, int64_t nD, int64_t nAL, hipStream_t s);
void launch_fill_values_a_f32(const Sell<float>& a, const int64_t* src, const double* D, const double* AL,
                              const double* AU, int64_t nD, int64_t nAL, hipStream_t s);
// 3x3 no-pivot LU with sigma on the pivots of D[row_src[t]] into 9-plane
// slices of `rows` rows (hecmw_precond_SSOR_33.F90:164-208,
// hecmw_precond_DIAG_33.F90:58-91)
void launch_block_lu(int32_t nslots, const int32_t* row_src, const double* D, double sigma, double* alu, int rows,
                     hipStream_t s);
void launch_block_lu_f32(int32_t nslots, const int32_t* row_src, const double* D, double sigma, float* alu, int rows,
                         hipStream_t s);

// ---- plain operators -------------------------------------------------------------
// Multipoint constraints of the MPC-CG SpMV y = T^T A T x (hecmw_TtmatTvec_33,
// hecmw_solver_las_33.F90:442-586; hecMESH%mpc): T is the identity except each
// slave dof's row, which is minus the constraint's master coefficients, so
// (T x)_slave = -sum_j val_j x_master_j and T^T zeroes the slave entries and
// subtracts val_j (A T x)_slave from each master.  Constraints with a dof above 3
// are skipped, as the reference skips them.
struct MpcOp {
  int32_t n3 = 0;                     // 3 N internal dofs
  const int32_t* slave_row = nullptr; // [3N] constraint whose slave this dof is, -1 if none
  const int32_t* mptr = nullptr;      // [nmpc+1] master terms of each constraint, in order
  const int32_t* mdof = nullptr;      //   master dof (0-based)
  const double* mval = nullptr;       //   coefficient
  const int32_t* tptr = nullptr;      // [3N+1] per dof, the constraints it is a master of,
  const int32_t* tslave = nullptr;    //   in constraint order: their slave dof
  const double* tval = nullptr;       //   and its coefficient there
  double* w1 = nullptr;               // T x   (3 NP)
  double* w2 = nullptr;               // A T x (3 NP)
};

// with mpc (n3 > 0) the SpMV launchers below multiply by T^T A T instead of A
void launch_spmv(const Sell<double>& a, const double* x, double* y, hipStream_t s);
void launch_spmv_mpc(const Sell<double>& a, const MpcOp& mpc, const double* x, double* y, hipStream_t s);
void launch_spmv_f32(const Sell<float>& a, const float* x, float* y, hipStream_t s);
void launch_matresid(const Sell<double>& a, const double* x, const double* b, double* r, hipStream_t s,
                     const MpcOp* mpc = nullptr);
void launch_dot_store(int64_t n, const double* x, const double* y, double* out, KrylovState* st, double* part,
                      hipStream_t s);
void launch_diag_apply(int32_t n, const double* alu, const double* r, double* z, const KrylovState* st,
                       hipStream_t s);
void launch_copy(int64_t n3, const double* r, double* z, const KrylovState* st, hipStream_t s);
void launch_zero(int64_t n, double* x, const KrylovState* st, hipStream_t s);
void launch_add(int64_t n, const double* x, double* y, const KrylovState* st, hipStream_t s);
void launch_matresid_guarded(const Sell<double>& a, const double* x, const double* b, double* r,
                             const KrylovState* st, hipStream_t s, const MpcOp* mpc = nullptr);
// z = M^-1 r, kind 0: copy, 1/2: multicolour SSOR, 3: DIAG (guarded by st->done)
void precond_apply(int32_t kind, const SsorOp<double>& op, const double* alu_diag, int32_t n, const double* r,
                   double* z, const KrylovState* st, hipStream_t s);

// ---- fused CG iteration pieces (hecmw_solver_CG.f90:164-279) ----------------------
// z = M^-1 r with rho = r.z and the rho/beta logic
void cg_precond_ssor_rho(const SsorOp<double>& op, const double* r, double* z, KrylovState* st, double* part,
                         hipStream_t s);
void cg_precond_diag_rho(int32_t n, const double* alu, const double* r, double* z, KrylovState* st, double* part,
                         hipStream_t s);
void cg_precond_none_rho(int64_t n3, const double* r, double* z, KrylovState* st, double* part, hipStream_t s);
// p = z | z + beta p
void cg_xpay(int64_t n3, const double* z, double* p, const KrylovState* st, hipStream_t s);
// q = A p, c1 = p.q, alpha
// the SpMV launchers with a fused dot record `after` (if given) between the SpMV
// kernel and its scalar tail, so TIMELOG's time_Ax is the SpMV kernel alone
void cg_spmv_dot(const Sell<double>& a, const double* p, double* q, KrylovState* st, double* part, hipStream_t s,
                 hipEvent_t after = nullptr, const MpcOp* mpc = nullptr);
// x += alpha p, r -= alpha q, |r|^2
void cg_update(int64_t n3, const double* p, const double* q, double* x, double* r, KrylovState* st, double* part,
               double* hist, hipStream_t s);
// conditional r = b - A x with |r|^2, convergence, end of iteration
void cg_true_resid(const Sell<double>& a, const double* x, const double* b, double* r, KrylovState* st, double* part,
                   double* hist, hipStream_t s, const MpcOp* mpc = nullptr);

// rho = r.z only (ITERPREMAX > 1 path)
void cg_rho(int64_t n3, const double* r, const double* z, KrylovState* st, double* part, hipStream_t s);

// ---- fused BiCGSTAB iteration pieces (hecmw_solver_BiCGSTAB.f90:164-275) ------------
void bi_rho(int64_t n3, const double* r, const double* rt, KrylovState* st, double* part, hipStream_t s);
void bi_p(int64_t n3, const double* r, const double* v, double* p, const KrylovState* st, hipStream_t s);
void bi_spmv_c2(const Sell<double>& a, const double* pt, const double* rt, double* v, KrylovState* st, double* part,
                hipStream_t s, hipEvent_t after = nullptr, const MpcOp* mpc = nullptr);
void bi_s(int64_t n3, const double* v, const double* r, double* sv, const KrylovState* st, hipStream_t s);
void bi_spmv_omega(const Sell<double>& a, const double* stv, const double* sv, double* t, KrylovState* st,
                   double* part, hipStream_t s, hipEvent_t after = nullptr, const MpcOp* mpc = nullptr);
void bi_update(int64_t n3, const double* pt, const double* stv, const double* t, const double* sv, double* x,
               double* r, KrylovState* st, double* part, double* hist, hipStream_t s);

// ---- flexible CG with FP32 operators and FP64 recurrences ------------------------
// z32 = M^-1 r32 in FP32 (kind 0 copy, 1 SSOR, 3 DIAG)
void precond_apply_f32(int32_t kind, const SsorOp<float>& op, const float* alu_diag, int32_t n, const float* r,
                       float* z, const KrylovState* st, hipStream_t s);
// z32 = M^-1 r32 fused with the FP64 dots r.z and z.q_prev; beta = -(z.q_prev)/(p.q)_prev
// (Notay's FCG(1))
void fcg_precond_dots(int32_t kind, const SsorOp<float>& op, const float* alu_diag, int32_t n, const float* r,
                      float* z, const float* q, KrylovState* st, double* part, hipStream_t s);
void fcg_p(int64_t n3, const float* z, float* p, const KrylovState* st, hipStream_t s);
// q = A32 p, pq = p.q, rp = r.p, alpha = rp / pq
void fcg_spmv(const Sell<float>& a, const float* p, const float* r, float* q, KrylovState* st, double* part,
              hipStream_t s, hipEvent_t after = nullptr);
// the same with the FP64 matrix values (FP32 vectors, FP64 accumulation)
void fcg_spmv_mixed(const Sell<double>& a, const float* p, const float* r, float* q, KrylovState* st, double* part,
                    hipStream_t s, hipEvent_t after = nullptr);
// x64 += alpha p, r32 -= alpha q, |r|^2, reliable-update decision
void fcg_update(int64_t n3, const float* p, const float* q, double* x, float* r, KrylovState* st, double* part,
                double* hist, hipStream_t s);
// conditional r32 = b - A64 x64 in FP64 with |r|^2, convergence, end of iteration
void fcg_reliable(const Sell<double>& a, const double* x, const double* b, float* r, KrylovState* st, double* part,
                  double* hist, hipStream_t s);
// r32 = b - A64 x64 and |r|^2 into *rr
void fcg_init_resid(const Sell<double>& a, const double* x, const double* b, float* r, double* rr, KrylovState* st,
                    double* part, hipStream_t s);
// ---- diagonal scaling (hecmw_solver_scaling_33.f90) ----------------------------------
void launch_make_scale(int32_t n, int32_t np, const double* D, double* sc, hipStream_t s);
void launch_scale_raw(bool fw, int32_t np, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                      const int32_t* itemU, double* D, double* AL, double* AU, const double* sc, hipStream_t s);
void launch_scale_vec(bool mul, int64_t n, double* v, const double* sc, hipStream_t s);

// learned flexible CG: beta from the four dots of net_apply (r.z, z.q, r.z0, z0.q), falling
// back to the plain SSOR output z0 when the learned z is not a descent direction
// the flexible CG's beta tail over nparts partials of (r.z, z.q_prev), and its
// alpha tail over (p.q, r.p)
void fcg_beta(KrylovState* st, const double* part, int nparts, hipStream_t s);
void fcg_alpha(KrylovState* st, const double* part, int nparts, hipStream_t s);
// y = A x for the flexible CG without dots: FP32 values or FP64 values (FP32 vectors)
void fcg_spmv_store(const Sell<float>& a32, const Sell<double>& a64, bool fp64_values, const float* x, float* y,
                    const KrylovState* st, hipStream_t s);

// Two-level correction (coarse.hip): piecewise rigid-body modes on geometric
// aggregates, z += Z E^-1 Z^T r.
struct CoarseOp {
  int32_t n_agg = 0, nc = 0, ldw = 0;  // aggregates, coarse unknowns (nm per aggregate), row stride of W
  int32_t nm = 6;                      // modes per aggregate: 6 rigid-body, or 12 with the learned ones
  const int32_t* agg = nullptr;        // [NP] aggregate of each node, -1 for external nodes
  const int32_t* aptr = nullptr;       // [n_agg+1] into anodes
  const int32_t* anodes = nullptr;     // nodes of each aggregate, ascending
  const float* dv = nullptr;           // [3N] scaled offsets from the aggregate centroid
  const float* xm = nullptr;           // nm = 12: [N][6][3] the learned modes' values at each node
  float* W = nullptr;                  // E^-1, FP32, nc x ldw row-major (six modes)
  float* cv = nullptr;                 // [ldw] Z^T r (entries nc..ldw-1 stay 0)
  float* yv = nullptr;                 // [nc] E^-1 Z^T r
  double* W64 = nullptr;               // twelve modes: L^-1 (E = L L^T) in FP64, mirrored into a full
  double* cv64 = nullptr;              // nc x ldw array, and c, L^-1 c and y in FP64
  double* tv64 = nullptr;
  double* yv64 = nullptr;
};
// E = Z^T A Z: the nm x nm block of every coupled aggregate pair (pa[p], pb[p]), summed
// over the pair's nodes pn[pn_ptr[p]..pn_ptr[p+1]), then scattered into the dense E
// (Eblk: nm*nm doubles per pair)
void coarse_assemble(const Sell<double>& A, const CoarseOp& op, int32_t npair, const int32_t* pa, const int32_t* pb,
                     const int32_t* pn_ptr, const int32_t* pn, double* Eblk, double* E, int32_t lda, hipStream_t s);
// the iterations' copies: E^-1 (potri's lower triangle) mirrored into FP32 W,
// or L^-1 (trtri's lower triangle) mirrored into FP64 W64
void coarse_inverse_f32(const CoarseOp& op, const double* Ei, int32_t lda, hipStream_t s);
void coarse_inverse_f64(const CoarseOp& op, const double* Ei, int32_t lda, hipStream_t s);
// The symmetric two-level step (W = A Z per node and touched aggregate):
// build W from the symbolic node-major lists ncnt/nagg into the compacted
// sc (per node: first slot, count) / aggW / W, dropping the entries that are
// zero in exact arithmetic, with kept[g] entries per 256-node workgroup;
// pre: z = Z E^-1 Z^T r, r1 = r - A z; add: z += t (the SSOR output of r1);
// post: with w = A z, z += Z E^-1 Z^T (r - w), az = w + A Z E^-1 Z^T (r - w) = A z, partials
// (r.z, q_prev.z); pq: p = z + beta p, q = az + beta q, partials (p.q, r.p).
void coarse_build_W(const Sell<double>& A, const CoarseOp& op, int32_t n, const int32_t* ncnt, const int32_t* nagg,
                    int2* sc, int32_t* aggW, float* W, int32_t* kept, hipStream_t s);
void coarse_sym_pre(const CoarseOp& op, int32_t n, const int2* sc, const int32_t* aggW, const float* W,
                    const float* r, float* z, float* r1, const KrylovState* st, hipStream_t s);
void coarse_sym_add(int64_t n3, const float* t, float* z, const KrylovState* st, hipStream_t s);
int coarse_sym_post(const CoarseOp& op, int32_t n, const int2* sc, const int32_t* aggW, const float* W,
                    const float* r, const float* w, float* z, float* az, const float* q, const KrylovState* st,
                    double* part, hipStream_t s);
int coarse_fcg_pq(int64_t n3, const float* z, const float* az, float* p, float* q, const float* r,
                  const KrylovState* st, double* part, hipStream_t s);
// z += Z E^-1 Z^T r; with dots also the beta partials (r.z, q.z) of the new z.
// Returns the number of partials.
int coarse_apply(const CoarseOp& op, int32_t n, const float* r, float* z, const float* q, const KrylovState* st,
                 double* part, bool dots, hipStream_t s);

// ---- learned coarse space (learned.hip) -------------------------------------------------
int net_weight_count();
void net_static_features(const Sell<double>& A, int64_t n, const float* mat, __bf16* X, hipStream_t s);
void net_forward(int64_t n, const __bf16* Wb, const __bf16* X, __bf16* H1, __bf16* H2, const float* z0, float* z,
                 hipStream_t s);
int net_bwd_parts(int64_t n);
void net_backward(int64_t n, const __bf16* Wb, const __bf16* X, const __bf16* H1, const __bf16* H2, const float* dY,
                  float* part, float* dW, hipStream_t s);
void net_adam(float* Wm, float* M, float* V, const float* dW, __bf16* Wb, float lr, float b1, float b2, float eps,
              int step, float scale, hipStream_t s);
void net_to_bf16(const float* Wm, __bf16* Wb, hipStream_t s);
void net_bce(int64_t n, const float* Y, const float* t, float* dY, double* part, double* out, hipStream_t s);
void net_mask(int64_t n, const float* Y, uint8_t* mask, hipStream_t s);
void net_material(int64_t n, const double* D, float* m, double* part, double* sum, hipStream_t s);

// learned coarse modes (per-aggregate linear autoencoders on the solver's slow
// errors): samples S[(3i+d)*16+s] = v - x; per (aggregate, sample) the rigid-body
// restriction Z^T e (rhs, 6 per pair) and per aggregate Z^T Z (ztz, 36); the part
// beyond the rigid-body modes R = S - Z c; the Gram matrices G_a = R_a^T R_a (16x16,
// BF16 MFMA, FP32 accumulation); the modes xm[18i+3j+d] = R V[agg(i)] (V: 16x6 per aggregate)
void modes_random(int64_t n3, uint64_t seed, double* v, hipStream_t st);
void modes_sample(int64_t n3, const double* v, const double* x, float* S, int s, hipStream_t st);
void modes_rigid(int32_t n_agg, int ns, const int32_t* aptr, const int32_t* anodes, const float* dv, const float* S,
                 double* rhs, double* ztz, hipStream_t st);
void modes_resid(int32_t n, const int32_t* agg, const float* dv, const double* c, float* S, hipStream_t st);
void modes_gram(int32_t n_agg, const int32_t* aptr, const int32_t* anodes, const float* S, float* G, hipStream_t st);
void modes_make(int32_t n, const int32_t* agg, const float* S, const float* V, float* xm, hipStream_t st);

// r32 = (float) r64 over n entries
void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t s);

}  // namespace hecmw33
