/*
 * CPU restatement of hecmw1's 3x3-block iterative solve (FrontISTR 5.9) in
 * plain C.  TEST INFRASTRUCTURE ONLY: the checker the GPU library is compared
 * with.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * use it; the product never links it.
 *
 * Every routine follows one reference routine step by step, in the same
 * floating-point operation order, single-threaded (dot products summed
 * serially as the reference's OpenMP-free loops do), so with a compiler that
 * does not contract a*b+c into FMA (x86-64 baseline) its results are bitwise
 * those of the reference build in oracle/_ref.  Arrays use the reference's
 * layout: 1-based item ids, 0-based cumulative index arrays, row-major 3x3
 * blocks, xyz-interleaved vectors.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t N, NP;
  const double *D, *AL, *AU;
  const int32_t *indexL, *itemL, *indexU, *itemU;
} rs_sys;

/* ---- ordering: hecmw1/src/solver/matrix/hecmw_matrix_ordering_CM.f90 ---- */

/* count_degrees (:166-179): L count + U items that are internal */
static void count_degrees(const rs_sys* A, int32_t* deg) {
  for (int32_t i = 1; i <= A->N; ++i) {
    int32_t d = A->indexL[i] - A->indexL[i - 1];
    for (int32_t j = A->indexU[i - 1]; j < A->indexU[i]; ++j)
      if (A->itemU[j] <= A->N) ++d;
    deg[i] = d;
  }
}

/* sort_nodes_by_degree (:207-233): recursive Hoare partition, middle pivot.
 * lv is 1-based (lv[1..]) so the middle index is computed as in Fortran. */
static void sort_by_degree(int32_t* lv, int32_t istart, int32_t iend, const int32_t* deg) {
  if (istart >= iend) return;
  const int32_t pivot = deg[lv[(istart + iend) / 2]];
  int32_t left = istart, right = iend;
  for (;;) {
    while (deg[lv[left]] < pivot) ++left;
    while (pivot < deg[lv[right]]) --right;
    if (left >= right) break;
    int32_t t = lv[left];
    lv[left] = lv[right];
    lv[right] = t;
    ++left;
    --right;
  }
  sort_by_degree(lv, istart, left - 1, deg);
  sort_by_degree(lv, right + 1, iend, deg);
}

/* ordering_CM_inner (:80-164): level structure from nstart; lv_item 1-based */
static int32_t cm_inner(const rs_sys* A, const int32_t* deg, int32_t nstart, int32_t* lv_index, int32_t* lv_item) {
  const int32_t N = A->N;
  int32_t* iwk = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  int32_t nlevel = 1;
  lv_index[0] = 0;
  iwk[nstart] = 1;
  int32_t cntall = 1;
  lv_item[1] = nstart;
  lv_index[1] = 1;
  if (cntall == N) nlevel = 1;
  for (int32_t level = 2; level <= N; ++level) {
    int32_t cnt = 0;
    for (int32_t j = lv_index[level - 2] + 1; j <= lv_index[level - 1]; ++j) {
      const int32_t jn = lv_item[j];
      for (int32_t k = A->indexL[jn - 1]; k < A->indexL[jn]; ++k) {
        const int32_t kn = A->itemL[k];
        if (iwk[kn] == 0) {
          iwk[kn] = level;
          ++cnt;
          lv_item[++cntall] = kn;
        }
      }
      for (int32_t k = A->indexU[jn - 1]; k < A->indexU[jn]; ++k) {
        const int32_t kn = A->itemU[k];
        if (kn > N) continue;
        if (iwk[kn] == 0) {
          iwk[kn] = level;
          ++cnt;
          lv_item[++cntall] = kn;
        }
      }
    }
    if (cnt == 0) {
      for (int32_t kn = 1; kn <= N; ++kn)
        if (iwk[kn] == 0) {
          iwk[kn] = level;
          ++cnt;
          lv_item[++cntall] = kn;
          break;
        }
    }
    lv_index[level] = cntall;
    sort_by_degree(lv_item, lv_index[level - 1] + 1, lv_index[level], deg);
    if (cntall == N) {
      nlevel = level;
      break;
    }
  }
  free(iwk);
  return nlevel;
}

/* hecmw_matrix_ordering_RCM (:65-78) = CM (:18-63) then reverse_ordering
 * (:235-245).  perm/iperm 1-based ids in 0-based arrays. */
void rs_ordering_rcm(const rs_sys* A, int32_t* perm, int32_t* iperm) {
  const int32_t N = A->N;
  int32_t* deg = (int32_t*)malloc(sizeof(int32_t) * ((size_t)N + 1));
  count_degrees(A, deg);
  /* find_minimum_degrees (:181-205) */
  int32_t mins[5], nmin = 0, degmin = N;
  for (int32_t i = 1; i <= N; ++i) {
    if (deg[i] == 0) continue;
    if (deg[i] < degmin) {
      degmin = deg[i];
      nmin = 1;
      mins[0] = i;
    } else if (deg[i] == degmin) {
      ++nmin;
      if (nmin <= 5) mins[nmin - 1] = i;
    }
  }
  if (nmin > 5) nmin = 5;
  int32_t best_level = -1;
  int32_t* lv_index = (int32_t*)malloc(sizeof(int32_t) * ((size_t)N + 2));
  int32_t* lv_item = (int32_t*)malloc(sizeof(int32_t) * ((size_t)N + 2));
  int32_t* best = (int32_t*)malloc(sizeof(int32_t) * ((size_t)N + 2));
  for (int32_t s = 0; s < nmin; ++s) {
    const int32_t nl = cm_inner(A, deg, mins[s], lv_index, lv_item);
    if (nl > best_level) { /* first maximum wins */
      best_level = nl;
      memcpy(best, lv_item, sizeof(int32_t) * ((size_t)N + 1));
    }
  }
  if (nmin == 0) /* the reference leaves this case undefined; identity */
    for (int32_t i = 1; i <= N; ++i) best[i] = i;
  /* CM permutation then reverse: iperm = N+1 - position */
  for (int32_t i = 1; i <= N; ++i) iperm[best[i] - 1] = i;
  for (int32_t i = 1; i <= N; ++i) {
    iperm[i - 1] = N + 1 - iperm[i - 1];
    perm[iperm[i - 1] - 1] = i;
  }
  free(deg);
  free(lv_index);
  free(lv_item);
  free(best);
}

/* hecmw_matrix_ordering_MC (hecmw_matrix_ordering_MC.f90:15-72).  Returns the
 * number of colours; colorindex has N+1 slots. */
int32_t rs_ordering_mc(const rs_sys* A, const int32_t* perm_cur, int32_t ncolor_in, int32_t* colorindex, int32_t* perm,
                       int32_t* iperm) {
  const int32_t N = A->N;
  int32_t* iwk = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  const int32_t nn_color = N / ncolor_in;
  int32_t cntall = 0, ncolor = 0;
  colorindex[0] = 0;
  for (int32_t color = 1; color <= N; ++color) {
    int32_t cnt = 0;
    for (int32_t i = 0; i < N; ++i) {
      const int32_t in = perm_cur[i];
      if (iwk[in] > 0 || iwk[in] == -1) continue;
      iwk[in] = color;
      perm[cntall++] = in;
      ++cnt;
      if (cnt == nn_color) break;
      if (cntall == N) break;
      for (int32_t j = A->indexL[in - 1]; j < A->indexL[in]; ++j)
        if (iwk[A->itemL[j]] == 0) iwk[A->itemL[j]] = -1;
      for (int32_t j = A->indexU[in - 1]; j < A->indexU[in]; ++j) {
        const int32_t jn = A->itemU[j];
        if (jn > N) continue;
        if (iwk[jn] == 0) iwk[jn] = -1;
      }
    }
    colorindex[color] = cntall;
    if (cntall == N) {
      ncolor = color;
      break;
    }
    for (int32_t i = 1; i <= N; ++i)
      if (iwk[i] == -1) iwk[i] = 0;
  }
  free(iwk);
  for (int32_t i = 0; i < N; ++i) iperm[perm[i] - 1] = i + 1;
  return ncolor;
}

/* ---- LAS: hecmw1/src/solver/las/hecmw_solver_las_33.F90 ----------------- */

/* hecmw_matvec_33_inner CPU branch (:276-309): D, then L, then U per row */
void rs_matvec(const rs_sys* A, const double* X, double* Y) {
  for (int32_t i = 0; i < A->N; ++i) {
    const double* d = A->D + 9 * (size_t)i;
    double x1 = X[3 * i], x2 = X[3 * i + 1], x3 = X[3 * i + 2];
    double y1 = d[0] * x1 + d[1] * x2 + d[2] * x3;
    double y2 = d[3] * x1 + d[4] * x2 + d[5] * x3;
    double y3 = d[6] * x1 + d[7] * x2 + d[8] * x3;
    for (int32_t j = A->indexL[i]; j < A->indexL[i + 1]; ++j) {
      const int32_t c = A->itemL[j] - 1;
      const double* a = A->AL + 9 * (size_t)j;
      x1 = X[3 * c];
      x2 = X[3 * c + 1];
      x3 = X[3 * c + 2];
      y1 = y1 + a[0] * x1 + a[1] * x2 + a[2] * x3;
      y2 = y2 + a[3] * x1 + a[4] * x2 + a[5] * x3;
      y3 = y3 + a[6] * x1 + a[7] * x2 + a[8] * x3;
    }
    for (int32_t j = A->indexU[i]; j < A->indexU[i + 1]; ++j) {
      const int32_t c = A->itemU[j] - 1;
      const double* a = A->AU + 9 * (size_t)j;
      x1 = X[3 * c];
      x2 = X[3 * c + 1];
      x3 = X[3 * c + 2];
      y1 = y1 + a[0] * x1 + a[1] * x2 + a[2] * x3;
      y2 = y2 + a[3] * x1 + a[4] * x2 + a[5] * x3;
      y3 = y3 + a[6] * x1 + a[7] * x2 + a[8] * x3;
    }
    Y[3 * i] = y1;
    Y[3 * i + 1] = y2;
    Y[3 * i + 2] = y3;
  }
}

/* hecmw_matresid_33 (:367-399): R = B - A X */
static void matresid(const rs_sys* A, const double* X, const double* B, double* R) {
  rs_matvec(A, X, R);
  for (int64_t i = 0; i < 3 * (int64_t)A->N; ++i) R[i] = B[i] - R[i];
}

/* hecmw_innerProduct_R (hecmw_solver_misc.F90:46-76): serial sum */
static double dot(int64_t n, const double* x, const double* y) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s = s + x[i] * y[i];
  return s;
}

/* ---- preconditioners -------------------------------------------------------- */

/* 3x3 LU with sigma on the pivots, inverted pivots stored on the diagonal
 * (hecmw_precond_DIAG_33.F90:58-91, same as hecmw_precond_SSOR_33.F90:164-208) */
static void block_lu(const double* d, double sigma, double* alu) {
  double a[3][3], pw[3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = d[3 * r + c];
  a[0][0] = a[0][0] * sigma;
  a[1][1] = a[1][1] * sigma;
  a[2][2] = a[2][2] * sigma;
  for (int k = 0; k < 3; ++k) {
    a[k][k] = 1.0 / a[k][k];
    for (int i = k + 1; i < 3; ++i) {
      a[i][k] = a[i][k] * a[k][k];
      for (int j = k + 1; j < 3; ++j) pw[j] = a[i][j] - a[i][k] * a[k][j];
      for (int j = k + 1; j < 3; ++j) a[i][j] = pw[j];
    }
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) alu[3 * r + c] = a[r][c];
}

/* forward/back substitution with those factors (hecmw_precond_DIAG_33.F90:124-133) */
static void lu_solve(const double* a, double* x) {
  double x1 = x[0], x2 = x[1], x3 = x[2];
  x2 = x2 - a[3] * x1;
  x3 = x3 - a[6] * x1 - a[7] * x2;
  x3 = a[8] * x3;
  x2 = a[4] * (x2 - a[5] * x3);
  x1 = a[0] * (x1 - a[2] * x3 - a[1] * x2);
  x[0] = x1;
  x[1] = x2;
  x[2] = x3;
}

typedef struct {
  int kind; /* 0 none, 1 SSOR, 3 DIAG */
  int32_t N, ncolor;
  int32_t *perm, *iperm, *colorindex;
  int32_t *indexL, *itemL, *indexU, *itemU; /* reordered, items renumbered to old ids */
  double *AL, *AU, *ALU;
} rs_pc;

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* hecmw_precond_SSOR_33_setup (hecmw_precond_SSOR_33.F90:57-218) with the
 * multicolour branch: RCM, MC, reorder_profile (sorted rows), reorder_values
 * (value lookup by binary search), renum_item, ALU of the reordered diagonal. */
static void ssor_setup(const rs_sys* A, double sigma, int32_t ncolor_in, rs_pc* pc) {
  const int32_t N = A->N;
  int32_t* prcm = (int32_t*)malloc(sizeof(int32_t) * N);
  int32_t* iprcm = (int32_t*)malloc(sizeof(int32_t) * N);
  pc->perm = (int32_t*)malloc(sizeof(int32_t) * N);
  pc->iperm = (int32_t*)malloc(sizeof(int32_t) * N);
  pc->colorindex = (int32_t*)malloc(sizeof(int32_t) * ((size_t)N + 1));
  rs_ordering_rcm(A, prcm, iprcm);
  pc->ncolor = rs_ordering_mc(A, prcm, ncolor_in, pc->colorindex, pc->perm, pc->iperm);
  free(prcm);
  free(iprcm);
  /* reorder_profile (hecmw_matrix_reorder.f90:20-66) */
  pc->indexL = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  pc->indexU = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  const int64_t cap = (int64_t)A->indexL[N] + A->indexU[N];
  pc->itemL = (int32_t*)malloc(sizeof(int32_t) * (cap + 1));
  pc->itemU = (int32_t*)malloc(sizeof(int32_t) * (cap + 1));
  int32_t cl = 0, cu = 0;
  for (int32_t inew = 1; inew <= N; ++inew) {
    const int32_t iold = pc->perm[inew - 1];
    for (int32_t j = A->indexL[iold - 1]; j < A->indexL[iold]; ++j) {
      const int32_t jnew = pc->iperm[A->itemL[j] - 1];
      if (jnew < inew) pc->itemL[cl++] = jnew; else pc->itemU[cu++] = jnew;
    }
    for (int32_t j = A->indexU[iold - 1]; j < A->indexU[iold]; ++j) {
      const int32_t jold = A->itemU[j];
      if (jold > N) continue;
      const int32_t jnew = pc->iperm[jold - 1];
      if (jnew < inew) pc->itemL[cl++] = jnew; else pc->itemU[cu++] = jnew;
    }
    pc->indexL[inew] = cl;
    pc->indexU[inew] = cu;
    qsort(pc->itemL + pc->indexL[inew - 1], (size_t)(cl - pc->indexL[inew - 1]), sizeof(int32_t), cmp_i32);
    qsort(pc->itemU + pc->indexU[inew - 1], (size_t)(cu - pc->indexU[inew - 1]), sizeof(int32_t), cmp_i32);
  }
  /* reorder_values: new (row, col) block = old (perm(row), perm(col)) block */
  pc->AL = (double*)calloc(9 * ((size_t)cl + 1), sizeof(double));
  pc->AU = (double*)calloc(9 * ((size_t)cu + 1), sizeof(double));
  for (int pass = 0; pass < 2; ++pass) {
    int32_t* idx = pass ? pc->indexU : pc->indexL;
    int32_t* item = pass ? pc->itemU : pc->itemL;
    double* dst = pass ? pc->AU : pc->AL;
    for (int32_t inew = 1; inew <= N; ++inew) {
      const int32_t iold = pc->perm[inew - 1];
      for (int32_t q = idx[inew - 1]; q < idx[inew]; ++q) {
        const int32_t kold = pc->perm[item[q] - 1];
        const double* src = NULL;
        if (kold < iold) {
          for (int32_t j = A->indexL[iold - 1]; j < A->indexL[iold]; ++j)
            if (A->itemL[j] == kold) src = A->AL + 9 * (size_t)j;
        } else {
          for (int32_t j = A->indexU[iold - 1]; j < A->indexU[iold]; ++j)
            if (A->itemU[j] == kold) src = A->AU + 9 * (size_t)j;
        }
        if (src) memcpy(dst + 9 * (size_t)q, src, 9 * sizeof(double));
      }
    }
  }
  /* renum_item (hecmw_matrix_reorder.f90:143-158): new column ids -> old ids */
  for (int32_t q = 0; q < cl; ++q) pc->itemL[q] = pc->perm[pc->itemL[q] - 1];
  for (int32_t q = 0; q < cu; ++q) pc->itemU[q] = pc->perm[pc->itemU[q] - 1];
  pc->ALU = (double*)malloc(sizeof(double) * 9 * (size_t)N);
  for (int32_t inew = 1; inew <= N; ++inew)
    block_lu(A->D + 9 * (size_t)(pc->perm[inew - 1] - 1), sigma, pc->ALU + 9 * (size_t)(inew - 1));
}

/* hecmw_precond_SSOR_33_apply (hecmw_precond_SSOR_33.F90:273-415) on ZP in place */
static void ssor_apply(const rs_pc* pc, double* ZP) {
  for (int32_t ic = 1; ic <= pc->ncolor; ++ic)
    for (int32_t i = pc->colorindex[ic - 1] + 1; i <= pc->colorindex[ic]; ++i) {
      const int32_t iold = pc->perm[i - 1];
      double s[3] = {ZP[3 * iold - 3], ZP[3 * iold - 2], ZP[3 * iold - 1]};
      for (int32_t j = pc->indexL[i - 1]; j < pc->indexL[i]; ++j) {
        const int32_t k = pc->itemL[j];
        const double* a = pc->AL + 9 * (size_t)j;
        const double x1 = ZP[3 * k - 3], x2 = ZP[3 * k - 2], x3 = ZP[3 * k - 1];
        s[0] = s[0] - a[0] * x1 - a[1] * x2 - a[2] * x3;
        s[1] = s[1] - a[3] * x1 - a[4] * x2 - a[5] * x3;
        s[2] = s[2] - a[6] * x1 - a[7] * x2 - a[8] * x3;
      }
      lu_solve(pc->ALU + 9 * (size_t)(i - 1), s);
      ZP[3 * iold - 3] = s[0];
      ZP[3 * iold - 2] = s[1];
      ZP[3 * iold - 1] = s[2];
    }
  for (int32_t ic = pc->ncolor; ic >= 1; --ic)
    for (int32_t i = pc->colorindex[ic]; i >= pc->colorindex[ic - 1] + 1; --i) {
      double s[3] = {0.0, 0.0, 0.0};
      for (int32_t j = pc->indexU[i] - 1; j >= pc->indexU[i - 1]; --j) {
        const int32_t k = pc->itemU[j];
        const double* a = pc->AU + 9 * (size_t)j;
        const double x1 = ZP[3 * k - 3], x2 = ZP[3 * k - 2], x3 = ZP[3 * k - 1];
        s[0] = s[0] + a[0] * x1 + a[1] * x2 + a[2] * x3;
        s[1] = s[1] + a[3] * x1 + a[4] * x2 + a[5] * x3;
        s[2] = s[2] + a[6] * x1 + a[7] * x2 + a[8] * x3;
      }
      lu_solve(pc->ALU + 9 * (size_t)(i - 1), s);
      const int32_t iold = pc->perm[i - 1];
      ZP[3 * iold - 3] = ZP[3 * iold - 3] - s[0];
      ZP[3 * iold - 2] = ZP[3 * iold - 2] - s[1];
      ZP[3 * iold - 1] = ZP[3 * iold - 1] - s[2];
    }
}

static void pc_free(rs_pc* pc) {
  free(pc->perm);
  free(pc->iperm);
  free(pc->colorindex);
  free(pc->indexL);
  free(pc->itemL);
  free(pc->indexU);
  free(pc->itemU);
  free(pc->AL);
  free(pc->AU);
  free(pc->ALU);
  memset(pc, 0, sizeof(*pc));
}

/* hecmw_precond_setup (hecmw_precond.f90:30-58) for PRECOND 1,2 / 3 */
static int pc_setup(const rs_sys* A, int32_t precond, int32_t iterpremax, double sigma, int32_t ncolor_in, rs_pc* pc) {
  memset(pc, 0, sizeof(*pc));
  pc->N = A->N;
  if (iterpremax <= 0) return 0;
  if (precond == 1 || precond == 2) {
    pc->kind = 1;
    ssor_setup(A, sigma, ncolor_in, pc);
  } else if (precond == 3) {
    pc->kind = 3;
    pc->ALU = (double*)malloc(sizeof(double) * 9 * (size_t)A->N);
    for (int32_t i = 0; i < A->N; ++i) block_lu(A->D + 9 * (size_t)i, sigma, pc->ALU + 9 * (size_t)i);
  } else {
    return 1001;
  }
  return 0;
}

/* hecmw_precond_apply (hecmw_precond.f90:86-172) */
static void pc_apply(const rs_sys* A, const rs_pc* pc, int32_t iterpremax, const double* R, double* Z, double* ZP) {
  const int64_t n3 = 3 * (int64_t)A->N, np3 = 3 * (int64_t)A->NP;
  if (iterpremax <= 0) {
    for (int64_t i = 0; i < n3; ++i) Z[i] = R[i];
    return;
  }
  for (int64_t i = 0; i < n3; ++i) ZP[i] = R[i];
  for (int64_t i = n3; i < np3; ++i) ZP[i] = 0.0;
  for (int64_t i = 0; i < np3; ++i) Z[i] = 0.0;
  for (int32_t it = 1; it <= iterpremax; ++it) {
    if (pc->kind == 1) {
      ssor_apply(pc, ZP);
    } else {
      for (int32_t i = 0; i < A->N; ++i) lu_solve(pc->ALU + 9 * (size_t)i, ZP + 3 * (size_t)i);
    }
    for (int64_t i = 0; i < n3; ++i) Z[i] = Z[i] + ZP[i];
    if (it == iterpremax) break;
    matresid(A, Z, R, ZP);
  }
}

/* z = M^-1 r, fresh setup */
int32_t rs_precond(const rs_sys* A, int32_t precond, int32_t iterpremax, double sigma, int32_t ncolor_in,
                   const double* R, double* Z, int32_t* ncolor_out) {
  rs_pc pc;
  int e = pc_setup(A, precond, iterpremax, sigma, ncolor_in, &pc);
  if (e) return e;
  double* ZP = (double*)calloc(3 * (size_t)A->NP, sizeof(double));
  pc_apply(A, &pc, iterpremax, R, Z, ZP);
  if (ncolor_out) *ncolor_out = pc.ncolor;
  free(ZP);
  pc_free(&pc);
  return 0;
}

/* ---- Krylov ---------------------------------------------------------------- */

/* hecmw_solve_CG (hecmw1/src/solver/iterative/hecmw_solver_CG.f90:19-320).
 * params: maxit, precond, iterpremax, ncolor_in; rparams: tol, sigma.
 * out: iter, resid; hist (may be NULL, >= maxit entries) receives ITERLOG values. */
int32_t rs_solve_cg(const rs_sys* A, const double* B, double* X, const int32_t* params, const double* rparams,
                    int32_t* iter_out, double* resid_out, double* hist) {
  const int32_t maxit_in = params[0], precond = params[1], iterpremax = params[2], ncolor_in = params[3];
  const double tol = rparams[0], sigma = rparams[1];
  const int64_t n3 = 3 * (int64_t)A->N, np3 = 3 * (int64_t)A->NP;
  double* R = (double*)calloc(np3, sizeof(double));
  double* Z = (double*)calloc(np3, sizeof(double));
  double* P = (double*)calloc(np3, sizeof(double));
  double* WK = (double*)calloc(np3, sizeof(double));
  double* Q = Z; /* the reference keeps q in z's column (hecmw_solver_CG.f90:49-51) */
  rs_pc pc;
  int32_t error = pc_setup(A, precond, iterpremax, sigma, ncolor_in, &pc);
  if (error) {
    free(R);
    free(Z);
    free(P);
    free(WK);
    return error;
  }
  matresid(A, X, B, R);
  double bnrm2 = dot(n3, B, B);
  int32_t maxit = maxit_in, iter = 0;
  double resid = tol, rho = 0.0, rho1 = 0.0, beta = 0.0, alpha, c1, dnrm2;
  int32_t n_indef = 0;
  if (bnrm2 == 0.0) {
    maxit = 0;
    resid = 0.0;
    for (int64_t i = 0; i < np3; ++i) X[i] = 0.0;
  }
  for (iter = 1; iter <= maxit; ++iter) {
    pc_apply(A, &pc, iterpremax, R, Z, WK);
    rho = dot(n3, R, Z);
    if (rho == 0.0) break;
    else if (iter > 1 && rho * rho1 <= 0) {
      ++n_indef;
      if (n_indef >= 3) {
        error = 3003;
        break;
      }
    } else if (rho != rho) {
      error = 3004;
      break;
    }
    if (iter == 1) {
      for (int64_t i = 0; i < n3; ++i) P[i] = Z[i];
    } else {
      beta = rho / rho1;
      for (int64_t i = 0; i < n3; ++i) P[i] = Z[i] + beta * P[i];
    }
    rs_matvec(A, P, Q);
    c1 = dot(n3, P, Q);
    if (c1 <= 0) {
      error = 3002;
      break;
    } else if (c1 != c1) {
      error = 3004;
      break;
    }
    alpha = rho / c1;
    for (int64_t i = 0; i < n3; ++i) X[i] = alpha * P[i] + X[i];
    if (iter % 50 == 0) {
      matresid(A, X, B, R);
    } else {
      for (int64_t i = 0; i < n3; ++i) R[i] = -alpha * Q[i] + R[i];
    }
    dnrm2 = dot(n3, R, R);
    resid = sqrt(dnrm2 / bnrm2);
    if (hist) hist[iter - 1] = resid;
    if (resid <= tol) {
      if (iter % 50 == 0) break;
      matresid(A, X, B, R);
      dnrm2 = dot(n3, R, R);
      resid = sqrt(dnrm2 / bnrm2);
      if (resid <= tol) break;
    }
    if (iter == maxit) error = 3001;
    rho1 = rho;
  }
  *iter_out = iter;
  *resid_out = resid;
  pc_free(&pc);
  free(R);
  free(Z);
  free(P);
  free(WK);
  return error;
}

/* hecmw_solve_BiCGSTAB (hecmw_solver_BiCGSTAB.f90:16-308) */
int32_t rs_solve_bicgstab(const rs_sys* A, const double* B, double* X, const int32_t* params, const double* rparams,
                          int32_t* iter_out, double* resid_out, double* hist) {
  const int32_t maxit_in = params[0], precond = params[1], iterpremax = params[2], ncolor_in = params[3];
  const double tol = rparams[0], sigma = rparams[1];
  const int64_t n3 = 3 * (int64_t)A->N, np3 = 3 * (int64_t)A->NP;
  double* W = (double*)calloc(8 * np3, sizeof(double));
  double *R = W, *RT = W + np3, *P = W + 2 * np3, *PT = W + 3 * np3, *S = W + 4 * np3, *ST = R, *T = W + 5 * np3,
         *V = W + 6 * np3, *WK = W + 7 * np3;
  rs_pc pc;
  int32_t error = pc_setup(A, precond, iterpremax, sigma, ncolor_in, &pc);
  if (error) {
    free(W);
    return error;
  }
  matresid(A, X, B, R);
  for (int64_t i = 0; i < n3; ++i) RT[i] = R[i];
  double bnrm2 = dot(n3, B, B);
  int32_t maxit = maxit_in, iter;
  double resid = tol, rho = 0.0, rho1 = 0.0, beta, alpha = 0.0, omega = 0.0, c2, dnrm2;
  if (bnrm2 == 0.0) {
    maxit = 0;
    resid = 0.0;
    for (int64_t i = 0; i < np3; ++i) X[i] = 0.0;
  }
  for (iter = 1; iter <= maxit; ++iter) {
    rho = dot(n3, R, RT);
    if (iter > 1) {
      beta = (rho / rho1) * (alpha / omega);
      for (int64_t i = 0; i < n3; ++i) P[i] = -omega * V[i] + P[i];
      for (int64_t i = 0; i < n3; ++i) P[i] = R[i] + beta * P[i];
    } else {
      for (int64_t i = 0; i < n3; ++i) P[i] = R[i];
    }
    pc_apply(A, &pc, iterpremax, P, PT, WK);
    rs_matvec(A, PT, V);
    c2 = dot(n3, RT, V);
    alpha = rho / c2;
    for (int64_t i = 0; i < n3; ++i) S[i] = -alpha * V[i] + R[i];
    pc_apply(A, &pc, iterpremax, S, ST, WK);
    rs_matvec(A, ST, T);
    const double ts = dot(n3, T, S), tt = dot(n3, T, T);
    omega = ts / tt;
    for (int64_t i = 0; i < n3; ++i) X[i] = alpha * PT[i] + X[i];
    for (int64_t i = 0; i < n3; ++i) X[i] = omega * ST[i] + X[i];
    if (iter % 100 == 0) {
      matresid(A, X, B, R);
    } else {
      for (int64_t i = 0; i < n3; ++i) R[i] = -omega * T[i] + S[i];
    }
    dnrm2 = dot(n3, R, R);
    resid = sqrt(dnrm2 / bnrm2);
    if (hist) hist[iter - 1] = resid;
    if (resid <= tol) {
      if (iter % 100 == 0) break;
      matresid(A, X, B, R);
      dnrm2 = dot(n3, R, R);
      resid = sqrt(dnrm2 / bnrm2);
      if (resid <= tol) break;
    }
    if (iter == maxit) error = 3001;
    rho1 = rho;
  }
  *iter_out = iter;
  *resid_out = resid;
  pc_free(&pc);
  free(W);
  return error;
}
