// Synthetic TET10 (fistr1 element type 342) linear-elastic box systems in
// hecmw1's 3x3 block layout, for the batched multi-RHS solves.
//
// Mesh: an nx x ny x nz grid of hexahedral cells on [0,Lx]x[0,Ly]x[0,Lz],
// each cell split into the 6 tetrahedra of the Kuhn (Freudenthal) split, which
// share the cell's main diagonal and conform across cells.  The quadratic
// nodes of every tetrahedron (edge midpoints) lie on the lattice of half
// cells, and every point of that (2nx+1)(2ny+1)(2nz+1) lattice is a node:
// cell corners, edge midpoints, face centres (the split's face diagonals) and
// cell centres (the main diagonal).
//
// Shape functions are fistr1's 342 (fistr1/src/lib/element/tet10n.f90:11-75),
// in their own node order: vertices 1-4, then midpoints 5=(1,2) 6=(2,3)
// 7=(3,1) 8=(1,4) 9=(2,4) 10=(3,4) (a mesh file lists 5-7 as edges 2-3, 3-1,
// 1-2, hecmw's order; the element matrix is the same); stiffness B^T D B by the
// 4-point rule fistr1 uses for 342 (element.f90 NumOfQuadPoints, gauss3d5,
// weight 1/24; points a=(5-sqrt5)/20, b=(5+3 sqrt5)/20 computed here in closed
// form).  Profile: hecmw_mat_con's insertion order (elements in order, local
// nodes 1..10 in order, hecmw_mat_con.f90:21-107).  Dirichlet u=0 on the
// z=0 face by row/column zeroing with a unit pivot (hecmw_mat_ass_bc,
// hecmw_mat_ass.f90:328-463; with a zero prescribed value the right-hand side
// of the other rows is unchanged).  Loads: nrhs independent columns of N(0,1)
// nodal forces (x, y, z) on the z=Lz face, seeded splitmix64 + Box-Muller.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../../include/hecmw33.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct Grid {
  int nx, ny, nz;
  int64_t X, Y, Z;  // lattice extents 2n+1
  int32_t id(int64_t x, int64_t y, int64_t z) const { return (int32_t)(x + X * (y + Y * z)); }
};

// The 6 tetrahedra of a cell as corner bit patterns (bit0 x, bit1 y, bit2 z),
// ordered so that the fistr1 vertex order has a positive Jacobian.
struct TetShape {
  int corner[4];
};
std::vector<TetShape> kuhn_tets() {
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::vector<TetShape> t(6);
  for (int p = 0; p < 6; ++p) {
    int c0 = 0, c1 = 1 << perm[p][0], c2 = c1 | (1 << perm[p][1]), c3 = 7;
    // orientation: det[v1-v0, v2-v0, v3-v0] > 0
    auto vec = [](int c, int d) { return (double)((c >> d) & 1); };
    double m[3][3];
    const int cs[3] = {c1, c2, c3};
    for (int r = 0; r < 3; ++r)
      for (int d = 0; d < 3; ++d) m[r][d] = vec(cs[r], d) - vec(c0, d);
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det < 0) std::swap(c1, c2);
    t[p].corner[0] = c0;
    t[p].corner[1] = c1;
    t[p].corner[2] = c2;
    t[p].corner[3] = c3;
  }
  return t;
}

const int kMid[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};  // nodes 5..10

// 10 lattice points (doubled coordinates relative to the cell's origin corner)
void tet_points(const TetShape& t, int (&pt)[10][3]) {
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) pt[a][d] = 2 * ((t.corner[a] >> d) & 1);
  for (int m = 0; m < 6; ++m)
    for (int d = 0; d < 3; ++d) pt[4 + m][d] = (pt[kMid[m][0]][d] + pt[kMid[m][1]][d]) / 2;
}

// fistr1 342 shape-function derivatives (tet10n.f90:29-75)
void dshape(double xi, double et, double ze, double (&g)[10][3]) {
  const double a = 1.0 - xi - et - ze;
  g[0][0] = 1 - 4 * a; g[0][1] = 1 - 4 * a; g[0][2] = 1 - 4 * a;
  g[1][0] = 4 * xi - 1; g[1][1] = 0; g[1][2] = 0;
  g[2][0] = 0; g[2][1] = 4 * et - 1; g[2][2] = 0;
  g[3][0] = 0; g[3][1] = 0; g[3][2] = 4 * ze - 1;
  g[4][0] = 4 * (1 - 2 * xi - et - ze); g[4][1] = -4 * xi; g[4][2] = -4 * xi;
  g[5][0] = 4 * et; g[5][1] = 4 * xi; g[5][2] = 0;
  g[6][0] = -4 * et; g[6][1] = 4 * (1 - xi - 2 * et - ze); g[6][2] = -4 * et;
  g[7][0] = -4 * ze; g[7][1] = -4 * ze; g[7][2] = 4 * (1 - xi - et - 2 * ze);
  g[8][0] = 4 * ze; g[8][1] = 0; g[8][2] = 4 * xi;
  g[9][0] = 0; g[9][1] = 4 * ze; g[9][2] = 4 * et;
}

// 30x30 stiffness (dof order node, xyz) of a tetrahedron with physical node
// coordinates xyz, isotropic (E, nu), 4-point rule.
void tet10_stiffness(const double (&xyz)[10][3], double E, double nu, double* K) {
  std::memset(K, 0, sizeof(double) * 900);
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  const double qa = (5.0 - std::sqrt(5.0)) / 20.0, qb = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
  const double qp[4][3] = {{qa, qa, qa}, {qb, qa, qa}, {qa, qb, qa}, {qa, qa, qb}};
  for (int q = 0; q < 4; ++q) {
    double g[10][3];
    dshape(qp[q][0], qp[q][1], qp[q][2], g);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[r][d] = d x_d / d xi_r
    for (int a = 0; a < 10; ++a)
      for (int r = 0; r < 3; ++r)
        for (int d = 0; d < 3; ++d) J[r][d] += g[a][r] * xyz[a][d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    double dN[10][3];  // d N_a / d x_d = sum_r Ji[d][r] g[a][r]
    for (int a = 0; a < 10; ++a)
      for (int d = 0; d < 3; ++d) dN[a][d] = Ji[d][0] * g[a][0] + Ji[d][1] * g[a][1] + Ji[d][2] * g[a][2];
    const double w = det / 24.0;
    for (int a = 0; a < 10; ++a)
      for (int b = 0; b < 10; ++b)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            // (B_a^T D B_b)_{rc} for isotropic D
            double v = lam * dN[a][r] * dN[b][c] + mu * dN[a][c] * dN[b][r];
            if (r == c) v += mu * (dN[a][0] * dN[b][0] + dN[a][1] * dN[b][1] + dN[a][2] * dN[b][2]);
            K[(3 * a + r) * 30 + 3 * b + c] += v * w;
          }
  }
}

struct Mesh {
  Grid g;
  int64_t ncell, nelem;
  std::vector<TetShape> shapes;
  int pts[6][10][3];
  std::vector<int64_t> ptr;   // node -> incident elements (ascending)
  std::vector<int64_t> elems;
};

int32_t elem_node(const Mesh& m, int64_t e, int a) {
  const int64_t cell = e / 6;
  const int t = (int)(e % 6);
  const int64_t cx = cell % m.g.nx, cy = (cell / m.g.nx) % m.g.ny, cz = cell / ((int64_t)m.g.nx * m.g.ny);
  return m.g.id(2 * cx + m.pts[t][a][0], 2 * cy + m.pts[t][a][1], 2 * cz + m.pts[t][a][2]);
}

int build_mesh(int nx, int ny, int nz, Mesh& m) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  m.g = Grid{nx, ny, nz, 2 * (int64_t)nx + 1, 2 * (int64_t)ny + 1, 2 * (int64_t)nz + 1};
  const int64_t N = m.g.X * m.g.Y * m.g.Z;
  if (N > (int64_t)INT32_MAX / 4) return -1;
  m.ncell = (int64_t)nx * ny * nz;
  m.nelem = 6 * m.ncell;
  m.shapes = kuhn_tets();
  for (int t = 0; t < 6; ++t) tet_points(m.shapes[t], m.pts[t]);
  // incidence: count, prefix, fill (elements visited in increasing order)
  std::vector<int64_t> cnt(N + 1, 0);
  for (int64_t e = 0; e < m.nelem; ++e)
    for (int a = 0; a < 10; ++a) ++cnt[elem_node(m, e, a) + 1];
  for (int64_t v = 0; v < N; ++v) cnt[v + 1] += cnt[v];
  m.ptr = cnt;
  m.elems.assign(cnt[N], 0);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < m.nelem; ++e)
    for (int a = 0; a < 10; ++a) m.elems[pos[elem_node(m, e, a)]++] = e;
  return 0;
}

// hecmw_mat_con order for row v: incident elements ascending, local nodes
// 1..10, first occurrence kept.  emit(w) per new neighbour w != v (0-based).
template <class F>
void row_profile(const Mesh& m, int32_t v, std::vector<int32_t>& seen, F emit) {
  seen.clear();
  for (int64_t q = m.ptr[v]; q < m.ptr[v + 1]; ++q) {
    const int64_t e = m.elems[q];
    for (int a = 0; a < 10; ++a) {
      const int32_t w = elem_node(m, e, a);
      if (w == v) continue;
      if (std::find(seen.begin(), seen.end(), w) != seen.end()) continue;
      seen.push_back(w);
      emit(w);
    }
  }
}

}  // namespace

extern "C" int32_t hecmw33_gen_tet10_count(int32_t nx, int32_t ny, int32_t nz, int32_t* N, int32_t* NPL,
                                           int32_t* NPU) {
  Mesh m;
  if (build_mesh(nx, ny, nz, m) != 0) return -1;
  const int32_t n = (int32_t)(m.g.X * m.g.Y * m.g.Z);
  int64_t l = 0, u = 0;
#pragma omp parallel reduction(+ : l, u)
  {
    std::vector<int32_t> seen;
#pragma omp for schedule(dynamic, 256)
    for (int32_t v = 0; v < n; ++v)
      row_profile(m, v, seen, [&](int32_t w) {
        if (w < v) ++l; else ++u;
      });
  }
  if (l > INT32_MAX || u > INT32_MAX) return -1;
  *N = n;
  *NPL = (int32_t)l;
  *NPU = (int32_t)u;
  return 0;
}

// Fills a caller-allocated system (sizes from hecmw33_gen_tet10_count).
// B holds nrhs columns of 3N entries (column k at B + 3N*k).
extern "C" int32_t hecmw33_gen_tet10(int32_t nx, int32_t ny, int32_t nz, const double* box, double E, double nu,
                                     int32_t nrhs, uint64_t seed, double* D, double* AL, double* AU,
                                     int32_t* indexL, int32_t* itemL, int32_t* indexU, int32_t* itemU, double* B) {
  Mesh m;
  if (build_mesh(nx, ny, nz, m) != 0 || nrhs < 0 || !box) return -1;
  const int32_t N = (int32_t)(m.g.X * m.g.Y * m.g.Z);
  // ---- profile ----------------------------------------------------------------
  std::vector<int32_t> cl(N), cu(N);
#pragma omp parallel
  {
    std::vector<int32_t> seen;
#pragma omp for schedule(dynamic, 256)
    for (int32_t v = 0; v < N; ++v) {
      int32_t l = 0, u = 0;
      row_profile(m, v, seen, [&](int32_t w) {
        if (w < v) ++l; else ++u;
      });
      cl[v] = l;
      cu[v] = u;
    }
  }
  indexL[0] = indexU[0] = 0;
  for (int32_t v = 0; v < N; ++v) {
    indexL[v + 1] = indexL[v] + cl[v];
    indexU[v + 1] = indexU[v] + cu[v];
  }
#pragma omp parallel
  {
    std::vector<int32_t> seen;
#pragma omp for schedule(dynamic, 256)
    for (int32_t v = 0; v < N; ++v) {
      int32_t l = indexL[v], u = indexU[v];
      row_profile(m, v, seen, [&](int32_t w) {
        if (w < v) itemL[l++] = w + 1; else itemU[u++] = w + 1;
      });
    }
  }
  const int64_t NPL = indexL[N], NPU = indexU[N];
  // ---- element matrices: 6 shapes, one cell size ----------------------------
  const double hx = box[0] / m.g.nx, hy = box[1] / m.g.ny, hz = box[2] / m.g.nz;
  std::vector<double> Ke(6 * 900);
  for (int t = 0; t < 6; ++t) {
    double xyz[10][3];
    for (int a = 0; a < 10; ++a) {
      xyz[a][0] = 0.5 * hx * m.pts[t][a][0];
      xyz[a][1] = 0.5 * hy * m.pts[t][a][1];
      xyz[a][2] = 0.5 * hz * m.pts[t][a][2];
    }
    tet10_stiffness(xyz, E, nu, &Ke[900 * t]);
  }
  // ---- assembly, row by row (each thread writes only its own row) ------------
  std::memset(D, 0, sizeof(double) * 9 * (size_t)N);
  std::memset(AL, 0, sizeof(double) * 9 * (size_t)NPL);
  std::memset(AU, 0, sizeof(double) * 9 * (size_t)NPU);
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t v = 0; v < N; ++v) {
    for (int64_t q = m.ptr[v]; q < m.ptr[v + 1]; ++q) {
      const int64_t e = m.elems[q];
      const double* K = &Ke[900 * (e % 6)];
      int32_t nid[10];
      int av = -1;
      for (int a = 0; a < 10; ++a) {
        nid[a] = elem_node(m, e, a);
        if (nid[a] == v) av = a;
      }
      for (int b = 0; b < 10; ++b) {
        const int32_t w = nid[b];
        double* blk;
        if (w == v) {
          blk = D + 9 * (size_t)v;
        } else if (w < v) {
          int32_t t = indexL[v];
          while (itemL[t] != w + 1) ++t;
          blk = AL + 9 * (size_t)t;
        } else {
          int32_t t = indexU[v];
          while (itemU[t] != w + 1) ++t;
          blk = AU + 9 * (size_t)t;
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) blk[3 * r + c] += K[(3 * av + r) * 30 + 3 * b + c];
      }
    }
  }
  // ---- loads --------------------------------------------------------------------
  const size_t n3 = 3 * (size_t)N;
  for (int32_t k = 0; k < nrhs; ++k) {
    double* b = B + n3 * k;
    std::memset(b, 0, sizeof(double) * n3);
    Rng rng(seed * 1000003ull + (uint64_t)k);
    for (int64_t y = 0; y < m.g.Y; ++y)
      for (int64_t x = 0; x < m.g.X; ++x) {
        const int32_t v = m.g.id(x, y, m.g.Z - 1);
        for (int d = 0; d < 3; ++d) b[3 * (size_t)v + d] = rng.normal();
      }
  }
  // ---- u = 0 on z = 0 ----------------------------------------------------------------
  for (int64_t y = 0; y < m.g.Y; ++y)
    for (int64_t x = 0; x < m.g.X; ++x) {
      const int32_t v = m.g.id(x, y, 0);
      double* Dv = D + 9 * (size_t)v;
      for (int idof = 0; idof < 3; ++idof) {
        for (int32_t k = 0; k < nrhs; ++k) B[n3 * k + 3 * (size_t)v + idof] = 0.0;
        for (int c = 0; c < 3; ++c) Dv[3 * idof + c] = 0.0;
        for (int r = 0; r < 3; ++r) Dv[3 * r + idof] = (r == idof) ? 1.0 : 0.0;
        for (int32_t q = indexL[v]; q < indexL[v + 1]; ++q) {
          for (int c = 0; c < 3; ++c) AL[9 * (size_t)q + 3 * idof + c] = 0.0;
          const int32_t w = itemL[q] - 1;
          for (int32_t t = indexU[w]; t < indexU[w + 1]; ++t)
            if (itemU[t] == v + 1) {
              for (int r = 0; r < 3; ++r) AU[9 * (size_t)t + 3 * r + idof] = 0.0;
              break;
            }
        }
        for (int32_t q = indexU[v]; q < indexU[v + 1]; ++q) {
          for (int c = 0; c < 3; ++c) AU[9 * (size_t)q + 3 * idof + c] = 0.0;
          const int32_t w = itemU[q] - 1;
          for (int32_t t = indexL[w]; t < indexL[w + 1]; ++t)
            if (itemL[t] == v + 1) {
              for (int r = 0; r < 3; ++r) AL[9 * (size_t)t + 3 * r + idof] = 0.0;
              break;
            }
        }
      }
    }
  return 0;
}
