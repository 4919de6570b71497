// Synthetic linear-elastic systems in hecmw1's 3x3 block layout.
//
// Produces exactly the arrays fistr1 hands to hecmw_solve for a structured
// HEX8 (type 361) cube: the node-pair profile built the way hecmw_mat_con does
// (hecmw1/src/solver/matrix/hecmw_mat_con.f90:21-107: elements in order, for
// each element every ordered node pair, item inserted into the row's list if
// not yet there, so items are in element-traversal order, not sorted), element
// stiffness by 2x2x2 Gauss full integration (FORM361=FI, the textbook B^T D B
// of STF_C3, fistr1/src/lib/static_LIB_3d.f90:50-208), block scatter into
// D/AL/AU (hecmw_mat_ass_elem, hecmw_mat_ass.f90:32-53), consistent pressure
// loads on the top face, and Dirichlet elimination of the bottom face with the
// row/column zeroing and unit pivot of hecmw_mat_ass_bc
// (hecmw_mat_ass.f90:328-463).
//
// Material fields: uniform (E, nu); per-element log-uniform E in
// [emin, emax] and uniform nu in [numin, numax]; or two-phase spherical
// inclusions of contrast `contrast` filling a target volume fraction.
// Random numbers come from splitmix64 so a seed reproduces the same field on
// any platform.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../../include/hecmw33.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Local node order of HEX8 type 361: bottom face (zeta=-1) counter-clockwise,
// then the top face.
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Element stiffness of a cube of edge h split as K = lambda*Kl + mu*Km, both
// 24x24 row-major, dof order (node, xyz).
void cube_stiffness_parts(double h, double* Kl, double* Km) {
  std::memset(Kl, 0, sizeof(double) * 576);
  std::memset(Km, 0, sizeof(double) * 576);
  const double g = 1.0 / std::sqrt(3.0);
  const double detJ = (h / 2) * (h / 2) * (h / 2);
  for (int gp = 0; gp < 8; ++gp) {
    const double xi[3] = {(gp & 1) ? g : -g, (gp & 2) ? g : -g, (gp & 4) ? g : -g};
    double dN[8][3];
    for (int a = 0; a < 8; ++a) {
      double s[3];
      for (int d = 0; d < 3; ++d) s[d] = kCorner[a][d] ? 1.0 : -1.0;
      const double f0 = 1 + s[0] * xi[0], f1 = 1 + s[1] * xi[1], f2 = 1 + s[2] * xi[2];
      // dN/dx = (2/h) dN/dxi
      dN[a][0] = 0.125 * s[0] * f1 * f2 * (2.0 / h);
      dN[a][1] = 0.125 * f0 * s[1] * f2 * (2.0 / h);
      dN[a][2] = 0.125 * f0 * f1 * s[2] * (2.0 / h);
    }
    // B (6 x 24): strains xx, yy, zz, xy, yz, zx (engineering shear)
    double B[6][24];
    std::memset(B, 0, sizeof(B));
    for (int a = 0; a < 8; ++a) {
      B[0][3 * a + 0] = dN[a][0];
      B[1][3 * a + 1] = dN[a][1];
      B[2][3 * a + 2] = dN[a][2];
      B[3][3 * a + 0] = dN[a][1];
      B[3][3 * a + 1] = dN[a][0];
      B[4][3 * a + 1] = dN[a][2];
      B[4][3 * a + 2] = dN[a][1];
      B[5][3 * a + 0] = dN[a][2];
      B[5][3 * a + 2] = dN[a][0];
    }
    // D = lambda * m m^T + mu * diag(2,2,2,1,1,1)
    for (int i = 0; i < 24; ++i) {
      const double ti = B[0][i] + B[1][i] + B[2][i];
      for (int j = 0; j < 24; ++j) {
        const double tj = B[0][j] + B[1][j] + B[2][j];
        Kl[i * 24 + j] += ti * tj * detJ;
        Km[i * 24 + j] += (2 * (B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j]) +
                           B[3][i] * B[3][j] + B[4][i] * B[4][j] + B[5][i] * B[5][j]) * detJ;
      }
    }
  }
}

}  // namespace

extern "C" int32_t hecmw33_gen_hex8_count(int32_t n, int32_t* N, int32_t* NPL, int32_t* NPU) {
  if (n < 1) return -1;
  const int64_t nn = (int64_t)n + 1;
  const int64_t nodes = nn * nn * nn;
  if (nodes > INT32_MAX) return -1;
  // every node couples to its clipped 3x3x3 neighbourhood
  auto c1 = [&](int64_t m) { return 3 * m - 2; };  // sum over a line of min(3, ...) neighbours
  const int64_t per_dim = c1(nn);
  const int64_t nnzb = per_dim * per_dim * per_dim;
  const int64_t off = (nnzb - nodes) / 2;
  if (off > INT32_MAX) return -1;
  *N = (int32_t)nodes;
  *NPL = (int32_t)off;
  *NPU = (int32_t)off;
  return 0;
}

// Fills a caller-allocated system for an n^3 HEX8 cube on [0,1]^3.
//   mat_kind 0: uniform E = p[0], nu = p[1]
//   mat_kind 1: per-element E = exp(U[ln p[0], ln p[1]]), nu = U[p[2], p[3]]
//   mat_kind 2: matrix E = p[0] with spherical inclusions of E = p[0]*p[2],
//               nu = p[1] everywhere, radius uniform in {p[3]..p[4]} elements,
//               placed until volume fraction p[5] is reached
// Loads: unit pressure on z=1 (consistent nodal forces, -z).
// BCs: every dof of every z=0 node fixed to 0 (hecmw_mat_ass_bc).
// Arrays use hecmw1 conventions: index*(0:N) cumulative, items 1-based.
// D == nullptr fills only the profile (indexL/itemL/indexU/itemU).
extern "C" int32_t hecmw33_gen_hex8(int32_t n, int32_t mat_kind, const double* p, uint64_t seed,
                                    double* D, double* AL, double* AU, int32_t* indexL,
                                    int32_t* itemL, int32_t* indexU, int32_t* itemU, double* B,
                                    float* elem_E) {
  int32_t N, NPL, NPU;
  if (hecmw33_gen_hex8_count(n, &N, &NPL, &NPU) != 0) return -1;
  const int nn = n + 1;
  const double h = 1.0 / n;
  auto node_id = [&](int i, int j, int k) -> int32_t { return i + nn * (j + nn * k); };  // 0-based
  const int64_t ne = (int64_t)n * n * n;

  // ---- profile, built per row in hecmw_mat_con's insertion order ----------
  // Row v's list receives, for elements e in increasing order that contain v,
  // and for that element's local nodes k = 1..8 in order, node k if not yet
  // present.  Lists for different rows are independent, so rows run in
  // parallel.
  std::vector<int32_t> cntL(N), cntU(N);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < N; ++v) {
    const int vi = v % nn, vj = (v / nn) % nn, vk = v / (nn * nn);
    int l = 0, u = 0;
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
          const int a = vi + di, b = vj + dj, c = vk + dk;
          if (a < 0 || b < 0 || c < 0 || a >= nn || b >= nn || c >= nn || (di == 0 && dj == 0 && dk == 0)) continue;
          const int32_t w = node_id(a, b, c);
          if (w < v) ++l; else ++u;
        }
    cntL[v] = l;
    cntU[v] = u;
  }
  indexL[0] = 0;
  indexU[0] = 0;
  for (int32_t v = 0; v < N; ++v) {
    indexL[v + 1] = indexL[v] + cntL[v];
    indexU[v + 1] = indexU[v] + cntU[v];
  }
  if (indexL[N] != NPL || indexU[N] != NPU) return -2;
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < N; ++v) {
    const int vi = v % nn, vj = (v / nn) % nn, vk = v / (nn * nn);
    int32_t* L = itemL + indexL[v];
    int32_t* U = itemU + indexU[v];
    int l = 0, u = 0;
    // incident elements in increasing element index: e = ex + n*(ey + n*ez)
    for (int ez = vk - 1; ez <= vk; ++ez) {
      if (ez < 0 || ez >= n) continue;
      for (int ey = vj - 1; ey <= vj; ++ey) {
        if (ey < 0 || ey >= n) continue;
        for (int ex = vi - 1; ex <= vi; ++ex) {
          if (ex < 0 || ex >= n) continue;
          for (int a = 0; a < 8; ++a) {
            const int32_t w = node_id(ex + kCorner[a][0], ey + kCorner[a][1], ez + kCorner[a][2]);
            if (w == v) continue;
            const int32_t w1 = w + 1;
            if (w < v) {
              bool seen = false;
              for (int q = 0; q < l; ++q) seen |= (L[q] == w1);
              if (!seen) L[l++] = w1;
            } else {
              bool seen = false;
              for (int q = 0; q < u; ++q) seen |= (U[q] == w1);
              if (!seen) U[u++] = w1;
            }
          }
        }
      }
    }
  }

  // structure only (D == nullptr): the profile is all the orderings need
  if (!D) return 0;

  // ---- material field ------------------------------------------------------
  std::vector<double> lam(ne), mu(ne);
  {
    std::vector<float> Ef(ne), nuf(ne);
    if (mat_kind == 0) {
      std::fill(Ef.begin(), Ef.end(), (float)p[0]);
      std::fill(nuf.begin(), nuf.end(), (float)p[1]);
    } else if (mat_kind == 1) {
      SplitMix64 rng(seed);
      const double l0 = std::log(p[0]), l1 = std::log(p[1]);
      for (int64_t e = 0; e < ne; ++e) {
        Ef[e] = (float)std::exp(l0 + (l1 - l0) * rng.uniform());
        nuf[e] = (float)(p[2] + (p[3] - p[2]) * rng.uniform());
      }
    } else if (mat_kind == 2) {
      SplitMix64 rng(seed);
      std::vector<uint8_t> phase(ne, 0);
      int64_t filled = 0;
      const int64_t target = (int64_t)std::llround(p[5] * (double)ne);
      const int rmin = (int)p[3], rmax = (int)p[4];
      int guard = 0;
      while (filled < target && guard < 10000000) {
        ++guard;
        const int r = rmin + (int)(rng.next() % (uint64_t)(rmax - rmin + 1));
        const double cx = rng.uniform() * n, cy = rng.uniform() * n, cz = rng.uniform() * n;
        const int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(n - 1, (int)std::ceil(cx + r));
        const int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(n - 1, (int)std::ceil(cy + r));
        const int z0 = std::max(0, (int)std::floor(cz - r)), z1 = std::min(n - 1, (int)std::ceil(cz + r));
        for (int z = z0; z <= z1 && filled < target; ++z)
          for (int y = y0; y <= y1 && filled < target; ++y)
            for (int x = x0; x <= x1 && filled < target; ++x) {
              const double dx = x + 0.5 - cx, dy = y + 0.5 - cy, dz = z + 0.5 - cz;
              if (dx * dx + dy * dy + dz * dz > (double)r * r) continue;
              const int64_t e = x + (int64_t)n * (y + (int64_t)n * z);
              if (!phase[e]) { phase[e] = 1; ++filled; }
            }
      }
      for (int64_t e = 0; e < ne; ++e) {
        Ef[e] = (float)(phase[e] ? p[0] * p[2] : p[0]);
        nuf[e] = (float)p[1];
      }
    } else {
      return -3;
    }
    for (int64_t e = 0; e < ne; ++e) {
      const double E = Ef[e], nu = nuf[e];
      lam[e] = E * nu / ((1 + nu) * (1 - 2 * nu));
      mu[e] = E / (2 * (1 + nu));
      if (elem_E) elem_E[e] = Ef[e];
    }
  }

  // ---- assembly: 8-colour element parity classes keep writes disjoint -----
  std::memset(D, 0, sizeof(double) * 9 * (size_t)N);
  std::memset(AL, 0, sizeof(double) * 9 * (size_t)NPL);
  std::memset(AU, 0, sizeof(double) * 9 * (size_t)NPU);
  double Kl[576], Km[576];
  cube_stiffness_parts(h, Kl, Km);
  for (int par = 0; par < 8; ++par) {
#pragma omp parallel for schedule(static) collapse(2)
    for (int ez = (par >> 2) & 1; ez < n; ez += 2)
      for (int ey = (par >> 1) & 1; ey < n; ey += 2)
        for (int ex = par & 1; ex < n; ex += 2) {
          const int64_t e = ex + (int64_t)n * (ey + (int64_t)n * ez);
          const double le = lam[e], me = mu[e];
          int32_t nid[8];
          for (int a = 0; a < 8; ++a) nid[a] = node_id(ex + kCorner[a][0], ey + kCorner[a][1], ez + kCorner[a][2]);
          for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
              double* blk;
              const int32_t va = nid[a], vb = nid[b];
              if (va == vb) {
                blk = D + 9 * (size_t)va;
              } else if (vb < va) {
                int32_t q = indexL[va];
                while (itemL[q] != vb + 1) ++q;
                blk = AL + 9 * (size_t)q;
              } else {
                int32_t q = indexU[va];
                while (itemU[q] != vb + 1) ++q;
                blk = AU + 9 * (size_t)q;
              }
              for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                  const int ij = (3 * a + r) * 24 + 3 * b + c;
                  blk[3 * r + c] += le * Kl[ij] + me * Km[ij];
                }
            }
        }
  }

  // ---- loads: unit pressure on z = 1, each face quarter to each corner -----
  std::memset(B, 0, sizeof(double) * 3 * (size_t)N);
  const double fq = 0.25 * h * h;
  for (int ey = 0; ey < n; ++ey)
    for (int ex = 0; ex < n; ++ex)
      for (int a = 0; a < 4; ++a) {
        const int32_t v = node_id(ex + kCorner[a][0], ey + kCorner[a][1], n);
        B[3 * (size_t)v + 2] -= fq;
      }

  // ---- Dirichlet u = 0 on z = 0 (hecmw_mat_ass_bc, value 0) ----------------
  for (int j = 0; j < nn; ++j)
    for (int i = 0; i < nn; ++i) {
      const int32_t v = node_id(i, j, 0);
      for (int idof = 0; idof < 3; ++idof) {
        const double rhs = 0.0;
        double* Dv = D + 9 * (size_t)v;
        B[3 * (size_t)v + idof] = rhs;
        for (int r = 0; r < 3; ++r)
          if (r != idof) B[3 * (size_t)v + r] -= Dv[3 * r + idof] * rhs;
        for (int c = 0; c < 3; ++c) Dv[3 * idof + c] = 0.0;
        for (int r = 0; r < 3; ++r) Dv[3 * r + idof] = (r == idof) ? 1.0 : 0.0;
        for (int32_t q = indexL[v]; q < indexL[v + 1]; ++q) {
          for (int c = 0; c < 3; ++c) AL[9 * (size_t)q + 3 * idof + c] = 0.0;
          const int32_t w = itemL[q] - 1;
          for (int32_t t = indexU[w]; t < indexU[w + 1]; ++t)
            if (itemU[t] == v + 1) {
              for (int r = 0; r < 3; ++r) {
                B[3 * (size_t)w + r] -= AU[9 * (size_t)t + 3 * r + idof] * rhs;
                AU[9 * (size_t)t + 3 * r + idof] = 0.0;
              }
              break;
            }
        }
        for (int32_t q = indexU[v]; q < indexU[v + 1]; ++q) {
          for (int c = 0; c < 3; ++c) AU[9 * (size_t)q + 3 * idof + c] = 0.0;
          const int32_t w = itemU[q] - 1;
          for (int32_t t = indexL[w]; t < indexL[w + 1]; ++t)
            if (itemL[t] == v + 1) {
              for (int r = 0; r < 3; ++r) {
                B[3 * (size_t)w + r] -= AL[9 * (size_t)t + 3 * r + idof] * rhs;
                AL[9 * (size_t)t + 3 * r + idof] = 0.0;
              }
              break;
            }
        }
      }
    }
  return 0;
}
