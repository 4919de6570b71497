// Host-side symbolic setup of the multicolour SSOR preconditioner.
//
// The colouring is a contract with hecmw1: the GPU sweeps must visit rows in
// exactly the order hecmw_precond_SSOR_33_setup builds
// (hecmw1/src/solver/precond/ssor/hecmw_precond_SSOR_33.F90:93-155), so every
// step below reproduces the reference's integer decisions:
//   * RCM (hecmw1/src/solver/matrix/hecmw_matrix_ordering_CM.f90:18-245):
//     degree = L count + U count of internal neighbours; up to five start
//     nodes = the first nodes of minimum non-zero degree in index order;
//     breadth-first level sets, each level sorted by degree with a recursive
//     Hoare partition around the middle element (not stable: the tie order it
//     leaves is part of the contract); the start with the most levels wins,
//     first one on ties; then reversed.
//   * multicolouring (hecmw_matrix_ordering_MC.f90:15-72): greedy in RCM
//     order, each colour capped at N/ncolor_in nodes; the scan for a colour
//     stops when the cap is hit or all nodes are coloured *before* the last
//     node's neighbours are marked.
//   * reordered profile (hecmw_matrix_reorder.f90:20-66): for new row i the
//     old row's internal neighbours split into L' (new index < i) and U'
//     (new index > i), each sorted ascending by new index.
// Everything runs on the CPU once per symbolic setup; the result is uploaded.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ordering.h"

namespace hecmw33 {

namespace {

// Hoare partition quicksort of idx[lo..hi] (inclusive) by key[idx[.]],
// pivot = key of the middle element.  Reproduces the tie order of
// sort_nodes_by_degree (hecmw_matrix_ordering_CM.f90:207-233); recursion
// is replaced by an explicit stack with the same visiting order of the two
// halves, which does not affect the result (the halves are disjoint).
void hoare_sort_by_key(int32_t* idx, int64_t lo, int64_t hi, const int32_t* key) {
  std::vector<std::pair<int64_t, int64_t>> stack;
  stack.emplace_back(lo, hi);
  while (!stack.empty()) {
    auto [s, e] = stack.back();
    stack.pop_back();
    if (s >= e) continue;
    const int32_t pivot = key[idx[(s + e) / 2]];
    int64_t l = s, r = e;
    for (;;) {
      while (key[idx[l]] < pivot) ++l;
      while (pivot < key[idx[r]]) --r;
      if (l >= r) break;
      std::swap(idx[l], idx[r]);
      ++l;
      --r;
    }
    stack.emplace_back(r + 1, e);
    stack.emplace_back(s, l - 1);
  }
}

// Same partition scheme on plain integer values (hecmw_qsort_int_array,
// hecmw1/src/solver/matrix/hecmw_array_util.f90:17-44).  Values inside one
// row are distinct, so any correct sort gives the same result; the
// partition scheme is kept anyway.
void hoare_sort_values(int32_t* a, int64_t lo, int64_t hi) {
  if (hi - lo < 16) {  // short rows: insertion sort gives the same (unique) order
    for (int64_t i = lo + 1; i <= hi; ++i) {
      int32_t v = a[i];
      int64_t j = i - 1;
      while (j >= lo && a[j] > v) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = v;
    }
    return;
  }
  std::sort(a + lo, a + hi + 1);
}

}  // namespace

// perm/iperm are 1-based node ids stored in 0-based arrays: perm[k] = old id of
// the (k+1)-th node in the new order.
void ordering_rcm(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                  const int32_t* itemU, int32_t* perm, int32_t* iperm) {
  // degrees (count_degrees)
  std::vector<int32_t> deg(N + 1, 0);
  for (int32_t i = 1; i <= N; ++i) {
    int32_t d = indexL[i] - indexL[i - 1];
    for (int32_t j = indexU[i - 1]; j < indexU[i]; ++j)
      if (itemU[j] <= N) ++d;
    deg[i] = d;
  }
  // candidate start nodes (find_minimum_degrees)
  const int kMaxStarts = 5;
  int32_t starts[kMaxStarts];
  int32_t nmin = 0, degmin = N;
  for (int32_t i = 1; i <= N; ++i) {
    if (deg[i] == 0) continue;
    if (deg[i] < degmin) {
      degmin = deg[i];
      nmin = 1;
      starts[0] = i;
    } else if (deg[i] == degmin) {
      ++nmin;
      if (nmin <= kMaxStarts) starts[nmin - 1] = i;
    }
  }
  if (nmin > kMaxStarts) nmin = kMaxStarts;

  // one Cuthill-McKee level structure per start (ordering_CM_inner)
  std::vector<std::vector<int32_t>> order(nmin);
  std::vector<int32_t> nlevel(nmin, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nmin; ++s) {
    std::vector<int32_t> visited(N + 1, 0);
    std::vector<int32_t>& item = order[s];
    item.assign(N, 0);
    int64_t prev_begin = 0, prev_end = 1;  // [begin, end) of the previous level in item
    visited[starts[s]] = 1;
    item[0] = starts[s];
    int64_t cntall = 1;
    int32_t level = 1;
    if (cntall == N) nlevel[s] = 1;
    for (level = 2; level <= N && cntall < N; ++level) {
      const int64_t level_begin = cntall;
      for (int64_t j = prev_begin; j < prev_end; ++j) {
        const int32_t jn = item[j];
        for (int32_t k = indexL[jn - 1]; k < indexL[jn]; ++k) {
          const int32_t kn = itemL[k];
          if (!visited[kn]) {
            visited[kn] = level;
            item[cntall++] = kn;
          }
        }
        for (int32_t k = indexU[jn - 1]; k < indexU[jn]; ++k) {
          const int32_t kn = itemU[k];
          if (kn > N) continue;
          if (!visited[kn]) {
            visited[kn] = level;
            item[cntall++] = kn;
          }
        }
      }
      if (cntall == level_begin) {  // disconnected: take the first unvisited node
        for (int32_t kn = 1; kn <= N; ++kn)
          if (!visited[kn]) {
            visited[kn] = level;
            item[cntall++] = kn;
            break;
          }
      }
      hoare_sort_by_key(item.data(), level_begin, cntall - 1, deg.data());
      prev_begin = level_begin;
      prev_end = cntall;
      if (cntall == N) {
        nlevel[s] = level;
        break;
      }
    }
  }
  int best = 0;
  for (int s = 1; s < nmin; ++s)
    if (nlevel[s] > nlevel[best]) best = s;
  // reverse (reverse_ordering): new position N+1-k
  for (int32_t k = 0; k < N; ++k) {
    const int32_t old = order[best][k];
    const int32_t newpos = N - k;  // 1-based
    iperm[old - 1] = newpos;
    perm[newpos - 1] = old;
  }
  if (nmin == 0) {  // no connected node at all: identity
    for (int32_t k = 0; k < N; ++k) {
      perm[k] = k + 1;
      iperm[k] = k + 1;
    }
  }
}

int32_t ordering_mc(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                    const int32_t* itemU, const int32_t* perm_cur, int32_t ncolor_in,
                    int32_t* colorindex, int32_t* perm, int32_t* iperm) {
  std::vector<int32_t> mark(N + 1, 0);  // >0 coloured, -1 blocked for this colour
  const int32_t cap = ncolor_in > 0 ? N / ncolor_in : 0;
  int32_t cntall = 0, ncolor = 0;
  colorindex[0] = 0;
  std::vector<int32_t> blocked;
  // nodes still uncoloured, kept in scan order so each colour pass skips the
  // coloured ones without changing the visiting order
  std::vector<int32_t> pending(perm_cur, perm_cur + N);
  std::vector<int32_t> next_pending;
  for (int32_t color = 1; color <= N; ++color) {
    int32_t cnt = 0;
    next_pending.clear();
    size_t p = 0;
    for (; p < pending.size(); ++p) {
      const int32_t in = pending[p];
      if (mark[in] > 0) continue;
      if (mark[in] == -1) {
        next_pending.push_back(in);
        continue;
      }
      mark[in] = color;
      perm[cntall++] = in;
      ++cnt;
      if (cnt == cap || cntall == N) {
        ++p;
        break;
      }
      for (int32_t j = indexL[in - 1]; j < indexL[in]; ++j) {
        const int32_t jn = itemL[j];
        if (mark[jn] == 0) {
          mark[jn] = -1;
          blocked.push_back(jn);
        }
      }
      for (int32_t j = indexU[in - 1]; j < indexU[in]; ++j) {
        const int32_t jn = itemU[j];
        if (jn > N) continue;
        if (mark[jn] == 0) {
          mark[jn] = -1;
          blocked.push_back(jn);
        }
      }
    }
    for (; p < pending.size(); ++p)
      if (mark[pending[p]] <= 0) next_pending.push_back(pending[p]);
    pending.swap(next_pending);
    colorindex[color] = cntall;
    if (cntall == N) {
      ncolor = color;
      break;
    }
    for (int32_t b : blocked)
      if (mark[b] == -1) mark[b] = 0;
    blocked.clear();
  }
  for (int32_t k = 0; k < N; ++k) iperm[perm[k] - 1] = k + 1;
  return ncolor;
}

void reorder_profile(int32_t N, const int32_t* perm, const int32_t* iperm, const int32_t* indexL,
                     const int32_t* itemL, const int32_t* indexU, const int32_t* itemU,
                     std::vector<int32_t>& indexLp, std::vector<int32_t>& itemLp,
                     std::vector<int32_t>& indexUp, std::vector<int32_t>& itemUp) {
  indexLp.assign(N + 1, 0);
  indexUp.assign(N + 1, 0);
  for (int32_t inew = 1; inew <= N; ++inew) {
    const int32_t iold = perm[inew - 1];
    int32_t l = 0, u = 0;
    for (int32_t j = indexL[iold - 1]; j < indexL[iold]; ++j) {
      if (iperm[itemL[j] - 1] < inew) ++l; else ++u;
    }
    for (int32_t j = indexU[iold - 1]; j < indexU[iold]; ++j) {
      if (itemU[j] > N) continue;
      if (iperm[itemU[j] - 1] < inew) ++l; else ++u;
    }
    indexLp[inew] = l;
    indexUp[inew] = u;
  }
  for (int32_t i = 0; i < N; ++i) {
    indexLp[i + 1] += indexLp[i];
    indexUp[i + 1] += indexUp[i];
  }
  itemLp.assign(indexLp[N], 0);
  itemUp.assign(indexUp[N], 0);
#pragma omp parallel for schedule(static, 1024)
  for (int32_t inew = 1; inew <= N; ++inew) {
    const int32_t iold = perm[inew - 1];
    int32_t l = indexLp[inew - 1], u = indexUp[inew - 1];
    for (int32_t j = indexL[iold - 1]; j < indexL[iold]; ++j) {
      const int32_t jnew = iperm[itemL[j] - 1];
      if (jnew < inew) itemLp[l++] = jnew; else itemUp[u++] = jnew;
    }
    for (int32_t j = indexU[iold - 1]; j < indexU[iold]; ++j) {
      if (itemU[j] > N) continue;
      const int32_t jnew = iperm[itemU[j] - 1];
      if (jnew < inew) itemLp[l++] = jnew; else itemUp[u++] = jnew;
    }
    if (indexLp[inew] - indexLp[inew - 1] > 1) hoare_sort_values(itemLp.data(), indexLp[inew - 1], indexLp[inew] - 1);
    if (indexUp[inew] - indexUp[inew - 1] > 1) hoare_sort_values(itemUp.data(), indexUp[inew - 1], indexUp[inew] - 1);
  }
}

}  // namespace hecmw33
