// Host-side symbolic setup of the multicolour SSOR ordering (see ordering.cpp).
#pragma once
#include <cstdint>
#include <vector>

namespace hecmw33 {

// Reverse Cuthill-McKee permutation of the internal rows (1-based ids stored in
// 0-based arrays): perm[k] = old id at new position k+1, iperm[old-1] = new id.
void ordering_rcm(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                  const int32_t* itemU, int32_t* perm, int32_t* iperm);

// Capped greedy multicolouring in the order perm_cur; colorindex has N+1
// slots, colour c holds new positions colorindex[c-1]+1 .. colorindex[c].
// Returns the number of colours.
int32_t ordering_mc(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                    const int32_t* itemU, const int32_t* perm_cur, int32_t ncolor_in,
                    int32_t* colorindex, int32_t* perm, int32_t* iperm);

// L'/U' profile of the permuted internal matrix, items are new ids (1-based),
// ascending within each row.
void reorder_profile(int32_t N, const int32_t* perm, const int32_t* iperm, const int32_t* indexL,
                     const int32_t* itemL, const int32_t* indexU, const int32_t* itemU,
                     std::vector<int32_t>& indexLp, std::vector<int32_t>& itemLp,
                     std::vector<int32_t>& indexUp, std::vector<int32_t>& itemUp);

}  // namespace hecmw33
