// Canonical relabelling of tet-10 / tet-4 elements.
//
// A row-gather pair (node i, element e) computes the row of node i in e's
// element matrix.  The element is relabelled first so that node i sits at local
// position 0 (vertex class) or 4 (edge (0,1), edge class); the physical integrals
// do not depend on the labelling, so only two specialised row kernels exist.
// PERM[li][c] = original FEDDLib local index (MeshStructured_def.hpp:906-987
// edge order 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(1,3) 9:(2,3)) found at canonical
// position c.  Vertices 0..3 of the relabelled element are the pressure (P1)
// vertices in the same order.
#pragma once
#include <cstdint>

#define FEDD_PERM_P2_INIT {{0, 1, 2, 3, 4, 5, 6, 7, 8, 9}, {1, 0, 2, 3, 4, 6, 5, 8, 7, 9}, {2, 1, 0, 3, 5, 4, 6, 9, 8, 7}, {3, 1, 2, 0, 8, 5, 9, 7, 4, 6}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9}, {1, 2, 0, 3, 5, 6, 4, 8, 9, 7}, {0, 2, 1, 3, 6, 5, 4, 7, 9, 8}, {0, 3, 1, 2, 7, 8, 4, 6, 9, 5}, {1, 3, 0, 2, 8, 7, 4, 5, 9, 6}, {2, 3, 0, 1, 9, 7, 6, 5, 8, 4}}
#define FEDD_PERM_P1_INIT {{0, 1, 2, 3}, {1, 0, 2, 3}, {2, 1, 0, 3}, {3, 1, 2, 0}}

// Relabelling acts on the affine map as B' = B Q (B[:,j] = x_{j+1} - x_0), so
// Binv' = QINV[li] Binv with the integer matrices below (rows of Binv' are signed
// sums of rows of Binv).
#define FEDD_QINV_INIT {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{-1, -1, -1}, {0, 1, 0}, {0, 0, 1}}, {{1, 0, 0}, {-1, -1, -1}, {0, 0, 1}}, {{1, 0, 0}, {0, 1, 0}, {-1, -1, -1}}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{0, 1, 0}, {-1, -1, -1}, {0, 0, 1}}, {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}}, {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}}, {{0, 0, 1}, {-1, -1, -1}, {0, 1, 0}}, {{0, 0, 1}, {-1, -1, -1}, {1, 0, 0}}}

namespace fedd {
static constexpr int8_t kPermP2[10][10] = FEDD_PERM_P2_INIT;
static constexpr int8_t kPermP1[4][4] = FEDD_PERM_P1_INIT;
static constexpr int8_t kQinv[10][3][3] = FEDD_QINV_INIT;
}  // namespace fedd
