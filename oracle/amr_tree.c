/*
 * TEST INFRASTRUCTURE (the oracle): the block table of a fixed block tree, restated from
 * src/amr/tree.ts (neighbor :293-324, writeRow :652-675). Given the blocks as (level, ix, iy, iz)
 * it builds per block: leaf flag, parent, the six face neighbours (a same-level block, -1 where
 * the region beyond is one level coarser, or -1 - boundary code at a closed face of the domain),
 * the eight children, and the coarser leaf beyond each -1 face. In 2D (ndim 2, one root block
 * along z, every iz 0) a block has four children (octant x | y << 1) and each z face names the
 * block itself (tree.ts writeRow :662-666: the single layer is its own neighbour in z).
 *
 * Parity: pinned by the tree invariants the reference's tests check (tests/unit/tree.test.ts:
 * 2:1 balance, every point of the domain covered by exactly one leaf) in tests/test_oracle_amr.py,
 * and by the GPU library building the same tables independently (compared entry by entry).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int find(const orc_amr* t, int level, int x, int y, int z) {
  for (int s = 0; s < t->nblocks; s++) {
    const int* b = t->blk + 4 * s;
    if (b[0] == level && b[1] == x && b[2] == y && b[3] == z) return s;
  }
  return -1;
}

int orc_amr_build(orc_amr* t, const int* blocks, int nblocks, const int root[3], const int bc_lo[3],
                  const int bc_hi[3]) {
  return orc_amr_build_nd(t, blocks, nblocks, root, bc_lo, bc_hi, 3);
}

int orc_amr_build_nd(orc_amr* t, const int* blocks, int nblocks, const int root[3], const int bc_lo[3],
                     const int bc_hi[3], int ndim) {
  memset(t, 0, sizeof *t);
  t->nblocks = nblocks;
  t->ndim = ndim == 2 ? 2 : 3;
  const int nchild = 1 << t->ndim, nface = 2 * t->ndim;
  for (int a = 0; a < 3; a++) { t->root[a] = root[a]; t->bc_lo[a] = bc_lo[a]; t->bc_hi[a] = bc_hi[a]; }
  t->blk = (int*)malloc(sizeof(int) * 4 * nblocks);
  t->tab = (int*)malloc(sizeof(int) * ORC_BT_STRIDE * nblocks);
  memcpy(t->blk, blocks, sizeof(int) * 4 * nblocks);
  for (int s = 0; s < nblocks; s++) {
    int* r = t->tab + ORC_BT_STRIDE * s;
    for (int k = 0; k < ORC_BT_STRIDE; k++) r[k] = -1;
    const int* b = t->blk + 4 * s;
    int L = b[0];
    r[0] = L; r[1] = b[1]; r[2] = b[2]; r[3] = b[3];
    r[5] = L > 0 ? find(t, L - 1, b[1] >> 1, b[2] >> 1, b[3] >> 1) : -1;
    int leaf = 1;
    for (int o = 0; o < nchild; o++) {
      int ch = find(t, L + 1, 2 * b[1] + (o & 1), 2 * b[2] + ((o >> 1) & 1), 2 * b[3] + ((o >> 2) & 1));
      r[12 + o] = ch;
      if (ch >= 0) leaf = 0;
    }
    r[4] = 1 | (leaf ? 2 : 0);
    for (int f = 0; f < 6; f++) {
      if (f >= nface) { r[6 + f] = s; continue; }
      int axis = f >> 1, up = f & 1;
      int c[3] = {b[1], b[2], b[3]};
      int n = root[axis] << L;
      int v = c[axis] + (up ? 1 : -1);
      if (v < 0 || v >= n) {
        int type = up ? bc_hi[axis] : bc_lo[axis];
        if (type != 0) { r[6 + f] = -1 - type; continue; }
        v = (v + n) % n;
      }
      c[axis] = v;
      int same = find(t, L, c[0], c[1], c[2]);
      if (same >= 0) { r[6 + f] = same; continue; }
      int coarse = L > 0 ? find(t, L - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1) : -1;
      if (coarse < 0) return -1 - s;     /* 2:1 balance broken */
      r[6 + f] = -1;
      r[20 + f] = coarse;
    }
  }
  return 0;
}

void orc_amr_free(orc_amr* t) {
  free(t->blk);
  free(t->tab);
  t->blk = t->tab = NULL;
}

const int* orc_amr_table(const orc_amr* t) { return t->tab; }
