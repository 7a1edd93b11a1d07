// C-ABI of libshx for the block grid (include/shx.h, shx_amr_*): tree validation, the block
// table and slot lists (built on the host once, as the reference's BlockTree and uploadTree do),
// device state and atlases, and the captured step graphs.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "amr.hpp"
#include "common.hpp"

using namespace shx;

namespace {

struct Range {
  int off = 0, count = 0;
};

}  // namespace

struct shx_amr {
  shx_config cfg;
  shx_amr_layout lay;
  int ndim = 3;
  int cells = AMR_CELLS;     // per block: 16^ndim cells, 20^ndim atlas texels
  int texels = AMR_TEXELS;
  int es = 4;
  int nblocks = 0;
  int nlevels = 0;
  size_t ncell = 0;
  std::vector<int> blocks;   // nblocks x 4
  std::vector<int> tab;      // nblocks x AMR_BT
  hipStream_t stream = nullptr;
  int* d_tab = nullptr;
  int* d_lists = nullptr;
  // ranges into d_lists: every block, leaves, leaves without / with a refined neighbour; the
  // same split per level
  Range all, leaf, plain, match;
  Range queue;   // the fused march's queue (-1 pads)
  Range all_lv[AMR_MAXLEV], leaf_lv[AMR_MAXLEV], plain_lv[AMR_MAXLEV], match_lv[AMR_MAXLEV];
  void* U1[2] = {nullptr, nullptr};
  void* U2[2] = {nullptr, nullptr};
  void *W1 = nullptr, *W2 = nullptr, *PA = nullptr, *PB = nullptr;
  // fused path: ghost shells (one per leaf, in leaf-list order) and each slot's shell index;
  // flux-register index per (slot, face) and the register
  void* shell = nullptr;
  int* d_shell_of = nullptr;
  // the shell's waves of work (shell index, slot, face, chunk): ghost texels, half steps
  int* d_shw[3] = {nullptr, nullptr, nullptr};   // the shell's work: texels, half steps, rim records
  int nshw[3] = {0, 0, 0};
  int* d_reg_of = nullptr;
  void* reg = nullptr;
  // coarse-fine faces seen from the coarse side: (slot, face) pairs by level, their buffers, and
  // the march's per-(slot, face) action (register write, coarse-face read, or none)
  int* d_act = nullptr;
  void* cf = nullptr;
  int ncf = 0;
  // the march queue's dependencies (per slot, the finer leaves it waits for), control words and
  // done flags
  int* d_dep_off = nullptr;
  int* d_dep = nullptr;
  int* d_qctl = nullptr;
  int* d_done = nullptr;
  int nreg = 0;
  int grid = 256;            // persistent workgroups of the fused kernel
  void* clk = nullptr;
  unsigned long long* slot = nullptr;
  double* diag_part = nullptr;
  double* diag_out = nullptr;
  int cur = 0;
  Params<float> pf;
  Params<double> pd;
  Levels<float> lf;
  Levels<double> ld;
  AmrOps ops{};
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int launches = 0;
  // per-block timesteps: bins (0 off), each slot's bin of the last cycle, each leaf's largest
  // rate / dx, the cycle's slot lists; the work done
  int kbins = 0;
  int* d_bins = nullptr;
  void* d_brate = nullptr;
  int* d_blists = nullptr;
  std::vector<int> bins;
  std::vector<int> all_host;   // the list of every block, as d_lists holds it
  std::vector<int> blist_host; // the last cycle's lists
  long long bin_cycles = 0, bin_updates = 0, bin_single = 0;
  int since_regrid = 0;        // steps since the last automatic re-mesh (layout regrid_every)
};

namespace {

uint64_t key_of(int L, int x, int y, int z) {
  return ((uint64_t)L << 60) | ((uint64_t)x << 40) | ((uint64_t)y << 20) | (uint64_t)z;
}

// Morton code of a block's lower corner at the finest level, for list order (neighbours close
// in time share cache lines of each other's rims)
uint64_t morton_of(const int* b, int maxlev) {
  uint64_t c[3];
  for (int a = 0; a < 3; a++) c[a] = (uint64_t)b[1 + a] << (maxlev - b[0]);
  uint64_t m = 0;
  for (int bit = 0; bit < 20; bit++)
    for (int a = 0; a < 3; a++) m |= ((c[a] >> bit) & 1ull) << (3 * bit + a);
  return m;
}

// The block table (tree.ts neighbor :293-324, writeRow :652-675); returns SHX_OK or SHX_EINVAL.
int build_table(shx_amr* s) {
  const int n = s->nblocks;
  const int* root = s->lay.root;
  // 2D: a quadtree (four children, octant x | y << 1); each z face names the block itself
  // (tree.ts writeRow :662-666: the single layer is its own neighbour in z)
  const int nchild = 1 << s->ndim, nface = 2 * s->ndim;
  std::unordered_map<uint64_t, int> at;
  at.reserve(2 * n);
  for (int i = 0; i < n; i++) {
    const int* b = &s->blocks[4 * i];
    if (b[0] < 0 || b[0] >= AMR_MAXLEV) return fail(SHX_EINVAL, "block %d: level %d outside [0, %d)", i, b[0], AMR_MAXLEV);
    for (int a = 0; a < 3; a++)
      if (b[1 + a] < 0 || b[1 + a] >= (root[a] << b[0])) return fail(SHX_EINVAL, "block %d: coordinate %d out of range", i, a);
    if (s->ndim == 2 && b[3] != 0) return fail(SHX_EINVAL, "block %d: a 2D block has iz 0", i);
    if (!at.emplace(key_of(b[0], b[1], b[2], b[3]), i).second) return fail(SHX_EINVAL, "block %d repeats an earlier block", i);
  }
  auto find = [&](int L, int x, int y, int z) {
    auto it = at.find(key_of(L, x, y, z));
    return it == at.end() ? -1 : it->second;
  };
  s->tab.assign((size_t)n * AMR_BT, -1);
  int maxlev = 0;
  for (int i = 0; i < n; i++) {
    const int* b = &s->blocks[4 * i];
    int* r = &s->tab[(size_t)i * AMR_BT];
    const int L = b[0];
    maxlev = std::max(maxlev, L);
    r[BT_LEVEL] = L; r[1] = b[1]; r[2] = b[2]; r[3] = b[3];
    r[BT_PARENT] = L > 0 ? find(L - 1, b[1] >> 1, b[2] >> 1, b[3] >> 1) : -1;
    if (L > 0 && r[BT_PARENT] < 0) return fail(SHX_EINVAL, "block %d (level %d) has no parent", i, L);
    int nch = 0;
    for (int o = 0; o < nchild; o++) {
      int ch = find(L + 1, 2 * b[1] + (o & 1), 2 * b[2] + ((o >> 1) & 1), 2 * b[3] + ((o >> 2) & 1));
      r[BT_CHILD + o] = ch;
      nch += ch >= 0;
    }
    if (nch != 0 && nch != nchild) return fail(SHX_EINVAL, "block %d has %d of its %d children", i, nch, nchild);
    r[BT_FLAGS] = 1 | (nch == 0 ? 2 : 0);
  }
  int nroot = 0;
  for (int i = 0; i < n; i++) nroot += s->blocks[4 * i] == 0;
  if (nroot != root[0] * root[1] * root[2]) return fail(SHX_EINVAL, "%d root blocks listed, the root grid has %d", nroot, root[0] * root[1] * root[2]);
  for (int i = 0; i < n; i++) {
    const int* b = &s->blocks[4 * i];
    int* r = &s->tab[(size_t)i * AMR_BT];
    const int L = b[0];
    for (int f = 0; f < 6; f++) {
      if (f >= nface) { r[BT_NB + f] = i; continue; }
      const int axis = f >> 1, up = f & 1;
      int c[3] = {b[1], b[2], b[3]};
      const int nn = root[axis] << L;
      int v = c[axis] + (up ? 1 : -1);
      if (v < 0 || v >= nn) {
        const int type = up ? s->cfg.bc_hi[axis] : s->cfg.bc_lo[axis];
        if (type != SHX_BC_PERIODIC) { r[BT_NB + f] = -1 - type; continue; }
        v = (v + nn) % nn;
      }
      c[axis] = v;
      const int same = find(L, c[0], c[1], c[2]);
      if (same >= 0) { r[BT_NB + f] = same; continue; }
      const int coarse = L > 0 ? find(L - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1) : -1;
      // the face region is covered at this level or one coarser (2:1 balance); with children
      // all-or-none, a block found one level coarser is a leaf
      if (coarse < 0) return fail(SHX_EINVAL, "block %d breaks 2:1 balance across face %d", i, f);
      r[BT_NB + f] = -1;
      r[BT_COARSE + f] = coarse;
    }
  }
  s->nlevels = maxlev + 1;
  return SHX_OK;
}

// BlockTree.adapt (tree.ts:489-641), the policy of amr_tree.adapt_tree: split the leaves above
// `hi` (below max_level), merge the parents whose leaf children are all below `lo`, rebuild the
// tree level by level from the roots with those decisions, then split whatever 2:1 face
// balance needs (the balanced closure is unique, so the order of the splits does not matter).
// out: (level, x, y, z) rows sorted by level, z, y, x.
void adapt_tree(const int* blocks, int n, const int root[3], int ndim, const double* ind, double hi, double lo,
                int max_level, const int periodic[3], std::vector<int>& out) {
  const int nchild = 1 << ndim;
  std::unordered_map<uint64_t, int> at;
  at.reserve(2 * n);
  for (int i = 0; i < n; i++) at[key_of(blocks[4 * i], blocks[4 * i + 1], blocks[4 * i + 2], blocks[4 * i + 3])] = i;
  auto child = [&](int L, int x, int y, int z, int o, int c[3]) {
    c[0] = 2 * x + (o & 1);
    c[1] = 2 * y + ((o >> 1) & 1);
    c[2] = ndim == 3 ? 2 * z + ((o >> 2) & 1) : 0;
    (void)L;
  };
  auto is_leaf = [&](int i) {
    const int* b = &blocks[4 * i];
    for (int o = 0; o < nchild; o++) {
      int c[3];
      child(b[0], b[1], b[2], b[3], o, c);
      if (at.count(key_of(b[0] + 1, c[0], c[1], c[2]))) return false;
    }
    return true;
  };
  std::unordered_set<uint64_t> keep;   // blocks of the current tree that stay refined or split
  for (int i = 0; i < n; i++) {
    const int* b = &blocks[4 * i];
    const uint64_t k = key_of(b[0], b[1], b[2], b[3]);
    if (is_leaf(i)) {
      if (b[0] < max_level && ind[i] > hi) keep.insert(k);
      continue;
    }
    bool merge = true;
    for (int o = 0; o < nchild && merge; o++) {
      int c[3];
      child(b[0], b[1], b[2], b[3], o, c);
      auto it = at.find(key_of(b[0] + 1, c[0], c[1], c[2]));
      merge = it != at.end() && is_leaf(it->second) && ind[it->second] < lo;
    }
    if (!merge) keep.insert(k);
  }
  // (level, x, y, z) of the new tree's leaves and refined blocks
  std::unordered_set<uint64_t> leaves, refined;
  auto split = [&](int L, int x, int y, int z) {
    const uint64_t k = key_of(L, x, y, z);
    if (!leaves.erase(k)) return;
    refined.insert(k);
    for (int o = 0; o < nchild; o++) {
      int c[3];
      child(L, x, y, z, o, c);
      leaves.insert(key_of(L + 1, c[0], c[1], c[2]));
    }
  };
  auto unkey = [](uint64_t k, int& L, int& x, int& y, int& z) {
    L = (int)(k >> 60);
    x = (int)((k >> 40) & 0xFFFFF);
    y = (int)((k >> 20) & 0xFFFFF);
    z = (int)(k & 0xFFFFF);
  };
  for (int z = 0; z < root[2]; z++)
    for (int y = 0; y < root[1]; y++)
      for (int x = 0; x < root[0]; x++) leaves.insert(key_of(0, x, y, z));
  for (int L = 0; L < max_level; L++) {
    std::vector<uint64_t> want;
    for (uint64_t k : leaves)
      if ((int)(k >> 60) == L && keep.count(k)) want.push_back(k);
    for (uint64_t k : want) {
      int l, x, y, z;
      unkey(k, l, x, y, z);
      split(l, x, y, z);
    }
  }
  for (bool changed = true; changed;) {
    changed = false;
    std::vector<uint64_t> snap(leaves.begin(), leaves.end());
    for (uint64_t k : snap) {
      if (!leaves.count(k)) continue;
      int L, c0, c1, c2;
      unkey(k, L, c0, c1, c2);
      if (L == 0) continue;
      for (int a = 0; a < ndim; a++) {
        const int nn = root[a] << L;
        for (int step = -1; step <= 1; step += 2) {
          int d[3] = {c0, c1, c2};
          int v = d[a] + step;
          if (v < 0 || v >= nn) {
            if (!periodic[a]) continue;
            v = (v + nn) % nn;
          }
          d[a] = v;
          const uint64_t dk = key_of(L, d[0], d[1], d[2]);
          if (leaves.count(dk) || refined.count(dk)) continue;
          if (leaves.count(key_of(L - 1, d[0] >> 1, d[1] >> 1, d[2] >> 1))) continue;
          for (int up = 2; up <= L; up++) {
            const uint64_t q = key_of(L - up, d[0] >> up, d[1] >> up, d[2] >> up);
            if (leaves.count(q)) {
              split(L - up, d[0] >> up, d[1] >> up, d[2] >> up);
              changed = true;
              break;
            }
          }
        }
      }
    }
  }
  std::vector<std::array<int, 4>> rows;
  rows.reserve(leaves.size() + refined.size());
  for (const auto* set : {&refined, &leaves})
    for (uint64_t k : *set) {
      std::array<int, 4> r;
      unkey(k, r[0], r[1], r[2], r[3]);
      rows.push_back(r);
    }
  std::sort(rows.begin(), rows.end(), [](const std::array<int, 4>& a, const std::array<int, 4>& b) {
    if (a[0] != b[0]) return a[0] < b[0];
    if (a[3] != b[3]) return a[3] < b[3];
    if (a[2] != b[2]) return a[2] < b[2];
    return a[1] < b[1];
  });
  out.clear();
  for (const auto& r : rows) out.insert(out.end(), r.begin(), r.end());
}

// Rim lane r (< 60) of face f of a leaf's ghost shell, beside a same-level leaf: the half-step
// item h it fills (faces 0-3: column f * 16 + i of the first ghost ring, plane j; faces 4 and 5:
// the plane z = -1 or 16, texel (i, j)) and the cell of the leaf across that is its source, as
// x | y << 4 | z << 8. The face positions (i, j) of the rim: the rows j = 0 and 15, then the
// columns i = 0 and 15 between them.
static void rim_ghost(int f, int r, int& h, int& cell) {
  int i, j;
  if (r < 16) { i = r; j = 0; }
  else if (r < 32) { i = r - 16; j = AMR_BS - 1; }
  else if (r < 46) { i = 0; j = r - 31; }
  else { i = AMR_BS - 1; j = r - 45; }
  int x, y, z;
  if (f < 4) {
    x = f == 0 ? -1 : (f == 1 ? AMR_BS : i);
    y = f == 2 ? -1 : (f == 3 ? AMR_BS : i);
    z = j;
    h = (f * AMR_BS + i) * AMR_BS + z;
  } else {
    x = i;
    y = j;
    z = f == 4 ? -1 : AMR_BS;
    h = 64 * AMR_BS + (f == 4 ? 0 : AMR_BS * AMR_BS) + y * AMR_BS + x;
  }
  // the ghost's source across a same-level face: the coordinate wraps by a block
  x = (x + AMR_BS) % AMR_BS;
  y = (y + AMR_BS) % AMR_BS;
  z = (z + AMR_BS) % AMR_BS;
  cell = x | (y << 4) | (z << 8);
}

int build_lists(shx_amr* s, std::vector<int>& lists) {
  const int n = s->nblocks;
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::vector<uint64_t> mort(n);
  for (int i = 0; i < n; i++) mort[i] = morton_of(&s->blocks[4 * i], s->nlevels - 1);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    int la = s->blocks[4 * a], lb = s->blocks[4 * b];
    return la != lb ? la > lb : mort[a] < mort[b];   // finest level first, Morton order within
  });
  auto is_leaf = [&](int i) { return (s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2) != 0; };
  auto has_fine = [&](int i) {
    for (int f = 0; f < 6; f++) {
      int nb = s->tab[(size_t)i * AMR_BT + BT_NB + f];
      if (nb >= 0 && !is_leaf(nb)) return true;
    }
    return false;
  };
  auto add = [&](Range& r, auto pred) {
    r.off = (int)lists.size();
    for (int i : order)
      if (pred(i)) lists.push_back(i);
    r.count = (int)lists.size() - r.off;
  };
  add(s->all, [&](int) { return true; });
  add(s->leaf, [&](int i) { return is_leaf(i); });
  add(s->plain, [&](int i) { return is_leaf(i) && !has_fine(i); });
  add(s->match, [&](int i) { return is_leaf(i) && has_fine(i); });
  // the fused march's queue, the chain of coarse-fine dependencies first: level by level from
  // the finest, the leaves beside a coarser leaf (whose fluxes it waits for) and the leaves
  // beside a refined block (which wait for the finer leaves, one part earlier); then every
  // other leaf, which fills the CUs while the chain runs. Each part is padded to an even
  // length, so a workgroup's pair never holds a leaf and a finer leaf it waits for.
  auto has_coarse = [&](int i) {
    for (int f = 0; f < 6; f++)
      if (s->tab[(size_t)i * AMR_BT + BT_NB + f] == -1) return true;
    return false;
  };
  s->queue.off = (int)lists.size();
  auto pad = [&] { if (((int)lists.size() - s->queue.off) & 1) lists.push_back(-1); };
  for (int L = s->nlevels - 1; L >= 0; L--) {
    for (int i : order)
      if (s->blocks[4 * i] == L && is_leaf(i) && (has_coarse(i) || has_fine(i))) lists.push_back(i);
    pad();
  }
  for (int i : order)
    if (is_leaf(i) && !has_coarse(i) && !has_fine(i)) lists.push_back(i);
  pad();
  s->queue.count = (int)lists.size() - s->queue.off;
  for (int L = 0; L < s->nlevels; L++) {
    auto lv = [&](int i) { return s->blocks[4 * i] == L; };
    add(s->all_lv[L], lv);
    add(s->leaf_lv[L], [&](int i) { return lv(i) && is_leaf(i); });
    add(s->plain_lv[L], [&](int i) { return lv(i) && is_leaf(i) && !has_fine(i); });
    add(s->match_lv[L], [&](int i) { return lv(i) && is_leaf(i) && has_fine(i); });
  }
  return SHX_OK;
}

AmrLaunch launch_of(shx_amr* s) {
  AmrLaunch L;
  L.stream = s->stream;
  L.params = s->es == 8 ? (const void*)&s->pd : (const void*)&s->pf;
  L.levels = s->es == 8 ? (const void*)&s->ld : (const void*)&s->lf;
  L.tab = s->d_tab;
  L.dx0 = s->cfg.dx;
  return L;
}

const int* lp(shx_amr* s, const Range& r) { return s->d_lists + r.off; }

// Pass groups of one step from parity p into 1 - p; `stage` selects one group (-1: all):
// 0 assemble, 1 clock, 2 predict, 3 update. Returns the launches enqueued.
int enqueue_step(shx_amr* s, int p, int stage = -1) {
  const AmrLaunch L = launch_of(s);
  const AmrOps& o = s->ops;
  const bool levels = s->lay.path == SHX_AMR_PATH_LEVELS;
  const bool comp = s->lay.compensated != 0;
  int n = 0;
  auto run = [&](int g) { return stage < 0 || stage == g; };
  if (s->lay.path == SHX_AMR_PATH_FUSED) {
    if (run(1)) { o.clock(L, s->clk, s->slot); n++; }
    if (run(2)) {
      o.shell(L, s->d_shw[0], s->nshw[0], s->d_shw[1], s->nshw[1], s->d_shw[2], s->nshw[2], s->U1[p], s->U2[p],
              s->shell, s->clk);
      n += (s->nshw[0] > 0) + (s->nshw[1] > 0) + (s->nshw[2] > 0);
    }
    if (run(3) && s->queue.count > 0) {
      AmrQueue Q{lp(s, s->queue), s->queue.count, s->grid, s->d_shell_of, s->d_act, s->reg, s->cf, s->d_reg_of,
                 comp ? 1 : 0, s->d_dep_off, s->d_dep, s->d_qctl, s->d_done};
      o.fused(L, Q, s->shell, s->U1[p], s->U2[p], s->U1[1 - p], s->U2[1 - p], s->clk, s->slot);
      n += 2;
    }
    return n;
  }
  if (run(0)) {
    // the leaves' atlases alone: with one timestep a parent's atlas feeds only the half steps of
    // its leaf neighbours' ghosts beside it, which the predict pass then skips (the faces there
    // are formed from the parent's children: match_faces), and ghosts on a parent are restricted
    // from its children's state
    if (levels) {
      for (int l = 0; l < s->nlevels; l++)
        if (s->leaf_lv[l].count) { o.assemble(L, lp(s, s->leaf_lv[l]), s->leaf_lv[l].count, s->U1[p], s->U2[p], s->W1, s->W2, s->slot); n++; }
    } else {
      o.assemble(L, lp(s, s->leaf), s->leaf.count, s->U1[p], s->U2[p], s->W1, s->W2, s->slot);
      n++;
    }
  }
  if (run(1)) { o.clock(L, s->clk, s->slot); n++; }
  if (run(2)) {
    if (levels) {
      for (int l = 0; l < s->nlevels; l++)
        if (s->leaf_lv[l].count) { o.predict(L, lp(s, s->leaf_lv[l]), s->leaf_lv[l].count, s->W1, s->W2, s->PA, s->PB, s->clk); n++; }
    } else {
      o.predict(L, lp(s, s->leaf), s->leaf.count, s->W1, s->W2, s->PA, s->PB, s->clk);
      n++;
    }
  }
  if (run(3)) {
    auto upd = [&](const Range& r, bool match) {
      if (!r.count) return;
      o.update(L, lp(s, r), r.count, match, comp, s->U1[p], s->U2[p], s->W1, s->W2, s->PA, s->PB, s->U1[1 - p],
               s->U2[1 - p], s->clk, false);
      n++;
    };
    if (levels) {
      for (int l = 0; l < s->nlevels; l++) { upd(s->plain_lv[l], false); upd(s->match_lv[l], true); }
    } else {
      upd(s->plain, false);
      upd(s->match, true);
    }
  }
  return n;
}

void destroy_graphs(shx_amr* s) {
  for (int p = 0; p < 2; p++)
    if (s->graph[p]) { (void)hipGraphExecDestroy(s->graph[p]); s->graph[p] = nullptr; }
}

int build_graphs(shx_amr* s) {
  destroy_graphs(s);
  for (int p = 0; p < 2; p++) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    s->launches = enqueue_step(s, p);
    HIPCHK(hipStreamEndCapture(s->stream, &g));
    HIPCHK(hipGraphInstantiate(&s->graph[p], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  return SHX_OK;
}

template <typename T>
Levels<T> levels_of(double dx0) {
  Levels<T> lv;
  for (int l = 0; l < AMR_MAXLEV; l++) lv.dx[l] = T(dx0 / (double)(1 << l));
  return lv;
}

template <typename T>
int set_state_t(shx_amr* s, const T* u1, const T* u2, size_t ncell, double time, int steps) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if ((int)sizeof(T) != s->es) return fail(SHX_EINVAL, "state precision %d does not match the sim's %d", (int)sizeof(T) * 8, s->es * 8);
  if (ncell != s->ncell) return fail(SHX_EINVAL, "ncell %zu != %zu (blocks x %d)", ncell, s->ncell, s->cells);
  for (int i = 0; i < s->nblocks; i++) {
    if (!(s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2)) continue;
    for (size_t o = (size_t)i * s->cells * 4; o < (size_t)(i + 1) * s->cells * 4; o++)
      if (!isfinite((double)u1[o]) || !isfinite((double)u2[o])) return fail(SHX_ESTATE, "non-finite state in leaf %d", i);
  }
  const size_t bytes = ncell * 4 * sizeof(T);
  s->cur = 0;
  HIPCHK(hipMemcpyAsync(s->U1[0], u1, bytes, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->U2[0], u2, bytes, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->U1[1], u1, bytes, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->U2[1], u2, bytes, hipMemcpyHostToDevice, s->stream));
  T clk[16] = {0};
  clk[1] = T(time);
  clk[3] = T(steps);
  clk[4] = clk[5] = clk[6] = T(1);
  HIPCHK(hipMemcpyAsync(s->clk, clk, sizeof clk, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemsetAsync(s->slot, 0, sizeof(unsigned long long), s->stream));
  if (s->lay.path == SHX_AMR_PATH_FUSED) {
    // the fused kernel leaves the next step's rate in the slot; the first step's comes from here
    s->ops.rates(launch_of(s), lp(s, s->leaf), s->leaf.count, s->U1[0], s->U2[0], s->slot);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return SHX_OK;
}

// After a synchronisation: a dependency wait of the fused march that ran out (the step's
// coarse-fine means were formed without the finer fluxes) is an error, reported once.
DiagGeo diag_geo(const shx_config& c) {
  DiagGeo G;
  G.dv = 0;   // per block: h^ndim
  for (int a = 0; a < 3; a++) G.org[a] = c.org[a];
  G.flat = c.n[2] == 1 || c.ndim < 3;
  G.gamma = c.gamma;
  G.rho_floor = c.rho_floor;
  G.p_floor = c.p_floor;
  return G;
}

int check_queue(shx_amr* s) {
  if (!s->d_qctl) return SHX_OK;
  int w[AMR_Q_WORDS];
  HIPCHK(hipMemcpy(w, s->d_qctl, sizeof w, hipMemcpyDeviceToHost));
  if (w[AMR_Q_ERROR]) {
    const int zero = 0;
    HIPCHK(hipMemcpy(s->d_qctl + AMR_Q_ERROR, &zero, sizeof zero, hipMemcpyHostToDevice));
    return fail(SHX_EDEVICE, "the block march's wait for finer leaves timed out; the state is unreliable");
  }
  return SHX_OK;
}

template <typename T>
int read_state_t(shx_amr* s, T* u1, T* u2) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if ((int)sizeof(T) != s->es) return fail(SHX_EINVAL, "state precision mismatch");
  const size_t bytes = s->ncell * 4 * sizeof(T);
  HIPCHK(hipMemcpyAsync(u1, s->U1[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(u2, s->U2[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return check_queue(s);
}

// writeParams / buildHydro on blocks: the run-time fields (and with `scheme` the limiter, solver
// and reconstruction) of *cfg; the tree, its storage, the boundaries, precision and device stay.
// The parameters are kernel arguments of the captured step graphs, so those are rebuilt; the
// fused march's first rate is taken again from the current state (it depends on gamma).
int amr_reconfigure(shx_amr* s, const shx_config* cfg, bool scheme) {
  if (!s || !cfg) return fail(SHX_EINVAL, "null argument");
  const shx_config& a = s->cfg;
  const shx_config& b = *cfg;
  bool same = a.ndim == b.ndim && a.precision == b.precision && a.device == b.device && a.dx == b.dx;
  for (int k = 0; k < 3; k++)
    same = same && a.n[k] == b.n[k] && a.bc_lo[k] == b.bc_lo[k] && a.bc_hi[k] == b.bc_hi[k];
  if (!scheme)
    same = same && a.limiter == b.limiter && a.riemann == b.riemann && (a.isentropic != 0) == (b.isentropic != 0);
  if (!same)
    return fail(SHX_EINVAL, scheme ? "the grid, boundaries, precision and device of a block sim are fixed"
                                   : "the grid, boundaries, precision, device and scheme of a block sim are fixed "
                                     "(shx_amr_build_hydro changes the scheme)");
  if (!(b.gamma > 1) || !(b.cfl > 0)) return fail(SHX_EINVAL, "gamma must exceed 1 and cfl be positive");
  AmrOps ops = s->ops;
  if (scheme) {
    ops = amr_ops(b.precision, b.limiter, b.riemann, b.isentropic ? 1 : 0, s->ndim);
    if (!ops.assemble)
      return fail(SHX_EINVAL, "no block kernels built for precision %d limiter %d riemann %d isentropic %d",
                  b.precision, b.limiter, b.riemann, b.isentropic);
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  int rc = check_queue(s);
  if (rc) return rc;
  s->ops = ops;
  s->cfg = b;
  s->pf = make_params<float>(b);
  s->pd = make_params<double>(b);
  if (s->lay.path == SHX_AMR_PATH_FUSED) {
    HIPCHK(hipMemsetAsync(s->slot, 0, sizeof(unsigned long long), s->stream));
    s->ops.rates(launch_of(s), lp(s, s->leaf), s->leaf.count, s->U1[s->cur], s->U2[s->cur], s->slot);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return build_graphs(s);
}

// ---------------- per-block timesteps (src/amr/bins_*.wgsl, amr_sim.ts encodeCycle :1509-1660) ----------------
// The bins of a cycle are planned on the host from each leaf's largest rate / dx and the
// cycle's dt, read back once per cycle (HIP has no indirect dispatch: every substep's lists are
// then exact launches): bins_init.wgsl, one DILATE sweep and (K + levels + 2) | 1 sweeps of
// bins_relax.wgsl (Jacobi, each reading the previous one's bins), in the precision of the run.
template <typename T>
void plan_bins(const shx_amr* s, const T* brate, T dt, T cfl, std::vector<int>& out) {
  const int n = s->nblocks, K = s->kbins;
  auto row = [&](int b) { return &s->tab[(size_t)b * AMR_BT]; };
  auto leaf = [&](int b) { return (row(b)[BT_FLAGS] & 3) == 3; };
  std::vector<int> A(n), B(n);
  for (int b = 0; b < n; b++) {
    int k = K;
    if (leaf(b) && dt > T(0)) {
      const T dtb = cfl / std::max(brate[b], T(1e-30));
      T f = std::floor(std::log2(dtb / dt));
      f = std::min(std::max(f, T(0)), T(K));
      k = (int)f;
    }
    A[b] = k;
  }
  const int sweeps = (K + s->nlevels + 2) | 1;
  for (int it = 0; it <= sweeps; it++) {
    const std::vector<int>& kin = (it % 2 == 0) ? A : B;
    std::vector<int>& kout = (it % 2 == 0) ? B : A;
    const int plus = it == 0 ? 0 : 1;
    for (int b = 0; b < n; b++) {
      const int* r = row(b);
      int k = kin[b];
      if (!leaf(b)) {
        k = K;
        for (int c = 0; c < 8; c++)
          if (r[BT_CHILD + c] >= 0) k = std::min(k, kin[r[BT_CHILD + c]]);
      } else {
        for (int f = 0; f < 6; f++) {
          const int nb = r[BT_NB + f];
          if (nb >= 0 && nb != b) k = std::min(k, kin[nb] + plus);
          else if (nb == -1) k = std::min(k, kin[r[BT_COARSE + f]]);
        }
      }
      kout[b] = k;
    }
  }
  out = A;
}

// faster_neighbor (bins_lists.wgsl:22-52)
bool faster_neighbor(const shx_amr* s, const std::vector<int>& k, int b) {
  const int* r = &s->tab[(size_t)b * AMR_BT];
  for (int f = 0; f < 6; f++) {
    const int nb = r[BT_NB + f];
    if (nb < 0 || nb == b) continue;
    const int* q = &s->tab[(size_t)nb * AMR_BT];
    if ((q[BT_FLAGS] & 3) == 3) {
      if (k[nb] < k[b]) return true;
      continue;
    }
    const int axis = f >> 1, want = 1 - (f & 1);
    for (int c = 0; c < 8; c++) {
      if (((c >> axis) & 1) != want) continue;
      const int ch = q[BT_CHILD + c];
      if (ch >= 0 && (s->tab[(size_t)ch * AMR_BT + BT_FLAGS] & 3) == 3 && k[ch] < k[b]) return true;
    }
  }
  return false;
}

// One cycle of 2^K substeps. Substep 0: bank the last step, assemble every block (its rates in
// the reduction slot), each leaf's own rate, the cycle's dt; then (after the one readback of the
// cycle) the bins and every substep's lists (bins_lists.wgsl): the blocks starting a step (LA,
// parents included: a ghost of a leaf beside a refined block is predicted from the parent's
// atlas, and a finer neighbour half way through the leaf's step interpolates from that
// prediction), the leaves among them (LL), those and the leaves half way through a step with a
// faster face neighbour (LU). Each substep (s > 0: bank, assemble LA) predicts LL,
// updates LU into the other buffer and copies those blocks back (the commit): the state stays
// in its buffer through the cycle.
int run_cycle(shx_amr* s) {
  const int n = 1 << s->kbins;
  const int p = s->cur, q = 1 - p;
  const AmrLaunch L = launch_of(s);
  const AmrOps& o = s->ops;
  o.bank(L, s->clk);
  o.assemble(L, lp(s, s->all), s->all.count, s->U1[p], s->U2[p], s->W1, s->W2, s->slot);
  o.block_rates(L, lp(s, s->leaf), s->leaf.count, s->U1[p], s->U2[p], s->d_brate);
  o.clock_cycle(L, s->clk, s->slot, n);
  HIPCHK(hipGetLastError());
  std::vector<unsigned char> br((size_t)s->nblocks * s->es);
  unsigned char ck[16 * 8];
  HIPCHK(hipMemcpyAsync(br.data(), s->d_brate, br.size(), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(ck, s->clk, 16 * s->es, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  // rates of parents were never written: leaves alone are read
  if (s->es == 8) plan_bins<double>(s, (const double*)br.data(), ((const double*)ck)[0], s->pd.cfl, s->bins);
  else plan_bins<float>(s, (const float*)br.data(), ((const float*)ck)[0], s->pf.cfl, s->bins);
  const std::vector<int>& k = s->bins;
  auto leaf = [&](int b) { return (s->tab[(size_t)b * AMR_BT + BT_FLAGS] & 3) == 3; };
  // (kept in the sim: the copy to the device reads it after this call returns)
  std::vector<int>& lists = s->blist_host;
  lists.clear();
  std::vector<Range> la(n), ll(n), lu(n);
  long long updates = 0;
  // in the slot order of the plain lists (finest level first, Morton order within)
  const std::vector<int>& all_order = s->all_host;
  for (int sub = 0; sub < n; sub++) {
    auto active = [&](int b) { return (sub & ((1 << k[b]) - 1)) == 0; };
    auto half = [&](int b) { return k[b] > 0 && (sub & ((1 << k[b]) - 1)) == (1 << (k[b] - 1)); };
    la[sub].off = (int)lists.size();
    for (int b : all_order)
      if (active(b)) lists.push_back(b);
    la[sub].count = (int)lists.size() - la[sub].off;
    ll[sub].off = (int)lists.size();
    for (int b : all_order)
      if (leaf(b) && active(b)) lists.push_back(b);
    ll[sub].count = (int)lists.size() - ll[sub].off;
    lu[sub].off = (int)lists.size();
    for (int b : all_order)
      if (leaf(b) && (active(b) || (half(b) && faster_neighbor(s, k, b)))) lists.push_back(b);
    lu[sub].count = (int)lists.size() - lu[sub].off;
    updates += lu[sub].count;
  }
  HIPCHK(hipMemcpyAsync(s->d_bins, k.data(), k.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->d_blists, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  const bool comp = s->lay.compensated != 0;
  for (int sub = 0; sub < n; sub++) {
    AmrLaunch Lb = L;
    Lb.bins.k = s->d_bins;
    Lb.bins.sub = sub;
    Lb.bins.pa = s->PA;
    Lb.bins.pb = s->PB;
    if (sub > 0) {
      o.bank(L, s->clk);
      // (the rates of these assemblies are not used: a scratch word of the slot takes them)
      o.assemble(Lb, s->d_blists + la[sub].off, la[sub].count, s->U1[p], s->U2[p], s->W1, s->W2, s->slot + 1);
    }
    o.predict(Lb, s->d_blists + ll[sub].off, ll[sub].count, s->W1, s->W2, s->PA, s->PB, s->clk);
    o.update(Lb, s->d_blists + lu[sub].off, lu[sub].count, true, comp, s->U1[p], s->U2[p], s->W1, s->W2, s->PA,
             s->PB, s->U1[q], s->U2[q], s->clk, false);
    o.commit(L, s->d_blists + lu[sub].off, lu[sub].count, s->U1[q], s->U2[q], s->U1[p], s->U2[p]);
  }
  HIPCHK(hipGetLastError());
  s->bin_cycles++;
  s->bin_updates += updates;
  s->bin_single += (long long)s->leaf.count << s->kbins;
  return SHX_OK;
}

}  // namespace

extern "C" {

int shx_amr_create(const shx_config* cfg, const shx_amr_layout* layout, const int* blocks, int nblocks,
                   shx_amr** out) {
  if (!cfg || !layout || !blocks || !out || nblocks < 1) return fail(SHX_EINVAL, "null argument or no blocks");
  *out = nullptr;
  const shx_config& c = *cfg;
  if (c.ndim != 2 && c.ndim != 3) return fail(SHX_EINVAL, "the block grid is 2D or 3D (ndim %d)", c.ndim);
  for (int a = 0; a < 3; a++) {
    if (layout->root[a] < 1) return fail(SHX_EINVAL, "root[%d] = %d", a, layout->root[a]);
    if (a == 2 && c.ndim == 2) {
      if (layout->root[2] != 1 || c.n[2] != 1) return fail(SHX_EINVAL, "a 2D block grid has root[2] = 1 and n[2] = 1");
    } else if (c.n[a] != AMR_BS * layout->root[a]) {
      return fail(SHX_EINVAL, "n[%d] = %d must be 16 x root[%d] = %d", a, c.n[a], a, AMR_BS * layout->root[a]);
    }
    if (c.bc_lo[a] < 0 || c.bc_lo[a] > 3 || c.bc_hi[a] < 0 || c.bc_hi[a] > 3) return fail(SHX_EINVAL, "boundary code on axis %d", a);
  }
  if (c.ndim == 2 && layout->path == SHX_AMR_PATH_FUSED)
    return fail(SHX_EINVAL, "2D block grids run on the passes and levels paths (the fused march steps 3D blocks)");
  if (c.precision != 32 && c.precision != 64) return fail(SHX_EINVAL, "precision %d", c.precision);
  if (!(c.gamma > 1) || !(c.dx > 0) || !(c.cfl > 0)) return fail(SHX_EINVAL, "gamma, dx and cfl must be positive (gamma > 1)");
  if (layout->path != SHX_AMR_PATH_FUSED && layout->path != SHX_AMR_PATH_PASSES && layout->path != SHX_AMR_PATH_LEVELS)
    return fail(SHX_EINVAL, "AMR path %d", layout->path);
  if (layout->time_bins < 0 || layout->time_bins > AMR_MAX_BINS)
    return fail(SHX_EINVAL, "time_bins %d (0 to %d)", layout->time_bins, AMR_MAX_BINS);
  if (layout->time_bins > 0 && layout->path == SHX_AMR_PATH_FUSED)
    return fail(SHX_EINVAL, "per-block timesteps run on the passes and levels paths");
  if (layout->regrid_every < 0 || (layout->regrid_every > 0 && (layout->max_level < 0 || layout->max_level >= AMR_MAXLEV)))
    return fail(SHX_EINVAL, "regrid_every %d, max_level %d (0 to %d)", layout->regrid_every, layout->max_level, AMR_MAXLEV - 1);
  AmrOps ops = amr_ops(c.precision, c.limiter, c.riemann, c.isentropic ? 1 : 0, c.ndim);
  if (!ops.assemble) return fail(SHX_EINVAL, "no block kernels built for precision %d limiter %d riemann %d isentropic %d", c.precision, c.limiter, c.riemann, c.isentropic);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c.device) return fail(SHX_EDEVICE, "no HIP device %d", c.device);

  shx_amr* s = new shx_amr();
  s->cfg = c;
  s->lay = *layout;
  s->es = c.precision == 64 ? 8 : 4;
  s->ndim = c.ndim;
  s->cells = c.ndim == 2 ? AMR_BS * AMR_BS : AMR_CELLS;
  s->texels = c.ndim == 2 ? AMR_PAD * AMR_PAD : AMR_TEXELS;
  s->nblocks = nblocks;
  s->ncell = (size_t)nblocks * s->cells;
  s->blocks.assign(blocks, blocks + 4 * (size_t)nblocks);
  int rc = build_table(s);
  if (rc) { delete s; return rc; }
  std::vector<int> lists;
  build_lists(s, lists);
  s->ops = ops;
  s->pf = make_params<float>(c);
  s->pd = make_params<double>(c);
  s->lf = levels_of<float>(c.dx);
  s->ld = levels_of<double>(c.dx);
  if (hipSetDevice(c.device) != hipSuccess) { delete s; return fail(SHX_EDEVICE, "hipSetDevice(%d)", c.device); }

  const bool fused = layout->path == SHX_AMR_PATH_FUSED;
  // flux registers: every leaf face on a coarser block
  std::vector<int> reg_of((size_t)nblocks * 6, -1);
  for (int i = 0; i < nblocks; i++) {
    if (!(s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2)) continue;
    for (int f = 0; f < 6; f++)
      if (s->tab[(size_t)i * AMR_BT + BT_NB + f] == -1) reg_of[(size_t)i * 6 + f] = s->nreg++;
  }
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus < 1) cus = 256;
    // one 12-wave workgroup (two blocks marched side by side, 102 KB of LDS in FP32) per CU
    s->grid = cus;
  }
  const size_t ubytes = s->ncell * 4 * s->es;
  const size_t abytes = fused ? 0 : (size_t)nblocks * s->texels * 4 * s->es;
  auto alloc = [&](void** p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
  bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
  for (int p = 0; ok && p < 2; p++) ok = alloc(&s->U1[p], ubytes) && alloc(&s->U2[p], ubytes);
  if (!fused) ok = ok && alloc(&s->W1, abytes) && alloc(&s->W2, abytes) && alloc(&s->PA, abytes) && alloc(&s->PB, abytes);
  // the march's action per (slot, face)
  std::vector<int> act((size_t)nblocks * 6, -1);
  for (int f = 0; f < nblocks * 6; f++)
    if (reg_of[f] >= 0) act[f] = 2 * reg_of[f];
  // per leaf with a refined neighbour: a coarse-face buffer per such face, and the finer leaves
  // across them (the four children of the neighbour on the face), which its march waits for
  std::vector<int> dep_off(nblocks + 1, 0), dep;
  for (int i = 0; i < nblocks; i++) {
    dep_off[i] = (int)dep.size();
    if (!(s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2)) continue;   // not a leaf
    for (int f = 0; f < 6; f++) {
      const int nb = s->tab[(size_t)i * AMR_BT + BT_NB + f];
      if (nb < 0 || (s->tab[(size_t)nb * AMR_BT + BT_FLAGS] & 2)) continue;   // not a refined neighbour
      act[(size_t)i * 6 + f] = 2 * s->ncf + 1;
      s->ncf++;
      const int axis = f >> 1, bit = (f & 1) ? 0 : 1;   // the neighbour's children on the shared face
      for (int c = 0; c < 8; c++) {
        if (((c >> axis) & 1) != bit) continue;
        const int ch = s->tab[(size_t)nb * AMR_BT + BT_CHILD + c];
        if (ch >= 0) dep.push_back(ch);
      }
    }
  }
  dep_off[nblocks] = (int)dep.size();
  ok = ok && alloc((void**)&s->d_reg_of, reg_of.size() * sizeof(int)) &&
       alloc(&s->reg, (size_t)(s->nreg > 0 ? s->nreg : 1) * 256 * 8 * s->es) &&
       alloc((void**)&s->d_act, act.size() * sizeof(int)) &&
       alloc((void**)&s->d_dep_off, dep_off.size() * sizeof(int)) &&
       alloc((void**)&s->d_dep, (dep.empty() ? 1 : dep.size()) * sizeof(int)) &&
       alloc((void**)&s->d_qctl, AMR_Q_WORDS * sizeof(int)) && alloc((void**)&s->d_done, nblocks * sizeof(int)) &&
       alloc(&s->cf, (size_t)(s->ncf > 0 ? s->ncf : 1) * 256 * 8 * s->es);
  // ghost shells: 5184 items of 8 values per leaf (the fused path)
  std::vector<int> shell_of(nblocks, -1);
  for (int i = 0; i < s->leaf.count; i++) shell_of[lists[s->leaf.off + i]] = i;
  if (fused) ok = ok && alloc(&s->shell, (size_t)s->leaf.count * 5184 * 8 * s->es);
  ok = ok && alloc((void**)&s->d_shell_of, shell_of.size() * sizeof(int));
  // the shell's work: a face whose neighbour is a same-level leaf needs only the half steps of
  // its rim (the march reads that leaf's cells and predicts the rest in place); a face on a
  // coarser or a refined block needs its ghost texels only; a boundary image its texels and all
  // its half steps. Items no work writes stay zero (the buffer is cleared at creation).
  std::vector<int> shw[3];
  // rim ghosts as (source slot, source cell x | y << 4 | z << 8, shell index, half-step item)
  std::vector<std::array<int, 4>> rim;
  for (int i = 0; fused && i < s->leaf.count; i++) {
    const int b = lists[s->leaf.off + i];
    for (int f = 0; f < 6; f++) {
      const int nb = s->tab[(size_t)b * AMR_BT + BT_NB + f];
      const bool same_leaf = nb >= 0 && (s->tab[(size_t)nb * AMR_BT + BT_FLAGS] & 3) == 3;
      auto add = [&](int part, int chunk) {
        shw[part].insert(shw[part].end(), {i, b, f, chunk});
      };
      if (same_leaf) {
        for (int r = 0; r < 60; r++) {
          int h, cell;
          rim_ghost(f, r, h, cell);
          rim.push_back({nb, cell, i, h});
        }
      } else {
        for (int c = 0; c < 8; c++) add(0, c);
        // half steps beside a boundary image only: beside a coarser block the march predicts
        // them in place, and beside a refined block the faces are the children's means (the
        // coarse-face buffer replaces what those half steps would give)
        if (nb < -1)
          for (int c = 0; c < 4; c++) add(1, c);
      }
    }
  }
  // one record per source cell, block by block (8 ints: slot, cell, then three (shell index,
  // item) destinations, -1 for none): an edge cell of a leaf is a rim ghost of the leaves across
  // both its faces, a corner cell of three
  std::sort(rim.begin(), rim.end());
  for (size_t a = 0; a < rim.size();) {
    size_t e = a;
    while (e < rim.size() && rim[e][0] == rim[a][0] && rim[e][1] == rim[a][1]) e++;
    for (size_t c = a; c < e; c += 3) {
      int rec[8] = {rim[a][0], rim[a][1], -1, -1, -1, -1, -1, -1};
      for (size_t d = c; d < e && d < c + 3; d++) {
        rec[2 + 2 * (d - c)] = rim[d][2];
        rec[3 + 2 * (d - c)] = rim[d][3];
      }
      shw[2].insert(shw[2].end(), rec, rec + 8);
    }
    a = e;
  }
  for (int part = 0; part < 3; part++) {
    s->nshw[part] = (int)(shw[part].size() / (part == 2 ? 8 : 4));
    ok = ok && alloc((void**)&s->d_shw[part], (shw[part].empty() ? 8 : shw[part].size()) * sizeof(int));
  }
  ok = ok && alloc((void**)&s->d_tab, s->tab.size() * sizeof(int)) && alloc((void**)&s->d_lists, lists.size() * sizeof(int));
  ok = ok && alloc(&s->clk, 16 * 8) && alloc((void**)&s->slot, 64) && alloc((void**)&s->diag_part, (size_t)nblocks * DIAG_CH * 8) &&
       alloc((void**)&s->diag_out, DIAG_CH * 8);
  s->kbins = layout->time_bins;
  s->all_host.assign(lists.begin() + s->all.off, lists.begin() + s->all.off + s->all.count);
  if (s->kbins > 0) {
    // every substep's three lists (blocks starting a step, the leaves among them, those leaves
    // and the leaves owed a deposit), each at most every block
    const size_t lists_len = (size_t)(1 << s->kbins) * 3 * nblocks;
    s->bins.assign(nblocks, 0);
    ok = ok && alloc((void**)&s->d_bins, (size_t)nblocks * sizeof(int)) && alloc(&s->d_brate, (size_t)nblocks * s->es) &&
         alloc((void**)&s->d_blists, lists_len * sizeof(int));
  }
  if (!ok) {
    shx_amr_destroy(s);
    return fail(SHX_ENOMEM, "device allocation failed for %d blocks (%.2f GB of state and atlases)", nblocks,
                (2 * ubytes * 2 + 4 * abytes) / 1e9);
  }
  // edge and corner texels of W are never assembled: zero, as a fresh atlas
  bool zeroed = (fused || (hipMemset(s->W1, 0, abytes) == hipSuccess && hipMemset(s->W2, 0, abytes) == hipSuccess &&
                            hipMemset(s->PA, 0, abytes) == hipSuccess && hipMemset(s->PB, 0, abytes) == hipSuccess)) &&
                (!fused || hipMemset(s->shell, 0, (size_t)s->leaf.count * 5184 * 8 * s->es) == hipSuccess) &&
                hipMemset(s->U1[0], 0, ubytes) == hipSuccess && hipMemset(s->U2[0], 0, ubytes) == hipSuccess &&
                hipMemset(s->U1[1], 0, ubytes) == hipSuccess && hipMemset(s->U2[1], 0, ubytes) == hipSuccess &&
                hipMemset(s->clk, 0, 16 * 8) == hipSuccess && hipMemset(s->slot, 0, 64) == hipSuccess;
  bool copied = hipMemcpy(s->d_tab, s->tab.data(), s->tab.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                hipMemcpy(s->d_reg_of, reg_of.data(), reg_of.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                hipMemcpy(s->d_act, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                hipMemcpy(s->d_dep_off, dep_off.data(), dep_off.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                (dep.empty() || hipMemcpy(s->d_dep, dep.data(), dep.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
                hipMemset(s->d_qctl, 0, AMR_Q_WORDS * sizeof(int)) == hipSuccess &&
                hipMemset(s->d_done, 0, nblocks * sizeof(int)) == hipSuccess &&
                hipMemcpy(s->d_shell_of, shell_of.data(), shell_of.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                (shw[0].empty() || hipMemcpy(s->d_shw[0], shw[0].data(), shw[0].size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
                (shw[1].empty() || hipMemcpy(s->d_shw[1], shw[1].data(), shw[1].size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
                (shw[2].empty() || hipMemcpy(s->d_shw[2], shw[2].data(), shw[2].size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
                hipMemcpy(s->d_lists, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  if (!zeroed || !copied || hipDeviceSynchronize() != hipSuccess) {
    shx_amr_destroy(s);
    return fail(SHX_EDEVICE, "initialising the block grid's device memory failed");
  }
  rc = build_graphs(s);
  if (rc) { shx_amr_destroy(s); return rc; }
  *out = s;
  return SHX_OK;
}

int shx_amr_destroy(shx_amr* s) {
  if (!s) return SHX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  destroy_graphs(s);
  for (int p = 0; p < 2; p++) { (void)hipFree(s->U1[p]); (void)hipFree(s->U2[p]); }
  (void)hipFree(s->W1); (void)hipFree(s->W2); (void)hipFree(s->PA); (void)hipFree(s->PB);
  (void)hipFree(s->d_reg_of); (void)hipFree(s->reg); (void)hipFree(s->shell); (void)hipFree(s->d_shell_of);
  (void)hipFree(s->d_shw[0]); (void)hipFree(s->d_shw[1]); (void)hipFree(s->d_shw[2]);
  (void)hipFree(s->d_act); (void)hipFree(s->cf);
  (void)hipFree(s->d_dep_off); (void)hipFree(s->d_dep); (void)hipFree(s->d_qctl); (void)hipFree(s->d_done);
  (void)hipFree(s->d_tab); (void)hipFree(s->d_lists); (void)hipFree(s->clk); (void)hipFree(s->slot);
  (void)hipFree(s->diag_part); (void)hipFree(s->diag_out);
  (void)hipFree(s->d_bins); (void)hipFree(s->d_brate); (void)hipFree(s->d_blists);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return SHX_OK;
}

int shx_amr_table(shx_amr* s, int* out, size_t len) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  if (len < s->tab.size()) return fail(SHX_EINVAL, "table needs %zu ints", s->tab.size());
  memcpy(out, s->tab.data(), s->tab.size() * sizeof(int));
  return SHX_OK;
}

int shx_amr_set_state(shx_amr* s, const float* u1, const float* u2, size_t ncell, double time, int steps) {
  return set_state_t<float>(s, u1, u2, ncell, time, steps);
}
int shx_amr_set_state_f64(shx_amr* s, const double* u1, const double* u2, size_t ncell, double time, int steps) {
  return set_state_t<double>(s, u1, u2, ncell, time, steps);
}
int shx_amr_read_state(shx_amr* s, float* u1, float* u2) { return read_state_t<float>(s, u1, u2); }
int shx_amr_read_state_f64(shx_amr* s, double* u1, double* u2) { return read_state_t<double>(s, u1, u2); }

// The automatic re-mesh of layout regrid_every (AmrSim's flag / adapt / applyRegrid cycle,
// amr_sim.ts:3020-3060 and :2648-2697): the indicator of every leaf, the next tree, and the
// re-mesh when the tree changes. Axes wrap for the balance where both faces are periodic.
static int auto_regrid(shx_amr* s) {
  std::vector<double> ind(s->nblocks);
  int rc = shx_amr_indicator(s, ind.data());
  if (rc) return rc;
  int periodic[3];
  for (int a = 0; a < 3; a++) periodic[a] = s->cfg.bc_lo[a] == SHX_BC_PERIODIC && s->cfg.bc_hi[a] == SHX_BC_PERIODIC;
  std::vector<int> next;
  adapt_tree(s->blocks.data(), s->nblocks, s->lay.root, s->ndim, ind.data(), s->lay.refine_above,
             s->lay.coarsen_below, std::min(s->lay.max_level, AMR_MAXLEV - 1), periodic, next);
  if (next == s->blocks) return SHX_OK;
  return shx_amr_regrid(s, next.data(), (int)(next.size() / 4));
}

// `count` steps (whole cycles with per-block timesteps) with no re-mesh between them
static int enqueue_steps(shx_amr* s, int count, int* taken) {
  if (s->kbins > 0) {
    // whole cycles only, as AmrSim.encodeSteps (amr_sim.ts:2783-2790): may run past `count`
    const int n = 1 << s->kbins;
    for (int done = 0; done < count; done += n) {
      int rc = run_cycle(s);
      if (rc) return rc;
      if (taken) *taken = done + n;
    }
    return SHX_OK;
  }
  for (int i = 0; i < count; i++) {
    HIPCHK(hipGraphLaunch(s->graph[s->cur], s->stream));
    s->cur = 1 - s->cur;
    if (taken) *taken = i + 1;
  }
  return SHX_OK;
}

int shx_amr_step(shx_amr* s, int count, int* taken) {
  if (taken) *taken = 0;
  if (!s || count < 0) return fail(SHX_EINVAL, "bad argument");
  if (s->lay.regrid_every <= 0) return enqueue_steps(s, count, taken);
  // steps up to each re-mesh, then the re-mesh (the sim may be a new tree's afterwards)
  int done = 0;
  while (done < count) {
    const int n = std::min(count - done, s->lay.regrid_every - s->since_regrid);
    int part = 0;
    int rc = enqueue_steps(s, n, &part);
    done += part;
    if (taken) *taken = done;
    if (rc) return rc;
    s->since_regrid += part;
    if (s->since_regrid >= s->lay.regrid_every) {
      s->since_regrid = 0;
      rc = auto_regrid(s);
      if (rc) return rc;
    }
  }
  return SHX_OK;
}

int shx_amr_adapt(const int* blocks, int nblocks, const int root[3], int ndim, const double* indicator,
                  double refine_above, double coarsen_below, int max_level, const int periodic[3], int* out, int cap,
                  int* nout) {
  if (!blocks || nblocks < 1 || !root || !indicator || !periodic || !nout || (ndim != 2 && ndim != 3))
    return fail(SHX_EINVAL, "bad argument");
  if (max_level < 0 || max_level >= AMR_MAXLEV) return fail(SHX_EINVAL, "max_level %d outside [0, %d)", max_level, AMR_MAXLEV);
  for (int a = 0; a < 3; a++)
    if (root[a] < 1 || (ndim == 2 && a == 2 && root[a] != 1)) return fail(SHX_EINVAL, "root[%d] = %d", a, root[a]);
  std::vector<int> next;
  adapt_tree(blocks, nblocks, root, ndim, indicator, refine_above, coarsen_below, max_level, periodic, next);
  *nout = (int)(next.size() / 4);
  if (!out || cap < *nout) return fail(SHX_EINVAL, "the next tree has %d blocks (room for %d)", *nout, cap);
  memcpy(out, next.data(), next.size() * sizeof(int));
  return SHX_OK;
}

int shx_amr_sync(shx_amr* s) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  HIPCHK(hipStreamSynchronize(s->stream));
  return check_queue(s);
}

int shx_amr_read_clock(shx_amr* s, double out[11]) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  unsigned char buf[16 * 8];
  HIPCHK(hipMemcpyAsync(buf, s->clk, sizeof buf, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int i = 0; i < 11; i++) out[i] = s->es == 8 ? ((double*)buf)[i] : (double)((float*)buf)[i];
  return SHX_OK;
}

int shx_amr_run_to_end(shx_amr* s, int batch, int max_steps, int* steps) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if (!(s->cfg.t_end < 1e30)) return fail(SHX_EINVAL, "run_to_end needs a finite t_end");
  int rc = run_to_end_loop([s](int n) { return shx_amr_step(s, n, nullptr); },
                           [s](double* c) { return shx_amr_read_clock(s, c); }, s->cfg.t_end, batch, max_steps, steps);
  if (rc) return rc;
  return check_queue(s);
}

int shx_amr_bins(shx_amr* s, int* out, size_t len, int64_t work[3]) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if (out) {
    if (len < (size_t)s->nblocks) return fail(SHX_EINVAL, "bins need %d ints", s->nblocks);
    for (int b = 0; b < s->nblocks; b++) out[b] = s->kbins > 0 ? s->bins[b] : 0;
  }
  if (work) {
    work[0] = s->bin_cycles;
    work[1] = s->bin_updates;
    work[2] = s->bin_single;
  }
  return SHX_OK;
}

int shx_amr_write_params(shx_amr* s, const shx_config* cfg) { return amr_reconfigure(s, cfg, false); }
int shx_amr_build_hydro(shx_amr* s, const shx_config* cfg) { return amr_reconfigure(s, cfg, true); }

int shx_amr_diagnostics(shx_amr* s, double out[24]) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  const AmrLaunch L = launch_of(s);
  s->ops.diag(L, diag_geo(s->cfg), lp(s, s->leaf), s->leaf.count, s->U1[s->cur], s->U2[s->cur], s->diag_part,
              s->diag_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->diag_out, DIAG_CH * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return check_queue(s);
}

int shx_amr_readback_begin(shx_amr* s, int what, shx_readback** out) {
  if (!s || !out) return fail(SHX_EINVAL, "null argument");
  if (what & SHX_READ_DIAG) {
    const AmrLaunch L = launch_of(s);
    s->ops.diag(L, diag_geo(s->cfg), lp(s, s->leaf), s->leaf.count, s->U1[s->cur], s->U2[s->cur], s->diag_part,
                s->diag_out);
    HIPCHK(hipGetLastError());
  }
  return readback_begin(s->stream, what, s->es, s->ncell, s->U1[s->cur], s->U2[s->cur], s->clk, s->diag_out, out);
}

int shx_amr_indicator(shx_amr* s, double* out) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, (size_t)s->nblocks * s->es));
  const AmrLaunch L = launch_of(s);
  s->ops.indicator(L, lp(s, s->leaf), s->leaf.count, s->U1[s->cur], s->U2[s->cur], d);
  int rc = hipGetLastError() == hipSuccess ? SHX_OK : SHX_EDEVICE;
  std::vector<unsigned char> h((size_t)s->nblocks * s->es);
  if (rc == SHX_OK && (hipMemcpyAsync(h.data(), d, h.size(), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                       hipStreamSynchronize(s->stream) != hipSuccess))
    rc = SHX_EDEVICE;
  (void)hipFree(d);
  if (rc != SHX_OK) return fail(rc, "the indicator pass failed");
  rc = check_queue(s);
  if (rc != SHX_OK) return rc;
  for (int i = 0; i < s->nblocks; i++) {
    const bool leaf = (s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2) != 0;
    const double v = s->es == 8 ? ((const double*)h.data())[i] : (double)((const float*)h.data())[i];
    out[i] = leaf ? v : NAN;
  }
  return SHX_OK;
}

int shx_amr_regrid(shx_amr* s, const int* blocks, int nblocks) {
  if (!s || !blocks || nblocks < 1) return fail(SHX_EINVAL, "bad argument");
  // what each block of the new tree takes from the current one (the regrid table of
  // src/amr/tree.ts:681-720): checked before anything is allocated
  std::unordered_map<uint64_t, int> cur;
  for (int i = 0; i < s->nblocks; i++) {
    const int* b = &s->blocks[4 * i];
    cur[key_of(b[0], b[1], b[2], b[3])] = i;
  }
  auto is_leaf = [&](int i) { return (s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2) != 0; };
  auto find = [&](int L, int x, int y, int z) {
    if (L < 0 || x < 0 || y < 0 || z < 0) return -1;
    auto it = cur.find(key_of(L, x, y, z));
    return it == cur.end() ? -1 : it->second;
  };
  std::unordered_map<uint64_t, int> nkeys;
  for (int j = 0; j < nblocks; j++) nkeys[key_of(blocks[4 * j], blocks[4 * j + 1], blocks[4 * j + 2], blocks[4 * j + 3])] = j;
  std::vector<int> map;   // (new slot, kind, old slot, octant) per new leaf
  for (int j = 0; j < nblocks; j++) {
    const int L = blocks[4 * j], x = blocks[4 * j + 1], y = blocks[4 * j + 2], z = blocks[4 * j + 3];
    bool new_leaf = true;
    const int nchild = 1 << s->ndim;
    for (int o = 0; o < nchild && new_leaf; o++)
      if (nkeys.count(key_of(L + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1)))) new_leaf = false;
    if (!new_leaf) continue;
    const int i = find(L, x, y, z);
    if (i >= 0 && is_leaf(i)) { map.insert(map.end(), {j, 1, i, 0}); continue; }
    if (i >= 0) {
      bool kids = true;
      for (int o = 0; o < nchild && kids; o++) {
        const int k = find(L + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1));
        kids = k >= 0 && is_leaf(k);
      }
      if (kids) { map.insert(map.end(), {j, 3, i, 0}); continue; }
    }
    const int p = L > 0 ? find(L - 1, x >> 1, y >> 1, z >> 1) : -1;
    if (p >= 0 && is_leaf(p)) { map.insert(map.end(), {j, 2, p, (x & 1) | ((y & 1) << 1) | ((z & 1) << 2)}); continue; }
    return fail(SHX_EINVAL, "new leaf (%d, %d, %d, %d) is more than one level from the current tree", L, x, y, z);
  }
  // a state the last march left unreliable is not carried onto the new tree
  HIPCHK(hipStreamSynchronize(s->stream));
  int rc = check_queue(s);
  if (rc != SHX_OK) return rc;
  // the new tree's sim (validates the tree and allocates its memory), then the state moved over
  shx_amr* ns = nullptr;
  rc = shx_amr_create(&s->cfg, &s->lay, blocks, nblocks, &ns);
  if (rc != SHX_OK) return rc;
  int* d_map = nullptr;
  const int count = (int)(map.size() / 4);
  bool ok = hipStreamSynchronize(s->stream) == hipSuccess &&
            hipMalloc((void**)&d_map, (map.empty() ? 4 : map.size()) * sizeof(int)) == hipSuccess &&
            hipMemcpy(d_map, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    const AmrLaunch Lold = launch_of(s);
    s->ops.regrid(Lold, d_map, count, s->U1[s->cur], s->U2[s->cur], ns->U1[0], ns->U2[0], ns->U1[1], ns->U2[1]);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(ns->clk, s->clk, 16 * s->es, hipMemcpyDeviceToDevice, s->stream) == hipSuccess &&
         hipStreamSynchronize(s->stream) == hipSuccess;
  }
  (void)hipFree(d_map);
  if (!ok) {
    shx_amr_destroy(ns);
    return fail(SHX_EDEVICE, "moving the state to the new tree failed");
  }
  ns->cur = 0;
  if (ns->lay.path == SHX_AMR_PATH_FUSED) {
    // the fused kernel takes each step's rate from the previous one: seed it for the new state
    ns->ops.rates(launch_of(ns), lp(ns, ns->leaf), ns->leaf.count, ns->U1[0], ns->U2[0], ns->slot);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ns->stream) != hipSuccess) {
      shx_amr_destroy(ns);
      return fail(SHX_EDEVICE, "seeding the new tree's rate failed");
    }
  }
  std::swap(*s, *ns);
  shx_amr_destroy(ns);
  return SHX_OK;
}

int shx_amr_profile(shx_amr* s, int steps, double ms[5]) {
  if (!s || !ms || steps < 1) return fail(SHX_EINVAL, "bad argument");
  hipEvent_t ev[5];
  for (int i = 0; i < 5; i++) HIPCHK(hipEventCreate(&ev[i]));
  double acc[5] = {0, 0, 0, 0, 0};
  for (int n = 0; n < steps; n++) {
    const int p = s->cur;
    HIPCHK(hipEventRecord(ev[0], s->stream));
    for (int g = 0; g < 4; g++) {
      enqueue_step(s, p, g);
      HIPCHK(hipEventRecord(ev[g + 1], s->stream));
    }
    HIPCHK(hipEventSynchronize(ev[4]));
    HIPCHK(hipGetLastError());
    s->cur = 1 - p;
    float t;
    for (int i = 0; i < 4; i++) { HIPCHK(hipEventElapsedTime(&t, ev[i], ev[i + 1])); acc[i] += t; }
    HIPCHK(hipEventElapsedTime(&t, ev[0], ev[4]));
    acc[4] += t;
  }
  for (int i = 0; i < 5; i++) ms[i] = acc[i] / steps;
  for (int i = 0; i < 5; i++) (void)hipEventDestroy(ev[i]);
  return SHX_OK;
}

int shx_amr_flux_mismatch(shx_amr* s, int64_t counts[3], double* worst_coarse_fine) {
  if (!s || !counts) return fail(SHX_EINVAL, "bad argument");
  // one step of the reference's passes with the update in DEBUG_FLUX form (the fused path gives
  // the same bits): atlases for a fused sim are made for the call and freed after it
  const bool own = s->W1 == nullptr;
  const size_t abytes = (size_t)s->nblocks * s->texels * 4 * s->es;
  if (own) {
    void** at[4] = {&s->W1, &s->W2, &s->PA, &s->PB};
    for (void** a : at) {
      if (hipMalloc(a, abytes) != hipSuccess) {
        for (void** b : at) { (void)hipFree(*b); *b = nullptr; }
        return fail(SHX_ENOMEM, "flux check: atlas allocation failed");
      }
      HIPCHK(hipMemsetAsync(*a, 0, abytes, s->stream));
    }
  }
  unsigned char saved[16 * 8];
  unsigned long long slot_saved;
  HIPCHK(hipMemcpyAsync(saved, s->clk, sizeof saved, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(&slot_saved, s->slot, 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const AmrLaunch L = launch_of(s);
  const int p = s->cur;
  HIPCHK(hipMemsetAsync(s->slot, 0, 8, s->stream));
  s->ops.assemble(L, lp(s, s->all), s->all.count, s->U1[p], s->U2[p], s->W1, s->W2, s->slot);
  s->ops.clock(L, s->clk, s->slot);
  s->ops.predict(L, lp(s, s->leaf), s->leaf.count, s->W1, s->W2, s->PA, s->PB, s->clk);
  const bool comp = s->lay.compensated != 0;
  s->ops.update(L, lp(s, s->plain), s->plain.count, false, comp, s->U1[p], s->U2[p], s->W1, s->W2, s->PA, s->PB,
                s->U1[1 - p], s->U2[1 - p], s->clk, true);
  s->ops.update(L, lp(s, s->match), s->match.count, true, comp, s->U1[p], s->U2[p], s->W1, s->W2, s->PA, s->PB,
                s->U1[1 - p], s->U2[1 - p], s->clk, true);
  HIPCHK(hipGetLastError());
  const size_t nc = s->ncell;
  std::vector<double> f1(4 * nc), f2(4 * nc);
  if (s->es == 8) {
    HIPCHK(hipMemcpyAsync(f1.data(), s->U1[1 - p], 32 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(f2.data(), s->U2[1 - p], 32 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    std::vector<float> a(4 * nc), b(4 * nc);
    HIPCHK(hipMemcpyAsync(a.data(), s->U1[1 - p], 16 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(b.data(), s->U2[1 - p], 16 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < 4 * nc; i++) { f1[i] = a[i]; f2[i] = b[i]; }
  }
  HIPCHK(hipMemcpyAsync(s->clk, saved, sizeof saved, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->slot, &slot_saved, 8, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (own) {
    void** at[4] = {&s->W1, &s->W2, &s->PA, &s->PB};
    for (void** a : at) { (void)hipFree(*a); *a = nullptr; }
  }
  // AmrSim.fluxMismatch (amr_sim.ts:4078-4160): same-level faces between blocks compared bit for
  // bit (once, from the lower block); a coarse face against the mean of the fine fluxes beyond it
  auto flux = [&](int b, int x, int y, int z, int f) {
    const size_t o = ((size_t)b * s->cells + ((size_t)z * AMR_BS + y) * AMR_BS + x) * 4;
    return f < 4 ? f1[o + f] : f2[o + f - 4];
  };
  // 2D: no z faces, one layer of cells, two fine faces per coarse face
  const int zc = s->ndim == 2 ? 1 : AMR_BS;
  auto is_leaf = [&](int i) { return (s->tab[(size_t)i * AMR_BT + BT_FLAGS] & 2) != 0; };
  int64_t faces = 0, mis = 0, cf = 0;
  double worst = 0;
  for (int b = 0; b < s->nblocks; b++) {
    if (!is_leaf(b)) continue;
    for (int axis = 0; axis < s->ndim; axis++)
      for (int side = 0; side < 2; side++) {
        const int nb = s->tab[(size_t)b * AMR_BT + BT_NB + 2 * axis + side];
        if (nb < 0) continue;
        const bool refined = !is_leaf(nb);
        if (!refined && side == 0) continue;
        for (int k = 0; k < zc; k++)
          for (int j = 0; j < AMR_BS; j++)
            for (int i = 0; i < AMR_BS; i++) {
              int l[3] = {i, j, k};
              if (l[axis] != (side ? AMR_BS - 1 : 0)) continue;
              const double mine = flux(b, i, j, k, 2 * axis + side);
              if (!refined) {
                int m[3] = {i, j, k};
                m[axis] = 0;
                const double other = flux(nb, m[0], m[1], m[2], 2 * axis);
                faces++;
                if (memcmp(&mine, &other, sizeof mine) != 0) mis++;
                continue;
              }
              const int t1 = axis == 2 ? 0 : axis + 1, t2 = axis == 0 ? 2 : axis - 1;
              const int nfine = s->ndim == 2 ? 2 : 4;
              double sum = 0;
              for (int sb = 0; sb < nfine; sb++) {
                int bits[3] = {0, 0, 0}, m[3] = {0, 0, 0};
                bits[axis] = side ? 0 : 1;
                m[axis] = side ? 0 : AMR_BS - 1;
                const int q[2] = {t1, t2};
                for (int n = 0, bit = 0; n < 2; n++) {
                  if (s->ndim == 2 && q[n] == 2) continue;   // the z index stays 0
                  bits[q[n]] = l[q[n]] >= AMR_BS / 2;
                  m[q[n]] = 2 * (l[q[n]] - (AMR_BS / 2) * bits[q[n]]) + ((sb >> bit++) & 1);
                }
                const int ch = s->tab[(size_t)nb * AMR_BT + BT_CHILD + (bits[0] | (bits[1] << 1) | (bits[2] << 2))];
                sum += flux(ch, m[0], m[1], m[2], 2 * axis + (side ? 0 : 1));
              }
              cf++;
              worst = fmax(worst, fabs(mine - sum / nfine));
            }
      }
  }
  counts[0] = faces;
  counts[1] = mis;
  counts[2] = cf;
  if (worst_coarse_fine) *worst_coarse_fine = worst;
  return SHX_OK;
}

int shx_amr_stats(shx_amr* s, int out[5]) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  out[0] = s->nblocks;
  out[1] = s->leaf.count;
  out[2] = s->match.count;
  out[3] = s->nlevels;
  out[4] = s->launches;
  return SHX_OK;
}

}  // extern "C"
