// Device layout of the edge-based operators (host side).
//
// rhs: nodes are grouped into blocks of consecutive ids; a block owns the
// edges and boundary faces whose smallest node id lies in it, and all nodes
// those touch that lie beyond the block are its halo.  One workgroup per
// block accumulates the block's nodal sums in LDS: the block's work items
// (edges, or boundary faces) are cut into chunks of at most chunk_slots
// flux slots; a chunk's fluxes go to LDS, then every node the chunk touches
// sums its slots in a fixed order.  Owned nodes are written directly; halo
// partial sums go to slots that a fix-up pass adds to their owners in block
// order.  Every sum is therefore taken in an order fixed here: results are
// bitwise reproducible, without atomics.
//
// The nodes a block touches are staged in LDS (primitive state, gradients,
// coordinates), so a chunk's flux evaluations read LDS only; a thread owns one
// local node and keeps its running sum in registers.
//
// The chorin operators use blocks of up to kBlockCap nodes with 1024-slot
// chunks (each node thread gathers the terms of the chunk edges that touch
// it); riemann::rhs builds its own layout of kRhsCap-node blocks and 512-slot
// chunks, two workgroups per CU.  riemann::grad, whose per-edge work is
// small, uses a gather layout (see build_layout): edges between two blocks
// are evaluated in both, so no block leaves partial sums behind.
//
// The layout is fastest when node ids are block-contiguous and spatially
// compact (xg_mesh_block_order), but is correct for any numbering.
#include "xg_layout.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace xg {

// Node blocks of consecutive ids for the node graph given as neighbour lists
// (uoff/up: the neighbours with larger ids of every node, or all neighbours
// for a gather layout), grown while owned <= block_nodes and owned + halo <=
// block_cap; bstart receives the block starts followed by npoin.
static int partition_blocks(int npoin, const std::vector<int64_t>& uoff,
                            const std::vector<int>& up, int block_nodes,
                            int block_cap, std::vector<int>& bstart,
                            std::string& err) {
  bstart.clear();
  std::vector<int> mark(npoin, -1);   // block id that has node in its halo
  int s = 0, b = 0;
  while (s < npoin) {
    int nown = 0, nhalo = 0, end = s;
    while (end < npoin && nown < block_nodes) {
      // adding node `end`: it leaves the halo if it was there
      // (a neighbour outside [s, end] is halo: beyond end, or before s)
      int add = 0;
      for (int64_t k = uoff[end]; k < uoff[end + 1]; ++k)
        if (mark[up[k]] != b && (up[k] > end || up[k] < s)) ++add;
      int was_halo = mark[end] == b ? 1 : 0;
      if (nown > 0 && nown + 1 + nhalo - was_halo + add > block_cap) break;
      for (int64_t k = uoff[end]; k < uoff[end + 1]; ++k)
        if (mark[up[k]] != b && (up[k] > end || up[k] < s)) { mark[up[k]] = b; ++nhalo; }
      if (was_halo) { --nhalo; mark[end] = -2; }
      ++nown; ++end;
    }
    if (nown + nhalo > block_cap) {
      err = "layout: a single node has more neighbours than the block capacity";
      return XG_EUNSUPPORTED;
    }
    bstart.push_back(s);
    s = end; ++b;
  }
  bstart.push_back(npoin);
  return XG_OK;
}

// Upper-neighbour lists of the graph with edges (p[e], q[e]).
static void upper_lists(int npoin, const std::vector<int>& p,
                        const std::vector<int>& q, std::vector<int64_t>& uoff,
                        std::vector<int>& up) {
  const int64_t ne = (int64_t)p.size();
  std::vector<int> ucnt(npoin + 1, 0);
  for (int64_t e = 0; e < ne; ++e) ++ucnt[std::min(p[e], q[e]) + 1];
  uoff.assign(npoin + 1, 0);
  for (int i = 0; i < npoin; ++i) uoff[i + 1] = uoff[i] + ucnt[i + 1];
  up.resize(uoff[npoin]);
  std::vector<int64_t> pos(uoff.begin(), uoff.end() - 1);
  for (int64_t e = 0; e < ne; ++e)
    up[pos[std::min(p[e], q[e])]++] = std::max(p[e], q[e]);
}

// Neighbour lists in both directions (CSR over all nodes).
static void full_lists(int npoin, const std::vector<int>& p, const std::vector<int>& q,
                       std::vector<int64_t>& off, std::vector<int>& nb) {
  const int64_t ne = (int64_t)p.size();
  off.assign(npoin + 1, 0);
  for (int64_t e = 0; e < ne; ++e) {
    ++off[p[e] + 1];
    ++off[q[e] + 1];
  }
  for (int i = 0; i < npoin; ++i) off[i + 1] += off[i];
  nb.resize(off[npoin]);
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < ne; ++e) {
    nb[pos[p[e]]++] = q[e];
    nb[pos[q[e]]++] = p[e];
  }
}

int build_layout(int npoin, const std::vector<int>& p, const std::vector<int>& q,
                 const std::vector<double>& d, int64_t ntri, const int* tri,
                 const uint8_t* besym, int block_nodes, int block_cap,
                 int chunk_slots, Layout& L, std::string& err, bool gather) {
  const int64_t ne = (int64_t)p.size();
  L = Layout();
  if (chunk_slots < 3 || chunk_slots > kChunkSlots) {
    err = "layout: chunk size out of range";
    return XG_EINVAL;
  }
  L.npoin = npoin;
  // neighbour lists for halo tracking: by the smaller end (a block's halo
  // lies beyond it), or both ways for a gather layout
  std::vector<int64_t> uoff;
  std::vector<int> up;
  if (gather)
    full_lists(npoin, p, q, uoff, up);
  else
    upper_lists(npoin, p, q, uoff, up);
  // (boundary faces are tet faces: their nodes are edge neighbours of the
  // face's smallest node, so the edge lists already cover them)
  std::vector<int> bstart;
  int prc = partition_blocks(npoin, uoff, up, block_nodes, block_cap, bstart, err);
  if (prc) return prc;
  const int nblk = (int)bstart.size() - 1;
  std::vector<int> blk_of(npoin);
  for (int b = 0; b < nblk; ++b)
    for (int i = bstart[b]; i < bstart[b + 1]; ++i) blk_of[i] = b;
  // work items per block (gather: an edge belongs to the blocks of both its
  // ends, and there are no boundary faces)
  std::vector<std::vector<int64_t>> bedges(nblk), bfaces(nblk);
  for (int64_t e = 0; e < ne; ++e) {
    const int b0 = blk_of[std::min(p[e], q[e])], b1 = blk_of[std::max(p[e], q[e])];
    bedges[b0].push_back(e);
    if (gather && b1 != b0) bedges[b1].push_back(e);
  }
  for (int64_t f = 0; f < (gather ? 0 : ntri); ++f) {
    int m = std::min(tri[f * 3], std::min(tri[f * 3 + 1], tri[f * 3 + 2]));
    bfaces[blk_of[m]].push_back(f);
  }
  L.blk.resize(nblk);
  std::vector<std::vector<std::pair<int,int>>> fix(npoin);   // node -> (block, slot)
  int nslot = 0;
  std::vector<int> local(npoin, -1);
  for (int b = 0; b < nblk; ++b) {
    Block& B = L.blk[b];
    B.own0 = bstart[b];
    B.nown = bstart[b + 1] - bstart[b];
    // halo: nodes touched by owned items outside the block, ascending
    std::vector<int> halo;
    for (int64_t e : bedges[b]) {
      for (int m : {p[e], q[e]})
        if (m >= bstart[b + 1] || m < bstart[b]) halo.push_back(m);
    }
    for (int64_t f : bfaces[b])
      for (int a = 0; a < 3; ++a)
        if (tri[f * 3 + a] >= bstart[b + 1]) halo.push_back(tri[f * 3 + a]);
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
    B.nloc = B.nown + (int)halo.size();
    if (B.nloc > block_cap) {
      err = "layout: block exceeds LDS capacity";
      return XG_EUNSUPPORTED;
    }
    B.halo_off = (int)L.halo.size();
    B.slot0 = nslot;
    for (size_t h = 0; h < halo.size(); ++h) {
      L.halo.push_back(halo[h]);
      if (!gather) fix[halo[h]].emplace_back(b, nslot + (int)h);
    }
    if (!gather) nslot += (int)halo.size();
    for (int i = 0; i < B.nown; ++i) local[B.own0 + i] = i;
    for (size_t h = 0; h < halo.size(); ++h) local[halo[h]] = B.nown + (int)h;
    // edges ordered by (smaller local end, larger local end) for locality
    auto& E = bedges[b];
    std::sort(E.begin(), E.end(), [&](int64_t a, int64_t c) {
      int a0 = std::min(local[p[a]], local[q[a]]), a1 = std::max(local[p[a]], local[q[a]]);
      int c0 = std::min(local[p[c]], local[q[c]]), c1 = std::max(local[p[c]], local[q[c]]);
      return a0 != c0 ? a0 < c0 : a1 < c1;
    });
    B.chunk0 = (int)L.chunks.size();
    // chunk builder: entries (local node, slot | sign) grouped by node, with
    // a dense table of per-node entry ends over all nloc local nodes
    auto emit = [&](int kind, int64_t first, int count,
                    const std::vector<std::pair<int, uint16_t>>& ent) {
      std::vector<std::pair<int, uint16_t>> s(ent);
      std::stable_sort(s.begin(), s.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      Chunk C;
      C.kind = kind;
      C.first = (int)first;
      C.count = count;
      C.tend_off = (int)L.tend.size();
      C.ent_off = (int)L.entries.size();
      size_t k = 0;
      for (int n = 0; n < B.nloc; ++n) {
        while (k < s.size() && s[k].first == n) L.entries.push_back(s[k++].second);
        L.tend.push_back((uint16_t)(L.entries.size() - C.ent_off));
      }
      L.chunks.push_back(C);
    };
    for (size_t i0 = 0; i0 < E.size(); i0 += (size_t)chunk_slots) {
      size_t i1 = std::min(E.size(), i0 + (size_t)chunk_slots);
      std::vector<std::pair<int, uint16_t>> ent;
      for (size_t i = i0; i < i1; ++i) {
        int64_t e = E[i];
        uint16_t slot = (uint16_t)(i - i0);
        // R(p) -= f, R(q) += f (gather: owned ends only)
        if (!gather || local[p[e]] < B.nown)
          ent.emplace_back(local[p[e]], (uint16_t)(slot | 0x8000));
        if (!gather || local[q[e]] < B.nown) ent.emplace_back(local[q[e]], slot);
      }
      emit(0, (int64_t)L.elpq.size(), (int)(i1 - i0), ent);
      for (size_t i = i0; i < i1; ++i) {
        int64_t e = E[i];
        L.elpq.push_back((uint32_t)local[p[e]] | ((uint32_t)local[q[e]] << 16));
        L.ep.push_back(p[e]);
        L.eq.push_back(q[e]);
        L.eidx.push_back(e);
        L.dx.push_back(d[e * 3 + 0]);
        L.dy.push_back(d[e * 3 + 1]);
        L.dz.push_back(d[e * 3 + 2]);
      }
    }
    auto& F = bfaces[b];
    const int fmax = chunk_slots / 3;
    for (size_t i0 = 0; i0 < F.size(); i0 += fmax) {
      size_t i1 = std::min(F.size(), i0 + (size_t)fmax);
      std::vector<std::pair<int, uint16_t>> ent;
      for (size_t i = i0; i < i1; ++i) {
        int64_t f = F[i];
        for (int a = 0; a < 3; ++a)
          ent.emplace_back(local[tri[f * 3 + a]], (uint16_t)(3 * (i - i0) + a));
      }
      emit(1, (int64_t)L.fab.size(), (int)(i1 - i0), ent);
      for (size_t i = i0; i < i1; ++i) {
        int64_t f = F[i];
        uint32_t sym = (besym[f * 3] ? 1u : 0u) | (besym[f * 3 + 1] ? 2u : 0u) |
                       (besym[f * 3 + 2] ? 4u : 0u);
        L.fab.push_back((uint32_t)local[tri[f * 3]] | ((uint32_t)local[tri[f * 3 + 1]] << 16));
        L.fcs.push_back((uint32_t)local[tri[f * 3 + 2]] | (sym << 16));
        for (int a = 0; a < 3; ++a) L.fglob.push_back(tri[f * 3 + a]);
        L.fidx.push_back(f);
      }
    }
    B.chunk1 = (int)L.chunks.size();
    for (int i = 0; i < B.nown; ++i) local[B.own0 + i] = -1;
    for (size_t h = 0; h < halo.size(); ++h) local[halo[h]] = -1;
  }
  L.nslot = nslot;
  // The block kernels prefetch a chunk's entries with a bound of 2 (edges)
  // or 3 (faces) entries per item.  A gather chunk holds fewer (only owned
  // ends), so that bound can reach past the last chunk: pad the list so the
  // prefetch stays inside the buffer.  The padding is never summed.
  L.entries.resize(L.entries.size() + 3 * (size_t)chunk_slots, 0);
  // fix-up lists: nodes receiving halo partials, slots in block order
  L.fx_off.push_back(0);
  for (int n = 0; n < npoin; ++n) {
    if (fix[n].empty()) continue;
    L.fx_node.push_back(n);
    for (auto& bs : fix[n]) L.fx_slot.push_back(bs.second);
    L.fx_off.push_back((int)L.fx_slot.size());
  }
  return XG_OK;
}

// kozak::rhs layout: blocks as above on the tets' node graph; a block owns the
// tets whose smallest node lies in it.  Tets are ordered by their local node
// ids and cut into chunks of tets_per_chunk; slot 4*i + a holds corner a of
// chunk tet i, and every local node sums its slots in slot order.
int build_tet_layout(int npoin, int64_t ntet, const size_t* inpoel,
                     int block_nodes, int block_cap, TetLayout& L,
                     std::string& err, int tets_per_chunk) {
  if (tets_per_chunk < 1 || 4 * tets_per_chunk > 32768) {
    err = "kozak layout: bad chunk size";
    return XG_EINVAL;
  }
  static const int kTetEdge[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  L = TetLayout();
  L.npoin = npoin;
  std::vector<int> p, q;
  p.reserve(ntet * 6);
  q.reserve(ntet * 6);
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      size_t a = inpoel[e * 4 + kTetEdge[k][0]], b = inpoel[e * 4 + kTetEdge[k][1]];
      if (a >= (size_t)npoin || b >= (size_t)npoin || a == b) {
        err = "kozak layout: bad tet connectivity";
        return XG_EINVAL;
      }
      p.push_back((int)a);
      q.push_back((int)b);
    }
  // unique edges suffice for the partition (duplicates only repeat marks)
  std::vector<int64_t> uoff;
  std::vector<int> up;
  upper_lists(npoin, p, q, uoff, up);
  for (int i = 0; i < npoin; ++i) {
    auto b = up.begin() + uoff[i], e = up.begin() + uoff[i + 1];
    std::sort(b, e);
  }
  {  // compact duplicates
    std::vector<int64_t> o2(npoin + 1, 0);
    std::vector<int> u2;
    u2.reserve(up.size() / 2);
    for (int i = 0; i < npoin; ++i) {
      for (int64_t k = uoff[i]; k < uoff[i + 1]; ++k)
        if (k == uoff[i] || up[k] != up[k - 1]) u2.push_back(up[k]);
      o2[i + 1] = (int64_t)u2.size();
    }
    uoff.swap(o2);
    up.swap(u2);
  }
  std::vector<int>().swap(p);
  std::vector<int>().swap(q);
  std::vector<int> bstart;
  int prc = partition_blocks(npoin, uoff, up, block_nodes, block_cap, bstart, err);
  if (prc) return prc;
  const int nblk = (int)bstart.size() - 1;
  std::vector<int> blk_of(npoin);
  for (int b = 0; b < nblk; ++b)
    for (int i = bstart[b]; i < bstart[b + 1]; ++i) blk_of[i] = b;
  std::vector<std::vector<int64_t>> btets(nblk);
  for (int64_t e = 0; e < ntet; ++e) {
    const size_t* N = inpoel + e * 4;
    size_t m = std::min(std::min(N[0], N[1]), std::min(N[2], N[3]));
    btets[blk_of[m]].push_back(e);
  }
  L.blk.resize(nblk);
  std::vector<std::vector<std::pair<int,int>>> fix(npoin);
  int nslot = 0;
  std::vector<int> local(npoin, -1);
  for (int b = 0; b < nblk; ++b) {
    Block& B = L.blk[b];
    B.own0 = bstart[b];
    B.nown = bstart[b + 1] - bstart[b];
    std::vector<int> halo;
    for (int64_t e : btets[b])
      for (int a = 0; a < 4; ++a)
        if ((int)inpoel[e * 4 + a] >= bstart[b + 1]) halo.push_back((int)inpoel[e * 4 + a]);
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
    B.nloc = B.nown + (int)halo.size();
    if (B.nloc > block_cap) {
      err = "kozak layout: block exceeds LDS capacity";
      return XG_EUNSUPPORTED;
    }
    B.halo_off = (int)L.halo.size();
    B.slot0 = nslot;
    for (size_t h = 0; h < halo.size(); ++h) {
      L.halo.push_back(halo[h]);
      fix[halo[h]].emplace_back(b, nslot + (int)h);
    }
    nslot += (int)halo.size();
    for (int i = 0; i < B.nown; ++i) local[B.own0 + i] = i;
    for (size_t h = 0; h < halo.size(); ++h) local[halo[h]] = B.nown + (int)h;
    auto& T = btets[b];
    auto key = [&](int64_t e) {
      int l[4];
      for (int a = 0; a < 4; ++a) l[a] = local[inpoel[e * 4 + a]];
      std::sort(l, l + 4);
      return ((uint64_t)l[0] << 48) | ((uint64_t)l[1] << 32) | ((uint64_t)l[2] << 16) |
             (uint64_t)l[3];
    };
    std::stable_sort(T.begin(), T.end(),
                     [&](int64_t a, int64_t c) { return key(a) < key(c); });
    B.chunk0 = (int)L.chunks.size();
    for (size_t i0 = 0; i0 < T.size(); i0 += (size_t)tets_per_chunk) {
      size_t i1 = std::min(T.size(), i0 + (size_t)tets_per_chunk);
      std::vector<std::pair<int, uint16_t>> ent;
      for (size_t i = i0; i < i1; ++i)
        for (int a = 0; a < 4; ++a)
          ent.emplace_back(local[inpoel[T[i] * 4 + a]], (uint16_t)(4 * (i - i0) + a));
      std::stable_sort(ent.begin(), ent.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      Chunk C;
      C.kind = 2;
      C.first = (int)L.tab.size();
      C.count = (int)(i1 - i0);
      C.tend_off = (int)L.tend.size();
      C.ent_off = (int)L.entries.size();
      size_t k = 0;
      for (int n = 0; n < B.nloc; ++n) {
        while (k < ent.size() && ent[k].first == n) L.entries.push_back(ent[k++].second);
        L.tend.push_back((uint16_t)(L.entries.size() - C.ent_off));
      }
      L.chunks.push_back(C);
      for (size_t i = i0; i < i1; ++i) {
        const size_t* N = inpoel + T[i] * 4;
        L.tab.push_back((uint32_t)local[N[0]] | ((uint32_t)local[N[1]] << 16));
        L.tcd.push_back((uint32_t)local[N[2]] | ((uint32_t)local[N[3]] << 16));
        for (int a = 0; a < 4; ++a) L.tglob.push_back((int)N[a]);
      }
    }
    B.chunk1 = (int)L.chunks.size();
    for (int i = 0; i < B.nown; ++i) local[B.own0 + i] = -1;
    for (size_t h = 0; h < halo.size(); ++h) local[halo[h]] = -1;
  }
  L.nslot = nslot;
  L.fx_off.push_back(0);
  for (int n = 0; n < npoin; ++n) {
    if (fix[n].empty()) continue;
    L.fx_node.push_back(n);
    for (auto& bs : fix[n]) L.fx_slot.push_back(bs.second);
    L.fx_off.push_back((int)L.fx_slot.size());
  }
  return XG_OK;
}

// Recursive coordinate bisection into cells of at most `leaf` nodes; the
// order of nodes inside a cell is their original order.
static void rcb(std::vector<int>& ids, size_t lo, size_t hi, const double* x,
                const double* y, const double* z, int leaf) {
  if (hi - lo <= (size_t)leaf) {
    std::sort(ids.begin() + lo, ids.begin() + hi);
    return;
  }
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  const double* X[3] = {x, y, z};
  for (size_t i = lo; i < hi; ++i)
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], X[a][ids[i]]);
      mx[a] = std::max(mx[a], X[a][ids[i]]);
    }
  int ax = 0;
  for (int a = 1; a < 3; ++a)
    if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
  // split so that the left part holds a whole number of leaves
  size_t n = hi - lo;
  size_t nleaf = (n + leaf - 1) / leaf;
  size_t mid = lo + ((nleaf / 2) * n) / nleaf;
  const double* C = X[ax];
  std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                   [&](int a, int b) { return C[a] < C[b] || (C[a] == C[b] && a < b); });
  rcb(ids, lo, mid, x, y, z, leaf);
  rcb(ids, mid, hi, x, y, z, leaf);
}

int block_order(int64_t npoin, const double* x, const double* y,
                const double* z, int leaf, size_t* perm) {
  std::vector<int> ids(npoin);
  std::iota(ids.begin(), ids.end(), 0);
  rcb(ids, 0, ids.size(), x, y, z, leaf);
  for (int64_t i = 0; i < npoin; ++i) perm[ids[i]] = (size_t)i;
  return XG_OK;
}

}  // namespace xg

extern "C" int xg_mesh_block_order(int64_t npoin, const double* x,
                                   const double* y, const double* z,
                                   int64_t cell_nodes, size_t* perm) {
  if (npoin <= 0 || npoin >= (1ll << 31) || !x || !y || !z || !perm ||
      cell_nodes < 1)
    return XG_EINVAL;
  return xg::block_order(npoin, x, y, z, (int)cell_nodes, perm);
}
