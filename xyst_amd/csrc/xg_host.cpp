// Host-side mesh preprocessing of the edge-based flow path (native C++, no
// device work): nodal volumes, edge integrals, locality renumbering.
//
// These replace the setup loops of the reference's scheme chares
// (Discretization::vol, RieCG::domint, the constructor renumbering) with
// sort/counting based O(T log deg) passes that scale to 10^8 tets; the
// arithmetic of each per-tet contribution is the reference's.
#include "xg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace xg {

static const int kLpoed[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};

static inline void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1]*b[2] - b[1]*a[2];
  r[1] = a[2]*b[0] - b[2]*a[0];
  r[2] = a[0]*b[1] - b[0]*a[1];
}

// Per-tet basis-gradient cross products, RieCG.cpp:360-368.
static inline void tet_grads(const double* x, const double* y, const double* z,
                             const size_t* N, double g[4][3]) {
  double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
  double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
  double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
  cross3(ca, da, g[1]);
  cross3(da, ba, g[2]);
  cross3(ba, ca, g[3]);
  for (int i = 0; i < 3; ++i) g[0][i] = -g[1][i] - g[2][i] - g[3][i];
}

int mesh_vol(int64_t npoin, const double* x, const double* y, const double* z,
             int64_t ntet, const size_t* inpoel, double* vol,
             std::string& err) {
  std::fill(vol, vol + npoin, 0.0);
  for (int64_t e = 0; e < ntet; ++e) {
    const size_t* N = inpoel + e*4;
    for (int j = 0; j < 4; ++j)
      if (N[j] >= (size_t)npoin) { err = "xg_mesh_vol: node id out of range"; return XG_EINVAL; }
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double c[3];
    cross3(ca, da, c);
    double J = (ba[0]*c[0] + ba[1]*c[1] + ba[2]*c[2]) / 24.0;
    if (!(J > 0)) {
      err = "Element Jacobian non-positive: tet " + std::to_string(e);
      return XG_EINVAL;
    }
    for (int j = 0; j < 4; ++j) vol[N[j]] += J;
  }
  return XG_OK;
}

// Edge integrals of RieCG::domint (stride 3) or ChoCG::domint (stride 5:
// also J/120 and grad_p.grad_q/J/6, ChoCG.cpp:399-445).  Records (lower
// node, upper node, tet-edge id) are bucketed by the lower node with a
// counting sort, each bucket is sorted by (upper node, tet-edge id) so that
// contributions are summed in tet order -- the order the reference's hash
// map sees them.
int mesh_domint(int64_t npoin, const double* x, const double* y,
                const double* z, int64_t ntet, const size_t* inpoel,
                std::vector<size_t>& edge, std::vector<double>& dint,
                bool count_only, std::string& err, int stride) {
  if (ntet * 6 >= (1ll << 32)) { err = "xg_mesh_domint: too many tets"; return XG_EINVAL; }
  const int64_t nrec = ntet * 6;
  std::vector<uint32_t> cnt((size_t)npoin + 1, 0);
  for (int64_t e = 0; e < ntet; ++e) {
    const size_t* N = inpoel + e*4;
    for (int j = 0; j < 4; ++j)
      if (N[j] >= (size_t)npoin) { err = "xg_mesh_domint: node id out of range"; return XG_EINVAL; }
    for (int k = 0; k < 6; ++k) {
      size_t a = N[kLpoed[k][0]], b = N[kLpoed[k][1]];
      ++cnt[std::min(a, b) + 1];
    }
  }
  std::vector<uint64_t> off((size_t)npoin + 1, 0);
  for (int64_t p = 0; p < npoin; ++p) off[p+1] = off[p] + cnt[p+1];
  // key = (upper node << 32) | tet-edge record id: sorting a bucket orders it
  // by upper node, then by tet (record ids grow with the tet index)
  std::vector<uint64_t> key((size_t)nrec);
  {
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ntet; ++e) {
      const size_t* N = inpoel + e*4;
      for (int k = 0; k < 6; ++k) {
        size_t a = N[kLpoed[k][0]], b = N[kLpoed[k][1]];
        key[pos[std::min(a, b)]++] =
            ((uint64_t)std::max(a, b) << 32) | (uint64_t)(e*6 + k);
      }
    }
  }
  edge.clear(); dint.clear();
  edge.reserve((size_t)nrec / 3); dint.reserve((size_t)nrec / 6 * stride);
  for (int64_t p = 0; p < npoin; ++p) {
    uint64_t* row = key.data() + off[p];
    size_t nr = (size_t)(off[p+1] - off[p]);
    std::sort(row, row + nr);
    for (size_t i = 0; i < nr; ++i) {
      uint64_t q = row[i] >> 32;
      if (i == 0 || (row[i-1] >> 32) != q) {
        edge.push_back((size_t)p); edge.push_back((size_t)q);
        for (int j = 0; j < stride; ++j) dint.push_back(0.0);
      }
      if (count_only) continue;
      uint32_t r = (uint32_t)(row[i] & 0xffffffffu);
      int64_t e = r / 6; int k = (int)(r % 6);
      const size_t* N = inpoel + e*4;
      double g[4][3];
      tet_grads(x, y, z, N, g);
      int a = kLpoed[k][0], b = kLpoed[k][1];
      double sig = N[a] > N[b] ? -1.0 : 1.0;
      double* d = dint.data() + dint.size() - stride;
      for (int j = 0; j < 3; ++j) d[j] += sig * (g[a][j] - g[b][j]) / 48.0;
      if (stride == 5) {
        const double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
        const double J = ba[0]*g[1][0] + ba[1]*g[1][1] + ba[2]*g[1][2];
        d[3] += J / 120.0;
        d[4] += (g[a][0]*g[b][0] + g[a][1]*g[b][1] + g[a][2]*g[b][2]) / J / 6.0;
      }
    }
  }
  return XG_OK;
}

// Node adjacency (CSR, neighbours sorted) from tets.
static void node_graph(int64_t npoin, int64_t ntet, const size_t* inpoel,
                       std::vector<int64_t>& off, std::vector<int32_t>& adj) {
  std::vector<int64_t> cnt((size_t)npoin + 1, 0);
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      size_t a = inpoel[e*4 + kLpoed[k][0]], b = inpoel[e*4 + kLpoed[k][1]];
      ++cnt[a+1]; ++cnt[b+1];
    }
  off.assign((size_t)npoin + 1, 0);
  for (int64_t p = 0; p < npoin; ++p) off[p+1] = off[p] + cnt[p+1];
  std::vector<int32_t> raw((size_t)off[npoin]);
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      size_t a = inpoel[e*4 + kLpoed[k][0]], b = inpoel[e*4 + kLpoed[k][1]];
      raw[pos[a]++] = (int32_t)b; raw[pos[b]++] = (int32_t)a;
    }
  std::vector<int64_t> noff((size_t)npoin + 1, 0);
  adj.clear(); adj.reserve(raw.size() / 4);
  for (int64_t p = 0; p < npoin; ++p) {
    auto b = raw.begin() + off[p], e = raw.begin() + off[p+1];
    std::sort(b, e);
    auto u = std::unique(b, e);
    adj.insert(adj.end(), b, u);
    noff[p+1] = (int64_t)adj.size();
  }
  off.swap(noff);
}

// BFS level structure from root; returns the last node visited and the
// eccentricity (number of levels).
static int64_t bfs_far(int64_t root, const std::vector<int64_t>& off,
                       const std::vector<int32_t>& adj,
                       std::vector<int32_t>& level, int& depth) {
  std::fill(level.begin(), level.end(), -1);
  std::vector<int64_t> q; q.reserve(level.size());
  q.push_back(root); level[root] = 0;
  int64_t last = root;
  for (size_t h = 0; h < q.size(); ++h) {
    int64_t p = q[h]; last = p;
    for (int64_t i = off[p]; i < off[p+1]; ++i) {
      int32_t r = adj[i];
      if (level[r] < 0) { level[r] = level[p] + 1; q.push_back(r); }
    }
  }
  depth = level[last];
  // among the last level prefer the minimum-degree node
  int64_t best = last;
  for (int64_t p : q)
    if (level[p] == depth && off[p+1]-off[p] < off[best+1]-off[best]) best = p;
  return best;
}

int mesh_rcm(int64_t npoin, int64_t ntet, const size_t* inpoel, size_t* perm,
             std::string& err) {
  for (int64_t i = 0; i < ntet*4; ++i)
    if (inpoel[i] >= (size_t)npoin) { err = "xg_mesh_rcm: node id out of range"; return XG_EINVAL; }
  std::vector<int64_t> off;
  std::vector<int32_t> adj;
  node_graph(npoin, ntet, inpoel, off, adj);
  std::vector<int32_t> level((size_t)npoin);
  std::vector<char> done((size_t)npoin, 0);
  std::vector<int64_t> order; order.reserve((size_t)npoin);
  for (int64_t s = 0; s < npoin; ++s) {
    if (done[s]) continue;
    // pseudo-peripheral start node (George-Liu): repeat BFS while the
    // eccentricity grows
    int64_t root = s; int depth = -1;
    for (int it = 0; it < 8; ++it) {
      int d;
      int64_t far = bfs_far(root, off, adj, level, d);
      if (d <= depth) break;
      depth = d; root = far;
    }
    // Cuthill-McKee: BFS visiting unvisited neighbours by increasing degree
    size_t head = order.size();
    order.push_back(root); done[root] = 1;
    std::vector<std::pair<int64_t,int32_t>> nb;
    for (size_t h = head; h < order.size(); ++h) {
      int64_t p = order[h];
      nb.clear();
      for (int64_t i = off[p]; i < off[p+1]; ++i) {
        int32_t r = adj[i];
        if (!done[r]) { nb.emplace_back(off[r+1]-off[r], r); done[r] = 1; }
      }
      std::sort(nb.begin(), nb.end());
      for (auto& v : nb) order.push_back(v.second);
    }
  }
  // reverse
  for (int64_t i = 0; i < npoin; ++i) perm[order[i]] = (size_t)(npoin - 1 - i);
  return XG_OK;
}

// Symmetry BC data of RieCG::bndint/bnorm/streamable (RieCG.cpp:281-337,
// 483-618): per symmetry side set, boundary point normals weighted by the
// inverse squared distance from each face centroid, normalised by the weight
// sum; one (node, normal) pair per node and symmetry set, sets in the given
// order; besym = 1 for boundary-face nodes on any symmetry set.
int mesh_symbc(int64_t npoin, const double* x, const double* y,
               const double* z, int64_t ntri, const size_t* tri,
               const int32_t* faceset, int32_t nsym, const int32_t* symsets,
               std::vector<size_t>& nodes, std::vector<double>& norms,
               uint8_t* besym, std::string& err) {
  for (int64_t i = 0; i < ntri*3; ++i)
    if (tri[i] >= (size_t)npoin) { err = "xg_mesh_symbc: node id out of range"; return XG_EINVAL; }
  // per (set, node): r*n (3) and r; keyed by set-major dense index
  std::vector<int32_t> slot((size_t)npoin, -1);    // node -> compact id
  std::vector<size_t> bn;                           // compact id -> node
  for (int64_t f = 0; f < ntri; ++f) {
    bool sym = false;
    for (int32_t s = 0; s < nsym; ++s) sym |= faceset[f] == symsets[s];
    if (!sym) continue;
    for (int a = 0; a < 3; ++a)
      if (slot[tri[f*3+a]] < 0) { slot[tri[f*3+a]] = (int32_t)bn.size(); bn.push_back(tri[f*3+a]); }
  }
  const size_t nb = bn.size();
  std::vector<double> acc((size_t)nsym * nb * 4, 0.0);
  std::vector<char> has((size_t)nsym * nb, 0);
  for (int32_t s = 0; s < nsym; ++s)
    for (int64_t f = 0; f < ntri; ++f) {
      if (faceset[f] != symsets[s]) continue;
      const size_t* N = tri + f*3;
      double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
      double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
      double n[3];
      cross3(ba, ca, n);
      double A2 = std::sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
      n[0] /= A2; n[1] /= A2; n[2] /= A2;
      double cx = (x[N[0]] + x[N[1]] + x[N[2]]) / 3.0;
      double cy = (y[N[0]] + y[N[1]] + y[N[2]]) / 3.0;
      double cz = (z[N[0]] + z[N[1]] + z[N[2]]) / 3.0;
      for (int a = 0; a < 3; ++a) {
        size_t p = N[a];
        double r = 1.0 / ((cx - x[p])*(cx - x[p]) + (cy - y[p])*(cy - y[p]) +
                          (cz - z[p])*(cz - z[p]));
        size_t k = (size_t)s * nb + (size_t)slot[p];
        double* b = acc.data() + k*4;
        b[0] += r*n[0]; b[1] += r*n[1]; b[2] += r*n[2]; b[3] += r;
        has[k] = 1;
      }
    }
  nodes.clear(); norms.clear();
  std::vector<size_t> sorted(bn);
  std::sort(sorted.begin(), sorted.end());
  for (size_t p : sorted)
    for (int32_t s = 0; s < nsym; ++s) {
      size_t k = (size_t)s * nb + (size_t)slot[p];
      if (!has[k]) continue;
      const double* b = acc.data() + k*4;
      nodes.push_back(p);
      norms.push_back(b[0] / b[3]);
      norms.push_back(b[1] / b[3]);
      norms.push_back(b[2] / b[3]);
    }
  for (int64_t i = 0; i < ntri*3; ++i) besym[i] = slot[tri[i]] >= 0 ? 1 : 0;
  return XG_OK;
}

// RieCG::domsuped (RieCG.cpp:620-736), ChoCG::domsuped (ChoCG.cpp:687-813):
// the unique edges (sorted by (p, q), p < q) grouped into tet superedges,
// then triangle superedges over the tets' faces in first-seen order, then
// the remaining edges.  An edge is looked up by a binary search within the
// row of its smaller node.  Faces are visited as (tet, face) in order: the
// first visit of a face is its first-seen place, and a later visit never
// qualifies (if it was taken its edges are used; if not, one of its edges
// was already used, and edges are never released), so the face set needs no
// deduplication.  The first three values of an integral flip their sign when
// the superedge orients the edge q -> p.
static const int kLpofa[4][3] = {{1, 2, 3}, {2, 0, 3}, {3, 0, 1}, {0, 2, 1}};
static const int kLpoet[3][2] = {{0, 1}, {1, 2}, {2, 0}};

struct Superedges {
  std::vector<size_t> sup[3];
  std::vector<double> val[3];
};

int mesh_domsuped(int64_t ntet, const size_t* inpoel, int64_t nedge, const size_t* edge,
                  const double* dint, int stride, Superedges& S, std::string& err) {
  size_t np = 0;
  for (int64_t m = 0; m < nedge; ++m) {
    if (edge[2 * m] >= edge[2 * m + 1] ||
        (m && (edge[2 * m] < edge[2 * m - 2] ||
               (edge[2 * m] == edge[2 * m - 2] && edge[2 * m + 1] <= edge[2 * m - 1])))) {
      err = "xg_mesh_domsuped: edges must be unique, p < q, sorted by (p, q)";
      return XG_EINVAL;
    }
    np = std::max(np, (size_t)edge[2 * m + 1] + 1);
  }
  for (int64_t e = 0; e < 4 * ntet; ++e) np = std::max(np, (size_t)inpoel[e] + 1);
  std::vector<int64_t> row(np + 1, 0);
  for (int64_t m = 0; m < nedge; ++m) ++row[edge[2 * m] + 1];
  for (size_t a = 0; a < np; ++a) row[a + 1] += row[a];
  auto find = [&](size_t a, size_t b) -> int64_t {
    if (a > b) std::swap(a, b);
    int64_t lo = row[a], hi = row[a + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (edge[2 * mid + 1] < b) lo = mid + 1; else hi = mid;
    }
    return (lo < row[a + 1] && edge[2 * lo + 1] == b) ? lo : -1;
  };
  std::vector<char> used(nedge, 0);
  auto take = [&](int k, const size_t* N, int nn, const int (*lp)[2], int ne) {
    int64_t id[6];
    for (int j = 0; j < ne; ++j) {
      id[j] = find(N[lp[j][0]], N[lp[j][1]]);
      if (id[j] < 0 || used[id[j]]) return;
    }
    for (int i = 0; i < nn; ++i) S.sup[k].push_back(N[i]);
    for (int j = 0; j < ne; ++j) {
      const double sg = N[lp[j][0]] < N[lp[j][1]] ? 1.0 : -1.0;
      for (int v = 0; v < stride; ++v)
        S.val[k].push_back((v < 3 ? sg : 1.0) * dint[id[j] * stride + v]);
      used[id[j]] = 1;
    }
  };
  for (int64_t e = 0; e < ntet; ++e) take(0, inpoel + 4 * e, 4, kLpoed, 6);
  for (int64_t e = 0; e < ntet; ++e)
    for (int f = 0; f < 4; ++f) {
      const size_t N[3] = {inpoel[4 * e + kLpofa[f][0]], inpoel[4 * e + kLpofa[f][1]],
                           inpoel[4 * e + kLpofa[f][2]]};
      take(1, N, 3, kLpoet, 3);
    }
  for (int64_t m = 0; m < nedge; ++m) {
    if (used[m]) continue;
    S.sup[2].push_back(edge[2 * m]);
    S.sup[2].push_back(edge[2 * m + 1]);
    for (int v = 0; v < stride; ++v) S.val[2].push_back(dint[m * stride + v]);
  }
  if (S.sup[0].size() / 4 * 6 + S.sup[1].size() / 3 * 3 + S.sup[2].size() / 2 !=
      (size_t)nedge) {
    err = "xg_mesh_domsuped: the tets do not hold every edge";   // RieCG.cpp:733-735
    return XG_EINVAL;
  }
  return XG_OK;
}

}  // namespace xg

// ------------------------------------------------------------- C-ABI ------

extern "C" int xg_mesh_domsuped(int64_t ntet, const size_t* inpoel, int64_t nedge,
                                const size_t* edge, const double* dint, int stride,
                                int64_t nsup[3], size_t* sup0, double* int0,
                                size_t* sup1, double* int1, size_t* sup2,
                                double* int2) {
  if (ntet < 0 || nedge < 0 || (ntet && !inpoel) || (nedge && (!edge || !dint)) ||
      (stride != 3 && stride != 5) || !nsup)
    return XG_EINVAL;
  xg::Superedges S;
  std::string err;
  int rc = xg::mesh_domsuped(ntet, inpoel, nedge, edge, dint, stride, S, err);
  if (rc) return rc;
  const int nn[3] = {4, 3, 2};
  for (int k = 0; k < 3; ++k) nsup[k] = (int64_t)(S.sup[k].size() / nn[k]);
  if (!sup0) return XG_OK;
  size_t* so[3] = {sup0, sup1, sup2};
  double* vo[3] = {int0, int1, int2};
  for (int k = 0; k < 3; ++k) {
    if (S.sup[k].empty()) continue;
    if (!so[k] || !vo[k]) return XG_EINVAL;
    std::memcpy(so[k], S.sup[k].data(), S.sup[k].size() * sizeof(size_t));
    std::memcpy(vo[k], S.val[k].data(), S.val[k].size() * sizeof(double));
  }
  return XG_OK;
}

extern "C" int xg_mesh_symbc(int64_t npoin, const double* x, const double* y,
                             const double* z, int64_t ntri, const size_t* tri,
                             const int32_t* faceset, int32_t nsym,
                             const int32_t* symsets, int64_t cap,
                             int64_t* nout, size_t* nodes, double* norms,
                             uint8_t* besym) {
  if (npoin <= 0 || ntri < 0 || nsym < 0 || !x || !y || !z || !nout ||
      (ntri && (!tri || !faceset || !besym)) || (nsym && !symsets) || cap < 0 ||
      (cap && (!nodes || !norms)))
    return XG_EINVAL;
  std::vector<size_t> nd;
  std::vector<double> nr;
  std::string err;
  int rc = xg::mesh_symbc(npoin, x, y, z, ntri, tri, faceset, nsym, symsets,
                          nd, nr, besym, err);
  if (rc) return rc;
  *nout = (int64_t)nd.size();
  if (cap == 0) return XG_OK;
  if (cap < *nout) return XG_EINVAL;
  std::memcpy(nodes, nd.data(), nd.size() * sizeof(size_t));
  std::memcpy(norms, nr.data(), nr.size() * sizeof(double));
  return XG_OK;
}

extern "C" int xg_mesh_neumann(int64_t npoin, const double* x, const double* y,
                               const double* z, int64_t ntri, const size_t* tri,
                               const uint8_t* besym, const double* pg, double* neubc) {
  if (npoin <= 0 || ntri < 0 || !x || !y || !z || !pg || !neubc ||
      (ntri && (!tri || !besym)))
    return XG_EINVAL;
  for (int64_t i = 0; i < npoin; ++i) neubc[i] = 0.0;
  for (int64_t e = 0; e < ntri; ++e) {
    if (!besym[e]) continue;
    const size_t* N = tri + 3 * e;
    if (N[0] >= (size_t)npoin || N[1] >= (size_t)npoin || N[2] >= (size_t)npoin)
      return XG_EINVAL;
    // one third of the face's outward area vector: (x1-x0) x (x2-x0) / 6
    const double ax = x[N[1]] - x[N[0]], ay = y[N[1]] - y[N[0]], az = z[N[1]] - z[N[0]];
    const double bx = x[N[2]] - x[N[0]], by = y[N[2]] - y[N[0]], bz = z[N[2]] - z[N[0]];
    const double n0 = (ay * bz - by * az) / 6.0;
    const double n1 = (az * bx - bz * ax) / 6.0;
    const double n2 = (ax * by - bx * ay) / 6.0;
    for (int k = 0; k < 3; ++k) {
      const double* g = pg + 3 * N[k];
      neubc[N[k]] -= n0 * g[0] + n1 * g[1] + n2 * g[2];
    }
  }
  return XG_OK;
}

extern "C" int xg_mesh_vol(int64_t npoin, const double* x, const double* y,
                           const double* z, int64_t ntet, const size_t* inpoel,
                           double* vol) {
  if (npoin < 0 || ntet < 0 || !x || !y || !z || (ntet && !inpoel) || !vol)
    return XG_EINVAL;
  std::string err;
  return xg::mesh_vol(npoin, x, y, z, ntet, inpoel, vol, err);
}

extern "C" int xg_mesh_domint(int64_t npoin, const double* x, const double* y,
                              const double* z, int64_t ntet,
                              const size_t* inpoel, int64_t cap,
                              int64_t* nedge, size_t* edge, double* dint) {
  if (npoin < 0 || npoin >= (1ll << 31) || ntet < 0 || !x || !y || !z ||
      (ntet && !inpoel) || !nedge || cap < 0 || (cap && (!edge || !dint)))
    return XG_EINVAL;
  std::vector<size_t> e;
  std::vector<double> d;
  std::string err;
  int rc = xg::mesh_domint(npoin, x, y, z, ntet, inpoel, e, d, cap == 0, err);
  if (rc) return rc;
  *nedge = (int64_t)(e.size() / 2);
  if (cap == 0) return XG_OK;
  if (cap < *nedge) return XG_EINVAL;
  std::memcpy(edge, e.data(), e.size() * sizeof(size_t));
  std::memcpy(dint, d.data(), d.size() * sizeof(double));
  return XG_OK;
}

extern "C" int xg_mesh_chocg_domint(int64_t npoin, const double* x, const double* y,
                                    const double* z, int64_t ntet,
                                    const size_t* inpoel, int64_t cap,
                                    int64_t* nedge, size_t* edge, double* dint) {
  if (npoin < 0 || npoin >= (1ll << 31) || ntet < 0 || !x || !y || !z ||
      (ntet && !inpoel) || !nedge || cap < 0 || (cap && (!edge || !dint)))
    return XG_EINVAL;
  std::vector<size_t> e;
  std::vector<double> d;
  std::string err;
  int rc = xg::mesh_domint(npoin, x, y, z, ntet, inpoel, e, d, cap == 0, err, 5);
  if (rc) return rc;
  *nedge = (int64_t)(e.size() / 2);
  if (cap == 0) return XG_OK;
  if (cap < *nedge) return XG_EINVAL;
  std::memcpy(edge, e.data(), e.size() * sizeof(size_t));
  std::memcpy(dint, d.data(), d.size() * sizeof(double));
  return XG_OK;
}

extern "C" int xg_mesh_rcm(int64_t npoin, int64_t ntet, const size_t* inpoel,
                           size_t* perm) {
  if (npoin <= 0 || npoin >= (1ll << 31) || ntet < 0 || (ntet && !inpoel) || !perm)
    return XG_EINVAL;
  std::string err;
  return xg::mesh_rcm(npoin, ntet, inpoel, perm, err);
}
