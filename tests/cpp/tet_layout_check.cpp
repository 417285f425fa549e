// CPU emulation of the kozak::rhs block layout (xg_layout.cpp,
// build_tet_layout) with integer corner fluxes: every tet corner's slot must
// reach exactly its node once (owned nodes directly, halo nodes through the
// fix-up slots), and every index the kernel forms must stay in bounds.
// Input file: int64 npoin, ntet; then int32 inpoel[4*ntet].  Optional second
// argument: tets per chunk (the four-lane kernels' 128, the two-threads-per-
// tet kozak::rhs's 256).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xyst_amd/csrc/xg_layout.hpp"

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

int main(int argc, char** argv) {
  CHECK(argc == 2 || argc == 3);
  const int tpc = argc == 3 ? std::atoi(argv[2]) : xg::kTetsPerPairChunk;
  CHECK(tpc >= 1 && tpc <= xg::kTetsPerPairChunk);
  FILE* f = std::fopen(argv[1], "rb");
  CHECK(f);
  int64_t hdr[2];
  CHECK(std::fread(hdr, 8, 2, f) == 2);
  const int npoin = (int)hdr[0];
  const int64_t ntet = hdr[1];
  std::vector<int32_t> t32(ntet * 4);
  CHECK(std::fread(t32.data(), 4, t32.size(), f) == t32.size());
  std::fclose(f);
  std::vector<size_t> inp(t32.begin(), t32.end());
  xg::TetLayout L;
  std::string err;
  CHECK(xg::build_tet_layout(npoin, ntet, inp.data(), xg::kKozNodes, xg::kKozCap, L, err,
                              tpc) == 0);
  // corner "flux" of tet e at corner a: a distinct integer
  auto flux = [&](int64_t e, int a) { return (int64_t)(e * 4 + a + 1); };
  std::vector<int64_t> R(npoin, 0), S(L.nslot, 0), expect(npoin, 0);
  for (int64_t e = 0; e < ntet; ++e)
    for (int a = 0; a < 4; ++a) expect[inp[e * 4 + a]] += flux(e, a);
  std::vector<int64_t> tet_of_pos(L.tab.size());
  // recover the global tet id of each layout position from its nodes
  {
    std::vector<int64_t> seen(ntet, 0);
    // tglob holds the nodes in the original corner order
    std::vector<std::vector<int64_t>> by_min(npoin);
    for (int64_t e = 0; e < ntet; ++e) {
      size_t m = inp[e * 4];
      for (int a = 1; a < 4; ++a) m = std::min(m, inp[e * 4 + a]);
      by_min[m].push_back(e);
    }
    for (size_t k = 0; k < L.tab.size(); ++k) {
      const int* g = &L.tglob[k * 4];
      int m = std::min(std::min(g[0], g[1]), std::min(g[2], g[3]));
      int64_t found = -1;
      for (int64_t e : by_min[m])
        if (!seen[e] && (int)inp[e * 4] == g[0] && (int)inp[e * 4 + 1] == g[1] &&
            (int)inp[e * 4 + 2] == g[2] && (int)inp[e * 4 + 3] == g[3]) { found = e; break; }
      CHECK(found >= 0);
      seen[found] = 1;
      tet_of_pos[k] = found;
    }
    for (int64_t e = 0; e < ntet; ++e) CHECK(seen[e]);
  }
  int maxloc = 0;
  for (size_t b = 0; b < L.blk.size(); ++b) {
    const auto& B = L.blk[b];
    CHECK(B.nloc <= xg::kKozCap && B.nown >= 1);
    maxloc = std::max(maxloc, B.nloc);
    std::vector<int> glob(B.nloc);
    for (int i = 0; i < B.nloc; ++i)
      glob[i] = i < B.nown ? B.own0 + i : L.halo[B.halo_off + i - B.nown];
    std::vector<int64_t> acc(B.nloc, 0), F(4 * tpc, 0);
    for (int ch = B.chunk0; ch < B.chunk1; ++ch) {
      const auto& C = L.chunks[ch];
      CHECK(C.kind == 2 && C.count >= 1 && C.count <= tpc);
      // the kernels load at most two entries per thread of 512
      CHECK(4 * C.count <= 2 * 512);
      for (int t = 0; t < 4 * C.count; ++t) {
        int i = t >> 2, a = t & 3;
        int64_t pos = C.first + i;
        CHECK(pos < (int64_t)L.tab.size());
        uint32_t ab = L.tab[pos], cd = L.tcd[pos];
        int N[4] = {(int)(ab & 0xffff), (int)(ab >> 16), (int)(cd & 0xffff), (int)(cd >> 16)};
        for (int k = 0; k < 4; ++k) CHECK(N[k] < B.nloc);
        CHECK(glob[N[a]] == (int)inp[tet_of_pos[pos] * 4 + a]);
        F[t] = flux(tet_of_pos[pos], a);
      }
      for (int n = 0; n < B.nloc; ++n) {
        int e0 = n ? L.tend[C.tend_off + n - 1] : 0, e1 = L.tend[C.tend_off + n];
        CHECK(e0 <= e1 && e1 <= 4 * C.count);
        for (int j = e0; j < e1; ++j) {
          int s = L.entries[C.ent_off + j];
          CHECK(s < 4 * C.count);
          acc[n] += F[s];
        }
      }
    }
    for (int n = 0; n < B.nloc; ++n) {
      if (n < B.nown) R[B.own0 + n] = acc[n];
      else S[B.slot0 + n - B.nown] = acc[n];
    }
  }
  for (size_t i = 0; i + 1 < L.fx_off.size(); ++i)
    for (int k = L.fx_off[i]; k < L.fx_off[i + 1]; ++k) R[L.fx_node[i]] += S[L.fx_slot[k]];
  for (int n = 0; n < npoin; ++n) CHECK(R[n] == expect[n]);
  std::printf("ok blocks=%zu chunks=%zu maxloc=%d halo/owned=%.3f\n", L.blk.size(),
              L.chunks.size(), maxloc, (double)L.halo.size() / npoin);
  return 0;
}
