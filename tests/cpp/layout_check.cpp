// CPU check of the rhs block layout (xyst_amd/csrc/xg_layout.cpp): emulates
// the block and fix-up kernels with integer-valued fluxes and compares with a
// direct scatter, bounds-checking every index the device kernels use.
// Input file (raw little-endian): int64 npoin, ne, ntri; int32 p[ne], q[ne];
// int32 tri[3*ntri]; uint8 besym[3*ntri].  Optional arguments: owned nodes,
// block capacity and chunk slots (default: the ChoCG blocks; the rhs blocks
// are kRhsNodes kRhsCap kRhsSlots), then 1 for a gather layout (the grad
// blocks: kGradNodes kGradCap kGradSlots 1), whose halo nodes must receive
// nothing and whose owned sums must be complete without a fix-up.  Prints
// "ok <stats>" or exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xyst_amd/csrc/xg_layout.hpp"

#define CHECK(c) do { if (!(c)) { std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); std::exit(1); } } while (0)

static double fval(long i, int c) { return (double)((i * 37 + c * 11) % 2001 - 1000); }

int main(int argc, char** argv) {
  CHECK(argc == 2 || argc == 5 || argc == 6);
  const bool gather = argc == 6 && std::atoi(argv[5]) != 0;
  const int nodes = argc >= 5 ? std::atoi(argv[2]) : xg::kBlockNodes;
  const int cap = argc >= 5 ? std::atoi(argv[3]) : xg::kBlockCap;
  const int slots = argc >= 5 ? std::atoi(argv[4]) : xg::kChunkSlots;
  FILE* f = std::fopen(argv[1], "rb");
  CHECK(f);
  int64_t hdr[3];
  CHECK(std::fread(hdr, 8, 3, f) == 3);
  int npoin = (int)hdr[0];
  int64_t ne = hdr[1], ntri = hdr[2];
  std::vector<int> p(ne), q(ne), tri(3 * ntri);
  std::vector<uint8_t> besym(3 * ntri);
  CHECK(std::fread(p.data(), 4, ne, f) == (size_t)ne);
  CHECK(std::fread(q.data(), 4, ne, f) == (size_t)ne);
  CHECK(std::fread(tri.data(), 4, 3 * ntri, f) == (size_t)(3 * ntri));
  CHECK(std::fread(besym.data(), 1, 3 * ntri, f) == (size_t)(3 * ntri));
  std::fclose(f);
  std::vector<double> d(3 * ne);
  for (int64_t i = 0; i < 3 * ne; ++i) d[i] = (double)((i * 13) % 17 - 8);
  // edge e's flux is identified by (p,q) so that layout reordering keeps it
  auto eflux = [&](int a, int b, int c) { return fval((long)a * 7919 + b, c); };
  auto fflux = [&](int a, int b, int cc, int k, int c) {
    return fval((long)a * 131 + (long)b * 17 + cc + k * 5, c);
  };
  xg::Layout L;
  std::string err;
  CHECK(xg::build_layout(npoin, p, q, d, ntri, tri.data(), besym.data(), nodes, cap,
                         slots, L, err, gather) == 0);
  // direct scatter
  std::vector<double> Rref((size_t)npoin * 5, 0.0), R((size_t)npoin * 5, -1e300);
  for (int64_t e = 0; e < ne; ++e)
    for (int c = 0; c < 5; ++c) {
      Rref[p[e] * 5 + c] -= eflux(p[e], q[e], c);
      Rref[q[e] * 5 + c] += eflux(p[e], q[e], c);
    }
  for (int64_t t = 0; t < (gather ? 0 : ntri); ++t)   // gather layouts carry no faces
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 5; ++c)
        Rref[tri[t * 3 + a] * 5 + c] += fflux(tri[t * 3], tri[t * 3 + 1], tri[t * 3 + 2], a, c);
  // emulate k_rhs_blocks: local node data, per-thread node accumulators
  std::vector<double> S((size_t)L.nslot * 5, -1e300);
  if (gather)   // edges between blocks twice, no faces, no halo partials
    CHECK((int64_t)L.elpq.size() >= ne && L.fab.empty() && L.nslot == 0 &&
          L.fx_node.empty());
  else
    CHECK((int64_t)L.elpq.size() == ne && (int64_t)L.fab.size() == ntri);
  std::vector<int> owned(npoin, 0);
  long maxloc = 0, halo = 0;
  for (size_t b = 0; b < L.blk.size(); ++b) {
    const auto& B = L.blk[b];
    CHECK(B.nloc <= cap && B.nown <= B.nloc && B.nown > 0);
    CHECK(B.nloc <= slots);                    // one local node per thread
    CHECK(B.own0 >= 0 && B.own0 + B.nown <= npoin);
    if (!gather) CHECK(B.slot0 + (B.nloc - B.nown) <= L.nslot);
    CHECK(B.halo_off + (B.nloc - B.nown) <= (int)L.halo.size());
    maxloc = std::max<long>(maxloc, B.nloc);
    halo += B.nloc - B.nown;
    for (int i = 0; i < B.nown; ++i) owned[B.own0 + i]++;
    std::vector<int> glob(B.nloc);
    for (int i = 0; i < B.nloc; ++i) {
      glob[i] = i < B.nown ? B.own0 + i : L.halo[B.halo_off + i - B.nown];
      CHECK(glob[i] >= 0 && glob[i] < npoin);
    }
    std::vector<double> acc((size_t)B.nloc * 5, 0.0), F((size_t)slots * 5);
    for (int ch = B.chunk0; ch < B.chunk1; ++ch) {
      const auto& C = L.chunks[ch];
      CHECK(C.count > 0 && C.count <= (C.kind ? slots / 3 : slots));
      for (int t = 0; t < C.count; ++t) {
        if (C.kind == 0) {
          int e = C.first + t;
          CHECK(e >= 0 && e < (int64_t)L.elpq.size());
          int lp = L.elpq[e] & 0xffff, lq = L.elpq[e] >> 16;
          CHECK(lp < B.nloc && lq < B.nloc);
          CHECK(glob[lp] == L.ep[e] && glob[lq] == L.eq[e]);
          for (int c = 0; c < 5; ++c) F[t * 5 + c] = eflux(glob[lp], glob[lq], c);
        } else {
          int ff = C.first + t;
          CHECK(ff >= 0 && ff < ntri);
          int la = L.fab[ff] & 0xffff, lb = L.fab[ff] >> 16, lc = L.fcs[ff] & 0xffff;
          CHECK(la < B.nloc && lb < B.nloc && lc < B.nloc);
          CHECK((int)(L.fcs[ff] >> 16) < 8);
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 5; ++c)
              F[(3 * t + a) * 5 + c] = fflux(glob[la], glob[lb], glob[lc], a, c);
        }
      }
      int nslots = C.kind ? 3 * C.count : C.count;
      CHECK(C.tend_off + B.nloc <= (int)L.tend.size());
      // the kernels' prefetch reads count * (2 or 3) entries from ent_off
      CHECK(C.ent_off + C.count * (C.kind ? 3 : 2) <= (int)L.entries.size());
      for (int n = 0; n < B.nloc; ++n) {       // thread n owns local node n
        int e0 = n ? L.tend[C.tend_off + n - 1] : 0, e1 = L.tend[C.tend_off + n];
        CHECK(e0 <= e1 && C.ent_off + e1 <= (int)L.entries.size());
        for (int j = e0; j < e1; ++j) {
          int v = L.entries[C.ent_off + j];
          int s = v & 0x3ff;
          CHECK(s < nslots);
          for (int c = 0; c < 5; ++c) acc[n * 5 + c] += (v & 0x8000) ? -F[s * 5 + c] : F[s * 5 + c];
        }
      }
    }
    for (int i = 0; i < B.nown * 5; ++i) R[(size_t)B.own0 * 5 + i] = acc[i];
    for (int i = 0; i < (B.nloc - B.nown) * 5; ++i) {
      if (gather)   // halo nodes receive nothing: the kernel stores no halo sums
        CHECK(acc[B.nown * 5 + i] == 0.0);
      else
        S[(size_t)B.slot0 * 5 + i] = acc[B.nown * 5 + i];
    }
  }
  for (int n = 0; n < npoin; ++n) CHECK(owned[n] == 1);
  // emulate k_rhs_fixup
  CHECK(L.fx_off.size() == L.fx_node.size() + 1);
  for (size_t i = 0; i < L.fx_node.size(); ++i) {
    int n = L.fx_node[i];
    CHECK(n >= 0 && n < npoin);
    for (int k = L.fx_off[i]; k < L.fx_off[i + 1]; ++k) {
      CHECK(L.fx_slot[k] >= 0 && L.fx_slot[k] < L.nslot);
      for (int c = 0; c < 5; ++c) R[n * 5 + c] += S[(size_t)L.fx_slot[k] * 5 + c];
    }
  }
  for (size_t i = 0; i < R.size(); ++i) CHECK(R[i] == Rref[i]);
  // gradient on the same blocks: node thread n gathers sign*D*(u_n + u_other)
  // over its chunk entries, the other end read from the edge's local pair;
  // integer-valued D and u make both sums exact
  {
    auto uval = [](int n) { return (double)((n * 29) % 101 - 50); };
    std::vector<double> gref((size_t)npoin * 3, 0.0), gs((size_t)npoin * 3, -1e300);
    std::vector<double> gsl((size_t)L.nslot * 3, -1e300);
    for (int64_t e = 0; e < ne; ++e)
      for (int j = 0; j < 3; ++j) {
        double tt = d[e * 3 + j] * (uval(q[e]) + uval(p[e]));
        gref[p[e] * 3 + j] -= tt;
        gref[q[e] * 3 + j] += tt;
      }
    for (size_t b = 0; b < L.blk.size(); ++b) {
      const auto& B = L.blk[b];
      std::vector<int> glob(B.nloc);
      for (int i = 0; i < B.nloc; ++i)
        glob[i] = i < B.nown ? B.own0 + i : L.halo[B.halo_off + i - B.nown];
      for (int n = 0; n < B.nloc; ++n) {
        double acc[3] = {0, 0, 0};
        for (int ch = B.chunk0; ch < B.chunk1; ++ch) {
          const auto& C = L.chunks[ch];
          if (C.kind) continue;
          int e0 = n ? L.tend[C.tend_off + n - 1] : 0, e1 = L.tend[C.tend_off + n];
          for (int j = e0; j < e1; ++j) {
            int v = L.entries[C.ent_off + j];
            int e = C.first + (v & 0x3ff);
            bool first = v & 0x8000;
            int lp = L.elpq[e] & 0xffff, lq = L.elpq[e] >> 16;
            CHECK((first ? lp : lq) == n);
            int o = first ? lq : lp;
            double D[3] = {L.dx[e], L.dy[e], L.dz[e]};
            for (int k = 0; k < 3; ++k)
              acc[k] += (first ? -D[k] : D[k]) * (uval(glob[o]) + uval(glob[n]));
          }
        }
        if (gather && n >= B.nown) {
          for (int k = 0; k < 3; ++k) CHECK(acc[k] == 0.0);
          continue;
        }
        double* out = n < B.nown ? &gs[(size_t)(B.own0 + n) * 3] : &gsl[(size_t)(B.slot0 + n - B.nown) * 3];
        for (int k = 0; k < 3; ++k) out[k] = acc[k];
      }
    }
    for (size_t i = 0; i < L.fx_node.size(); ++i)
      for (int k = L.fx_off[i]; k < L.fx_off[i + 1]; ++k)
        for (int j = 0; j < 3; ++j) gs[L.fx_node[i] * 3 + j] += gsl[(size_t)L.fx_slot[k] * 3 + j];
    for (size_t i = 0; i < gs.size(); ++i) CHECK(gs[i] == gref[i]);
  }
  std::printf("ok blocks=%zu chunks=%zu maxloc=%ld halo/owned=%.3f nfix=%zu edges/input=%.3f\n",
              L.blk.size(), L.chunks.size(), maxloc, (double)halo / npoin,
              L.fx_node.size(), (double)L.elpq.size() / (double)ne);
  return 0;
}
