// Host-side description of the device layout (see xg_layout.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "xg_internal.hpp"

namespace xg {

// ChoCG edge blocks: slots per chunk = threads per workgroup (also the
// largest chunk any layout may use: slot ids are 10-bit)
constexpr int kChunkSlots = 1024;
// nodes (owned + halo) a ChoCG block may touch; its operators stage their
// node fields for these in LDS beside one chunk of edge terms
constexpr int kBlockCap = 616;
// owned nodes per block (a block also stops growing at kBlockCap)
constexpr int kBlockNodes = 432;
// riemann::grad blocks (a gather layout): 5 primitive doubles per local
// node beside 512 slots of 8 doubles and the entry lists, 56,320 bytes of
// LDS, so two 512-thread workgroups share a CU
constexpr int kGradSlots = 512;
constexpr int kGradCap = 512;
constexpr int kGradNodes = 384;
// riemann::rhs has blocks of its own, small enough that two 512-thread
// workgroups share a CU: 23 doubles per local node, 512 flux slots of 5
// doubles and the entry lists take 80,560 of the 81,920 bytes of LDS each
constexpr int kRhsSlots = 512;
constexpr int kRhsCap = 312;
constexpr int kRhsNodes = 312;

struct Block {
  int own0 = 0, nown = 0, nloc = 0;  // owned ids [own0, own0+nown); local ids
  int halo_off = 0;                  // halo node ids at halo[halo_off ...]
  int slot0 = 0;                     // first halo output slot
  int chunk0 = 0, chunk1 = 0;        // chunks [chunk0, chunk1)
};

struct Chunk {
  int kind = 0;       // 0 edges, 1 boundary faces
  int first = 0;      // first edge / face of the chunk in layout order
  int count = 0;      // number of edges / faces
  int tend_off = 0;   // dense per-local-node entry ends (nloc of them)
  int ent_off = 0;    // entry list offset
};

struct Layout {
  int npoin = 0;
  // edges and faces in block/chunk order, end nodes as block-local ids
  std::vector<uint32_t> elpq;      // local p | local q << 16
  std::vector<double> dx, dy, dz;
  std::vector<int> ep, eq;         // global end nodes (grad stream, checks)
  std::vector<int64_t> eidx;       // input index of each layout edge
  std::vector<int64_t> fidx;       // input index of each layout face
  std::vector<uint32_t> fab;       // local a | local b << 16
  std::vector<uint32_t> fcs;       // local c | symmetry bits << 16
  std::vector<int> fglob;          // global face nodes (checks)
  std::vector<Block> blk;
  std::vector<Chunk> chunks;
  std::vector<int> halo;
  std::vector<uint16_t> tend;      // per chunk: nloc cumulative entry ends
  std::vector<uint16_t> entries;   // slot | 0x8000 for a negative sign
  int nslot = 0;
  std::vector<int> fx_node, fx_off, fx_slot;
};

// KozCG blocks: a block stages U (5) and coordinates (3) of up to kKozCap
// nodes in LDS beside one chunk's corner values (5 doubles per slot); a
// chunk holds kKozPairSlots / 4 tets, two threads per tet.  At 1024 slots
// and 512 nodes the largest of the KozCG block kernels takes 78,848 bytes of
// LDS, so two share a CU.
constexpr int kKozCap = 512;
constexpr int kKozNodes = 384;
constexpr int kKozPairSlots = 1024;
constexpr int kTetsPerPairChunk = kKozPairSlots / 4;

struct TetLayout {
  int npoin = 0;
  std::vector<uint32_t> tab, tcd;  // tets in block/chunk order: local a|b<<16, c|d<<16
  std::vector<int> tglob;          // their global nodes (4 per tet, checks)
  std::vector<Block> blk;
  std::vector<Chunk> chunks;       // kind 2
  std::vector<int> halo;
  std::vector<uint16_t> tend, entries;   // entry = slot 4*i+a
  int nslot = 0;
  std::vector<int> fx_node, fx_off, fx_slot;
};

int build_tet_layout(int npoin, int64_t ntet, const size_t* inpoel,
                     int block_nodes, int block_cap, TetLayout& L,
                     std::string& err,
                     int tets_per_chunk = kTetsPerPairChunk);

// chunk_slots: flux slots per edge chunk (<= kChunkSlots; boundary-face
// chunks hold chunk_slots / 3 faces).
// gather: a block takes every edge that touches one of its nodes (edges
// between two blocks are in both), its halo is every outside end of those,
// and its entries add each edge to its owned ends only, so every owned sum
// is complete in its block: no halo slots, no fix-up lists, and no boundary
// faces (an operator on such a layout adds boundary terms by itself)
int build_layout(int npoin, const std::vector<int>& p, const std::vector<int>& q,
                 const std::vector<double>& d, int64_t ntri, const int* tri,
                 const uint8_t* besym, int block_nodes, int block_cap,
                 int chunk_slots, Layout& L, std::string& err, bool gather = false);

int block_order(int64_t npoin, const double* x, const double* y,
                const double* z, int leaf, size_t* perm);

}  // namespace xg
