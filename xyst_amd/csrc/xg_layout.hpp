// Host-side description of the device layout (see xg_layout.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "xg_internal.hpp"

namespace xg {

// flux slots per chunk = threads per rhs workgroup
constexpr int kChunkSlots = 1024;
// nodes (owned + halo) a block may touch: its LDS accumulator holds
// kBlockCap x 5 doubles
constexpr int kBlockCap = 2560;
constexpr int kBlockNodes = 2048;

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
  int touch_off = 0;  // touched-node list offset
  int ntouch = 0;
  int ent_off = 0;    // entry list offset
};

struct Layout {
  int npoin = 0;
  // edges and faces in block/chunk order
  std::vector<int> ep, eq;
  std::vector<double> dx, dy, dz;
  std::vector<int> fa, fb, fc;
  std::vector<uint8_t> fsym;       // bit a: face node a is a symmetry node
  std::vector<Block> blk;
  std::vector<Chunk> chunks;
  std::vector<int> halo;
  std::vector<uint16_t> touched;   // local node id
  std::vector<uint16_t> tend;      // exclusive end of its entries (chunk-relative)
  std::vector<uint16_t> entries;   // slot | 0x8000 for a negative sign
  int nslot = 0;
  std::vector<int> fx_node, fx_off, fx_slot;
  std::vector<int> inc_off, inc_edge, inc_nb;
};

int build_layout(int npoin, const std::vector<int>& p, const std::vector<int>& q,
                 const std::vector<double>& d, int64_t ntri, const int* tri,
                 const uint8_t* besym, int block_nodes, int block_cap,
                 Layout& L, std::string& err);

int block_order(int64_t npoin, const double* x, const double* y,
                const double* z, int leaf, size_t* perm);

}  // namespace xg
