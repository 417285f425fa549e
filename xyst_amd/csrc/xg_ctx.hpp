// The context of one GPU-resident partition (shared by the device sources).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "xg_internal.hpp"

struct DevArr {
  void* p = nullptr;
  size_t bytes = 0;
};

struct CgSolver;   // xg_cg.hip
struct KozakMesh;  // xg_kozak.hip
struct ChoMesh;    // xg_chocg.hip
struct GnnNet;     // xg_gnn.hip
struct AmgDev;     // xg_amg.hip

struct xg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  xg_config cfg{1.4, 0, 0, 0.0};
  std::string err;
  std::vector<DevArr> mesh_allocs;     // freed on re-upload / destroy
  std::vector<DevArr> riecg_allocs;
  // mesh
  bool have_mesh = false;
  int npoin = 0;
  // riemann::rhs's own edge blocks (xg_layout.hpp: kRhsCap, kRhsSlots)
  struct EdgeBlocks {
    int nblk = 0, nfx = 0;
    int4 *blk0 = nullptr, *blk1 = nullptr, *chk = nullptr;
    int *halo = nullptr, *chk_ent = nullptr;
    int *fx_node = nullptr, *fx_off = nullptr, *fx_slot = nullptr;
    uint16_t *tend = nullptr, *entries = nullptr;
    uint32_t *elpq = nullptr, *fab = nullptr, *fcs = nullptr;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr;
    double* slots = nullptr;             // halo partials (XG_MAX_NCOMP per slot)
  } rb;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  int64_t nedge = 0;
  uint32_t* elpq = nullptr;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr;
  // grad blocks (a gather layout: no halo partials, no boundary faces)
  int nblk = 0, nchunk = 0;
  int4 *blk0 = nullptr, *blk1 = nullptr;     // (own0,nown,nloc,halo_off), (slot0,chunk0,chunk1,-)
  int *halo = nullptr;
  int4 *chk = nullptr;                       // (kind,first,count,tend_off)
  int *chk_ent = nullptr;                    // ent_off
  uint16_t *tend = nullptr, *entries = nullptr;
  // grad boundary faces
  int nbnode = 0;
  int *bnode = nullptr, *boff = nullptr, *bface = nullptr;
  int* tri = nullptr;
  // RieCG stepping
  bool have_riecg = false;
  int ncomp = 5;                         // components per node (xg_riecg_set_ncomp)
  double* vol = nullptr;
  // per node: 1 = boundary node, whose gradient k_grad_bnd completes and
  // divides by vol; the others are divided by k_grad_blocks (RK stages)
  uint8_t* gflag = nullptr;
  int nsymnode = 0;
  int *symnode = nullptr, *symoff = nullptr;
  double* symnorm = nullptr;
  // the other RieCG BCs (xg_riecg_bc): Dirichlet nodes with masks and
  // values, farfield (node, normal) pairs grouped by node, pressure
  // (node, density, pressure) records grouped by node
  std::vector<DevArr> bc_allocs;
  int ndirnode = 0, nfarnode = 0, nprenode = 0;
  int* dirnode = nullptr;
  uint8_t* dirmask = nullptr;
  double* dirval = nullptr;
  int *farnode = nullptr, *faroff = nullptr;
  double* farnorm = nullptr;
  double far[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
  int *prenode = nullptr, *preoff = nullptr;
  double* preval = nullptr;
  // source term of riemann::rhs and kozak::rhs (xg_set_problem)
  xg_problem problem{XG_PROBLEM_NONE, 0.0, 0.0, 0.0};
  double* red = nullptr;               // reduction partials + result
  // staging for the host-pointer entry points
  double *sU = nullptr, *sG = nullptr, *sR = nullptr;
  size_t sbytes = 0;                     // bytes of one component over all nodes
  int snc = 0;                           // components the buffers hold
  // pressure-Poisson solver (xg_cg.hip)
  CgSolver* cg = nullptr;
  // KozCG tet layout (xg_kozak.hip)
  KozakMesh* koz = nullptr;
  // ChoCG edge layout (xg_chocg.hip)
  ChoMesh* cho = nullptr;
  // learned preconditioner (xg_gnn.hip)
  GnnNet* gnn = nullptr;
  // diagnostics partial sums (xg_diag.hip, grow only)
  DevArr diagbuf;
  // algebraic multigrid preconditioner of the CG (xg_amg.hip)
  AmgDev* amg = nullptr;
};

void xg_kozak_free(xg_ctx* c);   // xg_kozak.hip
void xg_chocg_free(xg_ctx* c);   // xg_chocg.hip
int xg_diag_scratch_init(int device);   // xg_diag.hip, with `device` current
// rows of the context's CG matrix when it has one component per node, else -1
int64_t xg_cg_rows(const xg_ctx* c);   // xg_cg.hip

namespace xgd {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

inline int fail(xg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define XG_HIP(ctx, call)                                                    \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(ctx, XG_EHIP, std::string(#call) + ": " +                  \
                                    hipGetErrorString(e_));                  \
  } while (0)

inline int dev_alloc(xg_ctx* c, std::vector<DevArr>& list, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(c, XG_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  list.push_back({*p, bytes});
  return XG_OK;
}

inline void dev_free_all(std::vector<DevArr>& list) {
  for (auto& a : list) (void)hipFree(a.p);
  list.clear();
}

template <class T>
int upload(xg_ctx* c, std::vector<DevArr>& list, T** dst, const T* src,
           size_t n) {
  int rc = dev_alloc(c, list, (void**)dst, n * sizeof(T));
  if (rc) return rc;
  if (n) XG_HIP(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return XG_OK;
}

template <class T>
int upload(xg_ctx* c, std::vector<DevArr>& list, T** dst,
           const std::vector<T>& v) {
  return upload(c, list, dst, v.data(), v.size());
}

inline unsigned nblocks(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

}  // namespace xgd
