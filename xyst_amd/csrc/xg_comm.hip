// Partitions of one mesh on one device: the exchange of partial nodal sums
// at shared nodes that the reference does with chare messages
// (Discretization's NodeCommMap; RieCG::comgrad/comrhs, RieCG.cpp:871-916,
// 948-990; ConjugateGradients::comq, ConjugateGradients.cpp:611-627).
//
// The host finds, for every global node held by more than one partition,
// its (partition, local id) copies in partition order.  One kernel then
// gives each copy the sum of all copies: one thread per (shared node,
// component) reads the copies in partition order, adds them in that order
// and writes the sum back to every copy.  Deterministic, no atomics.  On
// several GPUs the same lists become the per-neighbour send/receive sets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/xg.h"

namespace {

constexpr int kMaxParts = 64;
constexpr int kThreads = 256;

struct FieldPtrs {
  double* p[kMaxParts];
};

// copy k of shared node i: ref[k] = (partition << 26) | local id
__global__ __launch_bounds__(kThreads) void k_comm_sum(
    int64_t nshared, int ncomp, const int* __restrict__ off,
    const uint32_t* __restrict__ ref, FieldPtrs F) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshared * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t - i * ncomp);
  double s = 0.0;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const uint32_t r = ref[k];
    s += F.p[r >> 26][(size_t)(r & 0x3ffffff) * ncomp + c];
  }
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const uint32_t r = ref[k];
    F.p[r >> 26][(size_t)(r & 0x3ffffff) * ncomp + c] = s;
  }
}

}  // namespace

struct xg_comm {
  int device = 0;
  int nparts = 0;
  std::vector<int64_t> nlocal;
  int64_t nshared = 0;
  int* off = nullptr;
  uint32_t* ref = nullptr;
  std::string err;
};

namespace {

int comm_fail(xg_comm* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

}  // namespace

extern "C" {

int xg_comm_create(int device, int nparts, const int64_t* nlocal,
                   const size_t* const* gid, xg_comm** out) {
  if (!out) return XG_EINVAL;
  *out = new xg_comm;
  xg_comm* m = *out;
  m->device = device;
  if (nparts < 1 || nparts > kMaxParts || !nlocal || !gid)
    return comm_fail(m, XG_EINVAL, "xg_comm_create: 1 to 64 partitions with node lists");
  // (global id, partition, local id) of every copy; sorted by global id then
  // partition, so the copies of a node are adjacent and in partition order
  struct Copy {
    size_t g;
    uint32_t part, lid;
  };
  std::vector<Copy> all;
  for (int k = 0; k < nparts; ++k) {
    if (nlocal[k] < 0 || nlocal[k] >= (1 << 26) || (nlocal[k] && !gid[k]))
      return comm_fail(m, XG_EINVAL, "xg_comm_create: bad partition size");
    for (int64_t i = 0; i < nlocal[k]; ++i)
      all.push_back({gid[k][i], (uint32_t)k, (uint32_t)i});
  }
  std::stable_sort(all.begin(), all.end(), [](const Copy& a, const Copy& b) {
    return a.g != b.g ? a.g < b.g : a.part < b.part;
  });
  std::vector<int> off(1, 0);
  std::vector<uint32_t> ref;
  for (size_t a = 0; a < all.size();) {
    size_t b = a + 1;
    while (b < all.size() && all[b].g == all[a].g) ++b;
    for (size_t k = a + 1; k < b; ++k)
      if (all[k].part == all[k - 1].part)
        return comm_fail(m, XG_EINVAL, "xg_comm_create: a global id repeats in a partition");
    if (b - a > 1) {
      for (size_t k = a; k < b; ++k) ref.push_back((all[k].part << 26) | all[k].lid);
      off.push_back((int)ref.size());
    }
    a = b;
  }
  m->nparts = nparts;
  m->nlocal.assign(nlocal, nlocal + nparts);
  m->nshared = (int64_t)off.size() - 1;
  if (hipSetDevice(device) != hipSuccess)
    return comm_fail(m, XG_EHIP, "xg_comm_create: hipSetDevice");
  if (hipMalloc((void**)&m->off, off.size() * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&m->ref, std::max<size_t>(1, ref.size()) * sizeof(uint32_t)) !=
          hipSuccess)
    return comm_fail(m, XG_EHIP, "xg_comm_create: hipMalloc");
  if (hipMemcpy(m->off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess ||
      (ref.size() && hipMemcpy(m->ref, ref.data(), ref.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice) != hipSuccess))
    return comm_fail(m, XG_EHIP, "xg_comm_create: hipMemcpy");
  return XG_OK;
}

int xg_comm_sum_dev(xg_comm* m, int ncomp, double* const* fields, void* stream) {
  if (!m) return XG_EINVAL;
  if (!m->off) return comm_fail(m, XG_EINVAL, "xg_comm_sum: communicator not created");
  if (ncomp < 1 || ncomp > 64 || !fields)
    return comm_fail(m, XG_EINVAL, "xg_comm_sum: bad arguments");
  FieldPtrs F{};
  for (int k = 0; k < m->nparts; ++k) {
    if (!fields[k] && m->nlocal[k])
      return comm_fail(m, XG_EINVAL, "xg_comm_sum: null field of a partition");
    F.p[k] = fields[k];
  }
  if (m->nshared == 0) return XG_OK;
  if (hipSetDevice(m->device) != hipSuccess)
    return comm_fail(m, XG_EHIP, "xg_comm_sum: hipSetDevice");
  const int64_t n = m->nshared * ncomp;
  hipLaunchKernelGGL(k_comm_sum, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, (hipStream_t)stream, m->nshared, ncomp, m->off,
                     m->ref, F);
  if (hipGetLastError() != hipSuccess)
    return comm_fail(m, XG_EHIP, "xg_comm_sum: kernel launch");
  return XG_OK;
}

int xg_comm_shared(const xg_comm* m, int64_t* nshared) {
  if (!m || !nshared) return XG_EINVAL;
  *nshared = m->nshared;
  return XG_OK;
}

const char* xg_comm_error(const xg_comm* m) { return m ? m->err.c_str() : "null communicator"; }

int xg_comm_destroy(xg_comm* m) {
  if (!m) return XG_EINVAL;
  if (m->off || m->ref) {
    (void)hipSetDevice(m->device);
    if (m->off) (void)hipFree(m->off);
    if (m->ref) (void)hipFree(m->ref);
  }
  delete m;
  return XG_OK;
}

}  // extern "C"
