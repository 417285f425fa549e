/* xg.h -- C-ABI of the MI355X-native edge-based flow path.
 *
 * The entry points replace the bodies behind the reference's own call
 * signatures; the reference-side binding (a C++ shim keeping those
 * signatures) is shown in INTEGRATION.md.  Conventions:
 *   - every function returns an xg_status; on failure xg_last_error(ctx)
 *     holds the message (no exception crosses this boundary);
 *   - node ids are 0-based local ids, passed as size_t like the reference's
 *     std::vector<std::size_t> (they must be < 2^31 on the device);
 *   - nodal fields are unknown-major AoS, U(i,c) = U[i*ncomp + c], exactly
 *     tk::Fields (Data.hpp:420-426);
 *   - entry points ending in _dev take device pointers (HBM-resident state);
 *     the others take host pointers and copy, like the reference's
 *     host-resident tk::Fields;
 *   - all device work is issued on the context's stream (xg_set_stream).
 */
#ifndef XG_H
#define XG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  XG_OK = 0,
  XG_EINVAL = 1,        /* bad argument: null pointer, size mismatch, range */
  XG_ENOMESH = 2,       /* no mesh uploaded / setup not done */
  XG_EHIP = 3,          /* a HIP runtime call failed */
  XG_EUNSUPPORTED = 4,  /* configuration not supported by the device path */
  XG_EBLOWUP = 5        /* time step below machine epsilon (RieCG.cpp:861) */
} xg_status;

typedef struct xg_ctx xg_ctx;

/* Physics configuration the reference reads from g_cfg. */
typedef struct {
  double gamma;       /* mat_spec_heat_ratio  (EOS.hpp:30-44) */
  int    flux;        /* 0 rusanov, 1 hllc    (Riemann.cpp:674-684) */
  int    stab2;       /* artificial viscosity (Riemann.cpp:449-458) */
  double stab2coef;
} xg_config;

/* ---- context ------------------------------------------------------------ */
int xg_ctx_create(int device, xg_ctx** ctx);
int xg_ctx_destroy(xg_ctx* ctx);
const char* xg_last_error(const xg_ctx* ctx);
/* Stream for all device work of this context (a hipStream_t; NULL = null
 * stream).  One context per partition/chare, as the reference runs one
 * entry method per chare at a time. */
int xg_set_stream(xg_ctx* ctx, void* hip_stream);
int xg_set_config(xg_ctx* ctx, const xg_config* cfg);

/* ---- host preprocessing (native C++, no device work) --------------------- */

/* Discretization::vol (Discretization.cpp:619-647): nodal volumes.
 * Returns XG_EINVAL on a non-positive element Jacobian. */
int xg_mesh_vol(int64_t npoin, const double* x, const double* y,
                const double* z, int64_t ntet, const size_t* inpoel,
                double* vol);

/* RieCG::domint (RieCG.cpp:339-382): unique edges p<q in (p,q) order and
 * their dual-face integrals D accumulated in tet order.  *nedge receives the
 * edge count; with cap = 0 only the count is computed, otherwise edge[2*cap]
 * and dint[3*cap] are filled (XG_EINVAL if cap < count). */
int xg_mesh_domint(int64_t npoin, const double* x, const double* y,
                   const double* z, int64_t ntet, const size_t* inpoel,
                   int64_t cap, int64_t* nedge, size_t* edge, double* dint);

/* Locality renumbering of the nodes, the job of the RieCG/ChoCG constructor
 * renumbering (RieCG.cpp:82-100) and tk::renumber (Reorder.cpp:215-260):
 * reverse Cuthill-McKee over the edge graph.  perm[old] = new. */
int xg_mesh_rcm(int64_t npoin, int64_t ntet, const size_t* inpoel,
                size_t* perm);

/* Block ordering for the device layout: recursive coordinate bisection of
 * the nodes into spatially compact cells of at most cell_nodes nodes, the
 * cells numbered consecutively and the nodes of a cell kept in their current
 * relative order (so an RCM order within each cell survives).  Renumbering
 * the mesh with perm[old] = new makes every rhs block a contiguous, compact
 * node range -- the layout the rhs kernel streams best (any numbering stays
 * correct). */
int xg_mesh_block_order(int64_t npoin, const double* x, const double* y,
                        const double* z, int64_t cell_nodes, size_t* perm);

/* Symmetry BC data of RieCG::bndint/bnorm/streamable (RieCG.cpp:281-337,
 * 483-618): boundary triangles tri[3*ntri] (outward) with side set ids
 * faceset[ntri]; symmetry side sets symsets[nsym].  Produces one
 * (node, unit normal) pair per node and symmetry set -- inverse-distance^2
 * weighted face normals -- in node order, sets in the order given, and the
 * besym flag of every boundary-face node (besym[3*ntri]).  cap = 0 queries
 * the pair count in *nout; otherwise nodes[cap], norms[3*cap] are filled. */
int xg_mesh_symbc(int64_t npoin, const double* x, const double* y,
                  const double* z, int64_t ntri, const size_t* tri,
                  const int32_t* faceset, int32_t nsym, const int32_t* symsets,
                  int64_t cap, int64_t* nout, size_t* nodes, double* norms,
                  uint8_t* besym);

/* ---- mesh on the device --------------------------------------------------- */

/* The mesh arguments of riemann::grad/rhs (Riemann.hpp:19-39): coordinates,
 * the three superedge groups (dsupedge[k], dsupint[k]: nsup[k] tets (4 ids,
 * 6x3 D), triangles (3 ids, 3x3 D), edges (2 ids, 3 D)), boundary triangles
 * and their per-node symmetry flags besym.  The device layout (edge list,
 * deterministic scatter order, boundary gather lists) is built here once. */
int xg_mesh_upload(xg_ctx* ctx, int64_t npoin,
                   const double* x, const double* y, const double* z,
                   const int64_t nsup[3],
                   const size_t* dsupedge0, const double* dsupint0,
                   const size_t* dsupedge1, const double* dsupint1,
                   const size_t* dsupedge2, const double* dsupint2,
                   int64_t ntri, const size_t* triinpoel,
                   const uint8_t* besym);

/* ---- riemann:: drop-ins ---------------------------------------------------- */

/* riemann::grad (Riemann.hpp:19-25, Riemann.cpp:229-367): G(npoin, 3*ncomp)
 * = undivided edge-based gradients of the primitive variables.  ncomp = 5. */
int xg_riemann_grad(xg_ctx* ctx, int ncomp, const double* U, double* G);
int xg_riemann_grad_dev(xg_ctx* ctx, int ncomp, const double* U, double* G);

/* riemann::rhs (Riemann.hpp:28-39, Riemann.cpp:909-946): R(npoin, ncomp)
 * from G (already divided by the nodal volume) and U.  No source term
 * (problems::SRC() is null for the problems this path serves). */
int xg_riemann_rhs(xg_ctx* ctx, int ncomp, const double* G, const double* U,
                   double* R);
int xg_riemann_rhs_dev(xg_ctx* ctx, int ncomp, const double* G,
                       const double* U, double* R);

/* ---- resident RieCG stepping ------------------------------------------------ */

/* Per-mesh data of the RieCG chare: nodal volumes Vol (RieCG.cpp:935,1017)
 * and the symmetry BC lists symbcnodes/symbcnorms (RieCG.cpp:574-589). */
int xg_riecg_setup(xg_ctx* ctx, const double* vol, int64_t nsym,
                   const size_t* symnodes, const double* symnorms);

/* RieCG::dt (RieCG.cpp:787-852), unsteady branch: *dt = cfl * min_p
 * cbrt(vol_p)/max(|u|+c, 1e-8).  U on the device. */
int xg_riecg_dt_dev(xg_ctx* ctx, const double* U, double cfl, double* dt);

/* One RK stage (RieCG::grad/rhs/solve, RieCG.cpp:871-1057): G = grad(U)/vol,
 * R = rhs(G,U); at stage 0 Un = U; U = Un - rkcoef[stage]*dt*R/vol with
 * rkcoef {1/3,1/2,1}; symmetry BCs.  All arrays on the device:
 * U, Un (npoin*5), G (npoin*15), R (npoin*5). */
int xg_riecg_stage_dev(xg_ctx* ctx, int stage, double dt, double* U,
                       double* Un, double* G, double* R);

/* Number of bytes the layout occupies on the device (for reporting). */
int xg_mesh_device_bytes(const xg_ctx* ctx, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
