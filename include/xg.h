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
  XG_EBLOWUP = 5,       /* time step below machine epsilon (RieCG.cpp:861) */
  XG_EIO = 6            /* a file could not be read or written (xg_exo_*) */
} xg_status;

typedef struct xg_ctx xg_ctx;

/* Physics configuration the reference reads from g_cfg. */
typedef struct {
  double gamma;       /* mat_spec_heat_ratio  (EOS.hpp:30-44) */
  int    flux;        /* 0 rusanov, 1 hllc    (Riemann.cpp:674-684) */
  int    stab2;       /* artificial viscosity (Riemann.cpp:449-458) */
  double stab2coef;
} xg_config;

/* The analytic problem whose source term the device adds (problems::SRC(),
 * Problems.cpp:1299-1325; g_cfg problem, problem_alpha/kappa/p0).  Problems
 * without a source (sod, sedov, userdef) are XG_PROBLEM_NONE.  The sources
 * served are independent of time. */
enum {
  XG_PROBLEM_NONE = 0,
  XG_PROBLEM_VORTICAL_FLOW = 1,   /* Problems.cpp:452-509 */
  XG_PROBLEM_TAYLOR_GREEN = 2,    /* Problems.cpp:407-450 */
  XG_PROBLEM_SLOT_CYL = 3         /* Problems.cpp:511-672: momentum source of
                                     the rotation carrying one scalar */
};
typedef struct {
  int    kind;
  double alpha, kappa, p0;        /* vortical_flow parameters */
} xg_problem;

/* ---- context ------------------------------------------------------------ */
int xg_ctx_create(int device, xg_ctx** ctx);
int xg_ctx_destroy(xg_ctx* ctx);
const char* xg_last_error(const xg_ctx* ctx);
/* Stream for all device work of this context (a hipStream_t; NULL = null
 * stream).  One context per partition/chare, as the reference runs one
 * entry method per chare at a time. */
int xg_set_stream(xg_ctx* ctx, void* hip_stream);
int xg_set_config(xg_ctx* ctx, const xg_config* cfg);
/* The problem of riemann::rhs's and kozak::rhs's source term (default
 * XG_PROBLEM_NONE); XG_EUNSUPPORTED for a kind the device does not have. */
int xg_set_problem(xg_ctx* ctx, const xg_problem* problem);
/* The problem set by xg_set_problem (a caller holding a source term checks
 * that the device adds the same one, as problems::SRC() is queried in
 * Riemann.cpp:895 and Kozak.cpp:62). */
int xg_get_problem(const xg_ctx* ctx, xg_problem* problem);

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

/* ChoCG::domint (ChoCG.cpp:399-445): as xg_mesh_domint with 5 values per
 * edge (dint[5*cap]): the dual-face integral, J/120 and the P1 stiffness
 * grad_p.grad_q/J/6, each accumulated in tet order. */
int xg_mesh_chocg_domint(int64_t npoin, const double* x, const double* y,
                         const double* z, int64_t ntet, const size_t* inpoel,
                         int64_t cap, int64_t* nedge, size_t* edge, double* dint);

/* RieCG::domsuped (RieCG.cpp:620-736) and ChoCG::domsuped (ChoCG.cpp:
 * 687-813): the unique edges of xg_mesh_domint / xg_mesh_chocg_domint
 * (edge[2*nedge] sorted by (p, q), p < q; stride 3 or 5 integral values per
 * edge) grouped into tet superedges (a tet whose six edges are all still
 * free: 4 ids, 6 x stride values), then triangle superedges over the tets'
 * faces in first-seen order (3 ids, 3 x stride), then the remaining edges
 * (2 ids, stride), each edge oriented as its superedge gives it (the first
 * three values change sign when that is q -> p).  The orientation matters
 * to the last bits: the MUSCL limiter's epsilon makes the edge flux not
 * exactly antisymmetric.  With sup0 = NULL only nsup[3] is computed;
 * otherwise the six arrays are filled (sized from nsup by the caller). */
int xg_mesh_domsuped(int64_t ntet, const size_t* inpoel, int64_t nedge,
                     const size_t* edge, const double* dint, int stride,
                     int64_t nsup[3], size_t* sup0, double* int0, size_t* sup1,
                     double* int1, size_t* sup2, double* int2);

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

/* The Neumann part of the pressure rhs of ChoCG::pinit (ChoCG.cpp:1070-1095):
 * neubc (npoin) zeroed, then for every boundary triangle e with besym[e]
 * (the faces of the pressure block's bc_sym side sets), n = (x1 - x0) x
 * (x2 - x0) / 6 and neubc[N_k] -= n . pg[N_k] for its three nodes, pg
 * (npoin*3) the problem's pressure gradient (PRESSURE_GRAD) at the nodes.
 * tri as in xg_mesh_symbc (outward); face order as given. */
int xg_mesh_neumann(int64_t npoin, const double* x, const double* y,
                    const double* z, int64_t ntri, const size_t* tri,
                    const uint8_t* besym, const double* pg, double* neubc);

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
 * = undivided edge-based gradients of the primitive variables (the
 * transported scalars as they are).  ncomp = 5 + scalars, 5 <= ncomp <=
 * XG_MAX_NCOMP; XG_EUNSUPPORTED beyond. */
#define XG_MAX_NCOMP 7
int xg_riemann_grad(xg_ctx* ctx, int ncomp, const double* U, double* G);
int xg_riemann_grad_dev(xg_ctx* ctx, int ncomp, const double* U, double* G);

/* riemann::rhs (Riemann.hpp:28-39, Riemann.cpp:909-946): R(npoin, ncomp)
 * = advdom + advbnd from G (already divided by the nodal volume) and U, then
 * the source term of the context's problem (riemann::src, Riemann.cpp:
 * 880-907): R(p,c) -= s_c(x_p, t) v[p], with v the nodal volumes of this
 * partition's own elements (may be null when the problem has no source) and
 * tp the per-node time of steady runs (null: t; the sources served do not
 * depend on time, so tp changes nothing).  Transported scalars (ncomp > 5)
 * get the MUSCL reconstruction and scalar fluxes of Riemann.cpp:145-209,
 * 455-466, 635-643 and no source.  XG_EUNSUPPORTED for ncomp outside
 * 5..XG_MAX_NCOMP. */
int xg_riemann_rhs(xg_ctx* ctx, int ncomp, const double* G, const double* U,
                   const double* v, double t, const double* tp, double* R);
int xg_riemann_rhs_dev(xg_ctx* ctx, int ncomp, const double* G,
                       const double* U, const double* v, double t,
                       const double* tp, double* R);

/* ---- resident RieCG stepping ------------------------------------------------ */

/* Per-mesh data of the RieCG chare: nodal volumes Vol (RieCG.cpp:935,1017)
 * and the symmetry BC lists symbcnodes/symbcnorms (RieCG.cpp:574-589). */
int xg_riecg_setup(xg_ctx* ctx, const double* vol, int64_t nsym,
                   const size_t* symnodes, const double* symnorms);

/* Components per node of the RieCG chare's fields (U.nprop(): 5 + the
 * problem's transported scalars, at most XG_MAX_NCOMP); default 5.  The
 * stepping entries below read and write U, Un, R with ncomp and G with
 * 3 ncomp components per node.  Call after xg_riecg_setup and before
 * xg_riecg_bc (it drops the Dirichlet lists, whose masks are per
 * component). */
int xg_riecg_set_ncomp(xg_ctx* ctx, int ncomp);

/* The other boundary conditions of the RieCG chare (setupBC and
 * streamable, RieCG.cpp:109-245, 527-618; applied by RieCG::BC,
 * RieCG.cpp:764-785, after the symmetry BCs of xg_riecg_setup in the order
 * dirbc, symbc, farbc, prebc):
 *   Dirichlet (physics::dirbc, BC.cpp:29-72): ndir nodes with ncomp mask
 *     entries each (xg_riecg_set_ncomp); a component with mask 1 takes
 *     dirvals[k*ncomp+c], the initial condition at the node, which the
 *     caller evaluates once: the values stay fixed in time (exact for the
 *     problems served: their flow fields do not depend on time, and
 *     slot_cyl's scalar changes only inside shapes that stay at distance
 *     <= 0.4 from the rotation axis, off the unit square's boundary);
 *     other masks leave it (the RieCG chare passes no BC values);
 *   farfield (physics::farbc, BC.cpp:152-221): nfar (node, unit normal)
 *     pairs and the free stream far = {density, u, v, w, pressure};
 *   pressure (physics::prebc, BC.cpp:223-241): npre nodes with (density,
 *     pressure) each.
 * Call after xg_riecg_setup; replaces any earlier lists. */
int xg_riecg_bc(xg_ctx* ctx, int64_t ndir, const size_t* dirnodes,
                const int32_t* dirmask, const double* dirvals, int64_t nfar,
                const size_t* farnodes, const double* farnorms,
                const double far[5], int64_t npre, const size_t* prenodes,
                const double* prevals);

/* RieCG::BC (RieCG.cpp:764-785) on device-resident U: dirbc, symbc, farbc,
 * prebc.  The chare applies it to the initial state too (RieCG.cpp:754). */
int xg_riecg_apply_bc_dev(xg_ctx* ctx, double* U);

/* RieCG::dt (RieCG.cpp:787-852), unsteady branch: *dt = cfl * min_p
 * cbrt(vol_p)/max(|u|+c, 1e-8).  U on the device. */
int xg_riecg_dt_dev(xg_ctx* ctx, const double* U, double cfl, double* dt);

/* RieCG::dt, steady branch (local time stepping, RieCG.cpp:810-824):
 * dtp[p] = cfl cbrt(vol_p)/max(|u|+c, 1e-8) (device, npoin) and *dt its
 * minimum. */
int xg_riecg_dt_steady_dev(xg_ctx* ctx, const double* U, double cfl,
                           double* dtp, double* dt);

/* One RK stage (RieCG::grad/rhs/solve, RieCG.cpp:871-1057): G = grad(U)/vol,
 * R = rhs(G,U) with the problem's source at time t (steady: at tp +
 * prev_rkcoef dtp); at stage 0 Un = U; U = Un - rkcoef[stage]*dt*R/vol with
 * rkcoef {1/3,1/2,1} (steady: dtp[i] for dt); then RieCG::BC.  dtp, tp
 * (device, npoin) null for time-accurate runs; with them, stage 2 also
 * advances tp += dtp (RieCG.cpp:1047-1050).  All arrays on the device:
 * U, Un (npoin*ncomp), G (npoin*3*ncomp), R (npoin*ncomp). */
int xg_riecg_stage_dev(xg_ctx* ctx, int stage, double t, double dt,
                       const double* dtp, double* tp, double* U, double* Un,
                       double* G, double* R);

/* The pieces of a stage for partitioned runs, where the partial sums of the
 * partitions are added between them (xg_comm_sum_dev):
 * xg_riecg_grad_div_dev divides G(npoin, 3 ncomp) by the summed nodal volume of
 * xg_riecg_setup (RieCG.cpp:935-939); xg_riecg_solve_dev is RieCG::solve
 * (RieCG.cpp:992-1057): Un = U at stage 0, the RK update with R, the BCs
 * at t + rkcoef dt, and tp += dtp after stage 2 when steady. */
int xg_riecg_grad_div_dev(xg_ctx* ctx, double* G);
int xg_riecg_solve_dev(xg_ctx* ctx, int stage, double t, double dt,
                       const double* dtp, double* tp, double* U, double* Un,
                       const double* R);

/* Number of bytes the layout occupies on the device (for reporting). */
int xg_mesh_device_bytes(const xg_ctx* ctx, int64_t* bytes);

/* ---- kozak:: drop-in (KozCG) ------------------------------------------------- */

/* The mesh arguments of kozak::rhs (Kozak.hpp:19-27): tet connectivity
 * inpoel (4 ids per tet) and coordinates.  The device layout (node blocks,
 * tets by block, deterministic scatter order) is built here once. */
int xg_kozak_upload(xg_ctx* ctx, int64_t npoin, const double* x,
                    const double* y, const double* z, int64_t ntet,
                    const size_t* inpoel);

/* kozak::rhs (Kozak.hpp:19-27, Kozak.cpp:29-180): R(npoin, ncomp) zeroed
 * and filled with the Taylor-Galerkin rhs of U for time step dt at time t,
 * with the source term of the context's problem (Kozak.cpp:124-136,
 * 163-173).  ncomp = 5 + transported scalars (advected with the flow
 * velocity, no source), at most XG_MAX_NCOMP.  Steady runs pass dtp
 * (npoin; device memory for _dev): each tet then uses the mean of its
 * corners' dtp (Kozak.cpp:105); tp is accepted and unused (the sources
 * served do not depend on time).  Null dtp: time-accurate with dt. */
int xg_kozak_rhs(xg_ctx* ctx, int ncomp, double t, double dt, const double* tp,
                 const double* dtp, const double* U, double* R);
int xg_kozak_rhs_dev(xg_ctx* ctx, int ncomp, double t, double dt,
                     const double* tp, const double* dtp, const double* U,
                     double* R);

/* ---- resident KozCG stepping ------------------------------------------------ */

/* KozCG's flux-corrected transport settings (InciterConfig.cpp:1751-1754). */
typedef struct {
  int    fct;         /* FCT on (default) or off */
  double fctdif;      /* antidiffusion coefficient, default 1.0 */
  int    fctclip;     /* allowed bounds from the low-order solution only */
  int    nsys;        /* number of fctsys components */
  int    sys[5];      /* fctsys, 0-based component ids limited together */
} xg_fctcfg;

/* Components per node of the KozCG chare's fields (5 + transported
 * scalars, at most XG_MAX_NCOMP; default 5).  Call before xg_kozcg_setup,
 * whose Dirichlet lists and buffers hold ncomp values per node (a change
 * drops an earlier setup). */
int xg_kozcg_set_ncomp(xg_ctx* ctx, int ncomp);

/* Per-mesh data of the KozCG chare on the mesh of xg_kozak_upload: nodal
 * volumes, the symmetry BC lists symbcnodes/symbcnorms (KozCG.cpp:524-540),
 * the Dirichlet nodes with a component mask and the values physics::dirbc
 * sets (KozCG.cpp:120-143, BC.cpp:30-69; values fixed in time, ndir*ncomp
 * each), and the FCT settings (fctsys ids below ncomp). */
int xg_kozcg_setup(xg_ctx* ctx, const double* vol, int64_t nsym,
                   const size_t* symnodes, const double* symnorms, int64_t ndir,
                   const size_t* dirnodes, const int* dirmask,
                   const double* dirvals, const xg_fctcfg* fct);

/* KozCG::dt (KozCG.cpp:614-700), unsteady branch:
 * *dt = cfl * min_p cbrt(vol_p)/max(|u|+c, 1e-8).  U on the device.  The
 * chare multiplies it by freezeflow once its time exceeds freezetime
 * (KozCG.cpp:669-674). */
int xg_kozcg_dt_dev(xg_ctx* ctx, const double* U, double cfl, double* dt);

/* KozCG::dt, steady branch (local time stepping, KozCG.cpp:636-649):
 * dtp[p] = cfl cbrt(vol_p)/max(|u|+c, 1e-8) (device, npoin) and *dt its
 * minimum. */
int xg_kozcg_dt_steady_dev(xg_ctx* ctx, const double* U, double cfl, double* dtp,
                           double* dt);

/* One KozCG time step (KozCG::rhs/aec/alw/lim/solve, KozCG.cpp:709-1197):
 * the Taylor-Galerkin rhs, with FCT the antidiffusive contributions (with
 * symmetry BCs on them), the low-order solution, allowed bounds, limit
 * coefficients and limited contributions, the new solution; without FCT
 * U + dt R/vol.  Then Dirichlet and symmetry BCs (KozCG::BC,
 * KozCG.cpp:590-611).  t is the time at the start of the step (the source
 * term's).  U (npoin*ncomp, device) is advanced in place; Un
 * (npoin*ncomp, device) receives the old solution. */
int xg_kozcg_step_dev(xg_ctx* ctx, double t, double dt, double* U, double* Un);

/* xg_kozcg_step_dev with local time stepping and freezeflow: dtp, tp
 * (device, npoin; both null for time-accurate runs) make kozak::rhs use
 * each tet's mean dtp, the low-order solution (or, without FCT, the update)
 * each node's own dtp (KozCG.cpp:889-890, 1155), and advance tp += dtp after
 * the step (:1189-1193); freeze (the chare's freezeflow > 1, KozCG.cpp:1141)
 * gives the flow components back their values of the step's start after
 * the BCs (:1168-1175), so only the scalars advance. */
int xg_kozcg_step2_dev(xg_ctx* ctx, double t, double dt, const double* dtp,
                       double* tp, int freeze, double* U, double* Un);

/* ---- chorin:: drop-ins (ChoCG) ------------------------------------------------ */

/* ChoCG's g_cfg reads of chorin::rhs (InciterConfig.cpp:1724-1757). */
typedef struct {
  int    flux;        /* 0 "damp2", 1 "damp4" (Chorin.cpp:862-866) */
  int    stab;        /* stab: |vn_L + vn_R|/2 */
  int    stab2;
  double stab2coef;
  double mu;          /* mat_dyn_viscosity */
  double diff;        /* mat_dyn_diffusivity: the transported scalars' */
} xg_chocfg;

/* The mesh arguments of the chorin:: operators (Chorin.hpp:19-70): the
 * superedges of ChoCG::domsuped with 5 values per edge integral
 * (ChoCG.cpp:398-445, 687-813: normal components, J/120,
 * dot(grad_p,grad_q)/J/6), coordinates and boundary triangles. */
int xg_chocg_upload(xg_ctx* ctx, int64_t npoin, const double* x,
                    const double* y, const double* z, const int64_t nsup[3],
                    const size_t* dsupedge0, const double* dsupint0,
                    const size_t* dsupedge1, const double* dsupint1,
                    const size_t* dsupedge2, const double* dsupint2,
                    int64_t ntri, const size_t* triinpoel);

/* In the chorin:: entry points `dev` = 1 means every field pointer is a
 * device pointer (enqueued on the context stream), 0 host pointers (copied
 * through device staging, synchronous). */

/* chorin::div (Chorin.hpp, Chorin.cpp:82-209): D += div of U (ldu columns,
 * 0..2 used); with stab the 4th-order pressure damping from P and its
 * gradient G (ldg columns, 0..2 used) with time step dt. */
int xg_chorin_div(xg_ctx* ctx, int dev, double dt, const double* P,
                  const double* G, int ldg, const double* U, int ldu,
                  double* D, int stab);

/* chorin::grad (Chorin.cpp:336-449): G(:, 0..2) += gradient of scalar U. */
int xg_chorin_grad(xg_ctx* ctx, int dev, const double* U, double* G, int ldg);

/* chorin::vgrad (Chorin.cpp:211-334): G(npoin, 3 ncomp) += gradients of the
 * velocity and transported scalars U(npoin, ncomp); ncomp 3 or 4 (one
 * transported scalar, the count of the ChoCG problems, Problems.cpp:531-545). */
int xg_chorin_vgrad(xg_ctx* ctx, int dev, int ncomp, const double* U, double* G);

/* chorin::flux (Chorin.cpp:451-638): F(npoin, 3) += momentum flux
 * divergence from U (ldu columns) and velocity gradients G (ldg >= 9). */
int xg_chorin_flux(xg_ctx* ctx, int dev, double mu, const double* U, int ldu,
                   const double* G, int ldg, double* F);

/* chorin::rhs (Chorin.cpp:1010-1044): R(npoin, ncomp) = advection rhs
 * (damp2 / damp4) of U (velocity, then the transported scalar: its flux
 * from the same normal velocities, diffusivity cfg->diff) with pressure P
 * and gradients G (npoin, 3 ncomp, damp4 only); ncomp 3 or 4.  Without the
 * source term (the ChoCG step adds it, below). */
int xg_chorin_rhs(xg_ctx* ctx, int dev, const xg_chocfg* cfg, int ncomp,
                  const double* P, const double* U, const double* G, double* R);

/* ---- resident ChoCG stepping ------------------------------------------------- */

/* ChoCG's g_cfg reads of its time step (InciterConfig.cpp): chorin::rhs
 * settings, cfl, rk stages (1-4, rkcoef of ChoCG.cpp:43-48) and the pressure
 * solve's pc (0 none, 1 jacobi; 2, without a reference counterpart: flexible
 * CG with the multigrid preconditioner of xg_cg_solve_flexible pc_kind 3),
 * iter and tol. */
typedef struct {
  xg_chocfg flux;
  double    cfl;
  int       rk;
  int       p_pc;
  int64_t   p_maxit;
  double    p_tol;
} xg_chostep;

/* The number of components of U for the ChoCG stepping of this context:
 * 3 (velocity) or 4 (velocity and one transported scalar, ChoCG's
 * Problems.cpp:531-545 count).  After xg_chocg_upload, before
 * xg_chocg_setup (a change drops the setup; default 3). */
int xg_chocg_set_ncomp(xg_ctx* ctx, int ncomp);

/* Per-mesh data of the ChoCG chare on the mesh of xg_chocg_upload, whose
 * pressure matrix the caller has given to xg_cg_matrix (ChoCG::prelhs):
 * nodal volumes; symmetry (node, normal) pairs and noslip nodes
 * (ChoCG.cpp:637-686); Dirichlet nodes with a per-component mask (ncomp
 * of xg_chocg_set_ncomp per node) and the values physics::dirbc sets
 * (setupDirBC, ChoCG.cpp:211-301); pressure
 * Dirichlet nodes (bc_dir of the pressure block and the hydrostat node,
 * ChoCG.cpp:1051-1112) with their initial-projection values (the pressure
 * increments of the time steps use 0). */
int xg_chocg_setup(xg_ctx* ctx, const double* vol, const xg_chostep* cfg,
                   int64_t nsym, const size_t* symnodes, const double* symnorms,
                   int64_t nnoslip, const size_t* noslipnodes, int64_t ndir,
                   const size_t* dirnodes, const int* dirmask,
                   const double* dirvals, int64_t npdir, const size_t* pdirnodes,
                   const double* pdirvals);

/* The ChoCG start-up (merge ... finpgrad, ChoCG.cpp:817-1323): BCs, the
 * divergence-free projection of U, the momentum-flux pressure solve that
 * sets P, and the pressure gradient.  U (npoin*ncomp), P (npoin) on the
 * device; pit[0..1] receive the two solves' iteration counts. */
int xg_chocg_initial_dev(xg_ctx* ctx, double* U, double* P, int64_t* pit);

/* ChoCG's run of the first pressure solve alone (nstep = 1 with a problem
 * that has a pressure rhs, the Poisson tests; ChoCG.cpp:1026-1124,
 * 1195-1225): BCs on U, b = div(U) -- replaced by pr * vol where pr (npoin,
 * PRESSURE_RHS at the nodes) is given -- plus the Neumann vector neubc
 * (npoin, from xg_mesh_neumann, or null), the CG solve with the pressure
 * Dirichlet nodes and values of xg_chocg_setup (mask 1: PRESSURE_IC at the
 * node), the projection U -= grad(p) with its BCs, and P = the solution.
 * pr, neubc, U (npoin*ncomp), P (npoin) on the device; *pit the
 * iterations. */
int xg_chocg_poisson_dev(xg_ctx* ctx, const double* pr, const double* neubc, double* U,
                         double* P, int64_t* pit);

/* ChoCG::dt (ChoCG.cpp:1357-1413) without a constant dt: cfl * min over
 * nodes of cbrt(vol)/max(|u|, 1e-8) and cbrt(vol)^2/max(mu, diff); the
 * caller multiplies by freezeflow once t > freezetime (:1396-1399). */
int xg_chocg_dt_dev(xg_ctx* ctx, const double* U, double* dt);

/* The theta-implicit momentum system of ChoCG (momlhs/lhs, ChoCG.cpp:
 * 191-209, 1434-1477): the 3-component CSR of the tets (inpoel, the mesh of
 * xg_chocg_upload) with per-entry mass and stiffness sums, and the momentum
 * solve's settings (pc 0 none / 1 jacobi, iter, tol).  With theta > 0 the
 * last RK stage of xg_chocg_step_dev solves A du = R, A = -(M/dt + theta mu
 * K), du = 0 on the dirbc-masked components and noslip nodes, by CG from the
 * previous du, and sets U = Un + du (ChoCG.cpp:1570-1646).  Call after
 * xg_chocg_setup; 3 components only (XG_EUNSUPPORTED with a transported
 * scalar). */
int xg_chocg_momentum_setup(xg_ctx* ctx, int64_t ntet, const size_t* inpoel,
                            double theta, int m_pc, int64_t m_maxit, double m_tol);
/* Iterations of the last momentum solve. */
int xg_chocg_momentum_its(xg_ctx* ctx, int64_t* its);

/* One ChoCG time step (rhs/solve/pred/corr for every RK stage, then
 * div/pinit/psolve/sgrad/psolved/pgrad, ChoCG.cpp:869-1695), the momentum
 * stages explicit or, after xg_chocg_momentum_setup with theta > 0, the
 * last one semi-implicit.  The rhs includes the source of the context's
 * problem (xg_set_problem: slot_cyl, with ncomp 4; the other device
 * problems have no ChoCG source and are refused).  U (npoin*ncomp) advanced
 * in place, Un receives the step's starting state, P += dP with dP (npoin)
 * the pressure increment; *pit the pressure solve's iterations. */
int xg_chocg_step_dev(xg_ctx* ctx, double dt, double* U, double* Un, double* P,
                      double* dP, int64_t* pit);
/* xg_chocg_step_dev with the flow frozen when freeze != 0 (the chare's
 * freezeflow > 1, ChoCG.cpp:1550-1568): every explicit stage puts the
 * velocity of its start back after its update, BCs and (damp4) gradients,
 * so only the transported scalar advances until the projection; with damp2
 * the last stage's divergence is taken before the velocity is put back,
 * as the reference's pred/corr order does. */
int xg_chocg_step2_dev(xg_ctx* ctx, double dt, int freeze, double* U, double* Un,
                       double* P, double* dP, int64_t* pit);

/* ---- node diagnostics -------------------------------------------------------- */

/* NodeDiagnostics::rhocompute (NodeDiagnostics.cpp:46-146) without the
 * analytic-solution error norms: out[c] = sum_i U(i,c)^2 V_i, out[ncomp+c] =
 * sum_i (U(i,c) - Un(i,c))^2 V_i, out[2 ncomp] = sum_i U(i,4) V_i (0 for
 * ncomp < 5).  U, Un (npoin*ncomp), V (npoin) on the device; out (host)
 * 2 ncomp + 1 values.  ncomp 1, 3 or 5..XG_MAX_NCOMP.  Deterministic
 * summation order. */
int xg_diag_rho_dev(xg_ctx* ctx, int64_t npoin, int ncomp, const double* U,
                    const double* Un, const double* V, double* out);

/* xg_diag_rho_dev for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops); it has no counterpart in
 * the reference.  Same sums in the same order as xg_diag_rho_dev, bit for
 * bit, but out (2 ncomp + 1 doubles) is a DEVICE pointer: the call only
 * enqueues two kernels on the context's stream, copies nothing to the host
 * and does not synchronize.  It allocates nothing: the workgroup partials
 * live in a fixed device array of 11 x 1024 doubles per GPU, held until the
 * process ends; larger sizes are refused (npoin never needs more, ncomp is
 * at most 5).  Calls that overlap on one GPU (two streams or threads) share
 * that array and race: serialize them. */
int xg_diag_rho_dev_noalloc(xg_ctx* ctx, int64_t npoin, int ncomp, const double* U,
                            const double* Un, const double* V, double* out);

/* The analytic-error sums of NodeDiagnostics::rhocompute
 * (NodeDiagnostics.cpp:79-128): with An (npoin*ncomp, device) the analytic
 * solution in primitive form (rho, u, v, w, specific internal energy, then
 * the scalars; the caller evaluates SOL()), out[c] = sum_i (nu_c - An_c)^2
 * V_i and out[ncomp+c] = sum_i |nu_c - An_c| V_i, nu the same form of U.
 * 5 <= ncomp <= XG_MAX_NCOMP; or the error sums of
 * NodeDiagnostics::precompute (NodeDiagnostics.cpp:186-250) without
 * conversion: ncomp = 1 with U the pressure and An PRESSURE_SOL() at the
 * nodes, ncomp 3 or 4 with U the velocity (and transported scalar) and An
 * SOL(); out (host) 2 ncomp values; deterministic summation order. */
int xg_diag_err_dev(xg_ctx* ctx, int64_t npoin, int ncomp, const double* U,
                    const double* V, const double* An, double* out);

/* NodeDiagnostics::precompute (NodeDiagnostics.cpp:148-264) without the
 * analytic-solution error norms: out = {sum p^2 V, sum u_c^2 V (3),
 * sum dp^2 V, sum (u_c - un_c)^2 V (3)}.  U, Un (npoin*3), P, dP, V (npoin)
 * on the device; out (host) 8 values. */
int xg_diag_pre_dev(xg_ctx* ctx, int64_t npoin, const double* U, const double* Un,
                    const double* P, const double* dP, const double* V, double* out);
/* xg_diag_pre_dev for U, Un of ncomp (3 or 4: with a transported scalar)
 * components per node: out = {sum p^2 V, sum u_c^2 V (ncomp), sum dp^2 V,
 * sum (u_c - un_c)^2 V (ncomp)}, 2 ncomp + 2 values. */
int xg_diag_pre_nc_dev(xg_ctx* ctx, int64_t npoin, int ncomp, const double* U,
                       const double* Un, const double* P, const double* dP,
                       const double* V, double* out);

/* ---- partitions on one device ------------------------------------------------
 * The reference runs one chare per mesh partition and adds the partitions'
 * partial nodal sums at shared nodes by messages (Discretization's
 * NodeCommMap; RieCG::comgrad/comrhs, RieCG.cpp:871-916, 948-990;
 * ConjugateGradients::comq, ConjugateGradients.cpp:611-627).  Here the
 * partitions are contexts on one device and the exchange is one kernel. */
typedef struct xg_comm xg_comm;
/* nparts partitions (at most 64); partition k has nlocal[k] nodes whose
 * global ids are gid[k][0 .. nlocal[k]).  Device memory on `device`.  A
 * handle is returned even when this fails (holding the message for
 * xg_comm_error); xg_comm_destroy releases it. */
int xg_comm_create(int device, int nparts, const int64_t* nlocal,
                   const size_t* const* gid, xg_comm** comm);
/* Every node held by more than one partition gets, in every copy, the sum
 * of the copies' values over the partitions in partition order.
 * fields[k]: partition k's device array (nlocal[k] x ncomp, AoS).  Enqueued
 * on `stream` (a hipStream_t, NULL = null stream); the caller orders it after
 * the work that wrote the partials. */
int xg_comm_sum_dev(xg_comm* comm, int ncomp, double* const* fields,
                    void* stream);
/* Number of nodes held by more than one partition. */
int xg_comm_shared(const xg_comm* comm, int64_t* nshared);
const char* xg_comm_error(const xg_comm* comm);
int xg_comm_destroy(xg_comm* comm);

/* The chares of one ConjugateGradients array that share a GPU
 * (ConjugateGradients.hpp:53-152, ConjugateGradients.cpp:105-823): nparts
 * partitions of one linear system with ncomp unknowns per node.  Partition
 * k holds npoin[k] nodes with global ids gid[k] (their overlap is the node
 * communication map), the partial matrix of its own elements in tk::CSR
 * form (ia[k], ja[k], a[k]; 1-based when one_based, as tk::CSR stores them;
 * Dirichlet rows set by CSR::dirichlet with 1/count on the diagonal,
 * CSR.cpp:107-152) and, at setup, b[k] with the full values at shared nodes
 * and x[k] (null: zero).  The exchanges (comres, comd, comq: partial sums
 * completed in partition order; comx: the copies averaged) run on the
 * device; dot products count a node in the highest partition holding it
 * (tk::slave, Reorder.cpp:393-412).  pc: 0 none (1/count per copy, summed),
 * 1 Jacobi.  xg_pcg_setup returns normb, xg_pcg_solve the recurrence's
 * ||r|| with the reference's stop test (||r|| < tol normb, |(p,q)| <
 * DBL_EPSILON, maxit).  A handle is returned even when create fails
 * (message in xg_pcg_error); xg_pcg_destroy releases it. */
typedef struct xg_pcg xg_pcg;
int xg_pcg_create(int device, int nparts, int ncomp, const int64_t* npoin,
                  const size_t* const* gid, const size_t* const* ia,
                  const size_t* const* ja, const double* const* a, int one_based,
                  xg_pcg** pcg);
int xg_pcg_set_stream(xg_pcg* pcg, void* hip_stream);
int xg_pcg_setup(xg_pcg* pcg, const double* const* x, const double* const* b, int pc,
                 double* normb);
int xg_pcg_solve(xg_pcg* pcg, int64_t maxit, double tol, double* normr, int64_t* it,
                 int* converged);
int xg_pcg_solution(xg_pcg* pcg, int part, double* x);
const char* xg_pcg_error(const xg_pcg* pcg);
int xg_pcg_destroy(xg_pcg* pcg);

/* ---- pressure-Poisson system (ChoCG) and conjugate gradients ---------------- */

/* tk::CSR(ncomp, psup) (CSR.hpp:29-31, CSR.cpp:19-84): the sparsity pattern
 * of the node graph of the tets, ncomp unknowns per node interleaved
 * (row i*ncomp+k), columns ascending, diagonal included; 0-based here (the
 * reference stores 1-based ia/ja).  cap = 0 only counts: *nnz. */
int xg_csr_pattern(int64_t npoin, int64_t ntet, const size_t* inpoel,
                   int ncomp, int64_t cap, int64_t* nnz, size_t* ia,
                   size_t* ja);

/* ChoCG::prelhs (ChoCG.cpp:146-188): the Laplacian A(p,q) -= grad_p.grad_q/J
 * summed over tets (J = 6 * tet volume), into the values a of the ncomp = 1
 * pattern above. */
int xg_poisson_matrix(int64_t npoin, const double* x, const double* y,
                      const double* z, int64_t ntet, const size_t* inpoel,
                      const size_t* ia, const size_t* ja, double* a);

/* ConjugateGradients(A, x, b) (ConjugateGradients.hpp:53-60) on one
 * partition: the matrix in CSR (base 0 or 1, ia has npoin*ncomp+1 entries,
 * rows with ascending columns and a structurally symmetric pattern, as
 * tk::CSR builds it) is copied to the device; x = b = 0. */
int xg_cg_matrix(xg_ctx* ctx, int ncomp, int64_t npoin, const size_t* ia,
                 const size_t* ja, const double* a, int base);

/* ConjugateGradients::init (ConjugateGradients.hpp:97-104,
 * ConjugateGradients.cpp:335-556); the caller then calls xg_cg_setup, as the
 * reference's init continues with setup.  x, b (host, npoin*ncomp; null
 * keeps the current ones); with apply: b += neubc (null: none), and for the
 * ndir nodes dirnode[k] with dirmask[k*ncomp+j] set, Dirichlet value
 * dirval[k*ncomp+j] applied to a copy of A as CSR::dirichlet does, node by
 * node in decreasing node order.  pc: 0 "none", 1 "jacobi". */
int xg_cg_init(xg_ctx* ctx, const double* x, const double* b,
               const double* neubc, int64_t ndir, const size_t* dirnode,
               const int32_t* dirmask, const double* dirval, int apply, int pc);

/* xg_cg_init with the rhs b on the device (npoin*ncomp) and x kept from the
 * previous solve (the warm start of ConjugateGradients::init with an empty
 * x, as ChoCG's pressure solves use it, ChoCG.cpp:1122-1124); no Neumann
 * term.  Same Dirichlet semantics as xg_cg_init. */
int xg_cg_init_dev(xg_ctx* ctx, const double* b, int64_t ndir,
                   const size_t* dirnode, const int32_t* dirmask,
                   const double* dirval, int apply, int pc);

/* ConjugateGradients::setup (ConjugateGradients.cpp:105-200): r = b - A x,
 * preconditioner, normb = ||b||, rho = (r,z). */
int xg_cg_setup(xg_ctx* ctx, double* normb);

/* ConjugateGradients::solve (ConjugateGradients.hpp:90-94,
 * ConjugateGradients.cpp:559-823): iterate until sqrt((r,r)) < tol*normb
 * (normb replaced by 1 if <= 1e-14), |(p,q)| < DBL_EPSILON, or maxit;
 * *normr = sqrt((r,r)) of the recurrence, *it iterations. */
int xg_cg_solve(xg_ctx* ctx, int64_t maxit, double tol, double* normr,
                int64_t* it, int* converged);

/* ConjugateGradients::solution (ConjugateGradients.hpp:152) to the host, or
 * the device address of x (valid until xg_cg_matrix/xg_cg_destroy). */
int xg_cg_solution(xg_ctx* ctx, double* x);
int xg_cg_device_x(xg_ctx* ctx, double** x);

/* Mixed precision with reliable updates (no counterpart in the reference):
 * the Jacobi-PCG recurrence of the system set up by xg_cg_init runs in FP32
 * (FP32 matrix values, vectors and SpMV, the solution increment kept in
 * FP32); whenever its residual has fallen by the factor delta (0 < delta <
 * 1) since the last update, the increment is added to the FP64 x, the FP64
 * residual b - A x replaces the FP32 one and the recurrence continues with
 * its search direction.  Stops when ||b - A x|| < tol*normb (FP64, checked
 * at every update) or after maxit iterations; *normr is that FP64 residual
 * norm, *nupdates the number of reliable updates. */
int xg_cg_solve_reliable(xg_ctx* ctx, int64_t maxit, double tol, double delta,
                         double* normr, int64_t* iters, int64_t* nupdates,
                         int* converged);

/* Mixed precision solve of the system set up by xg_cg_init (no counterpart
 * in the reference): FP64 residuals r = b - A x, FP32 preconditioned CG on
 * A e = r to ||r_e|| < inner_tol*||r|| (or inner_maxit), x += e, until the
 * true residual ||b - A x|| < tol*normb or maxit inner iterations in total.
 * *normr is the true FP64 residual norm. */
int xg_cg_solve_mixed(xg_ctx* ctx, int64_t maxit, double tol,
                      double inner_tol, int64_t inner_maxit, double* normr,
                      int64_t* iters, int64_t* nouter, int* converged);

/* Flexible CG (no counterpart in the reference): Notay's FCG(1) in FP64 on
 * the system set up by xg_cg_init, with preconditioner pc_kind 0 (Jacobi),
 * 1 (`inner` iterations of FP32 Jacobi-PCG from a zero guess, i.e. FP32
 * inner solves inside an FP64 outer Krylov recurrence), 2 (the learned
 * preconditioner, xg_gnn_*, judged by the residual: Jacobi for an
 * iteration whose learned z is no descent direction, and for the rest of the
 * solve once ||r|| has not improved on its best value for 20 iterations) or
 * 3 (one FP32 V-cycle of smoothed-aggregation algebraic multigrid on
 * sign(diag A) A, built on the host from the matrix with its Dirichlet rows
 * when first used or when the values change; one component per node).
 * Stops when ||r|| < tol*normb or after maxit outer iterations; *iters outer
 * iterations, *inner_total the inner iterations (pc 1), the iterations
 * that used Jacobi (pc 2) or the V-cycles (pc 3). */
int xg_cg_solve_flexible(xg_ctx* ctx, int pc_kind, int64_t inner, int64_t maxit,
                         double tol, double* normr, int64_t* iters,
                         int64_t* inner_total, int* converged);

/* The multigrid hierarchy of pc_kind 3: number of levels, operator
 * complexity (sum of the levels' nonzeros over the fine level's) and the
 * host setup time of its last build. */
int xg_amg_info(xg_ctx* ctx, int* levels, double* op_complexity, double* setup_ms);

/* ---- learned graph preconditioner (no counterpart in the reference) -------- */

/* A 3-layer message-passing network over the graph of the CG matrix (32
 * channels, per-node 64->32 layers on BF16 MFMA with FP32 accumulation) that
 * maps a residual r to an approximate error z; used by xg_cg_solve_flexible
 * with pc_kind 2, with the Jacobi z wherever (z, r) <= 0.  Parameters are a
 * flat array of XG_GNN_NPARAM = 6337 values: we[32] be[32], for each layer
 * W[32][64] b[32], wo[32] bo[1].  n, nnz: rows and nonzeros of the system
 * matrix it will precondition. */
int xg_gnn_create(xg_ctx* ctx, int64_t n, int64_t nnz);
int xg_gnn_set_params(xg_ctx* ctx, const double* params, int64_t nparam);
/* params and/or the accumulated gradient (either may be null) */
int xg_gnn_get_params(xg_ctx* ctx, double* params, double* grad, int64_t nparam);
/* Adam step with the gradient accumulated since the last step (averaged over
 * the pairs it holds), which is then cleared. */
int xg_gnn_adam(xg_ctx* ctx, double lr, double beta1, double beta2, double eps);
/* Normalised matrix of the current CG system (after xg_cg_init/xg_cg_setup
 * with pc = jacobi). */
int xg_gnn_prepare(xg_ctx* ctx);
/* z = network(r), device vectors of the CG system's size. */
int xg_gnn_apply_dev(xg_ctx* ctx, const double* r, double* z);
/* Loss of the pair (r, e), A e = r: the energy-norm error of z = network(r)
 * relative to that of e, (z-e)^T A (z-e) / (e^T r); its gradient is added to
 * the accumulated gradient; z receives the network output. */
int xg_gnn_loss_grad_dev(xg_ctx* ctx, const double* r, const double* e, double* z,
                         double* loss);
/* Trains on residual/error pairs of solves of the current CG system (see
 * xg_cg.hip): restarted Jacobi-PCG from x = 0 in runs of `stride`
 * iterations, one pair per restart, an Adam step every `batch` pairs. */
int xg_gnn_train_system(xg_ctx* ctx, int64_t stride, int64_t max_pairs, int batch,
                        double lr, double* mean_loss, int64_t* npairs);

/* Releases the solver's device memory. */
int xg_cg_destroy(xg_ctx* ctx);

/* ---- ExodusII files (NetCDF classic, CDF-1/CDF-2), host only ------------
 * The mesh input and field output of the path without the SEACAS/NetCDF
 * libraries.  A handle is returned by xg_exo_open/xg_exo_create even when
 * they fail (then holding the message for xg_exo_error); xg_exo_close
 * releases it.  Node ids are 0-based; steps are 0-based record indices. */
typedef struct xg_exo xg_exo;
const char* xg_exo_error(const xg_exo* exo);
void xg_exo_close(xg_exo* exo);
/* ExodusIIMeshReader(filename) (ExodusIIMeshReader.hpp:53-58); append != 0
 * opens a file written by xg_exo_create for further steps (ExoWriter::OPEN,
 * ExodusIIMeshWriter.hpp:32-35). */
int xg_exo_open(const char* path, int append, xg_exo** exo);
/* readHeader/nelem (ExodusIIMeshReader.hpp:136,245): nodes, tetrahedra and
 * triangles over all element blocks, side sets, node variables, steps. */
int xg_exo_sizes(xg_exo* exo, int64_t* npoin, int64_t* ntet, int64_t* ntri,
                 int* nsideset, int* nnodevar, int64_t* nstep);
/* readNodes (ExodusIIMeshReader.hpp:103): coordx/y/z (or coord). */
int xg_exo_coords(xg_exo* exo, double* x, double* y, double* z);
/* readElements of the tetrahedron blocks (ExodusIIMeshReader.hpp:109),
 * 4 ids per element, blocks in file order. */
int xg_exo_tets(xg_exo* exo, int64_t* inpoel);
/* readFaces (ExodusIIMeshReader.hpp:86): the triangle blocks, 3 ids each. */
int xg_exo_tris(xg_exo* exo, int64_t* triinpoel);
/* Side set k (0-based position in the file): its id (ss_prop1) and size. */
int xg_exo_sideset(xg_exo* exo, int k, int* id, int64_t* nface);
/* readSidesetFaces (ExodusIIMeshReader.hpp:82-86) resolved to node triples:
 * a side naming a triangle element gives that triangle, a side of a
 * tetrahedron gives that side's nodes in ExodusII side order. */
int xg_exo_sideset_faces(xg_exo* exo, int k, int64_t* tri);
/* readNodeVarNames / readTimeValues / readNodeScalars
 * (ExodusIIMeshReader.hpp:124-133). */
int xg_exo_node_var_name(xg_exo* exo, int var, char* buf, int len);
int xg_exo_times(xg_exo* exo, double* t);
int xg_exo_read_node_var(xg_exo* exo, int64_t step, int var, double* out);
/* writeMesh + writeNodeVarNames (ExodusIIMeshWriter.hpp:41,64): creates the
 * file with coordinates, a triangle block (ntri > 0, element ids 1..ntri), a
 * tetrahedron block, side sets listing triangles (ss_face[ss_off[s] ..
 * ss_off[s+1]) index triinpoel, like UnsMesh::bface) and the node-variable
 * names, which a NetCDF classic header must hold before any step. */
int xg_exo_create(const char* path, int64_t npoin, const double* x, const double* y,
                  const double* z, int64_t ntet, const int64_t* inpoel, int64_t ntri,
                  const int64_t* triinpoel, int nsideset, const int* ss_id,
                  const int64_t* ss_off, const int64_t* ss_face, int nnodevar,
                  const char* const* names, xg_exo** exo);
/* writeTimeStamp / writeNodeScalar (ExodusIIMeshWriter.hpp:58,75); steps are
 * written in order: step <= the current count (a new step is appended). */
int xg_exo_write_time(xg_exo* exo, int64_t step, double time);
int xg_exo_write_node_var(xg_exo* exo, int64_t step, int var, const double* vals);

#ifdef __cplusplus
}
#endif
#endif
