/* CPU oracle for the edge-based flow path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms, used by tests/, by the
 * smoke check and by the CPU-baseline leg of bench.py to check and to time
 * against.  The product path (xyst_amd/) never links or calls it.
 *
 * Conventions follow the reference: 0-based local node ids, FP64, nodal
 * fields stored unknown-major AoS, U(i,c) = U[i*ncomp + c] (Data.hpp:420-426).
 * Each function names the reference lines it restates.
 */
#ifndef XYST_ORACLE_H
#define XYST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics configuration read from g_cfg by the reference. */
typedef struct {
  double gamma;       /* mat_spec_heat_ratio (EOS.hpp:30-44) */
  int    flux;        /* 0 = rusanov, 1 = hllc (Riemann.cpp:674-684) */
  int    stab2;       /* artificial viscosity on/off (Riemann.cpp:449-458) */
  double stab2coef;
} orc_cfg;

/* Discretization::vol, Discretization.cpp:619-647: V/4 of every tet summed to
 * its nodes.  Returns 0, or -1 if a tet has a non-positive Jacobian. */
int orc_vol(int64_t npoin, const double* x, const double* y, const double* z,
            int64_t ntet, const int64_t* inpoel, double* vol);

/* RieCG::domint, RieCG.cpp:339-382.  Unique edges (p<q), sorted
 * lexicographically, with the edge integral D accumulated in tet order.
 * edge[2*cap], dint[3*cap].  Returns the number of edges (or -1 if cap is too
 * small). */
int64_t orc_domint(int64_t npoin, const double* x, const double* y,
                   const double* z, int64_t ntet, const int64_t* inpoel,
                   int64_t cap, int64_t* edge, double* dint);

/* RieCG::domsuped, RieCG.cpp:620-736.  Greedy superedges: tets whose six edges
 * are all free, then tet faces (first-seen order) whose three edges are free,
 * then the remaining edges in (p,q) order.  Outputs are sized by the caller:
 * sup0[4*ntet], int0[18*ntet], sup1[3*4*ntet], int1[9*4*ntet],
 * sup2[2*nedge], int2[3*nedge]; nsup[3] receives the group counts. */
int orc_domsuped(int64_t ntet, const int64_t* inpoel, int64_t nedge,
                 const int64_t* edge, const double* dint, int64_t* nsup,
                 int64_t* sup0, double* int0, int64_t* sup1, double* int1,
                 int64_t* sup2, double* int2);

/* RieCG::bndint + bnorm + streamable (RieCG.cpp:281-337, 483-618) restricted
 * to symmetry BCs: inverse-distance^2-weighted point normals per side set,
 * symmetry node list with one normal per (node, symmetry set) pair in the
 * order of symsets[], and the besym flag of every boundary-triangle node.
 * nfaces boundary triangles tri[3*nfaces] with side set ids faceset[nfaces].
 * symnodes[cap], symnorms[3*cap], besym[3*nfaces].  Returns the number of
 * (node, normal) pairs. */
int64_t orc_symbc_setup(int64_t npoin, const double* x, const double* y,
                        const double* z, int64_t nfaces, const int64_t* tri,
                        const int32_t* faceset, int32_t nsym,
                        const int32_t* symsets, int64_t cap,
                        int64_t* symnodes, double* symnorms, uint8_t* besym);

/* riemann::grad, Riemann.cpp:229-367.  G is (npoin, 3*ncomp), zeroed here.
 * A (nullable, same shape) receives, for every term added to G, the
 * magnitude of the products it is made of (|D|(|u_p|+|u_q|), ...) summed per
 * entry: the scale of round-off at each node, used as the parity yardstick. */
void orc_grad(int64_t npoin, int ncomp, const int64_t* nsup,
              const int64_t* sup0, const double* int0,
              const int64_t* sup1, const double* int1,
              const int64_t* sup2, const double* int2,
              const double* x, const double* y, const double* z,
              int64_t nbtri, const int64_t* triinpoel,
              const double* U, double* G, double* A);

/* riemann::rhs, Riemann.cpp:909-946 (advdom 652-766, advbnd 768-878; the
 * source term src() is null for the problems covered here).  G is the nodal
 * gradient already divided by the nodal volume.  R is zeroed here.  A
 * (nullable) receives, per entry of R, the summed magnitudes of the products
 * each added flux is made of (e.g. fw(|l|+|r|), |p n|): the round-off scale. */
int orc_rhs(const orc_cfg* cfg, int64_t npoin, int ncomp, const int64_t* nsup,
            const int64_t* sup0, const double* int0,
            const int64_t* sup1, const double* int1,
            const int64_t* sup2, const double* int2,
            const double* x, const double* y, const double* z,
            int64_t nbtri, const int64_t* triinpoel, const uint8_t* besym,
            const double* G, const double* U, double* R, double* A);

/* orc_rhs over nthread OpenMP threads, each on a contiguous share of the
 * superedges and boundary triangles, partial sums added per node in thread
 * order (the reference's per-chare rhs plus comrhs, RieCG.cpp:948-965). */
int orc_rhs_mt(int nthread, const orc_cfg* cfg, int64_t npoin, int ncomp,
               const int64_t* nsup, const int64_t* sup0, const double* int0,
               const int64_t* sup1, const double* int1,
               const int64_t* sup2, const double* int2,
               const double* x, const double* y, const double* z,
               int64_t nbtri, const int64_t* triinpoel, const uint8_t* besym,
               const double* G, const double* U, double* R);

/* RieCG::dt, RieCG.cpp:787-852 (unsteady branch): cfl * min over nodes of
 * cbrt(vol)/max(|u|+c, 1e-8). */
double orc_dt(const orc_cfg* cfg, int64_t npoin, int ncomp, const double* U,
              const double* vol, double cfl);

/* physics::symbc, BC.cpp:110-136 (pos = 1). */
void orc_symbc(int64_t npoin, int ncomp, double* U, int64_t nsym,
               const int64_t* symnodes, const double* symnorms);

/* One complete RieCG time step: rhs-stage loop of RieCG::advance/grad/rhs/
 * solve/stage (RieCG.cpp:854-1057, 1409-1421) with rkcoef {1/3,1/2,1}
 * (RieCG.cpp:41) and symmetry BCs after each stage.  dt is given (the caller
 * evaluates orc_dt and setdt truncation).  Un receives U at time n.
 * Work arrays: G (npoin*3*ncomp), R (npoin*ncomp). */
int orc_riecg_step(const orc_cfg* cfg, int64_t npoin, int ncomp,
                   const int64_t* nsup, const int64_t* sup0, const double* int0,
                   const int64_t* sup1, const double* int1,
                   const int64_t* sup2, const double* int2,
                   const double* x, const double* y, const double* z,
                   int64_t nbtri, const int64_t* triinpoel,
                   const uint8_t* besym, const double* vol,
                   int64_t nsym, const int64_t* symnodes,
                   const double* symnorms, double dt,
                   double* U, double* Un, double* G, double* R);

/* NodeDiagnostics::rhocompute, NodeDiagnostics.cpp:46-146 (no analytic
 * solution): out[0..ncomp) = sum U^2 V, out[ncomp..2ncomp) = sum (U-Un)^2 V,
 * out[2ncomp] = sum rhoE V. */
void orc_diag(int64_t npoin, int ncomp, const double* U, const double* Un,
              const double* v, double* out);

/* ---- problems, BCs, steady stepping, error norms (problems.c) ---------- */

/* An analytic problem of Problems.cpp: kind 0 none (sod, sedov, userdef:
 * SRC() and SOL() null), 1 vortical_flow (Problems.cpp:452-509), 2
 * taylor_green (Problems.cpp:407-450). */
typedef struct {
  int    kind;
  double gamma;                 /* mat_spec_heat_ratio */
  double alpha, kappa, p0;      /* problem_alpha, problem_kappa, problem_p0 */
} orc_problem;

/* IC() of the problem at (x, y, z, t): the conserved values (5, or 6 with
 * slot_cyl's scalar); returns how many, -1 for kind 0. */
int orc_problem_ic(const orc_problem* P, double x, double y, double z,
                   double t, double* s);
/* SRC() of the problem: the 5 flow values (the scalars have no source);
 * returns 0 (and zeros) if it has none. */
int orc_problem_src(const orc_problem* P, double x, double y, double z,
                    double t, double* s);

/* riemann::src, Riemann.cpp:880-907: R(p,c) -= s_c(x_p, t or tp[p]) v[p]. */
void orc_src(const orc_problem* P, int64_t npoin, int ncomp, const double* x,
             const double* y, const double* z, const double* v, double t,
             const double* tp, double* R);

/* physics::dirbc, BC.cpp:29-72: ncomp mask entries per node (1: IC at t, or
 * vals where the problem has no analytic IC; 2: vals). */
void orc_dirbc(const orc_problem* P, int64_t npoin, int ncomp, double* U,
               double t, const double* x, const double* y, const double* z,
               int64_t ndir, const int64_t* nodes, const int32_t* mask,
               const double* vals);

/* physics::farbc, BC.cpp:152-221; far = {density, u, v, w, pressure}. */
void orc_farbc(double g, int ncomp, double* U, int64_t nfar,
               const int64_t* nodes, const double* norms, const double* far);

/* physics::prebc, BC.cpp:223-241; vals = (density, pressure) per node. */
void orc_prebc(double g, int ncomp, double* U, int64_t npre,
               const int64_t* nodes, const double* vals);

/* RieCG::dt steady branch (RieCG.cpp:810-824): dtp[p] and its minimum. */
double orc_dt_steady(const orc_cfg* cfg, int64_t npoin, int ncomp,
                     const double* U, const double* vol, double cfl,
                     double* dtp);

/* The boundary-condition node lists of the RieCG chare (setupBC and
 * streamable, RieCG.cpp:109-245, 527-618). */
typedef struct {
  int64_t nsym; const int64_t* symnodes; const double* symnorms;
  int64_t ndir; const int64_t* dirnodes; const int32_t* dirmask;
  const double* dirvals;
  int64_t nfar; const int64_t* farnodes; const double* farnorms;
  double far[5];
  int64_t npre; const int64_t* prenodes; const double* prevals;
} orc_riecg_bc;

/* One RieCG time step with source, every BC and optional local time
 * stepping (dtp non-null: steady, tp advanced by dtp); RieCG.cpp:764-1057. */
int orc_riecg_step2(const orc_cfg* cfg, const orc_problem* P, int64_t npoin,
                    int ncomp, const int64_t* nsup, const int64_t* sup0,
                    const double* int0, const int64_t* sup1, const double* int1,
                    const int64_t* sup2, const double* int2, const double* x,
                    const double* y, const double* z, int64_t nbtri,
                    const int64_t* triinpoel, const uint8_t* besym,
                    const double* vol, const double* v, const orc_riecg_bc* bc,
                    double t, double dt, const double* dtp, double* tp,
                    double* U, double* Un, double* G, double* R);

/* NodeDiagnostics::rhocompute analytic-error sums (NodeDiagnostics.cpp:
 * 84-128): out[c] L2 sums, out[ncomp+c] L1 sums of the primitive error
 * against An (the analytic solution in primitive form). */
void orc_diag_err(int64_t npoin, int ncomp, const double* U, const double* v,
                  const double* An, double* out);

/* ---- KozCG (kozcg.c) ------------------------------------------------------ */

/* kozak::rhs, Kozak.cpp:29-180: Taylor-Galerkin rhs, R zeroed first.  dtp:
 * per-node time step of steady mode (null: dt for every tet). */
void orc_kozak_rhs(const orc_cfg* cfg, int64_t npoin, int ncomp,
                   const double* x, const double* y, const double* z,
                   int64_t ntet, const int64_t* inpoel, double dt,
                   const double* dtp, const double* U, double* R);

/* kozak::rhs with the problem's source term (Kozak.cpp:124-136, 163-173):
 * the half step adds dt/8 s(x_a, t) per corner (t = mean tp of the corners
 * when tp is given), the second step J/24 s(x_centroid, t + dt/2). */
void orc_kozak_rhs_src(const orc_cfg* cfg, const orc_problem* P,
                       int64_t npoin, int ncomp, const double* x,
                       const double* y, const double* z, int64_t ntet,
                       const int64_t* inpoel, double t, double dt,
                       const double* tp, const double* dtp, const double* U,
                       double* R);

/* The same evaluation with first-order round-off magnitudes: A receives, per
 * R entry, the sum over its added terms of m(term) (kozcg.c). */
void orc_kozak_rhs_mag(const orc_cfg* cfg, int64_t npoin, int ncomp,
                       const double* x, const double* y, const double* z,
                       int64_t ntet, const int64_t* inpoel, double dt,
                       const double* U, double* R, double* A);

/* KozCG flux-corrected transport settings (InciterConfig.cpp:1751-1754) */
typedef struct {
  int    fct;          /* FCT on (default) or off */
  double fctdif;       /* antidiffusion coefficient, default 1.0 */
  int    fctclip;      /* bounds from the low-order solution only */
  int    nsys;         /* components limited together (fctsys, 0-based) */
  int    sys[16];
} orc_fctcfg;

/* One KozCG step (KozCG.cpp:709-1197): kozak::rhs, then with FCT the
 * antidiffusive contributions (aec, with symmetry BCs on them), the low-order
 * solution and allowed bounds (alw), the limit coefficients and limited
 * contributions (lim), the new solution (solve); without FCT U + dt R/vol.
 * Then Dirichlet values (all components of the ndir nodes, dirvals
 * ndir*ncomp) and symmetry BCs.  U is advanced in place, Un receives the old
 * U.  P (nullable) adds its source term to kozak::rhs at time t.
 * Steady (local time stepping, dtp and tp nonnull): kozak::rhs with dtp/tp,
 * the low-order solution (or U + dt R/vol) with each node's dtp, and tp +=
 * dtp at the end.  dirmask (nullable: every component) selects the
 * Dirichlet components (nonzero).  freeze: the flow components keep their
 * old values after the BCs (freezeflow > 1, KozCG.cpp:1141, 1168-1175).
 * Returns 0, or -1 for ncomp outside 5..16. */
int orc_kozcg_step(const orc_cfg* cfg, const orc_fctcfg* f,
                   const orc_problem* P, double t, int64_t npoin,
                   int ncomp, const double* x, const double* y,
                   const double* z, int64_t ntet, const int64_t* inpoel,
                   const double* vol, double dt, const double* dtp, double* tp,
                   int freeze, int64_t nsym,
                   const int64_t* symnodes, const double* symnorms,
                   int64_t ndir, const int64_t* dirnodes, const int32_t* dirmask,
                   const double* dirvals, double* U, double* Un);

/* ---- ChoCG (chocg.c, riecg.c) -------------------------------------------- */

/* ChoCG configuration read from g_cfg (InciterConfig.cpp:1724-1757) */
typedef struct {
  int    flux;        /* 0 = damp2, 1 = damp4 (Chorin.cpp:862-866) */
  int    stab;        /* |vn_L + vn_R|/2 stabilization */
  int    stab2;
  double stab2coef;
  double mu;          /* mat_dyn_viscosity */
  double diff;        /* mat_dyn_diffusivity */
} orc_chocfg;

/* ChoCG::domint, ChoCG.cpp:398-445: as orc_domint with 5 values per edge
 * (signed normal components, J/120, dot(grad_p,grad_q)/J/6). */
int64_t orc_chocg_domint(int64_t npoin, const double* x, const double* y,
                         const double* z, int64_t ntet, const int64_t* inpoel,
                         int64_t cap, int64_t* edge, double* dint);

/* ChoCG::domsuped, ChoCG.cpp:687-813: orc_domsuped with stride-5 integrals
 * (only the first three components change sign with the orientation). */
int orc_domsuped5(int64_t ntet, const int64_t* inpoel, int64_t nedge,
                  const int64_t* edge, const double* dint, int64_t* nsup,
                  int64_t* sup0, double* int0, int64_t* sup1, double* int1,
                  int64_t* sup2, double* int2);

/* chorin::div, Chorin.cpp:34-209 (accumulates into D; P, G pressure and its
 * gradient (ldg columns, 0..2 used) enter only with stab) */
void orc_chorin_div(const int64_t* nsup, const int64_t* sup0, const double* int0,
                    const int64_t* sup1, const double* int1, const int64_t* sup2,
                    const double* int2, const double* x, const double* y,
                    const double* z, int64_t ntri, const int64_t* tri, double dt,
                    const double* P, const double* G, int ldg, const double* U,
                    int ldu, double* D, int stab, double* A);

/* chorin::vgrad, Chorin.cpp:211-334 (accumulates G(npoin, 3*ncomp)) */
void orc_chorin_vgrad(const int64_t* nsup, const int64_t* sup0, const double* int0,
                      const int64_t* sup1, const double* int1, const int64_t* sup2,
                      const double* int2, const double* x, const double* y,
                      const double* z, int64_t ntri, const int64_t* tri,
                      const double* U, int ncomp, double* G, double* A);

/* chorin::grad, Chorin.cpp:336-449 (scalar U; accumulates G columns 0..2) */
void orc_chorin_grad(const int64_t* nsup, const int64_t* sup0, const double* int0,
                     const int64_t* sup1, const double* int1, const int64_t* sup2,
                     const double* int2, const double* x, const double* y,
                     const double* z, int64_t ntri, const int64_t* tri,
                     const double* U, double* G, int ldg, double* A);
/* ChoCG::pinit's Neumann vector of the pressure rhs (ChoCG.cpp:1070-1095):
 * neubc (npoin) from the flagged boundary faces and the pressure gradient
 * pg (npoin x 3) at the nodes. */
void orc_neumann(int64_t npoin, const double* x, const double* y, const double* z,
                 int64_t ntri, const int64_t* tri, const unsigned char* besym,
                 const double* pg, double* neubc);


/* chorin::flux, Chorin.cpp:451-638 (accumulates F(npoin, 3)) */
void orc_chorin_flux(const int64_t* nsup, const int64_t* sup0, const double* int0,
                     const int64_t* sup1, const double* int1, const int64_t* sup2,
                     const double* int2, const double* x, const double* y,
                     const double* z, int64_t ntri, const int64_t* tri,
                     int64_t npoin, const double* U, int ldu, const double* G,
                     int ldg, double mu, double* F, double* A);

/* chorin::rhs, Chorin.cpp:1010-1044 (adv only: no source); R zeroed.
 * Returns -1 for an unknown flux, -2 for ncomp outside 3..16. */
int orc_chorin_rhs(const orc_chocfg* cfg, int64_t npoin, int ncomp,
                   const int64_t* nsup, const int64_t* sup0, const double* int0,
                   const int64_t* sup1, const double* int1, const int64_t* sup2,
                   const double* int2, const double* x, const double* y,
                   const double* z, int64_t ntri, const int64_t* tri,
                   const double* P, const double* U, const double* G, double* R,
                   double* A);

/* ---- linear solver (cg.c) ------------------------------------------------ */

/* tk::CSR constructor (CSR.cpp:19-84) from tets: ia[npoin*ncomp+1],
 * ja[cap]; 0-based, rows = node + edge neighbours ascending.  Returns nnz. */
int64_t orc_csr_build(int64_t npoin, int64_t ntet, const int64_t* inpoel,
                      int ncomp, int64_t* ia, int64_t* ja, int64_t cap);

/* ChoCG::prelhs (ChoCG.cpp:146-188): a = -K (ncomp = 1). */
int orc_prelhs(int64_t npoin, const double* x, const double* y,
               const double* z, int64_t ntet, const int64_t* inpoel,
               const int64_t* ia, const int64_t* ja, double* a);

/* +K of TestConjugateGradients.cpp:175-199, every component. */
int orc_laplace_test(int64_t npoin, const double* x, const double* y,
                     const double* z, int64_t ntet, const int64_t* inpoel,
                     int ncomp, const int64_t* ia, const int64_t* ja,
                     double* a);

/* CSR::dirichlet (CSR.cpp:106-152), one partition. */
void orc_csr_dirichlet(int64_t nrow, int ncomp, const int64_t* ia,
                       const int64_t* ja, double* a, int64_t i, double val,
                       double* b, int pos);

/* CSR::mult (CSR.cpp:154-172). */
void orc_csr_mult(int64_t nrow, const int64_t* ia, const int64_t* ja,
                  const double* a, const double* x, double* r);

/* ConjugateGradients init(apply)/setup/solve (ConjugateGradients.cpp:
 * 105-823), one partition.  dirnode ascending, dirmask/dirval ncomp per
 * node; pc 0 = "none", 1 = "jacobi".  x is the warm start and the result; b
 * is modified as the reference modifies it; a is restored when apply.
 * rhist (nullable, maxit entries) receives sqrt((r,r)) per iteration. */
int orc_cg(int64_t npoin, int ncomp, const int64_t* ia, const int64_t* ja,
           double* a, double* x, double* b, const double* neubc,
           int64_t ndir, const int64_t* dirnode, const int32_t* dirmask,
           const double* dirval, int apply, int pc, int64_t maxit,
           double tol, double* normb_out, double* normr_out, int64_t* it_out,
           int* converged_out, double* rhist);

#ifdef __cplusplus
}
#endif
#endif
