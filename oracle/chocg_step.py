"""CPU oracle of the ChoCG time step -- TEST INFRASTRUCTURE ONLY.

Restates the ChoCG chare's control flow around the chorin operators and the
pressure solve (ChoCG.cpp), on one partition, with the C oracle's operators
(chorin_div/grad/vgrad/flux/rhs, prelhs, ConjugateGradients):

  setup    ChoCG ctor/feop/streamable (ChoCG.cpp:55-146, 211-301, 575-686):
           pressure matrix -K (prelhs), Dirichlet velocity masks and values
           (setupDirBC), symmetry and noslip node lists
  initial  merge (817-839) -> div(u) -> pinit -> psolve -> sgrad -> psolved
           (m_np 0: project u -= grad(p), BC) -> velgrad -> flux -> div(flux)
           -> pinit -> psolve -> psolved (m_np 1: p = solution) -> pgrad
  step     dt (1357-1413) -> rhs (1480-1504) -> solve (1528-1607, explicit
           RK stages, rkcoef of ChoCG.cpp:43-48) -> pred (BC, velgrad for
           damp4) -> corr -> div(u) with the 4th-order pressure damping ->
           pinit (div/dt, hydrostat / Dirichlet pressure increment 0) ->
           psolve (warm start from the previous increment) -> sgrad ->
           psolved (u -= dt grad(dp), BC, p += dp) -> pgrad -> diag
  diag     NodeDiagnostics::precompute (NodeDiagnostics.cpp:148-264):
           sums of p^2 V, u_c^2 V, dp^2 V, (u_c - un_c)^2 V

Gradients are finalised as in ChoCG::fingrad (841-866): divided by the nodal
volume.  BCs (ChoCG::BC, 1341-1355): dirbc (mask 1: initial value, mask 2:
configured value), then symmetry (pos 0), then noslip.  The theta-implicit momentum solve
(theta > 0) is not restated: the cases served run explicit RK.
"""
import numpy as np

import oracle as O


class ChoCGRun:
    """One ChoCG partition.  Velocity U (npoin, 3), pressure P (npoin,)."""

    def __init__(self, coord, inpoel, tri, faceset, mu, flux="damp4", stab=True,
                 stab2=False, stab2coef=0.2, cfl=0.9, rk=1, noslip=(), sym=(),
                 dirbc=None, p_hydrostat=None, p_dirbc=None, p_tol=1e-3, p_iter=500,
                 p_pc="jacobi", ic=(0.0, 0.0, 0.0)):
        self.coord = np.ascontiguousarray(coord, dtype=np.float64)
        self.inpoel = np.ascontiguousarray(inpoel, dtype=np.int64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int64)
        self.faceset = np.asarray(faceset)
        n = self.coord.shape[1]
        self.npoin = n
        edge, dint = O.chocg_domint(self.coord, self.inpoel)
        self.se = O.domsuped(self.inpoel, edge, dint)
        self.vol = O.vol(self.coord, self.inpoel)
        self.A = O.CSR.from_tets(n, self.inpoel)
        O.prelhs(self.coord, self.inpoel, self.A)
        self.cfg = O.chocfg(flux, stab, stab2, stab2coef, mu)
        self.flux = flux
        self.mu = mu
        self.cfl = cfl
        self.rkcoef = [[1.0], [0.5, 1.0], [1 / 3, 0.5, 1.0],
                       [0.25, 1 / 3, 0.5, 1.0]][rk - 1]
        nodes_of = lambda sets: np.unique(self.tri[np.isin(self.faceset, list(sets))])
        self.noslip = nodes_of(noslip) if len(noslip) else np.zeros(0, np.int64)
        if len(sym):
            self.symnodes, self.symnorms, _ = O.symbc_setup(self.coord, self.tri,
                                                            self.faceset, list(sym))
        else:
            self.symnodes, self.symnorms = np.zeros(0, np.int64), np.zeros((0, 3))
        # setupDirBC: per node a mask (0/1/2) and values; masks of several sets
        # merge (0 -> nonzero), values are the last set's
        self.dir = {}
        for s, (mask, val) in (dirbc or {}).items():
            for p in nodes_of([s]):
                m, v = self.dir.get(int(p), ([0, 0, 0], None))
                m = [m[c] if m[c] else mask[c] for c in range(3)]
                self.dir[int(p)] = (m, list(val))
        # pressure Dirichlet nodes and the values of the start-up solves
        # (mask 1: initial value 0, mask 2: the configured value); the
        # time-step solves set the increment to 0 there
        self.pdir = {}
        for sset, (mask, val) in (p_dirbc or {}).items():
            for p in nodes_of([sset]):
                self.pdir[int(p)] = 0.0 if mask == 1 else float(val)
        self.hydrostat = p_hydrostat
        self.ic = np.asarray(ic, np.float64)
        self.p_tol, self.p_iter, self.p_pc = p_tol, p_iter, p_pc
        self.U = np.tile(np.asarray(ic, np.float64), (n, 1))
        self.Un = self.U.copy()
        self.P = np.zeros(n)
        self.pgrad = np.zeros((n, 3))
        self.vgrad = np.zeros((n, 9))
        self.x = np.zeros(n)          # the pressure solver's solution (warm start)
        self.div = np.zeros(n)
        self.t, self.dt, self.it = 0.0, 0.0, 0
        self.np_ = 0
        self.pit = []

    # ---- helpers --------------------------------------------------------
    def BC(self, U):
        for p, (m, v) in self.dir.items():
            for c in range(3):
                if m[c] == 1:
                    U[p, c] = self.ic[c]
                elif m[c] == 2:
                    U[p, c] = v[c]
        if self.symnodes.size:
            for k, p in enumerate(self.symnodes):
                nx, ny, nz = self.symnorms[k]
                vn = U[p, 0] * nx + U[p, 1] * ny + U[p, 2] * nz
                U[p, 0] -= vn * nx
                U[p, 1] -= vn * ny
                U[p, 2] -= vn * nz
        U[self.noslip, :] = 0.0

    def _grad(self, s):
        return O.chorin_grad(self.se, self.coord, self.tri, s) / self.vol[:, None]

    def _divergence(self, u, stab):
        return O.chorin_div(self.se, self.coord, self.tri, self.dt, self.P, self.pgrad,
                            u, stab=stab)

    def _psolve(self, b, initial=False):
        """pinit + psolve: Dirichlet pressure (the configured values in the
        start-up solves, increment 0 once stepping) on the configured sets
        and the hydrostat node, then CG warm-started from the previous
        solution (ConjugateGradients::init with empty x)."""
        dirbc = {p: [(1, v if initial else 0.0)] for p, v in self.pdir.items()}
        if self.hydrostat is not None and self.hydrostat not in dirbc:
            dirbc[self.hydrostat] = [(1, 0.0)]
        b = b.copy()
        normb, normr, it, conv = O.cg(self.A, self.x, b, dirbc=dirbc, apply=True,
                                      pc=self.p_pc, maxit=self.p_iter, tol=self.p_tol)
        self.pit.append(it)
        return self.x.copy()

    # ---- ChoCG control flow --------------------------------------------
    def initial(self):
        """merge ... start (ChoCG.cpp:817-839, 869-900, 1026-1323)."""
        self.BC(self.U)
        # m_np = 0: div(u), solve, project with dt = 1
        self.div = self._divergence(self.U, stab=False)
        sol = self._psolve(self.div, initial=True)
        self.U -= self._grad(sol)
        self.BC(self.U)
        self.np_ = 1
        # velgrad -> flux -> div(flux): initial pressure from the momentum flux
        self.vgrad = O.chorin_vgrad(self.se, self.coord, self.tri, self.U) / self.vol[:, None]
        F = O.chorin_flux(self.se, self.coord, self.tri, self.U, self.vgrad, self.mu)
        F /= self.vol[:, None]
        for k, p in enumerate(self.symnodes):
            nx, ny, nz = self.symnorms[k]
            vn = F[p, 0] * nx + F[p, 1] * ny + F[p, 2] * nz
            F[p] -= vn * np.array([nx, ny, nz])
        self.div = self._divergence(F, stab=False)
        sol = self._psolve(self.div, initial=True)
        self.np_ = 2
        self.P = sol.copy()
        self.pgrad = self._grad(self.P)

    def dt_cfl(self):
        """ChoCG::dt (1357-1413): cfl * min(L/max(|u|,1e-8), L^2/mu)."""
        u = self.U
        vel = np.sqrt(u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])
        L = np.cbrt(self.vol)
        m = np.min(L / np.maximum(vel, 1.0e-8))
        eps = np.finfo(float).eps
        if self.mu > eps:
            m = min(m, np.min(L * L / self.mu))
        return m * self.cfl

    def step(self, dt=None):
        self.dt = self.dt_cfl() if dt is None else dt
        for stage, c in enumerate(self.rkcoef):
            R = O.chorin_rhs(self.cfg, self.se, self.coord, self.tri, self.P, self.U,
                             self.vgrad)
            if stage == 0:
                self.Un = self.U.copy()
            self.U = self.Un - (c * self.dt) * R / self.vol[:, None]
            self.BC(self.U)
            if self.flux == "damp4":
                self.vgrad = (O.chorin_vgrad(self.se, self.coord, self.tri, self.U)
                              / self.vol[:, None])
        # projection
        div = self._divergence(self.U, stab=True)
        self.div = div / self.dt
        dp = self._psolve(self.div)
        self.U -= self.dt * self._grad(dp)
        self.BC(self.U)
        self.P += dp
        self.pgrad = self._grad(self.P)
        self.it += 1
        self.t += self.dt
        return dp

    def diag(self, dp):
        """Columns of the diagnostics file: it, t, dt, L2(p), L2(u,v,w),
        L2(dp), L2(du,dv,dw); sums over nodes of f^2 V, normalised by the
        total volume and square-rooted (Transporter::diagnostics)."""
        V = self.vol
        tot = V.sum()
        s = [np.sum(self.P * self.P * V)]
        s += [np.sum(self.U[:, c] * self.U[:, c] * V) for c in range(3)]
        s += [np.sum(dp * dp * V)]
        d = self.U - self.Un
        s += [np.sum(d[:, c] * d[:, c] * V) for c in range(3)]
        return np.concatenate([[self.it, self.t, self.dt], np.sqrt(np.array(s) / tot)])
