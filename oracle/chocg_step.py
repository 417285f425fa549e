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
  poisson  the run of the first pressure solve alone (nstep = 1, the
           Poisson problems): b = pr * vol + Neumann vector, Dirichlet
           values from the problem's pressure, p = solution; its
           diagnostics add the analytic-pressure error norms

Gradients are finalised as in ChoCG::fingrad (841-866): divided by the nodal
volume.  BCs (ChoCG::BC, 1341-1355): dirbc (mask 1: initial value, mask 2:
configured value), then symmetry (pos 0), then noslip.  With theta > 0 the
last RK stage solves the momentum system (ChoCG::lhs, solve, msolve,
msolved) with the 3-component ConjugateGradients.

Frozen flow (freezeflow > 1 after freezetime, ChoCG.cpp:1550-1568): the
velocity is put back after each stage's update, BCs and velocity
gradients, and with damp2 after the last stage's divergence.  The frozen
velocity staying constant is checked; the order of those restores (which
velocity the stage's gradients and divergence see) is parity unpinned:
the reference's frozen-flow case (SlotCyl) holds the velocity by its BCs,
so no fixture tells the orders apart.
"""
import numpy as np

import oracle as O


def pressure_problem(name, coord, beta=(0.0, 0.0, 0.0)):
    """The pressure functions of the Poisson problems at the nodes
    (Problems.cpp:841-996, PRESSURE_RHS/IC/GRAD at 1172-1264): the rhs pr,
    the initial/analytic pressure (the Dirichlet value of mask 1 and
    PRESSURE_SOL) and the gradient of the Neumann BCs (None where the
    problem has none)."""
    x, y, z = (np.asarray(coord[i], np.float64) for i in range(3))
    pi = np.pi
    pg = None
    if name == "poisson_const":
        pr = np.full(x.size, 6.0)
        ic = x * x + y * y + z * z
    elif name == "poisson_harmonic":
        pr = np.zeros(x.size)
        dx, dy, dz = x - beta[0], y - beta[1], z - beta[2]
        ic = 1.0 / np.sqrt(dx * dx + dy * dy + dz * dz)
    elif name == "poisson_sine":
        pr = -pi * pi * x * y * np.sin(pi * z)
        ic = x * y * np.sin(pi * z)
    elif name == "poisson_sine3":
        pr = -3.0 * pi * pi * np.sin(pi * x) * np.sin(pi * y) * np.sin(pi * z)
        ic = np.sin(pi * x) * np.sin(pi * y) * np.sin(pi * z)
    elif name == "poisson_neumann":
        pr = -3.0 * np.cos(2.0 * x) * np.exp(y)
        ic = np.cos(2.0 * x) * np.exp(y)
        pg = np.stack([-2.0 * np.sin(2.0 * x) * np.exp(y), np.cos(2.0 * x) * np.exp(y),
                       np.zeros(x.size)], 1)
    else:
        raise ValueError("unknown Poisson problem %r" % name)
    return pr, ic, pg


class ChoCGRun:
    """One ChoCG partition.  Velocity U (npoin, 3), pressure P (npoin,)."""

    def __init__(self, coord, inpoel, tri, faceset, mu, flux="damp4", stab=True,
                 stab2=False, stab2coef=0.2, cfl=0.9, rk=1, noslip=(), sym=(),
                 dirbc=None, p_hydrostat=None, p_dirbc=None, p_tol=1e-3, p_iter=500,
                 p_pc="jacobi", ic=(0.0, 0.0, 0.0), theta=0.0, m_tol=1e-3, m_iter=50,
                 m_pc="jacobi", p_ic=None, U0=None, src=None, diff=0.0, freezeflow=1.0,
                 freezetime=0.0):
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
        self.cfg = O.chocfg(flux, stab, stab2, stab2coef, mu, diff)
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
        # the initial state: velocity and the transported scalars (ncomp
        # columns), a constant ic or the field U0
        if U0 is None:
            U0 = np.tile(np.asarray(ic, np.float64), (n, 1))
        self.U0 = np.ascontiguousarray(U0, dtype=np.float64)
        nc = self.U0.shape[1]
        self.ncomp = nc
        # setupDirBC: per node a mask (0/1/2) per component and values; masks
        # of several sets merge (0 -> nonzero), values are the last set's
        self.dir = {}
        for s, (mask, val) in (dirbc or {}).items():
            for p in nodes_of([s]):
                m, v = self.dir.get(int(p), ([0] * nc, None))
                m = [m[c] if m[c] else mask[c] for c in range(nc)]
                self.dir[int(p)] = (m, list(val))
        # pressure Dirichlet nodes and the values of the start-up solves
        # (mask 1: the initial pressure p_ic at the node, 0 without one; mask
        # 2: the configured value); the time-step solves set the increment to
        # 0 there
        self.pdir = {}
        for sset, (mask, val) in (p_dirbc or {}).items():
            for p in nodes_of([sset]):
                if mask == 1:
                    self.pdir[int(p)] = 0.0 if p_ic is None else float(p_ic[p])
                else:
                    self.pdir[int(p)] = float(val)
        self.hydrostat = p_hydrostat
        # a time-independent source of chorin::src (Chorin.cpp:985-1008):
        # s (npoin, ncomp) at the nodes, R -= s vol
        self.src = None if src is None else np.asarray(src, np.float64)
        self.freezeflow, self.freezetime, self.ff = freezeflow, freezetime, 1.0
        self.p_tol, self.p_iter, self.p_pc = p_tol, p_iter, p_pc
        self.theta, self.m_tol, self.m_iter, self.m_pc = theta, m_tol, m_iter, m_pc
        self.mit = []
        if theta > np.finfo(float).eps:
            assert nc == 3, "the theta-implicit momentum system has 3 components"
            self._momentum_setup()
        self.U = self.U0.copy()
        self.Un = self.U.copy()
        self.P = np.zeros(n)
        self.pgrad = np.zeros((n, 3))
        self.vgrad = np.zeros((n, 3 * nc))
        self.x = np.zeros(n)          # the pressure solver's solution (warm start)
        self.div = np.zeros(n)
        self.t, self.dt, self.it = 0.0, 0.0, 0
        self.np_ = 0
        self.pit = []

    # ---- theta-implicit momentum system ---------------------------------
    def _momentum_setup(self):
        """ChoCG::momlhs (ChoCG.cpp:191-209): the 3-component CSR of the
        momentum system; per tet and node pair the mass and stiffness terms
        of ChoCG::lhs (1434-1477), with the CSR position of each
        (N_a, N_b, c) entry so that a step's matrix is assembled in the
        reference's element order."""
        n = self.npoin
        self.Am = O.CSR.from_tets(n, self.inpoel, ncomp=3)
        X = self.coord
        N = self.inpoel
        ba = (X[:, N[:, 1]] - X[:, N[:, 0]]).T
        ca = (X[:, N[:, 2]] - X[:, N[:, 0]]).T
        da = (X[:, N[:, 3]] - X[:, N[:, 0]]).T
        g1, g2, g3 = np.cross(ca, da), np.cross(da, ba), np.cross(ba, ca)
        J = np.einsum("ij,ij->i", ba, g1)
        g = np.stack([-g1 - g2 - g3, g1, g2, g3], 1)        # (T, 4, 3)
        ne = N.shape[0]
        rows, cols, mass, stiff = [], [], [], []
        for a in range(4):
            for b in range(4):
                dot = g[:, a, 0] * g[:, b, 0] + g[:, a, 1] * g[:, b, 1] + g[:, a, 2] * g[:, b, 2]
                for c in range(3):
                    rows.append(N[:, a] * 3 + c)
                    cols.append(N[:, b] * 3 + c)
                    mass.append((J, 2.0 if a == b else 1.0))
                    stiff.append(dot)
        # element-major order of the reference's loops: e, then a, b, c
        R = np.stack(rows, 1).reshape(-1)
        C = np.stack(cols, 1).reshape(-1)
        pos = np.empty(R.size, np.int64)
        ia, ja = self.Am.ia, self.Am.ja
        for k in range(R.size):
            r0, r1 = ia[R[k]], ia[R[k] + 1]
            pos[k] = r0 + np.searchsorted(ja[r0:r1], C[k])
        self._mpos = pos
        self._mJ = np.repeat(J, 48)
        self._mfac = np.tile(np.array([m[1] for m in mass]), ne)
        self._mdot = np.stack(stiff, 1).reshape(-1)

    def _momentum_matrix(self):
        """ChoCG::lhs: A(N_a,N_b,c) -= J/dt/120 (2 or 1) + theta mu
        grad_a.grad_b / J / 6, summed in element order."""
        v = self._mJ / self.dt / 120.0 * self._mfac
        v = v + self.theta * self.mu * self._mdot / self._mJ / 6.0
        a = np.zeros(self.Am.ja.size)
        np.add.at(a, self._mpos, -v)
        self.Am.a = a

    def _momentum_dirbc(self):
        """The momentum solve's Dirichlet data (ChoCG.cpp:1577-1600): the
        dirbc masks with value 0, noslip nodes on all three components."""
        bc = {}
        for p, (m, _) in self.dir.items():
            bc[p] = [(m[c], 0.0) for c in range(3)]
        for p in self.noslip:
            bc[int(p)] = [(1, 0.0)] * 3
        return bc

    # ---- helpers --------------------------------------------------------
    def BC(self, U):
        for p, (m, v) in self.dir.items():
            for c in range(self.ncomp):
                if m[c] == 1:
                    U[p, c] = self.U0[p, c]
                elif m[c] == 2:
                    U[p, c] = v[c]
        if self.symnodes.size:
            for k, p in enumerate(self.symnodes):
                nx, ny, nz = self.symnorms[k]
                vn = U[p, 0] * nx + U[p, 1] * ny + U[p, 2] * nz
                U[p, 0] -= vn * nx
                U[p, 1] -= vn * ny
                U[p, 2] -= vn * nz
        U[self.noslip, :3] = 0.0

    def _grad(self, s):
        return O.chorin_grad(self.se, self.coord, self.tri, s) / self.vol[:, None]

    def _divergence(self, u, stab):
        return O.chorin_div(self.se, self.coord, self.tri, self.dt, self.P, self.pgrad,
                            u, stab=stab)

    def _psolve(self, b, initial=False, neubc=None):
        """pinit + psolve: Dirichlet pressure (the configured values in the
        start-up solves, increment 0 once stepping) on the configured sets
        and the hydrostat node, then CG warm-started from the previous
        solution (ConjugateGradients::init with empty x)."""
        dirbc = {p: [(1, v if initial else 0.0)] for p, v in self.pdir.items()}
        if self.hydrostat is not None and self.hydrostat not in dirbc:
            dirbc[self.hydrostat] = [(1, 0.0)]
        b = b.copy()
        normb, normr, it, conv = O.cg(self.A, self.x, b, neubc=neubc, dirbc=dirbc,
                                      apply=True, pc=self.p_pc, maxit=self.p_iter,
                                      tol=self.p_tol)
        self.pit.append(it)
        return self.x.copy()

    # ---- ChoCG control flow --------------------------------------------
    def initial(self):
        """merge ... start (ChoCG.cpp:817-839, 869-900, 1026-1323)."""
        self.BC(self.U)
        # m_np = 0: div(u), solve, project with dt = 1
        self.div = self._divergence(self.U, stab=False)
        sol = self._psolve(self.div, initial=True)
        self.U[:, :3] -= self._grad(sol)
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

    def poisson(self, pr=None, neubc=None):
        """The first pressure solve alone, as ChoCG runs it with nstep = 1
        (ChoCG.cpp:1026-1124, 1195-1225): b = div(u), replaced by pr * vol
        where the problem has a pressure rhs (pr at the nodes), plus the
        Neumann vector neubc; the solution projects u (dt = 1) and becomes
        p.  Returns the solution."""
        self.BC(self.U)
        self.div = self._divergence(self.U, stab=False)
        if pr is not None:
            self.div = np.asarray(pr, np.float64) * self.vol
        sol = self._psolve(self.div, initial=True, neubc=neubc)
        self.U[:, :3] -= self._grad(sol)
        self.BC(self.U)
        self.P = sol.copy()
        return sol

    def pressure_diag(self, dp, ap):
        """NodeDiagnostics::precompute with an analytic pressure ap
        (NodeDiagnostics.cpp:186-241, Transporter.cpp:1570-1586): L2(p),
        L2(dp), L2(p - ap), L1(p - ap) over the total volume."""
        V = self.vol
        tot = V.sum()
        e = self.P - ap
        return np.array([np.sqrt(np.sum(self.P * self.P * V) / tot),
                         np.sqrt(np.sum(dp * dp * V) / tot),
                         np.sqrt(np.sum(e * e * V) / tot), np.sum(np.abs(e) * V) / tot])

    def dt_cfl(self):
        """ChoCG::dt (1357-1413): cfl * min(L/max(|u|,1e-8), L^2/mu), times
        freezeflow once t > freezetime (then for good)."""
        u = self.U
        vel = np.sqrt(u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])
        L = np.cbrt(self.vol)
        m = np.min(L / np.maximum(vel, 1.0e-8))
        eps = np.finfo(float).eps
        if self.mu > eps:
            m = min(m, np.min(L * L / self.mu))
        m = m * self.cfl
        if self.t > self.freezetime:
            self.ff = self.freezeflow
        return m * self.ff

    def _rhs(self):
        R = O.chorin_rhs(self.cfg, self.se, self.coord, self.tri, self.P, self.U,
                         self.vgrad)
        if self.src is not None:
            R -= self.src * self.vol[:, None]
        return R

    def step(self, dt=None):
        """One time step.  With the flow frozen (freezeflow > 1 once set by
        dt_cfl) every stage keeps the velocity it started with
        (ChoCG.cpp:1550-1568): the velocity is put back after the stage's
        update, BCs and -- damp4 -- velocity gradients, which therefore see
        the updated velocity; with damp2 the last stage's divergence does
        too (corr runs inside pred there, before the velocity is put
        back)."""
        self.dt = self.dt_cfl() if dt is None else dt
        implicit = self.theta > np.finfo(float).eps
        frozen = self.ff > 1.0
        if implicit:
            self._momentum_matrix()
        udiv = None
        for stage, c in enumerate(self.rkcoef):
            R = self._rhs()
            if stage == 0:
                self.Un = self.U.copy()
            u_in = self.U.copy() if frozen else None
            last = stage + 1 == len(self.rkcoef)
            if implicit and last:
                # semi-implicit solve (ChoCG.cpp:1570-1646): A du = R, u = un + du
                if not hasattr(self, "xm"):
                    self.xm = np.zeros(3 * self.npoin)
                b = R.reshape(-1).copy()
                _, _, it, _ = O.cg(self.Am, self.xm, b, dirbc=self._momentum_dirbc(),
                                   apply=True, pc=self.m_pc, maxit=self.m_iter,
                                   tol=self.m_tol)
                self.mit.append(it)
                self.U = self.Un + self.xm.reshape(-1, 3)
            else:
                self.U = self.Un - (c * self.dt) * R / self.vol[:, None]
            self.BC(self.U)
            if self.flux == "damp4":
                self.vgrad = (O.chorin_vgrad(self.se, self.coord, self.tri, self.U)
                              / self.vol[:, None])
            elif frozen and last:
                udiv = self.U.copy()
            if frozen:
                self.U[:, :3] = u_in[:, :3]
        # projection
        div = self._divergence(self.U if udiv is None else udiv, stab=True)
        self.div = div / self.dt
        dp = self._psolve(self.div)
        self.U[:, :3] -= self.dt * self._grad(dp)
        self.BC(self.U)
        self.P += dp
        self.pgrad = self._grad(self.P)
        self.it += 1
        self.t += self.dt
        return dp

    def diag(self, dp, An=None):
        """Columns of the diagnostics file: it, t, dt, L2(p), L2(u_c), L2(dp),
        L2(du_c), then with an analytic solution An (npoin, ncomp) L2(err_c)
        and L1(err_c); sums over nodes of f^2 V (|f| V for L1), normalised by
        the total volume, L2 square-rooted (NodeDiagnostics::precompute,
        Transporter::diagnostics)."""
        V = self.vol
        tot = V.sum()
        nc = self.ncomp
        s = [np.sum(self.P * self.P * V)]
        s += [np.sum(self.U[:, c] * self.U[:, c] * V) for c in range(nc)]
        s += [np.sum(dp * dp * V)]
        d = self.U - self.Un
        s += [np.sum(d[:, c] * d[:, c] * V) for c in range(nc)]
        row = [[self.it, self.t, self.dt], np.sqrt(np.array(s) / tot)]
        if An is not None:
            e = self.U - An
            row.append(np.sqrt(np.array([np.sum(e[:, c] * e[:, c] * V)
                                         for c in range(nc)]) / tot))
            row.append(np.array([np.sum(np.abs(e[:, c]) * V) for c in range(nc)]) / tot)
        return np.concatenate(row)
