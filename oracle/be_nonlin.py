"""TEST INFRASTRUCTURE ONLY -- not part of libptd. Imported by tests/ as the reference for the gradient of
the nonlinear (D = H^3) solves.

Exact backward-Euler steps of the system the reference's nonlinear PT iteration converges to, solved by
Newton's method with a sparse direct Jacobian, with the discrete adjoint (reverse mode) and the tangent
(forward mode) of L = 1/2 * sum over inner cells (H(T) - H_obs)^2 with respect to the initial condition H0.

The discrete system (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:37-50, and the 1D / 2D nonlinear
scripts' kernels): on the inner cells
    F_i(H) = (H_i - Hold_i)/dt + sum_axes (q_up - q_low)/h = 0,
    q_f = -c_f (H_b - H_a)/h,  c_f = s^3,  s = (H_a + H_b)/2   (face f between cells a < b),
the boundary cells fixed at their initial values (Dirichlet: no kernel writes them). Faces are those of the
reference's flux arrays (transverse indices inner). With G_f = c_f (H_b - H_a)/h:
    dG_f/dH_a = (1.5 s^2 (H_b - H_a) - c_f)/h,   dG_f/dH_b = (1.5 s^2 (H_b - H_a) + c_f)/h,
    J = dF/dH = I/dt + sum_f (e_b - e_a) (dG_f/dH)^T / h   (F_a gets -G_f/h, F_b gets +G_f/h).
Adjoint of nt steps (Hold of step n is H^{n-1}):
    J_nt^T lam_nt = H^nt - H_obs,   J_n^T lam_n = lam_{n+1}/dt,
    dL/dH0 = lam_1/dt on the inner cells,
    dL/dH0_b = sum_n sum_{faces f at boundary cell b} dG_f/dH_b (lam_a - lam_b)/h  (lam = 0 on the boundary).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class BENonlin:
    """Backward-Euler steps of the D = H^3 problem on a grid of `shape` cells with cell sizes h (one per
    axis) and time step dt. `cubic_deriv` = False drops the 1.5 s^2 term of dG/dH (a planted mistake the
    tests must catch)."""

    def __init__(self, shape, h, dt, cubic_deriv=True):
        self.shape = tuple(shape)
        self.nd = len(self.shape)
        self.h = list(h)
        self.dt = float(dt)
        self.cubic_deriv = cubic_deriv
        n = self.shape
        inner = np.zeros(n, dtype=bool)
        inner[tuple(slice(1, -1) for _ in range(self.nd))] = True
        self.idx = -np.ones(n, dtype=np.int64)
        self.idx[inner] = np.arange(int(inner.sum()))
        self.idxf = self.idx.ravel(order="F")
        self.nu = int(inner.sum())
        self.order = np.argsort(self.idxf[self.idxf >= 0])
        self.faces = []
        flat = np.arange(int(np.prod(n))).reshape(n, order="F")
        for ax in range(self.nd):
            sl_a = [slice(1, -1)] * self.nd
            sl_b = [slice(1, -1)] * self.nd
            sl_a[ax] = slice(0, -1)
            sl_b[ax] = slice(1, None)
            self.faces.append((flat[tuple(sl_a)].ravel(order="F"), flat[tuple(sl_b)].ravel(order="F"), self.h[ax]))

    # ---- pieces of F and its derivatives ----------------------------------------------------------
    def _face(self, Hf, ax):
        a, b, h = self.faces[ax]
        s = (Hf[a] + Hf[b]) * 0.5
        c = s * s * s
        dH = Hf[b] - Hf[a]
        G = c * dH / h
        k = 1.5 * s * s * dH if self.cubic_deriv else 0.0
        return a, b, h, G, (k - c) / h, (k + c) / h   # G, dG/dHa, dG/dHb

    def inner_of(self, Hfull):
        return Hfull.ravel(order="F")[self.idxf >= 0][self.order]

    def scatter(self, u, like):
        out = np.array(like, dtype=np.float64, order="F", copy=True)
        of = out.reshape(-1, order="F")
        m = self.idxf >= 0
        of[m] = u[self.idxf[m]]
        return out

    def residual(self, Hfull, Hold):
        """F(H) on the inner cells (inner-cell order)."""
        Hf = Hfull.ravel(order="F")
        out = np.zeros(int(np.prod(self.shape)))
        for ax in range(self.nd):
            a, b, h, G, _, _ = self._face(Hf, ax)
            np.add.at(out, a, -G / h)
            np.add.at(out, b, G / h)
        return (self.inner_of(Hfull) - self.inner_of(Hold)) / self.dt + out[self.idxf >= 0][self.order]

    def jacobian(self, Hfull):
        """dF/dH over the inner cells (sparse, CSC) and the boundary coupling dF/dH_bd as a function."""
        Hf = Hfull.ravel(order="F")
        rows, cols, vals = [np.arange(self.nu)], [np.arange(self.nu)], [np.full(self.nu, 1.0 / self.dt)]
        for ax in range(self.nd):
            a, b, h, _, ga, gb = self._face(Hf, ax)
            ia, ib = self.idxf[a], self.idxf[b]
            # F_a: -G/h, F_b: +G/h
            for row, sign in ((ia, -1.0), (ib, 1.0)):
                for col, g in ((ia, ga), (ib, gb)):
                    m = (row >= 0) & (col >= 0)
                    rows.append(row[m])
                    cols.append(col[m])
                    vals.append(sign * g[m] / h)
        return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(self.nu, self.nu))

    def boundary_jvp(self, Hfull, vfull):
        """(dF/dH_bd) v_bd: the inner equations' response to a change v of the boundary values."""
        Hf = Hfull.ravel(order="F")
        vf = vfull.ravel(order="F")
        out = np.zeros(int(np.prod(self.shape)))
        for ax in range(self.nd):
            a, b, h, _, ga, gb = self._face(Hf, ax)
            ia, ib = self.idxf[a], self.idxf[b]
            dG = np.where(ia < 0, ga * vf[a], 0.0) + np.where(ib < 0, gb * vf[b], 0.0)
            np.add.at(out, a, -dG / h)
            np.add.at(out, b, dG / h)
        return out[self.idxf >= 0][self.order]

    # ---- solves -------------------------------------------------------------------------------------
    def step(self, Hprev, tol=1e-13, maxit=60):
        """One physical step: Newton on F(H) = 0 from H = Hprev (boundary cells kept), each step halved
        until the residual norm decreases (rough fields overshoot otherwise)."""
        H = np.array(Hprev, dtype=np.float64, order="F", copy=True)
        F = self.residual(H, Hprev)
        fn = np.sqrt(np.mean(F * F))
        for _ in range(maxit):
            if fn <= tol:
                return H
            du = spla.spsolve(self.jacobian(H), -F)
            u = self.inner_of(H)
            alpha = 1.0
            while True:
                Hn = self.scatter(u + alpha * du, H)
                Fn = self.residual(Hn, Hprev)
                fnn = np.sqrt(np.mean(Fn * Fn))
                if fnn < fn or alpha < 1e-6:
                    break
                alpha *= 0.5
            H, F, fn = Hn, Fn, fnn
        if fn <= 10 * tol:
            return H
        raise RuntimeError(f"Newton did not converge (residual {fn:.3e})")

    def forward(self, H0, nt):
        Hs = [np.asfortranarray(np.asarray(H0, dtype=np.float64))]
        for _ in range(nt):
            Hs.append(self.step(Hs[-1]))
        return Hs

    def loss(self, HN, Hobs):
        r = self.inner_of(HN) - self.inner_of(Hobs)
        return 0.5 * float(r @ r)

    def gradient(self, H0, Hobs, nt):
        """(L, dL/dH0) by the discrete adjoint."""
        Hs = self.forward(H0, nt)
        r = self.inner_of(Hs[-1]) - self.inner_of(Hobs)
        L = 0.5 * float(r @ r)
        g = np.zeros(int(np.prod(self.shape)))
        rhs = r
        lam = None
        for n in range(nt, 0, -1):
            lam = spla.spsolve(self.jacobian(Hs[n]).T.tocsc(), rhs)
            lamf = np.zeros(int(np.prod(self.shape)))
            lamf[self.idxf >= 0] = lam[self.idxf[self.idxf >= 0]]
            Hf = Hs[n].ravel(order="F")
            for ax in range(self.nd):
                a, b, h, _, ga, gb = self._face(Hf, ax)
                ia, ib = self.idxf[a], self.idxf[b]
                dl = (lamf[a] - lamf[b]) / h
                m = (ia >= 0) & (ib < 0)
                np.add.at(g, b[m], gb[m] * dl[m])
                m = (ib >= 0) & (ia < 0)
                np.add.at(g, a[m], ga[m] * dl[m])
            rhs = lam / self.dt
        m = self.idxf >= 0
        g[m] += lam[self.idxf[m]] / self.dt
        return L, g.reshape(self.shape, order="F")

    def jvp(self, H0, Hobs, nt, vH0):
        """Directional derivative of L along vH0 in forward mode (tangent linear model)."""
        Hs = self.forward(H0, nt)
        vH0 = np.asfortranarray(np.asarray(vH0, dtype=np.float64))
        dH = vH0
        for n in range(1, nt + 1):
            src = self.inner_of(dH) / self.dt - self.boundary_jvp(Hs[n], vH0)
            dH = self.scatter(spla.spsolve(self.jacobian(Hs[n]), src), vH0)
        r = self.inner_of(Hs[-1]) - self.inner_of(Hobs)
        return float(r @ self.inner_of(dH))
