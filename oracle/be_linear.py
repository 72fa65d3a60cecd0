"""TEST INFRASTRUCTURE ONLY -- not part of libptd. Imported by tests/ and bench.py's checks as the
reference for the gradient of the linear / step-function solves.

Exact (sparse direct) backward-Euler solves of the system the reference's PT iteration converges to
in the lin / step cases, with its tangent (forward mode) and its adjoint (reverse mode) for
L = 1/2 * sum over inner cells (H(T) - H_obs)^2.

The discrete system (scripts/diff_3D/diff_3D_linstep_multixpu.jl:29-47): on the inner cells
    (H^n - H^{n-1})/dt + div q = 0,   q_x = -av_x(D) * d_x(H) / dx  (av_x(D) = (D[i] + D[i+1]) * 0.5),
i.e. A(D) H^n = H^{n-1}/dt with A = I/dt - div(av(D) grad .), the boundary cells fixed at H0 (Dirichlet:
no kernel writes them). Faces exist where the reference's flux arrays have them: an x face between
cells i and i+1 for every inner row (j, k). A is symmetric, so the adjoint system A^T mu = r is A mu = r.

    forward    H^n = A^{-1} (H^{n-1}/dt + boundary coupling)
    tangent    A dH^n = dH^{n-1}/dt + div(dD_f grad H^n) + boundary coupling of dH0 on the boundary
    adjoint    mu_N = A^{-1} (H^N - H_obs),  mu_n = A^{-1} (mu_{n+1}/dt)
    gradient   dL/dD_f = sum_n (H^n_b - H^n_a)(mu_n,a - mu_n,b)/h^2,  dL/dD_c = sum_{f at c} dL/dD_f / 2
               dL/dH0  = mu_1/dt (inner),  sum_n sum_f mu_n,i D_f/h^2 (boundary cell next to inner i)
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class BELinear:
    """Backward-Euler operator of one lin / step problem: D (cell field, column-major n-D array), cell
    sizes h (one per axis), time step dt."""

    def __init__(self, D, h, dt):
        D = np.asarray(D, dtype=np.float64)
        self.shape = D.shape
        self.nd = D.ndim
        self.h = list(h)
        self.dt = float(dt)
        self.D = D
        n = self.shape
        inner = np.zeros(n, dtype=bool)
        inner[tuple(slice(1, -1) for _ in range(self.nd))] = True
        self.inner = inner
        self.idx = -np.ones(n, dtype=np.int64)
        self.idx[inner] = np.arange(int(inner.sum()))
        self.nu = int(inner.sum())
        # faces: for each axis, the pairs (a, b) of flat (F-order) cell indices where the flux array has a face
        self.faces = []
        flat = np.arange(int(np.prod(n))).reshape(n, order="F")
        for ax in range(self.nd):
            sl_a = [slice(1, -1)] * self.nd
            sl_b = [slice(1, -1)] * self.nd
            sl_a[ax] = slice(0, -1)
            sl_b[ax] = slice(1, None)
            a = flat[tuple(sl_a)].ravel(order="F")
            b = flat[tuple(sl_b)].ravel(order="F")
            self.faces.append((a, b, 1.0 / (self.h[ax] * self.h[ax])))
        self._assemble()

    def face_d(self, D, ax):
        a, b, _ = self.faces[ax]
        Df = D.ravel(order="F")
        return (Df[a] + Df[b]) * 0.5

    def _assemble(self):
        rows, cols, vals = [], [], []
        diag = np.full(self.nu, 1.0 / self.dt)
        idxf = self.idx.ravel(order="F")
        for ax in range(self.nd):
            a, b, w = self.faces[ax]
            c = self.face_d(self.D, ax) * w
            ia, ib = idxf[a], idxf[b]
            for p, q in ((ia, ib), (ib, ia)):
                m = p >= 0
                np.add.at(diag, p[m], c[m])
                mm = m & (q >= 0)
                rows.append(p[mm])
                cols.append(q[mm])
                vals.append(-c[mm])
        rows.append(np.arange(self.nu))
        cols.append(np.arange(self.nu))
        vals.append(diag)
        self.A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(self.nu, self.nu))
        self.lu = spla.splu(self.A)

    def _boundary_rhs(self, Hfull, D=None):
        """Coupling of the boundary values of Hfull into the inner equations: sum D_f H_b / h^2."""
        D = self.D if D is None else D
        out = np.zeros(self.nu)
        idxf = self.idx.ravel(order="F")
        Hf = Hfull.ravel(order="F")
        for ax in range(self.nd):
            a, b, w = self.faces[ax]
            c = self.face_d(D, ax) * w
            ia, ib = idxf[a], idxf[b]
            m = (ia >= 0) & (ib < 0)
            np.add.at(out, ia[m], c[m] * Hf[b[m]])
            m = (ib >= 0) & (ia < 0)
            np.add.at(out, ib[m], c[m] * Hf[a[m]])
        return out

    def _inner(self, Hfull):
        return Hfull.ravel(order="F")[self.idx.ravel(order="F") >= 0][np.argsort(self.idx.ravel(order="F")[self.idx.ravel(order="F") >= 0])]

    def _scatter(self, u, like):
        out = np.array(like, dtype=np.float64, order="F", copy=True)
        of = out.reshape(-1, order="F")
        idxf = self.idx.ravel(order="F")
        m = idxf >= 0
        of[m] = u[idxf[m]]
        return out

    def step(self, Hprev):
        """One physical step: the converged solution of the PT solve with Hold = Hprev, H boundary = Hprev's."""
        rhs = self._inner(Hprev) / self.dt + self._boundary_rhs(Hprev)
        return self._scatter(self.lu.solve(rhs), Hprev)

    def forward(self, H0, nt):
        Hs = [np.asfortranarray(np.asarray(H0, dtype=np.float64))]
        for _ in range(nt):
            Hs.append(self.step(Hs[-1]))
        return Hs

    def loss(self, HN, Hobs):
        r = self._inner(HN) - self._inner(Hobs)
        return 0.5 * float(r @ r)

    def gradient(self, H0, Hobs, nt):
        """(L, dL/dD, dL/dH0) by the discrete adjoint."""
        Hs = self.forward(H0, nt)
        r = self._inner(Hs[-1]) - self._inner(Hobs)
        L = 0.5 * float(r @ r)
        gD = np.zeros(int(np.prod(self.shape)))
        gH0 = np.zeros(int(np.prod(self.shape)))
        idxf = self.idx.ravel(order="F")
        rhs = r
        mu1 = None
        for n in range(nt, 0, -1):
            mu = self.lu.solve(rhs)                      # A^T mu = rhs, A symmetric
            mufull = np.zeros(int(np.prod(self.shape)))
            mufull[idxf >= 0] = mu[idxf[idxf >= 0]]
            Hn = Hs[n].ravel(order="F")
            for ax in range(self.nd):
                a, b, w = self.faces[ax]
                gf = (Hn[b] - Hn[a]) * (mufull[a] - mufull[b]) * w
                np.add.at(gD, a, 0.5 * gf)
                np.add.at(gD, b, 0.5 * gf)
                c = self.face_d(self.D, ax) * w
                ia, ib = idxf[a], idxf[b]
                m = (ia >= 0) & (ib < 0)
                np.add.at(gH0, b[m], mufull[a[m]] * c[m])
                m = (ib >= 0) & (ia < 0)
                np.add.at(gH0, a[m], mufull[b[m]] * c[m])
            rhs = mu / self.dt
            mu1 = mufull
        gH0[idxf >= 0] += mu1[idxf >= 0] / self.dt
        return L, gD.reshape(self.shape, order="F"), gH0.reshape(self.shape, order="F")

    def jvp(self, H0, Hobs, nt, vD, vH0):
        """Directional derivative of L along (vD, vH0) in forward mode (tangent linear model)."""
        Hs = self.forward(H0, nt)
        vD = np.asarray(vD, dtype=np.float64)
        vH0 = np.asfortranarray(np.asarray(vH0, dtype=np.float64))
        dH = vH0
        idxf = self.idx.ravel(order="F")
        for n in range(1, nt + 1):
            Hn = Hs[n].ravel(order="F")
            src = self._inner(dH) / self.dt + self._boundary_rhs(vH0)
            extra = np.zeros(int(np.prod(self.shape)))
            for ax in range(self.nd):
                a, b, w = self.faces[ax]
                dc = self.face_d(vD, ax) * w
                g = dc * (Hn[b] - Hn[a])           # div(dD_f grad H): +g at a, -g at b
                np.add.at(extra, a, g)
                np.add.at(extra, b, -g)
            src = src + extra[idxf >= 0][np.argsort(idxf[idxf >= 0])]
            dH = self._scatter(self.lu.solve(src), vH0)
        r = self._inner(Hs[-1]) - self._inner(Hobs)
        return float(r @ self._inner(dH))
