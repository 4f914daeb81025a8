"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

CPU restatement of line-preconditioned flexible GMRES(m), the Krylov solver
the reference's paper wraps around the LIJ smoother (kirby-mavriplis-2025,
Table 6; the reference repository ships no GMRES code, so the algorithm is
the published one: Saad, "A flexible inner-outer preconditioned GMRES
algorithm", SIAM J. Sci. Comput. 14 (1993), right preconditioning, Arnoldi
with classical Gram-Schmidt applied twice, Givens rotations on the
Hessenberg matrix).

Operator and preconditioner come from the C++ oracle (oracle.cpp):
  J z      = lineRes with R preset to 0 (src/okl/linesmoothLU_TriBlock_v3.okl:228-336)
  P v      = `sweeps` LIJ sweeps on J z = v from z = 0 (solveDU + lineRes,
             the reference loop of src/main.cxx:197-229 with rhs = -v).
Solves J U = -rhs.  Only tests/ use this module, as the checker.

Pins (tests/test_gmres.py): the converged U equals a sparse direct solve of
J U = -rhs; with no off-line coupling the preconditioner is exact and one
iteration converges; the final estimate equals the true residual; it
converges where plain LIJ diverges.
"""
import numpy as np


class LinePreconditionedGmres:
    def __init__(self, oracle, mesh, jac):
        self.o, self.m, self.jac = oracle, mesh, jac
        self.Dia, self.DinvC = oracle.lineLU(mesh, jac["jacD"], jac["jacO1"], jac["jacO2"])
        self.A = oracle.assembleTriBlocks(mesh, jac["jacO1"], jac["jacO2"])
        self.N = mesh.nvar * mesh.nelem

    def residual(self, rhs, U):
        """rhs + J U."""
        return self.o.lineRes(self.m, self.jac["jacD"], self.jac["jacO1"], self.jac["jacO2"], U, rhs)

    def matvec(self, z):
        return self.residual(np.zeros(self.N), z)

    def precondition(self, v, sweeps):
        z = np.zeros(self.N)
        R = -v.copy()
        for s in range(sweeps):
            z = z + self.o.solveDU(self.m, self.Dia, self.DinvC, self.A, R)
            if s + 1 < sweeps:
                R = self.residual(-v, z)
        return z

    def solve(self, rhs, U0=None, restart=20, max_iters=100, rtol=1e-10, sweeps=1):
        x = np.zeros(self.N) if U0 is None else U0.copy()
        hist, it, beta0 = [], 0, None
        while True:
            R = self.residual(rhs, x)
            beta = np.linalg.norm(R)
            if beta0 is None:
                beta0 = beta
                hist.append(beta)
            if beta <= rtol * beta0 or beta == 0.0 or it >= max_iters:
                break
            V = [-R / beta]
            Z = []
            H = np.zeros((restart + 1, restart))
            cs, sn = np.zeros(restart), np.zeros(restart)
            g = np.zeros(restart + 1)
            g[0] = beta
            k = 0
            for j in range(restart):
                if it >= max_iters:
                    break
                z = self.precondition(V[j], sweeps)
                Z.append(z)
                w = self.matvec(z)
                Vm = np.array(V)
                h = Vm @ w
                w = w - Vm.T @ h
                h2 = Vm @ w
                w = w - Vm.T @ h2
                h = h + h2
                hn = np.linalg.norm(w)
                V.append(w / hn if hn > 0 else w)
                H[: j + 1, j] = h
                H[j + 1, j] = hn
                for i in range(j):
                    a, b = H[i, j], H[i + 1, j]
                    H[i, j] = cs[i] * a + sn[i] * b
                    H[i + 1, j] = -sn[i] * a + cs[i] * b
                a, b = H[j, j], H[j + 1, j]
                r = np.hypot(a, b)
                cs[j], sn[j] = (a / r, b / r) if r > 0 else (1.0, 0.0)
                H[j, j], H[j + 1, j] = r, 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                it += 1
                k = j + 1
                hist.append(abs(g[j + 1]))
                if abs(g[j + 1]) <= rtol * beta0 or hn == 0.0:
                    break
            y = np.zeros(k)
            for i in range(k - 1, -1, -1):
                y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
            for i in range(k):
                x = x + y[i] * Z[i]
        return x, np.array(hist), it
