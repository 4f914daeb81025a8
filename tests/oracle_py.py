"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: a plain restatement of the reference's
active path (see oracle/oracle.cpp).  Arrays are numpy, reference layout.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")


def _p(a):
    return C.c_void_p(None) if a is None else C.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        self.lib = C.CDLL(path)
        self.lib.orc_norm2.restype = C.c_double
        self.lib.orc_norm2.argtypes = [C.c_long, C.c_void_p]
        self.lib.orc_max_threads.restype = C.c_int

    # low-level kernels ----------------------------------------------------
    def LU(self, nv, A):
        A = np.array(A, dtype=np.float64, order="F").ravel(order="F").copy()
        self.lib.orc_LU(nv, _p(A))
        return A.reshape(nv, nv, order="F")

    def solveLU(self, nv, LUf, b):
        LUf = np.asfortranarray(LUf, dtype=np.float64).ravel(order="F").copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(nv)
        self.lib.orc_solveLU(nv, _p(LUf), _p(b), _p(x))
        return x

    def dgemvLU(self, nv, LU, x, nthreads=1):
        """b = L U x per block (LU: packed factors, blocks by element)."""
        LU = np.ascontiguousarray(LU, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.size // nv
        b = np.zeros(nv * n)
        self.lib.orc_dgemvLU(nv, n, _p(LU), _p(x), _p(b), nthreads)
        return b

    def assembleTriBlocks(self, m, jacO1, jacO2, nthreads=1):
        Aout = np.zeros(m.nvar * m.nvar * m.nelem)
        self.lib.orc_assembleTriBlocks(m.nvar, m.nline, _p(m.linesize), _p(m.linepoint), _p(m.lines),
                                       _p(m.lineface), _p(m.fc), _p(jacO1), _p(jacO2), _p(Aout))
        return Aout

    def lineLU(self, m, jacD, jacO1, jacO2, nthreads=1):
        Dia = jacD.copy()
        DinvC = np.zeros_like(jacD)
        self.lib.orc_lineLU(m.nvar, m.nelem, m.nintface, m.nline, _p(m.fc), _p(m.linesize), _p(m.linepoint),
                            _p(m.lines), _p(m.lineface), _p(Dia), _p(jacO1), _p(jacO2), _p(DinvC), nthreads)
        return Dia, DinvC

    def solveDU(self, m, Dia, DinvC, Amat, R, nthreads=1):
        dU = np.zeros(m.nvar * m.nelem)
        self.lib.orc_solveDU(m.nvar, m.nelem, m.nintface, m.eftot, m.nline, m.nlineelem, _p(m.linesize),
                             _p(m.linepoint), _p(m.lines), _p(Dia), _p(DinvC), _p(Amat), _p(dU), _p(R), nthreads)
        return dU

    def lineRes(self, m, jacD, jacO1, jacO2, U, R, nthreads=1):
        R = R.copy()
        self.lib.orc_lineRes(m.nvar, m.nelem, m.nintface, m.eftot, m.nline, m.nlineelem, _p(m.epoint),
                             _p(m.ef), _p(m.fc), _p(m.linesize), _p(m.linepoint), _p(m.lines), _p(jacD),
                             _p(jacO1), _p(jacO2), _p(U), _p(R), nthreads)
        return R

    def norm2(self, v):
        return self.lib.orc_norm2(v.size, _p(v))

    def lij_run(self, m, jac, iters, nthreads=1, norms=True):
        """Reference driver loop; returns dict(U, R, hist, dU, Dia, DinvC, A, t_factor, t_sweeps,
        t_each).  norms=False skips the history (t_each then times the sweeps alone)."""
        nv, n = m.nvar, m.nelem
        bs = nv * nv
        out = dict(jacDLU=np.zeros(bs * n), DinvC=np.zeros(bs * n), A=np.zeros(bs * n),
                   U=np.zeros(nv * n), dU=np.zeros(nv * n), R=np.zeros(nv * n), hist=np.zeros(iters + 1),
                   t_each=np.zeros(max(iters, 1)))
        tf, ts = C.c_double(), C.c_double()
        self.lib.orc_lij_run(nv, n, m.nintface, m.eftot, m.nline, m.nlineelem, _p(m.epoint), _p(m.ef),
                             _p(m.fc), _p(m.linesize), _p(m.linepoint), _p(m.lines), _p(m.lineface),
                             _p(jac["jacD"]), _p(jac["jacO1"]), _p(jac["jacO2"]), _p(jac["rhs"]),
                             _p(out["jacDLU"]), _p(out["DinvC"]), _p(out["A"]), _p(out["U"]), _p(out["dU"]),
                             _p(out["R"]), iters, _p(out["hist"]) if norms else None, nthreads, C.byref(tf),
                             C.byref(ts), _p(out["t_each"]))
        out["t_factor"], out["t_sweeps"] = tf.value, ts.value
        return out
