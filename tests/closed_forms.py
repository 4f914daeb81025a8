"""Independent closed forms used to pin the oracle (and, through it, the HIP path).

J  : the full block Jacobian assembled from (jacD, jacO1, jacO2, fc, ef) with scipy.
M  : its line block-tridiagonal part (diagonal + in-line neighbour blocks).
LIJ: U_{k+1} = U_k - M^{-1} (rhs + J U_k),  h_k = ||rhs + J U_k||_2, U_0 = 0.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _blk(arr, nv, b):
    return arr[nv * nv * b: nv * nv * (b + 1)].reshape(nv, nv, order="F")


def assemble_J(m, jac):
    nv, n = m.nvar, m.nelem
    rows, cols, vals = [], [], []
    ii, jj = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")

    def put(r, c, B):
        rows.append((nv * r + ii).ravel())
        cols.append((nv * c + jj).ravel())
        vals.append(B.ravel())

    for e in range(n):
        put(e, e, _blk(jac["jacD"], nv, e))
        for k2 in range(m.epoint[e], m.epoint[e + 1]):
            f = m.ef[k2]
            if f < 0:
                continue
            e1, e2 = m.fc[2 * f], m.fc[2 * f + 1]
            if e == e1:
                put(e, e2, _blk(jac["jacO2"], nv, f))
            else:
                put(e, e1, _blk(jac["jacO1"], nv, f))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(nv * n, nv * n))


def assemble_M(m, jac):
    nv, n = m.nvar, m.nelem
    rows, cols, vals = [], [], []
    ii, jj = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")

    def put(r, c, B):
        rows.append((nv * r + ii).ravel())
        cols.append((nv * c + jj).ravel())
        vals.append(B.ravel())

    for e in range(n):
        put(e, e, _blk(jac["jacD"], nv, e))
    for l in range(m.nline):
        for k in range(1, m.linesize[l]):
            mm = m.linepoint[l] + k
            e, ep, f = m.lines[mm], m.lines[mm - 1], m.lineface[mm]
            e1 = m.fc[2 * f]
            Ablk = _blk(jac["jacO2"] if e == e1 else jac["jacO1"], nv, f)
            Cblk = _blk(jac["jacO1"] if e == e1 else jac["jacO2"], nv, f)
            put(e, ep, Ablk)
            put(ep, e, Cblk)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(nv * n, nv * n))


def lij_closed_form(m, jac, iters):
    J = assemble_J(m, jac).tocsc()
    M = assemble_M(m, jac).tocsc()
    lu = spla.splu(M)
    rhs = jac["rhs"]
    U = np.zeros_like(rhs)
    R = rhs.copy()
    hist = [np.linalg.norm(R)]
    for _ in range(iters):
        U = U - lu.solve(R)
        R = rhs + J @ U
        hist.append(np.linalg.norm(R))
    return np.array(hist), U, R


def thomas_scalar(a, b, c, d):
    """Textbook scalar Thomas algorithm for a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i."""
    n = len(b)
    cp, dp = np.zeros(n), np.zeros(n)
    cp[0], dp[0] = c[0] / b[0], d[0] / b[0]
    for i in range(1, n):
        den = b[i] - a[i] * cp[i - 1]
        cp[i] = c[i] / den if i < n - 1 else 0.0
        dp[i] = (d[i] - a[i] * dp[i - 1]) / den
    x = np.zeros(n)
    x[-1] = dp[-1]
    for i in range(n - 2, -1, -1):
        x[i] = dp[i] - cp[i] * x[i + 1]
    return x


def history_close(h_test, h_ref, rel=1e-12, floor=1e-4):
    """Normalised-history parity: |h_k/h_0 - h'_k/h'_0| <= rel * max(h'_k/h'_0, floor)."""
    a = np.asarray(h_test) / h_test[0]
    b = np.asarray(h_ref) / h_ref[0]
    return np.all(np.abs(a - b) <= rel * np.maximum(b, floor)), np.max(np.abs(a - b) / np.maximum(b, floor))
