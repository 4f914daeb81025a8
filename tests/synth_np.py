"""Independent numpy statement of the workload generator and the structured
mesh rule, used to pin the product's generator (bit-for-bit) and mesh builder
(against data/ex05/gpuline.mesh.data.bin, whose copy is tests/golden/ex05_gpuline.mesh.data.bin).

Generator: key = mix64(seed*G + id*H + 1); u_i = (mix64(key + (i+1)*G) >> 11) * 2^-53;
value = a + (b - a) * u (separately rounded); ids 1 jacD, 2 jacO1, 3 jacO2, 4 rhs;
jacD diagonal entries get + shift.
"""
import numpy as np

G = np.uint64(0x9E3779B97F4A7C15)
H = np.uint64(0xD1B54A32D192ED03)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed, aid):
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) * G + np.uint64(aid) * H + np.uint64(1))


def uniform(seed, aid, n, a, b, start=0):
    key = stream_key(seed, aid)
    i = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = mix64(key + (i + np.uint64(1)) * G)
    u = (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return a + (b - a) * u


def generate(seed, nvar, nelem, nintface, shift):
    bs = nvar * nvar
    D = uniform(seed, 1, bs * nelem, -1.0, 1.0)
    q = np.arange(bs * nelem) % bs
    D[q % (nvar + 1) == 0] += shift
    return dict(jacD=D, jacO1=uniform(seed, 2, bs * nintface, -0.1, 0.1),
                jacO2=uniform(seed, 3, bs * nintface, -0.1, 0.1), rhs=uniform(seed, 4, nvar * nelem, -1.0, 1.0))


def structured(NL, NK):
    """0-based mesh arrays of the NL x NK structured grid (ex05 rule)."""
    n = NL * NK
    e = np.arange(n)
    l, k = e // NK, e % NK
    has_in = k < NK - 1
    has_cr = l < NL - 1
    cnt = has_in.astype(int) + has_cr.astype(int)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    nif = int(cnt.sum())
    f_in = np.where(has_in, first, -1)
    f_cr = np.where(has_cr, first + has_in, -1)
    fc = np.zeros((nif, 2), np.int64)
    fc[f_in[has_in], 0] = e[has_in] + 1
    fc[f_in[has_in], 1] = e[has_in]
    fc[f_cr[has_cr], 0] = e[has_cr] + NK
    fc[f_cr[has_cr], 1] = e[has_cr]
    bS, bE, bN, bW = 0, NK, NK + NL, NK + NL + NK

    def bnd(b):
        return -(b + 1) - 1

    ef = np.zeros((n, 4), np.int64)
    ef[:, 0] = np.where(l > 0, f_cr[np.maximum(e - NK, 0)], bnd(bS + k))
    ef[:, 1] = np.where(k < NK - 1, f_in, bnd(bE + l))
    ef[:, 2] = np.where(l < NL - 1, f_cr, bnd(bN + k))
    ef[:, 3] = np.where(k > 0, f_in[np.maximum(e - 1, 0)], bnd(bW + l))
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
    return dict(epoint=i32(4 * np.arange(n + 1)), ef=i32(ef), fc=i32(fc), linesize=i32(np.full(NL, NK)),
                linepoint=i32(NK * np.arange(NL + 1)), lines=i32(e), lineface=i32(ef[:, 3]))


def file_image(mesh, nvar=9):
    """int32 words of gpuline.mesh.data.bin for 0-based arrays (adds the +1)."""
    nelem = mesh["epoint"].size - 1
    hdr = [nvar, nelem, mesh["fc"].size // 2, mesh["ef"].size, mesh["linesize"].size, mesh["lines"].size]
    parts = [np.array(hdr, np.int32)]
    for f in ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface"):
        parts.append(mesh[f].astype(np.int32) + (0 if f == "linesize" else 1))
    return np.concatenate(parts).astype("<i4")
