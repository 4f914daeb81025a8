"""Helpers for the GPU parity tests: inputs on the device, mesh variants."""
import numpy as np

import synth_np
from triblocksolve_amd.solver import HostMesh


_MEM = []


def mctx():
    """Context used for test allocations (HBM pointers are valid for every context on the device)."""
    if not _MEM:
        import triblocksolve_amd as tbs
        _MEM.append(tbs.Context(0, 9))
    return _MEM[0]


def dev(a):
    """Host array -> DeviceBuffer (library-allocated HBM)."""
    return mctx().array(np.ascontiguousarray(a))


def zeros(n, dtype=np.float64):
    return mctx().array(size=n, dtype=dtype)


def host(buf):
    """DeviceBuffer -> host numpy copy, after a device-wide synchronisation."""
    return buf.numpy()


def reversed_lines(mesh):
    """Same elements and faces, every line traversed backwards: exercises the
    e == fc[2f+1] branch of the A/C selection."""
    lines = mesh.lines.copy()
    lineface = mesh.lineface.copy()
    for l in range(mesh.nline):
        a, n = mesh.linepoint[l], mesh.linesize[l]
        seg = mesh.lines[a:a + n][::-1].copy()
        lines[a:a + n] = seg
        lineface[a] = -2
        for k in range(1, n):
            lineface[a + k] = mesh.lineface[a + n - k]
    return HostMesh(mesh.nvar, mesh.epoint, mesh.ef, mesh.fc, mesh.linesize, mesh.linepoint, lines, lineface)


def cross_lines(NL, NK, nvar):
    """Lines along the slow index of the NL x NK grid: line k holds e = l*NK + k, so
    lines[] is a true permutation and each line's faces are the cross faces."""
    base = synth_np.structured(NL, NK)
    n = NL * NK
    e = np.arange(n)
    lines = (e % NL) * NK + e // NL
    lines = lines.astype(np.int32)
    ef = base["ef"].reshape(n, 4)
    lineface = np.empty(n, np.int32)
    for m in range(n):
        k, l = m // NL, m % NL
        lineface[m] = ef[lines[m], 0] if l > 0 else -2   # S face joins (e, e - NK)
    return HostMesh(nvar, base["epoint"], base["ef"], base["fc"], np.full(NK, NL, np.int32),
                    (NL * np.arange(NK + 1)).astype(np.int32), lines, lineface)


def ragged_lines(NK_list, nvar):
    """Stacked lines of different lengths; element (l, k) couples to (l+1, k) when both exist."""
    NL = len(NK_list)
    start = np.concatenate([[0], np.cumsum(NK_list)]).astype(np.int64)
    n = int(start[-1])
    fcs, ef = [], [[] for _ in range(n)]
    for l, NK in enumerate(NK_list):
        for k in range(NK):
            e = start[l] + k
            if k < NK - 1:
                f = len(fcs); fcs.append((e + 1, e)); ef[e].append(f); ef[e + 1].append(f)
            if l < NL - 1 and k < NK_list[l + 1]:
                f = len(fcs); fcs.append((start[l + 1] + k, e)); ef[e].append(f); ef[start[l + 1] + k].append(f)
    epoint = np.zeros(n + 1, np.int32)
    flat = []
    for e in range(n):
        row = sorted(ef[e]) + [-2]          # one boundary entry per element
        flat.extend(row)
        epoint[e + 1] = epoint[e] + len(row)
    lineface = np.full(n, -2, np.int32)
    for (a, b), f in zip(fcs, range(len(fcs))):
        if a == b + 1 and np.searchsorted(start, a, side="right") == np.searchsorted(start, b, side="right"):
            lineface[a] = f
    return HostMesh(nvar, epoint, np.array(flat, np.int32), np.array(fcs, np.int32).ravel(),
                    np.array(NK_list, np.int32), start.astype(np.int32), np.arange(n, dtype=np.int32), lineface)
