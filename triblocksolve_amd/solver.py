"""Host-side helpers over the C ABI: meshes, synthetic systems, contexts, plans.

Naming follows the reference host API (src/Mesh.cxx, src/Jacobian.cxx,
src/main.cxx): meshes carry epoint/ef/fc/linesize/linepoint/lines/lineface
(0-based), Jacobians jacD/jacO1/jacO2/rhs in the reference block layout.
Device arrays are DeviceBuffer objects (HBM allocated through the library's
own C ABI); all arithmetic happens in libtriblock_hip.so.
"""
import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _capi as A


def factor_plans(ctx, plans, jacD, jacO1, jacO2):
    """Factorise several plans of one context in one launch over all their lines
    (tbs_plans_factor); jacD[i], jacO1[i], jacO2[i] belong to plans[i].  Returns
    after the device has finished reading the Jacobians."""
    n = len(plans)
    P = C.c_void_p * n
    hs = P(*[p.handle for p in plans])
    d, o1, o2 = (P(*[_ptr(x).value for x in arrs]) for arrs in (jacD, jacO1, jacO2))
    A.check(ctx._lib.tbs_plans_factor(ctx.handle, hs, n, d, o1, o2))
    ctx.sync()


def variant_name(code):
    """Name of a Thomas variant from its code (100 * prefetch windows + ring
    cycles + 10000 * (lines per workgroup - 1)); the lines per workgroup are
    named when more than one."""
    code = int(code)
    name = f"pf{(code % 10000) // 100}_k{code % 100}"
    return name + (f"_l{code // 10000 + 1}" if code >= 10000 else "")


def _tuning_dict(t):
    return {"thomas_variant": t.thomas_variant, "thomas_prefetch": t.thomas_prefetch,
            "thomas_lds_windows": t.thomas_lds_windows, "thomas_lines_per_wg": t.thomas_lines_per_wg,
            "residual_variant": t.residual_variant, "residual_waves_per_wg": t.residual_waves_per_wg,
            "two_stream": bool(t.two_stream),
            "thomas_us": [float(x) for x in t.thomas_us[: t.n_thomas]],
            "residual_us": [float(x) for x in t.residual_us[: t.n_residual]],
            "residual_waves_per_wg_by_variant": [int(x) for x in t.residual_shapes[: t.n_residual]],
            "pass_us": [float(x) for x in t.pass_us]}


def _set_tuning(fn_get, fn_set, handle, kw):
    t = A.Tuning()
    A.check(fn_get(handle, C.byref(t)))
    for k, v in kw.items():
        if k not in ("thomas_variant", "residual_variant", "two_stream"):
            raise KeyError(f"{k} is not settable")
        setattr(t, k, int(v))
    A.check(fn_set(handle, C.byref(t)))


def _ptr(x):
    """c_void_p of a numpy array, a DeviceBuffer or an int address (None -> NULL)."""
    if x is None:
        return C.c_void_p(None)
    if isinstance(x, DeviceBuffer):
        return C.c_void_p(x.ptr)
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return C.c_void_p(x.ctypes.data)
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(x.data_ptr())
    if isinstance(x, int):
        return C.c_void_p(x)
    raise TypeError(f"cannot take a pointer of {type(x)}")


class DeviceBuffer:
    """HBM allocated through the library (tbs_malloc), freed on close/GC.

    Typed view for host transfers: `dtype` and `size` (elements)."""

    def __init__(self, ctx, size, dtype=np.float64, data=None, zero=False):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        A.check(ctx._lib.tbs_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        ctx._children.add(self)
        if data is not None:
            self.upload(data)
        elif zero:
            self.zero()

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"size {a.size} != buffer size {self.size}")
        A.check(self.ctx._lib.tbs_memcpy_h2d(self.ctx.handle, C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data),
                                             self.nbytes))

    def numpy(self, count=None, offset=0):
        count = self.size - offset if count is None else count
        out = np.empty(count, self.dtype)
        A.check(self.ctx._lib.tbs_device_synchronize(self.ctx.handle))
        A.check(self.ctx._lib.tbs_memcpy_d2h(self.ctx.handle, C.c_void_p(out.ctypes.data),
                                             C.c_void_p(self.ptr + offset * self.dtype.itemsize),
                                             count * self.dtype.itemsize))
        return out

    def zero(self):
        """Zero the buffer; complete on return, so work on any stream sees it."""
        A.check(self.ctx._lib.tbs_memset(self.ctx.handle, C.c_void_p(self.ptr), 0, self.nbytes))
        A.check(self.ctx._lib.tbs_ctx_synchronize(self.ctx.handle))

    def close(self):
        if self.ptr and self.ctx.handle:
            self.ctx._lib.tbs_free(self.ctx.handle, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MESH_FIELDS = ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface")


@dataclass
class HostMesh:
    """Mesh connectivity as host int32 arrays (the Mesh class of src/Mesh.cxx)."""
    nvar: int
    epoint: np.ndarray
    ef: np.ndarray
    fc: np.ndarray
    linesize: np.ndarray
    linepoint: np.ndarray
    lines: np.ndarray
    lineface: np.ndarray

    @property
    def nelem(self):
        return int(self.epoint.size - 1)

    @property
    def nintface(self):
        return int(self.fc.size // 2)

    @property
    def eftot(self):
        return int(self.ef.size)

    @property
    def nline(self):
        return int(self.linesize.size)

    @property
    def nlineelem(self):
        return int(self.lines.size)

    @property
    def max_line_nelem(self):
        return int(self.linesize.max()) if self.linesize.size else 0

    def sizes(self):
        return [self.nvar, self.nelem, self.nintface, self.eftot, self.nline, self.nlineelem]

    def as_struct(self):
        for f in MESH_FIELDS:
            a = getattr(self, f)
            if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
                setattr(self, f, np.ascontiguousarray(a, dtype=np.int32))
        s = A.MeshHost(self.nvar, self.nelem, self.nintface, self.eftot, self.nline, self.nlineelem,
                       *[_ptr(getattr(self, f)) for f in MESH_FIELDS])
        s._keep = self  # keep arrays alive
        return s

    def to_device(self, ctx):
        """Device copies of the connectivity arrays (dict of DeviceBuffer, int32)."""
        return {f: DeviceBuffer(ctx, getattr(self, f).size, np.int32, data=getattr(self, f)) for f in MESH_FIELDS}


def structured_mesh(NL, NK, nvar=9):
    """NL lines x NK elements structured grid (the ex05 topology rule)."""
    lib = A.load()
    s = (C.c_int * 5)()
    A.check(lib.tbs_structured_sizes(NL, NK, *[C.byref(s, 4 * i) for i in range(5)]))
    nelem, nif, eftot, nline, nle = list(s)
    arrs = dict(epoint=np.empty(nelem + 1, np.int32), ef=np.empty(eftot, np.int32),
                fc=np.empty(2 * nif, np.int32), linesize=np.empty(nline, np.int32),
                linepoint=np.empty(nline + 1, np.int32), lines=np.empty(nle, np.int32),
                lineface=np.empty(nle, np.int32))
    A.check(lib.tbs_structured_mesh(NL, NK, *[_ptr(arrs[f]) for f in MESH_FIELDS]))
    return HostMesh(nvar=nvar, **arrs)


StructuredMesh = structured_mesh


def ragged_mesh(lens, nvar=9):
    """Lines of the given lengths stacked like the structured grid."""
    lens = np.ascontiguousarray(lens, dtype=np.int32)
    lib = A.load()
    s = (C.c_int * 3)()
    A.check(lib.tbs_ragged_sizes(lens.size, _ptr(lens), C.byref(s, 0), C.byref(s, 4), C.byref(s, 8)))
    nelem, nif, eftot = list(s)
    nline = lens.size
    arrs = dict(epoint=np.empty(nelem + 1, np.int32), ef=np.empty(eftot, np.int32),
                fc=np.empty(2 * nif, np.int32), linesize=np.empty(nline, np.int32),
                linepoint=np.empty(nline + 1, np.int32), lines=np.empty(nelem, np.int32),
                lineface=np.empty(nelem, np.int32))
    A.check(lib.tbs_ragged_mesh(nline, _ptr(lens), *[_ptr(arrs[f]) for f in MESH_FIELDS]))
    return HostMesh(nvar=nvar, **arrs)


def batch_shapes(seed, nsys=64, total=384000, lmin=96, lmax=1280, nlmin=8, nlmax=40):
    """Seeded line lengths per system (list of int arrays)."""
    lib = A.load()
    nl = np.zeros(nsys, np.int32)
    cap = nsys * nlmax
    lens = np.zeros(cap, np.int32)
    cnt = C.c_int()
    A.check(lib.tbs_batch_shapes(seed, nsys, total, lmin, lmax, nlmin, nlmax, _ptr(nl), _ptr(lens), cap,
                                 C.byref(cnt)))
    out, o = [], 0
    for s in range(nsys):
        out.append(lens[o:o + nl[s]].copy())
        o += nl[s]
    return out


def generate_host(seed, nvar, nelem, nintface, shift):
    """Seeded synthetic Jacobian on the host (product generator)."""
    bs = nvar * nvar
    out = dict(jacD=np.empty(bs * nelem), jacO1=np.empty(bs * nintface),
               jacO2=np.empty(bs * nintface), rhs=np.empty(nvar * nelem))
    A.check(A.load().tbs_generate_host(seed, nvar, nelem, nintface, float(shift),
                                       *[_ptr(out[k]) for k in ("jacD", "jacO1", "jacO2", "rhs")]))
    return out


def read_mesh(path):
    lib = A.load()
    h = np.empty(6, np.int32)
    A.check(lib.tbs_read_mesh_header(str(path).encode(), _ptr(h)))
    nvar, nelem, nif, eftot, nline, nle = [int(v) for v in h]
    arrs = dict(epoint=np.empty(nelem + 1, np.int32), ef=np.empty(eftot, np.int32),
                fc=np.empty(2 * nif, np.int32), linesize=np.empty(nline, np.int32),
                linepoint=np.empty(nline + 1, np.int32), lines=np.empty(nle, np.int32),
                lineface=np.empty(nle, np.int32))
    A.check(lib.tbs_read_mesh(str(path).encode(), *[_ptr(arrs[f]) for f in MESH_FIELDS]))
    return HostMesh(nvar=nvar, **arrs)


def write_mesh(path, mesh):
    h = np.array(mesh.sizes(), np.int32)
    mesh.as_struct()
    A.check(A.load().tbs_write_mesh(str(path).encode(), _ptr(h), *[_ptr(getattr(mesh, f)) for f in MESH_FIELDS]))


def read_jacobian(path):
    lib = A.load()
    h = np.empty(3, np.int32)
    A.check(lib.tbs_read_jacobian_header(str(path).encode(), _ptr(h)))
    nvar, nelem, nif = [int(v) for v in h]
    bs = nvar * nvar
    out = dict(jacD=np.empty(bs * nelem), jacO1=np.empty(bs * nif), jacO2=np.empty(bs * nif),
               rhs=np.empty(nvar * nelem), U0=np.empty(nvar * nelem))
    A.check(lib.tbs_read_jacobian(str(path).encode(), *[_ptr(out[k]) for k in ("jacD", "jacO1", "jacO2", "rhs", "U0")]))
    return nvar, nelem, nif, out


def write_jacobian(path, nvar, nelem, nintface, jac, U0=None):
    A.check(A.load().tbs_write_jacobian(str(path).encode(), nvar, nelem, nintface,
                                        *[_ptr(np.ascontiguousarray(jac[k])) for k in ("jacD", "jacO1", "jacO2", "rhs")],
                                        _ptr(U0)))


def model_bytes(mesh, factor_bytes=8):
    """Algorithmic HBM bytes of one LIJ sweep: D', A, DinvC (factor precision),
    D, Of1, Of2 (FP64) and six FP64 vector passes."""
    bs = mesh.nvar * mesh.nvar
    return (factor_bytes * bs * (3 * mesh.nelem - 2 * mesh.nline) + 8 * bs * (mesh.nelem + 2 * mesh.nintface)
            + 6 * 8 * mesh.nvar * mesh.nelem)


def reference_model_bytes(mesh):
    """The reference's own per-sweep byte model, src/main.cxx:149-175 divided by iters."""
    nv, bs = mesh.nvar, mesh.nvar * mesh.nvar
    du = 8 * (bs * mesh.nelem + bs * mesh.nelem + bs * mesh.nlineelem + 4 * nv * mesh.nelem + nv * mesh.nelem) \
        + 4 * (mesh.nline + mesh.nline + 1 + 2 * mesh.nlineelem)
    cp = 5 * nv * mesh.nelem * 8
    lr = 8 * (bs * mesh.nelem + 2 * bs * mesh.nintface + 5 * mesh.nelem * nv + 2 * nv * mesh.nelem) \
        + 4 * (mesh.nelem + 1 + mesh.eftot + 2 * mesh.nintface + mesh.nline + mesh.nline + 1 + 2 * mesh.nlineelem)
    return du + cp + lr


class Context:
    """Device context (the Platform of include/Platform.hxx): device id, block size, stream."""

    def __init__(self, device=0, nvar=9):
        self._lib = A.load()
        h = C.c_void_p()
        A.check(self._lib.tbs_ctx_create(device, nvar, C.byref(h)))
        self.handle = h
        self.nvar = nvar
        self.device = device
        self._children = weakref.WeakSet()   # plans, batches, buffers: closed before the context

    def close(self):
        if self.handle:
            for ch in sorted(list(self._children), key=lambda o: 0 if isinstance(o, (Batch, Server)) else
                             (1 if isinstance(o, Plan) else 2)):
                ch.close()
            self._lib.tbs_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        A.check(self._lib.tbs_ctx_synchronize(self.handle))

    def stream(self):
        return self._lib.tbs_ctx_stream(self.handle)

    def array(self, data=None, size=None, dtype=np.float64, zero=True):
        """DeviceBuffer holding `data` (host array) or `size` zeroed elements."""
        if data is not None:
            data = np.ascontiguousarray(data)
            return DeviceBuffer(self, data.size, data.dtype, data=data)
        return DeviceBuffer(self, size, dtype, zero=zero)

    def event(self):
        e = C.c_void_p()
        A.check(self._lib.tbs_event_create(self.handle, C.byref(e)))
        return e

    def record(self, ev):
        A.check(self._lib.tbs_event_record(self.handle, ev))

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        A.check(self._lib.tbs_event_elapsed_ms(a, b, C.byref(ms)))
        return ms.value

    def destroy_event(self, ev):
        self._lib.tbs_event_destroy(self.handle, ev)

    def stream_ceiling(self, nbytes=2 << 30, reps=10):
        """Achievable HBM bandwidth (GB/s) of a read-only stream and of a copy."""
        out = np.zeros(2)
        A.check(self._lib.tbs_stream_ceiling(self.handle, int(nbytes), reps, _ptr(out)))
        return {"read_gbps": float(out[0]), "copy_gbps": float(out[1])}

    # --- drop-in kernels (device tensors) ---
    def copyAtoBjac(self, nelem, a, b):
        A.check(self._lib.tbs_copyAtoBjac(self.handle, nelem, _ptr(a), _ptr(b)))

    def copyAtoB(self, nelem, a, b):
        A.check(self._lib.tbs_copyAtoB(self.handle, nelem, _ptr(a), _ptr(b)))

    def addAtoB(self, nelem, a, b):
        A.check(self._lib.tbs_addAtoB(self.handle, nelem, _ptr(a), _ptr(b)))

    def assembleTriBlocks(self, m, jacO1, jacO2, Aout):
        A.check(self._lib.tbs_assembleTriBlocks(self.handle, m["linesize"].size, _ptr(m["linesize"]),
                                                _ptr(m["linepoint"]), _ptr(m["lines"]), _ptr(m["lineface"]),
                                                _ptr(m["fc"]), _ptr(jacO1), _ptr(jacO2), _ptr(Aout)))

    def lineLU(self, nelem, nintface, m, Dia, Of1, Of2, DinvC):
        A.check(self._lib.tbs_lineLU(self.handle, nelem, nintface, m["linesize"].size, _ptr(m["fc"]),
                                     _ptr(m["linesize"]), _ptr(m["linepoint"]), _ptr(m["lines"]),
                                     _ptr(m["lineface"]), _ptr(Dia), _ptr(Of1), _ptr(Of2), _ptr(DinvC)))

    def solveDU(self, nelem, nintface, m, Dia, DinvC, Amat, dU, R):
        A.check(self._lib.tbs_triblock_solveDU(self.handle, nelem, nintface, m["ef"].size, m["linesize"].size,
                                               m["lines"].size, _ptr(m["linesize"]), _ptr(m["linepoint"]),
                                               _ptr(m["lines"]), _ptr(Dia), _ptr(DinvC), _ptr(Amat), _ptr(dU), _ptr(R)))

    def lineRes(self, nelem, nintface, m, Dia, Of1, Of2, U, R):
        A.check(self._lib.tbs_triblock_lineRes(self.handle, nelem, nintface, m["ef"].size, m["linesize"].size,
                                               m["lines"].size, _ptr(m["epoint"]), _ptr(m["ef"]), _ptr(m["fc"]),
                                               _ptr(m["linesize"]), _ptr(m["linepoint"]), _ptr(m["lines"]),
                                               _ptr(Dia), _ptr(Of1), _ptr(Of2), _ptr(U), _ptr(R)))

    def dgemvLU(self, nelem, LU, x, b):
        A.check(self._lib.tbs_dgemvLU(self.handle, nelem, _ptr(LU), _ptr(x), _ptr(b)))

    def norm2(self, n, v):
        out = C.c_double()
        A.check(self._lib.tbs_norm2(self.handle, n, _ptr(v), C.byref(out)))
        return out.value

    def generate_device(self, seed, nelem, nintface, shift, jacD, jacO1, jacO2, rhs):
        A.check(self._lib.tbs_generate_device(self.handle, seed, self.nvar, nelem, nintface, float(shift),
                                              _ptr(jacD), _ptr(jacO1), _ptr(jacO2), _ptr(rhs)))


class Plan:
    """Fused LIJ solver for one system: re-laid-out, factored, graph-captured sweeps."""

    def __init__(self, ctx, mesh, precision=A.TBS_FP64):
        self.ctx = ctx
        self.mesh = mesh
        self._ms = mesh.as_struct()
        h = C.c_void_p()
        A.check(ctx._lib.tbs_plan_create(ctx.handle, C.byref(self._ms), precision, C.byref(h)))
        self.handle = h
        self.precision = precision
        ctx._children.add(self)

    @staticmethod
    def footprint(mesh, precision=A.TBS_FP64):
        ms = mesh.as_struct()
        out = C.c_size_t()
        A.check(A.load().tbs_plan_footprint(C.byref(ms), precision, C.byref(out)))
        return out.value

    def close(self):
        if self.handle and self.ctx.handle:
            self.ctx._lib.tbs_plan_destroy(self.handle)
        self.handle = None

    def internal(self, which):
        """(device address, bytes) of an internal array, for diagnostics:
        0 g, 1 d', 2 U, 3 rhs, 4 R, 5 W, 6 E windows, 7 Gamma windows, 8 residual blocks."""
        p, nb = C.c_void_p(), C.c_size_t()
        A.check(self.ctx._lib.tbs_plan_internal(self.handle, which, C.byref(p), C.byref(nb)))
        return p.value or 0, nb.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = A.PlanInfo()
        A.check(self.ctx._lib.tbs_plan_info_get(self.handle, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in A.PlanInfo._fields_}

    def autotune(self, reps=3):
        """Pick the fastest Thomas variant; returns (chosen index, {variant name: us})."""
        pf = self.variants()
        n = len(pf)
        us = np.zeros(n)
        ch = C.c_int()
        A.check(self.ctx._lib.tbs_plan_autotune(self.handle, reps, C.byref(ch), _ptr(us), n))
        return ch.value, {variant_name(p): float(t) for p, t in zip(pf, us)}

    def variants(self):
        """Codes of the compiled Thomas variants (see variant_name)."""
        n = self.ctx._lib.tbs_plan_variants(self.handle, None, 0)
        pf = np.zeros(n, np.int32)
        self.ctx._lib.tbs_plan_variants(self.handle, _ptr(pf), n)
        return [int(x) for x in pf]

    def set_variant(self, v):
        A.check(self.ctx._lib.tbs_plan_set_variant(self.handle, v))

    def tuning(self):
        """The chosen tuple (Thomas variant and its prefetch / LDS windows /
        lines per workgroup, residual waves per workgroup) and the last
        autotune's timings."""
        t = A.Tuning()
        A.check(self.ctx._lib.tbs_plan_tuning(self.handle, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, **kw):
        """Set thomas_variant and/or residual_variant."""
        _set_tuning(self.ctx._lib.tbs_plan_tuning, self.ctx._lib.tbs_plan_set_tuning, self.handle, kw)

    def factor(self, jacD, jacO1, jacO2):
        """Factorise; returns after the device has finished reading the Jacobian
        (the C call is asynchronous, and the caller's buffers may be temporaries)."""
        A.check(self.ctx._lib.tbs_plan_factor(self.handle, _ptr(jacD), _ptr(jacO1), _ptr(jacO2)))
        self.ctx.sync()

    def solve(self, rhs, U, max_sweeps, rtol=0.0, R_out=None):
        hist = np.zeros(max_sweeps + 1)
        ns = C.c_int()
        A.check(self.ctx._lib.tbs_plan_solve(self.handle, _ptr(rhs), _ptr(U), max_sweeps, float(rtol),
                                             _ptr(hist), C.byref(ns), _ptr(R_out)))
        return hist[: ns.value + 1], ns.value

    def begin(self, rhs, U0=None):
        A.check(self.ctx._lib.tbs_plan_begin(self.handle, _ptr(rhs), _ptr(U0)))

    def launch_thomas(self):
        A.check(self.ctx._lib.tbs_plan_launch_thomas(self.handle))

    def launch_residual(self):
        A.check(self.ctx._lib.tbs_plan_launch_residual(self.handle))

    def launch_kernel(self, step, ev_start=None, ev_stop=None):
        """One kernel of a sweep (STEP_THOMAS, STEP_RESIDUAL, STEP_NORM); with
        two events from Context.event(), they time that kernel alone."""
        A.check(self.ctx._lib.tbs_plan_launch_kernel(self.handle, step, ev_start, ev_stop))

    def launch_sweeps(self, n, use_graph=True):
        A.check(self.ctx._lib.tbs_plan_launch_sweeps(self.handle, n, 1 if use_graph else 0))

    def end(self, U=None, R=None):
        A.check(self.ctx._lib.tbs_plan_end(self.handle, _ptr(U), _ptr(R)))

    def history(self, cap=4096):
        hist = np.zeros(cap)
        cnt = C.c_int()
        A.check(self.ctx._lib.tbs_plan_history(self.handle, _ptr(hist), cap, C.byref(cnt)))
        return hist[: min(cnt.value, cap)]


class Batch:
    """Many independent plans advanced together (one launch per kernel per sweep)."""

    def __init__(self, ctx, plans):
        self.ctx = ctx
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = C.c_void_p()
        A.check(ctx._lib.tbs_batch_create(ctx.handle, arr, len(self.plans), C.byref(h)))
        self.handle = h
        ctx._children.add(self)

    def close(self):
        if self.handle and self.ctx.handle:
            self.ctx._lib.tbs_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=3, nvariants=8):
        """Tune Thomas variant, residual shape and the two-stream split over all
        the batch's systems; returns (Thomas variant, its timings per variant)."""
        us = np.zeros(nvariants)
        ch = C.c_int()
        A.check(self.ctx._lib.tbs_batch_autotune(self.handle, reps, C.byref(ch), _ptr(us), nvariants))
        return ch.value, us

    def tuning(self):
        t = A.Tuning()
        A.check(self.ctx._lib.tbs_batch_tuning(self.handle, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, **kw):
        """Set thomas_variant, residual_variant and/or two_stream."""
        _set_tuning(self.ctx._lib.tbs_batch_tuning, self.ctx._lib.tbs_batch_set_tuning, self.handle, kw)

    def set_plans(self, plans):
        """Replace the batch's systems (tables rewritten in place, stream-ordered)."""
        self.plans = list(plans)
        arr = (C.c_void_p * max(len(self.plans), 1))(*[p.handle.value for p in self.plans])
        A.check(self.ctx._lib.tbs_batch_set_plans(self.handle, arr, len(self.plans)))

    def launch_sweeps(self, n, max_sweeps, rtol=0.0):
        """Enqueue n sweeps of every system (each plan begun with Plan.begin)."""
        A.check(self.ctx._lib.tbs_batch_launch_sweeps(self.handle, n, max_sweeps, float(rtol)))

    def states(self):
        """Per-system [sweeps recorded, active, status, 0] after a device sync."""
        st = np.zeros((len(self.plans), 4), np.int32)
        A.check(self.ctx._lib.tbs_batch_states(self.handle, _ptr(st)))
        return st

    def solve(self, rhs_list, U_list, max_sweeps, rtol=0.0, sweeps_per_check=4):
        n = len(self.plans)
        rp = (C.c_void_p * n)(*[_ptr(r).value for r in rhs_list])
        up = (C.c_void_p * n)(*[_ptr(u).value for u in U_list])
        hist = np.zeros((n, max_sweeps + 1))
        ns = np.zeros(n, np.int32)
        A.check(self.ctx._lib.tbs_batch_solve(self.handle, rp, up, max_sweeps, float(rtol), sweeps_per_check,
                                              _ptr(hist), _ptr(ns)))
        return [hist[i, : ns[i] + 1] for i in range(n)], ns


class Gmres:
    """Flexible line-preconditioned GMRES(restart) on a factored plan (csrc/gmres.hip)."""

    def __init__(self, plan, restart=20):
        self.plan = plan
        self.ctx = plan.ctx
        h = C.c_void_p()
        A.check(self.ctx._lib.tbs_gmres_create(plan.handle, restart, C.byref(h)))
        self.handle = h
        self.ctx._children.add(self)

    def close(self):
        if self.handle and self.ctx.handle:
            self.ctx._lib.tbs_gmres_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, rhs, U, max_iters=100, rtol=1e-10, sweeps=1):
        """Solve J U = -rhs from the initial guess in U; returns (history, iterations)."""
        hist = np.zeros(max_iters + 2)
        it = C.c_int()
        A.check(self.ctx._lib.tbs_gmres_solve(self.handle, _ptr(rhs), _ptr(U), max_iters, float(rtol), sweeps,
                                              _ptr(hist), hist.size, C.byref(it)))
        return hist[: it.value + 1], it.value

    def stats(self):
        s = np.zeros(4)
        A.check(self.ctx._lib.tbs_gmres_stats(self.handle, _ptr(s), 4))
        return dict(operator_applications=int(s[0]), preconditioner_sweeps=int(s[1]), restarts=int(s[2]),
                    true_residual=float(s[3]))


class Server:
    """Continuous-batching solve service (the C engine in csrc/server.hip)."""

    def __init__(self, ctx, hbm_budget=0, precision=A.TBS_FP32, rtol=1e-10, max_sweeps=100, sweeps_per_step=4):
        self.ctx = ctx
        h = C.c_void_p()
        A.check(ctx._lib.tbs_server_create(ctx.handle, hbm_budget, precision, float(rtol), max_sweeps,
                                           sweeps_per_step, C.byref(h)))
        self.handle = h
        self.max_sweeps = max_sweeps
        self.nvar = ctx.nvar
        self.sizes = {}
        self._keep = {}       # device inputs of submitted systems, until released
        ctx._children.add(self)

    def close(self):
        if self.handle and self.ctx.handle:
            self.ctx._lib.tbs_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, NL, NK):
        req = A.ServerRequest(NL, NK, 0, 0.0)
        out = C.c_size_t()
        A.check(self.ctx._lib.tbs_server_footprint(self.handle, C.byref(req), C.byref(out)))
        return out.value

    def submit(self, NL, NK, seed, shift=12.0):
        req = A.ServerRequest(NL, NK, seed, float(shift))
        rid = C.c_long()
        A.check(self.ctx._lib.tbs_server_submit(self.handle, C.byref(req), C.byref(rid)))
        self.sizes[rid.value] = NL * NK
        return rid.value

    def submit_system(self, mesh, jacD, jacO1, jacO2, rhs):
        """Submit a caller's system: mesh (HostMesh), Jacobian blocks and rhs in
        the reference layout, all host numpy arrays (copied now) or all
        DeviceBuffers (read at admission; keep them until the id is polled)."""
        ms = mesh.as_struct()
        dev = isinstance(jacD, DeviceBuffer)
        if not dev:
            jacD, jacO1, jacO2, rhs = (np.ascontiguousarray(a, dtype=np.float64) for a in (jacD, jacO1, jacO2, rhs))
        rid = C.c_long()
        A.check(self.ctx._lib.tbs_server_submit_system(self.handle, C.byref(ms), _ptr(jacD), _ptr(jacO1), _ptr(jacO2),
                                                       _ptr(rhs), A.TBS_DEVICE if dev else A.TBS_HOST,
                                                       C.byref(rid)))
        self.sizes[rid.value] = mesh.nelem
        if dev:
            self._keep[rid.value] = (jacD, jacO1, jacO2, rhs)
        return rid.value

    def footprint_system(self, mesh):
        ms = mesh.as_struct()
        out = C.c_size_t()
        A.check(self.ctx._lib.tbs_server_footprint_system(self.handle, C.byref(ms), C.byref(out)))
        return out.value

    def timeline(self, rid):
        """(submitted, admitted, set-up done, converged, copies enqueued, done), server clock seconds."""
        t = np.zeros(6)
        A.check(self.ctx._lib.tbs_server_timeline(self.handle, rid, _ptr(t), 6))
        return t

    def step(self):
        a, q, m = C.c_int(), C.c_int(), C.c_int()
        A.check(self.ctx._lib.tbs_server_step(self.handle, C.byref(a), C.byref(q), C.byref(m)))
        return a.value, q.value, m.value

    def poll(self, cap=4096):
        ids = np.zeros(cap, np.int64)
        n = C.c_int()
        A.check(self.ctx._lib.tbs_server_poll(self.handle, _ptr(ids), cap, C.byref(n)))
        return [int(x) for x in ids[: n.value]]

    def result(self, rid, want_U=True):
        n = self.sizes.get(rid)           # unknown ids: the library reports the error
        U = np.zeros(self.nvar * n) if (want_U and n is not None) else None
        hist = np.zeros(self.max_sweeps + 1)
        ns, st = C.c_int(), C.c_int()
        times = np.zeros(3)
        A.check(self.ctx._lib.tbs_server_result(self.handle, rid, _ptr(U), 0 if U is None else U.size, _ptr(hist),
                                                hist.size, C.byref(ns), C.byref(st), _ptr(times)))
        out = dict(U=U, hist=hist[: ns.value + 1], nsweeps=ns.value, status=st.value, t_submit=times[0],
                   t_admit=times[1], t_done=times[2])
        if st.value == -1:                 # admission failed: the library's message says why
            out["error"] = self.ctx._lib.tbs_last_error().decode()
        return out

    def release(self, rid):
        A.check(self.ctx._lib.tbs_server_release(self.handle, rid))
        self.sizes.pop(rid, None)
        self._keep.pop(rid, None)

    def info(self):
        b, u, t = C.c_size_t(), C.c_size_t(), C.c_double()
        A.check(self.ctx._lib.tbs_server_info(self.handle, C.byref(b), C.byref(u), C.byref(t)))
        return dict(budget=b.value, used=u.value, now=t.value)

    def stats(self):
        """Cumulative host seconds per engine phase and event counts."""
        t = np.zeros(7)
        c = np.zeros(4, np.int64)
        A.check(self.ctx._lib.tbs_server_stats(self.handle, _ptr(t), 7, _ptr(c), 4))
        return dict(zip(("wait", "membership", "launch", "retire", "admit", "admit_plan_build", "admit_thread"),
                        t.tolist()),
                    **dict(zip(("steps", "admitted", "retired", "membership_updates"), c.tolist())))
