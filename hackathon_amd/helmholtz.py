"""Python host interface of the MI355X compact-Helmholtz solver.

Mirrors the reference's host pipeline (2D_Program/main_functions/main.c:4-46,
3D_Program/main_functions/main.c:4-39): coefficients -> rhs -> solver ->
residual -> normInf, with the same row-major complex layouts, over the C ABI
of libhh.so (include/hh_solver.h).  PyTorch is used only to hold device
memory and to supply the stream; every computation runs in the HIP kernels
of libhh.so, and this module refuses to run without that library.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhh.so")

HH_OK, HH_ERR_ARG, HH_ERR_DEVICE, HH_ERR_NOMEM, HH_ERR_UNSUPPORTED, HH_ERR_SINGULAR = range(6)
PREC_F64, PREC_MIXED = 0, 1

_lib = None


class HHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def lib():
    """Load libhh.so; raises if the HIP library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hackathon_amd.build`")
        # PyTorch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7) as the one libhh.so is linked against.  Loading
        # PyTorch's first makes libhh.so bind to it, so that the process has
        # one HIP runtime and the two share streams, events and memory.
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.hh_last_error.restype = ctypes.c_char_p
        L.hh_status_string.restype = ctypes.c_char_p
        L.hh_context_get_stream.restype = ctypes.c_void_p
        L.hh_context_bytes_in_use.restype = ctypes.c_size_t
        L.hh_plan_field_size.restype = ctypes.c_long
        L.hh_plan_workspace_bytes.restype = ctypes.c_size_t
        for name in ("hh_dst_rows", "hh_dst2d_layers"):
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def _check(rc):
    if rc != HH_OK:
        raise HHError(rc, lib().hh_last_error().decode())


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _hptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} complex values, got {a.size}")
    return a


def _torch():
    import torch
    return torch


def _require_cuda(t, name):
    if not (t.is_cuda and t.dtype == _torch().complex128 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous complex128 tensor on the GPU")


def _require_size(t, name, n, exact=True):
    """t is a contiguous complex128 GPU tensor of n (exact) or at least n values:
    the library reads and writes through raw device pointers."""
    _require_cuda(t, name)
    if (t.numel() != n) if exact else (t.numel() < n):
        raise ValueError(f"{name} holds {t.numel()} values, {'expected' if exact else 'needs at least'} {n}")


class Context:
    """Device, stream, twiddle tables and scratch of the library."""

    def __init__(self, device=0):
        torch = _torch()
        self.device = device
        torch.cuda.set_device(device)
        h = ctypes.c_void_p()
        _check(lib().hh_context_create(device, ctypes.byref(h)))
        self.h = h
        self.algo = -1
        self.bind_stream()

    def bind_stream(self, stream=None):
        """Run on `stream` (default: torch's current stream of this device)."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        _check(lib().hh_context_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)))

    def synchronize(self):
        _check(lib().hh_context_synchronize(self.h))

    def set_dst_algorithm(self, algo):
        """-1: automatic by length; else bit 0: 2(N+1)-point odd-extension FFT (else (N+1)-point with
        pre/post-pass), bit 1: 64-thread blocks."""
        _check(lib().hh_context_set_dst_algorithm(self.h, int(algo)))
        self.algo = int(algo)

    def set_memory_budget(self, nbytes):
        _check(lib().hh_context_set_memory_budget(self.h, ctypes.c_size_t(int(nbytes))))

    def bytes_in_use(self):
        return int(lib().hh_context_bytes_in_use(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().hh_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- transforms (forwardDST / reverseDST, DST of 3D layers) ----
    def dst_rows(self, x, out=None):
        """Orthonormal DST-I of every row (last axis) of a complex128 GPU tensor."""
        _require_cuda(x, "x")
        out = x.new_empty(x.shape) if out is None else out
        _require_size(out, "out", x.numel())
        N = x.shape[-1]
        self.bind_stream()
        _check(lib().hh_dst_rows(self.h, ctypes.c_int(N), ctypes.c_long(x.numel() // N), _dptr(x), _dptr(out)))
        return out

    def dst2d_layers(self, x, out=None):
        """DST-I along x then y of tensors shaped (..., Ny, Nx)."""
        _require_cuda(x, "x")
        out = x.new_empty(x.shape) if out is None else out
        _require_size(out, "out", x.numel())
        Ny, Nx = x.shape[-2], x.shape[-1]
        self.bind_stream()
        _check(lib().hh_dst2d_layers(self.h, ctypes.c_int(Nx), ctypes.c_int(Ny),
                                     ctypes.c_long(x.numel() // (Nx * Ny)), _dptr(x), _dptr(out)))
        return out

    def norms(self, x, y=None):
        """(max |x|, max |x - y|) over all entries."""
        _require_cuda(x, "x")
        out = (ctypes.c_double * 2)()
        self.bind_stream()
        _check(lib().hh_norms(self.h, ctypes.c_long(x.numel()), _dptr(x),
                              _dptr(y) if y is not None else ctypes.c_void_p(0), out))
        return out[0], out[1]

    def norms_l2(self, x, y=None):
        """(sqrt(sum |x|^2), sqrt(sum |x - y|^2)) over all entries, on the device."""
        _require_cuda(x, "x")
        out = (ctypes.c_double * 2)()
        self.bind_stream()
        _check(lib().hh_norms_l2(self.h, ctypes.c_long(x.numel()), _dptr(x),
                                 _dptr(y) if y is not None else ctypes.c_void_p(0), out))
        return out[0], out[1]


# ------------------------------------------------------------------ coefficients
def coefficients2d(order, Nx, Ny, hx, hy, k2e, k2y=None, k2yy=None):
    """The reference's six 2D stencil vectors [bp, ap, b, a, bm, am].

    k2e: k^2 at y0 + j*hy, j = 0..Ny+1; k2y, k2yy: analytic (k^2)_y, (k^2)_yy
    at the interior rows (sixth order; None = 0)."""
    k2e = _c128(k2e, Ny + 2)
    k2y = None if k2y is None else _c128(k2y, Ny)
    k2yy = None if k2yy is None else _c128(k2yy, Ny)
    out = [np.zeros(Ny, np.complex128) for _ in range(6)]
    null = ctypes.c_void_p(0)
    _check(lib().hh_coefficients2d(order, Nx, Ny, ctypes.c_double(hx), ctypes.c_double(hy), _hptr(k2e),
                                   _hptr(k2y) if k2y is not None else null, _hptr(k2yy) if k2yy is not None else null,
                                   *[_hptr(o) for o in out]))
    return out


def coefficients3d(order, Nx, Ny, Nz, hx, hy, hz, k2e, k2d=None):
    """The reference's twelve 3D vectors [a, b, c, d, ap, bp, cp, dp, am, bm, cm, dm]."""
    k2e = _c128(k2e, Nz + 2)
    k2d = None if k2d is None else _c128(k2d, 4)
    out = [np.zeros(Nz, np.complex128) for _ in range(12)]
    arr = (ctypes.c_void_p * 12)(*[o.ctypes.data for o in out])
    _check(lib().hh_coefficients3d(order, Nx, Ny, Nz, ctypes.c_double(hx), ctypes.c_double(hy),
                                   ctypes.c_double(hz), _hptr(k2e), _hptr(k2d) if k2d is not None else ctypes.c_void_p(0),
                                   arr))
    return out


def _robin(fn, dims, nvec, coef, h, lo, hi):
    n_tri = dims[-1]
    face = int(np.prod(dims[:-1]))
    out = [_c128(v, n_tri).copy() for v in _operators(coef, nvec, n_tri)[0]]
    sides, corr = [], []
    for s in (lo, hi):
        if s is None:
            sides += [ctypes.c_void_p(0), ctypes.c_void_p(0)]
            corr.append(None)
            continue
        alpha, g = s
        a = np.array([complex(alpha)], np.complex128)
        gg = None if g is None else _c128(g, face)
        c = np.zeros(face, np.complex128)
        corr.append((a, gg, c))
        sides += [_hptr(a), _hptr(gg) if gg is not None else ctypes.c_void_p(0)]
    arr = (ctypes.c_void_p * nvec)(*[o.ctypes.data for o in out])
    cptr = [_hptr(c[2]) if c is not None else ctypes.c_void_p(0) for c in corr]
    _check(fn(*dims, ctypes.c_double(h), *sides, arr, *cptr))
    shape = tuple(reversed(dims[:-1]))
    return out, [None if c is None else c[2].reshape(shape) for c in corr]


def robin2d(coef, Nx, Ny, h, lo=None, hi=None):
    """Robin boundary du/dn + alpha u = g on the first (lo) and/or last (hi)
    row of y, which then lie on the boundary; lo / hi = (alpha, g) with g
    the boundary data along x (Nx complex, None = 0), or None (Dirichlet).
    Returns (folded coefficient vectors, [corr_lo, corr_hi]): add corr to
    that row of the right-hand side.  Second-order closure (ghost-row
    elimination with central differences); no reference counterpart."""
    return _robin(lib().hh_robin2d, (Nx, Ny), 6, coef, h, lo, hi)


def robin3d(coef, Nx, Ny, Nz, h, lo=None, hi=None):
    """As robin2d on the first/last z-layer; g: (Ny, Nx) face data."""
    return _robin(lib().hh_robin3d, (Nx, Ny, Nz), 12, coef, h, lo, hi)


def _packed_operator(dim, Nx, Ny, Nz, coef):
    nvec, n = (6, Ny) if dim == 2 else (12, Nz)
    ops = _operators(coef, nvec, n)
    if len(ops) != 1:
        raise ValueError("one operator expected")
    return np.ascontiguousarray(np.stack(ops[0]))


def operator_modes(dim, Nx, Ny, Nz, coef, kappa=4096.0):
    """Host-only classification of one operator's modes: (near-resonant modes, singular?, largest
    condition estimate among them).  coef: 2D [bp, ap, b, a, bm, am], 3D the twelve vectors."""
    packed = _packed_operator(dim, Nx, Ny, Nz, coef)
    nf, sg, kw = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().hh_operator_modes(dim, Nx, Ny, Nz if dim == 3 else 1, _hptr(packed), ctypes.c_double(kappa),
                                   ctypes.byref(nf), None, 0, ctypes.byref(sg), ctypes.byref(kw)))
    buf = (ctypes.c_int * max(nf.value, 1))()
    _check(lib().hh_operator_modes(dim, Nx, Ny, Nz if dim == 3 else 1, _hptr(packed), ctypes.c_double(kappa),
                                   ctypes.byref(nf), buf, nf.value, ctypes.byref(sg), ctypes.byref(kw)))
    return list(buf[:nf.value]), bool(sg.value), kw.value


def operator_mode_factors(dim, Nx, Ny, Nz, coef, mode):
    """Host-only reference-order factors of one mode: complex array (n, 3) of (L, R, Up), R the
    back-substitution multiplier (2D 1/U, 3D the reference's reciprocal pivot).  Returns
    (factors, singular?)."""
    packed = _packed_operator(dim, Nx, Ny, Nz, coef)
    n = Ny if dim == 2 else Nz
    out = np.zeros((n, 3), np.complex128)
    rc = lib().hh_operator_mode_factors(dim, Nx, Ny, Nz if dim == 3 else 1, _hptr(packed), int(mode), _hptr(out))
    if rc not in (HH_OK, HH_ERR_SINGULAR):
        _check(rc)
    return out, rc == HH_ERR_SINGULAR


# ------------------------------------------------------------------ plans
def _operators(coef, nvec, n):
    """Normalise one operator (nvec vectors of length n) or a list of them."""
    a = coef
    if isinstance(a, np.ndarray) and a.ndim == 3:
        return [[_c128(v, n) for v in op] for op in a]
    if isinstance(a, np.ndarray) and a.ndim == 2:
        return [[_c128(v, n) for v in a]]
    a = list(a)
    if len(a) == nvec and np.ndim(a[0]) == 1:
        return [[_c128(v, n) for v in a]]
    return [[_c128(v, n) for v in (op.reshape(nvec, n) if np.ndim(op) == 1 else op)] for op in a]


class Plan:
    dim = 0

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = None

    @property
    def field_size(self):
        return int(lib().hh_plan_field_size(self.h))

    def workspace_bytes(self, batch=1):
        return int(lib().hh_plan_workspace_bytes(self.h, batch))

    def _batch_of(self, t):
        n = self.field_size
        if t.numel() % n:
            raise ValueError(f"tensor of {t.numel()} values is not a whole number of fields of {n}")
        return t.numel() // n

    def solve(self, rhs, out=None):
        """sol = A^{-1} rhs for a batch of right-hand sides (complex128, GPU)."""
        _require_cuda(rhs, "rhs")
        out = rhs.new_empty(rhs.shape) if out is None else out
        _require_size(out, "out", rhs.numel())
        self.ctx.bind_stream()
        _check(lib().hh_solve(self.h, _dptr(rhs), _dptr(out), self._batch_of(rhs)))
        return out

    def solve_host(self, rhs):
        rhs = _c128(rhs)
        sol = np.empty_like(rhs)
        self.ctx.bind_stream()
        _check(lib().hh_solve_host(self.h, _hptr(rhs), _hptr(sol), rhs.size // self.field_size))
        return sol

    def residual(self, sol, rhs, out=None):
        """res = A sol - rhs (FP64 compact operator)."""
        _require_cuda(sol, "sol")
        _require_size(rhs, "rhs", sol.numel())
        out = sol.new_empty(sol.shape) if out is None else out
        _require_size(out, "out", sol.numel())
        self.ctx.bind_stream()
        _check(lib().hh_residual(self.h, _dptr(sol), _dptr(rhs), _dptr(out), self._batch_of(sol)))
        return out

    def set_refine(self, tol=1e-12, max_iter=6):
        _check(lib().hh_plan_set_refine(self.h, ctypes.c_double(tol), ctypes.c_int(max_iter)))

    def last_refine(self):
        s, r = ctypes.c_int(), ctypes.c_double()
        _check(lib().hh_plan_last_refine(self.h, ctypes.byref(s), ctypes.byref(r)))
        return s.value, r.value

    def set_real_operator(self, mode):
        """-1: real-band mode solves whenever every coefficient is real (default);
        0: complex-band mode solves always."""
        _check(lib().hh_plan_set_real_operator(self.h, int(mode)))

    @property
    def is_real_operator(self):
        return bool(lib().hh_plan_is_real_operator(self.h))

    def set_resonance_threshold(self, kappa):
        """Varah-bound threshold above which a mode is near-resonant (default 4096): such modes
        run the reference's sequential substitution with host-computed reference-order factors."""
        self.ctx.bind_stream()
        _check(lib().hh_plan_set_resonance_threshold(self.h, ctypes.c_double(kappa)))

    def mode_info(self, op=0):
        """(number of near-resonant modes, singular?, largest condition estimate among them) of operator op."""
        f, sg, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().hh_plan_mode_info(self.h, int(op), ctypes.byref(f), ctypes.byref(sg), ctypes.byref(k)))
        return f.value, bool(sg.value), k.value

    def flagged_modes(self, op=0):
        n = lib().hh_plan_flagged_modes(self.h, int(op), None, 0)
        if n < 0:
            _check(n)
        buf = (ctypes.c_int * max(n, 1))()
        lib().hh_plan_flagged_modes(self.h, int(op), buf, n)
        return list(buf[:n])

    def set_operators(self, coef):
        """Replace the operators (at most as many as the plan was created with)."""
        nvec, n = (6, self.Ny) if self.dim == 2 else (12, self.Nz)
        ops = _operators(coef, nvec, n)
        packed = np.ascontiguousarray(np.stack([np.stack(o) for o in ops]))
        self.ctx.bind_stream()
        _check(lib().hh_plan_set_operators(self.h, len(ops), _hptr(packed)))
        self.nop, self.ops, self.coef = len(ops), ops, ops[0]

    def close(self):
        if self.h:
            lib().hh_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan2D(Plan):
    """2D operator(s): coefficient vectors [bp, ap, b, a, bm, am] of length Ny.

    `coef` is one operator (six vectors) or a list of operators; with nop
    operators, right-hand side b of a solve uses operator b % nop."""
    dim = 2

    def __init__(self, ctx, Nx, Ny, coef, precision=PREC_F64):
        super().__init__(ctx)
        self.Nx, self.Ny = Nx, Ny
        ops = _operators(coef, 6, Ny)
        self.nop = len(ops)
        self.coef = ops[0]
        self.ops = ops
        packed = np.ascontiguousarray(np.stack([np.stack(o) for o in ops]))
        h = ctypes.c_void_p()
        _check(lib().hh_plan2d_create_multi(ctx.h, Nx, Ny, self.nop, _hptr(packed), precision, ctypes.byref(h)))
        self.h = h
        self.precision = precision

    def rhs(self, order, hx, hy, f_e, u_e, k2e, fy_b=None, fxx_b=None, fyy_b=None, k2yb=None, out=None):
        """Compact right-hand side (rhs.c) from f and u on the extended grid."""
        ext = (self.Nx + 2) * (self.Ny + 2)
        _require_cuda(f_e, "f_e")
        if f_e.numel() % ext:
            raise ValueError(f"f_e holds {f_e.numel()} values, not a whole number of extended grids of {ext}")
        batch = f_e.numel() // ext
        _require_size(u_e, "u_e", f_e.numel())
        out = f_e.new_empty((batch, self.Ny, self.Nx)) if out is None else out
        _require_size(out, "out", batch * self.Nx * self.Ny)
        ring = 2 * (self.Nx + 2) + 2 * (self.Ny + 2)
        for t, name, n in ((fy_b, "fy_b", 2 * self.Nx), (fxx_b, "fxx_b", ring), (fyy_b, "fyy_b", ring)):
            if t is not None:
                _require_size(t, name, batch * n, exact=False)
        # one k^2 profile (and (k^2)_y pair) per operator; a single one is shared
        k2e = _c128(k2e)
        if k2e.size == self.Ny + 2:
            k2e = np.tile(k2e, self.nop)
        k2e = _c128(k2e, (self.Ny + 2) * self.nop)
        k2yb = _c128(np.zeros(2) if k2yb is None else k2yb)
        if k2yb.size == 2:
            k2yb = np.tile(k2yb, self.nop)
        k2yb = _c128(k2yb, 2 * self.nop)
        null = ctypes.c_void_p(0)
        dp = lambda t: _dptr(t) if t is not None else null
        self.ctx.bind_stream()
        _check(lib().hh_rhs2d(self.h, order, ctypes.c_double(hx), ctypes.c_double(hy), _dptr(f_e), _dptr(u_e),
                              dp(fy_b), dp(fxx_b), dp(fyy_b), _hptr(k2e), _hptr(k2yb), _dptr(out), batch))
        return out


class Plan3D(Plan):
    """3D operator: the twelve vectors [a, b, c, d, ap, bp, cp, dp, am, bm, cm, dm] of length Nz."""
    dim = 3

    def __init__(self, ctx, Nx, Ny, Nz, coef, precision=PREC_F64):
        super().__init__(ctx)
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        if np.ndim(coef) == 1 and len(coef) == 12 * Nz:
            coef = np.asarray(coef).reshape(12, Nz)
        ops = _operators(coef, 12, Nz)
        self.nop = len(ops)
        self.coef = ops[0]
        self.ops = ops
        packed = np.ascontiguousarray(np.stack([np.stack(o) for o in ops]))
        h = ctypes.c_void_p()
        _check(lib().hh_plan3d_create_multi(ctx.h, Nx, Ny, Nz, self.nop, _hptr(packed), precision,
                                            ctypes.byref(h)))
        self.h = h
        self.precision = precision

    def rhs(self, order, hx, hy, hz, f_e, u_e, k2e, fxx_f=None, fyy_f=None, fzz_f=None, k2d=None, out=None):
        X, Y, Z = self.Nx + 2, self.Ny + 2, self.Nz + 2
        _require_cuda(f_e, "f_e")
        if f_e.numel() % (X * Y * Z):
            raise ValueError(f"f_e holds {f_e.numel()} values, not a whole number of extended grids of {X * Y * Z}")
        batch = f_e.numel() // (X * Y * Z)
        _require_size(u_e, "u_e", f_e.numel())
        out = f_e.new_empty((batch, self.Nz, self.Ny, self.Nx)) if out is None else out
        _require_size(out, "out", batch * self.Nx * self.Ny * self.Nz)
        for t, name, n in ((fxx_f, "fxx_f", 2 * Y * Z), (fyy_f, "fyy_f", 2 * X * Z), (fzz_f, "fzz_f", 2 * X * Y)):
            if t is not None:
                _require_size(t, name, batch * n, exact=False)
        k2e = _c128(k2e)
        if k2e.size == self.Nz + 2:
            k2e = np.tile(k2e, self.nop)
        k2e = _c128(k2e, (self.Nz + 2) * self.nop)
        k2d = _c128(np.zeros(4) if k2d is None else k2d)
        if k2d.size == 4:
            k2d = np.tile(k2d, self.nop)
        k2d = _c128(k2d, 4 * self.nop)
        null = ctypes.c_void_p(0)
        dp = lambda t: _dptr(t) if t is not None else null
        self.ctx.bind_stream()
        _check(lib().hh_rhs3d(self.h, order, ctypes.c_double(hx), ctypes.c_double(hy), ctypes.c_double(hz),
                              _dptr(f_e), _dptr(u_e), dp(fxx_f), dp(fyy_f), dp(fzz_f), _hptr(k2e), _hptr(k2d),
                              _dptr(out), batch))
        return out


class Graph:
    """A fixed list of solves [(plan, rhs, out), ...] recorded into one HIP
    graph; launch() replays all of them with one call on the context's stream.
    The tensors must stay alive (and at the same addresses) while the graph
    exists."""

    def __init__(self, ctx, solves):
        self.ctx = ctx
        self.h = None
        n = len(solves)
        plans = (ctypes.c_void_p * n)(*[p.h.value for p, _, _ in solves])
        rhs = (ctypes.c_void_p * n)(*[r.data_ptr() for _, r, _ in solves])
        out = (ctypes.c_void_p * n)(*[o.data_ptr() for _, _, o in solves])
        batch = (ctypes.c_int * n)(*[p._batch_of(r) for p, r, _ in solves])
        self._keep = solves
        ctx.bind_stream()
        h = ctypes.c_void_p()
        _check(lib().hh_graph_create(ctx.h, n, plans, rhs, out, batch, ctypes.byref(h)))
        self.h = h

    @property
    def num_nodes(self):
        return int(lib().hh_graph_num_nodes(self.h))

    def launch(self):
        self.ctx.bind_stream()
        _check(lib().hh_graph_launch(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().hh_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ serving
class _Request(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("order", ctypes.c_int), ("precision", ctypes.c_int), ("k_re", ctypes.c_double),
                ("k_im", ctypes.c_double), ("hx", ctypes.c_double), ("hy", ctypes.c_double),
                ("hz", ctypes.c_double), ("h_rhs", ctypes.c_void_p), ("h_sol", ctypes.c_void_p)]


class _Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("queue_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("sweeps", ctypes.c_int), ("rel_residual", ctypes.c_double)]


class Server:
    """Continuous-batching request loop of libhh.so (hh_server_*).

    submit() takes a host right-hand side (complex128, Nx*Ny(*Nz) values) and
    returns a ticket; wait(ticket) blocks until the solution is in host memory
    and returns (solution, info) with the queueing time and the latency from
    arrival to result."""

    def __init__(self, device=0, hbm_budget=0, lanes=2, max_batch=8):
        h = ctypes.c_void_p()
        _check(lib().hh_server_create(device, ctypes.c_size_t(int(hbm_budget)), lanes, max_batch, ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, dim, N, order, k, rhs, precision=PREC_F64, h=None, out=None):
        """Queue one solve; `out` (optional, e.g. pinned memory for fast copies) receives the solution."""
        Nx, Ny, Nz = (N, N, N) if np.isscalar(N) else (tuple(N) + (1,))[:3]
        n = Nx * Ny * (Nz if dim == 3 else 1)
        rhs = _c128(rhs, n)   # the server copies whole fields from these buffers
        sol = np.empty_like(rhs) if out is None else out
        if sol.dtype != np.complex128 or sol.size != rhs.size or not sol.flags.c_contiguous:
            raise ValueError("out must be a contiguous complex128 array of the right-hand side's size")
        hx, hy, hz = (1.0 / (Nx + 1), 1.0 / (Ny + 1), 1.0 / (Nz + 1)) if h is None else h
        k = complex(k)
        r = _Request(dim, Nx, Ny, Nz if dim == 3 else 1, order, precision, k.real, k.imag, hx, hy, hz,
                     rhs.ctypes.data, sol.ctypes.data)
        t = ctypes.c_long()
        _check(lib().hh_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)))
        self._live[t.value] = (rhs, sol)
        return t.value

    def wait(self, ticket):
        res = _Result()
        rc = lib().hh_server_wait(self.h, ctypes.c_long(ticket), ctypes.byref(res))
        rhs, sol = self._live.pop(ticket)
        _check(rc)
        return sol, {"queue_s": res.queue_s, "latency_s": res.latency_s, "sweeps": res.sweeps,
                     "rel_residual": res.rel_residual}

    def stats(self):
        b, s, r = ctypes.c_long(), ctypes.c_long(), ctypes.c_size_t()
        _check(lib().hh_server_stats(self.h, ctypes.byref(b), ctypes.byref(s), ctypes.byref(r)))
        return {"batches": b.value, "solved": s.value, "reserved_bytes": r.value}

    def close(self):
        if getattr(self, "h", None):
            lib().hh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
