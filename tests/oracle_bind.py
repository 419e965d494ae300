"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker, never the thing under test: only tests/, the smoke
check and the CPU-baseline leg of bench.py load it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

PC = {"none": 0, "jacobi": 1, "ilu0": 2, "asm": 3, "ras": 4, "twolevel_asm": 5,
      "twolevel_ras": 6}

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        L = _lib
        L.or_poisson2d.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_poisson3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_spmv.argtypes = [C.c_int, _I, _I, _D, _D, _D]
        L.or_dot.argtypes = [C.c_int, _D, _D]
        L.or_dot.restype = C.c_double
        L.or_ilu0_factor.argtypes = [C.c_int, _I, _I, _D, _D, _I]
        L.or_ilu0_solve.argtypes = [C.c_int, _I, _I, _D, _I, _D, _D]
        L.or_overlap.argtypes = [C.c_int, _I, _I, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p]
        L.or_precond_new.argtypes = [C.c_int, C.c_int, _I, _I, _D, C.c_int, C.c_int,
                                     C.c_void_p, C.c_int]
        L.or_precond_new.restype = C.c_void_p
        L.or_precond_apply.argtypes = [C.c_void_p, _D, _D]
        _F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.or_unet_nparams.restype = C.c_int
        L.or_unet_param_offset.argtypes = [C.c_int]
        L.or_unet_param_offset.restype = C.c_int
        L.or_unet_forward.argtypes = [_F, C.c_int, C.c_int, C.c_int, _F, _F, C.c_int]
        L.or_unet_loss_grad.argtypes = [_F, C.c_int, C.c_int, C.c_int, _F, _F, C.c_int, _D]
        L.or_unet_loss_grad.restype = C.c_double
        L.or_unet_loss_grad_kind.argtypes = [_F, C.c_int, C.c_int, C.c_int, _F, _F, C.c_int,
                                             C.c_int, _D]
        L.or_unet_loss_grad_kind.restype = C.c_double
        L.or_precond_coarse.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_precond_free.argtypes = [C.c_void_p]
        L.or_lu_factor.argtypes = [C.c_int, _D, _I]
        L.or_lu_solve.argtypes = [C.c_int, _D, _I, _D, _D]
        L.or_solve.argtypes = [C.c_int, C.c_int, _I, _I, _D, _D, _D, C.c_void_p, C.c_int,
                               C.c_int, C.c_double, C.POINTER(C.c_double),
                               C.POINTER(C.c_int), _D, C.c_int, C.POINTER(C.c_int)]
        L.or_metrics_json.argtypes = [C.c_char_p, C.c_char_p] + [C.c_int] * 8 + \
            [C.c_double] * 6 + [C.c_char_p, C.c_int]
        L.or_bicgstab.argtypes = [C.c_int, _I, _I, _D, _D, _D, C.c_void_p, C.c_int, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int), _D, C.c_int,
                                  C.POINTER(C.c_int)]
        L.or_ilut_new.argtypes = [C.c_int, _I, _I, _D, C.c_double, C.c_int, C.c_int,
                                  C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.or_ilut_new.restype = C.c_void_p
        L.or_ilut_max_list.argtypes = [C.c_void_p]
        L.or_ilut_max_list.restype = C.c_longlong
        L.or_lanczos.argtypes = [C.c_int, _I, _I, _D, C.c_void_p, C.c_int, C.c_uint, _D, _D]
        L.or_tridiag_eigs.argtypes = [C.c_int, _D, _D, _D]
        L.or_condition.argtypes = [C.c_int, _I, _I, _D, C.c_void_p, C.c_int, C.c_uint, _D,
                                   C.POINTER(C.c_int)]
        L.or_spectral.argtypes = [C.c_int, _I, _I, _D, C.c_void_p, C.c_int, C.c_double,
                                  C.c_uint, C.POINTER(C.c_double)]
    return _lib


class Csr:
    """Host CSR triple (int32 row_ptr / col_idx, float64 vals)."""

    def __init__(self, rp, ci, v, ncols=None):
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.ci = np.ascontiguousarray(ci, dtype=np.int32)
        self.v = np.ascontiguousarray(v, dtype=np.float64)
        self.n = len(self.rp) - 1
        self.ncols = self.n if ncols is None else ncols

    @property
    def nnz(self):
        return int(self.rp[-1])

    def dense(self):
        A = np.zeros((self.n, self.ncols))
        for i in range(self.n):
            for j in range(self.rp[i], self.rp[i + 1]):
                A[i, self.ci[j]] += self.v[j]
        return A


def poisson2d(nx, ny):
    L = lib()
    nnz = L.or_poisson2d(nx, ny, None, None, None)
    rp = np.zeros(nx * ny + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    L.or_poisson2d(nx, ny, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
    return Csr(rp, ci, v)


def poisson3d(nx, ny, nz):
    L = lib()
    nnz = L.or_poisson3d(nx, ny, nz, None, None, None)
    rp = np.zeros(nx * ny * nz + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    L.or_poisson3d(nx, ny, nz, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
    return Csr(rp, ci, v)


def grid(spec):
    """"2d:NX:NY", "3d:NX:NY:NZ" or "rnd:N:SEED" (tests/rng_ref.rnd_matrix)."""
    if spec.startswith("rnd:"):
        from rng_ref import rnd_matrix
        n, seed = (int(t) for t in spec.split(":")[1:])
        return Csr(*rnd_matrix(n, seed))
    dims = [int(t) for t in spec.split(":")[1:]]
    return poisson2d(*dims) if len(dims) == 2 else poisson3d(*dims)


def spmv(A, x):
    y = np.zeros(A.n)
    lib().or_spmv(A.n, A.rp, A.ci, A.v, np.ascontiguousarray(x, np.float64), y)
    return y


def ilu0(A):
    lu = np.zeros(A.nnz)
    diag = np.zeros(A.n, np.int32)
    lib().or_ilu0_factor(A.n, A.rp, A.ci, A.v, lu, diag)
    return lu, diag


def ilu0_solve(A, lu, diag, x):
    y = np.zeros(A.n)
    lib().or_ilu0_solve(A.n, A.rp, A.ci, lu, diag, np.ascontiguousarray(x, np.float64), y)
    return y


def overlap_sets(A, part, nparts, overlap):
    L = lib()
    part = np.ascontiguousarray(part, np.int32)
    tot = L.or_overlap(A.n, A.rp, A.ci, part.ctypes.data, nparts, overlap, None, None)
    ptr = np.zeros(nparts + 1, np.int32)
    idx = np.zeros(tot, np.int32)
    L.or_overlap(A.n, A.rp, A.ci, part.ctypes.data, nparts, overlap, ptr.ctypes.data,
                 idx.ctypes.data)
    return [idx[ptr[s]:ptr[s + 1]] for s in range(nparts)]


class Precond:
    def __init__(self, kind, A, nparts=1, overlap=0, part=None, precision=0, drop_tol=1e-4,
                 fill_factor=10):
        self.A = A
        if kind == "ilut":
            ln, un = C.c_longlong(), C.c_longlong()
            self.h = lib().or_ilut_new(A.n, A.rp, A.ci, A.v, drop_tol, fill_factor, precision,
                                       C.byref(ln), C.byref(un))
            self.l_nnz, self.u_nnz = ln.value, un.value
            self.max_list = lib().or_ilut_max_list(self.h)
            self.nparts = 1
            self._part = None
            return
        self._part = None if part is None else np.ascontiguousarray(part, np.int32)
        self.h = lib().or_precond_new(PC[kind], A.n, A.rp, A.ci, A.v, nparts, overlap,
                                      None if self._part is None else self._part.ctypes.data,
                                      precision)
        if not self.h:
            raise RuntimeError("oracle preconditioner setup failed")
        self.nparts = nparts

    def apply(self, x):
        y = np.zeros(self.A.n)
        lib().or_precond_apply(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def coarse(self):
        nc = self.nparts
        A0 = np.zeros(nc * nc)
        LU = np.zeros(nc * nc)
        piv = np.zeros(nc, np.int32)
        lib().or_precond_coarse(self.h, A0.ctypes.data, LU.ctypes.data, piv.ctypes.data)
        return A0.reshape(nc, nc), LU.reshape(nc, nc), piv

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_precond_free(self.h)
            self.h = None


def lu_factor(A):
    n = A.shape[0]
    LU = np.ascontiguousarray(A, np.float64).ravel().copy()
    piv = np.zeros(n, np.int32)
    rc = lib().or_lu_factor(n, LU, piv)
    return rc, LU.reshape(n, n), piv


def lu_solve(LU, piv, b):
    n = LU.shape[0]
    x = np.zeros(n)
    lib().or_lu_solve(n, np.ascontiguousarray(LU).ravel(), piv, np.ascontiguousarray(b, np.float64), x)
    return x


def solve(method, A, b, pc=None, restart=30, max_iter=1000, tol=1e-10, x0=None):
    """Reference-semantics GMRES (method 'gmres') or FGMRES ('fgmres')."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    rn = C.c_double()
    conv = C.c_int()
    hl = C.c_int()
    cap = max_iter + restart + 8
    hist = np.zeros(cap)
    it = lib().or_solve(1 if method == "fgmres" else 0, A.n, A.rp, A.ci, A.v,
                        np.ascontiguousarray(b, np.float64), x, pc.h if pc else None,
                        restart, max_iter, tol, C.byref(rn), C.byref(conv), hist, cap,
                        C.byref(hl))
    return dict(iterations=it, residual_norm=rn.value, converged=bool(conv.value),
                hist=hist[:min(hl.value, cap)].copy(), x=x)


def metrics_json(**kw):
    buf = C.create_string_buffer(4096)
    f = dict(matrix="", method="", n=0, nnz=0, ranks=1, threads=1, gpus=0, overlap=0,
             restart=30, iterations=0, tol=1e-10, residual_norm=0.0, time_setup_s=0.0,
             time_solve_s=0.0, speedup=1.0, efficiency=1.0)
    f.update(kw)
    lib().or_metrics_json(f["matrix"].encode(), f["method"].encode(), f["n"], f["nnz"],
                          f["ranks"], f["threads"], f["gpus"], f["overlap"], f["restart"],
                          f["iterations"], f["tol"], f["residual_norm"], f["time_setup_s"],
                          f["time_solve_s"], f["speedup"], f["efficiency"], buf, 4096)
    return buf.value.decode()


# ---- learned-stage network (oracle/unet_oracle.cpp) ----

def unet_nparams():
    return lib().or_unet_nparams()


def unet_param_offset(i):
    return lib().or_unet_param_offset(i)


def unet_forward(params, x, round_bf16=True):
    """x: [B, 2, H, W] float32 -> [B, H, W]."""
    x = np.ascontiguousarray(x, np.float32)
    B, _, H, W = x.shape
    out = np.zeros((B, H, W), np.float32)
    lib().or_unet_forward(np.ascontiguousarray(params, np.float32), B, H, W, x, out,
                          1 if round_bf16 else 0)
    return out


def unet_loss_grad(params, x, t, round_bf16=True, kind=0):
    """kind 0: relative squared error against t; kind 1: residual loss (t unused);
    kind 2: kind 1 plus 0.01 ln(|Au|^2/|s|^2)^2 (t unused)."""
    x = np.ascontiguousarray(x, np.float32)
    B, _, H, W = x.shape
    g = np.zeros(unet_nparams())
    tt = np.ascontiguousarray(t if t is not None else np.zeros((B, H, W)), np.float32)
    loss = lib().or_unet_loss_grad_kind(np.ascontiguousarray(params, np.float32), B, H, W, x, tt,
                                        kind, 1 if round_bf16 else 0, g)
    return loss, g


def bicgstab(A, b, pc=None, max_iter=1000, tol=1e-10, x0=None):
    """Reference-semantics BiCGSTAB (relative tol)."""
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    rn, conv, hl = C.c_double(), C.c_int(), C.c_int()
    cap = max(max_iter, 1) + 2
    hist = np.zeros(cap)
    it = lib().or_bicgstab(A.n, A.rp, A.ci, A.v, np.ascontiguousarray(b, np.float64), x,
                           pc.h if pc else None, max_iter, tol, C.byref(rn), C.byref(conv),
                           hist, cap, C.byref(hl))
    return dict(iterations=it, residual_norm=rn.value, converged=bool(conv.value),
                hist=hist[:min(hl.value, cap)].copy(), x=x)


def lanczos(A, pc=None, k=80, seed=42):
    kk = max(2, min(k, A.n))
    al, be = np.zeros(kk), np.zeros(kk)
    m = lib().or_lanczos(A.n, A.rp, A.ci, A.v, pc.h if pc else None, k, seed, al, be)
    return al[:m].copy(), be[:max(m - 1, 0)].copy()


def tridiag_eigs(d, e):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(np.append(e, 0.0), np.float64)
    out = np.zeros(len(d))
    lib().or_tridiag_eigs(len(d), d, e, out)
    return out


def condition(A, pc=None, k=80, seed=42):
    out = np.zeros(3)
    valid = C.c_int()
    m = lib().or_condition(A.n, A.rp, A.ci, A.v, pc.h if pc else None, k, seed, out,
                           C.byref(valid))
    return dict(lambda_min=out[0], lambda_max=out[1], kappa=out[2], steps=m,
                valid=int(valid.value))


def spectral(A, pc=None, max_iter=200, tol=1e-6, seed=123):
    rho = C.c_double()
    it = lib().or_spectral(A.n, A.rp, A.ci, A.v, pc.h if pc else None, max_iter, tol, seed,
                           C.byref(rho))
    return dict(rho=rho.value, power_iters=it, ok=int(rho.value < 1.0))
