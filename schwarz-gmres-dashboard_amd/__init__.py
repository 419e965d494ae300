"""ctypes binding of libsgamd.so, the MI355X Schwarz-(F)GMRES C ABI (include/sg_capi.h).

The product path is the HIP library; this module only moves arrays and calls
the C entry points.  It never falls back to a CPU implementation: when the
library is missing or no GPU is visible, the calls raise.

Names mirror the reference's host interface (izbassarn7/schwarz-gmres-dashboard):
``Precond`` covers JacobiPrecond / ILU0Precond / ASMPrecond / RASPrecond /
TwoLevelSchwarzPrecond (src/precond/*.h), ``solve`` covers gmres / fgmres
(src/core/gmres.h:28-33, src/core/fgmres.h:42-55) with GMRESParams /
GMRESResult fields.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgamd.so")

PC_KIND = {"none": 0, "jacobi": 1, "ilu0": 2, "asm": 3, "ras": 4, "twolevel_asm": 5,
           "twolevel_ras": 6, "ilut": 8}
STATUS = {0: "SG_OK", 1: "SG_E_ARG", 2: "SG_E_HIP", 3: "SG_E_SINGULAR", 4: "SG_E_PARTITION",
          5: "SG_E_IO", 6: "SG_E_NOCONV", 7: "SG_E_STATE"}
K_FAMILY = {"spmv": 0, "precond": 1, "trsv": 2, "ortho": 3, "update": 4, "gram_dot": 5,
            "gram_update": 6}


class SgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class PrecondOpts(C.Structure):
    _fields_ = [("nparts", C.c_int), ("overlap", C.c_int), ("precision", C.c_int),
                ("partition", C.POINTER(C.c_int)), ("ufactor", C.c_int),
                ("coarse_solve", C.c_int), ("weighted_partition", C.c_int),
                ("drop_tol", C.c_double), ("fill_factor", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("method", C.c_int), ("restart", C.c_int), ("max_iter", C.c_int),
                ("tol_abs", C.c_double), ("rtol", C.c_double), ("ortho", C.c_int),
                ("lookahead", C.c_int), ("flags", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("residual_norm", C.c_double), ("rel_true_residual", C.c_double),
                ("time_solve_s", C.c_double), ("cycles", C.c_int),
                ("precond_applies", C.c_int), ("hist_len", C.c_int)]


class ConditionEstimate(C.Structure):
    _fields_ = [("lambda_min", C.c_double), ("lambda_max", C.c_double), ("kappa", C.c_double),
                ("lanczos_steps", C.c_int), ("valid", C.c_int)]


class SpectralResult(C.Structure):
    _fields_ = [("spectral_radius", C.c_double), ("convergence_ok", C.c_int),
                ("power_iters", C.c_int)]


class Convergence(C.Structure):
    _fields_ = [("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("asymptotic_rate", C.c_double), ("stagnation_start", C.c_int),
                ("converged", C.c_int), ("iterations", C.c_int)]


class LearnedOpts(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("min_gain", C.c_double), ("every", C.c_int),
                ("coarsen", C.c_int)]


class PrecondInfo(C.Structure):
    _fields_ = [("nsub", C.c_int), ("n_ext", C.c_longlong), ("nnz_ext", C.c_longlong),
                ("levels_fwd", C.c_longlong), ("levels_bwd", C.c_longlong),
                ("max_levels", C.c_int), ("z_in_lds", C.c_int), ("nc", C.c_int),
                ("max_ns", C.c_int), ("sweep_shape", C.c_int)]


_lib = None
_VP = C.c_void_p


def lib():
    """Load libsgamd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    sig = {
        "sg_context_create": ([C.c_int, C.POINTER(_VP)], C.c_int),
        "sg_context_destroy": ([_VP], C.c_int),
        "sg_context_synchronize": ([_VP], C.c_int),
        "sg_last_error": ([_VP], C.c_char_p),
        "sg_malloc": ([_VP, C.c_size_t, C.POINTER(_VP)], C.c_int),
        "sg_free": ([_VP, _VP], C.c_int),
        "sg_memcpy_h2d": ([_VP, _VP, _VP, C.c_size_t], C.c_int),
        "sg_memcpy_d2h": ([_VP, _VP, _VP, C.c_size_t], C.c_int),
        "sg_memcpy_d2d": ([_VP, _VP, _VP, C.c_size_t], C.c_int),
        "sg_memset_zero": ([_VP, _VP, C.c_size_t], C.c_int),
        "sg_csr_create": ([_VP, C.c_int, C.c_int, C.c_int, _VP, _VP, _VP, C.POINTER(_VP)], C.c_int),
        "sg_csr_destroy": ([_VP], C.c_int),
        "sg_csr_info": ([_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "sg_spmv": ([_VP, C.c_double, _VP, C.c_double, _VP], C.c_int),
        "sg_dot": ([_VP, C.c_int, _VP, _VP, C.POINTER(C.c_double)], C.c_int),
        "sg_nrm2": ([_VP, C.c_int, _VP, C.POINTER(C.c_double)], C.c_int),
        "sg_axpy": ([_VP, C.c_int, C.c_double, _VP, _VP], C.c_int),
        "sg_scal": ([_VP, C.c_int, C.c_double, _VP], C.c_int),
        "sg_precond_create": ([_VP, C.c_int, C.POINTER(PrecondOpts), C.POINTER(_VP)], C.c_int),
        "sg_precond_setup": ([_VP, _VP], C.c_int),
        "sg_precond_apply": ([_VP, _VP, _VP], C.c_int),
        "sg_precond_destroy": ([_VP], C.c_int),
        "sg_precond_create_callback": ([_VP, C.c_int, _VP, _VP, C.c_int, C.POINTER(_VP)], C.c_int),
        "sg_precond_partition": ([_VP, _VP], C.c_int),
        "sg_precond_subdomains": ([_VP, C.POINTER(C.c_int), _VP, _VP], C.c_int),
        "sg_precond_ilu_values": ([_VP, C.c_int, C.POINTER(C.c_int), _VP], C.c_int),
        "sg_precond_coarse": ([_VP, C.POINTER(C.c_int), _VP, _VP, _VP], C.c_int),
        "sg_solve": ([_VP, _VP, _VP, _VP, C.POINTER(SolveParams), C.POINTER(SolveResult), _VP,
                      C.c_int], C.c_int),
        "sg_partition": ([C.c_int, _VP, _VP, _VP, C.c_int, C.c_int, C.c_int, _VP], C.c_int),
        "sg_precond_create_learned": ([_VP, _VP, _VP, _VP, C.POINTER(LearnedOpts),
                                       C.POINTER(_VP)], C.c_int),
        "sg_precond_learned_stats": ([_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
                                     C.c_int),
        "sg_unet_nparams": ([], C.c_int),
        "sg_unet_create": ([_VP, C.POINTER(_VP)], C.c_int),
        "sg_unet_destroy": ([_VP], C.c_int),
        "sg_unet_set_params": ([_VP, _VP], C.c_int),
        "sg_unet_get_params": ([_VP, _VP], C.c_int),
        "sg_unet_forward": ([_VP, C.c_int, C.c_int, C.c_int, _VP, _VP], C.c_int),
        "sg_unet_loss_grad": ([_VP, C.c_int, C.c_int, C.c_int, _VP, _VP,
                               C.POINTER(C.c_double), _VP], C.c_int),
        "sg_unet_adam_step": ([_VP, C.c_float], C.c_int),
        "sg_unet_set_weight_decay": ([_VP, C.c_float], C.c_int),
        "sg_unet_set_loss": ([_VP, C.c_int], C.c_int),
        "sg_csr_create_batch": ([_VP, C.c_int, _VP, C.c_int, _VP, _VP, _VP, C.POINTER(_VP)],
                                C.c_int),
        "sg_csr_systems": ([_VP, C.POINTER(C.c_int)], C.c_int),
        "sg_solve_batch": ([_VP, _VP, _VP, _VP, C.POINTER(SolveParams), _VP, _VP, C.c_int],
                           C.c_int),
        "sg_precond_stats": ([_VP, C.POINTER(PrecondInfo)], C.c_int),
        "sg_precond_sweep_clock": ([_VP, _VP, _VP, _VP, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)], C.c_int),
        "sg_context_set_profiling": ([_VP, C.c_int], C.c_int),
        "sg_context_set_autotune": ([_VP, C.c_int], C.c_int),
        "sg_context_set_tune_cache": ([_VP, C.c_char_p], C.c_int),
        "sg_context_kernel_stats": ([_VP, C.c_int, C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_double)], C.c_int),
        "sg_context_reset_stats": ([_VP], C.c_int),
        "sg_context_kernel_units": ([_VP, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "sg_mm_read": ([C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                        C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int)),
                        C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_double))],
                       C.c_int),
        "sg_mm_free": ([_VP], None),
        "sg_precond_ilut_nnz": ([_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
                                C.c_int),
        "sg_estimate_condition": ([_VP, _VP, C.c_int, C.c_uint, C.POINTER(ConditionEstimate),
                                   _VP, _VP], C.c_int),
        "sg_estimate_spectral_radius": ([_VP, _VP, C.c_int, C.c_double, C.c_uint,
                                         C.POINTER(SpectralResult), _VP], C.c_int),
        "sg_convergence_analyze": ([_VP, C.c_int, C.c_int, C.POINTER(Convergence), _VP],
                                   C.c_int),
        "sg_convergence_write_jsonl": ([C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                        _VP, C.c_int, C.c_int], C.c_int),
        "sg_mm_write": ([C.c_char_p, C.c_int, C.c_int, _VP, _VP, _VP, C.c_int], C.c_int),
        "sg_solve_noalloc": ([_VP, _VP, _VP, _VP, C.POINTER(SolveParams), C.POINTER(SolveResult),
                              _VP, C.c_int], C.c_int),
        "sg_solve_last_basis": ([_VP, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), _VP, _VP, _VP], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/sg_capi.h that the binding knows."""
    return [n for n in dir(lib()) if n.startswith("sg_")]


def _p(a):
    return a.ctypes.data_as(_VP) if a is not None else None


class Context:
    """One HIP device + one stream (sg_context)."""

    def __init__(self, device=0):
        self.h = _VP()
        st = lib().sg_context_create(device, C.byref(self.h))
        if st:
            raise SgError(st, lib().sg_last_error(None).decode())

    def check(self, st):
        if st:
            raise SgError(st, lib().sg_last_error(self.h).decode())

    def sync(self):
        self.check(lib().sg_context_synchronize(self.h))

    def profiling(self, on=True):
        self.check(lib().sg_context_set_profiling(self.h, 1 if on else 0))

    def autotune(self, mode=1):
        """0 fixed shapes, 1 time shapes on first use, 2..11 force a shape (see
        sg_context_set_autotune)."""
        self.check(lib().sg_context_set_autotune(self.h, int(mode)))

    def tune_cache(self, path):
        self.check(lib().sg_context_set_tune_cache(self.h, str(path).encode() if path else None))

    def kernel_stats(self, family):
        n = C.c_longlong()
        ms = C.c_double()
        self.check(lib().sg_context_kernel_stats(self.h, K_FAMILY[family], C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_units(self, family):
        u = C.c_double()
        self.check(lib().sg_context_kernel_units(self.h, K_FAMILY[family], C.byref(u)))
        return u.value

    def reset_stats(self):
        self.check(lib().sg_context_reset_stats(self.h))

    def close(self):
        if self.h:
            lib().sg_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchCsr:
    """Block-diagonal batch of independent systems (sg_csr_create_batch).

    systems: list of (row_ptr, col_idx, vals) with local indices; they are
    concatenated with column offsets."""

    def __init__(self, ctx, systems):
        self.ctx = ctx
        rps, cis, vs, off = [], [], [], [0]
        nnz = 0
        for rp, ci, v in systems:
            rp = np.asarray(rp, np.int64)
            rps.append(rp[:-1] + nnz)
            cis.append(np.asarray(ci, np.int64) + off[-1])
            vs.append(np.asarray(v, np.float64))
            nnz += int(rp[-1])
            off.append(off[-1] + len(rp) - 1)
        self.rp = np.ascontiguousarray(np.concatenate(rps + [np.array([nnz])]), np.int32)
        self.ci = np.ascontiguousarray(np.concatenate(cis), np.int32)
        self.v = np.ascontiguousarray(np.concatenate(vs))
        self.sys_rows = np.asarray(off, np.int32)
        self.nsys = len(systems)
        self.n = int(off[-1])
        self.h = _VP()
        ctx.check(lib().sg_csr_create_batch(ctx.h, self.nsys, _p(self.sys_rows), nnz, _p(self.rp),
                                            _p(self.ci), _p(self.v), C.byref(self.h)))

    @property
    def nnz(self):
        return int(self.rp[-1])

    def spmv(self, x, y, alpha=1.0, beta=0.0):
        self.ctx.check(lib().sg_spmv(self.h, alpha, x.ptr, beta, y.ptr))

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().sg_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass


class UNet:
    """The learned stage's residual-to-error network (sg_unet)."""

    def __init__(self, ctx, params=None):
        self.ctx = ctx
        self.h = _VP()
        ctx.check(lib().sg_unet_create(ctx.h, C.byref(self.h)))
        self.nparams = lib().sg_unet_nparams()
        if params is not None:
            self.set_params(params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.nparams
        self.ctx.check(lib().sg_unet_set_params(self.h, _p(p)))

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx.check(lib().sg_unet_get_params(self.h, _p(p)))
        return p

    def forward(self, x):
        """x: host [B, 2, H, W] -> host [B, H, W]."""
        x = np.ascontiguousarray(x, np.float32)
        B, _, H, W = x.shape
        dx = DeviceArray.from_host(self.ctx, x.ravel())
        dy = DeviceArray(self.ctx, B * H * W, np.float32)
        self.forward_dev(B, H, W, dx, dy)
        return dy.get().reshape(B, H, W)

    def forward_dev(self, B, H, W, dx, dy):
        self.ctx.check(lib().sg_unet_forward(self.h, B, H, W, dx.ptr, dy.ptr))

    def set_loss(self, kind):
        """0: relative squared error against a target; 1: residual loss (no target);
        2: residual loss plus a term that holds the output's scale."""
        self.ctx.check(lib().sg_unet_set_loss(self.h, int(kind)))

    def loss_grad_dev(self, B, H, W, dx, dt, want_grad=False):
        loss = C.c_double()
        g = np.zeros(self.nparams, np.float32) if want_grad else None
        self.ctx.check(lib().sg_unet_loss_grad(self.h, B, H, W, dx.ptr,
                                               dt.ptr if dt is not None else None, C.byref(loss),
                                               _p(g)))
        return (loss.value, g) if want_grad else loss.value

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        B, _, H, W = x.shape
        dx = DeviceArray.from_host(self.ctx, x.ravel())
        dt = (DeviceArray.from_host(self.ctx, np.ascontiguousarray(t, np.float32).ravel())
              if t is not None else None)
        return self.loss_grad_dev(B, H, W, dx, dt, want_grad=True)

    def adam(self, lr):
        self.ctx.check(lib().sg_unet_adam_step(self.h, float(lr)))

    def set_weight_decay(self, wd):
        """Decoupled (AdamW) weight decay of the following Adam steps."""
        self.ctx.check(lib().sg_unet_set_weight_decay(self.h, float(wd)))

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().sg_unet_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LearnedStage:
    """Learned FGMRES stage (sg_precond_create_learned): base preconditioner +
    network correction with residual-reduction fallback.  base and net must
    outlive the stage (references are kept here)."""

    def __init__(self, ctx, base, net, log_kappa, nx, ny, min_gain=0.01, every=1, coarsen=1):
        self.ctx, self.base, self.net = ctx, base, net
        lk = np.ascontiguousarray(log_kappa, np.float64)
        assert lk.size == nx * ny
        o = LearnedOpts(nx, ny, min_gain, every, coarsen)
        self.h = _VP()
        ctx.check(lib().sg_precond_create_learned(ctx.h, base.h, net.h, _p(lk), C.byref(o),
                                                  C.byref(self.h)))
        self.n = nx * ny

    def setup(self, A):
        self.ctx.check(lib().sg_precond_setup(self.h, A.h))
        return self

    def apply(self, x, y):
        self.ctx.check(lib().sg_precond_apply(self.h, x.ptr, y.ptr))

    def counts(self):
        c, a = C.c_longlong(), C.c_longlong()
        self.ctx.check(lib().sg_precond_learned_stats(self.h, C.byref(c), C.byref(a)))
        return c.value, a.value

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().sg_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


def batch_partition(systems, nparts, ufactor=0, weighted=False):
    """Concatenated partition for a batch: system i gets parts
    [sum(nparts[:i]), sum(nparts[:i+1])) from the host partitioner."""
    out, base = [], 0
    for (rp, ci, v), k in zip(systems, nparts):
        out.append(partition(rp, ci, k, ufactor, vals=v, weighted=weighted) + base)
        base += k
    return np.ascontiguousarray(np.concatenate(out), np.int32)


class DeviceArray:
    """A device allocation holding a 1-D array (float64 by default)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = _VP()
        ctx.check(lib().sg_malloc(ctx.h, max(self.nbytes, 8), C.byref(self.ptr)))

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        self.ctx.check(lib().sg_memcpy_h2d(self.ctx.h, self.ptr, _p(a), self.nbytes))

    def get(self):
        out = np.empty(self.n, self.dtype)
        self.ctx.check(lib().sg_memcpy_d2h(self.ctx.h, _p(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ctx.check(lib().sg_memset_zero(self.ctx.h, self.ptr, self.nbytes))

    def free(self):
        if self.ptr:
            lib().sg_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.free()
        except Exception:
            pass


class Csr:
    """Device CSR matrix (sg_csr); host arrays are copied."""

    def __init__(self, ctx, rp, ci, v, ncols=None):
        self.ctx = ctx
        self.rp = np.ascontiguousarray(rp, np.int32)
        self.ci = np.ascontiguousarray(ci, np.int32)
        self.v = np.ascontiguousarray(v, np.float64)
        self.n = len(self.rp) - 1
        self.ncols = self.n if ncols is None else ncols
        self.h = _VP()
        ctx.check(lib().sg_csr_create(ctx.h, self.n, self.ncols, int(self.rp[-1]), _p(self.rp),
                                      _p(self.ci), _p(self.v), C.byref(self.h)))

    @property
    def nnz(self):
        return int(self.rp[-1])

    def spmv(self, x, y, alpha=1.0, beta=0.0):
        self.ctx.check(lib().sg_spmv(self.h, alpha, x.ptr, beta, y.ptr))

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().sg_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Precond:
    """Device preconditioner (sg_precond): kind in PC_KIND."""

    def __init__(self, ctx, kind, nparts=4, overlap=1, precision="fp64", partition=None,
                 ufactor=0, coarse_solve=0, weighted_partition=False, drop_tol=1e-4,
                 fill_factor=10):
        self.ctx = ctx
        self.kind = kind
        self._part = None if partition is None else np.ascontiguousarray(partition, np.int32)
        self.opts = PrecondOpts(nparts, overlap, 1 if precision == "fp32" else 0,
                                None if self._part is None else
                                self._part.ctypes.data_as(C.POINTER(C.c_int)),
                                ufactor, coarse_solve, 1 if weighted_partition else 0,
                                drop_tol, fill_factor)
        self.h = _VP()
        ctx.check(lib().sg_precond_create(ctx.h, PC_KIND[kind], C.byref(self.opts),
                                          C.byref(self.h)))
        self.n = None

    def setup(self, A):
        self.ctx.check(lib().sg_precond_setup(self.h, A.h))
        self.n = A.n
        return self

    def apply(self, x, y):
        self.ctx.check(lib().sg_precond_apply(self.h, x.ptr, y.ptr))

    def partition(self):
        out = np.zeros(self.n, np.int32)
        self.ctx.check(lib().sg_precond_partition(self.h, _p(out)))
        return out

    def subdomains(self):
        tot = C.c_int()
        self.ctx.check(lib().sg_precond_subdomains(self.h, C.byref(tot), None, None))
        nparts = self.opts.nparts if self.kind not in ("ilu0",) else 1
        ptr = np.zeros(nparts + 1, np.int32)
        idx = np.zeros(tot.value, np.int32)
        self.ctx.check(lib().sg_precond_subdomains(self.h, C.byref(tot), _p(ptr), _p(idx)))
        return [idx[ptr[s]:ptr[s + 1]] for s in range(nparts)]

    def ilut_nnz(self):
        """(l_nnz, u_nnz) of an ILUT preconditioner, as ILUTPrecond reports them."""
        a, b = C.c_longlong(), C.c_longlong()
        self.ctx.check(lib().sg_precond_ilut_nnz(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def ilu_values(self, s=0):
        nnz = C.c_int()
        self.ctx.check(lib().sg_precond_ilu_values(self.h, s, C.byref(nnz), None))
        out = np.zeros(nnz.value)
        self.ctx.check(lib().sg_precond_ilu_values(self.h, s, C.byref(nnz), _p(out)))
        return out

    def stats(self):
        info = PrecondInfo()
        self.ctx.check(lib().sg_precond_stats(self.h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in PrecondInfo._fields_}

    def sweep_clock(self, x, y):
        """One apply y = M^{-1} x with the staged sweep's workgroup clock armed.

        Returns a dict of per-workgroup arrays (launch order): start, gathered
        (right-hand side in LDS), swept (both triangular sweeps done) and end
        in microseconds from the first start, subdomain, hardware id (HIP's
        __smid: XCC, SE and CU ids packed), rows, forward and backward levels,
        chunks, owned rows;
        only the workgroups that recorded.  Raises SgError when the sweep in
        use is not the staged kernel."""
        nsub = self.stats()["nsub"]
        rec = np.zeros(12 * nsub, np.int64)
        nrec, khz = C.c_int(), C.c_double()
        self.ctx.check(lib().sg_precond_sweep_clock(self.h, x.ptr, y.ptr, _p(rec), len(rec),
                                                    C.byref(nrec), C.byref(khz)))
        if nrec.value == 0:
            raise SgError(7, "sweep clock: the staged sweep is not the one in use")
        r = rec.reshape(nsub, 12)
        r = r[r[:, 3] != 0]
        t0 = r[:, 0].min()
        us = 1e3 / khz.value
        return dict(start_us=(r[:, 0] - t0) * us, gathered_us=(r[:, 1] - t0) * us,
                    swept_us=(r[:, 2] - t0) * us, end_us=(r[:, 3] - t0) * us, sub=r[:, 4],
                    hwid=r[:, 5], rows=r[:, 6], flev=r[:, 7], blev=r[:, 8], chunks=r[:, 9],
                    owned=r[:, 10], rate_khz=khz.value)

    def coarse(self):
        nc = C.c_int()
        self.ctx.check(lib().sg_precond_coarse(self.h, C.byref(nc), None, None, None))
        k = nc.value
        A0 = np.zeros(k * k)
        LU = np.zeros(k * k)
        piv = np.zeros(k, np.int32)
        self.ctx.check(lib().sg_precond_coarse(self.h, C.byref(nc), _p(A0), _p(LU), _p(piv)))
        return A0.reshape(k, k), LU.reshape(k, k), piv

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().sg_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


def partition(rp, ci, nparts, ufactor=0, vals=None, weighted=False):
    """Host graph partition (sg_partition); no GPU needed."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, np.float64)
    n = len(rp) - 1
    part = np.zeros(n, np.int32)
    st = lib().sg_partition(n, _p(rp), _p(ci), _p(v), nparts, ufactor, 1 if weighted else 0,
                            _p(part))
    if st:
        raise SgError(st, lib().sg_last_error(None).decode())
    return part


SOLVE_GRAPH = 1
SOLVE_NO_GRAPH = 2


MM_SUM_DUPLICATES = 1
MM_STRICT = 2


def read_matrix_market(path, sum_duplicates=False, strict=False):
    """sg_mm_read: (nrows, ncols, row_ptr, col_idx, vals) as numpy arrays.
    Host only; no GPU needed."""
    L = lib()
    nr, nc, nz = C.c_int(), C.c_int(), C.c_int()
    rp, ci = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    v = C.POINTER(C.c_double)()
    flags = (MM_SUM_DUPLICATES if sum_duplicates else 0) | (MM_STRICT if strict else 0)
    st = L.sg_mm_read(os.fsencode(path), flags, C.byref(nr), C.byref(nc), C.byref(nz),
                      C.byref(rp), C.byref(ci), C.byref(v))
    if st:
        raise SgError(st, L.sg_last_error(None).decode())
    try:
        n, k = nr.value, nz.value
        out = (n, nc.value, np.ctypeslib.as_array(rp, (n + 1,)).copy(),
               np.ctypeslib.as_array(ci, (max(k, 1),))[:k].copy(),
               np.ctypeslib.as_array(v, (max(k, 1),))[:k].copy())
    finally:
        for a in (rp, ci, v):
            L.sg_mm_free(C.cast(a, _VP))
    return out


def write_matrix_market(path, rp, ci, v, ncols=None, digits=6):
    """sg_mm_write (digits=6 writes as the reference; 17 round-trips)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    st = lib().sg_mm_write(os.fsencode(path), n, n if ncols is None else ncols, _p(rp), _p(ci),
                           _p(v), digits)
    if st:
        raise SgError(st, lib().sg_last_error(None).decode())


def estimate_condition(A, M=None, k=80, seed=42):
    """sg_estimate_condition: Lanczos estimate of kappa(M^-1 A) on the device.
    Returns a dict with lambda_min, lambda_max, kappa, lanczos_steps, valid and
    the tridiagonal coefficients alpha, beta."""
    e = ConditionEstimate()
    kk = max(k, 2)
    al, be = np.zeros(kk), np.zeros(kk)
    A.ctx.check(lib().sg_estimate_condition(A.h, M.h if M is not None else None, k, seed,
                                            C.byref(e), _p(al), _p(be)))
    m = e.lanczos_steps
    return dict(lambda_min=e.lambda_min, lambda_max=e.lambda_max, kappa=e.kappa,
                lanczos_steps=e.lanczos_steps, valid=bool(e.valid), alpha=al[:m].copy(),
                beta=be[:max(m - 1, 0)].copy())


def estimate_spectral_radius(A, M=None, max_iter=200, tol=1e-6, seed=123, eigvec=None):
    """sg_estimate_spectral_radius: power iteration on I - M^-1 A on the device."""
    r = SpectralResult()
    A.ctx.check(lib().sg_estimate_spectral_radius(
        A.h, M.h if M is not None else None, max_iter, tol, seed, C.byref(r),
        eigvec.ptr if eigvec is not None else None))
    return dict(spectral_radius=r.spectral_radius, convergence_ok=bool(r.convergence_ok),
                power_iters=r.power_iters)


def convergence_metrics(hist, converged):
    """sg_convergence_analyze (host only)."""
    h = np.ascontiguousarray(hist, np.float64)
    c = Convergence()
    rates = np.zeros(max(len(h), 1))
    st = lib().sg_convergence_analyze(_p(h), len(h), 1 if converged else 0, C.byref(c),
                                      _p(rates))
    if st:
        raise SgError(st, lib().sg_last_error(None).decode())
    return dict(initial_residual=c.initial_residual, final_residual=c.final_residual,
                asymptotic_rate=c.asymptotic_rate, stagnation_start=c.stagnation_start,
                converged=bool(c.converged), iterations=c.iterations, rates=rates[:len(h)])


def convergence_write_jsonl(path, matrix, precond, nparts, overlap, hist, converged):
    """sg_convergence_write_jsonl (host only): appends the "convergence" line."""
    h = np.ascontiguousarray(hist, np.float64)
    st = lib().sg_convergence_write_jsonl(os.fsencode(path), matrix.encode(), precond.encode(),
                                          nparts, overlap, _p(h), len(h),
                                          1 if converged else 0)
    if st:
        raise SgError(st, lib().sg_last_error(None).decode())


METHODS = {"gmres": 0, "fgmres": 1, "bicgstab": 2}


def _result(r, hist, cap):
    return dict(iterations=r.iterations, converged=bool(r.converged),
                residual_norm=r.residual_norm, rel_true_residual=r.rel_true_residual,
                time_solve_s=r.time_solve_s, cycles=r.cycles,
                precond_applies=r.precond_applies,
                hist=hist[:min(r.hist_len, cap)].copy())


def solve(A, b, x, M=None, method="gmres", restart=30, max_iter=1000, tol=1e-10, rtol=0.0,
          ortho="mgs", lookahead=0, hist_cap=None, graph=False):
    """Run sg_solve.  b, x are DeviceArrays (x in/out).  Returns a dict shaped
    like GMRESResult plus the FP64 true relative residual.  method: "gmres"
    (left), "fgmres" or "bicgstab" (tol relative to ||r0|| there)."""
    p = SolveParams(METHODS[method], restart, max_iter, tol, rtol,
                    1 if ortho == "cgs2" else 0, lookahead, SOLVE_GRAPH if graph else 0)
    r = SolveResult()
    cap = hist_cap if hist_cap is not None else max_iter + restart + 2
    hist = np.zeros(max(cap, 1))
    A.ctx.check(lib().sg_solve(A.h, M.h if M is not None else None, b.ptr, x.ptr, C.byref(p),
                               C.byref(r), _p(hist), cap))
    return _result(r, hist, cap)


def solve_batch(A, b, x, M=None, method="fgmres", restart=30, max_iter=1000, tol=1e-10,
                rtol=0.0, ortho="cgs2", graph=True, hist_cap=None):
    """Run sg_solve_batch on a BatchCsr: one result dict per system."""
    p = SolveParams(METHODS[method], restart, max_iter, tol, rtol,
                    1 if ortho == "cgs2" else 0, 0, 0 if graph else SOLVE_NO_GRAPH)
    nsys = A.nsys
    rs = (SolveResult * nsys)()
    cap = hist_cap if hist_cap is not None else max_iter + restart + 2
    hist = np.zeros((nsys, max(cap, 1)))
    A.ctx.check(lib().sg_solve_batch(A.h, M.h if M is not None else None, b.ptr, x.ptr,
                                     C.byref(p), C.cast(rs, _VP), _p(hist), hist.shape[1]))
    return [_result(rs[q], hist[q], cap) for q in range(nsys)]


def last_basis(ctx, with_z=True):
    """The Krylov basis of the last cycle of the last single-system GMRES/FGMRES
    solve on ctx (sg_solve_last_basis): dict with V (n x kv), Z (n x k, or None),
    Hbar ((m+1) x m, before the Givens rotations), k and kv.  with_z=False
    copies V only (a transposed view; Z and Hbar None)."""
    n, m, k, kv = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    L = lib()
    ctx.check(L.sg_solve_last_basis(ctx.h, C.byref(n), C.byref(m), C.byref(k), C.byref(kv),
                                    None, None, None))
    V = np.zeros((kv.value, n.value))
    Z = np.zeros((k.value, n.value)) if with_z else None
    H = np.zeros((m.value, m.value + 1)) if with_z else None  # column-major (m+1) x m
    ctx.check(L.sg_solve_last_basis(ctx.h, C.byref(n), C.byref(m), C.byref(k), C.byref(kv),
                                    _p(V), _p(Z) if with_z else None, _p(H) if with_z else None))
    return dict(V=V.T.copy() if with_z else V.T, Z=Z.T.copy() if with_z else None,
                Hbar=H.T.copy() if with_z else None, k=k.value, kv=kv.value,
                m=m.value, n=n.value)
