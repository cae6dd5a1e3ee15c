"""ctypes binding of libsfdx.so and a Python mirror of the reference's host API.

The reference's Fortran entry points (sfd_excmg/src) map to:
  fwd3d::fwd_solver   (fdfe_main.f90:1565-1594) -> fwd_solver()
  fwd3d::post_process (fdfe_main.f90:549-716)   -> post_process()
  model_selection::mt1d_anl (model_selec.f90:266-401) -> mt1d_anl()
and the kernel-level pieces (zamux on the assembled operator, SSOR, div_correc,
extint_v) to apply_operator(), ssor_apply(), div_correct(), extint_v().

Nothing here computes on the CPU as a fallback: every call goes through the HIP
library, and importing without a built library raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libsfdx.so")

SOLVER_BICGSTAB, SOLVER_GPBICG, SOLVER_COCR = 1, 3, 5
PREC_FP64, PREC_MIXED, PREC_MIXED_IR, PREC_MIXED_SPLIT = 0, 1, 2, 3
NORM_REFERENCE, NORM_INTERIOR = 0, 1
PRECOND_SSOR, PRECOND_HYBRID, PRECOND_MG, PRECOND_AUTO = 0, 1, 2, 3
MAX_LEVELS = 8
WARN_NOCONV = -1


class SfdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sfdx error {code}: {msg}")
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nair", ctypes.c_int),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p)]


class Opts(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int), ("use_excmg", ctypes.c_int), ("solver", ctypes.c_int),
                ("eps", ctypes.c_double), ("corr_intv", ctypes.c_int), ("corr_tol", ctypes.c_double),
                ("ssor_omega", ctypes.c_double), ("precision", ctypes.c_int), ("norm", ctypes.c_int),
                ("maxit_plain", ctypes.c_int), ("maxit_excmg", ctypes.c_int), ("inner_tol", ctypes.c_double),
                ("max_refine", ctypes.c_int), ("verbose", ctypes.c_int), ("dc_restart", ctypes.c_int),
                ("dc_at_start", ctypes.c_int), ("inner_maxit", ctypes.c_int), ("precond", ctypes.c_int),
                ("sigma_floor32", ctypes.c_double), ("mg_levels", ctypes.c_int), ("mg_coarse_pairs", ctypes.c_int),
                ("restart_drop", ctypes.c_double), ("use_graph", ctypes.c_int),
                ("nodal_smoother", ctypes.c_int), ("coarse_solver", ctypes.c_int),
                ("full_background", ctypes.c_int), ("spmv_tile", ctypes.c_int), ("ssor_cross_mass", ctypes.c_int),
                ("surface_sigma", ctypes.c_double), ("pol_streams", ctypes.c_int),
                ("mg_merge_ratio", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("total_seconds", ctypes.c_double), ("fine_seconds", ctypes.c_double),
                ("fine_bytes", ctypes.c_double), ("iters", (ctypes.c_int * 2) * MAX_LEVELS),
                ("relres", (ctypes.c_double * 2) * MAX_LEVELS), ("dc_calls", (ctypes.c_int * 2) * MAX_LEVELS),
                ("pcg_iters", (ctypes.c_int * 2) * MAX_LEVELS), ("refine_steps", ctypes.c_int * 2),
                ("fine_iters_total", ctypes.c_int), ("fallback_fp64", ctypes.c_int),
                ("solver_switches", ctypes.c_int), ("spmv_tile", ctypes.c_int), ("mg_freqs", ctypes.c_int)]

    def to_dict(self, levels=MAX_LEVELS):
        return dict(status=self.status, total_seconds=self.total_seconds, fine_seconds=self.fine_seconds,
                    fine_bytes=self.fine_bytes,
                    iters=[[self.iters[i][p] for p in range(2)] for i in range(levels)],
                    relres=[[self.relres[i][p] for p in range(2)] for i in range(levels)],
                    dc_calls=[[self.dc_calls[i][p] for p in range(2)] for i in range(levels)],
                    pcg_iters=[[self.pcg_iters[i][p] for p in range(2)] for i in range(levels)],
                    refine_steps=[self.refine_steps[0], self.refine_steps[1]],
                    fine_iters_total=self.fine_iters_total, fallback_fp64=self.fallback_fp64,
                    solver_switches=self.solver_switches, spmv_tile=self.spmv_tile, mg_freqs=self.mg_freqs)


class Job(ctypes.Structure):
    _fields_ = [("grid", Grid), ("sigma", ctypes.c_void_p), ("nf", ctypes.c_int), ("freq", ctypes.c_void_p),
                ("ex", ctypes.c_void_p), ("level", ctypes.c_int)]


EXPORTS = ["sfdx_create", "sfdx_destroy", "sfdx_last_error", "sfdx_default_opts", "sfdx_fwd_solve",
           "sfdx_post_process", "sfdx_mt1d_anl", "sfdx_apply_operator", "sfdx_ssor_apply", "sfdx_hybrid_apply", "sfdx_div_correct",
           "sfdx_extint_v", "sfdx_bench_kernels", "sfdx_direct_solve", "sfdx_solve_batch", "sfdx_misfit",
           "sfdx_adjoint", "sfdx_fwd_response", "sfdx_post_process_topo", "sfdx_surface_planes",
           "sfdx_mesh_coarsen", "sfdx_apply_operator_noalloc", "sfdx_split_residual", "sfdx_mesh_coarsen_device"]

_lib = None


def load(path: str = LIB_PATH):
    """Load libsfdx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run `python -m sfd_excmg_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.sfdx_create.argtypes = [I, ctypes.POINTER(P)]
    L.sfdx_destroy.argtypes = [P]
    L.sfdx_last_error.restype = ctypes.c_char_p
    L.sfdx_default_opts.argtypes = [ctypes.POINTER(Opts)]
    L.sfdx_fwd_solve.argtypes = [P, ctypes.POINTER(Grid), P, I, P, ctypes.POINTER(Opts), P, ctypes.POINTER(Stats)]
    L.sfdx_post_process.argtypes = [ctypes.POINTER(Grid), I, P, I, P, P, I, I, P, P]
    L.sfdx_post_process_topo.argtypes = [ctypes.POINTER(Grid), P, I, P, I, P, P, I, I, P, P]
    L.sfdx_surface_planes.argtypes = [ctypes.POINTER(Grid), P, D, P]
    L.sfdx_mesh_coarsen.argtypes = [ctypes.POINTER(Grid), P, I, P, P, P, P]
    L.sfdx_mesh_coarsen_device.argtypes = [P, ctypes.POINTER(Grid), P, I, P, P, P, P]
    L.sfdx_mt1d_anl.argtypes = [I, P, P, D, P, P, P]
    L.sfdx_apply_operator.argtypes = [P, ctypes.POINTER(Grid), P, D, I, P, P]
    L.sfdx_apply_operator_noalloc.argtypes = [P, ctypes.POINTER(Grid), P, D, P, P, P]
    L.sfdx_split_residual.argtypes = [P, ctypes.POINTER(Grid), P, D, I, P, P, P, P]
    L.sfdx_ssor_apply.argtypes = [P, ctypes.POINTER(Grid), P, D, D, I, P, P]
    L.sfdx_hybrid_apply.argtypes = [P, ctypes.POINTER(Grid), P, D, D, I, I, P, P]
    L.sfdx_div_correct.argtypes = [P, ctypes.POINTER(Grid), P, P, D, ctypes.POINTER(I)]
    L.sfdx_extint_v.argtypes = [P, ctypes.POINTER(Grid), P, P, P]
    L.sfdx_bench_kernels.argtypes = [P, ctypes.POINTER(Grid), P, D, I, I, P]
    L.sfdx_direct_solve.argtypes = [ctypes.POINTER(Grid), P, D, P]
    L.sfdx_solve_batch.argtypes = [P, I, ctypes.POINTER(Job), ctypes.POINTER(Opts), I, ctypes.POINTER(Stats)]
    L.sfdx_fwd_response.argtypes = [P, ctypes.POINTER(Grid), P, I, P, ctypes.POINTER(Opts), I, P, I, I, P, P,
                                    ctypes.POINTER(Stats)]
    L.sfdx_misfit.argtypes = [ctypes.POINTER(Grid), I, P, I, P, I, P, P, ctypes.POINTER(D), P]
    L.sfdx_adjoint.argtypes = [P, ctypes.POINTER(Grid), P, I, P, I, P, I, P, ctypes.POINTER(Opts), ctypes.POINTER(D), P,
                               P, ctypes.POINTER(Stats)]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(code):
    if code > 0:
        raise SfdxError(code, load().sfdx_last_error().decode())
    return code


def sigma_flat(sigma):
    """sigma(ncell, 6) -> sigma[comp*ncell + ie] (Fortran column-major)."""
    s = np.asarray(sigma, dtype=np.float64)
    return np.ascontiguousarray(s.T).reshape(-1)


def nedges(nx, ny, nz):
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


class _GridHolder:
    def __init__(self, a, b, c, nair=0):
        self.a = np.ascontiguousarray(a, np.float64)
        self.b = np.ascontiguousarray(b, np.float64)
        self.c = np.ascontiguousarray(c, np.float64)
        self.g = Grid(len(self.a), len(self.b), len(self.c), int(nair), _ptr(self.a), _ptr(self.b), _ptr(self.c))


def default_opts(**kw) -> Opts:
    o = Opts()
    load().sfdx_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    """One device context (stream + workspace)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = ctypes.c_void_p()
        _check(self.lib.sfdx_create(device, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.sfdx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ solve
    def fwd_solve(self, a, b, c, nair, sigma, freqs, opts: Opts | None = None):
        G = _GridHolder(a, b, c, nair)
        sf = sigma_flat(sigma)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        nl = nedges(len(G.a), len(G.b), len(G.c))
        ex = np.zeros((len(fr), 2 * nl), np.complex128)
        st = Stats()
        o = opts if opts is not None else default_opts()
        code = self.lib.sfdx_fwd_solve(self.h, ctypes.byref(G.g), _ptr(sf), len(fr), _ptr(fr), ctypes.byref(o),
                                       _ptr(ex), ctypes.byref(st))
        _check(code)
        return ex.T.copy(), st   # ex(2*nl, nf) like the reference

    def solve_batch(self, models, opts: Opts | None = None, nstreams: int = 0):
        """Batched forward solves (sfdx_solve_batch): every (model, frequency)
        pair runs as an independent item on one of `nstreams` concurrent HIP
        streams.  `models`: objects with a, b, c, nair, sigma, freqs and an
        optional level.  Returns ([ex(2*nl, nf) per model], [Stats per model])."""
        models = list(models)
        n = len(models)
        jobs = (Job * n)()
        keep, outs = [], []
        for k, m in enumerate(models):
            G = _GridHolder(m.a, m.b, m.c, m.nair)
            sf = sigma_flat(m.sigma)
            fr = np.ascontiguousarray(np.atleast_1d(m.freqs), np.float64)
            ex = np.zeros((len(fr), 2 * nedges(len(G.a), len(G.b), len(G.c))), np.complex128)
            jobs[k].grid = G.g
            jobs[k].sigma = _ptr(sf)
            jobs[k].nf = len(fr)
            jobs[k].freq = _ptr(fr)
            jobs[k].ex = _ptr(ex)
            jobs[k].level = int(getattr(m, "level", 0) or 0)
            keep.append((G, sf, fr))
            outs.append(ex)
        st = (Stats * n)()
        o = opts if opts is not None else default_opts()
        _check(self.lib.sfdx_solve_batch(self.h, n, jobs, ctypes.byref(o), int(nstreams), st))
        return [ex.T.copy() for ex in outs], list(st)

    def fwd_response(self, a, b, c, nair, sigma, freqs, recv, opts: Opts | None = None, out_type: int = 1,
                     rd: int = 2):
        """Forward solve + station responses with only the surface fields
        leaving the device (sfdx_fwd_response); returns (results, tipper, stats)
        shaped like post_process."""
        G = _GridHolder(a, b, c, nair)
        sf = sigma_flat(sigma)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        rv = np.ascontiguousarray(recv, np.float64)
        res = np.zeros((len(fr) * len(rv), 4))
        tip = np.zeros((len(fr) * len(rv), 4))
        st = Stats()
        o = opts if opts is not None else default_opts()
        _check(self.lib.sfdx_fwd_response(self.h, ctypes.byref(G.g), _ptr(sf), len(fr), _ptr(fr), ctypes.byref(o),
                                          len(rv), _ptr(rv), int(out_type), int(rd), _ptr(res), _ptr(tip),
                                          ctypes.byref(st)))
        return res, tip, st

    def adjoint(self, a, b, c, nair, sigma, freqs, recv, obs, opts: Opts | None = None, rd: int = 2,
                return_fields: bool = False):
        """Misfit and d misfit / d log(s_c) for every cell (sfdx_adjoint).
        obs: (nf*nrec, 4) rows rho_xy, rho_yx, phi_xy, phi_yx (ohm-m, degrees)."""
        G = _GridHolder(a, b, c, nair)
        sf = sigma_flat(sigma)
        fr = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
        rv = np.ascontiguousarray(recv, np.float64)
        ob = np.ascontiguousarray(obs, np.float64)
        assert ob.shape == (len(fr) * len(rv), 4)
        ncell = len(G.a) * len(G.b) * len(G.c)
        grad = np.zeros(ncell)
        phi = ctypes.c_double(0.0)
        ex = np.zeros((len(fr), 2 * nedges(len(G.a), len(G.b), len(G.c))), np.complex128) if return_fields else None
        st = Stats()
        o = opts if opts is not None else default_opts()
        _check(self.lib.sfdx_adjoint(self.h, ctypes.byref(G.g), _ptr(sf), len(fr), _ptr(fr), len(rv), _ptr(rv), int(rd),
                                     _ptr(ob), ctypes.byref(o), ctypes.byref(phi), _ptr(grad),
                                     _ptr(ex) if ex is not None else None, ctypes.byref(st)))
        if return_fields:
            return phi.value, grad, ex.T.copy(), st
        return phi.value, grad, st

    def mesh_coarsen(self, a, b, c, sigma, k):
        """mesh_coarsen with sigma averaged on the device (sfdx_mesh_coarsen_device):
        widths and sigma(ncell, 6) of the grid coarsened by 2^k."""
        G = _GridHolder(a, b, c)
        f = 1 << int(k)
        ac, bc, cc = np.zeros(len(G.a) // f), np.zeros(len(G.b) // f), np.zeros(len(G.c) // f)
        sc = np.zeros(6 * len(ac) * len(bc) * len(cc))
        _check(self.lib.sfdx_mesh_coarsen_device(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), int(k), _ptr(ac),
                                                 _ptr(bc), _ptr(cc), _ptr(sc)))
        return ac, bc, cc, sc.reshape(6, -1).T.copy()

    def apply_operator(self, a, b, c, sigma, freq, x, precision=PREC_FP64):
        G = _GridHolder(a, b, c)
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros_like(x)
        _check(self.lib.sfdx_apply_operator(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq),
                                            int(precision), _ptr(x), _ptr(y)))
        return y

    def split_residual(self, a, b, c, sigma, freq, r, z, psi, precision=PREC_FP64):
        """r - A_s (z + (i/wmu) G psi) on interior edges (0 on boundary edges)."""
        G = _GridHolder(a, b, c)
        r = np.ascontiguousarray(r, np.complex128)
        z = np.ascontiguousarray(z, np.complex128)
        psi = np.ascontiguousarray(psi, np.complex128)
        nn = (len(a) + 1) * (len(b) + 1) * (len(c) + 1)
        if r.shape != z.shape or psi.shape != (nn,):
            raise ValueError("split_residual: r, z need nl entries and psi one per node")
        t = np.zeros_like(r)
        _check(self.lib.sfdx_split_residual(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq),
                                            int(precision), _ptr(r), _ptr(z), _ptr(psi), _ptr(t)))
        return t

    def ssor_apply(self, a, b, c, sigma, freq, x, omega=1.0, precision=PREC_FP64):
        G = _GridHolder(a, b, c)
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros_like(x)
        _check(self.lib.sfdx_ssor_apply(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq),
                                        float(omega), int(precision), _ptr(x), _ptr(y)))
        return y

    def hybrid_apply(self, a, b, c, sigma, freq, x, omega=1.0, nodal=1, precision=PREC_FP64):
        """The solvers' hybrid preconditioner applied to x (boundary entries 0)."""
        G = _GridHolder(a, b, c)
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros_like(x)
        _check(self.lib.sfdx_hybrid_apply(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq),
                                          float(omega), int(nodal), int(precision), _ptr(x), _ptr(y)))
        return y

    def div_correct(self, a, b, c, sigma, x, corr_tol=1e-3):
        G = _GridHolder(a, b, c)
        x = np.ascontiguousarray(x, np.complex128).copy()
        it = ctypes.c_int(0)
        _check(self.lib.sfdx_div_correct(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), _ptr(x),
                                         float(corr_tol), ctypes.byref(it)))
        return x, it.value

    def extint_v(self, a, b, c, nair, u0, u1):
        G = _GridHolder(a, b, c, nair)
        u0 = np.ascontiguousarray(u0, np.complex128)
        u1 = np.ascontiguousarray(u1, np.complex128)
        out = np.zeros(nedges(len(G.a), len(G.b), len(G.c)), np.complex128)
        _check(self.lib.sfdx_extint_v(self.h, ctypes.byref(G.g), _ptr(u0), _ptr(u1), _ptr(out)))
        return out

    def bench_kernels(self, a, b, c, sigma, freq, precision=PREC_FP64, reps=20):
        G = _GridHolder(a, b, c)
        ms = np.zeros(14)
        _check(self.lib.sfdx_bench_kernels(self.h, ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq),
                                           int(precision), int(reps), _ptr(ms)))
        return ms


def misfit(a, b, c, nair, freqs, recv, ex, obs, rd=2, with_grad=False):
    """Data misfit of the fields ex(2*nl, nf) and optionally its field gradient
    w (d misfit = Re(w^T d ex)), host FP64 (sfdx_misfit)."""
    G = _GridHolder(a, b, c, nair)
    fr = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
    rv = np.ascontiguousarray(recv, np.float64)
    ob = np.ascontiguousarray(obs, np.float64)
    e = np.ascontiguousarray(np.asarray(ex).T, np.complex128)   # (nf, 2*nl)
    w = np.zeros_like(e) if with_grad else None
    phi = ctypes.c_double(0.0)
    _check(load().sfdx_misfit(ctypes.byref(G.g), len(fr), _ptr(fr), len(rv), _ptr(rv), int(rd), _ptr(ob), _ptr(e),
                              ctypes.byref(phi), _ptr(w) if w is not None else None))
    return (phi.value, w.T.copy()) if with_grad else phi.value


# ---------------------------------------------------------------- reference-shaped API
def fwd_solver(modelname, a, b, c, nair, level, slv, eps, sigma, freq, ctx: Context | None = None, **opt):
    """Mirror of fwd3d::fwd_solver(modelname,a,b,c,nair,level,slv,eps,sigma,freq,ex)
    (fdfe_main.f90:1565): returns ex(2*nl, nf).  Extra keyword options fill
    sfdx_opts (use_excmg, corr_intv, precision, norm, ...)."""
    own = ctx is None
    ctx = ctx or Context()
    try:
        o = default_opts(level=int(level), solver=int(slv), eps=float(eps))
        for k, v in opt.items():
            setattr(o, k, v)
        ex, st = ctx.fwd_solve(a, b, c, nair, sigma, freq, o)
        return ex, st
    finally:
        if own:
            ctx.close()


def post_process(a, b, c, nair, freqs, recv, ex, out_type=1, rd=2, surf=None):
    """Mirror of fwd3d::post_process (fdfe_main.f90:549-716): returns
    results(nrec*nf, 4) and tipper(nrec*nf, 4).  surf (optional, (nx, ny)
    integer node planes, see surface_planes) is the topography branch
    (fdfe_main.f90:604-614, 656-666)."""
    G = _GridHolder(a, b, c, nair)
    fr = np.ascontiguousarray(np.atleast_1d(freqs), np.float64)
    recv = np.ascontiguousarray(recv, np.float64)
    exf = np.ascontiguousarray(np.asarray(ex).T, np.complex128)  # (nf, 2nl)
    nrec = recv.shape[0]
    res = np.zeros((nrec * len(fr), 4))
    tip = np.zeros((nrec * len(fr), 4))
    if surf is None:
        _check(load().sfdx_post_process(ctypes.byref(G.g), len(fr), _ptr(fr), nrec, _ptr(recv), _ptr(exf),
                                         int(out_type), int(rd), _ptr(res), _ptr(tip)))
    else:
        sp = np.ascontiguousarray(np.asarray(surf, np.int32).T).reshape(-1)   # index i + nx*j
        _check(load().sfdx_post_process_topo(ctypes.byref(G.g), _ptr(sp), len(fr), _ptr(fr), nrec, _ptr(recv),
                                              _ptr(exf), int(out_type), int(rd), _ptr(res), _ptr(tip)))
    return res, tip


def surface_planes(a, b, c, sigma, air_max):
    """Surface node plane of every cell column, (nx, ny) int32, of a model with
    topography carved into sigma (sfdx_surface_planes)."""
    G = _GridHolder(a, b, c)
    out = np.zeros(len(G.a) * len(G.b), np.int32)
    _check(load().sfdx_surface_planes(ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(air_max), _ptr(out)))
    return out.reshape(len(G.b), len(G.a)).T.copy()


def mesh_coarsen(a, b, c, sigma, k):
    """Mirror of mesh_coarsen (fdfe_main.f90:18-84): widths and sigma(ncell, 6)
    of the grid coarsened by 2^k (sfdx_mesh_coarsen)."""
    G = _GridHolder(a, b, c)
    f = 1 << int(k)
    ac, bc, cc = np.zeros(len(G.a) // f), np.zeros(len(G.b) // f), np.zeros(len(G.c) // f)
    sc = np.zeros(6 * len(ac) * len(bc) * len(cc))
    _check(load().sfdx_mesh_coarsen(ctypes.byref(G.g), _ptr(sigma_flat(sigma)), int(k), _ptr(ac), _ptr(bc), _ptr(cc),
                                    _ptr(sc)))
    return ac, bc, cc, sc.reshape(6, -1).T.copy()


def direct_solve(a, b, c, sigma, freq):
    """Host direct solve of one level (both polarisations): u(2*nl)."""
    G = _GridHolder(a, b, c)
    u = np.zeros(2 * nedges(len(G.a), len(G.b), len(G.c)), np.complex128)
    _check(load().sfdx_direct_solve(ctypes.byref(G.g), _ptr(sigma_flat(sigma)), float(freq), _ptr(u)))
    return u


def mt1d_anl(freq, conv, c):
    conv = np.asarray(conv, np.float64)
    nz = conv.shape[0]
    cf = np.ascontiguousarray(conv.T).reshape(-1)
    c = np.ascontiguousarray(c, np.float64)
    ex = np.zeros(2 * nz + 2, np.complex128)
    hx = np.zeros(2 * nz + 2, np.complex128)
    imp = np.zeros(4, np.complex128)
    _check(load().sfdx_mt1d_anl(nz, _ptr(cf), _ptr(c), float(freq), _ptr(ex), _ptr(hx), _ptr(imp)))
    return ex, hx, imp
