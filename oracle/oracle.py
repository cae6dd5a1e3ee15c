"""FP64 CPU oracle for the SFD_EXCMG forward solve -- TEST INFRASTRUCTURE ONLY.

This module restates the reference's CPU algorithms (CSR assembly, lexicographic
SSOR, BiCGStab/GPBiCG/COCR, divergence correction, EXCMG level loop, 1-D
background, post-process).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, as the checker or as the
timed CPU baseline.  The product package ``sfd_excmg_amd`` never imports it.

The heavy loops live in ``sfd_oracle.c`` (built by ``make -C oracle``); the
level loop, the coarse direct solves and the post-process live here.  Coarse
direct solves use SciPy's SuperLU where the reference calls MKL PARDISO
(``src/solvers.f90:765-875``); a direct solve is pinned by its residual.

Function map (reference ``sfd_excmg/src`` file:line):
  mt1d_anl        model_selec.f90:266-401
  build_system    fdfe_main.f90:226-546      (fd_assemb + dcfast)
  krylov          solvers.f90:522-719 (BiCGStab), 137-315 (GPBiCG), COCR new
  mesh_coarsen    fdfe_main.f90:18-84        (index bug fixed)
  extint_v        excmg.f90:9-77, 357-772
  solve_level     fdfe_main.f90:892-1063     (doublemode_nor)
  excmg           fdfe_main.f90:1228-1563    (u1u2 + doublemode_acc)
  fwd_solver      fdfe_main.f90:1565-1594
  post_process    fdfe_main.f90:549-790      (+ arho_interp2; topography branch :604-614, 656-666)
  surface_planes  fdfe_main.f90:604-614      (first cell below the ground, from carved sigma)
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
MU0 = 4.0 * np.pi * 1e-7

SOLVER_BICGSTAB, SOLVER_GPBICG, SOLVER_COCR = 1, 3, 5


def build(force: bool = False) -> str:
    """Compile the C part of the oracle with gcc (no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "-s"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.orc_nedges.restype = i64
        L.orc_nedges.argtypes = [i32, i32, i32]
        L.orc_nnodes.restype = i64
        L.orc_nnodes.argtypes = [i32, i32, i32]
        L.orc_build_system.restype = i64
        L.orc_build_system.argtypes = [i32, i32, i32, P, P, P, P, dbl, P, P, P, P, P, P, P]
        L.orc_build_system_full.restype = i64
        L.orc_build_system_full.argtypes = [i32, i32, i32, P, P, P, P, dbl, P, P, P, P, P, P, P, P]
        L.orc_mt1d_anl_full.argtypes = [i32, P, P, dbl, P, P]
        L.orc_row_scale.argtypes = [i32, i32, i32, P, P, P, P, P]
        L.orc_zamux.argtypes = [i64, P, P, P, P, P]
        L.orc_ssor_apply.argtypes = [i64, P, P, P, dbl, P, P]
        L.orc_mass_apply.argtypes = [i32, i32, i32, P, P, P, P, P, P]
        L.orc_grad.argtypes = [i32, i32, i32, P, P, P, P, P]
        L.orc_grad_T.argtypes = [i32, i32, i32, P, P, P, P, P]
        L.orc_dc_matrix.restype = i64
        L.orc_dc_matrix.argtypes = [i32, i32, i32, P, P, P, P, P, P, P, P]
        L.orc_div_correct.restype = i32
        L.orc_div_correct.argtypes = [i32, i32, i32, P, P, P, P, P, dbl]
        ks = [i64, P, P, P, P, P, dbl, i32, i32, i32, i32, P, P, P, P, i32, dbl, dbl, dbl, P]
        L.orc_bicgstab.restype = i32
        L.orc_bicgstab.argtypes = ks
        L.orc_gpbicg.restype = i32
        L.orc_gpbicg.argtypes = ks
        L.orc_cocr.restype = i32
        L.orc_cocr.argtypes = [i64, P, P, P, P, P, P, dbl, i32, i32, i32, i32, P, P, P, P, i32, dbl, dbl, dbl, P]
        L.orc_mt1d_anl.argtypes = [i32, P, P, dbl, P, P, P]
        L.orc_quad_interp_e.restype = i32
        L.orc_quad_interp_e.argtypes = [i32, i32, i32, i32, P, P, P, P, P]
        L.orc_cubic_interp.restype = i32
        L.orc_cubic_interp.argtypes = [i32, i32, i32, i32, P, P, P, P, P]
        L.orc_extint_v.restype = i32
        L.orc_extint_v.argtypes = [i32, i32, i32, i32, P, P, P, P, P, P]
        L.orc_mesh_coarsen.restype = i32
        L.orc_mesh_coarsen.argtypes = [i32, i32, i32, P, P, P, P, i32, P, P, P, P]
        _lib = L
    return _lib


def _p(arr: np.ndarray):
    assert arr.flags["C_CONTIGUOUS"] or arr.ndim == 1
    return arr.ctypes.data_as(ctypes.c_void_p)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _c128(x):
    return np.ascontiguousarray(x, dtype=np.complex128)


def sigma_flat(sigma: np.ndarray) -> np.ndarray:
    """sigma(ncell, 6) -> column-major flat buffer sigma[comp*ncell + ie]."""
    sigma = np.asarray(sigma, dtype=np.float64)
    assert sigma.ndim == 2 and sigma.shape[1] == 6
    return np.ascontiguousarray(sigma.T).reshape(-1)


def nedges(nx, ny, nz):
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


def nnodes(nx, ny, nz):
    return (nx + 1) * (ny + 1) * (nz + 1)


# ---------------------------------------------------------------- 1-D field
def mt1d_anl(freq: float, conv: np.ndarray, c: np.ndarray):
    """mt1d_anl, src/model_selec.f90:266-401.  conv is (nz, 6)."""
    conv = np.asarray(conv, dtype=np.float64)
    nz = conv.shape[0]
    cf = np.ascontiguousarray(conv.T).reshape(-1)
    ex = np.zeros(2 * nz + 2, np.complex128)
    hx = np.zeros(2 * nz + 2, np.complex128)
    imp = np.zeros(4, np.complex128)
    c = _f64(c)
    lib().orc_mt1d_anl(nz, _p(cf), _p(c), float(freq), _p(ex), _p(hx), _p(imp))
    return ex, hx, imp


def background_column(sigma: np.ndarray, nx: int, ny: int, nz: int, full: bool = False) -> np.ndarray:
    """cont1(i,1:3) = sigma((i-1)*nx*ny+1, 1:3), cols 4:6 zero
    (src/fdfe_main.f90:938-942, 1376-1380); with full, all 6 components of
    that column (full-tensor boundary, a deliberate deviation)."""
    conv = np.zeros((nz, 6))
    cols = 6 if full else 3
    conv[:, :cols] = np.asarray(sigma)[np.arange(nz) * nx * ny, :cols]
    return conv


def mt1d_anl_full(freq: float, conv: np.ndarray, c: np.ndarray):
    """mt1d_anl with the cross-polar components: returns (ex, exc), ex as
    mt1d_anl, exc = Ey of the xy mode then Ex of the yx mode (2*nz+2)."""
    conv = np.asarray(conv, dtype=np.float64)
    nz = conv.shape[0]
    cf = np.ascontiguousarray(conv.T).reshape(-1)
    ex = np.zeros(2 * nz + 2, np.complex128)
    exc = np.zeros(2 * nz + 2, np.complex128)
    c = _f64(c)
    lib().orc_mt1d_anl_full(nz, _p(cf), _p(c), float(freq), _p(ex), _p(exc))
    return ex, exc


# ---------------------------------------------------------------- system
@dataclass
class System:
    nx: int
    ny: int
    nz: int
    A: sp.csr_matrix          # 0-based copy for SciPy use
    ia: np.ndarray            # 1-based, as dcfast builds it
    ja: np.ndarray
    val: np.ndarray
    p1: np.ndarray
    p2: np.ndarray
    bnd: np.ndarray
    e_dc: np.ndarray


def build_system(a, b, c, sigma, freq, e_dc=None, full=False, e_x=None) -> System:
    """fd_assemb + dcfast (src/fdfe_main.f90:226-546).  With full, the 1-D
    background uses the whole tensor of the first column and the cross-polar
    boundary components e_x are set as well (orc_build_system_full)."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    nx, ny, nz = len(a), len(b), len(c)
    if e_dc is None:
        if full:
            e_dc, e_x = mt1d_anl_full(freq, background_column(sigma, nx, ny, nz, True), c)
        else:
            e_dc, _, _ = mt1d_anl(freq, background_column(sigma, nx, ny, nz), c)
    e_dc = _c128(e_dc)
    nl = nedges(nx, ny, nz)
    ia = np.zeros(nl + 1, np.int64)
    ja = np.zeros(13 * nl, np.int32)
    val = np.zeros(13 * nl, np.complex128)
    p1 = np.zeros(nl, np.complex128)
    p2 = np.zeros(nl, np.complex128)
    bnd = np.zeros(nl, np.uint8)
    sf = sigma_flat(sigma)
    if full:
        e_x = _c128(e_x)
        nnz = lib().orc_build_system_full(nx, ny, nz, _p(a), _p(b), _p(c), _p(sf), float(freq), _p(e_dc),
                                          _p(e_x), _p(ia), _p(ja), _p(val), _p(p1), _p(p2), _p(bnd))
    else:
        nnz = lib().orc_build_system(nx, ny, nz, _p(a), _p(b), _p(c), _p(sf), float(freq), _p(e_dc),
                                     _p(ia), _p(ja), _p(val), _p(p1), _p(p2), _p(bnd))
    if nnz < 0:
        raise MemoryError("orc_build_system failed")
    ja = ja[:nnz].copy()
    val = val[:nnz].copy()
    A = sp.csr_matrix((val, ja - 1, ia - 1), shape=(nl, nl))
    return System(nx, ny, nz, A, ia, ja, val, p1, p2, bnd.astype(bool), e_dc)


def row_scale(a, b, c, bnd) -> np.ndarray:
    a, b, c = _f64(a), _f64(b), _f64(c)
    nl = nedges(len(a), len(b), len(c))
    ve = np.zeros(nl)
    bb = np.ascontiguousarray(bnd, dtype=np.uint8)
    lib().orc_row_scale(len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(bb), _p(ve))
    return ve


def zamux(S: System, x):
    x = _c128(x)
    y = np.zeros_like(x)
    lib().orc_zamux(len(x), _p(S.ia), _p(S.ja), _p(S.val), _p(x), _p(y))
    return y


def ssor_apply(S: System, y, omega=1.0):
    y = _c128(y)
    x = np.zeros_like(y)
    lib().orc_ssor_apply(len(y), _p(S.ia), _p(S.ja), _p(S.val), float(omega), _p(y), _p(x))
    return x


def mass_apply(a, b, c, sigma, x):
    a, b, c, x = _f64(a), _f64(b), _f64(c), _c128(x)
    y = np.zeros_like(x)
    lib().orc_mass_apply(len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(sigma_flat(sigma)), _p(x), _p(y))
    return y


def grad(a, b, c, phi):
    a, b, c, phi = _f64(a), _f64(b), _f64(c), _c128(phi)
    ge = np.zeros(nedges(len(a), len(b), len(c)), np.complex128)
    lib().orc_grad(len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(phi), _p(ge))
    return ge


def grad_T(a, b, c, w):
    a, b, c, w = _f64(a), _f64(b), _f64(c), _c128(w)
    t = np.zeros(nnodes(len(a), len(b), len(c)), np.complex128)
    lib().orc_grad_T(len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(w), _p(t))
    return t


def dc_matrix(a, b, c, sigma):
    a, b, c = _f64(a), _f64(b), _f64(c)
    nx, ny, nz = len(a), len(b), len(c)
    np_ = nnodes(nx, ny, nz)
    pia = np.zeros(np_ + 1, np.int64)
    pja = np.zeros(27 * np_, np.int32)
    pv = np.zeros(27 * np_)
    dcn = np.zeros(np_, np.uint8)
    nnz = lib().orc_dc_matrix(nx, ny, nz, _p(a), _p(b), _p(c), _p(sigma_flat(sigma)), _p(pia), _p(pja), _p(pv), _p(dcn))
    return sp.csr_matrix((pv[:nnz], pja[:nnz], pia), shape=(np_, np_)), dcn.astype(bool)


def div_correct(a, b, c, sigma, x, tol=1e-3):
    a, b, c = _f64(a), _f64(b), _f64(c)
    x = _c128(x).copy()
    it = lib().orc_div_correct(len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(sigma_flat(sigma)), _p(x), float(tol))
    return x, it


# ---------------------------------------------------------------- Krylov
@dataclass
class SolveStats:
    iters: int = 0
    relres: float = 0.0
    n_dc: int = 0
    pcg_iters: int = 0
    status: int = 0
    seconds: float = 0.0


NORM_REFERENCE, NORM_INTERIOR = "reference", "interior"


def krylov(S: System, rhs, x0, a, b, c, sigma, solver=SOLVER_BICGSTAB, eps=1e-6, maxit=5000,
           corr_intv=200, corr_tol=1e-3, omega=1.0, norm=NORM_REFERENCE):
    """Preconditioned Krylov solve on the reference CSR.  solver: 1 BiCGStab
    (src/solvers.f90:522-719), 3 GPBiCG (src/solvers.f90:137-315), 5 COCR on the
    dual-volume scaled symmetric system.

    norm="reference": stop on ||b-Ax||/||b|| over all rows, b including the
    Dirichlet identity rows (src/solvers.f90:615-625).  norm="interior": the
    Dirichlet values are imposed in x first and the normaliser is ||b|| over the
    interior rows only."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    x = _c128(x0).copy()
    rhs = _c128(rhs)
    bnorm = 0.0
    if norm == NORM_INTERIOR:
        x[S.bnd] = rhs[S.bnd]
        bnorm = float(np.linalg.norm(rhs[~S.bnd]))
    st = np.zeros(4)
    sf = sigma_flat(sigma)
    n = len(rhs)
    t0 = time.perf_counter()
    common_tail = (len(a), len(b), len(c), _p(a), _p(b), _p(c), _p(sf), int(corr_intv), float(corr_tol), float(omega), bnorm, _p(st))
    if solver == SOLVER_BICGSTAB:
        status = lib().orc_bicgstab(n, _p(S.ia), _p(S.ja), _p(S.val), _p(rhs), _p(x), float(eps), int(maxit), *common_tail)
    elif solver == SOLVER_GPBICG:
        status = lib().orc_gpbicg(n, _p(S.ia), _p(S.ja), _p(S.val), _p(rhs), _p(x), float(eps), int(maxit), *common_tail)
    elif solver == SOLVER_COCR:
        ve = row_scale(a, b, c, S.bnd)
        status = lib().orc_cocr(n, _p(S.ia), _p(S.ja), _p(S.val), _p(rhs), _p(x), _p(ve), float(eps), int(maxit), *common_tail)
    else:
        raise ValueError(f"unsupported solver {solver}")
    dt = time.perf_counter() - t0
    return x, SolveStats(int(st[0]), float(st[1]), int(st[2]), int(st[3]), int(status), dt)


def direct_solve(S: System):
    """Coarse-level direct solve of both polarisations (the reference uses
    PARDISO, src/solvers.f90:765-875, refactoring per polarisation; here one
    SuperLU factorisation serves both right-hand sides)."""
    lu = spla.splu(S.A.tocsc())
    return lu.solve(S.p1), lu.solve(S.p2)


# ---------------------------------------------------------------- EXCMG
def mesh_coarsen(a, b, c, sigma, k):
    """mesh_coarsen, src/fdfe_main.f90:18-84 (index bug Q6 fixed)."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    nx, ny, nz = len(a), len(b), len(c)
    asf = 2 ** k
    mx, my, mz = nx // asf, ny // asf, nz // asf
    a1, b1, c1 = np.zeros(mx), np.zeros(my), np.zeros(mz)
    s1 = np.zeros(6 * mx * my * mz)
    st = lib().orc_mesh_coarsen(nx, ny, nz, _p(a), _p(b), _p(c), _p(sigma_flat(sigma)), int(k), _p(a1), _p(b1), _p(c1), _p(s1))
    if st != 0:
        raise ValueError("error in mesh_coarsen, please check the input mesh")
    return a1, b1, c1, s1.reshape(6, -1).T.copy()


def extint_v(a, b, c, nair, u0, u1):
    """extint_v, src/excmg.f90:9-77 on level-i widths a,b,c."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    nx, ny, nz = len(a), len(b), len(c)
    out = np.zeros(nedges(nx, ny, nz), np.complex128)
    st = lib().orc_extint_v(nx, ny, nz, int(nair), _p(a), _p(b), _p(c), _p(_c128(u0)), _p(_c128(u1)), _p(out))
    if st != 0:
        raise ValueError("error in interpolation, please check the input mesh")
    return out


def quad_interp_e(a, b, c, nair, uc):
    a, b, c = _f64(a), _f64(b), _f64(c)
    out = np.zeros(nedges(len(a), len(b), len(c)), np.complex128)
    st = lib().orc_quad_interp_e(len(a), len(b), len(c), int(nair), _p(a), _p(b), _p(c), _p(_c128(uc)), _p(out))
    if st:
        raise ValueError("mesh not divisible by 4")
    return out


def cubic_interp(a, b, c, nair, uc):
    a, b, c = _f64(a), _f64(b), _f64(c)
    out = np.zeros(nedges(len(a), len(b), len(c)), np.complex128)
    st = lib().orc_cubic_interp(len(a), len(b), len(c), int(nair), _p(a), _p(b), _p(c), _p(_c128(uc)), _p(out))
    if st:
        raise ValueError("mesh not divisible by 8")
    return out


@dataclass
class Options:
    solver: int = SOLVER_BICGSTAB
    corr_intv: int = 200
    corr_tol: float = 1e-3
    omega: float = 1.0
    use_excmg: bool = True
    maxit_plain: int = 5000      # src/fdfe_main.f90:993
    maxit_excmg: int = 1000      # src/fdfe_main.f90:1414
    norm: str = "reference"      # or "interior"
    full_background: bool = False  # full-tensor 1-D boundary with cross-polar components


@dataclass
class LevelReport:
    level: int
    dims: tuple
    tol: float
    stats: list = field(default_factory=list)
    seconds: float = 0.0


def solve_level(a, b, c, sigma, freq, tol, opts: Options, x0=None, direct=False):
    """doublemode_nor (src/fdfe_main.f90:892-1063) for one grid: 1-D background,
    assembly, then both polarisations (direct for the coarse levels)."""
    S = build_system(a, b, c, sigma, freq, full=opts.full_background)
    nl = len(S.p1)
    if direct:
        u1, u2 = direct_solve(S)
        return u1, u2, [], S
    x1 = np.zeros(nl, np.complex128) if x0 is None else x0[0]
    x2 = np.zeros(nl, np.complex128) if x0 is None else x0[1]
    maxit = opts.maxit_plain if x0 is None else opts.maxit_excmg
    u1, s1 = krylov(S, S.p1, x1, a, b, c, sigma, opts.solver, tol, maxit, opts.corr_intv, opts.corr_tol, opts.omega, opts.norm)
    u2, s2 = krylov(S, S.p2, x2, a, b, c, sigma, opts.solver, tol, maxit, opts.corr_intv, opts.corr_tol, opts.omega, opts.norm)
    return u1, u2, [s1, s2], S


def excmg(a, b, c, nair, level, sigma, freq, eps, opts: Options):
    """u1u2 + doublemode_acc (src/fdfe_main.f90:1228-1563)."""
    reports = []
    tol = eps * 10.0 ** (-level + 1)                     # :1312
    levs = {}
    for ng in (level - 1, level - 2):                    # u1u2, :1243-1251
        a1, b1, c1, s1 = mesh_coarsen(a, b, c, sigma, ng)
        t0 = time.perf_counter()
        u1, u2, _, _ = solve_level(a1, b1, c1, s1, freq, tol, opts, direct=True)
        levs[ng] = (u1, u2)
        reports.append(LevelReport(level - ng, (len(a1), len(b1), len(c1)), 0.0, [], time.perf_counter() - t0))
    u0 = levs[level - 1]
    u1 = levs[level - 2]
    tol *= 10                                            # :1323
    out = None
    for i in range(3, level + 1):                        # :1329
        tol *= 10
        if i != level:
            a1, b1, c1, s1 = mesh_coarsen(a, b, c, sigma, level - i)
        else:
            a1, b1, c1, s1 = _f64(a), _f64(b), _f64(c), np.asarray(sigma, np.float64)
        nair_i = nair // 2 ** (level - i)
        x01 = extint_v(a1, b1, c1, nair_i, u0[0], u1[0])  # :1365-1366
        x02 = extint_v(a1, b1, c1, nair_i, u0[1], u1[1])
        t0 = time.perf_counter()
        w1, w2, st, _ = solve_level(a1, b1, c1, s1, freq, tol, opts, x0=(x01, x02))
        reports.append(LevelReport(i, (len(a1), len(b1), len(c1)), tol, st, time.perf_counter() - t0))
        u0, u1 = u1, (w1, w2)
        out = (w1, w2)
    return out[0], out[1], reports


def fwd_solver(a, b, c, nair, level, eps, sigma, freqs, opts: Options = Options()):
    """fwd_solver, src/fdfe_main.f90:1565-1594: ex(2*nl, nf) with xy then yx."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    nl = nedges(len(a), len(b), len(c))
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    ex = np.zeros((2 * nl, len(freqs)), np.complex128)
    reports = []
    for f_i, ff in enumerate(freqs):
        if opts.use_excmg:
            u1, u2, rep = excmg(a, b, c, nair, level, sigma, ff, eps, opts)
        else:
            u1, u2, st, _ = solve_level(a, b, c, sigma, ff, eps, opts)
            rep = [LevelReport(level, (len(a), len(b), len(c)), eps, st)]
        ex[:nl, f_i] = u1
        ex[nl:, f_i] = u2
        reports.append(rep)
    return ex, reports


# ---------------------------------------------------------------- post-process
def _lagrange3(x, x0):
    """coef of arho_interp2 (src/fdfe_main.f90:779-789)."""
    return np.array([(x - x0[1]) * (x - x0[2]) / (x0[0] - x0[1]) / (x0[0] - x0[2]),
                     (x - x0[0]) * (x - x0[2]) / (x0[1] - x0[0]) / (x0[1] - x0[2]),
                     (x - x0[0]) * (x - x0[1]) / (x0[2] - x0[0]) / (x0[2] - x0[1])])


def arho_interp2(mode, rd, a, b, fields, recv):
    """arho_interp2, src/fdfe_main.f90:718-790.  fields = (ex,ey,hx,hy,hz) each
    (nx, ny, 2).  rd=2 reproduces the reference's transposed sampling (Q7); rd=1
    samples at (irec, jrec)."""
    ex, ey, hx, hy, hz = fields
    nx, ny = len(a), len(b)
    sa, sb = a.sum(), b.sum()
    aax = np.empty(nx)
    bby = np.empty(ny)
    aax[0] = (a[0] - sa) / 2
    bby[0] = (b[0] - sb) / 2
    for i in range(1, nx):
        aax[i] = -sa / 2 + a[:i].sum() + a[i] / 2
    for i in range(1, ny):
        bby[i] = -sb / 2 + b[:i].sum() + b[i] / 2
    nrec = recv.shape[0]
    out = np.zeros((nrec, 5), np.complex128)
    md = mode - 1
    for r in range(nrec):
        irec = next(j for j in range(nx - 1) if aax[j] <= recv[r, 0] <= aax[j + 1])
        jrec = next(j for j in range(ny - 1) if bby[j] <= recv[r, 1] <= bby[j + 1])
        c1 = _lagrange3(recv[r, 0], aax[irec - 1:irec + 2])
        c2 = _lagrange3(recv[r, 1], bby[jrec - 1:jrec + 2])
        acc = np.zeros(5, np.complex128)
        for m in range(3):
            for n in range(3):
                w = c1[m] * c2[n]
                if rd == 1:
                    I_, J_ = irec - 1 + m, jrec - 1 + n
                    acc += w * np.array([ex[I_, J_, md], ey[I_, J_, md], hx[I_, J_, md], hy[I_, J_, md], hz[I_, J_, md]])
                else:
                    I_, J_ = jrec - 1 + m, irec - 1 + n
                    Ih, Jh = irec - 1 + m, jrec - 1 + n
                    acc += w * np.array([ex[I_, J_, md], ey[I_, J_, md], hx[I_, J_, md], hy[I_, J_, md], hz[Ih, Jh, md]])
        out[r] = acc
    return out


def surface_planes(a, b, c, sigma, air_max):
    """Surface node plane (0-based, from the top of the air) of every cell
    column, (nx, ny): the top of the column's first cell, from the top, whose
    sigma_xx exceeds air_max -- with cells above the ground carved to air, the
    first cell whose centre lies below the ground, as the topography branch of
    post_process selects it (src/fdfe_main.f90:604-614)."""
    nx, ny, nz = len(a), len(b), len(c)
    sxx = np.asarray(sigma, np.float64)[:, 0].reshape(nz, ny, nx)
    earth = sxx > air_max
    if not earth.any(axis=0).all():
        raise ValueError("a cell column is air throughout")
    return np.argmax(earth, axis=0).T.copy()


def topo_surface_planes(c, nair, topo):
    """Surface plane of each column from ground elevations topo(nx, ny) (m above
    the datum): one less than the first cell k1 (1-based) whose centre depth
    ccz(k1) >= -topo, the test of src/fdfe_main.f90:606-611."""
    c = np.asarray(c, np.float64)
    ccz = -c[:nair].sum() + np.cumsum(c) - c / 2
    topo = np.asarray(topo, np.float64)
    below = ccz[None, None, :] >= -topo[:, :, None]
    return np.argmax(below, axis=2)


def surface_fields(a, b, c, nair, freq, x1, x2, surf=None):
    """E and H at the surface cell centres (src/fdfe_main.f90:597-683).  surf
    (nx, ny) gives each column's surface node plane (the topography branch,
    :604-614, 656-666); None is the flat datum nair.  The z difference spans
    the cell below the surface plane: c(nair+1) on the flat datum, as the
    reference."""
    nx, ny, nz = len(a), len(b), len(c)
    n1 = nx * (ny + 1) * (nz + 1)
    n2 = ny * (nz + 1) * (nx + 1)
    m = np.full((nx, ny), nair, np.int64) if surf is None else np.asarray(surf, np.int64)
    assert m.shape == (nx, ny) and m.min() >= 0 and m.max() < nz
    I = np.arange(1, nx + 1)[:, None]
    J = np.arange(1, ny + 1)[None, :]
    idx = m * nx * (ny + 1) + (J - 1) * nx + I - 1           # 0-based versions
    idx2 = n1 + m * ny * (nx + 1) + (J - 1) * (nx + 1) + I - 1
    idx3 = n1 + n2 + m * (nx + 1) * (ny + 1) + (J - 1) * (nx + 1) + I - 1
    cz = c[m]
    kap = 1j * 2.0 * np.pi * freq * MU0
    A = a[:, None]
    B = b[None, :]
    ex = np.zeros((nx, ny, 2), np.complex128)
    ey = np.zeros_like(ex)
    hx = np.zeros_like(ex)
    hy = np.zeros_like(ex)
    hz = np.zeros_like(ex)
    for md, x in enumerate((x1, x2)):
        ex[:, :, md] = (x[idx] + x[idx + nx]) / 2
        ey[:, :, md] = (x[idx2] + x[idx2 + 1]) / 2
        exz = ((x[idx + nx * (ny + 1)] + x[idx + nx * (ny + 1) + nx]) / 2 - ex[:, :, md]) / cz
        ezx = ((x[idx3 + 1] + x[idx3 + nx + 2]) / 2 - (x[idx3 + nx + 1] + x[idx3]) / 2) / A
        eyz = ((x[idx2 + ny * (nx + 1)] + x[idx2 + ny * (nx + 1) + 1]) / 2 - ey[:, :, md]) / cz
        ezy = (-(x[idx3] + x[idx3 + 1]) / 2 + (x[idx3 + nx + 1] + x[idx3 + nx + 2]) / 2) / B
        exy = (x[idx + nx] - x[idx]) / B
        eyx = (x[idx2 + 1] - x[idx2]) / A
        hy[:, :, md] = (exz - ezx) / kap
        hx[:, :, md] = (ezy - eyz) / kap
        hz[:, :, md] = (exy - eyx) / kap
    return ex, ey, hx, hy, hz


def post_process(a, b, c, nair, freqs, recv, ex_all, out_type=1, rd=2, surf=None):
    """post_process, src/fdfe_main.f90:549-716.  Returns results (nrec*nf, 4):
    out_type 1 -> (Re Zxy, Im Zxy, Re Zyx, Im Zyx); 2 -> (rho_xy, rho_yx,
    phi_xy, phi_yx); and the tipper array (nrec*nf, 4).  surf: per-column
    surface node planes (topography), see surface_fields."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    recv = np.asarray(recv, np.float64)
    nl = nedges(len(a), len(b), len(c))
    freqs = np.atleast_1d(freqs)
    nrec = recv.shape[0]
    res = np.zeros((nrec * len(freqs), 4))
    tip = np.zeros((nrec * len(freqs), 4))
    for k, ff in enumerate(freqs):
        flds = surface_fields(a, b, c, nair, ff, ex_all[:nl, k], ex_all[nl:, k], surf)
        ar1 = arho_interp2(1, rd, a, b, flds, recv)
        ar2 = arho_interp2(2, rd, a, b, flds, recv)
        tmp = ar1[:, 2] * ar2[:, 3] - ar2[:, 2] * ar1[:, 3]
        zxy = (ar2[:, 0] * ar1[:, 2] - ar1[:, 0] * ar2[:, 2]) / tmp
        zyx = (ar1[:, 1] * ar2[:, 3] - ar2[:, 1] * ar1[:, 3]) / tmp
        om = 2 * np.pi * ff * MU0
        sl = slice(k * nrec, (k + 1) * nrec)
        if out_type == 2:
            res[sl, 0] = np.abs(zxy ** 2 * 1j) / om
            res[sl, 1] = np.abs(zyx ** 2 * 1j) / om
            res[sl, 2] = -np.arctan(zxy.imag / zxy.real) * 180 / np.pi
            res[sl, 3] = -np.arctan(zyx.imag / zyx.real) * 180 / np.pi
        else:
            res[sl] = np.stack([zxy.real, zxy.imag, zyx.real, zyx.imag], 1)
        tx = (ar2[:, 3] * ar1[:, 4] - ar1[:, 3] * ar2[:, 4]) / tmp
        ty = (ar1[:, 2] * ar2[:, 4] - ar2[:, 2] * ar1[:, 4]) / tmp
        tip[sl] = np.stack([tx.real, tx.imag, ty.real, ty.imag], 1)
    return res, tip


# ---------------------------------------------------------------- misfit and its sensitivity
def misfit(a, b, c, nair, freqs, recv, ex_all, obs, rd=2):
    """1/2 sum (log10 rho_a - log10 rho_obs)^2 + (phi - phi_obs)^2, phase in
    radians, rho_a and phi from post_process out_type 2 (src/fdfe_main.f90:
    686-697); obs rows (rho_xy, rho_yx, phi_xy, phi_yx) in ohm-m and degrees."""
    res, _ = post_process(a, b, c, nair, freqs, recv, ex_all, out_type=2, rd=rd)
    rr = np.log10(res[:, :2]) - np.log10(obs[:, :2])
    rp = np.deg2rad(res[:, 2:] - obs[:, 2:])
    return 0.5 * float(np.sum(rr ** 2) + np.sum(rp ** 2))


def misfit_sensitivity(a, b, c, nair, sigma, freqs, recv, obs, cells, rd=2, h=1e-4):
    """d misfit / d log(s_c), s_c a scale factor on cell c's conductivity
    tensor, by tangent-linear sensitivities on the CPU in FP64: the assembled
    operator (fd_assemb + dcfast) is affine in s_c, so A(2 sigma_c) - A(sigma_c)
    is its exact derivative; the field sensitivity de = -A^-1 dA e comes from
    SuperLU solves, and the misfit derivative along de from a central
    difference of the (smooth) post-process.  The 1-D Dirichlet data are held
    fixed, as in the device adjoint."""
    freqs = np.atleast_1d(freqs)
    nrec = len(recv)
    out = np.zeros(len(cells))
    for fi, f in enumerate(freqs):
        S = build_system(a, b, c, sigma, f)
        lu = spla.splu(S.A.tocsc())
        e1, e2 = lu.solve(S.p1), lu.solve(S.p2)
        ex = np.stack([np.concatenate([e1, e2])], axis=1)
        ob = obs[fi * nrec:(fi + 1) * nrec]
        for q, cell in enumerate(cells):
            sig2 = np.array(sigma, dtype=np.float64, copy=True)
            sig2[cell] *= 2.0
            S2 = build_system(a, b, c, sig2, f, e_dc=S.e_dc)
            d1 = -lu.solve(S2.A @ e1 - S2.p1)
            d2 = -lu.solve(S2.A @ e2 - S2.p2)
            de = np.stack([np.concatenate([d1, d2])], axis=1)
            fp = misfit(a, b, c, nair, [f], recv, ex + h * de, ob, rd)
            fm = misfit(a, b, c, nair, [f], recv, ex - h * de, ob, rd)
            out[q] += (fp - fm) / (2 * h)
    return out
