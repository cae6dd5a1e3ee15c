"""Pins for the FP64 CPU oracle (oracle/).  No GPU.

The reference ships no tests, golden vectors or fixtures for this path
(.MISSING_LARGE_BLOBS: dtm1.fwd), so each oracle function is pinned by what the
mathematics fixes: closed forms, invariants, exact polynomial reproduction,
brute-force dense algebra on tiny grids, and formulas of the reference retyped
independently of the oracle's own construction.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from sfd_excmg_amd import synth

MU0 = 4e-7 * np.pi


def tiny_model(nx=6, ny=5, nz=7, nair=2, seed=0, aniso=False):
    rng = np.random.default_rng(seed)
    a = rng.uniform(50, 150, nx)
    b = rng.uniform(50, 150, ny)
    c = rng.uniform(30, 120, nz)
    ncell = nx * ny * nz
    sig = np.zeros((ncell, 6))
    sig[:, :3] = 10 ** rng.uniform(-3, -1, (ncell, 1))
    if aniso:
        for ie in range(ncell):
            rho = 10 ** rng.uniform(0, 3, 3)
            S = synth.euler_tensor(*(1 / rho), *np.deg2rad(rng.uniform(0, 180, 3)))
            sig[ie] = synth.tensor_to_cols(S)
    sig[: nair * nx * ny, :3] = 1e-8
    sig[: nair * nx * ny, 3:] = 0
    return a, b, c, nair, sig


# ------------------------------------------------------------------ assembly
@pytest.mark.parametrize("aniso", [False, True])
def test_row_scaled_operator_is_complex_symmetric(aniso):
    a, b, c, nair, sig = tiny_model(aniso=aniso)
    S = O.build_system(a, b, c, sig, 0.3)
    ve = O.row_scale(a, b, c, S.bnd)
    As = sp.diags(ve) @ S.A
    inn = ~S.bnd
    Ai = As[inn][:, inn].toarray()
    assert np.abs(Ai - Ai.T).max() <= 1e-13 * np.abs(Ai).max()
    # and unscaled it is not symmetric (the reference stores it unscaled)
    A0 = S.A[inn][:, inn].toarray()
    assert np.abs(A0 - A0.T).max() > 1e-6 * np.abs(A0).max()


def test_curl_of_grad_vanishes_at_zero_conductivity():
    a, b, c, nair, sig = tiny_model()
    sig[:] = 0
    S = O.build_system(a, b, c, sig, 1.0, e_dc=np.zeros(2 * len(c) + 2))
    rng = np.random.default_rng(3)
    nx, ny, nz = len(a), len(b), len(c)
    phi = rng.standard_normal(O.nnodes(nx, ny, nz)) + 1j * rng.standard_normal(O.nnodes(nx, ny, nz))
    P, dcn = O.dc_matrix(a, b, c, np.ones_like(sig))
    phi[dcn] = 0
    ge = O.grad(a, b, c, phi)
    assert np.all(ge[S.bnd] == 0)
    y = O.zamux(S, ge)
    scale = np.abs(S.A).max() * np.abs(ge).max()
    assert np.abs(y[~S.bnd]).max() < 1e-12 * scale


def test_stencil_matches_reference_formulas_on_one_row():
    """Ex row coefficients retyped from src/fdfe_main.f90:252-290 (1-based)."""
    a, b, c, nair, sig = tiny_model(aniso=True)
    nx, ny, nz = len(a), len(b), len(c)
    f = 0.7
    S = O.build_system(a, b, c, sig, f)
    kap = 1j * 2 * np.pi * f * MU0
    i, j, k = 3, 2, 4                      # 1-based, interior
    ind = (k - 1) * nx * (ny + 1) + (j - 1) * nx + i
    ie = (k - 2) * nx * ny + (j - 2) * nx + i
    s = lambda e, comp: sig[e - 1, comp - 1]
    bj1, bj, ck1, ck = b[j - 2], b[j - 1], c[k - 2], c[k - 1]
    den = (bj1 + bj) * (ck1 + ck)
    sigmas = (s(ie, 1) * bj1 * ck1 + s(ie + nx, 1) * bj * ck1 + s(ie + nx * ny, 1) * bj1 * ck + s(ie + nx * ny + nx, 1) * bj * ck) / den
    sv1 = (s(ie, 4) * bj1 * ck1 + s(ie + nx * ny, 4) * bj1 * ck) / 2 / den
    gy = 2 / (a[i - 1] * (bj1 + bj))
    xnl = nx * (ny + 1) * (nz + 1)
    row = S.A.getrow(ind - 1)
    col = dict(zip(row.indices + 1, row.data))
    v1 = -2 / (ck1 * (ck1 + ck)); v2 = -2 / (bj1 * (bj1 + bj)); v4 = -2 / (bj * (bj1 + bj)); v5 = -2 / (ck * (ck1 + ck))
    assert np.isclose(col[ind], -(v1 + v2 + v4 + v5) - kap * sigmas, rtol=1e-14)
    ey = xnl + (k - 1) * (nx + 1) * ny + (j - 2) * (nx + 1) + i
    assert np.isclose(col[ey], gy - kap * sv1, rtol=1e-14)
    assert np.isclose(col[ey + 1], -gy - kap * sv1, rtol=1e-14)


def _layered_rotated_model(nx=6, ny=6, nair=2):
    """Horizontally layered earth whose layers carry tensors rotated about z
    (azimuthal anisotropy: xx, yy, xy nonzero, xz = yz = 0), under air."""
    ce = np.array([200.0, 300, 500, 800, 1500, 3000])
    c = synth.with_air(ce, nair)
    nz = len(c)
    layers = [(0.01, 0.002, 0.01, 30), (0.05, 0.01, 0.05, -50), (0.003, 0.03, 0.003, 70)]
    sig = np.zeros((nx * ny * nz, 6))
    for k in range(nz):
        if k < nair:
            cols = np.array([1e-8, 1e-8, 1e-8, 0, 0, 0])
        else:
            sx, sy, sz, strike = layers[min((k - nair) // 2, 2)]
            cols = synth.tensor_to_cols(synth.euler_tensor(sx, sy, sz, np.deg2rad(strike), 0.0, 0.0))
        sig[k * nx * ny:(k + 1) * nx * ny] = cols
    return np.full(nx, 1000.0), np.full(ny, 1000.0), c, sig


@pytest.mark.parametrize("pol", [0, 1])
def test_full_tensor_boundary_reproduces_1d_field_in_layered_anisotropic_model(pol):
    """Full-tensor 1-D boundary (a deliberate deviation from the reference's
    diagonal-only background, src/fdfe_main.f90:938-942, 470-494): in a layered
    model with rotated tensors both horizontal components of the 3-D solution
    follow the 1-D field everywhere inside the grid (to the discretisation
    error), while the reference boundary, which zeroes the cross-polar
    component on the faces, leaves an O(1) error in it."""
    a, b, c, sig = _layered_rotated_model()
    nx, ny, nz = len(a), len(b), len(c)
    f = 1.0
    conv = O.background_column(sig, nx, ny, nz, True)
    ex, exc = O.mt1d_anl_full(f, conv, c)
    assert abs(exc[0]) < 1e-12 and abs(exc[nz + 1]) < 1e-12  # top normalisation E = (1,0), (0,1)
    co = ex[: nz + 1] if pol == 0 else ex[nz + 1:]            # co-polar 1-D component at the z-nodes
    cr = exc[: nz + 1] if pol == 0 else exc[nz + 1:]          # cross-polar
    assert np.abs(cr).max() > 0.05 * np.abs(co).max()        # the rotation couples the modes
    xnl = nx * (ny + 1) * (nz + 1)
    ynl = (nx + 1) * ny * (nz + 1)

    def errors(full):
        S = O.build_system(a, b, c, sig, f, full=full)
        x = spla.spsolve(S.A.tocsc(), S.p1 if pol == 0 else S.p2)
        exs = x[:xnl].reshape(nz + 1, ny + 1, nx)[1:nz, 1:ny, :]          # interior x-edges
        eys = x[xnl:xnl + ynl].reshape(nz + 1, ny, nx + 1)[1:nz, :, 1:nx]  # interior y-edges
        ref_x, ref_y = (co, cr) if pol == 0 else (cr, co)
        scale = np.abs(co).max()
        return (np.abs(exs - ref_x[1:nz, None, None]).max() / scale,
                np.abs(eys - ref_y[1:nz, None, None]).max() / scale)

    ex_full, ey_full = errors(True)
    ex_ref, ey_ref = errors(False)
    cross_full, cross_ref = (ey_full, ey_ref) if pol == 0 else (ex_full, ex_ref)
    assert max(ex_full, ey_full) < 0.02
    assert cross_ref > 10 * cross_full


def test_layered_model_equals_1d_fd_solution_and_converges_to_analytic():
    """In a horizontally layered isotropic model the xy-mode 3-D solution is
    Ex(z) only and equals the 3-point 1-D FD solution on the same z grid; the
    1-D FD solution approaches mt1d_anl at O(h^2)."""
    nx, ny = 4, 4
    f = 1.0
    errs = []
    for refine in (1, 2):
        nair = 2 * refine
        ce = np.repeat(np.array([200.0, 300, 500, 800, 1500, 3000]), refine) / refine
        c = synth.with_air(ce, nair)
        nz = len(c)
        layer_sig = np.concatenate([np.full(nair, 1e-8), np.repeat([0.01, 0.01, 0.1, 0.1, 0.003, 0.003], refine)])
        sig = np.zeros((nx * ny * nz, 6))
        for k in range(nz):
            sig[k * nx * ny:(k + 1) * nx * ny, :3] = layer_sig[k]
        a = np.full(nx, 1000.0)
        b = np.full(ny, 1000.0)
        edc, _, _ = O.mt1d_anl(f, O.background_column(sig, nx, ny, nz), c)
        # 1-D FD with identical coefficients
        kap = 1j * 2 * np.pi * f * MU0
        n = nz + 1
        M = np.zeros((n, n), complex)
        rhs = np.zeros(n, complex)
        M[0, 0] = 1; rhs[0] = edc[0]; M[-1, -1] = 1; rhs[-1] = edc[nz]
        for k in range(1, nz):
            cm, cp = c[k - 1], c[k]
            M[k, k - 1] = -2 / (cm * (cm + cp)); M[k, k + 1] = -2 / (cp * (cm + cp))
            se = (layer_sig[k - 1] * cm + layer_sig[k] * cp) / (cm + cp)
            M[k, k] = -(M[k, k - 1] + M[k, k + 1]) - kap * se
        e1 = np.linalg.solve(M, rhs)
        # with the 1-D FD field as Dirichlet data the 3-D solve must reproduce it
        S = O.build_system(a, b, c, sig, f, e_dc=np.concatenate([e1, e1]))
        x = spla.spsolve(S.A.tocsc(), S.p1)
        xs = x[: nx * (ny + 1) * (nz + 1)].reshape(nz + 1, ny + 1, nx)
        for jj in range(0, ny + 1):
            assert np.allclose(xs[:, jj, :], e1[:, None], rtol=1e-7, atol=1e-8)  # cond(A) ~ 1e13 with air at 1e-8 S/m
        # Ey, Ez vanish in the xy mode
        assert np.abs(x[nx * (ny + 1) * (nz + 1):]).max() < 1e-7
        errs.append(np.abs(e1 - edc[: nz + 1]).max() if refine == 1 else np.abs(e1[::2] - edc[: nz + 1:2]).max())
    assert errs[1] < errs[0] / 3.0  # second order: ratio ~4


# ------------------------------------------------------------------ 1-D field
def test_mt1d_uniform_earth_is_plane_wave():
    nz = 10
    c = np.full(nz, 100.0)
    s = 0.02
    conv = np.zeros((nz, 6)); conv[:, :3] = s
    f = 3.0
    ex, hx, imp = O.mt1d_anl(f, conv, c)
    z = np.concatenate([[0], np.cumsum(c)])
    k = np.sqrt(-1j * 2 * np.pi * f * MU0 * s)
    k = k if k.real > 0 else -k
    assert np.allclose(ex[: nz + 1], np.exp(-k * z), rtol=1e-10)
    assert np.allclose(ex[nz + 1:], np.exp(-k * z), rtol=1e-10)
    # half-space impedance magnitude: rho_a = |Z|^2/(w mu) = 1/sigma
    Z = 2 * np.pi * f * MU0 / k  # E/H for a decaying plane wave, |.| fixed
    assert np.isclose(abs(Z) ** 2 / (2 * np.pi * f * MU0), 1 / s, rtol=1e-12)


def test_mt1d_axial_anisotropy_decouples_modes():
    nz = 8
    c = np.full(nz, 250.0)
    conv = np.zeros((nz, 6)); conv[:, 0] = 0.1; conv[:, 1] = 0.001; conv[:, 2] = 0.01
    ex, _, _ = O.mt1d_anl(0.5, conv, c)
    z = np.concatenate([[0], np.cumsum(c)])
    for s, part in ((0.1, ex[: nz + 1]), (0.001, ex[nz + 1:])):
        k = np.sqrt(-1j * 2 * np.pi * 0.5 * MU0 * s)
        k = k if k.real > 0 else -k
        assert np.allclose(part, np.exp(-k * z), rtol=1e-9)


def test_mt1d_two_layers_matches_fine_fd():
    """Layer over a half-space: mt1d_anl against a very fine 1-D FD solve with a
    radiation (impedance) condition at the bottom."""
    f = 0.2
    c = np.array([300.0, 300, 300, 5000])
    sig = np.array([0.05, 0.05, 0.05, 0.002])
    conv = np.zeros((4, 6)); conv[:, :3] = sig[:, None]
    ex, _, _ = O.mt1d_anl(f, conv, c)
    kap = 1j * 2 * np.pi * f * MU0
    h = 1.0
    zs = np.arange(0, c.sum() + h / 2, h)
    n = len(zs)
    sz = np.where(zs[:-1] + h / 2 < 900, 0.05, 0.002)
    main = np.zeros(n, complex); lo = np.zeros(n - 1, complex); up = np.zeros(n - 1, complex)
    for q in range(1, n - 1):
        lo[q - 1] = -1 / h ** 2; up[q] = -1 / h ** 2
        main[q] = 2 / h ** 2 - kap * (sz[q - 1] + sz[q]) / 2
    main[0] = 1
    kb = np.sqrt(-kap * 0.002); kb = kb if kb.real > 0 else -kb
    # bottom: dE/dz = -k E  (outgoing)
    lo[-1] = -1 / h; main[-1] = 1 / h + kb
    M = sp.diags([lo, main, up], [-1, 0, 1]).tocsc()
    rhs = np.zeros(n, complex); rhs[0] = 1
    e = spla.spsolve(M, rhs)
    zn = np.concatenate([[0], np.cumsum(c)])
    idx = np.round(zn / h).astype(int)
    assert np.allclose(ex[:5], e[idx], rtol=2e-4, atol=1e-6)


# ------------------------------------------------------------------ SSOR / solvers
def test_ssor_equals_dense_factorisation():
    a, b, c, nair, sig = tiny_model(4, 3, 4, 1)
    S = O.build_system(a, b, c, sig, 1.0)
    A = S.A.toarray()
    D = np.diag(np.diag(A)); L = np.tril(A, -1); U = np.triu(A, 1)
    for w in (1.0, 1.3):
        M = (D / w + L) @ np.linalg.inv(D) @ (D / w + U) * (w / (2 - w))
        y = np.random.default_rng(1).standard_normal(len(A)) + 0j
        x = O.ssor_apply(S, y, w)
        assert np.allclose(M @ x, y, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("solver", [O.SOLVER_BICGSTAB, O.SOLVER_GPBICG, O.SOLVER_COCR])
@pytest.mark.parametrize("corr", [0, 7])
def test_krylov_solvers_reach_dense_solution(solver, corr):
    a, b, c, nair, sig = tiny_model(6, 6, 6, 2, aniso=True)
    sig[: nair * 36, :3] = 1e-2      # mild air contrast keeps cond(A) ~ 1e7
    S = O.build_system(a, b, c, sig, 2.0)
    xd = np.linalg.solve(S.A.toarray(), S.p1)
    # BiCGStab stagnates near 1e-6 on this system (SciPy's bicgstab does too)
    eps = 1e-5 if solver == O.SOLVER_BICGSTAB else 1e-12
    x, st = O.krylov(S, S.p1, np.zeros_like(S.p1), a, b, c, sig, solver, eps, 3000, corr, 1e-6,
                     norm="interior")
    assert st.status == 0
    r = S.p1 - S.A @ x
    # recursive and true residual drift apart by a small factor after DC restarts
    assert np.linalg.norm(r[~S.bnd]) <= 10 * eps * np.linalg.norm(S.p1[~S.bnd])
    tol = 1e-2 if solver == O.SOLVER_BICGSTAB else (1e-8 if corr == 0 else 1e-5)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < tol


@pytest.mark.parametrize("k", [1, 2, 3, 7])
def test_bicgstab_iterates_match_scipy(k):
    """Right-preconditioned BiCGStab (src/solvers.f90:621-684) is van der
    Vorst's algorithm; SciPy's bicgstab with the same SSOR operator produces the
    same iterates."""
    a, b, c, nair, sig = tiny_model(5, 4, 5, 1, aniso=True)
    sig[:20, :3] = 1e-2              # keep rounding growth small (cond ~ 1e7)
    S = O.build_system(a, b, c, sig, 1.0)
    x0 = np.zeros_like(S.p1)
    M = spla.LinearOperator(S.A.shape, matvec=lambda v: O.ssor_apply(S, v), dtype=complex)
    xs, _ = spla.bicgstab(S.A, S.p1, x0=x0, M=M, rtol=1e-30, atol=0, maxiter=k)
    x, st = O.krylov(S, S.p1, x0, a, b, c, sig, O.SOLVER_BICGSTAB, 1e-30, k, 0, 1e-3)
    assert st.iters == k
    # iterates agree to rounding; BiCGStab amplifies it quickly here, so compare
    # the iterates for the first steps and the residuals afterwards
    if k <= 3:
        assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    rn = lambda v: np.linalg.norm(S.p1 - S.A @ v)
    assert abs(rn(x) - rn(xs)) <= 1e-3 * rn(xs)


@pytest.mark.parametrize("dims", [(3, 2, 2), (3, 3, 2)])
def test_gpbicg_first_step_is_bicgstab_and_terminates(dims):
    """GPBiCG (src/solvers.f90:137-315) pinned at iterate level by two
    properties of the method (Zhang, SIAM J. Sci. Comput. 18 (1997) 537-551):
    its first step has eta = 0 and zeta = (At, t)/(At, At), i.e. it is the first
    BiCGStab step, checked against SciPy's bicgstab; and its residual is a
    product of the BiCG polynomial, so in exact arithmetic it vanishes within
    as many steps as there are interior unknowns (the boundary rows are the
    identity).  Dropping the beta*u term, flipping the sign of eta, transposing
    an inner product in zeta, or dropping a term of y, z or w leaves a relative
    interior residual of 1e-4 or more after those steps."""
    a, b, c, nair, sig = tiny_model(*dims, 1, aniso=True)
    S = O.build_system(a, b, c, sig, 1.0)
    inner = ~S.bnd
    x0 = np.zeros_like(S.p1)
    M = spla.LinearOperator(S.A.shape, matvec=lambda v: O.ssor_apply(S, v), dtype=complex)
    xs, _ = spla.bicgstab(S.A, S.p1, x0=x0, M=M, rtol=1e-30, atol=0, maxiter=1)
    x1, st = O.krylov(S, S.p1, x0, a, b, c, sig, O.SOLVER_GPBICG, 1e-300, 1, 0, 1e-3)
    assert st.iters == 1
    assert np.linalg.norm(x1 - xs) <= 1e-12 * np.linalg.norm(xs)
    n = int(inner.sum())
    x, st = O.krylov(S, S.p1, x0, a, b, c, sig, O.SOLVER_GPBICG, 1e-300, n + 1, 0, 1e-3)
    red = np.linalg.norm((S.p1 - S.A @ x)[inner]) / np.linalg.norm(S.p1[inner])
    assert red <= 1e-8, red


# ------------------------------------------------------------------ divergence correction
def reference_iso_PT(a, b, c, sig):
    """Isotropic P (7-point) and T (6 per node), retyped from
    src/div_correc.f90:593-615 and 709-750 (1-based)."""
    nx, ny, nz = len(a), len(b), len(c)
    n1 = nx * ny
    xnl = nx * (ny + 1) * (nz + 1); ynl = (nx + 1) * ny * (nz + 1)
    npn = (nx + 1) * (ny + 1) * (nz + 1)
    nl = xnl + ynl + (nx + 1) * (ny + 1) * nz
    S = lambda e, comp: sig[e - 1, comp - 1]
    A = lambda q: a[q - 1]; B = lambda q: b[q - 1]; C = lambda q: c[q - 1]
    Pr, Pc, Pv, Tr, Tc, Tv = [], [], [], [], [], []
    for k in range(2, nz + 1):
        for j in range(2, ny + 1):
            for i in range(2, nx + 1):
                ind = (k - 1) * (nx + 1) * (ny + 1) + (j - 1) * (nx + 1) + i
                ie = (k - 1) * nx * ny + (j - 1) * nx + i
                s = [(S(ie - n1 - nx - 1, 3) * A(i - 1) * B(j - 1) + S(ie - n1 - nx, 3) * A(i) * B(j - 1) + S(ie - n1 - 1, 3) * A(i - 1) * B(j) + S(ie - n1, 3) * A(i) * B(j)) / C(k - 1) * 2,
                     (S(ie - n1 - nx - 1, 2) * A(i - 1) * C(k - 1) + S(ie - n1 - nx, 2) * A(i) * C(k - 1) + S(ie - nx - 1, 2) * A(i - 1) * C(k) + S(ie - nx, 2) * A(i) * C(k)) / B(j - 1) * 2,
                     (S(ie - n1 - nx - 1, 1) * B(j - 1) * C(k - 1) + S(ie - n1 - 1, 1) * B(j) * C(k - 1) + S(ie - nx - 1, 1) * B(j - 1) * C(k) + S(ie - 1, 1) * B(j) * C(k)) / A(i - 1) * 2,
                     (S(ie - n1 - nx, 1) * B(j - 1) * C(k - 1) + S(ie - n1, 1) * B(j) * C(k - 1) + S(ie - nx, 1) * B(j - 1) * C(k) + S(ie, 1) * B(j) * C(k)) / A(i) * 2,
                     (S(ie - n1 - 1, 2) * A(i - 1) * C(k - 1) + S(ie - n1, 2) * A(i) * C(k - 1) + S(ie - 1, 2) * A(i - 1) * C(k) + S(ie, 2) * A(i) * C(k)) / B(j) * 2,
                     (S(ie - nx - 1, 3) * A(i - 1) * B(j - 1) + S(ie - nx, 3) * A(i) * B(j - 1) + S(ie - 1, 3) * A(i - 1) * B(j) + S(ie, 3) * A(i) * B(j)) / C(k) * 2]
                cols = [ind - (nx + 1) * (ny + 1), ind - nx - 1, ind - 1, ind, ind + 1, ind + nx + 1, ind + (nx + 1) * (ny + 1)]
                vals = [-s[0], -s[1], -s[2], sum(s), -s[3], -s[4], -s[5]]
                Pr += [ind - 1] * 7; Pc += [q - 1 for q in cols]; Pv += vals
                t = [-x * C(k - 1) * C(k) / 1 for x in ()]
                ts = [-s[0] * C(k - 1), -s[1] * B(j - 1), -s[2] * A(i - 1), s[3] * A(i), s[4] * B(j), s[5] * C(k)]
                idx = (2, 3, 1, 4, 0, 5)
                tv = [ts[q] for q in idx]
                tc = [(k - 1) * nx * (ny + 1) + (j - 1) * nx + i - 1, (k - 1) * nx * (ny + 1) + (j - 1) * nx + i,
                      xnl + (k - 1) * (nx + 1) * ny + (j - 2) * (nx + 1) + i, xnl + (k - 1) * (nx + 1) * ny + (j - 1) * (nx + 1) + i,
                      xnl + ynl + (k - 2) * (nx + 1) * (ny + 1) + (j - 1) * (nx + 1) + i, xnl + ynl + (k - 1) * (nx + 1) * (ny + 1) + (j - 1) * (nx + 1) + i]
                Tr += [ind - 1] * 6; Tc += [q - 1 for q in tc]; Tv += tv
    P = sp.csr_matrix((Pv, (Pr, Pc)), shape=(npn, npn))
    T = sp.csr_matrix((Tv, (Tr, Tc)), shape=(npn, nl))
    return P, T


def test_isotropic_dc_operators_equal_reference_up_to_factor_8():
    a, b, c, nair, sig = tiny_model(5, 4, 6, 2)
    nx, ny, nz = len(a), len(b), len(c)
    Pref, Tref = reference_iso_PT(a, b, c, sig)
    P, dcn = O.dc_matrix(a, b, c, sig)
    inner = ~dcn
    assert np.allclose((8 * P)[inner].toarray(), Pref[inner].toarray(), rtol=1e-12, atol=1e-12 * abs(Pref).max())
    x = np.random.default_rng(2).standard_normal(O.nedges(nx, ny, nz)) * (1 + 0.3j)
    tmine = O.grad_T(a, b, c, O.mass_apply(a, b, c, sig, x))
    tref = Tref @ x
    assert np.allclose(-8 * tmine[inner], tref[inner], rtol=1e-11, atol=1e-11 * abs(tref).max())


@pytest.mark.parametrize("aniso", [False, True])
def test_dc_removes_discrete_divergence(aniso):
    a, b, c, nair, sig = tiny_model(6, 5, 6, 2, aniso=aniso)
    nx, ny, nz = len(a), len(b), len(c)
    P, dcn = O.dc_matrix(a, b, c, sig)
    Pi = P[~dcn][:, ~dcn].toarray()
    assert np.abs(Pi - Pi.T).max() < 1e-12 * np.abs(Pi).max()
    assert np.linalg.eigvalsh(Pi).min() > 0
    rng = np.random.default_rng(5)
    x = rng.standard_normal(O.nedges(nx, ny, nz)) + 1j * rng.standard_normal(O.nedges(nx, ny, nz))
    div = lambda v: O.grad_T(a, b, c, O.mass_apply(a, b, c, sig, v))[~dcn]
    d0 = np.linalg.norm(div(x))
    x2, it = O.div_correct(a, b, c, sig, x, tol=1e-10)
    assert np.linalg.norm(div(x2)) < 1e-8 * d0
    S = O.build_system(a, b, c, sig, 1.0)
    assert np.array_equal(x2[S.bnd], x[S.bnd])


# ------------------------------------------------------------------ EXCMG transfer
def _ref_point_coords(w, nb_fine, m):
    """Along-edge sample positions used by cal_coef (src/excmg.f90:683-685):
    coarse edge q of a block sits at the midpoint of fine edge 2q."""
    e = np.concatenate([[0], np.cumsum(w)])
    fine_mid = (e[:-1] + e[1:]) / 2
    return e, fine_mid


@pytest.mark.parametrize("m", [2, 4])
def test_block_interp_reproduces_polynomials(m):
    rng = np.random.default_rng(7)
    B = 2 * m
    nx, ny, nz = 2 * B, B, 2 * B
    a = rng.uniform(50, 150, nx); b = rng.uniform(50, 150, ny); c = rng.uniform(50, 150, nz)
    nair = 2
    # node coordinates as cal_coef uses them
    ex_, ey_, ez_ = np.concatenate([[0], np.cumsum(a)]) - a.sum() / 2, np.concatenate([[0], np.cumsum(b)]) - b.sum() / 2, np.concatenate([[0], np.cumsum(c)]) - c[:nair].sum()
    co = [rng.standard_normal() for _ in range(40)]
    degT, degL = m, m - 1

    def f(x, y, z, d):
        # transverse polynomial of degree m in each transverse coordinate, along-edge degree m-1
        t = [y, z] if d == 0 else ([x, z] if d == 1 else [x, y])
        l = [x, y, z][d]
        s = 0
        q = 0
        for p1 in range(degT + 1):
            for p2 in range(degT + 1):
                for pl in range(degL + 1):
                    s += co[(q * 7 + d) % 40] * (t[0] / 500) ** p1 * (t[1] / 500) ** p2 * (l / 500) ** pl
                    q += 1
        return s

    def field(nodes, mids, coarse):
        X, Y, Z = nodes
        MX, MY, MZ = mids
        out = []
        nxl, nyl, nzl = len(X) - 1, len(Y) - 1, len(Z) - 1
        for k in range(nzl + 1):
            for j in range(nyl + 1):
                for i in range(nxl):
                    out.append(f(MX[i], Y[j], Z[k], 0))
        for k in range(nzl + 1):
            for j in range(nyl):
                for i in range(nxl + 1):
                    out.append(f(X[i], MY[j], Z[k], 1))
        for k in range(nzl):
            for j in range(nyl + 1):
                for i in range(nxl + 1):
                    out.append(f(X[i], Y[j], MZ[k], 2))
        return np.array(out, complex)

    fine_nodes = (ex_, ey_, ez_)
    fine_mids = tuple((v[:-1] + v[1:]) / 2 for v in fine_nodes)
    coarse_nodes = tuple(v[::2] for v in fine_nodes)
    # coarse along-edge sample points: midpoint of the first fine sub-edge
    coarse_pts = tuple((v[:-1:2] + v[1::2]) / 2 for v in fine_nodes)
    uc = field(coarse_nodes, coarse_pts, True)
    uf_exact = field(fine_nodes, fine_mids, False)
    interp = O.quad_interp_e if m == 2 else O.cubic_interp
    uf = interp(a, b, c, nair, uc)
    assert np.allclose(uf, uf_exact, rtol=1e-9, atol=1e-9 * np.abs(uf_exact).max())


def test_extint_v_extrapolation_constants():
    """w = 1.25 u1 - 0.25 I2 u0 (src/excmg.f90:40-41,60-63): with u0 = 0 the
    result is I4(1.25 u1); with u1 = I2 u0 it is I4(I2 u0)."""
    rng = np.random.default_rng(1)
    nx = ny = nz = 16
    a, b, c = rng.uniform(50, 100, nx), rng.uniform(50, 100, ny), rng.uniform(50, 100, nz)
    nair = 4
    a1, b1, c1 = a.reshape(-1, 2).sum(1), b.reshape(-1, 2).sum(1), c.reshape(-1, 2).sum(1)
    n1 = O.nedges(8, 8, 8); n0 = O.nedges(4, 4, 4)
    u1 = rng.standard_normal(n1) + 1j * rng.standard_normal(n1)
    u0 = rng.standard_normal(n0) + 1j * rng.standard_normal(n0)
    out = O.extint_v(a, b, c, nair, np.zeros(n0), u1)
    assert np.allclose(out, O.cubic_interp(a, b, c, nair, 1.25 * u1), rtol=1e-13)
    i2 = O.quad_interp_e(a1, b1, c1, nair // 2, u0)
    out2 = O.extint_v(a, b, c, nair, u0, i2)
    assert np.allclose(out2, O.cubic_interp(a, b, c, nair, i2), rtol=1e-12, atol=1e-13)


def test_mesh_coarsen_preserves_integral():
    a, b, c, nair, sig = tiny_model(8, 8, 8, 2, aniso=True)
    a1, b1, c1, s1 = O.mesh_coarsen(a, b, c, sig, 1)
    V = (c[:, None, None] * b[None, :, None] * a[None, None, :]).reshape(-1)
    V1 = (c1[:, None, None] * b1[None, :, None] * a1[None, None, :]).reshape(-1)
    assert np.allclose((sig * V[:, None]).sum(0), (s1 * V1[:, None]).sum(0), rtol=1e-13)
    sig2 = np.tile([0.3, 0.2, 0.1, 0.01, 0.02, 0.03], (len(V), 1))
    _, _, _, s2 = O.mesh_coarsen(a, b, c, sig2, 2)
    assert np.allclose(s2, sig2[0], rtol=1e-14)


# ------------------------------------------------------------------ post-process
def test_arho_interp_exact_on_quadratics_rd1():
    rng = np.random.default_rng(3)
    nx = ny = 10
    a = rng.uniform(80, 120, nx); b = rng.uniform(80, 120, ny)
    xc, yc = synth.cell_centres(a), synth.cell_centres(b)
    F = lambda x, y: 1 + 0.002 * x - 0.003 * y + 1e-6 * x * x + 2e-6 * x * y - 1e-6 * y * y
    grid = F(xc[:, None], yc[None, :])
    fields = tuple(np.stack([grid * (q + 1), grid * (q + 2)], -1).astype(complex) for q in range(5))
    recv = np.array([[rng.uniform(xc[2], xc[-3]), rng.uniform(yc[2], yc[-3]), 0] for _ in range(8)])
    out = O.arho_interp2(1, 1, a, b, fields, recv)
    exp = F(recv[:, 0], recv[:, 1])
    for q in range(5):
        assert np.allclose(out[:, q], exp * (q + 1), rtol=1e-12)


def test_arho_interp_rd2_samples_the_transposed_window():
    """rd=2 (src/fdfe_main.f90:763-768, Q7) reads E and H at cell (jrec+m,
    irec+n) with the x-weights of irec and the y-weights of jrec; hz is read at
    (irec+m, jrec+n).  On cells of one width h the Lagrange weights reproduce
    quadratics, and the window (jrec.., irec..) is the station's own window
    shifted by d = (jrec - irec) h in x and -d in y, so for a biquadratic field
    E, H come out at (x + d, y - d) and hz at (x, y).  Transposing the weights
    instead of the cells, or hz as well, misses these points."""
    rng = np.random.default_rng(4)
    nx, ny, h = 14, 11, 100.0
    a, b = np.full(nx, h), np.full(ny, h)
    xc, yc = synth.cell_centres(a), synth.cell_centres(b)
    F = lambda x, y: 1 + 0.002 * x - 0.003 * y + 1e-6 * x * x + 2e-6 * x * y - 1e-6 * y * y + 1e-9 * x * x * y * y
    grid = F(xc[:, None], yc[None, :])
    fields = tuple(np.stack([grid * (q + 1), grid * (q + 2)], -1).astype(complex) for q in range(5))
    recv, shift = [], []
    while len(recv) < 12:
        x, y = rng.uniform(xc[1], xc[-2]), rng.uniform(yc[1], yc[-2])
        irec = int(np.searchsorted(xc, x)) - 1     # xc[irec] <= x < xc[irec + 1]
        jrec = int(np.searchsorted(yc, y)) - 1
        # the transposed window must lie inside the grid as well
        if 1 <= jrec <= nx - 2 and 1 <= irec <= ny - 2:
            recv.append([x, y, 0]); shift.append((jrec - irec) * h)
    recv, d = np.array(recv), np.array(shift)
    assert np.any(d != 0)
    out = O.arho_interp2(1, 2, a, b, fields, recv)
    moved = F(recv[:, 0] + d, recv[:, 1] - d)
    for q in range(4):
        assert np.allclose(out[:, q], moved * (q + 1), rtol=1e-11)
    assert np.allclose(out[:, 4], F(recv[:, 0], recv[:, 1]) * 5, rtol=1e-11)


@pytest.mark.slow
def test_c1_fixture_is_the_oracle_solution():
    """tests/golden/c1_oracle_impedances.npz (written by
    tests/golden/make_c1_fixture.py) is what the oracle computes now: a change
    of the oracle or of the model generator shows up here (about 10 s)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_oracle_impedances.npz"))
    m = synth.make_block_halfspace_32()
    assert float(np.sum(m.sigma)) == float(g["sigma_sum"])
    ex, st = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs,
                          O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2)
    assert np.allclose(res[:, 0] + 1j * res[:, 1], g["zxy"], rtol=1e-10, atol=0)
    assert np.allclose(res[:, 2] + 1j * res[:, 3], g["zyx"], rtol=1e-10, atol=0)


@pytest.mark.slow
def test_halfspace_apparent_resistivity_and_phase():
    m = synth.make_block_halfspace_32()
    m.sigma[m.nair * m.nx * m.ny:, :3] = 0.01
    opts = O.Options(norm="interior")
    ex, _ = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-10, m.sigma, m.freqs, opts)
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=2)
    assert np.allclose(res[:, :2], 100.0, rtol=0.02)
    assert np.allclose(res[:, 2:], 45.0, atol=0.5)


def _rotated_model_pair(seed=5):
    """A symmetric grid and a full-tensor block model, plus the same model
    rotated by +90 degrees about z: sigma' = R sigma R^T at the rotated cells."""
    rng = np.random.default_rng(seed)
    a = synth.padded_axis(4, 100.0, 2, 1.3)
    c = synth.with_air(synth.earth_layers(3, 60.0, 2, 1.3), 2)
    n, nz = len(a), len(c)
    sig = np.zeros((n * n * nz, 6))
    sig[:, :3] = 0.01                       # horizontally isotropic background
    for k in range(3, 6):
        for j in range(2, 6):
            for i in range(3, 7):
                S = synth.euler_tensor(*(10 ** rng.uniform(-2.5, -0.5, 3)), *rng.uniform(0, np.pi, 3))
                sig[(k * n + j) * n + i] = synth.tensor_to_cols(S)
    synth.set_air(sig, n, n, 2, 1e-6)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    rot = np.zeros_like(sig)
    for k in range(nz):
        for jp in range(n):
            for ip in range(n):
                s = sig[(k * n + (n - 1 - ip)) * n + jp]     # original cell (i=j', j=n-1-i')
                S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
                rot[(k * n + jp) * n + ip] = synth.tensor_to_cols(R @ S @ R.T)
    return a, c, sig, rot


def test_full_tensor_operator_is_rotation_equivariant():
    """Pins the anisotropic off-diagonal conductivity averaging of fd_assemb
    (src/fdfe_main.f90:257-260,305-308,354-357) without retyping it: on a grid
    symmetric under a 90-degree rotation about z, the discrete solution of the
    rotated full-tensor model must be the rotated solution of the original
    (E'_pol1 = -R E_pol2, E'_pol2 = R E_pol1), which fails for a wrong cell
    index, a dropped coupling or a sign error in any xy/xz/yz term."""
    a, c, sig, rot = _rotated_model_pair()
    n, nz = len(a), len(c)
    f = 1.0
    S0 = O.build_system(a, a, c, sig, f)
    S1 = O.build_system(a, a, c, rot, f)
    x1, x2 = O.direct_solve(S0)
    y1, y2 = O.direct_solve(S1)
    xn = n * (n + 1) * (nz + 1)
    ex = lambda v: v[:xn].reshape(nz + 1, n + 1, n)          # [k, j, i]
    ey = lambda v: v[xn:2 * xn].reshape(nz + 1, n, n + 1)    # [k, j, i]
    ez = lambda v: v[2 * xn:].reshape(nz, n + 1, n + 1)
    flipT = lambda v: np.transpose(v[:, ::-1, :], (0, 2, 1))
    # E'(x') = R E(R^T x'):  E'_x[k,j',i'] = -E_y[k, n-1-i', j'],  E'_y[k,j',i'] = E_x[k, n-i', j'],
    # E'_z[k,j',i'] = E_z[k, n-i', j'];  E'_pol1 = -R E_pol2 and E'_pol2 = +R E_pol1
    for yp, xs, sgn in ((y1, x2, -1.0), (y2, x1, 1.0)):
        scale = np.abs(yp).max()
        assert np.abs(ex(yp) - sgn * -flipT(ey(xs))).max() / scale < 1e-6
        assert np.abs(ey(yp) - sgn * flipT(ex(xs))).max() / scale < 1e-6
        assert np.abs(ez(yp) - sgn * flipT(ez(xs))).max() / scale < 1e-6


def test_misfit_sensitivity_matches_brute_force_difference():
    """The oracle's tangent-linear misfit derivative (exact operator derivative,
    SuperLU solves) equals a central difference in log s_c of the full
    solve-and-misfit chain."""
    m = synth.make_iso_block_model("sens", 16, 16, 12, 4, seed=41, nblocks=3, freq=1.0, level=3, nstations=4)
    freqs = np.array([1.0])
    cols = []
    for f in freqs:
        x1, x2 = O.direct_solve(O.build_system(m.a, m.b, m.c, m.sigma, f))
        cols.append(np.concatenate([x1, x2]))
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, np.stack(cols, 1), out_type=2)
    obs = res * np.array([1.1, 0.9, 1.0, 1.0]) + np.array([0, 0, 2.0, -1.0])
    nx, ny = m.nx, m.ny
    cells = [(m.nair + 1) * nx * ny + 8 * nx + 8, (m.nair + 3) * nx * ny + 6 * nx + 9]
    sens = O.misfit_sensitivity(m.a, m.b, m.c, m.nair, m.sigma, freqs, m.recv, obs, cells)

    def phi_of(sig):
        S = O.build_system(m.a, m.b, m.c, sig, freqs[0], e_dc=O.build_system(m.a, m.b, m.c, m.sigma, freqs[0]).e_dc)
        x1, x2 = O.direct_solve(S)
        return O.misfit(m.a, m.b, m.c, m.nair, freqs, m.recv, np.stack([np.concatenate([x1, x2])], 1), obs)

    h = 1e-3   # O(h^2) truncation ~1e-6; smaller steps drown in the solves' rounding
    for q, cell in enumerate(cells):
        sp_, sm_ = m.sigma.copy(), m.sigma.copy()
        sp_[cell] *= np.exp(h)
        sm_[cell] *= np.exp(-h)
        fd = (phi_of(sp_) - phi_of(sm_)) / (2 * h)
        assert sens[q] == pytest.approx(fd, rel=2e-5)


# ------------------------------------------------------------------ topography and station algebra
def _linear_in_z_fields(a, b, c, nair, alpha, beta):
    """Edge fields of two polarisations that are linear in depth: Ex = alpha +
    beta z on every x-edge of node plane k (pol 0), Ey likewise (pol 1), all
    other components 0.  z is the depth of the node plane below the datum."""
    nx, ny, nz = len(a), len(b), len(c)
    zn = np.concatenate([[0.0], np.cumsum(c)]) - c[:nair].sum()
    nl = O.nedges(nx, ny, nz)
    n1 = nx * (ny + 1) * (nz + 1)
    n2 = (nx + 1) * ny * (nz + 1)
    x1 = np.zeros(nl, complex)
    x2 = np.zeros(nl, complex)
    x1[:n1] = np.repeat(alpha + beta * zn, nx * (ny + 1))
    x2[n1:n1 + n2] = np.repeat(alpha + beta * zn, (nx + 1) * ny)
    return x1, x2, zn


def test_topography_surface_samples_its_own_plane_and_cell():
    """The topography branch (src/fdfe_main.f90:604-614, 656-666) with a
    per-column surface plane m: on fields linear in depth, E at the surface is
    exactly alpha + beta z_m and the one-sided z difference over the cell below
    the surface gives dE/dz = beta exactly, so H = beta/(i w mu).  Sampling the
    plane below (the reference's k1) or the flat-datum spacing c(nair+1) on a
    grid whose layer thicknesses vary fails one of the two."""
    rng = np.random.default_rng(21)
    nx, ny, nair = 7, 6, 3
    a, b = rng.uniform(80, 120, nx), rng.uniform(80, 120, ny)
    c = rng.uniform(20, 200, 12)
    alpha, beta, f = 0.7 - 0.2j, 1e-3 + 4e-4j, 0.3
    x1, x2, zn = _linear_in_z_fields(a, b, c, nair, alpha, beta)
    surf = rng.integers(1, 9, size=(nx, ny))
    ex, ey, hx, hy, hz = O.surface_fields(a, b, c, nair, f, x1, x2, surf)
    kap = 1j * 2 * np.pi * f * MU0
    assert np.allclose(ex[:, :, 0], alpha + beta * zn[surf], rtol=1e-14)
    assert np.allclose(ey[:, :, 1], alpha + beta * zn[surf], rtol=1e-14)
    assert np.allclose(hy[:, :, 0], beta / kap, rtol=1e-12)
    assert np.allclose(hx[:, :, 1], -beta / kap, rtol=1e-12)
    assert np.all(hz == 0) and np.all(ey[:, :, 0] == 0)
    # the flat datum is the constant plane nair
    flat = O.surface_fields(a, b, c, nair, f, x1, x2)
    same = O.surface_fields(a, b, c, nair, f, x1, x2, np.full((nx, ny), nair))
    for u, v in zip(flat, same):
        assert np.array_equal(u, v)
    # planted: the plane below the surface cell's top, or the datum spacing
    wrong = O.surface_fields(a, b, c, nair, f, x1, x2, surf + 1)
    assert not np.allclose(wrong[0][:, :, 0], alpha + beta * zn[surf], rtol=1e-6)
    assert abs(c[nair] - c[surf]).max() > 1.0   # a datum-spacing slip would change hy


def test_surface_planes_from_sigma_match_elevation_test():
    """Planes read off the carved conductivity equal the reference's test of
    cell centres against the ground elevation (src/fdfe_main.f90:606-611),
    on the topography generator at 64x64x48 (ground up to ~950 m below the
    datum, so up to several carved layers per column)."""
    m = synth.make_aniso_topography_128(nx=64, nz_earth=40, nair=8, nstations=20)
    p1 = O.surface_planes(m.a, m.b, m.c, m.sigma, 1e-6)
    p2 = O.topo_surface_planes(m.c, m.nair, m.meta["topo"])
    assert np.array_equal(p1, p2)
    assert p1.min() == m.nair and p1.max() > m.nair + 3
    # the flat model's surface is the datum everywhere
    flat = synth.make_aniso_blocks_128(nx=64, nz_earth=40, nair=8, nstations=20)
    assert np.all(O.surface_planes(flat.a, flat.b, flat.c, flat.sigma, 1e-6) == flat.nair)


def test_impedance_and_tipper_solve_the_two_polarisation_system():
    """Z and T of post_process (src/fdfe_main.f90:686-705) are defined by
    E = Z H and Hz = T (Hx, Hy) over the two polarisations.  On random fields
    the oracle's closed forms must equal numpy's solve of those 2x2 systems at
    every station; a transposed index or a sign slip in zz or tx/ty fails."""
    rng = np.random.default_rng(23)
    m = synth.make_iso_block_model("zt", 16, 16, 12, 4, seed=3, nblocks=2, freq=1.0, level=3, nstations=5)
    ex_all = rng.standard_normal((2 * m.nl, 1)) + 1j * rng.standard_normal((2 * m.nl, 1))
    res, tip = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex_all, out_type=1, rd=1)
    nl = m.nl
    flds = O.surface_fields(m.a, m.b, m.c, m.nair, m.freqs[0], ex_all[:nl, 0], ex_all[nl:, 0])
    ar1 = O.arho_interp2(1, 1, m.a, m.b, flds, m.recv)
    ar2 = O.arho_interp2(2, 1, m.a, m.b, flds, m.recv)
    for r in range(len(m.recv)):
        H = np.array([[ar1[r, 2], ar1[r, 3]], [ar2[r, 2], ar2[r, 3]]])      # rows: pol, cols: hx, hy
        zrow_x = np.linalg.solve(H, np.array([ar1[r, 0], ar2[r, 0]]))       # Ex = Zxx hx + Zxy hy
        zrow_y = np.linalg.solve(H, np.array([ar1[r, 1], ar2[r, 1]]))       # Ey = Zyx hx + Zyy hy
        t = np.linalg.solve(H, np.array([ar1[r, 4], ar2[r, 4]]))            # hz = Tx hx + Ty hy
        assert np.isclose(res[r, 0] + 1j * res[r, 1], zrow_x[1], rtol=1e-10)
        assert np.isclose(res[r, 2] + 1j * res[r, 3], zrow_y[0], rtol=1e-10)
        assert np.isclose(tip[r, 0] + 1j * tip[r, 1], t[0], rtol=1e-10)
        assert np.isclose(tip[r, 2] + 1j * tip[r, 3], t[1], rtol=1e-10)


def test_layered_model_has_no_tipper():
    """A horizontally layered model has no vertical magnetic field: with the
    3-point 1-D FD field of the same z grid as Dirichlet data the discrete
    solution is 1-D (see the layered-model test above), Hz vanishes at every
    surface cell and T = 0 at every station, while Z stays finite.  A tipper
    formula reading hx or hy where hz belongs would return O(1)."""
    nx, ny, f = 12, 12, 1.0
    rng = np.random.default_rng(6)
    a, b = rng.uniform(600, 1400, nx), rng.uniform(600, 1400, ny)
    nair = 3
    ce = np.array([200.0, 300, 500, 800, 1500, 3000])
    c = synth.with_air(ce, nair)
    nz = len(c)
    layer_sig = np.concatenate([np.full(nair, 1e-8), [0.01, 0.03, 0.1, 0.05, 0.003, 0.003]])
    sig = np.zeros((nx * ny * nz, 6))
    for k in range(nz):
        sig[k * nx * ny:(k + 1) * nx * ny, :3] = layer_sig[k]
    edc, _, _ = O.mt1d_anl(f, O.background_column(sig, nx, ny, nz), c)
    kap = 1j * 2 * np.pi * f * MU0
    n = nz + 1
    M = np.zeros((n, n), complex)
    rhs = np.zeros(n, complex)
    M[0, 0] = 1; rhs[0] = edc[0]; M[-1, -1] = 1; rhs[-1] = edc[nz]
    for k in range(1, nz):
        cm, cp = c[k - 1], c[k]
        M[k, k - 1] = -2 / (cm * (cm + cp)); M[k, k + 1] = -2 / (cp * (cm + cp))
        se = (layer_sig[k - 1] * cm + layer_sig[k] * cp) / (cm + cp)
        M[k, k] = -(M[k, k - 1] + M[k, k + 1]) - kap * se
    e1 = np.linalg.solve(M, rhs)
    S = O.build_system(a, b, c, sig, f, e_dc=np.concatenate([e1, e1]))
    lu = spla.splu(S.A.tocsc())
    ex = np.concatenate([lu.solve(S.p1), lu.solve(S.p2)])[:, None]
    xc, yc = synth.cell_centres(a), synth.cell_centres(b)
    recv = np.array([[x, y, 0.0] for x in np.linspace(xc[2], xc[-3], 4) for y in np.linspace(yc[2], yc[-3], 4)])
    res, tip = O.post_process(a, b, c, nair, [f], recv, ex, out_type=1, rd=1)
    flds = O.surface_fields(a, b, c, nair, f, ex[:S.A.shape[0], 0], ex[S.A.shape[0]:, 0])
    assert np.abs(flds[4]).max() < 1e-6 * np.abs(flds[3]).max()
    assert np.abs(res).max() > 1e-3
    assert np.abs(tip).max() < 1e-6
