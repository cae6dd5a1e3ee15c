"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Tolerances: FP64 kernels 1e-12 relative (stencil/SSOR/prolongation are the
same arithmetic in a different order); the complex64 stencil 1e-6 relative,
the complex64 SSOR inside the hybrid preconditioner 1e-5;
impedances of full solves 1e-6 relative at an interior residual of 1e-10.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from sfd_excmg_amd import sfdx, synth
from test_oracle import tiny_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sfdx.Context(0)
    yield c
    c.close()


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


# shapes: one partial tile; a second tile in x and in y holding one column/row;
# several chunks of node planes per block column
@pytest.mark.parametrize("shape", [(9, 7, 8), (33, 9, 5), (40, 17, 26)])
@pytest.mark.parametrize("aniso", [False, True])
def test_operator_fp64_matches_oracle_csr(ctx, aniso, shape):
    a, b, c, nair, sig = tiny_model(*shape, 2, aniso=aniso)
    f = 0.37
    S = O.build_system(a, b, c, sig, f)
    ve = O.row_scale(a, b, c, S.bnd)
    x = _rand(len(S.p1), 1)
    x[S.bnd] = 0
    y = ctx.apply_operator(a, b, c, sig, f, x)
    yo = ve * (S.A @ x)
    yo[S.bnd] = 0
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    # boundary coupling: A_s applied to the Dirichlet data alone is -ve * rhs
    g = np.where(S.bnd, S.p1, 0)
    y2 = ctx.apply_operator(a, b, c, sig, f, g)
    ref = -ve * S.p1
    ref[S.bnd] = 0
    assert np.abs(y2 - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(9, 7, 8), (33, 9, 5), (40, 17, 26)])
@pytest.mark.parametrize("aniso", [False, True])
def test_split_residual_matches_oracle_operator(ctx, aniso, shape):
    """sfdx_split_residual, the multigrid V-cycle's residual of a correction
    held as z + (i/wmu) G psi (formed without the curl part acting on G psi),
    against the oracle's ve-scaled CSR applied to the assembled field
    u = z + (i/wmu) G psi, G from the oracle's gradient; psi vanishes on
    boundary nodes, as in the V-cycle.  complex128 to 1e-12, complex64 to
    2e-6 of the largest entry.  The psi part must move the result by far more
    than the tolerance, so a dropped or misplaced gradient term fails."""
    nx, ny, nz = shape
    a, b, c, nair, sig = tiny_model(nx, ny, nz, 2, aniso=aniso)
    f = 0.37
    wmu = 2 * np.pi * f * O.MU0
    S = O.build_system(a, b, c, sig, f)
    ve = O.row_scale(a, b, c, S.bnd)
    nn = O.nnodes(nx, ny, nz)
    r = _rand(len(S.p1), 3)
    z = _rand(len(S.p1), 4)
    r[S.bnd] = 0
    z[S.bnd] = 0
    psi = 0.1 * _rand(nn, 5)   # G psi part ~10-50% of the result; oracle rounding < 2e-13
    idx = np.arange(nn)
    ii, jj, kk = idx % (nx + 1), (idx // (nx + 1)) % (ny + 1), idx // ((nx + 1) * (ny + 1))
    psi[(ii == 0) | (ii == nx) | (jj == 0) | (jj == ny) | (kk == 0) | (kk == nz)] = 0
    u = z + (1j / wmu) * O.grad(a, b, c, psi)
    assert np.all(u[S.bnd] == 0)
    ref = r - ve * (S.A @ u)
    ref[S.bnd] = 0
    ref0 = r - ve * (S.A @ z)
    ref0[S.bnd] = 0
    assert np.abs(ref - ref0).max() > 2e-2 * np.abs(ref).max()
    for prec, tol in ((sfdx.PREC_FP64, 1e-12), (sfdx.PREC_MIXED, 2e-6)):
        t = ctx.split_residual(a, b, c, sig, f, r, z, psi, precision=prec)
        assert np.all(t[S.bnd] == 0)
        assert np.abs(t - ref).max() <= tol * np.abs(ref).max(), prec


@pytest.mark.parametrize("model", ["halfspace32", "aniso64"])
@pytest.mark.parametrize("f", [1.0, 0.1])
def test_operator_fp32_matches_fp64(ctx, model, f):
    """complex64 stencil against complex128 to 1e-6 of the largest entry
    (measured 1.3-1.5e-7 on these models)."""
    if model == "halfspace32":
        m = synth.make_block_halfspace_32()
    else:
        m = synth.make_aniso_blocks_128(nx=64, nz_earth=40, nair=8, nstations=20)
    x = _rand(m.nl, 2)
    y64 = ctx.apply_operator(m.a, m.b, m.c, m.sigma, f, x)
    y32 = ctx.apply_operator(m.a, m.b, m.c, m.sigma, f, x, precision=sfdx.PREC_MIXED)
    assert np.abs(y32 - y64).max() <= 1e-6 * np.abs(y64).max()


def _colour_order(nx, ny, nz, bnd):
    """Interior edges sorted by SSOR colour (orientation x parity)."""
    xnl = nx * (ny + 1) * (nz + 1)
    ynl = (nx + 1) * ny * (nz + 1)
    cols = []
    for e in range(len(bnd)):
        if e < xnl:
            i = e % nx; j = (e // nx) % (ny + 1); k = e // (nx * (ny + 1)); col = (j + k) & 1
        elif e < xnl + ynl:
            q = e - xnl; i = q % (nx + 1); j = (q // (nx + 1)) % ny; k = q // ((nx + 1) * ny); col = 2 + ((i + k) & 1)
        else:
            q = e - xnl - ynl; i = q % (nx + 1); j = (q // (nx + 1)) % (ny + 1); col = 4 + ((i + j) & 1)
        cols.append(col)
    cols = np.array(cols)
    idx = np.arange(len(bnd))[~bnd]
    return idx[np.argsort(cols[~bnd], kind="stable")], cols


@pytest.mark.parametrize("omega", [1.0, 1.2])
def test_multicolour_ssor_matches_dense(ctx, omega):
    a, b, c, nair, sig = tiny_model(5, 4, 5, 1, aniso=True)
    f = 1.3
    S = O.build_system(a, b, c, sig, f)
    ve = O.row_scale(a, b, c, S.bnd)
    As = (sp.diags(ve) @ S.A).toarray()
    order, cols = _colour_order(len(a), len(b), len(c), S.bnd)
    Ap = As[np.ix_(order, order)]
    cp = cols[order]
    D = np.diag(np.diag(Ap))
    Lw = np.where(cp[:, None] > cp[None, :], Ap, 0)
    Uw = np.where(cp[:, None] < cp[None, :], Ap, 0)
    # within one colour there is no coupling
    same = (cp[:, None] == cp[None, :]) & ~np.eye(len(cp), dtype=bool)
    assert np.abs(Ap[same]).max() == 0
    M = (D / omega + Lw) @ np.linalg.inv(D) @ (D / omega + Uw) * (omega / (2 - omega))
    x = _rand(len(S.p1), 3)
    x[S.bnd] = 0
    y = ctx.ssor_apply(a, b, c, sig, f, x, omega=omega)
    assert np.all(y[S.bnd] == 0)
    yo = np.linalg.solve(M, x[order])
    assert np.abs(y[order] - yo).max() <= 1e-11 * np.abs(yo).max()


# Larger and odd shapes than the dense check (a colour row of nx = 300 spans
# two 256-thread blocks; odd nx, ny, nz put both parities at the row ends);
# M^-1 applied by two sparse triangular solves.
@pytest.mark.parametrize("shape,omega", [((300, 5, 4), 1.0), ((23, 19, 18), 1.0), ((20, 17, 21), 1.3)])
def test_multicolour_ssor_matches_triangular_solves(ctx, shape, omega):
    import scipy.sparse.linalg as spla
    nx, ny, nz = shape
    a, b, c, nair, sig = tiny_model(nx, ny, nz, 2, aniso=True)
    f = 0.7
    S = O.build_system(a, b, c, sig, f)
    ve = O.row_scale(a, b, c, S.bnd)
    order, cols = _colour_order(nx, ny, nz, S.bnd)
    Ap = (sp.diags(ve) @ S.A).tocsr()[order][:, order].tocsr()
    cp = cols[order]
    d = Ap.diagonal()
    C = Ap.tocoo()
    low, up = cp[C.row] > cp[C.col], cp[C.row] < cp[C.col]
    n = len(order)
    Lw = sp.csr_matrix((C.data[low], (C.row[low], C.col[low])), shape=(n, n)) + sp.diags(d / omega)
    Uw = sp.csr_matrix((C.data[up], (C.row[up], C.col[up])), shape=(n, n)) + sp.diags(d / omega)
    x = _rand(len(S.p1), 11)
    x[S.bnd] = 0
    y = ctx.ssor_apply(a, b, c, sig, f, x, omega=omega)
    assert np.all(y[S.bnd] == 0)
    t = spla.spsolve_triangular(Lw.tocsr(), x[order], lower=True)
    yo = spla.spsolve_triangular(Uw.tocsr(), d * t, lower=False) * ((2 - omega) / omega)
    assert np.abs(y[order] - yo).max() <= 1e-11 * np.abs(yo).max()


def _ssor_colour_ref(nx, ny, nz, As, bnd, x, omega):
    """M^-1 x of the multicolour SSOR of As (colour order, two sparse triangular solves)."""
    import scipy.sparse.linalg as spla
    order, cols = _colour_order(nx, ny, nz, bnd)
    Ap = As.tocsr()[order][:, order].tocsr()
    cp = cols[order]
    d = Ap.diagonal()
    C = Ap.tocoo()
    low, up = cp[C.row] > cp[C.col], cp[C.row] < cp[C.col]
    n = len(order)
    Lw = sp.csr_matrix((C.data[low], (C.row[low], C.col[low])), shape=(n, n)) + sp.diags(d / omega)
    Uw = sp.csr_matrix((C.data[up], (C.row[up], C.col[up])), shape=(n, n)) + sp.diags(d / omega)
    t = spla.spsolve_triangular(Lw.tocsr(), x[order], lower=True)
    y = np.zeros_like(x)
    y[order] = spla.spsolve_triangular(Uw.tocsr(), d * t, lower=False) * ((2 - omega) / omega)
    return y


@pytest.mark.parametrize("shape", [(5, 4, 5), (20, 17, 9)])
@pytest.mark.parametrize("precision,nodal", [(sfdx.PREC_FP64, 1), (sfdx.PREC_MIXED, 1), (sfdx.PREC_FP64, 0)])
def test_hybrid_preconditioner_matches_cpu_construction(ctx, shape, precision, nodal):
    """The solvers' preconditioner z = SSOR_e(r) + (i/wmu) G phi on the interior
    nodes, against a CPU construction: the edge SSOR by triangular solves in
    colour order on the lumped-mass operator (the oracle's ve-scaled operator of
    the model with its off-diagonal conductivities set to 0: the same
    stiffness and diagonal mass, no cross-orientation mass couplings), G from
    the oracle's gradient, and P from G^T A_s G = -i wmu P
    (A_s the oracle's ve-scaled operator; the test checks that this P is real,
    i.e. that the curl part of A_s annihilates G).  phi: nodal 1, the two-step
    Jacobi polynomial w(2-w) u - w^2 D^-1 (P - D) u, u = D^-1 G^T r, w = 0.8;
    nodal 0, symmetric Gauss-Seidel on P in the node colours
    (i&1) + 2(j&1) + 4(k&1): (D + U)^-1 D (D + L)^-1 G^T r.  FP64 to 1e-10;
    the complex64 edge SSOR of the mixed path to 1e-5."""
    nx, ny, nz = shape
    a, b, c, nair, sig = tiny_model(nx, ny, nz, 2, aniso=True)
    f = 0.9
    wmu = 2 * np.pi * f * O.MU0
    S = O.build_system(a, b, c, sig, f)
    ve = O.row_scale(a, b, c, S.bnd)
    As = (sp.diags(ve) @ S.A).tocsr()
    nn = O.nnodes(nx, ny, nz)
    rows, cols, vals = [], [], []
    for m in range(nn):
        e = np.zeros(nn)
        e[m] = 1
        g = O.grad(a, b, c, e)
        nz_ = np.flatnonzero(g)
        rows.append(nz_); cols.append(np.full(len(nz_), m)); vals.append(g[nz_])
    Gm = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(S.p1), nn))
    idx = np.arange(nn)
    ii, jj, kk = idx % (nx + 1), (idx // (nx + 1)) % (ny + 1), idx // ((nx + 1) * (ny + 1))
    inner = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)
    Gi = Gm[:, inner]
    Pc = (Gi.T @ As @ Gi).toarray() / (-1j * wmu)
    assert np.abs(Pc.imag).max() <= 1e-9 * np.abs(Pc.real).max()
    P = Pc.real
    r = _rand(len(S.p1), 12)
    r[S.bnd] = 0
    t = (Gm.T @ r)[inner]
    D = np.diag(P)
    if nodal == 1:
        u = t / D
        phi_i = 0.8 * 1.2 * u - 0.64 * (P @ u - D * u) / D
    else:
        import scipy.linalg as sla
        col = ((ii & 1) + 2 * (jj & 1) + 4 * (kk & 1))[inner]
        o = np.argsort(col, kind="stable")
        Po, co = P[np.ix_(o, o)], col[o]
        lower = np.where(co[:, None] > co[None, :], Po, 0) + np.diag(np.diag(Po))
        upper = np.where(co[:, None] < co[None, :], Po, 0) + np.diag(np.diag(Po))
        yv = sla.solve_triangular(lower, t[o], lower=True)
        phi_i = np.zeros_like(t)
        phi_i[o] = sla.solve_triangular(upper, np.diag(Po) * yv, lower=False)
    phi = np.zeros(nn, complex)
    phi[inner] = phi_i
    sig_l = sig.copy()
    sig_l[:, 3:] = 0
    Al = (sp.diags(ve) @ O.build_system(a, b, c, sig_l, f).A).tocsr()
    assert abs(Al - As).max() > 1e-6 * abs(As).max()    # the lumping changes the operator
    z = _ssor_colour_ref(nx, ny, nz, Al, S.bnd, r, 1.0) + (1j / wmu) * (Gm @ phi)
    y = ctx.hybrid_apply(a, b, c, sig, f, r, omega=1.0, nodal=nodal, precision=precision)
    assert np.all(y[S.bnd] == 0)
    tol = 1e-10 if precision == sfdx.PREC_FP64 else 1e-5
    assert np.abs(y - z).max() <= tol * np.abs(z).max()


def test_extint_v_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    nx, ny, nz, nair = 16, 24, 16, 4
    a, b, c = rng.uniform(50, 150, nx), rng.uniform(50, 150, ny), rng.uniform(50, 150, nz)
    u0 = _rand(O.nedges(nx // 4, ny // 4, nz // 4), 6)
    u1 = _rand(O.nedges(nx // 2, ny // 2, nz // 2), 7)
    out = ctx.extint_v(a, b, c, nair, u0, u1)
    ref = O.extint_v(a, b, c, nair, u0, u1)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("aniso", [False, True])
def test_divergence_correction_matches_oracle(ctx, aniso):
    a, b, c, nair, sig = tiny_model(8, 7, 9, 2, aniso=aniso)
    x = _rand(O.nedges(len(a), len(b), len(c)), 8)
    xg, itg = ctx.div_correct(a, b, c, sig, x, corr_tol=1e-11)
    xo, ito = O.div_correct(a, b, c, sig, x, tol=1e-11)
    assert itg > 0
    assert np.abs(xg - xo).max() <= 1e-8 * np.abs(xo).max()


def _impedances(m, ex, rd=2):
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=rd)
    return res[:, 0] + 1j * res[:, 1], res[:, 2] + 1j * res[:, 3]


@pytest.mark.parametrize("k", [1, 2, 3])
def test_device_mesh_coarsen_equals_host_and_oracle(ctx, k):
    """sfdx_mesh_coarsen_device (k_mesh_coarsen: sigma averaged on the device,
    as sfdx_fwd_solve builds its coarse levels) gives the host coarsening's
    bits and the oracle's mesh_coarsen (fdfe_main.f90:18-84, index fixed) to
    1e-13, on a full-tensor model with non-uniform widths."""
    m = synth.make_aniso_blocks_128(nx=32, nz_earth=20, nair=4, nstations=8)
    a1, b1, c1, s1 = ctx.mesh_coarsen(m.a, m.b, m.c, m.sigma, k)
    a0, b0, c0, s0 = sfdx.mesh_coarsen(m.a, m.b, m.c, m.sigma, k)
    a2, b2, c2, s2 = O.mesh_coarsen(m.a, m.b, m.c, m.sigma, k)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0) and np.array_equal(c1, c0)
    assert np.array_equal(s1, s0)
    assert np.allclose(s1, s2, rtol=1e-13, atol=1e-15 * np.abs(s2).max())


def small_excmg_model(aniso=False):
    m = synth.make_iso_block_model("s16", 16, 16, 12, 4, seed=11, nblocks=3, freq=1.0, level=3, nstations=6)
    if aniso:
        rng = np.random.default_rng(3)
        e = slice(m.nair * m.nx * m.ny, None)
        n = m.sigma[e].shape[0]
        m.sigma[e, 3:] = 0.2 * m.sigma[e, :1] * rng.uniform(-1, 1, (n, 3))
    return m


@pytest.mark.parametrize("precond", [sfdx.PRECOND_HYBRID, sfdx.PRECOND_AUTO])
@pytest.mark.parametrize("solver", [sfdx.SOLVER_BICGSTAB, sfdx.SOLVER_GPBICG, sfdx.SOLVER_COCR])
@pytest.mark.parametrize("aniso", [False, True])
def test_fwd_solve_fp64_matches_oracle(ctx, solver, aniso, precond):
    """Each Krylov method on the hybrid preconditioner (whose solves run the
    periodic divergence correction every 50 iterations) and on the default
    choice (the V-cycle, converging before the first correction) reaches the
    oracle's impedances to 1e-6."""
    m = small_excmg_model(aniso)
    opts = sfdx.default_opts(level=3, solver=solver, eps=1e-10, corr_intv=50, norm=sfdx.NORM_INTERIOR,
                             precond=precond)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    if precond == sfdx.PRECOND_HYBRID:
        assert max(st.dc_calls[2]) >= 1
    oopt = O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior")
    exo, _ = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs, oopt)
    z1 = _impedances(m, ex)
    z2 = _impedances(m, exo)
    for p, q in zip(z1, z2):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("precond", [sfdx.PRECOND_HYBRID, sfdx.PRECOND_MG])
def test_fwd_solve_mixed_matches_fp64(ctx, precond):
    """complex64 preconditioner inside FP64 COCR with restarts from the true
    residual: same impedances as the FP64 path to 1e-6."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR, precond=precond)
    ex64, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(**kw))
    exmx, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs,
                             sfdx.default_opts(precision=sfdx.PREC_MIXED, restart_drop=1e-3, sigma_floor32=0.0, **kw))
    assert st.status == 0
    assert st.refine_steps[0] >= 1
    for p, q in zip(_impedances(m, exmx), _impedances(m, ex64)):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("solver", [sfdx.SOLVER_BICGSTAB, sfdx.SOLVER_GPBICG])
def test_fwd_solve_mixed_bicg_family_matches_fp64(ctx, solver):
    """BiCGStab and GPBiCG with the complex64 preconditioner (restarts from the
    true residual, convergence confirmed on it): same impedances as FP64."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=solver, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    ex64, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(**kw))
    exmx, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs,
                             sfdx.default_opts(precision=sfdx.PREC_MIXED, restart_drop=1e-3, sigma_floor32=0.0, **kw))
    assert st.status == 0
    for p, q in zip(_impedances(m, exmx), _impedances(m, ex64)):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("merge", [0.0, 3.0])
@pytest.mark.parametrize("use_excmg,mg_levels", [(1, 0), (0, 2), (0, 0)])
def test_fwd_solve_multigrid_matches_oracle(ctx, use_excmg, mg_levels, merge):
    """Multigrid-preconditioned COCR (split edge/nodal V-cycle) reaches the
    oracle's solution, over 2:1 grids (merge 0; the hierarchy may be deeper
    than the EXCMG levels) and over merge-ratio grids (3, the default)."""
    m = small_excmg_model(True)
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                             precond=sfdx.PRECOND_MG, use_excmg=use_excmg, mg_levels=mg_levels,
                             mg_merge_ratio=merge)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    oopt = O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior")
    exo, _ = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs, oopt)
    for p, q in zip(_impedances(m, ex), _impedances(m, exo)):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_device_coarse_levels_match_direct_solves(ctx, precision):
    """EXCMG levels 1-2 solved by the device Krylov method (coarse_solver 1,
    the default) and by the host direct solves (0, the reference's u1u2):
    same impedances to 1e-6; only the device path reports Krylov iterations
    on the two coarse levels."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
              precision=precision, sigma_floor32=0.0)
    assert sfdx.default_opts().coarse_solver == 1
    exd, sd = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(coarse_solver=0, **kw))
    exk, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(**kw))
    assert st.status == 0 and sd.status == 0
    assert min(st.iters[0][0], st.iters[0][1], st.iters[1][0], st.iters[1][1]) > 0
    assert max(sd.iters[0][0], sd.iters[0][1], sd.iters[1][0], sd.iters[1][1]) == 0
    for p, q in zip(_impedances(m, exk), _impedances(m, exd)):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("merge", [0.0, 3.0])
@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_auto_preconditioner_chooses_per_frequency(ctx, precision, merge):
    """SFDX_PRECOND_AUTO on a 32x32x24 full-tensor block model (4 km wide,
    smallest cells 50 m, first 2:1 coarse cells 100 m, most conductive cell
    0.84 S/m, median 0.01 S/m) at 1 mHz, 4 Hz and 1 kHz.  With merge-ratio
    grids every level takes the V-cycle except at 1 kHz, where the skin depth
    in the most conductive cell (17 m) is below half the cells of every level
    (50, 100, 200 m); with 2:1 grids (merge 0) only at 4 Hz (at 1 mHz the
    skin depth at the median conductivity, 159 km, exceeds the mesh).  Each
    frequency's field has the bits of the solve with its choice fixed, and
    at 4 Hz the V-cycle and hybrid impedances agree to 1e-6."""
    m = synth.make_aniso_blocks_128(nx=32, nz_earth=20, nair=4, nstations=8)
    m.freqs = np.array([1e-3, 4.0, 1e3])
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
              precision=precision, sigma_floor32=0.0, restart_drop=1e-3, mg_merge_ratio=merge)
    exa, sa = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs,
                            sfdx.default_opts(precond=sfdx.PRECOND_AUTO, **kw))
    choice = [True, True, False] if merge > 0 else [False, True, False]   # V-cycle per frequency
    assert sa.status == 0 and sa.mg_freqs == sum(choice)
    exh, sh = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs[~np.array(choice)],
                            sfdx.default_opts(precond=sfdx.PRECOND_HYBRID, **kw))
    exm, sm = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs[np.array(choice)],
                            sfdx.default_opts(precond=sfdx.PRECOND_MG, **kw))
    assert sh.status == 0 and sh.mg_freqs == 0 and sm.status == 0 and sm.mg_freqs == sum(choice)
    ih = im = 0
    for k, mg in enumerate(choice):
        if mg:
            assert np.array_equal(exa[:, k], exm[:, im]), k
            im += 1
        else:
            assert np.array_equal(exa[:, k], exh[:, ih]), k
            ih += 1
    exh, sh = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs[1:2],
                            sfdx.default_opts(precond=sfdx.PRECOND_HYBRID, **kw))
    exm, sm = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs[1:2],
                            sfdx.default_opts(precond=sfdx.PRECOND_MG, **kw))
    m1 = synth.make_aniso_blocks_128(nx=32, nz_earth=20, nair=4, nstations=8)
    m1.freqs = m.freqs[1:2]
    for p, q in zip(_impedances(m1, exm), _impedances(m1, exh)):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


def test_spmv_tiles_and_autotuner_give_the_same_solve(ctx):
    """Every SpMV kernel variant (0: the cell-plane march in complex128, 1..4
    the k-marching tiles) and the autotuner (-1) solve the same system:
    impedances agree to 1e-9 (only the rounding order of the rows and the
    grouping of the fused reductions differ); a fixed variant is reported as
    used, the autotuned one is a valid variant, a second autotuned solve on
    the same context reuses the cached choice, and variant 5 is rejected."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    ref, st0 = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(**kw))
    assert st0.status == 0 and st0.spmv_tile == 0
    zref = _impedances(m, ref)
    for tile in (1, 2, 3, 4, -1):
        ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(spmv_tile=tile, **kw))
        assert st.status == 0
        assert st.spmv_tile == tile if tile >= 0 else 0 <= st.spmv_tile <= 4
        for p, q in zip(_impedances(m, ex), zref):
            assert np.max(np.abs(p - q) / np.abs(q)) < 1e-9, tile
    ex2, st2 = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(spmv_tile=-1, **kw))
    assert st2.spmv_tile == st.spmv_tile
    for p, q in zip(_impedances(m, ex2), zref):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-9
    with pytest.raises(sfdx.SfdxError):
        ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(spmv_tile=5, **kw))


def test_complex64_krylov_option_reports_nonconvergence_with_air(ctx):
    """PREC_MIXED_IR (complex64 inner Krylov solves in FP64 refinement) cannot
    converge with 1e-8 S/m air: the solve must say so (SFDX_WARN_NOCONV)
    rather than return its fields as converged."""
    m = small_excmg_model(False)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
              precond=sfdx.PRECOND_HYBRID)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs,
                           sfdx.default_opts(precision=sfdx.PREC_MIXED_IR, **kw))
    assert st.status == -1      # SFDX_WARN_NOCONV


def test_full_tensor_boundary_matches_oracle(ctx):
    """Full-tensor 1-D boundary (cross-polar Dirichlet components, a deliberate
    deviation from the reference's diagonal-only background): GPU impedances
    equal the oracle's with the same boundary to 1e-6, and differ from the
    reference boundary's on a model whose first column is anisotropic."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    exf, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(full_background=1, **kw))
    assert st.status == 0
    exr, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(**kw))
    oopt = O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior", full_background=True)
    exo, _ = O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs, oopt)
    zf, zo, zr = _impedances(m, exf), _impedances(m, exo), _impedances(m, exr)
    for p, q in zip(zf, zo):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6
    assert max(np.max(np.abs(p - q) / np.abs(q)) for p, q in zip(zf, zr)) > 1e-4


@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_c1_workload_matches_oracle_fixture(ctx, precision):
    """The reference's CPU case at full size (32x32x32 block half-space, 3 EXCMG
    levels, 1 Hz, 10x10 stations): GPU impedances against the oracle's
    (tests/golden/c1_oracle_impedances.npz, interior residual 1e-11) to 1e-6."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "c1_oracle_impedances.npz"))
    m = synth.make_block_halfspace_32()
    assert float(np.sum(m.sigma)) == float(g["sigma_sum"]) and np.array_equal(np.asarray(m.recv, float), g["recv"])
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                             precision=precision)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    zxy, zyx = _impedances(m, ex)
    for p, q in ((zxy, g["zxy"]), (zyx, g["zyx"])):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_aniso64_workload_matches_oracle_fixture(ctx, precision):
    """The benchmark model's generator at half resolution (64x64x48 cells, 16
    seeded blocks of rotated full tensors, 0.1 Hz, 4 EXCMG levels, 20x20
    stations): GPU impedances against the oracle's
    (tests/golden/aniso64_oracle_impedances.npz, interior residual 1e-10) to
    1e-6."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "aniso64_oracle_impedances.npz"))
    m = synth.make_aniso_blocks_128(nx=64, nz_earth=40, nair=8, nstations=20)
    assert float(np.sum(m.sigma)) == float(g["sigma_sum"]) and np.array_equal(np.asarray(m.recv, float), g["recv"])
    opts = sfdx.default_opts(level=m.level, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0,
                             norm=sfdx.NORM_INTERIOR, precision=precision, maxit_excmg=8000)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    zxy, zyx = _impedances(m, ex)
    for p, q in ((zxy, g["zxy"]), (zyx, g["zyx"])):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6
    # the library's own post-process (host FP64) of the same fields
    rl, _ = sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2)
    for p, q in ((rl[:, 0] + 1j * rl[:, 1], g["zxy"]), (rl[:, 2] + 1j * rl[:, 3], g["zyx"])):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6
    # the device post-process (sfdx_fwd_response: surface fields on the GPU,
    # stations on the host) against the oracle's post-process of oracle fields
    rr, _, st2 = ctx.fwd_response(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, opts, out_type=1, rd=2)
    assert st2.status == 0
    for p, q in ((rr[:, 0] + 1j * rr[:, 1], g["zxy"]), (rr[:, 2] + 1j * rr[:, 3], g["zyx"])):
        assert np.max(np.abs(p - q) / np.abs(q)) < 1e-6


def test_halfspace_on_gpu(ctx):
    m = synth.make_block_halfspace_32()
    m.sigma[m.nair * m.nx * m.ny:, :3] = 0.01
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=200, norm=sfdx.NORM_INTERIOR)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    res, _ = sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=2)
    assert np.allclose(res[:, :2], 100.0, rtol=0.02)
    assert np.allclose(res[:, 2:], 45.0, atol=0.5)


def test_fortran_driver_matches_library(ctx, tmp_path):
    """The drop-in Fortran driver (reference command line and file formats)
    reads <name>.fwd/.dat, solves through the C-ABI and writes the same
    apparent resistivities/phases as the library called directly."""
    import subprocess
    drv = os.path.join(ROOT, "sfd_excmg_amd", "bin", "sfd_excmg_gpu")
    assert os.path.exists(drv), "Fortran driver not built"
    m0 = small_excmg_model(True)
    m0.out_type = 2
    prefix = str(tmp_path / "drv")
    synth.write_fwd(m0, prefix)
    synth.write_dat(m0, prefix)
    m = synth.read_fwd_dat(prefix, sigma_air=1e-8)
    env = dict(os.environ, SFDX_FWD_ACC="1", SFDX_SOLVER="5", SFDX_NORM="interior", SFDX_RES_FULL="1",
               SFDX_SIGMA_AIR="1e-8", SFDX_CORR_INTV="0")
    print("running", drv, prefix, flush=True)
    out = subprocess.run([drv, prefix, "3", "1e-10"], cwd=str(tmp_path), env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = np.loadtxt(prefix + ".resx")
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    ref, _ = sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=2, rd=2)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-6


def test_fortran_driver_reads_topography(ctx, tmp_path):
    """The driver with a .topo file (the corrected reader of file_opt.f90:97-128):
    from the uncarved model and the ground elevations it carves sigma itself,
    takes each column's ground surface in post_process, and writes what the
    library gives on the same carved model with the surface planes read off
    sigma."""
    import subprocess
    drv = os.path.join(ROOT, "sfd_excmg_amd", "bin", "sfd_excmg_gpu")
    assert os.path.exists(drv), "Fortran driver not built"
    carved = synth.make_aniso_topography_128(nx=32, nz_earth=20, nair=4, nstations=8)
    m0 = synth.make_aniso_blocks_128(nx=32, nz_earth=20, nair=4, nstations=8)   # the same model before carving
    m0.meta["topo"] = carved.meta["topo"]
    m0.out_type = 1
    prefix = str(tmp_path / "topo")
    synth.write_fwd(m0, prefix, topo_indic=1)
    synth.write_dat(m0, prefix)
    synth.write_topo(m0, prefix)
    m = synth.read_fwd_dat(prefix, sigma_air=1e-8)          # reads and carves like the driver
    assert np.allclose(m.sigma, carved.sigma, rtol=1e-9, atol=0)   # the text format keeps 11 digits
    env = dict(os.environ, SFDX_FWD_ACC="1", SFDX_SOLVER="5", SFDX_NORM="interior", SFDX_RES_FULL="1",
               SFDX_SIGMA_AIR="1e-8", SFDX_CORR_INTV="0")
    print("running", drv, prefix, flush=True)
    out = subprocess.run([drv, prefix, "3", "1e-10"], cwd=str(tmp_path), env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = np.loadtxt(prefix + ".resx")
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    assert st.status == 0
    surf = sfdx.surface_planes(m.a, m.b, m.c, m.sigma, 1e-6)
    assert surf.max() > m.nair
    ref, _ = sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=1, rd=2, surf=surf)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_graph_replay_matches_eager(ctx, precision):
    """The steady-state COCR iteration replayed as a captured HIP graph gives
    bit-identical iterates to eager launches (the reductions are deterministic),
    across hundreds of replays interleaved with eager restarts and copies."""
    m = small_excmg_model(True)
    kw = dict(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
              precision=precision, sigma_floor32=0.0)
    ex0, st0 = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(use_graph=0, **kw))
    ex1, st1 = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, sfdx.default_opts(use_graph=1, **kw))
    assert st0.status == 0 and st1.status == 0
    assert st0.fine_iters_total == st1.fine_iters_total
    assert np.array_equal(ex0, ex1)


@pytest.mark.parametrize("precision", [sfdx.PREC_FP64, sfdx.PREC_MIXED])
def test_multi_frequency_solve_runs_items_with_identical_fields(ctx, precision):
    """sfdx_fwd_solve over several frequencies runs them as concurrent batch
    items (pol_streams 2, the default): fields bit-identical to the frequency
    loop on the context's stream (pol_streams 1)."""
    m = small_excmg_model(True)
    freqs = np.array([0.3, 30.0, 3.0])
    out = {}
    for ps in (2, 1):
        o = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-9, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                              precision=precision, pol_streams=ps)
        ex, st = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, freqs, o)
        assert st.status == 0
        out[ps] = ex
    assert out[2].shape[1] == 3
    assert np.array_equal(out[1], out[2])


@pytest.mark.parametrize("use_excmg", [0, 1])
def test_batched_solve_matches_single_solves(ctx, use_excmg):
    """sfdx_solve_batch over models of uneven sizes and several frequencies,
    run concurrently on several streams, returns what one-model sfdx_fwd_solve
    calls return.  Without EXCMG the whole solve is device arithmetic and the
    fields are bit-identical.  With EXCMG the two coarsest levels are host
    PARDISO solves, which run in parallel threads and may round differently;
    the converged solutions then agree at the solver tolerance, checked on the
    impedances."""
    ms = [synth.make_iso_block_model("b0", 16, 16, 12, 4, seed=21, nblocks=2, freq=1.0, level=3, nstations=4),
          synth.make_iso_block_model("b1", 24, 16, 16, 8, seed=22, nblocks=2, freq=1.0, level=3, nstations=4),
          small_excmg_model(True)]
    ms[0].freqs = np.array([0.1, 10.0])
    ms[1].freqs = np.array([1.0, 3.0, 30.0])
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                             use_excmg=use_excmg)
    exb, stb = ctx.solve_batch(ms, opts, nstreams=4)
    for m, ex, st in zip(ms, exb, stb):
        ex1, st1 = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
        assert st.status == 0 and st1.status == 0
        assert ex.shape == ex1.shape
        if not use_excmg:
            assert np.array_equal(ex, ex1)
        # rd=1: the reference's transposed station sampling (rd=2) needs nx == ny
        for p, q in zip(_impedances(m, ex, rd=1), _impedances(m, ex1, rd=1)):
            assert np.max(np.abs(p - q) / np.abs(q)) < 1e-7


def _adjoint_case():
    m = small_excmg_model(True)
    m.freqs = np.array([1.0, 10.0])
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv,
                            O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs,
                                         O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))[0],
                            out_type=2)
    obs = res * np.array([1.15, 0.9, 1.0, 1.0]) + np.array([0, 0, 2.0, -1.5])
    nx, ny = m.nx, m.ny
    cells = [(m.nair + 1) * nx * ny + 8 * nx + 8, (m.nair + 2) * nx * ny + 6 * nx + 9,
             (m.nair + 4) * nx * ny + 9 * nx + 6]
    return m, obs, cells


@pytest.mark.parametrize("precond", [sfdx.PRECOND_HYBRID, sfdx.PRECOND_MG])
def test_adjoint_gradient_matches_oracle_sensitivity_and_fd(ctx, precond):
    """d misfit / d log(s_c) from the device adjoint (forward + adjoint COCR
    solves with the forward kernels, cell kernel over the 12 edges; the
    hybrid preconditioner, or the V-cycle over merge-ratio grids in the forward
    and the adjoint solves) against the FP64 CPU tangent-linear sensitivity of
    the oracle and against central differences of device forward solves."""
    m, obs, cells = _adjoint_case()
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-12, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                             maxit_plain=5000, maxit_excmg=5000, precond=precond)
    phi, grad, st = ctx.adjoint(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, obs, opts)
    assert st.status == 0
    sens = O.misfit_sensitivity(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, obs, cells)
    scale = np.max(np.abs(sens))
    assert np.max(np.abs(grad[cells] - sens)) <= 1e-4 * scale
    h = 1e-3
    for q, cell in enumerate(cells[:2]):
        sp_, sm_ = m.sigma.copy(), m.sigma.copy()
        sp_[cell] *= np.exp(h)
        sm_[cell] *= np.exp(-h)
        exp_, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, sp_, m.freqs, opts)
        exm_, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, sm_, m.freqs, opts)
        fd = (sfdx.misfit(m.a, m.b, m.c, m.nair, m.freqs, m.recv, exp_, obs)
              - sfdx.misfit(m.a, m.b, m.c, m.nair, m.freqs, m.recv, exm_, obs)) / (2 * h)
        assert abs(grad[cell] - fd) <= 1e-3 * scale


def test_adjoint_frequencies_as_concurrent_items_sum_single_frequency_calls(ctx):
    """sfdx_adjoint runs its frequencies as concurrent items on worker streams
    (higher frequencies first) and sums the gradient afterwards in frequency
    order: misfit and gradient over four unsorted frequencies equal the sums of
    one-frequency calls (which run on the context itself), to rounding."""
    m = small_excmg_model(True)
    m.freqs = np.array([1.0, 10.0, 3.0, 0.3])
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv,
                            O.fwd_solver(m.a, m.b, m.c, m.nair, 3, 1e-11, m.sigma, m.freqs,
                                         O.Options(solver=O.SOLVER_COCR, corr_intv=50, norm="interior"))[0],
                            out_type=2)
    obs = res * np.array([1.1, 0.95, 1.0, 1.0]) + np.array([0, 0, 1.0, -2.0])
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-12, corr_intv=0, norm=sfdx.NORM_INTERIOR,
                             maxit_plain=5000, maxit_excmg=5000)
    phi, grad, st = ctx.adjoint(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, obs, opts)
    assert st.status == 0
    nrec = len(m.recv)
    phis, grads = [], []
    for f, fr in enumerate(m.freqs):
        p1, g1, s1 = ctx.adjoint(m.a, m.b, m.c, m.nair, m.sigma, np.array([fr]), m.recv,
                                 obs[f * nrec:(f + 1) * nrec], opts)
        assert s1.status == 0
        phis.append(p1)
        grads.append(g1)
    assert abs(phi - sum(phis)) <= 1e-12 * abs(phi)
    gsum = np.sum(grads, axis=0)
    assert np.max(np.abs(grad - gsum)) <= 1e-12 * np.max(np.abs(gsum))
    assert np.max(np.abs(grad)) > 0


def test_autograd_function_returns_adjoint_gradient(ctx):
    import torch
    from sfd_excmg_amd.autograd import mt_misfit

    m, obs, cells = _adjoint_case()
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    ls = torch.zeros(m.nx * m.ny * m.nz, dtype=torch.float64, requires_grad=True)
    phi = mt_misfit(ls, m, obs, opts, context=ctx)
    phi.backward()
    phi2, grad, _ = ctx.adjoint(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, obs, opts)
    assert float(phi.detach()) == pytest.approx(phi2, rel=1e-12)
    assert np.allclose(ls.grad.numpy(), grad, rtol=0, atol=1e-12 * np.abs(grad).max())


@pytest.mark.parametrize("out_type,rd", [(1, 2), (2, 1)])
def test_fwd_response_matches_host_post_process(ctx, out_type, rd):
    """Surface fields evaluated on the device (only they leave it) give the same
    station responses as the host post-process of the full fields."""
    m = small_excmg_model(True)
    m.freqs = np.array([1.0, 4.0])
    opts = sfdx.default_opts(level=3, solver=sfdx.SOLVER_COCR, eps=1e-10, corr_intv=0, norm=sfdx.NORM_INTERIOR)
    res, tip, st = ctx.fwd_response(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, m.recv, opts, out_type=out_type, rd=rd)
    assert st.status == 0
    ex, _ = ctx.fwd_solve(m.a, m.b, m.c, m.nair, m.sigma, m.freqs, opts)
    ref, reft = sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, out_type=out_type, rd=rd)
    assert np.max(np.abs(res - ref) / np.abs(ref)) < 1e-10
    assert np.max(np.abs(tip - reft)) < 1e-10 * np.abs(reft).max()
