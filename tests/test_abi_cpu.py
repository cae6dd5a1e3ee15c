"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/sfdx.h, and its host-side FP64 parts (1-D background,
post-process, coarse direct solve) agree with the oracle."""
import os
import re

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from sfd_excmg_amd import sfdx, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sfdx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sfdx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = sfdx.load()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(sfdx.EXPORTS)
    # every entry point with parameters has its ctypes prototype (floats would
    # otherwise not convert to double)
    for n in names:
        if n != "sfdx_last_error":
            assert getattr(lib, n).argtypes is not None, n


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Size and every field offset of sfdx_grid, sfdx_opts, sfdx_stats and
    sfdx_job as gcc lays them out from include/sfdx.h equal the ctypes
    structures of the binding."""
    import ctypes
    import subprocess

    structs = {"sfdx_grid": sfdx.Grid, "sfdx_opts": sfdx.Opts, "sfdx_stats": sfdx.Stats, "sfdx_job": sfdx.Job}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sfdx.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_default_options_match_reference_parameters():
    o = sfdx.default_opts()
    assert o.corr_intv == 200 and o.corr_tol == 1e-3 and o.ssor_omega == 1.0   # global_para.f90:9-10, solvers.f90:14
    assert o.maxit_plain == 5000 and o.maxit_excmg == 1000                    # fdfe_main.f90:993,1414
    assert o.spmv_tile == 0      # fixed SpMV tile by default: repeated runs stay bit-identical
    assert o.precond == sfdx.PRECOND_AUTO and o.mg_merge_ratio == 3.0   # the V-cycle over merge-ratio grids


@pytest.mark.parametrize("aniso", [False, True])
def test_mt1d_matches_oracle(aniso):
    rng = np.random.default_rng(4)
    nz = 30
    c = rng.uniform(20, 800, nz)
    conv = np.zeros((nz, 6))
    conv[:, :3] = 10 ** rng.uniform(-3, -1, (nz, 3))
    if aniso:
        conv[:, 3:] = 0.1 * conv[:, :1] * rng.uniform(-0.5, 0.5, (nz, 3))
    for f in (1e-3, 0.1, 10.0):
        e1, h1, i1 = sfdx.mt1d_anl(f, conv, c)
        e2, h2, i2 = O.mt1d_anl(f, conv, c)
        assert np.allclose(e1, e2, rtol=1e-11, atol=1e-14 * abs(e2).max())
        assert np.allclose(h1, h2, rtol=1e-11, atol=1e-14 * abs(h2).max())
        assert np.allclose(i1, i2, rtol=1e-10)


def _te_profile(sig, c, f):
    """E at the nodes of a layered column (E = 1 at the top, the last layer
    continued as the half-space below), by the impedance recursion from the
    bottom and decaying exponentials only (the reference's time convention,
    k^2 = -i w mu sigma): an independent, overflow-free 1-D solution."""
    mu = 4e-7 * np.pi
    w = 2 * np.pi * f
    k = np.sqrt(-1j * w * mu * sig)
    k = np.where(k.real < 0, -k, k)
    zeta = -1j * w * mu / k
    n = len(sig)
    Z = np.zeros(n + 1, complex)
    Z[n] = zeta[n - 1]
    for L in range(n - 1, -1, -1):
        t = np.tanh(k[L] * c[L]) if abs(k[L] * c[L]) < 300 else 1.0
        Z[L] = zeta[L] * (Z[L + 1] + zeta[L] * t) / (zeta[L] + Z[L + 1] * t)
    E = np.ones(n + 1, complex)
    for L in range(n):
        a = zeta[L] / Z[L + 1]
        E[L + 1] = E[L] * 2 * np.exp(-k[L] * c[L]) / ((1 + a) + (1 - a) * np.exp(-2 * k[L] * c[L]))
    return E


def test_mt1d_thick_columns_match_a_stable_recursion():
    """A column with 20 thin layers over 10 layers growing to 41 km, from
    0.1 Hz to 3 kHz: the reference's propagator overflows or loses its top
    normalisation from ~300 Hz (summed decay ~700), the library's mt1d_anl
    cuts the column where the fields have decayed by e^-40 and matches the
    impedance recursion above (E_x with sigma_xx, E_y with sigma_yy for
    diagonal anisotropy) to 1e-13 of the top value at every node."""
    rng = np.random.default_rng(3)
    nz = 30
    c = np.concatenate([rng.uniform(20, 800, 20), 2000 * 1.4 ** np.arange(10)])
    conv = np.zeros((nz, 6))
    conv[:, 0] = 10 ** rng.uniform(-3, -1, nz)
    conv[:, 1] = conv[:, 0] * 10 ** rng.uniform(-0.5, 0.5, nz)
    conv[:, 2] = conv[:, 0]
    for f in (0.1, 3.0, 30.0, 300.0, 3000.0):
        e, h, imp = sfdx.mt1d_anl(f, conv, c)
        assert np.all(np.isfinite(e))
        assert np.abs(e[:nz + 1] - _te_profile(conv[:, 0], c, f)).max() <= 1e-13
        assert np.abs(e[nz + 1:] - _te_profile(conv[:, 1], c, f)).max() <= 1e-13


@pytest.mark.parametrize("rd,out_type", [(1, 1), (2, 1), (2, 2)])
def test_post_process_matches_oracle(rd, out_type):
    m = synth.make_block_halfspace_32()
    rng = np.random.default_rng(0)
    ex = rng.standard_normal((2 * m.nl, 2)) + 1j * rng.standard_normal((2 * m.nl, 2))
    freqs = np.array([1.0, 0.1])
    r1, t1 = sfdx.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=out_type, rd=rd)
    r2, t2 = O.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=out_type, rd=rd)
    assert np.allclose(r1, r2, rtol=1e-11, atol=1e-12 * abs(r2).max())
    assert np.allclose(t1, t2, rtol=1e-11, atol=1e-12 * abs(t2).max())


def test_direct_solve_matches_superlu():
    """PARDISO (complex symmetric) on the dual-volume scaled system against
    SuperLU on the reference's unscaled CSR; pinned by the interior residual."""
    m = synth.make_block_halfspace_32()
    a1, b1, c1, s1 = O.mesh_coarsen(m.a, m.b, m.c, m.sigma, 2)
    try:
        u = sfdx.direct_solve(a1, b1, c1, s1, 1.0)
    except sfdx.SfdxError as e:
        pytest.skip(f"MKL PARDISO unavailable: {e}")
    S = O.build_system(a1, b1, c1, s1, 1.0)
    nl = len(S.p1)
    for pol, rhs in ((0, S.p1), (1, S.p2)):
        x = u[pol * nl:(pol + 1) * nl]
        r = rhs - S.A @ x
        assert np.linalg.norm(r[~S.bnd]) <= 1e-10 * np.linalg.norm(rhs[~S.bnd])
        assert np.allclose(x[S.bnd], rhs[S.bnd])


@pytest.mark.parametrize("rd", [1, 2])
def test_misfit_field_gradient_matches_directional_difference(rd):
    """sfdx_misfit's field gradient w (the transposed post-process) against a
    central difference of the misfit along a random complex direction; the
    misfit itself against the oracle's."""
    m = synth.make_iso_block_model("mf", 16, 16, 12, 4, seed=31, nblocks=3, freq=1.0, level=3, nstations=6)
    freqs = np.array([0.5, 5.0])
    cols = []
    for f in freqs:
        S = O.build_system(m.a, m.b, m.c, m.sigma, f)
        x1, x2 = O.direct_solve(S)
        cols.append(np.concatenate([x1, x2]))
    ex = np.stack(cols, axis=1)
    res, _ = O.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=2, rd=rd)
    obs = res.copy()
    obs[:, :2] *= 1.2
    obs[:, 2:] += 3.0
    phi, w = sfdx.misfit(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, obs, rd=rd, with_grad=True)
    assert phi == pytest.approx(O.misfit(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, obs, rd=rd), rel=1e-12)
    rng = np.random.default_rng(7)
    dirn = rng.standard_normal(ex.shape) + 1j * rng.standard_normal(ex.shape)
    h = 1e-6 * np.abs(ex).max() / np.abs(dirn).max()
    fp = sfdx.misfit(m.a, m.b, m.c, m.nair, freqs, m.recv, ex + h * dirn, obs, rd=rd)
    fm = sfdx.misfit(m.a, m.b, m.c, m.nair, freqs, m.recv, ex - h * dirn, obs, rd=rd)
    fd = (fp - fm) / (2 * h)
    an = float(np.real(np.sum(w * dirn)))
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_post_process_topography_matches_oracle():
    """sfdx_post_process_topo / sfdx_surface_planes (the topography branch,
    fdfe_main.f90:604-614, 656-666) against the oracle on the carved
    topography generator at 32x32x24: planes from the carved sigma, then
    impedances and tippers on random fields, both station samplings."""
    m = synth.make_aniso_topography_128(nx=32, nz_earth=20, nair=4, nstations=8)
    p1 = sfdx.surface_planes(m.a, m.b, m.c, m.sigma, 1e-6)
    p2 = O.surface_planes(m.a, m.b, m.c, m.sigma, 1e-6)
    assert np.array_equal(p1, p2) and p1.max() > m.nair
    rng = np.random.default_rng(8)
    ex = rng.standard_normal((2 * m.nl, 2)) + 1j * rng.standard_normal((2 * m.nl, 2))
    freqs = np.array([0.1, 2.0])
    for rd in (1, 2):
        r1, t1 = sfdx.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=1, rd=rd, surf=p1)
        r2, t2 = O.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=1, rd=rd, surf=p2)
        assert np.allclose(r1, r2, rtol=1e-11, atol=1e-12 * abs(r2).max())
        assert np.allclose(t1, t2, rtol=1e-11, atol=1e-12 * abs(t2).max())
        # the flat datum differs on these fields
        r0, _ = O.post_process(m.a, m.b, m.c, m.nair, freqs, m.recv, ex, out_type=1, rd=rd)
        assert not np.allclose(r0, r2, rtol=1e-3)


def test_post_process_topography_rejects_bad_planes():
    m = synth.make_iso_block_model("bp", 16, 16, 12, 4, seed=1, nblocks=1, freq=1.0, level=3, nstations=4)
    ex = np.zeros((2 * m.nl, 1), complex)
    surf = np.full((m.nx, m.ny), m.nz, np.int32)
    with pytest.raises(sfdx.SfdxError):
        sfdx.post_process(m.a, m.b, m.c, m.nair, m.freqs, m.recv, ex, surf=surf)
    air = m.sigma.copy()
    air[:, :3] = 1e-8
    air[:, 3:] = 0
    with pytest.raises(sfdx.SfdxError):
        sfdx.surface_planes(m.a, m.b, m.c, air, 1e-6)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_mesh_coarsen_matches_oracle(k):
    """sfdx_mesh_coarsen (the product's coarsening, mesh_coarsen at
    fdfe_main.f90:18-84 with the index fixed) against the oracle on a
    full-tensor model with non-uniform widths."""
    m = synth.make_aniso_blocks_128(nx=32, nz_earth=20, nair=4, nstations=8)
    a1, b1, c1, s1 = sfdx.mesh_coarsen(m.a, m.b, m.c, m.sigma, k)
    a2, b2, c2, s2 = O.mesh_coarsen(m.a, m.b, m.c, m.sigma, k)
    assert np.allclose(a1, a2, rtol=1e-15) and np.allclose(b1, b2, rtol=1e-15) and np.allclose(c1, c2, rtol=1e-15)
    assert s1.shape == s2.shape
    assert np.allclose(s1, s2, rtol=1e-13, atol=1e-15 * np.abs(s2).max())


def test_direct_solve_is_independent_of_mkl_threads():
    """The coarse direct solve runs on a fixed number of MKL threads, so the
    caller's MKL_NUM_THREADS cannot change its result bit pattern (the
    factorisation's summation order depends on the thread count)."""
    import subprocess
    import sys
    import tempfile
    script = (
        "import sys, numpy as np\n"
        f"sys.path[:0] = [{ROOT!r}, {os.path.join(ROOT, 'oracle')!r}]\n"
        "import oracle as O\n"
        "from sfd_excmg_amd import sfdx, synth\n"
        "m = synth.make_aniso_blocks_128(nx=64, nz_earth=40, nair=8, nstations=8)\n"
        "a, b, c, s = O.mesh_coarsen(m.a, m.b, m.c, m.sigma, 2)\n"
        "np.save(sys.argv[1], sfdx.direct_solve(a, b, c, s, 0.1))\n")
    outs = []
    with tempfile.TemporaryDirectory() as td:
        for nt in ("1", "3"):
            path = os.path.join(td, f"u{nt}.npy")
            env = dict(os.environ, MKL_NUM_THREADS=nt, OMP_NUM_THREADS=nt)
            r = subprocess.run([sys.executable, "-c", script, path], env=env, capture_output=True, text=True)
            if r.returncode != 0 and "PARDISO" in r.stderr:
                pytest.skip("MKL PARDISO unavailable")
            assert r.returncode == 0, r.stderr
            outs.append(np.load(path))
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
