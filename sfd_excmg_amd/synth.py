"""Synthetic MT models and the reference's on-disk formats.

* ``Model`` holds what ``read_input`` returns in the reference
  (``sfd_excmg/src/file_opt.f90:8-95``): cell widths a(nx), b(ny), c(nz+nair)
  in metres with the air on top, nair, sigma(ncell, 6) in S/m (columns xx, yy,
  zz, xy, xz, yz; cell index (k-1)*nx*ny + (j-1)*nx + i, k=1 at the top of the
  air), frequencies and stations (x, y, z).
* ``write_fwd`` / ``read_fwd`` / ``write_dat`` / ``read_dat`` follow the text
  formats of ``file_opt.f90:8-95`` and ``input.py:6-118``.
* ``make_block_halfspace_32`` ... ``make_random_field_96`` build the five benchmark workloads with fixed
  seeds (numpy ``default_rng``); the air thicknesses mirror the deepest earth
  layers as ``file_opt.f90:52`` does.
* ``euler_tensor`` restates the Euler-rotated conductivity tensor of
  ``input.py:123-140`` (R = Rz(strike) Rx(dip) Rz(slant), sigma = R diag R^T).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SIGMA_AIR_DEFAULT = 1e-8


@dataclass
class Model:
    name: str
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray               # nz_total = nz_earth + nair entries, air first
    nair: int
    sigma: np.ndarray           # (ncell, 6)
    freqs: np.ndarray
    recv: np.ndarray            # (nrec, 3)
    level: int = 3
    eps: float = 1e-6
    solver: int = 1             # 1 BiCGStab, 3 GPBiCG, 5 COCR
    out_type: int = 1
    corr_intv: int = 200
    meta: dict = field(default_factory=dict)

    @property
    def nx(self):
        return len(self.a)

    @property
    def ny(self):
        return len(self.b)

    @property
    def nz(self):
        return len(self.c)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz

    @property
    def nl(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz

    def is_isotropic(self):
        s = self.sigma
        return bool(np.all(s[:, 3:] == 0) and np.all(s[:, 0] == s[:, 1]) and np.all(s[:, 0] == s[:, 2]))


# ----------------------------------------------------------------- helpers
def euler_tensor(sx, sy, sz, strike, dip, slant):
    """input.py:123-140 -- angles in radians."""
    rs = np.array([[math.cos(strike), -math.sin(strike), 0], [math.sin(strike), math.cos(strike), 0], [0, 0, 1]])
    rd = np.array([[1, 0, 0], [0, math.cos(dip), -math.sin(dip)], [0, math.sin(dip), math.cos(dip)]])
    rl = np.array([[math.cos(slant), -math.sin(slant), 0], [math.sin(slant), math.cos(slant), 0], [0, 0, 1]])
    R = rs @ rd @ rl
    return R @ np.diag([sx, sy, sz]) @ R.T


def tensor_to_cols(S):
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


def padded_axis(ncore, h, npad, growth):
    core = np.full(ncore, float(h))
    pad = h * growth ** np.arange(1, npad + 1)
    return np.concatenate([pad[::-1], core, pad])


def earth_layers(ncore, h, ngrow, growth):
    return np.concatenate([np.full(ncore, float(h)), h * growth ** np.arange(1, ngrow + 1)])


def with_air(c_earth, nair):
    """c(1:nair) = c(nz+nair:nz+1:-1) (file_opt.f90:52)."""
    c_earth = np.asarray(c_earth, np.float64)
    return np.concatenate([c_earth[::-1][:nair], c_earth])


def station_grid(n, half_extent):
    xs = np.linspace(-half_extent, half_extent, n)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1)


def cell_centres(w):
    e = np.concatenate([[0.0], np.cumsum(w)])
    return (e[:-1] + e[1:]) / 2 - e[-1] / 2


def _sigma_iso(ncell, s):
    sig = np.zeros((ncell, 6))
    sig[:, :3] = s
    return sig


def set_air(sig, nx, ny, nair, sigma_air):
    """sigma(1:nair*nx*ny,1:3) = sigma_air, off-diagonals 0 (file_opt.f90:75)."""
    sig[: nair * nx * ny, :3] = sigma_air
    sig[: nair * nx * ny, 3:] = 0


# ----------------------------------------------------------------- configs
def make_block_halfspace_32(seed: int = 1, sigma_air: float = SIGMA_AIR_DEFAULT) -> Model:
    """32^3 cells, 3 levels: 100 ohm-m half-space, one 10 ohm-m 8x8x6 block at a
    seeded position under the surface, 1 Hz, BiCGStab+SSOR to 1e-6, 10x10
    stations."""
    rng = np.random.default_rng(seed)
    a = padded_axis(20, 100.0, 6, 1.3)
    b = a.copy()
    nair = 8
    c = with_air(earth_layers(8, 50.0, 16, 1.3), nair)
    nx, ny, nz = len(a), len(b), len(c)
    sig = _sigma_iso(nx * ny * nz, 1.0 / 100.0)
    i0 = int(rng.integers(6, 6 + 20 - 8 + 1))
    j0 = int(rng.integers(6, 6 + 20 - 8 + 1))
    k0 = int(rng.integers(nair, nair + 12 - 6 + 1))
    for k in range(k0, k0 + 6):
        for j in range(j0, j0 + 8):
            sig[k * nx * ny + j * nx + i0: k * nx * ny + j * nx + i0 + 8, :3] = 1.0 / 10.0
    set_air(sig, nx, ny, nair, sigma_air)
    return Model("block_halfspace_32", a, b, c, nair, sig, np.array([1.0]), station_grid(10, 750.0), level=3, eps=1e-6,
                 solver=1, corr_intv=200, meta={"block": (i0, j0, k0)})


def _anisotropic_blocks(rng, sig, nx, ny, nair, nz_earth, core_lo, core_n, nblocks):
    blocks = []
    for _ in range(nblocks):
        sx, sy, sz = (int(rng.integers(4, 25)), int(rng.integers(4, 25)), int(rng.integers(4, 17)))
        i0 = int(rng.integers(core_lo, core_lo + core_n - sx + 1))
        j0 = int(rng.integers(core_lo, core_lo + core_n - sy + 1))
        k0 = int(rng.integers(nair, nair + min(48, nz_earth) - sz + 1))
        rho = 10.0 ** rng.uniform(0.0, 3.0, size=3)
        ang = np.deg2rad(rng.uniform(0.0, 180.0, size=3))
        cols = tensor_to_cols(euler_tensor(1 / rho[0], 1 / rho[1], 1 / rho[2], *ang))
        for k in range(k0, k0 + sz):
            for j in range(j0, j0 + sy):
                base = k * nx * ny + j * nx
                sig[base + i0: base + i0 + sx, :] = cols
        blocks.append(dict(ijk=(i0, j0, k0), size=(sx, sy, sz), rho=rho.tolist(), angles_deg=np.rad2deg(ang).tolist()))
    return blocks


def make_aniso_blocks_128(seed: int = 2, sigma_air: float = SIGMA_AIR_DEFAULT, nx: int = 128, nz_earth: int = 80,
            nair: int = 16, nstations: int = 40) -> Model:
    """128x128x96 cells (16 air + 80 earth), full anisotropic tensor: 16 seeded
    blocks with principal resistivities log-uniform in [1,1000] ohm-m and Euler
    angles uniform in [0,180) deg over 100 ohm-m; 0.1 Hz, 4 levels, COCR+SSOR to
    1e-8, DC every 50 iterations, 40x40 stations."""
    rng = np.random.default_rng(seed)
    npad = nx // 8
    ncore = nx - 2 * npad
    a = padded_axis(ncore, 100.0, npad, 1.3)
    b = a.copy()
    c = with_air(earth_layers(nz_earth * 3 // 5, 50.0, nz_earth - nz_earth * 3 // 5, 1.15), nair)
    ny, nz = nx, len(c)
    sig = _sigma_iso(nx * ny * nz, 1.0 / 100.0)
    blocks = _anisotropic_blocks(rng, sig, nx, ny, nair, nz_earth, npad, ncore, 16)
    set_air(sig, nx, ny, nair, sigma_air)
    half = ncore * 100.0 / 2 - 2 * 100.0
    return Model("aniso_blocks_128", a, b, c, nair, sig, np.array([0.1]), station_grid(nstations, half), level=4, eps=1e-8,
                 solver=5, corr_intv=50, meta={"blocks": blocks})


def topography_depth(rng, x, y, amplitude=500.0):
    """Sum of 4 seeded sinusoids with total amplitude `amplitude` and
    wavelengths 2-10 km; returns the depth of the ground surface below the
    flat datum (0 at the highest point)."""
    amps = rng.uniform(0.2, 1.0, 4)
    amps *= amplitude / amps.sum()
    lam = rng.uniform(2000.0, 10000.0, 4)
    th = rng.uniform(0, np.pi, 4)
    ph = rng.uniform(0, 2 * np.pi, 4)
    X, Y = np.meshgrid(x, y, indexing="ij")
    h = np.zeros_like(X)
    for k in range(4):
        h += amps[k] * np.sin(2 * np.pi * (X * np.cos(th[k]) + Y * np.sin(th[k])) / lam[k] + ph[k])
    return h.max() - h


def carve_topography(m: Model, elevation, sigma_air: float = SIGMA_AIR_DEFAULT):
    """Apply ground elevations elevation(nx, ny) (m above the flat datum) to
    m.sigma in place, as the driver's corrected .topo reader does: every cell
    whose centre lies above the ground becomes air (sigma_air diagonal,
    off-diagonals 0); every air cell whose centre lies below it takes the tensor
    of the column's first earth cell (reference: file_opt.f90:97-128)."""
    elevation = np.asarray(elevation, np.float64)
    nx, ny = m.nx, m.ny
    zc = np.cumsum(m.c) - m.c / 2 - m.c[: m.nair].sum()    # cell-centre depth below the datum
    sig = m.sigma.reshape(m.nz, ny, nx, 6)
    earth = sig[m.nair].copy()                              # (ny, nx, 6)
    ground = -elevation.T                                   # (ny, nx) depth of the ground
    for k in range(m.nz):
        air = zc[k] < ground
        sig[k][air, :3] = sigma_air
        sig[k][air, 3:] = 0
        if k < m.nair:
            raise_ = ~air
            sig[k][raise_] = earth[raise_]
    m.meta["topo"] = elevation.copy()
    return m


def make_aniso_topography_128(seed: int = 3, sigma_air: float = SIGMA_AIR_DEFAULT, **kw) -> Model:
    """The 128x128x96 anisotropic block model plus undulating topography: earth cells whose centre lies above the
    ground surface become air.  meta["topo"] holds the ground elevation (nx, ny) above the flat datum (<= 0 here:
    the highest ground point lies on the datum)."""
    m = make_aniso_blocks_128(sigma_air=sigma_air, **kw)
    rng = np.random.default_rng(seed)
    xc, yc = cell_centres(m.a), cell_centres(m.b)
    depth = topography_depth(rng, xc, yc)
    carve_topography(m, -depth, sigma_air)
    m.name = "aniso_topography_128"
    m.meta["topography_max_depth"] = float(depth.max())
    return m


def make_iso_block_model(name, nx, ny, nz_earth, nair, seed, nblocks=6, freq=1.0, level=3,
                         sigma_air=SIGMA_AIR_DEFAULT, nstations=10):
    """Isotropic model with seeded random blocks (resistivity log-uniform in
    [1,1000] ohm-m) over 100 ohm-m, used for the uneven batch models."""
    rng = np.random.default_rng(seed)
    npad = max(2, nx // 8)
    ncore = nx - 2 * npad
    a = padded_axis(ncore, 100.0, npad, 1.3)
    npady = max(2, ny // 8)
    b = padded_axis(ny - 2 * npady, 100.0, npady, 1.3)
    c = with_air(earth_layers(nz_earth * 3 // 5, 50.0, nz_earth - nz_earth * 3 // 5, 1.2), nair)
    nz = len(c)
    sig = _sigma_iso(nx * ny * nz, 1.0 / 100.0)
    for _ in range(nblocks):
        sx, sy, sz = (int(rng.integers(2, max(3, ncore // 3))), int(rng.integers(2, max(3, (ny - 2 * npady) // 3))),
                      int(rng.integers(2, max(3, nz_earth // 4))))
        i0 = int(rng.integers(npad, npad + ncore - sx + 1))
        j0 = int(rng.integers(npady, ny - npady - sy + 1))
        k0 = int(rng.integers(nair, nair + nz_earth // 2 - sz + 1))
        s = 1.0 / 10.0 ** rng.uniform(0, 3)
        for k in range(k0, k0 + sz):
            for j in range(j0, j0 + sy):
                sig[k * nx * ny + j * nx + i0: k * nx * ny + j * nx + i0 + sx, :3] = s
    set_air(sig, nx, ny, nair, sigma_air)
    half = min(ncore, ny - 2 * npady) * 100.0 / 2 - 200.0
    return Model(name, a, b, c, nair, sig, np.array([freq]), station_grid(nstations, half), level=level,
                 eps=1e-6, solver=1)


UNEVEN_ISO_SHAPES = [(48, 48, 48), (64, 64, 64), (64, 64, 48), (80, 80, 80), (96, 96, 64), (56, 56, 56), (72, 72, 72), (96, 96, 96)]


def max_levels(nx, ny, nz, nair):
    """Largest EXCMG level count L with every dimension (and nair) divisible by
    2^(L-1) and the level-L cubic blocks (8 cells) covering the grid."""
    L = 3
    while True:
        f = 2 ** L
        if nx % f or ny % f or nz % f or nair % (2 ** L):
            return L
        L += 1


def make_uneven_iso_models(seed: int = 4, sigma_air: float = SIGMA_AIR_DEFAULT):
    rng = np.random.default_rng(seed)
    out = []
    for idx, (nx, ny, nz) in enumerate(UNEVEN_ISO_SHAPES):
        nair = 8
        lv = 4 if (nx % 16 == 0 and ny % 16 == 0 and nz % 16 == 0) else 3
        m = make_iso_block_model(f"iso_{nx}x{ny}x{nz}", nx, ny, nz - nair, nair, int(rng.integers(1 << 30)),
                                 freq=1.0, level=lv, sigma_air=sigma_air)
        out.append(m)
    return out


def sounding_frequencies_16():
    return np.logspace(-3, 3, 16)


def gaussian_random_field(shape, corr_len, seed):
    """Stationary Gaussian field with Gaussian covariance of correlation length
    corr_len (in cells), unit variance, by spectral filtering."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    F = np.fft.fftn(w)
    ks = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    k2 = sum(k ** 2 for k in ks)
    filt = np.exp(-(np.pi ** 2) * corr_len ** 2 * k2)
    f = np.real(np.fft.ifftn(F * np.sqrt(filt)))
    return (f - f.mean()) / f.std()


def make_random_field_96(seed: int = 5, sigma_air: float = SIGMA_AIR_DEFAULT, nx=96, ny=96, nz_earth=56, nair=8,
            nstations=20) -> Model:
    """96x96x64 isotropic model, log10 rho = Gaussian random field (mean 2, sd
    0.5, correlation length 8 cells), 6 frequencies 1e-2..1e2 Hz, 20x20
    stations."""
    npad = nx // 8
    ncore = nx - 2 * npad
    a = padded_axis(ncore, 100.0, npad, 1.3)
    b = padded_axis(ny - 2 * npad, 100.0, npad, 1.3)
    c = with_air(earth_layers(nz_earth * 3 // 5, 50.0, nz_earth - nz_earth * 3 // 5, 1.2), nair)
    nz = len(c)
    g = gaussian_random_field((nz_earth, ny, nx), 8.0, seed)
    logrho = 2.0 + 0.5 * g
    sig = np.zeros((nx * ny * nz, 6))
    sig[nair * nx * ny:, :3] = (10.0 ** -logrho).reshape(-1)[:, None]
    set_air(sig, nx, ny, nair, sigma_air)
    half = ncore * 100.0 / 2 - 200.0
    return Model("random_field_96", a, b, c, nair, sig, np.logspace(-2, 2, 6), station_grid(nstations, half), level=4,
                 eps=1e-8, solver=5)


# ----------------------------------------------------------------- file IO
def write_fwd(m: Model, path_prefix: str, conduc_type: int = 3, topo_indic: int = 0):
    """<name>.fwd in the layout read by file_opt.f90:8-95."""
    nz_e = m.nz - m.nair
    with open(path_prefix + ".fwd", "w") as f:
        print(f"{m.name} written by sfd_excmg_amd.synth", file=f)
        print(f"{m.level} {m.eps:e} {m.solver} {m.out_type}", file=f)
        print(f"{m.nx} {m.ny} {nz_e} {m.nair} {topo_indic}", file=f)
        print(" ".join(f"{v:.10e}" for v in m.a), file=f)
        print(" ".join(f"{v:.10e}" for v in m.b), file=f)
        print(" ".join(f"{v:.10e}" for v in m.c[m.nair:]), file=f)
        print("1.3 1.15", file=f)
        print(f"1 {m.nx} 1 {m.ny} {m.nair + 1} {m.nz}", file=f)
        print(f"{conduc_type}", file=f)
        s = m.sigma[m.nair * m.nx * m.ny:]
        if conduc_type == 1:
            np.savetxt(f, s[:, :1], fmt="%.10e")
        elif conduc_type == 2:
            np.savetxt(f, s[:, :3], fmt="%.10e")
        else:
            np.savetxt(f, s, fmt="%.10e")


def write_dat(m: Model, path_prefix: str):
    """<name>.dat: nf, freqs, nrec, stations (file_opt.f90:80-93)."""
    with open(path_prefix + ".dat", "w") as f:
        print(len(m.freqs), file=f)
        for fr in m.freqs:
            print(f"{fr:.10e}", file=f)
        print(len(m.recv), file=f)
        for r in m.recv:
            print(f"{r[0]:.10e} {r[1]:.10e} {r[2]:.10e}", file=f)


def write_topo(m: Model, path_prefix: str):
    """<name>.topo for the driver's corrected reader: "1 n" then the n*n ground
    elevations of m.meta["topo"], i fastest (square grids)."""
    if m.nx != m.ny:
        raise ValueError("the .topo window is square")
    el = np.asarray(m.meta["topo"], np.float64)
    with open(path_prefix + ".topo", "w") as f:
        print(f"1 {m.nx}", file=f)
        for j in range(m.ny):
            for i in range(m.nx):
                print(f"{el[i, j]:.10e}", file=f)


def read_topo(path_prefix: str, nx: int, ny: int):
    """Ground elevations (nx, ny) from <name>.topo, 0 outside its window."""
    with open(path_prefix + ".topo") as f:
        toks = f.read().split()
    n1, n2 = int(toks[0]), int(toks[1])
    vals = np.array([float(t) for t in toks[2:2 + (n2 - n1 + 1) ** 2]]).reshape(n2 - n1 + 1, n2 - n1 + 1)
    el = np.zeros((nx, ny))
    el[n1 - 1:n2, n1 - 1:n2] = vals.T                      # vals[j, i]
    return el


def read_fwd_dat(path_prefix: str, sigma_air: float = SIGMA_AIR_DEFAULT) -> Model:
    """Parse <prefix>.fwd and <prefix>.dat like read_input (file_opt.f90:8-95)."""
    with open(path_prefix + ".fwd") as f:
        f.readline()
        toks = f.read().split()
    it = iter(toks)
    level, eps, solver, out_type = int(next(it)), float(next(it)), int(next(it)), int(next(it))
    nx, ny, nz, nair, topo = (int(next(it)) for _ in range(5))
    a = np.array([float(next(it)) for _ in range(nx)])
    b = np.array([float(next(it)) for _ in range(ny)])
    ce = np.array([float(next(it)) for _ in range(nz)])
    next(it); next(it)
    [int(next(it)) for _ in range(6)]
    ctype = int(next(it))
    ncol = {1: 1, 2: 3}.get(ctype, 6)
    vals = np.array([float(next(it)) for _ in range(nx * ny * nz * ncol)]).reshape(-1, ncol)
    c = with_air(ce, nair)
    sig = np.zeros((nx * ny * (nz + nair), 6))
    e = sig[nair * nx * ny:]
    if ctype == 1:
        e[:, :3] = vals
    elif ctype == 2:
        e[:, :3] = vals
    else:
        e[:, :] = vals
    set_air(sig, nx, ny, nair, sigma_air)
    with open(path_prefix + ".dat") as f:
        dt = f.read().split()
    jt = iter(dt)
    nf = int(next(jt))
    freqs = np.array([float(next(jt)) for _ in range(nf)])
    nrec = int(next(jt))
    recv = np.array([float(next(jt)) for _ in range(3 * nrec)]).reshape(nrec, 3)
    name = path_prefix.rstrip("/").split("/")[-1]
    m = Model(name, a, b, c, nair, sig, freqs, recv, level=level, eps=eps, solver=solver, out_type=out_type)
    if topo:
        carve_topography(m, read_topo(path_prefix, nx, ny), sigma_air)
    return m
