// sfdx_stencil.h -- grid indexing and the 13-point edge stencil, shared by the
// gfx950 kernels and the host-side coarse-level assembly.
//
// Operator.  The reference assembles, for every interior edge, a 13-entry row of
// the staggered-grid curl-curl operator A = C^T C - i*w*mu*S_sigma
// (sfd_excmg/src/fdfe_main.f90:226-396) and leaves it unscaled (line 292).
// Multiplying each row by its dual volume ve (lines 252, 301, 350) gives the
// complex-symmetric form used here:
//
//   A_s = K - kappa * M,   kappa = i*omega*mu0,
//
// K: curl-curl with face weights; for Ex(i,j,k) the same-orientation couplings
// are -a_i(b_{j-1}+b_j)/(2c_{k-1}) etc. and the cross couplings +-(c_{k-1}+c_k)/2
// (to Ey) and +-(b_{j-1}+b_j)/2 (to Ez).
// M: corner-lumped anisotropic edge mass; diagonal sum_cells sigma_dd V/4 over
// the 4 cells sharing the edge, off-diagonal sum sigma_{d1 d2} V/8 over the 2
// cells containing both edges (the reference's sigmav terms times ve).
//
// Layout (0-based, identical to the reference's 1-based edge numbering):
//   Ex(i,j,k) = (k*(ny+1) + j)*nx + i                 i<nx,  j<=ny, k<=nz
//   Ey(i,j,k) = XN + (k*ny + j)*(nx+1) + i            i<=nx, j<ny,  k<=nz
//   Ez(i,j,k) = XN + YN + (k*(ny+1) + j)*(nx+1) + i   i<=nx, j<=ny, k<nz
//   cell(i,j,k) = (k*ny + j)*nx + i ; node(i,j,k) = (k*(ny+1) + j)*(nx+1) + i
//   sigma[comp*ncell + cell], comp = xx, yy, zz, xy, xz, yz
// Dirichlet (boundary) edges: sfd_excmg/src/fdfe_main.f90:436-468.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SFDX_HD __host__ __device__ __forceinline__
#else
#define SFDX_HD inline
#endif

namespace sfdx {

struct Dims {
    int nx, ny, nz;
    int64_t XN, YN, ZN, NL, NC, NP;
};

SFDX_HD Dims make_dims(int nx, int ny, int nz)
{
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.XN = (int64_t)nx * (ny + 1) * (nz + 1);
    d.YN = (int64_t)(nx + 1) * ny * (nz + 1);
    d.ZN = (int64_t)(nx + 1) * (ny + 1) * nz;
    d.NL = d.XN + d.YN + d.ZN;
    d.NC = (int64_t)nx * ny * nz;
    d.NP = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    return d;
}

SFDX_HD int64_t ex_id(const Dims &d, int i, int j, int k) { return ((int64_t)k * (d.ny + 1) + j) * d.nx + i; }
SFDX_HD int64_t ey_id(const Dims &d, int i, int j, int k) { return d.XN + ((int64_t)k * d.ny + j) * (d.nx + 1) + i; }
SFDX_HD int64_t ez_id(const Dims &d, int i, int j, int k) { return d.XN + d.YN + ((int64_t)k * (d.ny + 1) + j) * (d.nx + 1) + i; }
SFDX_HD int64_t cell_id(const Dims &d, int i, int j, int k) { return ((int64_t)k * d.ny + j) * d.nx + i; }
SFDX_HD int64_t node_id(const Dims &d, int i, int j, int k) { return ((int64_t)k * (d.ny + 1) + j) * (d.nx + 1) + i; }

// orientation (0,1,2) and (i,j,k) of edge e
SFDX_HD int edge_decode(const Dims &d, int64_t e, int &i, int &j, int &k)
{
    if (e < d.XN) {
        i = (int)(e % d.nx); int64_t r = e / d.nx;
        j = (int)(r % (d.ny + 1)); k = (int)(r / (d.ny + 1));
        return 0;
    }
    if (e < d.XN + d.YN) {
        int64_t q = e - d.XN;
        i = (int)(q % (d.nx + 1)); int64_t r = q / (d.nx + 1);
        j = (int)(r % d.ny); k = (int)(r / d.ny);
        return 1;
    }
    int64_t q = e - d.XN - d.YN;
    i = (int)(q % (d.nx + 1)); int64_t r = q / (d.nx + 1);
    j = (int)(r % (d.ny + 1)); k = (int)(r / (d.ny + 1));
    return 2;
}

SFDX_HD bool edge_interior(const Dims &d, int o, int i, int j, int k)
{
    if (o == 0) return j > 0 && j < d.ny && k > 0 && k < d.nz;
    if (o == 1) return i > 0 && i < d.nx && k > 0 && k < d.nz;
    return i > 0 && i < d.nx && j > 0 && j < d.ny;
}

// SSOR colour of an edge: orientation x parity.  No edge couples to a
// same-orientation neighbour of equal parity, so each colour is independent.
SFDX_HD int edge_colour(int o, int i, int j, int k)
{
    int p = (o == 0) ? ((j + k) & 1) : (o == 1) ? ((i + k) & 1) : ((i + j) & 1);
    return 2 * o + p;
}

// One stencil row.  nb[0..3]: same-orientation neighbours, nb[4..7]: cross
// neighbours of the lower-numbered other orientation, nb[8..11]: of the
// higher-numbered one (for Ex: Ey then Ez; Ey: Ex then Ez; Ez: Ex then Ey).
// Coefficient of neighbour s: kg[s] - kappa*km[s]; diagonal kgd - kappa*kmd.
template <typename R>
struct Row {
    int64_t nb[12];
    R kg[12];
    R km[12];
    R kgd, kmd;
};

// sigma accessor: sig[comp*NC + cell]
// With RCP the reciprocal widths ra, rb, rc (1/a, 1/b, 1/c) replace the four
// divisions of a row by multiplications (device kernels); without, the widths
// are divided (host assembly).  With XM false the cross-orientation mass
// couplings of anisotropic cells (the sigma_xy, sigma_xz, sigma_yz terms) are
// left out (km[0..11] = 0) and only the row's own diagonal conductivity is
// read: the lumped-mass approximation some preconditioners use.
template <typename R, bool RCP = false, bool XM = true>
SFDX_HD void row_coeffs(const Dims &d, const R *__restrict__ a, const R *__restrict__ b,
                        const R *__restrict__ c, const R *__restrict__ sig, int o, int i, int j, int k,
                        Row<R> &r, const R *__restrict__ ra = nullptr, const R *__restrict__ rb = nullptr,
                        const R *__restrict__ rc = nullptr)
{
    auto ia = [&](int q) { return RCP ? ra[q] : (R)1 / a[q]; };
    auto ib = [&](int q) { return RCP ? rb[q] : (R)1 / b[q]; };
    auto ic = [&](int q) { return RCP ? rc[q] : (R)1 / c[q]; };
    const int64_t NC = d.NC;
    if (o == 0) {
        const R ai = a[i], bm = b[j - 1], bp = b[j], cm = c[k - 1], cp = c[k];
        const int64_t C00 = cell_id(d, i, j - 1, k - 1), C10 = cell_id(d, i, j, k - 1);
        const int64_t C01 = cell_id(d, i, j - 1, k), C11 = cell_id(d, i, j, k);
        const R hb = (R)0.5 * ai * (bm + bp), hc = (R)0.5 * ai * (cm + cp);
        const R wkm = hb * ic(k - 1), wkp = hb * ic(k);
        const R wjm = hc * ib(j - 1), wjp = hc * ib(j);
        const R hy = (cm + cp) / 2, hz = (bm + bp) / 2;
        r.nb[0] = ex_id(d, i, j, k - 1); r.kg[0] = -wkm;
        r.nb[1] = ex_id(d, i, j - 1, k); r.kg[1] = -wjm;
        r.nb[2] = ex_id(d, i, j + 1, k); r.kg[2] = -wjp;
        r.nb[3] = ex_id(d, i, j, k + 1); r.kg[3] = -wkp;
        r.kgd = wkm + wkp + wjm + wjp;
        const R *sxx = sig, *sxy = sig + 3 * NC, *sxz = sig + 4 * NC;
        r.kmd = ai * (sxx[C00] * bm * cm + sxx[C10] * bp * cm + sxx[C01] * bm * cp + sxx[C11] * bp * cp) / 4;
        for (int s = 0; s < 4; s++) r.km[s] = 0;
        const R my1 = XM ? ai * (sxy[C00] * bm * cm + sxy[C01] * bm * cp) / 8 : (R)0;
        const R my2 = XM ? ai * (sxy[C10] * bp * cm + sxy[C11] * bp * cp) / 8 : (R)0;
        const R mz1 = XM ? ai * (sxz[C00] * bm * cm + sxz[C10] * bp * cm) / 8 : (R)0;
        const R mz2 = XM ? ai * (sxz[C01] * bm * cp + sxz[C11] * bp * cp) / 8 : (R)0;
        r.nb[4] = ey_id(d, i, j - 1, k);     r.kg[4] = hy;  r.km[4] = my1;
        r.nb[5] = ey_id(d, i + 1, j - 1, k); r.kg[5] = -hy; r.km[5] = my1;
        r.nb[6] = ey_id(d, i, j, k);         r.kg[6] = -hy; r.km[6] = my2;
        r.nb[7] = ey_id(d, i + 1, j, k);     r.kg[7] = hy;  r.km[7] = my2;
        r.nb[8] = ez_id(d, i, j, k - 1);     r.kg[8] = hz;  r.km[8] = mz1;
        r.nb[9] = ez_id(d, i + 1, j, k - 1); r.kg[9] = -hz; r.km[9] = mz1;
        r.nb[10] = ez_id(d, i, j, k);        r.kg[10] = -hz; r.km[10] = mz2;
        r.nb[11] = ez_id(d, i + 1, j, k);    r.kg[11] = hz;  r.km[11] = mz2;
    } else if (o == 1) {
        const R bj = b[j], am = a[i - 1], ap = a[i], cm = c[k - 1], cp = c[k];
        const int64_t D00 = cell_id(d, i - 1, j, k - 1), D10 = cell_id(d, i, j, k - 1);
        const int64_t D01 = cell_id(d, i - 1, j, k), D11 = cell_id(d, i, j, k);
        const R ha = (R)0.5 * bj * (am + ap), hc = (R)0.5 * bj * (cm + cp);
        const R wkm = ha * ic(k - 1), wkp = ha * ic(k);
        const R wim = hc * ia(i - 1), wip = hc * ia(i);
        const R hx = (cm + cp) / 2, hz = (am + ap) / 2;
        r.nb[0] = ey_id(d, i, j, k - 1); r.kg[0] = -wkm;
        r.nb[1] = ey_id(d, i - 1, j, k); r.kg[1] = -wim;
        r.nb[2] = ey_id(d, i + 1, j, k); r.kg[2] = -wip;
        r.nb[3] = ey_id(d, i, j, k + 1); r.kg[3] = -wkp;
        r.kgd = wkm + wkp + wim + wip;
        const R *syy = sig + NC, *sxy = sig + 3 * NC, *syz = sig + 5 * NC;
        r.kmd = bj * (syy[D00] * am * cm + syy[D10] * ap * cm + syy[D01] * am * cp + syy[D11] * ap * cp) / 4;
        for (int s = 0; s < 4; s++) r.km[s] = 0;
        const R mx1 = XM ? bj * (sxy[D00] * am * cm + sxy[D01] * am * cp) / 8 : (R)0;
        const R mx2 = XM ? bj * (sxy[D10] * ap * cm + sxy[D11] * ap * cp) / 8 : (R)0;
        const R mz1 = XM ? bj * (syz[D00] * am * cm + syz[D10] * ap * cm) / 8 : (R)0;
        const R mz2 = XM ? bj * (syz[D01] * am * cp + syz[D11] * ap * cp) / 8 : (R)0;
        r.nb[4] = ex_id(d, i - 1, j, k);     r.kg[4] = hx;  r.km[4] = mx1;
        r.nb[5] = ex_id(d, i, j, k);         r.kg[5] = -hx; r.km[5] = mx2;
        r.nb[6] = ex_id(d, i - 1, j + 1, k); r.kg[6] = -hx; r.km[6] = mx1;
        r.nb[7] = ex_id(d, i, j + 1, k);     r.kg[7] = hx;  r.km[7] = mx2;
        r.nb[8] = ez_id(d, i, j, k - 1);     r.kg[8] = hz;  r.km[8] = mz1;
        r.nb[9] = ez_id(d, i, j + 1, k - 1); r.kg[9] = -hz; r.km[9] = mz1;
        r.nb[10] = ez_id(d, i, j, k);        r.kg[10] = -hz; r.km[10] = mz2;
        r.nb[11] = ez_id(d, i, j + 1, k);    r.kg[11] = hz;  r.km[11] = mz2;
    } else {
        const R ck = c[k], am = a[i - 1], ap = a[i], bm = b[j - 1], bp = b[j];
        const int64_t F00 = cell_id(d, i - 1, j - 1, k), F01 = cell_id(d, i - 1, j, k);
        const int64_t F10 = cell_id(d, i, j - 1, k), F11 = cell_id(d, i, j, k);
        const R ha = (R)0.5 * ck * (am + ap), hb = (R)0.5 * ck * (bm + bp);
        const R wjm = ha * ib(j - 1), wjp = ha * ib(j);
        const R wim = hb * ia(i - 1), wip = hb * ia(i);
        const R hx = (bm + bp) / 2, hy = (am + ap) / 2;
        r.nb[0] = ez_id(d, i, j - 1, k); r.kg[0] = -wjm;
        r.nb[1] = ez_id(d, i - 1, j, k); r.kg[1] = -wim;
        r.nb[2] = ez_id(d, i + 1, j, k); r.kg[2] = -wip;
        r.nb[3] = ez_id(d, i, j + 1, k); r.kg[3] = -wjp;
        r.kgd = wjm + wjp + wim + wip;
        const R *szz = sig + 2 * NC, *sxz = sig + 4 * NC, *syz = sig + 5 * NC;
        r.kmd = ck * (szz[F00] * bm * am + szz[F01] * bp * am + szz[F10] * bm * ap + szz[F11] * bp * ap) / 4;
        for (int s = 0; s < 4; s++) r.km[s] = 0;
        const R mx1 = XM ? ck * (sxz[F00] * bm * am + sxz[F01] * bp * am) / 8 : (R)0;
        const R mx2 = XM ? ck * (sxz[F10] * bm * ap + sxz[F11] * bp * ap) / 8 : (R)0;
        const R my1 = XM ? ck * (syz[F00] * bm * am + syz[F10] * bm * ap) / 8 : (R)0;
        const R my2 = XM ? ck * (syz[F01] * bp * am + syz[F11] * bp * ap) / 8 : (R)0;
        r.nb[4] = ex_id(d, i - 1, j, k);     r.kg[4] = hx;  r.km[4] = mx1;
        r.nb[5] = ex_id(d, i, j, k);         r.kg[5] = -hx; r.km[5] = mx2;
        r.nb[6] = ex_id(d, i - 1, j, k + 1); r.kg[6] = -hx; r.km[6] = mx1;
        r.nb[7] = ex_id(d, i, j, k + 1);     r.kg[7] = hx;  r.km[7] = mx2;
        r.nb[8] = ey_id(d, i, j - 1, k);     r.kg[8] = hy;  r.km[8] = my1;
        r.nb[9] = ey_id(d, i, j, k);         r.kg[9] = -hy; r.km[9] = my2;
        r.nb[10] = ey_id(d, i, j - 1, k + 1); r.kg[10] = -hy; r.km[10] = my1;
        r.nb[11] = ey_id(d, i, j, k + 1);    r.kg[11] = hy;  r.km[11] = my2;
    }
}

// Dual volume ve of an interior edge (the reference's row scale).
template <typename R>
SFDX_HD R dual_volume(const R *a, const R *b, const R *c, int o, int i, int j, int k)
{
    if (o == 0) return a[i] * (b[j - 1] + b[j]) * (c[k - 1] + c[k]) / 4;
    if (o == 1) return b[j] * (a[i - 1] + a[i]) * (c[k - 1] + c[k]) / 4;
    return c[k] * (b[j - 1] + b[j]) * (a[i - 1] + a[i]) / 4;
}

// Edge length and end nodes (sfd_excmg/src/div_correc.f90:831-855).
template <typename R>
SFDX_HD R edge_len_nodes(const Dims &d, const R *a, const R *b, const R *c, int o, int i, int j, int k,
                         int64_t &n0, int64_t &n1)
{
    n0 = node_id(d, i, j, k);
    if (o == 0) { n1 = n0 + 1; return a[i]; }
    if (o == 1) { n1 = n0 + (d.nx + 1); return b[j]; }
    n1 = n0 + (int64_t)(d.nx + 1) * (d.ny + 1);
    return c[k];
}

SFDX_HD bool node_boundary(const Dims &d, int i, int j, int k)
{
    return i == 0 || i == d.nx || j == 0 || j == d.ny || k == 0 || k == d.nz;
}

// Dirichlet value of boundary edge (o,i,j,k) for polarisation pol (0 = xy,
// 1 = yx) from the 1-D field e_dc[2*nz+2] (Ex at z-nodes, then Ey), following
// dcfast (sfd_excmg/src/fdfe_main.f90:470-494).  Returns false for 0.
// With full (full-tensor boundary, a deliberate deviation: the reference sets
// the cross-polar components to 0), e_dc holds 4*nz+4 values and the
// boundary y-edges of xy and x-edges of yx take the cross-polar components
// e_dc[2nz+2+k] (Ey of the xy mode) and e_dc[3nz+3+k] (Ex of the yx mode),
// 0 on the top plane where the modes are normalised to E = (1,0), (0,1).
template <typename C>
SFDX_HD bool dirichlet_value(const Dims &d, const C *e_dc, int pol, int o, int i, int j, int k, C &v,
                             bool full = false)
{
    if (pol == 0 && o == 0) {
        if (k == 0) { v = C(1.0); return true; }
        if (k == d.nz) { v = e_dc[d.nz]; return true; }
        if (j == 0 || j == d.ny) { v = e_dc[k]; return true; }
    }
    if (pol == 1 && o == 1) {
        if (k == 0) { v = C(1.0); return true; }
        if (k == d.nz) { v = e_dc[2 * d.nz + 1]; return true; }
        if (i == 0 || i == d.nx) { v = e_dc[d.nz + 1 + k]; return true; }
    }
    if (full) {
        const int n1 = d.nz + 1;
        if (pol == 0 && o == 1 && (k == 0 || k == d.nz || i == 0 || i == d.nx)) {
            v = (k == 0) ? C(0.0) : e_dc[2 * n1 + k];
            return true;
        }
        if (pol == 1 && o == 0 && (k == 0 || k == d.nz || j == 0 || j == d.ny)) {
            v = (k == 0) ? C(0.0) : e_dc[3 * n1 + k];
            return true;
        }
    }
    return false;
}

}  // namespace sfdx
