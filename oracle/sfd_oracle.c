/*
 * sfd_oracle.c -- FP64 CPU restatement of the SFD_EXCMG forward-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU library is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or as the timed CPU
 * baseline -- never as part of the product path (sfd_excmg_amd/), which has its
 * own independent implementation and fails loudly without its HIP library.
 *
 * Every routine names the reference file:line it restates (paths relative to
 * the reference repository, `src/` = `sfd_excmg/src/`).  Storage is 0-based
 * but the reference's 1-based layouts are kept exactly:
 *   cell (i,j,k)  ie   = (k*ny + j)*nx + i
 *   x-edge        (k*(ny+1) + j)*nx + i                     i<nx, j<=ny, k<=nz
 *   y-edge  xnl + (k*ny + j)*(nx+1) + i                     i<=nx, j<ny, k<=nz
 *   z-edge  xnl + ynl + (k*(ny+1) + j)*(nx+1) + i           i<=nx, j<=ny, k<nz
 *   node          (k*(ny+1) + j)*(nx+1) + i
 * sigma is column-major sigma(ncell,6): sigma[comp*ncell + ie], comp order
 * (xx, yy, zz, xy, xz, yz) (src/file_opt.f90:56-71).
 *
 * Deliberate deviations from the reference (each also fixed identically in the
 * GPU library):
 *   - SSOR apply returns +M^-1 y (reference returns -M^-1 y, src/solvers.f90:745).
 *   - mesh_coarsen uses the correct fine-cell index (src/fdfe_main.f90:64 is one
 *     layer and one row off).
 *   - divergence correction uses T = G^T M_sigma (M_sigma = the operator's own
 *     corner-lumped edge mass) and P = G^T M_sigma G.  For isotropic sigma this
 *     equals the reference's fddc_assemb iso branch up to the constant factor -8
 *     on both T and P (src/div_correc.f90:593-615,709-750), hence the same phi
 *     correction; the reference's anisotropic branch (src/div_correc.f90:616-664,
 *     751-819) is not restated because its P and T disagree in sign and scale.
 *
 * How each routine is pinned (tests/test_oracle.py, tests/test_abi_cpu.py):
 *   orc_build_system   row-scaled operator complex symmetric; curl(grad) = 0 at
 *                      sigma = 0; layered model = 1-D FD solution, O(h^2) to the
 *                      analytic field; full-tensor rotation equivariance (pins
 *                      the off-diagonal conductivity averaging: a flipped xy sign
 *                      or swapped xz/yz moves the result by 0.24 / 0.11 relative)
 *   orc_ssor_apply     dense (D+L) D^-1 (D+U) factorisation
 *   orc_bicgstab       iterates equal SciPy's BiCGStab; dense solution reached
 *   orc_gpbicg         dense solution reached; first step equals SciPy's BiCGStab
 *                      step (eta = 0); finite termination within the number of
 *                      interior unknowns (a dropped or sign-flipped term of u, z,
 *                      y, w, eta or zeta leaves >= 1e-4 of the residual)
 *   orc_cocr           dense solution reached (new solver: the reference's COCR
 *                      cannot run, so there is nothing to be at parity with)
 *   orc_dc_matrix      isotropic P, T equal the reference formulas x 8; DC drives
 *                      G^T M x to the PCG tolerance
 *   orc_mt1d_anl       plane wave in a uniform earth; axial anisotropy decouples
 *                      the modes; two layers against a fine 1-D FD solve
 *   orc_mt1d_anl_full, orc_build_system_full
 *                      layered model with a rotated (full) tensor: the 3-D
 *                      solve with the full-tensor boundary reproduces both
 *                      1-D components inside the grid to discretisation
 *                      error, the diagonal boundary does not
 *   orc_extint_v       polynomial reproduction of I2 / I4 on nonuniform nodes;
 *                      extrapolation constants from the text
 *   orc_mesh_coarsen   conductance integral preserved
 *   post_process       (oracle.py) rd = 1 exact on quadratic fields; a
 *                      half-space gives 100 ohm-m and 45 degrees; rd = 2 (the
 *                      reference's transposed station indexing) pinned in
 *                      closed form on biquadratic fields; Z and T equal numpy's
 *                      solve of E = Z H and Hz = T (Hx, Hy); a layered model has
 *                      T = 0; the topography branch samples each column's own
 *                      plane with its own cell thickness on fields linear in z
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

static const double PI = 3.141592653589793238462643;
static const double MU0 = 4.0 * 3.141592653589793238462643 * 1e-7; /* src/global_para.f90:14 */

/* ------------------------------------------------------------------ sizes */
static inline int64_t n_xedge(int nx, int ny, int nz) { return (int64_t)nx * (ny + 1) * (nz + 1); }
static inline int64_t n_yedge(int nx, int ny, int nz) { return (int64_t)(nx + 1) * ny * (nz + 1); }
static inline int64_t n_zedge(int nx, int ny, int nz) { return (int64_t)(nx + 1) * (ny + 1) * nz; }
int64_t orc_nedges(int nx, int ny, int nz)
{
    return n_xedge(nx, ny, nz) + n_yedge(nx, ny, nz) + n_zedge(nx, ny, nz);
}
int64_t orc_nnodes(int nx, int ny, int nz) { return (int64_t)(nx + 1) * (ny + 1) * (nz + 1); }

/* 1-based index helpers that mirror the reference formulas literally. */
#define EX1(i, j, k) ((int64_t)((k)-1) * nx * (ny + 1) + (int64_t)((j)-1) * nx + (i))
#define EY1(i, j, k) (xnl + (int64_t)((k)-1) * (nx + 1) * ny + (int64_t)((j)-1) * (nx + 1) + (i))
#define EZ1(i, j, k) (xnl + ynl + (int64_t)((k)-1) * (nx + 1) * (ny + 1) + (int64_t)((j)-1) * (nx + 1) + (i))
#define SIG(ie1, comp) sigma[(int64_t)(comp)*ncell + (ie1)-1]

/* --------------------------------------------------------------------------
 * fd_assemb + dcfast  (src/fdfe_main.f90:226-396, 399-546)
 *
 * Builds the 13-slot COO rows of the complex curl-curl operator for interior
 * edges, flags the Dirichlet edges, writes the Dirichlet data from the 1-D
 * field e_dc (2*nz+2 values: Ex at z-nodes then Ey at z-nodes), moves the
 * boundary columns to the right-hand sides p1 (xy) and p2 (yx), puts identity
 * rows on the boundary and compacts to CSR with 1-based ia (int64) / ja (int32)
 * exactly as dcfast does (column order = slot order, zero values dropped).
 * ia must hold nl+1, ja/val 13*nl, p1/p2/bnd nl entries.  Returns nnz.
 * -------------------------------------------------------------------------- */
static int64_t build_system_core(int nx, int ny, int nz, const double *a, const double *b,
                                 const double *c, const double *sigma, double freq,
                                 const zc *e_dc, const zc *e_x, int64_t *ia, int32_t *ja, zc *val,
                                 zc *p1, zc *p2, uint8_t *bnd);

int64_t orc_build_system(int nx, int ny, int nz, const double *a, const double *b,
                         const double *c, const double *sigma, double freq,
                         const zc *e_dc, int64_t *ia, int32_t *ja, zc *val, zc *p1,
                         zc *p2, uint8_t *bnd)
{
    return build_system_core(nx, ny, nz, a, b, c, sigma, freq, e_dc, NULL, ia, ja, val, p1, p2, bnd);
}

/* Full-tensor boundary (a deliberate deviation from the reference, which
 * builds the 1-D field from the diagonal of the first column only and sets the
 * cross-polar boundary components to 0, src/fdfe_main.f90:938-942, 470-494):
 * as orc_build_system, plus the cross-polar components of the 1-D field on
 * the boundary.  e_x holds 2*nz+2 values: Ey of the xy mode at the
 * z-nodes, then Ex of the yx mode; boundary y-edges of p1 and x-edges of p2
 * at node plane k get e_x[k] / e_x[nz+1+k] (0 on the top plane, where the
 * modes are normalised to E = (1,0) and (0,1)). */
int64_t orc_build_system_full(int nx, int ny, int nz, const double *a, const double *b,
                              const double *c, const double *sigma, double freq,
                              const zc *e_dc, const zc *e_x, int64_t *ia, int32_t *ja, zc *val,
                              zc *p1, zc *p2, uint8_t *bnd)
{
    return build_system_core(nx, ny, nz, a, b, c, sigma, freq, e_dc, e_x, ia, ja, val, p1, p2, bnd);
}

static int64_t build_system_core(int nx, int ny, int nz, const double *a, const double *b,
                                 const double *c, const double *sigma, double freq,
                                 const zc *e_dc, const zc *e_x, int64_t *ia, int32_t *ja, zc *val,
                                 zc *p1, zc *p2, uint8_t *bnd)
{
    const int64_t xnl = n_xedge(nx, ny, nz), ynl = n_yedge(nx, ny, nz);
    const int64_t nl = xnl + ynl + n_zedge(nx, ny, nz);
    const int64_t ncell = (int64_t)nx * ny * nz;
    const double w = 2 * PI * freq;
    const zc coef = I * w * MU0; /* src/fdfe_main.f90:237-238 */
    zc *v = calloc((size_t)nl * 13, sizeof(zc));
    int32_t *vc = calloc((size_t)nl * 13, sizeof(int32_t));
    if (!v || !vc) { free(v); free(vc); return -1; }

    /* Ex rows, src/fdfe_main.f90:248-295 */
    for (int k = 2; k <= nz; k++)
        for (int j = 2; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int64_t ie = (int64_t)(k - 2) * nx * ny + (int64_t)(j - 2) * nx + i;
                double bb = b[j - 2] + b[j - 1], cc = c[k - 2] + c[k - 1];
                double sigmas = (SIG(ie, 0) * b[j - 2] * c[k - 2] + SIG(ie + nx, 0) * b[j - 1] * c[k - 2] +
                                 SIG(ie + (int64_t)nx * ny, 0) * b[j - 2] * c[k - 1] +
                                 SIG(ie + (int64_t)nx * ny + nx, 0) * b[j - 1] * c[k - 1]) / (bb * cc);
                double sv1 = (SIG(ie, 3) * b[j - 2] * c[k - 2] + SIG(ie + (int64_t)nx * ny, 3) * b[j - 2] * c[k - 1]) / 2 / (bb * cc);
                double sv2 = (SIG(ie + nx, 3) * b[j - 1] * c[k - 2] + SIG(ie + (int64_t)nx * ny + nx, 3) * b[j - 1] * c[k - 1]) / 2 / (bb * cc);
                double sv3 = (SIG(ie, 4) * b[j - 2] * c[k - 2] + SIG(ie + nx, 4) * b[j - 1] * c[k - 2]) / 2 / (bb * cc);
                double sv4 = (SIG(ie + (int64_t)nx * ny, 4) * b[j - 2] * c[k - 1] + SIG(ie + nx + (int64_t)nx * ny, 4) * b[j - 1] * c[k - 1]) / 2 / (bb * cc);
                int64_t ind = EX1(i, j, k);
                zc *r = v + (ind - 1) * 13;
                int32_t *rc = vc + (ind - 1) * 13;
                r[0] = -2 / (c[k - 2] * cc);  rc[0] = (int32_t)(ind - (int64_t)nx * (ny + 1));
                r[1] = -2 / (b[j - 2] * bb);  rc[1] = (int32_t)(ind - nx);
                r[3] = -2 / (b[j - 1] * bb);  rc[3] = (int32_t)(ind + nx);
                r[4] = -2 / (c[k - 1] * cc);  rc[4] = (int32_t)(ind + (int64_t)nx * (ny + 1));
                r[2] = -(r[0] + r[1] + r[3] + r[4]) - coef * sigmas; rc[2] = (int32_t)ind;
                double gy = 2 / (a[i - 1] * bb);
                r[5] = gy - coef * sv1;   rc[5] = (int32_t)EY1(i, j - 1, k);
                r[6] = -gy - coef * sv1;  rc[6] = (int32_t)EY1(i + 1, j - 1, k);
                r[7] = -gy - coef * sv2;  rc[7] = (int32_t)EY1(i, j, k);
                r[8] = gy - coef * sv2;   rc[8] = (int32_t)EY1(i + 1, j, k);
                double gz = 2 / (a[i - 1] * cc);
                r[9] = gz - coef * sv3;   rc[9] = (int32_t)EZ1(i, j, k - 1);
                r[10] = -gz - coef * sv3; rc[10] = (int32_t)EZ1(i + 1, j, k - 1);
                r[11] = -gz - coef * sv4; rc[11] = (int32_t)EZ1(i, j, k);
                r[12] = gz - coef * sv4;  rc[12] = (int32_t)EZ1(i + 1, j, k);
            }
    /* Ey rows, src/fdfe_main.f90:297-344 */
    for (int k = 2; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 2; i <= nx; i++) {
                int64_t ie = (int64_t)(k - 2) * nx * ny + (int64_t)(j - 1) * nx + i - 1;
                int64_t nxy = (int64_t)nx * ny;
                double aa = a[i - 2] + a[i - 1], cc = c[k - 2] + c[k - 1];
                double sigmas = (SIG(ie, 1) * a[i - 2] * c[k - 2] + SIG(ie + 1, 1) * a[i - 1] * c[k - 2] +
                                 SIG(ie + nxy, 1) * a[i - 2] * c[k - 1] + SIG(ie + nxy + 1, 1) * a[i - 1] * c[k - 1]) / (aa * cc);
                double sv1 = (SIG(ie, 3) * a[i - 2] * c[k - 2] + SIG(ie + nxy, 3) * a[i - 2] * c[k - 1]) / 2 / (aa * cc);
                double sv2 = (SIG(ie + 1, 3) * a[i - 1] * c[k - 2] + SIG(ie + nxy + 1, 3) * a[i - 1] * c[k - 1]) / 2 / (aa * cc);
                double sv3 = (SIG(ie, 5) * a[i - 2] * c[k - 2] + SIG(ie + 1, 5) * a[i - 1] * c[k - 2]) / 2 / (aa * cc);
                double sv4 = (SIG(ie + nxy, 5) * a[i - 2] * c[k - 1] + SIG(ie + nxy + 1, 5) * a[i - 1] * c[k - 1]) / 2 / (aa * cc);
                int64_t ind = EY1(i, j, k);
                zc *r = v + (ind - 1) * 13;
                int32_t *rc = vc + (ind - 1) * 13;
                double gx = 2 / (b[j - 1] * aa);
                r[0] = gx - coef * sv1;  rc[0] = (int32_t)EX1(i - 1, j, k);
                r[1] = -gx - coef * sv2; rc[1] = (int32_t)EX1(i, j, k);
                r[2] = -gx - coef * sv1; rc[2] = (int32_t)EX1(i - 1, j + 1, k);
                r[3] = gx - coef * sv2;  rc[3] = (int32_t)EX1(i, j + 1, k);
                r[4] = -2 / (c[k - 2] * cc); rc[4] = (int32_t)(ind - (int64_t)(nx + 1) * ny);
                r[5] = -2 / (a[i - 2] * aa); rc[5] = (int32_t)(ind - 1);
                r[7] = -2 / (a[i - 1] * aa); rc[7] = (int32_t)(ind + 1);
                r[8] = -2 / (c[k - 1] * cc); rc[8] = (int32_t)(ind + (int64_t)(nx + 1) * ny);
                r[6] = -(r[4] + r[5] + r[7] + r[8]) - coef * sigmas; rc[6] = (int32_t)ind;
                double gz = 2 / (b[j - 1] * cc);
                r[9] = gz - coef * sv3;   rc[9] = (int32_t)EZ1(i, j, k - 1);
                r[10] = -gz - coef * sv3; rc[10] = (int32_t)EZ1(i, j + 1, k - 1);
                r[11] = -gz - coef * sv4; rc[11] = (int32_t)EZ1(i, j, k);
                r[12] = gz - coef * sv4;  rc[12] = (int32_t)EZ1(i, j + 1, k);
            }
    /* Ez rows, src/fdfe_main.f90:346-393 */
    for (int k = 1; k <= nz; k++)
        for (int j = 2; j <= ny; j++)
            for (int i = 2; i <= nx; i++) {
                int64_t ie = (int64_t)(k - 1) * nx * ny + (int64_t)(j - 2) * nx + i - 1;
                double aa = a[i - 2] + a[i - 1], bb = b[j - 2] + b[j - 1];
                double sigmas = (SIG(ie, 2) * b[j - 2] * a[i - 2] + SIG(ie + nx, 2) * b[j - 1] * a[i - 2] +
                                 SIG(ie + 1, 2) * b[j - 2] * a[i - 1] + SIG(ie + nx + 1, 2) * b[j - 1] * a[i - 1]) / (bb * aa);
                double sv1 = (SIG(ie, 4) * b[j - 2] * a[i - 2] + SIG(ie + nx, 4) * b[j - 1] * a[i - 2]) / 2 / (bb * aa);
                double sv2 = (SIG(ie + 1, 4) * b[j - 2] * a[i - 1] + SIG(ie + nx + 1, 4) * b[j - 1] * a[i - 1]) / 2 / (bb * aa);
                double sv3 = (SIG(ie, 5) * b[j - 2] * a[i - 2] + SIG(ie + 1, 5) * b[j - 2] * a[i - 1]) / 2 / (bb * aa);
                double sv4 = (SIG(ie + nx, 5) * b[j - 1] * a[i - 2] + SIG(ie + nx + 1, 5) * b[j - 1] * a[i - 1]) / 2 / (bb * aa);
                int64_t ind = EZ1(i, j, k);
                zc *r = v + (ind - 1) * 13;
                int32_t *rc = vc + (ind - 1) * 13;
                double gx = 2 / (c[k - 1] * aa);
                r[0] = gx - coef * sv1;  rc[0] = (int32_t)EX1(i - 1, j, k);
                r[1] = -gx - coef * sv2; rc[1] = (int32_t)EX1(i, j, k);
                r[2] = -gx - coef * sv1; rc[2] = (int32_t)EX1(i - 1, j, k + 1);
                r[3] = gx - coef * sv2;  rc[3] = (int32_t)EX1(i, j, k + 1);
                double gy = 2 / (c[k - 1] * bb);
                r[4] = gy - coef * sv3;  rc[4] = (int32_t)EY1(i, j - 1, k);
                r[5] = -gy - coef * sv4; rc[5] = (int32_t)EY1(i, j, k);
                r[6] = -gy - coef * sv3; rc[6] = (int32_t)EY1(i, j - 1, k + 1);
                r[7] = gy - coef * sv4;  rc[7] = (int32_t)EY1(i, j, k + 1);
                r[8] = -2 / (b[j - 2] * bb);  rc[8] = (int32_t)(ind - (nx + 1));
                r[9] = -2 / (a[i - 2] * aa);  rc[9] = (int32_t)(ind - 1);
                r[11] = -2 / (a[i - 1] * aa); rc[11] = (int32_t)(ind + 1);
                r[12] = -2 / (b[j - 1] * bb); rc[12] = (int32_t)(ind + (nx + 1));
                r[10] = -(r[8] + r[9] + r[11] + r[12]) - coef * sigmas; rc[10] = (int32_t)ind;
            }

    /* Dirichlet flags, src/fdfe_main.f90:436-468 */
    memset(bnd, 0, (size_t)nl);
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx; i++)
                if (k == 1 || k == nz + 1 || j == 1 || j == ny + 1) bnd[EX1(i, j, k) - 1] = 1;
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx + 1; i++)
                if (k == 1 || k == nz + 1 || i == 1 || i == nx + 1) bnd[EY1(i, j, k) - 1] = 1;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++)
                if (i == 1 || i == nx + 1 || j == 1 || j == ny + 1) bnd[EZ1(i, j, k) - 1] = 1;

    /* Dirichlet data, src/fdfe_main.f90:470-494 (e_dc 0-based: e_dc[k-1] = ex(k)) */
    for (int64_t e = 0; e < nl; e++) p1[e] = p2[e] = 0;
    for (int64_t e = 0; e < (int64_t)nx * (ny + 1); e++) p1[e] = 1.0;           /* top layer */
    for (int k = 1; k <= nz; k++) {
        int64_t lay = (int64_t)k * nx * (ny + 1);
        for (int i = 0; i < nx; i++) {
            p1[lay + i] = e_dc[k];                                                /* j = 1 row    */
            p1[(int64_t)(k + 1) * nx * (ny + 1) - nx + i] = e_dc[k];             /* j = ny+1 row */
        }
    }
    for (int64_t e = xnl - (int64_t)nx * (ny + 1); e < xnl; e++) p1[e] = e_dc[nz];
    for (int64_t e = xnl; e < xnl + (int64_t)ny * (nx + 1); e++) p2[e] = 1.0;
    for (int k = 1; k <= nz; k++)
        for (int jj = 1; jj <= ny; jj++) {
            p2[xnl + (int64_t)k * (nx + 1) * ny + (int64_t)(jj - 1) * (nx + 1)] = e_dc[nz + 1 + k];
            p2[xnl + (int64_t)k * (nx + 1) * ny + (int64_t)jj * (nx + 1) - 1] = e_dc[nz + 1 + k];
        }
    for (int64_t e = xnl + ynl - (int64_t)(nx + 1) * ny; e < xnl + ynl; e++) p2[e] = e_dc[2 * nz + 1];
    if (e_x) {
        /* cross-polar boundary values: y-edges of p1, x-edges of p2 */
        for (int k = 1; k <= nz + 1; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx + 1; i++)
                    if (k == 1 || k == nz + 1 || i == 1 || i == nx + 1)
                        p1[EY1(i, j, k) - 1] = (k == 1) ? 0 : e_x[k - 1];
        for (int k = 1; k <= nz + 1; k++)
            for (int j = 1; j <= ny + 1; j++)
                for (int i = 1; i <= nx; i++)
                    if (k == 1 || k == nz + 1 || j == 1 || j == ny + 1)
                        p2[EX1(i, j, k) - 1] = (k == 1) ? 0 : e_x[nz + k];
    }

    /* boundary elimination and CSR compaction, src/fdfe_main.f90:496-530 */
    int64_t kk = 0;
    ia[0] = 1;
    for (int64_t r = 0; r < nl; r++) {
        if (!bnd[r]) {
            zc t1 = 0, t2 = 0;
            for (int s = 0; s < 13; s++) {
                int64_t ind = r * 13 + s;
                if (vc[ind] > 0) {
                    int64_t col = vc[ind] - 1;
                    if (!bnd[col] && v[ind] != 0) {
                        val[kk] = v[ind];
                        ja[kk] = vc[ind];
                        kk++;
                    } else {
                        t1 += v[ind] * p1[col];
                        t2 += v[ind] * p2[col];
                    }
                }
            }
            p1[r] -= t1;
            p2[r] -= t2;
        } else {
            val[kk] = 1.0;
            ja[kk] = (int32_t)(r + 1);
            kk++;
        }
        ia[r + 1] = kk + 1;
    }
    free(v);
    free(vc);
    return kk;
}

/* Dual-volume row scale of src/fdfe_main.f90:252,301,350 (computed there but not
 * applied, line 292).  ve*A is complex symmetric on the interior rows.  Boundary
 * rows get scale 1. */
void orc_row_scale(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                   const uint8_t *bnd, double *ve)
{
    const int64_t xnl = n_xedge(nx, ny, nz), ynl = n_yedge(nx, ny, nz);
    const int64_t nl = xnl + ynl + n_zedge(nx, ny, nz);
    for (int64_t e = 0; e < nl; e++) ve[e] = 1.0;
    for (int k = 2; k <= nz; k++)
        for (int j = 2; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                ve[EX1(i, j, k) - 1] = a[i - 1] * (b[j - 2] + b[j - 1]) * (c[k - 2] + c[k - 1]) / 4;
    for (int k = 2; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 2; i <= nx; i++)
                ve[EY1(i, j, k) - 1] = b[j - 1] * (a[i - 2] + a[i - 1]) * (c[k - 2] + c[k - 1]) / 4;
    for (int k = 1; k <= nz; k++)
        for (int j = 2; j <= ny; j++)
            for (int i = 2; i <= nx; i++)
                ve[EZ1(i, j, k) - 1] = c[k - 1] * (b[j - 2] + b[j - 1]) * (a[i - 2] + a[i - 1]) / 4;
    for (int64_t e = 0; e < nl; e++)
        if (bnd[e]) ve[e] = 1.0;
}

/* zamux, src/div_correc.f90:431-449 (1-based CSR) */
void orc_zamux(int64_t n, const int64_t *ia, const int32_t *ja, const zc *a, const zc *x, zc *y)
{
    for (int64_t i = 0; i < n; i++) {
        zc t = 0;
        for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) t += a[k] * x[ja[k] - 1];
        y[i] = t;
    }
}

static double vnorm(int64_t n, const zc *z) /* src/model_selec.f90:631-643 */
{
    double s = 0;
    for (int64_t i = 0; i < n; i++) s += creal(z[i]) * creal(z[i]) + cimag(z[i]) * cimag(z[i]);
    return sqrt(s);
}
static zc cdot(int64_t n, const zc *x, const zc *y) /* Fortran dot_product: conj(x).y */
{
    zc s = 0;
    for (int64_t i = 0; i < n; i++) s += conj(x[i]) * y[i];
    return s;
}
static zc udot(int64_t n, const zc *x, const zc *y) /* unconjugated bilinear form x^T y */
{
    zc s = 0;
    for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* --------------------------------------------------------------------------
 * SSOR, src/div_correc.f90:171-220 (ssorfac) and src/solvers.f90:721-762 (lusol)
 * Lexicographic: forward sweep over rows 1..n with the strictly lower part,
 * scale by diag, backward sweep with the strictly upper part.  omega = 1 by
 * default (src/solvers.f90:14).  Returns +M^-1 y (sign deviation noted above).
 * -------------------------------------------------------------------------- */
typedef struct {
    int64_t n;
    const int64_t *ia;
    const int32_t *ja;
    const zc *a;
    zc *dinv; /* 1/a_ii */
    double omega;
} ssor_t;

static int ssor_setup(ssor_t *s, int64_t n, const int64_t *ia, const int32_t *ja, const zc *a, double omega)
{
    s->n = n; s->ia = ia; s->ja = ja; s->a = a; s->omega = omega;
    s->dinv = malloc((size_t)n * sizeof(zc));
    for (int64_t i = 0; i < n; i++) {
        zc d = 0;
        for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++)
            if (ja[k] - 1 == i) d = a[k];
        if (d == 0) return (int)(i + 1); /* zero pivot, src/div_correc.f90:208,217 */
        s->dinv[i] = 1.0 / d;
    }
    return 0;
}

static void ssor_apply(const ssor_t *s, const zc *y, zc *x)
{
    const double w = s->omega;
    for (int64_t i = 0; i < s->n; i++) {
        zc t = y[i];
        for (int64_t k = s->ia[i] - 1; k < s->ia[i + 1] - 1; k++) {
            int64_t j = s->ja[k] - 1;
            if (j < i) t -= s->a[k] * x[j];
        }
        x[i] = t * s->dinv[i] * w;
    }
    for (int64_t i = 0; i < s->n; i++) x[i] = x[i] / s->dinv[i] * (2 - w) / w;
    for (int64_t i = s->n - 1; i >= 0; i--) {
        zc t = x[i];
        for (int64_t k = s->ia[i] - 1; k < s->ia[i + 1] - 1; k++) {
            int64_t j = s->ja[k] - 1;
            if (j > i) t -= s->a[k] * x[j];
        }
        x[i] = s->dinv[i] * t * w;
    }
}

void orc_ssor_apply(int64_t n, const int64_t *ia, const int32_t *ja, const zc *a, double omega,
                    const zc *y, zc *x)
{
    ssor_t s;
    if (ssor_setup(&s, n, ia, ja, a, omega) == 0) ssor_apply(&s, y, x);
    free(s.dinv);
}

/* --------------------------------------------------------------------------
 * Divergence correction.
 *
 * Structure follows div_correc (src/div_correc.f90:11-147): t = T x, zero t on
 * boundary nodes, Jacobi-PCG on P (src/div_correc.f90:343-402) with tolerance
 * corr_tol and at most 2000 iterations, then x_e += (phi(end)-phi(start))/l_e on
 * every edge that does not have both end nodes on the boundary
 * (src/div_correc.f90:94-99).  Edge lengths and edge->node map as
 * src/div_correc.f90:831-855.
 *
 * Operators (see header): M_sigma is the corner-lumped edge mass that appears in
 * the assembled curl-curl rows: for each cell of volume V and each of its 8
 * corners, V/8 * sum_{d1,d2} sigma_{d1 d2} E_{d1} E_{d2} over the 3 cell edges
 * meeting at that corner.  T x = G^T M_sigma x (scaled by -8 this is the
 * reference's isotropic T), P = G^T M_sigma G (x8 = the reference's isotropic P).
 * -------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz;
    int64_t np, nl;
    int64_t *pia; int32_t *pja; double *pv; /* P, 0-based CSR, boundary rows identity */
    uint8_t *dcn;                          /* boundary node flags */
    const double *a, *b, *c, *sigma;
} dc_t;

static inline int64_t node0(int nx, int ny, int i, int j, int k)
{
    return ((int64_t)k * (ny + 1) + j) * (nx + 1) + i;
}

/* corner-lumped mass product y += M_sigma x over all edges (0-based grid). */
static void mass_apply(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                       const double *sigma, const zc *x, zc *y)
{
    const int64_t xnl = n_xedge(nx, ny, nz), ynl = n_yedge(nx, ny, nz);
    const int64_t ncell = (int64_t)nx * ny * nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t ie = ((int64_t)k * ny + j) * nx + i;
                double V8 = a[i] * b[j] * c[k] / 8;
                double s[3][3];
                s[0][0] = sigma[ie]; s[1][1] = sigma[ncell + ie]; s[2][2] = sigma[2 * ncell + ie];
                s[0][1] = s[1][0] = sigma[3 * ncell + ie];
                s[0][2] = s[2][0] = sigma[4 * ncell + ie];
                s[1][2] = s[2][1] = sigma[5 * ncell + ie];
                for (int cz = 0; cz < 2; cz++)
                    for (int cy = 0; cy < 2; cy++)
                        for (int cx = 0; cx < 2; cx++) {
                            int64_t e[3];
                            e[0] = ((int64_t)(k + cz) * (ny + 1) + (j + cy)) * nx + i;
                            e[1] = xnl + ((int64_t)(k + cz) * ny + j) * (nx + 1) + (i + cx);
                            e[2] = xnl + ynl + ((int64_t)k * (ny + 1) + (j + cy)) * (nx + 1) + (i + cx);
                            for (int d1 = 0; d1 < 3; d1++) {
                                zc acc = 0;
                                for (int d2 = 0; d2 < 3; d2++) acc += s[d1][d2] * x[e[d2]];
                                y[e[d1]] += V8 * acc;
                            }
                        }
            }
}

/* edge length, start node, end node (src/div_correc.f90:831-855) */
static void edge_info(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                      int64_t e, double *len, int64_t *n0, int64_t *n1)
{
    const int64_t xnl = n_xedge(nx, ny, nz), ynl = n_yedge(nx, ny, nz);
    if (e < xnl) {
        int i = (int)(e % nx); int64_t r = e / nx; int j = (int)(r % (ny + 1)); int k = (int)(r / (ny + 1));
        *len = a[i]; *n0 = node0(nx, ny, i, j, k); *n1 = *n0 + 1;
    } else if (e < xnl + ynl) {
        int64_t q = e - xnl;
        int i = (int)(q % (nx + 1)); int64_t r = q / (nx + 1); int j = (int)(r % ny); int k = (int)(r / ny);
        *len = b[j]; *n0 = node0(nx, ny, i, j, k); *n1 = *n0 + (nx + 1);
    } else {
        int64_t q = e - xnl - ynl;
        int i = (int)(q % (nx + 1)); int64_t r = q / (nx + 1); int j = (int)(r % (ny + 1)); int k = (int)(r / (ny + 1));
        *len = c[k]; *n0 = node0(nx, ny, i, j, k); *n1 = *n0 + (int64_t)(nx + 1) * (ny + 1);
    }
}

/* Apply G^T to an edge vector: t_n = sum_e (+-1/l_e) w_e.  G phi on edge e is
 * (phi(end)-phi(start))/l_e. */
static void grad_T(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                   const zc *w, zc *t)
{
    int64_t np = orc_nnodes(nx, ny, nz), nl = orc_nedges(nx, ny, nz);
    for (int64_t n = 0; n < np; n++) t[n] = 0;
    for (int64_t e = 0; e < nl; e++) {
        double l; int64_t n0, n1;
        edge_info(nx, ny, nz, a, b, c, e, &l, &n0, &n1);
        t[n1] += w[e] / l;
        t[n0] -= w[e] / l;
    }
}

/* Nodal operator P = G^T M_sigma G assembled cell by cell, stored as CSR with
 * up to 19 entries per interior node; boundary nodes carry identity rows
 * (src/div_correc.f90:665-670). */
static int dc_build(dc_t *d, int nx, int ny, int nz, const double *a, const double *b,
                    const double *c, const double *sigma)
{
    d->nx = nx; d->ny = ny; d->nz = nz; d->a = a; d->b = b; d->c = c; d->sigma = sigma;
    d->np = orc_nnodes(nx, ny, nz); d->nl = orc_nedges(nx, ny, nz);
    const int64_t np = d->np, ncell = (int64_t)nx * ny * nz;
    const int64_t sx = 1, sy = nx + 1, sz = (int64_t)(nx + 1) * (ny + 1);
    /* dense 27-slot band per node, compacted afterwards */
    double *band = calloc((size_t)np * 27, sizeof(double));
    d->dcn = calloc((size_t)np, 1);
    if (!band || !d->dcn) return -1;
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++)
                if (i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz)
                    d->dcn[node0(nx, ny, i, j, k)] = 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int64_t ie = ((int64_t)k * ny + j) * nx + i;
                double h[3] = {a[i], b[j], c[k]};
                double V8 = a[i] * b[j] * c[k] / 8;
                double s[3][3];
                s[0][0] = sigma[ie]; s[1][1] = sigma[ncell + ie]; s[2][2] = sigma[2 * ncell + ie];
                s[0][1] = s[1][0] = sigma[3 * ncell + ie];
                s[0][2] = s[2][0] = sigma[4 * ncell + ie];
                s[1][2] = s[2][1] = sigma[5 * ncell + ie];
                for (int cz = 0; cz < 2; cz++)
                    for (int cy = 0; cy < 2; cy++)
                        for (int cx = 0; cx < 2; cx++) {
                            /* gradient at the corner: node list and weights per direction */
                            int off[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                            (void)off;
                            int64_t nlo[3], nhi[3];
                            int64_t base = node0(nx, ny, i, j, k);
                            int64_t corner = base + cx * sx + cy * sy + cz * sz;
                            /* direction d: edge within the cell touching the corner goes
                               from the low node to the high node along d */
                            nlo[0] = base + cy * sy + cz * sz;          nhi[0] = nlo[0] + sx;
                            nlo[1] = base + cx * sx + cz * sz;          nhi[1] = nlo[1] + sy;
                            nlo[2] = base + cx * sx + cy * sy;          nhi[2] = nlo[2] + sz;
                            (void)corner;
                            for (int d1 = 0; d1 < 3; d1++)
                                for (int d2 = 0; d2 < 3; d2++) {
                                    double wgt = V8 * s[d1][d2] / (h[d1] * h[d2]);
                                    int64_t r1[2] = {nhi[d1], nlo[d1]};
                                    double g1[2] = {1, -1};
                                    int64_t r2[2] = {nhi[d2], nlo[d2]};
                                    double g2[2] = {1, -1};
                                    for (int p = 0; p < 2; p++)
                                        for (int q = 0; q < 2; q++) {
                                            int64_t row = r1[p], col = r2[q];
                                            int64_t dz = col / sz - row / sz;
                                            int64_t dy = (col % sz) / sy - (row % sz) / sy;
                                            int64_t dx = col % sy - row % sy;
                                            int slot = (int)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
                                            band[row * 27 + slot] += wgt * g1[p] * g2[q];
                                        }
                                }
                        }
            }
    /* compact: interior rows keep nonzero couplings to all columns, boundary rows identity */
    d->pia = malloc((size_t)(np + 1) * sizeof(int64_t));
    d->pja = malloc((size_t)np * 27 * sizeof(int32_t));
    d->pv = malloc((size_t)np * 27 * sizeof(double));
    int64_t kk = 0;
    d->pia[0] = 0;
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++) {
                int64_t n = node0(nx, ny, i, j, k);
                if (d->dcn[n]) {
                    d->pja[kk] = (int32_t)n; d->pv[kk] = 1.0; kk++;
                } else {
                    for (int s = 0; s < 27; s++) {
                        double vv = band[n * 27 + s];
                        if (vv == 0) continue;
                        int dz = s / 9 - 1, dy = (s / 3) % 3 - 1, dx = s % 3 - 1;
                        d->pja[kk] = (int32_t)(n + dz * sz + dy * sy + dx * sx);
                        d->pv[kk] = vv;
                        kk++;
                    }
                }
                d->pia[n + 1] = kk;
            }
    free(band);
    return 0;
}

static void dc_free(dc_t *d)
{
    free(d->pia); free(d->pja); free(d->pv); free(d->dcn);
    d->pia = NULL; d->pja = NULL; d->pv = NULL; d->dcn = NULL;
}

static void p_apply(const dc_t *d, const zc *x, zc *y)
{
    for (int64_t i = 0; i < d->np; i++) {
        zc t = 0;
        for (int64_t k = d->pia[i]; k < d->pia[i + 1]; k++) t += d->pv[k] * x[d->pja[k]];
        y[i] = t;
    }
}

/* Jacobi-PCG, src/div_correc.f90:343-402.  Returns the iteration count. */
static int dc_pcg(const dc_t *d, const zc *bvec, zc *u, double tol, int maxit)
{
    int64_t n = d->np;
    zc *r = malloc(n * sizeof(zc)), *rold = malloc(n * sizeof(zc)), *p = malloc(n * sizeof(zc)),
       *v = malloc(n * sizeof(zc)), *z = malloc(n * sizeof(zc)), *zold = malloc(n * sizeof(zc));
    double *mj = malloc(n * sizeof(double));
    for (int64_t i = 0; i < n; i++)
        for (int64_t k = d->pia[i]; k < d->pia[i + 1]; k++)
            if (d->pja[k] == i) mj[i] = d->pv[k];
    p_apply(d, u, r);
    for (int64_t i = 0; i < n; i++) r[i] = bvec[i] - r[i];
    double babs = vnorm(n, bvec);
    double err = babs > 0 ? vnorm(n, r) / babs : 0;
    for (int64_t i = 0; i < n; i++) p[i] = r[i] / mj[i];
    int it = 0;
    for (int i = 0; i < maxit; i++) {
        if (err < tol) break; /* reference also requires err > 0; an all-zero rhs stops here */
        memcpy(rold, r, n * sizeof(zc));
        p_apply(d, p, v);
        for (int64_t q = 0; q < n; q++) z[q] = r[q] / mj[q];
        zc alpha = cdot(n, z, r) / cdot(n, p, v);
        for (int64_t q = 0; q < n; q++) { u[q] += alpha * p[q]; r[q] -= alpha * v[q]; }
        for (int64_t q = 0; q < n; q++) { z[q] = r[q] / mj[q]; zold[q] = rold[q] / mj[q]; }
        zc beta = cdot(n, z, r) / cdot(n, zold, rold);
        for (int64_t q = 0; q < n; q++) p[q] = r[q] / mj[q] + beta * p[q];
        err = vnorm(n, r) / babs;
        it++;
    }
    free(r); free(rold); free(p); free(v); free(z); free(zold); free(mj);
    return it;
}

/* div_correc, src/div_correc.f90:11-100 */
static int dc_correct(const dc_t *d, zc *x, double tol)
{
    int64_t np = d->np, nl = d->nl;
    zc *w = calloc(nl, sizeof(zc)), *t = malloc(np * sizeof(zc)), *phi = calloc(np, sizeof(zc));
    mass_apply(d->nx, d->ny, d->nz, d->a, d->b, d->c, d->sigma, x, w);
    grad_T(d->nx, d->ny, d->nz, d->a, d->b, d->c, w, t);
    for (int64_t i = 0; i < np; i++)
        if (d->dcn[i]) t[i] = 0;
    /* P phi = G^T M x  and  x <- x - G phi  makes G^T M x vanish on interior nodes */
    int it = dc_pcg(d, t, phi, tol, 2000);
    for (int64_t e = 0; e < nl; e++) {
        double l; int64_t n0, n1;
        edge_info(d->nx, d->ny, d->nz, d->a, d->b, d->c, e, &l, &n0, &n1);
        if (!(d->dcn[n0] && d->dcn[n1])) x[e] -= (phi[n1] - phi[n0]) / l;
    }
    free(w); free(t); free(phi);
    return it;
}

/* exported pieces used by the oracle's own tests */
void orc_mass_apply(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                    const double *sigma, const zc *x, zc *y)
{
    int64_t nl = orc_nedges(nx, ny, nz);
    for (int64_t e = 0; e < nl; e++) y[e] = 0;
    mass_apply(nx, ny, nz, a, b, c, sigma, x, y);
}
void orc_grad(int nx, int ny, int nz, const double *a, const double *b, const double *c,
              const zc *phi, zc *ge)
{
    int64_t nl = orc_nedges(nx, ny, nz);
    for (int64_t e = 0; e < nl; e++) {
        double l; int64_t n0, n1;
        edge_info(nx, ny, nz, a, b, c, e, &l, &n0, &n1);
        ge[e] = (phi[n1] - phi[n0]) / l;
    }
}
void orc_grad_T(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                const zc *w, zc *t)
{
    grad_T(nx, ny, nz, a, b, c, w, t);
}
/* P as CSR (0-based), caller gives capacity 27*np for pja/pv; returns nnz */
int64_t orc_dc_matrix(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                      const double *sigma, int64_t *pia, int32_t *pja, double *pv, uint8_t *dcn)
{
    dc_t d;
    if (dc_build(&d, nx, ny, nz, a, b, c, sigma)) return -1;
    int64_t nnz = d.pia[d.np];
    memcpy(pia, d.pia, (d.np + 1) * sizeof(int64_t));
    memcpy(pja, d.pja, nnz * sizeof(int32_t));
    memcpy(pv, d.pv, nnz * sizeof(double));
    memcpy(dcn, d.dcn, d.np);
    dc_free(&d);
    return nnz;
}
int orc_div_correct(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                    const double *sigma, zc *x, double tol)
{
    dc_t d;
    if (dc_build(&d, nx, ny, nz, a, b, c, sigma)) return -1;
    int it = dc_correct(&d, x, tol);
    dc_free(&d);
    return it;
}

/* --------------------------------------------------------------------------
 * Krylov solvers.  All take the reference CSR (1-based), rhs b, initial guess x
 * (in/out), tolerance eps on ||b-Ax||/||b||, maxit, and the divergence
 * correction settings (corr_intv <= 0 disables it).  bnorm > 0 replaces ||b||
 * as the normaliser (used for the interior-residual convention, where the
 * caller also imposes the Dirichlet values in x).  stats[0] = iterations,
 * stats[1] = final relative residual (recursive), stats[2] = DC calls,
 * stats[3] = total PCG iterations.
 * -------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz;
    const double *a, *b, *c, *sigma;
    int corr_intv;
    double corr_tol;
} dcopt_t;

/* Bicg_stab_p, src/solvers.f90:522-719 (right-preconditioned BiCGStab, van der
 * Vorst 1992), divergence correction every corr_intv iterations
 * (src/solvers.f90:630-646) followed by r = b - Ax, rhat = r. */
int orc_bicgstab(int64_t n, const int64_t *ia, const int32_t *ja, const zc *A, const zc *b, zc *x,
                 double eps, int maxit, int nx, int ny, int nz, const double *ga, const double *gb,
                 const double *gc, const double *sigma, int corr_intv, double corr_tol,
                 double omega, double bnorm, double *stats)
{
    ssor_t s;
    if (ssor_setup(&s, n, ia, ja, A, omega)) { free(s.dinv); return -2; }
    dc_t d = {0};
    int use_dc = corr_intv > 0;
    if (use_dc && dc_build(&d, nx, ny, nz, ga, gb, gc, sigma)) return -1;
    zc *r = malloc(n * sizeof(zc)), *rhat = malloc(n * sizeof(zc)), *p = calloc(n, sizeof(zc)),
       *y = malloc(n * sizeof(zc)), *vk = calloc(n, sizeof(zc)), *sv = malloc(n * sizeof(zc)),
       *z = malloc(n * sizeof(zc)), *t = malloc(n * sizeof(zc));
    orc_zamux(n, ia, ja, A, x, r);
    for (int64_t i = 0; i < n; i++) { r[i] = b[i] - r[i]; rhat[i] = r[i]; }
    zc rho = 1, alpha = 1, om = 1;
    double babs = bnorm > 0 ? bnorm : vnorm(n, b);
    double err = vnorm(n, r) / babs;
    int it = 0, ndc = 0, npcg = 0, status = 0;
    for (int i = 1; i <= maxit; i++) {
        if (err < eps) break;
        if (use_dc && i % corr_intv == 0) {
            npcg += dc_correct(&d, x, corr_tol);
            ndc++;
            orc_zamux(n, ia, ja, A, x, r);
            for (int64_t q = 0; q < n; q++) { r[q] = b[q] - r[q]; rhat[q] = r[q]; }
        }
        zc rhoold = rho;
        rho = cdot(n, rhat, r);
        zc beta = (rho / rhoold) * (alpha / om);
        for (int64_t q = 0; q < n; q++) p[q] = r[q] + beta * (p[q] - om * vk[q]);
        ssor_apply(&s, p, y);
        orc_zamux(n, ia, ja, A, y, vk);
        alpha = rho / cdot(n, rhat, vk);
        for (int64_t q = 0; q < n; q++) sv[q] = r[q] - alpha * vk[q];
        ssor_apply(&s, sv, z);
        orc_zamux(n, ia, ja, A, z, t);
        om = cdot(n, t, sv) / cdot(n, t, t);
        for (int64_t q = 0; q < n; q++) { x[q] += alpha * y[q] + om * z[q]; r[q] = sv[q] - om * t[q]; }
        err = vnorm(n, r) / babs;
        it = i;
        if (isnan(err)) { status = -3; break; }
    }
    if (status == 0 && !(err < eps)) status = 1;
    stats[0] = it; stats[1] = err; stats[2] = ndc; stats[3] = npcg;
    free(r); free(rhat); free(p); free(y); free(vk); free(sv); free(z); free(t);
    free(s.dinv);
    if (use_dc) dc_free(&d);
    return status;
}

/* pgpbicg, src/solvers.f90:137-315 (GPBiCG, right preconditioning), DC every
 * corr_intv iterations followed by r = b - Ax, r0 = r (src/solvers.f90:223-235). */
int orc_gpbicg(int64_t n, const int64_t *ia, const int32_t *ja, const zc *A, const zc *b, zc *x,
               double eps, int maxit, int nx, int ny, int nz, const double *ga, const double *gb,
               const double *gc, const double *sigma, int corr_intv, double corr_tol,
               double omega_ssor, double bnorm, double *stats)
{
    ssor_t s;
    if (ssor_setup(&s, n, ia, ja, A, omega_ssor)) { free(s.dinv); return -2; }
    dc_t d = {0};
    int use_dc = corr_intv > 0;
    if (use_dc && dc_build(&d, nx, ny, nz, ga, gb, gc, sigma)) return -1;
    size_t sz = n * sizeof(zc);
    zc *r = malloc(sz), *r0 = malloc(sz), *rp = malloc(sz), *t = calloc(n, sizeof(zc)),
       *w = calloc(n, sizeof(zc)), *p = calloc(n, sizeof(zc)), *u = calloc(n, sizeof(zc)),
       *tp = calloc(n, sizeof(zc)), *z = calloc(n, sizeof(zc)), *y = malloc(sz), *ap = malloc(sz),
       *at = malloc(sz), *tmp = malloc(sz), *tpold = malloc(sz), *rold = malloc(sz);
    orc_zamux(n, ia, ja, A, x, r);
    for (int64_t q = 0; q < n; q++) { r[q] = b[q] - r[q]; r0[q] = r[q]; }
    ssor_apply(&s, r, rp);
    zc beta = 0, alpha, om, eta;
    double babs = bnorm > 0 ? bnorm : vnorm(n, b);
    double err = vnorm(n, r) / babs;
    int it = 0, ndc = 0, npcg = 0, status = 1;
    if (err < eps) status = 0;
    for (int i = 1; i <= maxit && status == 1; i++) {
        if (use_dc && i % corr_intv == 0) {
            npcg += dc_correct(&d, x, corr_tol);
            ndc++;
            orc_zamux(n, ia, ja, A, x, r);
            for (int64_t q = 0; q < n; q++) { r[q] = b[q] - r[q]; r0[q] = r[q]; }
        }
        memcpy(tpold, tp, sz);
        memcpy(rold, r, sz);
        for (int64_t q = 0; q < n; q++) p[q] = rp[q] + beta * (p[q] - u[q]);
        orc_zamux(n, ia, ja, A, p, ap);
        alpha = cdot(n, r0, r) / cdot(n, r0, ap);
        for (int64_t q = 0; q < n; q++) { y[q] = t[q] - r[q] - alpha * w[q] + alpha * ap[q]; t[q] = r[q] - alpha * ap[q]; }
        ssor_apply(&s, ap, tmp);
        for (int64_t q = 0; q < n; q++) tp[q] = rp[q] - alpha * tmp[q];
        orc_zamux(n, ia, ja, A, tp, at);
        if (i == 1) {
            om = cdot(n, at, t) / cdot(n, at, at);
            eta = 0;
        } else {
            zc yy = cdot(n, y, y), att = cdot(n, at, t), yt = cdot(n, y, t), aty = cdot(n, at, y),
               atat = cdot(n, at, at), yat = cdot(n, y, at);
            zc den = atat * yy - yat * aty;
            om = (yy * att - yt * aty) / den;
            eta = (atat * yt - yat * att) / den;
        }
        for (int64_t q = 0; q < n; q++) {
            u[q] = om * tmp[q] + eta * (tpold[q] - rp[q] + beta * u[q]);
            z[q] = om * rp[q] + eta * z[q] - alpha * u[q];
            x[q] += alpha * p[q] + z[q];
            r[q] = t[q] - eta * y[q] - om * at[q];
        }
        err = vnorm(n, r) / babs;
        it = i;
        if (isnan(err)) { status = -3; break; }
        if (err <= eps) { status = 0; break; }
        ssor_apply(&s, r, rp);
        beta = alpha / om * cdot(n, r0, r) / cdot(n, r0, rold);
        for (int64_t q = 0; q < n; q++) w[q] = at[q] + beta * ap[q];
    }
    stats[0] = it; stats[1] = err; stats[2] = ndc; stats[3] = npcg;
    free(r); free(r0); free(rp); free(t); free(w); free(p); free(u); free(tp); free(z); free(y);
    free(ap); free(at); free(tmp); free(tpold); free(rold);
    free(s.dinv);
    if (use_dc) dc_free(&d);
    return status;
}

/* COCR (Sogabe & Zhang, J. Comput. Appl. Math. 199 (2007) 297-303) with a
 * symmetric SSOR preconditioner.  The reference's pcocr (src/solvers.f90:58-135)
 * is unreachable and broken; this restates the published algorithm on the
 * dual-volume scaled system ve*A x = ve*b, which is complex symmetric
 * (ve from orc_row_scale).  Inner products are the unconjugated bilinear form.
 * Convergence is measured on the unscaled residual ||b - A x||/||b|| so the
 * stopping rule matches the other solvers.  DC every corr_intv iterations,
 * followed by a restart from the true residual. */
int orc_cocr(int64_t n, const int64_t *ia, const int32_t *ja, const zc *A, const zc *b, zc *x,
             const double *ve, double eps, int maxit, int nx, int ny, int nz, const double *ga,
             const double *gb, const double *gc, const double *sigma, int corr_intv,
             double corr_tol, double omega, double bnorm, double *stats)
{
    int64_t nnz = ia[n] - 1;
    zc *As = malloc(nnz * sizeof(zc)), *bs = malloc(n * sizeof(zc));
    for (int64_t i = 0; i < n; i++) {
        for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) As[k] = A[k] * ve[i];
        bs[i] = b[i] * ve[i];
    }
    ssor_t s;
    if (ssor_setup(&s, n, ia, ja, As, omega)) { free(s.dinv); free(As); free(bs); return -2; }
    dc_t d = {0};
    int use_dc = corr_intv > 0;
    if (use_dc && dc_build(&d, nx, ny, nz, ga, gb, gc, sigma)) return -1;
    size_t sz = n * sizeof(zc);
    zc *r = malloc(sz), *z = malloc(sz), *Az = malloc(sz), *p = malloc(sz), *Ap = malloc(sz),
       *q = malloc(sz), *ru = malloc(sz);
    double babs = bnorm > 0 ? bnorm : vnorm(n, b);
    int it = 0, ndc = 0, npcg = 0, status = 1;
    double err = 0;
    int restart = 1;
    zc rho = 0;
    for (int i = 0; i <= maxit; i++) {
        if (restart) {
            orc_zamux(n, ia, ja, As, x, r);
            for (int64_t k = 0; k < n; k++) r[k] = bs[k] - r[k];
            ssor_apply(&s, r, z);
            orc_zamux(n, ia, ja, As, z, Az);
            memcpy(p, z, sz);
            memcpy(Ap, Az, sz);
            rho = udot(n, z, Az);
            restart = 0;
        }
        for (int64_t k = 0; k < n; k++) ru[k] = r[k] / ve[k];
        err = vnorm(n, ru) / babs;
        if (isnan(err)) { status = -3; break; }
        if (err < eps) { status = 0; break; }
        if (i == maxit) break;
        it = i + 1;
        ssor_apply(&s, Ap, q);
        zc alpha = rho / udot(n, Ap, q);
        for (int64_t k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * Ap[k]; z[k] -= alpha * q[k]; }
        if (use_dc && it % corr_intv == 0) {
            npcg += dc_correct(&d, x, corr_tol);
            ndc++;
            restart = 1;
            continue;
        }
        orc_zamux(n, ia, ja, As, z, Az);
        zc rhonew = udot(n, z, Az);
        zc beta = rhonew / rho;
        rho = rhonew;
        for (int64_t k = 0; k < n; k++) { p[k] = z[k] + beta * p[k]; Ap[k] = Az[k] + beta * Ap[k]; }
    }
    stats[0] = it; stats[1] = err; stats[2] = ndc; stats[3] = npcg;
    free(r); free(z); free(Az); free(p); free(Ap); free(q); free(ru); free(As); free(bs);
    free(s.dinv);
    if (use_dc) dc_free(&d);
    return status;
}

/* --------------------------------------------------------------------------
 * 1-D anisotropic layered background, mt1d_anl, src/model_selec.f90:266-401
 * (Pek & Santos 2002 propagation).  conv is column-major conv(nz,6); out ex has
 * 2*nz+2 entries (Ex at the nz+1 z-nodes for the xy mode, then Ey for yx),
 * hx likewise, imp[4] the bottom impedance.
 * -------------------------------------------------------------------------- */
static void mat4mul(zc out[4][4], zc A[4][4], zc B[4][4])
{
    zc t[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            zc s = 0;
            for (int k = 0; k < 4; k++) s += A[i][k] * B[k][j];
            t[i][j] = s;
        }
    memcpy(out, t, sizeof(t));
}

static void mt1d_core(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp,
                      zc *exc);

void orc_mt1d_anl(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp)
{
    mt1d_core(nz, conv, c, freq, ex, hx, imp, NULL);
}

/* The same field with the cross-polar components as well (full-tensor
 * boundary): exc[0..nz] = Ey of the xy mode, exc[nz+1..2nz+1] = Ex of the yx
 * mode. */
void orc_mt1d_anl_full(int nz, const double *conv, const double *c, double freq, zc *ex, zc *exc)
{
    zc *hx = calloc((size_t)(2 * nz + 2), sizeof(zc));
    zc imp[4];
    mt1d_core(nz, conv, c, freq, ex, hx, imp, exc);
    free(hx);
}

static void mt1d_core(int nz, const double *conv, const double *c, double freq, zc *ex, zc *hx, zc *imp,
                      zc *exc)
{
    const double w = 2 * PI * freq;
    zc (*sprod)[4][4] = calloc((size_t)(nz + 1), sizeof(zc[4][4]));
    zc m[4][4] = {{0}};
    for (int i = 0; i < 4; i++) sprod[nz][i][i] = 1.0;
    for (int i = nz; i >= 1; i--) {
        const double *cv = conv; /* conv(i, comp) = conv[(comp-1)*nz + i-1] */
#define CV(ii, cc) cv[(int64_t)((cc)-1) * nz + (ii)-1]
        double axx = CV(i, 1) - CV(i, 5) * CV(i, 5) / CV(i, 3);
        double ayy = CV(i, 2) - CV(i, 6) * CV(i, 6) / CV(i, 3);
        double axy = CV(i, 4) - CV(i, 5) * CV(i, 6) / CV(i, 3);
#undef CV
        double ada = axx + ayy, add = axx - ayy;
        double da12 = add >= 0 ? sqrt(add * add + 4 * axy * axy) : -sqrt(add * add + 4 * axy * axy);
        double a1 = (ada + da12) / 2, a2 = (ada - da12) / 2;
        zc kp = csqrt(-I * w * MU0) * csqrt(a1 + 0 * I);
        zc km = csqrt(-I * w * MU0) * csqrt(a2 + 0 * I);
        zc xip = -kp / (I * w * MU0), xim = -km / (I * w * MU0);
        double qp, qm;
        if (axy == 0) { qp = 0; qm = 0; }
        else { qp = 2 * axy / (add + da12); qm = (add - da12) / axy / 2; }
        double qdq = qp * qm, dq = 1 - qdq;
        if (i == nz) {
            m[0][1] = 1.0; m[1][1] = qp; m[2][1] = -xip * qp; m[3][1] = xip;
            m[0][3] = qm; m[1][3] = 1.0; m[2][3] = -xim; m[3][3] = xim * qm;
        }
        zc tmp = kp * c[i - 1];
        zc sp = (cexp(tmp) - cexp(-tmp)) / 2.0, cp = (cexp(tmp) + cexp(-tmp)) / 2.0;
        tmp = km * c[i - 1];
        zc sm = (cexp(tmp) - cexp(-tmp)) / 2.0, cm = (cexp(tmp) + cexp(-tmp)) / 2.0;
        zc s[4][4] = {
            {cp - qdq * cm, -qm * (cp - cm), qm * (sp / xip - sm / xim), sp / xip - qdq * sm / xim},
            {qp * (cp - cm), -qdq * cp + cm, qdq * sp / xip - sm / xim, qp * (sp / xip - sm / xim)},
            {-qp * (xip * sp - xim * sm), qdq * xip * sp - xim * sm, -qdq * cp + cm, -qp * (cp - cm)},
            {xip * sp - qdq * xim * sm, -qm * (xip * sp - xim * sm), qm * (cp - cm), cp - qdq * cm}};
        for (int r = 0; r < 4; r++)
            for (int q = 0; q < 4; q++) s[r][q] /= dq;
        mat4mul(sprod[i - 1], s, sprod[i]);
    }
    zc s_m[4][4];
    mat4mul(s_m, sprod[0], m);
    zc atmp[8];
    for (int mode = 0; mode < 2; mode++) {
        zc e0 = mode == 0 ? 1.0 : I;
        zc g11 = s_m[0][1], g21 = s_m[1][1], g12 = s_m[0][3], g22 = s_m[1][3];
        zc detg = g11 * g22 - g12 * g21;
        zc cbot[4] = {0, (g22 * creal(e0) - g12 * cimag(e0)) / detg, 0, (-g21 * creal(e0) + g11 * cimag(e0)) / detg};
        zc mc[4];
        for (int r = 0; r < 4; r++) {
            mc[r] = 0;
            for (int q = 0; q < 4; q++) mc[r] += m[r][q] * cbot[q];
        }
        for (int i = 1; i <= nz + 1; i++) {
            zc f[4];
            for (int r = 0; r < 4; r++) {
                f[r] = 0;
                for (int q = 0; q < 4; q++) f[r] += sprod[i - 1][r][q] * mc[q];
            }
            if (mode == 0) { ex[i - 1] = f[0]; hx[i - 1] = f[3]; if (exc) exc[i - 1] = f[1]; }
            else { ex[i + nz] = f[1]; hx[i + nz] = f[2]; if (exc) exc[i + nz] = f[0]; }
            if (i == nz + 1)
                for (int r = 0; r < 4; r++) atmp[mode * 4 + r] = f[r];
        }
    }
    zc tmp = atmp[2] * atmp[7] - atmp[3] * atmp[6];
    imp[0] = (atmp[0] * atmp[7] - atmp[4] * atmp[3]) / tmp;
    imp[1] = (atmp[4] * atmp[2] - atmp[0] * atmp[6]) / tmp;
    imp[2] = (atmp[1] * atmp[7] - atmp[5] * atmp[3]) / tmp;
    imp[3] = (atmp[5] * atmp[2] - atmp[1] * atmp[6]) / tmp;
    free(sprod);
}

/* --------------------------------------------------------------------------
 * EXCMG transfer, src/excmg.f90
 * -------------------------------------------------------------------------- */
/* lbf, src/excmg.f90:754-772: Lagrange basis of order l_ord on points cx. */
static double lbf(int l_ord, const double *cx, int id, double x)
{
    double num = 1, den = 1;
    for (int i = 1; i <= l_ord + 1; i++)
        if (i != id) {
            num *= x - cx[i - 1];
            den *= cx[id - 1] - cx[i - 1];
        }
    return num / den;
}

/* One block of cal_coef (src/excmg.f90:662-752) for orientation `dir`
 * (0=x,1=y,2=z), merge m, block (n1,n2,n3) 1-based, node coordinates
 * aax/bby/ccz (1-based in the formulas, stored 0-based).  coef is
 * [(2m+1)^2*2m][(m+1)^2*m]. */
static void cal_coef_dir(const double *aax, const double *bby, const double *ccz, int m, int n1,
                         int n2, int n3, int dir, double *coef)
{
    const int nout = (2 * m + 1) * (2 * m + 1) * 2 * m, nin = (m + 1) * (m + 1) * m;
    double cx1[8], cx2[8], cx3[8];
#define AX(q) aax[(q)-1]
#define BY(q) bby[(q)-1]
#define CZ(q) ccz[(q)-1]
    /* transverse node sets (every other fine node) and along-edge points: the
       midpoints of fine edges 1,3,5,.. of the block, src/excmg.f90:683-685 */
    const double *T1, *T2, *L; int o1, o2, o3;
    if (dir == 0) { T1 = bby; o1 = (n2 - 1) * 2 * m; T2 = ccz; o2 = (n3 - 1) * 2 * m; L = aax; o3 = (n1 - 1) * 2 * m; }
    else if (dir == 1) { T1 = aax; o1 = (n1 - 1) * 2 * m; T2 = ccz; o2 = (n3 - 1) * 2 * m; L = bby; o3 = (n2 - 1) * 2 * m; }
    else { T1 = aax; o1 = (n1 - 1) * 2 * m; T2 = bby; o2 = (n2 - 1) * 2 * m; L = ccz; o3 = (n3 - 1) * 2 * m; }
    for (int q = 0; q <= m; q++) { cx1[q] = T1[o1 + 2 * q]; cx2[q] = T2[o2 + 2 * q]; }
    for (int q = 0; q < m; q++) cx3[q] = (L[o3 + 2 * q] + L[o3 + 2 * q + 1]) / 2;
#undef AX
#undef BY
#undef CZ
    for (int k = 1; k <= 2 * m; k++)
        for (int j = 1; j <= 2 * m + 1; j++)
            for (int i = 1; i <= 2 * m + 1; i++) {
                int ijk = (k - 1) * (2 * m + 1) * (2 * m + 1) + (j - 1) * (2 * m + 1) + i;
                double ct1, ct2, cl;
                /* output point: transverse node i (first transverse dir), node j
                   (second), along-edge midpoint of fine edge k */
                cl = (L[o3 + k - 1] + L[o3 + k]) / 2;
                if (dir == 0) { ct1 = bby[o1 + i - 1]; ct2 = ccz[o2 + j - 1]; }
                else if (dir == 1) { ct1 = aax[o1 + i - 1]; ct2 = ccz[o2 + j - 1]; }
                else { ct1 = aax[o1 + i - 1]; ct2 = bby[o2 + j - 1]; }
                for (int kk = 1; kk <= m; kk++)
                    for (int jj = 1; jj <= m + 1; jj++)
                        for (int ii = 1; ii <= m + 1; ii++) {
                            int ind = (kk - 1) * (m + 1) * (m + 1) + (jj - 1) * (m + 1) + ii;
                            coef[(int64_t)(ijk - 1) * nin + ind - 1] =
                                lbf(m, cx1, ii, ct1) * lbf(m, cx2, jj, ct2) * lbf(m - 1, cx3, kk, cl);
                        }
            }
    (void)nout;
}

/* node coordinates as quad_interp_e/cubic_interp build them
 * (src/excmg.f90:367-378, 524-535): x,y centred, z from the top of the air. */
static void node_coords(int nx, int ny, int nz, int nair, const double *a, const double *b,
                        const double *c, double *aax, double *bby, double *ccz)
{
    double sa = 0, sb = 0, sc = 0;
    for (int i = 0; i < nx; i++) sa += a[i];
    for (int i = 0; i < ny; i++) sb += b[i];
    for (int i = 0; i < nair; i++) sc += c[i];
    aax[0] = -sa / 2; bby[0] = -sb / 2; ccz[0] = -sc;
    for (int i = 0; i < nx; i++) aax[i + 1] = aax[i] + a[i];
    for (int i = 0; i < ny; i++) bby[i + 1] = bby[i] + b[i];
    for (int i = 0; i < nz; i++) ccz[i + 1] = ccz[i] + c[i];
}

/* Generic block prolongation: fine grid (nx,ny,nz) with widths a,b,c, coarse
 * grid (nx/2,...), merge m = 2 (quad_interp_e, src/excmg.f90:357-512: blocks of
 * 4^3 fine cells) or m = 4 (cubic_interp, src/excmg.f90:514-660: blocks of
 * 8^3 fine cells). */
static int block_interp(int m, int nx, int ny, int nz, int nair, const double *a, const double *b,
                        const double *c, const zc *uc, zc *uf)
{
    const int B = 2 * m; /* fine cells per block edge */
    if (nx % B || ny % B || nz % B) return -1; /* src/excmg.f90:19-22, Q17 */
    double *aax = malloc((nx + 1) * sizeof(double)), *bby = malloc((ny + 1) * sizeof(double)),
           *ccz = malloc((nz + 1) * sizeof(double));
    node_coords(nx, ny, nz, nair, a, b, c, aax, bby, ccz);
    const int nout = (B + 1) * (B + 1) * B, nin = (m + 1) * (m + 1) * m;
    double *coef = malloc((size_t)nout * nin * sizeof(double));
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const int64_t xnlc = (int64_t)cx * (cy + 1) * (cz + 1), ynlc = (int64_t)(cx + 1) * cy * (cz + 1);
    const int64_t xnlf = n_xedge(nx, ny, nz), ynlf = n_yedge(nx, ny, nz);
    for (int kb = 1; kb <= nz / B; kb++)
        for (int jb = 1; jb <= ny / B; jb++)
            for (int ib = 1; ib <= nx / B; ib++)
                for (int dir = 0; dir < 3; dir++) {
                    cal_coef_dir(aax, bby, ccz, m, ib, jb, kb, dir, coef);
                    /* fine/coarse 0-based starting cell of the block */
                    int fi = (ib - 1) * B, fj = (jb - 1) * B, fk = (kb - 1) * B;
                    int ci = fi / 2, cj = fj / 2, ck = fk / 2;
                    for (int k = 1; k <= B; k++)          /* along-edge fine edge */
                        for (int j = 1; j <= B + 1; j++)  /* second transverse node */
                            for (int i = 1; i <= B + 1; i++) {
                                int ijk = (k - 1) * (B + 1) * (B + 1) + (j - 1) * (B + 1) + i;
                                int64_t fe;
                                if (dir == 0) fe = ((int64_t)(fk + j - 1) * (ny + 1) + (fj + i - 1)) * nx + (fi + k - 1);
                                else if (dir == 1) fe = xnlf + ((int64_t)(fk + j - 1) * ny + (fj + k - 1)) * (nx + 1) + (fi + i - 1);
                                else fe = xnlf + ynlf + ((int64_t)(fk + k - 1) * (ny + 1) + (fj + j - 1)) * (nx + 1) + (fi + i - 1);
                                zc acc = 0;
                                for (int kk = 1; kk <= m; kk++)
                                    for (int jj = 1; jj <= m + 1; jj++)
                                        for (int ii = 1; ii <= m + 1; ii++) {
                                            int ind = (kk - 1) * (m + 1) * (m + 1) + (jj - 1) * (m + 1) + ii;
                                            int64_t ce;
                                            if (dir == 0) ce = ((int64_t)(ck + jj - 1) * (cy + 1) + (cj + ii - 1)) * cx + (ci + kk - 1);
                                            else if (dir == 1) ce = xnlc + ((int64_t)(ck + jj - 1) * cy + (cj + kk - 1)) * (cx + 1) + (ci + ii - 1);
                                            else ce = xnlc + ynlc + ((int64_t)(ck + kk - 1) * (cy + 1) + (cj + jj - 1)) * (cx + 1) + (ci + ii - 1);
                                            acc += coef[(int64_t)(ijk - 1) * nin + ind - 1] * uc[ce];
                                        }
                                uf[fe] = acc;
                            }
                }
    free(coef); free(aax); free(bby); free(ccz);
    return 0;
}

int orc_quad_interp_e(int nx, int ny, int nz, int nair, const double *a, const double *b,
                      const double *c, const zc *uc, zc *uf)
{
    return block_interp(2, nx, ny, nz, nair, a, b, c, uc, uf);
}
int orc_cubic_interp(int nx, int ny, int nz, int nair, const double *a, const double *b,
                     const double *c, const zc *uc, zc *uf)
{
    return block_interp(4, nx, ny, nz, nair, a, b, c, uc, uf);
}

/* extint_v, src/excmg.f90:9-77: level i (nx,ny,nz, widths a,b,c, nair) from the
 * solutions u0 (level i-2) and u1 (level i-1):  w = (1+2^-p) u1 - 2^-p I2 u0,
 * p = 2 (src/excmg.f90:5,40-41,60-63), then I4 w. */
int orc_extint_v(int nx, int ny, int nz, int nair, const double *a, const double *b,
                 const double *c, const zc *u0, const zc *u1, zc *out)
{
    if (nx % 8 || ny % 8 || nz % 8) return -1;
    int mx = nx / 2, my = ny / 2, mz = nz / 2;
    double *a2 = malloc(mx * sizeof(double)), *b2 = malloc(my * sizeof(double)), *c2 = malloc(mz * sizeof(double));
    for (int i = 0; i < mx; i++) a2[i] = a[2 * i] + a[2 * i + 1];
    for (int i = 0; i < my; i++) b2[i] = b[2 * i] + b[2 * i + 1];
    for (int i = 0; i < mz; i++) c2[i] = c[2 * i] + c[2 * i + 1];
    int64_t n1 = orc_nedges(mx, my, mz);
    zc *ut = calloc(n1, sizeof(zc));
    int st = block_interp(2, mx, my, mz, nair / 2, a2, b2, c2, u0, ut);
    const double co1 = 1.0 + 1.0 / 4.0, co2 = 1.0 / 4.0;
    for (int64_t e = 0; e < n1; e++) ut[e] = co1 * u1[e] - co2 * ut[e];
    if (st == 0) st = block_interp(4, nx, ny, nz, nair, a, b, c, ut, out);
    free(ut); free(a2); free(b2); free(c2);
    return st;
}

/* mesh_coarsen, src/fdfe_main.f90:18-84, with the fine-cell index corrected
 * (line 64 is one layer and one row off): volume-weighted average of all 6
 * components over each 2^k x 2^k x 2^k block. */
int orc_mesh_coarsen(int nx, int ny, int nz, const double *a, const double *b, const double *c,
                     const double *sigma, int k, double *a1, double *b1, double *c1, double *sig1)
{
    int asf = 1 << k;
    if (nx % asf || ny % asf || nz % asf) return -1;
    int mx = nx / asf, my = ny / asf, mz = nz / asf;
    for (int i = 0; i < mx; i++) { a1[i] = 0; for (int t = 0; t < asf; t++) a1[i] += a[i * asf + t]; }
    for (int i = 0; i < my; i++) { b1[i] = 0; for (int t = 0; t < asf; t++) b1[i] += b[i * asf + t]; }
    for (int i = 0; i < mz; i++) { c1[i] = 0; for (int t = 0; t < asf; t++) c1[i] += c[i * asf + t]; }
    int64_t ncf = (int64_t)nx * ny * nz, ncc = (int64_t)mx * my * mz;
    for (int m = 0; m < mz; m++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                double vbig = a1[i] * b1[j] * c1[m], tmp[6] = {0, 0, 0, 0, 0, 0};
                for (int r = 0; r < asf; r++)
                    for (int s = 0; s < asf; s++)
                        for (int t = 0; t < asf; t++) {
                            int fi = i * asf + t, fj = j * asf + s, fk = m * asf + r;
                            int64_t ind = ((int64_t)fk * ny + fj) * nx + fi;
                            double vs = a[fi] * b[fj] * c[fk];
                            for (int q = 0; q < 6; q++) tmp[q] += sigma[q * ncf + ind] * vs;
                        }
                int64_t ic = ((int64_t)m * my + j) * mx + i;
                for (int q = 0; q < 6; q++) sig1[q * ncc + ic] = tmp[q] / vbig;
            }
    return 0;
}
