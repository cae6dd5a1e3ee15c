// sfdx_host.h -- host-side FP64 helpers (1-D background, coarsening, EXCMG
// weight tables, coarse direct solve, post-process).
#pragma once
#include <array>
#include <complex>
#include <string>
#include <vector>

#include "sfdx_stencil.h"

namespace sfdx {

extern const double kMu0;

struct HostLevel {
    int nx = 0, ny = 0, nz = 0, nair = 0;
    std::vector<double> a, b, c;  // widths
    std::vector<double> sig;      // 6*ncell, sig[comp*ncell + cell] (empty when ext is set)
    const double *ext = nullptr;  // the caller's sigma in the same layout, not copied
    size_t ncell() const { return (size_t)nx * ny * nz; }
    const double *S() const { return ext ? ext : sig.data(); }
};

struct ProlongHost {
    std::vector<int> nb[3], eb[3];
    std::vector<double> nw[3], ew[3];
};

struct CoarseSystem {
    int n = 0;
    std::vector<int64_t> inner;  // interior edge ids
    std::vector<int> imap;       // edge -> interior row or -1
    std::vector<int> ia, ja;     // upper-triangular CSR, 0-based
    std::vector<std::complex<double>> val;
    std::vector<std::complex<double>> rhs;  // 2*n: xy then yx
};

struct PostGrid {
    int nx, ny, nz, nair;
    const double *a, *b, *c;
    // surface node plane of each cell column (index i + nx*j, 0-based plane
    // counted from the top of the air); null = the flat datum, plane nair
    const int *surf = nullptr;
    int plane(int i, int j) const { return surf ? surf[(size_t)i + (size_t)nx * j] : nair; }
};

void mt1d_anl(int nz, const double *conv, const double *c, double freq, std::complex<double> *ex,
              std::complex<double> *hx, std::complex<double> *imp, std::complex<double> *exc = nullptr);
std::vector<std::complex<double>> background_field(const HostLevel &L, double freq, bool full = false);
HostLevel coarsen(const HostLevel &F, int k);
// Surface node plane of every cell column of a model whose topography is carved
// into sigma: the top node plane of the first cell, counted from the top, whose
// sigma_xx exceeds air_max (S/m); nz if the column is all air.  surf[i + nx*j].
void surface_planes(const HostLevel &L, double air_max, std::vector<int> &surf);
void prolong_tables(const HostLevel &fine, int m, ProlongHost &T);
void coarse_system(const HostLevel &L, double freq, const std::vector<std::complex<double>> &edc, CoarseSystem &S);
bool direct_solve(const CoarseSystem &S, std::vector<std::complex<double>> &sol, std::string &err);
// threads every PARDISO call runs on (0 if MKL is unavailable)
int direct_solve_threads();
int post_process(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv,
                 const std::complex<double> *ex, int out_type, int rd, double *results, double *tipper);
// post_process from surface fields already evaluated (on the device):
// surf[((f*2 + pol)*5 + q)*nx*ny + i + nx*j], q = ex, ey, hx, hy, hz
int post_process_surface(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv,
                         const std::complex<double> *surf, int out_type, int rd, double *results, double *tipper);
// Misfit 1/2 sum over frequencies, stations and the xy/yx modes of
// (log10 rho_a - log10 rho_obs)^2 + (phi - phi_obs)^2 (phase in radians), obs
// given as post_process out_type 2 rows (rho_xy, rho_yx, phi_xy, phi_yx in
// ohm-m and degrees).  If w is non-null it receives dPhi = Re(w^T dex) for the
// fields ex (same layout as ex).  Returns a non-zero station number on error.
int misfit_fields(const PostGrid &g, int nf, const double *freq, int nrec, const double *recv, int rd,
                  const double *obs, const std::complex<double> *ex, double &misfit, std::complex<double> *w);

}  // namespace sfdx
