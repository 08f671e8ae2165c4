// rheogrid.cpp -- host-side init of the rheology_file and mask_file branches
// of isos_init (fastisostasy.f90:199-213, 262-290, 336-411): gridded inputs
// interpolated onto the padded domain, optionally smoothed.  The results enter
// the device init as the field inputs (lithosphere / mantle = fields).
#include <cmath>
#include <vector>

#include "../../include/fastiso.h"
#include "fi_kernels.h"

namespace fi {
namespace {

// bilinear_interpolation (isos_utils.f90:495-558) at one target point: within
// 1 m of a grid end (or beyond it) in either direction the value of the
// nearest edge cell is returned without interpolating in the other direction
double bilinear_at(int nxr, int nyr, const double* x, const double* y, const double* z, double xo, double yo) {
    constexpr double tol = 1.0;
    bool flat = false;
    int i = 0, j = 0;
    double tx = 0.0, ty = 0.0;
    auto locate = [&](const double* c, int nc, double v, int& k, double& t) {
        if (v <= c[0] + tol) { k = 0; flat = true; return; }
        if (v >= c[nc - 1] - tol) { k = nc - 1; flat = true; return; }
        for (k = 0; k < nc - 1; ++k)
            if (v > c[k] && v <= c[k + 1]) break;
        t = (v - c[k]) / (c[k + 1] - c[k]);
    };
    locate(x, nxr, xo, i, tx);
    locate(y, nyr, yo, j, ty);
    auto Z = [&](int a, int b) { return z[(size_t)b * nxr + a]; };
    if (flat) return Z(i, j);
    return Z(i, j) * (1.0 - tx) * (1.0 - ty) + Z(i + 1, j) * tx * (1.0 - ty) + Z(i, j + 1) * (1.0 - tx) * ty +
           Z(i + 1, j + 1) * tx * ty;
}

// bilinear_interpolation_array onto the n x n domain with coordinates xc (both axes)
void bilinear_domain(int nxr, int nyr, const double* x, const double* y, const double* z, int n,
                     const std::vector<double>& xc, double* out) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) out[(size_t)j * n + i] = bilinear_at(nxr, nyr, x, y, z, xc[i], xc[j]);
}

// smooth_gauss_2D (isos_utils.f90:696-817) with the kernel of gauss_values
// (:918-957): window half-width ceiling(2 sigma / dx), weights normalised (the
// reference normalises twice), field mirrored about its edges
void smooth_gauss(int n, double* var, double dx, double sigma) {
    const int h = (int)std::ceil(2.0 * sigma / dx), w = 2 * h + 1;
    const double pi_ref = (double)3.14159265359f;
    std::vector<double> f((size_t)w * w);
    for (int b = -h; b <= h; ++b)
        for (int a = -h; a <= h; ++a) {
            const double x = a * dx, y = b * dx;
            f[(size_t)(b + h) * w + (a + h)] =
                1.0 / (2.0 * pi_ref * (sigma * sigma)) * std::exp(-(x * x + y * y) / (2 * (sigma * sigma)));
        }
    for (int pass = 0; pass < 2; ++pass) {
        double sum = 0.0;
        for (double v : f) sum += v;
        for (double& v : f) v = v / sum;
    }
    const int m = n + 2 * h;
    std::vector<double> e((size_t)m * m);
    auto mirror = [&](int k) { return k < 0 ? -k - 1 : (k >= n ? 2 * n - 1 - k : k); };
    for (int j = -h; j < n + h; ++j)
        for (int i = -h; i < n + h; ++i) e[(size_t)(j + h) * m + (i + h)] = var[(size_t)mirror(j) * n + mirror(i)];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int b = 0; b < w; ++b)
                for (int a = 0; a < w; ++a) acc += e[(size_t)(j + b) * m + (i + a)] * f[(size_t)b * w + a];
            var[(size_t)j * n + i] = acc;
        }
}

}  // namespace

void grid_smooth(int n, double* var, double dx, double sigma) { smooth_gauss(n, var, dx, sigma); }

int grid_rheology(const fi_params& P, const fi_rheology_grid& g, int n, double dx, std::vector<double>& he,
                  std::vector<double>& eta) {
    const size_t nn = (size_t)n * n;
    std::vector<double> xc(n);
    for (int i = 0; i < n; ++i) xc[i] = (double)(i - n / 2) * dx;   // init_dims (isos_utils.f90:221-240)
    const size_t nr = (size_t)g.nx * g.ny;
    if (P.lithosphere == 4) {
        std::vector<double> T(nr);
        for (size_t q = 0; q < nr; ++q) T[q] = g.litho_thickness[q] * g.litho_thickness_scaling;
        he.resize(nn);
        bilinear_domain(g.nx, g.ny, g.xc, g.yc, T.data(), n, xc, he.data());
        if (g.rheo_smoothing_radius > 0.0) smooth_gauss(n, he.data(), dx, g.rheo_smoothing_radius);
    }
    if (P.mantle == 4) {
        const int nz = g.nz, nl = P.nl;
        // layer boundaries (lv_elva.f90:131-168) from the interpolated He_lith
        std::vector<double> bnd(nn * nl);
        for (size_t c = 0; c < nn; ++c) bnd[c] = he[c];
        for (int k = 1; k < nl; ++k)
            for (size_t c = 0; c < nn; ++c)
                bnd[k * nn + c] = (P.layering == 0) ? bnd[(k - 1) * nn + c] + P.dl : P.zl[k];
        std::vector<double> depth(nz);
        for (int i = 0; i < nz; ++i) depth[i] = g.zc[nz - 1] - g.zc[nz - 1 - i];
        std::vector<double> bil(nn * nz), tmp(nn);
        for (int i = 0; i < nz; ++i) {
            double* lev = bil.data() + (size_t)i * nn;
            bilinear_domain(g.nx, g.ny, g.xc, g.yc, g.log10_eta + (size_t)(nz - 1 - i) * nr, n, xc, lev);
            if (P.viscosity_scaling_method == 4 && g.stddev_log10_eta) {
                bilinear_domain(g.nx, g.ny, g.xc, g.yc, g.stddev_log10_eta + (size_t)(nz - 1 - i) * nr, n, xc,
                                tmp.data());
                for (size_t c = 0; c < nn; ++c) lev[c] = lev[c] + P.viscosity_scaling * tmp[c];
            }
        }
        for (double& v : bil) v = v < 16 ? 16 : (v > 23 ? 23 : v);
        // interp_0d along depth at each boundary (isos_utils.f90:423-457), then 10**
        eta.resize(nn * nl);
        std::vector<double> col(nz);
        for (int k = 0; k < nl; ++k)
            for (size_t c = 0; c < nn; ++c) {
                for (int i = 0; i < nz; ++i) col[i] = bil[(size_t)i * nn + c];
                const double xo = bnd[k * nn + c];
                double v;
                if (xo < depth[0]) v = col[0];
                else if (xo > depth[nz - 1]) v = col[nz - 1];
                else {
                    int j = 0;
                    while (j < nz && !(depth[j] >= xo)) ++j;
                    if (j == 0) v = col[0];
                    else if (j == nz) v = col[nz - 1];
                    else v = col[j - 1] + (xo - depth[j - 1]) / (depth[j] - depth[j - 1]) * (col[j] - col[j - 1]);
                }
                eta[k * nn + c] = std::pow(10.0, v);
            }
    }
    return 0;
}

void grid_mask(const fi_rheology_grid& g, int n, double dx, std::vector<uint8_t>& active) {
    const size_t nn = (size_t)n * n;
    std::vector<double> xc(n), m(nn);
    for (int i = 0; i < n; ++i) xc[i] = (double)(i - n / 2) * dx;
    bilinear_domain(g.mask_nx, g.mask_ny, g.mask_xc, g.mask_yc, g.mask, n, xc, m.data());
    active.resize(nn);
    for (size_t c = 0; c < nn; ++c) active[c] = m[c] > 0.5 ? 1 : 0;
}

}  // namespace fi
