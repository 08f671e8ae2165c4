/*
 * fi_oracle.c -- CPU FP64 oracle for the FastIsostasy LV-ELVA time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  The product
 * library never links or calls this file.
 *
 * A plain-C restatement of the reference Fortran hot path with wp = dp
 * (precision.f90:10 switched to dp).  Every function cites the reference
 * file:line it follows.  Quirks of the reference are reproduced on purpose:
 *   - single-precision literals in FP64 mode (pi, 9.81, 1.e-12, GE table),
 *   - separable 2-D Hartley transform (FFTW_DHT x FFTW_DHT),
 *   - kappa period (n-1)*dx and kappa(1,1) averaging,
 *   - convolution crop starting at n/2 (+1 cell shift), corner BC on the
 *     full linear convolution,
 *   - degenerate tsit54 (all six stages identical, scalar tol),
 *   - lagged canom_mantle, diagnostics cadence, init call without stepping.
 *
 * The reference's third-party arithmetic (FFTW3, version unpinned) is
 * replaced by the mixed-radix Stockham FFT below; the linear convolution
 * uses a 2n-point transform instead of FFTW's (2n-1)-point one, which gives
 * the same linear convolution (no wrap for any size >= 2n-1).
 */
#include "fi_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

typedef double complex cplx;

static char g_err[512];
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *or_last_error(void) { return g_err; }

/* single-precision literals of the reference, promoted to double */
#define F32(x) ((double)(float)(x))
static const double PI_REF = (double)3.14159265359f;     /* isostasy_defs.f90:8 */

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, forward (exp(-2 pi i nk/N)), unnormalised.
 * Replaces FFTW (isos_utils.f90:143-217); checked against numpy.fft in tests. */
typedef struct {
    int n, nf;
    int fac[64];
    cplx *tw;       /* tw[m] = exp(-2 pi i m / n) */
} fftplan;

static fftplan *plan_create(int n) {
    fftplan *p = (fftplan *)calloc(1, sizeof *p);
    p->n = n;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
        if (f * f > m && m > 1) { p->fac[p->nf++] = m; m = 1; }
    }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[k] = cos(a) + I * sin(a);
    }
    return p;
}
static void plan_destroy(fftplan *p) { if (p) { free(p->tw); free(p); } }

/* in-place forward FFT of x (work has n entries) */
static void fft_exec(const fftplan *p, cplx *x, cplx *work) {
    const int n = p->n;
    cplx *a = x, *b = work;
    int Ns = 1;
    for (int f = 0; f < p->nf; f++) {
        const int R = p->fac[f];
        const int m = n / R;
        const int tstep = n / (Ns * R);
        cplx v[R], u[R];
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            for (int r = 0; r < R; r++) v[r] = a[j + r * m] * p->tw[k * r * tstep];
            if (R == 2) {
                u[0] = v[0] + v[1]; u[1] = v[0] - v[1];
            } else if (R == 4) {
                cplx s0 = v[0] + v[2], s1 = v[0] - v[2];
                cplx s2 = v[1] + v[3], s3 = (v[1] - v[3]) * (-I);
                u[0] = s0 + s2; u[2] = s0 - s2; u[1] = s1 + s3; u[3] = s1 - s3;
            } else {
                for (int q = 0; q < R; q++) {
                    cplx acc = 0;
                    for (int r = 0; r < R; r++) acc += v[r] * p->tw[((r * q) % R) * m];
                    u[q] = acc;
                }
            }
            const int idxD = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; q++) b[idxD + q * Ns] = u[q];
        }
        cplx *t = a; a = b; b = t;
        Ns *= R;
    }
    if (a != x) memcpy(x, a, sizeof(cplx) * (size_t)n);
}

void or_fft(int n, const double *re_in, const double *im_in, double *re_out, double *im_out) {
    fftplan *p = plan_create(n);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n), *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int i = 0; i < n; i++) x[i] = re_in[i] + I * (im_in ? im_in[i] : 0.0);
    fft_exec(p, x, w);
    for (int i = 0; i < n; i++) { re_out[i] = creal(x[i]); im_out[i] = cimag(x[i]); }
    free(x); free(w); plan_destroy(p);
}

/* ------------------------------------------------------------------------ */
/* Separable 2-D DHT (FFTW r2r_2d with FFTW_DHT x FFTW_DHT, fastisostasy.f90:140-144,
 * executed at isos_utils.f90:151).  H[k] = Re F[k] - Im F[k] per dimension.
 * Two real lines are packed into one complex FFT. */
static void dht_lines(const fftplan *p, double *a, double *b, cplx *z, cplx *w) {
    const int n = p->n;
    for (int i = 0; i < n; i++) z[i] = a[i] + I * (b ? b[i] : 0.0);
    fft_exec(p, z, w);
    for (int k = 0; k < n; k++) {
        const int nk = (n - k) % n;
        const double zr = creal(z[k]), zi = cimag(z[k]);
        const double yr = creal(z[nk]), yi = cimag(z[nk]);
        a[k] = 0.5 * ((zr + yr) - (zi - yi));
        if (b) b[k] = 0.5 * ((zi + yi) + (zr - yr));
    }
}

static void dht2_plan(const fftplan *p, double *x, double *tmp2n, cplx *z, cplx *w) {
    const int n = p->n;
    for (int j = 0; j + 1 < n; j += 2) dht_lines(p, x + (size_t)j * n, x + (size_t)(j + 1) * n, z, w);
    if (n % 2) dht_lines(p, x + (size_t)(n - 1) * n, NULL, z, w);
    double *ca = tmp2n, *cb = tmp2n + n;
    for (int i = 0; i < n; i += 2) {
        const int two = (i + 1 < n);
        for (int j = 0; j < n; j++) { ca[j] = x[(size_t)j * n + i]; if (two) cb[j] = x[(size_t)j * n + i + 1]; }
        dht_lines(p, ca, two ? cb : NULL, z, w);
        for (int j = 0; j < n; j++) { x[(size_t)j * n + i] = ca[j]; if (two) x[(size_t)j * n + i + 1] = cb[j]; }
    }
}

void or_dht2(int n, const double *in, double *out) {
    fftplan *p = plan_create(n);
    cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)n), *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(out, in, sizeof(double) * (size_t)n * n);
    dht2_plan(p, out, t, z, w);
    free(z); free(w); free(t); plan_destroy(p);
}

/* 2-D complex FFT (rows then columns) of an m x m array, forward or inverse
 * (inverse unnormalised, via conjugation). */
static void fft2_plan(const fftplan *p, cplx *x, int inverse, cplx *col, cplx *w) {
    const int m = p->n;
    if (inverse) for (size_t i = 0; i < (size_t)m * m; i++) x[i] = conj(x[i]);
    for (int j = 0; j < m; j++) fft_exec(p, x + (size_t)j * m, w);
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < m; j++) col[j] = x[(size_t)j * m + i];
        fft_exec(p, col, w);
        for (int j = 0; j < m; j++) x[(size_t)j * m + i] = col[j];
    }
    if (inverse) for (size_t i = 0; i < (size_t)m * m; i++) x[i] = conj(x[i]);
}

/* ------------------------------------------------------------------------ */
/* Finite differences, finite_differences.f90:17-120 (verbatim edge formulas,
 * including uxx(nx,:) dividing by dx(1,:), :91). Arrays [ny][nx]. */
#define IX(i, j) ((size_t)(j) * nx + (i))
void or_deriv_x(int nx, int ny, const double *u, const double *dx, double *d) {
    for (int j = 0; j < ny; j++) {
        for (int i = 1; i < nx - 1; i++) d[IX(i, j)] = (u[IX(i + 1, j)] - u[IX(i - 1, j)]) / (2 * dx[IX(i, j)]);
        d[IX(nx - 1, j)] = (u[IX(nx - 1, j)] - u[IX(nx - 2, j)]) / dx[IX(nx - 1, j)];
        d[IX(0, j)] = (u[IX(1, j)] - u[IX(0, j)]) / dx[IX(0, j)];
    }
}
void or_deriv_y(int nx, int ny, const double *u, const double *dy, double *d) {
    for (int j = 1; j < ny - 1; j++)
        for (int i = 0; i < nx; i++) d[IX(i, j)] = (u[IX(i, j + 1)] - u[IX(i, j - 1)]) / (2 * dy[IX(i, j)]);
    for (int i = 0; i < nx; i++) {
        d[IX(i, ny - 1)] = (u[IX(i, ny - 1)] - u[IX(i, ny - 2)]) / dy[IX(i, ny - 1)];
        d[IX(i, 0)] = (u[IX(i, 1)] - u[IX(i, 0)]) / dy[IX(i, 0)];
    }
}
void or_deriv_xx(int nx, int ny, const double *u, const double *dx, double *d) {
    for (int j = 0; j < ny; j++) {
        for (int i = 1; i < nx - 1; i++)
            d[IX(i, j)] = (u[IX(i + 1, j)] - 2 * u[IX(i, j)] + u[IX(i - 1, j)]) / (dx[IX(i, j)] * dx[IX(i, j)]);
        d[IX(0, j)] = (u[IX(2, j)] - 2 * u[IX(1, j)] + u[IX(0, j)]) / (dx[IX(0, j)] * dx[IX(0, j)]);
        d[IX(nx - 1, j)] = (u[IX(nx - 1, j)] - 2 * u[IX(nx - 2, j)] + u[IX(nx - 3, j)]) / (dx[IX(0, j)] * dx[IX(0, j)]);
    }
}
void or_deriv_yy(int nx, int ny, const double *u, const double *dy, double *d) {
    for (int j = 1; j < ny - 1; j++)
        for (int i = 0; i < nx; i++)
            d[IX(i, j)] = (u[IX(i, j + 1)] - 2 * u[IX(i, j)] + u[IX(i, j - 1)]) / (dy[IX(i, j)] * dy[IX(i, j)]);
    for (int i = 0; i < nx; i++) {
        d[IX(i, 0)] = (u[IX(i, 2)] - 2 * u[IX(i, 1)] + u[IX(i, 0)]) / (dy[IX(i, 0)] * dy[IX(i, 0)]);
        d[IX(i, ny - 1)] = (u[IX(i, ny - 1)] - 2 * u[IX(i, ny - 2)] + u[IX(i, ny - 3)]) / (dy[IX(i, ny - 1)] * dy[IX(i, ny - 1)]);
    }
}
#undef IX

/* ------------------------------------------------------------------------ */
/* calc_kappa, lv_elva.f90:442-481 */
void or_kappa(int n, double dx, double *kappa) {
    const double mu = 2.0 * PI_REF / ((n - 1) * dx);
    const int ic = (n - 1) / 2 + 1;
    for (int j = 0; j < n; j++) {
        const int jj = j + 1;
        const double q = (jj <= ic) ? mu * (jj - 1) : mu * (n - jj + 1);
        for (int i = 0; i < n; i++) {
            const int ii = i + 1;
            const double p = (ii <= ic) ? mu * (ii - 1) : mu * (n - ii + 1);
            kappa[(size_t)j * n + i] = sqrt(p * p + q * q);
        }
    }
    kappa[0] = (kappa[(size_t)1 * n + 0] + kappa[1]) / 2.0;
}

/* calc_effective_viscosity + calc_viscosity_scaling_freqdomain, lv_elva.f90:207-306.
 * eta, bnd: [nl][n][n].  R must hold its incoming value (1 after isos_init). */
void or_lumping(int n, int nl, int lumping, const double *eta, const double *bnd,
                const double *kappa, double *eta_eff, double *R) {
    const size_t nn = (size_t)n * n;
    for (size_t c = 0; c < nn; c++) {
        double e;
        switch (lumping) {
        case 0: e = eta[(size_t)(nl - 1) * nn + c]; break;
        case 1: e = eta[c]; for (int k = 1; k < nl; k++) e = fmin(e, eta[k * nn + c]); break;
        case 2: e = eta[c]; for (int k = 1; k < nl; k++) e = fmax(e, eta[k * nn + c]); break;
        case 3: e = 0; for (int k = 0; k < nl; k++) e += eta[k * nn + c]; e = e / (double)nl; break;
        case 4: e = 0; for (int k = 0; k < nl; k++) e += log10(eta[k * nn + c]); e = pow(10.0, e / (double)nl); break;
        default: {
            e = eta[(size_t)(nl - 1) * nn + c];
            double r = R[c];
            for (int k = nl - 1; k >= 1; k--) {
                const double dz = bnd[k * nn + c] - bnd[(k - 1) * nn + c];
                const double eta_c = eta[(size_t)(k - 1) * nn + c];
                const double ratio = eta_c / e;
                const double inv = 1. / ratio;
                const double kap = kappa[c];
                const double ch = cosh(dz * kap), sh = sinh(dz * kap);
                const double dzk = dz * kap;
                const double num = 2 * ratio * ch * sh + (1 - ratio * ratio) * (dzk * dzk) + (ratio * sh) * (ratio * sh) + ch * ch;
                const double den = (ratio + inv) * ch * sh + (ratio - inv) * dz * kap + sh * sh + ch * ch;
                r = r * num / den;
            }
            R[c] = r;
        }
        }
        eta_eff[c] = e;
    }
}

/* calc_elastic_green + get_ge_value, green_functions.f90:43-91,131-168.
 * Tables are default-real literals in the reference, so they are rounded
 * through float even when wp = dp. */
static const float RN_F[42] = {0.0f, 0.011f, 0.111f, 1.112f, 2.224f, 3.336f, 4.448f, 6.672f, 8.896f, 11.12f,
    17.79f, 22.24f, 27.80f, 33.36f, 44.48f, 55.60f, 66.72f, 88.96f, 111.2f, 133.4f, 177.9f, 222.4f, 278.0f,
    333.6f, 444.8f, 556.0f, 667.2f, 778.4f, 889.6f, 1001.0f, 1112.0f, 1334.0f, 1779.0f, 2224.0f, 2780.0f,
    3336.0f, 4448.0f, 5560.0f, 6672.0f, 7784.0f, 8896.0f, 10008.0f};
static const float GE_F[42] = {-33.6488f, -33.64f, -33.56f, -32.75f, -31.86f, -30.98f, -30.12f, -28.44f,
    -26.87f, -25.41f, -21.80f, -20.02f, -18.36f, -17.18f, -15.71f, -14.91f, -14.41f, -13.69f, -13.01f,
    -12.31f, -10.95f, -9.757f, -8.519f, -7.533f, -6.131f, -5.237f, -4.660f, -4.272f, -3.999f, -3.798f,
    -3.640f, -3.392f, -2.999f, -2.619f, -2.103f, -1.530f, -0.292f, 0.848f, 1.676f, 2.083f, 2.057f, 1.643f};

static double ge_value(double r) {
    double rn[42], ge[42];
    for (int k = 0; k < 42; k++) { rn[k] = (double)(RN_F[k] * 1.e3f); ge[k] = (double)GE_F[k]; }
    if (r > rn[41]) return 0.;
    int k;
    for (k = 0; k < 41; k++) if (r >= rn[k] && r < rn[k + 1]) break;
    if (k == 41) k = 40;   /* r == rn(n): the reference would index past the table */
    return ge[k] + (r - rn[k]) / (rn[k + 1] - rn[k]) * (ge[k + 1] - ge[k]);
}

static int midindex(int n) { return (n % 2 == 0) ? n / 2 : (n - 1) / 2; }

void or_green_elastic(int n, double dx, double *gk) {
    const double dy = dx;
    const int mx = midindex(n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double x = (i - mx) * dx, y = (j - mx) * dy;
            const double r = sqrt(x * x + y * y);
            gk[(size_t)j * n + i] = ge_value(r) * F32(1.e-12) * (dx * dy) / (F32(9.81) * fmax(r, dx));
        }
}

/* calc_z_ss_green + calc_gn_value, green_functions.f90:94-128,171-181 */
void or_green_ssh(int n, double dx, double r_earth, double m_earth, double *gk) {
    const double dy = dx;
    const int mx = midindex(n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double x = (i - mx) * dx, y = (j - mx) * dy;
            const double r = fmax(dy / 2, sqrt(x * x + y * y));
            gk[(size_t)j * n + i] = r_earth / (2. * m_earth * sin(r / (2. * r_earth)));
        }
}

/* ------------------------------------------------------------------------ */
/* Kelvin function of order zero, Abramowitz & Stegun 9.9.10-9.9.12:
 *   ber x = sum_k (-1)^k (x^2/4)^(2k) / ((2k)!)^2
 *   bei x = sum_k (-1)^k (x^2/4)^(2k+1) / ((2k+1)!)^2
 *   kei x = -ln(x/2) bei x - (pi/4) ber x + sum_k (-1)^k psi(2k+2) (x^2/4)^(2k+1) / ((2k+1)!)^2
 * with psi(m+1) = -gamma + H_m.  Summed in long double: at x = 10 the terms
 * reach 1e3 for a result of 3e-4, which leaves ~13 correct digits. */
double or_kei(double xd) {
    const long double pi = 3.14159265358979323846264338327950288L;
    const long double euler_gamma = 0.57721566490153286060651209008240243L;
    if (xd == 0.0) return (double)(-pi / 4);
    const long double x = xd, q = x * x / 4;
    long double t = 1.0L, H = 0.0L, ber = 0.0L, bei = 0.0L, ser = 0.0L;
    for (int m = 0; m < 80; m++) {
        if (m > 0) { t = t * q / ((long double)m * (long double)m); H += 1.0L / m; }
        const long double sg = ((m / 2) % 2) ? -1.0L : 1.0L;
        if (m % 2 == 0) ber += sg * t;
        else { bei += sg * t; ser += sg * (H - euler_gamma) * t; }
    }
    return (double)(-logl(x / 2) * bei - pi / 4 * ber + ser);
}

/* load_kei_values (kelvin_function.f90:167-1190) tabulates kei at
 * x = 0.00, 0.01, ..., 10.00 to five significant digits as default-real
 * literals; regenerated here from or_kei with the same decimal rounding and
 * the same conversion to single precision. */
void or_kei_table(double *rn, double *kei) {
    char buf[64];
    for (int k = 0; k <= 1000; k++) {
        snprintf(buf, sizeof buf, "%.4f", k / 100.0);
        rn[k] = (double)strtof(buf, NULL);
        snprintf(buf, sizeof buf, "%.4e", or_kei(k / 100.0));
        kei[k] = (double)strtof(buf, NULL);
    }
}

/* get_kei_value, kelvin_function.f90:96-131.  A radius of exactly 10 L_w
 * makes the reference's search run past the table (k = n); the last entry is
 * returned there. */
static double kei_value(double r, double L_w, const double *rn, const double *kv) {
    const int n = 1001;
    const double x = r / L_w;
    if (x > rn[n - 1]) return kv[n - 1];
    int k;
    for (k = 0; k < n - 1; k++)
        if (x >= rn[k] && x < rn[k + 1]) break;
    if (k == n - 1) return kv[n - 1];
    return kv[k] + (x - rn[k]) / (rn[k + 1] - rn[k]) * (kv[k + 1] - kv[k]);
}

/* calc_kei_filter_2D (kelvin_function.f90:22-94) on the domain coordinates
 * of init_dims (isos_utils.f90:221-240), then calc_viscous_green
 * (green_functions.f90:28-41): GV = -L_w^2 / (2 pi D) * kei * (dx dy) */
void or_green_viscous(int n, double dx, double L_w, double D_lith, double *gv) {
    double rn[1001], kv[1001];
    or_kei_table(rn, kv);
    const double c = L_w * L_w / (2.0 * PI_REF * D_lith);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double x = (double)(i - n / 2) * dx, y = (double)(j - n / 2) * dx;
            const double kei = kei_value(sqrt(x * x + y * y), L_w, rn, kv);
            gv[(size_t)j * n + i] = -(c * kei * (dx * dx));
        }
}

/* ------------------------------------------------------------------------ */
/* precompute_kernel / precomputed_fftconvolution, convolutions.f90:29-108 */
typedef struct {
    int n, m;
    fftplan *p;
    cplx *buf, *col, *w;
} convctx;

static convctx *conv_create(int n) {
    convctx *c = (convctx *)calloc(1, sizeof *c);
    c->n = n; c->m = 2 * n;
    c->p = plan_create(c->m);
    c->buf = (cplx *)malloc(sizeof(cplx) * (size_t)c->m * c->m);
    c->col = (cplx *)malloc(sizeof(cplx) * (size_t)c->m);
    c->w = (cplx *)malloc(sizeof(cplx) * (size_t)c->m);
    return c;
}
static void conv_destroy(convctx *c) { if (c) { plan_destroy(c->p); free(c->buf); free(c->col); free(c->w); free(c); } }

static void conv_spectrum(convctx *c, const double *kernel, cplx *spec) {
    const int n = c->n, m = c->m;
    memset(spec, 0, sizeof(cplx) * (size_t)m * m);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) spec[(size_t)j * m + i] = kernel[(size_t)j * n + i];
    fft2_plan(c->p, spec, 0, c->col, c->w);
}

static void conv_apply(convctx *c, const cplx *spec, const double *in, double *out) {
    const int n = c->n, m = c->m;
    cplx *b = c->buf;
    memset(b, 0, sizeof(cplx) * (size_t)m * m);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) b[(size_t)j * m + i] = in[(size_t)j * n + i];
    fft2_plan(c->p, b, 0, c->col, c->w);
    for (size_t q = 0; q < (size_t)m * m; q++) b[q] *= spec[q];
    fft2_plan(c->p, b, 1, c->col, c->w);
    const double scale = (double)m * (double)m;
    /* the reference's full array is (2n-1)^2: its corners sit at 0 and 2n-2 */
    const int e = 2 * n - 2;
    const double c11 = creal(b[0]) / scale, cee = creal(b[(size_t)e * m + e]) / scale;
    const double c1e = creal(b[(size_t)e * m + 0]) / scale, ce1 = creal(b[e]) / scale;
    const double corr = 0.25 * (c11 + cee + c1e + ce1);
    const int i1 = (n % 2 == 0) ? n / 2 : n / 2 + 1;   /* 1-based, convolutions.f90:34-39 */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            out[(size_t)j * n + i] = creal(b[(size_t)(j + i1 - 1) * m + (i + i1 - 1)]) / scale - corr;
}

void or_convolve(int n, const double *kernel, const double *in, double *out) {
    convctx *c = conv_create(n);
    cplx *spec = (cplx *)malloc(sizeof(cplx) * (size_t)c->m * c->m);
    conv_spectrum(c, kernel, spec);
    conv_apply(c, spec, in, out);
    free(spec); conv_destroy(c);
}

/* ------------------------------------------------------------------------ */
/* apply_zerobc_at_corners, isos_utils.f90:104-118 (x(1,1)+x(nx,ny)+x(1,ny)+x(nx,1)) */
static void zerobc(int n, double *x) {
    const double c = 0.25 * (x[0] + x[(size_t)(n - 1) * n + (n - 1)] + x[(size_t)(n - 1) * n] + x[n - 1]);
    for (size_t q = 0; q < (size_t)n * n; q++) x[q] = x[q] - c;
}

/* calc_lvelva, lv_elva.f90:31-126 (with dx_matrix == dy_matrix since dx == dy) */
typedef struct {
    int n;
    fftplan *p;
    double *w_x, *w_xy, *w_xx, *w_yy, *Mxx, *Myy, *Mxy, *Mxy_x, *Mxy_xy, *Mx_xx, *My_yy, *f, *tmp2n;
    cplx *z, *wk;
} rhsctx;

static rhsctx *rhs_create(int n) {
    rhsctx *r = (rhsctx *)calloc(1, sizeof *r);
    const size_t nn = (size_t)n * n;
    r->n = n; r->p = plan_create(n);
    double **bufs[] = {&r->w_x, &r->w_xy, &r->w_xx, &r->w_yy, &r->Mxx, &r->Myy, &r->Mxy, &r->Mxy_x,
                       &r->Mxy_xy, &r->Mx_xx, &r->My_yy, &r->f};
    for (size_t b = 0; b < sizeof bufs / sizeof bufs[0]; b++) *bufs[b] = (double *)calloc(nn, sizeof(double));
    r->tmp2n = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    r->z = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    r->wk = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    return r;
}
static void rhs_destroy(rhsctx *r) {
    if (!r) return;
    double *bufs[] = {r->w_x, r->w_xy, r->w_xx, r->w_yy, r->Mxx, r->Myy, r->Mxy, r->Mxy_x, r->Mxy_xy, r->Mx_xx, r->My_yy, r->f};
    for (size_t b = 0; b < sizeof bufs / sizeof bufs[0]; b++) free(bufs[b]);
    free(r->tmp2n); free(r->z); free(r->wk); plan_destroy(r->p); free(r);
}

static void rhs_eval(rhsctx *r, const double *w, const double *canom_full, const double *maskactive,
                     double g, double nu, const double *D, const double *eta, const double *R,
                     const double *kappa, const double *dxm, double sec_per_year, double *dwdt) {
    const int n = r->n;
    const size_t nn = (size_t)n * n;
    or_deriv_x(n, n, w, dxm, r->w_x);
    or_deriv_y(n, n, r->w_x, dxm, r->w_xy);
    or_deriv_xx(n, n, w, dxm, r->w_xx);
    or_deriv_yy(n, n, w, dxm, r->w_yy);
    for (size_t c = 0; c < nn; c++) {
        r->Mxx[c] = -(D[c] * (r->w_xx[c] + nu * r->w_yy[c]));
        r->Myy[c] = -(D[c] * (r->w_yy[c] + nu * r->w_xx[c]));
        r->Mxy[c] = -(D[c] * (1.0 - nu) * r->w_xy[c]);
    }
    or_deriv_x(n, n, r->Mxy, dxm, r->Mxy_x);
    or_deriv_y(n, n, r->Mxy_x, dxm, r->Mxy_xy);
    or_deriv_xx(n, n, r->Mxx, dxm, r->Mx_xx);
    or_deriv_yy(n, n, r->Myy, dxm, r->My_yy);
    for (size_t c = 0; c < nn; c++) {
        const double p = (maskactive[c] != 0.0) ? -(g * canom_full[c]) : 0.0;
        r->f[c] = (p + r->Mx_xx[c] + 2.0 * r->Mxy_xy[c] + r->My_yy[c]) / (2.0 * eta[c]);
    }
    dht2_plan(r->p, r->f, r->tmp2n, r->z, r->wk);
    for (size_t c = 0; c < nn; c++) r->f[c] = r->f[c] / (R[c] * kappa[c]);
    dht2_plan(r->p, r->f, r->tmp2n, r->z, r->wk);
    const double scale = (double)(float)((double)n * (double)n);   /* (nx*ny*1.) default real */
    for (size_t c = 0; c < nn; c++) r->f[c] = r->f[c] / scale;
    zerobc(n, r->f);
    for (size_t c = 0; c < nn; c++) dwdt[c] = r->f[c] * sec_per_year;
}

void or_lvelva(int n, const double *w, const double *canom_full, const double *maskactive,
               double g, double nu, const double *D, const double *eta, const double *R,
               const double *kappa, const double *dxm, double sec_per_year, double *dwdt) {
    rhsctx *r = rhs_create(n);
    rhs_eval(r, w, canom_full, maskactive, g, nu, D, eta, R, kappa, dxm, sec_per_year, dwdt);
    rhs_destroy(r);
}

/* ------------------------------------------------------------------------ */
/* interp_0d, isos_utils.f90:423-457 */
double or_interp_0d(int n, const double *x, const double *y, double xout) {
    if (xout < x[0]) return y[0];
    if (xout > x[n - 1]) return y[n - 1];
    int j;
    for (j = 0; j < n; j++) if (x[j] >= xout) break;
    if (j == 0) return y[0];
    if (j == n) return y[n - 1];
    const double alpha = (xout - x[j - 1]) / (x[j] - x[j - 1]);
    return y[j - 1] + alpha * (y[j] - y[j - 1]);
}

/* ------------------------------------------------------------------------ */
/* bsl_class, barysealevel.f90:16-29, 40-252 (methods const, fastiso, file, mixed) */
struct or_bsl {
    int method;                 /* 0 const, 1 fastiso, 2 file, 3 mixed */
    double bsl_init, bsl_now, time, A_ocean_pd, A_ocean_now;
    int constant_ocean_surface, nvec, ns;
    double *bsl_vec, *A_vec;
    double *st, *sv;            /* series%time, series%var */
};

/* bsl_init_fastiso, barysealevel.f90:155-193 */
static void bsl_init_fastiso(or_bsl *b, double bsl_init, int nvec, const double *bsl_vec, const double *A_vec) {
    b->bsl_init = bsl_init;
    if (nvec <= 0) {
        b->constant_ocean_surface = 1; b->A_ocean_now = b->A_ocean_pd;
        return;
    }
    b->constant_ocean_surface = 0;
    b->nvec = nvec;
    b->bsl_vec = (double *)malloc(sizeof(double) * (size_t)nvec);
    b->A_vec = (double *)malloc(sizeof(double) * (size_t)nvec);
    memcpy(b->bsl_vec, bsl_vec, sizeof(double) * (size_t)nvec);
    memcpy(b->A_vec, A_vec, sizeof(double) * (size_t)nvec);
    const double biased = or_interp_0d(nvec, b->bsl_vec, b->A_vec, b->bsl_init);
    for (int i = 0; i < nvec; i++) b->A_vec[i] = b->A_vec[i] * b->A_ocean_pd / biased;
    b->A_ocean_now = or_interp_0d(nvec, b->bsl_vec, b->A_vec, b->bsl_init);
}

/* series_interp / interp_linear, barysealevel.f90:353-400 (the interp_0d rule) */
static double series_interp(const or_bsl *b, double year_bp) { return or_interp_0d(b->ns, b->st, b->sv, year_bp); }

/* bsl_init, barysealevel.f90:40-106; bsl_init_file :108-153 */
or_bsl *or_bsl_create_series(int method, double bsl_init, double A_ocean_pd, int nvec, const double *bsl_vec,
                             const double *A_vec, int ns, const double *st, const double *sv, double year_bp) {
    if (method < 0 || method > 3) { set_err("bsl: method not recognized"); return NULL; }
    if ((method == 2 || method == 3) && ns < 1) { set_err("bsl: file and mixed need a series"); return NULL; }
    or_bsl *b = (or_bsl *)calloc(1, sizeof *b);
    b->method = method; b->time = year_bp; b->A_ocean_pd = A_ocean_pd; b->A_ocean_now = A_ocean_pd;
    b->bsl_init = bsl_init;
    if (method == 2 || method == 3) {
        b->ns = ns;
        b->st = (double *)malloc(sizeof(double) * (size_t)ns);
        b->sv = (double *)malloc(sizeof(double) * (size_t)ns);
        memcpy(b->st, st, sizeof(double) * (size_t)ns);
        memcpy(b->sv, sv, sizeof(double) * (size_t)ns);
        b->bsl_now = series_interp(b, year_bp);
        b->bsl_init = b->bsl_now;
    }
    if (method == 1 || method == 3) bsl_init_fastiso(b, bsl_init, nvec, bsl_vec, A_vec);
    b->bsl_now = b->bsl_init;
    return b;
}

or_bsl *or_bsl_create(int method, double bsl_init, double A_ocean_pd, int nvec,
                      const double *bsl_vec, const double *A_vec, double year_bp) {
    if (method != 0 && method != 1) { set_err("bsl: file and mixed need a series"); return NULL; }
    return or_bsl_create_series(method, bsl_init, A_ocean_pd, nvec, bsl_vec, A_vec, 0, NULL, NULL, year_bp);
}

/* bsl_update, barysealevel.f90:195-252 */
void or_bsl_update(or_bsl *b, double year_bp) {
    b->time = year_bp;
    if (b->method == 0) { b->bsl_now = b->bsl_init; return; }
    if (b->method == 2 || b->method == 3) b->bsl_now = series_interp(b, year_bp);   /* bsl_update_file */
    if (b->method == 2) return;
    if (b->constant_ocean_surface) b->A_ocean_now = b->A_ocean_pd;                   /* bsl_update_fastiso */
    else b->A_ocean_now = or_interp_0d(b->nvec, b->bsl_vec, b->A_vec, b->bsl_now);
}
double or_bsl_now(const or_bsl *b) { return b->bsl_now; }
double or_bsl_A_ocean(const or_bsl *b) { return b->A_ocean_now; }
void or_bsl_destroy(or_bsl *b) { if (b) { free(b->bsl_vec); free(b->A_vec); free(b->st); free(b->sv); free(b); } }

/* ------------------------------------------------------------------------ */
/* isos_class (isostasy_defs.f90:202-209) */
typedef struct {
    double bsl, V_af, V_pov, V_den, deltaV_bsl, dbsl_total;
    double *z_bed, *dwdt, *w, *we, *Haf, *Hice, *z_ss, *dz_ss, *canom_ice, *canom_seawater,
           *canom_litho, *canom_mantle, *canom_load, *canom_full, *mass_anom, *w_eq;
    unsigned char *maskocean, *maskgrounded, *maskcontinent;
} state_t;

struct or_isos {
    or_params par;
    int n, nx_ice, ny_ice, icrop1, icrop2, jcrop1, jcrop2;  /* 1-based crop indices */
    double dx, dy, sec_per_year, compressibility_correction, Vden_factor;
    double time_diagnostics, time_prognostics;
    int ref_was_set, use_restart;
    int nan_stop;                /* the last integration stopped on a NaN error norm */
    double *K, *dxm, *A, *kappa, *R, *maskactive, *He_lith, *D_lith, *eta_eff;
    cplx *FGE, *FGN, *FGV;
    double tau, L_w;            /* domain%tau (uniform), par%L_w */
    state_t now, ref;
    double ode_t, ode_dt, *ode_x, *ode_x1, *ode_x2, *k;
    long attempts, accepts;
    rhsctx *rhs;
    convctx *conv;
};

static void state_alloc(state_t *s, size_t nn) {
    double **f[] = {&s->z_bed, &s->dwdt, &s->w, &s->we, &s->Haf, &s->Hice, &s->z_ss, &s->dz_ss, &s->canom_ice,
                    &s->canom_seawater, &s->canom_litho, &s->canom_mantle, &s->canom_load, &s->canom_full, &s->mass_anom,
                    &s->w_eq};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; i++) *f[i] = (double *)calloc(nn, sizeof(double));
    s->maskocean = (unsigned char *)calloc(nn, 1);
    s->maskgrounded = (unsigned char *)calloc(nn, 1);
    s->maskcontinent = (unsigned char *)calloc(nn, 1);
}
static void state_free(state_t *s) {
    double *f[] = {s->z_bed, s->dwdt, s->w, s->we, s->Haf, s->Hice, s->z_ss, s->dz_ss, s->canom_ice,
                   s->canom_seawater, s->canom_litho, s->canom_mantle, s->canom_load, s->canom_full, s->mass_anom,
                   s->w_eq};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; i++) free(f[i]);
    free(s->maskocean); free(s->maskgrounded); free(s->maskcontinent);
}
/* zero_init_state, fastisostasy.f90:506-548 */
static void state_zero(state_t *s, size_t nn, double z_bed_background) {
    for (size_t c = 0; c < nn; c++) s->z_bed[c] = z_bed_background;
    s->bsl = s->deltaV_bsl = s->dbsl_total = s->V_af = s->V_den = s->V_pov = 0.0;
}

/* init_domain_size, isos_utils.f90:285-354 */
static int domain_size(or_isos *s, int nx_ice, int ny_ice, double dx, double dy, double min_pad) {
    if (min_pad < 0.0) { set_err("Minimum padding must be positive."); return -1; }
    if (dx != dy) { set_err("Currently, only square grid cells are supported."); return -1; }
    s->dx = dx; s->dy = dy; s->nx_ice = nx_ice; s->ny_ice = ny_ice;
    int icrop1 = 1, jcrop1 = 1, nx_tmp = nx_ice, ny_tmp = ny_ice;
    if (nx_ice % 2 != 0) { icrop1++; nx_tmp++; }
    if (ny_ice % 2 != 0) { jcrop1++; ny_tmp++; }
    const int npad = (int)lround(min_pad / dx);
    int npx = npad, npy = npad;
    if (nx_tmp < ny_tmp) npx += (ny_tmp - nx_tmp) / 2;
    else if (nx_tmp > ny_tmp) npy += (nx_tmp - ny_tmp) / 2;
    const int n = nx_tmp + 2 * npx;
    if (n != ny_tmp + 2 * npy) { set_err("Failed to correctly pad domain."); return -1; }
    s->n = n;
    s->icrop1 = icrop1 + npx; s->icrop2 = n - npx;
    s->jcrop1 = jcrop1 + npy; s->jcrop2 = n - npy;
    return 0;
}

static void out2in(const or_isos *s, double *X_in, const double *X_out) {
    const int n = s->n, ni = s->icrop2 - s->icrop1 + 1, nj = s->jcrop2 - s->jcrop1 + 1;
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) X_in[(size_t)(j + s->jcrop1 - 1) * n + (i + s->icrop1 - 1)] = X_out[(size_t)j * ni + i];
}

/* calc_gaussian_thickness / calc_gaussian_viscosity, isos_utils.f90:960-1035 */
static void gaussian_shape(int n, const double *xc, const double *yc, double *f) {
    const double xmin = xc[0], xmax = xc[n - 1], ymin = yc[0], ymax = yc[n - 1];
    const double xcn = (xmax + xmin) / 2.0, ycn = (ymax + ymin) / 2.0;
    const double Lx = (xmax - xmin) / 2., Ly = (ymax - ymin) / 2.;
    const double sx = (4.0 / Lx) * (4.0 / Lx), sy = (4.0 / Ly) * (4.0 / Ly);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double ddx = xc[i] - xcn, ddy = yc[j] - ycn;
            f[(size_t)j * n + i] = exp(-0.5 * (ddx * (sx * ddx) + ddy * (sy * ddy)));
        }
}

/* bilinear_interpolation (isos_utils.f90:495-558): flat (nearest edge value in
 * both directions) when either coordinate lies within 1 m of or beyond the
 * grid ends; z is x fastest */
static double bilinear_point(int nxr, int nyr, const double *x, const double *y, const double *z,
                             double xo, double yo) {
    const double tol = 1.0;
    int i = 0, j = 0, flat = 0;
    double tx = 0.0, ty = 0.0;
    if (xo <= x[0] + tol) { i = 0; flat = 1; }
    else if (xo >= x[nxr - 1] - tol) { i = nxr - 1; flat = 1; }
    else {
        for (i = 0; i < nxr - 1; i++) if (xo > x[i] && xo <= x[i + 1]) break;
        tx = (xo - x[i]) / (x[i + 1] - x[i]);
    }
    if (yo <= y[0] + tol) { j = 0; flat = 1; }
    else if (yo >= y[nyr - 1] - tol) { j = nyr - 1; flat = 1; }
    else {
        for (j = 0; j < nyr - 1; j++) if (yo > y[j] && yo <= y[j + 1]) break;
        ty = (yo - y[j]) / (y[j + 1] - y[j]);
    }
    if (flat) return z[(size_t)j * nxr + i];
    const double z11 = z[(size_t)j * nxr + i], z21 = z[(size_t)j * nxr + i + 1];
    const double z12 = z[(size_t)(j + 1) * nxr + i], z22 = z[(size_t)(j + 1) * nxr + i + 1];
    return z11 * (1.0 - tx) * (1.0 - ty) + z21 * tx * (1.0 - ty) + z12 * (1.0 - tx) * ty + z22 * tx * ty;
}

void or_bilinear(int nxr, int nyr, const double *x, const double *y, const double *z,
                 int nxo, int nyo, const double *xo, const double *yo, double *out) {
    for (int j = 0; j < nyo; j++)
        for (int i = 0; i < nxo; i++) out[(size_t)j * nxo + i] = bilinear_point(nxr, nyr, x, y, z, xo[i], yo[j]);
}

/* smooth_gauss_2D (isos_utils.f90:696-817) with gauss_values (:918-957): a
 * (2 n2 + 1)^2 Gaussian window, n2 = ceiling(2 sigma / dx), normalised twice,
 * over the field mirrored at its edges (edge cells repeated) */
void or_smooth_gauss(int nx, int ny, double *var, double dx, double sigma) {
    const int n2 = (int)ceil(2.0 * sigma / dx), n = 2 * n2 + 1;
    double *f = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int j = -n2; j <= n2; j++)
        for (int i = -n2; i <= n2; i++) {
            const double x = i * dx, y = j * dx;
            f[(size_t)(j + n2) * n + (i + n2)] =
                1.0 / (2.0 * PI_REF * (sigma * sigma)) * exp(-(x * x + y * y) / (2 * (sigma * sigma)));
        }
    double fs = 0.0;
    for (size_t q = 0; q < (size_t)n * n; q++) fs += f[q];
    for (size_t q = 0; q < (size_t)n * n; q++) f[q] = f[q] / fs;
    fs = 0.0;
    for (size_t q = 0; q < (size_t)n * n; q++) fs += f[q];
    for (size_t q = 0; q < (size_t)n * n; q++) f[q] = f[q] / fs;
    const int ex = nx + 2 * n2, ey = ny + 2 * n2;
    double *e = (double *)malloc(sizeof(double) * (size_t)ex * ey);
#define EXT(i, j) e[(size_t)((j) + n2) * ex + ((i) + n2)]   /* 0-based i in [-n2, nx+n2) */
    for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) EXT(i, j) = var[(size_t)j * nx + i];
    for (int j = 0; j < ny; j++) {
        for (int i = -1; i >= -n2; i--) EXT(i, j) = EXT(-i - 1, j);
        for (int i = nx; i < nx + n2; i++) EXT(i, j) = EXT(2 * nx - 1 - i, j);
    }
    for (int i = -n2; i < nx + n2; i++) {
        for (int j = -1; j >= -n2; j--) EXT(i, j) = EXT(i, -j - 1);
        for (int j = ny; j < ny + n2; j++) EXT(i, j) = EXT(i, 2 * ny - 1 - j);
    }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            double acc = 0.0;
            for (int b = -n2; b <= n2; b++)
                for (int a = -n2; a <= n2; a++) acc += EXT(i + a, j + b) * f[(size_t)(b + n2) * n + (a + n2)];
            var[(size_t)j * nx + i] = acc;
        }
#undef EXT
    free(e); free(f);
}

static or_isos *create_impl(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                            const double *K, const double *rho_ice, const double *he_field,
                            const double *eta_field, const or_rheology_grid *g);

or_isos *or_isos_create_gridded(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                                const double *K, const double *rho_ice, const or_rheology_grid *g) {
    return create_impl(p, nx_ice, ny_ice, dx, dy, K, rho_ice, NULL, NULL, g);
}

or_isos *or_isos_create(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                        const double *K, const double *rho_ice,
                        const double *he_field, const double *eta_field) {
    return create_impl(p, nx_ice, ny_ice, dx, dy, K, rho_ice, he_field, eta_field, NULL);
}

static or_isos *create_impl(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                            const double *K, const double *rho_ice, const double *he_field,
                            const double *eta_field, const or_rheology_grid *g) {
    if (p->method < 0 || p->method > 3) { set_err("method must be 0 (none), 1 (LLRA), 2 (LV-ELRA) or 3 (LV-ELVA)"); return NULL; }
    if (p->nl < 1 || p->nl > OR_MAXL) { set_err("nl out of range"); return NULL; }
    or_isos *s = (or_isos *)calloc(1, sizeof *s);
    s->par = *p;
    if (K) s->par.correct_distortion = 1;
    if (rho_ice) s->par.rho_ice = *rho_ice;
    if (domain_size(s, nx_ice, ny_ice, dx, dy, p->min_pad)) { free(s); return NULL; }
    const int n = s->n;
    const size_t nn = (size_t)n * n;
    double *xc = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) xc[i] = (double)(i + 1 - n / 2 - 1) * dx;     /* init_dims, isos_utils.f90:221-240 */
    double **f[] = {&s->K, &s->dxm, &s->A, &s->kappa, &s->R, &s->maskactive, &s->He_lith, &s->D_lith, &s->eta_eff};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; i++) *f[i] = (double *)calloc(nn, sizeof(double));
    for (size_t c = 0; c < nn; c++) { s->R[c] = 1; s->K[c] = K ? K[c] : 1; }
    for (size_t c = 0; c < nn; c++) { s->dxm[c] = dx * s->K[c]; s->A[c] = s->dxm[c] * (dy * s->K[c]); }
    s->sec_per_year = (double)(3600.0f * 24.0f * 365.25f);        /* fastisostasy.f90:158 */
    s->compressibility_correction = s->par.correct_compression ? 1.5 / (1 + s->par.nu) : 1.0;
    s->Vden_factor = s->par.rho_ice / s->par.rho_water - s->par.rho_ice / s->par.rho_seawater;

    /* lithosphere, fastisostasy.f90:166-172,246-301 */
    const double nu = s->par.nu;
    for (size_t c = 0; c < nn; c++) s->He_lith[c] = s->par.zl[0];
    const double He0 = s->par.zl[0];
    const double D0 = s->par.E * 1e9 * (He0 * He0 * He0) / (12.0 * (1.0 - nu * nu));
    for (size_t c = 0; c < nn; c++) s->D_lith[c] = D0;

    /* Green's functions and spectra, fastisostasy.f90:174-183 */
    s->conv = conv_create(n);
    double *gk = (double *)malloc(sizeof(double) * nn);
    const size_t mm = (size_t)s->conv->m * s->conv->m;
    s->FGE = (cplx *)malloc(sizeof(cplx) * mm);
    s->FGN = (cplx *)malloc(sizeof(cplx) * mm);
    or_green_elastic(n, dx, gk); conv_spectrum(s->conv, gk, s->FGE);
    or_green_ssh(n, dx, s->par.r_earth, s->par.m_earth, gk); conv_spectrum(s->conv, gk, s->FGN);

    /* maskactive, fastisostasy.f90:185-213 */
    if (g && g->mask) {
        double *mr = (double *)malloc(sizeof(double) * nn);
        or_bilinear(g->mask_nx, g->mask_ny, g->mask_xc, g->mask_yc, g->mask, n, n, xc, xc, mr);
        for (size_t c = 0; c < nn; c++) s->maskactive[c] = mr[c] > 0.5 ? 1 : 0;
        free(mr);
    } else if (!s->par.interactive_sealevel) {
        for (size_t c = 0; c < nn; c++) s->maskactive[c] = 1;
    } else if (s->par.min_pad > 100.e3f) {
        for (int j = s->jcrop1 - 1; j < s->jcrop2; j++)
            for (int i = s->icrop1 - 1; i < s->icrop2; i++) s->maskactive[(size_t)j * n + i] = 1;
    } else {
        set_err("interactive_sealevel=true requires a mask file or at least 100km of padding.");
        free(gk); free(xc); or_isos_destroy(s); return NULL;
    }

    /* fastisostasy.f90:216-236,444-450 */
    if (s->par.method == 2) {
        s->L_w = pow(D0 / (s->par.rho_uppermantle * s->par.g), 0.25);   /* calc_flexural_lengthscale, isos_utils.f90:66-73 */
        or_green_viscous(n, dx, s->L_w, D0, gk);
        s->FGV = (cplx *)malloc(sizeof(cplx) * mm);
        conv_spectrum(s->conv, gk, s->FGV);
        s->tau = s->par.tau;
    } else if (s->par.method != 3) {
        s->tau = s->par.tau;
        for (size_t c = 0; c < nn; c++) s->maskactive[c] = 1;
        s->L_w = 0.0;
    }
    if (s->par.method == 3) {
        or_kappa(n, dx, s->kappa);
        switch (s->par.lithosphere) {
        case 0: break;
        case 1: case 2: {
            gaussian_shape(n, xc, xc, gk);
            double mx = gk[0];
            for (size_t c = 1; c < nn; c++) mx = fmax(mx, gk[c]);
            const double sign = (s->par.lithosphere == 1) ? 1.0 : -1.0;
            for (size_t c = 0; c < nn; c++) s->He_lith[c] = He0 + sign * F32(100e3) * (gk[c] / mx);
            break;
        }
        case 3:
            if (!he_field) { set_err("lithosphere=fields needs he_field"); free(gk); free(xc); or_isos_destroy(s); return NULL; }
            memcpy(s->He_lith, he_field, sizeof(double) * nn);
            break;
        case 4: {   /* rheology_file, fastisostasy.f90:262-290 */
            if (!g || !g->litho_thickness) { set_err("lithosphere=rheology_file needs the gridded thickness"); free(gk); free(xc); or_isos_destroy(s); return NULL; }
            const size_t nr = (size_t)g->nx * g->ny;
            double *T = (double *)malloc(sizeof(double) * nr);
            for (size_t q = 0; q < nr; q++) T[q] = g->litho_thickness[q] * g->litho_thickness_scaling;
            or_bilinear(g->nx, g->ny, g->xc, g->yc, T, n, n, xc, xc, s->He_lith);
            free(T);
            if (g->rheo_smoothing_radius > 0.0) or_smooth_gauss(n, n, s->He_lith, dx, g->rheo_smoothing_radius);
            break;
        }
        default: set_err("unknown lithosphere"); free(gk); free(xc); or_isos_destroy(s); return NULL;
        }
        if (s->par.lithosphere != 0)
            for (size_t c = 0; c < nn; c++) {
                const double h = s->He_lith[c];
                s->D_lith[c] = s->par.E * 1e9 * (h * h * h) / (12.0 * (1.0 - nu * nu));
            }
        /* layer boundaries, fastisostasy.f90:303-316 / lv_elva.f90:131-168 */
        const int nl = s->par.nl;
        double *bnd = (double *)malloc(sizeof(double) * nn * (size_t)nl);
        double *eta = (double *)malloc(sizeof(double) * nn * (size_t)nl);
        for (size_t c = 0; c < nn; c++) bnd[c] = s->He_lith[c];
        for (int k = 1; k < nl; k++)
            for (size_t c = 0; c < nn; c++)
                bnd[k * nn + c] = (s->par.layering == 0) ? bnd[(k - 1) * nn + c] + s->par.dl : s->par.zl[k];
        for (int k = 0; k < nl; k++) for (size_t c = 0; c < nn; c++) eta[k * nn + c] = s->par.viscosities[k];
        switch (s->par.mantle) {
        case 0: or_lumping(n, nl, s->par.lumping, eta, bnd, s->kappa, s->eta_eff, s->R); break;
        case 1: case 2: {
            gaussian_shape(n, xc, xc, gk);
            const double sign = (s->par.mantle == 1) ? 1.0 : -1.0;
            for (size_t c = 0; c < nn; c++) s->eta_eff[c] = s->par.viscosities[0] * pow(10.0, sign * gk[c]);
            break;
        }
        case 3:
            if (!eta_field) { set_err("mantle=fields needs eta_field"); free(bnd); free(eta); free(gk); free(xc); or_isos_destroy(s); return NULL; }
            memcpy(eta, eta_field, sizeof(double) * nn * (size_t)nl);
            or_lumping(n, nl, s->par.lumping, eta, bnd, s->kappa, s->eta_eff, s->R);
            break;
        case 4: {   /* rheology_file, fastisostasy.f90:336-411 */
            if (!g || !g->log10_eta || !g->zc || g->nz < 1) { set_err("mantle=rheology_file needs the gridded viscosity"); free(bnd); free(eta); free(gk); free(xc); or_isos_destroy(s); return NULL; }
            const int nz = g->nz;
            const size_t nr = (size_t)g->nx * g->ny;
            double *depth = (double *)malloc(sizeof(double) * (size_t)nz);
            for (int i = 0; i < nz; i++) depth[i] = g->zc[nz - 1] - g->zc[nz - 1 - i];
            double *bil = (double *)malloc(sizeof(double) * nn * (size_t)nz), *tmp = (double *)malloc(sizeof(double) * nn);
            for (int i = 0; i < nz; i++) {
                or_bilinear(g->nx, g->ny, g->xc, g->yc, g->log10_eta + (size_t)(nz - 1 - i) * nr, n, n, xc, xc, bil + (size_t)i * nn);
                if (s->par.viscosity_scaling_method == 4 && g->stddev_log10_eta) {
                    or_bilinear(g->nx, g->ny, g->xc, g->yc, g->stddev_log10_eta + (size_t)(nz - 1 - i) * nr, n, n, xc, xc, tmp);
                    for (size_t c = 0; c < nn; c++) bil[(size_t)i * nn + c] = bil[(size_t)i * nn + c] + s->par.viscosity_scaling * tmp[c];
                }
            }
            for (size_t q = 0; q < nn * (size_t)nz; q++) {
                if (bil[q] < 16) bil[q] = 16;
                if (bil[q] > 23) bil[q] = 23;
            }
            double *col = (double *)malloc(sizeof(double) * (size_t)nz);
            for (int k = 0; k < nl; k++)
                for (size_t c = 0; c < nn; c++) {
                    for (int i = 0; i < nz; i++) col[i] = bil[(size_t)i * nn + c];
                    eta[(size_t)k * nn + c] = or_interp_0d(nz, depth, col, bnd[(size_t)k * nn + c]);
                }
            for (size_t q = 0; q < nn * (size_t)nl; q++) eta[q] = pow(10.0, eta[q]);
            or_lumping(n, nl, s->par.lumping, eta, bnd, s->kappa, s->eta_eff, s->R);
            if (g->rheo_smoothing_radius > 0.0) or_smooth_gauss(n, n, s->eta_eff, dx, g->rheo_smoothing_radius);
            free(depth); free(bil); free(tmp); free(col);
            break;
        }
        default: set_err("unknown mantle"); free(bnd); free(eta); free(gk); free(xc); or_isos_destroy(s); return NULL;
        }
        free(bnd); free(eta);
        /* fastisostasy.f90:425-442 */
        if (s->par.viscosity_scaling_method == 1) {
            double m = s->eta_eff[0];
            for (size_t c = 1; c < nn; c++) m = fmin(m, s->eta_eff[c]);
            for (size_t c = 0; c < nn; c++) s->eta_eff[c] = m;
        } else if (s->par.viscosity_scaling_method == 2) {
            for (size_t c = 0; c < nn; c++) s->eta_eff[c] = s->eta_eff[c] * s->par.viscosity_scaling;
        } else if (s->par.viscosity_scaling_method == 3) {
            double m = s->eta_eff[0];
            for (size_t c = 1; c < nn; c++) m = fmax(m, s->eta_eff[c]);
            for (size_t c = 0; c < nn; c++) s->eta_eff[c] = m;
        }
        for (size_t c = 0; c < nn; c++) s->eta_eff[c] = s->eta_eff[c] * s->compressibility_correction;
        for (size_t c = 0; c < nn; c++) s->eta_eff[c] = exp(log10(s->par.eta_ref / s->eta_eff[c])) * s->eta_eff[c];
    }
    free(gk); free(xc);

    s->time_diagnostics = F32(1e10);
    s->time_prognostics = F32(1e10);
    state_alloc(&s->now, nn); state_alloc(&s->ref, nn);
    state_zero(&s->now, nn, 1e6); state_zero(&s->ref, nn, 1e6);
    s->ref_was_set = 0;
    s->ode_x = (double *)calloc(nn, sizeof(double));
    s->ode_x1 = (double *)calloc(nn, sizeof(double));
    s->ode_x2 = (double *)calloc(nn, sizeof(double));
    s->k = (double *)calloc(nn, sizeof(double));
    s->rhs = rhs_create(n);
    return s;
}

void or_isos_destroy(or_isos *s) {
    if (!s) return;
    double *f[] = {s->K, s->dxm, s->A, s->kappa, s->R, s->maskactive, s->He_lith, s->D_lith, s->eta_eff,
                   s->ode_x, s->ode_x1, s->ode_x2, s->k};
    for (size_t i = 0; i < sizeof f / sizeof f[0]; i++) free(f[i]);
    free(s->FGE); free(s->FGN); free(s->FGV);
    if (s->now.z_bed) state_free(&s->now);
    if (s->ref.z_bed) state_free(&s->ref);
    rhs_destroy(s->rhs); conv_destroy(s->conv);
    free(s);
}

/* isos_init_ref, fastisostasy.f90:550-575 */
int or_isos_init_ref(or_isos *s, const double *z_bed, const double *H_ice, const double *bsl, const double *dz_ss) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) { s->ref.z_bed[c] = 1e6; s->ref.Hice[c] = 0.0; }
    out2in(s, s->ref.z_bed, z_bed);
    out2in(s, s->ref.Hice, H_ice);
    if (bsl) s->ref.bsl = *bsl;
    if (dz_ss) memcpy(s->ref.dz_ss, dz_ss, sizeof(double) * nn);
    s->ref_was_set = 1;
    return 0;
}

/* sealevel.f90 */
static void calc_z_ss(size_t nn, double *z_ss, double bsl, const double *z_ss_ref, const double *dz_ss) {
    for (size_t c = 0; c < nn; c++) z_ss[c] = bsl + z_ss_ref[c] + dz_ss[c];
}
static void calc_Haf(const or_isos *s, state_t *st) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) {
        double h = (st->z_ss[c] - st->z_bed[c]) * (s->par.rho_seawater / s->par.rho_ice);
        if (h < 0.0) h = 0.0;
        h = st->Hice[c] - h;
        if (h < 0.0) h = 0.0;
        st->Haf[c] = h;
    }
}
static void calc_masks(const or_isos *s, state_t *st) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) {
        st->maskcontinent[c] = st->z_bed[c] > st->z_ss[c];
        st->maskgrounded[c] = (st->Hice[c] > 0.0) && (st->Haf[c] > 0.0);
        st->maskocean[c] = !(st->maskcontinent[c] || st->maskgrounded[c]);
    }
}
static void canom_ice(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) s->now.canom_ice[c] = 0.0 + s->par.rho_ice * (s->now.Haf[c] - s->ref.Haf[c]);
}
static void canom_seawater(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) {
        const double h = s->now.maskocean[c] ? s->par.rho_water * (s->now.z_ss[c] - s->ref.z_ss[c]) : 0.0;
        s->now.canom_seawater[c] = 0.0 + h;
    }
}
static void canom_load(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) {
        double v = s->now.canom_ice[c];
        if (s->par.interactive_sealevel) v = v + s->now.canom_seawater[c];
        s->now.canom_load[c] = s->maskactive[c] != 0.0 ? v : 0.0;
    }
}
static void canom_litho(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) s->now.canom_litho[c] = 0.0 + s->par.rho_litho * (s->now.we[c] - s->ref.we[c]);
}
static void canom_mantle(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) s->now.canom_mantle[c] = 0.0 + s->par.rho_uppermantle * (s->now.w[c] - s->ref.w[c]);
}
static void canom_full(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) {
        const double v = s->now.canom_load[c] + s->now.canom_litho[c] + s->now.canom_mantle[c];
        s->now.canom_full[c] = s->maskactive[c] != 0.0 ? v : 0.0;
    }
}
static void mass_anom(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    for (size_t c = 0; c < nn; c++) s->now.mass_anom[c] = s->maskactive[c] != 0.0 ? s->A[c] * s->now.canom_full[c] : 0.0;
}
/* calc_sl_contributions, sealevel.f90:189-205 */
static void sl_contributions(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    const double V_old = s->now.V_af + s->now.V_den + s->now.V_pov;
    double a = 0.0, b = 0.0;
    for (size_t c = 0; c < nn; c++) a += (s->now.Haf[c] - s->ref.Haf[c]) * s->A[c];
    for (size_t c = 0; c < nn; c++) b += (s->now.Hice[c] - s->ref.Hice[c]) * s->Vden_factor * s->A[c];
    s->now.V_af = a * (s->par.rho_ice / s->par.rho_seawater);
    s->now.V_den = b;
    s->now.V_pov = 0.0;
    s->now.deltaV_bsl = (s->now.V_af + s->now.V_den + s->now.V_pov) - V_old;
}

/* get_dwdt, fastisostasy.f90:677-729 (the w argument is ignored, now%w is used) */
static void get_dwdt(or_isos *s) {
    const size_t nn = (size_t)s->n * s->n;
    if (s->par.method == 0) {
        memset(s->now.w_eq, 0, sizeof(double) * nn);
        memset(s->now.dwdt, 0, sizeof(double) * nn);
        return;
    }
    if (s->par.method == 1 || s->par.method == 2) {
        if (s->par.method == 1) {
            /* calc_llra_equilibrium, lv_xlra.f90:29-38 */
            for (size_t c = 0; c < nn; c++) s->now.w_eq[c] = -s->now.canom_load[c] / s->par.rho_uppermantle;
        } else {
            /* fastisostasy.f90:702-712: convolution of -canom_load g K^2 with GV, corner BC */
            double *in = (double *)calloc(nn, sizeof(double));
            for (size_t c = 0; c < nn; c++) in[c] = -s->now.canom_load[c] * s->par.g * pow(s->K[c], 2.0);
            conv_apply(s->conv, s->FGV, in, s->now.w_eq);
            free(in);
            zerobc(s->n, s->now.w_eq);
        }
        /* calc_asthenosphere_relax, lv_xlra.f90:17-26 */
        for (size_t c = 0; c < nn; c++) s->now.dwdt[c] = (s->now.w_eq[c] - s->now.w[c]) / s->tau;
        return;
    }
    canom_full(s);
    rhs_eval(s->rhs, s->now.w, s->now.canom_full, s->maskactive, s->par.g, s->par.nu, s->D_lith, s->eta_eff,
             s->R, s->kappa, s->dxm, s->sec_per_year, s->now.dwdt);
}

static double maxabs(const double *x, size_t nn) {
    double m = 0.0;
    for (size_t c = 0; c < nn; c++) m = fmax(m, fabs(x[c]));
    return m;
}

/* integrators.f90:12-203.  All stages evaluate to the same k (the RHS reads
 * now%w and the a-matrix lower triangle is zero), so k is evaluated once per
 * attempt; the stage sums keep the reference's association order. */
static void integrate(or_isos *s, double tf) {
    const size_t nn = (size_t)s->n * s->n;
    double *x = s->ode_x, *x1 = s->ode_x1, *x2 = s->ode_x2, *k = s->now.dwdt;
    const double b[6] = {0.0845, 0.0, 0.0, 0.5179, 0.1276, 0.2700};
    const double bh[6] = {0.0845, 0.0, 0.0, 0.5179, 0.1276, 0.2700 - 1 / 5.0};
    while (s->ode_t < tf) {
        const int m = s->par.dt_method;
        if (m == 1) {               /* rk4 */
            const double h = fmin(s->ode_dt, tf - s->ode_t);
            get_dwdt(s); s->attempts++;
            for (size_t c = 0; c < nn; c++) x[c] = x[c] + (h / 6) * (((k[c] + 2 * k[c]) + 2 * k[c]) + k[c]);
            memcpy(s->now.w, x, sizeof(double) * nn);
            s->ode_t = s->ode_t + h; s->accepts++;
            continue;
        }
        s->ode_dt = fmin(s->ode_dt, tf - s->ode_t);
        if (m == 0) {               /* euler */
            get_dwdt(s); s->attempts++;
            for (size_t c = 0; c < nn; c++) x[c] = x[c] + s->ode_dt * k[c];
            memcpy(s->now.w, x, sizeof(double) * nn);
            s->ode_t = s->ode_t + s->ode_dt; s->accepts++;
            continue;
        }
        if (m == 2) s->ode_dt = fmin(s->ode_dt, s->par.dt_max);
        get_dwdt(s); s->attempts++;
        const double dt = s->ode_dt;
        if (m == 2) {               /* bs32: 7/24, 1/4, 1/3 are integer divisions (= 0) */
            for (size_t c = 0; c < nn; c++) {
                x1[c] = x[c] + dt / 9 * ((2 * k[c] + 3 * k[c]) + 4 * k[c]);
                x2[c] = x[c] + dt * (((0.0 * k[c] + 0.0 * k[c]) + 0.0 * k[c]) + 3 * k[c]);
            }
        } else {                    /* tsit54 */
            double cb[6], cbh[6];
            for (int i = 0; i < 6; i++) { cb[i] = dt * b[i]; cbh[i] = dt * bh[i]; }
            for (size_t c = 0; c < nn; c++) {
                const double kc = k[c];
                x1[c] = x[c] + cbh[0] * kc + cbh[1] * kc + cbh[2] * kc + cbh[3] * kc + cbh[4] * kc + cbh[5] * kc;
                x2[c] = x[c] + cb[0] * kc + cb[1] * kc + cb[2] * kc + cb[3] * kc + cb[4] * kc + cb[5] * kc;
            }
        }
        const double tol = s->par.atol + s->par.rtol * fmax(maxabs(x1, nn), maxabs(x2, nn));
        double acc = 0.0;
        for (size_t c = 0; c < nn; c++) { const double e = (x2[c] - x1[c]) / tol; acc += e * e; }
        const double error = sqrt(acc / (double)nn);
        /* A NaN error norm rejects every attempt and the reference halves dt
         * without end (integrators.f90:188-199); the update stops instead and
         * is reported by the NaN check. */
        if (isnan(error)) { s->nan_stop = 1; break; }
        if (error < 1) {
            memcpy(x, x2, sizeof(double) * nn);
            memcpy(s->now.w, x, sizeof(double) * nn);
            s->ode_t = s->ode_t + dt;
            s->ode_dt = fmin(dt * pow(1 / error, 1 / s->par.q), s->par.dt_max);
            s->accepts++;
        } else {
            s->ode_dt = dt * 0.5;
        }
    }
}

static int large_check(const double *x, size_t nn, double thr, const char *name) {
    if (maxabs(x, nn) > thr) {
        snprintf(g_err, sizeof g_err, "large_check:: Error: Large values detected in array %s", name);
        return -1;
    }
    return 0;
}

/* isos_update, fastisostasy.f90:731-872.  `stepping` is 0 for the call made
 * from isos_init_state, where the ODE state is not yet initialised and the
 * reference's integrator loop does not run. */
static int isos_update_impl(or_isos *s, const double *H_ice, double time, or_bsl *b, int stepping) {
    const int n = s->n;
    const size_t nn = (size_t)n * n;
    memset(s->now.Hice, 0, sizeof(double) * nn);
    out2in(s, s->now.Hice, H_ice);
    calc_Haf(s, &s->now);
    sl_contributions(s);
    if (b->method == 1) b->bsl_now = b->bsl_now - s->now.deltaV_bsl / b->A_ocean_now;
    if (b->method == 3) {       /* fastisostasy.f90:771-775 */
        s->now.dbsl_total = s->now.dbsl_total + s->now.deltaV_bsl / b->A_ocean_now;
        b->bsl_now = b->bsl_now - s->now.dbsl_total;
    }
    s->now.bsl = b->bsl_now;
    calc_z_ss(nn, s->now.z_ss, s->now.bsl, s->ref.z_ss, s->now.dz_ss);
    int diag = 0;
    if (time - s->time_diagnostics >= F32(1e-3)) { diag = 1; s->time_diagnostics = s->time_diagnostics + s->par.dt_diagnostics; }
    calc_masks(s, &s->now);
    canom_ice(s); canom_seawater(s); canom_load(s);
    double *tmp = (double *)malloc(sizeof(double) * nn);
    if (diag) {
        if (s->par.include_elastic) {
            for (size_t c = 0; c < nn; c++) tmp[c] = s->now.canom_load[c] * s->par.g * pow(s->K[c], 2.0);
            conv_apply(s->conv, s->FGE, tmp, s->now.we);
        } else memset(s->now.we, 0, sizeof(double) * nn);
    }
    canom_litho(s); canom_full(s); mass_anom(s);
    if (diag) {
        if (s->par.heterogeneous_ssh) conv_apply(s->conv, s->FGN, s->now.mass_anom, s->now.dz_ss);
        else memset(s->now.dz_ss, 0, sizeof(double) * nn);
    }
    free(tmp);
    calc_z_ss(nn, s->now.z_ss, s->now.bsl, s->ref.z_ss, s->now.dz_ss);
    calc_masks(s, &s->now);
    canom_seawater(s); canom_load(s);
    if (stepping) integrate(s, time);
    zerobc(n, s->now.w);
    canom_mantle(s);
    for (size_t c = 0; c < nn; c++) s->now.z_bed[c] = s->ref.z_bed[c] + s->now.w[c] + s->now.we[c];
    s->time_prognostics = time;
    /* check_isos_instabilities, fastisostasy.f90:876-884: w, we, w_equilibrium, dz_ss */
    if (large_check(s->now.w, nn, 1e5, "w") || large_check(s->now.we, nn, 1e5, "we") ||
        large_check(s->now.w_eq, nn, 1e5, "w_equilibrium") || large_check(s->now.dz_ss, nn, 1e3, "dz_ss"))
        return -2;
    int nan = s->nan_stop;
    s->nan_stop = 0;
    for (size_t c = 0; c < nn; c++)
        if (isnan(s->now.z_bed[c])) nan = 1;
    if (nan) { set_err("isos_update:: Error: NaNs detected."); return -3; }
    return 0;
}

int or_isos_update(or_isos *s, const double *H_ice, double time, or_bsl *b) {
    return isos_update_impl(s, H_ice, time, b, 1);
}

/* isos_init_state, fastisostasy.f90:596-660 */
/* isos_init_state, fastisostasy.f90:596-660 (restart branch :609-615) */
int or_isos_init_state(or_isos *s, const double *z_bed, const double *H_ice, double time, or_bsl *b) {
    const size_t nn = (size_t)s->n * s->n;
    s->now.bsl = b->bsl_now;
    if (s->use_restart) {
        /* the restart read set ref and now; extendice2isostasy, isos_utils.f90:407-418 */
        s->ref_was_set = 1;
        memset(s->now.Hice, 0, sizeof(double) * nn);
        out2in(s, s->now.z_bed, z_bed);
        out2in(s, s->now.Hice, H_ice);
    } else {
        out2in(s, s->now.z_bed, z_bed);
        out2in(s, s->now.Hice, H_ice);
        if (!s->ref_was_set) or_isos_init_ref(s, z_bed, H_ice, &s->now.bsl, s->now.dz_ss);
    }
    memset(s->ref.z_ss, 0, sizeof(double) * nn);
    calc_z_ss(nn, s->ref.z_ss, s->ref.bsl, s->ref.z_ss, s->ref.dz_ss);
    calc_z_ss(nn, s->now.z_ss, s->now.bsl, s->now.z_ss, s->now.dz_ss);
    calc_Haf(s, &s->ref); calc_Haf(s, &s->now);
    calc_masks(s, &s->ref); calc_masks(s, &s->now);
    s->time_prognostics = time; s->time_diagnostics = time;
    const int rc = isos_update_impl(s, H_ice, time, b, 0);
    s->ode_dt = s->par.dt_init;
    s->ode_t = time;
    memcpy(s->ode_x, s->now.w, sizeof(double) * nn);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Restart files.  The variables, units and grid of the NetCDF restarts
 * (isos_restart_write / _read, isostasy_io.f90:56-194; bsl_restart_write,
 * barysealevel.f90:426-450) in a binary container:
 *   "FIRESTRT", int32 version = 1, int32 nrec, double time,
 *   nrec x { char name[32], char units[16], int32 nx, int32 ny, double v[nx*ny] }
 * little-endian, fields in Fortran (nx, ny) order (x fastest). */
#define RST_MAGIC "FIRESTRT"

typedef struct { const char *name, *units; const double *v; int nx, ny; } rst_rec;

static int rst_write(const char *path, double time, const rst_rec *r, int nrec) {
    FILE *f = fopen(path, "wb");
    if (!f) { set_err("restart: cannot create file"); return -1; }
    const int32_t ver = 1, nr = nrec;
    fwrite(RST_MAGIC, 1, 8, f); fwrite(&ver, 4, 1, f); fwrite(&nr, 4, 1, f); fwrite(&time, 8, 1, f);
    for (int k = 0; k < nrec; k++) {
        char name[32] = {0}, units[16] = {0};
        strncpy(name, r[k].name, sizeof name - 1); strncpy(units, r[k].units, sizeof units - 1);
        const int32_t nx = r[k].nx, ny = r[k].ny;
        fwrite(name, 1, 32, f); fwrite(units, 1, 16, f); fwrite(&nx, 4, 1, f); fwrite(&ny, 4, 1, f);
        fwrite(r[k].v, 8, (size_t)nx * ny, f);
    }
    const int bad = ferror(f);
    fclose(f);
    if (bad) { set_err("restart: write failed"); return -1; }
    return 0;
}

/* the record `name` of shape nx x ny into out (time of the file into *time) */
static int rst_read(const char *path, const char *name, int nx, int ny, double *out, double *time) {
    FILE *f = fopen(path, "rb");
    if (!f) { set_err("restart: cannot open file"); return -1; }
    char magic[8]; int32_t ver, nr; double t;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, RST_MAGIC, 8) || fread(&ver, 4, 1, f) != 1 ||
        fread(&nr, 4, 1, f) != 1 || fread(&t, 8, 1, f) != 1 || ver != 1) {
        fclose(f); set_err("restart: not a restart file"); return -1;
    }
    if (time) *time = t;
    for (int k = 0; k < nr; k++) {
        char nm[32], un[16]; int32_t rx, ry;
        if (fread(nm, 1, 32, f) != 32 || fread(un, 1, 16, f) != 16 || fread(&rx, 4, 1, f) != 1 ||
            fread(&ry, 4, 1, f) != 1) break;
        const size_t cnt = (size_t)rx * ry;
        if (strncmp(nm, name, 32) == 0) {
            if (rx != nx || ry != ny) { fclose(f); set_err("restart: grid size differs"); return -1; }
            const size_t got = fread(out, 8, cnt, f);
            fclose(f);
            if (got != cnt) { set_err("restart: truncated file"); return -1; }
            return 0;
        }
        if (fseek(f, (long)(cnt * 8), SEEK_CUR)) break;
    }
    fclose(f);
    set_err("restart: variable missing");
    return -1;
}

int or_isos_restart_write(const or_isos *s, const char *path, double time) {
    const int n = s->n;
    const size_t nn = (size_t)n * n;
    double *he_km = (double *)malloc(sizeof(double) * nn), *lg = NULL;
    for (size_t c = 0; c < nn; c++) he_km[c] = s->He_lith[c] * (double)1e-3f;
    rst_rec r[13] = {
        {"z_bed", "m", s->now.z_bed, n, n}, {"z_ss", "m", s->now.z_ss, n, n}, {"dz_ss", "m", s->now.dz_ss, n, n},
        {"w", "m", s->now.w, n, n}, {"we", "m", s->now.we, n, n},
        {"dz_ss_ref", "m", s->ref.dz_ss, n, n}, {"H_ice_ref", "m", s->ref.Hice, n, n},
        {"z_bed_ref", "m", s->ref.z_bed, n, n}, {"w_ref", "m", s->ref.w, n, n}, {"we_ref", "m", s->ref.we, n, n},
        {"He_lith", "km", he_km, n, n}, {"mask_active", "1", s->maskactive, n, n}, {NULL, NULL, NULL, 0, 0}};
    int nrec = 12;
    if (s->par.method == 3) {
        lg = (double *)malloc(sizeof(double) * nn);
        for (size_t c = 0; c < nn; c++) lg[c] = log10(s->eta_eff[c]);
        r[12] = (rst_rec){"log10_eta_eff", "Pa s", lg, n, n};
        nrec = 13;
    }
    const int rc = rst_write(path, time, r, nrec);
    free(he_km); free(lg);
    return rc;
}

int or_isos_restart_read(or_isos *s, const char *path, double time) {
    (void)time;
    const int n = s->n;
    struct { const char *name; double *dst; } v[10] = {
        {"z_bed_ref", s->ref.z_bed}, {"H_ice_ref", s->ref.Hice}, {"dz_ss_ref", s->ref.dz_ss},
        {"w_ref", s->ref.w}, {"we_ref", s->ref.we}, {"z_bed", s->now.z_bed}, {"z_ss", s->now.z_ss},
        {"dz_ss", s->now.dz_ss}, {"w", s->now.w}, {"we", s->now.we}};
    for (int k = 0; k < 10; k++)
        if (rst_read(path, v[k].name, n, n, v[k].dst, NULL)) return -1;
    s->use_restart = 1;
    return 0;
}

int or_bsl_restart_write(const or_bsl *b, const char *path, double time) {
    rst_rec r[2] = {{"bsl", "m", &b->bsl_now, 1, 1}, {"A_ocean", "m^2", &b->A_ocean_now, 1, 1}};
    return rst_write(path, time, r, 2);
}

int or_bsl_restart_read(const char *path, double *bsl, double *A_ocean) {
    if (rst_read(path, "bsl", 1, 1, bsl, NULL)) return -1;
    return rst_read(path, "A_ocean", 1, 1, A_ocean, NULL);
}

static const double *field_ptr(const or_isos *s, int f, int *is_mask, const unsigned char **mask) {
    *is_mask = 0; *mask = NULL;
    switch (f) {
    case OR_Z_BED: return s->now.z_bed;
    case OR_Z_SS: return s->now.z_ss;
    case OR_DZ_SS: return s->now.dz_ss;
    case OR_W: return s->now.w;
    case OR_WE: return s->now.we;
    case OR_DWDT: return s->now.dwdt;
    case OR_CANOM_FULL: return s->now.canom_full;
    case OR_HICE: return s->now.Hice;
    case OR_MASKOCEAN: *is_mask = 1; *mask = s->now.maskocean; return NULL;
    case OR_MASKGROUNDED: *is_mask = 1; *mask = s->now.maskgrounded; return NULL;
    case OR_MASKCONTINENT: *is_mask = 1; *mask = s->now.maskcontinent; return NULL;
    case OR_CANOM_LOAD: return s->now.canom_load;
    case OR_CANOM_MANTLE: return s->now.canom_mantle;
    case OR_MASS_ANOM: return s->now.mass_anom;
    case OR_HAF: return s->now.Haf;
    case OR_ETA_EFF: return s->eta_eff;
    case OR_D_LITH: return s->D_lith;
    case OR_KAPPA: return s->kappa;
    case OR_R: return s->R;
    case OR_MASKACTIVE: return s->maskactive;
    case OR_HE_LITH: return s->He_lith;
    case OR_ODE_X: return s->ode_x;
    case OR_CANOM_SEAWATER: return s->now.canom_seawater;
    case OR_CANOM_ICE: return s->now.canom_ice;
    case OR_CANOM_LITHO: return s->now.canom_litho;
    case OR_Z_BED_REF: return s->ref.z_bed;
    case OR_W_REF: return s->ref.w;
    case OR_HAF_REF: return s->ref.Haf;
    case OR_Z_SS_REF: return s->ref.z_ss;
    case OR_W_EQUILIBRIUM: return s->now.w_eq;
    default: return NULL;
    }
}

int or_isos_get_full(const or_isos *s, int f, double *dst) {
    int is_mask; const unsigned char *m;
    const double *p = field_ptr(s, f, &is_mask, &m);
    const size_t nn = (size_t)s->n * s->n;
    if (is_mask) { for (size_t c = 0; c < nn; c++) dst[c] = m[c]; return 0; }
    if (!p) { set_err("unknown field"); return -1; }
    memcpy(dst, p, sizeof(double) * nn);
    return 0;
}

/* cropstate2out, isos_utils.f90:356-405 */
int or_isos_get(const or_isos *s, int f, double *dst) {
    int is_mask; const unsigned char *m;
    const double *p = field_ptr(s, f, &is_mask, &m);
    if (!p && !is_mask) { set_err("unknown field"); return -1; }
    const int n = s->n, ni = s->icrop2 - s->icrop1 + 1, nj = s->jcrop2 - s->jcrop1 + 1;
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
            const size_t c = (size_t)(j + s->jcrop1 - 1) * n + (i + s->icrop1 - 1);
            dst[(size_t)j * ni + i] = is_mask ? (double)m[c] : p[c];
        }
    return 0;
}

double or_isos_scalar(const or_isos *s, int which) {
    switch (which) {
    case OR_S_N: return s->n;
    case OR_S_T: return s->ode_t;
    case OR_S_DT: return s->ode_dt;
    case OR_S_BSL: return s->now.bsl;
    case OR_S_DELTAV: return s->now.deltaV_bsl;
    case OR_S_V_AF: return s->now.V_af;
    case OR_S_V_DEN: return s->now.V_den;
    case OR_S_ATTEMPTS: return (double)s->attempts;
    case OR_S_ACCEPTS: return (double)s->accepts;
    case OR_S_D_LITH11: return s->D_lith[0];
    case OR_S_ICROP1: return s->icrop1;
    case OR_S_JCROP1: return s->jcrop1;
    case OR_S_NX_ICE: return s->nx_ice;
    case OR_S_NY_ICE: return s->ny_ice;
    default: return NAN;
    }
}
