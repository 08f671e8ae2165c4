// init.hip -- init-time fields of isos_init on the GPU (FP64).
//
//   k_kappa       calc_kappa (lv_elva.f90:442-481): period (n-1) dx, kappa(1,1)
//                 averaged, pi as the reference's single-precision literal
//   k_gauss       Gaussian shape of calc_gaussian_{thickness,viscosity}
//                 (isos_utils.f90:960-1035)
//   k_rheo        lithosphere thickness, rigidity (isos_utils.f90:79-100),
//                 layer boundaries (lv_elva.f90:131-168) and viscosity lumping
//                 (lv_elva.f90:207-306)
//   k_vfinal      viscosity scaling, compressibility and the
//                 exp(log10(eta_ref/eta)) transform (fastisostasy.f90:425-442),
//                 then the RHS-precision copies D, eta, kappa*R
//   k_green       elastic (Farrell table) and sea-surface Green's functions
//                 (green_functions.f90:43-181)
#include "fi_device.hpp"
#include "fi_kernels.h"

namespace fi {

__device__ __constant__ float c_rn[42] = {0.0f, 0.011f, 0.111f, 1.112f, 2.224f, 3.336f, 4.448f, 6.672f, 8.896f,
    11.12f, 17.79f, 22.24f, 27.80f, 33.36f, 44.48f, 55.60f, 66.72f, 88.96f, 111.2f, 133.4f, 177.9f, 222.4f,
    278.0f, 333.6f, 444.8f, 556.0f, 667.2f, 778.4f, 889.6f, 1001.0f, 1112.0f, 1334.0f, 1779.0f, 2224.0f,
    2780.0f, 3336.0f, 4448.0f, 5560.0f, 6672.0f, 7784.0f, 8896.0f, 10008.0f};
__device__ __constant__ float c_ge[42] = {-33.6488f, -33.64f, -33.56f, -32.75f, -31.86f, -30.98f, -30.12f,
    -28.44f, -26.87f, -25.41f, -21.80f, -20.02f, -18.36f, -17.18f, -15.71f, -14.91f, -14.41f, -13.69f,
    -13.01f, -12.31f, -10.95f, -9.757f, -8.519f, -7.533f, -6.131f, -5.237f, -4.660f, -4.272f, -3.999f,
    -3.798f, -3.640f, -3.392f, -2.999f, -2.619f, -2.103f, -1.530f, -0.292f, 0.848f, 1.676f, 2.083f,
    2.057f, 1.643f};

__device__ __forceinline__ double kappa_at(int i, int j, int n, double dx) {
    const double pi_ref = (double)3.14159265359f;
    const double mu = 2.0 * pi_ref / ((n - 1) * dx);
    const int ic = (n - 1) / 2 + 1;
    const int ii = i + 1, jj = j + 1;
    const double p = (ii <= ic) ? mu * (ii - 1) : mu * (n - ii + 1);
    const double q = (jj <= ic) ? mu * (jj - 1) : mu * (n - jj + 1);
    return sqrt(p * p + q * q);
}

__global__ void k_kappa(int n, double dx, double* __restrict__ kappa) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)n * n) return;
    const int i = (int)(c % n), j = (int)(c / n);
    double v = kappa_at(i, j, n, dx);
    if (c == 0) v = (kappa_at(0, 1, n, dx) + kappa_at(1, 0, n, dx)) / 2.0;
    kappa[c] = v;
}

// exp(-0.5 (dx' S dx)) with S = diag((4/Lx)^2, (4/Ly)^2) on the padded grid
__global__ void k_gauss(int n, double dx, double* __restrict__ out) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)n * n) return;
    const int i = (int)(c % n), j = (int)(c / n);
    const double xmin = (double)(1 - n / 2 - 1) * dx, xmax = (double)(n - n / 2 - 1) * dx;
    const double xcn = (xmax + xmin) / 2.0;
    const double Lx = (xmax - xmin) / 2.;
    const double s = (4.0 / Lx) * (4.0 / Lx);
    const double ddx = (double)(i + 1 - n / 2 - 1) * dx - xcn, ddy = (double)(j + 1 - n / 2 - 1) * dx - xcn;
    out[c] = exp(-0.5 * (ddx * (s * ddx) + ddy * (s * ddy)));
}

// single-block reductions (init only)
__global__ void __launch_bounds__(1024) k_reduce(const double* __restrict__ x, size_t nn, int op, double* out) {
    __shared__ double red[16];
    double v = (op == 1) ? HUGE_VAL : -HUGE_VAL;
    for (size_t c = threadIdx.x; c < nn; c += blockDim.x) v = (op == 1) ? fmin(v, x[c]) : fmax(v, x[c]);
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (op == 1) ? fmin(v, w) : fmax(v, w);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) v = (op == 1) ? fmin(v, red[q]) : fmax(v, red[q]);
        *out = v;
    }
}

__global__ void k_rheo(InitArgs a, const double* __restrict__ kappa, const double* __restrict__ gshape,
                       const double* __restrict__ gmax, const double* __restrict__ he_f,
                       const double* __restrict__ eta_f, double* __restrict__ etaeff,
                       double* __restrict__ R, double* __restrict__ He, double* __restrict__ Dd) {
    const int n = a.n;
    const size_t nn = (size_t)n * n;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nn) return;
    // lithosphere (fastisostasy.f90:166-172, 246-301)
    double he = a.zl[0];
    if (a.lithosphere == 1 || a.lithosphere == 2) {
        const double sign = (a.lithosphere == 1) ? 1.0 : -1.0;
        he = a.zl[0] + sign * (double)100e3f * (gshape[c] / *gmax);
    } else if (a.lithosphere == 3) {
        he = he_f[c];
    }
    He[c] = he;
    Dd[c] = a.E * 1e9 * (he * he * he) / (12.0 * (1.0 - a.nu * a.nu));
    // layer boundaries and layer viscosities
    double bnd[8], eta[8];
    bnd[0] = he;
    for (int k = 1; k < a.nl; ++k) bnd[k] = (a.layering == 0) ? bnd[k - 1] + a.dl : a.zl[k];
    for (int k = 0; k < a.nl; ++k) eta[k] = (a.mantle == 3) ? eta_f[k * nn + c] : a.visc[k];
    double e, r = 1.0;
    if (a.mantle == 1 || a.mantle == 2) {
        const double sign = (a.mantle == 1) ? 1.0 : -1.0;
        e = a.visc[0] * pow(10.0, sign * gshape[c]);
    } else {
        const int nl = a.nl;
        switch (a.lumping) {
        case 0: e = eta[nl - 1]; break;
        case 1: e = eta[0]; for (int k = 1; k < nl; ++k) e = fmin(e, eta[k]); break;
        case 2: e = eta[0]; for (int k = 1; k < nl; ++k) e = fmax(e, eta[k]); break;
        case 3: e = 0; for (int k = 0; k < nl; ++k) e += eta[k]; e = e / (double)nl; break;
        case 4: e = 0; for (int k = 0; k < nl; ++k) e += log10(eta[k]); e = pow(10.0, e / (double)nl); break;
        default: {
            e = eta[nl - 1];
            const double kap = kappa[c];
            for (int k = nl - 1; k >= 1; --k) {
                const double dz = bnd[k] - bnd[k - 1];
                const double ratio = eta[k - 1] / e;
                const double inv = 1. / ratio;
                const double ch = cosh(dz * kap), sh = sinh(dz * kap);
                const double dzk = dz * kap;
                const double num = 2 * ratio * ch * sh + (1 - ratio * ratio) * (dzk * dzk) + (ratio * sh) * (ratio * sh) + ch * ch;
                const double den = (ratio + inv) * ch * sh + (ratio - inv) * dz * kap + sh * sh + ch * ch;
                r = r * num / den;
            }
        }
        }
    }
    etaeff[c] = e;
    R[c] = r;
}

template <typename T>
__global__ void k_vfinal(InitArgs a, const double* __restrict__ red, double* __restrict__ etaeff,
                         const double* __restrict__ kappa, const double* __restrict__ R,
                         const double* __restrict__ Dd, T* __restrict__ D_t, T* __restrict__ eta_t,
                         T* __restrict__ kR_t) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)a.n * a.n) return;
    double e = etaeff[c];
    if (a.vscale_method == 1 || a.vscale_method == 3) e = *red;
    else if (a.vscale_method == 2) e = e * a.vscale;
    e = e * a.comp_corr;
    e = exp(log10(a.eta_ref / e)) * e;
    etaeff[c] = e;
    D_t[c] = (T)Dd[c];
    if (sizeof(T) == 8) eta_t[c] = (T)e;
    else eta_t[c] = (T)(1.0 / (2.0 * e));
    // mixed: the reciprocal, so the FP32 column pass multiplies instead of dividing
    if (sizeof(T) == 8) kR_t[c] = (T)(R[c] * kappa[c]);
    else kR_t[c] = (T)(1.0 / (R[c] * kappa[c]));
}

__global__ void k_green(int which, int n, double dx, double r_earth, double m_earth, double* __restrict__ out) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)n * n) return;
    const int i = (int)(c % n), j = (int)(c / n);
    const int mx = (n % 2 == 0) ? n / 2 : (n - 1) / 2;
    const double dy = dx;
    const double x = (i - mx) * dx, y = (j - mx) * dy;
    const double r = sqrt(x * x + y * y);
    if (which == 0) {
        double ge;
        const double rmax = (double)(c_rn[41] * 1.e3f);
        if (r > rmax) ge = 0.;
        else {
            int k;
            for (k = 0; k < 41; ++k) {
                const double a = (double)(c_rn[k] * 1.e3f), b = (double)(c_rn[k + 1] * 1.e3f);
                if (r >= a && r < b) break;
            }
            if (k == 41) k = 40;
            const double rk = (double)(c_rn[k] * 1.e3f), rk1 = (double)(c_rn[k + 1] * 1.e3f);
            ge = (double)c_ge[k] + (r - rk) / (rk1 - rk) * ((double)c_ge[k + 1] - (double)c_ge[k]);
        }
        out[c] = ge * (double)1.e-12f * (dx * dy) / ((double)9.81f * fmax(r, dx));
    } else {
        const double rr = fmax(dy / 2, r);
        out[c] = r_earth / (2. * m_earth * sin(rr / (2. * r_earth)));
    }
}

// ------------------------------------------------------------------ launchers
static dim3 cells_grid(int n) { return dim3((unsigned)(((size_t)n * n + 255) / 256)); }

void launch_init_rheology(const Member* d_mem, const InitArgs& a, const double* d_he, const double* d_eta,
                          double* d_etaeff, double* d_kappa, double* d_R, double* d_He, double* d_Dd,
                          double* d_scratch, bool fp64, hipStream_t st) {
    (void)d_mem;
    const int n = a.n;
    const size_t nn = (size_t)n * n;
    double* gshape = d_scratch;           // nn
    double* gmax = d_scratch + nn;        // 1
    double* red = d_scratch + nn + 1;     // 1
    hipLaunchKernelGGL(k_kappa, cells_grid(n), dim3(256), 0, st, n, a.dx, d_kappa);
    if (a.lithosphere == 1 || a.lithosphere == 2 || a.mantle == 1 || a.mantle == 2) {
        hipLaunchKernelGGL(k_gauss, cells_grid(n), dim3(256), 0, st, n, a.dx, gshape);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, (const double*)gshape, nn, 2, gmax);
    }
    hipLaunchKernelGGL(k_rheo, cells_grid(n), dim3(256), 0, st, a, (const double*)d_kappa, (const double*)gshape,
                       (const double*)gmax, d_he, d_eta, d_etaeff, d_R, d_He, d_Dd);
    (void)red;
}

void launch_init_finalize(const InitArgs& a, const double* d_scratch, double* d_etaeff, const double* d_kappa,
                          const double* d_R, const double* d_Dd, void* D_t, void* eta_t, void* kR_t,
                          bool fp64, hipStream_t st) {
    const size_t nn = (size_t)a.n * a.n;
    double* red = const_cast<double*>(d_scratch) + nn + 1;
    // min / max scaling of the (possibly smoothed) eta_eff (fastisostasy.f90:413-436)
    if (a.vscale_method == 1 || a.vscale_method == 3)
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, (const double*)d_etaeff, nn,
                           a.vscale_method == 1 ? 1 : 2, red);
    if (fp64)
        hipLaunchKernelGGL(k_vfinal<double>, cells_grid(a.n), dim3(256), 0, st, a, red, d_etaeff, d_kappa, d_R,
                           d_Dd, (double*)D_t, (double*)eta_t, (double*)kR_t);
    else
        hipLaunchKernelGGL(k_vfinal<float>, cells_grid(a.n), dim3(256), 0, st, a, red, d_etaeff, d_kappa, d_R,
                           d_Dd, (float*)D_t, (float*)eta_t, (float*)kR_t);
}

void launch_green(int which, int n, double dx, double r_earth, double m_earth, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_green, cells_grid(n), dim3(256), 0, st, which, n, dx, r_earth, m_earth, d_out);
}

}  // namespace fi
