// kelvin.cpp -- host-side init of the LV-ELRA viscous Green's function
// (calc_kei_filter_2D + get_kei_value, kelvin_function.f90:22-131;
// calc_viscous_green, green_functions.f90:28-41).
//
// The reference interpolates a table of kei_0(x), x = 0.00 ... 10.00 in steps
// of 0.01, held to five significant digits as single-precision literals
// (load_kei_values, kelvin_function.f90:167-1190).  The table is rebuilt here
// from the Kelvin function itself -- the series of Abramowitz & Stegun
// 9.9.10-9.9.12 summed in long double -- rounded to the same five digits and
// converted to single precision the way a compiler converts the literal.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fi_kernels.h"

namespace fi {
namespace {

constexpr int KEI_N = 1001;

// kei_0(x) = -ln(x/2) bei(x) - (pi/4) ber(x)
//            + sum_j (-1)^j psi(2j+2) (x^2/4)^(2j+1) / ((2j+1)!)^2,  psi(m+1) = H_m - gamma
long double kelvin_kei(long double x) {
    const long double pi = 3.14159265358979323846264338327950288L;
    if (x == 0.0L) return -pi / 4;
    const long double gamma_e = 0.57721566490153286060651209008240243L;
    const long double q = x * x / 4;
    long double term = 1.0L, harmonic = 0.0L, ber = 1.0L, bei = 0.0L, tail = 0.0L;
    for (int m = 1; m < 80; ++m) {
        term *= q / ((long double)m * (long double)m);   // (x^2/4)^m / (m!)^2
        harmonic += 1.0L / m;
        const long double sign = ((m / 2) & 1) ? -1.0L : 1.0L;
        if (m & 1) {
            bei += sign * term;
            tail += sign * (harmonic - gamma_e) * term;
        } else {
            ber += sign * term;
        }
    }
    return -std::log(x / 2) * bei - pi / 4 * ber + tail;
}

struct KeiTable {
    double rn[KEI_N], kei[KEI_N];
    KeiTable() {
        char buf[64];
        for (int k = 0; k < KEI_N; ++k) {
            std::snprintf(buf, sizeof buf, "%.4f", k / 100.0);
            rn[k] = (double)std::strtof(buf, nullptr);
            std::snprintf(buf, sizeof buf, "%.4Le", kelvin_kei((long double)k / 100.0L));
            kei[k] = (double)std::strtof(buf, nullptr);
        }
    }
    // get_kei_value: linear interpolation in r / L_w, the last entry beyond the
    // table (also at exactly its end, where the reference's search runs past it)
    double at(double r, double L_w) const {
        const double x = r / L_w;
        if (x > rn[KEI_N - 1]) return kei[KEI_N - 1];
        int k = 0;
        while (k < KEI_N - 1 && !(x >= rn[k] && x < rn[k + 1])) ++k;
        if (k == KEI_N - 1) return kei[KEI_N - 1];
        return kei[k] + (x - rn[k]) / (rn[k + 1] - rn[k]) * (kei[k + 1] - kei[k]);
    }
};

}  // namespace

double kelvin_kei_value(double x) { return (double)kelvin_kei((long double)x); }

// GV(x, y) = -L_w^2 / (2 pi D) kei(r / L_w) dx dy on the domain coordinates
// (i - N/2) dx of init_dims (isos_utils.f90:221-240); pi is the reference's
// single-precision literal.
void green_viscous_host(int n, double dx, double L_w, double D_lith, double* out) {
    static const KeiTable table;
    const double pi_ref = (double)3.14159265359f;
    const double c = L_w * L_w / (2.0 * pi_ref * D_lith);
    for (int j = 0; j < n; ++j) {
        const double y = (double)(j - n / 2) * dx;
        for (int i = 0; i < n; ++i) {
            const double x = (double)(i - n / 2) * dx;
            out[(size_t)j * n + i] = -(c * table.at(std::sqrt(x * x + y * y), L_w) * (dx * dx));
        }
    }
}

}  // namespace fi
