"""Golden vectors: analytical ELVA disc displacement (Bueler et al. 2007).

Restates the known-answer benchmark of the reference
(src/isostasy_benchmarks.f90:74-127 and the integrand at :335-367):

  w(r, t) = int_0^kmax rho_i g h0 r0 (exp(-beta t / (2 eta k)) - 1)
                        J1(k r0) J0(k r) / beta dk,  beta = rho_m g + D k^4,

with t in seconds = t_yr * 365 * 24 * 3600 (365-day years, :349).  Evaluated
with scipy.special j0/j1 and scipy.integrate.quad for the plain-disc
parameters (r0 = 1000 km, h0 = 1000 m, eta = 1e21, He = 88 km, E = 66 GPa,
nu = 0.28, rho_i = 910, rho_m = 3300, g = 9.81).

    python tests/golden/make_elva_disc.py   # writes elva_disc_analytic.json
"""
import json
import os

import numpy as np
from scipy import integrate, special

E, nu, He = 66.0, 0.28, 88e3
D = E * 1e9 * He ** 3 / (12.0 * (1.0 - nu ** 2))
P = dict(r0=1000e3, h0=1000.0, eta=1e21, D=D, rho_i=910.0, rho_m=3300.0, g=9.81)


def w_disc(r, t_yr, r0, h0, eta, D, rho_i, rho_m, g):
    ts = t_yr * 365 * 24 * 3600

    def f(k):
        beta = rho_m * g + D * k ** 4
        return rho_i * g * h0 * r0 * (np.exp(-beta * ts / (2 * eta * k)) - 1.0) * \
            special.j1(k * r0) * special.j0(k * r) / beta
    pts = np.linspace(1e-7, 1e-4, 60)[1:-1]
    val, _ = integrate.quad(f, 1e-12, 1e-4, limit=4000, points=pts)
    return val


if __name__ == "__main__":
    times = [1000.0, 2000.0, 3000.0, 5000.0, 50000.0]
    radii = [0.0, 500e3, 900e3, 1100e3, 1500e3]
    table = {"params": P, "times_yr": times, "radii_m": radii,
             "w_m": [[w_disc(r, t, **P) for r in radii] for t in times]}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elva_disc_analytic.json")
    with open(out, "w") as fh:
        json.dump(table, fh, indent=1)
    print(out)
