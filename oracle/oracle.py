"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the CPU
baseline of bench.py.  Nothing in the product package imports this module.

Every function of fi_oracle.c restates the reference Fortran path
(FastIsostasy src/*.f90) with wp = dp; see the citations there.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MAXL = 8

# field ids (fi_oracle.h)
FIELDS = ["z_bed", "z_ss", "dz_ss", "w", "we", "dwdt", "canom_full", "Hice",
          "maskocean", "maskgrounded", "maskcontinent",
          "canom_load", "canom_mantle", "mass_anom", "Haf", "eta_eff", "D_lith",
          "kappa", "R", "maskactive", "He_lith", "ode_x", "canom_seawater",
          "canom_ice", "canom_litho", "z_bed_ref", "w_ref", "Haf_ref", "z_ss_ref", "w_equilibrium"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
SCALARS = ["n", "t", "dt", "bsl", "deltaV", "V_af", "V_den", "attempts", "accepts",
           "D_lith11", "icrop1", "jcrop1", "nx_ice", "ny_ice"]
SCALAR_ID = {name: i for i, name in enumerate(SCALARS)}

MANTLE = {"uniform": 0, "gaussian_plus": 1, "gaussian_minus": 2, "fields": 3, "rheology_file": 4}
LITHOSPHERE = MANTLE
LAYERING = {"parallel": 0, "equalized": 1}
LUMPING = {"bottom": 0, "min": 1, "max": 2, "mean": 3, "logmean": 4, "freqdomain": 5}
VSCALING = {"none": 0, "min": 1, "scale": 2, "max": 3, "stddev": 4}
DT_METHOD = {"euler": 0, "rk4": 1, "bs32": 2, "tsit54": 3}
BSL_METHOD = {"const": 0, "fastiso": 1, "file": 2, "mixed": 3}


class OrParams(ctypes.Structure):
    _fields_ = [
        ("heterogeneous_ssh", ctypes.c_int), ("interactive_sealevel", ctypes.c_int),
        ("include_elastic", ctypes.c_int), ("correct_compression", ctypes.c_int),
        ("correct_distortion", ctypes.c_int), ("method", ctypes.c_int),
        ("dt_diagnostics", ctypes.c_double), ("min_pad", ctypes.c_double),
        ("mantle", ctypes.c_int), ("lithosphere", ctypes.c_int), ("layering", ctypes.c_int),
        ("lumping", ctypes.c_int), ("viscosity_scaling_method", ctypes.c_int),
        ("viscosity_scaling", ctypes.c_double), ("nl", ctypes.c_int),
        ("zl", ctypes.c_double * MAXL), ("viscosities", ctypes.c_double * MAXL),
        ("dl", ctypes.c_double), ("eta_ref", ctypes.c_double),
        ("rho_water", ctypes.c_double), ("rho_ice", ctypes.c_double),
        ("rho_seawater", ctypes.c_double), ("rho_uppermantle", ctypes.c_double),
        ("rho_litho", ctypes.c_double), ("E", ctypes.c_double), ("nu", ctypes.c_double),
        ("g", ctypes.c_double), ("r_earth", ctypes.c_double), ("m_earth", ctypes.c_double),
        ("dt_method", ctypes.c_int), ("atol", ctypes.c_double), ("rtol", ctypes.c_double),
        ("dt_min", ctypes.c_double), ("dt_max", ctypes.c_double),
        ("dt_init", ctypes.c_double), ("q", ctypes.c_double), ("tau", ctypes.c_double),
    ]


class OrRheologyGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("xc", ctypes.POINTER(ctypes.c_double)), ("yc", ctypes.POINTER(ctypes.c_double)),
        ("litho_thickness", ctypes.POINTER(ctypes.c_double)), ("litho_thickness_scaling", ctypes.c_double),
        ("nz", ctypes.c_int), ("zc", ctypes.POINTER(ctypes.c_double)),
        ("log10_eta", ctypes.POINTER(ctypes.c_double)), ("stddev_log10_eta", ctypes.POINTER(ctypes.c_double)),
        ("rheo_smoothing_radius", ctypes.c_double),
        ("mask_nx", ctypes.c_int), ("mask_ny", ctypes.c_int),
        ("mask_xc", ctypes.POINTER(ctypes.c_double)), ("mask_yc", ctypes.POINTER(ctypes.c_double)),
        ("mask", ctypes.POINTER(ctypes.c_double)),
    ]


def make_grid(struct_type, xc=None, yc=None, litho_thickness=None, litho_thickness_scaling=1.0, zc=None,
              log10_eta=None, stddev_log10_eta=None, rheo_smoothing_radius=0.0, mask_xc=None, mask_yc=None,
              mask=None):
    """Gridded rheology / mask inputs (arrays [ny, nx] or [nz, ny, nx]) in a
    ctypes struct; returns (struct, arrays to keep alive)."""
    keep = {k: (None if v is None else np.ascontiguousarray(v, dtype=np.float64)) for k, v in dict(
        xc=xc, yc=yc, litho_thickness=litho_thickness, zc=zc, log10_eta=log10_eta,
        stddev_log10_eta=stddev_log10_eta, mask_xc=mask_xc, mask_yc=mask_yc, mask=mask).items()}
    g = struct_type()
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for k, v in keep.items():
        setattr(g, k, ptr(v))
    g.nx = 0 if xc is None else len(xc)
    g.ny = 0 if yc is None else len(yc)
    g.nz = 0 if zc is None else len(zc)
    g.mask_nx = 0 if mask_xc is None else len(mask_xc)
    g.mask_ny = 0 if mask_yc is None else len(mask_yc)
    g.litho_thickness_scaling = litho_thickness_scaling
    g.rheo_smoothing_radius = rheo_smoothing_radius
    return g, keep


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.or_isos_create.restype = vp
        L.or_isos_create.argtypes = [ctypes.POINTER(OrParams), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
        L.or_isos_init_ref.argtypes = [vp, dp, dp, dp, dp]
        L.or_isos_init_state.argtypes = [vp, dp, dp, ctypes.c_double, vp]
        L.or_isos_update.argtypes = [vp, dp, ctypes.c_double, vp]
        L.or_isos_get.argtypes = [vp, ctypes.c_int, dp]
        L.or_isos_get_full.argtypes = [vp, ctypes.c_int, dp]
        L.or_isos_scalar.argtypes = [vp, ctypes.c_int]
        L.or_isos_scalar.restype = ctypes.c_double
        L.or_isos_destroy.argtypes = [vp]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_bsl_create.restype = vp
        L.or_bsl_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    dp, dp, ctypes.c_double]
        L.or_bsl_create_series.restype = vp
        L.or_bsl_create_series.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           dp, dp, ctypes.c_int, dp, dp, ctypes.c_double]
        L.or_bsl_update.argtypes = [vp, ctypes.c_double]
        L.or_isos_restart_write.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
        L.or_isos_restart_read.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
        L.or_bsl_restart_write.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
        L.or_bsl_restart_read.argtypes = [ctypes.c_char_p, dp, dp]
        L.or_bsl_now.argtypes = [vp]
        L.or_bsl_now.restype = ctypes.c_double
        L.or_bsl_A_ocean.argtypes = [vp]
        L.or_bsl_A_ocean.restype = ctypes.c_double
        L.or_bsl_destroy.argtypes = [vp]
        L.or_fft.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.or_dht2.argtypes = [ctypes.c_int, dp, dp]
        for nm in ("or_deriv_x", "or_deriv_y", "or_deriv_xx", "or_deriv_yy"):
            getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.or_kappa.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.or_lumping.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]
        L.or_green_elastic.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.or_green_ssh.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.or_isos_create_gridded.restype = vp
        L.or_isos_create_gridded.argtypes = [ctypes.POINTER(OrParams), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, dp, dp,
                                             ctypes.POINTER(OrRheologyGrid)]
        L.or_bilinear.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.or_smooth_gauss.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double]
        L.or_kei.argtypes = [ctypes.c_double]
        L.or_kei.restype = ctypes.c_double
        L.or_kei_table.argtypes = [dp, dp]
        L.or_green_viscous.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.or_convolve.argtypes = [ctypes.c_int, dp, dp, dp]
        L.or_lvelva.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                dp, dp, dp, dp, dp, ctypes.c_double, dp]
        L.or_interp_0d.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double]
        L.or_interp_0d.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_params(**kw):
    """Build OrParams from keyword values (strings for the enumerated options,
    lengths in metres)."""
    p = OrParams()
    enum = {"mantle": MANTLE, "lithosphere": LITHOSPHERE, "layering": LAYERING,
            "lumping": LUMPING, "viscosity_scaling_method": VSCALING, "dt_method": DT_METHOD}
    for k, v in kw.items():
        if k in enum:
            v = enum[k][v] if isinstance(v, str) else v
        if k in ("zl", "viscosities"):
            arr = (ctypes.c_double * MAXL)(*([0.0] * MAXL))
            for i, x in enumerate(v):
                arr[i] = x
            setattr(p, k, arr)
        elif isinstance(getattr(p, k), int) and not isinstance(v, float):
            setattr(p, k, int(v))
        else:
            setattr(p, k, float(v))
    return p


# ---- building blocks ---------------------------------------------------
def fft(x):
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    re, im = _c(x.real), _c(x.imag)
    ro, io = np.empty(n), np.empty(n)
    lib().or_fft(n, _p(re), _p(im), _p(ro), _p(io))
    return ro + 1j * io


def dht2(a):
    a = _c(a)
    out = np.empty_like(a)
    lib().or_dht2(a.shape[0], _p(a), _p(out))
    return out


def deriv(kind, u, dx):
    u, dx = _c(u), _c(np.broadcast_to(dx, u.shape))
    out = np.zeros_like(u)
    getattr(lib(), "or_deriv_" + kind)(u.shape[1], u.shape[0], _p(u), _p(dx), _p(out))
    return out


def kappa(n, dx):
    out = np.empty((n, n))
    lib().or_kappa(n, dx, _p(out))
    return out


def lumping(kind, eta, bnd, kap):
    eta, bnd, kap = _c(eta), _c(bnd), _c(kap)
    nl, n = eta.shape[0], eta.shape[1]
    eff = np.empty((n, n))
    R = np.ones((n, n))
    lib().or_lumping(n, nl, LUMPING[kind], _p(eta), _p(bnd), _p(kap), _p(eff), _p(R))
    return eff, R


def green_elastic(n, dx):
    out = np.empty((n, n))
    lib().or_green_elastic(n, dx, _p(out))
    return out


def green_ssh(n, dx, r_earth, m_earth):
    out = np.empty((n, n))
    lib().or_green_ssh(n, dx, r_earth, m_earth, _p(out))
    return out


def bilinear(x, y, z, xo, yo):
    x, y, z, xo, yo = _c(x), _c(y), _c(z), _c(xo), _c(yo)
    out = np.empty((len(yo), len(xo)))
    lib().or_bilinear(len(x), len(y), _p(x), _p(y), _p(z), len(xo), len(yo), _p(xo), _p(yo), _p(out))
    return out


def smooth_gauss(var, dx, sigma):
    out = _c(var).copy()
    lib().or_smooth_gauss(out.shape[1], out.shape[0], _p(out), dx, sigma)
    return out


def kei(x):
    return lib().or_kei(float(x))


def kei_table():
    rn, kv = np.empty(1001), np.empty(1001)
    lib().or_kei_table(_p(rn), _p(kv))
    return rn, kv


def green_viscous(n, dx, L_w, D_lith):
    out = np.empty((n, n))
    lib().or_green_viscous(n, dx, L_w, D_lith, _p(out))
    return out


def convolve(kernel, x):
    kernel, x = _c(kernel), _c(x)
    out = np.empty_like(x)
    lib().or_convolve(x.shape[0], _p(kernel), _p(x), _p(out))
    return out


def lvelva(w, canom_full, maskactive, g, nu, D, eta, R, kap, dxm, sec_per_year):
    args = [_c(a) for a in (w, canom_full, maskactive, D, eta, R, kap, dxm)]
    out = np.empty_like(args[0])
    w_, c_, m_, D_, e_, R_, k_, d_ = args
    lib().or_lvelva(w_.shape[0], _p(w_), _p(c_), _p(m_), g, nu, _p(D_), _p(e_), _p(R_),
                    _p(k_), _p(d_), sec_per_year, _p(out))
    return out


def interp_0d(x, y, xout):
    x, y = _c(x), _c(y)
    return lib().or_interp_0d(x.size, _p(x), _p(y), xout)


# ---- isos_class / bsl_class mirrors -------------------------------------
class OracleBSL:
    def __init__(self, method="const", bsl_init=0.0, A_ocean_pd=3.625e14, bsl_vec=None,
                 A_vec=None, year_bp=0.0, series=None):
        n = 0 if bsl_vec is None else len(bsl_vec)
        bv = None if bsl_vec is None else _c(bsl_vec)
        av = None if A_vec is None else _c(A_vec)
        if series is None:
            self.h = lib().or_bsl_create(BSL_METHOD[method], bsl_init, A_ocean_pd, n, _p(bv), _p(av), year_bp)
        else:
            st, sv = _c(series[0]), _c(series[1])
            self.h = lib().or_bsl_create_series(BSL_METHOD[method], bsl_init, A_ocean_pd, n, _p(bv), _p(av),
                                                len(st), _p(st), _p(sv), year_bp)
        if not self.h:
            raise RuntimeError(lib().or_last_error().decode())

    def update(self, year_bp):
        lib().or_bsl_update(self.h, year_bp)

    def restart_write(self, path, time):
        if lib().or_bsl_restart_write(self.h, str(path).encode(), time):
            raise RuntimeError(lib().or_last_error().decode())

    @property
    def bsl_now(self):
        return lib().or_bsl_now(self.h)

    @property
    def A_ocean_now(self):
        return lib().or_bsl_A_ocean(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bsl_destroy(self.h)
            self.h = None


class OracleIsos:
    def __init__(self, params, nx_ice, ny_ice, dx, dy=None, K=None, rho_ice=None,
                 he_field=None, eta_field=None, grid=None):
        """grid: keyword arrays of the rheology_file / mask_file inputs (make_grid)."""
        dy = dx if dy is None else dy
        self._keep = [_c(a) if a is not None else None for a in (K, he_field, eta_field)]
        Kc, he, eta = self._keep
        rho = None if rho_ice is None else np.array([rho_ice], dtype=np.float64)
        self._rho = rho
        if grid is not None:
            self._grid = make_grid(OrRheologyGrid, **grid)
            self.h = lib().or_isos_create_gridded(ctypes.byref(params), nx_ice, ny_ice, dx, dy, _p(Kc), _p(rho),
                                                  ctypes.byref(self._grid[0]))
        else:
            self.h = lib().or_isos_create(ctypes.byref(params), nx_ice, ny_ice, dx, dy,
                                          _p(Kc), _p(rho), _p(he), _p(eta))
        if not self.h:
            raise RuntimeError(lib().or_last_error().decode())
        self.n = int(self.scalar("n"))
        self.nx_ice, self.ny_ice = nx_ice, ny_ice

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())

    def init_ref(self, z_bed, H_ice, bsl=None, dz_ss=None):
        b = None if bsl is None else np.array([bsl], dtype=np.float64)
        self._check(lib().or_isos_init_ref(self.h, _p(_c(z_bed)), _p(_c(H_ice)), _p(b),
                                           _p(None if dz_ss is None else _c(dz_ss))))

    def init_state(self, z_bed, H_ice, time, bsl):
        self._check(lib().or_isos_init_state(self.h, _p(_c(z_bed)), _p(_c(H_ice)), time, bsl.h))

    def update(self, H_ice, time, bsl):
        self._check(lib().or_isos_update(self.h, _p(_c(H_ice)), time, bsl.h))

    def restart_write(self, path, time):
        self._check(lib().or_isos_restart_write(self.h, str(path).encode(), time))

    def restart_read(self, path, time):
        self._check(lib().or_isos_restart_read(self.h, str(path).encode(), time))

    def get(self, name):
        out = np.empty((self.ny_ice, self.nx_ice))
        self._check(lib().or_isos_get(self.h, FIELD_ID[name], _p(out)))
        return out

    def get_full(self, name):
        out = np.empty((self.n, self.n))
        self._check(lib().or_isos_get_full(self.h, FIELD_ID[name], _p(out)))
        return out

    def scalar(self, name):
        return lib().or_isos_scalar(self.h, SCALAR_ID[name])

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_isos_destroy(self.h)
            self.h = None


def bsl_restart_read(path):
    """(bsl, A_ocean) of a BSL restart file."""
    b, a = ctypes.c_double(), ctypes.c_double()
    if lib().or_bsl_restart_read(str(path).encode(), ctypes.byref(b), ctypes.byref(a)):
        raise RuntimeError(lib().or_last_error().decode())
    return b.value, a.value


def read_restart_file(path):
    """All records of a restart container: {name: (units, array[ny, nx])}, time."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"FIRESTRT"
    ver, nrec = np.frombuffer(raw, dtype="<i4", count=2, offset=8)
    time = float(np.frombuffer(raw, dtype="<f8", count=1, offset=16)[0])
    assert ver == 1
    out, off = {}, 24
    for _ in range(nrec):
        name = raw[off:off + 32].split(b"\0")[0].decode()
        units = raw[off + 32:off + 48].split(b"\0")[0].decode()
        nx, ny = np.frombuffer(raw, dtype="<i4", count=2, offset=off + 48)
        off += 56
        v = np.frombuffer(raw, dtype="<f8", count=int(nx) * int(ny), offset=off).reshape(int(ny), int(nx))
        off += 8 * int(nx) * int(ny)
        out[name] = (units, v.copy())
    return out, time
