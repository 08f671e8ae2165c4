"""Python mirror of the reference's isos_class / bsl_class API.

Names, argument meaning and call order follow the reference's public Fortran
module (src/fastisostasy.f90:34-60 and src/barysealevel.f90:31-36):

    bsl = bsl_init(method="const", bsl_init=0.0, year_bp=t0)
    isos = isos_init(params, nx_ice, ny_ice, dx, dy)
    isos_init_state(isos, z_bed, H_ice, t0, bsl)
    for t in times:
        bsl_update(bsl, t)
        isos_update(isos, H_ice, t, bsl)
        z_bed = isos.out("z_bed")
    isos_end(isos)

The compute path is the HIP library (libfastiso.so) through its C ABI; a
reference `stop` becomes a raised FiError.  Arrays are float64 [ny][nx].
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FiParams, check, load

MANTLE = {"uniform": 0, "gaussian_plus": 1, "gaussian_minus": 2, "fields": 3, "rheology_file": 4}
LITHOSPHERE = MANTLE
LAYERING = {"parallel": 0, "equalized": 1}
LUMPING = {"bottom": 0, "min": 1, "max": 2, "mean": 3, "logmean": 4, "freqdomain": 5}
VSCALING = {"none": 0, "min": 1, "scale": 2, "max": 3, "stddev": 4}
DT_METHOD = {"euler": 0, "rk4": 1, "bs32": 2, "tsit54": 3}
PRECISION = {"fp64": 0, "mixed": 1}
BSL_METHOD = {"const": 0, "fastiso": 1, "file": 2, "mixed": 3}

_ENUMS = {"mantle": MANTLE, "lithosphere": LITHOSPHERE, "layering": LAYERING, "lumping": LUMPING,
          "viscosity_scaling_method": VSCALING, "dt_method": DT_METHOD, "precision": PRECISION}


def make_params(**kw):
    """FiParams from keyword values; enumerated options may be given as the
    reference's namelist strings.  Lengths are in metres."""
    p = FiParams()
    p.precision = 0
    for k, v in kw.items():
        if k in _ENUMS and isinstance(v, str):
            v = _ENUMS[k][v]
        if k in ("zl", "viscosities"):
            arr = (ctypes.c_double * _lib.MAXL)()
            for i, x in enumerate(v):
                arr[i] = float(x)
            setattr(p, k, arr)
        elif isinstance(getattr(p, k), int):
            setattr(p, k, int(v))
        else:
            setattr(p, k, float(v))
    return p


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def padded_size(nx_ice, ny_ice, dx, dy, min_pad):
    n = ctypes.c_int()
    check(load().fi_padded_size(nx_ice, ny_ice, dx, dy, min_pad, ctypes.byref(n)))
    return n.value


def rheology_grid(xc=None, yc=None, litho_thickness=None, litho_thickness_scaling=1.0, zc=None,
                  log10_eta=None, stddev_log10_eta=None, rheo_smoothing_radius=0.0, mask_xc=None,
                  mask_yc=None, mask=None):
    """fi_rheology_grid from arrays ([ny, nx] / [nz, ny, nx], coordinates in m);
    returns (struct, arrays that must outlive it)."""
    keep = {k: _c(v) for k, v in dict(xc=xc, yc=yc, litho_thickness=litho_thickness, zc=zc,
                                      log10_eta=log10_eta, stddev_log10_eta=stddev_log10_eta,
                                      mask_xc=mask_xc, mask_yc=mask_yc, mask=mask).items()}
    g = _lib.FiRheologyGrid()
    for k, v in keep.items():
        setattr(g, k, _dp(v))
    g.nx, g.ny = (0 if xc is None else len(xc)), (0 if yc is None else len(yc))
    g.nz = 0 if zc is None else len(zc)
    g.mask_nx, g.mask_ny = (0 if mask_xc is None else len(mask_xc)), (0 if mask_yc is None else len(mask_yc))
    g.litho_thickness_scaling = litho_thickness_scaling
    g.rheo_smoothing_radius = rheo_smoothing_radius
    return g, keep


class BslClass:
    """bsl_class (src/barysealevel.f90:16-29): methods const, fastiso, file and
    mixed; `series` = (time_yr, bsl_m) is the sea-level series of sl_path
    (read_series gives it from an ASCII file)."""

    def __init__(self, method="const", bsl_init=0.0, A_ocean_pd=3.625e14, bsl_vec=None, A_vec=None,
                 year_bp=0.0, series=None):
        n = 0 if bsl_vec is None else len(bsl_vec)
        self._bv, self._av = _c(bsl_vec), _c(A_vec)
        h = ctypes.c_void_p()
        m = BSL_METHOD[method] if isinstance(method, str) else int(method)
        if series is None:
            check(load().fi_bsl_create(m, bsl_init, A_ocean_pd, n, _dp(self._bv), _dp(self._av), year_bp,
                                       ctypes.byref(h)))
        else:
            st, sv = _c(series[0]), _c(series[1])
            check(load().fi_bsl_create_series(m, bsl_init, A_ocean_pd, n, _dp(self._bv), _dp(self._av),
                                              len(st), _dp(st), _dp(sv), year_bp, ctypes.byref(h)))
        self.h = h

    @property
    def bsl_now(self):
        v = ctypes.c_double()
        check(load().fi_bsl_get(self.h, ctypes.byref(v), None))
        return v.value

    @property
    def A_ocean_now(self):
        v = ctypes.c_double()
        check(load().fi_bsl_get(self.h, None, ctypes.byref(v)))
        return v.value

    def restart_write(self, path, time):
        """bsl_restart_write (src/barysealevel.f90:426-450)."""
        check(load().fi_bsl_restart_write(self.h, str(path).encode(), float(time)))

    def write_init(self, path, time_init):
        """bsl_write_init (src/barysealevel.f90:453-465): a new output file."""
        check(load().fi_bsl_write_init(self.h, str(path).encode(), float(time_init)))

    def write_step(self, path, time):
        """bsl_write_step (src/barysealevel.f90:467-490): bsl and A_ocean at `time`."""
        check(load().fi_bsl_write_step(self.h, str(path).encode(), float(time)))

    def close(self):
        if getattr(self, "h", None):
            load().fi_bsl_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class IsosClass:
    """isos_class (src/isostasy_defs.f90:202-209) held on the GPU."""

    def __init__(self, params, nx_ice, ny_ice, dx, dy=None, K=None, rho_ice=None, he_field=None,
                 eta_field=None, grid=None):
        """grid: the rheology_file / mask_file arrays as keywords of
        rheology_grid() (xc, yc, litho_thickness, zc, log10_eta, ...)."""
        dy = dx if dy is None else dy
        if isinstance(params, dict):
            params = make_params(**params)
        self.params = params
        keep = [_c(K), _c(he_field), _c(eta_field)]
        rho = None if rho_ice is None else np.array([rho_ice], dtype=np.float64)
        h = ctypes.c_void_p()
        if grid is not None:
            self._grid = rheology_grid(**grid)
            check(load().fi_isos_create_gridded(ctypes.byref(params), nx_ice, ny_ice, dx, dy, _dp(keep[0]),
                                                _dp(rho), ctypes.byref(self._grid[0]), ctypes.byref(h)))
        else:
            check(load().fi_isos_create(ctypes.byref(params), nx_ice, ny_ice, dx, dy, _dp(keep[0]), _dp(rho),
                                        _dp(keep[1]), _dp(keep[2]), ctypes.byref(h)))
        self.h = h
        self.nx_ice, self.ny_ice = nx_ice, ny_ice
        self.n = int(self.scalar("n"))

    def init_ref(self, z_bed, H_ice, bsl=None, dz_ss=None):
        b = None if bsl is None else np.array([bsl], dtype=np.float64)
        check(load().fi_isos_init_ref(self.h, _dp(_c(z_bed)), _dp(_c(H_ice)), _dp(b), _dp(_c(dz_ss))))

    def init_state(self, z_bed, H_ice, time, bsl):
        check(load().fi_isos_init_state(self.h, _dp(_c(z_bed)), _dp(_c(H_ice)), float(time), bsl.h))

    def update(self, H_ice, time, bsl):
        check(load().fi_isos_update(self.h, _dp(_c(H_ice)), float(time), bsl.h))

    def out(self, name):
        """out%<name> on the ice grid (masks as bool)."""
        if name.startswith("mask"):
            a = np.empty((self.ny_ice, self.nx_ice), dtype=np.int8)
            check(load().fi_isos_get_mask(self.h, _lib.FIELD_ID[name],
                                          a.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))))
            return a.astype(bool)
        a = np.empty((self.ny_ice, self.nx_ice))
        check(load().fi_isos_get(self.h, _lib.FIELD_ID[name], _dp(a)))
        return a

    def restart_write(self, path, time):
        """isos_restart_write (src/isostasy_io.f90:56-139)."""
        check(load().fi_isos_restart_write(self.h, str(path).encode(), float(time)))

    def restart_read(self, path, time):
        """isos_restart_read (src/isostasy_io.f90:141-194); the next init_state
        takes the restart branch."""
        check(load().fi_isos_restart_read(self.h, str(path).encode(), float(time)))

    def write_init(self, xc, yc, path, time_init):
        """isos_write_init (src/isostasy_io.f90:197-221): a new output file on
        the ice grid with axes xc, yc (m)."""
        check(load().fi_isos_write_init(self.h, _dp(_c(xc)), _dp(_c(yc)), str(path).encode(), float(time_init)))

    def write_step(self, path, time, H_ice=None, z_bed_bench=None):
        """isos_write_step (src/isostasy_io.f90:318-405): the out% fields and dt
        at `time` (an existing time is overwritten, a new one appended); a
        benchmark solution adds z_bed_bench and err_z_bed."""
        check(load().fi_isos_write_step(self.h, str(path).encode(), float(time), _dp(_c(H_ice)),
                                        _dp(_c(z_bed_bench))))

    def full(self, name):
        """Padded-grid (N x N) state field."""
        a = np.empty((self.n, self.n))
        check(load().fi_isos_get_full(self.h, _lib.FIELD_ID[name], _dp(a)))
        return a

    def scalar(self, name):
        v = ctypes.c_double()
        check(load().fi_isos_get_scalar(self.h, _lib.SCALAR_ID[name], ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            load().fi_isos_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Batch:
    """Many domains of one padded size advanced with one launch sequence."""

    def __init__(self, members):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        check(load().fi_batch_create(arr, len(self.members), ctypes.byref(h)))
        self.h = h

    def update(self, time, bsls, H_ice=None):
        n = len(self.members)
        if H_ice is None:
            # (the pointer arrays of an update without new ice are reused while
            # the bsl objects stay the same)
            key = tuple(b.h.value for b in bsls)
            if getattr(self, "_key", None) != key:
                self._key = key
                self._harr = (ctypes.POINTER(ctypes.c_double) * n)(*([None] * n))
                self._barr = (ctypes.c_void_p * n)(*key)
            check(load().fi_batch_update(self.h, self._harr, float(time), self._barr))
            return
        keep = [_c(a) for a in H_ice]
        Harr = (ctypes.POINTER(ctypes.c_double) * n)(*[_dp(a) for a in keep])
        barr = (ctypes.c_void_p * n)(*[b.h.value for b in bsls])
        check(load().fi_batch_update(self.h, Harr, float(time), barr))

    def close(self):
        if getattr(self, "h", None):
            load().fi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def read_output(path):
    """Read an output file of isos_write_* / bsl_write_*: dict with 'time',
    'xc', 'yc' (km), 'units' and 'long_name' per variable, and each variable
    as an array [time] or [time, ny, nx]."""
    import struct
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != b"FIOUTPUT":
        raise ValueError(f"{path} is not an output file")
    version, nx, ny, nvar, nrec = struct.unpack_from("<5i", raw, 8)
    off = 28
    names, kinds, units, long_names = [], [], {}, {}
    for _ in range(nvar):
        name = raw[off:off + 32].split(b"\0")[0].decode()
        units[name] = raw[off + 32:off + 48].split(b"\0")[0].decode()
        long_names[name] = raw[off + 48:off + 112].split(b"\0")[0].decode()
        kinds.append(struct.unpack_from("<i", raw, off + 112)[0])
        names.append(name)
        off += 116
    xc = np.frombuffer(raw, dtype="<f8", count=nx, offset=off)
    yc = np.frombuffer(raw, dtype="<f8", count=ny, offset=off + 8 * nx)
    off += 8 * (nx + ny)
    per = 1 + sum(nx * ny if k == 0 else 1 for k in kinds)
    recs = np.frombuffer(raw, dtype="<f8", count=per * nrec, offset=off).reshape(nrec, per)
    out = {"time": recs[:, 0].copy(), "xc": xc.copy(), "yc": yc.copy(), "units": units, "long_name": long_names}
    col = 1
    for name, k in zip(names, kinds):
        if k == 0:
            out[name] = recs[:, col:col + nx * ny].reshape(nrec, ny, nx).copy()
            col += nx * ny
        else:
            out[name] = recs[:, col].copy()
            col += 1
    return out


# ---- reference-named procedures (src/fastisostasy.f90:40-50, barysealevel.f90:32-33)
def bsl_init(method="const", bsl_init=0.0, A_ocean_pd=3.625e14, year_bp=0.0, bsl_vec=None, A_vec=None,
             series=None):
    return BslClass(method, bsl_init, A_ocean_pd, bsl_vec, A_vec, year_bp, series)


def bsl_restart_read(path):
    """(bsl, A_ocean) of a BSL restart file; bsl_init of the const method takes
    bsl from it (src/barysealevel.f90:72-73)."""
    b, a = ctypes.c_double(), ctypes.c_double()
    check(load().fi_bsl_restart_read(str(path).encode(), ctypes.byref(b), ctypes.byref(a)))
    return b.value, a.value


def read_series(path, max_n=10000):
    """read_series (src/barysealevel.f90:254-315): (time, value) arrays of an
    ASCII series file."""
    t, v, n = np.empty(max_n), np.empty(max_n), ctypes.c_int()
    check(load().fi_bsl_read_series(str(path).encode(), max_n, _dp(t), _dp(v), ctypes.byref(n)))
    return t[:n.value].copy(), v[:n.value].copy()


def bsl_update(bsl, year_bp):
    check(load().fi_bsl_update(bsl.h, float(year_bp)))


def isos_init(params, nx_ice, ny_ice, dx, dy=None, rho_ice=None, K=None, he_field=None, eta_field=None):
    return IsosClass(params, nx_ice, ny_ice, dx, dy, K=K, rho_ice=rho_ice, he_field=he_field,
                     eta_field=eta_field)


def isos_init_ref(isos, z_bed, H_ice, bsl=None, dz_ss=None):
    isos.init_ref(z_bed, H_ice, bsl, dz_ss)


def isos_init_state(isos, z_bed, H_ice, time, bsl):
    isos.init_state(z_bed, H_ice, time, bsl)


def isos_update(isos, H_ice, time, bsl):
    isos.update(H_ice, time, bsl)


def isos_end(isos):
    isos.close()
