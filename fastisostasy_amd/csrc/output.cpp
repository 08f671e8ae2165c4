// output.cpp -- the output writers isos_write_init / isos_write_step
// (src/isostasy_io.f90:197-405) and bsl_write_init / bsl_write_step
// (src/barysealevel.f90:453-494) on top of the public C ABI.
//
// The reference writes NetCDF through ncio; no NetCDF library is available
// here, so the same variables (names, units, long names, the unlimited time
// axis with its time index rule) go into a self-describing binary container:
//
//   char    magic[8]   "FIOUTPUT"
//   int32   version    1
//   int32   nx, ny     grid of the 2-D variables (0, 0 for a scalar series)
//   int32   nvar
//   int32   nrec       records on the time axis (updated in place)
//   nvar x { char name[32]; char units[16]; char long_name[64]; int32 kind }
//                      kind 0: nx * ny values per record (x fastest), 1: one value
//   double  xc[nx], yc[ny]   grid axes in km
//   nrec x { double time; the values of every variable, in declaration order }
//
// little-endian.  isos_write_init creates the axis with the initial time and
// no data (NaN, as a NetCDF fill value); isos_write_step writes at the index
// of `time` on the axis (an equal time is overwritten, a new time appended),
// which is nc_time_index's rule.  The first step with a benchmark solution
// adds z_bed_bench and err_z_bed to the file.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastiso.h"
#include "fi_kernels.h"

namespace {

constexpr char kMagic[9] = "FIOUTPUT";
constexpr long kNrecOffset = 8 + 4 * 4;   // magic, version, nx, ny, nvar

struct Var {
    const char* name;
    const char* units;
    const char* long_name;
    int kind;   // 0 field, 1 scalar
};

// isos_write_step (isostasy_io.f90:334-386): dt and the out%* fields
const Var kIsosVars[] = {
    {"dt", "yr", "Time step", 1},
    {"H_ice", "m", "Ice thickness", 0},
    {"z_ss", "m", "Sea-surface height", 0},
    {"z_bed", "m", "Bedrock elevation", 0},
    {"dwdt", "m/yr", "Bedrock elevation change", 0},
    {"w_viscous", "m", "Displacement (viscous)", 0},
    {"w_elastic", "m", "Displacement (elastic)", 0},
    {"z_ss_perturbation", "m", "Geoid displacement", 0},
    {"column_anomaly", "N m^-2", "Anomaly in column pressure", 0},
    {"ocean_mask", "1", "Ocean mask", 0},
    {"grounded_mask", "1", "Grounded mask", 0},
    {"continent_mask", "1", "Continent mask", 0},
};
// the C ABI field behind each 2-D variable above (after dt)
const int kIsosFields[] = {FI_HICE, FI_Z_SS, FI_Z_BED, FI_DWDT, FI_W, FI_WE, FI_DZ_SS, FI_CANOM_FULL,
                           FI_MASKOCEAN, FI_MASKGROUNDED, FI_MASKCONTINENT};
// with a benchmark solution (isostasy_io.f90:388-397)
const Var kBenchVars[] = {
    {"z_bed_bench", "m", "Benchmark bedrock elevation", 0},
    {"err_z_bed", "m", "Error in bedrock elevation", 0},
};
// bsl_write_step (barysealevel.f90:466-488)
const Var kBslVars[] = {
    {"bsl", "m", "Barystatic sea level", 1},
    {"A_ocean", "m^2", "Ocean surface area", 1},
};

// errors go to the library's fi_last_error()
int err(int code, const std::string& msg) { return fi::set_error(code, msg.c_str()); }

struct Header {
    int32_t nx = 0, ny = 0, nvar = 0, nrec = 0;
    std::vector<std::string> names;
    std::vector<int32_t> kind;
    long data_off = 0;       // first record
    size_t rec_doubles = 0;  // doubles per record, time included
};

int create(const char* path, int nx, int ny, const std::vector<Var>& vars, const double* xc, const double* yc,
           double time_init) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return err(FI_E_ARG, std::string("output: cannot create ") + path);
    const int32_t head[5] = {1, nx, ny, (int32_t)vars.size(), 1};
    std::fwrite(kMagic, 1, 8, f);
    std::fwrite(head, 4, 5, f);
    size_t per = 1;
    for (const Var& v : vars) {
        char name[32] = {}, units[16] = {}, ln[64] = {};
        std::strncpy(name, v.name, sizeof name - 1);
        std::strncpy(units, v.units, sizeof units - 1);
        std::strncpy(ln, v.long_name, sizeof ln - 1);
        const int32_t kind = v.kind;
        std::fwrite(name, 1, 32, f);
        std::fwrite(units, 1, 16, f);
        std::fwrite(ln, 1, 64, f);
        std::fwrite(&kind, 4, 1, f);
        per += kind == 0 ? (size_t)nx * ny : 1;
    }
    for (int i = 0; i < nx; ++i) { const double v = xc[i] * 1e-3; std::fwrite(&v, 8, 1, f); }
    for (int j = 0; j < ny; ++j) { const double v = yc[j] * 1e-3; std::fwrite(&v, 8, 1, f); }
    // the time axis starts with time_init and no data
    std::vector<double> rec(per, std::nan(""));
    rec[0] = time_init;
    std::fwrite(rec.data(), 8, rec.size(), f);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    return bad ? err(FI_E_ARG, std::string("output: write to ") + path + " failed") : 0;
}

int read_header(FILE* f, Header& h, const char* path) {
    char magic[8];
    int32_t head[5];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0 || std::fread(head, 4, 5, f) != 5 ||
        head[0] != 1 || head[1] < 0 || head[2] < 0 || head[3] < 0 || head[4] < 0)
        return err(FI_E_ARG, std::string("output: ") + path + " is not an output file");
    h.nx = head[1];
    h.ny = head[2];
    h.nvar = head[3];
    h.nrec = head[4];
    h.rec_doubles = 1;
    for (int k = 0; k < h.nvar; ++k) {
        char name[33] = {}, units[17] = {}, ln[65] = {};
        int32_t kind;
        if (std::fread(name, 1, 32, f) != 32 || std::fread(units, 1, 16, f) != 16 || std::fread(ln, 1, 64, f) != 64 ||
            std::fread(&kind, 4, 1, f) != 1 || (kind != 0 && kind != 1))
            return err(FI_E_ARG, std::string("output: ") + path + " is truncated");
        h.names.push_back(name);
        h.kind.push_back(kind);
        h.rec_doubles += kind == 0 ? (size_t)h.nx * h.ny : 1;
    }
    h.data_off = std::ftell(f) + (long)(8 * ((size_t)h.nx + h.ny));
    return 0;
}

// Write one record at the time index of `time` (nc_time_index): the record
// whose time equals it, else a new one at the end.
int write_record(const char* path, const std::vector<double>& rec, int nx, int ny, size_t nvar) {
    FILE* f = std::fopen(path, "r+b");
    if (!f) return err(FI_E_ARG, std::string("output: cannot open ") + path);
    Header h;
    int rc = read_header(f, h, path);
    if (rc) { std::fclose(f); return rc; }
    if (h.nx != nx || h.ny != ny || (size_t)h.nvar != nvar || h.rec_doubles != rec.size()) {
        std::fclose(f);
        return err(FI_E_ARG, std::string("output: ") + path + " was created for another grid or variable set");
    }
    const size_t rb = 8 * h.rec_doubles;
    int32_t idx = h.nrec;
    for (int32_t n = 0; n < h.nrec; ++n) {
        double t;
        if (std::fseek(f, h.data_off + (long)(rb * n), SEEK_SET) != 0 || std::fread(&t, 8, 1, f) != 1) {
            std::fclose(f);
            return err(FI_E_ARG, std::string("output: ") + path + " is truncated");
        }
        if (t == rec[0]) { idx = n; break; }
    }
    std::fseek(f, h.data_off + (long)(rb * idx), SEEK_SET);
    std::fwrite(rec.data(), 8, rec.size(), f);
    if (idx == h.nrec) {
        const int32_t nrec = h.nrec + 1;
        std::fseek(f, kNrecOffset, SEEK_SET);
        std::fwrite(&nrec, 4, 1, f);
    }
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    return bad ? err(FI_E_ARG, std::string("output: write to ") + path + " failed") : 0;
}

int peek_header(const char* path, Header& h) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return err(FI_E_ARG, std::string("output: cannot open ") + path);
    const int rc = read_header(f, h, path);
    std::fclose(f);
    return rc;
}

// Rewrite a file with `extra` variables appended to its declarations; the
// records already written hold NaN for them.
int add_vars(const char* path, const Header& h, const std::vector<Var>& extra) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return err(FI_E_ARG, std::string("output: cannot open ") + path);
    std::vector<char> decl((size_t)h.nvar * 116);
    std::vector<double> axes((size_t)h.nx + h.ny), recs(h.rec_doubles * h.nrec);
    const long decl_off = 8 + 5 * 4;
    bool ok = std::fseek(f, decl_off, SEEK_SET) == 0 && std::fread(decl.data(), 1, decl.size(), f) == decl.size() &&
              std::fread(axes.data(), 8, axes.size(), f) == axes.size() &&
              std::fread(recs.data(), 8, recs.size(), f) == recs.size();
    std::fclose(f);
    if (!ok) return err(FI_E_ARG, std::string("output: ") + path + " is truncated");
    size_t add = 0;
    for (const Var& v : extra) add += v.kind == 0 ? (size_t)h.nx * h.ny : 1;
    f = std::fopen(path, "wb");
    if (!f) return err(FI_E_ARG, std::string("output: cannot rewrite ") + path);
    const int32_t head[5] = {1, h.nx, h.ny, h.nvar + (int32_t)extra.size(), h.nrec};
    std::fwrite(kMagic, 1, 8, f);
    std::fwrite(head, 4, 5, f);
    std::fwrite(decl.data(), 1, decl.size(), f);
    for (const Var& v : extra) {
        char name[32] = {}, units[16] = {}, ln[64] = {};
        std::strncpy(name, v.name, sizeof name - 1);
        std::strncpy(units, v.units, sizeof units - 1);
        std::strncpy(ln, v.long_name, sizeof ln - 1);
        const int32_t kind = v.kind;
        std::fwrite(name, 1, 32, f);
        std::fwrite(units, 1, 16, f);
        std::fwrite(ln, 1, 64, f);
        std::fwrite(&kind, 4, 1, f);
    }
    std::fwrite(axes.data(), 8, axes.size(), f);
    const std::vector<double> fill(add, std::nan(""));
    for (int32_t n = 0; n < h.nrec; ++n) {
        std::fwrite(recs.data() + (size_t)n * h.rec_doubles, 8, h.rec_doubles, f);
        std::fwrite(fill.data(), 8, fill.size(), f);
    }
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    return bad ? err(FI_E_ARG, std::string("output: rewrite of ") + path + " failed") : 0;
}

int ice_grid(fi_isos* h, int& nx, int& ny) {
    double a, b;
    int rc = fi_isos_get_scalar(h, FI_S_NX_ICE, &a);
    if (!rc) rc = fi_isos_get_scalar(h, FI_S_NY_ICE, &b);
    if (rc) return rc;
    nx = (int)a;
    ny = (int)b;
    return 0;
}

}  // namespace

extern "C" {

int fi_isos_write_init(fi_isos* h, const double* xc, const double* yc, const char* path, double time_init) {
    if (!h || !xc || !yc || !path) return err(FI_E_ARG, "NULL argument");
    int nx, ny;
    int rc = ice_grid(h, nx, ny);
    if (rc) return rc;
    std::vector<Var> vars(std::begin(kIsosVars), std::end(kIsosVars));
    return create(path, nx, ny, vars, xc, yc, time_init);
}

int fi_isos_write_step(fi_isos* h, const char* path, double time, const double* H_ice, const double* z_bed_bench) {
    (void)H_ice;   // as in the reference, the file gets out%Hice (isostasy_io.f90:337-338)
    if (!h || !path) return err(FI_E_ARG, "NULL argument");
    int nx, ny;
    int rc = ice_grid(h, nx, ny);
    if (rc) return rc;
    const size_t nc = (size_t)nx * ny;
    const size_t nbase = sizeof kIsosVars / sizeof kIsosVars[0];
    Header hd;
    rc = peek_header(path, hd);
    if (rc) return rc;
    if (hd.nx != nx || hd.ny != ny)
        return err(FI_E_ARG, std::string("output: ") + path + " was created for another grid");
    bool bench = (size_t)hd.nvar == nbase + 2;
    if (z_bed_bench && !bench) {
        // the first benchmark solution adds its two variables (NetCDF would
        // define them at their first write): earlier records get NaN
        rc = add_vars(path, hd, std::vector<Var>(std::begin(kBenchVars), std::end(kBenchVars)));
        if (rc) return rc;
        bench = true;
        hd.nvar += 2;
    }
    std::vector<double> rec;
    rec.reserve(1 + 1 + nc * (nbase + 1));
    rec.push_back(time);
    double dt;
    rc = fi_isos_get_scalar(h, FI_S_DT, &dt);
    if (rc) return rc;
    rec.push_back(dt);
    std::vector<double> buf(nc), w(nc);
    for (int fld : kIsosFields) {
        rc = fi_isos_get(h, fld, buf.data());
        if (rc) return rc;
        if (fld == FI_W) w = buf;
        rec.insert(rec.end(), buf.begin(), buf.end());
    }
    if (bench) {
        for (size_t c = 0; c < nc; ++c) rec.push_back(z_bed_bench ? z_bed_bench[c] : std::nan(""));
        // err_z_bed = out%w - z_bed_bench (isostasy_io.f90:392-394)
        for (size_t c = 0; c < nc; ++c) rec.push_back(z_bed_bench ? w[c] - z_bed_bench[c] : std::nan(""));
    }
    return write_record(path, rec, nx, ny, (size_t)hd.nvar);
}

int fi_bsl_write_init(fi_bsl* b, const char* path, double time_init) {
    if (!b || !path) return err(FI_E_ARG, "NULL argument");
    std::vector<Var> vars(std::begin(kBslVars), std::end(kBslVars));
    return create(path, 0, 0, vars, nullptr, nullptr, time_init);
}

int fi_bsl_write_step(fi_bsl* b, const char* path, double time) {
    if (!b || !path) return err(FI_E_ARG, "NULL argument");
    double bsl, area;
    int rc = fi_bsl_get(b, &bsl, &area);
    if (rc) return rc;
    const std::vector<double> rec{time, bsl, area};
    return write_record(path, rec, 0, 0, 2);
}

}  // extern "C"
