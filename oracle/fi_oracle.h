/*
 * fi_oracle.h -- CPU FP64 oracle for the FastIsostasy LV-ELVA time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran path (compiled there with wp = dp) used as the parity checker for
 * the HIP library and as the CPU baseline in bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links or calls it.
 *
 * Layout: every 2-D field is row-major [ny][nx] with x fastest, i.e. the same
 * memory order as the reference's Fortran column-major (nx, ny) arrays.
 */
#ifndef FI_ORACLE_H
#define FI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 8

/* Mirrors isos_param_class (isostasy_defs.f90:34-100).  Lengths in metres
 * (the namelist km -> m conversion of isos_par_load, fastisostasy.f90:971-979,
 * is the caller's job). */
typedef struct {
    int heterogeneous_ssh, interactive_sealevel, include_elastic;
    int correct_compression, correct_distortion;
    int method;                 /* 0 none, 1 LLRA, 2 LV-ELRA, 3 LV-ELVA */
    double dt_diagnostics;      /* [yr] */
    double min_pad;             /* [m] */
    int mantle;                 /* 0 uniform, 1 gaussian_plus, 2 gaussian_minus, 3 fields, 4 rheology_file */
    int lithosphere;            /* 0 uniform, 1 gaussian_plus, 2 gaussian_minus, 3 fields, 4 rheology_file */
    int layering;               /* 0 parallel, 1 equalized */
    int lumping;                /* 0 bottom, 1 min, 2 max, 3 mean, 4 logmean, 5 freqdomain */
    int viscosity_scaling_method; /* 0 as-is, 1 min, 2 scale, 3 max, 4 stddev */
    double viscosity_scaling;
    int nl;
    double zl[OR_MAXL];         /* [m] */
    double viscosities[OR_MAXL];/* [Pa s] */
    double dl;                  /* [m] */
    double eta_ref;             /* [Pa s] */
    double rho_water, rho_ice, rho_seawater, rho_uppermantle, rho_litho;
    double E;                   /* [GPa] */
    double nu, g, r_earth, m_earth;
    int dt_method;              /* 0 euler, 1 rk4, 2 bs32, 3 tsit54 */
    double atol, rtol, dt_min, dt_max, dt_init, q;
    double tau;                 /* [yr] asthenosphere relaxation time (methods 0-2) */
} or_params;

typedef struct or_isos or_isos;
typedef struct or_bsl or_bsl;

/* field ids for or_isos_get / or_isos_get_full */
enum {
    OR_Z_BED = 0, OR_Z_SS, OR_DZ_SS, OR_W, OR_WE, OR_DWDT, OR_CANOM_FULL, OR_HICE,
    OR_MASKOCEAN, OR_MASKGROUNDED, OR_MASKCONTINENT,
    /* padded-grid only */
    OR_CANOM_LOAD, OR_CANOM_MANTLE, OR_MASS_ANOM, OR_HAF, OR_ETA_EFF, OR_D_LITH,
    OR_KAPPA, OR_R, OR_MASKACTIVE, OR_HE_LITH, OR_ODE_X, OR_CANOM_SEAWATER,
    OR_CANOM_ICE, OR_CANOM_LITHO, OR_Z_BED_REF, OR_W_REF, OR_HAF_REF, OR_Z_SS_REF,
    OR_W_EQUILIBRIUM, OR_NFIELDS
};
/* scalar ids */
enum {
    OR_S_N = 0, OR_S_T, OR_S_DT, OR_S_BSL, OR_S_DELTAV, OR_S_V_AF, OR_S_V_DEN,
    OR_S_ATTEMPTS, OR_S_ACCEPTS, OR_S_D_LITH11, OR_S_ICROP1, OR_S_JCROP1,
    OR_S_NX_ICE, OR_S_NY_ICE
};

/* --- lifecycle (fastisostasy.f90:64-473, 550-660, 731-872) --- */
or_isos *or_isos_create(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                        const double *K, const double *rho_ice,
                        const double *he_field, const double *eta_field);
/* The arrays of the rheology_file and mask_file NetCDF inputs of isos_init
 * (fastisostasy.f90:199-213, 262-290, 336-411), coordinates in metres,
 * fields x fastest (Fortran (nx, ny[, nz]) order). */
typedef struct {
    int nx, ny;
    const double *xc, *yc;
    const double *litho_thickness;      /* nx*ny [m], lithosphere = 4 */
    double litho_thickness_scaling;
    int nz;
    const double *zc;                   /* nz [m] */
    const double *log10_eta;            /* nx*ny*nz, mantle = 4 */
    const double *stddev_log10_eta;     /* nx*ny*nz, viscosity_scaling_method = stddev */
    double rheo_smoothing_radius;       /* [m] */
    int mask_nx, mask_ny;
    const double *mask_xc, *mask_yc, *mask;   /* mask_file, or NULL */
} or_rheology_grid;
or_isos *or_isos_create_gridded(const or_params *p, int nx_ice, int ny_ice, double dx, double dy,
                                const double *K, const double *rho_ice, const or_rheology_grid *g);
/* bilinear_interpolation_array (isos_utils.f90:495-575), smooth_gauss_2D (:696-817) */
void or_bilinear(int nxr, int nyr, const double *x, const double *y, const double *z,
                 int nxo, int nyo, const double *xo, const double *yo, double *out);
void or_smooth_gauss(int nx, int ny, double *var, double dx, double sigma);
int  or_isos_init_ref(or_isos *s, const double *z_bed, const double *H_ice,
                      const double *bsl, const double *dz_ss);
int  or_isos_init_state(or_isos *s, const double *z_bed, const double *H_ice,
                        double time, or_bsl *b);
int  or_isos_update(or_isos *s, const double *H_ice, double time, or_bsl *b);
int  or_isos_get(const or_isos *s, int field, double *dst);      /* ice grid */
int  or_isos_get_full(const or_isos *s, int field, double *dst); /* padded grid */
double or_isos_scalar(const or_isos *s, int which);
/* isos_restart_write / isos_restart_read (isostasy_io.f90:56-194): binary
 * container of the same variables; a read makes the next init_state take
 * the restart branch (fastisostasy.f90:609-615) */
int or_isos_restart_write(const or_isos *s, const char *path, double time);
int or_isos_restart_read(or_isos *s, const char *path, double time);
int or_bsl_restart_write(const or_bsl *b, const char *path, double time);
int or_bsl_restart_read(const char *path, double *bsl, double *A_ocean);
void or_isos_destroy(or_isos *s);
const char *or_last_error(void);

/* --- barystatic sea level (barysealevel.f90:40-252) --- */
or_bsl *or_bsl_create(int method, double bsl_init, double A_ocean_pd,
                      int nvec, const double *bsl_vec, const double *A_vec,
                      double year_bp);
or_bsl *or_bsl_create_series(int method, double bsl_init, double A_ocean_pd, int nvec, const double *bsl_vec,
                             const double *A_vec, int ns, const double *st, const double *sv, double year_bp);
void   or_bsl_update(or_bsl *b, double year_bp);
double or_bsl_now(const or_bsl *b);
double or_bsl_A_ocean(const or_bsl *b);
void   or_bsl_destroy(or_bsl *b);

/* --- building blocks exposed for unit tests --- */
void or_fft(int n, const double *re_in, const double *im_in, double *re_out, double *im_out);
void or_dht2(int n, const double *in, double *out);
void or_deriv_x(int nx, int ny, const double *u, const double *dx, double *out);
void or_deriv_y(int nx, int ny, const double *u, const double *dy, double *out);
void or_deriv_xx(int nx, int ny, const double *u, const double *dx, double *out);
void or_deriv_yy(int nx, int ny, const double *u, const double *dy, double *out);
void or_kappa(int n, double dx, double *kappa);
void or_lumping(int n, int nl, int lumping, const double *eta, const double *bnd,
                const double *kappa, double *eta_eff, double *R);
void or_green_elastic(int n, double dx, double *ge);
void or_green_ssh(int n, double dx, double r_earth, double m_earth, double *gn);
/* Kelvin function kei_0(x) (series of Abramowitz & Stegun 9.9.12) and the
 * tabulation kei(0.00 .. 10.00 step 0.01) of load_kei_values
 * (kelvin_function.f90:167-1190): 5 significant digits, single precision */
double or_kei(double x);
void or_kei_table(double *rn, double *kei);
/* calc_kei_filter_2D + calc_viscous_green (kelvin_function.f90:22-94,
 * green_functions.f90:28-41) on the n x n domain grid */
void or_green_viscous(int n, double dx, double L_w, double D_lith, double *gv);
void or_convolve(int n, const double *kernel, const double *in, double *out);
void or_lvelva(int n, const double *w, const double *canom_full, const double *maskactive,
               double g, double nu, const double *D, const double *eta, const double *R,
               const double *kappa, const double *dxm, double sec_per_year, double *dwdt);
double or_interp_0d(int n, const double *x, const double *y, double xout);

#ifdef __cplusplus
}
#endif
#endif
