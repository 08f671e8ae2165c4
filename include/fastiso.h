/*
 * fastiso.h -- C ABI of the MI355X-native LV-ELVA time step.
 *
 * Drop-in boundary for the reference's public Fortran API (module
 * fastisostasy, src/fastisostasy.f90:34-60).  A Fortran ice-sheet model binds
 * these entry points through iso_c_binding (fastisostasy_amd/fortran/) and keeps
 * the isos_class / bsl_class init / update call sequence.
 *
 * Conventions
 *  - Every 2-D array is [ny][nx] with x fastest, i.e. exactly the memory order
 *    of a Fortran (nx, ny) array.  Ice-grid arrays are nx_ice x ny_ice; the
 *    padded-grid arrays (K, rheology fields, *_full getters) are N x N with
 *    N from fi_padded_size().
 *  - Input pointers are borrowed for the duration of the call only.
 *  - All functions return 0 on success and a negative FI_E* code on failure;
 *    fi_last_error() returns a thread-local message.  Where the reference
 *    calls `stop` the library returns an error instead.
 *  - Handles own their device memory.  Calls on one handle must be
 *    serialised by the caller; all handles of a process share one HIP stream.
 */
#ifndef FASTISO_H
#define FASTISO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FI_MAX_LAYERS 8

enum fi_status {
    FI_OK = 0,
    FI_E_ARG = -1,        /* invalid argument / configuration */
    FI_E_UNSUPPORTED = -2,/* option or size not supported by this build */
    FI_E_HIP = -3,        /* HIP runtime error */
    FI_E_UNSTABLE = -4,   /* large_check threshold exceeded (fastisostasy.f90:874-899) */
    FI_E_NAN = -5,        /* NaN in z_bed (fastisostasy.f90:864-869) */
    FI_E_STATE = -6,      /* call sequence violated (e.g. update before init_state) */
    FI_E_CAPACITY = -7    /* serving: the domain does not fit the HBM budget */
};

/* Precision of the RHS (stencil output, Hartley passes, spectral divide,
 * convolutions).  The integrator state, error norm, sea level, masks and the
 * barystatic reduction are FP64 in both modes. */
enum fi_precision { FI_FP64 = 0, FI_MIXED = 1 };

/* isos_param_class (src/isostasy_defs.f90:34-100), after the km -> m
 * conversions of isos_par_load (src/fastisostasy.f90:971-979).  Strings of
 * the namelist are enumerations here. */
typedef struct {
    int heterogeneous_ssh, interactive_sealevel, include_elastic;
    int correct_compression, correct_distortion;
    int method;                   /* 0 no displacement, 1 LLRA, 2 LV-ELRA, 3 LV-ELVA */
    double dt_diagnostics;        /* [yr] */
    double min_pad;               /* [m]  namelist "pad" */
    int mantle;                   /* 0 uniform, 1 gaussian_plus, 2 gaussian_minus, 3 fields, 4 rheology_file */
    int lithosphere;              /* 0 uniform, 1 gaussian_plus, 2 gaussian_minus, 3 fields, 4 rheology_file */
    int layering;                 /* 0 parallel, 1 equalized */
    int lumping;                  /* 0 bottom, 1 min, 2 max, 3 mean, 4 logmean, 5 freqdomain */
    int viscosity_scaling_method; /* 0 as-is, 1 min, 2 scale, 3 max, 4 stddev (no-op) */
    double viscosity_scaling;
    int nl;
    double zl[FI_MAX_LAYERS];     /* [m] */
    double viscosities[FI_MAX_LAYERS]; /* [Pa s] */
    double dl;                    /* [m] */
    double eta_ref;               /* [Pa s] ref_viscosity */
    double rho_water, rho_ice, rho_seawater, rho_uppermantle, rho_litho;
    double E;                     /* [GPa] */
    double nu, g, r_earth, m_earth; /* r_earth in m */
    int dt_method;                /* 0 euler, 1 rk4, 2 bs32, 3 tsit54 */
    double atol, rtol, dt_min, dt_max, dt_init, q;
    int precision;                /* enum fi_precision */
    double tau;                   /* [yr] asthenosphere relaxation time (methods 0-2) */
} fi_params;

/* ice-grid output components of isos_out_class (isostasy_defs.f90:185-200) */
enum fi_field {
    FI_Z_BED = 0, FI_Z_SS, FI_DZ_SS, FI_W, FI_WE, FI_DWDT, FI_CANOM_FULL, FI_HICE,
    FI_MASKOCEAN, FI_MASKGROUNDED, FI_MASKCONTINENT,
    /* padded-grid extras (fi_isos_get_full) */
    FI_CANOM_MANTLE, FI_HAF, FI_ETA_EFF, FI_D_LITH, FI_KAPPA, FI_R, FI_MASKACTIVE, FI_HE_LITH,
    FI_ODE_X, FI_Z_SS_REF, FI_Z_BED_REF,
    FI_W_EQUILIBRIUM,       /* now%w_equilibrium (methods 0-2) */
    FI_NFIELDS
};

enum fi_scalar {
    FI_S_N = 0, FI_S_T, FI_S_DT, FI_S_BSL, FI_S_DELTAV, FI_S_V_AF, FI_S_V_DEN,
    FI_S_ATTEMPTS, FI_S_ACCEPTS, FI_S_D_LITH11, FI_S_ERR, FI_S_TOL,
    FI_S_HBM_BYTES,         /* device memory held by the handle (bytes) */
    FI_S_STATUS,            /* status of the handle's last update checks: 0, FI_E_UNSTABLE or FI_E_NAN
                               (tells which members of a failed fi_batch_update failed) */
    FI_S_NX_ICE, FI_S_NY_ICE  /* the ice grid */
};

typedef struct fi_isos fi_isos;
typedef struct fi_bsl fi_bsl;
typedef struct fi_batch fi_batch;

const char* fi_last_error(void);
int fi_version(void);

/* Padded FFT edge N of init_domain_size (src/isos_utils.f90:285-354). */
int fi_padded_size(int nx_ice, int ny_ice, double dx, double dy, double min_pad, int* n_out);

/* isos_init (src/fastisostasy.f90:64-473).  K: distortion on the padded grid
 * (N x N) or NULL; rho_ice: optional override; he_field (N x N, m) and
 * eta_field (nl x N x N, Pa s) replace the rheology_file NetCDF path when
 * lithosphere / mantle == 3. */
int fi_isos_create(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy,
                   const double* K, const double* rho_ice, const double* he_field,
                   const double* eta_field, fi_isos** out);
/* The contents of the rheology_file and mask_file NetCDF inputs of isos_init
 * (src/fastisostasy.f90:199-213, 262-290, 336-411), read by the caller:
 * coordinates in metres (as nc_read_and_scale_vec leaves them), fields x
 * fastest (Fortran (nx, ny[, nz]) order).  lithosphere = 4 interpolates
 * litho_thickness (m) x litho_thickness_scaling onto the domain; mantle = 4
 * (needs lithosphere = 4, as in the reference) interpolates log10_eta per
 * level, adds viscosity_scaling x stddev_log10_eta when
 * viscosity_scaling_method is stddev, clamps to [16, 23] and interpolates to
 * the layer boundaries; both smooth with rheo_smoothing_radius > 0.  mask
 * (or NULL) gives maskactive = interpolated mask > 0.5. */
typedef struct {
    int nx, ny;
    const double* xc;
    const double* yc;
    const double* litho_thickness;
    double litho_thickness_scaling;
    int nz;
    const double* zc;
    const double* log10_eta;
    const double* stddev_log10_eta;
    double rheo_smoothing_radius;
    int mask_nx, mask_ny;
    const double* mask_xc;
    const double* mask_yc;
    const double* mask;
} fi_rheology_grid;

/* isos_init with the rheology / mask file branches (src/fastisostasy.f90:64-473) */
int fi_isos_create_gridded(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy,
                           const double* K, const double* rho_ice, const fi_rheology_grid* grid,
                           fi_isos** out);
/* isos_init_ref (src/fastisostasy.f90:550-575); bsl and dz_ss (N x N) may be NULL */
int fi_isos_init_ref(fi_isos* h, const double* z_bed, const double* H_ice,
                     const double* bsl, const double* dz_ss);
/* isos_init_state (src/fastisostasy.f90:596-660) */
int fi_isos_init_state(fi_isos* h, const double* z_bed, const double* H_ice, double time, fi_bsl* b);
/* isos_update (src/fastisostasy.f90:731-872); dwdt_corr is accepted but, as in
 * the reference (:755-759), unused */
int fi_isos_update(fi_isos* h, const double* H_ice, double time, fi_bsl* b);
/* No counterpart in the reference.  fi_isos_update for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc / hipFree; the
 * per-update working memory comes from a static 13 MiB array per GPU instead
 * of the handle's arena, so the handle (created beforehand) is the only
 * allocation.  Same results as fi_isos_update.  Limit: N <= 512 in either
 * precision, with or without convolutions; larger N -> FI_E_CAPACITY.  Calls
 * on one GPU share the array: calls that overlap in time race. */
int fi_isos_update_noalloc(fi_isos* h, const double* H_ice, double time, fi_bsl* b);
/* out%<field> after the last update (cropstate2out, src/isos_utils.f90:386-405):
 * double values, masks as 0.0 / 1.0 */
int fi_isos_get(fi_isos* h, int field, double* dst);
/* out%mask* as int8 0/1 */
int fi_isos_get_mask(fi_isos* h, int field, int8_t* dst);
/* padded-grid state (N x N) */
int fi_isos_get_full(fi_isos* h, int field, double* dst);
int fi_isos_get_scalar(fi_isos* h, int which, double* out);
/* isos_restart_write / isos_restart_read (src/isostasy_io.f90:56-194): the
 * reference's restart variables and units (now and ref states, He_lith in km,
 * mask_active, log10_eta_eff) on the padded grid, in a binary container
 * ("FIRESTRT", int32 version 1, int32 nrec, double time, then per record
 * char name[32], char units[16], int32 nx, int32 ny, double values, x fastest).
 * After a read, the next fi_isos_init_state takes the reference's restart
 * branch (src/fastisostasy.f90:609-615). */
int fi_isos_restart_write(fi_isos* h, const char* path, double time);
int fi_isos_restart_read(fi_isos* h, const char* path, double time);
/* bsl_restart_write (src/barysealevel.f90:426-450): bsl (m), A_ocean (m^2).
 * fi_bsl_restart_read (host only) gives the value bsl_init takes from a
 * restart file (:72-73). */
int fi_bsl_restart_write(fi_bsl* b, const char* path, double time);
int fi_bsl_restart_read(const char* path, double* bsl, double* A_ocean);
/* isos_end (src/fastisostasy.f90:911-923) */
int fi_isos_destroy(fi_isos* h);

/* bsl_init / bsl_update (src/barysealevel.f90:40-106, 195-252).
 * method 0 = const, 1 = fastiso (2 and 3 need fi_bsl_create_series); nvec > 0 gives the A_ocean(bsl) table of
 * A_ocean_path (otherwise the ocean surface is constant). */
int fi_bsl_create(int method, double bsl_init, double A_ocean_pd, int nvec, const double* bsl_vec,
                  const double* A_vec, double year_bp, fi_bsl** out);
/* bsl_init for the series methods 2 = file and 3 = mixed
 * (src/barysealevel.f90:40-153): series_time (yr) / series_bsl (m) is the
 * sea-level series of sl_path; mixed also takes the fastiso arguments.
 * fi_bsl_update then interpolates the series at year_bp (interp_linear,
 * :370-400); mixed subtracts each domain's cumulative contribution in
 * isos_update (src/fastisostasy.f90:771-775). */
int fi_bsl_create_series(int method, double bsl_init, double A_ocean_pd, int nvec, const double* bsl_vec,
                         const double* A_vec, int nseries, const double* series_time,
                         const double* series_bsl, double year_bp, fi_bsl** out);
/* read_series (src/barysealevel.f90:254-315): ASCII [time, value] columns
 * after one header line; '#' / '!' comment lines skipped.  Host only. */
int fi_bsl_read_series(const char* path, int max_n, double* time, double* var, int* n_out);
int fi_bsl_update(fi_bsl* b, double year_bp);
int fi_bsl_get(fi_bsl* b, double* bsl_now, double* A_ocean_now);
int fi_bsl_destroy(fi_bsl* b);

/* Batched extension (no reference counterpart): advance many domains of the
 * same padded size and precision with one launch sequence per attempted
 * step.  Members keep their own handles (init_state, get, destroy); the
 * batch only groups them.  Domains sharing one fi_bsl are updated in member
 * order, as consecutive isos_update calls would be. */
int fi_batch_create(fi_isos* const* members, int nmem, fi_batch** out);
/* H_ice[m] may be NULL to keep member m's current ice thickness. */
int fi_batch_update(fi_batch* b, const double* const* H_ice, double time, fi_bsl* const* bsl);
int fi_batch_destroy(fi_batch* b);

/* Serving (no reference counterpart; the batch / serving extension of the
 * boundary): continuous batching of update requests over resident domains.
 *   fi_server_create   budget for the device memory of admitted domains (bytes)
 *                      and the largest number of domains per launch group
 *   fi_server_admit    make an initialised domain resident, with the bsl object
 *                      its updates use; FI_E_CAPACITY when its footprint
 *                      (FI_S_HBM_BYTES) does not fit the remaining budget
 *   fi_server_submit   queue one isos_update of a domain to `time` (H_ice may be
 *                      NULL to keep the current ice); tags order requests
 *   fi_server_run      one round: the oldest request of every domain with one,
 *                      grouped by N, precision, options and diagnostics phase,
 *                      one batched update per group (with max_batch > 1 the
 *                      groups run concurrently on separate streams unless two
 *                      share a bsl object); returns the count served
 *   fi_server_poll     continuous scheduling instead of rounds: follows up the
 *                      groups whose last launch round has completed (another
 *                      round, or retire), then launches the oldest requests of
 *                      idle domains, grouped as in fi_server_run, on idle
 *                      streams (one group per stream of the server's 8, one
 *                      group per launch key at a time, so that requests arriving
 *                      meanwhile gather into the next), without waiting for the
 *                      groups still running; with wait != 0 it
 *                      blocks for the oldest group when nothing else could be
 *                      done.  A domain whose bsl object another domain in flight
 *                      uses waits.  completed: requests retired by this call;
 *                      pending counts requests in flight.  fi_server_run refuses
 *                      (FI_E_STATE) while polled work is in flight.
 *   fi_server_completed  tags of served requests, in completion order
 *   fi_server_failed     tags of requests whose update ran but failed the
 *                      instability / NaN checks (the domain's FI_S_STATUS
 *                      says which); fi_server_run returned that error.  A
 *                      request leaves the queue once its update has run, so
 *                      the next fi_server_run never applies it twice.
 * A request's results are bitwise those of the same update run alone.
 * Handles must not be updated through fi_isos_update while admitted. */
typedef struct fi_server fi_server;
int fi_server_create(double hbm_budget_bytes, int max_batch, fi_server** out);
int fi_server_admit(fi_server* s, fi_isos* h, fi_bsl* b, int* slot);
int fi_server_release(fi_server* s, int slot);
int fi_server_submit(fi_server* s, int slot, const double* H_ice, double time, long long tag);
int fi_server_pending(fi_server* s, int* n);
int fi_server_run(fi_server* s, int* completed);
int fi_server_poll(fi_server* s, int wait, int* completed);
int fi_server_completed(fi_server* s, long long* tags, int max, int* n);
int fi_server_failed(fi_server* s, long long* tags, int max, int* n);
int fi_server_stats(fi_server* s, long long* rounds, long long* groups, long long* served, double* hbm_used);
int fi_server_destroy(fi_server* s);

/* Output writers (isos_write_init / isos_write_step, src/isostasy_io.f90:197-405;
 * bsl_write_init / bsl_write_step, src/barysealevel.f90:453-494): the
 * reference's variables, units and long names on the ice grid in a binary
 * container ("FIOUTPUT", no NetCDF library here), time axis starting at
 * time_init with no data; a step at a time already on the axis overwrites
 * it, a new time is appended (nc_time_index).  xc, yc: the ice grid's axes
 * (m, nx_ice / ny_ice values, written in km).  A non-NULL z_bed_bench adds
 * z_bed_bench and err_z_bed (= out%w - z_bed_bench) to the file; H_ice is
 * unused, as in the reference (the file gets out%Hice). */
int fi_isos_write_init(fi_isos* h, const double* xc, const double* yc, const char* path, double time_init);
int fi_isos_write_step(fi_isos* h, const char* path, double time, const double* H_ice, const double* z_bed_bench);
int fi_bsl_write_init(fi_bsl* b, const char* path, double time_init);
int fi_bsl_write_step(fi_bsl* b, const char* path, double time);

/* Autotuner (no reference counterpart): times candidate launch shapes of the
 * attempt kernels -- threads per row block, line pairs per column block, row
 * tiles of the inverse pass and concurrent sub-batches on the register-FFT
 * path, and for batches of up to 8 domains rows per fused stencil block and
 * the inverse pass's own block size; line pairs per row / column block on the
 * mixed-radix path -- on
 * `nmem` synthetic N x N domains (`reps` timed 10-yr updates each) and keeps
 * the fastest for the process: batches, sub-batches and serving groups of that
 * (N, precision, size) created afterwards use it.  Results do not depend on
 * the shape.  choice (may be NULL, else 8 ints) receives {threads_row, C,
 * row_tiles, K, L_generic, C_generic, srow_rows, threads_inv}, 0 = the
 * built-in choice. */
int fi_autotune(int n, int precision, int nmem, int reps, int* choice);

/* Synchronise the library stream. */
int fi_sync(void);

/* Benchmark instrumentation (no reference counterpart): time every launch of
 * one attempt kernel with HIP events on the library stream.  kernel: 0 FD
 * stencil, 1 forward row Hartley pass, 2 column Hartley pass with spectral
 * divide, 3 inverse row Hartley pass with the step update; -1 disables.
 * fi_prof_collect returns the summed device time and the launch count since
 * the last fi_prof_select. */
int fi_prof_select(int kernel);
int fi_prof_collect(double* total_ms, long long* launches);
/* 1 when the last attempt ran the stencil fused into the forward row pass
 * (uniform spacing, power-of-two N); kernel 0 then times the fused kernel. */
int fi_prof_fused(void);
/* Launch shape of the last attempt (11 ints): {N, fp64, fused, register-FFT
 * kernels, threads per row block, line pairs per column block, row tiles of
 * the inverse pass, rows per fused block (0: 2 x threads / (N/8)), threads of
 * the inverse pass (0: threads per row block), integrator (-1: run time),
 * LDS exchange buffers of the column pass (1 or 2)}. */
int fi_prof_shape(int* out);
/* kernel = 99 in fi_prof_select times every attempt kernel; this sums the
 * timed launches of one of them (0, 2, 3; -1: all timed launches) */
int fi_prof_collect_kernel(int kernel, double* total_ms, long long* launches);
/* Achieved device-to-device copy bandwidth (bytes read + written per second /
 * 1e9) of `reps` copies of `mib` MiB on the library stream: the practical
 * ceiling beside the HBM peak. */
int fi_bench_copy(double mib, int reps, double* gbs);

#ifdef __cplusplus
}
#endif
#endif
