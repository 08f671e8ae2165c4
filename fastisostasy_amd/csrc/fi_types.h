// fi_types.h -- plain structs shared by host and device code of the HIP library.
#pragma once
#include <stdint.h>

namespace fi {

// Maximum butterflies per thread per Stockham stage; the host sizes FFT
// blocks so that (N / R) * lines <= FFT_MAXIT * blockDim.x for every stage.
constexpr int FFT_MAXIT = 4;

// Factorisation of a transform length into Stockham stages.
struct FftPlan {
    int n;
    int nf;
    int radix[24];
};

// Integrator / bookkeeping state of one domain, resident in device memory.
// Mirrors ode_class (isostasy_defs.f90:11-32) plus the scalar members of
// isos_state_class and isos_param_class that change during a run.
struct Ctl {
    double t, dt, tf;            // ode%t, ode%dt, target time of the running update
    double err, tol;             // last attempt: error norm and scalar tolerance
    double V_af, V_den, V_pov, deltaV, dbsl_total;
    double bsl;                  // now%bsl
    double time_diag, time_prog; // par%time_diagnostics / time_prognostics
    double maxw, maxwe, maxdzss; // instability check values of the last update
    double maxweq;               // max |w_equilibrium| of the last update (0 for LV-ELVA)
    double wcorner;              // corner correction applied to now%w at the end of an update
    double ref_bsl;              // ref%bsl
    int cur;                     // index of the live ode%x buffer (0/1)
    int w_is_x;                  // now%w currently equals ode%x[cur]
    int k_valid;                 // dwdt holds the RHS of the current now%w
    int status;                  // ST_RUN stepping, ST_IDLE idle, ST_NAN stopped by a NaN error norm
    int diag;                    // diagnostics updated in the running update
    int nonfinite;               // a NaN was seen in z_bed
    int pending;                 // the last attempt's accept/reject is not applied yet
    int skip;                    // this attempt slot does nothing (its inputs were loaded from the wrong state)
    int fin;                     // the running update has been finalized
    int pad_;
    long long attempts, accepts;
};

// ST_NAN: stepping stopped by a NaN error norm (reported as FI_E_NAN)
enum { ST_IDLE = 0, ST_RUN = 1, ST_NAN = -1 };

// Host-visible summary of one domain, gathered after a group of launches.
struct CtlSummary {
    double t, dt, tf, maxw, maxwe, maxdzss, maxweq;
    int status, nonfinite;
};

// Barystatic sea level, resident in device memory (bsl_class,
// barysealevel.f90:16-29).  The sea-level series of the methods file and
// mixed stays on the host, which interpolates it at each bsl_update.
struct BslDev {
    double bsl_init, bsl_now, time, A_ocean_pd, A_ocean_now;
    int method;                  // 0 const, 1 fastiso, 2 file, 3 mixed
    int constant_ocean_surface;
    int nvec;
    int pad_;
    double bsl_vec[64];
    double A_vec[64];
};

// Scalar physics of a domain (host-computed at create time).
struct PhysDev {
    double g, nu, rho_ice, rho_water, rho_seawater, rho_uppermantle, rho_litho;
    double sec_per_year, Vden_factor, dt_diagnostics;
    double atol, rtol, dt_max, q;
    double dx;
    double inv_dx, inv_2dx, inv_dx2;   // 1/dx, 1/(2 dx), 1/(dx*dx), rounded once on the host
    double tau;                  // [yr] domain%tau of methods 0-2 (uniform, fastisostasy.f90:236,446)
    int dt_method;              // 0 euler, 1 rk4, 2 bs32, 3 tsit54
    int interactive_sealevel, include_elastic, heterogeneous_ssh;
    int uniform_K;               // K == 1 everywhere (dx_matrix scalar)
    int method;                  // 0 none, 1 LLRA, 2 LV-ELRA, 3 LV-ELVA
};

// Per-domain device pointers.  Fields are N x N row-major [y][x] (x fastest)
// unless noted.  `rp` fields hold the RHS precision (double or float).
struct Member {
    int n;                       // padded size N
    int icrop1, jcrop1;          // 0-based crop origin of the ice grid
    int nx_ice, ny_ice;
    // integrator state (FP64)
    double* X[2];                // ode%x double buffer
    double* W;                   // now%w when w_is_x == 0
    double* dwdt;                // RHS k (m/yr), FP64
    // state fields (FP64)
    double *z_bed, *z_ss, *dz_ss, *we, *Hice, *Haf, *canom_mantle, *canom_full, *canom_load;
    double* conv_in;             // N x N staging of a convolution input
    double *z_bed_ref, *z_ss_ref, *Haf_ref, *Hice_ref, *w_ref, *we_ref, *dz_ss_ref;
    double *A, *dxm, *K2;        // cell area, dx*K, K^2 (only when !uniform_K)
    uint8_t* mask;               // bit0 ocean, bit1 grounded, bit2 continent, bit3 active
    uint8_t* mask_ref;
    double* H_in;                // ice-grid staging buffer (nx_ice * ny_ice)
    // RHS inputs in RHS precision
    void *D, *eta, *kR, *cfull;  // D_lith, eta_eff (plain: eta, mixed: 1/(2 eta)),
                                 // kappa*R (plain; mixed: 1/(kappa*R)), canom_full
    void *g1, *g2;               // FFT work arrays (N x N)
    double* corners;             // per-column terms of the four Hartley corners [N][4]
    double* partials;            // error-norm partials [nblk][4]: one per row pair (power-of-two
                                 // kernels), per row (mixed-radix), per row pair (relaxation),
                                 // or per cell block (bookkeeping sums)
    double* conv_work;           // convolution work (complex, 2N x (N+1)) or null
    void* FGE;                   // elastic kernel half-spectrum, (2N) x (N+1) complex
    void* FGN;                   // sea-surface kernel half-spectrum
    void* FGV;                   // LV-ELRA viscous (Kelvin) kernel half-spectrum
    double kc_E[4], kc_N[4], kc_V[4];   // Green-kernel corner values (00, NN, N0, 0N)
    double* weq;                 // now%w_equilibrium (methods 0-2), N x N
    Ctl* ctl;
    Ctl* ctl_next;               // decision of the previous attempt, written by the first pass
    BslDev* bsl;
    PhysDev* phys;
    // domains sharing one bsl object: the first in batch order leads and
    // applies the group's updates in order (next: batch index, -1 at the end)
    int bsl_lead, bsl_next;
    double tf_req;               // per-domain target time of a serving round
};

}  // namespace fi
