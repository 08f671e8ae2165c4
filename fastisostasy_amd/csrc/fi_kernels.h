// fi_kernels.h -- host-callable launchers of the device kernels.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/fastiso.h"
#include "fi_types.h"

namespace fi {

// Launch geometry of one attempted step for a batch of equal-size domains.
struct RhsShape {
    int n;
    int L;            // complex lines (= 2L real rows) per row-pass block
    int C;            // complex lines (= 2C real columns) per column-pass block
    int threads_row, threads_col;
    FftPlan plan;
    const void* tw;   // twiddles in RHS precision
    const double2* tw64;
    bool fp64;
    bool uniform;     // K == 1
    bool fast;        // power-of-two N: register-FFT kernels (threads_row = block size)
    bool fused;       // fast && uniform: stencil fused into the forward row pass
    int row_tiles;    // row blocks per thread block of the inverse row pass (1 or 2)
    int srow_rows;    // fast && fused: rows per block of the fused stencil pass with one
                      // column per thread (4 or 8); 0: 2L rows in threads_row threads
    int threads_inv;  // fast: block size of the inverse row pass (0: threads_row)
    int method;       // 3 LV-ELVA (transform passes); 0-2 pointwise relaxation RHS
    int dt_method;    // integrator shared by the batch (0-3), -1 if not known
};

// Geometry of the 2N-point linear convolution.
struct ConvShape {
    int n, m;         // m = 2n
    int Lr, Cc;       // lines per row / column block
    int threads_row, threads_col;
    FftPlan plan;     // length m
    const void* tw;   // length-m twiddles in conv precision
    bool fp64;
    bool rows_twl;    // the row passes also hold the twiddle table in LDS (the column pass always does)
};

void launch_attempt(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st);
// applies a pending step decision left by the last attempt (no-op otherwise)
void launch_decide(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st);
// power-of-two passes: which = 1 row forward, 2 column, 3 row inverse
bool launch_fast_pass(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st, int which);
// LDS exchange buffers (1 or 2) of the column pass launched for shape s
int col_buffers(const RhsShape& s);
// event timing of one attempt kernel: 0 stencil (fused: stencil + row DHT), 1 row DHT, 2 column DHT+divide,
// 3 inverse row DHT+update; PROF_ALL times every one of them
constexpr int PROF_ALL = 99;
void prof_select(int kernel);
int prof_collect(double* total_ms, long long* launches);
// the launches of one kernel (-1: all timed launches)
int prof_collect_kernel(int kernel, double* total_ms, long long* launches);
// the library's error message (fi_last_error) for code `code`; returns code
int set_error(int code, const char* msg);
int prof_fused();
void prof_shape(int* out);   // 11 ints, fi_prof_shape
bool prof_active();
// a start/stop event pair when launches of `kernel` are being timed
bool prof_pair(int kernel, hipEvent_t* start, hipEvent_t* stop);

// Launch of an attempt kernel.  When its launches are being timed, the events
// ride on the dispatch itself (hipExtLaunchKernel), so the timing adds no
// packets to the stream and measures the kernel alone.
template <typename... P, typename... A>
inline void launch_k(int prof_id, void (*kern)(P...), dim3 grid, dim3 block, size_t lds, hipStream_t st,
                     A... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof_pair(prof_id, &e0, &e1))
        hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)lds, st, e0, e1, 0u, args...);
    else
        kern<<<grid, block, lds, st>>>(args...);
}

// bookkeeping of isos_update (sealevel.f90, fastisostasy.f90:750-833)
void launch_pre_haf(const Member* d_mem, int nmem, int n, bool new_H, hipStream_t st);
void launch_pre_bsl(const Member* d_mem, int nmem, int n, hipStream_t st);
// phase 0: z_ss/masks/loads and, without diagnostics, everything the RHS needs
// phase 1: (after the elastic convolution) litho/full/mass anomaly
// phase 2: (after the sea-surface convolution) z_ss/masks/loads/full
void launch_pre_fields(const Member* d_mem, int nmem, int n, int phase, bool diag, bool fp64, hipStream_t st);
void launch_begin(const Member* d_mem, int nmem, double tf, int stepping, bool per_member, hipStream_t st);
void launch_finalize(const Member* d_mem, int nmem, int n, hipStream_t st);
void launch_gather(const Member* d_mem, int nmem, CtlSummary* d_out, hipStream_t st);
void launch_bsl_update(BslDev* b, double year_bp, double file_value, hipStream_t st);
// methods 0-2: w_equilibrium of the update (lv_xlra.f90, fastisostasy.f90:686-714);
// method 2 first stages -canom_load g K^2 (launch_weq_input), then convolves it
void launch_weq_input(const Member* d_mem, int nmem, int n, hipStream_t st);
void launch_weq(const Member* d_mem, int nmem, int n, int method, hipStream_t st);
void launch_set_ref_zss(const Member* d_mem, int nmem, int n, hipStream_t st);
void launch_init_state_fields(const Member* d_mem, int nmem, int n, hipStream_t st);

// linear convolution with a precomputed kernel spectrum (convolutions.f90:59-91)
// src_kind: 0 elastic input (canom_load*g*K^2 taken from conv_work staging),
// which = 0 elastic (writes we), 1 sea surface (writes dz_ss), 2 LV-ELRA (writes weq)
void launch_convolution(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st);
void launch_kernel_spectrum(const Member* d_mem, int nmem, const ConvShape& s, int which,
                            const double* d_kernel, hipStream_t st);

// init-time fields (lv_elva.f90:442-481, 207-306; green_functions.f90:43-181)
struct InitArgs {
    int n;
    double dx;
    int nl;
    int lithosphere, mantle, layering, lumping, vscale_method;
    double vscale, zl[8], visc[8], dl, eta_ref, E, nu, comp_corr;
    double r_earth, m_earth;
};
void launch_init_rheology(const Member* d_mem, const InitArgs& a, const double* d_he, const double* d_eta,
                          double* d_etaeff, double* d_kappa, double* d_R, double* d_He, double* d_Dd,
                          double* d_scratch, bool fp64, hipStream_t st);
void launch_init_finalize(const InitArgs& a, const double* d_scratch, double* d_etaeff, const double* d_kappa,
                          const double* d_R, const double* d_Dd, void* D_t, void* eta_t, void* kR_t,
                          bool fp64, hipStream_t st);
void launch_green(int which, int n, double dx, double r_earth, double m_earth, double* d_out, hipStream_t st);
// rheology_file / mask_file inputs on the host (rheogrid.cpp)
void grid_smooth(int n, double* var, double dx, double sigma);
int grid_rheology(const fi_params& P, const fi_rheology_grid& g, int n, double dx, std::vector<double>& he,
                  std::vector<double>& eta);
void grid_mask(const fi_rheology_grid& g, int n, double dx, std::vector<uint8_t>& active);
// LV-ELRA viscous Green's function on the host (kelvin.cpp)
void green_viscous_host(int n, double dx, double L_w, double D_lith, double* out);
double kelvin_kei_value(double x);

}  // namespace fi
