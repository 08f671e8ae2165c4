// rhs_common.hpp -- helpers shared by the generic and the power-of-two
// attempt kernels: step-size clamp, stage sums and block reductions.
#pragma once
#include "fi_device.hpp"

namespace fi {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double attempt_dt(int method, double dt, double t, double tf, double dt_max) {
    // integrators.f90:63,100-101,165
    const double rem = tf - t;
    if (method == 1) return fmin(dt, rem);  // rk4: h, not persisted
    dt = fmin(dt, rem);
    if (method == 2) dt = fmin(dt, dt_max);
    return dt;
}
__device__ __forceinline__ double attempt_dt(const Ctl& c, const PhysDev& ph) {
    return attempt_dt(ph.dt_method, c.dt, c.t, c.tf, ph.dt_max);
}

// Spectral divide by kappa*R (lv_elva.f90:115): FP64 divides; the FP32 path
// holds 1/(kappa*R) and multiplies.
__device__ __forceinline__ double kdiv(double h, double kr) { return h / kr; }
__device__ __forceinline__ float kdiv(float h, float rkr) { return h * rkr; }

// Scalar-cache views of the control block and the physics scalars.  Both are
// read-only while an attempt pass runs (k_decide, a separate launch, writes
// Ctl), so their fields load with SMEM into SGPRs rather than with one vector
// load and wait per field.
template <typename T> using kptr = const __attribute__((address_space(4))) T*;
template <typename T> __device__ __forceinline__ kptr<T> kview(const T* p) { return (kptr<T>)p; }

template <typename T> __device__ __forceinline__ T ld(const void* p, size_t i) { return ((const T*)p)[i]; }

// Global-address-space views of the field pointers held in Member: accesses
// through them compile to global_* instructions, which count only against
// vmcnt.  Flat accesses also count against lgkmcnt, so every LDS wait after
// them would wait for all outstanding global loads as well.
template <typename T> using gptr = const __attribute__((address_space(1))) T*;
template <typename T> using gmut = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> gin(const void* p) { return (gptr<T>)p; }
template <typename T> __device__ __forceinline__ gmut<T> gout(void* p) { return (gmut<T>)p; }

// ------------------------------------------------------------------ epilogue
struct EpiOut { double x1max, x2max, d2; };

// Stage sums of integrators.f90:12-203 for one cell, in the reference's
// association order.  Returns x2 (the candidate / new state).
__device__ __forceinline__ double stage_sums(int method, double x, double k, double dt,
                                             double& x1) {
    if (method == 0) { x1 = x + dt * k; return x1; }                       // euler
    if (method == 1) { x1 = x + (dt / 6) * (((k + 2 * k) + 2 * k) + k); return x1; }  // rk4
    if (method == 2) {                                                     // bs32
        x1 = x + dt / 9 * ((2 * k + 3 * k) + 4 * k);
        return x + dt * (((0.0 * k + 0.0 * k) + 0.0 * k) + 3 * k);
    }
    // tsit54: b = (0.0845, 0, 0, 0.5179, 0.1276, 0.27), bh(6) = 0.27 - 1/5
    const double b1 = dt * 0.0845, b4 = dt * 0.5179, b5 = dt * 0.1276, b6 = dt * 0.2700;
    const double z = dt * 0.0, h6 = dt * (0.2700 - 1 / 5.0);
    x1 = x + b1 * k + z * k + z * k + b4 * k + b5 * k + h6 * k;
    return x + b1 * k + z * k + z * k + b4 * k + b5 * k + b6 * k;
}

__device__ __forceinline__ void block_reduce3(double a, double b, double s, double* out) {
    __shared__ double red[3][16];
    a = wave_max(a); b = wave_max(b); s = wave_sum(s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[0][wv] = a; red[1][wv] = b; red[2][wv] = s; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        double A = red[0][0], B = red[1][0], S = red[2][0];
        for (int q = 1; q < nw; ++q) { A = fmax(A, red[0][q]); B = fmax(B, red[1][q]); S += red[2][q]; }
        out[0] = A; out[1] = B; out[2] = S;
    }
}


// Error norm and step control of one domain (integrators.f90:186-199), one
// wave per domain, after every block of the attempt has written its partials.
// dwdt is stored only by attempts that can end the update (t + dt >= tf, the
// candidate whose k is out%dwdt); a rejected attempt reuses it when present
// and otherwise recomputes the RHS.
__device__ __forceinline__ void decide_wave(const Member& M, int nblk) {
    Ctl& c = *M.ctl;
    const PhysDev& ph = *M.phys;
    const int lane = threadIdx.x & 63;
    double a1 = 0.0, a2 = 0.0, s2 = 0.0;
    for (int b = lane; b < nblk; b += 64) {
        const double* p = M.partials + (size_t)b * 4;
        a1 = fmax(a1, p[0]);
        a2 = fmax(a2, p[1]);
        s2 += p[2];
    }
    a1 = wave_max(a1); a2 = wave_max(a2); s2 = wave_sum(s2);
    if (lane != 0) return;
    const double dt = attempt_dt(c, ph);
    const bool stored = c.k_valid || !(c.t + dt < c.tf);
    c.attempts++;
    const int m = ph.dt_method;
    bool accept = true;
    if (m == 2 || m == 3) {
        const double tol = ph.atol + ph.rtol * fmax(a1, a2);
        const double nn = (double)M.n * (double)M.n;
        const double err = sqrt(s2 / (tol * tol) / nn);
        c.tol = tol; c.err = err;
        accept = err < 1;
        if (accept) c.dt = fmin(dt * pow(1 / err, 1 / ph.q), ph.dt_max);
        else c.dt = dt * 0.5;
    } else if (m == 0) {
        c.dt = dt;
    }
    if (accept) {
        c.t = c.t + dt;
        c.cur ^= 1;
        c.w_is_x = 1;
        c.k_valid = 0;
        c.accepts++;
    } else {
        c.k_valid = stored ? 1 : 0;
    }
    if (!(c.t < c.tf)) c.status = ST_IDLE;
}

}  // namespace fi
