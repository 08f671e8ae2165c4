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

// Requires the given block-uniform values (pointers read from a Member) to be
// in scalar registers at this point: their scalar loads are then issued as
// one batch with one wait, ahead of the vector loads that use them, instead
// of each next to its first use behind the previous one's wait.
template <typename P> __device__ __forceinline__ void sgpr_ready(P p) { asm volatile("" : : "s"(p)); }
template <typename... P> __device__ __forceinline__ void scalars_ready(P... p) {
    const int order[] = {(sgpr_ready(p), 0)...};
    (void)order;
}

// ------------------------------------------------------------------ epilogue
struct EpiOut { double x1max, x2max, d2; };

// Stage sums of integrators.f90:12-203 for one cell, in the reference's
// association order.  Returns x2 (the candidate / new state).
template <int METHOD>
__device__ __forceinline__ double stage_sums_m(double x, double k, double dt, double& x1) {
    if constexpr (METHOD == 0) { x1 = x + dt * k; return x1; }                       // euler
    if constexpr (METHOD == 1) { x1 = x + (dt / 6) * (((k + 2 * k) + 2 * k) + k); return x1; }  // rk4
    if constexpr (METHOD == 2) {                                                     // bs32
        x1 = x + dt / 9 * ((2 * k + 3 * k) + 4 * k);
        return x + dt * (((0.0 * k + 0.0 * k) + 0.0 * k) + 3 * k);
    }
    // tsit54: b = (0.0845, 0, 0, 0.5179, 0.1276, 0.27), bh(6) = 0.27 - 1/5
    const double b1 = dt * 0.0845, b4 = dt * 0.5179, b5 = dt * 0.1276, b6 = dt * 0.2700;
    const double z = dt * 0.0, h6 = dt * (0.2700 - 1 / 5.0);
    x1 = x + b1 * k + z * k + z * k + b4 * k + b5 * k + h6 * k;
    return x + b1 * k + z * k + z * k + b4 * k + b5 * k + b6 * k;
}

// the integrator as a run-time value
__device__ __forceinline__ double stage_sums(int method, double x, double k, double dt, double& x1) {
    if (method == 0) return stage_sums_m<0>(x, k, dt, x1);
    if (method == 1) return stage_sums_m<1>(x, k, dt, x1);
    if (method == 2) return stage_sums_m<2>(x, k, dt, x1);
    return stage_sums_m<3>(x, k, dt, x1);
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
        out[0] = fmax(A, B); out[1] = S;
    }
}


// A block-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ double uni(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ long long uni(long long v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return ((long long)hi << 32) | (unsigned int)lo;
}

// The fields the step control reads and writes.
struct StepState {
    double t, dt, tf, err, tol;
    int cur, w_is_x, k_valid, status;
    long long attempts, accepts;
};

__device__ __forceinline__ StepState load_state(const kptr<Ctl> c) {
    StepState s;
    s.t = c->t; s.dt = c->dt; s.tf = c->tf; s.err = c->err; s.tol = c->tol;
    s.cur = c->cur; s.w_is_x = c->w_is_x; s.k_valid = c->k_valid; s.status = c->status;
    s.attempts = c->attempts; s.accepts = c->accepts;
    return s;
}

// Sum the per-row (or per-block) partials of the last attempt: lane-strided,
// then fixed butterflies -- an order fixed by the partial count alone.
// Partials hold (max(max|x1|, max|x2|), sum (x2-x1)^2) per row or block: the
// tolerance uses only the larger maximum, and max is exact in any order.
__device__ __forceinline__ void reduce_partials(gptr<double> part, int nblk, double& amax, double& s2) {
    const int lane = threadIdx.x & 63;
    amax = 0.0; s2 = 0.0;
    for (int b = lane; b < nblk; b += 64) {
        amax = fmax(amax, part[(size_t)b * 4 + 0]);
        s2 += part[(size_t)b * 4 + 1];
    }
    amax = wave_max(amax); s2 = wave_sum(s2);
}

// Error norm and step control of one domain (integrators.f90:186-199) after
// an attempt whose partials give max|x1|, max|x2| and sum (x2-x1)^2.  dwdt is
// stored only by attempts that can end the update (t + dt >= tf, the
// candidate whose k is out%dwdt); a rejected attempt reuses it when present
// and otherwise recomputes the RHS.
__device__ __forceinline__ StepState decide_step(StepState s, double amax, double s2, const kptr<PhysDev> ph, int n) {
    const int m = ph->dt_method;
    const double dt = attempt_dt(m, s.dt, s.t, s.tf, ph->dt_max);
    const bool stored = s.k_valid || !(s.t + dt < s.tf);
    s.attempts += 1;
    bool accept = true, nan_err = false;
    if (m == 2 || m == 3) {
        const double tol = ph->atol + ph->rtol * amax;
        const double nn = (double)n * (double)n;
        const double err = sqrt(s2 / (tol * tol) / nn);
        s.tol = tol; s.err = err;
        accept = err < 1;
        nan_err = err != err;
        // pow(x, 1) is x exactly (C99 F.9.4.4): the default q = 1 skips the
        // long FP64 pow sequence that every wave of the fused pass would run
        if (accept) {
            const double inv = 1 / err;
            s.dt = fmin(dt * (ph->q == 1.0 ? inv : pow(inv, 1 / ph->q)), ph->dt_max);
        }
        else s.dt = dt * 0.5;
    } else if (m == 0) {
        s.dt = dt;
    }
    if (accept) {
        s.t = s.t + dt;
        s.cur ^= 1;
        s.w_is_x = 1;
        s.k_valid = 0;
        s.accepts += 1;
    } else {
        s.k_valid = stored ? 1 : 0;
    }
    if (!(s.t < s.tf)) s.status = ST_IDLE;
    // A NaN error norm rejects every attempt (NaN < 1 is false), and the
    // reference then halves dt without end; the update stops instead and
    // reports the NaN (the finalize turns ST_NAN into the NaN flag).
    if (nan_err) s.status = ST_NAN;
    return s;
}

// Field-wise copy of a control block read through the scalar cache.
__device__ __forceinline__ void copy_ctl(Ctl* d, const kptr<Ctl> s) {
    d->t = s->t; d->dt = s->dt; d->tf = s->tf; d->err = s->err; d->tol = s->tol;
    d->V_af = s->V_af; d->V_den = s->V_den; d->V_pov = s->V_pov; d->deltaV = s->deltaV;
    d->dbsl_total = s->dbsl_total; d->bsl = s->bsl; d->time_diag = s->time_diag; d->time_prog = s->time_prog;
    d->maxw = s->maxw; d->maxwe = s->maxwe; d->maxdzss = s->maxdzss; d->maxweq = s->maxweq;
    d->wcorner = s->wcorner;
    d->ref_bsl = s->ref_bsl; d->cur = s->cur; d->w_is_x = s->w_is_x; d->k_valid = s->k_valid;
    d->status = s->status; d->diag = s->diag; d->nonfinite = s->nonfinite; d->pending = s->pending;
    d->skip = s->skip; d->fin = s->fin; d->pad_ = 0; d->attempts = s->attempts; d->accepts = s->accepts;
}

__device__ __forceinline__ void store_state(Ctl* c, const StepState& s) {
    c->t = s.t; c->dt = s.dt; c->tf = s.tf; c->err = s.err; c->tol = s.tol;
    c->cur = s.cur; c->w_is_x = s.w_is_x; c->k_valid = s.k_valid; c->status = s.status;
    c->attempts = s.attempts; c->accepts = s.accepts;
}

// Stand-alone decision of one domain (one wave): applies the pending
// decision of the last attempt, if any.
__device__ __forceinline__ void decide_wave(const Member& M, int nblk) {
    const kptr<Ctl> ck = kview(M.ctl);
    const StepState s0 = load_state(ck);
    const int pending = ck->pending;
    double amax, s2;
    reduce_partials(gin<double>(M.partials), nblk, amax, s2);
    if (!pending || (threadIdx.x & 63) != 0) return;
    const StepState s = decide_step(s0, amax, s2, kview(M.phys), M.n);
    store_state(M.ctl, s);
    M.ctl->pending = 0;
}

}  // namespace fi
