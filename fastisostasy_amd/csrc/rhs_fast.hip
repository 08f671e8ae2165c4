// rhs_fast.hip -- attempt kernels for power-of-two N (64 ... 1024) built on the
// register Stockham FFT of fft_reg.hpp; same contract as the generic kernels of
// rhs.hip (identical inputs, outputs, partial layouts and Ctl gating).
//
//   kf_row_fwd   rows of f -> row Hartley transform        (g1 -> g2)
//   kf_col       column Hartley transform, divide by kappa*R, column Hartley
//                transform, corner partial sums             (g2 -> g1)
//   kf_row_inv   row Hartley transform, /N^2, corner BC, * sec_per_year,
//                stage sums, reduction partials             (g1 -> dwdt, x')
//
// Rows: L complex lines (2L real rows: r0+l in the real part, r0+L+l in the
// imaginary part) per block, TPL = N/8 threads per line.  Columns: C complex
// lines (columns c0+l and c0+C+l), lane -> (line = tid % C, t = tid / C) so
// that consecutive lanes read consecutive columns.
#include "fft_reg.hpp"
#include "fi_kernels.h"
#include "rhs_common.hpp"

namespace fi {

template <typename T, int N>
__device__ __forceinline__ void load_tw(cplx<T>* tw_l, const cplx<T>* __restrict__ twg) {
    for (int q = threadIdx.x; q < N; q += blockDim.x) tw_l[q] = twg[q];
}

template <typename T, int N, int THREADS>
__global__ void __launch_bounds__(THREADS) kf_row_fwd(const Member* __restrict__ mem, const void* __restrict__ twv) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    using PL = P2Plan<N>;
    constexpr int TPL = PL::TPL, E = PL::E, L = THREADS / TPL;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    load_tw<T, N>(tw_l, (const cplx<T>*)twv);
    const int ra = blockIdx.x * 2 * L + line, rb = ra + L;
    const T* __restrict__ src = (const T*)M.g1;
    cplx<T> v[E], m[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = mk<T>(src[(size_t)ra * N + t + TPL * e], src[(size_t)rb * N + t + TPL * e]);
    __syncthreads();
    RowXch<T, N> xc(buf, line);
    reg_fft<T, N>(v, t, tw_l, xc);
    reg_mirror<T, N>(v, m, t, xc);
    T* __restrict__ dst = (T*)M.g2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        dst[(size_t)ra * N + t + TPL * e] = ha;
        dst[(size_t)rb * N + t + TPL * e] = hb;
    }
}

template <typename T, int N, int C>
__global__ void __launch_bounds__(C * (N / 8)) kf_col(const Member* __restrict__ mem, const void* __restrict__ twv,
                                                     const double2* __restrict__ tw64) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    using PL = P2Plan<N>;
    constexpr int TPL = PL::TPL, E = PL::E;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    __shared__ double sc[2][2 * C];
    const int tid = threadIdx.x, line = tid % C, t = tid / C;
    load_tw<T, N>(tw_l, (const cplx<T>*)twv);
    const int c0 = blockIdx.x * 2 * C;
    const int ca = c0 + line, cb = c0 + C + line;
    const T* __restrict__ src = (const T*)M.g2;
    const T* __restrict__ kR = (const T*)M.kR;
    cplx<T> v[E], m[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const size_t row = (size_t)(t + TPL * e) * N;
        v[e] = mk<T>(src[row + ca], src[row + cb]);
    }
    __syncthreads();
    ColXch<T, N> xc(buf, line, C);
    reg_fft<T, N>(v, t, tw_l, xc);
    reg_mirror<T, N>(v, m, t, xc);
    // Hartley values and the spectral divide (lv_elva.f90:115)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const size_t row = (size_t)(t + TPL * e) * N;
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        v[e] = mk<T>(ha / kR[row + ca], hb / kR[row + cb]);
    }
    reg_fft<T, N>(v, t, tw_l, xc);
    reg_mirror<T, N>(v, m, t, xc);
    T* __restrict__ dst = (T*)M.g1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TPL * e;
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        dst[(size_t)k * N + ca] = ha;
        dst[(size_t)k * N + cb] = hb;
        if (k == 0) { sc[0][line] = (double)ha; sc[0][C + line] = (double)hb; }
        if (k == N - 1) { sc[1][line] = (double)ha; sc[1][C + line] = (double)hb; }
    }
    __syncthreads();
    if (tid < 4) {
        const int r = tid >> 1;
        const bool cs = tid & 1;
        double acc = 0.0;
        for (int mc = 0; mc < 2 * C; ++mc) {
            const double val = sc[r][mc];
            const int col = c0 + mc;
            acc += cs ? val * (tw64[col].x + tw64[col].y) : val;
        }
        M.corners[(size_t)blockIdx.x * 4 + tid] = acc;
    }
}

template <typename T, int N, int THREADS>
__global__ void __launch_bounds__(THREADS) kf_row_inv(const Member* __restrict__ mem, const void* __restrict__ twv,
                                                      int ncolblk) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN) return;
    const PhysDev& ph = *M.phys;
    using PL = P2Plan<N>;
    constexpr int TPL = PL::TPL, E = PL::E, L = THREADS / TPL;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    __shared__ double corner_c;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    const int ra = blockIdx.x * 2 * L + line, rb = ra + L;
    const double dt = attempt_dt(c, ph);
    const int meth = ph.dt_method;
    const double* __restrict__ X = M.X[c.cur];
    double* __restrict__ Xn = M.X[c.cur ^ 1];
    double* __restrict__ dw = M.dwdt;
    double a1 = 0.0, a2 = 0.0, s2 = 0.0;
    auto update = [&](size_t cidx, double k) {
        double x1;
        const double x2 = stage_sums(meth, X[cidx], k, dt, x1);
        Xn[cidx] = x2;
        a1 = fmax(a1, fabs(x1));
        a2 = fmax(a2, fabs(x2));
        const double d = x2 - x1;
        s2 += d * d;
    };
    if (c.k_valid) {
        // rejected attempt before: the RHS is unchanged, only dt differs
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const size_t ca = (size_t)ra * N + t + TPL * e, cb = (size_t)rb * N + t + TPL * e;
            update(ca, dw[ca]);
            update(cb, dw[cb]);
        }
    } else {
        load_tw<T, N>(tw_l, (const cplx<T>*)twv);
        const double S = (double)(float)((double)N * (double)N);
        if (tid == 0) {
            double q0 = 0, q1 = 0, q2 = 0, q3 = 0;
            for (int b = 0; b < ncolblk; ++b) {
                q0 += M.corners[(size_t)b * 4 + 0]; q1 += M.corners[(size_t)b * 4 + 1];
                q2 += M.corners[(size_t)b * 4 + 2]; q3 += M.corners[(size_t)b * 4 + 3];
            }
            // x(1,1) + x(nx,ny) + x(1,ny) + x(nx,1)  (isos_utils.f90:116)
            corner_c = 0.25 * (q0 / S + q3 / S + q2 / S + q1 / S);
        }
        const T* __restrict__ src = (const T*)M.g1;
        cplx<T> v[E], m[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            v[e] = mk<T>(src[(size_t)ra * N + t + TPL * e], src[(size_t)rb * N + t + TPL * e]);
        __syncthreads();
        RowXch<T, N> xc(buf, line);
        reg_fft<T, N>(v, t, tw_l, xc);
        reg_mirror<T, N>(v, m, t, xc);
        const double cc = corner_c, spy = ph.sec_per_year;
        const bool store_k = !(c.t + dt < c.tf);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            T ha, hb;
            dht_pair(v[e], m[e], ha, hb);
            const size_t ca = (size_t)ra * N + t + TPL * e, cb = (size_t)rb * N + t + TPL * e;
            const double ka = ((double)ha / S - cc) * spy, kb = ((double)hb / S - cc) * spy;
            if (store_k) {
                dw[ca] = ka;
                dw[cb] = kb;
            }
            update(ca, ka);
            update(cb, kb);
        }
    }
    block_reduce3(a1, a2, s2, M.partials + (size_t)blockIdx.x * 4);
}

// ------------------------------------------------------------------ dispatch
template <typename T, int N>
static size_t row_lds(int threads) {
    const int L = threads / (N / 8);
    return ((size_t)N + (size_t)L * RowXch<T, N>::PITCH) * sizeof(cplx<T>);
}
template <typename T, int N>
static size_t col_lds(int C) { return ((size_t)N + (size_t)N * (C + 1)) * sizeof(cplx<T>); }

template <typename T, int N, int THREADS>
static void rows_fwd(const Member* d, int nmem, hipStream_t st, const void* tw) {
    constexpr int L = THREADS / (N / 8);
    const size_t lds = row_lds<T, N>(THREADS);
    kf_row_fwd<T, N, THREADS><<<dim3(N / (2 * L), nmem), dim3(THREADS), lds, st>>>(d, tw);
}
template <typename T, int N, int THREADS>
static void rows_inv(const Member* d, int nmem, hipStream_t st, const void* tw, int ncolblk) {
    constexpr int L = THREADS / (N / 8);
    const size_t lds = row_lds<T, N>(THREADS);
    kf_row_inv<T, N, THREADS><<<dim3(N / (2 * L), nmem), dim3(THREADS), lds, st>>>(d, tw, ncolblk);
}
template <typename T, int N, int C>
static void cols(const Member* d, int nmem, hipStream_t st, const void* tw, const double2* tw64) {
    const size_t lds = col_lds<T, N>(C);
    kf_col<T, N, C><<<dim3(N / (2 * C), nmem), dim3(C * (N / 8)), lds, st>>>(d, tw, tw64);
}

template <typename T, int N>
static bool fast_n(const Member* d, int nmem, const RhsShape& s, hipStream_t st, int which) {
    // which: 1 row forward, 2 column, 3 row inverse
    if (which == 2) {
        switch (s.C) {
        case 1: cols<T, N, 1>(d, nmem, st, s.tw, s.tw64); return true;
        case 2: if constexpr (2 * (N / 8) <= 1024) { cols<T, N, 2>(d, nmem, st, s.tw, s.tw64); return true; } break;
        case 4: if constexpr (4 * (N / 8) <= 1024) { cols<T, N, 4>(d, nmem, st, s.tw, s.tw64); return true; } break;
        case 8: if constexpr (8 * (N / 8) <= 1024) { cols<T, N, 8>(d, nmem, st, s.tw, s.tw64); return true; } break;
        case 16: if constexpr (16 * (N / 8) <= 1024) { cols<T, N, 16>(d, nmem, st, s.tw, s.tw64); return true; } break;
        }
        return false;
    }
    const int ncolblk = s.n / (2 * s.C);
    switch (s.threads_row) {
    case 64: if constexpr (N / 8 <= 64) { which == 1 ? rows_fwd<T, N, 64>(d, nmem, st, s.tw) : rows_inv<T, N, 64>(d, nmem, st, s.tw, ncolblk); return true; } break;
    case 128: if constexpr (N / 8 <= 128) { which == 1 ? rows_fwd<T, N, 128>(d, nmem, st, s.tw) : rows_inv<T, N, 128>(d, nmem, st, s.tw, ncolblk); return true; } break;
    case 256: if constexpr (N / 8 <= 256) { which == 1 ? rows_fwd<T, N, 256>(d, nmem, st, s.tw) : rows_inv<T, N, 256>(d, nmem, st, s.tw, ncolblk); return true; } break;
    }
    return false;
}

template <typename T>
static bool fast_t(const Member* d, int nmem, const RhsShape& s, hipStream_t st, int which) {
    switch (s.n) {
    case 64: return fast_n<T, 64>(d, nmem, s, st, which);
    case 128: return fast_n<T, 128>(d, nmem, s, st, which);
    case 256: return fast_n<T, 256>(d, nmem, s, st, which);
    case 512: return fast_n<T, 512>(d, nmem, s, st, which);
    case 1024: return fast_n<T, 1024>(d, nmem, s, st, which);
    }
    return false;
}

bool launch_fast_pass(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st, int which) {
    return s.fp64 ? fast_t<double>(d_mem, nmem, s, st, which) : fast_t<float>(d_mem, nmem, s, st, which);
}

}  // namespace fi
