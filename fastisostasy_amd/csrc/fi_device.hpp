// fi_device.hpp -- device-side building blocks shared by the LV-ELVA kernels.
//
//  * complex helpers on float2 / double2,
//  * an in-LDS mixed-radix Stockham FFT (radices 2,3,4,5,7,8) that transforms
//    L independent lines stored "line-fastest" (element (line, idx) at
//    idx * P + line), so that consecutive lanes touch consecutive LDS words
//    in every stage,
//  * the Hartley extraction H[k] = Re F[k] - Im F[k] for two real lines packed
//    into one complex line.
//
// Written for gfx950 (wave64).  All FFTs are forward, unnormalised,
// F[k] = sum_n x[n] exp(-2 pi i n k / N); the inverse DFT used by the
// convolution is obtained by conjugation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fi_types.h"

namespace fi {

template <typename T> struct c2;
template <> struct c2<float> { using t = float2; };
template <> struct c2<double> { using t = double2; };
template <typename T> using cplx = typename c2<T>::t;

template <typename T> __device__ __forceinline__ cplx<T> mk(T x, T y) { cplx<T> r; r.x = x; r.y = y; return r; }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// FP32 transforms (mixed precision) use fused multiply-adds for the twiddle
// products; FP64 keeps separate roundings like every other FP64 operation
__device__ __forceinline__ cplx<float> cmul(cplx<float> a, cplx<float> b) {
    cplx<float> r;
    r.x = __builtin_fmaf(a.x, b.x, -(a.y * b.y));
    r.y = __builtin_fmaf(a.x, b.y, a.y * b.x);
    return r;
}
// multiply by -i
template <typename C> __device__ __forceinline__ C cmi(C a) { C r; r.x = a.y; r.y = -a.x; return r; }
// multiply by +i
template <typename C> __device__ __forceinline__ C cpi(C a) { C r; r.x = -a.y; r.y = a.x; return r; }

// ---------------------------------------------------------------- butterflies
template <typename C> __device__ __forceinline__ void bfly2(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <typename C> __device__ __forceinline__ void bfly4(C* v) {
    C s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
    C s2 = cadd(v[1], v[3]), s3 = cmi(csub(v[1], v[3]));
    v[0] = cadd(s0, s2); v[2] = csub(s0, s2);
    v[1] = cadd(s1, s3); v[3] = csub(s1, s3);
}
template <typename T> __device__ __forceinline__ void bfly8(cplx<T>* v) {
    using C = cplx<T>;
    const T h = (T)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    bfly4(e); bfly4(o);
    // o[k] *= exp(-2 pi i k / 8)
    C o1 = mk<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    C o2 = cmi(o[2]);
    C o3 = mk<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}
template <typename T> __device__ __forceinline__ void bfly3(cplx<T>* v) {
    const T s = (T)0.86602540378443864676;
    auto a = v[0], b = v[1], c = v[2];
    auto sum = cadd(b, c);
    auto d = csub(b, c);
    auto t = mk<T>(a.x - (T)0.5 * sum.x, a.y - (T)0.5 * sum.y);
    auto u = mk<T>(s * d.y, -s * d.x);  // -i s (b - c)
    v[0] = cadd(a, sum);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
}
// generic prime radix (5, 7): direct DFT with compile-time roots of unity
template <typename T, int R> __device__ __forceinline__ void bflyP(cplx<T>* v) {
    cplx<T> out[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        T re = 0, im = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int m = (r * q) % R;
            const double ang = -2.0 * 3.14159265358979323846 * (double)m / (double)R;
            const T c = (T)__builtin_cos(ang), sn = (T)__builtin_sin(ang);
            re += v[r].x * c - v[r].y * sn;
            im += v[r].x * sn + v[r].y * c;
        }
        out[q] = mk<T>(re, im);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
}
template <typename T, int R> __device__ __forceinline__ void bfly(cplx<T>* v) {
    if constexpr (R == 2) bfly2(v);
    else if constexpr (R == 3) bfly3<T>(v);
    else if constexpr (R == 4) bfly4(v);
    else if constexpr (R == 8) bfly8<T>(v);
    else bflyP<T, R>(v);
}

// ---------------------------------------------------------------- LDS FFT

// One Stockham stage of radix R over L lines of length N in LDS (in place,
// register staged: all reads of the stage complete before any write).
template <typename T, int R>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, int N, int L, int P, int Ns,
                                          const cplx<T>* __restrict__ tw) {
    using C = cplx<T>;
    const int m = N / R;
    const int NB = m * L;
    const int tstep = N / (Ns * R);
    const int tid = threadIdx.x, nthr = blockDim.x;
    C v[FFT_MAXIT][R];
#pragma unroll
    for (int it = 0; it < FFT_MAXIT; ++it) {
        const int b = tid + it * nthr;
        if (b < NB) {
            const int line = b % L, j = b / L;
            const int k = j % Ns;
            v[it][0] = buf[j * P + line];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const C x = buf[(j + r * m) * P + line];
                v[it][r] = (k == 0) ? x : cmul(x, tw[k * r * tstep]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < FFT_MAXIT; ++it) {
        const int b = tid + it * nthr;
        if (b < NB) {
            const int line = b % L, j = b / L;
            const int k = j % Ns;
            bfly<T, R>(v[it]);
            const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[(idxD + q * Ns) * P + line] = v[it][q];
        }
    }
    __syncthreads();
}

// Copy `total` elements, element q read by ld(q) and written by st(q, v), with
// U loads in flight per thread ahead of their stores: a plain load-store loop
// waits for every load in turn.  Loads are unconditional (indices past the end
// re-read element `base`), so no branch splits the wait counts.
template <int U, class Ld, class St>
__device__ __forceinline__ void staged_copy(int total, Ld ld, St st) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int base = tid; base < total; base += U * nt) {
        decltype(ld(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = base + u * nt;
            v[u] = ld(q < total ? q : base);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = base + u * nt;
            if (q < total) st(q, v[u]);
        }
    }
}

// Copy the n-entry twiddle table into LDS right after the `lines` data lines
// of buf (the caller's barrier before the first stage publishes it), so the
// stages read their twiddles from LDS instead of issuing a global load, and a
// wait, per butterfly.
template <typename T>
__device__ __forceinline__ const cplx<T>* stage_twiddles(cplx<T>* buf, int n, int lines, const void* twv) {
    cplx<T>* twl = buf + (size_t)n * lines;
    const cplx<T>* __restrict__ g = (const cplx<T>*)twv;
    for (int q = threadIdx.x; q < n; q += blockDim.x) twl[q] = g[q];
    return twl;
}

// Full forward FFT of L lines (caller has synchronised after filling buf).
template <typename T>
__device__ __forceinline__ void lds_fft(cplx<T>* buf, const FftPlan& plan, int L, int P,
                                        const cplx<T>* __restrict__ tw) {
    int Ns = 1;
    for (int s = 0; s < plan.nf; ++s) {
        const int R = plan.radix[s];
        switch (R) {
        case 8: fft_stage<T, 8>(buf, plan.n, L, P, Ns, tw); break;
        case 4: fft_stage<T, 4>(buf, plan.n, L, P, Ns, tw); break;
        case 2: fft_stage<T, 2>(buf, plan.n, L, P, Ns, tw); break;
        case 3: fft_stage<T, 3>(buf, plan.n, L, P, Ns, tw); break;
        case 5: fft_stage<T, 5>(buf, plan.n, L, P, Ns, tw); break;
        default: fft_stage<T, 7>(buf, plan.n, L, P, Ns, tw); break;
        }
        Ns *= R;
    }
}

// Hartley values of the two real lines packed in complex line Z (a = Re, b = Im):
// Ha[k] = Re A[k] - Im A[k], Hb[k] = Re B[k] - Im B[k].
template <typename C>
__device__ __forceinline__ void dht_pair(C zk, C zm, decltype(C::x)& ha, decltype(C::x)& hb) {
    using T = decltype(C::x);
    ha = (T)0.5 * ((zk.x + zm.x) - (zk.y - zm.y));
    hb = (T)0.5 * ((zk.y + zm.y) + (zk.x - zm.x));
}

// wave64 reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace fi
