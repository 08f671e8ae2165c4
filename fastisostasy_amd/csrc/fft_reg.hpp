// fft_reg.hpp -- register-resident Stockham FFT for power-of-two lengths.
//
// A line of N complex values is processed by TPL = N / 8 threads that each hold
// E = 8 elements, element e of thread t being X[t + TPL * e].  Stage 0 reads
// its inputs in exactly that order, so it is fed straight from global memory;
// the last stage leaves its outputs in the same order in registers.  The
// stages in between exchange data through one LDS buffer (one barrier
// per exchange, plus one before the buffer is reused).  Radices: 8 for every stage but a possible final 4 or 2.
// Twiddles exp(-2 pi i m / N) are read from an LDS copy of the table.
#pragma once
#include "fi_device.hpp"

namespace fi {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

template <int N> struct P2Plan {
    static_assert((N & (N - 1)) == 0 && N >= 8, "power of two >= 8");
    static constexpr int E = 8;
    static constexpr int TPL = N / E;
    static constexpr int LOG = ilog2c(N);
    static constexpr int NF = (LOG + 2) / 3;                 // stages
    static constexpr int radix(int s) {
        return (s < LOG / 3) ? 8 : (LOG % 3 == 1 ? 2 : 4);
    }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * radix(s - 1); }
};

// Exchange buffers (single LDS buffer, two barriers per exchange): where
// element `pos` of the calling thread's line lives.  Rows: line-major with one
// pad element per 8 (spreads the stride-R stage writes over the banks);
// columns: position-major with the lines fastest and one pad per row.
template <typename T, int N> struct RowXch {
    static constexpr int PITCH = N + N / 8 + 1;
    cplx<T>* base;
    int line;
    __device__ __forceinline__ RowXch(cplx<T>* b, int l) : base(b), line(l) {}
    __device__ __forceinline__ int addr(int pos) const { return line * PITCH + pos + (pos >> 3); }
    __device__ __forceinline__ void put(int pos, cplx<T> v) { base[addr(pos)] = v; }
    __device__ __forceinline__ cplx<T> get(int pos) const { return base[addr(pos)]; }
};

template <typename T, int N> struct ColXch {
    cplx<T>* base;
    int line, pitch;
    __device__ __forceinline__ ColXch(cplx<T>* b, int l, int C) : base(b), line(l), pitch(C + 1) {}
    __device__ __forceinline__ int addr(int pos) const { return pos * pitch + line; }
    __device__ __forceinline__ void put(int pos, cplx<T> v) { base[addr(pos)] = v; }
    __device__ __forceinline__ cplx<T> get(int pos) const { return base[addr(pos)]; }
};

template <typename T, int N, int S, class X>
__device__ __forceinline__ void reg_stage(cplx<T>* v, int t, const cplx<T>* __restrict__ tw, X& xc) {
    using PL = P2Plan<N>;
    constexpr int E = PL::E, TPL = PL::TPL;
    constexpr int R = PL::radix(S), NS = PL::ns(S), B = E / R, TSTEP = N / (NS * R);
    constexpr bool LAST = (S == PL::NF - 1);
    cplx<T> y[E];
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + TPL * i;
        const int k = b & (NS - 1);
        cplx<T> u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cplx<T> x = v[i + r * B];
            if (NS > 1 && r > 0) x = cmul(x, tw[k * r * TSTEP]);
            u[r] = x;
        }
        bfly<T, R>(u);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) y[i + q * B] = u[q];
        } else {
            const int base = (b / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) xc.put(base + q * NS, u[q]);
        }
    }
    if constexpr (LAST) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = y[e];
    } else {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = xc.get(t + TPL * e);
        __syncthreads();
    }
}

template <typename T, int N, int S, class X>
__device__ __forceinline__ void reg_fft_from(cplx<T>* v, int t, const cplx<T>* __restrict__ tw, X& xc) {
    if constexpr (S < P2Plan<N>::NF) {
        reg_stage<T, N, S>(v, t, tw, xc);
        reg_fft_from<T, N, S + 1>(v, t, tw, xc);
    }
}

// Forward FFT of the thread's line; v[e] = X[t + TPL e] on entry and exit.
template <typename T, int N, class X>
__device__ __forceinline__ void reg_fft(cplx<T>* v, int t, const cplx<T>* __restrict__ tw, X& xc) {
    reg_fft_from<T, N, 0>(v, t, tw, xc);
}

// Fetch Z[(N - k) mod N] for every element k = t + TPL e of the thread
// (one LDS exchange).
template <typename T, int N, class X>
__device__ __forceinline__ void reg_mirror(const cplx<T>* v, cplx<T>* m, int t, X& xc) {
    constexpr int E = P2Plan<N>::E, TPL = P2Plan<N>::TPL;
#pragma unroll
    for (int e = 0; e < E; ++e) xc.put(t + TPL * e, v[e]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) m[e] = xc.get((N - (t + TPL * e)) & (N - 1));
    __syncthreads();
}

}  // namespace fi
