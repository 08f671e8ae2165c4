// fft_reg.hpp -- register-resident Stockham FFT for power-of-two lengths.
//
// A line of N complex values is processed by TPL = N / 8 threads that each hold
// E = 8 elements, element e of thread t being X[t + TPL * e].  Stage 0 reads
// its inputs in exactly that order, so it is fed straight from global memory;
// the last stage leaves its outputs in the same order in registers.  The
// stages in between exchange data through LDS: with one buffer, a barrier
// between the writes and the reads of an exchange and one before the buffer
// is written again; with two buffers used in turn (ColXch2), only the first
// -- the barrier of exchange j + 1 already orders exchange j's reads of a
// buffer before exchange j + 2's writes to it.  Radices: 8 for every stage
// but a possible final 4 or 2.
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

// Register layout of a line of N = 2^k or 3 * 2^k values: TPL threads with
// E = 8 elements each, element e of thread t being X[t + TPL e].
template <int N> struct RegPlan {
    static constexpr int E = 8;
    static constexpr int TPL = N / E;
    static_assert(N % E == 0, "line length");
};

// Exchange buffers: where element `pos` of the calling thread's line lives,
// and what ends an exchange once its reads are issued (`release`).  Rows:
// line-major with one pad element per 8 (spreads the stride-R stage writes
// over the banks); columns: position-major with the lines fastest and one pad
// per row.
template <typename T, int N> struct RowXch {
    static constexpr int PITCH = N + N / 8 + 1;
    cplx<T>* base;
    int line;
    __device__ __forceinline__ RowXch(cplx<T>* b, int l) : base(b), line(l) {}
    __device__ __forceinline__ int addr(int pos) const { return line * PITCH + pos + (pos >> 3); }
    __device__ __forceinline__ void put(int pos, cplx<T> v) { base[addr(pos)] = v; }
    __device__ __forceinline__ cplx<T> get(int pos) const { return base[addr(pos)]; }
    __device__ __forceinline__ void release() { __syncthreads(); }
    static constexpr bool TURNS = false;
};

template <typename T, int N> struct ColXch {
    cplx<T>* base;
    int line, pitch;
    __device__ __forceinline__ ColXch(cplx<T>* b, int l, int C) : base(b), line(l), pitch(C + 1) {}
    __device__ __forceinline__ int addr(int pos) const { return pos * pitch + line; }
    __device__ __forceinline__ void put(int pos, cplx<T> v) { base[addr(pos)] = v; }
    __device__ __forceinline__ cplx<T> get(int pos) const { return base[addr(pos)]; }
    __device__ __forceinline__ void release() { __syncthreads(); }
    static constexpr bool TURNS = false;
};

// The column layout in two buffers of N (C + 1) values each, used in turn:
// `release` switches buffers instead of waiting.  The switches happen in
// straight-line code from a known start, so `cur` folds to a constant.
template <typename T, int N> struct ColXch2 {
    cplx<T>* base;
    int line, pitch, cur;
    __device__ __forceinline__ ColXch2(cplx<T>* b, int l, int C) : base(b), line(l), pitch(C + 1), cur(0) {}
    __device__ __forceinline__ int addr(int pos) const { return cur * N * pitch + pos * pitch + line; }
    __device__ __forceinline__ void put(int pos, cplx<T> v) { base[addr(pos)] = v; }
    __device__ __forceinline__ cplx<T> get(int pos) const { return base[addr(pos)]; }
    __device__ __forceinline__ void release() { cur ^= 1; }
    static constexpr bool TURNS = true;
};

// TWS: stride of the twiddle table (1: the table of this length; 3: the
// table of length 3N, for the sub-transforms of the N = 3 * 2^k engine)
template <typename T, int N, int S, int TWS, class X>
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
            if (NS > 1 && r > 0) x = cmul(x, tw[k * r * TSTEP * TWS]);
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
        xc.release();
    }
}

template <typename T, int N, int S, int TWS, class X>
__device__ __forceinline__ void reg_fft_from(cplx<T>* v, int t, const cplx<T>* __restrict__ tw, X& xc) {
    if constexpr (S < P2Plan<N>::NF) {
        reg_stage<T, N, S, TWS>(v, t, tw, xc);
        reg_fft_from<T, N, S + 1, TWS>(v, t, tw, xc);
    }
}

// Forward FFT of the thread's line; v[e] = X[t + TPL e] on entry and exit.
template <typename T, int N, class X>
__device__ __forceinline__ void reg_fft_p2(cplx<T>* v, int t, const cplx<T>* __restrict__ tw, X& xc) {
    reg_fft_from<T, N, 0, 1>(v, t, tw, xc);
}

// Fetch Z[(N - k) mod N] for every element k = t + TPL e of the thread
// (one LDS exchange).  FINAL: the caller does not use the exchange buffer
// again, and the exchange is not closed.
template <typename T, int N, bool FINAL, class X>
__device__ __forceinline__ void reg_mirror_p2(const cplx<T>* v, cplx<T>* m, int t, X& xc) {
    constexpr int E = P2Plan<N>::E, TPL = P2Plan<N>::TPL;
#pragma unroll
    for (int e = 0; e < E; ++e) xc.put(t + TPL * e, v[e]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) m[e] = xc.get((N - (t + TPL * e)) & (N - 1));
    if constexpr (!FINAL) xc.release();
}

// ---------------------------------------------------------------- N = 3 * 2^k
// Lengths with one factor 3 (96, 192, 384, 768) run the same register
// layout: TPL = N / 8 threads per line, element e of thread T being
// X[T + TPL e].  With T = 3t + r that is X[3 (t + (N/24) e) + r]: element e
// of thread t of sub-sequence r, in exactly the layout of an M = N / 3 point
// register FFT.  The transform (decimation in time, N = 3 x M):
//   1. three M-point register FFTs Y_r (sub-line r at positions r M .. r M + M - 1
//      of the line's exchange buffer; their twiddles are every third entry of
//      the N-point table),
//   2. Y_r[k] *= exp(-2 pi i r k / N),
//   3. one exchange and per k the output q = r of a radix-3 butterfly:
//      thread (t, r) forms Z[k + M r] = sum_s exp(-2 pi i r s / 3) Y_s[k],
//   4. one exchange that puts every Z at its position and reads back the
//      layout X[T + TPL e] together with the Hartley mirrors Z[N - k].
template <int N> struct IsPow2 { static constexpr bool value = N >= 8 && (N & (N - 1)) == 0; };
template <int N> struct Is3Pow2 {
    static constexpr bool value = N % 3 == 0 && IsPow2<N / 3>::value && N / 3 >= 16;
};

// An exchange buffer seen from offset `off` (sub-line r at r M); holds a
// copy of the exchange object (a pointer and an index or two), so nothing
// needs an address in private memory.
template <class X> struct SubXch {
    X x;
    int off;
    __device__ __forceinline__ SubXch(const X& xx, int o) : x(xx), off(o) {}
    template <class V> __device__ __forceinline__ void put(int pos, V v) { x.put(off + pos, v); }
    __device__ __forceinline__ auto get(int pos) const { return x.get(off + pos); }
    // a copy cannot switch the caller's buffer: single-buffered exchanges only
    static_assert(!X::TURNS, "sub-line views of single-buffered exchanges only");
    __device__ __forceinline__ void release() { x.release(); }
};

template <typename T, int N, bool FINAL, class X>
__device__ __forceinline__ void r3_fft_mirror(cplx<T>* v, cplx<T>* m, int tt, const cplx<T>* __restrict__ tw, X& xc) {
    constexpr int M = N / 3, E = 8, TPL = N / 8, TM = M / 8;
    const int t = tt / 3, r = tt - 3 * t;
    SubXch<X> sx(xc, r * M);
    reg_fft_from<T, M, 0, 3>(v, t, tw, sx);   // v[e] = Y_r[t + TM e]
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TM * e;
        // r k < 2 M < N: no reduction of the table index
        if (r > 0) v[e] = cmul(v[e], tw[r * k]);
        xc.put(r * M + k, v[e]);
    }
    __syncthreads();
    // the one output q = r of the radix-3 butterfly:
    // Z = y0 + w y1 + w^2 y2, w = exp(-2 pi i r / 3) = c + i s, w^2 = c - i s (r > 0)
    const T c = r == 0 ? (T)1 : (T)-0.5;
    const T sn = r == 0 ? (T)0 : (r == 1 ? (T)-0.86602540378443864676 : (T)0.86602540378443864676);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TM * e;
        const cplx<T> y0 = xc.get(k), y1 = xc.get(M + k), y2 = xc.get(2 * M + k);
        // w y1 + w^2 y2 = c (y1 + y2) + i s (y1 - y2)
        const cplx<T> a = cadd(y1, y2), b = csub(y1, y2);
        v[e] = mk<T>(y0.x + (c * a.x - sn * b.y), y0.y + (c * a.y + sn * b.x));
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) xc.put(t + TM * e + M * r, v[e]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = tt + TPL * e;
        v[e] = xc.get(k);
        m[e] = xc.get(k == 0 ? 0 : N - k);
    }
    if constexpr (!FINAL) __syncthreads();
}

// Forward FFT of the thread's line with the Hartley mirrors, for N = 2^k or
// 3 * 2^k: on exit v[e] = Z[k], m[e] = Z[(N - k) mod N], k = t + TPL e.
// FINAL: the caller does not use the exchange buffer again (the last
// exchange's reads need no barrier behind them).
template <typename T, int N, bool FINAL = false, class X>
__device__ __forceinline__ void reg_fft_mirror(cplx<T>* v, cplx<T>* m, int t, const cplx<T>* __restrict__ tw, X& xc) {
    if constexpr (IsPow2<N>::value) {
        reg_fft_p2<T, N>(v, t, tw, xc);
        reg_mirror_p2<T, N, FINAL>(v, m, t, xc);
    } else {
        static_assert(Is3Pow2<N>::value, "register FFT lengths: 2^k or 3 * 2^k");
        r3_fft_mirror<T, N, FINAL>(v, m, t, tw, xc);
    }
}

}  // namespace fi
