// conv.hip -- linear convolution with a precomputed Green's-function spectrum
// (precomputed_fftconvolution / precompute_kernel, convolutions.f90:59-108).
//
// The reference zero-pads to (2N-1)^2 and uses FFTW r2c/c2r; any transform
// size >= 2N-1 gives the same linear convolution, so here M = 2N (smooth
// whenever N is).  Three passes per convolution, half-spectrum layout
// [ky = 0..M-1][kx = 0..N]:
//   c_rows_r2c   rows j < N of the input (rows >= N are zero): two real rows per
//                complex FFT of length M, split into their half spectra
//   c_cols       per kx column: FFT over ky (input rows >= N are zero),
//                multiply by the kernel spectrum, inverse FFT, keep only the
//                N output rows of the crop window [N/2-1, 3N/2-1)
// For M <= 1024 (FP64: M <= 256) a convolution runs the same three passes on
// the register FFT engine (c_rows_r2c_reg, c_cols_reg, c_rows_c2r_reg); the
// LDS Stockham kernels below serve other lengths and the kernel spectra.
//   c_rows_c2r   two cropped rows per inverse FFT, scale 1/M^2, subtract the
//                corner correction of apply_zerobc_at_corners_dp on the full
//                (2N-1)^2 array, keep the crop columns (the +1-cell shift of
//                convolutions.f90:34-39,89 is kept)
// The four corners of the full linear convolution are exact products of input
// and kernel corners, so the correction needs no extra pass.
#include "fft_reg.hpp"
#include "fi_device.hpp"
#include "fi_kernels.h"

namespace fi {

// pass 1: input rows -> half spectra.  src real N x N; dst complex N x (N+1)
// TWL: the twiddle table is held in LDS after the L lines.
template <typename T, bool TWL>
__global__ void __launch_bounds__(512) c_rows_r2c(const Member* __restrict__ mem, FftPlan plan,
                                                  const void* __restrict__ twv, int L, int which_src) {
    const Member& M = mem[blockIdx.y];
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    const int N = M.n, Mm = 2 * N, r0 = blockIdx.x * 2 * L;
    const cplx<T>* tw = TWL ? stage_twiddles<T>(buf, Mm, L, twv) : (const cplx<T>*)twv;
    const double* __restrict__ src = M.conv_in;
    (void)which_src;
    const int tid = threadIdx.x, nt = blockDim.x;
    T* bf = (T*)buf;
    // zero padding: rows >= N and columns >= N of the (2N)^2 input
    auto inside = [&](int q) { return q % Mm < N && r0 + q / Mm < N; };
    staged_copy<8>(2 * L * Mm,
                   [&](int q) { return inside(q) ? src[(size_t)(r0 + q / Mm) * N + q % Mm] : src[0]; },
                   [&](int q, double v) {
                       const int mr = q / Mm, i = q % Mm;
                       bf[2 * (i * L + mr % L) + mr / L] = inside(q) ? (T)v : (T)0;
                   });
    __syncthreads();
    lds_fft<T>(buf, plan, L, L, tw);
    cplx<T>* dst = (cplx<T>*)M.conv_work;
    const int H = N + 1;
    for (int q = tid; q < 2 * L * H; q += nt) {
        const int mr = q / H, k = q % H;
        const int row = r0 + mr;
        if (row >= N) continue;
        const int l = mr % L, part = mr / L;
        const cplx<T> zk = buf[k * L + l], zm = buf[((Mm - k) % Mm) * L + l];
        cplx<T> v;
        if (part == 0) v = mk<T>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y));
        else v = mk<T>((T)0.5 * (zk.y + zm.y), -(T)0.5 * (zk.x - zm.x));
        dst[(size_t)row * H + k] = v;
    }
}

// pass 2: columns.  mode 0 = spectrum of a kernel (write all M rows to spec);
// mode 1 = convolve with spec (which: 0 elastic, 1 sea surface, 2 viscous) and keep crop rows.
template <typename T>
__global__ void __launch_bounds__(512) c_cols(const Member* __restrict__ mem, FftPlan plan,
                                              const void* __restrict__ twv, int C, int mode, int which) {
    const Member& M = mem[blockIdx.y];
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    const int N = M.n, Mm = 2 * N, H = N + 1;
    const int k0 = blockIdx.x * C;
    const cplx<T>* tw = stage_twiddles<T>(buf, Mm, C, twv);   // the host sizes C to leave room
    const cplx<T>* __restrict__ w1 = (const cplx<T>*)M.conv_work;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto inside = [&](int q) { return q / C < N && k0 + q % C < H; };
    staged_copy<8>(Mm * C, [&](int q) { return inside(q) ? w1[(size_t)(q / C) * H + k0 + q % C] : w1[0]; },
                   [&](int q, cplx<T> v) { buf[q] = inside(q) ? v : mk<T>((T)0, (T)0); });
    __syncthreads();
    lds_fft<T>(buf, plan, C, C, tw);
    cplx<T>* spec = (cplx<T>*)(which == 0 ? M.FGE : which == 1 ? M.FGN : M.FGV);
    if (mode == 0) {
        for (int q = tid; q < Mm * C; q += nt) {
            const int r = q / C, l = q % C;
            const int kx = k0 + l;
            if (kx < H) spec[(size_t)r * H + kx] = buf[r * C + l];
        }
        return;
    }
    // product with the kernel spectrum, then conjugate for the inverse FFT
    staged_copy<8>(Mm * C,
                   [&](int q) { return k0 + q % C < H ? spec[(size_t)(q / C) * H + k0 + q % C] : spec[0]; },
                   [&](int q, cplx<T> sv) {
                       cplx<T> v = mk<T>((T)0, (T)0);
                       if (k0 + q % C < H) v = cmul(sv, buf[q]);
                       v.y = -v.y;
                       buf[q] = v;
                   });
    __syncthreads();
    lds_fft<T>(buf, plan, C, C, tw);
    cplx<T>* w2 = (cplx<T>*)M.conv_work + (size_t)N * H;
    const int c0 = N / 2 - 1 + (N % 2);   // 0-based crop start (convolutions.f90:34-39)
    for (int q = tid; q < N * C; q += nt) {
        const int jj = q / C, l = q % C;
        const int kx = k0 + l;
        if (kx >= H) continue;
        cplx<T> v = buf[(jj + c0) * C + l];
        v.y = -v.y;
        w2[(size_t)jj * H + kx] = v;
    }
}

// pass 2 of a convolution (mode 1 of c_cols) on the register engine, for
// power-of-two M = 2N <= 1024: C columns per block, TPL = M/8 threads per
// column, element e of thread t being row t + TPL e.  Input rows >= N are
// zero (elements 4..7 are not loaded); the product with the kernel spectrum
// is conjugated so that a second forward transform gives the conjugated
// inverse; only the N crop rows are stored.  Columns past kx = N (the last
// block's spare lines) transform a copy of column N and store nothing.
template <typename T, int M, int C>
__global__ void __launch_bounds__(C * (M / 8)) c_cols_reg(const Member* __restrict__ mem, const void* __restrict__ twv,
                                                          int which) {
    constexpr int N = M / 2, H = N + 1, TPL = M / 8, E = 8, THREADS = C * TPL;
    const Member& Mb = mem[blockIdx.y];
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + M;
    const int tid = threadIdx.x, line = tid % C, t = tid / C;
    const int kx = blockIdx.x * C + line;
    const bool live = kx < H;
    const int kc = live ? kx : H - 1;
    const cplx<T>* __restrict__ g = (const cplx<T>*)twv;
    cplx<T> twr[(M + THREADS - 1) / THREADS];
#pragma unroll
    for (int k = 0; k < (M + THREADS - 1) / THREADS; ++k)
        if (tid + THREADS * k < M) twr[k] = g[tid + THREADS * k];
    const cplx<T>* __restrict__ w1 = (const cplx<T>*)Mb.conv_work;
    const cplx<T>* __restrict__ spec = (const cplx<T>*)(which == 0 ? Mb.FGE : which == 1 ? Mb.FGN : Mb.FGV);
    cplx<T> v[E], sp[E];
#pragma unroll
    for (int e = 0; e < E / 2; ++e) v[e] = w1[(size_t)(t + TPL * e) * H + kc];
#pragma unroll
    for (int e = E / 2; e < E; ++e) v[e] = mk<T>((T)0, (T)0);
#pragma unroll
    for (int e = 0; e < E; ++e) sp[e] = spec[(size_t)(t + TPL * e) * H + kc];
#pragma unroll
    for (int k = 0; k < (M + THREADS - 1) / THREADS; ++k)
        if (tid + THREADS * k < M) tw_l[tid + THREADS * k] = twr[k];
    __syncthreads();
    ColXch<T, M> xc(buf, line, C);
    reg_fft_p2<T, M>(v, t, tw_l, xc);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        cplx<T> p = cmul(sp[e], v[e]);
        p.y = -p.y;
        v[e] = p;
    }
    // (the first transform's last exchange ended with a barrier: the buffer is free)
    reg_fft_p2<T, M>(v, t, tw_l, xc);
    if (!live) return;
    cplx<T>* __restrict__ w2 = (cplx<T>*)Mb.conv_work + (size_t)N * H;
    constexpr int c0 = N / 2 - 1;   // 0-based crop start (convolutions.f90:34-39), N even
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int jj = t + TPL * e - c0;
        if (jj >= 0 && jj < N) {
            cplx<T> o = v[e];
            o.y = -o.y;
            w2[(size_t)jj * H + kx] = o;
        }
    }
}

// pass 3: cropped rows back to real space, scale, corner correction, crop.
template <typename T, bool TWL>
__global__ void __launch_bounds__(512) c_rows_c2r(const Member* __restrict__ mem, FftPlan plan,
                                                  const void* __restrict__ twv, int L, int which) {
    const Member& M = mem[blockIdx.y];
    // w_equilibrium is only recomputed for domains that step in this update
    if (which == 2 && M.ctl->status != ST_RUN) return;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    __shared__ double corr;
    const int N = M.n, Mm = 2 * N, H = N + 1, r0 = blockIdx.x * 2 * L;
    const cplx<T>* tw = TWL ? stage_twiddles<T>(buf, Mm, L, twv) : (const cplx<T>*)twv;
    const cplx<T>* __restrict__ w2 = (const cplx<T>*)M.conv_work + (size_t)N * H;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        const double* u = M.conv_in;
        const double* kc = (which == 0) ? M.kc_E : (which == 1) ? M.kc_N : M.kc_V;
        const double f00 = u[0] * kc[0];
        const double fee = u[(size_t)N * N - 1] * kc[1];
        const double fe0 = u[(size_t)(N - 1) * N] * kc[2];
        const double f0e = u[N - 1] * kc[3];
        corr = 0.25 * (f00 + fee + fe0 + f0e);
    }
    // Hermitian extension of two half-spectrum rows, conjugated for the inverse
    struct Pair { cplx<T> a, b; };
    staged_copy<4>(L * Mm,
                   [&](int q) {
                       const int l = q / Mm, k = q % Mm, kk = k <= N ? k : Mm - k;
                       return Pair{w2[(size_t)(r0 + l) * H + kk], w2[(size_t)(r0 + L + l) * H + kk]};
                   },
                   [&](int q, Pair p) {
                       const int l = q / Mm, k = q % Mm;
                       cplx<T> xa = p.a, xb = p.b;
                       if (k > N) { xa.y = -xa.y; xb.y = -xb.y; }
                       // z = xa + i xb, conjugated
                       buf[k * L + l] = mk<T>(xa.x - xb.y, -(xa.y + xb.x));
                   });
    __syncthreads();
    lds_fft<T>(buf, plan, L, L, tw);
    const double scale = (double)Mm * (double)Mm;
    const double cc = corr;
    const int c0 = N / 2 - 1 + (N % 2);
    double* out = (which == 0) ? M.we : (which == 1) ? M.dz_ss : M.weq;
    for (int q = tid; q < 2 * L * N; q += nt) {
        const int mr = q / N, i = q % N;
        const int l = mr % L, part = mr / L;
        const cplx<T> z = buf[(i + c0) * L + l];
        // conj of the FFT: real part = row a, -(imag) = row b
        const double v = part ? -(double)z.y : (double)z.x;
        out[(size_t)(r0 + mr) * N + i] = v / scale - cc;
    }
}

// passes 1 and 3 on the register engine (M = 2N <= 1024): a block holds L
// complex lines of TPL = M/8 threads, line l packing rows 2j and 2j+1
// (j = block * L + l), element e of thread t being column t + TPL e.
//
// pass 1: columns >= N are zero (elements 4..7 are not loaded); the half
// spectra k = 0 .. N of the two rows from Z[k] and Z[M - k].
template <typename T, int M, int L>
__global__ void __launch_bounds__(L * (M / 8)) c_rows_r2c_reg(const Member* __restrict__ mem,
                                                              const void* __restrict__ twv) {
    constexpr int N = M / 2, H = N + 1, TPL = M / 8, E = 8, THREADS = L * TPL, TWQ = (M + THREADS - 1) / THREADS;
    const Member& Mb = mem[blockIdx.y];
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + M;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    const int ra = 2 * (blockIdx.x * L + line), rb = ra + 1;
    const cplx<T>* __restrict__ g = (const cplx<T>*)twv;
    cplx<T> twr[TWQ];
#pragma unroll
    for (int k = 0; k < TWQ; ++k)
        if (tid + THREADS * k < M) twr[k] = g[tid + THREADS * k];
    const double* __restrict__ src = Mb.conv_in;
    double xa[E / 2], xb[E / 2];
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
        xa[e] = src[(size_t)ra * N + t + TPL * e];
        xb[e] = src[(size_t)rb * N + t + TPL * e];
    }
#pragma unroll
    for (int k = 0; k < TWQ; ++k)
        if (tid + THREADS * k < M) tw_l[tid + THREADS * k] = twr[k];
    __syncthreads();
    cplx<T> v[E], m[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = e < E / 2 ? mk<T>((T)xa[e % (E / 2)], (T)xb[e % (E / 2)]) : mk<T>((T)0, (T)0);
    RowXch<T, M> xc(buf, line);
    reg_fft_mirror<T, M, true>(v, m, t, tw_l, xc);
    cplx<T>* __restrict__ dst = (cplx<T>*)Mb.conv_work;
#pragma unroll
    for (int e = 0; e <= E / 2; ++e) {
        const int k = t + TPL * e;
        if (e < E / 2 || t == 0) {   // k = 0 .. N
            const cplx<T> zk = v[e], zm = m[e];
            dst[(size_t)ra * H + k] = mk<T>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y));
            dst[(size_t)rb * H + k] = mk<T>((T)0.5 * (zk.y + zm.y), -(T)0.5 * (zk.x - zm.x));
        }
    }
}

// pass 3: the Hermitian extensions of two cropped half-spectrum rows as one
// conjugated complex line, one forward transform, scale, corner correction,
// crop columns [N/2 - 1, 3N/2 - 1).
template <typename T, int M, int L>
__global__ void __launch_bounds__(L * (M / 8)) c_rows_c2r_reg(const Member* __restrict__ mem,
                                                              const void* __restrict__ twv, int which) {
    constexpr int N = M / 2, H = N + 1, TPL = M / 8, E = 8, THREADS = L * TPL, TWQ = (M + THREADS - 1) / THREADS;
    const Member& Mb = mem[blockIdx.y];
    // w_equilibrium is only recomputed for domains that step in this update
    if (which == 2 && Mb.ctl->status != ST_RUN) return;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + M;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    const int ra = 2 * (blockIdx.x * L + line), rb = ra + 1;
    const cplx<T>* __restrict__ g = (const cplx<T>*)twv;
    cplx<T> twr[TWQ];
#pragma unroll
    for (int k = 0; k < TWQ; ++k)
        if (tid + THREADS * k < M) twr[k] = g[tid + THREADS * k];
    const cplx<T>* __restrict__ w2 = (const cplx<T>*)Mb.conv_work + (size_t)N * H;
    cplx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TPL * e, kk = k <= N ? k : M - k;
        cplx<T> xa = w2[(size_t)ra * H + kk], xb = w2[(size_t)rb * H + kk];
        if (k > N) { xa.y = -xa.y; xb.y = -xb.y; }
        // z = xa + i xb, conjugated
        v[e] = mk<T>(xa.x - xb.y, -(xa.y + xb.x));
    }
    // the corner correction of apply_zerobc_at_corners_dp on the full array
    const double* __restrict__ u = Mb.conv_in;
    const double* kc = (which == 0) ? Mb.kc_E : (which == 1) ? Mb.kc_N : Mb.kc_V;
    const double f00 = u[0] * kc[0];
    const double fee = u[(size_t)N * N - 1] * kc[1];
    const double fe0 = u[(size_t)(N - 1) * N] * kc[2];
    const double f0e = u[N - 1] * kc[3];
    const double cc = 0.25 * (f00 + fee + fe0 + f0e);
#pragma unroll
    for (int k = 0; k < TWQ; ++k)
        if (tid + THREADS * k < M) tw_l[tid + THREADS * k] = twr[k];
    __syncthreads();
    RowXch<T, M> xc(buf, line);
    reg_fft_p2<T, M>(v, t, tw_l, xc);
    const double scale = (double)M * (double)M;
    constexpr int c0 = N / 2 - 1;   // N even
    double* __restrict__ out = (which == 0) ? Mb.we : (which == 1) ? Mb.dz_ss : Mb.weq;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = t + TPL * e - c0;
        if (i >= 0 && i < N) {
            // conj of the FFT: real part = row a, -(imag) = row b
            out[(size_t)ra * N + i] = (double)v[e].x / scale - cc;
            out[(size_t)rb * N + i] = -(double)v[e].y / scale - cc;
        }
    }
}

template <typename T>
static size_t lbytes(int m, int lines) { return (size_t)m * lines * sizeof(cplx<T>); }

// lines per block of the register row passes: the most within 1024 threads and
// 64 KiB of LDS (M twiddles and L padded lines) that divide the N / 2 line pairs
template <typename T, int M>
constexpr int reg_row_lines() {
    constexpr auto fits = [](int L) {
        return L * (M / 8) <= 1024 && ((size_t)M + (size_t)L * RowXch<T, M>::PITCH) * sizeof(cplx<T>) <= 65536 &&
               (M / 4) % L == 0;
    };
    return fits(8) ? 8 : fits(4) ? 4 : fits(2) ? 2 : 1;
}
template <typename T, int M>
static void rows_reg(const Member* d_mem, int nmem, const void* tw, int which, hipStream_t st, bool inverse) {
    constexpr int L = reg_row_lines<T, M>();
    const size_t lds = ((size_t)M + (size_t)L * RowXch<T, M>::PITCH) * sizeof(cplx<T>);
    const dim3 g(M / 4 / L, nmem), b(L * (M / 8));
    if (inverse) hipLaunchKernelGGL((c_rows_c2r_reg<T, M, L>), g, b, lds, st, d_mem, tw, which);
    else hipLaunchKernelGGL((c_rows_r2c_reg<T, M, L>), g, b, lds, st, d_mem, tw);
}

template <typename T>
static void rows_r2c(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    const dim3 g((s.n + 2 * s.Lr - 1) / (2 * s.Lr), nmem), b(s.threads_row);
    if (s.rows_twl)
        hipLaunchKernelGGL((c_rows_r2c<T, true>), g, b, lbytes<T>(s.m, s.Lr + 1), st, d_mem, s.plan, s.tw, s.Lr, which);
    else
        hipLaunchKernelGGL((c_rows_r2c<T, false>), g, b, lbytes<T>(s.m, s.Lr), st, d_mem, s.plan, s.tw, s.Lr, which);
}

// the register-engine column pass: the widest column blocks within 1024
// threads and 64 KiB of LDS (M twiddles and the M x (C + 1) exchange buffer)
template <typename T, int M>
static void cols_reg(const Member* d_mem, int nmem, const void* tw, int which, hipStream_t st) {
    constexpr int H = M / 2 + 1;
    constexpr auto fits = [](int C) {
        return C * (M / 8) <= 1024 && ((size_t)M + (size_t)M * (C + 1)) * sizeof(cplx<T>) <= 65536;
    };
    constexpr int C = fits(8) ? 8 : fits(4) ? 4 : fits(2) ? 2 : 1;
    static_assert(fits(C), "register column pass does not fit");
    const size_t lds = ((size_t)M + (size_t)M * (C + 1)) * sizeof(cplx<T>);
    hipLaunchKernelGGL((c_cols_reg<T, M, C>), dim3((H + C - 1) / C, nmem), dim3(C * (M / 8)), lds, st, d_mem, tw,
                       which);
}

template <typename T>
static void conv_t(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    const int n = s.n, H = n + 1;
    // the register engine for M = 2N <= 1024 (FP64 lines of 512 and 1024
    // points would spill the register window: they stay on the LDS Stockham
    // kernels, as do other lengths)
    const bool reg = s.m == 128 || s.m == 256 || (sizeof(T) == 4 && (s.m == 512 || s.m == 1024));
    if (reg) {
        if (s.m == 128) rows_reg<T, 128>(d_mem, nmem, s.tw, which, st, false);
        else if (s.m == 256) rows_reg<T, 256>(d_mem, nmem, s.tw, which, st, false);
        else if (s.m == 512) rows_reg<float, 512>(d_mem, nmem, s.tw, which, st, false);
        else rows_reg<float, 1024>(d_mem, nmem, s.tw, which, st, false);
    } else {
        rows_r2c<T>(d_mem, nmem, s, which, st);
    }
    if (s.m == 128) cols_reg<T, 128>(d_mem, nmem, s.tw, which, st);
    else if (s.m == 256) cols_reg<T, 256>(d_mem, nmem, s.tw, which, st);
    else if (s.m == 512 && sizeof(T) == 4) cols_reg<float, 512>(d_mem, nmem, s.tw, which, st);
    else if (s.m == 1024 && sizeof(T) == 4) cols_reg<float, 1024>(d_mem, nmem, s.tw, which, st);
    else
        hipLaunchKernelGGL(c_cols<T>, dim3((H + s.Cc - 1) / s.Cc, nmem), dim3(s.threads_col),
                           lbytes<T>(s.m, s.Cc + 1), st, d_mem, s.plan, s.tw, s.Cc, 1, which);
    if (reg) {
        if (s.m == 128) rows_reg<T, 128>(d_mem, nmem, s.tw, which, st, true);
        else if (s.m == 256) rows_reg<T, 256>(d_mem, nmem, s.tw, which, st, true);
        else if (s.m == 512) rows_reg<float, 512>(d_mem, nmem, s.tw, which, st, true);
        else rows_reg<float, 1024>(d_mem, nmem, s.tw, which, st, true);
        return;
    }
    const dim3 g(n / (2 * s.Lr), nmem), b(s.threads_row);
    if (s.rows_twl)
        hipLaunchKernelGGL((c_rows_c2r<T, true>), g, b, lbytes<T>(s.m, s.Lr + 1), st, d_mem, s.plan, s.tw, s.Lr, which);
    else
        hipLaunchKernelGGL((c_rows_c2r<T, false>), g, b, lbytes<T>(s.m, s.Lr), st, d_mem, s.plan, s.tw, s.Lr, which);
}

void launch_convolution(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    if (s.fp64) conv_t<double>(d_mem, nmem, s, which, st);
    else conv_t<float>(d_mem, nmem, s, which, st);
}

// Kernel spectrum: the kernel must already be in conv_in.
template <typename T>
static void spec_t(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    const int n = s.n, H = n + 1;
    rows_r2c<T>(d_mem, nmem, s, which, st);
    hipLaunchKernelGGL(c_cols<T>, dim3((H + s.Cc - 1) / s.Cc, nmem), dim3(s.threads_col),
                       lbytes<T>(s.m, s.Cc + 1), st, d_mem, s.plan, s.tw, s.Cc, 0, which);
}

void launch_kernel_spectrum(const Member* d_mem, int nmem, const ConvShape& s, int which,
                            const double* d_kernel, hipStream_t st) {
    (void)d_kernel;
    if (s.fp64) spec_t<double>(d_mem, nmem, s, which, st);
    else spec_t<float>(d_mem, nmem, s, which, st);
}

}  // namespace fi
