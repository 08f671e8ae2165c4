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
//   c_rows_c2r   two cropped rows per inverse FFT, scale 1/M^2, subtract the
//                corner correction of apply_zerobc_at_corners_dp on the full
//                (2N-1)^2 array, keep the crop columns (the +1-cell shift of
//                convolutions.f90:34-39,89 is kept)
// The four corners of the full linear convolution are exact products of input
// and kernel corners, so the correction needs no extra pass.
#include "fi_device.hpp"
#include "fi_kernels.h"

namespace fi {

// pass 1: input rows -> half spectra.  src real N x N; dst complex N x (N+1)
template <typename T>
__global__ void __launch_bounds__(512) c_rows_r2c(const Member* __restrict__ mem, FftPlan plan,
                                                  const void* __restrict__ twv, int L, int which_src) {
    const Member& M = mem[blockIdx.y];
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    const int N = M.n, Mm = 2 * N, r0 = blockIdx.x * 2 * L;
    const cplx<T>* tw = (const cplx<T>*)twv;
    const double* __restrict__ src = M.conv_in;
    (void)which_src;
    const int tid = threadIdx.x, nt = blockDim.x;
    T* bf = (T*)buf;
    for (int q = tid; q < 2 * L * Mm; q += nt) {
        const int mr = q / Mm, i = q % Mm;
        const int l = mr % L, part = mr / L;
        const int row = r0 + mr;
        bf[2 * (i * L + l) + part] = (i < N && row < N) ? (T)src[(size_t)row * N + i] : (T)0;
    }
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
    const cplx<T>* tw = (const cplx<T>*)twv;
    const cplx<T>* __restrict__ w1 = (const cplx<T>*)M.conv_work;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < Mm * C; q += nt) {
        const int r = q / C, l = q % C;
        const int kx = k0 + l;
        buf[r * C + l] = (r < N && kx < H) ? w1[(size_t)r * H + kx] : mk<T>((T)0, (T)0);
    }
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
    for (int q = tid; q < Mm * C; q += nt) {
        const int r = q / C, l = q % C;
        const int kx = k0 + l;
        cplx<T> v = mk<T>((T)0, (T)0);
        if (kx < H) v = cmul(spec[(size_t)r * H + kx], buf[r * C + l]);
        v.y = -v.y;
        buf[r * C + l] = v;
    }
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

// pass 3: cropped rows back to real space, scale, corner correction, crop.
template <typename T>
__global__ void __launch_bounds__(512) c_rows_c2r(const Member* __restrict__ mem, FftPlan plan,
                                                  const void* __restrict__ twv, int L, int which) {
    const Member& M = mem[blockIdx.y];
    // w_equilibrium is only recomputed for domains that step in this update
    if (which == 2 && M.ctl->status != ST_RUN) return;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    __shared__ double corr;
    const int N = M.n, Mm = 2 * N, H = N + 1, r0 = blockIdx.x * 2 * L;
    const cplx<T>* tw = (const cplx<T>*)twv;
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
    for (int q = tid; q < L * Mm; q += nt) {
        const int l = q / Mm, k = q % Mm;
        const int ra = r0 + l, rb = r0 + L + l;
        cplx<T> xa, xb;
        if (k <= N) { xa = w2[(size_t)ra * H + k]; xb = w2[(size_t)rb * H + k]; }
        else {
            xa = w2[(size_t)ra * H + (Mm - k)]; xa.y = -xa.y;
            xb = w2[(size_t)rb * H + (Mm - k)]; xb.y = -xb.y;
        }
        // z = xa + i xb, conjugated
        buf[k * L + l] = mk<T>(xa.x - xb.y, -(xa.y + xb.x));
    }
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

template <typename T>
static size_t lbytes(int m, int lines) { return (size_t)m * lines * sizeof(cplx<T>); }

template <typename T>
static void conv_t(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    const int n = s.n, H = n + 1;
    hipLaunchKernelGGL(c_rows_r2c<T>, dim3((n + 2 * s.Lr - 1) / (2 * s.Lr), nmem), dim3(s.threads_row),
                       lbytes<T>(s.m, s.Lr), st, d_mem, s.plan, s.tw, s.Lr, which);
    hipLaunchKernelGGL(c_cols<T>, dim3((H + s.Cc - 1) / s.Cc, nmem), dim3(s.threads_col), lbytes<T>(s.m, s.Cc), st,
                       d_mem, s.plan, s.tw, s.Cc, 1, which);
    hipLaunchKernelGGL(c_rows_c2r<T>, dim3(n / (2 * s.Lr), nmem), dim3(s.threads_row), lbytes<T>(s.m, s.Lr), st,
                       d_mem, s.plan, s.tw, s.Lr, which);
}

void launch_convolution(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    if (s.fp64) conv_t<double>(d_mem, nmem, s, which, st);
    else conv_t<float>(d_mem, nmem, s, which, st);
}

// Kernel spectrum: the kernel must already be in conv_in.
template <typename T>
static void spec_t(const Member* d_mem, int nmem, const ConvShape& s, int which, hipStream_t st) {
    const int n = s.n, H = n + 1;
    hipLaunchKernelGGL(c_rows_r2c<T>, dim3((n + 2 * s.Lr - 1) / (2 * s.Lr), nmem), dim3(s.threads_row),
                       lbytes<T>(s.m, s.Lr), st, d_mem, s.plan, s.tw, s.Lr, which);
    hipLaunchKernelGGL(c_cols<T>, dim3((H + s.Cc - 1) / s.Cc, nmem), dim3(s.threads_col), lbytes<T>(s.m, s.Cc), st,
                       d_mem, s.plan, s.tw, s.Cc, 0, which);
}

void launch_kernel_spectrum(const Member* d_mem, int nmem, const ConvShape& s, int which,
                            const double* d_kernel, hipStream_t st) {
    (void)d_kernel;
    if (s.fp64) spec_t<double>(d_mem, nmem, s, which, st);
    else spec_t<float>(d_mem, nmem, s, which, st);
}

}  // namespace fi
