// rhs.hip -- LV-ELVA right-hand side and adaptive step on the GPU.
//
// One attempted step of the degenerate Tsitouras loop (integrators.f90:163-200)
// for a batch of domains of equal size N, one blockIdx.y per domain:
//
//   k_stencil      FD flexure + load -> f (lv_elva.f90:95-113)
//   k_row_fwd      Hartley transform along x (first half of DHT2)
//   k_col          along y: DHT, divide by kappa*R (lv_elva.f90:115), DHT;
//                  also the per-column terms that give the four corners of
//                  the inverse transform (apply_zerobc_at_corners_dp)
//   k_row_inv      along x: DHT, /N^2, corner BC, * sec_per_year
//                  (lv_elva.f90:118-123), then the stage sums x1/x2 and the
//                  per-row max|x1|, max|x2|, sum (x2-x1)^2
//                  (after a rejected attempt the RHS is unchanged because it
//                  reads now%w, fastisostasy.f90:720: only this epilogue runs)
//   k_decide       the scalar tol, error norm, accept / reject and dt control
//
// Every kernel reads the domain's Ctl and returns at once if the domain is not
// stepping, so the same launch sequence serves a batch whose members need
// different numbers of attempts.
#include <atomic>
#include <vector>

#include "fi_device.hpp"
#include "fi_kernels.h"
#include "rhs_common.hpp"

namespace fi {

// ------------------------------------------------------------------ stencil
constexpr int ST_TX = 64, ST_TY = 8;

template <typename T, bool UNIFORM>
__global__ void __launch_bounds__(256) k_stencil(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.z];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    const PhysDev& ph = *M.phys;
    const int N = M.n;
    const int i0 = blockIdx.x * ST_TX, j0 = blockIdx.y * ST_TY;
    const double* __restrict__ w = c.w_is_x ? M.X[c.cur] : M.W;
    const double* __restrict__ dxm = M.dxm;
    const double dxs = ph.dx;
    // uniform spacing: divisions by dx, 2 dx and dx^2 become multiplications
    const double inv1 = 1.0 / dxs, inv2 = 1.0 / (2 * dxs), invq = 1.0 / (dxs * dxs);
    auto DX = [&](int i, int j) -> double { return dxm[(size_t)j * N + i]; };
    auto Q1 = [&](double a, int i, int j) -> double { return UNIFORM ? a * inv1 : a / DX(i, j); };
    auto Q2 = [&](double a, int i, int j) -> double { return UNIFORM ? a * inv2 : a / (2 * DX(i, j)); };
    auto QQ = [&](double a, int i, int j) -> double { return UNIFORM ? a * invq : a / (DX(i, j) * DX(i, j)); };

    constexpr int W0 = ST_TX + 4, H0 = ST_TY + 4;   // w:   [i0-2, i0+TX+2) x [j0-2, j0+TY+2)
    constexpr int W2 = ST_TX + 2, H2 = ST_TY + 2;   // M**: [i0-1, i0+TX+1) x [j0-1, j0+TY+1)
    __shared__ double sw[H0 * W0];
    __shared__ double sMxx[H2 * W2], sMyy[H2 * W2], sMxy[H2 * W2];
    const int tid = threadIdx.x, nt = blockDim.x;
#define SW(i, j) sw[((j) - (j0 - 2)) * W0 + ((i) - (i0 - 2))]
#define S2(a, i, j) a[((j) - (j0 - 1)) * W2 + ((i) - (i0 - 1))]

    for (int q = tid; q < H0 * W0; q += nt) {
        const int i = i0 - 2 + q % W0, j = j0 - 2 + q / W0;
        if (i >= 0 && i < N && j >= 0 && j < N) SW(i, j) = w[(size_t)j * N + i];
    }
    __syncthreads();
    // w_x (finite_differences.f90:17-41) at (ii, jj)
    auto WX = [&](int ii, int jj) -> double {
        if (ii == 0) return Q1(SW(1, jj) - SW(0, jj), 0, jj);
        if (ii == N - 1) return Q1(SW(N - 1, jj) - SW(N - 2, jj), N - 1, jj);
        return Q2(SW(ii + 1, jj) - SW(ii - 1, jj), ii, jj);
    };
    // w_xx, w_yy, w_xy = Dy(w_x) and the plate moments (lv_elva.f90:95-104)
    const double nu = ph.nu;
    for (int q = tid; q < H2 * W2; q += nt) {
        const int i = i0 - 1 + q % W2, j = j0 - 1 + q / W2;
        if (i < 0 || i >= N || j < 0 || j >= N) continue;
        double wxx, wyy, wxy;
        if (i == 0) wxx = QQ(SW(2, j) - 2 * SW(1, j) + SW(0, j), 0, j);
        else if (i == N - 1) wxx = QQ(SW(N - 1, j) - 2 * SW(N - 2, j) + SW(N - 3, j), 0, j);
        else wxx = QQ(SW(i + 1, j) - 2 * SW(i, j) + SW(i - 1, j), i, j);
        if (j == 0) wyy = QQ(SW(i, 2) - 2 * SW(i, 1) + SW(i, 0), i, 0);
        else if (j == N - 1) wyy = QQ(SW(i, N - 1) - 2 * SW(i, N - 2) + SW(i, N - 3), i, N - 1);
        else wyy = QQ(SW(i, j + 1) - 2 * SW(i, j) + SW(i, j - 1), i, j);
        if (j == 0) wxy = Q1(WX(i, 1) - WX(i, 0), i, 0);
        else if (j == N - 1) wxy = Q1(WX(i, N - 1) - WX(i, N - 2), i, N - 1);
        else wxy = Q2(WX(i, j + 1) - WX(i, j - 1), i, j);
        const double D = (double)ld<T>(M.D, (size_t)j * N + i);
        S2(sMxx, i, j) = -(D * (wxx + nu * wyy));
        S2(sMyy, i, j) = -(D * (wyy + nu * wxx));
        S2(sMxy, i, j) = -(D * (1.0 - nu) * wxy);
    }
    __syncthreads();
    // Mxy_x = Dx(Mxy) at (ii, jj)
    auto MXYX = [&](int ii, int jj) -> double {
        if (ii == 0) return Q1(S2(sMxy, 1, jj) - S2(sMxy, 0, jj), 0, jj);
        if (ii == N - 1) return Q1(S2(sMxy, N - 1, jj) - S2(sMxy, N - 2, jj), N - 1, jj);
        return Q2(S2(sMxy, ii + 1, jj) - S2(sMxy, ii - 1, jj), ii, jj);
    };
    // f = (p + Mx_xx + 2 Mxy_xy + My_yy) / (2 eta)   (lv_elva.f90:107-113)
    const double g = ph.g;
    for (int q = tid; q < ST_TX * ST_TY; q += nt) {
        const int i = i0 + q % ST_TX, j = j0 + q / ST_TX;
        if (i >= N || j >= N) continue;
        double mxx, myy, mxyxy;
        if (i == 0) mxx = QQ(S2(sMxx, 2, j) - 2 * S2(sMxx, 1, j) + S2(sMxx, 0, j), 0, j);
        else if (i == N - 1) mxx = QQ(S2(sMxx, N - 1, j) - 2 * S2(sMxx, N - 2, j) + S2(sMxx, N - 3, j), 0, j);
        else mxx = QQ(S2(sMxx, i + 1, j) - 2 * S2(sMxx, i, j) + S2(sMxx, i - 1, j), i, j);
        if (j == 0) myy = QQ(S2(sMyy, i, 2) - 2 * S2(sMyy, i, 1) + S2(sMyy, i, 0), i, 0);
        else if (j == N - 1) myy = QQ(S2(sMyy, i, N - 1) - 2 * S2(sMyy, i, N - 2) + S2(sMyy, i, N - 3), i, N - 1);
        else myy = QQ(S2(sMyy, i, j + 1) - 2 * S2(sMyy, i, j) + S2(sMyy, i, j - 1), i, j);
        if (j == 0) mxyxy = Q1(MXYX(i, 1) - MXYX(i, 0), i, 0);
        else if (j == N - 1) mxyxy = Q1(MXYX(i, N - 1) - MXYX(i, N - 2), i, N - 1);
        else mxyxy = Q2(MXYX(i, j + 1) - MXYX(i, j - 1), i, j);
        const size_t cidx = (size_t)j * N + i;
        const double p = -(g * (double)ld<T>(M.cfull, cidx));
        const double num = p + mxx + 2.0 * mxyxy + myy;
        T f;
        if constexpr (sizeof(T) == 8) f = (T)(num / (2.0 * (double)ld<T>(M.eta, cidx)));
        else f = (T)(num * (double)ld<T>(M.eta, cidx));   // mixed: eta holds 1/(2 eta)
        ((T*)M.g1)[cidx] = f;
    }
#undef SW
#undef S2
}

// Uniform spacing (K == 1): every one-sided formula of finite_differences.f90
// is the interior formula at a clamped centre (second derivatives: centre
// clamped to [1, N-2]; first derivatives: (u(hi) - u(lo)) / ((hi - lo) dx) with
// lo = max(i-1, 0), hi = min(i+1, N-1)), so the tile runs without divergent
// boundary branches.  Halo cells outside the domain compute finite values
// from clamped loads that no in-domain output reads.
template <typename T>
__global__ void __launch_bounds__(256) k_stencil_u(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.z];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    const PhysDev& ph = *M.phys;
    const int N = M.n;
    const int i0 = blockIdx.x * ST_TX, j0 = blockIdx.y * ST_TY;
    const double* __restrict__ w = c.w_is_x ? M.X[c.cur] : M.W;
    const double inv1 = ph.inv_dx, inv2 = ph.inv_2dx, invq = ph.inv_dx2;
    constexpr int W0 = ST_TX + 4, H0 = ST_TY + 4;  // w:   [i0-2, i0+TX+2) x [j0-2, j0+TY+2)
    constexpr int W2 = ST_TX + 2, H2 = ST_TY + 2;   // M**: [i0-1, i0+TX+1) x [j0-1, j0+TY+1)
    __shared__ double sw[H0 * W0];
    __shared__ double sMxx[H2 * W2], sMyy[H2 * W2], sMxy[H2 * W2];
    const int tid = threadIdx.x, nt = blockDim.x;
    auto cl = [&](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
#define SW(i, j) sw[((j) - (j0 - 2)) * W0 + ((i) - (i0 - 2))]
#define S2(a, i, j) a[((j) - (j0 - 1)) * W2 + ((i) - (i0 - 1))]
    for (int q = tid; q < H0 * W0; q += nt) {
        const int i = cl(i0 - 2 + q % W0, 0, N - 1), j = cl(j0 - 2 + q / W0, 0, N - 1);
        sw[q] = w[(size_t)j * N + i];
    }
    __syncthreads();
    const double nu = ph.nu;
    for (int q = tid; q < H2 * W2; q += nt) {
        const int i = i0 - 1 + q % W2, j = j0 - 1 + q / W2;
        const int jc = cl(j, 0, N - 1), ic = cl(i, 0, N - 1);
        const int ii = cl(i, 1, N - 2), jj = cl(j, 1, N - 2);
        const double wxx = (SW(ii + 1, jc) - 2 * SW(ii, jc) + SW(ii - 1, jc)) * invq;
        const double wyy = (SW(ic, jj + 1) - 2 * SW(ic, jj) + SW(ic, jj - 1)) * invq;
        const int ilo = i - 1 < 0 ? 0 : i - 1, ihi = i + 1 > N - 1 ? N - 1 : i + 1;
        const int jlo = j - 1 < 0 ? 0 : j - 1, jhi = j + 1 > N - 1 ? N - 1 : j + 1;
        const double sx = (ihi - ilo == 2) ? inv2 : inv1, sy = (jhi - jlo == 2) ? inv2 : inv1;
        const double wx_hi = (SW(cl(ihi, i0 - 2, i0 + ST_TX + 1), jhi) - SW(cl(ilo, i0 - 2, i0 + ST_TX + 1), jhi)) * sx;
        const double wx_lo = (SW(cl(ihi, i0 - 2, i0 + ST_TX + 1), jlo) - SW(cl(ilo, i0 - 2, i0 + ST_TX + 1), jlo)) * sx;
        const double wxy = (wx_hi - wx_lo) * sy;
        const double D = (double)ld<T>(M.D, (size_t)jc * N + ic);
        sMxx[q] = -(D * (wxx + nu * wyy));
        sMyy[q] = -(D * (wyy + nu * wxx));
        sMxy[q] = -(D * (1.0 - nu) * wxy);
    }
    __syncthreads();
    const double g = ph.g;
    for (int q = tid; q < ST_TX * ST_TY; q += nt) {
        const int i = i0 + q % ST_TX, j = j0 + q / ST_TX;
        if (i >= N || j >= N) continue;
        const int ii = cl(i, 1, N - 2), jj = cl(j, 1, N - 2);
        const double mxx = (S2(sMxx, ii + 1, j) - 2 * S2(sMxx, ii, j) + S2(sMxx, ii - 1, j)) * invq;
        const double myy = (S2(sMyy, i, jj + 1) - 2 * S2(sMyy, i, jj) + S2(sMyy, i, jj - 1)) * invq;
        const int ilo = i - 1 < 0 ? 0 : i - 1, ihi = i + 1 > N - 1 ? N - 1 : i + 1;
        const int jlo = j - 1 < 0 ? 0 : j - 1, jhi = j + 1 > N - 1 ? N - 1 : j + 1;
        const double sx = (ihi - ilo == 2) ? inv2 : inv1, sy = (jhi - jlo == 2) ? inv2 : inv1;
        const double mxyx_hi = (S2(sMxy, ihi, jhi) - S2(sMxy, ilo, jhi)) * sx;
        const double mxyx_lo = (S2(sMxy, ihi, jlo) - S2(sMxy, ilo, jlo)) * sx;
        const double mxyxy = (mxyx_hi - mxyx_lo) * sy;
        const size_t cidx = (size_t)j * N + i;
        const double p = -(g * (double)ld<T>(M.cfull, cidx));
        const double num = p + mxx + 2.0 * mxyxy + myy;
        T f;
        if constexpr (sizeof(T) == 8) f = (T)(num / (2.0 * (double)ld<T>(M.eta, cidx)));
        else f = (T)(num * (double)ld<T>(M.eta, cidx));   // mixed: eta holds 1/(2 eta)
        ((T*)M.g1)[cidx] = f;
    }
#undef SW
#undef S2
}

// ------------------------------------------------------------------ row pass
// Block: 2L rows [r0, r0 + 2L); neighbouring rows r0+2l (real part) and
// r0+2l+1 (imaginary part) share complex line l, so a row's rounding does not
// depend on the block size.
template <typename T>
__device__ __forceinline__ void load_rows(cplx<T>* buf, const T* __restrict__ src, int N, int L, int r0) {
    T* bf = (T*)buf;
    staged_copy<8>(2 * L * N, [&](int q) { return src[(size_t)r0 * N + q]; },
                   [&](int q, T v) {
                       const int mr = q / N, i = q % N;
                       bf[2 * (i * L + (mr >> 1)) + (mr & 1)] = v;
                   });
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(512) k_row_fwd(const Member* __restrict__ mem, FftPlan plan,
                                                 const void* __restrict__ twv, int L) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    const int N = M.n, r0 = blockIdx.x * 2 * L;
    const cplx<T>* tw = stage_twiddles<T>(buf, N, L, twv);
    load_rows<T>(buf, (const T*)M.g1, N, L, r0);
    lds_fft<T>(buf, plan, L, L, tw);
    T* out = (T*)M.g2;
    for (int q = threadIdx.x; q < 2 * L * N; q += blockDim.x) {
        const int mr = q / N, k = q % N;
        const int l = mr >> 1, part = mr & 1;
        T ha, hb;
        dht_pair(buf[k * L + l], buf[((N - k) % N) * L + l], ha, hb);
        out[(size_t)(r0 + mr) * N + k] = part ? hb : ha;
    }
}

// ------------------------------------------------------------------ column pass
// Block: real columns [c0, c0 + 2C); neighbouring columns c0+2l and c0+2l+1
// share line l (a column's rounding does not depend on the block size).
template <typename T>
__global__ void __launch_bounds__(512) k_col(const Member* __restrict__ mem, FftPlan plan,
                                             const void* __restrict__ twv, const double2* __restrict__ tw64,
                                             int C) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN || c.k_valid) return;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    const int N = M.n, c0 = blockIdx.x * 2 * C;
    const int tid = threadIdx.x, nt = blockDim.x;
    const cplx<T>* tw = stage_twiddles<T>(buf, N, C, twv);
    const T* __restrict__ src = (const T*)M.g2;
    const T* __restrict__ kR = (const T*)M.kR;
    T* bf = (T*)buf;
    staged_copy<8>(2 * C * N, [&](int q) { return src[(size_t)(q / (2 * C)) * N + c0 + q % (2 * C)]; },
                   [&](int q, T v) {
                       const int k = q / (2 * C), mc = q % (2 * C);
                       bf[2 * (k * C + (mc >> 1)) + (mc & 1)] = v;
                   });
    __syncthreads();
    lds_fft<T>(buf, plan, C, C, tw);
    // Hartley values, spectral divide, re-pack (lv_elva.f90:115)
    for (int q = tid; q < (N / 2 + 1) * C; q += nt) {
        const int l = q % C, k = q / C, km = (N - k) % N;
        const cplx<T> zk = buf[k * C + l], zm = buf[km * C + l];
        T ha, hb, ham, hbm;
        dht_pair(zk, zm, ha, hb);
        dht_pair(zm, zk, ham, hbm);
        ha = kdiv(ha, kR[(size_t)k * N + c0 + 2 * l]);
        hb = kdiv(hb, kR[(size_t)k * N + c0 + 2 * l + 1]);
        buf[k * C + l] = mk<T>(ha, hb);
        if (km != k) {
            ham = kdiv(ham, kR[(size_t)km * N + c0 + 2 * l]);
            hbm = kdiv(hbm, kR[(size_t)km * N + c0 + 2 * l + 1]);
            buf[km * C + l] = mk<T>(ham, hbm);
        }
    }
    __syncthreads();
    lds_fft<T>(buf, plan, C, C, tw);
    T* dst = (T*)M.g1;
    for (int q = tid; q < 2 * C * N; q += nt) {
        const int k = q / (2 * C), mc = q % (2 * C);
        const int l = mc >> 1;
        T ha, hb;
        dht_pair(buf[k * C + l], buf[((N - k) % N) * C + l], ha, hb);
        const T val = (mc & 1) ? hb : ha;
        dst[(size_t)k * N + c0 + mc] = val;
        // per-column terms of the four corners of the inverse transform (rows 0
        // and N-1, weights 1 and cas(2 pi col (N-1) / N)); the row pass sums
        // them in an order set by N alone, whatever the blocking
        if (k == 0 || k == N - 1) {
            const int col = c0 + mc, qc = (k == 0) ? 0 : 2;
            const double v = (double)val;
            M.corners[(size_t)col * 4 + qc] = v;
            M.corners[(size_t)col * 4 + qc + 1] = v * (tw64[col].x + tw64[col].y);
        }
    }
}

// The epilogue runs one row per wave (rows wave, wave + waves, ...), each
// lane taking columns lane, lane + 64, ...; every row leaves one error partial
// (max(max|x1|, max|x2|), sum (x2-x1)^2) reduced by a fixed butterfly, so the
// error norm of a domain depends on N alone, not on the blocking.
template <typename T>
__global__ void __launch_bounds__(512) k_row_inv(const Member* __restrict__ mem, FftPlan plan,
                                                 const void* __restrict__ twv, int L) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN) return;
    const PhysDev& ph = *M.phys;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = (cplx<T>*)smem;
    __shared__ double corner_c;
    const int N = M.n, r0 = blockIdx.x * 2 * L;
    const int nw = blockDim.x >> 6, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double dt = attempt_dt(c, ph);
    const double* __restrict__ X = M.X[c.cur];
    double* __restrict__ Xn = M.X[c.cur ^ 1];
    auto update = [&](size_t cidx, double x, double k, double& a1, double& a2, double& s2) {
        double x1;
        const double x2 = stage_sums(ph.dt_method, x, k, dt, x1);
        Xn[cidx] = x2;
        a1 = fmax(a1, fabs(x1)); a2 = fmax(a2, fabs(x2));
        const double d = x2 - x1;
        s2 += d * d;
    };
    auto row_done = [&](int mr, double a1, double a2, double s2) {
        a1 = wave_max(a1); a2 = wave_max(a2); s2 = wave_sum(s2);
        if (lane == 0) {
            M.partials[(size_t)(r0 + mr) * 4 + 0] = fmax(a1, a2);
            M.partials[(size_t)(r0 + mr) * 4 + 1] = s2;
        }
    };
    // columns of a row in chunks of EU per lane, the chunk's loads first (the
    // stores of the previous chunk would otherwise enter every wait)
    constexpr int EU = 4;
    if (c.k_valid) {
        // rejected attempt before: the RHS is unchanged, only dt differs
        for (int mr = wv; mr < 2 * L; mr += nw) {
            const size_t rb = (size_t)(r0 + mr) * N;
            double a1 = 0.0, a2 = 0.0, s2 = 0.0;
            for (int i0 = lane; i0 < N; i0 += 64 * EU) {
                double xv[EU], kv[EU];
#pragma unroll
                for (int u = 0; u < EU; ++u) {
                    const int i = i0 + 64 * u < N ? i0 + 64 * u : i0;
                    xv[u] = X[rb + i];
                    kv[u] = M.dwdt[rb + i];
                }
#pragma unroll
                for (int u = 0; u < EU; ++u)
                    if (i0 + 64 * u < N) update(rb + i0 + 64 * u, xv[u], kv[u], a1, a2, s2);
            }
            row_done(mr, a1, a2, s2);
        }
    } else {
        const double S = (double)(float)((double)N * (double)N);
        if (wv == 0) {
            // the four corner sums over the N per-column terms of the column
            // pass: lane-strided, then a fixed butterfly (order set by N alone)
            double v[4] = {0, 0, 0, 0};
            for (int col = lane; col < N; col += 64)
                for (int q = 0; q < 4; ++q) v[q] += M.corners[(size_t)col * 4 + q];
            for (int q = 0; q < 4; ++q) v[q] = wave_sum(v[q]);
            // x(1,1) + x(nx,ny) + x(1,ny) + x(nx,1)  (isos_utils.f90:116)
            if (lane == 0) corner_c = 0.25 * (v[0] / S + v[3] / S + v[2] / S + v[1] / S);
        }
        const cplx<T>* tw = stage_twiddles<T>(buf, N, L, twv);
        load_rows<T>(buf, (const T*)M.g1, N, L, r0);   // also synchronises corner_c and tw
        lds_fft<T>(buf, plan, L, L, tw);
        const double cc = corner_c, spy = ph.sec_per_year;
        const bool store_k = !(c.t + dt < c.tf);
        for (int mr = wv; mr < 2 * L; mr += nw) {
            const int l = mr >> 1, part = mr & 1;
            const size_t rb = (size_t)(r0 + mr) * N;
            double a1 = 0.0, a2 = 0.0, s2 = 0.0;
            for (int i0 = lane; i0 < N; i0 += 64 * EU) {
                double xv[EU];
#pragma unroll
                for (int u = 0; u < EU; ++u) xv[u] = X[rb + (i0 + 64 * u < N ? i0 + 64 * u : i0)];
#pragma unroll
                for (int u = 0; u < EU; ++u) {
                    const int i = i0 + 64 * u;
                    if (i >= N) continue;
                    T ha, hb;
                    dht_pair(buf[i * L + l], buf[((N - i) % N) * L + l], ha, hb);
                    const double raw = (double)(part ? hb : ha);
                    const double k = (raw / S - cc) * spy;
                    if (store_k) M.dwdt[rb + i] = k;
                    update(rb + i, xv[u], k, a1, a2, s2);
                }
            }
            row_done(mr, a1, a2, s2);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) M.ctl->pending = 1;
}

// Attempt of methods 0-2: RHS dwdt = (w_equilibrium - now%w) / tau
// (calc_asthenosphere_relax, lv_xlra.f90:17-26; 0 for method 0), stage sums
// and error partials.  One block per pair of rows, so the partials have the
// layout of the power-of-two LV-ELVA path (one entry per row pair).
constexpr int XLRA_THREADS = 256;
__global__ void __launch_bounds__(XLRA_THREADS) k_xlra(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.y];
    const Ctl& c = *M.ctl;
    if (c.status != ST_RUN) return;
    const PhysDev& ph = *M.phys;
    const int N = M.n;
    const double dt = attempt_dt(c, ph);
    const double* __restrict__ X = M.X[c.cur];
    double* __restrict__ Xn = M.X[c.cur ^ 1];
    const double* __restrict__ w = c.w_is_x ? X : M.W;
    const double* __restrict__ weq = M.weq;
    const bool relax = ph.method != 0, store_k = !(c.t + dt < c.tf);
    const double tau = ph.tau;
    const int dtm = ph.dt_method;
    const int r0 = blockIdx.x * 2, rows = min(2, N - r0);
    double a1 = 0.0, a2 = 0.0, s2 = 0.0;
    for (int q = threadIdx.x; q < rows * N; q += XLRA_THREADS) {
        const size_t cidx = (size_t)r0 * N + q;
        const double k = relax ? (weq[cidx] - w[cidx]) / tau : 0.0;
        if (store_k) M.dwdt[cidx] = k;
        double x1;
        const double x2 = stage_sums(dtm, X[cidx], k, dt, x1);
        Xn[cidx] = x2;
        a1 = fmax(a1, fabs(x1)); a2 = fmax(a2, fabs(x2));
        const double d = x2 - x1;
        s2 += d * d;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) M.ctl->pending = 1;
    block_reduce3(a1, a2, s2, M.partials + (size_t)blockIdx.x * 4);
}

__global__ void __launch_bounds__(64) k_decide(const Member* __restrict__ mem, int nblk) {
    decide_wave(mem[blockIdx.x], nblk);
}

// ------------------------------------------------------------------ launchers
template <typename T>
static size_t lds_bytes(int n, int lines) { return (size_t)n * (lines + 1) * sizeof(cplx<T>); }   // + twiddles

// Optional HIP-event timing of the attempt kernels (for benchmarks): events
// ride on every launch of the selected kernel (or of every attempt kernel,
// PROF_ALL) and are summed per kernel on demand.
namespace {
struct Prof {
    int kernel = -1;              // -1 off, 0-3 one kernel, PROF_ALL every attempt kernel
    std::vector<hipEvent_t> ev;   // start/stop pairs
    std::vector<int> which;       // kernel of each pair
    size_t used = 0;
    std::atomic<bool> fused{false};   // last attempt ran the fused stencil + row kernel
    RhsShape last{};              // launch shape of the last attempt
};
Prof g_prof;

}  // namespace

bool prof_pair(int kernel, hipEvent_t* start, hipEvent_t* stop) {
    if (g_prof.kernel < 0 || (g_prof.kernel != kernel && g_prof.kernel != PROF_ALL)) return false;
    while (g_prof.used + 2 > g_prof.ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return false;
        g_prof.ev.push_back(e);
    }
    g_prof.which.resize(g_prof.ev.size() / 2);
    g_prof.which[g_prof.used / 2] = kernel;
    *start = g_prof.ev[g_prof.used++];
    *stop = g_prof.ev[g_prof.used++];
    return true;
}

void prof_select(int kernel) {
    g_prof.kernel = kernel;
    g_prof.used = 0;
}

int prof_fused() { return g_prof.fused.load(std::memory_order_relaxed) ? 1 : 0; }
void prof_shape(int* out) {
    const RhsShape& s = g_prof.last;
    const int v[11] = {s.n, s.fp64 ? 1 : 0, prof_fused(), s.fast ? 1 : 0, s.threads_row, s.C,
                       s.row_tiles, s.srow_rows, s.threads_inv, s.dt_method, s.fast ? col_buffers(s) : 1};
    for (int i = 0; i < 11; ++i) out[i] = v[i];
}
bool prof_active() { return g_prof.kernel >= 0; }

int prof_collect_kernel(int kernel, double* total_ms, long long* launches) {
    double tot = 0.0;
    long long n = 0;
    for (size_t i = 0; i + 1 < g_prof.used; i += 2) {
        if (kernel >= 0 && g_prof.which[i / 2] != kernel) continue;
        float ms = 0.f;
        if (hipEventSynchronize(g_prof.ev[i + 1]) != hipSuccess) return -1;
        if (hipEventElapsedTime(&ms, g_prof.ev[i], g_prof.ev[i + 1]) != hipSuccess) return -1;
        tot += ms;
        ++n;
    }
    *total_ms = tot;
    *launches = n;
    return 0;
}

int prof_collect(double* total_ms, long long* launches) { return prof_collect_kernel(-1, total_ms, launches); }

// methods 0-2: one pointwise kernel, decided at once
void launch_attempt_xlra(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st) {
    const int nblk = (s.n + 1) / 2;
    launch_k(3, k_xlra, dim3(nblk, nmem), dim3(XLRA_THREADS), 0, st, d_mem);
    hipLaunchKernelGGL(k_decide, dim3(nmem), dim3(64), 0, st, d_mem, nblk);
}

template <typename T>
void launch_attempt_t(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st) {
    dim3 gs((s.n + ST_TX - 1) / ST_TX, (s.n + ST_TY - 1) / ST_TY, nmem);
    // uniform spacing, power-of-two N: stencil and forward row pass in one kernel
    const bool fused = s.fused && s.fast && s.uniform && launch_fast_pass(d_mem, nmem, s, st, 0);
    g_prof.fused.store(fused, std::memory_order_relaxed);
    if (g_prof.kernel >= 0) g_prof.last = s;   // profiling runs (one host thread): the timed launches' shape
    if (!fused) {
        if (s.uniform) launch_k(0, k_stencil_u<T>, gs, dim3(256), 0, st, d_mem);
        else launch_k(0, k_stencil<T, false>, gs, dim3(256), 0, st, d_mem);
    }
    if (!fused && (!s.fast || !launch_fast_pass(d_mem, nmem, s, st, 1)))
        launch_k(1, k_row_fwd<T>, dim3(s.n / (2 * s.L), nmem), dim3(s.threads_row), lds_bytes<T>(s.n, s.L), st,
                 d_mem, s.plan, (const void*)s.tw, s.L);
    // after the fused pass the column pass commits its control block
    if (!s.fast || !launch_fast_pass(d_mem, nmem, s, st, fused ? 4 : 2))
        launch_k(2, k_col<T>, dim3(s.n / (2 * s.C), nmem), dim3(s.threads_col), lds_bytes<T>(s.n, s.C), st,
                 d_mem, s.plan, (const void*)s.tw, s.tw64, s.C);
    // partials of the error norm: one per line pair of rows (fast pass), one per row (generic)
    int nblk = s.n / 2;
    if (!s.fast || !launch_fast_pass(d_mem, nmem, s, st, 3)) {
        launch_k(3, k_row_inv<T>, dim3(s.n / (2 * s.L), nmem), dim3(s.threads_row), lds_bytes<T>(s.n, s.L), st,
                 d_mem, s.plan, (const void*)s.tw, s.L);
        nblk = s.n;
    }
    // the fused pass of the next attempt applies this attempt's decision; the
    // last one of a launch round is applied by launch_decide
    if (!fused) hipLaunchKernelGGL(k_decide, dim3(nmem), dim3(64), 0, st, d_mem, nblk);
}

void launch_decide(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st) {
    // partials: one per line pair of rows on the power-of-two path, one per row otherwise
    const int nblk = s.fast ? s.n / 2 : s.n;
    hipLaunchKernelGGL(k_decide, dim3(nmem), dim3(64), 0, st, d_mem, nblk);
}

void launch_attempt(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st) {
    if (s.method != 3) launch_attempt_xlra(d_mem, nmem, s, st);
    else if (s.fp64) launch_attempt_t<double>(d_mem, nmem, s, st);
    else launch_attempt_t<float>(d_mem, nmem, s, st);
}

}  // namespace fi
