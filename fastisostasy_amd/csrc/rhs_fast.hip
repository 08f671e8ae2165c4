// rhs_fast.hip -- attempt kernels for power-of-two N (64 ... 1024) built on the
// register Stockham FFT of fft_reg.hpp; same contract as the generic kernels of
// rhs.hip (identical inputs, outputs, partial layouts and Ctl gating).
//
//   kf_srow_fwd  stencil fused with the row Hartley transform (x -> g2)
//   kf_row_fwd   rows of f -> row Hartley transform        (g1 -> g2)
//   kf_col       column Hartley transform, divide by kappa*R, column Hartley
//                transform, per-column corner terms         (g2 -> g1)
//   kf_row_inv   row Hartley transform, /N^2, corner BC, * sec_per_year,
//                stage sums, reduction partials             (g1 -> dwdt, x')
//
// Rows: L complex lines (2L real rows: r0+2l in the real part, r0+2l+1 in the
// imaginary part) per block, TPL = N/8 threads per line.  Columns: C complex
// lines (columns c0+2l and c0+2l+1, read and written as one vector), lane ->
// (line = tid % C, t = tid / C) so that consecutive lanes read consecutive
// column pairs.  Packing neighbours (not r and r+L) makes every line's
// rounding independent of the block shape, hence of the batch size.
#include <cstdlib>
#include <type_traits>

#include "fft_reg.hpp"
#include "fi_kernels.h"
#include "rhs_common.hpp"

namespace fi {

template <typename T, int N>
__device__ __forceinline__ void load_tw(cplx<T>* tw_l, const cplx<T>* __restrict__ twg) {
    gptr<T> g = gin<T>(twg);
    T* l = (T*)tw_l;
    for (int q = threadIdx.x; q < 2 * N; q += blockDim.x) l[q] = g[q];
}

template <typename T> struct vec2;
template <> struct vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T, int N, int THREADS>
__global__ void __launch_bounds__(THREADS) kf_row_fwd(const Member* __restrict__ mem, const void* __restrict__ twv) {
    const Member& M = mem[blockIdx.y];
    const kptr<Ctl> c = kview(M.ctl);
    if ((c->status != ST_RUN) | (c->k_valid != 0)) return;
    using PL = RegPlan<N>;
    constexpr int TPL = PL::TPL, E = PL::E, L = THREADS / TPL;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    load_tw<T, N>(tw_l, (const cplx<T>*)twv);
    const int ra = blockIdx.x * 2 * L + 2 * line, rb = ra + 1;
    gptr<T> src = gin<T>(M.g1);
    cplx<T> v[E], m[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        v[e] = mk<T>(src[(size_t)ra * N + t + TPL * e], src[(size_t)rb * N + t + TPL * e]);
    __syncthreads();
    RowXch<T, N> xc(buf, line);
    reg_fft_mirror<T, N, true>(v, m, t, tw_l, xc);
    T* __restrict__ dst = (T*)M.g2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        dst[(size_t)ra * N + t + TPL * e] = ha;
        dst[(size_t)rb * N + t + TPL * e] = hb;
    }
}

// ------------------------------------------------------------------ fused stencil + row pass
// Flexure stencil for uniform spacing (lv_elva.f90:60-108 with the one-sided
// differences of isos_utils.f90 at the edges) fused with the forward row
// transform, so f never leaves the CU.  The block owns 2L rows [R0, R0+2L);
// thread `tid` owns columns tid + THREADS q.  Global inputs are loaded
// SR_LA march steps ahead of their use (w rows R0-2 .. R0+2L+1, D rows
// R0-1 .. R0+2L, canom_full and eta rows R0 .. R0+2L-1), and the block
// marches down the rows: step m (local row, -1 .. 2L+1) forms the moments of
// row m and f of row m - 2, with one barrier per step.  x-neighbours go
// through LDS rings (w 3 rows, Mxx 3, Mxy 2); each row's x-derivatives of w
// and Mxy are formed once and kept in registers; the y-direction (w and Myy
// of the owned column) stays in registers.  Edge rows pick the shifted
// windows the clamped centre rows of the reference select.
//
// Block order is XCD-aware: consecutive workgroup ids go to different XCDs,
// so the row blocks of one member are mapped to ids that share an XCD and
// its L2 holds the halo rows of the neighbouring blocks.
//
// LF: complex lines (2 LF rows) per block; by default every thread takes part
// in a line (LF = THREADS / TPL).  A smaller LF gives more, shorter row blocks
// for small batches, where the march of one block sets the launch's latency;
// the threads beyond LF lines then transform a copy of line 0, discarded.
template <typename T, int N, int THREADS, int LF = THREADS / (N / 8)> struct SRow {
    static constexpr int TPL = N / 8, L = LF, ROWS = 2 * L, LX = THREADS / TPL;
    static_assert(LF >= 1 && LF <= LX, "lines per block");
    static constexpr int CPT = N / THREADS;
    static constexpr size_t ring_bytes = (size_t)(3 + 3 + 2) * N * sizeof(double);
    static constexpr size_t stencil_bytes = ring_bytes + (size_t)ROWS * N * sizeof(T);
    static constexpr size_t xch_bytes = (size_t)LX * RowXch<T, N>::PITCH * sizeof(cplx<T>);
    static constexpr size_t lds = (size_t)N * sizeof(cplx<T>) + (stencil_bytes > xch_bytes ? stencil_bytes : xch_bytes);
};

// rows of canom_full / eta the march loads ahead of their use
// (4: measured 23.7 us per launch on the 64-member ensemble against 24.9 at 6
// and 25.7 at 8; nearer loads hold fewer registers)
constexpr int SR_LA = 4;

template <int B, int E, class F> __device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// (member, block-within-member) for a grid (blocks per member, members)
__device__ __forceinline__ void xcd_order(int& member, int& blk) {
    const int per = gridDim.x, total = gridDim.x * gridDim.y;
    int b = blockIdx.x + per * blockIdx.y;
    if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
    member = b / per;
    blk = b % per;
}

// The march and the row transform of one block.  issue(tau) issues the
// global loads due at march step tau (inputs arrive SR_LA steps ahead of use).
// Rows are counted in march order: first_edge / last_edge say that the march
// starts / ends at a domain edge; UP marches run against the row order.
template <typename T, int N, int THREADS, int LF, bool UP, class Issue>
__device__ __forceinline__ void srow_body(const Member& M, const kptr<PhysDev> ph, int R0, int tid, bool first_edge,
                                          bool last_edge, unsigned char* smem,
                                          double (&wv)[SRow<T, N, THREADS, LF>::ROWS + 4][SRow<T, N, THREADS, LF>::CPT],
                                          T (&dv)[SRow<T, N, THREADS, LF>::ROWS + 2][SRow<T, N, THREADS, LF>::CPT],
                                          T (&cv)[SRow<T, N, THREADS, LF>::ROWS][SRow<T, N, THREADS, LF>::CPT],
                                          T (&ev)[SRow<T, N, THREADS, LF>::ROWS][SRow<T, N, THREADS, LF>::CPT],
                                          Issue&& issue) {
    using SR = SRow<T, N, THREADS, LF>;
    constexpr int TPL = SR::TPL, E = RegPlan<N>::E, L = SR::L, ROWS = SR::ROWS, CPT = SR::CPT;
    cplx<T>* tw_l = (cplx<T>*)smem;
    double* rw = (double*)(tw_l + N);     // w    rows, slot (r + 3) % 3
    double* rxx = rw + 3 * N;             // Mxx  rows, slot (r + 3) % 3
    double* rxy = rxx + 3 * N;            // Mxy  rows, slot (r + 2) % 2
    T* sf = (T*)(rxy + 2 * N);            // f    [ROWS][N]
    const double invq = ph->inv_dx2, inv1 = ph->inv_dx, inv2 = ph->inv_2dx, nu = ph->nu, g = ph->g;
    constexpr auto s3 = [](int r) { return (r + 3) % 3; };
    constexpr auto s2 = [](int r) { return (r + 2) % 2; };
    // per owned column: clamped centre, x-neighbours and x-spacing factor
    int icol[CPT], ic[CPT], ilo[CPT], ihi[CPT];
    double sx[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = tid + THREADS * q;
        icol[q] = i;
        ic[q] = i < 1 ? 1 : (i > N - 2 ? N - 2 : i);
        ilo[q] = i - 1 < 0 ? 0 : i - 1;
        ihi[q] = i + 1 > N - 1 ? N - 1 : i + 1;
        sx[q] = (ihi[q] - ilo[q] == 2) ? inv2 : inv1;
    }
    // centred / one-sided x-derivative of a ring row at the owned columns
    auto xder = [&](const double* row, int q) { return (row[ihi[q]] - row[ilo[q]]) * sx[q]; };
    // x-derivatives of w (rows -2 .. ROWS+1) and of Mxy (rows -1 .. ROWS):
    // each row's is formed once and kept while the three-row window needs it
    double wx[ROWS + 4][CPT], mx[ROWS + 2][CPT];
    double my[ROWS + 2][CPT];   // Myy of local rows -1 .. ROWS
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        rw[s3(-2) * N + icol[q]] = wv[0][q];
        rw[s3(-1) * N + icol[q]] = wv[1][q];
        rw[s3(0) * N + icol[q]] = wv[2][q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        wx[0][q] = xder(rw + s3(-2) * N, q);
        wx[1][q] = xder(rw + s3(-1) * N, q);
    }
    // March step m (local row, -1 .. ROWS+1), one barrier per step: the
    // moments of row m (w rows m, m+1 in the ring), the x-derivative of Mxy
    // row m-1 and f of row m-2 (Mxx row m-2 in the ring); the step ends by
    // writing w row m+2 into the slot of row m-1, which no thread reads in
    // this step.  Row windows at the domain edges follow the clamped
    // centres of the reference (jlo = max(g-1, 0), jhi = min(g+1, N-1)).
    static_for<-1, ROWS + 2>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        issue(mc);
        __syncthreads();
        if constexpr (m <= ROWS) {
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int i = icol[q], ii = ic[q];
                wx[m + 3][q] = xder(rw + s3(m + 1) * N, q);
                const double* wr = rw + s3(m) * N;
                const double wxx = (wr[ii + 1] - 2 * wr[ii] + wr[ii - 1]) * invq;
                // centre row of the y-difference: clamp(gm, 1, N-2)
                double wa = wv[m + 3][q], wb = wv[m + 2][q], wc = wv[m + 1][q];
                double whi = wx[m + 3][q], wlo = wx[m + 1][q], sy = inv2;
                if constexpr (m <= 0) {
                    if (first_edge) {
                        wa = wv[2 + 2][q]; wb = wv[1 + 2][q]; wc = wv[0 + 2][q];
                        if constexpr (m == -1) { whi = wx[2][q]; wlo = wx[2][q]; }
                        else { wlo = wx[m + 2][q]; }
                        sy = inv1;
                    }
                }
                if constexpr (m >= ROWS - 1) {
                    if (last_edge) {
                        wa = wv[ROWS - 1 + 2][q]; wb = wv[ROWS - 2 + 2][q]; wc = wv[ROWS - 3 + 2][q];
                        if constexpr (m == ROWS - 1) { whi = wx[m + 2][q]; }
                        else { whi = wx[m + 1][q]; wlo = wx[m + 1][q]; }
                        sy = inv1;
                    }
                }
                // (an upward march meets row g + 1 before row g - 1: the
                // operands keep the reference's order, w(g+1) first)
                if constexpr (UP) {
                    const double t1 = wa; wa = wc; wc = t1;
                    const double t2 = whi; whi = wlo; wlo = t2;
                }
                const double wyy = (wa - 2 * wb + wc) * invq;
                const double wxy = (whi - wlo) * sy;
                const double D = (double)dv[m + 1][q];
                rxx[s3(m) * N + i] = -(D * (wxx + nu * wyy));
                my[m + 1][q] = -(D * (wyy + nu * wxx));
                rxy[s2(m) * N + i] = -(D * (1.0 - nu) * wxy);
            }
        }
        if constexpr (m >= 0) {
            // x-derivative of Mxy row m-1 (written in the previous step)
#pragma unroll
            for (int q = 0; q < CPT; ++q) mx[m][q] = xder(rxy + s2(m - 1) * N, q);
        }
        if constexpr (m - 2 >= 0) {
            // f of local row j = m - 2
            constexpr int j = m - 2;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int i = icol[q], ii = ic[q];
                const double* xr = rxx + s3(j) * N;
                const double mxx = (xr[ii + 1] - 2 * xr[ii] + xr[ii - 1]) * invq;
                double ya = my[j + 2][q], yb = my[j + 1][q], yc = my[j][q];
                double hi = mx[j + 2][q], lo = mx[j][q], sy = inv2;
                if constexpr (j == 0) {
                    if (first_edge) { ya = my[2 + 1][q]; yb = my[1 + 1][q]; yc = my[0 + 1][q]; lo = mx[j + 1][q]; sy = inv1; }
                }
                if constexpr (j == ROWS - 1) {
                    if (last_edge) {
                        ya = my[ROWS - 1 + 1][q]; yb = my[ROWS - 2 + 1][q]; yc = my[ROWS - 3 + 1][q];
                        hi = mx[j + 1][q]; sy = inv1;
                    }
                }
                if constexpr (UP) {
                    const double t1 = ya; ya = yc; yc = t1;
                    const double t2 = hi; hi = lo; lo = t2;
                }
                const double myy = (ya - 2 * yb + yc) * invq;
                const double mxyxy = (hi - lo) * sy;
                const double p = -(g * (double)cv[j][q]);
                const double num = p + mxx + 2.0 * mxyxy + myy;
                T f;
                if constexpr (sizeof(T) == 8) f = (T)(num / (2.0 * (double)ev[j][q]));
                else f = (T)(num * (double)ev[j][q]);   // mixed: eta holds 1/(2 eta)
                sf[(UP ? ROWS - 1 - j : j) * N + i] = f;
            }
        }
        if constexpr (m + 2 <= ROWS + 1) {
#pragma unroll
            for (int q = 0; q < CPT; ++q) rw[s3(m + 2) * N + icol[q]] = wv[m + 4][q];
        }
    });
    __syncthreads();
    const int line = tid / TPL, t = tid % TPL;
    const int ld = (L == SR::LX || line < L) ? line : 0;   // lines beyond L: a discarded copy of line 0
    cplx<T> v[E], mr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk<T>(sf[2 * ld * N + t + TPL * e], sf[(2 * ld + 1) * N + t + TPL * e]);
    __syncthreads();
    RowXch<T, N> xc((cplx<T>*)(tw_l + N), line);
    reg_fft_mirror<T, N, true>(v, mr, t, tw_l, xc);
    if (L < SR::LX && line >= L) return;
    gmut<T> dst = gout<T>(M.g2);
    const int ra = R0 + 2 * line, rb = ra + 1;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        T ha, hb;
        dht_pair(v[e], mr[e], ha, hb);
        dst[(size_t)ra * N + t + TPL * e] = ha;
        dst[(size_t)rb * N + t + TPL * e] = hb;
    }
}

// One row block of the fused pass.  UP: the block marches from its last row
// to its first.  Neighbouring blocks march in opposite directions, so the two
// blocks that share halo rows load them at the same point of their marches
// (both at the start or both at the end) and the second load finds them in
// L2; marching the same way, one block would load them first and the other
// a dozen steps later, after L2 had evicted them.
template <typename T, int N, int THREADS, int LF, bool UP>
__device__ __forceinline__ void srow_block(const Member* __restrict__ mem, const void* __restrict__ twv, int mi,
                                           int rblk) {
    using SR = SRow<T, N, THREADS, LF>;
    constexpr int ROWS = SR::ROWS, CPT = SR::CPT;
    const Member& M = mem[mi];
    // the member's pointers in one batch of scalar loads
    const Ctl* const p_ctl = M.ctl;
    const double *x0 = M.X[0], *x1 = M.X[1], *wW = M.W;
    const void *p_D = M.D, *p_C = M.cfull, *p_E = M.eta, *p_part = M.partials;
    const kptr<PhysDev> ph = kview(M.phys);
    const int tid = threadIdx.x;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    const int R0 = rblk * ROWS;
    const bool top = R0 == 0, bottom = R0 + ROWS == N;
    gptr<T> Dg = gin<T>(p_D), Cg = gin<T>(p_C), Eg = gin<T>(p_E);
    auto crow = [](int r) { return r < 0 ? 0 : (r > N - 1 ? N - 1 : r); };
    // Loads that do not depend on the domain's control block go first, so the
    // two dependent scalar loads of its test are not in front of them: the
    // twiddles and the previous attempt's partials (the LDS store and the
    // decision then wait only for their own loads), then the rheology rows
    // of the march prologue.  The buffers exist whatever the domain's state;
    // an idle domain's loads are dropped.
    constexpr int TWQ = (2 * N + THREADS - 1) / THREADS;
    T twr[TWQ];
    {
        // (no branch around the load when the threads cover the table evenly:
        // the join would wait for the load before any row load is issued)
        gptr<T> g = gin<T>(twv);
#pragma unroll
        for (int k = 0; k < TWQ; ++k) {
            if constexpr (THREADS * TWQ == 2 * N) twr[k] = g[tid + THREADS * k];
            else if (tid + THREADS * k < 2 * N) twr[k] = g[tid + THREADS * k];
        }
    }
    scalars_ready(p_ctl, x0, x1, wW, p_D, p_C, p_E, p_part);
    // (loaded whether or not a decision is pending: no branch for the
    // compiler to reorder the decision around)
    constexpr int NP = N / 2, PJ = (NP + 63) / 64;   // one partial per line
    double pa[PJ][2];
    {
        gptr<double> part = gin<double>(p_part);
        const int lane = tid & 63;
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const bool in = lane + 64 * j < NP;
#pragma unroll
            for (int k = 0; k < 2; ++k) pa[j][k] = in ? part[(size_t)(lane + 64 * j) * 4 + k] : 0.0;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    double wv[ROWS + 4][CPT];   // local rows -2 .. ROWS+1
    T dv[ROWS + 2][CPT];        // -1 .. ROWS
    T cv[ROWS][CPT], ev[ROWS][CPT];
    // Software-pipelined loads: at march step tau the block issues w row
    // tau + 2 + SR_LA, D row tau + SR_LA and canom_full / eta row
    // tau - 2 + SR_LA, each SR_LA steps before its first use.  Row bases are
    // uniform (scalar); the lane offset is one 32-bit value for every row.
    gptr<double> w = nullptr;   // the state buffer, known once the control block is read
    // global row of march row r (-2 .. ROWS+1)
    auto grow = [&](int r) { return UP ? R0 + ROWS - 1 - r : R0 + r; };
    auto issue_fields = [&](auto tc) {
        constexpr int tau = decltype(tc)::value;
        constexpr int rd = tau + SR_LA, rc = tau - 2 + SR_LA;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const unsigned col = tid + THREADS * q;
            if constexpr (rd >= -1 && rd <= ROWS) dv[rd + 1][q] = (Dg + (size_t)crow(grow(rd)) * N)[col];
            if constexpr (rc >= 0 && rc < ROWS) {
                cv[rc][q] = (Cg + (size_t)grow(rc) * N)[col];
                ev[rc][q] = (Eg + (size_t)grow(rc) * N)[col];
            }
        }
    };
    auto issue_w = [&](auto tc) {
        constexpr int tau = decltype(tc)::value;
        constexpr int rw_ = tau + 2 + SR_LA;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const unsigned col = tid + THREADS * q;
            if constexpr (rw_ >= -2 && rw_ <= ROWS + 1) wv[rw_ + 2][q] = (w + (size_t)crow(grow(rw_)) * N)[col];
        }
    };
    auto issue = [&](auto tc) {
        issue_w(tc);
        issue_fields(tc);
        // keep the issue order: each step waits only on loads SR_LA steps old
        __builtin_amdgcn_sched_barrier(0);
    };
    // the rheology rows due before the first march step
    static_for<-4 - SR_LA, -1>([&](auto tc) {
        issue_fields(tc);
        __builtin_amdgcn_sched_barrier(0);
    });
    // Apply the previous attempt's pending accept/reject here instead of in a
    // separate kernel: every wave recomputes it from the same partials (same
    // order, same result); block 0 writes the resulting control block to
    // ctl_next, which the column pass commits.  The state rows are loaded
    // before the decision is known, from the buffer an accepted step leaves
    // current; a rejected step without a stored RHS reloads the old state.
    const kptr<Ctl> c = kview(p_ctl);
    StepState st = load_state(c);
    const bool pending = c->pending != 0;
    // The twiddles go to LDS ahead of the tests, which keeps their loads ahead
    // of them too (after a test they would be issued behind the state rows,
    // and waiting for them would wait for those as well).
#pragma unroll
    for (int k = 0; k < TWQ; ++k) {
        if constexpr (THREADS * TWQ == 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
        else if (tid + THREADS * k < 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
    }
    if (!pending && ((st.status != ST_RUN) | (st.k_valid != 0))) {
        if (rblk == 0 && tid == 0) {
            copy_ctl(M.ctl_next, c);
            M.ctl_next->skip = 0;
        }
        return;
    }
    const double* wpred = pending ? (st.cur ? x0 : x1) : (st.w_is_x ? (st.cur ? x1 : x0) : wW);
    w = gin<double>(wpred);
    // the state rows due before the first march step
    static_for<-4 - SR_LA, -1>([&](auto tc) {
        issue_w(tc);
        __builtin_amdgcn_sched_barrier(0);
    });
    {
        // the order of reduce_partials: lane-strided, then butterflies
        double amax = 0.0, s2 = 0.0;
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            amax = fmax(amax, pa[j][0]);
            s2 += pa[j][1];
        }
        amax = wave_max(amax); s2 = wave_sum(s2);
        const StepState sd = decide_step(st, amax, s2, ph, N);
        if (pending) st = sd;
        // uniform across the block: keep the state in scalar registers
        st.cur = __builtin_amdgcn_readfirstlane(st.cur);
        st.w_is_x = __builtin_amdgcn_readfirstlane(st.w_is_x);
        st.k_valid = __builtin_amdgcn_readfirstlane(st.k_valid);
        st.status = __builtin_amdgcn_readfirstlane(st.status);
        st.t = uni(st.t); st.dt = uni(st.dt); st.tf = uni(st.tf); st.err = uni(st.err); st.tol = uni(st.tol);
        st.attempts = uni(st.attempts); st.accepts = uni(st.accepts);
    }
    // A step rejected without a stored RHS needs the old state, not the one
    // loaded: this attempt slot then does nothing for the domain (the column
    // and inverse row passes skip it) and the next attempt runs on the old
    // state.  Rare: about one attempt in a hundred.
    const double* wnow = st.w_is_x ? (st.cur ? x1 : x0) : wW;
    const int skip = __builtin_amdgcn_readfirstlane(wnow != wpred ? 1 : 0);
    // block 0 writes ctl_next after its last load (stores inside a one-lane
    // branch would make every later wait a full one)
    auto write_next = [&]() {
        if (rblk == 0 && tid == 0) {
            copy_ctl(M.ctl_next, c);
            store_state(M.ctl_next, st);
            M.ctl_next->pending = 0;
            M.ctl_next->skip = skip;
        }
    };
    if ((st.status != ST_RUN) | (st.k_valid != 0) | skip) {
        write_next();
        return;
    }
    // the march meets the domain edge it starts from first: top for a downward
    // block, bottom for an upward one
    srow_body<T, N, THREADS, LF, UP>(M, ph, R0, tid, UP ? bottom : top, UP ? top : bottom, smem, wv, dv, cv, ev,
                                     issue);
    write_next();
}

template <typename T, int N, int THREADS, int LF>
__global__ void __launch_bounds__(THREADS) kf_srow_fwd(const Member* __restrict__ mem, const void* __restrict__ twv) {
    using SR = SRow<T, N, THREADS, LF>;
    static_assert(SR::CPT >= 1 && SR::CPT * THREADS == N && SR::ROWS >= 2, "fused stencil shape");
    int mi, rblk;
    xcd_order(mi, rblk);
    if (rblk & 1) srow_block<T, N, THREADS, LF, true>(mem, twv, mi, rblk);
    else srow_block<T, N, THREADS, LF, false>(mem, twv, mi, rblk);
}

// COMMIT: the attempt began with the fused stencil pass, which left the
// control block in ctl_next; block 0 commits it (the other blocks read
// ctl_next, so nothing reads the block being written).
//
// Blocks whose rows are narrower than a 128-byte line (2C values of T) are
// ordered XCD-aware (as in the row passes): a domain's column blocks run on
// one XCD and the neighbours that split a line share its L2 (8 line pairs of
// FP32: 13.9 us per launch on the 64-member ensemble against 15.0 in the
// plain order).  Blocks of whole lines keep the plain order, which measured
// about 2% faster for them.
template <typename T, int N, int C, bool COMMIT, bool DB>
__global__ void __launch_bounds__(C * (N / 8)) kf_col(const Member* __restrict__ mem, const void* __restrict__ twv,
                                                     const double2* __restrict__ tw64) {
    int mi = blockIdx.y, cblk = blockIdx.x;
    if constexpr (2 * C * sizeof(T) < 128) xcd_order(mi, cblk);
    const Member& M = mem[mi];
    // the member's pointers in one batch of scalar loads
    const void* const p_src = M.g2;
    const void* const p_kr = M.kR;
    const Ctl* const p_ctl = COMMIT ? M.ctl_next : M.ctl;
    using PL = RegPlan<N>;
    constexpr int TPL = PL::TPL, E = PL::E;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    const int tid = threadIdx.x, line = tid % C, t = tid / C;
    // Loads in the order of use, each group behind a scheduling barrier so
    // that every wait covers only its own group: twiddles (stored to LDS
    // once the field loads are in flight), the spectrum (first transform),
    // then kappa*R (needed only after the first transform).  They are issued
    // before the control block is tested (the buffers exist whatever the
    // domain's state), so the test's load latency is not in front of them.
    constexpr int THREADS = C * (N / 8), TWQ = (2 * N + THREADS - 1) / THREADS;
    T twr[TWQ];
    {
        gptr<T> g = gin<T>(twv);
#pragma unroll
        for (int k = 0; k < TWQ; ++k) {
            if constexpr (THREADS * TWQ == 2 * N) twr[k] = g[tid + THREADS * k];
            else if (tid + THREADS * k < 2 * N) twr[k] = g[tid + THREADS * k];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    const int c0 = cblk * 2 * C;
    const int ca = c0 + 2 * line, cb = ca + 1;   // adjacent columns share a complex line
    using V2 = typename vec2<T>::type;
    // element e of the thread (row t + TPL e, columns ca and cb) lies lo bytes
    // past the uniform row base e * ESTEP: one 32-bit lane offset for all 8
    // loads of a field and scalar bases (global_load v, v_off, s_base), not 8
    // 64-bit lane addresses
    const unsigned lo = (unsigned)(t * N + ca) * (unsigned)sizeof(T);
    constexpr size_t ESTEP = (size_t)TPL * N / 2;
    // (the bases pass through an empty asm that holds them in scalar registers:
    // otherwise the compiler folds e * ESTEP into 64-bit lane addresses)
    auto at = [lo](auto base) {
        asm volatile("" : "+s"(base));
        return (decltype(base))((gptr<char>)base + lo);
    };
    gptr<V2> src = gin<V2>(p_src);
    gptr<V2> kR = gin<V2>(p_kr);
    cplx<T> v[E], m[E];
    T ka[E], kb[E];
    V2 xs[E], xk[E];
#pragma unroll
    for (int e = 0; e < E; ++e) xs[e] = *at(src + e * ESTEP);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < E; ++e) xk[e] = *at(kR + e * ESTEP);
    // cas weights of the corner terms (used by the threads holding rows 0 and N-1)
    const double2 wa = tw64[ca], wb = tw64[cb];
    __builtin_amdgcn_sched_barrier(0);
    const kptr<Ctl> c = kview(p_ctl);
    const bool idle = (c->status != ST_RUN) | (c->k_valid != 0) | (c->skip != 0);
    // The twiddles go to LDS ahead of the test, which keeps their loads ahead
    // of it too (after the test they would be issued behind the field loads,
    // and waiting for them would wait for kappa*R as well).  (No branch around
    // the store when the threads cover the table evenly: a join here would
    // make the compiler wait for every load in flight.)
#pragma unroll
    for (int k = 0; k < TWQ; ++k) {
        if constexpr (THREADS * TWQ == 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
        else if (tid + THREADS * k < 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
    }
    // block 0 commits after its last load: a one-lane store ahead of the loads
    // would make every later wait a full one
    const bool commit = COMMIT && cblk == 0 && threadIdx.x == 0;
    if (idle) {
        if (commit) copy_ctl(M.ctl, c);
        return;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        v[e] = mk<T>(xs[e].x, xs[e].y);
        ka[e] = xk[e].x;
        kb[e] = xk[e].y;
    }
    std::conditional_t<DB, ColXch2<T, N>, ColXch<T, N>> xc(buf, line, C);
    reg_fft_mirror<T, N>(v, m, t, tw_l, xc);
    // Hartley values and the spectral divide (lv_elva.f90:115)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        v[e] = mk<T>(kdiv(ha, ka[e]), kdiv(hb, kb[e]));
    }
    reg_fft_mirror<T, N, true>(v, m, t, tw_l, xc);
    gmut<V2> dst = gout<V2>(M.g1);
    gmut<double> cor = gout<double>(M.corners);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TPL * e;
        T ha, hb;
        dht_pair(v[e], m[e], ha, hb);
        V2 o;
        o.x = ha;
        o.y = hb;
        gmut<V2> row = dst + e * ESTEP;
        asm volatile("" : "+s"(row));
        *(gmut<V2>)((gmut<char>)row + lo) = o;
        // per-column terms of the four corners of the inverse transform (rows 0
        // and N-1, weights 1 and cas(2 pi col (N-1) / N)); the row pass sums them
        // in an order fixed by N alone
        if (k == 0 || k == N - 1) {
            const int q = (k == 0) ? 0 : 2;
            const double va = (double)ha, vb = (double)hb;
            cor[(size_t)ca * 4 + q] = va;
            cor[(size_t)ca * 4 + q + 1] = va * (wa.x + wa.y);
            cor[(size_t)cb * 4 + q] = vb;
            cor[(size_t)cb * 4 + q + 1] = vb * (wb.x + wb.y);
        }
    }
    if (commit) copy_ctl(M.ctl, c);
}

// Reduce the per-thread row statistics of two rows over the TPL threads of a
// line with a fixed butterfly (lines wider than a wave combine their wave
// halves through LDS in wave order; lines of 3 * 2^j threads combine their
// three 2^j-lane groups through LDS in group order), and store them as the
// row's partials.  The pattern depends on N alone, so the error norm of a
// domain does not depend on how many domains share the launch.
template <int TPL>
__device__ __forceinline__ void row_reduce_store(double a1a, double a2a, double s2a, double a1b, double a2b,
                                                 double s2b, int line, int t, gmut<double> pa) {
    constexpr bool P2 = (TPL & (TPL - 1)) == 0;
    constexpr int G = P2 ? TPL : TPL / 3;                  // lanes reduced by shuffles
    static_assert(P2 || (G * 3 == TPL && (G & (G - 1)) == 0 && G <= 64), "TPL = 2^j or 3 * 2^j");
    constexpr int W = G < 64 ? G : 64;
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        a1a = fmax(a1a, __shfl_xor(a1a, o, W)); a2a = fmax(a2a, __shfl_xor(a2a, o, W));
        s2a += __shfl_xor(s2a, o, W);
        a1b = fmax(a1b, __shfl_xor(a1b, o, W)); a2b = fmax(a2b, __shfl_xor(a2b, o, W));
        s2b += __shfl_xor(s2b, o, W);
    }
    constexpr int NG = P2 ? (TPL > 64 ? TPL / 64 : 1) : 3;   // groups combined through LDS
    if constexpr (NG > 1) {
        __shared__ double red[16][NG][6];
        if ((t % W) == 0) {
            double* r = red[line][t / W];
            r[0] = a1a; r[1] = a2a; r[2] = s2a; r[3] = a1b; r[4] = a2b; r[5] = s2b;
        }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < NG; ++w) {
                const double* r = red[line][w];
                a1a = fmax(a1a, r[0]); a2a = fmax(a2a, r[1]); s2a += r[2];
                a1b = fmax(a1b, r[3]); a2b = fmax(a2b, r[4]); s2b += r[5];
            }
        }
    }
    if (t == 0) {
        // one entry per line: rows 2j and 2j+1 combined in a fixed order
        pa[0] = fmax(fmax(a1a, a2a), fmax(a1b, a2b));
        pa[1] = s2a + s2b;
    }
}

// TILES row blocks per thread block: every load of all tiles is issued before
// the first transform, so the later tiles' data streams in while the first
// tile is transformed (the grid is otherwise a single occupancy round in
// which all blocks load, then compute, in lockstep).
// METH: the integrator (integrators.f90) compiled in, or -1 for a run-time one
template <typename T, int N, int THREADS, int TILES, int METH>
__global__ void __launch_bounds__(THREADS) kf_row_inv(const Member* __restrict__ mem, const void* __restrict__ twv,
                                                      int ncolblk) {
    // XCD-aware order: a domain's row blocks share an XCD, so its corner terms
    // and twiddles come from one L2 instead of eight
    int mi, rblk;
    xcd_order(mi, rblk);
    const Member& M = mem[mi];
    // the member's pointers in one batch of scalar loads
    const Ctl* const p_ctl = M.ctl;
    const void* const p_src = M.g1;
    const double* const p_cor = M.corners;
    scalars_ready(p_ctl, p_src, p_cor);
    using PL = RegPlan<N>;
    constexpr int TPL = PL::TPL, E = PL::E, L = THREADS / TPL;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* tw_l = (cplx<T>*)smem;
    cplx<T>* buf = tw_l + N;
    const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
    int ra[TILES];
#pragma unroll
    for (int k = 0; k < TILES; ++k) ra[k] = (rblk * TILES + k) * 2 * L + 2 * line;   // rb = ra + 1
    auto idx = [&](int r, int e) { return (size_t)r * N + t + TPL * e; };
    // The transform's loads, in the order of use, each group behind a
    // scheduling barrier so that every wait covers only the groups issued
    // before it: corner terms, twiddles, spectrum.  They are issued before the
    // control block is read (the buffers exist whatever the domain's state;
    // after a rejected attempt they go unused), so its load latency is not in
    // front of them; X (consumed last) follows once the live buffer is known.
    // The four corner sums over the N per-column terms of the column pass:
    // quantity q is summed lane-strided, then by a fixed butterfly, by wave
    // q (wave w of NW < 4 takes q = w, w + NW, ... below 4; with more, waves
    // 4.. repeat waves 0..3), so the order is set by N alone.
    constexpr int NW = THREADS / 64, QPW = NW >= 4 ? 1 : (4 + NW - 1) / NW, CPL = N / 64;
    static_assert(NW >= 1 && THREADS % 64 == 0 && N % 64 == 0, "corner-sum layout");
    const int wave = tid >> 6, lane = tid & 63;
    double cq[QPW][CPL];
    {
        gptr<double> cp = gin<double>(p_cor);
#pragma unroll
        for (int j = 0; j < QPW; ++j) {
            const int qi = NW >= 4 ? (wave & 3) : wave + NW * j;
            const int qc = qi < 4 ? qi : 0;   // (a wave past the four quantities loads one it discards)
#pragma unroll
            for (int i = 0; i < CPL; ++i) cq[j][i] = cp[(size_t)(lane + 64 * i) * 4 + qc];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    constexpr int TWQ = (2 * N + THREADS - 1) / THREADS;
    T twr[TWQ];
    {
        gptr<T> g = gin<T>(twv);
#pragma unroll
        for (int k = 0; k < TWQ; ++k) {
            if constexpr (THREADS * TWQ == 2 * N) twr[k] = g[tid + THREADS * k];
            else if (tid + THREADS * k < 2 * N) twr[k] = g[tid + THREADS * k];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    cplx<T> v[TILES][E];
    {
        gptr<T> src = gin<T>(p_src);
#pragma unroll
        for (int k = 0; k < TILES; ++k)
#pragma unroll
            for (int e = 0; e < E; ++e) v[k][e] = mk<T>(src[idx(ra[k], e)], src[idx(ra[k] + 1, e)]);
    }
    __builtin_amdgcn_sched_barrier(0);
    const kptr<Ctl> c = kview(p_ctl);
    const bool idle = (c->status != ST_RUN) | (c->skip != 0);
    // The twiddles go to LDS ahead of the test, which keeps their loads ahead
    // of it too (after it they would be issued behind the spectrum and X, and
    // waiting for them would wait for those as well).  (No branch around the
    // store when the threads cover the table evenly: a join here would make
    // the compiler wait for every load in flight.)
#pragma unroll
    for (int k = 0; k < TWQ; ++k) {
        if constexpr (THREADS * TWQ == 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
        else if (tid + THREADS * k < 2 * N) ((T*)tw_l)[tid + THREADS * k] = twr[k];
    }
    if (idle) return;
    const kptr<PhysDev> ph = kview(M.phys);
    const int meth = ph->dt_method;
    const double dt = attempt_dt(meth, c->dt, c->t, c->tf, ph->dt_max);
    const int cur = c->cur;
    const bool k_valid = c->k_valid != 0;
    // block 0 marks the decision pending (the step control reads this attempt
    // next) after its last load: a one-lane store ahead of the loads would
    // make every later wait a full one
    const bool mark = rblk == 0 && tid == 0;
    gptr<double> X = gin<double>(M.X[cur]);
    gmut<double> Xn = gout<double>(M.X[cur ^ 1]);
    gmut<double> dw = gout<double>(M.dwdt);
    gmut<double> part = gout<double>(M.partials);
    auto update = [&](size_t cidx, double x, double k, double& a1, double& a2, double& s2) {
        double x1;
        double x2;
        if constexpr (METH >= 0) x2 = stage_sums_m<METH>(x, k, dt, x1);
        else x2 = stage_sums(meth, x, k, dt, x1);
        Xn[cidx] = x2;
        a1 = fmax(a1, fabs(x1));
        a2 = fmax(a2, fabs(x2));
        const double d = x2 - x1;
        s2 += d * d;
    };
    double xa[TILES][E], xb[TILES][E];
    if (k_valid) {
        // rejected attempt before: the RHS is unchanged, only dt differs
        double ka[TILES][E], kb[TILES][E];
#pragma unroll
        for (int k = 0; k < TILES; ++k)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                ka[k][e] = dw[idx(ra[k], e)];
                kb[k][e] = dw[idx(ra[k] + 1, e)];
            }
#pragma unroll
        for (int k = 0; k < TILES; ++k)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                xa[k][e] = X[idx(ra[k], e)];
                xb[k][e] = X[idx(ra[k] + 1, e)];
            }
#pragma unroll
        for (int k = 0; k < TILES; ++k) {
            double a1a = 0.0, a2a = 0.0, s2a = 0.0, a1b = 0.0, a2b = 0.0, s2b = 0.0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                update(idx(ra[k], e), xa[k][e], ka[k][e], a1a, a2a, s2a);
                update(idx(ra[k] + 1, e), xb[k][e], kb[k][e], a1b, a2b, s2b);
            }
            row_reduce_store<TPL>(a1a, a2a, s2a, a1b, a2b, s2b, line, t, part + (size_t)(ra[k] / 2) * 4);
        }
        if (mark) M.ctl->pending = 1;
    } else {
        __shared__ double qs[4];
        // /(nx*ny*1.) (lv_elva.f90:117): for a power-of-two N, N*N is a power
        // of two and the division is a multiplication by its exact inverse
        const double S = (double)(float)((double)N * (double)N), rS = 1.0 / S;
        auto scl = [&](double x) {
            if constexpr (IsPow2<N>::value) return x * rS;
            else return x / S;
        };
        cplx<T> m[E];
#pragma unroll
        for (int k = 0; k < TILES; ++k)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                xa[k][e] = X[idx(ra[k], e)];
                xb[k][e] = X[idx(ra[k] + 1, e)];
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < QPW; ++j) {
            double q = 0;
#pragma unroll
            for (int i = 0; i < CPL; ++i) q += cq[j][i];
            q = wave_sum(q);
            const int qi = NW >= 4 ? (wave & 3) : wave + NW * j;
            if (lane == 0 && qi < 4) qs[qi] = q;
        }
        __syncthreads();
        // x(1,1) + x(nx,ny) + x(1,ny) + x(nx,1)  (isos_utils.f90:116)
        const double cc = 0.25 * (scl(qs[0]) + scl(qs[3]) + scl(qs[2]) + scl(qs[1]));
        const double spy = ph->sec_per_year;
        const bool store_k = !(c->t + dt < c->tf);
        RowXch<T, N> xc(buf, line);
#pragma unroll
        for (int k = 0; k < TILES; ++k) {
            if (k > 0) __syncthreads();   // the exchange buffer is reused
            reg_fft_mirror<T, N, true>(v[k], m, t, tw_l, xc);
            double a1a = 0.0, a2a = 0.0, s2a = 0.0, a1b = 0.0, a2b = 0.0, s2b = 0.0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                T ha, hb;
                dht_pair(v[k][e], m[e], ha, hb);
                const size_t ca = idx(ra[k], e), cb = idx(ra[k] + 1, e);
                const double ka = (scl((double)ha) - cc) * spy, kb = (scl((double)hb) - cc) * spy;
                if (store_k) {
                    dw[ca] = ka;
                    dw[cb] = kb;
                }
                update(ca, xa[k][e], ka, a1a, a2a, s2a);
                update(cb, xb[k][e], kb, a1b, a2b, s2b);
            }
            row_reduce_store<TPL>(a1a, a2a, s2a, a1b, a2b, s2b, line, t, part + (size_t)(ra[k] / 2) * 4);
        }
        if (mark) M.ctl->pending = 1;
    }
}

// ------------------------------------------------------------------ dispatch
template <typename T, int N>
static size_t row_lds(int threads) {
    const int L = threads / (N / 8);
    return ((size_t)N + (size_t)L * RowXch<T, N>::PITCH) * sizeof(cplx<T>);
}
template <typename T, int N>
constexpr size_t col_lds(int C, bool db = false) {
    return ((size_t)N + (size_t)N * (C + 1) * (db ? 2 : 1)) * sizeof(cplx<T>);
}

template <typename T, int N, int THREADS>
static void rows_fwd(const Member* d, int nmem, hipStream_t st, const void* tw) {
    constexpr int L = THREADS / (N / 8);
    const size_t lds = row_lds<T, N>(THREADS);
    launch_k(1, kf_row_fwd<T, N, THREADS>, dim3(N / (2 * L), nmem), dim3(THREADS), lds, st, d, tw);
}
// LDS of one workgroup on gfx950 (160 KiB per CU); launches above 64 KiB of
// dynamic LDS declare it to the runtime first
constexpr size_t FI_LDS_MAX = 160 * 1024;

template <typename T, int N, int THREADS, int LF = THREADS / (N / 8)>
static bool srow_fwd(const Member* d, int nmem, hipStream_t st, const void* tw) {
    using SR = SRow<T, N, THREADS, LF>;
    if constexpr (SR::CPT >= 1 && SR::CPT * THREADS == N && (SR::ROWS >= 4 || THREADS == N || THREADS == 384) &&
                  SR::lds <= FI_LDS_MAX) {
        if constexpr (SR::lds > 65536) {
            static const bool ok = hipFuncSetAttribute((const void*)kf_srow_fwd<T, N, THREADS, LF>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)SR::lds) == hipSuccess;
            if (!ok) return false;   // the separate stencil and row kernels run instead
        }
        launch_k(0, kf_srow_fwd<T, N, THREADS, LF>, dim3(N / SR::ROWS, nmem), dim3(THREADS), SR::lds, st, d, tw);
        return true;
    }
    return false;
}
// one column per thread (THREADS = N, or 384 at N = 384) with 4, 8 or 16
// rows per block (srow_rows): 4 and 8 for small batches, 16 where the
// row blocks' own thread count would give a thread several columns (FP64 at
// N = 512, whose 256-thread blocks spill the register window)
template <typename T, int N, int THREADS>
static bool srow_fwd_rows(const Member* d, int nmem, hipStream_t st, const void* tw, int rows) {
    if constexpr (THREADS / (N / 8) >= 4) {
        if (rows == 2) return srow_fwd<T, N, THREADS, 1>(d, nmem, st, tw);
        if (rows == 4) return srow_fwd<T, N, THREADS, 2>(d, nmem, st, tw);
        if (rows == 8) return srow_fwd<T, N, THREADS, 4>(d, nmem, st, tw);
    }
    if constexpr (THREADS / (N / 8) >= 8)
        if (rows == 16) return srow_fwd<T, N, THREADS, 8>(d, nmem, st, tw);
    return false;
}
template <typename T, int N, int THREADS, int METH>
static void rows_inv_m(const Member* d, int nmem, hipStream_t st, const void* tw, int ncolblk, int tiles) {
    constexpr int L = THREADS / (N / 8);
    const size_t lds = row_lds<T, N>(THREADS);
    if (tiles == 2 && (N / (2 * L)) % 2 == 0)
        launch_k(3, kf_row_inv<T, N, THREADS, 2, METH>, dim3(N / (4 * L), nmem), dim3(THREADS), lds, st, d, tw,
                 ncolblk);
    else
        launch_k(3, kf_row_inv<T, N, THREADS, 1, METH>, dim3(N / (2 * L), nmem), dim3(THREADS), lds, st, d, tw,
                 ncolblk);
}
// tsit54, the reference's default integrator, is compiled into the epilogue
template <typename T, int N, int THREADS>
static void rows_inv(const Member* d, int nmem, hipStream_t st, const void* tw, int ncolblk, int tiles,
                     int dt_method) {
    if (dt_method == 3) rows_inv_m<T, N, THREADS, 3>(d, nmem, st, tw, ncolblk, tiles);
    else rows_inv_m<T, N, THREADS, -1>(d, nmem, st, tw, ncolblk, tiles);
}
// Two exchange buffers in the column pass (one barrier per exchange instead
// of two) where the blocks of a full CU (16 waves) still fit its LDS: FP32
// lines of 2^k, 8 line pairs or more at N = 256 (64 x 256^2, 8 line pairs:
// 12.2 us per launch against 12.5-12.9 with one buffer, same build and box).
template <typename T, int N, int C>
constexpr bool col_two_buffers() {
    constexpr int blocks_per_cu = 8192 / (C * N) > 1 ? 8192 / (C * N) : 1;
    return IsPow2<N>::value && sizeof(T) == 4 && col_lds<T, N>(C, true) * blocks_per_cu <= FI_LDS_MAX;
}
template <typename T, int N, int C, bool DB>
static void cols_launch(const Member* d, int nmem, hipStream_t st, const void* tw, const double2* tw64, bool commit) {
    const size_t lds = col_lds<T, N>(C, DB);
    if (commit)
        launch_k(2, kf_col<T, N, C, true, DB>, dim3(N / (2 * C), nmem), dim3(C * (N / 8)), lds, st, d, tw, tw64);
    else
        launch_k(2, kf_col<T, N, C, false, DB>, dim3(N / (2 * C), nmem), dim3(C * (N / 8)), lds, st, d, tw, tw64);
}
// whether the column pass of this shape runs with two buffers (launches above
// 64 KiB of dynamic LDS declare it to the runtime first; if that fails, one)
template <typename T, int N, int C>
static bool col_two_buffers_ready() {
    if constexpr (col_two_buffers<T, N, C>()) {
        constexpr int lds = (int)col_lds<T, N>(C, true);
        static const bool ok =
            lds <= 65536 ||
            (hipFuncSetAttribute((const void*)kf_col<T, N, C, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 lds) == hipSuccess &&
             hipFuncSetAttribute((const void*)kf_col<T, N, C, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 lds) == hipSuccess);
        return ok;
    }
    return false;
}
template <typename T, int N, int C>
static void cols(const Member* d, int nmem, hipStream_t st, const void* tw, const double2* tw64, bool commit) {
    if constexpr (col_two_buffers<T, N, C>()) {
        if (col_two_buffers_ready<T, N, C>()) {
            cols_launch<T, N, C, true>(d, nmem, st, tw, tw64, commit);
            return;
        }
    }
    cols_launch<T, N, C, false>(d, nmem, st, tw, tw64, commit);
}

template <typename T, int N>
static bool fast_n(const Member* d, int nmem, const RhsShape& s, hipStream_t st, int which) {
    // which: 0 fused stencil + row forward, 1 row forward, 2 column (4: column
    // committing the control block of the fused pass), 3 row inverse
    if (which == 2 || which == 4) {
        const bool cm = which == 4;
        // (shapes whose LDS exceeds 64 KiB are never chosen by the host and not instantiated)
        constexpr auto fits = [](int C) { return C * (N / 8) <= 1024 && col_lds<T, N>(C) <= 65536; };
        switch (s.C) {
        case 1: cols<T, N, 1>(d, nmem, st, s.tw, s.tw64, cm); return true;
        case 2: if constexpr (fits(2)) { cols<T, N, 2>(d, nmem, st, s.tw, s.tw64, cm); return true; } break;
        case 4: if constexpr (fits(4)) { cols<T, N, 4>(d, nmem, st, s.tw, s.tw64, cm); return true; } break;
        case 8: if constexpr (fits(8)) { cols<T, N, 8>(d, nmem, st, s.tw, s.tw64, cm); return true; } break;
        case 16: if constexpr (fits(16)) { cols<T, N, 16>(d, nmem, st, s.tw, s.tw64, cm); return true; } break;
        }
        return false;
    }
    const int ncolblk = s.n / (2 * s.C);
    if constexpr (Is3Pow2<N>::value) {
        // N = 3 * 2^k: 384-thread row blocks (8 lines of 48 threads at N = 384,
        // 16 of 24 at N = 192)
        if (s.threads_row != 384) return false;
        if (which == 0)
            return s.srow_rows ? srow_fwd_rows<T, N, 384>(d, nmem, st, s.tw, s.srow_rows)
                               : srow_fwd<T, N, 384>(d, nmem, st, s.tw);
        if (which == 1) rows_fwd<T, N, 384>(d, nmem, st, s.tw);
        // the inverse pass also in 192-thread blocks (3 waves) for small batches
        else if (s.threads_inv == 192) rows_inv<T, N, 192>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method);
        else rows_inv<T, N, 384>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method);
        return true;
    } else {
    if (which == 0) {
        if (s.srow_rows) {
            // one column per thread, 4 or 8 rows per block
            if constexpr (N <= 512) return srow_fwd_rows<T, N, N>(d, nmem, st, s.tw, s.srow_rows);
            return false;
        }
        switch (s.threads_row) {
        case 64: if constexpr (N / 8 <= 64) return srow_fwd<T, N, 64>(d, nmem, st, s.tw); break;
        case 128: if constexpr (N / 8 <= 128) return srow_fwd<T, N, 128>(d, nmem, st, s.tw); break;
        case 256: if constexpr (N / 8 <= 256) return srow_fwd<T, N, 256>(d, nmem, st, s.tw); break;
        // one column per thread at N = 512 (two columns per thread spill)
        case 512: if constexpr (N == 512 && sizeof(T) == 4) return srow_fwd<T, N, 512>(d, nmem, st, s.tw); break;
        }
        return false;
    }
    // the inverse row pass may take its own block size (threads_inv)
    switch (which == 3 && s.threads_inv ? s.threads_inv : s.threads_row) {
    case 64: if constexpr (N / 8 <= 64) { which == 1 ? rows_fwd<T, N, 64>(d, nmem, st, s.tw) : rows_inv<T, N, 64>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method); return true; } break;
    case 128: if constexpr (N / 8 <= 128) { which == 1 ? rows_fwd<T, N, 128>(d, nmem, st, s.tw) : rows_inv<T, N, 128>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method); return true; } break;
    case 256: if constexpr (N / 8 <= 256) { which == 1 ? rows_fwd<T, N, 256>(d, nmem, st, s.tw) : rows_inv<T, N, 256>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method); return true; } break;
    case 512: if constexpr (N == 512 && sizeof(T) == 4) { which == 1 ? rows_fwd<T, N, 512>(d, nmem, st, s.tw) : rows_inv<T, N, 512>(d, nmem, st, s.tw, ncolblk, s.row_tiles, s.dt_method); return true; } break;
    }
    return false;
    }
}

template <typename T>
static bool fast_t(const Member* d, int nmem, const RhsShape& s, hipStream_t st, int which) {
    switch (s.n) {
    case 64: return fast_n<T, 64>(d, nmem, s, st, which);
    case 128: return fast_n<T, 128>(d, nmem, s, st, which);
    case 256: return fast_n<T, 256>(d, nmem, s, st, which);
    case 512: return fast_n<T, 512>(d, nmem, s, st, which);
    case 1024: return fast_n<T, 1024>(d, nmem, s, st, which);
    case 192: return fast_n<T, 192>(d, nmem, s, st, which);
    case 384: return fast_n<T, 384>(d, nmem, s, st, which);
    }
    return false;
}

template <typename T, int N, int C>
static int col_buffers_c() {
    // (the shapes fast_n launches: at most 1024 threads and 64 KiB with one buffer)
    if constexpr (C * (N / 8) <= 1024 && col_lds<T, N>(C) <= 65536) return col_two_buffers_ready<T, N, C>() ? 2 : 1;
    return 1;
}
template <typename T, int N>
static int col_buffers_n(int C) {
    switch (C) {
    case 1: return col_buffers_c<T, N, 1>();
    case 2: return col_buffers_c<T, N, 2>();
    case 4: return col_buffers_c<T, N, 4>();
    case 8: return col_buffers_c<T, N, 8>();
    case 16: return col_buffers_c<T, N, 16>();
    }
    return 1;
}
template <typename T>
static int col_buffers_t(int n, int C) {
    switch (n) {
    case 64: return col_buffers_n<T, 64>(C);
    case 128: return col_buffers_n<T, 128>(C);
    case 256: return col_buffers_n<T, 256>(C);
    case 512: return col_buffers_n<T, 512>(C);
    case 1024: return col_buffers_n<T, 1024>(C);
    }
    return 1;
}
int col_buffers(const RhsShape& s) { return s.fp64 ? col_buffers_t<double>(s.n, s.C) : col_buffers_t<float>(s.n, s.C); }

bool launch_fast_pass(const Member* d_mem, int nmem, const RhsShape& s, hipStream_t st, int which) {
    return s.fp64 ? fast_t<double>(d_mem, nmem, s, st, which) : fast_t<float>(d_mem, nmem, s, st, which);
}

}  // namespace fi
