// sealevel.hip -- per-update bookkeeping of isos_update on the GPU.
//
//   k_pre_haf     Hice embedding (out2in), calc_Haf and the two domain sums of
//                 calc_sl_contributions (sealevel.f90:132-145, 189-205)
//   k_pre_bsl     V_af / V_den / deltaV and the barystatic sea-level update
//                 (fastisostasy.f90:763-778).  One wave per group of domains
//                 sharing a bsl object walks the group in batch order, so the
//                 group sees the same sequence of updates as consecutive
//                 isos_update calls; independent groups run in parallel.
//   k_pre_fields  z_ss, masks and the column anomalies (sealevel.f90:25-187,
//                 fastisostasy.f90:778-833) in three phases around the two
//                 Green's-function convolutions
//   k_finalize    corner BC of now%w, canom_mantle, z_bed and the instability
//                 maxima (fastisostasy.f90:856-869)
// All bookkeeping is FP64; masks are bit-packed in one byte per cell.
#include "fi_device.hpp"
#include "fi_kernels.h"
#include "rhs_common.hpp"

namespace fi {

constexpr int PRE_THREADS = 256, PRE_CELLS = 1024;   // cells per block
// values per block of the finalize reduction: max |w|, |we|, |dz_ss|, NaN flag, max |w_equilibrium|
constexpr int FIN_P = 5;

enum { MB_OCEAN = 1, MB_GROUNDED = 2, MB_CONTINENT = 4, MB_ACTIVE = 8 };

__device__ __forceinline__ double calc_haf(double z_ss, double z_bed, double Hice, const PhysDev& ph) {
    double h = (z_ss - z_bed) * (ph.rho_seawater / ph.rho_ice);
    if (h < 0.0) h = 0.0;
    h = Hice - h;
    if (h < 0.0) h = 0.0;
    return h;
}

__device__ __forceinline__ uint8_t calc_mask(double z_bed, double z_ss, double Hice, double Haf, uint8_t keep) {
    const bool cont = z_bed > z_ss;
    const bool grd = (Hice > 0.0) && (Haf > 0.0);
    const bool ocean = !(cont || grd);
    return (uint8_t)((keep & MB_ACTIVE) | (ocean ? MB_OCEAN : 0) | (grd ? MB_GROUNDED : 0) | (cont ? MB_CONTINENT : 0));
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    return s;
}

__global__ void __launch_bounds__(PRE_THREADS) k_pre_haf(const Member* __restrict__ mem, int new_H) {
    const Member& M = mem[blockIdx.y];
    const kptr<PhysDev> ph = kview(M.phys);
    const int N = M.n;
    const size_t nn = (size_t)N * N;
    double s1 = 0.0, s2 = 0.0;
    const double A_uni = ph->dx * ph->dx, Vden = ph->Vden_factor, rsw_ri = ph->rho_seawater / ph->rho_ice;
    const bool uniform = ph->uniform_K;
    gptr<double> Hin = gin<double>(M.H_in), zss = gin<double>(M.z_ss), zbed = gin<double>(M.z_bed),
                 Haf_ref = gin<double>(M.Haf_ref), Hice_ref = gin<double>(M.Hice_ref);
    gmut<double> Hice = gout<double>(M.Hice), Haf = gout<double>(M.Haf);
    // every load of the thread's cells first (indices clamped, so no branch),
    // then the arithmetic and the guarded stores
    constexpr int IT = PRE_CELLS / PRE_THREADS;
    double Hv[IT], zs[IT], zb[IT], hr[IT], ir[IT], Av[IT];
    bool inside[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        c = c < nn ? c : nn - 1;
        if (new_H) {
            // out2in embedding of the ice grid (isos_utils.f90:330-345)
            const int i = (int)(c % N) - M.icrop1, j = (int)(c / N) - M.jcrop1;
            inside[it] = i >= 0 && i < M.nx_ice && j >= 0 && j < M.ny_ice;
            const int ic = i < 0 ? 0 : (i >= M.nx_ice ? M.nx_ice - 1 : i);
            const int jc = j < 0 ? 0 : (j >= M.ny_ice ? M.ny_ice - 1 : j);
            Hv[it] = Hin[(size_t)jc * M.nx_ice + ic];
        } else {
            inside[it] = true;
            Hv[it] = Hice[c];
        }
        zs[it] = zss[c]; zb[it] = zbed[c]; hr[it] = Haf_ref[c]; ir[it] = Hice_ref[c];
        Av[it] = uniform ? A_uni : M.A[c];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        if (c >= nn) break;
        const double H = inside[it] ? Hv[it] : 0.0;
        if (new_H) Hice[c] = H;
        // calc_Haf (sealevel.f90:132-145)
        double h = (zs[it] - zb[it]) * rsw_ri;
        if (h < 0.0) h = 0.0;
        h = H - h;
        if (h < 0.0) h = 0.0;
        Haf[c] = h;
        const double A = Av[it];
        s1 += (h - hr[it]) * A;
        s2 += (H - ir[it]) * Vden * A;
    }
    __shared__ double red[16];
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
        M.partials[(size_t)blockIdx.x * 4 + 0] = s1;
        M.partials[(size_t)blockIdx.x * 4 + 1] = s2;
    }
}

__global__ void __launch_bounds__(64) k_pre_bsl(const Member* __restrict__ mem, int nmem, int nblk) {
    const int lane = threadIdx.x;
    if (!mem[blockIdx.x].bsl_lead) return;
    for (int m = blockIdx.x; m >= 0 && m < nmem; m = mem[m].bsl_next) {
        const Member& M = mem[m];
        double s1 = 0.0, s2 = 0.0;
        for (int b = lane; b < nblk; b += 64) { s1 += M.partials[(size_t)b * 4]; s2 += M.partials[(size_t)b * 4 + 1]; }
        s1 = wave_sum(s1); s2 = wave_sum(s2);
        if (lane == 0) {
            const PhysDev& ph = *M.phys;
            Ctl& c = *M.ctl;
            const double V_old = c.V_af + c.V_den + c.V_pov;
            c.V_af = s1 * (ph.rho_ice / ph.rho_seawater);
            c.V_den = s2;
            c.V_pov = 0.0;
            c.deltaV = (c.V_af + c.V_den + c.V_pov) - V_old;
            BslDev& b = *M.bsl;
            if (b.method == 1) b.bsl_now = b.bsl_now - c.deltaV / b.A_ocean_now;
            if (b.method == 3) {
                // mixed: the domain's cumulative contribution comes off the series value (:771-775)
                c.dbsl_total = c.dbsl_total + c.deltaV / b.A_ocean_now;
                b.bsl_now = b.bsl_now - c.dbsl_total;
            }
            c.bsl = b.bsl_now;
        }
        __syncthreads();
    }
}

// Variant PH of the per-update field pass:
//   PF_FULL   phase 0 without diagnostics, and phase 2 (after the sea-surface
//             convolution): z_ss, masks, load and full column anomalies
//   PF_DIAG   phase 0 with diagnostics: z_ss, masks and canom_load, then the
//             input of the elastic convolution (or we = 0)
//   PF_LITHO  phase 1 (after the elastic convolution): canom_litho,
//             canom_full, mass_anom (fastisostasy.f90:811-813)
//   PF_ALL    phases 0, 1 and 2 without convolutions (we = 0, dz_ss = 0)
//   PF_NOSSH  phases 1 and 2 without the sea-surface convolution (dz_ss = 0)
// The fused passes apply the same operations to each cell in the same order.
// Each thread loads every input of its cells first and then computes and
// stores: a store between loads would make the next wait cover it too.
enum { PF_FULL = 0, PF_LITHO = 1, PF_ALL = 3, PF_NOSSH = 4, PF_DIAG = 5 };

template <typename T, int PH>
__global__ void __launch_bounds__(PRE_THREADS) k_pre_fields(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.y];
    const kptr<PhysDev> ph = kview(M.phys);
    const int N = M.n;
    const size_t nn = (size_t)N * N;
    const double bsl = kview(M.ctl)->bsl;
    const double rho_ice = ph->rho_ice, rho_water = ph->rho_water, rho_litho = ph->rho_litho, g = ph->g;
    const bool uniform = ph->uniform_K, interactive = ph->interactive_sealevel, elastic = ph->include_elastic,
               hssh = ph->heterogeneous_ssh;
    const bool keep_load = ph->method != 3;   // methods 0-2 relax towards an equilibrium of canom_load
    const double A_uni = ph->dx * ph->dx;
    gmut<uint8_t> mask = gout<uint8_t>(M.mask);
    gmut<double> zss = gout<double>(M.z_ss), dzss = gout<double>(M.dz_ss), we = gout<double>(M.we),
                 cload = gout<double>(M.canom_load), cfull = gout<double>(M.canom_full),
                 conv = gout<double>(M.conv_in);
    gptr<double> zss_ref = gin<double>(M.z_ss_ref), zbed = gin<double>(M.z_bed), Hice = gin<double>(M.Hice),
                 Haf = gin<double>(M.Haf), Haf_ref = gin<double>(M.Haf_ref), we_ref = gin<double>(M.we_ref),
                 cmantle = gin<double>(M.canom_mantle);
    gmut<T> cf_rhs = gout<T>(M.cfull);
    constexpr bool SEA = PH != PF_LITHO;                                     // z_ss, masks, load
    constexpr bool DZ = PH == PF_FULL || PH == PF_DIAG || PH == PF_ALL;       // reads dz_ss
    constexpr bool WE = PH == PF_FULL || PH == PF_LITHO || PH == PF_NOSSH;    // reads we
    constexpr bool FULL = PH != PF_DIAG;                                      // reads we_ref, canom_mantle
    struct In {
        uint8_t m;
        double zr, hf, zb, hi, hfr, dz, we, wr, cm, cl, aux;
    };
    constexpr int IT = PRE_CELLS / PRE_THREADS;
    In v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        c = c < nn ? c : nn - 1;
        v[it].m = mask[c];
        if constexpr (SEA) {
            v[it].zr = zss_ref[c]; v[it].hf = Haf[c]; v[it].zb = zbed[c]; v[it].hi = Hice[c]; v[it].hfr = Haf_ref[c];
        }
        if constexpr (DZ) v[it].dz = dzss[c];
        if constexpr (WE) v[it].we = we[c];
        if constexpr (FULL) { v[it].wr = we_ref[c]; v[it].cm = cmantle[c]; }
        if constexpr (PH == PF_LITHO) {
            v[it].cl = cload[c];
            v[it].aux = uniform ? A_uni : M.A[c];
        }
        if constexpr (PH == PF_DIAG) v[it].aux = uniform ? 1.0 : M.K2[c];
    }
    // calc_z_ss, calc_masks, seawater and load anomalies for one cell
    auto sea_load = [&](size_t c, const In& x, uint8_t mold, bool active, double dz, uint8_t& mk, bool store) {
        const double z_ss = bsl + x.zr + dz;
        const double haf = x.hf;
        mk = calc_mask(x.zb, z_ss, x.hi, haf, mold);
        if (store) {
            zss[c] = z_ss;
            mask[c] = mk;
        }
        const double ice = 0.0 + rho_ice * (haf - x.hfr);
        const double sw = 0.0 + ((mk & MB_OCEAN) ? rho_water * (z_ss - x.zr) : 0.0);
        double load = ice;
        if (interactive) load = load + sw;
        return active ? load : 0.0;
    };
    auto full_of = [&](size_t c, const In& x, bool active, double load, double wec) {
        if (keep_load) cload[c] = load;   // the last calc_columnanoms_load of the update (:833)
        const double litho = 0.0 + rho_litho * (wec - x.wr);
        const double full = active ? (load + litho + x.cm) : 0.0;
        cfull[c] = full;
        cf_rhs[c] = (T)full;
    };
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        if (c >= nn) break;
        const In& x = v[it];
        const uint8_t mold = x.m;
        const bool active = mold & MB_ACTIVE;
        uint8_t mk;
        if constexpr (PH == PF_LITHO) {
            const double litho = 0.0 + rho_litho * (x.we - x.wr);
            const double full = active ? (x.cl + litho + x.cm) : 0.0;
            if (hssh) conv[c] = active ? x.aux * full : 0.0;
            else dzss[c] = 0.0;
        } else if constexpr (PH == PF_NOSSH) {
            dzss[c] = 0.0;
            const double load = sea_load(c, x, mold, active, 0.0, mk, true);
            full_of(c, x, active, load, x.we);
        } else if constexpr (PH == PF_ALL) {
            // (z_ss and the masks are overwritten by the second pass: no stores in the first)
            const double load = sea_load(c, x, mold, active, x.dz, mk, false);
            cload[c] = load;
            we[c] = 0.0;
            dzss[c] = 0.0;
            const double load2 = sea_load(c, x, mk, active, 0.0, mk, true);
            full_of(c, x, active, load2, 0.0);
        } else if constexpr (PH == PF_DIAG) {
            const double load = sea_load(c, x, mold, active, x.dz, mk, true);
            cload[c] = load;
            if (elastic) conv[c] = load * g * x.aux;
            else we[c] = 0.0;
        } else {
            const double load = sea_load(c, x, mold, active, x.dz, mk, true);
            full_of(c, x, active, load, x.we);
        }
    }
}

// LV-ELRA input of the viscous convolution, -canom_load g K^2
// (fastisostasy.f90:702-708), for domains that step in this update.
__global__ void __launch_bounds__(PRE_THREADS) k_weq_input(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.y];
    if (kview(M.ctl)->status != ST_RUN) return;
    const kptr<PhysDev> ph = kview(M.phys);
    const size_t nn = (size_t)M.n * M.n;
    const double g = ph->g;
    const bool uniform = ph->uniform_K;
    gptr<double> cload = gin<double>(M.canom_load);
    gmut<double> conv = gout<double>(M.conv_in);
#pragma unroll
    for (int it = 0; it < PRE_CELLS / PRE_THREADS; ++it) {
        const size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        if (c >= nn) break;
        const double K2 = uniform ? 1.0 : M.K2[c];
        conv[c] = -(cload[c] * g * K2);
    }
}

// w_equilibrium of methods 0-2 for domains that step in this update (the
// reference recomputes it in every RHS call from fields that stay fixed
// during the update: get_dwdt, fastisostasy.f90:686-714).
//   0: 0;  1: -canom_load / rho_uppermantle (lv_xlra.f90:29-38);
//   2: corner BC (isos_utils.f90:104-118) of the viscous convolution in weq.
// One block per domain, so the corner values are read before any is changed.
constexpr int WEQ_THREADS = 1024;
__global__ void __launch_bounds__(WEQ_THREADS) k_weq(const Member* __restrict__ mem, int method) {
    const Member& M = mem[blockIdx.x];
    if (kview(M.ctl)->status != ST_RUN) return;
    const int N = M.n;
    const size_t nn = (size_t)N * N;
    gmut<double> weq = gout<double>(M.weq);
    if (method == 0) {
        for (size_t c = threadIdx.x; c < nn; c += WEQ_THREADS) weq[c] = 0.0;
        return;
    }
    if (method == 1) {
        const double rho_um = kview(M.phys)->rho_uppermantle;
        gptr<double> cload = gin<double>(M.canom_load);
        for (size_t c = threadIdx.x; c < nn; c += WEQ_THREADS) weq[c] = -cload[c] / rho_um;
        return;
    }
    const double* raw = M.weq;
    const double cc = 0.25 * (raw[0] + raw[nn - 1] + raw[(size_t)(N - 1) * N] + raw[N - 1]);
    __syncthreads();
    for (size_t c = threadIdx.x; c < nn; c += WEQ_THREADS) weq[c] = weq[c] - cc;
}

void launch_weq_input(const Member* d_mem, int nmem, int n, hipStream_t st) {
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(k_weq_input, dim3((unsigned)((nn + PRE_CELLS - 1) / PRE_CELLS), nmem), dim3(PRE_THREADS), 0, st,
                       d_mem);
}

void launch_weq(const Member* d_mem, int nmem, int n, int method, hipStream_t st) {
    (void)n;
    hipLaunchKernelGGL(k_weq, dim3(nmem), dim3(WEQ_THREADS), 0, st, d_mem, method);
}

// Corner correction of now%w (isos_utils.f90:104-110), one thread per domain.
// The finalize kernels run once per update and domain, for domains that have
// stopped stepping: the host may launch them before it knows that every
// domain is done, and again after further attempts.
__global__ void k_fin_corner(const Member* __restrict__ mem, int nmem) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    const Member& M = mem[m];
    Ctl& c = *M.ctl;
    if (c.status == ST_RUN || c.fin) return;
    const int N = M.n;
    const double* w = c.w_is_x ? M.X[c.cur] : M.W;
    c.wcorner = 0.25 * (w[0] + w[(size_t)N * N - 1] + w[(size_t)(N - 1) * N] + w[N - 1]);
}

__global__ void __launch_bounds__(PRE_THREADS) k_finalize(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.y];
    const kptr<PhysDev> ph = kview(M.phys);
    const kptr<Ctl> ctl = kview(M.ctl);
    if ((ctl->status == ST_RUN) | (ctl->fin != 0)) return;
    const int N = M.n;
    const size_t nn = (size_t)N * N;
    gptr<double> wsrc = gin<double>(ctl->w_is_x ? M.X[ctl->cur] : M.W);
    const double cw = ctl->wcorner, rho_um = ph->rho_uppermantle;
    gptr<double> w_ref = gin<double>(M.w_ref), wev = gin<double>(M.we), zb_ref = gin<double>(M.z_bed_ref),
                 dz = gin<double>(M.dz_ss);
    gmut<double> W = gout<double>(M.W), cm = gout<double>(M.canom_mantle), zbed = gout<double>(M.z_bed);
    // w_equilibrium exists for methods 0-2; LV-ELVA leaves it zero (the reference checks it too)
    const bool has_weq = M.weq != nullptr;
    gptr<double> weqv = gin<double>(has_weq ? M.weq : M.dz_ss);
    double mw = 0.0, mwe = 0.0, mdz = 0.0, nanf = 0.0, meq = 0.0;
    // every load of the thread's cells first (indices clamped into the field,
    // so no branch), then the arithmetic and the guarded stores: a store
    // between loads would make the next wait cover it too
    constexpr int IT = PRE_CELLS / PRE_THREADS;
    double ws[IT], wr[IT], wea[IT], zr[IT], dzv[IT], eqv[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        c = c < nn ? c : nn - 1;
        ws[it] = wsrc[c]; wr[it] = w_ref[c]; wea[it] = wev[c]; zr[it] = zb_ref[c]; dzv[it] = dz[c];
        eqv[it] = has_weq ? weqv[c] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const size_t c = (size_t)blockIdx.x * PRE_CELLS + it * PRE_THREADS + threadIdx.x;
        if (c >= nn) break;
        const double w = ws[it] - cw;
        W[c] = w;
        cm[c] = 0.0 + rho_um * (w - wr[it]);
        const double we = wea[it];
        const double zb = zr[it] + w + we;
        zbed[c] = zb;
        mw = fmax(mw, fabs(w));
        mwe = fmax(mwe, fabs(we));
        mdz = fmax(mdz, fabs(dzv[it]));
        meq = fmax(meq, fabs(eqv[it]));
        if (zb != zb || w != w) nanf = 1.0;
    }
    __shared__ double red[FIN_P][16];
    mw = wave_max(mw); mwe = wave_max(mwe); mdz = wave_max(mdz); nanf = wave_max(nanf); meq = wave_max(meq);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[0][wv] = mw; red[1][wv] = mwe; red[2][wv] = mdz; red[3][wv] = nanf; red[4][wv] = meq; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < PRE_THREADS / 64; ++q)
            for (int r = 0; r < FIN_P; ++r) red[r][0] = fmax(red[r][0], red[r][q]);
        for (int r = 0; r < FIN_P; ++r) M.partials[(size_t)blockIdx.x * FIN_P + r] = red[r][0];
    }
}

__global__ void __launch_bounds__(64) k_fin_check(const Member* __restrict__ mem, int nblk) {
    const Member& M = mem[blockIdx.x];
    Ctl& c = *M.ctl;
    if (c.status == ST_RUN || c.fin) return;
    double v[FIN_P] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 64)
        for (int r = 0; r < FIN_P; ++r) v[r] = fmax(v[r], M.partials[(size_t)b * FIN_P + r]);
    for (int r = 0; r < FIN_P; ++r) v[r] = wave_max(v[r]);
    if (threadIdx.x == 0) {
        c.maxw = v[0]; c.maxwe = v[1]; c.maxdzss = v[2]; c.maxweq = v[4];
        c.nonfinite = v[3] > 0.0 || c.status == ST_NAN;
        c.w_is_x = 0;
        c.status = ST_IDLE;
        c.fin = 1;
    }
}

__global__ void k_begin(const Member* __restrict__ mem, int nmem, double tf, int stepping, int per_member) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    Ctl& c = *mem[m].ctl;
    if (per_member) tf = mem[m].tf_req;
    c.tf = tf;
    c.k_valid = 0;
    c.pending = 0;
    c.skip = 0;
    c.fin = 0;
    c.status = (stepping && c.t < tf) ? ST_RUN : ST_IDLE;
}

__global__ void k_gather(const Member* __restrict__ mem, int nmem, CtlSummary* __restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    const Ctl& c = *mem[m].ctl;
    CtlSummary s;
    s.t = c.t; s.dt = c.dt; s.tf = c.tf; s.maxw = c.maxw; s.maxwe = c.maxwe; s.maxdzss = c.maxdzss; s.maxweq = c.maxweq;
    s.status = c.status; s.nonfinite = c.nonfinite;
    out[m] = s;
}

void launch_gather(const Member* d_mem, int nmem, CtlSummary* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3((nmem + 63) / 64), dim3(64), 0, st, d_mem, nmem, d_out);
}

// barysealevel.f90:195-252 (const / fastiso / file / mixed), one thread
__global__ void k_bsl_update(BslDev* b, double year_bp, double file_value) {
    b->time = year_bp;
    if (b->method == 0) { b->bsl_now = b->bsl_init; return; }
    // file and mixed: the series value, interpolated by the host (bsl_update_file, :230-238)
    if (b->method == 2 || b->method == 3) b->bsl_now = file_value;
    if (b->method == 2) return;
    if (b->constant_ocean_surface) { b->A_ocean_now = b->A_ocean_pd; return; }
    // interp_0d (isos_utils.f90:423-457)
    const int n = b->nvec;
    const double x = b->bsl_now;
    double y;
    if (x < b->bsl_vec[0]) y = b->A_vec[0];
    else if (x > b->bsl_vec[n - 1]) y = b->A_vec[n - 1];
    else {
        int j;
        for (j = 0; j < n; ++j) if (b->bsl_vec[j] >= x) break;
        if (j == 0) y = b->A_vec[0];
        else if (j == n) y = b->A_vec[n - 1];
        else {
            const double al = (x - b->bsl_vec[j - 1]) / (b->bsl_vec[j] - b->bsl_vec[j - 1]);
            y = b->A_vec[j - 1] + al * (b->A_vec[j] - b->A_vec[j - 1]);
        }
    }
    b->A_ocean_now = y;
}

// isos_init_state (fastisostasy.f90:628-637): ref%z_ss, now%z_ss, Haf and masks
// of both states, after z_bed / Hice / dz_ss have been embedded.
__global__ void __launch_bounds__(PRE_THREADS) k_init_state(const Member* __restrict__ mem) {
    const Member& M = mem[blockIdx.y];
    const PhysDev& ph = *M.phys;
    const Ctl& ctl = *M.ctl;
    const int N = M.n;
    const size_t nn = (size_t)N * N;
    for (int q = threadIdx.x; q < PRE_CELLS; q += PRE_THREADS) {
        const size_t c = (size_t)blockIdx.x * PRE_CELLS + q;
        if (c >= nn) break;
        const double zr = ctl.ref_bsl + 0.0 + M.dz_ss_ref[c];
        M.z_ss_ref[c] = zr;
        const double zn = ctl.bsl + M.z_ss[c] + M.dz_ss[c];
        M.z_ss[c] = zn;
        const double har = calc_haf(zr, M.z_bed_ref[c], M.Hice_ref[c], ph);
        M.Haf_ref[c] = har;
        const double han = calc_haf(zn, M.z_bed[c], M.Hice[c], ph);
        M.Haf[c] = han;
        M.mask_ref[c] = calc_mask(M.z_bed_ref[c], zr, M.Hice_ref[c], har, M.mask_ref[c]);
        M.mask[c] = calc_mask(M.z_bed[c], zn, M.Hice[c], han, M.mask[c]);
    }
}

// ------------------------------------------------------------------ launchers
static int pre_blocks(int n) { return (int)(((size_t)n * n + PRE_CELLS - 1) / PRE_CELLS); }

void launch_pre_haf(const Member* d_mem, int nmem, int n, bool new_H, hipStream_t st) {
    hipLaunchKernelGGL(k_pre_haf, dim3(pre_blocks(n), nmem), dim3(PRE_THREADS), 0, st, d_mem, (int)new_H);
}
void launch_pre_bsl(const Member* d_mem, int nmem, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_pre_bsl, dim3(nmem), dim3(64), 0, st, d_mem, nmem, pre_blocks(n));
}
template <typename T>
static void pre_fields_t(const Member* d_mem, int nmem, int n, int phase, bool diag, hipStream_t st) {
    const dim3 g(pre_blocks(n), nmem), b(PRE_THREADS);
    switch (phase) {
    case 0:
        if (diag) hipLaunchKernelGGL((k_pre_fields<T, PF_DIAG>), g, b, 0, st, d_mem);
        else hipLaunchKernelGGL((k_pre_fields<T, PF_FULL>), g, b, 0, st, d_mem);
        break;
    case 1: hipLaunchKernelGGL((k_pre_fields<T, PF_LITHO>), g, b, 0, st, d_mem); break;
    case 2: hipLaunchKernelGGL((k_pre_fields<T, PF_FULL>), g, b, 0, st, d_mem); break;
    case 3: hipLaunchKernelGGL((k_pre_fields<T, PF_ALL>), g, b, 0, st, d_mem); break;
    case 4: hipLaunchKernelGGL((k_pre_fields<T, PF_NOSSH>), g, b, 0, st, d_mem); break;
    }
}

void launch_pre_fields(const Member* d_mem, int nmem, int n, int phase, bool diag, bool fp64, hipStream_t st) {
    if (fp64) pre_fields_t<double>(d_mem, nmem, n, phase, diag, st);
    else pre_fields_t<float>(d_mem, nmem, n, phase, diag, st);
}
void launch_begin(const Member* d_mem, int nmem, double tf, int stepping, bool per_member, hipStream_t st) {
    hipLaunchKernelGGL(k_begin, dim3((nmem + 63) / 64), dim3(64), 0, st, d_mem, nmem, tf, stepping,
                       per_member ? 1 : 0);
}
void launch_finalize(const Member* d_mem, int nmem, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_fin_corner, dim3((nmem + 63) / 64), dim3(64), 0, st, d_mem, nmem);
    hipLaunchKernelGGL(k_finalize, dim3(pre_blocks(n), nmem), dim3(PRE_THREADS), 0, st, d_mem);
    hipLaunchKernelGGL(k_fin_check, dim3(nmem), dim3(64), 0, st, d_mem, pre_blocks(n));
}
void launch_bsl_update(BslDev* b, double year_bp, double file_value, hipStream_t st) {
    hipLaunchKernelGGL(k_bsl_update, dim3(1), dim3(1), 0, st, b, year_bp, file_value);
}
void launch_init_state_fields(const Member* d_mem, int nmem, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_init_state, dim3(pre_blocks(n), nmem), dim3(PRE_THREADS), 0, st, d_mem);
}

}  // namespace fi
