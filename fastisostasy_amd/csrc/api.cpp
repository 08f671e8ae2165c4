// api.cpp -- host side of the C ABI (include/fastiso.h).
//
// Owns device memory, FFT plans and twiddle tables, and sequences the kernels
// of isos_init / isos_init_state / isos_update (src/fastisostasy.f90) for one
// domain or a batch of domains.  There is no CPU compute path: every field
// operation runs on the GPU, and the library refuses to create a handle when
// no HIP device is present.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/fastiso.h"
#include "fi_kernels.h"

using namespace fi;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

namespace fi {
int set_error(int code, const char* msg) {
    g_err = msg;
    return code;
}
}  // namespace fi

namespace {

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(FI_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

std::mutex g_mu;
hipStream_t g_stream = nullptr;

}  // namespace

// Scratch of the allocation-free update (fi_isos_update_noalloc): one array
// per GPU, held until the process ends.  It carries the per-update working
// memory of one domain -- the launch descriptor, the per-column corner terms,
// the reduction partials, the two transform work arrays and the convolution
// work array -- which fi_isos_update takes from the domain's own arena.  Host
// code never uses the name `scratch` as an address: the device address comes
// from hipGetSymbolAddress, per device.
constexpr size_t FI_NOALLOC_SCRATCH_BYTES = (size_t)13 << 20;   // 13 MiB: N <= 512 in either precision
static __device__ __align__(256) unsigned char scratch[FI_NOALLOC_SCRATCH_BYTES];

namespace {

std::map<int, char*> g_scratch_ptr;   // device ordinal -> address of `scratch` on that device

// Address of `scratch` on the current device (resolved once per device; g_mu held by the caller).
int scratch_address_locked(char** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(FI_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        it = g_scratch_ptr.emplace(dev, (char*)p).first;
    }
    *out = it->second;
    return 0;
}

int ensure_device() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_stream) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(FI_E_HIP, "no HIP device available: the LV-ELVA library has no CPU path");
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // set-up of the allocation-free update on this device
    char* p = nullptr;
    return scratch_address_locked(&p);
}

bool make_plan(int n, FftPlan& p) {
    p.n = n;
    p.nf = 0;
    int m = n;
    const int order[] = {8, 4, 2, 3, 5, 7};
    for (int r : order)
        while (m % r == 0 && p.nf < 24) { p.radix[p.nf++] = r; m /= r; }
    return m == 1 && n >= 4;
}

struct Twiddles {
    double2* d64 = nullptr;
    float2* d32 = nullptr;
};
std::map<int, Twiddles> g_tw;

int get_twiddles(int n, Twiddles& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tw.find(n);
    if (it != g_tw.end()) { out = it->second; return 0; }
    std::vector<double2> h64(n);
    std::vector<float2> h32(n);
    for (int k = 0; k < n; ++k) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        const double c = (double)cosl(a), s = -(double)sinl(a);
        h64[k] = make_double2(c, s);
        h32[k] = make_float2((float)c, (float)s);
    }
    Twiddles t;
    HIPCHK(hipMalloc(&t.d64, sizeof(double2) * n));
    HIPCHK(hipMalloc(&t.d32, sizeof(float2) * n));
    HIPCHK(hipMemcpy(t.d64, h64.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.d32, h32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    g_tw[n] = t;
    out = t;
    return 0;
}

int fft_threads(const FftPlan& p, int lines, int& threads) {
    int nb = 0;
    for (int s = 0; s < p.nf; ++s) nb = std::max(nb, (p.n / p.radix[s]) * lines);
    threads = ((nb / 2 + 63) / 64) * 64;
    threads = std::min(512, std::max(64, threads));
    if (nb > FFT_MAXIT * threads) return fail(FI_E_UNSUPPORTED, "FFT block too large (N=%d, lines=%d)", p.n, lines);
    return 0;
}

// Launch shapes picked by the autotuner (fi_autotune) per (N, precision,
// group size), and the shape under trial while it times candidates (this
// thread only).  0 fields keep the built-in choice.
struct TunedShape {
    int threads_row = 0, C = 0, row_tiles = 0, K = 0;   // power-of-two kernels; K concurrent sub-batches
    int gL = 0, gC = 0;                                 // mixed-radix kernels: line pairs per row / column block
    int srow_rows = 0, threads_inv = 0;                 // register kernels: fused-pass rows, inverse-pass threads
};

// Rows per block of the fused stencil pass with one column per thread
// (FI_SROW_ROWS = 4 | 8, or the tuned value; 0 keeps 2L rows): more,
// shorter row blocks for small batches.  Block size of the inverse row pass
// (FI_INV_THREADS, or the tuned value; 0 keeps threads_row).  Untuned, a
// group whose 16-row fused blocks or default inverse blocks would not give
// every CU one (256) takes 4-row (2-row) fused blocks and 64-thread inverse
// blocks, the shapes the autotuner picks for a single 256^2 domain (0.95 ms
// per 10-yr update against 1.16); larger groups whose row blocks hold fewer
// threads than columns (FP64 at N = 512) take 16-row fused blocks of N threads.
void small_batch_shape(RhsShape& s, int nmem, bool tuned, const TunedShape& ts, bool one_col_ok, int def_inv,
                       const std::function<bool(int)>& inv_ok) {
    s.srow_rows = 0;
    s.threads_inv = 0;
    // fewer blocks than CUs: fused blocks of 16 (or 4) rows, inverse blocks of threads_row
    const bool few_rows = (long long)(s.n / 16) * nmem < 256, fewer_rows = (long long)(s.n / 4) * nmem < 256;
    const bool few = (long long)(s.n / (2 * s.L)) * nmem < 256;
    if (s.fused && one_col_ok) {
        // (2-row blocks while 4-row ones would still leave CUs idle: one 256^2
        // domain's fused pass 6.25 us against 6.92; 16 rows in N threads where
        // the row blocks hold fewer threads than columns: one column per thread
        // keeps the march's window in registers)
        int r = tuned && ts.srow_rows ? ts.srow_rows
                                      : (fewer_rows ? 2 : few_rows ? 4 : (s.threads_row < s.n ? 16 : 0));
        if (const char* e = std::getenv("FI_SROW_ROWS")) r = std::atoi(e);
        if (r == 2 || r == 4 || r == 8 || r == 16) s.srow_rows = r;
    }
    int th = tuned ? ts.threads_inv : (few ? def_inv : 0);
    if (const char* e = std::getenv("FI_INV_THREADS")) th = std::atoi(e);
    if (th && th != s.threads_row && inv_ok(th)) s.threads_inv = th;
}
std::map<std::tuple<int, int, int>, TunedShape> g_tuned;
thread_local const TunedShape* t_trial = nullptr;

bool tuned_for(int n, bool fp64, int nmem, TunedShape& out) {
    if (t_trial) { out = *t_trial; return true; }
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(std::make_tuple(n, (int)fp64, nmem));
    if (it == g_tuned.end()) return false;
    out = it->second;
    return true;
}

// Row / column blocking for a batch of `nmem` domains of size n.
int make_shape(int n, int nmem, bool fp64, bool uniform, RhsShape& s) {
    TunedShape ts;
    const bool tuned = tuned_for(n, fp64, nmem, ts);
    if (!make_plan(n, s.plan))
        return fail(FI_E_UNSUPPORTED, "padded size N=%d has a prime factor > 7; choose pad so that N is 2^a 3^b 5^c 7^d", n);
    s.n = n;
    s.fp64 = fp64;
    s.uniform = uniform;
    s.method = 3;
    s.dt_method = -1;
    const bool pow2 = (n & (n - 1)) == 0 && n >= 64 && n <= 1024;
    const bool r3 = n == 192 || n == 384;   // 3 * 2^k: register kernels with 384-thread row blocks
    s.fast = (pow2 || r3) && !std::getenv("FI_GENERIC_FFT");
    if (s.fast && r3) {
        // one row-block shape (384 threads); column blocks as for 2^k
        const int tpl = n / 8;
        const size_t cb = fp64 ? 16 : 8;
        const size_t pitch = n + n / 8 + 1;
        const int L = 384 / tpl;
        const bool rows = (n + L * pitch) * cb <= 65536 && n % (2 * L) == 0;
        auto col_ok = [&](int C) {
            return C * tpl <= 1024 && ((size_t)n + (size_t)n * (C + 1)) * cb <= 65536 && n % (2 * C) == 0;
        };
        int best_c = 0;
        // the largest blocks that still give every CU one; no block below a
        // wave when a wave-sized one exists (one 256^2 domain: 2 line pairs)
        for (int C : {16, 8, 4, 2, 1}) {
            if (!col_ok(C)) continue;
            if (best_c && C * tpl < 64) break;
            best_c = C;
            if ((size_t)(n / (2 * C)) * nmem >= 256) break;
        }
        if (tuned && ts.C && col_ok(ts.C)) best_c = ts.C;
        if (const char* e = std::getenv("FI_COL_LINES")) {
            const int C = std::atoi(e);
            if ((C == 1 || C == 2 || C == 4 || C == 8 || C == 16) && col_ok(C)) best_c = C;
        }
        if (rows && best_c) {
            s.fused = uniform && !std::getenv("FI_NO_FUSED_STENCIL");   // fused at N = 384 (one column per thread)
            s.row_tiles = 1;
            s.threads_row = 384;
            s.L = L;
            s.C = best_c;
            s.threads_col = best_c * tpl;
            // (inverse row blocks of 384 or 192 threads: the corner sums take whole waves)
            small_batch_shape(s, nmem, tuned, ts, n == 384, 192, [](int th) { return th == 192; });
            Twiddles t;
            int rc = get_twiddles(n, t);
            if (rc) return rc;
            s.tw = fp64 ? (const void*)t.d64 : (const void*)t.d32;
            s.tw64 = t.d64;
            return 0;
        }
        s.fast = false;
    }
    if (s.fast) {
        // register-FFT kernels: TPL = n/8 threads per line, LDS <= 64 KiB
        const int tpl = n / 8;
        const size_t cb = fp64 ? 16 : 8;
        const size_t pitch = n + n / 8 + 1;
        int best_t = 0;
        auto row_ok = [&](int th) {
            if (th < tpl) return false;
            const int L = th / tpl;
            return (n + L * pitch) * cb <= 65536 && n % (2 * L) == 0;
        };
        for (int th : {256, 128, 64}) {
            if (!row_ok(th)) continue;
            best_t = th;
            if ((size_t)(n / (2 * th / tpl)) * nmem >= 512) break;
        }
        // The fused stencil pass holds its row window in registers per owned
        // column, so several columns per thread spill; with one column per
        // thread (THREADS = N, possible up to N = 256, and N = 512 in the
        // mixed mode with 68 KiB of LDS per block) even a single domain
        // runs 6-15% faster per update than with more, smaller blocks.
        if (uniform && (n <= 256 || (n == 512 && !fp64)) && row_ok(n)) best_t = n;
        // FI_ROW_THREADS = 64 | 128 | 256 (| 512 for mixed N = 512) overrides the row block size (tuning, tests)
        if (const char* e = std::getenv("FI_ROW_THREADS")) {
            const int th = std::atoi(e);
            if ((th == 64 || th == 128 || th == 256 || (th == 512 && n == 512 && !fp64)) && row_ok(th)) best_t = th;
        }
        int best_c = 0;
        auto col_ok = [&](int C) {
            return C * tpl <= 1024 && ((size_t)n + (size_t)n * (C + 1)) * cb <= 65536 && n % (2 * C) == 0;
        };
        // the largest blocks that still give every CU one; no block below a
        // wave when a wave-sized one exists (one 256^2 domain: 2 line pairs)
        for (int C : {16, 8, 4, 2, 1}) {
            if (!col_ok(C)) continue;
            if (best_c && C * tpl < 64) break;
            best_c = C;
            if ((size_t)(n / (2 * C)) * nmem >= 256) break;
        }
        // FI_COL_LINES = 1 | 2 | 4 | 8 | 16 overrides the column block (tuning, tests)
        if (const char* e = std::getenv("FI_COL_LINES")) {
            const int C = std::atoi(e);
            if ((C == 1 || C == 2 || C == 4 || C == 8 || C == 16) && col_ok(C)) best_c = C;
        }
        if (tuned && ts.threads_row && row_ok(ts.threads_row)) best_t = ts.threads_row;
        if (tuned && ts.C && col_ok(ts.C)) best_c = ts.C;
        if (best_t && best_c) {
            s.fused = uniform && !std::getenv("FI_NO_FUSED_STENCIL");
            s.row_tiles = 1;
            if (const char* e = std::getenv("FI_ROW_TILES")) s.row_tiles = std::atoi(e) == 2 ? 2 : 1;
            if (tuned && (ts.row_tiles == 1 || ts.row_tiles == 2) && n % (4 * (best_t / tpl)) == 0)
                s.row_tiles = ts.row_tiles;
            s.threads_row = best_t;
            s.L = best_t / tpl;
            s.C = best_c;
            s.threads_col = best_c * tpl;
            // one column per thread in the fused pass: N threads (its LDS is
            // checked where it is launched; beyond 64 KiB it is declared)
            const bool one_col = n <= 512;
            small_batch_shape(s, nmem, tuned, ts, one_col, std::max(64, n / 8), [&](int th) {
                return (th == 64 || th == 128 || th == 256 || (th == 512 && n == 512 && !fp64)) && row_ok(th) &&
                       (s.row_tiles == 1 || n % (4 * (th / tpl)) == 0);
            });
            Twiddles t;
            int rc = get_twiddles(n, t);
            if (rc) return rc;
            s.tw = fp64 ? (const void*)t.d64 : (const void*)t.d32;
            s.tw64 = t.d64;
            return 0;
        }
        s.fast = false;
    }
    // Generic kernels: the error partials are per row and the corner terms per
    // column, summed in an order set by N alone, so a domain's results do not
    // depend on the blocking or on the batch it is launched in.  LDS per block
    // stays within 64 KiB.
    const size_t cbytes = fp64 ? 16 : 8;
    // data lines plus the twiddle table in LDS
    auto fits = [&](int lines) { return (size_t)n * (lines + 1) * cbytes <= 65536; };
    if (!fits(1)) return fail(FI_E_UNSUPPORTED, "padded size N=%d too large for one LDS line", n);
    int nbmax = 0;   // butterflies of the largest stage of one line
    for (int q = 0; q < s.plan.nf; ++q) nbmax = std::max(nbmax, n / s.plan.radix[q]);
    auto ok = [&](int lines) { return fits(lines) && nbmax * lines <= FFT_MAXIT * 512; };
    // Line pairs per block: the largest blocks that still give the batch at
    // least 96 row blocks and 48 column blocks (from N = 256 on; smaller N
    // keep 8 and 4 per domain).  One N = 384 domain then runs with L = 2,
    // C = 4, 18% faster per update than with 8 and 16, while large groups keep
    // the large blocks (16% more requests/s in the batched serving mix).
    int rb = 8, cb = 4, lmax = 8, cmax = fp64 ? 8 : 16;
    if (n >= 256) {
        rb = (int)std::max<long long>(1, (96 + nmem - 1) / nmem);
        cb = (int)std::max<long long>(1, (48 + nmem - 1) / nmem);
    }
    if (tuned && ts.gL >= 1 && ts.gC >= 1) {
        lmax = std::min(lmax, ts.gL);
        cmax = std::min(cmax, ts.gC);
        rb = cb = 1;
    }
    // FI_GEN_LINES = "L,C" asks for L / C line pairs per row / column block
    // where the LDS and the divisibility allow them (tuning)
    if (const char* e = std::getenv("FI_GEN_LINES")) {
        int l2 = 0, c2 = 0;
        if (std::sscanf(e, "%d,%d", &l2, &c2) == 2 && l2 >= 1 && c2 >= 1) {
            lmax = std::min(lmax, l2);
            cmax = std::min(cmax, c2);
            rb = cb = 1;
        }
    }
    int L = 1;
    while (L < lmax && (n / 2) % (2 * L) == 0 && ok(2 * L) && n / (4 * L) >= rb) L *= 2;
    int C = 1;
    while (C < cmax && (n / 2) % (2 * C) == 0 && ok(2 * C) && n / (4 * C) >= cb) C *= 2;
    s.L = L;
    s.C = C;
    int rc = fft_threads(s.plan, L, s.threads_row);
    if (rc) return rc;
    rc = fft_threads(s.plan, C, s.threads_col);
    if (rc) return rc;
    Twiddles t;
    rc = get_twiddles(n, t);
    if (rc) return rc;
    s.tw = fp64 ? (const void*)t.d64 : (const void*)t.d32;
    s.tw64 = t.d64;
    return 0;
}

int make_conv_shape(int n, bool fp64, ConvShape& s) {
    s.n = n;
    s.m = 2 * n;
    s.fp64 = fp64;
    if (!make_plan(s.m, s.plan)) return fail(FI_E_UNSUPPORTED, "convolution size 2N=%d not supported", s.m);
    const size_t cbytes = fp64 ? 16 : 8;
    auto fits = [&](int lines) { return (size_t)s.m * lines * cbytes <= 65536; };
    if (!fits(1)) return fail(FI_E_UNSUPPORTED, "convolution size 2N=%d too large for one LDS line", s.m);
    // lines per block: LDS-bounded and at most FFT_MAXIT * 512 butterflies per stage
    int nbmax = 0;
    for (int q = 0; q < s.plan.nf; ++q) nbmax = std::max(nbmax, s.m / s.plan.radix[q]);
    auto ok = [&](int lines) { return fits(lines) && nbmax * lines <= FFT_MAXIT * 512; };
    int L = 1;
    while (L < 4 && n % (4 * L) == 0 && ok(2 * L)) L *= 2;
    // the column pass's lines are independent (its blocking does not change
    // results): room for the twiddle table in LDS besides the lines; the row
    // passes pair rows r0 + l and r0 + L + l, so L stays as is and they hold
    // the table in LDS only where it fits
    int C = 1;
    while (C < 8 && ok(2 * C) && fits(2 * C + 1)) C *= 2;
    s.Lr = L;
    s.Cc = C;
    s.rows_twl = fits(L + 1);
    int rc = fft_threads(s.plan, L, s.threads_row);
    if (rc) return rc;
    rc = fft_threads(s.plan, C, s.threads_col);
    if (rc) return rc;
    Twiddles t;
    rc = get_twiddles(s.m, t);
    if (rc) return rc;
    s.tw = fp64 ? (const void*)t.d64 : (const void*)t.d32;
    return 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Summary buffer of a launch group: pinned, coherent host memory that the
// gather kernel stores to directly through its device address `d`, so a
// round's readback is that kernel alone (no blit copy queued behind it).
struct SumBuf {
    CtlSummary* d = nullptr;
    CtlSummary* h = nullptr;
    int n = 0;
};

// ------------------------------------------------------------------ handles
struct fi_bsl {
    BslDev* d = nullptr;
    BslDev h{};
    std::vector<double> st, sv;   // sea-level series of the methods file and mixed (time yr, bsl m)
};

struct fi_isos {
    fi_params par{};
    int n = 0, nx_ice = 0, ny_ice = 0, icrop1 = 0, jcrop1 = 0;  // 0-based crop origin
    double dx = 0;
    bool fp64 = true, uniform = true;
    bool ref_was_set = false, state_init = false;
    bool use_restart = false;    // a restart was read: the next init_state takes the restart branch
    double time_diag = 0, time_prog = 0;
    double last_t = 0, last_dt = 0;  // ode%t and ode%dt after the last update (host copy; 0 = unknown)
    int last_status = 0;         // status of the last update's instability and NaN checks
    char* arena = nullptr;
    size_t arena_bytes = 0;      // HBM footprint of the domain (one allocation)
    Member hm{};
    Member* d_mem = nullptr;
    double *eta_eff = nullptr, *kappa = nullptr, *R = nullptr, *He = nullptr, *Dd = nullptr;
    RhsShape shape{};
    ConvShape cshape{};
    bool has_conv = false;
    int part_blocks = 0;
    fi_bsl* bsl_bound = nullptr;
    SumBuf sb;
};

struct fi_batch {
    std::vector<fi_isos*> mem;
    Member* d_mem = nullptr;
    std::vector<Member> hm;
    RhsShape shape{};
    SumBuf sb;
};

static const ConvShape& conv_shape_of(const fi_isos* h) { return h->cshape; }

namespace {

int domain_size(int nx_ice, int ny_ice, double dx, double dy, double min_pad, int& n, int& icrop1, int& jcrop1) {
    // init_domain_size (src/isos_utils.f90:285-354); crop origin 0-based
    if (min_pad < 0.0) return fail(FI_E_ARG, "Minimum padding must be positive.");
    if (dx != dy) return fail(FI_E_ARG, "Currently, only square grid cells are supported.");
    if (nx_ice < 3 || ny_ice < 3) return fail(FI_E_ARG, "ice grid too small");
    int ic = 1, jc = 1, nxt = nx_ice, nyt = ny_ice;
    if (nx_ice % 2) { ic++; nxt++; }
    if (ny_ice % 2) { jc++; nyt++; }
    const int npad = (int)std::lround(min_pad / dx);
    int npx = npad, npy = npad;
    if (nxt < nyt) npx += (nyt - nxt) / 2;
    else if (nxt > nyt) npy += (nxt - nyt) / 2;
    n = nxt + 2 * npx;
    if (n != nyt + 2 * npy) return fail(FI_E_ARG, "Failed to correctly pad domain.");
    icrop1 = ic + npx - 1;
    jcrop1 = jc + npy - 1;
    return 0;
}

// Upload an ice-grid array embedded in an N x N background.
int upload_embedded(const fi_isos* h, double* dst, const double* src, double background) {
    const int n = h->n;
    std::vector<double> buf((size_t)n * n, background);
    for (int j = 0; j < h->ny_ice; ++j)
        for (int i = 0; i < h->nx_ice; ++i)
            buf[(size_t)(j + h->jcrop1) * n + (i + h->icrop1)] = src[(size_t)j * h->nx_ice + i];
    HIPCHK(hipMemcpyAsync(dst, buf.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int read_ctl(const Member& m, Ctl& c) {
    HIPCHK(hipMemcpyAsync(&c, m.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int write_ctl(const Member& m, const Ctl& c) {
    HIPCHK(hipMemcpyAsync(m.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int sum_alloc(SumBuf& b, int n) {
    HIPCHK(hipHostMalloc(&b.h, sizeof(CtlSummary) * n, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&b.d, b.h, 0));
    b.n = n;
    return 0;
}
void sum_free(SumBuf& b) {
    if (b.h) (void)hipHostFree(b.h);
    b.d = nullptr;
    b.h = nullptr;
}

int gather(Member* d_mem, int nmem, SumBuf& sb) {
    launch_gather(d_mem, nmem, sb.d, g_stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

// Concurrent sub-batches of the attempt loop (at most three streams besides
// the library stream, default priority).  One launch of an attempt kernel
// over a large batch is a single occupancy round in which all blocks load,
// then compute, in step; sub-batches on separate streams drift apart, so one
// sub-batch's transforms overlap another's loads.  Results do not depend on
// the split (reductions are fixed by N).
constexpr int MAX_SUB = 3;
struct SubStreams {
    hipStream_t s[MAX_SUB] = {};
    hipEvent_t done[MAX_SUB] = {};
    hipEvent_t start = nullptr;
    int n = 0;
};
SubStreams g_sub;

int ensure_sub(int k) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_sub.start) HIPCHK(hipEventCreateWithFlags(&g_sub.start, hipEventDisableTiming));
    for (; g_sub.n < k; ++g_sub.n) {
        HIPCHK(hipStreamCreateWithFlags(&g_sub.s[g_sub.n], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g_sub.done[g_sub.n], hipEventDisableTiming));
    }
    return 0;
}

// number of sub-batches for a batch of nmem domains (the autotuner's choice,
// else 2 from 32 domains; FI_SUBBATCHES overrides)
int sub_count(int n, bool fp64, int nmem) {
    int k = nmem >= 32 ? 2 : 1;
    TunedShape ts;
    if (tuned_for(n, fp64, nmem, ts) && ts.K >= 1) k = ts.K;
    if (const char* e = std::getenv("FI_SUBBATCHES")) k = std::atoi(e);
    return std::max(1, std::min({k, MAX_SUB, nmem}));
}

std::map<std::tuple<int, int, int, int, int>, RhsShape> g_sub_shapes;

thread_local RhsShape t_trial_sub[MAX_SUB];

int sub_shape(int n, int nsub, bool fp64, bool uniform, int dt_method, const RhsShape*& out, int k) {
    if (t_trial) {
        // a shape under trial is not cached
        RhsShape& s = t_trial_sub[k];
        s = RhsShape{};
        int rc = make_shape(n, nsub, fp64, uniform, s);
        if (rc) return rc;
        s.dt_method = dt_method;
        out = &s;
        return 0;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(n, nsub, (int)fp64, (int)uniform, dt_method);
    auto it = g_sub_shapes.find(key);
    if (it == g_sub_shapes.end()) {
        RhsShape s{};
        g_mu.unlock();
        int rc = make_shape(n, nsub, fp64, uniform, s);
        g_mu.lock();
        if (rc) return rc;
        s.dt_method = dt_method;
        it = g_sub_shapes.emplace(key, s).first;
    }
    out = &it->second;
    return 0;
}

// Attempt graphs: a block of GRAPH_REPS attempted steps captured once per
// (descriptor array, batch size, shape, stream) and replayed with one
// hipGraphLaunch, so the host issues one call instead of 4 x GRAPH_REPS
// kernel launches.  Kernel nodes only (no memset or copy nodes); the kernels
// read the descriptors and the control block at run time, so one graph
// serves every update of a batch.  Opt-in for batches and single domains
// (FI_GRAPHS=1): measured on the 64-member ensemble they are neutral (the
// GPU, not the host, bounds the attempt rate) and rocprofv3's kernel tracing
// crashed the host process inside a graph launch.  Benchmark kernel timing
// and round-based serving (whose groupings change every round) never use
// them; polled serving does for groups of one or two domains, where the host's
// launches would otherwise set the pace (FI_SERVER_GRAPHS=0 turns that off,
// e.g. under a kernel tracer).
constexpr int GRAPH_REPS = 8, GRAPH_CACHE = 128;
struct GraphEntry {
    const Member* d;
    int nmem;
    RhsShape s;
    hipStream_t st;
    hipGraphExec_t exec;
    unsigned long long use;
};
std::vector<GraphEntry> g_graphs;
unsigned long long g_graph_clock = 0;
std::mutex g_graph_mu;   // the cache is shared by every thread's batches and servers

bool same_shape(const RhsShape& a, const RhsShape& b) {
    return a.n == b.n && a.L == b.L && a.C == b.C && a.threads_row == b.threads_row &&
           a.threads_col == b.threads_col && a.fp64 == b.fp64 && a.uniform == b.uniform && a.fast == b.fast &&
           a.fused == b.fused && a.row_tiles == b.row_tiles && a.srow_rows == b.srow_rows &&
           a.threads_inv == b.threads_inv && a.tw == b.tw && a.tw64 == b.tw64 &&
           a.plan.n == b.plan.n && a.plan.nf == b.plan.nf && a.method == b.method && a.dt_method == b.dt_method;
}

int attempt_graph(const Member* d, int nmem, const RhsShape& s, hipStream_t st, hipGraphExec_t* out) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    ++g_graph_clock;
    for (auto& e : g_graphs)
        if (e.d == d && e.nmem == nmem && e.st == st && same_shape(e.s, s)) {
            e.use = g_graph_clock;
            *out = e.exec;
            return 0;
        }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < GRAPH_REPS; ++r) launch_attempt(d, nmem, s, st);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(FI_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    if ((int)g_graphs.size() >= GRAPH_CACHE) {
        auto lru = std::min_element(g_graphs.begin(), g_graphs.end(),
                                    [](const GraphEntry& a, const GraphEntry& b) { return a.use < b.use; });
        // a launch of the evicted exec may still be queued: let its stream drain first
        (void)hipStreamSynchronize(lru->st);
        (void)hipGraphExecDestroy(lru->exec);
        g_graphs.erase(lru);
    }
    g_graphs.push_back({d, nmem, s, st, ex, g_graph_clock});
    *out = ex;
    return 0;
}

// graphs over descriptor arrays in [lo, hi) (before that memory is freed)
void purge_graphs(const Member* lo, const Member* hi) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (size_t i = 0; i < g_graphs.size();) {
        if (g_graphs[i].d >= lo && g_graphs[i].d < hi) {
            (void)hipGraphExecDestroy(g_graphs[i].exec);
            g_graphs.erase(g_graphs.begin() + i);
        } else {
            ++i;
        }
    }
}

// `count` attempts: whole graphs, then the remainder as plain launches
// graphs: 0 plain launches, 1 attempt graphs when FI_GRAPHS is set, 2 attempt graphs
int launch_attempts(const Member* d, int nmem, const RhsShape& s, hipStream_t st, int count, int graphs) {
    int done = 0;
    const bool use = graphs == 2 || (graphs == 1 && std::getenv("FI_GRAPHS"));
    if (use && count >= GRAPH_REPS && !prof_active()) {
        hipGraphExec_t ex;
        int rc = attempt_graph(d, nmem, s, st, &ex);
        if (rc) return rc;
        for (; done + GRAPH_REPS <= count; done += GRAPH_REPS) HIPCHK(hipGraphLaunch(ex, st));
    }
    for (; done < count; ++done) launch_attempt(d, nmem, s, st);
    return 0;
}

// Diagnostics cadence of one isos_update call (fastisostasy.f90:782-787): a
// function of the call times only.
bool diag_due(const fi_isos* h, double time) { return time - h->time_diag >= (double)1e-3f; }

// One isos_update for a group of members that share N, precision and options,
// as a job bound to one stream: job_begin (the pre-update kernels), rounds of
// attempts each ending in the finalize and an asynchronous readback of the
// per-domain summaries (job_round), the host's look at the summaries
// (job_more), job_end (the instability checks).  All members advance to
// `time`, or member m to times[m] when `times` is given (the Member copies on
// the device then carry the targets in tf_req).  run_update drives one job on
// the library stream; the request server drives several on their own streams.
struct UpdateJob {
    std::vector<fi_isos*> hs;
    Member* d_mem = nullptr;
    const RhsShape* shape = nullptr;
    SumBuf sb;
    const double* times = nullptr;
    double time = 0.0;
    bool stepping = true;
    int graphs = 1;              // launch_attempts: 0 plain, 1 graphs under FI_GRAPHS, 2 graphs
    bool subbatch = true;        // may split into concurrent sub-batches (library stream only)
    hipStream_t st = nullptr;
    int K = 1, ahead = 1;
    long long total = 0;
    int sb_lo[MAX_SUB + 1] = {};
    const RhsShape* sshape[MAX_SUB] = {};
    int failed_member = -1;      // set by job_end: the first member that failed its checks
    double t_of(int m) const { return times ? times[m] : time; }
};

// Attempts launched beyond the first round's estimate: each costs a few
// microseconds of gated no-op kernels when unused, a missing one a host round
// trip.  One beats two (0.4%) and none (a rejection then forces a second
// round) on the 64-member ensemble.  FI_MARGIN overrides (tuning).
int first_round_margin() {
    static const int m = [] {
        const char* e = std::getenv("FI_MARGIN");
        return e ? std::max(0, std::atoi(e)) : 1;
    }();
    return m;
}

int job_begin(UpdateJob& j) {
    const int nmem = (int)j.hs.size();
    const fi_isos& h0 = *j.hs[0];
    const int n = h0.n;
    const RhsShape& shape = *j.shape;
    Member* d_mem = j.d_mem;
    hipStream_t st = j.st;
    const bool diag = diag_due(j.hs[0], j.t_of(0));
    for (int m = 1; m < nmem; ++m)
        if (diag_due(j.hs[m], j.t_of(m)) != diag) return fail(FI_E_STATE, "batch members disagree on the diagnostics cadence");
    if (diag)
        for (fi_isos* h : j.hs) h->time_diag = h->time_diag + h->par.dt_diagnostics;
    launch_pre_haf(d_mem, nmem, n, true, st);
    launch_pre_bsl(d_mem, nmem, n, st);
    if (!diag) {
        launch_pre_fields(d_mem, nmem, n, 0, false, h0.fp64, st);
    } else if (!h0.par.include_elastic && !h0.par.heterogeneous_ssh) {
        launch_pre_fields(d_mem, nmem, n, 3, true, h0.fp64, st);   // phases 0-2 in one pass
    } else {
        launch_pre_fields(d_mem, nmem, n, 0, true, h0.fp64, st);
        if (h0.par.include_elastic) launch_convolution(d_mem, nmem, conv_shape_of(&h0), 0, st);
        if (h0.par.heterogeneous_ssh) {
            launch_pre_fields(d_mem, nmem, n, 1, true, h0.fp64, st);
            launch_convolution(d_mem, nmem, conv_shape_of(&h0), 1, st);
            launch_pre_fields(d_mem, nmem, n, 2, true, h0.fp64, st);
        } else {
            launch_pre_fields(d_mem, nmem, n, 4, true, h0.fp64, st);   // phases 1-2 in one pass
        }
    }
    launch_begin(d_mem, nmem, j.time, j.stepping ? 1 : 0, j.times != nullptr, st);
    if (j.stepping && shape.method != 3) {
        if (shape.method == 2) {
            launch_weq_input(d_mem, nmem, n, st);
            launch_convolution(d_mem, nmem, conv_shape_of(&h0), 2, st);
        }
        launch_weq(d_mem, nmem, n, shape.method, st);
    }
    HIPCHK(hipGetLastError());
    if (!j.stepping) return 0;
    // Attempts until every member reached `time`.  Members that are done skip
    // the launches on the device, so the host may launch ahead: the first
    // round launches as many attempts as the slowest member needs at the dt it
    // ended its previous update with, plus a margin for rejections; after each
    // readback job_more corrects the estimate.
    double need = 0.0;
    for (int m = 0; m < nmem; ++m)
        if (j.hs[m]->last_dt > 0.0) need = std::max(need, std::ceil((j.t_of(m) - j.hs[m]->last_t) / j.hs[m]->last_dt));
    if (need > 0.0) need += (double)first_round_margin();
    j.ahead = (int)std::min(256.0, std::max(1.0, need));
    // concurrent sub-batches only on the library stream (their streams are shared)
    j.K = shape.method == 3 && j.st == g_stream && j.subbatch ? sub_count(n, h0.fp64, nmem) : 1;
    if (j.K > 1) {
        int rc = ensure_sub(j.K);
        if (rc) return rc;
        for (int k = 0; k <= j.K; ++k) j.sb_lo[k] = (int)((long long)nmem * k / j.K);
        for (int k = 0; k < j.K; ++k) {
            rc = sub_shape(n, j.sb_lo[k + 1] - j.sb_lo[k], h0.fp64, h0.uniform, shape.dt_method, j.sshape[k], k);
            if (rc) return rc;
        }
    }
    return 0;
}

// One round: the attempts (when stepping) and the decision of the last one,
// then the finalize of the members that are done (gated per member on the
// device, so that usually no second round follows) and the asynchronous
// readback of the summaries, all on the job's stream.
int job_round(UpdateJob& j) {
    const int nmem = (int)j.hs.size();
    const int n = j.hs[0]->n;
    const RhsShape& shape = *j.shape;
    Member* d_mem = j.d_mem;
    if (j.stepping) {
        if (j.K == 1) {
            int rc = launch_attempts(d_mem, nmem, shape, j.st, j.ahead, j.graphs);
            if (rc) return rc;
            launch_decide(d_mem, nmem, shape, j.st);
        } else {
            HIPCHK(hipEventRecord(g_sub.start, j.st));
            for (int k = 0; k < j.K; ++k) HIPCHK(hipStreamWaitEvent(g_sub.s[k], g_sub.start, 0));
            // interleave the sub-batches' launches so both queues fill together
            for (int a = 0; a < j.ahead; a += GRAPH_REPS)
                for (int k = 0; k < j.K; ++k) {
                    int rc = launch_attempts(d_mem + j.sb_lo[k], j.sb_lo[k + 1] - j.sb_lo[k], *j.sshape[k], g_sub.s[k],
                                             std::min(GRAPH_REPS, j.ahead - a), j.graphs);
                    if (rc) return rc;
                }
            for (int k = 0; k < j.K; ++k) {
                launch_decide(d_mem + j.sb_lo[k], j.sb_lo[k + 1] - j.sb_lo[k], *j.sshape[k], g_sub.s[k]);
                HIPCHK(hipEventRecord(g_sub.done[k], g_sub.s[k]));
                HIPCHK(hipStreamWaitEvent(j.st, g_sub.done[k], 0));
            }
        }
        j.total += j.ahead;
    }
    launch_finalize(d_mem, nmem, n, j.st);
    HIPCHK(hipGetLastError());
    launch_gather(d_mem, nmem, j.sb.d, j.st);   // stores into pinned host memory
    HIPCHK(hipGetLastError());
    return 0;
}

// After a round's readback has completed: whether another round is needed
// (as many attempts as the slowest running member needs at its current dt).
int job_more(UpdateJob& j, bool& more) {
    more = false;
    if (!j.stepping) return 0;
    double need = 0.0;
    bool running = false;
    for (int m = 0; m < (int)j.hs.size(); ++m) {
        const CtlSummary& c = j.sb.h[m];
        if (c.status != ST_RUN) continue;
        // rejections halve dt; it reaches zero only if every attempt is rejected
        if (!(c.dt > 0.0)) return fail(FI_E_STATE, "member %d: the step size collapsed to zero", m);
        running = true;
        need = std::max(need, std::ceil((c.tf - c.t) / std::max(c.dt, 1e-300)));
    }
    if (!running) return 0;
    j.ahead = (int)std::min(256.0, std::max(1.0, need + 1.0));
    if (j.total > 100000000LL) return fail(FI_E_STATE, "integrator did not reach the target time");
    more = true;
    return 0;
}

// check_isos_instabilities (fastisostasy.f90:876-884: w, we, w_equilibrium,
// dz_ss, in that order) and the NaN check of z_bed (:864-869) of one domain:
// 0, FI_E_UNSTABLE or FI_E_NAN, with the array named in `what`.
int instability_status(const CtlSummary& c, const char*& what) {
    if (c.maxw > 1e5) { what = "large_check:: Error: Large values detected in array w"; return FI_E_UNSTABLE; }
    if (c.maxwe > 1e5) { what = "large_check:: Error: Large values detected in array we"; return FI_E_UNSTABLE; }
    if (c.maxweq > 1e5) {
        what = "large_check:: Error: Large values detected in array w_equilibrium";
        return FI_E_UNSTABLE;
    }
    if (c.maxdzss > 1e3) { what = "large_check:: Error: Large values detected in array dz_ss"; return FI_E_UNSTABLE; }
    if (c.nonfinite) { what = "isos_update:: Error: NaNs detected."; return FI_E_NAN; }
    what = "";
    return 0;
}

// The end of an update: every member's times advance (its state has been
// stepped on the device, failed or not) and keeps its own status
// (FI_S_STATUS); the first failing member in batch order names the error.
int job_end(UpdateJob& j) {
    int first = 0;
    const int nmem = (int)j.hs.size();
    for (int m = 0; m < nmem; ++m) {
        const CtlSummary& c = j.sb.h[m];
        fi_isos* h = j.hs[m];
        h->time_prog = j.t_of(m);
        h->last_t = c.t;
        h->last_dt = j.stepping ? c.dt : 0.0;
        const char* what;
        h->last_status = instability_status(c, what);
        if (h->last_status && !first) {
            first = nmem > 1 ? fail(h->last_status, "member %d: %s", m, what) : fail(h->last_status, "%s", what);
            j.failed_member = m;
        }
    }
    return first;
}

int run_update(std::vector<fi_isos*>& hs, Member* d_mem, const RhsShape& shape, SumBuf& sb, double time,
               bool stepping, const double* times = nullptr, int graphs = 1) {
    UpdateJob j;
    j.hs = hs;
    j.d_mem = d_mem;
    j.shape = &shape;
    j.sb = sb;
    j.times = times;
    j.time = time;
    j.stepping = stepping;
    j.graphs = graphs;
    j.st = g_stream;
    int rc = job_begin(j);
    if (rc) return rc;
    for (;;) {
        rc = job_round(j);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(j.st));
        bool more = false;
        rc = job_more(j, more);
        if (rc) return rc;
        if (!more) break;
    }
    return job_end(j);
}

int bind_bsl(fi_isos* h, fi_bsl* b) {
    if (!b) return fail(FI_E_ARG, "bsl is NULL");
    if (h->bsl_bound != b) {
        h->hm.bsl = b->d;
        h->bsl_bound = b;
        HIPCHK(hipMemcpy(h->d_mem, &h->hm, sizeof(Member), hipMemcpyHostToDevice));
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* fi_last_error(void) { return g_err.c_str(); }
int fi_version(void) { return 1; }

int fi_padded_size(int nx_ice, int ny_ice, double dx, double dy, double min_pad, int* n_out) {
    int n, ic, jc;
    int rc = domain_size(nx_ice, ny_ice, dx, dy, min_pad, n, ic, jc);
    if (rc) return rc;
    *n_out = n;
    return 0;
}

int fi_sync(void) {
    int rc = ensure_device();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

namespace {

// interp_0d (isos_utils.f90:423-457) and interp_linear (barysealevel.f90:370-400):
// the same rule -- clamp outside [x(1), x(n)], else the first x(j) >= xout
double interp_first_ge(const double* x, const double* y, int n, double xout) {
    if (xout < x[0]) return y[0];
    if (xout > x[n - 1]) return y[n - 1];
    int j = 0;
    while (j < n && !(x[j] >= xout)) ++j;
    if (j == 0) return y[0];
    if (j == n) return y[n - 1];
    const double al = (xout - x[j - 1]) / (x[j] - x[j - 1]);
    return y[j - 1] + al * (y[j] - y[j - 1]);
}

// bsl_init_fastiso (barysealevel.f90:155-193): bsl_init from the parameters,
// the A_ocean(bsl) table corrected for its present-day bias at bsl_init
void bsl_init_fastiso(BslDev& d, double bsl_init, int nvec, const double* bsl_vec, const double* A_vec) {
    d.bsl_init = bsl_init;
    if (nvec == 0) {
        d.constant_ocean_surface = 1;
        d.A_ocean_now = d.A_ocean_pd;
        return;
    }
    d.nvec = nvec;
    for (int i = 0; i < nvec; ++i) { d.bsl_vec[i] = bsl_vec[i]; d.A_vec[i] = A_vec[i]; }
    const double biased = interp_first_ge(d.bsl_vec, d.A_vec, nvec, d.bsl_init);
    for (int i = 0; i < nvec; ++i) d.A_vec[i] = d.A_vec[i] * d.A_ocean_pd / biased;
    d.A_ocean_now = interp_first_ge(d.bsl_vec, d.A_vec, nvec, d.bsl_init);
}

int bsl_create_impl(int method, double bsl_init, double A_ocean_pd, int nvec, const double* bsl_vec,
                    const double* A_vec, int nseries, const double* st, const double* sv, double year_bp,
                    fi_bsl** out) {
    int rc = ensure_device();
    if (rc) return rc;
    if (!out) return fail(FI_E_ARG, "NULL argument");
    if (method < 0 || method > 3)
        return fail(FI_E_ARG, "bsl_init:: Error: method not recognized (0 const, 1 fastiso, 2 file, 3 mixed)");
    if (nvec < 0 || nvec > 64) return fail(FI_E_ARG, "A_ocean table must have 0..64 entries");
    if (nvec > 0 && (!bsl_vec || !A_vec)) return fail(FI_E_ARG, "A_ocean table is NULL");
    const bool series = method == 2 || method == 3;
    if (series && (nseries < 1 || !st || !sv)) return fail(FI_E_ARG, "bsl methods file and mixed need a sea-level series");
    fi_bsl* b = new fi_bsl();
    BslDev& d = b->h;
    d.method = method;
    d.time = year_bp;
    d.A_ocean_pd = A_ocean_pd;
    d.A_ocean_now = A_ocean_pd;
    d.bsl_init = bsl_init;
    if (series) {
        // bsl_init_file (barysealevel.f90:108-153): bsl_init = the series at year_bp
        b->st.assign(st, st + nseries);
        b->sv.assign(sv, sv + nseries);
        d.bsl_now = interp_first_ge(b->st.data(), b->sv.data(), nseries, year_bp);
        d.bsl_init = d.bsl_now;
    }
    // mixed re-reads bsl_init from the parameters in bsl_init_fastiso
    if (method == 1 || method == 3) bsl_init_fastiso(d, bsl_init, nvec, bsl_vec, A_vec);
    d.bsl_now = d.bsl_init;   // barysealevel.f90:100
    if (hipMalloc(&b->d, sizeof(BslDev)) != hipSuccess) { delete b; return fail(FI_E_HIP, "hipMalloc bsl"); }
    HIPCHK(hipMemcpy(b->d, &d, sizeof(BslDev), hipMemcpyHostToDevice));
    *out = b;
    return 0;
}

}  // namespace

int fi_bsl_create(int method, double bsl_init, double A_ocean_pd, int nvec, const double* bsl_vec,
                  const double* A_vec, double year_bp, fi_bsl** out) {
    if (method == 2 || method == 3)
        return fail(FI_E_ARG, "bsl methods file and mixed take a series: use fi_bsl_create_series");
    return bsl_create_impl(method, bsl_init, A_ocean_pd, nvec, bsl_vec, A_vec, 0, nullptr, nullptr, year_bp, out);
}

int fi_bsl_create_series(int method, double bsl_init, double A_ocean_pd, int nvec, const double* bsl_vec,
                         const double* A_vec, int nseries, const double* series_time, const double* series_bsl,
                         double year_bp, fi_bsl** out) {
    return bsl_create_impl(method, bsl_init, A_ocean_pd, nvec, bsl_vec, A_vec, nseries, series_time, series_bsl,
                           year_bp, out);
}

// read_series (barysealevel.f90:254-315): the first line is a header; data
// lines hold "time value"; lines starting with '#' or '!' and blank lines are
// skipped (the reference counts them as entries it never fills).
int fi_bsl_read_series(const char* path, int max_n, double* time, double* var, int* n_out) {
    if (!path || !n_out || max_n < 0 || (max_n > 0 && (!time || !var))) return fail(FI_E_ARG, "NULL argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(FI_E_ARG, "read_series:: cannot open %s", path);
    char line[4096];
    int n = 0;
    bool header = true;
    while (std::fgets(line, sizeof line, f)) {
        if (header) { header = false; continue; }
        const char* p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\0' || *p == '\n' || *p == '\r' || *p == '#' || *p == '!') continue;
        double a, v;
        if (std::sscanf(p, "%lf%*[ ,\t]%lf", &a, &v) != 2) {
            std::fclose(f);
            return fail(FI_E_ARG, "read_series:: bad line %d in %s", n + 2, path);
        }
        if (n >= max_n) {
            std::fclose(f);
            return fail(FI_E_ARG, "read_series:: more than %d entries in %s", max_n, path);
        }
        time[n] = a;
        var[n] = v;
        ++n;
    }
    std::fclose(f);
    *n_out = n;
    return 0;
}

int fi_bsl_update(fi_bsl* b, double year_bp) {
    if (!b) return fail(FI_E_ARG, "bsl is NULL");
    // bsl_update_file (barysealevel.f90:230-238): the series at the caller's time
    const double file_value =
        b->st.empty() ? 0.0 : interp_first_ge(b->st.data(), b->sv.data(), (int)b->st.size(), year_bp);
    launch_bsl_update(b->d, year_bp, file_value, g_stream);
    HIPCHK(hipGetLastError());
    return 0;
}

int fi_bsl_get(fi_bsl* b, double* bsl_now, double* A_ocean_now) {
    if (!b) return fail(FI_E_ARG, "bsl is NULL");
    HIPCHK(hipMemcpyAsync(&b->h, b->d, sizeof(BslDev), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (bsl_now) *bsl_now = b->h.bsl_now;
    if (A_ocean_now) *A_ocean_now = b->h.A_ocean_now;
    return 0;
}

int fi_bsl_destroy(fi_bsl* b) {
    if (!b) return 0;
    (void)hipStreamSynchronize(g_stream);
    (void)hipFree(b->d);
    delete b;
    return 0;
}

static int create_impl(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy, const double* K,
                       const double* rho_ice, const double* he_field, const double* eta_field,
                       const fi_rheology_grid* grid, fi_isos** out);

int fi_isos_create(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy, const double* K,
                   const double* rho_ice, const double* he_field, const double* eta_field, fi_isos** out) {
    return create_impl(p, nx_ice, ny_ice, dx, dy, K, rho_ice, he_field, eta_field, nullptr, out);
}

int fi_isos_create_gridded(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy, const double* K,
                           const double* rho_ice, const fi_rheology_grid* grid, fi_isos** out) {
    if (!grid) return fail(FI_E_ARG, "NULL rheology grid");
    return create_impl(p, nx_ice, ny_ice, dx, dy, K, rho_ice, nullptr, nullptr, grid, out);
}

static int create_impl(const fi_params* p, int nx_ice, int ny_ice, double dx, double dy, const double* K,
                       const double* rho_ice, const double* he_field, const double* eta_field,
                       const fi_rheology_grid* grid, fi_isos** out) {
    int rc = ensure_device();
    if (rc) return rc;
    if (!p || !out) return fail(FI_E_ARG, "NULL argument");
    if (p->lithosphere == 4 && !(grid && grid->litho_thickness && grid->xc && grid->yc && grid->nx > 1 && grid->ny > 1))
        return fail(FI_E_ARG, "lithosphere=rheology_file needs the gridded thickness (fi_isos_create_gridded)");
    if (p->mantle == 4 && !(grid && grid->log10_eta && grid->zc && grid->nz >= 1 && p->lithosphere == 4))
        return fail(FI_E_ARG, "mantle=rheology_file needs the gridded log10 viscosity and lithosphere=rheology_file");
    if (grid && grid->mask && !(grid->mask_xc && grid->mask_yc && grid->mask_nx > 1 && grid->mask_ny > 1))
        return fail(FI_E_ARG, "mask_file grid is incomplete");
    if (p->method < 0 || p->method > 3)
        return fail(FI_E_ARG, "method must be 0 (none), 1 (LLRA), 2 (LV-ELRA) or 3 (LV-ELVA)");
    if (p->nl < 1 || p->nl > FI_MAX_LAYERS) return fail(FI_E_ARG, "nl out of range");
    if (p->dt_method < 0 || p->dt_method > 3) return fail(FI_E_ARG, "unknown dt_method");
    if (p->lithosphere == 3 && !he_field) return fail(FI_E_ARG, "lithosphere=fields needs he_field");
    if (p->mantle == 3 && !eta_field) return fail(FI_E_ARG, "mantle=fields needs eta_field");
    if (p->layering != 0 && p->layering != 1) return fail(FI_E_UNSUPPORTED, "layering must be parallel or equalized");
    fi_isos* h = new fi_isos();
    h->par = *p;
    if (K) h->par.correct_distortion = 1;
    if (rho_ice) h->par.rho_ice = *rho_ice;
    rc = domain_size(nx_ice, ny_ice, dx, dy, p->min_pad, h->n, h->icrop1, h->jcrop1);
    if (rc) { delete h; return rc; }
    const bool mask_file = grid && grid->mask;
    if (p->interactive_sealevel && !mask_file && !(p->min_pad > (double)100.e3f)) {
        delete h;
        return fail(FI_E_ARG, "interactive_sealevel=true requires a mask file or at least 100km of padding.");
    }
    h->nx_ice = nx_ice;
    h->ny_ice = ny_ice;
    h->dx = dx;
    h->fp64 = (p->precision == FI_FP64);
    h->uniform = (K == nullptr);
    const int n = h->n;
    const size_t nn = (size_t)n * n;
    rc = make_shape(n, 1, h->fp64, h->uniform, h->shape);
    if (rc) { delete h; return rc; }
    h->shape.method = p->method;
    h->shape.dt_method = p->dt_method;
    h->has_conv = p->include_elastic || p->heterogeneous_ssh || p->method == 2;
    if (h->has_conv) {
        rc = make_conv_shape(n, h->fp64, h->cshape);
        if (rc) { delete h; return rc; }
    }
    const size_t rp = h->fp64 ? 8 : 4;
    h->part_blocks = (int)std::max<size_t>((nn + 1023) / 1024, (size_t)n);

    // ---- arena layout
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t f64 = nn * 8;
    size_t oX0 = take(f64), oX1 = take(f64), oW = take(f64), odw = take(f64);
    size_t oz = take(f64), ozss = take(f64), odz = take(f64), owe = take(f64), oH = take(f64), oHaf = take(f64);
    size_t ocm = take(f64), ocf = take(f64), ocl = take(f64), oci = take(f64);
    size_t ozr = take(f64), ozssr = take(f64), ohafr = take(f64), ohr = take(f64), owr = take(f64), ower = take(f64),
           odzr = take(f64);
    size_t oA = 0, odxm = 0, oK2 = 0;
    if (!h->uniform) { oA = take(f64); odxm = take(f64); oK2 = take(f64); }
    size_t omask = take(nn), omaskr = take(nn);
    size_t oHin = take((size_t)nx_ice * ny_ice * 8);
    size_t oD = take(nn * rp), oeta = take(nn * rp), okR = take(nn * rp), ocfr = take(nn * rp), og1 = take(nn * rp),
           og2 = take(nn * rp);
    // partials: 4 values per block for the attempt kernels, 5 for the finalize reduction
    size_t ocor = take((size_t)n * 4 * 8), opart = take((size_t)h->part_blocks * 5 * 8);
    size_t oee = take(f64), okap = take(f64), oR = take(f64), oHe = take(f64), oDd = take(f64);
    size_t oscr = take(f64 + 64);
    size_t oweq = (p->method != 3) ? take(f64) : 0;
    const size_t hc = (size_t)(n + 1) * (h->fp64 ? 16 : 8);
    size_t ocw = 0, oFGE = 0, oFGN = 0, oFGV = 0;
    if (h->has_conv) { ocw = take(2 * (size_t)n * hc); oFGE = take(2 * (size_t)n * hc); oFGN = take(2 * (size_t)n * hc); }
    if (p->method == 2) oFGV = take(2 * (size_t)n * hc);
    size_t octl = take(sizeof(Ctl)), octl2 = take(sizeof(Ctl)), ophys = take(sizeof(PhysDev)),
           omem = take(sizeof(Member));
    if (hipMalloc(&h->arena, off) != hipSuccess) { delete h; return fail(FI_E_HIP, "hipMalloc arena (%zu bytes)", off); }
    h->arena_bytes = off;
    HIPCHK(hipMemsetAsync(h->arena, 0, off, g_stream));
    rc = sum_alloc(h->sb, 1);
    if (rc) return rc;
    char* a = h->arena;
    Member& M = h->hm;
    M.n = n; M.icrop1 = h->icrop1; M.jcrop1 = h->jcrop1; M.nx_ice = nx_ice; M.ny_ice = ny_ice;
    M.X[0] = (double*)(a + oX0); M.X[1] = (double*)(a + oX1); M.W = (double*)(a + oW); M.dwdt = (double*)(a + odw);
    M.z_bed = (double*)(a + oz); M.z_ss = (double*)(a + ozss); M.dz_ss = (double*)(a + odz); M.we = (double*)(a + owe);
    M.Hice = (double*)(a + oH); M.Haf = (double*)(a + oHaf); M.canom_mantle = (double*)(a + ocm);
    M.canom_full = (double*)(a + ocf); M.canom_load = (double*)(a + ocl); M.conv_in = (double*)(a + oci);
    M.z_bed_ref = (double*)(a + ozr); M.z_ss_ref = (double*)(a + ozssr); M.Haf_ref = (double*)(a + ohafr);
    M.Hice_ref = (double*)(a + ohr); M.w_ref = (double*)(a + owr); M.we_ref = (double*)(a + ower);
    M.dz_ss_ref = (double*)(a + odzr);
    if (!h->uniform) { M.A = (double*)(a + oA); M.dxm = (double*)(a + odxm); M.K2 = (double*)(a + oK2); }
    M.mask = (uint8_t*)(a + omask); M.mask_ref = (uint8_t*)(a + omaskr);
    M.H_in = (double*)(a + oHin);
    M.D = a + oD; M.eta = a + oeta; M.kR = a + okR; M.cfull = a + ocfr; M.g1 = a + og1; M.g2 = a + og2;
    M.corners = (double*)(a + ocor); M.partials = (double*)(a + opart);
    M.ctl = (Ctl*)(a + octl); M.ctl_next = (Ctl*)(a + octl2); M.phys = (PhysDev*)(a + ophys);
    M.bsl_lead = 1;
    M.bsl_next = -1;
    h->d_mem = (Member*)(a + omem);
    h->eta_eff = (double*)(a + oee); h->kappa = (double*)(a + okap); h->R = (double*)(a + oR);
    h->He = (double*)(a + oHe); h->Dd = (double*)(a + oDd);
    double* scratch = (double*)(a + oscr);
    if (h->has_conv) { M.conv_work = (double*)(a + ocw); M.FGE = a + oFGE; M.FGN = a + oFGN; }
    if (p->method == 2) M.FGV = a + oFGV;
    if (p->method != 3) M.weq = (double*)(a + oweq);

    // ---- physics scalars (fastisostasy.f90:152-164)
    const fi_params& P = h->par;
    PhysDev ph{};
    ph.g = P.g; ph.nu = P.nu; ph.rho_ice = P.rho_ice; ph.rho_water = P.rho_water; ph.rho_seawater = P.rho_seawater;
    ph.rho_uppermantle = P.rho_uppermantle; ph.rho_litho = P.rho_litho;
    ph.sec_per_year = (double)(3600.0f * 24.0f * 365.25f);
    ph.Vden_factor = P.rho_ice / P.rho_water - P.rho_ice / P.rho_seawater;
    ph.dt_diagnostics = P.dt_diagnostics;
    ph.atol = P.atol; ph.rtol = P.rtol; ph.dt_max = P.dt_max; ph.q = P.q;
    ph.dx = dx;
    ph.inv_dx = 1.0 / dx;
    ph.inv_2dx = 1.0 / (2 * dx);
    ph.inv_dx2 = 1.0 / (dx * dx);
    ph.tau = P.tau;
    ph.dt_method = P.dt_method;
    ph.interactive_sealevel = P.interactive_sealevel; ph.include_elastic = P.include_elastic;
    ph.heterogeneous_ssh = P.heterogeneous_ssh;
    ph.uniform_K = h->uniform;
    ph.method = P.method;
    HIPCHK(hipMemcpyAsync(M.phys, &ph, sizeof ph, hipMemcpyHostToDevice, g_stream));

    // ---- distortion fields (fastisostasy.f90:153-155)
    if (!h->uniform) {
        std::vector<double> A(nn), dxm(nn), K2(nn);
        for (size_t c = 0; c < nn; ++c) {
            dxm[c] = dx * K[c];
            A[c] = dxm[c] * (dy * K[c]);
            K2[c] = std::pow(K[c], 2.0);
        }
        HIPCHK(hipMemcpy(M.A, A.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M.dxm, dxm.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M.K2, K2.data(), nn * 8, hipMemcpyHostToDevice));
    }

    // ---- maskactive (fastisostasy.f90:185-213) in bit 3 of both masks
    {
        std::vector<uint8_t> mk(nn, 0), from_file;
        if (mask_file && P.method >= 2) grid_mask(*grid, n, dx, from_file);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                // methods 0 and 1 set maskactive everywhere (fastisostasy.f90:444-447)
                bool act = true;
                if (!from_file.empty())
                    act = from_file[(size_t)j * n + i] != 0;
                else if (P.interactive_sealevel && P.method >= 2)
                    act = (i >= h->icrop1 && i < h->icrop1 + nx_ice && j >= h->jcrop1 && j < h->jcrop1 + ny_ice);
                mk[(size_t)j * n + i] = act ? 8 : 0;
            }
        HIPCHK(hipMemcpy(M.mask, mk.data(), nn, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M.mask_ref, mk.data(), nn, hipMemcpyHostToDevice));
    }

    // ---- rheology (fastisostasy.f90:166-172, 240-442)
    if (P.method == 3) {
        // rheology_file: interpolated on the host, then used as field inputs
        std::vector<double> he_grid, eta_grid;
        if (P.lithosphere == 4 || P.mantle == 4) {
            rc = grid_rheology(P, *grid, n, dx, he_grid, eta_grid);
            if (rc) return rc;
            if (P.lithosphere == 4) he_field = he_grid.data();
            if (P.mantle == 4) eta_field = eta_grid.data();
        }
        InitArgs ia{};
        ia.n = n; ia.dx = dx; ia.nl = P.nl;
        ia.lithosphere = P.lithosphere == 4 ? 3 : P.lithosphere;
        ia.mantle = P.mantle == 4 ? 3 : P.mantle;
        ia.layering = P.layering; ia.lumping = P.lumping; ia.vscale_method = P.viscosity_scaling_method;
        ia.vscale = P.viscosity_scaling;
        for (int k = 0; k < FI_MAX_LAYERS; ++k) { ia.zl[k] = P.zl[k]; ia.visc[k] = P.viscosities[k]; }
        ia.dl = P.dl; ia.eta_ref = P.eta_ref; ia.E = P.E; ia.nu = P.nu;
        ia.comp_corr = P.correct_compression ? 1.5 / (1 + P.nu) : 1.0;
        ia.r_earth = P.r_earth; ia.m_earth = P.m_earth;
        double *d_he = nullptr, *d_eta = nullptr;
        if (ia.lithosphere == 3) {
            HIPCHK(hipMalloc(&d_he, nn * 8));
            HIPCHK(hipMemcpy(d_he, he_field, nn * 8, hipMemcpyHostToDevice));
        }
        if (ia.mantle == 3) {
            HIPCHK(hipMalloc(&d_eta, nn * 8 * P.nl));
            HIPCHK(hipMemcpy(d_eta, eta_field, nn * 8 * P.nl, hipMemcpyHostToDevice));
        }
        launch_init_rheology(h->d_mem, ia, d_he, d_eta, h->eta_eff, h->kappa, h->R, h->He, h->Dd, scratch, h->fp64,
                             g_stream);
        if (P.mantle == 4 && grid->rheo_smoothing_radius > 0.0) {
            // smooth_gauss_2D of eta_eff before the scaling tail (fastisostasy.f90:404-407)
            std::vector<double> e(nn);
            HIPCHK(hipStreamSynchronize(g_stream));
            HIPCHK(hipMemcpy(e.data(), h->eta_eff, nn * 8, hipMemcpyDeviceToHost));
            grid_smooth(n, e.data(), dx, grid->rheo_smoothing_radius);
            HIPCHK(hipMemcpy(h->eta_eff, e.data(), nn * 8, hipMemcpyHostToDevice));
        }
        launch_init_finalize(ia, scratch, h->eta_eff, h->kappa, h->R, h->Dd, M.D, M.eta, M.kR, h->fp64, g_stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(g_stream));
        if (d_he) (void)hipFree(d_he);
        if (d_eta) (void)hipFree(d_eta);
    } else {
        // He_lith = zl(1), D_lith = D_lith_const (fastisostasy.f90:166-172)
        const double He0 = P.zl[0];
        const double D0 = P.E * 1e9 * (He0 * He0 * He0) / (12.0 * (1.0 - P.nu * P.nu));
        std::vector<double> he(nn, He0), dd(nn, D0);
        HIPCHK(hipMemcpy(h->He, he.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->Dd, dd.data(), nn * 8, hipMemcpyHostToDevice));
    }

    // ---- Green's functions and their spectra (fastisostasy.f90:174-183)
    if (h->has_conv) {
        HIPCHK(hipMemcpy(h->d_mem, &M, sizeof M, hipMemcpyHostToDevice));
        for (int which = 0; which < 2; ++which) {
            launch_green(which, n, dx, P.r_earth, P.m_earth, M.conv_in, g_stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(g_stream));
            double* kc = which == 0 ? M.kc_E : M.kc_N;
            HIPCHK(hipMemcpy(&kc[0], M.conv_in, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&kc[1], M.conv_in + nn - 1, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&kc[2], M.conv_in + (size_t)(n - 1) * n, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&kc[3], M.conv_in + (n - 1), 8, hipMemcpyDeviceToHost));
            launch_kernel_spectrum(h->d_mem, 1, h->cshape, which, M.conv_in, g_stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(g_stream));
        }
        if (P.method == 2) {
            // LV-ELRA: flexural length scale and viscous Green's function
            // (isos_utils.f90:66-73, fastisostasy.f90:218-236)
            const double He0 = P.zl[0];
            const double D0 = P.E * 1e9 * (He0 * He0 * He0) / (12.0 * (1.0 - P.nu * P.nu));
            const double L_w = std::pow(D0 / (P.rho_uppermantle * P.g), 0.25);
            std::vector<double> gv(nn);
            green_viscous_host(n, dx, L_w, D0, gv.data());
            M.kc_V[0] = gv[0];
            M.kc_V[1] = gv[nn - 1];
            M.kc_V[2] = gv[(size_t)(n - 1) * n];
            M.kc_V[3] = gv[n - 1];
            HIPCHK(hipMemcpy(M.conv_in, gv.data(), nn * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->d_mem, &M, sizeof M, hipMemcpyHostToDevice));
            launch_kernel_spectrum(h->d_mem, 1, h->cshape, 2, M.conv_in, g_stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(g_stream));
        }
        HIPCHK(hipMemsetAsync(M.conv_in, 0, nn * 8, g_stream));
    }

    // ---- zero_init_state (fastisostasy.f90:461-462, 506-548)
    {
        std::vector<double> bg(nn, 1e6);
        HIPCHK(hipMemcpy(M.z_bed, bg.data(), nn * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M.z_bed_ref, bg.data(), nn * 8, hipMemcpyHostToDevice));
    }
    Ctl c{};
    c.time_diag = (double)1e10f;
    c.time_prog = (double)1e10f;
    h->time_diag = c.time_diag;
    h->time_prog = c.time_prog;
    HIPCHK(hipMemcpy(M.ctl, &c, sizeof c, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_mem, &M, sizeof M, hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(g_stream));
    *out = h;
    return 0;
}

int fi_isos_init_ref(fi_isos* h, const double* z_bed, const double* H_ice, const double* bsl, const double* dz_ss) {
    if (!h || !z_bed || !H_ice) return fail(FI_E_ARG, "NULL argument");
    const size_t nn = (size_t)h->n * h->n;
    int rc = upload_embedded(h, h->hm.z_bed_ref, z_bed, 1e6);
    if (rc) return rc;
    rc = upload_embedded(h, h->hm.Hice_ref, H_ice, 0.0);
    if (rc) return rc;
    if (bsl) {
        Ctl c;
        rc = read_ctl(h->hm, c);
        if (rc) return rc;
        c.ref_bsl = *bsl;
        rc = write_ctl(h->hm, c);
        if (rc) return rc;
    }
    if (dz_ss) HIPCHK(hipMemcpy(h->hm.dz_ss_ref, dz_ss, nn * 8, hipMemcpyHostToDevice));
    h->ref_was_set = true;
    return 0;
}

int fi_isos_init_state(fi_isos* h, const double* z_bed, const double* H_ice, double time, fi_bsl* b) {
    if (!h || !z_bed || !H_ice || !b) return fail(FI_E_ARG, "NULL argument");
    int rc = bind_bsl(h, b);
    if (rc) return rc;
    const Member& M = h->hm;
    const size_t nn = (size_t)h->n * h->n;
    // now%bsl = bsl%bsl_now (fastisostasy.f90:606)
    rc = fi_bsl_get(b, nullptr, nullptr);
    if (rc) return rc;
    Ctl c;
    rc = read_ctl(M, c);
    if (rc) return rc;
    c.bsl = b->h.bsl_now;
    if (h->use_restart) {
        // restart branch (:609-615): ref and now were read from the file;
        // extendice2isostasy (isos_utils.f90:407-418) zeroes Hice, then embeds
        // z_bed and H_ice into the ice-grid window only
        h->ref_was_set = true;
        rc = upload_embedded(h, M.Hice, H_ice, 0.0);
        if (rc) return rc;
        HIPCHK(hipMemcpy2D(M.z_bed + (size_t)h->jcrop1 * h->n + h->icrop1, (size_t)h->n * 8, z_bed,
                           (size_t)h->nx_ice * 8, (size_t)h->nx_ice * 8, h->ny_ice, hipMemcpyHostToDevice));
    } else {
        // out2in of z_bed and Hice into the zero-initialised state (:618-619)
        rc = upload_embedded(h, M.z_bed, z_bed, 1e6);
        if (rc) return rc;
        rc = upload_embedded(h, M.Hice, H_ice, 0.0);
        if (rc) return rc;
    }
    if (!h->ref_was_set) {
        // isos_init_ref(z_bed, H_ice, now%bsl, now%dz_ss) (:621-624)
        rc = upload_embedded(h, M.z_bed_ref, z_bed, 1e6);
        if (rc) return rc;
        rc = upload_embedded(h, M.Hice_ref, H_ice, 0.0);
        if (rc) return rc;
        c.ref_bsl = c.bsl;
        HIPCHK(hipMemcpy(M.dz_ss_ref, M.dz_ss, nn * 8, hipMemcpyDeviceToDevice));
        h->ref_was_set = true;
    }
    rc = write_ctl(M, c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(M.H_in, H_ice, sizeof(double) * h->nx_ice * h->ny_ice, hipMemcpyHostToDevice, g_stream));
    launch_init_state_fields(h->d_mem, 1, h->n, g_stream);
    h->time_prog = time;
    h->time_diag = time;
    std::vector<fi_isos*> hs{h};
    // the update inside isos_init_state runs before the ODE state exists: no step (:644)
    rc = run_update(hs, h->d_mem, h->shape, h->sb, time, false);
    if (rc) return rc;
    // ODE initial state (:654-656): x = now%w, kept in X[0]
    rc = read_ctl(M, c);
    if (rc) return rc;
    HIPCHK(hipMemcpy(M.X[0], M.W, nn * 8, hipMemcpyDeviceToDevice));
    c.dt = h->par.dt_init;
    c.t = time;
    c.cur = 0;
    c.w_is_x = 0;
    c.k_valid = 0;
    rc = write_ctl(M, c);
    if (rc) return rc;
    h->state_init = true;
    h->use_restart = false;
    return 0;
}

int fi_isos_update(fi_isos* h, const double* H_ice, double time, fi_bsl* b) {
    if (!h || !H_ice || !b) return fail(FI_E_ARG, "NULL argument");
    if (!h->state_init) return fail(FI_E_STATE, "isos_update called before isos_init_state");
    int rc = bind_bsl(h, b);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->hm.H_in, H_ice, sizeof(double) * h->nx_ice * h->ny_ice, hipMemcpyHostToDevice, g_stream));
    std::vector<fi_isos*> hs{h};
    return run_update(hs, h->d_mem, h->shape, h->sb, time, true);
}

// isos_update without any device allocation: the domain's per-update working
// memory (see `scratch`) lives in the library's static scratch array instead
// of the handle's arena; the kernels receive its address as their launch
// descriptor.  Same kernels, same results as fi_isos_update.  The array is
// shared by every call on a GPU and keeps the last call's contents; nothing
// in it needs to start at zero (every word is written before it is read
// within one update), and two calls that overlap on one GPU race.
int fi_isos_update_noalloc(fi_isos* h, const double* H_ice, double time, fi_bsl* b) {
    if (!h || !H_ice || !b) return fail(FI_E_ARG, "NULL argument");
    if (!h->state_init) return fail(FI_E_STATE, "isos_update called before isos_init_state");
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int rc = scratch_address_locked(&base);
        if (rc) return rc;
    }
    const int n = h->n;
    const size_t nn = (size_t)n * n, rp = h->fp64 ? 8 : 4;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t o_mem = take(sizeof(Member)), o_cor = take((size_t)n * 4 * 8),
                 o_part = take((size_t)h->part_blocks * 5 * 8), o_g1 = take(nn * rp), o_g2 = take(nn * rp);
    const size_t o_cw = h->has_conv ? take(2 * (size_t)n * (size_t)(n + 1) * (h->fp64 ? 16 : 8)) : 0;
    if (off > FI_NOALLOC_SCRATCH_BYTES)
        return fail(FI_E_CAPACITY,
                    "fi_isos_update_noalloc: N=%d (%s%s) needs %zu bytes of scratch, the library holds %zu "
                    "(N <= 512); use fi_isos_update",
                    n, h->fp64 ? "fp64" : "mixed", h->has_conv ? ", convolutions" : "", off,
                    FI_NOALLOC_SCRATCH_BYTES);
    int rc = bind_bsl(h, b);
    if (rc) return rc;
    Member M = h->hm;
    M.corners = (double*)(base + o_cor);
    M.partials = (double*)(base + o_part);
    M.g1 = base + o_g1;
    M.g2 = base + o_g2;
    if (h->has_conv) M.conv_work = (double*)(base + o_cw);
    Member* d_mem = (Member*)(base + o_mem);
    HIPCHK(hipMemcpyAsync(d_mem, &M, sizeof(Member), hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(h->hm.H_in, H_ice, sizeof(double) * h->nx_ice * h->ny_ice, hipMemcpyHostToDevice, g_stream));
    std::vector<fi_isos*> hs{h};
    // plain launches: capturing an attempt graph would allocate
    return run_update(hs, d_mem, h->shape, h->sb, time, true, nullptr, false);
}

static const double* field_ptr(fi_isos* h, int f, bool& is_mask, int& bit) {
    const Member& M = h->hm;
    is_mask = false;
    bit = 0;
    switch (f) {
    case FI_Z_BED: return M.z_bed;
    case FI_Z_SS: return M.z_ss;
    case FI_DZ_SS: return M.dz_ss;
    case FI_W: return M.W;
    case FI_WE: return M.we;
    case FI_DWDT: return M.dwdt;
    case FI_CANOM_FULL: return M.canom_full;
    case FI_HICE: return M.Hice;
    case FI_MASKOCEAN: is_mask = true; bit = 1; return nullptr;
    case FI_MASKGROUNDED: is_mask = true; bit = 2; return nullptr;
    case FI_MASKCONTINENT: is_mask = true; bit = 4; return nullptr;
    case FI_MASKACTIVE: is_mask = true; bit = 8; return nullptr;
    case FI_CANOM_MANTLE: return M.canom_mantle;
    case FI_HAF: return M.Haf;
    case FI_ETA_EFF: return h->eta_eff;
    case FI_D_LITH: return h->Dd;
    case FI_KAPPA: return h->kappa;
    case FI_R: return h->R;
    case FI_HE_LITH: return h->He;
    case FI_ODE_X: return nullptr;
    case FI_Z_SS_REF: return M.z_ss_ref;
    case FI_Z_BED_REF: return M.z_bed_ref;
    case FI_W_EQUILIBRIUM: return M.weq;
    default: return nullptr;
    }
}

// ------------------------------------------------------------------ restart files
// The variables, units and grid of the reference's NetCDF restarts
// (isos_restart_write / _read, isostasy_io.f90:56-194; bsl_restart_write,
// barysealevel.f90:426-450) in a binary container:
//   "FIRESTRT", int32 version 1, int32 nrec, double time,
//   nrec x { char name[32], char units[16], int32 nx, int32 ny, double v[nx*ny] }
// little-endian, fields in Fortran (nx, ny) order.
namespace {

constexpr char kRestartMagic[9] = "FIRESTRT";

struct RestartRec {
    std::string name, units;
    int nx, ny;
    std::vector<double> v;
};

int restart_write_file(const char* path, double time, const std::vector<RestartRec>& recs) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(FI_E_ARG, "restart: cannot create %s", path);
    const int32_t head[2] = {1, (int32_t)recs.size()};
    std::fwrite(kRestartMagic, 1, 8, f);
    std::fwrite(head, 4, 2, f);
    std::fwrite(&time, 8, 1, f);
    for (const RestartRec& r : recs) {
        char name[32] = {}, units[16] = {};
        std::strncpy(name, r.name.c_str(), sizeof name - 1);
        std::strncpy(units, r.units.c_str(), sizeof units - 1);
        const int32_t dims[2] = {r.nx, r.ny};
        std::fwrite(name, 1, 32, f);
        std::fwrite(units, 1, 16, f);
        std::fwrite(dims, 4, 2, f);
        std::fwrite(r.v.data(), 8, r.v.size(), f);
    }
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) return fail(FI_E_ARG, "restart: write to %s failed", path);
    return 0;
}

int restart_read_file(const char* path, std::vector<RestartRec>& recs, double* time) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(FI_E_ARG, "restart: cannot open %s", path);
    char magic[8];
    int32_t head[2];
    double t;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kRestartMagic, 8) != 0 || std::fread(head, 4, 2, f) != 2 ||
        std::fread(&t, 8, 1, f) != 1 || head[0] != 1 || head[1] < 0) {
        std::fclose(f);
        return fail(FI_E_ARG, "restart: %s is not a restart file", path);
    }
    recs.clear();
    for (int k = 0; k < head[1]; ++k) {
        char name[33] = {}, units[17] = {};
        int32_t dims[2];
        if (std::fread(name, 1, 32, f) != 32 || std::fread(units, 1, 16, f) != 16 || std::fread(dims, 4, 2, f) != 2 ||
            dims[0] < 0 || dims[1] < 0) {
            std::fclose(f);
            return fail(FI_E_ARG, "restart: %s is truncated", path);
        }
        RestartRec r{name, units, dims[0], dims[1], std::vector<double>((size_t)dims[0] * dims[1])};
        if (std::fread(r.v.data(), 8, r.v.size(), f) != r.v.size()) {
            std::fclose(f);
            return fail(FI_E_ARG, "restart: %s is truncated", path);
        }
        recs.push_back(std::move(r));
    }
    std::fclose(f);
    if (time) *time = t;
    return 0;
}

int download(const double* d, size_t count, std::vector<double>& out) {
    out.resize(count);
    HIPCHK(hipMemcpy(out.data(), d, count * 8, hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

int fi_isos_restart_write(fi_isos* h, const char* path, double time) {
    if (!h || !path) return fail(FI_E_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(g_stream));
    const Member& M = h->hm;
    const int n = h->n;
    const size_t nn = (size_t)n * n;
    struct { const char* name; const char* units; const double* d; } f64[] = {
        {"z_bed", "m", M.z_bed}, {"z_ss", "m", M.z_ss}, {"dz_ss", "m", M.dz_ss}, {"w", "m", M.W}, {"we", "m", M.we},
        {"dz_ss_ref", "m", M.dz_ss_ref}, {"H_ice_ref", "m", M.Hice_ref}, {"z_bed_ref", "m", M.z_bed_ref},
        {"w_ref", "m", M.w_ref}, {"we_ref", "m", M.we_ref}};
    std::vector<RestartRec> recs;
    for (const auto& v : f64) {
        RestartRec r{v.name, v.units, n, n, {}};
        int rc = download(v.d, nn, r.v);
        if (rc) return rc;
        recs.push_back(std::move(r));
    }
    RestartRec he{"He_lith", "km", n, n, {}};
    int rc = download(h->He, nn, he.v);
    if (rc) return rc;
    for (double& x : he.v) x = x * (double)1e-3f;   // He_lith * 1e-3 (single-precision literal)
    recs.push_back(std::move(he));
    std::vector<uint8_t> mk(nn);
    HIPCHK(hipMemcpy(mk.data(), M.mask, nn, hipMemcpyDeviceToHost));
    RestartRec ma{"mask_active", "1", n, n, std::vector<double>(nn)};
    for (size_t c = 0; c < nn; ++c) ma.v[c] = (mk[c] & 8) ? 1.0 : 0.0;
    recs.push_back(std::move(ma));
    if (h->par.method == 3) {
        RestartRec lg{"log10_eta_eff", "Pa s", n, n, {}};
        rc = download(h->eta_eff, nn, lg.v);
        if (rc) return rc;
        for (double& x : lg.v) x = std::log10(x);
        recs.push_back(std::move(lg));
    }
    return restart_write_file(path, time, recs);
}

int fi_isos_restart_read(fi_isos* h, const char* path, double time) {
    (void)time;
    if (!h || !path) return fail(FI_E_ARG, "NULL argument");
    std::vector<RestartRec> recs;
    int rc = restart_read_file(path, recs, nullptr);
    if (rc) return rc;
    const Member& M = h->hm;
    struct { const char* name; double* d; } want[] = {
        {"z_bed_ref", M.z_bed_ref}, {"H_ice_ref", M.Hice_ref}, {"dz_ss_ref", M.dz_ss_ref}, {"w_ref", M.w_ref},
        {"we_ref", M.we_ref}, {"z_bed", M.z_bed}, {"z_ss", M.z_ss}, {"dz_ss", M.dz_ss}, {"w", M.W}, {"we", M.we}};
    HIPCHK(hipStreamSynchronize(g_stream));
    for (const auto& w : want) {
        const RestartRec* r = nullptr;
        for (const RestartRec& q : recs)
            if (q.name == w.name) r = &q;
        if (!r) return fail(FI_E_ARG, "restart: %s has no variable %s", path, w.name);
        if (r->nx != h->n || r->ny != h->n)
            return fail(FI_E_ARG, "restart: %s is %d x %d, the domain is %d x %d", w.name, r->nx, r->ny, h->n, h->n);
        HIPCHK(hipMemcpy(w.d, r->v.data(), r->v.size() * 8, hipMemcpyHostToDevice));
    }
    h->use_restart = true;
    return 0;
}

int fi_bsl_restart_write(fi_bsl* b, const char* path, double time) {
    if (!b || !path) return fail(FI_E_ARG, "NULL argument");
    int rc = fi_bsl_get(b, nullptr, nullptr);
    if (rc) return rc;
    std::vector<RestartRec> recs;
    recs.push_back({"bsl", "m", 1, 1, {b->h.bsl_now}});
    recs.push_back({"A_ocean", "m^2", 1, 1, {b->h.A_ocean_now}});
    return restart_write_file(path, time, recs);
}

int fi_bsl_restart_read(const char* path, double* bsl, double* A_ocean) {
    if (!path) return fail(FI_E_ARG, "NULL argument");
    std::vector<RestartRec> recs;
    int rc = restart_read_file(path, recs, nullptr);
    if (rc) return rc;
    bool got_bsl = false;
    for (const RestartRec& r : recs) {
        if (r.v.size() != 1) continue;
        if (r.name == "bsl") { if (bsl) *bsl = r.v[0]; got_bsl = true; }
        if (r.name == "A_ocean" && A_ocean) *A_ocean = r.v[0];
    }
    if (!got_bsl) return fail(FI_E_ARG, "restart: %s has no variable bsl", path);
    return 0;
}

int fi_isos_get_full(fi_isos* h, int field, double* dst) {
    if (!h || !dst) return fail(FI_E_ARG, "NULL argument");
    const size_t nn = (size_t)h->n * h->n;
    bool is_mask;
    int bit;
    const double* p = field_ptr(h, field, is_mask, bit);
    if (field == FI_ODE_X) {
        Ctl c;
        int rc = read_ctl(h->hm, c);
        if (rc) return rc;
        p = h->hm.X[c.cur];
    }
    HIPCHK(hipStreamSynchronize(g_stream));
    if (is_mask) {
        std::vector<uint8_t> m(nn);
        HIPCHK(hipMemcpy(m.data(), h->hm.mask, nn, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < nn; ++c) dst[c] = (m[c] & bit) ? 1.0 : 0.0;
        return 0;
    }
    if (!p) return fail(FI_E_ARG, "unknown field %d", field);
    HIPCHK(hipMemcpy(dst, p, nn * 8, hipMemcpyDeviceToHost));
    return 0;
}

int fi_isos_get(fi_isos* h, int field, double* dst) {
    if (!h || !dst) return fail(FI_E_ARG, "NULL argument");
    if (field < 0 || field > FI_MASKCONTINENT) return fail(FI_E_ARG, "field %d is not an out%% component", field);
    bool is_mask;
    int bit;
    const double* p = field_ptr(h, field, is_mask, bit);
    const int n = h->n, ni = h->nx_ice, nj = h->ny_ice;
    HIPCHK(hipStreamSynchronize(g_stream));
    if (is_mask) {
        std::vector<uint8_t> m((size_t)ni * nj);
        HIPCHK(hipMemcpy2D(m.data(), ni, h->hm.mask + (size_t)h->jcrop1 * n + h->icrop1, n, ni, nj,
                           hipMemcpyDeviceToHost));
        for (size_t c = 0; c < m.size(); ++c) dst[c] = (m[c] & bit) ? 1.0 : 0.0;
        return 0;
    }
    HIPCHK(hipMemcpy2D(dst, (size_t)ni * 8, p + (size_t)h->jcrop1 * n + h->icrop1, (size_t)n * 8, (size_t)ni * 8, nj,
                       hipMemcpyDeviceToHost));
    return 0;
}

int fi_isos_get_mask(fi_isos* h, int field, int8_t* dst) {
    if (field < FI_MASKOCEAN || field > FI_MASKCONTINENT) return fail(FI_E_ARG, "not a mask field");
    std::vector<double> tmp((size_t)h->nx_ice * h->ny_ice);
    int rc = fi_isos_get(h, field, tmp.data());
    if (rc) return rc;
    for (size_t c = 0; c < tmp.size(); ++c) dst[c] = (int8_t)(tmp[c] != 0.0);
    return 0;
}

int fi_isos_get_scalar(fi_isos* h, int which, double* out) {
    if (!h || !out) return fail(FI_E_ARG, "NULL argument");
    Ctl c;
    int rc = read_ctl(h->hm, c);
    if (rc) return rc;
    switch (which) {
    case FI_S_N: *out = h->n; break;
    case FI_S_T: *out = c.t; break;
    case FI_S_DT: *out = c.dt; break;
    case FI_S_BSL: *out = c.bsl; break;
    case FI_S_DELTAV: *out = c.deltaV; break;
    case FI_S_V_AF: *out = c.V_af; break;
    case FI_S_V_DEN: *out = c.V_den; break;
    case FI_S_ATTEMPTS: *out = (double)c.attempts; break;
    case FI_S_ACCEPTS: *out = (double)c.accepts; break;
    case FI_S_D_LITH11: HIPCHK(hipMemcpy(out, h->Dd, 8, hipMemcpyDeviceToHost)); break;
    case FI_S_ERR: *out = c.err; break;
    case FI_S_TOL: *out = c.tol; break;
    case FI_S_HBM_BYTES: *out = (double)h->arena_bytes; break;
    case FI_S_STATUS: *out = (double)h->last_status; break;
    case FI_S_NX_ICE: *out = h->nx_ice; break;
    case FI_S_NY_ICE: *out = h->ny_ice; break;
    default: return fail(FI_E_ARG, "unknown scalar %d", which);
    }
    return 0;
}

int fi_isos_destroy(fi_isos* h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(g_stream);
    purge_graphs(h->d_mem, h->d_mem + 1);
    (void)hipFree(h->arena);
    sum_free(h->sb);
    delete h;
    return 0;
}

int fi_batch_create(fi_isos* const* members, int nmem, fi_batch** out) {
    if (!members || nmem < 1 || !out) return fail(FI_E_ARG, "empty batch");
    fi_batch* b = new fi_batch();
    for (int m = 0; m < nmem; ++m) {
        fi_isos* h = members[m];
        if (!h || h->n != members[0]->n || h->fp64 != members[0]->fp64 || h->uniform != members[0]->uniform ||
            h->par.dt_method != members[0]->par.dt_method || h->par.include_elastic != members[0]->par.include_elastic ||
            h->par.heterogeneous_ssh != members[0]->par.heterogeneous_ssh || h->par.method != members[0]->par.method) {
            delete b;
            return fail(FI_E_ARG, "batch members must share N, precision, distortion, method and dt_method");
        }
        b->mem.push_back(h);
        b->hm.push_back(h->hm);
    }
    int rc = make_shape(members[0]->n, nmem, members[0]->fp64, members[0]->uniform, b->shape);
    if (rc) { delete b; return rc; }
    b->shape.method = members[0]->par.method;
    b->shape.dt_method = members[0]->par.dt_method;
    if (hipMalloc(&b->d_mem, sizeof(Member) * nmem) != hipSuccess) { delete b; return fail(FI_E_HIP, "hipMalloc"); }
    rc = sum_alloc(b->sb, nmem);
    if (rc) return rc;
    HIPCHK(hipMemcpy(b->d_mem, b->hm.data(), sizeof(Member) * nmem, hipMemcpyHostToDevice));
    *out = b;
    return 0;
}

int fi_batch_update(fi_batch* b, const double* const* H_ice, double time, fi_bsl* const* bsl) {
    if (!b || !bsl) return fail(FI_E_ARG, "NULL argument");
    const int nmem = (int)b->mem.size();
    bool dirty = false;
    for (int m = 0; m < nmem; ++m) {
        fi_isos* h = b->mem[m];
        if (!h->state_init) return fail(FI_E_STATE, "member %d: update before init_state", m);
        if (!bsl[m]) return fail(FI_E_ARG, "member %d: bsl is NULL", m);
        if (b->hm[m].bsl != bsl[m]->d) { b->hm[m].bsl = bsl[m]->d; dirty = true; }
        if (H_ice && H_ice[m])
            HIPCHK(hipMemcpyAsync(h->hm.H_in, H_ice[m], sizeof(double) * h->nx_ice * h->ny_ice,
                                  hipMemcpyHostToDevice, g_stream));
    }
    // groups of members sharing a bsl object, in batch order
    for (int m = 0; m < nmem; ++m) {
        int prev = -1, next = -1;
        for (int q = m - 1; q >= 0 && prev < 0; --q)
            if (b->hm[q].bsl == b->hm[m].bsl) prev = q;
        for (int q = m + 1; q < nmem && next < 0; ++q)
            if (b->hm[q].bsl == b->hm[m].bsl) next = q;
        const int lead = prev < 0 ? 1 : 0;
        if (b->hm[m].bsl_lead != lead || b->hm[m].bsl_next != next) {
            b->hm[m].bsl_lead = lead;
            b->hm[m].bsl_next = next;
            dirty = true;
        }
    }
    if (dirty) HIPCHK(hipMemcpy(b->d_mem, b->hm.data(), sizeof(Member) * nmem, hipMemcpyHostToDevice));
    return run_update(b->mem, b->d_mem, b->shape, b->sb, time, true);
}

// ------------------------------------------------------------------ serving
// Continuous batching over resident domains.  Each request advances one
// domain to a target time (one isos_update).  A round takes the oldest
// pending request of every domain that has one, groups the domains that can
// share launches (equal N, precision, spacing, options and diagnostics
// phase), splits groups at max_batch and runs one batched update per group.
// Requests that arrive while a round runs join the next round.  Domains are
// admitted while the sum of their device footprints fits the HBM budget.
// Results do not depend on the grouping (the attempt kernels' reductions and
// line pairings are fixed by N), so a request gives the same bits alone or
// in any mix.
struct fi_server {
    struct Req {
        long long tag;
        double time;
        std::vector<double> H;   // empty: keep the domain's ice thickness
    };
    struct Slot {
        fi_isos* h = nullptr;
        fi_bsl* b = nullptr;
        std::deque<Req> q;
    };
    size_t budget = 0, used = 0;
    int max_batch = 0;
    std::vector<Slot> slots;
    std::vector<long long> done;
    Member* d_mem = nullptr;     // one round's Member copies, grouped
    Member* h_stage = nullptr;   // pinned staging of the same
    int cap = 0;                 // capacity of d_mem / h_stage / sb
    SumBuf sb;
    // (n, fp64, uniform, nmem, method, dt_method)
    std::map<std::tuple<int, int, int, int, int, int>, RhsShape> shapes;
    long long rounds = 0, groups = 0, served = 0;
    // Groups of one round run concurrently, round-robin on the server's own
    // streams (default priority).  Eight streams share the 4 hardware queues;
    // capping the process at one stream per queue was measured slower on the
    // serving mix (5044-5277 against 5806 requests/s of batched capacity,
    // p99 54-59 against 45 ms at 0.8x capacity).
    static constexpr int NSTREAM = 8;
    hipStream_t st[NSTREAM] = {};
    std::vector<hipEvent_t> ev;   // one per concurrent group: its last readback
    hipEvent_t start = nullptr;   // library-stream work queued before the round
    std::vector<long long> failed;   // tags of requests whose update failed its checks
    // Continuous scheduling (fi_server_poll): stream k runs at most one group
    // at a time, staged in its own range of lane_cap Member copies and
    // summaries (d_mem / h_stage / sb from k * lane_cap); a domain whose
    // oldest request is in flight stays in its queue until the group retires.
    struct Lane {
        bool busy = false;
        std::vector<int> slots;
        std::vector<double> times;
        UpdateJob job;
        hipEvent_t ev = nullptr;
        long long seq = 0;
        std::tuple<int, int, int, int, int, int, int, int> key;   // launch key (fi_server_poll's grouping)
    };
    Lane lane[NSTREAM];
    int lane_cap = 0;
    long long lane_seq = 0;
    std::vector<int> inflight;    // per slot: lane + 1 while its front request is in flight, else 0
};

namespace {

// every stream the server launches on, drained
void server_drain(fi_server* s) {
    (void)hipStreamSynchronize(g_stream);
    for (hipStream_t t : s->st)
        if (t) (void)hipStreamSynchronize(t);
}

int server_reserve(fi_server* s, int n) {
    if (n <= s->cap) return 0;
    server_drain(s);
    purge_graphs(s->d_mem, s->d_mem + s->cap);
    if (s->d_mem) (void)hipFree(s->d_mem);
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    sum_free(s->sb);
    const int c = std::max(n, 2 * s->cap);
    s->d_mem = nullptr;
    s->h_stage = nullptr;
    s->cap = 0;
    HIPCHK(hipMalloc(&s->d_mem, sizeof(Member) * c));
    HIPCHK(hipHostMalloc(&s->h_stage, sizeof(Member) * c, hipHostMallocDefault));
    int rc = sum_alloc(s->sb, c);
    if (rc) return rc;
    s->cap = c;
    return 0;
}

}  // namespace

int fi_server_create(double hbm_budget_bytes, int max_batch, fi_server** out) {
    int rc = ensure_device();
    if (rc) return rc;
    if (!out || hbm_budget_bytes <= 0 || max_batch < 1) return fail(FI_E_ARG, "bad server parameters");
    fi_server* s = new fi_server();
    s->budget = (size_t)hbm_budget_bytes;
    s->max_batch = max_batch;
    *out = s;
    return 0;
}

int fi_server_admit(fi_server* s, fi_isos* h, fi_bsl* b, int* slot) {
    if (!s || !h || !b || !slot) return fail(FI_E_ARG, "NULL argument");
    if (!h->state_init) return fail(FI_E_STATE, "admit after isos_init_state");
    for (const auto& sl : s->slots)
        if (sl.h == h) return fail(FI_E_ARG, "domain already admitted");
    if (s->used + h->arena_bytes > s->budget)
        return fail(FI_E_CAPACITY, "domain needs %zu bytes of HBM, %zu of %zu in use", h->arena_bytes, s->used,
                    s->budget);
    int id = -1;
    for (int i = 0; i < (int)s->slots.size(); ++i)
        if (!s->slots[i].h) { id = i; break; }
    if (id < 0) { id = (int)s->slots.size(); s->slots.emplace_back(); }
    s->slots[id].h = h;
    s->slots[id].b = b;
    s->used += h->arena_bytes;
    *slot = id;
    return 0;
}

int fi_server_release(fi_server* s, int slot) {
    if (!s || slot < 0 || slot >= (int)s->slots.size() || !s->slots[slot].h) return fail(FI_E_ARG, "bad slot");
    if (!s->slots[slot].q.empty()) return fail(FI_E_STATE, "slot %d has pending requests", slot);
    s->used -= s->slots[slot].h->arena_bytes;
    s->slots[slot] = fi_server::Slot();
    return 0;
}

int fi_server_submit(fi_server* s, int slot, const double* H_ice, double time, long long tag) {
    if (!s || slot < 0 || slot >= (int)s->slots.size() || !s->slots[slot].h) return fail(FI_E_ARG, "bad slot");
    fi_server::Slot& sl = s->slots[slot];
    fi_server::Req r;
    r.tag = tag;
    r.time = time;
    if (H_ice) r.H.assign(H_ice, H_ice + (size_t)sl.h->nx_ice * sl.h->ny_ice);
    sl.q.push_back(std::move(r));
    return 0;
}

int fi_server_pending(fi_server* s, int* n) {
    if (!s || !n) return fail(FI_E_ARG, "NULL argument");
    size_t k = 0;
    for (const auto& sl : s->slots) k += sl.q.size();
    *n = (int)k;
    return 0;
}

int fi_server_run(fi_server* s, int* completed) {
    if (!s || !completed) return fail(FI_E_ARG, "NULL argument");
    *completed = 0;
    for (const auto& l : s->lane)
        if (l.busy) return fail(FI_E_STATE, "fi_server_poll work in flight: poll until it has retired");
    // oldest pending request of each domain, oldest first
    std::vector<int> cand;
    for (int i = 0; i < (int)s->slots.size(); ++i)
        if (s->slots[i].h && !s->slots[i].q.empty()) cand.push_back(i);
    if (cand.empty()) return 0;
    std::stable_sort(cand.begin(), cand.end(),
                     [&](int a, int b) { return s->slots[a].q.front().tag < s->slots[b].q.front().tag; });
    // group by launch compatibility, in order of each group's oldest request
    auto key_of = [&](int i) {
        const fi_isos* h = s->slots[i].h;
        const bool d = diag_due(h, s->slots[i].q.front().time);
        return std::make_tuple(h->n, (int)h->fp64, (int)h->uniform, h->par.dt_method, h->par.include_elastic,
                               h->par.heterogeneous_ssh, (int)d, h->par.method);
    };
    std::vector<std::vector<int>> grp;
    std::vector<decltype(key_of(0))> keys;
    for (int i : cand) {
        const auto k = key_of(i);
        size_t g = 0;
        for (; g < keys.size(); ++g)
            if (keys[g] == k && (int)grp[g].size() < s->max_batch) break;
        if (g == keys.size()) { keys.push_back(k); grp.emplace_back(); }
        grp[g].push_back(i);
    }
    int rc = server_reserve(s, (int)cand.size());
    if (rc) return rc;
    // Groups run concurrently, each on its own stream, when the server batches
    // (max_batch > 1) and no bsl object is shared by two groups (domains
    // sharing one are updated in batch order, which needs one stream).
    bool concurrent = s->max_batch > 1 && grp.size() > 1;
    if (concurrent) {
        std::map<const fi_bsl*, size_t> owner;
        for (size_t g = 0; g < grp.size() && concurrent; ++g)
            for (int i : grp[g]) {
                auto it = owner.emplace(s->slots[i].b, g).first;
                if (it->second != g) { concurrent = false; break; }
            }
    }
    if (concurrent) {
        for (int k = 0; k < fi_server::NSTREAM; ++k)
            if (!s->st[k]) HIPCHK(hipStreamCreateWithFlags(&s->st[k], hipStreamNonBlocking));
        while (s->ev.size() < grp.size()) {
            hipEvent_t e = nullptr;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            s->ev.push_back(e);
        }
        // the round starts after everything already queued on the library stream
        if (!s->start) HIPCHK(hipEventCreateWithFlags(&s->start, hipEventDisableTiming));
        HIPCHK(hipEventRecord(s->start, g_stream));
        for (hipStream_t t : s->st) HIPCHK(hipStreamWaitEvent(t, s->start, 0));
    }
    std::vector<UpdateJob> jobs(grp.size());
    std::vector<std::vector<double>> tvec(grp.size());
    // A group whose update has run on the device leaves the queues at once,
    // whether its checks passed or not, so that no retry updates it twice:
    // its requests complete (fi_server_completed) or fail (fi_server_failed).
    // The round's return value is the first error; the other groups still run.
    int first_rc = 0;
    auto retire = [&](size_t gi, int status) {
        for (int i : grp[gi]) {
            fi_server::Slot& sl = s->slots[i];
            // after failed checks only the members that failed them; after any other error the whole group
            const bool checks = status == FI_E_UNSTABLE || status == FI_E_NAN;
            const bool bad = status != 0 && (!checks || sl.h->last_status != 0);
            (bad ? s->failed : s->done).push_back(sl.q.front().tag);
            sl.q.pop_front();
            if (!bad) ++*completed;
            ++s->served;
        }
        ++s->groups;
        if (status && !first_rc) first_rc = status;
    };
    // an update whose error came from its checks has been applied
    auto applied = [](int status) { return status == 0 || status == FI_E_UNSTABLE || status == FI_E_NAN; };
    std::vector<char> launched(grp.size(), 0);
    int base = 0;
    for (size_t gi = 0; gi < grp.size(); ++gi) {
        const auto& g = grp[gi];
        hipStream_t gst = concurrent ? s->st[gi % fi_server::NSTREAM] : g_stream;
        const int nm = (int)g.size();
        std::vector<fi_isos*> hs;
        std::vector<double> times;
        for (int q = 0; q < nm; ++q) {
            fi_server::Slot& sl = s->slots[g[q]];
            fi_isos* h = sl.h;
            const fi_server::Req& r = sl.q.front();
            if (!r.H.empty())
                HIPCHK(hipMemcpyAsync(h->hm.H_in, r.H.data(), sizeof(double) * r.H.size(), hipMemcpyHostToDevice,
                                      gst));
            Member M = h->hm;
            M.bsl = sl.b->d;
            M.tf_req = r.time;
            // domains of this group sharing one bsl object: batch order
            M.bsl_lead = 1;
            M.bsl_next = -1;
            for (int p = q - 1; p >= 0; --p)
                if (s->slots[g[p]].b == sl.b) { M.bsl_lead = 0; break; }
            for (int p = q + 1; p < nm; ++p)
                if (s->slots[g[p]].b == sl.b) { M.bsl_next = p; break; }
            s->h_stage[base + q] = M;
            hs.push_back(h);
            times.push_back(r.time);
        }
        // the staged copies of earlier groups are still being read by queued copies;
        // each group uses its own stage range, so no wait is needed
        HIPCHK(hipMemcpyAsync(s->d_mem + base, s->h_stage + base, sizeof(Member) * nm, hipMemcpyHostToDevice,
                              gst));
        const fi_isos* h0 = hs[0];
        auto sk = std::make_tuple(h0->n, (int)h0->fp64, (int)h0->uniform, nm, h0->par.method, h0->par.dt_method);
        auto it = s->shapes.find(sk);
        if (it == s->shapes.end()) {
            RhsShape shp{};
            rc = make_shape(h0->n, nm, h0->fp64, h0->uniform, shp);
            if (rc) { first_rc = rc; break; }
            shp.method = h0->par.method;
            shp.dt_method = h0->par.dt_method;
            it = s->shapes.emplace(sk, shp).first;
        }
        if (!concurrent) {
            // groupings change from round to round: plain launches, no graph captures
            SumBuf sb{s->sb.d + base, s->sb.h + base, nm};
            rc = run_update(hs, s->d_mem + base, it->second, sb, 0.0, true, times.data(), false);
            if (!applied(rc)) { first_rc = rc; break; }
            retire(gi, rc);
        } else {
            tvec[gi] = times;
            UpdateJob& j = jobs[gi];
            j.hs = hs;
            j.d_mem = s->d_mem + base;
            j.shape = &it->second;
            j.sb = SumBuf{s->sb.d + base, s->sb.h + base, nm};
            j.times = tvec[gi].data();
            j.stepping = true;
            j.graphs = false;
            j.subbatch = false;
            j.st = gst;
            rc = job_begin(j);
            if (!rc) rc = job_round(j);
            if (!rc) rc = hipEventRecord(s->ev[gi], gst) == hipSuccess ? 0 : fail(FI_E_HIP, "hipEventRecord");
            if (rc) {
                // this group's launches may be partly queued: let them finish, launch no more groups
                first_rc = rc;
                server_drain(s);
                break;
            }
            launched[gi] = 1;
        }
        base += nm;
    }
    if (concurrent) {
        // rounds of every launched group until all are done; a group that needs
        // another round launches it while the others' work is still on the GPU;
        // a group is retired as soon as it has finished
        size_t left = 0;
        for (char l : launched) left += l;
        std::vector<char> fin(grp.size(), 0);
        while (left > 0) {
            for (size_t gi = 0; gi < grp.size(); ++gi) {
                if (!launched[gi] || fin[gi]) continue;
                if (hipEventSynchronize(s->ev[gi]) != hipSuccess) {
                    server_drain(s);
                    return fail(FI_E_HIP, "hipEventSynchronize: serving round lost");
                }
                bool more = false;
                rc = job_more(jobs[gi], more);
                if (!rc && more) {
                    rc = job_round(jobs[gi]);
                    if (!rc && hipEventRecord(s->ev[gi], jobs[gi].st) != hipSuccess) rc = fail(FI_E_HIP, "hipEventRecord");
                    if (rc) {
                        server_drain(s);
                        if (!first_rc) first_rc = rc;
                        more = false;
                    }
                }
                if (!more) {
                    fin[gi] = 1;
                    --left;
                    // job_end's status (checks) or the integrator's failure to finish
                    retire(gi, rc ? rc : job_end(jobs[gi]));
                }
            }
        }
    }
    if (first_rc) return first_rc;
    ++s->rounds;
    return 0;
}

// Continuous scheduling: finished rounds are followed up (another round, or
// the group retires), then the oldest requests of idle domains are grouped as
// in fi_server_run and launched on idle streams, without waiting for the
// groups already running.  A domain whose bsl object is shared with a domain
// in flight (or with one of an earlier group of this call) waits, so that
// shared objects see their updates in order.
int fi_server_poll(fi_server* s, int wait, int* completed) {
    if (!s || !completed) return fail(FI_E_ARG, "NULL argument");
    *completed = 0;
    constexpr int NS = fi_server::NSTREAM;
    s->inflight.resize(s->slots.size(), 0);
    int first_rc = 0;
    auto retire = [&](fi_server::Lane& l, int status) {
        for (int i : l.slots) {
            fi_server::Slot& sl = s->slots[i];
            const bool checks = status == FI_E_UNSTABLE || status == FI_E_NAN;
            const bool bad = status != 0 && (!checks || sl.h->last_status != 0);
            (bad ? s->failed : s->done).push_back(sl.q.front().tag);
            sl.q.pop_front();
            s->inflight[i] = 0;
            if (!bad) ++*completed;
            ++s->served;
        }
        ++s->groups;
        l.busy = false;
        l.slots.clear();
        if (status && !first_rc) first_rc = status;
    };
    // 1. follow up the lanes whose last round has completed (with `wait`, block
    //    on the oldest busy lane when no round has completed yet)
    bool lost = false;   // a launch failed: no new groups in this call
    auto follow = [&](fi_server::Lane& l) {
        bool more = false;
        int rc = job_more(l.job, more);
        if (!rc && more) {
            rc = job_round(l.job);
            if (!rc && hipEventRecord(l.ev, l.job.st) != hipSuccess) rc = fail(FI_E_HIP, "hipEventRecord");
            if (rc) {
                server_drain(s);
                more = false;
                lost = true;
            }
        }
        if (!more) retire(l, rc ? rc : job_end(l.job));
    };
    bool progressed = false;
    for (auto& l : s->lane) {
        if (!l.busy) continue;
        const hipError_t q = hipEventQuery(l.ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
            server_drain(s);
            return fail(FI_E_HIP, "hipEventQuery: serving lane lost");
        }
        follow(l);
        progressed = true;
    }
    // 2. new groups on idle lanes.  One group per launch key in flight at a
    // time: requests of a key that arrive while its group runs gather into
    // the next one, so groups grow with the load.  On the serving mix
    // (40 domains of 4 sizes, Poisson arrivals at 7100 requests/s) this gave
    // p50/p99 5.8/16.7 ms against 9.9/39.9 with a new group per idle lane, 8.4/30.5
    // with two per key, and 8.4/29.5 for rounds; at 1.2x the unbatched
    // capacity 1.85/4.2 against 1.9/4.4 ms.  FI_SERVER_LANES limits the
    // lanes in use, FI_SERVER_KEY_LANES the groups of one key (tuning).
    static const int max_lanes = [] {
        const char* e = std::getenv("FI_SERVER_LANES");
        const int v = e ? std::atoi(e) : fi_server::NSTREAM;
        return std::max(1, std::min(v, (int)fi_server::NSTREAM));
    }();
    static const int key_lanes = [] {
        const char* e = std::getenv("FI_SERVER_KEY_LANES");
        return e ? std::max(1, std::atoi(e)) : 1;
    }();
    int idle = 0;
    for (int k = 0; k < max_lanes; ++k) idle += !s->lane[k].busy;
    std::vector<int> cand;
    for (int i = 0; i < (int)s->slots.size(); ++i)
        if (s->slots[i].h && !s->slots[i].q.empty() && !s->inflight[i]) cand.push_back(i);
    if (idle > 0 && !cand.empty() && !lost) {
        // stage ranges: sized while no lane is busy (a reallocation drains the streams)
        bool any_busy = false;
        for (const auto& l : s->lane) any_busy |= l.busy;
        const int want = std::min<int>(s->max_batch, (int)s->slots.size());
        if (!any_busy && want > s->lane_cap) {
            int rc = server_reserve(s, NS * want);
            if (rc) return rc;
            s->lane_cap = want;
        }
        std::stable_sort(cand.begin(), cand.end(),
                         [&](int a, int b) { return s->slots[a].q.front().tag < s->slots[b].q.front().tag; });
        // bsl object -> lane in flight (0 .. NS-1), new group (NS + index), or
        // -1: a domain using it waits in this call, so later ones wait too
        std::map<const fi_bsl*, int> bsl_owner;
        for (int k = 0; k < NS; ++k)
            if (s->lane[k].busy)
                for (int i : s->lane[k].slots) bsl_owner[s->slots[i].b] = k;
        auto key_of = [&](int i) {
            const fi_isos* h = s->slots[i].h;
            const bool d = diag_due(h, s->slots[i].q.front().time);
            return std::make_tuple(h->n, (int)h->fp64, (int)h->uniform, h->par.dt_method, h->par.include_elastic,
                                   h->par.heterogeneous_ssh, (int)d, h->par.method);
        };
        std::vector<std::vector<int>> grp;
        std::vector<decltype(key_of(0))> keys;
        std::map<decltype(key_of(0)), int> per_key;   // groups of a key in flight or formed here
        for (const auto& l : s->lane)
            if (l.busy) ++per_key[l.key];
        for (int i : cand) {
            const auto k = key_of(i);
            size_t g = 0;
            for (; g < keys.size(); ++g)
                if (keys[g] == k && (int)grp[g].size() < s->lane_cap) break;
            auto it = bsl_owner.find(s->slots[i].b);
            if (it != bsl_owner.end() && it->second != (int)g + NS) {   // shared with other work: wait
                it->second = -1;
                continue;
            }
            if (g == keys.size()) {
                if ((int)keys.size() == idle || per_key[k] >= key_lanes) {   // no lane for another group
                    bsl_owner[s->slots[i].b] = -1;
                    continue;
                }
                keys.push_back(k);
                grp.emplace_back();
                ++per_key[k];
            }
            grp[g].push_back(i);
            bsl_owner[s->slots[i].b] = (int)g + NS;
        }
        if (!grp.empty()) {
            for (int k = 0; k < NS; ++k)
                if (!s->st[k]) HIPCHK(hipStreamCreateWithFlags(&s->st[k], hipStreamNonBlocking));
            // new groups start after everything already queued on the library stream
            if (!s->start) HIPCHK(hipEventCreateWithFlags(&s->start, hipEventDisableTiming));
            HIPCHK(hipEventRecord(s->start, g_stream));
        }
        int k = 0;
        for (size_t gi = 0; gi < grp.size(); ++gi) {
            while (s->lane[k].busy) ++k;
            fi_server::Lane& l = s->lane[k];
            if (!l.ev) HIPCHK(hipEventCreateWithFlags(&l.ev, hipEventDisableTiming));
            hipStream_t gst = s->st[k];
            HIPCHK(hipStreamWaitEvent(gst, s->start, 0));
            const auto& g = grp[gi];
            const int nm = (int)g.size(), base = k * s->lane_cap;
            std::vector<fi_isos*> hs;
            l.times.clear();
            for (int q = 0; q < nm; ++q) {
                fi_server::Slot& sl = s->slots[g[q]];
                fi_isos* h = sl.h;
                const fi_server::Req& r = sl.q.front();
                if (!r.H.empty())
                    HIPCHK(hipMemcpyAsync(h->hm.H_in, r.H.data(), sizeof(double) * r.H.size(), hipMemcpyHostToDevice,
                                          gst));
                Member M = h->hm;
                M.bsl = sl.b->d;
                M.tf_req = r.time;
                M.bsl_lead = 1;
                M.bsl_next = -1;
                for (int p = q - 1; p >= 0; --p)
                    if (s->slots[g[p]].b == sl.b) { M.bsl_lead = 0; break; }
                for (int p = q + 1; p < nm; ++p)
                    if (s->slots[g[p]].b == sl.b) { M.bsl_next = p; break; }
                s->h_stage[base + q] = M;
                hs.push_back(h);
                l.times.push_back(r.time);
            }
            HIPCHK(hipMemcpyAsync(s->d_mem + base, s->h_stage + base, sizeof(Member) * nm, hipMemcpyHostToDevice,
                                  gst));
            const fi_isos* h0 = hs[0];
            auto sk = std::make_tuple(h0->n, (int)h0->fp64, (int)h0->uniform, nm, h0->par.method, h0->par.dt_method);
            auto it = s->shapes.find(sk);
            if (it == s->shapes.end()) {
                RhsShape shp{};
                int rc = make_shape(h0->n, nm, h0->fp64, h0->uniform, shp);
                if (rc) { server_drain(s); return rc; }
                shp.method = h0->par.method;
                shp.dt_method = h0->par.dt_method;
                it = s->shapes.emplace(sk, shp).first;
            }
            UpdateJob& j = l.job;
            j = UpdateJob{};
            j.hs = hs;
            j.d_mem = s->d_mem + base;
            j.shape = &it->second;
            j.sb = SumBuf{s->sb.d + base, s->sb.h + base, nm};
            j.times = l.times.data();
            j.stepping = true;
            // Attempt graphs for groups of one or two domains: one host call per
            // 8 attempts instead of 24 launches, so that enqueueing a group does
            // not hold up the launch of the next (serving mix at 1.2x the
            // unbatched capacity: p50 1.42 ms against 1.92); each lane stages in
            // its own range, so a graph serves every later group of its size
            // and shape on that lane.  FI_SERVER_GRAPHS=0 keeps plain launches.
            static const bool lane_graphs = [] {
                const char* e = std::getenv("FI_SERVER_GRAPHS");
                return !(e && std::atoi(e) == 0);
            }();
            j.graphs = lane_graphs && nm <= 2 ? 2 : 0;   // (all group sizes: no gain at high load)
            j.subbatch = false;
            j.st = gst;
            int rc = job_begin(j);
            if (!rc) rc = job_round(j);
            if (!rc) rc = hipEventRecord(l.ev, gst) == hipSuccess ? 0 : fail(FI_E_HIP, "hipEventRecord");
            if (rc) {
                // partly queued launches: let them finish; the group stays queued
                server_drain(s);
                return rc;
            }
            l.busy = true;
            l.slots = g;
            l.key = keys[gi];
            l.seq = ++s->lane_seq;
            for (int i : g) s->inflight[i] = k + 1;
            progressed = true;
        }
    }
    // 3. nothing done and something in flight: wait for the oldest lane
    if (wait && !progressed) {
        fi_server::Lane* old = nullptr;
        for (auto& l : s->lane)
            if (l.busy && (!old || l.seq < old->seq)) old = &l;
        if (old) {
            if (hipEventSynchronize(old->ev) != hipSuccess) {
                server_drain(s);
                return fail(FI_E_HIP, "hipEventSynchronize: serving lane lost");
            }
            follow(*old);
        }
    }
    return first_rc;
}

int fi_server_failed(fi_server* s, long long* tags, int max, int* n) {
    if (!s || !n || (max > 0 && !tags)) return fail(FI_E_ARG, "NULL argument");
    const int k = std::min<int>(max, (int)s->failed.size());
    for (int i = 0; i < k; ++i) tags[i] = s->failed[i];
    s->failed.erase(s->failed.begin(), s->failed.begin() + k);
    *n = k;
    return 0;
}

int fi_server_completed(fi_server* s, long long* tags, int max, int* n) {
    if (!s || !n || (max > 0 && !tags)) return fail(FI_E_ARG, "NULL argument");
    const int k = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < k; ++i) tags[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return 0;
}

int fi_server_stats(fi_server* s, long long* rounds, long long* groups, long long* served, double* hbm_used) {
    if (!s) return fail(FI_E_ARG, "NULL argument");
    if (rounds) *rounds = s->rounds;
    if (groups) *groups = s->groups;
    if (served) *served = s->served;
    if (hbm_used) *hbm_used = (double)s->used;
    return 0;
}

int fi_server_destroy(fi_server* s) {
    if (!s) return 0;
    server_drain(s);
    purge_graphs(s->d_mem, s->d_mem + s->cap);
    if (s->d_mem) (void)hipFree(s->d_mem);
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    sum_free(s->sb);
    for (hipStream_t t : s->st)
        if (t) (void)hipStreamDestroy(t);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    for (auto& l : s->lane)
        if (l.ev) (void)hipEventDestroy(l.ev);
    if (s->start) (void)hipEventDestroy(s->start);
    delete s;
    return 0;
}

int fi_prof_select(int kernel) {
    int rc = ensure_device();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g_stream));
    prof_select(kernel);
    return 0;
}

int fi_prof_collect(double* total_ms, long long* launches) {
    if (!total_ms || !launches) return fail(FI_E_ARG, "NULL argument");
    if (prof_collect(total_ms, launches)) return fail(FI_E_HIP, "event timing failed");
    return 0;
}

int fi_prof_collect_kernel(int kernel, double* total_ms, long long* launches) {
    if (!total_ms || !launches) return fail(FI_E_ARG, "NULL argument");
    if (prof_collect_kernel(kernel, total_ms, launches)) return fail(FI_E_HIP, "event timing failed");
    return 0;
}

int fi_prof_fused(void) { return prof_fused(); }

int fi_prof_shape(int* out) {
    if (!out) return fail(FI_E_ARG, "NULL argument");
    prof_shape(out);
    return 0;
}

int fi_bench_copy(double mib, int reps, double* gbs) {
    int rc = ensure_device();
    if (rc) return rc;
    if (!gbs || mib <= 0 || reps < 1) return fail(FI_E_ARG, "bad copy benchmark parameters");
    const size_t bytes = (size_t)(mib * 1048576.0);
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return fail(FI_E_HIP, "hipMalloc"); }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a, 0, bytes, g_stream));
    for (int i = 0; i < 2; ++i) HIPCHK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, g_stream));
    HIPCHK(hipEventRecord(e0, g_stream));
    for (int i = 0; i < reps; ++i) HIPCHK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, g_stream));
    HIPCHK(hipEventRecord(e1, g_stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return 0;
}

// ------------------------------------------------------------------ autotuner
// Launch shapes of the attempt kernels for a group of `nmem` domains of
// size N, picked from timed trial updates on synthetic domains (uniform
// rheology, a dome load) and kept for the process: every batch, sub-batch and
// serving group of that (N, precision, size) created afterwards uses them.
// Results do not depend on the shape (line pairings and reduction orders are
// fixed by N), only the time does.
}  // extern "C"

namespace {

fi_params synthetic_params(int precision) {
    fi_params p{};
    p.method = 3;
    p.dt_diagnostics = 10.0;
    p.mantle = 0; p.lithosphere = 0; p.layering = 1; p.lumping = 5;
    p.viscosity_scaling_method = 2; p.viscosity_scaling = 1.0;
    p.nl = 1; p.zl[0] = 100e3; p.viscosities[0] = 1e21; p.eta_ref = 1e21;
    p.rho_water = 1000.0; p.rho_ice = 910.0; p.rho_seawater = 1028.0; p.rho_uppermantle = 3300.0;
    p.E = 66.0; p.nu = 0.28; p.g = 9.81; p.r_earth = 6.378e6; p.m_earth = 5.972e24;
    p.dt_method = 3; p.atol = 1e-2; p.rtol = 1e-4; p.dt_min = 1e-10; p.dt_max = 0.2; p.dt_init = 1.0; p.q = 1.0;
    p.precision = precision;
    p.tau = 3000.0;
    return p;
}

// candidate shapes, the built-in choice first
std::vector<TunedShape> tune_candidates(int n, int nmem) {
    std::vector<TunedShape> c;
    c.push_back(TunedShape{});
    const bool p2 = (n & (n - 1)) == 0 && n >= 64 && n <= 1024;
    const bool r3 = n == 192 || n == 384;
    if (p2) {
        for (int th : {64, 128, 256, 512})
            for (int C : {2, 4, 8, 16})
                for (int rt : {1, 2})
                    for (int K = 1; K <= std::min(MAX_SUB, nmem); ++K) {
                        if (th == 512 && n != 512) continue;   // 512-thread row blocks exist for N = 512
                        TunedShape t;
                        t.threads_row = th; t.C = C; t.row_tiles = rt; t.K = K;
                        c.push_back(t);
                    }
    } else if (r3) {
        for (int C : {1, 2, 4, 8, 16})
            for (int K = 1; K <= std::min(MAX_SUB, nmem); ++K) {
                TunedShape t;
                t.C = C; t.K = K;
                c.push_back(t);
            }
    } else {
        for (int L : {1, 2, 4, 8})
            for (int C : {1, 2, 4, 8, 16})
                for (int K = 1; K <= std::min(MAX_SUB, nmem); ++K) {
                    TunedShape t;
                    t.gL = L; t.gC = C; t.K = K;
                    c.push_back(t);
                }
    }
    // small batches, where the latency of one block's work sets the launch
    // time: shorter fused row blocks (one column per thread) and an inverse
    // row pass with its own block size
    if ((p2 || r3) && nmem <= 8)
        for (int rows : {2, 4, 8})
            for (int inv : {0, 64, 128, 192})
                for (int C : {1, 2, 4, 8}) {
                    if (r3 ? (inv != 0 && inv != 192) : inv == 192) continue;   // 3 * 2^k: 384 or 192 threads
                    TunedShape t;
                    t.threads_row = p2 ? std::min(n, 256) : 0; t.C = C; t.row_tiles = 1; t.K = 1;
                    t.srow_rows = rows; t.threads_inv = inv;
                    c.push_back(t);
                }
    return c;
}

}  // namespace

extern "C" {

int fi_autotune(int n, int precision, int nmem, int reps, int* choice) {
    int rc = ensure_device();
    if (rc) return rc;
    if (n < 16 || n % 2 || nmem < 1 || nmem > 1024 || reps < 1 || (precision != FI_FP64 && precision != FI_MIXED))
        return fail(FI_E_ARG, "fi_autotune: bad arguments");
    const bool fp64 = precision == FI_FP64;
    const fi_params p = synthetic_params(precision);
    const double dx = 16e3;
    std::vector<fi_isos*> hs;
    std::vector<fi_bsl*> bs;
    auto cleanup = [&]() {
        for (fi_isos* h : hs) fi_isos_destroy(h);
        for (fi_bsl* b : bs) fi_bsl_destroy(b);
    };
    const size_t nc = (size_t)n * n;
    std::vector<double> z(nc, 1e6), H0(nc, 0.0), H(nc, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double x = (i - 0.5 * (n - 1)) * dx, y = (j - 0.5 * (n - 1)) * dx, R = 0.3 * n * dx;
            H[(size_t)j * n + i] = std::max(0.0, 2000.0 * (1.0 - (x * x + y * y) / (R * R)));
        }
    for (int m = 0; m < nmem; ++m) {
        fi_isos* h = nullptr;
        fi_bsl* b = nullptr;
        rc = fi_bsl_create(0, 0.0, 3.625e14, 0, nullptr, nullptr, 0.0, &b);
        if (!rc) bs.push_back(b);
        if (!rc) rc = fi_isos_create(&p, n, n, dx, dx, nullptr, nullptr, nullptr, nullptr, &h);
        if (!rc) hs.push_back(h);
        if (!rc) rc = fi_isos_init_state(h, z.data(), H0.data(), 0.0, b);
        if (rc) { cleanup(); return rc; }
        if (hs.back()->n != n) { cleanup(); return fail(FI_E_ARG, "fi_autotune: N must be even"); }
    }
    std::vector<const double*> Hp(nmem, H.data());
    double t = 0.0;
    // seconds per attempted step of the group under candidate c, over `r` timed updates
    auto trial = [&](const TunedShape& c, int r, double& per) -> int {
        t_trial = &c;
        fi_batch* bt = nullptr;
        per = 1e300;
        if (fi_batch_create(hs.data(), nmem, &bt)) {   // a shape the kernels cannot take is skipped
            t_trial = nullptr;
            return 0;
        }
        // one untimed update (first-round estimates, first launch of the kernels)
        t += 10.0;
        int e = fi_batch_update(bt, Hp.data(), t, bs.data());
        long long a0 = 0, a1 = 0;
        for (fi_isos* h : hs) { Ctl c0{}; if (!e) e = read_ctl(h->hm, c0); a0 += c0.attempts; }
        const auto s0 = std::chrono::steady_clock::now();
        for (int q = 0; q < r && !e; ++q) {
            t += 10.0;
            e = fi_batch_update(bt, nullptr, t, bs.data());
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
        for (fi_isos* h : hs) { Ctl c1{}; if (!e) e = read_ctl(h->hm, c1); a1 += c1.attempts; }
        fi_batch_destroy(bt);
        t_trial = nullptr;
        if (!e) per = sec / (double)std::max(1LL, a1 - a0);
        return e;
    };
    // first pass: every candidate, `reps` updates; second pass: the built-in
    // shape and the four fastest, 4 x reps updates each, twice in turn, best
    // of the two;
    // a candidate replaces the built-in shape only when it is 1% faster
    const std::vector<TunedShape> cand = tune_candidates(n, nmem);
    std::vector<std::pair<double, size_t>> r1;
    for (size_t ci = 0; ci < cand.size() && !rc; ++ci) {
        double per;
        rc = trial(cand[ci], reps, per);
        if (!rc && per < 1e299) r1.emplace_back(per, ci);
    }
    std::sort(r1.begin(), r1.end());
    std::vector<size_t> fin{0};
    for (size_t k = 0; k < r1.size() && fin.size() < 5; ++k)
        if (r1[k].second != 0) fin.push_back(r1[k].second);
    std::vector<double> best2(fin.size(), 1e300);
    for (int pass = 0; pass < 2 && !rc; ++pass)
        for (size_t k = 0; k < fin.size() && !rc; ++k) {
            double per;
            rc = trial(cand[fin[k]], 4 * reps, per);
            best2[k] = std::min(best2[k], per);
        }
    TunedShape pick{};
    double best = best2[0];
    for (size_t k = 1; k < fin.size(); ++k)
        if (best2[k] < 0.99 * best2[0] && best2[k] < best) { best = best2[k]; pick = cand[fin[k]]; }
    t_trial = nullptr;
    cleanup();
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_tuned[std::make_tuple(n, (int)fp64, nmem)] = pick;
        // the sub-batches of a split group take the same blocking
        const int K = std::max(1, std::min({pick.K ? pick.K : (nmem >= 32 ? 2 : 1), MAX_SUB, nmem}));
        if (K > 1)
            for (int k = 0; k < K; ++k) {
                const int ns = (int)((long long)nmem * (k + 1) / K) - (int)((long long)nmem * k / K);
                TunedShape sub = pick;
                sub.K = 1;
                g_tuned[std::make_tuple(n, (int)fp64, ns)] = sub;
            }
        for (auto it = g_sub_shapes.begin(); it != g_sub_shapes.end();)
            it = std::get<0>(it->first) == n ? g_sub_shapes.erase(it) : std::next(it);
    }
    if (choice) {
        choice[0] = pick.threads_row; choice[1] = pick.C; choice[2] = pick.row_tiles; choice[3] = pick.K;
        choice[4] = pick.gL; choice[5] = pick.gC; choice[6] = pick.srow_rows; choice[7] = pick.threads_inv;
    }
    return 0;
}

int fi_batch_destroy(fi_batch* b) {
    if (!b) return 0;
    (void)hipStreamSynchronize(g_stream);
    purge_graphs(b->d_mem, b->d_mem + b->mem.size());
    (void)hipFree(b->d_mem);
    sum_free(b->sb);
    delete b;
    return 0;
}

}  // extern "C"
