// Runtime pieces shared by every gfx950 entry point.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <mutex>
#include <vector>

#include "mdgfx950.h"
#include "runtime.h"
#include "block_scan.h"

namespace mdg {

// ---------------------------------------------------------------- status
static int g_status = 0;
static char g_msg[512];

void set_error(int code, const char* fmt, ...)
{
    if (g_status) return;  // keep the first failure
    g_status = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    fprintf(stderr, "libmdgfx950: %s\n", g_msg);
}

bool failed() { return g_status != 0; }

// ---------------------------------------------------------------- stream and mode
static hipStream_t g_stream = nullptr;
static bool g_strict        = false;

hipStream_t stream()
{
    if (!g_stream) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            set_error(MDG_ERR_NODEVICE, "no HIP device visible");
            return nullptr;
        }
        // Default priority on purpose: streams of mixed priorities are avoided.
        MDG_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    }
    return g_stream;
}

bool strict_mode() { return g_strict; }

double wtime()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// ---------------------------------------------------------------- timing
struct Pending {
    hipEvent_t a, b;
    int fam;
};
static bool g_timing = false;
static unsigned g_timed_mask = ~0u;  // families that record events while timing is on
static int g_every = 1;              // events on every g_every-th entry of a family's region
static long long g_seen[F_NFAMILY];
static double g_ms[F_NFAMILY];
static long long g_launch[F_NFAMILY];
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_pool;
static const char* g_famname[F_NFAMILY] = { "force_lj", "force_eam", "neighbor", "binning", "pbc",
    "integrate", "thermo", "force_cp", "cluster", "surrogate" };

static hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    MDG_HIP(hipEventCreate(&e));
    return e;
}

static void harvest()
{
    for (auto& p : g_pending) {
        float ms = 0.f;
        MDG_HIP(hipEventSynchronize(p.b));
        MDG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        g_ms[p.fam] += ms;
        g_launch[p.fam] += 1;
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

TimedRegion::TimedRegion(Family f) : fam(f), slot(-1)
{
    if (!g_timing || !(g_timed_mask & (1u << f)) || failed()) return;
    if (g_seen[f]++ % g_every) return;
    Pending p;
    p.a   = take_event();
    p.b   = take_event();
    p.fam = f;
    MDG_HIP(hipEventRecord(p.a, stream()));
    g_pending.push_back(p);
    slot = (int)g_pending.size() - 1;
}

TimedRegion::~TimedRegion()
{
    if (slot < 0) return;
    MDG_HIP(hipEventRecord(g_pending[slot].b, stream()));
    if (g_pending.size() >= 1024) harvest();
}

// ---------------------------------------------------------------- buffers
bool grow(void** ptr, size_t* cap, size_t bytes, size_t keep)
{
    if (failed()) return false;
    if (bytes <= *cap && *ptr) return true;
    size_t nb = bytes + bytes / 4 + 256;
    void* np  = nullptr;
    if (hipMalloc(&np, nb) != hipSuccess) {
        set_error(MDG_ERR_CAPACITY, "hipMalloc of %zu bytes failed", nb);
        return false;
    }
    if (*ptr) {
        if (keep) MDG_HIP(hipMemcpyAsync(np, *ptr, keep < *cap ? keep : *cap,
                      hipMemcpyDeviceToDevice, stream()));
        MDG_HIP(hipStreamSynchronize(stream()));
        MDG_HIP(hipFree(*ptr));
    }
    *ptr = np;
    *cap = nb;
    return true;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------- fills
template <class T>
__global__ void k_fill(T* __restrict__ p, T v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class T>
static void fill_t(T* p, T v, size_t n)
{
    if (failed() || n == 0) return;
    unsigned g = blocks_for((long long)n, 256);
    if (g > 8192) g = 8192;
    k_fill<T><<<g, 256, 0, stream()>>>(p, v, n);
    MDG_LAUNCH_CHECK("fill");
}
void fill_i32(int* p, int v, size_t n) { fill_t<int>(p, v, n); }

// several int ranges in one launch (each range is one kernel boundary fewer)
__global__ void k_fill_multi(FillRanges f)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int q = 0; q < f.k; q++)
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.r[q].n; i += stride)
            f.r[q].p[i] = f.r[q].v;
}

void fill_i32_multi(std::initializer_list<FillRange> ranges)
{
    if (failed()) return;
    FillRanges f = {};
    size_t nmax  = 0;
    for (const FillRange& r : ranges) {
        if (f.k == FILL_MULTI_MAX) {
            set_error(1, "fill_i32_multi: more than %d ranges", FILL_MULTI_MAX);
            return;
        }
        if (r.n == 0) continue;
        f.r[f.k++] = r;
        if (r.n > nmax) nmax = r.n;
    }
    if (f.k == 0) return;
    unsigned g = blocks_for((long long)nmax, 256);
    if (g > 8192) g = 8192;
    k_fill_multi<<<g, 256, 0, stream()>>>(f);
    MDG_LAUNCH_CHECK("fill_multi");
}
void fill_f64(double* p, double v, size_t n) { fill_t<double>(p, v, n); }
void fill_f32(float* p, float v, size_t n) { fill_t<float>(p, v, n); }

// ---------------------------------------------------------------- exclusive scan
// Tiles of 256 threads x 8 items.  Pass 1 sums tiles, pass 2 scans tile sums in one block,
// pass 3 rescans each tile with its offset.
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__global__ void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ tile_sums)
{
    __shared__ int sh[33];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
    int s          = 0;
#pragma unroll
    for (int k = 0; k < SCAN_I; k++)
        if (base + k < n) s += in[base + k];
    int tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_tiles(int* __restrict__ tile_sums, int ntiles, int* __restrict__ d_total,
    int* __restrict__ d_total2)
{
    __shared__ int sh[33];
    // each thread owns a contiguous chunk of tile sums
    int per   = (ntiles + blockDim.x - 1) / blockDim.x;
    int begin = threadIdx.x * per;
    int s     = 0;
    for (int k = 0; k < per; k++)
        if (begin + k < ntiles) s += tile_sums[begin + k];
    int tot;
    int off = block_excl_scan(s, sh, &tot);
    for (int k = 0; k < per; k++)
        if (begin + k < ntiles) {
            int t                  = tile_sums[begin + k];
            tile_sums[begin + k]   = off;
            off += t;
        }
    if (threadIdx.x == 0 && d_total) *d_total = tot;
    if (threadIdx.x == 0 && d_total2) *d_total2 = tot;
}

__global__ void k_scan_apply(const int* __restrict__ in, int* __restrict__ out, int n,
    const int* __restrict__ tile_sums)
{
    __shared__ int sh[33];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
    int v[SCAN_I];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_I; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int off = block_excl_scan(s, sh, &tot) + tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_I; k++) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
}

// The whole scan in one block of SCAN1_T threads, each owning a contiguous run of at most
// SCAN1_I values (short arrays: one launch instead of three).  in and out may alias.
constexpr int SCAN1_T = 1024, SCAN1_I = 16;
__global__ void __launch_bounds__(SCAN1_T) k_scan_single(const int* in, int* out, int n,
    int* __restrict__ d_total, int* __restrict__ d_total2)
{
    __shared__ int sh[33];
    const int per = (n + SCAN1_T - 1) / SCAN1_T;
    const int b   = threadIdx.x * per;
    int v[SCAN1_I];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_I; k++) {
        v[k] = (k < per && b + k < n) ? in[b + k] : 0;
        s += v[k];
    }
    int tot;
    int off = block_excl_scan(s, sh, &tot);
#pragma unroll
    for (int k = 0; k < SCAN1_I; k++) {
        if (k < per && b + k < n) out[b + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0 && d_total) *d_total = tot;
    if (threadIdx.x == 0 && d_total2) *d_total2 = tot;
}

// k_scan_single that also clears its input (each value is read and cleared by the one thread
// that owns it); in and out must not alias.
__global__ void __launch_bounds__(SCAN1_T) k_scan_single_take(int* __restrict__ in,
    int* __restrict__ out, int n, int* __restrict__ d_total)
{
    __shared__ int sh[33];
    const int per = (n + SCAN1_T - 1) / SCAN1_T;
    const int b   = threadIdx.x * per;
    int v[SCAN1_I];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_I; k++) {
        v[k] = (k < per && b + k < n) ? in[b + k] : 0;
        s += v[k];
    }
    int tot;
    int off = block_excl_scan(s, sh, &tot);
#pragma unroll
    for (int k = 0; k < SCAN1_I; k++) {
        if (k < per && b + k < n) {
            out[b + k] = off;
            in[b + k]  = 0;
        }
        off += v[k];
    }
    if (threadIdx.x == 0 && d_total) *d_total = tot;
}

static int* g_tiles       = nullptr;
static size_t g_tiles_cap = 0;

void exclusive_scan(const int* in, int* out, int n, int* d_total, int* d_total2)
{
    if (failed()) return;
    if (n <= SCAN1_T * SCAN1_I) {
        k_scan_single<<<1, SCAN1_T, 0, stream()>>>(in, out, n, d_total, d_total2);
        MDG_LAUNCH_CHECK("exclusive_scan");
        return;
    }
    int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles < 1) ntiles = 1;
    if (!grow_n(&g_tiles, &g_tiles_cap, (size_t)ntiles, 0)) return;
    hipStream_t s = stream();
    k_scan_reduce<<<ntiles, SCAN_T, 0, s>>>(in, n, g_tiles);
    k_scan_tiles<<<1, 1024, 0, s>>>(g_tiles, ntiles, d_total, d_total2);
    k_scan_apply<<<ntiles, SCAN_T, 0, s>>>(in, out, n, g_tiles);
    MDG_LAUNCH_CHECK("exclusive_scan");
}

void exclusive_scan_take(int* in, int* out, int n, int* d_total)
{
    if (failed()) return;
    if (n <= SCAN1_T * SCAN1_I) {
        k_scan_single_take<<<1, SCAN1_T, 0, stream()>>>(in, out, n, d_total);
        MDG_LAUNCH_CHECK("exclusive_scan_take");
        return;
    }
    exclusive_scan(in, out, n, d_total);
    fill_i32(in, 0, (size_t)n);
}

}  // namespace mdg

// ---------------------------------------------------------------- C ABI
extern "C" {

int mdgfx950Status(void) { return mdg::g_status; }
const char* mdgfx950StatusMessage(void) { return mdg::g_status ? mdg::g_msg : ""; }
void mdgfx950ClearStatus(void)
{
    mdg::g_status = 0;
    mdg::g_msg[0] = 0;
}
void mdgfx950SetStrict(int strict) { mdg::g_strict = strict != 0; }

void mdgfx950EnableTiming(int on)
{
    if (on && !mdg::g_timing) {
        for (int f = 0; f < mdg::F_NFAMILY; f++) {
            mdg::g_ms[f]     = 0;
            mdg::g_launch[f] = 0;
            mdg::g_seen[f]   = 0;
        }
    }
    if (!on) mdg::harvest();
    mdg::g_timing = on != 0;
}

int mdgfx950TimeOnly(const char* family)
{
    if (!family) {
        mdg::g_timed_mask = ~0u;
        return 0;
    }
    for (int f = 0; f < mdg::F_NFAMILY; f++)
        if (strcmp(family, mdg::g_famname[f]) == 0) {
            mdg::g_timed_mask = 1u << f;
            return 0;
        }
    return -1;
}

void mdgfx950TimeEvery(int every) { mdg::g_every = every > 1 ? every : 1; }

int mdgfx950KernelTime(const char* family, double* total_ms, long long* launches)
{
    mdg::harvest();
    for (int f = 0; f < mdg::F_NFAMILY; f++)
        if (strcmp(family, mdg::g_famname[f]) == 0) {
            if (total_ms) *total_ms = mdg::g_ms[f];
            if (launches) *launches = mdg::g_launch[f];
            return 0;
        }
    return -1;
}

void mdgfx950Synchronize(void)
{
    if (mdg::g_stream) MDG_HIP(hipStreamSynchronize(mdg::g_stream));
}

}  // extern "C"
