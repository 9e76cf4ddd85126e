// Internal runtime shared by the gfx950 translation units: sticky status, the library stream,
// event-based kernel timing, device buffers and a three-pass exclusive scan.
#pragma once
#include <initializer_list>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>

namespace mdg {

// ---------------------------------------------------------------- status
// Codes match MDG_ERR_* of mdgfx950.h (that header cannot be included next to the
// cluster-pair structs, which reuse the reference's type names).
enum { E_HIP = 1, E_BIN = 2, E_CAPACITY = 3, E_ARG = 4, E_NODEVICE = 5 };
void set_error(int code, const char* fmt, ...);
bool failed();  // true once an error is recorded; entry points then do nothing

#define MDG_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            ::mdg::set_error(1, "%s:%d %s: %s", __FILE__, __LINE__, #call,                      \
                hipGetErrorString(e_));                                                         \
        }                                                                                       \
    } while (0)

#define MDG_LAUNCH_CHECK(name)                                                                  \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess) ::mdg::set_error(1, "launch %s: %s", name, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- stream
hipStream_t stream();
double wtime();  // host monotonic seconds
bool strict_mode();

// ---------------------------------------------------------------- timing
enum Family { F_FORCE_LJ = 0, F_FORCE_EAM, F_NEIGHBOR, F_BINNING, F_PBC, F_INTEGRATE, F_THERMO,
    F_FORCE_CP, F_CLUSTER, F_SURROGATE, F_NFAMILY };
// Brackets one or more launches with events on the library stream when timing is on.
struct TimedRegion {
    explicit TimedRegion(Family f);
    ~TimedRegion();
    Family fam;
    int slot;
};

// ---------------------------------------------------------------- buffers
// Grows a device buffer to at least `bytes` (1.25x headroom); preserves `keep` leading bytes.
bool grow(void** ptr, size_t* cap, size_t bytes, size_t keep);
template <class T>
bool grow_n(T** ptr, size_t* cap_elems, size_t n, size_t keep_n)
{
    size_t capb = *cap_elems * sizeof(T);
    bool ok     = grow((void**)ptr, &capb, n * sizeof(T), keep_n * sizeof(T));
    *cap_elems  = capb / sizeof(T);
    return ok;
}
void dfree(void* p);

// Fill with a kernel (graph-safe; no hipMemsetAsync on captured streams).
void fill_i32(int* p, int v, size_t n);
void fill_f64(double* p, double v, size_t n);
void fill_f32(float* p, float v, size_t n);
// Up to FILL_MULTI_MAX int ranges filled in one launch.
constexpr int FILL_MULTI_MAX = 4;
struct FillRange {
    int* p;
    int v;
    size_t n;
};
struct FillRanges {
    FillRange r[FILL_MULTI_MAX];
    int k;
};
void fill_i32_multi(std::initializer_list<FillRange> ranges);

// Exclusive scan of n ints into out (in == out allowed); writes the total to *d_total (device).
// out[i] = in[0] + ... + in[i-1]; the total goes to *d_total and, if given, *d_total2.
void exclusive_scan(const int* in, int* out, int n, int* d_total, int* d_total2 = nullptr);
// The same (in and out distinct, one total) and in[0 .. n) cleared to 0 afterwards.
void exclusive_scan_take(int* in, int* out, int n, int* d_total);

// Grid size for n items at `block` threads per block.
inline unsigned blocks_for(long long n, int block)
{
    long long b = (n + block - 1) / block;
    return (unsigned)(b > 0 ? b : 1);
}

}  // namespace mdg
