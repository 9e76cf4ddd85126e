// Cluster-pair (GROMACS NBNXM-style) scheme on gfx950.
//
// Rebuild chain (every reneigh_every steps), all on the device and reproducing the reference's
// structures exactly: updateSingleAtoms -> wrap -> 2D binning (stable counting sort + the
// reference's selection sort by z, emulated one wave per bin) -> i-clusters of M -> j-clusters
// of N -> ghost j-clusters -> j-cluster bins (stable by bbminz) -> cluster-pair lists (one wave
// per i-cluster, ordered compaction with ballots, masked entries first).
//
// Force kernels:
//   k_cp_force_f64          one thread per i-slot, the reference kernel's arithmetic
//                           (computeForceLJRef, src/clusterpair/force_lj.c:22-189); full lists
//                           are bit-identical to it.
//   k_cp_force_f32_full     FP32-mixed, full lists: one wavefront per four i-clusters, one
//                           j-atom per lane against the group's four i-atoms, j-atoms read as
//                           16-byte FP32 records (offset in a power-of-two grid cell, cell bytes
//                           and type), the cutoff decided as in FP64 (a band re-check); with the
//                           run loop's final kick, next half kick and drift fused in.
//   k_cp_force_half_tiled   FP32-mixed, half lists: one block per tile of 16 i-clusters, the
//   + k_cp_gather_half      tile's j-forces summed in LDS as 64-bit fixed point and written once
//                           per tile, then added per atom in record order (no atomics).
//   k_cp_force_f32          FP32-mixed, one wavefront per i-cluster of 4 atoms (lane = i-atom x
//                           j-slot): half lists whose tiles outgrow their tables (one atomic per
//                           j-atom and component) and compensated half lists.
#include <type_traits>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "mdgfx950_cp.h"
#include "block_scan.h"
#include "runtime.h"

extern "C" int mdgfx950Status(void);  // mdgfx950.h (not includable beside the cluster structs)

using namespace mdg;

// Partial sums of kineticSumClusterNoAllocGFX950: a fixed device array, so that entry point
// allocates nothing.  Every device the library runs on holds its own copy until the process
// ends.  Its host-side name is not a device address: host code uses only the address that
// hipGetSymbolAddress returned for the current device (kept per device at
// initDeviceClusterGFX950), never `scratch` itself.
constexpr int KSUM_PARTS = 512;
static __device__ __align__(256) double scratch[KSUM_PARTS];

namespace {

constexpr int BLOCK       = 256;
constexpr int MAX_DEVICES = 64;
double* g_scratch[MAX_DEVICES];  // &scratch on each device, resolved once per device
constexpr int MAXT2       = 64;   // ntypes <= 8
constexpr int BIN_CAP     = 2048; // atoms per 2D column handled in LDS
constexpr int JBIN_CAP    = 1024; // j-clusters per column handled in LDS
constexpr double PAD_BASE = 1.0e12, PAD_STEP = 1.0e3;
constexpr unsigned IMASK_ALL = 0xffffffffu;

enum { ST_BIN_RANGE = 1, ST_NB_OVERFLOW = 2, ST_BIN_CAP = 4, ST_JBIN_CAP = 8, ST_REC_RANGE = 16 };
// status word of the record builders (kept across rebuilds, which clear words 0..7; words
// 12..13 hold the list build's largest cluster height)
constexpr int ST_REC_WORD = 14;

// Grid of the FP32 j-records (see make_rec): spacing h (a power of two), cell index k =
// rint(x / h) + bias stored in one byte per axis.
struct RecGrid {
    double h, inv_h;
    float hf;
    int bias;
};

struct CGrid {
    double xprd, yprd, zprd;
    double bininvx, bininvy;
    int nbinx, nbiny, mbinxlo, mbinylo, mbinx, mbiny, mbins;
};

struct Tab64 {
    double eps[MAXT2], sig6[MAXT2], cutf[MAXT2];
};
struct Tab32 {
    // F/r = r^-8 (c12 r^-6 - c6): c12 = 48 eps sigma^12, c6 = 24 eps sigma^6
    float c12[MAXT2], c6[MAXT2], eps4[MAXT2], sig6[MAXT2], cut2[MAXT2];
    double cut2d[MAXT2];  // cutoffs for the FP64 inclusion test
};

struct CPState {
    bool ready = false;
    int M = 4, N = 4, rule = 4, fma_dist = 1;
    int precision = 0, comp = 0;
    CGrid g{};
    double binsizex = 0, binsizey = 0, cutneigh = 0, cutneighsq = 0, rbb_sq = 0, cutforce = 0;
    int nstencil = 0;
    int* d_stencil = nullptr;
    size_t cap_stencil = 0;
    int natoms = 0;
    // atoms (reference order)
    double *d_ax = nullptr, *d_av = nullptr;
    int* d_at  = nullptr;
    size_t cap_ax = 0, cap_av = 0, cap_at = 0;
    int *d_abin_of = nullptr, *d_abin_cnt = nullptr, *d_abin_start = nullptr,
        *d_abin_atoms = nullptr;
    size_t cap_abin_of = 0, cap_abin_cnt = 0, cap_abin_start = 0, cap_abin_atoms = 0;
    // clusters
    int nci = 0, ncj = 0, nghost = 0, dummy_cj = 0, nghost_atoms = 0;
    // after a binning, until the ghost setup (or any other entry point) reads the cluster count
    // back: nci and ncj hold upper bounds, the count itself is status word 8 on the device
    bool nci_pending = false;
    int *d_ncl = nullptr, *d_cstart = nullptr;
    size_t cap_ncl = 0, cap_cstart = 0;
    int *d_ci_natoms = nullptr, *d_ci_bin = nullptr, *d_cj_natoms = nullptr;
    double *d_ci_bbox = nullptr, *d_cj_bbox = nullptr;
    size_t cap_ci_natoms = 0, cap_ci_bin = 0, cap_cj_natoms = 0, cap_ci_bbox = 0, cap_cj_bbox = 0;
    double *d_sx = nullptr, *d_sv = nullptr, *d_sf64 = nullptr;
    // second slot-position buffer: the force call that also integrates the next step writes
    // the new positions there (its other waves still read the current ones) and the two swap
    double* d_sxn   = nullptr;
    size_t cap_sxn  = 0;
    bool sxn_valid  = false;  // padding, dummy and ghost slots of d_sxn match d_sx
    // FP32 j-records of d_sx and d_sxn (make_rec), for the FP32 full-list kernel
    float4 *d_sr = nullptr, *d_srn = nullptr;
    size_t cap_sr = 0, cap_srn = 0;
    bool rec_valid = false;  // d_sr matches d_sx in every slot
    bool cj_from_fill = false;  // the last cluster fill also wrote the j-cluster descriptors
    // updateSingleAtoms / updateAtomsPbc requested and not yet run (see flush_atoms)
    bool atoms_copy_pending = false, atoms_wrap_pending = false;
    double wrap_prd[3] = { 0.0, 0.0, 0.0 };
    // half lists without atomics: tile tables and the records' reverse map, built from the
    // current list on the first FP32 force call after a build or prune (tiles_valid); tiles_ok
    // false when a tile outgrew its tables (that list then runs the atomic kernel)
    bool tiles_valid = false, tiles_ok = false;
    bool tiles_wait = false;  // tile tables launched, their status not yet read
    int ntiles = 0, tile_ucap = 0, tile_nmax = 0;
    int *d_tile_cj = nullptr, *d_tile_n = nullptr, *d_eslot = nullptr, *d_rec_off = nullptr,
        *d_rcount = nullptr, *d_rstart = nullptr, *d_rcur = nullptr, *d_rev = nullptr;
    size_t cap_tile_cj = 0, cap_tile_n = 0, cap_eslot = 0, cap_rec_off = 0, cap_rcount = 0,
           cap_rstart = 0, cap_rcur = 0, cap_rev = 0;
    float *d_rec = nullptr, *d_fown = nullptr;
    size_t cap_rec = 0, cap_fown = 0;
    RecGrid rg{ 1.0, 1.0, 1.0f, 0 };
    float* d_sf32 = nullptr;
    int* d_st     = nullptr;
    size_t cap_sx = 0, cap_sv = 0, cap_sf64 = 0, cap_sf32 = 0, cap_st = 0;
    // ghosts
    int *d_gcnt = nullptr, *d_goff = nullptr, *d_border = nullptr, *d_pbc = nullptr,
        *d_gbin = nullptr;
    size_t cap_gcnt = 0, cap_goff = 0, cap_border = 0, cap_pbc = 0, cap_gbin = 0;
    // j-cluster bins
    int *d_jcnt = nullptr, *d_jstart = nullptr, *d_jlist = nullptr, *d_gcur = nullptr;
    size_t cap_jcnt = 0, cap_jstart = 0, cap_jlist = 0, cap_gcur = 0;
    double* d_jkey = nullptr;  // bounding box of jlist[m] (6 per entry), in column order
    size_t cap_jkey = 0;
    int* d_zidx = nullptr;     // per column, first jlist position at or above each z boundary
    size_t cap_zidx = 0;
    double z0 = 0, inv_dz = 1;
    int *d_gbin_start = nullptr, *d_gbin_list = nullptr;
    size_t cap_gbin_start = 0, cap_gbin_list = 0;
    // lists
    int maxneighs = 200;
    int *d_nb = nullptr, *d_numneigh = nullptr, *d_nmasked = nullptr;
    size_t cap_nb = 0, cap_numneigh = 0, cap_nmasked = 0;
    // parameters
    int ntypes = 1;
    bool uniform = true;
    Tab64 t64{};
    Tab32 t32{};
    double dt = 0, dtforce = 0, mass = 1.0;
    bool forces_are_f32 = false;
    // reductions
    double *d_red = nullptr, *d_ev = nullptr, *d_evp = nullptr;  // d_evp: per-wavefront pairs
    size_t cap_red = 0, cap_evp = 0;
    int* d_status = nullptr;
    int* h_status = nullptr;
    double* h_red = nullptr;
    bool ev_on    = false;
};

CPState C;

// ---------------------------------------------------------------- device helpers
__device__ inline void c2b2d(const CGrid& g, double x, double y, int* ix, int* iy)
{
    if (x >= g.xprd) *ix = (int)((x - g.xprd) * g.bininvx) + g.nbinx - g.mbinxlo;
    else if (x >= 0.0) *ix = (int)(x * g.bininvx) - g.mbinxlo;
    else *ix = (int)(x * g.bininvx) - g.mbinxlo - 1;
    if (y >= g.yprd) *iy = (int)((y - g.yprd) * g.bininvy) + g.nbiny - g.mbinylo;
    else if (y >= 0.0) *iy = (int)(y * g.bininvy) - g.mbinylo;
    else *iy = (int)(y * g.bininvy) - g.mbinylo - 1;
}

__device__ inline unsigned imask_of(int rule, int ci, int cj)
{
    // get_imask_simd_j2/j4/j8 with rdiag = 1 (src/clusterpair/neighbor.c:240-280)
    if (rule == 2) return (ci * 2 == cj) ? 0x0002u : ((ci * 2 + 1 == cj) ? 0x002fu : IMASK_ALL);
    if (rule == 4) return (ci == cj) ? 0x08ceu : IMASK_ALL;
    return (ci == cj * 2) ? 0xf0f8fcfeu : ((ci == cj * 2 + 1) ? 0x0080c0e0u : IMASK_ALL);
}

__device__ inline double pad_pos(int slot, int d) { return PAD_BASE + (double)slot * PAD_STEP + d; }

// The FP32 j-record of a slot, one 16-byte load for the FP32 force kernels: x = (k - bias) h + o
// per axis, with k = rint(x / h) + bias an integer in [0, 255] (one byte of .w each for x, y,
// z; the atom type in the fourth byte) and o = x - (k - bias) h, exact in FP64 (|o| <= h/2, the
// subtraction of a nearby multiple of a power of two) and rounded once to FP32 (error at most
// h 2^-26).  The difference of two cell indices times h is exact in FP32, so a pair
// displacement built from records carries only the roundings of the offsets and of the final
// sums.  Far sentinel slots (padding, dummy) keep their FP32 position with k = 0: they stay far
// from everything.  A position outside the grid is clamped and flags ST_REC_RANGE.
__device__ inline float4 make_rec(double x, double y, double z, int type, const RecGrid& rg,
    int* status)
{
    const double p[3] = { x, y, z };
    float o[3];
    unsigned w = (unsigned)type << 24;
    if (fabs(x) > 0.5 * PAD_BASE || fabs(y) > 0.5 * PAD_BASE || fabs(z) > 0.5 * PAD_BASE) {
#pragma unroll
        for (int d = 0; d < 3; d++)
            o[d] = (float)p[d];
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double k = rint(p[d] * rg.inv_h) + rg.bias;
            if (!(k >= 0.0 && k <= 255.0)) {
                atomicOr(status, ST_REC_RANGE);
                k = k < 0.0 ? 0.0 : 255.0;
            }
            o[d] = (float)(p[d] - (k - rg.bias) * rg.h);
            w |= (unsigned)k << (8 * d);
        }
    }
    return make_float4(o[0], o[1], o[2], __uint_as_float(w));
}

// records of every slot from the FP64 positions
__global__ void k_cp_records(int nslot, const double* __restrict__ sx, const int* __restrict__ st,
    RecGrid rg, float4* __restrict__ sr, int* __restrict__ status)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    sr[s] = make_rec(sx[3 * s], sx[3 * s + 1], sx[3 * s + 2], st[s], rg, status);
}

// 26 ghost images in the reference's order (src/clusterpair/pbc.c:248-317): side per axis,
// 1 = low face (shift +1), 2 = high face (shift -1).
__constant__ signed char c_side[26][3] = { { 1, 0, 0 }, { 2, 0, 0 }, { 0, 1, 0 }, { 0, 2, 0 },
    { 0, 0, 1 }, { 0, 0, 2 }, { 1, 1, 1 }, { 1, 2, 1 }, { 1, 1, 2 }, { 1, 2, 2 }, { 2, 1, 1 },
    { 2, 2, 1 }, { 2, 1, 2 }, { 2, 2, 2 }, { 1, 0, 1 }, { 1, 0, 2 }, { 2, 0, 1 }, { 2, 0, 2 },
    { 0, 1, 1 }, { 0, 1, 2 }, { 0, 2, 1 }, { 0, 2, 2 }, { 1, 1, 0 }, { 2, 1, 0 }, { 1, 2, 0 },
    { 2, 2, 0 } };

__device__ inline unsigned ghost_mask_bb(const double* bb, double xprd, double yprd, double zprd,
    double cut)
{
    const double prd[3] = { xprd, yprd, zprd };
    bool lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = bb[2 * d] < cut;
        hi[d] = bb[2 * d + 1] >= (prd[d] - cut);
    }
    unsigned m = 0;
#pragma unroll
    for (int c = 0; c < 26; c++) {
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (c_side[c][d] == 1) ok = ok && lo[d];
            if (c_side[c][d] == 2) ok = ok && hi[d];
        }
        if (ok) m |= 1u << c;
    }
    return m;
}

// ---------------------------------------------------------------- updateSingleAtoms + wrap
// The atoms of cluster ci in the cluster order with the padding dropped start at its bin's first
// atom plus M per cluster before it in the bin (k_cp_fill_clusters fills only a bin's last
// cluster partly), from the binning that made the clusters.
// With wrap, each position is also wrapped into the box as k_cp_wrap does (updateAtomsPbc); with
// cnt, each atom is also binned and counted as k_cp_bin_count does (the start of the next
// binning), the counts summed per wavefront and bin before the atomic.
__global__ void k_cp_slots_to_atoms(int nci, int M, const int* __restrict__ ci_natoms,
    const int* __restrict__ ci_bin, const int* __restrict__ cstart,
    const int* __restrict__ abin_start, const double* __restrict__ sx,
    const double* __restrict__ sv, const int* __restrict__ st, double* __restrict__ ax,
    double* __restrict__ av, int* __restrict__ at, int wrap, double xprd, double yprd,
    double zprd, CGrid g, int* __restrict__ bin_of, int* __restrict__ cnt,
    int* __restrict__ status)
{
    const int s    = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int ci = s / M, cii = s % M;
    int bn       = -1;
    if (ci < nci && cii < ci_natoms[ci]) {
        const int b         = ci_bin[ci];
        const int a         = abin_start[b] + (ci - cstart[b]) * M + cii;
        const double prd[3] = { xprd, yprd, zprd };
        double x[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double xx = sx[3 * s + d];
            if (wrap) {
                if (xx < 0.0) xx += prd[d];
                else if (xx >= prd[d]) xx -= prd[d];
            }
            x[d]          = xx;
            ax[3 * a + d] = xx;
            av[3 * a + d] = sv[3 * s + d];
        }
        at[a] = st[s];
        if (cnt) {
            int ix, iy;
            c2b2d(g, x[0], x[1], &ix, &iy);
            bn = iy * g.mbinx + ix + 1;
            if (bn < 0 || bn >= g.mbins) {
                atomicOr(status, ST_BIN_RANGE);
                bn = -1;
            }
            bin_of[a] = bn;
        }
    }
    if (!cnt) return;  // uniform
    unsigned long long act = __ballot(bn >= 0);
    while (act) {
        const int l                  = __ffsll((long long)act) - 1;
        const int lb                 = __shfl(bn, l, 64);
        const unsigned long long grp = __ballot(bn == lb);
        if (lane == l) atomicAdd(&cnt[lb], __popcll(grp));
        act &= ~grp;
    }
}

__global__ void k_cp_wrap(int n3, double* __restrict__ x, double xprd, double yprd, double zprd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    int d      = i % 3;
    double prd = d == 0 ? xprd : (d == 1 ? yprd : zprd);
    double xx  = x[i];
    if (xx < 0.0) xx += prd;
    else if (xx >= prd) xx -= prd;
    x[i] = xx;
}

// ---------------------------------------------------------------- binAtoms (neighbor.c:1409-1449)
// Atoms arrive in cluster order after the first build, so a wavefront's atoms mostly share one
// column: the lanes of each distinct bin are counted with a ballot and one lane adds the count
// (64 lanes hitting one counter would serialise).  Every lane runs the loop; it is wave-uniform.
__global__ void k_cp_bin_count(int natoms, const double* __restrict__ ax, CGrid g,
    int* __restrict__ bin_of, int* __restrict__ cnt, int* __restrict__ status)
{
    const int i    = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int b          = -1;
    if (i < natoms) {
        int ix, iy;
        c2b2d(g, ax[3 * i], ax[3 * i + 1], &ix, &iy);
        b = iy * g.mbinx + ix + 1;
        if (b < 0 || b >= g.mbins) {
            atomicOr(status, ST_BIN_RANGE);
            b = -1;
        }
        bin_of[i] = b;
    }
    unsigned long long act = __ballot(b >= 0);
    while (act) {
        const int l                  = __ffsll((long long)act) - 1;
        const int lb                 = __shfl(b, l, 64);
        const unsigned long long grp = __ballot(b == lb);
        if (lane == l) atomicAdd(&cnt[lb], __popcll(grp));
        act &= ~grp;
    }
}

// Same grouping; lanes of one bin take consecutive places in lane (atom-index) order.
__global__ void k_cp_bin_scatter(int natoms, const int* __restrict__ bin_of,
    const int* __restrict__ start, int* __restrict__ cursor, int* __restrict__ atoms)
{
    const int i    = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int b    = i < natoms ? bin_of[i] : -1;
    unsigned long long act = __ballot(b >= 0);
    while (act) {
        const int l                  = __ffsll((long long)act) - 1;
        const int lb                 = __shfl(b, l, 64);
        const unsigned long long grp = __ballot(b == lb);
        int base                     = 0;
        if (lane == l) base = atomicAdd(&cursor[lb], __popcll(grp));
        base = __shfl(base, l, 64);
        if (b == lb) atoms[start[lb] + base + __popcll(grp & lt)] = i;
        act &= ~grp;
    }
}

// One block per bin.  The reference fills a bin in atom-index order and then runs
// sortAtomsByZCoord (neighbor.c:1452-1481): at step a, the first position holding the minimum z
// among positions a..c-1 swaps with position a.  With all z distinct that selection sort ends in
// the unique z order, so each atom's place is its z rank.  Only a column with equal z values (a
// perfect lattice) restores index order and replays the selection sort, whose swaps decide the
// order of ties.  Also counts the bin's clusters.
// The z rank: the atoms go into SORT_NBK z buckets (floor(z inv_dz), clamped: non-decreasing in
// z, so an atom in a lower bucket has a smaller z); an atom's rank is the count of the lower
// buckets plus the atoms of its own bucket with a smaller z (mostly none or one).
constexpr int SORT_T = 256, SORT_NBK = 1024, SORT_PER = BIN_CAP / SORT_T;
__device__ inline int zsort_bucket(double z, double inv_dz)
{
    const double t = floor(z * inv_dz);
    return t < 0.0 ? 0 : (t > (double)(SORT_NBK - 1) ? SORT_NBK - 1 : (int)t);
}
__global__ void __launch_bounds__(SORT_T) k_cp_bin_sort(const int* __restrict__ start,
    int* __restrict__ atoms, const double* __restrict__ ax, int M, int N, int* __restrict__ ncl,
    double inv_dz, int* __restrict__ status)
{
    __shared__ int s_bk[SORT_NBK];
    __shared__ __align__(16) int s_idx[BIN_CAP];
    __shared__ __align__(16) int s_tmp[BIN_CAP];
    __shared__ __align__(16) double s_z[BIN_CAP];
    __shared__ int s_tie;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63;
    const int s0 = start[b], c = start[b + 1] - s0;
    if (c > BIN_CAP) {
        if (t == 0) {
            atomicOr(status, ST_BIN_CAP);
            ncl[b] = 0;
        }
        return;
    }
    if (t == 0) {
        s_tie  = 0;
        int nc = (c + M - 1) / M;
        if (N > M && (nc % 2)) nc++;
        ncl[b] = nc;
    }
    for (int p = t; p < c; p += SORT_T) {
        const int i = atoms[s0 + p];
        s_idx[p]    = i;
        s_z[p]      = ax[3 * i + 2];
    }
    for (int q = t; q < SORT_NBK; q += SORT_T)
        s_bk[q] = 0;
    __syncthreads();
    for (int p = t; p < c; p += SORT_T)
        atomicAdd(&s_bk[zsort_bucket(s_z[p], inv_dz)], 1);
    __syncthreads();
    {  // bucket starts: an exclusive scan, SORT_NBK / SORT_T buckets per thread
        __shared__ int sh[33];
        constexpr int PB = SORT_NBK / SORT_T;
        int v[PB], sum = 0;
#pragma unroll
        for (int k = 0; k < PB; k++) {
            v[k] = s_bk[t * PB + k];
            sum += v[k];
        }
        int tot;
        int off = mdg::block_excl_scan(sum, sh, &tot);
#pragma unroll
        for (int k = 0; k < PB; k++) {
            s_bk[t * PB + k] = off;
            off += v[k];
        }
    }
    __syncthreads();
    // the atoms in bucket order (s_bk then holds each bucket's end)
    for (int p = t; p < c; p += SORT_T)
        s_tmp[atomicAdd(&s_bk[zsort_bucket(s_z[p], inv_dz)], 1)] = p;
    __syncthreads();
    bool tie = false;
    int rk[SORT_PER];
#pragma unroll
    for (int k = 0; k < SORT_PER; k++) {
        const int p = t + k * SORT_T;
        rk[k]       = 0;
        if (p < c) {
            const double zp = s_z[p];
            const int bk    = zsort_bucket(zp, inv_dz);
            const int q1    = s_bk[bk];
            int r = bk ? s_bk[bk - 1] : 0, e = 0;
            for (int q = r; q < q1; q++) {
                const double zq = s_z[s_tmp[q]];
                r += zq < zp;
                e += zq == zp;
            }
            tie |= e > 1;
            rk[k] = r;
        }
    }
    if (__any(tie) && lane == 0) s_tie = 1;
    __syncthreads();
    if (!s_tie) {
#pragma unroll
        for (int k = 0; k < SORT_PER; k++) {
            const int p = t + k * SORT_T;
            if (p < c) atoms[s0 + rk[k]] = s_idx[p];
        }
        return;
    }
    __syncthreads();  // (s_tmp is reused below)
    // ties: index order first (indices are distinct), then the selection sort on one wavefront
    for (int p = t; p < c; p += SORT_T) {
        const int v = s_idx[p];
        int r = 0, q = 0;
        for (; q + 4 <= c; q += 4) {
            const int4 w = *reinterpret_cast<const int4*>(&s_idx[q]);
            r += (w.x < v) + (w.y < v) + (w.z < v) + (w.w < v);
        }
        for (; q < c; q++)
            r += s_idx[q] < v;
        s_tmp[r] = v;
    }
    __syncthreads();
    for (int p = t; p < c; p += SORT_T) {
        s_idx[p] = s_tmp[p];
        s_z[p]   = ax[3 * s_tmp[p] + 2];
    }
    __syncthreads();
    if (t < 64) {
        for (int a = 0; a < c; a++) {
            double bz = INFINITY;
            int bp    = INT_MAX;
            for (int p = a + lane; p < c; p += 64) {
                double zz = s_z[p];
                if (zz < bz || (zz == bz && p < bp)) {
                    bz = zz;
                    bp = p;
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                double oz = __shfl_xor(bz, o, 64);
                int op    = __shfl_xor(bp, o, 64);
                if (oz < bz || (oz == bz && op < bp)) {
                    bz = oz;
                    bp = op;
                }
            }
            if (lane == 0) {
                int mp    = (bp == INT_MAX) ? a : bp;
                int ti    = s_idx[a];
                double tz = s_z[a];
                s_idx[a]  = s_idx[mp];
                s_z[a]    = s_z[mp];
                s_idx[mp] = ti;
                s_z[mp]   = tz;
            }
            // one wavefront: its LDS operations complete in order
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    for (int p = t; p < c; p += SORT_T)
        atoms[s0 + p] = s_idx[p];
}

// Scans inside one block of ONE_T threads, each owning a contiguous run of at most ONE_I bins.
constexpr int ONE_T = 1024, ONE_I = 16;

// buildClustersCPU (neighbor.c:1513-1611): cluster ci holds atoms cl*M .. cl*M+M-1 of its bin.
// One block per bin, one lane per slot of the bin's clusters (contiguous from cstart[b] M on;
// each wavefront takes 64 of them at a time); a cluster's M lanes are adjacent (M divides 64)
// and reduce its bounding box and atom count with cross-lane reads.  Also each cluster's bin,
// and with cj_natoms (M == N) the j-cluster descriptors, which are then the i-clusters'.
__global__ void k_cp_fill_clusters(int mbins, int M, const int* __restrict__ ncl,
    const int* __restrict__ cstart, const int* __restrict__ abin_start,
    const int* __restrict__ abin_atoms, const double* __restrict__ ax,
    const double* __restrict__ av, const int* __restrict__ at, double* __restrict__ sx,
    double* __restrict__ sv, int* __restrict__ st, int* __restrict__ ci_natoms,
    double* __restrict__ ci_bbox, int* __restrict__ ci_bin, int* __restrict__ cj_natoms,
    double* __restrict__ cj_bbox)
{
    const int lane = threadIdx.x & 63;
    const int b    = blockIdx.x;
    if (b >= mbins) return;
    const int c0 = cstart[b], nsl = ncl[b] * M;
    const int s0 = abin_start[b], c = abin_start[b + 1] - s0;
    for (int base = threadIdx.x & ~63; base < nsl; base += blockDim.x) {
        const int k     = base + lane;  // slot within the bin: the atom index within it
        const bool live = k < nsl;
        const int s     = c0 * M + k;
        const int ci = c0 + k / M, cii = k % M;
        bool has        = false;
        double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
        if (live) {
            has = k < c;
            if (has) {
                const int i = abin_atoms[s0 + k];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double xv = ax[3 * i + d];
                    sx[3 * s + d]   = xv;
                    sv[3 * s + d]   = av[3 * i + d];
                    lo[d] = hi[d] = xv;
                }
                st[s] = at[i];
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    sx[3 * s + d] = pad_pos(s, d);
                    sv[3 * s + d] = 0.0;
                }
                st[s] = 0;
            }
        }
        for (int o = 1; o < M; o <<= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
                hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
            }
        const unsigned long long bal = __ballot(has);
        if (live && cii == 0) {
            const int na  = __popcll((bal >> (lane & ~(M - 1))) & ((M >= 64) ? ~0ull : ((1ull << M) - 1)));
            ci_natoms[ci] = na;
            ci_bin[ci]    = b;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                ci_bbox[6 * ci + 2 * d]     = lo[d];
                ci_bbox[6 * ci + 2 * d + 1] = hi[d];
            }
            // M == N: the j-clusters are the i-clusters (defineJClusters, fused)
            if (cj_natoms) {
                cj_natoms[ci] = na;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    cj_bbox[6 * ci + 2 * d]     = lo[d];
                    cj_bbox[6 * ci + 2 * d + 1] = hi[d];
                }
            }
        }
    }
}

// defineJClusters (neighbor.c:1807-1927), shapes M == N and N == 2M.  The j-cluster count is
// nci * M / N with nci read from the device (the grid covers an upper bound).
__global__ void k_cp_define_j(const int* __restrict__ d_nci, int M, int N,
    const int* __restrict__ ci_natoms, const double* __restrict__ ci_bbox,
    int* __restrict__ cj_natoms, double* __restrict__ cj_bbox)
{
    int cj = blockIdx.x * blockDim.x + threadIdx.x;
    if (cj >= *d_nci * M / N) return;
    if (M == N) {
        cj_natoms[cj] = ci_natoms[cj];
        for (int k = 0; k < 6; k++)
            cj_bbox[6 * cj + k] = ci_bbox[6 * cj + k];
    } else {
        int c0 = 2 * cj, c1 = 2 * cj + 1;
        for (int d = 0; d < 3; d++) {
            cj_bbox[6 * cj + 2 * d]     = fmin(ci_bbox[6 * c0 + 2 * d], ci_bbox[6 * c1 + 2 * d]);
            cj_bbox[6 * cj + 2 * d + 1] = fmax(ci_bbox[6 * c0 + 2 * d + 1], ci_bbox[6 * c1 + 2 * d + 1]);
        }
        cj_natoms[cj] = ci_natoms[c0] + ci_natoms[c1];
    }
}

// ---------------------------------------------------------------- setupPbc (pbc.c:188-331)
// ghost clusters per j-cluster, and the ghost atoms they hold summed into *natoms_ghost: one
// atomic per block of GC_T threads (same-address atomics serialise at the memory side).  The
// grid covers ncj_cap counters; those past the j-cluster count (nci * M / N, nci read from the
// device) are set to 0, so that a scan over all ncj_cap gives the same offsets and total.
constexpr int GC_T = 1024;
__global__ void __launch_bounds__(GC_T) k_cp_ghost_count(int ncj_cap, const int* __restrict__ d_nci,
    int M, int N, const int* __restrict__ cj_natoms, const double* __restrict__ cj_bbox,
    double xprd, double yprd, double zprd, double cut, int* __restrict__ cnt,
    int* __restrict__ natoms_ghost)
{
    __shared__ int s_n[GC_T / 64];
    const int cj  = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncj = *d_nci * M / N;
    int nadd      = 0;
    if (cj >= ncj && cj < ncj_cap) cnt[cj] = 0;
    if (cj < ncj) {
        const int na = cj_natoms[cj];
        const int c  = na > 0 ? __popc(ghost_mask_bb(&cj_bbox[6 * cj], xprd, yprd, zprd, cut)) : 0;
        cnt[cj]      = c;
        nadd         = c * na;
    }
    for (int o = 32; o > 0; o >>= 1)
        nadd += __shfl_xor(nadd, o, 64);
    if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = nadd;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < GC_T / 64; w++)
            t += s_n[w];
        if (t) atomicAdd(natoms_ghost, t);
    }
}

__global__ void k_cp_ghost_fill(int ncj, int N, const int* __restrict__ cj_natoms_in,
    const double* __restrict__ cj_bbox, double xprd, double yprd, double zprd, double cut,
    const int* __restrict__ off, int* __restrict__ border, int* __restrict__ pbc,
    int* __restrict__ cj_natoms, int* __restrict__ st)
{
    const int cj = blockIdx.x * blockDim.x + threadIdx.x;
    if (cj < ncj) {
        const int na = cj_natoms_in[cj];
        if (na > 0) {
            unsigned m = ghost_mask_bb(&cj_bbox[6 * cj], xprd, yprd, zprd, cut);
            int g      = off[cj];
            for (int c = 0; c < 26; c++) {
                if (!(m & (1u << c))) continue;
                border[g] = cj;
                for (int d = 0; d < 3; d++)
                    pbc[3 * g + d] = c_side[c][d] == 1 ? 1 : (c_side[c][d] == 2 ? -1 : 0);
                int cg        = ncj + g;
                cj_natoms[cg] = na;
                for (int cjj = 0; cjj < na; cjj++)
                    st[cg * N + cjj] = st[cj * N + cjj];
                g++;
            }
        }
    }
}

// updatePbcCPU with firstUpdate = 1: positions, padding and bounding boxes of ghost clusters.
__global__ void k_cp_ghost_first(int ncj, int nghost, int N, const int* __restrict__ border,
    const int* __restrict__ pbc, const int* __restrict__ cj_natoms, double xprd, double yprd,
    double zprd, double* __restrict__ sx, int* __restrict__ st, double* __restrict__ cj_bbox)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > nghost) return;
    int cg = ncj + g;
    if (g == nghost) {  // the dummy cluster
        for (int cjj = 0; cjj < N; cjj++) {
            int s = cg * N + cjj;
            for (int d = 0; d < 3; d++)
                sx[3 * s + d] = pad_pos(s, d);
            st[s] = 0;
        }
        return;
    }
    const double prd[3] = { xprd, yprd, zprd };
    int b  = border[g];
    int na = cj_natoms[cg];
    double bb[6] = { INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY };
    for (int cjj = 0; cjj < N; cjj++) {
        int s = cg * N + cjj;
        if (cjj < na) {
            int sb = b * N + cjj;
            for (int d = 0; d < 3; d++) {
                double xv    = sx[3 * sb + d] + pbc[3 * g + d] * prd[d];
                sx[3 * s + d] = xv;
                if (bb[2 * d] > xv) bb[2 * d] = xv;
                if (bb[2 * d + 1] < xv) bb[2 * d + 1] = xv;
            }
        } else {
            for (int d = 0; d < 3; d++)
                sx[3 * s + d] = pad_pos(s, d);
            st[s] = 0;
        }
    }
    for (int k = 0; k < 6; k++)
        cj_bbox[6 * cg + k] = bb[k];
}

// updatePbcCPU with firstUpdate = 0: one thread per ghost slot; with sr, its FP32 record too
// (the ghost's own, from its shifted FP64 position).
template <int N>
__global__ void k_cp_ghost_update(int ncj, int nghost, const int* __restrict__ border,
    const int* __restrict__ pbc, const int* __restrict__ cj_natoms, double xprd, double yprd,
    double zprd, double* __restrict__ sx, const int* __restrict__ st, RecGrid rg,
    float4* __restrict__ sr, int* __restrict__ status)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nghost * N) return;
    int g = s / N, cjj = s % N;
    // every load is issued before the atom-count test, whose only effect is on the stores (a
    // padding slot's source is a padding slot of the source cluster, in bounds): one chain of
    // two dependent loads per slot instead of three
    const int na   = cj_natoms[ncj + g];
    const int sb   = border[g] * N + cjj, sg = (ncj + g) * N + cjj;
    const int tg   = st[sg];
    const double x = sx[3 * sb] + pbc[3 * g] * xprd;
    const double y = sx[3 * sb + 1] + pbc[3 * g + 1] * yprd;
    const double z = sx[3 * sb + 2] + pbc[3 * g + 2] * zprd;
    asm volatile("" ::"v"(x), "v"(y), "v"(z), "v"(tg));  // (the compiler would sink the loads)
    if (cjj >= na) return;
    sx[3 * sg]     = x;
    sx[3 * sg + 1] = y;
    sx[3 * sg + 2] = z;
    if (sr) sr[sg] = make_rec(x, y, z, tg, rg, status);
}

// ---------------------------------------------------------------- binJClusters (neighbor.c:1929-2094)
// The bin of ghost cluster k (binJClusters' rule for its atoms), -1 for an empty cluster or
// (flagged) one outside the grid.
__device__ inline int ghost_bin_of(int ncj, int k, int N, const CGrid& g,
    const int* __restrict__ cj_natoms, const double* __restrict__ sx,
    const int* __restrict__ pbc, int* __restrict__ status)
{
    int cj = ncj + k, na = cj_natoms[cj];
    if (na <= 0) return -1;
    int ix, iy;
    c2b2d(g, sx[3 * cj * N], sx[3 * cj * N + 1], &ix, &iy);
    ix = max(min(ix, g.mbinx - 1), 0);
    iy = max(min(iy, g.mbiny - 1), 0);
    for (int cjj = 1; cjj < na; cjj++) {
        int nix, niy;
        int s = cj * N + cjj;
        c2b2d(g, sx[3 * s], sx[3 * s + 1], &nix, &niy);
        nix = max(min(nix, g.mbinx - 1), 0);
        niy = max(min(niy, g.mbiny - 1), 0);
        if (pbc[3 * k] > 0 && ix > nix) ix = nix;
        if (pbc[3 * k] < 0 && ix < nix) ix = nix;
        if (pbc[3 * k + 1] > 0 && iy > niy) iy = niy;
        if (pbc[3 * k + 1] < 0 && iy < niy) iy = niy;
    }
    int b = iy * g.mbinx + ix + 1;
    if (b < 0 || b >= g.mbins) {
        atomicOr(status, ST_BIN_RANGE);
        return -1;
    }
    return b;
}

__global__ void k_cp_ghost_bin(int ncj, int nghost, int N, CGrid g, const int* __restrict__ cj_natoms,
    const double* __restrict__ sx, const int* __restrict__ pbc, int* __restrict__ gbin,
    int* __restrict__ gcnt, int* __restrict__ status)
{
    const int k    = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int b    = k < nghost ? ghost_bin_of(ncj, k, N, g, cj_natoms, sx, pbc, status) : -1;
    if (k < nghost) gbin[k] = b;
    // one atomic per wavefront and bin (consecutive ghosts mostly share a bin)
    unsigned long long act = __ballot(b >= 0);
    while (act) {
        const int l                  = __ffsll((long long)act) - 1;
        const int lb                 = __shfl(b, l, 64);
        const unsigned long long grp = __ballot(b == lb);
        if (lane == l) atomicAdd(&gcnt[lb], __popcll(grp));
        act &= ~grp;
    }
}


__global__ void k_cp_jbin_count(int mbins, int M, int N, const int* __restrict__ ncl,
    const int* __restrict__ gcnt, int* __restrict__ jcnt)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= mbins) return;
    jcnt[b] = (M == N ? ncl[b] : ncl[b] / 2) + gcnt[b];
}

// k_cp_jbin_count between the two bin scans it sits between, in one block: the ghost clusters'
// bin starts (gstart, cursors gcnt cleared for the scatter), the j-cluster bin starts (jstart).
__global__ void __launch_bounds__(ONE_T) k_cp_jbin_starts(int mbins, int M, int N,
    const int* __restrict__ ncl, int* __restrict__ gcnt, int* __restrict__ gstart,
    int* __restrict__ jstart)
{
    __shared__ int sh[33];
    const int per = (mbins + ONE_T - 1) / ONE_T;
    const int b0  = threadIdx.x * per;
    int g[ONE_I];
    int gsum = 0, jsum = 0;
#pragma unroll
    for (int k = 0; k < ONE_I; k++) {
        const bool in = k < per && b0 + k < mbins;
        g[k]          = in ? gcnt[b0 + k] : 0;
        gsum += g[k];
        jsum += in ? (M == N ? ncl[b0 + k] : ncl[b0 + k] / 2) + g[k] : 0;
    }
    int gtot, jtot;
    int goff = mdg::block_excl_scan(gsum, sh, &gtot);
    int joff = mdg::block_excl_scan(jsum, sh, &jtot);
#pragma unroll
    for (int k = 0; k < ONE_I; k++) {
        if (k < per && b0 + k < mbins) {
            gstart[b0 + k] = goff;
            jstart[b0 + k] = joff;
            gcnt[b0 + k]   = 0;
            goff += g[k];
            joff += (M == N ? ncl[b0 + k] : ncl[b0 + k] / 2) + g[k];
        }
    }
    if (threadIdx.x == 0) {
        gstart[mbins] = gtot;
        jstart[mbins] = jtot;
    }
}

__global__ void k_cp_ghost_scatter(int nghost, const int* __restrict__ gbin,
    const int* __restrict__ gstart, int* __restrict__ gcur, int* __restrict__ glist)
{
    // as k_cp_bin_scatter: one atomic per wavefront and bin (the order inside a bin is the
    // merge's to fix)
    const int k    = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int b    = k < nghost ? gbin[k] : -1;
    unsigned long long act = __ballot(b >= 0);
    while (act) {
        const int l                  = __ffsll((long long)act) - 1;
        const int lb                 = __shfl(b, l, 64);
        const unsigned long long grp = __ballot(b == lb);
        int base                     = 0;
        if (lane == l) base = atomicAdd(&gcur[lb], __popcll(grp));
        base = __shfl(base, l, 64);
        if (b == lb) glist[gstart[lb] + base + __popcll(grp & lt)] = k;
        act &= ~grp;
    }
}

// One wavefront per bin: locals (in cluster order) then ghosts, stably ordered by bbminz --
// the order the reference's sorted insertion of ghosts produces.  Ties go to the smaller
// j-cluster index (locals before ghosts, ghosts in creation order).  A bin's local clusters
// take its z-sorted atoms in order, so their keys (their first atoms' z) come sorted: each
// local's place is then its own index plus the ghosts below it, and a ghost's the locals at or
// below it plus the ghosts before it (interior bins, without ghosts, are a copy).  Unsorted
// locals (checked) fall back to counting every entry below each entry.
__global__ void __launch_bounds__(64) k_cp_jbin_merge(int M, int N, int ncj,
    const int* __restrict__ ncl, const int* __restrict__ cstart, const int* __restrict__ gstart,
    const int* __restrict__ glist, const double* __restrict__ cj_bbox,
    const int* __restrict__ jstart, int* __restrict__ jlist, double* __restrict__ jkey,
    int* __restrict__ status)
{
    __shared__ int s_cj[JBIN_CAP];
    __shared__ double s_key[JBIN_CAP];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int nloc = M == N ? ncl[b] : ncl[b] / 2;
    const int j0   = M == N ? cstart[b] : cstart[b] / 2;
    const int g0 = gstart[b], ng = gstart[b + 1] - g0;
    const int n  = nloc + ng;
    if (n > JBIN_CAP) {
        if (lane == 0) atomicOr(status, ST_JBIN_CAP);
        return;
    }
    for (int e = lane; e < n; e += 64) {
        int cj    = e < nloc ? j0 + e : ncj + glist[g0 + e - nloc];
        s_cj[e]   = cj;
        s_key[e]  = cj_bbox[6 * cj + 4];
    }
    __syncthreads();
    bool unsorted = false;
    for (int e = lane; e + 1 < nloc; e += 64)
        unsorted = unsorted || !(s_key[e] <= s_key[e + 1]);
    unsorted = __ballot(unsorted) != 0;
    const int out = jstart[b];
    for (int e = lane; e < n; e += 64) {
        double k = s_key[e];
        int cj   = s_cj[e], r = 0;
        if (unsorted) {
            for (int f = 0; f < n; f++)
                r += (s_key[f] < k) || (s_key[f] == k && s_cj[f] < cj);
        } else if (e < nloc) {
            // locals before ghosts on equal keys (smaller index)
            r = e;
            for (int f = nloc; f < n; f++)
                r += s_key[f] < k;
        } else {
            for (int f = 0; f < nloc; f++)
                r += s_key[f] <= k;
            for (int f = nloc; f < n; f++)
                r += (s_key[f] < k) || (s_key[f] == k && s_cj[f] < cj);
        }
        jlist[out + r] = cj;
#pragma unroll
        for (int q = 0; q < 6; q++)
            jkey[6 * (out + r) + q] = cj_bbox[6 * cj + q];
    }
}

// Coarse z index of every column: zidx[b * NZB + k] is the first position of column b (sorted by
// bbminz) whose key is at least z0 + k dz.  The list build reads two entries per stencil column
// instead of running two binary searches.
constexpr int NZB = 256;
// Bucket of z, clamped to [-4, NZB + 4] (NaN to -4) before the conversion to int.
__device__ inline int zbucket(double z, double z0, double inv_dz)
{
    const double t = floor((z - z0) * inv_dz);
    if (!(t >= -4.0)) return -4;
    return t > (double)(NZB + 4) ? NZB + 4 : (int)t;
}
__global__ void k_cp_zindex(int mbins, const int* __restrict__ jstart,
    const double* __restrict__ jkey, double z0, double dz, int* __restrict__ zidx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mbins * NZB) return;
    const int b = t / NZB, k = t % NZB;
    const double zb = z0 + k * dz;
    int lo = jstart[b], hi = jstart[b + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jkey[6 * mid + 4] < zb) lo = mid + 1;
        else hi = mid;
    }
    zidx[t] = lo;
}

// ---------------------------------------------------------------- buildNeighborCPU (neighbor.c:282-766)
// Largest z-extent of any j-cluster (local or ghost): bounds bbmaxz from bbminz in the column
// search below.  Positive doubles order like their bit patterns.
__global__ void k_cp_jheight(int n, const double* __restrict__ cj_bbox,
    unsigned long long* __restrict__ hmax)
{
    __shared__ double s_h[BLOCK / 64];
    double h = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        h = fmax(h, cj_bbox[6 * k + 5] - cj_bbox[6 * k + 4]);
    for (int o = 32; o > 0; o >>= 1)
        h = fmax(h, __shfl_xor(h, o, 64));
    if ((threadIdx.x & 63) == 0) s_h[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++)
            h = fmax(h, s_h[w]);
        if (h > 0.0) atomicMax(hmax, (unsigned long long)__double_as_longlong(h));
    }
}


// One wavefront per i-cluster.  Each stencil column is sorted by bbminz, so the j-clusters
// whose z-gap to the i-cluster alone exceeds the list cutoff form a prefix and a suffix of the
// column; the column's coarse z index cuts most of them off (a j-cluster's bbmaxz is at most its
// bbminz plus the largest cluster height).  Dropped candidates are exactly ones the
// bounding-box test would reject, so the list is the reference's, in the reference's order
// (stencil columns in order, clusters in column order).  The remaining candidates of up to 64
// columns are flattened and handled in segments of NB_SEG, in three passes:
//   1. the bounding-box test of every candidate (its list entry and its box, stored in column
//      order, load side by side), 64 at a time: reject, accept, or "decided by the atoms";
//   2. the atom-pair test of the undecided ones only, queued in LDS and taken 64 at a time
//      (all N positions of a candidate load at once);
//   3. the accepted ones in candidate order, ballots giving each its place in the list.
// The undecided third or so of the candidates then costs one atom-pair pass instead of one per
// batch in which any lane needed it.
// 4x4: six wavefronts per SIMD (80 VGPRs, a few spilled) ran the 1M-atom build in 442-446 us
// against 465-469 us at the five the compiler picks by itself (91 VGPRs); at seven (72 VGPRs,
// more spilled) it took 557 us.  4x8 (143 VGPRs) forced to six spilled its eight j-positions
// and doubled its time (667 -> 1258 us); it keeps the compiler's choice.
constexpr int NB_SEG = 256;
// FMA: the atom-pair distance formed with fused multiply-adds (1, the 4xN SIMD reference's
// test) or without (0, the plain-C one), fixed at compile time so that the pair loop has no
// branch; 2: chosen by fma_dist at run time.  With the fixed form the 4x8 step ran 2-3 %
// faster; the 4x4 build keeps the run-time flag (fixed, the compiler keeps more values in
// scalar registers and spills them into the candidate loops: 1 % slower per step).
template <int M, int N, int FMA>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((M == 4 && N == 4) ? 6 : 1))) k_cp_build_nb(int nci, int rule, int fma_dist, int half,
    const int* __restrict__ stencil, int nstencil, int mbins, const int* __restrict__ ci_bin,
    const double* __restrict__ ci_bbox, const double* __restrict__ jbox,
    const int* __restrict__ jstart, const int* __restrict__ jlist,
    const int* __restrict__ zidx, double z0, double inv_dz, const double* __restrict__ sx,
    double cutneighsq, double R, double rbb_sq, const unsigned long long* __restrict__ hmax_bits,
    int maxneighs, int* __restrict__ nb, int* __restrict__ numneigh, int* __restrict__ nmasked,
    int* __restrict__ status)
{
    enum : unsigned char { C_REJECT = 0, C_ACCEPT = 1, C_ATOMS = 2 };
    __shared__ int s_st[1024];
    __shared__ int s_mpos[BLOCK / 64][4];
    __shared__ int s_cj[BLOCK / 64][NB_SEG];
    __shared__ unsigned char s_fl[BLOCK / 64][NB_SEG];
    __shared__ short s_q[BLOCK / 64][NB_SEG];
    __shared__ int s_m[BLOCK / 64][NB_SEG];  // the segment's candidates as column entries
    for (int k = threadIdx.x; k < nstencil; k += blockDim.x)
        s_st[k] = stencil[k];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ci = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64) + w);
    if (ci >= nci) return;
    const int ci_cj0 = (M == N) ? ci : ci >> 1;
    const int ibin   = ci_bin[ci];
    double ib[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
        ib[k] = ci_bbox[6 * ci + k];
    double xi[M][3];  // the i-cluster's slots (empty ones sit at far sentinel positions)
#pragma unroll
    for (int cii = 0; cii < M; cii++)
#pragma unroll
        for (int d = 0; d < 3; d++)
            xi[cii][d] = sx[3 * (ci * M + cii) + d];
    // search window in bbminz (R: the list cutoff with a margin, from the host; the margin only
    // widens the window)
    const double hmax = __longlong_as_double((long long)*hmax_bits);
    const double zlo  = ib[4] - R - hmax * (1.0 + 1e-12) - 1e-9;
    const double zhi  = ib[5] + R;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int* out = nb + (size_t)ci * maxneighs;
    int n = 0, nm = 0;
    for (int s0 = 0; s0 < nstencil; s0 += 64) {
        const int s = s0 + lane;
        int lo = 0, cnt = 0;
        if (s < nstencil) {
            const int jb = ibin + s_st[s];
            if (jb < 0 || jb >= mbins) {
                atomicOr(status, ST_BIN_RANGE);
            } else {
                // bucket boundaries one below zlo and one above zhi (a rounding of the bucket
                // index can only move a boundary by one): the window only widens
                const int kl = zbucket(zlo, z0, inv_dz) - 1;
                const int kh = zbucket(zhi, z0, inv_dz) + 2;
                lo           = kl < 0 ? jstart[jb] : (kl >= NZB ? jstart[jb + 1] : zidx[jb * NZB + kl]);
                const int hi = kh >= NZB ? jstart[jb + 1] : (kh < 0 ? jstart[jb] : zidx[jb * NZB + kh]);
                cnt          = hi > lo ? hi - lo : 0;
            }
        }
        // exclusive scan of the per-column counts
        int inc = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int v = __shfl_up(inc, o, 64);
            if (lane >= o) inc += v;
        }
        const int total = __shfl(inc, 63, 64);
        const int pre   = inc - cnt;
        for (int g0 = 0; g0 < total; g0 += NB_SEG) {
            const int segn = min(NB_SEG, total - g0);
            int nq         = 0;
            // the segment's candidates: candidate pre + i of this lane's column (i < cnt) is its
            // entry lo + i; each column lane writes those inside the segment, once, instead of
            // every candidate searching for its column
            {
                const int t1 = min(pre + cnt, g0 + segn);
                for (int t = max(pre, g0); t < t1; t++)
                    s_m[w][t - g0] = lo + (t - pre);
            }
            // one wavefront: its LDS operations complete in order
            __builtin_amdgcn_wave_barrier();
            // 1. bounding-box tests
            for (int b0 = 0; b0 < segn; b0 += 64) {
                const int k = b0 + lane;
                unsigned char fl = C_REJECT;
                if (k < segn) {
                    const int m = s_m[w][k];
                    // the box in column order, loaded beside the entry rather than after it
                    const double* jbb = &jbox[6 * m];
                    double d2         = 0.0;
#pragma unroll
                    for (int d = 2; d >= 0; d--) {
                        double dl  = ib[2 * d] - jbb[2 * d + 1];
                        double dh  = jbb[2 * d] - ib[2 * d + 1];
                        double dm0 = fmax(fmax(dl, dh), 0.0);
                        d2 += dm0 * dm0;
                    }
                    const int cj = jlist[m];
                    if (!(half && ci_cj0 > cj) && d2 < cutneighsq)
                        fl = d2 < rbb_sq ? C_ACCEPT : C_ATOMS;
                    s_cj[w][k] = cj;
                    s_fl[w][k] = fl;
                }
                const unsigned long long qb = __ballot(fl == C_ATOMS);
                if (fl == C_ATOMS) s_q[w][nq + __popcll(qb & lt)] = (short)k;
                nq += __popcll(qb);
            }
            // one wavefront: its LDS operations complete in order
            __builtin_amdgcn_wave_barrier();
            // 2. atom-pair tests of the undecided candidates
            for (int q0 = 0; q0 < nq; q0 += 64) {
                if (q0 + lane < nq) {
                    const int k  = s_q[w][q0 + lane];
                    const int cj = s_cj[w][k];
                    // any atom pair inside the cutoff: all N positions load at once
                    double xj[N][3];
#pragma unroll
                    for (int cjj = 0; cjj < N; cjj++)
#pragma unroll
                        for (int d = 0; d < 3; d++)
                            xj[cjj][d] = sx[3 * (cj * N + cjj) + d];
                    bool pass = false;
#pragma unroll
                    for (int cii = 0; cii < M; cii++)
#pragma unroll
                        for (int cjj = 0; cjj < N; cjj++) {
                            double dx = xi[cii][0] - xj[cjj][0];
                            double dy = xi[cii][1] - xj[cjj][1];
                            double dz = xi[cii][2] - xj[cjj][2];
                            const bool fm = FMA == 2 ? fma_dist != 0 : FMA == 1;
                            double rsq = fm ? fma(dx, dx, fma(dy, dy, dz * dz))
                                            : dx * dx + dy * dy + dz * dz;
                            pass = pass || rsq < cutneighsq;
                        }
                    s_fl[w][k] = pass ? C_ACCEPT : C_REJECT;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // 3. the accepted candidates in order
            for (int b0 = 0; b0 < segn; b0 += 64) {
                const int k       = b0 + lane;
                const bool pass   = k < segn && s_fl[w][k] == C_ACCEPT;
                const int cj      = pass ? s_cj[w][k] : -1;
                const bool masked = pass && imask_of(rule, ci, cj) != IMASK_ALL;
                const unsigned long long bal  = __ballot(pass);
                const unsigned long long mbal = __ballot(masked);
                if (pass) {
                    int pos = n + __popcll(bal & lt);
                    if (pos < maxneighs) out[pos] = cj;
                    if (masked) {
                        int kk = nm + __popcll(mbal & lt);
                        if (kk < 4) s_mpos[w][kk] = pos;
                    }
                }
                n += __popcll(bal);
                nm += __popcll(mbal);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // single-lane epilogue: masked entries move to the front, in order of arrival.  In a full
    // list the only masked entry is the i-cluster's own j-cluster, so for every cluster with
    // atoms it becomes entry 0: k_cp_force_f32_full tests for the distance-0 pair in its first
    // iteration only and relies on this (checked by test_own_jcluster_is_entry_zero)
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        numneigh[ci] = n;
        nmasked[ci]  = nm;
        if (n > maxneighs || nm > 4) {
            atomicOr(status, ST_NB_OVERFLOW);
            atomicMax(&status[1], n);
        } else {
            for (int k = 0; k < nm; k++) {
                int p   = s_mpos[w][k];
                int t   = out[k];
                out[k]  = out[p];
                out[p]  = t;
            }
        }
    }
}

// ---------------------------------------------------------------- pruneNeighborCPU (neighbor.c:1099-1264)
// One wavefront per i-cluster.  Whether a listed j-cluster still has an atom pair inside the
// list cutoff is a property of the entry alone, so the lanes test all entries first (64 at a
// time, all M x N pairs each; empty slots sit at far sentinel positions).  The reference then
// walks the list once: an entry out of range is overwritten by the current last entry, which is
// tested in its place, and every removal at a position inside the masked prefix shortens that
// prefix.  One lane replays exactly that walk on the entries staged in LDS.  The i-cluster's
// own j-cluster (entry 0 of a full list) holds pairs at distance 0, so it is never dropped and
// never moves: k_cp_force_f32_full's first-iteration-only distance-0 test relies on that.
constexpr int PRUNE_CAP = 512;  // list entries per i-cluster staged in LDS
template <int M, int N>
__global__ void __launch_bounds__(BLOCK) k_cp_prune(int nci, int fma_dist,
    const double* __restrict__ sx, double cutneighsq, int maxneighs, int* __restrict__ nb,
    int* __restrict__ numneigh, int* __restrict__ nmasked, int* __restrict__ status)
{
    __shared__ int s_e[BLOCK / 64][PRUNE_CAP];
    __shared__ unsigned char s_in[BLOCK / 64][PRUNE_CAP];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ci = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64) + w);
    if (ci >= nci) return;
    const int nn = numneigh[ci];
    if (nn > PRUNE_CAP) {
        if (lane == 0) atomicOr(status, ST_NB_OVERFLOW);
        return;
    }
    double xi[M][3];
#pragma unroll
    for (int cii = 0; cii < M; cii++)
#pragma unroll
        for (int d = 0; d < 3; d++)
            xi[cii][d] = sx[3 * (ci * M + cii) + d];
    int* lst = nb + (size_t)ci * maxneighs;
    for (int k = lane; k < nn; k += 64) {
        const int cj = lst[k];
        double xj[N][3];
#pragma unroll
        for (int cjj = 0; cjj < N; cjj++)
#pragma unroll
            for (int d = 0; d < 3; d++)
                xj[cjj][d] = sx[3 * (cj * N + cjj) + d];
        bool in = false;
#pragma unroll
        for (int cii = 0; cii < M; cii++)
#pragma unroll
            for (int cjj = 0; cjj < N; cjj++) {
                double dx = xi[cii][0] - xj[cjj][0];
                double dy = xi[cii][1] - xj[cjj][1];
                double dz = xi[cii][2] - xj[cjj][2];
                double rsq = fma_dist ? fma(dx, dx, fma(dy, dy, dz * dz))
                                      : dx * dx + dy * dy + dz * dz;
                in = in || rsq < cutneighsq;
            }
        s_e[w][k]  = cj;
        s_in[w][k] = in ? 1 : 0;
    }
    __builtin_amdgcn_wave_barrier();
    __shared__ int s_n[BLOCK / 64];
    if (lane == 0) {
        int n = nn, nm = nmasked[ci], k = 0;
        while (k < n) {
            if (s_in[w][k]) {
                k++;
            } else {
                n--;
                if (k < nm) nm--;
                s_e[w][k]  = s_e[w][n];
                s_in[w][k] = s_in[w][n];
            }
        }
        numneigh[ci] = n;
        nmasked[ci]  = nm;
        s_n[w]       = n;
    }
    __builtin_amdgcn_wave_barrier();
    const int n = s_n[w];
    for (int k = lane; k < n; k += 64)
        lst[k] = s_e[w][k];
}

// ---------------------------------------------------------------- energy and virial sums
// Every wavefront of a force launch leaves its energy and virial sums in its own pair of slots
// (cleared before the launch, so a wavefront that has no i-cluster leaves zeros); k_cp_sum_ev
// adds the slots in a fixed order.  No atomics: one FP64 atomic pair per wavefront on one
// address cost 1.4 ms per 1M-atom call, and the sum now does not depend on timing.
__device__ __forceinline__ void ev_store(double e, double w, double* __restrict__ part)
{
    for (int o = 32; o > 0; o >>= 1) {
        e += __shfl_xor(e, o, 64);
        w += __shfl_xor(w, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        const size_t k = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        part[2 * k]     = e;
        part[2 * k + 1] = w;
    }
}

// ev[0..1] = the sums of the nw slot pairs: per thread a strided sum, then a fixed tree
constexpr int EVR_T = 1024;
__global__ void __launch_bounds__(EVR_T) k_cp_sum_ev(size_t nw, const double* __restrict__ part,
    double* __restrict__ ev)
{
    __shared__ double s[2][EVR_T];
    const int t = threadIdx.x;
    double e = 0.0, w = 0.0;
    for (size_t k = t; k < nw; k += EVR_T) {
        e += part[2 * k];
        w += part[2 * k + 1];
    }
    s[0][t] = e;
    s[1][t] = w;
    __syncthreads();
    for (int o = EVR_T / 2; o > 0; o >>= 1) {
        if (t < o) {
            s[0][t] += s[0][t + o];
            s[1][t] += s[1][t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        ev[0] = s[0][0];
        ev[1] = s[1][0];
    }
}

// ---------------------------------------------------------------- FP64 force, reference arithmetic
template <int M, int N, bool HALF, bool EV>
__global__ void __launch_bounds__(BLOCK) k_cp_force_f64(int nci, int ncj, int ntypes, int uniform,
    Tab64 tab, const double* __restrict__ sx, const int* __restrict__ st,
    const int* __restrict__ nb, const int* __restrict__ numneigh, int maxneighs,
    double* __restrict__ f, double* __restrict__ ev)
{
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    const int ci = si / M;
    double e = 0.0, w = 0.0;
    if (ci < nci) {
        const int ci_cj0 = (M == N) ? ci : ci >> 1;
        const double xt = sx[3 * si], yt = sx[3 * si + 1], zt = sx[3 * si + 2];
        const int ti    = st[si];
        const int* lst  = nb + (size_t)ci * maxneighs;
        const int nn    = numneigh[ci];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int k = 0; k < nn; k++) {
            const int cj = lst[k];
            double fix = 0, fiy = 0, fiz = 0;
            for (int cjj = 0; cjj < N; cjj++) {
                const int sj = cj * N + cjj;
                bool cond    = HALF ? (ci_cj0 != cj || si < sj) : (ci_cj0 != cj || si != sj);
                if (!cond) continue;
                double delx = xt - sx[3 * sj];
                double dely = yt - sx[3 * sj + 1];
                double delz = zt - sx[3 * sj + 2];
                double rsq  = delx * delx + dely * dely + delz * delz;
                int tij     = uniform ? 0 : ti * ntypes + st[sj];
                if (rsq < tab.cutf[tij]) {
                    double sr2   = 1.0 / rsq;
                    double sr6   = sr2 * sr2 * sr2 * tab.sig6[tij];
                    double force = 48.0 * sr6 * (sr6 - 0.5) * sr2 * tab.eps[tij];
                    if (HALF && cj < ncj) {
                        atomicAdd(&f[3 * sj], -(delx * force));
                        atomicAdd(&f[3 * sj + 1], -(dely * force));
                        atomicAdd(&f[3 * sj + 2], -(delz * force));
                    }
                    fix += delx * force;
                    fiy += dely * force;
                    fiz += delz * force;
                    if (EV) {
                        double wgt = (HALF && cj < ncj) ? 1.0 : 0.5;
                        e += wgt * 4.0 * tab.eps[tij] * sr6 * (sr6 - 1.0);
                        w += wgt * rsq * force;
                    }
                }
            }
            fx += fix;
            fy += fiy;
            fz += fiz;
        }
        if (HALF) {
            atomicAdd(&f[3 * si], fx);
            atomicAdd(&f[3 * si + 1], fy);
            atomicAdd(&f[3 * si + 2], fz);
        } else {
            f[3 * si]     = fx;
            f[3 * si + 1] = fy;
            f[3 * si + 2] = fz;
        }
    }
    if (EV) ev_store(e, w, ev);
}

// ---------------------------------------------------------------- FP32-mixed force kernel
// f(integral_constant<int, I>) for I = B .. E-1 in order, until f returns false
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        if (f(std::integral_constant<int, B>{})) static_for<B + 1, E>(f);
    }
}

__device__ inline void kahan_add(float& s, float& c, float t)
{
    float y  = t - c;
    float tt = s + y;
    c        = (tt - s) - y;
    s        = tt;
}

// Lane i of a 16-lane DPP row reads a lane of the same row chosen by CTRL (a DPP move, no LDS):
// DPP_XOR1 / DPP_XOR2 (quad permutations: lane i^1, i^2), DPP_HALF_MIRROR (7-i inside each half
// row), DPP_MIRROR (15-i).
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_MIRROR = 0x140, DPP_HALF_MIRROR = 0x141;
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// (s, c) += (s2, c2) with the rounding error of s + s2 kept (TwoSum).
__device__ inline void two_sum_merge(float& s, float& c, float s2, float c2)
{
    float t  = s + s2;
    float bp = t - s;
    float er = (s - (t - bp)) + (s2 - bp);
    s        = t;
    c        = c + c2 + er;
}

// Pair set.  A pair is inside when its FP64 r^2 (the FP64 kernel's own expression) is below the
// cutoff: a pair whose r differs from r_c by less than an FP32 rounding would otherwise be inside
// for one kernel and outside for the other and change two forces by the full F(r_c).  The FP32
// r^2 of the rounded displacement is within 3e-7 of it (relative), so it decides every pair
// outside a +-2e-6 band around r_c^2; a wavefront evaluates the FP64 r^2 only when one of its
// pairs falls into the band.  Full lists exclude the self pair by r^2 > 0 (the same slot, the
// only pair at distance 0); padded slots and list entries past the end (the dummy j-cluster)
// sit at far sentinel positions and never pass.
constexpr float CUT_BAND = 2e-6f;

struct JPos {
    int cj, sj;
    double x, y, z;
};

template <int N, bool HALF, bool EV, bool UNIFORM, bool COMP>
struct CPAcc {
    float fx = 0.f, fy = 0.f, fz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, e = 0.f, w = 0.f;

    __device__ __forceinline__ void pair(const JPos& J, double xi, double yi, double zi, int si,
        int ci_cj0, int ncj, int a, int ti, int ntypes, const int* __restrict__ st,
        const float (*s_tab)[MAXT2], const double* s_cut, float c12u, float nc6u, float eps4u,
        float sig6u, float clo_u, float chi_u, double cut2u, float* __restrict__ f)
    {
        const double dxd = xi - J.x;
        const double dyd = yi - J.y;
        const double dzd = zi - J.z;
        const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
        const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        float c12 = c12u, nc6 = nc6u, eps4 = eps4u, sig6 = sig6u, clo = clo_u, chi = chi_u;
        double cut2 = cut2u;
        if (!UNIFORM) {
            const int tij = ti * ntypes + st[J.sj];
            c12           = s_tab[0][tij];
            nc6           = -s_tab[1][tij];
            eps4          = s_tab[2][tij];
            sig6          = s_tab[3][tij];
            cut2          = s_cut[tij];
            clo           = (float)cut2 * (1.0f - CUT_BAND);
            chi           = (float)cut2 * (1.0f + CUT_BAND);
        }
        bool in          = rsq < clo;
        const bool below = rsq < chi;
        // lanes in the band: below and not in, from the two lane masks
        if (__builtin_expect((__ballot(below) & ~__ballot(in)) != 0, 0)) {
            const double rsqd = dxd * dxd + dyd * dyd + dzd * dzd;
            if (below && !in) in = rsqd < cut2;
        }
        if (HALF) in = in && !(J.cj == ci_cj0 && !(si < J.sj));
        else in = in && rsq > 0.0f;
        const float r2i = __builtin_amdgcn_rcpf(rsq);  // no Newton step (see the full-list kernel)
        const float r6i = r2i * r2i * r2i;
        float fpair     = r6i * fmaf(c12, r6i, nc6) * r2i;
        fpair           = in ? fpair : 0.0f;
        if (EV) {
            const float sr6 = r6i * sig6;
            const float wgt = (HALF && J.cj < ncj) ? 1.0f : 0.5f;
            e += in ? wgt * eps4 * sr6 * (sr6 - 1.0f) : 0.0f;
            w += wgt * rsq * fpair;
        }
        if (COMP) {
            kahan_add(fx, cx, dx * fpair);
            kahan_add(fy, cy, dy * fpair);
            kahan_add(fz, cz, dz * fpair);
        } else {
            fx = fmaf(dx, fpair, fx);
            fy = fmaf(dy, fpair, fy);
            fz = fmaf(dz, fpair, fz);
        }
        if (HALF) {
            // j-force: reduce over the four i-lanes that share this j-atom, one atomic per atom
            float jx = -dx * fpair, jy = -dy * fpair, jz = -dz * fpair;
            jx += __shfl_xor(jx, 1, 64);
            jy += __shfl_xor(jy, 1, 64);
            jz += __shfl_xor(jz, 1, 64);
            jx += __shfl_xor(jx, 2, 64);
            jy += __shfl_xor(jy, 2, 64);
            jz += __shfl_xor(jz, 2, 64);
            if (a == 0 && J.cj < ncj && (jx != 0.f || jy != 0.f || jz != 0.f)) {
                atomicAdd(&f[3 * J.sj], jx);
                atomicAdd(&f[3 * J.sj + 1], jy);
                atomicAdd(&f[3 * J.sj + 2], jz);
            }
        }
    }
};

template <int N, bool HALF, bool EV, bool UNIFORM, bool COMP>
__global__ void __launch_bounds__(BLOCK) k_cp_force_f32(int nci, int ncj, int dummy_cj,
    int ntypes, Tab32 tab, const double* __restrict__ sx, const int* __restrict__ st,
    const int* __restrict__ nb, const int* __restrict__ numneigh,
    const int* __restrict__ ci_natoms, int maxneighs, float* __restrict__ f,
    double* __restrict__ ev)
{
    constexpr int CPI = 16 / N;  // j-clusters per iteration
    __shared__ float s_tab[4][MAXT2];
    __shared__ double s_cut[MAXT2];
    if (!UNIFORM) {
        for (int k = threadIdx.x; k < ntypes * ntypes; k += blockDim.x) {
            s_tab[0][k] = tab.c12[k];
            s_tab[1][k] = tab.c6[k];
            s_tab[2][k] = tab.eps4[k];
            s_tab[3][k] = tab.sig6[k];
            s_cut[k]    = tab.cut2d[k];
        }
        __syncthreads();
    }
    // XCD-aware block order: the 8 XCDs receive blocks round-robin; give each XCD a contiguous
    // range of i-clusters so that neighbouring i-clusters share that XCD's L2.  The grid is a
    // multiple of 8 blocks, which makes the remapping a permutation.
    const int per = gridDim.x >> 3;
    const int lb  = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    // wave-uniform values live in scalar registers, so loop control stays scalar
    const int ci = __builtin_amdgcn_readfirstlane(lb * (BLOCK / 64) + (threadIdx.x >> 6));
    if (ci >= nci) return;
    const int lane = threadIdx.x & 63;
    const int a = lane & 3, s = lane >> 2;
    const int g = s / N, jj = s % N;
    const int g4 = g << 2;
    const unsigned jj24 = 24u * (unsigned)jj;
    const int si     = ci * 4 + a;
    const bool ilive = a < ci_natoms[ci];
    const double xi = sx[3 * si], yi = sx[3 * si + 1], zi = sx[3 * si + 2];
    const int ci_cj0 = (N == 4) ? ci : ci >> 1;
    const int ti     = UNIFORM ? 0 : st[si];
    const int nn     = __builtin_amdgcn_readfirstlane(numneigh[ci]);
    const int* lst   = nb + (size_t)ci * maxneighs;
    const float c12u = tab.c12[0], eps4u = tab.eps4[0], sig6u = tab.sig6[0];
    // -c6 in a vector register once (a VALU op reads at most one scalar operand)
    float nc6u = -tab.c6[0];
    asm volatile("" : "+v"(nc6u));
    const double cut2u = tab.cut2d[0];
    const float clo_u  = (float)cut2u * (1.0f - CUT_BAND), chi_u = (float)cut2u * (1.0f + CUT_BAND);
    CPAcc<N, HALF, EV, UNIFORM, COMP> acc;
    // The list of this i-cluster is held in registers, 128 entries at a time (two per lane), and
    // handed out with cross-lane reads (in every lane: a lane outside a branch supplies no data);
    // entries past the end read as the dummy j-cluster.  Two register sets alternate: the
    // positions of the next iteration load while the current ones are used.  Every load inside
    // the loop is unconditional (a fetch past the chunk reads valid entries and is discarded),
    // so the compiler's wait before a use waits only for that use's own loads.
    auto use = [&](const JPos& J) {
        acc.pair(J, xi, yi, zi, si, ci_cj0, ncj, a, ti, ntypes, st, s_tab, s_cut, c12u, nc6u,
            eps4u, sig6u, clo_u, chi_u, cut2u, f);
    };
    for (int c0 = 0; c0 < nn; c0 += 128) {
        const int e0   = c0 + lane < nn ? lst[c0 + lane] : dummy_cj;
        const int e1   = c0 + 64 + lane < nn ? lst[c0 + 64 + lane] : dummy_cj;
        const int cend = min(nn, c0 + 128);
        auto fetch = [&](int kb, JPos& J) {
            const int src = (kb & 64) ? e1 : e0;  // kb is wave-uniform
            // kb is a multiple of CPI and g < CPI, so (kb & 63) + g is the source lane
            J.cj = __builtin_amdgcn_ds_bpermute(((kb & 63) << 2) + g4, src);
            J.sj = J.cj * N + jj;
            // 32-bit byte offset from the base, one 24-bit multiply-add (the host keeps every
            // j-cluster index below 2^24 and every offset below 2^32)
            const double* p = reinterpret_cast<const double*>(
                reinterpret_cast<const char*>(sx) + (__umul24(J.cj, 24u * N) + jj24));
            J.x = p[0];
            J.y = p[1];
            J.z = p[2];
        };
        JPos A, B;
        fetch(c0, A);
        for (int k0 = c0; k0 < cend;) {
            fetch(k0 + CPI, B);
            use(A);
            k0 += CPI;
            if (k0 >= cend) break;
            fetch(k0 + CPI, A);
            use(B);
            k0 += CPI;
        }
    }
    float fx = acc.fx, fy = acc.fy, fz = acc.fz, cx = acc.cx, cy = acc.cy, cz = acc.cz;
    const float e = acc.e, w = acc.w;
    // reduce the 16 j-slot lanes of each i-atom (lane bits 2..5)
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
        float ox = __shfl_xor(fx, o, 64), oy = __shfl_xor(fy, o, 64), oz = __shfl_xor(fz, o, 64);
        if (COMP) {
            float ocx = __shfl_xor(cx, o, 64), ocy = __shfl_xor(cy, o, 64),
                  ocz = __shfl_xor(cz, o, 64);
            two_sum_merge(fx, cx, ox, ocx);
            two_sum_merge(fy, cy, oy, ocy);
            two_sum_merge(fz, cz, oz, ocz);
        } else {
            fx += ox;
            fy += oy;
            fz += oz;
        }
    }
    if (s == 0 && ilive) {
        const float rx = COMP ? fx + cx : fx, ry = COMP ? fy + cy : fy, rz = COMP ? fz + cz : fz;
        if (HALF) {
            atomicAdd(&f[3 * si], rx);
            atomicAdd(&f[3 * si + 1], ry);
            atomicAdd(&f[3 * si + 2], rz);
        } else {
            f[3 * si]     = rx;
            f[3 * si + 1] = ry;
            f[3 * si + 2] = rz;
        }
    }
    if (EV) ev_store(e, w, ev);
}

// Full lists, j-atom lanes.  One wavefront works on four i-clusters: the 16 lanes of group g
// take i-cluster 4w+g, hold its four i-atoms, and each lane takes one j-atom per iteration
// (16/N j-clusters of the group's list per iteration) and computes that j-atom's pairs with all
// four i-atoms.  Every j-atom is loaded by one lane only, as one 16-byte record (make_rec), and
// serves four pairs; each lane keeps the forces of its group's four i-atoms and the 16 lanes of
// a group reduce them at the end.  The four lists of a wavefront run to the longest; shorter
// ones read the dummy j-cluster, whose slots sit at far sentinel positions.  Pair set, cutoff
// band and the arithmetic after the displacement are those of k_cp_force_f32.
//
// Displacements from the records:
// - one-type variants: the j-atom in the frame of the i-cluster's grid cell kI (the cell of its
//   bounding-box centre), y = (k_j - kI) h + o_j, once per cluster pair (the cell difference
//   exact, one rounding), the i-atoms the same way once per launch; a pair's displacement
//   z_a - y: three subtractions per pair;
// - type-table variants (the Kob-Andersen mixture, whose force error the frame's extra
//   rounding pushed past 1e-6): d = (k_a - k_j) h + (o_a - o_j) per pair, the cell difference
//   exact and one rounding at d, nine operations per pair (the cells held as FP32 values:
//   one conversion per j-atom, one subtraction per pair); the j type is the record's fourth
//   byte.
// The one-type variant is held to 80 VGPRs, six wavefronts per SIMD, the type-table variant to
// 96, five, both without spills; the energy and compensated variants keep their register count.
template <int N, bool EV, bool UNIFORM, bool COMP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((COMP || EV) ? 1 : (UNIFORM ? 6 : 5))))
k_cp_force_f32_full(int nci, int ncj, int dummy_cj,
    int ntypes, Tab32 tab, const double* __restrict__ sx, const float4* __restrict__ sr,
    const int* __restrict__ st, const int* __restrict__ nb, const int* __restrict__ numneigh,
    const int* __restrict__ ci_natoms, const double* __restrict__ ci_bbox, RecGrid rg,
    int maxneighs, float* __restrict__ f, double* __restrict__ ev, double* __restrict__ v,
    double dtforce, double* __restrict__ xn, float4* __restrict__ rn, double dt,
    int* __restrict__ status)
{
    constexpr int CPG    = 16 / N;  // j-clusters per group per iteration
    constexpr bool SPLIT = !UNIFORM;
    __shared__ float s_tab[4][MAXT2];
    __shared__ double s_cut[MAXT2];
    // per type pair, read with one 16-byte LDS load: c12, -c6 and the FP32 band edges
    __shared__ float4 s_pk[MAXT2];
    if (!UNIFORM) {
        for (int k = threadIdx.x; k < ntypes * ntypes; k += blockDim.x) {
            s_tab[0][k]  = tab.c12[k];
            s_tab[1][k]  = tab.c6[k];
            s_tab[2][k]  = tab.eps4[k];
            s_tab[3][k]  = tab.sig6[k];
            s_cut[k]     = tab.cut2d[k];
            const float c = (float)tab.cut2d[k];
            s_pk[k]      = make_float4(tab.c12[k], -tab.c6[k], c * (1.0f - CUT_BAND),
                     c * (1.0f + CUT_BAND));
        }
        __syncthreads();
    }
    // XCD-aware block order, as in k_cp_force_f32 (the grid is a multiple of 8 blocks)
    const int per  = gridDim.x >> 3;
    const int lb   = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int wv   = __builtin_amdgcn_readfirstlane(lb * (BLOCK / 64) + (threadIdx.x >> 6));
    if (wv * 4 >= nci) return;
    const int lane = threadIdx.x & 63, grp = lane >> 4, sub = lane & 15;
    const int jc = sub / N, jj = sub % N;
    const int ci    = wv * 4 + grp;
    const bool live = ci < nci;
    // the i-cluster's grid cell: that of the centre of its bounding box (as of the list build;
    // an empty cluster's box is inverted and its centre taken at 0)
    float kI[3] = { 0.f, 0.f, 0.f };
    if (!SPLIT) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double lo = live ? ci_bbox[6 * ci + 2 * d] : 0.0;
            const double hi = live ? ci_bbox[6 * ci + 2 * d + 1] : 0.0;
            const double c  = lo <= hi ? 0.5 * (lo + hi) : 0.0;
            kI[d]           = (float)(rint(c * rg.inv_h) + rg.bias);
        }
    }
    const float hf = rg.hf;
    // i-atoms from their own records, by the same operations as a j-atom's, so that the self
    // pair's displacement is exactly 0: one-type, z = (k_a - kI) h + o_a; type-table, the cell
    // index and the offset.  Groups past the last i-cluster sit far away (their force sums are
    // never written; far enough that no pair is inside a cutoff, near enough that r^2 stays
    // finite, so the energy variant's virial term r^2 * 0 stays 0).
    float zi[4][3], ki[4][3];
    int ti[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float4 r   = live ? sr[ci * 4 + a] : make_float4(1e15f, 1e15f, 1e15f, 0.f);
        const unsigned w = __float_as_uint(r.w);
        ki[a][0] = (float)(w & 0xffu);
        ki[a][1] = (float)((w >> 8) & 0xffu);
        ki[a][2] = (float)((w >> 16) & 0xffu);
        if (SPLIT) {
            zi[a][0] = r.x;
            zi[a][1] = r.y;
            zi[a][2] = r.z;
        } else {
            zi[a][0] = fmaf(ki[a][0] - kI[0], hf, r.x);
            zi[a][1] = fmaf(ki[a][1] - kI[1], hf, r.y);
            zi[a][2] = fmaf(ki[a][2] - kI[2], hf, r.z);
        }
        ti[a] = UNIFORM ? 0 : (int)(w >> 24);
        // type-table: the cell indices stay FP32 values (opaque to the compiler, which would
        // otherwise turn each pair's ki - kj back into a byte subtraction and a conversion)
        if (SPLIT) asm volatile("" : "+v"(ki[a][0]), "+v"(ki[a][1]), "+v"(ki[a][2]));
    }
    const int nn = live ? numneigh[ci] : 0;
    // the group's atom count, for the tail, loaded here rather than after the loop
    const int nat = live ? ci_natoms[ci] : 0;
    int nmax     = max(nn, __shfl_xor(nn, 16, 64));
    nmax         = max(nmax, __shfl_xor(nmax, 32, 64));
    nmax         = __builtin_amdgcn_readfirstlane(nmax);
    const int* lst = nb + (size_t)(live ? ci : 0) * maxneighs;
    const float c12u = tab.c12[0], eps4u = tab.eps4[0], sig6u = tab.sig6[0];
    float nc6u = -tab.c6[0];
    asm volatile("" : "+v"(nc6u));
    const double cut2u = tab.cut2d[0];
    const float clo_u  = (float)cut2u * (1.0f - CUT_BAND), chi_u = (float)cut2u * (1.0f + CUT_BAND);
    float fx[4] = {}, fy[4] = {}, fz[4] = {}, cx[4] = {}, cy[4] = {}, cz[4] = {};
    float e = 0.f, w = 0.f;
    struct JL {
        float4 r;
        unsigned off;  // byte offset of the record from sr: cj * 16 N + 16 jj
    };
    // SELF: the cluster pair may hold pairs at distance 0 (see the list loop below)
    auto use = [&](const JL& J, auto self) {
        constexpr bool SELF = decltype(self)::value;
        const unsigned wj = __float_as_uint(J.r.w);
        float kx = (float)(wj & 0xffu), ky = (float)((wj >> 8) & 0xffu),
              kz = (float)((wj >> 16) & 0xffu);
        // one conversion per j-atom, one FP32 subtraction per pair (see the i-atoms' cells)
        if (SPLIT) asm volatile("" : "+v"(kx), "+v"(ky), "+v"(kz));
        const int tj = UNIFORM ? 0 : (int)(wj >> 24);
        float yx = 0.f, yy = 0.f, yz = 0.f;
        if (!SPLIT) {
            yx = fmaf(kx - kI[0], hf, J.r.x);
            yy = fmaf(ky - kI[1], hf, J.r.y);
            yz = fmaf(kz - kI[2], hf, J.r.z);
        }
#pragma unroll
        for (int a = 0; a < 4; a++) {
            float dx, dy, dz;
            if (SPLIT) {
                dx = fmaf(ki[a][0] - kx, hf, zi[a][0] - J.r.x);
                dy = fmaf(ki[a][1] - ky, hf, zi[a][1] - J.r.y);
                dz = fmaf(ki[a][2] - kz, hf, zi[a][2] - J.r.z);
            } else {
                dx = zi[a][0] - yx;
                dy = zi[a][1] - yy;
                dz = zi[a][2] - yz;
            }
            const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            float c12 = c12u, nc6 = nc6u, eps4 = eps4u, sig6 = sig6u, clo = clo_u, chi = chi_u;
            const int tij = UNIFORM ? 0 : ti[a] * ntypes + tj;
            if (!UNIFORM) {
                const float4 pk = s_pk[tij];
                c12             = pk.x;
                nc6             = pk.y;
                clo             = pk.z;
                chi             = pk.w;
                if (EV) {
                    eps4 = s_tab[2][tij];
                    sig6 = s_tab[3][tij];
                }
            }
            bool in          = rsq < clo;
            const bool below = rsq < chi;
            if (__builtin_expect((__ballot(below) & ~__ballot(in)) != 0, 0)) {
                // FP64 re-check from the FP64 positions of both atoms, read again: the i-atom's
                // slot from the wavefront index and the lane id (computed here by volatile asm,
                // which keeps the compiler from hoisting the address out of the loop and holding
                // it in a register or spill slot for this rare branch), clamped to the last
                // i-cluster for lanes of groups past it
                int lr;
                asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0"
                             : "=v"(lr));
                const int cir     = min(wv * 4 + (lr >> 4), nci - 1);
                const double* pi  = sx + 3 * (cir * 4 + a);
                const double* pj  = sx + 3 * (J.off >> 4);  // slot cj * N + jj
                const double dxd  = pi[0] - pj[0];
                const double dyd  = pi[1] - pj[1];
                const double dzd  = pi[2] - pj[2];
                const double rsqd = dxd * dxd + dyd * dyd + dzd * dzd;
                const double cut2 = UNIFORM ? cut2u : s_cut[tij];
                if (below && !in) in = rsqd < cut2;
            }
            if (SELF || EV) in = in && rsq > 0.0f;  // the self pair is the only one at distance 0
            // v_rcp_f32 alone: with a Newton step the 1M-atom force error was 5.7e-7, without it
            // 6.0e-7 (measured with the rounded FP64 displacement); the step cost two of the ~25
            // instructions of a pair
            const float r2i = __builtin_amdgcn_rcpf(rsq);
            const float r6i = r2i * r2i * r2i;
            float fpair     = r6i * fmaf(c12, r6i, nc6) * r2i;
            fpair           = in ? fpair : 0.0f;
            if (COMP) {
                kahan_add(fx[a], cx[a], dx * fpair);
                kahan_add(fy[a], cy[a], dy * fpair);
                kahan_add(fz[a], cz[a], dz * fpair);
            } else {
                fx[a] = fmaf(dx, fpair, fx[a]);
                fy[a] = fmaf(dy, fpair, fy[a]);
                fz[a] = fmaf(dz, fpair, fz[a]);
            }
            if (EV) {
                const float sr6 = r6i * sig6;
                e += in ? 0.5f * eps4 * sr6 * (sr6 - 1.0f) : 0.0f;
                w += 0.5f * rsq * fpair;
            }
        }
    };
    // Entries of the group's list are held 32 at a time in two registers per lane (16 each) and
    // handed out with cross-lane reads inside the group; entries past a list's end read as the
    // dummy j-cluster.  Records of the next iteration load while the current ones are used, and
    // every load in the loop is unconditional (see k_cp_force_f32).
    // The list builder puts the entries that need the diagonal mask first, and in a full list
    // the only such entry is the j-cluster holding the i-cluster itself: of a cluster with atoms
    // it is entry 0, and an empty cluster has no entries.  Only the first iteration can meet a
    // pair at distance 0 and only it tests for one; elsewhere the test is dropped (one compare
    // and one scalar op per pair).  Without the test a padding i-slot whose FP32 offset equals
    // that of a padding j-slot (sentinels near 1e12 round to multiples of 2^16) gets a NaN
    // force, in accumulators of its own that are never written out; the energy variant, whose
    // sum would take the NaN, keeps the test everywhere.
    // Entries are handed out in chunks of 16, one per lane of the group, the next chunk's entry
    // loaded while the current chunk runs.  Entries are held as the byte offsets of their
    // j-cluster's records (cj * 16 N; the host keeps every slot offset below 2^32), so a lane
    // adds only its own record's 16 jj.  Records of the next iteration load while the current
    // ones are used (across chunk boundaries too), and every load in the loop is unconditional
    // (see k_cp_force_f32).
    const unsigned jj16 = 16u * (unsigned)jj;
    const int bp_base   = ((grp << 4) + jc) << 2;
    auto entry = [&](int k) { return 16u * N * (unsigned)(k < nn ? lst[k] : dummy_cj); };
    unsigned ec = entry(sub), en = entry(16 + sub);
    // kb: the wave-uniform index of the next iteration's first entry (a multiple of CPG)
    auto fetch = [&](int kb, JL& J) {
        const int kk       = kb & 15;
        const unsigned src = kk ? ec : en;  // kk == 0: the next chunk's first entry
        const unsigned cjo = (unsigned)__builtin_amdgcn_ds_bpermute(bp_base + (kk << 2), (int)src);
        J.off = cjo + jj16;
        J.r   = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sr) + J.off);
    };
    JL A, B;
    {
        const unsigned cjo = (unsigned)__builtin_amdgcn_ds_bpermute(bp_base, (int)ec);
        A.off = cjo + jj16;
        A.r   = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sr) + A.off);
    }
    int k0 = 0;
    auto advance = [&]() {  // after the iteration at k0, when k0 + CPG starts a new chunk
        k0 += CPG;
        if (!(k0 & 15)) {
            ec = en;
            en = entry(k0 + 16 + sub);
        }
    };
    fetch(CPG, B);
    use(A, std::true_type{});
    advance();
    while (k0 < nmax) {
        fetch(k0 + CPG, A);
        use(B, std::false_type{});
        advance();
        if (k0 >= nmax) break;
        fetch(k0 + CPG, B);
        use(A, std::false_type{});
        advance();
    }
    // the tail's loads of this lane's atom (the lanes sub < 4 of a group), issued before the
    // group sums so that their latency overlaps them
    const bool tail = sub < 4 && sub < nat;
    const int si_t  = (live ? ci : 0) * 4 + sub;
    double px = 0.0, py = 0.0, pz = 0.0, vx0 = 0.0, vy0 = 0.0, vz0 = 0.0;
    if (tail) {
        if (xn) {
            px = sx[3 * si_t];
            py = sx[3 * si_t + 1];
            pz = sx[3 * si_t + 2];
        }
        if (v) {
            vx0 = v[3 * si_t];
            vy0 = v[3 * si_t + 1];
            vz0 = v[3 * si_t + 2];
        }
    }
    // reduce over the 16 lanes of each group (a DPP row)
    if (!COMP) {
        // by DPP moves inside the row, no LDS round trips: partners i^1, i^2, then 7-i inside
        // each half row, then 15-i; every lane of the group ends with the group's sums
        auto row_step = [&](auto ctrl) {
            constexpr int C = decltype(ctrl)::value;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                fx[a] += dpp_row<C>(fx[a]);
                fy[a] += dpp_row<C>(fy[a]);
                fz[a] += dpp_row<C>(fz[a]);
            }
        };
        row_step(std::integral_constant<int, DPP_XOR1>{});
        row_step(std::integral_constant<int, DPP_XOR2>{});
        row_step(std::integral_constant<int, DPP_HALF_MIRROR>{});
        row_step(std::integral_constant<int, DPP_MIRROR>{});
    } else {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const float ox = __shfl_xor(fx[a], o, 64), oy = __shfl_xor(fy[a], o, 64),
                            oz = __shfl_xor(fz[a], o, 64);
                const float ocx = __shfl_xor(cx[a], o, 64), ocy = __shfl_xor(cy[a], o, 64),
                            ocz = __shfl_xor(cz[a], o, 64);
                two_sum_merge(fx[a], cx[a], ox, ocx);
                two_sum_merge(fy[a], cy[a], oy, ocy);
                two_sum_merge(fz[a], cz[a], oz, ocz);
            }
        }
    }
    if (tail) {
        float rx = 0.f, ry = 0.f, rz = 0.f;
#pragma unroll
        for (int a = 0; a < 4; a++)
            if (a == sub) {
                rx = COMP ? fx[a] + cx[a] : fx[a];
                ry = COMP ? fy[a] + cy[a] : fy[a];
                rz = COMP ? fz[a] + cz[a] : fz[a];
            }
        const int si  = si_t;
        f[3 * si]     = rx;
        f[3 * si + 1] = ry;
        f[3 * si + 2] = rz;
        if (v) {
            // finalIntegrate (integrate.c:61-86) of this atom, fused: its force is complete
            double vx = vx0 + dtforce * (double)rx;
            double vy = vy0 + dtforce * (double)ry;
            double vz = vz0 + dtforce * (double)rz;
            if (xn) {
                // and the next step's initialIntegrate (integrate.c:29-59), which kicks with
                // the same force: half kick, then the drift into the other position buffer,
                // with its record
                vx             = vx + dtforce * (double)rx;
                vy             = vy + dtforce * (double)ry;
                vz             = vz + dtforce * (double)rz;
                const double nxp = px + dt * vx, nyp = py + dt * vy, nzp = pz + dt * vz;
                xn[3 * si]     = nxp;
                xn[3 * si + 1] = nyp;
                xn[3 * si + 2] = nzp;
                rn[si]         = make_rec(nxp, nyp, nzp, st[si], rg, status);
            }
            v[3 * si]     = vx;
            v[3 * si + 1] = vy;
            v[3 * si + 2] = vz;
        }
    }
    if (EV) ev_store(e, w, ev);
    (void)ncj;
}

// ---------------------------------------------------------------- half lists without atomics
// A tile is TILE_CI consecutive i-clusters.  Its table lists, in increasing order, the distinct
// local j-clusters its lists name (at most TILE_UCAP j-atom slots' worth), and every list entry
// gets its slot in that table (-1 for a ghost j-cluster: in a half list a pair with a ghost acts
// on its local atom only, its mirror pair sits in the ghost owner's list).  The force pass
// evaluates every listed cluster pair once: i-forces in registers, j-forces summed in the
// tile's table in LDS and written once per tile as records; a gather pass then gives each atom
// its i-force plus the records of its j-cluster, in record order.  The table sums 64-bit fixed
// point (2^-40 units, range +-2^23): integer adds are exact, so the sums do not depend on the
// order in which the wavefronts of a tile add (and LDS integer atomics run at full rate, FP32
// ones did not: 0.57 against 0.21 ms per 1M-atom call).  No global atomics anywhere.
constexpr int TILE_CI    = 16;    // i-clusters per tile
constexpr int TILE_ECAP  = 2048;  // list entries per tile (more in any tile: the atomic kernel)
constexpr int TILE_W64   = 7680;  // table capacity in 64-bit words (60 KB): 640 slots of 4x3,
                                  // 320 of 8x3 (the force pass's LDS holds the largest table
                                  // of the current list, not the capacity)
constexpr int TILE_WAVES = 4;     // wavefronts per tile block, four i-clusters each per round
constexpr double FIX_ONE = 1099511627776.0;  // 2^40, one force unit in table units

// v * 2^40 rounded to the nearest integer (ties to even) as a 64-bit integer: the value of
// __double2ll_rn((double)v * FIX_ONE), in FP32 operations, each exact (the scalings by powers
// of two; the low part, which holds at most the 24 bits of y below 2^32); |v| < 2^23.
__device__ __forceinline__ long long fix40(float v)
{
    const float y  = __builtin_rintf(v * 1099511627776.0f);
    const float hi = floorf(y * 2.3283064365386963e-10f);
    const float lo = fmaf(-hi, 4294967296.0f, y);
    return (long long)(int)hi * 4294967296ll + (long long)(unsigned)lo;
}
static_assert(TILE_CI % (4 * TILE_WAVES) == 0, "whole rounds per tile");

constexpr int TILE_HASH = 2 * TILE_ECAP;  // open-addressing set of a tile's j-clusters (LDS)
constexpr int TILE_UNI  = TILE_ECAP;  // the distinct j-clusters, at most one per entry
static_assert(TILE_W64 / (3 * 2) <= TILE_HASH, "the sorted table fits the set's array");
constexpr int TILE_TT   = 256;   // threads of the tile-table kernel
__global__ void __launch_bounds__(TILE_TT) k_cp_tile_table(int nci, int ncj, int ucap,
    const int* __restrict__ nb, const int* __restrict__ numneigh, int maxneighs,
    int* __restrict__ tile_cj, int* __restrict__ tile_n, int* __restrict__ entry_slot,
    int* __restrict__ status, int* __restrict__ nmax, int* __restrict__ rcount)
{
    static_assert(TILE_CI <= 64, "the list-length prefix is one wavefront scan");
    __shared__ int s_hash[TILE_HASH];
    __shared__ int s_uni[TILE_UNI];
    __shared__ int s_pre[TILE_CI + 1];
    __shared__ int s_sh[33];
    const int t = blockIdx.x, c0 = t * TILE_CI, nc = min(nci - c0, TILE_CI);
    const int tid = threadIdx.x;
    // entries of the tile: prefix of the list lengths (clamped: the kernel may run on a list
    // whose build overflowed, before the host has seen that and rebuilt it)
    if (tid < 64) {
        int v = tid < nc ? min(numneigh[c0 + tid], maxneighs) : 0;
        int inc = v;
        for (int o = 1; o < 64; o <<= 1) {
            int u = __shfl_up(inc, o, 64);
            if (tid >= o) inc += u;
        }
        if (tid < TILE_CI) s_pre[tid + 1] = inc;
        if (tid == 0) s_pre[0] = 0;
    }
    __syncthreads();
    const int E = s_pre[TILE_CI];
    if (E > TILE_ECAP) {
        if (tid == 0) atomicOr(status, 1);
        return;
    }
    // the set's size: the power of two at least twice the entry count
    int HS = 64;
    while (HS < 2 * E)
        HS <<= 1;
    for (int k = tid; k < HS; k += TILE_TT)
        s_hash[k] = -1;
    __syncthreads();
    auto entry_at = [&](int e) {
        int c = 0;
        for (int step = TILE_CI / 2; step > 0; step >>= 1)
            if (s_pre[c + step] <= e) c += step;
        return (size_t)(c0 + c) * maxneighs + (e - s_pre[c]);
    };
    // the distinct local j-clusters into the set (linear probing; at most half full, as
    // 2 E <= HS <= TILE_HASH, so every probe sequence ends)
    for (int e = tid; e < E; e += TILE_TT) {
        const int cj = nb[entry_at(e)];
        if (cj >= ncj) continue;
        unsigned h = ((unsigned)cj * 2654435761u) & (HS - 1);
        for (;;) {
            const int old = atomicCAS(&s_hash[h], -1, cj);
            if (old == -1 || old == cj) break;
            h = (h + 1) & (HS - 1);
        }
    }
    __syncthreads();
    // compaction: each thread a contiguous run of the set, block scan of the counts
    const int RUN = (HS + TILE_TT - 1) / TILE_TT, r0 = min(tid * RUN, HS), r1 = min(r0 + RUN, HS);
    int cnt = 0;
    for (int q = r0; q < r1; q++)
        cnt += s_hash[q] >= 0;
    int tot;
    {
        const int lane = tid & 63, w = tid >> 6, nw = TILE_TT >> 6;
        int inc = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            int u = __shfl_up(inc, o, 64);
            if (lane >= o) inc += u;
        }
        if (lane == 63) s_sh[w] = inc;
        __syncthreads();
        if (w == 0) {
            int v  = lane < nw ? s_sh[lane] : 0;
            int vi = v;
            for (int o = 1; o < 64; o <<= 1) {
                int u = __shfl_up(vi, o, 64);
                if (lane >= o) vi += u;
            }
            if (lane < nw) s_sh[lane] = vi - v;
            if (lane == nw - 1) s_sh[32] = vi;
        }
        __syncthreads();
        int off = inc - cnt + s_sh[w];
        tot     = s_sh[32];
        for (int q = r0; q < r1; q++) {
            const int k = s_hash[q];
            if (k >= 0) s_uni[off++] = k;
        }
    }
    if (tot > ucap) {
        if (tid == 0) atomicOr(status, 2);
        return;
    }
    // the distinct keys in increasing order: each key goes to its rank, the number of keys below
    // it (every thread reads the same key at once, a broadcast; no barrier per step as a
    // sorting network needs), into the set's array, which the compaction no longer reads
    __syncthreads();
    int* s_srt = s_hash;
    for (int u = tid; u < tot; u += TILE_TT) {
        const int key = s_uni[u];
        int r         = 0;
        for (int q = 0; q < tot; q++)
            r += s_uni[q] < key;
        s_srt[r] = key;
    }
    __syncthreads();
    // the table, and each j-cluster's record count for the reverse map (cleared before the launch)
    for (int u = tid; u < tot; u += TILE_TT) {
        tile_cj[(size_t)t * ucap + u] = s_srt[u];
        atomicAdd(&rcount[s_srt[u]], 1);
    }
    if (tid == 0) {
        tile_n[t] = tot;
        atomicMax(nmax, tot);  // the largest table, which sizes the force pass's LDS
    }
    // every entry's slot: its j-cluster's place in the table (binary search)
    for (int e = tid; e < E; e += TILE_TT) {
        const size_t at = entry_at(e);
        const int cj    = nb[at];
        int slot        = -1;
        if (cj < ncj) {
            int lo = 0, hi = tot - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_srt[mid] < cj) lo = mid + 1;
                else hi = mid;
            }
            slot = lo;
        }
        entry_slot[at] = slot;
    }
}

// records of every local j-cluster (counted by the tile-table kernel): after a scan of the
// counts, fill and sort each run so the gather visits a cluster's records in increasing record
// order (w threads per tile: w is the largest table of the list)
__global__ void k_cp_rev_fill(int ntiles, int ucap, int w, const int* __restrict__ tile_cj,
    const int* __restrict__ tile_n, const int* __restrict__ rec_off,
    const int* __restrict__ rstart, int* __restrict__ rcur, int* __restrict__ rev)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(k / w), u = (int)(k % w);
    if (t >= ntiles || u >= tile_n[t]) return;
    const int cj = tile_cj[(size_t)t * ucap + u];
    rev[rstart[cj] + atomicAdd(&rcur[cj], 1)] = rec_off[t] + u;
}

// A run of at most R records sorted in registers: R independent loads (padded with INT_MAX),
// a bitonic network with compile-time indices, the n values stored back.
template <int R>
__device__ __forceinline__ void rev_sort_regs(int* __restrict__ run, int n)
{
    int v[R];
#pragma unroll
    for (int i = 0; i < R; i++)
        v[i] = i < n ? run[i] : INT_MAX;
#pragma unroll
    for (int k = 2; k <= R; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const int lo = min(v[i], v[l]), hi = max(v[i], v[l]);
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
#pragma unroll
    for (int i = 0; i < R; i++)
        if (i < n) run[i] = v[i];
}
__global__ void k_cp_rev_sort(int ncj, const int* __restrict__ rstart, int* __restrict__ rev)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncj) return;
    const int b = rstart[c], e = rstart[c + 1];
    // most runs hold ~10 records; longer ones (rare) by insertion in memory
    if (e - b <= 16) {
        rev_sort_regs<16>(rev + b, e - b);
        return;
    }
    if (e - b <= 32) {
        rev_sort_regs<32>(rev + b, e - b);
        return;
    }
    for (int i = b + 1; i < e; i++) {
        const int v = rev[i];
        int j       = i - 1;
        while (j >= b && rev[j] > v) {
            rev[j + 1] = rev[j];
            j--;
        }
        rev[j + 1] = v;
    }
}

// Half-list force pass over one tile per block (TILE_WAVES wavefronts, each taking four
// i-clusters per round).  The lane work is that of k_cp_force_f32_full (one j-atom per lane
// against the four i-atoms of its group, records, cutoff band, FP32 pair math), with the half
// list's pair rule in the first iteration (entry 0 is the i-cluster's own j-cluster: its pairs
// with s_j > s_i only) and the j-atom's force summed over its four pairs and added to the
// tile's table in LDS.  A pair with a ghost j-cluster acts on the i-atom only and, in the
// energy sums, weighs one half.
template <int N, bool EV, bool UNIFORM>
__global__ void __launch_bounds__(TILE_WAVES * 64) k_cp_force_half_tiled(int nci, int ncj,
    int dummy_cj, int ntypes, Tab32 tab, const double* __restrict__ sx,
    const float4* __restrict__ sr, const int* __restrict__ nb, const int* __restrict__ slots,
    const int* __restrict__ numneigh, const double* __restrict__ ci_bbox, RecGrid rg,
    int maxneighs, int ucap, const int* __restrict__ tile_n, const int* __restrict__ rec_off,
    float* __restrict__ rec, float* __restrict__ fown, double* __restrict__ ev)
{
    constexpr int CPG    = 16 / N;
    constexpr bool SPLIT = !UNIFORM;
    constexpr int RF     = 3 * N;  // 64-bit words per table slot
    extern __shared__ unsigned long long s_table[];
    __shared__ float s_tab[2][MAXT2];
    __shared__ double s_cut[MAXT2];
    __shared__ float4 s_pk[MAXT2];
    const int t = blockIdx.x;
    const int ntab = tile_n[t];
    for (int k = threadIdx.x; k < ntab * RF; k += blockDim.x)
        s_table[k] = 0ull;
    if (!UNIFORM) {
        for (int k = threadIdx.x; k < ntypes * ntypes; k += blockDim.x) {
            s_tab[0][k]   = tab.eps4[k];
            s_tab[1][k]   = tab.sig6[k];
            s_cut[k]      = tab.cut2d[k];
            const float c = (float)tab.cut2d[k];
            s_pk[k] = make_float4(tab.c12[k], -tab.c6[k], c * (1.0f - CUT_BAND), c * (1.0f + CUT_BAND));
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, grp = lane >> 4, sub = lane & 15;
    const int jc = sub / N, jj = sub % N;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const float c12u = tab.c12[0], eps4u = tab.eps4[0], sig6u = tab.sig6[0];
    float nc6u = -tab.c6[0];
    asm volatile("" : "+v"(nc6u));
    const double cut2u = tab.cut2d[0];
    const float clo_u  = (float)cut2u * (1.0f - CUT_BAND), chi_u = (float)cut2u * (1.0f + CUT_BAND);
    const float hf     = rg.hf;
    const unsigned jj16 = 16u * (unsigned)jj;
    const int bp_base   = ((grp << 4) + jc) << 2;
    float e = 0.f, w = 0.f;
    for (int round = 0; round < TILE_CI / (4 * TILE_WAVES); round++) {
        // the four i-clusters of this wavefront and round, a quarter tile apart: groups of one
        // wavefront then rarely meet the same j-atom in one iteration, whose LDS adds to one
        // address would serialise
        const int base  = __builtin_amdgcn_readfirstlane(t * TILE_CI + round * TILE_WAVES + wave);
        if (base >= nci) break;
        const int ci    = base + grp * (TILE_CI / 4);
        const bool live = ci < nci;
        const int ci_cj0 = (N == 4) ? ci : ci >> 1;
        float kI[3] = { 0.f, 0.f, 0.f };
        if (!SPLIT) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double lo = live ? ci_bbox[6 * ci + 2 * d] : 0.0;
                const double hi = live ? ci_bbox[6 * ci + 2 * d + 1] : 0.0;
                const double c  = lo <= hi ? 0.5 * (lo + hi) : 0.0;
                kI[d]           = (float)(rint(c * rg.inv_h) + rg.bias);
            }
        }
        float zi[4][3], ki[4][3];
        int ti[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const float4 r   = live ? sr[ci * 4 + a] : make_float4(1e15f, 1e15f, 1e15f, 0.f);
            const unsigned u = __float_as_uint(r.w);
            ki[a][0] = (float)(u & 0xffu);
            ki[a][1] = (float)((u >> 8) & 0xffu);
            ki[a][2] = (float)((u >> 16) & 0xffu);
            if (SPLIT) {
                zi[a][0] = r.x;
                zi[a][1] = r.y;
                zi[a][2] = r.z;
            } else {
                zi[a][0] = fmaf(ki[a][0] - kI[0], hf, r.x);
                zi[a][1] = fmaf(ki[a][1] - kI[1], hf, r.y);
                zi[a][2] = fmaf(ki[a][2] - kI[2], hf, r.z);
            }
            ti[a] = UNIFORM ? 0 : (int)(u >> 24);
        }
        const int nn = live ? numneigh[ci] : 0;
        int nmax     = max(nn, __shfl_xor(nn, 16, 64));
        nmax         = max(nmax, __shfl_xor(nmax, 32, 64));
        nmax         = __builtin_amdgcn_readfirstlane(nmax);
        const size_t lbase = (size_t)(live ? ci : 0) * maxneighs;
        float fx[4] = {}, fy[4] = {}, fz[4] = {};
        struct JL {
            float4 r;
            unsigned off;
            int slot;
        };
        auto use = [&](const JL& J, bool first) {
            const unsigned wj = __float_as_uint(J.r.w);
            const float kx = (float)(wj & 0xffu), ky = (float)((wj >> 8) & 0xffu),
                        kz = (float)((wj >> 16) & 0xffu);
            const int tj   = UNIFORM ? 0 : (int)(wj >> 24);
            const int sj   = (int)(J.off >> 4);
            const int cj   = sj / N;
            float yx = 0.f, yy = 0.f, yz = 0.f;
            if (!SPLIT) {
                yx = fmaf(kx - kI[0], hf, J.r.x);
                yy = fmaf(ky - kI[1], hf, J.r.y);
                yz = fmaf(kz - kI[2], hf, J.r.z);
            }
            float gx = 0.f, gy = 0.f, gz = 0.f;  // the j-atom's force, over its four pairs
#pragma unroll
            for (int a = 0; a < 4; a++) {
                float dx, dy, dz;
                if (SPLIT) {
                    dx = fmaf(ki[a][0] - kx, hf, zi[a][0] - J.r.x);
                    dy = fmaf(ki[a][1] - ky, hf, zi[a][1] - J.r.y);
                    dz = fmaf(ki[a][2] - kz, hf, zi[a][2] - J.r.z);
                } else {
                    dx = zi[a][0] - yx;
                    dy = zi[a][1] - yy;
                    dz = zi[a][2] - yz;
                }
                const float rsq = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                float c12 = c12u, nc6 = nc6u, eps4 = eps4u, sig6 = sig6u, clo = clo_u, chi = chi_u;
                const int tij = UNIFORM ? 0 : ti[a] * ntypes + tj;
                if (!UNIFORM) {
                    const float4 pk = s_pk[tij];
                    c12             = pk.x;
                    nc6             = pk.y;
                    clo             = pk.z;
                    chi             = pk.w;
                    if (EV) {
                        eps4 = s_tab[0][tij];
                        sig6 = s_tab[1][tij];
                    }
                }
                bool in          = rsq < clo;
                const bool below = rsq < chi;
                if (__builtin_expect((__ballot(below) & ~__ballot(in)) != 0, 0)) {
                    int lr;
                    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0"
                                 : "=v"(lr));
                    const int cir     = min(base + (lr >> 4) * (TILE_CI / 4), nci - 1);
                    const double* pi  = sx + 3 * (cir * 4 + a);
                    const double* pj  = sx + 3 * sj;
                    const double dxd  = pi[0] - pj[0];
                    const double dyd  = pi[1] - pj[1];
                    const double dzd  = pi[2] - pj[2];
                    const double rsqd = dxd * dxd + dyd * dyd + dzd * dzd;
                    const double cut2 = UNIFORM ? cut2u : s_cut[tij];
                    if (below && !in) in = rsqd < cut2;
                }
                // half-list rule in the i-cluster's own j-cluster (entry 0): s_j > s_i only; the
                // energy variant also drops pairs at distance 0 everywhere (padding slots whose
                // FP32 sentinels coincide), which would put a NaN into its sums
                if (first) in = in && !(cj == ci_cj0 && !(ci * 4 + a < sj));
                if (EV) in = in && rsq > 0.0f;
                const float r2i = __builtin_amdgcn_rcpf(rsq);
                const float r6i = r2i * r2i * r2i;
                float fpair     = r6i * fmaf(c12, r6i, nc6) * r2i;
                fpair           = in ? fpair : 0.0f;
                fx[a] = fmaf(dx, fpair, fx[a]);
                fy[a] = fmaf(dy, fpair, fy[a]);
                fz[a] = fmaf(dz, fpair, fz[a]);
                gx    = fmaf(-dx, fpair, gx);
                gy    = fmaf(-dy, fpair, gy);
                gz    = fmaf(-dz, fpair, gz);
                if (EV) {
                    const float sr6 = r6i * sig6;
                    const float wgt = cj < ncj ? 1.0f : 0.5f;
                    e += in ? wgt * eps4 * sr6 * (sr6 - 1.0f) : 0.0f;
                    w += wgt * rsq * fpair;
                }
            }
            if (J.slot >= 0) {
                unsigned long long* q = s_table + (J.slot * N + jj) * 3;
                atomicAdd(q, (unsigned long long)fix40(gx));
                atomicAdd(q + 1, (unsigned long long)fix40(gy));
                atomicAdd(q + 2, (unsigned long long)fix40(gz));
            }
        };
        auto entry = [&](int k) {
            return k < nn ? make_int2((int)(16u * N * (unsigned)nb[lbase + k]), slots[lbase + k])
                          : make_int2((int)(16u * N * (unsigned)dummy_cj), -1);
        };
        int2 ec = entry(sub), en = entry(16 + sub);
        auto fetch = [&](int kb, JL& J) {
            const int kk    = kb & 15;
            const int2 src  = kk ? ec : en;
            const int addr  = bp_base + (kk << 2);
            J.off  = (unsigned)__builtin_amdgcn_ds_bpermute(addr, src.x) + jj16;
            J.slot = __builtin_amdgcn_ds_bpermute(addr, src.y);
            J.r    = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sr) + J.off);
        };
        JL A, B;
        {
            A.off  = (unsigned)__builtin_amdgcn_ds_bpermute(bp_base, ec.x) + jj16;
            A.slot = __builtin_amdgcn_ds_bpermute(bp_base, ec.y);
            A.r    = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sr) + A.off);
        }
        int k0 = 0;
        auto advance = [&]() {
            k0 += CPG;
            if (!(k0 & 15)) {
                ec = en;
                en = entry(k0 + 16 + sub);
            }
        };
        fetch(CPG, B);
        use(A, true);
        advance();
        while (k0 < nmax) {
            fetch(k0 + CPG, A);
            use(B, false);
            advance();
            if (k0 >= nmax) break;
            fetch(k0 + CPG, B);
            use(A, false);
            advance();
        }
        // the i-forces of the round: reduce over the 16 lanes of each group (a DPP row), as in
        // k_cp_force_f32_full
        auto row_step = [&](auto ctrl) {
            constexpr int C = decltype(ctrl)::value;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                fx[a] += dpp_row<C>(fx[a]);
                fy[a] += dpp_row<C>(fy[a]);
                fz[a] += dpp_row<C>(fz[a]);
            }
        };
        row_step(std::integral_constant<int, DPP_XOR1>{});
        row_step(std::integral_constant<int, DPP_XOR2>{});
        row_step(std::integral_constant<int, DPP_HALF_MIRROR>{});
        row_step(std::integral_constant<int, DPP_MIRROR>{});
        if (live && sub < 4) {
            float rx = 0.f, ry = 0.f, rz = 0.f;
#pragma unroll
            for (int a = 0; a < 4; a++)
                if (a == sub) {
                    rx = fx[a];
                    ry = fy[a];
                    rz = fz[a];
                }
            const int si      = ci * 4 + sub;
            fown[3 * si]      = rx;
            fown[3 * si + 1]  = ry;
            fown[3 * si + 2]  = rz;
        }
    }
    __syncthreads();
    // the tile's records, written once (FP32)
    float* out = rec + (size_t)rec_off[t] * RF;
    for (int k = threadIdx.x; k < ntab * RF; k += blockDim.x)
        out[k] = (float)((double)(long long)s_table[k] * (1.0 / FIX_ONE));
    if (EV) ev_store(e, w, ev);
    (void)ucap;
}

// Each local atom's total force: its i-force plus the records of its j-cluster in record
// order; then, when v is given, the final kick and, when xn is given, the next step's half
// kick and drift (the same operations as the tail of k_cp_force_f32_full).
template <int N>
__global__ void k_cp_gather_half(int nslot_loc, int M, const int* __restrict__ ci_natoms,
    const float* __restrict__ fown, const float* __restrict__ rec,
    const int* __restrict__ rstart, const int* __restrict__ rev, float* __restrict__ f,
    const double* __restrict__ sx, const int* __restrict__ st, double* __restrict__ v,
    double dtforce, double* __restrict__ xn, float4* __restrict__ rn, double dt, RecGrid rg,
    int* __restrict__ status)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot_loc || s % M >= ci_natoms[s / M]) return;
    const int c = s / N, jj = s - c * N;
    float rx = fown[3 * s], ry = fown[3 * s + 1], rz = fown[3 * s + 2];
    const int e = rstart[c + 1];
    int p       = rstart[c];
    // four records at a time: their indices, then their values load together; the sums keep
    // record order
    for (; p + 4 <= e; p += 4) {
        const int4 r4 = make_int4(rev[p], rev[p + 1], rev[p + 2], rev[p + 3]);
        const float* q0 = rec + ((size_t)r4.x * N + jj) * 3;
        const float* q1 = rec + ((size_t)r4.y * N + jj) * 3;
        const float* q2 = rec + ((size_t)r4.z * N + jj) * 3;
        const float* q3 = rec + ((size_t)r4.w * N + jj) * 3;
        const float a0 = q0[0], b0 = q0[1], c0 = q0[2], a1 = q1[0], b1 = q1[1], c1 = q1[2];
        const float a2 = q2[0], b2 = q2[1], c2 = q2[2], a3 = q3[0], b3 = q3[1], c3 = q3[2];
        rx = (((rx + a0) + a1) + a2) + a3;
        ry = (((ry + b0) + b1) + b2) + b3;
        rz = (((rz + c0) + c1) + c2) + c3;
    }
    for (; p < e; p++) {
        const float* q = rec + ((size_t)rev[p] * N + jj) * 3;
        rx += q[0];
        ry += q[1];
        rz += q[2];
    }
    f[3 * s]     = rx;
    f[3 * s + 1] = ry;
    f[3 * s + 2] = rz;
    if (v) {
        double vx = v[3 * s] + dtforce * (double)rx;
        double vy = v[3 * s + 1] + dtforce * (double)ry;
        double vz = v[3 * s + 2] + dtforce * (double)rz;
        if (xn) {
            vx = vx + dtforce * (double)rx;
            vy = vy + dtforce * (double)ry;
            vz = vz + dtforce * (double)rz;
            const double nxp = sx[3 * s] + dt * vx, nyp = sx[3 * s + 1] + dt * vy,
                         nzp = sx[3 * s + 2] + dt * vz;
            xn[3 * s]     = nxp;
            xn[3 * s + 1] = nyp;
            xn[3 * s + 2] = nzp;
            rn[s]         = make_rec(nxp, nyp, nzp, st[s], rg, status);
        }
        v[3 * s]     = vx;
        v[3 * s + 1] = vy;
        v[3 * s + 2] = vz;
    }
}

// ---------------------------------------------------------------- integrate (integrate.c:29-86)
template <class FT>
__global__ void k_cp_initial(int n3, int M, const int* __restrict__ ci_natoms, double* __restrict__ x,
    double* __restrict__ v, const FT* __restrict__ f, double dt, double dtforce)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    int s = k / 3;
    if (s % M >= ci_natoms[s / M]) return;
    double vv = v[k] + dtforce * (double)f[k];
    v[k]      = vv;
    x[k]      = x[k] + dt * vv;
}

template <class FT>
__global__ void k_cp_final(int n3, int M, const int* __restrict__ ci_natoms, double* __restrict__ v,
    const FT* __restrict__ f, double dtforce)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    int s = k / 3;
    if (s % M >= ci_natoms[s / M]) return;
    v[k] = v[k] + dtforce * (double)f[k];
}

__global__ void __launch_bounds__(BLOCK) k_cp_ksum(int nslot, int M, const int* __restrict__ ci_natoms,
    const double* __restrict__ v, double mass, double* __restrict__ part)
{
    __shared__ double sh[BLOCK / 64];
    double t = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslot; s += gridDim.x * blockDim.x)
        if (s % M < ci_natoms[s / M])
            t += (v[3 * s] * v[3 * s] + v[3 * s + 1] * v[3 * s + 1] + v[3 * s + 2] * v[3 * s + 2]) *
                 mass;
    for (int o = 32; o > 0; o >>= 1)
        t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int k = 0; k < BLOCK / 64; k++)
            r += sh[k];
        part[blockIdx.x] = r;
    }
}

// out[0] = the sum of part[0..n): per thread a strided sum, then a fixed tree (the order does
// not depend on timing).  One block of BLOCK threads; out may be part (every read precedes the
// first barrier).  A single thread took 31 us for 512 partials, one dependent load at a time.
__global__ void __launch_bounds__(BLOCK) k_cp_sum_partials(int n, const double* part, double* out)
{
    __shared__ double s[BLOCK];
    const int t = threadIdx.x;
    double r = 0.0;
    for (int k = t; k < n; k += BLOCK)
        r += part[k];
    s[t] = r;
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) out[0] = s[0];
}

// ---------------------------------------------------------------- host helpers
void settle_counts();
// Entry points that need the host-side cluster counts call ok(), which reads a pending count
// back first; the rebuild steps between the binning and the ghost setup call queue_ok().
bool queue_ok() { return !failed() && C.ready; }
bool ok()
{
    if (!queue_ok()) return false;
    if (C.nci_pending) settle_counts();
    return !failed();
}

// The address of `scratch` on the current device, resolved on the first call per device.
double* scratch_ptr()
{
    int dev = 0;
    MDG_HIP(hipGetDevice(&dev));
    if (failed()) return nullptr;
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error(E_ARG, "device %d: the scratch table holds devices 0..%d", dev, MAX_DEVICES - 1);
        return nullptr;
    }
    if (!g_scratch[dev]) {
        void* p      = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

void read_status()
{
    MDG_HIP(hipMemcpyAsync(C.h_status, C.d_status, 16 * sizeof(int), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
}

int read_int(const int* d)
{
    MDG_HIP(hipMemcpyAsync(C.h_status + 8, d, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return C.h_status[8];
}

size_t nslots_all() { return ((size_t)C.ncj + C.nghost + 1) * C.N; }

bool ensure_slots(size_t nslot, size_t keep_slots)
{
    C.sxn_valid = false;
    C.rec_valid = false;
    grow_n(&C.d_sx, &C.cap_sx, 3 * nslot, 3 * keep_slots);
    grow_n(&C.d_st, &C.cap_st, nslot, keep_slots);
    return !failed();
}

}  // namespace

// ================================================================ C ABI
extern "C" {

int setClusterShapeGFX950(int M, int N)
{
    if (!((M == 4 && N == 4) || (M == 4 && N == 8) || (M == 1 && N == 2))) {
        set_error(E_ARG, "cluster shape %dx%d not supported (4x4, 4x8, 1x2)", M, N);
        return -1;
    }
    C.M        = M;
    C.N        = N;
    C.rule     = (M == 1) ? 2 : N;  // get_imask_simd_j2/j4/j8 of the matching reference build
    C.fma_dist = (M == 4) ? 1 : 0;  // the 4xN SIMD list test uses FMA; the plain-C one does not
    return 0;
}

void setClusterPrecisionGFX950(int precision, int compensated)
{
    C.precision = precision;
    C.comp      = compensated;
}

void initDeviceClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor)
{
    if (failed() || !stream()) return;
    if (C.precision == 0 && C.M != 4) {
        set_error(E_ARG, "the FP32 kernel needs i-clusters of 4 atoms");
        return;
    }
    if (!C.d_status) {
        MDG_HIP(hipMalloc(&C.d_status, 16 * sizeof(int)));
        MDG_HIP(hipHostMalloc(&C.h_status, 16 * sizeof(int), hipHostMallocDefault));
        MDG_HIP(hipMalloc(&C.d_ev, 4 * sizeof(double)));
        MDG_HIP(hipHostMalloc(&C.h_red, 4 * sizeof(double), hipHostMallocDefault));
        if (failed()) return;
    }
    fill_i32(C.d_status, 0, 16);
    C.natoms   = atom->Nlocal;
    C.ntypes   = atom->ntypes;
    C.cutforce = param->cutforce;
    C.cutneigh = param->cutneigh;
    C.dt       = param->dt;
    C.dtforce  = param->dtforce;
    C.mass     = param->mass;
    if (C.ntypes * C.ntypes > MAXT2) {
        set_error(E_ARG, "ntypes %d > 8 not supported", C.ntypes);
        return;
    }
    C.uniform = true;
    for (int k = 0; k < C.ntypes * C.ntypes; k++) {
        C.t64.eps[k]    = atom->epsilon[k];
        C.t64.sig6[k]   = atom->sigma6[k];
        C.t64.cutf[k]   = atom->cutforcesq[k];
        C.t32.c12[k]    = (float)(48.0 * atom->epsilon[k] * atom->sigma6[k] * atom->sigma6[k]);
        C.t32.c6[k]     = (float)(24.0 * atom->epsilon[k] * atom->sigma6[k]);
        C.t32.eps4[k]   = (float)(4.0 * atom->epsilon[k]);
        C.t32.sig6[k]   = (float)atom->sigma6[k];
        C.t32.cut2[k]   = (float)atom->cutforcesq[k];
        C.t32.cut2d[k]  = atom->cutforcesq[k];
        if (atom->epsilon[k] != atom->epsilon[0] || atom->sigma6[k] != atom->sigma6[0] ||
            atom->cutforcesq[k] != atom->cutforcesq[0])
            C.uniform = false;
    }
    C.maxneighs = neighbor->maxneighs > 0 ? neighbor->maxneighs : 200;
    {
        // record grid: the finest power-of-two spacing whose 256 cells per axis cover the box
        // plus a margin on both sides (the ghost layer and the drift between rebuilds)
        const double margin = 2.0 * param->cutneigh + 4.0;
        const double ext =
            std::max(param->xprd, std::max(param->yprd, param->zprd)) + 2.0 * margin;
        double h = 1.0 / 1024.0;
        while (ext / h > 254.0)
            h *= 2.0;
        C.rg = RecGrid{ h, 1.0 / h, (float)h, (int)std::ceil(margin / h) };
    }
    if (!scratch_ptr()) return;  // set-up of kineticSumClusterNoAllocGFX950, once per device
    C.ready = true;
    copyAtomsToClusterGFX950Device(atom);
}

void freeDeviceClusterGFX950(void)
{
    if (C.d_status) (void)hipStreamSynchronize(stream());
    void* ptrs[] = { C.d_stencil, C.d_ax, C.d_av, C.d_at, C.d_abin_of, C.d_abin_cnt,
        C.d_abin_start, C.d_abin_atoms, C.d_ncl, C.d_cstart, C.d_ci_natoms, C.d_ci_bin,
        C.d_cj_natoms, C.d_ci_bbox, C.d_cj_bbox, C.d_sx, C.d_sv, C.d_sf64, C.d_sf32, C.d_st,
        C.d_gcnt, C.d_goff, C.d_border, C.d_pbc, C.d_gbin, C.d_jcnt, C.d_jstart,
        C.d_jlist, C.d_jkey, C.d_zidx, C.d_gcur, C.d_gbin_start, C.d_gbin_list, C.d_nb,
        C.d_numneigh, C.d_nmasked, C.d_red, C.d_ev, C.d_evp, C.d_status, C.d_sxn, C.d_sr, C.d_srn,
        C.d_tile_cj, C.d_tile_n, C.d_eslot, C.d_rec_off, C.d_rcount, C.d_rstart, C.d_rcur,
        C.d_rev, C.d_rec, C.d_fown };
    for (void* p : ptrs)
        dfree(p);
    if (C.h_status) (void)hipHostFree(C.h_status);
    if (C.h_red) (void)hipHostFree(C.h_red);
    int M = C.M, N = C.N, rule = C.rule, fm = C.fma_dist, prec = C.precision, comp = C.comp;
    C = CPState();
    C.M = M;
    C.N = N;
    C.rule = rule;
    C.fma_dist = fm;
    C.precision = prec;
    C.comp = comp;
}

namespace {
void flush_atoms();
}

void copyAtomsToClusterGFX950Device(Atom* atom)
{
    if (!ok()) return;
    flush_atoms();  // requested before this upload, so run before it
    size_t n = (size_t)atom->Nlocal;
    C.natoms = atom->Nlocal;
    grow_n(&C.d_ax, &C.cap_ax, 3 * n + 3, 0);
    grow_n(&C.d_av, &C.cap_av, 3 * n + 3, 0);
    grow_n(&C.d_at, &C.cap_at, n + 1, 0);
    if (failed()) return;
    double* v = (double*)malloc(sizeof(double) * 3 * n + 8);
    for (size_t i = 0; i < n; i++) {
        v[3 * i]     = atom->vx[i];
        v[3 * i + 1] = atom->vy[i];
        v[3 * i + 2] = atom->vz[i];
    }
    MDG_HIP(hipMemcpyAsync(C.d_ax, atom->x, 3 * n * sizeof(double), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(C.d_av, v, 3 * n * sizeof(double), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(C.d_at, atom->type, n * sizeof(int), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    free(v);
}

void copyAtomsFromClusterGFX950Device(Atom* atom)
{
    if (!ok()) return;
    flush_atoms();
    size_t n = (size_t)C.natoms;
    double* v = (double*)malloc(sizeof(double) * 3 * n + 8);
    MDG_HIP(hipMemcpyAsync(atom->x, C.d_ax, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipMemcpyAsync(v, C.d_av, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipMemcpyAsync(atom->type, C.d_at, n * sizeof(int), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    for (size_t i = 0; i < n; i++) {
        atom->vx[i] = v[3 * i];
        atom->vy[i] = v[3 * i + 1];
        atom->vz[i] = v[3 * i + 2];
    }
    free(v);
}

// setupNeighbor (src/clusterpair/neighbor.c:108-215), restated on the host.
void setupNeighborClusterGFX950(Parameter* param, Atom* atom)
{
    if (!ok()) return;
    const double SMALL = 1.0e-6, FACTOR = 0.999;
    double xprd = param->xprd, yprd = param->yprd, zprd = param->zprd;
    double xlo = 0.0, xhi = xprd, ylo = 0.0, yhi = yprd, zlo = 0.0, zhi = zprd;
    double atom_density  = ((double)(atom->Natoms)) / ((xhi - xlo) * (yhi - ylo) * (zhi - zlo));
    double atoms_in_cell = C.M > C.N ? C.M : C.N;
    double tsx = cbrt(atoms_in_cell / atom_density), tsy = cbrt(atoms_in_cell / atom_density);
    CGrid& g   = C.g;
    g.xprd = xprd;
    g.yprd = yprd;
    g.zprd = zprd;
    int t  = (int)ceil((xhi - xlo) / tsx);
    g.nbinx = t > 1 ? t : 1;
    t       = (int)ceil((yhi - ylo) / tsy);
    g.nbiny = t > 1 ? t : 1;
    C.binsizex   = (xhi - xlo) / g.nbinx;
    C.binsizey   = (yhi - ylo) / g.nbiny;
    g.bininvx    = 1.0 / C.binsizex;
    g.bininvy    = 1.0 / C.binsizey;
    C.cutneigh   = param->cutneigh;
    C.cutneighsq = C.cutneigh * C.cutneigh;
    double coord = xlo - C.cutneigh - SMALL * xprd;
    g.mbinxlo    = (int)(coord * g.bininvx);
    if (coord < 0.0) g.mbinxlo = g.mbinxlo - 1;
    coord       = xhi + C.cutneigh + SMALL * xprd;
    int mbinxhi = (int)(coord * g.bininvx);
    coord       = ylo - C.cutneigh - SMALL * yprd;
    g.mbinylo   = (int)(coord * g.bininvy);
    if (coord < 0.0) g.mbinylo = g.mbinylo - 1;
    coord       = yhi + C.cutneigh + SMALL * yprd;
    int mbinyhi = (int)(coord * g.bininvy);
    g.mbinxlo   = g.mbinxlo - 1;
    mbinxhi     = mbinxhi + 1;
    g.mbinx     = mbinxhi - g.mbinxlo + 1;
    g.mbinylo   = g.mbinylo - 1;
    mbinyhi     = mbinyhi + 1;
    g.mbiny     = mbinyhi - g.mbinylo + 1;
    g.mbins     = g.mbinx * g.mbiny;
    int nextx   = (int)(C.cutneigh * g.bininvx);
    int nexty   = (int)(C.cutneigh * g.bininvy);
    if (nextx * C.binsizex < FACTOR * C.cutneigh) nextx++;
    if (nexty * C.binsizey < FACTOR * C.cutneigh) nexty++;
    int st[1024], ns = 0;
    for (int j = -nexty; j <= nexty; j++)
        for (int i = -nextx; i <= nextx; i++) {
            double dx = i > 0 ? (i - 1) * C.binsizex : (i == 0 ? 0.0 : (i + 1) * C.binsizex);
            double dy = j > 0 ? (j - 1) * C.binsizey : (j == 0 ? 0.0 : (j + 1) * C.binsizey);
            if (dx * dx + dy * dy < C.cutneighsq) {
                if (ns >= 1024) {
                    set_error(E_ARG, "cluster stencil larger than 1024 columns");
                    return;
                }
                st[ns++] = j * g.mbinx + i;
            }
        }
    C.nstencil = ns;
    double bbx = 0.5 * (C.binsizex + C.binsizex), bby = 0.5 * (C.binsizey + C.binsizey);
    double r   = C.cutneigh - 0.5 * sqrt(bbx * bbx + bby * bby);
    r          = r > 0.0 ? r : 0.0;
    C.rbb_sq   = r * r;
    grow_n(&C.d_stencil, &C.cap_stencil, (size_t)ns, 0);
    if (failed()) return;
    MDG_HIP(hipMemcpyAsync(C.d_stencil, st, ns * sizeof(int), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
}

// updateSingleAtoms and updateAtomsPbc are deferred: the next binning does both in its first
// kernel (one pass over the slots: copy, wrap, bin count); anything else that reads or writes
// the atom arrays first runs what is pending (flush_atoms).
namespace {
void launch_slots_to_atoms(bool wrap, bool count)
{
    const int ns = C.nci * C.M;
    k_cp_slots_to_atoms<<<blocks_for(ns, BLOCK), BLOCK, 0, stream()>>>(C.nci, C.M,
        C.d_ci_natoms, C.d_ci_bin, C.d_cstart, C.d_abin_start, C.d_sx, C.d_sv, C.d_st, C.d_ax,
        C.d_av, C.d_at, wrap ? 1 : 0, C.wrap_prd[0], C.wrap_prd[1], C.wrap_prd[2], C.g,
        count ? C.d_abin_of : nullptr, count ? C.d_abin_cnt : nullptr, C.d_status);
    MDG_LAUNCH_CHECK("slots_to_atoms");
}

void flush_atoms()
{
    const bool copy = C.atoms_copy_pending, wrap = C.atoms_wrap_pending;
    C.atoms_copy_pending = C.atoms_wrap_pending = false;
    if (failed() || !(copy || wrap)) return;
    if (copy) {
        launch_slots_to_atoms(wrap, false);
    } else {
        const int n3 = 3 * C.natoms;
        k_cp_wrap<<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, C.d_ax, C.wrap_prd[0],
            C.wrap_prd[1], C.wrap_prd[2]);
        MDG_LAUNCH_CHECK("cp_wrap");
    }
}
}  // namespace

void updateSingleAtomsGFX950(Parameter* param, Atom* atom)
{
    (void)param;
    (void)atom;
    if (!ok() || C.nci == 0) return;
    flush_atoms();  // a wrap requested before this copy runs before it
    C.atoms_copy_pending = true;
}

void updateAtomsPbcClusterGFX950(Atom* atom, Parameter* param, bool dummy)
{
    (void)dummy;
    (void)atom;
    if (!ok()) return;
    if (C.atoms_wrap_pending) flush_atoms();
    C.atoms_wrap_pending = true;
    C.wrap_prd[0]        = param->xprd;
    C.wrap_prd[1]        = param->yprd;
    C.wrap_prd[2]        = param->zprd;
}

namespace {
// The status words written since the binning: binning flags (word 0), the cluster count (word
// 8), the record range flag.  Called after the status read that follows the binning.
void apply_cluster_counts()
{
    if (!C.nci_pending) return;
    C.nci_pending = false;
    if (C.h_status[0] & (ST_BIN_RANGE | ST_BIN_CAP)) {
        set_error(E_BIN, "cluster binning failed (flags %d): an atom outside the grid or a "
                         "column above %d atoms", C.h_status[0], BIN_CAP);
        return;
    }
    if (C.h_status[ST_REC_WORD] & ST_REC_RANGE) {
        set_error(E_BIN, "FP32 j-records: a position left the record grid (box +- %g)",
            C.rg.bias * C.rg.h);
        return;
    }
    C.nci = C.h_status[8];
    C.ncj = C.nci * C.M / C.N;
}

void settle_counts()
{
    read_status();
    if (!failed()) apply_cluster_counts();
}
}  // namespace

// The cluster count is left on the device (status word 8): the j-cluster and ghost-count
// kernels queue behind the binning without a host wait, sized for an upper bound of the count
// (each bin holds ceil(atoms / M) clusters, one more when N = 2M rounds the count up to even),
// and the ghost setup reads the count with the ghost totals.
void buildClustersGFX950(Atom* atom)
{
    if (!ok()) return;
    C.sxn_valid = false;
    C.rec_valid = false;
    if (C.g.mbins == 0) {
        set_error(E_ARG, "buildClustersGFX950 before setupNeighborClusterGFX950");
        return;
    }
    TimedRegion tr(F_CLUSTER);
    hipStream_t s = stream();
    const int na = C.natoms, mb = C.g.mbins;
    grow_n(&C.d_abin_of, &C.cap_abin_of, (size_t)na + 1, 0);
    grow_n(&C.d_abin_cnt, &C.cap_abin_cnt, (size_t)mb + 1, 0);
    if (failed()) return;
    fill_i32_multi({ { C.d_status, 0, 8 }, { C.d_abin_cnt, 0, (size_t)mb + 1 } });
    if (C.atoms_copy_pending) {
        // the deferred updateSingleAtoms (and updateAtomsPbc) with the bin count, in one pass
        // over the slots; launched before anything below can move the slots or the last
        // binning's arrays it reads
        const bool wrap      = C.atoms_wrap_pending;
        C.atoms_copy_pending = C.atoms_wrap_pending = false;
        launch_slots_to_atoms(wrap, true);
    } else {
        flush_atoms();
        k_cp_bin_count<<<blocks_for(na, BLOCK), BLOCK, 0, s>>>(na, C.d_ax, C.g, C.d_abin_of,
            C.d_abin_cnt, C.d_status);
    }
    grow_n(&C.d_abin_start, &C.cap_abin_start, (size_t)mb + 1, 0);
    grow_n(&C.d_abin_atoms, &C.cap_abin_atoms, (size_t)na + 1, 0);
    grow_n(&C.d_ncl, &C.cap_ncl, (size_t)mb + 1, 0);
    grow_n(&C.d_cstart, &C.cap_cstart, (size_t)mb + 1, 0);
    if (failed()) return;
    const long long cap = ((long long)na + C.M - 1) / C.M + (long long)mb * (C.N > C.M ? 2 : 1);
    if (cap * C.M + C.N >= (1ll << 31)) {
        set_error(E_ARG, "cluster binning: %lld cluster slots exceed 32-bit indexing", cap * C.M);
        return;
    }
    const int nci_cap = (int)cap;
    grow_n(&C.d_ci_natoms, &C.cap_ci_natoms, (size_t)nci_cap + 1, 0);
    grow_n(&C.d_ci_bin, &C.cap_ci_bin, (size_t)nci_cap + 1, 0);
    grow_n(&C.d_ci_bbox, &C.cap_ci_bbox, 6 * (size_t)nci_cap + 6, 0);
    grow_n(&C.d_sv, &C.cap_sv, 3 * (size_t)nci_cap * C.M + 3, 0);
    ensure_slots((size_t)nci_cap * C.M + C.N, 0);
    // M == N: the j-cluster descriptors are written with the i-clusters' (defineJClusters
    // then has nothing to do); the ghosts' follow after them
    C.cj_from_fill = C.M == C.N;
    if (C.cj_from_fill) {
        grow_n(&C.d_cj_natoms, &C.cap_cj_natoms, (size_t)nci_cap + 1, 0);
        grow_n(&C.d_cj_bbox, &C.cap_cj_bbox, 6 * (size_t)nci_cap + 6, 0);
    }
    if (failed()) return;
    // the bin starts; the counts cleared for the scatter's cursors
    exclusive_scan_take(C.d_abin_cnt, C.d_abin_start, mb, C.d_abin_start + mb);
    k_cp_bin_scatter<<<blocks_for(na, BLOCK), BLOCK, 0, s>>>(na, C.d_abin_of, C.d_abin_start,
        C.d_abin_cnt, C.d_abin_atoms);
    k_cp_bin_sort<<<mb, SORT_T, 0, s>>>(C.d_abin_start, C.d_abin_atoms, C.d_ax, C.M, C.N, C.d_ncl,
        (double)SORT_NBK / C.g.zprd,
        C.d_status);
    // the cluster starts and the cluster count (to status word 8)
    exclusive_scan(C.d_ncl, C.d_cstart, mb, C.d_cstart + mb, C.d_status + 8);
    MDG_LAUNCH_CHECK("cp_binning");
    C.nci_pending = true;
    C.nci         = nci_cap;
    C.ncj         = nci_cap * C.M / C.N;
    k_cp_fill_clusters<<<mb, BLOCK, 0, s>>>(mb, C.M, C.d_ncl, C.d_cstart,
        C.d_abin_start, C.d_abin_atoms, C.d_ax, C.d_av, C.d_at, C.d_sx, C.d_sv, C.d_st,
        C.d_ci_natoms, C.d_ci_bbox, C.d_ci_bin, C.cj_from_fill ? C.d_cj_natoms : nullptr,
        C.d_cj_bbox);
    MDG_LAUNCH_CHECK("cp_fill_clusters");
    (void)atom;
}

void defineJClustersGFX950(Parameter* param, Atom* atom)
{
    (void)param;
    (void)atom;
    if (!queue_ok()) return;
    if (C.cj_from_fill) {  // written by the cluster fill (M == N)
        C.cj_from_fill = false;
        return;
    }
    TimedRegion tr(F_CLUSTER);
    // ghosts follow: reserve their descriptors later, now the local ones (C.ncj is an upper
    // bound while the cluster count is pending)
    grow_n(&C.d_cj_natoms, &C.cap_cj_natoms, (size_t)C.ncj + 1, 0);
    grow_n(&C.d_cj_bbox, &C.cap_cj_bbox, 6 * (size_t)C.ncj + 6, 0);
    if (failed()) return;
    if (!C.nci_pending) fill_i32(C.d_status + 8, C.nci, 1);  // the kernel reads the count there
    k_cp_define_j<<<blocks_for(C.ncj, BLOCK), BLOCK, 0, stream()>>>(C.d_status + 8, C.M, C.N,
        C.d_ci_natoms, C.d_ci_bbox, C.d_cj_natoms, C.d_cj_bbox);
    MDG_LAUNCH_CHECK("cp_define_j");
}

void setupPbcClusterGFX950(Atom* atom, Parameter* param)
{
    C.sxn_valid = false;
    C.rec_valid = false;
    if (!queue_ok()) return;
    TimedRegion tr(F_PBC);
    hipStream_t s = stream();
    const int ncj_cap = C.ncj;  // the count itself, or an upper bound while it is pending
    grow_n(&C.d_gcnt, &C.cap_gcnt, (size_t)ncj_cap + 1, 0);
    grow_n(&C.d_goff, &C.cap_goff, (size_t)ncj_cap + 1, 0);
    if (failed()) return;
    // ghost clusters (status word 2) and ghost atoms (word 3), read in one copy with the
    // cluster count (word 8) and the binning flags
    if (!C.nci_pending) fill_i32(C.d_status + 8, C.nci, 1);
    fill_i32(C.d_status + 3, 0, 1);
    k_cp_ghost_count<<<blocks_for(ncj_cap, GC_T), GC_T, 0, s>>>(ncj_cap, C.d_status + 8, C.M, C.N,
        C.d_cj_natoms, C.d_cj_bbox, param->xprd, param->yprd, param->zprd, param->cutneigh,
        C.d_gcnt, C.d_status + 3);
    exclusive_scan(C.d_gcnt, C.d_goff, ncj_cap, C.d_status + 2);
    read_status();
    if (failed()) return;
    apply_cluster_counts();
    if (failed()) return;
    const int ncj = C.ncj;
    const int ng = C.h_status[2];
    C.nghost     = ng;
    C.dummy_cj   = ncj + ng;
    size_t nj    = (size_t)ncj + ng + 1;
    grow_n(&C.d_cj_natoms, &C.cap_cj_natoms, nj, (size_t)ncj);
    grow_n(&C.d_cj_bbox, &C.cap_cj_bbox, 6 * nj, 6 * (size_t)ncj);
    grow_n(&C.d_border, &C.cap_border, (size_t)ng + 1, 0);
    grow_n(&C.d_pbc, &C.cap_pbc, 3 * (size_t)ng + 3, 0);
    ensure_slots(nj * C.N, (size_t)ncj * C.N);
    if (failed()) return;
    k_cp_ghost_fill<<<blocks_for(ncj, BLOCK), BLOCK, 0, s>>>(ncj, C.N, C.d_cj_natoms, C.d_cj_bbox,
        param->xprd, param->yprd, param->zprd, param->cutneigh, C.d_goff, C.d_border, C.d_pbc,
        C.d_cj_natoms, C.d_st);
    k_cp_ghost_first<<<blocks_for(ng + 1, BLOCK), BLOCK, 0, s>>>(ncj, ng, C.N, C.d_border, C.d_pbc,
        C.d_cj_natoms, param->xprd, param->yprd, param->zprd, C.d_sx, C.d_st, C.d_cj_bbox);
    MDG_LAUNCH_CHECK("cp_setup_pbc");
    C.nghost_atoms        = C.h_status[3];
    atom->Nclusters_local = C.nci;
    atom->ncj             = C.ncj;
    atom->Nclusters_ghost = ng;
    atom->dummy_cj        = C.dummy_cj;
    atom->Nghost          = C.nghost_atoms;
}

void updatePbcClusterGFX950(Atom* atom, Parameter* param, bool firstUpdate)
{
    (void)atom;
    if (!ok() || C.nghost == 0) return;
    TimedRegion tr(F_PBC);
    if (firstUpdate) {
        C.sxn_valid = false;
        C.rec_valid = false;
        k_cp_ghost_first<<<blocks_for(C.nghost + 1, BLOCK), BLOCK, 0, stream()>>>(C.ncj, C.nghost,
            C.N, C.d_border, C.d_pbc, C.d_cj_natoms, param->xprd, param->yprd, param->zprd, C.d_sx,
            C.d_st, C.d_cj_bbox);
    } else {
        // the records of the ghost slots too, while the other records are current (each
        // record buffer then holds every slot)
        long long n = (long long)C.nghost * C.N;
#define MDG_GU(NN)                                                                               \
    k_cp_ghost_update<NN><<<blocks_for(n, BLOCK), BLOCK, 0, stream()>>>(C.ncj, C.nghost,         \
        C.d_border, C.d_pbc, C.d_cj_natoms, param->xprd, param->yprd, param->zprd, C.d_sx, C.d_st, \
        C.rg, C.rec_valid ? C.d_sr : nullptr, C.d_status + ST_REC_WORD)
        if (C.N == 4) MDG_GU(4);
        else if (C.N == 8) MDG_GU(8);
        else MDG_GU(2);
#undef MDG_GU
    }
    MDG_LAUNCH_CHECK("cp_update_pbc");
}

void binJClustersGFX950(Parameter* param, Atom* atom)
{
    (void)param;
    (void)atom;
    if (!ok()) return;
    TimedRegion tr(F_CLUSTER);
    hipStream_t s = stream();
    const int mb = C.g.mbins, ng = C.nghost;
    grow_n(&C.d_gbin, &C.cap_gbin, (size_t)ng + 1, 0);
    grow_n(&C.d_gcur, &C.cap_gcur, (size_t)mb + 1, 0);
    grow_n(&C.d_gbin_start, &C.cap_gbin_start, (size_t)mb + 1, 0);
    grow_n(&C.d_gbin_list, &C.cap_gbin_list, (size_t)ng + 1, 0);
    grow_n(&C.d_jcnt, &C.cap_jcnt, (size_t)mb + 1, 0);
    grow_n(&C.d_jstart, &C.cap_jstart, (size_t)mb + 1, 0);
    grow_n(&C.d_jlist, &C.cap_jlist, (size_t)C.ncj + ng + 1, 0);
    grow_n(&C.d_jkey, &C.cap_jkey, 6 * ((size_t)C.ncj + ng + 1), 0);
    if (failed()) return;
    fill_i32_multi({ { C.d_status, 0, 1 }, { C.d_gcur, 0, (size_t)mb + 1 } });
    k_cp_ghost_bin<<<blocks_for(ng, BLOCK), BLOCK, 0, s>>>(C.ncj, ng, C.N, C.g, C.d_cj_natoms,
        C.d_sx, C.d_pbc, C.d_gbin, C.d_gcur, C.d_status);
    if (mb <= ONE_T * ONE_I) {
        k_cp_jbin_starts<<<1, ONE_T, 0, s>>>(mb, C.M, C.N, C.d_ncl, C.d_gcur, C.d_gbin_start,
            C.d_jstart);
    } else {
        exclusive_scan(C.d_gcur, C.d_gbin_start, mb, C.d_gbin_start + mb);
        k_cp_jbin_count<<<blocks_for(mb, BLOCK), BLOCK, 0, s>>>(mb, C.M, C.N, C.d_ncl, C.d_gcur,
            C.d_jcnt);
        exclusive_scan(C.d_jcnt, C.d_jstart, mb, C.d_jstart + mb);
        fill_i32(C.d_gcur, 0, (size_t)mb + 1);
    }
    k_cp_ghost_scatter<<<blocks_for(ng, BLOCK), BLOCK, 0, s>>>(ng, C.d_gbin, C.d_gbin_start,
        C.d_gcur, C.d_gbin_list);
    k_cp_jbin_merge<<<mb, 64, 0, s>>>(C.M, C.N, C.ncj, C.d_ncl, C.d_cstart, C.d_gbin_start,
        C.d_gbin_list, C.d_cj_bbox, C.d_jstart, C.d_jlist, C.d_jkey, C.d_status);
    // z boundaries spanning the ghost layers below and above the box; keys outside the span
    // only make the windows wider
    grow_n(&C.d_zidx, &C.cap_zidx, (size_t)mb * NZB, 0);
    if (failed()) return;
    {
        const double zlo_all = -(C.cutneigh + 4.0), zhi_all = param->zprd + C.cutneigh + 4.0;
        const double dz      = (zhi_all - zlo_all) / (NZB - 1);
        C.z0                 = zlo_all;
        C.inv_dz             = 1.0 / dz;
        k_cp_zindex<<<blocks_for((long long)mb * NZB, BLOCK), BLOCK, 0, s>>>(mb, C.d_jstart,
            C.d_jkey, C.z0, dz, C.d_zidx);
    }
    MDG_LAUNCH_CHECK("cp_bin_j");
}

static void tiles_launch();
static void tiles_finish();
static void ensure_records();

void buildNeighborClusterGFX950(Atom* atom, Neighbor* neighbor)
{
    if (!ok()) return;
    C.tiles_valid = false;
    TimedRegion tr(F_NEIGHBOR);
    hipStream_t s = stream();
    const int nci = C.nci;
    if (neighbor->maxneighs > C.maxneighs) C.maxneighs = neighbor->maxneighs;
    // largest j-cluster height (status words 12-13 hold it as a 64-bit pattern)
    static_assert(ST_REC_WORD != 12 && ST_REC_WORD != 13, "status words 12..13 hold hmax");
    unsigned long long* d_hmax = reinterpret_cast<unsigned long long*>(C.d_status + 12);
    // hmax and the first attempt's flag and count words
    fill_i32_multi({ { C.d_status + 12, 0, 2 }, { C.d_status, 0, 2 } });
    const int ncj_all = C.ncj + C.nghost;
    if (ncj_all > 0)
        k_cp_jheight<<<std::min(blocks_for(ncj_all, BLOCK), 512u), BLOCK, 0, s>>>(ncj_all,
            C.d_cj_bbox, d_hmax);
    const double Rwin = sqrt(C.cutneighsq) * (1.0 + 1e-12) + 1e-9;
    for (int attempt = 0; attempt < 8; attempt++) {
        grow_n(&C.d_nb, &C.cap_nb, (size_t)nci * C.maxneighs + 1, 0);
        grow_n(&C.d_numneigh, &C.cap_numneigh, (size_t)nci + 1, 0);
        grow_n(&C.d_nmasked, &C.cap_nmasked, (size_t)nci + 1, 0);
        if (failed()) return;
        if (attempt > 0) fill_i32(C.d_status, 0, 2);
        unsigned grid = blocks_for(nci, BLOCK / 64);
#define MDG_NB1(MM, NN, FM)                                                                      \
    k_cp_build_nb<MM, NN, FM><<<grid, BLOCK, 0, s>>>(nci, C.rule, C.fma_dist, neighbor->half_neigh, \
        C.d_stencil, C.nstencil, C.g.mbins, C.d_ci_bin, C.d_ci_bbox, C.d_jkey, C.d_jstart,         \
        C.d_jlist, C.d_zidx, C.z0, C.inv_dz, C.d_sx, C.cutneighsq, Rwin, C.rbb_sq, d_hmax,        \
        C.maxneighs, C.d_nb, C.d_numneigh, C.d_nmasked, C.d_status)
#define MDG_NB(MM, NN)                                                                           \
    do {                                                                                         \
        if (C.fma_dist) MDG_NB1(MM, NN, 1);                                                      \
        else MDG_NB1(MM, NN, 0);                                                                 \
    } while (0)
        if (C.M == 4 && C.N == 4) MDG_NB1(4, 4, 2);
        else if (C.M == 4 && C.N == 8) MDG_NB(4, 8);
        else MDG_NB(1, 2);
#undef MDG_NB
#undef MDG_NB1
        MDG_LAUNCH_CHECK("cp_build_nb");
        // the tile tables of an FP32 half list and the records of the new slots are launched
        // before the status read, which then serves them too (the force call would otherwise
        // launch them after it, each launch then waiting on the host)
        const bool half = neighbor->half_neigh != 0;
        if (C.precision == 0 && half && !C.comp) tiles_launch();
        if (C.precision == 0 && !(half && C.comp)) ensure_records();
        read_status();
        if (failed()) return;
        if (C.h_status[0] & (ST_BIN_RANGE | ST_JBIN_CAP)) {
            set_error(E_BIN, "cluster-pair list build failed (flags %d)", C.h_status[0]);
            return;
        }
        if (!(C.h_status[0] & ST_NB_OVERFLOW)) {
            neighbor->maxneighs = C.maxneighs;
            tiles_finish();
            (void)atom;
            return;
        }
        C.tiles_valid = false;
        C.tiles_wait  = false;
        // neighbor.c:705-720: grow to 1.2x the largest count and rebuild
        C.maxneighs = (int)(C.h_status[1] * 1.2) + 1;
    }
    set_error(E_CAPACITY, "cluster-pair list did not converge");
}

void pruneNeighborClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor)
{
    (void)param;
    (void)atom;
    if (!ok()) return;
    C.tiles_valid = false;
    TimedRegion tr(F_NEIGHBOR);
    hipStream_t s = stream();
    const int nci = C.nci;
    if (nci <= 0) return;
    fill_i32(C.d_status, 0, 2);
    unsigned grid = blocks_for(nci, BLOCK / 64);
#define MDG_PRUNE(MM, NN)                                                                        \
    k_cp_prune<MM, NN><<<grid, BLOCK, 0, s>>>(nci, C.fma_dist, C.d_sx, C.cutneighsq, C.maxneighs,  \
        C.d_nb, C.d_numneigh, C.d_nmasked, C.d_status)
    if (C.M == 4 && C.N == 4) MDG_PRUNE(4, 4);
    else if (C.M == 4 && C.N == 8) MDG_PRUNE(4, 8);
    else MDG_PRUNE(1, 2);
#undef MDG_PRUNE
    MDG_LAUNCH_CHECK("cp_prune");
    if (C.precision == 0 && neighbor->half_neigh && !C.comp) tiles_launch();  // as in the build
    read_status();
    if (failed()) return;
    if (C.h_status[0] & ST_NB_OVERFLOW)
        set_error(E_CAPACITY, "cluster-pair prune: a list longer than %d entries", PRUNE_CAP);
    neighbor->maxneighs = C.maxneighs;
    tiles_finish();
}

// Tile tables and the reverse map of the current half list (see k_cp_tile_table).  A tile that
// outgrows its tables leaves tiles_ok false: that list runs the atomic half-list kernel.
constexpr int ST_TILE_WORD = 15;  // status word of the tile-table kernel
constexpr int ST_TILE_NMAX = 11;  // the largest tile table of the current list (slots)
// The tile-table kernel of the current list, before a status read: the list build and the
// prune launch it before their own status read, so that one read serves both.
static void tiles_launch()
{
    hipStream_t s = stream();
    const int nci = C.nci, ncj = C.ncj, N = C.N;
    const int ntiles = (nci + TILE_CI - 1) / TILE_CI, ucap = TILE_W64 / (3 * N);
    C.ntiles      = ntiles;
    C.tile_ucap   = ucap;
    C.tiles_valid = true;
    C.tiles_ok    = false;
    C.tiles_wait  = false;
    if (ntiles == 0) return;
    const size_t nrec = (size_t)ntiles * ucap;
    grow_n(&C.d_tile_cj, &C.cap_tile_cj, nrec, 0);
    grow_n(&C.d_tile_n, &C.cap_tile_n, (size_t)ntiles + 1, 0);
    grow_n(&C.d_eslot, &C.cap_eslot, (size_t)nci * C.maxneighs + 1, 0);
    grow_n(&C.d_rec_off, &C.cap_rec_off, (size_t)ntiles + 1, 0);
    grow_n(&C.d_rcount, &C.cap_rcount, (size_t)ncj + 1, 0);
    grow_n(&C.d_rstart, &C.cap_rstart, (size_t)ncj + 1, 0);
    grow_n(&C.d_rcur, &C.cap_rcur, (size_t)ncj + 1, 0);
    grow_n(&C.d_rev, &C.cap_rev, nrec, 0);
    grow_n(&C.d_rec, &C.cap_rec, nrec * 3 * N, 0);
    grow_n(&C.d_fown, &C.cap_fown, 3 * (size_t)nci * C.M, 0);
    if (failed()) return;
    fill_i32_multi({ { C.d_status + ST_TILE_WORD, 0, 1 }, { C.d_status + ST_TILE_NMAX, 0, 1 },
        { C.d_rcount, 0, (size_t)ncj + 1 } });
    k_cp_tile_table<<<ntiles, TILE_TT, 0, s>>>(nci, ncj, ucap, C.d_nb, C.d_numneigh, C.maxneighs,
        C.d_tile_cj, C.d_tile_n, C.d_eslot, C.d_status + ST_TILE_WORD, C.d_status + ST_TILE_NMAX,
        C.d_rcount);
    MDG_LAUNCH_CHECK("cp_tile_table");
    C.tiles_wait = true;
}

// After the status read that followed tiles_launch: the reverse map, unless a tile outgrew its
// tables.
static void tiles_finish()
{
    if (!C.tiles_wait) return;
    C.tiles_wait = false;
    if (failed() || C.h_status[ST_TILE_WORD]) return;
    C.tile_nmax = C.h_status[ST_TILE_NMAX];
    hipStream_t s = stream();
    const int ntiles = C.ntiles, ucap = C.tile_ucap, ncj = C.ncj;
    const int w = std::max(C.tile_nmax, 1);  // threads per tile of the reverse-map kernels
    const size_t nrec = (size_t)ntiles * w;
    exclusive_scan(C.d_tile_n, C.d_rec_off, ntiles, C.d_rec_off + ntiles);
    fill_i32(C.d_rcur, 0, (size_t)ncj + 1);
    exclusive_scan(C.d_rcount, C.d_rstart, ncj, C.d_rstart + ncj);
    k_cp_rev_fill<<<blocks_for((long long)nrec, BLOCK), BLOCK, 0, s>>>(ntiles, ucap, w,
        C.d_tile_cj, C.d_tile_n, C.d_rec_off, C.d_rstart, C.d_rcur, C.d_rev);
    k_cp_rev_sort<<<blocks_for(ncj, BLOCK), BLOCK, 0, s>>>(ncj, C.d_rstart, C.d_rev);
    MDG_LAUNCH_CHECK("cp_rev_map");
    C.tiles_ok = !failed();
}

static void build_tiles()
{
    tiles_launch();
    if (!C.tiles_wait) return;
    read_status();
    tiles_finish();
}

// The records of every slot (the FP32 full-list and tiled half-list kernels read them), rebuilt
// from the positions whenever something other than the integrating force launch and the ghost
// update has moved them.
static void ensure_records()
{
    const size_t nsl = nslots_all(), cap0 = C.cap_sr;
    grow_n(&C.d_sr, &C.cap_sr, nsl, 0);
    if (failed()) return;
    if (C.cap_sr != cap0) C.rec_valid = false;
    if (!C.rec_valid) {
        k_cp_records<<<blocks_for((long long)nsl, BLOCK), BLOCK, 0, stream()>>>((int)nsl, C.d_sx,
            C.d_st, C.rg, C.d_sr, C.d_status + ST_REC_WORD);
        MDG_LAUNCH_CHECK("cp_records");
        C.rec_valid = true;
    }
}

// The force call; with fuse_final the FP32 full-list kernel also applies finalIntegrate to each
// atom whose total force it has just written, and returns true (false: not fused, the caller
// integrates).
// What a force call also does in the same kernel; the caller runs the rest separately.
enum { FUSED_FINAL = 1, FUSED_INITIAL = 2 };
static int force_cluster(Parameter* param, Neighbor* neighbor, bool fuse_final,
    bool fuse_initial, double* secs);

double computeForceLJClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats)
{
    (void)atom;
    (void)stats;
    double secs = 0.0;
    force_cluster(param, neighbor, false, false, &secs);
    return secs;
}

double computeForceFinalClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats)
{
    (void)stats;
    double secs = 0.0;
    if (!(force_cluster(param, neighbor, true, false, &secs) & FUSED_FINAL))
        finalIntegrateClusterGFX950(param, atom);
    return secs;
}

double computeForceFinalInitialClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats)
{
    (void)stats;
    double secs   = 0.0;
    const int got = force_cluster(param, neighbor, true, true, &secs);
    if (!(got & FUSED_FINAL)) finalIntegrateClusterGFX950(param, atom);
    if (!(got & FUSED_INITIAL)) initialIntegrateClusterGFX950(param, atom);
    return secs;
}

static int force_cluster(Parameter* param, Neighbor* neighbor, bool fuse_final,
    bool fuse_initial, double* secs)
{
    if (!ok()) return 0;
    flush_atoms();  // a pending copy reads the slots before this call moves them
    const double t0 = wtime();
    int fused       = 0;
    hipStream_t s   = stream();
    const int nci = C.nci, half = neighbor->half_neigh;
    const size_t nloc = (size_t)nci * C.M;
    // FP32 half lists (plain sums): tile tables of the current list, then the tiled kernels
    if (C.precision == 0 && half && !C.comp && !C.tiles_valid) build_tiles();
    if (failed()) return 0;
    const bool tiled = C.precision == 0 && half && !C.comp && C.tiles_ok;
    // the FP32 full-list and tiled half-list kernels read records
    const bool recs = C.precision == 0 && (!half || tiled);
    if (recs) ensure_records();
    if (failed()) return 0;
    // the next step's positions and their records go to the second buffers, whose padding,
    // dummy and ghost slots must match the current ones (copied once after each rebuild or
    // upload)
    double* xn = nullptr;
    float4* rn = nullptr;
    if (fuse_initial && fuse_final && recs) {
        const size_t nsl = nslots_all(), n3 = 3 * nsl, cap0 = C.cap_sxn, capr0 = C.cap_srn;
        grow_n(&C.d_sxn, &C.cap_sxn, n3, 0);
        grow_n(&C.d_srn, &C.cap_srn, nsl, 0);
        if (failed()) return 0;
        if (C.cap_sxn != cap0 || C.cap_srn != capr0) C.sxn_valid = false;
        if (!C.sxn_valid) {
            MDG_HIP(hipMemcpyAsync(C.d_sxn, C.d_sx, n3 * sizeof(double),
                hipMemcpyDeviceToDevice, s));
            MDG_HIP(hipMemcpyAsync(C.d_srn, C.d_sr, nsl * sizeof(float4),
                hipMemcpyDeviceToDevice, s));
            C.sxn_valid = true;
        }
        xn = C.d_sxn;
        rn = C.d_srn;
    }
    TimedRegion tr(F_FORCE_CP);
    if (C.precision == 0 &&
        ((size_t)C.ncj + C.nghost + 1 >= (1u << 24) || nslots_all() * 24 >= (1ull << 32))) {
        set_error(E_ARG, "FP32 cluster kernel: %zu slots exceed its 32-bit addressing",
            nslots_all());
        return 0;
    }
    // energy and virial: one slot pair per wavefront of the force launch, cleared before it
    size_t ev_nw = 0;
    auto ev_parts = [&](size_t nw) -> double* {
        if (!C.ev_on) return C.d_ev;
        grow_n(&C.d_evp, &C.cap_evp, 2 * nw, 0);
        if (failed()) return nullptr;
        fill_f64(C.d_evp, 0.0, 2 * nw);
        ev_nw = nw;
        return C.d_evp;
    };
    if (C.precision == 1) {
        grow_n(&C.d_sf64, &C.cap_sf64, 3 * nslots_all(), 0);
        if (failed()) return 0;
        if (half) fill_f64(C.d_sf64, 0.0, 3 * nslots_all());
        unsigned grid = blocks_for((long long)nloc, BLOCK);
        double* evp   = ev_parts((size_t)grid * (BLOCK / 64));
        if (failed()) return 0;
        const int uni = C.uniform ? 1 : 0;
#define MDG_F64(MM, NN, H, E)                                                                    \
    k_cp_force_f64<MM, NN, H, E><<<grid, BLOCK, 0, s>>>(nci, C.ncj, C.ntypes, uni, C.t64, C.d_sx, \
        C.d_st, C.d_nb, C.d_numneigh, C.maxneighs, C.d_sf64, evp)
#define MDG_F64_SHAPE(MM, NN)                                                                    \
    do {                                                                                         \
        if (half) {                                                                              \
            if (C.ev_on) MDG_F64(MM, NN, true, true);                                            \
            else MDG_F64(MM, NN, true, false);                                                   \
        } else {                                                                                 \
            if (C.ev_on) MDG_F64(MM, NN, false, true);                                           \
            else MDG_F64(MM, NN, false, false);                                                  \
        }                                                                                        \
    } while (0)
        if (C.M == 4 && C.N == 4) MDG_F64_SHAPE(4, 4);
        else if (C.M == 4 && C.N == 8) MDG_F64_SHAPE(4, 8);
        else MDG_F64_SHAPE(1, 2);
#undef MDG_F64_SHAPE
#undef MDG_F64
        C.forces_are_f32 = false;
    } else {
        grow_n(&C.d_sf32, &C.cap_sf32, 3 * nslots_all(), 0);
        if (failed()) return 0;
        if (half && !tiled) fill_f32(C.d_sf32, 0.0f, 3 * nslots_all());
        unsigned grid = (blocks_for(nci, BLOCK / 64) + 7) & ~7u;  // multiple of 8, see kernel
        // full lists: four i-clusters per wavefront; tiled half lists: the gather pass (both with
        // the fused final integrate)
        fused          = (fuse_final && (!half || tiled)) ? FUSED_FINAL | (xn ? FUSED_INITIAL : 0) : 0;
        double* fuse_v = fused ? C.d_sv : nullptr;
        unsigned grid4 = (blocks_for((nci + 3) / 4, BLOCK / 64) + 7) & ~7u;
        double* evp    = ev_parts(tiled ? (size_t)C.ntiles * TILE_WAVES
                                        : (size_t)(half ? grid : grid4) * (BLOCK / 64));
        if (failed()) return 0;
#define MDG_F32(NN, H, E, U, CO)                                                                 \
    do {                                                                                         \
        if (H)                                                                                   \
            k_cp_force_f32<NN, true, E, U, CO><<<grid, BLOCK, 0, s>>>(nci, C.ncj, C.dummy_cj,     \
                C.ntypes, C.t32, C.d_sx, C.d_st, C.d_nb, C.d_numneigh, C.d_ci_natoms,             \
                C.maxneighs, C.d_sf32, evp);                                                      \
        else                                                                                     \
            k_cp_force_f32_full<NN, E, U, CO><<<grid4, BLOCK, 0, s>>>(nci, C.ncj, C.dummy_cj,     \
                C.ntypes, C.t32, C.d_sx, C.d_sr, C.d_st, C.d_nb, C.d_numneigh, C.d_ci_natoms,     \
                C.d_ci_bbox, C.rg, C.maxneighs, C.d_sf32, evp, fuse_v, param->dtforce, xn,       \
                rn, param->dt, C.d_status + ST_REC_WORD);                                        \
    } while (0)
#define MDG_F32_U(NN, H, E)                                                                      \
    do {                                                                                         \
        if (C.uniform) {                                                                         \
            if (C.comp) MDG_F32(NN, H, E, true, true);                                           \
            else MDG_F32(NN, H, E, true, false);                                                 \
        } else {                                                                                 \
            if (C.comp) MDG_F32(NN, H, E, false, true);                                          \
            else MDG_F32(NN, H, E, false, false);                                                \
        }                                                                                        \
    } while (0)
#define MDG_F32_N(NN)                                                                            \
    do {                                                                                         \
        if (half) {                                                                              \
            if (C.ev_on) MDG_F32_U(NN, true, true);                                              \
            else MDG_F32_U(NN, true, false);                                                     \
        } else {                                                                                 \
            if (C.ev_on) MDG_F32_U(NN, false, true);                                             \
            else MDG_F32_U(NN, false, false);                                                    \
        }                                                                                        \
    } while (0)
        if (tiled) {
            // the kernel clears and uses the first tile_n[t] slots of its table
            const size_t lds = (size_t)std::max(C.tile_nmax, 1) * 3 * C.N * sizeof(unsigned long long);
            const unsigned gg = blocks_for((long long)nloc, BLOCK);
#define MDG_HT(NN, E, U)                                                                         \
    k_cp_force_half_tiled<NN, E, U><<<C.ntiles, TILE_WAVES * 64, lds, s>>>(nci, C.ncj, C.dummy_cj, \
        C.ntypes, C.t32, C.d_sx, C.d_sr, C.d_nb, C.d_eslot, C.d_numneigh, C.d_ci_bbox, C.rg,      \
        C.maxneighs, C.tile_ucap, C.d_tile_n, C.d_rec_off, C.d_rec, C.d_fown, evp)
#define MDG_HT_N(NN)                                                                             \
    do {                                                                                         \
        if (C.ev_on) {                                                                           \
            if (C.uniform) MDG_HT(NN, true, true);                                               \
            else MDG_HT(NN, true, false);                                                        \
        } else {                                                                                 \
            if (C.uniform) MDG_HT(NN, false, true);                                              \
            else MDG_HT(NN, false, false);                                                       \
        }                                                                                        \
        k_cp_gather_half<NN><<<gg, BLOCK, 0, s>>>((int)nloc, C.M, C.d_ci_natoms, C.d_fown,       \
            C.d_rec, C.d_rstart, C.d_rev, C.d_sf32, C.d_sx, C.d_st, fuse_v, param->dtforce, xn,  \
            rn, param->dt, C.rg, C.d_status + ST_REC_WORD);                                      \
    } while (0)
            if (C.N == 4) MDG_HT_N(4);
            else MDG_HT_N(8);
#undef MDG_HT_N
#undef MDG_HT
        } else if (C.N == 4) MDG_F32_N(4);
        else MDG_F32_N(8);
#undef MDG_F32_N
#undef MDG_F32_U
#undef MDG_F32
        C.forces_are_f32 = true;
    }
    MDG_LAUNCH_CHECK("cp_force");
    if (ev_nw) {
        k_cp_sum_ev<<<1, EVR_T, 0, s>>>(ev_nw, C.d_evp, C.d_ev);
        MDG_LAUNCH_CHECK("cp_sum_ev");
    }
    if (fused & FUSED_INITIAL) {
        std::swap(C.d_sx, C.d_sxn);
        std::swap(C.cap_sx, C.cap_sxn);
        std::swap(C.d_sr, C.d_srn);
        std::swap(C.cap_sr, C.cap_srn);
    }
    *secs = wtime() - t0;
    return fused;
}

void initialIntegrateClusterGFX950(Parameter* param, Atom* atom)
{
    (void)atom;
    if (!ok()) return;
    flush_atoms();
    TimedRegion tr(F_INTEGRATE);
    int n3 = 3 * C.nci * C.M;
    if (!(C.forces_are_f32 ? (void*)C.d_sf32 : (void*)C.d_sf64)) {
        set_error(E_ARG, "initialIntegrateClusterGFX950 before the first force call");
        return;
    }
    C.rec_valid = false;  // the records follow at the next FP32 force call
    if (C.forces_are_f32)
        k_cp_initial<float><<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, C.M, C.d_ci_natoms,
            C.d_sx, C.d_sv, C.d_sf32, param->dt, param->dtforce);
    else
        k_cp_initial<double><<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, C.M, C.d_ci_natoms,
            C.d_sx, C.d_sv, C.d_sf64, param->dt, param->dtforce);
    MDG_LAUNCH_CHECK("cp_initial");
}

void finalIntegrateClusterGFX950(Parameter* param, Atom* atom)
{
    (void)atom;
    if (!ok()) return;
    flush_atoms();
    TimedRegion tr(F_INTEGRATE);
    int n3 = 3 * C.nci * C.M;
    if (!(C.forces_are_f32 ? (void*)C.d_sf32 : (void*)C.d_sf64)) {
        set_error(E_ARG, "finalIntegrateClusterGFX950 before the first force call");
        return;
    }
    if (C.forces_are_f32)
        k_cp_final<float><<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, C.M, C.d_ci_natoms,
            C.d_sv, C.d_sf32, param->dtforce);
    else
        k_cp_final<double><<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, C.M, C.d_ci_natoms,
            C.d_sv, C.d_sf64, param->dtforce);
    MDG_LAUNCH_CHECK("cp_final");
}

double kineticSumClusterGFX950(Parameter* param, Atom* atom)
{
    (void)atom;
    if (!ok()) return 0.0;
    TimedRegion tr(F_THERMO);
    const int nb = KSUM_PARTS;
    grow_n(&C.d_red, &C.cap_red, (size_t)nb, 0);
    if (failed()) return 0.0;
    k_cp_ksum<<<nb, BLOCK, 0, stream()>>>(C.nci * C.M, C.M, C.d_ci_natoms, C.d_sv, param->mass, C.d_red);
    k_cp_sum_partials<<<1, BLOCK, 0, stream()>>>(nb, C.d_red, C.d_red);
    MDG_LAUNCH_CHECK("cp_ksum");
    MDG_HIP(hipMemcpyAsync(C.h_red, C.d_red, sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? 0.0 : C.h_red[0];
}

int kineticSumClusterNoAllocGFX950(Parameter* param, Atom* atom, double* d_out)
{
    (void)atom;
    if (!ok()) return failed() ? mdgfx950Status() : E_ARG;
    if (!d_out) {
        set_error(E_ARG, "kineticSumClusterNoAllocGFX950: no output pointer");
        return E_ARG;
    }
    if ((long long)C.nci * C.M > INT32_MAX) {
        set_error(E_ARG, "kineticSumClusterNoAllocGFX950: %lld slots exceed its 32-bit indexing",
            (long long)C.nci * C.M);
        return E_ARG;
    }
    int dev = 0;
    MDG_HIP(hipGetDevice(&dev));
    if (failed()) return mdgfx950Status();
    double* part = (dev >= 0 && dev < MAX_DEVICES) ? g_scratch[dev] : nullptr;
    if (!part) {
        set_error(E_ARG, "kineticSumClusterNoAllocGFX950: device %d was not set up by "
                         "initDeviceClusterGFX950", dev);
        return E_ARG;
    }
    // the same kernels and partial count as kineticSumClusterGFX950, so the same sum; every
    // partial is written before it is read, so nothing is cleared
    k_cp_ksum<<<KSUM_PARTS, BLOCK, 0, stream()>>>(C.nci * C.M, C.M, C.d_ci_natoms, C.d_sv,
        param->mass, part);
    k_cp_sum_partials<<<1, BLOCK, 0, stream()>>>(KSUM_PARTS, part, d_out);
    MDG_LAUNCH_CHECK("cp_ksum_noalloc");
    return mdgfx950Status();
}

void requestEnergyVirialClusterGFX950(int on) { C.ev_on = on != 0; }

void energyVirialClusterGFX950(double* epot, double* virial)
{
    double h[2] = { 0, 0 };
    if (!failed() && C.d_ev) {
        MDG_HIP(hipMemcpyAsync(h, C.d_ev, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
        MDG_HIP(hipStreamSynchronize(stream()));
    }
    if (epot) *epot = h[0];
    if (virial) *virial = h[1];
}

void clusterInfoGFX950(int* info)
{
    (void)ok();  // reads a pending cluster count back
    int v[11] = { C.nci, C.ncj, C.nghost, C.dummy_cj, C.nghost_atoms, C.g.mbins, C.nstencil,
        C.maxneighs, C.M, C.N, C.natoms };
    memcpy(info, v, sizeof(v));
}

long long uploadClusterArrayGFX950(const char* name, const void* in, long long n)
{
    if (!ok()) return -3;
    flush_atoms();  // a pending copy reads the slots as they were
    void* dst       = nullptr;
    long long avail = 0;
    if (!strcmp(name, "slot_x")) { dst = C.d_sx; avail = 3LL * (long long)nslots_all(); }
    else if (!strcmp(name, "slot_v")) { dst = C.d_sv; avail = 3LL * C.nci * C.M; }
    if (!dst) return -1;
    if (n > avail || n < 0) return -2;
    if (dst == C.d_sx) C.sxn_valid = C.rec_valid = false;
    MDG_HIP(hipMemcpyAsync(dst, in, (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? -3 : n;
}

long long downloadClusterArrayGFX950(const char* name, void* out, long long cap)
{
    if (!ok()) return -3;
    flush_atoms();  // the atom arrays as the calls so far leave them
    hipStream_t s = stream();
    MDG_HIP(hipStreamSynchronize(s));
    const size_t nsl = nslots_all(), nj = (size_t)C.ncj + C.nghost;
    const void* src  = nullptr;
    long long n = 0, es = 4;
    if (!strcmp(name, "slot_x")) { src = C.d_sx; n = 3 * (long long)nsl; es = 8; }
    else if (!strcmp(name, "slot_v")) { src = C.d_sv; n = 3LL * C.nci * C.M; es = 8; }
    else if (!strcmp(name, "slot_t")) { src = C.d_st; n = (long long)nsl; }
    else if (!strcmp(name, "ci_natoms")) { src = C.d_ci_natoms; n = C.nci; }
    else if (!strcmp(name, "ci_bin")) { src = C.d_ci_bin; n = C.nci; }
    else if (!strcmp(name, "ci_bbox")) { src = C.d_ci_bbox; n = 6LL * C.nci; es = 8; }
    else if (!strcmp(name, "cj_natoms")) { src = C.d_cj_natoms; n = (long long)nj; }
    else if (!strcmp(name, "cj_bbox")) { src = C.d_cj_bbox; n = 6LL * nj; es = 8; }
    else if (!strcmp(name, "ghost_map")) { src = C.d_border; n = C.nghost; }
    else if (!strcmp(name, "ghost_pbc")) { src = C.d_pbc; n = 3LL * C.nghost; }
    else if (!strcmp(name, "numneigh")) { src = C.d_numneigh; n = C.nci; }
    else if (!strcmp(name, "numneigh_masked")) { src = C.d_nmasked; n = C.nci; }
    else if (!strcmp(name, "cbin_start")) { src = C.d_jstart; n = C.g.mbins + 1; }
    else if (!strcmp(name, "cbin_list")) { src = C.d_jlist; n = (long long)nj; }
    else if (!strcmp(name, "atoms_x")) { src = C.d_ax; n = 3LL * C.natoms; es = 8; }
    else if (!strcmp(name, "atoms_v")) { src = C.d_av; n = 3LL * C.natoms; es = 8; }
    else if (!strcmp(name, "atoms_type")) { src = C.d_at; n = C.natoms; }
    else if (!strcmp(name, "slot_f")) {
        n = 3LL * C.nci * C.M;
        if (n > cap) return -2;
        if (C.forces_are_f32) {
            float* tmp = (float*)malloc(n * sizeof(float) + 4);
            MDG_HIP(hipMemcpy(tmp, C.d_sf32, n * sizeof(float), hipMemcpyDeviceToHost));
            for (long long k = 0; k < n; k++)
                ((double*)out)[k] = tmp[k];
            free(tmp);
        } else {
            MDG_HIP(hipMemcpy(out, C.d_sf64, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        return failed() ? -3 : n;
    } else if (!strcmp(name, "neighbors")) {
        int* nn   = (int*)malloc(sizeof(int) * (C.nci + 1));
        int* full = (int*)malloc(sizeof(int) * ((size_t)C.nci * C.maxneighs + 1));
        MDG_HIP(hipMemcpy(nn, C.d_numneigh, C.nci * sizeof(int), hipMemcpyDeviceToHost));
        MDG_HIP(hipMemcpy(full, C.d_nb, (size_t)C.nci * C.maxneighs * sizeof(int),
            hipMemcpyDeviceToHost));
        long long k = 0;
        for (int ci = 0; ci < C.nci && k >= 0; ci++)
            for (int m = 0; m < nn[ci]; m++) {
                if (k >= cap) {
                    k = -2;
                    break;
                }
                ((int*)out)[k++] = full[(size_t)ci * C.maxneighs + m];
            }
        free(nn);
        free(full);
        return failed() ? -3 : k;
    } else {
        return -1;
    }
    if (n > cap) return -2;
    if (n) MDG_HIP(hipMemcpy(out, src, n * es, hipMemcpyDeviceToHost));
    return failed() ? -3 : n;
}

}  // extern "C"
