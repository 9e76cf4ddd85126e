// Verlet-list scheme on gfx950: ghost map, binning, neighbour build, LJ (full/half), EAM,
// velocity Verlet and the kinetic-energy reduction.  All arrays stay on the device between
// calls.  Layouts: x, v, f AoS (3 doubles per atom, the reference's ATOM_POSITION_AOS);
// neighbour list SoA (entry k of atom i at k*Nlocal + i), the reference's GPU layout; bins CSR.
//
// Strict kernels keep the reference's evaluation order and never fuse multiply-adds (the whole
// file is built with -ffp-contract=off), so FP64 results are bit-identical to MD-Bench compiled
// the same way.  Fast kernels spread each atom's list over LPA lanes and use explicit fma().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "mdgfx950.h"
#include "runtime.h"

using namespace mdg;

namespace {

constexpr int BLOCK     = 256;
constexpr int MAXTYPES2 = 64;  // ntypes <= 8
constexpr int FAST_LPA  = 4;   // lanes per atom in the fast force kernels

struct Grid {
    double off[3];     // pad * binsize + 1e-9 (coord2bin, neighbor.c:386-400)
    double bininv[3];
    int mbinx, mbiny, mbinz, mbins;
};

struct Tables {
    double eps[MAXTYPES2], sig6[MAXTYPES2], cutf[MAXTYPES2], cutn[MAXTYPES2];
};

struct State {
    bool ready = false;
    // grid (setupNeighborGFX950)
    Grid g{};
    double binsize[3]{};
    int pad[3]{}, nstencil = 0;
    int* d_stencil = nullptr;
    size_t stencil_cap = 0;
    double prd[3]{}, cutneigh = 0;
    // atoms
    size_t cap_all = 0, cap_ghost = 0, cap_v = 0;
    // second position buffer: a force call that also integrates the next step writes the new
    // local positions there (its other lanes still read the current ones) and the two swap
    double* d_xn  = nullptr;
    size_t cap_xn = 0;
    // atom sort: the bin-ordered copies of x, v and f (9 doubles per atom) and of the types
    double* d_sort    = nullptr;
    size_t cap_sort   = 0;
    int* d_sort_t     = nullptr;
    size_t cap_sort_t = 0;
    int ntypes = 0;
    int ncu    = 0;  // compute units, rounded up to a multiple of 8 (persistent EAM grids)
    bool uniform = true;  // every type pair has the same LJ parameters
    Tables tab{};
    // bins
    int* d_bin_of = nullptr;
    size_t cap_bin_of = 0;
    int* d_bincount = nullptr;
    int* d_binstart = nullptr;
    size_t cap_bins = 0, cap_binstart = 0;
    int* d_binatoms = nullptr;
    size_t cap_binatoms = 0;
    bool binned = false;
    // per-atom EAM energies (training labels for the learned potential)
    double* d_eatom = nullptr;
    size_t cap_eatom = 0;
    bool eatom_on = false;
    // ghosts
    int* d_gcnt = nullptr;
    int* d_goff = nullptr;
    size_t cap_gcnt = 0, cap_goff = 0;
    int* d_pbc = nullptr;
    size_t cap_pbc = 0;
    // neighbour list capacity (elements)
    size_t cap_nb = 0, cap_numneigh = 0;
    // EAM
    double* d_fp = nullptr;
    size_t cap_fp = 0;
    double *d_rhor = nullptr, *d_frho = nullptr, *d_z2r = nullptr;
    int nr = 0, nrho = 0;
    double rdr = 0, rdrho = 0;
    // reductions and status
    double* d_red = nullptr;
    size_t cap_red = 0;
    double* d_ev = nullptr;
    int* d_status = nullptr;  // [0] flags, [1] max numneigh, [2] scan total
    int* h_status = nullptr;  // pinned
    double* h_red = nullptr;  // pinned
    bool ev_on = false;
    double ev_last[2]{};
};

State S;

enum { ST_BIN_RANGE = 1, ST_NB_OVERFLOW = 2 };

// ---------------------------------------------------------------- helpers
__device__ inline int coord2bin(const Grid& g, double x, double y, double z)
{
    int ix = (int)((x + g.off[0]) * g.bininv[0]);
    int iy = (int)((y + g.off[1]) * g.bininv[1]);
    int iz = (int)((z + g.off[2]) * g.bininv[2]);
    return iz * g.mbiny * g.mbinx + iy * g.mbinx + ix;
}

template <int W>
__device__ inline double group_sum(double v)
{
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, W);
    return v;
}

__device__ inline double block_sum(double v, double* sh)
{
    v                = group_sum<64>(v);
    const int lane   = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++)
            r += sh[k];
    __syncthreads();
    return r;  // valid in thread 0
}

// ---------------------------------------------------------------- integrate (integrate.c:21-42)
__global__ void k_initial_integrate(int n3, double* __restrict__ x, double* __restrict__ v,
    const double* __restrict__ f, double dt, double dtforce)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double vv = v[i] + dtforce * f[i];
    v[i]      = vv;
    x[i]      = x[i] + dt * vv;
}

__global__ void k_final_integrate(int n3, double* __restrict__ v, const double* __restrict__ f,
    double dtforce)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    v[i] = v[i] + dtforce * f[i];
}

// ---------------------------------------------------------------- pbc (pbc.c:42-232)
__global__ void k_wrap(int n3, double* __restrict__ x, double xprd, double yprd, double zprd)
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

// The 26 ghost images in the reference's order: 6 planes, 8 corners, 12 edges.  Each entry:
// required side per axis (0 none, 1 low -> shift +1, 2 high -> shift -1).
__constant__ signed char c_ghost_side[26][3] = {
    { 1, 0, 0 }, { 2, 0, 0 }, { 0, 1, 0 }, { 0, 2, 0 }, { 0, 0, 1 }, { 0, 0, 2 },
    { 1, 1, 1 }, { 1, 2, 1 }, { 1, 1, 2 }, { 1, 2, 2 }, { 2, 1, 1 }, { 2, 2, 1 }, { 2, 1, 2 },
    { 2, 2, 2 },
    { 1, 0, 1 }, { 1, 0, 2 }, { 2, 0, 1 }, { 2, 0, 2 }, { 0, 1, 1 }, { 0, 1, 2 }, { 0, 2, 1 },
    { 0, 2, 2 }, { 1, 1, 0 }, { 2, 1, 0 }, { 1, 2, 0 }, { 2, 2, 0 } };

__device__ inline unsigned ghost_mask(const double* xi, const double* prd, double cut)
{
    unsigned lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = xi[d] < cut;
        hi[d] = xi[d] >= (prd[d] - cut);
    }
    unsigned m = 0;
#pragma unroll
    for (int c = 0; c < 26; c++) {
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int s = c_ghost_side[c][d];
            if (s == 1) ok = ok && lo[d];
            if (s == 2) ok = ok && hi[d];
        }
        if (ok) m |= 1u << c;
    }
    return m;
}

struct Prd {
    double p[3];
};

__global__ void k_ghost_count(int nlocal, const double* __restrict__ x, Prd prd, double cut,
    int* __restrict__ cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlocal) return;
    double xi[3] = { x[3 * i], x[3 * i + 1], x[3 * i + 2] };
    cnt[i]       = __popc(ghost_mask(xi, prd.p, cut));
}

__global__ void k_ghost_fill(int nlocal, const double* __restrict__ x, Prd prd, double cut,
    const int* __restrict__ off, int* __restrict__ border_map, int* __restrict__ pbc,
    int* __restrict__ type)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlocal) return;
    double xi[3] = { x[3 * i], x[3 * i + 1], x[3 * i + 2] };
    unsigned m   = ghost_mask(xi, prd.p, cut);
    int g        = off[i];
    int ti       = type[i];
    for (int c = 0; c < 26; c++) {
        if (!(m & (1u << c))) continue;
        border_map[g] = i;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int s          = c_ghost_side[c][d];
            pbc[3 * g + d] = s == 1 ? 1 : (s == 2 ? -1 : 0);
        }
        type[nlocal + g] = ti;
        g++;
    }
}

__global__ void k_update_pbc(int nlocal, int nghost, double* __restrict__ x,
    const int* __restrict__ border_map, const int* __restrict__ pbc, Prd prd)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * nghost) return;
    int g               = k / 3, d = k % 3;
    int b               = border_map[g];
    x[3 * (nlocal + g) + d] = x[3 * b + d] + pbc[3 * g + d] * prd.p[d];
}

// ---------------------------------------------------------------- binning (neighbor.c:386-464)
__global__ void k_bin_count(int nall, const double* __restrict__ x, Grid g, int* __restrict__ bin_of,
    int* __restrict__ count, int* __restrict__ status)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    int b = coord2bin(g, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    if (b < 0 || b >= g.mbins) {
        atomicOr(&status[0], ST_BIN_RANGE);
        bin_of[i] = -1;
        return;
    }
    bin_of[i] = b;
    atomicAdd(&count[b], 1);
}

__global__ void k_bin_scatter(int nall, const int* __restrict__ bin_of,
    const int* __restrict__ binstart, int* __restrict__ cursor, int* __restrict__ binatoms)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    int b = bin_of[i];
    if (b < 0) return;
    int slot                  = atomicAdd(&cursor[b], 1);
    binatoms[binstart[b] + slot] = i;
}

// Atomics scatter in arbitrary order; the reference fills bins in atom-index order, so each
// bin (a handful of atoms) is insertion-sorted by index.
__global__ void k_bin_sort(int mbins, const int* __restrict__ binstart, int* __restrict__ binatoms)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= mbins) return;
    int s = binstart[b], e = binstart[b + 1];
    for (int p = s + 1; p < e; p++) {
        int v = binatoms[p];
        int q = p - 1;
        while (q >= s && binatoms[q] > v) {
            binatoms[q + 1] = binatoms[q];
            q--;
        }
        binatoms[q + 1] = v;
    }
}

// sortAtom (neighbor.c:466-536): new slot p takes the atom at position p of the bin-major list of
// the local atoms (in index order within a bin), with its position, velocity, force and type.
__global__ void k_sort_gather(int n, const int* __restrict__ order, const double* __restrict__ x,
    const double* __restrict__ v, const double* __restrict__ f, const int* __restrict__ type,
    double* __restrict__ xs, double* __restrict__ vs, double* __restrict__ fs,
    int* __restrict__ ts)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int o = order[p];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        xs[3 * p + d] = x[3 * o + d];
        vs[3 * p + d] = v[3 * o + d];
        fs[3 * p + d] = f[3 * o + d];
    }
    ts[p] = type[o];
}

// ---------------------------------------------------------------- neighbour build
// (neighbor.c:260-352): stencil bins in order, atoms in bin order, rsq <= cutneighsq.
// One wavefront per bin.  The bin's local atoms (32 at a time, bin order) sit in LDS; the
// candidates of the bin's stencil, flattened in the reference's loop order (stencil bins in order,
// atoms in bin order), are tested 64 at a time, one per lane, against every local atom of the
// bin, and a ballot gives each accepted candidate its place in that atom's list.  A lane's two
// loads per candidate are independent of the previous candidate's, and lanes of a wavefront read
// neighbouring entries of the bin arrays.
__global__ void __launch_bounds__(BLOCK) k_build_neighbor(int nlocal, const double* __restrict__ x,
    const int* __restrict__ type, int ntypes, Tables tab, int uniform, int mbins,
    const int* __restrict__ binstart, const int* __restrict__ binatoms,
    const int* __restrict__ stencil, int nstencil, int half, int maxneighs,
    int* __restrict__ neighbors, int* __restrict__ numneigh, int* __restrict__ status)
{
    __shared__ int s_st[1024];
    __shared__ double s_x[BLOCK / 64][32][3];
    __shared__ int s_id[BLOCK / 64][32], s_t[BLOCK / 64][32];
    for (int k = threadIdx.x; k < nstencil; k += blockDim.x)
        s_st[k] = stencil[k];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64) + w);
    if (b >= mbins) return;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int beg = binstart[b], end = binstart[b + 1];
    for (int q0 = beg; q0 < end; q0 += 32) {
        // this chunk's local atoms, in bin order
        const int m0 = q0 + lane;
        int i0 = -1;
        if (lane < 32 && m0 < end) i0 = binatoms[m0];
        const bool isloc = i0 >= 0 && i0 < nlocal;
        const unsigned long long lb = __ballot(isloc);
        const int nq = __popcll(lb);
        if (nq == 0) continue;
        if (isloc) {
            const int q  = __popcll(lb & lt);
            s_id[w][q]   = i0;
            s_t[w][q]    = type[i0];
            s_x[w][q][0] = x[3 * i0];
            s_x[w][q][1] = x[3 * i0 + 1];
            s_x[w][q][2] = x[3 * i0 + 2];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int cnt = 0;  // lane q < nq: length of atom q's list so far
        for (int s0 = 0; s0 < nstencil; s0 += 64) {
            const int sidx = s0 + lane;
            int lo = 0, cc = 0;
            if (sidx < nstencil) {
                const int jb = b + s_st[sidx];
                if (jb < 0 || jb >= mbins) {
                    atomicOr(&status[0], ST_BIN_RANGE);
                } else {
                    lo = binstart[jb];
                    cc = binstart[jb + 1] - lo;
                }
            }
            int inc = cc;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(inc, o, 64);
                if (lane >= o) inc += v;
            }
            const int total = __shfl(inc, 63, 64);
            const int pre   = inc - cc;
            for (int t0 = 0; t0 < total; t0 += 64) {
                const int t = t0 + lane;
                // largest column c with pre[c] <= t (cross-lane reads in every lane)
                int c = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int pc = __shfl(pre, c + step, 64);
                    if (pc <= t) c += step;
                }
                const int m     = __shfl(lo, c, 64) + (t - __shfl(pre, c, 64));
                const bool cand = t < total;
                int j = -1, tj = 0;
                double xj = 0.0, yj = 0.0, zj = 0.0;
                if (cand) {
                    j  = binatoms[m];
                    xj = x[3 * j];
                    yj = x[3 * j + 1];
                    zj = x[3 * j + 2];
                    if (!uniform) tj = type[j];
                }
                for (int q = 0; q < nq; q++) {
                    const int i     = s_id[w][q];
                    const double dx = s_x[w][q][0] - xj;
                    const double dy = s_x[w][q][1] - yj;
                    const double dz = s_x[w][q][2] - zj;
                    const double rsq = dx * dx + dy * dy + dz * dz;
                    const double cn  = uniform ? tab.cutn[0] : tab.cutn[s_t[w][q] * ntypes + tj];
                    const bool pass  = cand && j != i && !(half && j < i) && rsq <= cn;
                    const unsigned long long bq = __ballot(pass);
                    const int base = __builtin_amdgcn_readlane(cnt, q);
                    if (pass) {
                        const int pos = base + __popcll(bq & lt);
                        if (pos < maxneighs) neighbors[(size_t)pos * nlocal + i] = j;
                    }
                    if (lane == q) cnt += __popcll(bq);
                }
            }
        }
        if (lane < nq) {
            numneigh[s_id[w][lane]] = cnt;
            if (cnt > maxneighs) {
                atomicOr(&status[0], ST_NB_OVERFLOW);
                atomicMax(&status[1], cnt);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// The run loop's integration inside a full-list force kernel: with v set, the lane that writes
// an atom's total force also applies finalIntegrate (integrate.c:61-86) and the next step's
// initialIntegrate (integrate.c:29-59) with it, the same FP64 operations in the same order as
// k_final_integrate then k_initial_integrate, and writes the new position to xn.
struct Fuse {
    double* v;
    double* xn;
    double dt, dtforce;
};

// v0 (loaded at the kernel's start, so its latency hides behind the neighbour loop) and x0 are
// the atom's velocity and position.
__device__ inline void fused_integrate(const Fuse& fu, int i, const double* v0, const double* x0,
    double fx, double fy, double fz)
{
    const double fr[3] = { fx, fy, fz };
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double vv         = v0[d] + fu.dtforce * fr[d];
        vv                = vv + fu.dtforce * fr[d];
        fu.v[3 * i + d]   = vv;
        fu.xn[3 * i + d]  = x0[d] + fu.dt * vv;
    }
}

// ---------------------------------------------------------------- LJ (force_lj.c:17-206)
// One atom per LPA lanes.  STRICT: LPA == 1, reference order, no fma.
template <int LPA, bool STRICT, bool HALF, bool EV>
__global__ void __launch_bounds__(BLOCK) k_force_lj(int nlocal, const double* __restrict__ x,
    const int* __restrict__ type, int ntypes, Tables tab, int uniform,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh, double* __restrict__ f,
    double* __restrict__ ev, Fuse fu)
{
    __shared__ double s_tab[3][MAXTYPES2];
    __shared__ double s_red[16];
    const int nt2 = ntypes * ntypes;
    for (int k = threadIdx.x; k < nt2; k += blockDim.x) {
        s_tab[0][k] = tab.eps[k];
        s_tab[1][k] = tab.sig6[k];
        s_tab[2][k] = tab.cutf[k];
    }
    __syncthreads();
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int i    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = i < nlocal;
    double fix = 0, fiy = 0, fiz = 0, e = 0, w = 0;
    double v0[3] = {}, x0[3] = {};
    if (!HALF && fu.v && act && lane == 0)
        for (int d = 0; d < 3; d++)
            v0[d] = fu.v[3 * i + d];
    if (act) {
        const double xt = x[3 * i], yt = x[3 * i + 1], zt = x[3 * i + 2];
        x0[0] = xt;
        x0[1] = yt;
        x0[2] = zt;
        const int ti    = type[i];
        const int nn    = numneigh[i];
        double eps0 = s_tab[0][0], sig0 = s_tab[1][0], cut0 = s_tab[2][0];
        for (int k = lane; k < nn; k += LPA) {
            const int j = neighbors[(size_t)k * nlocal + i];
            double dx   = xt - x[3 * j];
            double dy   = yt - x[3 * j + 1];
            double dz   = zt - x[3 * j + 2];
            double rsq;
            if (STRICT) rsq = dx * dx + dy * dy + dz * dz;
            else rsq = fma(dx, dx, fma(dy, dy, dz * dz));
            double eps = eps0, sig6 = sig0, cutf = cut0;
            if (!uniform) {
                int tij = ti * ntypes + type[j];
                eps     = s_tab[0][tij];
                sig6    = s_tab[1][tij];
                cutf    = s_tab[2][tij];
            }
            if (rsq < cutf) {
                double sr2   = 1.0 / rsq;
                double sr6   = sr2 * sr2 * sr2 * sig6;
                double force = 48.0 * sr6 * (sr6 - 0.5) * sr2 * eps;
                if (STRICT) {
                    fix += dx * force;
                    fiy += dy * force;
                    fiz += dz * force;
                } else {
                    fix = fma(dx, force, fix);
                    fiy = fma(dy, force, fiy);
                    fiz = fma(dz, force, fiz);
                }
                if (HALF && j < nlocal) {
                    atomicAdd(&f[3 * j], -(dx * force));
                    atomicAdd(&f[3 * j + 1], -(dy * force));
                    atomicAdd(&f[3 * j + 2], -(dz * force));
                }
                if (EV) {
                    double wgt = (HALF && j < nlocal) ? 1.0 : 0.5;
                    e += wgt * 4.0 * eps * sr6 * (sr6 - 1.0);
                    w += wgt * rsq * force;
                }
            }
        }
    }
    if (LPA > 1) {
        fix = group_sum<LPA>(fix);
        fiy = group_sum<LPA>(fiy);
        fiz = group_sum<LPA>(fiz);
    }
    if (act && lane == 0) {
        if (HALF) {
            atomicAdd(&f[3 * i], fix);
            atomicAdd(&f[3 * i + 1], fiy);
            atomicAdd(&f[3 * i + 2], fiz);
        } else {
            f[3 * i]     = fix;
            f[3 * i + 1] = fiy;
            f[3 * i + 2] = fiz;
            if (fu.v) fused_integrate(fu, i, v0, x0, fix, fiy, fiz);
        }
    }
    if (EV) {
        double be = block_sum(e, s_red);
        double bw = block_sum(w, s_red);
        if (threadIdx.x == 0) {
            atomicAdd(&ev[0], be);
            atomicAdd(&ev[1], bw);
        }
    }
}

// ---------------------------------------------------------------- EAM (force_eam.c:19-228)
// Spline knot of a pair distance (force_eam.c:72-76, 154-158): row m and offset p in [0, 1].
__device__ inline int eam_knot(double r, double rdr, int nr, double* p)
{
    double pp = r * rdr + 1.0;
    int m     = (int)pp;
    m         = m < nr - 1 ? m : nr - 1;
    pp -= m;
    *p = pp < 1.0 ? pp : 1.0;
    return m;
}

// rho_j(r) from the value coefficients (c3..c6) of its row (force_eam.c:85-88)
__device__ inline double eam_rho(const double* c, double p)
{
    return ((c[0] * p + c[1]) * p + c[2]) * p + c[3];
}

// The pair force over r (force_eam.c:188-206) from the derivative coefficients c0..c2 of the rho
// row (cr) and the full z2r row (cz); phi for the energy
__device__ inline double eam_fpair(const double* cr, const double* cz, double p, double r,
    double fpi, double fpj, double* phi_out)
{
    double rhoip = (cr[0] * p + cr[1]) * p + cr[2];
    double z2p   = (cz[0] * p + cz[1]) * p + cz[2];
    double z2    = ((cz[3] * p + cz[4]) * p + cz[5]) * p + cz[6];
    double recip = 1.0 / r;
    double phi   = z2 * recip;
    double phip  = z2p * recip - phi * recip;
    double psip  = fpi * rhoip + fpj * rhoip + phip;
    *phi_out     = phi;
    return -psip * recip;
}

template <int LPA, bool STRICT, bool EV>
__global__ void __launch_bounds__(BLOCK) k_eam_density(int nlocal, const double* __restrict__ x,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh, double cutforcesq,
    const double* __restrict__ rs, const double* __restrict__ fs, int nr, int nrho, double rdr,
    double rdrho, double* __restrict__ fp, double* __restrict__ ev, double* __restrict__ eatom)
{
    __shared__ double s_red[16];
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int i    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = i < nlocal;
    double rhoi    = 0, e = 0;
    if (act) {
        const double xt = x[3 * i], yt = x[3 * i + 1], zt = x[3 * i + 2];
        const int nn    = numneigh[i];
        for (int k = lane; k < nn; k += LPA) {
            const int j = neighbors[(size_t)k * nlocal + i];
            double dx   = xt - x[3 * j];
            double dy   = yt - x[3 * j + 1];
            double dz   = zt - x[3 * j + 2];
            double rsq  = STRICT ? dx * dx + dy * dy + dz * dz : fma(dx, dx, fma(dy, dy, dz * dz));
            if (rsq < cutforcesq) {
                double p;
                const int m = eam_knot(sqrt(rsq), rdr, nr, &p);
                rhoi += eam_rho(rs + m * 7 + 3, p);
            }
        }
    }
    if (LPA > 1) rhoi = group_sum<LPA>(rhoi);
    if (act && lane == 0) {
        double p = 1.0 * rhoi * rdrho + 1.0;
        int m    = (int)p;
        m        = m < nrho - 1 ? m : nrho - 1;
        m        = m > 1 ? m : 1;
        p -= m;
        p               = p < 1.0 ? p : 1.0;
        const double* c = fs + m * 7;
        fp[i]           = (c[0] * p + c[1]) * p + c[2];
        if (EV) {
            e = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
            if (eatom) eatom[i] = e;
        }
    }
    if (EV) {
        double be = block_sum(e, s_red);
        if (threadIdx.x == 0) atomicAdd(&ev[0], be);
    }
}

__global__ void k_ghost_copy(int nlocal, int nghost, double* __restrict__ a,
    const int* __restrict__ border_map)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nghost) return;
    a[nlocal + g] = a[border_map[g]];
}

template <int LPA, bool STRICT, bool EV>
__global__ void __launch_bounds__(BLOCK) k_eam_force(int nlocal, const double* __restrict__ x,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh, double cutforcesq,
    const double* __restrict__ rs, const double* __restrict__ zs, int nr, double rdr,
    const double* __restrict__ fp, double* __restrict__ f, double* __restrict__ ev,
    double* __restrict__ eatom, Fuse fu)
{
    __shared__ double s_red[16];
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int i    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = i < nlocal;
    double fix = 0, fiy = 0, fiz = 0, e = 0, w = 0;
    double v0[3] = {}, x0[3] = {};
    if (fu.v && act && lane == 0)
        for (int d = 0; d < 3; d++)
            v0[d] = fu.v[3 * i + d];
    if (act) {
        const double xt = x[3 * i], yt = x[3 * i + 1], zt = x[3 * i + 2];
        x0[0] = xt;
        x0[1] = yt;
        x0[2] = zt;
        const double fpi = fp[i];
        const int nn     = numneigh[i];
        for (int k = lane; k < nn; k += LPA) {
            const int j = neighbors[(size_t)k * nlocal + i];
            double dx   = xt - x[3 * j];
            double dy   = yt - x[3 * j + 1];
            double dz   = zt - x[3 * j + 2];
            double rsq  = STRICT ? dx * dx + dy * dy + dz * dz : fma(dx, dx, fma(dy, dy, dz * dz));
            if (rsq < cutforcesq) {
                double r = sqrt(rsq);
                double p, phi;
                const int m        = eam_knot(r, rdr, nr, &p);
                const double fpair = eam_fpair(rs + m * 7, zs + m * 7, p, r, fpi, fp[j], &phi);
                if (STRICT) {
                    fix += dx * fpair;
                    fiy += dy * fpair;
                    fiz += dz * fpair;
                } else {
                    fix = fma(dx, fpair, fix);
                    fiy = fma(dy, fpair, fiy);
                    fiz = fma(dz, fpair, fiz);
                }
                if (EV) {
                    e += 0.5 * phi;
                    w += 0.5 * rsq * fpair;
                }
            }
        }
    }
    if (LPA > 1) {
        fix = group_sum<LPA>(fix);
        fiy = group_sum<LPA>(fiy);
        fiz = group_sum<LPA>(fiz);
    }
    double ea = 0.0;
    if (EV && eatom) ea = LPA > 1 ? group_sum<LPA>(e) : e;
    if (act && lane == 0) {
        f[3 * i]     = fix;
        f[3 * i + 1] = fiy;
        f[3 * i + 2] = fiz;
        if (EV && eatom) eatom[i] += ea;
        if (fu.v) fused_integrate(fu, i, v0, x0, fix, fiy, fiz);
    }
    if (EV) {
        double be = block_sum(e, s_red);
        double bw = block_sum(w, s_red);
        if (threadIdx.x == 0) {
            atomicAdd(&ev[0], be);
            atomicAdd(&ev[1], bw);
        }
    }
}

// The same two passes with the spline rows staged in LDS (no energy): the per-pair table rows
// (32 B in the density pass, 80 B in the force pass) then come from LDS, and the texture path
// serves only the neighbour gathers.  The tables take 64 KB and 160,000 B for nr <= TAB_NR, so
// the blocks are persistent: TAB_BLOCK threads, one block per 64 KB / 160 KB of a CU's LDS,
// each over a contiguous range of atoms (ranges of one XCD adjacent, as in the cluster-pair
// kernels).  Operations and their order are those of k_eam_density / k_eam_force.
constexpr int TAB_BLOCK = 1024;
constexpr int TAB_NR    = 2000;

template <int LPA, bool STRICT>
__global__ void __launch_bounds__(TAB_BLOCK) k_eam_density_tab(int nlocal, int per_block,
    const double* __restrict__ x, const int* __restrict__ neighbors,
    const int* __restrict__ numneigh, double cutforcesq, const double* __restrict__ rs,
    const double* __restrict__ fs, int nr, int nrho, double rdr, double rdrho,
    double* __restrict__ fp)
{
    __shared__ double s_rs[TAB_NR * 4];  // value coefficients c3..c6 of each rho row
    for (int k = threadIdx.x; k < nr * 4; k += TAB_BLOCK)
        s_rs[k] = rs[(k >> 2) * 7 + 3 + (k & 3)];
    __syncthreads();
    const int per   = gridDim.x >> 3;
    const int lb    = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int begin = lb * per_block;
    const int end   = min(nlocal, begin + per_block);
    const int lane  = threadIdx.x % LPA;
    for (int base = begin; base < end; base += TAB_BLOCK / LPA) {
        const int i    = base + threadIdx.x / LPA;
        const bool act = i < end;
        double rhoi    = 0;
        if (act) {
            const double xt = x[3 * i], yt = x[3 * i + 1], zt = x[3 * i + 2];
            const int nn    = numneigh[i];
            for (int k = lane; k < nn; k += LPA) {
                const int j = neighbors[(size_t)k * nlocal + i];
                double dx   = xt - x[3 * j];
                double dy   = yt - x[3 * j + 1];
                double dz   = zt - x[3 * j + 2];
                double rsq  = STRICT ? dx * dx + dy * dy + dz * dz : fma(dx, dx, fma(dy, dy, dz * dz));
                if (rsq < cutforcesq) {
                    double p;
                    const int m = eam_knot(sqrt(rsq), rdr, nr, &p);
                    rhoi += eam_rho(s_rs + m * 4, p);
                }
            }
        }
        if (LPA > 1) rhoi = group_sum<LPA>(rhoi);
        if (act && lane == 0) {
            double p = 1.0 * rhoi * rdrho + 1.0;
            int m    = (int)p;
            m        = m < nrho - 1 ? m : nrho - 1;
            m        = m > 1 ? m : 1;
            p -= m;
            p               = p < 1.0 ? p : 1.0;
            const double* c = fs + m * 7;
            fp[i]           = (c[0] * p + c[1]) * p + c[2];
        }
    }
}

template <int LPA, bool STRICT>
__global__ void __launch_bounds__(TAB_BLOCK) k_eam_force_tab(int nlocal, int per_block,
    const double* __restrict__ x, const int* __restrict__ neighbors,
    const int* __restrict__ numneigh, double cutforcesq, const double* __restrict__ rs,
    const double* __restrict__ zs, int nr, double rdr, const double* __restrict__ fp,
    double* __restrict__ f, Fuse fu)
{
    __shared__ double s_t[TAB_NR * 10];  // per row: rho c0..c2, then z2r c0..c6
    for (int k = threadIdx.x; k < nr * 10; k += TAB_BLOCK) {
        const int row = k / 10, c = k - 10 * row;
        s_t[k]        = c < 3 ? rs[row * 7 + c] : zs[row * 7 + c - 3];
    }
    __syncthreads();
    const int per   = gridDim.x >> 3;
    const int lb    = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int begin = lb * per_block;
    const int end   = min(nlocal, begin + per_block);
    const int lane  = threadIdx.x % LPA;
    for (int base = begin; base < end; base += TAB_BLOCK / LPA) {
        const int i    = base + threadIdx.x / LPA;
        const bool act = i < end;
        double fix = 0, fiy = 0, fiz = 0;
        double v0[3] = {}, x0[3] = {};
        if (fu.v && act && lane == 0)
            for (int d = 0; d < 3; d++)
                v0[d] = fu.v[3 * i + d];
        if (act) {
            const double xt = x[3 * i], yt = x[3 * i + 1], zt = x[3 * i + 2];
            x0[0]            = xt;
            x0[1]            = yt;
            x0[2]            = zt;
            const double fpi = fp[i];
            const int nn     = numneigh[i];
            for (int k = lane; k < nn; k += LPA) {
                const int j = neighbors[(size_t)k * nlocal + i];
                double dx   = xt - x[3 * j];
                double dy   = yt - x[3 * j + 1];
                double dz   = zt - x[3 * j + 2];
                double rsq  = STRICT ? dx * dx + dy * dy + dz * dz : fma(dx, dx, fma(dy, dy, dz * dz));
                if (rsq < cutforcesq) {
                    double r = sqrt(rsq);
                    double p, phi;
                    const int m        = eam_knot(r, rdr, nr, &p);
                    const double* row  = s_t + m * 10;
                    const double fpair = eam_fpair(row, row + 3, p, r, fpi, fp[j], &phi);
                    if (STRICT) {
                        fix += dx * fpair;
                        fiy += dy * fpair;
                        fiz += dz * fpair;
                    } else {
                        fix = fma(dx, fpair, fix);
                        fiy = fma(dy, fpair, fiy);
                        fiz = fma(dz, fpair, fiz);
                    }
                }
            }
        }
        if (LPA > 1) {
            fix = group_sum<LPA>(fix);
            fiy = group_sum<LPA>(fiy);
            fiz = group_sum<LPA>(fiz);
        }
        if (act && lane == 0) {
            f[3 * i]     = fix;
            f[3 * i + 1] = fiy;
            f[3 * i + 2] = fiz;
            if (fu.v) fused_integrate(fu, i, v0, x0, fix, fiy, fiz);
        }
    }
}

// ---------------------------------------------------------------- thermo (thermo.c:76-81)
__global__ void __launch_bounds__(BLOCK) k_ksum_partial(int n, const double* __restrict__ v,
    double mass, double* __restrict__ part)
{
    __shared__ double s_red[16];
    double t = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        t += (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * mass;
    double b = block_sum(t, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void k_sum_partials(int n, double* __restrict__ part)
{
    __shared__ double s_red[16];
    double t = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x)
        t += part[k];
    double b = block_sum(t, s_red);
    if (threadIdx.x == 0) part[0] = b;
}

// ---------------------------------------------------------------- host helpers
bool ensure_status()
{
    if (!S.d_status) {
        MDG_HIP(hipMalloc(&S.d_status, 16 * sizeof(int)));
        MDG_HIP(hipHostMalloc(&S.h_status, 16 * sizeof(int), hipHostMallocDefault));
        MDG_HIP(hipMalloc(&S.d_ev, 4 * sizeof(double)));
        MDG_HIP(hipHostMalloc(&S.h_red, 4 * sizeof(double), hipHostMallocDefault));
        if (failed()) return false;
        fill_i32(S.d_status, 0, 16);
    }
    return !failed();
}

void read_status()
{
    MDG_HIP(hipMemcpyAsync(S.h_status, S.d_status, 16 * sizeof(int), hipMemcpyDeviceToHost,
        stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
}

void launch_ghost_copy(double* a, Atom* atom)
{
    if (atom->Nghost <= 0) return;
    k_ghost_copy<<<blocks_for(atom->Nghost, BLOCK), BLOCK, 0, stream()>>>(atom->Nlocal,
        atom->Nghost, a, atom->d_atom.border_map);
    MDG_LAUNCH_CHECK("ghost_copy");
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void initDeviceGFX950(Parameter* param, Atom* atom, Neighbor* neighbor)
{
    if (failed() || !stream()) return;
    if (!ensure_status()) return;
    {
        int dev = 0, cu = 0;
        MDG_HIP(hipGetDevice(&dev));
        MDG_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        if (failed()) return;
        S.ncu = (cu + 7) / 8 * 8;
    }
    DeviceAtom* d   = &atom->d_atom;
    memset(d, 0, sizeof(*d));
    S.cap_all = S.cap_v = 0;
    size_t n  = (size_t)(atom->Nmax > atom->Nlocal + atom->Nghost ? atom->Nmax
                                                                    : atom->Nlocal + atom->Nghost);
    size_t c  = 0;
    grow_n(&d->x, &c, 3 * n, 0);
    c = 0;
    grow_n(&d->fx, &c, 3 * n, 0);
    c = 0;
    grow_n(&d->type, &c, n, 0);
    c = 0;
    grow_n(&d->vx, &c, 3 * (size_t)atom->Nlocal + 3, 0);
    S.cap_v   = 3 * (size_t)atom->Nlocal + 3;
    S.cap_all = n;
    if (failed()) return;
    fill_f64(d->fx, 0.0, 3 * n);

    // type-pair tables (atom.c:125-142)
    S.ntypes = atom->ntypes;
    if (S.ntypes * S.ntypes > MAXTYPES2) {
        set_error(MDG_ERR_ARG, "ntypes %d > 8 not supported", S.ntypes);
        return;
    }
    S.uniform = true;
    for (int k = 0; k < S.ntypes * S.ntypes; k++) {
        S.tab.eps[k]  = atom->epsilon[k];
        S.tab.sig6[k] = atom->sigma6[k];
        S.tab.cutf[k] = atom->cutforcesq[k];
        S.tab.cutn[k] = atom->cutneighsq[k];
        if (atom->epsilon[k] != atom->epsilon[0] || atom->sigma6[k] != atom->sigma6[0] ||
            atom->cutforcesq[k] != atom->cutforcesq[0] || atom->cutneighsq[k] != atom->cutneighsq[0])
            S.uniform = false;
    }
    size_t t2 = (size_t)S.ntypes * S.ntypes;
    MDG_HIP(hipMalloc(&d->epsilon, t2 * sizeof(double)));
    MDG_HIP(hipMalloc(&d->sigma6, t2 * sizeof(double)));
    MDG_HIP(hipMalloc(&d->cutforcesq, t2 * sizeof(double)));
    MDG_HIP(hipMalloc(&d->cutneighsq, t2 * sizeof(double)));
    if (failed()) return;
    MDG_HIP(hipMemcpyAsync(d->epsilon, atom->epsilon, t2 * 8, hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d->sigma6, atom->sigma6, t2 * 8, hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d->cutforcesq, atom->cutforcesq, t2 * 8, hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d->cutneighsq, atom->cutneighsq, t2 * 8, hipMemcpyHostToDevice, stream()));

    // neighbour list
    DeviceNeighbor* dn = &neighbor->d_neighbor;
    dn->neighbors      = nullptr;
    dn->numneigh       = nullptr;
    S.cap_nb = S.cap_numneigh = 0;
    grow_n(&dn->neighbors, &S.cap_nb, (size_t)atom->Nlocal * neighbor->maxneighs, 0);
    grow_n(&dn->numneigh, &S.cap_numneigh, (size_t)atom->Nlocal, 0);
    S.binned = false;
    S.ready  = true;
    copyDataToGFX950Device(param, atom);
}

void freeDeviceGFX950(Atom* atom, Neighbor* neighbor)
{
    if (S.d_status) (void)hipStreamSynchronize(stream());
    if (atom) {
        DeviceAtom* d = &atom->d_atom;
        dfree(d->x);
        dfree(d->vx);
        dfree(d->fx);
        dfree(d->type);
        dfree(d->border_map);
        dfree(d->epsilon);
        dfree(d->sigma6);
        dfree(d->cutforcesq);
        dfree(d->cutneighsq);
        memset(d, 0, sizeof(*d));
    }
    if (neighbor) {
        dfree(neighbor->d_neighbor.neighbors);
        dfree(neighbor->d_neighbor.numneigh);
        neighbor->d_neighbor.neighbors = nullptr;
        neighbor->d_neighbor.numneigh  = nullptr;
    }
    dfree(S.d_xn);
    dfree(S.d_sort);
    dfree(S.d_sort_t);
    dfree(S.d_stencil);
    dfree(S.d_bin_of);
    dfree(S.d_bincount);
    dfree(S.d_binstart);
    dfree(S.d_binatoms);
    dfree(S.d_eatom);
    dfree(S.d_gcnt);
    dfree(S.d_goff);
    dfree(S.d_pbc);
    dfree(S.d_fp);
    dfree(S.d_rhor);
    dfree(S.d_frho);
    dfree(S.d_z2r);
    dfree(S.d_red);
    dfree(S.d_ev);
    dfree(S.d_status);
    if (S.h_status) (void)hipHostFree(S.h_status);
    if (S.h_red) (void)hipHostFree(S.h_red);
    S = State();
}

void uploadNeighborGFX950(Atom* atom, Neighbor* neighbor)
{
    if (failed() || !S.ready) return;
    const int nlocal = atom->Nlocal, mn = neighbor->maxneighs;
    grow_n(&neighbor->d_neighbor.neighbors, &S.cap_nb, (size_t)nlocal * mn, 0);
    grow_n(&neighbor->d_neighbor.numneigh, &S.cap_numneigh, (size_t)nlocal, 0);
    if (failed()) return;
    // the host list is AoS (neighbors[i * maxneighs + k]); the device list is SoA
    int* soa = (int*)malloc((size_t)nlocal * mn * sizeof(int) + 4);
    if (!soa) {
        set_error(MDG_ERR_CAPACITY, "uploadNeighborGFX950: host buffer");
        return;
    }
    for (int i = 0; i < nlocal; i++)
        for (int k = 0; k < neighbor->numneigh[i]; k++)
            soa[(size_t)k * nlocal + i] = neighbor->neighbors[(size_t)i * mn + k];
    MDG_HIP(hipMemcpyAsync(neighbor->d_neighbor.neighbors, soa, (size_t)nlocal * mn * sizeof(int),
        hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(neighbor->d_neighbor.numneigh, neighbor->numneigh,
        (size_t)nlocal * sizeof(int), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    free(soa);
}

void copyDataToGFX950Device(Parameter* param, Atom* atom)
{
    (void)param;
    if (failed() || !S.ready) return;
    size_t nall = (size_t)atom->Nlocal + atom->Nghost;
    DeviceAtom* d = &atom->d_atom;
    MDG_HIP(hipMemcpyAsync(d->x, atom->x, 3 * nall * sizeof(double), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d->vx, atom->vx, 3 * (size_t)atom->Nlocal * sizeof(double),
        hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d->type, atom->type, nall * sizeof(int), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
}

void copyDataFromGFX950Device(Parameter* param, Atom* atom)
{
    (void)param;
    if (failed() || !S.ready) return;
    DeviceAtom* d = &atom->d_atom;
    size_t nl     = (size_t)atom->Nlocal;
    MDG_HIP(hipMemcpyAsync(atom->x, d->x, 3 * nl * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipMemcpyAsync(atom->vx, d->vx, 3 * nl * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipMemcpyAsync(atom->fx, d->fx, 3 * nl * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipMemcpyAsync(atom->type, d->type, nl * sizeof(int), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
}

// setupNeighbor (neighbor.c:101-258), restated for the host; the stencil goes to the device.
void setupNeighborGFX950(Parameter* param)
{
    if (failed()) return;
    if (!ensure_status()) return;
    const double FACTOR = 0.999;
    double cutneigh     = param->cutneigh;
    double xprd = param->nx * param->lattice, yprd = param->ny * param->lattice,
           zprd = param->nz * param->lattice;
    double bs[3], inv[3];
    int nb[3];
    if (param->input_file != NULL) {
        double lo[3] = { param->xlo, param->ylo, param->zlo };
        double hi[3] = { param->xhi, param->yhi, param->zhi };
        xprd = param->xprd;
        yprd = param->yprd;
        zprd = param->zprd;
        for (int d = 0; d < 3; d++) {
            bs[d] = cutneigh * 0.5;
            nb[d] = (int)((hi[d] - lo[d]) / bs[d]);
            if (nb[d] == 0) nb[d] = 1;
            inv[d] = nb[d] / (hi[d] - lo[d]);
        }
    } else {
        const double neighscale = 5.0 / 6.0;
        int nxyz[3]             = { param->nx, param->ny, param->nz };
        double prd[3]           = { xprd, yprd, zprd };
        for (int d = 0; d < 3; d++) {
            double t = neighscale * nxyz[d];
            nb[d]    = (int)(t > 1 ? t : 1);
            bs[d]    = prd[d] / nb[d];
            inv[d]   = 1.0 / bs[d];
        }
    }
    int pad[3], next[3];
    for (int d = 0; d < 3; d++) {
        pad[d] = (int)(cutneigh * inv[d]);
        while (pad[d] * bs[d] < FACTOR * cutneigh)
            pad[d]++;
    }
    for (int d = 0; d < 3; d++) {
        next[d] = (int)(cutneigh * inv[d]);
        while (next[d] * bs[d] < FACTOR * cutneigh) {
            next[d]++;
            pad[d]++;
        }
    }
    int mb[3];
    for (int d = 0; d < 3; d++) {
        mb[d] = nb[d] + 4 * pad[d];
        if (mb[d] < 1) mb[d] = 1;
    }
    double cutneighsq = cutneigh * cutneigh;
    int st[1024];
    int ns = 0;
    for (int k = -next[2]; k <= next[2]; k++)
        for (int j = -next[1]; j <= next[1]; j++)
            for (int i = -next[0]; i <= next[0]; i++) {
                int ijk[3] = { i, j, k };
                double dd[3];
                for (int d = 0; d < 3; d++)
                    dd[d] = ijk[d] > 0 ? (ijk[d] - 1) * bs[d] : (ijk[d] == 0 ? 0.0 : (ijk[d] + 1) * bs[d]);
                double dist = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
                if (dist < cutneighsq) {
                    if (ns >= 1024) {
                        set_error(MDG_ERR_ARG, "stencil larger than 1024 bins");
                        return;
                    }
                    st[ns++] = k * mb[1] * mb[0] + j * mb[0] + i;
                }
            }
    for (int d = 0; d < 3; d++) {
        S.binsize[d]  = bs[d];
        S.g.bininv[d] = inv[d];
        S.pad[d]      = pad[d];
        S.g.off[d]    = fabs(0.0 - pad[d] * bs[d]) + 1e-9;
    }
    S.g.mbinx  = mb[0];
    S.g.mbiny  = mb[1];
    S.g.mbinz  = mb[2];
    S.g.mbins  = mb[0] * mb[1] * mb[2];
    S.nstencil = ns;
    S.prd[0]   = param->xprd;
    S.prd[1]   = param->yprd;
    S.prd[2]   = param->zprd;
    S.cutneigh = cutneigh;
    grow_n(&S.d_stencil, &S.stencil_cap, (size_t)ns, 0);
    if (failed()) return;
    MDG_HIP(hipMemcpyAsync(S.d_stencil, st, ns * sizeof(int), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    S.binned = false;
}

void binatomsGFX950(Atom* atom)
{
    if (failed() || !S.ready || S.g.mbins == 0) {
        if (!failed()) set_error(MDG_ERR_ARG, "binatomsGFX950 before setupNeighborGFX950");
        return;
    }
    TimedRegion tr(F_BINNING);
    int nall = atom->Nlocal + atom->Nghost;
    int mb   = S.g.mbins;
    grow_n(&S.d_bin_of, &S.cap_bin_of, (size_t)nall, 0);
    grow_n(&S.d_bincount, &S.cap_bins, (size_t)mb + 1, 0);
    grow_n(&S.d_binstart, &S.cap_binstart, (size_t)mb + 1, 0);
    grow_n(&S.d_binatoms, &S.cap_binatoms, (size_t)nall + 1, 0);
    if (failed()) return;
    hipStream_t s = stream();
    fill_i32(S.d_bincount, 0, (size_t)mb + 1);
    k_bin_count<<<blocks_for(nall, BLOCK), BLOCK, 0, s>>>(nall, atom->d_atom.x, S.g, S.d_bin_of,
        S.d_bincount, S.d_status);
    exclusive_scan(S.d_bincount, S.d_binstart, mb, S.d_binstart + mb);
    fill_i32(S.d_bincount, 0, (size_t)mb);
    k_bin_scatter<<<blocks_for(nall, BLOCK), BLOCK, 0, s>>>(nall, S.d_bin_of, S.d_binstart,
        S.d_bincount, S.d_binatoms);
    k_bin_sort<<<blocks_for(mb, BLOCK), BLOCK, 0, s>>>(mb, S.d_binstart, S.d_binatoms);
    MDG_LAUNCH_CHECK("binatoms");
    S.binned = true;
}

void sortAtomGFX950(Atom* atom)
{
    // sortAtom (neighbor.c:466-536) with the caller's Nghost = 0 of main.c:92-95 and 372-381
    // done here: the local atoms are binned alone and re-stored in bin order
    if (failed() || !S.ready) return;
    const int n  = atom->Nlocal;
    atom->Nghost = 0;
    if (n <= 0) return;
    fill_i32(S.d_status, 0, 2);
    binatomsGFX950(atom);
    if (failed()) return;
    read_status();
    if (failed()) return;
    if (S.h_status[0] & ST_BIN_RANGE) {
        set_error(MDG_ERR_BIN, "sortAtomGFX950: an atom lies outside the bin grid");
        return;
    }
    grow_n(&S.d_sort, &S.cap_sort, 9 * (size_t)n, 0);
    grow_n(&S.d_sort_t, &S.cap_sort_t, (size_t)n, 0);
    if (failed()) return;
    DeviceAtom* d     = &atom->d_atom;
    hipStream_t st    = stream();
    double* xs        = S.d_sort;
    double* vs        = S.d_sort + 3 * (size_t)n;
    double* fs        = S.d_sort + 6 * (size_t)n;
    k_sort_gather<<<blocks_for(n, BLOCK), BLOCK, 0, st>>>(n, S.d_binatoms, d->x, d->vx, d->fx,
        d->type, xs, vs, fs, S.d_sort_t);
    MDG_LAUNCH_CHECK("sort_gather");
    MDG_HIP(hipMemcpyAsync(d->x, xs, 24 * (size_t)n, hipMemcpyDeviceToDevice, st));
    MDG_HIP(hipMemcpyAsync(d->vx, vs, 24 * (size_t)n, hipMemcpyDeviceToDevice, st));
    MDG_HIP(hipMemcpyAsync(d->fx, fs, 24 * (size_t)n, hipMemcpyDeviceToDevice, st));
    MDG_HIP(hipMemcpyAsync(d->type, S.d_sort_t, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
    S.binned = false;  // the bins name the old indices
}

void buildNeighborGFX950(Atom* atom, Neighbor* neighbor)
{
    if (failed() || !S.ready) return;
    binatomsGFX950(atom);
    if (failed()) return;
    const int nlocal = atom->Nlocal;
    for (int attempt = 0; attempt < 8; attempt++) {
        grow_n(&neighbor->d_neighbor.neighbors, &S.cap_nb, (size_t)nlocal * neighbor->maxneighs, 0);
        grow_n(&neighbor->d_neighbor.numneigh, &S.cap_numneigh, (size_t)nlocal, 0);
        if (failed()) return;
        fill_i32(S.d_status, 0, 2);
        {
            TimedRegion tr(F_NEIGHBOR);
            k_build_neighbor<<<blocks_for(S.g.mbins, BLOCK / 64), BLOCK, 0, stream()>>>(nlocal,
                atom->d_atom.x, atom->d_atom.type, S.ntypes, S.tab, S.uniform ? 1 : 0, S.g.mbins,
                S.d_binstart, S.d_binatoms, S.d_stencil, S.nstencil, neighbor->half_neigh,
                neighbor->maxneighs, neighbor->d_neighbor.neighbors,
                neighbor->d_neighbor.numneigh, S.d_status);
            MDG_LAUNCH_CHECK("build_neighbor");
        }
        read_status();
        if (failed()) return;
        if (S.h_status[0] & ST_BIN_RANGE) {
            set_error(MDG_ERR_BIN, "an atom lies outside the bin grid (coord2bin out of range)");
            return;
        }
        if (!(S.h_status[0] & ST_NB_OVERFLOW)) return;
        // neighbor.c:336-345: grow to 1.2x the largest count and rebuild
        neighbor->maxneighs = (int)(S.h_status[1] * 1.2) + 1;
    }
    set_error(MDG_ERR_CAPACITY, "neighbour list did not converge");
}

void setupPbcGFX950(Atom* atom, Parameter* param)
{
    if (failed() || !S.ready) return;
    TimedRegion tr(F_PBC);
    const int nlocal = atom->Nlocal;
    Prd prd{ { param->xprd, param->yprd, param->zprd } };
    grow_n(&S.d_gcnt, &S.cap_gcnt, (size_t)nlocal + 1, 0);
    grow_n(&S.d_goff, &S.cap_goff, (size_t)nlocal + 1, 0);
    if (failed()) return;
    hipStream_t s = stream();
    k_ghost_count<<<blocks_for(nlocal, BLOCK), BLOCK, 0, s>>>(nlocal, atom->d_atom.x, prd,
        param->cutneigh, S.d_gcnt);
    exclusive_scan(S.d_gcnt, S.d_goff, nlocal, S.d_status + 2);
    read_status();
    if (failed()) return;
    int nghost = S.h_status[2];
    // grow atom arrays (keep the locals), ghost map and shifts
    size_t need = (size_t)nlocal + nghost;
    if (need > S.cap_all) {
        int keep_ghost = atom->Nghost;
        atom->Nghost   = 0;  // only locals carry data worth keeping
        DeviceAtom* d  = &atom->d_atom;
        size_t c       = S.cap_all * 3;
        grow_n(&d->x, &c, 3 * need, 3 * (size_t)nlocal);
        c = S.cap_all * 3;
        grow_n(&d->fx, &c, 3 * need, 0);
        c = S.cap_all;
        grow_n(&d->type, &c, need, (size_t)nlocal);
        S.cap_all    = need;
        atom->Nghost = keep_ghost;
    }
    size_t gc = S.cap_ghost;
    grow_n(&atom->d_atom.border_map, &gc, (size_t)nghost + 1, 0);
    S.cap_ghost = gc;
    grow_n(&S.d_pbc, &S.cap_pbc, 3 * (size_t)nghost + 3, 0);
    if (failed()) return;
    k_ghost_fill<<<blocks_for(nlocal, BLOCK), BLOCK, 0, s>>>(nlocal, atom->d_atom.x, prd,
        param->cutneigh, S.d_goff, atom->d_atom.border_map, S.d_pbc, atom->d_atom.type);
    MDG_LAUNCH_CHECK("ghost_fill");
    atom->Nghost = nghost;
}

void updatePbcGFX950(Atom* atom, Parameter* param, bool doReneighbor)
{
    (void)doReneighbor;
    if (failed() || !S.ready || atom->Nghost <= 0) return;
    TimedRegion tr(F_PBC);
    Prd prd{ { param->xprd, param->yprd, param->zprd } };
    k_update_pbc<<<blocks_for(3LL * atom->Nghost, BLOCK), BLOCK, 0, stream()>>>(atom->Nlocal,
        atom->Nghost, atom->d_atom.x, atom->d_atom.border_map, S.d_pbc, prd);
    MDG_LAUNCH_CHECK("update_pbc");
}

void updateAtomsPbcGFX950(Atom* atom, Parameter* param, bool doReneighbor)
{
    (void)doReneighbor;
    if (failed() || !S.ready) return;
    TimedRegion tr(F_PBC);
    k_wrap<<<blocks_for(3LL * atom->Nlocal, BLOCK), BLOCK, 0, stream()>>>(3 * atom->Nlocal,
        atom->d_atom.x, param->xprd, param->yprd, param->zprd);
    MDG_LAUNCH_CHECK("wrap");
}

static double force_lj(Parameter* param, Atom* atom, Neighbor* neighbor, Fuse fu)
{
    (void)param;
    if (failed() || !S.ready) return 0.0;
    const double t0 = wtime();
    const int nlocal = atom->Nlocal;
    const int nall   = nlocal + atom->Nghost;
    const bool strict = strict_mode();
    const int lpa    = strict ? 1 : FAST_LPA;
    unsigned grid    = blocks_for((long long)nlocal * lpa, BLOCK);
    hipStream_t s    = stream();
    if (S.ev_on) fill_f64(S.d_ev, 0.0, 2);
    const int half = neighbor->half_neigh;
    if (half) fill_f64(atom->d_atom.fx, 0.0, 3 * (size_t)nall);
    {
        TimedRegion tr(F_FORCE_LJ);
#define MDG_LJ(L, ST, H, E)                                                                      \
    k_force_lj<L, ST, H, E><<<grid, BLOCK, 0, s>>>(nlocal, atom->d_atom.x, atom->d_atom.type,    \
        S.ntypes, S.tab, S.uniform ? 1 : 0, neighbor->d_neighbor.neighbors,                      \
        neighbor->d_neighbor.numneigh, atom->d_atom.fx, S.d_ev, fu)
        if (strict) {
            if (half) {
                if (S.ev_on) MDG_LJ(1, true, true, true);
                else MDG_LJ(1, true, true, false);
            } else {
                if (S.ev_on) MDG_LJ(1, true, false, true);
                else MDG_LJ(1, true, false, false);
            }
        } else {
            if (half) {
                if (S.ev_on) MDG_LJ(FAST_LPA, false, true, true);
                else MDG_LJ(FAST_LPA, false, true, false);
            } else {
                if (S.ev_on) MDG_LJ(FAST_LPA, false, false, true);
                else MDG_LJ(FAST_LPA, false, false, false);
            }
        }
#undef MDG_LJ
        MDG_LAUNCH_CHECK("force_lj");
    }
    return wtime() - t0;
}

double computeForceLJGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats)
{
    (void)stats;
    return force_lj(param, atom, neighbor, Fuse{ nullptr, nullptr, 0.0, 0.0 });
}

void setEamTablesGFX950(int nr, int nrho, double rdr, double rdrho, const double* rhor_spline,
    const double* frho_spline, const double* z2r_spline, int nr_tot, int nrho_tot)
{
    if (failed() || !stream()) return;
    dfree(S.d_rhor);
    dfree(S.d_frho);
    dfree(S.d_z2r);
    S.d_rhor = S.d_frho = S.d_z2r = nullptr;
    MDG_HIP(hipMalloc(&S.d_rhor, (size_t)nr_tot * 8));
    MDG_HIP(hipMalloc(&S.d_z2r, (size_t)nr_tot * 8));
    MDG_HIP(hipMalloc(&S.d_frho, (size_t)nrho_tot * 8));
    if (failed()) return;
    MDG_HIP(hipMemcpy(S.d_rhor, rhor_spline, (size_t)nr_tot * 8, hipMemcpyHostToDevice));
    MDG_HIP(hipMemcpy(S.d_z2r, z2r_spline, (size_t)nr_tot * 8, hipMemcpyHostToDevice));
    MDG_HIP(hipMemcpy(S.d_frho, frho_spline, (size_t)nrho_tot * 8, hipMemcpyHostToDevice));
    S.nr    = nr;
    S.nrho  = nrho;
    S.rdr   = rdr;
    S.rdrho = rdrho;
}

static double force_eam(Parameter* param, Atom* atom, Neighbor* neighbor, Fuse fu)
{
    if (failed() || !S.ready) return 0.0;
    if (!S.d_rhor) {
        set_error(MDG_ERR_ARG, "computeForceEamGFX950 before setEamTablesGFX950");
        return 0.0;
    }
    const double t0 = wtime();
    const int nlocal = atom->Nlocal;
    const int nall   = nlocal + atom->Nghost;
    grow_n(&S.d_fp, &S.cap_fp, (size_t)nall + 1, 0);
    if (failed()) return 0.0;
    const bool strict = strict_mode();
    const int lpa     = strict ? 1 : FAST_LPA;
    unsigned grid     = blocks_for((long long)nlocal * lpa, BLOCK);
    hipStream_t s     = stream();
    double cutf       = param->cutforce * param->cutforce;
    double* eatom = nullptr;
    if (S.eatom_on) {
        grow_n(&S.d_eatom, &S.cap_eatom, (size_t)nlocal + 1, 0);
        if (failed()) return 0.0;
        eatom = S.d_eatom;
    }
    const bool ev = S.ev_on || S.eatom_on;
    if (ev) fill_f64(S.d_ev, 0.0, 2);
    TimedRegion tr(F_FORCE_EAM);
#define MDG_EAM(L, ST, E)                                                                        \
    do {                                                                                         \
        k_eam_density<L, ST, E><<<grid, BLOCK, 0, s>>>(nlocal, atom->d_atom.x,                   \
            neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh, cutf, S.d_rhor,       \
            S.d_frho, S.nr, S.nrho, S.rdr, S.rdrho, S.d_fp, S.d_ev, eatom);                      \
        launch_ghost_copy(S.d_fp, atom);                                                         \
        k_eam_force<L, ST, E><<<grid, BLOCK, 0, s>>>(nlocal, atom->d_atom.x,                     \
            neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh, cutf, S.d_rhor,       \
            S.d_z2r, S.nr, S.rdr, S.d_fp, atom->d_atom.fx, S.d_ev, eatom, fu);                   \
    } while (0)
    // without the energy and for tables of up to TAB_NR rows: the LDS-table passes
#define MDG_EAM_TAB(L, ST)                                                                       \
    do {                                                                                         \
        const unsigned gd = 2u * (unsigned)S.ncu, gf = (unsigned)S.ncu;                          \
        const int pd = (int)((nlocal + gd - 1) / gd), pf = (int)((nlocal + gf - 1) / gf);        \
        k_eam_density_tab<L, ST><<<gd, TAB_BLOCK, 0, s>>>(nlocal, pd, atom->d_atom.x,            \
            neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh, cutf, S.d_rhor,       \
            S.d_frho, S.nr, S.nrho, S.rdr, S.rdrho, S.d_fp);                                     \
        launch_ghost_copy(S.d_fp, atom);                                                         \
        k_eam_force_tab<L, ST><<<gf, TAB_BLOCK, 0, s>>>(nlocal, pf, atom->d_atom.x,              \
            neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh, cutf, S.d_rhor,       \
            S.d_z2r, S.nr, S.rdr, S.d_fp, atom->d_atom.fx, fu);                                  \
    } while (0)
    const bool tab = !ev && S.nr <= TAB_NR && S.ncu > 0 && nlocal > 0;
    if (strict) {
        if (ev) MDG_EAM(1, true, true);
        else if (tab) MDG_EAM_TAB(1, true);
        else MDG_EAM(1, true, false);
    } else {
        if (ev) MDG_EAM(FAST_LPA, false, true);
        else if (tab) MDG_EAM_TAB(FAST_LPA, false);
        else MDG_EAM(FAST_LPA, false, false);
    }
#undef MDG_EAM_TAB
#undef MDG_EAM
    MDG_LAUNCH_CHECK("force_eam");
    return wtime() - t0;
}

double computeForceEamGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats)
{
    (void)stats;
    return force_eam(param, atom, neighbor, Fuse{ nullptr, nullptr, 0.0, 0.0 });
}

double computeForceFinalInitialGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats)
{
    if (failed() || !S.ready) return 0.0;
    const bool eam = param->force_field == FF_EAM;
    if (strict_mode() || (!eam && neighbor->half_neigh)) {
        // the strict kernels and half lists (atomic forces) run the three calls in order
        const double t = eam ? computeForceEamGFX950(param, atom, neighbor, stats)
                             : computeForceLJGFX950(param, atom, neighbor, stats);
        finalIntegrateGFX950(false, param, atom);
        initialIntegrateGFX950(false, param, atom);
        return t;
    }
    grow_n(&S.d_xn, &S.cap_xn, 3 * S.cap_all, 0);
    if (failed()) return 0.0;
    const Fuse fu{ atom->d_atom.vx, S.d_xn, param->dt, param->dtforce };
    const double t = eam ? force_eam(param, atom, neighbor, fu) : force_lj(param, atom, neighbor, fu);
    // the new local positions become the current ones; ghost positions follow at updatePbc
    std::swap(atom->d_atom.x, S.d_xn);
    S.cap_xn = 3 * S.cap_all;
    return t;
}

void requestEamAtomEnergyGFX950(int on)
{
    S.eatom_on = on != 0;
}

int downloadEamAtomEnergyGFX950(Atom* atom, double* out)
{
    if (failed() || !S.ready) return -1;
    if (!S.eatom_on || !S.d_eatom) {
        set_error(MDG_ERR_ARG, "per-atom EAM energies were not requested before the force call");
        return -1;
    }
    MDG_HIP(hipMemcpyAsync(out, S.d_eatom, (size_t)atom->Nlocal * 8, hipMemcpyDeviceToHost,
        stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? -1 : 0;
}

void initialIntegrateGFX950(bool reneigh, Parameter* param, Atom* atom)
{
    (void)reneigh;
    if (failed() || !S.ready) return;
    TimedRegion tr(F_INTEGRATE);
    int n3 = 3 * atom->Nlocal;
    k_initial_integrate<<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, atom->d_atom.x,
        atom->d_atom.vx, atom->d_atom.fx, param->dt, param->dtforce);
    MDG_LAUNCH_CHECK("initial_integrate");
}

void finalIntegrateGFX950(bool reneigh, Parameter* param, Atom* atom)
{
    (void)reneigh;
    if (failed() || !S.ready) return;
    TimedRegion tr(F_INTEGRATE);
    int n3 = 3 * atom->Nlocal;
    k_final_integrate<<<blocks_for(n3, BLOCK), BLOCK, 0, stream()>>>(n3, atom->d_atom.vx,
        atom->d_atom.fx, param->dtforce);
    MDG_LAUNCH_CHECK("final_integrate");
}

double kineticSumGFX950(Parameter* param, Atom* atom)
{
    if (failed() || !S.ready) return 0.0;
    TimedRegion tr(F_THERMO);
    const int nb = 512;
    grow_n(&S.d_red, &S.cap_red, (size_t)nb, 0);
    if (failed()) return 0.0;
    k_ksum_partial<<<nb, BLOCK, 0, stream()>>>(atom->Nlocal, atom->d_atom.vx, param->mass, S.d_red);
    k_sum_partials<<<1, BLOCK, 0, stream()>>>(nb, S.d_red);
    MDG_LAUNCH_CHECK("ksum");
    MDG_HIP(hipMemcpyAsync(S.h_red, S.d_red, sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? 0.0 : S.h_red[0];
}

void requestEnergyVirialGFX950(int on)
{
    if (!ensure_status()) return;
    S.ev_on = on != 0;
}

void energyVirialGFX950(double* epot, double* virial)
{
    double h[2] = { 0, 0 };
    if (!failed() && S.d_ev) {
        MDG_HIP(hipMemcpyAsync(h, S.d_ev, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
        MDG_HIP(hipStreamSynchronize(stream()));
    }
    if (epot) *epot = h[0];
    if (virial) *virial = h[1];
}

int downloadVLStateGFX950(Atom* atom, Neighbor* neighbor, double* x, double* f, double* v,
    int* border_map, int* pbc, int* numneigh, int* neighbors_csr, long long csr_cap)
{
    if (failed() || !S.ready) return -1;
    hipStream_t s = stream();
    size_t nl = atom->Nlocal, nall = nl + atom->Nghost;
    if (x) MDG_HIP(hipMemcpyAsync(x, atom->d_atom.x, 3 * nall * 8, hipMemcpyDeviceToHost, s));
    if (f) MDG_HIP(hipMemcpyAsync(f, atom->d_atom.fx, 3 * nl * 8, hipMemcpyDeviceToHost, s));
    if (v) MDG_HIP(hipMemcpyAsync(v, atom->d_atom.vx, 3 * nl * 8, hipMemcpyDeviceToHost, s));
    if (border_map && atom->Nghost)
        MDG_HIP(hipMemcpyAsync(border_map, atom->d_atom.border_map, atom->Nghost * 4,
            hipMemcpyDeviceToHost, s));
    if (pbc && atom->Nghost)
        MDG_HIP(hipMemcpyAsync(pbc, S.d_pbc, 3 * (size_t)atom->Nghost * 4, hipMemcpyDeviceToHost, s));
    int rc = 0;
    if (numneigh && neighbors_csr) {
        int* nn = (int*)malloc(nl * sizeof(int));
        size_t tot = nl * (size_t)neighbor->maxneighs;
        int* soa   = (int*)malloc(tot * sizeof(int) + 4);
        MDG_HIP(hipMemcpyAsync(nn, neighbor->d_neighbor.numneigh, nl * 4, hipMemcpyDeviceToHost, s));
        MDG_HIP(hipMemcpyAsync(soa, neighbor->d_neighbor.neighbors, tot * 4, hipMemcpyDeviceToHost, s));
        MDG_HIP(hipStreamSynchronize(s));
        long long k = 0;
        for (size_t i = 0; i < nl && rc == 0; i++) {
            numneigh[i] = nn[i];
            for (int m = 0; m < nn[i]; m++) {
                if (k >= csr_cap) {
                    rc = -2;
                    break;
                }
                neighbors_csr[k++] = soa[(size_t)m * nl + i];
            }
        }
        free(nn);
        free(soa);
        if (rc == 0) rc = (int)(k > 0x7fffffff ? -3 : k);
    }
    MDG_HIP(hipStreamSynchronize(s));
    return failed() ? -1 : rc;
}

int downloadBinsGFX950(int* binstart, int mbins_cap, int* binatoms, int nall_cap)
{
    if (failed() || !S.binned) return -1;
    int mb = S.g.mbins;
    if (mbins_cap < mb + 1) return -2;
    MDG_HIP(hipMemcpy(binstart, S.d_binstart, (mb + 1) * sizeof(int), hipMemcpyDeviceToHost));
    int tot = binstart[mb];
    if (tot > nall_cap) return -3;
    MDG_HIP(hipMemcpy(binatoms, S.d_binatoms, tot * sizeof(int), hipMemcpyDeviceToHost));
    return failed() ? -1 : mb;
}

}  // extern "C"
