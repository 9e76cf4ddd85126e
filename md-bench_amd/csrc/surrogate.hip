// Learned EAM surrogate on gfx950.
//
// Per MD step (all local atoms, Verlet list of the EAM path):
//   k_sur_desc    LPA lanes per atom walk the list; per pair inside rc = 4.95 A the 24 Gaussians
//                 come from a multiplicative recurrence started at the middle centre (three
//                 exponentials per pair), and the four angular channels accumulate the moments
//                 A_c = sum w_c u and T_c = sum w_c u u^T.  Writes the normalised descriptors
//                 (24 radial, |A_c|^2, 1.5 |T~_c|^2) and the moments (A, traceless T~).
//   k_sur_mlp     one wavefront per 16 atoms: the MLP 32 -> 64 -> 64 -> 1 and its input
//                 gradient as 24 v_mfma_f32_16x16x32_bf16 (weights as the A operand, atoms as the
//                 16 columns).  A layer's accumulator layout is the next layer's B operand once
//                 the hidden features are numbered to match, so activations never leave
//                 registers.  The epilogue turns dE/dD into 60 force coefficients per atom.
//   k_sur_force   LPA lanes per atom, one pair per lane: F_k = sum_j (dE_k/dx_j - dE_j/dx_k),
//                 the j side from the neighbour's coefficient row (a ghost reads its owner's
//                 row), the atom's own row staged in LDS.
// Training adds the force adjoint (k_sur_adjoint), the exact FP32 weight gradient of the loss
// (k_sur_wgrad: forward mode along the adjoint, then reverse mode, block partials reduced in a
// fixed order) and Adam (k_sur_adam).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "mdgfx950.h"
#include "mdgfx950_surrogate.h"
#include "runtime.h"

using namespace mdg;

namespace {

constexpr int BLOCK = 256;
constexpr int LPA   = 4;  // lanes per atom in the list passes
constexpr int ND = 32, NR = 24, NA = 4, NMOM = 36, NCOEF = 64, HID = 64;
constexpr int NPARAM = MDG_SUR_NPARAM;
// theta offsets
constexpr int O_W1 = 0, O_B1 = O_W1 + HID * ND, O_W2 = O_B1 + HID, O_B2 = O_W2 + HID * HID,
              O_W3 = O_B2 + HID, O_B3 = O_W3 + HID;
static_assert(O_B3 + 1 == NPARAM, "parameter count");
// descriptor constants (the oracle's)
constexpr float RC = 4.95f, RC2 = RC * RC;
constexpr float R_MU0 = 1.9f, R_DMU = 0.13f;
constexpr float R_ETA = 19.53125f;        // 1 / (2 * 0.16^2)
constexpr float A_MU0 = 2.5f, A_DMU = 0.6f;
constexpr float A_ETA = 4.0816326530612246f;  // 1 / (2 * 0.35^2)
constexpr int RMID = 12;
constexpr int NFRAG = 24;  // MFMA A-operand fragments: 4 (L1) + 8 (L2) + 8 (W2^T) + 4 (W1^T)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct SurModel {
    float* theta  = nullptr;  // device, NPARAM
    bf16x8* frag  = nullptr;  // device, NFRAG * 64
    float* norm   = nullptr;  // device, mean[32], inv_std[32]
    float e0 = 0.f, s_e = 1.f;
    float r_core = 0.f, k_core = 0.f;  // short-range core U(r) = k (r_core - r)^3, r < r_core
    bool ready = false;
    // per-atom work arrays of the drop-in force call
    float *dn = nullptr, *mom = nullptr;
    __half* coef = nullptr;
    int *pairs = nullptr, *npairs = nullptr;
    size_t cap_dn = 0, cap_mom = 0, cap_coef = 0, cap_pairs = 0, cap_npairs = 0;
    double* fsave = nullptr;
    size_t cap_fsave = 0;
    // scratch of the weight gradient (block partials) and of reductions
    float* part     = nullptr;
    size_t cap_part = 0;
    double* red     = nullptr;
    size_t cap_red  = 0;
    double* h_red   = nullptr;
};
SurModel M;

// ---------------------------------------------------------------- pair geometry
// The LPA = 4 lanes of an atom walk its list together, one pair at a time, and split the
// features: lane q owns radial descriptors 6 q .. 6 q + 5 and angular channel q.  Per-atom data
// is laid out in matching lane blocks, so the four lanes read one contiguous row:
//   coefficients (64 halves): block q = [c_{6q} .. c_{6q+5}, a_q (3), M_q (6: xx yy zz xy xz yz), 0]
//   moments      (36 floats): block q = [A_q (3), T~_q (6)]
constexpr int RPL = NR / LPA;  // radial descriptors per lane (6)

// exp(-eta (r - mu_n)^2) for the lane's six centres by a multiplicative recurrence from the
// lane's third centre: the ratio of neighbours is exp(+-2 eta D (r - mu_n) - eta D^2) and changes
// by exp(-2 eta D^2) per step.  Three exponentials give six values; a value can underflow only
// when it is below 1e-20 and negligible.
__device__ inline void gaussians6(float r, int q, float* g)
{
    const float t   = r - (R_MU0 + (RPL * q + 3) * R_DMU);
    const float e2  = R_ETA * R_DMU * R_DMU;
    const float c   = __expf(-2.0f * e2);
    float up        = __expf(2.0f * R_ETA * R_DMU * t - e2);
    float dn        = __expf(-2.0f * R_ETA * R_DMU * t - e2);
    g[3]            = __expf(-R_ETA * t * t);
    g[4]            = g[3] * up;
    up *= c;
    g[5]            = g[4] * up;
    g[2]            = g[3] * dn;
    dn *= c;
    g[1]            = g[2] * dn;
    dn *= c;
    g[0]            = g[1] * dn;
}

__device__ inline void cutoff_fn(float r2, float r, float* fc, float* dfc)
{
    const float t = 1.0f - r2 * (1.0f / RC2);
    *fc           = t * t * t;
    *dfc          = -6.0f * r * (1.0f / RC2) * t * t;
}

__device__ inline float rcp_refined(float x)
{
    float y = __builtin_amdgcn_rcpf(x);
    return fmaf(y, fmaf(-x, y, 1.0f), y);
}

__device__ inline int coef_row(int j, int nlocal, const int* __restrict__ border_map)
{
    return j < nlocal ? j : border_map[j - nlocal];
}

template <int L>
__device__ inline float gsum(float v)
{
#pragma unroll
    for (int o = 1; o < L; o <<= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Geometry of one pair, evaluated by each of the atom's lanes.
struct Pair {
    float r, ir, ux, uy, uz, fc, dfc;
};

__device__ inline bool pair_geom(const double* __restrict__ x, double xi, double yi, double zi,
    int j, Pair* p)
{
    const float dx = (float)(x[3 * j] - xi);
    const float dy = (float)(x[3 * j + 1] - yi);
    const float dz = (float)(x[3 * j + 2] - zi);
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < RC2)) return false;
    p->r  = __builtin_sqrtf(r2);
    p->ir = rcp_refined(p->r);
    p->ux = dx * p->ir;
    p->uy = dy * p->ir;
    p->uz = dz * p->ir;
    cutoff_fn(r2, p->r, &p->fc, &p->dfc);
    return true;
}

// ---------------------------------------------------------------- descriptors
// Here the atom's lanes split the pairs (lane q takes list entries q, q + 4, ...) and each
// accumulates all 60 moments; the lanes then add theirs.  The pairs inside rc are also written,
// in list order, to a compact list (SoA, cl[k * n + a]) for the force and adjoint passes.
__device__ inline void gaussians24(float r, float* g)
{
    const float t  = r - (R_MU0 + RMID * R_DMU);
    const float e2 = R_ETA * R_DMU * R_DMU;
    const float c  = __expf(-2.0f * e2);
    float up       = __expf(2.0f * R_ETA * R_DMU * t - e2);
    float dn       = __expf(-2.0f * R_ETA * R_DMU * t - e2);
    g[RMID]        = __expf(-R_ETA * t * t);
#pragma unroll
    for (int m = RMID + 1; m < NR; m++) {
        g[m] = g[m - 1] * up;
        up *= c;
    }
#pragma unroll
    for (int m = RMID - 1; m >= 0; m--) {
        g[m] = g[m + 1] * dn;
        dn *= c;
    }
}

__global__ void __launch_bounds__(BLOCK) k_sur_desc(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ neighbors,
    const int* __restrict__ numneigh, const float* __restrict__ norm, float* __restrict__ dn_out,
    float* __restrict__ mom_out, int cap, int* __restrict__ cl, int* __restrict__ ncl)
{
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int q    = gid % LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
    const int lane = threadIdx.x & 63;
    const int gbase = lane & ~(LPA - 1);
    float G[NR], A[NA][3], T[NA][6];
#pragma unroll
    for (int k = 0; k < NR; k++)
        G[k] = 0.f;
#pragma unroll
    for (int c = 0; c < NA; c++) {
        A[c][0] = A[c][1] = A[c][2] = 0.f;
#pragma unroll
        for (int k = 0; k < 6; k++)
            T[c][k] = 0.f;
    }
    int nc = 0;  // compact-list length (the same in the atom's four lanes)
    const int nn = act ? numneigh[i] : 0;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    // the wavefront walks in step so that the ballot below sees every lane
    int nmax = nn;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
        nmax = max(nmax, __shfl_xor(nmax, o, 64));
    for (int k0 = 0; k0 < nmax; k0 += LPA) {
        const int k = k0 + q;
        Pair p;
        bool in = false;
        int j   = 0;
        if (k < nn) {
            j  = neighbors[(size_t)k * nlocal + i];
            in = pair_geom(x, xi, yi, zi, j, &p);
        }
        const unsigned long long bal = __ballot(in);
        const unsigned grp = (unsigned)(bal >> gbase) & ((1u << LPA) - 1u);
        if (in) {
            const int slot = nc + __popc(grp & ((1u << q) - 1u));
            if (slot < cap) cl[(size_t)slot * n + a] = j;
            float g[NR];
            gaussians24(p.r, g);
#pragma unroll
            for (int m = 0; m < NR; m++)
                G[m] = fmaf(g[m], p.fc, G[m]);
            const float uu[6] = { p.ux * p.ux, p.uy * p.uy, p.uz * p.uz, p.ux * p.uy,
                p.ux * p.uz, p.uy * p.uz };
#pragma unroll
            for (int c = 0; c < NA; c++) {
                const float dr = p.r - (A_MU0 + c * A_DMU);
                const float w  = __expf(-A_ETA * dr * dr) * p.fc;
                A[c][0] = fmaf(w, p.ux, A[c][0]);
                A[c][1] = fmaf(w, p.uy, A[c][1]);
                A[c][2] = fmaf(w, p.uz, A[c][2]);
#pragma unroll
                for (int m = 0; m < 6; m++)
                    T[c][m] = fmaf(w, uu[m], T[c][m]);
            }
        }
        nc += __popc(grp);
    }
#pragma unroll
    for (int k = 0; k < NR; k++)
        G[k] = gsum<LPA>(G[k]);
#pragma unroll
    for (int c = 0; c < NA; c++) {
#pragma unroll
        for (int m = 0; m < 3; m++)
            A[c][m] = gsum<LPA>(A[c][m]);
#pragma unroll
        for (int m = 0; m < 6; m++)
            T[c][m] = gsum<LPA>(T[c][m]);
    }
    if (!act || q != 0) return;
    ncl[a] = nc;
    float D[ND];
#pragma unroll
    for (int k = 0; k < NR; k++)
        D[k] = G[k];
    float* mo = mom_out + (size_t)a * NMOM;
#pragma unroll
    for (int c = 0; c < NA; c++) {
        D[NR + c] = A[c][0] * A[c][0] + A[c][1] * A[c][1] + A[c][2] * A[c][2];
        const float tr = (T[c][0] + T[c][1] + T[c][2]) * (1.0f / 3.0f);
        T[c][0] -= tr;
        T[c][1] -= tr;
        T[c][2] -= tr;
        D[NR + NA + c] = 1.5f * (T[c][0] * T[c][0] + T[c][1] * T[c][1] + T[c][2] * T[c][2] +
                                    2.0f * (T[c][3] * T[c][3] + T[c][4] * T[c][4] +
                                               T[c][5] * T[c][5]));
        mo[9 * c]     = A[c][0];
        mo[9 * c + 1] = A[c][1];
        mo[9 * c + 2] = A[c][2];
#pragma unroll
        for (int m = 0; m < 6; m++)
            mo[9 * c + 3 + m] = T[c][m];
    }
    float* o = dn_out + (size_t)a * ND;
#pragma unroll
    for (int k = 0; k < ND; k += 4) {
        float4 v;
        v.x = (D[k] - norm[k]) * norm[ND + k];
        v.y = (D[k + 1] - norm[k + 1]) * norm[ND + k + 1];
        v.z = (D[k + 2] - norm[k + 2]) * norm[ND + k + 2];
        v.w = (D[k + 3] - norm[k + 3]) * norm[ND + k + 3];
        *reinterpret_cast<float4*>(o + k) = v;
    }
}

// ---------------------------------------------------------------- MFMA fragments
// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds A[m = l % 16][k = 8 (l / 16)
// + i] and B[k = 8 (l / 16) + i][n = l % 16] for i = 0..7; the result D[m = 4 (l / 16) + i]
// [n = l % 16] for i = 0..3.  The network computes H^T = W X^T with 16 atoms as the columns, so
// a 16-row output tile t leaves features 16 t + 4 g + i (g = l / 16) in lane l.  The K index of
// the next product is numbered so that K-step s takes lane values i < 4 from tile 2 s and i >= 4
// from tile 2 s + 1: k = 8 g + i  <->  feature phi_s(g, i) below.
__device__ __host__ inline int phi(int s, int g, int i)
{
    return i < 4 ? 32 * s + 4 * g + i : 32 * s + 16 + 4 * g + (i - 4);
}

// frag[f * 64 + lane]: f = t (layer 1, W1 rows 16t..), 4 + 2 t + s (layer 2, W2), 12 + 2 t + s
// (W2^T for the input gradient), 20 + 2 t + s (W1^T).
__global__ void k_sur_pack(const float* __restrict__ th, bf16x8* __restrict__ frag)
{
    const int f = blockIdx.x, l = threadIdx.x;
    const int m = l & 15, g = l >> 4;
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        float w;
        if (f < 4) {
            w = th[O_W1 + (16 * f + m) * ND + 8 * g + i];
        } else if (f < 12) {
            const int t = (f - 4) >> 1, s = (f - 4) & 1;
            w = th[O_W2 + (16 * t + m) * HID + phi(s, g, i)];
        } else if (f < 20) {
            const int t = (f - 12) >> 1, s = (f - 12) & 1;
            w = th[O_W2 + phi(s, g, i) * HID + 16 * t + m];
        } else {
            const int t = (f - 20) >> 1, s = (f - 20) & 1;
            w = th[O_W1 + phi(s, g, i) * ND + 16 * t + m];
        }
        v[i] = (__bf16)w;
    }
    frag[f * 64 + l] = v;
}

// tanh(x) = 1 - 2 / (1 + exp(2x)): one exponential and one reciprocal (1 ulp); saturates at
// both ends (exp overflow: rcp(inf) = 0 gives 1; underflow gives -1)
__device__ inline float tanh_fast(float x)
{
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * x));
}

__device__ inline bf16x8 pack8(const float* a, const float* b)
{
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i]     = (__bf16)a[i];
        v[i + 4] = (__bf16)b[i];
    }
    return v;
}

__device__ inline floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- network
__global__ void __launch_bounds__(BLOCK) k_sur_mlp(int n, const float* __restrict__ dn,
    const float* __restrict__ mom, const bf16x8* __restrict__ frag, const float* __restrict__ th,
    const float* __restrict__ norm, float s_e, float e0, float* __restrict__ e_out,
    __half* __restrict__ coef)
{
    __shared__ bf16x8 s_frag[NFRAG * 64];
    __shared__ __align__(16) __half s_co[BLOCK / 64][16 * NCOEF];  // a wavefront's 16 rows
    for (int k = threadIdx.x; k < NFRAG * 64; k += blockDim.x)
        s_frag[k] = frag[k];
    __syncthreads();
    const int l = threadIdx.x & 63, m = l & 15, g = l >> 4;
    __half* wco = s_co[threadIdx.x >> 6];
    // per-lane constants: biases and output weights of the lane's features 16 t + 4 g + i
    float b1[4][4], b2[4][4], w3[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            b1[t][i] = th[O_B1 + 16 * t + 4 * g + i];
            b2[t][i] = th[O_B2 + 16 * t + 4 * g + i];
            w3[t][i] = th[O_W3 + 16 * t + 4 * g + i];
        }
    const float b3 = th[O_B3];
    float istd[2][4];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            istd[t][i] = norm[ND + 16 * t + 4 * g + i];
    const int ntile  = (n + 15) >> 4;
    const int wave   = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (BLOCK / 64);
    for (int tile = wave; tile < ntile; tile += nwaves) {
        const int atom = tile * 16 + m;
        const bool act = atom < n;
        const int ar   = act ? atom : n - 1;
        // layer 1: B operand = descriptors 8 g .. 8 g + 7 of this lane's atom
        const float4 x0 = *reinterpret_cast<const float4*>(dn + (size_t)ar * ND + 8 * g);
        const float4 x1 = *reinterpret_cast<const float4*>(dn + (size_t)ar * ND + 8 * g + 4);
        bf16x8 X;
        X[0] = (__bf16)x0.x; X[1] = (__bf16)x0.y; X[2] = (__bf16)x0.z; X[3] = (__bf16)x0.w;
        X[4] = (__bf16)x1.x; X[5] = (__bf16)x1.y; X[6] = (__bf16)x1.z; X[7] = (__bf16)x1.w;
        float h1[4][4], h2[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = mfma(s_frag[t * 64 + l], X, acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                h1[t][i] = tanh_fast(acc[i] + b1[t][i]);
        }
        // layer 2
        const bf16x8 B0 = pack8(h1[0], h1[1]), B1 = pack8(h1[2], h1[3]);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = mfma(s_frag[(4 + 2 * t) * 64 + l], B0, acc);
            acc = mfma(s_frag[(5 + 2 * t) * 64 + l], B1, acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                h2[t][i] = tanh_fast(acc[i] + b2[t][i]);
        }
        // energy: w3 . h2 over the 64 features, summed over the four lane groups
        float part = 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                part = fmaf(w3[t][i], h2[t][i], part);
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (e_out && act && g == 0) e_out[atom] = s_e * (part + b3) + e0;
        // input gradient: d2 = (1 - h2^2) w3, gamma = W2^T d2, d1 = (1 - h1^2) gamma, g~ = W1^T d1
        float d2[4][4], d1[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                d2[t][i] = (1.0f - h2[t][i] * h2[t][i]) * w3[t][i];
        const bf16x8 C0 = pack8(d2[0], d2[1]), C1 = pack8(d2[2], d2[3]);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = mfma(s_frag[(12 + 2 * t) * 64 + l], C0, acc);
            acc = mfma(s_frag[(13 + 2 * t) * 64 + l], C1, acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                d1[t][i] = (1.0f - h1[t][i] * h1[t][i]) * acc[i];
        }
        const bf16x8 E0 = pack8(d1[0], d1[1]), E1 = pack8(d1[2], d1[3]);
        float gd[2][4];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = mfma(s_frag[(20 + 2 * t) * 64 + l], E0, acc);
            acc = mfma(s_frag[(21 + 2 * t) * 64 + l], E1, acc);
#pragma unroll
            for (int i = 0; i < 4; i++)
                gd[t][i] = s_e * acc[i] * istd[t][i];  // dE/dD_d, d = 16 t + 4 g + i
        }
        // force coefficients in lane blocks (see the layout above), staged in LDS and stored as
        // whole rows: radial d -> block d / 6, slot d % 6; channel c: 2 g_q1 A_c at 16 c + 6,
        // 3 g_q2 T~_c at 16 c + 9, zero at 16 c + 15.  FP16: the force pass gathers a
        // neighbour's row per pair, and half the bytes is half its time; 11 significant bits keep
        // the pair terms to 5e-4.
        __half* co      = wco + m * NCOEF;
        const float* mo = mom + (size_t)ar * NMOM;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int d = 4 * g + i;
            co[16 * (d / RPL) + d % RPL] = __float2half(gd[0][i]);
        }
        if (g < 2) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int d = 16 + 4 * g + i;
                co[16 * (d / RPL) + d % RPL] = __float2half(gd[1][i]);
            }
        } else if (g == 2) {
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int k = 0; k < 3; k++)
                    co[16 * c + 6 + k] = __float2half(2.0f * gd[1][c] * mo[9 * c + k]);
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int k = 0; k < 6; k++)
                    co[16 * c + 9 + k] = __float2half(3.0f * gd[1][c] * mo[9 * c + 3 + k]);
                co[16 * c + 15] = __float2half(0.f);
            }
        }
        // the wavefront's own LDS writes are complete before its reads (in-order LDS queue;
        // the fence keeps the compiler from moving them)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // 16 rows x 128 B = 128 chunks of 16 B: two per lane
        const uint4* src = reinterpret_cast<const uint4*>(wco);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int ch  = l + 64 * h;
            const int row = ch >> 3;
            if (tile * 16 + row < n)
                reinterpret_cast<uint4*>(coef + (size_t)(tile * 16 + row) * NCOEF)[ch & 7] = src[ch];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------- forces
// F_i = sum_j (dE_i/dx_j - dE_j/dx_i).  With u = (x_j - x_i) / r, the radial parts combine to
// (S_i + S_j) u with S = sum_n c_n G_n'(r); an angular channel of atom i contributes
// dE_i/dx_j = w'(u.a + u.M u) u + (w/r)(a - (u.a) u) + (2 w/r)(M u - (u.M u) u), and atom j's
// the same with -u, so the difference needs only da = a_i - a_j and sM = M_i + M_j:
//   w'(u.da + u.sM u) u + (w/r)(da - (u.da) u) + (2 w/r)(sM u - (u.sM u) u).
// 16 halves of a coefficient lane block as floats (two 16-byte loads)
__device__ inline void load_block(const __half* __restrict__ p, float* c)
{
    const uint4* v = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint4 w     = v[h];
        const unsigned u[4] = { w.x, w.y, w.z, w.w };
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const __half2 hh = *reinterpret_cast<const __half2*>(&u[k]);
            const float2 ff  = __half22float2(hh);
            c[8 * h + 2 * k]     = ff.x;
            c[8 * h + 2 * k + 1] = ff.y;
        }
    }
}

// Lane per pair: the atom's LPA lanes take its in-cutoff pairs in turn (lane q: entries q,
// q + LPA, ...) and each evaluates all four coefficient blocks of its pair, so the pair's
// geometry, the neighbour lookup, the core term and the force accumulation run once per pair.
// The atom's own coefficients sit in LDS (rows padded to 68 floats: the 16 atoms of a wavefront
// read different banks).  Block b: radial descriptors 6b .. 6b+5 and angular channel b.
constexpr int CI_STRIDE = 68;
__global__ void __launch_bounds__(BLOCK) k_sur_force(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ border_map, const int* __restrict__ cl,
    const int* __restrict__ ncl, const __half* __restrict__ coef, float rcore, float kcore,
    double* __restrict__ f)
{
    __shared__ float s_ci[BLOCK / LPA][CI_STRIDE];
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int q    = gid % LPA;
    const int la   = threadIdx.x / LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
    {
        float c[16];
        load_block(coef + (size_t)(act ? a : 0) * NCOEF + 16 * q, c);
#pragma unroll
        for (int m = 0; m < 16; m++)
            s_ci[la][16 * q + m] = c[m];
    }
    __syncthreads();
    const float* ci = s_ci[la];
    const float dk  = 2.0f * R_ETA * R_DMU;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (act) {
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const int nn = ncl[a];
        for (int k = q; k < nn; k += LPA) {
            const int j = cl[(size_t)k * n + a];
            Pair p;
            if (!pair_geom(x, xi, yi, zi, j, &p)) continue;
            const int rj = coef_row(j, nlocal, border_map) - i0;
            if (rj < 0 || rj >= n) continue;  // owner outside the range: the caller's error
            // sum over blocks of the coefficient of u and of the vector terms.  Every product
            // with a neighbour coefficient is an explicit fma with the FP16 value as a factor
            // (one v_fma_mix_f32 each, no conversion).
            float su = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
            auto block = [&](int b, const float* g) {
                const __half2* hj = reinterpret_cast<const __half2*>(coef + (size_t)rj * NCOEF + 16 * b);
                __half2 h[8];
                {
                    const uint4* v = reinterpret_cast<const uint4*>(hj);
                    const uint4 w0 = v[0], w1 = v[1];
                    const unsigned u[8] = { w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w };
#pragma unroll
                    for (int k2 = 0; k2 < 8; k2++)
                        h[k2] = *reinterpret_cast<const __half2*>(&u[k2]);
                }
                auto cj = [&](int m) {
                    return (m & 1) ? __high2float(h[m >> 1]) : __low2float(h[m >> 1]);
                };
                const float* cb = ci + 16 * b;
                // radial: sum_n (c_i + c_j)_n g_n (fc k_n + dfc), k_n = -2 eta (r - mu_n)
                float tk          = fmaf(p.fc,
                             fmaf(-2.0f * R_ETA, p.r, 2.0f * R_ETA * (R_MU0 + RPL * b * R_DMU)), p.dfc);
                const float tstep = p.fc * dk;
#pragma unroll
                for (int m = 0; m < RPL; m++) {
                    const float gt = g[m] * tk;
                    su             = fmaf(cb[m], gt, su);
                    su             = fmaf(cj(m), gt, su);
                    tk += tstep;
                }
                // angular channel b, both atoms at once: a = a_i - a_j, S = S_i + S_j
                const float dr = p.r - (A_MU0 + b * A_DMU);
                const float ex = __expf(-A_ETA * dr * dr);
                const float dw = ex * fmaf(-2.0f * A_ETA * dr, p.fc, p.dfc);
                const float wr = ex * p.fc * p.ir;
                const float w2 = 2.0f * wr;
                float ua       = fmaf(p.uz, cb[8], fmaf(p.uy, cb[7], p.ux * cb[6]));
                ua             = fmaf(-p.uz, cj(8), fmaf(-p.uy, cj(7), fmaf(-p.ux, cj(6), ua)));
                // M = S u, S symmetric: (xx, yy, zz, xy, xz, yz) at 9..14
                float mx = fmaf(cb[9], p.ux, fmaf(cb[12], p.uy, cb[13] * p.uz));
                mx       = fmaf(cj(9), p.ux, fmaf(cj(12), p.uy, fmaf(cj(13), p.uz, mx)));
                float my = fmaf(cb[12], p.ux, fmaf(cb[10], p.uy, cb[14] * p.uz));
                my       = fmaf(cj(12), p.ux, fmaf(cj(10), p.uy, fmaf(cj(14), p.uz, my)));
                float mz = fmaf(cb[13], p.ux, fmaf(cb[14], p.uy, cb[11] * p.uz));
                mz       = fmaf(cj(13), p.ux, fmaf(cj(14), p.uy, fmaf(cj(11), p.uz, mz)));
                const float uMu = fmaf(p.ux, mx, fmaf(p.uy, my, p.uz * mz));
                su              = fmaf(dw, ua + uMu, su);
                su              = fmaf(-wr, ua, su);
                su              = fmaf(-w2, uMu, su);
                vx = fmaf(w2, mx, fmaf(-wr, cj(6), fmaf(wr, cb[6], vx)));
                vy = fmaf(w2, my, fmaf(-wr, cj(7), fmaf(wr, cb[7], vy)));
                vz = fmaf(w2, mz, fmaf(-wr, cj(8), fmaf(wr, cb[8], vz)));
            };
            // the 24 Gaussians by one recurrence from the middle one (as gaussians24): blocks 2
            // and 3 upwards, then 1 and 0 downwards, six values live at a time
            {
                const float t  = p.r - (R_MU0 + RMID * R_DMU);
                const float e2 = R_ETA * R_DMU * R_DMU;
                const float c  = __expf(-2.0f * e2);
                float up       = __expf(2.0f * R_ETA * R_DMU * t - e2);
                float dn       = __expf(-2.0f * R_ETA * R_DMU * t - e2);
                const float gm = __expf(-R_ETA * t * t);
                float g[RPL];
                g[0] = gm;
#pragma unroll
                for (int m = 1; m < RPL; m++) {
                    g[m] = g[m - 1] * up;
                    up *= c;
                }
                block(2, g);
                float gl = g[RPL - 1];
#pragma unroll
                for (int m = 0; m < RPL; m++) {
                    gl   = gl * up;
                    up *= c;
                    g[m] = gl;
                }
                block(3, g);
                gl = gm;
#pragma unroll
                for (int m = RPL - 1; m >= 0; m--) {
                    gl   = gl * dn;
                    dn *= c;
                    g[m] = gl;
                }
                block(1, g);
                gl = g[0];
#pragma unroll
                for (int m = RPL - 1; m >= 0; m--) {
                    gl   = gl * dn;
                    dn *= c;
                    g[m] = gl;
                }
                block(0, g);
            }
            // the core pair term: F_i += U'(r) u, U' = -3 k (rc - r)^2
            if (p.r < rcore) su -= 3.0f * kcore * (rcore - p.r) * (rcore - p.r);
            fx += su * p.ux + vx;
            fy += su * p.uy + vy;
            fz += su * p.uz + vz;
        }
    }
    fx = gsum<LPA>(fx);
    fy = gsum<LPA>(fy);
    fz = gsum<LPA>(fz);
    if (act && q == 0) {
        f[3 * (size_t)a]     = fx;
        f[3 * (size_t)a + 1] = fy;
        f[3 * (size_t)a + 2] = fz;
    }
}

// ---------------------------------------------------------------- the core pair term alone
// Per-atom energy E_i = 1/2 sum_j U(r_ij) and force F_i = sum_j U'(r_ij) u_ij over the full
// Verlet list (FP64), U(r) = k (r_core - r)^3 for r < r_core: the part of the model the network
// does not learn (training labels have it subtracted).
__global__ void k_sur_core(int i0, int n, const double* __restrict__ x,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh, int nlocal, double rcore,
    double kcore, double* __restrict__ e, double* __restrict__ f)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int i = i0 + a;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    double ea = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    const int nn = numneigh[i];
    for (int k = 0; k < nn; k++) {
        const int j    = neighbors[(size_t)k * nlocal + i];
        const double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
        const double r  = sqrt(dx * dx + dy * dy + dz * dz);
        if (!(r < rcore) || r == 0.0) continue;
        const double t  = rcore - r;
        ea += 0.5 * kcore * t * t * t;
        const double du = -3.0 * kcore * t * t / r;  // U'(r) / r
        fx += du * dx;
        fy += du * dy;
        fz += du * dz;
    }
    if (e) e[a] = ea;
    if (f) {
        f[3 * (size_t)a]     = fx;
        f[3 * (size_t)a + 1] = fy;
        f[3 * (size_t)a + 2] = fz;
    }
}

// ---------------------------------------------------------------- training: loss terms
__global__ void k_sur_loss(int n, const float* __restrict__ e, const double* __restrict__ f,
    const double* __restrict__ e_ref, const double* __restrict__ f_ref, float ce_scale,
    float* __restrict__ cE, float* __restrict__ dF, double* __restrict__ red)
{
    __shared__ double s_e[BLOCK / 64], s_f[BLOCK / 64];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double de2 = 0.0, df2 = 0.0;
    if (a < n) {
        const double de = (double)e[a] - e_ref[a];
        cE[a]           = (float)(ce_scale * de);
        de2             = de * de;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double d = f[3 * (size_t)a + c] - f_ref[3 * (size_t)a + c];
            dF[3 * (size_t)a + c] = (float)d;
            df2 += d * d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        de2 += __shfl_xor(de2, o, 64);
        df2 += __shfl_xor(df2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_e[threadIdx.x >> 6] = de2;
        s_f[threadIdx.x >> 6] = df2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tf = 0.0;
        for (int w = 0; w < BLOCK / 64; w++) {
            te += s_e[w];
            tf += s_f[w];
        }
        red[2 * blockIdx.x]     = te;
        red[2 * blockIdx.x + 1] = tf;
    }
}

// ---------------------------------------------------------------- training: force adjoint
// U_i = scale * 2 sum_j J_ij^T e_ij, e_ij = dF_i - dF_owner(j): the derivative of atom i's
// descriptors when neighbour j moves by e_ij.  Lane q produces the lane's own descriptors
// (radial 6 q .. 6 q + 5, q1_q, q2_q).
__global__ void __launch_bounds__(BLOCK) k_sur_adjoint(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ border_map, const int* __restrict__ cl,
    const int* __restrict__ ncl, const float* __restrict__ mom, const float* __restrict__ dF,
    float scale, float* __restrict__ U)
{
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int q    = gid % LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
    float u[RPL], u1 = 0.f, u2 = 0.f;
#pragma unroll
    for (int k = 0; k < RPL; k++)
        u[k] = 0.f;
    const float amu = A_MU0 + q * A_DMU;
    if (act) {
        const float* mo = mom + (size_t)a * NMOM + 9 * q;
        const float Ax = mo[0], Ay = mo[1], Az = mo[2];
        const float Txx = mo[3], Tyy = mo[4], Tzz = mo[5], Txy = mo[6], Txz = mo[7], Tyz = mo[8];
        const float ex0 = dF[3 * (size_t)a], ey0 = dF[3 * (size_t)a + 1], ez0 = dF[3 * (size_t)a + 2];
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const int nn = ncl[a];
        for (int k = 0; k < nn; k++) {
            const int j = cl[(size_t)k * n + a];
            Pair p;
            if (!pair_geom(x, xi, yi, zi, j, &p)) continue;
            const int rj = coef_row(j, nlocal, border_map) - i0;
            if (rj < 0 || rj >= n) continue;  // owner outside the range: the caller's error
            const float ex = ex0 - dF[3 * (size_t)rj];
            const float ey = ey0 - dF[3 * (size_t)rj + 1];
            const float ez = ez0 - dF[3 * (size_t)rj + 2];
            const float ue = p.ux * ex + p.uy * ey + p.uz * ez;
            float g[RPL];
            gaussians6(p.r, q, g);
            float kn       = -2.0f * R_ETA * (p.r - (R_MU0 + RPL * q * R_DMU));
            const float dk = 2.0f * R_ETA * R_DMU;
#pragma unroll
            for (int m = 0; m < RPL; m++) {
                u[m] = fmaf(g[m] * fmaf(p.fc, kn, p.dfc), ue, u[m]);
                kn += dk;
            }
            const float dr  = p.r - amu;
            const float exr = __expf(-A_ETA * dr * dr);
            const float w   = exr * p.fc;
            const float dw  = exr * fmaf(-2.0f * A_ETA * dr, p.fc, p.dfc);
            const float wr  = w * p.ir;
            const float uA  = p.ux * Ax + p.uy * Ay + p.uz * Az;
            const float eA  = ex * Ax + ey * Ay + ez * Az;
            u1 += 2.0f * (dw * ue * uA + wr * (eA - ue * uA));
            const float mx  = Txx * p.ux + Txy * p.uy + Txz * p.uz;
            const float my  = Txy * p.ux + Tyy * p.uy + Tyz * p.uz;
            const float mz  = Txz * p.ux + Tyz * p.uy + Tzz * p.uz;
            const float uTu = p.ux * mx + p.uy * my + p.uz * mz;
            const float eTu = ex * mx + ey * my + ez * mz;
            u2 += 3.0f * (dw * uTu * ue + 2.0f * wr * (eTu - uTu * ue));
        }
    }
    if (!act) return;
    float* o = U + (size_t)a * ND;
#pragma unroll
    for (int m = 0; m < RPL; m++)
        o[RPL * q + m] = 2.0f * scale * u[m];
    o[NR + q]      = 2.0f * scale * u1;
    o[NR + NA + q] = 2.0f * scale * u2;
}

// ---------------------------------------------------------------- training: weight gradient
// Per tile of 16 atoms, FP32 and exact (no bfloat16): with u~ = U / std,
//   z1 = W1 Dn + b1, h1 = tanh z1, zd1 = W1 u~, hd1 = (1 - h1^2) zd1
//   z2 = W2 h1 + b2, h2 = tanh z2, zd2 = W2 hd1
//   a2 = s (1 - h2^2) w3,  b2 = -2 s w3 zd2 h2 (1 - h2^2)
//   g1 = W2^T a2, q = W2^T b2, a1 = (1 - h1^2) g1, b1 = -2 h1 (1 - h1^2) zd1 g1 + (1 - h1^2) q
//   dW1 += (b1 + cE a1) Dn^T + a1 u~^T, db1 += b1 + cE a1
//   dW2 += (b2 + cE a2) h1^T + a2 hd1^T, db2 += b2 + cE a2
//   dw3 += s ((1 - h2^2) zd2 + cE h2), db3 += s cE
// Each block accumulates its tiles in registers (thread t owns parameters t, t + 256, ...) and
// writes one partial row; k_sur_wsum adds the rows in block order.
constexpr int WT = 16;  // atoms per tile (keeps the block's LDS under 64 KB)
constexpr int PPT = (NPARAM + BLOCK - 1) / BLOCK;  // parameters per thread (25)

__global__ void __launch_bounds__(BLOCK) k_sur_wgrad(int n, const float* __restrict__ dn,
    const float* __restrict__ U, const float* __restrict__ cE, const float* __restrict__ th,
    const float* __restrict__ norm, float s_e, float* __restrict__ part)
{
    __shared__ float sW1[HID * ND], sW2[HID * HID], sb1[HID], sb2[HID], sw3[HID], sistd[ND];
    __shared__ float sD[WT][ND], sU[WT][ND], sc[WT];
    __shared__ float sH1[WT][HID], sZd1[WT][HID], sH2[WT][HID], sZd2[WT][HID];
    __shared__ float sA2[WT][HID], sB2[WT][HID], sG1[WT][HID], sQ[WT][HID];
    const int tid = threadIdx.x;
    for (int k = tid; k < HID * ND; k += BLOCK)
        sW1[k] = th[O_W1 + k];
    for (int k = tid; k < HID * HID; k += BLOCK)
        sW2[k] = th[O_W2 + k];
    for (int k = tid; k < HID; k += BLOCK) {
        sb1[k] = th[O_B1 + k];
        sb2[k] = th[O_B2 + k];
        sw3[k] = th[O_W3 + k];
    }
    for (int k = tid; k < ND; k += BLOCK)
        sistd[k] = norm[ND + k];
    float acc[PPT];
#pragma unroll
    for (int q = 0; q < PPT; q++)
        acc[q] = 0.f;
    const int ntile = (n + WT - 1) / WT;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        __syncthreads();
        // u~ = U / std (U is dL/dg in raw descriptor units; the network sees Dn = (D - mean) / std)
        for (int k = tid; k < WT * ND; k += BLOCK) {
            const int a = k / ND, d = k % ND, atom = tile * WT + a;
            const bool ok = atom < n;
            sD[a][d] = ok ? dn[(size_t)atom * ND + d] : 0.f;
            sU[a][d] = ok ? U[(size_t)atom * ND + d] * sistd[d] : 0.f;
        }
        for (int a = tid; a < WT; a += BLOCK) {
            const int atom = tile * WT + a;
            sc[a]          = atom < n ? cE[atom] : 0.f;
        }
        __syncthreads();
        // z1, h1, zd1
        for (int k = tid; k < WT * HID; k += BLOCK) {
            const int a = k / HID, o = k % HID;
            float z = sb1[o], zd = 0.f;
            for (int d = 0; d < ND; d++) {
                z  = fmaf(sW1[o * ND + d], sD[a][d], z);
                zd = fmaf(sW1[o * ND + d], sU[a][d], zd);
            }
            sH1[a][o]  = tanhf(z);
            sZd1[a][o] = zd;
        }
        __syncthreads();
        // z2, h2, zd2 (from hd1 = (1 - h1^2) zd1)
        for (int k = tid; k < WT * HID; k += BLOCK) {
            const int a = k / HID, o = k % HID;
            float z = sb2[o], zd = 0.f;
            for (int h = 0; h < HID; h++) {
                const float w = sW2[o * HID + h];
                const float hv = sH1[a][h];
                z  = fmaf(w, hv, z);
                zd = fmaf(w, (1.0f - hv * hv) * sZd1[a][h], zd);
            }
            const float h2 = tanhf(z);
            const float t2 = 1.0f - h2 * h2;
            sH2[a][o]  = h2;
            sZd2[a][o] = zd;
            sA2[a][o]  = s_e * t2 * sw3[o];
            sB2[a][o]  = -2.0f * s_e * sw3[o] * zd * h2 * t2;
        }
        __syncthreads();
        // g1 = W2^T a2, q = W2^T b2  ->  a1, b1 (stored in sG1, sQ)
        for (int k = tid; k < WT * HID; k += BLOCK) {
            const int a = k / HID, h = k % HID;
            float g1 = 0.f, q = 0.f;
            for (int o = 0; o < HID; o++) {
                const float w = sW2[o * HID + h];
                g1 = fmaf(w, sA2[a][o], g1);
                q  = fmaf(w, sB2[a][o], q);
            }
            const float h1 = sH1[a][h];
            const float t1 = 1.0f - h1 * h1;
            sG1[a][h] = t1 * g1;
            sQ[a][h]  = -2.0f * h1 * t1 * sZd1[a][h] * g1 + t1 * q;
        }
        __syncthreads();
        // accumulate this tile into the thread's parameters
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int p = tid + q * BLOCK;
            if (p >= NPARAM) break;
            float s = 0.f;
            if (p < O_B1) {
                const int o = p / ND, d = p % ND;
                for (int a = 0; a < WT; a++)
                    s += (sQ[a][o] + sc[a] * sG1[a][o]) * sD[a][d] + sG1[a][o] * sU[a][d];
            } else if (p < O_W2) {
                const int o = p - O_B1;
                for (int a = 0; a < WT; a++)
                    s += sQ[a][o] + sc[a] * sG1[a][o];
            } else if (p < O_B2) {
                const int o = (p - O_W2) / HID, h = (p - O_W2) % HID;
                for (int a = 0; a < WT; a++) {
                    const float h1 = sH1[a][h];
                    s += (sB2[a][o] + sc[a] * sA2[a][o]) * h1 +
                         sA2[a][o] * (1.0f - h1 * h1) * sZd1[a][h];
                }
            } else if (p < O_W3) {
                const int o = p - O_B2;
                for (int a = 0; a < WT; a++)
                    s += sB2[a][o] + sc[a] * sA2[a][o];
            } else if (p < O_B3) {
                const int o = p - O_W3;
                for (int a = 0; a < WT; a++) {
                    const float h2 = sH2[a][o];
                    s += s_e * ((1.0f - h2 * h2) * sZd2[a][o] + sc[a] * h2);
                }
            } else {
                for (int a = 0; a < WT; a++)
                    s += s_e * sc[a];
            }
            acc[q] += s;
        }
    }
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        const int p = tid + q * BLOCK;
        if (p < NPARAM) part[(size_t)blockIdx.x * NPARAM + p] = acc[q];
    }
}

__global__ void k_sur_wsum(int nblocks, const float* __restrict__ part, float* __restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++)
        s += part[(size_t)b * NPARAM + p];
    grad[p] += s;
}

__global__ void k_sur_adam(int np, float* __restrict__ th, const float* __restrict__ gr,
    float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
    float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float g  = gr[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p]           = mm;
    v[p]           = vv;
    th[p] -= lr * (mm / c1) / (__builtin_sqrtf(vv / c2) + eps);
}

__global__ void k_sur_sqdiff(int n3, const double* __restrict__ a, const double* __restrict__ b,
    double* __restrict__ red)
{
    __shared__ double s[BLOCK / 64];
    double t = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += gridDim.x * blockDim.x) {
        const double d = a[k] - b[k];
        t += d * d;
    }
    for (int o = 32; o > 0; o >>= 1)
        t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ void k_sur_probe(const float* __restrict__ A, const float* __restrict__ B,
    float* __restrict__ C)
{
    const int l = threadIdx.x, m = l & 15, g = l >> 4;
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16)A[m * 32 + 8 * g + i];
        b[i] = (__bf16)B[(8 * g + i) * 16 + m];
    }
    floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
    acc         = mfma(a, b, acc);
#pragma unroll
    for (int i = 0; i < 4; i++)
        C[(4 * g + i) * 16 + m] = acc[i];
}

bool model_ok()
{
    if (failed()) return false;
    if (!M.ready) {
        set_error(E_ARG, "surrogate: no model (surSetModelGFX950 first)");
        return false;
    }
    return true;
}

void pack()
{
    k_sur_pack<<<NFRAG, 64, 0, stream()>>>(M.theta, M.frag);
    MDG_LAUNCH_CHECK("sur_pack");
}

bool alloc_model()
{
    if (M.theta) return true;
    MDG_HIP(hipMalloc(&M.theta, NPARAM * sizeof(float)));
    MDG_HIP(hipMalloc(&M.frag, NFRAG * 64 * sizeof(bf16x8)));
    MDG_HIP(hipMalloc(&M.norm, 2 * ND * sizeof(float)));
    MDG_HIP(hipHostMalloc(&M.h_red, 4096 * sizeof(double)));
    return !failed();
}

double reduce_blocks(double* d, int nblk)
{
    MDG_HIP(hipMemcpyAsync(M.h_red, d, nblk * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    double s = 0.0;
    for (int b = 0; b < nblk; b++)
        s += M.h_red[b];
    return s;
}

}  // namespace

// ==================================================================== C entry points
extern "C" {

void* mdgDeviceAlloc(size_t bytes)
{
    if (failed() || !stream()) return nullptr;
    void* p = nullptr;
    MDG_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return failed() ? nullptr : p;
}

void mdgDeviceFree(void* p)
{
    if (p) (void)hipFree(p);
}

int mdgCopyToDevice(void* dst, const void* src, size_t bytes)
{
    if (failed() || !stream()) return -1;
    MDG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? -1 : 0;
}

int mdgCopyFromDevice(void* dst, const void* src, size_t bytes)
{
    if (failed() || !stream()) return -1;
    MDG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    return failed() ? -1 : 0;
}

int mdgDeviceZero(void* p, size_t bytes)
{
    if (failed() || !stream()) return -1;
    if (bytes % 4) {
        set_error(E_ARG, "mdgDeviceZero: size %zu is not a multiple of 4", bytes);
        return -1;
    }
    fill_f32((float*)p, 0.0f, bytes / 4);
    return failed() ? -1 : 0;
}

int surSetModelGFX950(const float* theta, const float* d_mean, const float* d_std, float e0,
    float s_e)
{
    if (failed() || !stream() || !alloc_model()) return -1;
    float norm[2 * ND];
    for (int k = 0; k < ND; k++) {
        norm[k]      = d_mean[k];
        norm[ND + k] = d_std[k] > 0.f ? 1.0f / d_std[k] : 1.0f;
    }
    MDG_HIP(hipMemcpyAsync(M.theta, theta, NPARAM * sizeof(float), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(M.norm, norm, sizeof(norm), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    M.e0  = e0;
    M.s_e = s_e;
    pack();
    M.ready = !failed();
    return M.ready ? 0 : -1;
}

int surSetParamsDeviceGFX950(const float* theta_dev)
{
    if (!model_ok()) return -1;
    if (theta_dev != M.theta)
        MDG_HIP(hipMemcpyAsync(M.theta, theta_dev, NPARAM * sizeof(float),
            hipMemcpyDeviceToDevice, stream()));
    pack();
    return failed() ? -1 : 0;
}

int surGetParamsGFX950(float* theta_host)
{
    if (!model_ok()) return -1;
    return mdgCopyFromDevice(theta_host, M.theta, NPARAM * sizeof(float));
}

int surDescriptorsGFX950(const MdgSurSystem* sys, int i0, int n, float* dn, float* mom, int cap,
    int* pairs, int* npairs)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surDescriptorsGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    TimedRegion tr(F_SURROGATE);
    k_sur_desc<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0, n,
        sys->x, sys->neighbors, sys->numneigh, M.norm, dn, mom, cap, pairs, npairs);
    MDG_LAUNCH_CHECK("sur_desc");
    return failed() ? -1 : 0;
}

int surEvaluateGFX950(int n, const float* dn, const float* mom, float* e, void* coef)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    TimedRegion tr(F_SURROGATE);
    const int ntile = (n + 15) / 16;
    int grid        = (ntile + 3) / 4;
    if (grid > 2048) grid = 2048;
    k_sur_mlp<<<grid, BLOCK, 0, stream()>>>(n, dn, mom, M.frag, M.theta, M.norm, M.s_e, M.e0, e,
        (__half*)coef);
    MDG_LAUNCH_CHECK("sur_mlp");
    return failed() ? -1 : 0;
}

int surForcesGFX950(const MdgSurSystem* sys, int i0, int n, const int* pairs, const int* npairs,
    const void* coef, double* f)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surForcesGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    TimedRegion tr(F_SURROGATE);
    k_sur_force<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0, n,
        sys->x, sys->border_map, pairs, npairs, (const __half*)coef, M.r_core, M.k_core, f);
    MDG_LAUNCH_CHECK("sur_force");
    return failed() ? -1 : 0;
}

int surSetCoreGFX950(float r_core, float k_core)
{
    if (!(r_core >= 0.f) || !(r_core < RC) || !(k_core >= 0.f)) {
        set_error(E_ARG, "surSetCoreGFX950: r_core %g (must be in [0, %g)), k_core %g", r_core,
            RC, k_core);
        return -1;
    }
    M.r_core = r_core;
    M.k_core = k_core;
    return 0;
}

int surCoreGFX950(const MdgSurSystem* sys, int i0, int n, float r_core, float k_core, double* e,
    double* f)
{
    if (failed() || !stream()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surCoreGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    k_sur_core<<<blocks_for(n, BLOCK), BLOCK, 0, stream()>>>(i0, n, sys->x, sys->neighbors,
        sys->numneigh, sys->nlocal, (double)r_core, (double)k_core, e, f);
    MDG_LAUNCH_CHECK("sur_core");
    return failed() ? -1 : 0;
}

int surLossTermsGFX950(int n, const float* e, const double* f, const double* e_ref,
    const double* f_ref, float wE, int n_norm, float* cE, float* dF, double* sums)
{
    if (failed() || !stream() || !alloc_model()) return -1;
    if (n <= 0) {
        if (sums) sums[0] = sums[1] = 0.0;
        return 0;
    }
    const int nblk = (int)blocks_for(n, BLOCK);
    grow_n(&M.red, &M.cap_red, 2 * (size_t)nblk, 0);
    if (failed()) return -1;
    k_sur_loss<<<nblk, BLOCK, 0, stream()>>>(n, e, f, e_ref, f_ref, 2.0f * wE / (float)n_norm, cE,
        dF, M.red);
    MDG_LAUNCH_CHECK("sur_loss");
    if (sums) {
        if (nblk > 2048) {
            set_error(E_ARG, "surLossTermsGFX950: more than %d atoms per call", 2048 * BLOCK);
            return -1;
        }
        MDG_HIP(hipMemcpyAsync(M.h_red, M.red, 2 * nblk * sizeof(double), hipMemcpyDeviceToHost,
            stream()));
        MDG_HIP(hipStreamSynchronize(stream()));
        double se = 0.0, sf = 0.0;
        for (int b = 0; b < nblk; b++) {
            se += M.h_red[2 * b];
            sf += M.h_red[2 * b + 1];
        }
        sums[0] = se;
        sums[1] = sf;
    }
    return failed() ? -1 : 0;
}

int surAdjointGFX950(const MdgSurSystem* sys, int i0, int n, const int* pairs, const int* npairs,
    const float* mom, const float* dF, float scale, float* U)
{
    if (failed() || !stream()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surAdjointGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    k_sur_adjoint<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0,
        n, sys->x, sys->border_map, pairs, npairs, mom, dF, scale, U);
    MDG_LAUNCH_CHECK("sur_adjoint");
    return failed() ? -1 : 0;
}

int surWeightGradGFX950(int n, const float* dn, const float* U, const float* cE, float* grad)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    const int ntile = (n + WT - 1) / WT;
    const int nblk  = ntile < 512 ? ntile : 512;
    grow_n(&M.part, &M.cap_part, (size_t)nblk * NPARAM, 0);
    if (failed()) return -1;
    k_sur_wgrad<<<nblk, BLOCK, 0, stream()>>>(n, dn, U, cE, M.theta, M.norm, M.s_e, M.part);
    k_sur_wsum<<<blocks_for(NPARAM, BLOCK), BLOCK, 0, stream()>>>(nblk, M.part, grad);
    MDG_LAUNCH_CHECK("sur_wgrad");
    return failed() ? -1 : 0;
}

int surAdamGFX950(float* theta, const float* grad, float* m, float* v, int t, float lr,
    float beta1, float beta2, float eps)
{
    if (failed() || !stream()) return -1;
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    k_sur_adam<<<blocks_for(NPARAM, BLOCK), BLOCK, 0, stream()>>>(NPARAM, theta, grad, m, v, lr,
        beta1, beta2, eps, c1, c2);
    MDG_LAUNCH_CHECK("sur_adam");
    return failed() ? -1 : 0;
}

double surForceRmseGFX950(int n, const double* fa, const double* fb)
{
    if (failed() || !stream() || !alloc_model()) return -1.0;
    if (n <= 0) return 0.0;
    const int nblk = 1024;
    grow_n(&M.red, &M.cap_red, nblk, 0);
    if (failed()) return -1.0;
    k_sur_sqdiff<<<nblk, BLOCK, 0, stream()>>>(3 * n, fa, fb, M.red);
    MDG_LAUNCH_CHECK("sur_sqdiff");
    const double s = reduce_blocks(M.red, nblk);
    return failed() ? -1.0 : sqrt(s / (3.0 * n));
}

int surMfmaProbeGFX950(const float* A, const float* B, float* C)
{
    if (failed() || !stream()) return -1;
    float* d = nullptr;
    MDG_HIP(hipMalloc(&d, (512 + 512 + 256) * sizeof(float)));
    if (failed()) return -1;
    MDG_HIP(hipMemcpyAsync(d, A, 512 * sizeof(float), hipMemcpyHostToDevice, stream()));
    MDG_HIP(hipMemcpyAsync(d + 512, B, 512 * sizeof(float), hipMemcpyHostToDevice, stream()));
    k_sur_probe<<<1, 64, 0, stream()>>>(d, d + 512, d + 1024);
    MDG_LAUNCH_CHECK("sur_probe");
    MDG_HIP(hipMemcpyAsync(C, d + 1024, 256 * sizeof(float), hipMemcpyDeviceToHost, stream()));
    MDG_HIP(hipStreamSynchronize(stream()));
    (void)hipFree(d);
    return failed() ? -1 : 0;
}

double computeForceSurrogateGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats)
{
    (void)param;
    (void)stats;
    if (!model_ok()) return 0.0;
    const double t0 = wtime();
    const int nl    = atom->Nlocal;
    grow_n(&M.dn, &M.cap_dn, (size_t)nl * ND + 4, 0);
    grow_n(&M.mom, &M.cap_mom, (size_t)nl * NMOM + 4, 0);
    grow_n(&M.coef, &M.cap_coef, (size_t)nl * NCOEF + 8, 0);
    const int cap = neighbor->maxneighs;
    grow_n(&M.pairs, &M.cap_pairs, (size_t)nl * cap + 4, 0);
    grow_n(&M.npairs, &M.cap_npairs, (size_t)nl + 4, 0);
    if (failed()) return 0.0;
    MdgSurSystem sys{ nl, atom->Nghost, atom->d_atom.x, atom->d_atom.border_map,
        neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh };
    surDescriptorsGFX950(&sys, 0, nl, M.dn, M.mom, cap, M.pairs, M.npairs);
    surEvaluateGFX950(nl, M.dn, M.mom, nullptr, M.coef);
    surForcesGFX950(&sys, 0, nl, M.pairs, M.npairs, M.coef, atom->d_atom.fx);
    return wtime() - t0;
}

double surrogateJudgeGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats,
    double tol, int* fallback)
{
    if (!model_ok()) return -1.0;
    const int nl = atom->Nlocal;
    grow_n(&M.fsave, &M.cap_fsave, 3 * (size_t)nl + 4, 0);
    if (failed()) return -1.0;
    MDG_HIP(hipMemcpyAsync(M.fsave, atom->d_atom.fx, 3 * (size_t)nl * sizeof(double),
        hipMemcpyDeviceToDevice, stream()));
    computeForceEamGFX950(param, atom, neighbor, stats);
    const double rmse = surForceRmseGFX950(nl, M.fsave, atom->d_atom.fx);
    if (failed()) return -1.0;
    if (rmse <= tol) {
        MDG_HIP(hipMemcpyAsync(atom->d_atom.fx, M.fsave, 3 * (size_t)nl * sizeof(double),
            hipMemcpyDeviceToDevice, stream()));
    } else if (fallback) {
        *fallback = 1;
    }
    return rmse;
}

}  // extern "C"
