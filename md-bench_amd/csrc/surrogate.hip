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
//   k_sur_force   LPA lanes per atom: F_k = sum_j (dE_k/dx_j - dE_j/dx_k), the j side from the
//                 neighbour's coefficient row (a ghost reads its owner's row).
// Training adds the force adjoint (k_sur_adjoint), the exact FP32 weight gradient of the loss
// (k_sur_wgrad: forward mode along the adjoint, then reverse mode, block partials reduced in a
// fixed order) and Adam (k_sur_adam).
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
    bool ready = false;
    // per-atom work arrays of the drop-in force call
    float *dn = nullptr, *mom = nullptr, *coef = nullptr;
    size_t cap_dn = 0, cap_mom = 0, cap_coef = 0;
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
// Gaussians exp(-eta (r - mu_n)^2) by a multiplicative recurrence from the middle centre: the
// ratio of neighbours is exp(+-2 eta D (r - mu_n) - eta D^2) and changes by exp(-2 eta D^2) per
// step, so three exponentials give all 24 values, and no intermediate over- or underflows
// spuriously (values shrink monotonically away from the middle when they are small).
__device__ inline void gaussians(float r, float* g)
{
    const float t  = r - (R_MU0 + RMID * R_DMU);
    const float c  = __expf(-2.0f * R_ETA * R_DMU * R_DMU);
    float up       = __expf(2.0f * R_ETA * R_DMU * t - R_ETA * R_DMU * R_DMU);
    float dn       = __expf(-2.0f * R_ETA * R_DMU * t - R_ETA * R_DMU * R_DMU);
    g[RMID]        = __expf(-R_ETA * t * t);
#pragma unroll
    for (int n = RMID + 1; n < NR; n++) {
        g[n] = g[n - 1] * up;
        up *= c;
    }
#pragma unroll
    for (int n = RMID - 1; n >= 0; n--) {
        g[n] = g[n + 1] * dn;
        dn *= c;
    }
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

// ---------------------------------------------------------------- descriptors
__global__ void __launch_bounds__(BLOCK) k_sur_desc(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ neighbors,
    const int* __restrict__ numneigh, const float* __restrict__ norm, float* __restrict__ dn_out,
    float* __restrict__ mom_out)
{
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
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
    if (act) {
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const int nn = numneigh[i];
        for (int k = lane; k < nn; k += LPA) {
            const int j    = neighbors[(size_t)k * nlocal + i];
            const float dx = (float)(x[3 * j] - xi);
            const float dy = (float)(x[3 * j + 1] - yi);
            const float dz = (float)(x[3 * j + 2] - zi);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < RC2)) continue;
            const float r    = __builtin_sqrtf(r2);
            const float ir   = rcp_refined(r);
            const float ux = dx * ir, uy = dy * ir, uz = dz * ir;
            float fc, dfc;
            cutoff_fn(r2, r, &fc, &dfc);
            float g[NR];
            gaussians(r, g);
#pragma unroll
            for (int m = 0; m < NR; m++)
                G[m] = fmaf(g[m], fc, G[m]);
            const float uu[6] = { ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz };
#pragma unroll
            for (int c = 0; c < NA; c++) {
                const float dr = r - (A_MU0 + c * A_DMU);
                const float w  = __expf(-A_ETA * dr * dr) * fc;
                A[c][0] = fmaf(w, ux, A[c][0]);
                A[c][1] = fmaf(w, uy, A[c][1]);
                A[c][2] = fmaf(w, uz, A[c][2]);
#pragma unroll
                for (int q = 0; q < 6; q++)
                    T[c][q] = fmaf(w, uu[q], T[c][q]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NR; k++)
        G[k] = gsum<LPA>(G[k]);
#pragma unroll
    for (int c = 0; c < NA; c++) {
#pragma unroll
        for (int q = 0; q < 3; q++)
            A[c][q] = gsum<LPA>(A[c][q]);
#pragma unroll
        for (int q = 0; q < 6; q++)
            T[c][q] = gsum<LPA>(T[c][q]);
    }
    if (!act || lane != 0) return;
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
        mo[3 * c]     = A[c][0];
        mo[3 * c + 1] = A[c][1];
        mo[3 * c + 2] = A[c][2];
#pragma unroll
        for (int q = 0; q < 6; q++)
            mo[12 + 6 * c + q] = T[c][q];
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
    float* __restrict__ coef)
{
    __shared__ bf16x8 s_frag[NFRAG * 64];
    for (int k = threadIdx.x; k < NFRAG * 64; k += blockDim.x)
        s_frag[k] = frag[k];
    __syncthreads();
    const int l = threadIdx.x & 63, m = l & 15, g = l >> 4;
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
                h1[t][i] = tanhf(acc[i] + b1[t][i]);
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
                h2[t][i] = tanhf(acc[i] + b2[t][i]);
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
        if (!act) continue;
        // force coefficients: [24 radial][4 x 3: 2 g_q1 A_c][4 x 6: 3 g_q2 T~_c][4 zero]
        float* co       = coef + (size_t)atom * NCOEF;
        const float* mo = mom + (size_t)atom * NMOM;
#pragma unroll
        for (int i = 0; i < 4; i++)
            co[4 * g + i] = gd[0][i];
        if (g < 2) {
#pragma unroll
            for (int i = 0; i < 4; i++)
                co[16 + 4 * g + i] = gd[1][i];
        } else if (g == 2) {
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int q = 0; q < 3; q++)
                    co[NR + 3 * c + q] = 2.0f * gd[1][c] * mo[3 * c + q];
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int q = 0; q < 6; q++)
                    co[NR + 12 + 6 * c + q] = 3.0f * gd[1][c] * mo[12 + 6 * c + q];
#pragma unroll
            for (int q = 60; q < 64; q++)
                co[q] = 0.f;
        }
    }
}

// ---------------------------------------------------------------- forces
// dE_i/dx_j (u = (x_j - x_i) / r) from atom i's angular coefficients, channel by channel:
//   w'(u.a) u + (w/r)(a - (u.a) u) + w'(u.M u) u + (2 w/r)(M u - (u.M u) u)
__device__ inline void ang_grad(const float* co, float ux, float uy, float uz, float dw, float wr,
    float* s, float* vx, float* vy, float* vz)
{
    const float ax = co[0], ay = co[1], az = co[2];
    const float ua = ux * ax + uy * ay + uz * az;
    *s += dw * ua;
    *vx += wr * (ax - ua * ux);
    *vy += wr * (ay - ua * uy);
    *vz += wr * (az - ua * uz);
}

__device__ inline void sym_grad(const float* B, float ux, float uy, float uz, float dw, float wr,
    float* s, float* vx, float* vy, float* vz)
{
    // M u with M = [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
    const float mx  = B[0] * ux + B[3] * uy + B[4] * uz;
    const float my  = B[3] * ux + B[1] * uy + B[5] * uz;
    const float mz  = B[4] * ux + B[5] * uy + B[2] * uz;
    const float uMu = ux * mx + uy * my + uz * mz;
    *s += dw * uMu;
    *vx += 2.0f * wr * (mx - uMu * ux);
    *vy += 2.0f * wr * (my - uMu * uy);
    *vz += 2.0f * wr * (mz - uMu * uz);
}

__global__ void __launch_bounds__(BLOCK) k_sur_force(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ border_map,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh,
    const float* __restrict__ coef, double* __restrict__ f)
{
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
    float ci[60];
    {
        const float4* r4 = reinterpret_cast<const float4*>(coef + (size_t)(act ? a : 0) * NCOEF);
#pragma unroll
        for (int q = 0; q < 15; q++) {
            const float4 v = r4[q];
            ci[4 * q] = v.x; ci[4 * q + 1] = v.y; ci[4 * q + 2] = v.z; ci[4 * q + 3] = v.w;
        }
    }
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (act) {
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const int nn = numneigh[i];
        for (int k = lane; k < nn; k += LPA) {
            const int j    = neighbors[(size_t)k * nlocal + i];
            const float dx = (float)(x[3 * j] - xi);
            const float dy = (float)(x[3 * j + 1] - yi);
            const float dz = (float)(x[3 * j + 2] - zi);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < RC2)) continue;
            const int rj = coef_row(j, nlocal, border_map) - i0;
            if (rj < 0 || rj >= n) continue;  // owner outside the range: the caller's error
            const float4* cj4 = reinterpret_cast<const float4*>(coef + (size_t)rj * NCOEF);
            float cj[60];
#pragma unroll
            for (int q = 0; q < 15; q++) {
                const float4 v = cj4[q];
                cj[4 * q] = v.x; cj[4 * q + 1] = v.y; cj[4 * q + 2] = v.z; cj[4 * q + 3] = v.w;
            }
            const float r  = __builtin_sqrtf(r2);
            const float ir = rcp_refined(r);
            const float ux = dx * ir, uy = dy * ir, uz = dz * ir;
            float fc, dfc;
            cutoff_fn(r2, r, &fc, &dfc);
            float g[NR];
            gaussians(r, g);
            // radial: (S_i + S_j) u with S = sum_n c_n G_n'(r), G_n' = g_n (fc k_n + dfc),
            // k_n = -2 eta (r - mu_n)
            float S     = 0.f;
            float kn    = -2.0f * R_ETA * (r - R_MU0);
            const float dk = 2.0f * R_ETA * R_DMU;
#pragma unroll
            for (int q = 0; q < NR; q++) {
                S = fmaf((ci[q] + cj[q]) * g[q], fmaf(fc, kn, dfc), S);
                kn += dk;
            }
            float vx = 0.f, vy = 0.f, vz = 0.f;  // v_ij - v_ji minus its radial part S u
#pragma unroll
            for (int c = 0; c < NA; c++) {
                const float dr = r - (A_MU0 + c * A_DMU);
                const float ex = __expf(-A_ETA * dr * dr);
                const float w  = ex * fc;
                const float dw = ex * fmaf(-2.0f * A_ETA * dr, fc, dfc);
                const float wr = w * ir;
                // i side, direction u
                float si = 0.f, pix = 0.f, piy = 0.f, piz = 0.f;
                ang_grad(ci + NR + 3 * c, ux, uy, uz, dw, wr, &si, &pix, &piy, &piz);
                sym_grad(ci + NR + 12 + 6 * c, ux, uy, uz, dw, wr, &si, &pix, &piy, &piz);
                // j side, direction -u
                float sj = 0.f, pjx = 0.f, pjy = 0.f, pjz = 0.f;
                ang_grad(cj + NR + 3 * c, -ux, -uy, -uz, dw, wr, &sj, &pjx, &pjy, &pjz);
                sym_grad(cj + NR + 12 + 6 * c, -ux, -uy, -uz, dw, wr, &sj, &pjx, &pjy, &pjz);
                // v_ij = si u + p_i ; v_ji = sj (-u) + p_j
                S += si + sj;
                vx += pix - pjx;
                vy += piy - pjy;
                vz += piz - pjz;
            }
            fx += S * ux + vx;
            fy += S * uy + vy;
            fz += S * uz + vz;
        }
    }
    fx = gsum<LPA>(fx);
    fy = gsum<LPA>(fy);
    fz = gsum<LPA>(fz);
    if (act && lane == 0) {
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
// descriptors when neighbour j moves by e_ij.
__global__ void __launch_bounds__(BLOCK) k_sur_adjoint(int nlocal, int i0, int n,
    const double* __restrict__ x, const int* __restrict__ border_map,
    const int* __restrict__ neighbors, const int* __restrict__ numneigh,
    const float* __restrict__ mom, const float* __restrict__ dF, float scale,
    float* __restrict__ U)
{
    const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const int a    = gid / LPA;
    const int lane = gid % LPA;
    const bool act = a < n;
    const int i    = i0 + (act ? a : 0);
    float u[ND];
#pragma unroll
    for (int k = 0; k < ND; k++)
        u[k] = 0.f;
    if (act) {
        const float* mo = mom + (size_t)a * NMOM;
        float Am[NMOM];
#pragma unroll
        for (int k = 0; k < NMOM; k++)
            Am[k] = mo[k];
        const float ex0 = dF[3 * (size_t)a], ey0 = dF[3 * (size_t)a + 1], ez0 = dF[3 * (size_t)a + 2];
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const int nn = numneigh[i];
        for (int k = lane; k < nn; k += LPA) {
            const int j    = neighbors[(size_t)k * nlocal + i];
            const float dx = (float)(x[3 * j] - xi);
            const float dy = (float)(x[3 * j + 1] - yi);
            const float dz = (float)(x[3 * j + 2] - zi);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < RC2)) continue;
            const int rj   = coef_row(j, nlocal, border_map) - i0;
            if (rj < 0 || rj >= n) continue;  // owner outside the range: the caller's error
            const float ex = ex0 - dF[3 * (size_t)rj];
            const float ey = ey0 - dF[3 * (size_t)rj + 1];
            const float ez = ez0 - dF[3 * (size_t)rj + 2];
            const float r  = __builtin_sqrtf(r2);
            const float ir = rcp_refined(r);
            const float ux = dx * ir, uy = dy * ir, uz = dz * ir;
            const float ue = ux * ex + uy * ey + uz * ez;
            float fc, dfc;
            cutoff_fn(r2, r, &fc, &dfc);
            float g[NR];
            gaussians(r, g);
            float kn       = -2.0f * R_ETA * (r - R_MU0);
            const float dk = 2.0f * R_ETA * R_DMU;
#pragma unroll
            for (int q = 0; q < NR; q++) {
                u[q] = fmaf(g[q] * fmaf(fc, kn, dfc), ue, u[q]);
                kn += dk;
            }
#pragma unroll
            for (int c = 0; c < NA; c++) {
                const float drr = r - (A_MU0 + c * A_DMU);
                const float exr = __expf(-A_ETA * drr * drr);
                const float w   = exr * fc;
                const float dw  = exr * fmaf(-2.0f * A_ETA * drr, fc, dfc);
                const float wr  = w * ir;
                const float* Ac = Am + 3 * c;
                const float uA  = ux * Ac[0] + uy * Ac[1] + uz * Ac[2];
                const float eA  = ex * Ac[0] + ey * Ac[1] + ez * Ac[2];
                u[NR + c] += 2.0f * (dw * ue * uA + wr * (eA - ue * uA));
                const float* B  = Am + 12 + 6 * c;
                const float mx  = B[0] * ux + B[3] * uy + B[4] * uz;
                const float my  = B[3] * ux + B[1] * uy + B[5] * uz;
                const float mz  = B[4] * ux + B[5] * uy + B[2] * uz;
                const float uTu = ux * mx + uy * my + uz * mz;
                const float eTu = ex * mx + ey * my + ez * mz;
                u[NR + NA + c] += 3.0f * (dw * uTu * ue + 2.0f * wr * (eTu - uTu * ue));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < ND; k++)
        u[k] = gsum<LPA>(u[k]);
    if (act && lane == 0) {
        float* o = U + (size_t)a * ND;
#pragma unroll
        for (int k = 0; k < ND; k++)
            o[k] = 2.0f * scale * u[k];
    }
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

int surDescriptorsGFX950(const MdgSurSystem* sys, int i0, int n, float* dn, float* mom)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surDescriptorsGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    TimedRegion tr(F_SURROGATE);
    k_sur_desc<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0, n,
        sys->x, sys->neighbors, sys->numneigh, M.norm, dn, mom);
    MDG_LAUNCH_CHECK("sur_desc");
    return failed() ? -1 : 0;
}

int surEvaluateGFX950(int n, const float* dn, const float* mom, float* e, float* coef)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    TimedRegion tr(F_SURROGATE);
    const int ntile = (n + 15) / 16;
    int grid        = (ntile + 3) / 4;
    if (grid > 2048) grid = 2048;
    k_sur_mlp<<<grid, BLOCK, 0, stream()>>>(n, dn, mom, M.frag, M.theta, M.norm, M.s_e, M.e0, e,
        coef);
    MDG_LAUNCH_CHECK("sur_mlp");
    return failed() ? -1 : 0;
}

int surForcesGFX950(const MdgSurSystem* sys, int i0, int n, const float* coef, double* f)
{
    if (!model_ok()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surForcesGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    TimedRegion tr(F_SURROGATE);
    k_sur_force<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0, n,
        sys->x, sys->border_map, sys->neighbors, sys->numneigh, coef, f);
    MDG_LAUNCH_CHECK("sur_force");
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

int surAdjointGFX950(const MdgSurSystem* sys, int i0, int n, const float* mom, const float* dF,
    float scale, float* U)
{
    if (failed() || !stream()) return -1;
    if (n <= 0) return 0;
    if (i0 < 0 || i0 + n > sys->nlocal) {
        set_error(E_ARG, "surAdjointGFX950: range %d+%d outside %d atoms", i0, n, sys->nlocal);
        return -1;
    }
    k_sur_adjoint<<<blocks_for((long long)n * LPA, BLOCK), BLOCK, 0, stream()>>>(sys->nlocal, i0,
        n, sys->x, sys->border_map, sys->neighbors, sys->numneigh, mom, dF, scale, U);
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
    grow_n(&M.coef, &M.cap_coef, (size_t)nl * NCOEF + 4, 0);
    if (failed()) return 0.0;
    MdgSurSystem sys{ nl, atom->Nghost, atom->d_atom.x, atom->d_atom.border_map,
        neighbor->d_neighbor.neighbors, neighbor->d_neighbor.numneigh };
    surDescriptorsGFX950(&sys, 0, nl, M.dn, M.mom);
    surEvaluateGFX950(nl, M.dn, M.mom, nullptr, M.coef);
    surForcesGFX950(&sys, 0, nl, M.coef, atom->d_atom.fx);
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
