/* Learned EAM surrogate on gfx950: per-atom descriptors (24 Gaussian radial + 8 angular) into an
 * MLP 32 -> 64 -> 64 -> 1 (bfloat16 MFMA, FP32 accumulation), analytic forces, and the training
 * kernels (force adjoint, weight gradients, Adam).
 *
 * MD-Bench has no learned potential: none of these entry points replaces a reference function.
 * computeForceSurrogateGFX950 has the signature of the reference's computeForce pointer
 * (src/verletlist/force.h:22-24) so a driver can bind it where it binds computeForceEam; the
 * plain EAM kernel (computeForceEamGFX950) stays its fallback and judge.
 *
 * Pointers marked (device) are device memory (mdgDeviceAlloc or the library's own arrays); all
 * work runs on the library's stream.  Errors set the library status (mdgfx950Status).
 */
#ifndef MDGFX950_SURROGATE_H
#define MDGFX950_SURROGATE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDG_SUR_ND 32      /* descriptors per atom */
#define MDG_SUR_NMOM 36    /* angular moments per atom: 4 x A (3), 4 x traceless T (6) */
#define MDG_SUR_NCOEF 64   /* force coefficients per atom */
#define MDG_SUR_NPARAM 6337 /* W1[64x32] b1[64] W2[64x64] b2[64] w3[64] b3 */

/* A system the kernels read: nlocal atoms with Verlet lists (SoA, neighbors[k * nlocal + i]),
   nghost ghost images after them in x, ghost g an image of local border_map[g].  The fields are
   the reference's own (Atom.d_atom.x, Atom.d_atom.border_map, Neighbor.d_neighbor.*). */
typedef struct {
    int nlocal;
    int nghost;
    const double* x;          /* (device) 3 * (nlocal + nghost) */
    const int* border_map;    /* (device) nghost */
    const int* neighbors;     /* (device) maxneighs * nlocal */
    const int* numneigh;      /* (device) nlocal */
} MdgSurSystem;

/* Device memory helpers (plain hipMalloc / hipFree / synchronous copies on the library stream). */
void* mdgDeviceAlloc(size_t bytes);
void mdgDeviceFree(void* p);
int mdgCopyToDevice(void* dst, const void* src, size_t bytes);
int mdgCopyFromDevice(void* dst, const void* src, size_t bytes);
int mdgDeviceZero(void* p, size_t bytes);

/* Model: theta (host, MDG_SUR_NPARAM floats in the order above), descriptor normalisation
   (host, 32 means and 32 standard deviations), energy offset and scale:
   E_i = s_e * (w3 . h2 + b3) + e0. */
int surSetModelGFX950(const float* theta, const float* d_mean, const float* d_std, float e0,
    float s_e);
/* Same, with theta already on the device (training updates it in place). */
int surSetParamsDeviceGFX950(const float* theta_dev);
int surGetParamsGFX950(float* theta_host);

/* Descriptors of atoms i0 .. i0+n-1: dn (device, n x 32) normalised with the model's
   statistics (pass identity statistics to get raw values), mom (device, n x 36: per angular
   channel A (3) and the traceless T (6)), and the list of pairs inside the descriptor cutoff
   (device: pairs[k * n + a], k < npairs[a] <= cap; cap >= the longest Verlet list). */
int surDescriptorsGFX950(const MdgSurSystem* sys, int i0, int n, float* dn, float* mom, int cap,
    int* pairs, int* npairs);
/* MLP forward and input gradient (bfloat16 MFMA): per-atom energies e (device, n, may be NULL)
   and force coefficients coef (device, n x 64 IEEE halves) from dn and mom. */
int surEvaluateGFX950(int n, const float* dn, const float* mom, float* e, void* coef);
/* Forces of atoms i0 .. i0+n-1 into f (device, n x 3 doubles) from coef (rows i - i0; a ghost
   neighbour uses its owner's row, which must lie in the same range: pairs whose row falls
   outside it are skipped) over the pair list of surDescriptorsGFX950. */
int surForcesGFX950(const MdgSurSystem* sys, int i0, int n, const int* pairs, const int* npairs,
    const void* coef, double* f);

/* Short-range core pair term U(r) = k_core (r_core - r)^3 for r < r_core (0 beyond), added to
   the network's energy: it keeps atoms apart below the closest approach the training data
   holds, where the network has no data.  surSetCoreGFX950 sets it for surForcesGFX950 and the
   drop-in force call (default k_core = 0: no core); surCoreGFX950 computes its per-atom
   energies e (device, n) and forces f (device, n x 3) alone, in FP64, over the Verlet list
   (training labels have it subtracted).  e or f may be NULL. */
int surSetCoreGFX950(float r_core, float k_core);
int surCoreGFX950(const MdgSurSystem* sys, int i0, int n, float r_core, float k_core, double* e,
    double* f);

/* The drop-in force call computeForceSurrogateGFX950 and the judge surrogateJudgeGFX950 work on
   the Verlet-list structs and are declared in mdgfx950.h. */

/* Training.  Residual terms of the loss wE mean (E - E_ref)^2 + wF mean |F - F_ref|^2 over
   n_norm atoms: cE = dL/dE (device, n), dF = F - F_ref (device, n x 3 floats); sums (host, 2)
   receives sum (E - E_ref)^2 and sum |F - F_ref|^2 of this range. */
int surLossTermsGFX950(int n, const float* e, const double* f, const double* e_ref,
    const double* f_ref, float wE, int n_norm, float* cE, float* dF, double* sums);
/* U = scale * 2 sum_j J_ij^T (dF_i - dF_owner(j)) (device, n x 32): dL/dg for the force term. */
int surAdjointGFX950(const MdgSurSystem* sys, int i0, int n, const int* pairs, const int* npairs,
    const float* mom, const float* dF, float scale, float* U);
/* grad (device, MDG_SUR_NPARAM) += d/dtheta sum_i [cE_i E_i + U_i . dE_i/dD_i] over n atoms
   (FP32, exact network arithmetic). */
int surWeightGradGFX950(int n, const float* dn, const float* U, const float* cE, float* grad);
/* One Adam step on theta (device) with gradient grad (device); m, v (device) are the moment
   estimates, t counts from 1. */
int surAdamGFX950(float* theta, const float* grad, float* m, float* v, int t, float lr,
    float beta1, float beta2, float eps);
/* sqrt(mean |f_a - f_b|^2) over n atoms' force components (device, n x 3 doubles each). */
double surForceRmseGFX950(int n, const double* fa, const double* fb);

/* Self-check of the matrix-core operand layout the network kernel assumes: C (host, 16 x 16,
   row-major) = A (host, 16 x 32, row-major) x B (host, 32 x 16, row-major), with A and B
   rounded to bfloat16 and one v_mfma_f32_16x16x32_bf16. */
int surMfmaProbeGFX950(const float* A, const float* B, float* C);

#ifdef __cplusplus
}
#endif

#endif
