/*
 * libmdgfx950 -- cluster-pair (GROMACS NBNXM-style) entry points.  C ABI over MD-Bench's own
 * cluster-pair structs (mdbench_cp.h).  They bind where MD-Bench's cluster-pair build binds its
 * CUDA path: computeForce (src/clusterpair/force.c:13-42), buildNeighbor / buildClusters
 * (src/clusterpair/neighbor.c:98-105), updatePbc / updateAtomsPbc (src/clusterpair/pbc.c:20-28),
 * initialIntegrate / finalIntegrate (src/clusterpair/integrate.c:15-27).
 *
 * Cluster shapes: i-clusters of M atoms against j-clusters of N atoms, chosen at run time
 * (reference: compile time, src/clusterpair/force.h:68-123): 4x4 and 4x8 for both force
 * kernels, 1x2 for the FP64 kernel (the shape of MD-Bench's plain-C reference kernel).
 *
 * Device layout: atom "slots" in cluster order, i-cluster ci owns slots [ci*M, ci*M+M) and
 * j-cluster cj owns [cj*N, cj*N+N); ghost j-clusters follow the local ones, then one dummy.
 * Positions, velocities: FP64 AoS (x,y,z) per slot.  Empty slots hold distinct far-away
 * positions (never within any cutoff).  Lists: entry k of i-cluster ci at ci*maxneighs + k,
 * entries whose j-cluster holds part of ci first (numneigh_masked), as in the reference.
 *
 * Status, timing and strict mode are shared with the Verlet-list entry points (mdgfx950.h).
 */
#ifndef MDGFX950_CP_H
#define MDGFX950_CP_H

#include <stdbool.h>

#include "mdbench_cp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Cluster shape and force-kernel precision, before initDeviceClusterGFX950 (no reference
   counterpart: compile-time CLUSTER_M/CLUSTER_N there).  precision 0 = FP32 pair math with FP64
   integrator state (default), 1 = FP64 (reference arithmetic, computeForceLJRef semantics).
   compensated 1 = Kahan-compensated FP32 per-lane force sums. */
int setClusterShapeGFX950(int M, int N);
void setClusterPrecisionGFX950(int precision, int compensated);

/* replaces initDevice + copyDataToCUDADevice (src/clusterpair/force_lj_cuda.cu:613-700):
   uploads the Nlocal atoms (x AoS, vx/vy/vz, type) and the type-pair tables */
void initDeviceClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor);
void freeDeviceClusterGFX950(void);
/* atoms (not clusters) host <-> device: x AoS, vx, vy, vz, type of Nlocal atoms */
void copyAtomsToClusterGFX950Device(Atom* atom);
void copyAtomsFromClusterGFX950Device(Atom* atom);
/* replaces setupNeighbor (src/clusterpair/neighbor.c:108-215): 2D bins, stencil */
void setupNeighborClusterGFX950(Parameter* param, Atom* atom);
/* replaces buildClustersCPU (neighbor.c:1513-1611) with binAtoms and sortAtomsByZCoord
   (neighbor.c:1409-1481): same cluster membership and order.  The cluster count stays on the
   device until setupPbcClusterGFX950 reads it with the ghost totals and sets
   atom->Nclusters_local and atom->ncj (MD-Bench's setup and reneighbour,
   src/clusterpair/main.c:85-93 and 106-115, read neither in between); the other entry points
   read a pending count back themselves */
void buildClustersGFX950(Atom* atom);
/* replaces defineJClusters (neighbor.c:1807-1927) */
void defineJClustersGFX950(Parameter* param, Atom* atom);
/* replaces setupPbc (src/clusterpair/pbc.c:188-331): ghost j-clusters and the dummy */
void setupPbcClusterGFX950(Atom* atom, Parameter* param);
/* replaces binJClusters (neighbor.c:1929-2094) */
void binJClustersGFX950(Parameter* param, Atom* atom);
/* replaces buildNeighborCPU (neighbor.c:282-766); grows neighbor->maxneighs on overflow */
void buildNeighborClusterGFX950(Atom* atom, Neighbor* neighbor);
/* replaces pruneNeighborCPU (neighbor.c:1099-1264; PruneNeighborFunction, neighbor.h:57):
   drops list entries with no atom pair inside cutneigh, in the reference's order (an entry out
   of range is overwritten by the current last entry); main.c:329-332 calls it every
   prune_every steps between rebuilds, before the ghost update */
void pruneNeighborClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor);
/* replace updatePbcCUDA / updatePbcCPU (force_lj_cuda.cu:583-611, pbc.c:43-116) and
   updateAtomsPbcCPU (pbc.c:120-145) */
void updatePbcClusterGFX950(Atom* atom, Parameter* param, bool firstUpdate);
void updateAtomsPbcClusterGFX950(Atom* atom, Parameter* param, bool dummy);
/* replaces updateSingleAtoms (neighbor.c:2096-2150): clusters -> atoms in cluster order.
   This and updateAtomsPbcClusterGFX950 are deferred: the next buildClustersGFX950 runs both in
   its first kernel; any other call that reads or writes the atoms or the slots first runs what
   is pending, so every call sees the state the two would have left. */
void updateSingleAtomsGFX950(Parameter* param, Atom* atom);
/* replaces computeForceLJCuda (force_lj_cuda.cu:603-700) and computeForceLJRef /
   computeForceLJ4xn{Full,Half}Neigh (src/clusterpair/force_lj.c): full or half list by
   neighbor->half_neigh, precision by setClusterPrecisionGFX950 */
double computeForceLJClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats);
/* computeForce then finalIntegrate (main.c:349-350) in one call: the FP32 full-list kernel
   applies the final velocity kick to each atom whose total force it has just written (one pass
   fewer over v and f); other configurations run the two calls in order (no reference
   counterpart: a fused form of the two reference calls) */
double computeForceFinalClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats);
/* computeForce, finalIntegrate and the NEXT step's initialIntegrate (main.c:349-350, 326) in one
   call: the FP32 full-list kernel kicks each atom twice with the force it has just written and
   drifts it into a second position buffer, which then becomes the current one; the same FP64
   operations in the same order as the three calls, so the trajectory is bit-identical.  The
   velocities it leaves are the next half step's: callers that read full-step velocities after a
   step (thermo, the last step) use computeForceFinalClusterGFX950 there.  Ghost slots keep
   positions one step older than after the separate calls until updatePbcClusterGFX950, so a
   pruneNeighborClusterGFX950 between the two differs; callers that prune there use the separate
   calls for the step before.  Other configurations
   run the calls in order (no reference counterpart: a fused form of three reference calls) */
double computeForceFinalInitialClusterGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats);
/* replace cudaInitialIntegrate / cudaFinalIntegrate (force_lj_cuda.cu:456-542; this one uses
   dtforce in the final kick, where the reference's CUDA final kick uses dt) */
void initialIntegrateClusterGFX950(Parameter* param, Atom* atom);
void finalIntegrateClusterGFX950(Parameter* param, Atom* atom);

/* ------------------------------------------------------------ additions */
double kineticSumClusterGFX950(Parameter* param, Atom* atom);
/* No counterpart in the reference.  kineticSumClusterGFX950 for callers that cannot allocate
   device memory or wait on the device (code under HIP graph capture, tight serving loops): it
   allocates nothing, copies nothing to the host and does not synchronise; it queues the sum
   (sum over local atoms of mass * |v|^2, the same kernels and summation order as
   kineticSumClusterGFX950, so the same value) on the library stream and writes it to d_out[0],
   device memory the caller owns.  Its 512 partial sums live in a fixed 4 KiB device array that
   each GPU holds until the process ends (any system size: the kernel strides over the slots);
   its address is resolved per device in initDeviceClusterGFX950.  Calls that overlap on one
   GPU share that array and race.  Returns 0, or the status code (MDG_ERR_ARG for a NULL
   d_out or a device not set up). */
int kineticSumClusterNoAllocGFX950(Parameter* param, Atom* atom, double* d_out);
void requestEnergyVirialClusterGFX950(int on);
void energyVirialClusterGFX950(double* epot, double* virial);
/* Copies a named device array to host; returns element count, -1 unknown name, -2 too small.
   Names: slot_x slot_v (double, 3/slot), slot_f (double, 3/local slot), slot_t, ci_natoms,
   ci_bin, ci_bbox (6/ci), cj_natoms, cj_bbox, ghost_map, ghost_pbc, numneigh,
   numneigh_masked, neighbors (CSR), cbin_start, cbin_list, atoms_x, atoms_v, atoms_type. */
long long downloadClusterArrayGFX950(const char* name, void* out, long long cap);
/* Overwrites the first n values of a named device array from host (test hook; names: slot_x,
   slot_v).  Returns n, -1 unknown name, -2 n larger than the array.  The lists and the cluster
   bounding boxes keep their last build's values (the one-type FP32 kernels take positions
   relative to the i-cluster's box centre, so positions far from their build-time boxes lose
   FP32 precision until the next rebuild). */
long long uploadClusterArrayGFX950(const char* name, const void* in, long long n);
/* nci ncj nghost_clusters dummy_cj nghost_atoms mbins nstencil maxneighs M N natoms */
void clusterInfoGFX950(int* info);

#ifdef __cplusplus
}
#endif
#endif
