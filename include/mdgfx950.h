/*
 * libmdgfx950 -- MI355X (gfx950) short-range force step for MD-Bench.  C ABI.
 *
 * The Verlet-list entry points take MD-Bench's own structs (mdbench_vl.h) and keep its
 * signatures, so MD-Bench's host code binds them exactly where it binds its CUDA variants:
 * the function-pointer globals computeForce (src/verletlist/force.c:13-42), buildNeighbor
 * (src/verletlist/neighbor.c:23-27), updatePbc/updateAtomsPbc (src/verletlist/pbc.c:20-26) and
 * initialIntegrate/finalIntegrate (src/verletlist/integrate.c:13-19).  All state stays on the
 * GPU; the host copies are touched only by copyDataFromGFX950Device.
 *
 * Differences from the reference's conventions, by design:
 *   - no exit() inside the library: failures set a sticky status (mdgfx950Status) and print one
 *     line to stderr; entry points become no-ops until mdgfx950ClearStatus;
 *   - calls are asynchronous on the library's stream; the double returned by computeForce* is
 *     host seconds spent in the call (enqueue time).  Kernel time comes from
 *     mdgfx950KernelTime when timing is enabled;
 *   - like the reference's file-static device state, one simulation per process at a time.
 */
#ifndef MDGFX950_H
#define MDGFX950_H

#include <stdbool.h>
#include <stddef.h>

#include "mdbench_vl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status, mode, timing (new) */
enum {
    MDG_OK            = 0,
    MDG_ERR_HIP       = 1, /* a HIP runtime call failed */
    MDG_ERR_BIN       = 2, /* an atom fell outside the bin grid */
    MDG_ERR_CAPACITY  = 3, /* a device buffer could not be grown */
    MDG_ERR_ARG       = 4, /* invalid argument or call order */
    MDG_ERR_NODEVICE  = 5
};
int mdgfx950Status(void);
const char* mdgfx950StatusMessage(void);
void mdgfx950ClearStatus(void);
/* 1 = strict kernels (one lane per atom, reference summation order, no FMA): FP64 results
   bit-identical to the reference; 0 = fast kernels (default). */
void mdgfx950SetStrict(int strict);
void mdgfx950EnableTiming(int on);
/* Sum of kernel time (ms) and launch count of one kernel family since timing was enabled.
   Families: "force_lj", "force_eam", "neighbor", "binning", "pbc", "integrate", "thermo",
   "force_cp", "cluster", "surrogate". */
int mdgfx950KernelTime(const char* family, double* total_ms, long long* launches);
/* Restricts event recording to one family (NULL: all families, the default), so that timing a
   kernel adds no events around the others; returns -1 for an unknown family. */
int mdgfx950TimeOnly(const char* family);
/* Records the events on every `every`-th entry of a timed kernel family only (default 1: all),
   so that a long run is sampled at a fraction of the event overhead; launches counts the sampled
   entries. */
void mdgfx950TimeEvery(int every);
/* Wait for all queued work. */
void mdgfx950Synchronize(void);

/* ------------------------------------------------------------ Verlet-list scheme (drop-in) */
/* replaces initDevice (src/verletlist/device_spec.c:12-52) */
void initDeviceGFX950(Parameter* param, Atom* atom, Neighbor* neighbor);
/* releases everything initDeviceGFX950 and later calls allocated (no reference counterpart) */
void freeDeviceGFX950(Atom* atom, Neighbor* neighbor);
/* Copies a host-built Verlet list (the reference's CPU layout, neighbors[i * maxneighs + k],
   numneigh[i]) to the device list (memcpyToGPU of the list, src/common/device.h:30-36; used
   where lists do not come from buildNeighborGFX950, e.g. the synthetic-pattern benchmark). */
void uploadNeighborGFX950(Atom* atom, Neighbor* neighbor);
/* replace copyDataToCUDADevice / copyDataFromCUDADevice (src/verletlist/forceCuda.cu:327-345):
   x, v, type of Nlocal atoms; From also copies f */
void copyDataToGFX950Device(Parameter* param, Atom* atom);
void copyDataFromGFX950Device(Parameter* param, Atom* atom);
/* replaces setupNeighbor (src/verletlist/neighbor.c:101-258): bin grid and stencil */
void setupNeighborGFX950(Parameter* param);
/* replaces binatoms (src/verletlist/neighbor.c:431-464): counting sort, atoms in index order
   inside each bin */
void binatomsGFX950(Atom* atom);
/* replaces sortAtom (src/verletlist/neighbor.c:466-536), called under SORT_ATOMS at setup and
   every resort_every steps (src/verletlist/main.c:92-95, 372-381): the local atoms re-stored in
   bin order (bin-major, index order within a bin).  Sets Nghost = 0 as those callers do; ghosts
   and the neighbour list must be rebuilt before the next force call.  Unlike the reference, which
   moves positions and velocities only, each atom's force and type move with it, so that the kick
   that follows uses the atom's own force. */
void sortAtomGFX950(Atom* atom);
/* replaces buildNeighborCUDA (src/verletlist/neighborCuda.cu:259-366) / buildNeighborCPU
   (src/verletlist/neighbor.c:260-352); grows neighbor->maxneighs on overflow */
void buildNeighborGFX950(Atom* atom, Neighbor* neighbor);
/* replaces setupPbc (src/verletlist/pbc.c:103-232), on the device: same ghost order */
void setupPbcGFX950(Atom* atom, Parameter* param);
/* replace updatePbcCUDA / updateAtomsPbcCUDA (src/verletlist/pbcCuda.cu:78-142) */
void updatePbcGFX950(Atom* atom, Parameter* param, bool doReneighbor);
void updateAtomsPbcGFX950(Atom* atom, Parameter* param, bool doReneighbor);
/* replaces computeForceLJCUDA (src/verletlist/forceCuda.cu:227-303): full or half list by
   neighbor->half_neigh */
double computeForceLJGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats);
/* replaces computeForceEam (src/verletlist/force_eam.c:19-228); tables from setEamTablesGFX950 */
double computeForceEamGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats);
/* No reference counterpart: computeForce (LJ or EAM by param->force_field), finalIntegrate and
   the NEXT step's initialIntegrate (src/verletlist/main.c loop) in one call.  The fast full-list
   kernels kick each atom twice with the force they have just written and drift it into a second
   position buffer, which becomes atom->d_atom.x; the same FP64 operations in the same order as
   the three calls.  Ghost positions are stale until updatePbcGFX950, as after initialIntegrate.
   Velocities are left at the next half step: callers reading full-step velocities after a step
   (thermo, the last step) call the three separately there.  Strict kernels and half lists run
   the three calls in order. */
double computeForceFinalInitialGFX950(Parameter* param, Atom* atom, Neighbor* neighbor,
    Stats* stats);
/* uploads one pair table in the reference's 7-coefficient layout (src/common/eam.h:20-29,
   eam_utils.c:222-284); replaces the global `Eam eam` the reference kernel reads */
void setEamTablesGFX950(int nr, int nrho, double rdr, double rdrho, const double* rhor_spline,
    const double* frho_spline, const double* z2r_spline, int nr_tot, int nrho_tot);
/* No reference counterpart (the reference computes no EAM energy): while on, every
   computeForceEamGFX950 call also leaves per-atom energies F(rho_i) + 1/2 sum_j phi(r_ij) on the
   device, and downloadEamAtomEnergyGFX950 copies the Nlocal values to `out` (returns 0, or -1 with
   the status set).  Used to label training snapshots for the learned potential. */
void requestEamAtomEnergyGFX950(int on);
int downloadEamAtomEnergyGFX950(Atom* atom, double* out);
/* No reference counterpart: the learned EAM surrogate (mdgfx950_surrogate.h) as a computeForce
   (signature of src/verletlist/force.h:22-24): descriptors, network and forces of all local atoms
   into atom->d_atom.fx, with the model set by surSetModelGFX950.  Returns seconds spent. */
double computeForceSurrogateGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats);
/* The judge: with atom->d_atom.fx holding surrogate forces, computes the plain EAM forces of the
   same positions and returns the RMS difference per force component (eV/A).  At or below tol the
   surrogate forces are put back; above it the EAM forces stay and *fallback is set to 1. */
double surrogateJudgeGFX950(Parameter* param, Atom* atom, Neighbor* neighbor, Stats* stats,
    double tol, int* fallback);
/* replace initialIntegrateCUDA / finalIntegrateCUDA (src/verletlist/forceCuda.cu:192-225) */
void initialIntegrateGFX950(bool reneigh, Parameter* param, Atom* atom);
void finalIntegrateGFX950(bool reneigh, Parameter* param, Atom* atom);

/* ------------------------------------------------------------ additions (no counterpart) */
/* sum over local atoms of mass * |v|^2 (the reduction inside computeThermo,
   src/common/thermo.c:76-81), reduced on the device in FP64 */
double kineticSumGFX950(Parameter* param, Atom* atom);
/* potential energy and virial sum (sum r.F) of the last computeForce*GFX950 call, FP64 */
void energyVirialGFX950(double* epot, double* virial);
/* ask the next force calls to accumulate energy and virial (off by default) */
void requestEnergyVirialGFX950(int on);
/* device -> host copies for tests: positions of all Nlocal+Nghost atoms, ghost map, bins,
   neighbour list (as CSR), EAM fp */
int downloadVLStateGFX950(Atom* atom, Neighbor* neighbor, double* x, double* f, double* v,
    int* border_map, int* pbc, int* numneigh, int* neighbors_csr, long long csr_cap);
int downloadBinsGFX950(int* binstart, int mbins_cap, int* binatoms, int nall_cap);

#ifdef __cplusplus
}
#endif
#endif
