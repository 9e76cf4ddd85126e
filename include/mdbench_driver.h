/*
 * Host driver of the gfx950 build -- the setup() and timestep loop of MD-Bench's main.c
 * (src/verletlist/main.c:46-113, 371-419) over the device entry points of mdgfx950.h.
 *
 * The driver owns the reference structs; tests and bench.py reach the drop-in entry points
 * through the Parameter/Atom/Neighbor pointers it hands out.
 */
#ifndef MDBENCH_DRIVER_H
#define MDBENCH_DRIVER_H

#include "mdbench_config.h"
#include "mdbench_vl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdb_vl mdb_vl;

/* setup(): parameters, lattice, thermo adjust, device upload, ghosts, neighbour list, first
   force.  Returns 0 or an MDG_ERR_* code. */
int mdbVLCreate(const mdb_config* cfg, mdb_vl** out);
/* nsteps of the timestep loop.  ksum (optional, nsteps entries) receives sum m v^2 after each
   step; asking for it synchronises every step. */
int mdbVLRun(mdb_vl* s, int nsteps, double* ksum);
void mdbVLDestroy(mdb_vl* s);
Parameter* mdbVLParam(mdb_vl* s);
Atom* mdbVLAtom(mdb_vl* s);
Neighbor* mdbVLNeighbor(mdb_vl* s);
Stats* mdbVLStats(mdb_vl* s);
/* Temperature and kinetic pressure from sum m v^2 (computeThermo, src/common/thermo.c:85-86). */
void mdbThermo(const mdb_vl* s, double ksum, double* temp, double* press);
/* Atoms after setup, before the device upload: x (3N), v (3N, adjusted), type (N). */
int mdbVLInitialState(const mdb_vl* s, double* x, double* v, int* type);
/* EAM runs only: forces from the learned surrogate (model set with surSetModelGFX950), judged
   against plain EAM every judge_every steps; above tol (RMS per force component, eV/A) the run
   falls back to plain EAM for good.  on = 0 returns to plain EAM. */
int mdbVLUseSurrogate(mdb_vl* s, int on, int judge_every, double tol);
/* out[5]: judges run, last RMS, largest RMS, step of the fallback (-1: none), fallback flag */
void mdbVLSurrogateStats(const mdb_vl* s, double* out);

/* MD-Bench's synthetic-pattern force benchmark (src/verletlist/main-stub.c) on the device:
   pattern 0 = seq, 1 = fix, 2 = rand; natoms on a line, nneighs per list replicated nreps times,
   ntimes LJ force calls (full or half list, strict or fast kernel).  out[5]: seconds, atom
   updates per second, ms per force call, f_x of the first and of the last atom. */
int mdbStubRun(int pattern, int natoms, int nneighs, int nreps, int ntimes, int half, int strict,
    double* out);

#ifdef __cplusplus
}
#endif
#endif
