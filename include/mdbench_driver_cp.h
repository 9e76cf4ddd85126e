/*
 * Host driver of the gfx950 cluster-pair build: MD-Bench's clusterpair setup() and timestep
 * loop (src/clusterpair/main.c:43-135, 326-379) over the entry points of mdgfx950_cp.h.
 */
#ifndef MDBENCH_DRIVER_CP_H
#define MDBENCH_DRIVER_CP_H

#include "mdbench_config.h"
#include "mdbench_cp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdb_cp mdb_cp;

/* setup(): lattice, thermo adjust, device upload, clusters, ghosts, lists, first force. */
int mdbCPCreate(const mdb_config* cfg, mdb_cp** out);
/* nsteps of the loop; ksum (optional) receives sum m v^2 after each step (synchronises). */
int mdbCPRun(mdb_cp* s, int nsteps, double* ksum);
/* Replace the atoms (x AoS, v AoS, type; n must equal Natoms) and rebuild clusters, ghosts,
   bins and lists from them without wrapping -- the reneighbour of main.c:99-118. */
int mdbCPSetAtoms(mdb_cp* s, const double* x, const double* v, const int* type, int n);
void mdbCPDestroy(mdb_cp* s);
Parameter* mdbCPParam(mdb_cp* s);
Atom* mdbCPAtom(mdb_cp* s);
Neighbor* mdbCPNeighbor(mdb_cp* s);
Stats* mdbCPStats(mdb_cp* s);
void mdbCPThermo(const mdb_cp* s, double ksum, double* temp, double* press);

#ifdef __cplusplus
}
#endif
#endif
