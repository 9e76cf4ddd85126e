/*
 * Host driver of the gfx950 build -- the setup() and timestep loop of MD-Bench's main.c
 * (src/verletlist/main.c:46-113, 371-419) over the device entry points of mdgfx950.h.
 *
 * The driver owns the reference structs; tests and bench.py reach the drop-in entry points
 * through the Parameter/Atom/Neighbor pointers it hands out.
 */
#ifndef MDBENCH_DRIVER_H
#define MDBENCH_DRIVER_H

#include "mdbench_vl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    int force_field;   /* FF_LJ or FF_EAM */
    int half_neigh;
    int ntimes, nstat, reneigh_every;
    double rho;        /* > 0 overrides the reduced density */
    double lattice_a;  /* > 0 overrides the fcc lattice constant (rho = 4 / a^3) */
    double temp;       /* > 0 overrides the initial temperature */
    double dt;         /* > 0 overrides the time step */
    double cutforce;   /* > 0 overrides the force cutoff */
    double skin;       /* >= 0 overrides the skin (use -1 for the default) */
    int ntypes;        /* > 0 overrides the number of types */
    const double* epsilon_per_type; /* optional, ntypes values (geometric mixing) */
    const double* sigma_per_type;
    const char* eam_file; /* FF_EAM: funcfl file; NULL = synthetic Cu-like tables */
    int strict;        /* 1 = strict kernels (bit-identical FP64) */
} mdb_config;

/* Defaults of initParameter (src/common/parameter.c:72-124) with -1/0 "keep default" markers. */
void mdbDefaultConfig(mdb_config* cfg);

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

/* Host setup pieces, exported for tests (no device work). */
/* Parameters, fcc lattice and thermo adjustment of setup() only: fills x, v (3N) and type (N)
   of up to cap atoms; returns the atom count or a negative error. */
int mdbCreateLattice(const mdb_config* cfg, double* x, double* v, int* type, int cap);
double mdbMyrandom(int* seed); /* Park-Miller, src/common/util.c:24-33 */
/* funcfl tables -> 7-coefficient splines (src/common/eam_utils.c:42-284), one element. */
typedef struct {
    int nr, nrho, nr_tot, nrho_tot;
    double dr, drho, rdr, rdrho, cut, mass;
    double *rhor_spline, *frho_spline, *z2r_spline;
} mdb_eam_tables;
int mdbEamLoad(const char* path, mdb_eam_tables* t);
void mdbEamFree(mdb_eam_tables* t);
/* Writes the synthetic Cu-like funcfl file (rho = exp(-3(r/2.556-1)) with a smooth cutoff,
   Morse-like Z, F = -sqrt(rho); nr = nrho = 2000, cutoff 4.95). */
int mdbWriteFuncflSynthetic(const char* path);

#ifdef __cplusplus
}
#endif
#endif
