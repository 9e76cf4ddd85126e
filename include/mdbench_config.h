/*
 * Run configuration and host-side setup helpers shared by the Verlet-list and cluster-pair
 * drivers (plain C types only, so either scheme's structs can be included next to it).
 */
#ifndef MDBENCH_CONFIG_H
#define MDBENCH_CONFIG_H

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
    /* cluster-pair scheme only */
    int cluster_m, cluster_n; /* i- and j-cluster sizes: 4x4 (default), 4x8, 1x2 */
    int precision;     /* 0 = FP32 pair math with FP64 state (default), 1 = FP64 */
    int compensated;   /* 1 = Kahan-compensated FP32 force sums */
    /* explicit type-pair tables (LJ_COMB_RULE none, src/common/parameter.h:60): ntypes^2
       epsilon, sigma and force cutoff per pair; NULL = geometric mixing of per-type values */
    const double* pair_epsilon;
    const double* pair_sigma;
    const double* pair_cutoff;
    /* > 0: types drawn per atom from a Park-Miller stream seeded with type_seed, type 1 with
       this probability, type 0 otherwise (replaces rand() % ntypes) */
    double type1_fraction;
    int type_seed;
    /* Verlet-list runs: atoms (positions, velocities, types and box) from MD-Bench's dump format
       instead of the lattice, as its -i option does (no thermo adjustment); NULL = lattice */
    const char* input_file;
    /* cluster-pair runs: prune the lists every this many steps between rebuilds (pruneNeighbor,
       src/clusterpair/main.c:329-332); 0 = the reference's default, 1000 */
    int prune_every;
    /* Verlet-list runs: re-sort the local atoms by bin at setup and every this many steps
       (SORT_ATOMS, src/verletlist/main.c:92-95, 372-381; the reference's default period is
       400); 0 = no sorting, the reference's default build (config.mk:29) */
    int resort_every;
    /* > 0: initial velocities drawn from the Maxwell-Boltzmann distribution (independent
       normal components from a generator seeded with this value), then centre-of-mass motion
       removed and rescaled to temp as adjustThermo does (src/common/thermo.c:102-159);
       0 = createAtom's uniform Park-Miller velocities (src/verletlist/atom.c:190-205) */
    int maxwell_seed;
} mdb_config;

/* Defaults of initParameter (src/common/parameter.c:72-124) with -1/0 "keep default" markers. */
void mdbDefaultConfig(mdb_config* cfg);

/* Host setup pieces, exported for tests (no device work). */
/* Parameters, fcc lattice and thermo adjustment of setup() only: fills x, v (3N) and type (N)
   of up to cap atoms; returns the atom count or a negative error. */
int mdbCreateLattice(const mdb_config* cfg, double* x, double* v, int* type, int cap);
double mdbMyrandom(int* seed); /* Park-Miller, src/common/util.c:24-33 */
/* n x 3 independent standard normal values (Box-Muller on a splitmix64 stream seeded with
   seed), the velocity draw behind mdb_config.maxwell_seed. */
void mdbMaxwellNormals(int n, unsigned long long seed, double* v);
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
/* MD-Bench's dump format (src/verletlist/atom.c:529-656).  box = xlo xhi ylo yhi zlo zhi.
   mdbWriteDmp writes 17 significant digits (exact round trip; 0 or < 0 on error); mdbReadDmp
   reads the first frame into arrays of cap atoms and returns the atom count (< 0 on error;
   -4: more than cap atoms), *ntypes = largest type + 1. */
int mdbWriteDmp(const char* path, int natoms, const double* x, const double* v, const int* type,
    const double* box, int timestep);
int mdbReadDmp(const char* path, int cap, double* x, double* v, int* type, double* box,
    int* ntypes);

#ifdef __cplusplus
}
#endif
#endif
