/*
 * MD-Bench Verlet-list data model, PRECISION=2 (double), ATOM_POSITION_AOS, LJ_COMB_RULE
 * geometric -- the layout the drop-in GFX950 entry points read and write.
 *
 * These structs must be layout-identical to MD-Bench's own so that its host code
 * (src/verletlist/main.c and friends) can hand its Parameter/Atom/Neighbor/Stats to the entry
 * points in mdgfx950.h unchanged.  tests/test_abi_layout.py compiles both and compares every
 * offset.  Reference definitions:
 *   Parameter        src/common/parameter.h:76-119
 *   Box              src/common/box.h:11-16
 *   DeviceAtom/Atom  src/verletlist/atom.h:16-54
 *   DeviceNeighbor/Neighbor  src/verletlist/neighbor.h:24-46
 *   Stats            src/verletlist/stats.h:11-16
 */
#ifndef MDBENCH_VL_H
#define MDBENCH_VL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double MD_FLOAT;

/* force_field values (src/verletlist/force.h:19) */
enum { FF_LJ = 0, FF_EAM = 1 };

typedef struct {
    int force_field;
    char* param_file;
    char* input_file;
    char* vtk_file;
    char* xtc_file;
    char* write_atom_file;
    char* types_file;
    MD_FLOAT epsilon;
    MD_FLOAT sigma;
    MD_FLOAT sigma6;
    MD_FLOAT temp;
    MD_FLOAT rho;
    MD_FLOAT mass;
    int ntypes;
    MD_FLOAT* epsilon_per_type;
    MD_FLOAT* sigma_per_type;
    int ntimes;
    int nstat;
    int reneigh_every;
    int resort_every;
    int prune_every;
    int x_out_every;
    int v_out_every;
    int half_neigh;
    MD_FLOAT dt;
    MD_FLOAT dtforce;
    MD_FLOAT skin;
    MD_FLOAT cutforce;
    MD_FLOAT cutneigh;
    int nx, ny, nz;
    int pbc_x, pbc_y, pbc_z;
    MD_FLOAT lattice;
    MD_FLOAT xlo, xhi, ylo, yhi, zlo, zhi;
    MD_FLOAT xprd, yprd, zprd;
    double proc_freq;
    int super_clustering;
    char* eam_file;
    int balance;
    int method;
    int balance_every;
    int setup;
} Parameter;

typedef struct {
    int id;
    MD_FLOAT xprd, yprd, zprd;
    MD_FLOAT lo[3];
    MD_FLOAT hi[3];
} Box;

/* Device mirrors of the per-atom arrays; x, vx, fx are AoS (3 per atom). */
typedef struct {
    MD_FLOAT *x, *y, *z;
    MD_FLOAT *vx, *vy, *vz;
    MD_FLOAT *fx, *fy, *fz;
    int* border_map;
    int* type;
    MD_FLOAT* epsilon;
    MD_FLOAT* sigma6;
    MD_FLOAT* cutforcesq;
    MD_FLOAT* cutneighsq;
    MD_FLOAT* sqrt_epsilon;
    MD_FLOAT* sigma3;
} DeviceAtom;

typedef struct {
    int Natoms, Nlocal, Nghost, Nmax;
    MD_FLOAT *x, *y, *z;
    MD_FLOAT *vx, *vy, *vz;
    MD_FLOAT *fx, *fy, *fz;
    int* border_map;
    int* type;
    int ntypes;
    MD_FLOAT* epsilon;
    MD_FLOAT* sigma6;
    MD_FLOAT* cutforcesq;
    MD_FLOAT* cutneighsq;
    MD_FLOAT* sqrt_epsilon;
    MD_FLOAT* sigma3;
    MD_FLOAT* sqrt_epsilon_per_type;
    MD_FLOAT* sigma3_per_type;
    DeviceAtom d_atom;
    Box mybox;
} Atom;

/* Device neighbour list: SoA, entry k of local atom i at neighbors[k * Nlocal + i]. */
typedef struct {
    int* neighbors;
    int* numneigh;
} DeviceNeighbor;

typedef struct {
    int every;
    int ncalls;
    int maxneighs;
    int half_neigh;
    int* neighbors;
    int* numneigh;
    DeviceNeighbor d_neighbor;
    int half_stencil;
    int Nshell;
    int* numNeighShell;
    int* neighshell;
    int* listshell;
} Neighbor;

typedef struct {
    long long int total_force_neighs;
    long long int total_force_iters;
    long long int atoms_within_cutoff;
    long long int atoms_outside_cutoff;
} Stats;

/* AoS accessors of the reference (src/verletlist/atom.h:78-89) */
#define atom_x(i)  atom->x[(i) * 3 + 0]
#define atom_y(i)  atom->x[(i) * 3 + 1]
#define atom_z(i)  atom->x[(i) * 3 + 2]
#define atom_vx(i) atom->vx[(i) * 3 + 0]
#define atom_vy(i) atom->vx[(i) * 3 + 1]
#define atom_vz(i) atom->vx[(i) * 3 + 2]
#define atom_fx(i) atom->fx[(i) * 3 + 0]
#define atom_fy(i) atom->fx[(i) * 3 + 1]
#define atom_fz(i) atom->fx[(i) * 3 + 2]

#ifdef __cplusplus
}
#endif
#endif
