/*
 * MD-Bench cluster-pair data model, PRECISION=2 (double), ATOM_POSITION_AOS -- the layout the
 * cluster-pair GFX950 entry points (mdgfx950_cp.h) read and write.  Layout-identical to
 * MD-Bench's own (tests/test_abi_layout.py compares every offset):
 *   Cluster, SuperCluster, Atom   src/clusterpair/atom.h:38-90
 *   Neighbor                      src/clusterpair/neighbor.h:36-51
 *   Parameter, Box                src/common/parameter.h:76-119, src/common/box.h:11-16
 * Include either this header or mdbench_vl.h in one translation unit, never both (the two
 * schemes share type names, as they do in the reference, which builds them separately).
 */
#ifndef MDBENCH_CP_H
#define MDBENCH_CP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MDBENCH_COMMON_TYPES
#define MDBENCH_COMMON_TYPES
typedef double MD_FLOAT;
typedef unsigned long long int MD_UINT;
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

typedef struct {
    long long int calculated_forces;
    long long int num_neighs;
    long long int force_iters;
    long long int atoms_within_cutoff;
    long long int atoms_outside_cutoff;
    long long int clusters_within_cutoff;
    long long int clusters_outside_cutoff;
} Stats;
#endif

typedef struct {
    int natoms;
    MD_FLOAT bbminx, bbmaxx;
    MD_FLOAT bbminy, bbmaxy;
    MD_FLOAT bbminz, bbmaxz;
} Cluster;

typedef struct {
    int nclusters;
    MD_FLOAT bbminx, bbmaxx;
    MD_FLOAT bbminy, bbmaxy;
    MD_FLOAT bbminz, bbmaxz;
} SuperCluster;

typedef struct {
    int Natoms, Nlocal, Nghost, Nmax;
    int Nclusters_local, Nclusters_ghost, Nclusters_max, NmaxGhost, ncj;
    MD_FLOAT *x, *y, *z;
    MD_FLOAT *vx, *vy, *vz;
    int* border_map;
    int* type;
    int ntypes;
    MD_FLOAT* epsilon;
    MD_FLOAT* sigma6;
    MD_FLOAT* cutforcesq;
    MD_FLOAT* cutneighsq;
    MD_FLOAT* sqrt_epsilon_per_type;
    MD_FLOAT* sigma3_per_type;
    int *PBCx, *PBCy, *PBCz;
    MD_FLOAT* cl_x;
    MD_FLOAT* cl_v;
    MD_FLOAT* cl_f;
    int* cl_t;
    MD_FLOAT* cl_sqrt_epsilon;
    MD_FLOAT* cl_sigma3;
    Cluster *iclusters, *jclusters;
    SuperCluster* siclusters;
    int* cluster_bin;
    int dummy_cj;
    MD_UINT* exclusion_filter;
    MD_FLOAT* diagonal_4xn_j_minus_i;
    MD_FLOAT* diagonal_2xnn_j_minus_i;
    unsigned int masks_2xnn_hn[8];
    unsigned int masks_2xnn_fn[8];
    unsigned int masks_4xn_hn[16];
    unsigned int masks_4xn_fn[16];
    Box mybox;
} Atom;

typedef struct {
    int every;
    int ncalls;
    int maxneighs;
    int* numneigh;
    int* numneigh_masked;
    int half_neigh;
    int* neighbors;
    unsigned int* neighbors_imask;
    int Nshell;
    int* numNeighShell;
    int* neighshell;
    int* listshell;
} Neighbor;

#ifdef __cplusplus
}
#endif
#endif
