/*
 * Cluster-pair driver: MD-Bench's clusterpair setup() and timestep loop
 * (src/clusterpair/main.c:43-135 setup/reneighbour/updateAtoms, :326-379 loop) with every
 * per-step operation and every rebuild on the device.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdbench_driver_cp.h"
#include "mdgfx950_cp.h"

/* shared with the Verlet-list side (mdgfx950.h cannot be included next to these structs) */
int mdgfx950Status(void);
const char* mdgfx950StatusMessage(void);
void mdgfx950SetStrict(int strict);

struct mdb_cp {
    Parameter param;
    Atom atom;
    Neighbor neighbor;
    Stats stats;
    double t_scale, dof_boltz, p_scale;
    int device_ready;
};

Parameter* mdbCPParam(mdb_cp* s) { return &s->param; }
Atom* mdbCPAtom(mdb_cp* s) { return &s->atom; }
Neighbor* mdbCPNeighbor(mdb_cp* s) { return &s->neighbor; }
Stats* mdbCPStats(mdb_cp* s) { return &s->stats; }

void mdbCPThermo(const mdb_cp* s, double ksum, double* temp, double* press)
{
    double t = ksum * s->t_scale;
    if (temp) *temp = t;
    if (press) *press = (t * s->dof_boltz) * s->p_scale;
}

static int reneighbour(mdb_cp* s)
{
    buildClustersGFX950(&s->atom);
    defineJClustersGFX950(&s->param, &s->atom);
    setupPbcClusterGFX950(&s->atom, &s->param);
    binJClustersGFX950(&s->param, &s->atom);
    buildNeighborClusterGFX950(&s->atom, &s->neighbor);
    return mdgfx950Status();
}

int mdbCPCreate(const mdb_config* cfg, mdb_cp** out)
{
    *out      = NULL;
    mdb_cp* s = calloc(1, sizeof(mdb_cp));
    if (!s) return 3;
    Parameter* p = &s->param;
    /* initParameter defaults (src/common/parameter.c:72-124) for the LJ solid */
    p->force_field   = FF_LJ;
    p->epsilon       = 1.0;
    p->sigma         = 1.0;
    p->sigma6        = 1.0;
    p->rho           = cfg->rho > 0 ? cfg->rho : 0.8442;
    p->ntypes        = cfg->ntypes > 0 ? cfg->ntypes : 4;
    p->ntimes        = cfg->ntimes;
    p->dt            = cfg->dt > 0 ? cfg->dt : 0.005;
    p->nx            = cfg->nx;
    p->ny            = cfg->ny;
    p->nz            = cfg->nz;
    p->pbc_x = p->pbc_y = p->pbc_z = 1;
    p->cutforce      = cfg->cutforce > 0 ? cfg->cutforce : 2.5;
    p->skin          = cfg->skin >= 0 ? cfg->skin : 0.3;
    p->cutneigh      = p->cutforce + p->skin;
    p->temp          = cfg->temp > 0 ? cfg->temp : 1.44;
    p->nstat         = cfg->nstat > 0 ? cfg->nstat : 100;
    p->mass          = 1.0;
    p->dtforce       = 0.5 * p->dt;
    p->reneigh_every = cfg->reneigh_every > 0 ? cfg->reneigh_every : 20;
    p->prune_every   = cfg->prune_every > 0 ? cfg->prune_every : 1000;
    p->half_neigh    = cfg->half_neigh;
    if (cfg->lattice_a > 0) p->rho = 4.0 / (cfg->lattice_a * cfg->lattice_a * cfg->lattice_a);
    p->lattice = pow((4.0 / p->rho), (1.0 / 3.0));
    p->xprd    = p->nx * p->lattice;
    p->yprd    = p->ny * p->lattice;
    p->zprd    = p->nz * p->lattice;

    /* atoms: the same fcc lattice, Park-Miller velocities and types as the Verlet-list setup */
    Atom* a   = &s->atom;
    int natom = 4 * p->nx * p->ny * p->nz;
    a->Natoms = a->Nlocal = a->Nmax = natom;
    a->ntypes = p->ntypes;
    a->x    = malloc(sizeof(double) * 3 * natom);
    a->vx   = malloc(sizeof(double) * natom);
    a->vy   = malloc(sizeof(double) * natom);
    a->vz   = malloc(sizeof(double) * natom);
    a->type = malloc(sizeof(int) * natom);
    double* v = malloc(sizeof(double) * 3 * natom);
    mdb_config lc = *cfg;
    lc.force_field = 0;
    lc.rho = p->rho;
    lc.lattice_a = 0;
    lc.temp = p->temp;
    lc.ntypes = p->ntypes;
    int n = mdbCreateLattice(&lc, a->x, v, a->type, natom);
    if (n != natom) {
        free(v);
        mdbCPDestroy(s);
        return 4;
    }
    for (int i = 0; i < natom; i++) {
        a->vx[i] = v[3 * i];
        a->vy[i] = v[3 * i + 1];
        a->vz[i] = v[3 * i + 2];
    }
    free(v);
    if (cfg->type1_fraction > 0.0) {
        /* seeded binary mixture: one Park-Miller stream in atom order */
        int seed = cfg->type_seed > 0 ? cfg->type_seed : 1;
        for (int i = 0; i < natom; i++)
            a->type[i] = mdbMyrandom(&seed) < cfg->type1_fraction ? 1 : 0;
    }
    /* type-pair tables: explicit pairs, or geometric mixing of per-type parameters
       (parameter.c:31-70) */
    int nt = a->ntypes;
    a->epsilon    = malloc(sizeof(double) * nt * nt);
    a->sigma6     = malloc(sizeof(double) * nt * nt);
    a->cutforcesq = malloc(sizeof(double) * nt * nt);
    a->cutneighsq = malloc(sizeof(double) * nt * nt);
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++) {
            double ei = (cfg->epsilon_per_type && i < cfg->ntypes) ? cfg->epsilon_per_type[i] : 1.0;
            double ej = (cfg->epsilon_per_type && j < cfg->ntypes) ? cfg->epsilon_per_type[j] : 1.0;
            double si = (cfg->sigma_per_type && i < cfg->ntypes) ? cfg->sigma_per_type[i] : 1.0;
            double sj = (cfg->sigma_per_type && j < cfg->ntypes) ? cfg->sigma_per_type[j] : 1.0;
            a->epsilon[i * nt + j]    = sqrt(ei) * sqrt(ej);
            a->sigma6[i * nt + j]     = (si * si * si) * (sj * sj * sj);
            a->cutforcesq[i * nt + j] = p->cutforce * p->cutforce;
            a->cutneighsq[i * nt + j] = p->cutneigh * p->cutneigh;
            if (cfg->pair_epsilon && cfg->pair_sigma) {
                double e = cfg->pair_epsilon[i * nt + j], sg = cfg->pair_sigma[i * nt + j];
                double c = cfg->pair_cutoff ? cfg->pair_cutoff[i * nt + j] : p->cutforce;
                a->epsilon[i * nt + j]    = e;
                a->sigma6[i * nt + j]     = sg * sg * sg * sg * sg * sg;
                a->cutforcesq[i * nt + j] = c * c;
            }
        }
    /* thermo scaling (thermo.c:35-58, LJ units) */
    s->dof_boltz = (natom * 3 - 3);
    s->t_scale   = 1.0 / s->dof_boltz;
    s->p_scale   = 1.0 / 3 / p->xprd / p->yprd / p->zprd;

    Neighbor* nb   = &s->neighbor;
    nb->maxneighs  = 200; /* clusterpair initNeighbor, neighbor.c:80-84 */
    nb->half_neigh = p->half_neigh;

    mdgfx950SetStrict(cfg->strict);
    int M = cfg->cluster_m > 0 ? cfg->cluster_m : 4, N = cfg->cluster_n > 0 ? cfg->cluster_n : 4;
    if (setClusterShapeGFX950(M, N)) {
        mdbCPDestroy(s);
        return mdgfx950Status() ? mdgfx950Status() : 4;
    }
    setClusterPrecisionGFX950(cfg->precision, cfg->compensated);
    initDeviceClusterGFX950(p, a, nb);
    s->device_ready = 1;
    setupNeighborClusterGFX950(p, a);
    int rc = reneighbour(s);
    if (!rc) computeForceLJClusterGFX950(p, a, nb, &s->stats);
    rc = mdgfx950Status();
    if (rc) {
        mdbCPDestroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int mdbCPSetAtoms(mdb_cp* s, const double* x, const double* v, const int* type, int n)
{
    Atom* a = &s->atom;
    if (n != a->Nlocal) return 4;
    memcpy(a->x, x, sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) {
        a->vx[i] = v[3 * i];
        a->vy[i] = v[3 * i + 1];
        a->vz[i] = v[3 * i + 2];
    }
    memcpy(a->type, type, sizeof(int) * n);
    copyAtomsToClusterGFX950Device(a);
    int rc = reneighbour(s);
    if (!rc) computeForceLJClusterGFX950(&s->param, a, &s->neighbor, &s->stats);
    return mdgfx950Status();
}

int mdbCPRun(mdb_cp* s, int nsteps, double* ksum)
{
    Parameter* p = &s->param;
    Atom* a      = &s->atom;
    Neighbor* nb = &s->neighbor;
    int integrated = 0; /* this step's initialIntegrate ran inside the last force call */
    for (int n = 0; n < nsteps; n++) {
        if (!integrated) initialIntegrateClusterGFX950(p, a);
        if ((n + 1) % p->reneigh_every) {
            if (!((n + 1) % p->prune_every)) pruneNeighborClusterGFX950(p, a, nb);
            updatePbcClusterGFX950(a, p, false);
        } else {
            updateSingleAtomsGFX950(p, a);
            updateAtomsPbcClusterGFX950(a, p, false);
            if (reneighbour(s)) return mdgfx950Status();
        }
        /* the next step's initialIntegrate rides with this force call unless this step's
           full-step velocities are read (after the last step, or every step when traced) or
           the next step prunes before its ghost update: the fused call leaves the ghost slots
           one step older than the separate calls do, and the prune reads them */
        const int m      = n + 2; /* the next step, counted as (n + 1) above */
        const int prunes = (m % p->reneigh_every) && !(m % p->prune_every);
        integrated       = n + 1 < nsteps && !ksum && !prunes;
        if (integrated) computeForceFinalInitialClusterGFX950(p, a, nb, &s->stats);
        else computeForceFinalClusterGFX950(p, a, nb, &s->stats);
        if (ksum) ksum[n] = kineticSumClusterGFX950(p, a);
        if (mdgfx950Status()) return mdgfx950Status();
    }
    return mdgfx950Status();
}

void mdbCPDestroy(mdb_cp* s)
{
    if (!s) return;
    if (s->device_ready) freeDeviceClusterGFX950();
    Atom* a = &s->atom;
    free(a->x);
    free(a->vx);
    free(a->vy);
    free(a->vz);
    free(a->type);
    free(a->epsilon);
    free(a->sigma6);
    free(a->cutforcesq);
    free(a->cutneighsq);
    free(s);
}
