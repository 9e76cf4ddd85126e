/*
 * Synthetic-pattern force benchmark: MD-Bench's main-stub (src/verletlist/main-stub.c:31-110,
 * 224-275) on the device.  Atoms lie on a line (x = y = z = i * 1e-5), the cutoff is so large
 * that every listed pair interacts, and the lists follow a pattern:
 *   seq   atom i: i+1, i+2, ... (mod N)       - contiguous gathers
 *   fix   every atom: 0, 1, ..., nneighs-1    - the same few atoms, cache resident
 *   rand  distinct random atoms (glibc rand)  - scattered gathers
 * each list replicated nreps times.  Only the LJ force kernel runs, ntimes times; the result is
 * the force kernel's gather rate in isolation from list quality.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

enum { P_SEQ = 0, P_FIX = 1, P_RAND = 2 };

static void stub_parameter(Parameter* p)
{
    mdb_init_parameter(p);
    p->input_file  = NULL;
    p->force_field = FF_LJ;
    p->epsilon     = 1.0;
    p->sigma6      = 1.0;
    p->rho         = 0.8442;
    p->ntypes      = 4;
    p->nx = p->ny = p->nz = 1;
    p->lattice  = 1.0;
    p->cutforce = 1000000.0;
    p->cutneigh = p->cutforce;
    p->mass     = 1.0;
    p->dt       = 0.005;
    p->dtforce  = 0.5 * p->dt;
}

int mdbStubRun(int pattern, int natoms, int nneighs, int nreps, int ntimes, int half, int strict,
    double* out)
{
    if (pattern < P_SEQ || pattern > P_RAND || natoms <= 0 || nneighs <= 0 || nreps <= 0)
        return MDG_ERR_ARG;
    if (pattern == P_RAND && natoms <= nneighs) return MDG_ERR_ARG;
    /* every listed index must name an atom: the fixed pattern lists atoms 0..nneighs-1 */
    if (pattern == P_FIX && natoms < nneighs) return MDG_ERR_ARG;
    Parameter p;
    Atom a;
    Neighbor nb;
    Stats st;
    memset(&a, 0, sizeof(a));
    memset(&nb, 0, sizeof(nb));
    memset(&st, 0, sizeof(st));
    stub_parameter(&p);
    p.half_neigh = half;
    mdgfx950SetStrict(strict);
    /* atoms and type tables (main-stub.c:213-236); types from a fresh glibc rand() */
    srand(1);
    a.ntypes = p.ntypes;
    a.Natoms = a.Nlocal = a.Nmax = natoms;
    const int nt2 = a.ntypes * a.ntypes;
    a.epsilon    = malloc(sizeof(double) * nt2);
    a.sigma6     = malloc(sizeof(double) * nt2);
    a.cutforcesq = malloc(sizeof(double) * nt2);
    a.cutneighsq = malloc(sizeof(double) * nt2);
    a.x          = calloc(3 * (size_t)natoms, sizeof(double));
    a.vx         = calloc(3 * (size_t)natoms, sizeof(double));
    a.fx         = calloc(3 * (size_t)natoms, sizeof(double));
    a.type       = calloc((size_t)natoms, sizeof(int));
    const int maxn = nneighs * nreps;
    nb.maxneighs   = maxn;
    nb.half_neigh  = half;
    nb.numneigh    = malloc(sizeof(int) * (size_t)natoms);
    nb.neighbors   = malloc(sizeof(int) * (size_t)natoms * maxn);
    int rc         = 0;
    if (!a.epsilon || !a.sigma6 || !a.cutforcesq || !a.cutneighsq || !a.x || !a.vx || !a.fx ||
        !a.type || !nb.numneigh || !nb.neighbors) {
        rc = MDG_ERR_CAPACITY;
        goto done;
    }
    for (int k = 0; k < nt2; k++) {
        a.epsilon[k]    = p.epsilon;
        a.sigma6[k]     = p.sigma6;
        a.cutneighsq[k] = p.cutneigh * p.cutneigh;
        a.cutforcesq[k] = p.cutforce * p.cutforce;
    }
    for (int i = 0; i < natoms; i++) {
        a.type[i]      = rand() % a.ntypes;
        a.x[3 * i]     = (double)i * 0.00001;
        a.x[3 * i + 1] = (double)i * 0.00001;
        a.x[3 * i + 2] = (double)i * 0.00001;
    }
    /* createNeighbors (main-stub.c:60-110).  main-stub.c:74 starts the last atom's sequential
     * list at index Nlocal, one past the last atom, which reads whatever lies beyond the position
     * array; the start wraps to atom 0 here, so every listed index names an atom. */
    for (int i = 0; i < natoms; i++) {
        int* li = nb.neighbors + (size_t)i * maxn;
        int j   = (pattern == P_SEQ) ? (i + 1) % natoms : 0;
        int m   = (pattern == P_SEQ) ? natoms : nneighs;
        for (int k = 0; k < nneighs; k++) {
            if (pattern == P_RAND) {
                int found;
                do {
                    j     = rand() % natoms;
                    li[k] = j;
                    found = (i == j);
                    for (int l = 0; l < k; l++)
                        if (li[l] == j) found = 1;
                } while (found);
            } else {
                li[k] = j;
                j     = (j + 1) % m;
            }
        }
        for (int r = 1; r < nreps; r++)
            memcpy(li + r * nneighs, li, sizeof(int) * nneighs);
        nb.numneigh[i] = maxn;
    }
    initDeviceGFX950(&p, &a, &nb);
    uploadNeighborGFX950(&a, &nb);
    computeForceLJGFX950(&p, &a, &nb, &st); /* warm-up */
    mdgfx950Synchronize();
    double t0 = mdb_wtime();
    for (int n = 0; n < ntimes; n++)
        computeForceLJGFX950(&p, &a, &nb, &st);
    mdgfx950Synchronize();
    double t = mdb_wtime() - t0;
    if (mdgfx950Status()) {
        rc = mdgfx950Status();
    } else if (out) {
        out[0] = t;
        out[1] = (double)natoms * ntimes / t;
        out[2] = 1e3 * t / ntimes;
        /* forces of the last call, for checks */
        copyDataFromGFX950Device(&p, &a);
        out[3] = a.fx[0];
        out[4] = a.fx[3 * (natoms - 1)];
    }
    freeDeviceGFX950(&a, &nb);
done:
    free(a.epsilon);
    free(a.sigma6);
    free(a.cutforcesq);
    free(a.cutneighsq);
    free(a.x);
    free(a.vx);
    free(a.fx);
    free(a.type);
    free(nb.numneigh);
    free(nb.neighbors);
    return rc;
}
