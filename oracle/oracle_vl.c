/*
 * CPU oracle, Verlet-list path -- TEST INFRASTRUCTURE ONLY (see oracle_md.h).
 * Strict IEEE double arithmetic; every expression keeps the reference's evaluation order so
 * that results are bit-identical to MD-Bench compiled without FMA contraction.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_md.h"

#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))

/* bindist: src/verletlist/neighbor.c:355-384 */
static double vl_bindist(const orc_vl_grid* g, int i, int j, int k)
{
    double delx, dely, delz;
    if (i > 0) delx = (i - 1) * g->binsizex;
    else if (i == 0) delx = 0.0;
    else delx = (i + 1) * g->binsizex;
    if (j > 0) dely = (j - 1) * g->binsizey;
    else if (j == 0) dely = 0.0;
    else dely = (j + 1) * g->binsizey;
    if (k > 0) delz = (k - 1) * g->binsizez;
    else if (k == 0) delz = 0.0;
    else delz = (k + 1) * g->binsizez;
    return delx * delx + dely * dely + delz * delz;
}

/* initNeighbor (neighbor.c:60-99, nbin = (int)(5/6 n)) + setupNeighbor for generated lattices
   (neighbor.c:101-258, input_file == NULL branch, full stencil). */
static int grid_tail(orc_vl_grid* g, double cutneigh);

int orc_vl_setup_grid(orc_vl_grid* g, int nx, int ny, int nz, double lattice, double cutneigh)
{
    const double neighscale = 5.0 / 6.0;
    memset(g, 0, sizeof(*g));
    g->xprd       = nx * lattice;
    g->yprd       = ny * lattice;
    g->zprd       = nz * lattice;
    g->cutneigh   = cutneigh;
    g->cutneighsq = cutneigh * cutneigh;
    g->nbinx      = (int)ORC_MAX(1, neighscale * nx);
    g->nbiny      = (int)ORC_MAX(1, neighscale * ny);
    g->nbinz      = (int)ORC_MAX(1, neighscale * nz);
    g->binsizex   = g->xprd / g->nbinx;
    g->binsizey   = g->yprd / g->nbiny;
    g->binsizez   = g->zprd / g->nbinz;
    g->bininvx    = 1.0 / g->binsizex;
    g->bininvy    = 1.0 / g->binsizey;
    g->bininvz    = 1.0 / g->binsizez;
    return grid_tail(g, cutneigh);
}

/* setupNeighbor with an input file (neighbor.c:107-140): bins of cutneigh / 2 over the file's
   box, bininv = nbin / length. */
int orc_vl_setup_grid_box(orc_vl_grid* g, const double* box, double cutneigh)
{
    memset(g, 0, sizeof(*g));
    g->xprd       = box[1] - box[0];
    g->yprd       = box[3] - box[2];
    g->zprd       = box[5] - box[4];
    g->cutneigh   = cutneigh;
    g->cutneighsq = cutneigh * cutneigh;
    g->binsizex = g->binsizey = g->binsizez = cutneigh * 0.5;
    g->nbinx = (int)((box[1] - box[0]) / g->binsizex);
    g->nbiny = (int)((box[3] - box[2]) / g->binsizey);
    g->nbinz = (int)((box[5] - box[4]) / g->binsizez);
    if (g->nbinx == 0) g->nbinx = 1;
    if (g->nbiny == 0) g->nbiny = 1;
    if (g->nbinz == 0) g->nbinz = 1;
    g->bininvx = g->nbinx / (box[1] - box[0]);
    g->bininvy = g->nbiny / (box[3] - box[2]);
    g->bininvz = g->nbinz / (box[5] - box[4]);
    return grid_tail(g, cutneigh);
}

static int grid_tail(orc_vl_grid* g, double cutneigh)
{
    const double FACTOR = 0.999;
    g->pad_x = (int)(cutneigh * g->bininvx);
    while (g->pad_x * g->binsizex < FACTOR * cutneigh)
        g->pad_x++;
    g->pad_y = (int)(cutneigh * g->bininvy);
    while (g->pad_y * g->binsizey < FACTOR * cutneigh)
        g->pad_y++;
    g->pad_z = (int)(cutneigh * g->bininvz);
    while (g->pad_z * g->binsizez < FACTOR * cutneigh)
        g->pad_z++;

    /* The stencil extent loops bump the pads a second time (neighbor.c:190-206). */
    int nextx = (int)(cutneigh * g->bininvx);
    while (nextx * g->binsizex < FACTOR * cutneigh) {
        nextx++;
        g->pad_x++;
    }
    int nexty = (int)(cutneigh * g->bininvy);
    while (nexty * g->binsizey < FACTOR * cutneigh) {
        nexty++;
        g->pad_y++;
    }
    int nextz = (int)(cutneigh * g->bininvz);
    while (nextz * g->binsizez < FACTOR * cutneigh) {
        nextz++;
        g->pad_z++;
    }
    g->mbinx = ORC_MAX(1, g->nbinx + 4 * g->pad_x);
    g->mbiny = ORC_MAX(1, g->nbiny + 4 * g->pad_y);
    g->mbinz = ORC_MAX(1, g->nbinz + 4 * g->pad_z);
    g->mbins = g->mbinx * g->mbiny * g->mbinz;

    g->nstencil = 0;
    for (int k = -nextz; k <= nextz; k++)
        for (int j = -nexty; j <= nexty; j++)
            for (int i = -nextx; i <= nextx; i++)
                if (vl_bindist(g, i, j, k) < g->cutneighsq) {
                    if (g->nstencil >= 1024) return -1;
                    g->stencil[g->nstencil++] = k * g->mbiny * g->mbinx + j * g->mbinx + i;
                }
    return 0;
}

/* coord2bin: neighbor.c:386-429 */
int orc_vl_coord2bin(const orc_vl_grid* g, double xin, double yin, double zin)
{
    double eps = 1e-9;
    double xlo = fabs(0.0 - g->pad_x * g->binsizex) + eps;
    double ylo = fabs(0.0 - g->pad_y * g->binsizey) + eps;
    double zlo = fabs(0.0 - g->pad_z * g->binsizez) + eps;
    int ix     = (int)((xin + xlo) * g->bininvx);
    int iy     = (int)((yin + ylo) * g->bininvy);
    int iz     = (int)((zin + zlo) * g->bininvz);
    return iz * g->mbiny * g->mbinx + iy * g->mbinx + ix;
}

/* binatoms: neighbor.c:431-464.  Atoms enter their bin in index order. */
void orc_vl_binatoms(const orc_vl_grid* g, int nall, const double* x, int* binstart, int* binatoms)
{
    int* cnt = calloc(g->mbins + 1, sizeof(int));
    int* bin = malloc(sizeof(int) * (nall + 1));
    for (int i = 0; i < nall; i++) {
        bin[i] = orc_vl_coord2bin(g, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        cnt[bin[i]]++;
    }
    binstart[0] = 0;
    for (int b = 0; b < g->mbins; b++)
        binstart[b + 1] = binstart[b] + cnt[b];
    memset(cnt, 0, sizeof(int) * g->mbins);
    for (int i = 0; i < nall; i++)
        binatoms[binstart[bin[i]] + cnt[bin[i]]++] = i;
    free(cnt);
    free(bin);
}

/* sortAtom: neighbor.c:466-536, with the caller's Nghost = 0 (main.c:92-95, 372-381): the local
   atoms binned alone and re-stored bin by bin, in binatoms order within a bin.  The reference
   moves x and v only; here f and type move too, so that the following kick uses each atom's own
   force (the product does the same). */
void orc_vl_sort_atoms(const orc_vl_grid* g, int nlocal, double* x, double* v, double* f,
    int* type)
{
    int* binstart = malloc(sizeof(int) * (g->mbins + 1));
    int* order    = malloc(sizeof(int) * (nlocal + 1));
    double* tmp   = malloc(sizeof(double) * 3 * (nlocal + 1));
    int* ttmp     = malloc(sizeof(int) * (nlocal + 1));
    orc_vl_binatoms(g, nlocal, x, binstart, order);
    double* arr[3] = { x, v, f };
    for (int a = 0; a < 3; a++) {
        if (!arr[a]) continue;
        for (int p = 0; p < nlocal; p++)
            for (int d = 0; d < 3; d++)
                tmp[3 * p + d] = arr[a][3 * order[p] + d];
        memcpy(arr[a], tmp, sizeof(double) * 3 * nlocal);
    }
    if (type) {
        for (int p = 0; p < nlocal; p++)
            ttmp[p] = type[order[p]];
        memcpy(type, ttmp, sizeof(int) * nlocal);
    }
    free(binstart);
    free(order);
    free(tmp);
    free(ttmp);
}

/* buildNeighborCPU: neighbor.c:260-352 (half_stencil = 0, no shell method). */
long long orc_vl_build_neighbor(const orc_vl_grid* g, int nlocal, int nall, const double* x,
    const int* type, int ntypes, const double* cutneighsq, int half, int* numneigh,
    int* neighbors, long long cap)
{
    int* binstart = malloc(sizeof(int) * (g->mbins + 1));
    int* binatoms = malloc(sizeof(int) * (nall + 1));
    orc_vl_binatoms(g, nall, x, binstart, binatoms);
    long long tot = 0;
    int err       = 0;
    for (int i = 0; i < nlocal && !err; i++) {
        int n       = 0;
        double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
        int ibin    = orc_vl_coord2bin(g, xtmp, ytmp, ztmp);
        int type_i  = type[i];
        for (int k = 0; k < g->nstencil; k++) {
            int jbin = ibin + g->stencil[k];
            if (jbin < 0 || jbin >= g->mbins) {
                err = 1;
                break;
            }
            for (int m = binstart[jbin]; m < binstart[jbin + 1]; m++) {
                int j = binatoms[m];
                if (i == j) continue;
                if (half && j < i) continue;
                double delx = xtmp - x[3 * j];
                double dely = ytmp - x[3 * j + 1];
                double delz = ztmp - x[3 * j + 2];
                double rsq  = delx * delx + dely * dely + delz * delz;
                if (rsq <= cutneighsq[type_i * ntypes + type[j]]) {
                    if (tot >= cap) {
                        err = 1;
                        break;
                    }
                    neighbors[tot++] = j;
                    n++;
                }
            }
            if (err) break;
        }
        numneigh[i] = n;
    }
    free(binstart);
    free(binatoms);
    return err ? -1 : tot;
}

/* setupPbc: src/verletlist/pbc.c:103-232 (all directions periodic).  pbc[3g..3g+2] = shift. */
int orc_vl_setup_pbc(int nlocal, const double* x, double xprd, double yprd, double zprd,
    double cutneigh, int* border_map, int* pbc, int cap)
{
    int ng = 0;
#define ORC_ADDGHOST(dx, dy, dz)                                                                 \
    do {                                                                                         \
        if (ng >= cap) return -1;                                                                \
        border_map[ng] = i;                                                                      \
        pbc[3 * ng]    = (dx);                                                                   \
        pbc[3 * ng + 1] = (dy);                                                                  \
        pbc[3 * ng + 2] = (dz);                                                                  \
        ng++;                                                                                    \
    } while (0)
    for (int i = 0; i < nlocal; i++) {
        double xx = x[3 * i], yy = x[3 * i + 1], zz = x[3 * i + 2];
        int xl = xx < cutneigh, xh = xx >= (xprd - cutneigh);
        int yl = yy < cutneigh, yh = yy >= (yprd - cutneigh);
        int zl = zz < cutneigh, zh = zz >= (zprd - cutneigh);
        /* 6 planes */
        if (xl) ORC_ADDGHOST(+1, 0, 0);
        if (xh) ORC_ADDGHOST(-1, 0, 0);
        if (yl) ORC_ADDGHOST(0, +1, 0);
        if (yh) ORC_ADDGHOST(0, -1, 0);
        if (zl) ORC_ADDGHOST(0, 0, +1);
        if (zh) ORC_ADDGHOST(0, 0, -1);
        /* 8 corners */
        if (xl && yl && zl) ORC_ADDGHOST(+1, +1, +1);
        if (xl && yh && zl) ORC_ADDGHOST(+1, -1, +1);
        if (xl && yl && zh) ORC_ADDGHOST(+1, +1, -1);
        if (xl && yh && zh) ORC_ADDGHOST(+1, -1, -1);
        if (xh && yl && zl) ORC_ADDGHOST(-1, +1, +1);
        if (xh && yh && zl) ORC_ADDGHOST(-1, -1, +1);
        if (xh && yl && zh) ORC_ADDGHOST(-1, +1, -1);
        if (xh && yh && zh) ORC_ADDGHOST(-1, -1, -1);
        /* 12 edges */
        if (xl && zl) ORC_ADDGHOST(+1, 0, +1);
        if (xl && zh) ORC_ADDGHOST(+1, 0, -1);
        if (xh && zl) ORC_ADDGHOST(-1, 0, +1);
        if (xh && zh) ORC_ADDGHOST(-1, 0, -1);
        if (yl && zl) ORC_ADDGHOST(0, +1, +1);
        if (yl && zh) ORC_ADDGHOST(0, +1, -1);
        if (yh && zl) ORC_ADDGHOST(0, -1, +1);
        if (yh && zh) ORC_ADDGHOST(0, -1, -1);
        if (yl && xl) ORC_ADDGHOST(+1, +1, 0);
        if (yl && xh) ORC_ADDGHOST(-1, +1, 0);
        if (yh && xl) ORC_ADDGHOST(+1, -1, 0);
        if (yh && xh) ORC_ADDGHOST(-1, -1, 0);
    }
#undef ORC_ADDGHOST
    return ng;
}

/* updatePbcCPU: pbc.c:42-58 */
void orc_vl_update_pbc(int nlocal, int nghost, double* x, const int* border_map, const int* pbc,
    double xprd, double yprd, double zprd)
{
    for (int i = 0; i < nghost; i++) {
        int b                    = border_map[i];
        x[3 * (nlocal + i)]      = x[3 * b] + pbc[3 * i] * xprd;
        x[3 * (nlocal + i) + 1]  = x[3 * b + 1] + pbc[3 * i + 1] * yprd;
        x[3 * (nlocal + i) + 2]  = x[3 * b + 2] + pbc[3 * i + 2] * zprd;
    }
}

/* updateAtomsPbcCPU: pbc.c:62-87 (same body in clusterpair/pbc.c:120-145) */
void orc_update_atoms_pbc(int nlocal, double* x, double xprd, double yprd, double zprd)
{
    const double prd[3] = { xprd, yprd, zprd };
    for (int i = 0; i < nlocal; i++)
        for (int d = 0; d < 3; d++) {
            if (x[3 * i + d] < 0.0) x[3 * i + d] += prd[d];
            else if (x[3 * i + d] >= prd[d]) x[3 * i + d] -= prd[d];
        }
}

/* computeForceLJFullNeigh (force_lj.c:17-110) / computeForceLJHalfNeigh (:112-206). */
void orc_vl_force_lj(int nlocal, int nall, const double* x, const int* type, int ntypes,
    const double* epsilon, const double* sigma6, const double* cutforcesq, const int* numneigh,
    const int* neighbors, int half, double* f, double* energy, double* virial)
{
    const double num1 = 1.0, num48 = 48.0, num05 = 0.5;
    int zero_n        = half ? nall : nlocal;
    for (int i = 0; i < 3 * zero_n; i++)
        f[i] = 0.0;
    double e = 0.0, w = 0.0;
    long long off = 0;
    for (int i = 0; i < nlocal; i++) {
        int nn      = numneigh[i];
        double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
        double fix = 0, fiy = 0, fiz = 0;
        int type_i = type[i];
        for (int k = 0; k < nn; k++) {
            int j       = neighbors[off + k];
            double delx = xtmp - x[3 * j];
            double dely = ytmp - x[3 * j + 1];
            double delz = ztmp - x[3 * j + 2];
            double rsq  = delx * delx + dely * dely + delz * delz;
            int tij     = type_i * ntypes + type[j];
            if (rsq < cutforcesq[tij]) {
                double sr2   = num1 / rsq;
                double sr6   = sr2 * sr2 * sr2 * sigma6[tij];
                double force = num48 * sr6 * (sr6 - num05) * sr2 * epsilon[tij];
                fix += delx * force;
                fiy += dely * force;
                fiz += delz * force;
                if (half && j < nlocal) {
                    f[3 * j] -= delx * force;
                    f[3 * j + 1] -= dely * force;
                    f[3 * j + 2] -= delz * force;
                }
                /* Pair energy and virial are new (the reference has neither). A half list
                   holds a local-ghost pair once from each side, so it is counted half like
                   every full-list pair; only local-local half pairs count whole. */
                double wgt = (half && j < nlocal) ? 1.0 : 0.5;
                e += wgt * 4.0 * epsilon[tij] * sr6 * (sr6 - 1.0);
                w += wgt * rsq * force;
            }
        }
        f[3 * i] += fix;
        f[3 * i + 1] += fiy;
        f[3 * i + 2] += fiz;
        off += nn;
    }
    if (energy) *energy = e;
    if (virial) *virial = w;
}

/* Velocity Verlet: integrate.c:21-42 */
void orc_initial_integrate(int n, double* x, double* v, const double* f, double dt, double dtforce)
{
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            v[3 * i + d] += dtforce * f[3 * i + d];
            x[3 * i + d] = x[3 * i + d] + dt * v[3 * i + d];
        }
}

void orc_final_integrate(int n, double* v, const double* f, double dtforce)
{
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++)
            v[3 * i + d] += dtforce * f[3 * i + d];
}

/* computeThermo's serial sum (thermo.c:76-81). */
double orc_kinetic_sum(int n, const double* v, double mass)
{
    double t = 0.0;
    for (int i = 0; i < n; i++)
        t += (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) *
             mass;
    return t;
}

/* Timestep loop: src/verletlist/main.c:46-113 (setup after adjustThermo) and :371-419. */
static int run_on_grid(const orc_vl_grid* gp, int nlocal, double* x, double* v,
    const int* type, int ntypes, const double* epsilon, const double* sigma6,
    const double* cutforcesq, const double* cutneighsq, double cutneigh, int half,
    int force_field, const orc_eam* eam, double dt, double dtforce, int ntimes,
    int reneigh_every, double mass, double* f, double* ksum, int resort_every, int* type_out);

int orc_vl_run(int nx, int ny, int nz, double lattice, int nlocal, double* x, double* v,
    const int* type, int ntypes, const double* epsilon, const double* sigma6,
    const double* cutforcesq, const double* cutneighsq, double cutneigh, int half,
    int force_field, const orc_eam* eam, double dt, double dtforce, int ntimes,
    int reneigh_every, double mass, double* f, double* ksum)
{
    orc_vl_grid g;
    if (orc_vl_setup_grid(&g, nx, ny, nz, lattice, cutneigh)) return -1;
    return run_on_grid(&g, nlocal, x, v, type, ntypes, epsilon, sigma6, cutforcesq, cutneighsq,
        cutneigh, half, force_field, eam, dt, dtforce, ntimes, reneigh_every, mass, f, ksum, 0,
        NULL);
}

int orc_vl_run_box(const double* box, int nlocal, double* x, double* v, const int* type,
    int ntypes, const double* epsilon, const double* sigma6, const double* cutforcesq,
    const double* cutneighsq, double cutneigh, int half, int force_field, const orc_eam* eam,
    double dt, double dtforce, int ntimes, int reneigh_every, double mass, double* f,
    double* ksum)
{
    orc_vl_grid g;
    if (orc_vl_setup_grid_box(&g, box, cutneigh)) return -1;
    return run_on_grid(&g, nlocal, x, v, type, ntypes, epsilon, sigma6, cutforcesq, cutneighsq,
        cutneigh, half, force_field, eam, dt, dtforce, ntimes, reneigh_every, mass, f, ksum, 0,
        NULL);
}

/* The run loop with SORT_ATOMS (main.c:92-95, 372-381): sorted at setup and every resort_every
   steps before the step's initialIntegrate, and that step rebuilds.  x, v (and f, type_out when
   given) come back in the final order. */
int orc_vl_run_resort(int nx, int ny, int nz, double lattice, int nlocal, double* x, double* v,
    const int* type, int ntypes, const double* epsilon, const double* sigma6,
    const double* cutforcesq, const double* cutneighsq, double cutneigh, int half,
    int force_field, const orc_eam* eam, double dt, double dtforce, int ntimes,
    int reneigh_every, double mass, double* f, double* ksum, int resort_every, int* type_out)
{
    orc_vl_grid g;
    if (orc_vl_setup_grid(&g, nx, ny, nz, lattice, cutneigh)) return -1;
    return run_on_grid(&g, nlocal, x, v, type, ntypes, epsilon, sigma6, cutforcesq, cutneighsq,
        cutneigh, half, force_field, eam, dt, dtforce, ntimes, reneigh_every, mass, f, ksum,
        resort_every, type_out);
}

static int run_on_grid(const orc_vl_grid* gp, int nlocal, double* x, double* v,
    const int* type, int ntypes, const double* epsilon, const double* sigma6,
    const double* cutforcesq, const double* cutneighsq, double cutneigh, int half,
    int force_field, const orc_eam* eam, double dt, double dtforce, int ntimes,
    int reneigh_every, double mass, double* f, double* ksum, int resort_every, int* type_out)
{
    const orc_vl_grid g = *gp;
    int gcap         = 27 * nlocal + 1024;
    int nmax         = nlocal + gcap;
    double* xa       = malloc(sizeof(double) * 3 * nmax);
    double* fa       = malloc(sizeof(double) * 3 * nmax);
    int* ta          = malloc(sizeof(int) * nmax);
    int* bmap        = malloc(sizeof(int) * gcap);
    int* pbc         = malloc(sizeof(int) * 3 * gcap);
    int* numneigh    = malloc(sizeof(int) * nlocal);
    long long ncap   = (long long)nlocal * 400;
    int* nbl         = malloc(sizeof(int) * ncap);
    double* fp       = malloc(sizeof(double) * nmax);
    int rc           = 0;
    memcpy(xa, x, sizeof(double) * 3 * nlocal);
    memcpy(ta, type, sizeof(int) * nlocal);

    int ng = -1;
#define ORC_RENEIGH()                                                                            \
    do {                                                                                         \
        ng = orc_vl_setup_pbc(nlocal, xa, g.xprd, g.yprd, g.zprd, cutneigh, bmap, pbc, gcap);    \
        if (ng < 0) {                                                                            \
            rc = -2;                                                                             \
            goto done;                                                                           \
        }                                                                                        \
        for (int q = 0; q < ng; q++)                                                             \
            ta[nlocal + q] = ta[bmap[q]];                                                        \
        orc_vl_update_pbc(nlocal, ng, xa, bmap, pbc, g.xprd, g.yprd, g.zprd);                    \
        if (orc_vl_build_neighbor(&g, nlocal, nlocal + ng, xa, ta, ntypes, cutneighsq, half,     \
                numneigh, nbl, ncap) < 0) {                                                      \
            rc = -3;                                                                             \
            goto done;                                                                           \
        }                                                                                        \
    } while (0)
#define ORC_FORCE()                                                                              \
    do {                                                                                         \
        if (force_field == 0)                                                                    \
            orc_vl_force_lj(nlocal, nlocal + ng, xa, ta, ntypes, epsilon, sigma6, cutforcesq,    \
                numneigh, nbl, half, fa, NULL, NULL);                                            \
        else                                                                                     \
            orc_vl_force_eam(eam, nlocal, ng, xa, bmap, cutforcesq[0], numneigh, nbl, fp, fa,    \
                NULL);                                                                           \
    } while (0)

    if (resort_every > 0) orc_vl_sort_atoms(&g, nlocal, xa, v, NULL, ta);
    ORC_RENEIGH();
    ORC_FORCE();
    if (ksum) ksum[0] = orc_kinetic_sum(nlocal, v, mass);
    for (int n = 0; n < ntimes; n++) {
        const int resort = resort_every > 0 && (n + 1) % resort_every == 0;
        int reneigh      = resort || (n + 1) % reneigh_every == 0;
        if (resort) orc_vl_sort_atoms(&g, nlocal, xa, v, fa, ta);
        orc_initial_integrate(nlocal, xa, v, fa, dt, dtforce);
        if (reneigh) {
            orc_update_atoms_pbc(nlocal, xa, g.xprd, g.yprd, g.zprd);
            ORC_RENEIGH();
        } else {
            orc_vl_update_pbc(nlocal, ng, xa, bmap, pbc, g.xprd, g.yprd, g.zprd);
        }
        ORC_FORCE();
        orc_final_integrate(nlocal, v, fa, dtforce);
        if (ksum) ksum[n + 1] = orc_kinetic_sum(nlocal, v, mass);
    }
    memcpy(x, xa, sizeof(double) * 3 * nlocal);
    if (f) memcpy(f, fa, sizeof(double) * 3 * nlocal);
    if (type_out) memcpy(type_out, ta, sizeof(int) * nlocal);
#undef ORC_RENEIGH
#undef ORC_FORCE
done:
    free(xa);
    free(fa);
    free(ta);
    free(bmap);
    free(pbc);
    free(numneigh);
    free(nbl);
    free(fp);
    return rc;
}
