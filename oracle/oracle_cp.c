/*
 * CPU oracle, cluster-pair (GROMACS NBNXM-style) path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates src/clusterpair/{neighbor.c,pbc.c,force_lj.c,integrate.c} for an M x N cluster
 * shape (1x2, 4x4, 4x8), strict IEEE double arithmetic.  Layout: positions per "slot"
 * (AoS x,y,z).  i-cluster ci owns slots [ci*M, ci*M+M); j-cluster cj owns slots [cj*N, cj*N+N);
 * ghost j-clusters follow the local ones, then one dummy cluster of INF positions.
 *
 * Pinning: cluster membership, bounding boxes, ghost map, list contents and order, and (for 1x2)
 * forces and whole trajectories are compared bit for bit with MD-Bench compiled from its sources
 * (strict builds in oracle/_ref; tests/golden/cp_*.npz).  For 4x4 and 4x8 the reference's force
 * kernels are SIMD with approximate reciprocals, so the force of this restatement is pinned
 * instead by the Verlet-list oracle on the same positions (same pair set, FP64) -- see
 * tests/test_oracle_cp.py.
 *
 * Reference choices kept on purpose:
 *   - the list-build distance test uses fma() when the reference's 4xN SIMD build does
 *     (fma_dist = 1, src/clusterpair/neighbor.c:452-470), plain arithmetic otherwise;
 *   - the "masked first" ordering uses the reference's mask rule for its vector width
 *     (get_imask_simd_j2/j4/j8, neighbor.c:240-280), which for the 1x2 build flags cj = 2ci and
 *     2ci+1.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_md.h"

#define CMAX(a, b) ((a) > (b) ? (a) : (b))
#define CMIN(a, b) ((a) < (b) ? (a) : (b))

struct orc_cp {
    int M, N, imask_rule, fma_dist;
    double xprd, yprd, zprd, cutneigh, cutneighsq;
    int nbinx, nbiny, mbinxlo, mbinylo, mbinx, mbiny, mbins;
    double binsizex, binsizey, bininvx, bininvy;
    int nstencil;
    int stencil[4096];
    /* atoms (reference atom order) */
    int natoms;
    double *ax, *av;
    int* at;
    /* atom bins (2D columns) */
    int *abin_start, *abin_atoms;
    /* clusters */
    int nci, ncj, nghost, dummy_cj, slot_cap, cl_cap;
    double *sx, *sv, *sf;
    int* st;
    int *ci_natoms, *ci_bin;
    double* ci_bbox;
    int* cj_natoms;
    double* cj_bbox;
    int *border_map, *pbc;
    /* j-cluster bins */
    int *cbin_start, *cbin_list;
    /* neighbour list, CSR */
    int *numneigh, *numneigh_masked, *nb;
    unsigned* imask;
    long long nb_total;
};

static double cp_bindist(const orc_cp* g, int i, int j)
{
    double delx, dely;
    if (i > 0) delx = (i - 1) * g->binsizex;
    else if (i == 0) delx = 0.0;
    else delx = (i + 1) * g->binsizex;
    if (j > 0) dely = (j - 1) * g->binsizey;
    else if (j == 0) dely = 0.0;
    else dely = (j + 1) * g->binsizey;
    return delx * delx + dely * dely;
}

/* coord2bin2D: neighbor.c:1393-1407 */
static void cp_coord2bin2d(const orc_cp* g, double xin, double yin, int* ix, int* iy)
{
    if (xin >= g->xprd) *ix = (int)((xin - g->xprd) * g->bininvx) + g->nbinx - g->mbinxlo;
    else if (xin >= 0.0) *ix = (int)(xin * g->bininvx) - g->mbinxlo;
    else *ix = (int)(xin * g->bininvx) - g->mbinxlo - 1;
    if (yin >= g->yprd) *iy = (int)((yin - g->yprd) * g->bininvy) + g->nbiny - g->mbinylo;
    else if (yin >= 0.0) *iy = (int)(yin * g->bininvy) - g->mbinylo;
    else *iy = (int)(yin * g->bininvy) - g->mbinylo - 1;
}

/* coord2bin: neighbor.c:1367-1388 (note the +1) */
static int cp_coord2bin(const orc_cp* g, double x, double y)
{
    int ix, iy;
    cp_coord2bin2d(g, x, y, &ix, &iy);
    return iy * g->mbinx + ix + 1;
}

/* initNeighbor (neighbor.c:57-106) + setupNeighbor (neighbor.c:108-215) */
static void cp_setup_grid(orc_cp* g, int natoms)
{
    const double SMALL = 1.0e-6, FACTOR = 0.999;
    double xlo = 0.0, xhi = g->xprd, ylo = 0.0, yhi = g->yprd, zlo = 0.0, zhi = g->zprd;
    double atom_density  = ((double)natoms) / ((xhi - xlo) * (yhi - ylo) * (zhi - zlo));
    double atoms_in_cell = CMAX(g->M, g->N);
    double targetsizex   = cbrt(atoms_in_cell / atom_density);
    double targetsizey   = cbrt(atoms_in_cell / atom_density);
    g->nbinx             = CMAX(1, (int)ceil((xhi - xlo) / targetsizex));
    g->nbiny             = CMAX(1, (int)ceil((yhi - ylo) / targetsizey));
    g->binsizex          = (xhi - xlo) / g->nbinx;
    g->binsizey          = (yhi - ylo) / g->nbiny;
    g->bininvx           = 1.0 / g->binsizex;
    g->bininvy           = 1.0 / g->binsizey;
    g->cutneighsq        = g->cutneigh * g->cutneigh;
    double coord         = xlo - g->cutneigh - SMALL * g->xprd;
    g->mbinxlo           = (int)(coord * g->bininvx);
    if (coord < 0.0) g->mbinxlo = g->mbinxlo - 1;
    coord       = xhi + g->cutneigh + SMALL * g->xprd;
    int mbinxhi = (int)(coord * g->bininvx);
    coord       = ylo - g->cutneigh - SMALL * g->yprd;
    g->mbinylo  = (int)(coord * g->bininvy);
    if (coord < 0.0) g->mbinylo = g->mbinylo - 1;
    coord       = yhi + g->cutneigh + SMALL * g->yprd;
    int mbinyhi = (int)(coord * g->bininvy);
    g->mbinxlo  = g->mbinxlo - 1;
    mbinxhi     = mbinxhi + 1;
    g->mbinx    = mbinxhi - g->mbinxlo + 1;
    g->mbinylo  = g->mbinylo - 1;
    mbinyhi     = mbinyhi + 1;
    g->mbiny    = mbinyhi - g->mbinylo + 1;
    int nextx   = (int)(g->cutneigh * g->bininvx);
    int nexty   = (int)(g->cutneigh * g->bininvy);
    if (nextx * g->binsizex < FACTOR * g->cutneigh) nextx++;
    if (nexty * g->binsizey < FACTOR * g->cutneigh) nexty++;
    g->nstencil = 0;
    for (int j = -nexty; j <= nexty; j++)
        for (int i = -nextx; i <= nextx; i++)
            if (cp_bindist(g, i, j) < g->cutneighsq && g->nstencil < 4096)
                g->stencil[g->nstencil++] = j * g->mbinx + i;
    g->mbins = g->mbinx * g->mbiny;
}

orc_cp* orc_cp_create(int M, int N, int imask_rule, int fma_dist, double xprd, double yprd,
    double zprd, double cutneigh, int natoms, const double* x, const double* v, const int* type)
{
    orc_cp* g = calloc(1, sizeof(orc_cp));
    g->M = M;
    g->N = N;
    g->imask_rule = imask_rule;
    g->fma_dist   = fma_dist;
    g->xprd = xprd;
    g->yprd = yprd;
    g->zprd = zprd;
    g->cutneigh = cutneigh;
    g->natoms   = natoms;
    g->ax = malloc(sizeof(double) * 3 * natoms);
    g->av = malloc(sizeof(double) * 3 * natoms);
    g->at = malloc(sizeof(int) * natoms);
    memcpy(g->ax, x, sizeof(double) * 3 * natoms);
    memcpy(g->av, v, sizeof(double) * 3 * natoms);
    memcpy(g->at, type, sizeof(int) * natoms);
    cp_setup_grid(g, natoms);
    g->abin_start = malloc(sizeof(int) * (g->mbins + 1));
    g->abin_atoms = malloc(sizeof(int) * (natoms + 1));
    /* capacities: every atom its own cluster plus ghosts (26 images) and padding */
    g->cl_cap   = 64 * (natoms / (M < N ? M : N) + 16);
    g->slot_cap = g->cl_cap * CMAX(M, N);
    g->sx = calloc(3 * (size_t)g->slot_cap, sizeof(double));
    g->sv = calloc(3 * (size_t)g->slot_cap, sizeof(double));
    g->sf = calloc(3 * (size_t)g->slot_cap, sizeof(double));
    g->st = calloc((size_t)g->slot_cap, sizeof(int));
    g->ci_natoms = calloc(g->cl_cap, sizeof(int));
    g->ci_bin    = calloc(g->cl_cap, sizeof(int));
    g->ci_bbox   = calloc(6 * (size_t)g->cl_cap, sizeof(double));
    g->cj_natoms = calloc(g->cl_cap, sizeof(int));
    g->cj_bbox   = calloc(6 * (size_t)g->cl_cap, sizeof(double));
    g->border_map = calloc(g->cl_cap, sizeof(int));
    g->pbc        = calloc(3 * (size_t)g->cl_cap, sizeof(int));
    g->cbin_start = calloc(g->mbins + 1, sizeof(int));
    g->cbin_list  = calloc(g->cl_cap, sizeof(int));
    g->numneigh        = calloc(g->cl_cap, sizeof(int));
    g->numneigh_masked = calloc(g->cl_cap, sizeof(int));
    g->nb    = NULL;
    g->imask = NULL;
    return g;
}

void orc_cp_free(orc_cp* g)
{
    if (!g) return;
    free(g->ax);
    free(g->av);
    free(g->at);
    free(g->abin_start);
    free(g->abin_atoms);
    free(g->sx);
    free(g->sv);
    free(g->sf);
    free(g->st);
    free(g->ci_natoms);
    free(g->ci_bin);
    free(g->ci_bbox);
    free(g->cj_natoms);
    free(g->cj_bbox);
    free(g->border_map);
    free(g->pbc);
    free(g->cbin_start);
    free(g->cbin_list);
    free(g->numneigh);
    free(g->numneigh_masked);
    free(g->nb);
    free(g->imask);
    free(g);
}

/* binAtoms (neighbor.c:1409-1449) + sortAtomsByZCoord (neighbor.c:1452-1481) +
   buildClustersCPU (neighbor.c:1513-1611) */
int orc_cp_build_clusters(orc_cp* g)
{
    const int M = g->M, N = g->N;
    int* cnt = calloc(g->mbins + 1, sizeof(int));
    int* bin = malloc(sizeof(int) * (g->natoms + 1));
    for (int i = 0; i < g->natoms; i++) {
        int b = cp_coord2bin(g, g->ax[3 * i], g->ax[3 * i + 1]);
        if (b < 0 || b >= g->mbins) {
            free(cnt);
            free(bin);
            return -1;
        }
        bin[i] = b;
        cnt[b]++;
    }
    g->abin_start[0] = 0;
    for (int b = 0; b < g->mbins; b++)
        g->abin_start[b + 1] = g->abin_start[b] + cnt[b];
    memset(cnt, 0, sizeof(int) * g->mbins);
    for (int i = 0; i < g->natoms; i++)
        g->abin_atoms[g->abin_start[bin[i]] + cnt[bin[i]]++] = i;
    free(cnt);
    free(bin);
    /* selection sort by z inside every bin, strict <, swap with the first minimum */
    for (int b = 0; b < g->mbins; b++) {
        int* bp = &g->abin_atoms[g->abin_start[b]];
        int c   = g->abin_start[b + 1] - g->abin_start[b];
        for (int a = 0; a < c; a++) {
            int i = bp[a], min_ac = a, min_idx = i;
            double min_z = g->ax[3 * i + 2];
            for (int q = a + 1; q < c; q++) {
                int j     = bp[q];
                double zj = g->ax[3 * j + 2];
                if (zj < min_z) {
                    min_ac  = q;
                    min_idx = j;
                    min_z   = zj;
                }
            }
            bp[a]      = min_idx;
            bp[min_ac] = i;
        }
    }
    g->nci = 0;
    for (int b = 0; b < g->mbins; b++) {
        int c         = g->abin_start[b + 1] - g->abin_start[b];
        int nclusters = (c + M - 1) / M;
        if (N > M && nclusters % 2) nclusters++;
        int ac = 0;
        for (int cl = 0; cl < nclusters; cl++) {
            int ci = g->nci;
            if ((ci + 2) * M > g->slot_cap || ci + 2 > g->cl_cap) return -2;
            double bb[6] = { INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY };
            g->ci_natoms[ci] = 0;
            for (int cii = 0; cii < M; cii++) {
                int s = ci * M + cii;
                if (ac < c) {
                    int i = g->abin_atoms[g->abin_start[b] + ac];
                    for (int d = 0; d < 3; d++) {
                        double xv = g->ax[3 * i + d];
                        g->sx[3 * s + d] = xv;
                        g->sv[3 * s + d] = g->av[3 * i + d];
                        if (bb[2 * d] > xv) bb[2 * d] = xv;
                        if (bb[2 * d + 1] < xv) bb[2 * d + 1] = xv;
                    }
                    g->st[s] = g->at[i];
                    g->ci_natoms[ci]++;
                } else {
                    for (int d = 0; d < 3; d++) {
                        g->sx[3 * s + d] = INFINITY;
                        g->sv[3 * s + d] = 0.0;
                    }
                    g->st[s] = 0;
                }
                ac++;
            }
            g->ci_bin[ci] = b;
            memcpy(&g->ci_bbox[6 * ci], bb, sizeof(bb));
            g->nci++;
        }
    }
    return 0;
}

/* defineJClusters: neighbor.c:1807-1927 (M == N and M < N shapes) */
int orc_cp_define_jclusters(orc_cp* g)
{
    const int M = g->M, N = g->N;
    if (M > N) return -1;
    g->ncj = g->nci * M / N;
    for (int ci = 0; ci < g->nci; ci++) {
        if (M == N) {
            memcpy(&g->cj_bbox[6 * ci], &g->ci_bbox[6 * ci], 6 * sizeof(double));
            g->cj_natoms[ci] = g->ci_natoms[ci];
        } else if (ci % 2 == 0) {
            int cj = ci >> 1, c1 = ci + 1;
            double* o = &g->cj_bbox[6 * cj];
            double* a = &g->ci_bbox[6 * ci];
            double* b = &g->ci_bbox[6 * c1];
            for (int d = 0; d < 3; d++) {
                o[2 * d]     = CMIN(a[2 * d], b[2 * d]);
                o[2 * d + 1] = CMAX(a[2 * d + 1], b[2 * d + 1]);
            }
            g->cj_natoms[cj] = g->ci_natoms[ci] + g->ci_natoms[c1];
        }
    }
    return 0;
}

/* updatePbcCPU: src/clusterpair/pbc.c:43-116 */
void orc_cp_update_pbc(orc_cp* g, int first)
{
    const int N = g->N;
    for (int cg = 0; cg < g->nghost; cg++) {
        int cj = g->ncj + cg, b = g->border_map[cg];
        double bb[6] = { INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY };
        double prd[3] = { g->xprd, g->yprd, g->zprd };
        for (int cjj = 0; cjj < g->cj_natoms[cj]; cjj++) {
            int s = cj * N + cjj, sb = b * N + cjj;
            for (int d = 0; d < 3; d++) {
                double xv        = g->sx[3 * sb + d] + g->pbc[3 * cg + d] * prd[d];
                g->sx[3 * s + d] = xv;
                if (first) {
                    if (bb[2 * d] > xv) bb[2 * d] = xv;
                    if (bb[2 * d + 1] < xv) bb[2 * d + 1] = xv;
                }
            }
            g->st[s] = g->st[sb];
        }
        if (first) {
            for (int cjj = g->cj_natoms[cj]; cjj < N; cjj++) {
                int s = cj * N + cjj;
                for (int d = 0; d < 3; d++)
                    g->sx[3 * s + d] = INFINITY;
                g->st[s] = 0;
            }
            memcpy(&g->cj_bbox[6 * cj], bb, sizeof(bb));
        }
    }
}

/* setupPbc: src/clusterpair/pbc.c:188-331.  Same 26 images in the same order as the
   Verlet-list ghost map, decided on the j-cluster bounding box. */
int orc_cp_setup_pbc(orc_cp* g)
{
    static const signed char side[26][3] = { { 1, 0, 0 }, { 2, 0, 0 }, { 0, 1, 0 }, { 0, 2, 0 },
        { 0, 0, 1 }, { 0, 0, 2 }, { 1, 1, 1 }, { 1, 2, 1 }, { 1, 1, 2 }, { 1, 2, 2 }, { 2, 1, 1 },
        { 2, 2, 1 }, { 2, 1, 2 }, { 2, 2, 2 }, { 1, 0, 1 }, { 1, 0, 2 }, { 2, 0, 1 }, { 2, 0, 2 },
        { 0, 1, 1 }, { 0, 1, 2 }, { 0, 2, 1 }, { 0, 2, 2 }, { 1, 1, 0 }, { 2, 1, 0 }, { 1, 2, 0 },
        { 2, 2, 0 } };
    const int N   = g->N;
    double prd[3] = { g->xprd, g->yprd, g->zprd };
    int ng        = 0;
    for (int cj = 0; cj < g->ncj; cj++) {
        if (g->cj_natoms[cj] <= 0) continue;
        double* bb = &g->cj_bbox[6 * cj];
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = bb[2 * d] < g->cutneigh;
            hi[d] = bb[2 * d + 1] >= (prd[d] - g->cutneigh);
        }
        for (int c = 0; c < 26; c++) {
            int ok = 1;
            for (int d = 0; d < 3; d++) {
                if (side[c][d] == 1) ok = ok && lo[d];
                if (side[c][d] == 2) ok = ok && hi[d];
            }
            if (!ok) continue;
            int cg = g->ncj + ng;
            if (cg + 2 > g->cl_cap) return -1;
            g->border_map[ng] = cj;
            for (int d = 0; d < 3; d++)
                g->pbc[3 * ng + d] = side[c][d] == 1 ? 1 : (side[c][d] == 2 ? -1 : 0);
            g->cj_natoms[cg] = g->cj_natoms[cj];
            for (int cjj = 0; cjj < g->cj_natoms[cj]; cjj++)
                g->st[cg * N + cjj] = g->st[cj * N + cjj];
            ng++;
        }
    }
    g->nghost   = ng;
    g->dummy_cj = g->ncj + ng;
    for (int cjj = 0; cjj < N; cjj++) {
        int s = g->dummy_cj * N + cjj;
        for (int d = 0; d < 3; d++)
            g->sx[3 * s + d] = INFINITY;
        g->st[s] = 0;
    }
    orc_cp_update_pbc(g, 1);
    return ng;
}

/* binJClusters: src/clusterpair/neighbor.c:1929-2094 */
int orc_cp_bin_jclusters(orc_cp* g)
{
    const int M = g->M, N = g->N;
    int* cnt  = calloc(g->mbins + 1, sizeof(int));
    int* gbin = malloc(sizeof(int) * (g->nghost + 1));
    for (int ci = 0; ci < g->nci; ci++)
        if (M >= N || ci % 2 == 0) cnt[g->ci_bin[ci]]++;
    for (int cg = 0; cg < g->nghost; cg++) {
        int cj  = g->ncj + cg;
        gbin[cg] = -1;
        if (g->cj_natoms[cj] <= 0) continue;
        int ix, iy;
        int s0 = cj * N;
        cp_coord2bin2d(g, g->sx[3 * s0], g->sx[3 * s0 + 1], &ix, &iy);
        ix = CMAX(CMIN(ix, g->mbinx - 1), 0);
        iy = CMAX(CMIN(iy, g->mbiny - 1), 0);
        for (int cjj = 1; cjj < g->cj_natoms[cj]; cjj++) {
            int nix, niy;
            int s = cj * N + cjj;
            cp_coord2bin2d(g, g->sx[3 * s], g->sx[3 * s + 1], &nix, &niy);
            nix = CMAX(CMIN(nix, g->mbinx - 1), 0);
            niy = CMAX(CMIN(niy, g->mbiny - 1), 0);
            if (g->pbc[3 * cg] > 0 && ix > nix) ix = nix;
            if (g->pbc[3 * cg] < 0 && ix < nix) ix = nix;
            if (g->pbc[3 * cg + 1] > 0 && iy > niy) iy = niy;
            if (g->pbc[3 * cg + 1] < 0 && iy < niy) iy = niy;
        }
        int b = iy * g->mbinx + ix + 1;
        if (b < 0 || b >= g->mbins) {
            free(cnt);
            free(gbin);
            return -1;
        }
        gbin[cg] = b;
        cnt[b]++;
    }
    g->cbin_start[0] = 0;
    for (int b = 0; b < g->mbins; b++)
        g->cbin_start[b + 1] = g->cbin_start[b] + cnt[b];
    memset(cnt, 0, sizeof(int) * (g->mbins + 1));
    for (int ci = 0; ci < g->nci; ci++)
        if (M >= N || ci % 2 == 0) {
            int b = g->ci_bin[ci];
            g->cbin_list[g->cbin_start[b] + cnt[b]++] = (M == N) ? ci : ci >> 1;
        }
    /* ghosts: insert before the first cluster whose bbminz is larger */
    for (int cg = 0; cg < g->nghost; cg++) {
        int b = gbin[cg];
        if (b < 0) continue;
        int cj       = g->ncj + cg;
        double zmin  = g->cj_bbox[6 * cj + 4];
        int* lst     = &g->cbin_list[g->cbin_start[b]];
        int c        = cnt[b];
        int pos      = c;
        for (int i = 0; i < c; i++)
            if (g->cj_bbox[6 * lst[i] + 4] > zmin) {
                pos = i;
                break;
            }
        for (int i = c; i > pos; i--)
            lst[i] = lst[i - 1];
        lst[pos] = cj;
        cnt[b]++;
    }
    free(cnt);
    free(gbin);
    return 0;
}

static unsigned cp_imask(const orc_cp* g, int ci, int cj)
{
    /* neighbor.c:240-280 with rdiag = 1; only "all" (0xffffffff) versus not matters here */
    switch (g->imask_rule) {
    case 2:
        return (ci * 2 == cj) ? 0x0002u : ((ci * 2 + 1 == cj) ? 0x002fu : 0xffffffffu);
    case 4:
        return (ci == cj) ? 0x08ceu : 0xffffffffu;
    default:
        return (ci == cj * 2) ? 0xf0f8fcfeu : ((ci == cj * 2 + 1) ? 0x0080c0e0u : 0xffffffffu);
    }
}

/* buildNeighborCPU: src/clusterpair/neighbor.c:282-766.  Returns total entries or -1. */
long long orc_cp_build_neighbor(orc_cp* g, int half)
{
    const int M = g->M, N = g->N;
    double bbx    = 0.5 * (g->binsizex + g->binsizex);
    double bby    = 0.5 * (g->binsizey + g->binsizey);
    double rbb_sq = CMAX(0.0, g->cutneigh - 0.5 * sqrt(bbx * bbx + bby * bby));
    rbb_sq        = rbb_sq * rbb_sq;
    long long cap = (long long)g->nci * 512 + 64;
    free(g->nb);
    free(g->imask);
    g->nb         = malloc(sizeof(int) * cap);
    g->imask      = malloc(sizeof(unsigned) * cap);
    long long off = 0;
    for (int ci = 0; ci < g->nci; ci++) {
        int ci_cj0       = (M == N) ? ci : ci >> 1;
        int n = 0, nmasked = 0;
        int ibin         = g->ci_bin[ci];
        double* ib       = &g->ci_bbox[6 * ci];
        int* out         = &g->nb[off];
        unsigned* om     = &g->imask[off];
        if (off + 4096 > cap) return -1;
        for (int k = 0; k < g->nstencil; k++) {
            int jbin = ibin + g->stencil[k];
            if (jbin < 0 || jbin >= g->mbins) return -2;
            for (int m = g->cbin_start[jbin]; m < g->cbin_start[jbin + 1]; m++) {
                int cj = g->cbin_list[m];
                if (half && ci_cj0 > cj) continue;
                double* jb = &g->cj_bbox[6 * cj];
                double d2  = 0.0;
                /* z, y, x in the reference's order */
                for (int d = 2; d >= 0; d--) {
                    double dl  = ib[2 * d] - jb[2 * d + 1];
                    double dh  = jb[2 * d] - ib[2 * d + 1];
                    double dm  = CMAX(dl, dh);
                    double dm0 = CMAX(dm, 0.0);
                    d2 += dm0 * dm0;
                }
                if (!(d2 < g->cutneighsq)) continue;
                int is_neighbor = d2 < rbb_sq;
                if (!is_neighbor) {
                    for (int cii = 0; cii < M && !is_neighbor; cii++) {
                        const double* xi = &g->sx[3 * (ci * M + cii)];
                        for (int cjj = 0; cjj < N; cjj++) {
                            const double* xj = &g->sx[3 * (cj * N + cjj)];
                            double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                            double rsq = g->fma_dist ? fma(dx, dx, fma(dy, dy, dz * dz))
                                                     : dx * dx + dy * dy + dz * dz;
                            if (rsq < g->cutneighsq) {
                                is_neighbor = 1;
                                break;
                            }
                        }
                    }
                }
                if (!is_neighbor) continue;
                unsigned im = cp_imask(g, ci, cj);
                if (im == 0xffffffffu) {
                    out[n] = cj;
                    om[n]  = im;
                } else {
                    out[n]       = out[nmasked];
                    om[n]        = om[nmasked];
                    out[nmasked] = cj;
                    om[nmasked]  = im;
                    nmasked++;
                }
                n++;
            }
        }
        g->numneigh[ci]        = n;
        g->numneigh_masked[ci] = nmasked;
        off += n;
    }
    g->nb_total = off;
    return off;
}

/* pruneNeighborCPU: src/clusterpair/neighbor.c:1099-1264 (the plain branch; with every shape
   here CLUSTER_N equals the reference's VECTOR_WIDTH, so no dummy padding is removed or added).
   An entry with no atom pair inside cutneigh is overwritten by the current last entry, which is
   tested in its place; a removal at a position inside the masked prefix shortens the prefix.
   As in the reference, the interaction masks stay at their positions.  The CSR storage is
   compacted afterwards (order within each list unchanged).  Returns the entries removed. */
long long orc_cp_prune(orc_cp* g)
{
    const int M = g->M, N = g->N;
    long long off = 0, out = 0, removed = 0;
    for (int ci = 0; ci < g->nci; ci++) {
        int n  = g->numneigh[ci];
        int nm = g->numneigh_masked[ci];
        int* lst      = &g->nb[off];
        unsigned* msk = &g->imask[off];
        int k = 0;
        while (k < n) {
            int cj = lst[k], in = 0;
            for (int cii = 0; cii < M && !in; cii++) {
                const double* xi = &g->sx[3 * (ci * M + cii)];
                for (int cjj = 0; cjj < N; cjj++) {
                    const double* xj = &g->sx[3 * (cj * N + cjj)];
                    double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                    double rsq = g->fma_dist ? fma(dx, dx, fma(dy, dy, dz * dz))
                                             : dx * dx + dy * dy + dz * dz;
                    if (rsq < g->cutneighsq) {
                        in = 1;
                        break;
                    }
                }
            }
            if (in) {
                k++;
            } else {
                n--;
                if (k < nm) nm--;
                lst[k] = lst[n];
            }
        }
        removed += g->numneigh[ci] - n;
        off += g->numneigh[ci];
        memmove(&g->nb[out], lst, sizeof(int) * n);
        memmove(&g->imask[out], msk, sizeof(unsigned) * n);
        out += n;
        g->numneigh[ci]        = n;
        g->numneigh_masked[ci] = nm;
    }
    g->nb_total = out;
    return removed;
}

/* computeForceLJRef: src/clusterpair/force_lj.c:22-189 (single type: eps, sigma6 given). f is
   per slot; energy/virial (may be NULL) are pair sums counted once per pair (new). */
void orc_cp_force_lj(orc_cp* g, int half, double epsilon, double sigma6, double cutforcesq,
    double* energy, double* virial)
{
    const int M = g->M, N = g->N;
    int nslot   = (g->ncj + g->nghost + 1) * N;
    memset(g->sf, 0, sizeof(double) * 3 * nslot);
    double e = 0.0, w = 0.0;
    long long off = 0;
    for (int ci = 0; ci < g->nci; ci++) {
        int ci_cj0 = (M == N) ? ci : ci >> 1;
        for (int k = 0; k < g->numneigh[ci]; k++) {
            int cj = g->nb[off + k];
            for (int cii = 0; cii < M; cii++) {
                int si           = ci * M + cii;
                const double* xi = &g->sx[3 * si];
                double fix = 0, fiy = 0, fiz = 0;
                for (int cjj = 0; cjj < N; cjj++) {
                    int sj = cj * N + cjj;
                    int cond;
                    if (M == N)
                        cond = half ? (ci_cj0 != cj || cii < cjj) : (ci_cj0 != cj || cii != cjj);
                    else
                        cond = half ? (ci_cj0 != cj || cii + M * (ci & 0x1) < cjj)
                                    : (ci_cj0 != cj || cii + M * (ci & 0x1) != cjj);
                    if (!cond) continue;
                    const double* xj = &g->sx[3 * sj];
                    double delx = xi[0] - xj[0];
                    double dely = xi[1] - xj[1];
                    double delz = xi[2] - xj[2];
                    double rsq  = delx * delx + dely * dely + delz * delz;
                    if (rsq < cutforcesq) {
                        double sr2   = 1.0 / rsq;
                        double sr6   = sr2 * sr2 * sr2 * sigma6;
                        double force = 48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon;
                        if (half) {
                            g->sf[3 * sj] -= delx * force;
                            g->sf[3 * sj + 1] -= dely * force;
                            g->sf[3 * sj + 2] -= delz * force;
                        }
                        fix += delx * force;
                        fiy += dely * force;
                        fiz += delz * force;
                        int local_j = cj < g->ncj;
                        double wgt  = (half && local_j) ? 1.0 : 0.5;
                        e += wgt * 4.0 * epsilon * sr6 * (sr6 - 1.0);
                        w += wgt * rsq * force;
                    }
                }
                g->sf[3 * si] += fix;
                g->sf[3 * si + 1] += fiy;
                g->sf[3 * si + 2] += fiz;
            }
        }
        off += g->numneigh[ci];
    }
    if (energy) *energy = e;
    if (virial) *virial = w;
}

/* initialIntegrateCPU / finalIntegrateCPU: src/clusterpair/integrate.c:29-86 */
void orc_cp_initial_integrate(orc_cp* g, double dt, double dtforce)
{
    for (int ci = 0; ci < g->nci; ci++)
        for (int cii = 0; cii < g->ci_natoms[ci]; cii++) {
            int s = ci * g->M + cii;
            for (int d = 0; d < 3; d++) {
                g->sv[3 * s + d] += dtforce * g->sf[3 * s + d];
                g->sx[3 * s + d] += dt * g->sv[3 * s + d];
            }
        }
}

void orc_cp_final_integrate(orc_cp* g, double dtforce)
{
    for (int ci = 0; ci < g->nci; ci++)
        for (int cii = 0; cii < g->ci_natoms[ci]; cii++) {
            int s = ci * g->M + cii;
            for (int d = 0; d < 3; d++)
                g->sv[3 * s + d] += dtforce * g->sf[3 * s + d];
        }
}

/* updateSingleAtoms (neighbor.c:2096-2150) + updateAtomsPbcCPU (pbc.c:120-145) */
void orc_cp_update_atoms(orc_cp* g)
{
    int n = 0;
    for (int ci = 0; ci < g->nci; ci++)
        for (int cii = 0; cii < g->ci_natoms[ci]; cii++) {
            int s = ci * g->M + cii;
            for (int d = 0; d < 3; d++) {
                g->ax[3 * n + d] = g->sx[3 * s + d];
                g->av[3 * n + d] = g->sv[3 * s + d];
            }
            g->at[n] = g->st[s];
            n++;
        }
    orc_update_atoms_pbc(n, g->ax, g->xprd, g->yprd, g->zprd);
}

double orc_cp_kinetic_sum(const orc_cp* g, double mass)
{
    double t = 0.0;
    for (int ci = 0; ci < g->nci; ci++)
        for (int cii = 0; cii < g->ci_natoms[ci]; cii++) {
            const double* v = &g->sv[3 * (ci * g->M + cii)];
            t += (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * mass;
        }
    return t;
}

int orc_cp_reneighbour(orc_cp* g, int half)
{
    if (orc_cp_build_clusters(g)) return -1;
    if (orc_cp_define_jclusters(g)) return -2;
    if (orc_cp_setup_pbc(g) < 0) return -3;
    if (orc_cp_bin_jclusters(g)) return -4;
    if (orc_cp_build_neighbor(g, half) < 0) return -5;
    return 0;
}

/* Timestep loop of src/clusterpair/main.c:326-379 (single type LJ); prune_every > 0 prunes the
   lists on steps between rebuilds, before the ghost update, as main.c:329-332 does. */
int orc_cp_run_prune(orc_cp* g, int half, double epsilon, double sigma6, double cutforcesq,
    double dt, double dtforce, int ntimes, int reneigh_every, int prune_every, double mass,
    double* ksum)
{
    int rc = orc_cp_reneighbour(g, half);
    if (rc) return rc;
    orc_cp_force_lj(g, half, epsilon, sigma6, cutforcesq, NULL, NULL);
    if (ksum) ksum[0] = orc_cp_kinetic_sum(g, mass);
    for (int n = 0; n < ntimes; n++) {
        orc_cp_initial_integrate(g, dt, dtforce);
        if ((n + 1) % reneigh_every) {
            if (prune_every > 0 && !((n + 1) % prune_every)) orc_cp_prune(g);
            orc_cp_update_pbc(g, 0);
        } else {
            orc_cp_update_atoms(g);
            rc = orc_cp_reneighbour(g, half);
            if (rc) return rc;
        }
        orc_cp_force_lj(g, half, epsilon, sigma6, cutforcesq, NULL, NULL);
        orc_cp_final_integrate(g, dtforce);
        if (ksum) ksum[n + 1] = orc_cp_kinetic_sum(g, mass);
    }
    return 0;
}

int orc_cp_run(orc_cp* g, int half, double epsilon, double sigma6, double cutforcesq, double dt,
    double dtforce, int ntimes, int reneigh_every, double mass, double* ksum)
{
    return orc_cp_run_prune(g, half, epsilon, sigma6, cutforcesq, dt, dtforce, ntimes,
        reneigh_every, 0, mass, ksum);
}

/* ------------------------------------------------------------------ accessors for tests */
/* info[]: M N nci ncj nghost dummy_cj mbins nstencil nbinx nbiny mbinx mbiny mbinxlo mbinylo
   natoms  (nb_total in info2) */
void orc_cp_info(const orc_cp* g, int* info, long long* nb_total)
{
    int v[15] = { g->M, g->N, g->nci, g->ncj, g->nghost, g->dummy_cj, g->mbins, g->nstencil,
        g->nbinx, g->nbiny, g->mbinx, g->mbiny, g->mbinxlo, g->mbinylo, g->natoms };
    memcpy(info, v, sizeof(v));
    if (nb_total) *nb_total = g->nb_total;
}

/* Copies a named array; returns the element count (or -1 for an unknown name, -2 if cap is too
   small).  Names: slot_x slot_v slot_f (double, 3 per slot), slot_t, ci_natoms, ci_bin,
   ci_bbox (6 per ci), cj_natoms, cj_bbox, ghost_map, ghost_pbc, numneigh, numneigh_masked,
   neighbors, imask, cbin_start, cbin_list, atoms_x, atoms_v, atoms_type, stencil. */
long long orc_cp_get(const orc_cp* g, const char* name, void* out, long long cap)
{
    int nslot = (g->ncj + g->nghost + 1) * g->N;
    int nj    = g->ncj + g->nghost;
    const void* src = NULL;
    long long n = 0, es = 4;
#define ORC_GET(nm, ptr, count, size)                                                          \
    if (!strcmp(name, nm)) {                                                                   \
        src = (ptr);                                                                           \
        n   = (count);                                                                         \
        es  = (size);                                                                          \
    }
    ORC_GET("slot_x", g->sx, 3LL * nslot, 8)
    ORC_GET("slot_v", g->sv, 3LL * g->nci * g->M, 8)
    ORC_GET("slot_f", g->sf, 3LL * nslot, 8)
    ORC_GET("slot_t", g->st, nslot, 4)
    ORC_GET("ci_natoms", g->ci_natoms, g->nci, 4)
    ORC_GET("ci_bin", g->ci_bin, g->nci, 4)
    ORC_GET("ci_bbox", g->ci_bbox, 6LL * g->nci, 8)
    ORC_GET("cj_natoms", g->cj_natoms, nj, 4)
    ORC_GET("cj_bbox", g->cj_bbox, 6LL * nj, 8)
    ORC_GET("ghost_map", g->border_map, g->nghost, 4)
    ORC_GET("ghost_pbc", g->pbc, 3LL * g->nghost, 4)
    ORC_GET("numneigh", g->numneigh, g->nci, 4)
    ORC_GET("numneigh_masked", g->numneigh_masked, g->nci, 4)
    ORC_GET("neighbors", g->nb, g->nb_total, 4)
    ORC_GET("imask", g->imask, g->nb_total, 4)
    ORC_GET("cbin_start", g->cbin_start, g->mbins + 1, 4)
    ORC_GET("cbin_list", g->cbin_list, g->cbin_start[g->mbins], 4)
    ORC_GET("atoms_x", g->ax, 3LL * g->natoms, 8)
    ORC_GET("atoms_v", g->av, 3LL * g->natoms, 8)
    ORC_GET("atoms_type", g->at, g->natoms, 4)
    ORC_GET("stencil", g->stencil, g->nstencil, 4)
#undef ORC_GET
    if (!src) return -1;
    if (n > cap) return -2;
    if (n) memcpy(out, src, n * es);
    return n;
}

/* Overwrites slot positions (3 per slot, first n slots), e.g. to evaluate forces on given
   coordinates. */
void orc_cp_set_slot_x(orc_cp* g, const double* x, int nslots)
{
    memcpy(g->sx, x, sizeof(double) * 3 * nslots);
}
