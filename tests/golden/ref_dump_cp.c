/*
 * Golden-vector writer for the cluster-pair path.
 *
 * Links one of the strict MD-Bench cluster-pair builds of oracle/Makefile (M1xN2 plain-C
 * reference kernel, 4x4 or 4x8 SIMD variants) and drives the reference's setup and loop
 * (src/clusterpair/main.c:43-97 setup, :99-135 reneighbour/updateAtoms, :326-379 loop).
 * Writes raw blobs into <outdir>; make_golden.py packs them.
 *
 * usage: ref_dump_cp <outdir> <nx> <ny> <nz> <half> <nsteps> <reneigh_every> <dump_force> [prune]
 *
 * With prune = 1 the run stops after the first build: the local slots are displaced by a fixed
 * integer pattern (up to 0.25 per coordinate), the ghosts refreshed (updatePbc) and the lists
 * pruned (pruneNeighbor, src/clusterpair/neighbor.c:1099-1264); tag "pr" holds the result.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atom.h>
#include <force.h>
#include <integrate.h>
#include <neighbor.h>
#include <parameter.h>
#include <pbc.h>
#include <stats.h>
#include <thermo.h>
#include <util.h>

static char g_out[1024];
static FILE* g_meta;

static void dump(const char* name, const void* p, size_t bytes)
{
    char path[1200];
    snprintf(path, sizeof(path), "%s/%s.bin", g_out, name);
    FILE* fp = fopen(path, "wb");
    if (!fp) {
        perror(path);
        exit(2);
    }
    if (bytes) fwrite(p, 1, bytes, fp);
    fclose(fp);
}
static void meta_i(const char* k, long long v) { fprintf(g_meta, "%s int %lld\n", k, v); }
static void meta_d(const char* k, double v) { fprintf(g_meta, "%s float %.17g\n", k, v); }

/* Positions of every j-cluster slot (local, ghost, dummy) as AoS xyz per slot. */
static void dump_clusters(const char* tag, Atom* atom, Neighbor* nb, int with_nb, int with_f)
{
    char name[256], key[256];
    int ncj   = get_ncj_from_nci(atom->Nclusters_local);
    int njall = ncj + atom->Nclusters_ghost + 1; /* + dummy */
    int nslot = njall * CLUSTER_N;
    double* xs = malloc(sizeof(double) * 3 * nslot);
    int* ts    = malloc(sizeof(int) * nslot);
    for (int cj = 0; cj < njall; cj++) {
        double* cx = &atom->cl_x[CJ_VECTOR3_BASE_INDEX(cj)];
        int* ct    = &atom->cl_t[CJ_SCALAR_BASE_INDEX(cj)];
        for (int jj = 0; jj < CLUSTER_N; jj++) {
            int s          = cj * CLUSTER_N + jj;
            xs[3 * s]      = cx[CL_X_INDEX_3D(jj)];
            xs[3 * s + 1]  = cx[CL_Y_INDEX_3D(jj)];
            xs[3 * s + 2]  = cx[CL_Z_INDEX_3D(jj)];
            ts[s]          = ct[jj];
        }
    }
    snprintf(name, sizeof(name), "%s_slot_x", tag);
    dump(name, xs, sizeof(double) * 3 * nslot);
    snprintf(name, sizeof(name), "%s_slot_t", tag);
    dump(name, ts, sizeof(int) * nslot);
    free(xs);
    free(ts);
    /* velocities and forces of local i-clusters, per i-slot (ci*M + cii) */
    int nci    = atom->Nclusters_local;
    double* vs = malloc(sizeof(double) * 3 * nci * CLUSTER_M);
    double* fs = malloc(sizeof(double) * 3 * nci * CLUSTER_M);
    for (int ci = 0; ci < nci; ci++) {
        double* cv = &atom->cl_v[CI_VECTOR3_BASE_INDEX(ci)];
        double* cf = &atom->cl_f[CI_VECTOR3_BASE_INDEX(ci)];
        for (int ii = 0; ii < CLUSTER_M; ii++) {
            int s = ci * CLUSTER_M + ii;
            int live = ii < atom->iclusters[ci].natoms;
            vs[3 * s]     = live ? cv[CL_X_INDEX_3D(ii)] : 0.0;
            vs[3 * s + 1] = live ? cv[CL_Y_INDEX_3D(ii)] : 0.0;
            vs[3 * s + 2] = live ? cv[CL_Z_INDEX_3D(ii)] : 0.0;
            fs[3 * s]     = live ? cf[CL_X_INDEX_3D(ii)] : 0.0;
            fs[3 * s + 1] = live ? cf[CL_Y_INDEX_3D(ii)] : 0.0;
            fs[3 * s + 2] = live ? cf[CL_Z_INDEX_3D(ii)] : 0.0;
        }
    }
    snprintf(name, sizeof(name), "%s_slot_v", tag);
    dump(name, vs, sizeof(double) * 3 * nci * CLUSTER_M);
    if (with_f) {
        snprintf(name, sizeof(name), "%s_slot_f", tag);
        dump(name, fs, sizeof(double) * 3 * nci * CLUSTER_M);
    }
    free(vs);
    free(fs);
    /* cluster descriptors */
    double* ib = malloc(sizeof(double) * 6 * (nci + 1));
    int* in    = malloc(sizeof(int) * (nci + 1));
    for (int ci = 0; ci < nci; ci++) {
        Cluster* c    = &atom->iclusters[ci];
        in[ci]        = c->natoms;
        ib[6 * ci]    = c->bbminx;
        ib[6 * ci + 1] = c->bbmaxx;
        ib[6 * ci + 2] = c->bbminy;
        ib[6 * ci + 3] = c->bbmaxy;
        ib[6 * ci + 4] = c->bbminz;
        ib[6 * ci + 5] = c->bbmaxz;
    }
    snprintf(name, sizeof(name), "%s_ci_natoms", tag);
    dump(name, in, sizeof(int) * nci);
    snprintf(name, sizeof(name), "%s_ci_bbox", tag);
    dump(name, ib, sizeof(double) * 6 * nci);
    snprintf(name, sizeof(name), "%s_ci_bin", tag);
    dump(name, atom->cluster_bin, sizeof(int) * nci);
    free(ib);
    free(in);
    int nj     = ncj + atom->Nclusters_ghost;
    double* jb = malloc(sizeof(double) * 6 * (nj + 1));
    int* jn    = malloc(sizeof(int) * (nj + 1));
    for (int cj = 0; cj < nj; cj++) {
        Cluster* c     = &atom->jclusters[cj];
        jn[cj]         = c->natoms;
        jb[6 * cj]     = c->bbminx;
        jb[6 * cj + 1] = c->bbmaxx;
        jb[6 * cj + 2] = c->bbminy;
        jb[6 * cj + 3] = c->bbmaxy;
        jb[6 * cj + 4] = c->bbminz;
        jb[6 * cj + 5] = c->bbmaxz;
    }
    snprintf(name, sizeof(name), "%s_cj_natoms", tag);
    dump(name, jn, sizeof(int) * nj);
    snprintf(name, sizeof(name), "%s_cj_bbox", tag);
    dump(name, jb, sizeof(double) * 6 * nj);
    free(jb);
    free(jn);
    int ng   = atom->Nclusters_ghost;
    int* pbc = malloc(sizeof(int) * 3 * (ng + 1));
    for (int g = 0; g < ng; g++) {
        pbc[3 * g]     = atom->PBCx[g];
        pbc[3 * g + 1] = atom->PBCy[g];
        pbc[3 * g + 2] = atom->PBCz[g];
    }
    snprintf(name, sizeof(name), "%s_ghost_pbc", tag);
    dump(name, pbc, sizeof(int) * 3 * ng);
    snprintf(name, sizeof(name), "%s_ghost_map", tag);
    dump(name, atom->border_map, sizeof(int) * ng);
    free(pbc);
    snprintf(key, sizeof(key), "%s_nci", tag);
    meta_i(key, nci);
    snprintf(key, sizeof(key), "%s_ncj", tag);
    meta_i(key, ncj);
    snprintf(key, sizeof(key), "%s_nghost_clusters", tag);
    meta_i(key, ng);
    snprintf(key, sizeof(key), "%s_nghost_atoms", tag);
    meta_i(key, atom->Nghost);
    snprintf(key, sizeof(key), "%s_dummy_cj", tag);
    meta_i(key, atom->dummy_cj);
    if (with_nb) {
        long long tot = 0;
        for (int ci = 0; ci < nci; ci++)
            tot += nb->numneigh[ci];
        int* csr = malloc(sizeof(int) * (tot + 1));
        unsigned* msk = malloc(sizeof(unsigned) * (tot + 1));
        long long k = 0;
        for (int ci = 0; ci < nci; ci++)
            for (int n = 0; n < nb->numneigh[ci]; n++) {
                csr[k]   = nb->neighbors[ci * nb->maxneighs + n];
                msk[k++] = nb->neighbors_imask[ci * nb->maxneighs + n];
            }
        snprintf(name, sizeof(name), "%s_numneigh", tag);
        dump(name, nb->numneigh, sizeof(int) * nci);
        snprintf(name, sizeof(name), "%s_numneigh_masked", tag);
        dump(name, nb->numneigh_masked, sizeof(int) * nci);
        snprintf(name, sizeof(name), "%s_neighbors", tag);
        dump(name, csr, sizeof(int) * tot);
        snprintf(name, sizeof(name), "%s_imask", tag);
        dump(name, msk, sizeof(unsigned) * tot);
        free(csr);
        free(msk);
    }
}

static double ksum(Atom* atom, Parameter* param)
{
    /* sum m v^2 over local cluster atoms in cluster order */
    double t = 0.0;
    for (int ci = 0; ci < atom->Nclusters_local; ci++) {
        double* cv = &atom->cl_v[CI_VECTOR3_BASE_INDEX(ci)];
        for (int ii = 0; ii < atom->iclusters[ci].natoms; ii++)
            t += (cv[CL_X_INDEX_3D(ii)] * cv[CL_X_INDEX_3D(ii)] +
                     cv[CL_Y_INDEX_3D(ii)] * cv[CL_Y_INDEX_3D(ii)] +
                     cv[CL_Z_INDEX_3D(ii)] * cv[CL_Z_INDEX_3D(ii)]) *
                 param->mass;
    }
    return t;
}

int main(int argc, char** argv)
{
    if (argc < 9) {
        fprintf(stderr, "usage: see source\n");
        return 1;
    }
    snprintf(g_out, sizeof(g_out), "%s", argv[1]);
    char mpath[1200];
    snprintf(mpath, sizeof(mpath), "%s/meta.txt", g_out);
    g_meta = fopen(mpath, "w");
    Parameter param;
    Atom atom;
    Neighbor neighbor;
    Stats stats;
    initParameter(&param);
    param.nx            = atoi(argv[2]);
    param.ny            = atoi(argv[3]);
    param.nz            = atoi(argv[4]);
    param.half_neigh    = atoi(argv[5]);
    param.ntimes        = atoi(argv[6]);
    param.reneigh_every = atoi(argv[7]);
    int dump_force      = atoi(argv[8]);
    int prune_test      = argc > 9 ? atoi(argv[9]) : 0;
    param.cutneigh      = param.cutforce + param.skin;
    param.lattice       = pow((4.0 / param.rho), (1.0 / 3.0));
    param.xprd          = param.nx * param.lattice;
    param.yprd          = param.ny * param.lattice;
    param.zprd          = param.nz * param.lattice;
    meta_i("cluster_m", CLUSTER_M);
    meta_i("cluster_n", CLUSTER_N);
    meta_d("lattice", param.lattice);
    meta_d("cutneigh", param.cutneigh);
    meta_d("dt", param.dt);
    meta_d("dtforce", param.dtforce);

    initAtom(&atom);
    initForce(&param);
    initPbc(&atom);
    initStats(&stats);
    initNeighbor(&neighbor, &param);
    createAtom(&atom, &param);
    setupNeighbor(&param, &atom);
    setupThermo(&param, atom.Natoms);
    adjustThermo(&param, &atom);
    double* ax = malloc(sizeof(double) * 3 * atom.Nlocal);
    double* av = malloc(sizeof(double) * 3 * atom.Nlocal);
    for (int i = 0; i < atom.Nlocal; i++) {
        ax[3 * i]     = atom.x[3 * i];
        ax[3 * i + 1] = atom.x[3 * i + 1];
        ax[3 * i + 2] = atom.x[3 * i + 2];
        av[3 * i]     = atom.vx[i];
        av[3 * i + 1] = atom.vy[i];
        av[3 * i + 2] = atom.vz[i];
    }
    dump("atoms_x", ax, sizeof(double) * 3 * atom.Nlocal);
    dump("atoms_v", av, sizeof(double) * 3 * atom.Nlocal);
    dump("atoms_type", atom.type, sizeof(int) * atom.Nlocal);
    meta_i("natoms", atom.Natoms);

    buildClusters(&atom);
    defineJClusters(&param, &atom);
    setupPbc(&atom, &param);
    binJClusters(&param, &atom);
    buildNeighbor(&atom, &neighbor);
    computeForce(&param, &atom, &neighbor, &stats);
    dump_clusters("s0", &atom, &neighbor, 1, dump_force);
    if (prune_test) {
        for (int ci = 0; ci < atom.Nclusters_local; ci++) {
            double* cx = &atom.cl_x[CI_VECTOR3_BASE_INDEX(ci)];
            for (int ii = 0; ii < atom.iclusters[ci].natoms; ii++) {
                int s = ci * CLUSTER_M + ii;
                cx[CL_X_INDEX_3D(ii)] += 0.25 * (double)((s * 7) % 9 - 4) / 4.0;
                cx[CL_Y_INDEX_3D(ii)] += 0.25 * (double)((s * 5) % 7 - 3) / 3.0;
                cx[CL_Z_INDEX_3D(ii)] += 0.25 * (double)((s * 3) % 5 - 2) / 2.0;
            }
        }
        updatePbc(&atom, &param, false);
        pruneNeighbor(&param, &atom, &neighbor);
        dump_clusters("pr", &atom, &neighbor, 1, 0);
        fclose(g_meta);
        return 0;
    }

    double* ks = malloc(sizeof(double) * (param.ntimes + 1));
    ks[0]      = ksum(&atom, &param);
    for (int n = 0; n < param.ntimes; n++) {
        initialIntegrate(&param, &atom);
        if ((n + 1) % param.reneigh_every) {
            updatePbc(&atom, &param, false);
        } else {
            updateSingleAtoms(&param, &atom);
            updateAtomsPbc(&atom, &param, false);
            if (n + 1 == param.reneigh_every) {
                /* atoms handed to the first rebuild (cluster order, wrapped) */
                for (int i = 0; i < atom.Nlocal; i++) {
                    ax[3 * i]     = atom.x[3 * i];
                    ax[3 * i + 1] = atom.x[3 * i + 1];
                    ax[3 * i + 2] = atom.x[3 * i + 2];
                    av[3 * i]     = atom.vx[i];
                    av[3 * i + 1] = atom.vy[i];
                    av[3 * i + 2] = atom.vz[i];
                }
                dump("pre1_x", ax, sizeof(double) * 3 * atom.Nlocal);
                dump("pre1_v", av, sizeof(double) * 3 * atom.Nlocal);
                dump("pre1_type", atom.type, sizeof(int) * atom.Nlocal);
            }
            buildClusters(&atom);
            defineJClusters(&param, &atom);
            setupPbc(&atom, &param);
            binJClusters(&param, &atom);
            buildNeighbor(&atom, &neighbor);
        }
        computeForce(&param, &atom, &neighbor, &stats);
        finalIntegrate(&param, &atom);
        ks[n + 1] = ksum(&atom, &param);
        if (n + 1 == param.reneigh_every) dump_clusters("r1", &atom, &neighbor, 1, dump_force);
    }
    dump("traj_ksum", ks, sizeof(double) * (param.ntimes + 1));
    dump_clusters("end", &atom, &neighbor, 0, dump_force);
    meta_i("ntimes", param.ntimes);
    fclose(g_meta);
    return 0;
}
