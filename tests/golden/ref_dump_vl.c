/*
 * Golden-vector writer for the Verlet-list path.
 *
 * Links the MD-Bench Verlet-list objects compiled by oracle/Makefile (strict IEEE build,
 * oracle/_ref/libmdref_vl.a) and drives the reference's own setup and timestep sequence
 * (src/verletlist/main.c:46-113 setup, :371-419 loop).  Every array the hot path produces is
 * written as a raw little-endian blob into <outdir>; make_golden.py packs them into .npz.
 *
 * usage: ref_dump_vl <outdir> <nx> <ny> <nz> <half> <nsteps> <reneigh_every>
 *                    [lj <types_file|->] | [eam <funcfl_file> <lattice_a> <temp>]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <allocate.h>
#include <atom.h>
#include <device.h>
#include <eam.h>
#include <force.h>
#include <integrate.h>
#include <neighbor.h>
#include <parameter.h>
#include <pbc.h>
#include <stats.h>
#include <thermo.h>
#include <util.h>

/* File-scope state of src/verletlist/neighbor.c and pbc.c (non-static there). */
extern int* bincount;
extern int* bins;
extern int mbins;
extern int atoms_per_bin;
extern int nstencil;
extern int* stencil;
extern int mbinx, mbiny, mbinz;
extern int pad_x, pad_y, pad_z;
extern MD_FLOAT binsizex, binsizey, binsizez;
extern MD_FLOAT bininvx, bininvy, bininvz;
extern int *PBCx, *PBCy, *PBCz;

static char g_out[1024];

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

static FILE* g_meta;
static void meta_i(const char* k, long long v) { fprintf(g_meta, "%s int %lld\n", k, v); }
static void meta_d(const char* k, double v) { fprintf(g_meta, "%s float %.17g\n", k, v); }

static double kinetic_sum(Atom* atom, Parameter* param)
{
    double t = 0.0;
    for (int i = 0; i < atom->Nlocal; i++) {
        t += (atom_vx(i) * atom_vx(i) + atom_vy(i) * atom_vy(i) + atom_vz(i) * atom_vz(i)) *
             param->mass;
    }
    return t;
}

static void dump_state(const char* tag, Atom* atom, Neighbor* neighbor, int with_nb)
{
    char name[256];
    int nall = atom->Nlocal + atom->Nghost;
    snprintf(name, sizeof(name), "%s_x", tag);
    dump(name, atom->x, sizeof(double) * 3 * nall);
    snprintf(name, sizeof(name), "%s_v", tag);
    dump(name, atom->vx, sizeof(double) * 3 * atom->Nlocal);
    snprintf(name, sizeof(name), "%s_f", tag);
    dump(name, atom->fx, sizeof(double) * 3 * atom->Nlocal);
    snprintf(name, sizeof(name), "%s_type", tag);
    dump(name, atom->type, sizeof(int) * nall);
    snprintf(name, sizeof(name), "%s_border_map", tag);
    dump(name, atom->border_map, sizeof(int) * atom->Nghost);
    int* pbc = malloc(sizeof(int) * 3 * (atom->Nghost + 1));
    for (int g = 0; g < atom->Nghost; g++) {
        pbc[3 * g + 0] = PBCx[g];
        pbc[3 * g + 1] = PBCy[g];
        pbc[3 * g + 2] = PBCz[g];
    }
    snprintf(name, sizeof(name), "%s_pbc", tag);
    dump(name, pbc, sizeof(int) * 3 * atom->Nghost);
    free(pbc);
    char key[256];
    snprintf(key, sizeof(key), "%s_nlocal", tag);
    meta_i(key, atom->Nlocal);
    snprintf(key, sizeof(key), "%s_nghost", tag);
    meta_i(key, atom->Nghost);
    if (with_nb) {
        snprintf(name, sizeof(name), "%s_numneigh", tag);
        dump(name, neighbor->numneigh, sizeof(int) * atom->Nlocal);
        /* Compact the AoS list (stride maxneighs) to a CSR layout. */
        long long tot = 0;
        for (int i = 0; i < atom->Nlocal; i++)
            tot += neighbor->numneigh[i];
        int* csr = malloc(sizeof(int) * (tot + 1));
        long long k = 0;
        for (int i = 0; i < atom->Nlocal; i++)
            for (int n = 0; n < neighbor->numneigh[i]; n++)
                csr[k++] = neighbor->neighbors[i * neighbor->maxneighs + n];
        snprintf(name, sizeof(name), "%s_neighbors", tag);
        dump(name, csr, sizeof(int) * tot);
        free(csr);
        snprintf(name, sizeof(name), "%s_bincount", tag);
        dump(name, bincount, sizeof(int) * mbins);
        int* bcsr = malloc(sizeof(int) * (nall + 1));
        k = 0;
        for (int b = 0; b < mbins; b++)
            for (int m = 0; m < bincount[b]; m++)
                bcsr[k++] = bins[b * atoms_per_bin + m];
        snprintf(name, sizeof(name), "%s_bins", tag);
        dump(name, bcsr, sizeof(int) * k);
        free(bcsr);
    }
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
    int is_eam          = strcmp(argv[8], "eam") == 0;
    if (!is_eam && argc > 9 && strcmp(argv[9], "-") != 0) {
        param.types_file = strdup(argv[9]);
        readTypesFile(&param);
    }
    param.cutneigh = param.cutforce + param.skin;

    if (is_eam) {
        param.force_field = FF_EAM;
        param.eam_file    = strdup(argv[9]);
        initEam(&param);
        /* initEam hard-codes rho/temp (eam_utils.c:22-40); take the lattice constant and
           temperature of the requested case instead, as a .dmp input would. */
        double a   = atof(argv[10]);
        param.rho  = 4.0 / (a * a * a);
        param.temp = atof(argv[11]);
        dump("eam_frho_spline", eam.frho_spline, sizeof(double) * eam.nrho_tot);
        dump("eam_rhor_spline", eam.rhor_spline, sizeof(double) * eam.nr_tot);
        dump("eam_z2r_spline", eam.z2r_spline, sizeof(double) * eam.nr_tot);
        meta_i("eam_nr", eam.nr);
        meta_i("eam_nrho", eam.nrho);
        meta_i("eam_nr_tot", eam.nr_tot);
        meta_i("eam_nrho_tot", eam.nrho_tot);
        meta_d("eam_dr", eam.dr);
        meta_d("eam_drho", eam.drho);
        meta_d("eam_rdr", eam.rdr);
        meta_d("eam_rdrho", eam.rdrho);
        meta_d("eam_mass", param.mass);
        meta_d("eam_cut", param.cutforce);
    }

    param.lattice = pow((4.0 / param.rho), (1.0 / 3.0));
    param.xprd    = param.nx * param.lattice;
    param.yprd    = param.ny * param.lattice;
    param.zprd    = param.nz * param.lattice;

    initAtom(&atom);
    initPbc(&atom);
    initStats(&stats);
    initNeighbor(&neighbor, &param);
    createAtom(&atom, &param);
    dump("create_x", atom.x, sizeof(double) * 3 * atom.Nlocal);
    dump("create_v", atom.vx, sizeof(double) * 3 * atom.Nlocal);
    dump("create_type", atom.type, sizeof(int) * atom.Nlocal);
    setupNeighbor(&param);
    setupThermo(&param, atom.Natoms);
    adjustThermo(&param, &atom);
    dump("adjust_v", atom.vx, sizeof(double) * 3 * atom.Nlocal);

    meta_d("lattice", param.lattice);
    meta_d("xprd", param.xprd);
    meta_d("yprd", param.yprd);
    meta_d("zprd", param.zprd);
    meta_d("cutforce", param.cutforce);
    meta_d("cutneigh", param.cutneigh);
    meta_d("dt", param.dt);
    meta_d("dtforce", param.dtforce);
    meta_d("mass", param.mass);
    meta_d("temp", param.temp);
    meta_i("ntypes", atom.ntypes);
    dump("epsilon", atom.epsilon, sizeof(double) * atom.ntypes * atom.ntypes);
    dump("sigma6", atom.sigma6, sizeof(double) * atom.ntypes * atom.ntypes);
    dump("cutforcesq", atom.cutforcesq, sizeof(double) * atom.ntypes * atom.ntypes);
    dump("cutneighsq", atom.cutneighsq, sizeof(double) * atom.ntypes * atom.ntypes);
    meta_i("mbinx", mbinx);
    meta_i("mbiny", mbiny);
    meta_i("mbinz", mbinz);
    meta_i("pad_x", pad_x);
    meta_i("pad_y", pad_y);
    meta_i("pad_z", pad_z);
    meta_d("binsizex", binsizex);
    meta_d("binsizey", binsizey);
    meta_d("binsizez", binsizez);
    meta_d("bininvx", bininvx);
    meta_d("bininvy", bininvy);
    meta_d("bininvz", bininvz);
    meta_i("nstencil", nstencil);
    dump("stencil", stencil, sizeof(int) * nstencil);

    setupPbc(&atom, &param);
    updatePbc(&atom, &param, true);
    buildNeighbor(&atom, &neighbor);
    initForce(&param);
    computeForce(&param, &atom, &neighbor, &stats);
    dump_state("s0", &atom, &neighbor, 1);

    /* Timestep loop of src/verletlist/main.c:371-419 (serial: forward = updatePbc,
       reverse = no-op). */
    double* tsum = malloc(sizeof(double) * (param.ntimes + 1));
    tsum[0]      = kinetic_sum(&atom, &param);
    for (int n = 0; n < param.ntimes; n++) {
        int reneigh = (n + 1) % param.reneigh_every == 0;
        initialIntegrate(reneigh, &param, &atom);
        if (reneigh) {
            updateAtomsPbc(&atom, &param, true);
            setupPbc(&atom, &param);
            updatePbc(&atom, &param, true);
            buildNeighbor(&atom, &neighbor);
        } else {
            updatePbc(&atom, &param, false);
        }
        computeForce(&param, &atom, &neighbor, &stats);
        finalIntegrate(reneigh, &param, &atom);
        tsum[n + 1] = kinetic_sum(&atom, &param);
    }
    dump("traj_ksum", tsum, sizeof(double) * (param.ntimes + 1));
    dump_state("end", &atom, &neighbor, 0);
    meta_i("ntimes", param.ntimes);
    meta_i("natoms", atom.Natoms);
    fclose(g_meta);
    return 0;
}
