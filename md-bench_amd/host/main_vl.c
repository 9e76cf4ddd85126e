/*
 * mdbench-gfx950-vl: the Verlet-list benchmark driver on one MI355X.
 * Command line and output follow src/verletlist/main.c:206-318 (options) and
 * src/common/thermo.c:97 / src/verletlist/main.c:442-529 (thermo table, performance report).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mdbench_driver.h"
#include "mdgfx950.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void usage(void)
{
    printf("mdbench-gfx950-vl: MD-Bench Verlet-list force step on gfx950\n"
           "-n / --nsteps <int>   timesteps (200)\n"
           "-nx/-ny/-nz <int>     fcc unit cells per dimension (32)\n"
           "-half <int>           half (1) or full (0) neighbour lists\n"
           "-f <lj|eam>           force field\n"
           "-e <file>             EAM funcfl file (default: synthetic Cu-like tables)\n"
           "-a <real>             fcc lattice constant (overrides the density)\n"
           "-t <real>             initial temperature\n"
           "-r / --radius <real>  force cutoff\n"
           "-s / --skin <real>    skin\n"
           "--reneigh <int>       reneighbour every n steps\n"
           "--nstat <int>         thermo output every n steps\n"
           "--strict              strict (bit-reproducible FP64) kernels\n");
}

int main(int argc, char** argv)
{
    mdb_config cfg;
    mdbDefaultConfig(&cfg);
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        int more      = i + 1 < argc;
        if ((!strcmp(a, "-n") || !strcmp(a, "--nsteps")) && more) cfg.ntimes = atoi(argv[++i]);
        else if (!strcmp(a, "-nx") && more) cfg.nx = atoi(argv[++i]);
        else if (!strcmp(a, "-ny") && more) cfg.ny = atoi(argv[++i]);
        else if (!strcmp(a, "-nz") && more) cfg.nz = atoi(argv[++i]);
        else if (!strcmp(a, "-half") && more) cfg.half_neigh = atoi(argv[++i]);
        else if (!strcmp(a, "-f") && more) {
            const char* ff = argv[++i];
            if (!strncmp(ff, "lj", 2)) cfg.force_field = FF_LJ;
            else if (!strncmp(ff, "eam", 3)) cfg.force_field = FF_EAM;
            else {
                fprintf(stderr, "Invalid force field!\n");
                return 1;
            }
        } else if (!strcmp(a, "-e") && more) cfg.eam_file = argv[++i];
        else if (!strcmp(a, "-a") && more) cfg.lattice_a = atof(argv[++i]);
        else if (!strcmp(a, "-t") && more) cfg.temp = atof(argv[++i]);
        else if ((!strcmp(a, "-r") || !strcmp(a, "--radius")) && more) cfg.cutforce = atof(argv[++i]);
        else if ((!strcmp(a, "-s") || !strcmp(a, "--skin")) && more) cfg.skin = atof(argv[++i]);
        else if (!strcmp(a, "--reneigh") && more) cfg.reneigh_every = atoi(argv[++i]);
        else if (!strcmp(a, "--nstat") && more) cfg.nstat = atoi(argv[++i]);
        else if (!strcmp(a, "--strict")) cfg.strict = 1;
        else if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
            usage();
            return 0;
        } else {
            fprintf(stderr, "unknown option %s\n", a);
            usage();
            return 1;
        }
    }

    double t_setup = now();
    mdb_vl* sim;
    int rc = mdbVLCreate(&cfg, &sim);
    if (rc) {
        fprintf(stderr, "setup failed (%d): %s\n", rc, mdgfx950StatusMessage());
        return 2;
    }
    mdgfx950Synchronize();
    t_setup        = now() - t_setup;
    Parameter* p   = mdbVLParam(sim);
    Atom* atom     = mdbVLAtom(sim);
    printf("Parameters: %s, %d x %d x %d cells, %d atoms, lattice %.6f, cutforce %.4f, "
           "cutneigh %.4f, dt %g, %s list, %s kernels\n",
        p->force_field == FF_EAM ? "eam" : "lj", p->nx, p->ny, p->nz, atom->Natoms, p->lattice,
        p->cutforce, p->cutneigh, p->dt, p->half_neigh ? "half" : "full",
        cfg.strict ? "strict" : "fast");
    printf("\nSIMULATION PROGRESS\n");
    printf("-------------------------------------------------------------------------------\n");
    printf("  %-10s %15s %15s\n", "Step", "Temperature", "Pressure");
    double temp, press;
    mdbThermo(sim, kineticSumGFX950(p, atom), &temp, &press);
    printf("  %-10d %15.6e %15.6e\n", 0, temp, press);

    mdgfx950EnableTiming(1);
    double t_total = now();
    int done = 0;
    while (done < p->ntimes && rc == 0) {
        int chunk = p->nstat - (done % p->nstat);
        if (done + chunk > p->ntimes) chunk = p->ntimes - done;
        rc = mdbVLRun(sim, chunk, NULL);
        done += chunk;
        if (done % p->nstat == 0 && done < p->ntimes) {
            mdbThermo(sim, kineticSumGFX950(p, atom), &temp, &press);
            printf("  %-10d %15.6e %15.6e\n", done, temp, press);
        }
    }
    mdgfx950Synchronize();
    t_total = now() - t_total;
    if (rc) {
        fprintf(stderr, "run failed (%d): %s\n", rc, mdgfx950StatusMessage());
        mdbVLDestroy(sim);
        return 3;
    }
    mdbThermo(sim, kineticSumGFX950(p, atom), &temp, &press);
    printf("  %-10d %15.6e %15.6e\n", p->ntimes, temp, press);
    printf("-------------------------------------------------------------------------------\n");

    const char* fam[] = { p->force_field == FF_EAM ? "force_eam" : "force_lj", "neighbor",
        "binning", "pbc", "integrate", "thermo" };
    const char* lab[] = { "Force", "Neighbor", "Binning", "Ghosts/PBC", "Integration", "Thermo" };
    printf("\nPERFORMANCE REPORT (device kernel time from HIP events)\n");
    for (int k = 0; k < 6; k++) {
        double ms    = 0;
        long long nl = 0;
        mdgfx950KernelTime(fam[k], &ms, &nl);
        printf("    %-20s %10.4f s  %7.2f%%  (%lld launches)\n", lab[k], ms * 1e-3,
            100.0 * ms * 1e-3 / t_total, nl);
    }
    printf("\n  System: %d atoms (%d ghost) | %d timesteps | setup %.2f s\n", atom->Natoms,
        atom->Nghost, p->ntimes, t_setup);
    printf("  Performance: %.4fs total | %.2f atom updates/us | %.2f steps/s\n", t_total,
        (double)atom->Natoms * p->ntimes / (t_total * 1e6), p->ntimes / t_total);
    mdbVLDestroy(sim);
    return 0;
}
