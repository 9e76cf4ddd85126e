/*
 * Host-side setup of the gfx950 driver: parameters, fcc lattice with Park-Miller velocities,
 * thermo scaling.  Runs once per simulation; everything per-step lives on the device.
 *
 * Results must be bit-identical to MD-Bench's setup (same lattice order, same velocities, same
 * glibc rand() type sequence), so the arithmetic follows the reference expression by expression:
 *   initParameter       src/common/parameter.c:72-124
 *   per-type / pair LJ  src/common/parameter.c:31-70, src/verletlist/atom.c:121-142
 *   createAtom          src/verletlist/atom.c:104-246
 *   myrandom            src/common/util.c:24-33
 *   setupThermo/adjustThermo  src/common/thermo.c:35-58, 102-159
 */
#include <math.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

/* Park-Miller minimal standard generator, Schrage factorisation (no masking). */
double mdbMyrandom(int* seed)
{
    const int a = 16807, m = 2147483647, q = 127773, r = 2836;
    int k = (*seed) / q;
    *seed = a * (*seed - k * q) - r * k;
    if (*seed < 0) *seed += m;
    return (1.0 / m) * (*seed);
}

/* splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number generators", OOPSLA
   2014): 64-bit state, one output per call */
static unsigned long long splitmix64(unsigned long long* s)
{
    unsigned long long z = (*s += 0x9e3779b97f4a7c15ULL);
    z                    = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z                    = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* uniform in (0, 1]: the top 53 bits, shifted off zero so that the logarithm is finite */
static double unit_open0(unsigned long long* s)
{
    return ((double)(splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

void mdbMaxwellNormals(int n, unsigned long long seed, double* v)
{
    /* Box-Muller (Box and Muller, Ann. Math. Statist. 29, 1958): two uniforms give two
       independent standard normals */
    unsigned long long s = seed;
    const long long m    = 3LL * n;
    for (long long k = 0; k < m; k += 2) {
        const double u1 = unit_open0(&s), u2 = unit_open0(&s);
        const double r  = sqrt(-2.0 * log(u1));
        const double t  = 6.283185307179586 * u2;
        v[k]            = r * cos(t);
        if (k + 1 < m) v[k + 1] = r * sin(t);
    }
}

void mdbDefaultConfig(mdb_config* c)
{
    memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 32;
    c->force_field   = FF_LJ;
    c->ntimes        = 200;
    c->nstat         = 100;
    c->reneigh_every = 20;
    c->skin          = -1.0;
}

void mdb_init_parameter(Parameter* p)
{
    memset(p, 0, sizeof(*p));
    p->force_field      = FF_LJ;
    p->epsilon          = 1.0;
    p->sigma            = 1.0;
    p->sigma6           = 1.0;
    p->rho              = 0.8442;
    p->ntypes           = 4;
    p->ntimes           = 200;
    p->dt               = 0.005;
    p->nx = p->ny = p->nz = 32;
    p->pbc_x = p->pbc_y = p->pbc_z = 1;
    p->cutforce         = 2.5;
    p->skin             = 0.3;
    p->cutneigh         = p->cutforce + p->skin;
    p->temp             = 1.44;
    p->nstat            = 100;
    p->mass             = 1.0;
    p->dtforce          = 0.5 * p->dt;
    p->reneigh_every    = 20;
    p->resort_every     = 400;
    p->prune_every      = 1000;
    p->x_out_every      = 20;
    p->v_out_every      = 5;
    p->half_neigh       = 0;
    p->proc_freq        = 2.4;
    p->balance_every    = p->reneigh_every;
    p->setup            = 1;
}

/* Per-type sqrt(eps), sigma^3 and the geometric type-pair tables. */
static void type_tables(const Parameter* p, Atom* a)
{
    int nt                     = a->ntypes;
    a->sqrt_epsilon_per_type   = malloc(sizeof(double) * nt);
    a->sigma3_per_type         = malloc(sizeof(double) * nt);
    a->epsilon                 = malloc(sizeof(double) * nt * nt);
    a->sigma6                  = malloc(sizeof(double) * nt * nt);
    a->cutforcesq              = malloc(sizeof(double) * nt * nt);
    a->cutneighsq              = malloc(sizeof(double) * nt * nt);
    for (int t = 0; t < nt; t++) {
        int have = t < p->ntypes;
        double e = (have && p->epsilon_per_type) ? p->epsilon_per_type[t] : p->epsilon;
        double s = (have && p->sigma_per_type) ? p->sigma_per_type[t] : p->sigma;
        a->sqrt_epsilon_per_type[t] = sqrt(e);
        a->sigma3_per_type[t]       = s * s * s;
    }
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++) {
            a->epsilon[i * nt + j] = a->sqrt_epsilon_per_type[i] * a->sqrt_epsilon_per_type[j];
            a->sigma6[i * nt + j]  = a->sigma3_per_type[i] * a->sigma3_per_type[j];
        }
    for (int k = 0; k < nt * nt; k++) {
        a->cutneighsq[k] = p->cutneigh * p->cutneigh;
        a->cutforcesq[k] = p->cutforce * p->cutforce;
    }
}

/* fcc lattice in 8^3 sub-box order with per-site Park-Miller velocities. */
int mdb_create_atoms(Parameter* p, Atom* a)
{
    const int sub = 8;
    a->Natoms     = 4 * p->nx * p->ny * p->nz;
    a->Nlocal     = 0;
    a->Nghost     = 0;
    a->Nmax       = a->Natoms;
    a->ntypes     = p->ntypes;
    type_tables(p, a);
    a->x            = calloc(3 * (size_t)a->Nmax, sizeof(double));
    a->vx           = calloc(3 * (size_t)a->Nmax, sizeof(double));
    a->fx           = calloc(3 * (size_t)a->Nmax, sizeof(double));
    a->type         = calloc((size_t)a->Nmax, sizeof(int));
    a->sqrt_epsilon = calloc((size_t)a->Nmax, sizeof(double));
    a->sigma3       = calloc((size_t)a->Nmax, sizeof(double));
    if (!a->x || !a->vx || !a->fx || !a->type) return -1;

    double xlo = 0.0, xhi = p->xprd, ylo = 0.0, yhi = p->yprd, zlo = 0.0, zhi = p->zprd;
    double alat = pow((4.0 / p->rho), (1.0 / 3.0));
    int ilo = (int)(xlo / (0.5 * alat) - 1), ihi = (int)(xhi / (0.5 * alat) + 1);
    int jlo = (int)(ylo / (0.5 * alat) - 1), jhi = (int)(yhi / (0.5 * alat) + 1);
    int klo = (int)(zlo / (0.5 * alat) - 1), khi = (int)(zhi / (0.5 * alat) + 1);
    if (ilo < 0) ilo = 0;
    if (jlo < 0) jlo = 0;
    if (klo < 0) klo = 0;
    if (ihi > 2 * p->nx - 1) ihi = 2 * p->nx - 1;
    if (jhi > 2 * p->ny - 1) jhi = 2 * p->ny - 1;
    if (khi > 2 * p->nz - 1) khi = 2 * p->nz - 1;

    /* The type sequence is glibc rand() from its initial state, as in a fresh process. */
    srand(1);
    for (int oz = 0; oz * sub <= khi; oz++)
        for (int oy = 0; oy * sub <= jhi; oy++)
            for (int ox = 0; ox * sub <= ihi; ox++)
                for (int sz = 0; sz < sub; sz++)
                    for (int sy = 0; sy < sub; sy++)
                        for (int sx = 0; sx < sub; sx++) {
                            int k = oz * sub + sz, j = oy * sub + sy, i = ox * sub + sx;
                            if ((i + j + k) % 2 != 0) continue;
                            if (i < ilo || i > ihi || j < jlo || j > jhi || k < klo || k > khi)
                                continue;
                            double xt = 0.5 * alat * i, yt = 0.5 * alat * j, zt = 0.5 * alat * k;
                            if (!(xt >= xlo && xt < xhi && yt >= ylo && yt < yhi && zt >= zlo &&
                                    zt < zhi))
                                continue;
                            int seed = k * (2 * p->ny) * (2 * p->nx) + j * (2 * p->nx) + i + 1;
                            double vv[3];
                            for (int d = 0; d < 3; d++) {
                                for (int w = 0; w < 5; w++)
                                    mdbMyrandom(&seed);
                                vv[d] = mdbMyrandom(&seed);
                            }
                            if (a->Nlocal >= a->Nmax) return -2;
                            int n            = a->Nlocal++;
                            a->x[3 * n]      = xt;
                            a->x[3 * n + 1]  = yt;
                            a->x[3 * n + 2]  = zt;
                            a->vx[3 * n]     = vv[0];
                            a->vx[3 * n + 1] = vv[1];
                            a->vx[3 * n + 2] = vv[2];
                            int t            = rand() % a->ntypes;
                            a->type[n]       = t;
                            a->sqrt_epsilon[n] = a->sqrt_epsilon_per_type[t];
                            a->sigma3[n]       = a->sigma3_per_type[t];
                        }
    return 0;
}

/* readAtom_dmp (src/verletlist/atom.c:529-656): atoms, box and types from a dump file; the
   box replaces x/y/zprd, ntypes becomes the largest type + 1. */
int mdb_load_atoms(Parameter* p, Atom* a, const char* path)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    /* count for the allocation: the atom count line follows its item line */
    char line[256];
    int n = -1;
    while (fgets(line, sizeof(line), fp)) {
        if (strncmp(line, "ITEM: NUMBER OF ATOMS", 21) == 0) {
            if (fgets(line, sizeof(line), fp)) n = atoi(line);
            break;
        }
    }
    fclose(fp);
    if (n <= 0) return -3;
    a->Natoms = n;
    a->Nlocal = 0;
    a->Nghost = 0;
    a->Nmax   = n;
    a->x      = calloc(3 * (size_t)n, sizeof(double));
    a->vx     = calloc(3 * (size_t)n, sizeof(double));
    a->fx     = calloc(3 * (size_t)n, sizeof(double));
    a->type   = calloc((size_t)n, sizeof(int));
    a->sqrt_epsilon = calloc((size_t)n, sizeof(double));
    a->sigma3       = calloc((size_t)n, sizeof(double));
    if (!a->x || !a->vx || !a->fx || !a->type || !a->sqrt_epsilon || !a->sigma3) return -2;
    double box[6];
    int nt    = 0;
    int got   = mdbReadDmp(path, n, a->x, a->vx, a->type, box, &nt);
    if (got != n) return got < 0 ? got : -3;
    a->Nlocal = n;
    p->xlo    = box[0];
    p->xhi    = box[1];
    p->ylo    = box[2];
    p->yhi    = box[3];
    p->zlo    = box[4];
    p->zhi    = box[5];
    p->xprd   = p->xhi - p->xlo;
    p->yprd   = p->yhi - p->ylo;
    p->zprd   = p->zhi - p->zlo;
    a->ntypes = nt;
    type_tables(p, a);
    for (int i = 0; i < n; i++) {
        a->sqrt_epsilon[i] = a->sqrt_epsilon_per_type[a->type[i]];
        a->sigma3[i]       = a->sigma3_per_type[a->type[i]];
    }
    return 0;
}

void mdb_setup_thermo(mdb_thermo* th, Parameter* p, int natoms)
{
    if (p->force_field == FF_EAM) {
        th->mvv2e     = 1.036427e-04;
        th->dof_boltz = (natoms * 3 - 3) * 8.617343e-05;
        th->t_scale   = th->mvv2e / th->dof_boltz;
        th->p_scale   = 1.602176e+06 / 3 / p->xprd / p->yprd / p->zprd;
        p->dtforce /= th->mvv2e;
    } else {
        th->mvv2e     = 1.0;
        th->dof_boltz = (natoms * 3 - 3);
        th->t_scale   = th->mvv2e / th->dof_boltz;
        th->p_scale   = 1.0 / 3 / p->xprd / p->yprd / p->zprd;
    }
}

/* Zero the centre-of-mass velocity and rescale to the target temperature. */
void mdb_adjust_thermo(const mdb_thermo* th, const Parameter* p, Atom* a)
{
    double vt[3] = { 0.0, 0.0, 0.0 };
    for (int i = 0; i < a->Nlocal; i++)
        for (int d = 0; d < 3; d++)
            vt[d] += a->vx[3 * i + d];
    for (int d = 0; d < 3; d++)
        vt[d] = vt[d] / a->Natoms;
    for (int i = 0; i < a->Nlocal; i++)
        for (int d = 0; d < 3; d++)
            a->vx[3 * i + d] -= vt[d];
    double t = 0.0;
    for (int i = 0; i < a->Nlocal; i++)
        t += (a->vx[3 * i] * a->vx[3 * i] + a->vx[3 * i + 1] * a->vx[3 * i + 1] +
                 a->vx[3 * i + 2] * a->vx[3 * i + 2]) *
             p->mass;
    t *= th->t_scale;
    double factor = sqrt(p->temp / t);
    for (int i = 0; i < 3 * a->Nlocal; i++)
        a->vx[i] *= factor;
}

double mdb_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
