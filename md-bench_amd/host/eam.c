/*
 * EAM tables for the gfx950 driver: funcfl reader, resampling to the working grid and the
 * 7-coefficient cubic splines the force kernels evaluate.  One element, so every type pair
 * shares one table.  Arithmetic follows src/common/eam_utils.c (readEamFile :42-93,
 * file2array :95-220, array2spline :222-251, interpolate :253-284) so the tables match the
 * reference's bit for bit.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

/* Reads n whitespace-separated numbers, whole lines at a time. */
static int read_values(FILE* fp, int n, double* out)
{
    char line[4096];
    int got = 0;
    while (got < n) {
        if (!fgets(line, sizeof(line), fp)) return -1;
        for (char* tok = strtok(line, " \t\n\r\f"); tok; tok = strtok(NULL, " \t\n\r\f")) {
            if (got < n) out[got] = atof(tok);
            got++;
        }
    }
    return 0;
}

/* Lagrange four-point resampling of a 1-based table with spacing h at abscissa r. */
static double lagrange4(const double* tab, int n, double h, double r)
{
    const double sixth = 1.0 / 6.0;
    double p           = r / h + 1.0;
    int k              = (int)(p);
    if (k > n - 2) k = n - 2;
    if (k < 2) k = 2;
    p -= k;
    if (p > 2.0) p = 2.0;
    double c1 = -sixth * p * (p - 1.0) * (p - 2.0);
    double c2 = 0.5 * (p * p - 1.0) * (p - 2.0);
    double c3 = -0.5 * p * (p + 1.0) * (p - 2.0);
    double c4 = sixth * p * (p * p - 1.0);
    return c1 * tab[k - 1] + c2 * tab[k] + c3 * tab[k + 1] + c4 * tab[k + 2];
}

/* Knot m holds {3c3/h, 2c2/h, c1/h, c3, c2, c1, f(m)}: value ((c3 p + c2) p + c1) p + f and
   derivative (3c3/h p + 2c2/h) p + c1/h for p in [0,1]. */
static void spline7(int n, double h, const double* f, double* s)
{
#define S(m, c) s[(m) * 7 + (c)]
    for (int m = 1; m <= n; m++)
        S(m, 6) = f[m];
    S(1, 5)     = S(2, 6) - S(1, 6);
    S(2, 5)     = 0.5 * (S(3, 6) - S(1, 6));
    S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
    S(n, 5)     = S(n, 6) - S(n - 1, 6);
    for (int m = 3; m <= n - 2; m++)
        S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;
    for (int m = 1; m <= n - 1; m++) {
        S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
        S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
    }
    S(n, 4) = 0.0;
    S(n, 3) = 0.0;
    for (int m = 1; m <= n; m++) {
        S(m, 2) = S(m, 5) / h;
        S(m, 1) = 2.0 * S(m, 4) / h;
        S(m, 0) = 3.0 * S(m, 3) / h;
    }
#undef S
}

int mdbEamLoad(const char* path, mdb_eam_tables* t)
{
    memset(t, 0, sizeof(*t));
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char line[4096];
    int ielem, nrho_f, nr_f;
    double drho_f, dr_f;
    int rc = -2;
    double *F = NULL, *Z = NULL, *R = NULL;
    if (!fgets(line, sizeof(line), fp) || !fgets(line, sizeof(line), fp)) goto out;
    if (sscanf(line, "%d %lg", &ielem, &t->mass) != 2) goto out;
    if (!fgets(line, sizeof(line), fp)) goto out;
    if (sscanf(line, "%d %lg %d %lg %lg", &nrho_f, &drho_f, &nr_f, &dr_f, &t->cut) != 5) goto out;
    if (nrho_f < 5 || nr_f < 5) goto out;
    /* 1-based copies (index 0 unused), a little slack for whole-line reads */
    F = calloc(nrho_f + 16, sizeof(double));
    Z = calloc(nr_f + 16, sizeof(double));
    R = calloc(nr_f + 16, sizeof(double));
    if (read_values(fp, nrho_f, F + 1) || read_values(fp, nr_f, Z + 1) || read_values(fp, nr_f, R + 1))
        goto out;

    t->dr   = dr_f;
    t->drho = drho_f;
    t->nr   = (int)(((nr_f - 1) * dr_f) / t->dr + 0.5);
    t->nrho = (int)(((nrho_f - 1) * drho_f) / t->drho + 0.5);
    double* frho = calloc(t->nrho + 2, sizeof(double));
    double* rhor = calloc(t->nr + 2, sizeof(double));
    double* z2r  = calloc(t->nr + 2, sizeof(double));
    for (int m = 1; m <= t->nrho; m++)
        frho[m] = lagrange4(F, nrho_f, drho_f, (m - 1) * t->drho);
    for (int m = 1; m <= t->nr; m++) {
        double r = (m - 1) * t->dr;
        rhor[m]  = lagrange4(R, nr_f, dr_f, r);
        double z = lagrange4(Z, nr_f, dr_f, r);
        z2r[m]   = 27.2 * 0.529 * z * z; /* Hartree*Bohr: phi(r) r = Z_i(r) Z_j(r) e^2 */
    }
    t->rdr      = 1.0 / t->dr;
    t->rdrho    = 1.0 / t->drho;
    t->nrho_tot = (t->nrho + 1) * 7 + 64;
    t->nr_tot   = (t->nr + 1) * 7 + 64;
    t->nrho_tot -= t->nrho_tot % 64;
    t->nr_tot -= t->nr_tot % 64;
    t->frho_spline = calloc(t->nrho_tot, sizeof(double));
    t->rhor_spline = calloc(t->nr_tot, sizeof(double));
    t->z2r_spline  = calloc(t->nr_tot, sizeof(double));
    spline7(t->nrho, t->drho, frho, t->frho_spline);
    spline7(t->nr, t->dr, rhor, t->rhor_spline);
    spline7(t->nr, t->dr, z2r, t->z2r_spline);
    free(frho);
    free(rhor);
    free(z2r);
    rc = 0;
out:
    free(F);
    free(Z);
    free(R);
    fclose(fp);
    return rc;
}

void mdbEamFree(mdb_eam_tables* t)
{
    free(t->rhor_spline);
    free(t->frho_spline);
    free(t->z2r_spline);
    memset(t, 0, sizeof(*t));
}

/* Synthetic copper-like element (see mdbench_config.h).  taper(r) = x^4/(1+x^4),
   x = (r - rc)/h, vanishing with three derivatives at rc. */
#define SYN_RE 2.556
#define SYN_RC 4.95
#define SYN_H 0.5
#define SYN_Z0 0.1776
#define SYN_BETA 1.5
#define SYN_E0 1.0
#define SYN_N 2000
#define SYN_DR 0.0025
#define SYN_DRHO 0.02

static double syn_taper(double r)
{
    if (r >= SYN_RC) return 0.0;
    double x  = (r - SYN_RC) / SYN_H;
    double x4 = x * x * x * x;
    return x4 / (1.0 + x4);
}

int mdbWriteFuncflSynthetic(const char* path)
{
    FILE* fp = fopen(path, "w");
    if (!fp) return -1;
    fprintf(fp, "synthetic Cu-like funcfl: rho=exp(-3(r/2.556-1))*taper, Z Morse-like, "
                "F=-sqrt(rho)\n");
    fprintf(fp, "%d %.6f %.6f FCC\n", 29, 63.55, 3.615);
    fprintf(fp, "%d %.16e %d %.16e %.16e\n", SYN_N, SYN_DRHO, SYN_N, SYN_DR, SYN_RC);
    for (int which = 0; which < 3; which++) {
        for (int i = 0; i < SYN_N; i++) {
            double v;
            if (which == 0) v = -SYN_E0 * sqrt(i * SYN_DRHO);
            else {
                double r = i * SYN_DR;
                if (which == 1) v = SYN_Z0 * exp(-SYN_BETA * (r - SYN_RE)) * syn_taper(r);
                else v = exp(-3.0 * (r / SYN_RE - 1.0)) * syn_taper(r);
            }
            fprintf(fp, "%.16e%s", v, (i % 5 == 4 || i == SYN_N - 1) ? "\n" : " ");
        }
    }
    fclose(fp);
    return 0;
}
