/*
 * CPU oracle, EAM tables and force -- TEST INFRASTRUCTURE ONLY (see oracle_md.h).
 * Restates src/common/eam_utils.c (funcfl read, file2array, array2spline/interpolate) and
 * src/verletlist/force_eam.c for one element (all type pairs share one table, as the reference's
 * replication loop at eam_utils.c:240-250 makes them).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_md.h"

#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))

/* grab: eam_utils.c:287-300 -- whitespace-separated values, whole lines at a time. */
static int orc_grab(FILE* fp, int n, double* list)
{
    char line[4096];
    int i = 0;
    while (i < n) {
        if (!fgets(line, sizeof(line), fp)) return -1;
        char* p = strtok(line, " \t\n\r\f");
        while (p) {
            if (i < n) list[i] = atof(p);
            i++;
            p = strtok(NULL, " \t\n\r\f");
        }
    }
    return 0;
}

/* readEamFile: eam_utils.c:42-93 */
int orc_eam_read_funcfl(const char* path, orc_funcfl* file)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char line[4096];
    int tmp;
    if (!fgets(line, sizeof(line), fp)) goto bad;
    if (!fgets(line, sizeof(line), fp)) goto bad;
    if (sscanf(line, "%d %lg", &tmp, &file->mass) != 2) goto bad;
    if (!fgets(line, sizeof(line), fp)) goto bad;
    if (sscanf(line, "%d %lg %d %lg %lg", &file->nrho, &file->drho, &file->nr, &file->dr,
            &file->cut) != 5)
        goto bad;
    file->frho = calloc(file->nrho + 1 + 8, sizeof(double));
    file->rhor = calloc(file->nr + 1 + 8, sizeof(double));
    file->zr   = calloc(file->nr + 1 + 8, sizeof(double));
    if (orc_grab(fp, file->nrho, file->frho) || orc_grab(fp, file->nr, file->zr) ||
        orc_grab(fp, file->nr, file->rhor))
        goto bad;
    for (int i = file->nrho; i > 0; i--)
        file->frho[i] = file->frho[i - 1];
    for (int i = file->nr; i > 0; i--)
        file->rhor[i] = file->rhor[i - 1];
    for (int i = file->nr; i > 0; i--)
        file->zr[i] = file->zr[i - 1];
    fclose(fp);
    return 0;
bad:
    fclose(fp);
    return -2;
}

/* Four-point Lagrange resampling used three times by file2array (eam_utils.c:95-220). */
static double orc_resample(const double* arr, int n, double dfile, double r)
{
    const double sixth = 1.0 / 6.0;
    double p           = r / dfile + 1.0;
    int k              = (int)(p);
    k                  = ORC_MIN(k, n - 2);
    k                  = ORC_MAX(k, 2);
    p -= k;
    p           = ORC_MIN(p, 2.0);
    double cof1 = -sixth * p * (p - 1.0) * (p - 2.0);
    double cof2 = 0.5 * (p * p - 1.0) * (p - 2.0);
    double cof3 = -0.5 * p * (p + 1.0) * (p - 2.0);
    double cof4 = sixth * p * (p * p - 1.0);
    return cof1 * arr[k - 1] + cof2 * arr[k] + cof3 * arr[k + 1] + cof4 * arr[k + 2];
}

/* interpolate: eam_utils.c:253-284 -- 7 coefficients per knot. */
static void orc_interpolate(int n, double delta, const double* f, double* s)
{
    for (int m = 1; m <= n; m++)
        s[m * 7 + 6] = f[m];
    s[1 * 7 + 5]       = s[2 * 7 + 6] - s[1 * 7 + 6];
    s[2 * 7 + 5]       = 0.5 * (s[3 * 7 + 6] - s[1 * 7 + 6]);
    s[(n - 1) * 7 + 5] = 0.5 * (s[n * 7 + 6] - s[(n - 2) * 7 + 6]);
    s[n * 7 + 5]       = s[n * 7 + 6] - s[(n - 1) * 7 + 6];
    for (int m = 3; m <= n - 2; m++)
        s[m * 7 + 5] = ((s[(m - 2) * 7 + 6] - s[(m + 2) * 7 + 6]) +
                           8.0 * (s[(m + 1) * 7 + 6] - s[(m - 1) * 7 + 6])) /
                       12.0;
    for (int m = 1; m <= n - 1; m++) {
        s[m * 7 + 4] = 3.0 * (s[(m + 1) * 7 + 6] - s[m * 7 + 6]) - 2.0 * s[m * 7 + 5] -
                       s[(m + 1) * 7 + 5];
        s[m * 7 + 3] = s[m * 7 + 5] + s[(m + 1) * 7 + 5] - 2.0 * (s[(m + 1) * 7 + 6] - s[m * 7 + 6]);
    }
    s[n * 7 + 4] = 0.0;
    s[n * 7 + 3] = 0.0;
    for (int m = 1; m <= n; m++) {
        s[m * 7 + 2] = s[m * 7 + 5] / delta;
        s[m * 7 + 1] = 2.0 * s[m * 7 + 4] / delta;
        s[m * 7 + 0] = 3.0 * s[m * 7 + 3] / delta;
    }
}

/* file2array (eam_utils.c:95-220) + array2spline (eam_utils.c:222-251), one element. */
int orc_eam_setup(const orc_funcfl* file, orc_eam* eam)
{
    double rmax   = (file->nr - 1) * file->dr;
    double rhomax = (file->nrho - 1) * file->drho;
    eam->dr       = ORC_MAX(0.0, file->dr);
    eam->drho     = ORC_MAX(0.0, file->drho);
    rmax          = ORC_MAX(0.0, rmax);
    rhomax        = ORC_MAX(0.0, rhomax);
    eam->nr       = (int)(rmax / eam->dr + 0.5);
    eam->nrho     = (int)(rhomax / eam->drho + 0.5);

    double* frho = calloc(eam->nrho + 2, sizeof(double));
    double* rhor = calloc(eam->nr + 2, sizeof(double));
    double* z2r  = calloc(eam->nr + 2, sizeof(double));
    for (int m = 1; m <= eam->nrho; m++)
        frho[m] = orc_resample(file->frho, file->nrho, file->drho, (m - 1) * eam->drho);
    for (int m = 1; m <= eam->nr; m++)
        rhor[m] = orc_resample(file->rhor, file->nr, file->dr, (m - 1) * eam->dr);
    for (int m = 1; m <= eam->nr; m++) {
        double r   = (m - 1) * eam->dr;
        double zri = orc_resample(file->zr, file->nr, file->dr, r);
        double zrj = orc_resample(file->zr, file->nr, file->dr, r);
        z2r[m]     = 27.2 * 0.529 * zri * zrj;
    }

    eam->rdr      = 1.0 / eam->dr;
    eam->rdrho    = 1.0 / eam->drho;
    eam->nrho_tot = (eam->nrho + 1) * 7 + 64;
    eam->nr_tot   = (eam->nr + 1) * 7 + 64;
    eam->nrho_tot -= eam->nrho_tot % 64;
    eam->nr_tot -= eam->nr_tot % 64;
    eam->frho_spline = calloc(eam->nrho_tot, sizeof(double));
    eam->rhor_spline = calloc(eam->nr_tot, sizeof(double));
    eam->z2r_spline  = calloc(eam->nr_tot, sizeof(double));
    orc_interpolate(eam->nrho, eam->drho, frho, eam->frho_spline);
    orc_interpolate(eam->nr, eam->dr, rhor, eam->rhor_spline);
    orc_interpolate(eam->nr, eam->dr, z2r, eam->z2r_spline);
    free(frho);
    free(rhor);
    free(z2r);
    return 0;
}

void orc_eam_free(orc_funcfl* file, orc_eam* eam)
{
    if (file) {
        free(file->frho);
        free(file->rhor);
        free(file->zr);
    }
    if (eam) {
        free(eam->frho_spline);
        free(eam->rhor_spline);
        free(eam->z2r_spline);
    }
}

/* computeForceEam: src/verletlist/force_eam.c:19-228 (CSR list, one table).  eatom (may be NULL)
   receives the per-atom energy F(rho_i) + 1/2 sum_j phi(r_ij) (new: the reference computes no
   energy). */
static void force_eam(const orc_eam* eam, int nlocal, int nghost, const double* x,
    const int* border_map, double cutforcesq, const int* numneigh, const int* neighbors,
    double* fp, double* f, double* energy, double* eatom)
{
    const double* rs = eam->rhor_spline;
    const double* fs = eam->frho_spline;
    const double* zs = eam->z2r_spline;
    const double rdr = eam->rdr, rdrho = eam->rdrho;
    const int nr = eam->nr, nrho = eam->nrho;
    double e      = 0.0;
    long long off = 0;
    /* pass 1: density and embedding derivative (:43-113) */
    for (int i = 0; i < nlocal; i++) {
        int nn      = numneigh[i];
        double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
        double rhoi = 0;
        for (int k = 0; k < nn; k++) {
            int j       = neighbors[off + k];
            double delx = xtmp - x[3 * j];
            double dely = ytmp - x[3 * j + 1];
            double delz = ztmp - x[3 * j + 2];
            double rsq  = delx * delx + dely * dely + delz * delz;
            if (rsq < cutforcesq) {
                double p = sqrt(rsq) * rdr + 1.0;
                int m    = (int)(p);
                m        = m < nr - 1 ? m : nr - 1;
                p -= m;
                p = p < 1.0 ? p : 1.0;
                rhoi += ((rs[m * 7 + 3] * p + rs[m * 7 + 4]) * p + rs[m * 7 + 5]) * p + rs[m * 7 + 6];
            }
        }
        double p = 1.0 * rhoi * rdrho + 1.0;
        int m    = (int)(p);
        m        = ORC_MAX(1, ORC_MIN(m, nrho - 1));
        p -= m;
        p     = ORC_MIN(p, 1.0);
        fp[i] = (fs[m * 7 + 0] * p + fs[m * 7 + 1]) * p + fs[m * 7 + 2];
        /* Embedding energy F(rho_i) from the value coefficients (new: the reference computes no
           energy). */
        const double emb = ((fs[m * 7 + 3] * p + fs[m * 7 + 4]) * p + fs[m * 7 + 5]) * p + fs[m * 7 + 6];
        e += emb;
        if (eatom) eatom[i] = emb;
        off += nn;
    }
    /* ghost fp copy (:115-118) */
    for (int g = 0; g < nghost; g++)
        fp[nlocal + g] = fp[border_map[g]];
    /* pass 2: force (:120-224) */
    off = 0;
    for (int i = 0; i < nlocal; i++) {
        int nn      = numneigh[i];
        double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
        double fix = 0, fiy = 0, fiz = 0;
        for (int k = 0; k < nn; k++) {
            int j       = neighbors[off + k];
            double delx = xtmp - x[3 * j];
            double dely = ytmp - x[3 * j + 1];
            double delz = ztmp - x[3 * j + 2];
            double rsq  = delx * delx + dely * dely + delz * delz;
            if (rsq < cutforcesq) {
                double r = sqrt(rsq);
                double p = r * rdr + 1.0;
                int m    = (int)(p);
                m        = m < nr - 1 ? m : nr - 1;
                p -= m;
                p = p < 1.0 ? p : 1.0;
                double rhoip = (rs[m * 7 + 0] * p + rs[m * 7 + 1]) * p + rs[m * 7 + 2];
                double z2p   = (zs[m * 7 + 0] * p + zs[m * 7 + 1]) * p + zs[m * 7 + 2];
                double z2 = ((zs[m * 7 + 3] * p + zs[m * 7 + 4]) * p + zs[m * 7 + 5]) * p + zs[m * 7 + 6];
                double recip = 1.0 / r;
                double phi   = z2 * recip;
                double phip  = z2p * recip - phi * recip;
                double psip  = fp[i] * rhoip + fp[j] * rhoip + phip;
                double fpair = -psip * recip;
                fix += delx * fpair;
                fiy += dely * fpair;
                fiz += delz * fpair;
                e += 0.5 * phi;
                if (eatom) eatom[i] += 0.5 * phi;
            }
        }
        f[3 * i]     = fix;
        f[3 * i + 1] = fiy;
        f[3 * i + 2] = fiz;
        off += nn;
    }
    if (energy) *energy = e;
}

void orc_vl_force_eam(const orc_eam* eam, int nlocal, int nghost, const double* x,
    const int* border_map, double cutforcesq, const int* numneigh, const int* neighbors,
    double* fp, double* f, double* energy)
{
    force_eam(eam, nlocal, nghost, x, border_map, cutforcesq, numneigh, neighbors, fp, f, energy,
        NULL);
}

void orc_vl_force_eam_atom(const orc_eam* eam, int nlocal, int nghost, const double* x,
    const int* border_map, double cutforcesq, const int* numneigh, const int* neighbors,
    double* fp, double* f, double* eatom)
{
    force_eam(eam, nlocal, nghost, x, border_map, cutforcesq, numneigh, neighbors, fp, f, NULL,
        eatom);
}
