/*
 * Atom files in MD-Bench's dump format (src/verletlist/atom.c:529-656, readAtom_dmp):
 *
 *   ITEM: TIMESTEP / <int>
 *   ITEM: NUMBER OF ATOMS / <int>
 *   ITEM: BOX BOUNDS pp pp pp / xlo xhi / ylo yhi / zlo zhi
 *   ITEM: ATOMS id type x y z vx vy vz / one line per atom, id from 1
 *
 * The writer prints every double with 17 significant digits, so MD-Bench's atof() (and the reader
 * here) recovers the identical values: a state written here starts MD-Bench and this library
 * from the same bits.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

int mdbWriteDmp(const char* path, int natoms, const double* x, const double* v, const int* type,
    const double* box, int timestep)
{
    FILE* fp = fopen(path, "w");
    if (!fp) return -1;
    fprintf(fp, "ITEM: TIMESTEP\n%d\n", timestep);
    fprintf(fp, "ITEM: NUMBER OF ATOMS\n%d\n", natoms);
    fprintf(fp, "ITEM: BOX BOUNDS pp pp pp\n");
    for (int d = 0; d < 3; d++)
        fprintf(fp, "%.17g %.17g\n", box[2 * d], box[2 * d + 1]);
    fprintf(fp, "ITEM: ATOMS id type x y z vx vy vz\n");
    for (int i = 0; i < natoms; i++)
        fprintf(fp, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g\n", i + 1, type[i], x[3 * i],
            x[3 * i + 1], x[3 * i + 2], v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    int rc = ferror(fp) ? -2 : 0;
    if (fclose(fp)) rc = -2;
    return rc;
}

static int next_line(FILE* fp, char* line, int n)
{
    return fgets(line, n, fp) != NULL;
}

int mdbReadDmp(const char* path, int cap, double* x, double* v, int* type, double* box,
    int* ntypes)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char line[4096];
    int natoms = 0, read = 0, ts = -1, nt = 0, rc = 0;
    unsigned char* seen = NULL; /* one flag per id: a repeated id is an error */
    while (!read && next_line(fp, line, sizeof(line))) {
        if (strncmp(line, "ITEM: ", 6) != 0) {
            rc = -3; /* the reference rejects anything but an item line here */
            break;
        }
        const char* item = line + 6;
        if (strncmp(item, "TIMESTEP", 8) == 0) {
            if (!next_line(fp, line, sizeof(line))) break;
            ts = atoi(line);
        } else if (strncmp(item, "NUMBER OF ATOMS", 15) == 0) {
            if (!next_line(fp, line, sizeof(line))) break;
            natoms = atoi(line);
            if (natoms > cap) {
                rc = -4;
                break;
            }
        } else if (strncmp(item, "BOX BOUNDS pp pp pp", 19) == 0) {
            for (int d = 0; d < 3 && rc == 0; d++) {
                if (!next_line(fp, line, sizeof(line))) {
                    rc = -3;
                    break;
                }
                char* tok    = strtok(line, " ");
                box[2 * d]   = tok ? atof(tok) : 0.0;
                tok          = strtok(NULL, " ");
                box[2 * d + 1] = tok ? atof(tok) : 0.0;
            }
        } else if (strncmp(item, "ATOMS id type x y z vx vy vz", 28) == 0) {
            free(seen);
            seen = (unsigned char*)calloc(natoms > 0 ? natoms : 1, 1);
            if (!seen) {
                rc = -5;
                break;
            }
            for (int i = 0; i < natoms; i++) {
                if (!next_line(fp, line, sizeof(line))) {
                    rc = -3;
                    break;
                }
                double f[6];
                char* tok = strtok(line, " ");
                int id    = tok ? atoi(tok) - 1 : -1;
                tok       = strtok(NULL, " ");
                int t     = tok ? atoi(tok) : 0;
                for (int k = 0; k < 6; k++) {
                    tok  = strtok(NULL, " ");
                    f[k] = tok ? atof(tok) : 0.0;
                }
                /* ids from 1 to natoms, each once; types from 0 (they index the per-type and
                   type-pair parameter tables) */
                if (id < 0 || id >= natoms || seen[id] || t < 0 || t >= MDB_MAX_TYPES) {
                    rc = -3;
                    break;
                }
                seen[id] = 1;
                type[id] = t;
                for (int d = 0; d < 3; d++) {
                    x[3 * id + d] = f[d];
                    v[3 * id + d] = f[3 + d];
                }
                if (t + 1 > nt) nt = t + 1;
                read++;
            }
        } else {
            rc = -3;
            break;
        }
    }
    fclose(fp);
    free(seen);
    if (rc) return rc;
    if (ts < 0 || !natoms || read != natoms) return -3;
    if (ntypes) *ntypes = nt;
    return natoms;
}
