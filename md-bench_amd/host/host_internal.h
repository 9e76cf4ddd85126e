/* Internal declarations of the host driver (not part of the C ABI). */
#ifndef MDB_HOST_INTERNAL_H
#define MDB_HOST_INTERNAL_H

#include "mdbench_driver.h"
#include "mdgfx950.h"

/* largest atom type a dump file may hold (the device kernels take up to 8 types and reject
   more themselves) */
#define MDB_MAX_TYPES 8

typedef struct {
    double mvv2e, dof_boltz, t_scale, p_scale;
} mdb_thermo;

void mdb_init_parameter(Parameter* p);
int mdb_create_atoms(Parameter* p, Atom* a);
int mdb_load_atoms(Parameter* p, Atom* a, const char* path);
double mdb_wtime(void);
void mdb_setup_thermo(mdb_thermo* th, Parameter* p, int natoms);
void mdb_adjust_thermo(const mdb_thermo* th, const Parameter* p, Atom* a);
void mdb_free_atom_host(Atom* a);

#endif
