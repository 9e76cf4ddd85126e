/*
 * Link-level binding of MD-Bench's own Verlet-list build to libmdgfx950 -- the drop-in shown
 * end to end, with no source file of MD-Bench changed (oracle/Makefile, target dropin; a GPU
 * test runs the result against MD-Bench's own CPU build and its regression values).
 *
 * MD-Bench binds its kernels through function-pointer globals (buildNeighbor, neighbor.c:23-27;
 * updatePbc / updateAtomsPbc, pbc.c:20-26; initialIntegrate / finalIntegrate, integrate.c:13-19;
 * computeForce, set by initForce, force.c:12-41).  A constructor points them at the gfx950 entry
 * points before main() runs.  Four plain calls of main.c need device work as well and are
 * wrapped at link time (GNU ld --wrap):
 *   initDevice    (main.c:98, a no-op in the CPU build, device.c:14): upload the atoms, set up
 *                 the bins and the device ghost map;
 *   setupPbc      (main.c:97 before initDevice: the host version; main.c:127 in reneighbour: the
 *                 device version);
 *   initForce     (main.c:110): MD-Bench's own choice, then the gfx950 force (and the EAM tables
 *                 that initEam, main.c:55, read);
 *   computeThermo (main.c:350, 412, 422): the host thermo of MD-Bench after the positions and
 *                 velocities come back from the device.
 * MDGFX950_STRICT=1 selects the kernels whose arithmetic equals MD-Bench's scalar build.
 * Library failures end the run with the library's message (the library itself never exits).
 */
#include <stdio.h>
#include <stdlib.h>

#include "mdgfx950.h"

/* MD-Bench's kernel-selection globals (force.h:16, neighbor.h:80, pbc.h:15, integrate.h:12) */
typedef double (*ComputeForceFunction)(Parameter*, Atom*, Neighbor*, Stats*);
typedef void (*BuildNeighborFunction)(Atom*, Neighbor*);
typedef void (*UpdatePbcFunction)(Atom*, Parameter*, bool);
typedef void (*IntegrationFunction)(bool, Parameter*, Atom*);
extern ComputeForceFunction computeForce;
extern BuildNeighborFunction buildNeighbor;
extern UpdatePbcFunction updatePbc, updateAtomsPbc;
extern IntegrationFunction initialIntegrate, finalIntegrate;

/* the leading fields of MD-Bench's Eam (eam.h:20-29), the global that initEam fills */
typedef struct {
    MD_FLOAT* fp;
    int nmax;
    int nrho, nr;
    int nrho_tot, nr_tot;
    MD_FLOAT dr, rdr, drho, rdrho;
    MD_FLOAT *frho, *rhor, *z2r;
    MD_FLOAT *rhor_spline, *frho_spline, *z2r_spline;
} EamTables;
extern EamTables eam;

void __real_initDevice(Parameter*, Atom*, Neighbor*);
void __real_setupPbc(Atom*, Parameter*);
void __real_initForce(Parameter*);
void __real_computeThermo(int, Parameter*, Atom*);

static int device_ready = 0;

static void check(const char* where)
{
    if (mdgfx950Status()) {
        fprintf(stderr, "libmdgfx950 (%s): %s\n", where, mdgfx950StatusMessage());
        exit(EXIT_FAILURE);
    }
}

__attribute__((constructor)) static void bindGfx950(void)
{
    buildNeighbor    = buildNeighborGFX950;
    updatePbc        = updatePbcGFX950;
    updateAtomsPbc   = updateAtomsPbcGFX950;
    initialIntegrate = initialIntegrateGFX950;
    finalIntegrate   = finalIntegrateGFX950;
    const char* strict = getenv("MDGFX950_STRICT");
    mdgfx950SetStrict(strict != NULL && atoi(strict) != 0);
}

void __wrap_initDevice(Parameter* param, Atom* atom, Neighbor* neighbor)
{
    initDeviceGFX950(param, atom, neighbor);
    setupNeighborGFX950(param);
    setupPbcGFX950(atom, param);
    check("initDevice");
    device_ready = 1;
}

void __wrap_setupPbc(Atom* atom, Parameter* param)
{
    if (!device_ready) {
        __real_setupPbc(atom, param);
        return;
    }
    setupPbcGFX950(atom, param);
    check("setupPbc");
}

void __wrap_initForce(Parameter* param)
{
    __real_initForce(param);
    if (param->force_field == FF_EAM) {
        setEamTablesGFX950(eam.nr, eam.nrho, eam.rdr, eam.rdrho, eam.rhor_spline,
            eam.frho_spline, eam.z2r_spline, eam.nr_tot, eam.nrho_tot);
        computeForce = computeForceEamGFX950;
    } else {
        computeForce = computeForceLJGFX950;
    }
    check("initForce");
}

void __wrap_computeThermo(int iflag, Parameter* param, Atom* atom)
{
    if (device_ready) {
        copyDataFromGFX950Device(param, atom);
        check("computeThermo");
    }
    __real_computeThermo(iflag, param, atom);
}
