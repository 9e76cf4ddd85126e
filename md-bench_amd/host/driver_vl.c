/*
 * Verlet-list driver: MD-Bench's setup() and timestep loop (src/verletlist/main.c:46-113,
 * 371-419) with every per-step operation on the device.  The host touches the GPU only for the
 * thermo reductions it is asked for and for the reneighbour status word.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "host_internal.h"

struct mdb_vl {
    Parameter param;
    Atom atom;
    Neighbor neighbor;
    Stats stats;
    mdb_thermo thermo;
    mdb_eam_tables eam;
    double* x0;
    double* v0;
    int* t0;
    int device_ready;
    int sort_on; /* SORT_ATOMS: re-sort by bin at setup and every param.resort_every steps */
    /* learned EAM surrogate (force_field FF_EAM only): judged by plain EAM every sur_every steps */
    int sur_on, sur_every, sur_fallback, sur_judges;
    double sur_tol, sur_last, sur_max;
    long long steps_done, sur_fallback_step;
};

Parameter* mdbVLParam(mdb_vl* s) { return &s->param; }
Atom* mdbVLAtom(mdb_vl* s) { return &s->atom; }
Neighbor* mdbVLNeighbor(mdb_vl* s) { return &s->neighbor; }
Stats* mdbVLStats(mdb_vl* s) { return &s->stats; }

void mdbThermo(const mdb_vl* s, double ksum, double* temp, double* press)
{
    double t = ksum * s->thermo.t_scale;
    if (temp) *temp = t;
    if (press) *press = (t * s->thermo.dof_boltz) * s->thermo.p_scale;
}

void mdb_free_atom_host(Atom* a)
{
    free(a->x);
    free(a->vx);
    free(a->fx);
    free(a->type);
    free(a->sqrt_epsilon);
    free(a->sigma3);
    free(a->epsilon);
    free(a->sigma6);
    free(a->cutforcesq);
    free(a->cutneighsq);
    free(a->sqrt_epsilon_per_type);
    free(a->sigma3_per_type);
    free(a->border_map);
}

static double compute_force(mdb_vl* s)
{
    if (s->param.force_field == FF_EAM) {
        if (s->sur_on && !s->sur_fallback)
            return computeForceSurrogateGFX950(&s->param, &s->atom, &s->neighbor, &s->stats);
        return computeForceEamGFX950(&s->param, &s->atom, &s->neighbor, &s->stats);
    }
    return computeForceLJGFX950(&s->param, &s->atom, &s->neighbor, &s->stats);
}

int mdbVLCreate(const mdb_config* cfg, mdb_vl** out)
{
    *out     = NULL;
    mdb_vl* s = calloc(1, sizeof(mdb_vl));
    if (!s) return MDG_ERR_CAPACITY;
    Parameter* p = &s->param;
    mdb_init_parameter(p);
    p->nx            = cfg->nx;
    p->ny            = cfg->ny;
    p->nz            = cfg->nz;
    p->force_field   = cfg->force_field;
    p->half_neigh    = cfg->half_neigh;
    p->ntimes        = cfg->ntimes;
    p->nstat         = cfg->nstat > 0 ? cfg->nstat : p->nstat;
    p->reneigh_every = cfg->reneigh_every > 0 ? cfg->reneigh_every : p->reneigh_every;
    if (cfg->resort_every > 0) {
        p->resort_every = cfg->resort_every;
        s->sort_on      = 1;
    }
    if (cfg->cutforce > 0) p->cutforce = cfg->cutforce;
    if (cfg->skin >= 0) p->skin = cfg->skin;
    if (cfg->ntypes > 0) p->ntypes = cfg->ntypes;
    if (cfg->epsilon_per_type && cfg->sigma_per_type) {
        p->epsilon_per_type = malloc(sizeof(double) * p->ntypes);
        p->sigma_per_type   = malloc(sizeof(double) * p->ntypes);
        memcpy(p->epsilon_per_type, cfg->epsilon_per_type, sizeof(double) * p->ntypes);
        memcpy(p->sigma_per_type, cfg->sigma_per_type, sizeof(double) * p->ntypes);
    }
    p->cutneigh = p->cutforce + p->skin; /* main.c:330 */
    mdgfx950SetStrict(cfg->strict);

    if (p->force_field == FF_EAM) {
        /* initEam (eam_utils.c:22-40) */
        char tmp[256];
        const char* path = cfg->eam_file;
        if (!path) {
            snprintf(tmp, sizeof(tmp), "/tmp/mdb_synth_%d.eam", (int)getpid());
            if (mdbWriteFuncflSynthetic(tmp)) {
                free(s);
                return MDG_ERR_ARG;
            }
            path = tmp;
        }
        int rc = mdbEamLoad(path, &s->eam);
        if (!cfg->eam_file) unlink(tmp);
        if (rc) {
            fprintf(stderr, "mdbVLCreate: cannot read EAM file %s\n", path);
            free(s);
            return MDG_ERR_ARG;
        }
        p->mass     = s->eam.mass;
        p->cutforce = s->eam.cut;
        p->cutneigh = p->cutforce + 1.0;
        p->temp     = 600.0;
        p->dt       = 0.001;
        p->rho      = 0.07041125;
        p->dtforce  = 0.5 * p->dt / p->mass;
    }
    if (cfg->rho > 0) p->rho = cfg->rho;
    if (cfg->lattice_a > 0) p->rho = 4.0 / (cfg->lattice_a * cfg->lattice_a * cfg->lattice_a);
    if (cfg->temp > 0) p->temp = cfg->temp;
    if (cfg->dt > 0) {
        p->dt      = cfg->dt;
        p->dtforce = p->force_field == FF_EAM ? 0.5 * p->dt / p->mass : 0.5 * p->dt;
    }

    /* setup() */
    p->lattice = pow((4.0 / p->rho), (1.0 / 3.0));
    p->xprd    = p->nx * p->lattice;
    p->yprd    = p->ny * p->lattice;
    p->zprd    = p->nz * p->lattice;
    Neighbor* nb   = &s->neighbor;
    nb->maxneighs  = 100; /* initNeighbor, neighbor.c:75 */
    nb->half_neigh = p->half_neigh;
    if (cfg->input_file) {
        /* readAtom (main.c:69-73): the file's atoms and box, no thermo adjustment; setupNeighbor
           then sizes the bins from the box (neighbor.c:107-140) */
        p->input_file = strdup(cfg->input_file);
        if (!p->input_file || mdb_load_atoms(p, &s->atom, cfg->input_file)) {
            fprintf(stderr, "mdbVLCreate: cannot read atoms from %s\n", cfg->input_file);
            mdbVLDestroy(s);
            return MDG_ERR_ARG;
        }
        mdb_setup_thermo(&s->thermo, p, s->atom.Natoms);
    } else {
        if (mdb_create_atoms(p, &s->atom)) {
            mdbVLDestroy(s);
            return MDG_ERR_CAPACITY;
        }
        mdb_setup_thermo(&s->thermo, p, s->atom.Natoms);
        if (cfg->maxwell_seed > 0)
            mdbMaxwellNormals(s->atom.Nlocal, (unsigned long long)cfg->maxwell_seed,
                s->atom.vx);
        mdb_adjust_thermo(&s->thermo, p, &s->atom);
    }
    int n = s->atom.Nlocal;
    s->x0 = malloc(sizeof(double) * 3 * n);
    s->v0 = malloc(sizeof(double) * 3 * n);
    s->t0 = malloc(sizeof(int) * n);
    memcpy(s->x0, s->atom.x, sizeof(double) * 3 * n);
    memcpy(s->v0, s->atom.vx, sizeof(double) * 3 * n);
    memcpy(s->t0, s->atom.type, sizeof(int) * n);

    initDeviceGFX950(p, &s->atom, nb);
    s->device_ready = 1;
    setupNeighborGFX950(p);
    if (s->sort_on) sortAtomGFX950(&s->atom); /* main.c:92-95, before setupPbc */
    if (p->force_field == FF_EAM)
        setEamTablesGFX950(s->eam.nr, s->eam.nrho, s->eam.rdr, s->eam.rdrho, s->eam.rhor_spline,
            s->eam.frho_spline, s->eam.z2r_spline, s->eam.nr_tot, s->eam.nrho_tot);
    setupPbcGFX950(&s->atom, p);
    updatePbcGFX950(&s->atom, p, true);
    buildNeighborGFX950(&s->atom, nb);
    compute_force(s);
    int st = mdgfx950Status();
    if (st) {
        mdbVLDestroy(s);
        return st;
    }
    *out = s;
    return 0;
}

int mdbCreateLattice(const mdb_config* cfg, double* x, double* v, int* type, int cap)
{
    if (cfg->force_field != FF_LJ) return -1;
    Parameter p;
    Atom a;
    mdb_thermo th;
    memset(&a, 0, sizeof(a));
    mdb_init_parameter(&p);
    p.nx = cfg->nx;
    p.ny = cfg->ny;
    p.nz = cfg->nz;
    if (cfg->ntypes > 0) p.ntypes = cfg->ntypes;
    if (cfg->rho > 0) p.rho = cfg->rho;
    if (cfg->lattice_a > 0) p.rho = 4.0 / (cfg->lattice_a * cfg->lattice_a * cfg->lattice_a);
    if (cfg->temp > 0) p.temp = cfg->temp;
    p.lattice = pow((4.0 / p.rho), (1.0 / 3.0));
    p.xprd    = p.nx * p.lattice;
    p.yprd    = p.ny * p.lattice;
    p.zprd    = p.nz * p.lattice;
    int rc    = mdb_create_atoms(&p, &a);
    if (rc == 0) {
        mdb_setup_thermo(&th, &p, a.Natoms);
        if (cfg->maxwell_seed > 0)
            mdbMaxwellNormals(a.Nlocal, (unsigned long long)cfg->maxwell_seed, a.vx);
        mdb_adjust_thermo(&th, &p, &a);
        rc = a.Nlocal;
        if (rc > cap) rc = -3;
        else {
            memcpy(x, a.x, sizeof(double) * 3 * a.Nlocal);
            memcpy(v, a.vx, sizeof(double) * 3 * a.Nlocal);
            memcpy(type, a.type, sizeof(int) * a.Nlocal);
        }
    }
    mdb_free_atom_host(&a);
    return rc;
}

int mdbVLInitialState(const mdb_vl* s, double* x, double* v, int* type)
{
    int n = s->atom.Nlocal;
    if (x) memcpy(x, s->x0, sizeof(double) * 3 * n);
    if (v) memcpy(v, s->v0, sizeof(double) * 3 * n);
    if (type) memcpy(type, s->t0, sizeof(int) * n);
    return n;
}

int mdbVLRun(mdb_vl* s, int nsteps, double* ksum)
{
    Parameter* p = &s->param;
    Atom* a      = &s->atom;
    Neighbor* nb = &s->neighbor;
    int integrated = 0; /* this step's initialIntegrate ran inside the last force call */
    for (int n = 0; n < nsteps; n++) {
        /* SORT_ATOMS (main.c:372-381): the sort comes before the step's initialIntegrate (so the
           step before it does not fuse that integration into its force call) and the step then
           rebuilds ghosts and lists, which the reference's default periods (400 and 20) give */
        const int resort = s->sort_on && (n + 1) % p->resort_every == 0;
        int reneigh      = resort || (n + 1) % p->reneigh_every == 0;
        if (resort) sortAtomGFX950(a);
        if (!integrated) initialIntegrateGFX950(reneigh, p, a);
        if (reneigh) {
            updateAtomsPbcGFX950(a, p, true);
            setupPbcGFX950(a, p);
            updatePbcGFX950(a, p, true);
            buildNeighborGFX950(a, nb);
        } else {
            updatePbcGFX950(a, p, false);
        }
        /* plain LJ/EAM steps carry the next step's initialIntegrate in the force call, unless
           this step's full-step velocities are read (after the last step, or every step when
           traced) or the surrogate runs (its judge reads this step's positions) */
        integrated = n + 1 < nsteps && !ksum && !s->sur_on &&
                     !(s->sort_on && (n + 2) % p->resort_every == 0);
        if (integrated) {
            computeForceFinalInitialGFX950(p, a, nb, &s->stats);
            s->steps_done++;
            if (mdgfx950Status()) return mdgfx950Status();
            continue;
        }
        compute_force(s);
        s->steps_done++;
        if (s->sur_on && !s->sur_fallback && s->sur_every > 0 &&
            s->steps_done % s->sur_every == 0) {
            int fb   = 0;
            double r = surrogateJudgeGFX950(p, a, nb, &s->stats, s->sur_tol, &fb);
            s->sur_judges++;
            s->sur_last = r;
            if (r > s->sur_max) s->sur_max = r;
            if (fb) {
                s->sur_fallback      = 1;
                s->sur_fallback_step = s->steps_done;
            }
        }
        finalIntegrateGFX950(reneigh, p, a);
        if (ksum) ksum[n] = kineticSumGFX950(p, a);
        if (mdgfx950Status()) return mdgfx950Status();
    }
    return mdgfx950Status();
}

int mdbVLUseSurrogate(mdb_vl* s, int on, int judge_every, double tol)
{
    if (s->param.force_field != FF_EAM) return MDG_ERR_ARG;
    s->sur_on            = on != 0;
    s->sur_every         = judge_every;
    s->sur_tol           = tol;
    s->sur_fallback      = 0;
    s->sur_judges        = 0;
    s->sur_last          = 0.0;
    s->sur_max           = 0.0;
    s->sur_fallback_step = -1;
    if (s->sur_on) compute_force(s); /* forces of the current positions from the surrogate */
    return mdgfx950Status();
}

void mdbVLSurrogateStats(const mdb_vl* s, double* out)
{
    out[0] = s->sur_judges;
    out[1] = s->sur_last;
    out[2] = s->sur_max;
    out[3] = (double)s->sur_fallback_step;
    out[4] = s->sur_fallback;
}

void mdbVLDestroy(mdb_vl* s)
{
    if (!s) return;
    if (s->device_ready) freeDeviceGFX950(&s->atom, &s->neighbor);
    mdb_free_atom_host(&s->atom);
    free(s->param.epsilon_per_type);
    free(s->param.sigma_per_type);
    free(s->param.input_file);
    mdbEamFree(&s->eam);
    free(s->x0);
    free(s->v0);
    free(s->t0);
    free(s);
}
