/* CPU restatement of Mjolnir's force-and-step path (TEST INFRASTRUCTURE ONLY).
 *
 * See mjo_oracle.h.  Citations are to the Mjolnir source tree.  The code is restated, not
 * copied: data structures are plain C arrays, but the arithmetic keeps the reference's
 * operation order so results agree bit for bit with the reference built with the same flags.
 */
#include "mjo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------- */
/* units: mjolnir/core/Unit.hpp:23-44,150-159 and mjolnir/input/read_units.hpp:32-92      */
/* ------------------------------------------------------------------------------------- */
static const double U_KB = 1.38064852e-23, U_NA = 6.022140857e23;
static const double U_E = 1.6021766208e-19, U_EPS0 = 8.854187817e-12;

double orc_kB(void)
{
    const double J_to_cal = 1.0 / 4.184;
    return U_KB * (J_to_cal / 1000.0) * U_NA;
}
double orc_eps0(void)
{
    const double J_to_cal = 1.0 / 4.184;
    double eps0 = (U_EPS0 / U_E) / U_E;
    eps0 = eps0 * (1000.0 / J_to_cal) / U_NA;
    return eps0 / 1e+10;
}
double orc_L_to_volume(void) { return 1e-3 * pow(1e+10, 3); }

/* ------------------------------------------------------------------------------------- */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,  */
/* 3", SC'11.  Round: (hi0,lo0)=M0*x0, (hi1,lo1)=M1*x2, x' = {hi1^x1^k0, lo1, hi0^x3^k1,   */
/* lo0}; key bumped by the Weyl constants between rounds.                                 */
/* ------------------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for(int r = 0; r < 10; ++r)
    {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        x0 = hi1 ^ x1 ^ k0;
        x1 = lo1;
        x2 = hi0 ^ x3 ^ k1;
        x3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* The engine's Langevin noise: three N(0,1) draws per (seed, step, particle), from one
 * Philox block through two Box-Muller transforms.  Uniforms are (r + 0.5) 2^-32, in (0,1). */
void orc_noise3(uint64_t seed, uint64_t step, uint32_t particle, double out[3])
{
    const uint32_t ctr[4] = {particle, (uint32_t)step, (uint32_t)(step >> 32), 0x4D4A4721u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    const double s = 2.3283064365386963e-10; /* 2^-32 */
    const double two_pi = 6.283185307179586;
    const double u0 = ((double)r[0] + 0.5) * s, u1 = ((double)r[1] + 0.5) * s;
    const double u2 = ((double)r[2] + 0.5) * s, u3 = ((double)r[3] + 0.5) * s;
    const double a = sqrt(-2.0 * log(u0)), b = sqrt(-2.0 * log(u2));
    out[0] = a * cos(two_pi * u1);
    out[1] = a * sin(two_pi * u1);
    out[2] = b * cos(two_pi * u3);
}

/* ------------------------------------------------------------------------------------- */
/* small containers                                                                       */
/* ------------------------------------------------------------------------------------- */
typedef struct { size_t* a; size_t n, cap; } zvec;
static void zpush(zvec* v, size_t x)
{
    if(v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 16; v->a = (size_t*)realloc(v->a, v->cap * sizeof(size_t)); }
    v->a[v->n++] = x;
}
static int cmp_size(const void* a, const void* b)
{
    const size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return (x < y) ? -1 : (x > y);
}
static size_t sort_unique(size_t* a, size_t n)
{
    if(n == 0) return 0;
    qsort(a, n, sizeof(size_t), cmp_size);
    size_t m = 1;
    for(size_t i = 1; i < n; ++i) if(a[i] != a[m - 1]) a[m++] = a[i];
    return m;
}

/* ------------------------------------------------------------------------------------- */
/* engine state                                                                            */
/* ------------------------------------------------------------------------------------- */
#define MAX_KINDS 16
#define MAX_WITHIN 8

typedef struct { size_t node; int kind; } edge;
typedef struct { edge* a; size_t n, cap; } edgevec;

typedef struct {
    int kind, topo_kind, arity, npar;
    size_t n;
    uint32_t* idx;
    double* par;
    /* Contact: active list (ContactInteraction.hpp:159-183) */
    double margin, cutoff, current_margin;
    zvec active;
} local_t;

typedef struct {
    int pot, partition, ignore_mol;
    double cutoff_ratio, eps, margin;
    size_t n_within;
    int within_kind[MAX_WITHIN];
    size_t within_dist[MAX_WITHIN];
    size_t np;
    size_t* part;          /* participants, input order */
    int npar;              /* per-particle parameter columns */
    double* ppar;          /* per particle index (size N*npar), zero for non-participants */
    /* potential-derived */
    double coef_at_cutoff, max_cutoff;
    double debye, inv_debye, abs_cutoff, inv4pieps; /* DH */
    /* exclusion list (ExclusionList.hpp:86-234) */
    size_t* ex_start;      /* N+1 */
    size_t* ex;            /* sorted ignored particle indices per i */
    size_t* mol;           /* molecule id per particle */
    /* neighbour list: CSR over leading participants (NeighborList.hpp:151-266) */
    size_t* rng_start;     /* np (ranges per leading participant index k, k+1) */
    zvec nbr;              /* partner index */
    double* npp; size_t npp_cap; /* pair parameters, npp_w per entry */
    int npp_w;
    /* spatial partition */
    double cutoff, current_margin;
    /* periodic grid (PeriodicGridCellList.hpp) */
    double r_x, r_y, r_z, sys_lo[3], sys_size[3];
    size_t dim_x, dim_y, dim_z;
    size_t* adj;           /* 27 per cell */
    double r_cell;         /* unlimited grid */
} global_t;

struct orc_engine {
    size_t n;
    int periodic;
    double lower[3], upper[3], width[3], halfw[3];
    double *mass, *rmass, *pos, *vel, *force;
    double temperature, ionic_strength;
    int has_T, has_I;
    local_t* loc; size_t n_loc;
    global_t* glo; size_t n_glo;
    char* kinds[MAX_KINDS]; int n_kinds;
    edgevec* adjs;
    size_t* mol; size_t n_mol;
    /* integrator */
    int integ; double dt, halfdt; uint64_t seed; uint64_t step;
    int rm_translation, rm_rotation, rm_rescale;   /* SystemMotionRemover flags */
    double *coef, *c1, *c2, *c3; /* BAOAB: exp(-g dt), noise; GJF: b, beta, a;
                                    UnderdampedLangevin: -, noise_coef sqrt(gamma / m) */
    double* accel;                 /* UnderdampedLangevin: f/m + noise of the last force call */
    size_t rebuilds;
    /* constraints (g-BAOAB) */
    size_t n_con; uint32_t* con_idx; double* con_v0; int con_kind; size_t con_maxit; double con_tol;
    double *con_sq_v0, *con_r_sq_v0, *con_redm;   /* v0^2, 1 / v0^2, 1 / (1/m1 + 1/m2) */
    double *old_pos, *old_rattle;                  /* position at the last step end / before an A step */
    double dt_corr, r_dt_corr;
};

static int kind_id(orc_engine* e, const char* name)
{
    for(int i = 0; i < e->n_kinds; ++i) if(strcmp(e->kinds[i], name) == 0) return i;
    if(e->n_kinds == MAX_KINDS) return -1;
    e->kinds[e->n_kinds] = (char*)malloc(strlen(name) + 1);
    strcpy(e->kinds[e->n_kinds], name);
    return e->n_kinds++;
}

/* BoundaryCondition.hpp:81-91 */
static void adjust_direction(const orc_engine* e, const double* from, const double* to, double* dr)
{
    dr[0] = to[0] - from[0]; dr[1] = to[1] - from[1]; dr[2] = to[2] - from[2];
    if(!e->periodic) return;
    for(int d = 0; d < 3; ++d)
    {
        if(dr[d] < -e->halfw[d]) dr[d] += e->width[d];
        else if(dr[d] >= e->halfw[d]) dr[d] -= e->width[d];
    }
}
/* BoundaryCondition.hpp:93-102 */
static void adjust_position(const orc_engine* e, double* p)
{
    if(!e->periodic) return;
    for(int d = 0; d < 3; ++d)
    {
        if(p[d] < e->lower[d]) p[d] += e->width[d];
        else if(p[d] >= e->upper[d]) p[d] -= e->width[d];
    }
}
/* math/Vector.hpp:85-88 */
static double length_sq(const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
static double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double rsqrt(double x) { return 1.0 / sqrt(x); }        /* math/functions.hpp:29-33 */
static double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

orc_engine* orc_create(int periodic, const double* lower, const double* upper, size_t n)
{
    orc_engine* e = (orc_engine*)calloc(1, sizeof(orc_engine));
    e->n = n;
    e->periodic = periodic;
    if(periodic)
    {
        /* CuboidalPeriodicBoundary::set_boundary (BoundaryCondition.hpp:115-123) */
        for(int d = 0; d < 3; ++d)
        {
            e->lower[d] = lower[d]; e->upper[d] = upper[d];
            e->width[d] = upper[d] - lower[d];
            e->halfw[d] = e->width[d] * 0.5;
        }
    }
    e->mass = (double*)calloc(n, sizeof(double));
    e->rmass = (double*)calloc(n, sizeof(double));
    e->pos = (double*)calloc(3 * n, sizeof(double));
    e->vel = (double*)calloc(3 * n, sizeof(double));
    e->force = (double*)calloc(3 * n, sizeof(double));
    e->adjs = (edgevec*)calloc(n, sizeof(edgevec));
    return e;
}

void orc_destroy(orc_engine* e)
{
    if(!e) return;
    for(size_t i = 0; i < e->n_loc; ++i) { free(e->loc[i].idx); free(e->loc[i].par); free(e->loc[i].active.a); }
    for(size_t g = 0; g < e->n_glo; ++g)
    {
        global_t* G = &e->glo[g];
        free(G->part); free(G->ppar); free(G->ex_start); free(G->ex); free(G->mol);
        free(G->rng_start); free(G->nbr.a); free(G->npp); free(G->adj);
    }
    free(e->loc); free(e->glo);
    for(int i = 0; i < e->n_kinds; ++i) free(e->kinds[i]);
    for(size_t i = 0; i < e->n; ++i) free(e->adjs[i].a);
    free(e->adjs); free(e->mol);
    free(e->mass); free(e->rmass); free(e->pos); free(e->vel); free(e->force);
    free(e->coef); free(e->c1); free(e->c2); free(e->c3); free(e->accel);
    free(e->con_idx); free(e->con_v0); free(e->con_sq_v0); free(e->con_r_sq_v0); free(e->con_redm);
    free(e->old_pos); free(e->old_rattle);
    free(e);
}

int orc_set_state(orc_engine* e, const double* mass, const double* pos, const double* vel)
{
    for(size_t i = 0; i < e->n; ++i)
    {
        e->mass[i] = mass[i];
        e->rmass[i] = 1.0 / mass[i];
    }
    memcpy(e->pos, pos, 3 * e->n * sizeof(double));
    memcpy(e->vel, vel, 3 * e->n * sizeof(double));
    /* System::initialize wraps positions into the boundary (System.hpp:63-67) */
    for(size_t i = 0; i < e->n; ++i) adjust_position(e, e->pos + 3 * i);
    memset(e->force, 0, 3 * e->n * sizeof(double));
    return 0;
}

int orc_get_state(const orc_engine* e, double* pos, double* vel, double* force)
{
    if(pos) memcpy(pos, e->pos, 3 * e->n * sizeof(double));
    if(vel) memcpy(vel, e->vel, 3 * e->n * sizeof(double));
    if(force) memcpy(force, e->force, 3 * e->n * sizeof(double));
    return 0;
}

int orc_set_attribute(orc_engine* e, const char* key, double value)
{
    if(strcmp(key, "temperature") == 0) { e->temperature = value; e->has_T = 1; return 0; }
    if(strcmp(key, "ionic_strength") == 0) { e->ionic_strength = value; e->has_I = 1; return 0; }
    return -1;
}

int orc_add_local(orc_engine* e, int kind, const char* topology, size_t n_terms,
                  const uint32_t* indices, const double* params, double margin)
{
    int arity, npar;
    switch(kind)
    {
        case ORC_BOND_HARMONIC: case ORC_BOND_GOCONTACT: case ORC_CONTACT_GOCONTACT: arity = 2; npar = 2; break;
        case ORC_ANGLE_HARMONIC: arity = 3; npar = 2; break;
        case ORC_ANGLE_GAUSSIAN: arity = 3; npar = 3; break;
        case ORC_DIHEDRAL_CLEMENTI: arity = 4; npar = 3; break;
        case ORC_BOND_3SPN2: arity = 2; npar = 2; break;
        case ORC_DIHEDRAL_PGAUSSIAN: arity = 4; npar = 3; break;
        case ORC_3SPN2_STACKING: arity = 6; npar = 1; break;
        default: return -1;
    }
    for(size_t t = 0; t < n_terms * (size_t)arity; ++t)
    {
        /* the phosphates of a stacking term may be absent */
        const int optional = (kind == ORC_3SPN2_STACKING) && (t % 6 == 3 || t % 6 == 5);
        if(indices[t] >= e->n && !(optional && indices[t] == ORC_NIL)) return -1;
    }
    e->loc = (local_t*)realloc(e->loc, (e->n_loc + 1) * sizeof(local_t));
    local_t* L = &e->loc[e->n_loc];
    memset(L, 0, sizeof(*L));
    L->kind = kind; L->arity = arity; L->npar = npar; L->n = n_terms; L->margin = margin;
    L->topo_kind = (topology == NULL || topology[0] == '\0' || strcmp(topology, "none") == 0) ? -1 : kind_id(e, topology);
    L->idx = (uint32_t*)malloc(n_terms * arity * sizeof(uint32_t) + 1);
    L->par = (double*)malloc(n_terms * npar * sizeof(double) + 1);
    memcpy(L->idx, indices, n_terms * arity * sizeof(uint32_t));
    memcpy(L->par, params, n_terms * npar * sizeof(double));
    return (int)(e->n_loc++);
}

int orc_add_global(orc_engine* e, int potential, double cutoff_ratio, double epsilon,
                   int partition, double margin, int ignore_molecule,
                   size_t n_within, const char* const* within_kinds, const uint32_t* within_dist,
                   size_t n_participants, const uint32_t* participants, const double* params)
{
    int npar;
    switch(potential)
    {
        case ORC_PAIR_LJ: npar = 2; break;
        case ORC_PAIR_EV: case ORC_PAIR_DH: case ORC_PAIR_3SPN2_EV: npar = 1; break;
        case ORC_3SPN2_BASEPAIR: npar = 2; break;
        case ORC_3SPN2_CROSSSTACK: npar = 4; break;
        default: return -1;
    }
    if(n_within > MAX_WITHIN || n_participants < 1) return -1;
    for(size_t t = 0; t < n_participants; ++t) if(participants[t] >= e->n) return -1;
    e->glo = (global_t*)realloc(e->glo, (e->n_glo + 1) * sizeof(global_t));
    global_t* G = &e->glo[e->n_glo];
    memset(G, 0, sizeof(*G));
    G->pot = potential; G->partition = partition; G->ignore_mol = ignore_molecule;
    G->cutoff_ratio = cutoff_ratio; G->eps = epsilon; G->margin = margin;
    /* the reference keeps ignore.particles_within in a std::map, i.e. sorted by kind name
       (ExclusionList.hpp:39-45); the union of the ignored sets does not depend on it */
    G->n_within = n_within;
    for(size_t k = 0; k < n_within; ++k) { G->within_kind[k] = kind_id(e, within_kinds[k]); G->within_dist[k] = within_dist[k]; }
    G->np = n_participants;
    G->part = (size_t*)malloc(n_participants * sizeof(size_t));
    G->npar = npar;
    G->ppar = (double*)calloc(e->n * npar, sizeof(double));
    for(size_t t = 0; t < n_participants; ++t)
    {
        G->part[t] = participants[t];
        for(int c = 0; c < npar; ++c) G->ppar[participants[t] * npar + c] = params[t * npar + c];
    }
    G->npp_w = (potential == ORC_PAIR_LJ) ? 2 : 1;
    G->current_margin = -1.0;
    return (int)(e->n_glo++);
}

/* ------------------------------------------------------------------------------------- */
/* topology: Topology.hpp:161-191 (add_connection), 243-264 (list_adjacent_within),         */
/* 280-317 (construct_molecules); write_topology of each local interaction                 */
/* ------------------------------------------------------------------------------------- */
static void add_connection(orc_engine* e, size_t i, size_t j, int kind)
{
    size_t ij[2][2] = {{i, j}, {j, i}};
    for(int s = 0; s < 2; ++s)
    {
        edgevec* v = &e->adjs[ij[s][0]];
        int found = 0;
        for(size_t k = 0; k < v->n; ++k) if(v->a[k].node == ij[s][1] && v->a[k].kind == kind) { found = 1; break; }
        if(found) continue;
        if(v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 4; v->a = (edge*)realloc(v->a, v->cap * sizeof(edge)); }
        v->a[v->n].node = ij[s][1]; v->a[v->n].kind = kind; v->n++;
    }
}

static void adjacent_within_rec(const orc_engine* e, size_t node, size_t dist, int kind, zvec* out)
{
    if(dist == 0) { zpush(out, node); return; }
    const edgevec* v = &e->adjs[node];
    for(size_t k = 0; k < v->n; ++k)
    {
        if(v->a[k].kind != kind) continue;
        zpush(out, v->a[k].node);
        adjacent_within_rec(e, v->a[k].node, dist - 1, kind, out);
    }
}

static void write_topology(orc_engine* e)
{
    for(size_t l = 0; l < e->n_loc; ++l)
    {
        local_t* L = &e->loc[l];
        if(L->topo_kind < 0) continue;
        if(L->kind == ORC_3SPN2_STACKING)
        {
            /* ThreeSPN2BaseStackingInteraction::write_topology (:125-179): every bead of the 5'
               nucleotide to every bead of the 3' one, phosphates when present */
            for(size_t t = 0; t < L->n; ++t)
            {
                const uint32_t* ix = L->idx + 6 * t;
                const uint32_t S5 = ix[0], B5 = ix[1], B3 = ix[2], P5 = ix[3], S3 = ix[4], P3 = ix[5];
                add_connection(e, S5, S3, L->topo_kind);
                add_connection(e, S5, B3, L->topo_kind);
                add_connection(e, B5, S3, L->topo_kind);
                add_connection(e, B5, B3, L->topo_kind);
                if(P5 != ORC_NIL) { add_connection(e, P5, S3, L->topo_kind); add_connection(e, P5, B3, L->topo_kind); }
                if(P3 != ORC_NIL) { add_connection(e, S5, P3, L->topo_kind); add_connection(e, B5, P3, L->topo_kind); }
                if(P5 != ORC_NIL && P3 != ORC_NIL) add_connection(e, P5, P3, L->topo_kind);
            }
            continue;
        }
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + t * L->arity;
            for(int a = 0; a < L->arity; ++a)
                for(int b = a + 1; b < L->arity; ++b)
                    add_connection(e, ix[a], ix[b], L->topo_kind);
        }
    }
    /* ConstraintForceField::write_topology (ConstraintForceField.hpp:49-60), after the locals
       (ForceField.hpp:53-56) */
    if(e->con_kind >= 0)
        for(size_t k = 0; k < e->n_con; ++k) add_connection(e, e->con_idx[2 * k], e->con_idx[2 * k + 1], e->con_kind);
    /* construct_molecules: only "bond" edges join molecules */
    const int bond = kind_id(e, "bond");
    const size_t UNINIT = (size_t)-1;
    e->mol = (size_t*)malloc(e->n * sizeof(size_t) + 1);
    for(size_t i = 0; i < e->n; ++i) e->mol[i] = UNINIT;
    size_t next = 0;
    for(size_t i = 0; i < e->n; ++i)
    {
        const edgevec* v = &e->adjs[i];
        for(size_t k = 0; k < v->n; ++k)
        {
            if(v->a[k].kind != bond) continue;
            if(e->mol[v->a[k].node] != UNINIT) { e->mol[i] = e->mol[v->a[k].node]; break; }
        }
        if(e->mol[i] == UNINIT) e->mol[i] = next++;
    }
    e->n_mol = next;
}

/* ExclusionList::make (ExclusionList.hpp:86-234) for one global interaction.  Groups are
 * all equal and no group is ignored, so the group test never excludes. */
static void make_exclusions(orc_engine* e, global_t* G)
{
    const size_t N = e->n;
    G->ex_start = (size_t*)malloc((N + 1) * sizeof(size_t));
    zvec all = {0};
    zvec tmp = {0};
    for(size_t i = 0; i < N; ++i)
    {
        G->ex_start[i] = all.n;
        tmp.n = 0;
        zpush(&tmp, i);
        for(size_t w = 0; w < G->n_within; ++w)
        {
            /* list_adjacent_within(i, dist, kind): {i} + recursive walk, sorted, unique */
            zvec adj = {0};
            zpush(&adj, i);
            if(G->within_dist[w] != 0)
            {
                const edgevec* v = &e->adjs[i];
                for(size_t k = 0; k < v->n; ++k)
                {
                    if(v->a[k].kind != G->within_kind[w]) continue;
                    zpush(&adj, v->a[k].node);
                    adjacent_within_rec(e, v->a[k].node, G->within_dist[w] - 1, G->within_kind[w], &adj);
                }
            }
            adj.n = sort_unique(adj.a, adj.n);
            for(size_t k = 0; k < adj.n; ++k) zpush(&tmp, adj.a[k]);
            free(adj.a);
        }
        tmp.n = sort_unique(tmp.a, tmp.n);
        for(size_t k = 0; k < tmp.n; ++k) zpush(&all, tmp.a[k]);
    }
    G->ex_start[N] = all.n;
    G->ex = all.a;
    free(tmp.a);
    G->mol = (size_t*)malloc(N * sizeof(size_t) + 1);
    memcpy(G->mol, e->mol, N * sizeof(size_t));
}

/* ------------------------------------------------------------------------------------- */
/* 3SPN2 parameters (the reference's default tables, kJ/mol and degrees as published, converted */
/* as the reference's initialize() does: epsilon * (1/4.184), degrees * (pi/180))            */
/* ------------------------------------------------------------------------------------- */
static const double TSP_PI = 3.14159265358979323846264338;       /* math/constants.hpp:13 */
static const double TSP_TWO_PI = 6.28318530717958647692528677;   /* math/constants.hpp:15 */
static const double TSP_TOL = 1e-8;                              /* math::abs_tolerance<double> */
static double tsp_J_to_cal(void) { return 1.0 / 4.184; }        /* Unit.hpp:43-44 */

/* ThreeSPN2BaseStackingPotentialParameter (ThreeSPN2BaseStackingPotential.hpp:298-326) */
static const double TSP_BS_EPS[16] = {14.39, 14.34, 13.25, 14.51, 10.37, 13.36, 10.34, 12.89,
                                      14.81, 15.57, 14.93, 15.39, 11.42, 12.79, 10.52, 13.24};
static const double TSP_BS_R0[16] = {3.716, 3.675, 3.827, 3.744, 4.238, 3.984, 4.416, 4.141,
                                     3.576, 3.598, 3.664, 3.635, 4.038, 3.798, 4.208, 3.935};
static const double TSP_BS_TH0[16] = {101.15, 85.94, 105.26, 89.00, 101.59, 89.50, 104.31, 91.28,
                                      100.89, 84.83, 105.48, 88.28, 106.49, 93.31, 109.54, 95.46};
/* ThreeSPN2CrossStackingGlobalPotentialParameter (ThreeSPN2CrossStackingPotential.hpp:516-566) */
static const double TSP_CS_EPS[32] = {2.186, 2.774, 2.833, 1.951, 2.774, 2.186, 2.539, 2.980,
                                      2.833, 2.539, 3.774, 1.129, 1.951, 2.980, 1.129, 4.802,
                                      2.186, 2.774, 2.980, 2.539, 2.774, 2.186, 1.951, 2.833,
                                      2.980, 1.951, 4.802, 1.129, 2.539, 2.833, 1.129, 3.774};
static const double TSP_CS_R0[32] = {6.208, 6.876, 6.072, 6.811, 6.876, 7.480, 6.771, 7.453,
                                     6.072, 6.771, 5.921, 6.688, 6.811, 7.453, 6.688, 7.409,
                                     5.435, 6.295, 5.183, 6.082, 6.295, 7.195, 6.028, 6.981,
                                     5.183, 6.028, 4.934, 5.811, 6.082, 6.981, 5.811, 6.757};
static const double TSP_CS_TH0[32] = {154.38, 159.10, 152.46, 158.38, 147.10, 153.79, 144.44, 151.48,
                                      154.69, 157.83, 153.43, 158.04, 152.99, 159.08, 150.53, 157.17,
                                      116.88, 121.74, 114.23, 119.06, 109.42, 112.95, 107.32, 110.56,
                                      119.34, 124.72, 116.51, 121.98, 114.60, 118.26, 112.45, 115.88};
/* ThreeSPN2ExcludedVolumePotentialParameter (ThreeSPN2ExcludedVolumePotential.hpp:437-446),
   bead kinds P, S, A, T, G, C */
static double tsp_ev_sigma(int kind)
{
    static const double sigma[6] = {4.5, 6.2, 5.4, 7.1, 4.9, 6.4};
    return sigma[kind];
}
static double tsp_bs_eps(int k) { return TSP_BS_EPS[k] * tsp_J_to_cal(); }
static double tsp_bs_theta0(int k) { return TSP_BS_TH0[k] * (TSP_PI / 180.0); }
static double tsp_cs_eps(int k) { return TSP_CS_EPS[k] * tsp_J_to_cal(); }
static double tsp_cs_theta0(int k) { return TSP_CS_TH0[k] * (TSP_PI / 180.0); }

/* base-pair kinds AT 0, TA 1, GC 2, CG 3 from bases A 0, T 1, G 2, C 3
   (ThreeSPN2BasePairParameterList::bp_kind, ThreeSPN2BasePairPotential.hpp:275-285) */
static int tsp_bp_kind(int a)
{
    switch(a) { case 0: return 0; case 1: return 1; case 3: return 3; default: return 2; }
}
/* ThreeSPN2BasePairGlobalPotentialParameter (ThreeSPN2BasePairPotential.hpp:530-560) */
static double tsp_bp_eps(int bp) { return (bp < 2 ? 16.73 : 21.18) * tsp_J_to_cal(); }
static double tsp_bp_r0(int bp) { return bp < 2 ? 5.941 : 5.530; }
static double tsp_bp_theta1(int bp)
{
    switch(bp)
    {
        case 0: return 156.54 / 180.0 * TSP_PI;
        case 1: return 135.78 / 180.0 * TSP_PI;
        case 2: return 159.81 / 180.0 * TSP_PI;
        default: return 141.16 / 180.0 * TSP_PI;
    }
}
static double tsp_bp_theta2(int bp)
{
    switch(bp)
    {
        case 0: return 135.78 / 180.0 * TSP_PI;
        case 1: return 156.54 / 180.0 * TSP_PI;
        case 2: return 141.16 / 180.0 * TSP_PI;
        default: return 159.81 / 180.0 * TSP_PI;
    }
}
static double tsp_bp_phi0(int bp) { return (bp < 2 ? -38.35 : -42.98) / 180.0 * TSP_PI; }
static double tsp_cs_theta3(int bp) { return (bp < 2 ? 116.09 : 124.94) / 180.0 * TSP_PI; }

/* f(theta) and df/dtheta of the 3SPN2 angular modulations (ThreeSPN2BasePairPotential.hpp:
   60-95, identical in the base-stacking and cross-stacking potentials) */
static double tsp_f(double K, double pi_over_K, double theta, double theta0)
{
    const double dtheta = theta - theta0;
    const double abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * 0.5) return 1.0;
    if(abs_dtheta < pi_over_K)
    {
        const double cos_Kdtheta = cos(K * dtheta);
        return 1.0 - cos_Kdtheta * cos_Kdtheta;
    }
    return 0.0;
}
static double tsp_df(double K, double pi_over_K, double theta, double theta0)
{
    const double dtheta = theta - theta0;
    const double abs_dtheta = fabs(dtheta);
    if(abs_dtheta < pi_over_K * 0.5) return 0.0;
    if(abs_dtheta < pi_over_K) return K * sin(2 * K * dtheta);
    return 0.0;
}

/* ExclusionList::is_excluded (ExclusionList.hpp:56-84) */
static int is_excluded(const global_t* G, size_t i, size_t j)
{
    const size_t mi = G->mol[i], mj = G->mol[j];
    if(G->ignore_mol == ORC_IGNORE_SELF && mi == mj) return 1;
    if(G->ignore_mol == ORC_IGNORE_OTHERS && mi != mj) return 1;
    for(size_t k = G->ex_start[i]; k < G->ex_start[i + 1]; ++k)
    {
        if(G->ex[k] > j) break;
        if(G->ex[k] == j) return 1;
    }
    return 0;
}
/* complementary bases A-T, G-C (codes A 0, T 1, G 2, C 3) */
static int complementary(int a, int b)
{
    return (a == 0 && b == 1) || (a == 1 && b == 0) || (a == 2 && b == 3) || (a == 3 && b == 2);
}
/* ParameterList.hpp:285-289; 3SPN2 EV: ThreeSPN2ExcludedVolumePotential.hpp:326-364 (base pairs that
 * could pair never interact; its within-3-nucleotides exception is keyed to a connection kind
 * no interaction writes, so it never applies); 3SPN2 BP / CS: ThreeSPN2BasePairPotential.hpp:
 * 445-459, ThreeSPN2CrossStackingPotential.hpp:437-451 (complementary bases only) */
static int has_interaction(const global_t* G, size_t i, size_t j)
{
    if(!((i < j) && !is_excluded(G, i, j))) return 0;
    if(G->pot == ORC_PAIR_3SPN2_EV)
    {
        const int ki = (int)G->ppar[i], kj = (int)G->ppar[j];
        if(ki < 2 || kj < 2) return 1;   /* a phosphate or a sugar */
        return !complementary(ki - 2, kj - 2);
    }
    if(G->pot == ORC_3SPN2_BASEPAIR || G->pot == ORC_3SPN2_CROSSSTACK)
    {
        const int w = G->npar;
        return complementary((int)G->ppar[i * w], (int)G->ppar[j * w]);
    }
    return 1;
}

/* ------------------------------------------------------------------------------------- */
/* potentials' derived constants                                                           */
/* ------------------------------------------------------------------------------------- */
static int setup_potential(orc_engine* e, global_t* G)
{
    if(G->pot == ORC_PAIR_LJ)
    {
        /* LennardJonesPotential.hpp:34-38, max_cutoff :82-98 */
        G->coef_at_cutoff = pow(1.0 / G->cutoff_ratio, 12) - pow(1.0 / G->cutoff_ratio, 6);
        double ms = 0;
        for(size_t i = 0; i < e->n; ++i) ms = fmax(ms, G->ppar[2 * i]);
        G->max_cutoff = ms * G->cutoff_ratio;
    }
    else if(G->pot == ORC_PAIR_EV)
    {
        /* ExcludedVolumePotential.hpp:44-48, max_cutoff :90-106 */
        G->coef_at_cutoff = pow(1.0 / G->cutoff_ratio, 12);
        double mr = 0;
        for(size_t i = 0; i < e->n; ++i) mr = fmax(mr, G->ppar[i]);
        G->max_cutoff = mr * 2 * G->cutoff_ratio;
    }
    else if(G->pot == ORC_PAIR_3SPN2_EV)
    {
        /* ThreeSPN2ExcludedVolumePotential.hpp:120-134 (epsilon 1 kJ/mol in kcal/mol), the list
           cutoff is the largest sigma of the table (:176-178) */
        G->eps = 1.0 * tsp_J_to_cal();
        G->max_cutoff = 7.1;
    }
    else if(G->pot == ORC_3SPN2_BASEPAIR || G->pot == ORC_3SPN2_CROSSSTACK)
    {
        G->max_cutoff = 18.0;   /* ThreeSPN2BasePairGlobalPotentialParameter::cutoff */
    }
    else
    {
        /* DebyeHuckelPotential.hpp:126-174 */
        if(!e->has_T || !e->has_I) return -1;
        const double pi = 3.14159265358979323846;
        const double kB = orc_kB(), NA = U_NA, eps0 = orc_eps0();
        const double T = e->temperature, c = e->ionic_strength;
        const double epsk = (249.4 - 0.788 * T + 7.2e-4 * T * T) *
                            (1. - 2.551e-1 * c + 5.151e-2 * c * c - 6.889e-3 * c * c * c);
        G->inv4pieps = 1.0 / (4 * pi * eps0 * epsk);
        const double I = c / orc_L_to_volume();
        G->debye = sqrt((eps0 * epsk * kB * T) / (2 * NA * I));
        G->inv_debye = 1. / G->debye;
        G->abs_cutoff = G->cutoff_ratio * G->debye;
        G->coef_at_cutoff = exp(-G->cutoff_ratio) / (G->debye * G->cutoff_ratio);
        G->max_cutoff = G->debye * G->cutoff_ratio;
    }
    return 0;
}

/* prepare_params: LorentzBerthelotRule (ParameterList.hpp:240-252), EV (ExcludedVolume-
 * Potential.hpp:192-195), DH (DebyeHuckelPotential.hpp:249-254) */
static void prepare_params(const global_t* G, size_t i, size_t j, double* out)
{
    if(G->pot == ORC_PAIR_LJ)
    {
        const double s1 = G->ppar[2 * i], e1 = G->ppar[2 * i + 1];
        const double s2 = G->ppar[2 * j], e2 = G->ppar[2 * j + 1];
        out[0] = (s1 + s2) / 2;
        out[1] = (e1 == e2) ? e1 : sqrt(e1 * e2);
    }
    else if(G->pot == ORC_PAIR_EV)
    {
        out[0] = G->ppar[i] + G->ppar[j];
    }
    else if(G->pot == ORC_PAIR_3SPN2_EV)
    {
        /* ThreeSPN2ExcludedVolumeParameterList::prepare_params (:198-202) */
        out[0] = (tsp_ev_sigma((int)G->ppar[i]) + tsp_ev_sigma((int)G->ppar[j])) * 0.5;
    }
    else if(G->pot == ORC_3SPN2_BASEPAIR || G->pot == ORC_3SPN2_CROSSSTACK)
    {
        out[0] = 0.0;   /* the pair's kinds are read from the per-base parameters */
    }
    else
    {
        out[0] = G->inv4pieps * G->ppar[i] * G->ppar[j];
    }
}

/* ------------------------------------------------------------------------------------- */
/* neighbour lists                                                                         */
/* ------------------------------------------------------------------------------------- */
static void nl_add(global_t* G, size_t j, const double* pp)
{
    zpush(&G->nbr, j);
    if(G->nbr.n * G->npp_w > G->npp_cap)
    {
        G->npp_cap = 2 * G->nbr.n * G->npp_w + 16;
        G->npp = (double*)realloc(G->npp, G->npp_cap * sizeof(double));
    }
    for(int c = 0; c < G->npp_w; ++c) G->npp[(G->nbr.n - 1) * G->npp_w + c] = pp[c];
}

typedef struct { size_t idx, cell; } pc_pair;
static int cmp_cell(const void* a, const void* b)
{
    const pc_pair* x = (const pc_pair*)a; const pc_pair* y = (const pc_pair*)b;
    if(x->cell != y->cell) return (x->cell < y->cell) ? -1 : 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx); /* grouping only; order inside a cell
                                                           is irrelevant after partner sort */
}
typedef struct { size_t j; double p[2]; } partner_t;
static int cmp_partner(const void* a, const void* b)
{
    const partner_t* x = (const partner_t*)a; const partner_t* y = (const partner_t*)b;
    return (x->j < y->j) ? -1 : (x->j > y->j);
}

/* PeriodicGridCellList::calc_index (PeriodicGridCellList.hpp:133-145) */
static size_t pgrid_index(const global_t* G, const double* pos)
{
    double ofs[3] = {pos[0] - G->sys_lo[0], pos[1] - G->sys_lo[1], pos[2] - G->sys_lo[2]};
    double fx = floor(ofs[0] * G->r_x), fy = floor(ofs[1] * G->r_y), fz = floor(ofs[2] * G->r_z);
    size_t ix = (size_t)fx, iy = (size_t)fy, iz = (size_t)fz;
    if(ix > G->dim_x - 1) ix = G->dim_x - 1;
    if(iy > G->dim_y - 1) iy = G->dim_y - 1;
    if(iz > G->dim_z - 1) iz = G->dim_z - 1;
    return ix + G->dim_x * iy + G->dim_x * G->dim_y * iz;
}
/* UnlimitedGridCellList::calc_index (UnlimitedGridCellList.hpp:105-119) */
static size_t ugrid_index(const global_t* G, const double* pos)
{
    const int64_t d = 8;
    int64_t x = ((int64_t)floor(pos[0] * G->r_cell)) % d;
    int64_t y = ((int64_t)floor(pos[1] * G->r_cell)) % d;
    int64_t z = ((int64_t)floor(pos[2] * G->r_cell)) % d;
    if(x < 0) x += d;
    if(y < 0) y += d;
    if(z < 0) z += d;
    return (size_t)(x + 8 * y + 64 * z);
}

/* cell connectivity: PeriodicGridCellList::construct_cells (:354-437) and
 * UnlimitedGridCellList::initialize (:257-319); 27 cells in the same order */
static void build_adjacency(global_t* G, size_t dx, size_t dy, size_t dz)
{
    free(G->adj);
    G->adj = (size_t*)malloc(dx * dy * dz * 27 * sizeof(size_t));
    for(size_t x = 0; x < dx; ++x)
    for(size_t y = 0; y < dy; ++y)
    for(size_t z = 0; z < dz; ++z)
    {
        const size_t c = x + dx * y + dx * dy * z;
        const size_t xs[3] = {(x == 0) ? dx - 1 : x - 1, x, (x == dx - 1) ? 0 : x + 1};
        const size_t ys[3] = {(y == 0) ? dy - 1 : y - 1, y, (y == dy - 1) ? 0 : y + 1};
        const size_t zs[3] = {(z == 0) ? dz - 1 : z - 1, z, (z == dz - 1) ? 0 : z + 1};
        int k = 0;
        for(int c2 = 0; c2 < 3; ++c2)
        for(int b2 = 0; b2 < 3; ++b2)
        for(int a2 = 0; a2 < 3; ++a2)
            G->adj[27 * c + (k++)] = xs[a2] + dx * ys[b2] + dx * dy * zs[c2];
    }
}

/* cell-list make: PeriodicGridCellList::make (:198-351) / UnlimitedGridCellList::make
 * (:152-255).  Both yield, for each leading participant i, the sorted list of participants
 * j in the 27 surrounding cells with has_interaction(i, j) and |r_ij|^2 < (r_c(1+m))^2. */
static void make_celllist(orc_engine* e, global_t* G)
{
    size_t ncell;
    if(e->periodic)
    {
        int changed = 0;
        for(int d = 0; d < 3; ++d)
            if(G->sys_lo[d] != e->lower[d] || G->sys_size[d] != e->width[d]) changed = 1;
        if(changed)
        {
            for(int d = 0; d < 3; ++d) { G->sys_lo[d] = e->lower[d]; G->sys_size[d] = e->width[d]; }
            const double me = 1e-6;
            G->cutoff = G->max_cutoff;
            G->r_x = G->r_y = G->r_z = 1 / (G->cutoff * (1 + G->margin) + me);
            double fx = floor(G->sys_size[0] * G->r_x), fy = floor(G->sys_size[1] * G->r_y), fz = floor(G->sys_size[2] * G->r_z);
            size_t dx = (size_t)fx < 3 ? 3 : (size_t)fx;
            size_t dy = (size_t)fy < 3 ? 3 : (size_t)fy;
            size_t dz = (size_t)fz < 3 ? 3 : (size_t)fz;
            if(dx != G->dim_x || dy != G->dim_y || dz != G->dim_z)
            {
                G->dim_x = dx; G->dim_y = dy; G->dim_z = dz;
                build_adjacency(G, dx, dy, dz);
            }
            G->r_x = 1.0 / (G->sys_size[0] / G->dim_x);
            G->r_y = 1.0 / (G->sys_size[1] / G->dim_y);
            G->r_z = 1.0 / (G->sys_size[2] / G->dim_z);
        }
        ncell = G->dim_x * G->dim_y * G->dim_z;
    }
    else
    {
        ncell = 512;
    }

    pc_pair* ibc = (pc_pair*)malloc(G->np * sizeof(pc_pair) + 1);
    for(size_t k = 0; k < G->np; ++k)
    {
        ibc[k].idx = G->part[k];
        ibc[k].cell = e->periodic ? pgrid_index(G, e->pos + 3 * G->part[k]) : ugrid_index(G, e->pos + 3 * G->part[k]);
    }
    qsort(ibc, G->np, sizeof(pc_pair), cmp_cell);
    size_t* cs = (size_t*)calloc(ncell + 1, sizeof(size_t));
    for(size_t k = 0; k < G->np; ++k) cs[ibc[k].cell + 1]++;
    for(size_t c = 0; c < ncell; ++c) cs[c + 1] += cs[c];

    const double r_c = G->cutoff * (1. + G->margin);
    const double r_c2 = r_c * r_c;
    G->nbr.n = 0;
    free(G->rng_start);
    G->rng_start = (size_t*)malloc((G->np + 1) * sizeof(size_t));
    partner_t* buf = NULL; size_t bcap = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        G->rng_start[k] = G->nbr.n;
        const size_t i = G->part[k];
        const double* ri = e->pos + 3 * i;
        const size_t cell = e->periodic ? pgrid_index(G, ri) : ugrid_index(G, ri);
        size_t nb = 0;
        for(int a = 0; a < 27; ++a)
        {
            const size_t c = G->adj[27 * cell + a];
            for(size_t q = cs[c]; q < cs[c + 1]; ++q)
            {
                const size_t j = ibc[q].idx;
                if(!has_interaction(G, i, j)) continue;
                double d[3];
                adjust_direction(e, ri, e->pos + 3 * j, d);
                if(length_sq(d) < r_c2)
                {
                    if(nb == bcap) { bcap = bcap ? 2 * bcap : 64; buf = (partner_t*)realloc(buf, bcap * sizeof(partner_t)); }
                    buf[nb].j = j;
                    prepare_params(G, i, j, buf[nb].p);
                    nb++;
                }
            }
        }
        qsort(buf, nb, sizeof(partner_t), cmp_partner);
        for(size_t q = 0; q < nb; ++q) nl_add(G, buf[q].j, buf[q].p);
    }
    G->rng_start[nlead] = G->nbr.n;
    free(buf); free(ibc); free(cs);
    G->current_margin = G->cutoff * G->margin;
}

/* VerletList::make (VerletList.hpp:98-134) and NaivePairCalculation::make
 * (NaivePairCalculation.hpp:66-90): partners taken from the participants after k */
static void make_pairwise(orc_engine* e, global_t* G, int use_distance)
{
    const double rc = G->cutoff * (1. + G->margin);
    const double rc2 = rc * rc;
    G->nbr.n = 0;
    free(G->rng_start);
    G->rng_start = (size_t*)malloc((G->np + 1) * sizeof(size_t));
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        G->rng_start[k] = G->nbr.n;
        const size_t i = G->part[k];
        for(size_t q = k + 1; q < G->np; ++q)
        {
            const size_t j = G->part[q];
            if(!has_interaction(G, i, j)) continue;
            if(use_distance)
            {
                double d[3];
                adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
                if(!(length_sq(d) < rc2)) continue;
            }
            double pp[2];
            prepare_params(G, i, j, pp);
            nl_add(G, j, pp);
        }
    }
    G->rng_start[nlead] = G->nbr.n;
    G->current_margin = use_distance ? G->cutoff * G->margin : INFINITY;
}

static void make_list(orc_engine* e, global_t* G)
{
    e->rebuilds++;
    if(G->partition == ORC_PART_CELLLIST) make_celllist(e, G);
    else if(G->partition == ORC_PART_VERLET) make_pairwise(e, G, 1);
    else make_pairwise(e, G, 0);
}

/* Contact active list: ContactInteraction.hpp:159-183 */
static void contact_make_list(orc_engine* e, local_t* L)
{
    L->active.n = 0;
    const double abs_margin = L->cutoff * L->margin;
    for(size_t t = 0; t < L->n; ++t)
    {
        double d[3];
        adjust_direction(e, e->pos + 3 * L->idx[2 * t], e->pos + 3 * L->idx[2 * t + 1], d);
        const double len2 = length_sq(d);
        const double rc = L->par[2 * t + 1] * 2.5 + abs_margin;
        if(len2 < rc * rc) zpush(&L->active, t);
    }
    L->current_margin = L->cutoff * L->margin;
}

int orc_add_constraint(orc_engine* e, const char* topology, size_t n, const uint32_t* indices, const double* v0,
                       size_t max_iteration, double tolerance)
{
    if(e->con_idx != NULL) return -1;
    for(size_t k = 0; k < 2 * n; ++k) if(indices[k] >= e->n) return -1;
    e->n_con = n;
    e->con_idx = (uint32_t*)malloc(2 * n * sizeof(uint32_t) + 1);
    e->con_v0 = (double*)malloc(n * sizeof(double) + 1);
    memcpy(e->con_idx, indices, 2 * n * sizeof(uint32_t));
    memcpy(e->con_v0, v0, n * sizeof(double));
    e->con_kind = (topology == NULL || topology[0] == '\0' || strcmp(topology, "none") == 0) ? -1 : kind_id(e, topology);
    e->con_maxit = max_iteration;
    e->con_tol = tolerance;
    return 0;
}

int orc_initialize(orc_engine* e)
{
    write_topology(e);
    for(size_t l = 0; l < e->n_loc; ++l)
    {
        local_t* L = &e->loc[l];
        if(L->kind == ORC_CONTACT_GOCONTACT && L->n > 0)
        {
            double c = 0;
            for(size_t t = 0; t < L->n; ++t) c = fmax(c, L->par[2 * t + 1] * 2.5);
            L->cutoff = c;
            contact_make_list(e, L);
        }
    }
    for(size_t g = 0; g < e->n_glo; ++g)
    {
        global_t* G = &e->glo[g];
        if(setup_potential(e, G) != 0) return -1;
        make_exclusions(e, G);
        if(G->partition == ORC_PART_CELLLIST)
        {
            if(e->periodic)
            {
                for(int d = 0; d < 3; ++d) { G->sys_lo[d] = NAN; G->sys_size[d] = NAN; }
            }
            else
            {
                G->cutoff = G->max_cutoff;
                G->r_cell = 1 / (G->cutoff * (1 + G->margin) * (1 + 1e-6));
                build_adjacency(G, 8, 8, 8);
            }
        }
        else
        {
            G->cutoff = G->max_cutoff;
        }
        make_list(e, G);
    }
    e->rebuilds = 0;
    return 0;
}

size_t orc_num_interactions(const orc_engine* e) { return e->n_loc + e->n_glo; }
size_t orc_rebuild_count(const orc_engine* e) { return e->rebuilds; }
uint64_t orc_step_count(const orc_engine* e) { return e->step; }

size_t orc_pairs_of(const orc_engine* e, int gid, uint32_t* out, size_t cap)
{
    if(gid < 0 || (size_t)gid >= e->n_glo) return 0;
    const global_t* G = &e->glo[gid];
    if(out)
    {
        size_t w = 0;
        for(size_t k = 0; k + 1 < G->np; ++k)
            for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
            {
                if(w < cap) { out[2 * w] = (uint32_t)G->part[k]; out[2 * w + 1] = (uint32_t)G->nbr.a[q]; }
                w++;
            }
    }
    return G->nbr.n;
}

/* ------------------------------------------------------------------------------------- */
/* forces and energies                                                                     */
/* ------------------------------------------------------------------------------------- */
/* need_energy is a compile-time constant after inlining (as the reference's NeedEnergy
 * template argument is), so the force-only variant carries no energy-only libm calls and the
 * compiler pairs sin/cos into sincos exactly where the reference's instantiations do. */
static inline __attribute__((always_inline))
double local_force_impl(orc_engine* e, const local_t* L, double* F, const int need_energy)
{
    double energy = 0;
    if(L->kind == ORC_BOND_HARMONIC)
    {
        /* BondLengthInteraction.hpp:126-148; HarmonicPotential.hpp:23-32 */
        for(size_t t = 0; t < L->n; ++t)
        {
            const size_t i0 = L->idx[2 * t], i1 = L->idx[2 * t + 1];
            const double k = L->par[2 * t], v0 = L->par[2 * t + 1];
            double d[3];
            adjust_direction(e, e->pos + 3 * i0, e->pos + 3 * i1, d);
            const double len2 = length_sq(d);
            const double rlen = rsqrt(len2);
            const double len = len2 * rlen;
            const double force = -1 * (2 * k * (len - v0));
            if(need_energy) { const double dv = len - v0; energy += k * dv * dv; }
            const double c = force * rlen;
            for(int a = 0; a < 3; ++a) { F[3 * i0 + a] -= d[a] * c; F[3 * i1 + a] += d[a] * c; }
        }
    }
    else if(L->kind == ORC_BOND_GOCONTACT || L->kind == ORC_CONTACT_GOCONTACT)
    {
        /* BondLengthGoContactInteraction.hpp:126-167, GoContactInteraction.hpp:185-228 */
        const size_t nt = (L->kind == ORC_CONTACT_GOCONTACT) ? L->active.n : L->n;
        for(size_t q = 0; q < nt; ++q)
        {
            const size_t t = (L->kind == ORC_CONTACT_GOCONTACT) ? L->active.a[q] : q;
            const size_t i0 = L->idx[2 * t], i1 = L->idx[2 * t + 1];
            const double k = L->par[2 * t], v0 = L->par[2 * t + 1];
            double d[3];
            adjust_direction(e, e->pos + 3 * i0, e->pos + 3 * i1, d);
            const double len2 = length_sq(d);
            const double cut = v0 * 2.5;
            if(cut * cut <= len2) continue;
            const double r2 = 1.0 / len2;
            const double v0r_2 = v0 * v0 * r2;
            const double v0r_6 = v0r_2 * v0r_2 * v0r_2;
            const double v0r_10 = v0r_6 * v0r_2 * v0r_2;
            const double v0r_12 = v0r_10 * v0r_2;
            if(need_energy) energy += k * (5 * v0r_12 - 6 * v0r_10);
            const double coef = -60 * k * r2 * (v0r_10 - v0r_12);
            for(int a = 0; a < 3; ++a) { F[3 * i0 + a] -= coef * d[a]; F[3 * i1 + a] += coef * d[a]; }
        }
    }
    else if(L->kind == ORC_ANGLE_HARMONIC || L->kind == ORC_ANGLE_GAUSSIAN)
    {
        /* BondAngleInteraction.hpp:129-189 */
        const double abs_tol = 1e-8; /* math::abs_tolerance<double> (math/constants.hpp) */
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 3 * t;
            const double* p = L->par + L->npar * t;
            const double* p0 = e->pos + 3 * ix[0];
            const double* p1 = e->pos + 3 * ix[1];
            const double* p2 = e->pos + 3 * ix[2];
            double r_ij[3], r_kj[3], eij[3], ekj[3];
            adjust_direction(e, p1, p0, r_ij);
            const double inv_ij = rsqrt(length_sq(r_ij));
            for(int a = 0; a < 3; ++a) eij[a] = r_ij[a] * inv_ij;
            adjust_direction(e, p1, p2, r_kj);
            const double inv_kj = rsqrt(length_sq(r_kj));
            for(int a = 0; a < 3; ++a) ekj[a] = r_kj[a] * inv_kj;
            const double cos_theta = clampd(dot(eij, ekj), -1, 1);
            const double theta = acos(cos_theta);
            double coef, pot;
            if(L->kind == ORC_ANGLE_HARMONIC)
            {
                coef = -(2 * p[0] * (theta - p[1]));
                const double dv = theta - p[1];
                pot = p[0] * dv * dv;
            }
            else
            {
                /* GaussianPotential.hpp:20-38 */
                const double k = p[0], sigma = p[1], v0 = p[2];
                const double inv_sigma2 = -1. / (2 * sigma * sigma);
                const double dval = theta - v0;
                coef = -(2 * inv_sigma2 * dval * k * exp(inv_sigma2 * dval * dval));
                pot = k * exp(inv_sigma2 * dval * dval);
            }
            if(need_energy) energy += pot;
            const double sin_theta = sin(theta);
            const double cis = coef / fmax(sin_theta, abs_tol);
            double Fi[3], Fk[3];
            const double ci = cis * inv_ij, ck = cis * inv_kj;
            for(int a = 0; a < 3; ++a)
            {
                Fi[a] = ci * (cos_theta * eij[a] - ekj[a]);
                Fk[a] = ck * (cos_theta * ekj[a] - eij[a]);
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * ix[0] + a] += Fi[a];
                F[3 * ix[1] + a] += (-Fi[a] - Fk[a]);
                F[3 * ix[2] + a] += Fk[a];
            }
        }
    }
    else if(L->kind == ORC_DIHEDRAL_CLEMENTI)
    {
        /* DihedralAngleInteraction.hpp:157-227; ClementiDihedralPotential.hpp:27-42 */
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 4 * t;
            const double* p = L->par + 3 * t;
            const double *r_i = e->pos + 3 * ix[0], *r_j = e->pos + 3 * ix[1];
            const double *r_k = e->pos + 3 * ix[2], *r_l = e->pos + 3 * ix[3];
            double r_ij[3], r_kj[3], r_lk[3], r_kl[3], m[3], n[3];
            adjust_direction(e, r_j, r_i, r_ij);
            adjust_direction(e, r_j, r_k, r_kj);
            adjust_direction(e, r_k, r_l, r_lk);
            for(int a = 0; a < 3; ++a) r_kl[a] = -1.0 * r_lk[a];
            const double r_kj_lensq = length_sq(r_kj);
            const double r_kj_rlen = rsqrt(r_kj_lensq);
            const double r_kj_rlensq = r_kj_rlen * r_kj_rlen;
            const double r_kj_len = r_kj_rlen * r_kj_lensq;
            cross(r_ij, r_kj, m);
            cross(r_kj, r_kl, n);
            const double m_lensq = length_sq(m), n_lensq = length_sq(n);
            const double dot_mn = dot(m, n) * rsqrt(m_lensq * n_lensq);
            const double cos_phi = clampd(dot_mn, -1, 1);
            const double phi = copysign(acos(cos_phi), dot(r_ij, n));
            const double k1 = p[0], k3 = p[1], v0 = p[2];
            const double dphi = phi - v0;
            const double sin1 = sin(dphi);
            const double sin3 = sin1 * (3.0 - 4.0 * sin1 * sin1);
            const double coef = -(k1 * sin1 + 3.0 * k3 * sin3);
            if(need_energy)
            {
                const double cos1 = cos(dphi);
                const double cos3 = cos1 * (4.0 * cos1 * cos1 - 3.0);
                energy += k1 * (1.0 - cos1) + k3 * (1.0 - cos3);
            }
            double Fi[3], Fl[3], Fj[3], Fk[3];
            const double ci = coef * r_kj_len / m_lensq;
            const double cl = -coef * r_kj_len / n_lensq;
            for(int a = 0; a < 3; ++a) { Fi[a] = ci * m[a]; Fl[a] = cl * n[a]; }
            const double coef_ijk = dot(r_ij, r_kj) * r_kj_rlensq;
            const double coef_jkl = dot(r_kl, r_kj) * r_kj_rlensq;
            for(int a = 0; a < 3; ++a)
            {
                Fj[a] = (coef_ijk - 1.0) * Fi[a] - coef_jkl * Fl[a];
                Fk[a] = (coef_jkl - 1.0) * Fl[a] - coef_ijk * Fi[a];
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * ix[0] + a] += Fi[a];
                F[3 * ix[1] + a] += Fj[a];
                F[3 * ix[2] + a] += Fk[a];
                F[3 * ix[3] + a] += Fl[a];
            }
        }
    }
    else if(L->kind == ORC_BOND_3SPN2)
    {
        /* BondLengthInteraction.hpp:126-148; ThreeSPN2BondPotential.hpp:22-35 */
        for(size_t t = 0; t < L->n; ++t)
        {
            const size_t i0 = L->idx[2 * t], i1 = L->idx[2 * t + 1];
            const double k = L->par[2 * t], v0 = L->par[2 * t + 1];
            double d[3];
            adjust_direction(e, e->pos + 3 * i0, e->pos + 3 * i1, d);
            const double len2 = length_sq(d);
            const double rlen = rsqrt(len2);
            const double len = len2 * rlen;
            const double dv = len - v0;
            const double force = -1 * (2 * k * (dv + 200 * dv * dv * dv));
            if(need_energy) { const double dv2 = dv * dv; energy += k * (dv2 + 100 * dv2 * dv2); }
            const double c = force * rlen;
            for(int a = 0; a < 3; ++a) { F[3 * i0 + a] -= d[a] * c; F[3 * i1 + a] += d[a] * c; }
        }
    }
    else if(L->kind == ORC_DIHEDRAL_PGAUSSIAN)
    {
        /* DihedralAngleInteraction.hpp:157-227; PeriodicGaussianPotential.hpp:21-44 */
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 4 * t;
            const double* p = L->par + 3 * t;
            const double *r_i = e->pos + 3 * ix[0], *r_j = e->pos + 3 * ix[1];
            const double *r_k = e->pos + 3 * ix[2], *r_l = e->pos + 3 * ix[3];
            double r_ij[3], r_kj[3], r_lk[3], r_kl[3], m[3], n[3];
            adjust_direction(e, r_j, r_i, r_ij);
            adjust_direction(e, r_j, r_k, r_kj);
            adjust_direction(e, r_k, r_l, r_lk);
            for(int a = 0; a < 3; ++a) r_kl[a] = -1.0 * r_lk[a];
            const double r_kj_lensq = length_sq(r_kj);
            const double r_kj_rlen = rsqrt(r_kj_lensq);
            const double r_kj_rlensq = r_kj_rlen * r_kj_rlen;
            const double r_kj_len = r_kj_rlen * r_kj_lensq;
            cross(r_ij, r_kj, m);
            cross(r_kj, r_kl, n);
            const double m_lensq = length_sq(m), n_lensq = length_sq(n);
            const double dot_mn = dot(m, n) * rsqrt(m_lensq * n_lensq);
            const double cos_phi = clampd(dot_mn, -1, 1);
            const double phi = copysign(acos(cos_phi), dot(r_ij, n));
            const double k = p[0], sigma = p[1], v0 = p[2];
            const double inv_sigma2 = -1. / (2 * sigma * sigma);
            double dval = phi - v0;
            if(dval < -TSP_PI) { dval += TSP_TWO_PI; } else if(TSP_PI < dval) { dval -= TSP_TWO_PI; }
            const double coef = -(2 * inv_sigma2 * dval * k * exp(inv_sigma2 * dval * dval));
            if(need_energy)
            {
                double dv = phi - v0;
                if(dv < -TSP_PI) { dv += TSP_TWO_PI; } else if(TSP_PI < dv) { dv -= TSP_TWO_PI; }
                energy += k * exp(inv_sigma2 * dv * dv);
            }
            double Fi[3], Fl[3], Fj[3], Fk[3];
            const double ci = coef * r_kj_len / m_lensq;
            const double cl = -coef * r_kj_len / n_lensq;
            for(int a = 0; a < 3; ++a) { Fi[a] = ci * m[a]; Fl[a] = cl * n[a]; }
            const double coef_ijk = dot(r_ij, r_kj) * r_kj_rlensq;
            const double coef_jkl = dot(r_kl, r_kj) * r_kj_rlensq;
            for(int a = 0; a < 3; ++a)
            {
                Fj[a] = (coef_ijk - 1.0) * Fi[a] - coef_jkl * Fl[a];
                Fk[a] = (coef_jkl - 1.0) * Fl[a] - coef_ijk * Fi[a];
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * ix[0] + a] += Fi[a];
                F[3 * ix[1] + a] += Fj[a];
                F[3 * ix[2] + a] += Fk[a];
                F[3 * ix[3] + a] += Fl[a];
            }
        }
    }
    else if(L->kind == ORC_3SPN2_STACKING)
    {
        /* ThreeSPN2BaseStackingInteraction::calc_force_energy_virial_impl (:281-428) with
           ThreeSPN2BaseStackingPotential (:172-294) */
        const double K = 6.0, piK = TSP_PI / 6.0, alp = 3.0;
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 6 * t;
            const size_t Si = ix[0], Bi = ix[1], Bj = ix[2];
            const int bs = (int)L->par[t];
            const double eps = tsp_bs_eps(bs), r0 = TSP_BS_R0[bs], th0 = tsp_bs_theta0(bs);
            double Bji[3], SBi[3], Bji_reg[3], SBi_reg[3];
            adjust_direction(e, e->pos + 3 * Bj, e->pos + 3 * Bi, Bji);
            adjust_direction(e, e->pos + 3 * Si, e->pos + 3 * Bi, SBi);
            const double lBji_sq = length_sq(Bji);
            const double rlBji = rsqrt(lBji_sq);
            const double lBji = lBji_sq * rlBji;
            for(int a = 0; a < 3; ++a) Bji_reg[a] = Bji[a] * rlBji;
            double f_Bi[3] = {0, 0, 0}, f_Bj[3] = {0, 0, 0}, f_Si[3] = {0, 0, 0};
            double dU_rep = 0.0;
            if(!(r0 <= lBji)) { const double term = exp(-alp * (lBji - r0)); dU_rep = 2 * alp * eps * term * (1.0 - term); }
            if(dU_rep != 0.0)
            {
                for(int a = 0; a < 3; ++a) { f_Bi[a] -= dU_rep * Bji_reg[a]; f_Bj[a] += dU_rep * Bji_reg[a]; }
            }
            if(need_energy)
            {
                if(!(r0 <= lBji)) { const double term = 1.0 - exp(-alp * (lBji - r0)); energy += eps * term * term; }
            }
            const double lSBi_sq = length_sq(SBi);
            const double rlSBi = rsqrt(lSBi_sq);
            for(int a = 0; a < 3; ++a) SBi_reg[a] = SBi[a] * rlSBi;
            const double cos_theta = dot(SBi_reg, Bji_reg);
            const double theta = acos(clampd(cos_theta, -1, 1));
            const double f_theta = tsp_f(K, piK, theta, th0);
            if(f_theta == 0.0)
            {
                for(int a = 0; a < 3; ++a) { F[3 * Bi + a] += f_Bi[a]; F[3 * Bj + a] += f_Bj[a]; }
                continue;
            }
            const double df_theta = tsp_df(K, piK, theta, th0);
            double U_attr, dU_attr;
            if(lBji <= r0) { U_attr = -eps; dU_attr = 0.0; }
            else
            {
                const double term = exp(-alp * (lBji - r0));
                U_attr = eps * (term * term - 2 * term);
                dU_attr = 2 * alp * eps * term * (1.0 - term);
            }
            if(need_energy) energy += U_attr * f_theta;
            if(df_theta != 0.0)
            {
                const double coef = -df_theta * U_attr;
                const double sin_theta = sin(theta);
                const double coef_rsin = (sin_theta > TSP_TOL) ? coef / sin_theta : coef / TSP_TOL;
                double fSi[3], fBj[3];
                for(int a = 0; a < 3; ++a)
                {
                    fSi[a] = (coef_rsin * rlSBi) * (Bji_reg[a] - cos_theta * SBi_reg[a]);
                    fBj[a] = (coef_rsin * rlBji) * (SBi_reg[a] - cos_theta * Bji_reg[a]);
                }
                for(int a = 0; a < 3; ++a) { f_Si[a] += fSi[a]; f_Bi[a] -= (fSi[a] + fBj[a]); f_Bj[a] += fBj[a]; }
            }
            if(dU_attr != 0.0)
            {
                const double coef = f_theta * dU_attr;
                for(int a = 0; a < 3; ++a) { f_Bi[a] -= coef * Bji_reg[a]; f_Bj[a] += coef * Bji_reg[a]; }
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * Bi + a] += f_Bi[a];
                F[3 * Bj + a] += f_Bj[a];
                F[3 * Si + a] += f_Si[a];
            }
        }
    }
    return energy;
}

static double local_force_f(orc_engine* e, const local_t* L, double* F) { return local_force_impl(e, L, F, 0); }
static double local_force_fe(orc_engine* e, const local_t* L, double* F) { return local_force_impl(e, L, F, 1); }
static double local_force(orc_engine* e, const local_t* L, double* F, int need_energy)
{
    return need_energy ? local_force_fe(e, L, F) : local_force_f(e, L, F);
}

/* calc_energy of local interactions: BondLengthInteraction.hpp:112-124 (and the GoContact
 * specialisation :46-55), GoContactInteraction.hpp:48-58, BondAngleInteraction.hpp:191-217,
 * DihedralAngleInteraction.hpp:125-155 */
static double local_energy(orc_engine* e, const local_t* L)
{
    double E = 0;
    if(L->kind == ORC_BOND_HARMONIC || L->kind == ORC_BOND_GOCONTACT || L->kind == ORC_CONTACT_GOCONTACT)
    {
        const size_t nt = (L->kind == ORC_CONTACT_GOCONTACT) ? L->active.n : L->n;
        for(size_t q = 0; q < nt; ++q)
        {
            const size_t t = (L->kind == ORC_CONTACT_GOCONTACT) ? L->active.a[q] : q;
            double d[3];
            adjust_direction(e, e->pos + 3 * L->idx[2 * t], e->pos + 3 * L->idx[2 * t + 1], d);
            const double r = sqrt(length_sq(d));
            const double k = L->par[2 * t], v0 = L->par[2 * t + 1];
            if(L->kind == ORC_BOND_HARMONIC)
            {
                const double dv = r - v0;
                E += k * dv * dv;
            }
            else
            {
                /* GoContactPotential.hpp:26-35 */
                if(v0 * 2.5 <= r) { E += 0.0; continue; }
                const double rd = v0 / r;
                const double rd3 = rd * rd * rd;
                const double rd9 = rd3 * rd3 * rd3;
                E += k * (5 * rd9 * rd3 - 6 * rd9 * rd);
            }
        }
    }
    else if(L->kind == ORC_ANGLE_HARMONIC || L->kind == ORC_ANGLE_GAUSSIAN)
    {
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 3 * t;
            const double* p = L->par + L->npar * t;
            double rji[3], rjk[3];
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[0], rji);
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[2], rjk);
            const double lsq_rji = length_sq(rji), lsq_rjk = length_sq(rjk);
            const double dot_ijk = dot(rji, rjk);
            const double reg = dot_ijk * rsqrt(lsq_rji * lsq_rjk);
            const double theta = acos(clampd(reg, -1, 1));
            if(L->kind == ORC_ANGLE_HARMONIC)
            {
                const double dv = theta - p[1];
                E += p[0] * dv * dv;
            }
            else
            {
                const double inv_sigma2 = -1. / (2 * p[1] * p[1]);
                const double dval = theta - p[2];
                E += p[0] * exp(inv_sigma2 * dval * dval);
            }
        }
    }
    else if(L->kind == ORC_DIHEDRAL_CLEMENTI)
    {
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 4 * t;
            const double* p = L->par + 3 * t;
            double r_ij[3], r_kj[3], r_lk[3], mr_lk[3], n[3], R[3], S[3];
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[0], r_ij);
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[2], r_kj);
            adjust_direction(e, e->pos + 3 * ix[2], e->pos + 3 * ix[3], r_lk);
            const double inv = 1.0 / length_sq(r_kj);
            for(int a = 0; a < 3; ++a) mr_lk[a] = -1.0 * r_lk[a];
            cross(r_kj, mr_lk, n);
            const double cij = dot(r_ij, r_kj) * inv, clk = dot(r_lk, r_kj) * inv;
            for(int a = 0; a < 3; ++a) { R[a] = r_ij[a] - cij * r_kj[a]; S[a] = r_lk[a] - clk * r_kj[a]; }
            const double Rl = length_sq(R), Sl = length_sq(S);
            const double dRS = dot(R, S) * rsqrt(Rl * Sl);
            const double phi = copysign(acos(clampd(dRS, -1.0, 1.0)), dot(r_ij, n));
            const double dphi = phi - p[2];
            const double cos1 = cos(dphi);
            const double cos3 = cos1 * (4.0 * cos1 * cos1 - 3.0);
            E += p[0] * (1.0 - cos1) + p[1] * (1.0 - cos3);
        }
    }
    else if(L->kind == ORC_BOND_3SPN2)
    {
        for(size_t t = 0; t < L->n; ++t)
        {
            double d[3];
            adjust_direction(e, e->pos + 3 * L->idx[2 * t], e->pos + 3 * L->idx[2 * t + 1], d);
            const double r = sqrt(length_sq(d));
            const double dv = r - L->par[2 * t + 1];
            const double dv2 = dv * dv;
            E += L->par[2 * t] * (dv2 + 100 * dv2 * dv2);
        }
    }
    else if(L->kind == ORC_DIHEDRAL_PGAUSSIAN)
    {
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 4 * t;
            const double* p = L->par + 3 * t;
            double r_ij[3], r_kj[3], r_lk[3], mr_lk[3], n[3], R[3], S[3];
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[0], r_ij);
            adjust_direction(e, e->pos + 3 * ix[1], e->pos + 3 * ix[2], r_kj);
            adjust_direction(e, e->pos + 3 * ix[2], e->pos + 3 * ix[3], r_lk);
            const double inv = 1.0 / length_sq(r_kj);
            for(int a = 0; a < 3; ++a) mr_lk[a] = -1.0 * r_lk[a];
            cross(r_kj, mr_lk, n);
            const double cij = dot(r_ij, r_kj) * inv, clk = dot(r_lk, r_kj) * inv;
            for(int a = 0; a < 3; ++a) { R[a] = r_ij[a] - cij * r_kj[a]; S[a] = r_lk[a] - clk * r_kj[a]; }
            const double Rl = length_sq(R), Sl = length_sq(S);
            const double dRS = dot(R, S) * rsqrt(Rl * Sl);
            const double phi = copysign(acos(clampd(dRS, -1.0, 1.0)), dot(r_ij, n));
            const double inv_sigma2 = -1. / (2 * p[1] * p[1]);
            double dval = phi - p[2];
            if(dval < -TSP_PI) { dval += TSP_TWO_PI; } else if(TSP_PI < dval) { dval -= TSP_TWO_PI; }
            E += p[0] * exp(inv_sigma2 * dval * dval);
        }
    }
    else if(L->kind == ORC_3SPN2_STACKING)
    {
        /* ThreeSPN2BaseStackingInteraction::calc_energy (:206-276) */
        const double K = 6.0, piK = TSP_PI / 6.0, alp = 3.0;
        for(size_t t = 0; t < L->n; ++t)
        {
            const uint32_t* ix = L->idx + 6 * t;
            const int bs = (int)L->par[t];
            const double eps = tsp_bs_eps(bs), r0 = TSP_BS_R0[bs];
            double Bji[3], SBi[3];
            adjust_direction(e, e->pos + 3 * ix[2], e->pos + 3 * ix[1], Bji);
            adjust_direction(e, e->pos + 3 * ix[0], e->pos + 3 * ix[1], SBi);
            const double lBji_sq = length_sq(Bji);
            const double rlBji = rsqrt(lBji_sq);
            const double lBji = lBji_sq * rlBji;
            if(!(r0 <= lBji)) { const double term = 1.0 - exp(-alp * (lBji - r0)); E += eps * term * term; }
            const double lSBi_sq = length_sq(SBi);
            const double rlSBi = rsqrt(lSBi_sq);
            const double dot_SBij = dot(SBi, Bji);
            const double cos_SBij = dot_SBij * rlBji * rlSBi;
            const double theta = acos(clampd(cos_SBij, -1, 1));
            const double f_theta = tsp_f(K, piK, theta, tsp_bs_theta0(bs));
            if(f_theta == 0.0) continue;
            double U_attr;
            if(lBji <= r0) U_attr = -eps;
            else { const double term = exp(-alp * (lBji - r0)); U_attr = eps * (term * term - 2 * term); }
            E += U_attr * f_theta;
        }
    }
    return E;
}


/* ------------------------------------------------------------------------------------- */
/* 3SPN2 base pairing and cross stacking                                                  */
/* ------------------------------------------------------------------------------------- */
static inline double tsp_U_rep(double eps, double alpha, double r, double r0)
{
    if(r0 < r) return 0.0;
    const double term = 1.0 - exp(-alpha * (r - r0));
    return eps * term * term;
}
static inline double tsp_dU_rep(double eps, double alpha, double r, double r0)
{
    if(r0 < r) return 0.0;
    const double term = exp(-alpha * (r - r0));
    return 2 * alpha * eps * term * (1.0 - term);
}
static inline void tsp_U_dU_attr(double eps, double alpha, double r, double r0, double* U, double* dU)
{
    if(r <= r0) { *U = -eps; *dU = 0.0; return; }
    const double term1 = exp(-alpha * (r - r0));
    const double term2 = 1.0 - term1;
    *U = eps * (term2 * term2 - 1.0);
    *dU = 2 * alpha * eps * term1 * term2;
}
static inline double tsp_U_attr(double eps, double alpha, double r, double r0)
{
    if(r <= r0) return -eps;
    const double term = 1.0 - exp(-alpha * (r - r0));
    return eps * (term * term - 1.0);
}

/* ThreeSPN2BasePairInteraction::calc_force_energy_virial_impl (ThreeSPN2BasePairInteraction.hpp:
 * 289-514).  Per base: ppar {base, sugar index}. */
static inline __attribute__((always_inline))
double basepair_force_impl(orc_engine* e, const global_t* G, double* F, const int need_energy)
{
    const double alpha = 2.0, K = 12.0, piK = TSP_PI / 12.0, cutoff_sq = 18.0 * 18.0;
    double energy = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        const size_t Bi = G->part[k];
        const double* rBi = e->pos + 3 * Bi;
        for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
        {
            const size_t Bj = G->nbr.a[q];
            const double* rBj = e->pos + 3 * Bj;
            double Bij[3];
            adjust_direction(e, rBi, rBj, Bij);
            const double lBij_sq = length_sq(Bij);
            if(lBij_sq > cutoff_sq) continue;
            const double rlBij = rsqrt(lBij_sq);
            const double lBij = lBij_sq * rlBij;
            double Bij_reg[3], Bji_reg[3];
            for(int a = 0; a < 3; ++a) { Bij_reg[a] = rlBij * Bij[a]; Bji_reg[a] = -rlBij * Bij[a]; }
            const int bp = tsp_bp_kind((int)G->ppar[2 * Bi]);
            const double eps = tsp_bp_eps(bp), r0 = tsp_bp_r0(bp);
            if(need_energy) energy += tsp_U_rep(eps, alpha, lBij, r0);
            const double dU_rep = tsp_dU_rep(eps, alpha, lBij, r0);
            if(dU_rep != 0.0)
            {
                double f[3];
                for(int a = 0; a < 3; ++a) f[a] = -dU_rep * Bji_reg[a];
                for(int a = 0; a < 3; ++a) { F[3 * Bi + a] += f[a]; F[3 * Bj + a] -= f[a]; }
            }
            const size_t Si = (size_t)G->ppar[2 * Bi + 1], Sj = (size_t)G->ppar[2 * Bj + 1];
            double SBi[3], SBj[3], BSi_reg[3], BSj_reg[3];
            adjust_direction(e, e->pos + 3 * Si, rBi, SBi);
            adjust_direction(e, e->pos + 3 * Sj, rBj, SBj);
            const double lSBi_sq = length_sq(SBi), lSBj_sq = length_sq(SBj);
            const double rlSBi = rsqrt(lSBi_sq), rlSBj = rsqrt(lSBj_sq);
            for(int a = 0; a < 3; ++a) { BSi_reg[a] = -rlSBi * SBi[a]; BSj_reg[a] = -rlSBj * SBj[a]; }
            const double dot_SBiBj = -dot(SBi, Bij);
            const double dot_SBjBi = dot(SBj, Bij);
            const double cos_theta1 = dot_SBiBj * rlSBi * rlBij;
            const double cos_theta2 = dot_SBjBi * rlSBj * rlBij;
            const double theta1 = acos(clampd(cos_theta1, -1, 1));
            const double theta2 = acos(clampd(cos_theta2, -1, 1));
            const double f1 = tsp_f(K, piK, theta1, tsp_bp_theta1(bp));
            const double f2 = tsp_f(K, piK, theta2, tsp_bp_theta2(bp));
            if(f1 == 0.0 || f2 == 0.0) continue;
            const double df1 = tsp_df(K, piK, theta1, tsp_bp_theta1(bp));
            const double df2 = tsp_df(K, piK, theta2, tsp_bp_theta2(bp));
            double mSBi[3], m[3], n[3];
            for(int a = 0; a < 3; ++a) mSBi[a] = -SBi[a];
            cross(mSBi, Bij, m);
            cross(Bij, SBj, n);
            const double m_lsq = length_sq(m), n_lsq = length_sq(n);
            const double dot_phi = dot(m, n) * rsqrt(m_lsq * n_lsq);
            const double cos_phi = clampd(dot_phi, -1, 1);
            const double phi = copysign(acos(cos_phi), -dot(SBi, n));
            double dphi = phi - tsp_bp_phi0(bp);
            if(dphi < -TSP_PI) { dphi += TSP_TWO_PI; }
            if(TSP_PI <= dphi) { dphi -= TSP_TWO_PI; }
            const double cos_dphi = cos(dphi);
            const double sin_dphi = sin(dphi);
            double U, dU;
            tsp_U_dU_attr(eps, alpha, lBij, r0, &U, &dU);
            if(need_energy) energy += 0.5 * (1 + cos_dphi) * f1 * f2 * U;
            if(cos_dphi == -1.0) continue;
            double f_Si[3] = {0, 0, 0}, f_Sj[3] = {0, 0, 0}, f_Bi[3] = {0, 0, 0}, f_Bj[3] = {0, 0, 0};
            if(sin_dphi != 0.0)
            {
                const double coef = 0.5 * sin_dphi * f1 * f2 * U;
                const double rlBij_sq = rlBij * rlBij;
                double fSi[3], fSj[3];
                for(int a = 0; a < 3; ++a)
                {
                    fSi[a] = (coef * lBij / m_lsq) * m[a];
                    fSj[a] = (-coef * lBij / n_lsq) * n[a];
                }
                const double coef_Bi = dot_SBiBj * rlBij_sq;
                const double coef_Bj = dot_SBjBi * rlBij_sq;
                for(int a = 0; a < 3; ++a)
                {
                    f_Si[a] += fSi[a];
                    f_Bi[a] += (coef_Bi - 1.0) * fSi[a] - coef_Bj * fSj[a];
                    f_Bj[a] += (coef_Bj - 1.0) * fSj[a] - coef_Bi * fSi[a];
                    f_Sj[a] += fSj[a];
                }
            }
            const double dihd_term = 0.5 * (1.0 + cos_dphi);
            if(df1 != 0.0)
            {
                const double coef = -dihd_term * df1 * f2 * U;
                const double sin_theta1 = sin(theta1);
                const double coef_rsin = (sin_theta1 > TSP_TOL) ? (coef / sin_theta1) : (coef / TSP_TOL);
                double fSi[3], fBj[3];
                for(int a = 0; a < 3; ++a)
                {
                    fSi[a] = (coef_rsin * rlSBi) * (cos_theta1 * BSi_reg[a] - Bij_reg[a]);
                    fBj[a] = (coef_rsin * rlBij) * (cos_theta1 * Bij_reg[a] - BSi_reg[a]);
                }
                for(int a = 0; a < 3; ++a) { f_Si[a] += fSi[a]; f_Bi[a] -= (fSi[a] + fBj[a]); f_Bj[a] += fBj[a]; }
            }
            if(df2 != 0.0)
            {
                const double coef = -dihd_term * f1 * df2 * U;
                const double sin_theta2 = sin(theta2);
                const double coef_rsin = (sin_theta2 > TSP_TOL) ? (coef / sin_theta2) : (coef / TSP_TOL);
                double fBi[3], fSj[3];
                for(int a = 0; a < 3; ++a)
                {
                    fBi[a] = (coef_rsin * rlBij) * (cos_theta2 * Bji_reg[a] - BSj_reg[a]);
                    fSj[a] = (coef_rsin * rlSBj) * (cos_theta2 * BSj_reg[a] - Bji_reg[a]);
                }
                for(int a = 0; a < 3; ++a) { f_Bi[a] += fBi[a]; f_Bj[a] -= (fBi[a] + fSj[a]); f_Sj[a] += fSj[a]; }
            }
            if(dU != 0.0)
            {
                const double coef = -dihd_term * f1 * f2 * dU;
                for(int a = 0; a < 3; ++a) { f_Bi[a] += coef * Bji_reg[a]; f_Bj[a] += coef * Bij_reg[a]; }
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * Si + a] += f_Si[a];
                F[3 * Bi + a] += f_Bi[a];
                F[3 * Bj + a] += f_Bj[a];
                F[3 * Sj + a] += f_Sj[a];
            }
        }
    }
    return energy;
}

/* ThreeSPN2BasePairInteraction::calc_energy (ThreeSPN2BasePairInteraction.hpp:133-284) */
static double basepair_energy(orc_engine* e, const global_t* G)
{
    const double alpha = 2.0, K = 12.0, piK = TSP_PI / 12.0, cutoff_sq = 18.0 * 18.0;
    double E = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        const size_t Bi = G->part[k];
        const double* rBi = e->pos + 3 * Bi;
        for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
        {
            const size_t Bj = G->nbr.a[q];
            const double* rBj = e->pos + 3 * Bj;
            const int bp = tsp_bp_kind((int)G->ppar[2 * Bi]);
            double Bij[3];
            adjust_direction(e, rBi, rBj, Bij);
            const double lBij_sq = length_sq(Bij);
            if(lBij_sq > cutoff_sq) continue;
            const double rlBij = rsqrt(lBij_sq);
            const double lBij = lBij_sq * rlBij;
            const double eps = tsp_bp_eps(bp), r0 = tsp_bp_r0(bp);
            E += tsp_U_rep(eps, alpha, lBij, r0);
            const size_t Si = (size_t)G->ppar[2 * Bi + 1], Sj = (size_t)G->ppar[2 * Bj + 1];
            double SBi[3], SBj[3];
            adjust_direction(e, e->pos + 3 * Si, rBi, SBi);
            adjust_direction(e, e->pos + 3 * Sj, rBj, SBj);
            const double lSBi_sq = length_sq(SBi), lSBj_sq = length_sq(SBj);
            const double rlSBi = rsqrt(lSBi_sq), rlSBj = rsqrt(lSBj_sq);
            const double dot_SBiBj = -dot(SBi, Bij);
            const double dot_SBjBi = dot(SBj, Bij);
            const double cos_theta1 = dot_SBiBj * rlSBi * rlBij;
            const double cos_theta2 = dot_SBjBi * rlSBj * rlBij;
            const double theta1 = acos(clampd(cos_theta1, -1, 1));
            const double theta2 = acos(clampd(cos_theta2, -1, 1));
            const double f1 = tsp_f(K, piK, theta1, tsp_bp_theta1(bp));
            const double f2 = tsp_f(K, piK, theta2, tsp_bp_theta2(bp));
            if(f1 == 0.0 || f2 == 0.0) continue;
            double Bij_reg[3], R[3], S[3], n[3];
            for(int a = 0; a < 3; ++a) Bij_reg[a] = Bij[a] * rlBij;
            const double dR = dot(SBi, Bij_reg), dS = dot(SBj, Bij_reg);
            for(int a = 0; a < 3; ++a) { R[a] = dR * Bij_reg[a] - SBi[a]; S[a] = dS * Bij_reg[a] - SBj[a]; }
            const double dot_RS = dot(R, S) * rsqrt(length_sq(R) * length_sq(S));
            const double cos_phi = clampd(dot_RS, -1, 1);
            cross(Bij, SBj, n);
            const double sign = -dot(SBi, n);
            const double phi = copysign(acos(cos_phi), sign);
            double dphi = phi - tsp_bp_phi0(bp);
            if(TSP_PI < dphi) { dphi -= TSP_TWO_PI; }
            else if(dphi < -TSP_PI) { dphi += TSP_TWO_PI; }
            const double cos_dphi = cos(dphi);
            const double U = tsp_U_attr(eps, alpha, lBij, r0);
            E += 0.5 * (1 + cos_dphi) * f1 * f2 * U;
        }
    }
    return E;
}

/* one cross-stacking element of ThreeSPN2CrossStackingInteraction::calc_force_energy_virial_impl
 * (ThreeSPN2CrossStackingInteraction.hpp:445-526) */
static inline __attribute__((always_inline))
double cs_force_element(orc_engine* e, int cs, double f3, double df3, const double* rB1, const double* rBn,
                        const double* SB1, double rlSB1, const double* BS1_reg,
                        const double* fS1_t3, const double* fB1_t3, const double* fS2_t3, const double* fB2_t3,
                        double* f_S1, double* f_B1, double* f_S2, double* f_B2, double* f_Bn, const int need_energy)
{
    const double K_CS = 8.0, piK_CS = TSP_PI / 8.0, alpha = 4.0;
    const double eps = tsp_cs_eps(cs), r0 = TSP_CS_R0[cs];
    double Bn1[3], Bn1_reg[3];
    adjust_direction(e, rBn, rB1, Bn1);
    const double lBn1_sq = length_sq(Bn1);
    const double rlBn1 = rsqrt(lBn1_sq);
    const double dot_theta_CS = dot(SB1, Bn1);
    const double cos_theta_CS = dot_theta_CS * rlSB1 * rlBn1;
    const double theta_CS = acos(clampd(cos_theta_CS, -1, 1));
    const double th0 = tsp_cs_theta0(cs);
    const double fCS = tsp_f(K_CS, piK_CS, theta_CS, th0);
    const double dfCS = tsp_df(K_CS, piK_CS, theta_CS, th0);
    if(fCS == 0.0) return 0.0;
    const double lBn1 = lBn1_sq * rlBn1;
    double U, dU;
    tsp_U_dU_attr(eps, alpha, lBn1, r0, &U, &dU);
    for(int a = 0; a < 3; ++a) Bn1_reg[a] = rlBn1 * Bn1[a];
    if(df3 != 0.0)
    {
        const double coef = -0.5 * df3 * fCS * U;
        for(int a = 0; a < 3; ++a)
        {
            f_S1[a] += coef * fS1_t3[a];
            f_S2[a] += coef * fS2_t3[a];
            f_B1[a] += coef * fB1_t3[a];
            f_B2[a] += coef * fB2_t3[a];
        }
    }
    if(dfCS != 0.0)
    {
        const double coef = -0.5 * f3 * dfCS * U;
        const double sin_theta_CS = sin(theta_CS);
        const double coef_rsin = (sin_theta_CS > TSP_TOL) ? (coef / sin_theta_CS) : (coef / TSP_TOL);
        double fS1[3], fBn[3];
        for(int a = 0; a < 3; ++a)
        {
            fS1[a] = coef_rsin * rlSB1 * (cos_theta_CS * BS1_reg[a] + Bn1_reg[a]);
            fBn[a] = -coef_rsin * rlBn1 * (cos_theta_CS * Bn1_reg[a] + BS1_reg[a]);
        }
        for(int a = 0; a < 3; ++a) { f_S1[a] += fS1[a]; f_B1[a] -= (fS1[a] + fBn[a]); f_Bn[a] += fBn[a]; }
    }
    if(dU != 0.0)
    {
        const double coef = -0.5 * f3 * fCS * dU;
        for(int a = 0; a < 3; ++a) { f_B1[a] += coef * Bn1_reg[a]; f_Bn[a] -= coef * Bn1_reg[a]; }
    }
    return need_energy ? 0.5 * f3 * fCS * U : 0.0;
}

/* ThreeSPN2CrossStackingInteraction::calc_force_energy_virial_impl (:340-640).  Per base:
 * ppar {base, sugar index, 5' base (-1: none), 3' base (-1: none)}; cross-stacking kinds
 * cs5 = b1 << 2 | b2, cs3 = (b1 << 2 | b2) + 16 (ThreeSPN2CrossStackingPotential.hpp:290-307) */
static inline __attribute__((always_inline))
double crossstack_force_impl(orc_engine* e, const global_t* G, double* F, const int need_energy)
{
    const double K_BP = 12.0, piK_BP = TSP_PI / 12.0, cutoff_sq = 18.0 * 18.0;
    double energy = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        const size_t Bi = G->part[k];
        const double* rBi = e->pos + 3 * Bi;
        const double* pi_ = G->ppar + 4 * Bi;
        const long Bi5 = (long)pi_[2], Bi3 = (long)pi_[3];
        const size_t Si = (size_t)pi_[1];
        const int Bi_base = (int)pi_[0];
        const int Bi5_exists = Bi5 >= 0, Bi3_exists = Bi3 >= 0;
        for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
        {
            const size_t Bj = G->nbr.a[q];
            const double* rBj = e->pos + 3 * Bj;
            const int bp = tsp_bp_kind(Bi_base);
            double Bij[3];
            adjust_direction(e, rBi, rBj, Bij);
            const double lBij_sq = length_sq(Bij);
            if(lBij_sq > cutoff_sq) continue;
            const double* pj = G->ppar + 4 * Bj;
            const long Bj5 = (long)pj[2], Bj3 = (long)pj[3];
            const size_t Sj = (size_t)pj[1];
            const int Bj5_exists = Bj5 >= 0, Bj3_exists = Bj3 >= 0;
            if(!Bi3_exists && !Bi5_exists && !Bj3_exists && !Bj5_exists) continue;
            const int Bj_base = (int)pj[0];
            double SBi[3], SBj[3], BSi_reg[3], BSj_reg[3];
            adjust_direction(e, e->pos + 3 * Si, rBi, SBi);
            adjust_direction(e, e->pos + 3 * Sj, rBj, SBj);
            const double lSBi_sq = length_sq(SBi), lSBj_sq = length_sq(SBj);
            const double rlSBi = rsqrt(lSBi_sq), rlSBj = rsqrt(lSBj_sq);
            for(int a = 0; a < 3; ++a) { BSi_reg[a] = -rlSBi * SBi[a]; BSj_reg[a] = -rlSBj * SBj[a]; }
            const double cos_theta3 = dot(BSi_reg, BSj_reg);
            const double theta3 = acos(clampd(cos_theta3, -1, 1));
            const double th3_0 = tsp_cs_theta3(bp);
            const double f3 = tsp_f(K_BP, piK_BP, theta3, th3_0);
            if(f3 == 0.0) continue;
            const double df3 = tsp_df(K_BP, piK_BP, theta3, th3_0);
            const double sin_theta3 = sin(theta3);
            const double rsin_theta3 = sin_theta3 > TSP_TOL ? 1.0 / sin_theta3 : 1.0 / TSP_TOL;
            double fBi_t3[3], fBj_t3[3], fSi_t3[3], fSj_t3[3];
            for(int a = 0; a < 3; ++a)
            {
                fBi_t3[a] = rsin_theta3 * rlSBi * (BSj_reg[a] - cos_theta3 * BSi_reg[a]);
                fBj_t3[a] = rsin_theta3 * rlSBj * (BSi_reg[a] - cos_theta3 * BSj_reg[a]);
            }
            for(int a = 0; a < 3; ++a) { fSi_t3[a] = -1.0 * fBi_t3[a]; fSj_t3[a] = -1.0 * fBj_t3[a]; }
            double f_Si[3] = {0, 0, 0}, f_Sj[3] = {0, 0, 0}, f_Bi[3] = {0, 0, 0}, f_Bi3[3] = {0, 0, 0};
            double f_Bi5[3] = {0, 0, 0}, f_Bj[3] = {0, 0, 0}, f_Bj3[3] = {0, 0, 0}, f_Bj5[3] = {0, 0, 0};
            if(Bj5_exists)
            {
                const int cs = (Bi_base << 2 | (int)G->ppar[4 * Bj5]);
                energy += cs_force_element(e, cs, f3, df3, rBi, e->pos + 3 * Bj5, SBi, rlSBi, BSi_reg,
                                           fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj5, need_energy);
            }
            if(Bi5_exists)
            {
                const int cs = (Bj_base << 2 | (int)G->ppar[4 * Bi5]);
                energy += cs_force_element(e, cs, f3, df3, rBj, e->pos + 3 * Bi5, SBj, rlSBj, BSj_reg,
                                           fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi5, need_energy);
            }
            if(Bi3_exists)
            {
                const int cs = (Bj_base << 2 | (int)G->ppar[4 * Bi3]) + 16;
                energy += cs_force_element(e, cs, f3, df3, rBj, e->pos + 3 * Bi3, SBj, rlSBj, BSj_reg,
                                           fSj_t3, fBj_t3, fSi_t3, fBi_t3, f_Sj, f_Bj, f_Si, f_Bi, f_Bi3, need_energy);
            }
            if(Bj3_exists)
            {
                const int cs = (Bi_base << 2 | (int)G->ppar[4 * Bj3]) + 16;
                energy += cs_force_element(e, cs, f3, df3, rBi, e->pos + 3 * Bj3, SBi, rlSBi, BSi_reg,
                                           fSi_t3, fBi_t3, fSj_t3, fBj_t3, f_Si, f_Bi, f_Sj, f_Bj, f_Bj3, need_energy);
            }
            for(int a = 0; a < 3; ++a)
            {
                F[3 * Si + a] += f_Si[a];
                F[3 * Sj + a] += f_Sj[a];
                F[3 * Bi + a] += f_Bi[a];
                F[3 * Bj + a] += f_Bj[a];
            }
            if(Bi3_exists) for(int a = 0; a < 3; ++a) F[3 * Bi3 + a] += f_Bi3[a];
            if(Bi5_exists) for(int a = 0; a < 3; ++a) F[3 * Bi5 + a] += f_Bi5[a];
            if(Bj3_exists) for(int a = 0; a < 3; ++a) F[3 * Bj3 + a] += f_Bj3[a];
            if(Bj5_exists) for(int a = 0; a < 3; ++a) F[3 * Bj5 + a] += f_Bj5[a];
        }
    }
    return energy;
}

static double cs_energy_element(orc_engine* e, int cs, double f3, const double* rB1, const double* rB2,
                                const double* SB1, double rlSB1)
{
    const double K_CS = 8.0, piK_CS = TSP_PI / 8.0, alpha = 4.0;
    const double eps = tsp_cs_eps(cs), r0 = TSP_CS_R0[cs];
    double B21[3];
    adjust_direction(e, rB2, rB1, B21);
    const double lB21_sq = length_sq(B21);
    const double rlB21 = rsqrt(lB21_sq);
    const double cos_theta_CS = dot(SB1, B21) * rlSB1 * rlB21;
    const double theta_CS = acos(clampd(cos_theta_CS, -1, 1));
    const double fCS = tsp_f(K_CS, piK_CS, theta_CS, tsp_cs_theta0(cs));
    if(fCS == 0.0) return 0.0;
    const double lB21 = lB21_sq * rlB21;
    const double U = tsp_U_attr(eps, alpha, lB21, r0);
    return 0.5 * f3 * fCS * U;
}

/* ThreeSPN2CrossStackingInteraction::calc_energy (:133-337) */
static double crossstack_energy(orc_engine* e, const global_t* G)
{
    const double K_BP = 12.0, piK_BP = TSP_PI / 12.0, cutoff_sq = 18.0 * 18.0;
    double E = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        const size_t Bi = G->part[k];
        const double* rBi = e->pos + 3 * Bi;
        const double* pi_ = G->ppar + 4 * Bi;
        const long Bi5 = (long)pi_[2], Bi3 = (long)pi_[3];
        const size_t Si = (size_t)pi_[1];
        const int Bi_base = (int)pi_[0];
        const int Bi5_exists = Bi5 >= 0, Bi3_exists = Bi3 >= 0;
        for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
        {
            const size_t Bj = G->nbr.a[q];
            const double* rBj = e->pos + 3 * Bj;
            const int bp = tsp_bp_kind(Bi_base);
            const double* pj = G->ppar + 4 * Bj;
            const int Bj_base = (int)pj[0];
            double Bij[3];
            adjust_direction(e, rBi, rBj, Bij);
            const double lBij_sq = length_sq(Bij);
            if(lBij_sq > cutoff_sq) continue;
            const long Bj5 = (long)pj[2], Bj3 = (long)pj[3];
            const size_t Sj = (size_t)pj[1];
            const int Bj5_exists = Bj5 >= 0, Bj3_exists = Bj3 >= 0;
            if(!Bi3_exists && !Bi5_exists && !Bj3_exists && !Bj5_exists) continue;
            double SBi[3], SBj[3];
            adjust_direction(e, e->pos + 3 * Si, rBi, SBi);
            adjust_direction(e, e->pos + 3 * Sj, rBj, SBj);
            const double lSBi_sq = length_sq(SBi), lSBj_sq = length_sq(SBj);
            const double rlSBi = rsqrt(lSBi_sq), rlSBj = rsqrt(lSBj_sq);
            const double dot_SBiSBj = dot(SBi, SBj);
            const double cos_theta3 = dot_SBiSBj * rlSBi * rlSBj;
            const double theta3 = acos(clampd(cos_theta3, -1, 1));
            const double f3 = tsp_f(K_BP, piK_BP, theta3, tsp_cs_theta3(bp));
            if(f3 == 0.0) continue;
            if(Bj5_exists)
                E += cs_energy_element(e, (Bi_base << 2 | (int)G->ppar[4 * Bj5]), f3, rBi, e->pos + 3 * Bj5, SBi, rlSBi);
            if(Bi5_exists)
                E += cs_energy_element(e, (Bj_base << 2 | (int)G->ppar[4 * Bi5]), f3, rBj, e->pos + 3 * Bi5, SBj, rlSBj);
            if(Bi3_exists)
                E += cs_energy_element(e, (Bj_base << 2 | (int)G->ppar[4 * Bi3]) + 16, f3, rBj, e->pos + 3 * Bi3, SBj, rlSBj);
            if(Bj3_exists)
                E += cs_energy_element(e, (Bi_base << 2 | (int)G->ppar[4 * Bj3]) + 16, f3, rBi, e->pos + 3 * Bj3, SBi, rlSBi);
        }
    }
    return E;
}

static inline __attribute__((always_inline))
double global_force_impl(orc_engine* e, const global_t* G, double* F, const int need_energy)
{
    double energy = 0;
    const size_t nlead = G->np - 1;
    if(G->pot == ORC_3SPN2_BASEPAIR) return basepair_force_impl(e, G, F, need_energy);
    if(G->pot == ORC_3SPN2_CROSSSTACK) return crossstack_force_impl(e, G, F, need_energy);
    if(G->pot == ORC_PAIR_3SPN2_EV)
    {
        /* GlobalPairInteraction.hpp:148-189 with ThreeSPN2ExcludedVolumePotential.hpp:47-64 */
        const double eps = G->eps;
        for(size_t k = 0; k < nlead; ++k)
        {
            const size_t i = G->part[k];
            for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
            {
                const size_t j = G->nbr.a[q];
                const double sigma = G->npp[q];
                double d[3];
                adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
                const double l2 = length_sq(d);
                const double rl = rsqrt(l2);
                const double l = l2 * rl;
                double f_mag = 0.0;
                if(!(sigma <= l))
                {
                    const double rinv = 1 / l;
                    const double r1s1 = sigma * rinv;
                    const double r3s3 = r1s1 * r1s1 * r1s1;
                    const double r6s6 = r3s3 * r3s3;
                    f_mag = 12.0 * eps * rinv * r6s6 * (1.0 - r6s6);
                }
                if(f_mag == 0.0) continue;
                const double c = f_mag * rl;
                for(int a = 0; a < 3; ++a) { F[3 * i + a] += d[a] * c; F[3 * j + a] -= d[a] * c; }
                if(need_energy)
                {
                    if(!(sigma <= l))
                    {
                        const double r1s1 = sigma / l;
                        const double r3s3 = r1s1 * r1s1 * r1s1;
                        const double r6s6 = r3s3 * r3s3;
                        energy += eps * (r6s6 * (r6s6 - 2.0) + 1.0);
                    }
                }
            }
        }
        return energy;
    }
    if(G->pot == ORC_PAIR_LJ)
    {
        /* GlobalPairLennardJonesInteraction.hpp:148-198 */
        const double coef = G->coef_at_cutoff, cr = G->cutoff_ratio, cr2 = cr * cr;
        for(size_t k = 0; k < nlead; ++k)
        {
            const size_t i = G->part[k];
            for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
            {
                const size_t j = G->nbr.a[q];
                const double sigma = G->npp[2 * q], eps = G->npp[2 * q + 1];
                double d[3];
                adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
                const double l_sq = length_sq(d);
                const double sigma_sq = sigma * sigma;
                if(sigma_sq * cr2 < l_sq) continue;
                const double rcp = 1 / l_sq;
                const double s2l2 = sigma_sq * rcp;
                const double s6l6 = s2l2 * s2l2 * s2l2;
                const double c = 24 * eps * (s6l6 - 2 * s6l6 * s6l6) * rcp;
                for(int a = 0; a < 3; ++a) { F[3 * i + a] += d[a] * c; F[3 * j + a] -= d[a] * c; }
                if(need_energy) energy += 4 * eps * (s6l6 * s6l6 - s6l6 - coef);
            }
        }
    }
    else if(G->pot == ORC_PAIR_EV)
    {
        /* GlobalPairExcludedVolumeInteraction.hpp:148-195 */
        const double coef = G->coef_at_cutoff, cr = G->cutoff_ratio, cr2 = cr * cr;
        const double eps = G->eps, eps12 = eps * 12;
        for(size_t k = 0; k < nlead; ++k)
        {
            const size_t i = G->part[k];
            for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
            {
                const size_t j = G->nbr.a[q];
                const double radius = G->npp[q];
                double d[3];
                adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
                const double l_sq = length_sq(d);
                const double sigma_sq = radius * radius;
                if(sigma_sq * cr2 < l_sq) continue;
                const double rcp = 1.0 / l_sq;
                const double s2l2 = sigma_sq * rcp;
                const double s6l6 = s2l2 * s2l2 * s2l2;
                const double c = -eps12 * s6l6 * s6l6 * rcp;
                for(int a = 0; a < 3; ++a) { F[3 * i + a] += d[a] * c; F[3 * j + a] -= d[a] * c; }
                if(need_energy) energy += eps * (s6l6 * s6l6 - coef);
            }
        }
    }
    else
    {
        /* GlobalPairInteraction.hpp:148-189 with DebyeHuckelPotential.hpp:47-65 */
        for(size_t k = 0; k < nlead; ++k)
        {
            const size_t i = G->part[k];
            for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
            {
                const size_t j = G->nbr.a[q];
                const double qq = G->npp[q];
                double d[3];
                adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
                const double l2 = length_sq(d);
                const double rl = rsqrt(l2);
                const double l = l2 * rl;
                double f_mag = 0.0;
                if(!(G->abs_cutoff <= l))
                    f_mag = -qq * (G->debye + l) * G->inv_debye / (l * l) * exp(-l * G->inv_debye);
                if(f_mag == 0.0) continue;
                const double c = f_mag * rl;
                for(int a = 0; a < 3; ++a) { F[3 * i + a] += d[a] * c; F[3 * j + a] -= d[a] * c; }
                if(need_energy)
                    energy += (G->abs_cutoff <= l) ? 0.0 : qq * (exp(-l * G->inv_debye) / l - G->coef_at_cutoff);
            }
        }
    }
    return energy;
}

static double global_force_f(orc_engine* e, const global_t* G, double* F) { return global_force_impl(e, G, F, 0); }
static double global_force_fe(orc_engine* e, const global_t* G, double* F) { return global_force_impl(e, G, F, 1); }
static double global_force(orc_engine* e, const global_t* G, double* F, int need_energy)
{
    return need_energy ? global_force_fe(e, G, F) : global_force_f(e, G, F);
}

/* calc_energy of the pair interactions: GlobalPairLennardJonesInteraction.hpp:98-132,
 * GlobalPairExcludedVolumeInteraction.hpp:100-132, GlobalPairInteraction.hpp:191-213 */
static double global_energy(orc_engine* e, const global_t* G)
{
    if(G->pot == ORC_3SPN2_BASEPAIR) return basepair_energy(e, G);
    if(G->pot == ORC_3SPN2_CROSSSTACK) return crossstack_energy(e, G);
    double E = 0;
    const size_t nlead = G->np - 1;
    for(size_t k = 0; k < nlead; ++k)
    {
        const size_t i = G->part[k];
        for(size_t q = G->rng_start[k]; q < G->rng_start[k + 1]; ++q)
        {
            const size_t j = G->nbr.a[q];
            double d[3];
            adjust_direction(e, e->pos + 3 * i, e->pos + 3 * j, d);
            if(G->pot == ORC_PAIR_LJ)
            {
                const double sigma = G->npp[2 * q], eps = G->npp[2 * q + 1];
                const double cr2 = G->cutoff_ratio * G->cutoff_ratio;
                const double l_sq = length_sq(d);
                const double sigma_sq = sigma * sigma;
                if(sigma_sq * cr2 < l_sq) continue;
                const double s2l2 = sigma_sq / l_sq;
                const double s6l6 = s2l2 * s2l2 * s2l2;
                E += 4 * eps * (s6l6 * s6l6 - s6l6 - G->coef_at_cutoff);
            }
            else if(G->pot == ORC_PAIR_EV)
            {
                const double cr2 = G->cutoff_ratio * G->cutoff_ratio;
                const double l_sq = length_sq(d);
                const double sigma_sq = G->npp[q] * G->npp[q];
                if(sigma_sq * cr2 < l_sq) continue;
                const double s2l2 = sigma_sq / l_sq;
                const double s6l6 = s2l2 * s2l2 * s2l2;
                E += G->eps * (s6l6 * s6l6 - G->coef_at_cutoff);
            }
            else if(G->pot == ORC_PAIR_3SPN2_EV)
            {
                const double l = sqrt(length_sq(d));
                const double sigma = G->npp[q];
                if(sigma <= l) continue;
                const double r1s1 = sigma / l;
                const double r3s3 = r1s1 * r1s1 * r1s1;
                const double r6s6 = r3s3 * r3s3;
                E += G->eps * (r6s6 * (r6s6 - 2.0) + 1.0);
            }
            else
            {
                const double l = sqrt(length_sq(d));
                E += (G->abs_cutoff <= l) ? 0.0 : G->npp[q] * (exp(-l * G->inv_debye) / l - G->coef_at_cutoff);
            }
        }
    }
    return E;
}

double orc_calc_energy_of(orc_engine* e, int id)
{
    if(id < 0) return NAN;
    if((size_t)id < e->n_loc) return local_energy(e, &e->loc[id]);
    if((size_t)id < e->n_loc + e->n_glo) return global_energy(e, &e->glo[id - e->n_loc]);
    return NAN;
}

int orc_calc_force_of(orc_engine* e, int id, double* force_out)
{
    if(id < 0) return -1;
    if((size_t)id < e->n_loc) { local_force(e, &e->loc[id], force_out, 0); return 0; }
    if((size_t)id < e->n_loc + e->n_glo) { global_force(e, &e->glo[id - e->n_loc], force_out, 0); return 0; }
    return -1;
}

/* ForceField::calc_force / calc_force_and_energy (ForceField.hpp:101-135) */
static double calc_force_impl(orc_engine* e, int need_energy)
{
    double E = 0;
    for(size_t l = 0; l < e->n_loc; ++l) E += local_force(e, &e->loc[l], e->force, need_energy);
    for(size_t g = 0; g < e->n_glo; ++g) E += global_force(e, &e->glo[g], e->force, need_energy);
    return E;
}
double orc_calc_force_and_energy(orc_engine* e)
{
    memset(e->force, 0, 3 * e->n * sizeof(double));
    return calc_force_impl(e, 1);
}

/* ForceField::reduce_margin (ForceField.hpp:75-92): local Contact lists and every global
 * partition (PeriodicGridCellList.hpp:100-110, UnlimitedGridCellList.hpp:71-81,
 * VerletList.hpp:55-65; Naive never rebuilds) */
int orc_reduce_margin(orc_engine* e, double dmargin)
{
    int rebuilt = 0;
    for(size_t l = 0; l < e->n_loc; ++l)
    {
        local_t* L = &e->loc[l];
        if(L->kind != ORC_CONTACT_GOCONTACT || L->n == 0) continue;
        L->current_margin -= dmargin;
        if(L->current_margin < 0) contact_make_list(e, L);
    }
    for(size_t g = 0; g < e->n_glo; ++g)
    {
        global_t* G = &e->glo[g];
        if(G->partition == ORC_PART_NAIVE) continue;
        G->current_margin -= dmargin;
        if(G->current_margin < 0) { make_list(e, G); rebuilt++; }
    }
    return rebuilt;
}

/* ------------------------------------------------------------------------------------- */
/* integrators                                                                             */
/* ------------------------------------------------------------------------------------- */
int orc_set_integrator(orc_engine* e, int type, double dt, const double* coefs, uint64_t seed)
{
    if(type != ORC_INTEG_BAOAB && type != ORC_INTEG_GJF && type != ORC_INTEG_GBAOAB &&
       type != ORC_INTEG_UNDERDAMPED) return -1;
    e->integ = type; e->dt = dt; e->halfdt = dt / 2; e->seed = seed; e->step = 0;
    free(e->coef); free(e->c1); free(e->c2); free(e->c3);
    e->coef = (double*)malloc(e->n * sizeof(double) + 1);
    e->c1 = (double*)malloc(e->n * sizeof(double) + 1);
    e->c2 = (double*)malloc(e->n * sizeof(double) + 1);
    e->c3 = (double*)malloc(e->n * sizeof(double) + 1);
    memcpy(e->coef, coefs, e->n * sizeof(double));
    return 0;
}

/* UnderdampedLangevinIntegrator (:459-542): a = f/m + noise (counter `ctr`), and with `kick`
 * the closing half-kick v += (dt/2)(1 - gamma dt/2) a */
static void ul_accel(orc_engine* e, uint64_t ctr, int kick)
{
    const double h = e->halfdt;
    for(size_t i = 0; i < e->n; ++i)
    {
        double R[3];
        orc_noise3(e->seed, ctr, (uint32_t)i, R);
        const double rm = e->rmass[i], nc = e->c2[i];
        double* a = e->accel + 3 * i;
        for(int c = 0; c < 3; ++c) a[c] = e->force[3 * i + c] * rm + nc * R[c];
        if(kick)
        {
            const double k = h * (1 - e->coef[i] * h);
            for(int c = 0; c < 3; ++c) e->vel[3 * i + c] += k * a[c];
        }
    }
}

/* BAOABLangevinIntegrator::initialize/reset_parameters (:73-99,136-168) and
 * GJFNVTLangevinIntegrator::initialize/reset_parameters (:44-75,106-134) */
int orc_integrator_initialize(orc_engine* e)
{
    if(!e->has_T) return -1;
    const double kBT = orc_kB() * e->temperature;
    for(size_t i = 0; i < e->n; ++i)
    {
        if(e->integ == ORC_INTEG_BAOAB || e->integ == ORC_INTEG_GBAOAB)
        {
            /* gBAOABLangevinIntegrator::reset_parameter (:82-107) is BAOAB's */
            const double gamma_dt = -1 * e->coef[i] * e->dt;
            e->c1[i] = exp(gamma_dt);
            e->c2[i] = sqrt(kBT * (1 - exp(2 * gamma_dt)) * e->rmass[i]);
        }
        else if(e->integ == ORC_INTEG_UNDERDAMPED)
        {
            /* UnderdampedLangevinIntegrator::update (:320-350): noise_coef sqrt(gamma / m) */
            const double noise_coef = sqrt(2 * orc_kB() * e->temperature / e->dt);
            e->c1[i] = 0.0;
            e->c2[i] = noise_coef * sqrt(e->coef[i] * e->rmass[i]);
        }
        else
        {
            const double alpha = e->coef[i], m = e->mass[i];
            e->c1[i] = 1.0 / (1.0 + alpha * e->dt * 0.5 / m);        /* b */
            e->c2[i] = sqrt(2 * alpha * kBT * e->dt);                  /* beta */
            e->c3[i] = 1.0 - alpha * e->c1[i] * e->dt / m;             /* a */
        }
    }
    memset(e->force, 0, 3 * e->n * sizeof(double));
    calc_force_impl(e, 0);
    if(e->integ == ORC_INTEG_UNDERDAMPED)
    {
        /* UnderdampedLangevinIntegrator::initialize (:404-457): a = f/m + noise, the noise of
         * the current step count */
        free(e->accel);
        e->accel = (double*)malloc(3 * e->n * sizeof(double) + 1);
        ul_accel(e, e->step, 0);
    }
    if(e->integ == ORC_INTEG_GBAOAB)
    {
        /* gBAOABLangevinIntegrator::initialize (:290-345), one correction step per A step */
        const size_t m = e->n_con;
        free(e->con_sq_v0); free(e->con_r_sq_v0); free(e->con_redm); free(e->old_pos); free(e->old_rattle);
        e->con_sq_v0 = (double*)malloc(m * sizeof(double) + 1);
        e->con_r_sq_v0 = (double*)malloc(m * sizeof(double) + 1);
        e->con_redm = (double*)malloc(m * sizeof(double) + 1);
        for(size_t k = 0; k < m; ++k)
        {
            e->con_sq_v0[k] = pow(e->con_v0[k], 2);
            e->con_r_sq_v0[k] = 1.0 / e->con_sq_v0[k];
            e->con_redm[k] = 1.0 / (e->rmass[e->con_idx[2 * k]] + e->rmass[e->con_idx[2 * k + 1]]);
        }
        if(e->con_idx == NULL) e->con_tol = INFINITY;   /* ConstraintForceField() default */
        e->dt_corr = e->dt * 0.5 / 1;
        e->r_dt_corr = 1. / e->dt_corr;
        e->old_pos = (double*)malloc(3 * e->n * sizeof(double) + 1);
        e->old_rattle = (double*)malloc(3 * e->n * sizeof(double) + 1);
        memcpy(e->old_pos, e->pos, 3 * e->n * sizeof(double));
    }
    return 0;
}

/* gBAOABLangevinIntegrator::correct_velocity (:177-225): Gauss-Seidel sweeps over the
 * constraints in input order until no correction exceeds the tolerance */
static void gb_correct_velocity(orc_engine* e)
{
    for(size_t it = 0; it < e->con_maxit; ++it)
    {
        int corrected = 0;
        for(size_t k = 0; k < e->n_con; ++k)
        {
            const uint32_t a = e->con_idx[2 * k], b = e->con_idx[2 * k + 1];
            double* v1 = e->vel + 3 * a; double* v2 = e->vel + 3 * b;
            double pd[3], vd[3];
            adjust_direction(e, e->pos + 3 * a, e->pos + 3 * b, pd);
            for(int c = 0; c < 3; ++c) vd[c] = v2[c] - v1[c];
            const double lambda = dot(pd, vd) * e->con_redm[k] * e->con_r_sq_v0[k];
            if(e->con_tol < fabs(lambda))
            {
                const double rm1 = e->rmass[a], rm2 = e->rmass[b];
                double cv[3];
                for(int c = 0; c < 3; ++c) cv[c] = lambda * pd[c];
                for(int c = 0; c < 3; ++c) { v1[c] += cv[c] * rm1; v2[c] -= cv[c] * rm2; }
                corrected = 1;
            }
        }
        if(!corrected) break;
    }
}

/* gBAOABLangevinIntegrator::correct_coordinate (:115-175) */
static void gb_correct_coordinate(orc_engine* e)
{
    for(size_t it = 0; it < e->con_maxit; ++it)
    {
        int corrected = 0;
        for(size_t k = 0; k < e->n_con; ++k)
        {
            const uint32_t a = e->con_idx[2 * k], b = e->con_idx[2 * k + 1];
            double* p1 = e->pos + 3 * a; double* p2 = e->pos + 3 * b;
            double* v1 = e->vel + 3 * a; double* v2 = e->vel + 3 * b;
            double dp[3];
            adjust_direction(e, p1, p2, dp);
            const double dp2 = length_sq(dp);
            const double miss = e->con_sq_v0[k] - dp2;
            if(e->con_tol < fabs(miss))
            {
                double odp[3];
                adjust_direction(e, e->old_rattle + 3 * a, e->old_rattle + 3 * b, odp);
                const double dot_on = dot(odp, dp);
                const double lambda = 0.5 * miss * e->con_redm[k] / dot_on;
                const double rm1 = e->rmass[a], rm2 = e->rmass[b];
                double cf[3], c1[3], c2[3];
                for(int c = 0; c < 3; ++c) cf[c] = lambda * odp[c];
                for(int c = 0; c < 3; ++c) { c1[c] = cf[c] * rm1; c2[c] = cf[c] * rm2; }
                for(int c = 0; c < 3; ++c)
                {
                    p1[c] -= c1[c];
                    p2[c] += c2[c];
                    v1[c] -= c1[c] * e->r_dt_corr;
                    v2[c] += c2[c] * e->r_dt_corr;
                }
                corrected = 1;
            }
        }
        if(!corrected) break;
    }
}

/* the A step with its corrections (gBAOABLangevinIntegrator.hpp:363-372) */
static void gb_drift(orc_engine* e)
{
    for(size_t i = 0; i < e->n; ++i)
        for(int c = 0; c < 3; ++c)
        {
            e->old_rattle[3 * i + c] = e->pos[3 * i + c];
            e->pos[3 * i + c] += e->dt_corr * e->vel[3 * i + c];
        }
    gb_correct_coordinate(e);
    gb_correct_velocity(e);
}

static void gbaoab_step(orc_engine* e)
{
    /* gBAOABLangevinIntegrator::step (:348-433) with the engine's noise stream */
    const double h = e->halfdt;
    for(size_t i = 0; i < e->n; ++i)
    {
        const double rm = e->rmass[i];
        for(int c = 0; c < 3; ++c) e->vel[3 * i + c] += h * rm * e->force[3 * i + c];
    }
    gb_correct_velocity(e);
    gb_drift(e);
    for(size_t i = 0; i < e->n; ++i)
    {
        double R[3];
        orc_noise3(e->seed, e->step, (uint32_t)i, R);
        for(int c = 0; c < 3; ++c) e->vel[3 * i + c] *= e->c1[i];
        for(int c = 0; c < 3; ++c) e->vel[3 * i + c] += e->c2[i] * R[c];
    }
    gb_correct_velocity(e);
    gb_drift(e);
    double largest = 0.0;
    for(size_t i = 0; i < e->n; ++i)
    {
        double d[3];
        for(int c = 0; c < 3; ++c) d[c] = e->pos[3 * i + c] - e->old_pos[3 * i + c];
        largest = fmax(largest, length_sq(d));
        adjust_position(e, e->pos + 3 * i);
        for(int c = 0; c < 3; ++c) e->force[3 * i + c] = 0;
    }
    orc_reduce_margin(e, 2 * sqrt(largest));
    calc_force_impl(e, 0);
    for(size_t i = 0; i < e->n; ++i)
    {
        const double rm = e->rmass[i];
        for(int c = 0; c < 3; ++c) e->vel[3 * i + c] += h * rm * e->force[3 * i + c];
    }
    gb_correct_velocity(e);
}

static void baoab_step(orc_engine* e)
{
    /* BAOABLangevinIntegrator::step (:170-240) with the engine's noise stream */
    double largest = 0.0;
    const double h = e->halfdt;
    for(size_t i = 0; i < e->n; ++i)
    {
        double R[3];
        orc_noise3(e->seed, e->step, (uint32_t)i, R);
        const double rm = e->rmass[i];
        double* p = e->pos + 3 * i; double* v = e->vel + 3 * i; double* f = e->force + 3 * i;
        const double expgt = e->c1[i], nc = e->c2[i];
        double dp[3] = {0, 0, 0};
        for(int a = 0; a < 3; ++a) v[a] += h * rm * f[a];
        for(int a = 0; a < 3; ++a) dp[a] += h * v[a];
        for(int a = 0; a < 3; ++a) v[a] *= expgt;
        for(int a = 0; a < 3; ++a) v[a] += nc * R[a];
        for(int a = 0; a < 3; ++a) dp[a] += h * v[a];
        double np[3] = {p[0] + dp[0], p[1] + dp[1], p[2] + dp[2]};
        adjust_position(e, np);
        p[0] = np[0]; p[1] = np[1]; p[2] = np[2];
        f[0] = f[1] = f[2] = 0;
        const double l2 = length_sq(dp);
        largest = fmax(largest, l2);
    }
    orc_reduce_margin(e, 2 * sqrt(largest));
    calc_force_impl(e, 0);
    for(size_t i = 0; i < e->n; ++i)
    {
        const double rm = e->rmass[i];
        for(int a = 0; a < 3; ++a) e->vel[3 * i + a] += h * rm * e->force[3 * i + a];
    }
}

static void gjf_step(orc_engine* e)
{
    /* GJFNVTLangevinIntegrator::step (:166-244) with the engine's noise stream */
    double largest = 0;
    const double dt = e->dt;
    for(size_t i = 0; i < e->n; ++i)
    {
        const double rm = e->rmass[i];
        const double r2m = 0.5 * rm;
        double* p = e->pos + 3 * i; double* v = e->vel + 3 * i; double* f = e->force + 3 * i;
        const double b = e->c1[i], a = e->c3[i];
        double R[3], beta[3], dp[3];
        orc_noise3(e->seed, e->step, (uint32_t)i, R);
        for(int c = 0; c < 3; ++c) beta[c] = R[c] * e->c2[i];
        for(int c = 0; c < 3; ++c) dp[c] = (b * dt) * (v[c] + r2m * (dt * f[c] + beta[c]));
        double np[3] = {p[0] + dp[0], p[1] + dp[1], p[2] + dp[2]};
        adjust_position(e, np);
        p[0] = np[0]; p[1] = np[1]; p[2] = np[2];
        for(int c = 0; c < 3; ++c) v[c] = a * (v[c] + (dt * r2m) * f[c]) + b * rm * beta[c];
        f[0] = f[1] = f[2] = 0;
        largest = fmax(largest, length_sq(dp));
    }
    orc_reduce_margin(e, 2 * sqrt(largest));
    calc_force_impl(e, 0);
    for(size_t i = 0; i < e->n; ++i)
    {
        const double rm = e->rmass[i];
        for(int c = 0; c < 3; ++c) e->vel[3 * i + c] += (dt * 0.5 * rm) * e->force[3 * i + c];
    }
}

static void ul_step(orc_engine* e)
{
    /* UnderdampedLangevinIntegrator::step (:459-542); the noise of the closing acceleration
     * is the one of step count + 1, so a restart at step s (initialize at s) continues the
     * same stream */
    double largest = 0;
    const double dt = e->dt, h = e->halfdt, h2 = e->dt * e->dt / 2;
    for(size_t i = 0; i < e->n; ++i)
    {
        double* p = e->pos + 3 * i; double* v = e->vel + 3 * i; double* f = e->force + 3 * i;
        const double* a = e->accel + 3 * i;
        const double g2 = e->coef[i] * h;
        const double omg = 1 - g2;
        double dp[3];
        for(int c = 0; c < 3; ++c) dp[c] = (dt * omg) * v[c] + h2 * a[c];
        double np[3] = {p[0] + dp[0], p[1] + dp[1], p[2] + dp[2]};
        adjust_position(e, np);
        p[0] = np[0]; p[1] = np[1]; p[2] = np[2];
        const double sv = omg * (omg * omg + g2);
        for(int c = 0; c < 3; ++c) v[c] *= sv;
        for(int c = 0; c < 3; ++c) v[c] += (h * omg) * a[c];
        f[0] = f[1] = f[2] = 0;
        largest = fmax(largest, length_sq(dp));
    }
    orc_reduce_margin(e, 2 * sqrt(largest));
    calc_force_impl(e, 0);
    ul_accel(e, e->step + 1, 1);
}

/* SystemMotionRemover::remove (SystemMotionRemover.hpp:42-148): sequential sums in particle
 * order; the inertia tensor is inverted by math::determinant / math::inverse
 * (math/Matrix.hpp:340-375). */
static void remove_motion(orc_engine* e)
{
    const size_t n = e->n;
    double E_pre = 0;
    if(e->rm_rescale)
    {
        for(size_t i = 0; i < n; ++i) E_pre += e->mass[i] * length_sq(e->vel + 3 * i);
        E_pre *= 0.5;
    }
    if(e->rm_translation)
    {
        double m_tot = 0, t[3] = {0, 0, 0};
        for(size_t i = 0; i < n; ++i)
        {
            m_tot += e->mass[i];
            for(int a = 0; a < 3; ++a) t[a] += e->mass[i] * e->vel[3 * i + a];
        }
        const double inv = 1.0 / m_tot;
        for(int a = 0; a < 3; ++a) t[a] *= inv;
        for(size_t i = 0; i < n; ++i)
            for(int a = 0; a < 3; ++a) e->vel[3 * i + a] -= t[a];
    }
    if(e->rm_rotation)
    {
        double m_tot = 0, com[3] = {0, 0, 0};
        for(size_t i = 0; i < n; ++i)
        {
            m_tot += e->mass[i];
            for(int a = 0; a < 3; ++a) com[a] += e->mass[i] * e->pos[3 * i + a];
        }
        const double inv = 1.0 / m_tot;
        for(int a = 0; a < 3; ++a) com[a] *= inv;
        double L[3] = {0, 0, 0};
        double I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for(size_t i = 0; i < n; ++i)
        {
            const double m = e->mass[i];
            const double r[3] = {e->pos[3 * i] - com[0], e->pos[3 * i + 1] - com[1], e->pos[3 * i + 2] - com[2]};
            const double* v = e->vel + 3 * i;
            const double c[3] = {r[1] * v[2] - r[2] * v[1], r[2] * v[0] - r[0] * v[2], r[0] * v[1] - r[1] * v[0]};
            for(int a = 0; a < 3; ++a) L[a] += m * c[a];
            const double Ixx = m * r[0] * r[0], Iyy = m * r[1] * r[1], Izz = m * r[2] * r[2];
            const double Ixy = m * r[0] * r[1], Iyz = m * r[1] * r[2], Izx = m * r[2] * r[0];
            I[0][0] += Iyy + Izz;
            I[1][1] += Izz + Ixx;
            I[2][2] += Ixx + Iyy;
            I[0][1] -= Ixy;
            I[0][2] -= Izx;
            I[1][2] -= Iyz;
        }
        I[1][0] = I[0][1];
        I[2][0] = I[0][2];
        I[2][1] = I[1][2];
        const double det = I[0][0] * I[1][1] * I[2][2] + I[1][0] * I[2][1] * I[0][2] + I[2][0] * I[0][1] * I[1][2] -
                           I[0][0] * I[2][1] * I[1][2] - I[2][0] * I[1][1] * I[0][2] - I[1][0] * I[0][1] * I[2][2];
        if(det != 0.0)
        {
            const double di = 1.0 / det;
            double J[3][3];
            J[0][0] = di * (I[1][1] * I[2][2] - I[1][2] * I[2][1]);
            J[1][1] = di * (I[0][0] * I[2][2] - I[0][2] * I[2][0]);
            J[2][2] = di * (I[0][0] * I[1][1] - I[0][1] * I[1][0]);
            J[0][1] = di * (I[0][2] * I[2][1] - I[0][1] * I[2][2]);
            J[0][2] = di * (I[0][1] * I[1][2] - I[0][2] * I[1][1]);
            J[1][2] = di * (I[0][2] * I[1][0] - I[0][0] * I[1][2]);
            J[1][0] = di * (I[1][2] * I[2][0] - I[1][0] * I[2][2]);
            J[2][0] = di * (I[1][0] * I[2][1] - I[2][0] * I[1][1]);
            J[2][1] = di * (I[2][0] * I[0][1] - I[0][0] * I[2][1]);
            /* Matrix * Vector (math/Matrix.hpp operator*): row-wise dot products */
            double w[3];
            for(int a = 0; a < 3; ++a) w[a] = J[a][0] * L[0] + J[a][1] * L[1] + J[a][2] * L[2];
            for(size_t i = 0; i < n; ++i)
            {
                const double r[3] = {e->pos[3 * i] - com[0], e->pos[3 * i + 1] - com[1], e->pos[3 * i + 2] - com[2]};
                e->vel[3 * i + 0] -= w[1] * r[2] - w[2] * r[1];
                e->vel[3 * i + 1] -= w[2] * r[0] - w[0] * r[2];
                e->vel[3 * i + 2] -= w[0] * r[1] - w[1] * r[0];
            }
        }
    }
    if(e->rm_rescale)
    {
        double E_post = 0;
        for(size_t i = 0; i < n; ++i) E_post += e->mass[i] * length_sq(e->vel + 3 * i);
        E_post *= 0.5;
        const double scale = sqrt(E_pre / E_post);
        for(size_t i = 0; i < 3 * n; ++i) e->vel[i] *= scale;
    }
}

int orc_set_remover(orc_engine* e, int translation, int rotation, int rescale)
{
    e->rm_translation = translation != 0;
    e->rm_rotation = rotation != 0;
    e->rm_rescale = rescale != 0;
    return 0;
}

int orc_step(orc_engine* e, size_t nsteps)
{
    for(size_t s = 0; s < nsteps; ++s)
    {
        if(e->integ == ORC_INTEG_BAOAB) baoab_step(e);
        else if(e->integ == ORC_INTEG_GJF) gjf_step(e);
        else if(e->integ == ORC_INTEG_GBAOAB) gbaoab_step(e);
        else if(e->integ == ORC_INTEG_UNDERDAMPED) ul_step(e);
        else return -1;
        if(e->rm_translation || e->rm_rotation) remove_motion(e);
        if(e->integ == ORC_INTEG_GBAOAB) memcpy(e->old_pos, e->pos, 3 * e->n * sizeof(double));
        e->step++;
    }
    return 0;
}
