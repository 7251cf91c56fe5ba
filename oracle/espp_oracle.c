/*
 * espp_oracle.c -- FP64 CPU restatement of ESPResSo++'s short-range MD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product path (espressopp_amd, libespp_amd.so) never
 * links, loads or calls it.
 *
 * What it restates (reference paths are relative to the ESPResSo++ tree):
 *   - cell grid geometry and particle->cell mapping
 *       src/storage/CellGrid.cpp:40-60 (cellSize = (right-left)/n, inv = 1/cellSize)
 *       src/storage/CellGrid.cpp:98-136 (mapPositionToCellChecked, truncation + clip)
 *       src/esutil/Grid.hpp:68-71 (linear index x + nx*(y + ny*z))
 *   - folding: src/bc/OrthorhombicBC.cpp:144-161 (tmp=floor(x*invL); img+=tmp; x-=tmp*L)
 *   - minimum image for bonds: src/bc/OrthorhombicBC.cpp:86-105
 *   - half-shell cell neighbours (halfCellInt stencil): src/storage/DomainDecomposition.cpp:325-374
 *       (useForAllPairs = cell2Idx < cellIdx; the list uses the others)
 *   - ghost shift: src/storage/DomainDecomposition.cpp:1320-1343 with
 *       src/storage/NodeGrid.cpp:117-118 (right ghost frame = left cells + L)
 *       and src/Particle.hpp:93-98 (ghost position = p + shift)
 *   - Verlet list: src/VerletList.cpp:44-67 (cutsq=(rc+skin)^2),
 *       src/VerletList.cpp:123-279 (_rebuildUsingBuffers: self cell p2>p1,
 *       then half-shell buffer; include iff d.sqr() <= cutsq; exclusions both ways)
 *   - Lennard-Jones: src/interaction/LennardJones.hpp:79-160,
 *       src/interaction/Potential.hpp:190-235,360-430 (cut test distSqr>cutoffSqr)
 *   - FENE: src/interaction/FENE.hpp:91-112
 *   - Tabulated: src/interaction/Tabulated.hpp:43-115 with its interpolation
 *       tables InterpolationLinear.cpp:35-150 (.hpp:58-75), InterpolationAkima.cpp:136-214
 *       (.hpp:86-120) and InterpolationCubic.cpp:58-215 (.hpp:89-120)
 *   - pair loops with Newton-3, energy, virial, virial tensor:
 *       src/interaction/VerletListInteractionTemplate.hpp:116-331
 *       src/interaction/FixedPairListInteractionTemplate.hpp:112-262,400-462
 *   - velocity Verlet loop and skin trigger: src/integrator/VelocityVerlet.cpp:63-169,236-305,307-371
 *   - Langevin: src/integrator/LangevinThermostat.cpp:141-214 (uniform noise,
 *       pref1=-gamma, pref2=sqrt(24 T gamma/dt), heatUp *sqrt(3) at run() start)
 *   - observables: src/analysis/Temperature.hpp:60-125, src/analysis/Pressure.cpp:41-121,
 *       src/analysis/RadialDistrF.cpp:47-140
 *
 * Arithmetic follows the reference's operation order with no FMA contraction
 * (the reference builds with -std=c++17, ISO mode, baseline x86-64): build this
 * file with -ffp-contract=off.
 *
 * Deliberate differences, documented because they cannot be pinned:
 *   - RNG: the reference uses boost::lagged_fibonacci607 (src/esutil/RNG.cpp:40-66).
 *     Here the Langevin noise comes from Philox4x32-10 (Salmon et al., SC'11)
 *     keyed (seed, thermostat call, particle id), the same stream the GPU path
 *     uses, so thermostatted GPU and oracle runs can be compared step by step.
 *   - particle order inside a cell is ascending insertion index; the reference
 *     keeps migration order.  Pair sets do not depend on it; FP64 force sums
 *     differ only by rounding.
 *   - type pairs without a potential are skipped (the reference evaluates a
 *     default LennardJones(eps=0, sigma=0) that yields zero force).
 *   - table index: clamped to the last interval [N-2] where the reference
 *     clamps to N-1 and then reads past its coefficient arrays (distances at or
 *     beyond the last grid point); the Akima table keeps the reference's
 *     integer truncation of z*z.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXVL 16
#define OR_MAXFPL 16
#define OR_MAXIA 32
#define OR_MAXT 64

typedef struct {
    int set;
    double eps, sig, rc, shift;
    double ff1, ff2, rc2;
} or_lj;

typedef struct {
    double cut, cutsq;
    int64_t npairs, cap;
    int32_t *p1, *p2;
    int8_t *sh;             /* 3 per pair: image shift applied to p2 */
    int64_t builds;
    uint64_t *excl;          /* sorted, packed (min id, max id) */
    int64_t nexcl;
    int max_type;
    int64_t *cell_pstart;    /* pairs of inner cell c are [cell_pstart[c], cell_pstart[c+1]) */
} or_vl;

typedef struct {
    int64_t n, cap;
    int32_t *a, *b;          /* bond ids, stored as given (a, b) */
} or_fpl;

/* Interpolation table of interaction.Tabulated (src/interaction/Tabulated.hpp:43-115).
 * itype 1 linear (src/interaction/InterpolationLinear.cpp:35-150), 2 Akima
 * (src/interaction/InterpolationAkima.cpp, .hpp:86-125), 3 cubic spline
 * (src/interaction/InterpolationCubic.cpp, .hpp:89-120). */
typedef struct {
    int set, itype, N;
    double inner, outer, delta, invdelta, deltasq6;
    double cutoff, cutoffSqr, shift;
    double *radius;
    double *ce[4], *cf[4];   /* per-interval coefficients of energy and force */
} or_tab;

typedef struct {
    int kind;                /* 0: VerletList LJ, 1: FixedPairList FENE,
                                2: VerletList Tabulated, 3: FixedPairList Tabulated */
    int list;
    or_lj lj[OR_MAXT * OR_MAXT];
    double K, r0, rMax, rMaxSqr, cutoff, cutoffSqr, shift;
    or_tab *tab[OR_MAXT * OR_MAXT];   /* kind 2: per type pair; kind 3: tab[0] */
} or_ia;

typedef struct or_sim {
    double L[3], invL[3], halfL[3];
    int n[3];
    int hci;                       /* halfCellInt: cell neighbours within +-hci cells */
    int nstencil;                  /* neighbour offsets with a larger framed index */
    int stencil[2456][3];
    double cs[3], ics[3];
    int64_t ncell;
    double skin;
    uint64_t seed;

    int64_t np, cap;
    int32_t *id, *type, *img;
    double *x, *v, *f, *m;
    int32_t *loc_of_id;
    int64_t nid;

    int32_t *cell_of;
    int64_t *cstart;
    int32_t *clist;

    or_vl vl[OR_MAXVL];
    int nvl;
    or_fpl fpl[OR_MAXFPL];
    int nfpl;
    or_ia *ia[OR_MAXIA];
    int nia;

    double dt;
    int64_t step;
    double maxDist;
    int resortFlag;
    int64_t nResorts;

    int lang_on;
    double gamma, temperature, pref1, pref2;
    int32_t *lang_excl;
    int64_t nlang_excl;
    uint64_t therm_calls;

    int nthreads;
    double t_force, t_resort, t_int1, t_int2, t_run;
    char err[256];
} or_sim;

static void assign_cells(or_sim *s, int64_t i0, int64_t i1);
static void build_cell_lists(or_sim *s);

static void or_seterr(or_sim *s, const char *msg) {
    if (s) { strncpy(s->err, msg, sizeof(s->err) - 1); s->err[sizeof(s->err) - 1] = 0; }
}

const char *or_last_error(or_sim *s) { return s ? s->err : "null handle"; }

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11.  Multipliers and Weyl increments are the published ones. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        if (r < 9) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Three uniforms in (0,1) for (seed, call, particle id); shared with the GPU path. */
static inline void langevin_uniforms(uint64_t seed, uint64_t call, int32_t pid, double u[3]) {
    uint32_t ctr[4] = {(uint32_t)pid, (uint32_t)call, (uint32_t)(call >> 32), 0x4C414E47u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    for (int k = 0; k < 3; ++k) u[k] = ((double)o[k] + 0.5) * 2.3283064365386963e-10;
}

void or_langevin_uniforms(uint64_t seed, uint64_t call, int32_t pid, double *u3) {
    langevin_uniforms(seed, call, pid, u3);
}

static void make_stencil(or_sim *s);

/* ---------------------------------------------------------------- lifecycle */
or_sim *or_create(const double *L, const int *cells, double skin, uint64_t seed) {
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    if (!s) return NULL;
    for (int d = 0; d < 3; ++d) {
        s->L[d] = L[d];
        s->invL[d] = 1.0 / L[d];
        s->halfL[d] = 0.5 * L[d];
        s->n[d] = cells[d];
        /* CellGrid: cellSize = (myRight - myLeft) / n with myLeft = 0, myRight = L */
        s->cs[d] = (L[d] - 0.0) / (double)cells[d];
        s->ics[d] = 1.0 / s->cs[d];
    }
    s->ncell = (int64_t)cells[0] * cells[1] * cells[2];
    s->hci = 1;
    make_stencil(s);
    s->skin = skin;
    s->seed = seed;
    s->dt = 0.005;
    s->resortFlag = 1;
    s->nthreads = 1;
    s->cstart = (int64_t *)calloc((size_t)s->ncell + 1, sizeof(int64_t));
    return s;
}

/* DomainDecomposition's halfCellInt: the cells given to or_create are the fine
 * grid (tools.decomp.cellGrid(..., halfCellInt)), neighbours within +-h cells. */
int or_set_halfcell(or_sim *s, int h) {
    if (h < 1 || h > 8) return -1;
    for (int d = 0; d < 3; ++d) if (s->n[d] < 2 * h) return -1;
    s->hci = h;
    make_stencil(s);
    return 0;
}

static void tab_free(or_tab *t) {
    if (!t) return;
    free(t->radius);
    for (int k = 0; k < 4; ++k) { free(t->ce[k]); free(t->cf[k]); }
    free(t);
}

static void ia_free(or_ia *a) {
    if (!a) return;
    for (int k = 0; k < OR_MAXT * OR_MAXT; ++k) {
        or_tab *t = a->tab[k];
        if (!t) continue;
        /* a table is shared by (t1, t2) and (t2, t1): clear every alias before freeing */
        for (int q = k; q < OR_MAXT * OR_MAXT; ++q) if (a->tab[q] == t) a->tab[q] = NULL;
        tab_free(t);
    }
    free(a);
}

static void vl_free(or_vl *v) {
    free(v->p1); free(v->p2); free(v->sh); free(v->excl); free(v->cell_pstart);
    memset(v, 0, sizeof(*v));
}

void or_destroy(or_sim *s) {
    if (!s) return;
    free(s->id); free(s->type); free(s->img); free(s->x); free(s->v); free(s->f); free(s->m);
    free(s->loc_of_id); free(s->cell_of); free(s->cstart); free(s->clist);
    for (int i = 0; i < s->nvl; ++i) vl_free(&s->vl[i]);
    for (int i = 0; i < s->nfpl; ++i) { free(s->fpl[i].a); free(s->fpl[i].b); }
    for (int i = 0; i < s->nia; ++i) ia_free(s->ia[i]);
    free(s->lang_excl);
    free(s);
}

void or_set_threads(or_sim *s, int n) { s->nthreads = n > 0 ? n : 1; }
void or_set_skin(or_sim *s, double skin) { s->skin = skin; }
void or_set_dt(or_sim *s, double dt) { s->dt = dt; }
int64_t or_get_step(or_sim *s) { return s->step; }
void or_set_step(or_sim *s, int64_t st) { s->step = st; }
int64_t or_num_resorts(or_sim *s) { return s->nResorts; }
int64_t or_num_particles(or_sim *s) { return s->np; }

/* fold one coordinate: src/bc/OrthorhombicBC.cpp:144-161 */
static inline void fold_coord(const or_sim *s, double *x, int32_t *img, int d) {
    int tmp = (int)floor(x[d] * s->invL[d]);
    img[d] += tmp;
    x[d] -= tmp * s->L[d];
}

/* inner cell coordinate of a folded position: CellGrid::mapPositionToCellChecked
 * for a single node (myLeft = 0), returning the clipped inner coordinate. */
static inline int cell_coord(const or_sim *s, double xd, int d) {
    double lpos = xd - 0.0;
    int c = (int)(lpos * s->ics[d]) + 1;           /* + frame */
    if (c < 1) c = 1;
    else if (c >= s->n[d] + 2 - 1) c = s->n[d] + 2 - 1 - 1;
    return c - 1;
}

static int ensure_cap(or_sim *s, int64_t need) {
    if (need <= s->cap) return 0;
    int64_t nc = s->cap ? s->cap : 1024;
    while (nc < need) nc *= 2;
#define GROW(p, T, k) do { void *q = realloc(p, (size_t)nc * (k) * sizeof(T)); if (!q) return -1; p = (T *)q; } while (0)
    GROW(s->id, int32_t, 1); GROW(s->type, int32_t, 1); GROW(s->img, int32_t, 3);
    GROW(s->x, double, 3); GROW(s->v, double, 3); GROW(s->f, double, 3); GROW(s->m, double, 1);
    GROW(s->cell_of, int32_t, 1); GROW(s->clist, int32_t, 1);
#undef GROW
    s->cap = nc;
    return 0;
}

static int rebuild_id_map(or_sim *s) {
    int64_t maxid = -1;
    for (int64_t i = 0; i < s->np; ++i) if (s->id[i] > maxid) maxid = s->id[i];
    free(s->loc_of_id);
    s->nid = maxid + 1;
    s->loc_of_id = (int32_t *)malloc((size_t)(s->nid > 0 ? s->nid : 1) * sizeof(int32_t));
    if (!s->loc_of_id) return -1;
    for (int64_t i = 0; i < s->nid; ++i) s->loc_of_id[i] = -1;
    for (int64_t i = 0; i < s->np; ++i) s->loc_of_id[s->id[i]] = (int32_t)i;
    return 0;
}

/* Storage::addParticle (src/storage/Storage.cpp:285-313): positions are folded on insertion. */
int or_add_particles(or_sim *s, int64_t n, const int32_t *id, const int32_t *type,
                     const double *pos, const double *vel, const double *mass) {
    if (ensure_cap(s, s->np + n)) { or_seterr(s, "out of memory"); return -1; }
    for (int64_t k = 0; k < n; ++k) {
        if (id[k] < 0) { or_seterr(s, "negative particle id"); return -2; }
        if (id[k] < s->nid && s->loc_of_id[id[k]] >= 0) { or_seterr(s, "particle id already exists"); return -3; }
        int64_t i = s->np + k;
        s->id[i] = id[k];
        s->type[i] = type ? type[k] : 0;
        for (int d = 0; d < 3; ++d) {
            s->x[3 * i + d] = pos[3 * k + d];
            s->v[3 * i + d] = vel ? vel[3 * k + d] : 0.0;
            s->f[3 * i + d] = 0.0;
            s->img[3 * i + d] = 0;
        }
        s->m[i] = mass ? mass[k] : 1.0;
        for (int d = 0; d < 3; ++d) fold_coord(s, &s->x[3 * i], &s->img[3 * i], d);
    }
    int64_t first = s->np;
    s->np += n;
    assign_cells(s, first, s->np);
    build_cell_lists(s);
    if (rebuild_id_map(s)) { or_seterr(s, "out of memory"); return -1; }
    return 0;
}

/* ---------------------------------------------------------------- cells */
static void assign_cells(or_sim *s, int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
        const double *x = &s->x[3 * i];
        int cx = cell_coord(s, x[0], 0), cy = cell_coord(s, x[1], 1), cz = cell_coord(s, x[2], 2);
        s->cell_of[i] = (int32_t)(cx + s->n[0] * (cy + s->n[1] * cz));
    }
}

/* cell member lists from cell_of, members in ascending insertion index */
static void build_cell_lists(or_sim *s) {
    int64_t nc = s->ncell;
    memset(s->cstart, 0, (size_t)(nc + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < s->np; ++i) s->cstart[s->cell_of[i] + 1]++;
    for (int64_t c = 0; c < nc; ++c) s->cstart[c + 1] += s->cstart[c];
    int64_t *fill = (int64_t *)malloc((size_t)nc * sizeof(int64_t));
    memcpy(fill, s->cstart, (size_t)nc * sizeof(int64_t));
    for (int64_t i = 0; i < s->np; ++i) s->clist[fill[s->cell_of[i]]++] = (int32_t)i;
    free(fill);
}

static void bin_particles(or_sim *s) {
    assign_cells(s, 0, s->np);
    build_cell_lists(s);
}

/* Half-shell offsets in the order DomainDecomposition::initCellInteractions
 * enumerates the neighbours within +-halfCellInt cells (z outer, then y, then x;
 * src/storage/DomainDecomposition.cpp:342-358), keeping those whose framed
 * linear index is larger than the cell's own (useForAllPairs == false): the
 * lexicographically positive offsets.  For halfCellInt 1 these are 13 cells. */
static void make_stencil(or_sim *s) {
    const int h = s->hci;
    s->nstencil = 0;
    for (int p = -h; p <= h; ++p)
        for (int q = -h; q <= h; ++q)
            for (int r = -h; r <= h; ++r)
                if (p > 0 || (p == 0 && (q > 0 || (q == 0 && r > 0)))) {
                    s->stencil[s->nstencil][0] = r;
                    s->stencil[s->nstencil][1] = q;
                    s->stencil[s->nstencil][2] = p;
                    s->nstencil++;
                }
}

static int is_excluded(const or_vl *v, int32_t a, int32_t b) {
    if (!v->nexcl) return 0;
    uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    uint64_t key = ((uint64_t)lo << 32) | hi;
    int64_t l = 0, r = v->nexcl - 1;
    while (l <= r) {
        int64_t mid = (l + r) / 2;
        if (v->excl[mid] == key) return 1;
        if (v->excl[mid] < key) l = mid + 1; else r = mid - 1;
    }
    return 0;
}

typedef struct { int64_t n, cap; int32_t *p1, *p2; int8_t *sh; int max_type; } pairbuf;

static void pb_push(pairbuf *b, int32_t i, int32_t j, const int sh[3]) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 256;
        b->p1 = (int32_t *)realloc(b->p1, (size_t)b->cap * sizeof(int32_t));
        b->p2 = (int32_t *)realloc(b->p2, (size_t)b->cap * sizeof(int32_t));
        b->sh = (int8_t *)realloc(b->sh, (size_t)b->cap * 3);
    }
    b->p1[b->n] = i; b->p2[b->n] = j;
    b->sh[3 * b->n] = (int8_t)sh[0]; b->sh[3 * b->n + 1] = (int8_t)sh[1]; b->sh[3 * b->n + 2] = (int8_t)sh[2];
    b->n++;
}

/* VerletList::_rebuildUsingBuffers for one real cell */
static void vl_build_cell(const or_sim *s, const or_vl *v, int64_t c, pairbuf *out) {
    int cx = (int)(c % s->n[0]), cy = (int)((c / s->n[0]) % s->n[1]), cz = (int)(c / ((int64_t)s->n[0] * s->n[1]));
    int64_t b0 = s->cstart[c], b1 = s->cstart[c + 1];
    const int zero[3] = {0, 0, 0};
    for (int64_t a = b0; a < b1; ++a) {
        int32_t i = s->clist[a];
        const double *x1 = &s->x[3 * i];
        /* self loop: CheckPair(p1, p2), d = p1 - p2 */
        for (int64_t b = a + 1; b < b1; ++b) {
            int32_t j = s->clist[b];
            const double *x2 = &s->x[3 * j];
            double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > v->cutsq) continue;
            if (is_excluded(v, s->id[i], s->id[j])) continue;
            pb_push(out, i, j, zero);
            int t = s->type[i] > s->type[j] ? s->type[i] : s->type[j];
            if (t > out->max_type) out->max_type = t;
        }
        /* neighbour loop over the half-shell buffer (ghost copies carry p + shift) */
        for (int k = 0; k < s->nstencil; ++k) {
            int q[3] = {cx + s->stencil[k][0], cy + s->stencil[k][1], cz + s->stencil[k][2]};
            int sh[3];
            for (int d = 0; d < 3; ++d) {
                sh[d] = 0;
                if (q[d] < 0) { q[d] += s->n[d]; sh[d] = -1; }
                else if (q[d] >= s->n[d]) { q[d] -= s->n[d]; sh[d] = 1; }
            }
            int64_t c2 = q[0] + (int64_t)s->n[0] * (q[1] + (int64_t)s->n[1] * q[2]);
            double shift[3] = {sh[0] * s->L[0], sh[1] * s->L[1], sh[2] * s->L[2]};
            for (int64_t b = s->cstart[c2]; b < s->cstart[c2 + 1]; ++b) {
                int32_t j = s->clist[b];
                const double *x2 = &s->x[3 * j];
                double g0 = x2[0] + shift[0], g1 = x2[1] + shift[1], g2 = x2[2] + shift[2];
                double dx = x1[0] - g0, dy = x1[1] - g1, dz = x1[2] - g2;
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > v->cutsq) continue;
                if (is_excluded(v, s->id[i], s->id[j])) continue;
                int t = s->type[i] > s->type[j] ? s->type[i] : s->type[j];
                if (t > out->max_type) out->max_type = t;
                pb_push(out, i, j, sh);
            }
        }
    }
}

static int vl_rebuild(or_sim *s, or_vl *v) {
    double cutVerlet = v->cut + s->skin;
    v->cutsq = cutVerlet * cutVerlet;
    int64_t nc = s->ncell;
    pairbuf *bufs = (pairbuf *)calloc((size_t)nc, sizeof(pairbuf));
    if (!bufs) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(s->nthreads)
    for (int64_t c = 0; c < nc; ++c) vl_build_cell(s, v, c, &bufs[c]);
    int64_t tot = 0;
    for (int64_t c = 0; c < nc; ++c) tot += bufs[c].n;
    if (tot > v->cap) {
        v->cap = tot + tot / 4 + 16;
        v->p1 = (int32_t *)realloc(v->p1, (size_t)v->cap * sizeof(int32_t));
        v->p2 = (int32_t *)realloc(v->p2, (size_t)v->cap * sizeof(int32_t));
        v->sh = (int8_t *)realloc(v->sh, (size_t)v->cap * 3);
    }
    free(v->cell_pstart);
    v->cell_pstart = (int64_t *)malloc((size_t)(nc + 1) * sizeof(int64_t));
    int64_t o = 0;
    v->max_type = 0;
    for (int64_t c = 0; c < nc; ++c) {
        v->cell_pstart[c] = o;
        if (bufs[c].n) {
            memcpy(v->p1 + o, bufs[c].p1, (size_t)bufs[c].n * sizeof(int32_t));
            memcpy(v->p2 + o, bufs[c].p2, (size_t)bufs[c].n * sizeof(int32_t));
            memcpy(v->sh + 3 * o, bufs[c].sh, (size_t)bufs[c].n * 3);
        }
        if (bufs[c].max_type > v->max_type) v->max_type = bufs[c].max_type;
        o += bufs[c].n;
        free(bufs[c].p1); free(bufs[c].p2); free(bufs[c].sh);
    }
    v->cell_pstart[nc] = o;
    v->npairs = o;
    v->builds++;
    free(bufs);
    return 0;
}

/* Storage::decompose for one node: fold x, y, z, re-sort into cells, then the
 * onParticlesChanged signal rebuilds every Verlet list. */
int or_decompose(or_sim *s) {
    for (int64_t i = 0; i < s->np; ++i) {
        double *x = &s->x[3 * i];
        if (x[0] != x[0] || x[1] != x[1] || x[2] != x[2]) {
            or_seterr(s, "particle has moved to outer space (one or more coordinates are nan)");
            return -4;
        }
        for (int d = 0; d < 3; ++d) fold_coord(s, x, &s->img[3 * i], d);
    }
    bin_particles(s);
    for (int k = 0; k < s->nvl; ++k)
        if (vl_rebuild(s, &s->vl[k])) { or_seterr(s, "out of memory"); return -1; }
    return 0;
}

/* ---------------------------------------------------------------- lists */
int or_vl_create(or_sim *s, double cutoff) {
    if (s->nvl >= OR_MAXVL) { or_seterr(s, "too many verlet lists"); return -1; }
    or_vl *v = &s->vl[s->nvl];
    memset(v, 0, sizeof(*v));
    v->cut = cutoff;
    if (vl_rebuild(s, v)) return -1;
    return s->nvl++;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int or_vl_exclude(or_sim *s, int vi, int64_t n, const int32_t *pairs) {
    or_vl *v = &s->vl[vi];
    v->excl = (uint64_t *)realloc(v->excl, (size_t)(v->nexcl + n) * sizeof(uint64_t));
    for (int64_t k = 0; k < n; ++k) {
        uint32_t a = (uint32_t)pairs[2 * k], b = (uint32_t)pairs[2 * k + 1];
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        v->excl[v->nexcl++] = ((uint64_t)lo << 32) | hi;
    }
    qsort(v->excl, (size_t)v->nexcl, sizeof(uint64_t), cmp_u64);
    return vl_rebuild(s, v);
}

int64_t or_vl_size(or_sim *s, int vi) { return s->vl[vi].npairs; }
int64_t or_vl_builds(or_sim *s, int vi) { return s->vl[vi].builds; }

/* pairs as particle ids in the reference's orientation (first = real particle) */
int64_t or_vl_pairs(or_sim *s, int vi, int32_t *out, int64_t cap) {
    or_vl *v = &s->vl[vi];
    int64_t n = v->npairs < cap ? v->npairs : cap;
    for (int64_t k = 0; k < n; ++k) { out[2 * k] = s->id[v->p1[k]]; out[2 * k + 1] = s->id[v->p2[k]]; }
    return v->npairs;
}

int or_fpl_create(or_sim *s) {
    if (s->nfpl >= OR_MAXFPL) { or_seterr(s, "too many fixed pair lists"); return -1; }
    memset(&s->fpl[s->nfpl], 0, sizeof(or_fpl));
    return s->nfpl++;
}

int or_fpl_add(or_sim *s, int li, int64_t n, const int32_t *pairs) {
    or_fpl *p = &s->fpl[li];
    if (p->n + n > p->cap) {
        p->cap = (p->n + n) * 2;
        p->a = (int32_t *)realloc(p->a, (size_t)p->cap * sizeof(int32_t));
        p->b = (int32_t *)realloc(p->b, (size_t)p->cap * sizeof(int32_t));
    }
    for (int64_t k = 0; k < n; ++k) { p->a[p->n] = pairs[2 * k]; p->b[p->n] = pairs[2 * k + 1]; p->n++; }
    return 0;
}

/* ---------------------------------------------------------------- interactions */
static void lj_preset(or_lj *p) {
    double sig2 = p->sig * p->sig;
    double sig6 = sig2 * sig2 * sig2;
    p->ff1 = 48.0 * p->eps * sig6 * sig6;
    p->ff2 = 24.0 * p->eps * sig6;
    p->rc2 = p->rc * p->rc;
}

static double lj_energy_raw(const or_lj *p, double distSqr) {
    double frac2 = p->sig * p->sig / distSqr;
    double frac6 = frac2 * frac2 * frac2;
    return 4.0 * p->eps * (frac6 * frac6 - frac6);
}

int or_ia_lj_create(or_sim *s, int vl) {
    if (s->nia >= OR_MAXIA) { or_seterr(s, "too many interactions"); return -1; }
    or_ia *a = (or_ia *)calloc(1, sizeof(or_ia));
    a->kind = 0; a->list = vl;
    s->ia[s->nia] = a;
    return s->nia++;
}

/* shift: NaN selects 'auto' (LennardJones::setAutoShift, src/interaction/Potential.hpp:214-222) */
int or_ia_lj_set(or_sim *s, int ii, int t1, int t2, double eps, double sig, double rc, double shift) {
    if (t1 < 0 || t2 < 0 || t1 >= OR_MAXT || t2 >= OR_MAXT) { or_seterr(s, "type out of range"); return -1; }
    or_lj p;
    memset(&p, 0, sizeof(p));
    p.set = 1; p.eps = eps; p.sig = sig; p.rc = rc;
    lj_preset(&p);
    if (shift != shift) p.shift = isinf(rc) ? 0.0 : lj_energy_raw(&p, p.rc2);
    else p.shift = shift;
    s->ia[ii]->lj[t1 * OR_MAXT + t2] = p;
    s->ia[ii]->lj[t2 * OR_MAXT + t1] = p;
    return 0;
}

double or_lj_energy(double eps, double sig, double rc, double shift, double r) {
    or_lj p; memset(&p, 0, sizeof(p));
    p.eps = eps; p.sig = sig; p.rc = rc; lj_preset(&p);
    if (shift != shift) shift = isinf(rc) ? 0.0 : lj_energy_raw(&p, p.rc2);
    double d2 = r * r;
    if (d2 > p.rc2) return 0.0;
    return lj_energy_raw(&p, d2) - shift;
}

static inline int lj_force(const or_lj *p, const double d[3], double F[3]) {
    double distSqr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (distSqr > p->rc2) return 0;
    double frac2 = 1.0 / distSqr;
    double frac6 = frac2 * frac2 * frac2;
    double ffactor = frac6 * (p->ff1 * frac6 - p->ff2) * frac2;
    F[0] = d[0] * ffactor; F[1] = d[1] * ffactor; F[2] = d[2] * ffactor;
    return 1;
}

void or_lj_force(double eps, double sig, double rc, const double *d, double *F) {
    or_lj p; memset(&p, 0, sizeof(p));
    p.eps = eps; p.sig = sig; p.rc = rc; lj_preset(&p);
    if (!lj_force(&p, d, F)) F[0] = F[1] = F[2] = 0.0;
}

int or_ia_fene_create(or_sim *s, int fpl, double K, double r0, double rMax, double cutoff, double shift) {
    if (s->nia >= OR_MAXIA) { or_seterr(s, "too many interactions"); return -1; }
    or_ia *a = (or_ia *)calloc(1, sizeof(or_ia));
    a->kind = 1; a->list = fpl;
    a->K = K; a->r0 = r0; a->rMax = rMax; a->rMaxSqr = rMax * rMax;
    a->cutoff = cutoff; a->cutoffSqr = cutoff * cutoff; a->shift = shift;
    s->ia[s->nia] = a;
    return s->nia++;
}

static double fene_energy_raw(const or_ia *a, double distSqr) {
    double t = (sqrt(distSqr) - a->r0) / a->rMax;
    return -0.5 * a->rMaxSqr * a->K * log(1 - pow(t, 2));
}

static inline int fene_force(const or_ia *a, const double d[3], double F[3]) {
    double distSqr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (distSqr > a->cutoffSqr) return 0;
    double ffactor;
    if (a->r0 > 2.220446049250313e-16) {
        double r = sqrt(distSqr);
        ffactor = -a->K * (r - a->r0) / (r * (1 - ((r - a->r0) * (r - a->r0) / a->rMaxSqr)));
    } else {
        ffactor = -a->K / (1.0 - distSqr / a->rMaxSqr);
    }
    F[0] = d[0] * ffactor; F[1] = d[1] * ffactor; F[2] = d[2] * ffactor;
    return 1;
}

double or_fene_energy(double K, double r0, double rMax, double r) {
    or_ia a; memset(&a, 0, sizeof(a));
    a.K = K; a.r0 = r0; a.rMax = rMax; a.rMaxSqr = rMax * rMax;
    return fene_energy_raw(&a, r * r);
}

void or_fene_force(double K, double r0, double rMax, const double *d, double *F) {
    or_ia a; memset(&a, 0, sizeof(a));
    a.K = K; a.r0 = r0; a.rMax = rMax; a.rMaxSqr = rMax * rMax; a.cutoffSqr = INFINITY;
    fene_force(&a, d, F);
}

/* ---------------------------------------------------------------- Tabulated */
/* Coefficients of the linear table: a r + b on interval i
 * (InterpolationLinear::spline, src/interaction/InterpolationLinear.cpp:139-150). */
static void tab_linear(const double *x, const double *y, int N, double *a, double *b) {
    for (int i = 0; i < N - 1; i++) {
        a[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        b[i] = y[i] - a[i] * x[i];
    }
    a[N - 1] = a[N - 2 >= 0 ? N - 2 : 0];
    b[N - 1] = b[N - 2 >= 0 ? N - 2 : 0];
}

/* Second derivatives of the clamped cubic spline: the tridiagonal sweep of
 * InterpolationCubic::spline (src/interaction/InterpolationCubic.cpp:170-215);
 * an end slope above 0.99e30 selects the natural end. */
static void tab_cubic(const double *x, const double *y, int n, double yp1, double ypn, double *y2) {
    double *u = (double *)malloc((size_t)n * sizeof(double));
    if (yp1 > 0.99e30) { y2[0] = 0.0; u[0] = 0.0; }
    else {
        y2[0] = -0.5;
        u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn, un;
    if (ypn > 0.99e30) { qn = 0.0; un = 0.0; }
    else {
        qn = 0.5;
        un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    }
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
    free(u);
}

/* Akima slope from four neighbouring secants (InterpolationAkima::getSlope,
 * src/interaction/InterpolationAkima.hpp:108-120). */
static double akima_slope(double m1, double m2, double m3, double m4) {
    if ((m1 == m2) && (m3 == m4)) return (m2 + m3) / 2.0;
    return (fabs(m4 - m3) * m2 + fabs(m2 - m1) * m3) / (fabs(m4 - m3) + fabs(m2 - m1));
}

/* Akima coefficients p0..p3 per interval, with two points extrapolated on each
 * side by a parabola (InterpolationAkima::spline, src/interaction/InterpolationAkima.cpp:136-214).
 * Needs N >= 4. */
static void tab_akima(const double *x, const double *y, int N, double *p0, double *p1, double *p2, double *p3) {
    double m1, m2, m3, m4, m5, temp, g0, g1, g2, x1, x2, y1, y2, x4, x5, y4, y5;
    double *t = (double *)malloc((size_t)N * sizeof(double));
    temp = (x[1] - x[0]) / (x[2] - x[0]);
    temp = temp * temp;
    g0 = y[0];
    g1 = ((y[1] - y[0]) - temp * (y[2] - y[0])) / ((x[1] - x[0]) - temp * (x[2] - x[0]));
    g2 = ((y[2] - y[0]) - g1 * (x[2] - x[0])) / ((x[2] - x[0]) * (x[2] - x[0]));
    x1 = x[0] - (x[2] - x[0]);
    x2 = x[1] - (x[2] - x[0]);
    y1 = g0 + g1 * (x1 - x[0]) + g2 * (x1 - x[0]) * (x1 - x[0]);
    y2 = g0 + g1 * (x2 - x[0]) + g2 * (x2 - x[0]) * (x2 - x[0]);
    m1 = (y2 - y1) / (x2 - x1);
    m2 = (y[0] - y2) / (x[0] - x2);
    m3 = (y[1] - y[0]) / (x[1] - x[0]);
    m4 = (y[2] - y[1]) / (x[2] - x[1]);
    t[0] = akima_slope(m1, m2, m3, m4);
    m5 = (y[3] - y[2]) / (x[3] - x[2]);
    t[1] = akima_slope(m2, m3, m4, m5);
    temp = (x[N - 2] - x[N - 1]) / (x[N - 3] - x[N - 1]);
    temp = temp * temp;
    g0 = y[N - 1];
    g1 = ((y[N - 2] - y[N - 1]) - temp * (y[N - 3] - y[N - 1])) /
         ((x[N - 2] - x[N - 1]) - temp * (x[N - 3] - x[N - 1]));
    g2 = ((y[N - 3] - y[N - 1]) - g1 * (x[N - 3] - x[N - 1])) /
         ((x[N - 3] - x[N - 1]) * (x[N - 3] - x[N - 1]));
    x4 = x[N - 2] + (x[N - 1] - x[N - 3]);
    x5 = x[N - 1] + (x[N - 1] - x[N - 3]);
    y4 = g0 + g1 * (x4 - x[N - 1]) + g2 * (x4 - x[N - 1]) * (x4 - x[N - 1]);
    y5 = g0 + g1 * (x5 - x[N - 1]) + g2 * (x5 - x[N - 1]) * (x5 - x[N - 1]);
    m1 = (y[N - 3] - y[N - 4]) / (x[N - 3] - x[N - 4]);
    m2 = (y[N - 2] - y[N - 3]) / (x[N - 2] - x[N - 3]);
    m3 = (y[N - 1] - y[N - 2]) / (x[N - 1] - x[N - 2]);
    m4 = (y4 - y[N - 1]) / (x4 - x[N - 1]);
    m5 = (y5 - y4) / (x5 - x4);
    t[N - 2] = akima_slope(m1, m2, m3, m4);
    t[N - 1] = akima_slope(m2, m3, m4, m5);
    for (int i = 2; i < N - 2; i++) {
        m1 = (y[i - 1] - y[i - 2]) / (x[i - 1] - x[i - 2]);
        m2 = (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        m3 = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        m4 = (y[i + 2] - y[i + 1]) / (x[i + 2] - x[i + 1]);
        t[i] = akima_slope(m1, m2, m3, m4);
    }
    for (int i = 0; i < N - 1; i++) {
        p0[i] = y[i];
        p1[i] = t[i];
        p2[i] = (3.0 * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - 2.0 * t[i] - t[i + 1]) / (x[i + 1] - x[i]);
        p3[i] = (t[i] + t[i + 1] - 2.0 * (y[i + 1] - y[i]) / (x[i + 1] - x[i])) /
                ((x[i + 1] - x[i]) * (x[i + 1] - x[i]));
    }
    p0[N - 1] = p0[N - 2]; p1[N - 1] = p1[N - 2]; p2[N - 1] = p2[N - 2]; p3[N - 1] = p3[N - 2];
    free(t);
}

/* Table from rows (r, e, f) on a uniform grid.  The grid spacing is the mean
 * spacing (outer - inner) / (N - 1) as in the reference's readRaw. */
static or_tab *tab_create(int itype, int N, const double *r, const double *e, const double *f,
                          double cutoff) {
    if (itype < 1 || itype > 3 || N < 2 || (itype == 2 && N < 4)) return NULL;
    or_tab *t = (or_tab *)calloc(1, sizeof(or_tab));
    t->set = 1; t->itype = itype; t->N = N;
    t->radius = (double *)malloc((size_t)N * sizeof(double));
    memcpy(t->radius, r, (size_t)N * sizeof(double));
    for (int k = 0; k < 4; ++k) {
        t->ce[k] = (double *)calloc((size_t)N, sizeof(double));
        t->cf[k] = (double *)calloc((size_t)N, sizeof(double));
    }
    t->inner = r[0];
    t->outer = r[N - 1];
    t->delta = (t->outer - t->inner) / (N - 1);
    t->invdelta = 1.0 / t->delta;
    t->deltasq6 = t->delta * t->delta / 6.0;
    t->cutoff = cutoff; t->cutoffSqr = cutoff * cutoff; t->shift = 0.0;
    if (itype == 1) {
        tab_linear(r, e, N, t->ce[0], t->ce[1]);
        tab_linear(r, f, N, t->cf[0], t->cf[1]);
    } else if (itype == 2) {
        tab_akima(r, e, N, t->ce[0], t->ce[1], t->ce[2], t->ce[3]);
        tab_akima(r, f, N, t->cf[0], t->cf[1], t->cf[2], t->cf[3]);
    } else {
        memcpy(t->ce[0], e, (size_t)N * sizeof(double));
        memcpy(t->cf[0], f, (size_t)N * sizeof(double));
        /* energy end slopes are minus the tabulated forces; force end slopes are secants */
        tab_cubic(r, e, N, -f[0], -f[N - 1], t->ce[1]);
        tab_cubic(r, f, N, (f[1] - f[0]) / (r[1] - r[0]), (f[N - 1] - f[N - 2]) / (r[N - 1] - r[N - 2]), t->cf[1]);
    }
    return t;
}

/* Interval index of r: truncation of (r - inner)/delta, clamped to the table's
 * intervals [0, N-2].  The reference clamps only to [0, N-1] and then reads
 * past its coefficient arrays for r at or beyond the last grid point; the last
 * interval is extended instead. */
static inline int tab_index(const or_tab *t, double r) {
    int index = (int)((r - t->inner) * t->invdelta);
    if (index < 0) index = 0;
    if (index > t->N - 2) index = t->N - 2;
    return index;
}

static double tab_eval(const or_tab *t, double *const c[4], double r) {
    int i = tab_index(t, r);
    if (t->itype == 1) return c[0][i] * r + c[1][i];
    if (t->itype == 2) {
        double z = r - t->radius[i];
        /* the reference stores z*z in an int (src/interaction/InterpolationAkima.hpp:105);
           kept so the values are the reference's */
        int zz2 = (int)(z * z);
        return c[0][i] + c[1][i] * z + c[2][i] * zz2 + c[3][i] * zz2 * z;
    }
    double b = (r - t->radius[i]) * t->invdelta;
    double a = 1.0 - b;
    return a * c[0][i] + b * c[0][i + 1] +
           ((a * a * a - a) * c[1][i] + (b * b * b - b) * c[1][i + 1]) * t->deltasq6;
}

double or_tab_energy_at(const or_tab *t, double r) { return tab_eval(t, (double *const *)t->ce, r); }
double or_tab_force_at(const or_tab *t, double r) { return tab_eval(t, (double *const *)t->cf, r); }

/* Tabulated::_computeForceRaw (src/interaction/Tabulated.hpp:100-115) behind
 * PotentialTemplate's cut test (distSqr > cutoffSqr -> no force). */
static inline int tab_force(const or_tab *t, const double d[3], double F[3]) {
    if (!t || !t->set) return 0;
    double distSqr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (distSqr > t->cutoffSqr) return 0;
    double distrt = sqrt(distSqr);
    double ffactor = or_tab_force_at(t, distrt);
    ffactor /= distrt;
    F[0] = d[0] * ffactor; F[1] = d[1] * ffactor; F[2] = d[2] * ffactor;
    return 1;
}

static inline double tab_energy(const or_tab *t, double distSqr) {
    if (!t || !t->set || distSqr > t->cutoffSqr) return 0.0;
    return or_tab_energy_at(t, sqrt(distSqr)) - t->shift;
}

/* stand-alone table for known-answer checks (handle freed with or_tab_destroy) */
or_tab *or_tab_new(int itype, int N, const double *r, const double *e, const double *f, double cutoff) {
    return tab_create(itype, N, r, e, f, cutoff);
}
void or_tab_destroy(or_tab *t) { tab_free(t); }
void or_tab_eval(const or_tab *t, int n, const double *r, double *e, double *f) {
    for (int k = 0; k < n; ++k) { e[k] = or_tab_energy_at(t, r[k]); f[k] = or_tab_force_at(t, r[k]); }
}

int or_ia_tab_create(or_sim *s, int vl) {
    if (s->nia >= OR_MAXIA) { or_seterr(s, "too many interactions"); return -1; }
    or_ia *a = (or_ia *)calloc(1, sizeof(or_ia));
    a->kind = 2; a->list = vl;
    s->ia[s->nia] = a;
    return s->nia++;
}

/* VerletListTabulated::setPotential(type1, type2, Tabulated(itype, file, cutoff)) */
int or_ia_tab_set(or_sim *s, int ii, int t1, int t2, int itype, int N, const double *r, const double *e,
                  const double *f, double cutoff) {
    if (t1 < 0 || t2 < 0 || t1 >= OR_MAXT || t2 >= OR_MAXT) { or_seterr(s, "type out of range"); return -1; }
    or_tab *t = tab_create(itype, N, r, e, f, cutoff);
    if (!t) { or_seterr(s, "illegal table"); return -1; }
    or_ia *a = s->ia[ii];
    or_tab *old1 = a->tab[t1 * OR_MAXT + t2], *old2 = a->tab[t2 * OR_MAXT + t1];
    a->tab[t1 * OR_MAXT + t2] = t;
    a->tab[t2 * OR_MAXT + t1] = t;
    int used = 0;
    for (int k = 0; k < OR_MAXT * OR_MAXT; ++k) used |= (a->tab[k] == old1);
    if (old1 && !used) tab_free(old1);
    used = 0;
    for (int k = 0; k < OR_MAXT * OR_MAXT; ++k) used |= (a->tab[k] == old2);
    if (old2 && old2 != old1 && !used) tab_free(old2);
    return 0;
}

/* FixedPairListTabulated(system, fpl, Tabulated(itype, file, cutoff)) */
int or_ia_tabbond_create(or_sim *s, int fpl, int itype, int N, const double *r, const double *e,
                         const double *f, double cutoff) {
    if (s->nia >= OR_MAXIA) { or_seterr(s, "too many interactions"); return -1; }
    or_tab *t = tab_create(itype, N, r, e, f, cutoff);
    if (!t) { or_seterr(s, "illegal table"); return -1; }
    or_ia *a = (or_ia *)calloc(1, sizeof(or_ia));
    a->kind = 3; a->list = fpl;
    a->tab[0] = t;
    s->ia[s->nia] = a;
    return s->nia++;
}

/* OrthorhombicBC::getMinimumImageVectorBox */
static inline void min_image_box(const or_sim *s, const double *p1, const double *p2, double d[3]) {
    for (int k = 0; k < 3; ++k) {
        d[k] = p1[k] - p2[k];
        if (d[k] < -s->halfL[k]) d[k] += s->L[k];
        else if (d[k] > s->halfL[k]) d[k] -= s->L[k];
    }
}

static inline void pair_dist(const or_sim *s, const or_vl *v, int64_t k, double d[3]) {
    const double *x1 = &s->x[3 * v->p1[k]], *x2 = &s->x[3 * v->p2[k]];
    const int8_t *sh = &v->sh[3 * k];
    for (int c = 0; c < 3; ++c) d[c] = x1[c] - (x2[c] + sh[c] * s->L[c]);
}

/* Pair force of a VerletList interaction: LennardJones (kind 0) or Tabulated (kind 2). */
static inline int vl_pair_force(const or_sim *s, const or_ia *a, int32_t i, int32_t j, const double d[3],
                                double F[3]) {
    const int tp = s->type[i] * OR_MAXT + s->type[j];
    if (a->kind == 0) {
        const or_lj *p = &a->lj[tp];
        return p->set ? lj_force(p, d, F) : 0;
    }
    return tab_force(a->tab[tp], d, F);
}

static inline double vl_pair_energy(const or_sim *s, const or_ia *a, int32_t i, int32_t j, double d2) {
    const int tp = s->type[i] * OR_MAXT + s->type[j];
    if (a->kind == 0) {
        const or_lj *p = &a->lj[tp];
        if (!p->set) return 0.0;
        return (d2 > p->rc2) ? 0.0 : lj_energy_raw(p, d2) - p->shift;
    }
    return tab_energy(a->tab[tp], d2);
}

static inline int bond_force(const or_ia *a, const double d[3], double F[3]) {
    return a->kind == 1 ? fene_force(a, d, F) : tab_force(a->tab[0], d, F);
}

static inline double bond_energy(const or_ia *a, double d2) {
    if (a->kind == 1) return (d2 > a->cutoffSqr) ? 0.0 : fene_energy_raw(a, d2) - a->shift;
    return tab_energy(a->tab[0], d2);
}

/* Pair interactions over the half list with Newton-3.  Cells are processed in z-layer phases
 * (even layers, odd layers, then a lone last layer when nz is odd) so that
 * threads never write the same particle; the per-particle summation order is
 * the same for every thread count. */
static void lj_add_forces(or_sim *s, const or_ia *a) {
    const or_vl *v = &s->vl[a->list];
    int nz = s->n[2];
    int64_t per_layer = (int64_t)s->n[0] * s->n[1];
    int nph = (nz % 2) ? 3 : 2;
    for (int ph = 0; ph < nph; ++ph) {
        int zfirst = ph == 2 ? nz - 1 : ph;
        int zstep = ph == 2 ? 1 : 2;
        int zend = (ph < 2 && (nz % 2)) ? nz - 1 : nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
        for (int z = zfirst; z < zend; z += zstep) {
            for (int64_t c = z * per_layer; c < (z + 1) * per_layer; ++c) {
                for (int64_t k = v->cell_pstart[c]; k < v->cell_pstart[c + 1]; ++k) {
                    int32_t i = v->p1[k], j = v->p2[k];
                    double d[3], F[3];
                    pair_dist(s, v, k, d);
                    if (vl_pair_force(s, a, i, j, d, F)) {
                        for (int c3 = 0; c3 < 3; ++c3) { s->f[3 * i + c3] += F[c3]; s->f[3 * j + c3] -= F[c3]; }
                    }
                }
            }
        }
    }
}

static int fene_bond_dist(or_sim *s, const or_fpl *p, int64_t k, int32_t *i, int32_t *j, double d[3]) {
    int32_t a = p->a[k], b = p->b[k];
    if (a >= s->nid || b >= s->nid || s->loc_of_id[a] < 0 || s->loc_of_id[b] < 0) return 0;
    *i = s->loc_of_id[a]; *j = s->loc_of_id[b];
    min_image_box(s, &s->x[3 * *i], &s->x[3 * *j], d);
    return 1;
}

static void fene_add_forces(or_sim *s, const or_ia *a) {
    const or_fpl *p = &s->fpl[a->list];
    for (int64_t k = 0; k < p->n; ++k) {
        int32_t i, j; double d[3], F[3];
        if (!fene_bond_dist(s, p, k, &i, &j, d)) continue;
        if (bond_force(a, d, F))
            for (int c = 0; c < 3; ++c) { s->f[3 * i + c] += F[c]; s->f[3 * j + c] -= F[c]; }
    }
}

static void thermalize(or_sim *s) {
    uint64_t call = s->therm_calls++;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int64_t i = 0; i < s->np; ++i) {
        int32_t pid = s->id[i];
        if (s->nlang_excl) {
            int64_t l = 0, r = s->nlang_excl - 1; int hit = 0;
            while (l <= r) { int64_t mid = (l + r) / 2; if (s->lang_excl[mid] == pid) { hit = 1; break; } if (s->lang_excl[mid] < pid) l = mid + 1; else r = mid - 1; }
            if (hit) continue;
        }
        double u[3];
        langevin_uniforms(s->seed, call, pid, u);
        double mass = s->m[i], massf = sqrt(mass);
        for (int c = 0; c < 3; ++c) {
            double ran = u[c] - 0.5;
            s->f[3 * i + c] += (s->pref1 * s->v[3 * i + c]) * mass + (s->pref2 * ran) * massf;
        }
    }
}

/* VelocityVerlet::calcForces + updateForces + aftCalcF (thermostat) */
static void update_forces(or_sim *s) {
    memset(s->f, 0, (size_t)s->np * 3 * sizeof(double));
    for (int k = 0; k < s->nia; ++k) {
        if (s->ia[k]->kind == 0 || s->ia[k]->kind == 2) lj_add_forces(s, s->ia[k]);
        else fene_add_forces(s, s->ia[k]);
    }
    if (s->lang_on) thermalize(s);
}

/* forces of all interactions, no thermostat */
int or_calc_forces(or_sim *s) {
    memset(s->f, 0, (size_t)s->np * 3 * sizeof(double));
    for (int k = 0; k < s->nia; ++k) {
        if (s->ia[k]->kind == 0 || s->ia[k]->kind == 2) lj_add_forces(s, s->ia[k]);
        else fene_add_forces(s, s->ia[k]);
    }
    return 0;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return x < y ? -1 : x > y;
}

int or_langevin(or_sim *s, int on, double gamma, double T, int64_t nexcl, const int32_t *excl) {
    s->lang_on = on; s->gamma = gamma; s->temperature = T;
    if (nexcl > 0) {
        s->lang_excl = (int32_t *)realloc(s->lang_excl, (size_t)(s->nlang_excl + nexcl) * sizeof(int32_t));
        memcpy(s->lang_excl + s->nlang_excl, excl, (size_t)nexcl * sizeof(int32_t));
        s->nlang_excl += nexcl;
        qsort(s->lang_excl, (size_t)s->nlang_excl, sizeof(int32_t), cmp_i32);
    }
    return 0;
}

void or_set_therm_calls(or_sim *s, uint64_t c) { s->therm_calls = c; }
uint64_t or_get_therm_calls(or_sim *s) { return s->therm_calls; }

static double integrate1(or_sim *s) {
    double maxSq = 0.0;
    double dt = s->dt;
#pragma omp parallel for schedule(static) reduction(max : maxSq) num_threads(s->nthreads)
    for (int64_t i = 0; i < s->np; ++i) {
        double dtfm = 0.5 * dt / s->m[i];
        double *v = &s->v[3 * i], *x = &s->x[3 * i];
        const double *f = &s->f[3 * i];
        double dp[3];
        for (int c = 0; c < 3; ++c) v[c] += dtfm * f[c];
        for (int c = 0; c < 3; ++c) { dp[c] = v[c]; dp[c] *= dt; x[c] += dp[c]; }
        double sq = 0.0;
        sq += dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        if (sq > maxSq) maxSq = sq;
    }
    return sqrt(maxSq);
}

static void integrate2(or_sim *s) {
    double half_dt = 0.5 * s->dt;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int64_t i = 0; i < s->np; ++i) {
        double dtfm = half_dt / s->m[i];
        for (int c = 0; c < 3; ++c) s->v[3 * i + c] += dtfm * s->f[3 * i + c];
    }
    s->step++;
}

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* VelocityVerlet::run (src/integrator/VelocityVerlet.cpp:63-169) */
int or_run(or_sim *s, int nsteps) {
    if (s->dt == 0.0) { or_seterr(s, "timestep must be non-zero"); return -5; }
    double t0 = now_s();
    double skinHalf = 0.5 * s->skin;
    s->nResorts = 0;
    if (s->lang_on) { s->pref1 = -s->gamma; s->pref2 = sqrt(24.0 * s->temperature * s->gamma / s->dt); }
    if (s->resortFlag) {
        int rc = or_decompose(s); if (rc) return rc;
        s->maxDist = 0.0; s->resortFlag = 0;
    }
    double pref2buffer = s->pref2;
    s->pref2 *= sqrt(3.0);
    update_forces(s);
    s->pref2 = pref2buffer;
    for (int i = 0; i < nsteps; ++i) {
        double ta = now_s();
        s->maxDist += integrate1(s);
        double tb = now_s();
        s->t_int1 += tb - ta;
        if (s->maxDist > skinHalf) s->resortFlag = 1;
        if (s->resortFlag) {
            int rc = or_decompose(s); if (rc) return rc;
            s->maxDist = 0.0; s->resortFlag = 0; s->nResorts++;
        }
        double tc = now_s();
        s->t_resort += tc - tb;
        update_forces(s);
        double td = now_s();
        s->t_force += td - tc;
        integrate2(s);
        s->t_int2 += now_s() - td;
    }
    s->t_run += now_s() - t0;
    return 0;
}

/* ---------------------------------------------------------------- observables */
double or_ia_energy(or_sim *s, int ii) {
    const or_ia *a = s->ia[ii];
    double es = 0.0;
    if (a->kind == 0 || a->kind == 2) {
        const or_vl *v = &s->vl[a->list];
        for (int64_t k = 0; k < v->npairs; ++k) {
            double d[3]; pair_dist(s, v, k, d);
            double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            es += vl_pair_energy(s, a, v->p1[k], v->p2[k], d2);
        }
    } else {
        const or_fpl *p = &s->fpl[a->list];
        for (int64_t k = 0; k < p->n; ++k) {
            int32_t i, j; double d[3];
            if (!fene_bond_dist(s, p, k, &i, &j, d)) continue;
            double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            es += bond_energy(a, d2);
        }
    }
    return es;
}

/* w6 = xx, yy, zz, xy, xz, yz (Tensor order, src/Tensor.hpp) */
double or_ia_virial(or_sim *s, int ii, double *w6) {
    const or_ia *a = s->ia[ii];
    double w = 0.0, t[6] = {0, 0, 0, 0, 0, 0};
    if (a->kind == 0 || a->kind == 2) {
        const or_vl *v = &s->vl[a->list];
        for (int64_t k = 0; k < v->npairs; ++k) {
            double d[3], F[3]; pair_dist(s, v, k, d);
            if (vl_pair_force(s, a, v->p1[k], v->p2[k], d, F)) {
                w = w + (d[0] * F[0] + d[1] * F[1] + d[2] * F[2]);
                t[0] += d[0] * F[0]; t[1] += d[1] * F[1]; t[2] += d[2] * F[2];
                t[3] += d[0] * F[1]; t[4] += d[0] * F[2]; t[5] += d[1] * F[2];
            }
        }
    } else {
        const or_fpl *p = &s->fpl[a->list];
        for (int64_t k = 0; k < p->n; ++k) {
            int32_t i, j; double d[3], F[3];
            if (!fene_bond_dist(s, p, k, &i, &j, d)) continue;
            if (bond_force(a, d, F)) {
                w = w + (d[0] * F[0] + d[1] * F[1] + d[2] * F[2]);
                t[0] += d[0] * F[0]; t[1] += d[1] * F[1]; t[2] += d[2] * F[2];
                t[3] += d[0] * F[1]; t[4] += d[0] * F[2]; t[5] += d[1] * F[2];
            }
        }
    }
    if (w6) memcpy(w6, t, sizeof(t));
    return w;
}

/* RadialDistrF::computeArray (src/analysis/RadialDistrF.cpp:47-140): all pairs,
 * minimum image by one +-L correction per component, bin = (int)(|d| / dr) with
 * dr = (L_y / 2) / nbins, normalised by 2 pi dr rho N (r^2 + dr^2 / 12), the
 * prefactor formed in long double as the reference does with M_PIl. */
int or_rdf(or_sim *s, int nbins, double *out) {
    if (nbins <= 0) { or_seterr(s, "nbins must be positive"); return -1; }
    double *hist = (double *)calloc((size_t)nbins, sizeof(double));
    if (!hist) { or_seterr(s, "out of memory"); return -1; }
    const double dr = s->halfL[1] / (double)nbins;
    const int64_t n = s->np;
#pragma omp parallel num_threads(s->nthreads)
    {
        double *h = (double *)calloc((size_t)nbins, sizeof(double));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            const double *p1 = &s->x[3 * i];
            for (int64_t j = i + 1; j < n; ++j) {
                const double *p2 = &s->x[3 * j];
                double d[3];
                for (int c = 0; c < 3; ++c) {
                    d[c] = p1[c] - p2[c];
                    if (d[c] < -s->halfL[c]) d[c] += s->L[c];
                    if (d[c] > s->halfL[c]) d[c] -= s->L[c];
                }
                double q = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / dr;
                if (q < (double)nbins) h[(int)q] += 1.0;
            }
        }
#pragma omp critical
        for (int b = 0; b < nbins; ++b) hist[b] += h[b];
        free(h);
    }
    const double rho = (double)n / (s->L[0] * s->L[1] * s->L[2]);
    const double factor = (double)(2.0L * 3.141592653589793238462643383279502884L * (long double)dr *
                                   (long double)rho * 1.0L * (long double)n);
    for (int b = 0; b < nbins; ++b) {
        double radius = (b + 0.5) * dr;
        out[b] = hist[b] / (factor * (radius * radius + dr * dr / 12.0));
    }
    free(hist);
    return 0;
}

double or_temperature(or_sim *s) {
    double v2sum = 0.0;
    for (int64_t i = 0; i < s->np; ++i) {
        const double *v = &s->v[3 * i];
        v2sum += s->m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    return v2sum / (3.0 * (double)s->np);
}

double or_pressure(or_sim *s) {
    double tripleV = 3.0 * s->L[0] * s->L[1] * s->L[2];
    double v2 = 0.0;
    for (int64_t i = 0; i < s->np; ++i) {
        const double *v = &s->v[3 * i];
        v2 += s->m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double w = 0.0;
    for (int k = 0; k < s->nia; ++k) w += or_ia_virial(s, k, NULL);
    return (v2 + w) / tripleV;
}

/* ---------------------------------------------------------------- accessors */
/* field: 0 pos, 1 vel, 2 force, 3 image(int as double) ; output in ascending id order */
int or_get_by_id(or_sim *s, int field, double *out) {
    int64_t o = 0;
    for (int64_t pid = 0; pid < s->nid; ++pid) {
        int32_t i = s->loc_of_id[pid];
        if (i < 0) continue;
        for (int c = 0; c < 3; ++c) {
            double val = field == 0 ? s->x[3 * i + c] : field == 1 ? s->v[3 * i + c]
                       : field == 2 ? s->f[3 * i + c] : (double)s->img[3 * i + c];
            out[3 * o + c] = val;
        }
        o++;
    }
    return 0;
}

int or_get_ids(or_sim *s, int32_t *out) {
    int64_t o = 0;
    for (int64_t pid = 0; pid < s->nid; ++pid) if (s->loc_of_id[pid] >= 0) out[o++] = (int32_t)pid;
    return 0;
}

/* field: 0 pos, 1 vel; input in ascending id order */
int or_set_by_id(or_sim *s, int field, const double *in) {
    int64_t o = 0;
    for (int64_t pid = 0; pid < s->nid; ++pid) {
        int32_t i = s->loc_of_id[pid];
        if (i < 0) continue;
        for (int c = 0; c < 3; ++c) {
            if (field == 0) s->x[3 * i + c] = in[3 * o + c];
            else s->v[3 * i + c] = in[3 * o + c];
        }
        o++;
    }
    return 0;
}

void or_timers(or_sim *s, double *t5) {
    t5[0] = s->t_run; t5[1] = s->t_force; t5[2] = s->t_int1; t5[3] = s->t_int2; t5[4] = s->t_resort;
}
