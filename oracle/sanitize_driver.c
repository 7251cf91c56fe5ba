/*
 * Sanitizer driver of the FP64 oracle (test infrastructure only): built together
 * with espp_oracle.c under AddressSanitizer and UndefinedBehaviorSanitizer by
 * `make -C oracle sanitize`, it runs the oracle's whole step on small systems --
 * a jittered LJ lattice with Langevin, a bonded WCA + FENE chain system with
 * exclusions, tabulated pair and bond tables, the RDF and the observables -- so
 * that an out-of-bounds access, a leak or undefined arithmetic in the oracle
 * aborts the run.  Exit status 0 means a clean run.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct or_sim or_sim;
or_sim *or_create(const double *L, const int *cells, double skin, uint64_t seed);
void or_destroy(or_sim *s);
void or_set_threads(or_sim *s, int n);
void or_set_dt(or_sim *s, double dt);
int or_add_particles(or_sim *s, int64_t n, const int32_t *id, const int32_t *type, const double *pos,
                     const double *vel, const double *mass);
int or_decompose(or_sim *s);
int or_vl_create(or_sim *s, double cutoff);
int or_vl_exclude(or_sim *s, int vi, int64_t n, const int32_t *pairs);
int64_t or_vl_size(or_sim *s, int vi);
int or_fpl_create(or_sim *s);
int or_fpl_add(or_sim *s, int li, int64_t n, const int32_t *pairs);
int or_ia_lj_create(or_sim *s, int vl);
int or_ia_lj_set(or_sim *s, int ii, int t1, int t2, double eps, double sig, double rc, double shift);
int or_ia_fene_create(or_sim *s, int fpl, double K, double r0, double rMax, double cutoff, double shift);
int or_ia_tab_create(or_sim *s, int vl);
int or_ia_tab_set(or_sim *s, int ii, int t1, int t2, int itype, int N, const double *r, const double *e,
                  const double *f, double cutoff);
int or_ia_tabbond_create(or_sim *s, int fpl, int itype, int N, const double *r, const double *e, const double *f,
                         double cutoff);
int or_langevin(or_sim *s, int on, double gamma, double T, int64_t nexcl, const int32_t *excl);
int or_run(or_sim *s, int nsteps);
int or_calc_forces(or_sim *s);
double or_ia_energy(or_sim *s, int ii);
double or_ia_virial(or_sim *s, int ii, double *w6);
int or_rdf(or_sim *s, int nbins, double *out);
double or_temperature(or_sim *s);
double or_pressure(or_sim *s);

static double uni(uint64_t *x) {
  *x = *x * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(*x >> 11) * (1.0 / 9007199254740992.0);
}

static int fail(const char *what) {
  fprintf(stderr, "sanitize driver: %s failed\n", what);
  return 1;
}

static int lj_fluid(void) {
  const int a = 8, n = a * a * a;
  const double L = cbrt(n / 0.8442);
  const double box[3] = {L, L, L};
  const int cells[3] = {(int)(L / 2.8), (int)(L / 2.8), (int)(L / 2.8)};
  double *pos = malloc(sizeof(double) * 3 * n), *vel = malloc(sizeof(double) * 3 * n);
  int32_t *id = malloc(sizeof(int32_t) * n), *type = malloc(sizeof(int32_t) * n);
  uint64_t seed = 7;
  for (int k = 0; k < n; ++k) {
    const int i = k % a, j = (k / a) % a, l = k / (a * a);
    pos[3 * k] = (i + 0.5) * L / a + 0.1 * (uni(&seed) - 0.5);
    pos[3 * k + 1] = (j + 0.5) * L / a + 0.1 * (uni(&seed) - 0.5);
    pos[3 * k + 2] = (l + 0.5) * L / a + 0.1 * (uni(&seed) - 0.5);
    for (int c = 0; c < 3; ++c) vel[3 * k + c] = uni(&seed) - 0.5;
    id[k] = k;
    type[k] = k % 2;
  }
  or_sim *s = or_create(box, cells, 0.3, 11);
  or_set_threads(s, 2);
  if (or_add_particles(s, n, id, type, pos, vel, NULL)) return fail("add");
  if (or_decompose(s)) return fail("decompose");
  const int vl = or_vl_create(s, 2.5);
  const int ia = or_ia_lj_create(s, vl);
  or_ia_lj_set(s, ia, 0, 0, 1.0, 1.0, 2.5, NAN);
  or_ia_lj_set(s, ia, 0, 1, 0.5, 0.9, 2.5, NAN);
  or_ia_lj_set(s, ia, 1, 1, 1.2, 1.1, 2.5, 0.0);
  or_langevin(s, 1, 1.0, 1.0, 0, NULL);
  or_set_dt(s, 0.005);
  if (or_run(s, 200)) return fail("run");
  double w6[6], rdf[50];
  const double e = or_ia_energy(s, ia), w = or_ia_virial(s, ia, w6);
  if (or_rdf(s, 50, rdf)) return fail("rdf");
  const double T = or_temperature(s), P = or_pressure(s);
  printf("lj: pairs %lld, E %.6f, W %.6f, T %.4f, P %.4f\n", (long long)or_vl_size(s, vl), e, w, T, P);
  or_destroy(s);
  free(pos); free(vel); free(id); free(type);
  return !(isfinite(e) && isfinite(w) && isfinite(T) && isfinite(P));
}

static int chains(void) {
  const int nc = 20, nb = 16, n = nc * nb;
  const double L = cbrt(n / 0.3);
  const double box[3] = {L, L, L};
  const int g = (int)(L / 1.4);
  const int cells[3] = {g, g, g};
  double *pos = malloc(sizeof(double) * 3 * n);
  int32_t *id = malloc(sizeof(int32_t) * n), *bonds = malloc(sizeof(int32_t) * 2 * n);
  uint64_t seed = 3;
  int nbond = 0;
  for (int c = 0; c < nc; ++c) {
    double x = uni(&seed) * L, y = uni(&seed) * L, z = uni(&seed) * L;
    for (int b = 0; b < nb; ++b) {
      const int k = c * nb + b;
      if (b) {
        const double u = 2 * uni(&seed) - 1, ph = 2 * M_PI * uni(&seed), rr = sqrt(1 - u * u);
        x += 0.97 * rr * cos(ph); y += 0.97 * rr * sin(ph); z += 0.97 * u;
        bonds[2 * nbond] = k - 1; bonds[2 * nbond + 1] = k; ++nbond;
      }
      pos[3 * k] = x; pos[3 * k + 1] = y; pos[3 * k + 2] = z;
      id[k] = k;
    }
  }
  or_sim *s = or_create(box, cells, 0.3, 5);
  if (or_add_particles(s, n, id, NULL, pos, NULL, NULL)) return fail("add chains");
  if (or_decompose(s)) return fail("decompose chains");
  const double rc = pow(2.0, 1.0 / 6.0);
  const int vl = or_vl_create(s, rc);
  or_vl_exclude(s, vl, nbond, bonds);
  const int fpl = or_fpl_create(s);
  or_fpl_add(s, fpl, nbond, bonds);
  const int fene = or_ia_fene_create(s, fpl, 30.0, 0.0, 1.5, INFINITY, 0.0);
  // a tabulated soft repulsion over the list and a tabulated bond beside FENE
  enum { N = 64 };
  double r[N], e[N], f[N], rb[N], eb[N], fb[N];
  for (int k = 0; k < N; ++k) {
    r[k] = rc * k / (N - 1);
    e[k] = 5.0 * (1 + cos(M_PI * r[k] / rc));
    f[k] = 5.0 * M_PI / rc * sin(M_PI * r[k] / rc);
    rb[k] = 0.5 + 1.0 * k / (N - 1);
    eb[k] = 2.0 * (rb[k] - 0.97) * (rb[k] - 0.97);
    fb[k] = -4.0 * (rb[k] - 0.97);
  }
  const int tab = or_ia_tab_create(s, vl);
  for (int it = 1; it <= 3; ++it) or_ia_tab_set(s, tab, 0, 0, it, N, r, e, f, rc);
  const int tb = or_ia_tabbond_create(s, fpl, 2, N, rb, eb, fb, 1.5);
  or_langevin(s, 1, 0.5, 1.0, 0, NULL);
  or_set_dt(s, 0.002);
  if (or_run(s, 100)) return fail("run chains");
  const double ef = or_ia_energy(s, fene), et = or_ia_energy(s, tab), eb2 = or_ia_energy(s, tb);
  printf("chains: E_fene %.6f, E_tab %.6f, E_bondtab %.6f\n", ef, et, eb2);
  or_destroy(s);
  free(pos); free(id); free(bonds);
  return !(isfinite(ef) && isfinite(et) && isfinite(eb2));
}

int main(void) {
  int bad = lj_fluid();
  bad |= chains();
  if (!bad) printf("sanitize driver: clean\n");
  return bad;
}
