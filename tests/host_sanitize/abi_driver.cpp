// Drives the library's C ABI end to end with its host code built under
// AddressSanitizer and UndefinedBehaviorSanitizer (host side only; the device
// code is built as usual).  Test infrastructure: tests/test_gpu_host_sanitized.py
// runs it on a GPU, and any invalid access, leak of the library's own host
// allocations or undefined arithmetic in the host code aborts it.
//
// Covered: handle life cycle, particles in and out (by slot and by id), Verlet
// lists with exclusions, LJ on both precisions and several types, FENE bonds,
// tabulated pair and bond potentials, Langevin runs, energies, virials, pairs,
// the RDF, removal and replacement of interactions, the allocation-free energy,
// checkpoint state, pruned lists, the learned pair force (evaluation, training,
// Adam), the batched replicas (load, autotune, run, energies, pairs), and the
// error paths a caller can hit (NULL arguments, bad handles, NaN positions).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "espp_amd.h"
#include "espp_amd_prof.h"

namespace {

int failures = 0;

void check(int rc, const char* what) {
  if (rc != ESPP_OK) {
    std::fprintf(stderr, "FAIL %s: %d %s\n", what, rc, espp_last_error());
    ++failures;
  }
}

void expect_error(int rc, const char* what) {
  if (rc == ESPP_OK) {
    std::fprintf(stderr, "FAIL %s: expected an error\n", what);
    ++failures;
  }
}

#define CHECK(x) check((x), #x)
#define EXPECT_ERROR(x) expect_error((x), #x)

// Simple-cubic lattice of n sites at density rho, jittered by a fixed LCG.
void lattice(int n, double rho, std::vector<double>& pos, double& L) {
  L = std::cbrt(n / rho);
  int a = (int)std::lround(std::cbrt((double)n));
  if (a * a * a < n) ++a;
  const double sp = L / a;
  unsigned long long s = 12345;
  auto uni = [&] {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  pos.resize((size_t)3 * n);
  for (int k = 0; k < n; ++k) {
    const int i = k / (a * a), j = (k / a) % a, l = k % a;
    pos[3 * k] = (i + 0.5) * sp + 0.1 * (uni() - 0.5);
    pos[3 * k + 1] = (j + 0.5) * sp + 0.1 * (uni() - 0.5);
    pos[3 * k + 2] = (l + 0.5) * sp + 0.1 * (uni() - 0.5);
  }
}

void lj_fluid(int precision) {
  const int n = 4000;
  std::vector<double> pos, vel((size_t)3 * n), mass((size_t)n, 1.0);
  double L;
  lattice(n, 0.8442, pos, L);
  for (int k = 0; k < 3 * n; ++k) vel[(size_t)k] = 0.5 * std::sin(0.37 * k);
  std::vector<int32_t> id((size_t)n), type((size_t)n);
  for (int k = 0; k < n; ++k) { id[(size_t)k] = 3 * k + 1; type[(size_t)k] = k % 2; }
  const double box[3] = {L, L, L};
  const double rc = 2.5, skin = 0.3;
  const int32_t cg = (int32_t)(L / (rc + skin));
  const int32_t cells[3] = {cg, cg, cg};
  espp_sim* s = nullptr;
  CHECK(espp_create(box, cells, 1, skin, 7, precision, &s));
  if (!s) return;
  CHECK(espp_add_particles(s, n, id.data(), type.data(), pos.data(), vel.data(), mass.data()));
  CHECK(espp_decompose(s));
  int32_t vl = -1, lj = -1, fpl = -1, fene = -1, tab = -1, tabb = -1;
  CHECK(espp_vl_create(s, rc, &vl));
  const int32_t excl[4] = {1, 4, 7, 10};
  CHECK(espp_vl_exclude(s, vl, 2, excl));
  CHECK(espp_lj_create(s, vl, &lj));
  CHECK(espp_lj_set(s, lj, 0, 0, 1.0, 1.0, rc, NAN));
  CHECK(espp_lj_set(s, lj, 0, 1, 1.1, 1.05, rc, NAN));
  CHECK(espp_lj_set(s, lj, 1, 1, 0.9, 0.95, rc, 0.0));
  CHECK(espp_add_interaction(s, lj));
  CHECK(espp_fpl_create(s, &fpl));
  CHECK(espp_fpl_add(s, fpl, 2, excl));
  CHECK(espp_fene_create(s, fpl, 30.0, 0.0, 1.5, INFINITY, 0.0, &fene));
  CHECK(espp_add_interaction(s, fene));
  CHECK(espp_langevin(s, 1, 1.0, 1.0, 1, &id[5]));
  CHECK(espp_run(s, 40, 0.005));
  double e = 0, w = 0, t6[6];
  CHECK(espp_ia_energy(s, lj, &e));
  CHECK(espp_ia_virial(s, lj, &w, t6));
  CHECK(espp_ia_energy_noalloc(s, lj, &e, &w, t6));
  int64_t np = 0, got = 0;
  CHECK(espp_vl_size(s, vl, &np));
  std::vector<int32_t> pairs((size_t)2 * (size_t)(np + 1));
  CHECK(espp_vl_pairs(s, vl, pairs.data(), np, &got));
  std::vector<double> out((size_t)3 * n), rdf(64);
  for (int f : {ESPP_FIELD_POS, ESPP_FIELD_VEL, ESPP_FIELD_FORCE, ESPP_FIELD_IMAGE}) CHECK(espp_get_particles(s, f, out.data()));
  CHECK(espp_get_particles_by_id(s, ESPP_FIELD_POS, 3, &id[10], out.data()));
  CHECK(espp_get_particles_by_id(s, ESPP_FIELD_VEL, 2, &id[20], out.data()));
  CHECK(espp_set_particles(s, ESPP_FIELD_VEL, 2, &id[20], out.data()));
  CHECK(espp_rdf(s, 64, rdf.data()));
  double mv2 = 0;
  CHECK(espp_kinetic(s, &mv2));
  // tabulated LJ-like table over the same list, and a tabulated bond
  std::vector<double> r(200), te(200), tf(200);
  for (int k = 0; k < 200; ++k) {
    r[(size_t)k] = 0.8 + k * (rc - 0.8) / 199.0;
    const double x = 1.0 / std::pow(r[(size_t)k], 6);
    te[(size_t)k] = 4.0 * (x * x - x);
    tf[(size_t)k] = 24.0 * (2.0 * x * x - x) / r[(size_t)k];
  }
  CHECK(espp_tab_create(s, vl, &tab));
  for (int it = 1; it <= 3; ++it) CHECK(espp_tab_set(s, tab, 0, 0, it, 200, r.data(), te.data(), tf.data(), rc));
  CHECK(espp_tabbond_create(s, fpl, 2, 200, r.data(), te.data(), tf.data(), rc, &tabb));
  CHECK(espp_add_interaction(s, tab));
  CHECK(espp_add_interaction(s, tabb));
  CHECK(espp_run(s, 10, 0.002));
  CHECK(espp_remove_interaction(s, 3));
  CHECK(espp_remove_interaction(s, 2));
  CHECK(espp_fene_set(s, fene, 25.0, 0.0, 1.6, INFINITY, 0.0));
  // pruned lists, then windows of steps, on the mixed path (the hot start from the
  // doubled LJ of the tables needs the smaller time step); checkpoint state
  CHECK(espp_set_prune_shell(s, -1.0));
  CHECK(espp_run(s, 30, 0.002));
  int64_t prunes = 0, walks = 0, entries = 0;
  CHECK(espp_prune_stats(s, lj, &prunes, &walks, &entries));
  CHECK(espp_set_prune_shell(s, 0.0));
  CHECK(espp_set_step_window(s, 8));
  CHECK(espp_run(s, 30, 0.002));
  int64_t windows = 0, launched = 0, ran = 0;
  CHECK(espp_step_window_stats(s, &windows, &launched, &ran));
  EXPECT_ERROR(espp_set_step_window(s, -1));
  uint64_t seed = 0, calls = 0;
  double maxdist = 0;
  int32_t resort = 0;
  CHECK(espp_get_rng_state(s, &seed, &calls));
  CHECK(espp_set_rng_state(s, seed, calls));
  CHECK(espp_get_run_state(s, &maxdist, &resort));
  CHECK(espp_set_run_state(s, maxdist, resort));
  double timers[10];
  CHECK(espp_timers(s, timers));
  // error paths
  EXPECT_ERROR(espp_vl_pairs(s, vl, nullptr, 10, &got));
  EXPECT_ERROR(espp_get_particles_by_id(s, ESPP_FIELD_POS, 1, &id[0] + 0, nullptr));
  const int32_t missing = 2;   // ids are 3k + 1
  EXPECT_ERROR(espp_get_particles_by_id(s, ESPP_FIELD_POS, 1, &missing, out.data()));
  EXPECT_ERROR(espp_lj_set(s, 99, 0, 0, 1.0, 1.0, rc, 0.0));
  EXPECT_ERROR(espp_run(s, -1, 0.005));
  const double nan3[3] = {NAN, 1.0, 1.0};
  CHECK(espp_set_particles(s, ESPP_FIELD_POS, 1, &id[3], nan3));
  EXPECT_ERROR(espp_run(s, 1, 0.005));
  CHECK(espp_destroy(s));
}

void learned_force() {
  const int n = 3000;
  std::vector<double> pos;
  double L;
  lattice(n, 0.2, pos, L);
  const double box[3] = {L, L, L};
  const int32_t cg = (int32_t)(L / 3.3);
  const int32_t cells[3] = {cg, cg, cg};
  espp_sim* s = nullptr;
  CHECK(espp_create(box, cells, 1, 0.3, 3, ESPP_FP64, &s));
  if (!s) return;
  std::vector<int32_t> id((size_t)n);
  for (int k = 0; k < n; ++k) id[(size_t)k] = k;
  CHECK(espp_add_particles(s, n, id.data(), nullptr, pos.data(), nullptr, nullptr));
  CHECK(espp_decompose(s));
  int32_t vl = -1, ia = -1;
  CHECK(espp_vl_create(s, 3.0, &vl));
  const double spec[7] = {3.0, 1.5, 3.0, 1.0, 1.0, 5.0, 2.0};
  CHECK(espp_mlp_create(s, vl, spec, &ia));
  std::vector<float> theta(10625), grad(10625);
  for (size_t k = 0; k < theta.size(); ++k) theta[k] = 0.05f * (float)std::sin(0.7 * (double)k);
  CHECK(espp_mlp_set_params(s, ia, theta.data(), (int64_t)theta.size()));
  CHECK(espp_add_interaction(s, ia));
  CHECK(espp_compute_forces(s));
  std::vector<double> fref((size_t)3 * n, 0.1);
  double loss = 0;
  CHECK(espp_mlp_zero_grad(s, ia));
  CHECK(espp_mlp_train_frame(s, ia, fref.data(), 1.0, &loss));
  CHECK(espp_mlp_grad(s, ia, grad.data(), (int64_t)grad.size()));
  CHECK(espp_mlp_adam(s, ia, 1e-3, 0.9, 0.999, 1e-8));
  CHECK(espp_mlp_get_params(s, ia, theta.data(), (int64_t)theta.size()));
  double e = 0;
  EXPECT_ERROR(espp_ia_energy(s, ia, &e));       // the learned force has no energy
  EXPECT_ERROR(espp_mlp_set_params(s, ia, theta.data(), 5));
  CHECK(espp_destroy(s));
}

void batch() {
  const int R = 6;
  const int32_t n[R] = {2048, 3000, 5000, 2500, 4096, 3500};
  std::vector<double> box(R), temp(R);
  std::vector<uint64_t> seed(R);
  std::vector<std::vector<double>> pos(R);
  for (int r = 0; r < R; ++r) {
    lattice(n[r], 0.8442, pos[(size_t)r], box[(size_t)r]);
    temp[(size_t)r] = 0.7 + 0.1 * r;
    seed[(size_t)r] = 100 + r;
  }
  espp_batch* b = nullptr;
  CHECK(espp_batch_create(R, n, box.data(), 2.5, 0.3, 1.0, 1.0, 1.0, temp.data(), seed.data(), &b));
  if (!b) return;
  for (int r = 0; r < R; ++r) {
    std::vector<int32_t> id((size_t)n[r]);
    for (int k = 0; k < n[r]; ++k) id[(size_t)k] = k;
    CHECK(espp_batch_load(b, r, id.data(), pos[(size_t)r].data(), nullptr, nullptr));
  }
  std::vector<int32_t> lpp(R);
  std::vector<double> table(16);
  double before = 0, after = 0;
  CHECK(espp_batch_autotune(b, 2, lpp.data(), table.data(), &before, &after));
  CHECK(espp_batch_run(b, 30, 0.005));
  CHECK(espp_batch_set_temperature(b, 1, 1.3));
  CHECK(espp_batch_set_graph(b, 0));
  CHECK(espp_batch_run(b, 7, 0.005));
  std::vector<double> e(R), w(R), mv2(R), out((size_t)3 * 5000);
  CHECK(espp_batch_energy(b, e.data(), w.data()));
  CHECK(espp_batch_kinetic(b, mv2.data()));
  CHECK(espp_batch_get(b, 2, ESPP_FIELD_POS, out.data()));
  int64_t total = 0;
  CHECK(espp_batch_pairs(b, 0, nullptr, 0, &total));
  std::vector<int32_t> pairs((size_t)2 * (size_t)(total + 1));
  CHECK(espp_batch_pairs(b, 0, pairs.data(), total, &total));
  int64_t resorts = 0, builds = 0;
  int32_t maxcount = 0;
  CHECK(espp_batch_stats(b, 0, &resorts, &builds, &maxcount));
  EXPECT_ERROR(espp_batch_get(b, R, ESPP_FIELD_POS, out.data()));
  CHECK(espp_batch_destroy(b));
}

}  // namespace

int main() {
  int32_t ndev = 0;
  if (espp_device_count(&ndev) != ESPP_OK || ndev < 1) {
    std::fprintf(stderr, "no GPU\n");
    return 2;
  }
  lj_fluid(ESPP_MIXED);
  lj_fluid(ESPP_FP64);
  learned_force();
  batch();
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("host sanitize driver: clean\n");
  return 0;
}
