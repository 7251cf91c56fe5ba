/*
 * espp_amd.h -- C ABI of the MI355X-native short-range MD step of ESPResSo++.
 *
 * One handle (espp_sim) owns every device buffer of one periodic system: the
 * cell-sorted particles, the Verlet lists, the bond lists, the interactions and
 * the integrator state.  Caller arrays are borrowed and copied; no device
 * pointer escapes.  Every call returns 0 on success and a negative code on
 * failure; espp_last_error() then describes the failure (thread-local).
 * Calls on one handle are not re-entrant; all work of a handle runs on one HIP
 * stream of the current device.
 *
 * Each entry point names the reference interface it replaces (Boost.Python
 * registration and C++ implementation, paths relative to the ESPResSo++ tree).
 */
#ifndef ESPP_AMD_H
#define ESPP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct espp_sim espp_sim;

/* error codes */
#define ESPP_OK 0
#define ESPP_ERR_ARG (-1)          /* invalid argument (ValueError in Python) */
#define ESPP_ERR_STATE (-2)        /* call not valid in the current state */
#define ESPP_ERR_DEVICE (-3)       /* HIP runtime failure or no GPU */
#define ESPP_ERR_NAN (-4)          /* a particle position became NaN */
#define ESPP_ERR_OOM (-5)          /* device allocation failed */

/* precision of the pair path */
#define ESPP_FP64 0                /* all-FP64 pair forces */
#define ESPP_MIXED 1               /* pair forces from 32-bit fixed-point positions:
                                      exact FP64 r^2, FP32-seeded reciprocal, FP64
                                      LJ factor and accumulation; FP64 integration,
                                      list test and reductions */

/* particle fields for espp_get_particles / espp_set_particles */
#define ESPP_FIELD_POS 0
#define ESPP_FIELD_VEL 1
#define ESPP_FIELD_FORCE 2
#define ESPP_FIELD_IMAGE 3
#define ESPP_FIELD_TYPE 4
#define ESPP_FIELD_MASS 5

/* Last error message of the calling thread ("" when none). */
const char *espp_last_error(void);

/* Library version string. */
const char *espp_version(void);

/* Number of usable GPUs (0 when no HIP device is present). */
int espp_device_count(int32_t *count);

/* System + OrthorhombicBC + storage::DomainDecomposition for one node.
 * Replaces System() (src/System.cpp:206-226), bc::OrthorhombicBC (src/bc/OrthorhombicBC.cpp:35-50)
 * and DomainDecomposition::DomainDecomposition(system, nodeGrid, cellGrid, halfCellInt) /
 * createCellGrid / initCellInteractions (src/storage/DomainDecomposition.cpp:70-170,325-374).
 * cell_grid is the inner cell grid (tools.decomp.cellGrid(box, nodeGrid, rc, skin,
 * halfCellInt)), half_cell_int the halfCellInt (1 full cells, 2 half cells, ...; 1..8),
 * skin is System.skin, seed the RNG seed.  cell_grid / half_cell_int must be at least 2 in
 * every dimension (ESPP_ERR_ARG otherwise).  precision: ESPP_FP64 or ESPP_MIXED. */
int espp_create(const double box[3], const int32_t cell_grid[3], int32_t half_cell_int, double skin,
                uint64_t seed, int32_t precision, espp_sim **out);

/* Frees the handle and all its device memory. */
int espp_destroy(espp_sim *sim);

/* System::setSkin / getSkin (src/System.cpp:207). Takes effect at the next rebuild. */
int espp_set_skin(espp_sim *sim, double skin);

/* Storage::addParticle for n particles (src/storage/Storage.cpp:285-313, bound at
 * src/storage/Storage.cpp:1037; Python loop at src/storage/Storage.py:233-330).
 * Positions are folded on insertion.  type, vel and mass may be NULL (0, 0, 1). */
int espp_add_particles(espp_sim *sim, int64_t n, const int32_t *id, const int32_t *type,
                       const double *pos, const double *vel, const double *mass);

/* Storage::decompose (src/storage/Storage.cpp:598-604, bound at :1049): fold,
 * re-sort into cells and fire onParticlesChanged (rebuild every Verlet list). */
int espp_decompose(espp_sim *sim);

/* Number of particles (Storage::getNRealParticles). */
int espp_num_particles(espp_sim *sim, int64_t *n);

/* Ids of all particles in ascending order (Storage::getRealParticleIDs,
 * src/storage/Storage.cpp:1050).  out holds espp_num_particles entries. */
int espp_get_ids(espp_sim *sim, int32_t *out);

/* Particle properties in ascending id order (ParticleAccess getters of
 * src/Particle.py: pos, v, f, imageBox, type, mass).  out holds 3 doubles per
 * particle for POS/VEL/FORCE/IMAGE and 1 for TYPE/MASS. */
int espp_get_particles(espp_sim *sim, int32_t field, double *out);

/* Properties of n particles given by id (Storage::lookupRealParticle + the
 * ParticleAccess getters, src/storage/Storage.cpp:1037-1050, src/Particle.py), in the
 * order of ids: 3 doubles per particle for POS/VEL/FORCE/IMAGE, 1 for TYPE/MASS.  Costs
 * O(n), not O(N).  Fails (ESPP_ERR_ARG) for an id that does not exist. */
int espp_get_particles_by_id(espp_sim *sim, int32_t field, int64_t n, const int32_t *ids, double *out);

/* Set a property of n particles given by id (Particle.pos/.v/.type/.mass setters,
 * src/Particle.py).  values: 3 per particle for POS/VEL, 1 for TYPE/MASS.  Costs O(n). */
int espp_set_particles(espp_sim *sim, int32_t field, int64_t n, const int32_t *ids,
                       const double *values);

/* VerletList(system, cutoff) (src/VerletList.cpp:44-67, bound at :373-385):
 * list of all pairs with d^2 <= (cutoff+skin)^2, rebuilt on every decompose. */
int espp_vl_create(espp_sim *sim, double cutoff, int32_t *vl);

/* VerletList::exclude for n id pairs (src/VerletList.cpp:327-332), then rebuild. */
int espp_vl_exclude(espp_sim *sim, int32_t vl, int64_t n, const int32_t *pairs);

/* VerletList::rebuild (src/VerletList.cpp:86-117). */
int espp_vl_rebuild(espp_sim *sim, int32_t vl);

/* VerletList::totalSize: number of distinct pairs (src/VerletList.cpp:302-310). */
int espp_vl_size(espp_sim *sim, int32_t vl, int64_t *npairs);

/* VerletList.builds (src/VerletList.cpp:378). */
int espp_vl_builds(espp_sim *sim, int32_t vl, int64_t *builds);

/* All pairs as (id1, id2) with id1 < id2 (VerletList::getPair / getAllPairs,
 * src/VerletList.cpp:314-325).  Writes min(cap, total) pairs and returns the
 * total in *n. */
int espp_vl_pairs(espp_sim *sim, int32_t vl, int32_t *out, int64_t cap, int64_t *n);

/* VerletList.getVerletCutoff (src/VerletList.cpp:69). */
int espp_vl_cutoff(espp_sim *sim, int32_t vl, double *cut_verlet);

/* FixedPairList(storage) (src/FixedPairList.cpp:52-66, bound at :346-355). */
int espp_fpl_create(espp_sim *sim, int32_t *fpl);

/* FixedPairList::add for n id pairs (src/FixedPairList.cpp:88-141). */
int espp_fpl_add(espp_sim *sim, int32_t fpl, int64_t n, const int32_t *pairs);

/* FixedPairList::totalSize (src/FixedPairList.cpp:326-333). */
int espp_fpl_size(espp_sim *sim, int32_t fpl, int64_t *n);

/* interaction.VerletListLennardJones(vl) (src/interaction/LennardJones.cpp:105-109). */
int espp_lj_create(espp_sim *sim, int32_t vl, int32_t *ia);

/* VerletListInteractionTemplate::setPotential(type1, type2, LennardJones(eps,
 * sigma, cutoff, shift)) (src/interaction/VerletListInteractionTemplate.hpp:65-75,
 * src/interaction/LennardJones.hpp:60-75).  shift = NaN selects 'auto'. */
int espp_lj_set(espp_sim *sim, int32_t ia, int32_t type1, int32_t type2, double epsilon,
                double sigma, double cutoff, double shift);

/* interaction.FixedPairListFENE(system, fpl, FENE(K, r0, rMax, cutoff, shift))
 * (src/interaction/FENE.cpp:39-53, src/interaction/FENE.hpp:40-112). */
int espp_fene_create(espp_sim *sim, int32_t fpl, double K, double r0, double rMax,
                     double cutoff, double shift, int32_t *ia);

/* interaction.VerletListTabulated(vl) (src/interaction/Tabulated.cpp:73,100-103,
 * VerletListInteractionTemplate<Tabulated>). */
int espp_tab_create(espp_sim *sim, int32_t vl, int32_t *ia);

/* VerletListTabulated::setPotential(type1, type2, Tabulated(itype, filename, cutoff))
 * (src/interaction/Tabulated.hpp:56-115, Tabulated.cpp:60-80).  The table is given as
 * its n rows (r, e, f) on a uniform grid, as the reference reads them from the file
 * (InterpolationLinear.cpp:89-125); itype 1 linear, 2 Akima, 3 cubic spline. */
int espp_tab_set(espp_sim *sim, int32_t ia, int32_t type1, int32_t type2, int32_t itype, int64_t n,
                 const double *r, const double *e, const double *f, double cutoff);

/* interaction.FixedPairListTabulated(system, fpl, Tabulated(itype, filename, cutoff))
 * (src/interaction/Tabulated.cpp:85,182-190, FixedPairListInteractionTemplate<Tabulated>). */
int espp_tabbond_create(espp_sim *sim, int32_t fpl, int32_t itype, int64_t n, const double *r,
                        const double *e, const double *f, double cutoff, int32_t *ia);

/* espp_ia_energy + espp_ia_virial in one call that allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight monitoring loops, harnesses that forbid hipMalloc).  The per-workgroup
 * partial sums and their reduction use a static 2 MiB device scratch per GPU (kept until the
 * process ends), so the system may have at most 262,136 / 8 = 32,767 force workgroups
 * (about 8.4 M particles with 256-particle tiles, 2.1 M with 64-particle tiles); larger
 * systems get ESPP_ERR_ARG.  The Verlet lists, tile plan and parameter tables must be
 * current (any espp_run or espp_compute_forces leaves them so); otherwise ESPP_ERR_STATE.
 * All calls on one GPU share the scratch: calls that overlap on one GPU race.  virial and
 * tensor6 (xx, yy, zz, xy, xz, yz) may be NULL. */
int espp_ia_energy_noalloc(espp_sim *sim, int32_t ia, double *energy, double *virial, double *tensor6);

/* System::addInteraction (src/System.cpp:75-85, bound at :221): appends the
 * interaction to the short-range list whose order is the force summation order. */
int espp_add_interaction(espp_sim *sim, int32_t ia);

/* System::removeInteraction (src/System.cpp:108-135, bound at :222): removes the
 * interaction at `position` of the short-range list (0-based, addInteraction order). */
int espp_remove_interaction(espp_sim *sim, int32_t position);

/* FixedPairListFENE::setPotential(FENE(K, r0, rMax, cutoff, shift))
 * (src/interaction/FENE.py:114-116, FixedPairListInteractionTemplate::setPotential). */
int espp_fene_set(espp_sim *sim, int32_t ia, double K, double r0, double rMax, double cutoff, double shift);

/* setFixedPairList of a FixedPairList interaction, FENE or Tabulated
 * (src/interaction/FENE.py:122-124, FixedPairListInteractionTemplate::setFixedPairList). */
int espp_ia_set_fpl(espp_sim *sim, int32_t ia, int32_t fpl);

/* Interaction::computeEnergy (src/interaction/Interaction.cpp:46). */
int espp_ia_energy(espp_sim *sim, int32_t ia, double *energy);

/* Interaction::computeVirial and computeVirialTensor (src/interaction/Interaction.cpp:52;
 * src/interaction/VerletListInteractionTemplate.hpp:274-331).  tensor6 (may be NULL)
 * receives xx, yy, zz, xy, xz, yz. */
int espp_ia_virial(espp_sim *sim, int32_t ia, double *virial, double *tensor6);

/* VelocityVerlet::calcForces without thermostat (src/integrator/VelocityVerlet.cpp:307-330). */
int espp_compute_forces(espp_sim *sim);

/* LangevinThermostat gamma/temperature/addExclusions and addExtension/disconnect
 * (src/integrator/LangevinThermostat.cpp:40-110,224-233).  on=0 disconnects. */
int espp_langevin(espp_sim *sim, int32_t on, double gamma, double temperature, int64_t nexcl,
                  const int32_t *excl);

/* MDIntegrator::run -> VelocityVerlet::run (src/integrator/MDIntegrator.cpp:103,
 * src/integrator/VelocityVerlet.cpp:63-169) with time step dt. */
int espp_run(espp_sim *sim, int32_t nsteps, double dt);

/* MDIntegrator.step property (src/integrator/MDIntegrator.cpp:101). */
int espp_get_step(espp_sim *sim, int64_t *step);
int espp_set_step(espp_sim *sim, int64_t step);

/* VelocityVerlet::getNumResorts (src/integrator/VelocityVerlet.cpp:434). */
int espp_num_resorts(espp_sim *sim, int64_t *n);

/* VelocityVerlet::getTimers (src/integrator/VelocityVerlet.cpp:165-204): Run,
 * ForceComp[0..2], UpdateGhosts, CollectGhostForces, Integrate1, Integrate2,
 * Resort, Lost (seconds of the last run()). */
int espp_timers(espp_sim *sim, double out10[10]);

/* esutil::RNG saveState / loadState (src/esutil/RNG.cpp:86-143).  The Langevin
 * noise here is Philox4x32-10 keyed by (seed, thermostat call, particle id), so
 * its whole state is the seed and the number of thermostat calls made. */
int espp_get_rng_state(espp_sim *sim, uint64_t *seed, uint64_t *calls);
int espp_set_rng_state(espp_sim *sim, uint64_t seed, uint64_t calls);

/* VelocityVerlet's skin bookkeeping between runs (src/integrator/VelocityVerlet.cpp:63-169):
 * the accumulated displacement maxDist and whether a resort is pending. */
int espp_get_run_state(espp_sim *sim, double *maxdist, int32_t *resort);
int espp_set_run_state(espp_sim *sim, double maxdist, int32_t resort);

/* analysis.RadialDistrF(system).compute(nbins) (src/analysis/RadialDistrF.cpp:47-140):
 * histogram of all pair distances up to L_y / 2 in nbins bins, normalised to g(r). */
int espp_rdf(espp_sim *sim, int32_t nbins, double *out);

/* Sum of m v^2 over all particles (src/analysis/Temperature.hpp:112-122). */
int espp_kinetic(espp_sim *sim, double *mv2);

/* Kinetic sums over particles whose type t < 64 has bit t set in type_mask
 * (Temperature with add_type, src/analysis/Temperature.hpp:100-122; kinetic part of
 * PressureTensor, src/analysis/PressureTensor.hpp:51-82): sum m v^2, the particle
 * count and the tensor xx, yy, zz, xy, xz, yz.  Each output may be NULL. */
int espp_kinetic_tensor(espp_sim *sim, uint64_t type_mask, double *mv2, int64_t *count,
                        double *tensor6);


/* ------------------------------------------------------------------------
 * Learned coarse-grained pair force.
 *
 * No counterpart in the reference (ESPResSo++ has no learned potentials): a
 * density-dependent pair MLP over a Verlet list, the learned alternative to
 * interaction.VerletListTabulated (src/interaction/Tabulated.hpp:43-115), which
 * stays the plain fallback.  For a listed pair at distance r < rc the inputs are
 * 32 Gaussian RBFs (centres k rc / 31, width rc / 31) and the pair means of two
 * smooth local densities (radii R_a, R_b; cosine weights), normalised by
 * (x - mean) / scale; three 64-wide SiLU layers and a linear output give s, the
 * pair force is g = s (1 + cos(pi r / rc)) / 2 along d / r.  Parameters: 10625
 * FP32 values W1 [34][64], b1, W2 [64][64], b2, W3 [64][64], b3, w4 [64], b4.
 * Evaluation and training run on BF16 matrix cores with FP32 accumulation.
 * ------------------------------------------------------------------------ */

/* Learned pair force over Verlet list vl: spec = {rc, R_a, R_b, mean_a, scale_a,
 * mean_b, scale_b} (rc, R_a, R_b within the list cutoff).  Added to the force
 * loop with espp_add_interaction like any interaction; it has no energy. */
int espp_mlp_create(espp_sim *sim, int32_t vl, const double spec[7], int32_t *ia);

/* Parameters (n = 10625 FP32 values) in and out. */
int espp_mlp_set_params(espp_sim *sim, int32_t ia, const float *theta, int64_t n);
int espp_mlp_get_params(espp_sim *sim, int32_t ia, float *theta, int64_t n);

/* Force matching on the current configuration: adds the gradient of
 * weight * sum_i |F_i - fref_i|^2 / (3 N) to the accumulated gradient and
 * returns that loss.  fref: 3 doubles per particle in ascending id order.  The
 * particles' forces are left holding the model's forces. */
int espp_mlp_train_frame(espp_sim *sim, int32_t ia, const double *fref, double weight, double *loss);

/* Accumulated gradient (n = 10625) and its reset. */
int espp_mlp_grad(espp_sim *sim, int32_t ia, float *grad, int64_t n);
int espp_mlp_zero_grad(espp_sim *sim, int32_t ia);

/* One Adam step (Kingma & Ba) with the accumulated gradient, which is then cleared. */
int espp_mlp_adam(espp_sim *sim, int32_t ia, double lr, double beta1, double beta2, double eps);

/* ------------------------------------------------------------------------
 * Batched replicas (parallel tempering on one GPU).
 *
 * One espp_batch holds many periodic cubic Lennard-Jones systems of uneven size
 * (one particle type, LJ(eps, sigma, rc) with the automatic shift, Verlet skin,
 * optional Langevin thermostat) packed into shared device arrays.  Each step of
 * every replica is one fixed sequence of launches over all replicas (integrate,
 * per-replica skin check and binning, list build, forces) with the rebuild
 * decisions kept on the device; the sequence is captured once into a HIP graph
 * and replayed, so a run makes no host round trip per step.  Replaces the
 * replica loop of ParallelTempering.run / MultiSystem.runIntegrator
 * (src/ParallelTempering.py:150-162, src/MultiSystem.py:34-81) over systems built
 * as System + VerletList + VerletListLennardJones + VelocityVerlet +
 * LangevinThermostat (src/VerletList.cpp:44-67, src/interaction/LennardJones.hpp,
 * src/integrator/VelocityVerlet.cpp:63-169, src/integrator/LangevinThermostat.cpp).
 * ------------------------------------------------------------------------ */
typedef struct espp_batch espp_batch;

/* nrep replicas: n[r] particles in a cubic box of edge box[r], cell grid
 * int(box / (rc + skin)) per dimension (at least 3, at most 8192 cells);
 * thermostat temperature[r] with friction gamma (gamma = 0: NVE) and Philox seed
 * seed[r].  Replaces building nrep Systems as in src/ParallelTempering.py:114-148. */
int espp_batch_create(int32_t nrep, const int32_t *n, const double *box, double rc, double skin, double eps,
                      double sigma, double gamma, const double *temperature, const uint64_t *seed,
                      espp_batch **out);

int espp_batch_destroy(espp_batch *b);

/* Particles of replica rep (Storage::addParticles, src/storage/Storage.py:233-330):
 * n[rep] unique non-negative ids, positions (3 per particle), velocities and
 * masses (may be NULL: 0 and 1). */
int espp_batch_load(espp_batch *b, int32_t rep, const int32_t *ids, const double *pos, const double *vel,
                    const double *mass);

/* LangevinThermostat.temperature of one replica (the swap of
 * src/ParallelTempering.py:194-196); takes effect at the next run. */
int espp_batch_set_temperature(espp_batch *b, int32_t rep, double temperature);
int espp_batch_get_temperature(espp_batch *b, int32_t rep, double *temperature);

/* VelocityVerlet::run(nsteps) of every replica (src/integrator/VelocityVerlet.cpp:63-169). */
int espp_batch_run(espp_batch *b, int32_t nsteps, double dt);

/* Interaction::computeEnergy / computeVirial of each replica's LJ interaction
 * (src/interaction/VerletListInteractionTemplate.hpp:116-198,274-331): nrep values
 * each; virial may be NULL. */
int espp_batch_energy(espp_batch *b, double *energy, double *virial);

/* Sum of m v^2 per replica (src/analysis/Temperature.hpp:112-122): nrep values. */
int espp_batch_kinetic(espp_batch *b, double *mv2);

/* Particle fields of one replica in ascending id order, 3 doubles per particle:
 * ESPP_FIELD_POS, _VEL, _FORCE or _IMAGE (src/Particle.py getters). */
int espp_batch_get(espp_batch *b, int32_t rep, int32_t field, double *out);

/* Pairs (id1 < id2) of one replica's Verlet list (VerletList::getAllPairs,
 * src/VerletList.cpp:314-325); min(cap, total) pairs are written, total returned. */
int espp_batch_pairs(espp_batch *b, int32_t rep, int32_t *out, int64_t cap, int64_t *total);

/* Lanes per particle (1, 2, 4 or 8) of the force launch, per replica.  No
 * reference counterpart: the reference evaluates one pair at a time per rank. */
int espp_batch_set_lpp(espp_batch *b, const int32_t *lpp);
int espp_batch_get_lpp(espp_batch *b, int32_t *lpp);

/* Average time (ms) of one force launch over all replicas, over reps launches. */
int espp_batch_force_time(espp_batch *b, int32_t reps, double *ms);

/* Average time (ms) of one list build of all replicas, over reps builds from one
 * binning: every replica is rebuilt (as a requested rebuild does, counted in its
 * builds) and binned once first.  No reference counterpart: a measurement of the
 * batched engine's list kernel. */
int espp_batch_build_time(espp_batch *b, int32_t reps, double *ms);

/* Autotuner of the force launch: replicas in four size classes (quartiles of
 * N), each class trying 1, 2, 4 and 8 lanes per particle with HIP-event timing;
 * the fastest shape is kept.  lpp_out (nrep), table (16 times, class-major),
 * before and after (ms) may be NULL.  No reference counterpart. */
int espp_batch_autotune(espp_batch *b, int32_t reps, int32_t *lpp_out, double *table, double *before,
                        double *after);

/* Replay each run's steps from captured HIP graphs (on = 1, the default) or launch
 * the same kernels directly (on = 0); both give identical results. */
int espp_batch_set_graph(espp_batch *b, int32_t on);

/* Per-replica counters: skin-triggered resorts (VelocityVerlet::getNumResorts,
 * src/integrator/VelocityVerlet.cpp:434, summed over runs), list builds, and the
 * longest neighbour list built so far. */
int espp_batch_stats(espp_batch *b, int32_t rep, int64_t *resorts, int64_t *builds, int32_t *maxcount);

#ifdef __cplusplus
}
#endif

#endif /* ESPP_AMD_H */
