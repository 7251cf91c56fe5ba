// Launch wrappers of kernels.hip, used by the host runtime (sim.hip).
#pragma once
#include "espp_dev.h"

namespace espp {

void launch_fold_cell(hipStream_t s, int n, double4* pos, int4* img, int* cell_of, int* rank,
                      int* cell_count, const Geom& g, Ctl* ctl, int do_fold);
void launch_scan_multi(hipStream_t s, int ncell, const int* count, int* start, int* blocksum, int* blockoff);
int scan_blocks(int ncell);
// blockmax of a needs nblocks(n, 256) entries
void launch_step_fused(hipStream_t s, int n, double* fx, double* fy, double* fz, const StepArgs& a, const Geom& g,
                       Ctl* ctl, unsigned long long* host_out);
void launch_scatter(hipStream_t s, int n, const int* cell_of, const int* rank, const int* start,
                    int* order);
void launch_sort_cells(hipStream_t s, int ncell, const int* start, int* order, const double4* pos, const Ctl* ctl);
void launch_permute(hipStream_t s, int n, const int* order, const int* cell_of, const double4* pos,
                    const double4* vel, const double* fx, const double* fy, const double* fz,
                    const int* id, const int4* img, double4* pos2, double4* vel2, double* fx2,
                    double* fy2, double* fz2, int* id2, int4* img2, int* cell_of_slot,
                    int* slot_of_id, uint4* posq, float4* posf, uint2* posh, const Geom& g);
void launch_posf(hipStream_t s, int n, const double4* pos, float4* posf, uint2* posh, double q16inv);
// x sub-bins of every cell for the tiled list build (ncell * (kXBins + 1) ints),
// from the positions the binning has just sorted
void launch_xbins(hipStream_t s, int n, int ncell, const int* cstart, const int* cell_of_slot, const float4* posf,
                  const Geom& g, int* xbin);
void launch_fill_int(hipStream_t s, int n, int* a, int v);
void launch_reset_ctl(hipStream_t s, Ctl* ctl, int fields = kCtlAll);
void launch_nbr_build(hipStream_t s, int n, const double4* pos, const int* id,
                      const int* cell_of_slot, const int* cstart, const Geom& g, double cutsq,
                      const int* ex_start, const int* ex_ids, int* nbr, int* cnt, int cap,
                      Ctl* ctl);
int lj_blocks(int n);
void launch_lj(hipStream_t s, bool mixed, bool onetype, bool energy, bool accum, int n,
               const uint4* posq, const double4* pos, const int* nbr, const int* cnt, int cap,
               const Geom& g, const LJParam* prm, int ntypes, double* fx, double* fy, double* fz,
               double* partial, const int* stop = nullptr);
void launch_tile_plan(hipStream_t s, int n, int T, int ntiles, const int* cell_of_slot, const int* cstart,
                      const Geom& g, int budget, TileRun* runs, int* nruns, Ctl* ctl);
void launch_nbr_build_tiled(hipStream_t s, int n, int T, const float4* posf, const double4* pos,
                            const int* id, const int* cell_of_slot, const int* cstart, const Geom& g, double cutsq,
                            float cut_lo, float cut_hi, const TileRun* runs, const int* nruns, const int* ex_start,
                            const int* ex_ids, uint4* nbr, int* cnt, int capg, int* lane_slot,
                            float clip_margin, float row_margin, const int* xbin, const uint2* posh,
                            double q16inv, int q16_tlo, int q16_w, Ctl* ctl);
void launch_lj_tiled(hipStream_t s, bool mixed, bool onetype, bool energy, bool accum, int n, int T, int lds_bytes,
                     const uint4* posq, const double4* pos, const uint4* nbr, const int* cnt, int capg,
                     const Geom& g, const TileRun* runs, const int* nruns, const int* lane_slot, const LJParam* prm,
                     int ntypes, double* fx, double* fy, double* fz, double* partial,
                     const PruneOut* prune = nullptr, const int* stop = nullptr);
int lj_tiled_blocks(int n, int T);
// FENE bonds as the last force of a step, finishing it per particle as
// launch_lj_tiled_step does (sa.blockmax needs nblocks(n, 256) entries).
void launch_fene_step(hipStream_t s, bool mixed, bool accum, int n, const double4* pos, const uint4* posq,
                      const int* id, const int* slot_of_id, int nid, const int* bstart, const int* bpartner,
                      const Geom& g, const FENEParam& fp, double* fx, double* fy, double* fz, const StepArgs& sa,
                      Ctl* ctl, unsigned long long* host_out, SpecState* spec = nullptr);
// Mixed tiled LJ as the last force of a step, finishing the step per particle
// (thermostat, integrate2, integrate1 into sa.pos_out / sa.posq_out), then the
// step's max |dx|^2 (sa.blockmax needs lj_tiled_blocks(n, T) entries) into
// ctl and host_out.  With prune, it also writes the pruned list *prune (same
// capacity capg as nbr, lane order of n entries).
void launch_lj_tiled_step(hipStream_t s, bool onetype, bool accum, int n, int T, int lds_bytes, const uint4* posq,
                          const double4* pos, const uint4* nbr, const int* cnt, int capg, const Geom& g,
                          const TileRun* runs, const int* nruns, const int* lane_slot, const LJParam* prm,
                          int ntypes, double* fx, double* fy, double* fz, const StepArgs& sa, Ctl* ctl,
                          unsigned long long* host_out, const PruneOut* prune = nullptr,
                          SpecState* spec = nullptr);
// Start of a window of steps with the device-side skin check (SpecState).
void launch_spec_begin(hipStream_t s, SpecState* sp, double dmax, double skin_half);
void launch_fene(hipStream_t s, bool mixed, bool energy, bool accum, int n, const double4* pos,
                 const uint4* posq, const int* id, const int* slot_of_id, int nid,
                 const int* bstart, const int* bpartner, const Geom& g, const FENEParam& fp,
                 double* fx, double* fy, double* fz, double* partial, const int* stop = nullptr);
void launch_tab_tiled(hipStream_t s, bool mixed, bool energy, bool accum, int n, int T, int lds_bytes,
                      const uint4* posq, const double4* pos, const uint4* nbr, const int* cnt, int capg,
                      const Geom& g, const TileRun* runs, const int* nruns, const int* lane_slot, const TabParam* prm,
                      const double* pool, int ntypes, double* fx, double* fy, double* fz, double* partial);
void launch_tab(hipStream_t s, bool energy, bool accum, int n, const double4* pos, const int* nbr, const int* cnt,
                int cap, const Geom& g, const TabParam* prm, const double* pool, int ntypes, double* fx, double* fy,
                double* fz, double* partial);
void launch_bond_tab(hipStream_t s, bool energy, bool accum, int n, const double4* pos, const int* id,
                     const int* slot_of_id, int nid, const int* bstart, const int* bpartner, const Geom& g,
                     const TabParam* prm, const double* pool, double* fx, double* fy, double* fz, double* partial);
constexpr int kRdfLdsBins = 12288;   // histograms up to this many bins are kept in LDS
void launch_rdf(hipStream_t s, int n, const double4* pos, const Geom& g, double dr, int nbins,
                unsigned long long* counts);
void launch_zero3(hipStream_t s, int n, double* a, double* b, double* c);
// fields of m particles given by id (3 doubles each) and their slots (-1: no such id)
void launch_gather_ids(hipStream_t s, int m, const int* ids, const int* slot_of_id, int nid, int field,
                       const double4* pos, const double4* vel, const double* fx, const double* fy, const double* fz,
                       const int4* img, double* out, int* slots_out);
void launch_scatter_slots(hipStream_t s, int m, const int* slots, int field, const double* val, double4* pos,
                          double4* vel);
void launch_thermo_integrate2(hipStream_t s, int n, const int* id, double4* vel, double* fx,
                              double* fy, double* fz, int lang, double pref1, double pref2,
                              unsigned long long seed, unsigned long long call,
                              const unsigned char* excl, int nid, int do_int2, double half_dt);
int kinetic_blocks(int n);
void launch_kinetic(hipStream_t s, int n, const double4* vel, const double4* pos,
                    unsigned long long type_mask, double* partial);
void launch_reduce_partials(hipStream_t s, int nblk, int nv, const double* partial, double* out);

}  // namespace espp
