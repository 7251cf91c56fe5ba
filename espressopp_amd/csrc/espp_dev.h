// Shared host/device definitions of the MI355X short-range MD library.
//
// Particle storage on the device is cell-sorted ("slot" order): after every
// decompose the particles of inner cell 0 come first, then cell 1, ... and
// inside a cell in the order of their previous slots (insertion order at the
// first decompose).  Cells use the reference's inner-cell
// numbering x + nx*(y + ny*z) (src/esutil/Grid.hpp:68-71).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace espp {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kMaxTypes = 16;      // particle types with a device parameter table

// x sub-bins per cell of the list builds' row clipping (a cell's table holds
// kXBins + 1 slots)
constexpr int kXBins = 16;

// Box and cell-grid geometry, passed by value to kernels.
struct Geom {
  double L[3], invL[3], halfL[3];
  int n[3];                        // inner cells per dimension
  double ics[3];                   // 1 / cellSize, cellSize = L / n (src/storage/CellGrid.cpp:55-56)
  double qinv[3];                  // 2^32 / L: FP64 position -> 32-bit fixed point
  float qs[3];                     // L / 2^32: fixed-point difference -> FP32 length
  double qsd[3];                   // L / 2^32 in FP64
  double qs2;                      // (L / 2^32)^2 for cubic boxes (tiled mixed kernel)
  float qsf;                       // L / 2^32 as FP32 (cubic boxes)
  double q16inv;                   // 1 / delta of the list build's 16-bit copy at binning (0: none)
  // The caller's cell grid with halfCellInt h (src/storage/DomainDecomposition.cpp:325-374):
  // nh = h * n cells of size >= (rc+skin)/h, stencil +-h.  The binning above uses the
  // grid n = nh / h; the caller's grid only fixes which of a pair's two particles the
  // reference takes as the first, i.e. the orientation of d for the list test.
  int hci;                         // halfCellInt (1: the two grids coincide)
  int nh[3];                       // the caller's cells per dimension
  double hics[3];                  // 1 / (L / nh)
  int* hcell;                      // per slot: linear cell of the caller's grid at the last binning (hci > 1)
  const int* cslot;                // per slot: binning cell (cell_of_slot), for the reference's orientation
};

// Lennard-Jones coefficients of one type pair (src/interaction/LennardJones.hpp:79-87).
struct LJParam {
  float ff1, ff2, rc2, shift;      // FP32 copies (untiled mixed kernel)
  double dff1, dff2, drc2, dshift; // FP64, length units
  double sig2, eps4;               // energy: 4 eps ((sig2/r2)^6 - (sig2/r2)^3)
  double dlo, dhi;                 // window of r^2 (length units) settled from FP64 positions
  // the same in fixed-point units of a cubic box (r^2 unit = (L / 2^(32+k))^2)
  double uff1, uff2, urc2, ulo, uhi, usig2;
  int set;
  int pad;
};

struct FENEParam {
  double K, r0, rMax, rMaxSqr, cutoffSqr, shift;
};

// One interaction.Tabulated table (src/interaction/Tabulated.hpp:43-115).  Its
// rows live in a coefficient pool at `off`: N force rows then N energy rows,
// each row kTabRow doubles {r_i, c0, c1, c2, c3}:
//   itype 1 (linear)  c0 r + c1
//   itype 2 (Akima)   c0 + c1 z + c2 int(z z) + c3 int(z z) z,  z = r - r_i
//   itype 3 (cubic)   c0 = y_i, c1 = y''_i, c2 = y_{i+1}, c3 = y''_{i+1}
struct TabParam {
  double inner, invdelta, deltasq6, rc2, shift;
  double dlo, dhi;                 // cutoff window settled from FP64 positions (mixed path)
  int itype, N, off, set;
};
constexpr int kTabRow = 5;

// One contiguous range of cells of a tile's staged region: the particles of
// cells [cell_a, cell_b] are slots [slot_start, slot_start + count) and sit at
// LDS offset lds_off of the tile's staging buffer.
struct TileRun {
  int slot_start, count, lds_off, cell_a, cell_b, pad0, pad1, pad2;
};
constexpr int kMaxRuns = 64;      // runs per tile

// Integrator arguments of a step applied per particle: the thermostat on the
// force (LangevinThermostat, aftCalcF), integrate2 when do_int2 is set, then
// integrate1 (VelocityVerlet).  pos is read and pos_out / posq_out written
// (the same buffers when the step runs in place).
struct StepArgs {
  const int* id;
  double4* vel;
  const double4* pos;
  double4* pos_out;
  uint4* posq_out;                 // fixed-point mirror of pos_out (null: none)
  const unsigned char* excl;       // thermostat exclusion flag per id (null: none)
  double pref1, pref2, dt;
  unsigned long long seed, call;
  int lang, nid, do_int2, pad;
  double* blockmax;                // per-block max |dx|^2
  const int* stop;                 // when set and nonzero at launch, the kernel does nothing (null: none)
};

// Device-side skin check of a window of steps launched without waiting on the
// host: each step's max |dx| is added to dmax (VelocityVerlet's maxDist), done
// counts the steps that ran, and stop is raised once dmax exceeds skin/2, after
// which the window's remaining kernels return at once.
struct SpecState {
  double dmax, skin_half;
  int stop, done;
};

// The fixed-point mirror has one scale in all three directions.
inline bool cubic_box(const Geom& g) { return g.qsd[0] == g.qsd[1] && g.qsd[1] == g.qsd[2]; }

// Pruned copy of a tiled list, written by a step kernel while it walks the full
// list: the entries with r^2 below r2 (the kernel's units) at the step's
// positions, in their order, in the full list's layout and capacity.
struct PruneOut {
  uint4* nbr;
  int* cnt;                        // lengths, by slot
  int* lane;                       // lane -> slot of the pruned list (ordered by its lengths)
  double r2;
};

// Field groups of Ctl that launch_reset_ctl clears: the binning's (max |dx|^2,
// nanpos), the list build's (overflow, maxcount) and the tile plan's
// (plan_fail, max_region).
constexpr int kCtlBin = 1, kCtlBuild = 2, kCtlPlan = 4, kCtlAll = 7;

// Per-call control block kept on the device.
struct Ctl {
  unsigned long long maxsq_bits;   // max |dx|^2 of the last integrate1 (non-negative double bits)
  int overflow;                    // neighbour list capacity exceeded
  int maxcount;                    // largest neighbour count of the last build
  int nanpos;                      // a NaN position was met while folding
  int plan_fail;                   // a tile exceeded the run or LDS limits
  int max_region;                  // largest staged region (particles) of the last plan
};

}  // namespace espp
