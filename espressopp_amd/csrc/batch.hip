// Batched replicas: many periodic cubic Lennard-Jones systems of uneven size
// (the replicas of a parallel-tempering run, src/ParallelTempering.py:90-200,
// src/MultiSystem.py:34-81) packed into shared device arrays and advanced
// together.  Every launch covers all replicas; a replica's per-step decisions
// (the skin/2 rebuild trigger of VelocityVerlet::run,
// src/integrator/VelocityVerlet.cpp:121-148, and its thermostat call counter)
// live in device memory, so a step is a fixed sequence of launches with
// no host round trip, captured once into a HIP graph and replayed:
//
//   kb_step       thermostat + integrate2 of the previous step and integrate1
//                 (src/integrator/LangevinThermostat.cpp:141-177,
//                 src/integrator/VelocityVerlet.cpp:236-305); per-replica
//                 max |dx|^2 by one atomic per wavefront
//   kb_check      one workgroup per replica: accumulate maxDist, decide the
//                 rebuild (and clear the rebuilding replicas' cell counts)
//   binning       kb_fold_count, kb_scan, kb_scatter, kb_sort_cells,
//                 kb_permute, kb_copy_back over all particles (or cells) of
//                 every replica; workgroups of replicas that do not rebuild
//                 return at once (Storage::decompose for one node,
//                 src/storage/DomainDecomposition.cpp:684-1179)
//   kb_build      full Verlet list of the replicas that rebuilt
//                 (src/VerletList.cpp:123-298, pair test in the reference's
//                 orientation, FP64 without contraction)
//   kb_force      mixed-precision LJ (fixed-point mirror differences, FP64
//                 r^2 and force sum), `lpp` lanes per particle per replica
//
// Replica r owns slots [slot0, slot0 + n) (slot0 a multiple of 64, so a 64-slot
// neighbour-list group never straddles two replicas) and cells
// [cell0, cell0 + nc^3] of the cell-start array.  Workgroups of the particle
// kernels are described by a block map {replica, first slot, count, lanes per
// particle} so that no workgroup straddles two replicas and the per-replica
// tuning (lanes per particle) is a per-workgroup constant.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "espp_amd.h"
#include "espp_dev.h"
#include "espp_devfn.h"
#include "host_util.h"

namespace espp {
namespace batch {

constexpr int kBinThreads = 1024;
constexpr int kMaxBinCells = 8192;      // nc^3 per replica held in LDS by the binning
constexpr int kMaxCellFill = 256;       // particles per cell the in-cell ordering handles
constexpr int kPartThreads = 256;

struct RepDesc {
  double L, invL, ics, qinv, qsd;
  double skin_half, cutsq;              // skin / 2 and (rc + skin)^2
  double ff1, ff2, rc2, shift, sig2, eps4, dlo, dhi;
  double cflo, cfhi;                    // (rc + skin)^2 window of the list build's FP32 prefilter
  double ulo, uhi, uff1, uff2, usig2;   // dlo, dhi, ff1, ff2, sig2 in fixed-point units (force launch)
  double pref1, pref2;                  // -gamma, sqrt(24 T gamma / dt)
  unsigned long long seed;
  double q16inv;                        // 1 / delta of the list build's 16-bit copies (0: FP32 prefilter)
  int slot0, n, cell0, nc;
  int lang, q16_tlo, q16_w, pad0;       // r^2 thresholds of the 16-bit prefilter (q16_thresholds)
};

struct RepState {
  unsigned long long maxsq;             // bits of the step's max |dx|^2 (non-negative doubles order as integers)
  double maxDist;
  unsigned long long call;              // thermostat applications so far
  long long resorts, builds;
  int rebuild, first, force, overflow, maxcount, nanpos, cellfull, pad;
};

struct BlockMap {
  int rep, s0, count, lpp;
};

// ------------------------------------------------------------------ device


__device__ __forceinline__ void thermostat(int pid, const double4& v, double& f0, double& f1, double& f2,
                                           double pref1, double pref2, unsigned long long seed,
                                           unsigned long long call) {
#pragma clang fp contract(off)
  uint32_t o[4];
  philox((uint32_t)pid, (uint32_t)call, (uint32_t)(call >> 32), 0x4C414E47u, (uint32_t)seed,
         (uint32_t)(seed >> 32), o);
  const double u0 = ((double)o[0] + 0.5) * 2.3283064365386963e-10;
  const double u1 = ((double)o[1] + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)o[2] + 0.5) * 2.3283064365386963e-10;
  const double m = v.w, massf = sqrt(m);
  f0 += (pref1 * v.x) * m + (pref2 * (u0 - 0.5)) * massf;
  f1 += (pref1 * v.y) * m + (pref2 * (u1 - 0.5)) * massf;
  f2 += (pref1 * v.z) * m + (pref2 * (u2 - 0.5)) * massf;
}

// Thermostat (when the replica has one and this is not the run's first step),
// integrate2 (not on the first step) and integrate1, per particle.  Each
// wavefront adds its max |dx|^2 to the replica's maximum with one atomic.
__global__ __launch_bounds__(kPartThreads) void kb_step(const BlockMap* __restrict__ bm,
                                                        const RepDesc* __restrict__ rd, RepState* __restrict__ st,
                                                        double4* __restrict__ pos, double4* __restrict__ vel,
                                                        uint4* __restrict__ posq, const double* __restrict__ fx,
                                                        const double* __restrict__ fy,
                                                        const double* __restrict__ fz, const int* __restrict__ id,
                                                        double dt) {
#pragma clang fp contract(off)
  const BlockMap m = bm[blockIdx.x];
  const int t = threadIdx.x;
  const int r = m.rep;
  const bool first = st[r].first != 0;
  const bool lang = rd[r].lang && !first;
  double sq = 0.0;
  if (t < m.count) {
    const int s = m.s0 + t;
    double4 v = vel[s];
    double4 p = pos[s];
    double f0 = fx[s], f1 = fy[s], f2 = fz[s];
    if (lang) thermostat(id[s], v, f0, f1, f2, rd[r].pref1, rd[r].pref2, rd[r].seed, st[r].call);
    const double dtfm = 0.5 * dt / v.w;
    if (!first) { v.x += dtfm * f0; v.y += dtfm * f1; v.z += dtfm * f2; }
    v.x += dtfm * f0; v.y += dtfm * f1; v.z += dtfm * f2;
    const double d0 = v.x * dt, d1 = v.y * dt, d2 = v.z * dt;
    p.x += d0; p.y += d1; p.z += d2;
    vel[s] = v;
    pos[s] = p;
    const double qinv = rd[r].qinv;
    posq[s] = make_uint4(fixq(p.x, qinv), fixq(p.y, qinv), fixq(p.z, qinv), (uint32_t)(int)p.w);
    sq = d0 * d0 + d1 * d1 + d2 * d2;
  }
  sq = wave_max(sq);
  if ((t & 63) == 0) atomicMax(&st[r].maxsq, (unsigned long long)__double_as_longlong(sq));
}

// Thermostat on the current forces (run start with the heat-up factor, or the
// run's last step) and optionally integrate2; the call counters advance in
// kb_rep_op.
__global__ __launch_bounds__(kPartThreads) void kb_thermo_int2(const BlockMap* __restrict__ bm,
                                                               const RepDesc* __restrict__ rd,
                                                               const RepState* __restrict__ st,
                                                               double4* __restrict__ vel, double* __restrict__ fx,
                                                               double* __restrict__ fy, double* __restrict__ fz,
                                                               const int* __restrict__ id, double pref2_factor,
                                                               int do_int2, double half_dt) {
#pragma clang fp contract(off)
  const BlockMap m = bm[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= m.count) return;
  const int r = m.rep, s = m.s0 + t;
  double4 v = vel[s];
  double f0 = fx[s], f1 = fy[s], f2 = fz[s];
  if (rd[r].lang) {
    thermostat(id[s], v, f0, f1, f2, rd[r].pref1, rd[r].pref2 * pref2_factor, rd[r].seed, st[r].call);
    fx[s] = f0; fy[s] = f1; fz[s] = f2;
  }
  if (do_int2) {
    const double dtfm = half_dt / v.w;
    v.x += dtfm * f0; v.y += dtfm * f1; v.z += dtfm * f2;
    vel[s] = v;
  }
}

// One thread per replica: op 0 marks the run's first step, op 1 counts a
// thermostat application, op 2 requests a rebuild at the next binning.
__global__ void kb_rep_op(int nrep, const RepDesc* __restrict__ rd, RepState* __restrict__ st, int op) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  if (op == 0) st[r].first = 1;
  else if (op == 1) { if (rd[r].lang) st[r].call++; }
  else { st[r].force = 1; st[r].maxsq = 0; }
}

// Exclusive scan of cnt[0, nc) in LDS by one workgroup of kBinThreads threads;
// cnt[nc] receives the total.
__device__ void block_scan_lds(int* cnt, int nc, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (nc + kBinThreads - 1) / kBinThreads;
  const int b = tid * per, e = min(b + per, nc);
  int local = 0;
  for (int c = b; c < e; ++c) local += cnt[c];
  int incl = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < kBinThreads / 64; ++q) { const int x = wsum[q]; wsum[q] = run; run += x; }
    wsum[kBinThreads / 64] = run;
  }
  __syncthreads();
  int off = wsum[w] + incl - local;
  for (int c = b; c < e; ++c) { const int x = cnt[c]; cnt[c] = off; off += x; }
  if (tid == 0) cnt[nc] = wsum[kBinThreads / 64];
  __syncthreads();
}

// One workgroup per replica.  Thread 0 accumulates the step's displacement into
// maxDist and decides the rebuild (maxDist > skin/2, a NaN, or a requested
// one); a rebuilding replica's cell counts are then cleared for kb_fold_count.
__global__ __launch_bounds__(kBinThreads) void kb_check(const RepDesc* __restrict__ rd, RepState* __restrict__ st,
                                                        int use_step, int* __restrict__ ccount) {
  __shared__ int flag;
  const int r = blockIdx.x;
  const RepDesc& d = rd[r];
  if (threadIdx.x == 0) {
    RepState& s = st[r];
    int rb = s.force;
    if (use_step) {
      const bool first = s.first != 0;
      if (d.lang && !first) s.call++;      // kb_step applied the thermostat with this call
      s.first = 0;
      const double sq = __longlong_as_double((long long)s.maxsq);
      s.maxsq = 0ull;
      s.maxDist += sqrt(sq);
      if (!(s.maxDist <= d.skin_half)) {
        if (!rb) s.resorts++;
        rb = 1;
      }
    }
    if (rb) { s.maxDist = 0.0; s.force = 0; s.builds++; }
    s.rebuild = rb;
    flag = rb;
  }
  __syncthreads();
  if (!flag) return;
  const int ncell = d.nc * d.nc * d.nc;
  for (int c = threadIdx.x; c <= ncell; c += kBinThreads) ccount[d.cell0 + c] = 0;
}

// Fold every position (OrthorhombicBC::foldCoordinate, src/bc/OrthorhombicBC.cpp:144-161),
// map it to its cell (CellGrid::mapPositionToCellChecked, src/storage/CellGrid.cpp:98-136)
// and count it there; lanes of a wavefront in one cell share one atomic.
__global__ __launch_bounds__(kPartThreads) void kb_fold_count(const BlockMap* __restrict__ bm,
                                                              const RepDesc* __restrict__ rd,
                                                              RepState* __restrict__ st, double4* __restrict__ pos,
                                                              int4* __restrict__ img, int* __restrict__ ccount,
                                                              int* __restrict__ tcell, int* __restrict__ trank) {
#pragma clang fp contract(off)
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild) return;
  const RepDesc& d = rd[r];
  const int t = threadIdx.x;
  const bool live = t < m.count;
  const int s = m.s0 + (live ? t : 0);
  double4 p = pos[s];
  int4 im = img[s];
  double x[3] = {p.x, p.y, p.z};
  int iv[3] = {im.x, im.y, im.z};
  int c[3];
  bool bad = false, moved = false;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(fabs(x[q] * d.invL) < 1.0e9)) { bad = true; c[q] = 0; continue; }
    const int tmp = (int)floor(x[q] * d.invL);
    iv[q] += tmp;
    x[q] -= (double)tmp * d.L;
    moved |= tmp != 0;
    int cc = (int)((x[q] - 0.0) * d.ics) + 1;
    if (cc < 1) cc = 1;
    else if (cc >= d.nc + 1) cc = d.nc;
    c[q] = cc - 1;
  }
  const int cell = c[0] + d.nc * (c[1] + d.nc * c[2]);
  const int rank = wave_aggregated_rank(cell, live, ccount + d.cell0);
  if (!live) return;
  if (bad) st[r].nanpos = 1;
  if (moved) {
    p.x = x[0]; p.y = x[1]; p.z = x[2];
    pos[s] = p;
    img[s] = make_int4(iv[0], iv[1], iv[2], 0);
  }
  tcell[s] = cell;
  trank[s] = rank;
}

// Exclusive scan of a rebuilding replica's cell counts into its cell starts
// (one workgroup per replica, the counts in LDS).
__global__ __launch_bounds__(kBinThreads) void kb_scan(const RepDesc* __restrict__ rd,
                                                       const RepState* __restrict__ st,
                                                       const int* __restrict__ ccount, int* __restrict__ cstart) {
  __shared__ int cnt[kMaxBinCells + 1];
  __shared__ int wsum[kBinThreads / 64 + 1];
  const int r = blockIdx.x;
  if (!st[r].rebuild) return;
  const RepDesc& d = rd[r];
  const int ncell = d.nc * d.nc * d.nc;
  for (int c = threadIdx.x; c < ncell; c += kBinThreads) cnt[c] = ccount[d.cell0 + c];
  __syncthreads();
  block_scan_lds(cnt, ncell, wsum);
  for (int c = threadIdx.x; c <= ncell; c += kBinThreads) cstart[d.cell0 + c] = cnt[c];
}

// Slot order of the new binning: local index of each particle at its cell's
// start plus its rank.
__global__ __launch_bounds__(kPartThreads) void kb_scatter(const BlockMap* __restrict__ bm,
                                                           const RepDesc* __restrict__ rd,
                                                           const RepState* __restrict__ st,
                                                           const int* __restrict__ cstart,
                                                           const int* __restrict__ tcell,
                                                           const int* __restrict__ trank, int* __restrict__ order) {
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild || threadIdx.x >= (unsigned)m.count) return;
  const RepDesc& d = rd[r];
  const int s = m.s0 + threadIdx.x;
  order[d.slot0 + cstart[d.cell0 + tcell[s]] + trank[s]] = s - d.slot0;
}

// Each cell ordered by x, ties by particle id, so that the slot order does not
// depend on the atomics' order and a row of cells is x-monotone for the list
// build's clipping: one wavefront per cell, up to four entries per lane, ranks
// by comparison.  A cell map block covers kSortCells cells of one replica.  An
// empty cell's x sub-bin table (kb_xbins fills the others) is its start.
constexpr int kSortCells = 16;
__global__ __launch_bounds__(kPartThreads) void kb_sort_cells(const BlockMap* __restrict__ cm,
                                                              const RepDesc* __restrict__ rd,
                                                              RepState* __restrict__ st,
                                                              const int* __restrict__ cstart,
                                                              const double4* __restrict__ pos,
                                                              const int* __restrict__ id, int* __restrict__ order,
                                                              int* __restrict__ xbin) {
  const BlockMap m = cm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild) return;
  const RepDesc& d = rd[r];
  const int s0 = d.slot0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int cc = wv; cc < m.count; cc += kPartThreads / 64) {
    const int c = m.s0 + cc;                      // cell index within the replica
    const int b = cstart[d.cell0 + c], mlen = cstart[d.cell0 + c + 1] - b;
    if (mlen == 0 && lane <= kXBins) xbin[(size_t)(d.cell0 + c) * (kXBins + 1) + lane] = s0 + b;
    if (mlen <= 1) continue;
    if (mlen > kMaxCellFill) {
      if (lane == 0) st[r].cellfull = 1;
      continue;
    }
    int ent[4], key[4];
    long long kx[4];                              // x_key of the folded x
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = lane + 64 * u;
      ent[u] = k < mlen ? order[s0 + b + k] : 0;
      key[u] = k < mlen ? id[s0 + ent[u]] : INT_MAX;
      kx[u] = k < mlen ? x_key(pos[s0 + ent[u]].x) : LLONG_MAX;
    }
    int rk[4] = {0, 0, 0, 0};
    if (mlen <= 64) {
      // one entry per lane (the common case)
      for (int q = 0; q < mlen; ++q) {
        const int other = __shfl(key[0], q, 64);
        const long long ox = __shfl(kx[0], q, 64);
        rk[0] += ox < kx[0] || (ox == kx[0] && (other < key[0] || (other == key[0] && q < lane)));
      }
    } else {
#pragma unroll
      for (int u2 = 0; u2 < 4; ++u2) {
        if (64 * u2 >= mlen) break;     // wavefront-uniform
        for (int q = 0; q < 64; ++q) {
          const int other = __shfl(key[u2], q, 64);
          const long long ox = __shfl(kx[u2], q, 64);
          const int ok = q + 64 * u2;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = lane + 64 * u;
            // ids are unique within a replica; ties (padding) keep the entry index order
            rk[u] += (ok < mlen) &&
                     (ox < kx[u] || (ox == kx[u] && (other < key[u] || (other == key[u] && ok < k))));
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (lane + 64 * u < mlen) order[s0 + b + rk[u]] = ent[u];
  }
}

// Gather the particle arrays into the new slot order (second buffers).
__global__ __launch_bounds__(kPartThreads) void kb_permute(const BlockMap* __restrict__ bm,
                                                           const RepDesc* __restrict__ rd,
                                                           const RepState* __restrict__ st,
                                                           const int* __restrict__ order,
                                                           const double4* __restrict__ pos,
                                                           const double4* __restrict__ vel,
                                                           const int* __restrict__ id, const int4* __restrict__ img,
                                                           const int* __restrict__ tcell, double4* __restrict__ pos2,
                                                           double4* __restrict__ vel2, int* __restrict__ id2,
                                                           int4* __restrict__ img2, int* __restrict__ cell_of_slot) {
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild || threadIdx.x >= (unsigned)m.count) return;
  const int s = m.s0 + threadIdx.x;
  const int o = rd[r].slot0 + order[s];
  pos2[s] = pos[o];
  vel2[s] = vel[o];
  id2[s] = id[o];
  img2[s] = img[o];
  cell_of_slot[s] = tcell[o];
}

// ... and back into the primary buffers, with the fixed-point mirror, the FP32
// copy and, where the replica's box allows them, the 16-bit copies.
__global__ __launch_bounds__(kPartThreads) void kb_copy_back(const BlockMap* __restrict__ bm,
                                                             const RepDesc* __restrict__ rd,
                                                             const RepState* __restrict__ st,
                                                             const double4* __restrict__ pos2,
                                                             const double4* __restrict__ vel2,
                                                             const int* __restrict__ id2,
                                                             const int4* __restrict__ img2, double4* __restrict__ pos,
                                                             double4* __restrict__ vel, int* __restrict__ id,
                                                             int4* __restrict__ img, uint4* __restrict__ posq,
                                                             float4* __restrict__ posf, uint2* __restrict__ posh) {
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild || threadIdx.x >= (unsigned)m.count) return;
  const int s = m.s0 + threadIdx.x;
  const double4 p = pos2[s];
  const double qinv = rd[r].qinv;
  pos[s] = p;
  posf[s] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
  if (rd[r].q16inv > 0.0) posh[s] = q16_pack(p, rd[r].q16inv);
  vel[s] = vel2[s];
  id[s] = id2[s];
  img[s] = img2[s];
  posq[s] = make_uint4(fixq(p.x, qinv), fixq(p.y, qinv), fixq(p.z, qinv), (uint32_t)(int)p.w);
}

// Per cell of a rebuilding replica, xbin[(cell0 + c) (kXBins+1) + k] is the
// first slot of the cell whose x sub-bin is at least k (k < kXBins), and the
// cell's end for k = kXBins (absolute slots).  Cells are sorted by x, so the
// sub-bins rise along a cell: slot s writes the entries from its predecessor's
// sub-bin (exclusive) to its own, the last slot of a cell the entries above
// its own (kb_sort_cells fills the empty cells).
__global__ __launch_bounds__(kPartThreads) void kb_xbins(const BlockMap* __restrict__ bm,
                                                         const RepDesc* __restrict__ rd,
                                                         const RepState* __restrict__ st,
                                                         const int* __restrict__ cell_of_slot,
                                                         const float4* __restrict__ posf, int* __restrict__ xbin) {
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild || threadIdx.x >= (unsigned)m.count) return;
  const RepDesc& d = rd[r];
  const int s = m.s0 + threadIdx.x;
  const int nc = d.nc;
  const float inv_cs = 1.0f / (float)(d.L / nc);
  const int c = cell_of_slot[s];
  const int cx = c % nc;
  int* out = xbin + (size_t)(d.cell0 + c) * (kXBins + 1);
  const int k = xsub_bin(posf[s].x, inv_cs, cx);
  const int kp = (s > d.slot0 && cell_of_slot[s - 1] == c) ? xsub_bin(posf[s - 1].x, inv_cs, cx) : -1;
  for (int q = kp + 1; q <= k; ++q) out[q] = s;
  if (s + 1 == d.slot0 + d.n || cell_of_slot[s + 1] != c)
    for (int q = k + 1; q <= kXBins; ++q) out[q] = s + 1;
}

// Full Verlet list of every replica that rebuilt this step.  Membership is the
// reference's half-list test (src/VerletList.cpp:247-276): d^2 <= (rc+skin)^2 in
// FP64 without contraction, d in the orientation of the half-shell visit (cell
// offset of j lexicographically positive, z most significant: d = xi - (xj + sL),
// else d = xj - (xi - sL)).  Entry k of slot s: nbr[((s/64)*cap + k)*64 + s%64].
//
// A particle scans 9 neighbour x-rows of three cells, each one contiguous slot
// range (two where the row wraps).  The binning has just folded every position
// into its cell, so a row whose (y, z) range lies beyond the prefilter's top is
// skipped and the others keep the x reach the y/z distance leaves; cells are
// sorted by x, and the x sub-bin table (kb_xbins) cuts each range to a stretch
// around x_i +- reach.  Sixteen candidates at a time, the FP32 copies form two
// bit masks without a branch (r^2 below the window: in; below its top: in or
// undecided); the undecided take the reference's FP64 test.
__global__ __launch_bounds__(kPartThreads) __attribute__((amdgpu_waves_per_eu(6))) void kb_build(
    const BlockMap* __restrict__ bm, const RepDesc* __restrict__ rd, RepState* __restrict__ st,
    const double4* __restrict__ pos, const float4* __restrict__ posf, const int* __restrict__ cell_of_slot,
    const int* __restrict__ cstart, const int* __restrict__ xbin, const uint2* __restrict__ posh,
    int* __restrict__ nbr, int* __restrict__ cnt, int cap) {
  constexpr int kChunk = 16;
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  if (!st[r].rebuild || st[r].nanpos || st[r].cellfull) return;
  const RepDesc& d = rd[r];
  const int t = threadIdx.x;
  int k = 0;
  if (t < m.count) {
    const int i = m.s0 + t;
    const double4 pi = pos[i];
    const int nc = d.nc, s0 = d.slot0;
    const int c = cell_of_slot[i];
    const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    int* out = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    const double L = d.L, cutsq = d.cutsq;
    const float cflo = (float)d.cflo, cfhi = (float)d.cfhi, win = cfhi - cflo;
    const float cs = (float)(L / nc), inv_cs = 1.0f / cs;
    // 1e-4 of a cell covers the FP32 rounding of the bounds and of the sub-bins
    const float bmargin = 1e-4f * cs;
    const float top = sqrtf(cfhi) * 1.0001f, top2 = top * top;
    // the 16-bit prefilter (replicas whose box allows it): particle i's copy, the
    // same for every row, and the thresholds
    const bool q16 = d.q16inv > 0.0;
    v2u16 lxy = {0, 0}, lz = {0, 0};
    if (q16) {
      const uint2 h = q16_pack(pi, d.q16inv);
      lxy = __builtin_bit_cast(v2u16, h.x);
      lz = __builtin_bit_cast(v2u16, h.y);
    }
    const int neg_tlo = -d.q16_tlo, q16_w = d.q16_w;
    for (int oz = -1; oz <= 1; ++oz) {
      int qz = cz + oz, sz = 0;
      if (qz < 0) { qz += nc; sz = -1; } else if (qz >= nc) { qz -= nc; sz = 1; }
      // particle i moved by minus the candidate cells' periodic shift, in FP32
      const float zs = (float)(pi.z - (double)sz * L);
      const float zlo = (float)qz * cs;
      const float dz_lo = fmaxf(fmaxf(zlo - zs, zs - (zlo + cs)) - bmargin, 0.0f);
      for (int oy = -1; oy <= 1; ++oy) {
        int qy = cy + oy, sy = 0;
        if (qy < 0) { qy += nc; sy = -1; } else if (qy >= nc) { qy -= nc; sy = 1; }
        const float ys = (float)(pi.y - (double)sy * L);
        const float ylo = (float)qy * cs;
        const float dy_lo = fmaxf(fmaxf(ylo - ys, ys - (ylo + cs)) - bmargin, 0.0f);
        const float rx2 = top2 - __builtin_fmaf(dy_lo, dy_lo, dz_lo * dz_lo);
        if (!(rx2 > 0.0f)) continue;
        const float reach = sqrtf(rx2) * 1.0001f;
        const int row = d.cell0 + (qy + nc * qz) * nc;
        // x pieces of the row segment (cx-1 .. cx+1), split where the row wraps
        int pa0, pb0, pa1 = 0, pb1 = -1;
        if (cx >= 1 && cx + 1 < nc) { pa0 = cx - 1; pb0 = cx + 1; }
        else if (cx == 0) { pa0 = nc - 1; pb0 = nc - 1; pa1 = 0; pb1 = 1; }
        else { pa0 = cx - 1; pb0 = nc - 1; pa1 = 0; pb1 = 0; }
        for (int p = 0; p < 2; ++p) {
          const int xa = p ? pa1 : pa0, xb = p ? pb1 : pb0;
          if (xb < xa) break;
          const int sxp = (cx == 0 && p == 0) ? -1 : ((cx == nc - 1 && p == 1) ? 1 : 0);
          const float xs = (float)(pi.x - (double)sxp * L);
          const int jb = s0 + cstart[row + xa], je0 = s0 + cstart[row + xb + 1];
          int ja, je;
          {
            const float ulo = (xs - reach - bmargin) * inv_cs;
            const float uhi = (xs + reach + bmargin) * inv_cs;
            if (!(ulo > (float)xa)) {
              ja = jb;
            } else if (ulo >= (float)(xb + 1)) {
              ja = je0;
            } else {
              const int cc = min((int)floorf(ulo), xb);
              const int kb = min(max((int)floorf((ulo - (float)cc) * (float)kXBins), 0), kXBins - 1);
              ja = xbin[(size_t)(row + cc) * (kXBins + 1) + kb];
            }
            if (!(uhi < (float)(xb + 1))) {
              je = je0;
            } else if (uhi < (float)xa) {
              je = jb;
            } else {
              const int cc = max((int)floorf(uhi), xa);
              const int kb = min(max((int)floorf((uhi - (float)cc) * (float)kXBins), 0) + 1, kXBins);
              je = xbin[(size_t)(row + cc) * (kXBins + 1) + kb];
            }
            je = max(je, ja);
          }
          for (int j0 = ja; j0 < je; j0 += kChunk) {
            const int jn = je - j0;
            uint32_t in_bits = 0, maybe_bits = 0;
            // candidates kChunk-1 .. 0, so that candidate u ends at bit u; each test
            // shifts in the sign bit of r^2 below the window (in) and below its top
            if (q16) {
              // 16-bit units: two packed differences and two dot products
#pragma unroll
              for (int u = kChunk - 4; u >= 0; u -= 4) {
                uint2 q[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) q[v] = posh[j0 + u + v];
#pragma unroll
                for (int v = 3; v >= 0; --v) {
                  const v2i16 dxy = __builtin_bit_cast(v2i16, lxy - __builtin_bit_cast(v2u16, q[v].x));
                  const v2i16 dz = __builtin_bit_cast(v2i16, lz - __builtin_bit_cast(v2u16, q[v].y));
                  const int f = sdot2_acc(dz, sdot2_sacc(dxy, neg_tlo));
                  in_bits = __builtin_amdgcn_alignbit(in_bits, (uint32_t)f, 31);
                  maybe_bits = __builtin_amdgcn_alignbit(maybe_bits, (uint32_t)(f - q16_w), 31);
                }
              }
            } else {
#pragma unroll
              for (int u = kChunk - 4; u >= 0; u -= 4) {
                float4 q[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) q[v] = posf[j0 + u + v];
#pragma unroll
                for (int v = 3; v >= 0; --v) {
                  const float dx = xs - q[v].x, dy = ys - q[v].y, dz = zs - q[v].z;
                  const float f = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, -cflo)));
                  in_bits = __builtin_amdgcn_alignbit(in_bits, __float_as_uint(f), 31);
                  maybe_bits = __builtin_amdgcn_alignbit(maybe_bits, __float_as_uint(f - win), 31);
                }
              }
            }
            uint32_t valid = jn >= kChunk ? 0xFFFFu : ((1u << jn) - 1u);
            const int self = i - j0;
            if (self >= 0 && self < kChunk) valid &= ~(1u << self);
            in_bits &= valid;
            uint32_t border = maybe_bits & valid & ~in_bits;
            while (border) {
              // the reference's FP64 test in the reference's orientation (rare)
              const int tb = __ffs(border) - 1;
              border &= border - 1u;
              const int j = j0 + tb;
              const int cjx = cell_of_slot[j] % nc;
              int ox = cjx - cx, sx = 0;
              if (ox > 1) { ox = -1; sx = -1; } else if (ox < -1) { ox = 1; sx = 1; }
              const bool fwd = (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
              const double4 pj = pos[j];
              const double shx = (double)sx * L, shy = (double)sy * L, shz = (double)sz * L;
              double ex, ey, ez;
              if (fwd) { ex = pi.x - (pj.x + shx); ey = pi.y - (pj.y + shy); ez = pi.z - (pj.z + shz); }
              else { ex = pj.x - (pi.x - shx); ey = pj.y - (pi.y - shy); ez = pj.z - (pi.z - shz); }
              if (sqr_nc(ex, ey, ez) <= cutsq) in_bits |= 1u << tb;
            }
            // up to four entries per pass, so the wavefront's passes follow the
            // largest count of its lanes divided by four
            while (in_bits) {
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                if (in_bits) {
                  const int tb = __ffs(in_bits) - 1;
                  in_bits &= in_bits - 1u;
                  if (k < cap) out[(size_t)k * 64] = j0 + tb;
                  ++k;
                }
              }
            }
          }
        }
      }
    }
    cnt[i] = k;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) k = max(k, __shfl_xor(k, o, 64));
  if ((t & 63) == 0 && k > 0) {
    atomicMax(&st[r].maxcount, k);
    if (k > cap) st[r].overflow = 1;
  }
}

// |d|^2 of a listed pair as the reference's list and force loops evaluate it
// (orientation of the half-shell visit, ghost at x + sL, no contraction).  Used
// only for pairs within a few fixed-point quanta of the cutoff.
__device__ double ref_r2(const double4& pi, const double4& pj, int ci, int cj, const RepDesc& d) {
#pragma clang fp contract(off)
  const int sx = (int)__builtin_rint((pi.x - pj.x) * d.invL);
  const int sy = (int)__builtin_rint((pi.y - pj.y) * d.invL);
  const int sz = (int)__builtin_rint((pi.z - pj.z) * d.invL);
  const int nc = d.nc;
  const int ox = cj % nc + sx * nc - ci % nc;
  const int oy = (cj / nc) % nc + sy * nc - (ci / nc) % nc;
  const int oz = cj / (nc * nc) + sz * nc - ci / (nc * nc);
  const bool fwd = (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
  const double shx = (double)sx * d.L, shy = (double)sy * d.L, shz = (double)sz * d.L;
  double ex, ey, ez;
  if (fwd) { ex = pi.x - (pj.x + shx); ey = pi.y - (pj.y + shy); ez = pi.z - (pj.z + shz); }
  else { ex = pj.x - (pi.x - shx); ey = pj.y - (pi.y - shy); ez = pj.z - (pi.z - shz); }
  return ex * ex + ey * ey + ez * ez;
}

// Mixed-precision LJ over the full list (src/interaction/LennardJones.hpp:129-136;
// energy :109-114), lpp lanes per particle: lane q of a particle takes entries
// q, q + lpp, ...; the lanes' FP64 sums are combined by shuffles in a fixed
// order.  Distances are exact integer differences of the 32-bit fixed-point
// mirror (minimum image for free); pairs within the cutoff window re-decide the
// test from the FP64 positions in the reference's orientation.  ENERGY writes
// per-workgroup partial sums {energy, virial} instead of forces.
template <bool ENERGY>
__global__ __launch_bounds__(kPartThreads) void kb_force(const BlockMap* __restrict__ bm,
                                                         const RepDesc* __restrict__ rd,
                                                         const RepState* __restrict__ st,
                                                         const uint4* __restrict__ posq,
                                                         const double4* __restrict__ pos,
                                                         const int* __restrict__ cell_of_slot,
                                                         const int* __restrict__ nbr, const int* __restrict__ cnt,
                                                         int cap, double* __restrict__ fx, double* __restrict__ fy,
                                                         double* __restrict__ fz, double* __restrict__ partial) {
  const BlockMap m = bm[blockIdx.x];
  const int r = m.rep;
  const RepDesc& d = rd[r];
  const int t = threadIdx.x;
  const int lpp = m.lpp;
  const int p = t / lpp, q = t - p * lpp;
  const bool live = p < m.count;
  const int i = m.s0 + (live ? p : 0);
  double Ax = 0.0, Ay = 0.0, Az = 0.0, E = 0.0, W = 0.0;
  if (live) {
    const uint4 qi = posq[i];
    const int* np = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    const int c = min(cnt[i], cap);
    // In fixed-point units: u = mirror difference (exact), u2 = |u|^2, F = u q ff_u.
    // u2 below ulo is inside the cutoff and u2 above uhi outside; a pair in
    // between is left out of the pass and decided by the reference's FP64 test
    // after it (rare).
    const double uff1 = d.uff1, uff2 = d.uff2, ulo = d.ulo, uhi = d.uhi, rc2 = d.rc2;
    auto contrib = [&](double ux, double uy, double uz, double u2, double ff, double inv) {
      Ax = __builtin_fma(ux, ff, Ax);
      Ay = __builtin_fma(uy, ff, Ay);
      Az = __builtin_fma(uz, ff, Az);
      if (ENERGY) {
        const double f2 = d.usig2 * inv;
        const double f6 = f2 * f2 * f2;
        E += d.eps4 * (f6 * f6 - f6) - d.shift;
        W += u2 * ff;
      }
    };
    auto pair = [&](const uint4& qj, bool& near) {
      const double ux = (double)(int)(qi.x - qj.x);
      const double uy = (double)(int)(qi.y - qj.y);
      const double uz = (double)(int)(qi.z - qj.z);
      const double u2 = ux * ux + uy * uy + uz * uz;
      const bool in = u2 < ulo;
      near |= (u2 <= uhi) & !in;
      double inv;
      const double ff = lj_ff_masked(u2, in, uff1, uff2, inv);
      Ax = __builtin_fma(ux, ff, Ax);
      Ay = __builtin_fma(uy, ff, Ay);
      Az = __builtin_fma(uz, ff, Az);
      if (ENERGY && in) {
        const double f2 = d.usig2 * inv;
        const double f6 = f2 * f2 * f2;
        E += d.eps4 * (f6 * f6 - f6) - d.shift;
        W += u2 * ff;
      }
    };
    auto settle = [&](int j, const uint4& qj) {
      const double ux = (double)(int)(qi.x - qj.x);
      const double uy = (double)(int)(qi.y - qj.y);
      const double uz = (double)(int)(qi.z - qj.z);
      const double u2 = ux * ux + uy * uy + uz * uz;
      if (u2 < ulo || !(u2 <= uhi)) return;
      if (!(ref_r2(pos[i], pos[j], cell_of_slot[i], cell_of_slot[j], d) <= rc2)) return;
      double inv;
      const double ff = lj_ff_seeded(u2, uff1, uff2, inv);
      contrib(ux, uy, uz, u2, ff, inv);
    };
    // entries four at a time (indices, then positions, issued together); pairs at
    // the cutoff are added after their group of four
    // (the next group's indices are loaded while this group's positions arrive)
    int k = q;
    int jn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jn[u] = k + 3 * lpp < c ? np[(size_t)(k + u * lpp) * 64] : 0;
    for (; k + 3 * lpp < c; k += 4 * lpp) {
      int jj[4];
      uint4 qq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) jj[u] = jn[u];
#pragma unroll
      for (int u = 0; u < 4; ++u) qq[u] = posq[jj[u]];
      const int kn = k + 4 * lpp;
#pragma unroll
      for (int u = 0; u < 4; ++u) jn[u] = kn + 3 * lpp < c ? np[(size_t)(kn + u * lpp) * 64] : 0;
      bool near = false;
#pragma unroll
      for (int u = 0; u < 4; ++u) pair(qq[u], near);
      if (near) {
#pragma unroll
        for (int u = 0; u < 4; ++u) settle(jj[u], qq[u]);
      }
    }
    for (; k < c; k += lpp) {
      const int j = np[(size_t)k * 64];
      const uint4 qj = posq[j];
      bool near = false;
      pair(qj, near);
      if (near) settle(j, qj);
    }
    Ax *= d.qsd; Ay *= d.qsd; Az *= d.qsd;
    if (ENERGY) W *= d.qsd * d.qsd;
  }
  // combine the lpp lanes of a particle (consecutive lanes, power of two)
  for (int o = 1; o < lpp; o <<= 1) {
    Ax += __shfl_xor(Ax, o, 64);
    Ay += __shfl_xor(Ay, o, 64);
    Az += __shfl_xor(Az, o, 64);
  }
  if (!ENERGY) {
    if (live && q == 0) { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    return;
  }
  __shared__ double red[2][kPartThreads / 64];
  E = wave_sum(E);
  W = wave_sum(W);
  if ((t & 63) == 0) { red[0][t >> 6] = E; red[1][t >> 6] = W; }
  __syncthreads();
  if (t == 0) {
    double e = 0.0, w = 0.0;
    for (int u = 0; u < (int)(blockDim.x >> 6); ++u) { e += red[0][u]; w += red[1][u]; }
    partial[2 * (size_t)blockIdx.x] = 0.5 * e;
    partial[2 * (size_t)blockIdx.x + 1] = 0.5 * w;
  }
}

// Kinetic partials per workgroup: sum m v^2 (src/analysis/Temperature.hpp:112-122).
__global__ __launch_bounds__(kPartThreads) void kb_kinetic(const BlockMap* __restrict__ bm,
                                                           const double4* __restrict__ vel,
                                                           double* __restrict__ partial) {
  const BlockMap m = bm[blockIdx.x];
  const int t = threadIdx.x;
  double e = 0.0;
  if (t < m.count) {
    const double4 v = vel[m.s0 + t];
    e = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
  }
  __shared__ double red[kPartThreads / 64];
  e = wave_sum(e);
  if ((t & 63) == 0) red[t >> 6] = e;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int u = 0; u < (int)(blockDim.x >> 6); ++u) s += red[u];
    partial[(size_t)blockIdx.x * 2] = s;
    partial[(size_t)blockIdx.x * 2 + 1] = 0.0;
  }
}

// Per replica, the sum of its workgroups' partials in workgroup order.
__global__ void kb_reduce_rep(int nrep, const int* __restrict__ blk0, const int* __restrict__ nblk,
                              const double* __restrict__ partial, double* __restrict__ out) {
  const int r = blockIdx.x;
  if (r >= nrep) return;
  const int lane = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int k = lane; k < nblk[r]; k += 64) {
    a += partial[2 * (size_t)(blk0[r] + k)];
    b += partial[2 * (size_t)(blk0[r] + k) + 1];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane == 0) { out[2 * r] = a; out[2 * r + 1] = b; }
}

__global__ void kb_clear_overflow(int nrep, RepState* __restrict__ st) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrep) { st[r].overflow = 0; st[r].maxcount = 0; }
}

// Device-to-device copy of `bytes` bytes as 16-byte words (snapshots).
__global__ void kb_copy16(size_t words, const uint4* __restrict__ src, uint4* __restrict__ dst) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < words; k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

}  // namespace batch
}  // namespace espp

// ------------------------------------------------------------------ host

using namespace espp;
using namespace espp::batch;
using namespace espp_host;

namespace {

int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

struct espp_batch {
  int R = 0;
  double rc = 2.5, skin = 0.3;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<RepDesc> hd;
  std::vector<double> temperature;     // thermostat temperature per replica
  double gamma = 0.0;
  std::vector<int> lpp;                // lanes per particle of the force kernel, per replica
  std::vector<std::vector<int>> ids;   // ids loaded per replica (for id-ordered access)
  std::vector<bool> loaded;
  int64_t total_slots = 0, total_cells = 0;
  int cap = 0;

  DBuf<RepDesc> drd;
  DBuf<RepState> dst, dst_snap;
  DBuf<BlockMap> bm_part, bm_force;
  int nb_part = 0, nb_force = 0;
  DBuf<int> fblk0, fnblk, pblk0, pnblk;   // per replica: first block and block count (force / particle maps)
  DBuf<double4> pos, vel, pos2, vel2, pos_snap, vel_snap;
  DBuf<int> id, id2, id_snap;
  DBuf<int4> img, img2, img_snap;
  DBuf<uint4> posq;
  DBuf<float4> posf;                      // FP32 copy of the positions at the last binning (list build)
  DBuf<uint2> posh;                       // 16-bit copies of the same (q16_pack; replicas with q16inv > 0)
  DBuf<double> fx, fy, fz;
  DBuf<int> cell_of_slot, cstart, ccount, tcell, trank, order;
  DBuf<int> xbin;                         // x sub-bin table of every cell ((kXBins + 1) slots each)
  DBuf<BlockMap> bm_cell;                 // cell blocks of kb_sort_cells
  int nb_cell = 0;
  DBuf<int> nbr, cnt;
  DBuf<double> partial, red;
  double* hred = nullptr;

  bool need_bin = true;       // particles were loaded or moved from the host since the last binning
  bool use_graph = true;      // replay captured steps (false: launch each step's kernels directly)
  double graph_dt = -1.0;
  hipGraphExec_t g1 = nullptr, g10 = nullptr;

  ~espp_batch() {
    if (g1) (void)hipGraphExecDestroy(g1);
    if (g10) (void)hipGraphExecDestroy(g10);
    if (hred) (void)hipHostFree(hred);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void sync() { hip_check(hipStreamSynchronize(stream), "batch stream synchronize"); }

  void drop_graphs() {
    if (g1) (void)hipGraphExecDestroy(g1);
    if (g10) (void)hipGraphExecDestroy(g10);
    g1 = g10 = nullptr;
  }

  void upload_desc() {
    hip_check(hipMemcpyAsync(drd.p, hd.data(), hd.size() * sizeof(RepDesc), hipMemcpyHostToDevice, stream),
              "descriptor upload");
    sync();
  }

  // Block maps: particle kernels take up to 256 slots of one replica per
  // workgroup; the force kernel 256 / lpp particles of one replica.
  void build_maps() {
    std::vector<BlockMap> mp, mf;
    std::vector<int> pb0(R), pnb(R), fb0(R), fnb(R);
    for (int r = 0; r < R; ++r) {
      const int n = hd[r].n, s0 = hd[r].slot0;
      pb0[r] = (int)mp.size();
      for (int k = 0; k < n; k += kPartThreads) mp.push_back({r, s0 + k, std::min(kPartThreads, n - k), 1});
      pnb[r] = (int)mp.size() - pb0[r];
      const int per = kPartThreads / lpp[r];
      fb0[r] = (int)mf.size();
      for (int k = 0; k < n; k += per) mf.push_back({r, s0 + k, std::min(per, n - k), lpp[r]});
      fnb[r] = (int)mf.size() - fb0[r];
    }
    std::vector<BlockMap> mc;
    for (int r = 0; r < R; ++r) {
      const int nc3 = hd[r].nc * hd[r].nc * hd[r].nc;
      for (int c = 0; c < nc3; c += kSortCells) mc.push_back({r, c, std::min(kSortCells, nc3 - c), 1});
    }
    nb_cell = (int)mc.size();
    bm_cell.resize(mc.size());
    nb_part = (int)mp.size();
    nb_force = (int)mf.size();
    bm_part.resize(mp.size());
    bm_force.resize(mf.size());
    pblk0.resize(R); pnblk.resize(R); fblk0.resize(R); fnblk.resize(R);
    partial.resize(2 * (size_t)std::max(nb_part, nb_force));
    auto up = [&](void* d, const void* h, size_t b) {
      hip_check(hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, stream), "block map upload");
    };
    up(bm_part.p, mp.data(), mp.size() * sizeof(BlockMap));
    up(bm_force.p, mf.data(), mf.size() * sizeof(BlockMap));
    up(bm_cell.p, mc.data(), mc.size() * sizeof(BlockMap));
    up(pblk0.p, pb0.data(), R * sizeof(int));
    up(pnblk.p, pnb.data(), R * sizeof(int));
    up(fblk0.p, fb0.data(), R * sizeof(int));
    up(fnblk.p, fnb.data(), R * sizeof(int));
    sync();
    drop_graphs();
  }

  // the skin check and, for the replicas that rebuild, the binning
  void launch_check_bin(int use_step) {
    hipLaunchKernelGGL(kb_check, dim3(R), dim3(kBinThreads), 0, stream, drd.p, dst.p, use_step, ccount.p);
    hipLaunchKernelGGL(kb_fold_count, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, pos.p,
                       img.p, ccount.p, tcell.p, trank.p);
    hipLaunchKernelGGL(kb_scan, dim3(R), dim3(kBinThreads), 0, stream, drd.p, dst.p, ccount.p, cstart.p);
    hipLaunchKernelGGL(kb_scatter, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, cstart.p,
                       tcell.p, trank.p, order.p);
    hipLaunchKernelGGL(kb_sort_cells, dim3(nb_cell), dim3(kPartThreads), 0, stream, bm_cell.p, drd.p, dst.p,
                       cstart.p, pos.p, id.p, order.p, xbin.p);
    hipLaunchKernelGGL(kb_permute, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, order.p,
                       pos.p, vel.p, id.p, img.p, tcell.p, pos2.p, vel2.p, id2.p, img2.p, cell_of_slot.p);
    hipLaunchKernelGGL(kb_copy_back, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, pos2.p,
                       vel2.p, id2.p, img2.p, pos.p, vel.p, id.p, img.p, posq.p, posf.p, posh.p);
    hipLaunchKernelGGL(kb_xbins, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p,
                       cell_of_slot.p, posf.p, xbin.p);
  }
  void launch_build() {
    hipLaunchKernelGGL(kb_build, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, pos.p,
                       posf.p, cell_of_slot.p, cstart.p, xbin.p, posh.p, nbr.p, cnt.p, cap);
  }
  void launch_force() {
    hipLaunchKernelGGL(kb_force<false>, dim3(nb_force), dim3(kPartThreads), 0, stream, bm_force.p, drd.p, dst.p,
                       posq.p, pos.p, cell_of_slot.p, nbr.p, cnt.p, cap, fx.p, fy.p, fz.p, nullptr);
  }
  void launch_step(double dt) {
    hipLaunchKernelGGL(kb_step, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, pos.p, vel.p,
                       posq.p, fx.p, fy.p, fz.p, id.p, dt);
  }
  void launch_rep_op(int op) {
    hipLaunchKernelGGL(kb_rep_op, dim3(nblk(R, 64)), dim3(64), 0, stream, R, drd.p, dst.p, op);
  }
  void launch_thermo_int2(double factor, int do_int2, double half_dt) {
    hipLaunchKernelGGL(kb_thermo_int2, dim3(nb_part), dim3(kPartThreads), 0, stream, bm_part.p, drd.p, dst.p, vel.p,
                       fx.p, fy.p, fz.p, id.p, factor, do_int2, half_dt);
  }
  // one velocity-Verlet step of every replica
  void enqueue_step(double dt) {
    launch_step(dt);
    launch_check_bin(1);
    launch_build();
    launch_force();
  }

  hipGraphExec_t capture(int steps, double dt) {
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int k = 0; k < steps; ++k) enqueue_step(dt);
    hip_check(hipStreamEndCapture(stream, &g), "end capture");
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(err, "graph instantiate");
    return e;
  }

  void ensure_graphs(double dt) {
    if (g1 && g10 && graph_dt == dt) return;
    drop_graphs();
    sync();
    g1 = capture(1, dt);
    g10 = capture(10, dt);
    graph_dt = dt;
  }

  void check_all_loaded() {
    for (int r = 0; r < R; ++r)
      if (!loaded[(size_t)r]) fail(ESPP_ERR_STATE, "replica " + std::to_string(r) + " has no particles loaded");
  }

  std::vector<RepState> read_state() {
    std::vector<RepState> h((size_t)R);
    hip_check(hipMemcpyAsync(h.data(), dst.p, R * sizeof(RepState), hipMemcpyDeviceToHost, stream), "state read");
    sync();
    return h;
  }

  void copy16(const void* src, void* dst_, size_t bytes) {
    const size_t words = bytes / 16;
    if (words) hipLaunchKernelGGL(kb_copy16, dim3(1024), dim3(256), 0, stream, words, (const uint4*)src, (uint4*)dst_);
  }

  void snapshot(bool restore) {
    const size_t S = (size_t)total_slots;
    auto cp = [&](void* live, void* snap, size_t bytes) { restore ? copy16(snap, live, bytes) : copy16(live, snap, bytes); };
    cp(pos.p, pos_snap.p, S * sizeof(double4));
    cp(vel.p, vel_snap.p, S * sizeof(double4));
    cp(img.p, img_snap.p, S * sizeof(int4));
    cp(id.p, id_snap.p, ((S * sizeof(int) + 15) / 16) * 16);
    cp(dst.p, dst_snap.p, ((R * sizeof(RepState) + 15) / 16) * 16);
  }

  // Forces at run start (rebinning first when particles came from the host).
  void start_forces() {
    if (need_bin) {
      launch_rep_op(2);
      launch_check_bin(0);
      launch_build();
      need_bin = false;
    }
    launch_force();
  }

  void grow_cap(const std::vector<RepState>& h) {
    int mx = 0;
    for (auto& s : h) mx = std::max(mx, s.maxcount);
    cap = ((int)std::ceil(mx * 1.25) + 8 + 7) & ~7;
    nbr.resize((size_t)(total_slots / 64) * (size_t)cap * 64);
    drop_graphs();
  }

  void check_state(const std::vector<RepState>& h) {
    for (int r = 0; r < R; ++r) {
      if (h[(size_t)r].nanpos)
        fail(ESPP_ERR_NAN, "replica " + std::to_string(r) +
                               ": particle has moved to outer space (one or more coordinates are nan or infinite)");
      if (h[(size_t)r].cellfull)
        fail(ESPP_ERR_STATE, "replica " + std::to_string(r) + ": more than " + std::to_string(kMaxCellFill) +
                                 " particles in one cell");
    }
  }

  // VelocityVerlet::run for every replica (src/integrator/VelocityVerlet.cpp:63-169):
  // forces with the thermostat's heat-up, nsteps graph-replayed steps, the
  // thermostat and integrate2 of the last step.  A list that overflowed its
  // capacity is detected after the run; the run is then repeated from a
  // snapshot with a larger capacity (the Philox noise makes it identical).
  void run(int nsteps, double dt) {
    check_all_loaded();
    for (int attempt = 0; attempt < 4; ++attempt) {
      if (use_graph && nsteps > 0) ensure_graphs(dt);
      snapshot(false);
      start_forces();
      launch_thermo_int2(std::sqrt(3.0), 0, 0.5 * dt);
      launch_rep_op(1);
      launch_rep_op(0);
      if (nsteps > 0) {
        if (use_graph) {
          ensure_graphs(dt);
          int k = 0;
          for (; k + 10 <= nsteps; k += 10) hip_check(hipGraphLaunch(g10, stream), "graph launch");
          for (; k < nsteps; ++k) hip_check(hipGraphLaunch(g1, stream), "graph launch");
        } else {
          for (int k = 0; k < nsteps; ++k) enqueue_step(dt);
        }
        launch_thermo_int2(1.0, 1, 0.5 * dt);
        launch_rep_op(1);
      }
      hip_check(hipGetLastError(), "batch launch");
      auto h = read_state();
      check_state(h);
      bool over = false;
      for (auto& s : h) over |= s.overflow != 0;
      if (!over) return;
      // list capacity exceeded: restore, enlarge, run again
      snapshot(true);
      need_bin = true;
      grow_cap(h);
      hipLaunchKernelGGL(kb_clear_overflow, dim3(nblk(R, 64)), dim3(64), 0, stream, R, dst.p);
    }
    fail(ESPP_ERR_STATE, "neighbour list capacity could not be sized");
  }
};

namespace {

// Time of one force launch over every replica (HIP events on the batch stream,
// averaged over reps launches after one untimed launch).
double time_force(espp_batch* b, int reps) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  b->launch_force();
  hip_check(hipEventRecord(e0, b->stream), "hipEventRecord");
  for (int k = 0; k < reps; ++k) b->launch_force();
  hip_check(hipEventRecord(e1, b->stream), "hipEventRecord");
  hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

// Time of one list build over every replica: every replica is rebuilt (a
// requested rebuild) and binned once, one build is untimed, then reps builds
// from that binning are timed with HIP events.
double time_build(espp_batch* b, int reps) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  b->launch_rep_op(2);
  b->launch_check_bin(0);
  b->launch_build();
  hip_check(hipEventRecord(e0, b->stream), "hipEventRecord");
  for (int k = 0; k < reps; ++k) b->launch_build();
  hip_check(hipEventRecord(e1, b->stream), "hipEventRecord");
  hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

espp_batch* checked(espp_batch* b) {
  if (!b) fail(ESPP_ERR_ARG, "NULL batch handle");
  return b;
}

void check_rep(espp_batch* b, int32_t rep) {
  if (rep < 0 || rep >= b->R) fail(ESPP_ERR_ARG, "replica index out of range");
}

}  // namespace

extern "C" {

int espp_batch_create(int32_t nrep, const int32_t* n, const double* box, double rc, double skin, double eps,
                      double sigma, double gamma, const double* temperature, const uint64_t* seed, espp_batch** out) {
  return guarded([&] {
    need(out && n && box && temperature && seed, "NULL argument");
    *out = nullptr;
    need(nrep >= 1, "at least one replica");
    need(rc > 0.0 && std::isfinite(rc) && skin >= 0.0 && std::isfinite(skin), "cutoff must be positive, skin >= 0");
    need(gamma >= 0.0, "gamma must be non-negative");
    int dc = 0;
    if (hipGetDeviceCount(&dc) != hipSuccess || dc == 0) {
      (void)hipGetLastError();
      fail(ESPP_ERR_DEVICE, "no HIP device available: the MI355X path needs a GPU");
    }
    std::unique_ptr<espp_batch> b(new espp_batch());
    b->R = nrep;
    b->rc = rc;
    b->skin = skin;
    b->gamma = gamma;
    hip_check(hipGetDevice(&b->device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    b->hd.resize((size_t)nrep);
    b->temperature.assign(temperature, temperature + nrep);
    b->lpp.assign((size_t)nrep, 1);
    b->ids.resize((size_t)nrep);
    b->loaded.assign((size_t)nrep, false);
    const double cutV = rc + skin;
    // LennardJones::preset and setAutoShift (src/interaction/LennardJones.hpp:79-87,
    // src/interaction/Potential.hpp:214-222)
    const double sig2 = sigma * sigma, sig6 = sig2 * sig2 * sig2;
    const double frac2 = sig2 / (rc * rc), frac6 = frac2 * frac2 * frac2;
    int64_t slot = 0, cell = 0;
    double max_density = 0.0;
    for (int r = 0; r < nrep; ++r) {
      need(n[r] >= 1, "every replica needs at least one particle");
      need(box[r] > 0.0 && std::isfinite(box[r]), "box lengths must be positive and finite");
      RepDesc& d = b->hd[(size_t)r];
      std::memset(&d, 0, sizeof(d));
      d.L = box[r];
      d.invL = 1.0 / box[r];
      // cellGrid = int(L / (rc + skin)) (src/tools/decomp.py:328-342)
      d.nc = (int)(box[r] / cutV);
      if (d.nc < 3)
        fail(ESPP_ERR_ARG, "replica " + std::to_string(r) + ": the box must hold three cells of cutoff + skin");
      if ((int64_t)d.nc * d.nc * d.nc > kMaxBinCells)
        fail(ESPP_ERR_ARG, "replica " + std::to_string(r) + ": more than " + std::to_string(kMaxBinCells) +
                               " cells (the binning keeps a replica's cells in LDS)");
      d.ics = 1.0 / ((box[r] - 0.0) / (double)d.nc);
      d.qinv = 4294967296.0 / box[r];
      d.qsd = box[r] / 4294967296.0;
      d.skin_half = 0.5 * skin;
      d.cutsq = cutV * cutV;
      d.ff1 = 48.0 * eps * sig6 * sig6;
      d.ff2 = 24.0 * eps * sig6;
      d.rc2 = rc * rc;
      d.shift = 4.0 * eps * (frac6 * frac6 - frac6);
      d.sig2 = sig2;
      d.eps4 = 4.0 * eps;
      const double win = 4.0 * std::sqrt(3.0) * d.qsd * rc + 1e-300;
      d.dlo = d.rc2 - win;
      d.dhi = d.rc2 + win;
      // the list build's FP32 prefilter: its coordinates (a folded position, or one
      // shifted by a box length) are below M = 2L in magnitude, each rounded once,
      // so a difference is off by at most one FP32 ulp of M and r^2 by at most
      // 2 sqrt(3) (rc+skin) ulp(M); the window takes twice that plus 1e-6 relative
      const double ulpM = std::ldexp(1.0, std::ilogb(2.0 * box[r]) - 23);
      const double cwin = 1e-6 + 4.0 * std::sqrt(3.0) * ulpM / cutV;
      d.cflo = d.cutsq * (1.0 - cwin);
      d.cfhi = d.cutsq * (1.0 + cwin);
      // the 16-bit prefilter where the box allows it (the list is built right after
      // the binning, so with no drift)
      d.q16inv = q16_scale_cubic(box[r], d.nc, 0.0, std::sqrt((double)(float)d.cfhi) * 1.0001);
      d.q16_tlo = d.q16_w = 0;
      if (d.q16inv > 0.0) q16_thresholds(cutV, d.q16inv, d.q16_tlo, d.q16_w);
      // the force launch works in fixed-point units: r^2 = u^2 q^2 (q = L / 2^32),
      // so ff(r^2) = u^-8 (ff1 q^-14 u^-6 - ff2 q^-8) and F = u q ff
      const double q2 = d.qsd * d.qsd;
      d.ulo = d.dlo / q2;
      d.uhi = d.dhi / q2;
      d.uff1 = d.ff1 / std::pow(q2, 7);
      d.uff2 = d.ff2 / std::pow(q2, 4);
      d.usig2 = d.sig2 / q2;
      d.pref1 = -gamma;
      d.seed = seed[r];
      d.lang = gamma > 0.0 ? 1 : 0;
      d.slot0 = (int)slot;
      d.n = n[r];
      d.cell0 = (int)cell;
      slot += ((int64_t)n[r] + 63) / 64 * 64;
      cell += (int64_t)d.nc * d.nc * d.nc + 1;
      need(slot < ((int64_t)1 << 31) - 64, "too many particles");
      max_density = std::max(max_density, n[r] / (box[r] * box[r] * box[r]));
    }
    b->total_slots = slot;
    b->total_cells = cell;
    // list capacity from the densest replica: 1.6 x the mean count within rc + skin, + 32
    const double nbar = 4.0 / 3.0 * M_PI * cutV * cutV * cutV * max_density;
    b->cap = ((int)std::ceil(1.6 * nbar + 32.0) + 7) & ~7;
    const size_t S = (size_t)slot;
    b->drd.resize((size_t)nrep);
    b->dst.resize((size_t)nrep + 1);
    b->dst_snap.resize((size_t)nrep + 1);
    b->pos.resize(S); b->vel.resize(S); b->pos2.resize(S); b->vel2.resize(S);
    b->pos_snap.resize(S); b->vel_snap.resize(S);
    b->id.resize(S + 4); b->id2.resize(S); b->id_snap.resize(S + 4);
    b->img.resize(S); b->img2.resize(S); b->img_snap.resize(S);
    b->posq.resize(S);
    b->posf.resize(S + 16);                // the list build reads up to 15 entries past a range
    b->posh.resize(S + 16);
    b->fx.resize(S); b->fy.resize(S); b->fz.resize(S);
    b->cell_of_slot.resize(S); b->tcell.resize(S); b->trank.resize(S); b->order.resize(S);
    b->cstart.resize((size_t)cell);
    b->ccount.resize((size_t)cell);
    b->xbin.resize((size_t)cell * (kXBins + 1));
    b->cnt.resize(S);
    b->nbr.resize((S / 64) * (size_t)b->cap * 64);
    b->red.resize(2 * (size_t)nrep);
    hip_check(hipHostMalloc((void**)&b->hred, 2 * (size_t)nrep * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc");
    std::vector<RepState> st((size_t)nrep + 1);
    std::memset(st.data(), 0, st.size() * sizeof(RepState));
    hip_check(hipMemcpyAsync(b->dst.p, st.data(), st.size() * sizeof(RepState), hipMemcpyHostToDevice, b->stream),
              "state upload");
    // slots past a replica's particles (up to the next multiple of 64) are never read
    hip_check(hipMemsetAsync(b->vel.p, 0, S * sizeof(double4), b->stream), "clear");
    hip_check(hipMemsetAsync(b->fx.p, 0, S * sizeof(double), b->stream), "clear");
    hip_check(hipMemsetAsync(b->fy.p, 0, S * sizeof(double), b->stream), "clear");
    hip_check(hipMemsetAsync(b->fz.p, 0, S * sizeof(double), b->stream), "clear");
    b->upload_desc();
    b->build_maps();
    *out = b.release();
  });
}

int espp_batch_destroy(espp_batch* b) {
  return guarded([&] { delete b; });
}

int espp_batch_load(espp_batch* b, int32_t rep, const int32_t* ids, const double* pos, const double* vel,
                    const double* mass) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    need(ids && pos, "NULL argument");
    const RepDesc& d = b->hd[(size_t)rep];
    const int n = d.n;
    std::vector<int> sorted(ids, ids + n);
    std::sort(sorted.begin(), sorted.end());
    for (int k = 0; k < n; ++k) {
      need(sorted[(size_t)k] >= 0, "particle ids must be non-negative");
      if (k && sorted[(size_t)k] == sorted[(size_t)k - 1]) fail(ESPP_ERR_ARG, "duplicate particle id in a replica");
    }
    std::vector<double4> p((size_t)n), v((size_t)n);
    std::vector<int4> im((size_t)n);
    for (int k = 0; k < n; ++k) {
      p[(size_t)k] = make_double4(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], 0.0);
      const double m = mass ? mass[k] : 1.0;
      need(m > 0.0, "masses must be positive");
      v[(size_t)k] = vel ? make_double4(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2], m) : make_double4(0, 0, 0, m);
      im[(size_t)k] = make_int4(0, 0, 0, 0);
    }
    b->sync();
    const size_t o = (size_t)d.slot0;
    hip_check(hipMemcpy(b->pos.p + o, p.data(), n * sizeof(double4), hipMemcpyHostToDevice), "load");
    hip_check(hipMemcpy(b->vel.p + o, v.data(), n * sizeof(double4), hipMemcpyHostToDevice), "load");
    hip_check(hipMemcpy(b->img.p + o, im.data(), n * sizeof(int4), hipMemcpyHostToDevice), "load");
    hip_check(hipMemcpy(b->id.p + o, ids, n * sizeof(int), hipMemcpyHostToDevice), "load");
    b->ids[(size_t)rep] = sorted;
    b->loaded[(size_t)rep] = true;
    b->need_bin = true;
  });
}

int espp_batch_set_temperature(espp_batch* b, int32_t rep, double temperature) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    need(temperature >= 0.0, "temperature must be non-negative");
    b->temperature[(size_t)rep] = temperature;
  });
}

int espp_batch_get_temperature(espp_batch* b, int32_t rep, double* temperature) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    need(temperature != nullptr, "NULL argument");
    *temperature = b->temperature[(size_t)rep];
  });
}

int espp_batch_run(espp_batch* b, int32_t nsteps, double dt) {
  return guarded([&] {
    checked(b);
    need(nsteps >= 0, "number of integration steps must be non-negative");
    need(dt > 0.0 && std::isfinite(dt), "time step must be positive");
    // LangevinThermostat::initialize: pref2 = sqrt(24 T gamma / dt)
    // (src/integrator/LangevinThermostat.cpp:187-188)
    bool changed = false;
    for (int r = 0; r < b->R; ++r) {
      const double p2 = std::sqrt(24.0 * b->temperature[(size_t)r] * b->gamma / dt);
      if (p2 != b->hd[(size_t)r].pref2) { b->hd[(size_t)r].pref2 = p2; changed = true; }
    }
    if (changed) b->upload_desc();
    b->run(nsteps, dt);
  });
}

// Per-replica potential energy and virial of the LJ interaction
// (VerletListInteractionTemplate::computeEnergy / computeVirial,
// src/interaction/VerletListInteractionTemplate.hpp:116-198,274-331).
int espp_batch_energy(espp_batch* b, double* energy, double* virial) {
  return guarded([&] {
    checked(b);
    need(energy != nullptr, "NULL argument");
    b->check_all_loaded();
    if (b->need_bin) {
      b->start_forces();
      auto h = b->read_state();
      b->check_state(h);
    }
    hipLaunchKernelGGL(kb_force<true>, dim3(b->nb_force), dim3(kPartThreads), 0, b->stream, b->bm_force.p, b->drd.p,
                       b->dst.p, b->posq.p, b->pos.p, b->cell_of_slot.p, b->nbr.p, b->cnt.p, b->cap, b->fx.p,
                       b->fy.p, b->fz.p, b->partial.p);
    hipLaunchKernelGGL(kb_reduce_rep, dim3(b->R), dim3(64), 0, b->stream, b->R, b->fblk0.p, b->fnblk.p, b->partial.p,
                       b->red.p);
    hip_check(hipMemcpyAsync(b->hred, b->red.p, 2 * (size_t)b->R * sizeof(double), hipMemcpyDeviceToHost, b->stream),
              "energy readback");
    b->sync();
    for (int r = 0; r < b->R; ++r) {
      energy[r] = b->hred[2 * r];
      if (virial) virial[r] = b->hred[2 * r + 1];
    }
  });
}

// Per-replica sum of m v^2 (src/analysis/Temperature.hpp:112-122).
int espp_batch_kinetic(espp_batch* b, double* mv2) {
  return guarded([&] {
    checked(b);
    need(mv2 != nullptr, "NULL argument");
    hipLaunchKernelGGL(kb_kinetic, dim3(b->nb_part), dim3(kPartThreads), 0, b->stream, b->bm_part.p, b->vel.p,
                       b->partial.p);
    hipLaunchKernelGGL(kb_reduce_rep, dim3(b->R), dim3(64), 0, b->stream, b->R, b->pblk0.p, b->pnblk.p, b->partial.p,
                       b->red.p);
    hip_check(hipMemcpyAsync(b->hred, b->red.p, 2 * (size_t)b->R * sizeof(double), hipMemcpyDeviceToHost, b->stream),
              "kinetic readback");
    b->sync();
    for (int r = 0; r < b->R; ++r) mv2[r] = b->hred[2 * r];
  });
}

// Particle fields of one replica in ascending id order: 3 doubles per particle
// (ESPP_FIELD_POS, _VEL, _FORCE, _IMAGE).
int espp_batch_get(espp_batch* b, int32_t rep, int32_t field, double* out) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    need(out != nullptr, "NULL argument");
    need(field == ESPP_FIELD_POS || field == ESPP_FIELD_VEL || field == ESPP_FIELD_FORCE || field == ESPP_FIELD_IMAGE,
         "field must be POS, VEL, FORCE or IMAGE");
    const RepDesc& d = b->hd[(size_t)rep];
    const int n = d.n;
    const size_t o = (size_t)d.slot0;
    b->sync();
    std::vector<int> id((size_t)n);
    hip_check(hipMemcpy(id.data(), b->id.p + o, n * sizeof(int), hipMemcpyDeviceToHost), "get");
    std::vector<double> val((size_t)n * 3);
    if (field == ESPP_FIELD_POS || field == ESPP_FIELD_VEL) {
      std::vector<double4> v((size_t)n);
      hip_check(hipMemcpy(v.data(), (field == ESPP_FIELD_POS ? b->pos.p : b->vel.p) + o, n * sizeof(double4),
                          hipMemcpyDeviceToHost), "get");
      for (int k = 0; k < n; ++k) { val[3 * k] = v[k].x; val[3 * k + 1] = v[k].y; val[3 * k + 2] = v[k].z; }
    } else if (field == ESPP_FIELD_FORCE) {
      std::vector<double> f((size_t)n);
      double* src[3] = {b->fx.p, b->fy.p, b->fz.p};
      for (int c = 0; c < 3; ++c) {
        hip_check(hipMemcpy(f.data(), src[c] + o, n * sizeof(double), hipMemcpyDeviceToHost), "get");
        for (int k = 0; k < n; ++k) val[3 * k + c] = f[k];
      }
    } else {
      std::vector<int4> v((size_t)n);
      hip_check(hipMemcpy(v.data(), b->img.p + o, n * sizeof(int4), hipMemcpyDeviceToHost), "get");
      for (int k = 0; k < n; ++k) { val[3 * k] = v[k].x; val[3 * k + 1] = v[k].y; val[3 * k + 2] = v[k].z; }
    }
    std::vector<int> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](int a, int c) { return id[(size_t)a] < id[(size_t)c]; });
    for (int k = 0; k < n; ++k)
      for (int c = 0; c < 3; ++c) out[3 * k + c] = val[3 * (size_t)perm[(size_t)k] + c];
  });
}

// Pairs (id1 < id2) of one replica's current list (VerletList::getAllPairs,
// src/VerletList.cpp:314-325).
int espp_batch_pairs(espp_batch* b, int32_t rep, int32_t* out, int64_t cap, int64_t* total) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    need(total != nullptr, "NULL argument");
    b->check_all_loaded();
    if (b->need_bin) {
      b->start_forces();
      auto h = b->read_state();
      b->check_state(h);
    }
    b->sync();
    const RepDesc& d = b->hd[(size_t)rep];
    const int n = d.n;
    const size_t o = (size_t)d.slot0;
    std::vector<int> id((size_t)n), cnt((size_t)n);
    hip_check(hipMemcpy(id.data(), b->id.p + o, n * sizeof(int), hipMemcpyDeviceToHost), "pairs");
    hip_check(hipMemcpy(cnt.data(), b->cnt.p + o, n * sizeof(int), hipMemcpyDeviceToHost), "pairs");
    const size_t groups = ((size_t)n + 63) / 64;
    std::vector<int> nb(groups * (size_t)b->cap * 64);
    hip_check(hipMemcpy(nb.data(), b->nbr.p + (o / 64) * (size_t)b->cap * 64, nb.size() * sizeof(int),
                        hipMemcpyDeviceToHost), "pairs");
    std::vector<std::pair<int, int>> pr;
    for (int i = 0; i < n; ++i) {
      const int c = std::min(cnt[(size_t)i], b->cap);
      for (int k = 0; k < c; ++k) {
        const int j = nb[((size_t)(i >> 6) * b->cap + k) * 64 + (i & 63)] - (int)o;
        if (j < 0 || j >= n) fail(ESPP_ERR_STATE, "list entry outside the replica");
        const int a = id[(size_t)i], c2 = id[(size_t)j];
        if (a < c2) pr.push_back({a, c2});
      }
    }
    std::sort(pr.begin(), pr.end());
    *total = (int64_t)pr.size();
    if (out)
      for (size_t k = 0; k < pr.size() && (int64_t)k < cap; ++k) { out[2 * k] = pr[k].first; out[2 * k + 1] = pr[k].second; }
  });
}

int espp_batch_set_lpp(espp_batch* b, const int32_t* lpp) {
  return guarded([&] {
    checked(b);
    need(lpp != nullptr, "NULL argument");
    for (int r = 0; r < b->R; ++r)
      need(lpp[r] == 1 || lpp[r] == 2 || lpp[r] == 4 || lpp[r] == 8, "lanes per particle must be 1, 2, 4 or 8");
    b->sync();
    b->lpp.assign(lpp, lpp + b->R);
    b->build_maps();
  });
}

int espp_batch_get_lpp(espp_batch* b, int32_t* lpp) {
  return guarded([&] {
    checked(b);
    need(lpp != nullptr, "NULL argument");
    for (int r = 0; r < b->R; ++r) lpp[r] = b->lpp[(size_t)r];
  });
}

// Average time of one force launch over all replicas (ms).
int espp_batch_force_time(espp_batch* b, int32_t reps, double* ms) {
  return guarded([&] {
    checked(b);
    need(ms != nullptr && reps >= 1, "bad argument");
    b->check_all_loaded();
    if (b->need_bin) b->start_forces();
    *ms = time_force(b, reps);
  });
}

// Average time of one list build over all replicas (ms), from one binning.
int espp_batch_build_time(espp_batch* b, int32_t reps, double* ms) {
  return guarded([&] {
    checked(b);
    need(ms != nullptr && reps >= 1, "bad argument");
    b->check_all_loaded();
    if (b->need_bin) b->start_forces();
    *ms = time_build(b, reps);
  });
}

// Autotuner of the force launch's shape.  Replicas fall into four size classes
// (quartiles of N); starting from one lane per particle everywhere, each class
// in turn tries 1, 2, 4 and 8 lanes per particle (256 / lpp particles per
// workgroup) with the other classes fixed, keeping the fastest whole-launch
// time measured with HIP events.  table (may be NULL) receives the 4 x 4
// measured times (class-major, ms); before / after the times of the starting
// and the chosen shape.
int espp_batch_autotune(espp_batch* b, int32_t reps, int32_t* lpp_out, double* table, double* before,
                        double* after) {
  return guarded([&] {
    checked(b);
    need(reps >= 1, "reps must be positive");
    b->check_all_loaded();
    if (b->need_bin) {
      b->start_forces();
      auto h = b->read_state();
      b->check_state(h);
    }
    const int R = b->R;
    std::vector<int> idx((size_t)R);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return b->hd[(size_t)a].n < b->hd[(size_t)c].n; });
    std::vector<int> cls((size_t)R);
    for (int k = 0; k < R; ++k) cls[(size_t)idx[(size_t)k]] = std::min(3, 4 * k / std::max(R, 1));
    std::vector<int> cur((size_t)R, 1);
    b->lpp = cur;
    b->build_maps();
    double best = time_force(b, reps);
    if (before) *before = best;
    const int cand[4] = {1, 2, 4, 8};
    for (int c = 0; c < 4; ++c) {
      int choice = 1;
      for (int q = 0; q < 4; ++q) {
        std::vector<int> trial = cur;
        for (int r = 0; r < R; ++r) if (cls[(size_t)r] == c) trial[(size_t)r] = cand[q];
        b->lpp = trial;
        b->build_maps();
        const double t = time_force(b, reps);
        if (table) table[4 * c + q] = t;
        if (t < best) { best = t; choice = cand[q]; }
      }
      for (int r = 0; r < R; ++r) if (cls[(size_t)r] == c) cur[(size_t)r] = choice;
    }
    b->lpp = cur;
    b->build_maps();
    if (after) *after = time_force(b, reps);
    if (lpp_out) for (int r = 0; r < R; ++r) lpp_out[r] = cur[(size_t)r];
  });
}

// Steps replayed from captured HIP graphs (on, the default) or launched directly.
int espp_batch_set_graph(espp_batch* b, int32_t on) {
  return guarded([&] {
    checked(b);
    b->use_graph = on != 0;
  });
}

// Per-replica counters: resorts (rebuilds triggered by the skin criterion),
// list builds, and the longest list of the last build.
int espp_batch_stats(espp_batch* b, int32_t rep, int64_t* resorts, int64_t* builds, int32_t* maxcount) {
  return guarded([&] {
    checked(b);
    check_rep(b, rep);
    auto h = b->read_state();
    if (resorts) *resorts = h[(size_t)rep].resorts;
    if (builds) *builds = h[(size_t)rep].builds;
    if (maxcount) *maxcount = h[(size_t)rep].maxcount;
  });
}

}  // extern "C"
