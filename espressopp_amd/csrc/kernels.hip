// HIP kernels of the MI355X short-range MD step (gfx950, wave64).
//
//   binning          k_fold_cell, k_scan_cells, k_scatter, k_sort_cells, k_permute
//   tiles            k_tile_plan (staged LDS region of each tile of particles)
//   neighbour build  k_nbr_build / k_nbr_build_tiled (full list, wavefront-padded
//                                   ragged SoA; 16-bit LDS offsets when tiled)
//   pair forces      k_lj_tiled_mixed / k_lj_mixed (32-bit fixed-point positions,
//                                   exact FP64 r^2, FP32-seeded reciprocal,
//                                   FP64 LJ factor and force accumulation; the
//                                   step variant also finishes the step)
//                    k_lj_tiled_fp64 / k_lj_fp64 (all-FP64 comparison path)
//                    k_tab_tiled / k_tab (tabulated potentials)
//   bonds            k_fene, k_bond_tab (per-particle bond adjacency, no atomics)
//   integrator       k_step_fused (thermostat + integrate2 + integrate1),
//                    k_thermo_integrate2, k_finish_max
//   observables      k_kinetic, k_rdf, k_reduce_partials
//
// Launch wrappers at the bottom are the only entry points used by the host
// runtime (sim.hip).
#include <climits>
#include <type_traits>

#include "espp_dev.h"
#include "espp_devfn.h"
#include "kernels.h"

namespace espp {

// ------------------------------------------------------------------ helpers

template <bool CUBIC>
__device__ __forceinline__ double pair_r2(const uint4& qi, int ix, int iy, int iz, const Geom& g, double& ex,
                                          double& ey, double& ez) {
  ex = (double)ix; ey = (double)iy; ez = (double)iz;
  if (!CUBIC) { ex *= g.qsd[0]; ey *= g.qsd[1]; ez *= g.qsd[2]; }
  return ex * ex + ey * ey + ez * ez;
}


// ------------------------------------------------------------------ binning

// Fold every coordinate (OrthorhombicBC::foldCoordinate, src/bc/OrthorhombicBC.cpp:144-161)
// and map the folded position to its inner cell (CellGrid::mapPositionToCellChecked
// for one node, src/storage/CellGrid.cpp:98-136).  Cell ranks come from atomics;
// k_sort_cells makes the order inside a cell deterministic afterwards.
__global__ void k_fold_cell(int n, double4* __restrict__ pos, int4* __restrict__ img,
                            int* __restrict__ cell_of, int* __restrict__ rank,
                            int* __restrict__ cell_count, Geom g, Ctl* ctl, int do_fold) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  if (!live) i = n - 1;          // inactive lanes mirror the last particle and write nothing
  double4 p = pos[i];
  int4 im = img[i];
  double x[3] = {p.x, p.y, p.z};
  int iv[3] = {im.x, im.y, im.z};
  int c[3];
  bool bad = false, moved = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // NaN, infinite, or so far outside that the image count would overflow
    if (!(fabs(x[d] * g.invL[d]) < 1.0e9)) { bad = true; c[d] = 0; continue; }
    if (do_fold) {
      int tmp = (int)floor(x[d] * g.invL[d]);
      iv[d] += tmp;
      x[d] -= (double)tmp * g.L[d];
      moved |= tmp != 0;
    }
    int cc = (int)((x[d] - 0.0) * g.ics[d]) + 1;
    if (cc < 1) cc = 1;
    else if (cc >= g.n[d] + 1) cc = g.n[d];
    c[d] = cc - 1;
  }
  int cell = c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
  const int r = wave_aggregated_rank(cell, live, cell_count);
  if (!live) return;
  if (bad) ctl->nanpos = 1;
  // a particle that stayed inside the box keeps its position and image bits
  if (moved) {
    p.x = x[0]; p.y = x[1]; p.z = x[2];
    pos[i] = p;
    img[i] = make_int4(iv[0], iv[1], iv[2], 0);
  }
  cell_of[i] = cell;
  rank[i] = r;
}

// Exclusive scan of per-cell counts in one workgroup of 1024 threads: tiles of
// 4096 cells, int4 loads per thread, wavefront shuffles and an LDS pass over the
// 16 wave totals, carried from tile to tile.
__global__ __launch_bounds__(1024) void k_scan_cells(int ncell, const int* __restrict__ count,
                                                     int* __restrict__ start) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < ncell; base += 4096) {
    int idx = base + tid * 4;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (idx + k < ncell) ? count[idx + k] : 0;
    int local = v[0] + v[1] + v[2] + v[3];
    int incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (tid < 64) {
      int w = (tid < 16) ? wsum[tid] : 0;
      int wi = w;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        int t = __shfl_up(wi, o, 64);
        if (tid >= o) wi += t;
      }
      if (tid < 16) wsum[tid] = wi - w;   // exclusive wave offsets
    }
    __syncthreads();
    int run = carry_s + wsum[wid] + incl - local;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (idx + k < ncell) start[idx + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (tid == 1023) carry_s = run;
    __syncthreads();
  }
  if (tid == 0) start[ncell] = carry_s;
}


// Multi-block exclusive scan of the cell counts: per-chunk totals, a scan of the
// totals (k_scan_cells on the short array) and a per-chunk scan with its offset.
constexpr int kScanChunk = 4096;   // 1024 threads x 4 cells

__device__ __forceinline__ int block_sum_1024(int v, int* wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 16; ++w) t += wsum[w];
  return t;
}

__global__ __launch_bounds__(1024) void k_scan_reduce(int ncell, const int* __restrict__ count,
                                                      int* __restrict__ blocksum) {
  __shared__ int wsum[16];
  const int idx = blockIdx.x * kScanChunk + threadIdx.x * 4;
  int v = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) v += (idx + k < ncell) ? count[idx + k] : 0;
  const int t = block_sum_1024(v, wsum);
  if (threadIdx.x == 0) blocksum[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_scan_down(int ncell, const int* __restrict__ count,
                                                    const int* __restrict__ blockoff, int* __restrict__ start) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int idx = blockIdx.x * kScanChunk + tid * 4;
  int v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = (idx + k < ncell) ? count[idx + k] : 0;
  const int local = v[0] + v[1] + v[2] + v[3];
  int incl = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  if (tid < 64) {
    const int w = (tid < 16) ? wsum[tid] : 0;
    int wi = w;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int t = __shfl_up(wi, o, 64);
      if (tid >= o) wi += t;
    }
    if (tid < 16) wsum[tid] = wi - w;
  }
  __syncthreads();
  int run = blockoff[blockIdx.x] + wsum[wid] + incl - local;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (idx + k < ncell) start[idx + k] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && tid == 0) start[ncell] = blockoff[gridDim.x];
}


__global__ void k_scatter(int n, const int* __restrict__ cell_of, const int* __restrict__ rank,
                          const int* __restrict__ start, int* __restrict__ order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  order[start[cell_of[i]] + rank[i]] = i;
}

// Inside each cell, order members by x, ties by previous slot.  A cell is
// handled by 32 lanes: its members' keys are staged in LDS and each member's
// place is the number of members with a smaller key (cells hold a few tens of
// particles at liquid densities; larger cells take several members per lane).
// The order is deterministic, and a row of consecutive cells is sorted by x,
// which lets the list build clip a row to the candidates within reach in x.
constexpr int kSortLanes = 32, kSortCellsPerBlock = 8, kSortMax = 96;

// The sort key is x_key of the coordinate (a total order, NaNs included), so
// the ranks below always form a permutation.

__global__ __launch_bounds__(kSortLanes * kSortCellsPerBlock) void k_sort_cells(
    int ncell, const int* __restrict__ start, int* __restrict__ order, const double4* __restrict__ pos) {
  __shared__ long long sx[kSortCellsPerBlock][kSortMax];
  __shared__ int so[kSortCellsPerBlock][kSortMax];
  const int w = threadIdx.x / kSortLanes, l = threadIdx.x % kSortLanes;
  const int c = blockIdx.x * kSortCellsPerBlock + w;
  const bool live = c < ncell;
  const int b = live ? start[c] : 0, m = live ? start[c + 1] - b : 0;
  const bool fits = m <= kSortMax;
  if (live && fits) {
    for (int k = l; k < m; k += kSortLanes) {
      const int o = order[b + k];
      so[w][k] = o;
      sx[w][k] = x_key(pos[o].x);
    }
  }
  __syncthreads();
  if (live && fits) {
    for (int k = l; k < m; k += kSortLanes) {
      const long long xk = sx[w][k];
      const int ok = so[w][k];
      int rank = 0;
      for (int q = 0; q < m; ++q) {
        const long long xq = sx[w][q];
        rank += (xq < xk || (xq == xk && so[w][q] < ok)) ? 1 : 0;
      }
      order[b + rank] = ok;
    }
  }
}

// The same ordering for cells with more than kSortMax members (one thread per
// cell, insertion sort in global memory); every other cell returns at once.
__global__ void k_sort_cells_big(int ncell, const int* __restrict__ start, int* __restrict__ order,
                                 const double4* __restrict__ pos, const Ctl* __restrict__ ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  // a binning that met a lost particle (all of them land in cell 0) is abandoned
  // by the host; skip the quadratic sort of that cell
  if (c >= ncell || ctl->nanpos) return;
  const int b = start[c], e = start[c + 1];
  if (e - b <= kSortMax) return;
  for (int k = b + 1; k < e; ++k) {
    const int o = order[k];
    const long long xo = x_key(pos[o].x);
    int m = k - 1;
    while (m >= b) {
      const int om = order[m];
      const long long xm = x_key(pos[om].x);
      if (!(xm > xo || (xm == xo && om > o))) break;
      order[m + 1] = om;
      --m;
    }
    order[m + 1] = o;
  }
}

// Cell of the caller's grid (halfCellInt > 1) holding a folded position:
// CellGrid::mapPositionToCellClipped for one node (src/storage/CellGrid.cpp:84-97),
// truncation toward zero, then clipped to the inner cells.
__device__ __forceinline__ int caller_cell(const double4& p, const Geom& g) {
  const double x[3] = {p.x, p.y, p.z};
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int cc = (int)(x[d] * g.hics[d]);
    if (cc < 0) cc = 0;
    else if (cc >= g.nh[d]) cc = g.nh[d] - 1;
    c[d] = cc;
  }
  return c[0] + g.nh[0] * (c[1] + g.nh[1] * c[2]);
}

// Orientation of the reference's list test for halfCellInt > 1: the pair is
// visited from the cell with the smaller framed index (neighbours within +-h
// cells, src/storage/DomainDecomposition.cpp:342-358), so d = xi - (xj + s L)
// iff the offset of j's ghost cell (its cell shifted by s grid lengths) from
// i's cell is lexicographically positive, z most significant.
__device__ __forceinline__ bool caller_forward(int ci, int cj, int sx, int sy, int sz, const Geom& g) {
  const int nx = g.nh[0], ny = g.nh[1], nz = g.nh[2];
  const int ox = cj % nx + sx * nx - ci % nx;
  const int oy = (cj / nx) % ny + sy * ny - (ci / nx) % ny;
  const int oz = cj / (nx * ny) + sz * nz - ci / (nx * ny);
  return (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
}

// |d|^2 of a listed pair (slots i, j) as the reference's list and force loops
// evaluate it: d in the orientation of the half-shell visit, with the second
// particle's ghost at x + s L (s: j's periodic image nearest to i; see k_nbr_build),
// summed left to right without contraction.  The force kernels form d their own
// way and use this only for pairs within a few ulp of a cutoff, so that such a
// pair falls on the reference's side of it.
__device__ double ref_pair_r2(int i, int j, const double4& pi, const double4& pj, const Geom& g) {
#pragma clang fp contract(off)
  const int sx = (int)__builtin_rint((pi.x - pj.x) * g.invL[0]);
  const int sy = (int)__builtin_rint((pi.y - pj.y) * g.invL[1]);
  const int sz = (int)__builtin_rint((pi.z - pj.z) * g.invL[2]);
  bool fwd;
  if (g.hci > 1) {
    fwd = caller_forward(g.hcell[i], g.hcell[j], sx, sy, sz, g);
  } else {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int ci = g.cslot[i], cj = g.cslot[j];
    const int ox = cj % nx + sx * nx - ci % nx;
    const int oy = (cj / nx) % ny + sy * ny - (ci / nx) % ny;
    const int oz = cj / (nx * ny) + sz * nz - ci / (nx * ny);
    fwd = (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
  }
  const double shx = (double)sx * g.L[0], shy = (double)sy * g.L[1], shz = (double)sz * g.L[2];
  double ex, ey, ez;
  if (fwd) { ex = pi.x - (pj.x + shx); ey = pi.y - (pj.y + shy); ez = pi.z - (pj.z + shz); }
  else { ex = pj.x - (pi.x - shx); ey = pj.y - (pi.y - shy); ez = pj.z - (pi.z - shz); }
  return ex * ex + ey * ey + ez * ez;
}

__global__ void k_permute(int n, const int* __restrict__ order, const int* __restrict__ cell_of,
                          const double4* __restrict__ pos, const double4* __restrict__ vel,
                          const double* __restrict__ fx, const double* __restrict__ fy,
                          const double* __restrict__ fz, const int* __restrict__ id,
                          const int4* __restrict__ img, double4* __restrict__ pos2,
                          double4* __restrict__ vel2, double* __restrict__ fx2,
                          double* __restrict__ fy2, double* __restrict__ fz2,
                          int* __restrict__ id2, int4* __restrict__ img2,
                          int* __restrict__ cell_of_slot, int* __restrict__ slot_of_id,
                          uint4* __restrict__ posq, float4* __restrict__ posf, uint2* __restrict__ posh,
                          Geom g) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = order[s];
  double4 p = pos[o];
  pos2[s] = p;
  vel2[s] = vel[o];
  if (fx2) {   // null when the forces are recomputed right after the binning
    fx2[s] = fx[o];
    fy2[s] = fy[o];
    fz2[s] = fz[o];
  }
  int pid = id[o];
  id2[s] = pid;
  img2[s] = img[o];
  cell_of_slot[s] = cell_of[o];
  slot_of_id[pid] = s;
  if (posq)
    posq[s] = make_uint4(fixq(p.x, g.qinv[0]), fixq(p.y, g.qinv[1]), fixq(p.z, g.qinv[2]),
                         (uint32_t)(int)p.w);
  posf[s] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
  if (posh) posh[s] = q16_pack(p, g.q16inv);
  if (g.hci > 1) g.hcell[s] = caller_cell(p, g);
}


// Per cell c, xbin[c (kXBins+1) + k] is the first slot of the cell whose x
// sub-bin is at least k (k < kXBins), and the cell's end for k = kXBins.  The
// binning has just sorted every cell by x, so the sub-bins rise along a cell:
// slot s writes the entries from its predecessor's sub-bin (exclusive) to its
// own, the last slot of a cell the entries above its sub-bin, and an empty
// cell all of its row (threads past n take the cells).
__global__ void k_xbins(int n, int ncell, const int* __restrict__ cstart, const int* __restrict__ cell_of_slot,
                        const float4* __restrict__ posf, Geom g, int* __restrict__ xbin) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.n[0];
  const float inv_cs = 1.0f / (float)(g.L[0] / nx);
  if (t < n) {
    const int c = cell_of_slot[t];
    const int cx = c % nx;
    int* out = xbin + (size_t)c * (kXBins + 1);
    const int k = xsub_bin(posf[t].x, inv_cs, cx);
    const int kp = (t > 0 && cell_of_slot[t - 1] == c) ? xsub_bin(posf[t - 1].x, inv_cs, cx) : -1;
    for (int q = kp + 1; q <= k; ++q) out[q] = t;
    if (t + 1 == n || cell_of_slot[t + 1] != c)
      for (int q = k + 1; q <= kXBins; ++q) out[q] = t + 1;
  }
  if (t < ncell && cstart[t] == cstart[t + 1]) {
    int* out = xbin + (size_t)t * (kXBins + 1);
    for (int q = 0; q <= kXBins; ++q) out[q] = cstart[t];
  }
}

__global__ void k_posf(int n, const double4* __restrict__ pos, float4* __restrict__ posf, uint2* __restrict__ posh,
                       double q16inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = pos[i];
  posf[i] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
  if (posh) posh[i] = q16_pack(p, q16inv);
}

__global__ void k_fill_int(int n, int* __restrict__ a, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_reset_ctl(Ctl* ctl, int fields) {
  if (threadIdx.x == 0) {
    if (fields & kCtlBin) {
      ctl->maxsq_bits = 0ull;
      ctl->nanpos = 0;
    }
    if (fields & kCtlBuild) {
      ctl->overflow = 0;
      ctl->maxcount = 0;
    }
    if (fields & kCtlPlan) {
      ctl->plan_fail = 0;
      ctl->max_region = 0;
    }
  }
}

// ------------------------------------------------------------------ neighbour list

__device__ bool is_excluded(int a, int b, const int* __restrict__ ex_start,
                            const int* __restrict__ ex_ids) {
  int lo = ex_start[a], hi = ex_start[a + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = ex_ids[mid];
    if (v == b) return true;
    if (v < b) lo = mid + 1; else hi = mid - 1;
  }
  return false;
}

// The workgroup's maximum of v into *dst with one atomic per workgroup (every
// thread calls it): same-address atomics from each wavefront of a large grid
// would queue behind each other.
__device__ void block_max_to(int v, int* dst) {
  __shared__ int wmax[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = wmax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = max(m, wmax[w]);
    atomicMax(dst, m);
  }
}

// Full (both-direction) Verlet list.  Membership reproduces the reference's
// half-list test exactly (src/VerletList.cpp:247-276): a pair is kept iff
// d.sqr() <= (rc+skin)^2 in FP64 with no contraction, where d is evaluated in
// the orientation the reference uses.  The reference evaluates a pair from the
// cell whose framed index is smaller (half shell, src/storage/DomainDecomposition.cpp:354-357),
// with the second particle's ghost copy at p + shift (src/Particle.hpp:93-98).
// So for cell offset o of j relative to i: if o is lexicographically positive
// (z most significant) d = xi - (xj + s L), else d = xj - (xi - s L).
//
// Layout: particle i's k-th neighbour is nbr[((i/64)*cap + k)*64 + i%64].
template <bool HAS_EXCL>
__global__ __launch_bounds__(256) void k_nbr_build(int n, const double4* __restrict__ pos,
                                                   const int* __restrict__ id,
                                                   const int* __restrict__ cell_of_slot,
                                                   const int* __restrict__ cstart, Geom g,
                                                   double cutsq, const int* __restrict__ ex_start,
                                                   const int* __restrict__ ex_ids,
                                                   int* __restrict__ nbr, int* __restrict__ cnt,
                                                   int cap, Ctl* ctl) {
#pragma clang fp contract(off)
  // a lost particle: the binning put it in cell 0 and the run ends in an error;
  // do not scan what may be one huge cell
  if (ctl->nanpos) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int k = 0;
  if (i < n) {
    const double4 pi = pos[i];
    const int c = cell_of_slot[i];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    const int myid = HAS_EXCL ? id[i] : 0;
    int* out = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    for (int oz = -1; oz <= 1; ++oz) {
      int qz = cz + oz, sz = 0;
      if (qz < 0) { qz += nz; sz = -1; } else if (qz >= nz) { qz -= nz; sz = 1; }
      for (int oy = -1; oy <= 1; ++oy) {
        int qy = cy + oy, sy = 0;
        if (qy < 0) { qy += ny; sy = -1; } else if (qy >= ny) { qy -= ny; sy = 1; }
        for (int ox = -1; ox <= 1; ++ox) {
          int qx = cx + ox, sx = 0;
          if (qx < 0) { qx += nx; sx = -1; } else if (qx >= nx) { qx -= nx; sx = 1; }
          const bool fwd = (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
          const bool unshifted = (sx | sy | sz) == 0;
          const double shx = (double)sx * g.L[0], shy = (double)sy * g.L[1], shz = (double)sz * g.L[2];
          const int c2 = qx + nx * (qy + ny * qz);
          const int jb = cstart[c2], je = cstart[c2 + 1];
          for (int j = jb; j < je; ++j) {
            if (j == i && unshifted) continue;
            const double4 pj = pos[j];
            double dx, dy, dz;
            if (g.hci > 1 ? caller_forward(g.hcell[i], g.hcell[j], sx, sy, sz, g) : fwd) {
              dx = pi.x - (pj.x + shx);
              dy = pi.y - (pj.y + shy);
              dz = pi.z - (pj.z + shz);
            } else {
              dx = pj.x - (pi.x - shx);
              dy = pj.y - (pi.y - shy);
              dz = pj.z - (pi.z - shz);
            }
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= cutsq) {
              if (HAS_EXCL && is_excluded(myid, id[j], ex_start, ex_ids)) continue;
              if (k < cap) out[(size_t)k * 64] = j;
              ++k;
            }
          }
        }
      }
    }
    cnt[i] = k;
    if (k > cap) ctl->overflow = 1;
  }
  block_max_to(k, &ctl->maxcount);
}

// ------------------------------------------------------------------ LJ forces

// Energy/virial partials per block: [e, w, wxx, wyy, wzz, wxy, wxz, wyz]
constexpr int kNPart = 8;

template <int NV>
__device__ void block_reduce_store(double (&v)[NV], double* __restrict__ partial) {
  __shared__ double red[NV][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    int q = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[q][w];
    partial[(size_t)blockIdx.x * NV + q] = s;
  }
}

// Mixed-precision LJ over the full list (grids with fewer than 3 cells in a
// dimension).  Pair distances come from exact integer differences of the 32-bit
// fixed-point mirror (minimum image for free), r^2 is formed in FP64, the
// reciprocal is seeded by v_rcp_f32 and refined by one FP64 Newton step, and the
// force sum is FP64.  Pairs whose r^2 lies in a small window around rc^2
// re-decide the cutoff test from the FP64 positions so the set of interacting
// pairs equals the FP64 path's.
// Potential: src/interaction/LennardJones.hpp:129-136 (force), :109-114 (energy).
template <bool ONETYPE, bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_lj_mixed(
    int n, const uint4* __restrict__ posq, const double4* __restrict__ pos,
    const int* __restrict__ nbr, const int* __restrict__ cnt, int cap, Geom g,
    const LJParam* __restrict__ prm, int ntypes, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ partial, const int* stop) {
  if (stop && *reinterpret_cast<const volatile int*>(stop)) return;   // a window's step after the stop
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (i < n) {
    const uint4 qi = posq[i];
    const int* np = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    const int c = min(cnt[i], cap);
    const int rowi = (int)qi.w * ntypes;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    for (int k = 0; k < c; ++k) {
      const int j = np[(size_t)k * 64];
      const uint4 qj = posq[j];
      const LJParam& p = prm[ONETYPE ? 0 : rowi + (int)qj.w];
      double ex, ey, ez;
      const double r2 = pair_r2<false>(qi, (int)(qi.x - qj.x), (int)(qi.y - qj.y), (int)(qi.z - qj.z), g, ex, ey, ez);
      bool in = p.set && r2 <= p.drc2;
      if (p.set && r2 >= p.dlo && r2 <= p.dhi) {
        // fixed-point quantisation could flip the cutoff test: decide it as the reference does
        in = ref_pair_r2(i, j, pos[i], pos[j], g) <= p.drc2;
      }
      double inv;
      double ff = lj_ff_seeded(r2, p.dff1, p.dff2, inv);
      ff = in ? ff : 0.0;
      Ax = __builtin_fma(ex, ff, Ax);
      Ay = __builtin_fma(ey, ff, Ay);
      Az = __builtin_fma(ez, ff, Az);
      if (ENERGY && in) {
        const double f2 = p.sig2 * inv;
        const double f6 = f2 * f2 * f2;
        acc[0] += p.eps4 * (f6 * f6 - f6) - p.dshift;
        acc[1] += r2 * ff;
        acc[2] += ex * ex * ff; acc[3] += ey * ey * ff; acc[4] += ez * ez * ff;
        acc[5] += ex * ey * ff; acc[6] += ex * ez * ff; acc[7] += ey * ez * ff;
      }
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// All-FP64 LJ over the same full list (the comparison path).
template <bool ONETYPE, bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_lj_fp64(
    int n, const double4* __restrict__ pos, const int* __restrict__ nbr,
    const int* __restrict__ cnt, int cap, Geom g, const LJParam* __restrict__ prm, int ntypes,
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    double* __restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (i < n) {
    const double4 pi = pos[i];
    const int* np = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    const int c = min(cnt[i], cap);
    const int rowi = (int)pi.w * ntypes;
    double ff1 = prm[0].dff1, ff2 = prm[0].dff2, rc2 = prm[0].drc2, shift = prm[0].dshift;
    double sig2 = prm[0].sig2, eps4 = prm[0].eps4, lo = prm[0].dlo, hi = prm[0].dhi;
    int set = prm[0].set;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    for (int k = 0; k < c; ++k) {
      const int j = np[(size_t)k * 64];
      const double4 pj = pos[j];
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      dx = __builtin_fma(-g.L[0], __builtin_rint(dx * g.invL[0]), dx);
      dy = __builtin_fma(-g.L[1], __builtin_rint(dy * g.invL[1]), dy);
      dz = __builtin_fma(-g.L[2], __builtin_rint(dz * g.invL[2]), dz);
      if (!ONETYPE) {
        const LJParam& p = prm[rowi + (int)pj.w];
        ff1 = p.dff1; ff2 = p.dff2; rc2 = p.drc2; shift = p.dshift; sig2 = p.sig2; eps4 = p.eps4;
        set = p.set; lo = p.dlo; hi = p.dhi;
      }
      const double r2 = sqr_nc(dx, dy, dz);
      bool in = set && r2 <= rc2;
      if (set && r2 >= lo && r2 <= hi) in = ref_pair_r2(i, j, pi, pj, g) <= rc2;   // at the cutoff (rare)
      const double inv = rcp_nr(r2);
      const double r6 = inv * inv * inv;
      double ff = r6 * (ff1 * r6 - ff2) * inv;
      ff = in ? ff : 0.0;
      Ax = __builtin_fma(dx, ff, Ax);
      Ay = __builtin_fma(dy, ff, Ay);
      Az = __builtin_fma(dz, ff, Az);
      if (ENERGY) {
        const double f2 = sig2 * inv;
        const double f6 = f2 * f2 * f2;
        acc[0] += in ? eps4 * (f6 * f6 - f6) - shift : 0.0;
        acc[1] += r2 * ff;
        acc[2] += dx * dx * ff; acc[3] += dy * dy * ff; acc[4] += dz * dz * ff;
        acc[5] += dx * dy * ff; acc[6] += dx * dz * ff; acc[7] += dy * dz * ff;
      }
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// ------------------------------------------------------------------ tiles
//
// A tile is T consecutive particles (T = blockDim.x of the tiled kernels, one
// particle per thread).  Its staged region is the union of the 27 neighbour
// cells of every cell the tile touches, kept as runs of consecutive cell
// indices; the particles of a run are consecutive slots, so staging is a set of
// coalesced copies into LDS.  Neighbour entries of the tiled list are 16-bit
// offsets into that LDS buffer, two per 32-bit word.

__device__ __forceinline__ int wrap_dim(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// One wavefront per tile, four tiles per workgroup.  A tile touches at most
// three consecutive cell rows (more and the plan fails, and the host tries a
// smaller tile); each of those rows has 9 neighbour rows, each of which gives a
// piece of at most 3 cells, split in two where it wraps: lane
// (row * 9 + neighbour) * 2 + part holds one piece.  The pieces are sorted by
// first cell across the wavefront (bitonic, by shuffles), a piece that does not
// touch the ones before it starts a run (a prefix maximum of the piece ends),
// and each run's slot range and LDS offset follow from the cell starts and a
// prefix sum.  Runs are sorted by cell index and never overlap.  The largest
// region goes to ctl by one atomic per workgroup (same-address atomics from
// every wavefront of the grid would serialise).
constexpr int kPlanTilesPerBlock = 16;
__global__ __launch_bounds__(64 * kPlanTilesPerBlock) void k_tile_plan(
    int n, int T, int ntiles, const int* __restrict__ cell_of_slot, const int* __restrict__ cstart, Geom g,
    int budget, TileRun* __restrict__ runs, int* __restrict__ nruns, Ctl* ctl) {
  // a lost particle: the run ends in an error (no plan is needed)
  if (ctl->nanpos) return;
  __shared__ int run_a[kPlanTilesPerBlock][64], run_b[kPlanTilesPerBlock][64];
  __shared__ int region_max[kPlanTilesPerBlock];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t = blockIdx.x * kPlanTilesPerBlock + wv;
  const bool live = t < ntiles;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int s0 = live ? t * T : 0, s1 = live ? min(s0 + T, n) - 1 : 0;
  const int c0 = cell_of_slot[s0], c1 = cell_of_slot[s1];
  const int r0 = c0 / nx, r1 = c1 / nx;
  const bool bad = r1 - r0 + 1 > 3;
  constexpr int kNone = 0x7FFFFFFF;
  int pa = kNone, pb = -1;
  if (live && !bad && lane < 54) {
    const int ri = lane / 18, rem = lane % 18, combo = rem >> 1, part = rem & 1;
    const int r = r0 + ri;
    if (r <= r1) {
      const int y = r % ny, z = r / ny;
      const int xa = (r == r0) ? c0 % nx : 0, xb = (r == r1) ? c1 % nx : nx - 1;
      const int lo = xa - 1, hi = xb + 1;
      const int zz = wrap_dim(z + combo / 3 - 1, nz), yy = wrap_dim(y + combo % 3 - 1, ny);
      const int rb = (yy + ny * zz) * nx;
      if (hi - lo + 1 >= nx) {
        if (part == 0) { pa = rb; pb = rb + nx - 1; }
      } else if (lo < 0) {
        if (part == 0) { pa = rb + lo + nx; pb = rb + nx - 1; } else { pa = rb; pb = rb + hi; }
      } else if (hi >= nx) {
        if (part == 0) { pa = rb + lo; pb = rb + nx - 1; } else { pa = rb; pb = rb + hi - nx; }
      } else if (part == 0) {
        pa = rb + lo; pb = rb + hi;
      }
    }
  }
  // bitonic sort of the pieces by (first cell, lane) across the wavefront
  int key = pa, val = pb, tie = lane;
  for (int size = 2; size <= 64; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int ok = __shfl_xor(key, stride, 64), ov = __shfl_xor(val, stride, 64), ot = __shfl_xor(tie, stride, 64);
      const bool up = (lane & size) == 0;
      const bool lower = (lane & stride) == 0;
      const bool other_less = ok < key || (ok == key && ot < tie);
      // the lower lane of an ascending pair keeps the smaller element
      if (other_less == (lower == up)) { key = ok; val = ov; tie = ot; }
    }
  }
  const bool valid = key != kNone;
  int incl = valid ? val : -2;                   // inclusive prefix maximum of the piece ends
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl = max(incl, y);
  }
  int excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = -2;
  const bool start = valid && key > excl + 1;
  const unsigned long long starts = __ballot(start);
  const int id = __popcll(starts & ((2ull << lane) - 1ull)) - 1;   // run of this piece
  const bool next_start = lane < 63 && ((starts >> (lane + 1)) & 1ull);
  const bool next_valid = __shfl_down(valid ? 1 : 0, 1, 64) != 0 && lane < 63;
  const bool last = valid && (!next_valid || next_start);
  if (start) run_a[wv][id] = key;
  if (last) run_b[wv][id] = incl;
  __syncthreads();
  const int nr = __popcll(starts);
  int slot_start = 0, count = 0, ca = 0, cb = 0;
  if (live && !bad && lane < nr) {
    ca = run_a[wv][lane];
    cb = run_b[wv][lane];
    slot_start = cstart[ca];
    count = cstart[cb + 1] - slot_start;
  }
  int off = count;                               // inclusive prefix sum of the counts
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(off, o, 64);
    if (lane >= o) off += y;
  }
  const int total = __shfl(off, 63, 64);
  const bool fail = bad || nr > kMaxRuns || total > budget || total > 65535;
  if (lane == 0) region_max[wv] = (live && !fail) ? total : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int w = 0; w < kPlanTilesPerBlock; ++w) m = max(m, region_max[w]);
    atomicMax(&ctl->max_region, m);
  }
  if (!live) return;
  if (!fail && lane < nr) {
    TileRun R;
    R.slot_start = slot_start;
    R.count = count;
    R.lds_off = off - count;
    R.cell_a = ca;
    R.cell_b = cb;
    R.pad0 = R.pad1 = R.pad2 = 0;
    runs[(size_t)t * kMaxRuns + lane] = R;
  }
  if (lane == 0) {
    if (fail) ctl->plan_fail = 1;
    nruns[t] = fail ? 0 : nr;
  }
}

// run holding cell c (runs sorted by cell_a, c guaranteed inside one of them)
__device__ __forceinline__ int find_run_by_cell(const TileRun* sr, int nr, int c) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sr[mid].cell_a <= c) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The part of a TileRun the force kernels keep in LDS: 8 bytes per run, with a
// closing entry whose lds_off is the region size (a run's count is the next
// run's lds_off minus its own).  The small table keeps the LDS of a force
// workgroup low enough for four workgroups per CU.
struct RunRef {
  int slot_start, lds_off;
};

// slot of LDS entry e (runs sorted by lds_off)
__device__ __forceinline__ int slot_of_lds(const RunRef* rr, int nr, int e) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rr[mid].lds_off <= e) lo = mid; else hi = mid - 1;
  }
  return rr[lo].slot_start + (e - rr[lo].lds_off);
}

// Stages tile t's region into LDS (entry e <- src of its slot, through conv) and
// leaves the run table in rr for slot_of_lds.  Every wavefront holds the table
// in its lanes (run k in lane k, kMaxRuns == 64) and stages 64 consecutive
// entries at a time; the run of each entry follows from two ballots over the
// runs' LDS offsets and scalar reads of the runs that start inside the 64, so
// the loads of all of a thread's entries go out back to back, with no LDS
// round trip before them.
// A tile's run table as each wavefront holds it: run k (slot_start, count,
// lds_off) in lane k.  The plan fills the first nr entries of the tile's table;
// the rest are loaded but never used.
struct RunTable {
  int4 R;
  int nr;
};

__device__ __forceinline__ RunTable load_run_table(const TileRun* __restrict__ runs, const int* __restrict__ nruns,
                                                   int t) {
  static_assert(kMaxRuns == 64, "one run per lane");
  RunTable rt;
  rt.R = *reinterpret_cast<const int4*>(runs + (size_t)t * kMaxRuns + ((int)threadIdx.x & 63));
  rt.nr = nruns[t];
  return rt;
}

// slot - entry for entry e = e0 + lane of a wavefront: the run of e0 is the last
// one starting at or before it (off holds run lane's LDS offset, rising with the
// lane, and INT_MAX past the table); later runs starting inside the
// 64 entries take over from their first entry on.  Up to three of them are
// resolved without a loop.
__device__ __forceinline__ int stage_delta(int e0, int e, int off, int delta) {
  const int r0 = __popcll(__ballot(off <= e0)) - 1;
  int d = __builtin_amdgcn_readlane(delta, r0);
#pragma unroll
  for (int b = 1; b <= 3; ++b) {
    const int k = min(r0 + b, 63);
    if (e >= __builtin_amdgcn_readlane(off, k)) d = __builtin_amdgcn_readlane(delta, k);
  }
  // done unless a fourth run starts inside the 64 entries
  if (r0 + 4 > 63 || __builtin_amdgcn_readlane(off, r0 + 4) > e0 + 63) return d;
  // more than three runs start inside the 64 entries (short runs)
  unsigned long long inside = __ballot(off > e0 && off <= e0 + 63);
  while (inside) {
    const int k = __ffsll((long long)inside) - 1;
    inside &= inside - 1ull;
    if (e >= __builtin_amdgcn_readlane(off, k)) d = __builtin_amdgcn_readlane(delta, k);
  }
  return d;
}

// One pass of stage_region over entries [base, base + kB * blockDim.x): source
// slots first (no memory access), then all loads, then all stores.  The loads
// are unconditional (entries past the region read slot 0 and are dropped), so
// the waits before the stores can count them; after_loads() issues the
// caller's own loads behind them.
template <int kB, typename V, typename S, typename Conv, typename After>
__device__ __forceinline__ void stage_pass(int base, int total, int off, int delta, const V* __restrict__ src,
                                           S* lds, Conv conv, After after_loads) {
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int bs = blockDim.x;
  int js[kB];
#pragma unroll
  for (int u = 0; u < kB; ++u) {
    const int e0 = base + u * bs + wv * 64;                         // first entry of the wavefront
    const int e = e0 + lane;
    js[u] = 0;
    if (e0 < total) {
      const int d = stage_delta(e0, e, off, delta);
      if (e < total) js[u] = e + d;
    }
  }
  V v[kB];
#pragma unroll
  for (int u = 0; u < kB; ++u) v[u] = src[js[u]];
  after_loads();
#pragma unroll
  for (int u = 0; u < kB; ++u) {
    const int e = base + u * bs + wv * 64 + lane;
    if (e < total) lds[e] = conv(v[u]);
  }
}

// Stages a tile's region into LDS (entry e <- src of its slot, through conv) and
// leaves the run table in rr for slot_of_lds.  Every wavefront holds the table
// in its lanes and stages 64 consecutive entries at a time, finding their runs
// from ballots and scalar reads, so the loads of all of a thread's entries go
// out back to back with no LDS round trip before them.  after_loads() runs once,
// after the first pass's loads are issued.
template <typename V, typename S, typename Conv, typename After>
__device__ __forceinline__ int stage_region(const RunTable& rt, const V* __restrict__ src, S* lds, RunRef* rr,
                                            Conv conv, After after_loads) {
  const int lane = (int)threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int4 R = rt.R;
  const int nr = rt.nr;
  // region size: end of the last run
  const int total = nr > 0 ? __builtin_amdgcn_readlane(R.z + R.y, nr - 1) : 0;
  // lanes past the table start beyond every entry, so no ballot or test takes them
  const int off = lane < nr ? R.z : INT_MAX;
  const int delta = lane < nr ? R.x - R.z : 0;                      // slot = entry + delta
  if (wv == 0) {
    if (lane < nr) rr[lane] = RunRef{R.x, R.z};
    if (lane == 0) rr[nr] = RunRef{0, total};
  }
  constexpr int kB = sizeof(V) <= 16 ? 16 : 8;                     // entries in flight per thread
  const int step = kB * (int)blockDim.x;
  stage_pass<kB>(0, total, off, delta, src, lds, conv, after_loads);
  for (int base = step; base < total; base += step)
    stage_pass<kB>(base, total, off, delta, src, lds, conv, [] {});
  __syncthreads();
  return nr;
}

template <typename V>
__device__ __forceinline__ int stage_tile(const TileRun* __restrict__ runs, const int* __restrict__ nruns,
                                          const V* __restrict__ src, V* lds, RunRef* rr, int t) {
  return stage_region(load_run_table(runs, nruns, t), src, lds, rr, [](const V& v) { return v; }, [] {});
}

template <typename S>
__device__ __forceinline__ S narrow_slot(const uint4& v);
template <>
__device__ __forceinline__ uint4 narrow_slot<uint4>(const uint4& v) { return v; }
template <>
__device__ __forceinline__ uint3 narrow_slot<uint3>(const uint4& v) { return make_uint3(v.x, v.y, v.z); }

// stage_tile for the fixed-point mirror, storing S (uint3 or uint4) per particle
template <typename S, typename After>
__device__ __forceinline__ int stage_tile_q(const RunTable& rt, const uint4* __restrict__ src, S* lds, RunRef* rr,
                                            After after_loads) {
  return stage_region(rt, src, lds, rr, [](const uint4& v) { return narrow_slot<S>(v); }, after_loads);
}

template <typename S>
__device__ __forceinline__ int stage_tile_q(const TileRun* __restrict__ runs, const int* __restrict__ nruns,
                                            const uint4* __restrict__ src, S* lds, RunRef* rr, int t) {
  return stage_tile_q<S>(load_run_table(runs, nruns, t), src, lds, rr, [] {});
}

// d.sqr() <= cutsq of Real3D (src/Real3D.hpp), evaluated left to right without contraction
__device__ __forceinline__ bool exact_sqr_le(double ex, double ey, double ez, double cutsq) {
#pragma clang fp contract(off)
  return ex * ex + ey * ey + ez * ez <= cutsq;
}

// Rank of each thread of the workgroup in order of (cls, thread index), for small
// class values 0 <= cls < 64 and at most 256 threads: the number of threads with
// a smaller class plus the lower-indexed threads of the same class.  Lanes of a
// wavefront with equal classes are counted by one ballot per distinct class.
__device__ int block_class_rank(int cls) {
  __shared__ int wcount[4][64];                  // per wavefront, threads of each class
  __shared__ int cbase[64];                      // threads of all smaller classes
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int q = threadIdx.x; q < 4 * 64; q += blockDim.x) wcount[q >> 6][q & 63] = 0;
  __syncthreads();
  unsigned long long todo = __ballot(1);
  int within = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lc = __shfl(cls, leader, 64);
    const unsigned long long m = __ballot(cls == lc);
    if (cls == lc) within = __popcll(m & ((1ull << ln) - 1ull));
    if (ln == leader) wcount[wv][lc] = __popcll(m);
    todo &= ~m;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    const int tot = wcount[0][c] + wcount[1][c] + wcount[2][c] + wcount[3][c];
    int incl = tot;                              // inclusive scan over the 64 classes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (c >= o) incl += y;
    }
    cbase[c] = incl - tot;
  }
  __syncthreads();
  int r = cbase[cls] + within;
  for (int w = 0; w < wv; ++w) r += wcount[w][cls];
  return r;
}

// Workgroup -> tile map that keeps consecutive tiles on one XCD: workgroups are
// dispatched round-robin over the 8 XCDs, so XCD x receives workgroups x, x+8,
// ...; those take tiles [x*q + min(x, r) ...) in order (ntiles = 8q + r).
// Neighbouring tiles stage overlapping regions, which then hit the same L2.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int q = ntiles / kXcd, r = ntiles % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * q + min(x, r) + k;
}

// lds[e] for the list entry e held in the low (HI false) or high half of the
// word w: one v_mad_u32_u16 reads the half and forms the LDS byte address
// e * sizeof(S) + lds, where a mask or shift and a multiply-add would take two.
template <bool HI, typename S>
__device__ __forceinline__ S lds_entry(const S* lds, uint32_t w) {
  typedef __attribute__((address_space(3))) const uint32_t* lds_word;
  static_assert(sizeof(S) % 4 == 0, "slots of whole words");
  const uint32_t base = (uint32_t)(uintptr_t)(lds_word)(const void*)lds;
  uint32_t a;
  if (HI) asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(a) : "v"(w), "n"((int)sizeof(S)), "v"(base));
  else asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(a) : "v"(w), "n"((int)sizeof(S)), "v"(base));
  const lds_word p = (lds_word)(uintptr_t)a;
  S r;
  uint32_t* rw = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(S) / 4); ++k) rw[k] = p[k];
  return r;
}

// Tiled list layout: entry q of particle i is 16-bit half (q & 1) of word
// ((q >> 1) & 3) of the uint4 at ((i >> 6) * capg + (q >> 3)) * 64 + (i & 63);
// a lane reads or writes eight entries with one 16-byte access and the 64 lanes
// of a wavefront touch 1 KiB of consecutive memory.
__device__ __forceinline__ size_t tiled_group(int i, int capg, int g) {
  return ((size_t)(i >> 6) * (size_t)capg + (size_t)g) * 64 + (size_t)(i & 63);
}

// candidates tested per block of the list build (bit masks of this width); a
// short block wastes less of a row's clipped range at its end
constexpr int kCandChunk = 16;
constexpr uint32_t kCandAll = (1u << kCandChunk) - 1u;

// entries of a lane's LDS staging buffer in the list build (pending and new entries
// of one block)
constexpr int kRing = 8 + kCandChunk;

// the 8 list entries at positions p0 .. p0+7 of a thread's staging buffer (position
// p at r[p * 256]) as one 16-byte group
__device__ __forceinline__ uint4 ring_group(const unsigned short* r, int p0) {
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    w[q] = (uint32_t)r[(p0 + 2 * q) * 256] | ((uint32_t)r[(p0 + 2 * q + 1) * 256] << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// Tiled Verlet-list build (cell grids with at least 3 cells per dimension).
// For each of its 9 neighbour x-rows a particle scans the candidates of the
// three cells (cx-1, cx, cx+1), one contiguous slot range or two when the row
// wraps.  A row whose y/z range lies beyond the prefilter's top is skipped, the
// others keep the x reach the y/z distance leaves (when the binning folded the
// positions), and the range is cut to that reach with the cells' x sub-bins
// (cells are sorted by x).  Candidates come from a copy of the positions padded
// so that reads past a range stay inside it: with Q16 the 16-bit modular copy
// (q16_pack; the packed differences pick the periodic image themselves), else
// FP32 with particle i shifted by each piece's periodic image once.  For
// kCandChunk candidates at a time two bit masks are formed without branches:
// r^2 below the window (surely in) and below its top (in or undecided).
// Undecided pairs, within a small window of (rc+skin)^2, are settled by the
// reference's FP64 test in the reference's orientation (see k_nbr_build).
// Accepted candidates are appended, without a branch, as their offset in the
// tile's staged region, eight entries per 16-byte store.
template <bool HAS_EXCL, bool Q16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_nbr_build_tiled(
    int n, int ntiles, const float4* __restrict__ posf, const double4* __restrict__ pos, const int* __restrict__ id,
    const int* __restrict__ cell_of_slot, const int* __restrict__ cstart, Geom g, double cutsq, float cut_lo,
    float cut_hi, const TileRun* __restrict__ runs, const int* __restrict__ nruns,
    const int* __restrict__ ex_start, const int* __restrict__ ex_ids, uint4* __restrict__ nbr,
    int* __restrict__ cnt, int capg, int* __restrict__ lane_slot, float clip_margin, float row_margin,
    const int* __restrict__ xbin, const uint2* __restrict__ posh, double q16inv, int q16_tlo, int q16_w,
    Ctl* ctl) {
  // a lost particle: the binning put it in cell 0 and the run ends in an error;
  // do not scan what may be one huge cell.  A failed tile plan: the host plans
  // smaller tiles and builds again.
  if (ctl->nanpos || ctl->plan_fail) return;
  __shared__ TileRun sr[kMaxRuns];
  __shared__ unsigned short build_ring[kRing * 256];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int nr = nruns[tile];
  if ((int)threadIdx.x < nr) sr[threadIdx.x] = runs[(size_t)tile * kMaxRuns + threadIdx.x];
  __syncthreads();
  // threads in slot order (lanes of one cell read the same candidates); each
  // particle's list is stored at its slot
  const int s0 = tile * blockDim.x;
  const int i = s0 + threadIdx.x;        // particle slot
  int k = 0;
  if (i < n && nr > 0) {
    const double4 pi = pos[i];
    const int c = cell_of_slot[i];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    const int myid = HAS_EXCL ? id[i] : 0;
    // Q16: particle i's 16-bit copy (the same for every row: the 16-bit
    // differences pick the periodic image themselves)
    v2u16 lxy = {0, 0}, lz = {0, 0};
    if (Q16) {
      const uint2 h = q16_pack(pi, q16inv);
      lxy = __builtin_bit_cast(v2u16, h.x);
      lz = __builtin_bit_cast(v2u16, h.y);
    }
    uint4* out = nbr + tiled_group(i, capg, 0);
    const int cap = 8 * capg;
    // entries not yet stored, in LDS: the pending part of the current 8-entry
    // group at positions [0, k & 7), a block's new entries after it; after each
    // block the groups it completed are stored, one 16-byte write each, and the
    // partial group moves back to positions 0..7 (a block adds at most kCandChunk
    // entries to fewer than 8 pending, so they fit).  Position p of this thread is
    // ring[p * 256]: the threads of a wavefront sit side by side at every
    // position, so their 16-bit accesses spread over all banks whatever
    // positions they address.
    unsigned short* ring = build_ring + threadIdx.x;
    // the prefilter's top as a length, with a relative margin for FP32 rounding
    const float top = sqrtf(cut_hi) * 1.0001f;
    const float top2 = top * top;
    const float cs_x = (float)(g.L[0] / nx), inv_cs = 1.0f / cs_x;
    const float bmargin = 1e-4f * cs_x;
    const float cs_y = (float)(g.L[1] / ny), cs_z = (float)(g.L[2] / nz);
    // With folded positions (row_margin >= 0) a particle binned into row (qy, qz)
    // lay inside the row's y and z range and has since moved by less than
    // row_margin, so every candidate of the row is at least the particle's
    // distance to that range (less the margin) away in y and z: the row's reach
    // in x shrinks to the rest of the prefilter's top, and a row beyond the top
    // is skipped.  1e-4 of a cell covers the FP32 rounding of the bounds.
    const bool rowclip = row_margin >= 0.0f;
    const float ymar = row_margin + 1e-4f * cs_y, zmar = row_margin + 1e-4f * cs_z;
    for (int oz = -1; oz <= 1; ++oz) {
      int qz = cz + oz, sz = 0;
      if (qz < 0) { qz += nz; sz = -1; } else if (qz >= nz) { qz -= nz; sz = 1; }
      // particle i moved by minus the candidate cells' periodic shift, in FP32
      const float zs = (float)(pi.z - (double)sz * g.L[2]);
      const float zlo = (float)qz * cs_z;
      const float dz_lo = fmaxf(fmaxf(zlo - zs, zs - (zlo + cs_z)) - zmar, 0.0f);
      for (int oy = -1; oy <= 1; ++oy) {
        int qy = cy + oy, sy = 0;
        if (qy < 0) { qy += ny; sy = -1; } else if (qy >= ny) { qy -= ny; sy = 1; }
        const float ys = (float)(pi.y - (double)sy * g.L[1]);
        const float ylo = (float)qy * cs_y;
        const float dy_lo = fmaxf(fmaxf(ylo - ys, ys - (ylo + cs_y)) - ymar, 0.0f);
        const float rx2 = rowclip ? top2 - __builtin_fmaf(dy_lo, dy_lo, dz_lo * dz_lo) : top2;
        if (!(rx2 > 0.0f)) continue;
        // reach in x: the row's share of the prefilter's top, plus clip_margin when
        // the x order inside the cells may be stale: a particle moved by at most d
        // since its cell was sorted leaves the order wrong by at most 2d, and a
        // stretch with that much more reach keeps every candidate within the top
        const float reach = sqrtf(rx2) * 1.0001f + clip_margin;
        const int row = (qy + ny * qz) * nx;
        // x pieces of the row segment (cx-1 .. cx+1), split where the row wraps
        int pa0, pb0, pa1 = 0, pb1 = -1;
        if (cx >= 1 && cx + 1 < nx) { pa0 = cx - 1; pb0 = cx + 1; }
        else if (cx == 0) { pa0 = nx - 1; pb0 = nx - 1; pa1 = 0; pb1 = 1; }
        else { pa0 = cx - 1; pb0 = nx - 1; pa1 = 0; pb1 = 0; }
        for (int p = 0; p < 2; ++p) {
          const int xa = p ? pa1 : pa0, xb = p ? pb1 : pb0;
          if (xb < xa) break;
          // the piece lies across the periodic boundary from i: cell nx-1 seen from
          // cx = 0, cell 0 seen from cx = nx-1
          const int sxp = (cx == 0 && p == 0) ? -1 : ((cx == nx - 1 && p == 1) ? 1 : 0);
          const float xs = (float)(pi.x - (double)sxp * g.L[0]);
          const int ca = row + xa;
          const int jb = cstart[ca], je0 = cstart[row + xb + 1];
          const TileRun R = sr[find_run_by_cell(sr, nr, ca)];
          const int base = R.lds_off + (jb - R.slot_start);
          // the piece is sorted by x, so x_i - x_j falls along it: keep a stretch
          // holding every j with |x_i - x_j| <= reach.  ja and je come from the
          // cells' x sub-bins (k_xbins): a bin boundary at or below xs - reach and
          // one above xs + reach (bmargin covers the FP32 rounding of the bin
          // arithmetic), so the stretch is that of an exact search widened to the
          // enclosing sub-bins.
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
              const int k = min(max((int)floorf((ulo - (float)cc) * (float)kXBins), 0), kXBins - 1);
              ja = xbin[(size_t)(row + cc) * (kXBins + 1) + k];
            }
            if (!(uhi < (float)(xb + 1))) {
              je = je0;
            } else if (uhi < (float)xa) {
              je = jb;
            } else {
              const int cc = max((int)floorf(uhi), xa);
              const int k = min(max((int)floorf((uhi - (float)cc) * (float)kXBins), 0) + 1, kXBins);
              je = xbin[(size_t)(row + cc) * (kXBins + 1) + k];
            }
            je = max(je, ja);
          }
          for (int j0 = ja; j0 < je; j0 += kCandChunk) {
            const int jn = je - j0;
            uint32_t in_bits = 0, maybe_bits = 0;
            // candidates kCandChunk-1 .. 0, so that candidate t ends at bit t of the masks;
            // each test shifts in the sign bit of r^2 - cut_lo (surely in) and of
            // r^2 - cut_hi (in or undecided: the window's top lies above every r^2
            // that may be in)
            if (Q16) {
              // r^2 - cut_lo in 16-bit units: two packed differences and two dot
              // products; the sums stay below 2^31 for every candidate of a row
              const uint2* cand = posh + j0;
              const int neg_tlo = -q16_tlo;
#pragma unroll
              for (int u = kCandChunk - 4; u >= 0; u -= 4) {
                uint2 q[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) q[v] = cand[u + v];
#pragma unroll
                for (int v = 3; v >= 0; --v) {
                  const v2i16 dxy = __builtin_bit_cast(v2i16, lxy - __builtin_bit_cast(v2u16, q[v].x));
                  const v2i16 dz = __builtin_bit_cast(v2i16, lz - __builtin_bit_cast(v2u16, q[v].y));
                  const int t = sdot2_acc(dz, sdot2_sacc(dxy, neg_tlo));
                  in_bits = __builtin_amdgcn_alignbit(in_bits, (uint32_t)t, 31);
                  maybe_bits = __builtin_amdgcn_alignbit(maybe_bits, (uint32_t)(t - q16_w), 31);
                }
              }
            } else {
              const float4* cand = posf + j0;
              const float win = cut_hi - cut_lo;
#pragma unroll
              for (int u = kCandChunk - 4; u >= 0; u -= 4) {
                float4 q[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) q[v] = cand[u + v];
#pragma unroll
                for (int v = 3; v >= 0; --v) {
                  const float dx = xs - q[v].x, dy = ys - q[v].y, dz = zs - q[v].z;
                  // r^2 - cut_lo with the cut folded into the chain of FMAs
                  const float t = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, -cut_lo)));
                  in_bits = __builtin_amdgcn_alignbit(in_bits, __float_as_uint(t), 31);
                  maybe_bits = __builtin_amdgcn_alignbit(maybe_bits, __float_as_uint(t - win), 31);
                }
              }
            }
            uint32_t valid = jn >= kCandChunk ? kCandAll : ((1u << jn) - 1u);
            const int self = i - j0;
            if (self >= 0 && self < kCandChunk) valid &= ~(1u << self);
            in_bits &= valid;
            uint32_t border = maybe_bits & valid & ~in_bits;
            while (border) {
              // the reference's FP64 test in the reference's orientation (rare)
              const int t = __ffs(border) - 1;
              border &= border - 1u;
              const int j = j0 + t;
              const int cjx = cell_of_slot[j] % nx;
              int ox = cjx - cx, sx = 0;
              if (ox > 1) { ox = -1; sx = -1; } else if (ox < -1) { ox = 1; sx = 1; }
              const bool fwd = g.hci > 1 ? caller_forward(g.hcell[i], g.hcell[j], sx, sy, sz, g)
                                         : (oz > 0) || (oz == 0 && (oy > 0 || (oy == 0 && ox >= 0)));
              const double4 pj = pos[j];
              const double shx = (double)sx * g.L[0], shy = (double)sy * g.L[1], shz = (double)sz * g.L[2];
              double ex, ey, ez;
              if (fwd) { ex = pi.x - (pj.x + shx); ey = pi.y - (pj.y + shy); ez = pi.z - (pj.z + shz); }
              else { ex = pj.x - (pi.x - shx); ey = pj.y - (pi.y - shy); ez = pj.z - (pi.z - shz); }
              if (exact_sqr_le(ex, ey, ez, cutsq)) in_bits |= 1u << t;
            }
            if (HAS_EXCL) {
              // drop the excluded pairs (one lookup per accepted candidate)
              for (uint32_t b = in_bits; b; b &= b - 1u) {
                const int t = __ffs(b) - 1;
                if (is_excluded(myid, id[j0 + t], ex_start, ex_ids)) in_bits &= ~(1u << t);
              }
            }
            const uint32_t ebase = (uint32_t)(base + (j0 - jb));   // entry of candidate 0
            // Every candidate t writes its entry at the ring position of its rank
            // among the accepted ones, the k & 7 pending entries first; a rejected
            // one writes to the next accepted one's position (overwritten by it,
            // as later writes of a lane land after earlier ones) or past the last.
            // No branch, so no lane waits on another's count.
            const int kp = k & 7;
#pragma unroll
            for (int t = 0; t < kCandChunk; ++t)
              ring[(kp + __popc(in_bits & ((1u << t) - 1u))) * 256] = (unsigned short)(ebase + (uint32_t)t);
            const int kn = kp + __popc(in_bits);
            // the groups this block completed (at most two), inside the capacity,
            // then the partial group back to positions 0..7
            const int g0 = k >> 3;
            for (int gq = 0; gq < min(kn >> 3, capg - g0); ++gq) out[(size_t)(g0 + gq) * 64] = ring_group(ring, 8 * gq);
            if (kn >= 8) {
              const int src = 8 * (kn >> 3);
#pragma unroll
              for (int e = 0; e < 8; ++e) ring[e * 256] = ring[(src + e) * 256];
            }
            k += kn - kp;
          }
        }
      }
    }
    // last partial group: the entries past the count are zero (offset 0 stages a
    // valid slot; the pair kernels read but never use them)
    if ((k & 7) != 0 && k < cap) {
      for (int r = k & 7; r < 8; ++r) ring[r * 256] = 0;
      out[(size_t)(k >> 3) * 64] = ring_group(ring, 0);
    }
    cnt[i] = k;
    if (k > cap) ctl->overflow = 1;
  } else if (i < n) {
    cnt[i] = 0;
  }
  // The pair kernels take the tile's particles in order of this list length,
  // rounded up to their 16-entry iterations, and in slot order within one
  // length class (lane_slot[s0 + r] is the slot of rank r).  The lanes of a
  // wavefront then run the same number of iterations, and neighbouring
  // particles, whose lists are alike, share a wavefront, so their gathers of
  // the staged positions fall on nearby LDS addresses.  Slots past n rank last.
  const int rank = block_class_rank(i < n ? min((k + 15) >> 4, 62) : 63);
  if (i < n) lane_slot[s0 + rank] = i;
  block_max_to(k, &ctl->maxcount);
}

// ------------------------------------------------------------------ step helpers

// Thermostat on particle i's force (LangevinThermostat::frictionThermo,
// src/integrator/LangevinThermostat.cpp:141-177); false when it does not apply.
__device__ __forceinline__ bool step_thermostat(int i, const double4& v, double& f0, double& f1, double& f2,
                                                const StepArgs& a) {
#pragma clang fp contract(off)
  if (!a.lang) return false;
  const int pid = a.id[i];
  if (a.excl && pid < a.nid && a.excl[pid]) return false;
  uint32_t o[4];
  philox((uint32_t)pid, (uint32_t)a.call, (uint32_t)(a.call >> 32), 0x4C414E47u, (uint32_t)a.seed,
         (uint32_t)(a.seed >> 32), o);
  const double u0 = ((double)o[0] + 0.5) * 2.3283064365386963e-10;
  const double u1 = ((double)o[1] + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)o[2] + 0.5) * 2.3283064365386963e-10;
  const double m = v.w, massf = sqrt(m);
  f0 += (a.pref1 * v.x) * m + (a.pref2 * (u0 - 0.5)) * massf;
  f1 += (a.pref1 * v.y) * m + (a.pref2 * (u1 - 0.5)) * massf;
  f2 += (a.pref1 * v.z) * m + (a.pref2 * (u2 - 0.5)) * massf;
  return true;
}

// integrate2 (when do_int2) and integrate1 of particle i (velocity v, position
// p) on force f (src/integrator/VelocityVerlet.cpp:236-305); returns |dx|^2.
__device__ __forceinline__ double step_kick_drift(int i, double4 v, double4 p, double f0, double f1, double f2,
                                                  const StepArgs& a, const Geom& g) {
#pragma clang fp contract(off)
  const double dt = a.dt;
  const double dtfm = 0.5 * dt / v.w;
  if (a.do_int2) {
    v.x += dtfm * f0; v.y += dtfm * f1; v.z += dtfm * f2;
  }
  v.x += dtfm * f0; v.y += dtfm * f1; v.z += dtfm * f2;
  const double d0 = v.x * dt, d1 = v.y * dt, d2 = v.z * dt;
  p.x += d0; p.y += d1; p.z += d2;
  a.vel[i] = v;
  a.pos_out[i] = p;
  if (a.posq_out)
    a.posq_out[i] = make_uint4(fixq(p.x, g.qinv[0]), fixq(p.y, g.qinv[1]), fixq(p.z, g.qinv[2]), (uint32_t)(int)p.w);
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// The block's maximum of sq into a.blockmax[blockIdx.x] (every thread calls it).
__device__ __forceinline__ void step_block_max(double sq, const StepArgs& a) {
  __shared__ double red[4];
  sq = wave_max(sq);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    a.blockmax[blockIdx.x] = m;
  }
}

// Mixed-precision LJ over the tiled list.  Neighbour positions come from the
// 32-bit fixed-point mirror staged in LDS; their differences are exact
// integers.  r^2 is formed exactly in FP64 from them, the reciprocal is seeded
// by v_rcp_f32 and refined by one FP64 Newton step, and the LJ factor and the
// force sum are FP64 (on CDNA3/4 FP64 FMA issues at the FP32 rate, only the
// reciprocal is cheaper in FP32).  For cubic boxes everything stays in
// fixed-point units (LJParam.u*) and the force is scaled once at the end.
// Pairs within a tiny window of the cutoff (fixed-point quantisation could flip
// the test) are left out of the main loop and settled from the FP64 positions
// afterwards, so the interacting set equals the FP64 path's.
// STEP: the kernel is the last force of the step and finishes it per particle:
// thermostat, integrate2 of this step and integrate1 of the next on the
// completed force (step_thermostat / step_kick_drift, the arithmetic of
// k_step_fused), writing the new positions to separate buffers (sa.pos_out,
// sa.posq_out) because other workgroups still stage the current ones.  The
// force itself is not stored: the next step recomputes it, and the last step
// of a run takes the unfused path.  The particle's velocity and position are
// loaded before the pair loop so that the loads overlap it.
// PRUNE (not with ENERGY): the walk also writes the pruned list po, the entries with
// r^2 < po.r2 at this step's positions in list order (the entries it drops have
// r^2 above the cutoff window, so they add exactly zero here), each lane
// appending to a 128-bit shift register as the list build does; then the
// pruned lengths rank the tile's lanes for the walks of the pruned list.
template <bool ONETYPE, bool ENERGY, bool ACCUM, bool CUBIC, bool STEP, bool PRUNE = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(ONETYPE && STEP && !PRUNE && !ENERGY ? 4 : 1))) void k_lj_tiled_mixed(
    int n, int ntiles, const uint4* __restrict__ posq, const double4* __restrict__ pos,
    const uint4* __restrict__ nbr, const int* __restrict__ cnt, int capg, Geom g,
    const TileRun* __restrict__ runs, const int* __restrict__ nruns, const int* __restrict__ lane_slot,
    const LJParam* __restrict__ prm, int ntypes, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ partial, StepArgs sa, PruneOut po) {
  static_assert(!(STEP && ENERGY), "the step variant computes forces only");
  static_assert(!(PRUNE && ENERGY), "the list is pruned by force walks");
  if (sa.stop && *reinterpret_cast<const volatile int*>(sa.stop)) return;   // a window's step after the stop
  // one type: LDS holds x, y, z (12 bytes); several types: x, y, z, type
  using Slot = typename std::conditional<ONETYPE, uint3, uint4>::type;
  extern __shared__ uint4 lds_raw[];
  Slot* lds = reinterpret_cast<Slot*>(lds_raw);
  __shared__ RunRef sr[kMaxRuns + 1];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const RunTable rt = load_run_table(runs, nruns, tile);
  const int L = tile * blockDim.x + threadIdx.x;   // list lane
  const bool live = L < n;
  // kGroups 16-byte groups (8 entries each) per iteration, the next ones loaded
  // during it; with several types one group, so the per-entry type-pair
  // coefficients of a batch stay within the register budget
  constexpr int kGroups = ONETYPE ? 2 : 1;
  constexpr int kBatch = 8 * kGroups;              // entries per iteration
  // The lane's own data and its first list groups do not depend on the staged
  // region: their loads go out before the staging, unconditionally (lanes past
  // n read lane n-1's and drop it), so that later waits can count them.
  const int Lc = live ? L : n - 1;
  const int i = lane_slot[Lc];                     // particle slot, which also indexes its list
  const int clist = cnt[i];                        // list length
  const uint4* np = nbr + tiled_group(i, capg, 0);
  uint4 wn[kGroups];
#pragma unroll
  for (int q = 0; q < kGroups; ++q) wn[q] = np[(size_t)min(q, capg - 1) * 64];
  // the particle's own mirror, velocity and position are loaded behind the
  // staging loads, so the staging does not wait for them
  uint4 qi;
  double4 step_v, step_p;
  const int nr = stage_tile_q<Slot>(rt, posq, lds, sr, [&] {
    qi = posq[i];
    if (STEP) { step_v = sa.vel[i]; step_p = sa.pos[i]; }
  });
  double stepsq = 0.0;                             // |dx|^2 of the drift (STEP)
  // PRUNE: entries not yet stored (the newest at the top of pw3) and their count
  uint32_t pw0 = 0, pw1 = 0, pw2 = 0, pw3 = 0;
  int kp = 0;
  uint4* pout = PRUNE ? po.nbr + tiled_group(i, capg, 0) : nullptr;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (live) {
    const int c = nr > 0 ? min(clist, 8 * capg) : 0;
    const int rowi = (int)qi.w * ntypes;
    const LJParam& p0 = prm[0];
    double ff1 = CUBIC ? p0.uff1 : p0.dff1, ff2 = CUBIC ? p0.uff2 : p0.dff2;
    // r^2 below lo is inside the cutoff; r^2 in [lo, hi] may be on either side
    // after fixed-point quantisation and is settled after the batch from the
    // FP64 positions; r^2 above hi is outside
    double lo = CUBIC ? p0.ulo : p0.dlo, hi = CUBIC ? p0.uhi : p0.dhi;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    const int ng = (c + 7) >> 3;
    const int nit = (ng + kGroups - 1) / kGroups;    // iterations
#pragma unroll
    for (int q = 0; q < kGroups; ++q)
      if (q >= ng) wn[q] = make_uint4(0u, 0u, 0u, 0u);
    // One iteration over entries [kBatch * it, kBatch * it + kBatch).  Only the
    // last iteration can hold entries past the list length c (zero offsets, which
    // stage a valid slot), so only CHECKED tests each entry against c.  The
    // positions of all the batch's entries are read from LDS before the first is
    // used, so the gathers overlap the arithmetic of the batch.
    auto batch = [&](int it, auto checked) {
      constexpr bool CHECKED = decltype(checked)::value;
      const int kg = it * kGroups;
      uint32_t w[4 * kGroups];
#pragma unroll
      for (int q = 0; q < kGroups; ++q) {
        w[4 * q] = wn[q].x; w[4 * q + 1] = wn[q].y; w[4 * q + 2] = wn[q].z; w[4 * q + 3] = wn[q].w;
        wn[q] = (kg + kGroups + q < ng) ? np[(size_t)(kg + kGroups + q) * 64] : make_uint4(0u, 0u, 0u, 0u);
      }
      Slot qa[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) qa[u] = (u & 1) ? lds_entry<true>(lds, w[u >> 1]) : lds_entry<false>(lds, w[u >> 1]);
      bool anynear = false;
      uint32_t keep = 0;                           // PRUNE: entries of the batch that stay
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const bool valid = !CHECKED || kBatch * it + u < c;
        const Slot qj = qa[u];
        if (!ONETYPE) {
          int tj = 0;
          if constexpr (!ONETYPE) tj = (int)qj.w;
          const LJParam& p = prm[rowi + tj];
          ff1 = CUBIC ? p.uff1 : p.dff1; ff2 = CUBIC ? p.uff2 : p.dff2;
          lo = p.set ? (CUBIC ? p.ulo : p.dlo) : -1.0;
          hi = p.set ? (CUBIC ? p.uhi : p.dhi) : -1.0;
        }
        double ex, ey, ez;
        const double r2 = pair_r2<CUBIC>(qi, (int)(qi.x - qj.x), (int)(qi.y - qj.y), (int)(qi.z - qj.z), g, ex, ey, ez);
        const bool in = valid & (r2 < lo);
        // in the window: at most hi and not in (in implies at most hi, hence the xor)
        anynear |= (valid & (r2 <= hi)) ^ in;
        if (PRUNE) keep |= (valid && r2 < po.r2) ? (1u << u) : 0u;
        double inv;
        const double ff = lj_ff_masked(r2, in, ff1, ff2, inv);
        Ax = __builtin_fma(ex, ff, Ax);
        Ay = __builtin_fma(ey, ff, Ay);
        Az = __builtin_fma(ez, ff, Az);
        if (ENERGY && in) {
          int tj = 0;
          if constexpr (!ONETYPE) tj = (int)qj.w;
          const LJParam& p = prm[ONETYPE ? 0 : rowi + tj];
          const double f2 = (CUBIC ? p.usig2 : p.sig2) * inv;
          const double f6 = f2 * f2 * f2;
          acc[0] += p.eps4 * (f6 * f6 - f6) - p.dshift;
          acc[1] += r2 * ff;
          acc[2] += ex * ex * ff; acc[3] += ey * ey * ff; acc[4] += ez * ez * ff;
          acc[5] += ex * ey * ff; acc[6] += ex * ez * ff; acc[7] += ey * ez * ff;
        }
      }
      if (PRUNE) {
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          if (keep & (1u << u)) {
            const uint32_t e = (u & 1) ? (w[u >> 1] >> 16) : (w[u >> 1] & 0xFFFFu);
            pw0 = __builtin_amdgcn_alignbit(pw1, pw0, 16);
            pw1 = __builtin_amdgcn_alignbit(pw2, pw1, 16);
            pw2 = __builtin_amdgcn_alignbit(pw3, pw2, 16);
            pw3 = __builtin_amdgcn_alignbit(e, pw3, 16);
            ++kp;
            if ((kp & 7) == 0) pout[(size_t)((kp >> 3) - 1) * 64] = make_uint4(pw0, pw1, pw2, pw3);
          }
        }
      }
      if (anynear) {
        // rare: the batch holds a pair in the cutoff window; decide it from the FP64 positions
        for (int u = 0; u < 8 * kGroups && 8 * kg + u < c; ++u) {
          const int e = (u & 1) ? (int)(w[u >> 1] >> 16) : (int)(w[u >> 1] & 0xFFFFu);
          const Slot qj = lds[e];
          int tj = 0;
          if constexpr (!ONETYPE) tj = (int)qj.w;
          const LJParam& p = prm[ONETYPE ? 0 : rowi + tj];
          if (!p.set) continue;
          double ex, ey, ez;
          const double r2 = pair_r2<CUBIC>(qi, (int)(qi.x - qj.x), (int)(qi.y - qj.y), (int)(qi.z - qj.z), g, ex, ey, ez);
          if (r2 < (CUBIC ? p.ulo : p.dlo) || !(r2 <= (CUBIC ? p.uhi : p.dhi))) continue;
          const int j = slot_of_lds(sr, nr, e);
          if (!(ref_pair_r2(i, j, pos[i], pos[j], g) <= p.drc2)) continue;
          double inv;
          const double ff = lj_ff_seeded(r2, CUBIC ? p.uff1 : p.dff1, CUBIC ? p.uff2 : p.dff2, inv);
          Ax = __builtin_fma(ex, ff, Ax);
          Ay = __builtin_fma(ey, ff, Ay);
          Az = __builtin_fma(ez, ff, Az);
          if (ENERGY) {
            const double f2 = (CUBIC ? p.usig2 : p.sig2) * inv;
            const double f6 = f2 * f2 * f2;
            acc[0] += p.eps4 * (f6 * f6 - f6) - p.dshift;
            acc[1] += r2 * ff;
            acc[2] += ex * ex * ff; acc[3] += ey * ey * ff; acc[4] += ez * ez * ff;
            acc[5] += ex * ey * ff; acc[6] += ex * ez * ff; acc[7] += ey * ez * ff;
          }
        }
      }
    };
    for (int it = 0; it + 1 < nit; ++it) batch(it, std::false_type{});
    if (nit > 0) batch(nit - 1, std::true_type{});
    if (CUBIC) {
      // back from fixed-point units: F = d_units * s * ff_units; w = r2_units * s^2 * ff_units
      Ax *= g.qsd[0]; Ay *= g.qsd[0]; Az *= g.qsd[0];
      if (ENERGY) {
#pragma unroll
        for (int q = 1; q < kNPart; ++q) acc[q] *= g.qs2;
      }
    }
    if (STEP) {
      double f0 = Ax, f1 = Ay, f2 = Az;
      if (ACCUM) { f0 = fx[i] + Ax; f1 = fy[i] + Ay; f2 = fz[i] + Az; }
      step_thermostat(i, step_v, f0, f1, f2, sa);
      stepsq = step_kick_drift(i, step_v, step_p, f0, f1, f2, sa, g);
    } else if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
    if (PRUNE) {
      // the last partial group shifted down to its place, zero entries above it
      if ((kp & 7) != 0) {
        for (int r = kp & 7; r < 8; ++r) {
          pw0 = __builtin_amdgcn_alignbit(pw1, pw0, 16);
          pw1 = __builtin_amdgcn_alignbit(pw2, pw1, 16);
          pw2 = __builtin_amdgcn_alignbit(pw3, pw2, 16);
          pw3 = pw3 >> 16;
        }
        pout[(size_t)(kp >> 3) * 64] = make_uint4(pw0, pw1, pw2, pw3);
      }
      po.cnt[i] = kp;
    }
  }
  if (PRUNE) {
    // the tile's lanes in order of the pruned lengths, as the list build orders them
    const int rank = block_class_rank(live ? min((kp + 15) >> 4, 62) : 63);
    if (live) po.lane[tile * blockDim.x + rank] = i;
  }
  if (STEP) step_block_max(stepsq, sa);
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// All-FP64 LJ over the tiled list; LDS holds FP64 positions with the type in w.
template <bool ONETYPE, bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_lj_tiled_fp64(
    int n, int ntiles, const double4* __restrict__ pos, const uint4* __restrict__ nbr,
    const int* __restrict__ cnt, int capg, Geom g, const TileRun* __restrict__ runs,
    const int* __restrict__ nruns, const int* __restrict__ lane_slot, const LJParam* __restrict__ prm,
    int ntypes, double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    double* __restrict__ partial) {
  extern __shared__ double4 ldsd[];
  __shared__ RunRef sr[kMaxRuns + 1];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int nr = stage_tile<double4>(runs, nruns, pos, ldsd, sr, tile);
  const int L = tile * blockDim.x + threadIdx.x;   // list lane
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (L < n) {
    const int i = lane_slot[L];                      // particle slot
    const double4 pi = pos[i];
    const uint4* np = nbr + tiled_group(i, capg, 0);
    const int c = nr > 0 ? min(cnt[i], 8 * capg) : 0;
    const int rowi = (int)pi.w * ntypes;
    double ff1 = prm[0].dff1, ff2 = prm[0].dff2, rc2 = prm[0].drc2, shift = prm[0].dshift;
    double sig2 = prm[0].sig2, eps4 = prm[0].eps4, lo = prm[0].dlo, hi = prm[0].dhi;
    int set = prm[0].set;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    const int ng = (c + 7) >> 3;
    uint4 wn = ng > 0 ? np[0] : make_uint4(0u, 0u, 0u, 0u);
    for (int kg = 0; kg < ng; ++kg) {
      const uint32_t w[4] = {wn.x, wn.y, wn.z, wn.w};
      wn = (kg + 1 < ng) ? np[(size_t)(kg + 1) * 64] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = 8 * kg + u;
        if (k >= c) break;
        const int e = (u & 1) ? (int)(w[u >> 1] >> 16) : (int)(w[u >> 1] & 0xFFFFu);
        const double4 pj = ldsd[e];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx = __builtin_fma(-g.L[0], __builtin_rint(dx * g.invL[0]), dx);
        dy = __builtin_fma(-g.L[1], __builtin_rint(dy * g.invL[1]), dy);
        dz = __builtin_fma(-g.L[2], __builtin_rint(dz * g.invL[2]), dz);
        if (!ONETYPE) {
          const LJParam& p = prm[rowi + (int)pj.w];
          ff1 = p.dff1; ff2 = p.dff2; rc2 = p.drc2; shift = p.dshift; sig2 = p.sig2; eps4 = p.eps4;
          set = p.set; lo = p.dlo; hi = p.dhi;
        }
        const double r2 = sqr_nc(dx, dy, dz);
        bool in = set && r2 <= rc2;
        if (set && r2 >= lo && r2 <= hi)   // at the cutoff (rare)
          in = ref_pair_r2(i, slot_of_lds(sr, nr, e), pi, pos[slot_of_lds(sr, nr, e)], g) <= rc2;
        const double inv = rcp_nr(r2);
        const double r6 = inv * inv * inv;
        double ff = r6 * (ff1 * r6 - ff2) * inv;
        ff = in ? ff : 0.0;
        Ax = __builtin_fma(dx, ff, Ax);
        Ay = __builtin_fma(dy, ff, Ay);
        Az = __builtin_fma(dz, ff, Az);
        if (ENERGY) {
          const double f2 = sig2 * inv;
          const double f6 = f2 * f2 * f2;
          acc[0] += in ? eps4 * (f6 * f6 - f6) - shift : 0.0;
          acc[1] += r2 * ff;
          acc[2] += dx * dx * ff; acc[3] += dy * dy * ff; acc[4] += dz * dz * ff;
          acc[5] += dx * dy * ff; acc[6] += dx * dz * ff; acc[7] += dy * dz * ff;
        }
      }
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// All-FP64 LJ for one particle type, two lanes per particle: a tile of T
// particles runs 2T threads (the two lanes of a particle take alternate
// 8-entry groups of its list and add their sums with one shuffle), and the
// staged region holds x, y, z only (24 bytes per particle instead of 32).  A
// group's eight LDS reads are issued before its arithmetic.  Per pair the
// arithmetic is the single-lane kernel's (minimum image by rint, FP64 r^2
// without contraction, the reference's test for pairs at the cutoff).
struct P3 {
  double x, y, z;
};

template <bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_lj_tiled_fp64_pair(
    int n, int ntiles, int T, const double4* __restrict__ pos, const uint4* __restrict__ nbr,
    const int* __restrict__ cnt, int capg, Geom g, const TileRun* __restrict__ runs, const int* __restrict__ nruns,
    const int* __restrict__ lane_slot, const LJParam* __restrict__ prm, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ partial) {
  extern __shared__ uint4 lds_raw[];
  P3* lds = reinterpret_cast<P3*>(lds_raw);
  __shared__ RunRef sr[kMaxRuns + 1];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int nr = stage_region(load_run_table(runs, nruns, tile), pos, lds, sr,
                              [](const double4& v) { return P3{v.x, v.y, v.z}; }, [] {});
  const int h = (int)threadIdx.x & 1;               // which alternate groups this lane takes
  const int L = tile * T + ((int)threadIdx.x >> 1); // list lane
  const bool live = L < n;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  double Ax = 0.0, Ay = 0.0, Az = 0.0;
  int i = 0;
  if (live) {
    i = lane_slot[L];
    const double4 pi = pos[i];
    const uint4* np = nbr + tiled_group(i, capg, 0);
    const int c = nr > 0 ? min(cnt[i], 8 * capg) : 0;
    const LJParam& p = prm[0];
    const double ff1 = p.dff1, ff2 = p.dff2, rc2 = p.drc2, lo = p.dlo, hi = p.dhi;
    const int set = p.set;
    const int ng = (c + 7) >> 3;
    uint4 wn = h < ng ? np[(size_t)h * 64] : make_uint4(0u, 0u, 0u, 0u);
    for (int kg = h; kg < ng; kg += 2) {
      const uint32_t w[4] = {wn.x, wn.y, wn.z, wn.w};
      wn = (kg + 2 < ng) ? np[(size_t)(kg + 2) * 64] : make_uint4(0u, 0u, 0u, 0u);
      P3 qa[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = (u & 1) ? (int)(w[u >> 1] >> 16) : (int)(w[u >> 1] & 0xFFFFu);
        qa[u] = lds[e];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (8 * kg + u >= c) break;
        double dx = pi.x - qa[u].x, dy = pi.y - qa[u].y, dz = pi.z - qa[u].z;
        dx = __builtin_fma(-g.L[0], __builtin_rint(dx * g.invL[0]), dx);
        dy = __builtin_fma(-g.L[1], __builtin_rint(dy * g.invL[1]), dy);
        dz = __builtin_fma(-g.L[2], __builtin_rint(dz * g.invL[2]), dz);
        const double r2 = sqr_nc(dx, dy, dz);
        bool in = set && r2 <= rc2;
        if (set && r2 >= lo && r2 <= hi) {   // at the cutoff (rare)
          const int e = (u & 1) ? (int)(w[u >> 1] >> 16) : (int)(w[u >> 1] & 0xFFFFu);
          const int j = slot_of_lds(sr, nr, e);
          in = ref_pair_r2(i, j, pi, pos[j], g) <= rc2;
        }
        const double inv = rcp_nr(r2);
        const double r6 = inv * inv * inv;
        double ff = r6 * (ff1 * r6 - ff2) * inv;
        ff = in ? ff : 0.0;
        Ax = __builtin_fma(dx, ff, Ax);
        Ay = __builtin_fma(dy, ff, Ay);
        Az = __builtin_fma(dz, ff, Az);
        if (ENERGY) {
          const double f2 = p.sig2 * inv;
          const double f6 = f2 * f2 * f2;
          acc[0] += in ? p.eps4 * (f6 * f6 - f6) - p.dshift : 0.0;
          acc[1] += r2 * ff;
          acc[2] += dx * dx * ff; acc[3] += dy * dy * ff; acc[4] += dz * dz * ff;
          acc[5] += dx * dy * ff; acc[6] += dx * dz * ff; acc[7] += dy * dz * ff;
        }
      }
    }
  }
  if (!ENERGY) {
    Ax += __shfl_xor(Ax, 1, 64);
    Ay += __shfl_xor(Ay, 1, 64);
    Az += __shfl_xor(Az, 1, 64);
    if (live && h == 0) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
  block_reduce_store<kNPart>(acc, partial);
}

// ------------------------------------------------------------------ Tabulated
//
// interaction.Tabulated over a Verlet list or a fixed pair list
// (src/interaction/Tabulated.hpp:100-115: F = d * table.getForce(r) / r,
// E = table.getEnergy(r) - shift, cut test distSqr > cutoffSqr).  Distances and
// table evaluation are FP64 in both precisions; the mixed path forms d from the
// exact fixed-point differences and settles pairs in the cutoff window from the
// FP64 positions.

// Interpolation of one table row at r (InterpolationLinear.hpp:58-75,
// InterpolationAkima.hpp:86-106, InterpolationCubic.hpp:89-120), without FMA
// contraction, as the reference evaluates it.
__device__ __forceinline__ double tab_row_eval(const double* __restrict__ row, int itype, double r,
                                               double invdelta, double deltasq6) {
#pragma clang fp contract(off)
  if (itype == 1) return row[1] * r + row[2];
  if (itype == 2) {
    const double z = r - row[0];
    const int zz2 = (int)(z * z);   // integer square, as the reference stores it
    return row[1] + row[2] * z + row[3] * zz2 + row[4] * zz2 * z;
  }
  const double b = (r - row[0]) * invdelta;
  const double a = 1.0 - b;
  return a * row[1] + b * row[3] + ((a * a * a - a) * row[2] + (b * b * b - b) * row[4]) * deltasq6;
}

// row index: truncation of (r - inner) / delta, kept inside the table's intervals
__device__ __forceinline__ int tab_index(const TabParam& t, double r) {
  int index = (int)((r - t.inner) * t.invdelta);
  index = index < 0 ? 0 : index;
  return index > t.N - 2 ? t.N - 2 : index;
}

// force factor F(r)/r of a pair at distance^2 r2 inside the cutoff
__device__ __forceinline__ double tab_ffactor(const TabParam& t, const double* __restrict__ pool, double r2) {
#pragma clang fp contract(off)
  const double r = sqrt(r2);
  const int idx = tab_index(t, r);
  double ff = tab_row_eval(pool + (size_t)t.off + (size_t)idx * kTabRow, t.itype, r, t.invdelta, t.deltasq6);
  return ff / r;
}

__device__ __forceinline__ double tab_energy(const TabParam& t, const double* __restrict__ pool, double r2) {
#pragma clang fp contract(off)
  const double r = sqrt(r2);
  const int idx = tab_index(t, r);
  return tab_row_eval(pool + (size_t)t.off + (size_t)(t.N + idx) * kTabRow, t.itype, r, t.invdelta, t.deltasq6) -
         t.shift;
}

// Adds the contribution of one pair (distance vector d, r2 = |d|^2) to the
// particle's force or to its energy/virial partials.
// r2c is the |d|^2 the cut test takes (the reference's own near the cutoff).
template <bool ENERGY>
__device__ __forceinline__ void tab_pair(const TabParam& t, const double* __restrict__ pool, double dx, double dy,
                                         double dz, double r2, double r2c, double& Ax, double& Ay, double& Az,
                                         double (&acc)[kNPart]) {
#pragma clang fp contract(off)
  if (!t.set || r2c > t.rc2) return;
  const double ff = tab_ffactor(t, pool, r2);
  const double Fx = dx * ff, Fy = dy * ff, Fz = dz * ff;
  if (!ENERGY) {
    Ax += Fx; Ay += Fy; Az += Fz;
  } else {
    acc[0] += tab_energy(t, pool, r2);
    acc[1] += dx * Fx + dy * Fy + dz * Fz;
    acc[2] += dx * Fx; acc[3] += dy * Fy; acc[4] += dz * Fz;
    acc[5] += dx * Fy; acc[6] += dx * Fz; acc[7] += dy * Fz;
  }
}

// FP64 minimum-image distance of slots i and j from the FP64 positions
__device__ __forceinline__ double min_image_fp64(const double4& pi, const double4& pj, const Geom& g, double& dx,
                                                 double& dy, double& dz) {
#pragma clang fp contract(off)
  dx = pi.x - pj.x; dy = pi.y - pj.y; dz = pi.z - pj.z;
  dx = __builtin_fma(-g.L[0], __builtin_rint(dx * g.invL[0]), dx);
  dy = __builtin_fma(-g.L[1], __builtin_rint(dy * g.invL[1]), dy);
  dz = __builtin_fma(-g.L[2], __builtin_rint(dz * g.invL[2]), dz);
  return dx * dx + dy * dy + dz * dz;
}

// Tabulated over the tiled list.  MIXED stages the fixed-point mirror (with
// type), FP64 stages the FP64 positions (type in w).
template <bool MIXED, bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_tab_tiled(
    int n, int ntiles, const uint4* __restrict__ posq, const double4* __restrict__ pos, const uint4* __restrict__ nbr,
    const int* __restrict__ cnt, int capg, Geom g, const TileRun* __restrict__ runs,
    const int* __restrict__ nruns, const int* __restrict__ lane_slot, const TabParam* __restrict__ prm,
    const double* __restrict__ pool, int ntypes, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ partial) {
  extern __shared__ uint4 lds_raw[];
  __shared__ RunRef sr[kMaxRuns + 1];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  int nr;
  if (MIXED) nr = stage_tile_q<uint4>(runs, nruns, posq, lds_raw, sr, tile);
  else nr = stage_tile<double4>(runs, nruns, pos, reinterpret_cast<double4*>(lds_raw), sr, tile);
  const int L = tile * blockDim.x + threadIdx.x;   // list lane
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (L < n) {
    const int i = lane_slot[L];                      // particle slot
    const double4 pi = pos[i];
    const uint4 qi = MIXED ? posq[i] : make_uint4(0u, 0u, 0u, 0u);
    const uint32_t* np = reinterpret_cast<const uint32_t*>(nbr + tiled_group(i, capg, 0));
    const int c = nr > 0 ? min(cnt[i], 8 * capg) : 0;
    const int rowi = (int)pi.w * ntypes;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    for (int k = 0; k < c; ++k) {
      const uint32_t w = np[(size_t)(k >> 3) * 256 + ((k >> 1) & 3)];
      const int e = (k & 1) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
      double dx, dy, dz, r2, r2c;
      int tj;
      if (MIXED) {
        const uint4 qj = lds_raw[e];
        tj = (int)qj.w;
        dx = (double)(int)(qi.x - qj.x) * g.qsd[0];
        dy = (double)(int)(qi.y - qj.y) * g.qsd[1];
        dz = (double)(int)(qi.z - qj.z) * g.qsd[2];
        r2 = dx * dx + dy * dy + dz * dz;
      } else {
        const double4 pj = reinterpret_cast<const double4*>(lds_raw)[e];
        tj = (int)pj.w;
        r2 = min_image_fp64(pi, pj, g, dx, dy, dz);
      }
      r2c = r2;
      const TabParam& t = prm[rowi + tj];
      if (r2 >= t.dlo && r2 <= t.dhi) {
        // at the cutoff, where quantisation or the rounding of d could flip the cut test:
        // the pair from the FP64 positions, and the reference's own |d|^2 for the test
        const int j = slot_of_lds(sr, nr, e);
        const double4 pj = pos[j];
        r2 = min_image_fp64(pi, pj, g, dx, dy, dz);
        r2c = ref_pair_r2(i, j, pi, pj, g);
      }
      tab_pair<ENERGY>(t, pool, dx, dy, dz, r2, r2c, Ax, Ay, Az, acc);
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// Tabulated over the untiled full list (neighbour entries are slots).
template <bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_tab(
    int n, const double4* __restrict__ pos, const int* __restrict__ nbr, const int* __restrict__ cnt, int cap,
    Geom g, const TabParam* __restrict__ prm, const double* __restrict__ pool, int ntypes, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (i < n) {
    const double4 pi = pos[i];
    const int* np = nbr + (size_t)(i >> 6) * (size_t)cap * 64 + (i & 63);
    const int c = min(cnt[i], cap);
    const int rowi = (int)pi.w * ntypes;
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    for (int k = 0; k < c; ++k) {
      const int j = np[(size_t)k * 64];
      const double4 pj = pos[j];
      double dx, dy, dz;
      const double r2 = min_image_fp64(pi, pj, g, dx, dy, dz);
      const TabParam& t = prm[rowi + (int)pj.w];
      const double r2c = (r2 >= t.dlo && r2 <= t.dhi) ? ref_pair_r2(i, j, pi, pj, g) : r2;
      tab_pair<ENERGY>(t, pool, dx, dy, dz, r2, r2c, Ax, Ay, Az, acc);
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// FixedPairListTabulated (src/interaction/FixedPairListInteractionTemplate.hpp:112-200
// with Tabulated): per-particle bond adjacency, FP64 minimum image of
// OrthorhombicBC::getMinimumImageVectorBox (src/bc/OrthorhombicBC.cpp:86-105).
template <bool ENERGY, bool ACCUM>
__global__ __launch_bounds__(256) void k_bond_tab(
    int n, const double4* __restrict__ pos, const int* __restrict__ id, const int* __restrict__ slot_of_id, int nid,
    const int* __restrict__ bstart, const int* __restrict__ bpartner, Geom g, const TabParam* __restrict__ prm,
    const double* __restrict__ pool, double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (i < n) {
    const int pid = id[i];
    const double4 pi = pos[i];
    const TabParam t = prm[0];
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    const int b0 = bstart[pid], b1 = bstart[pid + 1];
    for (int b = b0; b < b1; ++b) {
      const int q = bpartner[b];
      const int j = (q >= 0 && q < nid) ? slot_of_id[q] : -1;
      if (j < 0) continue;
      const double4 pj = pos[j];
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      if (dx < -g.halfL[0]) dx += g.L[0]; else if (dx > g.halfL[0]) dx -= g.L[0];
      if (dy < -g.halfL[1]) dy += g.L[1]; else if (dy > g.halfL[1]) dy -= g.L[1];
      if (dz < -g.halfL[2]) dz += g.L[2]; else if (dz > g.halfL[2]) dz -= g.L[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      tab_pair<ENERGY>(t, pool, dx, dy, dz, r2, r2, Ax, Ay, Az, acc);
    }
    if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// ------------------------------------------------------------------ radial distribution
//
// Pair-distance histogram of RadialDistrF::computeArray
// (src/analysis/RadialDistrF.cpp:47-125): every pair i < j, minimum image by a
// single +-L correction per component, bin = (int)(|d| / dr), kept when
// bin < nbins.  FP64 without contraction and with IEEE sqrt and division, so
// every pair falls into the reference's bin.  A workgroup takes 256 particles i
// and a chunk of particles j staged through LDS; counts gather in an LDS
// histogram (LDS_HIST) or directly in global memory.
template <bool LDS_HIST>
__global__ __launch_bounds__(256) void k_rdf(int n, const double4* __restrict__ pos, Geom g, double dr, int nbins,
                                             int jchunk, unsigned long long* __restrict__ counts) {
#pragma clang fp contract(off)
  extern __shared__ unsigned int hist[];
  __shared__ double4 pj_s[256];
  const int i0 = blockIdx.x * 256;
  const int j0 = blockIdx.y * jchunk;
  if (j0 + jchunk - 1 <= i0) return;               // no pair j > i in this block (uniform)
  if (LDS_HIST) {
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0u;
  }
  const int i = i0 + threadIdx.x;
  const double4 pi = i < n ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const int jend = min(j0 + jchunk, n);
  for (int jt = j0; jt < jend; jt += 256) {
    __syncthreads();
    if (jt + (int)threadIdx.x < jend) pj_s[threadIdx.x] = pos[jt + threadIdx.x];
    __syncthreads();
    const int m = min(256, jend - jt);
    if (i < n) {
      for (int k = 0; k < m; ++k) {
        const int j = jt + k;
        if (j <= i) continue;
        const double4 pj = pj_s[k];
        double d[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (d[c] < -g.halfL[c]) d[c] += g.L[c];
          if (d[c] > g.halfL[c]) d[c] -= g.L[c];
        }
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double q = r / dr;
        if (q < (double)nbins) {
          const int bin = (int)q;
          if (LDS_HIST) atomicAdd(&hist[bin], 1u);
          else atomicAdd(&counts[bin], 1ull);
        }
      }
    }
  }
  if (LDS_HIST) {
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
      if (hist[b]) atomicAdd(&counts[b], (unsigned long long)hist[b]);
  }
}

// ------------------------------------------------------------------ FENE bonds

// Bonds are stored per particle id in both directions (bstart/bpartner), so
// each particle sums F(x_i - x_j) over its partners and no atomics are needed;
// this equals the reference's p1 += F, p2 -= F (src/interaction/FixedPairListInteractionTemplate.hpp:112-200).
// Distance: minimum image as OrthorhombicBC::getMinimumImageVectorBox
// (src/bc/OrthorhombicBC.cpp:86-105) in FP64; in mixed mode from the fixed-point
// mirror, which is the same minimum image for |d| < L/2.
// STEP: the bonds are the last force of the step and the kernel finishes it
// per particle, as the step variant of k_lj_tiled_mixed does (new positions to
// sa.pos_out / sa.posq_out, since other threads still read the current ones).
template <bool MIXED, bool ENERGY, bool ACCUM, bool STEP>
__global__ __launch_bounds__(256) void k_fene(
    int n, const double4* __restrict__ pos, const uint4* __restrict__ posq,
    const int* __restrict__ id, const int* __restrict__ slot_of_id, int nid,
    const int* __restrict__ bstart, const int* __restrict__ bpartner, Geom g, FENEParam fp,
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    double* __restrict__ partial, StepArgs sa) {
#pragma clang fp contract(off)
  static_assert(!(STEP && ENERGY), "the step variant computes forces only");
  if (sa.stop && *reinterpret_cast<const volatile int*>(sa.stop)) return;   // a window's step after the stop
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double stepsq = 0.0;
  double acc[kNPart];
#pragma unroll
  for (int q = 0; q < kNPart; ++q) acc[q] = 0.0;
  if (i < n) {
    const int pid = id[i];
    // the step's own loads go out before the bond chain (id -> bonds -> partner slot)
    double4 step_v, step_p;
    if (STEP) { step_v = sa.vel[i]; step_p = sa.pos[i]; }
    double Ax = 0.0, Ay = 0.0, Az = 0.0;
    const int b0 = bstart[pid], b1 = bstart[pid + 1];
    auto bond = [&](int j) {
#pragma clang fp contract(off)
      double dx, dy, dz;
      if (MIXED) {
        const uint4 qi = posq[i], qj = posq[j];
        dx = (double)((float)(int)(qi.x - qj.x) * g.qs[0]);
        dy = (double)((float)(int)(qi.y - qj.y) * g.qs[1]);
        dz = (double)((float)(int)(qi.z - qj.z) * g.qs[2]);
      } else {
        const double4 pi = pos[i], pj = pos[j];
        dx = pi.x - pj.x; dy = pi.y - pj.y; dz = pi.z - pj.z;
        if (dx < -g.halfL[0]) dx += g.L[0]; else if (dx > g.halfL[0]) dx -= g.L[0];
        if (dy < -g.halfL[1]) dy += g.L[1]; else if (dy > g.halfL[1]) dy -= g.L[1];
        if (dz < -g.halfL[2]) dz += g.L[2]; else if (dz > g.halfL[2]) dz -= g.L[2];
      }
      double ff, e = 0.0;
      double r2;
      if (MIXED) {
        const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
        const float fr2 = fdx * fdx + fdy * fdy + fdz * fdz;
        r2 = fr2;
        float f32;
        if (fp.r0 > 2.220446049250313e-16) {
          const float r = __builtin_sqrtf(fr2), rr = r - (float)fp.r0;
          f32 = -(float)fp.K * rr / (r * (1.0f - rr * rr / (float)fp.rMaxSqr));
        } else {
          f32 = -(float)fp.K / (1.0f - fr2 / (float)fp.rMaxSqr);
        }
        ff = f32;
      } else {
        r2 = dx * dx + dy * dy + dz * dz;
        if (fp.r0 > 2.220446049250313e-16) {
          const double r = sqrt(r2);
          ff = -fp.K * (r - fp.r0) / (r * (1 - ((r - fp.r0) * (r - fp.r0) / fp.rMaxSqr)));
        } else {
          ff = -fp.K / (1.0 - r2 / fp.rMaxSqr);
        }
      }
      if (r2 > fp.cutoffSqr) ff = 0.0;
      Ax += dx * ff; Ay += dy * ff; Az += dz * ff;
      if (ENERGY) {
        if (r2 <= fp.cutoffSqr) {
          const double t = (sqrt(r2) - fp.r0) / fp.rMax;
          e = -0.5 * fp.rMaxSqr * fp.K * log(1 - t * t) - fp.shift;
        }
        acc[0] += e;
        acc[1] += r2 * ff;
        acc[2] += dx * dx * ff; acc[3] += dy * dy * ff; acc[4] += dz * dz * ff;
        acc[5] += dx * dy * ff; acc[6] += dx * dz * ff; acc[7] += dy * dz * ff;
      }
    };
    // bonds two at a time: both partners' lookups are in flight together (a chain
    // bead has two bonds), and the bonds are summed in their stored order
    for (int b = b0; b < b1; b += 2) {
      const int q0 = bpartner[b];
      const int q1 = b + 1 < b1 ? bpartner[b + 1] : -1;
      const int j0 = (q0 >= 0 && q0 < nid) ? slot_of_id[q0] : -1;
      const int j1 = (q1 >= 0 && q1 < nid) ? slot_of_id[q1] : -1;
      if (j0 >= 0) bond(j0);
      if (j1 >= 0) bond(j1);
    }
    if (STEP) {
      double f0 = Ax, f1 = Ay, f2 = Az;
      if (ACCUM) { f0 = fx[i] + Ax; f1 = fy[i] + Ay; f2 = fz[i] + Az; }
      step_thermostat(i, step_v, f0, f1, f2, sa);
      stepsq = step_kick_drift(i, step_v, step_p, f0, f1, f2, sa, g);
    } else if (!ENERGY) {
      if (ACCUM) { fx[i] += Ax; fy[i] += Ay; fz[i] += Az; }
      else { fx[i] = Ax; fy[i] = Ay; fz[i] = Az; }
    }
  }
  if (STEP) step_block_max(stepsq, sa);
  if (ENERGY) {
#pragma unroll
    for (int q = 0; q < kNPart; ++q) acc[q] *= 0.5;
    block_reduce_store<kNPart>(acc, partial);
  }
}

// Properties of m particles given by id (field numbering of espp_amd.h: 0 pos,
// 1 vel, 2 force, 3 image, 4 type, 5 mass); out holds 3 doubles per particle.
// slots_out[k] is the particle's slot, or -1 for an id that does not exist.
__global__ void k_gather_ids(int m, const int* __restrict__ ids, const int* __restrict__ slot_of_id, int nid,
                             int field, const double4* __restrict__ pos, const double4* __restrict__ vel,
                             const double* __restrict__ fx, const double* __restrict__ fy,
                             const double* __restrict__ fz, const int4* __restrict__ img,
                             double* __restrict__ out, int* __restrict__ slots_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int pid = ids[k];
  const int s = (pid >= 0 && pid < nid) ? slot_of_id[pid] : -1;
  slots_out[k] = s;
  if (s < 0) return;
  double a = 0.0, b = 0.0, c = 0.0;
  if (field == 0 || field == 4) {
    const double4 p = pos[s];
    if (field == 0) { a = p.x; b = p.y; c = p.z; } else { a = p.w; }
  } else if (field == 1 || field == 5) {
    const double4 v = vel[s];
    if (field == 1) { a = v.x; b = v.y; c = v.z; } else { a = v.w; }
  } else if (field == 2) {
    a = fx[s]; b = fy[s]; c = fz[s];
  } else {
    const int4 q = img[s];
    a = (double)q.x; b = (double)q.y; c = (double)q.z;
  }
  out[3 * k] = a;
  out[3 * k + 1] = b;
  out[3 * k + 2] = c;
}

// Set a property of m particles at the given slots (pos / vel: 3 values per
// particle; type / mass: 1).
__global__ void k_scatter_slots(int m, const int* __restrict__ slots, int field, const double* __restrict__ val,
                                double4* __restrict__ pos, double4* __restrict__ vel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int s = slots[k];
  if (field == 0) {
    double4 p = pos[s];
    p.x = val[3 * k]; p.y = val[3 * k + 1]; p.z = val[3 * k + 2];
    pos[s] = p;
  } else if (field == 4) {
    pos[s].w = val[k];
  } else if (field == 1) {
    double4 v = vel[s];
    v.x = val[3 * k]; v.y = val[3 * k + 1]; v.z = val[3 * k + 2];
    vel[s] = v;
  } else {
    vel[s].w = val[k];
  }
}

__global__ void k_zero3(int n, double* __restrict__ a, double* __restrict__ b, double* __restrict__ c) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { a[i] = 0.0; b[i] = 0.0; c[i] = 0.0; }
}

// ------------------------------------------------------------------ integrator

// Langevin friction + uniform noise (LangevinThermostat::frictionThermo,
// src/integrator/LangevinThermostat.cpp:141-177) followed, when do_int2 is set,
// by VelocityVerlet::integrate2 (src/integrator/VelocityVerlet.cpp:289-305).
__global__ __launch_bounds__(256) void k_thermo_integrate2(
    int n, const int* __restrict__ id, double4* __restrict__ vel, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz, int lang, double pref1, double pref2,
    unsigned long long seed, unsigned long long call, const unsigned char* __restrict__ excl,
    int nid, int do_int2, double half_dt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 v = vel[i];
  double f0 = fx[i], f1 = fy[i], f2 = fz[i];
  if (lang) {
    const int pid = id[i];
    const bool skip = excl && pid < nid && excl[pid];
    if (!skip) {
      uint32_t o[4];
      philox((uint32_t)pid, (uint32_t)call, (uint32_t)(call >> 32), 0x4C414E47u,
             (uint32_t)seed, (uint32_t)(seed >> 32), o);
      const double u0 = ((double)o[0] + 0.5) * 2.3283064365386963e-10;
      const double u1 = ((double)o[1] + 0.5) * 2.3283064365386963e-10;
      const double u2 = ((double)o[2] + 0.5) * 2.3283064365386963e-10;
      const double m = v.w, massf = sqrt(m);
      f0 += (pref1 * v.x) * m + (pref2 * (u0 - 0.5)) * massf;
      f1 += (pref1 * v.y) * m + (pref2 * (u1 - 0.5)) * massf;
      f2 += (pref1 * v.z) * m + (pref2 * (u2 - 0.5)) * massf;
      fx[i] = f0; fy[i] = f1; fz[i] = f2;
    }
  }
  if (do_int2) {
    const double dtfm = half_dt / v.w;
    v.x += dtfm * f0;
    v.y += dtfm * f1;
    v.z += dtfm * f2;
    vel[i] = v;
  }
}

// ------------------------------------------------------------------ fused step
//
// Thermostat and integrate2 of step s-1 followed by integrate1 of step s on the
// same forces (VelocityVerlet::run, src/integrator/VelocityVerlet.cpp:121-165:
// aftCalcF -> integrate2 -> befIntP -> integrate1), one pass over the particles,
// in place.  Writes the block's maximum |dx|^2; k_finish_max reduces them.
__global__ __launch_bounds__(256) void k_step_fused(int n, double* __restrict__ fx, double* __restrict__ fy,
                                                    double* __restrict__ fz, StepArgs a, Geom g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    const double4 v = a.vel[i];
    double f0 = fx[i], f1 = fy[i], f2 = fz[i];
    if (step_thermostat(i, v, f0, f1, f2, a)) { fx[i] = f0; fy[i] = f1; fz[i] = f2; }
    sq = step_kick_drift(i, v, a.pos[i], f0, f1, f2, a, g);
  }
  step_block_max(sq, a);
}

// host_out (pinned host memory, may be null) receives the same bits, so the host
// reads the step's displacement without a copy or a stream synchronisation.
__global__ __launch_bounds__(1024) void k_finish_max(int nblk, const double* __restrict__ blockmax, Ctl* ctl,
                                                     unsigned long long* host_out, SpecState* sp) {
  __shared__ double red[16];
  // a window's step that did not run (stop raised before it) has no displacement
  if (sp && *reinterpret_cast<volatile int*>(&sp->stop)) return;
  double m = 0.0;
  // sixteen independent loads in flight per thread, not one round trip per load
  for (int b0 = 0; b0 < nblk; b0 += 16 * (int)blockDim.x) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int b = b0 + u * (int)blockDim.x + (int)threadIdx.x;
      v[u] = b < nblk ? blockmax[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) m = fmax(m, v[u]);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, red[w]);
    const unsigned long long bits = (unsigned long long)__double_as_longlong(t);
    ctl->maxsq_bits = bits;
    if (host_out) *reinterpret_cast<volatile unsigned long long*>(host_out) = bits;
    if (sp) {
      // the skin check of VelocityVerlet::run (src/integrator/VelocityVerlet.cpp:121-136)
      sp->dmax += sqrt(t);
      sp->done += 1;
      if (!(sp->dmax <= sp->skin_half)) sp->stop = 1;
    }
  }
}

__global__ void k_spec_begin(SpecState* sp, double dmax, double skin_half) {
  sp->dmax = dmax;
  sp->skin_half = skin_half;
  sp->stop = 0;
  sp->done = 0;
}

// Kinetic sums per block: [sum m v^2, count, m vx vx, m vy vy, m vz vz, m vx vy,
// m vx vz, m vy vz] over particles whose type bit is set in type_mask
// (src/analysis/Temperature.hpp:100-122, src/analysis/PressureTensor.hpp:51-82).
__global__ __launch_bounds__(256) void k_kinetic(int n, const double4* __restrict__ vel,
                                                 const double4* __restrict__ pos,
                                                 unsigned long long type_mask,
                                                 double* __restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const double4 v = vel[i];
    const int t = (int)pos[i].w;
    if (t < 64 && ((type_mask >> t) & 1ull)) {
      acc[0] = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
      acc[1] = 1.0;
      acc[2] = v.w * v.x * v.x; acc[3] = v.w * v.y * v.y; acc[4] = v.w * v.z * v.z;
      acc[5] = v.w * v.x * v.y; acc[6] = v.w * v.x * v.z; acc[7] = v.w * v.y * v.z;
    }
  }
  block_reduce_store<8>(acc, partial);
}

// Sum nblk rows of nv partials in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_reduce_partials(int nblk, int nv,
                                                         const double* __restrict__ partial,
                                                         double* __restrict__ out) {
  __shared__ double red[4];
  for (int q = 0; q < nv; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(size_t)b * nv + q];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
      out[q] = t;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launchers

static inline int nblocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_fold_cell(hipStream_t s, int n, double4* pos, int4* img, int* cell_of, int* rank,
                      int* cell_count, const Geom& g, Ctl* ctl, int do_fold) {
  if (n > 0) hipLaunchKernelGGL(k_fold_cell, dim3(nblocks(n, 256)), dim3(256), 0, s, n, pos, img, cell_of, rank, cell_count, g, ctl, do_fold);
}


void launch_scan_multi(hipStream_t s, int ncell, const int* count, int* start, int* blocksum, int* blockoff) {
  const int nb = (ncell + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(1024), 0, s, ncell, count, blocksum);
  hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, nb, blocksum, blockoff);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(1024), 0, s, ncell, count, blockoff, start);
}

int scan_blocks(int ncell) { return (ncell + kScanChunk - 1) / kScanChunk; }

void launch_step_fused(hipStream_t s, int n, double* fx, double* fy, double* fz, const StepArgs& a, const Geom& g,
                       Ctl* ctl, unsigned long long* host_out) {
  if (n <= 0) return;
  const int nb = nblocks(n, 256);
  hipLaunchKernelGGL(k_step_fused, dim3(nb), dim3(256), 0, s, n, fx, fy, fz, a, g);
  hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, nb, a.blockmax, ctl, host_out, nullptr);
}


void launch_scatter(hipStream_t s, int n, const int* cell_of, const int* rank, const int* start, int* order) {
  if (n > 0) hipLaunchKernelGGL(k_scatter, dim3(nblocks(n, 256)), dim3(256), 0, s, n, cell_of, rank, start, order);
}

void launch_sort_cells(hipStream_t s, int ncell, const int* start, int* order, const double4* pos, const Ctl* ctl) {
  hipLaunchKernelGGL(k_sort_cells, dim3(nblocks(ncell, kSortCellsPerBlock)), dim3(kSortLanes * kSortCellsPerBlock), 0,
                     s, ncell, start, order, pos);
  hipLaunchKernelGGL(k_sort_cells_big, dim3(nblocks(ncell, 256)), dim3(256), 0, s, ncell, start, order, pos, ctl);
}

void launch_permute(hipStream_t s, int n, const int* order, const int* cell_of, const double4* pos,
                    const double4* vel, const double* fx, const double* fy, const double* fz,
                    const int* id, const int4* img, double4* pos2, double4* vel2, double* fx2,
                    double* fy2, double* fz2, int* id2, int4* img2, int* cell_of_slot,
                    int* slot_of_id, uint4* posq, float4* posf, uint2* posh, const Geom& g) {
  if (n > 0)
    hipLaunchKernelGGL(k_permute, dim3(nblocks(n, 256)), dim3(256), 0, s, n, order, cell_of, pos, vel, fx, fy, fz, id, img,
                       pos2, vel2, fx2, fy2, fz2, id2, img2, cell_of_slot, slot_of_id, posq, posf,
                       g.q16inv > 0.0 ? posh : nullptr, g);
}

void launch_spec_begin(hipStream_t s, SpecState* sp, double dmax, double skin_half) {
  hipLaunchKernelGGL(k_spec_begin, dim3(1), dim3(1), 0, s, sp, dmax, skin_half);
}

void launch_xbins(hipStream_t s, int n, int ncell, const int* cstart, const int* cell_of_slot, const float4* posf,
                  const Geom& g, int* xbin) {
  const int m = std::max(n, ncell);
  if (m <= 0) return;
  hipLaunchKernelGGL(k_xbins, dim3(nblocks(m, 256)), dim3(256), 0, s, n, ncell, cstart, cell_of_slot, posf, g, xbin);
}

void launch_posf(hipStream_t s, int n, const double4* pos, float4* posf, uint2* posh, double q16inv) {
  if (n > 0)
    hipLaunchKernelGGL(k_posf, dim3(nblocks(n, 256)), dim3(256), 0, s, n, pos, posf, q16inv > 0.0 ? posh : nullptr,
                       q16inv);
}

void launch_fill_int(hipStream_t s, int n, int* a, int v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(nblocks(n, 256)), dim3(256), 0, s, n, a, v);
}

void launch_reset_ctl(hipStream_t s, Ctl* ctl, int fields) {
  hipLaunchKernelGGL(k_reset_ctl, dim3(1), dim3(64), 0, s, ctl, fields);
}

void launch_nbr_build(hipStream_t s, int n, const double4* pos, const int* id, const int* cell_of_slot,
                      const int* cstart, const Geom& g, double cutsq, const int* ex_start,
                      const int* ex_ids, int* nbr, int* cnt, int cap, Ctl* ctl) {
  if (n <= 0) return;
  if (ex_start)
    hipLaunchKernelGGL(k_nbr_build<true>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, pos, id, cell_of_slot, cstart, g,
                       cutsq, ex_start, ex_ids, nbr, cnt, cap, ctl);
  else
    hipLaunchKernelGGL(k_nbr_build<false>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, pos, id, cell_of_slot, cstart, g,
                       cutsq, ex_start, ex_ids, nbr, cnt, cap, ctl);
}

int lj_blocks(int n) { return nblocks(n, 256); }

void launch_lj(hipStream_t s, bool mixed, bool onetype, bool energy, bool accum, int n,
               const uint4* posq, const double4* pos, const int* nbr, const int* cnt, int cap,
               const Geom& g, const LJParam* prm, int ntypes, double* fx, double* fy, double* fz,
               double* partial, const int* stop) {
  if (n <= 0) return;
  dim3 grid(nblocks(n, 256)), block(256);
#define LJM(OT, EN, AC) hipLaunchKernelGGL((k_lj_mixed<OT, EN, AC>), grid, block, 0, s, n, posq, pos, nbr, cnt, cap, g, prm, ntypes, fx, fy, fz, partial, stop)
#define LJD(OT, EN, AC) hipLaunchKernelGGL((k_lj_fp64<OT, EN, AC>), grid, block, 0, s, n, pos, nbr, cnt, cap, g, prm, ntypes, fx, fy, fz, partial)
  if (mixed) {
    if (energy) { if (onetype) LJM(true, true, false); else LJM(false, true, false); }
    else if (onetype) { if (accum) LJM(true, false, true); else LJM(true, false, false); }
    else { if (accum) LJM(false, false, true); else LJM(false, false, false); }
  } else {
    if (energy) { if (onetype) LJD(true, true, false); else LJD(false, true, false); }
    else if (onetype) { if (accum) LJD(true, false, true); else LJD(true, false, false); }
    else { if (accum) LJD(false, false, true); else LJD(false, false, false); }
  }
#undef LJM
#undef LJM1
#undef LJD
}


void launch_tile_plan(hipStream_t s, int n, int T, int ntiles, const int* cell_of_slot, const int* cstart,
                      const Geom& g, int budget, TileRun* runs, int* nruns, Ctl* ctl) {
  if (ntiles > 0)
    hipLaunchKernelGGL(k_tile_plan, dim3(nblocks(ntiles, kPlanTilesPerBlock)), dim3(64 * kPlanTilesPerBlock), 0, s, n,
                       T, ntiles, cell_of_slot, cstart, g, budget, runs, nruns, ctl);
}



void launch_nbr_build_tiled(hipStream_t s, int n, int T, const float4* posf, const double4* pos,
                            const int* id, const int* cell_of_slot, const int* cstart, const Geom& g, double cutsq,
                            float cut_lo, float cut_hi, const TileRun* runs, const int* nruns, const int* ex_start,
                            const int* ex_ids, uint4* nbr, int* cnt, int capg, int* lane_slot,
                            float clip_margin, float row_margin, const int* xbin, const uint2* posh,
                            double q16inv, int q16_tlo, int q16_w, Ctl* ctl) {
  if (n <= 0) return;
  const int ntiles = nblocks(n, T);
  dim3 grid(ntiles), block(T);
#define NB(EX, Q) hipLaunchKernelGGL((k_nbr_build_tiled<EX, Q>), grid, block, 0, s, n, ntiles, posf, pos, id, cell_of_slot, cstart, g, cutsq, cut_lo, cut_hi, runs, nruns, ex_start, ex_ids, nbr, cnt, capg, lane_slot, clip_margin, row_margin, xbin, posh, q16inv, q16_tlo, q16_w, ctl)
  const bool q16 = q16inv > 0.0 && posh;
  if (ex_start) { if (q16) NB(true, true); else NB(true, false); }
  else { if (q16) NB(false, true); else NB(false, false); }
#undef NB
}

void launch_lj_tiled(hipStream_t s, bool mixed, bool onetype, bool energy, bool accum, int n, int T, int lds_bytes,
                     const uint4* posq, const double4* pos, const uint4* nbr, const int* cnt, int capg,
                     const Geom& g, const TileRun* runs, const int* nruns, const int* lane_slot, const LJParam* prm,
                     int ntypes, double* fx, double* fy, double* fz, double* partial, const PruneOut* prune,
                     const int* stop) {
  if (n <= 0) return;
  const int ntiles = nblocks(n, T);
  dim3 grid(ntiles), block(T);
  const bool cubic = cubic_box(g);
  StepArgs sa{};
  sa.stop = stop;
  // only mixed force walks prune
  if (!mixed || energy) prune = nullptr;
  const PruneOut po = prune ? *prune : PruneOut{};
#define LJM1(OT, EN, AC, CU, PR) hipLaunchKernelGGL((k_lj_tiled_mixed<OT, EN, AC, CU, false, PR>), grid, block, lds_bytes, s, n, ntiles, posq, pos, nbr, cnt, capg, g, runs, nruns, lane_slot, prm, ntypes, fx, fy, fz, partial, sa, po)
#define LJM(OT, EN, AC) do { if (cubic) { if (!EN && prune) LJM1(OT, false, AC, true, true); else LJM1(OT, EN, AC, true, false); } \
                             else { if (!EN && prune) LJM1(OT, false, AC, false, true); else LJM1(OT, EN, AC, false, false); } } while (0)
#define LJD(OT, EN, AC) hipLaunchKernelGGL((k_lj_tiled_fp64<OT, EN, AC>), grid, block, lds_bytes, s, n, ntiles, pos, nbr, cnt, capg, g, runs, nruns, lane_slot, prm, ntypes, fx, fy, fz, partial)
  if (mixed) {
    if (energy) { if (onetype) LJM(true, true, false); else LJM(false, true, false); }
    else if (onetype) { if (accum) LJM(true, false, true); else LJM(true, false, false); }
    else { if (accum) LJM(false, false, true); else LJM(false, false, false); }
  } else if (onetype && 2 * T <= 256) {
    // two lanes per particle, 24-byte slots: the staged region needs 3/4 of lds_bytes
    const int lb = lds_bytes / 32 * 24;
    dim3 block2(2 * T);
#define LJP(EN, AC) hipLaunchKernelGGL((k_lj_tiled_fp64_pair<EN, AC>), grid, block2, lb, s, n, ntiles, T, pos, nbr, cnt, capg, g, runs, nruns, lane_slot, prm, fx, fy, fz, partial)
    if (energy) LJP(true, false);
    else if (accum) LJP(false, true);
    else LJP(false, false);
#undef LJP
  } else {
    if (energy) { if (onetype) LJD(true, true, false); else LJD(false, true, false); }
    else if (onetype) { if (accum) LJD(true, false, true); else LJD(true, false, false); }
    else { if (accum) LJD(false, false, true); else LJD(false, false, false); }
  }
#undef LJM
#undef LJM1
#undef LJD
}

int lj_tiled_blocks(int n, int T) { return nblocks(n, T); }

void launch_lj_tiled_step(hipStream_t s, bool onetype, bool accum, int n, int T, int lds_bytes, const uint4* posq,
                          const double4* pos, const uint4* nbr, const int* cnt, int capg, const Geom& g,
                          const TileRun* runs, const int* nruns, const int* lane_slot, const LJParam* prm,
                          int ntypes, double* fx, double* fy, double* fz, const StepArgs& sa, Ctl* ctl,
                          unsigned long long* host_out, const PruneOut* prune, SpecState* spec) {
  if (n <= 0) return;
  const int ntiles = nblocks(n, T);
  dim3 grid(ntiles), block(T);
  const bool cubic = cubic_box(g);
  const PruneOut po = prune ? *prune : PruneOut{};
#define LJS1(OT, AC, CU, PR) hipLaunchKernelGGL((k_lj_tiled_mixed<OT, false, AC, CU, true, PR>), grid, block, lds_bytes, s, n, ntiles, posq, pos, nbr, cnt, capg, g, runs, nruns, lane_slot, prm, ntypes, fx, fy, fz, nullptr, sa, po)
#define LJS(OT, AC) do { if (cubic) { if (prune) LJS1(OT, AC, true, true); else LJS1(OT, AC, true, false); } \
                         else { if (prune) LJS1(OT, AC, false, true); else LJS1(OT, AC, false, false); } } while (0)
  if (onetype) { if (accum) LJS(true, true); else LJS(true, false); }
  else { if (accum) LJS(false, true); else LJS(false, false); }
#undef LJS
#undef LJS1
  hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, ntiles, sa.blockmax, ctl, host_out, spec);
}

void launch_fene(hipStream_t s, bool mixed, bool energy, bool accum, int n, const double4* pos,
                 const uint4* posq, const int* id, const int* slot_of_id, int nid, const int* bstart,
                 const int* bpartner, const Geom& g, const FENEParam& fp, double* fx, double* fy,
                 double* fz, double* partial, const int* stop) {
  if (n <= 0) return;
  dim3 grid(nblocks(n, 256)), block(256);
  StepArgs sa{};
  sa.stop = stop;
#define FN(MX, EN, AC) hipLaunchKernelGGL((k_fene<MX, EN, AC, false>), grid, block, 0, s, n, pos, posq, id, slot_of_id, nid, bstart, bpartner, g, fp, fx, fy, fz, partial, sa)
  if (mixed) {
    if (energy) FN(true, true, false); else if (accum) FN(true, false, true); else FN(true, false, false);
  } else {
    if (energy) FN(false, true, false); else if (accum) FN(false, false, true); else FN(false, false, false);
  }
#undef FN
}

void launch_fene_step(hipStream_t s, bool mixed, bool accum, int n, const double4* pos, const uint4* posq,
                      const int* id, const int* slot_of_id, int nid, const int* bstart, const int* bpartner,
                      const Geom& g, const FENEParam& fp, double* fx, double* fy, double* fz, const StepArgs& sa,
                      Ctl* ctl, unsigned long long* host_out, SpecState* spec) {
  if (n <= 0) return;
  const int nb = nblocks(n, 256);
#define FN(MX, AC) hipLaunchKernelGGL((k_fene<MX, false, AC, true>), dim3(nb), dim3(256), 0, s, n, pos, posq, id, slot_of_id, nid, bstart, bpartner, g, fp, fx, fy, fz, nullptr, sa)
  if (mixed) { if (accum) FN(true, true); else FN(true, false); }
  else { if (accum) FN(false, true); else FN(false, false); }
#undef FN
  hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, nb, sa.blockmax, ctl, host_out, spec);
}

void launch_tab_tiled(hipStream_t s, bool mixed, bool energy, bool accum, int n, int T, int lds_bytes,
                      const uint4* posq, const double4* pos, const uint4* nbr, const int* cnt, int capg,
                      const Geom& g, const TileRun* runs, const int* nruns, const int* lane_slot, const TabParam* prm,
                      const double* pool, int ntypes, double* fx, double* fy, double* fz, double* partial) {
  if (n <= 0) return;
  const int ntiles = nblocks(n, T);
  dim3 grid(ntiles), block(T);
#define TT(MX, EN, AC) hipLaunchKernelGGL((k_tab_tiled<MX, EN, AC>), grid, block, lds_bytes, s, n, ntiles, posq, pos, nbr, cnt, capg, g, runs, nruns, lane_slot, prm, pool, ntypes, fx, fy, fz, partial)
  if (mixed) {
    if (energy) TT(true, true, false); else if (accum) TT(true, false, true); else TT(true, false, false);
  } else {
    if (energy) TT(false, true, false); else if (accum) TT(false, false, true); else TT(false, false, false);
  }
#undef TT
}

void launch_tab(hipStream_t s, bool energy, bool accum, int n, const double4* pos, const int* nbr, const int* cnt,
                int cap, const Geom& g, const TabParam* prm, const double* pool, int ntypes, double* fx, double* fy,
                double* fz, double* partial) {
  if (n <= 0) return;
  dim3 grid(nblocks(n, 256)), block(256);
#define TB(EN, AC) hipLaunchKernelGGL((k_tab<EN, AC>), grid, block, 0, s, n, pos, nbr, cnt, cap, g, prm, pool, ntypes, fx, fy, fz, partial)
  if (energy) TB(true, false); else if (accum) TB(false, true); else TB(false, false);
#undef TB
}

void launch_bond_tab(hipStream_t s, bool energy, bool accum, int n, const double4* pos, const int* id,
                     const int* slot_of_id, int nid, const int* bstart, const int* bpartner, const Geom& g,
                     const TabParam* prm, const double* pool, double* fx, double* fy, double* fz, double* partial) {
  if (n <= 0) return;
  dim3 grid(nblocks(n, 256)), block(256);
#define BT(EN, AC) hipLaunchKernelGGL((k_bond_tab<EN, AC>), grid, block, 0, s, n, pos, id, slot_of_id, nid, bstart, bpartner, g, prm, pool, fx, fy, fz, partial)
  if (energy) BT(true, false); else if (accum) BT(false, true); else BT(false, false);
#undef BT
}

void launch_rdf(hipStream_t s, int n, const double4* pos, const Geom& g, double dr, int nbins,
                unsigned long long* counts) {
  if (n <= 1 || nbins <= 0) return;
  // j chunks keep gridDim.y within its limit and give every workgroup a few thousand pairs per i
  int jchunk = 4096;
  while ((n + jchunk - 1) / jchunk > 32768) jchunk *= 2;
  dim3 grid(nblocks(n, 256), (n + jchunk - 1) / jchunk), block(256);
  if (nbins <= kRdfLdsBins)
    hipLaunchKernelGGL(k_rdf<true>, grid, block, (size_t)nbins * sizeof(unsigned int), s, n, pos, g, dr, nbins, jchunk,
                       counts);
  else
    hipLaunchKernelGGL(k_rdf<false>, grid, block, 0, s, n, pos, g, dr, nbins, jchunk, counts);
}

void launch_gather_ids(hipStream_t s, int m, const int* ids, const int* slot_of_id, int nid, int field,
                       const double4* pos, const double4* vel, const double* fx, const double* fy, const double* fz,
                       const int4* img, double* out, int* slots_out) {
  if (m > 0)
    hipLaunchKernelGGL(k_gather_ids, dim3(nblocks(m, 256)), dim3(256), 0, s, m, ids, slot_of_id, nid, field, pos, vel,
                       fx, fy, fz, img, out, slots_out);
}

void launch_scatter_slots(hipStream_t s, int m, const int* slots, int field, const double* val, double4* pos,
                          double4* vel) {
  if (m > 0)
    hipLaunchKernelGGL(k_scatter_slots, dim3(nblocks(m, 256)), dim3(256), 0, s, m, slots, field, val, pos, vel);
}

void launch_zero3(hipStream_t s, int n, double* a, double* b, double* c) {
  if (n > 0) hipLaunchKernelGGL(k_zero3, dim3(nblocks(n, 256)), dim3(256), 0, s, n, a, b, c);
}

void launch_thermo_integrate2(hipStream_t s, int n, const int* id, double4* vel, double* fx,
                              double* fy, double* fz, int lang, double pref1, double pref2,
                              unsigned long long seed, unsigned long long call,
                              const unsigned char* excl, int nid, int do_int2, double half_dt) {
  if (n > 0)
    hipLaunchKernelGGL(k_thermo_integrate2, dim3(nblocks(n, 256)), dim3(256), 0, s, n, id, vel, fx, fy, fz, lang, pref1,
                       pref2, seed, call, excl, nid, do_int2, half_dt);
}


int kinetic_blocks(int n) { return nblocks(n, 256); }

void launch_kinetic(hipStream_t s, int n, const double4* vel, const double4* pos,
                    unsigned long long type_mask, double* partial) {
  if (n > 0) hipLaunchKernelGGL(k_kinetic, dim3(nblocks(n, 256)), dim3(256), 0, s, n, vel, pos, type_mask, partial);
}

void launch_reduce_partials(hipStream_t s, int nblk, int nv, const double* partial, double* out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, nblk, nv, partial, out);
}

}  // namespace espp
