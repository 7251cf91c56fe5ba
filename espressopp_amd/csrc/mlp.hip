// Learned coarse-grained pair force: a density-dependent pair MLP evaluated and
// trained with BF16 matrix-core (MFMA) kernels and FP32 accumulation on gfx950.
// The model has no counterpart in the reference; it is the learned interaction
// of the coarse-grained workload, used beside interaction.Tabulated
// (src/interaction/Tabulated.hpp:43-115), which stays the plain fallback.
//
// For a listed pair (i, j), d = r_i - r_j, r = |d| < rc:
//   x[0:32] Gaussian RBFs exp(-(r - k rc/31)^2 / (2 (rc/31)^2)),
//   x[32+a] ((rho_i^a + rho_j^a)/2 - mean_a) / scale_a with the smooth local
//           densities rho_i^a = sum_j (1 + cos(pi r_ij / R_a)) / 2 (r_ij < R_a),
//   h1..h3 = silu(. W + b) (64 wide), s = h3 . w4 + b4, g = fc(r) s with
//   fc(r) = (1 + cos(pi r / rc)) / 2, and F_i = sum_j g_ij d_ij / r_ij.
// Training minimises sum_i |F_i - F_i^ref|^2 / (3N) (force matching).
//
// Kernel shape.  A workgroup of four wavefronts owns one 64-particle group of
// the full (untiled) Verlet list and walks it by neighbour index: step k is a
// tile of 64 pairs (owner t, its k-th neighbour).  Per tile the 64 x 64 GEMMs
// of the three hidden layers (and, when training, the two back-propagation
// GEMMs and the three weight-gradient GEMMs) run as v_mfma_f32_16x16x32_bf16
// on operands staged in LDS; wavefront w computes the 16-column block w of
// every product.  Weight gradients stay in each wavefront's registers across
// all tiles and are added to the global FP32 gradient once per workgroup.  A
// particle's force is summed by the thread that owns it, so no atomics touch
// forces and the result does not depend on scheduling.
//
// Rounding: inputs, weight matrices, hidden activations, silu'(z) and the
// back-propagated dz are rounded to bfloat16 (nearest, ties to even) where they
// become MFMA operands; sums, biases, w4, b4, forces and the loss are FP32.
#include <hip/hip_runtime.h>

#include "espp_dev.h"
#include "espp_devfn.h"
#include "mlp.h"

namespace espp {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kS = 72;                 // LDS row stride (bf16): 64 + 8, rows stay 16-byte aligned
constexpr int kMat = 64 * kS;          // one 64 x 64 operand in LDS
constexpr int kThreads = 256;

// FP32 -> bfloat16 bits, rounded to nearest even (one v_cvt_pk_bf16_f32 on gfx950;
// FP32 denormals are kept, so it rounds every finite value as the oracle does)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float bf16_val(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }
__device__ __forceinline__ float bf16_round(float f) { return bf16_val(bf16_bits(f)); }

// sigmoid with the hardware reciprocal (about 1 ulp): the results are rounded to
// bfloat16 right after, so an IEEE division would buy nothing
__device__ __forceinline__ float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float sg = sigmoid_f(z);
  return sg * (1.0f + z * (1.0f - sg));
}
__device__ __forceinline__ float cosine_cut(float r, float R) {
  return r < R ? 0.5f * (1.0f + __cosf(3.14159265358979f * r / R)) : 0.0f;
}

// 16 output columns (block nb) of C[64 x 64] = A[64 x 32 ksteps] * B, A row-major
// [m][k] and Bt = B transposed, [n][k], both in LDS with stride kS.  Lane l gets
// acc[mb][v] = C[16 mb + 4 (l / 16) + v][16 nb + l % 16].  Lane l supplies row l % 16
// of A and column l % 16 of B with the same eight k of each 32-wide step, so the
// products pair up whatever order the unit reduces k in.
__device__ __forceinline__ void gemm_cols(const unsigned short* A, const unsigned short* Bt, int nb, int ksteps,
                                          int mblocks, f32x4 acc[4]) {
  const int l = threadIdx.x & 63;
  const int row = l & 15, kq = (l >> 4) * 8;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < ksteps; ++ks) {
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bt + (nb * 16 + row) * kS + ks * 32 + kq);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      if (mb >= mblocks) break;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + (mb * 16 + row) * kS + ks * 32 + kq);
      acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[mb], 0, 0, 0);
    }
  }
}

// Same, accumulating into acc (weight gradients summed over tiles).
__device__ __forceinline__ void gemm_cols_acc(const unsigned short* A, const unsigned short* Bt, int nb, int ksteps,
                                              int mblocks, f32x4 acc[4]) {
  const int l = threadIdx.x & 63;
  const int row = l & 15, kq = (l >> 4) * 8;
  for (int ks = 0; ks < ksteps; ++ks) {
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bt + (nb * 16 + row) * kS + ks * 32 + kq);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      if (mb >= mblocks) break;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + (mb * 16 + row) * kS + ks * 32 + kq);
      acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[mb], 0, 0, 0);
    }
  }
}

// Weights in LDS: three forward operands W_l^T ([out][in]) and, for training,
// W2 and W3 as stored ([in][out]) for the back-propagation products.
__device__ void load_weights(const unsigned short* __restrict__ wpack, unsigned short* lds, int nmat) {
  for (int e = threadIdx.x; e < nmat * 64 * 8; e += blockDim.x) {
    const int m = e / 512, rem = e % 512, row = rem / 8, c8 = rem % 8;
    const uint4 v = reinterpret_cast<const uint4*>(wpack)[(size_t)m * 512 + row * 8 + c8];
    *reinterpret_cast<uint4*>(lds + m * kMat + row * kS + c8 * 8) = v;
  }
}

struct PairCtx {
  float dx, dy, dz, r;
  int valid;
};

// The neighbour data of one tile (read by the thread with q == 0): its slot, its
// position and its densities.  Entries past the owner's list read the owner itself.
struct Nbr {
  int j;
  double4 p;
  float2 rho;
};

__device__ __forceinline__ int nbr_slot(int i, bool live, int k, int cnt_i, const int* __restrict__ nbr, int cap) {
  return (live && k < cnt_i) ? nbr[((size_t)(i >> 6) * cap + k) * 64 + (i & 63)] : i;
}

// Geometry of the pair (owner i at pi, neighbour nb); a pair past the list, at
// r >= rc or at r = 0 is invalid and gets r = rc so that its features stay finite.
__device__ __forceinline__ PairCtx pair_geom(const double4& pi, int i, const Nbr& nb, const Geom& g, float rc) {
  PairCtx c{0.f, 0.f, 0.f, rc, 0};
  if (nb.j != i) {
    double dx = pi.x - nb.p.x, dy = pi.y - nb.p.y, dz = pi.z - nb.p.z;
    dx -= g.L[0] * __builtin_rint(dx * g.invL[0]);
    dy -= g.L[1] * __builtin_rint(dy * g.invL[1]);
    dz -= g.L[2] * __builtin_rint(dz * g.invL[2]);
    const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
    if (r < rc && r > 0.f) {
      c.dx = (float)dx; c.dy = (float)dy; c.dz = (float)dz; c.r = r; c.valid = 1;
    }
  }
  return c;
}

// Sixteen inputs [16 q, 16 q + 16) of one pair into X[p][.] (and X^T when T is set).
template <bool T>
__device__ __forceinline__ void write_features(int p, int q, float r, float u0, float u1, const Spec& sp,
                                               unsigned short* X, unsigned short* Xt) {
  unsigned short v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int k = 16 * q + e;
    float x = 0.f;
    if (k < kRbf) {
      const float t = r - sp.mu0 - sp.dmu * (float)k;
      x = __expf(-t * t * sp.inv2s2);
    } else if (k == kRbf) {
      x = u0;
    } else if (k == kRbf + 1) {
      x = u1;
    }
    v[e] = bf16_bits(x);
  }
  uint4* dst = reinterpret_cast<uint4*>(X + p * kS + 16 * q);
  dst[0] = make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
  dst[1] = make_uint4(v[8] | (v[9] << 16), v[10] | (v[11] << 16), v[12] | (v[13] << 16), v[14] | (v[15] << 16));
  if (T) {
#pragma unroll
    for (int e = 0; e < 16; ++e) Xt[(16 * q + e) * kS + p] = v[e];
  }
}

// Local densities rho^a of every particle over its list (one thread per particle).
__global__ __launch_bounds__(256) void k_mlp_density(int n, const double4* __restrict__ pos,
                                                     const int* __restrict__ nbr, const int* __restrict__ cnt,
                                                     int cap, Geom g, Spec sp, float2* __restrict__ rho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int c = min(cnt[i], cap);
  float a = 0.f, b = 0.f;
  for (int k = 0; k < c; ++k) {
    const int j = nbr[((size_t)(i >> 6) * cap + k) * 64 + (i & 63)];
    const double4 pj = pos[j];
    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    dx -= g.L[0] * __builtin_rint(dx * g.invL[0]);
    dy -= g.L[1] * __builtin_rint(dy * g.invL[1]);
    dz -= g.L[2] * __builtin_rint(dz * g.invL[2]);
    const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
    a += cosine_cut(r, sp.R[0]);
    b += cosine_cut(r, sp.R[1]);
  }
  rho[i] = make_float2(a, b);
}

// Forward pass; TRAIN adds the backward pass for the force-matching loss.
// Forward: the owner's force sum, added to (accum) or stored in fx, fy, fz.
// TRAIN: reads the forward's forces fpred (FP32, per slot) and the reference
// forces fref (per particle id), adds the weight gradient of
// weight * sum_i |F_i - F_i^ref|^2 / (3 N) to grad and the loss to *loss.
template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_mlp(int n, const double4* __restrict__ pos,
                                                  const int* __restrict__ id, const int* __restrict__ nbr,
                                                  const int* __restrict__ cnt, int cap, Geom g, Spec sp,
                                                  const float2* __restrict__ rho,
                                                  const unsigned short* __restrict__ wpack,
                                                  const float* __restrict__ theta, int accum,
                                                  double* __restrict__ fx, double* __restrict__ fy,
                                                  double* __restrict__ fz, const float* __restrict__ fpred,
                                                  const double* __restrict__ fref, double scale,
                                                  float* __restrict__ grad, double* __restrict__ loss) {
  // LDS: weights (3 forward + 2 backward), then the tile buffers
  extern __shared__ __align__(16) unsigned short lds[];
  unsigned short* Wt1 = lds;
  unsigned short* Wt2 = lds + kMat;
  unsigned short* Wt3 = lds + 2 * kMat;
  unsigned short* W2 = lds + 3 * kMat;          // TRAIN only
  unsigned short* W3 = lds + 4 * kMat;          // TRAIN only
  unsigned short* buf = lds + (TRAIN ? 5 : 3) * kMat;
  unsigned short* B0 = buf;                      // X, then H2          [p][.]
  unsigned short* B1 = buf + kMat;               // H1, then H3         [p][.]
  unsigned short* Xt = buf + 2 * kMat;           // TRAIN: X^T, H1^T, H2^T, S1, S2, S3, dZ, dZ^T
  unsigned short* H1t = buf + 3 * kMat;
  unsigned short* H2t = buf + 4 * kMat;
  unsigned short* S1 = buf + 5 * kMat;
  unsigned short* S2 = buf + 6 * kMat;
  unsigned short* S3 = buf + 7 * kMat;
  unsigned short* DZ = buf + 8 * kMat;
  unsigned short* DZt = buf + 9 * kMat;
  __shared__ float sb[4][64];                    // b1, b2, b3, w4
  __shared__ float sds[64];                      // d loss / d s of the tile's pairs
  __shared__ int skmax;

  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  const int p = t >> 2, q = t & 3;               // pair (= owner) of this thread, quarter of its inputs
  const int i = blockIdx.x * 64 + p;             // owner slot
  const bool live = i < n;
  load_weights(wpack, lds, TRAIN ? 5 : 3);
  {
    const float* b1 = theta + kOffB1;
    const float* b2 = theta + kOffB2;
    const float* b3 = theta + kOffB3;
    const float* w4 = theta + kOffW4;
    if (t < 64) { sb[0][t] = b1[t]; sb[1][t] = b2[t]; sb[2][t] = b3[t]; sb[3][t] = w4[t]; }
  }
  if (t == 0) skmax = 0;
  __syncthreads();
  const float b4 = theta[kOffB4];
  const int cnt_i = live ? min(cnt[i], cap) : 0;
  if (q == 0) atomicMax(&skmax, cnt_i);
  const float2 rho_i = live ? rho[i] : make_float2(0.f, 0.f);
  // owner's force residual (TRAIN)
  float rfx = 0.f, rfy = 0.f, rfz = 0.f;
  if (TRAIN && live && q == 0) {
    const int pid = id[i];
    rfx = fpred[3 * (size_t)i] - (float)fref[3 * (size_t)pid];
    rfy = fpred[3 * (size_t)i + 1] - (float)fref[3 * (size_t)pid + 1];
    rfz = fpred[3 * (size_t)i + 2] - (float)fref[3 * (size_t)pid + 2];
    atomicAdd(loss, scale * ((double)rfx * rfx + (double)rfy * rfy + (double)rfz * rfz));
  }
  __syncthreads();
  const int kmax = skmax;
  float Fx = 0.f, Fy = 0.f, Fz = 0.f;
  // gradient accumulators of this wavefront's output columns
  f32x4 gW1[4], gW2[4], gW3[4], tmp[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) { gW1[mb] = gW2[mb] = gW3[mb] = f32x4{0.f, 0.f, 0.f, 0.f}; }
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f;
  const int col = 16 * w + (l & 15);             // output column of this lane in the GEMMs
  const int rq = 4 * (l >> 4);                   // first of its four rows in each 16-row block

  // Neighbour data is software-pipelined: the slot two tiles ahead and the
  // position and densities one tile ahead are loaded while the current tile's
  // GEMMs run, so the dependent global loads do not stall each tile.
  const bool loader = live && q == 0;
  const int ic = live ? i : 0;
  const double4 pi = pos[ic];
  Nbr cur, nxt;
  cur.j = nbr_slot(ic, loader, 0, cnt_i, nbr, cap);
  int j2 = nbr_slot(ic, loader, 1, cnt_i, nbr, cap);
  cur.p = pos[cur.j];
  cur.rho = rho[cur.j];
  for (int k = 0; k < kmax; ++k) {
    nxt.j = j2;
    nxt.p = pos[j2];
    nxt.rho = rho[j2];
    j2 = nbr_slot(ic, loader, k + 2, cnt_i, nbr, cap);
    PairCtx pc = pair_geom(pi, ic, cur, g, sp.rc);
    if (!loader) pc.valid = 0;
    // the thread with q == 0 computed the geometry; share it with the pair's other three
    const int src = l & ~3;
    pc.dx = __shfl(pc.dx, src, 64); pc.dy = __shfl(pc.dy, src, 64); pc.dz = __shfl(pc.dz, src, 64);
    pc.r = __shfl(pc.r, src, 64); pc.valid = __shfl(pc.valid, src, 64);
    float2 rho_j;
    rho_j.x = __shfl(cur.rho.x, src, 64);
    rho_j.y = __shfl(cur.rho.y, src, 64);
    if (!pc.valid) rho_j = rho_i;
    cur = nxt;
    const float u0 = (0.5f * (rho_i.x + rho_j.x) - sp.mean[0]) * sp.inv_scale[0];
    const float u1 = (0.5f * (rho_i.y + rho_j.y) - sp.mean[1]) * sp.inv_scale[1];
    write_features<TRAIN>(p, q, pc.r, u0, u1, sp, B0, Xt);
    __syncthreads();
    // layer 1: B0 (X) -> B1 (H1)
    gemm_cols(B0, Wt1, w, 2, 4, tmp);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + rq + v;
        const float z = tmp[mb][v] + sb[0][col];
        const unsigned short h = bf16_bits(silu_f(z));
        B1[row * kS + col] = h;
        if (TRAIN) { H1t[col * kS + row] = h; S1[row * kS + col] = bf16_bits(dsilu_f(z)); }
      }
    __syncthreads();
    // layer 2: B1 (H1) -> B0 (H2)
    gemm_cols(B1, Wt2, w, 2, 4, tmp);
    __syncthreads();                              // every wavefront has read X before H2 overwrites it
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + rq + v;
        const float z = tmp[mb][v] + sb[1][col];
        const unsigned short h = bf16_bits(silu_f(z));
        B0[row * kS + col] = h;
        if (TRAIN) { H2t[col * kS + row] = h; S2[row * kS + col] = bf16_bits(dsilu_f(z)); }
      }
    __syncthreads();
    // layer 3: B0 (H2) -> B1 (H3)
    gemm_cols(B0, Wt3, w, 2, 4, tmp);
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + rq + v;
        const float z = tmp[mb][v] + sb[2][col];
        B1[row * kS + col] = bf16_bits(silu_f(z));
        if (TRAIN) S3[row * kS + col] = bf16_bits(dsilu_f(z));
      }
    __syncthreads();
    // output: s = h3 . w4 + b4 over the pair's four threads
    float part = 0.f;
#pragma unroll
    for (int e = 0; e < 16; ++e) part += bf16_val(B1[p * kS + 16 * q + e]) * sb[3][16 * q + e];
    part += __shfl_xor(part, 1, 64);
    part += __shfl_xor(part, 2, 64);
    const float s = part + b4;
    const float fc = pc.valid ? cosine_cut(pc.r, sp.rc) : 0.f;
    const float gpair = fc * s;
    const float inv_r = 1.0f / pc.r;
    if (!TRAIN) {
      Fx += gpair * pc.dx * inv_r;
      Fy += gpair * pc.dy * inv_r;
      Fz += gpair * pc.dz * inv_r;
      __syncthreads();                            // B0 / B1 are rewritten by the next tile
      continue;
    }
    // d loss / d s = 2 scale (F_i - F_i^ref) . d / r  fc(r)
    float rx = __shfl(rfx, src, 64), ry = __shfl(rfy, src, 64), rz = __shfl(rfz, src, 64);
    const float ds = pc.valid ? (float)(2.0 * scale) * (rx * pc.dx + ry * pc.dy + rz * pc.dz) * inv_r * fc : 0.f;
    if (q == 0) sds[p] = ds;
    __syncthreads();
    // output-layer gradients and dz3 = ds w4 silu'(z3)
    {
      const int o = t & 63, pq = t >> 6;
      float a = 0.f, b = 0.f;
      for (int e = 0; e < 16; ++e) {
        const int pp = 16 * pq + e;
        a += sds[pp] * bf16_val(B1[pp * kS + o]);
        b += sds[pp];
      }
      gw4 += a;
      if (o == 0) gb4 += b;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int o = 16 * q + e;
      const unsigned short dz = bf16_bits(ds * sb[3][o] * bf16_val(S3[p * kS + o]));
      DZ[p * kS + o] = dz;
      DZt[o * kS + p] = dz;
    }
    __syncthreads();
    // dW3 += H2^T dZ3 ; dH2 = dZ3 W3^T ; dz2 = dH2 silu'(z2)
    gemm_cols_acc(H2t, DZt, w, 2, 4, gW3);
    gemm_cols(DZ, W3, w, 2, 4, tmp);
    {
      const int o = t & 63, pq = t >> 6;
      float a = 0.f;
      for (int e = 0; e < 16; ++e) a += bf16_val(DZt[o * kS + 16 * pq + e]);
      gb3 += a;
    }
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + rq + v;
        const unsigned short dz = bf16_bits(tmp[mb][v] * bf16_val(S2[row * kS + col]));
        DZ[row * kS + col] = dz;
        DZt[col * kS + row] = dz;
      }
    __syncthreads();
    // dW2 += H1^T dZ2 ; dH1 = dZ2 W2^T ; dz1 = dH1 silu'(z1)
    gemm_cols_acc(H1t, DZt, w, 2, 4, gW2);
    gemm_cols(DZ, W2, w, 2, 4, tmp);
    {
      const int o = t & 63, pq = t >> 6;
      float a = 0.f;
      for (int e = 0; e < 16; ++e) a += bf16_val(DZt[o * kS + 16 * pq + e]);
      gb2 += a;
    }
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + rq + v;
        DZt[col * kS + row] = bf16_bits(tmp[mb][v] * bf16_val(S1[row * kS + col]));
      }
    __syncthreads();
    // dW1 += X^T dZ1 (input rows 0..47 hold the 34 inputs)
    gemm_cols_acc(Xt, DZt, w, 2, 3, gW1);
    {
      const int o = t & 63, pq = t >> 6;
      float a = 0.f;
      for (int e = 0; e < 16; ++e) a += bf16_val(DZt[o * kS + 16 * pq + e]);
      gb1 += a;
    }
    __syncthreads();
  }

  if (!TRAIN) {
    if (live && q == 0) {
      if (accum) { fx[i] += Fx; fy[i] += Fy; fz[i] += Fz; }
      else { fx[i] = Fx; fy[i] = Fy; fz[i] = Fz; }
    }
    return;
  }
  // gradient of this workgroup into the global FP32 gradient
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int in = 16 * mb + rq + v;
      if (in < kIn && mb < 3) atomicAdd(&grad[kOffW1 + in * kHid + col], gW1[mb][v]);
      atomicAdd(&grad[kOffW2 + in * kHid + col], gW2[mb][v]);
      atomicAdd(&grad[kOffW3 + in * kHid + col], gW3[mb][v]);
    }
  {
    const int o = t & 63;
    // the four quarter-sums of each column
    atomicAdd(&grad[kOffB1 + o], gb1);
    atomicAdd(&grad[kOffB2 + o], gb2);
    atomicAdd(&grad[kOffB3 + o], gb3);
    atomicAdd(&grad[kOffW4 + o], gw4);
    if (o == 0) atomicAdd(&grad[kOffB4], gb4);
  }
}

// Operand layouts of the weights from the FP32 parameters: W1^T (inputs padded
// to 64), W2^T, W3^T, W2, W3, each 64 x 64 bfloat16, row-major.
__global__ void k_mlp_pack(const float* __restrict__ theta, unsigned short* __restrict__ wpack) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 5 * 4096) return;
  const int m = e / 4096, r = (e / 64) % 64, c = e % 64;
  float v = 0.f;
  if (m == 0) v = c < kIn ? theta[kOffW1 + c * kHid + r] : 0.f;      // [out r][in c]
  else if (m == 1) v = theta[kOffW2 + c * kHid + r];
  else if (m == 2) v = theta[kOffW3 + c * kHid + r];
  else if (m == 3) v = theta[kOffW2 + r * kHid + c];                 // [in r][out c]
  else v = theta[kOffW3 + r * kHid + c];
  wpack[e] = bf16_bits(v);
}

// Forces of the forward pass as FP32 per slot (input of the training pass).
__global__ void k_mlp_f32(int n, const double* __restrict__ fx, const double* __restrict__ fy,
                          const double* __restrict__ fz, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[3 * (size_t)i] = (float)fx[i];
  out[3 * (size_t)i + 1] = (float)fy[i];
  out[3 * (size_t)i + 2] = (float)fz[i];
}

// Adam (Kingma & Ba 2015) on the FP32 parameters; clears the gradient.
__global__ void k_mlp_adam(int np, float* __restrict__ theta, float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float beta1, float beta2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  const float gr = grad[e];
  const float a = beta1 * m1[e] + (1.f - beta1) * gr;
  const float b = beta2 * m2[e] + (1.f - beta2) * gr * gr;
  m1[e] = a;
  m2[e] = b;
  theta[e] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
  grad[e] = 0.f;
}

__global__ void k_mlp_zero(int np, float* __restrict__ a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < np) a[e] = 0.f;
}

__global__ void k_mlp_zero_d(double* a) { *a = 0.0; }

// ------------------------------------------------------------------ launchers

static inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

size_t lds_bytes(bool train) { return (size_t)(train ? 15 : 5) * kMat * sizeof(unsigned short); }

void launch_pack(hipStream_t s, const float* theta, unsigned short* wpack) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(5 * 4096, 256)), dim3(256), 0, s, theta, wpack);
}

void launch_density(hipStream_t s, int n, const double4* pos, const int* nbr, const int* cnt, int cap, const Geom& g,
                    const Spec& sp, float2* rho) {
  if (n > 0) hipLaunchKernelGGL(k_mlp_density, dim3(nblk(n, 256)), dim3(256), 0, s, n, pos, nbr, cnt, cap, g, sp, rho);
}

void launch_forces(hipStream_t s, int n, const double4* pos, const int* nbr, const int* cnt, int cap, const Geom& g,
                   const Spec& sp, const float2* rho, const unsigned short* wpack, const float* theta, bool accum,
                   double* fx, double* fy, double* fz) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp<false>, dim3(nblk(n, 64)), dim3(kThreads), lds_bytes(false), s, n, pos, nullptr, nbr,
                     cnt, cap, g, sp, rho, wpack, theta, accum ? 1 : 0, fx, fy, fz, nullptr, nullptr, 0.0, nullptr,
                     nullptr);
}

void launch_train(hipStream_t s, int n, const double4* pos, const int* id, const int* nbr, const int* cnt, int cap,
                  const Geom& g, const Spec& sp, const float2* rho, const unsigned short* wpack, const float* theta,
                  const double* fx, const double* fy, const double* fz, float* fpred, const double* fref,
                  double scale, float* grad, double* loss) {
  if (n <= 0) return;
  static bool lds_set = false;
  if (!lds_set) {
    // the training kernel's LDS exceeds the 64 KiB a workgroup gets without asking
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes(true));
    lds_set = true;
  }
  hipLaunchKernelGGL(k_mlp_f32, dim3(nblk(n, 256)), dim3(256), 0, s, n, fx, fy, fz, fpred);
  hipLaunchKernelGGL(k_mlp<true>, dim3(nblk(n, 64)), dim3(kThreads), lds_bytes(true), s, n, pos, id, nbr, cnt, cap,
                     g, sp, rho, wpack, theta, 0, nullptr, nullptr, nullptr, fpred, fref, scale, grad, loss);
}

void launch_adam(hipStream_t s, float* theta, float* grad, float* m1, float* m2, float lr, float beta1, float beta2,
                 float eps, float c1, float c2) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(kParams, 256)), dim3(256), 0, s, kParams, theta, grad, m1, m2, lr, beta1,
                     beta2, eps, c1, c2);
}

void launch_zero(hipStream_t s, float* a, int np) {
  hipLaunchKernelGGL(k_mlp_zero, dim3(nblk(np, 256)), dim3(256), 0, s, np, a);
}

void launch_zero_d(hipStream_t s, double* a) { hipLaunchKernelGGL(k_mlp_zero_d, dim3(1), dim3(1), 0, s, a); }

}  // namespace mlp
}  // namespace espp
