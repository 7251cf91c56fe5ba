// Device helper functions shared by the kernel files: exact FP64 squares, the
// fixed-point position mirror, reciprocals, wavefront reductions and ranks, the
// Philox4x32-10 counter-based generator and the Lennard-Jones force factor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "espp_dev.h"

namespace espp {

// |d|^2 summed left to right without contraction (Real3D::sqr, src/Real3D.hpp)
__device__ __forceinline__ double sqr_nc(double x, double y, double z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}

__device__ __forceinline__ uint32_t fixq(double x, double qinv) {
  // position -> 32-bit fixed point with L <-> 2^32; the int64 detour makes the
  // mapping modular, so unfolded positions wrap onto the same periodic image.
  return (uint32_t)(unsigned long long)(long long)__builtin_rint(x * qinv);
}

// The list build's 16-bit copy of a position: each coordinate to round(x / delta)
// modulo 2^16 (q16inv = 1 / delta, delta = L / (2^16 m) for a whole m), x and y
// in the low and high half of the first word, z in the low half of the second
// (its high half zero).  16-bit differences of two copies are the coordinate
// differences of the periodic image that lies within 2^15 delta.
__device__ __forceinline__ uint2 q16_pack(const double4& p, double q16inv) {
  const uint32_t ux = (uint32_t)(unsigned long long)(long long)__builtin_rint(p.x * q16inv) & 0xFFFFu;
  const uint32_t uy = (uint32_t)(unsigned long long)(long long)__builtin_rint(p.y * q16inv) & 0xFFFFu;
  const uint32_t uz = (uint32_t)(unsigned long long)(long long)__builtin_rint(p.z * q16inv) & 0xFFFFu;
  return make_uint2(ux | (uy << 16), uz);
}

typedef short v2i16 __attribute__((ext_vector_type(2)));
typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));

// a.lo^2 + a.hi^2 + c (signed 16-bit halves), the accumulator in a scalar or a
// vector register: the three-operand v_dot2_i32_i16, which the compiler would
// replace by a copy of c and the accumulating two-operand form
__device__ __forceinline__ int sdot2_sacc(v2i16 a, int c) {
  int r;
  asm("v_dot2_i32_i16 %0, %1, %1, %2" : "=v"(r) : "v"(a), "s"(c));
  return r;
}
__device__ __forceinline__ int sdot2_acc(v2i16 a, int c) {
  int r;
  asm("v_dot2_i32_i16 %0, %1, %1, %2" : "=v"(r) : "v"(a), "v"(c));
  return r;
}

// Order-preserving integer image of a double (negative values reflected), a
// total order with NaN included, so ranks by comparison always form a permutation
__device__ __forceinline__ long long x_key(double x) {
  const long long b = __double_as_longlong(x);
  return b ^ ((b >> 63) & 0x7fffffffffffffffll);
}

// x sub-bin of a candidate position for the list builds' row clipping: cell
// units u = x / cs from the FP32 copy, k = floor((u - cx) kXBins) clamped to the
// cell (the same arithmetic as the builds' lookups, so both order alike)
__device__ __forceinline__ int xsub_bin(float x, float inv_cs, int cx) {
  const float u = x * inv_cs;
  return min(max((int)floorf((u - (float)cx) * (float)kXBins), 0), kXBins - 1);
}

__device__ __forceinline__ float rcp_nr(float x) {
  // v_rcp_f32 followed by one Newton step: ~0.5 ulp, three VALU ops
  float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return r;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}


// Rank of this lane among all particles counted into cell `key`: lanes of one
// wavefront holding the same key share one atomicAdd (a wavefront spans only a
// few cells because particles arrive in cell order) and are ranked in lane order.
__device__ __forceinline__ int wave_aggregated_rank(int key, bool live, int* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(live);
  int rank = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lk = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(live && key == lk);
    int base = 0;
    if (lane == leader) base = atomicAdd(&count[lk], __popcll(same));
    base = __shfl(base, leader, 64);
    if (live && key == lk) rank = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return rank;
}

// Philox4x32-10 (Salmon et al., SC'11; Random123 constants): ten rounds of the
// multiply-xor bijection on the counter (c0..c3) with a Weyl-sequence key.
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double lj_ff_seeded(double r2, double ff1, double ff2, double& inv) {
  double u = (double)__builtin_amdgcn_rcpf((float)r2);
  u = u * __builtin_fma(-r2, u, 2.0);          // one Newton step: ~1e-14 relative
  inv = u;
  const double r6 = u * u * u;
  return r6 * __builtin_fma(ff1, r6, -ff2) * u;
}

// lj_ff_seeded for a pair that may be out of range: such a pair seeds the
// reciprocal with rcp(+inf) = 0, after which u, r6 and the factor are exactly 0
// (r2 is finite), so no select on the FP64 result is needed.
__device__ __forceinline__ double lj_ff_masked(double r2, bool in, double ff1, double ff2, double& inv) {
  const float rf = in ? (float)r2 : __builtin_inff();
  double u = (double)__builtin_amdgcn_rcpf(rf);
  u = u * __builtin_fma(-r2, u, 2.0);
  inv = u;
  const double r6 = u * u * u;
  return r6 * __builtin_fma(ff1, r6, -ff2) * u;
}

}  // namespace espp
