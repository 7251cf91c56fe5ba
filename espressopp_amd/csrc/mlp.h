// Learned coarse-grained pair force (mlp.hip): model constants, parameter
// layout and launchers used by the host runtime.
#pragma once
#include "espp_dev.h"

namespace espp {
namespace mlp {

constexpr int kRbf = 32;                 // Gaussian radial basis functions
constexpr int kIn = kRbf + 2;            // + two pair-density inputs
constexpr int kHid = 64;                 // width of the three hidden layers
// FP32 parameter vector: W1 [kIn][kHid], b1, W2 [kHid][kHid], b2, W3, b3, w4 [kHid], b4
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kHid;
constexpr int kParams = kOffB4 + 1;      // 10625
constexpr int kPackElems = 5 * 64 * 64;  // bfloat16 operand matrices

// Cutoff, RBF grid and the density features' radii and normalisation.
struct Spec {
  float rc;                // pair cutoff of the model
  float mu0, dmu, inv2s2;  // RBF centres mu0 + k dmu, exp(-(r - mu)^2 inv2s2)
  float R[2];              // radii of the two local densities
  float mean[2], inv_scale[2];
};

size_t lds_bytes(bool train);
void launch_pack(hipStream_t s, const float* theta, unsigned short* wpack);
void launch_density(hipStream_t s, int n, const double4* pos, const int* nbr, const int* cnt, int cap, const Geom& g,
                    const Spec& sp, float2* rho);
void launch_forces(hipStream_t s, int n, const double4* pos, const int* nbr, const int* cnt, int cap, const Geom& g,
                   const Spec& sp, const float2* rho, const unsigned short* wpack, const float* theta, bool accum,
                   double* fx, double* fy, double* fz);
// fpred: 3 n floats of scratch; fref: 3 doubles per particle id; adds to grad and *loss
void launch_train(hipStream_t s, int n, const double4* pos, const int* id, const int* nbr, const int* cnt, int cap,
                  const Geom& g, const Spec& sp, const float2* rho, const unsigned short* wpack, const float* theta,
                  const double* fx, const double* fy, const double* fz, float* fpred, const double* fref,
                  double scale, float* grad, double* loss);
void launch_adam(hipStream_t s, float* theta, float* grad, float* m1, float* m2, float lr, float beta1, float beta2,
                 float eps, float c1, float c2);
void launch_zero(hipStream_t s, float* a, int np);
void launch_zero_d(hipStream_t s, double* a);

}  // namespace mlp
}  // namespace espp
