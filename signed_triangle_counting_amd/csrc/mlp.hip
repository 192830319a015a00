// Training of the dispatcher MLP on the device.
//
// No reference counterpart (the reference dispatches with a fixed 4x skew
// rule, Signed_TC/GPU Implementations/GlobalMEM/signed_tc.hxx:48,65).  The
// model maps a bucket's features to logits over {merge, hash, bitmap, pull};
// the counter picks the argmax, so inference needs no exponential.  Training
// minimises, over samples i, w_i * sum_s softmax(z_i)_s * r_is where r_is is
// strategy s's cost normalised to [0, 1] within the sample and w_i the time at
// stake (max - min cost).  One wave64 runs full-batch Adam: each lane
// accumulates gradients of a fixed sample subset in its LDS row, then lane p
// sums column p over lanes in a fixed order and updates parameter p, so the
// result is deterministic.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace stc {

namespace {

constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kMlpHidden * kMlpIn;
constexpr int kW2 = kB1 + kMlpHidden;
constexpr int kB2 = kW2 + kMlpOut * kMlpHidden;
constexpr int kParams = kB2 + kMlpOut;  // 179
constexpr int kLanes = 64;
constexpr int kOwned = (kParams + kLanes - 1) / kLanes;

__global__ __launch_bounds__(kLanes) void k_train_mlp(const float *__restrict__ x,
                                                      const float *__restrict__ r,
                                                      const float *__restrict__ w, int nsamp,
                                                      int iters, float lr,
                                                      float *__restrict__ params) {
  extern __shared__ float part[];        // kLanes rows of kParams gradients
  __shared__ float P[kParams];
  const int tid = threadIdx.x;
  for (int p = tid; p < kParams; p += kLanes) P[p] = params[p];
  float m1[kOwned], m2[kOwned];
#pragma unroll
  for (int k = 0; k < kOwned; k++) m1[k] = m2[k] = 0.f;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  float b1t = 1.f, b2t = 1.f;
  __syncthreads();
  float *g = part + tid * kParams;
  for (int it = 0; it < iters; it++) {
    for (int p = 0; p < kParams; p++) g[p] = 0.f;
    for (int i = tid; i < nsamp; i += kLanes) {
      float xi[kMlpIn];
#pragma unroll
      for (int k = 0; k < kMlpIn; k++) xi[k] = x[i * kMlpIn + k];
      float h[kMlpHidden];
#pragma unroll
      for (int j = 0; j < kMlpHidden; j++) {
        float a = P[kB1 + j];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) a += P[kW1 + j * kMlpIn + k] * xi[k];
        h[j] = a > 0.f ? a : 0.f;
      }
      float z[kMlpOut], zmax = -3.0e38f;
#pragma unroll
      for (int s = 0; s < kMlpOut; s++) {
        float a = P[kB2 + s];
#pragma unroll
        for (int j = 0; j < kMlpHidden; j++) a += P[kW2 + s * kMlpHidden + j] * h[j];
        z[s] = a;
        zmax = a > zmax ? a : zmax;
      }
      float pr[kMlpOut], den = 0.f;
#pragma unroll
      for (int s = 0; s < kMlpOut; s++) {
        pr[s] = expf(z[s] - zmax);
        den += pr[s];
      }
      float expect = 0.f;
#pragma unroll
      for (int s = 0; s < kMlpOut; s++) {
        pr[s] /= den;
        expect += pr[s] * r[i * kMlpOut + s];
      }
      float dz[kMlpOut];
#pragma unroll
      for (int s = 0; s < kMlpOut; s++) dz[s] = w[i] * pr[s] * (r[i * kMlpOut + s] - expect);
#pragma unroll
      for (int s = 0; s < kMlpOut; s++) {
        g[kB2 + s] += dz[s];
#pragma unroll
        for (int j = 0; j < kMlpHidden; j++) g[kW2 + s * kMlpHidden + j] += dz[s] * h[j];
      }
#pragma unroll
      for (int j = 0; j < kMlpHidden; j++) {
        if (h[j] <= 0.f) continue;
        float dh = 0.f;
#pragma unroll
        for (int s = 0; s < kMlpOut; s++) dh += P[kW2 + s * kMlpHidden + j] * dz[s];
        g[kB1 + j] += dh;
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) g[kW1 + j * kMlpIn + k] += dh * xi[k];
      }
    }
    __syncthreads();
    b1t *= beta1;
    b2t *= beta2;
#pragma unroll
    for (int k = 0; k < kOwned; k++) {
      const int p = tid + k * kLanes;
      if (p >= kParams) break;
      float grad = 0.f;
      for (int t = 0; t < kLanes; t++) grad += part[t * kParams + p];
      m1[k] = beta1 * m1[k] + (1.f - beta1) * grad;
      m2[k] = beta2 * m2[k] + (1.f - beta2) * grad * grad;
      const float mh = m1[k] / (1.f - b1t), vh = m2[k] / (1.f - b2t);
      P[p] -= lr * mh / (sqrtf(vh) + eps);
    }
    __syncthreads();
  }
  for (int p = tid; p < kParams; p += kLanes) params[p] = P[p];
}

}  // namespace

hipError_t train_mlp_device(const float *x, const float *r, const float *w, int nsamp,
                            int iters, float lr, float *params_io, hipStream_t st) {
  float *d = nullptr;
  const size_t nx = (size_t)nsamp * kMlpIn, nr = (size_t)nsamp * kMlpOut;
  hipError_t e = hipMalloc(&d, (nx + nr + nsamp + kParams) * sizeof(float));
  if (e != hipSuccess) return e;
  float *dx = d, *dr = d + nx, *dw = dr + nr, *dp = dw + nsamp;
  e = hipMemcpyAsync(dx, x, nx * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dr, r, nr * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dw, w, (size_t)nsamp * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dp, params_io, kParams * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_train_mlp, dim3(1), dim3(kLanes), kLanes * kParams * sizeof(float), st,
                       dx, dr, dw, nsamp, iters, lr, dp);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(params_io, dp, kParams * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  return e;
}

int mlp_param_count() { return kParams; }

void mlp_unpack(const float *params, MlpModel &m) {
  for (int j = 0; j < kMlpHidden; j++) {
    for (int k = 0; k < kMlpIn; k++) m.w1[j][k] = params[kW1 + j * kMlpIn + k];
    m.b1[j] = params[kB1 + j];
  }
  for (int s = 0; s < kMlpOut; s++) {
    for (int j = 0; j < kMlpHidden; j++) m.w2[s][j] = params[kW2 + s * kMlpHidden + j];
    m.b2[s] = params[kB2 + s];
  }
}

}  // namespace stc
