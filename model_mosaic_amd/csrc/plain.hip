// Reference-structured comparator for the benchmark's plain one-slice path:
// the polar gather and its adjoint as one thread per polar sample that reads
// (or atomically adds) its 81 taps directly in global memory, as the
// reference's kernels do (processing/kernels.py:274-349 gather_rfft_full and
// :16-93 gather dir=1; launched one slice per call from processing/tomo.py:138
// and :172).  The FFTs, pads and sign passes around them are left to the
// caller (plain_path.py runs them through torch.fft, i.e. rocFFT), so a call
// sequence has the reference's structure: one plan-less slice at a time, a
// full 2n x 2n complex grid for the adjoint, float atomics.  Not on the
// product path (TomoReal / TomoLargeReal use radon.hip).
#include <cmath>

#include "mm_common.h"

namespace mm {
namespace plain {

constexpr int M = 4;

__device__ __forceinline__ int wrap2n(int v, int twon) {
  const int r = v % twon;
  return r < 0 ? r + twon : r;
}

// F: rfft2 of one padded slice, (2n, n+1) complex, raw fftfreq order.
// g[th, r] = (1/n) sum_taps (pi/mu) exp(-pi^2/mu |l/2n - (x0, y0)|^2) (-1)^(k0+k1) F~[k1, k0]
// with F~ the conjugate-reflected value for k0 > n.
__global__ void gather_kernel(const float2* __restrict__ F, float2* __restrict__ g, const float* __restrict__ theta,
                              int n, int ntheta, float mu) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, th = blockIdx.y;
  if (r >= n || th >= ntheta) return;
  float s, c;
  sincosf(theta[th], &s, &c);
  const float x0 = ((float)r - 0.5f * n) / (float)n * c, y0 = -((float)r - 0.5f * n) / (float)n * s;
  const int twon = 2 * n;
  const float inv2n = 1.0f / (float)twon, c0 = (float)M_PI / mu, c1 = -(float)(M_PI * M_PI) / mu;
  const int bx = (int)rintf(twon * x0) - M, by = (int)rintf(twon * y0) - M;
  float2 acc = make_float2(0.f, 0.f);
  for (int i1 = 0; i1 <= 2 * M; ++i1) {
    const int l1 = by + i1, ky = wrap2n(l1, twon);
    const float w1 = (float)l1 * inv2n - y0;
    const float ey = c0 * expf(c1 * w1 * w1);
    for (int i0 = 0; i0 <= 2 * M; ++i0) {
      const int l0 = bx + i0, k0 = wrap2n(l0, twon);
      const float w0 = (float)l0 * inv2n - x0;
      float w = ey * expf(c1 * w0 * w0);
      if ((k0 + ky) & 1) w = -w;
      const bool refl = k0 > n;
      float2 v = F[(long long)(refl ? (ky == 0 ? 0 : twon - ky) : ky) * (n + 1) + (refl ? twon - k0 : k0)];
      if (refl) v.y = -v.y;
      acc.x += w * v.x;
      acc.y += w * v.y;
    }
  }
  g[(long long)th * n + r] = make_float2(acc.x / n, acc.y / n);
}

// G (2n, 2n) complex, centred index (n + l + 2n) % 2n, accumulated with float atomics.
__global__ void scatter_kernel(const float2* __restrict__ g, float2* __restrict__ G, const float* __restrict__ theta,
                               int n, int ntheta, float mu) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, th = blockIdx.y;
  if (r >= n || th >= ntheta) return;
  float s, c;
  sincosf(theta[th], &s, &c);
  const float x0 = ((float)r - 0.5f * n) / (float)n * c, y0 = -((float)r - 0.5f * n) / (float)n * s;
  const int twon = 2 * n;
  const float inv2n = 1.0f / (float)twon, c0 = (float)M_PI / mu, c1 = -(float)(M_PI * M_PI) / mu;
  const int bx = (int)rintf(twon * x0) - M, by = (int)rintf(twon * y0) - M;
  const float2 v = g[(long long)th * n + r];
  for (int i1 = 0; i1 <= 2 * M; ++i1) {
    const int l1 = by + i1, fy = wrap2n(n + l1, twon);
    const float w1 = (float)l1 * inv2n - y0;
    const float ey = c0 * expf(c1 * w1 * w1);
    for (int i0 = 0; i0 <= 2 * M; ++i0) {
      const int l0 = bx + i0, fx = wrap2n(n + l0, twon);
      const float w0 = (float)l0 * inv2n - x0;
      const float w = ey * expf(c1 * w0 * w0);
      float2* dst = G + (long long)fy * twon + fx;
      atomicAdd(&dst->x, w * v.x);
      atomicAdd(&dst->y, w * v.y);
    }
  }
}

}  // namespace plain
}  // namespace mm

extern "C" {

int mm_plain_gather(const void* F, void* g, const float* theta_dev, int n, int ntheta, void* stream) {
  MM_ARG_CHECK(F && g && theta_dev, "mm_plain_gather: null pointer");
  MM_ARG_CHECK(n >= 16 && n % 2 == 0 && ntheta >= 1 && ntheta < 65536, "mm_plain_gather: bad sizes");
  const float mu = (float)(-std::log(1e-3) / (2.0 * n * n));
  hipLaunchKernelGGL(mm::plain::gather_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)ntheta), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)F, (float2*)g, theta_dev, n, ntheta, mu);
  return mm::check_launch("plain gather");
}

int mm_plain_scatter(const void* g, void* G, const float* theta_dev, int n, int ntheta, void* stream) {
  MM_ARG_CHECK(g && G && theta_dev, "mm_plain_scatter: null pointer");
  MM_ARG_CHECK(n >= 16 && n % 2 == 0 && ntheta >= 1 && ntheta < 65536, "mm_plain_scatter: bad sizes");
  const float mu = (float)(-std::log(1e-3) / (2.0 * n * n));
  hipLaunchKernelGGL(mm::plain::scatter_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)ntheta), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)g, (float2*)G, theta_dev, n, ntheta, mu);
  return mm::check_launch("plain scatter");
}

}  // extern "C"
