// Error state, twiddle bank and the element-wise kernels of step3_propagation.py.
#include <cmath>
#include <mutex>
#include <vector>

#include "mm_common.h"

namespace mm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return MM_OK;
}

// One bank per device, built once and kept until the process ends.
static std::mutex g_bank_mu;
static std::vector<float2*> g_bank;

int twiddle_bank(const float2** out) {
  int dev = 0;
  MM_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_bank_mu);
  if ((int)g_bank.size() <= dev) g_bank.resize(dev + 1, nullptr);
  if (!g_bank[dev]) {
    std::vector<float2> host(kTwiddleBank, make_float2(0.f, 0.f));
    auto fill = [&](int L) {
      for (int t = 0; t < L; ++t) {
        double a = -2.0 * M_PI * (double)t / (double)L;
        host[tw_offset(L) + t] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
    };
    for (int L = 2; L <= kMaxFFT; L *= 2) fill(L);
    for (int L = 12; L <= kMaxFFT; L *= 2) fill(L);
    for (int e = 0; e < 16; ++e)
      for (int k = 0; k < 16; ++k) {
        const double a = -2.0 * M_PI * (double)(e * k) / 256.0;
        host[kTwGrid256 + 16 * e + k] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
    float2* d = nullptr;
    hipError_t e = hipMalloc(&d, host.size() * sizeof(float2));
    if (e != hipSuccess) return fail(MM_ERR_OOM, std::string("twiddle bank: ") + hipGetErrorString(e));
    e = hipMemcpy(d, host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(MM_ERR_HIP, std::string("twiddle bank copy: ") + hipGetErrorString(e));
    }
    g_bank[dev] = d;
  }
  *out = g_bank[dev];
  return MM_OK;
}

// psi = exp(-phase * inv_beta_ratio) * (cos phase + i sin phase), phase = p * inv_norm
// (step3_propagation.py:94-103).  HIP's precise expf / sincosf.
__global__ void psi_from_proj_kernel(const float* __restrict__ p, float2* __restrict__ psi, long long count,
                                     float inv_norm, float inv_beta_ratio) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    psi[i] = psi_of(p[i], inv_norm, inv_beta_ratio);
  }
}

// out = re^2 + im^2 (step3_propagation.py:105-110).
__global__ void abs2_kernel(const float2* __restrict__ z, float* __restrict__ out, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    out[i] = abs2_of(z[i]);
  }
}

static unsigned grid_for(long long count, int block) {
  long long g = (count + block - 1) / block;
  if (g > 65535LL * 16) g = 65535LL * 16;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace mm

extern "C" {

const char* mm_last_error(void) { return mm::g_last_error.c_str(); }

int mm_version(void) { return 100; }

int mm_psi_from_proj(const float* proj_dev, void* psi_dev, long long count, float inv_norm, float inv_beta_ratio,
                     void* stream) {
  MM_ARG_CHECK(count >= 0, "mm_psi_from_proj: count < 0");
  MM_ARG_CHECK(count == 0 || (proj_dev && psi_dev), "mm_psi_from_proj: null pointer");
  if (count == 0) return MM_OK;
  hipLaunchKernelGGL(mm::psi_from_proj_kernel, dim3(mm::grid_for(count, 256)), dim3(256), 0, (hipStream_t)stream,
                     proj_dev, (float2*)psi_dev, count, inv_norm, inv_beta_ratio);
  return mm::check_launch("psi_from_proj");
}

int mm_abs2(const void* z_dev, float* out_dev, long long count, void* stream) {
  MM_ARG_CHECK(count >= 0, "mm_abs2: count < 0");
  MM_ARG_CHECK(count == 0 || (z_dev && out_dev), "mm_abs2: null pointer");
  if (count == 0) return MM_OK;
  hipLaunchKernelGGL(mm::abs2_kernel, dim3(mm::grid_for(count, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)z_dev, out_dev, count);
  return mm::check_launch("abs2");
}

}  // extern "C"
