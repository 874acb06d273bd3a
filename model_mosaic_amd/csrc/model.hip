// Differentiable two-physics forward model (delta, beta) -> |D(exp(i k R delta - k R beta))|^2
// with an L2 misfit and a hand-written backward (not in the reference; the
// reference pipeline runs R in step2_radon.py:191 and psi/D/|.|^2 in
// step3_propagation.py:220-225 with no adjoint).
//
// Forward:  p = R delta, b = R beta, psi = exp(i k p - k b), u = D psi,
//           L = sum (|u|^2 - d)^2   (FP64 block partials + one-block reduce).
// Backward: psi and u are recomputed (only the two float32 sinograms p, b and
//           one complex64 sinogram-sized buffer are held),
//           r = 4 (|u|^2 - d) u,  G = DT r,
//           dL/dp = -k Im(conj(G) psi),  dL/db = -k Re(conj(G) psi),
//           dL/ddelta = RT dL/dp,  dL/dbeta = RT dL/db.
#include <vector>

#include "mm_common.h"

namespace mm {
namespace model {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 4096;

__global__ void psi_kernel(const float* __restrict__ p, const float* __restrict__ b, float2* __restrict__ u,
                           long long count, float k) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
    float s, c;
    sincosf(k * p[i], &s, &c);
    const float at = expf(-k * b[i]);
    u[i] = make_float2(at * c, at * s);
  }
}

__global__ void loss_partial_kernel(const float2* __restrict__ u, const float* __restrict__ d, long long count,
                                    double* __restrict__ partial) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
    const float2 v = u[i];
    const double r = (double)(v.x * v.x + v.y * v.y) - (double)d[i];
    acc += r * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void reduce_kernel(const double* __restrict__ partial, int nparts, double* __restrict__ out) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void residual_kernel(float2* __restrict__ u, const float* __restrict__ d, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
    const float2 v = u[i];
    const float r = 4.0f * (v.x * v.x + v.y * v.y - d[i]);
    u[i] = make_float2(r * v.x, r * v.y);
  }
}

// gp = -k Im(conj(G) psi) * gscale (over p), gb = -k Re(conj(G) psi) * gscale (over b)
__global__ void chain_kernel(const float2* __restrict__ G, float* __restrict__ p, float* __restrict__ b,
                             long long count, float k, float gscale) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
    float s, c;
    sincosf(k * p[i], &s, &c);
    const float at = expf(-k * b[i]);
    const float2 psi = make_float2(at * c, at * s);
    const float2 g = G[i];
    const float re = g.x * psi.x + g.y * psi.y;  // Re(conj(g) psi)
    const float im = g.x * psi.y - g.y * psi.x;  // Im(conj(g) psi)
    p[i] = -k * im * gscale;
    b[i] = -k * re * gscale;
  }
}

static unsigned blocks_for(long long count) {
  long long g = (count + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, kMaxBlocks));
}

}  // namespace model
}  // namespace mm

using namespace mm;
using namespace mm::model;

struct mm_model {
  mm_radon* radon = nullptr;
  mm_fresnel* fresnel = nullptr;
  int n = 0, nz = 0, ntheta = 0;
  float k = 0.f;
  float* p = nullptr;        // ntheta x nz x n
  float* b = nullptr;        // ntheta x nz x n
  float2* u = nullptr;       // ntheta x nz x n
  double* partial = nullptr; // kMaxBlocks
  size_t bytes = 0;
};

static void model_free(mm_model* m) {
  if (!m) return;
  void* ptrs[] = {m->p, m->b, m->u, m->partial};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete m;
}

extern "C" {

int mm_model_create(mm_model** model, mm_radon* radon, mm_fresnel* fresnel, int nz, float k) {
  MM_ARG_CHECK(model != nullptr, "mm_model_create: model is NULL");
  *model = nullptr;
  MM_ARG_CHECK(radon && fresnel, "mm_model_create: null plan");
  int rn, ntheta, nzmax, fn, fnz, ndist;
  radon_dims(radon, &rn, &ntheta, &nzmax);
  fresnel_dims(fresnel, &fn, &fnz, &ndist);
  MM_ARG_CHECK(nz >= 1 && nz <= nzmax, "mm_model_create: nz outside the Radon plan's [1, nz_max]");
  MM_ARG_CHECK(fn == rn && fnz == nz,
               "mm_model_create: Fresnel frame (" + std::to_string(fnz) + ", " + std::to_string(fn) +
                   ") must equal (nz, n) = (" + std::to_string(nz) + ", " + std::to_string(rn) + ")");
  mm_model* m = new mm_model();
  m->radon = radon;
  m->fresnel = fresnel;
  m->n = rn;
  m->nz = nz;
  m->ntheta = ntheta;
  m->k = k;
  const size_t cnt = (size_t)ntheta * nz * rn;
  if (hipMalloc(&m->p, cnt * sizeof(float)) != hipSuccess || hipMalloc(&m->b, cnt * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->u, cnt * sizeof(float2)) != hipSuccess ||
      hipMalloc(&m->partial, kMaxBlocks * sizeof(double)) != hipSuccess) {
    model_free(m);
    return fail(MM_ERR_OOM, "mm_model_create: workspace allocation failed");
  }
  m->bytes = cnt * (2 * sizeof(float) + sizeof(float2)) + kMaxBlocks * sizeof(double);
  *model = m;
  return MM_OK;
}

static int model_forward_common(mm_model* m, const float* delta, const float* beta, int j, hipStream_t st) {
  int rc;
  if ((rc = mm_radon_fwd(m->radon, delta, m->p, m->nz, st))) return rc;
  if ((rc = mm_radon_fwd(m->radon, beta, m->b, m->nz, st))) return rc;
  const long long cnt = (long long)m->ntheta * m->nz * m->n;
  hipLaunchKernelGGL(psi_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->p, m->b, m->u, cnt, m->k);
  if ((rc = check_launch("model psi"))) return rc;
  return mm_fresnel_fwd(m->fresnel, m->u, m->u, m->ntheta, j, st);
}

int mm_model_fwd(mm_model* m, const float* delta, const float* beta, const float* data, int j, double* loss_dev,
                 void* stream) {
  MM_ARG_CHECK(m != nullptr, "mm_model_fwd: model is NULL");
  MM_ARG_CHECK(delta && beta && data && loss_dev, "mm_model_fwd: null pointer");
  hipStream_t st = (hipStream_t)stream;
  int rc = model_forward_common(m, delta, beta, j, st);
  if (rc) return rc;
  const long long cnt = (long long)m->ntheta * m->nz * m->n;
  const unsigned nb = blocks_for(cnt);
  hipLaunchKernelGGL(loss_partial_kernel, dim3(nb), dim3(kBlock), 0, st, m->u, data, cnt, m->partial);
  if ((rc = check_launch("model loss"))) return rc;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kBlock), 0, st, m->partial, (int)nb, loss_dev);
  return check_launch("model reduce");
}

int mm_model_bwd(mm_model* m, const float* delta, const float* beta, const float* data, int j, float gscale,
                 float* gdelta, float* gbeta, void* stream) {
  MM_ARG_CHECK(m != nullptr, "mm_model_bwd: model is NULL");
  MM_ARG_CHECK(delta && beta && data && gdelta && gbeta, "mm_model_bwd: null pointer");
  hipStream_t st = (hipStream_t)stream;
  int rc = model_forward_common(m, delta, beta, j, st);
  if (rc) return rc;
  const long long cnt = (long long)m->ntheta * m->nz * m->n;
  hipLaunchKernelGGL(residual_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->u, data, cnt);
  if ((rc = check_launch("model residual"))) return rc;
  if ((rc = mm_fresnel_adj(m->fresnel, m->u, m->u, m->ntheta, j, st))) return rc;
  hipLaunchKernelGGL(chain_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->u, m->p, m->b, cnt, m->k, gscale);
  if ((rc = check_launch("model chain"))) return rc;
  if ((rc = mm_radon_adj(m->radon, m->p, 0, gdelta, m->nz, st))) return rc;
  return mm_radon_adj(m->radon, m->b, 0, gbeta, m->nz, st);
}

size_t mm_model_workspace_bytes(const mm_model* m) { return m ? m->bytes : 0; }

void mm_model_destroy(mm_model* m) { model_free(m); }

}  // extern "C"
