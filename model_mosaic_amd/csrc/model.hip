// Differentiable two-physics forward model (delta, beta) -> |D(exp(i k R delta - k R beta))|^2
// with an L2 misfit and a hand-written backward (not in the reference; the
// reference pipeline runs R in step2_radon.py:191 and psi/D/|.|^2 in
// step3_propagation.py:220-225 with no adjoint).
//
// Forward:  p = R delta, b = R beta, psi = exp(i k p - k b), u = D psi,
//           L = sum (|u|^2 - d)^2   (FP64 block partials + one-block reduce).
// Backward: r = 4 (|u|^2 - d) u,  G = DT r,
//           dL/dp = -k Im(conj(G) psi),  dL/db = -k Re(conj(G) psi),
//           dL/ddelta = RT dL/dp,  dL/dbeta = RT dL/db.
// The forward leaves p, b and u in the model's workspace under a token; a
// backward given that token (no other forward since) starts from them and
// recomputes only psi, otherwise it reruns the forward first.  psi itself is
// never stored.  With a mirror-padded Fresnel plan on the DCT form (the
// default sizes) psi is built inside D's column pass, the loss partials come
// from its row pass, 4 (|u|^2 - d) u is formed inside DT's column pass and the
// chain rule inside its row pass (fresnel_model_fwd / fresnel_model_bwd);
// other plans run the element-wise kernels below around D and DT.
#include <algorithm>
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
  int n = 0, nz = 0, ntheta = 0, ndist = 0;
  float k = 0.f;
  bool fused = false;        // Fresnel plan on the DCT form: element-wise steps inside its passes
  float* p = nullptr;        // ntheta x nz x n
  float* b = nullptr;        // ntheta x nz x n
  float2* u = nullptr;       // ntheta x nz x n
  double* partial = nullptr; // loss partials
  size_t bytes = 0;
  unsigned long long issued = 0;  // tokens handed out so far
  unsigned long long state = 0;   // token of the forward whose p, b, u the workspace holds (0: none)
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
  m->ndist = ndist;
  m->k = k;
  const size_t cnt = (size_t)ntheta * nz * rn;
  m->fused = fresnel_fused_ok(fresnel);
  long long nparts = kMaxBlocks;
  if (m->fused) {
    long long fp = 0;
    int rc = fresnel_model_parts(fresnel, ntheta, &fp);
    if (rc) {
      model_free(m);
      return rc;
    }
    nparts = std::max<long long>(nparts, fp);
  }
  if (hipMalloc(&m->p, cnt * sizeof(float)) != hipSuccess || hipMalloc(&m->b, cnt * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->u, cnt * sizeof(float2)) != hipSuccess ||
      hipMalloc(&m->partial, nparts * sizeof(double)) != hipSuccess) {
    model_free(m);
    return fail(MM_ERR_OOM, "mm_model_create: workspace allocation failed");
  }
  m->bytes = cnt * (2 * sizeof(float) + sizeof(float2)) + nparts * sizeof(double);
  *model = m;
  return MM_OK;
}

// p = R delta, b = R beta, u = D psi and the loss partials (*nparts of them in m->partial).
static int model_forward_state(mm_model* m, const float* delta, const float* beta, const float* data, int j,
                               int* nparts, hipStream_t st) {
  int rc;
  m->state = 0;
  if ((rc = mm_radon_fwd(m->radon, delta, m->p, m->nz, st))) return rc;
  if ((rc = mm_radon_fwd(m->radon, beta, m->b, m->nz, st))) return rc;
  if (m->fused) return fresnel_model_fwd(m->fresnel, m->p, m->b, m->k, m->u, data, m->partial, nparts, m->ntheta, j, st);
  const long long cnt = (long long)m->ntheta * m->nz * m->n;
  hipLaunchKernelGGL(psi_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->p, m->b, m->u, cnt, m->k);
  if ((rc = check_launch("model psi"))) return rc;
  if ((rc = mm_fresnel_fwd(m->fresnel, m->u, m->u, m->ntheta, j, st))) return rc;
  const unsigned nb = blocks_for(cnt);
  hipLaunchKernelGGL(loss_partial_kernel, dim3(nb), dim3(kBlock), 0, st, m->u, data, cnt, m->partial);
  *nparts = (int)nb;
  return check_launch("model loss");
}

int mm_model_fwd(mm_model* m, const float* delta, const float* beta, const float* data, int j, double* loss_dev,
                 unsigned long long* token, void* stream) {
  MM_ARG_CHECK(m != nullptr, "mm_model_fwd: model is NULL");
  MM_ARG_CHECK(delta && beta && data && loss_dev, "mm_model_fwd: null pointer");
  MM_ARG_CHECK(j >= 0 && j < m->ndist, "mm_model_fwd: distance index " + std::to_string(j) + " outside [0, " +
                                           std::to_string(m->ndist) + ")");
  hipStream_t st = (hipStream_t)stream;
  int nparts = 0;
  int rc = model_forward_state(m, delta, beta, data, j, &nparts, st);
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kBlock), 0, st, m->partial, nparts, loss_dev);
  if ((rc = check_launch("model reduce"))) return rc;
  m->state = ++m->issued;
  if (token) *token = m->state;
  return MM_OK;
}

int mm_model_bwd(mm_model* m, const float* delta, const float* beta, const float* data, int j, float gscale,
                 unsigned long long token, float* gdelta, float* gbeta, void* stream) {
  MM_ARG_CHECK(m != nullptr, "mm_model_bwd: model is NULL");
  MM_ARG_CHECK(delta && beta && data && gdelta && gbeta, "mm_model_bwd: null pointer");
  MM_ARG_CHECK(j >= 0 && j < m->ndist, "mm_model_bwd: distance index " + std::to_string(j) + " outside [0, " +
                                           std::to_string(m->ndist) + ")");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if (token == 0 || token != m->state) {  // the workspace does not hold this forward's p, b, u
    int nparts = 0;
    if ((rc = model_forward_state(m, delta, beta, data, j, &nparts, st))) return rc;
  }
  m->state = 0;  // u, p, b are overwritten below
  if (m->fused) {
    if ((rc = fresnel_model_bwd(m->fresnel, m->u, data, m->p, m->b, m->k, gscale, m->ntheta, j, st))) return rc;
  } else {
    const long long cnt = (long long)m->ntheta * m->nz * m->n;
    hipLaunchKernelGGL(residual_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->u, data, cnt);
    if ((rc = check_launch("model residual"))) return rc;
    if ((rc = mm_fresnel_adj(m->fresnel, m->u, m->u, m->ntheta, j, st))) return rc;
    hipLaunchKernelGGL(chain_kernel, dim3(blocks_for(cnt)), dim3(kBlock), 0, st, m->u, m->p, m->b, cnt, m->k, gscale);
    if ((rc = check_launch("model chain"))) return rc;
  }
  if ((rc = mm_radon_adj(m->radon, m->p, 0, gdelta, m->nz, st))) return rc;
  return mm_radon_adj(m->radon, m->b, 0, gbeta, m->nz, st);
}

size_t mm_model_workspace_bytes(const mm_model* m) { return m ? m->bytes : 0; }

void mm_model_destroy(mm_model* m) { model_free(m); }

}  // extern "C"
