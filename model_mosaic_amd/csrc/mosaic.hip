// Uneven mosaic batch: R, psi, Fresnel D and |.|^2 over tiles of different
// detector widths, angle counts and z stacks, enqueued as one HIP graph.
//
// The reference runs the same chain tile by tile from its pipeline scripts:
// R per z chunk of each tile (step2_radon.py:147,191), then psi, D and |.|^2
// per angle chunk (step3_propagation.py:175,223 with _psi_from_proj and
// _abs2_c64_to_f32 at :94-110), one host loop iteration and a round of
// launches per chunk.  Here the whole batch is one graph: plans are shared by
// tiles of equal geometry, the half-spectrum, representative, sinogram and psi
// buffers are shared by all tiles (sized for the largest; the tiles run one
// after another on one stream), Radon plans are forward-only (no scatter
// tables), and the graph is captured once per set of input / output pointers
// and replayed as a single hipGraphLaunch.  With mirror padding psi is built
// inside the Fresnel column pass and |.|^2 taken inside its row pass
// (fresnel_intensity), so a tile is R plus two Fresnel passes.
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "mm_common.h"

using namespace mm;

struct mm_mosaic {
  struct Tile {
    int n = 0, nz = 0, ntheta = 0, rplan = -1, fplan = -1;
    long long count = 0;  // ntheta * nz * n
  };
  int dev = 0;
  unsigned flags = 0;
  float k = 0.f, beta_ratio = 0.f;
  std::vector<Tile> tiles;
  std::vector<mm_radon*> rplans;
  std::vector<int> rplan_first;  // first tile of each Radon plan (autotune input)
  std::vector<mm_fresnel*> fplans;
  float2* fde = nullptr;
  float2* gh = nullptr;
  float* p = nullptr;
  float2* psi = nullptr;
  hipStream_t st = nullptr;  // capturable stream of the batch (the caller's may be the legacy default stream)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> cap_ptrs;  // inputs and outputs the graph was captured with
  int captures = 0;
  bool tuned = false;
  size_t bytes = 0;
};

static void mosaic_free(mm_mosaic* m) {
  if (!m) return;
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (mm_radon* r : m->rplans) mm_radon_destroy(r);
  for (mm_fresnel* f : m->fplans) mm_fresnel_destroy(f);
  void* ptrs[] = {m->fde, m->gh, m->p, m->psi};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (m->ev_in) (void)hipEventDestroy(m->ev_in);
  if (m->ev_out) (void)hipEventDestroy(m->ev_out);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

// The chain of every tile, in tile order, on stream st.
static int mosaic_enqueue(mm_mosaic* m, const float* const* delta, float* const* out, hipStream_t st) {
  int rc;
  for (size_t t = 0; t < m->tiles.size(); ++t) {
    const mm_mosaic::Tile& T = m->tiles[t];
    if ((rc = mm_radon_fwd(m->rplans[T.rplan], delta[t], m->p, T.nz, st))) return rc;
    mm_fresnel* fp = m->fplans[T.fplan];
    if (fresnel_fused_ok(fp)) {  // psi built in D's column pass, |.|^2 taken in its row pass
      if ((rc = fresnel_intensity(fp, m->p, m->k, m->beta_ratio, m->psi, out[t], T.ntheta, 0, st))) return rc;
    } else if ((rc = mm_psi_from_proj(m->p, m->psi, T.count, m->k, m->beta_ratio, st)) ||
               (rc = mm_fresnel_fwd(fp, m->psi, m->psi, T.ntheta, 0, st)) ||
               (rc = mm_abs2(m->psi, out[t], T.count, st))) {
      return rc;
    }
  }
  return MM_OK;
}

extern "C" {

int mm_mosaic_create(mm_mosaic** out, int ntiles, const int* n, const int* nz, const int* ntheta,
                     const double* theta_all, double wavelength, double voxelsize, double distance, float k,
                     float beta_ratio, int pad_mode, unsigned flags) {
  MM_ARG_CHECK(out != nullptr, "mm_mosaic_create: plan is NULL");
  *out = nullptr;
  MM_ARG_CHECK(ntiles >= 1 && n && nz && ntheta && theta_all, "mm_mosaic_create: need ntiles >= 1 and all arrays");
  MM_ARG_CHECK((flags & ~(MM_MOSAIC_NO_GRAPH | MM_MOSAIC_AUTOTUNE)) == 0, "mm_mosaic_create: unknown flags");
  MM_ARG_CHECK(pad_mode == 0 || pad_mode == 1, "mm_mosaic_create: pad_mode must be 0 or 1");
  mm_mosaic* m = new mm_mosaic();
  auto bail = [&](int code) {
    mosaic_free(m);
    return code;
  };
  if (hipGetDevice(&m->dev) != hipSuccess) return bail(fail(MM_ERR_HIP, "hipGetDevice"));
  m->flags = flags;
  m->k = k;
  m->beta_ratio = beta_ratio;
  // group the tiles: Radon plans by (n, angles), Fresnel plans by (n, nz)
  std::map<std::tuple<int, std::vector<double>>, int> rkey;
  std::map<std::pair<int, int>, int> fkey;
  std::vector<int> rnz;                     // nz_max of each Radon plan
  std::vector<std::pair<int, const double*>> rgeo;  // (n, theta) of each Radon plan
  std::vector<int> rnt;
  long long th_off = 0, max_count = 0;
  size_t max_fde = 0, max_gh = 0;
  for (int t = 0; t < ntiles; ++t) {
    MM_ARG_CHECK(n[t] >= 16 && nz[t] >= 1 && ntheta[t] >= 1, "mm_mosaic_create: bad tile " + std::to_string(t));
    mm_mosaic::Tile T;
    T.n = n[t];
    T.nz = nz[t];
    T.ntheta = ntheta[t];
    T.count = (long long)T.ntheta * T.nz * T.n;
    const double* th = theta_all + th_off;
    th_off += ntheta[t];
    auto rk = std::make_tuple(T.n, std::vector<double>(th, th + T.ntheta));
    auto it = rkey.find(rk);
    if (it == rkey.end()) {
      it = rkey.emplace(rk, (int)rgeo.size()).first;
      rgeo.push_back({T.n, th});
      rnt.push_back(T.ntheta);
      rnz.push_back(0);
      m->rplan_first.push_back(t);
    }
    T.rplan = it->second;
    rnz[T.rplan] = std::max(rnz[T.rplan], T.nz);
    auto fk = std::make_pair(T.n, T.nz);
    auto jt = fkey.find(fk);
    if (jt == fkey.end()) jt = fkey.emplace(fk, (int)fkey.size()).first;
    T.fplan = jt->second;
    max_count = std::max(max_count, T.count);
    m->tiles.push_back(T);
  }
  for (size_t r = 0; r < rgeo.size(); ++r) {
    max_fde = std::max(max_fde, radon_fde_elems(rgeo[r].first, rnz[r]));
    max_gh = std::max(max_gh, radon_gh_elems(rgeo[r].first, rnz[r], rnt[r]));
  }
  auto alloc = [&](void** q, size_t nb, const char* what) {
    if (hipMalloc(q, nb) != hipSuccess) {
      *q = nullptr;
      return fail(MM_ERR_OOM, std::string("mm_mosaic_create: ") + what + " (" + std::to_string(nb) + " bytes)");
    }
    m->bytes += nb;
    return MM_OK;
  };
  int rc;
  if ((rc = alloc((void**)&m->fde, sizeof(float2) * max_fde, "half-spectrum workspace")) ||
      (rc = alloc((void**)&m->gh, sizeof(float2) * max_gh, "representative workspace")) ||
      (rc = alloc((void**)&m->p, sizeof(float) * (size_t)max_count, "sinogram workspace")) ||
      (rc = alloc((void**)&m->psi, sizeof(float2) * (size_t)max_count, "psi workspace")))
    return bail(rc);
  for (size_t r = 0; r < rgeo.size(); ++r) {
    mm_radon* rp = nullptr;
    if ((rc = radon_create_impl(&rp, rgeo[r].first, rnz[r], rnt[r], rgeo[r].second, 0, m->fde, m->gh, true)))
      return bail(rc);
    m->rplans.push_back(rp);
    m->bytes += mm_radon_workspace_bytes(rp);
  }
  m->fplans.resize(fkey.size(), nullptr);
  for (const auto& kv : fkey) {
    mm_fresnel* fp = nullptr;
    if ((rc = mm_fresnel_create(&fp, kv.first.first, kv.first.second, 1, wavelength, voxelsize, &distance, pad_mode)))
      return bail(rc);
    m->fplans[kv.second] = fp;
    m->bytes += mm_fresnel_workspace_bytes(fp);
  }
  if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MM_ERR_HIP, "mm_mosaic_create: stream / events"));
  *out = m;
  return MM_OK;
}

int mm_mosaic_run(mm_mosaic* m, const float* const* delta, float* const* out, void* stream) {
  MM_ARG_CHECK(m != nullptr, "mm_mosaic_run: plan is NULL");
  MM_ARG_CHECK(delta && out, "mm_mosaic_run: null pointer array");
  const size_t nt = m->tiles.size();
  for (size_t t = 0; t < nt; ++t) MM_ARG_CHECK(delta[t] && out[t], "mm_mosaic_run: null pointer for tile " + std::to_string(t));
  hipStream_t caller = (hipStream_t)stream;
  MM_HIP_TRY(hipEventRecord(m->ev_in, caller));
  MM_HIP_TRY(hipStreamWaitEvent(m->st, m->ev_in, 0));
  int rc;
  if ((m->flags & MM_MOSAIC_AUTOTUNE) && !m->tuned) {
    for (size_t r = 0; r < m->rplans.size(); ++r) {
      const int t = m->rplan_first[r];
      if ((rc = mm_radon_autotune(m->rplans[r], delta[t], m->p, nullptr, m->tiles[t].nz, m->st))) return rc;
    }
    m->tuned = true;
  }
  if (m->flags & MM_MOSAIC_NO_GRAPH) {
    if ((rc = mosaic_enqueue(m, delta, out, m->st))) return rc;
  } else {
    std::vector<const void*> ptrs(delta, delta + nt);
    ptrs.insert(ptrs.end(), out, out + nt);
    if (!m->exec || ptrs != m->cap_ptrs) {
      if (m->exec) (void)hipGraphExecDestroy(m->exec);
      if (m->graph) (void)hipGraphDestroy(m->graph);
      m->exec = nullptr;
      m->graph = nullptr;
      MM_HIP_TRY(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
      rc = mosaic_enqueue(m, delta, out, m->st);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(m->st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(MM_ERR_HIP, std::string("mm_mosaic_run: capture: ") + hipGetErrorString(e));
      m->graph = g;
      MM_HIP_TRY(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
      m->cap_ptrs = ptrs;
      ++m->captures;
    }
    MM_HIP_TRY(hipGraphLaunch(m->exec, m->st));
  }
  MM_HIP_TRY(hipEventRecord(m->ev_out, m->st));
  MM_HIP_TRY(hipStreamWaitEvent(caller, m->ev_out, 0));
  return MM_OK;
}

int mm_mosaic_info(const mm_mosaic* m, int what) {
  if (!m) return MM_ERR_ARG;
  switch (what) {
    case MM_MOSAIC_INFO_RADON_PLANS: return (int)m->rplans.size();
    case MM_MOSAIC_INFO_FRESNEL_PLANS: return (int)m->fplans.size();
    case MM_MOSAIC_INFO_CAPTURES: return m->captures;
    case MM_MOSAIC_INFO_TILES: return (int)m->tiles.size();
    default: return MM_ERR_ARG;
  }
}

int mm_mosaic_tile_knob(const mm_mosaic* m, int tile, int knob) {
  if (!m || tile < 0 || tile >= (int)m->tiles.size()) return MM_ERR_ARG;
  return mm_radon_get_tune(m->rplans[m->tiles[tile].rplan], knob);
}

size_t mm_mosaic_workspace_bytes(const mm_mosaic* m) { return m ? m->bytes : 0; }

void mm_mosaic_destroy(mm_mosaic* m) { mosaic_free(m); }

}  // extern "C"
