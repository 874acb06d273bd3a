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
//
// The tiles are split over B concurrent branches of the graph (default 2):
// each branch has its own half-spectrum, representative, sinogram and psi
// workspaces and its own stream, forked from and joined back to the batch
// stream, and runs its tiles one after another; the plans are shared (R runs
// on the branch's workspace through radon_fwd_ws).  Tiles go to branches by
// longest-first greedy balancing of their HBM bytes, so one tile's launch
// tails overlap another's work.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "mm_common.h"

using namespace mm;

struct mm_mosaic {
  struct Tile {
    int n = 0, nz = 0, ntheta = 0, rplan = -1, fplan = -1, branch = 0;
    long long count = 0;  // ntheta * nz * n
  };
  struct Branch {
    float2* fde = nullptr;
    float2* gh = nullptr;
    float* p = nullptr;
    float2* psi = nullptr;
    hipStream_t st = nullptr;  // branch 0 runs on the batch stream itself
    hipEvent_t ev_join = nullptr;
    std::vector<int> tiles;
  };
  int dev = 0;
  unsigned flags = 0;
  float k = 0.f, beta_ratio = 0.f;
  std::vector<Tile> tiles;
  std::vector<mm_radon*> rplans;
  std::vector<int> rplan_first;  // first tile of each Radon plan (autotune input)
  std::vector<mm_fresnel*> fplans;
  std::vector<Branch> branches;
  hipStream_t st = nullptr;  // capturable stream of the batch (the caller's may be the legacy default stream)
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_fork = nullptr;
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
  for (mm_mosaic::Branch& b : m->branches) {
    void* ptrs[] = {b.fde, b.gh, b.p, b.psi};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if (b.st) (void)hipStreamDestroy(b.st);
    if (b.ev_join) (void)hipEventDestroy(b.ev_join);
  }
  if (m->ev_in) (void)hipEventDestroy(m->ev_in);
  if (m->ev_out) (void)hipEventDestroy(m->ev_out);
  if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

// The chain of every tile: forked from st onto the branches, each branch's tiles in tile order on its own
// stream and workspaces, joined back to st.
static int mosaic_enqueue(mm_mosaic* m, const float* const* delta, float* const* out, hipStream_t st) {
  MM_HIP_TRY(hipEventRecord(m->ev_fork, st));
  for (size_t b = 1; b < m->branches.size(); ++b) MM_HIP_TRY(hipStreamWaitEvent(m->branches[b].st, m->ev_fork, 0));
  int rc = MM_OK;
  for (size_t b = 0; b < m->branches.size() && rc == MM_OK; ++b) {
    const mm_mosaic::Branch& B = m->branches[b];
    hipStream_t bs = b == 0 ? st : B.st;
    for (size_t i = 0; i < B.tiles.size() && rc == MM_OK; ++i) {
      const int t = B.tiles[i];
      const mm_mosaic::Tile& T = m->tiles[t];
      if ((rc = radon_fwd_ws(m->rplans[T.rplan], delta[t], B.p, T.nz, bs, B.fde, B.gh))) break;
      mm_fresnel* fp = m->fplans[T.fplan];
      if (fresnel_fused_ok(fp))  // psi built in D's column pass, |.|^2 taken in its row pass
        rc = fresnel_intensity(fp, B.p, m->k, m->beta_ratio, B.psi, out[t], T.ntheta, 0, bs);
      else if (!(rc = mm_psi_from_proj(B.p, B.psi, T.count, m->k, m->beta_ratio, bs)) &&
               !(rc = mm_fresnel_fwd(fp, B.psi, B.psi, T.ntheta, 0, bs)))
        rc = mm_abs2(B.psi, out[t], T.count, bs);
    }
  }
  // the branches are joined back even after an error, so that no stream is left forked off a capture
  for (size_t b = 1; b < m->branches.size(); ++b) {
    MM_HIP_TRY(hipEventRecord(m->branches[b].ev_join, m->branches[b].st));
    MM_HIP_TRY(hipStreamWaitEvent(st, m->branches[b].ev_join, 0));
  }
  return rc;
}

extern "C" {

int mm_mosaic_create(mm_mosaic** out, int ntiles, const int* n, const int* nz, const int* ntheta,
                     const double* theta_all, double wavelength, double voxelsize, double distance, float k,
                     float beta_ratio, int pad_mode, unsigned flags) {
  MM_ARG_CHECK(out != nullptr, "mm_mosaic_create: plan is NULL");
  *out = nullptr;
  MM_ARG_CHECK(ntiles >= 1 && n && nz && ntheta && theta_all, "mm_mosaic_create: need ntiles >= 1 and all arrays");
  MM_ARG_CHECK((flags & ~(MM_MOSAIC_NO_GRAPH | MM_MOSAIC_AUTOTUNE | MM_MOSAIC_BRANCHES(15))) == 0,
               "mm_mosaic_create: unknown flags");
  const int nbranch_req = (int)((flags >> 8) & 15u);
  MM_ARG_CHECK(nbranch_req <= 8, "mm_mosaic_create: at most 8 branches");
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
  long long th_off = 0;
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
    m->tiles.push_back(T);
  }
  // branches: longest-first greedy balance of the tiles' HBM bytes (object + 32 B per projection pixel)
  const int nbr = std::min(nbranch_req ? nbranch_req : 2, ntiles);
  m->branches.resize(nbr);
  {
    std::vector<int> order(ntiles);
    for (int t = 0; t < ntiles; ++t) order[t] = t;
    auto cost = [&](int t) {
      const mm_mosaic::Tile& T = m->tiles[t];
      return 4.0 * T.nz * (double)T.n * T.n + 32.0 * (double)T.count;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    std::vector<double> load(nbr, 0.0);
    for (int t : order) {
      const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[b] += cost(t);
      m->tiles[t].branch = b;
    }
    for (int t = 0; t < ntiles; ++t) m->branches[m->tiles[t].branch].tiles.push_back(t);
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
  for (mm_mosaic::Branch& B : m->branches) {  // each branch's workspaces fit the largest of its tiles
    size_t bf = 0, bg = 0;
    long long bc = 0;
    for (int t : B.tiles) {
      const mm_mosaic::Tile& T = m->tiles[t];
      bf = std::max(bf, radon_fde_elems(T.n, T.nz));
      bg = std::max(bg, radon_gh_elems(T.n, T.nz, T.ntheta));
      bc = std::max(bc, T.count);
    }
    if ((rc = alloc((void**)&B.fde, sizeof(float2) * bf, "half-spectrum workspace")) ||
        (rc = alloc((void**)&B.gh, sizeof(float2) * bg, "representative workspace")) ||
        (rc = alloc((void**)&B.p, sizeof(float) * (size_t)bc, "sinogram workspace")) ||
        (rc = alloc((void**)&B.psi, sizeof(float2) * (size_t)bc, "psi workspace")))
      return bail(rc);
  }
  // the plans' own half-spectrum / representative pointers are branch 0's (R always runs through
  // radon_fwd_ws with the branch's workspaces; the autotuner runs on branch 0)
  for (size_t r = 0; r < rgeo.size(); ++r) {
    mm_radon* rp = nullptr;
    if ((rc = radon_create_impl(&rp, rgeo[r].first, rnz[r], rnt[r], rgeo[r].second, 0, m->branches[0].fde,
                                m->branches[0].gh, true)))
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
      hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess)
    return bail(fail(MM_ERR_HIP, "mm_mosaic_create: stream / events"));
  for (size_t b = 1; b < m->branches.size(); ++b)  // default-priority streams, like the batch stream
    if (hipStreamCreateWithFlags(&m->branches[b].st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->branches[b].ev_join, hipEventDisableTiming) != hipSuccess)
      return bail(fail(MM_ERR_HIP, "mm_mosaic_create: branch streams / events"));
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
      const size_t before = mm_radon_workspace_bytes(m->rplans[r]);  // the gather tables may be rebuilt
      if ((rc = mm_radon_autotune(m->rplans[r], delta[t], m->branches[0].p, nullptr, m->tiles[t].nz, m->st)))
        return rc;
      m->bytes = m->bytes - before + mm_radon_workspace_bytes(m->rplans[r]);
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
    case MM_MOSAIC_INFO_BRANCHES: return (int)m->branches.size();
    default: return MM_ERR_ARG;
  }
}

int mm_mosaic_tile_knob(const mm_mosaic* m, int tile, int knob) {
  if (!m || tile < 0 || tile >= (int)m->tiles.size()) return MM_ERR_ARG;
  if (knob == MM_MOSAIC_TILE_BRANCH) return m->tiles[tile].branch;
  return mm_radon_get_tune(m->rplans[m->tiles[tile].rplan], knob);
}

size_t mm_mosaic_workspace_bytes(const mm_mosaic* m) { return m ? m->bytes : 0; }

void mm_mosaic_destroy(mm_mosaic* m) { mosaic_free(m); }

}  // extern "C"
