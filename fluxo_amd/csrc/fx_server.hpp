// Batched serving of independent solves (fx_server_* in include/fluxo_gpu.h).
//
// Compiled as part of fx_engine.hip, after fx_launch.hpp.
//
// Layout: one Pool per (N, parabolic).  A pool owns device arrays with room for Ecap
// elements, Scap sides and Rcap problems; a request occupies a contiguous element range,
// a contiguous side range and one problem slot r (its DtState st[r] and its element range
// rfirst[r]/rcount[r]).  Connectivity is stored with pool-global ids, so the kernels of one
// step run over [0, Ehw) elements and [0, Shw) sides of all requests at once; unused
// element slots carry eprob = -1 and unused side slots a zeroed SideToElem, which every
// kernel skips.  A request never references another request's slots (its mesh is
// periodic and self-contained).
//
// The worker thread is the only thread that calls HIP.  Each pool steps on its own stream,
// independently of the others: when a pool's step completes (its event), the worker counts
// the step, reads every active problem's DtState (pinned read-back queued behind the step),
// retires finished requests (result copied to the caller's U_out, slots freed), admits queued
// requests into it (HBM budget, pool slots; pools grow by reallocation) and launches its
// next step graph.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace fx {
namespace srv {

using Clock = std::chrono::steady_clock;

// first-fit allocator of contiguous [start, start+len) ranges in [0, cap)
struct RangeAlloc {
  std::map<int, int> free_;  // start -> len
  int cap = 0;
  void grow(int newcap)
  {
    // extend (or create) the free range that ends at the old capacity
    int start = cap, len = newcap - cap;
    auto it = free_.empty() ? free_.end() : std::prev(free_.end());
    if (it != free_.end() && it->first + it->second == cap) { start = it->first; len += it->second; free_.erase(it); }
    free_[start] = len;
    cap = newcap;
  }
  int alloc(int len)
  {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= len) {
        int s = it->first, l = it->second;
        free_.erase(it);
        if (l > len) free_[s + len] = l - len;
        return s;
      }
    return -1;
  }
  void release(int s, int len)
  {
    auto it = free_.emplace(s, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && s + len == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
  // smallest index above every allocated range
  int high_water() const
  {
    if (free_.empty()) return cap;
    auto last = std::prev(free_.end());
    return (last->first + last->second == cap) ? last->first : cap;
  }
};

struct Request {
  fx_request req{};
  int id = -1;
  int state = FX_REQ_QUEUED;
  int errType = 0, steps = 0;
  double t_end = 0.0, dt_last = 0.0;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  int pool = -1, e0 = -1, s0 = -1, r = -1;
  long long bytes = 0;
};

struct Pool {
  int N = 0, parabolic = 0;
  Batch b;  // device arrays; b.cfg.nElems/nSides/nProb = current launch extents
  RangeAlloc ea, sa, ra;
  int nActive = 0;
  hipGraphExec_t graph = nullptr;
  int gE = -1, gS = -1, gR = -1;  // extents the graph was captured with
  bool ptrs_changed = true;
  DtState *st_host = nullptr;     // pinned: the per-problem step state read back after each step
  int st_cap = 0;
  hipEvent_t done = nullptr;      // recorded after a step's graph and state read-back
  bool inflight = false;          // a step is queued on the stream and not yet seen complete
  std::vector<int> active_ids;
  long long bytes_alloc = 0;
};

}  // namespace srv
}  // namespace fx

struct fx_server {
  fx_server_config cfg{};
  fx::srv::Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int> queue;                 // ids waiting for admission (FIFO)
  std::vector<fx::srv::Request> reqs;    // by id
  bool stop = false;
  std::thread worker;
  fx::srv::Pool pools[2 * (fx::MAXN + 1)];
  long long bytes_in_use = 0, bytes_peak = 0, steps_launched = 0, graph_captures = 0, graph_updates = 0;
  long long nodes_active = 0;      // volume nodes of the admitted requests (admitKNodes cap)
  long long bytes_allocated = 0;   // sum of the pools' device arrays
  int nActive = 0, nDone = 0, nFailed = 0;
  char worker_err[512] = "";
};

namespace fx {
namespace srv {

inline double seconds_since(const fx_server *s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

inline long long request_bytes(const fx_server_config &c, int N, int par, long long E, long long S)
{
  const long long n = N + 1, n3 = n * n * n, n2 = n * n;
  long long per_e = 2 * 5 * n3 * 8 + 9 * n3 * 8 + n3 * 8 + 12 * 4 + 4 + 8 + 4 + 6 * n2 * 8;  // ... face-point table
  if (c.precision == FX_PREC_FP32_SPLIT) per_e += 9 * n3 * 4;
  if (par) per_e += 4 * n3 * 8;
  const long long per_s = 5 * n2 * 8 + n2 * (long long)sizeof(double4) + 5 * 4 + (par ? 8 * n2 * 8 : 0);
  return per_e * E + per_s * S + (long long)sizeof(DtState) + 8;
}

// (Re)allocate every device array of a pool with the given capacities, keeping contents.
inline int pool_realloc(Pool &p, const fx_server_config &c, int Ecap, int Scap, int Rcap)
{
  Batch &b = p.b;
  const size_t n3 = b.n3, n2 = b.n2;
  const size_t oE = p.ea.cap, oS = p.sa.cap, oR = p.ra.cap;
  if (!face_table_fits(p.N, Ecap, Scap))
    return set_err(FX_ENOMEM, "pool of N=%d: %d elements / %d sides beyond 32-bit face-table indices", p.N, Ecap, Scap);
  long long bytes = 0;
  auto grow = [&](auto *&ptr, size_t per, size_t oldn, size_t newn, int fill) -> int {
    using T = std::remove_reference_t<decltype(*ptr)>;
    T *np = nullptr;
    hipError_t e = hipMalloc((void **)&np, per * newn * sizeof(T));
    if (e != hipSuccess)
      return set_err(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP, "pool allocation: %s", hipGetErrorString(e));
    bytes += (long long)(per * newn * sizeof(T));
    if (fill >= 0) HIPCHK(hipMemsetAsync(np, fill, per * newn * sizeof(T), b.stream));
    if (ptr && oldn) HIPCHK(hipMemcpyAsync(np, ptr, per * oldn * sizeof(T), hipMemcpyDeviceToDevice, b.stream));
    HIPCHK(hipStreamSynchronize(b.stream));
    if (ptr) (void)hipFree(ptr);
    ptr = np;
    return FX_OK;
  };
  int rc;
#define GROW(ptr, per, oldn, newn, fill) \
  if ((rc = grow(ptr, per, oldn, newn, fill)) != FX_OK) return rc;
  GROW(b.U, 5 * n3, oE, Ecap, -1);
  GROW(b.Uttmp, 5 * n3, oE, Ecap, -1);
  GROW(b.Mf, 3 * n3, oE, Ecap, -1);
  GROW(b.Mg, 3 * n3, oE, Ecap, -1);
  GROW(b.Mh, 3 * n3, oE, Ecap, -1);
  GROW(b.sJ, n3, oE, Ecap, -1);
  if (c.precision == FX_PREC_FP32_SPLIT) {
    GROW(b.Mf32, 3 * n3, oE, Ecap, -1);
    GROW(b.Mg32, 3 * n3, oE, Ecap, -1);
    GROW(b.Mh32, 3 * n3, oE, Ecap, -1);
  }
  if (p.parabolic) GROW(b.Uv, 4 * n3, oE, Ecap, -1);
  GROW(b.e2s, 12, oE, Ecap, 0);
  GROW(b.ftab, 6 * n2, oE, Ecap, 0);
  GROW(b.eprob, 1, oE, Ecap, 0xFF);  // -1: unused
  GROW(b.dtElem, 1, oE, Ecap, -1);
  GROW(b.errElem, 1, oE, Ecap, 0);
  GROW(b.Flux, 5 * n2, oS, Scap, -1);
  GROW(b.nvse, n2, oS, Scap, -1);
  if (p.parabolic) GROW(b.FvN, 8 * n2, oS, Scap, -1);
  GROW(b.s2e, 5, oS, Scap, 0);       // zero SideToElem: unused side slot
  GROW(b.st, 1, oR, Rcap, 0);
  GROW(b.rfirst, 1, oR, Rcap, 0);
  GROW(b.rcount, 1, oR, Rcap, 0);    // count 0: unused problem slot
#undef GROW
  if (Ecap > (int)oE) p.ea.grow(Ecap);
  if (Scap > (int)oS) p.sa.grow(Scap);
  if (Rcap > (int)oR) p.ra.grow(Rcap);
  if (Rcap > p.st_cap) {
    DtState *h = nullptr;
    HIPCHK(hipHostMalloc((void **)&h, sizeof(DtState) * Rcap, hipHostMallocDefault));
    if (p.st_host) (void)hipHostFree(p.st_host);
    p.st_host = h;
    p.st_cap = Rcap;
  }
  p.bytes_alloc = bytes;
  p.ptrs_changed = true;
  return FX_OK;
}

inline void pool_free(Pool &p)
{
  Batch &b = p.b;
  if (b.stream) (void)hipStreamSynchronize(b.stream);
  if (p.graph) (void)hipGraphExecDestroy(p.graph);
  p.graph = nullptr;
  void *ptrs[] = {b.U, b.Uttmp, b.Mf, b.Mg, b.Mh, b.sJ, b.Mf32, b.Mg32, b.Mh32, b.Uv, b.FvN, b.e2s, b.ftab, b.eprob,
                  b.dtElem, b.errElem, b.Flux, b.nvse, b.s2e, b.st, b.rfirst, b.rcount};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (b.stream) (void)hipStreamDestroy(b.stream);
  if (p.st_host) (void)hipHostFree(p.st_host);
  if (p.done) (void)hipEventDestroy(p.done);
  p.st_host = nullptr; p.st_cap = 0; p.done = nullptr; p.inflight = false;
  b = Batch();
}

inline int pool_init(Pool &p, const fx_server_config &c, int N, int par)
{
  p.N = N;
  p.parabolic = par;
  Batch &b = p.b;
  b.cfg.N = N; b.cfg.nVar = 5; b.cfg.volumeFlux = c.volumeFlux; b.cfg.riemann = c.riemann;
  b.cfg.parabolic = par; b.cfg.lifting = c.lifting; b.cfg.precision = c.precision;
  b.cfg.kappa = c.kappa; b.cfg.mu0 = c.mu0; b.cfg.Pr = c.Pr;
  b.set_degree(N);
  b.vol_tile = tuned_tile(N, par, c.precision);
  b.cfg.nElems = 0; b.cfg.nSides = 0; b.nProb = 0;
  int rc = ensure_ops(N);
  if (rc) return rc;
  HIPCHK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&p.done, hipEventDisableTiming));
  return FX_OK;
}

// Copy a request's mesh and state into its pool slots (ids shifted to pool-global).
inline int upload_request(Pool &p, Request &q)
{
  Batch &b = p.b;
  const fx_request &r = q.req;
  const int E = r.nElems, S = r.nSides;
  const size_t n3 = b.n3, n2 = b.n2, n3E = n3 * E, nf = n2 * S;
  const size_t eo = (size_t)q.e0, so = (size_t)q.s0;
  std::vector<int> e2s(12 * (size_t)E), s2e(5 * (size_t)S), ep(E, q.r);
  for (size_t i = 0; i < 12 * (size_t)E; i += 2) { e2s[i] = r.mesh.ElemToSide[i] + q.s0; e2s[i + 1] = r.mesh.ElemToSide[i + 1]; }
  for (size_t s = 0; s < (size_t)S; s++) {
    const int *t = r.mesh.SideToElem + 5 * s;
    s2e[5 * s] = t[0] + q.e0; s2e[5 * s + 1] = t[1] + q.e0;
    s2e[5 * s + 2] = t[2]; s2e[5 * s + 3] = t[3]; s2e[5 * s + 4] = t[4];
  }
  std::vector<double4> nvse(nf);
  for (size_t f = 0; f < nf; f++)
    nvse[f] = make_double4(r.mesh.NormVec[3 * f], r.mesh.NormVec[3 * f + 1], r.mesh.NormVec[3 * f + 2],
                           r.mesh.SurfElem[f]);
  std::vector<float> m32;
  HIPCHK(hipMemcpyAsync(b.U + 5 * n3 * eo, r.U0, 5 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Mf + 3 * n3 * eo, r.mesh.Metrics_fTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Mg + 3 * n3 * eo, r.mesh.Metrics_gTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Mh + 3 * n3 * eo, r.mesh.Metrics_hTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.sJ + n3 * eo, r.mesh.sJ, n3E * 8, hipMemcpyHostToDevice, b.stream));
  if (b.cfg.precision == FX_PREC_FP32_SPLIT) {
    m32.resize(9 * n3E);
    const double *src[3] = {r.mesh.Metrics_fTilde, r.mesh.Metrics_gTilde, r.mesh.Metrics_hTilde};
    float *dst[3] = {b.Mf32, b.Mg32, b.Mh32};
    for (int d = 0; d < 3; d++) {
      for (size_t i = 0; i < 3 * n3E; i++) m32[d * 3 * n3E + i] = (float)src[d][i];
      HIPCHK(hipMemcpyAsync(dst[d] + 3 * n3 * eo, m32.data() + d * 3 * n3E, 3 * n3E * 4, hipMemcpyHostToDevice,
                            b.stream));
    }
  }
  HIPCHK(hipMemcpyAsync(b.nvse + n2 * so, nvse.data(), nf * sizeof(double4), hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.e2s + 12 * eo, e2s.data(), 12 * (size_t)E * 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.s2e + 5 * so, s2e.data(), 5 * (size_t)S * 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.eprob + eo, ep.data(), (size_t)E * 4, hipMemcpyHostToDevice, b.stream));
  FX_DISPATCH_N(p.N, launch_face_table_n, b, q.e0, E);
  HIPCHK(hipGetLastError());
  DtState z{};
  z.errType = 0; z.errStep = -1; z.dt = 0.0; z.t = 0.0; z.step = 0;
  z.glm_init = 0; z.glm_ch = 0.5; z.glm_dtch1 = 0.0;
  z.tEnd = 1.0e300; z.done = 0; z.pad = 0;
  HIPCHK(hipMemcpyAsync(b.st + q.r, &z, sizeof(z), hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.rfirst + q.r, &q.e0, 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.rcount + q.r, &E, 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipStreamSynchronize(b.stream));
  return FX_OK;
}

// Mark a request's slots unused on the device (kernels skip them) and free them.
inline int release_request(Pool &p, Request &q)
{
  Batch &b = p.b;
  const int E = q.req.nElems, S = q.req.nSides, zero = 0;
  std::vector<int> ep(E, -1), s2e(5 * (size_t)S, 0);
  HIPCHK(hipMemcpyAsync(b.eprob + q.e0, ep.data(), (size_t)E * 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.s2e + 5 * (size_t)q.s0, s2e.data(), 5 * (size_t)S * 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.rcount + q.r, &zero, 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipStreamSynchronize(b.stream));
  p.ea.release(q.e0, E);
  p.sa.release(q.s0, S);
  p.ra.release(q.r, 1);
  p.nActive--;
  p.active_ids.erase(std::find(p.active_ids.begin(), p.active_ids.end(), q.id));
  return FX_OK;
}

// Place a request in its pool (growing the pool if needed) and upload it.  Returns FX_OK,
// or a negative code (the request fails).
inline int admit(fx_server *s, Request &q)
{
  const int N = q.req.N, par = q.req.parabolic ? 1 : 0;
  Pool &p = s->pools[2 * N + par];
  int rc;
  if (!p.b.stream && (rc = pool_init(p, s->cfg, N, par)) != FX_OK) return rc;
  const int E = q.req.nElems, S = q.req.nSides;
  int e0 = p.ea.alloc(E), s0 = p.sa.alloc(S), r = p.ra.alloc(1);
  if (e0 < 0 || s0 < 0 || r < 0) {
    if (e0 >= 0) p.ea.release(e0, E);
    if (s0 >= 0) p.sa.release(s0, S);
    if (r >= 0) p.ra.release(r, 1);
    const int Ecap = std::max(2 * p.ea.cap, p.ea.cap + E + 1024);
    const int Scap = std::max(2 * p.sa.cap, p.sa.cap + S + 3072);
    const int Rcap = std::max(2 * p.ra.cap, p.ra.cap + 16);
    if ((rc = pool_realloc(p, s->cfg, Ecap, Scap, Rcap)) != FX_OK) return rc;
    e0 = p.ea.alloc(E); s0 = p.sa.alloc(S); r = p.ra.alloc(1);
    if (e0 < 0 || s0 < 0 || r < 0) return set_err(FX_ENOMEM, "pool slot allocation failed");
  }
  q.pool = 2 * N + par; q.e0 = e0; q.s0 = s0; q.r = r;
  if ((rc = upload_request(p, q)) != FX_OK) {
    p.ea.release(e0, E); p.sa.release(s0, S); p.ra.release(r, 1);
    return rc;
  }
  p.nActive++;
  p.active_ids.push_back(q.id);
  return FX_OK;
}

// Step graph of a pool over its current extents: CALCTIMESTEP (per problem) + 5 stages.  When
// the extents or arrays change, the step is recorded again and the pool's executable graph is
// updated in place (same kernels in the same order, only grid sizes and arguments differ);
// instantiation, the costly part, happens only for a pool's first graph or if an update is
// refused.
inline int pool_graph(fx_server *s, Pool &p)
{
  const int E = p.ea.high_water(), S = p.sa.high_water(), R = p.ra.high_water();
  if (p.graph && !p.ptrs_changed && E == p.gE && S == p.gS && R == p.gR) return FX_OK;
  Batch &b = p.b;
  b.cfg.nElems = E; b.cfg.nSides = S; b.nProb = R;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
  q_dt(b, cfl_number(p.N, s->cfg.CFLscale), p.parabolic ? dfl_number(p.N, s->cfg.DFLscale) : 0.0, 1);
  q_rk_step(b);
  HIPCHK(hipStreamEndCapture(b.stream, &g));
  hipError_t e = hipSuccess;
  bool updated = false;
  if (p.graph) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    e = hipGraphExecUpdate(p.graph, g, &bad, &res);
    if (e == hipSuccess && res == hipGraphExecUpdateSuccess) {
      updated = true;
    } else {
      (void)hipGetLastError();  // a refused update leaves no error behind
      (void)hipGraphExecDestroy(p.graph);
      p.graph = nullptr;
    }
  }
  if (!updated) e = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { p.graph = nullptr; return set_err(FX_EHIP, "graph instantiate: %s", hipGetErrorString(e)); }
  p.gE = E; p.gS = S; p.gR = R;
  p.ptrs_changed = false;
  if (updated) s->graph_updates++;
  else s->graph_captures++;
  return FX_OK;
}

// completion of an admitted request (called with the server lock held)
inline long long request_nodes(const Request &q)
{
  const long long n = q.req.N + 1;
  return n * n * n * q.req.nElems;
}

inline void finish(fx_server *s, Request &q, int state)
{
  q.state = state;
  q.t_done = seconds_since(s);
  if (state == FX_REQ_DONE) s->nDone++; else s->nFailed++;
  s->bytes_in_use -= q.bytes;
  s->nodes_active -= request_nodes(q);
  s->nActive--;
}

// Bookkeeping after a pool's step has completed (called with the server lock held; the lock
// is released around device copies).  Requests that reached nSteps or failed are finished.
inline void complete_step(fx_server *s, Pool &p, std::unique_lock<std::mutex> &lk)
{
  s->steps_launched++;
  for (int id : std::vector<int>(p.active_ids)) {
    Request &q = s->reqs[id];
    const DtState st = p.st_host[q.r];
    q.steps++;
    q.errType = st.errType;
    q.t_end = st.t;
    q.dt_last = st.dt;
    if (q.steps < q.req.nSteps && st.errType == 0) continue;
    lk.unlock();
    int rc2 = FX_OK;
    if (st.errType == 0 && q.req.U_out) {
      hipError_t e = hipMemcpyAsync(q.req.U_out, p.b.U + (size_t)5 * p.b.n3 * q.e0,
                                    (size_t)5 * p.b.n3 * q.req.nElems * 8, hipMemcpyDeviceToHost, p.b.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(p.b.stream);
      if (e != hipSuccess) rc2 = set_err(FX_EHIP, "result copy: %s", hipGetErrorString(e));
    }
    int rc3 = release_request(p, q);
    lk.lock();
    finish(s, q, (st.errType == 0 && rc2 == FX_OK && rc3 == FX_OK) ? FX_REQ_DONE : FX_REQ_FAILED);
  }
}

// A device failure fails every running request; the pools are drained and left as they are.
inline void fail_running(fx_server *s, std::unique_lock<std::mutex> &lk)
{
  snprintf(s->worker_err, sizeof(s->worker_err), "%s", err_buf());
  for (auto &p : s->pools) {
    if (p.b.stream) (void)hipStreamSynchronize(p.b.stream);
    p.inflight = false;
    for (int id : std::vector<int>(p.active_ids)) {
      Request &q = s->reqs[id];
      lk.unlock();
      (void)release_request(p, q);
      lk.lock();
      finish(s, q, FX_REQ_FAILED);
    }
  }
  s->cv_done.notify_all();
}

// The scheduler.  Every pool (one per polynomial degree and equation) steps on its own stream
// as soon as its previous step has completed, so a pool of small requests is not held back by
// a pool of large ones.  Admission into a pool, request completion and graph re-capture
// happen only between that pool's steps.
inline void worker_loop(fx_server *s)
{
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    int rc = FX_OK;
    for (auto &p : s->pools) {
      if (!p.inflight) continue;
      hipError_t e = hipEventQuery(p.done);
      if (e == hipErrorNotReady) continue;
      p.inflight = false;
      progressed = true;
      if (e != hipSuccess) { rc = set_err(FX_EHIP, "step: %s", hipGetErrorString(e)); break; }
      complete_step(s, p, lk);
    }
    if (rc != FX_OK) { fail_running(s, lk); continue; }
    if (progressed) s->cv_done.notify_all();
    // admission in FIFO order; a request that does not fit the budget, or whose pool is
    // between launch and completion of a step, waits; later ones may pass it.  With an
    // active-work cap (admitKNodes), a request beyond the cap waits and no later one passes
    // it (strict FIFO at the cap, so a large request is not starved by small ones; it then
    // runs with a bounded number of others).  Submissions may arrive while the lock is
    // released.
    const std::vector<int> waiting(s->queue.begin(), s->queue.end());
    const long long cap = 1000LL * s->cfg.admitKNodes;
    for (int qid : waiting) {
      Request &q = s->reqs[qid];
      if (cap > 0 && s->nodes_active > 0 && s->nodes_active + request_nodes(q) > cap) break;
      if (s->bytes_in_use + q.bytes > s->cfg.hbmBudget) continue;
      if (s->pools[2 * q.req.N + (q.req.parabolic ? 1 : 0)].inflight) continue;
      lk.unlock();
      int rc2 = admit(s, q);
      lk.lock();
      s->queue.erase(std::find(s->queue.begin(), s->queue.end(), qid));
      long long a = 0;
      for (auto &p : s->pools) a += p.bytes_alloc;
      s->bytes_allocated = a;
      if (rc2 != FX_OK) {
        q.state = FX_REQ_FAILED; q.t_done = seconds_since(s); s->nFailed++;
        snprintf(s->worker_err, sizeof(s->worker_err), "%s", err_buf());
        s->cv_done.notify_all();
        continue;
      }
      q.state = FX_REQ_RUNNING;
      q.t_admit = seconds_since(s);
      s->nActive++;
      s->nodes_active += request_nodes(q);
      s->bytes_in_use += q.bytes;
      s->bytes_peak = std::max(s->bytes_peak, s->bytes_in_use);
    }
    if (s->nActive == 0) {  // no pool holds a request, so none is in flight
      if (s->stop) break;
      s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      continue;
    }
    if (s->stop) {
      for (auto &p : s->pools)
        if (p.inflight) { (void)hipStreamSynchronize(p.b.stream); p.inflight = false; }
      break;
    }
    lk.unlock();
    // next step of every pool that holds requests and has no step outstanding
    bool launched = false;
    for (auto &p : s->pools) {
      if (p.nActive == 0 || p.inflight) continue;
      if ((rc = pool_graph(s, p)) != FX_OK) break;
      hipError_t e = hipGraphLaunch(p.graph, p.b.stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(p.st_host, p.b.st, sizeof(DtState) * p.gR, hipMemcpyDeviceToHost, p.b.stream);
      if (e == hipSuccess) e = hipEventRecord(p.done, p.b.stream);
      if (e != hipSuccess) { rc = set_err(FX_EHIP, "step launch: %s", hipGetErrorString(e)); break; }
      p.inflight = true;
      launched = true;
    }
    // nothing to do until a step completes: poll at a period well below one step
    if (rc == FX_OK && !progressed && !launched) std::this_thread::sleep_for(std::chrono::microseconds(10));
    lk.lock();
    if (rc != FX_OK) fail_running(s, lk);
  }
  // stopping: whatever has not completed fails
  for (auto &q : s->reqs)
    if (q.state == FX_REQ_QUEUED || q.state == FX_REQ_RUNNING) {
      q.state = FX_REQ_FAILED;
      q.t_done = seconds_since(s);
      s->nFailed++;
    }
  s->queue.clear();
  s->cv_done.notify_all();
}

}  // namespace srv
}  // namespace fx

extern "C" {

long long fx_request_bytes(const fx_server_config *cfg, int N, int parabolic, int nElems, int nSides)
{
  if (!cfg) return -1;
  return fx::srv::request_bytes(*cfg, N, parabolic, nElems, nSides);
}

int fx_server_create(const fx_server_config *cfg, fx_server **out)
{
  using namespace fx;
  if (!cfg || !out) return set_err(FX_EINVAL, "null argument");
  *out = nullptr;
  int rc = check_physics(cfg->volumeFlux, cfg->riemann, cfg->precision, cfg->kappa, cfg->Pr, 1);
  if (rc) return rc;
  if (cfg->lifting != 1) return set_err(FX_EINVAL, "only BR1 lifting (lifting=1)");
  if (cfg->hbmBudget <= 0) return set_err(FX_EINVAL, "hbmBudget must be > 0");
  if (cfg->maxRequests < 1) return set_err(FX_EINVAL, "maxRequests must be >= 1");
  if (cfg->admitKNodes < 0) return set_err(FX_EINVAL, "admitKNodes must be >= 0");
  if (!(cfg->CFLscale > 0.0)) return set_err(FX_EINVAL, "CFLscale must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(FX_ENODEV, "no HIP device");
  fx_server *s = new fx_server();
  s->cfg = *cfg;
  s->t0 = srv::Clock::now();
  s->reqs.reserve(cfg->maxRequests);
  s->worker = std::thread(srv::worker_loop, s);
  *out = s;
  return FX_OK;
}

int fx_server_submit(fx_server *s, const fx_request *r, int *id)
{
  using namespace fx;
  if (!s || !r || !id) return set_err(FX_EINVAL, "null argument");
  *id = -1;
  if (r->N < 2 || r->N > 5) return set_err(FX_EINVAL, "N=%d not supported (2..5)", r->N);
  if (r->nElems < 1 || r->nSides < 1 || r->nSteps < 1) return set_err(FX_EINVAL, "empty request");
  const fx_mesh &m = r->mesh;
  if (!m.ElemToSide || !m.SideToElem || !m.Metrics_fTilde || !m.Metrics_gTilde || !m.Metrics_hTilde || !m.sJ ||
      !m.NormVec || !m.SurfElem || !r->U0)
    return set_err(FX_EINVAL, "missing request array");
  int rc = check_connectivity(m.ElemToSide, m.SideToElem, r->nElems, r->nSides);
  if (rc) return rc;
  const long long bytes = srv::request_bytes(s->cfg, r->N, r->parabolic ? 1 : 0, r->nElems, r->nSides);
  if (bytes > s->cfg.hbmBudget)
    return set_err(FX_ENOMEM, "request needs %lld bytes, budget is %lld", bytes, s->cfg.hbmBudget);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_err(FX_ESTATE, "server stopping");
    if ((int)s->reqs.size() >= s->cfg.maxRequests) return set_err(FX_ENOMEM, "maxRequests reached");
    srv::Request q;
    q.req = *r;
    q.id = (int)s->reqs.size();
    q.bytes = bytes;
    q.t_submit = srv::seconds_since(s);
    s->reqs.push_back(q);
    s->queue.push_back(q.id);
    *id = q.id;
  }
  s->cv_work.notify_one();
  return FX_OK;
}

int fx_server_status(fx_server *s, int id, fx_request_status *st)
{
  using namespace fx;
  if (!s || !st) return set_err(FX_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int)s->reqs.size()) return set_err(FX_EINVAL, "unknown request %d", id);
  const srv::Request &q = s->reqs[id];
  st->state = q.state; st->errType = q.errType; st->steps = q.steps; st->reserved = 0;
  st->t_end = q.t_end; st->dt_last = q.dt_last;
  st->t_submit = q.t_submit; st->t_admit = q.t_admit; st->t_done = q.t_done;
  return FX_OK;
}

int fx_server_wait(fx_server *s, int id, double timeout_s)
{
  using namespace fx;
  if (!s) return set_err(FX_EINVAL, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (id >= (int)s->reqs.size()) return set_err(FX_EINVAL, "unknown request %d", id);
  auto done = [&] {
    if (id >= 0) return s->reqs[id].state >= FX_REQ_DONE;
    return s->nDone + s->nFailed == (int)s->reqs.size();
  };
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return set_err(FX_ESTATE, "wait timed out");
  return FX_OK;
}

int fx_server_get_stats(fx_server *s, fx_server_stats *o)
{
  using namespace fx;
  if (!s || !o) return set_err(FX_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  o->steps_launched = s->steps_launched;
  o->graph_captures = s->graph_captures;
  o->graph_updates = s->graph_updates;
  o->bytes_in_use = s->bytes_in_use;
  o->bytes_peak = s->bytes_peak;
  o->bytes_allocated = s->bytes_allocated;
  o->active = s->nActive;
  o->queued = (int)s->queue.size();
  o->done = s->nDone;
  o->failed = s->nFailed;
  return FX_OK;
}

const char *fx_server_last_error(fx_server *s)
{
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  return s->worker_err;
}

void fx_server_destroy(fx_server *s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto &p : s->pools) fx::srv::pool_free(p);
  delete s;
}

}  // extern "C"
