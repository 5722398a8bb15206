// MI355X DGSEM right-hand-side engine: the single-problem context and the C ABI declared
// in include/fluxo_gpu.h.  This file is the library's one device-code unit: the launch
// layer (fx_launch.hpp) and the batched server (fx_server.hpp) are compiled with it.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

// The non-template kernels (CFL reductions, NFVSE propagation, the no-allocation helpers) and
// their launches are compiled here; the per-degree kernels are in fx_tu.hip units.
#define FX_KERNEL_HOME 1
#include "fx_kernels.hpp"
#include "fx_launch.hpp"
#include "fx_analyze.hpp"

namespace fx {

// Load-time registry of the units' constant-memory uploaders (see ensure_ops).
static std::vector<OpsUploader> &ops_uploaders()
{
  static std::vector<OpsUploader> v;
  return v;
}
int register_ops_uploader(OpsUploader f)
{
  ops_uploaders().push_back(f);
  return 1;
}
int upload_ops_everywhere(const Ops &o, int N)
{
  if (ops_uploaders().empty()) return set_err(FX_EHIP, "no kernel units registered");
  for (OpsUploader f : ops_uploaders()) {
    const int rc = f(o, N);
    if (rc) return rc;
  }
  return FX_OK;
}

}  // namespace fx

using namespace fx;

struct fx_ctx {
  Batch b;                       // nElems/nSides of the mesh, eprob == nullptr
  double *part = nullptr, *red = nullptr;
  double *xgp = nullptr;                                   // Elem_xGP (fx_elem_xgp_upload)
  double *ana_part = nullptr, *ana_ops = nullptr, *ana_out = nullptr;  // fx_analyze, on first use
  double *api_fs = nullptr;                                // Flux_slave of fx_surfint (nVar 5), on first use
  int ana_na = -1;
  bool mesh_ok = false, state_ok = false;
  hipGraphExec_t graph = nullptr;
  double graph_cfl = -1.0, graph_dfl = -1.0;
};

namespace {

size_t vol_doubles(const fx_config &c) { size_t n = c.N + 1; return (size_t)c.nVar * n * n * n * c.nElems; }

void q_entropy(fx_ctx *c) { FX_DISPATCH_N(c->b.cfg.N, launch_entropy_n, c->b, c->part, c->red); }

__global__ void k_set_dt(DtState *st, double dt, double t, int resetErr, double tEnd)
{
  st->dt = dt;
  st->t = t;
  st->tEnd = tEnd;
  st->done = 0;
  if (resetErr) { st->errType = 0; st->errStep = -1; st->step = 0; }
}

constexpr double NO_END_TIME = 1.0e300;

int check_config(const fx_config *c)
{
  if (!c) return set_err(FX_EINVAL, "null config");
  if (c->N < 2 || c->N > 5) return set_err(FX_EINVAL, "N=%d not supported (2..5)", c->N);
  if (c->nVar != 5 && c->nVar != 9) return set_err(FX_EINVAL, "nVar=%d not supported (5, 9 GLM-MHD)", c->nVar);
  if (!face_table_fits(c->N, c->nElems, c->nSides))
    return set_err(FX_EINVAL, "%d elements / %d sides at N=%d: volume nodes or side points beyond 32-bit indices",
                   c->nElems, c->nSides, c->N);
  if (c->nVar == 9) {
    if (c->volumeFlux != 10)
      return set_err(FX_EINVAL, "GLM-MHD: VolumeFlux=10 (Derigs) only");
    if (c->riemann != 0 && c->riemann != 1 && c->riemann != 10 && c->riemann != 11)
      return set_err(FX_EINVAL, "GLM-MHD: Riemann=%d not supported (0, 1, 10, 11)", c->riemann);
    if (c->nfvse) return set_err(FX_EINVAL, "GLM-MHD: no shock capturing");
    if (!(c->kappa > 1.0) || !(c->mu_0 > 0.0) || !(c->GLM_scale > 0.0))
      return set_err(FX_EINVAL, "GLM-MHD: kappa > 1, mu_0 > 0, GLM_scale > 0");
    if (c->parabolic && ((c->lifting != 1 && c->lifting != 2) || !(c->Pr > 0.0) || !(c->mu0 >= 0.0) ||
                         !(c->eta >= 0.0)))
      return set_err(FX_EINVAL, "resistive GLM-MHD: lifting 1 (BR1) or 2 (BR2), Pr > 0, mu >= 0, eta >= 0");
    if (c->parabolic && c->lifting == 2 && !(c->etaBR2 > 0.0)) return set_err(FX_EINVAL, "BR2 needs etaBR2 > 0");
    if (c->precision != FX_PREC_FP64 && c->precision != FX_PREC_FP32_SPLIT)
      return set_err(FX_EINVAL, "precision %d not supported", c->precision);
    return FX_OK;
  }
  if (c->nElems < 1 || c->nSides < 1) return set_err(FX_EINVAL, "empty mesh");
  if (c->parabolic && c->lifting != 1 && c->lifting != 2)
    return set_err(FX_EINVAL, "lifting=%d not supported (1 BR1, 2 BR2)", c->lifting);
  if (c->parabolic && c->lifting == 2 && !(c->etaBR2 > 0.0)) return set_err(FX_EINVAL, "BR2 needs etaBR2 > 0");
  const int rc = check_physics(c->volumeFlux, c->riemann, c->precision, c->kappa, c->Pr, c->parabolic);
  return rc ? rc : check_nfvse(c);
}

}  // namespace

extern "C" {

const char *fx_last_error(void) { return err_buf(); }

long long fx_ctx_bytes(const fx_config *cfg)
{
  if (!cfg) return -1;
  const long long n = cfg->N + 1, n3 = n * n * n, E = cfg->nElems, S = cfg->nSides;
  long long b = 0;
  b += 3LL * 5 * n3 * E * 8;         // U, Ut, Ut_tmp
  b += 9LL * n3 * E * 8 + n3 * E * 8; // metrics, sJ
  if (cfg->precision == FX_PREC_FP32_SPLIT) b += 9LL * n3 * E * 4;  // FP32 metrics
  b += 5LL * n * n * S * 8;          // Flux
  b += 4LL * n * n * S * 8;          // nvse
  b += (12LL + 1) * E * 4 + 5LL * S * 4;
  b += 6LL * n * n * E * 8;          // face-point table
  if (cfg->parabolic) {  // Uv, FvN, gradients
    const long long nvv = cfg->nVar == 9 ? 7 : 4;
    b += nvv * n3 * E * 8 + 2 * nvv * n * n * S * 8 + 3LL * cfg->nVar * n3 * E * 8;
  }
  if (cfg->nfvse) b += 5LL * n3 * E * 8 + 2LL * E * 8;  // Ffv, alpha buffers
  if (cfg->nVar == 9) b += 4LL * 3 * n3 * E * 8 + (9LL + 4) * n * n * S * 8 * 2 + E * 8;  // MHD extra
  b += 2LL * ((n3 * E + 255) / 256 + 1) * 8 + 64;
  return b;
}

int fx_ctx_create(const fx_config *cfg, fx_ctx **out)
{
  if (!out) return set_err(FX_EINVAL, "null out");
  *out = nullptr;
  int rc = check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(FX_ENODEV, "no HIP device");
  rc = ensure_ops(cfg->N);
  if (rc) return rc;
  fx_ctx *c = new fx_ctx();
  Batch &b = c->b;
  b.cfg = *cfg;
  b.set_degree(cfg->N);
  b.vol_tile = tuned_tile(cfg->N, cfg->parabolic, cfg->precision);
  const size_t nv = vol_doubles(*cfg), n3E = (size_t)b.n3 * cfg->nElems;
  const size_t nf = (size_t)b.n2 * cfg->nSides;
  auto fail = [&](hipError_t e, const char *what) {
    set_err(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP, "%s: %s", what, hipGetErrorString(e));
    fx_ctx_destroy(c);
    return e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
#define ALLOC(p, bytes)                                               \
  if ((e = hipMalloc((void **)&(p), (bytes))) != hipSuccess) return fail(e, "hipMalloc " #p);
  ALLOC(b.U, nv * 8);
  ALLOC(b.Ut, nv * 8);
  ALLOC(b.Uttmp, nv * 8);
  ALLOC(b.Mf, 3 * n3E * 8);
  ALLOC(b.Mg, 3 * n3E * 8);
  ALLOC(b.Mh, 3 * n3E * 8);
  ALLOC(b.sJ, n3E * 8);
  if (cfg->precision == FX_PREC_FP32_SPLIT) {
    ALLOC(b.Mf32, 3 * n3E * 4);
    ALLOC(b.Mg32, 3 * n3E * 4);
    ALLOC(b.Mh32, 3 * n3E * 4);
  }
  ALLOC(b.Flux, (size_t)cfg->nVar * nf * 8);
  if (cfg->nVar == 9) {
    ALLOC(b.FluxS, (size_t)cfg->nVar * nf * 8);
    ALLOC(b.dtch1Elem, (size_t)cfg->nElems * 8);
  }
  ALLOC(b.nvse, nf * sizeof(double4));
  ALLOC(b.e2s, (size_t)12 * cfg->nElems * 4);
  ALLOC(b.s2e, (size_t)5 * cfg->nSides * 4);
  ALLOC(b.ftab, (size_t)6 * b.n2 * cfg->nElems * sizeof(int2));
  if (cfg->parabolic) {
    const size_t nvv = cfg->nVar == 9 ? 7 : 4;  // nonzero viscous components
    ALLOC(b.Uv, nvv * n3E * 8);
    ALLOC(b.FvN, 2 * nvv * nf * 8);
    ALLOC(b.grad, 3 * nv * 8);
  }
  if (cfg->nfvse) {
    ALLOC(b.alpha, (size_t)cfg->nElems * 8);
    ALLOC(b.alpha2, (size_t)cfg->nElems * 8);
    ALLOC(b.Ffv, nv * 8);
  }
  const size_t npart = (n3E + 255) / 256 + 1;
  ALLOC(c->part, 2 * npart * 8);
  ALLOC(b.dtElem, (size_t)cfg->nElems * 8);
  ALLOC(b.errElem, (size_t)cfg->nElems * 4);
  ALLOC(c->red, 2 * 8);
  ALLOC(b.st, sizeof(DtState));
#undef ALLOC
  DtState z{};
  z.errType = 0; z.errStep = -1; z.dt = 0.0; z.t = 0.0; z.step = 0;
  z.glm_init = 0; z.glm_ch = 0.5; z.glm_dtch1 = 0.0;
  z.tEnd = 1.0e300; z.done = 0; z.pad = 0;
  if ((e = hipMemcpy(b.st, &z, sizeof(z), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "init");
  *out = c;
  return FX_OK;
}

void fx_ctx_destroy(fx_ctx *c)
{
  if (!c) return;
  Batch &b = c->b;
  if (b.stream) (void)hipStreamSynchronize(b.stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  void *ptrs[] = {b.U, b.Ut, b.Uttmp, b.Mf, b.Mg, b.Mh, b.Mf32, b.Mg32, b.Mh32, b.sJ, b.Flux, b.Uv, b.FvN, b.grad,
                  c->part, c->red, b.nvse, b.e2s, b.s2e, b.ftab, b.st, b.dtElem, b.errElem, b.alpha, b.alpha2,
                  b.Ffv, b.FluxS, b.dtch1Elem, c->xgp, c->ana_part, c->ana_ops, c->ana_out, c->api_fs,
                  b.mort, b.mside, b.mops, b.Umort, b.LiftS, b.LiftM, b.GradM};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b.stream) (void)hipStreamDestroy(b.stream);
  delete c;
}

int fx_mesh_upload(fx_ctx *c, const fx_mesh *m)
{
  if (!c || !m) return set_err(FX_EINVAL, "null argument");
  if (!m->ElemToSide || !m->SideToElem || !m->Metrics_fTilde || !m->Metrics_gTilde ||
      !m->Metrics_hTilde || !m->sJ || !m->NormVec || !m->SurfElem)
    return set_err(FX_EINVAL, "missing mesh array");
  Batch &b = c->b;
  const int E = b.cfg.nElems, S = b.cfg.nSides, nm = m->nMortarSides;
  std::vector<int> mort;
  int rc = check_connectivity(m->ElemToSide, m->SideToElem, E, S, nm, m->MortarType, m->MortarInfo, &mort);
  if (rc) return rc;
  if (nm > 0 && b.cfg.whichMortar != 0 && b.cfg.whichMortar != 1)
    return set_err(FX_EINVAL, "whichMortar=%d (0 projection, 1 collocation)", b.cfg.whichMortar);
  const size_t n3E = (size_t)b.n3 * E, nf = (size_t)b.n2 * S;
  if (nm > 0) {
    HIPCHK(hipStreamSynchronize(b.stream));
    if (b.nMortar != nm) {  // the mortar buffers are sized by the mesh
      for (void *p : {(void *)b.mort, (void *)b.mside, (void *)b.mops, (void *)b.Umort, (void *)b.LiftS,
                      (void *)b.LiftM, (void *)b.GradM})
        if (p) HIPCHK(hipFree(p));
      b.mort = nullptr; b.mside = nullptr; b.mops = nullptr; b.Umort = nullptr; b.nMortar = 0;
      b.LiftS = nullptr; b.LiftM = nullptr; b.GradM = nullptr;
      HIPCHK(hipMalloc((void **)&b.mort, mort.size() * sizeof(int)));
      HIPCHK(hipMalloc((void **)&b.mside, (size_t)S * sizeof(int)));
      HIPCHK(hipMalloc((void **)&b.mops, 4 * (size_t)b.n2 * 8));
      const size_t nv = (size_t)b.cfg.nVar;
      HIPCHK(hipMalloc((void **)&b.Umort, nv * nf * 8));
      if (b.cfg.parabolic) {
        HIPCHK(hipMalloc((void **)&b.LiftS, 3 * nv * nf * 8));
        HIPCHK(hipMalloc((void **)&b.LiftM, 3 * nv * nm * b.n2 * 8));
        HIPCHK(hipMalloc((void **)&b.GradM, 3 * nv * nm * b.n2 * 8));
      }
    }
    double mops[4 * (MAXN + 1) * (MAXN + 1)];
    make_mortar_ops(b.cfg.N, b.cfg.whichMortar, mops);
    std::vector<int> mside(S, -1);
    for (int m = 0; m < nm; m++)
      for (int im = 0; im < (mort[MORT_INTS * m] == 1 ? 4 : 2); im++) mside[mort[MORT_INTS * m + 3 + im] - 1] = m;
    HIPCHK(hipMemcpy(b.mort, mort.data(), mort.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.mside, mside.data(), (size_t)S * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.mops, mops, 4 * (size_t)b.n2 * 8, hipMemcpyHostToDevice));
    b.nMortar = nm;
  } else {
    b.nMortar = 0;
  }
  if (c->graph) {  // a step graph captured for another mesh's launches
    HIPCHK(hipStreamSynchronize(b.stream));
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  std::vector<double4> nvse(nf);
  for (size_t f = 0; f < nf; f++)
    nvse[f] = make_double4(m->NormVec[3 * f], m->NormVec[3 * f + 1], m->NormVec[3 * f + 2], m->SurfElem[f]);
  HIPCHK(hipMemcpyAsync(b.Mf, m->Metrics_fTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Mg, m->Metrics_gTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Mh, m->Metrics_hTilde, 3 * n3E * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.sJ, m->sJ, n3E * 8, hipMemcpyHostToDevice, b.stream));
  std::vector<float> m32;
  if (b.cfg.precision == FX_PREC_FP32_SPLIT) {
    m32.resize(9 * n3E);
    const double *src[3] = {m->Metrics_fTilde, m->Metrics_gTilde, m->Metrics_hTilde};
    float *dst[3] = {b.Mf32, b.Mg32, b.Mh32};
    for (int d = 0; d < 3; d++) {
      for (size_t i = 0; i < 3 * n3E; i++) m32[d * 3 * n3E + i] = (float)src[d][i];
      HIPCHK(hipMemcpyAsync(dst[d], m32.data() + d * 3 * n3E, 3 * n3E * 4, hipMemcpyHostToDevice, b.stream));
    }
  }
  HIPCHK(hipMemcpyAsync(b.nvse, nvse.data(), nf * sizeof(double4), hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.e2s, m->ElemToSide, (size_t)12 * E * 4, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.s2e, m->SideToElem, (size_t)5 * S * 4, hipMemcpyHostToDevice, b.stream));
  FX_DISPATCH_N(b.cfg.N, launch_face_table_n, b, 0, E);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b.stream));
  c->mesh_ok = true;
  return FX_OK;
}

int fx_state_upload(fx_ctx *c, const double *U)
{
  if (!c || !U) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(c->b.U, U, vol_doubles(c->b.cfg) * 8, hipMemcpyHostToDevice, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  c->state_ok = true;
  return FX_OK;
}

int fx_state_download(fx_ctx *c, double *U)
{
  if (!c || !U) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(U, c->b.U, vol_doubles(c->b.cfg) * 8, hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

int fx_ut_download(fx_ctx *c, double *Ut)
{
  if (!c || !Ut) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(Ut, c->b.Ut, vol_doubles(c->b.cfg) * 8, hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

static int ready(fx_ctx *c)
{
  if (!c) return set_err(FX_EINVAL, "null context");
  if (!c->mesh_ok) return set_err(FX_ESTATE, "mesh not uploaded");
  if (!c->state_ok) return set_err(FX_ESTATE, "state not uploaded");
  return FX_OK;
}

int fx_rhs(fx_ctx *c, double t, double *Ut_host)
{
  (void)t;
  int rc = ready(c);
  if (rc) return rc;
  RKArgs rk{};
  q_stage(c->b, VMODE_RHS, rk);
  HIPCHK(hipGetLastError());
  if (Ut_host) return fx_ut_download(c, Ut_host);
  return FX_OK;
}

int fx_volint(fx_ctx *c, double *Ut_host)
{
  int rc = ready(c);
  if (rc) return rc;
  RKArgs rk{};
  q_volume(c->b, VMODE_VOLINT, rk);
  HIPCHK(hipGetLastError());
  if (Ut_host) return fx_ut_download(c, Ut_host);
  return FX_OK;
}

int fx_lifting(fx_ctx *c, double *gx, double *gy, double *gz)
{
  int rc = ready(c);
  if (rc) return rc;
  if (!c->b.cfg.parabolic) return set_err(FX_ESTATE, "lifting needs parabolic=1");
  q_lifting(c->b, 0);
  HIPCHK(hipGetLastError());
  const size_t nv = vol_doubles(c->b.cfg);
  double *dst[3] = {gx, gy, gz};
  for (int d = 0; d < 3; d++)
    if (dst[d]) HIPCHK(hipMemcpyAsync(dst[d], c->b.grad + d * nv, nv * 8, hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

int fx_fillflux(fx_ctx *c, double *Flux_master, double *Flux_slave)
{
  int rc = ready(c);
  if (rc) return rc;
  Batch &b = c->b;
  if (b.cfg.parabolic) q_lifting(b, 1);  // the sides' face-normal viscous fluxes
  q_surface(b);
  HIPCHK(hipGetLastError());
  const size_t nf = (size_t)b.cfg.nVar * b.n2 * b.cfg.nSides;
  if (Flux_master) HIPCHK(hipMemcpyAsync(Flux_master, b.Flux, nf * 8, hipMemcpyDeviceToHost, b.stream));
  if (Flux_slave)  // conforming sides: Flux_slave = Flux_master, plus the slave's Powell term for MHD
    HIPCHK(hipMemcpyAsync(Flux_slave, b.mhd() ? b.FluxS : b.Flux, nf * 8, hipMemcpyDeviceToHost, b.stream));
  HIPCHK(hipStreamSynchronize(b.stream));
  return FX_OK;
}

int fx_surfint(fx_ctx *c, const double *Flux_master, const double *Flux_slave, double *Ut)
{
  if (!c || !Flux_master || !Flux_slave || !Ut) return set_err(FX_EINVAL, "null argument");
  if (!c->mesh_ok) return set_err(FX_ESTATE, "mesh not uploaded");
  Batch &b = c->b;
  const size_t nf = (size_t)b.cfg.nVar * b.n2 * b.cfg.nSides;
  double *fs = b.FluxS;
  if (!fs) {
    if (!c->api_fs) HIPCHK(hipMalloc((void **)&c->api_fs, nf * 8));
    fs = c->api_fs;
  }
  // the device face-flux buffers are scratch of the right-hand side: they take the caller's fluxes
  HIPCHK(hipMemcpyAsync(b.Flux, Flux_master, nf * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(fs, Flux_slave, nf * 8, hipMemcpyHostToDevice, b.stream));
  HIPCHK(hipMemcpyAsync(b.Ut, Ut, vol_doubles(b.cfg) * 8, hipMemcpyHostToDevice, b.stream));
  FX_DISPATCH_N(b.cfg.N, launch_surfint_n, b, b.Flux, fs);
  HIPCHK(hipGetLastError());
  return fx_ut_download(c, Ut);
}

int fx_volint_nfvse(fx_ctx *c, double *Ut_host)
{
  int rc = ready(c);
  if (rc) return rc;
  Batch &b = c->b;
  if (!b.cfg.nfvse) return set_err(FX_ESTATE, "VolInt_NFVSE needs nfvse=1");
  const double *a = q_nfvse_alpha(b);
  FX_DISPATCH_N(b.cfg.N, launch_nfvse_fv_n, b, a);
  RKArgs rk{};
  rk.alpha = a;
  rk.Ffv = b.Ffv;
  q_volume(b, VMODE_VOLINT, rk);
  HIPCHK(hipGetLastError());
  if (Ut_host) return fx_ut_download(c, Ut_host);
  return FX_OK;
}

int fx_blending(fx_ctx *c, double *alpha)
{
  int rc = ready(c);
  if (rc) return rc;
  if (!c->b.cfg.nfvse) return set_err(FX_ESTATE, "blending needs nfvse=1");
  if (!alpha) return set_err(FX_EINVAL, "null alpha");
  const double *a = q_nfvse_alpha(c->b);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(alpha, a, (size_t)c->b.cfg.nElems * 8, hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

int fx_calc_dt(fx_ctx *c, double CFLscale, double DFLscale, double *dt, int *errType)
{
  int rc = ready(c);
  if (rc) return rc;
  const int N = c->b.cfg.N;
  const double cfl = cfl_number(N, CFLscale), dfl = dfl_number(N, DFLscale);
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->b.stream, c->b.st, 0.0, 0.0, 1, NO_END_TIME);
  q_dt(c->b, cfl, dfl, 0);
  HIPCHK(hipGetLastError());
  DtState h;
  HIPCHK(hipMemcpyAsync(&h, c->b.st, sizeof(h), hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  if (dt) *dt = h.dt;
  if (errType) *errType = h.errType;
  return FX_OK;
}

int fx_rk_step(fx_ctx *c, double t, double dt)
{
  int rc = ready(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->b.stream, c->b.st, dt, t, 0, NO_END_TIME);
  q_rk_step(c->b);
  HIPCHK(hipGetLastError());
  return FX_OK;
}

// The step graph of the context (CALCTIMESTEP + the five stages), captured for cfl/dfl.
// Navier-Stokes: the first stage's lifting kernel forms the per-element CALCTIMESTEP from the
// U and metrics it reads anyway, and the step needs no CALCTIMESTEP pass of its own.
static int step_graph(fx_ctx *c, double cfl, double dfl)
{
  if (!c->graph || c->graph_cfl != cfl || c->graph_dfl != dfl) {
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->b.stream, hipStreamCaptureModeThreadLocal));
    if (dt_fusable(c->b)) {
      const DtFuse f = dt_fuse(c->b, cfl, dfl);
      q_rk_step(c->b, &f);
    } else {
      q_dt(c->b, cfl, dfl, 1);
      q_rk_step(c->b);
    }
    HIPCHK(hipStreamEndCapture(c->b.stream, &g));
    hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->graph = nullptr; return set_err(FX_EHIP, "graph instantiate: %s", hipGetErrorString(e)); }
    c->graph_cfl = cfl; c->graph_dfl = dfl;
  }
  return FX_OK;
}

int fx_run(fx_ctx *c, int nSteps, double CFLscale, double DFLscale, double *t, double *dt, int *errType)
{
  int rc = ready(c);
  if (rc) return rc;
  if (nSteps < 0) return set_err(FX_EINVAL, "nSteps < 0");
  const int N = c->b.cfg.N;
  const double cfl = cfl_number(N, CFLscale), dfl = dfl_number(N, DFLscale);
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->b.stream, c->b.st, 0.0, t ? *t : 0.0, 1, NO_END_TIME);
  HIPCHK(hipGetLastError());
  if ((rc = step_graph(c, cfl, dfl))) return rc;
  for (int s = 0; s < nSteps; s++) HIPCHK(hipGraphLaunch(c->graph, c->b.stream));
  DtState h;
  HIPCHK(hipMemcpyAsync(&h, c->b.st, sizeof(h), hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  if (t) *t = h.t;
  if (dt) *dt = h.dt;
  if (errType) *errType = h.errType;
  return FX_OK;
}

int fx_run_to(fx_ctx *c, int maxSteps, double CFLscale, double DFLscale, double tEnd, double *t, double *dt,
              int *steps, int *errType)
{
  int rc = ready(c);
  if (rc) return rc;
  if (!t) return set_err(FX_EINVAL, "null t");
  if (!(tEnd == tEnd)) return set_err(FX_EINVAL, "tEnd is NaN");
  const int N = c->b.cfg.N;
  const double cfl = cfl_number(N, CFLscale), dfl = dfl_number(N, DFLscale);
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->b.stream, c->b.st, 0.0, *t, 1, tEnd);
  HIPCHK(hipGetLastError());
  if ((rc = step_graph(c, cfl, dfl))) return rc;
  DtState h{};
  h.t = *t;
  int done = 0;
  if (!(*t < tEnd) || maxSteps == 0) done = 1;  // nothing to do (timedisc.f90:236)
  // launch in chunks sized from the last dt, so that few steps run after tEnd (those have
  // dt = 0 and change nothing); stop on the end time, a CALCTIMESTEP error or maxSteps
  int chunk = 1;
  while (!done) {
    if (maxSteps > 0) chunk = std::min(chunk, maxSteps - h.step);
    for (int s = 0; s < chunk; s++) HIPCHK(hipGraphLaunch(c->graph, c->b.stream));
    HIPCHK(hipMemcpyAsync(&h, c->b.st, sizeof(h), hipMemcpyDeviceToHost, c->b.stream));
    HIPCHK(hipStreamSynchronize(c->b.stream));
    if (h.done || h.errType != 0 || (maxSteps > 0 && h.step >= maxSteps)) break;
    const double left = (tEnd - h.t) / (h.dt > 0.0 ? h.dt : 1.0);
    chunk = (int)std::max(1.0, std::min(100.0, std::floor(left)));
  }
  *t = h.t;
  if (dt) *dt = h.dt;
  if (steps) *steps = h.step;
  if (errType) *errType = h.errType;
  return FX_OK;
}

int fx_glm(fx_ctx *c, double *glm_ch, double *glm_dtch1)
{
  if (!c) return set_err(FX_EINVAL, "null context");
  if (!c->b.mhd()) return set_err(FX_ESTATE, "GLM state needs nVar=9");
  DtState h;
  HIPCHK(hipMemcpyAsync(&h, c->b.st, sizeof(h), hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  if (glm_ch) *glm_ch = h.glm_ch;
  if (glm_dtch1) *glm_dtch1 = h.glm_dtch1;
  return FX_OK;
}

int fx_entropy(fx_ctx *c, double *S, double *dSdU_Ut)
{
  int rc = ready(c);
  if (rc) return rc;
  if (c->b.mhd()) return set_err(FX_EINVAL, "CalcEntropy is implemented for Euler/Navier-Stokes");
  q_entropy(c);
  HIPCHK(hipGetLastError());
  double h[2];
  HIPCHK(hipMemcpyAsync(h, c->red, sizeof(h), hipMemcpyDeviceToHost, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  if (S) *S = h[0] / (c->b.cfg.kappa - 1.0);
  if (dSdU_Ut) *dSdU_Ut = h[1];
  return FX_OK;
}

int fx_profile_kernel(fx_ctx *c, int which, int reps, double *avg_ms)
{
  int rc = ready(c);
  if (rc) return rc;
  if (reps < 1 || !avg_ms) return set_err(FX_EINVAL, "reps < 1 or null avg_ms");
  if (which == FX_KERNEL_LIFTING && !c->b.cfg.parabolic) return set_err(FX_EINVAL, "lifting needs parabolic=1");
  if (which < 0 || which > 6) return set_err(FX_EINVAL, "unknown kernel %d", which);
  if ((which == FX_KERNEL_NFVSE_ALPHA || which == FX_KERNEL_NFVSE_FV) && !c->b.cfg.nfvse)
    return set_err(FX_EINVAL, "NFVSE kernels need nfvse=1");
  // make the inputs of the timed kernel valid: viscous fluxes, face fluxes and a time step
  if (c->b.cfg.parabolic) q_lifting(c->b, 1);
  q_surface(c->b);
  if (which == FX_KERNEL_DT || which == FX_KERNEL_VOLUME_RK)
    q_dt(c->b, cfl_number(c->b.cfg.N, 0.9), 0.0, 0);
  RKArgs rk;
  rk.st = c->b.st; rk.eprob = nullptr; rk.A = c->b.rkA[1]; rk.b = c->b.rkb[1]; rk.first = 0;
  rk.alpha = nullptr; rk.Ffv = nullptr;
  if (which == FX_KERNEL_NFVSE_FV) {  // the FV kernel reads the blending coefficient
    const double *a = q_nfvse_alpha(c->b);
    if (a != c->b.alpha)
      HIPCHK(hipMemcpyAsync(c->b.alpha, a, (size_t)c->b.cfg.nElems * 8, hipMemcpyDeviceToDevice, c->b.stream));
  }
  if (which == FX_KERNEL_VOLUME_RK) {  // Ut_tmp must hold a previous stage: run stage 1 once
    RKArgs r1 = rk; r1.A = 0.0; r1.b = c->b.rkb[0]; r1.first = 1;
    q_volume(c->b, VMODE_RK, r1);
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->b.stream));
  for (int r = 0; r < reps; r++) {
    switch (which) {
    case FX_KERNEL_VOLUME_RK: q_volume(c->b, VMODE_RK, rk); break;
    case FX_KERNEL_SURFACE: q_surface(c->b); break;
    case FX_KERNEL_LIFTING: q_lifting(c->b, 1); break;
    case FX_KERNEL_DT: q_dt(c->b, 0.9, 0.9, 0); break;
    case FX_KERNEL_NFVSE_ALPHA: (void)q_nfvse_alpha(c->b); break;
    case FX_KERNEL_NFVSE_FV: FX_DISPATCH_N(c->b.cfg.N, launch_nfvse_fv_n, c->b, c->b.alpha); break;
    default: q_volume(c->b, VMODE_RHS, rk); break;
    }
  }
  HIPCHK(hipEventRecord(e1, c->b.stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *avg_ms = (double)ms / reps;
  return FX_OK;
}

int fx_volume_tile(int N, int parabolic, int precision, int *epb)
{
  if (N < 2 || N > 5 || !epb) return set_err(FX_EINVAL, "N=%d not supported or null epb", N);
  *epb = tile_epb(N, tuned_tile(N, parabolic, precision));
  return FX_OK;
}

int fx_use_volume_tile(fx_ctx *c, int epb)
{
  if (!c) return set_err(FX_EINVAL, "null context");
  const int N = c->b.cfg.N, nt = (c->b.cfg.volumeFlux == 32) ? 3 : 1;
  for (int t = 0; t < nt; t++)
    if (tile_epb(N, t) == epb) {
      c->b.vol_tile = t;
      if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
      return FX_OK;
    }
  return set_err(FX_EINVAL, "no volume tile with %d elements per workgroup for N=%d", epb, N);
}

int fx_tune_volume(fx_ctx *c, int reps, int *best_epb, int *tiles, double *ms, int *ntiles)
{
  int rc = ready(c);
  if (rc) return rc;
  if (reps < 1) return set_err(FX_EINVAL, "reps < 1");
  Batch &b = c->b;
  const int N = b.cfg.N;
  const int nt = (b.cfg.volumeFlux == 32) ? 3 : 1;  // only the Ranocha LSERK form has alternatives
  const size_t nv = vol_doubles(b.cfg);
  double *save = nullptr;
  HIPCHK(hipMalloc((void **)&save, 2 * nv * 8));
  auto cleanup = [&]() { (void)hipFree(save); };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best = 0;
  double bestms = 1e300;
  hipError_t err = hipSuccess;
#define TCHK(x) if ((err = (x)) != hipSuccess) goto fail;
  TCHK(hipMemcpyAsync(save, b.U, nv * 8, hipMemcpyDeviceToDevice, b.stream));
  TCHK(hipMemcpyAsync(save + nv, b.Uttmp, nv * 8, hipMemcpyDeviceToDevice, b.stream));
  TCHK(hipEventCreate(&e0));
  TCHK(hipEventCreate(&e1));
  // valid inputs of the timed kernel: viscous and face fluxes, a time step, a previous stage
  if (b.cfg.parabolic) q_lifting(b, 1);
  q_surface(b);
  q_dt(b, cfl_number(N, 0.9), 0.0, 0);
  for (int t = 0; t < nt; t++) {
    RKArgs rk;
    rk.st = b.st; rk.eprob = nullptr; rk.A = b.rkA[0]; rk.b = b.rkb[0]; rk.first = 1;
    rk.alpha = nullptr; rk.Ffv = nullptr;
    b.vol_tile = t;
    q_volume(b, VMODE_RK, rk);  // stage 1 form once (warm-up; fills Ut_tmp)
    rk.A = b.rkA[1]; rk.b = b.rkb[1]; rk.first = 0;
    TCHK(hipEventRecord(e0, b.stream));
    for (int r = 0; r < reps; r++) q_volume(b, VMODE_RK, rk);
    TCHK(hipEventRecord(e1, b.stream));
    TCHK(hipEventSynchronize(e1));
    float f = 0.0f;
    TCHK(hipEventElapsedTime(&f, e0, e1));
    const double avg = (double)f / reps;
    if (tiles && t < 4) tiles[t] = tile_epb(N, t);
    if (ms && t < 4) ms[t] = avg;
    if (avg < bestms) { bestms = avg; best = t; }
  }
  TCHK(hipMemcpyAsync(b.U, save, nv * 8, hipMemcpyDeviceToDevice, b.stream));
  TCHK(hipMemcpyAsync(b.Uttmp, save + nv, nv * 8, hipMemcpyDeviceToDevice, b.stream));
  TCHK(hipStreamSynchronize(b.stream));
  TCHK(hipGetLastError());
#undef TCHK
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cleanup();
  b.vol_tile = best;
  tuned_tile(N, b.cfg.parabolic, b.cfg.precision) = best;
  if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
  if (best_epb) *best_epb = tile_epb(N, best);
  if (ntiles) *ntiles = nt;
  return FX_OK;
fail:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(b.stream);
  cleanup();
  return set_err(FX_EHIP, "tuning: %s", hipGetErrorString(err));
}

int fx_elem_xgp_upload(fx_ctx *c, const double *Elem_xGP)
{
  if (!c || !Elem_xGP) return set_err(FX_EINVAL, "null argument");
  const size_t bytes = (size_t)3 * c->b.n3 * c->b.cfg.nElems * 8;
  if (!c->xgp) HIPCHK(hipMalloc(&c->xgp, bytes));
  HIPCHK(hipMemcpyAsync(c->xgp, Elem_xGP, bytes, hipMemcpyHostToDevice, c->b.stream));
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

int fx_analyze(fx_ctx *c, double t, const fx_analyze_config *cfg, fx_analysis *out)
{
  int rc = ready(c);
  if (rc) return rc;
  if (!cfg || !out) return set_err(FX_EINVAL, "null argument");
  Batch &b = c->b;
  const int N = b.cfg.N, NV = b.cfg.nVar, n = N + 1;
  if (cfg->exactFunc < 0 || cfg->exactFunc > 2 || (cfg->exactFunc == 2 && NV != 5))
    return set_err(FX_EINVAL, "AnalyzeExactFunc %d not supported (0, 1; 2 for 5 variables)", cfg->exactFunc);
  if (cfg->exactFunc > 0 && !c->xgp) return set_err(FX_ESTATE, "error norms need fx_elem_xgp_upload");
  const int NA = cfg->nAnalyze > 0 ? cfg->nAnalyze : 2 * (N + 1);
  if (NA > ANA_MAXNA - 1) return set_err(FX_EINVAL, "NAnalyze=%d above 16", NA);
  const int L = NV == 9 ? AnaRec<9>::COUNT : AnaRec<5>::COUNT;
  if ((size_t)(NV + 4) * (NA + 1) * n * n * 8 + 4 * 2 * NV * 8 > 65536)
    return set_err(FX_EINVAL, "NAnalyze=%d needs more than 64 KiB of LDS at N=%d", NA, N);
  if (!c->ana_part) {
    HIPCHK(hipMalloc(&c->ana_part, (size_t)b.cfg.nElems * L * 8));
    HIPCHK(hipMalloc(&c->ana_ops, (size_t)(ANA_MAXNA * (MAXN + 1) + ANA_MAXNA) * 8));
    HIPCHK(hipMalloc(&c->ana_out, (size_t)L * 8));
  }
  if (c->ana_na != NA) {
    std::vector<double> h(ANA_MAXNA * (MAXN + 1) + ANA_MAXNA, 0.0);
    analyze_basis(N, NA, h.data(), h.data() + ANA_MAXNA * (MAXN + 1));
    HIPCHK(hipMemcpyAsync(c->ana_ops, h.data(), h.size() * 8, hipMemcpyHostToDevice, b.stream));
    HIPCHK(hipStreamSynchronize(b.stream));
    c->ana_na = NA;
  }
  AnaArgs a{};
  a.exactFunc = cfg->exactFunc;
  a.na = NA + 1;
  for (int v = 0; v < 9; v++) a.ref[v] = cfg->refState[v];
  for (int d = 0; d < 3; d++) { a.adv[d] = cfg->advVel[d]; a.wave[d] = cfg->waveNumber[d]; }
  a.amp = cfg->amplitude;
  a.t = t;
  a.kappa = b.cfg.kappa;
  a.s2mu0 = NV == 9 ? 0.5 / b.cfg.mu_0 : 0.0;
  a.R = cfg->R;
  // DGTimeDerivative of the state (Ut), and the primitive gradients when parabolic
  RKArgs rk{};
  q_stage(b, VMODE_RHS, rk);
  if (b.cfg.parabolic) q_lifting(b, 0);
  FX_DISPATCH_N(N, launch_analyze_n, b, a, c->xgp, c->ana_ops, c->ana_part, c->ana_out);
  HIPCHK(hipGetLastError());
  double h[AnaRec<9>::COUNT];
  HIPCHK(hipMemcpyAsync(h, c->ana_out, (size_t)L * 8, hipMemcpyDeviceToHost, b.stream));
  HIPCHK(hipStreamSynchronize(b.stream));
  // normalisation of CalcErrorNorms and AnalyzeTestCase
  std::memset(out, 0, sizeof(*out));
  const int L2 = 1, L2C = 1 + NV, LINF = 1 + 2 * NV, BULK = 1 + 3 * NV, BULKT = 1 + 4 * NV, MAXUT = 1 + 5 * NV,
            TC = 1 + 6 * NV;
  const double vol = h[0], rho0 = 1.0;
  out->vol = vol;
  for (int v = 0; v < NV; v++) {
    out->L2[v] = std::sqrt(h[L2 + v] / vol);
    out->L2colloc[v] = std::sqrt(h[L2C + v] / vol);
    out->Linf[v] = h[LINF + v];
    out->bulk[v] = h[BULK + v];
    out->bulk_t[v] = h[BULKT + v];
    out->maxabs_Ut[v] = h[MAXUT + v];
  }
  out->uprime = std::sqrt(h[TC + 0] / vol * 2.0 / 3.0);
  out->Ekin = h[TC + 0] / vol;
  out->Ekin_comp = h[TC + 1] / (rho0 * vol);
  out->Emag_comp = h[TC + 2] / (rho0 * vol);
  out->Enstrophy_comp = h[TC + 3] / (rho0 * vol);
  out->Mag_Enstrophy = h[TC + 4] / (rho0 * vol);
  out->Entropy_comp = h[TC + 5] / (rho0 * vol);
  out->DR_u = h[TC + 6] * cfg->dissFactor / vol;
  out->DR_S = h[TC + 7] * 2.0 * cfg->dissFactor / (rho0 * vol);
  out->DR_Sd = h[TC + 8] * 2.0 * cfg->dissFactor / (rho0 * vol);
  out->DR_p = -h[TC + 9] / (rho0 * vol);
  out->mean_temperature = h[TC + 10] / vol;
  out->max_Vorticity = h[TC + 11];
  out->max_Current = h[TC + 12];
  return FX_OK;
}

int fx_synchronize(fx_ctx *c)
{
  if (!c) return set_err(FX_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->b.stream));
  return FX_OK;
}

}  // extern "C"

// ======================================================================================
// fx_rhs_noalloc: the right-hand side of fx_rhs for callers that cannot allocate device
// memory.  The face fluxes, packed normals, viscous terms and NFVSE buffers that a context
// keeps live in `scratch`, one file-scope device array per GPU, held until the process
// ends.  Host code never passes `scratch` itself: its name on the host is a stand-in, so
// each device's address is looked up once (hipGetSymbolAddress) and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[FX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAXDEV = 64;
std::mutex g_scratch_mu;
void *g_scratch[NOALLOC_MAXDEV] = {};

// The current device's address of `scratch` (looked up on its first call on that device).
int scratch_ptr(unsigned char **out)
{
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAXDEV) return set_err(FX_ENODEV, "device %d beyond %d", dev, NOALLOC_MAXDEV);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(FX_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_scratch[dev]);
  return FX_OK;
}

// Scratch layout of one call, each part 256-byte aligned; returns the bytes needed.
struct NoallocLayout {
  size_t nvse = 0, ftab = 0, flux = 0, uv = 0, fvn = 0, alpha = 0, alpha2 = 0, ffv = 0, flag = 0, total = 0;
};
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
NoallocLayout noalloc_layout(const fx_config &c)
{
  const size_t n = c.N + 1, n3E = n * n * n * c.nElems, nf = n * n * c.nSides;
  NoallocLayout L;
  size_t o = 0;
  auto take = [&](size_t &at, size_t bytes) { at = o; o += align256(bytes); };
  take(L.nvse, nf * sizeof(double4));
  take(L.flux, 5 * nf * 8);
  if (c.parabolic) {
    take(L.ftab, (size_t)6 * n * n * c.nElems * sizeof(int2));
    take(L.uv, 4 * n3E * 8);
    take(L.fvn, 8 * nf * 8);
  }
  if (c.nfvse) {
    take(L.alpha, (size_t)c.nElems * 8);
    take(L.alpha2, (size_t)c.nElems * 8);
    take(L.ffv, 5 * n3E * 8);
  }
  take(L.flag, sizeof(int));
  L.total = o;
  return L;
}

__global__ void k_noalloc_flag_clear(int *flag) { flag[0] = 0; }

// Device-side form of check_connectivity: every index the kernels derive from the caller's
// ElemToSide / SideToElem must stay inside the mesh; any violation sets *flag.
__global__ void k_noalloc_check(const int *__restrict__ e2s, const int *__restrict__ s2e, int E, int S,
                                int *__restrict__ flag)
{
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int bad = 0;
  if (g < S) {
    const int s = (int)g;
    const int *t = s2e + 5 * (size_t)s;
    if (t[0] < 1 || t[0] > E || t[1] < 1 || t[1] > E || t[2] < 1 || t[2] > 6 || t[3] < 1 || t[3] > 6 ||
        t[4] < 1 || t[4] > 4) {
      bad = 1;
    } else {
      const int *am = e2s + 2 * (6 * (size_t)(t[0] - 1) + t[2] - 1);
      const int *as = e2s + 2 * (6 * (size_t)(t[1] - 1) + t[3] - 1);
      if (am[0] != s + 1 || am[1] != 0 || as[0] != s + 1 || as[1] != t[4]) bad = 1;
    }
  } else if (g < (long long)S + 6LL * E) {
    const int *a = e2s + 2 * (size_t)(g - S);
    if (a[0] < 1 || a[0] > S || a[1] < 0 || a[1] > 4) bad = 1;
  }
  if (bad) atomicOr(flag, 1);
}

// (nx, ny, nz, SurfElem) per face point, the layout the kernels read
__global__ void k_noalloc_pack_nvse(const double *__restrict__ nv, const double *__restrict__ se,
                                    double4 *__restrict__ out, long long nf)
{
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) out[f] = make_double4(nv[3 * f], nv[3 * f + 1], nv[3 * f + 2], se[f]);
}

}  // namespace

extern "C" {

long long fx_rhs_noalloc_bytes(const fx_config *cfg)
{
  if (!cfg || cfg->N < 2 || cfg->N > 5 || cfg->nElems < 1 || cfg->nSides < 1) return -1;
  return (long long)noalloc_layout(*cfg).total;
}

int fx_rhs_noalloc(const fx_config *cfg, const fx_mesh *mesh, const double *U, double *Ut, void *stream)
{
  if (!cfg || !mesh || !U || !Ut) return set_err(FX_EINVAL, "null argument");
  int rc = check_config(cfg);
  if (rc) return rc;
  if (cfg->nVar != 5)
    return set_err(FX_EINVAL, "fx_rhs_noalloc: Euler / Navier-Stokes (nVar = 5) only; GLM-MHD keeps GLM_ch "
                              "state between calls and needs a context");
  if (cfg->precision != FX_PREC_FP64)
    return set_err(FX_EINVAL, "fx_rhs_noalloc: FP64 only (FP32-split needs FP32 metric copies)");
  if (!mesh->ElemToSide || !mesh->SideToElem || !mesh->Metrics_fTilde || !mesh->Metrics_gTilde ||
      !mesh->Metrics_hTilde || !mesh->sJ || !mesh->NormVec || !mesh->SurfElem)
    return set_err(FX_EINVAL, "missing mesh array");
  if (mesh->nMortarSides != 0) return set_err(FX_EINVAL, "fx_rhs_noalloc: conforming meshes only (no mortars)");
  const NoallocLayout L = noalloc_layout(*cfg);
  if (L.total > (size_t)FX_NOALLOC_SCRATCH_BYTES)
    return set_err(FX_EINVAL, "fx_rhs_noalloc: N=%d, %d elements, %d sides need %zu bytes of scratch, above "
                              "the maximum %lld", cfg->N, cfg->nElems, cfg->nSides, L.total,
                   (long long)FX_NOALLOC_SCRATCH_BYTES);
  unsigned char *sp = nullptr;
  if ((rc = scratch_ptr(&sp))) return rc;
  if ((rc = ensure_ops(cfg->N))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int *flag = reinterpret_cast<int *>(sp + L.flag);
  // connectivity check on the device, read back before any kernel indexes through it
  hipLaunchKernelGGL(k_noalloc_flag_clear, dim3(1), dim3(1), 0, st, flag);
  const long long nchk = (long long)cfg->nSides + 6LL * cfg->nElems;
  hipLaunchKernelGGL(k_noalloc_check, dim3((unsigned)((nchk + 255) / 256)), dim3(256), 0, st,
                     (const int *)mesh->ElemToSide, (const int *)mesh->SideToElem, cfg->nElems, cfg->nSides, flag);
  HIPCHK(hipGetLastError());
  int hflag = 0;
  HIPCHK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (hflag) return set_err(FX_EINVAL, "fx_rhs_noalloc: ElemToSide / SideToElem invalid (only inner conforming "
                                       "sides are supported)");
  Batch b;
  b.cfg = *cfg;
  b.set_degree(cfg->N);
  b.stream = st;
  b.U = const_cast<double *>(U);  // read only in the right-hand-side form
  b.Ut = Ut;
  b.Mf = const_cast<double *>(mesh->Metrics_fTilde);
  b.Mg = const_cast<double *>(mesh->Metrics_gTilde);
  b.Mh = const_cast<double *>(mesh->Metrics_hTilde);
  b.sJ = const_cast<double *>(mesh->sJ);
  b.e2s = const_cast<int *>(mesh->ElemToSide);
  b.s2e = const_cast<int *>(mesh->SideToElem);
  b.nvse = reinterpret_cast<double4 *>(sp + L.nvse);
  b.Flux = reinterpret_cast<double *>(sp + L.flux);
  if (cfg->parabolic) {
    b.ftab = reinterpret_cast<int2 *>(sp + L.ftab);
    b.Uv = reinterpret_cast<double *>(sp + L.uv);
    b.FvN = reinterpret_cast<double *>(sp + L.fvn);
  }
  if (cfg->nfvse) {
    b.alpha = reinterpret_cast<double *>(sp + L.alpha);
    b.alpha2 = reinterpret_cast<double *>(sp + L.alpha2);
    b.Ffv = reinterpret_cast<double *>(sp + L.ffv);
  }
  const long long nf = (long long)b.n2 * cfg->nSides;
  hipLaunchKernelGGL(k_noalloc_pack_nvse, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, mesh->NormVec,
                     mesh->SurfElem, b.nvse, nf);
  if (cfg->parabolic) FX_DISPATCH_N(b.cfg.N, launch_face_table_n, b, 0, cfg->nElems);
  RKArgs rk{};
  q_stage(b, VMODE_RHS, rk);
  HIPCHK(hipGetLastError());
  return FX_OK;
}

}  // extern "C"

// The batched server (fx_server_*) shares this device-code unit.
#include "fx_server.hpp"
