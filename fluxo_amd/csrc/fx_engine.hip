// MI355X DGSEM right-hand-side engine: context, device memory, launches and the C ABI
// declared in include/fluxo_gpu.h.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/fluxo_gpu.h"
#include "fx_kernels.hpp"

namespace fx {
__constant__ Ops c_ops[MAXN + 1];
}
using namespace fx;

static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t _e = (x);                                                                   \
    if (_e != hipSuccess)                                                                  \
      return set_err(FX_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, \
                     __LINE__);                                                            \
  } while (0)

// --------------------------------------------------------------------------------------
// Host-side LGL basis (the engine's InitDGbasis, dg.f90:181-252, with the node/weight
// Newton iteration of LegGaussLobNodesAndWeights basis.f90:544-608 and the barycentric
// derivative matrix of PolynomialDerivativeMatrix basis.f90:641-663).
namespace {

void lgl_nodes(int N, double *x, double *w)
{
  const double pi = std::acos(-1.0);
  x[0] = -1.0; x[N] = 1.0;
  w[0] = 2.0 / (double)(N * (N + 1)); w[N] = w[0];
  auto qL = [N](double xx, double &q, double &qd, double &L) {
    double Lm2 = 1.0, Lm1 = xx, Ld, Ldm2 = 0.0, Ldm1 = 1.0;
    L = 0.0;
    for (int i = 2; i <= N; i++) {
      L = ((double)(2 * i - 1) * xx * Lm1 - (double)(i - 1) * Lm2) / (double)i;
      Ld = Ldm2 + (double)(2 * i - 1) * Lm1;
      Lm2 = Lm1; Lm1 = L; Ldm2 = Ldm1; Ldm1 = Ld;
    }
    q = (double)(2 * N + 1) / (double)(N + 1) * (xx * L - Lm2);
    qd = (double)(2 * N + 1) * L;
  };
  if (N > 1) {
    double c1 = pi / (double)N, c2 = 3.0 / ((double)(8 * N) * pi);
    for (int i = 1; i <= (N + 1) / 2 - 1; i++) {
      double xi = -std::cos(c1 * ((double)i + 0.25) - c2 / ((double)i + 0.25)), q, qd, L;
      for (int it = 0; it <= 10; it++) {
        qL(xi, q, qd, L);
        double dx = -q / qd;
        xi += dx;
        if (std::fabs(dx) < 1e-15 * std::fabs(xi)) break;
      }
      qL(xi, q, qd, L);
      x[i] = xi; x[N - i] = -xi;
      w[i] = w[0] / (L * L); w[N - i] = w[i];
    }
  }
  if (N % 2 == 0) {
    double q, qd, L;
    x[N / 2] = 0.0;
    qL(0.0, q, qd, L);
    w[N / 2] = w[0] / (L * L);
  }
}

void make_ops(int N, Ops &o)
{
  std::memset(&o, 0, sizeof(o));
  const int n = N + 1;
  double x[MAXN + 1], w[MAXN + 1], wb[MAXN + 1];
  lgl_nodes(N, x, w);
  for (int i = 0; i < n; i++) wb[i] = 1.0;
  for (int i = 1; i <= N; i++)
    for (int j = 0; j < i; j++) { wb[j] *= (x[j] - x[i]); wb[i] *= (x[i] - x[j]); }
  for (int i = 0; i < n; i++) wb[i] = 1.0 / wb[i];
  for (int l = 0; l < n; l++)
    for (int i = 0; i < n; i++)
      if (l != i) {
        o.D[i][l] = wb[l] / (wb[i] * (x[i] - x[l]));
        o.D[i][i] -= o.D[i][l];
      }
  // Lagrange polynomials at -1, +1 are unit vectors on LGL nodes
  double Lm[MAXN + 1] = {0}, Lp[MAXN + 1] = {0};
  Lm[0] = 1.0; Lp[N] = 1.0;
  for (int i = 0; i < n; i++) {
    o.w[i] = w[i];
    for (int j = 0; j < n; j++) {
      o.Dhat[i][j] = -((1.0 / w[i]) * (o.D[j][i] * w[j]));
      double vbv = -Lm[i] * Lm[j] + Lp[i] * Lp[j];
      o.Ds[i][j] = 2.0 * o.D[i][j] - (1.0 / w[i]) * vbv;
    }
  }
  o.L0 = Lm[0] / w[0];
  for (int i = 0; i <= MAXN; i++)
    for (int j = 0; j <= MAXN; j++) {
      o.Ds32[i][j] = (float)o.Ds[i][j];
      o.Dhat32[i][j] = (float)o.Dhat[i][j];
    }
}

bool g_ops_ready[MAXN + 1] = {false};

int ensure_ops(int N)
{
  if (g_ops_ready[N]) return FX_OK;
  Ops o;
  make_ops(N, o);
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ops), &o, sizeof(Ops), sizeof(Ops) * N, hipMemcpyHostToDevice));
  g_ops_ready[N] = true;
  return FX_OK;
}

}  // namespace

// --------------------------------------------------------------------------------------
struct fx_ctx {
  fx_config cfg;
  int n = 0, n2 = 0, n3 = 0;
  hipStream_t stream = nullptr;
  double *U = nullptr, *Ut = nullptr, *Uttmp = nullptr;
  double *Mf = nullptr, *Mg = nullptr, *Mh = nullptr, *sJ = nullptr;
  float *Mf32 = nullptr, *Mg32 = nullptr, *Mh32 = nullptr;  // split precision only
  double *Flux = nullptr, *Fv = nullptr, *grad = nullptr, *part = nullptr, *red = nullptr;
  double *dtElem = nullptr;
  int *errElem = nullptr;
  double4 *nvse = nullptr;
  int *e2s = nullptr, *s2e = nullptr;
  DtState *dts = nullptr;
  bool mesh_ok = false, state_ok = false;
  hipGraphExec_t graph = nullptr;
  double graph_cfl = -1.0, graph_dfl = -1.0;
  double rkA[5], rkb[5], rkc[5];
};

namespace {

size_t vol_doubles(const fx_config &c) { size_t n = c.N + 1; return (size_t)c.nVar * n * n * n * c.nElems; }

void rk_coefs(double *A, double *b, double *c)
{
  // Carpenter & Kennedy RK4(5) 2N-storage (timedisc_vars.f90:219-233)
  A[0] = 0.0;
  A[1] = 567301805773.0 / 1357537059087.0;
  A[2] = 2404267990393.0 / 2016746695238.0;
  A[3] = 3550918686646.0 / 2091501179385.0;
  A[4] = 1275806237668.0 / 842570457699.0;
  b[0] = 1432997174477.0 / 9575080441755.0;
  b[1] = 5161836677717.0 / 13612068292357.0;
  b[2] = 1720146321549.0 / 2090206949498.0;
  b[3] = 3134564353537.0 / 4481467310338.0;
  b[4] = 2277821191437.0 / 14882151754819.0;
  c[0] = 0.0;
  c[1] = 1432997174477.0 / 9575080441755.0;
  c[2] = 2526269341429.0 / 6820363962896.0;
  c[3] = 2006345519317.0 / 3224310063776.0;
  c[4] = 2802321613138.0 / 2924317926251.0;
}

// CFLScaleAlpha / DFLScaleAlpha for LGL and Carpenter RK4-5 (timedisc_vars.f90:79-84, 212-216)
double cfl_alpha(int N)
{
  static const double a[15] = {4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830,
                               1.5682, 1.4692, 1.3849, 1.3106, 1.2454, 1.1880, 1.1362};
  return a[(N < 15 ? N : 15) - 1];
}
double dfl_alpha(int N)
{
  static const double a[10] = {4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17};
  return a[(N < 10 ? N : 10) - 1] * 1.853;
}

// ---- template dispatch ---------------------------------------------------------------
template <int N, int VF, bool PARA, int MODE>
void launch_volume_t(fx_ctx *c, RKArgs rk)
{
  const int nb = (c->cfg.nElems + VolGeom<N>::EPB - 1) / VolGeom<N>::EPB;
  if (c->cfg.precision == FX_PREC_FP32_SPLIT)
    hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, float>), dim3(nb), dim3(3 * VolGeom<N>::W), 0,
                       c->stream, c->U, c->Mf32, c->Mg32, c->Mh32, c->sJ, c->e2s, c->Flux, c->Fv, c->Ut,
                       c->Uttmp, rk, c->cfg.nElems, c->cfg.kappa);
  else
    hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, double>), dim3(nb), dim3(3 * VolGeom<N>::W), 0,
                       c->stream, c->U, c->Mf, c->Mg, c->Mh, c->sJ, c->e2s, c->Flux, c->Fv, c->Ut,
                       c->Uttmp, rk, c->cfg.nElems, c->cfg.kappa);
}

template <int N, int VF>
void launch_volume_v(fx_ctx *c, int mode, RKArgs rk)
{
  const bool para = c->cfg.parabolic != 0;
  if (mode == VMODE_VOLINT) launch_volume_t<N, VF, false, VMODE_VOLINT>(c, rk);
  else if (mode == VMODE_RK) {
    if (para) launch_volume_t<N, VF, true, VMODE_RK>(c, rk);
    else launch_volume_t<N, VF, false, VMODE_RK>(c, rk);
  } else {
    if (para) launch_volume_t<N, VF, true, VMODE_RHS>(c, rk);
    else launch_volume_t<N, VF, false, VMODE_RHS>(c, rk);
  }
}

template <int N>
void launch_volume_n(fx_ctx *c, int mode, RKArgs rk)
{
  switch (c->cfg.volumeFlux) {
  case 5: launch_volume_v<N, 5>(c, mode, rk); break;
  case 32: launch_volume_v<N, 32>(c, mode, rk); break;
  default: launch_volume_v<N, 0>(c, mode, rk); break;
  }
}

template <int N, int RS>
void launch_surface_t(fx_ctx *c)
{
  const long long tot = (long long)c->cfg.nSides * c->n2;
  const int nb = (int)((tot + 255) / 256);
  if (c->cfg.parabolic)
    hipLaunchKernelGGL((k_surface<N, RS, true>), dim3(nb), dim3(256), 0, c->stream, c->U, c->Fv, c->s2e,
                       c->nvse, c->Flux, c->cfg.nSides, c->cfg.kappa);
  else
    hipLaunchKernelGGL((k_surface<N, RS, false>), dim3(nb), dim3(256), 0, c->stream, c->U, c->Fv, c->s2e,
                       c->nvse, c->Flux, c->cfg.nSides, c->cfg.kappa);
}

template <int N>
void launch_surface_n(fx_ctx *c)
{
  switch (c->cfg.riemann) {
  case 0: launch_surface_t<N, 0>(c); break;
  case 9: launch_surface_t<N, 9>(c); break;
  case 32: launch_surface_t<N, 32>(c); break;
  default: launch_surface_t<N, 1>(c); break;
  }
}

template <int N>
void launch_lifting_n(fx_ctx *c, int out)
{
  const int epb = LiftGeom<N>::EPB;
  const int nb = (c->cfg.nElems + epb - 1) / epb;
  const double KsPr = c->cfg.kappa / c->cfg.Pr;
  const size_t nv = vol_doubles(c->cfg);
  if (out == 0)
    hipLaunchKernelGGL((k_lifting<N, 0>), dim3(nb), dim3(256), 0, c->stream, c->U, c->Mf, c->Mg, c->Mh,
                       c->sJ, c->e2s, c->s2e, c->nvse, c->grad, c->grad + nv, c->grad + 2 * nv,
                       c->cfg.nElems, c->cfg.kappa, c->cfg.mu0, KsPr);
  else
    hipLaunchKernelGGL((k_lifting<N, 1>), dim3(nb), dim3(256), 0, c->stream, c->U, c->Mf, c->Mg, c->Mh,
                       c->sJ, c->e2s, c->s2e, c->nvse, c->Fv, nullptr, nullptr, c->cfg.nElems,
                       c->cfg.kappa, c->cfg.mu0, KsPr);
}

template <int N>
void launch_dt_n(fx_ctx *c, double cfl, double dfl)
{
  const int epb = DtGeom<N>::EPB;
  const int nb = (c->cfg.nElems + epb - 1) / epb;
  const double KsPr = c->cfg.kappa / c->cfg.Pr;
  const double KsPr_max = (4.0 / 3.0 > KsPr) ? 4.0 / 3.0 : KsPr;
  hipLaunchKernelGGL((k_calc_dt<N>), dim3(nb), dim3(DtGeom<N>::T), 0, c->stream, c->U, c->Mf, c->Mg, c->Mh, c->sJ,
                     c->dtElem, c->errElem, c->cfg.nElems, c->cfg.kappa, cfl, dfl, c->cfg.mu0, KsPr_max,
                     c->cfg.parabolic);
}

template <int N>
void launch_entropy_n(fx_ctx *c)
{
  const long long nn = (long long)c->n3 * c->cfg.nElems;
  const int nb = (int)((nn + 255) / 256);
  hipLaunchKernelGGL((k_entropy<N>), dim3(nb), dim3(256), 0, c->stream, c->U, c->Ut, c->sJ, c->part, nn,
                     c->cfg.kappa);
  hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, c->stream, c->part, nb, c->red);
}

#define FX_DISPATCH_N(N_, fn, ...)             \
  switch (N_) {                                \
  case 2: fn<2>(__VA_ARGS__); break;           \
  case 3: fn<3>(__VA_ARGS__); break;           \
  case 4: fn<4>(__VA_ARGS__); break;           \
  case 5: fn<5>(__VA_ARGS__); break;           \
  default: break;                              \
  }

void q_volume(fx_ctx *c, int mode, RKArgs rk) { FX_DISPATCH_N(c->cfg.N, launch_volume_n, c, mode, rk); }
void q_surface(fx_ctx *c) { FX_DISPATCH_N(c->cfg.N, launch_surface_n, c); }
void q_lifting(fx_ctx *c, int out) { FX_DISPATCH_N(c->cfg.N, launch_lifting_n, c, out); }
// CALCTIMESTEP into c->dts (advance=1 also moves t and the step counter)
void q_dt(fx_ctx *c, double cfl, double dfl, int advance)
{
  FX_DISPATCH_N(c->cfg.N, launch_dt_n, c, cfl, dfl);
  hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, c->stream, c->dtElem, c->errElem, c->cfg.nElems,
                     c->dts, advance);
}
void q_entropy(fx_ctx *c) { FX_DISPATCH_N(c->cfg.N, launch_entropy_n, c); }

// one RHS evaluation, followed by the RK stage update when rk != nullptr
void q_stage(fx_ctx *c, int mode, RKArgs rk)
{
  if (c->cfg.parabolic) q_lifting(c, 1);
  q_surface(c);
  q_volume(c, mode, rk);
}

void q_rk_step(fx_ctx *c)
{
  for (int s = 0; s < 5; s++) {
    RKArgs rk;
    rk.dt = &c->dts->dt;
    rk.A = c->rkA[s];
    rk.b = c->rkb[s];
    rk.first = (s == 0);
    q_stage(c, VMODE_RK, rk);
  }
}

__global__ void k_set_dt(DtState *st, double dt, double t, int resetErr)
{
  st->dt = dt;
  st->t = t;
  if (resetErr) { st->errType = 0; st->errStep = -1; st->step = 0; }
}

int check_config(const fx_config *c)
{
  if (!c) return set_err(FX_EINVAL, "null config");
  if (c->N < 2 || c->N > 5) return set_err(FX_EINVAL, "N=%d not supported (2..5)", c->N);
  if (c->nVar != 5) return set_err(FX_EINVAL, "nVar=%d not supported (5)", c->nVar);
  if (c->nElems < 1 || c->nSides < 1) return set_err(FX_EINVAL, "empty mesh");
  if (c->volumeFlux != 0 && c->volumeFlux != 5 && c->volumeFlux != 32)
    return set_err(FX_EINVAL, "VolumeFlux=%d not supported (0,5,32)", c->volumeFlux);
  if (c->riemann != 0 && c->riemann != 1 && c->riemann != 9 && c->riemann != 32)
    return set_err(FX_EINVAL, "Riemann=%d not supported (0,1,9,32)", c->riemann);
  if (c->parabolic && c->lifting != 1) return set_err(FX_EINVAL, "only BR1 lifting (lifting=1)");
  if (c->precision != FX_PREC_FP64 && c->precision != FX_PREC_FP32_SPLIT)
    return set_err(FX_EINVAL, "precision %d not supported (0 FP64, 1 FP32-split)", c->precision);
  if (!(c->kappa > 1.0)) return set_err(FX_EINVAL, "kappa must be > 1");
  if (c->parabolic && !(c->Pr > 0.0)) return set_err(FX_EINVAL, "Pr must be > 0");
  return FX_OK;
}

}  // namespace

// ======================================================================================
extern "C" {

const char *fx_last_error(void) { return g_err; }

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
  if (cfg->parabolic) b += 12LL * n3 * E * 8 + 15LL * n3 * E * 8;
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
  c->cfg = *cfg;
  c->n = cfg->N + 1; c->n2 = c->n * c->n; c->n3 = c->n2 * c->n;
  rk_coefs(c->rkA, c->rkb, c->rkc);
  const size_t nv = vol_doubles(*cfg), n3E = (size_t)c->n3 * cfg->nElems;
  const size_t nf = (size_t)c->n2 * cfg->nSides;
  auto fail = [&](hipError_t e, const char *what) {
    set_err(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP, "%s: %s", what, hipGetErrorString(e));
    fx_ctx_destroy(c);
    return e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
#define ALLOC(p, bytes)                                               \
  if ((e = hipMalloc((void **)&(p), (bytes))) != hipSuccess) return fail(e, "hipMalloc " #p);
  ALLOC(c->U, nv * 8);
  ALLOC(c->Ut, nv * 8);
  ALLOC(c->Uttmp, nv * 8);
  ALLOC(c->Mf, 3 * n3E * 8);
  ALLOC(c->Mg, 3 * n3E * 8);
  ALLOC(c->Mh, 3 * n3E * 8);
  ALLOC(c->sJ, n3E * 8);
  if (cfg->precision == FX_PREC_FP32_SPLIT) {
    ALLOC(c->Mf32, 3 * n3E * 4);
    ALLOC(c->Mg32, 3 * n3E * 4);
    ALLOC(c->Mh32, 3 * n3E * 4);
  }
  ALLOC(c->Flux, 5 * nf * 8);
  ALLOC(c->nvse, nf * sizeof(double4));
  ALLOC(c->e2s, (size_t)12 * cfg->nElems * 4);
  ALLOC(c->s2e, (size_t)5 * cfg->nSides * 4);
  if (cfg->parabolic) {
    ALLOC(c->Fv, 12 * n3E * 8);
    ALLOC(c->grad, 3 * nv * 8);
  }
  const size_t npart = (n3E + 255) / 256 + 1;
  ALLOC(c->part, 2 * npart * 8);
  ALLOC(c->dtElem, (size_t)cfg->nElems * 8);
  ALLOC(c->errElem, (size_t)cfg->nElems * 4);
  ALLOC(c->red, 2 * 8);
  ALLOC(c->dts, sizeof(DtState));
#undef ALLOC
  DtState z{};
  z.errType = 0; z.errStep = -1; z.dt = 0.0; z.t = 0.0; z.step = 0;
  if ((e = hipMemcpy(c->dts, &z, sizeof(z), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "init");
  *out = c;
  return FX_OK;
}

void fx_ctx_destroy(fx_ctx *c)
{
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  void *ptrs[] = {c->U, c->Ut, c->Uttmp, c->Mf, c->Mg, c->Mh, c->Mf32, c->Mg32, c->Mh32, c->sJ, c->Flux, c->Fv, c->grad,
                  c->part, c->red, c->nvse, c->e2s, c->s2e, c->dts, c->dtElem, c->errElem};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int fx_mesh_upload(fx_ctx *c, const fx_mesh *m)
{
  if (!c || !m) return set_err(FX_EINVAL, "null argument");
  if (!m->ElemToSide || !m->SideToElem || !m->Metrics_fTilde || !m->Metrics_gTilde ||
      !m->Metrics_hTilde || !m->sJ || !m->NormVec || !m->SurfElem)
    return set_err(FX_EINVAL, "missing mesh array");
  const int E = c->cfg.nElems, S = c->cfg.nSides, N = c->cfg.N, n = c->n, n2 = c->n2;
  // connectivity checks: every index the kernels derive must stay in range
  for (int s = 0; s < S; s++) {
    const int *t = m->SideToElem + 5 * s;
    if (t[0] < 1 || t[0] > E || t[1] < 1 || t[1] > E || t[2] < 1 || t[2] > 6 || t[3] < 1 || t[3] > 6 ||
        t[4] < 1 || t[4] > 4)
      return set_err(FX_EINVAL, "SideToElem(:,%d) invalid (only inner conforming sides are supported)", s + 1);
    const int *am = m->ElemToSide + 2 * (6 * (t[0] - 1) + t[2] - 1);
    const int *as = m->ElemToSide + 2 * (6 * (t[1] - 1) + t[3] - 1);
    if (am[0] != s + 1 || am[1] != 0 || as[0] != s + 1 || as[1] != t[4])
      return set_err(FX_EINVAL, "ElemToSide inconsistent with SideToElem at side %d", s + 1);
  }
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 6; l++) {
      const int *a = m->ElemToSide + 2 * (6 * e + l);
      if (a[0] < 1 || a[0] > S || a[1] < 0 || a[1] > 4)
        return set_err(FX_EINVAL, "ElemToSide(:,%d,%d) invalid", l + 1, e + 1);
    }
  (void)N;
  const size_t n3E = (size_t)c->n3 * E, nf = (size_t)n2 * S;
  std::vector<double4> nvse(nf);
  for (size_t f = 0; f < nf; f++)
    nvse[f] = make_double4(m->NormVec[3 * f], m->NormVec[3 * f + 1], m->NormVec[3 * f + 2], m->SurfElem[f]);
  (void)n;
  HIPCHK(hipMemcpyAsync(c->Mf, m->Metrics_fTilde, 3 * n3E * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Mg, m->Metrics_gTilde, 3 * n3E * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->Mh, m->Metrics_hTilde, 3 * n3E * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->sJ, m->sJ, n3E * 8, hipMemcpyHostToDevice, c->stream));
  std::vector<float> m32;
  if (c->cfg.precision == FX_PREC_FP32_SPLIT) {
    m32.resize(9 * n3E);
    const double *src[3] = {m->Metrics_fTilde, m->Metrics_gTilde, m->Metrics_hTilde};
    float *dst[3] = {c->Mf32, c->Mg32, c->Mh32};
    for (int d = 0; d < 3; d++) {
      for (size_t i = 0; i < 3 * n3E; i++) m32[d * 3 * n3E + i] = (float)src[d][i];
      HIPCHK(hipMemcpyAsync(dst[d], m32.data() + d * 3 * n3E, 3 * n3E * 4, hipMemcpyHostToDevice, c->stream));
    }
  }
  HIPCHK(hipMemcpyAsync(c->nvse, nvse.data(), nf * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->e2s, m->ElemToSide, (size_t)12 * E * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->s2e, m->SideToElem, (size_t)5 * S * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->mesh_ok = true;
  return FX_OK;
}

int fx_state_upload(fx_ctx *c, const double *U)
{
  if (!c || !U) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(c->U, U, vol_doubles(c->cfg) * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->state_ok = true;
  return FX_OK;
}

int fx_state_download(fx_ctx *c, double *U)
{
  if (!c || !U) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(U, c->U, vol_doubles(c->cfg) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FX_OK;
}

int fx_ut_download(fx_ctx *c, double *Ut)
{
  if (!c || !Ut) return set_err(FX_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(Ut, c->Ut, vol_doubles(c->cfg) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
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
  q_stage(c, VMODE_RHS, rk);
  HIPCHK(hipGetLastError());
  if (Ut_host) return fx_ut_download(c, Ut_host);
  return FX_OK;
}

int fx_volint(fx_ctx *c, double *Ut_host)
{
  int rc = ready(c);
  if (rc) return rc;
  RKArgs rk{};
  q_volume(c, VMODE_VOLINT, rk);
  HIPCHK(hipGetLastError());
  if (Ut_host) return fx_ut_download(c, Ut_host);
  return FX_OK;
}

int fx_lifting(fx_ctx *c, double *gx, double *gy, double *gz)
{
  int rc = ready(c);
  if (rc) return rc;
  if (!c->cfg.parabolic) return set_err(FX_ESTATE, "lifting needs parabolic=1");
  q_lifting(c, 0);
  HIPCHK(hipGetLastError());
  const size_t nv = vol_doubles(c->cfg);
  double *dst[3] = {gx, gy, gz};
  for (int d = 0; d < 3; d++)
    if (dst[d]) HIPCHK(hipMemcpyAsync(dst[d], c->grad + d * nv, nv * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FX_OK;
}

int fx_calc_dt(fx_ctx *c, double CFLscale, double DFLscale, double *dt, int *errType)
{
  int rc = ready(c);
  if (rc) return rc;
  const int N = c->cfg.N;
  const double cfl = CFLscale * cfl_alpha(N) / (2.0 * N + 1.0);
  const double dfl = DFLscale * dfl_alpha(N) / ((2.0 * N + 1.0) * (2.0 * N + 1.0));
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->stream, c->dts, 0.0, 0.0, 1);
  q_dt(c, cfl, dfl, 0);
  HIPCHK(hipGetLastError());
  DtState h;
  HIPCHK(hipMemcpyAsync(&h, c->dts, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt) *dt = h.dt;
  if (errType) *errType = h.errType;
  return FX_OK;
}

int fx_rk_step(fx_ctx *c, double t, double dt)
{
  int rc = ready(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->stream, c->dts, dt, t, 0);
  q_rk_step(c);
  HIPCHK(hipGetLastError());
  return FX_OK;
}

int fx_run(fx_ctx *c, int nSteps, double CFLscale, double DFLscale, double *t, double *dt, int *errType)
{
  int rc = ready(c);
  if (rc) return rc;
  if (nSteps < 0) return set_err(FX_EINVAL, "nSteps < 0");
  const int N = c->cfg.N;
  const double cfl = CFLscale * cfl_alpha(N) / (2.0 * N + 1.0);
  const double dfl = DFLscale * dfl_alpha(N) / ((2.0 * N + 1.0) * (2.0 * N + 1.0));
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, c->stream, c->dts, 0.0, t ? *t : 0.0, 1);
  HIPCHK(hipGetLastError());
  if (!c->graph || c->graph_cfl != cfl || c->graph_dfl != dfl) {
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    q_dt(c, cfl, dfl, 1);
    q_rk_step(c);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->graph = nullptr; return set_err(FX_EHIP, "graph instantiate: %s", hipGetErrorString(e)); }
    c->graph_cfl = cfl; c->graph_dfl = dfl;
  }
  for (int s = 0; s < nSteps; s++) HIPCHK(hipGraphLaunch(c->graph, c->stream));
  DtState h;
  HIPCHK(hipMemcpyAsync(&h, c->dts, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (t) *t = h.t;
  if (dt) *dt = h.dt;
  if (errType) *errType = h.errType;
  return FX_OK;
}

int fx_entropy(fx_ctx *c, double *S, double *dSdU_Ut)
{
  int rc = ready(c);
  if (rc) return rc;
  q_entropy(c);
  HIPCHK(hipGetLastError());
  double h[2];
  HIPCHK(hipMemcpyAsync(h, c->red, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (S) *S = h[0] / (c->cfg.kappa - 1.0);
  if (dSdU_Ut) *dSdU_Ut = h[1];
  return FX_OK;
}

int fx_profile_kernel(fx_ctx *c, int which, int reps, double *avg_ms)
{
  int rc = ready(c);
  if (rc) return rc;
  if (reps < 1 || !avg_ms) return set_err(FX_EINVAL, "reps < 1 or null avg_ms");
  if (which == FX_KERNEL_LIFTING && !c->cfg.parabolic) return set_err(FX_EINVAL, "lifting needs parabolic=1");
  if (which < 0 || which > 4) return set_err(FX_EINVAL, "unknown kernel %d", which);
  // make the inputs of the timed kernel valid: viscous fluxes, face fluxes and a time step
  if (c->cfg.parabolic) q_lifting(c, 1);
  q_surface(c);
  if (which == FX_KERNEL_DT || which == FX_KERNEL_VOLUME_RK)
    q_dt(c, 0.9 * cfl_alpha(c->cfg.N) / (2.0 * c->cfg.N + 1.0), 0.0, 0);
  RKArgs rk;
  rk.dt = &c->dts->dt; rk.A = c->rkA[1]; rk.b = c->rkb[1]; rk.first = 0;
  if (which == FX_KERNEL_VOLUME_RK) {  // Ut_tmp must hold a previous stage: run stage 1 once
    RKArgs r1 = rk; r1.A = 0.0; r1.b = c->rkb[0]; r1.first = 1;
    q_volume(c, VMODE_RK, r1);
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++) {
    switch (which) {
    case FX_KERNEL_VOLUME_RK: q_volume(c, VMODE_RK, rk); break;
    case FX_KERNEL_SURFACE: q_surface(c); break;
    case FX_KERNEL_LIFTING: q_lifting(c, 1); break;
    case FX_KERNEL_DT: q_dt(c, 0.9, 0.9, 0); break;
    default: q_volume(c, VMODE_RHS, rk); break;
    }
  }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *avg_ms = (double)ms / reps;
  return FX_OK;
}

int fx_synchronize(fx_ctx *c)
{
  if (!c) return set_err(FX_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return FX_OK;
}

}  // extern "C"
