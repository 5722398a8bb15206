// Host-side launch layer shared by the single-problem context and the batched server:
// basis operators in constant memory, error reporting, and the per-stage kernel sequence
// over a "Batch" (device arrays of one or more independent meshes sharing launches).
//
// Compiled as part of fx_engine.hip (one device-code unit, so every kernel sees the one
// c_ops table).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/fluxo_gpu.h"
#include "fx_kernels.hpp"
#include "fx_nfvse.hpp"
#include "fx_mhd.hpp"

namespace fx {

// ---- errors ---------------------------------------------------------------------------
inline char *err_buf()
{
  static thread_local char buf[512] = "";
  return buf;
}

inline int set_err(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err_buf(), 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t _e = (x);                                                                       \
    if (_e != hipSuccess)                                                                      \
      return ::fx::set_err(FX_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, \
                           __LINE__);                                                          \
  } while (0)

// ---- basis ----------------------------------------------------------------------------
// Host-side LGL basis (the engine's InitDGbasis, dg.f90:181-252, with the node/weight
// Newton iteration of LegGaussLobNodesAndWeights basis.f90:544-608 and the barycentric
// derivative matrix of PolynomialDerivativeMatrix basis.f90:641-663).
inline void lgl_nodes(int N, double *x, double *w)
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

// Legendre polynomial of degree N normalised by sqrt(N+1/2) (LegendrePolynomialAndDerivative
// basis.f90) and the Gauss nodes/weights (LegendreGaussNodesAndWeights basis.f90:430-497).
inline double legendre_normalized(int N, double x)
{
  if (N == 0) return std::sqrt(0.5);
  double Lm2 = 1.0, Lm1 = x, L = x;
  for (int i = 2; i <= N; i++) {
    L = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
    Lm2 = Lm1; Lm1 = L;
  }
  return L * std::sqrt((double)N + 0.5);
}

inline void gauss_nodes(int N, double *x, double *w)
{
  const double pi = std::acos(-1.0);
  const int n = N + 1;
  for (int i = 0; i < n; i++) {
    double xi = -std::cos((2.0 * i + 1.0) / (2.0 * N + 2.0) * pi), L = 0.0, Ld = 1.0;
    for (int it = 0; it < 50; it++) {
      double Lm2 = 1.0, Lm1 = xi, Ldm2 = 0.0, Ldm1 = 1.0;
      L = xi; Ld = 1.0;
      for (int k = 2; k <= n; k++) {
        L = ((double)(2 * k - 1) * xi * Lm1 - (double)(k - 1) * Lm2) / (double)k;
        Ld = Ldm2 + (double)(2 * k - 1) * Lm1;
        Lm2 = Lm1; Lm1 = L; Ldm2 = Ldm1; Ldm1 = Ld;
      }
      const double dx = -L / Ld;
      xi += dx;
      if (std::fabs(dx) < 1e-15 * std::fabs(xi) + 1e-300) break;
    }
    x[i] = xi;
    w[i] = 2.0 / ((1.0 - xi * xi) * Ld * Ld);
  }
}

inline void make_ops(int N, Ops &o)
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
  // NFVSE: nodal -> modal Legendre projection by Gauss quadrature with 3N/2+2 points
  // (BuildLegendreVdm basis.f90:196-252), sub-cell face metric weights, 1/w
  const int Ng = 3 * N / 2 + 1;
  double xg[3 * MAXN / 2 + 2], wg[3 * MAXN / 2 + 2];
  gauss_nodes(Ng, xg, wg);
  for (int jm = 0; jm < n; jm++)
    for (int c = 0; c < n; c++) {
      double sum = 0.0;
      for (int g = 0; g <= Ng; g++) {
        // barycentric Lagrange polynomial of node c at xg (a node itself: unit vector)
        int hit = -1;
        for (int k = 0; k < n; k++)
          if (std::fabs(xg[g] - x[k]) <= 4.5e-16 * std::fmax(1.0, std::fabs(x[k]))) hit = k;
        double lc;
        if (hit >= 0) {
          lc = (hit == c) ? 1.0 : 0.0;
        } else {
          double den = 0.0;
          for (int k = 0; k < n; k++) den += wb[k] / (xg[g] - x[k]);
          lc = (wb[c] / (xg[g] - x[c])) / den;
        }
        sum += legendre_normalized(jm, xg[g]) * wg[g] * lc;
      }
      o.sVdm[jm][c] = sum;
    }
  for (int f = 0; f < n; f++)
    for (int m = 0; m < n; m++) {
      double c = 0.0;
      for (int l = 0; l <= f; l++) c += w[l] * o.D[l][m];
      o.Csub[f][m] = c;
    }
  for (int i = 0; i < n; i++) o.sw[i] = 1.0 / w[i];
  for (int i = 0; i < n; i++) {
    double r = 0.0;
    for (int l = 0; l < n; l++) r += o.Ds[i][l];
    o.DsRow[i] = r;
  }
}

// Operators of degree N into constant memory of the current device (once per process).
inline int ensure_ops(int N)
{
  static bool ready[MAXN + 1] = {false};
  if (ready[N]) return FX_OK;
  Ops o;
  make_ops(N, o);
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ops), &o, sizeof(Ops), sizeof(Ops) * N, hipMemcpyHostToDevice));
  ready[N] = true;
  return FX_OK;
}

// ---- time integration constants ---------------------------------------------------------
inline void rk_coefs(double *A, double *b, double *c)
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
inline double cfl_alpha(int N)
{
  static const double a[15] = {4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830,
                               1.5682, 1.4692, 1.3849, 1.3106, 1.2454, 1.1880, 1.1362};
  return a[(N < 15 ? N : 15) - 1];
}
inline double dfl_alpha(int N)
{
  static const double a[10] = {4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17};
  return a[(N < 10 ? N : 10) - 1] * 1.853;
}
// FillCFL_DFL (timedisc.f90:697-741): the scaled CFL and DFL numbers of CALCTIMESTEP
inline double cfl_number(int N, double CFLscale) { return CFLscale * cfl_alpha(N) / (2.0 * N + 1.0); }
inline double dfl_number(int N, double DFLscale)
{
  return DFLscale * dfl_alpha(N) / ((2.0 * N + 1.0) * (2.0 * N + 1.0));
}

// Process-wide tile choice (index into VolTiles<N>::T) of the LSERK volume kernel for each
// degree, Euler/Navier-Stokes and precision: set by the tuner, read when a context or server
// pool is created.
inline int &tuned_tile(int N, int parabolic, int precision)
{
  static int t[MAXN + 1][2][2] = {};
  return t[N][parabolic ? 1 : 0][precision == FX_PREC_FP32_SPLIT ? 1 : 0];
}
inline int tile_epb(int N, int idx)
{
  switch (N) {
  case 2: return VolTiles<2>::T[idx];
  case 3: return VolTiles<3>::T[idx];
  case 4: return VolTiles<4>::T[idx];
  case 5: return VolTiles<5>::T[idx];
  default: return 0;
  }
}

// ---- batch ----------------------------------------------------------------------------
// Device arrays of one launch domain.  cfg.nElems / cfg.nSides are the launch extents;
// with eprob == nullptr all elements belong to problem 0 (st[0]), otherwise eprob maps
// elements to problems (-1: unused slot) and rfirst/rcount give each problem's
// contiguous element range for the per-problem time-step minimum.
struct Batch {
  fx_config cfg{};
  int n = 0, n2 = 0, n3 = 0;
  hipStream_t stream = nullptr;
  double *U = nullptr, *Ut = nullptr, *Uttmp = nullptr;
  double *Mf = nullptr, *Mg = nullptr, *Mh = nullptr, *sJ = nullptr;
  float *Mf32 = nullptr, *Mg32 = nullptr, *Mh32 = nullptr;  // split precision only
  // viscous terms formed by the lifting kernel: Uv (NVV per node), FvN (2 NVV per face point)
  double *Flux = nullptr, *Uv = nullptr, *FvN = nullptr, *grad = nullptr;
  double4 *nvse = nullptr;
  int *e2s = nullptr, *s2e = nullptr;
  int *eprob = nullptr;
  double *dtElem = nullptr;
  int *errElem = nullptr;
  DtState *st = nullptr;
  int *rfirst = nullptr, *rcount = nullptr;
  int nProb = 1;
  int vol_tile = 0;  // index into VolTiles<N>::T for the LSERK volume kernel
  // NFVSE (cfg.nfvse): blending coefficient buffers (two for the propagation sweeps) and
  // the sub-cell finite-volume volume term
  double *alpha = nullptr, *alpha2 = nullptr, *Ffv = nullptr;
  // GLM-MHD (cfg.nVar == 9): slave-side face fluxes (Flux holds the master side) and the
  // per-element dt of a unit wave speed
  double *FluxS = nullptr, *dtch1Elem = nullptr;

  bool mhd() const { return cfg.nVar == 9; }
  MhdArgs mhd_args() const
  {
    MhdArgs a;
    a.kappa = cfg.kappa;
    a.smu0 = 1.0 / cfg.mu_0;
    return a;
  }
  double rkA[5], rkb[5], rkc[5];

  void set_degree(int N)
  {
    cfg.N = N;
    n = N + 1; n2 = n * n; n3 = n2 * n;
    rk_coefs(rkA, rkb, rkc);
  }
};

template <int N, int VF, bool PARA, int MODE, int EPB>
void launch_volume_e(const Batch &b, RKArgs rk)
{
  using G = VolGeom<N, EPB>;
  const int nb = (b.cfg.nElems + EPB - 1) / EPB;
  if (nb <= 0) return;
  if constexpr (has_fp32(VF)) {
    if (b.cfg.precision == FX_PREC_FP32_SPLIT) {
      hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, float, EPB>), dim3(nb), dim3(3 * G::W), 0, b.stream, b.U,
                         b.Mf32, b.Mg32, b.Mh32, b.sJ, b.e2s, b.Flux, b.Uv, b.Ut, b.Uttmp, rk, b.cfg.nElems,
                         b.cfg.kappa);
      return;
    }
  }
  hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, double, EPB>), dim3(nb), dim3(3 * G::W), 0, b.stream, b.U,
                       b.Mf, b.Mg, b.Mh, b.sJ, b.e2s, b.Flux, b.Uv, b.Ut, b.Uttmp, rk, b.cfg.nElems,
                       b.cfg.kappa);
}

// The LSERK form with the Ranocha flux has every tile of VolTiles<N>; other forms use the
// default tile.
template <int N, int VF, bool PARA, int MODE>
void launch_volume_t(const Batch &b, RKArgs rk)
{
  constexpr int T0 = VolTiles<N>::T[0], T1 = VolTiles<N>::T[1], T2 = VolTiles<N>::T[2];
  if (MODE == VMODE_RK && VF == 32) {
    if (b.vol_tile == 1) { launch_volume_e<N, VF, PARA, MODE, T1>(b, rk); return; }
    if (b.vol_tile == 2) { launch_volume_e<N, VF, PARA, MODE, T2>(b, rk); return; }
  }
  launch_volume_e<N, VF, PARA, MODE, T0>(b, rk);
}

template <int N, int VF>
void launch_volume_v(const Batch &b, int mode, RKArgs rk)
{
  const bool para = b.cfg.parabolic != 0;
  if (mode == VMODE_VOLINT) launch_volume_t<N, VF, false, VMODE_VOLINT>(b, rk);
  else if (mode == VMODE_RK) {
    if (para) launch_volume_t<N, VF, true, VMODE_RK>(b, rk);
    else launch_volume_t<N, VF, false, VMODE_RK>(b, rk);
  } else {
    if (para) launch_volume_t<N, VF, true, VMODE_RHS>(b, rk);
    else launch_volume_t<N, VF, false, VMODE_RHS>(b, rk);
  }
}

template <int N>
void launch_volume_n(const Batch &b, int mode, RKArgs rk)
{
  switch (b.cfg.volumeFlux) {
  case 1: launch_volume_v<N, 1>(b, mode, rk); break;
  case 2: launch_volume_v<N, 2>(b, mode, rk); break;
  case 3: launch_volume_v<N, 3>(b, mode, rk); break;
  case 4: launch_volume_v<N, 4>(b, mode, rk); break;
  case 5: launch_volume_v<N, 5>(b, mode, rk); break;
  case 6: launch_volume_v<N, 6>(b, mode, rk); break;
  case 7: launch_volume_v<N, 7>(b, mode, rk); break;
  case 8: launch_volume_v<N, 8>(b, mode, rk); break;
  case 9: launch_volume_v<N, 9>(b, mode, rk); break;
  case 10: launch_volume_v<N, 10>(b, mode, rk); break;
  case 32: launch_volume_v<N, 32>(b, mode, rk); break;
  default: launch_volume_v<N, 0>(b, mode, rk); break;
  }
}

template <int N, int RS>
void launch_surface_t(const Batch &b)
{
  constexpr int T = SurfGeom<N>::T;
  const long long tot = (long long)b.cfg.nSides * b.n2;
  const int nb = (int)((tot + T - 1) / T);
  if (nb <= 0) return;
  if (b.cfg.parabolic)
    hipLaunchKernelGGL((k_surface<N, RS, true>), dim3(nb), dim3(T), 0, b.stream, b.U, b.FvN, b.s2e,
                       b.nvse, b.Flux, b.cfg.nSides, b.cfg.kappa);
  else
    hipLaunchKernelGGL((k_surface<N, RS, false>), dim3(nb), dim3(T), 0, b.stream, b.U, b.FvN, b.s2e,
                       b.nvse, b.Flux, b.cfg.nSides, b.cfg.kappa);
}

template <int N>
void launch_surface_n(const Batch &b)
{
  switch (b.cfg.riemann) {
  case 0: launch_surface_t<N, 0>(b); break;
  case 4: launch_surface_t<N, 4>(b); break;
  case 2: launch_surface_t<N, 2>(b); break;
  case 22: launch_surface_t<N, 22>(b); break;
  case 23: launch_surface_t<N, 23>(b); break;
  case 24: launch_surface_t<N, 24>(b); break;
  case 3:  // Roe's central part follows the volume flux (riemann.f90:836-856)
    switch (b.cfg.volumeFlux) {
    case 2: launch_surface_t<N, 302>(b); break;
    case 3: launch_surface_t<N, 303>(b); break;
    case 4: launch_surface_t<N, 308>(b); break;
    default: launch_surface_t<N, 3>(b); break;
    }
    break;
  case 5: launch_surface_t<N, 5>(b); break;
  case 6: launch_surface_t<N, 6>(b); break;
  case 7: launch_surface_t<N, 7>(b); break;
  case 8: launch_surface_t<N, 8>(b); break;
  case 10: launch_surface_t<N, 10>(b); break;
  case 11: launch_surface_t<N, 11>(b); break;
  case 12: launch_surface_t<N, 12>(b); break;
  case 13: launch_surface_t<N, 13>(b); break;
  case 14: launch_surface_t<N, 14>(b); break;
  case 15: launch_surface_t<N, 15>(b); break;
  case 17: launch_surface_t<N, 17>(b); break;
  case 18: launch_surface_t<N, 18>(b); break;
  case 19: launch_surface_t<N, 19>(b); break;
  case 20: launch_surface_t<N, 20>(b); break;
  case 21: launch_surface_t<N, 21>(b); break;
  case 9: launch_surface_t<N, 9>(b); break;
  case 16: launch_surface_t<N, 16>(b); break;
  case 32: launch_surface_t<N, 32>(b); break;
  case 33: launch_surface_t<N, 33>(b); break;
  case 44: launch_surface_t<N, 44>(b); break;
  default: launch_surface_t<N, 1>(b); break;
  }
}

template <int N, class EQ, int LIFT>
void launch_lifting_l(const Batch &b, int out, const EQ &eq)
{
  const int epb = LiftGeom<N>::EPB;
  const int nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  const size_t nv = (size_t)EQ::NV * b.n3 * b.cfg.nElems;
  if (out == 0)
    hipLaunchKernelGGL((k_lifting<N, 0, EQ, LIFT>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                       b.e2s, b.s2e, b.nvse, b.grad, b.grad + nv, b.grad + 2 * nv, b.eprob, b.cfg.nElems, eq,
                       b.cfg.etaBR2);
  else
    hipLaunchKernelGGL((k_lifting<N, 1, EQ, LIFT>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                       b.e2s, b.s2e, b.nvse, b.Uv, b.FvN, nullptr, b.eprob, b.cfg.nElems, eq, b.cfg.etaBR2);
}

template <int N, class EQ>
void launch_lifting_e(const Batch &b, int out, const EQ &eq)
{
  if (b.cfg.lifting == 2) launch_lifting_l<N, EQ, 2>(b, out, eq);
  else launch_lifting_l<N, EQ, 1>(b, out, eq);
}

template <int N>
void launch_lifting_n(const Batch &b, int out)
{
  if (b.mhd()) {
    MhdLift eq;
    eq.kappa = b.cfg.kappa;
    eq.smu0 = 1.0 / b.cfg.mu_0;
    eq.mu = b.cfg.mu0;
    eq.etasmu0 = b.cfg.eta * eq.smu0;
    eq.KappasPr = b.cfg.kappa / b.cfg.Pr;
    launch_lifting_e<N>(b, out, eq);
  } else {
    NsLift eq;
    eq.kappa = b.cfg.kappa;
    eq.mu0 = b.cfg.mu0;
    eq.KappasPr = b.cfg.kappa / b.cfg.Pr;
    launch_lifting_e<N>(b, out, eq);
  }
}

template <int N>
void launch_dt_n(const Batch &b, double cfl, double dfl)
{
  const int epb = DtGeom<N>::EPB;
  const int nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  const double KsPr = b.cfg.kappa / b.cfg.Pr;
  const double KsPr_max = (4.0 / 3.0 > KsPr) ? 4.0 / 3.0 : KsPr;
  hipLaunchKernelGGL((k_calc_dt<N>), dim3(nb), dim3(DtGeom<N>::T), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                     b.dtElem, b.errElem, b.eprob, b.cfg.nElems, b.cfg.kappa, cfl, dfl, b.cfg.mu0, KsPr_max,
                     b.cfg.parabolic);
}

template <int N, int MODE, bool PARA>
void launch_mhd_volume_p(const Batch &b, RKArgs rk)
{
  constexpr int EPB = MhdGeom<N>::EPB, W = MhdGeom<N>::W;
  const int nb = (b.cfg.nElems + EPB - 1) / EPB;
  if (nb <= 0) return;
  if (b.cfg.precision == FX_PREC_FP32_SPLIT)
    hipLaunchKernelGGL((k_mhd_volume<N, MODE, float, PARA>), dim3(nb), dim3(3 * W), 0, b.stream, b.U, b.Mf32,
                       b.Mg32, b.Mh32, b.sJ, b.e2s, b.Flux, b.FluxS, b.Uv, b.Ut, b.Uttmp, rk, b.cfg.nElems,
                       b.mhd_args(), b.st);
  else
    hipLaunchKernelGGL((k_mhd_volume<N, MODE, double, PARA>), dim3(nb), dim3(3 * W), 0, b.stream, b.U, b.Mf, b.Mg,
                       b.Mh, b.sJ, b.e2s, b.Flux, b.FluxS, b.Uv, b.Ut, b.Uttmp, rk, b.cfg.nElems, b.mhd_args(),
                       b.st);
}

template <int N, int MODE>
void launch_mhd_volume_t(const Batch &b, RKArgs rk)
{
  if (b.cfg.parabolic) launch_mhd_volume_p<N, MODE, true>(b, rk);
  else launch_mhd_volume_p<N, MODE, false>(b, rk);
}

template <int N>
void launch_mhd_volume_n(const Batch &b, int mode, RKArgs rk)
{
  if (mode == VMODE_RK) launch_mhd_volume_t<N, VMODE_RK>(b, rk);
  else if (mode == VMODE_VOLINT) launch_mhd_volume_p<N, VMODE_VOLINT, false>(b, rk);
  else launch_mhd_volume_t<N, VMODE_RHS>(b, rk);
}

template <int N, int RS>
void launch_mhd_surface_r(const Batch &b)
{
  const long long tot = (long long)b.cfg.nSides * b.n2;
  const int nb = (int)((tot + 255) / 256);
  if (nb <= 0) return;
  if (b.cfg.parabolic)
    hipLaunchKernelGGL((k_mhd_surface<N, RS, true>), dim3(nb), dim3(256), 0, b.stream, b.U, b.FvN, b.s2e, b.nvse,
                       b.Flux, b.FluxS, b.cfg.nSides, b.mhd_args(), b.st);
  else
    hipLaunchKernelGGL((k_mhd_surface<N, RS, false>), dim3(nb), dim3(256), 0, b.stream, b.U, b.FvN, b.s2e, b.nvse,
                       b.Flux, b.FluxS, b.cfg.nSides, b.mhd_args(), b.st);
}

template <int N>
void launch_mhd_surface_n(const Batch &b)
{
  switch (b.cfg.riemann) {
    case 0: launch_mhd_surface_r<N, 0>(b); break;
    case 10: launch_mhd_surface_r<N, 10>(b); break;
    case 11: launch_mhd_surface_r<N, 11>(b); break;
    default: launch_mhd_surface_r<N, 1>(b); break;
  }
}

template <int N>
void launch_mhd_dt_n(const Batch &b, double cfl, double dfl)
{
  const int epb = DtGeom<N>::EPB, nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  MhdDtArgs da;
  da.cfl = cfl;
  da.dfl = dfl;
  const double KsPr = b.cfg.kappa / b.cfg.Pr;
  da.muk = b.cfg.mu0 * ((4.0 / 3.0 > KsPr) ? 4.0 / 3.0 : KsPr);
  da.diffc = b.cfg.eta * (1.0 / b.cfg.mu_0);
  da.parabolic = b.cfg.parabolic;
  hipLaunchKernelGGL((k_mhd_calc_dt<N>), dim3(nb), dim3(DtGeom<N>::T), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                     b.dtElem, b.dtch1Elem, b.errElem, b.cfg.nElems, b.mhd_args(), da);
}

#define FX_DISPATCH_N(N_, fn, ...)             \
  switch (N_) {                                \
  case 2: fn<2>(__VA_ARGS__); break;           \
  case 3: fn<3>(__VA_ARGS__); break;           \
  case 4: fn<4>(__VA_ARGS__); break;           \
  case 5: fn<5>(__VA_ARGS__); break;           \
  default: break;                              \
  }

inline void q_volume(const Batch &b, int mode, RKArgs rk)
{
  if (b.mhd()) { FX_DISPATCH_N(b.cfg.N, launch_mhd_volume_n, b, mode, rk); }
  else { FX_DISPATCH_N(b.cfg.N, launch_volume_n, b, mode, rk); }
}
inline void q_surface(const Batch &b)
{
  if (b.mhd()) { FX_DISPATCH_N(b.cfg.N, launch_mhd_surface_n, b); }
  else { FX_DISPATCH_N(b.cfg.N, launch_surface_n, b); }
}
inline void q_lifting(const Batch &b, int out) { FX_DISPATCH_N(b.cfg.N, launch_lifting_n, b, out); }

// CALCTIMESTEP into the problem states (advance=1 also moves t and the step counter)
inline void q_dt(const Batch &b, double cfl, double dfl, int advance)
{
  if (b.mhd()) {  // single problem only
    FX_DISPATCH_N(b.cfg.N, launch_mhd_dt_n, b, cfl, dfl);
    hipLaunchKernelGGL(k_mhd_dt_reduce, dim3(1), dim3(1024), 0, b.stream, b.dtElem, b.dtch1Elem, b.errElem,
                       b.cfg.nElems, b.st, advance, b.cfg.GLM_scale);
    return;
  }
  FX_DISPATCH_N(b.cfg.N, launch_dt_n, b, cfl, dfl);
  if (b.eprob) {
    if (b.nProb > 0)
      hipLaunchKernelGGL(k_dt_reduce_seg, dim3(b.nProb), dim3(256), 0, b.stream, b.dtElem, b.errElem,
                         b.rfirst, b.rcount, b.st, advance);
  } else {
    hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, b.stream, b.dtElem, b.errElem, b.cfg.nElems,
                       b.st, advance);
  }
}

inline NfvseArgs nfvse_args(const fx_config &c)
{
  NfvseArgs a{};
  a.computeAlpha = c.computeAlpha;
  a.threshold = 0.5 * std::pow(10.0, -1.8 * std::pow(c.N + 1.0, 0.25));
  a.sharpness = std::log((1.0 - 1.e-4) / 1.e-4);
  a.alpha_min = c.alpha_min;
  a.alpha_max = c.alpha_max;
  a.shockBlendCoef = c.shockBlendCoef;
  return a;
}

template <int N>
void launch_nfvse_alpha_n(const Batch &b)
{
  const int epb = LiftGeom<N>::EPB, nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  hipLaunchKernelGGL((k_nfvse_alpha<N>), dim3(nb), dim3(256), 0, b.stream, b.U, b.eprob, b.alpha2,
                     b.cfg.nElems, b.cfg.kappa, nfvse_args(b.cfg));
}

template <int N, int RS>
void launch_nfvse_fv_t(const Batch &b, const double *alpha)
{
  const int nb = (b.cfg.nElems + FvGeom<N>::EPB - 1) / FvGeom<N>::EPB;
  if (nb <= 0) return;
  hipLaunchKernelGGL((k_nfvse_fv<N, RS>), dim3(nb), dim3(3 * FvGeom<N>::W), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh,
                     alpha, b.eprob, b.Ffv, b.cfg.nElems, b.cfg.kappa);
}

template <int N>
void launch_nfvse_fv_n(const Batch &b, const double *alpha)
{
  switch (b.cfg.riemann) {
  case 0: launch_nfvse_fv_t<N, 0>(b, alpha); break;
  case 2: launch_nfvse_fv_t<N, 2>(b, alpha); break;
  case 3:
    switch (b.cfg.volumeFlux) {
    case 2: launch_nfvse_fv_t<N, 302>(b, alpha); break;
    case 3: launch_nfvse_fv_t<N, 303>(b, alpha); break;
    case 4: launch_nfvse_fv_t<N, 308>(b, alpha); break;
    default: launch_nfvse_fv_t<N, 3>(b, alpha); break;
    }
    break;
  case 4: launch_nfvse_fv_t<N, 4>(b, alpha); break;
  case 5: launch_nfvse_fv_t<N, 5>(b, alpha); break;
  case 6: launch_nfvse_fv_t<N, 6>(b, alpha); break;
  case 7: launch_nfvse_fv_t<N, 7>(b, alpha); break;
  case 8: launch_nfvse_fv_t<N, 8>(b, alpha); break;
  case 10: launch_nfvse_fv_t<N, 10>(b, alpha); break;
  case 11: launch_nfvse_fv_t<N, 11>(b, alpha); break;
  case 12: launch_nfvse_fv_t<N, 12>(b, alpha); break;
  case 20: launch_nfvse_fv_t<N, 20>(b, alpha); break;
  case 22: launch_nfvse_fv_t<N, 22>(b, alpha); break;
  case 23: launch_nfvse_fv_t<N, 23>(b, alpha); break;
  case 24: launch_nfvse_fv_t<N, 24>(b, alpha); break;
  case 33: launch_nfvse_fv_t<N, 33>(b, alpha); break;
  case 44: launch_nfvse_fv_t<N, 44>(b, alpha); break;
  case 9: launch_nfvse_fv_t<N, 9>(b, alpha); break;
  case 32: launch_nfvse_fv_t<N, 32>(b, alpha); break;
  default: launch_nfvse_fv_t<N, 1>(b, alpha); break;
  }
}

// CalcBlendingCoefficient + propagation sweeps; returns the buffer holding the final alpha.
inline const double *q_nfvse_alpha(const Batch &b)
{
  FX_DISPATCH_N(b.cfg.N, launch_nfvse_alpha_n, b);
  double *src = b.alpha2, *dst = b.alpha;
  const int nb = (b.cfg.nElems + 255) / 256;
  for (int sw = 0; sw < b.cfg.spacePropSweeps && nb > 0; sw++) {
    hipLaunchKernelGGL(k_nfvse_prop, dim3(nb), dim3(256), 0, b.stream, src, dst, b.e2s, b.s2e, b.eprob,
                       b.cfg.nElems, b.cfg.spacePropFactor);
    double *t = src; src = dst; dst = t;
  }
  return src;
}

// one RHS evaluation, followed by the RK stage update in VMODE_RK
inline void q_stage(const Batch &b, int mode, RKArgs rk)
{
  if (b.cfg.parabolic) q_lifting(b, 1);
  if (b.cfg.nfvse && mode != VMODE_VOLINT) {
    const double *a = q_nfvse_alpha(b);
    FX_DISPATCH_N(b.cfg.N, launch_nfvse_fv_n, b, a);
    rk.alpha = a;
    rk.Ffv = b.Ffv;
  }
  q_surface(b);
  q_volume(b, mode, rk);
}

// the five stages of one TimeStepByLSERKW2 step with dt from the problem states
inline void q_rk_step(const Batch &b)
{
  for (int s = 0; s < 5; s++) {
    RKArgs rk;
    rk.st = b.st;
    rk.eprob = b.eprob;
    rk.A = b.rkA[s];
    rk.b = b.rkb[s];
    rk.first = (s == 0);
    rk.alpha = nullptr;
    rk.Ffv = nullptr;
    q_stage(b, VMODE_RK, rk);
  }
}

// Connectivity check of one conforming periodic mesh (every side an inner side): every
// index the kernels derive from ElemToSide/SideToElem must stay inside the mesh.
inline int check_connectivity(const int *ElemToSide, const int *SideToElem, int E, int S)
{
  for (int s = 0; s < S; s++) {
    const int *t = SideToElem + 5 * s;
    if (t[0] < 1 || t[0] > E || t[1] < 1 || t[1] > E || t[2] < 1 || t[2] > 6 || t[3] < 1 || t[3] > 6 ||
        t[4] < 1 || t[4] > 4)
      return set_err(FX_EINVAL, "SideToElem(:,%d) invalid (only inner conforming sides are supported)", s + 1);
    const int *am = ElemToSide + 2 * (6 * (t[0] - 1) + t[2] - 1);
    const int *as = ElemToSide + 2 * (6 * (t[1] - 1) + t[3] - 1);
    if (am[0] != s + 1 || am[1] != 0 || as[0] != s + 1 || as[1] != t[4])
      return set_err(FX_EINVAL, "ElemToSide inconsistent with SideToElem at side %d", s + 1);
  }
  for (int e = 0; e < E; e++)
    for (int l = 0; l < 6; l++) {
      const int *a = ElemToSide + 2 * (6 * e + l);
      if (a[0] < 1 || a[0] > S || a[1] < 0 || a[1] > 4)
        return set_err(FX_EINVAL, "ElemToSide(:,%d,%d) invalid", l + 1, e + 1);
    }
  return FX_OK;
}

// Physics options shared by contexts and servers.
inline int check_physics(int volumeFlux, int riemann, int precision, double kappa, double Pr, int parabolic)
{
  if (volumeFlux < 0 || (volumeFlux > 10 && volumeFlux != 32))
    return set_err(FX_EINVAL, "VolumeFlux=%d not supported (0..10, 32)", volumeFlux);
  if (precision == FX_PREC_FP32_SPLIT && !has_fp32(volumeFlux))
    return set_err(FX_EINVAL, "VolumeFlux=%d runs in FP64 only (FP32-split: 0, 5, 10, 32)", volumeFlux);
  if (riemann < 0 || (riemann > 24 && riemann != 32 && riemann != 33 && riemann != 44))
    return set_err(FX_EINVAL, "Riemann=%d not supported (0-24,32,33,44)", riemann);
  if (precision != FX_PREC_FP64 && precision != FX_PREC_FP32_SPLIT)
    return set_err(FX_EINVAL, "precision %d not supported (0 FP64, 1 FP32-split)", precision);
  if (!(kappa > 1.0)) return set_err(FX_EINVAL, "kappa must be > 1");
  if (parabolic && !(Pr > 0.0)) return set_err(FX_EINVAL, "Pr must be > 0");
  return FX_OK;
}

inline int check_nfvse(const fx_config *c)
{
  if (!c->nfvse) return FX_OK;
  if (c->nfvse != 1) return set_err(FX_EINVAL, "nfvse=%d (0 off, 1 first-order sub-cell FV)", c->nfvse);
  // the sub-cell interfaces take point-wise solvers; those with a face-wide wave speed
  // (13-19, 21) would need it over each sub-cell interface plane
  if (c->riemann >= 13 && c->riemann <= 21 && c->riemann != 20)
    return set_err(FX_EINVAL, "NFVSE sub-cell fluxes: Riemann 0-12, 20, 22-24, 32, 33, 44 (not 13-19, 21)");
  if (c->computeAlpha != 1 && c->computeAlpha != 3)
    return set_err(FX_EINVAL, "ComputeAlpha=%d not supported (1 Persson-Peraire, 3 fixed)", c->computeAlpha);
  if (c->spacePropSweeps < 0 || c->spacePropSweeps > 16) return set_err(FX_EINVAL, "SpacePropSweeps out of 0..16");
  if (!(c->alpha_max > 0.0) || c->alpha_max > 1.0 || c->alpha_min < 0.0)
    return set_err(FX_EINVAL, "alpha_max must be in (0,1], alpha_min >= 0");
  return FX_OK;
}

}  // namespace fx
