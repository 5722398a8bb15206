// Host-side launch layer shared by the single-problem context and the batched server:
// basis operators in constant memory, error reporting, and the per-stage kernel sequence
// over a "Batch" (device arrays of one or more independent meshes sharing launches).
//
// Included by every compilation unit of the library; the kernel launches themselves are in
// fx_launch_n.hpp (per degree) and, for the few non-template kernels, in fx_engine.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

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

// Mortar operators of one degree on LGL nodes (InitMortarBase, mortar.f90:63-110, with
// MortarBasis_BigToSmall and MortarBasis_SmallToBig_Projection / _Collocation,
// mortar_vars.f90:70-196) as the kernels read them: MInt[k][l][p] = l_l(xi^k_p) (Lagrange
// polynomial l of the big side at node p of the half k: 0 = [-1,0], 1 = [0,1]), then
// MProj[k][l][p], the weight of node l of half k in node p of the big side: whichMortar 0 the
// L2 projection through the orthonormal Legendre modes with N+1 Gauss points, 1 collocation
// (w_l / w_p l_p(xi^k_l)).
inline void make_mortar_ops(int N, int whichMortar, double *mops)
{
  const int n = N + 1;
  double x[MAXN + 1], w[MAXN + 1], wb[MAXN + 1];
  lgl_nodes(N, x, w);
  for (int i = 0; i < n; i++) wb[i] = 1.0;
  for (int i = 1; i <= N; i++)
    for (int j = 0; j < i; j++) { wb[j] *= (x[j] - x[i]); wb[i] *= (x[i] - x[j]); }
  for (int i = 0; i < n; i++) wb[i] = 1.0 / wb[i];
  // barycentric Lagrange polynomials of the LGL nodes at y (a node itself: unit vector,
  // LagrangeInterpolationPolys basis.f90:732-767)
  auto lagrange = [&](double y, double *L) {
    int hit = -1;
    for (int k = 0; k < n; k++) {
      const double a = y, b = x[k], tol = 2.220446049250313e-16;
      const bool eq = (a == 0.0 || b == 0.0) ? std::fabs(a - b) <= 2.0 * tol
                                             : (std::fabs(a - b) <= tol * std::fabs(a) && std::fabs(a - b) <= tol * std::fabs(b));
      if (eq) hit = k;
    }
    if (hit >= 0) {
      for (int k = 0; k < n; k++) L[k] = k == hit ? 1.0 : 0.0;
      return;
    }
    double sum = 0.0;
    for (int k = 0; k < n; k++) { L[k] = wb[k] / (y - x[k]); sum += L[k]; }
    for (int k = 0; k < n; k++) L[k] = L[k] / sum;
  };
  double V[2][MAXN + 1][MAXN + 1];  // V[k][p][l] = l_l(xi^k_p)
  for (int k = 0; k < 2; k++)
    for (int p = 0; p < n; p++) {
      const double y = k == 0 ? 0.5 * (x[p] - 1.0) : 0.5 * (x[p] + 1.0);
      lagrange(y, V[k][p]);
    }
  double *MI = mops, *MP = mops + 2 * n * n;
  for (int k = 0; k < 2; k++)
    for (int l = 0; l < n; l++)
      for (int p = 0; p < n; p++) MI[(k * n + l) * n + p] = V[k][p][l];
  if (whichMortar == 1) {
    for (int k = 0; k < 2; k++)
      for (int l = 0; l < n; l++)
        for (int p = 0; p < n; p++) MP[(k * n + l) * n + p] = w[l] / w[p] * V[k][l][p];
    return;
  }
  double xg[MAXN + 1], wg[MAXN + 1];
  gauss_nodes(N, xg, wg);
  double VGP[MAXN + 1][MAXN + 1];  // wg_g l_j(xg_g)
  for (int g = 0; g < n; g++) {
    double L[MAXN + 1];
    lagrange(xg[g], L);
    for (int j = 0; j < n; j++) VGP[g][j] = wg[g] * L[j];
  }
  for (int k = 0; k < 2; k++) {
    // T(jm, l) = sum_g phi_jm(xg_g on half k) VGP(g, l);  M_k0(p, l) = sum_jm phi_jm(x_p) T(jm, l)
    double T[MAXN + 1][MAXN + 1];
    for (int jm = 0; jm < n; jm++)
      for (int l = 0; l < n; l++) {
        double sum = 0.0;
        for (int g = 0; g < n; g++)
          sum += legendre_normalized(jm, k == 0 ? 0.5 * (xg[g] - 1.0) : 0.5 * (xg[g] + 1.0)) * VGP[g][l];
        T[jm][l] = sum;
      }
    for (int p = 0; p < n; p++)
      for (int l = 0; l < n; l++) {
        double sum = 0.0;
        for (int jm = 0; jm < n; jm++) sum += legendre_normalized(jm, x[p]) * T[jm][l];
        MP[(k * n + l) * n + p] = sum;
      }
  }
}

// Every compilation unit with kernels keeps its own c_ops and registers, at load time, a
// function that fills its copy for one degree on the current device.
using OpsUploader = int (*)(const Ops &o, int N);
int register_ops_uploader(OpsUploader f);  // fx_engine.hip
int upload_ops_everywhere(const Ops &o, int N);

// Operators of degree N into constant memory of the current device (once per process).
inline int ensure_ops(int N)
{
  static bool ready[MAXN + 1] = {false};
  if (ready[N]) return FX_OK;
  Ops o;
  make_ops(N, o);
  const int rc = upload_ops_everywhere(o, N);
  if (rc) return rc;
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

// blocks of a lifting launch (LiftGeom<N>::EPB elements each)
inline int lift_blocks(int N, int nElems)
{
  const int n3 = (N + 1) * (N + 1) * (N + 1), epb = 256 / n3 > 0 ? 256 / n3 : 1;
  return (nElems + epb - 1) / epb;
}

// max(4/3, kappa/Pr): the viscous eigenvalue factor of CALCTIMESTEP (calctimestep.f90:85)
inline double ks_pr_max(const fx_config &c)
{
  const double KsPr = c.kappa / c.Pr;
  return (4.0 / 3.0 > KsPr) ? 4.0 / 3.0 : KsPr;
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
  // non-conforming interfaces: nMortar big mortar sides (sides 1..nMortar), their description
  // (MORT_INTS ints each), the interpolation / projection operators and the master states of
  // the small sides (face-point layout of Flux)
  int nMortar = 0;
  int *mort = nullptr, *mside = nullptr;  // mside: mortar of each small side, -1 elsewhere
  double *mops = nullptr, *Umort = nullptr;
  // BR1 lifting with mortars: the small sides' lifting fluxes (3 nVar per face point), their
  // projection onto the big sides, and the big sides' primitive gradients (3 nVar each)
  double *LiftS = nullptr, *LiftM = nullptr, *GradM = nullptr;

  bool mhd() const { return cfg.nVar == 9; }
  MhdArgs mhd_args() const
  {
    MhdArgs a;
    a.kappa = cfg.kappa;
    a.smu0 = 1.0 / cfg.mu_0;
    return a;
  }
  double rkA[5], rkb[5], rkc[5];
  int2 *ftab = nullptr;  // face-point table of the lifting kernel (k_face_table)

  void set_degree(int N)
  {
    cfg.N = N;
    n = N + 1; n2 = n * n; n3 = n2 * n;
    rk_coefs(rkA, rkb, rkc);
  }
};

// The face-point table holds volume nodes and 2 x side points as 32-bit integers.
inline bool face_table_fits(int N, long long nElems, long long nSides)
{
  const long long n = N + 1;
  return nElems * n * n * n < 2147483647LL && 2 * nSides * n * n + 1 < 2147483647LL;
}

// ---- per-degree launchers ---------------------------------------------------------------
// Defined in fx_launch_n.hpp and instantiated for one degree per compilation unit
// (fx_tu.hip, built once per N and kernel group), so the kernels of different degrees
// compile in parallel.  Each unit holds its own constant-memory copy of the operators
// (c_ops); ensure_ops fills every unit's copy.
template <int N> void launch_volume_n(const Batch &b, int mode, RKArgs rk);
template <int N> void launch_surface_n(const Batch &b);
// the face-point table of elements e0 .. e0 + nE - 1 from b.e2s / b.s2e (after every upload of
// their connectivity)
template <int N> void launch_face_table_n(const Batch &b, int e0, int nE);
// U_Mortar into b.Umort, Flux_Mortar into the big sides of b.Flux (meshes with mortars)
template <int N> void launch_mortar_u_n(const Batch &b);
template <int N> void launch_flux_mortar_n(const Batch &b);
// dtf (Navier-Stokes, viscous form): the kernel also forms CALCTIMESTEP's dtElem / errElem
template <int N> void launch_lifting_n(const Batch &b, int out, const DtFuse *dtf = nullptr);
template <int N> void launch_dt_n(const Batch &b, double cfl, double dfl);
template <int N> void launch_mhd_volume_n(const Batch &b, int mode, RKArgs rk);
template <int N> void launch_mhd_surface_n(const Batch &b);
template <int N> void launch_mhd_dt_n(const Batch &b, double cfl, double dfl);
template <int N> void launch_nfvse_alpha_n(const Batch &b);
template <int N> void launch_nfvse_fv_n(const Batch &b, const double *alpha);
// SurfInt alone: Ut (b.Ut) += the surface integral of the given master/slave fluxes
template <int N> void launch_surfint_n(const Batch &b, const double *Fm, const double *Fs);
// CalcEntropy partial sums (2 per 256 nodes) into part, their sum into red[0..1]
template <int N> void launch_entropy_n(const Batch &b, double *part, double *red);

#define FX_DISPATCH_N(N_, fn, ...)             \
  switch (N_) {                                \
  case 2: fn<2>(__VA_ARGS__); break;           \
  case 3: fn<3>(__VA_ARGS__); break;           \
  case 4: fn<4>(__VA_ARGS__); break;           \
  case 5: fn<5>(__VA_ARGS__); break;           \
  default: break;                              \
  }

#ifdef FX_KERNEL_HOME  // the launches of the non-template kernels live in one unit
inline void q_volume(const Batch &b, int mode, RKArgs rk)
{
  if (b.mhd()) { FX_DISPATCH_N(b.cfg.N, launch_mhd_volume_n, b, mode, rk); }
  else { FX_DISPATCH_N(b.cfg.N, launch_volume_n, b, mode, rk); }
}
inline void q_surface(const Batch &b)
{
  if (b.nMortar > 0) { FX_DISPATCH_N(b.cfg.N, launch_mortar_u_n, b); }
  if (b.mhd()) { FX_DISPATCH_N(b.cfg.N, launch_mhd_surface_n, b); }
  else { FX_DISPATCH_N(b.cfg.N, launch_surface_n, b); }
  if (b.nMortar > 0) { FX_DISPATCH_N(b.cfg.N, launch_flux_mortar_n, b); }
}
inline void q_lifting(const Batch &b, int out, const DtFuse *dtf = nullptr)
{
  FX_DISPATCH_N(b.cfg.N, launch_lifting_n, b, out, dtf);
}

// The minimum of the per-element time steps into the problem states (advance=1 also moves t
// and the step counter); count > 0: of the first count entries (per-block minima) instead
inline void q_dt_reduce(const Batch &b, int advance, int count = 0)
{
  if (b.eprob) {
    if (b.nProb > 0)
      hipLaunchKernelGGL(k_dt_reduce_seg, dim3(b.nProb), dim3(256), 0, b.stream, b.dtElem, b.errElem,
                         b.rfirst, b.rcount, b.st, advance);
  } else {
    hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, b.stream, b.dtElem, b.errElem,
                       count > 0 ? count : b.cfg.nElems, b.st, advance);
  }
}

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
  q_dt_reduce(b, advance);
}

#endif  // FX_KERNEL_HOME

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

#ifdef FX_KERNEL_HOME
// CalcBlendingCoefficient + propagation sweeps; returns the buffer holding the final alpha.
inline const double *q_nfvse_alpha(const Batch &b)
{
  FX_DISPATCH_N(b.cfg.N, launch_nfvse_alpha_n, b);
  double *src = b.alpha2, *dst = b.alpha;
  const int nb = (b.cfg.nElems + 255) / 256;
  for (int sw = 0; sw < b.cfg.spacePropSweeps && nb > 0; sw++) {
    hipLaunchKernelGGL(k_nfvse_prop, dim3(nb), dim3(256), 0, b.stream, src, dst, b.e2s, b.s2e, b.eprob,
                       b.cfg.nElems, b.cfg.spacePropFactor, b.mort, b.mside, b.nMortar);
    double *t = src; src = dst; dst = t;
  }
  return src;
}

// CALCTIMESTEP inside the first stage of a step: a single Navier-Stokes problem, whose
// first-stage lifting kernel reads the step's U and metrics at every node anyway
inline bool dt_fusable(const Batch &b) { return !b.mhd() && !b.eprob && b.cfg.parabolic; }
inline DtFuse dt_fuse(const Batch &b, double cfl, double dfl)
{
  DtFuse f;
  f.dtBlock = b.dtElem; f.errBlock = b.errElem;  // one entry per lifting block (fewer than elements)
  f.cfl = cfl; f.dfl = dfl; f.mu0 = b.cfg.mu0; f.KsPr_max = ks_pr_max(b.cfg);
  return f;
}

// one RHS evaluation, followed by the RK stage update in VMODE_RK; dtf (non-null): the
// lifting kernel also forms the step's CALCTIMESTEP, whose minimum (advancing t and the
// step counter) follows it
inline void q_stage(const Batch &b, int mode, RKArgs rk, const DtFuse *dtf = nullptr)
{
  if (b.cfg.parabolic) {
    q_lifting(b, 1, dtf);
    if (dtf) q_dt_reduce(b, 1, lift_blocks(b.cfg.N, b.cfg.nElems));
  }
  if (b.cfg.nfvse && mode != VMODE_VOLINT) {
    const double *a = q_nfvse_alpha(b);
    FX_DISPATCH_N(b.cfg.N, launch_nfvse_fv_n, b, a);
    rk.alpha = a;
    rk.Ffv = b.Ffv;
  }
  q_surface(b);
  q_volume(b, mode, rk);
}

// the five stages of one TimeStepByLSERKW2 step with dt from the problem states, or, with
// dtf (dt_fusable batches), with the step's CALCTIMESTEP formed in its first stage
inline void q_rk_step(const Batch &b, const DtFuse *dtf = nullptr)
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
    q_stage(b, VMODE_RK, rk, s == 0 ? dtf : nullptr);
  }
}

#endif  // FX_KERNEL_HOME

// Connectivity check of one periodic mesh: every index the kernels derive from
// ElemToSide/SideToElem (and, with mortars, from MortarType/MortarInfo) must stay inside the
// mesh.  Without mortars every side is a conforming inner side.  With nMortar > 0 the sides
// 1..nMortar are big mortar sides (master element only, MortarType(2,s) = s) whose 4 (type 1)
// or 2 (types 2, 3) small sides each have a slave element only and belong to exactly one
// mortar, with flip 0 towards it; mort receives the kernels' description of each mortar.
inline int check_connectivity(const int *ElemToSide, const int *SideToElem, int E, int S, int nMortar = 0,
                              const int *MortarType = nullptr, const int *MortarInfo = nullptr,
                              std::vector<int> *mort = nullptr)
{
  if (nMortar < 0 || nMortar >= S || (nMortar > 0 && (!MortarType || !MortarInfo)))
    return set_err(FX_EINVAL, "nMortarSides=%d invalid or mortar arrays missing", nMortar);
  std::vector<int> owner(S, -1);  // mortar of each small side
  if (mort) mort->assign((size_t)MORT_INTS * nMortar, 0);
  for (int m = 0; m < nMortar; m++) {
    const int type = MortarType[2 * m], *mi = MortarInfo + 10 * m, *t = SideToElem + 5 * m;
    if (type < 1 || type > 3 || MortarType[2 * m + 1] != m + 1 || mi[0] != m + 1 || mi[1] != 0)
      return set_err(FX_EINVAL, "side %d: not a big mortar side in MortarType / MortarInfo", m + 1);
    if (t[0] < 1 || t[0] > E || t[2] < 1 || t[2] > 6 || t[1] != -1)
      return set_err(FX_EINVAL, "SideToElem(:,%d) invalid for a big mortar side", m + 1);
    const int *am = ElemToSide + 2 * (6 * (t[0] - 1) + t[2] - 1);
    if (am[0] != m + 1 || am[1] != 0) return set_err(FX_EINVAL, "ElemToSide inconsistent at mortar side %d", m + 1);
    const int ns = type == 1 ? 4 : 2;
    int *d = mort ? mort->data() + (size_t)MORT_INTS * m : nullptr;
    if (d) { d[0] = type; d[1] = t[0] - 1; d[2] = t[2]; d[7] = m + 1; }
    for (int im = 1; im <= 4; im++) {
      const int sid = mi[2 * im], fl = mi[2 * im + 1];
      if (im > ns) continue;
      if (sid <= nMortar || sid > S || fl != 0 || MortarType[2 * (sid - 1)] != 0 || owner[sid - 1] >= 0)
        return set_err(FX_EINVAL, "mortar side %d: small side %d invalid (id %d, flip %d)", m + 1, im, sid, fl);
      owner[sid - 1] = m;
      if (d) d[2 + im] = sid;
    }
  }
  for (int s = nMortar; s < S; s++) {
    const int *t = SideToElem + 5 * s;
    const bool small = owner[s] >= 0;
    if (t[1] < 1 || t[1] > E || t[3] < 1 || t[3] > 6 || t[4] < 1 || t[4] > 4 ||
        (small ? t[0] != -1 : (t[0] < 1 || t[0] > E || t[2] < 1 || t[2] > 6)))
      return set_err(FX_EINVAL, "SideToElem(:,%d) invalid (inner sides and the small sides of mortars only)",
                     s + 1);
    const int *as = ElemToSide + 2 * (6 * (t[1] - 1) + t[3] - 1);
    if (as[0] != s + 1 || as[1] != t[4]) return set_err(FX_EINVAL, "ElemToSide inconsistent with SideToElem at side %d", s + 1);
    if (!small) {
      const int *am = ElemToSide + 2 * (6 * (t[0] - 1) + t[2] - 1);
      if (am[0] != s + 1 || am[1] != 0)
        return set_err(FX_EINVAL, "ElemToSide inconsistent with SideToElem at side %d", s + 1);
    }
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
