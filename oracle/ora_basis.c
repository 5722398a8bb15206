/*
 * Oracle (test infrastructure only): 1D polynomial bases, Vandermonde matrices and
 * the DG operators on Legendre-Gauss-Lobatto nodes.  Restates
 * src/interpolation/basis.f90, interpolation.f90, changeBasis.f90 and
 * src/dg/dg.f90 (InitDGbasis).  Matrices are column-major: A(r,c) -> A[r + rows*c].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "ora.h"

/* LegendrePolynomialAndDerivative, normalised by sqrt(N+1/2) (basis.f90:286-323) */
void ora_legendre(int N, double x, double *L, double *Lder)
{
  double l, ld;
  if (N == 0) { l = 1.0; ld = 0.0; }
  else if (N == 1) { l = x; ld = 1.0; }
  else {
    double Lm2 = 1.0, Lm1 = x, Ldm2 = 0.0, Ldm1 = 1.0;
    l = 0; ld = 0;
    for (int i = 2; i <= N; i++) {
      l = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
      ld = Ldm2 + (double)(2 * i - 1) * Lm1;
      Lm2 = Lm1; Lm1 = l; Ldm2 = Ldm1; Ldm1 = ld;
    }
  }
  *L = l * sqrt((double)N + 0.5);
  *Lder = ld * sqrt((double)N + 0.5);
}

/* LegendreGaussNodesAndWeights (basis.f90:430-497) */
void ora_gauss_nodes(int N, double *x, double *w)
{
  if (N == 0) { x[0] = 0.0; if (w) w[0] = 2.0; return; }
  if (N == 1) {
    x[0] = -sqrt(1.0 / 3.0); x[1] = -x[0];
    if (w) { w[0] = 1.0; w[1] = 1.0; }
    return;
  }
  double cheb = 2.0 * atan(1.0) / (double)(N + 1);
  for (int i = 0; i <= (N + 1) / 2 - 1; i++) {
    double xi = -cos(cheb * (double)(2 * i + 1)), L, Ld;
    for (int it = 0; it <= 10; it++) {
      ora_legendre(N + 1, xi, &L, &Ld);
      double dx = -L / Ld;
      xi += dx;
      if (fabs(dx) < 1e-15 * fabs(xi)) break;
    }
    ora_legendre(N + 1, xi, &L, &Ld);
    x[i] = xi; x[N - i] = -xi;
    if (w) { w[i] = (2.0 * N + 3) / ((1.0 - xi * xi) * Ld * Ld); w[N - i] = w[i]; }
  }
  if (N % 2 == 0) {
    double L, Ld;
    x[N / 2] = 0.0;
    ora_legendre(N + 1, 0.0, &L, &Ld);
    if (w) w[N / 2] = (2.0 * N + 3) / (Ld * Ld);
  }
}

/* qAndLEvaluation (basis.f90:505-534) */
static void q_and_L(int N, double x, double *q, double *qder, double *Lout)
{
  double Lm2 = 1.0, Lm1 = x, Ldm2 = 0.0, Ldm1 = 1.0, L = 0.0, Ld;
  for (int i = 2; i <= N; i++) {
    L = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
    Ld = Ldm2 + (double)(2 * i - 1) * Lm1;
    Lm2 = Lm1; Lm1 = L; Ldm2 = Ldm1; Ldm1 = Ld;
  }
  *q = (double)(2 * N + 1) / (double)(N + 1) * (x * L - Lm2);
  *qder = (double)(2 * N + 1) * L;
  *Lout = L;
}

/* LegGaussLobNodesAndWeights (basis.f90:544-608) */
void ora_lgl_nodes(int N, double *x, double *w)
{
  const double pi = acos(-1.0);
  x[0] = -1.0; x[N] = 1.0;
  if (w) { w[0] = 2.0 / (double)(N * (N + 1)); w[N] = w[0]; }
  if (N > 1) {
    double c1 = pi / (double)N, c2 = 3.0 / ((double)(8 * N) * pi);
    for (int i = 1; i <= (N + 1) / 2 - 1; i++) {
      double xi = -cos(c1 * ((double)i + 0.25) - c2 / ((double)i + 0.25)), q, qd, L;
      for (int it = 0; it <= 10; it++) {
        q_and_L(N, xi, &q, &qd, &L);
        double dx = -q / qd;
        xi += dx;
        if (fabs(dx) < 1e-15 * fabs(xi)) break;
      }
      q_and_L(N, xi, &q, &qd, &L);
      x[i] = xi; x[N - i] = -xi;
      if (w) { w[i] = w[0] / (L * L); w[N - i] = w[i]; }
    }
  }
  if (N % 2 == 0) {
    double q, qd, L;
    x[N / 2] = 0.0;
    q_and_L(N, 0.0, &q, &qd, &L);
    if (w) w[N / 2] = w[0] / (L * L);
  }
}

/* GetNodesAndWeights (interpolation.f90:190-265), the node types used on the path */
void ora_nodes(int N, int type, double *x, double *w)
{
  switch (type) {
  case ORA_NODE_GAUSS: ora_gauss_nodes(N, x, w); break;
  case ORA_NODE_GL: ora_lgl_nodes(N, x, w); break;
  default:
    for (int i = 0; i <= N; i++) x[i] = 2.0 * (double)i / (double)N - 1.0;
    if (w) {
      for (int i = 0; i <= N; i++) w[i] = 2.0 / (double)N;
      w[0] *= 0.5; w[N] *= 0.5;
    }
  }
}

/* BarycentricWeights (basis.f90:615-634) */
void ora_barycentric_weights(int N, const double *x, double *wb)
{
  for (int i = 0; i <= N; i++) wb[i] = 1.0;
  for (int i = 1; i <= N; i++)
    for (int j = 0; j <= i - 1; j++) {
      wb[j] *= (x[j] - x[i]);
      wb[i] *= (x[i] - x[j]);
    }
  for (int i = 0; i <= N; i++) wb[i] = 1.0 / wb[i];
}

/* PolynomialDerivativeMatrix (basis.f90:641-663); D(i,l) at D[i + n*l] */
void ora_derivative_matrix(int N, const double *x, double *D)
{
  int n = N + 1;
  double wb[64];
  ora_barycentric_weights(N, x, wb);
  memset(D, 0, sizeof(double) * n * n);
  for (int l = 0; l <= N; l++)
    for (int i = 0; i <= N; i++)
      if (l != i) {
        D[i + n * l] = wb[l] / (wb[i] * (x[i] - x[l]));
        D[i + n * i] -= D[i + n * l];
      }
}

/* ALMOSTEQUAL (basis.f90:672-689) */
static int almost_equal(double x, double y)
{
  const double tol = DBL_EPSILON;
  if (x == 0.0 || y == 0.0) return fabs(x - y) <= 2.0 * tol;
  return (fabs(x - y) <= tol * fabs(x)) && (fabs(x - y) <= tol * fabs(y));
}

/* LagrangeInterpolationPolys (basis.f90:732-767) */
void ora_lagrange_polys(double x, int N, const double *xGP, const double *wb, double *L)
{
  int hit = 0;
  for (int i = 0; i <= N; i++) {
    L[i] = 0.0;
    if (almost_equal(x, xGP[i])) { L[i] = 1.0; hit = 1; }
  }
  if (hit) return;
  double s = 0.0;
  for (int i = 0; i <= N; i++) { L[i] = wb[i] / (x - xGP[i]); s += L[i]; }
  for (int i = 0; i <= N; i++) L[i] /= s;
}

/* InitializeVandermonde (basis.f90:259-277): Vdm(0:Nout,0:Nin) */
static void init_vandermonde(int Nin, int Nout, const double *wb, const double *xin,
                             const double *xout, double *Vdm)
{
  double L[64];
  for (int r = 0; r <= Nout; r++) {
    ora_lagrange_polys(xout[r], Nin, xin, wb, L);
    for (int c = 0; c <= Nin; c++) Vdm[r + (Nout + 1) * c] = L[c];
  }
}

/* BuildLegendreVdm (basis.f90:196-252) */
void ora_legendre_vdm(int N, const double *xi, double *Vdm_Leg, double *sVdm_Leg)
{
  int n = N + 1, Ng = 3 * N / 2 + 1, ng = Ng + 1;
  double wb[64], xg[128], wg[128], dummy;
  double *VlT = (double *)malloc(sizeof(double) * n * ng);   /* (0:N,0:Ng) */
  double *Vgp = (double *)malloc(sizeof(double) * ng * n);   /* (0:Ng,0:N) */
  ora_barycentric_weights(N, xi, wb);
  ora_gauss_nodes(Ng, xg, wg);
  for (int i = 0; i <= Ng; i++)
    for (int j = 0; j <= N; j++) ora_legendre(j, xg[i], &VlT[j + n * i], &dummy);
  for (int i = 0; i <= Ng; i++)
    for (int j = 0; j <= N; j++) VlT[j + n * i] *= wg[i];
  init_vandermonde(N, Ng, wb, xi, xg, Vgp);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) {
      double s = 0.0;
      for (int k = 0; k < ng; k++) s += VlT[r + n * k] * Vgp[k + ng * c];
      sVdm_Leg[r + n * c] = s;
    }
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N; j++) ora_legendre(j, xi[i], &Vdm_Leg[i + n * j], &dummy);
  free(VlT); free(Vgp);
}

/* GetVandermonde (interpolation.f90:271-339), In->Out only */
void ora_vandermonde(int Nin, int typeIn, int Nout, int typeOut, int modal, double *Vdm)
{
  int ni = Nin + 1, no = Nout + 1;
  if (typeIn == typeOut && Nin == Nout) {
    memset(Vdm, 0, sizeof(double) * ni * no);
    for (int i = 0; i < ni; i++) Vdm[i + no * i] = 1.0;
    return;
  }
  double xin[64], xout[64], wbin[64];
  ora_nodes(Nin, typeIn, xin, NULL);
  ora_barycentric_weights(Nin, xin, wbin);
  ora_nodes(Nout, typeOut, xout, NULL);
  if (modal && Nout < Nin) {
    double *VLi = malloc(sizeof(double) * ni * ni), *sVLi = malloc(sizeof(double) * ni * ni);
    double *VLo = malloc(sizeof(double) * no * no), *sVLo = malloc(sizeof(double) * no * no);
    ora_legendre_vdm(Nin, xin, VLi, sVLi);
    ora_legendre_vdm(Nout, xout, VLo, sVLo);
    /* Vdm = Vdm_Leg_Out(0:Nout,0:Nout) * sVdm_Leg_In(0:Nout,0:Nin) */
    for (int r = 0; r < no; r++)
      for (int c = 0; c < ni; c++) {
        double s = 0.0;
        for (int k = 0; k < no; k++) s += VLo[r + no * k] * sVLi[k + ni * c];
        Vdm[r + no * c] = s;
      }
    free(VLi); free(sVLi); free(VLo); free(sVLo);
  } else {
    init_vandermonde(Nin, Nout, wbin, xin, xout, Vdm);
  }
}

/* ChangeBasis3D_Single (changeBasis.f90:225-275): X(dim1,0:Nin^3) -> X(dim1,0:Nout^3) */
void ora_change_basis_3d(int dim1, int Nin, int Nout, const double *Vdm, const double *Xin, double *Xout)
{
  int ni = Nin + 1, no = Nout + 1;
  double *b1 = calloc((size_t)dim1 * no * ni * ni, sizeof(double));
  double *b2 = calloc((size_t)dim1 * no * no * ni, sizeof(double));
#define IDX(d, i, j, k, n1, n2) ((d) + dim1 * ((i) + (n1) * ((j) + (n2) * (k))))
  for (int k = 0; k < ni; k++)
    for (int j = 0; j < ni; j++)
      for (int i = 0; i < ni; i++)
        for (int io = 0; io < no; io++)
          for (int d = 0; d < dim1; d++)
            b1[IDX(d, io, j, k, no, ni)] += Vdm[io + no * i] * Xin[IDX(d, i, j, k, ni, ni)];
  for (int k = 0; k < ni; k++)
    for (int j = 0; j < ni; j++)
      for (int jo = 0; jo < no; jo++)
        for (int io = 0; io < no; io++)
          for (int d = 0; d < dim1; d++)
            b2[IDX(d, io, jo, k, no, no)] += Vdm[jo + no * j] * b1[IDX(d, io, j, k, no, ni)];
  memset(Xout, 0, sizeof(double) * dim1 * no * no * no);
  for (int k = 0; k < ni; k++)
    for (int ko = 0; ko < no; ko++)
      for (int jo = 0; jo < no; jo++)
        for (int io = 0; io < no; io++)
          for (int d = 0; d < dim1; d++)
            Xout[IDX(d, io, jo, ko, no, no)] += Vdm[ko + no * k] * b2[IDX(d, io, jo, k, no, no)];
#undef IDX
  free(b1); free(b2);
}

/* InitDGbasis for LGL nodes (src/dg/dg.f90:181-252):
   D_Hat = -M^-1 D^T M, L_Hat = M^-1 L, Dvolsurf = 2D - M^-1 Vf^T B Vf. */
void ora_dg_operators(int N, double *xGP, double *wGP, double *D, double *D_hat,
                      double *Dvolsurf, double *L_hatminus, double *L_hatplus)
{
  int n = N + 1;
  double wb[64], Lm[64], Lp[64];
  ora_lgl_nodes(N, xGP, wGP);
  ora_barycentric_weights(N, xGP, wb);
  ora_derivative_matrix(N, xGP, D);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      D_hat[i + n * j] = -((1.0 / wGP[i]) * (D[j + n * i] * wGP[j]));
  ora_lagrange_polys(1.0, N, xGP, wb, Lp);
  ora_lagrange_polys(-1.0, N, xGP, wb, Lm);
  for (int i = 0; i < n; i++) { L_hatplus[i] = Lp[i] / wGP[i]; L_hatminus[i] = Lm[i] / wGP[i]; }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      /* (Vf^T B Vf)(i,j) = -Lm(i)Lm(j) + Lp(i)Lp(j) */
      double vbv = -Lm[i] * Lm[j] + Lp[i] * Lp[j];
      Dvolsurf[i + n * j] = 2.0 * D[i + n * j] - (1.0 / wGP[i]) * vbv;
    }
}
