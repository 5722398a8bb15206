/*
 * CPU oracle: NFVSE shock capturing (native finite-volume sub-cells blended with the
 * split-form DG volume integral), first-order sub-cell fluxes (SubFVMethod=1) and the
 * Persson-Peraire indicator (ComputeAlpha=1, ModalThreshold=1).
 *
 * TEST INFRASTRUCTURE ONLY (see ora.h).  Restates src/shockcapturing/nfvse.f90,
 * nfvse_mpi.f90 (serial, conforming), nfvse_vars.f90 and src/indicators/indicators.f90.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

#define NV 5

/* threshold of ModalThreshold=1 (nfvse.f90:208-209) and the sharpness (nfvse_vars.f90:74) */
double ora_nfvse_threshold(int N) { return 0.5 * pow(10.0, -1.8 * pow(N + 1.0, 0.25)); }
static double nfvse_sharpness(void) { return log((1.0 - 1.e-4) / 1.e-4); }

/* Sub-cell normal and tangent vectors of the inner faces of one element
   (ComputeSubcellMetrics nfvse.f90:261-374).  For direction d (0 xi, 1 eta, 2 zeta) and
   inner face f = 0..N-1 between sub-cells f and f+1, at the tangential node (a, b):
     m(f) = M_d(node 0) + sum_{m=0}^{N} sum_{l=0}^{f} w_l D(l,m) M_d(node m)
   along the line, with M_d the d-th contravariant metric (Metrics_f/g/hTilde) and the
   first tangent built from the next metric (f->g, g->h, h->f).  out layout:
   [d][f][b][a][comp], comp 0-2 nv, 3-5 t1, 6-8 t2, 9 norm. */
void ora_subcell_metrics(int N, const double *Mf, const double *Mg, const double *Mh, double *out)
{
  const int n = N + 1;
  double x[16], w[16], D[256];
  ora_lgl_nodes(N, x, w);
  ora_derivative_matrix(N, x, D);  /* D[i + n*j] = D(i,j) */
  const double *Md[3] = {Mf, Mg, Mh}, *Mt[3] = {Mg, Mh, Mf};
  for (int d = 0; d < 3; d++)
    for (int f = 0; f < N; f++)
      for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++) {
          /* volume node of line position m at tangential (a,b) */
#define NODE(m) (d == 0 ? ((m) + n * ((a) + n * (b))) : (d == 1 ? ((a) + n * ((m) + n * (b))) : ((a) + n * ((b) + n * (m)))))
          double nv[3], t1[3];
          for (int c = 0; c < 3; c++) { nv[c] = Md[d][3 * NODE(0) + c]; t1[c] = Mt[d][3 * NODE(0) + c]; }
          for (int m = 0; m < n; m++)
            for (int l = 0; l <= f; l++) {
              const double wd = w[l] * D[l + n * m];
              for (int c = 0; c < 3; c++) {
                nv[c] = nv[c] + wd * Md[d][3 * NODE(m) + c];
                t1[c] = t1[c] + wd * Mt[d][3 * NODE(m) + c];
              }
            }
#undef NODE
          const double norm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          for (int c = 0; c < 3; c++) nv[c] = nv[c] / norm;
          const double tn = t1[0] * nv[0] + t1[1] * nv[1] + t1[2] * nv[2];
          for (int c = 0; c < 3; c++) t1[c] = t1[c] - tn * nv[c];
          const double t1n = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
          for (int c = 0; c < 3; c++) t1[c] = t1[c] / t1n;
          double *o = out + 10 * (a + n * (b + n * (f + N * d)));
          o[0] = nv[0]; o[1] = nv[1]; o[2] = nv[2];
          o[3] = t1[0]; o[4] = t1[1]; o[5] = t1[2];
          o[6] = nv[1] * t1[2] - nv[2] * t1[1];
          o[7] = nv[2] * t1[0] - nv[0] * t1[2];
          o[8] = nv[0] * t1[1] - nv[1] * t1[0];
          o[9] = norm;
        }
}

/* Persson-Peraire indicator on rho*p (PerssonPeraire_Compute indicators.f90:108-149,
   Get_DensityTimesPressure equation_vars.f90:752-775), one value per element. */
void ora_persson_peraire(const ora_problem *P, const double *U, double *eta)
{
  const int N = P->N, n = N + 1, n3 = n * n * n;
  double x[16], w[16], Vdm[256], sVdm[256];
  ora_lgl_nodes(N, x, w);
  ora_legendre_vdm(N, x, Vdm, sVdm);
  const double km1 = P->kappa - 1.0;
  double *ind = malloc(sizeof(double) * n3), *mod = malloc(sizeof(double) * n3);
  for (int e = 0; e < P->nElems; e++) {
    for (int p = 0; p < n3; p++) {
      const double *u = U + NV * ((size_t)e * n3 + p);
      const double pr = km1 * (u[4] - 0.5 * ((u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]));
      ind[p] = u[0] * pr;
    }
    ora_change_basis_3d(1, N, N, sVdm, ind, mod);
    double LU = 0.0, LUM1 = 0.0, LUM2 = 0.0;
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          const double m2 = mod[i + n * (j + n * k)] * mod[i + n * (j + n * k)];
          LU += m2;
          if (i < N && j < N && k < N) LUM1 += m2;
          if (i < N - 1 && j < N - 1 && k < N - 1) LUM2 += m2;
        }
    const double LU_N = LU - LUM1, LU_NM1 = LUM1 - LUM2;
    double et = fmax(LU_N / LU, LU_NM1 / LUM1);
    if (et < DBL_EPSILON) et = DBL_EPSILON;
    eta[e] = et;
  }
  free(ind); free(mod);
}

/* CalcBlendingCoefficient (nfvse.f90:1537-1615; ComputeAlpha=1 with TimeRelFactor=0, or
   ComputeAlpha=3: alpha = ShockBlendCoef) followed by
   the propagation sweeps (ProlongBlendingCoeffToFaces / PropagateBlendingCoeff
   nfvse_mpi.f90:31-153, conforming sides). */
void ora_blending_coefficient(const ora_problem *P, const double *U, double *alpha)
{
  const int E = P->nElems;
  if (P->computeAlpha == 1) ora_persson_peraire(P, U, alpha);
  const double T = ora_nfvse_threshold(P->N), s = nfvse_sharpness();
  for (int e = 0; e < E; e++) {
    double a = (P->computeAlpha == 1) ? 1.0 / (1.0 + exp(-s * (alpha[e] - T) / T)) : P->shockBlendCoef;
    if (a < P->alpha_min) a = 0.0;
    else if (a >= P->alpha_max) a = P->alpha_max;
    alpha[e] = a;
  }
  double *am = malloc(sizeof(double) * P->nSides), *as = malloc(sizeof(double) * P->nSides);
  for (int sweep = 0; sweep < P->spacePropSweeps; sweep++) {
    for (int sd = 0; sd < P->nSides; sd++) {
      am[sd] = alpha[P->SideToElem[5 * sd] - 1];
      as[sd] = alpha[P->SideToElem[5 * sd + 1] - 1];
    }
    for (int sd = 0; sd < P->nSides; sd++) {
      const double m = P->spacePropFactor * fmax(am[sd], as[sd]);
      const int em = P->SideToElem[5 * sd] - 1, es = P->SideToElem[5 * sd + 1] - 1;
      if (alpha[em] < m) alpha[em] = m;
      if (alpha[es] < m) alpha[es] = m;
    }
  }
  free(am); free(as);
}

/* VolInt_NFVSE (nfvse.f90:475-561) with Compute_FVFluxes_1st_Order (:566-670) and
   Compute_FVFluxes_1D (:1389-1443): on elements with alpha > 0, the volume part of Ut is
   blended, Ut = (1-alpha) Ut + alpha F_FV, where F_FV(i,j,k) = sum_d (f_d(i) - f_d(i-1))/w_i
   and f_d(f) = norm * AdvRiemann(U_f, U_f+1; nv, t1, t2) at the inner sub-cell faces
   (zero on the element faces, whose flux the surface integral adds). */
void ora_volint_nfvse(const ora_problem *P, const double *U, const double *alpha, double *Ut)
{
  const int N = P->N, n = N + 1, n3 = n * n * n;
  double x[16], w[16];
  ora_lgl_nodes(N, x, w);
  double *scm = malloc(sizeof(double) * 10 * 3 * N * n * n);
  double *fl = malloc(sizeof(double) * NV * 3 * (n + 1) * n * n);
  for (int e = 0; e < P->nElems; e++) {
    if (alpha[e] == 0.0) continue;
    const size_t eo = (size_t)e * n3;
    ora_subcell_metrics(N, P->Mf + 3 * eo, P->Mg + 3 * eo, P->Mh + 3 * eo, scm);
    /* fl[d][f+1][b][a][v], f = -1..N: zero at f = -1 and f = N */
    memset(fl, 0, sizeof(double) * NV * 3 * (n + 1) * n * n);
    for (int d = 0; d < 3; d++)
      for (int f = 0; f < N; f++)
        for (int b = 0; b < n; b++)
          for (int a = 0; a < n; a++) {
#define NODE(m) (d == 0 ? ((m) + n * ((a) + n * (b))) : (d == 1 ? ((a) + n * ((m) + n * (b))) : ((a) + n * ((b) + n * (m)))))
            const double *g = scm + 10 * (a + n * (b + n * (f + N * d)));
            double F[5];
            /* point-wise solvers only (no face-wide wave speed on sub-cell interfaces) */
            ora_riemann_lam_vf(P->riemann, P->volflux, P->kappa, U + NV * (eo + NODE(f)), U + NV * (eo + NODE(f + 1)),
                               g, g + 3, g + 6, 0.0, F);
#undef NODE
            double *o = fl + NV * (a + n * (b + n * ((f + 1) + (n + 1) * d)));
            for (int v = 0; v < NV; v++) o[v] = F[v] * g[9];
          }
#define FL(d, f, a, b, v) fl[NV * ((a) + n * ((b) + n * (((f) + 1) + (n + 1) * (d)))) + (v)]
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double *ut = Ut + NV * (eo + i + n * (j + n * k));
          for (int v = 0; v < NV; v++) {
            const double ffv = (1.0 / w[i]) * (FL(0, i, j, k, v) - FL(0, i - 1, j, k, v)) +
                               (1.0 / w[j]) * (FL(1, j, i, k, v) - FL(1, j - 1, i, k, v)) +
                               (1.0 / w[k]) * (FL(2, k, i, j, v) - FL(2, k - 1, i, j, v));
            ut[v] = (1.0 - alpha[e]) * ut[v] + alpha[e] * ffv;
          }
        }
#undef FL
  }
  free(scm); free(fl);
}
