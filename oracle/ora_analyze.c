/*
 * CPU oracle: the reference's analysis of a state.
 *   CalcErrorNorms (src/analyze/analyze.f90:557-628): L2 error on NAnalyze Gauss-Lobatto
 *     points (solution, coordinates and Jacobian interpolated with Vdm_GaussN_NAnalyze,
 *     InitAnalyzeBasis :267-296), L2 error on the collocation points, Linf error on the
 *     analyze points, against AnalyzeExactFunc at time t;
 *   the domain volume Vol (InitAnalyze :213-221);
 *   CalcBulk (:633-675): integrals of U and Ut, max|Ut|;
 *   the Taylor-Green vortex testcase analysis (src/testcase/taylorgreenvortex/
 *     testcase_analyze.f90:127-401): kinetic and magnetic energy, enstrophies, entropy,
 *     dissipation rates, maximum vorticity and current, mean temperature, normalised as
 *     there.  The reference's testcase is GLM-MHD (9 variables); for 5 variables the same
 *     formulas run with B = psi = 0.
 *
 * TEST INFRASTRUCTURE ONLY (see ora.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

/* ExactFunc 1 (constant RefState) and 2 (density wave) of
   src/equation/navierstokes/equation.f90:607-651; ExactFunc 1 of src/equation/mhd/equation.f90 */
static void exact_state(const ora_analyze_cfg *c, int nVar, double kappa, double t, const double *x, double *u)
{
  if (c->exactFunc == 1) {
    for (int v = 0; v < nVar; v++) u[v] = c->refState[v];
    return;
  }
  const double pi = acos(-1.0);
  double cent[3], arg = 0.0;
  for (int d = 0; d < 3; d++) { cent[d] = x[d] - c->advVel[d] * t; arg += c->waveNumber[d] * cent[d]; }
  const double rho = 1.0 * (1.0 + c->amplitude * sin(pi * arg));
  u[0] = rho;
  for (int d = 0; d < 3; d++) u[1 + d] = rho * c->advVel[d];
  u[4] = 1.0 / (kappa - 1.0) + 0.5 * rho *
         (c->advVel[0] * c->advVel[0] + c->advVel[1] * c->advVel[1] + c->advVel[2] * c->advVel[2]);
}

void ora_analyze(int N, int nVar, int nElems, const double *Elem_xGP, const double *sJ, const double *U,
                 const double *Ut, const double *gx, const double *gy, const double *gz, double kappa,
                 double mu_0, double t, const ora_analyze_cfg *c, ora_analysis *o)
{
  const int n = N + 1, n3 = n * n * n;
  const int NA = c->nAnalyze > 0 ? c->nAnalyze : 2 * (N + 1), na = NA + 1, na3 = na * na * na;
  const double km1 = kappa - 1.0;
  const double s2mu0 = nVar == 9 ? 0.5 / mu_0 : 0.0;
  double xGP[32], wGP[32], xA[32], wA[32];
  ora_nodes(N, ORA_NODE_GL, xGP, wGP);
  ora_nodes(NA, ORA_NODE_GL, xA, wA);
  double *Vdm = malloc(sizeof(double) * na * n);
  ora_vandermonde(N, ORA_NODE_GL, NA, ORA_NODE_GL, 0, Vdm);
  double *xa = malloc(sizeof(double) * 3 * na3), *ja = malloc(sizeof(double) * na3);
  double *ua = malloc(sizeof(double) * nVar * na3), *jn = malloc(sizeof(double) * n3);
  memset(o, 0, sizeof(*o));
  for (int v = 0; v < nVar; v++) { o->Linf[v] = c->exactFunc > 0 ? -1e10 : 0.0; o->maxabsUt[v] = -1e10; }
  o->max_Vorticity = -1.0;
  o->max_Current = -1.0;
  double uex[9];
  for (int e = 0; e < nElems; e++) {
    const double *ue = U + (size_t)nVar * n3 * e, *xe = Elem_xGP + (size_t)3 * n3 * e, *se = sJ + (size_t)n3 * e;
    /* CalcErrorNorms: analyze points */
    if (c->exactFunc > 0) {
      for (int p = 0; p < n3; p++) jn[p] = 1.0 / se[p];
      ora_change_basis_3d(3, N, NA, Vdm, xe, xa);
      ora_change_basis_3d(1, N, NA, Vdm, jn, ja);
      ora_change_basis_3d(nVar, N, NA, Vdm, ue, ua);
      for (int m = 0; m < na; m++)
        for (int l = 0; l < na; l++)
          for (int k = 0; k < na; k++) {
            const int p = k + na * (l + na * m);
            exact_state(c, nVar, kappa, t, xa + 3 * p, uex);
            const double w = wA[k] * wA[l] * wA[m] * ja[p];
            for (int v = 0; v < nVar; v++) {
              const double d = ua[nVar * p + v] - uex[v];
              if (fabs(d) > o->Linf[v]) o->Linf[v] = fabs(d);
              o->L2[v] += d * d * w;
            }
          }
    }
    for (int kk = 0; kk < n; kk++)
      for (int jj = 0; jj < n; jj++)
        for (int ii = 0; ii < n; ii++) {
          const int p = ii + n * (jj + n * kk);
          const size_t g = (size_t)n3 * e + p;
          const double *u = U + nVar * g, *ut = Ut + nVar * g;
          const double w = wGP[ii] * wGP[jj] * wGP[kk] / sJ[g];
          o->vol += w;
          if (c->exactFunc > 0) {
            exact_state(c, nVar, kappa, t, Elem_xGP + 3 * g, uex);
            for (int v = 0; v < nVar; v++) o->L2colloc[v] += (u[v] - uex[v]) * (u[v] - uex[v]) * w;
          }
          for (int v = 0; v < nVar; v++) {
            o->bulk[v] += u[v] * w;
            o->bulk_t[v] += ut[v] * w;
            if (fabs(ut[v]) > o->maxabsUt[v]) o->maxabsUt[v] = fabs(ut[v]);
          }
          /* AnalyzeTestCase integrands */
          const double rho = u[0], srho = 1.0 / rho;
          double vel[3], B[3] = {0.0, 0.0, 0.0}, G[3][3], GB[3][3];
          for (int d = 0; d < 3; d++) vel[d] = u[1 + d] * srho;
          if (nVar == 9)
            for (int d = 0; d < 3; d++) B[d] = u[5 + d];
          const double psi = nVar == 9 ? u[8] : 0.0;
          for (int q = 0; q < 3; q++) {
            const double *gq = q == 0 ? gx : (q == 1 ? gy : gz);
            for (int r = 0; r < 3; r++) {
              G[r][q] = gq ? gq[nVar * g + 1 + r] : 0.0;
              GB[r][q] = (gq && nVar == 9) ? gq[nVar * g + 5 + r] : 0.0;
            }
          }
          const double BB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
          const double pres = km1 * (u[4] - 0.5 * (u[1] * vel[0] + u[2] * vel[1] + u[3] * vel[2]) -
                                     s2mu0 * (BB + psi * psi));
          const double divU = G[0][0] + G[1][1] + G[2][2];
          double S[3][3], Sd[3][3];
          for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) { S[r][q] = 0.5 * (G[r][q] + G[q][r]); Sd[r][q] = S[r][q]; }
          for (int r = 0; r < 3; r++) Sd[r][r] -= 1.0 / 3.0 * divU;
          const double kE = 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
          const double mag = s2mu0 * BB;
          const double om[3] = {G[2][1] - G[1][2], G[0][2] - G[2][0], G[1][0] - G[0][1]};
          const double omn = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
          if (omn > o->max_Vorticity) o->max_Vorticity = omn;
          const double cj[3] = {GB[2][1] - GB[1][2], GB[0][2] - GB[2][0], GB[1][0] - GB[0][1]};
          const double cjn = sqrt(cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2]);
          if (cjn > o->max_Current) o->max_Current = cjn;
          double ut2 = 0.0, st2 = 0.0, sd2 = 0.0;
          for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
              ut2 += G[r][q] * G[r][q];
              st2 += S[r][q] * S[r][q];
              sd2 += Sd[r][q] * Sd[r][q];
            }
          o->Ekin += kE * w;
          o->Ekin_comp += rho * kE * w;
          o->Emag_comp += mag * w;
          o->Enstrophy_comp += 0.5 * rho * (om[0] * om[0] + om[1] * om[1] + om[2] * om[2]) * w;
          o->Mag_Enstrophy += s2mu0 * (cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2]) * w;
          o->Entropy_comp += -rho * (log(pres) - kappa * log(rho)) / km1 * w;
          o->DR_u += ut2 * w;
          o->DR_S += st2 * w;
          o->DR_Sd += sd2 * w;
          o->DR_p += pres * divU * w;
          o->mean_temperature += km1 / c->R * (u[4] * srho - kE - mag * srho - s2mu0 * psi * psi) * w;
        }
  }
  const double Vol = o->vol, rho0 = 1.0;
  for (int v = 0; v < nVar; v++) {
    o->L2[v] = sqrt(o->L2[v] / Vol);
    o->L2colloc[v] = sqrt(o->L2colloc[v] / Vol);
  }
  o->uprime = sqrt(o->Ekin / Vol * 2.0 / 3.0);
  o->Ekin = o->Ekin / Vol;
  o->Ekin_comp = o->Ekin_comp / (rho0 * Vol);
  o->Emag_comp = o->Emag_comp / (rho0 * Vol);
  o->Enstrophy_comp = o->Enstrophy_comp / (rho0 * Vol);
  o->Mag_Enstrophy = o->Mag_Enstrophy / (rho0 * Vol);
  o->Entropy_comp = o->Entropy_comp / (rho0 * Vol);
  o->DR_u = o->DR_u * c->dissFactor / Vol;
  o->DR_S = o->DR_S * 2.0 * c->dissFactor / (rho0 * Vol);
  o->DR_Sd = o->DR_Sd * 2.0 * c->dissFactor / (rho0 * Vol);
  o->DR_p = -o->DR_p / (rho0 * Vol);
  o->mean_temperature = o->mean_temperature / Vol;
  free(Vdm); free(xa); free(ja); free(ua); free(jn);
}
