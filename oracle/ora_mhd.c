/*
 * CPU oracle: GLM-MHD split-form DGSEM on LGL nodes (9 variables: rho, rho v, E, B, psi),
 * with the Derigs et al. entropy- and kinetic-energy-conserving two-point flux
 * (VolumeFlux=10), Powell non-conservative terms (NONCONS=1), the central, Rusanov,
 * Derigs entropy-stable LLF or Derigs EC surface flux (Riemann=0, 1, 10, 11), CALCTIMESTEP with the GLM cleaning speed from the time step, and the
 * Carpenter RK4-5 step.  Resistive case (PARABOLIC): BR1 lifting of the entropy variables,
 * isotropic viscosity, resistivity and heat conduction (EvalDiffFlux3D), the weak-form
 * VolInt_visc and the BR1 average of the diffusive fluxes at faces.
 *
 * TEST INFRASTRUCTURE ONLY (see ora.h).  Restates src/equation/mhd/{flux_average,flux,
 * riemann,equation_vars,calctimestep}.f90 and the NONCONS branches of src/dg/{fillflux,
 * surfint,volint}.f90 for periodic conforming meshes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

#define MV 9

typedef struct { double sr, v[3], pt, v2, B2, vB; } aux_t;

/* EvalUaux (mhd/flux_average.f90:223-259) */
static void mhd_uaux(const ora_mhd *M, const double *U, aux_t *a)
{
  const double km1 = M->kappa - 1.0, km2 = M->kappa - 2.0, s2mu0 = 0.5 * (1.0 / M->mu_0);
  a->sr = 1.0 / U[0];
  for (int d = 0; d < 3; d++) a->v[d] = U[1 + d] * a->sr;
  a->v2 = a->v[0] * a->v[0] + a->v[1] * a->v[1] + a->v[2] * a->v[2];
  a->B2 = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
  a->vB = a->v[0] * U[5] + a->v[1] * U[6] + a->v[2] * U[7];
  a->pt = km1 * (U[4] - 0.5 * U[0] * a->v2 - s2mu0 * U[8] * U[8]) - km2 * s2mu0 * a->B2;
}

/* LN_MEAN of mhd/flux_average.f90:1091-1111 is the Navier-Stokes one (ora_ln_mean) */
#define ln_mean ora_ln_mean

/* EntropyAndKinEnergyConservingFluxVec_Derigs (mhd/flux_average.f90:761-848) */
void ora_mhd_derigs(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, const double *mL,
                    const double *mR, double *F)
{
  aux_t aL, aR;
  mhd_uaux(M, UL, &aL);
  mhd_uaux(M, UR, &aR);
  const double smu0 = 1.0 / M->mu_0, s2mu0 = 0.5 * smu0, skm1 = 1.0 / (M->kappa - 1.0);
  double m[3];
  for (int d = 0; d < 3; d++) m[d] = 0.5 * (mL[d] + mR[d]);
  const double betaL = 0.5 * UL[0] / (aL.pt - s2mu0 * aL.B2), betaR = 0.5 * UR[0] / (aR.pt - s2mu0 * aR.B2);
  const double rhoLN = ln_mean(UL[0], UR[0]), betaLN = ln_mean(betaL, betaR);
  double vA[3], BA[3];
  for (int d = 0; d < 3; d++) { vA[d] = 0.5 * (aL.v[d] + aR.v[d]); BA[d] = 0.5 * (UL[5 + d] + UR[5 + d]); }
  const double pTilde = 0.5 * ((UL[0] + UR[0]) / (betaL + betaR) + smu0 * 0.5 * (aL.B2 + aR.B2));
  const double vm = vA[0] * m[0] + vA[1] * m[1] + vA[2] * m[2];
  const double Bm = BA[0] * m[0] + BA[1] * m[1] + BA[2] * m[2];
  const double psiA = 0.5 * (UL[8] + UR[8]);
  F[0] = rhoLN * vm;
  for (int d = 0; d < 3; d++) F[1 + d] = F[0] * vA[d] - (smu0 * Bm) * BA[d] + pTilde * m[d];
  for (int d = 0; d < 3; d++) F[5 + d] = vm * BA[d] - Bm * vA[d] + (glm_ch * psiA) * m[d];
  F[8] = glm_ch * Bm;
  double vF = vA[0] * F[1] + vA[1] * F[2] + vA[2] * F[3];
  double BF = BA[0] * F[5] + BA[1] * F[6] + BA[2] * F[7];
  double B2vm = 0.0, psiBm = 0.0;
  for (int d = 0; d < 3; d++) {
    B2vm += (aL.B2 * aL.v[d] + aR.B2 * aR.v[d]) * m[d];
    psiBm += (UL[8] * UL[5 + d] + UR[8] * UR[5 + d]) * m[d];
  }
  F[4] = F[0] * 0.5 * (skm1 / betaLN - 0.5 * (aL.v2 + aR.v2)) + vF +
         smu0 * (BF - 0.25 * B2vm + 0.5 * (aL.vB + aR.vB) * Bm + F[8] * psiA - (glm_ch * 0.5) * psiBm);
}

/* EvalAdvectionFlux1D (mhd/flux.f90:391-442), x-direction of a rotated state */
static void adv_flux_1d(const ora_mhd *M, double glm_ch, const double *U, double *F)
{
  const double km1 = M->kappa - 1.0, km2 = M->kappa - 2.0, smu0 = 1.0 / M->mu_0, s2mu0 = 0.5 * smu0;
  const double sr = 1.0 / U[0], v1 = U[1] * sr, v2 = U[2] * sr, v3 = U[3] * sr;
  const double b1 = U[5], b2 = U[6], b3 = U[7];
  const double Et = U[4] - s2mu0 * U[8] * U[8];
  const double vb = b1 * v1 + b2 * v2 + b3 * v3;
  const double pt = km1 * (Et - 0.5 * U[0] * (v1 * v1 + v2 * v2 + v3 * v3)) - km2 * s2mu0 * (b1 * b1 + b2 * b2 + b3 * b3);
  F[0] = U[1];
  F[1] = U[1] * v1 + pt - smu0 * b1 * b1;
  F[2] = U[1] * v2 - smu0 * b1 * b2;
  F[3] = U[1] * v3 - smu0 * b1 * b3;
  F[4] = (Et + pt) * v1 - smu0 * b1 * vb;
  F[5] = 0.0;
  F[6] = v1 * b2 - b1 * v2;
  F[7] = v1 * b3 - b1 * v3;
  F[4] = F[4] + smu0 * glm_ch * b1 * U[8];
  F[5] = F[5] + glm_ch * U[8];
  F[8] = glm_ch * b1;
}

/* ConsToPrim (mhd/equation_vars.f90:229-251) */
static void cons_to_prim(const ora_mhd *M, const double *c, double *p)
{
  const double sr = 1.0 / c[0];
  p[0] = c[0];
  for (int d = 1; d < 4; d++) p[d] = c[d] * sr;
  double b2 = 0.0;
  for (int v = 5; v < MV; v++) b2 += c[v] * c[v];
  p[4] = (M->kappa - 1.0) * (c[4] - 0.5 * (c[1] * p[1] + c[2] * p[2] + c[3] * p[3]) - 0.5 * (1.0 / M->mu_0) * b2);
  for (int v = 5; v < MV; v++) p[v] = c[v];
}

/* FastestWave1D (mhd/equation_vars.f90:663-684), rotated primitive state */
static double fastest_wave_1d(const ora_mhd *M, const double *p)
{
  const double sr = 1.0 / p[0], smu0 = 1.0 / M->mu_0;
  const double c2 = M->kappa * p[4] * sr;
  const double ca2 = p[5] * p[5] * sr * smu0;
  const double va2 = (p[6] * p[6] + p[7] * p[7]) * sr * smu0 + ca2;
  const double astar = sqrt((c2 + va2) * (c2 + va2) - 4.0 * c2 * ca2);
  return sqrt(0.5 * (c2 + va2 + astar));
}

/* EntropyAndKinEnergyConservingFlux_Derigs (mhd/flux_average.f90:660-751): x-direction of a
   rotated state */
static void derigs_1d(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, double *F)
{
  const double smu0 = 1.0 / M->mu_0, km1 = M->kappa - 1.0, skm1 = 1.0 / km1;
  const double EL = UL[4] - 0.5 * smu0 * UL[8] * UL[8], ER = UR[4] - 0.5 * smu0 * UR[8] * UR[8];
  double vL[3], vR[3], vA[3], BA[3];
  for (int d = 0; d < 3; d++) { vL[d] = UL[1 + d] / UL[0]; vR[d] = UR[1 + d] / UR[0]; }
  const double v2L = vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2];
  const double v2R = vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2];
  const double B2L = UL[5] * UL[5] + UL[6] * UL[6] + UL[7] * UL[7];
  const double B2R = UR[5] * UR[5] + UR[6] * UR[6] + UR[7] * UR[7];
  const double pL = km1 * (EL - 0.5 * (UL[0] * v2L + smu0 * B2L));
  const double pR = km1 * (ER - 0.5 * (UR[0] * v2R + smu0 * B2R));
  const double betaL = 0.5 * UL[0] / pL, betaR = 0.5 * UR[0] / pR;
  const double rhoLN = ln_mean(UL[0], UR[0]), betaLN = ln_mean(betaL, betaR);
  for (int d = 0; d < 3; d++) { vA[d] = 0.5 * (vL[d] + vR[d]); BA[d] = 0.5 * (UL[5 + d] + UR[5 + d]); }
  const double v1B2 = 0.5 * (vL[0] * B2L + vR[0] * B2R);
  const double vB = 0.5 * ((vL[0] * UL[5] + vL[1] * UL[6] + vL[2] * UL[7]) +
                           (vR[0] * UR[5] + vR[1] * UR[6] + vR[2] * UR[7]));
  const double pTilde = 0.5 * ((UL[0] + UR[0]) / (betaL + betaR) + smu0 * 0.5 * (B2L + B2R));
  const double psiA = 0.5 * (UL[8] + UR[8]);
  F[0] = rhoLN * vA[0];
  F[1] = F[0] * vA[0] - smu0 * BA[0] * BA[0] + pTilde;
  F[2] = F[0] * vA[1] - smu0 * BA[0] * BA[1];
  F[3] = F[0] * vA[2] - smu0 * BA[0] * BA[2];
  F[6] = vA[0] * BA[1] - BA[0] * vA[1];
  F[7] = vA[0] * BA[2] - BA[0] * vA[2];
  F[5] = glm_ch * psiA;
  F[8] = glm_ch * BA[0];
  F[4] = F[0] * 0.5 * (skm1 / betaLN - 0.5 * (v2L + v2R)) + (vA[0] * F[1] + vA[1] * F[2] + vA[2] * F[3]) +
         smu0 * ((BA[0] * F[5] + BA[1] * F[6] + BA[2] * F[7]) - 0.5 * v1B2 + BA[0] * vB + F[8] * psiA -
                 glm_ch * 0.5 * (UL[8] * UL[5] + UR[8] * UR[5]));
}

/* EntropyStableByLLF (mhd/riemann.f90:1027-1211), rotated states: the Derigs flux minus
   LambdaMax_s2 H (v_R - v_L), H the symmetric matrix of the reference (Euler block, tau on
   B and psi), written out as the full 9x9 product (the reference skips its zero blocks) */
static void es_llf_1d(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, double *F)
{
  const double smu0 = 1.0 / M->mu_0, km1 = M->kappa - 1.0, skm1 = 1.0 / km1, kappa = M->kappa;
  derigs_1d(M, glm_ch, UL, UR, F);
  const double EL = UL[4] - 0.5 * smu0 * UL[8] * UL[8], ER = UR[4] - 0.5 * smu0 * UR[8] * UR[8];
  const double srL = 1.0 / UL[0], srR = 1.0 / UR[0];
  double uL[3], uR[3], uA[3], BA[3];
  for (int d = 0; d < 3; d++) { uL[d] = UL[1 + d] * srL; uR[d] = UR[1 + d] * srR; }
  const double u2L = uL[0] * uL[0] + uL[1] * uL[1] + uL[2] * uL[2];
  const double u2R = uR[0] * uR[0] + uR[1] * uR[1] + uR[2] * uR[2];
  const double B2L = UL[5] * UL[5] + UL[6] * UL[6] + UL[7] * UL[7];
  const double B2R = UR[5] * UR[5] + UR[6] * UR[6] + UR[7] * UR[7];
  const double pL = km1 * (EL - 0.5 * (UL[0] * u2L + smu0 * B2L));
  const double pR = km1 * (ER - 0.5 * (UR[0] * u2R + smu0 * B2R));
  const double bL = 0.5 * UL[0] / pL, bR = 0.5 * UR[0] / pR;
  const double rhoLN = ln_mean(UL[0], UR[0]), sbetaLN = 1.0 / ln_mean(bL, bR);
  for (int d = 0; d < 3; d++) { uA[d] = 0.5 * (uL[d] + uR[d]); BA[d] = 0.5 * (UL[5 + d] + UR[5 + d]); }
  const double u2Avg = 0.5 * (u2L + u2R), B2Avg = 0.5 * (B2L + B2R);
  const double pAvg = 0.5 * (UL[0] + UR[0]) / (bL + bR);
  const double psiA = 0.5 * (UL[8] + UR[8]);
  const double pLN = 0.5 * rhoLN * sbetaLN, srhoLN = 1.0 / rhoLN;
  const double a2 = kappa * pAvg * srhoLN, va2 = B2Avg * srhoLN, ca2 = BA[0] * BA[0] * srhoLN;
  const double cf = sqrt(0.5 * ((a2 + va2) + sqrt((a2 + va2) * (a2 + va2) - 4.0 * a2 * ca2)));
  const double lam = 0.5 * (fabs(uA[0]) + cf);
  const double uA2 = uA[0] * uA[0] + uA[1] * uA[1] + uA[2] * uA[2];
  const double Eavg = pLN * skm1 + 0.5 * rhoLN * (2.0 * uA2 - u2Avg);
  const double tau = 1.0 / (bL + bR);
  double H[9][9] = {{0}};
  H[0][0] = rhoLN;
  for (int d = 0; d < 3; d++) { H[1 + d][0] = H[0][1 + d] = rhoLN * uA[d]; }
  H[4][0] = H[0][4] = Eavg;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) H[1 + r][1 + c] = rhoLN * uA[r] * uA[c] + (r == c ? pAvg : 0.0);
  for (int d = 0; d < 3; d++) H[1 + d][4] = H[4][1 + d] = (Eavg + pAvg) * uA[d];
  H[4][4] = (pLN * pLN * skm1 + Eavg * Eavg) * srhoLN + uA2 * pAvg +
            tau * (BA[0] * BA[0] + BA[1] * BA[1] + BA[2] * BA[2] + psiA * psiA);
  for (int d = 0; d < 3; d++) { H[5 + d][4] = H[4][5 + d] = tau * BA[d]; H[5 + d][5 + d] = tau; }
  H[8][4] = H[4][8] = tau * psiA;
  H[8][8] = tau;
  double V[9];
  V[0] = (kappa * log(UR[0] * srL) - log(pR / pL)) * skm1 - (bR * u2R - bL * u2L);
  for (int d = 0; d < 3; d++) V[1 + d] = 2.0 * (bR * uR[d] - bL * uL[d]);
  V[4] = -2.0 * (bR - bL);
  for (int v = 5; v < MV; v++) V[v] = 2.0 * (bR * UR[v] - bL * UL[v]);
  for (int r = 0; r < MV; r++) {
    double s = 0.0;
    for (int c = 0; c < MV; c++) s += H[r][c] * V[c];
    F[r] -= lam * s;
  }
}

/* AdvRiemann (mhd/riemann.f90:174-246) with the surface flux M->riemann: StandardDGFlux
   (mhd/flux_average.f90:421-484; 0), RiemannSolverByRusanov (:273-308; 1), EntropyStableByLLF
   (10), EntropyAndKinEnergyConservingFlux_Derigs (11) -- mhd/equation.f90:327-386 */
void ora_mhd_riemann(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, const double *nv,
                     const double *t1, const double *t2, double *F)
{
  double cL[MV], cR[MV], FL[MV], FR[MV], pL[MV], pR[MV], f[MV];
  memcpy(cL, UL, sizeof(cL));
  memcpy(cR, UR, sizeof(cR));
  for (int b = 1; b <= 5; b += 4) {  /* momentum (1:3) and B (5:7) */
    cL[b] = UL[b] * nv[0] + UL[b + 1] * nv[1] + UL[b + 2] * nv[2];
    cL[b + 1] = UL[b] * t1[0] + UL[b + 1] * t1[1] + UL[b + 2] * t1[2];
    cL[b + 2] = UL[b] * t2[0] + UL[b + 1] * t2[1] + UL[b + 2] * t2[2];
    cR[b] = UR[b] * nv[0] + UR[b + 1] * nv[1] + UR[b + 2] * nv[2];
    cR[b + 1] = UR[b] * t1[0] + UR[b + 1] * t1[1] + UR[b + 2] * t1[2];
    cR[b + 2] = UR[b] * t2[0] + UR[b + 1] * t2[1] + UR[b + 2] * t2[2];
  }
  if (M->riemann == 10) {
    es_llf_1d(M, glm_ch, cL, cR, f);
  } else if (M->riemann == 11) {
    derigs_1d(M, glm_ch, cL, cR, f);
  } else {
    adv_flux_1d(M, glm_ch, cL, FL);
    adv_flux_1d(M, glm_ch, cR, FR);
    cons_to_prim(M, cL, pL);
    cons_to_prim(M, cR, pR);
    const double cfL = fastest_wave_1d(M, pL), cfR = fastest_wave_1d(M, pR);
    const double lam = M->riemann == 1 ? fmax(fabs(pL[1]) + cfL, fabs(pR[1]) + cfR) : 0.0;
    for (int v = 0; v < MV; v++) f[v] = 0.5 * ((FL[v] + FR[v]) - lam * (cR[v] - cL[v]));
  }
  for (int v = 0; v < MV; v++) F[v] = f[v];
  for (int b = 1; b <= 5; b += 4)
    for (int d = 0; d < 3; d++) F[b + d] = nv[d] * f[b] + t1[d] * f[b + 1] + t2[d] * f[b + 2];
}

/* AddNonConsFlux, Powell (mhd/riemann.f90:230-267): FL(2:8) += 1/2 (B_L+B_R).n (B_L, v_L.B_L, v_L) */
/* ConsToEntropy (mhd/equation_vars.f90:330-361) */
static void mhd_cons_to_entropy(const ora_mhd *M, const double *c, double *w)
{
  const double km1 = M->kappa - 1.0, s2mu0 = 0.5 * (1.0 / M->mu_0);
  const double srho = 1.0 / c[0];
  const double u = c[1] * srho, v = c[2] * srho, ww = c[3] * srho;
  const double v2s2 = 0.5 * (u * u + v * v + ww * ww);
  double b2 = 0.0;
  for (int i = 5; i < MV; i++) b2 += c[i] * c[i];
  const double rho_sp = c[0] / (km1 * (c[4] - c[0] * v2s2 - s2mu0 * b2));
  const double s = -log(rho_sp * pow(c[0], km1));
  w[0] = (M->kappa - s) * (1.0 / km1) - rho_sp * v2s2;
  w[1] = rho_sp * u;
  w[2] = rho_sp * v;
  w[3] = rho_sp * ww;
  w[4] = -rho_sp;
  for (int i = 5; i < MV; i++) w[i] = rho_sp * c[i];
}

/* ConvertToGradPrimVec, entropy-variable lifting (mhd/equation_vars.f90:547-628): gradient of
   the entropy variables in, gradient of (rho, v, p, B, psi) out */
static void mhd_grad_prim(const ora_mhd *M, const double *c, double *g)
{
  const double km1 = M->kappa - 1.0, s2mu0 = 0.5 * (1.0 / M->mu_0);
  const double srho = 1.0 / c[0];
  const double u = c[1] * srho, v = c[2] * srho, w = c[3] * srho;
  const double ekin = 0.5 * (c[1] * u + c[2] * v + c[3] * w);
  double b2 = 0.0;
  for (int i = 5; i < MV; i++) b2 += c[i] * c[i];
  const double p = km1 * (c[4] - ekin - s2mu0 * b2);
  const double rho_sp = c[0] / p, p_srho = p * srho;
  const double g1 = g[0], g5 = g[4];
  const double gu = p_srho * (g[1] + u * g5), gv = p_srho * (g[2] + v * g5), gw = p_srho * (g[3] + w * g5);
  g[0] = c[0] * g1 - g5 * (ekin - p * (1.0 / km1)) + rho_sp * (c[1] * gu + c[2] * gv + c[3] * gw);
  g[1] = gu;
  g[2] = gv;
  g[3] = gw;
  g[4] = p_srho * (g[0] + p * g5);
  for (int i = 5; i < MV; i++) g[i] = p_srho * (g[i] + c[i] * g5);
}

/* EvalDiffFlux3D, isotropic heat flux (mhd/flux.f90:449-573), one point: Cartesian diffusive
   fluxes f, g, h from the state and its primitive gradients */
static void mhd_diff_flux(const ora_mhd *M, const double *c, const double *gx, const double *gy,
                          const double *gz, double *f, double *g, double *h)
{
  const double km1 = M->kappa - 1.0, skm1 = 1.0 / km1, smu0 = 1.0 / M->mu_0, s2mu0 = 0.5 * smu0;
  const double s23 = 2.0 / 3.0, mu = M->mu, etasmu0 = M->eta * smu0;
  const double srho = 1.0 / c[0];
  const double v1 = c[1] * srho, v2 = c[2] * srho, v3 = c[3] * srho;
  double b2 = 0.0;
  for (int i = 5; i < MV; i++) b2 += c[i] * c[i];
  const double p = km1 * (c[4] - 0.5 * (c[0] * (v1 * v1 + v2 * v2 + v3 * v3)) - s2mu0 * b2);
  const double b1 = c[5], bb2 = c[6], b3 = c[7];
  const double divv = gx[1] + gy[2] + gz[3];
  const double tx = skm1 * srho * (gx[4] - srho * p * gx[0]);
  const double ty = skm1 * srho * (gy[4] - srho * p * gy[0]);
  const double tz = skm1 * srho * (gz[4] - srho * p * gz[0]);
  const double lambda = mu * (M->kappa / M->Pr);
  const double qx = lambda * tx, qy = lambda * ty, qz = lambda * tz;
  f[0] = 0.0;
  f[1] = -mu * (2 * gx[1] - s23 * divv);
  f[2] = -mu * (gx[2] + gy[1]);
  f[3] = -mu * (gx[3] + gz[1]);
  f[5] = 0.0;
  f[6] = -etasmu0 * (gx[6] - gy[5]);
  f[7] = -etasmu0 * (gx[7] - gz[5]);
  f[4] = v1 * f[1] + v2 * f[2] + v3 * f[3] + smu0 * (b1 * f[5] + bb2 * f[6] + b3 * f[7]) - qx;
  g[0] = 0.0;
  g[1] = f[2];
  g[2] = -mu * (2 * gy[2] - s23 * divv);
  g[3] = -mu * (gy[3] + gz[2]);
  g[5] = -f[6];
  g[6] = 0.0;
  g[7] = -etasmu0 * (gy[7] - gz[6]);
  g[4] = v1 * g[1] + v2 * g[2] + v3 * g[3] + smu0 * (b1 * g[5] + bb2 * g[6] + b3 * g[7]) - qy;
  h[0] = 0.0;
  h[1] = f[3];
  h[2] = g[3];
  h[3] = -mu * (2 * gz[3] - s23 * divv);
  h[5] = -f[7];
  h[6] = -g[7];
  h[7] = 0.0;
  h[4] = v1 * h[1] + v2 * h[2] + v3 * h[3] + smu0 * (b1 * h[5] + bb2 * h[6] + b3 * h[7]) - qz;
  f[8] = g[8] = h[8] = 0.0;
}

/* BR1 lifting of the entropy variables (br1.f90:154-364 with mhd/flux.f90:718-793) followed
   by ConvertToGradPrimVec */
/* BR2 (br2.f90:167-336): volume gradients and, when fg is given, each side's master (fg[0],
   fg[2], fg[4]) and slave (fg[1], fg[3], fg[5]) face gradients (MV,0:N,0:N,nSides), all
   converted to primitive gradients with the node / face state */
static void mhd_br2_gradients(const ora_mhd *M, const double *U, double *gx, double *gy, double *gz,
                              double *const fg[6])
{
  const ora_problem *P = M->P;
  const int N = P->N, n = N + 1, n3 = n * n * n;
  const size_t n3E = (size_t)n3 * P->nElems, nface = (size_t)MV * n * n * P->nSides;
  double *w = calloc((size_t)MV * n3E, sizeof(double));
  double *f[6];
  for (int d = 0; d < 6; d++) f[d] = fg ? fg[d] : malloc(sizeof(double) * nface);
  for (size_t p = 0; p < n3E; p++) mhd_cons_to_entropy(M, U + MV * p, w + MV * p);
  ora_lift_br2(P, MV, w, M->etaBR2, gx, gy, gz, f[0], f[2], f[4], f[1], f[3], f[5]);
  for (size_t p = 0; p < n3E; p++) {
    mhd_grad_prim(M, U + MV * p, gx + MV * p);
    mhd_grad_prim(M, U + MV * p, gy + MV * p);
    mhd_grad_prim(M, U + MV * p, gz + MV * p);
  }
  for (int s = 0; s < P->nSides; s++) {  /* face states of LGL: the sides' own volume nodes */
    const int *t = P->SideToElem + 5 * s;
    for (int q = 0; q < n; q++)
      for (int pp = 0; pp < n; pp++) {
        int im[3], is[3];
        ora_side_to_vol(N, 0, pp, q, 0, t[2], im);
        ora_side_to_vol(N, 0, pp, q, t[4], t[3], is);
        const double *um = U + MV * ((size_t)(t[0] - 1) * n3 + im[0] + n * (im[1] + n * im[2]));
        const double *us = U + MV * ((size_t)(t[1] - 1) * n3 + is[0] + n * (is[1] + n * is[2]));
        const size_t fi = MV * ((size_t)s * n * n + pp + n * q);
        for (int d = 0; d < 3; d++) {
          mhd_grad_prim(M, um, f[2 * d] + fi);
          mhd_grad_prim(M, us, f[2 * d + 1] + fi);
        }
      }
  }
  if (!fg)
    for (int d = 0; d < 6; d++) free(f[d]);
  free(w);
}

void ora_mhd_lifting(const ora_mhd *M, const double *U, double *gx, double *gy, double *gz)
{
  const ora_problem *P = M->P;
  const size_t n3E = (size_t)(P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  if (M->lifting == 2) {  /* BR2: the volume gradients */
    mhd_br2_gradients(M, U, gx, gy, gz, NULL);
    return;
  }
  double *w = calloc((size_t)MV * n3E, sizeof(double));
  for (size_t p = 0; p < n3E; p++) mhd_cons_to_entropy(M, U + MV * p, w + MV * p);
  ora_lift(P, MV, w, gx, gy, gz);
  for (size_t p = 0; p < n3E; p++) {
    mhd_grad_prim(M, U + MV * p, gx + MV * p);
    mhd_grad_prim(M, U + MV * p, gy + MV * p);
    mhd_grad_prim(M, U + MV * p, gz + MV * p);
  }
  free(w);
}

/* VolInt_visc (volint.f90:228-271) with EvalDiffFluxTilde3D (mhd/flux.f90:579-712) */
static void mhd_volint_visc(const ora_mhd *M, const double *U, const double *gx, const double *gy,
                            const double *gz, const double *Dh, double *Ut)
{
  const ora_problem *P = M->P;
  const int n = P->N + 1, n3 = n * n * n;
  double *ft = malloc(sizeof(double) * MV * n3 * 3);
  double *gt = ft + MV * n3, *ht = gt + MV * n3;
  for (int e = 0; e < P->nElems; e++) {
    for (int p = 0; p < n3; p++) {
      const size_t c = (size_t)n3 * e + p;
      double f[MV], g[MV], h[MV];
      mhd_diff_flux(M, U + MV * c, gx + MV * c, gy + MV * c, gz + MV * c, f, g, h);
      const double *mf = P->Mf + 3 * c, *mg = P->Mg + 3 * c, *mh = P->Mh + 3 * c;
      for (int v = 0; v < MV; v++) {
        ft[MV * p + v] = f[v] * mf[0] + g[v] * mf[1] + h[v] * mf[2];
        gt[MV * p + v] = f[v] * mg[0] + g[v] * mg[1] + h[v] * mg[2];
        ht[MV * p + v] = f[v] * mh[0] + g[v] * mh[1] + h[v] * mh[2];
      }
    }
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double *ut = Ut + MV * ((size_t)n3 * e + i + n * (j + n * k));
          for (int l = 0; l < n; l++)
            for (int v = 0; v < MV; v++)
              ut[v] = ut[v] + Dh[i + n * l] * ft[MV * (l + n * (j + n * k)) + v]
                            + Dh[j + n * l] * gt[MV * (i + n * (l + n * k)) + v]
                            + Dh[k + n * l] * ht[MV * (i + n * (j + n * l)) + v];
        }
  }
  free(ft);
}

static void add_noncons_surface(const double *UL, const double *UR, const double *nv, double *F)
{
  double vL[3];
  for (int d = 0; d < 3; d++) vL[d] = UL[1 + d] / UL[0];
  const double bn = 0.5 * ((UL[5] + UR[5]) * nv[0] + (UL[6] + UR[6]) * nv[1] + (UL[7] + UR[7]) * nv[2]);
  const double phi[7] = {UL[5], UL[6], UL[7], UL[5] * vL[0] + UL[6] * vL[1] + UL[7] * vL[2], vL[0], vL[1], vL[2]};
  for (int v = 0; v < 7; v++) F[1 + v] = F[1 + v] + bn * phi[v];
}

/* DGTimeDerivative_weakForm (dg.f90:282-452) for ideal GLM-MHD with NONCONS=1 on LGL
   periodic conforming meshes: VolInt_SplitForm with EvalAdvFluxAverage3D +
   AddNonConsFluxTilde3D, FillFlux (Riemann + AddNonConsFlux on both sides), SurfInt, -sJ. */
/* FillFlux for GLM-MHD (src/dg/fillflux.f90:41-110 with NONCONS): the surface flux, in the
   resistive case plus the mean of the two sides' diffusive fluxes (mhd/riemann.f90:116-138) from
   the volume gradients (BR1) or the sides' face gradients fg (BR2), then each side's Powell term
   (AddNonConsFlux), times SurfElem -> FluxM (Flux_master), FluxS (Flux_slave) */
static void mhd_fill_flux(const ora_mhd *M, const double *U, double glm_ch, const double *gx, const double *gy,
                          const double *gz, double *const fg[6], double *FluxM, double *FluxS)
{
  const ora_problem *P = M->P;
  const int N = P->N, n = N + 1, n3 = n * n * n;
  const int br2 = M->parabolic && M->lifting == 2;
  for (int s = 0; s < P->nSides; s++) {
    const int *t = P->SideToElem + 5 * s;
    const int em = t[0] - 1, es = t[1] - 1, lm = t[2], ls = t[3], fl = t[4];
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        int im[3], is[3];
        ora_side_to_vol(N, 0, p, q, 0, lm, im);
        ora_side_to_vol(N, 0, p, q, fl, ls, is);
        const size_t om = MV * ((size_t)em * n3 + im[0] + n * (im[1] + n * im[2]));
        const size_t os = MV * ((size_t)es * n3 + is[0] + n * (is[1] + n * is[2]));
        const size_t f = (size_t)s * n * n + p + n * q;
        const double *nv = P->NormVec + 3 * f, *t1 = P->TangVec1 + 3 * f, *t2 = P->TangVec2 + 3 * f;
        double Fm[MV], Fs[MV];
        ora_mhd_riemann(M, glm_ch, U + om, U + os, nv, t1, t2, Fm);
        if (M->parabolic) {
          double kL[MV], gL[MV], jL[MV], kR[MV], gR[MV], jR[MV];
          if (br2) {
            const size_t fo = MV * f;
            mhd_diff_flux(M, U + om, fg[0] + fo, fg[2] + fo, fg[4] + fo, kL, gL, jL);
            mhd_diff_flux(M, U + os, fg[1] + fo, fg[3] + fo, fg[5] + fo, kR, gR, jR);
          } else {
            mhd_diff_flux(M, U + om, gx + om, gy + om, gz + om, kL, gL, jL);
            mhd_diff_flux(M, U + os, gx + os, gy + os, gz + os, kR, gR, jR);
          }
          for (int v = 1; v < MV; v++)
            Fm[v] = Fm[v] + 0.5 * (nv[0] * (kL[v] + kR[v]) + nv[1] * (gL[v] + gR[v]) + nv[2] * (jL[v] + jR[v]));
        }
        memcpy(Fs, Fm, sizeof(Fs));
        add_noncons_surface(U + om, U + os, nv, Fm);
        add_noncons_surface(U + os, U + om, nv, Fs);
        for (int v = 0; v < MV; v++) {
          FluxM[MV * f + v] = Fm[v] * P->SurfElem[f];
          FluxS[MV * f + v] = Fs[v] * P->SurfElem[f];
        }
      }
  }
}

void ora_mhd_fill_flux(const ora_mhd *M, const double *U, double glm_ch, double *FluxM, double *FluxS)
{
  const ora_problem *P = M->P;
  const int n = P->N + 1;
  const size_t nt = (size_t)MV * n * n * n * P->nElems;
  double *gx = NULL, *gy = NULL, *gz = NULL, *fg[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  if (M->parabolic) {
    gx = malloc(sizeof(double) * nt); gy = malloc(sizeof(double) * nt); gz = malloc(sizeof(double) * nt);
    if (M->lifting == 2) {
      for (int d = 0; d < 6; d++) fg[d] = malloc(sizeof(double) * MV * n * n * P->nSides);
      mhd_br2_gradients(M, U, gx, gy, gz, fg);
    } else {
      ora_mhd_lifting(M, U, gx, gy, gz);
    }
  }
  mhd_fill_flux(M, U, glm_ch, gx, gy, gz, fg, FluxM, FluxS);
  free(gx); free(gy); free(gz);
  for (int d = 0; d < 6; d++) free(fg[d]);
}

void ora_mhd_rhs(const ora_mhd *M, const double *U, double glm_ch, double *Ut)
{
  const ora_problem *P = M->P;
  const int N = P->N, n = N + 1, n3 = n * n * n;
  double x[16], w[16], D[256], Dh[256], Ds[256], Lm[16], Lp[16];
  ora_dg_operators(N, x, w, D, Dh, Ds, Lm, Lp);  /* Ds[i + n*l] = Dvolsurf(i,l) */
  const size_t E = P->nElems;
  memset(Ut, 0, sizeof(double) * MV * n3 * E);
  double *ft = malloc(sizeof(double) * MV * n * n3);
  const double *Met[3] = {P->Mf, P->Mg, P->Mh};
  for (size_t e = 0; e < E; e++) {
    const double *Ue = U + MV * n3 * e;
    double *ute = Ut + MV * n3 * e;
    for (int d = 0; d < 3; d++) {
      const double *Me = Met[d] + 3 * n3 * e;
      /* ft[(l, node)] = flux between node and the node l positions along direction d */
      for (int node = 0; node < n3; node++) {
        const int i = node % n, j = (node / n) % n, k = node / (n * n);
        const int a = (d == 0) ? i : (d == 1 ? j : k);
        for (int l = 0; l < n; l++) {
          const int nl = (d == 0) ? (l + n * (j + n * k)) : (d == 1 ? (i + n * (l + n * k)) : (i + n * (j + n * l)));
          double *f = ft + MV * (l + n * node);
          /* the symmetric two-point flux is evaluated with the lower line index as left state
             and copied to the other ordering, as EvalAdvFluxAverage3D does */
          if (l == a) { for (int v = 0; v < MV; v++) f[v] = 0.0; }
          else if (a < l) ora_mhd_derigs(M, glm_ch, Ue + MV * node, Ue + MV * nl, Me + 3 * node, Me + 3 * nl, f);
          else ora_mhd_derigs(M, glm_ch, Ue + MV * nl, Ue + MV * node, Me + 3 * nl, Me + 3 * node, f);
          /* AddNonConsFluxTilde3D (mhd/flux_average.f90:268-366), Powell, at node `node` */
          const double *ui = Ue + MV * node, *ul = Ue + MV * nl, *mi = Me + 3 * node, *ml = Me + 3 * nl;
          double vi[3];
          const double sri = 1.0 / ui[0];
          for (int c = 0; c < 3; c++) vi[c] = ui[1 + c] * sri;
          const double phi[7] = {0.5 * ui[5], 0.5 * ui[6], 0.5 * ui[7],
                                 0.5 * (vi[0] * ui[5] + vi[1] * ui[6] + vi[2] * ui[7]),
                                 0.5 * vi[0], 0.5 * vi[1], 0.5 * vi[2]};
          const double bhat = mi[0] * ui[5] + mi[1] * ui[6] + mi[2] * ui[7];
          const double bl = 0.5 * (mi[0] + ml[0]) * ul[5] + 0.5 * (mi[1] + ml[1]) * ul[6] + 0.5 * (mi[2] + ml[2]) * ul[7];
          for (int v = 0; v < 7; v++) f[1 + v] = f[1 + v] + phi[v] * (bhat + bl);
        }
      }
      for (int node = 0; node < n3; node++) {
        const int i = node % n, j = (node / n) % n, k = node / (n * n);
        const int a = (d == 0) ? i : (d == 1 ? j : k);
        for (int l = 0; l < n; l++)
          for (int v = 0; v < MV; v++) ute[MV * node + v] += Ds[a + n * l] * ft[MV * (l + n * node) + v];
      }
    }
  }
  free(ft);
  double *gx = NULL, *gy = NULL, *gz = NULL, *fg[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  const int br2 = M->parabolic && M->lifting == 2;
  if (M->parabolic) {
    const size_t nt = (size_t)MV * n3 * E;
    gx = malloc(sizeof(double) * nt); gy = malloc(sizeof(double) * nt); gz = malloc(sizeof(double) * nt);
    if (br2) {
      for (int d = 0; d < 6; d++) fg[d] = malloc(sizeof(double) * MV * n * n * P->nSides);
      mhd_br2_gradients(M, U, gx, gy, gz, fg);
    } else {
      ora_mhd_lifting(M, U, gx, gy, gz);
    }
    mhd_volint_visc(M, U, gx, gy, gz, Dh, Ut);
  }
  /* surface: FillFlux + SurfInt (LGL) */
  const size_t nface = (size_t)MV * n * n * P->nSides;
  double *FluxM = malloc(sizeof(double) * nface), *FluxS = malloc(sizeof(double) * nface);
  mhd_fill_flux(M, U, glm_ch, gx, gy, gz, fg, FluxM, FluxS);
  ora_surf_int(P, MV, FluxM, FluxS, Ut);
  free(FluxM); free(FluxS);
  for (size_t p = 0; p < (size_t)n3 * E; p++) {
    const double sj = -P->sJ[p];
    for (int v = 0; v < MV; v++) Ut[MV * p + v] = Ut[MV * p + v] * sj;
  }
  free(gx); free(gy); free(gz);
  for (int d = 0; d < 6; d++) free(fg[d]);
}

/* InitTimeStep_GLM (mhd/calctimestep.f90:186-215): dt of a unit wave speed */
double ora_mhd_glm_dtch1(const ora_mhd *M, double CFLscale)
{
  const ora_problem *P = M->P;
  const int n = P->N + 1, n3 = n * n * n;
  const double cfl = CFLscale * ora_cfl_alpha(P->N) / (2.0 * P->N + 1.0);
  double dtch1 = 1.79e308;
  const double *Met[3] = {P->Mf, P->Mg, P->Mh};
  for (int e = 0; e < P->nElems; e++) {
    double lam[3] = {0, 0, 0};
    for (int p = 0; p < n3; p++) {
      const size_t c = (size_t)e * n3 + p;
      for (int d = 0; d < 3; d++) {
        const double *m = Met[d] + 3 * c;
        lam[d] = fmax(lam[d], P->sJ[c] * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
      }
    }
    dtch1 = fmin(dtch1, cfl * 2.0 / (lam[0] + lam[1] + lam[2]));
  }
  return dtch1;
}

/* CALCTIMESTEP (mhd/calctimestep.f90:39-179): returns dt; *glm_ch receives
   GLM_scale * GLM_dtch1 / dt.  Resistive case: the element's dt is also limited by
   DFL*4/max_node(max(eta/mu_0, mu max(4/3, kappa/Pr)/rho) sum_d |sJ m_d|^2). */
double ora_mhd_calc_timestep(const ora_mhd *M, const double *U, double CFLscale, double DFLscale,
                             double glm_dtch1, double *glm_ch, int *errType)
{
  const ora_problem *P = M->P;
  const int n = P->N + 1, n3 = n * n * n;
  const double cfl = CFLscale * ora_cfl_alpha(P->N) / (2.0 * P->N + 1.0);
  const double dfl = DFLscale * ora_dfl_alpha(P->N) / ((2.0 * P->N + 1.0) * (2.0 * P->N + 1.0));
  const double muk = M->mu * fmax(4.0 / 3.0, M->kappa / M->Pr), diffc = M->eta * (1.0 / M->mu_0);
  double dt = 1.79e308, dtv = 1.79e308;
  int err = 0;
  const double *Met[3] = {P->Mf, P->Mg, P->Mh};
  for (int e = 0; e < P->nElems && !err; e++) {
    double lam[3] = {0, 0, 0}, lamv = 0.0;
    for (int p = 0; p < n3; p++) {
      const size_t c = (size_t)e * n3 + p;
      const double *u = U + MV * c;
      if (isnan(u[0])) err = 1;
      if (M->parabolic) {
        double lv = 0.0;
        for (int d = 0; d < 3; d++) {
          const double *m = Met[d] + 3 * c;
          const double a = m[0] * P->sJ[c], b = m[1] * P->sJ[c], cc = m[2] * P->sJ[c];
          lv += a * a + b * b + cc * cc;
        }
        lamv = fmax(lamv, fmax(diffc, (1.0 / u[0]) * muk) * lv);
      }
      double prim[MV];
      cons_to_prim(M, u, prim);
      if (prim[4] <= 0.0) err = 2;
      const double cf = sqrt((M->kappa * prim[4] + (prim[5] * prim[5] + prim[6] * prim[6] + prim[7] * prim[7]) * (1.0 / M->mu_0)) / prim[0]);
      for (int d = 0; d < 3; d++) {
        const double *m = Met[d] + 3 * c;
        lam[d] = fmax(lam[d], P->sJ[c] * (fabs(m[0] * prim[1] + m[1] * prim[2] + m[2] * prim[3]) +
                                          cf * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])));
      }
    }
    const double dte = cfl * 2.0 / (lam[0] + lam[1] + lam[2]);
    dt = fmin(dt, dte);
    if (isnan(dt)) err = 3;
    if (M->parabolic) {
      if (lamv > 0.0) dtv = fmin(dtv, dfl * 4.0 / lamv);
      if (isnan(dtv)) err = 4;
    }
  }
  dt = fmin(dt, dtv);
  if (errType) *errType = err;
  if (glm_ch) *glm_ch = M->glm_scale * (glm_dtch1 / dt);
  return dt;
}

/* TimeStepByLSERKW2 (timedisc.f90:424-480) with the cleaning speed of this step */
void ora_mhd_lserk_step(const ora_mhd *M, double *U, double *Ut_tmp, double glm_ch, double dt)
{
  const ora_problem *P = M->P;
  const size_t nt = (size_t)MV * (P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  double A[5], b[5], c[5];
  ora_rk_coefs(A, b, c);
  double *Ut = malloc(sizeof(double) * nt);
  for (int s = 0; s < 5; s++) {
    ora_mhd_rhs(M, U, glm_ch, Ut);
    for (size_t i = 0; i < nt; i++) {
      Ut_tmp[i] = (s == 0) ? Ut[i] : Ut_tmp[i] * (-A[s]) + Ut[i];
      U[i] = U[i] + Ut_tmp[i] * (b[s] * dt);
    }
  }
  free(Ut);
}

/* ExactFunc 312, 3D Orszag-Tang-type vortex (mhd/equation.f90:1257-1266) with PrimToCons
   (mhd/equation_vars.f90:253-266) */
void ora_mhd_orszag_tang(const ora_mhd *M, int npts, const double *x, double *U)
{
  const double pi = acos(-1.0), k = M->kappa;
  for (int p = 0; p < npts; p++) {
    const double *xp = x + 3 * p;
    double prim[MV] = {0};
    prim[0] = 1.0;
    prim[1] = -(1.0 + 0.2 * sin(2.0 * pi * xp[2])) * sin(2.0 * pi * xp[1]);
    prim[2] = (1.0 + 0.2 * sin(2.0 * pi * xp[2])) * sin(2.0 * pi * xp[0]);
    prim[3] = 0.2 * sin(2.0 * pi * xp[2]);
    prim[4] = 1.0 / k;
    prim[5] = -sin(2.0 * pi * xp[1]) / k;
    prim[6] = sin(4.0 * pi * xp[0]) / k;
    double *c = U + MV * p;
    c[0] = prim[0];
    for (int d = 1; d < 4; d++) c[d] = prim[d] * prim[0];
    double b2 = 0.0;
    for (int v = 5; v < MV; v++) b2 += prim[v] * prim[v];
    c[4] = (1.0 / (k - 1.0)) * prim[4] + 0.5 * (c[1] * prim[1] + c[2] * prim[2] + c[3] * prim[3]) + 0.5 * (1.0 / M->mu_0) * b2;
    for (int v = 5; v < MV; v++) c[v] = prim[v];
  }
}
