/*
 * Oracle (test infrastructure only): compressible Euler / Navier-Stokes physics and
 * the split-form DGSEM operator on LGL nodes, restating
 *   src/equation/navierstokes/{flux_average,riemann,flux,equation_vars,equation,
 *                              calctimestep,analyze_equation}.f90,
 *   src/dg/{dg,volint,fillflux,surfint}.f90, src/dg/br1/br1{,_volint,_fillflux,_surfint}.f90,
 *   src/timedisc/{timedisc,timedisc_vars}.f90.
 * Each function names the routine and lines it follows.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora.h"

#define NV 5

/* LN_MEAN (flux_average.f90:1945-1965) */
double ora_ln_mean(double aL, double aR)
{
  const double eps = 1.0e-4;
  double xi = aR / aL;
  double u = (xi * (xi - 2.0) + 1.0) / (xi * (xi + 2.0) + 1.0);
  if (u < eps) return (aL + aR) * 52.5 / (105.0 + u * (35.0 + u * (21.0 + u * 15.0)));
  return (aR - aL) / log(xi);
}

/* EvalUaux (flux_average.f90:305-332): aux = (1/rho, v1, v2, v3, p, |v|^2) */
void ora_eval_uaux(double kappa, const double *U, double *aux)
{
  double srho = 1.0 / U[0];
  double v1 = U[1] * srho, v2 = U[2] * srho, v3 = U[3] * srho;
  double vv = v1 * v1 + v2 * v2 + v3 * v3;
  aux[0] = srho; aux[1] = v1; aux[2] = v2; aux[3] = v3; aux[5] = vv;
  aux[4] = (kappa - 1.0) * (U[4] - 0.5 * U[0] * vv);
}

/* StandardDGFluxVec (flux_average.f90:394-439) */
static void flux_standard(const double *UL, const double *UR, const double *aL, const double *aR,
                          const double *mL, const double *mR, double *F)
{
  double qL = aL[1] * mL[0] + aL[2] * mL[1] + aL[3] * mL[2];
  double qR = aR[1] * mR[0] + aR[2] * mR[1] + aR[3] * mR[2];
  F[0] = 0.5 * (UL[0] * qL + UR[0] * qR);
  F[1] = 0.5 * (UL[1] * qL + UR[1] * qR + mL[0] * aL[4] + mR[0] * aR[4]);
  F[2] = 0.5 * (UL[2] * qL + UR[2] * qR + mL[1] * aL[4] + mR[1] * aR[4]);
  F[3] = 0.5 * (UL[3] * qL + UR[3] * qR + mL[2] * aL[4] + mR[2] * aR[4]);
  F[4] = 0.5 * ((UL[4] + aL[4]) * qL + (UR[4] + aR[4]) * qR);
}

/* EntropyAndEnergyConservingFluxVec, Chandrashekar EC-KEP (flux_average.f90:1277-1356) */
static void flux_chandrashekar(double kappa, const double *UL, const double *UR, const double *aL,
                               const double *aR, const double *mL, const double *mR, double *F)
{
  double skm1 = 1.0 / (kappa - 1.0);
  double m[3] = {0.5 * (mL[0] + mR[0]), 0.5 * (mL[1] + mR[1]), 0.5 * (mL[2] + mR[2])};
  double bL = 0.5 * UL[0] / aL[4], bR = 0.5 * UR[0] / aR[4];
  double rho_mean = 0.5 * (UL[0] + UR[0]);
  double rhoHat = ora_ln_mean(UL[0], UR[0]);
  double beta_mean = 0.5 * (bL + bR);
  double beta_hat = ora_ln_mean(bL, bR);
  double uH = 0.5 * (aL[1] + aR[1]), vH = 0.5 * (aL[2] + aR[2]), wH = 0.5 * (aL[3] + aR[3]);
  double p1 = 0.5 * rho_mean / beta_mean;
  double qH = uH * m[0] + vH * m[1] + wH * m[2];
  F[0] = rhoHat * qH;
  F[1] = F[0] * uH + m[0] * p1;
  F[2] = F[0] * vH + m[1] * p1;
  F[3] = F[0] * wH + m[2] * p1;
  F[4] = F[0] * 0.5 * (skm1 / beta_hat - 0.5 * (aL[5] + aR[5])) + uH * F[1] + vH * F[2] + wH * F[3];
}

/* RanochaFluxVec (flux_average.f90:1876-1929) */
static void flux_ranocha(double kappa, const double *UL, const double *UR, const double *aL,
                         const double *aR, const double *mL, const double *mR, double *F)
{
  double skm1 = 1.0 / (kappa - 1.0);
  double m[3] = {0.5 * (mL[0] + mR[0]), 0.5 * (mL[1] + mR[1]), 0.5 * (mL[2] + mR[2])};
  double rhoLN = ora_ln_mean(UL[0], UR[0]);
  double va[3] = {0.5 * (aL[1] + aR[1]), 0.5 * (aL[2] + aR[2]), 0.5 * (aL[3] + aR[3])};
  double v2zip = 0.5 * (aL[1] * aR[1] + aL[2] * aR[2] + aL[3] * aR[3]);
  double pavg = 0.5 * (aL[4] + aR[4]);
  double vmL = aL[1] * m[0] + aL[2] * m[1] + aL[3] * m[2];
  double vmR = aR[1] * m[0] + aR[2] * m[1] + aR[3] * m[2];
  double rspL = UL[0] / aL[4], rspR = UR[0] / aR[4];
  double rspM = ora_ln_mean(rspL, rspR);
  F[0] = rhoLN * 0.5 * (vmL + vmR);
  F[1] = F[0] * va[0] + m[0] * pavg;
  F[2] = F[0] * va[1] + m[1] * pavg;
  F[3] = F[0] * va[2] + m[2] * pavg;
  F[4] = F[0] * (v2zip + 1.0 / rspM * skm1) + 0.5 * (aL[4] * vmR + aR[4] * vmL);
}

/* TwoPointEntropyConservingFluxVec, Ismail-Roe (flux_average.f90:574-673) */
static void flux_ismail_roe(double kappa, const double *UL, const double *UR, const double *aL,
                            const double *aR, const double *mL, const double *mR, double *F)
{
  const double skm1 = 1.0 / (kappa - 1.0), kp1 = kappa + 1.0;
  double m[3] = {0.5 * (mL[0] + mR[0]), 0.5 * (mL[1] + mR[1]), 0.5 * (mL[2] + mR[2])};
  const double z1L = sqrt(UL[0] / aL[4]), z1R = sqrt(UR[0] / aR[4]);
  const double z1M = 0.5 * (z1L + z1R), z1LN = ora_ln_mean(z1L, z1R);
  const double z2M = 0.5 * (z1L * aL[1] + z1R * aR[1]);
  const double z3M = 0.5 * (z1L * aL[2] + z1R * aR[2]);
  const double z4M = 0.5 * (z1L * aL[3] + z1R * aR[3]);
  const double z5L = z1L * aL[4], z5R = z1R * aR[4];
  const double z5M = 0.5 * (z5L + z5R), z5LN = ora_ln_mean(z5L, z5R);
  const double rhoHat = z1M * z5LN, sz1 = 1.0 / z1M;
  const double uH = z2M * sz1, vH = z3M * sz1, wH = z4M * sz1, p1 = z5M * sz1;
  const double qH = uH * m[0] + vH * m[1] + wH * m[2];
  F[0] = rhoHat * qH;
  F[1] = F[0] * uH + m[0] * p1;
  F[2] = F[0] * vH + m[1] * p1;
  F[3] = F[0] * wH + m[2] * p1;
  F[4] = 0.5 * (qH * kp1 * skm1 * (z5LN / z1LN) + uH * F[1] + vH * F[2] + wH * F[3]);
}

/* The averaged metric {{m}} and the contravariant velocity {{v}}.{{m}} shared by the fluxes
   below (each starts with metric = 0.5*(metric_L+metric_R)). */
static void mean_metric(const double *mL, const double *mR, const double *aL, const double *aR, double *m,
                        double *vh, double *qh)
{
  for (int d = 0; d < 3; d++) {
    m[d] = 0.5 * (mL[d] + mR[d]);
    vh[d] = 0.5 * (aL[1 + d] + aR[1 + d]);
  }
  *qh = vh[0] * m[0] + vh[1] * m[1] + vh[2] * m[2];
}

/* StandardDGFluxDealiasedMetricVec (flux_average.f90:446-494) */
static void flux_standard_dealiased(const double *UL, const double *UR, const double *aL, const double *aR,
                                    const double *mL, const double *mR, double *F)
{
  double m[3] = {0.5 * (mL[0] + mR[0]), 0.5 * (mL[1] + mR[1]), 0.5 * (mL[2] + mR[2])};
  double qL = aL[1] * m[0] + aL[2] * m[1] + aL[3] * m[2];
  double qR = aR[1] * m[0] + aR[2] * m[1] + aR[3] * m[2];
  F[0] = 0.5 * (UL[0] * qL + UR[0] * qR);
  for (int d = 0; d < 3; d++) F[1 + d] = 0.5 * (UL[1 + d] * qL + UR[1 + d] * qR + m[d] * (aL[4] + aR[4]));
  F[4] = 0.5 * ((UL[4] + aL[4]) * qL + (UR[4] + aR[4]) * qR);
}

/* KennedyAndGruberFluxVec1 (flux_average.f90:731-803) and KennedyAndGruberFluxVec2
   (flux_average.f90:857-929): the energy flux with {{e}} = {{rhoE/rho}} (KG1) or the
   enthalpy {{H}} = {{(rhoE+p)/rho}} (KG2). */
static void flux_kennedy_gruber(int which, const double *UL, const double *UR, const double *aL,
                                const double *aR, const double *mL, const double *mR, double *F)
{
  double m[3], vh[3], qh;
  mean_metric(mL, mR, aL, aR, m, vh, &qh);
  double rhoHat = 0.5 * (UL[0] + UR[0]), p1 = 0.5 * (aL[4] + aR[4]);
  F[0] = rhoHat * qh;
  for (int d = 0; d < 3; d++) F[1 + d] = F[0] * vh[d] + m[d] * p1;
  if (which == 2)
    F[4] = F[0] * (0.5 * (aL[0] * UL[4] + aR[0] * UR[4])) + p1 * qh;
  else
    F[4] = F[0] * (0.5 * ((UL[4] + aL[4]) * aL[0] + (UR[4] + aR[4]) * aR[0]));
}

/* DucrosFluxVec (flux_average.f90:986-1059) */
static void flux_ducros(const double *UL, const double *UR, const double *aL, const double *aR, const double *mL,
                        const double *mR, double *F)
{
  double m[3], vh[3], qh;
  mean_metric(mL, mR, aL, aR, m, vh, &qh);
  double p1 = 0.5 * (aL[4] + aR[4]);
  F[0] = 0.5 * (UL[0] + UR[0]) * qh;
  for (int d = 0; d < 3; d++) F[1 + d] = 0.5 * (UL[1 + d] + UR[1 + d]) * qh + m[d] * p1;
  F[4] = (0.5 * (UL[4] + UR[4]) + p1) * qh;
}

/* MorinishiFluxVec (flux_average.f90:1121-1205) */
static void flux_morinishi(double kappa, const double *UL, const double *UR, const double *aL, const double *aR,
                           const double *mL, const double *mR, double *F)
{
  double m[3], vh[3], qh;
  mean_metric(mL, mR, aL, aR, m, vh, &qh);
  double p1 = 0.5 * (aL[4] + aR[4]);
  double qL = aL[1] * m[0] + aL[2] * m[1] + aL[3] * m[2];
  double qR = aR[1] * m[0] + aR[2] * m[1] + aR[3] * m[2];
  double rqL = 0.5 * UL[0] * qL, rqR = 0.5 * UR[0] * qR;
  F[0] = rqL + rqR;
  for (int d = 0; d < 3; d++) F[1 + d] = F[0] * vh[d] + m[d] * p1;
  F[4] = 0.5 * (kappa / (kappa - 1.0) * (aL[4] * qL + aR[4] * qR) - (rqL * aL[5] + rqR * aR[5])) +
         rqL * (aL[1] * vh[0] + aL[2] * vh[1] + aL[3] * vh[2]) + rqR * (aR[1] * vh[0] + aR[2] * vh[1] + aR[3] * vh[2]);
}

/* EntropyAndEnergyConservingFluxVec2 (flux_average.f90:1427-1504) with arithmetic means of rho
   and beta = rho/(2p), and ggFluxVec (flux_average.f90:1580-1653) with the log mean of beta. */
static void flux_ec2_gg(int which, double kappa, const double *UL, const double *UR, const double *aL,
                        const double *aR, const double *mL, const double *mR, double *F)
{
  double m[3], vh[3], qh;
  mean_metric(mL, mR, aL, aR, m, vh, &qh);
  const double skm1 = 1.0 / (kappa - 1.0);
  double bL = 0.5 * UL[0] / aL[4], bR = 0.5 * UR[0] / aR[4];
  double rho_mean = 0.5 * (UL[0] + UR[0]);
  double sbeta, p1;
  if (which == 6) {
    sbeta = 2.0 / (bL + bR);
    p1 = 0.5 * rho_mean * sbeta;
  } else {
    sbeta = 1.0 / ora_ln_mean(bL, bR);
    p1 = 0.5 * (aL[4] + aR[4]);
  }
  F[0] = rho_mean * qh;
  for (int d = 0; d < 3; d++) F[1 + d] = F[0] * vh[d] + m[d] * p1;
  F[4] = F[0] * 0.5 * (skm1 * sbeta - 0.5 * (aL[5] + aR[5])) + vh[0] * F[1] + vh[1] * F[2] + vh[2] * F[3];
}

/* GassnerWintersWalchFluxVec (flux_average.f90:1724-1799): averages of sqrt(rho)-weighted
   variables */
static void flux_gww(double kappa, const double *UL, const double *UR, const double *aL, const double *aR,
                     const double *mL, const double *mR, double *F)
{
  double m[3] = {0.5 * (mL[0] + mR[0]), 0.5 * (mL[1] + mR[1]), 0.5 * (mL[2] + mR[2])};
  double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
  double rhoHat = 0.5 * (sL + sR), h[3];
  for (int d = 0; d < 3; d++) h[d] = 0.5 * (sL * aL[1 + d] + sR * aR[1 + d]);
  double Hh = 0.5 * (sL * aL[0] * (UL[4] + aL[4]) + sR * aR[0] * (UR[4] + aR[4]));
  double p1 = (kappa - 1.0) / kappa * (rhoHat * Hh - 0.5 * (h[0] * h[0] + h[1] * h[1] + h[2] * h[2]));
  double qh = h[0] * m[0] + h[1] * m[1] + h[2] * m[2];
  F[0] = rhoHat * qh;
  for (int d = 0; d < 3; d++) F[1 + d] = qh * h[d] + m[d] * p1;
  F[4] = qh * Hh;
}

void ora_volflux(int which, double kappa, const double *UL, const double *UR, const double *auxL,
                 const double *auxR, const double *mL, const double *mR, double *F)
{
  switch (which) {
  case 1: flux_standard_dealiased(UL, UR, auxL, auxR, mL, mR, F); break;
  case 2: flux_kennedy_gruber(2, UL, UR, auxL, auxR, mL, mR, F); break;
  case 8: flux_kennedy_gruber(8, UL, UR, auxL, auxR, mL, mR, F); break;
  case 3: flux_ducros(UL, UR, auxL, auxR, mL, mR, F); break;
  case 4: flux_morinishi(kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 6: flux_ec2_gg(6, kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 7: flux_ec2_gg(7, kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 9: flux_gww(kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 5: flux_chandrashekar(kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 10: flux_ismail_roe(kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  case 32: flux_ranocha(kappa, UL, UR, auxL, auxR, mL, mR, F); break;
  default: flux_standard(UL, UR, auxL, auxR, mL, mR, F); break;
  }
}

/* EvalEulerFlux1D (flux.f90:298-329) */
static void euler_flux_1d(double kappa, const double *U, double *F)
{
  double srho = 1.0 / U[0];
  double v1 = U[1] * srho, v2 = U[2] * srho, v3 = U[3] * srho;
  double p = (kappa - 1.0) * (U[4] - 0.5 * (U[1] * v1 + U[2] * v2 + U[3] * v3));
  F[0] = U[1];
  F[1] = U[1] * v1 + p;
  F[2] = U[1] * v2;
  F[3] = U[1] * v3;
  F[4] = (U[4] + p) * v1;
}

/* SoundSpeed2 (equation_vars.f90:342-356) */
static double sound_speed2(double kappa, const double *U)
{
  return kappa * (kappa - 1.0) * (U[4] * U[0] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3])) / (U[0] * U[0]);
}

/* 1D (already rotated) two-point fluxes used as EC surface fluxes:
   EntropyAndEnergyConservingFlux (flux_average.f90:1214-1270) and RanochaFlux (:1806-1866) */
static void flux1d_chandrashekar(double kappa, const double *UL, const double *UR, double *F)
{
  double km1 = kappa - 1.0, skm1 = 1.0 / km1;
  double sL = 1.0 / UL[0], sR = 1.0 / UR[0];
  double uL = sL * UL[1], vL = sL * UL[2], wL = sL * UL[3];
  double uR = sR * UR[1], vR = sR * UR[2], wR = sR * UR[3];
  double v2s2L = 0.5 * (uL * uL + vL * vL + wL * wL), v2s2R = 0.5 * (uR * uR + vR * vR + wR * wR);
  double pL = km1 * (UL[4] - UL[0] * v2s2L), pR = km1 * (UR[4] - UR[0] * v2s2R);
  double bL = 0.5 * UL[0] / pL, bR = 0.5 * UR[0] / pR;
  double rho_mean = 0.5 * (UL[0] + UR[0]);
  double rhoHat = ora_ln_mean(UL[0], UR[0]);
  double beta_mean = 0.5 * (bL + bR);
  double beta_hat = ora_ln_mean(bL, bR);
  double uH = 0.5 * (uL + uR), vH = 0.5 * (vL + vR), wH = 0.5 * (wL + wR);
  double p1 = 0.5 * rho_mean / beta_mean;
  double v2m = v2s2L + v2s2R;
  F[0] = rhoHat * uH;
  F[1] = F[0] * uH + p1;
  F[2] = F[0] * vH;
  F[3] = F[0] * wH;
  F[4] = F[0] * 0.5 * (skm1 / beta_hat - v2m) + uH * F[1] + vH * F[2] + wH * F[3];
}

static void flux1d_ranocha(double kappa, const double *UL, const double *UR, double *F)
{
  double km1 = kappa - 1.0, skm1 = 1.0 / km1;
  double sL = 1.0 / UL[0], sR = 1.0 / UR[0];
  double vL[3] = {UL[1] * sL, UL[2] * sL, UL[3] * sL}, vR[3] = {UR[1] * sR, UR[2] * sR, UR[3] * sR};
  double v2L = vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2];
  double v2R = vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2];
  double pL = km1 * (UL[4] - 0.5 * UL[0] * v2L), pR = km1 * (UR[4] - 0.5 * UR[0] * v2R);
  double rhoHat = ora_ln_mean(UL[0], UR[0]);
  double uH = 0.5 * (vL[0] + vR[0]), vH = 0.5 * (vL[1] + vR[1]), wH = 0.5 * (vL[2] + vR[2]);
  double v2zip = 0.5 * (vL[0] * vR[0] + vL[1] * vR[1] + vL[2] * vR[2]);
  double p1 = 0.5 * (pL + pR);
  double rspM = ora_ln_mean(UL[0] / pL, UR[0] / pR);
  F[0] = rhoHat * uH;
  F[1] = F[0] * uH + p1;
  F[2] = F[0] * vH;
  F[3] = F[0] * wH;
  F[4] = F[0] * (v2zip + 1.0 / rspM * skm1) + 0.5 * (pL * vR[0] + pR * vL[0]);
}

/* The 1D EC fluxes with the averaged quantities the entropy-stable dissipation uses
   (uHat, vHat, wHat, aHat, HHat, p1Hat, rhoHat as in the reference's signatures):
   TwoPointEntropyConservingFlux / Ismail-Roe (flux_average.f90:500-572),
   EntropyAndEnergyConservingFlux / Chandrashekar (:1214-1269), RanochaFlux (:1806-1866). */
typedef struct { double u, v, w, a, H, p1, rho; } hat_t;

static void flux1d_hat(int which, double kappa, const double *UL, const double *UR, double *F, hat_t *h)
{
  const double km1 = kappa - 1.0, skm1 = 1.0 / km1;
  const double sL = 1.0 / UL[0], sR = 1.0 / UR[0];
  const double uL = sL * UL[1], vL = sL * UL[2], wL = sL * UL[3];
  const double uR = sR * UR[1], vR = sR * UR[2], wR = sR * UR[3];
  if (which == 4) {  /* Ismail-Roe */
    const double pL = km1 * (UL[4] - 0.5 * (UL[1] * uL + UL[2] * vL + UL[3] * wL));
    const double pR = km1 * (UR[4] - 0.5 * (UR[1] * uR + UR[2] * vR + UR[3] * wR));
    const double z1L = sqrt(UL[0] / pL), z1R = sqrt(UR[0] / pR);
    const double z1M = 0.5 * (z1L + z1R), z1LN = ora_ln_mean(z1L, z1R);
    const double z2M = 0.5 * (z1L * uL + z1R * uR), z3M = 0.5 * (z1L * vL + z1R * vR), z4M = 0.5 * (z1L * wL + z1R * wR);
    const double z5L = pL * z1L, z5R = pR * z1R;
    const double z5M = 0.5 * (z5L + z5R), z5LN = ora_ln_mean(z5L, z5R);
    h->rho = z1M * z5LN;
    const double sz1 = 1.0 / z1M;
    h->u = z2M * sz1; h->v = z3M * sz1; h->w = z4M * sz1;
    h->p1 = z5M * sz1;
    const double kp2 = 0.5 * ((kappa + 1.0) * (z5LN / z1LN) + km1 * h->p1);
    h->a = sqrt(kp2 / h->rho);
    h->H = kp2 / (h->rho * km1) + 0.5 * (h->u * h->u + h->v * h->v + h->w * h->w);
    F[0] = h->rho * h->u;
    F[1] = F[0] * h->u + h->p1;
    F[2] = F[0] * h->v;
    F[3] = F[0] * h->w;
    F[4] = F[0] * h->H;
  } else if (which == 44 || which == 16) {  /* Chandrashekar */
    const double v2s2L = 0.5 * (uL * uL + vL * vL + wL * wL), v2s2R = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double pL = km1 * (UL[4] - UL[0] * v2s2L), pR = km1 * (UR[4] - UR[0] * v2s2R);
    const double bL = 0.5 * UL[0] / pL, bR = 0.5 * UR[0] / pR;
    const double rho_mean = 0.5 * (UL[0] + UR[0]);
    h->rho = ora_ln_mean(UL[0], UR[0]);
    const double beta_mean = 0.5 * (bL + bR), beta_hat = ora_ln_mean(bL, bR);
    h->u = 0.5 * (uL + uR); h->v = 0.5 * (vL + vR); h->w = 0.5 * (wL + wR);
    h->a = sqrt(kappa * 0.5 * (pL + pR) / h->rho);
    h->H = kappa / (2.0 * km1 * beta_hat) + 0.5 * (uL * uR + vL * vR + wL * wR);
    h->p1 = 0.5 * rho_mean / beta_mean;
    F[0] = h->rho * h->u;
    F[1] = F[0] * h->u + h->p1;
    F[2] = F[0] * h->v;
    F[3] = F[0] * h->w;
    F[4] = F[0] * 0.5 * (skm1 / beta_hat - (v2s2L + v2s2R)) + h->u * F[1] + h->v * F[2] + h->w * F[3];
  } else {  /* Ranocha (33) */
    const double v2L = uL * uL + vL * vL + wL * wL, v2R = uR * uR + vR * vR + wR * wR;
    const double pL = km1 * (UL[4] - 0.5 * UL[0] * v2L), pR = km1 * (UR[4] - 0.5 * UR[0] * v2R);
    h->rho = ora_ln_mean(UL[0], UR[0]);
    h->u = 0.5 * (uL + uR); h->v = 0.5 * (vL + vR); h->w = 0.5 * (wL + wR);
    const double v2zip = 0.5 * (uL * uR + vL * vR + wL * wR);
    h->p1 = 0.5 * (pL + pR);
    h->a = sqrt(kappa * h->p1 / h->rho);
    h->H = kappa * h->p1 * skm1 / h->rho + 0.5 * (h->u * h->u + h->v * h->v + h->w * h->w);
    const double rspM = ora_ln_mean(UL[0] / pL, UR[0] / pR);
    F[0] = h->rho * h->u;
    F[1] = F[0] * h->u + h->p1;
    F[2] = F[0] * h->v;
    F[3] = F[0] * h->w;
    F[4] = F[0] * (v2zip + 1.0 / rspM * skm1) + 0.5 * (pL * uR + pR * uL);
  }
}

/* RiemannSolver_EntropyStable (riemann.f90:874-931) with
   RiemannSolver_EntropyStable_VolFluxAndDissipMatrices (:936-975): EC flux minus
   1/2 R D R^T [v], v the entropy variables, R the right eigenvectors, D the scaled
   eigenvalues; rotated frame. */
static void riemann_es(int which, double kappa, const double *uLL, const double *uRR, double *f)
{
  const double km1 = kappa - 1.0;
  double Fc[5];
  hat_t h;
  flux1d_hat(which, kappa, uLL, uRR, Fc, &h);
  double R[5][5] = {{1.0, 1.0, 0.0, 0.0, 1.0},
                    {h.u - h.a, h.u, 0.0, 0.0, h.u + h.a},
                    {h.v, h.v, 1.0, 0.0, h.v},
                    {h.w, h.w, 0.0, 1.0, h.w},
                    {h.H - h.u * h.a, 0.5 * (h.u * h.u + h.v * h.v + h.w * h.w), h.v, h.w, h.H + h.u * h.a}};
  double D[5];
  D[0] = 0.5 * fabs(h.u - h.a) * h.rho / kappa;
  D[1] = fabs(h.u) * (km1 / kappa) * h.rho;
  D[2] = fabs(h.u) * h.p1;
  D[3] = D[2];
  D[4] = 0.5 * fabs(h.u + h.a) * h.rho / kappa;
  const double sL = 1.0 / uLL[0], sR = 1.0 / uRR[0];
  const double VL[3] = {uLL[1] * sL, uLL[2] * sL, uLL[3] * sL}, VR[3] = {uRR[1] * sR, uRR[2] * sR, uRR[3] * sR};
  const double pL = km1 * (uLL[4] - 0.5 * (uLL[1] * VL[0] + uLL[2] * VL[1] + uLL[3] * VL[2]));
  const double pR = km1 * (uRR[4] - 0.5 * (uRR[1] * VR[0] + uRR[2] * VR[1] + uRR[3] * VR[2]));
  const double sRs = log(pR) - kappa * log(uRR[0]), sLs = log(pL) - kappa * log(uLL[0]);
  const double rpR = uRR[0] / pR, rpL = uLL[0] / pL;
  double vL[5], vR[5], dv[5], d[5], diss[5];
  vL[0] = (kappa - sLs) / km1 - 0.5 * rpL * (VL[0] * VL[0] + VL[1] * VL[1] + VL[2] * VL[2]);
  vR[0] = (kappa - sRs) / km1 - 0.5 * rpR * (VR[0] * VR[0] + VR[1] * VR[1] + VR[2] * VR[2]);
  for (int c = 0; c < 3; c++) { vL[1 + c] = rpL * VL[c]; vR[1 + c] = rpR * VR[c]; }
  vR[4] = -rpR;
  vL[4] = -rpL;
  for (int c = 0; c < 5; c++) dv[c] = vR[c] - vL[c];
  for (int c = 0; c < 5; c++)  /* R^T [v] */
    d[c] = R[0][c] * dv[0] + R[1][c] * dv[1] + R[2][c] * dv[2] + R[3][c] * dv[3] + R[4][c] * dv[4];
  for (int c = 0; c < 5; c++) d[c] = D[c] * d[c];
  for (int r = 0; r < 5; r++)
    diss[r] = R[r][0] * d[0] + R[r][1] * d[1] + R[r][2] * d[2] + R[r][3] * d[3] + R[r][4] * d[4];
  for (int c = 0; c < 5; c++) f[c] = Fc[c] - 0.5 * diss[c];
}

/* RiemannSolver_ECKEP_LLF (riemann.f90:1157-1209): Chandrashekar flux with a local
   Lax-Friedrichs dissipation whose energy part is consistent with the entropy variables;
   lam is the largest wave speed over the whole face (LambdaMaxMax). */
static void riemann_eckep_llf(double kappa, const double *uLL, const double *uRR, double lam, double *f)
{
  const double km1 = kappa - 1.0;
  double Fc[5];
  hat_t h;
  flux1d_hat(16, kappa, uLL, uRR, Fc, &h);
  const double sL = 1.0 / uLL[0], sR = 1.0 / uRR[0];
  const double VL[3] = {uLL[1] * sL, uLL[2] * sL, uLL[3] * sL}, VR[3] = {uRR[1] * sR, uRR[2] * sR, uRR[3] * sR};
  const double pL = km1 * (uLL[4] - 0.5 * (uLL[1] * VL[0] + uLL[2] * VL[1] + uLL[3] * VL[2]));
  const double pR = km1 * (uRR[4] - 0.5 * (uRR[1] * VR[0] + uRR[2] * VR[1] + uRR[3] * VR[2]));
  const double bL = 0.5 * uLL[0] / pL, bR = 0.5 * uRR[0] / pR;
  const double rhoa = 0.5 * (uRR[0] + uLL[0]);
  const double bLN = ora_ln_mean(bL, bR);
  for (int c = 0; c < 4; c++) f[c] = Fc[c] - 0.5 * lam * (uRR[c] - uLL[c]);
  f[4] = Fc[4] - 0.5 * lam * ((0.5 / (km1 * bLN) + 0.5 * (VL[0] * VR[0] + VL[1] * VR[1] + VL[2] * VR[2])) * (uRR[0] - uLL[0]) +
                              rhoa * (h.u * (VR[0] - VL[0]) + h.v * (VR[1] - VL[1]) + h.w * (VR[2] - VL[2]) +
                                      0.5 / km1 * (1.0 / bR - 1.0 / bL)));
}

/* The 1D flux of RiemannSolver_VolumeFluxAverage (riemann.f90:980-1005) and
   RiemannSolver_VolumeFluxAverage_LLF (:1118-1151) in the rotated frame, selected as
   equation.f90:292-404 does: 5 Ismail-Roe (TwoPointEntropyConservingFlux), 6 / 14
   KennedyAndGruberFlux1, 7 / 15 DucrosFlux, 8 / 19 MorinishiFlux, 10
   EntropyAndEnergyConservingFlux2, 11 / 17 ggflux, 12 / 18 KennedyAndGruberFlux2, 20 / 21
   GassnerWintersWalchFlux.  These equal the metric forms of the volume fluxes with the
   metric (1,0,0), except ggflux (flux_average.f90:1513-1574), whose mass flux uses the log
   mean of rho where ggFluxVec uses the arithmetic mean. */
static int riemann_volflux_of(int which)
{
  switch (which) {
  case 5: return 10;
  case 6: case 14: return 2;
  case 7: case 15: return 3;
  case 8: case 19: return 4;
  case 10: return 6;
  case 11: case 17: return 7;
  case 12: case 18: return 8;
  case 20: case 21: return 9;
  default: return -1;
  }
}

static void riemann_volflux_avg(int which, double kappa, const double *uLL, const double *uRR, double lam,
                                double *f)
{
  const double e1[3] = {1.0, 0.0, 0.0};
  double aL[6], aR[6];
  ora_eval_uaux(kappa, uLL, aL);
  ora_eval_uaux(kappa, uRR, aR);
  if (which == 5) {  /* TwoPointEntropyConservingFlux with its energy flux F1*HHat */
    hat_t h;
    flux1d_hat(4, kappa, uLL, uRR, f, &h);
    return;
  }
  ora_volflux(riemann_volflux_of(which), kappa, uLL, uRR, aL, aR, e1, e1, f);
  if (which == 11 || which == 17) {  /* ggflux: F1 = LN_MEAN(rho) uHat */
    const double uH = 0.5 * (aL[1] + aR[1]), vH = 0.5 * (aL[2] + aR[2]), wH = 0.5 * (aL[3] + aR[3]);
    const double bL = 0.5 * uLL[0] / aL[4], bR = 0.5 * uRR[0] / aR[4];
    const double p1 = 0.5 * (aL[4] + aR[4]);
    f[0] = ora_ln_mean(uLL[0], uRR[0]) * uH;
    f[1] = f[0] * uH + p1;
    f[2] = f[0] * vH;
    f[3] = f[0] * wH;
    f[4] = f[0] * 0.5 * (1.0 / ((kappa - 1.0) * ora_ln_mean(bL, bR)) - 0.5 * (aL[5] + aR[5])) + uH * f[1] +
           vH * f[2] + wH * f[3];
  }
  if (which == 14 || which == 15 || which == 17 || which == 18 || which == 19 || which == 21)
    for (int v = 0; v < 5; v++) f[v] = f[v] - 0.5 * lam * (uRR[v] - uLL[v]);
}

/* RiemannSolver_EC_LLF (riemann.f90:1011-1112): the Ismail-Roe flux minus 1/2 lam H [v],
   H = dU/dv at the Ismail-Roe averages, [v] the jump of the entropy variables, lam the
   face's LambdaMaxMax; rotated frame. */
static void riemann_ec_llf(double kappa, const double *uLL, const double *uRR, double lam, double *f)
{
  const double km1 = kappa - 1.0;
  double Fc[5], dv[5], H[5][5];
  hat_t h;
  flux1d_hat(4, kappa, uLL, uRR, Fc, &h);
  /* [v] with the reference's evaluation order: s = ln p - kappa ln rho */
  const double *U2[2] = {uLL, uRR};
  for (int c = 0; c < 5; c++) dv[c] = 0.0;
  for (int k = 0; k < 2; k++) {
    const double *u = U2[k];
    const double sr = 1.0 / u[0], vel[3] = {u[1] * sr, u[2] * sr, u[3] * sr};
    const double p = km1 * (u[4] - 0.5 * (u[1] * vel[0] + u[2] * vel[1] + u[3] * vel[2]));
    const double se = log(p) - kappa * log(u[0]);
    const double w1 = (kappa - se) / km1 - 0.5 * u[0] * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) / p;
    dv[0] = k == 0 ? -w1 : dv[0] + w1;
    for (int d = 0; d < 3; d++) dv[1 + d] = k == 0 ? -u[0] * vel[d] / p : dv[1 + d] + u[0] * vel[d] / p;
    dv[4] = k == 0 ? u[0] / p : dv[4] - u[0] / p;
  }
  const double va[3] = {h.u, h.v, h.w};
  const double rhoea = h.p1 / km1 + 0.5 * h.rho * (h.u * h.u + h.v * h.v + h.w * h.w);
  H[0][0] = h.rho;
  H[4][4] = h.rho * h.H * h.H - h.a * h.a * h.p1 / km1;
  H[0][4] = H[4][0] = rhoea;
  for (int i = 0; i < 3; i++) {
    H[0][1 + i] = H[1 + i][0] = h.rho * va[i];
    H[4][1 + i] = H[1 + i][4] = h.rho * va[i] * h.H;
    for (int j = 0; j < 3; j++) H[1 + i][1 + j] = h.rho * va[i] * va[j] + (i == j ? h.p1 : 0.0);
  }
  for (int r = 0; r < 5; r++) {
    double d = 0.0;
    for (int c = 0; c < 5; c++) d += H[r][c] * dv[c];
    f[r] = Fc[r] - 0.5 * lam * d;
  }
}

/* HLL-type and Roe solvers in the rotated frame: RiemannSolverByHLL (riemann.f90:320-388),
   ByHLLE (:396-462), ByHLLEM (:468-547), ByHLLC (:552-615) and ByRoe (:746-869), whose central
   part follows the volume flux (PP_DiscType 2: VolumeFlux 2 KennedyAndGruberFlux1, 3
   DucrosFlux, 4 KennedyAndGruberFlux2 (sic), otherwise the average of the Euler fluxes; with
   2 and 4 the acoustic waves take max(|u-c|,|u+c|)). */
static void riemann_hll_roe(int which, int volflux, double kappa, const double *uLL, const double *uRR, double *f)
{
  const double km1 = kappa - 1.0;
  double FL[5], FR[5];
  euler_flux_1d(kappa, uLL, FL);
  euler_flux_1d(kappa, uRR, FR);
  const double sL = 1.0 / uLL[0], sR = 1.0 / uRR[0];
  const double vL[3] = {uLL[1] * sL, uLL[2] * sL, uLL[3] * sL}, vR[3] = {uRR[1] * sR, uRR[2] * sR, uRR[3] * sR};
  const double pL = km1 * (uLL[4] - 0.5 * (uLL[1] * vL[0] + uLL[2] * vL[1] + uLL[3] * vL[2]));
  const double pR = km1 * (uRR[4] - 0.5 * (uRR[1] * vR[0] + uRR[2] * vR[1] + uRR[3] * vR[2]));
  if (which == 2) {  /* HLLC */
    const double Ssl = vL[0] - sqrt(kappa * pL * sL), Ssr = vR[0] + sqrt(kappa * pR * sR);
    if (Ssl >= 0.0) { for (int v = 0; v < 5; v++) f[v] = FL[v]; return; }
    if (Ssr <= 0.0) { for (int v = 0; v < 5; v++) f[v] = FR[v]; return; }
    const double sMuL = Ssl - vL[0], sMuR = Ssr - vR[0];
    const double SStar = (pR - pL + uLL[1] * sMuL - uRR[1] * sMuR) / (uLL[0] * sMuL - uRR[0] * sMuR);
    const int left = (Ssl <= 0.0) && (SStar >= 0.0);
    const double *u = left ? uLL : uRR, *Fk = left ? FL : FR, *vk = left ? vL : vR;
    const double sk = left ? sL : sR, sMu = left ? sMuL : sMuR, Ss = left ? Ssl : Ssr, pk = left ? pL : pR;
    const double c = u[0] * sMu / (Ss - SStar);
    const double Us[5] = {c, c * SStar, c * u[2] * sk, c * u[3] * sk,
                          c * (u[4] * sk + (SStar - vk[0]) * (SStar + pk * sk / sMu))};
    for (int v = 0; v < 5; v++) f[v] = Fk[v] + Ss * (Us[v] - u[v]);
    return;
  }
  const double HL = (uLL[4] + pL) * sL, HR = (uRR[4] + pR) * sR;
  const double qL = sqrt(uLL[0]), qR = sqrt(uRR[0]);
  if (which == 22) {  /* HLL */
    const double Ht = (qL * HL + qR * HR) / (qL + qR), vt = (qL * vL[0] + qR * vR[0]) / (qL + qR);
    const double at = sqrt((kappa - 1.0) * (Ht - 0.5 * vt * vt));
    const double Ssl = vt - at, Ssr = vt + at;
    if (Ssl >= 0.0 && Ssr > 0.0) { for (int v = 0; v < 5; v++) f[v] = FL[v]; return; }
    if (Ssr <= 0.0 && Ssl < 0.0) { for (int v = 0; v < 5; v++) f[v] = FR[v]; return; }
    for (int v = 0; v < 5; v++) f[v] = (Ssr * FL[v] - Ssl * FR[v] + Ssl * Ssr * (uRR[v] - uLL[v])) / (Ssr - Ssl);
    return;
  }
  const double sq = 1.0 / (qL + qR);
  double rv[3];
  for (int d = 0; d < 3; d++) rv[d] = (qR * vR[d] + qL * vL[d]) * sq;
  const double rH = (qR * HR + qL * HL) * sq, av = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
  const double rc = sqrt(km1 * (rH - 0.5 * av));
  if (which == 23 || which == 24) {  /* HLLE, HLLEM */
    const double beta = sqrt(0.5 * km1 / kappa);
    const double aL = sqrt(kappa * pL * sL), aR = sqrt(kappa * pR * sR);
    double Ssl = rv[0] - rc, Ssr = rv[0] + rc;
    if (vL[0] - beta * aL < Ssl) Ssl = vL[0] - beta * aL;
    if (Ssl > 0.0) Ssl = 0.0;
    if (vR[0] + beta * aR > Ssr) Ssr = vR[0] + beta * aR;
    if (Ssr < 0.0) Ssr = 0.0;
    if (Ssl >= 0.0 && Ssr > 0.0) { for (int v = 0; v < 5; v++) f[v] = FL[v]; return; }
    if (Ssr <= 0.0 && Ssl < 0.0) { for (int v = 0; v < 5; v++) f[v] = FR[v]; return; }
    double anti[5] = {0, 0, 0, 0, 0};
    if (which == 24) {
      const double delta = rc / (rc + fabs(0.5 * (Ssl + Ssr)));
      const double rd = sqrt(uLL[0] * uRR[0]);
      const double A2 = (uRR[0] - uLL[0]) - (pR - pL) / (rc * rc);
      const double A3 = rd * (vR[1] - vL[1]), A4 = rd * (vR[2] - vL[2]);
      const double r2[5] = {1.0, rv[0], rv[1], rv[2], 0.5 * av};
      const double r3[5] = {0.0, 0.0, 1.0, 0.0, rv[1]}, r4[5] = {0.0, 0.0, 0.0, 1.0, rv[2]};
      for (int v = 0; v < 5; v++) anti[v] = delta * (r2[v] * A2 + r3[v] * A3 + r4[v] * A4);
    }
    for (int v = 0; v < 5; v++)
      f[v] = (Ssr * FL[v] - Ssl * FR[v] + Ssl * Ssr * (uRR[v] - uLL[v] - anti[v])) / (Ssr - Ssl);
    return;
  }
  /* Roe */
  double a1 = rv[0] - rc, a5 = rv[0] + rc;
  const double a2 = rv[0];
  const double r1[5] = {1.0, a1, rv[1], rv[2], rH - rv[0] * rc};
  const double r2[5] = {1.0, rv[0], rv[1], rv[2], 0.5 * av};
  const double r3[5] = {0.0, 0.0, 1.0, 0.0, rv[1]}, r4[5] = {0.0, 0.0, 0.0, 1.0, rv[2]};
  const double r5[5] = {1.0, a5, rv[1], rv[2], rH + rv[0] * rc};
  double dU[6];
  for (int v = 0; v < 5; v++) dU[v] = uRR[v] - uLL[v];
  dU[5] = dU[4] - (dU[2] - rv[1] * dU[0]) * rv[1] - (dU[3] - rv[2] * dU[0]) * rv[2];
  const double A3 = dU[2] - rv[1] * dU[0], A4 = dU[3] - rv[2] * dU[0];
  const double A2 = km1 / (rc * rc) * (dU[0] * (rH - rv[0] * rv[0]) - dU[5] + rv[0] * dU[1]);
  const double A1 = 0.5 / rc * (dU[0] * (rv[0] + rc) - dU[1] - rc * A2);
  const double A5 = dU[0] - A1 - A2;
  double Fc[5];
  if (volflux == 2 || volflux == 3 || volflux == 4) {
    const double e1[3] = {1.0, 0.0, 0.0};
    double xL[6], xR[6];
    ora_eval_uaux(kappa, uLL, xL);
    ora_eval_uaux(kappa, uRR, xR);
    ora_volflux(volflux == 4 ? 8 : volflux, kappa, uLL, uRR, xL, xR, e1, e1, Fc);
    if (volflux != 3) {
      const double m = fabs(rv[0] - rc) > fabs(rv[0] + rc) ? fabs(rv[0] - rc) : fabs(rv[0] + rc);
      a1 = m;
      a5 = m;
    }
  } else {
    for (int v = 0; v < 5; v++) Fc[v] = 0.5 * (FL[v] + FR[v]);
  }
  for (int v = 0; v < 5; v++)
    f[v] = Fc[v] - 0.5 * (A1 * fabs(a1) * r1[v] + A2 * fabs(a2) * r2[v] + A3 * fabs(a2) * r3[v] + A4 * fabs(a2) * r4[v] +
                          A5 * fabs(a5) * r5[v]);
}

/* the largest wave speed of one face point in the rotated frame (the LambdaMaxMax term) */
double ora_face_lambda(double kappa, const double *UL, const double *UR, const double *nv)
{
  const double mL = UL[1] * nv[0] + UL[2] * nv[1] + UL[3] * nv[2];
  const double mR = UR[1] * nv[0] + UR[2] * nv[1] + UR[3] * nv[2];
  const double a = fabs(mL / UL[0]), b = fabs(mR / UR[0]);
  const double c2L = sound_speed2(kappa, UL), c2R = sound_speed2(kappa, UR);
  return (a > b ? a : b) + sqrt(c2L > c2R ? c2L : c2R);
}

/* Riemann -> AdvRiemann: rotate, solve, rotate back (riemann.f90:174-246); solvers:
   0 central (:251-272), 1 Rusanov/LLF (:279-313), 9 EC-KEP and 32 Ranocha through
   RiemannSolver_VolumeFluxAverage (:980-1005), 4 / 44 / 33 entropy stable with the
   Ismail-Roe / Chandrashekar / Ranocha flux (:874-975), 16 EC-KEP + LLF (:1157-1209) and
   13 Ismail-Roe + LLF (:1011-1112) with the face's LambdaMaxMax in lam_face, and the
   volume-flux averages 5-8, 10-12, 20 and their LLF forms 14, 15, 17-19, 21.  Point-wise
   (one face node). */
void ora_riemann_lam(int which, double kappa, const double *UL, const double *UR, const double *nv,
                     const double *t1, const double *t2, double lam_face, double *F)
{
  ora_riemann_lam_vf(which, 0, kappa, UL, UR, nv, t1, t2, lam_face, F);
}

/* the same with the compile-time VolumeFlux, which selects the central part of Roe (3) */
void ora_riemann_lam_vf(int which, int volflux, double kappa, const double *UL, const double *UR, const double *nv,
                        const double *t1, const double *t2, double lam_face, double *F)
{
  double uLL[5], uRR[5];
  uLL[0] = UL[0]; uRR[0] = UR[0];
  uLL[1] = UL[1] * nv[0] + UL[2] * nv[1] + UL[3] * nv[2];
  uLL[2] = UL[1] * t1[0] + UL[2] * t1[1] + UL[3] * t1[2];
  uLL[3] = UL[1] * t2[0] + UL[2] * t2[1] + UL[3] * t2[2];
  uRR[1] = UR[1] * nv[0] + UR[2] * nv[1] + UR[3] * nv[2];
  uRR[2] = UR[1] * t1[0] + UR[2] * t1[1] + UR[3] * t1[2];
  uRR[3] = UR[1] * t2[0] + UR[2] * t2[1] + UR[3] * t2[2];
  uLL[4] = UL[4]; uRR[4] = UR[4];
  double f[5];
  if (which == 0 || which == 1) {
    double FL[5], FR[5];
    euler_flux_1d(kappa, uLL, FL);
    euler_flux_1d(kappa, uRR, FR);
    if (which == 0) {
      for (int v = 0; v < 5; v++) f[v] = 0.5 * (FL[v] + FR[v]);
    } else {
      double a = fabs(uLL[1] / uLL[0]), b = fabs(uRR[1] / uRR[0]);
      double c2L = sound_speed2(kappa, uLL), c2R = sound_speed2(kappa, uRR);
      double lam = (a > b ? a : b) + sqrt(c2L > c2R ? c2L : c2R);
      for (int v = 0; v < 5; v++) f[v] = 0.5 * ((FL[v] + FR[v]) - lam * (uRR[v] - uLL[v]));
    }
  } else if (which == 9) {
    flux1d_chandrashekar(kappa, uLL, uRR, f);
  } else if (which == 4 || which == 33 || which == 44) {
    riemann_es(which, kappa, uLL, uRR, f);
  } else if (which == 16) {
    riemann_eckep_llf(kappa, uLL, uRR, lam_face, f);
  } else if (which == 13) {
    riemann_ec_llf(kappa, uLL, uRR, lam_face, f);
  } else if (which == 2 || which == 3 || (which >= 22 && which <= 24)) {
    riemann_hll_roe(which, volflux, kappa, uLL, uRR, f);
  } else if (riemann_volflux_of(which) >= 0) {
    riemann_volflux_avg(which, kappa, uLL, uRR, lam_face, f);
  } else {
    flux1d_ranocha(kappa, uLL, uRR, f);
  }
  F[0] = f[0];
  F[1] = nv[0] * f[1] + t1[0] * f[2] + t2[0] * f[3];
  F[2] = nv[1] * f[1] + t1[1] * f[2] + t2[1] * f[3];
  F[3] = nv[2] * f[1] + t1[2] * f[2] + t2[2] * f[3];
  F[4] = f[4];
}

void ora_riemann(int which, double kappa, const double *UL, const double *UR, const double *nv,
                 const double *t1, const double *t2, double *F)
{
  ora_riemann_lam(which, kappa, UL, UR, nv, t1, t2, ora_face_lambda(kappa, UL, UR, nv), F);
}

/* ConsToEntropy (equation_vars.f90:362-391) */
void ora_cons_to_entropy(double kappa, const double *U, double *w)
{
  double km1 = kappa - 1.0;
  double srho = 1.0 / U[0];
  double u = U[1] * srho, v = U[2] * srho, ww = U[3] * srho;
  double v2s2 = 0.5 * (u * u + v * v + ww * ww);
  double rho_sp = U[0] / (km1 * (U[4] - U[0] * v2s2));
  double s = -log(rho_sp * pow(U[0], km1));
  w[0] = (kappa - s) / km1 - rho_sp * v2s2;
  w[1] = rho_sp * u;
  w[2] = rho_sp * v;
  w[3] = rho_sp * ww;
  w[4] = -rho_sp;
}

/* ConvertToGradPrimVec for entropy-variable lifting (equation_vars.f90:571-647) */
static void convert_to_grad_prim(double kappa, const double *U, double *g)
{
  double km1 = kappa - 1.0, skm1 = 1.0 / km1;
  double srho = 1.0 / U[0];
  double u = U[1] * srho, v = U[2] * srho, w = U[3] * srho;
  double ekin = 0.5 * (U[1] * u + U[2] * v + U[3] * w);
  double p = km1 * (U[4] - ekin);
  double rho_sp = U[0] / p;
  double p_srho = p * srho;
  double g1 = g[0];
  double gu = p_srho * (g[1] + u * g[4]);
  double gv = p_srho * (g[2] + v * g[4]);
  double gw = p_srho * (g[3] + w * g[4]);
  double g5 = g[4];
  g[0] = U[0] * g1 - g5 * (ekin - p * skm1) + rho_sp * (U[1] * gu + U[2] * gv + U[3] * gw);
  g[1] = gu; g[2] = gv; g[3] = gw;
  g[4] = p_srho * (g[0] + p * g5);
}

/* EvalDiffFlux3D for constant viscosity (flux.f90:407-517), one point */
static void diff_flux_3d(const ora_problem *P, const double *U, const double *gx, const double *gy,
                         const double *gz, double *f, double *g, double *h)
{
  double kappa = P->kappa, km1 = kappa - 1.0, skm1 = 1.0 / km1;
  double KappasPr = kappa / P->Pr, s23 = 2.0 / 3.0;
  double srho = 1.0 / U[0];
  double v1 = U[1] * srho, v2 = U[2] * srho, v3 = U[3] * srho;
  double p = km1 * (U[4] - 0.5 * (U[1] * v1 + U[2] * v2 + U[3] * v3));
  double muS = P->mu0, lambda = muS * KappasPr;
  double divv = gx[1] + gy[2] + gz[3];
  double cTx = skm1 * srho * (gx[4] - srho * p * gx[0]);
  double cTy = skm1 * srho * (gy[4] - srho * p * gy[0]);
  double cTz = skm1 * srho * (gz[4] - srho * p * gz[0]);
  f[0] = 0.0;
  f[1] = -muS * (2 * gx[1] - s23 * divv);
  f[2] = -muS * (gx[2] + gy[1]);
  f[3] = -muS * (gx[3] + gz[1]);
  f[4] = v1 * f[1] + v2 * f[2] + v3 * f[3] - lambda * cTx;
  g[0] = 0.0;
  g[1] = f[2];
  g[2] = -muS * (2 * gy[2] - s23 * divv);
  g[3] = -muS * (gy[3] + gz[2]);
  g[4] = v1 * g[1] + v2 * g[2] + v3 * g[3] - lambda * cTy;
  h[0] = 0.0;
  h[1] = f[3];
  h[2] = g[3];
  h[3] = -muS * (2 * gz[3] - s23 * divv);
  h[4] = v1 * h[1] + v2 * h[2] + v3 * h[3] - lambda * cTz;
}

/* ---------------------------------------------------------------------------- */
typedef struct {
  int N, n, n3, E, S;
  double xGP[16], wGP[16], D[256], Dhat[256], Ds[256], Lhm[16], Lhp[16];
} ops_t;

static void make_ops(const ora_problem *P, ops_t *o)
{
  o->N = P->N; o->n = P->N + 1; o->n3 = o->n * o->n * o->n; o->E = P->nElems; o->S = P->nSides;
  ora_dg_operators(P->N, o->xGP, o->wGP, o->D, o->Dhat, o->Ds, o->Lhm, o->Lhp);
}

#define VOL(e, i, j, k) ((((size_t)(e) * n + (k)) * n + (j)) * n + (i))
#define FACE(s, p, q) (((size_t)(s) * n + (q)) * n + (p))

/* VolInt_SplitForm + EvalAdvFluxAverage3D (volint.f90:173-218, flux_average.f90:198-269).
   Dvolsurf_T(l,i) = Dvolsurf(i,l). */
void ora_volint_splitform(const ora_problem *P, const double *U, double *Ut)
{
  ops_t o; make_ops(P, &o);
  int n = o.n;
  double *aux = malloc(sizeof(double) * 6 * o.n3);
  double *ft = malloc(sizeof(double) * NV * n * o.n3), *gt = malloc(sizeof(double) * NV * n * o.n3),
         *ht = malloc(sizeof(double) * NV * n * o.n3);
  const double *Mm[3] = {P->Mf, P->Mg, P->Mh};
#define T4(l, i, j, k) ((((size_t)(k) * n + (j)) * n + (i)) * n + (l))
  for (int e = 0; e < o.E; e++) {
    const double *Ue = U + (size_t)NV * o.n3 * e;
    for (int p = 0; p < o.n3; p++) ora_eval_uaux(P->kappa, Ue + NV * p, aux + 6 * p);
    memset(ft, 0, sizeof(double) * NV * n * o.n3);
    memset(gt, 0, sizeof(double) * NV * n * o.n3);
    memset(ht, 0, sizeof(double) * NV * n * o.n3);
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          int a = i + n * (j + n * k);
          for (int l = i + 1; l < n; l++) {
            int b = l + n * (j + n * k);
            double F[5];
            ora_volflux(P->volflux, P->kappa, Ue + NV * a, Ue + NV * b, aux + 6 * a, aux + 6 * b,
                        Mm[0] + 3 * VOL(e, i, j, k), Mm[0] + 3 * VOL(e, l, j, k), F);
            for (int v = 0; v < NV; v++) { ft[NV * T4(l, i, j, k) + v] = F[v]; ft[NV * T4(i, l, j, k) + v] = F[v]; }
          }
          for (int l = j + 1; l < n; l++) {
            int b = i + n * (l + n * k);
            double F[5];
            ora_volflux(P->volflux, P->kappa, Ue + NV * a, Ue + NV * b, aux + 6 * a, aux + 6 * b,
                        Mm[1] + 3 * VOL(e, i, j, k), Mm[1] + 3 * VOL(e, i, l, k), F);
            for (int v = 0; v < NV; v++) { gt[NV * T4(l, i, j, k) + v] = F[v]; gt[NV * T4(j, i, l, k) + v] = F[v]; }
          }
          for (int l = k + 1; l < n; l++) {
            int b = i + n * (j + n * l);
            double F[5];
            ora_volflux(P->volflux, P->kappa, Ue + NV * a, Ue + NV * b, aux + 6 * a, aux + 6 * b,
                        Mm[2] + 3 * VOL(e, i, j, k), Mm[2] + 3 * VOL(e, i, j, l), F);
            for (int v = 0; v < NV; v++) { ht[NV * T4(l, i, j, k) + v] = F[v]; ht[NV * T4(k, i, j, l) + v] = F[v]; }
          }
        }
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double *ut = Ut + NV * VOL(e, i, j, k);
          for (int l = 0; l < n; l++)
            for (int v = 0; v < NV; v++)
              ut[v] = ut[v] + o.Ds[i + n * l] * ft[NV * T4(l, i, j, k) + v]
                            + o.Ds[j + n * l] * gt[NV * T4(l, i, j, k) + v]
                            + o.Ds[k + n * l] * ht[NV * T4(l, i, j, k) + v];
        }
  }
#undef T4
  free(aux); free(ft); free(gt); free(ht);
}

/* ProlongToFace for LGL: face node = volume node S2V(:,0,p,q,flip,locSide) (prolongtoface.f90:97-102) */
static void prolong_to_face(const ora_problem *P, const ops_t *o, int nvar, const double *Uv,
                            double *Um, double *Us)
{
  int n = o->n;
  for (int s = 0; s < o->S; s++) {
    const int *se = P->SideToElem + 5 * s;
    int em = se[0] - 1, es = se[1] - 1, lm = se[2], ls = se[3], fl = se[4];
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        int ijk[3];
        ora_side_to_vol(o->N, 0, p, q, 0, lm, ijk);
        memcpy(Um + nvar * FACE(s, p, q), Uv + nvar * VOL(em, ijk[0], ijk[1], ijk[2]), sizeof(double) * nvar);
        ora_side_to_vol(o->N, 0, p, q, fl, ls, ijk);
        memcpy(Us + nvar * FACE(s, p, q), Uv + nvar * VOL(es, ijk[0], ijk[1], ijk[2]), sizeof(double) * nvar);
      }
  }
}

/* SurfInt for LGL (surfint.f90:177-184, 253-259): master +F*L_hatMinus0, slave -F*L_hatMinus0 */
static void surf_int(const ora_problem *P, const ops_t *o, int nvar, const double *Fm, const double *Fs,
                     double sign_slave, double *Ut)
{
  int n = o->n;
  double l0 = o->Lhm[0];
  for (int s = 0; s < o->S; s++) {
    const int *se = P->SideToElem + 5 * s;
    int em = se[0] - 1, es = se[1] - 1, lm = se[2], ls = se[3], fl = se[4];
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        int ijk[3];
        ora_side_to_vol(o->N, 0, p, q, 0, lm, ijk);
        double *u = Ut + nvar * VOL(em, ijk[0], ijk[1], ijk[2]);
        for (int v = 0; v < nvar; v++) u[v] = u[v] + Fm[nvar * FACE(s, p, q) + v] * l0;
        ora_side_to_vol(o->N, 0, p, q, fl, ls, ijk);
        u = Ut + nvar * VOL(es, ijk[0], ijk[1], ijk[2]);
        for (int v = 0; v < nvar; v++) u[v] = u[v] + sign_slave * Fs[nvar * FACE(s, p, q) + v] * l0;
      }
  }
}

/* BR1 lifting (br1.f90:154-364, br1_volint.f90:45-98, br1_fillflux.f90:49-91,
   br1_surfint.f90:52-131) of nvar lifting variables w given at the volume nodes: returns the
   physical gradients of w (Lifting_VolInt + strong surface flux 1/2(w_s-w_m) SurfElem nv on
   both sides + V2D_M_V1D(sJ)); the equation converts them to primitive gradients. */
void ora_lift(const ora_problem *P, int nvar, const double *w, double *gx, double *gy, double *gz)
{
  ops_t o; make_ops(P, &o);
  int n = o.n, N = o.N;
  size_t nvol = (size_t)nvar * o.n3 * o.E, nface = (size_t)nvar * n * n * o.S;
  double *wm = malloc(sizeof(double) * nface), *ws = malloc(sizeof(double) * nface);
  double *fx = malloc(sizeof(double) * nface), *fy = malloc(sizeof(double) * nface),
         *fz = malloc(sizeof(double) * nface);
  memset(gx, 0, sizeof(double) * nvol);
  memset(gy, 0, sizeof(double) * nvol);
  memset(gz, 0, sizeof(double) * nvol);
  /* Lifting_VolInt: strong-form D contraction with the node's metrics */
  for (int e = 0; e < o.E; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double gxi[16], geta[16], gzeta[16];
          for (int v = 0; v < nvar; v++) {
            gxi[v] = o.D[i + n * 0] * w[nvar * VOL(e, 0, j, k) + v];
            geta[v] = o.D[j + n * 0] * w[nvar * VOL(e, i, 0, k) + v];
            gzeta[v] = o.D[k + n * 0] * w[nvar * VOL(e, i, j, 0) + v];
          }
          for (int l = 1; l <= N; l++)
            for (int v = 0; v < nvar; v++) {
              gxi[v] = gxi[v] + o.D[i + n * l] * w[nvar * VOL(e, l, j, k) + v];
              geta[v] = geta[v] + o.D[j + n * l] * w[nvar * VOL(e, i, l, k) + v];
              gzeta[v] = gzeta[v] + o.D[k + n * l] * w[nvar * VOL(e, i, j, l) + v];
            }
          size_t c = VOL(e, i, j, k);
          const double *mf = P->Mf + 3 * c, *mg = P->Mg + 3 * c, *mh = P->Mh + 3 * c;
          for (int v = 0; v < nvar; v++) {
            gx[nvar * c + v] = gx[nvar * c + v] + (mf[0] * gxi[v] + mg[0] * geta[v] + mh[0] * gzeta[v]);
            gy[nvar * c + v] = gy[nvar * c + v] + (mf[1] * gxi[v] + mg[1] * geta[v] + mh[1] * gzeta[v]);
            gz[nvar * c + v] = gz[nvar * c + v] + (mf[2] * gxi[v] + mg[2] * geta[v] + mh[2] * gzeta[v]);
          }
        }
  /* Lifting_FillFlux: strong flux 1/2(w_s-w_m)*SurfElem*nv (flux.f90:677-716) */
  prolong_to_face(P, &o, nvar, w, wm, ws);
  for (int s = 0; s < o.S; s++)
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        size_t f = FACE(s, p, q);
        const double *nv = P->NormVec + 3 * f;
        double se = P->SurfElem[f];
        for (int v = 0; v < nvar; v++) {
          double fl = 0.5 * (ws[nvar * f + v] - wm[nvar * f + v]) * se;
          fx[nvar * f + v] = fl * nv[0];
          fy[nvar * f + v] = fl * nv[1];
          fz[nvar * f + v] = fl * nv[2];
        }
      }
  /* Lifting_SurfInt: both sides add +Flux*L_hatMinus0 */
  surf_int(P, &o, nvar, fx, fx, 1.0, gx);
  surf_int(P, &o, nvar, fy, fy, 1.0, gy);
  surf_int(P, &o, nvar, fz, fz, 1.0, gz);
  for (size_t p = 0; p < (size_t)o.n3 * o.E; p++) {
    double sj = P->sJ[p];
    for (int v = 0; v < nvar; v++) {
      gx[nvar * p + v] = gx[nvar * p + v] * sj;
      gy[nvar * p + v] = gy[nvar * p + v] * sj;
      gz[nvar * p + v] = gz[nvar * p + v] * sj;
    }
  }
  free(wm); free(ws); free(fx); free(fy); free(fz);
}

/* BR2 lifting (br2.f90:167-336, br2_volint.f90:45-98, br2_fillflux.f90:49-91,
   br2_surfint.f90:52-131) of nvar lifting variables w: volume gradients g* (sJ applied inside
   the volume integral and to every surface term) and, per side, the face gradients of the
   master (fm*) and slave (fs*) element: the prolonged volume-only gradient plus etaBR2 times
   that face's own surface term.  Face arrays (nvar,0:N,0:N,nSides). */
void ora_lift_br2(const ora_problem *P, int nvar, const double *w, double eta, double *gx, double *gy,
                  double *gz, double *fxm, double *fym, double *fzm, double *fxs, double *fys, double *fzs)
{
  ops_t o; make_ops(P, &o);
  int n = o.n, N = o.N;
  size_t nvol = (size_t)nvar * o.n3 * o.E, nface = (size_t)nvar * n * n * o.S;
  double *wm = malloc(sizeof(double) * nface), *ws = malloc(sizeof(double) * nface);
  double *fl[3] = {malloc(sizeof(double) * nface), malloc(sizeof(double) * nface), malloc(sizeof(double) * nface)};
  double *g[3] = {gx, gy, gz}, *fmst[3] = {fxm, fym, fzm}, *fslv[3] = {fxs, fys, fzs};
  for (int d = 0; d < 3; d++) memset(g[d], 0, sizeof(double) * nvol);
  /* Lifting_FillFlux: strong flux 1/2(w_s-w_m)*SurfElem*nv */
  prolong_to_face(P, &o, nvar, w, wm, ws);
  for (int s = 0; s < o.S; s++)
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        size_t f = FACE(s, p, q);
        const double *nv = P->NormVec + 3 * f;
        double se = P->SurfElem[f];
        for (int v = 0; v < nvar; v++) {
          double flx = 0.5 * (ws[nvar * f + v] - wm[nvar * f + v]) * se;
          for (int d = 0; d < 3; d++) fl[d][nvar * f + v] = flx * nv[d];
        }
      }
  /* Lifting_VolInt with sJ inside */
  for (int e = 0; e < o.E; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double gxi[16], geta[16], gzeta[16];
          for (int v = 0; v < nvar; v++) {
            gxi[v] = o.D[i + n * 0] * w[nvar * VOL(e, 0, j, k) + v];
            geta[v] = o.D[j + n * 0] * w[nvar * VOL(e, i, 0, k) + v];
            gzeta[v] = o.D[k + n * 0] * w[nvar * VOL(e, i, j, 0) + v];
          }
          for (int l = 1; l <= N; l++)
            for (int v = 0; v < nvar; v++) {
              gxi[v] = gxi[v] + o.D[i + n * l] * w[nvar * VOL(e, l, j, k) + v];
              geta[v] = geta[v] + o.D[j + n * l] * w[nvar * VOL(e, i, l, k) + v];
              gzeta[v] = gzeta[v] + o.D[k + n * l] * w[nvar * VOL(e, i, j, l) + v];
            }
          size_t c = VOL(e, i, j, k);
          const double *mf = P->Mf + 3 * c, *mg = P->Mg + 3 * c, *mh = P->Mh + 3 * c;
          const double sj = P->sJ[c];
          for (int d = 0; d < 3; d++)
            for (int v = 0; v < nvar; v++)
              g[d][nvar * c + v] = g[d][nvar * c + v] + (mf[d] * gxi[v] + mg[d] * geta[v] + mh[d] * gzeta[v]) * sj;
        }
  /* ProlongToFace of the volume gradients, then BR2_SurfInt per direction */
  for (int d = 0; d < 3; d++) {
    prolong_to_face(P, &o, nvar, g[d], fmst[d], fslv[d]);
    const double l0 = o.Lhm[0];
    for (int s = 0; s < o.S; s++) {
      const int *se = P->SideToElem + 5 * s;
      int em = se[0] - 1, es = se[1] - 1, lm = se[2], ls = se[3], flp = se[4];
      for (int q = 0; q < n; q++)
        for (int p = 0; p < n; p++) {
          int ijk[3];
          size_t f = FACE(s, p, q);
          ora_side_to_vol(N, 0, p, q, 0, lm, ijk);
          size_t cm = VOL(em, ijk[0], ijk[1], ijk[2]);
          for (int v = 0; v < nvar; v++) {
            double floc = P->sJ[cm] * l0 * fl[d][nvar * f + v];
            g[d][nvar * cm + v] = g[d][nvar * cm + v] + floc;
            fmst[d][nvar * f + v] = fmst[d][nvar * f + v] + eta * floc;
          }
        }
      for (int q = 0; q < n; q++)
        for (int p = 0; p < n; p++) {
          int ijk[3];
          size_t f = FACE(s, p, q);
          ora_side_to_vol(N, 0, p, q, flp, ls, ijk);
          size_t cs = VOL(es, ijk[0], ijk[1], ijk[2]);
          for (int v = 0; v < nvar; v++) {
            double floc = P->sJ[cs] * l0 * fl[d][nvar * f + v];
            g[d][nvar * cs + v] = g[d][nvar * cs + v] + floc;
            fslv[d][nvar * f + v] = fslv[d][nvar * f + v] + eta * floc;
          }
        }
    }
  }
  free(wm); free(ws); free(fl[0]); free(fl[1]); free(fl[2]);
}

/* BR1 lifting of the Navier-Stokes entropy variables; output: primitive gradients at the
   volume nodes. */
void ora_lifting(const ora_problem *P, const double *U, double *gx, double *gy, double *gz)
{
  size_t n3E = (size_t)(P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  double *w = calloc((size_t)NV * n3E, sizeof(double));
  for (size_t p = 0; p < n3E; p++) ora_cons_to_entropy(P->kappa, U + NV * p, w + NV * p);
  if (P->lifting == 2) {  /* BR2: volume gradients; the face gradients are not returned */
    size_t nf = (size_t)NV * (P->N + 1) * (P->N + 1) * P->nSides;
    double *fa = malloc(sizeof(double) * 6 * nf);
    ora_lift_br2(P, NV, w, P->etaBR2, gx, gy, gz, fa, fa + nf, fa + 2 * nf, fa + 3 * nf, fa + 4 * nf, fa + 5 * nf);
    free(fa);
  } else {
    ora_lift(P, NV, w, gx, gy, gz);
  }
  for (size_t p = 0; p < n3E; p++) {
    convert_to_grad_prim(P->kappa, U + NV * p, gx + NV * p);
    convert_to_grad_prim(P->kappa, U + NV * p, gy + NV * p);
    convert_to_grad_prim(P->kappa, U + NV * p, gz + NV * p);
  }
  free(w);
}

/* FillFlux + SurfInt (fillflux.f90:41-110, riemann.f90:104-169, surfint.f90) with the face
   gradients (primitive, (5,0:N,0:N,nSides) master/slave) given, or NULL for Euler */
/* FillFlux (src/dg/fillflux.f90:41-110): the numerical flux of every side (Riemann,
   riemann.f90:104-169, with the BR1/BR2 viscous mean from the sides' face gradients fg when
   parabolic) times SurfElem -> Fm (Flux_master; Flux_slave is the same for these equations) */
static void fill_flux_faces(const ora_problem *P, const double *U, double *const fg[6], double *Fm)
{
  ops_t o; make_ops(P, &o);
  int n = o.n;
  size_t nface = (size_t)NV * n * n * o.S;
  double *Um = malloc(sizeof(double) * nface), *Us = malloc(sizeof(double) * nface);
  prolong_to_face(P, &o, NV, U, Um, Us);
  for (int s = 0; s < o.S; s++) {
    double lam = 0.0;  /* LambdaMaxMax of the face (Riemann 16) */
    for (int f = s * n * n; f < (s + 1) * n * n; f++) {
      const double l = ora_face_lambda(P->kappa, Um + NV * f, Us + NV * f, P->NormVec + 3 * f);
      lam = lam > l ? lam : l;
    }
    for (int q = 0; q < n; q++)
      for (int p = 0; p < n; p++) {
        size_t f = FACE(s, p, q);
        double *F = Fm + NV * f;
        ora_riemann_lam_vf(P->riemann, P->volflux, P->kappa, Um + NV * f, Us + NV * f, P->NormVec + 3 * f,
                           P->TangVec1 + 3 * f, P->TangVec2 + 3 * f, lam, F);
        if (P->parabolic) {
          double kL[5], gL[5], jL[5], kR[5], gR[5], jR[5];
          const double *nv = P->NormVec + 3 * f;
          diff_flux_3d(P, Um + NV * f, fg[0] + NV * f, fg[2] + NV * f, fg[4] + NV * f, kL, gL, jL);
          diff_flux_3d(P, Us + NV * f, fg[1] + NV * f, fg[3] + NV * f, fg[5] + NV * f, kR, gR, jR);
          for (int v = 1; v < NV; v++)
            F[v] = F[v] + 0.5 * (nv[0] * (kL[v] + kR[v]) + nv[1] * (gL[v] + gR[v]) + nv[2] * (jL[v] + jR[v]));
        }
        double se = P->SurfElem[f];
        for (int v = 0; v < NV; v++) F[v] = F[v] * se;
      }
  }
  free(Um); free(Us);
}

/* FillFlux + SurfInt with the sides' face gradients fg */
static void surface_integral_faces(const ora_problem *P, const double *U, double *const fg[6], double *Ut)
{
  ops_t o; make_ops(P, &o);
  double *Fm = malloc(sizeof(double) * (size_t)NV * o.n * o.n * o.S);
  fill_flux_faces(P, U, fg, Fm);
  surf_int(P, &o, NV, Fm, Fm, -1.0, Ut);
  free(Fm);
}

void ora_surf_int(const ora_problem *P, int nvar, const double *Fm, const double *Fs, double *Ut)
{
  ops_t o; make_ops(P, &o);
  surf_int(P, &o, nvar, Fm, Fs, -1.0, Ut);
}

/* FillFlux + SurfInt with BR1 face gradients: the volume gradients prolonged to the faces
   (br1.f90:154-364) */
void ora_surface_integral(const ora_problem *P, const double *U, const double *gx, const double *gy,
                          const double *gz, double *Ut)
{
  ops_t o; make_ops(P, &o);
  size_t nface = (size_t)NV * o.n * o.n * o.S;
  double *fg[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  if (P->parabolic) {
    const double *g[3] = {gx, gy, gz};
    for (int d = 0; d < 3; d++) {
      fg[2 * d] = malloc(sizeof(double) * nface);
      fg[2 * d + 1] = malloc(sizeof(double) * nface);
      prolong_to_face(P, &o, NV, g[d], fg[2 * d], fg[2 * d + 1]);
    }
  }
  surface_integral_faces(P, U, fg, Ut);
  for (int d = 0; d < 6; d++) free(fg[d]);
}

/* BR2 lifting of the Navier-Stokes entropy variables with the conversion to primitive
   gradients of the volume gradients (with U) and the face gradients (with the face states,
   the volume nodes for LGL) (br2.f90:264-300) */
static void br2_gradients(const ora_problem *P, const double *U, double *gx, double *gy, double *gz,
                          double *const fg[6])
{
  ops_t o; make_ops(P, &o);
  int n = o.n;
  size_t n3E = (size_t)o.n3 * o.E, nface = (size_t)NV * n * n * o.S;
  double *w = calloc((size_t)NV * n3E, sizeof(double));
  for (size_t p = 0; p < n3E; p++) ora_cons_to_entropy(P->kappa, U + NV * p, w + NV * p);
  ora_lift_br2(P, NV, w, P->etaBR2, gx, gy, gz, fg[0], fg[2], fg[4], fg[1], fg[3], fg[5]);
  for (size_t p = 0; p < n3E; p++) {
    convert_to_grad_prim(P->kappa, U + NV * p, gx + NV * p);
    convert_to_grad_prim(P->kappa, U + NV * p, gy + NV * p);
    convert_to_grad_prim(P->kappa, U + NV * p, gz + NV * p);
  }
  double *Um = malloc(sizeof(double) * nface), *Us = malloc(sizeof(double) * nface);
  prolong_to_face(P, &o, NV, U, Um, Us);
  for (size_t f = 0; f < (size_t)n * n * o.S; f++)
    for (int d = 0; d < 3; d++) {
      convert_to_grad_prim(P->kappa, Um + NV * f, fg[2 * d] + NV * f);
      convert_to_grad_prim(P->kappa, Us + NV * f, fg[2 * d + 1] + NV * f);
    }
  free(Um); free(Us); free(w);
}

/* VolInt_visc + EvalDiffFluxTilde3D (volint.f90:228-271, flux.f90:523-635) */
static void volint_visc(const ora_problem *P, const double *U, const double *gx, const double *gy,
                        const double *gz, double *Ut)
{
  ops_t o; make_ops(P, &o);
  int n = o.n;
  double *ftd = malloc(sizeof(double) * NV * o.n3), *gtd = malloc(sizeof(double) * NV * o.n3),
         *htd = malloc(sizeof(double) * NV * o.n3);
  for (int e = 0; e < o.E; e++) {
    for (int p = 0; p < o.n3; p++) {
      size_t c = (size_t)o.n3 * e + p;
      double f[5], g[5], h[5];
      diff_flux_3d(P, U + NV * c, gx + NV * c, gy + NV * c, gz + NV * c, f, g, h);
      const double *mf = P->Mf + 3 * c, *mg = P->Mg + 3 * c, *mh = P->Mh + 3 * c;
      for (int v = 0; v < NV; v++) {
        ftd[NV * p + v] = f[v] * mf[0] + g[v] * mf[1] + h[v] * mf[2];
        gtd[NV * p + v] = f[v] * mg[0] + g[v] * mg[1] + h[v] * mg[2];
        htd[NV * p + v] = f[v] * mh[0] + g[v] * mh[1] + h[v] * mh[2];
      }
    }
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double *ut = Ut + NV * VOL(e, i, j, k);
          for (int l = 0; l < n; l++)
            for (int v = 0; v < NV; v++)
              ut[v] = ut[v] + o.Dhat[i + n * l] * ftd[NV * (l + n * (j + n * k)) + v]
                            + o.Dhat[j + n * l] * gtd[NV * (i + n * (l + n * k)) + v]
                            + o.Dhat[k + n * l] * htd[NV * (i + n * (j + n * l)) + v];
        }
  }
  free(ftd); free(gtd); free(htd);
}

/* DGTimeDerivative_weakForm (dg.f90:282-452) for periodic conforming LGL meshes */
void ora_dg_time_derivative(const ora_problem *P, const double *U, double t, double *Ut)
{
  (void)t;
  size_t n3E = (size_t)(P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  memset(Ut, 0, sizeof(double) * NV * n3E);
  double *alpha = NULL;
  if (P->nfvse) {  /* CalcBlendingCoefficient before the volume integral (dg.f90:358-361) */
    alpha = P->alpha ? P->alpha : malloc(sizeof(double) * P->nElems);
    ora_blending_coefficient(P, U, alpha);
  }
  ora_volint_splitform(P, U, Ut);
  if (P->nfvse) {  /* VolInt_NFVSE blends the volume integral only (dg.f90:381-383) */
    ora_volint_nfvse(P, U, alpha, Ut);
    if (alpha != P->alpha) free(alpha);
  }
  double *gx = NULL, *gy = NULL, *gz = NULL;
  if (P->parabolic) {
    gx = malloc(sizeof(double) * NV * n3E); gy = malloc(sizeof(double) * NV * n3E);
    gz = malloc(sizeof(double) * NV * n3E);
  }
  if (P->parabolic && P->lifting == 2) {
    size_t nface = (size_t)NV * (P->N + 1) * (P->N + 1) * P->nSides;
    double *fg[6];
    for (int d = 0; d < 6; d++) fg[d] = malloc(sizeof(double) * nface);
    br2_gradients(P, U, gx, gy, gz, fg);
    volint_visc(P, U, gx, gy, gz, Ut);
    surface_integral_faces(P, U, fg, Ut);
    for (int d = 0; d < 6; d++) free(fg[d]);
  } else {
    if (P->parabolic) {
      ora_lifting(P, U, gx, gy, gz);
      volint_visc(P, U, gx, gy, gz, Ut);
    }
    ora_surface_integral(P, U, gx, gy, gz, Ut);
  }
  for (size_t p = 0; p < n3E; p++) {
    double s = -P->sJ[p];
    for (int v = 0; v < NV; v++) Ut[NV * p + v] = Ut[NV * p + v] * s;
  }
  if (P->parabolic) { free(gx); free(gy); free(gz); }
}

void ora_fill_flux(const ora_problem *P, const double *U, double *Fm, double *Fs)
{
  const size_t n3E = (size_t)(P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  const size_t nface = (size_t)NV * (P->N + 1) * (P->N + 1) * P->nSides;
  double *fg[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  if (P->parabolic) {
    double *gx = malloc(sizeof(double) * NV * n3E), *gy = malloc(sizeof(double) * NV * n3E);
    double *gz = malloc(sizeof(double) * NV * n3E);
    for (int d = 0; d < 6; d++) fg[d] = malloc(sizeof(double) * nface);
    if (P->lifting == 2) {
      br2_gradients(P, U, gx, gy, gz, fg);
    } else {  /* BR1: the volume gradients prolonged to the faces */
      ops_t o; make_ops(P, &o);
      ora_lifting(P, U, gx, gy, gz);
      const double *g[3] = {gx, gy, gz};
      for (int d = 0; d < 3; d++) prolong_to_face(P, &o, NV, g[d], fg[2 * d], fg[2 * d + 1]);
    }
    free(gx); free(gy); free(gz);
  }
  fill_flux_faces(P, U, fg, Fm);
  memcpy(Fs, Fm, sizeof(double) * nface);
  for (int d = 0; d < 6; d++) free(fg[d]);
}

/* Carpenter RK4-5 (timedisc_vars.f90:204-233) */
void ora_rk_coefs(double *A, double *b, double *c)
{
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

/* CFLScaleAlpha for LGL + Carpenter RK4-5 (timedisc_vars.f90:212-213) */
double ora_cfl_alpha(int N)
{
  static const double a[15] = {4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830,
                               1.5682, 1.4692, 1.3849, 1.3106, 1.2454, 1.1880, 1.1362};
  return a[(N < 15 ? N : 15) - 1];
}

/* DFLScaleAlpha(LGL) * RelativeDFL(RK4-5 = 1.853) (timedisc_vars.f90:79-84, 214-216) */
double ora_dfl_alpha(int N)
{
  static const double a[10] = {4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17};
  return a[(N < 10 ? N : 10) - 1] * 1.853;
}

/* TimeStepByLSERKW2 (timedisc.f90:424-480) with VCopy/VAXPBY (vector.f90:68-117) */
void ora_lserk_step(const ora_problem *P, double *U, double *Ut_tmp, double t, double dt)
{
  double A[5], b[5], c[5];
  ora_rk_coefs(A, b, c);
  size_t nt = (size_t)NV * (P->N + 1) * (P->N + 1) * (P->N + 1) * P->nElems;
  double *Ut = malloc(sizeof(double) * nt);
  for (int s = 0; s < 5; s++) {
    double bdt = b[s] * dt;
    ora_dg_time_derivative(P, U, t + (s ? dt * c[s] : 0.0), Ut);
    if (s == 0) {
      for (size_t i = 0; i < nt; i++) { Ut_tmp[i] = Ut[i]; U[i] = U[i] + Ut[i] * bdt; }
    } else {
      for (size_t i = 0; i < nt; i++) {
        Ut_tmp[i] = Ut_tmp[i] * (-A[s]) + Ut[i];
        U[i] = U[i] + Ut_tmp[i] * bdt;
      }
    }
  }
  free(Ut);
}

/* CALCTIMESTEP (calctimestep.f90:39-170) with FillCFL_DFL (timedisc.f90:697-741) */
double ora_calc_timestep(const ora_problem *P, const double *U, double CFLscale, double DFLscale,
                         int *errType)
{
  int N = P->N, n = N + 1, n3 = n * n * n;
  double kappa = P->kappa, km1 = kappa - 1.0;
  double cfl = CFLscale * ora_cfl_alpha(N) / (2.0 * N + 1.0);
  double dfl = DFLscale * ora_dfl_alpha(N) / ((2.0 * N + 1.0) * (2.0 * N + 1.0));
  double KsPr = kappa / P->Pr, KsPr_max = (4.0 / 3.0 > KsPr) ? 4.0 / 3.0 : KsPr;
  double dtc = 1e308, dtv = 1e308;
  *errType = 0;
  for (int e = 0; e < P->nElems; e++) {
    double l1 = 1e-12, l2 = 1e-12, l3 = 1e-12, lv1 = 1e-12, lv2 = 1e-12, lv3 = 1e-12;
    for (int p = 0; p < n3; p++) {
      size_t c = (size_t)n3 * e + p;
      const double *u = U + NV * c;
      if (isnan(u[0])) *errType = 1;
      double srho = 1.0 / u[0];
      double v[3] = {u[1] * srho, u[2] * srho, u[3] * srho};
      double pr = km1 * (u[4] - 0.5 * u[0] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      if (pr <= 0.0) *errType = 2;
      double cs = sqrt(kappa * pr * srho), sj = P->sJ[c];
      const double *mf = P->Mf + 3 * c, *mg = P->Mg + 3 * c, *mh = P->Mh + 3 * c;
      double a1 = sj * (fabs(mf[0] * v[0] + mf[1] * v[1] + mf[2] * v[2]) + cs * sqrt(mf[0] * mf[0] + mf[1] * mf[1] + mf[2] * mf[2]));
      double a2 = sj * (fabs(mg[0] * v[0] + mg[1] * v[1] + mg[2] * v[2]) + cs * sqrt(mg[0] * mg[0] + mg[1] * mg[1] + mg[2] * mg[2]));
      double a3 = sj * (fabs(mh[0] * v[0] + mh[1] * v[1] + mh[2] * v[2]) + cs * sqrt(mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2]));
      if (a1 > l1) l1 = a1;
      if (a2 > l2) l2 = a2;
      if (a3 > l3) l3 = a3;
      if (P->parabolic) {
        double muX = srho * KsPr_max * P->mu0;
        double b1 = muX * ((mf[0] * sj) * (mf[0] * sj) + (mf[1] * sj) * (mf[1] * sj) + (mf[2] * sj) * (mf[2] * sj));
        double b2 = muX * ((mg[0] * sj) * (mg[0] * sj) + (mg[1] * sj) * (mg[1] * sj) + (mg[2] * sj) * (mg[2] * sj));
        double b3 = muX * ((mh[0] * sj) * (mh[0] * sj) + (mh[1] * sj) * (mh[1] * sj) + (mh[2] * sj) * (mh[2] * sj));
        if (b1 > lv1) lv1 = b1;
        if (b2 > lv2) lv2 = b2;
        if (b3 > lv3) lv3 = b3;
      }
    }
    double dte = cfl * 2.0 / (l1 + l2 + l3);
    if (dte < dtc) dtc = dte;
    if (isnan(dtc)) *errType = 3;
    if (P->parabolic) {
      double dve = dfl * 4.0 / (lv1 + lv2 + lv3);
      if (dve < dtv) dtv = dve;
      if (isnan(dtv)) *errType = 4;
    }
    if (*errType) break;
  }
  return dtc < dtv ? dtc : dtv;
}

/* CalcEntropy (analyze_equation.f90:549-599) */
void ora_calc_entropy(const ora_problem *P, const double *U, const double *Ut, double *S, double *dSdU_Ut)
{
  int N = P->N, n = N + 1, n3 = n * n * n;
  double xg[16], wg[16];
  ora_lgl_nodes(N, xg, wg);
  double kappa = P->kappa, km1 = kappa - 1.0;
  double ent = 0.0, dsu = 0.0;
  for (int e = 0; e < P->nElems; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          size_t c = (size_t)n3 * e + i + n * (j + n * k);
          const double *u = U + NV * c;
          double srho = 1.0 / u[0];
          double v1 = u[1] * srho, v2 = u[2] * srho, v3 = u[3] * srho;
          double pr = km1 * (u[4] - 0.5 * (u[1] * v1 + u[2] * v2 + u[3] * v3));
          double wv = wg[i] * wg[j] * wg[k];
          double el = -u[0] * (log(pr) - kappa * log(u[0]));
          ent = ent + el * wv / P->sJ[c];
          double w[5];
          ora_cons_to_entropy(kappa, u, w);
          double s = 0.0;
          for (int v = 0; v < NV; v++) s += w[v] * Ut[NV * c + v];
          dsu = dsu + s * wv / P->sJ[c];
        }
  *S = ent / km1;
  *dSdU_Ut = dsu;
}

/* ExactFunc 2 (density wave) and 7 (Taylor-Green vortex) (equation.f90:632-651, 802-811) */
void ora_exactfunc(int which, double kappa, const double *advvel, double amp, double t,
                   int npts, const double *x, double *U)
{
  const double pi = acos(-1.0);
  double skm1 = 1.0 / (kappa - 1.0);
  for (int p = 0; p < npts; p++) {
    const double *xp = x + 3 * p;
    double *u = U + NV * p;
    if (which == 2) {
      double prim[5] = {1.0, advvel[0], advvel[1], advvel[2], 1.0};
      double cent[3] = {xp[0] - advvel[0] * t, xp[1] - advvel[1] * t, xp[2] - advvel[2] * t};
      prim[0] = prim[0] * (1.0 + amp * sin(pi * (1.0 * cent[0] + 1.0 * cent[1] + 1.0 * cent[2])));
      u[0] = prim[0];
      u[1] = prim[0] * prim[1]; u[2] = prim[0] * prim[2]; u[3] = prim[0] * prim[3];
      u[4] = prim[4] * skm1 + 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3]);
    } else {
      double A = 1.0, Ms = 0.1, prim[5];
      prim[0] = 1.0;
      prim[1] = A * sin(xp[0]) * cos(xp[1]) * cos(xp[2]);
      prim[2] = -A * cos(xp[0]) * sin(xp[1]) * cos(xp[2]);
      prim[3] = 0.0;
      prim[4] = (A / Ms * A / Ms / kappa * prim[0]);
      prim[4] = prim[4] + 1.0 / 16.0 * A * A * prim[0] *
                (cos(2 * xp[0]) * cos(2.0 * xp[2]) + 2.0 * cos(2.0 * xp[1]) + 2.0 * cos(2.0 * xp[0]) +
                 cos(2 * xp[1]) * cos(2.0 * xp[2]));
      /* PrimToCons (equation_vars.f90) */
      u[0] = prim[0];
      u[1] = prim[1] * prim[0]; u[2] = prim[2] * prim[0]; u[3] = prim[3] * prim[0];
      u[4] = skm1 * prim[4] + 0.5 * (u[1] * prim[1] + u[2] * prim[2] + u[3] * prim[3]);
    }
  }
}
