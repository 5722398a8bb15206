// GLM-MHD (9 variables: rho, rho v, E, B, psi) on gfx950: the Derigs et al. entropy- and
// kinetic-energy-conserving two-point flux, Powell non-conservative terms, the Rusanov
// surface flux, CALCTIMESTEP with the GLM cleaning speed from the time step, and the LSERK
// stage update (src/equation/mhd/{flux_average,flux,riemann,equation_vars,calctimestep}.f90,
// NONCONS branches of src/dg/{volint,fillflux,surfint}.f90).  Resistive case (PARABOLIC):
// k_lifting with MhdLift (BR1 lifting of the entropy variables) forms the diffusive fluxes
// and writes the weak-form viscous volume term (VolInt_visc) and each side's face-normal
// diffusive flux; the surface kernel adds their BR1 mean, the volume kernel the volume
// term, the time step the viscous limit.
//
// Kernels per RK stage: k_mhd_surface (side-centric: conservative flux plus the
// non-conservative term of each side -> FluxM, FluxS) and k_mhd_volume (element-centric like
// k_volume: node fields in LDS, one thread per (direction, element, line) evaluating each
// pair's two-point flux once, the Powell terms from per-node line sums, direction sum in
// LDS, surface gather +FluxM L0 / -FluxS L0, -sJ, RK update).  Reformulations exact in real
// arithmetic: per-node logs for LN_MEAN, the Cartesian frame for the surface flux (the
// reference rotates momentum and B into (n,t1,t2) and back), the Powell volume term at node
// i as 1/2 Phi_i (B_i.m_i sum_l Dvolsurf(i,l) + sum_l Dvolsurf(i,l) B_l.{{m}}_il).
#pragma once
#include "fx_kernels.hpp"

namespace fx {

constexpr int MV = 9;

struct MhdArgs {
  double kappa, smu0;   // kappa, 1/mu_0
};

// fields staged per node (EvalUaux mhd/flux_average.f90:223-259 plus beta and the logs)
enum { MF_RHO = 0, MF_V1, MF_V2, MF_V3, MF_B1, MF_B2, MF_B3, MF_PSI, MF_BETA, MF_LNRHO, MF_LNBETA,
       MF_VV, MF_BB, MF_VB, MF_COUNT };

__device__ __forceinline__ void mhd_node_fields(const double *u, const MhdArgs &a, double *f)
{
  const double km1 = a.kappa - 1.0, km2 = a.kappa - 2.0, s2mu0 = 0.5 * a.smu0;
  const double sr = 1.0 / u[0];
  const double v1 = u[1] * sr, v2 = u[2] * sr, v3 = u[3] * sr;
  const double vv = v1 * v1 + v2 * v2 + v3 * v3;
  const double BB = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  const double pt = km1 * (u[4] - 0.5 * u[0] * vv - s2mu0 * u[8] * u[8]) - km2 * s2mu0 * BB;
  const double beta = 0.5 * u[0] / (pt - s2mu0 * BB);
  f[MF_RHO] = u[0]; f[MF_V1] = v1; f[MF_V2] = v2; f[MF_V3] = v3;
  f[MF_B1] = u[5]; f[MF_B2] = u[6]; f[MF_B3] = u[7]; f[MF_PSI] = u[8];
  f[MF_BETA] = beta; f[MF_LNRHO] = log(u[0]); f[MF_LNBETA] = log(beta);
  f[MF_VV] = vv; f[MF_BB] = BB; f[MF_VB] = v1 * u[5] + v2 * u[6] + v3 * u[7];
}

// EntropyAndKinEnergyConservingFluxVec_Derigs (mhd/flux_average.f90:761-848)
template <typename R>
__device__ __forceinline__ void derigs_flux(const R *L, const R *Rt, const R *mL, const R *mR, R ch,
                                            R smu0, R skm1, R *F)
{
  const R m0 = R(0.5) * (mL[0] + mR[0]), m1 = R(0.5) * (mL[1] + mR[1]), m2 = R(0.5) * (mL[2] + mR[2]);
  const R rhoLN = ln_mean(L[MF_RHO], Rt[MF_RHO], L[MF_LNRHO], Rt[MF_LNRHO]);
  const R sbetaLN = inv_ln_mean(L[MF_BETA], Rt[MF_BETA], L[MF_LNBETA], Rt[MF_LNBETA]);
  const R vA0 = R(0.5) * (L[MF_V1] + Rt[MF_V1]), vA1 = R(0.5) * (L[MF_V2] + Rt[MF_V2]),
          vA2 = R(0.5) * (L[MF_V3] + Rt[MF_V3]);
  const R BA0 = R(0.5) * (L[MF_B1] + Rt[MF_B1]), BA1 = R(0.5) * (L[MF_B2] + Rt[MF_B2]),
          BA2 = R(0.5) * (L[MF_B3] + Rt[MF_B3]);
  const R pTilde = R(0.5) * ((L[MF_RHO] + Rt[MF_RHO]) * rcp_nr(L[MF_BETA] + Rt[MF_BETA]) +
                             smu0 * R(0.5) * (L[MF_BB] + Rt[MF_BB]));
  const R vm = vA0 * m0 + vA1 * m1 + vA2 * m2;
  const R Bm = BA0 * m0 + BA1 * m1 + BA2 * m2;
  const R psiA = R(0.5) * (L[MF_PSI] + Rt[MF_PSI]);
  F[0] = rhoLN * vm;
  F[1] = F[0] * vA0 - (smu0 * Bm) * BA0 + pTilde * m0;
  F[2] = F[0] * vA1 - (smu0 * Bm) * BA1 + pTilde * m1;
  F[3] = F[0] * vA2 - (smu0 * Bm) * BA2 + pTilde * m2;
  F[5] = vm * BA0 - Bm * vA0 + (ch * psiA) * m0;
  F[6] = vm * BA1 - Bm * vA1 + (ch * psiA) * m1;
  F[7] = vm * BA2 - Bm * vA2 + (ch * psiA) * m2;
  F[8] = ch * Bm;
  const R B2vm = (L[MF_BB] * L[MF_V1] + Rt[MF_BB] * Rt[MF_V1]) * m0 +
                 (L[MF_BB] * L[MF_V2] + Rt[MF_BB] * Rt[MF_V2]) * m1 +
                 (L[MF_BB] * L[MF_V3] + Rt[MF_BB] * Rt[MF_V3]) * m2;
  const R psiBm = (L[MF_PSI] * L[MF_B1] + Rt[MF_PSI] * Rt[MF_B1]) * m0 +
                  (L[MF_PSI] * L[MF_B2] + Rt[MF_PSI] * Rt[MF_B2]) * m1 +
                  (L[MF_PSI] * L[MF_B3] + Rt[MF_PSI] * Rt[MF_B3]) * m2;
  F[4] = F[0] * R(0.5) * (skm1 * sbetaLN - R(0.5) * (L[MF_VV] + Rt[MF_VV])) + (vA0 * F[1] + vA1 * F[2] + vA2 * F[3]) +
         smu0 * ((BA0 * F[5] + BA1 * F[6] + BA2 * F[7]) - R(0.25) * B2vm + R(0.5) * (L[MF_VB] + Rt[MF_VB]) * Bm +
                 F[8] * psiA - (ch * R(0.5)) * psiBm);
}

// physical GLM-MHD flux along the unit normal n (EvalAdvectionFlux1D mhd/flux.f90:391-442 in
// the Cartesian frame) and the fast magnetosonic speed along n (FastestWave1D
// mhd/equation_vars.f90:663-684)
__device__ __forceinline__ void mhd_flux_n(const double *U, const double *n, const MhdArgs &a, double ch,
                                           double *F, double &vn, double &cf)
{
  const double km1 = a.kappa - 1.0, km2 = a.kappa - 2.0, smu0 = a.smu0, s2mu0 = 0.5 * smu0;
  const double sr = 1.0 / U[0];
  const double v1 = U[1] * sr, v2 = U[2] * sr, v3 = U[3] * sr;
  const double Et = U[4] - s2mu0 * U[8] * U[8];
  const double BB = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
  const double pt = km1 * (Et - 0.5 * U[0] * (v1 * v1 + v2 * v2 + v3 * v3)) - km2 * s2mu0 * BB;
  vn = v1 * n[0] + v2 * n[1] + v3 * n[2];
  const double Bn = U[5] * n[0] + U[6] * n[1] + U[7] * n[2];
  const double vb = U[5] * v1 + U[6] * v2 + U[7] * v3;
  F[0] = U[0] * vn;
  F[1] = U[1] * vn + pt * n[0] - smu0 * Bn * U[5];
  F[2] = U[2] * vn + pt * n[1] - smu0 * Bn * U[6];
  F[3] = U[3] * vn + pt * n[2] - smu0 * Bn * U[7];
  F[4] = (Et + pt) * vn - smu0 * Bn * vb + smu0 * ch * Bn * U[8];
  F[5] = vn * U[5] - Bn * v1 + ch * U[8] * n[0];
  F[6] = vn * U[6] - Bn * v2 + ch * U[8] * n[1];
  F[7] = vn * U[7] - Bn * v3 + ch * U[8] * n[2];
  F[8] = ch * Bn;
  // ConsToPrim pressure (mhd/equation_vars.f90:248) and FastestWave1D with B_n, |B_t|^2
  const double p = km1 * (U[4] - 0.5 * (U[1] * v1 + U[2] * v2 + U[3] * v3) - s2mu0 * (BB + U[8] * U[8]));
  const double c2 = a.kappa * p * sr, ca2 = Bn * Bn * sr * smu0, va2 = (BB - Bn * Bn) * sr * smu0 + ca2;
  const double astar = sqrt((c2 + va2) * (c2 + va2) - 4.0 * c2 * ca2);
  cf = sqrt(0.5 * (c2 + va2 + astar));
}

// Components 1..7 of f*a0 + g*a1 + h*a2 for the packed diffusive values s (MhdLift); the
// mass and psi components vanish.
template <typename R>
__device__ __forceinline__ void mhd_visc_dir(const R *s, const R *a, R *o)
{
  o[1] = s[0] * a[0] + s[1] * a[1] + s[2] * a[2];
  o[2] = s[1] * a[0] + s[3] * a[1] + s[4] * a[2];
  o[3] = s[2] * a[0] + s[4] * a[1] + s[5] * a[2];
  o[4] = s[9] * a[0] + s[10] * a[1] + s[11] * a[2];
  o[5] = -(s[6] * a[1]) - s[7] * a[2];
  o[6] = s[6] * a[0] - s[8] * a[2];
  o[7] = s[7] * a[0] + s[8] * a[1];
}

// Resistive GLM-MHD lifting physics: entropy variables (ConsToEntropy
// mhd/equation_vars.f90:330-361), the conversion of their gradients to gradients of
// (rho, v, p, B, psi) (ConvertToGradPrimVec, :547-628) and the isotropic diffusive fluxes
// (EvalDiffFlux3D mhd/flux.f90:449-573).  The 18 nonzero Cartesian components of f, g, h
// hold 12 independent values, written per node as
//   s = (f1, f2, f3, g2, g3, h3, f6, f7, g7, f4, g4, h4)
// (f1..f3 the viscous stress row x, f6/f7/g7 the resistive terms, f4/g4/h4 the energy
// fluxes); g1 = f2, h1 = f3, h2 = g3, g5 = -f6, h5 = -f7, h6 = -g7, the rest vanish.
struct MhdLift {
  static constexpr int NV = 9, NVV = 7;
  double kappa, smu0, mu, etasmu0, KappasPr;
  // components 1..7 of f a0 + g a1 + h a2
  __device__ __forceinline__ static void vis_dir(const double *s, double a0, double a1, double a2, double *o)
  {
    const double a[3] = {a0, a1, a2};
    double t[8];
    mhd_visc_dir<double>(s, a, t);
#pragma unroll
    for (int v = 0; v < 7; v++) o[v] = t[v + 1];
  }
  __device__ __forceinline__ void entropy(const double *c, double *w) const
  {
    const double km1 = kappa - 1.0, s2mu0 = 0.5 * smu0;
    const double srho = 1.0 / c[0];
    const double u = c[1] * srho, v = c[2] * srho, ww = c[3] * srho;
    const double v2s2 = 0.5 * (u * u + v * v + ww * ww);
    const double b2 = c[5] * c[5] + c[6] * c[6] + c[7] * c[7] + c[8] * c[8];
    const double rho_sp = c[0] / (km1 * (c[4] - c[0] * v2s2 - s2mu0 * b2));
    const double s = -log(rho_sp * pow(c[0], km1));
    w[0] = (kappa - s) * (1.0 / km1) - rho_sp * v2s2;
    w[1] = rho_sp * u;
    w[2] = rho_sp * v;
    w[3] = rho_sp * ww;
    w[4] = -rho_sp;
#pragma unroll
    for (int i = 5; i < NV; i++) w[i] = rho_sp * c[i];
  }
  __device__ __forceinline__ void grad_prim(const double *c, double *g) const
  {
    const double km1 = kappa - 1.0, s2mu0 = 0.5 * smu0;
    const double srho = 1.0 / c[0];
    const double u = c[1] * srho, v = c[2] * srho, w = c[3] * srho;
    const double ekin = 0.5 * (c[1] * u + c[2] * v + c[3] * w);
    const double b2 = c[5] * c[5] + c[6] * c[6] + c[7] * c[7] + c[8] * c[8];
    const double p = km1 * (c[4] - ekin - s2mu0 * b2);
    const double rho_sp = c[0] / p, p_srho = p * srho;
    const double g1 = g[0], g5 = g[4];
    const double gu = p_srho * (g[1] + u * g5), gv = p_srho * (g[2] + v * g5), gw = p_srho * (g[3] + w * g5);
    g[0] = c[0] * g1 - g5 * (ekin - p * (1.0 / km1)) + rho_sp * (c[1] * gu + c[2] * gv + c[3] * gw);
    g[1] = gu;
    g[2] = gv;
    g[3] = gw;
    g[4] = p_srho * (g[0] + p * g5);
#pragma unroll
    for (int i = 5; i < NV; i++) g[i] = p_srho * (g[i] + c[i] * g5);
  }
  __device__ __forceinline__ void fluxes(const double *c, const double *gx, const double *gy, const double *gz,
                                         double *sv) const
  {
    const double km1 = kappa - 1.0, skm1 = 1.0 / km1, s2mu0 = 0.5 * smu0, s23 = 2.0 / 3.0;
    const double srho = 1.0 / c[0];
    const double v1 = c[1] * srho, v2 = c[2] * srho, v3 = c[3] * srho;
    const double b2 = c[5] * c[5] + c[6] * c[6] + c[7] * c[7] + c[8] * c[8];
    const double p = km1 * (c[4] - 0.5 * (c[0] * (v1 * v1 + v2 * v2 + v3 * v3)) - s2mu0 * b2);
    const double b1 = c[5], bb2 = c[6], b3 = c[7];
    const double divv = gx[1] + gy[2] + gz[3];
    const double lambda = mu * KappasPr;
    const double qx = lambda * (skm1 * srho * (gx[4] - srho * p * gx[0]));
    const double qy = lambda * (skm1 * srho * (gy[4] - srho * p * gy[0]));
    const double qz = lambda * (skm1 * srho * (gz[4] - srho * p * gz[0]));
    const double f1 = -mu * (2 * gx[1] - s23 * divv), f2 = -mu * (gx[2] + gy[1]), f3 = -mu * (gx[3] + gz[1]);
    const double g2 = -mu * (2 * gy[2] - s23 * divv), g3 = -mu * (gy[3] + gz[2]);
    const double h3 = -mu * (2 * gz[3] - s23 * divv);
    const double f6 = -etasmu0 * (gx[6] - gy[5]), f7 = -etasmu0 * (gx[7] - gz[5]);
    const double g7 = -etasmu0 * (gy[7] - gz[6]);
    sv[0] = f1; sv[1] = f2; sv[2] = f3; sv[3] = g2; sv[4] = g3; sv[5] = h3;
    sv[6] = f6; sv[7] = f7; sv[8] = g7;
    sv[9] = v1 * f1 + v2 * f2 + v3 * f3 + smu0 * (bb2 * f6 + b3 * f7) - qx;
    sv[10] = v1 * f2 + v2 * g2 + v3 * g3 + smu0 * (b1 * (-f6) + b3 * g7) - qy;
    sv[11] = v1 * f3 + v2 * g3 + v3 * h3 + smu0 * (b1 * (-f7) + bb2 * (-g7)) - qz;
  }
};

// Powell non-conservative surface term of the left state (AddNonConsFlux mhd/riemann.f90:230-267)
__device__ __forceinline__ void mhd_noncons_surface(const double *UL, const double *UR, const double *n, double *F)
{
  const double vL0 = UL[1] / UL[0], vL1 = UL[2] / UL[0], vL2 = UL[3] / UL[0];
  const double bn = 0.5 * ((UL[5] + UR[5]) * n[0] + (UL[6] + UR[6]) * n[1] + (UL[7] + UR[7]) * n[2]);
  F[1] += bn * UL[5]; F[2] += bn * UL[6]; F[3] += bn * UL[7];
  F[4] += bn * (UL[5] * vL0 + UL[6] * vL1 + UL[7] * vL2);
  F[5] += bn * vL0; F[6] += bn * vL1; F[7] += bn * vL2;
}

// Entropy-conserving surface fluxes along the unit normal n: the Derigs flux
// (EntropyAndKinEnergyConservingFlux_Derigs mhd/flux_average.f90:660-751; Riemann=11) and
// EntropyStableByLLF (mhd/riemann.f90:1027-1211; Riemann=10), which subtracts
// LambdaMax_s2 H [v_R - v_L] with the jump of the entropy variables and the reference's
// symmetric H (Euler block in the log-mean averages, tau = 1/(beta_L+beta_R) on B and psi).
// The reference evaluates both on states rotated to (n, t1, t2) and rotates the flux back; here
// the same quantities are formed in the Cartesian frame: scalars and the products u_avg.n,
// B_avg.n are rotation invariant, the momentum row of H is u_avg (rhoLN v1 + rhoLN u_avg.vv +
// (Eavg+pAvg) v5) + pAvg vv, which is what rotating the rotated-frame rows back gives.
template <int RS>
__device__ __forceinline__ void mhd_riemann_ec(const double *UL, const double *UR, const double *n, const MhdArgs &a,
                                               double ch, double *F)
{
  const double km1 = a.kappa - 1.0, skm1 = 1.0 / km1;
  double fL[MF_COUNT], fR[MF_COUNT];
  mhd_node_fields(UL, a, fL);
  mhd_node_fields(UR, a, fR);
  derigs_flux<double>(fL, fR, n, n, ch, a.smu0, skm1, F);
  if (RS != 10) return;
  const double rhoL = fL[MF_RHO], rhoR = fR[MF_RHO], bL = fL[MF_BETA], bR = fR[MF_BETA];
  const double rhoLN = ln_mean(rhoL, rhoR, fL[MF_LNRHO], fR[MF_LNRHO]);
  const double sbetaLN = inv_ln_mean(bL, bR, fL[MF_LNBETA], fR[MF_LNBETA]);
  const double tau = 1.0 / (bL + bR);
  const double pAvg = 0.5 * (rhoL + rhoR) * tau;
  const double pLN = 0.5 * rhoLN * sbetaLN, srhoLN = 1.0 / rhoLN;
  const double uA[3] = {0.5 * (fL[MF_V1] + fR[MF_V1]), 0.5 * (fL[MF_V2] + fR[MF_V2]), 0.5 * (fL[MF_V3] + fR[MF_V3])};
  const double BA[3] = {0.5 * (fL[MF_B1] + fR[MF_B1]), 0.5 * (fL[MF_B2] + fR[MF_B2]), 0.5 * (fL[MF_B3] + fR[MF_B3])};
  const double psiA = 0.5 * (fL[MF_PSI] + fR[MF_PSI]);
  const double u2Avg = 0.5 * (fL[MF_VV] + fR[MF_VV]), B2Avg = 0.5 * (fL[MF_BB] + fR[MF_BB]);
  const double un = uA[0] * n[0] + uA[1] * n[1] + uA[2] * n[2];
  const double Bn = BA[0] * n[0] + BA[1] * n[1] + BA[2] * n[2];
  // wave speed estimate (the reference's va2Avg carries no 1/mu_0)
  const double a2 = a.kappa * pAvg * srhoLN, va2 = B2Avg * srhoLN, ca2 = Bn * Bn * srhoLN;
  const double cf = sqrt(0.5 * ((a2 + va2) + sqrt((a2 + va2) * (a2 + va2) - 4.0 * a2 * ca2)));
  const double lam = 0.5 * (fabs(un) + cf);
  const double uA2 = uA[0] * uA[0] + uA[1] * uA[1] + uA[2] * uA[2];
  const double Eavg = pLN * skm1 + 0.5 * rhoLN * (2.0 * uA2 - u2Avg);
  const double H55 = (pLN * pLN * skm1 + Eavg * Eavg) * srhoLN + uA2 * pAvg +
                     tau * ((BA[0] * BA[0] + BA[1] * BA[1] + BA[2] * BA[2]) + psiA * psiA);
  // jump of the entropy variables (ConsToEntropy, logs of the ratios as in the reference)
  const double pL = 0.5 * rhoL / bL, pR = 0.5 * rhoR / bR;
  const double j1 = (a.kappa * log(rhoR * (1.0 / rhoL)) - log(pR / pL)) * skm1 - (bR * fR[MF_VV] - bL * fL[MF_VV]);
  const double jv[3] = {2.0 * (bR * fR[MF_V1] - bL * fL[MF_V1]), 2.0 * (bR * fR[MF_V2] - bL * fL[MF_V2]),
                        2.0 * (bR * fR[MF_V3] - bL * fL[MF_V3])};
  const double j5 = -2.0 * (bR - bL);
  const double jB[3] = {2.0 * (bR * fR[MF_B1] - bL * fL[MF_B1]), 2.0 * (bR * fR[MF_B2] - bL * fL[MF_B2]),
                        2.0 * (bR * fR[MF_B3] - bL * fL[MF_B3])};
  const double j9 = 2.0 * (bR * fR[MF_PSI] - bL * fL[MF_PSI]);
  const double ujv = uA[0] * jv[0] + uA[1] * jv[1] + uA[2] * jv[2];
  const double Ep = Eavg + pAvg;
  F[0] -= lam * (rhoLN * j1 + rhoLN * ujv + Eavg * j5);
  const double cm = rhoLN * j1 + rhoLN * ujv + Ep * j5;
#pragma unroll
  for (int d = 0; d < 3; d++) F[1 + d] -= lam * (uA[d] * cm + pAvg * jv[d]);
  F[4] -= lam * (Eavg * j1 + Ep * ujv + H55 * j5 + tau * ((BA[0] * jB[0] + BA[1] * jB[1] + BA[2] * jB[2]) + psiA * j9));
  const double lt = lam * tau;
#pragma unroll
  for (int d = 0; d < 3; d++) F[5 + d] -= lt * (BA[d] * j5 + jB[d]);
  F[8] -= lt * (psiA * j5 + j9);
}

// FillFlux for GLM-MHD (fillflux.f90:41-110 with NONCONS): the surface flux RS -- standard
// DG average 0 (StandardDGFlux mhd/flux_average.f90:421-484), Rusanov 1
// (RiemannSolverByRusanov mhd/riemann.f90:273-308), entropy-stable LLF 10, Derigs EC 11 --
// in the resistive case plus the BR1 average of the diffusive fluxes (Riemann
// mhd/riemann.f90:83-142), then each side's Powell term, times SurfElem.
// Um (meshes with mortars, else null): the master states of the small sides of mortars; a big
// mortar side has no slave element and is skipped (its fluxes come from Flux_Mortar).
template <int N, int RS, bool PARA>
__global__ void __launch_bounds__(256)
k_mhd_surface(const double *__restrict__ U, const double *__restrict__ FvN, const int *__restrict__ s2e,
              const double4 *__restrict__ nvse, double *__restrict__ FluxM, double *__restrict__ FluxS, int nSides,
              MhdArgs a, const DtState *__restrict__ st, const double *__restrict__ Um)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nSides * n2) return;
  const int s = gid / n2, pq = gid - s * n2, p = pq % n, q = pq / n;
  const int em = s2e[5 * s] - 1, es = s2e[5 * s + 1] - 1;
  if (es < 0 || !(em >= 0 || (em == -2 && Um != nullptr))) return;
  const int vs = side_to_vol<N>(s2e[5 * s + 3], s2e[5 * s + 4], p, q);
  const double *um = em >= 0 ? U + ((size_t)em * n3 + side_to_vol<N>(s2e[5 * s + 2], 0, p, q)) * MV
                             : Um + (size_t)gid * MV;
  const double *us = U + ((size_t)es * n3 + vs) * MV;
  double UL[MV], UR[MV];
#pragma unroll
  for (int v = 0; v < MV; v++) { UL[v] = um[v]; UR[v] = us[v]; }
  const double4 g = nvse[gid];
  const double nv[3] = {g.x, g.y, g.z};
  const double ch = st->glm_ch;
  double Fm[MV], Fs[MV];
  if (RS == 0 || RS == 1) {
    double FL[MV], FR[MV], vnL, vnR, cfL, cfR;
    mhd_flux_n(UL, nv, a, ch, FL, vnL, cfL);
    mhd_flux_n(UR, nv, a, ch, FR, vnR, cfR);
    const double lam = RS == 1 ? fmax(fabs(vnL) + cfL, fabs(vnR) + cfR) : 0.0;
#pragma unroll
    for (int v = 0; v < MV; v++) Fm[v] = 0.5 * ((FL[v] + FR[v]) - lam * (UR[v] - UL[v]));
  } else {
    mhd_riemann_ec<RS>(UL, UR, nv, a, ch, Fm);
  }
  if (PARA) {  // BR1: mean of the two sides' normal diffusive fluxes
    const double *fn = FvN + (size_t)gid * 14;
#pragma unroll
    for (int v = 1; v < 8; v++) Fm[v] = Fm[v] + 0.5 * (fn[v - 1] + fn[6 + v]);
  }
#pragma unroll
  for (int v = 0; v < MV; v++) Fs[v] = Fm[v];
  mhd_noncons_surface(UL, UR, nv, Fm);
  mhd_noncons_surface(UR, UL, nv, Fs);
  double *om = FluxM + (size_t)gid * MV, *os = FluxS + (size_t)gid * MV;
#pragma unroll
  for (int v = 0; v < MV; v++) { om[v] = Fm[v] * g.w; os[v] = Fs[v] * g.w; }
}

// Element-centric GLM-MHD volume kernel (same phases as k_volume).
// Launch geometry of the GLM-MHD volume kernel: EPB elements per block, one thread per
// (element, direction, node), T threads rounded up to whole wavefronts.
template <int N> struct MhdGeom;
template <> struct MhdGeom<2> { static constexpr int EPB = 3; };
template <> struct MhdGeom<3> { static constexpr int EPB = 2; };
template <> struct MhdGeom<4> { static constexpr int EPB = 1; };
template <> struct MhdGeom<5> { static constexpr int EPB = 1; };
template <int N> struct MhdVolGeom {
  static constexpr int n3 = (N + 1) * (N + 1) * (N + 1), EPB = MhdGeom<N>::EPB;
  static constexpr int T = ((EPB * 3 * n3 + 63) / 64) * 64;
};

// Element-centric GLM-MHD volume kernel (VolInt_SplitForm with EvalAdvFluxAverage3D and
// AddNonConsFluxTilde3D, mhd/flux_average.f90:112-186, 268-366; VolInt_visc; SurfInt; -sJ; the
// LSERK stage update).  One thread per (element, direction d, node k at position pos of its
// line in direction d).  Each pair of a line is evaluated once, by one of its two nodes: in
// slot s = 1..n/2 node pos takes the pair (pos, pos + s mod n) (for even n the last slot only
// for pos < n/2), so every node evaluates at most n/2 two-point Derigs fluxes, the lower line
// index always the left state.  The thread adds Dvolsurf(pos,l) F and its Powell term to its
// own sums and leaves Dvolsurf(l,pos) F and the partner's Powell term in an LDS exchange
// buffer, from which the partner adds them after a barrier (a thread keeps one node's sums).
//   phase 1  one thread per (element, direction, node): the node fields (direction 0) and the
//            node's metric of its direction into LDS
//   phase 2  the flux differencing as above, sums into LDS over the fields (after a barrier)
//   phase 3  face gather +FluxM L0 / -FluxS L0 into LDS (the memory of the exchange buffer),
//            then one thread per node: direction sums, viscous volume term, surface terms,
//            -sJ, Ut or the RK stage update
template <int N, int MODE, typename VR, bool PARA>
__global__ void __launch_bounds__(MhdVolGeom<N>::T)
k_mhd_volume(double *__restrict__ U, const VR *__restrict__ Mf, const VR *__restrict__ Mg, const VR *__restrict__ Mh,
             const double *__restrict__ sJ, const int *__restrict__ e2s, const double *__restrict__ FluxM,
             const double *__restrict__ FluxS, const double *__restrict__ Uv, double *__restrict__ Ut,
             double *__restrict__ Uttmp, RKArgs rk, int nElems, MhdArgs a, const DtState *__restrict__ st)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr int EPB = MhdVolGeom<N>::EPB, T = MhdVolGeom<N>::T, NP = Lds<N>::NP;
  // per element: the node fields and the three metric vectors (phases 1-2), later the three
  // direction sums (MV values per node each)
  constexpr int FSLOTS = MF_COUNT + 9, ASLOTS = 3 * MV, SLOTS = FSLOTS > ASLOTS ? FSLOTS : ASLOTS;
  __shared__ VR s_buf[EPB][SLOTS * NP];
#define MF_AT(el_, c_, l_) s_buf[el_][(c_) * NP + (l_)]
#define MM_AT(el_, d_, c_, l_) s_buf[el_][(MF_COUNT + 3 * (d_) + (c_)) * NP + (l_)]
#define MA_AT(el_, d_, v_, l_) s_buf[el_][(MV * (d_) + (v_)) * NP + (l_)]
  const Ops &op = c_ops[N];
  const int tid = threadIdx.x, e0 = blockIdx.x * EPB;
  const VR ch = VR(st->glm_ch), smu0 = VR(a.smu0), skm1 = VR(1.0 / (a.kappa - 1.0));

  const int el = tid / (3 * n3), rr = tid - el * (3 * n3), dir = rr / n3, node = rr - dir * n3;
  const int e = e0 + el;
  const bool active = (el < EPB) && (e < nElems) && elem_problem(rk.eprob, e) >= 0;
  const int i = node % n, j = (node / n) % n, k = node / n2;
  const int ln = Lds<N>::li(node);
  // the face gather's fluxes (phase 3), loaded before phase 1 and kept in registers through the
  // pair slots, whose exchange buffer the face terms later reuse
#ifndef FX_MHDVOL_GPRE
#define FX_MHDVOL_GPRE 1
#endif
  constexpr int NFP = 6 * n2;
  constexpr bool GPRE = FX_MHDVOL_GPRE && MODE != VMODE_VOLINT && (EPB * NFP <= T);
  double gpre[MV];
  double gsg = 0.0;
  if (GPRE && tid < EPB * NFP) {
    const int fe = tid / NFP, r = tid - fe * NFP, face = r / n2, ab = r - face * n2;
    const int ee = e0 + fe;
    if (ee < nElems && elem_problem(rk.eprob, ee) >= 0) {
      const int loc = face + 1, fnode = face_node<N>(loc, ab % n, ab / n);
      const int sd = e2s[(ee * 6 + face) * 2] - 1, fl = e2s[(ee * 6 + face) * 2 + 1];
      int p, q;
      vol_to_side<N>(loc, fl, fnode % n, (fnode / n) % n, fnode / n2, p, q);
      const size_t f = ((size_t)(sd * n + q) * n + p) * MV;
      const double *F = (fl == 0) ? FluxM + f : FluxS + f;
      gsg = (fl == 0) ? op.L0 : -op.L0;
#pragma unroll
      for (int v = 0; v < MV; v++) gpre[v] = F[v];
    }
  }
  if (active) {
    if (dir == 0) {
      const double *u = U + ((size_t)e * n3 + node) * MV;
      double uu[MV], f[MF_COUNT];
#pragma unroll
      for (int v = 0; v < MV; v++) uu[v] = u[v];
      mhd_node_fields(uu, a, f);
#pragma unroll
      for (int c = 0; c < MF_COUNT; c++) MF_AT(el, c, ln) = VR(f[c]);
    }
    const VR *M = (dir == 0) ? Mf : ((dir == 1) ? Mg : Mh);
    const VR *mp = M + ((size_t)e * n3 + node) * 3;
#pragma unroll
    for (int c = 0; c < 3; c++) MM_AT(el, dir, c, ln) = mp[c];
  }
  __syncthreads();

  // exchange buffer of phase 2 (MV flux values and the Powell term per thread), later the
  // face terms of phase 3
  constexpr int NF = 6 * n2, NX = (MV + 1) * T, NFACE = MODE != VMODE_VOLINT ? EPB * NF * MV : 1;
  __shared__ double s_xf[NX > NFACE ? NX : NFACE];
  VR acc[MV];
#pragma unroll
  for (int v = 0; v < MV; v++) acc[v] = VR(0);
  VR snc = VR(0);
  const int pos = (dir == 0) ? i : ((dir == 1) ? j : k);  // the node's index along its line
  const int stride = (dir == 0) ? 1 : ((dir == 1) ? (n + 1) : (n + 1) * n);  // LDS step along the line
  const int l0 = ln - pos * stride;                        // LDS index of the line's node 0
  const int tstep = (dir == 0) ? 1 : ((dir == 1) ? n : n2);  // thread step along the line
  VR fk[MF_COUNT], mk[3];
  if (active) {
#pragma unroll
    for (int c = 0; c < MF_COUNT; c++) fk[c] = MF_AT(el, c, ln);
#pragma unroll
    for (int c = 0; c < 3; c++) mk[c] = MM_AT(el, dir, c, ln);
  }
  constexpr int SLOT_UNROLL = n / 2 <= 2 ? n / 2 : 1;  // N=5 unrolled spills
#pragma unroll SLOT_UNROLL
  for (int sl = 1; sl <= n / 2; sl++) {
    const bool half = (n % 2 == 0) && (sl == n / 2);  // the pairs (p, p + n/2): once each
    if (active && (!half || pos < n / 2)) {
      const int l = pos + sl < n ? pos + sl : pos + sl - n;
      const int ll = l0 + l * stride;
      VR fl[MF_COUNT], ml[3], F[MV];
#pragma unroll
      for (int c = 0; c < MF_COUNT; c++) fl[c] = MF_AT(el, c, ll);
#pragma unroll
      for (int c = 0; c < 3; c++) ml[c] = MM_AT(el, dir, c, ll);
      if (pos < l) derigs_flux<VR>(fk, fl, mk, ml, ch, smu0, skm1, F);
      else derigs_flux<VR>(fl, fk, ml, mk, ch, smu0, skm1, F);
      const VR mb0 = VR(0.5) * (mk[0] + ml[0]), mb1 = VR(0.5) * (mk[1] + ml[1]), mb2 = VR(0.5) * (mk[2] + ml[2]);
      const VR dkl = OpsOf<VR>::ds(op, pos, l), dlk = OpsOf<VR>::ds(op, l, pos);
#pragma unroll
      for (int v = 0; v < MV; v++) {
        acc[v] += dkl * F[v];
        s_xf[v * T + tid] = dlk * F[v];
      }
      // Powell: B_l.{{m}} towards node pos, B_pos.{{m}} towards node l
      snc += dkl * (mb0 * fl[MF_B1] + mb1 * fl[MF_B2] + mb2 * fl[MF_B3]);
      s_xf[MV * T + tid] = dlk * (mb0 * fk[MF_B1] + mb1 * fk[MF_B2] + mb2 * fk[MF_B3]);
    }
    __syncthreads();
    if (active && (!half || pos >= n / 2)) {  // the pair (m, pos) taken by node m in this slot
      const int m = pos - sl >= 0 ? pos - sl : pos - sl + n;
      const int ts = tid + (m - pos) * tstep;
#pragma unroll
      for (int v = 0; v < MV; v++) acc[v] += s_xf[v * T + ts];
      snc += s_xf[MV * T + ts];
    }
    __syncthreads();  // the next slot overwrites the exchange buffer
  }
  if (active) {
    const VR b1 = fk[MF_B1], b2 = fk[MF_B2], b3 = fk[MF_B3];
    const VR bhat = mk[0] * b1 + mk[1] * b2 + mk[2] * b3;
    const VR t = VR(0.5) * (bhat * VR(op.DsRow[pos]) + snc);
    acc[1] += t * b1; acc[2] += t * b2; acc[3] += t * b3;
    acc[4] += t * fk[MF_VB];
    acc[5] += t * fk[MF_V1];
    acc[6] += t * fk[MF_V2];
    acc[7] += t * fk[MF_V3];
  }
  __syncthreads();  // the sums overwrite the fields: every thread must have read them
  if (active) {
#pragma unroll
    for (int v = 0; v < MV; v++) MA_AT(el, dir, v, ln) = acc[v];
  }

  // the element's SurfInt terms +FluxM L0 (master) / -FluxS L0 (slave) into LDS in
  // element-local face coordinates, one work item per face point (the exchange buffer is no
  // longer read: the last slot ended with a barrier)
  const double L0 = op.L0;
  auto s_face = reinterpret_cast<double (*)[NF * MV]>(s_xf);
  if (GPRE) {
    if (tid < EPB * NF && gsg != 0.0) {
      const int fe = tid / NF, r = tid - fe * NF;
#pragma unroll
      for (int v = 0; v < MV; v++) s_face[fe][r * MV + v] = gpre[v] * gsg;
    }
  } else if (MODE != VMODE_VOLINT) {
    for (int it = tid; it < EPB * NF; it += T) {
      const int fe = it / NF, r = it - fe * NF, face = r / n2, ab = r - face * n2;
      const int ee = e0 + fe;
      if (ee >= nElems || elem_problem(rk.eprob, ee) < 0) continue;
      const int loc = face + 1, fnode = face_node<N>(loc, ab % n, ab / n);
      const int sd = e2s[(ee * 6 + face) * 2] - 1, fl = e2s[(ee * 6 + face) * 2 + 1];
      int p, q;
      vol_to_side<N>(loc, fl, fnode % n, (fnode / n) % n, fnode / n2, p, q);
      const size_t f = ((size_t)(sd * n + q) * n + p) * MV;
      const double *F = (fl == 0) ? FluxM + f : FluxS + f;
      const double sg = (fl == 0) ? L0 : -L0;
#pragma unroll
      for (int v = 0; v < MV; v++) s_face[fe][r * MV + v] = F[v] * sg;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < EPB * n3; idx += T) {
    const int el2 = idx / n3, node2 = idx - el2 * n3, e2 = e0 + el2;
    if (e2 >= nElems) continue;
    const int prob = elem_problem(rk.eprob, e2);
    if (prob < 0) continue;
    const int i2 = node2 % n, j2 = (node2 / n) % n, k2 = node2 / n2, l2 = Lds<N>::li(node2);
    double tot[MV];
#pragma unroll
    for (int v = 0; v < MV; v++) {
      VR sd = MA_AT(el2, 0, v, l2);
      sd = sd + MA_AT(el2, 1, v, l2);
      sd = sd + MA_AT(el2, 2, v, l2);
      tot[v] = double(sd);
    }
    if (PARA && MODE != VMODE_VOLINT) {  // VolInt_visc (volint.f90:228-271), formed by the lifting kernel
      const double *uv = Uv + ((size_t)e2 * n3 + node2) * 7;
#pragma unroll
      for (int v = 1; v < 8; v++) tot[v] = tot[v] + uv[v - 1];
    }
    if (MODE != VMODE_VOLINT) {
      const NodeFaces nf = node_faces<N>(i2, j2, k2);
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int r = nf.get(t);
        if (r < 0) continue;
#pragma unroll
        for (int v = 0; v < MV; v++) tot[v] = tot[v] + s_face[el2][r * MV + v];
      }
      const double msj = -sJ[(size_t)e2 * n3 + node2];
#pragma unroll
      for (int v = 0; v < MV; v++) tot[v] = tot[v] * msj;
    }
    const size_t o = ((size_t)e2 * n3 + node2) * MV;
    if (MODE == VMODE_RK) {
      const double bdt = rk.b * rk.st[prob].dt;
#pragma unroll
      for (int v = 0; v < MV; v++) {
        const double ut = rk.first ? tot[v] : Uttmp[o + v] * (-rk.A) + tot[v];
        Uttmp[o + v] = ut;
        U[o + v] = U[o + v] + ut * bdt;
      }
    } else {
#pragma unroll
      for (int v = 0; v < MV; v++) Ut[o + v] = tot[v];
    }
  }
#undef MF_AT
#undef MM_AT
#undef MA_AT
}

// The line form of the GLM-MHD volume kernel, used for the split-precision path: one thread
// per (direction, element, line) evaluates each pair's flux once and keeps the n nodes' sums
// in registers (in FP32 these fit at two waves per SIMD, and evaluating every pair once is
// faster there than the node form above).  Same phases and sums otherwise.
template <int N> struct MhdLineGeom;
template <> struct MhdLineGeom<2> { static constexpr int EPB = 7, W = 64; };
template <> struct MhdLineGeom<3> { static constexpr int EPB = 4, W = 64; };
template <> struct MhdLineGeom<4> { static constexpr int EPB = 2, W = 64; };
template <> struct MhdLineGeom<5> { static constexpr int EPB = 1, W = 64; };

template <int N, int MODE, typename VR, bool PARA>
__global__ void __launch_bounds__(3 * MhdLineGeom<N>::W)
k_mhd_volume_line(double *__restrict__ U, const VR *__restrict__ Mf, const VR *__restrict__ Mg, const VR *__restrict__ Mh,
             const double *__restrict__ sJ, const int *__restrict__ e2s, const double *__restrict__ FluxM,
             const double *__restrict__ FluxS, const double *__restrict__ Uv, double *__restrict__ Ut,
             double *__restrict__ Uttmp, RKArgs rk, int nElems, MhdArgs a, const DtState *__restrict__ st)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr int EPB = MhdLineGeom<N>::EPB, W = MhdLineGeom<N>::W, NP = Lds<N>::NP;
  static_assert(EPB * n2 <= W, "MHD volume launch geometry");
  __shared__ VR s_f[EPB][MF_COUNT * NP];
  __shared__ VR s_acc[EPB][MV * NP];
  const Ops &op = c_ops[N];
  const int tid = threadIdx.x, e0 = blockIdx.x * EPB;
  const VR ch = VR(st->glm_ch), smu0 = VR(a.smu0), skm1 = VR(1.0 / (a.kappa - 1.0));

  for (int idx = tid; idx < EPB * n3; idx += 3 * W) {
    const int el = idx / n3, node = idx - el * n3, e = e0 + el;
    if (e < nElems && elem_problem(rk.eprob, e) >= 0) {
      const double *u = U + ((size_t)e * n3 + node) * MV;
      double uu[MV], f[MF_COUNT];
#pragma unroll
      for (int v = 0; v < MV; v++) uu[v] = u[v];
      mhd_node_fields(uu, a, f);
      const int l = Lds<N>::li(node);
#pragma unroll
      for (int c = 0; c < MF_COUNT; c++) s_f[el][c * NP + l] = VR(f[c]);
    }
  }
  __syncthreads();

  const int dir = tid / W, lt = tid - dir * W;
  const int el = lt / n2, line = lt - el * n2, e = e0 + el;
  const bool active = (el < EPB) && (e < nElems) && elem_problem(rk.eprob, e) >= 0;
  const int aa = line % n, bb = line / n;
  int nodes[n], lds[n];
#pragma unroll
  for (int l = 0; l < n; l++) {
    const int i = (dir == 0) ? l : aa, j = (dir == 1) ? l : (dir == 0 ? aa : bb), k = (dir == 2) ? l : bb;
    nodes[l] = i + n * (j + n * k);
    lds[l] = Lds<N>::li(i, j, k);
  }
  VR acc[n][MV], snc[n];
#pragma unroll
  for (int i = 0; i < n; i++) {
    snc[i] = VR(0);
#pragma unroll
    for (int v = 0; v < MV; v++) acc[i][v] = VR(0);
  }
  if (active) {
    const VR *M = (dir == 0) ? Mf : ((dir == 1) ? Mg : Mh);
    VR m[n][3];
#pragma unroll
    for (int l = 0; l < n; l++) {
      const VR *mp = M + ((size_t)e * n3 + nodes[l]) * 3;
      m[l][0] = mp[0]; m[l][1] = mp[1]; m[l][2] = mp[2];
    }
#pragma unroll
    for (int i = 0; i < n; i++)
#pragma unroll
      for (int l = i + 1; l < n; l++) {
        VR fi[MF_COUNT], fl[MF_COUNT], F[MV];
#pragma unroll
        for (int c = 0; c < MF_COUNT; c++) { fi[c] = s_f[el][c * NP + lds[i]]; fl[c] = s_f[el][c * NP + lds[l]]; }
        derigs_flux<VR>(fi, fl, m[i], m[l], ch, smu0, skm1, F);
        const VR dil = OpsOf<VR>::ds(op, i, l), dli = OpsOf<VR>::ds(op, l, i);
#pragma unroll
        for (int v = 0; v < MV; v++) {
          acc[i][v] += dil * F[v];
          acc[l][v] += dli * F[v];
        }
        // Powell terms: B_l.{{m}} towards node i, B_i.{{m}} towards node l
        const VR mb0 = VR(0.5) * (m[i][0] + m[l][0]), mb1 = VR(0.5) * (m[i][1] + m[l][1]),
                 mb2 = VR(0.5) * (m[i][2] + m[l][2]);
        snc[i] += dil * (mb0 * fl[MF_B1] + mb1 * fl[MF_B2] + mb2 * fl[MF_B3]);
        snc[l] += dli * (mb0 * fi[MF_B1] + mb1 * fi[MF_B2] + mb2 * fi[MF_B3]);
      }
#pragma unroll
    for (int i = 0; i < n; i++) {
      const VR b1 = s_f[el][MF_B1 * NP + lds[i]], b2 = s_f[el][MF_B2 * NP + lds[i]], b3 = s_f[el][MF_B3 * NP + lds[i]];
      const VR bhat = m[i][0] * b1 + m[i][1] * b2 + m[i][2] * b3;
      const VR t = VR(0.5) * (bhat * VR(op.DsRow[i]) + snc[i]);
      acc[i][1] += t * b1; acc[i][2] += t * b2; acc[i][3] += t * b3;
      acc[i][4] += t * s_f[el][MF_VB * NP + lds[i]];
      acc[i][5] += t * s_f[el][MF_V1 * NP + lds[i]];
      acc[i][6] += t * s_f[el][MF_V2 * NP + lds[i]];
      acc[i][7] += t * s_f[el][MF_V3 * NP + lds[i]];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (active && dir == d) {
#pragma unroll
      for (int l = 0; l < n; l++)
#pragma unroll
        for (int v = 0; v < MV; v++)
          s_acc[el][v * NP + lds[l]] = (d == 0) ? acc[l][v] : s_acc[el][v * NP + lds[l]] + acc[l][v];
    }
    __syncthreads();
  }

  // the element's SurfInt terms +FluxM L0 (master) / -FluxS L0 (slave) into LDS in
  // element-local face coordinates, one work item per face point
  const double L0 = op.L0;
  constexpr int NF = 6 * n2;
  __shared__ double s_face[MODE != VMODE_VOLINT ? EPB : 1][MODE != VMODE_VOLINT ? NF * MV : 1];
  if (MODE != VMODE_VOLINT) {
    for (int it = tid; it < EPB * NF; it += 3 * W) {
      const int fe = it / NF, r = it - fe * NF, face = r / n2, ab = r - face * n2;
      const int ee = e0 + fe;
      if (ee >= nElems || elem_problem(rk.eprob, ee) < 0) continue;
      const int loc = face + 1, fnode = face_node<N>(loc, ab % n, ab / n);
      const int sd = e2s[(ee * 6 + face) * 2] - 1, fl = e2s[(ee * 6 + face) * 2 + 1];
      int p, q;
      vol_to_side<N>(loc, fl, fnode % n, (fnode / n) % n, fnode / n2, p, q);
      const size_t f = ((size_t)(sd * n + q) * n + p) * MV;
      const double *F = (fl == 0) ? FluxM + f : FluxS + f;
      const double sg = (fl == 0) ? L0 : -L0;
#pragma unroll
      for (int v = 0; v < MV; v++) s_face[fe][r * MV + v] = F[v] * sg;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < EPB * n3; idx += 3 * W) {
    const int el2 = idx / n3, node = idx - el2 * n3, e2 = e0 + el2;
    if (e2 >= nElems) continue;
    const int prob = elem_problem(rk.eprob, e2);
    if (prob < 0) continue;
    const int i = node % n, j = (node / n) % n, k = node / n2, l = Lds<N>::li(node);
    double tot[MV];
#pragma unroll
    for (int v = 0; v < MV; v++) tot[v] = double(s_acc[el2][v * NP + l]);
    if (PARA && MODE != VMODE_VOLINT) {  // VolInt_visc (volint.f90:228-271), formed by the lifting kernel
      const double *uv = Uv + ((size_t)e2 * n3 + node) * 7;
#pragma unroll
      for (int v = 1; v < 8; v++) tot[v] = tot[v] + uv[v - 1];
    }
    if (MODE != VMODE_VOLINT) {
      const NodeFaces nf = node_faces<N>(i, j, k);
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int r = nf.get(t);
        if (r < 0) continue;
#pragma unroll
        for (int v = 0; v < MV; v++) tot[v] = tot[v] + s_face[el2][r * MV + v];
      }
      const double msj = -sJ[(size_t)e2 * n3 + node];
#pragma unroll
      for (int v = 0; v < MV; v++) tot[v] = tot[v] * msj;
    }
    const size_t o = ((size_t)e2 * n3 + node) * MV;
    if (MODE == VMODE_RK) {
      const double bdt = rk.b * rk.st[prob].dt;
#pragma unroll
      for (int v = 0; v < MV; v++) {
        const double ut = rk.first ? tot[v] : Uttmp[o + v] * (-rk.A) + tot[v];
        Uttmp[o + v] = ut;
        U[o + v] = U[o + v] + ut * bdt;
      }
    } else {
#pragma unroll
      for (int v = 0; v < MV; v++) Ut[o + v] = tot[v];
    }
  }
}

// CALCTIMESTEP for GLM-MHD (mhd/calctimestep.f90:39-179) and InitTimeStep_GLM (:186-215):
// per element the convective dt (FastestWave3D), in the resistive case limited by
// DFL*4/max_node(max(eta/mu_0, mu max(4/3, kappa/Pr)/rho) sum_d |sJ m_d|^2), and the dt of a
// unit wave speed.
struct MhdDtArgs {
  double cfl, dfl, muk, diffc;  // CFLScale, DFLScale, mu*max(4/3, kappa/Pr), eta/mu_0
  int parabolic;
};

template <int N>
__global__ void __launch_bounds__(DtGeom<N>::T)
k_mhd_calc_dt(const double *__restrict__ U, const double *__restrict__ Mf, const double *__restrict__ Mg,
              const double *__restrict__ Mh, const double *__restrict__ sJ, double *__restrict__ dtElem,
              double *__restrict__ dtch1Elem, int *__restrict__ errElem, int nElems, MhdArgs a, MhdDtArgs da)
{
  constexpr int n3 = DtGeom<N>::n3, G = DtGeom<N>::G, EPB = DtGeom<N>::EPB, T = DtGeom<N>::T;
  __shared__ double s_l[7][T];
  __shared__ int s_err[T];
  const int tid = threadIdx.x, el = tid / G, node = tid - el * G, e = blockIdx.x * EPB + el;
  const double cfl = da.cfl;
  double l[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int err = 0;
  if (node < n3 && e < nElems) {
    const size_t c = (size_t)e * n3 + node;
    const double *u = U + MV * c;
    if (isnan(u[0])) err = 1;
    const double sr = 1.0 / u[0];
    const double v0 = u[1] * sr, v1 = u[2] * sr, v2 = u[3] * sr;
    double b2 = 0.0;
#pragma unroll
    for (int v = 5; v < MV; v++) b2 += u[v] * u[v];
    const double p = (a.kappa - 1.0) * (u[4] - 0.5 * (u[1] * v0 + u[2] * v1 + u[3] * v2) - 0.5 * a.smu0 * b2);
    if (p <= 0.0) err = 2;
    const double cf = sqrt((a.kappa * p + (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) * a.smu0) / u[0]);
    const double sj = sJ[c];
    const double *mf = Mf + 3 * c, *mg = Mg + 3 * c, *mh = Mh + 3 * c;
    const double nf = sqrt(mf[0] * mf[0] + mf[1] * mf[1] + mf[2] * mf[2]);
    const double ng = sqrt(mg[0] * mg[0] + mg[1] * mg[1] + mg[2] * mg[2]);
    const double nh = sqrt(mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2]);
    l[0] = sj * (fabs(mf[0] * v0 + mf[1] * v1 + mf[2] * v2) + cf * nf);
    l[1] = sj * (fabs(mg[0] * v0 + mg[1] * v1 + mg[2] * v2) + cf * ng);
    l[2] = sj * (fabs(mh[0] * v0 + mh[1] * v1 + mh[2] * v2) + cf * nh);
    l[3] = sj * nf; l[4] = sj * ng; l[5] = sj * nh;
    if (da.parabolic) {
      const double lf = (mf[0] * sj) * (mf[0] * sj) + (mf[1] * sj) * (mf[1] * sj) + (mf[2] * sj) * (mf[2] * sj);
      const double lg = (mg[0] * sj) * (mg[0] * sj) + (mg[1] * sj) * (mg[1] * sj) + (mg[2] * sj) * (mg[2] * sj);
      const double lh = (mh[0] * sj) * (mh[0] * sj) + (mh[1] * sj) * (mh[1] * sj) + (mh[2] * sj) * (mh[2] * sj);
      l[6] = fmax(da.diffc, sr * da.muk) * (lf + lg + lh);
    }
  }
#pragma unroll
  for (int d = 0; d < 7; d++) s_l[d][tid] = l[d];
  s_err[tid] = err;
  __syncthreads();
  for (int stp = G / 2; stp > 0; stp >>= 1) {
    if (node < stp) {
#pragma unroll
      for (int d = 0; d < 7; d++) s_l[d][tid] = fmax(s_l[d][tid], s_l[d][tid + stp]);
      s_err[tid] = max(s_err[tid], s_err[tid + stp]);
    }
    __syncthreads();
  }
  if (node == 0 && e < nElems) {
    int er = s_err[tid];
    double dte = cfl * 2.0 / (s_l[0][tid] + s_l[1][tid] + s_l[2][tid]);
    if (isnan(dte)) er = max(er, 3);
    if (da.parabolic && s_l[6][tid] > 0.0) dte = fmin(dte, da.dfl * 4.0 / s_l[6][tid]);
    dtElem[e] = dte;
    dtch1Elem[e] = cfl * 2.0 / (s_l[3][tid] + s_l[4][tid] + s_l[5][tid]);
    errElem[e] = er;
  }
}

#ifdef FX_KERNEL_HOME  // non-template kernel: compiled in one unit only
// minimum dt (and, on the first call, the unit-speed dt), GLM_ch = scale * dtch1 / dt
__global__ void __launch_bounds__(1024)
k_mhd_dt_reduce(const double *__restrict__ dtElem, const double *__restrict__ dtch1Elem,
                const int *__restrict__ errElem, int nElems, DtState *st, int advance, double glm_scale)
{
  __shared__ double s_dt[1024], s_c[1024];
  __shared__ int s_er[1024];
  const int tid = threadIdx.x;
  double m = 1.79e308, mc = 1.79e308;
  int er = 0;
  for (int e = tid; e < nElems; e += 1024) {
    const double d = dtElem[e], dc = dtch1Elem[e];
    if (d < m) m = d;
    if (dc < mc) mc = dc;
    er = max(er, errElem[e]);
  }
  s_dt[tid] = m; s_c[tid] = mc; s_er[tid] = er;
  __syncthreads();
  for (int stp = 512; stp > 0; stp >>= 1) {
    if (tid < stp) {
      s_dt[tid] = fmin(s_dt[tid], s_dt[tid + stp]);
      s_c[tid] = fmin(s_c[tid], s_c[tid + stp]);
      s_er[tid] = max(s_er[tid], s_er[tid + stp]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    const bool was_done = st->done != 0;
    if (!was_done) {
      // GLM_ch from the CFL time step itself (calctimestep.f90:175-177), before the end-time cut
      if (!st->glm_init) { st->glm_dtch1 = s_c[0]; st->glm_init = 1; }
      st->glm_ch = glm_scale * (st->glm_dtch1 / s_dt[0]);
    }
    st->dt = end_time_step(*st, s_dt[0], advance);
    st->errType = max(st->errType, s_er[0]);
    if (st->errType != 0 && st->errStep < 0) st->errStep = st->step;
    if (advance && !was_done) { st->t = st->t + st->dt; st->step = st->step + 1; }
  }
}

#endif  // FX_KERNEL_HOME

}  // namespace fx
