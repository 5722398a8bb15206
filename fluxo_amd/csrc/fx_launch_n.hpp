// Per-degree kernel launchers declared in fx_launch.hpp.  Included only by fx_tu.hip, which
// instantiates them for one degree and one kernel group per compilation unit.
#pragma once
#include "fx_launch.hpp"
#include "fx_analyze.hpp"

namespace fx {

template <int N, int VF, bool PARA, int MODE, int EPB>
void launch_volume_e(const Batch &b, RKArgs rk)
{
  using G = VolGeom<N, EPB>;
  const int nb = (b.cfg.nElems + EPB - 1) / EPB;
  if (nb <= 0) return;
  if constexpr (has_fp32(VF)) {
    if (b.cfg.precision == FX_PREC_FP32_SPLIT) {
      hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, float, EPB>), dim3(nb), dim3(G::TB), 0, b.stream, b.U,
                         b.Mf32, b.Mg32, b.Mh32, b.sJ, b.e2s, b.Flux, b.Uv, b.Ut, b.Uttmp, rk, b.cfg.nElems,
                         b.cfg.kappa);
      return;
    }
  }
  hipLaunchKernelGGL((k_volume<N, VF, PARA, MODE, double, EPB>), dim3(nb), dim3(G::TB), 0, b.stream, b.U,
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
                       b.nvse, b.Flux, b.cfg.nSides, b.cfg.kappa, b.Umort);
  else
    hipLaunchKernelGGL((k_surface<N, RS, false>), dim3(nb), dim3(T), 0, b.stream, b.U, b.FvN, b.s2e,
                       b.nvse, b.Flux, b.cfg.nSides, b.cfg.kappa, b.Umort);
}

template <int N>
void launch_mortar_u_n(const Batch &b)
{
  const long long tot = (long long)b.nMortar * 4 * b.n2;
  const int nb = (int)((tot + 255) / 256);
  if (nb <= 0) return;
  if (b.mhd())
    hipLaunchKernelGGL((k_mortar_u<N, 9>), dim3(nb), dim3(256), 0, b.stream, b.U, b.mort, b.mops, b.Umort, b.nMortar);
  else
    hipLaunchKernelGGL((k_mortar_u<N, 5>), dim3(nb), dim3(256), 0, b.stream, b.U, b.mort, b.mops, b.Umort, b.nMortar);
}

template <int N>
void launch_flux_mortar_n(const Batch &b)
{
  const long long tot = (long long)b.nMortar * b.n2;
  const int nb = (int)((tot + 255) / 256);
  if (nb <= 0) return;
  if (b.mhd())  // the small sides' master fluxes (with their Powell terms) onto the big side
    hipLaunchKernelGGL((k_flux_mortar<N, 9>), dim3(nb), dim3(256), 0, b.stream, b.Flux, b.Flux, b.mort, b.mops,
                       b.nMortar);
  else
    hipLaunchKernelGGL((k_flux_mortar<N, 5>), dim3(nb), dim3(256), 0, b.stream, b.Flux, b.Flux, b.mort, b.mops,
                       b.nMortar);
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

template <int N>
void launch_face_table_n(const Batch &b, int e0, int nE)
{
  const long long cnt = (long long)nE * 6 * b.n2;
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_face_table<N>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, b.stream, b.e2s, b.s2e, e0, nE,
                     b.nMortar, b.ftab);
}

template <int N, class EQ, int LIFT>
void launch_lifting_l(const Batch &b, int out, const EQ &eq, const DtFuse *dtf)
{
  const int epb = LiftGeom<N>::EPB;
  const int nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  const size_t nv = (size_t)EQ::NV * b.n3 * b.cfg.nElems;
  const DtFuse none{};
  if (out == 0)
    hipLaunchKernelGGL((k_lifting<N, 0, EQ, LIFT>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                       b.ftab, b.nvse, b.grad, b.grad + nv, b.grad + 2 * nv, b.eprob, b.cfg.nElems, eq,
                       b.cfg.etaBR2, b.Umort, b.LiftM, b.GradM, b.nMortar, none);
  else if constexpr (std::is_same<EQ, NsLift>::value) {
    if (dtf) {
      hipLaunchKernelGGL((k_lifting<N, 1, EQ, LIFT, true>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh,
                         b.sJ, b.ftab, b.nvse, b.Uv, b.FvN, nullptr, b.eprob, b.cfg.nElems, eq, b.cfg.etaBR2,
                         b.Umort, b.LiftM, b.GradM, b.nMortar, *dtf);
      return;
    }
  }
  if (out != 0)
    hipLaunchKernelGGL((k_lifting<N, 1, EQ, LIFT>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                       b.ftab, b.nvse, b.Uv, b.FvN, nullptr, b.eprob, b.cfg.nElems, eq, b.cfg.etaBR2,
                       b.Umort, b.LiftM, b.GradM, b.nMortar, none);
}

// Lifting (BR1 or BR2) on a mesh with mortars: U_Mortar of the state, the small sides' lifting
// fluxes and their Flux_Cons_Mortar onto the big sides before the lifting kernel; after it
// (viscous form) U_Mortar of the big sides' gradients and the master viscous flux of the small
// sides.
template <int N, class EQ>
void launch_lifting_mortar(const Batch &b, int out, const EQ &eq, const DtFuse *dtf)
{
  const int nbs = (int)(((long long)b.nMortar * 4 * b.n2 + 255) / 256);
  const int nbb = (int)(((long long)b.nMortar * b.n2 + 255) / 256);
  hipLaunchKernelGGL((k_mortar_u<N, EQ::NV>), dim3(nbs), dim3(256), 0, b.stream, b.U, b.mort, b.mops, b.Umort,
                     b.nMortar);
  hipLaunchKernelGGL((k_lift_mortar_small<N, EQ>), dim3(nbs), dim3(256), 0, b.stream, b.U, b.Umort, b.s2e, b.nvse,
                     b.mort, b.LiftS, b.nMortar, eq);
  hipLaunchKernelGGL((k_flux_mortar<N, 3 * EQ::NV>), dim3(nbb), dim3(256), 0, b.stream, b.LiftS, b.LiftM, b.mort,
                     b.mops, b.nMortar);
  if (b.cfg.lifting == 2) {
    launch_lifting_l<N, EQ, 2>(b, out, eq, dtf);
    if (out == 1)
      hipLaunchKernelGGL((k_mortar_fv<N, EQ, true>), dim3(nbs), dim3(256), 0, b.stream, b.Umort, b.GradM, b.nvse,
                         b.mort, b.mops, b.FvN, b.nMortar, eq);
  } else {
    launch_lifting_l<N, EQ, 1>(b, out, eq, dtf);
    if (out == 1)
      hipLaunchKernelGGL((k_mortar_fv<N, EQ, false>), dim3(nbs), dim3(256), 0, b.stream, b.Umort, b.GradM, b.nvse,
                         b.mort, b.mops, b.FvN, b.nMortar, eq);
  }
}

template <int N, class EQ>
void launch_lifting_e(const Batch &b, int out, const EQ &eq, const DtFuse *dtf)
{
  if (b.cfg.lifting == 2) launch_lifting_l<N, EQ, 2>(b, out, eq, dtf);
  else launch_lifting_l<N, EQ, 1>(b, out, eq, dtf);
}

template <int N>
void launch_lifting_n(const Batch &b, int out, const DtFuse *dtf)
{
  if (b.mhd()) {
    MhdLift eq;
    eq.kappa = b.cfg.kappa;
    eq.smu0 = 1.0 / b.cfg.mu_0;
    eq.mu = b.cfg.mu0;
    eq.etasmu0 = b.cfg.eta * eq.smu0;
    eq.KappasPr = b.cfg.kappa / b.cfg.Pr;
    if (b.nMortar > 0) launch_lifting_mortar<N>(b, out, eq, dtf);
    else launch_lifting_e<N>(b, out, eq, dtf);
  } else {
    NsLift eq;
    eq.kappa = b.cfg.kappa;
    eq.mu0 = b.cfg.mu0;
    eq.KappasPr = b.cfg.kappa / b.cfg.Pr;
    if (b.nMortar > 0) launch_lifting_mortar<N>(b, out, eq, dtf);
    else launch_lifting_e<N>(b, out, eq, dtf);
  }
}

template <int N>
void launch_dt_n(const Batch &b, double cfl, double dfl)
{
  const int epb = DtGeom<N>::EPB;
  const int nb = (b.cfg.nElems + epb - 1) / epb;
  if (nb <= 0) return;
  const double KsPr_max = ks_pr_max(b.cfg);
  hipLaunchKernelGGL((k_calc_dt<N>), dim3(nb), dim3(DtGeom<N>::T), 0, b.stream, b.U, b.Mf, b.Mg, b.Mh, b.sJ,
                     b.dtElem, b.errElem, b.eprob, b.cfg.nElems, b.cfg.kappa, cfl, dfl, b.cfg.mu0, KsPr_max,
                     b.cfg.parabolic);
}

template <int N, int MODE, bool PARA>
void launch_mhd_volume_p(const Batch &b, RKArgs rk)
{
  constexpr int EPB = MhdVolGeom<N>::EPB, T = MhdVolGeom<N>::T;
  const int nb = (b.cfg.nElems + EPB - 1) / EPB;
  if (nb <= 0) return;
  if (b.cfg.precision == FX_PREC_FP32_SPLIT) {  // the line form (faster in FP32)
    constexpr int EPBL = MhdLineGeom<N>::EPB, WL = MhdLineGeom<N>::W;
    const int nbl = (b.cfg.nElems + EPBL - 1) / EPBL;
    hipLaunchKernelGGL((k_mhd_volume_line<N, MODE, float, PARA>), dim3(nbl), dim3(3 * WL), 0, b.stream, b.U,
                       b.Mf32, b.Mg32, b.Mh32, b.sJ, b.e2s, b.Flux, b.FluxS, b.Uv, b.Ut, b.Uttmp, rk,
                       b.cfg.nElems, b.mhd_args(), b.st);
  } else
    hipLaunchKernelGGL((k_mhd_volume<N, MODE, double, PARA>), dim3(nb), dim3(T), 0, b.stream, b.U, b.Mf, b.Mg,
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
                       b.Flux, b.FluxS, b.cfg.nSides, b.mhd_args(), b.st, b.Umort);
  else
    hipLaunchKernelGGL((k_mhd_surface<N, RS, false>), dim3(nb), dim3(256), 0, b.stream, b.U, b.FvN, b.s2e, b.nvse,
                       b.Flux, b.FluxS, b.cfg.nSides, b.mhd_args(), b.st, b.Umort);
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


template <int N>
void launch_surfint_n(const Batch &b, const double *Fm, const double *Fs)
{
  const long long nn = (long long)b.n3 * b.cfg.nElems;
  const int nb = (int)((nn + 255) / 256);
  if (nb <= 0) return;
  if (b.mhd())
    hipLaunchKernelGGL((k_surfint<N, 9>), dim3(nb), dim3(256), 0, b.stream, Fm, Fs, b.e2s, b.Ut, b.cfg.nElems);
  else
    hipLaunchKernelGGL((k_surfint<N, 5>), dim3(nb), dim3(256), 0, b.stream, Fm, Fs, b.e2s, b.Ut, b.cfg.nElems);
}

template <int N>
void launch_entropy_n(const Batch &b, double *part, double *red)
{
  const long long nn = (long long)b.n3 * b.cfg.nElems;
  const int nb = (int)((nn + 255) / 256);
  hipLaunchKernelGGL((k_entropy<N>), dim3(nb), dim3(256), 0, b.stream, b.U, b.Ut, b.sJ, part, nn, b.cfg.kappa);
  hipLaunchKernelGGL((k_sum_pairs<N>), dim3(1), dim3(256), 0, b.stream, part, nb, red);
}

template <int N, int NV>
void launch_analyze_nv(const Batch &b, const AnaArgs &a, const double *xgp, const double *ana_ops, double *ana_part,
                       double *ana_out)
{
  constexpr int n = N + 1;
  const size_t lds = (size_t)(NV + 4) * a.na * n * n * 8;
  const double *g = b.cfg.parabolic ? b.grad : nullptr;
  const size_t nv = (size_t)NV * b.n3 * b.cfg.nElems;
  hipLaunchKernelGGL((k_analyze_nodes<N, NV>), dim3(b.cfg.nElems), dim3(256), 0, b.stream, b.U, b.Ut, b.sJ, xgp, g,
                     g ? g + nv : nullptr, g ? g + 2 * nv : nullptr, a, ana_part);
  if (a.exactFunc > 0)
    hipLaunchKernelGGL((k_analyze_points<N, NV>), dim3(b.cfg.nElems), dim3(256), lds, b.stream, b.U, b.sJ, xgp,
                       ana_ops, ana_ops + ANA_MAXNA * (MAXN + 1), a, ana_part);
  else  // no error norms: zero the L2 and LINF columns
    hipLaunchKernelGGL((k_analyze_zero_norms<N, NV>), dim3((b.cfg.nElems + 255) / 256), dim3(256), 0, b.stream,
                       ana_part, b.cfg.nElems);
  hipLaunchKernelGGL((k_analyze_reduce<N, NV>), dim3(AnaRec<NV>::COUNT), dim3(256), 0, b.stream, ana_part,
                     b.cfg.nElems, ana_out);
}

template <int N>
void launch_analyze_n(const Batch &b, const AnaArgs &a, const double *xgp, const double *ana_ops, double *ana_part,
                      double *ana_out)
{
  if (b.mhd()) launch_analyze_nv<N, 9>(b, a, xgp, ana_ops, ana_part, ana_out);
  else launch_analyze_nv<N, 5>(b, a, xgp, ana_ops, ana_part, ana_out);
}

}  // namespace fx
