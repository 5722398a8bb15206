// NFVSE shock capturing on gfx950: blending coefficient and the sub-cell finite-volume
// volume term (src/shockcapturing/nfvse.f90, nfvse_mpi.f90, src/indicators/indicators.f90).
//
// Per RHS evaluation (before the volume kernel, which blends the FV term into the DG volume
// integral of every element with alpha > 0):
//   k_nfvse_alpha  one group of n^3 threads per element: rho*p per node, modal Legendre
//                  coefficients by three 1D transforms in LDS, the Persson-Peraire energy
//                  ratios, the logistic blending function and the alpha_min/alpha_max clamp
//   k_nfvse_prop   one thread per element: alpha = max(alpha, f max(alpha, alpha_nb)) over
//                  the six faces (one propagation sweep; sweeps alternate two buffers)
//   k_nfvse_fv     element-centric like the volume kernel: one thread per (direction,
//                  element, line) forms the N inner sub-cell faces of its line, their metrics
//                  on the fly (normal = M(0) + sum_m Csub(f,m) M(m), length = its norm), the
//                  numerical flux there, and the per-node differences (f(i)-f(i-1))/w_i;
//                  the three directions are summed in LDS and written as Ffv.
// Sub-cell fluxes are evaluated in the Cartesian frame with the unit normal (the reference
// rotates into (n,t1,t2) and back, an orthogonal change of basis for these fluxes).
#pragma once
#include "fx_kernels.hpp"

namespace fx {

struct NfvseArgs {
  int computeAlpha;      // 1 Persson-Peraire indicator, 3 fixed shockBlendCoef
  int pad;
  double threshold;      // 0.5*10^(-1.8 (N+1)^0.25)  (nfvse.f90:208-209)
  double sharpness;      // ln((1-1e-4)/1e-4)          (nfvse_vars.f90:74)
  double alpha_min, alpha_max, shockBlendCoef;
};

// CalcBlendingCoefficient without the propagation (nfvse.f90:1537-1609) with
// PerssonPeraire_Compute (indicators.f90:108-149) on Get_DensityTimesPressure.
template <int N>
__global__ void __launch_bounds__(256)
k_nfvse_alpha(const double *__restrict__ U, const int *__restrict__ eprob, double *__restrict__ alpha,
              int nElems, double kappa, NfvseArgs a)
{
  constexpr int n = N + 1, n3 = n * n * n;
  constexpr int EPB = LiftGeom<N>::EPB;
  __shared__ double s_a[EPB][n3], s_b[EPB][n3];
  const Ops &op = c_ops[N];
  const int tid = threadIdx.x, el = tid / n3, node = tid - el * n3;
  const int e = blockIdx.x * EPB + el;
  const bool active = el < EPB && e < nElems && elem_problem(eprob, e) >= 0;
  const int i = node % n, j = (node / n) % n, k = node / (n * n);
  if (a.computeAlpha == 1) {
    if (active) {
      const double *u = U + ((size_t)e * n3 + node) * 5;
      const double p = (kappa - 1.0) * (u[4] - 0.5 * ((u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]));
      s_a[el][node] = u[0] * p;
    }
    __syncthreads();
    // ChangeBasis3D: xi, then eta, then zeta
    if (active) {
      double t = 0.0;
#pragma unroll
      for (int l = 0; l < n; l++) t += op.sVdm[i][l] * s_a[el][l + n * (j + n * k)];
      s_b[el][node] = t;
    }
    __syncthreads();
    if (active) {
      double t = 0.0;
#pragma unroll
      for (int l = 0; l < n; l++) t += op.sVdm[j][l] * s_b[el][i + n * (l + n * k)];
      s_a[el][node] = t;
    }
    __syncthreads();
    if (active) {
      double t = 0.0;
#pragma unroll
      for (int l = 0; l < n; l++) t += op.sVdm[k][l] * s_a[el][i + n * (j + n * l)];
      s_b[el][node] = t * t;
    }
    __syncthreads();
  }
  if (active && node == 0) {
    double al;
    if (a.computeAlpha == 1) {
      double LU = 0.0, LUM1 = 0.0, LUM2 = 0.0;
      for (int kk = 0; kk < n; kk++)
        for (int jj = 0; jj < n; jj++)
          for (int ii = 0; ii < n; ii++) {
            const double m2 = s_b[el][ii + n * (jj + n * kk)];
            LU += m2;
            if (ii < N && jj < N && kk < N) LUM1 += m2;
            if (ii < N - 1 && jj < N - 1 && kk < N - 1) LUM2 += m2;
          }
      double eta = fmax((LU - LUM1) / LU, (LUM1 - LUM2) / LUM1);
      if (eta < 2.220446049250313e-16) eta = 2.220446049250313e-16;
      al = 1.0 / (1.0 + exp(-a.sharpness * (eta - a.threshold) / a.threshold));
    } else {
      al = a.shockBlendCoef;
    }
    if (al < a.alpha_min) al = 0.0;
    else if (al >= a.alpha_max) al = a.alpha_max;
    alpha[e] = al;
  }
}

#ifdef FX_KERNEL_HOME  // non-template kernel: compiled in one unit only
// One sweep of PropagateBlendingCoeff (nfvse_mpi.f90:78-153): every side raises both its
// elements to factor * max(alpha_master, alpha_slave), from the values before the sweep, so
// the order of the sides does not matter.  Across a mortar (Alpha_Mortar nfvse_mpi.f90:
// 192-240 gives each small side the big element's alpha as its master value) the big element
// takes the maximum over its small sides and each small element the big element's value;
// mside maps a small side to its mortar (-1 elsewhere), mort describes the mortars.
__global__ void __launch_bounds__(256)
k_nfvse_prop(const double *__restrict__ ain, double *__restrict__ aout, const int *__restrict__ e2s,
             const int *__restrict__ s2e, const int *__restrict__ eprob, int nElems, double factor,
             const int *__restrict__ mort, const int *__restrict__ mside, int nMortar)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nElems || elem_problem(eprob, e) < 0) return;
  const double ae = ain[e];
  double r = ae;
  for (int l = 0; l < 6; l++) {
    const int s = e2s[(e * 6 + l) * 2] - 1, fl = e2s[(e * 6 + l) * 2 + 1];
    if (fl == 0 && s < nMortar) {  // big mortar side: its small sides' elements
      const int *mi = mort + MORT_INTS * s;
      const int ns = mi[0] == 1 ? 4 : 2;
      for (int im = 0; im < ns; im++) {
        const double m = factor * fmax(ae, ain[s2e[5 * (mi[3 + im] - 1) + 1] - 1]);
        if (r < m) r = m;
      }
      continue;
    }
    int nb = (fl == 0 ? s2e[5 * s + 1] : s2e[5 * s]) - 1;
    if (nb < 0) nb = mort[MORT_INTS * mside[s] + 1];  // small side: the big element
    const double m = factor * fmax(ae, ain[nb]);
    if (r < m) r = m;
  }
  aout[e] = r;
}
#endif  // FX_KERNEL_HOME

// Compute_FVFluxes_1st_Order + Compute_FVFluxes_1D + the F_FV sum of VolInt_NFVSE
// (nfvse.f90:536-548, 566-670, 1389-1443) for elements with alpha > 0.
template <int N> struct FvGeom { static constexpr int EPB = VolGeom<N, VolTiles<N>::T[0]>::EPB,
                                                      W = VolGeom<N, VolTiles<N>::T[0]>::W; };

template <int N, int RS>
__global__ void __launch_bounds__(3 * FvGeom<N>::W)
k_nfvse_fv(const double *__restrict__ U, const double *__restrict__ Mf, const double *__restrict__ Mg,
           const double *__restrict__ Mh, const double *__restrict__ alpha, const int *__restrict__ eprob,
           double *__restrict__ Ffv, int nElems, double kappa)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr int EPB = FvGeom<N>::EPB, W = FvGeom<N>::W;
  __shared__ double s_f[EPB][5][n3];
  const Ops &op = c_ops[N];
  const int tid = threadIdx.x, e0 = blockIdx.x * EPB;
  const int dir = tid / W, lt = tid - dir * W;
  const int el = lt / n2, line = lt - el * n2;
  const int e = e0 + el;
  const bool active = el < EPB && e < nElems && elem_problem(eprob, e) >= 0 && alpha[e] != 0.0;
  const int a = line % n, b = line / n;
  double acc[n][5];
#pragma unroll
  for (int i = 0; i < n; i++)
#pragma unroll
    for (int v = 0; v < 5; v++) acc[i][v] = 0.0;
  int nodes[n];
#pragma unroll
  for (int l = 0; l < n; l++) {
    const int i = (dir == 0) ? l : a, j = (dir == 1) ? l : (dir == 0 ? a : b), k = (dir == 2) ? l : b;
    nodes[l] = i + n * (j + n * k);
  }
  if (active) {
    const double *M = (dir == 0) ? Mf : ((dir == 1) ? Mg : Mh);
    double m[n][3];
#pragma unroll
    for (int l = 0; l < n; l++) {
      const double *mp = M + ((size_t)e * n3 + nodes[l]) * 3;
      m[l][0] = mp[0]; m[l][1] = mp[1]; m[l][2] = mp[2];
    }
    double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double *u0 = U + ((size_t)e * n3 + nodes[0]) * 5;
    double UL[5] = {u0[0], u0[1], u0[2], u0[3], u0[4]};
#pragma unroll
    for (int f = 0; f < N; f++) {
      double nv[3] = {m[0][0], m[0][1], m[0][2]};
#pragma unroll
      for (int mm = 0; mm < n; mm++) {
        const double c = op.Csub[f][mm];
        nv[0] = nv[0] + c * m[mm][0];
        nv[1] = nv[1] + c * m[mm][1];
        nv[2] = nv[2] + c * m[mm][2];
      }
      const double norm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      nv[0] = nv[0] / norm; nv[1] = nv[1] / norm; nv[2] = nv[2] / norm;
      const double *ur = U + ((size_t)e * n3 + nodes[f + 1]) * 5;
      double UR[5] = {ur[0], ur[1], ur[2], ur[3], ur[4]}, F[5];
      if constexpr (RS == 4 || RS == 5 || RS == 33 || RS == 44) riemann_es<RS>(UL, UR, nv, kappa, 0.0, F);
      else riemann<RS>(UL, UR, nv, kappa, F);
#pragma unroll
      for (int v = 0; v < 5; v++) {
        F[v] = F[v] * norm;
        acc[f][v] = op.sw[f] * (F[v] - fprev[v]);
        fprev[v] = F[v];
        UL[v] = UR[v];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) acc[N][v] = op.sw[N] * (0.0 - fprev[v]);
  }
  // xi + eta + zeta, in this order, into LDS
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (active && dir == d) {
#pragma unroll
      for (int l = 0; l < n; l++)
#pragma unroll
        for (int v = 0; v < 5; v++)
          s_f[el][v][nodes[l]] = (d == 0) ? acc[l][v] : s_f[el][v][nodes[l]] + acc[l][v];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < EPB * n3; idx += 3 * W) {
    const int el2 = idx / n3, node = idx - el2 * n3, e2 = e0 + el2;
    if (e2 >= nElems || elem_problem(eprob, e2) < 0 || alpha[e2] == 0.0) continue;
    double *o = Ffv + ((size_t)e2 * n3 + node) * 5;
#pragma unroll
    for (int v = 0; v < 5; v++) o[v] = s_f[el2][v][node];
  }
}

}  // namespace fx
