// Analysis of a state on the device (fx_analyze in include/fluxo_gpu.h):
//   CalcErrorNorms (src/analyze/analyze.f90:557-628) -- L2 and Linf errors on NAnalyze+1
//     Gauss-Lobatto points per direction (U, the coordinates and J = 1/sJ interpolated with
//     the barycentric Vandermonde of InitAnalyzeBasis :267-296) and the L2 error on the
//     collocation points, against AnalyzeExactFunc at time t;
//   the domain volume (InitAnalyze :213-221) and CalcBulk (:633-675);
//   the Taylor-Green vortex testcase quantities (src/testcase/taylorgreenvortex/
//     testcase_analyze.f90:127-401) from U and the lifted primitive gradients.
// Two kernels with one workgroup per element fill disjoint columns of a per-element record of
// partial sums and maxima (k_analyze_nodes: collocation points; k_analyze_points: the
// NAnalyze points, LDS (nVar+4)(NAnalyze+1)(N+1)^2 doubles); k_analyze_reduce
// combines the records column by column in a fixed order; the host normalises.
//
// Compiled as part of fx_engine.hip, after fx_launch.hpp.
#pragma once
#include "fx_launch.hpp"

namespace fx {

constexpr int ANA_MAXNA = 17;  // NAnalyze <= 16

struct AnaArgs {
  int exactFunc;    // 0 none, 1 constant ref, 2 density wave
  int na;           // NAnalyze + 1
  double ref[9], adv[3], wave[3], amp, t, kappa, s2mu0, R;
};

// record layout per element: volume, six per-variable blocks, 11 testcase sums, 2 maxima
template <int NV>
struct AnaRec {
  static constexpr int VOL = 0, L2 = 1, L2C = 1 + NV, LINF = 1 + 2 * NV, BULK = 1 + 3 * NV, BULKT = 1 + 4 * NV,
                       MAXUT = 1 + 5 * NV, TC = 1 + 6 * NV, NSUM_TC = 11, MAXV = TC + 11, MAXC = TC + 12,
                       COUNT = TC + 13;
  __host__ __device__ static constexpr bool is_max(int c)
  {
    return (c >= LINF && c < LINF + NV) || (c >= MAXUT && c < MAXUT + NV) || c == MAXV || c == MAXC;
  }
};

// ExactFunc 1 (RefState) and 2 (density wave, navierstokes/equation.f90:632-651)
template <int NV>
__device__ __forceinline__ void ana_exact(const AnaArgs &a, const double *x, double *u)
{
  if (a.exactFunc == 1) {
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = a.ref[v];
    return;
  }
  const double arg = a.wave[0] * (x[0] - a.adv[0] * a.t) + a.wave[1] * (x[1] - a.adv[1] * a.t) +
                     a.wave[2] * (x[2] - a.adv[2] * a.t);
  const double rho = 1.0 * (1.0 + a.amp * sin(M_PI * arg));
  u[0] = rho;
  u[1] = rho * a.adv[0];
  u[2] = rho * a.adv[1];
  u[3] = rho * a.adv[2];
  u[4] = 1.0 / (a.kappa - 1.0) + 0.5 * rho * (a.adv[0] * a.adv[0] + a.adv[1] * a.adv[1] + a.adv[2] * a.adv[2]);
#pragma unroll
  for (int v = 5; v < NV; v++) u[v] = 0.0;
}

// wavefront reduction (sum or max) of one value; lane 0 holds the result
__device__ __forceinline__ double wave_reduce(double v, bool mx)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_down(v, o, 64);
    v = mx ? fmax(v, w) : v + w;
  }
  return v;
}

// Collocation points of one element per workgroup: volume, L2 on the nodes, bulk, max|Ut| and
// the testcase integrands (columns VOL .. MAXC except L2 and LINF).
template <int N, int NV>
__global__ void __launch_bounds__(256)
k_analyze_nodes(const double *__restrict__ U, const double *__restrict__ Ut, const double *__restrict__ sJ,
                const double *__restrict__ xgp, const double *__restrict__ gx, const double *__restrict__ gy,
                const double *__restrict__ gz, AnaArgs a, double *__restrict__ part)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  using Rc = AnaRec<NV>;
  // columns handled here, packed: VOL, L2C[NV], BULK[NV], BULKT[NV], MAXUT[NV], TC[13]
  constexpr int P_L2C = 1, P_BULK = 1 + NV, P_BULKT = 1 + 2 * NV, P_MAXUT = 1 + 3 * NV, P_TC = 1 + 4 * NV,
                CNT = P_TC + 13;
  __shared__ double s_red[4 * CNT];
  const Ops &op = c_ops[N];
  const int e = blockIdx.x, tid = threadIdx.x;
  double acc[CNT];
#pragma unroll
  for (int c = 0; c < CNT; c++) acc[c] = 0.0;
  if (tid < n3) {
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    const size_t g = (size_t)e * n3 + tid;
    double u[NV], ut[NV], x[3];
#pragma unroll
    for (int v = 0; v < NV; v++) { u[v] = U[NV * g + v]; ut[v] = Ut[NV * g + v]; }
#pragma unroll
    for (int d = 0; d < 3; d++) x[d] = xgp ? xgp[3 * g + d] : 0.0;
    const double w = op.w[i] * op.w[j] * op.w[k] / sJ[g];
    acc[0] = w;
    if (a.exactFunc > 0) {
      double ue[NV];
      ana_exact<NV>(a, x, ue);
#pragma unroll
      for (int v = 0; v < NV; v++) acc[P_L2C + v] = (u[v] - ue[v]) * (u[v] - ue[v]) * w;
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      acc[P_BULK + v] = u[v] * w;
      acc[P_BULKT + v] = ut[v] * w;
      acc[P_MAXUT + v] = fabs(ut[v]);
    }
    const double rho = u[0], srho = 1.0 / rho, km1 = a.kappa - 1.0;
    const double vel[3] = {u[1] * srho, u[2] * srho, u[3] * srho};
    double G[3][3], GB[3][3];
    const double *gq[3] = {gx, gy, gz};
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int r = 0; r < 3; r++) {
        G[r][q] = gq[q] ? gq[q][NV * g + 1 + r] : 0.0;
        GB[r][q] = (NV == 9 && gq[q]) ? gq[q][NV * g + (5 + r) % NV] : 0.0;
      }
    const double B0 = NV == 9 ? u[5 % NV] : 0.0, B1 = NV == 9 ? u[6 % NV] : 0.0, B2 = NV == 9 ? u[7 % NV] : 0.0;
    const double psi = NV == 9 ? u[8 % NV] : 0.0;
    const double BB = B0 * B0 + B1 * B1 + B2 * B2;
    const double pres =
        km1 * (u[4] - 0.5 * (u[1] * vel[0] + u[2] * vel[1] + u[3] * vel[2]) - a.s2mu0 * (BB + psi * psi));
    const double divU = G[0][0] + G[1][1] + G[2][2];
    double ut2 = 0.0, st2 = 0.0, sd2 = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const double sv = 0.5 * (G[r][q] + G[q][r]), sd = r == q ? sv - 1.0 / 3.0 * divU : sv;
        ut2 += G[r][q] * G[r][q];
        st2 += sv * sv;
        sd2 += sd * sd;
      }
    const double kE = 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double mag = a.s2mu0 * BB;
    const double o0 = G[2][1] - G[1][2], o1 = G[0][2] - G[2][0], o2 = G[1][0] - G[0][1];
    const double c0 = GB[2][1] - GB[1][2], c1 = GB[0][2] - GB[2][0], c2 = GB[1][0] - GB[0][1];
    const double om2 = o0 * o0 + o1 * o1 + o2 * o2, cj2 = c0 * c0 + c1 * c1 + c2 * c2;
    acc[P_TC + 0] = kE * w;
    acc[P_TC + 1] = rho * kE * w;
    acc[P_TC + 2] = mag * w;
    acc[P_TC + 3] = 0.5 * rho * om2 * w;
    acc[P_TC + 4] = a.s2mu0 * cj2 * w;
    acc[P_TC + 5] = -rho * (log(pres) - a.kappa * log(rho)) / km1 * w;
    acc[P_TC + 6] = ut2 * w;
    acc[P_TC + 7] = st2 * w;
    acc[P_TC + 8] = sd2 * w;
    acc[P_TC + 9] = pres * divU * w;
    acc[P_TC + 10] = km1 / a.R * (u[4] * srho - kE - mag * srho - a.s2mu0 * psi * psi) * w;
    acc[P_TC + 11] = sqrt(om2);
    acc[P_TC + 12] = sqrt(cj2);
  }
  // store: VOL | L2C | (L2..LINF skipped) | BULK BULKT MAXUT TC are contiguous in the record
  static_assert(Rc::L2C == 1 + NV && Rc::BULK == 1 + 3 * NV && Rc::COUNT == 1 + 6 * NV + 13, "record layout");
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int c = 0; c < CNT; c++) {
    const int col = c == 0 ? Rc::VOL : (c < P_BULK ? Rc::L2C + c - P_L2C : Rc::BULK + c - P_BULK);
    const double r = wave_reduce(acc[c], Rc::is_max(col));
    if (lane == 0) s_red[wv * CNT + c] = r;
  }
  __syncthreads();
  if (tid < CNT) {
    const int col = tid == 0 ? Rc::VOL : (tid < P_BULK ? Rc::L2C + tid - P_L2C : Rc::BULK + tid - P_BULK);
    const bool mx = Rc::is_max(col);
    double r = s_red[tid];
#pragma unroll
    for (int q = 1; q < 4; q++) r = mx ? fmax(r, s_red[q * CNT + tid]) : r + s_red[q * CNT + tid];
    part[(size_t)e * Rc::COUNT + col] = r;
  }
}

// Error norms on the NAnalyze points of one element per workgroup (columns L2, LINF): the
// element's U, coordinates and J = 1/sJ interpolated along xi into LDS (from the nodes, in
// the caches after k_analyze_nodes), then along eta and zeta per point.
template <int N, int NV>
__global__ void __launch_bounds__(256)
k_analyze_points(const double *__restrict__ U, const double *__restrict__ sJ, const double *__restrict__ xgp,
                 const double *__restrict__ vdm, const double *__restrict__ wA, AnaArgs a, double *__restrict__ part)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n, NF = NV + 4;  // fields: U, x, y, z, J
  using Rc = AnaRec<NV>;
  extern __shared__ double s_t[];  // [NF][na n2]
  __shared__ double s_red[4 * 2 * NV];
  const int e = blockIdx.x, tid = threadIdx.x, na = a.na, nt = na * n2;
  const size_t g0 = (size_t)e * n3;
  for (int idx = tid; idx < NF * nt; idx += 256) {
    const int f = idx / nt, r = idx - f * nt, a1 = r % na, jk = r / na;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < n; i++) {
      const size_t g = g0 + i + n * jk;
      const double v = f < NV ? U[NV * g + f] : (f < NV + 3 ? xgp[3 * g + f - NV] : 1.0 / sJ[g]);
      s += vdm[a1 * n + i] * v;
    }
    s_t[f * nt + a1 + na * jk] = s;
  }
  __syncthreads();
  double acc[2 * NV];
#pragma unroll
  for (int c = 0; c < 2 * NV; c++) acc[c] = 0.0;
  const int na3 = na * na * na;
  for (int pt = tid; pt < na3; pt += 256) {
    const int a1 = pt % na, a2 = (pt / na) % na, a3 = pt / (na * na);
    double val[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const double *src = s_t + f * nt + a1;
      double s = 0.0;
#pragma unroll 1
      for (int k = 0; k < n; k++) {
        double t2 = 0.0;
#pragma unroll
        for (int j = 0; j < n; j++) t2 += vdm[a2 * n + j] * src[na * (j + n * k)];
        s += vdm[a3 * n + k] * t2;
      }
      val[f] = s;
    }
    double ue[NV];
    ana_exact<NV>(a, val + NV, ue);
    const double w = wA[a1] * wA[a2] * wA[a3] * val[NV + 3];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double d = val[v] - ue[v];
      acc[v] += d * d * w;
      acc[NV + v] = fmax(acc[NV + v], fabs(d));
    }
  }
  static_assert(Rc::LINF == Rc::L2 + 2 * NV, "record layout");
  // columns L2 .. L2+NV-1 are sums; LINF .. are maxima: pack (L2, LINF) as acc[0..2NV)
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int c = 0; c < 2 * NV; c++) {
    const double r = wave_reduce(acc[c], c >= NV);
    if (lane == 0) s_red[wv * 2 * NV + c] = r;
  }
  __syncthreads();
  if (tid < 2 * NV) {
    const bool mx = tid >= NV;
    double r = s_red[tid];
#pragma unroll
    for (int q = 1; q < 4; q++) r = mx ? fmax(r, s_red[q * 2 * NV + tid]) : r + s_red[q * 2 * NV + tid];
    part[(size_t)e * Rc::COUNT + (mx ? Rc::LINF + tid - NV : Rc::L2 + tid)] = r;
  }
}

// (these two are also templated on the caller's degree N, so that each compilation unit
// instantiates its own)
template <int N, int NV>
__global__ void k_analyze_zero_norms(double *__restrict__ part, int nElems)
{
  using Rc = AnaRec<NV>;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nElems) return;
  for (int c = Rc::L2; c < Rc::LINF + NV; c++) part[(size_t)e * Rc::COUNT + c] = 0.0;
}

// column c of the element records: sum or max in a fixed order (one workgroup per column)
template <int N, int NV>
__global__ void __launch_bounds__(256) k_analyze_reduce(const double *__restrict__ part, int nElems, double *__restrict__ out)
{
  using Rc = AnaRec<NV>;
  __shared__ double s[256];
  const int c = blockIdx.x, tid = threadIdx.x;
  const bool mx = Rc::is_max(c);
  double r = 0.0;
  for (int e = tid; e < nElems; e += 256) {
    const double v = part[(size_t)e * Rc::COUNT + c];
    r = mx ? fmax(r, v) : r + v;
  }
  s[tid] = r;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) s[tid] = mx ? fmax(s[tid], s[tid + st]) : s[tid] + s[tid + st];
    __syncthreads();
  }
  if (tid == 0) out[c] = s[0];
}

// barycentric Lagrange interpolation matrix from the LGL nodes of N to the Gauss-Lobatto
// nodes of NA (InitializeVandermonde basis.f90 with LagrangeInterpolationPolys), [a][i]
inline void analyze_basis(int N, int NA, double *vdm, double *wA)
{
  double x[MAXN + 1], w[MAXN + 1], wb[MAXN + 1], xa[ANA_MAXNA];
  lgl_nodes(N, x, w);
  lgl_nodes(NA, xa, wA);
  for (int i = 0; i <= N; i++) {
    wb[i] = 1.0;
    for (int j = 0; j <= N; j++)
      if (j != i) wb[i] *= x[i] - x[j];
    wb[i] = 1.0 / wb[i];
  }
  for (int q = 0; q <= NA; q++) {
    int hit = -1;
    for (int i = 0; i <= N; i++)
      if (std::fabs(xa[q] - x[i]) <= 1e-14 * std::fmax(1.0, std::fabs(x[i]))) hit = i;
    if (hit >= 0) {
      for (int i = 0; i <= N; i++) vdm[q * (N + 1) + i] = i == hit ? 1.0 : 0.0;
      continue;
    }
    double den = 0.0;
    for (int i = 0; i <= N; i++) den += wb[i] / (xa[q] - x[i]);
    for (int i = 0; i <= N; i++) vdm[q * (N + 1) + i] = wb[i] / (xa[q] - x[i]) / den;
  }
}

// Analysis kernels of degree N (fx_launch_n.hpp): per-element records into ana_part, their
// reduction into ana_out.
template <int N>
void launch_analyze_n(const Batch &b, const AnaArgs &a, const double *xgp, const double *ana_ops, double *ana_part,
                      double *ana_out);

}  // namespace fx
