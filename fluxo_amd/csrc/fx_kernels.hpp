// HIP kernels of the split-form DGSEM right-hand side for gfx950 (CDNA4).
//
// Data in HBM keeps the reference's layout (variables fastest, then i, j, k, element):
//   U, Ut, Ut_tmp : nVar x (N+1)^3 x nElems doubles
//   Mf, Mg, Mh    : 3 x (N+1)^3 x nElems       sJ : (N+1)^3 x nElems
//   nvse          : per side face point (nx, ny, nz, SurfElem)
//   Flux          : nVar x (N+1)^2 x nSides (Flux_master * SurfElem; slave uses -Flux)
//   Uv            : NVV x (N+1)^3 x nElems, weak-form viscous volume term (components 1..NVV)
//   FvN           : NVV x 2 x (N+1)^2 x nSides, each side's face-normal viscous flux
//
// Kernel roles per RHS (one LSERK stage):
//   k_lifting  element-centric BR1 gradients -> viscous fluxes -> Uv, FvN (Navier-Stokes,
//              resistive GLM-MHD)
//   k_surface  side-centric numerical flux at every face point  -> Flux
//   k_volume   element-centric flux differencing + viscous volume term + surface
//              gather + Jacobian + (optionally) the low-storage RK stage update.
#pragma once
#include <hip/hip_runtime.h>
#include "fx_physics.hpp"

namespace fx {

constexpr int MAXN = 7;  // highest supported polynomial degree

// Basis operators for one N, in constant memory (uniform loads go through the scalar cache).
struct Ops {
  double Ds[MAXN + 1][MAXN + 1];    // Dvolsurf(i,l)  (dg.f90:241-249)
  double Dhat[MAXN + 1][MAXN + 1];  // D_Hat(i,l)     (dg.f90:214-216)
  double D[MAXN + 1][MAXN + 1];     // D(i,l)         (dg.f90:210)
  double w[MAXN + 1];               // LGL weights
  double L0;                        // L_HatMinus0 = L_HatMinus(0) (dg.f90:228-229)
  float Ds32[MAXN + 1][MAXN + 1];   // FP32 copies for the split-precision volume integral
  float Dhat32[MAXN + 1][MAXN + 1];
  // NFVSE shock capturing
  double DsRow[MAXN + 1];           // sum_l Dvolsurf(i,l) (non-conservative MHD terms)
  double sVdm[MAXN + 1][MAXN + 1];  // nodal -> Legendre modal (BuildLegendreVdm basis.f90:196-252)
  double Csub[MAXN + 1][MAXN + 1];  // Csub(f,m) = sum_{l<=f} w_l D(l,m): sub-cell face metrics
  double sw[MAXN + 1];              // 1/w_i (sWGP, nfvse.f90:219)
};

template <typename VR> struct OpsOf;
template <> struct OpsOf<double> {
  __device__ static __forceinline__ double ds(const Ops &o, int i, int l) { return o.Ds[i][l]; }
  __device__ static __forceinline__ double dhat(const Ops &o, int i, int l) { return o.Dhat[i][l]; }
};
template <> struct OpsOf<float> {
  __device__ static __forceinline__ float ds(const Ops &o, int i, int l) { return o.Ds32[i][l]; }
  __device__ static __forceinline__ float dhat(const Ops &o, int i, int l) { return o.Dhat32[i][l]; }
};
// One copy per compilation unit (internal linkage): the library is built from several units
// that each hold kernels, and every unit's copy is filled by ensure_ops (fx_launch.hpp).
static __constant__ Ops c_ops[MAXN + 1];

// local side ids (defines.h:68-73)
enum { ZETA_MINUS = 1, ETA_MINUS = 2, XI_PLUS = 3, ETA_PLUS = 4, XI_MINUS = 5, ZETA_PLUS = 6 };

// Flip_S2M (mappings.f90:262-284); an involution, so it also maps master->slave.
__device__ __forceinline__ void flip_pq(int N, int flip, int &p, int &q)
{
  int a = p, b = q;
  switch (flip) {
  case 1: p = b; q = a; break;
  case 2: p = N - a; q = b; break;
  case 3: p = N - b; q = N - a; break;
  case 4: p = a; q = N - b; break;
  default: break;
  }
}

// Volume node (i + n*(j + n*k)) of side point (p,q) given in the master frame:
// SideToVol(N,0,p,q,flip,locSide) (mappings.f90:459-472).
template <int N>
__device__ __forceinline__ int side_to_vol(int loc, int flip, int p, int q)
{
  constexpr int n = N + 1;
  flip_pq(N, flip, p, q);
  int i, j, k;
  switch (loc) {
  case XI_MINUS: i = 0; j = q; k = p; break;
  case XI_PLUS: i = N; j = p; k = q; break;
  case ETA_MINUS: i = p; j = 0; k = q; break;
  case ETA_PLUS: i = N - p; j = N; k = q; break;
  case ZETA_MINUS: i = q; j = p; k = 0; break;
  default: i = p; j = q; k = N; break;
  }
  return i + n * (j + n * k);
}

// Master-frame side point of volume node (i,j,k): VolToSide (mappings.f90:412-426).
template <int N>
__device__ __forceinline__ void vol_to_side(int loc, int flip, int i, int j, int k, int &p, int &q)
{
  switch (loc) {
  case XI_MINUS: p = k; q = j; break;
  case XI_PLUS: p = j; q = k; break;
  case ETA_MINUS: p = i; q = k; break;
  case ETA_PLUS: p = N - i; q = k; break;
  case ZETA_MINUS: p = j; q = i; break;
  default: p = i; q = j; break;
  }
  flip_pq(N, flip, p, q);
}

// Node (i + n(j + n k)) at element-local face coordinates (a, b) of local side loc: xi sides
// (a, b) = (j, k), eta sides (i, k), zeta sides (i, j).
template <int N>
__device__ __forceinline__ int face_node(int loc, int a, int b)
{
  constexpr int n = N + 1;
  switch (loc) {
  case XI_MINUS: return 0 + n * (a + n * b);
  case XI_PLUS: return N + n * (a + n * b);
  case ETA_MINUS: return a + n * (0 + n * b);
  case ETA_PLUS: return a + n * (N + n * b);
  case ZETA_MINUS: return a + n * (b + n * 0);
  default: return a + n * (b + n * N);
  }
}

// The faces of node (i,j,k) in the reference's order of SurfInt (ZETA_MINUS, ETA_MINUS,
// XI_PLUS, ETA_PLUS, XI_MINUS, ZETA_PLUS; at most 3), each as the face-point slot
// (locSide - 1) n^2 + a + n b of the element, -1 where there is none.  Kept in scalars (a
// dynamically indexed array would live in scratch memory).
struct NodeFaces {
  int r0 = -1, r1 = -1, r2 = -1;
  __device__ __forceinline__ void push(int r)
  {
    if (r0 < 0) r0 = r;
    else if (r1 < 0) r1 = r;
    else r2 = r;
  }
  __device__ __forceinline__ int get(int t) const { return t == 0 ? r0 : (t == 1 ? r1 : r2); }
};

template <int N>
__device__ __forceinline__ NodeFaces node_faces(int i, int j, int k)
{
  constexpr int n = N + 1, n2 = n * n;
  NodeFaces f;
  if (k == 0) f.push((ZETA_MINUS - 1) * n2 + i + n * j);
  if (j == 0) f.push((ETA_MINUS - 1) * n2 + i + n * k);
  if (i == N) f.push((XI_PLUS - 1) * n2 + j + n * k);
  if (j == N) f.push((ETA_PLUS - 1) * n2 + i + n * k);
  if (i == 0) f.push((XI_MINUS - 1) * n2 + j + n * k);
  if (k == N) f.push((ZETA_PLUS - 1) * n2 + i + n * j);
  return f;
}

// Face-point table of the lifting kernel, one int2 per element face point
// r = (locSide - 1) n^2 + a + n b (the element-local face coordinates of face_node), built once
// per mesh from ElemToSide / SideToElem so that the kernel reads the neighbour's state and the
// side point with one coalesced load instead of the connectivity chain and its index arithmetic:
//   .x  the neighbour's volume node at that point, enb n^3 + vnb (SideToVol of the neighbour's
//       side and flip); -1 on a big mortar side (no neighbour element: its lifting flux is
//       projected from the small sides); -2 on the small side of a mortar (its master state is
//       the big side's interpolated state at the side point)
//   .y  2 fpt + (0 master, 1 slave), fpt = (s n + q) n + p the side point in the master frame
// (VolToSide mappings.f90:412-426, SideToVol :459-472).
template <int N>
__global__ void __launch_bounds__(256)
k_face_table(const int *__restrict__ e2s, const int *__restrict__ s2e, int e0, int nE, int nMortar,
             int2 *__restrict__ ftab)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n, NF = 6 * n2;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nE * NF) return;
  const int e = e0 + (int)(g / NF), r = (int)(g % NF), face = r / n2, ab = r - face * n2;
  const int loc = face + 1, fnode = face_node<N>(loc, ab % n, ab / n);
  const int sd = e2s[(e * 6 + face) * 2] - 1, fl = e2s[(e * 6 + face) * 2 + 1];
  int p, q;
  vol_to_side<N>(loc, fl, fnode % n, (fnode / n) % n, fnode / n2, p, q);
  const int fpt = (sd * n + q) * n + p;
  int nb = -1;
  if (!(fl == 0 && sd < nMortar)) {
    const int enb = (fl == 0) ? s2e[5 * sd + 1] - 1 : s2e[5 * sd] - 1;
    if (enb >= 0) {
      const int vnb = (fl == 0) ? side_to_vol<N>(s2e[5 * sd + 3], s2e[5 * sd + 4], p, q)
                                : side_to_vol<N>(s2e[5 * sd + 2], 0, p, q);
      nb = enb * n3 + vnb;
    } else {
      nb = -2;
    }
  }
  ftab[(size_t)e * NF + r] = make_int2(nb, 2 * fpt + (fl == 0 ? 0 : 1));
}

// -----------------------------------------------------------------------------------
// Launch geometry of the volume kernel: EPB elements per block and W = EPB*(N+1)^2
// threads per direction rounded up to a multiple of 32 (3*W threads in the line phase).  Tiles[0]
// is the default EPB of each N; the others are the candidates the volume-kernel tuner
// times (fx_tune_volume) for the LSERK form with the Ranocha flux.
template <int N> struct VolTiles;
template <> struct VolTiles<2> { static constexpr int n = 3, T[3] = {7, 14, 3}; };
template <> struct VolTiles<3> { static constexpr int n = 3, T[3] = {4, 2, 8}; };
template <> struct VolTiles<4> { static constexpr int n = 3, T[3] = {5, 2, 3}; };
template <> struct VolTiles<5> { static constexpr int n = 3, T[3] = {3, 1, 2}; };
// TB threads per block: the 3 W of the line phase, plus one wavefront that sits out the line
// phase where it gives every node of the block its own thread in the node passes (N=3 with
// 4 elements: 256 nodes on 256 threads instead of two rounds on 192).
template <int N, int EPB_> struct VolGeom {
  static constexpr int EPB = EPB_;
  static constexpr int W = ((EPB * (N + 1) * (N + 1) + 31) / 32) * 32;
  static constexpr int NODES = EPB * (N + 1) * (N + 1) * (N + 1);
  static constexpr int TB = 3 * W + ((3 * W < NODES && 3 * W + 64 >= NODES) ? 64 : 0);
};

// LDS node index with the xi-rows padded to n+1 so xi-line owners (stride n) do not
// hit one bank.
template <int N> struct Lds {
  static constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  static constexpr int NP = (n + 1) * n2;  // padded nodes per element
  __device__ static __forceinline__ int li(int i, int j, int k) { return i + (n + 1) * (j + n * k); }
  __device__ static __forceinline__ int li(int node)
  {
    int i = node % n, jk = node / n;
    return i + (n + 1) * jk;
  }
};

enum { VMODE_RHS = 0, VMODE_RK = 1, VMODE_VOLINT = 2 };

// Time-step state of one problem (device resident).
struct DtState {
  int errType;                 // max error code seen
  int errStep;                 // first step with an error (-1 none)
  double dt;                   // dt of the current step
  double t;                    // simulation time
  int step;                    // steps done
  int glm_init;                // GLM-MHD: dtch1 computed (InitTimeStep_GLM done)
  double glm_ch;               // GLM-MHD cleaning speed (0.5 before the first CALCTIMESTEP)
  double glm_dtch1;            // GLM-MHD time step of a unit wave speed
  double tEnd;                 // end time (huge: none); the step reaching it is shortened
  int done;                    // tEnd reached: later steps have dt = 0 and do not advance
  int pad;
};

// The reference's end-of-run time step (timedisc.f90:313-317): when the CFL time step dt
// reaches tEnd within 1e-4 of dt, the step is shortened to end exactly at tEnd and it is the
// last one.  Returns the step's dt and sets S.done; once done, dt = 0 (the stage updates
// leave U unchanged) and time and step count stay where they are.
__device__ __forceinline__ double end_time_step(DtState &S, double dt, int advance)
{
  if (S.done) return 0.0;
  if (S.tEnd - S.t <= dt * (1.0 + 1.0e-4)) {
    dt = S.tEnd - S.t;
    if (advance) S.done = 1;
  }
  return dt;
}

// Element -> problem map of a batch of independent meshes sharing one launch: eprob[e] is
// the problem (index into the per-problem time-step states) of element e, or -1 for an
// unused slot, which every kernel skips.  nullptr means one problem owning all elements.
__device__ __forceinline__ int elem_problem(const int *eprob, int e) { return eprob ? eprob[e] : 0; }

struct RKArgs {
  const DtState *st;  // per-problem time-step states (dt of the current step)
  const int *eprob;   // element -> problem (nullptr: one problem)
  double A, b;        // RKA(stage) (0 for stage 1) and RKb(stage)
  int first;          // 1 on stage 1 (Ut_tmp = Ut)
  // NFVSE blending of the volume integral (nullptr: off): alpha per element and the
  // sub-cell finite-volume volume term Ffv (5 x (N+1)^3 per element, written for alpha > 0)
  const double *alpha;
  const double *Ffv;
};

// Element-centric volume kernel (VolInt_SplitForm volint.f90:173-218 + VolInt_visc
// volint.f90:228-271 + SurfInt gather surfint.f90 + V2D_M_V1D(-sJ) dg.f90:446 +
// the LSERK stage update timedisc.f90:454-470).
//   phase 1  one thread per node stages the flux fields of EPB elements in LDS
//   phase 2  one thread per (direction, element, line) evaluates the n(n-1)/2 symmetric
//            two-point fluxes of its line (each pair once), accumulating both Dvolsurf
//            contributions in registers
//   phase 3  the three direction partial sums are added in LDS (reusing the field
//            storage), then one thread per node adds the weak-form viscous volume term
//            (formed by the lifting kernel), the surface fluxes, applies -sJ and writes
//            Ut or performs the RK stage update (coalesced by node).
// VR is the arithmetic type of phases 1-3a (the flux differencing and its accumulation):
// double for the FP64 path, float for the split-precision path, whose metrics are FP32
// copies.  The per-node fields are formed in FP64 and rounded once; the viscous term, the
// surface gather, the Jacobian and the RK update are FP64 in both.
#ifndef FX_VOL_WAVES
#define FX_VOL_WAVES 1
#endif
#ifndef FX_VOL_ND_LDS
#define FX_VOL_ND_LDS 1
#endif
#ifndef FX_VOL_WAVES32
#define FX_VOL_WAVES32 1
#endif
template <int N, int VF, bool PARA, int MODE, typename VR, int EPB_>
__global__ void __launch_bounds__((VolGeom<N, EPB_>::TB), (sizeof(VR) == 4 ? FX_VOL_WAVES32 : FX_VOL_WAVES))
k_volume(double *__restrict__ U, const VR *__restrict__ Mf, const VR *__restrict__ Mg,
         const VR *__restrict__ Mh, const double *__restrict__ sJ, const int *__restrict__ e2s,
         const double *__restrict__ Flux, const double *__restrict__ Uv, double *__restrict__ Ut,
         double *__restrict__ Uttmp, RKArgs rk, int nElems, double kappa)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr int EPB = EPB_, W = VolGeom<N, EPB_>::W, TB = VolGeom<N, EPB_>::TB, NP = Lds<N>::NP;
  constexpr int NFV = n_fields<VF>();
  constexpr int SLOTS = NFV > 5 ? NFV : 5;
  static_assert(EPB * n2 <= W, "volume launch geometry");
#ifndef FX_VOL_EPAD
#define FX_VOL_EPAD 0
#endif
#ifndef FX_VOL_PRE
#define FX_VOL_PRE 1
#endif
  // per-element LDS blocks padded by FX_VOL_EPAD doubles to spread elements over banks
  __shared__ VR s_lds_raw[EPB][SLOTS * NP + FX_VOL_EPAD];  // node fields
  auto s_lds = reinterpret_cast<VR (*)[SLOTS * NP + FX_VOL_EPAD]>(s_lds_raw);
#define SL(el_, slot_, l_) s_lds[el_][(slot_) * NP + (l_)]
// the direction sums reuse the node-field LDS (one more barrier, 40 % less LDS per block)
#ifndef FX_VOL_ACC_ALIAS
#define FX_VOL_ACC_ALIAS 1
#endif
#if FX_VOL_ND_LDS && !FX_VOL_ACC_ALIAS
  __shared__ VR s_acc_raw[EPB][5 * NP + FX_VOL_EPAD];
  auto s_acc = reinterpret_cast<VR (*)[5 * NP + FX_VOL_EPAD]>(s_acc_raw);
#define SA(el_, v_, l_) s_acc[el_][(v_) * NP + (l_)]
#else
#define SA(el_, v_, l_) SL(el_, v_, l_)
#endif

  const Ops &op = c_ops[N];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  const double kappaM1 = kappa - 1.0;
  const VR skappaM1 = VR(1.0 / kappaM1);

  // phase 2's thread: one per (direction, element, line); an extra wavefront (dir 3) sits it
  // out.  Its line's metric vectors are loaded before phase 1, so that they are in flight
  // together with phase 1's state loads instead of after the barrier.
  const int dir = tid / W, lt = tid - dir * W;
  const int el = lt / n2, line = lt - el * n2;
  const int e = e0 + el;
  const bool active = (dir < 3) && (el < EPB) && (e < nElems) && elem_problem(rk.eprob, e) >= 0;
  const int a = line % n, b = line / n;
  int nodes[n], lds[n];
#pragma unroll
  for (int l = 0; l < n; l++) {
    const int i = (dir == 0) ? l : a, j = (dir == 1) ? l : (dir == 0 ? a : b), k = (dir == 2) ? l : b;
    nodes[l] = i + n * (j + n * k);
    lds[l] = Lds<N>::li(i, j, k);
  }
  VR m[n][3];
  if (active) {
    const VR *M = (dir == 0) ? Mf : ((dir == 1) ? Mg : Mh);
#pragma unroll
    for (int l = 0; l < n; l++) {
      const VR *mp = M + ((size_t)e * n3 + nodes[l]) * 3;
      m[l][0] = mp[0]; m[l][1] = mp[1]; m[l][2] = mp[2];
    }
  }

  // the face gather's (side, flip) pairs of this thread's face points (phase 3b), loaded here so
  // that only the face flux load follows the direction-sum barriers
#ifndef FX_VOL_E2SPRE
#define FX_VOL_E2SPRE 1
#endif
  constexpr int NFG = 6 * n2, NGI = (EPB * NFG + TB - 1) / TB;
  constexpr bool GPRE = FX_VOL_E2SPRE && MODE != VMODE_VOLINT;
  constexpr int NGP = GPRE ? 1 : 0;  // prefetched items (the first; more cost an occupancy level)
  int2 gsd[1];
  if (GPRE) {
#pragma unroll
    for (int g = 0; g < NGP; g++) {
      const int it = tid + g * TB, fe = it / NFG, face = (it - fe * NFG) / n2;
      const int q = ((e0 + fe) * 6 + face) * 2;
      gsd[g] = (it < EPB * NFG && e0 + fe < nElems) ? make_int2(e2s[q], e2s[q + 1]) : make_int2(0, 0);
    }
  }

  // phase 1: per-node fields into LDS
  for (int idx = tid; idx < EPB * n3; idx += TB) {
    const int el1 = idx / n3, node = idx - el1 * n3;
    const int e1 = e0 + el1;
    if (e1 < nElems && elem_problem(rk.eprob, e1) >= 0) {
      const double *u = U + ((size_t)e1 * n3 + node) * 5;
      double uu[5] = {u[0], u[1], u[2], u[3], u[4]};
      double f[NF_COUNT];
      node_fields<VF>(uu, kappaM1, f);
      const int l = Lds<N>::li(node);
#pragma unroll
      for (int c = 0; c < NF_COUNT; c++)
        if (uses_field<VF>(c)) SL(el1, field_slot<VF>(c), l) = VR(f[c]);
    }
  }
  __syncthreads();

  // phase 2: the line's two-point fluxes
#if !FX_VOL_ND_LDS
  VR nd[n][NF_COUNT];
#pragma unroll
  for (int l = 0; l < n; l++)
#pragma unroll
    for (int c = 0; c < NF_COUNT; c++)
      nd[l][c] = (active && uses_field<VF>(c)) ? SL(active ? el : 0, field_slot<VF>(c), lds[l]) : VR(0);
  __syncthreads();  // field storage is free from here on
#endif

  VR acc[n][5];
#pragma unroll
  for (int i = 0; i < n; i++)
#pragma unroll
    for (int v = 0; v < 5; v++) acc[i][v] = VR(0);
  if (active) {
#pragma unroll
    for (int i = 0; i < n; i++)
#pragma unroll
      for (int l = i + 1; l < n; l++) {
        VR F[5];
#if FX_VOL_ND_LDS
        VR ndi[NF_COUNT], ndl[NF_COUNT];
#pragma unroll
        for (int c = 0; c < NF_COUNT; c++) {
          ndi[c] = uses_field<VF>(c) ? SL(el, field_slot<VF>(c), lds[i]) : VR(0);
          ndl[c] = uses_field<VF>(c) ? SL(el, field_slot<VF>(c), lds[l]) : VR(0);
        }
        volume_flux<VF, VR>(ndi, ndl, m[i], m[l], skappaM1, F);
#else
        volume_flux<VF, VR>(nd[i], nd[l], m[i], m[l], skappaM1, F);
#endif
        const VR dil = OpsOf<VR>::ds(op, i, l), dli = OpsOf<VR>::ds(op, l, i);
#pragma unroll
        for (int v = 0; v < 5; v++) {
          acc[i][v] += dil * F[v];
          acc[l][v] += dli * F[v];
        }
      }
  }
  // where every node of the block has its own thread, the node pass's inputs from memory are
  // loaded here, before the direction-sum barriers, so that their latency overlaps them
  constexpr bool PRE = (TB >= EPB * n3) && MODE != VMODE_VOLINT && FX_VOL_PRE;
  double pre_uv[4] = {0.0, 0.0, 0.0, 0.0}, pre_tt[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double pre_u[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, pre_sj = 0.0;
  if (PRE && tid < EPB * n3) {
    const int el2 = tid / n3, e2 = e0 + el2;
    if (e2 < nElems && elem_problem(rk.eprob, e2) >= 0) {
      const size_t c = (size_t)e2 * n3 + (tid - el2 * n3);
      if (PARA) {
#pragma unroll
        for (int q = 0; q < 4; q++) pre_uv[q] = Uv[c * 4 + q];
      }
      pre_sj = sJ[c];
      if (MODE == VMODE_RK) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          pre_u[v] = U[c * 5 + v];
          if (!rk.first) pre_tt[v] = Uttmp[c * 5 + v];
        }
      }
    }
  }
  // phase 3a: xi + eta + zeta partial sums, in this order, into LDS
#if FX_VOL_ND_LDS && FX_VOL_ACC_ALIAS
  __syncthreads();  // the sums overwrite the node fields: every line must have read them
#endif
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (active && dir == d) {
#pragma unroll
      for (int l = 0; l < n; l++)
#pragma unroll
        for (int v = 0; v < 5; v++)
          SA(el, v, lds[l]) = (d == 0) ? acc[l][v] : SA(el, v, lds[l]) + acc[l][v];
    }
    __syncthreads();
  }

  // phase 3b: one work item per face point of the block's elements gathers the element's
  // SurfInt term +-Flux L_HatMinus0 (master +, slave -; surfint.f90:177-184, 253-259) into
  // LDS in element-local face coordinates, so that the node pass reads it without side and
  // flip index arithmetic
  const double L0 = op.L0;
  constexpr int NF = 6 * n2;
  __shared__ double s_face[MODE != VMODE_VOLINT ? EPB : 1][MODE != VMODE_VOLINT ? NF * 5 : 1];
  if (MODE != VMODE_VOLINT) {
#pragma unroll
    for (int g = 0; g < NGI; g++) {
      const int it = tid + g * TB;
      if (it >= EPB * NF) break;
      const int fe = it / NF, r = it - fe * NF, face = r / n2, ab = r - face * n2;
      const int ee = e0 + fe;
      if (ee >= nElems || elem_problem(rk.eprob, ee) < 0) continue;
      const int loc = face + 1, fnode = face_node<N>(loc, ab % n, ab / n);
      const int2 sf = (g < NGP) ? gsd[0] : make_int2(e2s[(ee * 6 + face) * 2], e2s[(ee * 6 + face) * 2 + 1]);
      const int sd = sf.x - 1, fl = sf.y;
      int p, q;
      vol_to_side<N>(loc, fl, fnode % n, (fnode / n) % n, fnode / n2, p, q);
      const double *F = Flux + ((size_t)(sd * n + q) * n + p) * 5;
      const double sg = (fl == 0) ? L0 : -L0;
#pragma unroll
      for (int v = 0; v < 5; v++) s_face[fe][r * 5 + v] = F[v] * sg;
    }
    __syncthreads();
  }
  // phase 3c: one thread per node
  for (int idx = tid; idx < EPB * n3; idx += TB) {
    const int el2 = idx / n3, node = idx - el2 * n3;
    const int e2 = e0 + el2;
    if (e2 >= nElems) continue;
    const int prob = elem_problem(rk.eprob, e2);
    if (prob < 0) continue;
    const int i = node % n, j = (node / n) % n, k = node / n2;
    const int l = Lds<N>::li(node);
    double tot[5];
#pragma unroll
    for (int v = 0; v < 5; v++) tot[v] = double(SA(el2, v, l));
    if (rk.alpha) {
      // VolInt_NFVSE (nfvse.f90:550): Ut = (1 - alpha) Ut + alpha F_FV on the volume part
      const double a = rk.alpha[e2];
      if (a != 0.0) {
        const double *ff = rk.Ffv + ((size_t)e2 * n3 + node) * 5;
#pragma unroll
        for (int v = 0; v < 5; v++) tot[v] = (1.0 - a) * tot[v] + a * ff[v];
      }
    }
    if (PARA && MODE != VMODE_VOLINT) {  // VolInt_visc, formed by the lifting kernel
      const double *uv = Uv + ((size_t)e2 * n3 + node) * 4;
#pragma unroll
      for (int c = 0; c < 4; c++) tot[c + 1] = tot[c + 1] + (PRE ? pre_uv[c] : uv[c]);
    }
    if (MODE != VMODE_VOLINT) {
      // SurfInt: the faces touched by node (i,j,k), in local side order
      const NodeFaces nf = node_faces<N>(i, j, k);
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int r = nf.get(t);
        if (r < 0) continue;
#pragma unroll
        for (int v = 0; v < 5; v++) tot[v] = tot[v] + s_face[el2][r * 5 + v];
      }
      const double msj = -(PRE ? pre_sj : sJ[(size_t)e2 * n3 + node]);
#pragma unroll
      for (int v = 0; v < 5; v++) tot[v] = tot[v] * msj;
    }
    const size_t o = ((size_t)e2 * n3 + node) * 5;
    if (MODE == VMODE_RK) {
      const double bdt = rk.b * rk.st[prob].dt;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const double ut = rk.first ? tot[v] : (PRE ? pre_tt[v] : Uttmp[o + v]) * (-rk.A) + tot[v];
        Uttmp[o + v] = ut;
        U[o + v] = (PRE ? pre_u[v] : U[o + v]) + ut * bdt;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; v++) Ut[o + v] = tot[v];
    }
  }
}

#undef SL
#undef SA

// SurfInt alone (surfint.f90:48-264, Gauss-Lobatto branch :177-184, 253-259): Ut +=
// Flux_master L_HatMinus0 at the master element's face nodes and Ut -= Flux_slave L_HatMinus0
// at the slave element's, one thread per volume node.  A node on several faces (one xi, one
// eta, one zeta face at most) adds them in increasing side id, the order of the reference's
// loop over sides.
template <int N, int NV>
__global__ void __launch_bounds__(256)
k_surfint(const double *__restrict__ Fm, const double *__restrict__ Fs, const int *__restrict__ e2s,
          double *__restrict__ Ut, int nElems)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr int NONE = 0x7fffffff;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nElems * n3) return;
  const int e = (int)(g / n3), node = (int)(g - (long long)e * n3);
  const int i = node % n, j = (node / n) % n, k = node / n2;
  // slot d: the node's face normal to direction d (side id, flux offset, master?)
  int sid[3] = {NONE, NONE, NONE}, off[3] = {0, 0, 0}, mas[3] = {0, 0, 0};
  const int locs[3] = {i == 0 ? XI_MINUS : (i == N ? XI_PLUS : 0), j == 0 ? ETA_MINUS : (j == N ? ETA_PLUS : 0),
                       k == 0 ? ZETA_MINUS : (k == N ? ZETA_PLUS : 0)};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int loc = locs[d];
    if (loc == 0) continue;
    const int sd = e2s[(e * 6 + loc - 1) * 2] - 1, fl = e2s[(e * 6 + loc - 1) * 2 + 1];
    int p, q;
    vol_to_side<N>(loc, fl, i, j, k, p, q);
    sid[d] = sd;
    off[d] = ((sd * n + q) * n + p) * NV;
    mas[d] = (fl == 0);
  }
  // sort the three slots by side id (compare-exchange network)
  auto cx = [&](int a, int b) {
    if (sid[b] < sid[a]) {
      int t = sid[a]; sid[a] = sid[b]; sid[b] = t;
      t = off[a]; off[a] = off[b]; off[b] = t;
      t = mas[a]; mas[a] = mas[b]; mas[b] = t;
    }
  };
  cx(0, 1); cx(1, 2); cx(0, 1);
  const double L0 = c_ops[N].L0;
  double *u = Ut + (size_t)g * NV;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    double r = u[v];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      if (sid[t] == NONE) continue;
      const double f = (mas[t] ? Fm[off[t] + v] : Fs[off[t] + v]) * L0;
      r = mas[t] ? r + f : r - f;
    }
    u[v] = r;
  }
}

// Side-centric numerical flux: Flux(:,p,q,s) = F*(U_master, U_slave; n) * SurfElem, with
// the BR1 viscous average for Navier-Stokes (FillFlux fillflux.f90:41-110, Riemann
// riemann.f90:104-169).
// Launch geometry of the surface kernel: whole sides per block (SPB sides of n^2 points), so
// that a face-wide maximum (Riemann 13-19, 21) is a block-local reduction.
template <int N> struct SurfGeom {
  static constexpr int n2 = (N + 1) * (N + 1), SPB = 256 / n2, T = SPB * n2;
};

// Um (meshes with mortars, else null): the master states of the small sides of mortars
// (k_mortar_u), which have no master element; a big mortar side has no slave element and is
// skipped here (k_flux_mortar fills its flux).
template <int N, int RS, bool PARA>
__global__ void __launch_bounds__(256)
k_surface(const double *__restrict__ U, const double *__restrict__ FvN, const int *__restrict__ s2e,
          const double4 *__restrict__ nvse, double *__restrict__ Flux, int nSides, double kappa,
          const double *__restrict__ Um)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  constexpr bool ES = (RS == 4 || RS == 5 || RS == 13 || RS == 16 || RS == 33 || RS == 44);
  const int gid = blockIdx.x * SurfGeom<N>::T + threadIdx.x;
  const int s = gid / n2, pq = gid - s * n2, p = pq % n, q = pq / n;
  const bool valid = threadIdx.x < SurfGeom<N>::T && gid < nSides * n2;
  const int em = valid ? s2e[5 * s] - 1 : -1;
  const int es = valid ? s2e[5 * s + 1] - 1 : -1;
  // unused side slots of a batch have SideToElem zeroed
  const bool used = valid && es >= 0 && (em >= 0 || (em == -2 && Um != nullptr));
  double UL[5], UR[5], nv[3] = {0.0, 0.0, 0.0}, se = 0.0;
  if (used) {
    const int lm = s2e[5 * s + 2], ls = s2e[5 * s + 3], fl = s2e[5 * s + 4];
    const int vs = side_to_vol<N>(ls, fl, p, q);
    const double *um = em >= 0 ? U + ((size_t)em * n3 + side_to_vol<N>(lm, 0, p, q)) * 5 : Um + (size_t)gid * 5;
    const double *us = U + ((size_t)es * n3 + vs) * 5;
#pragma unroll
    for (int v = 0; v < 5; v++) { UL[v] = um[v]; UR[v] = us[v]; }
    const double4 g = nvse[gid];
    nv[0] = g.x; nv[1] = g.y; nv[2] = g.z; se = g.w;
  }
  double lam = 0.0;
  if (riemann_face_lambda(RS)) {  // LambdaMaxMax of the face
    __shared__ double s_lam[SurfGeom<N>::T];
    if (threadIdx.x < SurfGeom<N>::T) s_lam[threadIdx.x] = used ? face_lambda(UL, UR, nv, kappa) : 0.0;
    __syncthreads();
    if (used) {
      const int b0 = threadIdx.x - pq;
#pragma unroll 4
      for (int k = 0; k < n2; k++) lam = fmax(lam, s_lam[b0 + k]);
    }
  }
  if (!used) return;
  double F[5];
  if (ES) riemann_es<ES ? RS : 33>(UL, UR, nv, kappa, lam, F);
  else riemann<RS>(UL, UR, nv, kappa, F, lam);
  if (PARA) {  // BR1: mean of the two sides' normal viscous fluxes
    const double *fn = FvN + (size_t)gid * 8;
#pragma unroll
    for (int c = 0; c < 4; c++) F[c + 1] = F[c + 1] + 0.5 * (fn[c] + fn[4 + c]);
  }
  double *o = Flux + (size_t)gid * 5;
#pragma unroll
  for (int v = 0; v < 5; v++) o[v] = F[v] * se;
}

// ---------------------------------------------------------------------------------
// Non-conforming interfaces (mortars) on one GPU.  mort holds per mortar [type, big element
// (0-based), big element's local side, small side ids (1-based, 4 or 2), big side id]:
// type 1 = 1->4, 2 = 1->2 in eta, 3 = 1->2 in xi (MortarType, MortarInfo of
// prepare_mesh.f90:700-718).  mops holds MInt[k][l][p] = l_l(xi^k_p) and then
// MProj[k][l][p] (the reference's MInt(l,p,k), MProj(l,p,k); InitMortarBase
// mortar.f90:63-110), k = 0 for the half [-1,0], 1 for [0,1].
constexpr int MORT_INTS = 8;

// Flux_Mortar (fillmortar.f90:239-368; Flux_Cons_Mortar :377-393 for the lifting fluxes):
// the big side's values as the projection of its small sides' values, in the reference's
// order of summation, NC components per face point.  src holds the small sides' values (the
// master values of the small sides: the mortar is their master side), dst receives the big
// side's, both in the face-point layout of Flux; src and dst may be one array (the sides
// differ).  One thread per (mortar, p, q) of the big side.
template <int N, int NC>
__global__ void __launch_bounds__(256)
k_flux_mortar(const double *src, double *dst, const int *__restrict__ mort, const double *__restrict__ mops,
              int nMortar)
{
  constexpr int n = N + 1, n2 = n * n;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = gid / n2, pq = gid - m * n2, p = pq % n, q = pq / n;
  if (m >= nMortar) return;
  const int *mi = mort + MORT_INTS * m;
  const int type = mi[0];
  const double *MP = mops + 2 * n2;
  double o[NC];
#pragma unroll
  for (int v = 0; v < NC; v++) o[v] = 0.0;
  if (type == 1) {  // xi first into Flux_tmp(p,l), then eta
    for (int jNb = 0; jNb < 2; jNb++)
      for (int l = 0; l < n; l++) {
        double t[NC];
#pragma unroll
        for (int v = 0; v < NC; v++) t[v] = 0.0;
        for (int iNb = 0; iNb < 2; iNb++) {
          const double *fs = src + ((size_t)(mi[3 + iNb + 2 * jNb] - 1) * n2 + l * n) * NC;
          for (int l2 = 0; l2 < n; l2++) {
            const double c = MP[(iNb * n + l2) * n + p];
#pragma unroll
            for (int v = 0; v < NC; v++) t[v] = t[v] + c * fs[l2 * NC + v];
          }
        }
        const double c = MP[(jNb * n + l) * n + q];
#pragma unroll
        for (int v = 0; v < NC; v++) o[v] = o[v] + c * t[v];
      }
  } else if (type == 2) {  // halves in eta
    for (int jNb = 0; jNb < 2; jNb++) {
      const double *fs = src + (size_t)(mi[3 + jNb] - 1) * n2 * NC;
      for (int l = 0; l < n; l++) {
        const double c = MP[(jNb * n + l) * n + q];
#pragma unroll
        for (int v = 0; v < NC; v++) o[v] = o[v] + c * fs[(l * n + p) * NC + v];
      }
    }
  } else {  // halves in xi
    for (int iNb = 0; iNb < 2; iNb++) {
      const double *fs = src + (size_t)(mi[3 + iNb] - 1) * n2 * NC;
      for (int l = 0; l < n; l++) {
        const double c = MP[(iNb * n + l) * n + p];
#pragma unroll
        for (int v = 0; v < NC; v++) o[v] = o[v] + c * fs[(q * n + l) * NC + v];
      }
    }
  }
  double *d = dst + ((size_t)(mi[7] - 1) * n2 + pq) * NC;
#pragma unroll
  for (int v = 0; v < NC; v++) d[v] = o[v];
}

// Big side face data -> one small side's face point (InterpolateBigToSmall, as k_mortar_u),
// NC components; big(a, b) returns the pointer to the big side's values at (a, b).
template <int N, int NC, class BIG>
__device__ __forceinline__ void mortar_interp(int type, int im, int p, int q, const double *MI, BIG big, double *o)
{
  constexpr int n = N + 1;
#pragma unroll
  for (int v = 0; v < NC; v++) o[v] = 0.0;
  if (type == 1) {
    const int iNb = im & 1, jNb = im >> 1;
    for (int l = 0; l < n; l++) {
      double t[NC];
#pragma unroll
      for (int v = 0; v < NC; v++) t[v] = 0.0;
      for (int mm = 0; mm < n; mm++) {
        const double c = MI[(jNb * n + mm) * n + q];
        const double *u = big(l, mm);
#pragma unroll
        for (int v = 0; v < NC; v++) t[v] = t[v] + c * u[v];
      }
      const double c = MI[(iNb * n + l) * n + p];
#pragma unroll
      for (int v = 0; v < NC; v++) o[v] = o[v] + c * t[v];
    }
  } else if (type == 2) {
    for (int l = 0; l < n; l++) {
      const double c = MI[(im * n + l) * n + q];
      const double *u = big(p, l);
#pragma unroll
      for (int v = 0; v < NC; v++) o[v] = o[v] + c * u[v];
    }
  } else {
    for (int l = 0; l < n; l++) {
      const double c = MI[(im * n + l) * n + p];
      const double *u = big(l, q);
#pragma unroll
      for (int v = 0; v < NC; v++) o[v] = o[v] + c * u[v];
    }
  }
}

// U_Mortar (fillmortar.f90:72-129) with InterpolateBigToSmall (mortar_vars.f90:203-266): the
// master state of each small side, interpolated from the big element's face (its master side,
// read from the volume as ProlongToFace would), in the reference's order of summation, into
// Um at the small side's face points.  One thread per (mortar, small side, p, q).
template <int N, int NV>
__global__ void __launch_bounds__(256)
k_mortar_u(const double *__restrict__ U, const int *__restrict__ mort, const double *__restrict__ mops,
           double *__restrict__ Um, int nMortar)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = gid / (4 * n2), r = gid - m * (4 * n2), im = r / n2, pq = r - im * n2;
  if (m >= nMortar) return;
  const int *mi = mort + MORT_INTS * m;
  const int type = mi[0], lb = mi[2];
  if (type != 1 && im >= 2) return;
  const double *ub = U + (size_t)mi[1] * n3 * NV;
  double o[NV];
  mortar_interp<N, NV>(type, im, pq % n, pq / n, mops,
                       [&](int a, int b) { return ub + side_to_vol<N>(lb, 0, a, b) * NV; }, o);
  double *dst = Um + ((size_t)(mi[3 + im] - 1) * n2 + pq) * NV;
#pragma unroll
  for (int v = 0; v < NV; v++) dst[v] = o[v];
}

// BR1 lifting with entropy variables (Lifting br1.f90:154-364: Lifting_VolInt
// br1_volint.f90:45-98, Lifting_FillFlux br1_fillflux.f90:49-91 with EvalLiftingSurfFlux
// flux.f90:677-716, Lifting_SurfInt br1_surfint.f90:52-131, V2D_M_V1D(sJ),
// ConvertToGradPrimVec).  Element-centric: the surface term of every face is formed from
// the neighbour's face node directly.  OUT=0 writes the primitive gradients (gradPx/y/z);
// OUT=1 forms the Cartesian viscous fluxes (EvalDiffFlux3D) and writes what the volume and
// surface kernels need of them: the weak-form viscous volume term Uv and the face-normal
// fluxes FvN.  EQ supplies the equation's entropy variables, gradient conversion and
// diffusive fluxes (NsLift; MhdLift for resistive GLM-MHD).
// CALCTIMESTEP of one node (navierstokes/calctimestep.f90:99-136): the convective
// eigenvalue bounds l[0..2] of the xi/eta/zeta directions and, with parabolic terms, the
// viscous ones l[3..5], each at least 1e-12 (a NaN bound becomes 1e-12); err 1 for a NaN
// density, 2 for a non-positive pressure.  Shared by the CALCTIMESTEP kernel and the
// first-stage lifting kernel, so that both form the same bits.
__device__ __forceinline__ void node_dt_bounds(const double *u, const double *mf, const double *mg,
                                               const double *mh, double sj, double kappa, double mu0,
                                               double KsPr_max, int parabolic, double *l, int &err)
{
  if (isnan(u[0])) err = 1;
  const double srho = 1.0 / u[0];
  const double v0 = u[1] * srho, v1 = u[2] * srho, v2 = u[3] * srho;
  const double p = (kappa - 1.0) * (u[4] - 0.5 * u[0] * (v0 * v0 + v1 * v1 + v2 * v2));
  if (p <= 0.0) err = 2;
  const double cs = sqrt(kappa * p * srho);
  l[0] = sj * (fabs(mf[0] * v0 + mf[1] * v1 + mf[2] * v2) + cs * sqrt(mf[0] * mf[0] + mf[1] * mf[1] + mf[2] * mf[2]));
  l[1] = sj * (fabs(mg[0] * v0 + mg[1] * v1 + mg[2] * v2) + cs * sqrt(mg[0] * mg[0] + mg[1] * mg[1] + mg[2] * mg[2]));
  l[2] = sj * (fabs(mh[0] * v0 + mh[1] * v1 + mh[2] * v2) + cs * sqrt(mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2]));
  l[3] = l[4] = l[5] = 1.0e-12;
  if (parabolic) {
    const double muX = srho * KsPr_max * mu0;
    l[3] = muX * ((mf[0] * sj) * (mf[0] * sj) + (mf[1] * sj) * (mf[1] * sj) + (mf[2] * sj) * (mf[2] * sj));
    l[4] = muX * ((mg[0] * sj) * (mg[0] * sj) + (mg[1] * sj) * (mg[1] * sj) + (mg[2] * sj) * (mg[2] * sj));
    l[5] = muX * ((mh[0] * sj) * (mh[0] * sj) + (mh[1] * sj) * (mh[1] * sj) + (mh[2] * sj) * (mh[2] * sj));
  }
#pragma unroll
  for (int d = 0; d < 6; d++) l[d] = fmax(l[d], 1.0e-12);
}

// The element's time step from the maxima of its node bounds (calctimestep.f90:137-156);
// err rises to 3 / 4 for a NaN convective / viscous time step.
__device__ __forceinline__ double elem_dt(const double *lmax, double cfl, double dfl, int parabolic, int &err)
{
  double dte = cfl * 2.0 / (lmax[0] + lmax[1] + lmax[2]);
  if (isnan(dte)) err = max(err, 3);
  if (parabolic) {
    const double dve = dfl * 4.0 / (lmax[3] + lmax[4] + lmax[5]);
    if (isnan(dve)) err = max(err, 4);
    dte = fmin(dte, dve);
  }
  return dte;
}

// CALCTIMESTEP formed inside the first stage's lifting kernel: the minimum time step of each
// block's elements (a NaN element time step skipped, as by the CALCTIMESTEP minimum) and their
// maximum error code into dtBlock / errBlock, one entry per block of the launch.
struct DtFuse {
  double *dtBlock;
  int *errBlock;
  double cfl, dfl, mu0, KsPr_max;
};

template <int N> struct LiftGeom { static constexpr int EPB = (256 / ((N + 1) * (N + 1) * (N + 1))) > 0 ? (256 / ((N + 1) * (N + 1) * (N + 1))) : 1; };

// Navier-Stokes lifting physics: entropy variables (ConsToEntropy), the primitive-gradient
// conversion and the 12 viscous flux values per node (f, g, h of momentum and energy);
// vis_dir gives the NVV = 4 nonzero components (momentum, energy) of f a0 + g a1 + h a2.
struct NsLift {
  static constexpr int NV = 5, NVV = 4;
  double kappa, mu0, KappasPr;
  __device__ __forceinline__ void entropy(const double *u, double *w) const { cons_to_entropy(u, kappa, w); }
  __device__ __forceinline__ void grad_prim(const double *u, double *g) const { grad_entropy_to_prim(u, kappa, g); }
  __device__ __forceinline__ void fluxes(const double *u, const double *gx, const double *gy, const double *gz,
                                         double *fv) const
  {
    diff_flux(u, gx, gy, gz, kappa, mu0, KappasPr, fv);
  }
  __device__ __forceinline__ static void vis_dir(const double *s, double a0, double a1, double a2, double *o)
  {
#pragma unroll
    for (int c = 0; c < 4; c++) o[c] = s[c] * a0 + s[4 + c] * a1 + s[8 + c] * a2;
  }
};

// Outputs of the lifting kernel's viscous form (OUT=1), per element and side:
//   Uv(NVV, n^3, E)      weak-form viscous volume term sum_l D_hat(i,l) (f~, g~, h~)(l) of
//                        VolInt_visc (volint.f90:228-271), components 1..NVV;
//   FvN(NVV, 2, n^2, S)  n.(f, g, h) of each side's element at every face point (master
//                        normal; [0] master, [1] slave), whose mean is the BR1 viscous
//                        surface flux (riemann.f90:155-168).
// The element's transformed fluxes pass through LDS: s_ft holds 3 NVV values per node.
#ifndef FX_LIFT_WAVES
#define FX_LIFT_WAVES 1
#endif
// LIFT 2 is BR2 (br2.f90:167-336, br2_volint.f90:45-98, br2_surfint.f90:52-131): sJ is
// applied inside the volume integral and to every surface term, and the gradient each side
// sees at its face is the node's volume-only gradient plus etaBR2 times that face's own
// surface term (instead of the full gradient of BR1); the volume gradient is the same sum.
// DT (Navier-Stokes, first stage of a step): the kernel also forms the step's CALCTIMESTEP
// from the U and metrics it reads (dtf), a segmented maximum over each wavefront's lanes of
// one element, then one LDS maximum per segment; the bounds are positive and not NaN, so
// their bit patterns order as the values and the maxima are exact in any order.
template <int N, int OUT, class EQ, int LIFT = 1, bool DT = false>
// (Navier-Stokes BR1 at N = 3 asks for 4 waves per SIMD: the matrix-core operands then stay in
// 96-109 VGPRs instead of taking accumulation registers that cost an occupancy level.)
__global__ void __launch_bounds__(256, (N == 3 && EQ::NV == 5 && LIFT == 1) ? 4 : FX_LIFT_WAVES)
k_lifting(const double *__restrict__ U, const double *__restrict__ Mf, const double *__restrict__ Mg,
          const double *__restrict__ Mh, const double *__restrict__ sJ, const int2 *__restrict__ ftab,
          const double4 *__restrict__ nvse, double *__restrict__ out0,
          double *__restrict__ out1, double *__restrict__ out2, const int *__restrict__ eprob,
          int nElems, EQ eq, double etaBR2, const double *__restrict__ Um, const double *__restrict__ LiftM,
          double *__restrict__ GradM, int nMortar, DtFuse dtf)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n, NV = EQ::NV, NVV = EQ::NVV;
  constexpr int EPB = LiftGeom<N>::EPB, NF = 6 * n2;  // face points per element
  // one LDS buffer per element: the entropy variables (WSZ) and after them the face data of
  // phase F ((NV + 3) per face point: the BR1 lifting flux 1/2 (w_slave - w_master) SurfElem
  // and the normal) until the gradients are formed; then (after a barrier) from its start the
  // node's 12 Cartesian viscous flux values, then its transformed fluxes (3 NVV)
  constexpr int FTW = 3 * NVV > 12 ? 3 * NVV : 12;
  constexpr int WSZ = NV * n3, FSZ = (NV + 3) * NF;
  constexpr int USZ = WSZ + FSZ > FTW * n3 ? WSZ + FSZ : FTW * n3;
  __shared__ double s_u[EPB][USZ];
  // node slot of a per-node LDS array: at N = 3 bits 1 and 3 are flipped where bit 5 is set,
  // which leaves the node's own, the xi- and the eta-transposed reads of the matrix-core
  // contractions free of bank conflicts (each 32-lane half of a wavefront hits 32 distinct
  // 8-byte slots modulo 32)
#ifndef FX_LIFT_SWZ
#define FX_LIFT_SWZ 1
#endif
  auto SW = [](int x) { return (N == 3 && FX_LIFT_SWZ) ? (x ^ (((x >> 5) & 1) * 10)) : x; };
#define LW(el_, v_, node_) s_u[el_][(v_) * n3 + SW(node_)]
  __shared__ int s_fpt[EPB][NF];  // 2 * face point ((s n + q) n + p) + (0 master, 1 slave)
  const Ops &op = c_ops[N];
  const int tid = threadIdx.x;
  const int el = tid / n3, node = tid - el * n3;
  const int e = blockIdx.x * EPB + el;
  const bool active = (el < EPB) && (e < nElems) && elem_problem(eprob, e) >= 0;
  const int i = node % n, j = (node / n) % n, k = node / n2;
  const size_t c = (size_t)e * n3 + node;
  static_assert(!DT || (NV == 5 && OUT == 1), "CALCTIMESTEP in the Navier-Stokes stage lifting only");
  __shared__ unsigned long long s_dtl[DT ? EPB : 1][6];  // bit patterns of the element maxima
  __shared__ int s_dterr[DT ? EPB : 1];
  if (DT && tid < EPB * 7) {
    if (tid < EPB * 6) s_dtl[tid / 6][tid % 6] = 0ull;
    else s_dterr[tid - EPB * 6] = 0;
  }
  // the node's metric vectors and Jacobian, loaded ahead of the face phase so that their latency
  // overlaps it instead of following its barrier (Navier-Stokes BR1 at N = 3, whose 4 waves per
  // SIMD leave room for them; elsewhere they would cost an occupancy level)
#ifndef FX_LIFT_PREM
#define FX_LIFT_PREM 1
#endif
  constexpr bool PREM = FX_LIFT_PREM && N == 3 && NV == 5 && LIFT == 1;
  double mreg[9], sjreg = 0.0;
  if (PREM && active) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
      mreg[q] = Mf[3 * c + q];
      mreg[3 + q] = Mg[3 * c + q];
      mreg[6 + q] = Mh[3 * c + q];
    }
    sjreg = sJ[c];
  }
  // the thread's first face point (phase F): table entry, neighbour state and normal, loaded
  // before the element's own state so that both are in flight together (where the metric
  // prefetch above is made too)
  constexpr bool PREF = PREM;
  // EARLYF (BR1): the node's viscous fluxes, its face-normal fluxes (FvN, with the normals that
  // phase F left in LDS) and its transformed fluxes are formed before the LDS is handed over, so
  // that only the 3 NVV transformed values cross the barrier (one barrier and the normals' global
  // loads fewer than storing the 12 Cartesian values first)
#ifndef FX_LIFT_EARLYF
#define FX_LIFT_EARLYF 1
#endif
  constexpr bool EARLYF = FX_LIFT_EARLYF && LIFT == 1 && OUT == 1;
  int2 ft0 = make_int2(-1, -1);
  double un0[NV];
  double4 g0 = make_double4(0.0, 0.0, 0.0, 0.0);
  if (PREF && tid < EPB * NF) {
    const int ee = blockIdx.x * EPB + tid / NF;
    if (ee < nElems && elem_problem(eprob, ee) >= 0) {
      ft0 = ftab[(size_t)blockIdx.x * EPB * NF + tid];
      if (ft0.x != -1) {
        const int fpt = ft0.y >> 1;
        const double *unbp = (ft0.x >= 0) ? U + (size_t)ft0.x * NV : Um + (size_t)fpt * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) un0[v] = unbp[v];
        g0 = nvse[fpt];
      }
    }
  }
  double u[NV], w[NV];
  if (active) {
    const double *up = U + c * NV;
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = up[v];
    eq.entropy(u, w);
#pragma unroll
    for (int v = 0; v < NV; v++) LW(el, v, node) = w[v];
  }
  __syncthreads();
  // phase F: one work item per face point of the block's elements (the element-local
  // coordinates keep a wavefront on few sides): neighbour state, its entropy variables and
  // the strong lifting flux, Lifting_FillFlux (br1_fillflux.f90:49-91)
  for (int it = tid; it < EPB * NF; it += 256) {
    const int fe = it / NF, r = it - fe * NF;
    const int ee = blockIdx.x * EPB + fe;
    if (ee >= nElems || elem_problem(eprob, ee) < 0) continue;
    const bool pre = PREF && it == tid;
    const int2 t = pre ? ft0 : ftab[(size_t)blockIdx.x * EPB * NF + it];
    s_fpt[fe][r] = t.y;
    // a big mortar side has no neighbour element: its lifting flux comes projected (LiftM); the
    // small side of a mortar has the big side's interpolated state (Um) as its master
    if (t.x == -1) {
      if (EARLYF) {  // the normal alone, for the face-normal viscous flux
        const double4 gm = nvse[t.y >> 1];
        double *d = &s_u[fe][WSZ + r * (NV + 3)];
        d[NV] = gm.x; d[NV + 1] = gm.y; d[NV + 2] = gm.z;
      }
      continue;
    }
    const int fpt = t.y >> 1;
    const bool master = (t.y & 1) == 0;
    double unb[NV], wnb[NV];
    double4 g;
    if (pre) {
#pragma unroll
      for (int v = 0; v < NV; v++) unb[v] = un0[v];
      g = g0;
    } else {
      const double *unbp = (t.x >= 0) ? U + (size_t)t.x * NV : Um + (size_t)fpt * NV;
#pragma unroll
      for (int v = 0; v < NV; v++) unb[v] = unbp[v];
      g = nvse[fpt];
    }
    eq.entropy(unb, wnb);
    const int face = r / n2, ab = r - face * n2, fnode = face_node<N>(face + 1, ab % n, ab / n);
    double *d = &s_u[fe][WSZ + r * (NV + 3)];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double wo = LW(fe, v, fnode);
      const double ws = master ? wnb[v] : wo;
      const double wm = master ? wo : wnb[v];
      d[v] = 0.5 * (ws - wm) * g.w;
    }
    d[NV] = g.x; d[NV + 1] = g.y; d[NV + 2] = g.z;
  }
  __syncthreads();
  double gx[NV], gy[NV], gz[NV];
  if (active) {
    double gxi[NV], geta[NV], gzeta[NV];
    if constexpr (N == 3 && LIFT == 1) {
      // The strong-form D contractions on the matrix core (BR1 at N = 3: 5 % faster than the
      // FMA loop below on the Navier-Stokes benchmark; with BR2 the matrix-core registers would
      // cost an occupancy level): a wavefront is one element, lane = node i + 4 j + 16 k.  v_mfma_f64_4x4x4_4b_f64 forms 4 blocks b of D(m,n) = sum_q A(m,q)
      // B(q,n) with A(m,q) in lane m + 4 b + 16 q, B(q,n) in lane n + 4 b + 16 q and D(m,n) in
      // lane n + 4 b + 16 m.  Every contraction takes the operator entry D(r,q) in lane
      // r + 4 b + 16 q (dop):
      //   zeta: A = D(k, q) [m=k, b=j], B = W(i, j, q) = the lane's own entropy variables, D in
      //         the node's lane;
      //   xi:   A = W(q, j, k) [m=k, b=j], read from LDS at node q + 4 j + 16 k, B = D(i, q);
      //   eta:  A = D(j, q) [m=j, b=k], B = W(i, q, k) from LDS; D(j, i; b=k) lands in lane
      //         i + 4 k + 16 j and is fetched from there (ds_bpermute).
      const int L = node;
      const double dop = op.D[L & 3][L >> 4];
      const int nx = (L >> 4) + (L & 12) + 16 * (L & 3);      // lane z+4j+16q <- node q+4j+16z
      const int ny = (L & 3) + 4 * (L >> 4) + 16 * ((L >> 2) & 3);  // lane i+4k+16q <- node i+4q+16k
#pragma unroll
      for (int v = 0; v < NV; v++) {
        gzeta[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(dop, w[v], 0.0, 0, 0, 0);
        gxi[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(LW(el, v, nx), dop, 0.0, 0, 0, 0);
        const double te = __builtin_amdgcn_mfma_f64_4x4x4f64(dop, LW(el, v, ny), 0.0, 0, 0, 0);
        const long long tb = __double_as_longlong(te);
        const int lo = __builtin_amdgcn_ds_bpermute(4 * ny, (int)(tb & 0xffffffffLL));
        const int hi = __builtin_amdgcn_ds_bpermute(4 * ny, (int)(tb >> 32));
        geta[v] = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        gxi[v] = op.D[i][0] * LW(el, v, 0 + n * (j + n * k));
        geta[v] = op.D[j][0] * LW(el, v, i + n * (0 + n * k));
        gzeta[v] = op.D[k][0] * LW(el, v, i + n * (j + n * 0));
      }
      // not unrolled: all n NV LDS operands in flight at once would cost an occupancy level
#pragma unroll 1
      for (int l = 1; l < n; l++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          gxi[v] = gxi[v] + op.D[i][l] * LW(el, v, l + n * (j + n * k));
          geta[v] = geta[v] + op.D[j][l] * LW(el, v, i + n * (l + n * k));
          gzeta[v] = gzeta[v] + op.D[k][l] * LW(el, v, i + n * (j + n * l));
        }
    }
    const double *mf = PREM ? mreg : Mf + 3 * c, *mg = PREM ? mreg + 3 : Mg + 3 * c, *mh = PREM ? mreg + 6 : Mh + 3 * c;
    const double sj = PREM ? sjreg : sJ[c];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      gx[v] = mf[0] * gxi[v] + mg[0] * geta[v] + mh[0] * gzeta[v];
      gy[v] = mf[1] * gxi[v] + mg[1] * geta[v] + mh[1] * gzeta[v];
      gz[v] = mf[2] * gxi[v] + mg[2] * geta[v] + mh[2] * gzeta[v];
      if (LIFT == 2) { gx[v] = gx[v] * sj; gy[v] = gy[v] * sj; gz[v] = gz[v] * sj; }
    }
    // BR2: the volume-only gradient, the base of every face gradient (in gxi.. storage)
    if (LIFT == 2) {
#pragma unroll
      for (int v = 0; v < NV; v++) { gxi[v] = gx[v]; geta[v] = gy[v]; gzeta[v] = gz[v]; }
    }
    // Lifting_SurfInt: + flux n L_HatMinus0 of every face of the node (br1_surfint.f90:52-131)
    const NodeFaces nf = node_faces<N>(i, j, k);
    const double L0 = op.L0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int r = nf.get(t);
      if (r < 0) continue;
      const double *d = &s_u[el][WSZ + r * (NV + 3)];
      const double nx = d[NV], ny = d[NV + 1], nz = d[NV + 2];
      const int mcode = s_fpt[el][r];
      if (LIFT == 1 && (mcode >> 1) < nMortar * n2) {
        // big mortar side: Flux_Cons_Mortar of the small sides' FluxX, FluxY, FluxZ
        const double *lm = LiftM + (size_t)(mcode >> 1) * 3 * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) {
          gx[v] = gx[v] + lm[v] * L0;
          gy[v] = gy[v] + lm[NV + v] * L0;
          gz[v] = gz[v] + lm[2 * NV + v] * L0;
        }
      } else if (LIFT == 1) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double f = d[v];
          gx[v] = gx[v] + (f * nx) * L0;
          gy[v] = gy[v] + (f * ny) * L0;
          gz[v] = gz[v] + (f * nz) * L0;
        }
      } else {
        // F_loc = sJ L_HatMinus0 Flux: into the volume gradient, and etaBR2 F_loc on the
        // volume-only gradient is this face's gradient
        const double sl = sj * L0;
        const bool mbig = (mcode >> 1) < nMortar * n2;  // big mortar side: the projected fluxes
        double fx[NV], fy[NV], fz[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          double lx, ly, lz;
          if (mbig) {
            const double *lm = LiftM + (size_t)(mcode >> 1) * 3 * NV;
            lx = sl * lm[v]; ly = sl * lm[NV + v]; lz = sl * lm[2 * NV + v];
          } else {
            const double f = d[v];
            lx = sl * (f * nx); ly = sl * (f * ny); lz = sl * (f * nz);
          }
          gx[v] = gx[v] + lx;
          gy[v] = gy[v] + ly;
          gz[v] = gz[v] + lz;
          fx[v] = gxi[v] + etaBR2 * lx;
          fy[v] = geta[v] + etaBR2 * ly;
          fz[v] = gzeta[v] + etaBR2 * lz;
        }
        if (mbig) {
          // the big side's face gradients of the entropy variables, for U_Mortar (converted to
          // primitive gradients with the small sides' interpolated states, br2.f90:271-290)
          double *gm = GradM + (size_t)(mcode >> 1) * 3 * NV;
#pragma unroll
          for (int v = 0; v < NV; v++) { gm[v] = fx[v]; gm[NV + v] = fy[v]; gm[2 * NV + v] = fz[v]; }
        } else if (OUT == 1) {
          double fv[12], o[NVV];
          eq.grad_prim(u, fx);
          eq.grad_prim(u, fy);
          eq.grad_prim(u, fz);
          eq.fluxes(u, fx, fy, fz, fv);
          EQ::vis_dir(fv, nx, ny, nz, o);
          double *dst = out1 + (size_t)s_fpt[el][r] * NVV;
#pragma unroll
          for (int v = 0; v < NVV; v++) dst[v] = o[v];
        }
      }
    }
    if (LIFT == 1) {
#pragma unroll
      for (int v = 0; v < NV; v++) { gx[v] = gx[v] * sj; gy[v] = gy[v] * sj; gz[v] = gz[v] * sj; }
    }
    eq.grad_prim(u, gx);
    eq.grad_prim(u, gy);
    eq.grad_prim(u, gz);
    if (LIFT == 1 && nMortar > 0) {  // the big side's prolonged gradients, for U_Mortar of gradPx/y/z
      const NodeFaces nfm = node_faces<N>(i, j, k);
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int r = nfm.get(t);
        if (r < 0) continue;
        const int code = s_fpt[el][r];
        if ((code >> 1) >= nMortar * n2) continue;
        double *gm = GradM + (size_t)(code >> 1) * 3 * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) { gm[v] = gx[v]; gm[NV + v] = gy[v]; gm[2 * NV + v] = gz[v]; }
      }
    }
    if (OUT == 0) {
#pragma unroll
      for (int v = 0; v < NV; v++) { out0[NV * c + v] = gx[v]; out1[NV * c + v] = gy[v]; out2[NV * c + v] = gz[v]; }
    }
  }
  if constexpr (DT) {  // every lane of the (full, 256-thread) block takes part
    double l[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int err = 0;
    const int key = active ? el : -1;
    if (active)
      node_dt_bounds(u, PREM ? mreg : Mf + 3 * c, PREM ? mreg + 3 : Mg + 3 * c, PREM ? mreg + 6 : Mh + 3 * c,
                     PREM ? sjreg : sJ[c], eq.kappa, dtf.mu0, dtf.KsPr_max, 1, l, err);
    const int lane = tid & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {  // suffix maximum within the lanes of one element
      const int ko = __shfl_down(key, s);
      const int eo = __shfl_down(err, s);
      double lo[6];
#pragma unroll
      for (int d = 0; d < 6; d++) lo[d] = __shfl_down(l[d], s);
      if (lane + s < 64 && ko == key) {
#pragma unroll
        for (int d = 0; d < 6; d++) l[d] = fmax(l[d], lo[d]);
        err = max(err, eo);
      }
    }
    const int kprev = __shfl_up(key, 1);
    if (key >= 0 && (lane == 0 || kprev != key)) {
#pragma unroll
      for (int d = 0; d < 6; d++) atomicMax(&s_dtl[el][d], (unsigned long long)__double_as_longlong(l[d]));
      atomicMax(&s_dterr[el], err);
    }
  }
  double ftr[EARLYF ? 3 * NVV : 1];
  if (EARLYF && active) {
    double fv[12], o[NVV];
    eq.fluxes(u, gx, gy, gz, fv);
    const NodeFaces nf = node_faces<N>(i, j, k);
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int r = nf.get(t);
      if (r < 0) continue;
      const double *d = &s_u[el][WSZ + r * (NV + 3)];
      EQ::vis_dir(fv, d[NV], d[NV + 1], d[NV + 2], o);
      double *dst = out1 + (size_t)s_fpt[el][r] * NVV;
#pragma unroll
      for (int v = 0; v < NVV; v++) dst[v] = o[v];
    }
    const double *mf = PREM ? mreg : Mf + 3 * c, *mg = PREM ? mreg + 3 : Mg + 3 * c, *mh = PREM ? mreg + 6 : Mh + 3 * c;
    EQ::vis_dir(fv, mf[0], mf[1], mf[2], ftr);
    EQ::vis_dir(fv, mg[0], mg[1], mg[2], ftr + NVV);
    EQ::vis_dir(fv, mh[0], mh[1], mh[2], ftr + 2 * NVV);
  }
  if (OUT == 0) return;
  __syncthreads();  // the entropy variables and face data are no longer read; the fluxes take their memory
  if constexpr (DT) {  // the block's minimum time step (NaN skipped) and maximum error code
    if (tid == 0) {
      double m = 1.79e308;
      int er = 0;
      for (int q = 0; q < EPB; q++) {
        const int ee = blockIdx.x * EPB + q;
        if (ee >= nElems || elem_problem(eprob, ee) < 0) continue;
        double lmax[6];
#pragma unroll
        for (int d = 0; d < 6; d++) lmax[d] = __longlong_as_double((long long)s_dtl[q][d]);
        int eq_er = s_dterr[q];
        const double dte = elem_dt(lmax, dtf.cfl, dtf.dfl, 1, eq_er);
        if (dte < m) m = dte;
        er = max(er, eq_er);
      }
      dtf.dtBlock[blockIdx.x] = m;
      dtf.errBlock[blockIdx.x] = er;
    }
  }
  if (EARLYF) {
    if (active) {
#pragma unroll
      for (int m = 0; m < 3 * NVV; m++) s_u[el][m * n3 + SW(node)] = ftr[m];
    }
  } else {
    if (active) {
      // the 12 Cartesian viscous flux values of this node into its own slots
      double fv[12];
      eq.fluxes(u, gx, gy, gz, fv);
#pragma unroll
      for (int m = 0; m < 12; m++) s_u[el][m * n3 + SW(node)] = fv[m];
    }
    __syncthreads();
    if (active) {
      double fv[12], o[NVV];
#pragma unroll
      for (int m = 0; m < 12; m++) fv[m] = s_u[el][m * n3 + SW(node)];
      // face-normal viscous flux of this element at each of its face points (BR1: from the
      // node's full gradient; BR2 wrote them from the face gradients above)
      if (LIFT == 1) {
        const NodeFaces nf = node_faces<N>(i, j, k);
#pragma unroll
        for (int t = 0; t < 3; t++) {
          const int r = nf.get(t);
          if (r < 0) continue;
          const int code = s_fpt[el][r];
          const double4 g = nvse[code >> 1];
          EQ::vis_dir(fv, g.x, g.y, g.z, o);
          double *dst = out1 + (size_t)code * NVV;
#pragma unroll
          for (int v = 0; v < NVV; v++) dst[v] = o[v];
        }
      }
      // transformed fluxes f~, g~, h~ (EvalDiffFluxTilde3D) over the node's own slots
      const double *mf = PREM ? mreg : Mf + 3 * c, *mg = PREM ? mreg + 3 : Mg + 3 * c, *mh = PREM ? mreg + 6 : Mh + 3 * c;
      EQ::vis_dir(fv, mf[0], mf[1], mf[2], o);
#pragma unroll
      for (int v = 0; v < NVV; v++) s_u[el][v * n3 + SW(node)] = o[v];
      EQ::vis_dir(fv, mg[0], mg[1], mg[2], o);
#pragma unroll
      for (int v = 0; v < NVV; v++) s_u[el][(NVV + v) * n3 + SW(node)] = o[v];
      EQ::vis_dir(fv, mh[0], mh[1], mh[2], o);
#pragma unroll
      for (int v = 0; v < NVV; v++) s_u[el][(2 * NVV + v) * n3 + SW(node)] = o[v];
    }
  }
  __syncthreads();
  if (active) {
    double uv[NVV];
    if constexpr (N == 3 && LIFT == 1) {
      // the D_hat contractions on the matrix core, with the lane layout of the D contractions
      // above: xi and zeta accumulate in the node's lane (zeta reads the node's own h~), eta
      // lands in lane i + 4 k + 16 j and is fetched from there
      const int L = node;
      const double dh = op.Dhat[L & 3][L >> 4];
      const int nx = (L >> 4) + (L & 12) + 16 * (L & 3);
      const int ny = (L & 3) + 4 * (L >> 4) + 16 * ((L >> 2) & 3);
#pragma unroll
      for (int v = 0; v < NVV; v++) {
        double acc = __builtin_amdgcn_mfma_f64_4x4x4f64(s_u[el][v * n3 + SW(nx)], dh, 0.0, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_4x4x4f64(dh, s_u[el][(2 * NVV + v) * n3 + SW(L)], acc, 0, 0, 0);
        const double te = __builtin_amdgcn_mfma_f64_4x4x4f64(dh, s_u[el][(NVV + v) * n3 + SW(ny)], 0.0, 0, 0, 0);
        const long long tb = __double_as_longlong(te);
        const int lo = __builtin_amdgcn_ds_bpermute(4 * ny, (int)(tb & 0xffffffffLL));
        const int hi = __builtin_amdgcn_ds_bpermute(4 * ny, (int)(tb >> 32));
        uv[v] = acc + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
      }
    } else {
#pragma unroll
      for (int v = 0; v < NVV; v++) uv[v] = 0.0;
#pragma unroll 1
      for (int l = 0; l < n; l++) {
        const double di = op.Dhat[i][l], dj = op.Dhat[j][l], dk = op.Dhat[k][l];
#pragma unroll
        for (int v = 0; v < NVV; v++)
          uv[v] = uv[v] + di * s_u[el][v * n3 + SW(l + n * (j + n * k))] +
                  dj * s_u[el][(NVV + v) * n3 + SW(i + n * (l + n * k))] +
                  dk * s_u[el][(2 * NVV + v) * n3 + SW(i + n * (j + n * l))];
      }
    }
#pragma unroll
    for (int v = 0; v < NVV; v++) out0[NVV * c + v] = uv[v];
  }
}
#undef LW


// BR1 lifting on the small sides of mortars (Lifting_FillFlux br1_fillflux.f90:49-91 with
// EvalLiftingSurfFlux flux.f90:677-716): F = 1/2 (w_slave - w_master) SurfElem with the master
// state interpolated from the big side (Um), times each normal component into LiftS (3 NV
// per face point: FluxX, FluxY, FluxZ), for Flux_Cons_Mortar onto the big side.  One thread
// per (mortar, small side, p, q).
template <int N, class EQ>
__global__ void __launch_bounds__(256)
k_lift_mortar_small(const double *__restrict__ U, const double *__restrict__ Um, const int *__restrict__ s2e,
                    const double4 *__restrict__ nvse, const int *__restrict__ mort, double *__restrict__ LiftS,
                    int nMortar, EQ eq)
{
  constexpr int n = N + 1, n2 = n * n, n3 = n2 * n, NV = EQ::NV;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = gid / (4 * n2), r = gid - m * (4 * n2), im = r / n2, pq = r - im * n2;
  if (m >= nMortar) return;
  const int *mi = mort + MORT_INTS * m;
  if (mi[0] != 1 && im >= 2) return;
  const int s = mi[3 + im] - 1;
  const size_t fpt = (size_t)s * n2 + pq;
  const int es = s2e[5 * s + 1] - 1, ls = s2e[5 * s + 3], fl = s2e[5 * s + 4];
  const double *us = U + ((size_t)es * n3 + side_to_vol<N>(ls, fl, pq % n, pq / n)) * NV;
  const double *um = Um + fpt * NV;
  double u[NV], ws[NV], wm[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = us[v];
  eq.entropy(u, ws);
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = um[v];
  eq.entropy(u, wm);
  const double4 g = nvse[fpt];
  double *o = LiftS + fpt * 3 * NV;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double f = 0.5 * (ws[v] - wm[v]) * g.w;
    o[v] = f * g.x;
    o[NV + v] = f * g.y;
    o[2 * NV + v] = f * g.z;
  }
}

// The master side's viscous flux on the small sides of mortars: U_Mortar of the big side's
// gradients (BR1: primitive gradients, br1.f90:245-322; BR2 (CONV): the face gradients of the
// entropy variables, converted here with the interpolated state, br2.f90:271-290) and of its
// state (Um), EvalDiffFlux3D and the small side's normal into the master slot of FvN (as the
// lifting kernel writes it for an element's own faces), whose mean with the small element's
// is the viscous surface flux (riemann.f90:155-168).  One thread per (mortar, small side, p, q).
template <int N, class EQ, bool CONV>
__global__ void __launch_bounds__(256)
k_mortar_fv(const double *__restrict__ Um, const double *__restrict__ GradM, const double4 *__restrict__ nvse,
            const int *__restrict__ mort, const double *__restrict__ mops, double *__restrict__ FvN, int nMortar,
            EQ eq)
{
  constexpr int n = N + 1, n2 = n * n, NV = EQ::NV, NVV = EQ::NVV;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = gid / (4 * n2), r = gid - m * (4 * n2), im = r / n2, pq = r - im * n2;
  if (m >= nMortar) return;
  const int *mi = mort + MORT_INTS * m;
  if (mi[0] != 1 && im >= 2) return;
  const size_t fpt = (size_t)(mi[3 + im] - 1) * n2 + pq;
  const double *gb = GradM + (size_t)(mi[7] - 1) * n2 * 3 * NV;
  double gr[3 * NV];
  mortar_interp<N, 3 * NV>(mi[0], im, pq % n, pq / n, mops,
                           [&](int a, int b) { return gb + (size_t)(b * n + a) * 3 * NV; }, gr);
  double u[NV], fv[12], o[NVV];
#pragma unroll
  for (int v = 0; v < NV; v++) u[v] = Um[fpt * NV + v];
  if (CONV) {
    eq.grad_prim(u, gr);
    eq.grad_prim(u, gr + NV);
    eq.grad_prim(u, gr + 2 * NV);
  }
  eq.fluxes(u, gr, gr + NV, gr + 2 * NV, fv);
  const double4 g = nvse[fpt];
  EQ::vis_dir(fv, g.x, g.y, g.z, o);
  double *dst = FvN + (2 * fpt) * NVV;
#pragma unroll
  for (int v = 0; v < NVV; v++) dst[v] = o[v];
}

// ---------------------------------------------------------------------------------
// CALCTIMESTEP (navierstokes/calctimestep.f90:39-170): per-element convective and
// viscous eigenvalue bounds into dtElem (as the reference's dtElem), then a one-block
// minimum in fixed order (deterministic, no atomics).

// one group of G = 2^ceil(log2 n^3) threads per element, tree max-reduction in LDS
template <int N> struct DtGeom {
  static constexpr int n3 = (N + 1) * (N + 1) * (N + 1);
  static constexpr int G = n3 <= 32 ? 32 : (n3 <= 64 ? 64 : (n3 <= 128 ? 128 : (n3 <= 256 ? 256 : 512)));
  static constexpr int EPB = 256 / G > 0 ? 256 / G : 1;
  static constexpr int T = EPB * G;
};

template <int N>
__global__ void __launch_bounds__(DtGeom<N>::T)
k_calc_dt(const double *__restrict__ U, const double *__restrict__ Mf, const double *__restrict__ Mg,
          const double *__restrict__ Mh, const double *__restrict__ sJ, double *__restrict__ dtElem,
          int *__restrict__ errElem, const int *__restrict__ eprob, int nElems, double kappa,
          double cfl, double dfl, double mu0, double KsPr_max, int parabolic)
{
  constexpr int n3 = DtGeom<N>::n3, G = DtGeom<N>::G, EPB = DtGeom<N>::EPB, T = DtGeom<N>::T;
  __shared__ double s_l[6][T];
  __shared__ int s_err[T];
  const int tid = threadIdx.x;
  const int el = tid / G, node = tid - el * G;
  const int e = blockIdx.x * EPB + el;
  const bool used = e < nElems && elem_problem(eprob, e) >= 0;
  double l[6] = {1.0e-12, 1.0e-12, 1.0e-12, 1.0e-12, 1.0e-12, 1.0e-12};
  int err = 0;
  if (node < n3 && used) {
    const size_t c = (size_t)e * n3 + node;
    node_dt_bounds(U + 5 * c, Mf + 3 * c, Mg + 3 * c, Mh + 3 * c, sJ[c], kappa, mu0, KsPr_max, parabolic, l, err);
  }
#pragma unroll
  for (int d = 0; d < 6; d++) s_l[d][tid] = l[d];
  s_err[tid] = err;
  __syncthreads();
  for (int stp = G / 2; stp > 0; stp >>= 1) {
    if (node < stp) {
#pragma unroll
      for (int d = 0; d < 6; d++) s_l[d][tid] = fmax(s_l[d][tid], s_l[d][tid + stp]);
      s_err[tid] = max(s_err[tid], s_err[tid + stp]);
    }
    __syncthreads();
  }
  if (node == 0 && used) {
    int er = s_err[tid];
    const double lmax[6] = {s_l[0][tid], s_l[1][tid], s_l[2][tid], s_l[3][tid], s_l[4][tid], s_l[5][tid]};
    dtElem[e] = elem_dt(lmax, cfl, dfl, parabolic, er);
    errElem[e] = er;
  }
}

#ifdef FX_KERNEL_HOME  // non-template kernels: compiled in one unit only
// Minimum of dtElem (NaN entries skipped) and maximum error code, one block, fixed order.
__global__ void __launch_bounds__(1024)
k_dt_reduce(const double *__restrict__ dtElem, const int *__restrict__ errElem, int nElems, DtState *st,
            int advance)
{
  __shared__ double s_dt[1024];
  __shared__ int s_er[1024];
  const int tid = threadIdx.x;
  double m = 1.79e308;
  int er = 0;
  for (int e = tid; e < nElems; e += 1024) {
    const double d = dtElem[e];
    if (d < m) m = d;
    er = max(er, errElem[e]);
  }
  s_dt[tid] = m;
  s_er[tid] = er;
  __syncthreads();
  for (int stp = 512; stp > 0; stp >>= 1) {
    if (tid < stp) {
      s_dt[tid] = fmin(s_dt[tid], s_dt[tid + stp]);
      s_er[tid] = max(s_er[tid], s_er[tid + stp]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    const bool was_done = st->done != 0;
    st->dt = end_time_step(*st, s_dt[0], advance);
    st->errType = max(st->errType, s_er[0]);
    if (st->errType != 0 && st->errStep < 0) st->errStep = st->step;
    if (advance && !was_done) { st->t = st->t + st->dt; st->step = st->step + 1; }
  }
}

// Per-problem CALCTIMESTEP minimum for a batch: block r reduces the contiguous element
// range [first[r], first[r] + count[r]) of problem r into st[r] (count 0: unused slot).
// The minimum is exact in any summation order, so a problem's dt in a batch equals its dt
// computed alone (k_dt_reduce).
__global__ void __launch_bounds__(256)
k_dt_reduce_seg(const double *__restrict__ dtElem, const int *__restrict__ errElem,
                const int *__restrict__ first, const int *__restrict__ count, DtState *st, int advance)
{
  __shared__ double s_dt[256];
  __shared__ int s_er[256];
  const int r = blockIdx.x, tid = threadIdx.x;
  const int e0 = first[r], ne = count[r];
  if (ne <= 0) return;
  double m = 1.79e308;
  int er = 0;
  for (int e = e0 + tid; e < e0 + ne; e += 256) {
    const double d = dtElem[e];
    if (d < m) m = d;
    er = max(er, errElem[e]);
  }
  s_dt[tid] = m;
  s_er[tid] = er;
  __syncthreads();
  for (int stp = 128; stp > 0; stp >>= 1) {
    if (tid < stp) {
      s_dt[tid] = fmin(s_dt[tid], s_dt[tid + stp]);
      s_er[tid] = max(s_er[tid], s_er[tid + stp]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    DtState &S = st[r];
    S.dt = s_dt[0];
    S.errType = max(S.errType, s_er[0]);
    if (S.errType != 0 && S.errStep < 0) S.errStep = S.step;
    if (advance) { S.t = S.t + S.dt; S.step = S.step + 1; }
  }
}
#endif  // FX_KERNEL_HOME

// CalcEntropy (analyze_equation.f90:549-599): per-block partial sums (fixed order),
// reduced by a second pass in block order.
template <int N>
__global__ void __launch_bounds__(256)
k_entropy(const double *__restrict__ U, const double *__restrict__ Ut, const double *__restrict__ sJ,
          double *__restrict__ part, long long nNodes, double kappa)
{
  constexpr int n = N + 1, n3 = n * n * n;
  __shared__ double s0[256], s1[256];
  const Ops &op = c_ops[N];
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (gid < nNodes) {
    const int node = (int)(gid % n3);
    const int i = node % n, j = (node / n) % n, k = node / (n * n);
    const double *u = U + 5 * gid;
    double uu[5] = {u[0], u[1], u[2], u[3], u[4]};
    const double km1 = kappa - 1.0;
    double srho = 1.0 / uu[0];
    double v1 = uu[1] * srho, v2 = uu[2] * srho, v3 = uu[3] * srho;
    double p = km1 * (uu[4] - 0.5 * (uu[1] * v1 + uu[2] * v2 + uu[3] * v3));
    double wv = op.w[i] * op.w[j] * op.w[k];
    double el = -uu[0] * (log(p) - kappa * log(uu[0]));
    a = el * wv / sJ[gid];
    double w[5];
    cons_to_entropy(uu, kappa, w);
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < 5; v++) s += w[v] * Ut[5 * gid + v];
    b = s * wv / sJ[gid];
  }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) { s0[threadIdx.x] += s0[threadIdx.x + st]; s1[threadIdx.x] += s1[threadIdx.x + st]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

// (templated on the degree of the caller so that each unit instantiates its own)
template <int TAG>
__global__ void k_sum_pairs(const double *__restrict__ part, int nPart, double *__restrict__ out)
{
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nPart; i += 256) { a += part[2 * i]; b += part[2 * i + 1]; }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) { s0[threadIdx.x] += s0[threadIdx.x + st]; s1[threadIdx.x] += s1[threadIdx.x + st]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

}  // namespace fx
