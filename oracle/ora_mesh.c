/*
 * Oracle (test infrastructure only): periodic box meshes, the reference's
 * side/volume index mappings, and the curvilinear metric terms.
 *
 * Mappings restate src/mesh/mappings.f90:262-515 (Flip_S2M, CGNS_SideToVol,
 * CGNS_SideToVol2, SideToVol).  Metrics restate src/mesh/metrics.f90:78-483
 * (CalcMetrics with curl-form metrics and the Jacobian built on NgeoRef=3*Ngeo,
 * CalcSurfMetrics, SurfMetricsFromJa).  The box generator is this project's own
 * stand-in for HOPR (no mesh files exist for the benchmark configurations):
 * element e=(ix,iy,iz) -> ix+Kx*(iy+Ky*iz), side 3e+d is the +d face of
 * element e (master, flip 0) and the -d face of its periodic neighbour (slave).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ora.h"

/* Flip_S2M (mappings.f90:262-284) */
void ora_flip_s2m(int N, int p, int q, int flip, int pq[2])
{
  switch (flip) {
  case 0: pq[0] = p; pq[1] = q; break;
  case 1: pq[0] = q; pq[1] = p; break;
  case 2: pq[0] = N - p; pq[1] = q; break;
  case 3: pq[0] = N - q; pq[1] = N - p; break;
  default: pq[0] = p; pq[1] = N - q; break;
  }
}

/* CGNS_SideToVol (mappings.f90:345-369) */
void ora_cgns_side_to_vol(int N, int l, int p, int q, int locSide, int ijk[3])
{
  switch (locSide) {
  case ORA_XI_MINUS:   ijk[0] = l;     ijk[1] = q;     ijk[2] = p; break;
  case ORA_XI_PLUS:    ijk[0] = N - l; ijk[1] = p;     ijk[2] = q; break;
  case ORA_ETA_MINUS:  ijk[0] = p;     ijk[1] = l;     ijk[2] = q; break;
  case ORA_ETA_PLUS:   ijk[0] = N - p; ijk[1] = N - l; ijk[2] = q; break;
  case ORA_ZETA_MINUS: ijk[0] = q;     ijk[1] = p;     ijk[2] = l; break;
  default:             ijk[0] = p;     ijk[1] = q;     ijk[2] = N - l; break;
  }
}

/* CGNS_SideToVol2 (mappings.f90:379-403) */
static void cgns_side_to_vol2(int N, int p, int q, int locSide, int pq[2])
{
  switch (locSide) {
  case ORA_XI_MINUS: case ORA_ZETA_MINUS: pq[0] = q; pq[1] = p; break;
  case ORA_ETA_PLUS: pq[0] = N - p; pq[1] = q; break;
  default: pq[0] = p; pq[1] = q; break;
  }
}

/* SideToVol = CGNS_SideToVol(Flip_M2S(...)) (mappings.f90:459-472) */
void ora_side_to_vol(int N, int l, int p, int q, int flip, int locSide, int ijk[3])
{
  int pq[2];
  ora_flip_s2m(N, p, q, flip, pq);
  ora_cgns_side_to_vol(N, l, pq[0], pq[1], locSide, ijk);
}

static const int plus_side[3] = {ORA_XI_PLUS, ORA_ETA_PLUS, ORA_ZETA_PLUS};
static const int minus_side[3] = {ORA_XI_MINUS, ORA_ETA_MINUS, ORA_ZETA_MINUS};

/* flip of the slave (-d face) so that master point (p,q) and slave point hit the same
   physical node; checked on a degree-2 face so that every flip is distinguishable */
static int find_flip(int d)
{
  const int N = 2;
  for (int f = 1; f <= 4; f++) {
    int ok = 1;
    for (int q = 0; q <= N && ok; q++)
      for (int p = 0; p <= N && ok; p++) {
        int m[3], s[3];
        ora_side_to_vol(N, 0, p, q, 0, plus_side[d], m);
        ora_side_to_vol(N, 0, p, q, f, minus_side[d], s);
        m[d] = 0; /* the neighbour's node sits on its minus face */
        if (m[0] != s[0] || m[1] != s[1] || m[2] != s[2]) ok = 0;
      }
    if (ok) return f;
  }
  return -1;
}

int ora_box_connectivity(int Kx, int Ky, int Kz, int *e2s, int *s2e)
{
  int K[3] = {Kx, Ky, Kz};
  int nE = Kx * Ky * Kz;
  int flips[3];
  for (int d = 0; d < 3; d++) { flips[d] = find_flip(d); if (flips[d] < 0) return -1; }
  for (int iz = 0; iz < Kz; iz++)
    for (int iy = 0; iy < Ky; iy++)
      for (int ix = 0; ix < Kx; ix++) {
        int e = ix + Kx * (iy + Ky * iz);
        int c[3] = {ix, iy, iz};
        for (int d = 0; d < 3; d++) {
          int cn[3] = {c[0], c[1], c[2]};
          cn[d] = (cn[d] + 1) % K[d];
          int en = cn[0] + Kx * (cn[1] + Ky * cn[2]);
          int s = 3 * e + d;
          s2e[s * 5 + 0] = e + 1;
          s2e[s * 5 + 1] = en + 1;
          s2e[s * 5 + 2] = plus_side[d];
          s2e[s * 5 + 3] = minus_side[d];
          s2e[s * 5 + 4] = flips[d];
          e2s[(e * 6 + plus_side[d] - 1) * 2 + 0] = s + 1;
          e2s[(e * 6 + plus_side[d] - 1) * 2 + 1] = 0;
          e2s[(en * 6 + minus_side[d] - 1) * 2 + 0] = s + 1;
          e2s[(en * 6 + minus_side[d] - 1) * 2 + 1] = flips[d];
        }
      }
  return nE;
}

/* Equidistant Ngeo nodes of a Kx x Ky x Kz box; warp=1 adds the periodic sine
   displacement amp*prod_k sin(2 pi (x_k-x0_k)/len_k) to every coordinate, which on
   [0,2]^3 with amp=0.1 is the reference's meshdeform (src/mesh/mesh.f90:238-245). */
void ora_box_nodecoords(int Kx, int Ky, int Kz, const double *x0, const double *len, int Ngeo,
                        int warp, double amp, double *X)
{
  const double pi = acos(-1.0);
  int K[3] = {Kx, Ky, Kz};
  int ng = Ngeo + 1;
  for (int iz = 0; iz < Kz; iz++)
    for (int iy = 0; iy < Ky; iy++)
      for (int ix = 0; ix < Kx; ix++) {
        int e = ix + Kx * (iy + Ky * iz);
        int c[3] = {ix, iy, iz};
        for (int k = 0; k < ng; k++)
          for (int j = 0; j < ng; j++)
            for (int i = 0; i < ng; i++) {
              int loc[3] = {i, j, k};
              double x[3];
              for (int d = 0; d < 3; d++) {
                double t = (double)(c[d] * Ngeo + loc[d]) / (double)(K[d] * Ngeo);
                x[d] = x0[d] + t * len[d];
              }
              if (warp) {
                double s = amp;
                for (int d = 0; d < 3; d++) s *= sin(2.0 * pi * (x[d] - x0[d]) / len[d]);
                for (int d = 0; d < 3; d++) x[d] += s;
              }
              double *o = X + 3 * (i + ng * (j + ng * (k + (size_t)ng * e)));
              o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
            }
      }
}

static const int normal_dirs[6] = {3, 2, 1, 2, 1, 3};  /* mesh_vars.f90:91 */
static const int tang_dirs[6] = {1, 3, 2, 3, 2, 1};    /* mesh_vars.f90:92 */
static const double normal_signs[6] = {-1., -1., 1., 1., -1., 1.}; /* mesh_vars.f90:93 */

/* CalcMetrics + CalcSurfMetrics for LGL volume nodes and EQUI mesh nodes */
int ora_calc_metrics(int N, int Ngeo, int nElems, int nSides, const double *NodeCoords,
                     const int *e2s, double *Mf, double *Mg, double *Mh, double *sJ,
                     double *NormVec, double *TangVec1, double *TangVec2, double *SurfElem,
                     double *Elem_xGP)
{
  (void)nSides;
  int n = N + 1, ng = Ngeo + 1, NgeoRef = 3 * Ngeo, nr = NgeoRef + 1;
  int bad = 0;
  double *V_eq_gl = malloc(sizeof(double) * ng * ng);
  double *DGLgeo = malloc(sizeof(double) * ng * ng);
  double *V_gl_ref = malloc(sizeof(double) * nr * ng);
  double *V_ref_N = malloc(sizeof(double) * n * nr);
  double *V_gl_glN = malloc(sizeof(double) * n * ng);
  double *DGLN = malloc(sizeof(double) * n * n);
  double xg[64];
  ora_vandermonde(Ngeo, ORA_NODE_EQUI, Ngeo, ORA_NODE_GL, 0, V_eq_gl);
  ora_nodes(Ngeo, ORA_NODE_GL, xg, NULL);
  ora_derivative_matrix(Ngeo, xg, DGLgeo);
  ora_vandermonde(Ngeo, ORA_NODE_GL, NgeoRef, ORA_NODE_GL, 0, V_gl_ref);
  ora_vandermonde(NgeoRef, ORA_NODE_GL, N, ORA_NODE_GL, 1, V_ref_N);
  ora_vandermonde(Ngeo, ORA_NODE_GL, N, ORA_NODE_GL, 0, V_gl_glN);
  double xN[64];
  ora_nodes(N, ORA_NODE_GL, xN, NULL);
  ora_derivative_matrix(N, xN, DGLN);

  double *XGLgeo = malloc(sizeof(double) * 3 * ng * ng * ng);
  double *XGLN = malloc(sizeof(double) * 3 * n * n * n);
  double *dXgeo = malloc(sizeof(double) * 9 * ng * ng * ng);   /* (dd,nn,i,j,k) */
  double *slice_geo = malloc(sizeof(double) * 3 * ng * ng * ng);
  double *slice_ref = malloc(sizeof(double) * 3 * nr * nr * nr);
  double *dXref = malloc(sizeof(double) * 9 * nr * nr * nr);
  double *detRef = malloc(sizeof(double) * nr * nr * nr);
  double *detN = malloc(sizeof(double) * n * n * n);
  double *slice_N = malloc(sizeof(double) * 3 * n * n * n);
  double *dXN = malloc(sizeof(double) * 9 * n * n * n);
  double *R = malloc(sizeof(double) * 9 * n * n * n);
  double *Ja = malloc(sizeof(double) * 9 * n * n * n);
#define G3(i, j, k) ((i) + ng * ((j) + ng * (k)))
#define R3(i, j, k) ((i) + nr * ((j) + nr * (k)))
#define N3(i, j, k) ((i) + n * ((j) + n * (k)))
  for (int e = 0; e < nElems; e++) {
    const double *NC = NodeCoords + (size_t)3 * ng * ng * ng * e;
    ora_change_basis_3d(3, Ngeo, Ngeo, V_eq_gl, NC, XGLgeo);
    ora_change_basis_3d(3, Ngeo, N, V_gl_glN, XGLgeo, XGLN);
    if (Elem_xGP) memcpy(Elem_xGP + (size_t)3 * n * n * n * e, XGLN, sizeof(double) * 3 * n * n * n);
    /* 1.b) dXGL_Ngeo(dd,nn,i,j,k) */
    memset(dXgeo, 0, sizeof(double) * 9 * ng * ng * ng);
    for (int k = 0; k < ng; k++)
      for (int j = 0; j < ng; j++)
        for (int i = 0; i < ng; i++) {
          double *o = dXgeo + 9 * G3(i, j, k);
          for (int l = 0; l < ng; l++)
            for (int nn = 0; nn < 3; nn++) {
              o[0 + 3 * nn] += DGLgeo[i + ng * l] * XGLgeo[nn + 3 * G3(l, j, k)];
              o[1 + 3 * nn] += DGLgeo[j + ng * l] * XGLgeo[nn + 3 * G3(i, l, k)];
              o[2 + 3 * nn] += DGLgeo[k + ng * l] * XGLgeo[nn + 3 * G3(i, j, l)];
            }
        }
    /* 1.c) Jacobian on NgeoRef */
    for (int nn = 0; nn < 3; nn++) {
      for (int p = 0; p < ng * ng * ng; p++)
        for (int dd = 0; dd < 3; dd++) slice_geo[dd + 3 * p] = dXgeo[dd + 3 * nn + 9 * p];
      ora_change_basis_3d(3, Ngeo, NgeoRef, V_gl_ref, slice_geo, slice_ref);
      for (int p = 0; p < nr * nr * nr; p++)
        for (int dd = 0; dd < 3; dd++) dXref[dd + 3 * nn + 9 * p] = slice_ref[dd + 3 * p];
    }
    for (int p = 0; p < nr * nr * nr; p++) {
      const double *a = dXref + 9 * p;
#define A(r, c) a[((r)-1) + 3 * ((c)-1)]
      detRef[p] = 0.0 + A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3))
                  + A(2, 1) * (A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3))
                  + A(3, 1) * (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3));
#undef A
    }
    ora_change_basis_3d(1, NgeoRef, N, V_ref_N, detRef, detN);
    double jmin = 1e300, jmax = -1e300;
    for (int p = 0; p < n * n * n; p++) {
      sJ[(size_t)n * n * n * e + p] = 1.0 / detN[p];
      if (detN[p] < jmin) jmin = detN[p];
      if (detN[p] > jmax) jmax = detN[p];
    }
    if (jmin <= 0.0 || jmin / jmax < 0.001) bad = 1;
    /* 2.a) dXGL_N */
    if (N >= Ngeo) {
      for (int nn = 0; nn < 3; nn++) {
        for (int p = 0; p < ng * ng * ng; p++)
          for (int dd = 0; dd < 3; dd++) slice_geo[dd + 3 * p] = dXgeo[dd + 3 * nn + 9 * p];
        ora_change_basis_3d(3, Ngeo, N, V_gl_glN, slice_geo, slice_N);
        for (int p = 0; p < n * n * n; p++)
          for (int dd = 0; dd < 3; dd++) dXN[dd + 3 * nn + 9 * p] = slice_N[dd + 3 * p];
      }
    } else {
      memset(dXN, 0, sizeof(double) * 9 * n * n * n);
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            double *o = dXN + 9 * N3(i, j, k);
            for (int l = 0; l < n; l++)
              for (int nn = 0; nn < 3; nn++) {
                o[0 + 3 * nn] += DGLN[i + n * l] * XGLN[nn + 3 * N3(l, j, k)];
                o[1 + 3 * nn] += DGLN[j + n * l] * XGLN[nn + 3 * N3(i, l, k)];
                o[2 + 3 * nn] += DGLN[k + n * l] * XGLN[nn + 3 * N3(i, j, l)];
              }
          }
    }
    /* curl form: R(dd,nn) ; Ja(dd,nn) */
    for (int p = 0; p < n * n * n; p++) {
      const double *dx = dXN + 9 * p, *X = XGLN + 3 * p;
      double *r = R + 9 * p;
      for (int dd = 0; dd < 3; dd++) {
        r[dd + 0] = 0.5 * (X[2] * dx[dd + 3] - X[1] * dx[dd + 6]);
        r[dd + 3] = 0.5 * (X[0] * dx[dd + 6] - X[2] * dx[dd + 0]);
        r[dd + 6] = 0.5 * (X[1] * dx[dd + 0] - X[0] * dx[dd + 3]);
      }
    }
    memset(Ja, 0, sizeof(double) * 9 * n * n * n);
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double *o = Ja + 9 * N3(i, j, k);
          for (int q = 0; q < n; q++)
            for (int nn = 0; nn < 3; nn++) {
              o[0 + 3 * nn] -= DGLN[j + n * q] * R[2 + 3 * nn + 9 * N3(i, q, k)];
              o[1 + 3 * nn] -= DGLN[k + n * q] * R[0 + 3 * nn + 9 * N3(i, j, q)];
              o[2 + 3 * nn] -= DGLN[i + n * q] * R[1 + 3 * nn + 9 * N3(q, j, k)];
            }
          for (int q = 0; q < n; q++)
            for (int nn = 0; nn < 3; nn++) {
              o[0 + 3 * nn] += DGLN[k + n * q] * R[1 + 3 * nn + 9 * N3(i, j, q)];
              o[1 + 3 * nn] += DGLN[i + n * q] * R[2 + 3 * nn + 9 * N3(q, j, k)];
              o[2 + 3 * nn] += DGLN[j + n * q] * R[0 + 3 * nn + 9 * N3(i, q, k)];
            }
        }
    /* volume metrics (Vdm_GLN_N is the identity for LGL) */
    for (int p = 0; p < n * n * n; p++)
      for (int nn = 0; nn < 3; nn++) {
        size_t o = (size_t)3 * ((size_t)n * n * n * e + p) + nn;
        Mf[o] = Ja[0 + 3 * nn + 9 * p];
        Mg[o] = Ja[1 + 3 * nn + 9 * p];
        Mh[o] = Ja[2 + 3 * nn + 9 * p];
      }
    /* CalcSurfMetrics for master sides (metrics.f90:334-449) */
    for (int loc = 1; loc <= 6; loc++) {
      int SideID = e2s[(e * 6 + loc - 1) * 2 + 0] - 1;
      int flip = e2s[(e * 6 + loc - 1) * 2 + 1];
      if (flip != 0) continue;
      int nd = normal_dirs[loc - 1] - 1, td = tang_dirs[loc - 1] - 1;
      double ns = normal_signs[loc - 1];
      for (int q = 0; q < n; q++)
        for (int p = 0; p < n; p++) {
          int pq[2], ijk[3];
          cgns_side_to_vol2(N, p, q, loc, pq);
          /* face slice index (a,b) -> volume node */
          switch (loc) {
          case ORA_XI_MINUS: ijk[0] = 0; ijk[1] = pq[0]; ijk[2] = pq[1]; break;
          case ORA_XI_PLUS: ijk[0] = N; ijk[1] = pq[0]; ijk[2] = pq[1]; break;
          case ORA_ETA_MINUS: ijk[0] = pq[0]; ijk[1] = 0; ijk[2] = pq[1]; break;
          case ORA_ETA_PLUS: ijk[0] = pq[0]; ijk[1] = N; ijk[2] = pq[1]; break;
          case ORA_ZETA_MINUS: ijk[0] = pq[0]; ijk[1] = pq[1]; ijk[2] = 0; break;
          default: ijk[0] = pq[0]; ijk[1] = pq[1]; ijk[2] = N; break;
          }
          const double *ja = Ja + 9 * N3(ijk[0], ijk[1], ijk[2]);
          double jn[3] = {ja[nd + 0], ja[nd + 3], ja[nd + 6]};
          double jt[3] = {ja[td + 0], ja[td + 3], ja[td + 6]};
          size_t fo = (size_t)(SideID * n + q) * n + p;
          double se = sqrt(jn[0] * jn[0] + jn[1] * jn[1] + jn[2] * jn[2]);
          double nv[3], t1[3], t2[3];
          for (int d = 0; d < 3; d++) nv[d] = ns * jn[d] / se;
          double dot = jt[0] * nv[0] + jt[1] * nv[1] + jt[2] * nv[2];
          for (int d = 0; d < 3; d++) t1[d] = jt[d] - dot * nv[d];
          double tn = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
          for (int d = 0; d < 3; d++) t1[d] = t1[d] / tn;
          t2[0] = nv[1] * t1[2] - nv[2] * t1[1];
          t2[1] = nv[2] * t1[0] - nv[0] * t1[2];
          t2[2] = nv[0] * t1[1] - nv[1] * t1[0];
          SurfElem[fo] = se;
          for (int d = 0; d < 3; d++) {
            NormVec[3 * fo + d] = nv[d];
            TangVec1[3 * fo + d] = t1[d];
            TangVec2[3 * fo + d] = t2[d];
          }
        }
    }
  }
#undef G3
#undef R3
#undef N3
  free(V_eq_gl); free(DGLgeo); free(V_gl_ref); free(V_ref_N); free(V_gl_glN); free(DGLN);
  free(XGLgeo); free(XGLN); free(dXgeo); free(slice_geo); free(slice_ref); free(dXref);
  free(detRef); free(detN); free(slice_N); free(dXN); free(R); free(Ja);
  return bad ? -1 : 0;
}
