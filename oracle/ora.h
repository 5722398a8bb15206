/*
 * CPU oracle for the FLUXO split-form DGSEM right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (project-fluxo/fluxo, Fortran) used as the parity checker for the
 * HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the reported CPU baseline.  The
 * product path never links or calls anything in this directory.
 *
 * Every routine cites the reference file:line it restates.  Array layouts follow
 * the reference (Fortran column-major, first index fastest):
 *   U(nVar,0:N,0:N,0:N,nElems)        -> U[(((e*n+k)*n+j)*n+i)*nVar+v]
 *   Metrics_fTilde(3,0:N,0:N,0:N,E)   -> M[(((e*n+k)*n+j)*n+i)*3+d]
 *   sJ(0:N,0:N,0:N,E)                 -> sJ[((e*n+k)*n+j)*n+i]
 *   NormVec(3,0:N,0:N,nSides)         -> nv[((s*n+q)*n+p)*3+d]
 *   SurfElem(0:N,0:N,nSides)          -> se[(s*n+q)*n+p]
 *   ElemToSide(2,6,nElems)            -> e2s[(e*6+locSide-1)*2+{0:SideID,1:flip}]
 *   SideToElem(5,nSides)              -> s2e[s*5+{0:Elem,1:nbElem,2:locSide,3:nbLocSide,4:flip}]
 * Element and side IDs stored in the connectivity arrays are 1-based as in the
 * reference; -1 marks "none".
 *
 * Parity status of each part is listed in oracle/README.md.
 */
#ifndef FLUXO_ORACLE_H
#define FLUXO_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* node types of GetNodesAndWeights (src/interpolation/interpolation.f90:190-265) */
enum { ORA_NODE_EQUI = 0, ORA_NODE_GL = 1, ORA_NODE_GAUSS = 2 };

/* local side ids (src/defines.h:68-73) */
enum { ORA_ZETA_MINUS = 1, ORA_ETA_MINUS = 2, ORA_XI_PLUS = 3, ORA_ETA_PLUS = 4,
       ORA_XI_MINUS = 5, ORA_ZETA_PLUS = 6 };

/* ---- basis (src/interpolation/basis.f90, interpolation.f90, changeBasis.f90) ---- */
void ora_legendre(int N, double x, double *L, double *Lder);
void ora_gauss_nodes(int N, double *x, double *w);
void ora_lgl_nodes(int N, double *x, double *w);
void ora_nodes(int N, int type, double *x, double *w);
void ora_barycentric_weights(int N, const double *x, double *wb);
void ora_derivative_matrix(int N, const double *x, double *D);
void ora_lagrange_polys(double x, int N, const double *xGP, const double *wb, double *L);
void ora_legendre_vdm(int N, const double *xi, double *Vdm_Leg, double *sVdm_Leg);
void ora_vandermonde(int Nin, int typeIn, int Nout, int typeOut, int modal, double *Vdm);
void ora_change_basis_3d(int dim1, int Nin, int Nout, const double *Vdm, const double *Xin, double *Xout);
/* D, D_Hat, Dvolsurf, L_HatMinus, L_HatPlus for LGL nodes (src/dg/dg.f90:181-252) */
void ora_dg_operators(int N, double *xGP, double *wGP, double *D, double *D_hat,
                      double *Dvolsurf, double *L_hatminus, double *L_hatplus);

/* ---- mesh (periodic box, my generator) and mappings (src/mesh/mappings.f90) ---- */
void ora_cgns_side_to_vol(int N, int l, int p, int q, int locSide, int ijk[3]);
void ora_flip_s2m(int N, int p, int q, int flip, int pq[2]);
void ora_side_to_vol(int N, int l, int p, int q, int flip, int locSide, int ijk[3]);
int  ora_box_connectivity(int Kx, int Ky, int Kz, int *ElemToSide, int *SideToElem);
void ora_box_nodecoords(int Kx, int Ky, int Kz, const double *x0, const double *len, int Ngeo,
                        int warp, double amp, double *NodeCoords);
int  ora_calc_metrics(int N, int Ngeo, int nElems, int nSides, const double *NodeCoords,
                      const int *ElemToSide, double *Mf, double *Mg, double *Mh, double *sJ,
                      double *NormVec, double *TangVec1, double *TangVec2, double *SurfElem,
                      double *Elem_xGP);

/* ---- equation options (ints and reals), mirrors parameter.ini keys ---- */
typedef struct {
  int N, nVar, nElems, nSides;
  int volflux;       /* VolumeFlux: 0 standard DG, 1 standard DG with averaged metrics,
                        2/8 Kennedy-Gruber, 3 Ducros, 4 Morinishi, 5 Chandrashekar EC-KEP,
                        6 EC with arithmetic means, 7 gg, 9 Gassner-Winters-Walch,
                        10 Ismail-Roe EC, 32 Ranocha EC-KEP */
  int riemann;       /* Riemann: 0-24, 32, 33, 44 (navierstokes/equation.f90:292-404) */
  int parabolic;     /* 0 Euler, 1 Navier-Stokes */
  int lifting;       /* 1 BR1, 2 BR2 */
  double kappa, mu0, Pr, etaBR2;
  /* mesh */
  const int *ElemToSide, *SideToElem;
  const double *Mf, *Mg, *Mh, *sJ, *NormVec, *TangVec1, *TangVec2, *SurfElem;
  /* NFVSE shock capturing (SHOCK_NFVSE, SubFVMethod=1, ComputeAlpha=1, ModalThreshold=1,
     TimeRelFactor=0, indicator rho*p): nfvse=0 off */
  int nfvse, spacePropSweeps;
  double alpha_max, alpha_min, spacePropFactor;
  double *alpha;     /* optional out (nElems): blending coefficient of the last RHS */
  int computeAlpha;  /* 1 Persson-Peraire indicator, 3 fixed alpha = shockBlendCoef */
  double shockBlendCoef;
} ora_problem;

/* two-point and surface fluxes; U/aux layout per point: U(5), aux(6)=(1/rho,v1,v2,v3,p,|v|^2) */
double ora_ln_mean(double aL, double aR);
void ora_eval_uaux(double kappa, const double *U, double *aux);
void ora_volflux(int which, double kappa, const double *UL, const double *UR, const double *auxL,
                 const double *auxR, const double *mL, const double *mR, double *F);
/* numerical flux at one face point (Riemann 0-24, 32, 33, 44); Riemann 13-19 and 21 take the
   face's largest wave speed lam_face (ora_riemann uses the point's own); Roe (3) takes its
   central part from volflux (ora_riemann_lam_vf; the others pass 0) */
void ora_riemann(int which, double kappa, const double *UL, const double *UR, const double *nv,
                 const double *t1, const double *t2, double *F);
void ora_riemann_lam(int which, double kappa, const double *UL, const double *UR, const double *nv,
                     const double *t1, const double *t2, double lam_face, double *F);
void ora_riemann_lam_vf(int which, int volflux, double kappa, const double *UL, const double *UR,
                        const double *nv, const double *t1, const double *t2, double lam_face, double *F);
double ora_face_lambda(double kappa, const double *UL, const double *UR, const double *nv);
void ora_cons_to_entropy(double kappa, const double *U, double *w);

/* ---- DG operator ---- */
/* full DGTimeDerivative_weakForm (src/dg/dg.f90:282-452) */
void ora_dg_time_derivative(const ora_problem *P, const double *U, double t, double *Ut);
/* individual contributions (no sJ scaling): volume split form, surface integral */
void ora_volint_splitform(const ora_problem *P, const double *U, double *Ut);
void ora_surface_integral(const ora_problem *P, const double *U, const double *gx, const double *gy,
                          const double *gz, double *Ut);
/* FillFlux (src/dg/fillflux.f90:41-110) of the state: Flux_master and Flux_slave
   (5,0:N,0:N,nSides), the numerical flux (with the BR1/BR2 viscous mean when parabolic) times
   SurfElem; Flux_slave = Flux_master */
void ora_fill_flux(const ora_problem *P, const double *U, double *Fm, double *Fs);
/* SurfInt (src/dg/surfint.f90:48-264, Gauss-Lobatto branch :177-184, 253-259): Ut +=
   Flux_master L_HatMinus0 (master face nodes), -= Flux_slave L_HatMinus0 (slave), side by side */
void ora_surf_int(const ora_problem *P, int nvar, const double *Fm, const double *Fs, double *Ut);
/* BR1 lifting of nvar lifting variables w (volume nodes): physical gradients of w */
void ora_lift(const ora_problem *P, int nvar, const double *w, double *gx, double *gy, double *gz);
/* BR2 lifting of nvar lifting variables w: volume gradients and each side's master/slave
   face gradients (nvar,0:N,0:N,nSides) */
void ora_lift_br2(const ora_problem *P, int nvar, const double *w, double eta, double *gx, double *gy,
                  double *gz, double *fxm, double *fym, double *fzm, double *fxs, double *fys, double *fzs);
/* BR1 (lifting 1) or BR2 (lifting 2, volume gradients) lifting of the Navier-Stokes entropy
   variables: primitive gradients gradPx/y/z(5,n,n,n,E) */
void ora_lifting(const ora_problem *P, const double *U, double *gx, double *gy, double *gz);
/* Carpenter-Kennedy RK4(5) 2N low-storage step (src/timedisc/timedisc.f90:424-480) */
void ora_lserk_step(const ora_problem *P, double *U, double *Ut_tmp, double t, double dt);
/* CALCTIMESTEP (src/equation/navierstokes/calctimestep.f90:39-170) with FillCFL_DFL scaling */
double ora_calc_timestep(const ora_problem *P, const double *U, double CFLscale, double DFLscale,
                         int *errType);
/* CalcEntropy (src/equation/navierstokes/analyze_equation.f90:549-599) */
void ora_calc_entropy(const ora_problem *P, const double *U, const double *Ut, double *S, double *dSdU_Ut);
/* exact functions 2 (density wave) and 7 (Taylor-Green) on the volume nodes
   (src/equation/navierstokes/equation.f90:632-651, 802-811) */
void ora_exactfunc(int which, double kappa, const double *advvel, double amp, double t,
                   int npts, const double *x, double *U);

/* ---- NFVSE shock capturing (src/shockcapturing/nfvse.f90, src/indicators/indicators.f90) ---- */
double ora_nfvse_threshold(int N);
/* sub-cell face metrics of one element: [d][f][b][a][nv(3) t1(3) t2(3) norm], f = 0..N-1 */
void ora_subcell_metrics(int N, const double *Mf, const double *Mg, const double *Mh, double *out);
void ora_persson_peraire(const ora_problem *P, const double *U, double *eta);
void ora_blending_coefficient(const ora_problem *P, const double *U, double *alpha);
void ora_volint_nfvse(const ora_problem *P, const double *U, const double *alpha, double *Ut);

/* ---- GLM-MHD (src/equation/mhd), 9 variables (rho, rho v, E, B, psi): ideal, or resistive
   (viscosity mu, resistivity eta, heat conduction mu kappa/Pr) with BR1 lifting of the
   entropy variables ---- */
typedef struct {
  const ora_problem *P;   /* mesh and N (nVar, volflux, riemann of P are not used) */
  double kappa, mu_0, glm_scale;
  int parabolic;
  int riemann;            /* surface flux: 0 central, 1 Rusanov, 10 Derigs + ES-LLF, 11 Derigs */
  double mu, eta, Pr;
  int lifting, reserved2; /* resistive case: 1 BR1 (default when 0), 2 BR2 with etaBR2 */
  double etaBR2;
} ora_mhd;
/* primitive gradients (rho, v, p, B, psi) of the resistive case, gradPx/y/z(9,n,n,n,E) */
void ora_mhd_lifting(const ora_mhd *M, const double *U, double *gx, double *gy, double *gz);
void ora_mhd_derigs(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, const double *mL,
                    const double *mR, double *F);
void ora_mhd_riemann(const ora_mhd *M, double glm_ch, const double *UL, const double *UR, const double *nv,
                     const double *t1, const double *t2, double *F);
void ora_mhd_rhs(const ora_mhd *M, const double *U, double glm_ch, double *Ut);
/* FillFlux of GLM-MHD (fillflux.f90:41-110 with NONCONS): Flux_master, Flux_slave (9,0:N,0:N,nSides),
   which differ by each side's own Powell term */
void ora_mhd_fill_flux(const ora_mhd *M, const double *U, double glm_ch, double *FluxM, double *FluxS);
double ora_mhd_glm_dtch1(const ora_mhd *M, double CFLscale);
/* CALCTIMESTEP; DFLscale sets the viscous/resistive limit of the resistive case */
double ora_mhd_calc_timestep(const ora_mhd *M, const double *U, double CFLscale, double DFLscale,
                             double glm_dtch1, double *glm_ch, int *errType);
void ora_mhd_lserk_step(const ora_mhd *M, double *U, double *Ut_tmp, double glm_ch, double dt);
void ora_mhd_orszag_tang(const ora_mhd *M, int npts, const double *x, double *U);

/* ---- analysis of a state (src/analyze/analyze.f90:557-675, src/testcase/taylorgreenvortex/
   testcase_analyze.f90:127-401) ---- */
typedef struct {
  int exactFunc;          /* AnalyzeExactFunc: 0 none (no error norms), 1 constant refState,
                             2 density wave (5 variables) */
  int nAnalyze;           /* NAnalyze; 0 = the default 2(N+1) */
  double refState[9];     /* conservative state of ExactFunc 1 */
  double advVel[3], waveNumber[3], amplitude;   /* AdvVel, IniWaveNumber, IniAmplitude */
  double R;               /* gas constant 'R' (mean temperature) */
  double dissFactor;      /* factor of DR_u, DR_S, DR_Sd (the MHD testcase uses mu_0) */
} ora_analyze_cfg;
typedef struct {
  double vol;
  double L2[9], L2colloc[9], Linf[9];
  double bulk[9], bulk_t[9], maxabsUt[9];
  double Ekin, Ekin_comp, Emag_comp, Enstrophy_comp, Mag_Enstrophy, Entropy_comp;
  double DR_u, DR_S, DR_Sd, DR_p, max_Vorticity, max_Current, mean_temperature, uprime;
} ora_analysis;
/* nVar 5 (Euler / Navier-Stokes) or 9 (GLM-MHD, mu_0 its permeability); gx/gy/gz the lifted
   primitive gradients or NULL (zero velocity and B gradients, as without PARABOLIC) */
void ora_analyze(int N, int nVar, int nElems, const double *Elem_xGP, const double *sJ, const double *U,
                 const double *Ut, const double *gx, const double *gy, const double *gz, double kappa,
                 double mu_0, double t, const ora_analyze_cfg *c, ora_analysis *o);

/* Carpenter RK4-5 coefficients and the LGL CFL/DFL alpha tables
   (src/timedisc/timedisc_vars.f90:79-90, 204-233) */
void ora_rk_coefs(double *A, double *b, double *c);
double ora_cfl_alpha(int N);
double ora_dfl_alpha(int N);

#ifdef __cplusplus
}
#endif
#endif
