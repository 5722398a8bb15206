/*
 * fluxo_gpu.h -- C ABI of the MI355X split-form DGSEM right-hand-side engine.
 *
 * The engine replaces the FP64 CPU path of the reference's spatial operator and
 * low-storage Runge-Kutta stage.  Each entry point names the reference routine
 * (project-fluxo/fluxo, file:line) whose work it takes over.  The Fortran host
 * binds these symbols with ISO_C_BINDING (fluxo_amd/fortran/fluxo_gpu_bind.f90).
 *
 * Conventions
 *  - Return 0 on success or a negative FX_E* code; fx_last_error() gives text.
 *    The library never aborts the process; the Fortran wrapper maps a non-zero
 *    code to Abort() like the reference's CollectiveStop/Abort (globals.f90:161-201).
 *  - Host arrays use the reference's layout (Fortran column-major, first index
 *    fastest) and are borrowed only for the duration of the call.
 *  - Device memory belongs to the context.  A context is used by one host thread
 *    at a time; work is ordered on the context's HIP stream.
 *  - Element and side ids inside ElemToSide/SideToElem are 1-based, as in the
 *    reference (mesh_vars.f90; S2E_ and E2S_ positions in defines.h:75-84).
 */
#ifndef FLUXO_GPU_H
#define FLUXO_GPU_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
  FX_OK = 0,
  FX_EINVAL = -1,    /* invalid argument or unsupported option */
  FX_EHIP = -2,      /* HIP runtime error (message in fx_last_error) */
  FX_ENOMEM = -3,    /* device allocation failed */
  FX_ESTATE = -4,    /* call out of order (e.g. rhs before mesh upload) */
  FX_ENODEV = -5     /* no usable gfx950 device */
};

enum { FX_PREC_FP64 = 0, FX_PREC_FP32_SPLIT = 1 };

/* Options that the reference reads from parameter.ini / CMake defines. */
typedef struct {
  int N;            /* polynomial degree, 'N' (interpolation.f90:125-126); 2..5 */
  int nVar;         /* PP_nVar: 5 (Euler / Navier-Stokes) or 9 (ideal GLM-MHD) */
  int nElems;       /* local elements */
  int nSides;       /* sides: the inner sides, after the big mortar sides of a mesh with mortars */
  int volumeFlux;   /* 'VolumeFlux' (navierstokes/equation.f90:414-460): 0 standard DG, 1
                       standard DG with averaged metrics, 2 / 8 Kennedy-Gruber, 3 Ducros,
                       4 Morinishi, 5 Chandrashekar, 6 EC with arithmetic means, 7 gg,
                       9 Gassner-Winters-Walch, 10 Ismail-Roe, 32 Ranocha (FP32-split:
                       0, 5, 10, 32 only); MHD 10 (Derigs) */
  int riemann;      /* 'Riemann' (navierstokes/equation.f90:291-404): 0 central, 1 LLF,
                       9 / 32 EC (Chandrashekar / Ranocha), 4 / 44 / 33 entropy stable
                       (Ismail-Roe / Chandrashekar / Ranocha + matrix dissipation),
                       16 EC-KEP + face LLF, 5-8 / 10-12 / 20 a two-point flux in the
                       normal direction (Ismail-Roe, Kennedy-Gruber 1, Ducros, Morinishi,
                       EC2, gg, Kennedy-Gruber 2, Gassner-Winters-Walch), 14 / 15 / 17 / 18 /
                       19 / 21 the same + face LLF, 13 Ismail-Roe + entropy-Jacobian face
                       LLF, 2 HLLC, 22 HLL, 23 HLLE, 24 HLLEM, 3 Roe (central part from
                       VolumeFlux 2 / 3 / 4 as the reference's split form does); MHD
                       (mhd/equation.f90:327-386): 0 central,
                       1 Rusanov, 10 Derigs EC + entropy-stable LLF, 11 Derigs EC */
  int parabolic;    /* PARABOLIC (CMakeLists.txt:253): 0 Euler / ideal MHD, 1 Navier-Stokes /
                       resistive MHD */
  int lifting;      /* PP_Lifting with entropy-variable lifting: 1 = BR1, 2 = BR2
                       (src/dg/br1, src/dg/br2; Navier-Stokes and resistive GLM-MHD) */
  int precision;    /* FX_PREC_FP64 or FX_PREC_FP32_SPLIT */
  int whichMortar;  /* 'whichMortar' (mortar.f90:88-104), meshes with mortars: 0 L2 projection of
                       the small sides' fluxes, 1 collocation */
  double kappa;     /* 'kappa' (equation.f90:218) */
  double mu0;       /* constant dynamic viscosity: 'mu0' (navierstokes/equation.f90:229-230),
                       'mu' for MHD (mhd/equation.f90:190) */
  double Pr;        /* 'Pr' (equation.f90:226) */
  double etaBR2;    /* 'etaBR2' BR2 penalty (br2.f90:116, default 2) */
  /* NFVSE shock capturing (SHOCK_NFVSE, CMakeLists.txt:316-331; parameters of InitNFVSE
     shockcapturing/nfvse.f90:138-146), Euler or Navier-Stokes: sub-cell first-order finite volumes
     (SubFVMethod=1) blended into the volume integral of each element with the coefficient
     of CalcBlendingCoefficient (nfvse.f90:1537-1615) and TimeRelFactor=0; with
     parabolic = 1 the viscous volume term is added unblended (dg.f90:368-398).  The sub-cell
     interfaces take any point-wise Riemann solver (not 13-19, 21, whose face-wide wave speed
     would span a sub-cell interface plane). */
  int nfvse;           /* 0 off, 1 on */
  int computeAlpha;    /* 'ComputeAlpha': 1 Persson-Peraire indicator on rho*p
                          (ModalThreshold=1), 3 fixed 'ShockBlendCoef' */
  int spacePropSweeps; /* 'SpacePropSweeps' */
  int reserved3;
  double alpha_max;       /* 'alpha_max' */
  double alpha_min;       /* 'alpha_min' */
  double spacePropFactor; /* 'SpacePropFactor' */
  double shockBlendCoef;  /* 'ShockBlendCoef' */
  /* GLM-MHD (nVar = 9: FLUXO_EQNSYSNAME=mhd with FLUXO_EQN_GLM and Powell
     non-conservative terms, CMakeLists.txt:174-209; mhd/equation.f90:174-208):
     VolumeFlux 10 (Derigs), Riemann 0, 1, 10 or 11, no shock capturing; parabolic = 1 adds
     viscosity mu0, resistivity eta and heat conduction mu0*kappa/Pr with BR1 or BR2 lifting
     of the entropy variables (mhd/flux.f90:449-573). */
  double mu_0;            /* 'mu_0' magnetic permeability */
  double GLM_scale;       /* 'GLM_scale' */
  double eta;             /* 'eta' resistivity (mhd/equation.f90:187) */
  double reserved4;
} fx_config;

/* Mesh data produced by the host's InitMesh/CalcMetrics (mesh.f90:100-300,
   metrics.f90:78-483), all FP64 except the connectivity. */
typedef struct {
  const int *ElemToSide;          /* (2,6,nElems)  [SideID, flip] per local side */
  const int *SideToElem;          /* (5,nSides)    [Elem, nbElem, locSide, nbLocSide, flip] */
  const double *Metrics_fTilde;   /* (3,0:N,0:N,0:N,nElems) */
  const double *Metrics_gTilde;   /* (3,0:N,0:N,0:N,nElems) */
  const double *Metrics_hTilde;   /* (3,0:N,0:N,0:N,nElems) */
  const double *sJ;               /* (0:N,0:N,0:N,nElems) */
  const double *NormVec;          /* (3,0:N,0:N,nSides) */
  const double *TangVec1;         /* (3,0:N,0:N,nSides) (accepted for interface parity) */
  const double *TangVec2;         /* (3,0:N,0:N,nSides) (accepted for interface parity) */
  const double *SurfElem;         /* (0:N,0:N,nSides) */
  /* Non-conforming interfaces (fillMeshInfo prepare_mesh.f90:700-718; fxio_mesh_mortars):
     the sides 1..nMortarSides are big mortar sides, each with 4 (MortarType 1) or 2 (2: halves
     in eta, 3: halves in xi) small sides, whose master data are interpolated from the big
     side (U_Mortar) and whose fluxes are projected back onto it (Flux_Mortar).  NULL / 0 on
     a conforming mesh.  Euler, Navier-Stokes and GLM-MHD (the small sides' master fluxes
     with their Powell terms are projected), BR1 or BR2 lifting (Flux_Cons_Mortar of the
     lifting fluxes, U_Mortar of the gradients: br1.f90:217-322, br2.f90:220-290), NFVSE
     shock capturing (the blending coefficient propagates across mortars as Alpha_Mortar /
     PropagateBlendingCoeff do, nfvse_mpi.f90:78-240). */
  const int *MortarType;          /* (2,nSides) */
  const int *MortarInfo;          /* (2,0:4,nMortarSides) */
  int nMortarSides;
  int reserved_m;
} fx_mesh;

typedef struct fx_ctx fx_ctx;

/* InitDG/InitEquation/InitLifting equivalent (dg.f90:61-175, equation.f90:162-470). */
int fx_ctx_create(const fx_config *cfg, fx_ctx **out);
void fx_ctx_destroy(fx_ctx *ctx);
const char *fx_last_error(void);

/* Mesh arrays the reference keeps in MOD_Mesh_Vars (mesh_vars.f90). */
int fx_mesh_upload(fx_ctx *ctx, const fx_mesh *mesh);

/* U(nVar,0:N,0:N,0:N,nElems) of MOD_DG_Vars (dg_vars.f90:75-80). */
int fx_state_upload(fx_ctx *ctx, const double *U);
int fx_state_download(fx_ctx *ctx, double *U);
int fx_ut_download(fx_ctx *ctx, double *Ut);

/* DGTimeDerivative_weakForm(tIn) (dg.f90:282-452): Ut = RHS(U).  Ut is kept on the
   device; if Ut_host is not NULL it is also copied back. */
int fx_rhs(fx_ctx *ctx, double t, double *Ut_host);

/* VolInt_SplitForm (volint.f90:173-218) contribution alone, before Jacobian
   scaling: Ut_host = sum_l Dvolsurf_T * {f~,g~,h~}. */
int fx_volint(fx_ctx *ctx, double *Ut_host);

/* Lifting (br1.f90:154-364): primitive-variable gradients gradPx/y/z
   (nVar,0:N,0:N,0:N,nElems). */
int fx_lifting(fx_ctx *ctx, double *gradPx, double *gradPy, double *gradPz);

/* FillFlux(Flux_master, Flux_slave, doMPISides=.FALSE.) (dg/fillflux.f90:41-110) on the device
   state: the numerical flux of every side times SurfElem, Flux_master and Flux_slave
   (nVar,0:N,0:N,nSides) (all sides are inner sides on a periodic conforming mesh, so
   firstSlaveSide = 1 and LastSlaveSide = nSides).  Flux_slave = Flux_master, plus each side's
   own Powell term for GLM-MHD (AddNonConsFlux).  With parabolic = 1 the lifting runs first and
   the BR1 viscous flux is included, as in the reference's FillFlux after Lifting.  Either
   output may be NULL. */
int fx_fillflux(fx_ctx *ctx, double *Flux_master, double *Flux_slave);

/* SurfInt(Flux_master, Flux_slave, Ut, doMPISides=.FALSE.) (dg/surfint.f90:48-264, Gauss-Lobatto
   branch :177-184, 253-259): Ut(nVar,0:N,0:N,0:N,nElems) += Flux_master L_HatMinus0 at the
   master element's face nodes and -= Flux_slave L_HatMinus0 at the slave element's, nodes on
   several faces in the reference's side order.  Ut is read and written (host arrays). */
int fx_surfint(fx_ctx *ctx, const double *Flux_master, const double *Flux_slave, double *Ut);

/* VolInt_NFVSE(Ut) (shockcapturing/nfvse.f90:475-561) after VolInt on the device state: the
   blending coefficient (CalcBlendingCoefficient + PropagateBlendingCoeff), the first-order
   sub-cell finite-volume term F_FV, and Ut = (1 - alpha) VolInt_SplitForm + alpha F_FV per
   element (the advective volume integral, before -sJ); needs nfvse = 1. */
int fx_volint_nfvse(fx_ctx *ctx, double *Ut_host);

/* CalcBlendingCoefficient + PropagateBlendingCoeff (shockcapturing/nfvse.f90:1537-1615,
   nfvse_mpi.f90:31-153) on the device state: alpha(nElems), as the RHS uses it. */
int fx_blending(fx_ctx *ctx, double *alpha);

/* CALCTIMESTEP (navierstokes/calctimestep.f90:39-170) with the CFL/DFL scaling of
   FillCFL_DFL (timedisc.f90:697-741): CFLscale/DFLscale are the parameter.ini
   values.  errType as in the reference: 1 NaN density, 2 p<=0, 3/4 NaN dt. */
int fx_calc_dt(fx_ctx *ctx, double CFLscale, double DFLscale, double *dt, int *errType);

/* One TimeStepByLSERKW2 step (timedisc.f90:424-480), Carpenter RK4-5, fixed dt. */
int fx_rk_step(fx_ctx *ctx, double t, double dt);

/* The TimeDisc loop body (timedisc.f90:274-413) for nSteps steps with dt from
   CALCTIMESTEP before every step, entirely on the device (one HIP graph per step).
   On return *t and *dt hold the final time and last time step, *errType the first
   CALCTIMESTEP error (0 if none). */
int fx_run(fx_ctx *ctx, int nSteps, double CFLscale, double DFLscale, double *t, double *dt,
           int *errType);

/* The TimeDisc loop (timedisc.f90:274-413) up to the end time tEnd, on the device: like fx_run,
   with the reference's last step (timedisc.f90:313-317): when the CFL time step reaches tEnd
   within 1e-4 of itself, the step is shortened to end exactly at tEnd and is the last one.
   Runs at most maxSteps steps (maxSteps < 0: no limit, the reference's maxIter = -1) and stops
   early on a CALCTIMESTEP error.  *t is the start time on entry and the final time on return
   (tEnd when it was reached); *steps the steps taken, *dt the last step's dt. */
int fx_run_to(fx_ctx *ctx, int maxSteps, double CFLscale, double DFLscale, double tEnd, double *t, double *dt,
              int *steps, int *errType);

/* CalcEntropy (navierstokes/analyze_equation.f90:549-599) on the device state and
   the last Ut: S = int -rho s/(kappa-1), dSdU_Ut = int dS/dU . Ut. */
int fx_entropy(fx_ctx *ctx, double *S, double *dSdU_Ut);

/* Analysis of the device state (Analyze, src/analyze/analyze.f90:370-540).  Options follow
   the reference's parameters. */
typedef struct {
  int exactFunc;          /* 'AnalyzeExactFunc' (analyze.f90:113): 0 no error norms, 1 the
                             constant refState, 2 density wave (ExactFunc 2 of
                             navierstokes/equation.f90:632-651; 5 variables only) */
  int nAnalyze;           /* 'NAnalyze' (analyze.f90:122); 0 = the default 2(N+1); at most 16 */
  double refState[9];     /* ExactFunc 1: conservative state (RefStateCons) */
  double advVel[3];       /* 'AdvVel' */
  double waveNumber[3];   /* 'IniWaveNumber' */
  double amplitude;       /* 'IniAmplitude' */
  double R;               /* 'R' gas constant, for the mean temperature (reference 287.058) */
  double dissFactor;      /* factor of DR_u, DR_S, DR_Sd (the reference's MHD Taylor-Green
                             testcase multiplies by mu_0) */
} fx_analyze_config;

/* Results, normalised as the reference prints them: CalcErrorNorms (analyze.f90:557-628: L2
   on the NAnalyze Gauss-Lobatto points, L2 on the collocation points, Linf; zero when
   exactFunc = 0), the domain volume, CalcBulk (:633-675), and the Taylor-Green vortex
   testcase analysis (testcase/taylorgreenvortex/testcase_analyze.f90:127-401; magnetic
   terms zero for 5 variables; gradient terms zero without PARABOLIC). */
typedef struct {
  double vol;
  double L2[9], L2colloc[9], Linf[9];
  double bulk[9], bulk_t[9], maxabs_Ut[9];
  double Ekin, Ekin_comp, Emag_comp, Enstrophy_comp, Mag_Enstrophy, Entropy_comp;
  double DR_u, DR_S, DR_Sd, DR_p, max_Vorticity, max_Current, mean_temperature, uprime;
} fx_analysis;

/* Elem_xGP(3,0:N,0:N,0:N,nElems) of CalcMetrics (metrics.f90), kept on the device for the
   exact functions of fx_analyze. */
int fx_elem_xgp_upload(fx_ctx *ctx, const double *Elem_xGP);

/* Analyze (analyze.f90:370-540) at time t: evaluates the right-hand side of the device state
   (DGTimeDerivative, and the lifted gradients when parabolic) and reduces everything on the
   device; only the fx_analysis record crosses to the host. */
int fx_analyze(fx_ctx *ctx, double t, const fx_analyze_config *cfg, fx_analysis *out);

/* GLM-MHD cleaning state: GLM_ch (0.5 until the first CALCTIMESTEP, then
   GLM_scale*GLM_dtch1/dt of the last one, mhd/calctimestep.f90:175-177) and GLM_dtch1. */
int fx_glm(fx_ctx *ctx, double *glm_ch, double *glm_dtch1);

/* Wait for all work queued on the context's stream. */
int fx_synchronize(fx_ctx *ctx);

/* Average duration (ms) of one launch of a single engine kernel, timed with HIP events on
   the context's stream over `reps` back-to-back launches on the current device state.
   which: FX_KERNEL_*.  The volume kernel is timed in its LSERK form (stage >= 2 traffic)
   and leaves U advanced by `reps` partial stages.  No counterpart in the reference; it
   gives the benchmark per-kernel durations for the roofline figures. */
enum { FX_KERNEL_VOLUME_RK = 0, FX_KERNEL_SURFACE = 1, FX_KERNEL_LIFTING = 2, FX_KERNEL_DT = 3,
       FX_KERNEL_VOLUME_RHS = 4, FX_KERNEL_NFVSE_ALPHA = 5, FX_KERNEL_NFVSE_FV = 6 };
int fx_profile_kernel(fx_ctx *ctx, int which, int reps, double *avg_ms);

/* Tile tuner of the volume kernel (no counterpart in the reference).  Times each compiled
   tile (elements per workgroup) of the LSERK volume kernel on this context's mesh and
   state, restores U and Ut_tmp, and keeps the fastest for this context and as the
   process-wide choice for its (N, parabolic, precision), used by contexts and server pools
   created afterwards.
   tiles[t] / ms[t] (room for 4) receive each candidate's elements per workgroup and average
   launch time, *ntiles their number, *best_epb the choice.  Only the Ranocha (VolumeFlux
   32) form has alternatives; other fluxes report their single tile. */
int fx_tune_volume(fx_ctx *ctx, int reps, int *best_epb, int *tiles, double *ms, int *ntiles);
/* The process-wide volume-kernel tile (elements per workgroup) for (N, parabolic, precision). */
int fx_volume_tile(int N, int parabolic, int precision, int *epb);
/* Use the tile with epb elements per workgroup for this context (FX_EINVAL if none). */
int fx_use_volume_tile(fx_ctx *ctx, int epb);

/* Device bytes held by a context with this configuration (for HBM admission). */
long long fx_ctx_bytes(const fx_config *cfg);

/* ---- Right-hand side without allocation ---------------------------------------------
   No counterpart in the reference.  For callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
   right-hand side of fx_rhs (DGTimeDerivative_weakForm, dg.f90:282-452) with every input and
   output a DEVICE pointer owned by the caller -- the fx_mesh arrays (TangVec1/2 unused), U and
   Ut, in the layouts above -- and no hipMalloc, hipFree or other allocation.  The face
   fluxes, packed normals, viscous terms and NFVSE buffers live in one library-owned static
   device array of FX_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; configurations
   that need more (fx_rhs_noalloc_bytes) are rejected with FX_EINVAL.  Every call on one GPU
   shares that array and overwrites it: calls that overlap on one GPU (other streams or host
   threads) race.  Euler / Navier-Stokes (nVar 5), FP64, with or without NFVSE; GLM-MHD and
   FP32-split are rejected.  Work is queued on `stream` (a hipStream_t; NULL the default
   stream); the call waits once on that stream for the device-side connectivity check, then
   returns with the right-hand side queued. */
#define FX_NOALLOC_SCRATCH_BYTES (64LL * 1024 * 1024)
long long fx_rhs_noalloc_bytes(const fx_config *cfg);
int fx_rhs_noalloc(const fx_config *cfg, const fx_mesh *mesh_dev, const double *U_dev, double *Ut_dev,
                   void *stream);

/* ---- Batched serving ---------------------------------------------------------------
   No counterpart in the reference, which runs one solve per process (fluxo.f90:180 ->
   TimeDisc, timedisc.f90:144-415); a request here is that whole run: nSteps steps of
   TimeStepByLSERKW2 with CALCTIMESTEP before each, on one periodic conforming mesh.
   Requests of mixed N, mesh size and physics share per-(N, parabolic) element pools in
   HBM; each pool advances all its requests with one HIP graph per time step (every kernel
   of the step covers the pool's elements/sides of all requests; each request keeps its own
   dt from its own CALCTIMESTEP minimum).  A worker thread admits queued requests between
   steps as the HBM budget and pool slots allow, so requests arriving while others run join
   the running batch.  A request's result is bitwise the same as the same run through
   fx_ctx_create/fx_run. */
typedef struct fx_server fx_server;

typedef struct {
  int volumeFlux;      /* as fx_config */
  int riemann;
  int precision;       /* FX_PREC_FP64 or FX_PREC_FP32_SPLIT */
  int lifting;         /* 1 = BR1 for the parabolic requests */
  double kappa, mu0, Pr;
  double CFLscale, DFLscale;
  long long hbmBudget; /* device bytes the admitted requests may hold (fx_request_bytes) */
  int maxRequests;     /* request ids available over the server's life */
  int admitKNodes;     /* 0: admit as the HBM budget allows; > 0: the admitted requests hold at
                          most this many thousand volume nodes (a request alone is always
                          admitted), and a request beyond the cap is not passed by later ones */
} fx_server_config;

typedef struct {
  int N;               /* 2..5 */
  int parabolic;       /* 0 Euler, 1 Navier-Stokes */
  int nElems, nSides;
  int nSteps;          /* time steps to run */
  int reserved;
  fx_mesh mesh;        /* host arrays as for fx_mesh_upload; borrowed until the request completes */
  const double *U0;    /* (5,0:N,0:N,0:N,nElems), borrowed until the request completes */
  double *U_out;       /* (5,0:N,0:N,0:N,nElems), written when the request completes */
} fx_request;

enum { FX_REQ_QUEUED = 0, FX_REQ_RUNNING = 1, FX_REQ_DONE = 2, FX_REQ_FAILED = 3 };

typedef struct {
  int state;           /* FX_REQ_* */
  int errType;         /* first CALCTIMESTEP error (0 none); a request with an error fails */
  int steps;           /* steps done */
  int reserved;
  double t_end, dt_last;                /* simulation time and last dt */
  double t_submit, t_admit, t_done;     /* seconds since fx_server_create (steady clock) */
} fx_request_status;

typedef struct {
  long long steps_launched;   /* pool step graphs launched */
  long long graph_captures;   /* step graphs instantiated (a pool's first, or an update refused) */
  long long bytes_in_use;     /* fx_request_bytes of the admitted requests */
  long long bytes_peak;
  long long bytes_allocated;  /* device memory held by the pools */
  int active, queued, done, failed;
  long long graph_updates;    /* executable step graphs updated in place after pool extent changes */
} fx_server_stats;

int fx_server_create(const fx_server_config *cfg, fx_server **out);
/* Queue a request (never waits for GPU work); *id identifies it. */
int fx_server_submit(fx_server *srv, const fx_request *req, int *id);
int fx_server_status(fx_server *srv, int id, fx_request_status *st);
/* Wait until request id (or every submitted request if id < 0) has completed or failed,
   at most timeout_s seconds; FX_ESTATE on time-out. */
int fx_server_wait(fx_server *srv, int id, double timeout_s);
int fx_server_get_stats(fx_server *srv, fx_server_stats *out);
/* Text of the last error met by the server's worker thread ("" if none): why a request
   failed other than through CALCTIMESTEP's errType. */
const char *fx_server_last_error(fx_server *srv);
/* Stops the worker after its current step; requests not completed are marked failed. */
void fx_server_destroy(fx_server *srv);
/* Device bytes one request holds in its pool (the admission charge). */
long long fx_request_bytes(const fx_server_config *cfg, int N, int parabolic, int nElems, int nSides);

#ifdef __cplusplus
}
#endif
#endif
