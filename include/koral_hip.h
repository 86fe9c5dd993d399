/*
 * koral_hip.h -- C-ABI of the MI355X (gfx950) implementation of KORAL's
 * per-timestep hot path.  Link libkoral_hip.so; no HIP or torch types cross
 * this boundary (streams are passed as void*).
 *
 * Two families of entry points:
 *
 *  1. Batched per-point operators on device-resident SoA arrays.  Point i of
 *     an n-point batch keeps variable iv at x[iv*n + i]; metric component c at
 *     g20[c*n + i].  These replace the per-cell functions the reference calls
 *     in loops (p2u, u2p/calc_primitives, u2p_rad, solve_implicit_lab).
 *
 *  2. Grid operators bound to a koral_hip_grid (device-resident state with
 *     ghosts), replacing op_explicit / op_implicit / calc_u2p and the RK2IMEX
 *     step body of ko.c.  See the grid section below.
 *
 * All calls return 0 on success and a negative KORAL_E_* code on error; the
 * text of the last error is available from koral_hip_last_error().  Physics
 * failures are never errors: they surface as cell flags and counters, as in
 * the reference (mnemonics.h:105-119).
 */
#ifndef KORAL_HIP_H
#define KORAL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define KORAL_E_ARG -1
#define KORAL_E_HIP -2
#define KORAL_E_STATE -3

/* Resolved compile-time configuration of a KORAL build (choices.h defaults and
   the problem's define.h).  Mirrors koral_amd/params.py. */
typedef struct koral_params {
  int nv, mag, rad, pad0;           /* NV, MAGNFIELD, RADIATION */
  double gamma;                     /* GAMMA */
  double k_boltz, m_proton, sigma_rad, mu_gas; /* ko.h:40-46 in code units */
  double ggg, ccc, masscm;          /* ko.h:9-15 */
  int kappa_model;                  /* 0: kappa=a*rho; 1: Kramers (cgs coefficient a) */
  int polar_nc;                     /* NCCORRECTPOLAR with CORRECT_POLARAXIS (spherical-like
                                       coordinates), 0 without */
  double kappa_a_coef, kappa_es_coef;
  double rhofloor, uurhoratiomin, uurhoratiomax, b2rhoratiomax;
  double gammamaxhd, gammamaxrad, eradfloor;
  double eerhoratiomin, eerhoratiomax, eeuuratiomin, eeuuratiomax;
  double u2pconv;                   /* U2PCONV */
  double radimpconv, radimpentrconv, radimpeps;
  int radimpmaxiter, convrel_enabled; /* RADIMPMAXITER, BHDISK_PROBLEMTYPE */
  double dampingfactor, maxenchangedown, maxenchangeup, maxradimpdamping;
  double radimplicitthreshold;
  double radimpconvrel, radimpconvrelerr, radimpconvrelentr, radimpconvrelentrerr;
  int int_order, fluxmethod;        /* INT_ORDER (1 minmod, 2 PPM, 4 MP5), FLUXMETHOD */
  double tsteplim, minmod_theta;
  int dofixups, dou2pmhdfixups, dou2pradfixups, doradimpfixups;
} koral_params;

/* Library and device information. */
const char *koral_hip_last_error(void);
int koral_hip_version(void);
int koral_hip_params_size(void);

/* ---- batched per-point operators (device pointers, SoA, asynchronous on
   `stream`; synchronise before reading results) ---- */

/* p2u (p2u.c:40-49): uu = U(pp) */
int koral_hip_p2u_batch(const koral_params *P, long n, const double *pp, double *uu,
                        const double *g20, const double *G20, void *stream);

/* calc_primitives (u2p.c:13-133) on independent points: u2p (u2p.c:176-485)
   with hot/entropy fallback, M1 inversion, check_floors_mhd/rad, then p2u.
   uu is updated to follow floored primitives; pp holds the guess on entry.
   flags[f*n+i] for f in 0..6 (mnemonics.h cell flags); ret[i] = u2p return. */
int koral_hip_u2p_batch(const koral_params *P, long n, double *uu, double *pp,
                        const double *g20, const double *G20, int *ret, int *flags,
                        void *stream);

/* u2p_solver (u2p.c:1219-1235 -> u2p_solver_W u2p.c:1965-2352), etype 0 hot
   or 1 entropy; pp is the guess on entry; ret[i] is the solver return code. */
int koral_hip_u2p_solver_batch(const koral_params *P, long n, const double *uu, double *pp,
                               const double *g20, const double *G20, int etype, int *ret,
                               void *stream);

/* u2p_solver_W from a learned W start (no reference counterpart): W0[i] > 0
   starts the Newton iteration in W instead of the start computed from pp
   (u2p.c:2094-2112), under the unchanged tolerance; a start that fails is
   retried from the plain start.  iters[i] (may be NULL) counts the Newton
   iterations of both tries, guessed[i] (may be NULL) is 1 when the learned
   start solved the point.  W0 NULL: the plain solver, with iteration counts. */
int koral_hip_u2p_solver_batch_from(const koral_params *P, long n, const double *uu, double *pp,
                                    const double *g20, const double *G20, int etype, const double *W0, int *ret,
                                    int *iters, int *guessed, void *stream);

/* u2p_rad (u2prad.c:485-511): M1 inversion of R^t_mu; corr[i] = corrected */
int koral_hip_u2p_rad_batch(const koral_params *P, long n, const double *uu, double *pp,
                            const double *g20, const double *G20, int *ret, int *corr,
                            void *stream);

/* calc_Gi (rad.c:1787-1888): contravariant four-force, out[k*n+i] */
int koral_hip_Gi_batch(const koral_params *P, long n, const double *pp, const double *g20,
                       const double *G20, int labframe, double *Gi, void *stream);

/* implicit_lab_rad_source_term / solve_implicit_lab (rad.c:2758-2795,
   rad.c:1225-1570) on independent cells: full six-attempt fallback chain of
   the damped 4-D Newton solve_implicit_lab_4dprim (rad.c:376-1218).
   uu/pp updated in place on success; flags as above (RADIMPFIXUPFLAG,
   RADFIXUPFLAG, RADSOURCETYPEFLAG); info[k*n+i]: k=0 iterations of the
   successful attempt, 1 attempt index (-1 failed), 2 whichprim, 3 whicheq,
   4 whichframe, 5 iterations over all attempts. */
int koral_hip_implicit_batch(const koral_params *P, long n, double *uu, double *pp,
                             const double *g20, const double *G20, const double *dt,
                             double global_time, int *flags, int *info, void *stream);

/* koral_hip_implicit_batch without any allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same arguments, kernels and results as koral_hip_implicit_batch, but the
   work lists and counters live in a fixed device array of the library (16 MiB
   per GPU, held until the process ends), so n may not exceed
   KORAL_HIP_NOALLOC_MAX_CELLS (larger batches return KORAL_E_ARG).  Every call
   on one GPU shares that array: calls that overlap on one GPU (other streams
   or threads) race and must be serialised by the caller; calls in stream
   order are safe (each clears the counters it uses). */
#define KORAL_HIP_NOALLOC_MAX_CELLS (1L << 20)
int koral_hip_implicit_batch_noalloc(const koral_params *P, long n, double *uu, double *pp,
                                     const double *g20, const double *G20, const double *dt,
                                     double global_time, int *flags, int *info, void *stream);

/* ---- learned initial guess of the implicit solve (no reference
   counterpart: KORAL starts the Newton solve from the old state) ----

   A small network (11 -> 128 -> 128 -> 128 -> 4, GELU) maps the local state
   of a cell to corrections of the Newton start.  kind 0 (gas) corrects
   (u, u~^1..3), kind 1 (radiation) corrects (E, F~^1..3), kind 2 predicts the
   W start of u2p (below), kind 3 (gate) scores in its first output whether
   the learned tries will solve a cell (+1 yes, -1 no, least squares on the
   gas features).  Parameters are
   float32 [koral_hip_mlp_nparam()] in the order W1[128][11], b1[128],
   W2[128][128], b2, W3[128][128], b3, W4[4][128], b4 (W[out][in]); norm is
   float32 [30]: feature mean[11], std[11], target mean[4], std[4].
   Features (feat[k*n+i], standardised when a network is given): log10 rho,
   log10(u/rho), log10(E/u), gas u~^1..3, radiation u~^1..3,
   log10(dt kappa_abs), log10(dt kappa_es).  Targets (targ[k*n+i]) between
   converged (pp1) and initial (pp0) primitives x of the kind:
   log(x1_0/x0_0), asinh((x1_k - x0_k)/1e-6) for k = 1..3.  Training batches
   are [11][B] / [4][B] with B <= 4096. */
typedef struct koral_hip_mlp koral_hip_mlp;
int koral_hip_mlp_nparam(void);
int koral_hip_mlp_create(int kind, koral_hip_mlp **out);
int koral_hip_mlp_destroy(koral_hip_mlp *m);
int koral_hip_mlp_kind(const koral_hip_mlp *m);
int koral_hip_mlp_set(koral_hip_mlp *m, const float *theta, const float *norm);  /* host arrays; resets Adam */
int koral_hip_mlp_get(koral_hip_mlp *m, float *theta, float *norm);              /* host arrays, may be NULL */
int koral_hip_mlp_features(const koral_params *P, const koral_hip_mlp *m, long n, const double *pp,
                           const double *dt, float *feat, void *stream);
int koral_hip_mlp_targets(int kind, long n, int nv, const double *pp0, const double *pp1, float *targ,
                          void *stream);
/* The W network (kind 2; no reference counterpart): a start for the W
   inversion of u2p from the conserved state uu and the guess pp.  Features:
   log10(u/rho), log10(b^2/rho + 1e-30), u~^i sqrt(g_ii) (3), log(D/D0),
   asinh((U_1 - U0_1)/|U0_1|), asinh((U_k - U0_k)/(|U0_0| + |U0_1|)) for
   k = 2..4, log10(W0/rho), where U0 = p2u(pp) and W0 = (rho + Gamma u)
   gamma^2 of pp.  Targets: log(W1/W0) between the solution pp1 and the guess
   pp0, and three zeros.  The guess is W0 exp(y0) (0, the plain start, where
   not finite or not positive). */
int koral_hip_mlp_u2p_features(const koral_params *P, const koral_hip_mlp *m, long n, const double *uu,
                               const double *pp, const double *g20, const double *G20, float *feat,
                               void *stream);
int koral_hip_mlp_u2p_targets(const koral_params *P, long n, const double *pp0, const double *pp1,
                              const double *g20, const double *G20, float *targ, void *stream);
int koral_hip_mlp_u2p_guess(const koral_params *P, const koral_hip_mlp *m, long n, const double *pp,
                            const double *g20, const double *G20, const float *Y, double *W0, void *stream);
/* FP32 forward pass of a training batch, Y [4][B] */
int koral_hip_mlp_forward(koral_hip_mlp *m, const float *X, int B, float *Y, void *stream);
/* loss (host, may be NULL; synchronises) and gradient (device [nparam], may be NULL) of the
   mean squared error over the batch */
int koral_hip_mlp_grad(koral_hip_mlp *m, const float *X, const float *T, int B, float *grad, float *loss,
                       void *stream);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on the mean squared error */
int koral_hip_mlp_train_step(koral_hip_mlp *m, const float *X, const float *T, int B, float lr, float *loss,
                             void *stream);
/* BF16 matrix-core inference: standardised outputs Y [4][n] */
int koral_hip_mlp_infer(koral_hip_mlp *m, const float *feat, long n, float *Y, void *stream);
/* Gate (kind 3) outputs Y applied to a gas start: gxg[i] (u of cell i) becomes NaN where the
   de-normalised score is below thr, so the learned solve skips cell i's learned tries */
int koral_hip_mlp_gate(koral_hip_mlp *m, long n, const float *Y, float thr, double *gxg, void *stream);
/* Newton start gx [4][n] of the network's kind from the outputs (the old state where not
   finite or not positive) */
int koral_hip_mlp_guess(koral_hip_mlp *m, long n, int nv, const double *pp, const float *Y, double *gx,
                        void *stream);
/* koral_hip_implicit_batch with a first attempt from the learned guess (energy equation, lab
   frame) under the unchanged convergence tolerance.  Without a radiation network (rad may be
   NULL) every cell iterates the gas primitives from the gas guess.  With one, a cell first
   iterates the primitives attempt 0 of the chain would (rad.c:1381 on the old state): the
   radiation primitives from the radiation guess (gas part from the gas guess) or the gas
   primitives from the gas guess; a cell that try does not solve tries the other primitives
   the same way.  Each try stops after maxiter iterations (RADIMPMAXITER when maxiter <= 0
   or larger).  Cells the learned tries do not solve run the unchanged six-attempt chain
   from the old state.  info[1*n+i] = 6 marks a cell solved by a learned try; info[5*n+i]
   counts the iterations of all attempts. */
/* The same first attempt from starts the caller provides: gxg [4][n] (u, u~^1..3) and gxr
   [4][n] (E, F~^1..3, may be NULL for gas primitives only), device arrays; a cell whose
   gxg[0*n+i] is NaN has no learned try (the chain starts at once, as in the plain solve). */
int koral_hip_implicit_batch_from(const koral_params *P, long n, double *uu, double *pp, const double *g20,
                                  const double *G20, const double *dt, double global_time, const double *gxg,
                                  const double *gxr, int maxiter, int *flags, int *info, void *stream);
/* koral_hip_u2p_solver_batch_from with the W start of a kind-2 network */
int koral_hip_u2p_solver_batch_learned(const koral_params *P, koral_hip_mlp *m, long n, const double *uu,
                                       double *pp, const double *g20, const double *G20, int etype, int *ret,
                                       int *iters, int *guessed, void *stream);
int koral_hip_implicit_batch_learned(const koral_params *P, koral_hip_mlp *gas, koral_hip_mlp *rad, long n,
                                     double *uu, double *pp, const double *g20, const double *G20,
                                     const double *dt, double global_time, int maxiter, int *flags, int *info,
                                     void *stream);
/* koral_hip_implicit_batch_learned with a gate network (kind 3): cells it scores below thr skip
   the learned tries and run the plain chain from the old state at once (their result is the
   plain solve's, bit for bit). */
int koral_hip_implicit_batch_learned_gated(const koral_params *P, koral_hip_mlp *gas, koral_hip_mlp *rad,
                                           koral_hip_mlp *gate, float thr, long n, double *uu, double *pp,
                                           const double *g20, const double *G20, const double *dt,
                                           double global_time, int maxiter, int *flags, int *info,
                                           void *stream);


/* ---- grid operators (device-resident state, one GPU holding the whole
   domain that KORAL's MPI tiling would split) ---- */

/* Grid description: interior sizes (1 = collapsed dimension, no ghosts), NG
   ghosts, boundary kind per side (0 copy, 1 periodic, 2 polar reflection of
   the normal vector components) ordered xlo,xhi,ylo,yhi,zlo,zhi, face
   coordinates xb[d][i+ng] for i in [-ng, n+ng], and host SoA metric tables
   over the ghost-inclusive (x1,x2) plane (KORAL's calc_metric output,
   metric.c:3866-4080, for an x3-independent metric):
     gc, Gc [20][SY][SX]   cell centres (struct geometry gg/GG)
     kr     [64][SY][SX]   Christoffel symbols Gamma^i_jk at centres
     gfx,Gfx [20][SY][SX+1], gfy,Gfy [20][SY+1][SX]   face metrics */
typedef struct koral_grid_desc {
  int nx, ny, nz, ng;
  int bc[6];
  const double *xb[3];
  const double *gc, *Gc, *kr, *gfx, *Gfx, *gfy, *Gfy;
} koral_grid_desc;

typedef struct koral_hip_grid koral_hip_grid;

int koral_hip_grid_create(const koral_params *P, const koral_grid_desc *D, koral_hip_grid **out);
int koral_hip_grid_destroy(koral_hip_grid *g);
int koral_hip_grid_set_stream(koral_hip_grid *g, void *stream);

/* Parity-mode boundary: KORAL's host arrays (ko.h:505-520 layout,
   [iz][iy][ix][iv] over the ghost-inclusive box; cellflag [..][7] ints).
   u and flags may be NULL. */
int koral_hip_grid_upload(koral_hip_grid *g, const double *p_aos, const double *u_aos, const int *flags_aos);
int koral_hip_grid_download(koral_hip_grid *g, double *p_aos, double *u_aos, int *flags_aos);
/* Resident mode: device SoA state x[iv][cell], flags int8 [7][cell]. */
int koral_hip_grid_device_state(koral_hip_grid *g, double **p, double **u, signed char **flags);

/* Diagnostics: the last op_explicit's work arrays, device SoA copied to the
   host.  which = 0: wave speeds [18][cell] (per dimension d: gas min/max at
   2d, 2d+1, radiation min/max at 6+2d, 7+2d, gas and radiation max |a| at
   12+d, 15+d); which = 1, 2, 3: face fluxes of direction x, y, z after
   flux-CT, [nv][face] with faces (SX+1)*SY*SZ, SX*(SY+1)*SZ, SX*SY*(SZ+1).
   No reference counterpart; used to localise parity differences. */
int koral_hip_grid_download_work(koral_hip_grid *g, int which, double *out);

/* Per-operator launch timing.  enable != 0 records an event pair around
   every following launch on the grid's stream.  When any output is given the
   call synchronises and returns, per operator kind k < nkinds, the summed
   milliseconds, the threads launched (cells, faces or edges) and the number
   of launches since the last call; the records are then cleared.  Kind names
   come from koral_hip_grid_op_name (14 kinds).  No reference counterpart
   (KORAL times whole steps only, ko.c:771). */
int koral_hip_grid_timing(koral_hip_grid *g, int enable, double *ms, long *units, int *launches, int nkinds);
/* Restricts the timing events to the operator kinds whose bit (1 << kind) is set in mask
   (all kinds by default), so one operator can be timed inside a timed sequence of steps. */
int koral_hip_grid_timing_mask(koral_hip_grid *g, unsigned mask);
const char *koral_hip_grid_op_name(int kind);

/* u = p2u(p) in every cell (as set_initial_profile leaves the state) */
int koral_hip_grid_p2u(koral_hip_grid *g);
/* set_bc (finite.c:2752-3700): ghost cells from the boundary kinds, then the
   domain corners */
int koral_hip_set_bc(koral_hip_grid *g);
/* calc_wavespeeds + save_wavespeeds reduction (finite.c:296-398, 568-599):
   returns max over the domain of the time-step density */
int koral_hip_tstepden(koral_hip_grid *g, double *tstepdenmax);
/* op_explicit (finite.c:647-1234); tstepdenmax may be NULL */
int koral_hip_op_explicit(koral_hip_grid *g, double t, double dt, double *tstepdenmax);
/* calc_u2p (finite.c:502-561) including cell_fixup and set_bc */
int koral_hip_calc_u2p(koral_hip_grid *g);
/* do_correct (finite.c:605-636): with koral_params.polar_nc > 0 (CORRECT_POLARAXIS,
   spherical-like coordinates) correct_polaraxis (finite.c:4924-5060) on the
   polar_nc cells next to each polar face; nothing otherwise */
int koral_hip_do_correct(koral_hip_grid *g);
/* op_implicit (finite.c:1242-1292) including cell_fixup(FIXUP_RADIMP) */
int koral_hip_op_implicit(koral_hip_grid *g, double t, double dt);
/* Body of one RK2IMEX step of ko.c:444-570, do_correct after every calc_u2p
   as ko.c:469-565 calls it.  tstepdenmax (next dt = 1/it)
   and counters[6] = {implicit successes, implicit Newton iterations,
   implicit failures, hd fixup requests, rad fixup requests, entropy
   inversions} may be NULL. */
int koral_hip_step_rk2imex(koral_hip_grid *g, double t, double dt, double *tstepdenmax, int *counters);
/* nsteps RK2IMEX steps of ko.c's loop (ko.c:399-570) with the time kept on the device: each
   step runs at (*t, *dt), then t += dt and dt = 1 / tstepdenmax of that step, as the
   reference's loop does, with no host round trip between steps.  On return *t, *dt are the
   time and step of the next step; counters[6] (may be NULL) sum the step counters. */
int koral_hip_run_rk2imex(koral_hip_grid *g, int nsteps, double *t, double *dt, int *counters);
/* fprint_restartfile_bin (fileop.c:911-982): folder/res%04d.head ("## nout1 nout2 t PROBLEM
   NX NY NZ", t as %10.6e) and folder/res%04d.dat (NX*NY*NZ int index triples, ix outer and iz
   inner, then NV doubles of primitives per cell in the same order) from the device state;
   folder/reslast.{head,dat} are linked to them. */
int koral_hip_grid_write_restart(koral_hip_grid *g, const char *folder, int nout1, int nout2, double t,
                                 int problem);
/* fread_restartfile_bin (fileop.c:1009-1108): reads res%04d (nout1 >= 0) or reslast
   (nout1 < 0) into the interior primitives and sets u = p2u(p) there; ghosts are left for
   set_bc.  Returns 1 without touching the grid when the header file does not exist (the
   reference then starts from scratch).  *t from the header; *nfout1 = file number + 1 and
   *nfout2 (may be NULL) as the reference sets nfout1, nfout2. */
int koral_hip_grid_read_restart(koral_hip_grid *g, const char *folder, int nout1, double *t, int *nfout1,
                                int *nfout2);
/* With enable != 0 (a stream must be set), koral_hip_run_rk2imex captures one step into a
   HIP graph on first use and replays it; enable = 0 drops the graph.  Per-operator timing,
   when on, runs the steps without the graph. */
int koral_hip_grid_use_graph(koral_hip_grid *g, int enable);

/* ---- launch choices and their autotuner (no reference counterpart: KORAL's
   CPU loops have no launch shape) ----

   The choices of one grid: threads per block of each operator kind (index as
   koral_hip_grid_op_name; 64, 128 or 256), per sweep direction whether the
   reconstruction runs inside the flux kernel (fused = 1) or writes the face
   states to HBM for a separate flux kernel (0), and the mapping of the
   deferred passes (the few faces needing the b^2/rho floor re-inversion and
   the cells needing the entropy inversion or that floor): one wave per face
   or cell (floor_wave = 1, the inversion's trials one per lane) or one lane
   per face or cell (0), with floor_nw / u2p_nw faces or cells in flight.
   Every choice computes the same results bit for bit. */
typedef struct koral_grid_tune {
  int block[32];
  int fused[3];
  int floor_wave;
  int update_pre;   /* update kernel: 1 loads the cell's u and p with the fluxes */
  int comb_nv;        /* 1: the stage-algebra kernel is compiled for the
                         grid's variable count */
  int u2p_budget;     /* residual evaluations the common u2p pass spends on
                         a cell's hot inversion before leaving the cell to
                         the deferred (wave-parallel) pass; 0: no bound */
  int u2p_group;      /* lanes per cell of the deferred u2p pass with
                         floor_wave = 1: 64 (one cell per wave) or 32 */
  long floor_nw, u2p_nw;
} koral_grid_tune;
int koral_hip_grid_get_tune(koral_hip_grid *g, koral_grid_tune *t);
/* Lengths of the deferred work lists (no reference counterpart; a
   diagnostic of the launch choices): counts[0..2] the faces of the x1, x2,
   x3 sweeps of the last koral_hip_op_explicit left to the b^2/rho floor
   pass, counts[3] the cells of the last u2p left to the full
   calc_primitives pass (also after a step; a step reuses the face counts).
   Synchronises the grid's stream. */
int koral_hip_grid_deferred_counts(koral_hip_grid *g, unsigned counts[4]);
/* Learned Newton starts for the grid's implicit operator (no reference
   counterpart; for callers that train the networks of the learned batch
   solve): every later op_implicit, and the implicit stages of the RK2IMEX
   step, solve each interior cell as koral_hip_implicit_batch_learned_gated
   does -- starts from the gas (and radiation) network on the old state,
   cells the gate (may be NULL) scores below thr start the plain chain at
   once, the learned tries limited to maxiter iterations, the unchanged
   chain for cells the tries leave.  gas NULL restores the plain solve.  The
   grid needs radiation; the networks must outlive their use. */
int koral_hip_grid_set_learned(koral_hip_grid *g, koral_hip_mlp *gas, koral_hip_mlp *rad, koral_hip_mlp *gate,
                               float thr, int maxiter);
/* The starts of the last learned implicit operator, [4][n] over the
   interior cells in x-fastest order (gxr may be NULL; NaN gas starts mark
   cells the gate left to the plain chain).  Synchronises. */
int koral_hip_grid_download_guess(koral_hip_grid *g, double *gxg, double *gxr);
int koral_hip_grid_set_tune(koral_hip_grid *g, const koral_grid_tune *t);
/* Coordinate search over the choices above on the grid's current state at
   time t with step dt: each candidate runs 1 + steps RK2IMEX steps from a
   snapshot of the state (restored afterwards, so the grid ends as it began),
   timed over the last `steps` with events; a change is kept when it is at
   least 1 % faster.  The chosen set stays on the grid.  report (may be NULL)
   receives one line per candidate: "name value ms_per_step". */
int koral_hip_grid_autotune(koral_hip_grid *g, double t, double dt, int steps, char *report, int report_len);

#ifdef __cplusplus
}
#endif
#endif
