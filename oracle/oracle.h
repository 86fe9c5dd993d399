/*
 * KORAL hot-path oracle: a plain-C, FP64, single-threaded restatement of the
 * reference algorithm for one GRRMHD+M1 timestep.
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU baseline).  The product path (koral_amd/) never links, loads or
 * calls anything in this directory.
 *
 * The reference (pseudotensor/koral) cannot be compiled in this image: every
 * translation unit includes ko.h, which includes 17 GSL headers (ko.h:77-94)
 * and GSL is not installed.  The oracle is therefore a restatement written from
 * the reference sources, pinned by the checks in tests/test_oracle_*.py
 * (closed forms, invariants, limits, brute force).  Each function cites the
 * reference file:line it follows.
 *
 * Conventions (all from the reference):
 *   primitive order  RHO,UU,VX,VY,VZ,ENTR,[B1,B2,B3],[EE,FX,FY,FZ]  mnemonics.h:13-44
 *   velocities       VELR (relative 4-velocity u~^i) for gas and radiation  choices.h
 *   metric per point gg[4][5]: g_{mu nu}, gg[i][4]=d ln sqrt(-g)/dx^i (i<3),
 *                    gg[3][4]=sqrt(-g);  GG[4][5]: g^{mu nu}, GG[3][4]=g_tt
 *                    perturbation (gttpert)   ko.h:336-353, metric.c:3866-3935
 *   GDETIN=1 (sqrt(-g) inside the conserved quantities)  choices.h
 */
#ifndef KORAL_ORACLE_H
#define KORAL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define O_RHO 0
#define O_UU 1
#define O_VX 2
#define O_VY 3
#define O_VZ 4
#define O_ENTR 5
#define O_NVMAX 13

/* radiation/MHD mnemonics (mnemonics.h:4-5) */
#define O_RAD 1
#define O_MHD 2

/* u2p types (mnemonics.h:65-69) */
#define O_U2P_HOT 0
#define O_U2P_ENTROPY 1

/* implicit solver equation/frame selectors (mnemonics.h:138-143) */
#define O_ENERGYEQ 0
#define O_ENTROPYEQ 1
#define O_LAB 0
#define O_FF 1

/* cell flags (mnemonics.h:105-113) */
#define O_NFLAGS 7
#define O_ENTROPYFLAG 0
#define O_RADIMPFIXUPFLAG 1
#define O_HDFIXUPFLAG 2
#define O_RADFIXUPFLAG 3
#define O_ENTROPYFLAG2 4
#define O_ENTROPYFLAG3 5
#define O_RADSOURCETYPEFLAG 6

/*
 * Resolved compile-time configuration of one KORAL build (choices.h defaults,
 * overridable per problem).  Field order mirrors koral_params in
 * include/koral_hip.h; tests check both layouts agree byte for byte.
 */
typedef struct {
  /* layout */
  int nv, mag, rad, pad0;
  /* equation of state and units (ko.h:4-48) */
  double gamma;
  double k_boltz, m_proton, sigma_rad, mu_gas;
  double ggg, ccc, masscm;
  /* opacities: model 0 -> kappa_abs = kappa_a_coef*rho (RADTUBE kappa.c);
     model 1 -> Kramers free-free in cgs (KATOTORUS kappa.c);
     kappa_es = kappa_es_coef*rho */
  int kappa_model;
  /* CORRECT_POLARAXIS: NCCORRECTPOLAR cells next to each polar face are
     overwritten by do_correct (finite.c:605-636); 0 = not defined */
  int polar_nc;
  double kappa_a_coef, kappa_es_coef;
  /* floors (choices.h:52-118,319-337) */
  double rhofloor, uurhoratiomin, uurhoratiomax, b2rhoratiomax;
  double gammamaxhd, gammamaxrad, eradfloor;
  double eerhoratiomin, eerhoratiomax, eeuuratiomin, eeuuratiomax;
  /* MHD inversion (choices.h:457-459) */
  double u2pconv;
  /* implicit source (choices.h:461-507,596-598) */
  double radimpconv, radimpentrconv, radimpeps;
  int radimpmaxiter, convrel_enabled;
  double dampingfactor, maxenchangedown, maxenchangeup, maxradimpdamping;
  double radimplicitthreshold;
  double radimpconvrel, radimpconvrelerr, radimpconvrelentr, radimpconvrelentrerr;
  /* explicit operator */
  int int_order, fluxmethod;
  double tsteplim, minmod_theta;
  int dofixups, dou2pmhdfixups, dou2pradfixups, doradimpfixups;
} oracle_params;

typedef struct {
  double gg[4][5];
  double GG[4][5];
  double gdet, alpha, gttpert;
} o_geom;

/* geometry from the 40 tabulated numbers of one point (fill_geometry,
   metric.c:1643-1676) */
void o_fill_geom(o_geom *g, const double *g20, const double *G20);

/* --- algebra (relele.c, frames.c, magn.c, p2u.c) --- */
void o_relr_to_4vel(const double *ur, double *ucon, double *ucov, const o_geom *g);
double o_utp1(const double *vr, const o_geom *g);
void o_bcon(const double *pp, const double *ucon, const double *ucov, double *bcon, int b1);
void o_lower(const double *a, double *b, const double gg[4][5]);
void o_raise(const double *a, double *b, const double GG[4][5]);

/* --- state conversions --- */
int o_p2u(const oracle_params *P, const double *pp, double *uu, const o_geom *g);
int o_p2u_mhd(const oracle_params *P, const double *pp, double *uu, const o_geom *g);
int o_p2u_rad(const oracle_params *P, const double *pp, double *uu, const o_geom *g);
double o_calc_Sfromu(const oracle_params *P, double rho, double u);
int o_u2p_solver_W(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int etype);
int o_u2p_solver_W_from(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int etype,
                        double Wstart, int *iters);
int o_u2p(const oracle_params *P, const double *uu0, double *pp, const o_geom *g,
          int corrected[3], int fixups[2]);
int o_u2p_rad(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int *corrected);
int o_check_floors_mhd(const oracle_params *P, double *pp, const o_geom *g);
int o_check_floors_rad(const oracle_params *P, double *pp, const o_geom *g);

/* --- radiation --- */
void o_calc_Rij(const oracle_params *P, const double *pp, const o_geom *g, double R[4][4]);
double o_kappa_abs(const oracle_params *P, const double *pp);
double o_kappa_es(const oracle_params *P, const double *pp);
void o_calc_Gi(const oracle_params *P, const double *pp, const o_geom *g, double Gi[4], int labframe);
void o_calc_ff_Rtt(const oracle_params *P, const double *pp, const o_geom *g, double *Rtt, double *ucon);
double o_TfromE(const oracle_params *P, double E);

typedef struct {
  int iters;       /* Newton iterations of the successful attempt (0 on failure) */
  int attempt;     /* 0..5 index of the successful attempt, -1 if all failed */
  int whichprim;   /* O_RAD or O_MHD of the successful attempt */
  int whicheq;
  int whichframe;
  int total_iters; /* iterations over all attempts (cost) */
} o_imp_info;

int o_f_implicit(const oracle_params *P, const double *ppin, const double *uu0, double *pp0,
                 double dt, const o_geom *g, double *f, const int *params, double *err0);
int o_solve_implicit_4dprim(const oracle_params *P, const double *uu00, const double *pp00,
                            const o_geom *g, double dt, const int *params, double *pp,
                            double global_time, int *iters_out);
int o_solve_implicit_lab(const oracle_params *P, double *uu, double *pp, const o_geom *g,
                         double dt, double global_time, int *flags, o_imp_info *info);
int o_solve_implicit_4dprim_from(const oracle_params *P, const double *uu00, const double *pp00,
                                 const o_geom *g, double dt, const int *params, double *pp,
                                 double global_time, int *iters_out, const double *xg, const double *xr,
                                 int maxiter);
int o_solve_implicit_lab_guess(const oracle_params *P, double *uu, double *pp, const o_geom *g,
                               double dt, double global_time, int *flags, o_imp_info *info,
                               const double *gxg, const double *gxr, int maxiter);

/* --- batched entry points used by tests and the CPU baseline --- */
int oracle_params_size(void);
int oracle_params_offset(int field);
void oracle_p2u_batch(const oracle_params *P, long n, const double *pp, double *uu,
                      const double *g20, const double *G20);
void oracle_u2p_batch(const oracle_params *P, long n, double *uu, double *pp,
                      const double *g20, const double *G20, int *ret, int *flags);
void oracle_u2p_solver_batch(const oracle_params *P, long n, const double *uu, double *pp,
                             const double *g20, const double *G20, int etype, int *ret);
void oracle_u2p_solver_batch_from(const oracle_params *P, long n, const double *uu, double *pp,
                                  const double *g20, const double *G20, int etype, const double *W0, int *ret,
                                  int *iters, int *guessed);
void oracle_u2p_rad_batch(const oracle_params *P, long n, const double *uu, double *pp,
                          const double *g20, const double *G20, int *ret, int *corr);
void oracle_floors_batch(const oracle_params *P, long n, double *pp,
                         const double *g20, const double *G20, int *ret);
void oracle_Gi_batch(const oracle_params *P, long n, const double *pp,
                     const double *g20, const double *G20, int labframe, double *Gi);
void oracle_ff_Rtt_batch(const oracle_params *P, long n, const double *pp,
                         const double *g20, const double *G20, double *Rtt);
void oracle_implicit_batch(const oracle_params *P, long n, double *uu, double *pp,
                           const double *g20, const double *G20, const double *dt,
                           double global_time, int *flags, int *info, int nthreads);
void oracle_implicit_batch_guess(const oracle_params *P, long n, double *uu, double *pp, const double *g20,
                                 const double *G20, const double *dt, double global_time, int *flags, int *info,
                                 int nthreads, const double *gxg, const double *gxr, int maxiter);
int oracle_inverse44(const double *a, double *ia);

#ifdef __cplusplus
}
#endif
#endif
