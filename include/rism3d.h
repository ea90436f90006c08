/*
 * rism3d.h -- C ABI of the MI355X-native 3D-RISM solute-in-solvent solver.
 *
 * This is the drop-in boundary for the reference's 3D-RISM solute path:
 *
 *   void hnc3d_solute_solve (const ProblemData *PD,
 *                            const int m, const Site solvent[m],
 *                            const int n, const Site solute[n],
 *                            void (*density)(int, const real[][3], real[]),
 *                            SCM *dict, Vec g[m], const real *chi_fft_buf,
 *                            Context **medium, Restart **restart);
 *                                                  (hnc3d.h:13-21)
 *
 * PETSc Vecs, Guile SCM values and the global option fluid are replaced by
 * plain pointers, sizes and an explicit options struct.  No torch types.
 * All floating point is IEEE double on the boundary.
 */
#ifndef RISM3D_H
#define RISM3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Closure numbering of ClosureEnum (bgy3d.h:211-223, foreign.f90:42-57). */
enum rism3d_closure {
  RISM3D_CLOSURE_HNC = 0,
  RISM3D_CLOSURE_KH = 1,
  RISM3D_CLOSURE_PY = 2,
  RISM3D_CLOSURE_PSE0 = 3,
  RISM3D_CLOSURE_PSE1 = 4,
  RISM3D_CLOSURE_PSE2 = 5,
  RISM3D_CLOSURE_PSE3 = 6,
  RISM3D_CLOSURE_PSE4 = 7,
  RISM3D_CLOSURE_PSE5 = 8,
  RISM3D_CLOSURE_PSE6 = 9,
  RISM3D_CLOSURE_PSE7 = 10
};

/* Field-for-field the reference ProblemData (bgy3d.h:228-256).  L is the
   full box edge and h = L/N, as bgy3d-guile.c:284-312 derives them from
   the CLI half-width --L. */
typedef struct rism3d_problem {
  int N[3];          /* grid size; N[0] is the stride-1 (x) axis        */
  double L[3];       /* box edge [A]                                    */
  double h[3];       /* mesh width [A], L/N                             */
  double rmax;       /* radial extent of the 1D chi table [A]           */
  int nrad;          /* number of 1D radial points                      */
  double beta;       /* inverse temperature [mol/kcal]                  */
  double rho;        /* solvent site number density [A^-3]              */
  double lambda;     /* mixing parameter of the fixed-point solvers     */
  double damp;       /* unused on this path, kept for layout            */
  int max_iter;      /* iteration limit                                 */
  double norm_tol;   /* convergence threshold, max-norm of the residual */
  int closure;       /* enum rism3d_closure                             */
} rism3d_problem;

/* Field-for-field the reference Site (bgy3d-solutes.h:9-16); the Guile
   SCM handle becomes an opaque user pointer at the same offset. */
typedef struct rism3d_site {
  char name[5];
  double x[3];       /* position [A]                                    */
  double sigma;      /* LJ sigma [A]                                    */
  double epsilon;    /* LJ epsilon [kcal/mol]                           */
  double charge;     /* [e]                                             */
  void *user;
} rism3d_site;

enum rism3d_status {
  RISM3D_OK = 0,
  RISM3D_NONCONVERGED = 1,  /* max_iter reached (bgy3d-snes.c:339-345)  */
  RISM3D_NAN = 2,           /* residual not finite (diverged closure)   */
  RISM3D_BAD_ARG = 3,
  RISM3D_NO_DEVICE = 4,
  RISM3D_OOM = 5,
  RISM3D_UNSUPPORTED = 6,   /* grid size or site count outside limits   */
  RISM3D_NO_SCRATCH = 7     /* rism3d_residual_noalloc: the problem needs
                               more than the fixed device scratch         */
};

/* The reference's snes-solver choices (bgy3d_snes_default,
   bgy3d-snes.c:16-35) plus MDIIS.  Every solver stops at max|F| <=
   norm_tol.  NEWTON, JAGER and TRIAL run in FP64 only. */
enum rism3d_solver {
  RISM3D_SOLVER_MDIIS = 0,  /* default: MDIIS / Anderson mixing         */
  RISM3D_SOLVER_PICARD = 1, /* x += lambda F(x) (bgy3d-snes.c:348-384)  */
  RISM3D_SOLVER_NEWTON = 2, /* Newton, GMRES(30) on the matrix-free
                               Jacobian, backtracking line search
                               (bgy3d_snes_newton, bgy3d-snes.c:161-346) */
  RISM3D_SOLVER_JAGER = 3,  /* Picard with the adaptive step schedule of
                               bgy3d_snes_jager (bgy3d-snes.c:402-519)   */
  RISM3D_SOLVER_TRIAL = 4   /* 100 Picard steps at lambda = 0.02, then
                               Newton (bgy3d_snes_trial, :387-399)       */
};

enum rism3d_precision {
  RISM3D_FP64 = 0,          /* every kernel in FP64                     */
  RISM3D_MIXED = 1,         /* FP32 passes on a correction around an FP64
                               base (defect correction): FP64 residual at
                               every rebase, then polish_iters FP64 steps;
                               no FP32 floor.  The FP32 MDIIS history is
                               kept as the newest u in FP32, BF16
                               differences of u and FP8 residuals;
                               solutes ready for a rebase wait for one
                               another (grouped FP64 passes)             */
  RISM3D_FP32 = 2           /* FP32 passes only, around the starting
                               point: they stop at the FP32 floor (no
                               rebase, no polish); the FP64 residual there
                               is reported (precision studies)          */
};

/* The reference reads these from the Guile *settings* fluid
   (hnc3d.c:1575, 1600, 1783-1784; bgy3d-snes.c:20).  Here they are
   explicit.  Zero-initialise and call rism3d_options_default(). */
typedef struct rism3d_options {
  int renorm;          /* 1: tau renormalisation (default), 0: --no-renorm */
  int solver;          /* enum rism3d_solver                               */
  int precision;       /* enum rism3d_precision                            */
  int polish_iters;    /* MIXED: FP64 steps after the FP32 phase, towards
                          norm_tol; NONCONVERGED if the last one is above
                          mixed_tol.  0: OK when the FP32 phase reached
                          mixed_tol or its floor (residual then reports
                          the FP64 residual of the FP32 iterate)          */
  double mixed_tol;    /* MIXED: max-norm residual at which FP32 hands over
                          (or its FP32 floor, detected as a stall below
                          1e-3)                                            */
  int diis_depth;      /* MDIIS history length (1..16)                     */
  double diis_step;    /* MDIIS step lambda applied to residuals           */
  int verbose;         /* print per-iteration norms to stderr              */
  int device;          /* HIP device ordinal                               */
  double dielectric;   /* chi_k == NULL only: DRISM target epsilon of the
                          internal 1D chi producer ("dielectric" setting,
                          rism.f90:346); 0 selects plain RISM              */
  int continuation;    /* HNC, PSE-n >= 2: a solute whose iteration
                          diverges (NAN) restarts from t = 0 along KH,
                          PSE-2, ... to the target closure, each rung
                          loosely converged (default 1; 0: report NAN).
                          The reference has no such fallback; the fixed
                          point is unchanged                               */
  const double *chi3d; /* NULL, or the --solvent-3d susceptibility: chi - 1
                          of a 3D solvent solve on this grid
                          (rism3d_solvent_solve's chi_out, real part),
                          double[m(m+1)/2][N2][N1][N0/2+1] by pair
                          i(i+1)/2 + j, j <= i (solvent_kernel_file,
                          hnc3d.c:1434-1437, 1572-1590).  Replaces chi_k;
                          no tau renormalisation, as in the reference      */
  const double *site_rho; /* NULL (every solvent site at pd->rho, as the
                          reference, hnc3d.c:90-92), or double[m]: the
                          number density of each solvent site of a
                          mixture (no counterpart in the reference).  chi_k
                          is then rism3d_solvent_chi_mixture's
                          chi_ij = w_ij + rho_i h_ij, contracted as
                          h_i = sum_j c_j chi_ji, tau_i = sum_j q_j chi_ji,
                          and mu / the PMV factor weight site i by rho_i.
                          Requires chi_k; no gradients (UNSUPPORTED)       */
  double coarse_tol;   /* 0: start from t = 0 (hnc3d.c:1845).  > 0 (no
                          counterpart in the reference): first solve the
                          same problem on the grid N/2 in the same box to
                          max|F| <= coarse_tol, then start from that t,
                          trigonometrically interpolated to N.  Same root,
                          fewer residuals on N; iterations counts those on
                          N only.  Applies when every N[d] / 2 is a
                          supported size and neither chi3d nor a restart
                          token is given; otherwise ignored.               */
  int tune;            /* 0: the built-in launch configuration.  1 (no
                          counterpart in the reference): before iterating,
                          time the candidate configurations of the residual
                          chain on this batch (FP32 chain on xy planes or in
                          five launches, z-pass block order, z-pass tile of
                          16 or 32 kx columns, the Nyquist column in a
                          launch of its own or not) and use the fastest;
                          choices are kept per process by grid, sites,
                          batch size and precision, and appended to the
                          file RISM3D_TUNE_FILE names, if any.  Only the
                          launch configuration changes (in FP32 the planes
                          change the rounding order, not the fixed point)  */
} rism3d_options;

typedef struct rism3d_result {
  double mu;           /* excess chemical potential, native closure [kcal/mol]
                          ('free-energy, hnc3d.c:2232-2237)                */
  double mu_hnc, mu_kh, mu_gf;  /* print_chempot functionals (hnc3d.c:669) */
  double pmv_factor;   /* 1 - rho int c d3r ('partial-molar-volume-factor) */
  double residual;     /* final max-norm of F(t)                            */
  int iterations;      /* residual evaluations, all precisions              */
  int iterations_fp32; /* of which in FP32                                  */
  int status;          /* enum rism3d_status                                */
  double *g;           /* in: NULL or caller buffer [m][N2][N1][N0]; out:
                          g_i = h_i + 1 on the reference grid, r = i h - L/2 */
  double *t;           /* in: NULL or caller buffer [m][N2][N1][N0]; out:
                          converged indirect correlation t_i                */
  double *gradient;    /* in: NULL or [n][3]; out: the reference's TPT
                          gradient rho h^3 sum_i <g_i | dv_i/dx_a>
                          ('free-energy-gradient, --derivatives;
                          gradients()/energy1() hnc3d.c:1666-1742)          */
  double *dmu_dsite;   /* in: NULL or [n][6]; out: exact d mu / d(x, y, z,
                          q, sigma, epsilon) of every solute site by the
                          adjoint of the residual, on the tau and the Ng
                          path; its linear solve runs in FP64, or by FP32
                          defect correction after a MIXED solve (no
                          counterpart in the reference)                     */
  int adjoint_iterations;  /* linear MDIIS steps of the adjoint solve      */
  double *response;    /* in: NULL or [n][3]; out: the reference's linear-
                          response gradient ('free-energy-response,
                          --response; hnc3d.c:1366-1421, 1969-2134): the
                          first order change of mu along each site's x, y,
                          z through v_short at fixed u_c and tau, with the
                          full capped Coulomb derivative in dv.  Equal to
                          the x, y, z of dmu_dsite when the solute is
                          uncharged.  Either path; NaN for PY, whose
                          differential the reference lacks
                          (closures.f90:131-153)                          */
  struct rism3d_info *info;  /* in: NULL or caller struct; out: the
                          observables print_info prints (below)           */
  double *g_device;    /* in: NULL or a device buffer [m][N2][N1][N0] on
                          options.device; out: g written there on the
                          device (no host transfer; g above, if also
                          given, receives a copy)                         */
} rism3d_result;

/* The post-solve observables of the solute in its solvent: info() and
   print_info() (bgy3d-potential.c:341-491) with cage = false, as
   hnc3d_solute_solve calls it (hnc3d.c:1774, 2325).  Fields:
   rho_v = rho sum_i q_i g_i, the solvent charge density; U_v its potential
   in kcal from the FFT Poisson solver (bgy3d_poisson,
   bgy3d-poisson.c:103-190, zero mean); rho_u the Gaussian solute cores
   (G = 1.2, bgy3d-solutes.c:555-577); U_u = eps0^-1 sum_a q_a erf(G r)/r
   the long-range solute field; rho_N the point nuclei.  Moments are
   bgy3d_moments (bgy3d-vec.c:573-593): about the grid centre, h^3 sum
   v (i - N/2) h ..., Q[a][b] the second moments. */
typedef struct rism3d_info {
  double q_u, q_v;            /* charges of rho_u and rho_v                */
  double d_n[3];              /* sum_a q_a x_a (dipole of the nuclei)      */
  double d_u[3], d_v[3];      /* first moments of rho_u, rho_v             */
  double Q_u[3][3], Q_v[3][3];/* second moments                            */
  double U_v_rho_N;           /* <U_v|rho_N> = sum_a q_a U_v(x_a), U_v at
                                 x_a by trigonometric interpolation
                                 (bgy3d_fft_interp, bgy3d-fftw3.c:504-612) */
  double U_v_rho_u;           /* <U_v|rho_u> = h^3 sum U_v rho_u           */
  double rho_v_U_u;           /* <rho_v|U_u> = h^3 sum rho_v U_u           */
  double *v_sites;            /* in: NULL or [n]; out: U_v(x_a)            */
  double *potential;          /* in: NULL or [N2][N1][N0]; out: U_v on the
                                 grid, the field of the Context the
                                 reference returns as *medium              */
} rism3d_info;

/* Opaque restart token: the converged t (hnc3d.c:1826-1845, 2299-2316). */
typedef struct rism3d_restart rism3d_restart;

void rism3d_options_default(rism3d_options *opt);

/* Replaces hnc3d_solute_solve (hnc3d.c:1762-2350).
   chi_k: double[m][m][nrad], solvent susceptibility chi(k) (not offset by
   one) on k_j = (j + 1/2) pi / rmax, as produced by rism_solvent
   (rism.f90:57-102) and checked at bgy3d-guile.c:1136-1147; or NULL to
   compute it with the 1D DRISM producer on the "upscaled" grid
   rmax = 2 max(L), nrad = 16 max(N) (hnc3d.c:772-783, 1471-1495).
   restart: NULL, or in/out token (consumed on entry, replaced on exit).
   The density callback of the reference is not supported (it aborts on
   the HNC3D path in the reference too, bgy3d-solutes.c:964-968). */
int rism3d_solute_solve(const rism3d_problem *pd,
                        int m, const rism3d_site *solvent,
                        int n, const rism3d_site *solute,
                        const double *chi_k,
                        const rism3d_options *opt,
                        rism3d_result *res,
                        rism3d_restart **restart);

/* Many solutes sharing one solvent, chi and grid, solved together: every
   kernel launch covers all unconverged solutes (one launch per pass).
   solutes: n_solutes arrays of sites, n_sites[b] sites each.
   res[b] receives each result. */
int rism3d_batch_solve(const rism3d_problem *pd,
                       int m, const rism3d_site *solvent,
                       int n_solutes, const int *n_sites,
                       const rism3d_site *const *solutes,
                       const double *chi_k,
                       const rism3d_options *opt,
                       rism3d_result *res);

/* Solutes on uneven grids in one call (no counterpart in the reference):
   pds[b] is solute b's problem (grid, box, closure, tolerances; the same
   radial grid rmax / nrad for all when chi_k is given).  Solutes with
   identical problems form a size class solved as one batch
   (rism3d_batch_solve); the classes run at the same time, each on its own
   stream (one after another while kernel timing is on).  res[b] receives
   each result; returns OK, or the first failure of a class. */
int rism3d_batch_solve_grids(const rism3d_problem *pds,
                             int m, const rism3d_site *solvent,
                             int n_solutes, const int *n_sites,
                             const rism3d_site *const *solutes,
                             const double *chi_k,
                             const rism3d_options *opt,
                             rism3d_result *res);

/* One residual evaluation F(t) = (chi - 1) * c(t) - t, replacing
   iterate_t1 (hnc3d.c:1184-1285) as called by the reference's SNES.
   t, r: double[m][N2][N1][N0].  Uses the FP32 kernels when
   opt->precision == RISM3D_MIXED (t and r stay FP64).  chi_k required. */
int rism3d_residual(const rism3d_problem *pd,
                    int m, const rism3d_site *solvent,
                    int n, const rism3d_site *solute,
                    const double *chi_k,
                    const rism3d_options *opt,
                    const double *t, double *r);

/* Capacity of the fixed device scratch of rism3d_residual_noalloc, per
   device. */
#define RISM3D_NOALLOC_SCRATCH_MB 96

/* rism3d_residual without device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (serving
   loops and harnesses that must not call hipMalloc / hipFree).  Same
   arguments and result as rism3d_residual (t and r are host buffers,
   copied through pageable transfers); options.coarse_tol is ignored.
   Every device buffer of the call is taken from one fixed array of
   RISM3D_NOALLOC_SCRATCH_MB MiB that each device holds from the library's
   load until the process ends: the call makes no hipMalloc, hipFree or
   pinned-host allocation, and replays no graph.  Its first call on a
   device looks up that device's copy (hipGetSymbolAddress) and creates the
   stream it keeps.  A problem that needs more scratch returns
   RISM3D_NO_SCRATCH before any device work: 64^3 with 3 solvent sites
   needs about 61 MB in FP64 and 82 MB mixed; 96^3 does not fit.  The
   scratch keeps the last call's contents; each call writes what it reads.
   Calls from several threads of one process share a device's scratch and
   would race on it: the library's lock serialises them. */
int rism3d_residual_noalloc(const rism3d_problem *pd,
                            int m, const rism3d_site *solvent,
                            int n, const rism3d_site *solute,
                            const double *chi_k,
                            const rism3d_options *opt,
                            const double *t, double *r);

/* One Jacobian application at t, replacing jacobian_t1
   (hnc3d.c:1290-1362): jdt = (chi - 1) * (dc/dt dt) - dt, or with
   flag != 0 its flag form (chi - 1) * (dc/dt dt + dt) that the response
   path uses (hnc3d.c:1310-1320, 1387-1391).  FP64; chi_k required;
   RISM3D_UNSUPPORTED for PY, whose differential the reference lacks
   (closures.f90:131-153).  t, dt, jdt: double[m][N2][N1][N0]. */
int rism3d_jacobian(const rism3d_problem *pd,
                    int m, const rism3d_site *solvent,
                    int n, const rism3d_site *solute,
                    const double *chi_k,
                    const rism3d_options *opt,
                    const double *t, const double *dt, int flag,
                    double *jdt);

/* The 3D solvent-solvent solve, replacing hnc3d_solvent_solve
   (hnc3d.c:804-1073): site-site pair correlations on the 3D grid, centred
   at the box centre, from the OZ equation with the intramolecular
   omega(k); m <= 8 solvent sites (RISM3D_BAD_ARG beyond) whose
   m(m+1)/2 pair spectra of two z lines fit 64 KB (RISM3D_UNSUPPORTED
   otherwise: 6 sites up to N2 = 96), cubic box, FP64, MDIIS or Picard.
   Pairs are indexed p = i(i+1)/2 + j, j <= i.
   t_init: NULL (start from t = 0) or double[m(m+1)/2][N2][N1][N0];
   g_out:  NULL or double[m(m+1)/2][N2][N1][N0], g_ij = 1 + h_ij (the
           reference's g%d%d.bin);
   chi_out: NULL or double[m(m+1)/2][N2][N1][N0/2+1][2], complex
           chi - 1 at the grid corner (the reference's x%d%d-fft.bin; its
           real part is what rism3d_options.chi3d takes);
   t_out:  NULL or double[m(m+1)/2][N2][N1][N0], the converged t_ij;
   res:    mu (native closure), mu_hnc/_kh/_gf (print_chempot of the m x m
           pairs, hnc3d.c:669-687), residual, iterations, status. */
int rism3d_solvent_solve(const rism3d_problem *pd,
                         int m, const rism3d_site *solvent,
                         const rism3d_options *opt,
                         const double *t_init,
                         double *g_out, double *chi_out, double *t_out,
                         rism3d_result *res);

/* Replaces bgy3d_restart_destroy (bgy3d-impure.c:593-600). */
void rism3d_restart_destroy(rism3d_restart *restart);

/* 1D XRISM susceptibility of a solvent mixture (no counterpart in the
   reference, whose solvent is one molecule at one density): molecule[i] the
   molecule of site i (intramolecular w_ij only within a molecule),
   rho_sites[i] its number density [A^-3] (equal within a molecule).  Per k,
   H = (1 - W C D)^-1 W C W with D = diag(rho_sites), closure and
   long-range handling as rism3d_solvent_chi, no dielectric correction;
   pd->rho is unused.  chi_out double[m][m][nrad],
   chi_ij = w_ij + rho_i h_ij (not symmetric when densities differ): the
   chi_k of rism3d_solute_solve with options.site_rho = rho_sites. */
int rism3d_solvent_chi_mixture(const rism3d_problem *pd,
                               int m, const rism3d_site *solvent,
                               const int *molecule, const double *rho_sites,
                               double *chi_out);

/* 1D DRISM solvent susceptibility, replaces rism_solvent
   (rism.f90:57-102) with t_buf == NULL: chi_out is double[m][m][nrad].
   dielectric == 0 selects plain XRISM (rism.f90:346-350). */
int rism3d_solvent_chi(const rism3d_problem *pd,
                       int m, const rism3d_site *solvent,
                       double dielectric,
                       double *chi_out);

/* Kernel timing (no counterpart in the reference, which has no in-code
   profiling; for benchmarks).  After rism3d_profile_start, every residual
   evaluation brackets its kernels with HIP events on the solver's stream
   (launched one by one, not as a graph):  0 closure_r2c_x, 1 fft_y
   forward, 2 zstar (z FFTs + chi contraction), 3 fft_y inverse,
   4 c2r_x_resid, and in the three-launch FP32 chain on xy planes
   5 closure_r2c_xy (closure + x and y transforms), 2 zstar,
   6 c2r_xy_resid.  rism3d_profile_read returns, per kernel and working
   precision, the summed event time [ms], the summed algorithmic HBM bytes
   (each array the kernel must read or write, counted once) and the launch
   count, then stops profiling.  Arrays of 16: entries 0-7 the FP64
   launches of kernels 0-7, entries 8-15 their FP32 launches (7 unused). */
int rism3d_profile_start(int device);
int rism3d_profile_read(double *ms, double *bytes, long *launches);

/* Device memory is cached across calls: a call's buffers are kept when it
   returns and reused by the next call of the same shape; at the end of
   each call the buffers it did not use are freed, so the cache holds at
   most the last call's working set.  With the environment variable
   RISM3D_CACHE_MB set, every released buffer is kept instead, up to that
   many MB (0 disables the cache).  This frees every cached buffer on every
   device.  No counterpart in the reference, which allocates and frees per
   solve. */
int rism3d_cache_release(void);

/* Library build/version string, for diagnostics. */
const char *rism3d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RISM3D_H */
