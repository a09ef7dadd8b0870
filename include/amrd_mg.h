/*
 * amrd_mg.h -- AMRD's multigrid driver (amrd/mg.c) over pamr.h, on one
 * MI355X (gfx950).
 *
 * The application supplies the MG hooks of amrd()'s argument list
 * (include/amrd.h:17-34,137-159): app_MG_residual, app_MG_relax, app_L_op.
 * As in AMRD they are called once per grid of a level inside a PAMR
 * sequential iterator (call_app / call_rr_app, amrd/util.c:39-103); a hook
 * fetches its grid with PAMR_get_g_attribs and its grid functions with
 * PAMR_get_gfn -- device pointers here, so a hook launches kernels (on
 * amrd_gpu_stream()) and returns its norm as a host double.  The norms of a
 * level are folded with max, as call_rr_app's MPI_MAX on one rank.
 *
 * Return convention: 1 success, 0 failure with a message, as AMRD / PAMR.
 */
#ifndef AMRD_MG_GPU_H
#define AMRD_MG_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double (*amrd_real_hook)(void);
typedef void (*amrd_void_hook)(void);
typedef void (*amrd_evolve_hook)(int iter);
typedef double (*amrd_level_relax_hook)(int L, int comm_only);
typedef double (*amrd_level_real_hook)(int L);
typedef void (*amrd_level_void_hook)(int L);

/* The MG hooks of amrd() (include/amrd.h:137-159): app_MG_residual
 * (RMS of res = L[f] - rhs where cmask is on), app_MG_relax (one sweep,
 * RMS of the residuals it saw), app_L_op (f_lop = L[f] where cmask is on). */
void amrd_set_app_MG_hooks(amrd_real_hook app_MG_residual, amrd_real_hook app_MG_relax,
                           amrd_void_hook app_L_op);

/* Optional level hooks, one call per MG level instead of one per grid; no
 * counterpart in the reference, for applications that batch all grids of a
 * level into one launch.  relax(L, comm_only) relaxes every grid of L (only
 * those with PAMR_get_g_comm != 0 when comm_only) and returns the max norm;
 * residual(L) / lop(L) likewise.  NULL entries keep the per-grid hooks. */
void amrd_set_app_MG_level_hooks(amrd_level_relax_hook relax, amrd_level_real_hook residual,
                                 amrd_level_void_hook lop);

/* AMRD's MG variable wiring (amrd/io.c:843-937,988-1005): for every
 * elliptic variable (names separated by spaces or commas, each already
 * defined with PAMR_def_var_full, in_mgh = 1) define <v>_res (mg_inject
 * half-weight, full-weight when <v>_res is in mg_fw_restr; mg_interp second
 * order when <v> is in mg_interp2), <v>_lop and <v>_rhs (MG only), then
 * cmask (and chr when do_ex).  Call before PAMR_compose_hierarchy. */
int amrd_mg_def_vars(const char *elliptic_vars, const char *mg_fw_restr, const char *mg_interp2,
                     int do_ex);

/* MG extrapolation in time for constrained evolution.
 *
 * amrd_mg_def_extrap_vars (amrd/io.c:898-934): for every elliptic variable v
 * of amrd_mg_def_vars, "<v>_brs", "<v>_extrap_tm1" and "<v>_extrap_tm2" in
 * the AMR hierarchy with one time level (the extrapolation variables
 * straight-injected, fourth-order interpolated and regridded, second order
 * when listed in amr_interp2 / amr_transfer2).  v itself needs two AMR time
 * levels.  Call after amrd_mg_def_vars, before PAMR_compose_hierarchy.
 *
 * amrd_mg_extrap (amrd/evolve.c:346-368): on AMR level lev > 1 (level 1 is
 * left alone), v(tl 1) = f1 tm1 + f2 tm2 with f1 = (tnp1 - (tnp1_lm1 -
 * 2 dt_lm1)) / dt_lm1, f2 = -(tnp1 - (tnp1_lm1 - dt_lm1)) / dt_lm1 (method
 * 0, AMRD_MG_EXTRAP_2ND_FROM_ETM1), or v(tl 1) = v(tl 2) + dt/dt_lm1 (tm2 -
 * tm1) (method 1, AMRD_MG_EXTRAP_2ND_FROM_TN).
 *
 * amrd_set_mg_brs_vars (amrd/evolve.c:1624-1639): brs = v(tl 2) on levels
 * L1..L2.
 *
 * amrd_update_mg_extrap_vars (amrd/evolve.c:1661-1702): on levels L1..Lf
 * (L1+1..Lf when L1 > 1), tm2 = tm1, tm1 = v(tl 2), and on levels > 1 with
 * eps_c != 0 (AMRD_MG_eps_c) tm2 += (tm1 - brs) (1 - eps_c / net_rho), net_rho
 * the time refinement product the reference accumulates.
 *
 * All return 1 / 0 and enqueue on amrd_gpu_stream(). */
int amrd_mg_def_extrap_vars(const char *amr_interp2, const char *amr_transfer2);
int amrd_mg_extrap(int lev, int method, double tnp1, double dt, double tnp1_lm1, double dt_lm1);
int amrd_set_mg_brs_vars(int L1, int L2);
int amrd_update_mg_extrap_vars(int L1, int Lf, double eps_c);

/* The t = 0 branches of solve_elliptics (amrd/mg.c:28-29,45-68).
 *
 * amrd_mg_def_vars_t0: AMRD's elliptic_vars_t0 -- MG variables of the
 * application solved alongside the elliptic variables while is_t0 is set
 * (every MG routine then runs over both lists, amrd/mg.c:248,279,327,612);
 * their <v>_res, <v>_lop, <v>_rhs as amrd_mg_def_vars defines them.
 * amrd_mg_def_cnst_data_vars: MG_cnst_data_vars, MG variables the
 * application fills in its app_t0_cnst_data hook.
 * amrd_set_is_t0: AMRD_is_t0 (default 1).
 * amrd_set_elliptic_vars_t0_init (amrd/amrd.c:109-112) and
 * amrd_set_app_t0_cnst_data (amrd()'s argument of that name): per-grid hooks
 * over the MG hierarchy.  With is_t0, solve_elliptics calls the t0 init hook
 * on every MG level from the coarsest up, syncing its variables on the level
 * and interpolating them (fourth order) to the next, then injects them
 * (straight) back down; after every V-cycle it does the same with the
 * constraint-data hook and variables.  Definitions return 1 / 0. */
int amrd_mg_def_vars_t0(const char *elliptic_vars_t0, const char *mg_fw_restr, const char *mg_interp2);
int amrd_mg_def_cnst_data_vars(const char *vars);
void amrd_set_is_t0(int is_t0);
void amrd_set_elliptic_vars_t0_init(amrd_void_hook app_elliptic_vars_t0_init);
void amrd_set_app_t0_cnst_data(amrd_void_hook app_t0_cnst_data);

/* tstep's iteration of the evolution and elliptic equations on AMR level
 * lev (amrd/evolve.c:449-500; the dissipation and i-1 boundary steps around
 * it are the application's): on the finest level with elliptic variables it
 * builds the MG hierarchy PAMR_build_mgh(Lf, Lf, 1) -- here only when the
 * finest level is level 1 --, then while (iter < evo_min_iter || evo_res >
 * evo_tol || mg_res > MG tol) && iter < evo_max_iter: app_evolve(iter+1) on
 * every grid of lev, PAMR_sync(lev, tl 1, AMR_bdy_width 0 first then 1),
 * evo_res = max over the grids of app_evo_residual(), and (unless evo_ssc
 * says it already converged) mg_res = one V-cycle.  Returns the iteration
 * count (negative on error) and the last residuals. */
int amrd_evo_iterate(int lev, amrd_evolve_hook app_evolve, amrd_real_hook app_evo_residual, int evo_min_iter,
                     int evo_max_iter, double evo_tol, int evo_ssc, double *evo_res, double *mg_res);

/* AMRD's MG parameters (amrd/io.c:97-106,269-285); defaults 3, 3, 1, 50,
 * 0.0, 1e-3, 5000, 1.0, 1.0 */
void amrd_mg_set_params(int pre_swp, int pst_swp, int min_iter, int max_iter, double tol,
                        double crtol, int max_citer, double w0_r, double w0_i);

/* AMRD's truncation-error variables (amrd/io.c:810-838): for every name in
 * tre_vars (AMRD_TRE_vars) "<name>_tre" is defined in the AMR hierarchy
 * with one time level; the first one also lives in the MG hierarchy and
 * receives the MG estimate.  Call before PAMR_compose_hierarchy.
 * Returns 1 / 0. */
int amrd_def_tre_vars(const char *tre_vars);

/* vcycle (amrd/mg.c:84-201) on the current MG hierarchy; returns the
 * residual norm of the finest level (negative on error).  amrd_vcycle is
 * vcycle(giter, 0). */
double amrd_vcycle(int giter);

/* vcycle(giter, save_tre) (amrd/mg.c:84): with save_tre set,
 * compute_MG_rhs (amrd/mg.c:271-314) also writes max_i |lop_i| *
 * sqrt(num_f_tre_vars) into the first "<name>_tre" variable on each coarse
 * level, syncs it and interpolates it (second order) to the next finer
 * level, so that every level above the coarsest ends with the estimate of
 * the level below it.  An error (negative return) without tre variables. */
double amrd_vcycle_tre(int giter, int save_tre);

/* solve_elliptics (amrd/mg.c:20-79): PAMR_build_mgh(L, finest AMR level,
 * tl), V-cycles while (iter < min_iter || res > tol) && iter < max_iter,
 * PAMR_destroy_mgh.  Returns the number of cycles (negative on error);
 * res_out (may be NULL) the last residual. */
int amrd_solve_elliptics(int L, int tl, double *res_out);
/* solve_elliptics(L, tl, save_tre) (amrd/mg.c:20): the same with every
 * V-cycle run as amrd_vcycle_tre(iter, save_tre) */
int amrd_solve_elliptics_tre(int L, int tl, int save_tre, double *res_out);

/* The reference's MG driver under its own names and signatures
 * (amrd/mg.h:32,37), for AMRD code that calls them unchanged
 * (amrd/evolve.c:285,483,980,1156,1192,1257,1294).
 * solve_elliptics(L, tl, save_tre) replaces amrd/mg.c:20-79: it is
 * amrd_solve_elliptics_tre, and like the reference (AMRD_stop,
 * amrd/util.c:27-34) it prints a message, frees the context and exits the
 * program with status 1 when the MG hierarchy cannot be built or a device
 * step fails.  vcycle(giter, save_tre) replaces amrd/mg.c:84-201: it is
 * amrd_vcycle_tre (the residual of the finest level; negative on error). */
void solve_elliptics(int L, int tl, int save_tre);
double vcycle(int giter, int save_tre);

/* AMRD's MG tracing parameters (amrd/io.c:80-83,256-259): MG_trace 1-4
 * prints, on stdout and in the reference's format, the solve's level range
 * and cycle count (1), the residual after every cycle (2), every level's
 * norm after its sweeps (3) and after every sweep (4) (amrd/mg.c:35-41,71,76,
 * 138,142,164,195).  MG_DV_trace != 0 makes every V-cycle dump its MG
 * grid functions with mg_dump at its four stages (amrd/mg.c:133,144,159,197)
 * while MG_DV_trace_t_on <= t <= MG_DV_trace_t_off, t the time of the finest
 * AMR level.  Defaults 0, -1, 1e-15, 1e-15. */
void amrd_mg_set_trace(int MG_trace, int MG_DV_trace, double MG_DV_trace_t_on, double MG_DV_trace_t_off);

/* The variable lists mg_dump writes with which = 0 after the elliptic
 * variables and their _lop/_rhs/_res (amrd/mg.c:757-763): hyperbolic_vars,
 * then "cmask", then MGH_work_vars, then AMRH_work_in_MGH_vars
 * (AMRD_hyperbolic_vars, AMRD_MGH_work_vars, AMRD_AMRH_work_in_MGH_vars,
 * read from the parameter file at amrd/io.c:318-352).  Space- or
 * comma-separated names; NULL or "" for none.  Returns 1. */
int amrd_mg_def_dump_vars(const char *hyperbolic_vars, const char *MGH_work_vars, const char *AMRH_work_in_MGH_vars);

/* mg_dump(L, tag, giter, which) (amrd/mg.c:707-766, the reference's name
 * and signature): PAMR_save_gfn of the MG grid functions of level L with
 * pre_tag `tag` at time lcount + giter / (max_iter + 2), lcount counting the
 * solves (giter returning to 0); which 2: the elliptic variables, 1: also
 * their _lop, _rhs, _res, 0: also the amrd_mg_def_dump_vars lists.  Does
 * nothing unless MG_DV_trace is set and the finest AMR level's time is in
 * its window. */
void mg_dump(int L, char *tag, int giter, int which);

/* last norms of amrd_vcycle: lev_norms[L-1] for MG levels 1..Lf (the
 * reference's res[] array) and the coarsest-level sweep count */
int amrd_mg_last_norms(double *lev_norms, int *citer);

/* the stream every device operation of pamr.h / amrd_mg.h is enqueued on
 * (a hipStream_t); NULL before PAMR_compose_hierarchy */
void *amrd_gpu_stream(void);

#ifdef __cplusplus
}
#endif
#endif
