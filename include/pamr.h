/*
 * pamr.h -- the PAMR calls of the multigrid path, over device-resident grid
 * functions on one MI355X (gfx950).
 *
 * Same names, arguments and return convention as the reference's
 * include/pamr.h (1 = success, 0 = failure with a message on stdout; `real`
 * is double), so an application written against PAMR links unchanged.  The
 * differences a caller sees:
 *
 *  - grid-function pointers handed out by PAMR_get_g_gfs / PAMR_get_g_attribs
 *    are DEVICE pointers (FP64, Fortran order, x fastest); coordinate arrays
 *    (PAMR_get_g_x) are host arrays;
 *  - device work of these calls is enqueued on amrd_gpu_stream() (amrd_mg.h),
 *    and application hooks launch their kernels on that stream.
 *
 * Supported scope: one context, dim 3, one rank (ghost width 0), refinement
 * ratio 2 on every level, AMR level 1 a single grid over the whole domain,
 * MG hierarchies built from AMR level 1 upwards (min_lev 1), restriction
 * ops 1-3, second- and fourth-order interpolation, AMR_bdy_width 0 or 1 in
 * PAMR_sync, excision (PAMR_excision_on).
 * Anything else fails with a message (return 0).  Not reentrant, as the
 * reference (one current context, src/gh.c:23-24).
 */
#ifndef PAMR_GPU_COMPAT_H
#define PAMR_GPU_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

/* hierarchies (include/pamr.h:118-119 of the reference) */
#define PAMR_AMRH 1
#define PAMR_MGH 2

/* injection / interpolation / sync operators (include/pamr.h:93-110) */
#define PAMR_NO_INJECT 0
#define PAMR_STRAIGHT_INJECT 1
#define PAMR_HW_RESTR 2
#define PAMR_FW_RESTR 3
#define PAMR_NO_INTERP 0
#define PAMR_SECOND_ORDER 2
#define PAMR_FOURTH_ORDER 4
#define PAMR_NO_TRANSFER 0
#define PAMR_NO_SYNC 0
#define PAMR_SYNC 1

/* physical boundary types (include/pamr.h:113-115) */
#define PAMR_UNKNOWN 0
#define PAMR_EVEN 1
#define PAMR_ODD 2

/* transfer-bit stages (include/pamr.h:131-139) */
#define PAMR_TF_SYNC 0
#define PAMR_TF_COMPOSE 1
#define PAMR_TF_INJECT 2
#define PAMR_TF_INJECT_TO_MG_LEV 3
#define PAMR_TF_INTERP 4
#define PAMR_TF_BDY_INTERP 5
#define PAMR_TF_MGH_INIT 6

/* ---- context ----------------------------------------------------------- */

/* include/pamr.h:16 (src/pamr.c:46-196): q must be NULL / 0 and cp_file
 * NULL; returns the context number (1) */
int PAMR_init_context(double *q, int q_size, int dim, int *shape, double *bbox, char *cp_file,
                      int cp_rank);
/* include/pamr.h:20 */
void PAMR_free_context(int num);
/* include/pamr.h:17 (src/pamr.c:257-347): boxes of levels min_lev..max_lev
 * (2*dim doubles each, lev[i] the level of box i); composing once per
 * context; device memory for every variable is allocated here */
int PAMR_compose_hierarchy(int min_lev, int max_lev, int num, int *lev, double *bbox, double t);
/* include/pamr.h:18 (src/pamr.c:354-378, src/gh.c:1222-1365): returns the
 * finest MG level */
int PAMR_build_mgh(int min_lev, int max_lev, int tl);
/* include/pamr.h:19 (src/gh.c:1371-1416) */
int PAMR_destroy_mgh(void);

/* include/pamr.h:24-41 */
void PAMR_set_lambda(double lambda);
void PAMR_get_lambda(double *lambda);
void PAMR_set_rho(int *rho_sp, int *rho_tm, int num);
void PAMR_get_rho(int *rho_sp, int *rho_tm, int num);
void PAMR_get_dxdt(int lev, double *dx, double *dt);
void PAMR_set_ghost_width(int *ghost_width);
void PAMR_get_ghost_width(int *ghost_width);
void PAMR_set_MG_coarse_width(int *min_width);
void PAMR_get_MG_coarse_width(int *min_width);
void PAMR_get_global_bbox(double *bbox);

/* ---- variables (include/pamr.h:74-86; src/pamr.c:1120-1300) ------------- */
int PAMR_def_var_brief(char *name);
int PAMR_def_var_full(char *name, int in_amrh, int in_mgh, int num_tl, int amr_inject,
                      int amr_interp, int amr_bdy_interp, int amr_sync, int mg_inject,
                      int mg_interp, int mg_sync, int mg_noinj_to_amr, int regrid_transfer,
                      int *phys_bdy_type);
int PAMR_get_var_attribs(char *name, int *in_amrh, int *in_mgh, int *num_tl, int *amr_inject,
                         int *amr_interp, int *amr_bdy_interp, int *amr_sync, int *mg_inject,
                         int *mg_interp, int *mg_sync, int *mg_noinj_to_amr,
                         int *regrid_transfer, int *phys_bdy_type);
/* include/pamr.h:188 (src/pamr.c:1735-1760) */
int PAMR_get_gfn(char *name, int hier, int tl);

/* ---- transfer bits (include/pamr.h:124-128; src/pamr.c:1767-1903) ------- */
void PAMR_freeze_tf_bits(void);
void PAMR_thaw_tf_bits(void);
void PAMR_clear_tf_bits(void);
void PAMR_set_tf_bit(int gf, int val);
void PAMR_set_tf_bits(int tl, int hierarchy, int stage);

/* ---- transfers (include/pamr.h:144-146; src/pamr.c:601-685) ------------- */
int PAMR_sync(int l, int tl, int hierarchy, int AMR_bdy_width);
int PAMR_inject(int lf, int tl, int hierarchy);
int PAMR_interp(int lc, int tl, int hierarchy);

/* ---- excision (include/pamr.h:160-163; src/pamr.c:1983-2037) ------------
 * ex_mask_var: a variable in both hierarchies with one time level; the
 * application's fill function writes `excised` into a HOST array of the
 * grid's shape at excised points (it is called per grid, by
 * PAMR_compose_hierarchy, by PAMR_build_mgh for the new MG levels, and now
 * when initialize_now; the library uploads the result).  While excision is
 * on, PAMR_inject and PAMR_interp transfer as local_transfer does with
 * do_ex (src/transfer.c:946-1163). */
int PAMR_excision_on(char *ex_mask_var,
                     void (*app_fill_ex_mask_fnc)(double *mask, int dim, int *shape, double *bbox, double excised),
                     double excised, int initialize_now);
void PAMR_excision_off(void);

/* ---- time (include/pamr.h:153-154) -------------------------------------- */
void PAMR_set_time(int lev, double t);
double PAMR_get_time(int lev);

/* ---- sequential grid iterator and grid getters (include/pamr.h:172-195;
 *      src/pamr.c:1372-1733) ---------------------------------------------- */
int PAMR_init_s_iter(int l, int hier, int all);
int PAMR_next_g(void);
int PAMR_push_iter(void);
int PAMR_pop_iter(void);
int PAMR_get_g_attribs(int *rank, int *dim, int *shape, double *bbox, int *ghost_width, double *t,
                       int *ngfs, double **x, double **gfs);
int PAMR_get_g_rank(int *rank);
int PAMR_get_g_dim(int *dim);
int PAMR_get_g_shape(int *shape);
int PAMR_get_g_bbox(double *bbox);
int PAMR_get_g_ghost_width(int *ghost_width);
int PAMR_get_g_t(double *t);
int PAMR_get_g_ngfs(int *ngfs);
int PAMR_get_g_x(double **x);
int PAMR_get_g_gfs(double **gfs);
int PAMR_get_g_coarsest(int *coarsest);
int PAMR_get_g_level(int *L);
int PAMR_get_max_lev(int hier);
int PAMR_get_min_lev(int hier);
int PAMR_get_g_comm(int *comm);
int PAMR_set_g_comm(int comm);

/* ---- grid-function output (include/pamr.h:202-203; src/pamr.c:1908-1978)
 * The grid function `name` of time level tl (AMR hierarchy) or the MG
 * hierarchy, every grid of level L (L < 0: every level), copied from the
 * device after the stream's earlier work and written as one record per grid
 * to "<pre_tag><name>_tl<tl><post_tag>_0" (AMR) or
 * "<pre_tag><name>_MG<post_tag>_0" (MG), with time t (-1.0: the level's
 * time).  The _close form closes that output afterwards, so the next save
 * to the name starts it again.  Records go to the writer registered with
 * PAMR_set_gft_writer; without one to the built-in record file
 * "<output name>.pgf" (below).  Returns 1 / 0. */
int PAMR_save_gfn(char *name, int hier, int tl, int L, double t, char *pre_tag, char *post_tag);
int PAMR_save_gfn_close(char *name, int hier, int tl, int L, double t, char *pre_tag, char *post_tag);

/* No counterpart in the reference, which calls bbhutil's SDF writer
 * gft_out_bbox / gft_close directly (src/pamr.c:1946,1951).  An application
 * that links bbhutil passes those two functions here and gets the
 * reference's SDF files; close may be NULL.  out_bbox(NULL, NULL) restores
 * the built-in writer: records appended to "<output name>.pgf" (the file is
 * truncated by its first record after the process starts or after a close),
 * each one, in the host's byte order,
 *   char magic[4] = "PGF1"; int32 rank; float64 time; int32 shape[rank];
 *   float64 bbox[2 * rank]; float64 data[shape[0] * ... * shape[rank-1]]
 * with data in Fortran order (x fastest), as gft_out_bbox takes it. */
typedef int (*PAMR_gft_out_bbox_fn)(char *name, double time, int *shape, int rank, double *bbox, double *data);
typedef int (*PAMR_gft_close_fn)(char *name);
void PAMR_set_gft_writer(PAMR_gft_out_bbox_fn out_bbox, PAMR_gft_close_fn close);

#ifdef __cplusplus
}
#endif
#endif
