/*
 * pamr_amr_gpu.h -- C ABI of the batched multigrid solve over an AMR
 * hierarchy (several grids per level) on one MI355X (gfx950).
 *
 * This is the device-resident counterpart of PAMR's MG hierarchy with
 * AMR levels: PAMR_compose_hierarchy + PAMR_build_mgh (src/pamr.c:257-378,
 * src/gh.c:1222-1365) build it, PAMR_sync / PAMR_inject / PAMR_interp
 * (src/pamr.c:601-685) move data with the reference's ownership and
 * transfer-list rules (src/transfer.c:73-575, 946-1163), and AMRD's
 * vcycle (amrd/mg.c:84-201) runs over it.  Every grid function is a device
 * FP64 array in Fortran order (x fastest), one per grid, as the reference's
 * gfs[] (src/gh.c:884-900).
 *
 * Execution: the transfer lists are resolved once per hierarchy into
 * disjoint index boxes (later transfers of a list win where they overlap,
 * as sequential execution gives), and each level operation -- a sweep over
 * every grid of the level, a sync, an inject, an interp -- is one kernel
 * launch.  A whole V-cycle is captured in one hipGraph.
 *
 * Scope: single MPI rank (ghost_width 0), refinement ratio 2, one base grid;
 * FP64, with FP32 inner cycles in pmg_amr_solve_rel mode 1; excision in the
 * transfers of pmg_amr_*_fields_ex only.  Return codes: 0 ok, < 0 error (as pamr_mg_gpu.h).
 */
#ifndef PAMR_AMR_GPU_H
#define PAMR_AMR_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmg_amr pmg_amr;

/* PAMR_compose_hierarchy + PAMR_build_mgh (src/pamr.c:257-378).
 * base_shape: points of the base grid over bbox (AMR level 1); namr refined
 * AMR levels; nbox[l] boxes of AMR level l+2; boxes: 6 doubles per box
 * (x0,x1,y0,y1,z0,z1), all levels concatenated in the caller's order (grids
 * are then ordered by insert_grid, src/gh.c:235-257).  op: 0 Poisson
 * (w/examples/nps/num_w.f), 1 div(a grad u) + c u. */
pmg_amr *pmg_amr_create(const int *base_shape, const double *bbox, int namr, const int *nbox,
                        const double *boxes, int min_mg_cwidth, int op);
void pmg_amr_destroy(pmg_amr *A);
/* PAMR_build_mgh(amr_lev, finest, tl) for amr_lev >= 1 (src/gh.c:1222-1365;
 * AMRD's tstep builds PAMR_build_mgh(Lf, Lf, 1), amrd/evolve.c:454): an MG
 * hierarchy whose finest levels are A's AMR levels amr_lev .. finest -- the
 * same grids, sharing A's grid functions and stream -- and whose coarser
 * levels are coarsened from the boxes given for AMR level amr_lev (in the
 * order given to pmg_amr_create): for cf = 2, 4, ... a level of the boxes
 * whose calc_island_no (src/gh.c:1427-1467) is nonzero, grids flagged
 * coarsest where it is negative, until no box survives; the grids of level
 * 1 are all coarsest.  New levels have every field A has (zero).  The
 * transfers and pmg_amr_field work on it; the built-in cycle
 * (pmg_amr_vcycle / _solve / _residual) does not (-1).  Destroy it before A;
 * NULL on failure. */
pmg_amr *pmg_amr_create_mgh(pmg_amr *A, int amr_lev);
/* 1 when grid g of level L is flagged coarsest, 0 when not, -1 on error */
int pmg_amr_grid_coarsest(pmg_amr *A, int L, int g);
/* PAMR_get_max_lev(PAMR_MGH) */
int pmg_amr_nlev(pmg_amr *A);
/* MG level that holds AMR level 1 */
int pmg_amr_base_level(pmg_amr *A);
int pmg_amr_ngrids(pmg_amr *A, int L);
/* PAMR_get_g_attribs subset: shape and bbox of grid g (0-based) of level L */
int pmg_amr_grid_info(pmg_amr *A, int L, int g, int *shape, double *bbox);
/* device pointer of a grid function (which: PMG_F_U .. PMG_F_C of
 * pamr_mg_gpu.h).  After writing u through it, call pmg_amr_commit_u. */
double *pmg_amr_field(pmg_amr *A, int L, int g, int which);
int pmg_amr_commit_u(pmg_amr *A);
int pmg_amr_upload(pmg_amr *A, int L, int g, int which, const double *host);
int pmg_amr_download(pmg_amr *A, int L, int g, int which, double *host);
/* V(pre,pst) cycle parameters; rop_*: restriction ops (1 straight, 2
 * half-weight, 3 full-weight) of u, u_res and the sources; interp_ord: the
 * coarse-grid correction's interpolation order (u_res's mg_interp), 2 or 4 */
int pmg_amr_set_params(pmg_amr *A, int pre, int pst, int max_citer, double crtol, double w0_r,
                       double w0_i, int rop_u, int rop_res, int rop_src, int interp_ord);
/* TF_MGH_INIT injection into the levels below the base (src/gh.c:1334) */
int pmg_amr_init(pmg_amr *A);
/* one V-cycle (amrd/mg.c:84-201); res_out = res[Lf] (max over grids),
 * lev_norms (nlev doubles) = last norm of every level */
int pmg_amr_vcycle(pmg_amr *A, double *res_out, double *lev_norms, int *citer);
/* solve_elliptics loop (amrd/mg.c:56-72) */
int pmg_amr_solve(pmg_amr *A, int min_iter, int max_iter, double tol, double *norms, int *iters);
/* relative solve ||r_k|| <= rtol ||r_0|| over the whole hierarchy (r: the
 * finest level's residual, max over its grids; no counterpart in the
 * reference, whose solve_elliptics loop has an absolute tolerance,
 * amrd/mg.c:56-72).  mode 0: FP64 V-cycles.  mode 1: mixed precision --
 * FP32 V-cycles of the same hierarchy on A e = -r from e = 0 (the FP64
 * defect r = L[u] of every AMR level after u is injected down the levels as
 * the FAS cycle does, written negated as the FP32 source), u += e in FP64 on
 * every AMR level, the FP64 defect as the test; while ||r|| > frac ||r0||
 * (pmg_amr_set_mixed_switch, default 1e-2) FP64 V-cycles are run instead.
 * hist[k] = ||r_{k+1}||; r0_out as pmg_solve_rel (a value > 0 on entry
 * replaces ||r_0||). */
int pmg_amr_solve_rel(pmg_amr *A, int mode, double rtol, int max_iter, double *hist, int *iters, double *r0_out);
/* pmg_amr_solve_rel mode 1: run FP64 V-cycles while the residual is above
 * frac (0..1, default 1e-2; 1 = FP32 from the start) times the reference
 * norm, FP32 inner cycles below.  A fine grid's AMR-boundary values are set
 * only through coarse-grid corrections, which no later defect measures, so
 * the FP32 rounding of the large early corrections stays in them; starting
 * FP32 late bounds that by frac.  No counterpart in the reference (a
 * precision policy of this implementation). */
int pmg_amr_set_mixed_switch(pmg_amr *A, double frac);
/* residual of the finest level (max over grids) */
int pmg_amr_residual(pmg_amr *A, double *norm_out);
/* single transfers of one field: PAMR_sync(L, AMR_bdy_width = bdy),
 * PAMR_inject(Lf -> Lf-1) with restriction op 1/2/3, PAMR_interp(Lc ->
 * Lc+1) second order; pmg_amr_interp_ord at order 2 (PAMR_SECOND_ORDER) or
 * 4 (PAMR_FOURTH_ORDER: dvi4q1d on the transfer's buffered temporary,
 * src/transfer.c:1007-1051) */
int pmg_amr_sync(pmg_amr *A, int L, int bdy, int which);
int pmg_amr_inject(pmg_amr *A, int Lf, int which, int op);
int pmg_amr_interp(pmg_amr *A, int Lc, int which);
int pmg_amr_interp_ord(pmg_amr *A, int Lc, int which, int ord);
/* set_cmask_bdy (which 0) / set_cmask_child (which 1) of grid g as the
 * reference's double mask (1.0 ON, 0.0 OFF), host output */
int pmg_amr_cmask(pmg_amr *A, int L, int g, int which, double *host_out);
/* number of disjoint transfer boxes per level: out[0] sync(bdy 0),
 * out[1] sync(bdy 1), out[2] inject into L, out[3] interp into L */
int pmg_amr_plan_stats(pmg_amr *A, int L, int *out);
/* average duration (ms) of one batched sweep launch over all grids of L */
int pmg_amr_time_sweep(pmg_amr *A, int L, int nrep, double *ms_avg);
int pmg_amr_set_graph(pmg_amr *A, int on);
void *pmg_amr_stream(pmg_amr *A);

/* Further grid functions on every grid of every level (zeroed), for callers
 * that hold more variables than the built-in ones (the PAMR variable table,
 * src/pamr.c:1120-1175).  Returns the new field id (PMG_F_EXTRA, +1, ...)
 * or -1; at most PMG_MAX_EXTRA fields. */
#define PMG_F_EXTRA 16
#define PMG_MAX_EXTRA 64
int pmg_amr_add_field(pmg_amr *A);
/* points of level L's field pool: the grids of L lie back to back from
 * pmg_amr_field(A, L, 0, which), each padded to a multiple of 64 points
 * (padding stays zero unless written) */
long long pmg_amr_pool_points(pmg_amr *A, int L);
/* PAMR_sync / PAMR_inject / PAMR_interp (src/pamr.c:601-685) of several
 * fields at once (the tf-bit-selected gfns of communicate(), src/pamr.c:
 * 533-593); ops[] per field 1 straight, 2 half-weight, 3 full-weight.
 * Enqueued on pmg_amr_stream without a host synchronisation.  sync: bdy
 * (AMR_bdy_width) 0 or 1, otherwise -2. */
int pmg_amr_sync_fields(pmg_amr *A, int L, int bdy, int nf, const int *which);
/* PAMR_sync(L, AMR_bdy_width 1) while only the coarsest grids communicate
 * (set_comm_coarsest, amrd/mg.c:571-586; the transfer lists then skip grids
 * with comm = 0, src/transfer.c:438,482) */
int pmg_amr_sync_fields_coarsest(pmg_amr *A, int L, int nf, const int *which);
int pmg_amr_inject_fields(pmg_amr *A, int Lf, int nf, const int *which, const int *ops);
int pmg_amr_interp_fields(pmg_amr *A, int Lc, int nf, const int *which);
/* the same at order ord (2 or 4, otherwise -2) */
int pmg_amr_interp_fields_ord(pmg_amr *A, int Lc, int ord, int nf, const int *which);
/* The same transfers with excision, as local_transfer does them when
 * PAMR's excision is on (src/transfer.c:946-1163): one transfer of the
 * level's list at a time, in the reference's order, with the device
 * m_util_r8 routines (restriction renormalised over the fine grid's
 * non-excised points; interpolation dminterp3d with do_ex on the buffered
 * temporary, then clear_ex of the whole source and destination grids).
 * mask: the field holding the excision mask (value ex = excised). */
int pmg_amr_inject_fields_ex(pmg_amr *A, int Lf, int nf, const int *which, const int *ops, int mask, double ex);
int pmg_amr_interp_fields_ex(pmg_amr *A, int Lc, int ord, int nf, const int *which, int mask, double ex);
/* set_cmask_bdy (which 0) / set_cmask_child (which 1) of every grid of L
 * (amrd/mg.c:362-488) written into a field as 1.0 / 0.0; enqueued on
 * pmg_amr_stream */
int pmg_amr_fill_cmask(pmg_amr *A, int L, int which, int field);
/* Sweep tiling of level L for the autotuner: tile variant (0 .. count-1:
 * 128x16, 64x16, 32x16, 32x8, 16x16 columns) and z-chunk length.  Results
 * are identical for every tiling (only the order of the norm's partial sums
 * changes); -3 when the level would have more tiles than the norm
 * workspace holds (the previous tiling is kept). */
int pmg_amr_num_sweep_variants(void);
int pmg_amr_set_level_sweep_config(pmg_amr *A, int L, int variant, int zchunk);
int pmg_amr_sweep_config(pmg_amr *A, int L, int *variant, int *zchunk);

#ifdef __cplusplus
}
#endif
#endif
