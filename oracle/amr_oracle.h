/*
 * amr_oracle.h -- CPU oracle for the multigrid solve over an AMR hierarchy
 * (several grids per level).
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of amrd/mg.c driving
 * PAMR's sync / inject / interp (src/pamr.c:533-685) with the grid-segment
 * bookkeeping of src/transfer.c executed literally: transfer lists are built
 * in the reference's order and applied one after the other, so overlapping
 * writes resolve exactly as in the reference.  Used only by tests/ as the
 * checker of the batched GPU path.
 *
 * Hierarchy model (single MPI rank, ghost_width 0, refinement ratio 2):
 *   AMR level 1 is one base grid; AMR level l > 1 holds nbox[l-2] grids
 *   given as physical bounding boxes; the MG hierarchy adds coarsened copies
 *   of the base below it (build_mgh, src/gh.c:1222-1365).  Grids of a level
 *   are ordered by insert_grid (src/gh.c:235-257), applied as build_sgh and
 *   compose_cgh apply it.
 */
#ifndef AMR_ORACLE_H
#define AMR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_amr or_amr;

/* field ids of or_amr_field */
enum { OA_U = 0, OA_RHO, OA_RHS, OA_RES, OA_LOP, OA_CMASK, OA_A, OA_C, OA_TRE, OA_CHR, OA_NUM };

/* base_shape: points of the base grid over bbox; namr: number of refined
 * AMR levels; nbox[l]: boxes of refined level l+2; boxes: 6 doubles per box
 * (x0,x1,y0,y1,z0,z1), all levels concatenated in the caller's order. */
or_amr *or_amr_create(const int *base_shape, const double *bbox, int namr,
                      const int *nbox, const double *boxes, int mcw, int op);
void or_amr_destroy(or_amr *m);
/* the MG hierarchy PAMR_build_mgh(amr_lev, finest, tl) builds from AMR level
 * amr_lev > 1 upward (src/gh.c:1222-1365): islands of that level's boxes
 * coarsened below it, its own levels sharing m's grid functions (destroy it
 * before m) */
or_amr *or_amr_create_mgh(or_amr *m, int amr_lev);
/* 1 when grid g of level L is flagged coarsest */
int or_amr_grid_coarsest(const or_amr *m, int L, int g);
int or_amr_nlev(const or_amr *m);
/* MG level of AMR level 1 (the base grid) */
int or_amr_base_level(const or_amr *m);
int or_amr_ngrids(const or_amr *m, int L);
/* grid g (0-based, level order) of MG level L: shape and bbox */
void or_amr_grid_info(const or_amr *m, int L, int g, int *shape, double *bbox);
double *or_amr_field(or_amr *m, int L, int g, int which);
void or_amr_set_params(or_amr *m, int pre, int pst, int max_citer,
                       double crtol, double w0_r, double w0_i, int rop_u,
                       int rop_res, int rop_src, int interp_ord);
/* TF_MGH_INIT injection into the MG levels below the base */
void or_amr_init(or_amr *m);
/* compute_MG_rhs's truncation-error estimate (amrd/mg.c:288-295,306-313):
 * with save_tre set, every V-cycle writes |lop| * sqrt(ntre) into OA_TRE
 * on each coarse level, syncs it and interpolates it to the next finer
 * level.  ntre is AMRD_num_f_tre_vars. */
void or_amr_set_tre(or_amr *m, int save_tre, int ntre);
/* PAMR's excision (src/pamr.c:1983-2021): with do_ex set, OA_CHR of every
 * grid is the mask (ex = excised; the caller fills it on every level), the
 * transfers take local_transfer's do_ex paths (renormalised restriction,
 * dminterp3d with the source mask, clear_ex of source and destination,
 * src/transfer.c:946-1163) and the hooks skip excised points */
void or_amr_set_excision(or_amr *m, int do_ex, double ex);
/* one FAS V-cycle over the hierarchy; returns res[Lf] (max over the grids
 * of Lf); lev_norms (nlev doubles) gets the last norm of every level */
double or_amr_vcycle(or_amr *m, double *lev_norms, int *citer);
/* finest-level residual (max over grids), as the final residual call */
double or_amr_residual_fine(or_amr *m);

/* ---- bookkeeping exports (for tests) ------------------------------------ */
/* owned segments of level L (build_owned_gsl, src/transfer.c:415-458):
 * writes grid index and bbox of each segment; returns the count */
int or_amr_owned(or_amr *m, int L, int interior, int extend, int *gidx,
                 double *bboxes, int cap);
/* transfer list of one communicate() call: kind 0 sync(L, AMR_bdy_width =
 * arg), 1 inject(L+1 -> L), 2 interp(L-1 -> L).  Per pair 16 ints:
 * src grid, dst grid, delta, irho, src ibbox[6], dst ibbox[6] (1-based,
 * after align_src_bbox); returns the number of pairs */
int or_amr_transfers(or_amr *m, int kind, int L, int arg, int *out, int cap);
/* the cmask rules of one level: which 0 set_cmask_bdy, 1 set_cmask_child */
void or_amr_set_cmask(or_amr *m, int L, int which);
/* one transfer of one field: PAMR_sync(L, AMR_bdy_width = bdy),
 * PAMR_inject(Lf -> Lf-1) with restriction op (1, 2, 3), and
 * PAMR_interp(Lc -> Lc+1) (second order) */
void or_amr_sync_field(or_amr *m, int L, int bdy, int which);
void or_amr_inject_field(or_amr *m, int Lf, int which, int op);
void or_amr_interp_field(or_amr *m, int Lc, int which);
/* PAMR_interp at order ord: 2 (PAMR_SECOND_ORDER) or 4 (PAMR_FOURTH_ORDER) */
void or_amr_interp_field_ord(or_amr *m, int Lc, int which, int ord);

#ifdef __cplusplus
}
#endif
#endif
