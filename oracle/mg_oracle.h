/*
 * mg_oracle.h -- CPU oracle for the PAMR/AMRD multigrid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker and as the timed CPU baseline.  No product
 * code links or calls it.
 *
 * Conventions
 *   - every grid function is a double array in Fortran order (x fastest);
 *   - the m_util_r8 restatements keep the Fortran argument meaning: offsets
 *     i1,j1,k1,... are 1-based, strides and counts as in src/m_util_r8.f;
 *   - everything is compiled with -O2 -ffp-contract=off (no FMA), matching
 *     the reference's x86-64 build, so results are bit-reproducible.
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXLEV 16

/* ---- m_util_r8 restatements (src/m_util_r8.f) -------------------------- */
void or_dmcopy3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                 int nx2, int ny2, int nz2, int i1, int j1, int k1,
                 int i2, int j2, int k2, int ni, int nj, int nk,
                 int si1, int sj1, int sk1, int si2, int sj2, int sk2);
void or_dvi4q1d(double *v1, double *v2, int i1, int i2, int n1, int n, int irho, int s1, int s2);
void or_dvi4q1d_ex(double *v1, double *v2, int i1, int i2, int n1, int n, int irho, int s1, int s2,
                   const double *chr1, int i3, int s3, double ex);
void or_dvi2q1d(double *v1, double *v2, int i1, int i2, int n, int irho,
                int s1, int s2);
void or_dvi2q1d_ex(double *v1, double *v2, int i1, int i2, int n1, int n,
                   int irho, int s1, int s2, const double *chr1, int i3,
                   int s3, double ex);
void or_dminterp3d(double *u1, double *u2, int nx1, int ny1, int nz1,
                   int nx2, int ny2, int nz2, int i1, int j1, int k1,
                   int i2, int j2, int k2, int ni, int nj, int nk,
                   int si1, int sj1, int sk1, int irho, int jrho, int krho,
                   int ord, const double *chr1, double ex, int do_ex);
void or_dmhwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                int nx2, int ny2, int nz2, int i1, int j1, int k1,
                int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                const double *chr1, double ex, int do_ex);
void or_dmfwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                int nx2, int ny2, int nz2, int i1, int j1, int k1,
                int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                const double *chr1, double ex, int do_ex);

/* ---- application hooks (w/examples/nps/num_w.f, examples/nbs/num.f) ---- */
/* op = 0: plain Poisson  L[V] = lap_h V - rho            (nps form)
 * op = 1: variable coefficient L[V] = div(a grad V) + c V - rho
 * nx,ny,nz are point counts; the hooks act on points 2..n-1 (1-based)
 * where cmask == 1.0 and (if chr != NULL) chr != ex. */
void or_lop(int op, double *LV, const double *V, const double *rho,
            const double *a, const double *c, const double *cmask,
            const double *chr, double ex, int nx, int ny, int nz,
            double dx, double dy, double dz);
double or_residual(int op, double *res, const double *rhs, const double *V,
                   const double *rho, const double *a, const double *c,
                   const double *cmask, const double *chr, double ex,
                   int nx, int ny, int nz, double dx, double dy, double dz);
double or_relax(int op, double *V, const double *rhs, const double *rho,
                const double *a, const double *c, const double *cmask,
                const double *chr, double ex, int nx, int ny, int nz,
                double dx, double dy, double dz);

/* ---- mg.c mask helpers on one grid (amrd/mg.c:362-488) ----------------- */
void or_set_cmask_bdy(double *mask, const int *shape, const double *gbbox,
                      const double *base_bbox, const double *gdx,
                      const int *periodic);
void or_set_cmask_child(double *mask, const int *shape, const double *gbbox,
                        const double *base_bbox, const double *gdx,
                        const int *periodic, int nchild,
                        const double *child_bboxes);
/* calc_island_no (src/gh.c:1427-1467) */
void or_calc_island_no(const double *bboxes, int num, double *island_no,
                       const double *dx, int cf, int dim,
                       const int *min_cwidth);
/* calc_gs_ibbox (src/transfer.c:73-95) for one segment of one grid */
void or_calc_gs_ibbox(const double *seg_bbox, const double *g_bbox,
                      const int *g_shape, int dim, int *ibbox, double *dx);

/* ---- uniform-hierarchy FAS driver (amrd/mg.c:20-201) -------------------- */
typedef struct or_mg or_mg;
/* Build the MG hierarchy of a single base grid nx x ny x nz over bbox,
 * coarsening by 2 while calc_island_no allows (min coarse width mcw). */
or_mg *or_mg_create(int nx, int ny, int nz, const double *bbox, int mcw,
                    int op, int do_ex, double ex);
void or_mg_destroy(or_mg *m);
int or_mg_nlev(const or_mg *m);
void or_mg_shape(const or_mg *m, int lev, int *shape);
/* which: 0 u, 1 rho, 2 rhs, 3 res, 4 lop, 5 cmask, 6 a, 7 c, 8 chr */
double *or_mg_field(or_mg *m, int lev, int which);
/* pre, pst, max_citer, crtol, w0_r, w0_i, restriction tf types (1,2,3) for
 * u, u_res and the sources/coefficients, interpolation order of u_res */
void or_mg_set_params(or_mg *m, int pre, int pst, int max_citer,
                      double crtol, double w0_r, double w0_i, int rop_u,
                      int rop_res, int rop_src, int interp_ord);
/* PAMR_build_mgh initial injection (TF_MGH_INIT) of u, rho, a, c, chr */
void or_mg_init(or_mg *m);
/* one V-cycle; returns res[Lf]; lev_norms (nlev+1 doubles, may be NULL)
 * receives the last relax/residual norm per level; citer receives the
 * iteration count of the coarsest loop. */
double or_mg_vcycle(or_mg *m, double *lev_norms, int *citer);
/* solve_elliptics loop; norms_out receives one norm per cycle */
int or_mg_solve(or_mg *m, int min_iter, int max_iter, double tol,
                double *norms_out);
/* finest-level residual (writes res on Lf) */
double or_mg_residual_fine(or_mg *m);

/* ---- learned-smoother network oracle (3x3x3 convs, FP64) ---------------- */
/* Channels: 1 -> H -> H -> 1, GELU (tanh form) after layers 1 and 2,
 * zero padding.  Weights layout: w1[H][1][27], b1[H], w2[H][H][27], b2[H],
 * w3[1][H][27], b3[1].  x, y are n0 x n1 x n2 (x fastest). */
void or_net_forward(int H, const double *w1, const double *b1,
                    const double *w2, const double *b2, const double *w3,
                    const double *b3, const double *x, double *y,
                    int n0, int n1, int n2);
/* Loss = 0.5 * mean((y - t)^2) over all points; fills gradients. */
double or_net_loss_grad(int H, const double *w1, const double *b1,
                        const double *w2, const double *b2, const double *w3,
                        const double *b3, const double *x, const double *t,
                        int n0, int n1, int n2, double *gw1, double *gb1,
                        double *gw2, double *gb2, double *gw3, double *gb3);

#ifdef __cplusplus
}
#endif
#endif
