/*
 * pamr_mg_gpu.h -- C ABI of the MI355X (gfx950) multigrid hot path.
 *
 * Three layers, all plain C with device pointers and sizes:
 *
 *  1. m_util_r8 transfer routines on device memory.  Same arguments and
 *     meaning as the Fortran routines (include/m_util_r8.h:10-43 of the
 *     reference), passed by value, plus a HIP stream.  The reference's silent
 *     bounds-check returns are kept (return 0, nothing written).
 *
 *  2. The application MG hooks (include/amrd.h:137-159): red-black
 *     Gauss-Seidel relax, residual and L_op for the plain Poisson operator
 *     (op 0, w/examples/nps/num_w.f:18-157) and a variable-coefficient
 *     operator (op 1).  cmask/chr are double grid functions as in AMRD.
 *
 *  3. pmg_solver: a device-resident MG hierarchy over one base grid with
 *     the FAS V-cycle of amrd/mg.c:84-201, graph-captured, in FP64 (the
 *     reference method) or with FP32 inner cycles (mixed precision).
 *
 * Return codes: 0 success, negative on error (a message goes to stderr).
 * The shared library is libpamr_mg.so (package directory pamr_amrd_amd/).
 */
#ifndef PAMR_MG_GPU_H
#define PAMR_MG_GPU_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. m_util_r8 on device ------------------------------------------- */

/* replaces dmcopy3d_ (src/m_util_r8.f:20-50; include/m_util_r8.h:10-13) */
int pmg_dmcopy3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                 int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj,
                 int nk, int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                 hipStream_t stream);

/* replaces dmfwr3d_ incl. dmfwr3d_ex (src/m_util_r8.f:733-949;
 * include/m_util_r8.h:27-31) */
int pmg_dmfwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2, const double *chr1,
                double ex, int do_ex, hipStream_t stream);

/* replaces dmhwr3d_ (src/m_util_r8.f:649-723; include/m_util_r8.h:21-25) */
int pmg_dmhwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2, const double *chr1,
                double ex, int do_ex, hipStream_t stream);

/* replaces dminterp3d_ (src/m_util_r8.f:473-637; include/m_util_r8.h:15-19):
 * ord = 2 (dvi2q1d / dvi2q1d_ex, :332-421) and ord = 4 (dvi4q1d /
 * dvi4q1d_ex, :65-319; with excision rho <= 20, else -2). */
int pmg_dminterp3d(double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2, int nz2,
                   int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                   int si1, int sj1, int sk1, int irho, int jrho, int krho, int ord,
                   const double *chr1, double ex, int do_ex, hipStream_t stream);

/* ---- 2. application MG hooks on device --------------------------------- */

/* One red-black GS sweep; replaces the app_MG_relax hook body
 * (include/amrd.h:148-152; w/examples/nps/num_w.f:101-157).  The RMS of the
 * residuals evaluated during the sweep is written to *norm_dev. */
int pmg_hook_relax(int op, double *V, const double *rhs, const double *rho, const double *a,
                   const double *c, const double *cmask, const double *chr, double ex, int nx,
                   int ny, int nz, double dx, double dy, double dz, double *norm_dev,
                   hipStream_t stream);

/* No counterpart in the reference: pmg_hook_relax for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, same results bit for
 * bit; it calls no allocating routine.  Its partial-norm scratch is a fixed
 * 16 MiB device array of the library (one per GPU, held until the process
 * ends, located on the first call on each device: call it with the stream's
 * device current), which limits the grid to (nx-2)(ny-2)(nz-2) <= 2^27
 * interior points (-2 beyond, nothing written).  Every call on a GPU uses
 * that GPU's one array: calls that overlap on one GPU (other streams or
 * threads) race. */
int pmg_hook_relax_noalloc(int op, double *V, const double *rhs, const double *rho, const double *a,
                           const double *c, const double *cmask, const double *chr, double ex, int nx,
                           int ny, int nz, double dx, double dy, double dz, double *norm_dev,
                           hipStream_t stream);

/* res = L[V] - rhs where cmask ON; replaces app_MG_residual
 * (include/amrd.h:137-141; w/examples/nps/num_w.f:59-90) */
int pmg_hook_residual(int op, double *res, const double *rhs, const double *V, const double *rho,
                      const double *a, const double *c, const double *cmask, const double *chr,
                      double ex, int nx, int ny, int nz, double dx, double dy, double dz,
                      double *norm_dev, hipStream_t stream);

/* LV = L[V] where cmask ON; replaces app_L_op (include/amrd.h:155-159;
 * w/examples/nps/num_w.f:18-52) */
int pmg_hook_lop(int op, double *LV, const double *V, const double *rho, const double *a,
                 const double *c, const double *cmask, const double *chr, double ex, int nx,
                 int ny, int nz, double dx, double dy, double dz, hipStream_t stream);

/* ---- 3. uniform-hierarchy solver --------------------------------------- */

typedef struct pmg_solver pmg_solver;

/* grid-function selectors for pmg_field / pmg_upload / pmg_download */
#define PMG_F_U 0   /* the elliptic variable f (amrd: AMRD_MG_f) */
#define PMG_F_RHO 1 /* source term of the operator (nps: rho) */
#define PMG_F_RHS 2 /* f_rhs */
#define PMG_F_RES 3 /* f_res */
#define PMG_F_LOP 4 /* f_lop */
#define PMG_F_A 5   /* op 1: coefficient a */
#define PMG_F_C 6   /* op 1: coefficient c */

/* Builds the MG hierarchy of one nx x ny x nz vertex grid over bbox
 * (x0,x1,y0,y1,z0,z1): coarsening by 2 while calc_island_no allows it with
 * minimum coarse width min_cwidth (PAMR_build_mgh, src/gh.c:1222-1365,
 * 1427-1467).  Replaces the MG-hierarchy part of PAMR_build_mgh
 * (src/pamr.c:354-378) for a single-grid base level.  op: 0 Poisson,
 * 1 variable coefficient; do_ex/ex: excision with chr == ex excised. */
pmg_solver *pmg_create(int nx, int ny, int nz, const double *bbox, int min_cwidth, int op,
                       int do_ex, double ex);
/* PAMR_destroy_mgh (src/gh.c:1371-1416) */
void pmg_destroy(pmg_solver *S);
int pmg_nlev(const pmg_solver *S);
/* shape of level L (1 = coarsest) */
int pmg_level_shape(const pmg_solver *S, int L, int *shape);
/* AMRD MG parameters (amrd/io.c:97-108,269-285): MG_pre_swp, MG_pst_swp,
 * MG_max_citer, MG_crtol, MG_w0_r, MG_w0_i; restriction tf type (1 straight,
 * 2 half weight, 3 full weight) of f, f_res and the sources; interp order 2 */
int pmg_set_params(pmg_solver *S, int pre, int pst, int max_citer, double crtol, double w0_r,
                   double w0_i, int rop_u, int rop_res, int rop_src, int interp_ord);
/* device pointer of a grid function on level L */
double *pmg_field(pmg_solver *S, int L, int which);
int pmg_upload(pmg_solver *S, int L, int which, const double *host);
int pmg_download(pmg_solver *S, int L, int which, double *host);
/* chr grid function of level L (values == ex are excised) */
int pmg_set_chr(pmg_solver *S, int L, const double *chr_host);
/* materialise AMRD's cmask of level L: kind 0 set_cmask_bdy, 1 set_cmask_child */
int pmg_download_cmask(pmg_solver *S, int L, int kind, double *host);
/* PAMR_build_mgh's initial injection (TF_MGH_INIT) and MG variable reset */
int pmg_init(pmg_solver *S);
/* one V-cycle (amrd/mg.c:84-201); *res_out = RMS residual on the finest
 * level; lev_norms[L-1] = last relax norm per level (nlev doubles);
 * *citer = iterations of the coarsest-level loop */
int pmg_vcycle(pmg_solver *S, double *res_out, double *lev_norms, int *citer);
int pmg_vcycle_async(pmg_solver *S);
int pmg_read_norms(pmg_solver *S, double *res_out, double *lev_norms, int *citer);
/* RMS residual of the finest level (writes f_res there) */
int pmg_residual(pmg_solver *S, double *norm_out);
/* solve_elliptics loop (amrd/mg.c:56-72), absolute tolerance */
int pmg_solve(pmg_solver *S, int min_iter, int max_iter, double tol, double *norms, int *iters);
/* relative solve ||r_k|| <= rtol ||r_0||; mode 0 FP64 V-cycles, 1 FP32
 * inner V-cycles in FP64 defect correction, 2 as 1 with the learned
 * correction (pmg_set_net; -1 on a solver with excision), 3 as 1 with the
 * learned band correction next to the excised region (pmg_band_setup,
 * pmg_set_band_net); hist[k] = ||r_{k+1}||.  r0_out: on entry, a
 * value > 0 is taken as the reference norm instead of ||r_0|| of the
 * initial u (warm starts measure against the residual of u = 0); on exit
 * the reference norm used. */
int pmg_solve_rel(pmg_solver *S, int mode, double rtol, int max_iter, double *hist, int *iters,
                  double *r0_out);
/* Relative solve as pmg_solve_rel with restarted flexible GMRES(m)
 * (1 <= m <= 15; no counterpart in the reference): the operator is the
 * finest level's FP64 residual hook with a zero source, the right
 * preconditioner one FP32 V-cycle from zero (the mixed mode's inner cycle).
 * hist[k]: the residual estimate after preconditioner application k+1, the
 * FP64 residual at the end of each restart; iters counts preconditioner
 * applications.  Holds 2m + 2 finest-level FP64 vectors. */
int pmg_solve_fgmres(pmg_solver *S, int m, double rtol, int max_iter, double *hist, int *iters,
                     double *r0_out);
int pmg_last_citer(pmg_solver *S, int which);
/* FP32 V-cycles per FGMRES preconditioner application (1..4, default 1);
 * with the band correction on (and more than one cycle) it runs between
 * the cycles instead of after the single one */
int pmg_set_fgmres_prec_cycles(pmg_solver *S, int cycles);
/* initial guess u = 0 on the finest level (device-side) */
int pmg_zero_solution(pmg_solver *S);
/* average duration (ms) of one launch of a solver kernel on level L, timed
 * with HIP events on the solver's stream: which 0 = red-black sweep,
 * 1 = residual, 2 = the finest residual fused with the full-weight
 * injection into level L-1 (L the finest level; -1 where that fusion does
 * not apply; it overwrites the coarse u and f_res), 3 = the finest FP64 sweep
 * inside nrep ungraphed V-cycles (events around each pre- and post-smoothing
 * group; the cycles advance u); prec 0 = FP64, 1 = FP32 */
int pmg_time_kernel(pmg_solver *S, int which, int L, int prec, int nrep, double *ms_avg);
/* execution tuning: sweep tile variant and z chunk; graph capture on/off */
int pmg_set_sweep_config(pmg_solver *S, int variant, int zchunk);
int pmg_set_level_sweep_config(pmg_solver *S, int L, int variant, int zchunk);
/* the same for one precision only: prec 0 the FP64 sweeps, 1 the FP32 sweeps
 * of the mixed-precision inner cycles */
int pmg_set_level_sweep_config_prec(pmg_solver *S, int L, int prec, int variant, int zchunk);
int pmg_num_sweep_variants(void);
int pmg_set_graph(pmg_solver *S, int on);
/* run the small bottom levels (all dimensions <= 17 points, below the
 * finest level, no excision) as one single-workgroup kernel (default on);
 * on = 0 off, 1 the default, 2..17 the largest side fused (9: only the
 * levels of 9 points and below); results are identical either way */
int pmg_set_bottom_fusion(pmg_solver *S, int on);
/* keep every array of the fused bottom levels in LDS for the whole bottom V
 * when they fit in one workgroup's LDS (on = 1, the default) or run it from
 * global memory (0); results are identical either way.  No counterpart in
 * the reference (a tuning switch of this implementation). */
int pmg_set_bottom_lds(pmg_solver *S, int on);
/* FP32 inner cycles (mixed precision, FGMRES preconditioner) on the plain
 * Poisson operator without excision: solve a 17^3 top bottom level exactly
 * by the discrete sine transform instead of running the V-cycle below it
 * (on = 1, the default; no counterpart in the reference, which relaxes
 * there; the FP64 cycles keep the reference's method) */
int pmg_set_fp32_exact_bottom(pmg_solver *S, int on);
/* compute the finest-level residual (of the FP64 cycle and of the FP32 inner
 * cycle) inside the full-weight injection of u and f_res instead of storing
 * it (default on; no excision, full weighting for both); u, the norms and
 * every coarse-level field are identical either way, but with the fusion on
 * the finest level's f_res (PMG_F_RES) is not stored by the V-cycle.
 * on = 1: a ring of three fine residual planes per workgroup; on = 2 (the
 * default): the restriction carried through the z march in registers; on =
 * 3: the same with two fine planes of u and rho in flight instead of one
 * (same results) */
int pmg_set_restrict_fusion(pmg_solver *S, int on);
/* coarse planes per workgroup of that fused kernel (prec 0 FP64 cycle, 1 FP32
 * inner cycle; 1..64; default 6 / 4 when the coarse level has 129 to 256
 * planes, else 8 / 8); results are identical for every value.  -3 (value kept)
 * when the launch would need more workgroups than the solver's partial-norm
 * workspace holds (small values on large grids) */
int pmg_set_resid_restrict_zchunk(pmg_solver *S, int prec, int czchunk);
/* run pairs of red-black sweeps on levels of at least 65 points per side as
 * one temporally blocked pass (default off: at 257^3 a pair takes 0.26 ms
 * against 2 x 0.11 ms for single sweeps; zchunk > 0 sets its z depth);
 * results are identical either way */
int pmg_set_sweep_blocking(pmg_solver *S, int on, int zchunk);
void *pmg_stream(pmg_solver *S);

/* ---- 4. learned smoother / warm-start network (no reference counterpart:
 * the learned part of the accelerated solve).  3x3x3 convolutions 1 -> 16 -> 16 -> 1,
 * GELU, BF16 MFMA with FP32 accumulation. ---------------------------------- */
typedef struct pmg_net pmg_net;
pmg_net *pmg_net_create(int hidden, unsigned long long seed);
void pmg_net_destroy(pmg_net *N);
/* weights in the oracle layout: w1[16][1][27], b1[16], w2[16][16][27],
 * b2[16], w3[1][16][27], b3[1] (host FP64) */
int pmg_net_set_weights(pmg_net *N, const double *w1, const double *b1, const double *w2,
                        const double *b2, const double *w3, const double *b3);
int pmg_net_get_weights(pmg_net *N, double *w1, double *b1, double *w2, double *b2, double *w3,
                        double *b3);
/* y = N(x) on nb boxes of nx*ny*nz points (device FP32, zero padding) */
int pmg_net_forward(pmg_net *N, const float *x, float *y, int nx, int ny, int nz, int nb);
/* loss 0.5*mean((N(x)-t)^2) over points at least `margin` from a box face,
 * and its gradients (host FP64, oracle layout) */
int pmg_net_loss_grad(pmg_net *N, const float *x, const float *t, int nx, int ny, int nz, int nb,
                      int margin, double *loss, double *gw1, double *gb1, double *gw2, double *gb2,
                      double *gw3, double *gb3);
/* one Adam step (lr) on a batch; *loss = loss before the step */
int pmg_net_train_step(pmg_net *N, const float *x, const float *t, int nx, int ny, int nz, int nb,
                       int margin, float lr, double *loss);
void *pmg_net_stream(pmg_net *N);
/* learned correction on one grid, on the caller's stream: y = N(scale*x),
 * u += coef*y at interior points; pmg_net_recorrect repeats/undoes it */
int pmg_net_correct(pmg_net *N, const float *x, float scale, double *u, double coef, int nx,
                    int ny, int nz, void *stream);
int pmg_net_recorrect(pmg_net *N, double *u, double coef, int nx, int ny, int nz, void *stream);
/* the inference path pmg_net_correct uses: y = N(scale * x) on nb boxes
 * (device fp32 [nb][P] in and out), layer 1 read straight from x, no
 * pre-activations stored; equal to pmg_net_forward to FP32 summation order */
int pmg_net_infer(pmg_net *N, const float *x, float scale, float *y, int nx, int ny, int nz, int nb);
/* allocates the inference buffers for `points` points and packs the weight
 * fragments ahead of time (pmg_net_correct would do it on first use) */
int pmg_net_reserve(pmg_net *N, long long points);
/* training crops: out[c] = scale*(a - b) on side^3 cubes at host origins */
int pmg_gather_crops(const double *a, const double *b, int nx, int ny, int nz, const int *origins_host,
                     int ncrop, int side, double scale, float *out, void *stream);
/* attach a network to a solver (pmg_solve_rel mode 2) and read its stats.
 * Mode 2 applies the network before an FP32 cycle and keeps the correction
 * only when its measured contraction per unit time beats the cycle's;
 * after a rejection it is not tried for 20 solves, a period that doubles
 * with every further rejection (at most 1280) and returns to 20 on an
 * acceptance or a new pmg_set_net. */
int pmg_set_net(pmg_solver *S, pmg_net *N);
int pmg_net_stats(pmg_solver *S, int *accepted, int *rejected, double *t_net_ms, double *t_vc_ms);
/* finest-level solution u = ca*a + cb*b (device FP64; b may be NULL) */
int pmg_set_solution_lincomb(pmg_solver *S, const double *a, const double *b, double ca, double cb);
/* Warm start from earlier solves of a sequence (no counterpart in the
 * reference): F[0..K-1] earlier finest-level sources (the rho field's
 * convention, newest first) and U[k] their solutions, device FP64 arrays of
 * the finest level's shape, 1 <= K <= 8.  Chooses c minimising
 * ||rho - sum_k c_k F[k]||_2 (rho: the solver's current finest source) and
 * sets the finest u (both sweep buffers) to sum_k c_k U[k]; L being linear,
 * the new residual is that misfit plus sum_k c_k r_k (r_k: the residuals the
 * earlier solves stopped at).  c_out: the K coefficients (may be NULL);
 * rel_out: when not NULL, ||rho - sum c_k F[k]|| / ||rho|| measured by one
 * more pass over the sources.  Synchronises the solver's stream; call
 * pmg_init afterwards, as after any new initial guess. */
int pmg_history_warm_start(pmg_solver *S, int K, const double *const *F, const double *const *U, double *c_out,
                           double *rel_out);
/* ---- 4b. learned correction next to excised regions (no reference
 * counterpart).  With excision the reference's transfers (renormalised
 * restriction, excised interpolation, clear_ex: src/m_util_r8.f:835-949,
 * 368-421, src/transfer.c:930-935) leave the coarse-grid correction short
 * in a thin layer of fine points next to each excised region.  The network
 * corrects u only there: the "band" of updatable points within Euclidean
 * distance `depth` (grid points) of an excised point, covered by cubic
 * cores; a correction gathers a crop of side core + 6 around each core that
 * holds band points, runs the network on the crops as one batch, and adds
 * ts * ||r|| * N(r / ||r||) to u at the band points of each core. */
/* band mask on the host (no GPU): band[p] = 1 at interior, non-excised
 * points (chr != ex) within distance depth of an excised point; returns the
 * number of band points (-1 on bad arguments) */
long long pmg_band_mask(const double *chr, int nx, int ny, int nz, double ex, int depth, unsigned char *band);
/* crop origins (3 ints each, core corner - margin) of the side-`core` cubes
 * tiling the grid from index 0 that hold band points; returns their number */
int pmg_band_cores(const unsigned char *band, int nx, int ny, int nz, int core, int margin, int *org, int cap);
/* band plan of a solver with excision, from its finest level's chr (set
 * pmg_set_chr first); returns the number of cores (or -1) */
int pmg_band_setup(pmg_solver *S, int depth, int core);
int pmg_band_info(pmg_solver *S, long long *npts, int *ncore, int *side);
int pmg_band_download(pmg_solver *S, unsigned char *host);
/* attach the network (NULL detaches): use_dc enables it in pmg_solve_rel
 * mode 3, use_fgmres in pmg_solve_fgmres's preconditioner.  Mode 3 judges
 * every iteration by the FP64 residual and drops the correction for the rest
 * of the solve when an iteration's contraction per unit time is worse than
 * the plain iteration's; FGMRES compares whole solves (the first solve runs
 * plain to measure the plain rate; after a worse solve with the correction
 * the next four run plain). */
int pmg_set_band_net(pmg_solver *S, pmg_net *N, double ts, int use_dc, int use_fgmres);
/* switch the two uses on or off without resetting the measurements (plain
 * FGMRES solves measure the plain rate whenever a network is attached);
 * use_fgmres = 2 applies the correction in every FGMRES solve regardless of
 * the measured rates (to measure its effect) */
int pmg_band_enable(pmg_solver *S, int use_dc, int use_fgmres);
/* place the correction inside the FP32 V-cycle, after the finest level's
 * coarse-grid correction and before its post-smoothing (on = 1), instead of
 * after the cycle (0, the default): u32 += ts * |src| * N(r32 / |src|) with
 * r32 the FP32 residual at that moment and |src| the RMS of the cycle's
 * finest source (the outer defect, or 1/sqrt(updatable points) for the unit
 * Krylov vectors of FGMRES) */
int pmg_band_set_inner(pmg_solver *S, int on);
/* the network of the in-cycle placement and its output scale (NULL detaches) */
int pmg_set_band_net_inner(pmg_solver *S, pmg_net *N, double ts);
/* training crops for the in-cycle placement: one mixed iteration whose FP32
 * cycle writes, after the finest coarse-grid correction, x = r32 / |r| and
 * t = (ustar - u - u32) / |r| (|r| the outer defect's RMS), w = band, on
 * crops centred on seeded random band points; *rnorm = the outer defect's
 * RMS after the iteration.  Needs the state of a mode-1 pmg_solve_rel. */
int pmg_band_sample_inner(pmg_solver *S, const double *ustar, int ncrop, unsigned long long seed, float *x, float *t,
                          float *w, double *rnorm);
/* accepted / rejected trials; out6 = plain and corrected contraction and ms
 * per defect-correction iteration, plain and corrected FGMRES rates
 * (log10 of the residual reduction per ms) */
int pmg_band_stats(pmg_solver *S, int *accepted, int *rejected, double *out6);
/* training crops (device fp32 [ncrop][side^3] each) centred on seeded random
 * band points: x = xs * (the finest FP32 source, -r in defect correction),
 * t = ts * (ustar - u), w = 1 at band points, 0 elsewhere (zero outside the
 * grid); ustar device FP64 of the finest level's shape */
int pmg_band_sample(pmg_solver *S, const double *ustar, double xs, double ts, int ncrop, unsigned long long seed,
                    float *x, float *t, float *w);
/* one correction now: which 0 the defect-correction form (needs the state
 * a mode-1/3 pmg_solve_rel leaves: FP32 source -r, its norm), 1 the
 * preconditioner form (FP32 u against the FP32 source) */
int pmg_band_correct(pmg_solver *S, int which);
/* FP32 arrays of the mixed-precision cycles: which 0 u, 1 rho (the source),
 * 2 rhs, 3 res, 4 lop (download); 0 and 1 (upload) */
int pmg_download32(pmg_solver *S, int L, int which, float *host);
int pmg_upload32(pmg_solver *S, int L, int which, const float *host);
/* weighted forms of pmg_net_loss_grad / pmg_net_train_step: the mean runs
 * over points with w != 0 (device fp32 [nb][P]) past the margin; w NULL in
 * the train step = every point past the margin */
int pmg_net_loss_grad_w(pmg_net *N, const float *x, const float *t, const float *w, int nx, int ny, int nz,
                        int nb, int margin, double *loss, double *gw1, double *gb1, double *gw2, double *gb2,
                        double *gw3, double *gb3);
int pmg_net_train_step_w(pmg_net *N, const float *x, const float *t, const float *w, int nx, int ny, int nz,
                         int nb, int margin, float lr, double *loss);
/* y = N(scale * x) on nb boxes, enqueued on `stream` without
 * synchronisation; returns the network's output buffer (fp32 [nb][P], valid
 * until its next inference) or NULL */
float *pmg_net_infer_async(pmg_net *N, const float *x, float scale, int nx, int ny, int nz, int nb, void *stream);
/* device launchers of the band correction on raw pointers (org: device
 * crop origins): gather out = mul / *norm * src (mul alone when norm is
 * NULL), zero outside the grid; scatter u += mul * *norm * y at band points
 * of each core (into u and u2, or into u32 when u is NULL); resid: -L64[u]
 * in FP32 at the updatable points of each core grown by one point */
int pmg_band_gather_f32(const float *src, int nx, int ny, int nz, const int *org, int ncrop, int side, float mul,
                        const double *norm, float *out, void *stream);
int pmg_band_scatter(double *u, double *u2, float *u32, const float *y, const unsigned char *band, int nx, int ny,
                     int nz, const int *org, int ncrop, int side, int margin, double mul, const double *norm,
                     void *stream);
int pmg_band_resid(int op, int do_ex, const double *u, const double *rho, const double *a, const double *c,
                   const unsigned char *chr, float *neg32, int nx, int ny, int nz, const int *org, int ncrop, int side,
                   int margin, double dx, double dy, double dz, void *stream);
int pmg_band_pairs(const float *xsrc, float xs, const double *ustar, const double *u, double ts,
                   const unsigned char *band, int nx, int ny, int nz, const int *org, int ncrop, int side, float *xo,
                   float *to, float *wo, void *stream);
/* the same inside the FP32 cycle: gather x = mul / *norm * r32 (r32 the
 * FP32 residual of u with source rho, rhs 0; 0 off the updatable points);
 * pairs x = r32 / *norm, t = (ustar - u64 - u) / *norm, w = band */
int pmg_band_gather_res32(int op, int do_ex, const float *u, const float *rho, const float *a, const float *c,
                          const unsigned char *chr, int nx, int ny, int nz, const int *org, int ncrop, int side,
                          float mul, const double *norm, double dx, double dy, double dz, float *out, void *stream);
int pmg_band_pairs_res32(int op, int do_ex, const float *u, const float *rho, const float *a, const float *c,
                         const unsigned char *chr, const double *ustar, const double *u64, const unsigned char *band,
                         int nx, int ny, int nz, const int *org, int ncrop, int side, const double *norm, double dx,
                         double dy, double dz, float *xo, float *to, float *wo, void *stream);
/* lane-layout self-test of the BF16 16x16x16 MFMA used by the network:
 * D = bf16(A) * bf16(B) (host 16x16 row-major) */
int pmg_net_mfma_selftest(const float *A, const float *B, float *D_out);
/* the same for the K = 32 instruction the inference kernels use: A 16x32,
 * B 32x16 (row-major host FP32, rounded to BF16), D = A*B 16x16 */
int pmg_net_mfma32_selftest(const float *A, const float *B, float *D_out);

#ifdef __cplusplus
}
#endif
#endif
