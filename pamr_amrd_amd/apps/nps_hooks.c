/*
 * nps_hooks.c -- the MG hooks of the nps example application
 * (w/examples/nps/num_w.f:18-157) written against pamr.h / amrd_mg.h for
 * device-resident grid functions.
 *
 * Glue as the reference applications do it (examples/nbs/nbs.c:71-128):
 * ldptr() fetches the current grid with PAMR_get_g_attribs and picks the
 * grid functions out of gfs[] with the gfns PAMR_get_gfn returned.  The
 * hooks then launch the library's device hook kernels (pmg_hook_relax /
 * pmg_hook_residual / pmg_hook_lop) on amrd_gpu_stream(); the two that
 * return a norm copy it to the host.
 *
 * Variables the application defines: V (the elliptic variable), rho (its
 * source; L[V] = lap V + rho, the "rhs on the lhs" convention of
 * include/amrd.h:201-203), and for op 1 the coefficients a, c
 * (L[V] = div(a grad V) + c V + rho).  amrd_mg_def_vars adds V_res, V_lop,
 * V_rhs and cmask (and chr with excision, nps_define_vars_ex).
 *
 * The level hooks (nps_set_level_hooks) launch the same kernels for every
 * grid of a level and read the norms back once per level.
 */
#include <stdio.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "amrd_mg.h"
#include "pamr.h"
#include "pamr_mg_gpu.h"

#define NPS_MAX_GRIDS 4096

static int op;
static int V_gfn, rho_gfn, a_gfn, c_gfn, V_res_gfn, V_rhs_gfn, V_lop_gfn, cmask_gfn, chr_gfn;
static double *V, *rho, *a, *c, *V_res, *V_rhs, *V_lop, *cmask, *chr;
static int do_ex;      /* excision: the hooks skip points where chr == ex_val */
static double ex_val;
static int shape[3];
static double dx, dy, dz;
static double *d_norm; /* device norms, one per grid of a level */
static double *h_norm; /* pinned host copy */

static hipStream_t stream(void) { return (hipStream_t)amrd_gpu_stream(); }

static int ldptr(void)
{
  int rank, dim, ngfs, gw[6];
  double bbox[6], t, *x[3], *gfs[512];

  if (!PAMR_get_g_attribs(&rank, &dim, shape, bbox, gw, &t, &ngfs, x, gfs)) {
    printf("ldptr: PAMR_get_g_attribs failed\n");
    return 0;
  }
  V = gfs[V_gfn - 1];
  rho = gfs[rho_gfn - 1];
  a = op == 1 ? gfs[a_gfn - 1] : NULL;
  c = op == 1 ? gfs[c_gfn - 1] : NULL;
  V_res = gfs[V_res_gfn - 1];
  V_rhs = gfs[V_rhs_gfn - 1];
  V_lop = gfs[V_lop_gfn - 1];
  cmask = gfs[cmask_gfn - 1];
  chr = do_ex ? gfs[chr_gfn - 1] : NULL;
  dx = x[0][1] - x[0][0];
  dy = x[1][1] - x[1][0];
  dz = x[2][1] - x[2][0];
  return 1;
}

static double read_norm(int n)
{
  int i;
  double r = 0;
  if (hipMemcpyAsync(h_norm, d_norm, sizeof(double) * n, hipMemcpyDeviceToHost, stream()) != hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess) {
    printf("nps_hooks: reading the norm failed\n");
    return -1.0;
  }
  for (i = 0; i < n; i++) r = (i == 0 || h_norm[i] > r) ? h_norm[i] : r;
  return r;
}

static int relax_grid(double *norm)
{
  if (!ldptr()) return -1;
  return pmg_hook_relax(op, V, V_rhs, rho, a, c, cmask, chr, ex_val, shape[0], shape[1], shape[2], dx, dy, dz,
                        norm, stream());
}

static int residual_grid(double *norm)
{
  if (!ldptr()) return -1;
  return pmg_hook_residual(op, V_res, V_rhs, V, rho, a, c, cmask, chr, ex_val, shape[0], shape[1], shape[2], dx,
                           dy, dz, norm, stream());
}

/* ---- the per-grid hooks of amrd() (include/amrd.h:137-159) ---------------- */
double app_MG_relax(void)
{
  if (relax_grid(d_norm)) return -1.0;
  return read_norm(1);
}

double app_MG_residual(void)
{
  if (residual_grid(d_norm)) return -1.0;
  return read_norm(1);
}

void app_L_op(void)
{
  if (!ldptr()) return;
  if (pmg_hook_lop(op, V_lop, V, rho, a, c, cmask, chr, ex_val, shape[0], shape[1], shape[2], dx, dy, dz, stream()))
    printf("app_L_op: hook kernel failed\n");
}

/* ---- level hooks: every grid of a level, one norm read-back --------------- */
static double level_relax(int L, int comm_only)
{
  int valid = PAMR_init_s_iter(L, PAMR_MGH, 0), n = 0, comm;
  while (valid && n < NPS_MAX_GRIDS) {
    PAMR_get_g_comm(&comm);
    if (comm || !comm_only) {
      if (relax_grid(d_norm + n)) return -1.0;
      n++;
    }
    valid = PAMR_next_g();
  }
  return n ? read_norm(n) : 0.0;
}

static double level_residual(int L)
{
  int valid = PAMR_init_s_iter(L, PAMR_MGH, 0), n = 0;
  while (valid && n < NPS_MAX_GRIDS) {
    if (residual_grid(d_norm + n)) return -1.0;
    n++;
    valid = PAMR_next_g();
  }
  return n ? read_norm(n) : 0.0;
}

static void level_lop(int L)
{
  int valid = PAMR_init_s_iter(L, PAMR_MGH, 0);
  while (valid) {
    app_L_op();
    valid = PAMR_next_g();
  }
}

/* ---- set-up --------------------------------------------------------------- */

/* defines V, rho (and a, c for op 1) with the attributes AMRD gives an
 * elliptic variable and its sources (amrd/io.c:700-790): V synchronised
 * and restricted with rop_u, the sources restricted with rop_src into
 * MG-only levels; then the MG variables via amrd_mg_def_vars.  Call after
 * PAMR_init_context, before PAMR_compose_hierarchy.  Returns 1 / 0. */
int nps_define_vars_tl(int op_, int rop_u, int rop_res, int rop_src, int num_tl);

int nps_define_vars(int op_, int rop_u, int rop_res, int rop_src)
{
  return nps_define_vars_tl(op_, rop_u, rop_res, rop_src, 1);
}

/* the same with num_tl AMR time levels of V (2 or more for the MG
 * extrapolation of a constrained evolution, amrd/evolve.c:346-368); the MG
 * hierarchy then works on the time level given to PAMR_build_mgh */
int nps_define_vars_tl(int op_, int rop_u, int rop_res, int rop_src, int num_tl)
{
  int pbt[6] = {0, 0, 0, 0, 0, 0};
  op = op_;
  if (!PAMR_def_var_full("V", 1, 1, num_tl, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC, rop_u,
                         PAMR_NO_INTERP, PAMR_SYNC, 0, PAMR_NO_TRANSFER, pbt))
    return 0;
  if (!PAMR_def_var_full("rho", 1, 1, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC, rop_src,
                         PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_NO_TRANSFER, pbt))
    return 0;
  if (op == 1) {
    if (!PAMR_def_var_full("a", 1, 1, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC, rop_src,
                           PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_NO_TRANSFER, pbt))
      return 0;
    if (!PAMR_def_var_full("c", 1, 1, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC, rop_src,
                           PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_NO_TRANSFER, pbt))
      return 0;
  }
  return amrd_mg_def_vars("V", rop_res == PAMR_FW_RESTR ? "V_res" : "", "V", do_ex);
}

/* the same with excision: AMRD's chr variable is defined too (amrd/io.c:
 * 988-1005) and the hooks skip the points where chr == ex; the application
 * then turns PAMR's excision on with PAMR_excision_on("chr", fill, ex, ...) */
int nps_define_vars_ex(int op_, int rop_u, int rop_res, int rop_src, double ex)
{
  do_ex = 1;
  ex_val = ex;
  return nps_define_vars_tl(op_, rop_u, rop_res, rop_src, 1);
}

/* looks up the gfns, allocates the norm buffers and registers the per-grid
 * hooks; call after PAMR_compose_hierarchy.  Returns 1 / 0. */
int nps_init(void)
{
  if ((V_gfn = PAMR_get_gfn("V", PAMR_MGH, 0)) <= 0) return 0;
  if ((rho_gfn = PAMR_get_gfn("rho", PAMR_MGH, 0)) <= 0) return 0;
  if (op == 1) {
    if ((a_gfn = PAMR_get_gfn("a", PAMR_MGH, 0)) <= 0) return 0;
    if ((c_gfn = PAMR_get_gfn("c", PAMR_MGH, 0)) <= 0) return 0;
  }
  if ((V_res_gfn = PAMR_get_gfn("V_res", PAMR_MGH, 0)) <= 0) return 0;
  if ((V_rhs_gfn = PAMR_get_gfn("V_rhs", PAMR_MGH, 0)) <= 0) return 0;
  if ((V_lop_gfn = PAMR_get_gfn("V_lop", PAMR_MGH, 0)) <= 0) return 0;
  if ((cmask_gfn = PAMR_get_gfn("cmask", PAMR_MGH, 0)) <= 0) return 0;
  if (do_ex && (chr_gfn = PAMR_get_gfn("chr", PAMR_MGH, 0)) <= 0) return 0;
  if (!d_norm && hipMalloc((void **)&d_norm, sizeof(double) * NPS_MAX_GRIDS) != hipSuccess) return 0;
  if (!h_norm && hipHostMalloc((void **)&h_norm, sizeof(double) * NPS_MAX_GRIDS, hipHostMallocDefault) != hipSuccess)
    return 0;
  amrd_set_app_MG_hooks(app_MG_residual, app_MG_relax, app_L_op);
  amrd_set_app_MG_level_hooks(NULL, NULL, NULL);
  return 1;
}

void nps_set_level_hooks(int on)
{
  if (on) amrd_set_app_MG_level_hooks(level_relax, level_residual, level_lop);
  else amrd_set_app_MG_level_hooks(NULL, NULL, NULL);
}

void nps_release(void)
{
  if (d_norm) hipFree(d_norm);
  if (h_norm) hipHostFree(h_norm);
  d_norm = h_norm = NULL;
  do_ex = 0;
  ex_val = 0.0;
}

/* host <-> device copies of n doubles on amrd_gpu_stream(), synchronous */
int nps_upload(double *dev, const double *host, long n)
{
  if (hipMemcpyAsync(dev, host, sizeof(double) * n, hipMemcpyHostToDevice, stream()) != hipSuccess) return 0;
  return hipStreamSynchronize(stream()) == hipSuccess;
}

int nps_download(double *host, const double *dev, long n)
{
  if (hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, stream()) != hipSuccess) return 0;
  return hipStreamSynchronize(stream()) == hipSuccess;
}
