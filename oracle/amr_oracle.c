/*
 * amr_oracle.c -- CPU oracle for the multigrid solve over an AMR hierarchy.
 *
 * TEST INFRASTRUCTURE ONLY (see amr_oracle.h).  Every function cites the
 * reference code it restates.  Grid segments carry physical bounding boxes
 * as in src/transfer.c, and all comparisons use the reference's fuzz rules
 * (src/misc.c:14-37), so the transfer lists -- their pieces and their
 * order -- are the reference's.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "amr_oracle.h"
#include "mg_oracle.h"

#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define INTERP_BUFFER 2 /* context default, src/pamr.c:157 */
#define SMALL_DX 1e-20  /* PAMR_SMALL_DX, include/internal_opts.h:19 */

/* src/misc.c:14-37 */
static int fz_eq(double a, double b, double f) { return fabs(b - a) < f; }
static int fz_lt(double a, double b, double f) { return a < -f + b; }
static int fz_gt(double a, double b, double f) { return a > f + b; }

typedef struct {
  int shape[3];
  double bbox[6];
  double hdx[3]; /* x(2)-x(1): the hooks' dx */
  double gdx[3]; /* x[2]-x[1]: mg_ldptr's AMRD_g_dx (amrd/mg.c:220) */
  int coarsest, comm;
  double *f[OA_NUM];
} ag_grid;

typedef struct {
  int ng;
  ag_grid *g;
  int in_amrh;
  int shared; /* grid functions belong to another hierarchy (MG over AMR levels) */
} ag_level;

struct or_amr {
  int nlev, nbase, op;
  ag_level lev[OR_MAXLEV];
  double cbbox[6]; /* context bbox = AMRD_base_bbox */
  int periodic[3];
  int pre, pst, max_citer, rop_u, rop_res, rop_src, iord;
  double crtol, w0_r, w0_i;
  int save_tre, ntre;
  int do_ex;
  double ex;
  int tf[OA_NUM]; /* transfer bits per field for the current communicate() */
  /* the boxes given for each AMR level l (index l-1) in the caller's order
   * (save_sgh_bboxes, src/gh.c:273-298); level 1 holds the base box */
  int nsgh[OR_MAXLEV];
  double *sgh[OR_MAXLEV];
  int namr_lev; /* AMR levels (base included) */
  double dx0[3];
  int mcw;
};

/* ------------------------------------------------------------------------ */
/* grid-segment lists                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int g;
  double bb[6];
} seg;
typedef struct {
  seg *s;
  int n, cap;
} segl;

static void sl_push(segl *l, int g, const double *bb)
{
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 16;
    l->s = (seg *)realloc(l->s, sizeof(seg) * l->cap);
  }
  l->s[l->n].g = g;
  memcpy(l->s[l->n].bb, bb, 6 * sizeof(double));
  l->n++;
}
static void sl_cat(segl *a, const segl *b)
{
  int i;
  for (i = 0; i < b->n; i++) sl_push(a, b->s[i].g, b->s[i].bb);
}
static void sl_free(segl *l)
{
  free(l->s);
  l->s = 0;
  l->n = l->cap = 0;
}

static double gdxd(const ag_grid *g, int d)
{
  return g->shape[d] > 1 ? (g->bbox[2 * d + 1] - g->bbox[2 * d]) / (g->shape[d] - 1) : SMALL_DX;
}

/* alloc_gsl with ghost_width 0 (src/transfer.c:132-188): the whole grid,
 * reduced by AMR_reduce points along AMR (non-physical) boundaries.
 * Returns 0 when the segment is empty (the reference frees it). */
static int seg_alloc(const or_amr *m, const ag_grid *g, int gi, int interior, int AMR_reduce, seg *out)
{
  int d, ib[6];
  double dxt[3];
  out->g = gi;
  for (d = 0; d < 3; d++) {
    double dx = gdxd(g, d);
    out->bb[2 * d] = g->bbox[2 * d];
    out->bb[2 * d + 1] = g->bbox[2 * d + 1];
    if (AMR_reduce > 0) {
      if (fz_gt(out->bb[2 * d], m->cbbox[2 * d], dx / 2) || m->periodic[d]) out->bb[2 * d] += AMR_reduce * dx;
      if (fz_lt(out->bb[2 * d + 1], m->cbbox[2 * d + 1], dx / 2) || m->periodic[d])
        out->bb[2 * d + 1] += -AMR_reduce * dx;
    }
    if (interior || AMR_reduce) {
      or_calc_gs_ibbox(out->bb, g->bbox, g->shape, 3, ib, dxt);
      if ((ib[2 * d + 1] - ib[2 * d] + 1) < 1) return 0;
    }
  }
  return 1;
}

/* gs_subtract (src/transfer.c:218-283): C = A - B.  The reference prepends
 * each new piece to its result list; `out` receives the pieces in that
 * final list order. */
static void gs_subtract(const ag_level *lv, const seg *A, const seg *B, segl *out)
{
  double cut[6], dx[3];
  seg made[6];
  int nm = 0, i, j;
  const ag_grid *g = &lv->g[A->g];
  for (i = 0; i < 3; i++) {
    cut[2 * i] = A->bb[2 * i];
    cut[2 * i + 1] = A->bb[2 * i + 1];
    dx[i] = gdxd(g, i);
    if (fz_gt(B->bb[2 * i], A->bb[2 * i + 1], dx[i] / 2) || fz_gt(A->bb[2 * i], B->bb[2 * i + 1], dx[i] / 2)) {
      sl_push(out, A->g, A->bb);
      return;
    }
  }
  for (i = 0; i < 3; i++) {
    cut[2 * i] = MAX(A->bb[2 * i], B->bb[2 * i]);
    if (fz_gt(cut[2 * i], A->bb[2 * i], dx[i] / 2)) {
      seg *C = &made[nm++];
      C->g = A->g;
      for (j = 0; j < 3; j++) {
        if (i == j) {
          C->bb[2 * i] = A->bb[2 * i];
          C->bb[2 * i + 1] = MAX(C->bb[2 * i], cut[2 * i] - dx[i]);
        } else {
          C->bb[2 * j] = cut[2 * j];
          C->bb[2 * j + 1] = cut[2 * j + 1];
        }
      }
    }
    cut[2 * i + 1] = MIN(A->bb[2 * i + 1], B->bb[2 * i + 1]);
    if (fz_lt(cut[2 * i + 1], A->bb[2 * i + 1], dx[i] / 2)) {
      seg *C = &made[nm++];
      C->g = A->g;
      for (j = 0; j < 3; j++) {
        if (i == j) {
          C->bb[2 * i + 1] = A->bb[2 * i + 1];
          C->bb[2 * i] = MIN(C->bb[2 * i + 1], cut[2 * i + 1] + dx[i]);
        } else {
          C->bb[2 * j] = cut[2 * j];
          C->bb[2 * j + 1] = cut[2 * j + 1];
        }
      }
    }
  }
  for (i = nm - 1; i >= 0; i--) sl_push(out, made[i].g, made[i].bb);
}

/* gsl_subtract (src/transfer.c:285-314) */
static void gsl_subtract(const ag_level *lv, const segl *A, const segl *B, segl *out)
{
  segl C = {0, 0, 0}, C0;
  int b, c;
  sl_cat(&C, A);
  for (b = 0; b < B->n; b++) {
    C0.s = 0;
    C0.n = C0.cap = 0;
    for (c = 0; c < C.n; c++) gs_subtract(lv, &C.s[c], &B->s[b], &C0);
    sl_free(&C);
    C = C0;
  }
  sl_cat(out, &C);
  sl_free(&C);
}

/* build_owned_gsl (src/transfer.c:415-458), one rank */
static void build_owned(const or_amr *m, int L, int interior, int extend, segl *owned)
{
  const ag_level *lv = &m->lev[L - 1];
  int cint = interior ? 1 : 0, cred = interior > 1 ? interior - 1 : 0, gi, s, d;
  segl ints = {0, 0, 0};
  for (gi = 0; gi < lv->ng; gi++) {
    seg one;
    segl A = {0, 0, 0};
    if (!lv->g[gi].comm) continue;
    if (!seg_alloc(m, &lv->g[gi], gi, cint, cred, &one)) continue;
    sl_push(&A, one.g, one.bb);
    gsl_subtract(lv, &A, &ints, owned);
    sl_push(&ints, one.g, one.bb);
    sl_free(&A);
  }
  sl_free(&ints);
  if (extend) /* extend_gsl (src/transfer.c:364-386) */
    for (s = 0; s < owned->n; s++) {
      const ag_grid *g = &lv->g[owned->s[s].g];
      for (d = 0; d < 3; d++)
        if (g->shape[d] > 1) {
          double dx = gdxd(g, d);
          owned->s[s].bb[2 * d] = MAX(owned->s[s].bb[2 * d] - dx, g->bbox[2 * d]);
          owned->s[s].bb[2 * d + 1] = MIN(owned->s[s].bb[2 * d + 1] + dx, g->bbox[2 * d + 1]);
        }
    }
}

/* build_complete_gsl with AMR_bdy_only = 0 (src/transfer.c:465-504) */
static void build_complete(const or_amr *m, int L, segl *out)
{
  const ag_level *lv = &m->lev[L - 1];
  int gi;
  for (gi = 0; gi < lv->ng; gi++) {
    seg one;
    if (!lv->g[gi].comm) continue;
    if (seg_alloc(m, &lv->g[gi], gi, 0, 0, &one)) sl_push(out, one.g, one.bb);
  }
}

typedef struct {
  seg src, dst;
} tpair;

/* build_gstl (src/transfer.c:514-540) with gs_and (src/transfer.c:325-356) */
static int build_gstl(const or_amr *m, int Ls, int Ld, const segl *src, const segl *dst, tpair **out)
{
  int n = 0, cap = 0, a, b, i;
  tpair *p = 0;
  for (a = 0; a < src->n; a++)
    for (b = 0; b < dst->n; b++) {
      const ag_grid *gA = &m->lev[Ls - 1].g[src->s[a].g], *gB = &m->lev[Ld - 1].g[dst->s[b].g];
      double nb[6];
      int ok = 1;
      if (Ls == Ld && src->s[a].g == dst->s[b].g) continue;
      for (i = 0; i < 3; i++) {
        double dxm = MAX(gdxd(gA, i), gdxd(gB, i));
        nb[2 * i] = MAX(src->s[a].bb[2 * i], dst->s[b].bb[2 * i]);
        nb[2 * i + 1] = MIN(src->s[a].bb[2 * i + 1], dst->s[b].bb[2 * i + 1]);
        if (fz_lt(nb[2 * i + 1], nb[2 * i], dxm / 2)) ok = 0;
      }
      if (!ok) continue;
      if (n == cap) {
        cap = cap ? 2 * cap : 32;
        p = (tpair *)realloc(p, sizeof(tpair) * cap);
      }
      p[n].src.g = src->s[a].g;
      p[n].dst.g = dst->s[b].g;
      memcpy(p[n].src.bb, nb, sizeof nb);
      memcpy(p[n].dst.bb, nb, sizeof nb);
      n++;
    }
  *out = p;
  return n;
}

/* calc_delta (src/transfer.c:545-575) */
static int calc_delta(const ag_grid *gs, const ag_grid *gd, int *irho)
{
  double dxsm = 0, dxdm = 0;
  int i;
  *irho = 0;
  for (i = 0; i < 3; i++)
    if (gs->shape[i] > 1 && gd->shape[i] > 1) {
      dxsm = MAX(dxsm, (gs->bbox[2 * i + 1] - gs->bbox[2 * i]) / (gs->shape[i] - 1));
      dxdm = MAX(dxdm, (gd->bbox[2 * i + 1] - gd->bbox[2 * i]) / (gd->shape[i] - 1));
    }
  if (dxsm == 0 || dxdm == 0) return 0;
  if (fz_eq(dxsm, dxdm, dxsm / 100)) return 0;
  if (fz_gt(dxsm, dxdm, dxsm / 100)) {
    *irho = (int)(dxsm / dxdm + 0.5);
    return -1;
  }
  *irho = (int)(dxdm / dxsm + 0.5);
  return 1;
}

/* align_src_bbox (src/transfer.c:101-121) */
static void align_src(const ag_grid *gs, const ag_grid *gd, seg *src, seg *dst)
{
  int i;
  for (i = 0; i < 3; i++)
    if (gs->shape[i] > 1 && gd->shape[i] > 1) {
      double dxs = (gs->bbox[2 * i + 1] - gs->bbox[2 * i]) / (gs->shape[i] - 1);
      double dxd = (gd->bbox[2 * i + 1] - gd->bbox[2 * i]) / (gd->shape[i] - 1);
      src->bb[2 * i] = ((int)((src->bb[2 * i] - gd->bbox[2 * i] + dxs / 2) / dxd)) * dxd + gd->bbox[2 * i];
      if (fz_lt(src->bb[2 * i], gs->bbox[2 * i], dxs / 2)) src->bb[2 * i] += dxd;
      dst->bb[2 * i] = src->bb[2 * i];
      src->bb[2 * i + 1] =
          ((int)((src->bb[2 * i + 1] - gd->bbox[2 * i] + dxd - dxs / 2) / dxd)) * dxd + gd->bbox[2 * i];
      if (fz_gt(src->bb[2 * i + 1], gs->bbox[2 * i + 1], dxs / 2)) src->bb[2 * i + 1] -= dxd;
      dst->bb[2 * i + 1] = src->bb[2 * i + 1];
    }
}

/* local_transfer (src/transfer.c:946-1163) for the fields whose transfer
 * bit is set; no excision. */
static void local_transfer(or_amr *m, int Ls, int Ld, tpair *pr, int np)
{
  int q, w;
  for (q = 0; q < np; q++) {
    ag_grid *g1 = &m->lev[Ls - 1].g[pr[q].src.g], *g2 = &m->lev[Ld - 1].g[pr[q].dst.g];
    int irho, ib[6], ibd[6], delta = calc_delta(g1, g2, &irho);
    double dx[3], dxd[3];
    int nx1 = g1->shape[0], ny1 = g1->shape[1], nz1 = g1->shape[2];
    int nx2 = g2->shape[0], ny2 = g2->shape[1], nz2 = g2->shape[2];
    /* excision masks of both grids (src/transfer.c:981-991) */
    const double *chr1 = m->do_ex ? g1->f[OA_CHR] : 0, *chr2 = m->do_ex ? g2->f[OA_CHR] : 0;
    if (delta == 1) align_src(g1, g2, &pr[q].src, &pr[q].dst);
    or_calc_gs_ibbox(pr[q].src.bb, g1->bbox, g1->shape, 3, ib, dx);
    or_calc_gs_ibbox(pr[q].dst.bb, g2->bbox, g2->shape, 3, ibd, dxd);
    for (w = 0; w < OA_NUM; w++) {
      int tb = m->tf[w];
      int ni = ib[1] - ib[0] + 1, nj = ib[3] - ib[2] + 1, nk = ib[5] - ib[4] + 1;
      int nib = ibd[1] - ibd[0] + 1, njb = ibd[3] - ibd[2] + 1, nkb = ibd[5] - ibd[4] + 1;
      if (!tb || !g1->f[w] || !g2->f[w]) continue;
      if (delta == 0) {
        or_dmcopy3d(g1->f[w], g2->f[w], nx1, ny1, nz1, nx2, ny2, nz2, ib[0], ib[2], ib[4], ibd[0], ibd[2], ibd[4],
                    ni, nj, nk, 1, 1, 1, 1, 1, 1);
      } else if (delta == -1) {
        /* interpolate a buffered coarse region into a temporary, then
         * copy the wanted fine box (src/transfer.c:1007-1051,1089-1115) */
        int b2[3], b2i[3], b1i[3], nbi[3], nb1b[3], nbd[3] = {nib, njb, nkb}, n1[3] = {nx1, ny1, nz1}, d;
        double *data;
        for (d = 0; d < 3; d++) {
          b2[d] = (int)((pr[q].src.bb[2 * d] - g1->bbox[2 * d]) / dxd[d] + 1.5);
          b2i[d] = MAX(1, (((b2[d] - 1) / irho) - INTERP_BUFFER) * irho + 1);
          b1i[d] = (b2i[d] - 1) / irho + 1;
          nbi[d] = MIN(n1[d], (b2[d] + nbd[d] - 1) / irho + 1 + INTERP_BUFFER + 1) - b1i[d] + 1;
          nb1b[d] = (nbi[d] - 1) * irho + 1;
          b2[d] = b2[d] - b2i[d] + 1;
        }
        data = (double *)malloc(sizeof(double) * (size_t)nb1b[0] * nb1b[1] * nb1b[2]);
        or_dminterp3d(g1->f[w], data, nx1, ny1, nz1, nb1b[0], nb1b[1], nb1b[2], b1i[0], b1i[1], b1i[2], 1, 1, 1,
                      nbi[0], nbi[1], nbi[2], 1, 1, 1, irho, irho, irho, tb, m->do_ex ? chr1 : g1->f[w], m->ex,
                      m->do_ex);
        or_dmcopy3d(data, g2->f[w], nb1b[0], nb1b[1], nb1b[2], nx2, ny2, nz2, b2[0], b2[1], b2[2], ibd[0], ibd[2],
                    ibd[4], nib, njb, nkb, 1, 1, 1, 1, 1, 1);
        free(data);
        if (m->do_ex) {
          /* clear_ex of both whole grids (src/transfer.c:930-935,1110-1114) */
          size_t j, n1 = (size_t)nx1 * ny1 * nz1, n2 = (size_t)nx2 * ny2 * nz2;
          for (j = 0; j < n1; j++) if (chr1[j] == m->ex) g1->f[w][j] = 0;
          for (j = 0; j < n2; j++) if (chr2[j] == m->ex) g2->f[w][j] = 0;
        }
      } else {
        if (tb == 1)
          or_dmcopy3d(g1->f[w], g2->f[w], nx1, ny1, nz1, nx2, ny2, nz2, ib[0], ib[2], ib[4], ibd[0], ibd[2], ibd[4],
                      nib, njb, nkb, irho, irho, irho, 1, 1, 1);
        else if (tb == 2)
          or_dmhwr3d(g1->f[w], g2->f[w], nx1, ny1, nz1, nx2, ny2, nz2, ib[0], ib[2], ib[4], ibd[0], ibd[2], ibd[4],
                     nib, njb, nkb, irho, irho, irho, 1, 1, 1, m->do_ex ? chr1 : g1->f[w], m->ex, m->do_ex);
        else
          or_dmfwr3d(g1->f[w], g2->f[w], nx1, ny1, nz1, nx2, ny2, nz2, ib[0], ib[2], ib[4], ibd[0], ibd[2], ibd[4],
                     nib, njb, nkb, irho, irho, irho, 1, 1, 1, m->do_ex ? chr1 : g1->f[w], m->ex, m->do_ex);
      }
    }
  }
}

/* communicate (src/pamr.c:533-593), one rank, AMR_bdy_only = 0 */
static int make_transfers(or_amr *m, int l1, int l2, int interior, int extend, tpair **pr)
{
  segl src = {0, 0, 0}, dst = {0, 0, 0};
  int np;
  build_owned(m, l1, interior, extend, &src);
  build_complete(m, l2, &dst);
  np = build_gstl(m, l1, l2, &src, &dst, pr);
  sl_free(&src);
  sl_free(&dst);
  return np;
}
static void communicate(or_amr *m, int l1, int l2, int interior, int extend)
{
  tpair *pr = 0;
  int np = make_transfers(m, l1, l2, interior, extend, &pr);
  local_transfer(m, l1, l2, pr, np);
  free(pr);
}

static void tf_clear(or_amr *m) { memset(m->tf, 0, sizeof m->tf); }

/* PAMR_sync (src/pamr.c:601-617) of the fields selected by the caller */
static void do_sync(or_amr *m, int L, int bdy) { communicate(m, L, L, 1 + bdy, 0); }

/* PAMR_inject (src/pamr.c:667-685) with the tf bits of src/pamr.c:1886-1897:
 * u and u_res always (mg_noinj_to_amr = 0), the sources only into levels
 * that are not in the AMR hierarchy (their mg_noinj_to_amr = 1). */
static void do_inject(or_amr *m, int Lf)
{
  tf_clear(m);
  m->tf[OA_U] = m->rop_u;
  m->tf[OA_RES] = m->rop_res;
  if (!m->lev[Lf - 2].in_amrh) {
    m->tf[OA_RHO] = m->rop_src;
    if (m->op == 1) m->tf[OA_A] = m->tf[OA_C] = m->rop_src;
  }
  communicate(m, Lf, Lf - 1, 1, 0);
}

/* ------------------------------------------------------------------------ */
/* hierarchy construction                                                    */
/* ------------------------------------------------------------------------ */

/* insert_grid (src/gh.c:235-257): position of a new bbox in a level list */
static void insert_sorted(double (*list)[6], int *idx, int *n, const double *bb, int id)
{
  int p = 0, k;
  while (p < *n && list[p][0] < bb[0] && list[p][2] < bb[2] && list[p][4] < bb[4] && list[p][1] < bb[1] &&
         list[p][3] < bb[3] && list[p][5] < bb[5])
    p++;
  for (k = *n; k > p; k--) {
    memcpy(list[k], list[k - 1], 6 * sizeof(double));
    idx[k] = idx[k - 1];
  }
  memcpy(list[p], bb, 6 * sizeof(double));
  idx[p] = id;
  (*n)++;
}

static void grid_setup(or_amr *m, ag_grid *g, const double *bb, const double *ldx)
{
  int d, w;
  size_t n;
  memcpy(g->bbox, bb, 6 * sizeof(double));
  for (d = 0; d < 3; d++) {
    double dx, x0, x1, x2;
    g->shape[d] = (int)((bb[2 * d + 1] - bb[2 * d]) / ldx[d] + 1.5);
    /* coordinates by repeated addition (src/gh.c:905-906) */
    dx = (bb[2 * d + 1] - bb[2 * d]) / (g->shape[d] - 1);
    x0 = bb[2 * d];
    x1 = x0 + dx;
    x2 = x1 + dx;
    g->hdx[d] = x1 - x0;
    g->gdx[d] = x2 - x1;
  }
  g->comm = 1;
  n = (size_t)g->shape[0] * g->shape[1] * g->shape[2];
  for (w = 0; w < OA_NUM; w++) {
    if ((w == OA_A || w == OA_C) && m->op == 0) continue;
    g->f[w] = (double *)calloc(n, sizeof(double));
  }
}

/* a level's grids from a bbox list: build_sgh then compose_cgh each apply
 * insert_grid (src/gh.c:331-378, 602-928) */
static void level_from_boxes_perm(or_amr *m, ag_level *lv, int nb, const double *boxes, const double *ldx,
                                 int *perm);
static void level_from_boxes(or_amr *m, ag_level *lv, int nb, const double *boxes, const double *ldx)
{
  level_from_boxes_perm(m, lv, nb, boxes, ldx, 0);
}
/* perm (may be NULL): perm[k] = index in `boxes` of grid k */
static void level_from_boxes_perm(or_amr *m, ag_level *lv, int nb, const double *boxes, const double *ldx,
                                 int *perm)
{
  double(*l1)[6] = malloc(sizeof(double[6]) * (nb + 1)), (*l2)[6] = malloc(sizeof(double[6]) * (nb + 1));
  int *i1 = malloc(sizeof(int) * (nb + 1)), *i2 = malloc(sizeof(int) * (nb + 1));
  int n1 = 0, n2 = 0, k;
  for (k = 0; k < nb; k++) insert_sorted(l1, i1, &n1, boxes + 6 * k, k);
  for (k = 0; k < n1; k++) insert_sorted(l2, i2, &n2, l1[k], i1[k]);
  lv->ng = nb;
  lv->g = (ag_grid *)calloc(nb > 0 ? (size_t)nb : 1, sizeof(ag_grid));
  for (k = 0; k < nb; k++) grid_setup(m, &lv->g[k], l2[k], ldx);
  if (perm)
    for (k = 0; k < nb; k++) perm[k] = i2[k];
  free(l1);
  free(l2);
  free(i1);
  free(i2);
}

or_amr *or_amr_create(const int *base_shape, const double *bbox, int namr, const int *nbox, const double *boxes,
                      int mcw, int op)
{
  or_amr *m = (or_amr *)calloc(1, sizeof(or_amr));
  double dx0[3], island, ldx[3];
  int mcwv[3] = {mcw, mcw, mcw}, ncoarse = 0, cf, d, L, l, off = 0;
  memcpy(m->cbbox, bbox, 6 * sizeof(double));
  m->op = op;
  for (d = 0; d < 3; d++) dx0[d] = (bbox[2 * d + 1] - bbox[2 * d]) / (base_shape[d] - 1);
  /* levels below the base: coarsen by 2 while the island number of the base
   * box is nonzero (build_mgh, src/gh.c:1275-1340) */
  for (cf = 2;; cf *= 2) {
    or_calc_island_no(bbox, 1, &island, dx0, cf, 3, mcwv);
    if (island == 0 || ncoarse + 1 + namr >= OR_MAXLEV) break;
    ncoarse++;
  }
  m->nbase = ncoarse + 1;
  m->nlev = ncoarse + 1 + namr;
  m->namr_lev = namr + 1;
  m->mcw = mcw;
  memcpy(m->dx0, dx0, sizeof dx0);
  m->nsgh[0] = 1;
  m->sgh[0] = (double *)malloc(6 * sizeof(double));
  memcpy(m->sgh[0], bbox, 6 * sizeof(double));
  for (l = 0, off = 0; l < namr; l++) {
    m->nsgh[l + 1] = nbox[l];
    m->sgh[l + 1] = (double *)malloc(6 * sizeof(double) * (nbox[l] > 0 ? nbox[l] : 1));
    memcpy(m->sgh[l + 1], boxes + 6 * off, 6 * sizeof(double) * nbox[l]);
    off += nbox[l];
  }
  off = 0;
  for (L = 1; L <= ncoarse; L++) {
    int c = 1 << (ncoarse + 1 - L);
    for (d = 0; d < 3; d++) ldx[d] = dx0[d] * c;
    level_from_boxes(m, &m->lev[L - 1], 1, bbox, ldx);
    m->lev[L - 1].in_amrh = 0;
  }
  level_from_boxes(m, &m->lev[m->nbase - 1], 1, bbox, dx0);
  m->lev[m->nbase - 1].in_amrh = 1;
  for (l = 0; l < namr; l++) {
    for (d = 0; d < 3; d++) ldx[d] = dx0[d] / (double)(2 << l);
    level_from_boxes(m, &m->lev[m->nbase + l], nbox[l], boxes + 6 * off, ldx);
    m->lev[m->nbase + l].in_amrh = 1;
    off += nbox[l];
  }
  /* the coarsest level's grids are flagged coarsest (src/gh.c:1353-1354) */
  for (d = 0; d < m->lev[0].ng; d++) m->lev[0].g[d].coarsest = 1;
  m->pre = m->pst = 2;
  m->max_citer = 5000;
  m->crtol = 1e-3;
  m->w0_r = m->w0_i = 1.0;
  m->rop_u = m->rop_res = m->rop_src = 3;
  m->iord = 2;
  return m;
}

void or_amr_destroy(or_amr *m)
{
  int L, g, w;
  if (!m) return;
  for (L = 0; L < m->nlev; L++) {
    if (!m->lev[L].shared)
      for (g = 0; g < m->lev[L].ng; g++)
        for (w = 0; w < OA_NUM; w++) free(m->lev[L].g[g].f[w]);
    free(m->lev[L].g);
  }
  for (L = 0; L < OR_MAXLEV; L++) free(m->sgh[L]);
  free(m);
}

/* PAMR_build_mgh(min_lev, max_lev, tl) for min_lev = amr_lev (src/gh.c:
 * 1222-1365): AMR levels amr_lev .. finest replicated -- their grid
 * functions are m's; the MG-only ones are zeroed, as amr_to_mgh allocates
 * them fresh (src/gh.c:1150-1157) -- and below them, for cf = 2, 4, ..., a
 * level of the boxes given for amr_lev whose calc_island_no(cf) is nonzero,
 * spacing dx * cf, coarsest where the island number is negative, until no box
 * survives (the n < 0 branch, src/gh.c:1281-1340); the grids of level 1 are
 * all coarsest (:1352-1354).  The excision mask of a new level is left to
 * the caller, as init_ex_mask calls the application (:1322). */
or_amr *or_amr_create_mgh(or_amr *m, int amr_lev)
{
  or_amr *q;
  int L, l, k, cf, j, nc = 0, nsh, mcwv[3];
  int num = m->nsgh[amr_lev - 1];
  const double *bb = m->sgh[amr_lev - 1];
  double dx[3], *isl, *keep;
  int *kisl, *perm;
  ag_level coarse[OR_MAXLEV];
  if (amr_lev < 1 || amr_lev > m->namr_lev) return 0;
  q = (or_amr *)calloc(1, sizeof(or_amr));
  memcpy(q->cbbox, m->cbbox, sizeof q->cbbox);
  memcpy(q->periodic, m->periodic, sizeof q->periodic);
  q->op = m->op; q->pre = m->pre; q->pst = m->pst; q->max_citer = m->max_citer;
  q->crtol = m->crtol; q->w0_r = m->w0_r; q->w0_i = m->w0_i;
  q->rop_u = m->rop_u; q->rop_res = m->rop_res; q->rop_src = m->rop_src; q->iord = m->iord;
  q->do_ex = m->do_ex; q->ex = m->ex; q->mcw = m->mcw;
  for (j = 0; j < 3; j++) {
    dx[j] = m->dx0[j] / (double)(1 << (amr_lev - 1));
    mcwv[j] = m->mcw;
  }
  nsh = m->nlev - (m->nbase + amr_lev - 1) + 1;
  isl = (double *)malloc(sizeof(double) * num);
  keep = (double *)malloc(sizeof(double) * 6 * num);
  kisl = (int *)malloc(sizeof(int) * num);
  perm = (int *)malloc(sizeof(int) * num);
  for (cf = 2; nc + nsh < OR_MAXLEV; cf *= 2) {
    int n = 0;
    double ldx[3];
    ag_level *lv = &coarse[nc];
    or_calc_island_no(bb, num, isl, dx, cf, 3, mcwv);
    for (k = 0; k < num; k++)
      if (isl[k] != 0) {
        memcpy(keep + 6 * n, bb + 6 * k, 6 * sizeof(double));
        kisl[n++] = isl[k] < 0;
      }
    if (!n) break;
    for (j = 0; j < 3; j++) ldx[j] = dx[j] * cf;
    memset(lv, 0, sizeof *lv);
    level_from_boxes_perm(q, lv, n, keep, ldx, perm);
    for (k = 0; k < n; k++) lv->g[k].coarsest = kisl[perm[k]];
    nc++;
  }
  free(isl);
  free(keep);
  free(kisl);
  free(perm);
  q->nbase = nc + 1;
  q->nlev = nc + nsh;
  for (L = 1; L <= nc; L++) q->lev[L - 1] = coarse[nc - L];
  for (l = 0; l < nsh; l++) {
    const ag_level *src = &m->lev[m->nbase + amr_lev - 2 + l];
    ag_level *dst = &q->lev[nc + l];
    int g, w;
    dst->ng = src->ng;
    dst->in_amrh = 1;
    dst->shared = 1;
    dst->g = (ag_grid *)malloc(sizeof(ag_grid) * (src->ng > 0 ? src->ng : 1));
    memcpy(dst->g, src->g, sizeof(ag_grid) * src->ng);
    for (g = 0; g < dst->ng; g++) {
      dst->g[g].coarsest = 0;
      dst->g[g].comm = 1;
      for (w = 0; w < OA_NUM; w++) /* MG-only grid functions start at zero */
        if (dst->g[g].f[w] && (w == OA_RHS || w == OA_RES || w == OA_LOP || w == OA_CMASK))
          memset(dst->g[g].f[w], 0, sizeof(double) * (size_t)dst->g[g].shape[0] * dst->g[g].shape[1] *
                                         dst->g[g].shape[2]);
    }
  }
  for (k = 0; k < q->lev[0].ng; k++) q->lev[0].g[k].coarsest = 1;
  return q;
}

int or_amr_grid_coarsest(const or_amr *m, int L, int g) { return m->lev[L - 1].g[g].coarsest; }

int or_amr_nlev(const or_amr *m) { return m->nlev; }
int or_amr_base_level(const or_amr *m) { return m->nbase; }
int or_amr_ngrids(const or_amr *m, int L) { return (L < 1 || L > m->nlev) ? 0 : m->lev[L - 1].ng; }
void or_amr_grid_info(const or_amr *m, int L, int g, int *shape, double *bbox)
{
  const ag_grid *gr = &m->lev[L - 1].g[g];
  memcpy(shape, gr->shape, 3 * sizeof(int));
  memcpy(bbox, gr->bbox, 6 * sizeof(double));
}
double *or_amr_field(or_amr *m, int L, int g, int which)
{
  if (L < 1 || L > m->nlev || g < 0 || g >= m->lev[L - 1].ng || which < 0 || which >= OA_NUM) return 0;
  return m->lev[L - 1].g[g].f[which];
}
void or_amr_set_params(or_amr *m, int pre, int pst, int max_citer, double crtol, double w0_r, double w0_i,
                       int rop_u, int rop_res, int rop_src, int interp_ord)
{
  m->pre = pre; m->pst = pst; m->max_citer = max_citer; m->crtol = crtol;
  m->w0_r = w0_r; m->w0_i = w0_i; m->rop_u = rop_u; m->rop_res = rop_res;
  m->rop_src = rop_src; m->iord = interp_ord;
}

void or_amr_set_excision(or_amr *m, int do_ex, double ex)
{
  m->do_ex = do_ex;
  m->ex = ex;
}

void or_amr_set_tre(or_amr *m, int save_tre, int ntre)
{
  m->save_tre = save_tre;
  m->ntre = ntre;
}

/* TF_MGH_INIT (src/gh.c:1237-1239,1334; src/pamr.c:1898): the in_amrh
 * fields with mg_inject -- u and the sources -- into each new MG level */
void or_amr_init(or_amr *m)
{
  int L;
  for (L = m->nbase; L > 1; L--) {
    tf_clear(m);
    m->tf[OA_U] = m->rop_u;
    m->tf[OA_RHO] = m->rop_src;
    if (m->op == 1) m->tf[OA_A] = m->tf[OA_C] = m->rop_src;
    communicate(m, L, L - 1, 1, 0);
  }
}

/* ------------------------------------------------------------------------ */
/* V-cycle                                                                   */
/* ------------------------------------------------------------------------ */
static size_t gsize(const ag_grid *g) { return (size_t)g->shape[0] * g->shape[1] * g->shape[2]; }

/* amrd/mg.c:242-262 */
static void clear_mg_vars(or_amr *m, int L)
{
  int gi;
  size_t j;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    for (j = 0; j < gsize(g); j++) g->f[OA_RES][j] = g->f[OA_RHS][j] = g->f[OA_LOP][j] = 0;
  }
}

/* amrd/mg.c:362-488 for every grid of L */
void or_amr_set_cmask(or_amr *m, int L, int which)
{
  int gi, k;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    if (which == 0) {
      or_set_cmask_bdy(g->f[OA_CMASK], g->shape, g->bbox, m->cbbox, g->gdx, m->periodic);
    } else {
      int nc = L < m->nlev ? m->lev[L].ng : 0;
      double *cb = (double *)malloc(sizeof(double) * 6 * (nc + 1));
      for (k = 0; k < nc; k++) memcpy(cb + 6 * k, m->lev[L].g[k].bbox, 6 * sizeof(double));
      or_set_cmask_child(g->f[OA_CMASK], g->shape, g->bbox, m->cbbox, g->gdx, m->periodic, nc, cb);
      free(cb);
    }
  }
}

/* call_rr_app (amrd/util.c:73-103): hook on every (comm) grid, max-fold */
static double hook_relax_level(or_amr *m, int L, int comm_only)
{
  int gi, first = 1;
  double lres = 0, r;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    if (!(g->comm || !comm_only)) continue;
    r = or_relax(m->op, g->f[OA_U], g->f[OA_RHS], g->f[OA_RHO], g->f[OA_A], g->f[OA_C], g->f[OA_CMASK],
                 m->do_ex ? g->f[OA_CHR] : 0, m->ex,
                 g->shape[0], g->shape[1], g->shape[2], g->hdx[0], g->hdx[1], g->hdx[2]);
    if (first) lres = r;
    lres = MAX(r, lres);
    first = 0;
  }
  return lres;
}
static double hook_residual_level(or_amr *m, int L)
{
  int gi, first = 1;
  double lres = 0, r;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    r = or_residual(m->op, g->f[OA_RES], g->f[OA_RHS], g->f[OA_U], g->f[OA_RHO], g->f[OA_A], g->f[OA_C],
                    g->f[OA_CMASK], m->do_ex ? g->f[OA_CHR] : 0, m->ex, g->shape[0], g->shape[1], g->shape[2],
                    g->hdx[0], g->hdx[1], g->hdx[2]);
    if (first) lres = r;
    lres = MAX(r, lres);
    first = 0;
  }
  return lres;
}
static void hook_lop_level(or_amr *m, int L)
{
  int gi;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    or_lop(m->op, g->f[OA_LOP], g->f[OA_U], g->f[OA_RHO], g->f[OA_A], g->f[OA_C], g->f[OA_CMASK],
           m->do_ex ? g->f[OA_CHR] : 0, m->ex,
           g->shape[0], g->shape[1], g->shape[2], g->hdx[0], g->hdx[1], g->hdx[2]);
  }
}

/* amrd/mg.c:271-314; one elliptic variable, so the TRE's running max
 * starts from 0 at every point (i == 0, :290) */
static void compute_mg_rhs(or_amr *m, int L)
{
  int gi;
  size_t j;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    for (j = 0; j < gsize(g); j++) {
      if (m->save_tre) {
        double t = 0, v = fabs(g->f[OA_LOP][j]) * sqrt((double)m->ntre);
        g->f[OA_TRE][j] = t > v ? t : v; /* max (amrd/util.h:12) */
      }
      if (g->f[OA_CMASK][j] == 1.0)
        g->f[OA_RHS][j] = -g->f[OA_RES][j] * m->w0_r + g->f[OA_LOP][j];
      else
        g->f[OA_RHS][j] = 0;
      g->f[OA_LOP][j] = g->f[OA_U][j];
    }
  }
}

/* amrd/mg.c:323-355 */
static void compute_cgc(or_amr *m, int L)
{
  int gi;
  size_t j;
  for (gi = 0; gi < m->lev[L - 2].ng; gi++) {
    ag_grid *g = &m->lev[L - 2].g[gi];
    for (j = 0; j < gsize(g); j++) g->f[OA_RES][j] = (g->f[OA_U][j] - g->f[OA_LOP][j]) / m->w0_i;
  }
  tf_clear(m);
  m->tf[OA_RES] = m->iord; /* f_res carries mg_interp (amrd/io.c:860-861) */
  communicate(m, L - 1, L, 0, 1);
  for (gi = 0; gi < m->lev[L - 1].ng; gi++) {
    ag_grid *g = &m->lev[L - 1].g[gi];
    for (j = 0; j < gsize(g); j++) g->f[OA_U][j] += g->f[OA_RES][j];
  }
}

static int is_coarsest(const or_amr *m, int L)
{
  int gi;
  for (gi = 0; gi < m->lev[L - 1].ng; gi++)
    if (m->lev[L - 1].g[gi].coarsest) return 1;
  return 0;
}

/* sync of the mg_sync fields: u (the elliptic variable) */
static void sync_u(or_amr *m, int L, int bdy)
{
  tf_clear(m);
  m->tf[OA_U] = 1;
  do_sync(m, L, bdy);
}

/* amrd/mg.c:84-201 */
double or_amr_vcycle(or_amr *m, double *lev_norms, int *citer)
{
  int L, Lf = m->nlev, iter, gi;
  double res[OR_MAXLEV + 2], ctol;
  for (L = 0; L < OR_MAXLEV + 2; L++) res[L] = 0;
  clear_mg_vars(m, Lf);
  for (L = Lf; L > 0; L--) {
    if (L < Lf) {
      res[L + 1] = hook_residual_level(m, L + 1);
      tf_clear(m); /* set_res_sync (amrd/mg.c:606-616) */
      m->tf[OA_RES] = 1;
      do_sync(m, L + 1, 0);
      clear_mg_vars(m, L);
      do_inject(m, L + 1);
      sync_u(m, L, 0);
      or_amr_set_cmask(m, L, 1);
      hook_lop_level(m, L);
      compute_mg_rhs(m, L);
      if (m->save_tre) { /* amrd/mg.c:306-313: sync on L, interp to L+1 */
        tf_clear(m);
        m->tf[OA_TRE] = 1;
        do_sync(m, L, 0);
        tf_clear(m);
        m->tf[OA_TRE] = 2;
        communicate(m, L, L + 1, 0, 1);
      }
    }
    or_amr_set_cmask(m, L, 0);
    for (iter = 0; iter < m->pre; iter++) {
      res[L] = hook_relax_level(m, L, 0);
      sync_u(m, L, iter == 0 ? 0 : 1);
    }
  }
  for (L = 1; L <= Lf; L++) {
    ctol = res[L] * m->crtol;
    if (L > 1) compute_cgc(m, L);
    for (iter = 0; iter < m->pst; iter++) {
      res[L] = hook_relax_level(m, L, 0);
      sync_u(m, L, iter == 0 ? 0 : 1);
    }
    if (is_coarsest(m, L)) {
      res[L] = ctol + 1;
      for (gi = 0; gi < m->lev[L - 1].ng; gi++) m->lev[L - 1].g[gi].comm = m->lev[L - 1].g[gi].coarsest;
      while (iter < m->max_citer && res[L] > ctol) {
        res[L] = hook_relax_level(m, L, 1);
        sync_u(m, L, 1);
        iter++;
      }
      for (gi = 0; gi < m->lev[L - 1].ng; gi++) m->lev[L - 1].g[gi].comm = 1;
      if (citer) *citer = iter;
    }
    if (L == Lf) res[Lf] = hook_residual_level(m, L);
  }
  if (lev_norms)
    for (L = 1; L <= Lf; L++) lev_norms[L - 1] = res[L];
  return res[Lf];
}

double or_amr_residual_fine(or_amr *m)
{
  or_amr_set_cmask(m, m->nlev, 0);
  return hook_residual_level(m, m->nlev);
}

/* ------------------------------------------------------------------------ */
/* bookkeeping exports                                                       */
/* ------------------------------------------------------------------------ */
int or_amr_owned(or_amr *m, int L, int interior, int extend, int *gidx, double *bboxes, int cap)
{
  segl o = {0, 0, 0};
  int k, n;
  build_owned(m, L, interior, extend, &o);
  n = o.n;
  for (k = 0; k < n && k < cap; k++) {
    gidx[k] = o.s[k].g;
    memcpy(bboxes + 6 * k, o.s[k].bb, 6 * sizeof(double));
  }
  sl_free(&o);
  return n;
}

int or_amr_transfers(or_amr *m, int kind, int L, int arg, int *out, int cap)
{
  tpair *pr = 0;
  int l1, l2, interior, extend, np, q;
  if (kind == 0) { l1 = l2 = L; interior = 1 + arg; extend = 0; }
  else if (kind == 1) { l1 = L + 1; l2 = L; interior = 1; extend = 0; }
  else { l1 = L - 1; l2 = L; interior = 0; extend = 1; }
  np = make_transfers(m, l1, l2, interior, extend, &pr);
  for (q = 0; q < np && q < cap; q++) {
    ag_grid *g1 = &m->lev[l1 - 1].g[pr[q].src.g], *g2 = &m->lev[l2 - 1].g[pr[q].dst.g];
    int irho, delta = calc_delta(g1, g2, &irho), *o = out + 16 * q;
    double dx[3];
    if (delta == 1) align_src(g1, g2, &pr[q].src, &pr[q].dst);
    o[0] = pr[q].src.g;
    o[1] = pr[q].dst.g;
    o[2] = delta;
    o[3] = irho;
    or_calc_gs_ibbox(pr[q].src.bb, g1->bbox, g1->shape, 3, o + 4, dx);
    or_calc_gs_ibbox(pr[q].dst.bb, g2->bbox, g2->shape, 3, o + 10, dx);
  }
  free(pr);
  return np;
}

/* single transfers on a chosen field, for tests */
void or_amr_sync_field(or_amr *m, int L, int bdy, int which)
{
  tf_clear(m);
  m->tf[which] = 1;
  do_sync(m, L, bdy);
}
void or_amr_inject_field(or_amr *m, int Lf, int which, int op)
{
  tf_clear(m);
  m->tf[which] = op;
  communicate(m, Lf, Lf - 1, 1, 0);
}
void or_amr_interp_field(or_amr *m, int Lc, int which) { or_amr_interp_field_ord(m, Lc, which, 2); }

void or_amr_interp_field_ord(or_amr *m, int Lc, int which, int ord)
{
  tf_clear(m);
  m->tf[which] = ord;
  communicate(m, Lc, Lc + 1, 0, 1);
}
