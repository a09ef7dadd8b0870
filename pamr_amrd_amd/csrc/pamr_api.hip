// pamr_api.hip -- PAMR's multigrid-path calls (include/pamr.h) and AMRD's
// multigrid driver (include/amrd_mg.h) over the device-resident hierarchy of
// pmg_amr.
//
// Storage.  One current context, as the reference (src/gh.c:23-24).  Every
// PAMR variable owns pmg_amr fields: one per AMR time level when it is in
// the AMR hierarchy, one MG field when it is in the MG hierarchy only.  The
// MG grid function of a variable that lives in both hierarchies is its AMR
// field of the time level given to PAMR_build_mgh: the reference shares that
// memory between the hierarchies (src/gh.c:1139-1157).  pmg_amr level
// nbase holds AMR level 1; the levels below it are the MG levels built by
// coarsening (src/gh.c:1275-1340) and belong to the MG hierarchy only.
//
// gfn numbering follows PAMR_def_var_brief (src/pamr.c:1126-1170): a
// variable takes num_tl AMR gfns (when in_amrh) and then one MG gfn (when
// in_mgh), numbered from 1 in definition order.  A grid's gfs[] has one
// entry per gfn: a device pointer, or NULL where that grid function does
// not exist (AMR-only variables on MG-only levels; MG variables while no MG
// hierarchy exists), as compose_cgh / amr_to_mgh leave them
// (src/gh.c:884-900,1139-1157).
//
// Transfers select their grid functions with the reference's transfer bits
// (src/pamr.c:1767-1903) and run as the batched transfer plans of pmg_amr:
// one launch per operator and up to five fields.  All device work goes to
// pmg_amr's stream; hooks are expected to use the same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amrd_mg.h"
#include "../../include/pamr.h"
#include "../../include/pamr_amr_gpu.h"
#include "../../include/pamr_mg_gpu.h"

namespace {

constexpr int kMaxLevs = 50;    // PAMR_MAX_LEVS (include/internal_opts.h:11)
constexpr int kMaxGfns = 512;   // PAMR_MAX_GFNS (include/internal_opts.h:17)
constexpr int kIterStack = 10;  // ITER_G_STACKSIZE (src/gh.h:11)
constexpr int kNT = 256;

struct Attr {
  int in_amrh = 0, in_mgh = 0, num_tl = 0;
  int amr_inject = 0, amr_interp = 0, amr_bdy_interp = 0, amr_sync = 0;
  int mg_inject = 0, mg_interp = 0, mg_sync = 0, mg_noinj_to_amr = 0, regrid_transfer = 0;
  int pbt[6] = {0, 0, 0, 0, 0, 0};
};

struct Var {
  std::string name;
  Attr a;
  int sgfn = 0;
  std::vector<int> amr_field;  // pmg field per AMR time level
  int mg_field = -1;           // MG-only storage
};

struct View {
  std::vector<double *> gfs;
  std::vector<double> x[3];
  int comm = 1;
};

struct Iter {
  int hier = 0, L = 0, g = -1;
};

struct Ctx {
  int shape[3] = {0, 0, 0};
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  double lambda = 1.0;
  int rho_sp[kMaxLevs], rho_tm[kMaxLevs];
  int mcw[3] = {5, 5, 5};
  double dx[kMaxLevs][3], dt[kMaxLevs];
  Attr staged;  // attributes the next PAMR_def_var_brief uses
  std::vector<Var> vars;
  int gfns = 0;
  pmg_amr *A = nullptr;
  int namr = 0, nbase = 0, nlev = 0;  // AMR levels; pmg level of AMR level 1; pmg levels
  double t[kMaxLevs];
  std::vector<int> tf;
  bool frozen = false;
  bool mgh = false;
  int mgh_tl = 1;
  std::vector<std::vector<View>> views;  // [pmg level - 1][grid]
  Iter it;
  Iter stack[kIterStack];
  int sp = 0;
  hipStream_t stream = nullptr;
  // excision (src/pamr.c:1983-2037): the mask variable and the application's fill
  bool ex_on = false;
  double excised = 0.0;
  std::string ex_var;
  void (*ex_fill)(double *, int, int *, double *, double) = nullptr;
  // the MG hierarchy's levels: A's when it is built from AMR level 1, else an
  // island hierarchy over AMR levels mg_min .. finest (kept in subs for reuse)
  pmg_amr *M = nullptr;
  int mg_min = 1;
  std::vector<std::pair<int, pmg_amr *>> subs;
  std::vector<std::vector<View>> mviews;  // views of M's levels when M != A
  int comm_coarsest = 0;  // MG level where only the coarsest grids communicate (0: none)
};

Ctx *C = nullptr;

bool need_ctx(const char *who) {
  if (!C) std::printf("%s: error ... no current context\n", who);
  return C != nullptr;
}

// recalc_lev_info (src/gh.c:55-79)
void recalc_lev_info() {
  for (int i = 0; i < 3; i++) {
    int j = 0;
    C->dx[0][i] = C->shape[i] > 1 ? (C->bbox[2 * i + 1] - C->bbox[2 * i]) / (C->shape[i] - 1) : 1e-10;
    if (i == 0 || C->dt[0] > C->lambda * C->dx[0][i]) C->dt[0] = C->lambda * C->dx[0][i];
    while (j < kMaxLevs - 1) {
      j++;
      C->dx[j][i] = C->dx[j - 1][i] / C->rho_sp[j - 1];
      C->dt[j] = C->dt[j - 1] / C->rho_tm[j - 1];
    }
  }
}

Var *find_var(const char *name) {
  if (!name) return nullptr;
  for (Var &v : C->vars)
    if (v.name == name) return &v;
  return nullptr;
}

// the variable and role of a gfn: tl >= 1 an AMR time level, 0 the MG gfn
Var *var_of_gfn(int gfn, int *tl) {
  for (Var &v : C->vars) {
    const int namr = v.a.in_amrh ? v.a.num_tl : 0;
    const int n = namr + (v.a.in_mgh ? 1 : 0);
    if (gfn >= v.sgfn && gfn < v.sgfn + n) {
      *tl = gfn - v.sgfn < namr ? gfn - v.sgfn + 1 : 0;
      return &v;
    }
  }
  return nullptr;
}

// pmg field behind a gfn (-1: none)
int field_of_gfn(int gfn) {
  int tl = 0;
  Var *v = var_of_gfn(gfn, &tl);
  if (!v) return -1;
  if (tl > 0) return v->amr_field[tl - 1];
  if (v->a.in_amrh) {  // the MG gfn aliases AMR time level min(tl, num_tl) (src/gh.c:1143-1146)
    if (C->mgh_tl < 1 || v->amr_field.empty()) return -1;
    return v->amr_field[std::min<size_t>(C->mgh_tl, v->amr_field.size()) - 1];
  }
  return v->mg_field;
}

// the device hierarchy behind a PAMR hierarchy
pmg_amr *hobj(int hier) { return (hier == PAMR_MGH && C->M) ? C->M : C->A; }
int mg_nlev() { return (C && C->mgh && C->M) ? pmg_amr_nlev(C->M) : 0; }
// MG level holding AMR level mg_min
int mg_nbase() { return (C && C->mgh && C->M) ? (C->M == C->A ? C->nbase : pmg_amr_base_level(C->M)) : 0; }

// pmg level of level l of a hierarchy (0: none)
int plevel(int l, int hier) {
  if (!C || !C->A) return 0;
  if (hier == PAMR_AMRH) return (l >= 1 && l <= C->namr + 1) ? C->nbase + l - 1 : 0;
  if (hier == PAMR_MGH) return (C->mgh && l >= 1 && l <= mg_nlev()) ? l : 0;
  return 0;
}

// time of level l of a hierarchy: MG-only levels carry the time of the AMR
// level the MG hierarchy starts at (amr_to_mgh / build_mgh copy the AMR
// level's t, src/gh.c:1108,1312)
double level_time(int l, int hier) {
  const int Lp = plevel(l, hier);
  const int amr_l = hier == PAMR_MGH ? C->mg_min + std::max(0, Lp - mg_nbase()) : Lp - C->nbase + 1;
  return C->t[std::max(1, amr_l) - 1];
}

std::vector<std::vector<View>> &views_of(int hier) {
  return (hier == PAMR_MGH && C->M && C->M != C->A) ? C->mviews : C->views;
}

// grid views of the levels of P: AMR gfns on levels >= Lamr0, MG gfns on
// levels >= Lmg0 (amr_to_mgh shares the AMR grid's gfs[], src/gh.c:1128-1146)
void fill_views(pmg_amr *P, int Lamr0, int Lmg0, std::vector<std::vector<View>> &views) {
  const int nl = pmg_amr_nlev(P);
  views.assign(nl, {});
  for (int L = 1; L <= nl; L++) {
    const int ng = pmg_amr_ngrids(P, L);
    std::vector<View> &vs = views[L - 1];
    vs.resize(ng);
    for (int g = 0; g < ng; g++) {
      View &w = vs[g];
      w.gfs.assign(C->gfns, nullptr);
      for (int gfn = 1; gfn <= C->gfns; gfn++) {
        int tl = 0;
        Var *v = var_of_gfn(gfn, &tl);
        if (!v) continue;
        const bool present = tl > 0 ? L >= Lamr0 : L >= Lmg0;
        if (!present) continue;
        const int f = field_of_gfn(gfn);
        if (f >= 0) w.gfs[gfn - 1] = pmg_amr_field(P, L, g, f);
      }
      int shp[3];
      double bb[6];
      pmg_amr_grid_info(P, L, g, shp, bb);
      for (int d = 0; d < 3; d++) {  // compose_cgh's accumulation (src/gh.c:902-906)
        w.x[d].resize(shp[d]);
        const double h = (bb[2 * d + 1] - bb[2 * d]) / (shp[d] - 1);
        double xv = bb[2 * d];
        for (int m = 0; m < shp[d]; m++, xv += h) w.x[d][m] = xv;
      }
    }
  }
}

void build_views() {
  const int big = 1 << 30;
  fill_views(C->A, C->nbase, !C->mgh ? big : (C->M == C->A ? 1 : C->nbase + C->mg_min - 1), C->views);
  if (C->mgh && C->M && C->M != C->A) fill_views(C->M, mg_nbase(), 1, C->mviews);
  else C->mviews.clear();
}

View *cur_view(const char *who) {
  if (!need_ctx(who)) return nullptr;
  if (C->it.g < 0) {
    std::printf("%s: error ... iterator not initialized\n", who);
    return nullptr;
  }
  const int Lp = plevel(C->it.L, C->it.hier);
  std::vector<std::vector<View>> &vw = views_of(C->it.hier);
  if (!Lp || Lp > (int)vw.size() || C->it.g >= (int)vw[Lp - 1].size()) {
    std::printf("%s: error ... iterator not initialized\n", who);
    return nullptr;
  }
  return &vw[Lp - 1][C->it.g];
}

// the pmg field of the excision mask (-1: excision off or no hierarchy)
int mask_field() {
  if (!C->ex_on || !C->A) return -1;
  Var *v = find_var(C->ex_var.c_str());
  return (v && !v->amr_field.empty()) ? v->amr_field[0] : -1;
}

// init_ex_mask (src/gh.c:1555-1577) on levels L0..L1 of P: the application's
// fill function on a host array of each grid, uploaded into the mask field
bool init_ex_mask(pmg_amr *P, int L0, int L1) {
  const int f = mask_field();
  if (f < 0 || !C->ex_fill) return false;
  if (hipStreamSynchronize(C->stream) != hipSuccess) return false;
  for (int L = L0; L <= L1; L++)
    for (int g = 0; g < pmg_amr_ngrids(P, L); g++) {
      int shp[3];
      double bb[6];
      pmg_amr_grid_info(P, L, g, shp, bb);
      std::vector<double> h((size_t)shp[0] * shp[1] * shp[2], 0.0);
      C->ex_fill(h.data(), 3, shp, bb, C->excised);
      if (hipMemcpy(pmg_amr_field(P, L, g, f), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess)
        return false;
    }
  return true;
}

// fields of the gfns whose transfer bit is set, with their operator
void selected(std::vector<int> &fields, std::vector<int> &ops) {
  fields.clear();
  ops.clear();
  for (int gfn = 1; gfn <= C->gfns; gfn++) {
    const int op = C->tf[gfn - 1];
    if (!op) continue;
    const int f = field_of_gfn(gfn);
    if (f < 0) continue;
    fields.push_back(f);
    ops.push_back(op);
  }
}

// ---------------------------------------------------------------------------
// elementwise level kernels of amrd/mg.c (whole field pools of one level)
// ---------------------------------------------------------------------------
// compute_MG_rhs (amrd/mg.c:286-303, save_tre = 0)
__global__ __launch_bounds__(kNT) void k_mg_rhs(double *rhs, double *lop, const double *res, const double *f,
                                                const double *cmask, double w0_r, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  rhs[i] = cmask[i] == 1.0 ? -res[i] * w0_r + lop[i] : 0.0;
  lop[i] = f[i];
}

// compute_MG_rhs with save_tre (amrd/mg.c:286-303): the running max of
// |lop| * sqrt(num_f_tre_vars) over the elliptic variables, taken before lop
// is overwritten, restarted at 0 by the first variable
__global__ __launch_bounds__(kNT) void k_mg_rhs_tre(double *rhs, double *lop, const double *res, const double *f,
                                                    const double *cmask, double w0_r, double *tre, int first,
                                                    double scale, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  const double t = first ? 0.0 : tre[i], v = fabs(lop[i]) * scale;
  tre[i] = t > v ? t : v;  // max (amrd/util.h:12)
  rhs[i] = cmask[i] == 1.0 ? -res[i] * w0_r + lop[i] : 0.0;
  lop[i] = f[i];
}

// MG extrapolation in time (amrd/evolve.c:346-368): the tl = 1 value of an
// elliptic variable from its extrapolation variables
__global__ __launch_bounds__(kNT) void k_extrap_etm1(double *v1, const double *etm1, const double *etm2, double f1,
                                                     double f2, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  v1[i] = f1 * etm1[i] + f2 * etm2[i];
}
__global__ __launch_bounds__(kNT) void k_extrap_tn(double *v1, const double *v2, const double *etm1,
                                                   const double *etm2, double r, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  v1[i] = v2[i] + r * (etm2[i] - etm1[i]);
}
// update_mg_extrap_vars (amrd/evolve.c:1684-1697): tm2 <- tm1, tm1 <- v(tl 2),
// then (eps) tm2 += (tm1 - brs) * f
__global__ __launch_bounds__(kNT) void k_update_extrap(double *tm1, double *tm2, const double *v2,
                                                       const double *brs, int eps, double f, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  double t2 = tm1[i];
  const double t1 = v2[i];
  if (eps) t2 += (t1 - brs[i]) * f;
  tm2[i] = t2;
  tm1[i] = t1;
}

// compute_cgc, coarse part (amrd/mg.c:336-339)
__global__ __launch_bounds__(kNT) void k_cgc_res(double *res, const double *f, const double *lop, double w0_i,
                                                 long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  res[i] = (f[i] - lop[i]) / w0_i;
}

// compute_cgc, fine part (amrd/mg.c:348-351)
__global__ __launch_bounds__(kNT) void k_add_res(double *f, const double *res, long long n) {
  const long long i = (long long)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  f[i] += res[i];
}

// ---------------------------------------------------------------------------
// AMRD side
// ---------------------------------------------------------------------------
struct Amrd {
  amrd_real_hook residual = nullptr, relax = nullptr;
  amrd_void_hook lop = nullptr;
  amrd_level_relax_hook lrelax = nullptr;
  amrd_level_real_hook lresidual = nullptr;
  amrd_level_void_hook llop = nullptr;
  std::vector<std::string> ell;
  std::vector<std::string> ell_t0;  // elliptic_vars_t0: solved alongside while is_t0
  std::vector<std::string> cnst;    // MG_cnst_data_vars
  int is_t0 = 1;                    // AMRD_is_t0 (amrd/globals.c:248)
  amrd_void_hook t0_init = nullptr, t0_cnst = nullptr;
  std::string tre;  // "<v>_tre" of the first truncation-error variable (also in the MG hierarchy)
  int ntre = 0;     // AMRD_num_f_tre_vars
  int do_ex = 0;
  int pre = 3, pst = 3, min_iter = 1, max_iter = 50, max_citer = 5000;
  double tol = 0.0, crtol = 1e-3, w0_r = 1.0, w0_i = 1.0;
  double last[kMaxLevs + 2];
  int last_nlev = 0, last_citer = 0;
  int cwarn = 10;
  // tracing (amrd/io.c:80-83): MG_trace, MG_DV_trace and its time window;
  // mg_dump's solve counter and last giter (static in amrd/mg.c:713)
  int trace = 0, dv_trace = -1;
  double dv_t_on = 1e-15, dv_t_off = 1e-15;
  double dv_lcount = -1, dv_pgiter = -1;
  std::vector<std::string> hyp, mgh_work, amrh_work_mg;  // mg_dump's which = 0 lists
};
Amrd D;

// the reference's indentation of the per-level trace lines (amrd/mg.c:138)
void trace_indent(int L, int Lf) {
  for (int i = L - 1; i < Lf; i++) std::printf("   ");
}


std::vector<std::string> split_list(const char *s) {
  std::vector<std::string> out;
  if (!s) return out;
  std::string cur;
  for (const char *p = s;; p++) {
    if (*p == 0 || *p == ' ' || *p == ',' || *p == '\t' || *p == '\n') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
      if (*p == 0) break;
    } else {
      cur += *p;
    }
  }
  return out;
}

bool in_list(const std::string &n, const std::vector<std::string> &l) {
  return std::find(l.begin(), l.end(), n) != l.end();
}

// ---------------------------------------------------------------------------
// grid-function output behind PAMR_save_gfn: the application's writer (for
// example bbhutil's gft_out_bbox / gft_close) or the built-in record file
// ---------------------------------------------------------------------------
PAMR_gft_out_bbox_fn gft_out_app = nullptr;
PAMR_gft_close_fn gft_close_app = nullptr;
std::vector<std::string> pgf_open;  // built-in outputs written since their last close

int pgf_out_bbox(char *name, double t, int *shape, int rank, double *bbox, double *data) {
  const std::string path = std::string(name) + ".pgf";
  const bool fresh = !in_list(name, pgf_open);
  FILE *f = std::fopen(path.c_str(), fresh ? "wb" : "ab");
  if (!f) {
    std::printf("PAMR_save_gfn: error ... cannot open %s\n", path.c_str());
    return 0;
  }
  if (fresh) pgf_open.push_back(name);
  size_t n = 1;
  for (int d = 0; d < rank; d++) n *= (size_t)shape[d];
  const int32_t r = rank;
  bool ok = std::fwrite("PGF1", 1, 4, f) == 4 && std::fwrite(&r, 4, 1, f) == 1 && std::fwrite(&t, 8, 1, f) == 1;
  for (int d = 0; d < rank && ok; d++) {
    const int32_t s = shape[d];
    ok = std::fwrite(&s, 4, 1, f) == 1;
  }
  ok = ok && std::fwrite(bbox, 8, 2 * (size_t)rank, f) == 2 * (size_t)rank && std::fwrite(data, 8, n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) std::printf("PAMR_save_gfn: error ... writing %s failed\n", path.c_str());
  return ok ? 1 : 0;
}

int pgf_close(char *name) {
  pgf_open.erase(std::remove(pgf_open.begin(), pgf_open.end(), std::string(name)), pgf_open.end());
  return 1;
}

struct EllGfns {
  int f, res, lop, rhs;
};

bool ell_gfns(std::vector<EllGfns> &out, int &cmask) {
  out.clear();
  // num = AMRD_num_elliptic_vars (+ AMRD_num_elliptic_vars_t0 while is_t0),
  // amrd/mg.c:248,279,327,612
  std::vector<std::string> all = D.ell;
  if (D.is_t0) all.insert(all.end(), D.ell_t0.begin(), D.ell_t0.end());
  for (const std::string &v : all) {
    EllGfns e;
    e.f = PAMR_get_gfn((char *)v.c_str(), PAMR_MGH, 0);
    e.res = PAMR_get_gfn((char *)(v + "_res").c_str(), PAMR_MGH, 0);
    e.lop = PAMR_get_gfn((char *)(v + "_lop").c_str(), PAMR_MGH, 0);
    e.rhs = PAMR_get_gfn((char *)(v + "_rhs").c_str(), PAMR_MGH, 0);
    if (e.f <= 0 || e.res <= 0 || e.lop <= 0 || e.rhs <= 0) return false;
    out.push_back(e);
  }
  cmask = PAMR_get_gfn((char *)"cmask", PAMR_MGH, 0);
  return cmask > 0;
}

// a level's field pool; L an MG level, or a pmg level of A for hier AMRH
double *pool(int L, int gfn, long long &n, int hier = PAMR_MGH) {
  pmg_amr *P = hobj(hier);
  n = pmg_amr_pool_points(P, L);
  const int f = field_of_gfn(gfn);
  return f < 0 ? nullptr : pmg_amr_field(P, L, 0, f);
}

int nblocks(long long n) { return (int)((n + kNT - 1) / kNT); }

// clear_MG_vars (amrd/mg.c:242-262)
bool clear_mg_vars(int L, const std::vector<EllGfns> &E) {
  for (const EllGfns &e : E)
    for (int gfn : {e.res, e.rhs, e.lop}) {
      long long n;
      double *p = pool(L, gfn, n);
      if (!p || hipMemsetAsync(p, 0, n * sizeof(double), C->stream) != hipSuccess) return false;
    }
  return true;
}

// compute_MG_rhs (amrd/mg.c:271-314)
bool compute_mg_rhs(int L, const std::vector<EllGfns> &E, int cmask, int save_tre) {
  long long n;
  const double *cm = pool(L, cmask, n);
  int tre_gfn = 0;
  double *tre = nullptr;
  if (save_tre) {
    if (!D.ntre || (tre_gfn = PAMR_get_gfn((char *)D.tre.c_str(), PAMR_MGH, 0)) <= 0 ||
        !(tre = pool(L, tre_gfn, n))) {
      std::printf("compute_MG_rhs: error ... save_tre=1, but no tre variables\n");
      return false;
    }
  }
  bool first = true;
  for (const EllGfns &e : E) {
    double *rhs = pool(L, e.rhs, n), *lop = pool(L, e.lop, n), *res = pool(L, e.res, n), *f = pool(L, e.f, n);
    if (!cm || !rhs || !lop || !res || !f) return false;
    if (save_tre)
      hipLaunchKernelGGL(k_mg_rhs_tre, dim3(nblocks(n)), dim3(kNT), 0, C->stream, rhs, lop, res, f, cm, D.w0_r, tre,
                         first ? 1 : 0, std::sqrt((double)D.ntre), n);
    else
      hipLaunchKernelGGL(k_mg_rhs, dim3(nblocks(n)), dim3(kNT), 0, C->stream, rhs, lop, res, f, cm, D.w0_r, n);
    if (hipGetLastError() != hipSuccess) return false;
    first = false;
  }
  if (save_tre) {  // set_gfn_sync, sync on L; set_gfn_in, interp to L+1 (amrd/mg.c:306-313)
    PAMR_freeze_tf_bits();
    PAMR_clear_tf_bits();
    PAMR_set_tf_bit(tre_gfn, PAMR_SYNC);
    if (!PAMR_sync(L, 0, PAMR_MGH, 0)) return false;
    PAMR_freeze_tf_bits();
    PAMR_clear_tf_bits();
    PAMR_set_tf_bit(tre_gfn, PAMR_SECOND_ORDER);
    if (!PAMR_interp(L, 0, PAMR_MGH)) return false;
    PAMR_thaw_tf_bits();
  }
  return true;
}

// compute_cgc (amrd/mg.c:323-355)
bool compute_cgc(int L, const std::vector<EllGfns> &E) {
  long long n;
  for (const EllGfns &e : E) {
    double *res = pool(L - 1, e.res, n), *f = pool(L - 1, e.f, n), *lop = pool(L - 1, e.lop, n);
    if (!res || !f || !lop) return false;
    hipLaunchKernelGGL(k_cgc_res, dim3(nblocks(n)), dim3(kNT), 0, C->stream, res, f, lop, D.w0_i, n);
    if (hipGetLastError() != hipSuccess) return false;
  }
  if (!PAMR_interp(L - 1, 0, PAMR_MGH)) return false;
  for (const EllGfns &e : E) {
    double *res = pool(L, e.res, n), *f = pool(L, e.f, n);
    if (!res || !f) return false;
    hipLaunchKernelGGL(k_add_res, dim3(nblocks(n)), dim3(kNT), 0, C->stream, f, res, n);
    if (hipGetLastError() != hipSuccess) return false;
  }
  return true;
}

// set_cmask_bdy (which 0, amrd/mg.c:432-488) / set_cmask_child (which 1,
// amrd/mg.c:362-426) on every grid of MG level L
bool set_cmask(int L, int which, int cmask) {
  const int f = field_of_gfn(cmask);
  return f >= 0 && pmg_amr_fill_cmask(hobj(PAMR_MGH), L, which, f) == 0;
}

// call_app / call_rr_app (amrd/util.c:39-103) over the MG grids of L
double call_rr(amrd_real_hook h, int L, int comm_only) {
  double lres = 0;
  bool first = true;
  int valid = PAMR_init_s_iter(L, PAMR_MGH, 0);
  while (valid) {
    int comm = 1;
    PAMR_get_g_comm(&comm);
    if (comm || !comm_only) {
      const double r = h();
      if (first) lres = r;
      lres = std::max(r, lres);
      first = false;
    }
    valid = PAMR_next_g();
  }
  return lres;
}

void call_app(amrd_void_hook h, int L) {
  int valid = PAMR_init_s_iter(L, PAMR_MGH, 0);
  while (valid) {
    h();
    valid = PAMR_next_g();
  }
}

double relax_level(int L, int comm_only) {
  return D.lrelax ? D.lrelax(L, comm_only) : call_rr(D.relax, L, comm_only);
}
double residual_level(int L) { return D.lresidual ? D.lresidual(L) : call_rr(D.residual, L, 0); }
void lop_level(int L) {
  if (D.llop) D.llop(L);
  else call_app(D.lop, L);
}

// is_coarsest (amrd/mg.c:553-565): a grid of MG level L is flagged coarsest
bool is_coarsest(int L) {
  for (int g = 0; g < pmg_amr_ngrids(hobj(PAMR_MGH), L); g++)
    if (pmg_amr_grid_coarsest(hobj(PAMR_MGH), L, g) == 1) return true;
  return false;
}

// set_comm_coarsest / reset_comm (amrd/mg.c:571-600): only the coarsest grids
// of L communicate, or all of them again
void set_comm(int L, bool coarsest_only) {
  std::vector<std::vector<View>> &vw = views_of(PAMR_MGH);
  if (L < 1 || L > (int)vw.size()) return;
  for (int g = 0; g < (int)vw[L - 1].size(); g++)
    vw[L - 1][g].comm = coarsest_only ? pmg_amr_grid_coarsest(hobj(PAMR_MGH), L, g) == 1 : 1;
  C->comm_coarsest = coarsest_only ? L : 0;
}

// set_res_sync (amrd/mg.c:606-616)
void set_res_sync(const std::vector<EllGfns> &E) {
  PAMR_freeze_tf_bits();
  PAMR_clear_tf_bits();
  for (const EllGfns &e : E) PAMR_set_tf_bit(e.res, PAMR_SYNC);
}

struct ExtrapGfns {
  int v1, v2, tm1, tm2, brs;
};
bool extrap_gfns(std::vector<ExtrapGfns> &out, const char *who) {
  out.clear();
  for (const std::string &v : D.ell) {
    ExtrapGfns e;
    e.v1 = PAMR_get_gfn((char *)v.c_str(), PAMR_AMRH, 1);
    e.v2 = PAMR_get_gfn((char *)v.c_str(), PAMR_AMRH, 2);
    e.tm1 = PAMR_get_gfn((char *)(v + "_extrap_tm1").c_str(), PAMR_AMRH, 1);
    e.tm2 = PAMR_get_gfn((char *)(v + "_extrap_tm2").c_str(), PAMR_AMRH, 1);
    e.brs = PAMR_get_gfn((char *)(v + "_brs").c_str(), PAMR_AMRH, 1);
    if (e.v1 <= 0 || e.v2 <= 0 || e.tm1 <= 0 || e.tm2 <= 0 || e.brs <= 0) {
      std::printf("%s: error ... <%s> needs two AMR time levels and its extrapolation variables "
                  "(amrd_mg_def_extrap_vars)\n", who, v.c_str());
      return false;
    }
    out.push_back(e);
  }
  return true;
}

// <v>_res, <v>_lop, <v>_rhs of each elliptic variable (amrd/io.c:850-897)
bool def_mg_work_vars(const char *who, const std::vector<std::string> &ell, const std::vector<std::string> &fw,
                      const std::vector<std::string> &i2) {
  int pbt[6] = {PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN};
  for (const std::string &v : ell) {
    Var *f = find_var(v.c_str());
    if (!f || !f->a.in_mgh) {
      std::printf("%s: error ... elliptic variable <%s> is not defined in the MG hierarchy\n", who, v.c_str());
      return false;
    }
    const int mg_inject = in_list(v + "_res", fw) ? PAMR_FW_RESTR : PAMR_HW_RESTR;
    const int mg_interp = in_list(v, i2) ? PAMR_SECOND_ORDER : PAMR_NO_INTERP;
    if (!PAMR_def_var_full((char *)(v + "_res").c_str(), 0, 1, 0, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP,
                           PAMR_NO_SYNC, mg_inject, mg_interp, PAMR_NO_SYNC, 0, PAMR_NO_TRANSFER, pbt))
      return false;
    for (const char *suf : {"_lop", "_rhs"})
      if (!PAMR_def_var_full((char *)(v + suf).c_str(), 0, 1, 0, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP,
                             PAMR_NO_SYNC, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 0, PAMR_NO_TRANSFER, pbt))
        return false;
  }
  return true;
}

// sync / fourth-order interp / straight inject of a list of MG variables
// (sync_cnst_data ... inject_elliptic_vars_t0_data, amrd/mg.c:622-698)
bool mg_list_transfer(const std::vector<std::string> &vars, int what, int L) {
  PAMR_freeze_tf_bits();
  PAMR_clear_tf_bits();
  const int bit = what == 0 ? PAMR_SYNC : what == 1 ? PAMR_FOURTH_ORDER : PAMR_STRAIGHT_INJECT;
  for (const std::string &v : vars) PAMR_set_tf_bit(PAMR_get_gfn((char *)v.c_str(), PAMR_MGH, 0), bit);
  const int ok = what == 0 ? PAMR_sync(L, 0, PAMR_MGH, 0) : what == 1 ? PAMR_interp(L, 0, PAMR_MGH)
                                                                    : PAMR_inject(L, 0, PAMR_MGH);
  PAMR_thaw_tf_bits();
  return ok != 0;
}

// the t = 0 hook pass of solve_elliptics (amrd/mg.c:45-52, 60-68): the hook on
// every MG level from the coarsest up, each level synchronised and
// interpolated to the next, then injected back down
bool t0_pass(amrd_void_hook h, const std::vector<std::string> &vars, int Lf) {
  for (int l = 1; l <= Lf; l++) {
    call_app(h, l);
    if (!mg_list_transfer(vars, 0, l)) return false;
    if (l < Lf && !mg_list_transfer(vars, 1, l)) return false;
  }
  for (int l = Lf; l > 1; l--)
    if (!mg_list_transfer(vars, 2, l)) return false;
  return true;
}

}  // namespace

// ===========================================================================
// PAMR
// ===========================================================================
extern "C" {

int PAMR_init_context(double *q, int q_size, int dim, int *shape, double *bbox, char *cp_file, int cp_rank) {
  (void)cp_rank;
  if (dim != 3) {
    std::printf("PAMR_init_context: error ... only dim = 3 is supported\n");
    return 0;
  }
  if (q || q_size) {
    std::printf("PAMR_init_context: external-memory-block option currently not supported\n");
    return 0;
  }
  if (cp_file) {
    std::printf("PAMR_init_context: error ... restoring a checkpoint is not supported\n");
    return 0;
  }
  if (C) {
    std::printf("PAMR_init_context: error ... only one context is supported\n");
    return 0;
  }
  if (!shape || !bbox) return 0;
  C = new Ctx();
  for (int i = 0; i < 3; i++) {
    C->shape[i] = shape[i];
    C->bbox[2 * i] = bbox[2 * i];
    C->bbox[2 * i + 1] = bbox[2 * i + 1];
  }
  for (int l = 0; l < kMaxLevs; l++) {
    C->rho_sp[l] = C->rho_tm[l] = 2;
    C->t[l] = 0.0;
  }
  recalc_lev_info();
  return 1;
}

void PAMR_free_context(int num) {
  if (!C || num != 1) return;
  for (auto &sub : C->subs) pmg_amr_destroy(sub.second);
  if (C->A) pmg_amr_destroy(C->A);
  delete C;
  C = nullptr;
  D.ell.clear();  // AMRD's variable tables belong to the context's variables
  D.ell_t0.clear();
  D.cnst.clear();
  D.tre.clear();
  D.ntre = 0;
}

void PAMR_set_lambda(double lambda) {
  if (!need_ctx("PAMR_set_lambda")) return;
  C->lambda = lambda;
  recalc_lev_info();
}
void PAMR_get_lambda(double *lambda) {
  if (need_ctx("PAMR_get_lambda")) *lambda = C->lambda;
}

void PAMR_set_rho(int *rho_sp, int *rho_tm, int num) {
  if (!need_ctx("PAMR_set_rho")) return;
  if (num > kMaxLevs) {
    std::printf("PAMR_set_rho: error ... number of levels specified (%i) is greater than the maximum allowed (%i)\n",
                num, kMaxLevs);
    return;
  }
  for (int i = 0; i < num; i++) {
    if (C->A && C->rho_sp[i] != rho_sp[i])
      std::printf("PAMR_set_rho: warning ... cannot change rho_sp after  the initial hierarchy has been composed\n");
    else
      C->rho_sp[i] = rho_sp[i];
    C->rho_tm[i] = rho_tm[i];
  }
  recalc_lev_info();
}
void PAMR_get_rho(int *rho_sp, int *rho_tm, int num) {
  if (!need_ctx("PAMR_get_rho")) return;
  for (int i = 0; i < num && i < kMaxLevs; i++) {
    rho_sp[i] = C->rho_sp[i];
    rho_tm[i] = C->rho_tm[i];
  }
}
void PAMR_get_dxdt(int lev, double *dx, double *dt) {
  if (!need_ctx("PAMR_get_dxdt") || lev < 1 || lev > kMaxLevs) return;
  for (int i = 0; i < 3; i++) dx[i] = C->dx[lev - 1][i];
  *dt = C->dt[lev - 1];
}
void PAMR_set_ghost_width(int *ghost_width) {
  if (!need_ctx("PAMR_set_ghost_width")) return;
  for (int i = 0; i < 3; i++)
    if (ghost_width[i] != 0) std::printf("PAMR_set_ghost_width: warning ... one rank: ghost width stays 0\n");
}
void PAMR_get_ghost_width(int *ghost_width) {
  if (!need_ctx("PAMR_get_ghost_width")) return;
  for (int i = 0; i < 3; i++) ghost_width[i] = 0;
}
void PAMR_set_MG_coarse_width(int *min_width) {
  if (!need_ctx("PAMR_set_MG_coarse_width")) return;
  if (C->A) {
    std::printf("PAMR_set_MG_coarse_width: error ... cannot change after  the initial hierarchy has been composed\n");
    return;
  }
  for (int i = 0; i < 3; i++) {
    if (min_width[i] < 2) std::printf("PAMR_set_MG_coarse_width: error ... width[%i]<2\n", i);
    C->mcw[i] = std::max(2, min_width[i]);
  }
}
void PAMR_get_MG_coarse_width(int *min_width) {
  if (!need_ctx("PAMR_get_MG_coarse_width")) return;
  for (int i = 0; i < 3; i++) min_width[i] = C->mcw[i];
}
void PAMR_get_global_bbox(double *bbox) {
  if (!need_ctx("PAMR_get_global_bbox")) return;
  std::memcpy(bbox, C->bbox, 6 * sizeof(double));
}

// ---- variables ------------------------------------------------------------
int PAMR_def_var_brief(char *name) {
  if (!C) {
    std::printf("PAMR_def_var_[]: error ... no current context\n");
    return 0;
  }
  const Attr &s = C->staged;
  if (s.in_amrh && s.num_tl < 0) {
    std::printf("PAMR_def_var_[]: error ... for AMR variable,  num_tl=%i must be greater than 0\n", s.num_tl);
    return 0;
  }
  if (!s.in_amrh && s.num_tl > 0) {
    std::printf("PAMR_def_var_[]: error ... for MG variable,  num_tl=%i must be 0\n", s.num_tl);
    return 0;
  }
  if (C->A) {
    std::printf("PAMR_def_var_[]: error ... cannot define a new variable after  the initial hierarchy has been composed\n");
    return 0;
  }
  const int num = (s.in_amrh ? s.num_tl : 0) + (s.in_mgh ? 1 : 0);
  if (C->gfns + num > kMaxGfns) {
    std::printf("PAMR_def_var_[]: maximum number of grid functions exceeded\n");
    return 0;
  }
  if (!name || find_var(name)) {
    std::printf("PAMR_def_var_brief: error ... variable %s already exists\n", name ? name : "(null)");
    return 0;
  }
  Var v;
  v.name = name;
  v.a = s;
  v.sgfn = C->gfns + 1;
  C->gfns += num;
  C->vars.push_back(v);
  return v.sgfn;
}

int PAMR_def_var_full(char *name, int in_amrh, int in_mgh, int num_tl, int amr_inject, int amr_interp,
                      int amr_bdy_interp, int amr_sync, int mg_inject, int mg_interp, int mg_sync, int mg_noinj_to_amr,
                      int regrid_transfer, int *phys_bdy_type) {
  if (!C) {
    std::printf("PAMR_def_var_[]: error ... no current context\n");
    return 0;
  }
  Attr &s = C->staged;
  s.in_amrh = in_amrh; s.in_mgh = in_mgh; s.num_tl = num_tl;
  s.amr_inject = amr_inject; s.amr_interp = amr_interp; s.amr_bdy_interp = amr_bdy_interp; s.amr_sync = amr_sync;
  s.mg_inject = mg_inject; s.mg_interp = mg_interp; s.mg_sync = mg_sync;
  s.regrid_transfer = regrid_transfer; s.mg_noinj_to_amr = mg_noinj_to_amr;
  for (int i = 0; i < 6; i++) s.pbt[i] = phys_bdy_type ? phys_bdy_type[i] : 0;
  return PAMR_def_var_brief(name);
}

int PAMR_get_var_attribs(char *name, int *in_amrh, int *in_mgh, int *num_tl, int *amr_inject, int *amr_interp,
                         int *amr_bdy_interp, int *amr_sync, int *mg_inject, int *mg_interp, int *mg_sync,
                         int *mg_noinj_to_amr, int *regrid_transfer, int *phys_bdy_type) {
  if (!C) {
    std::printf("PAMR_get_var_attribs[]: error ... no current context\n");
    return 0;
  }
  Var *v = find_var(name);
  if (!v) {
    std::printf("PAMR_get_var_attribs[]: error ... variable <%s> does not exist\n", name ? name : "(null)");
    return 0;
  }
  *in_amrh = v->a.in_amrh; *in_mgh = v->a.in_mgh; *num_tl = v->a.num_tl;
  *amr_inject = v->a.amr_inject; *amr_interp = v->a.amr_interp; *amr_bdy_interp = v->a.amr_bdy_interp;
  *amr_sync = v->a.amr_sync; *mg_inject = v->a.mg_inject; *mg_interp = v->a.mg_interp; *mg_sync = v->a.mg_sync;
  *mg_noinj_to_amr = v->a.mg_noinj_to_amr; *regrid_transfer = v->a.regrid_transfer;
  for (int i = 0; i < 6; i++) phys_bdy_type[i] = v->a.pbt[i];
  return v->sgfn;
}

int PAMR_get_gfn(char *name, int hier, int tl) {
  if (!C) {
    std::printf("PAMR_get_gfn: error ... no current context\n");
    return 0;
  }
  Var *v = find_var(name);
  if (!v) {
    std::printf("PAMR_get_gfn: error ... variable <%s> does not exist\n", name ? name : "(null)");
    return 0;
  }
  int offset;
  if (hier == PAMR_MGH && !v->a.in_mgh) {
    std::printf("PAMR_get_gfn: error ... variable <%s> is not in the MG hierarchy\n", name);
    return 0;
  } else if (hier == PAMR_MGH) {
    offset = v->a.in_amrh ? v->a.num_tl : 0;
  } else if (hier == PAMR_AMRH && !v->a.in_amrh) {
    std::printf("PAMR_get_gfn: error ... variable <%s> is not in the AMR hierarchy\n", name);
    return 0;
  } else if (hier != PAMR_AMRH) {
    std::printf("PAMR_get_gfn: error ... unknown hierarhcy %i\n", hier);
    return 0;
  } else if (tl < 1 || tl > v->a.num_tl) {
    std::printf("PAMR_get_gfn: error ... invalid time level %i\n", tl);
    return 0;
  } else {
    offset = tl - 1;
  }
  return v->sgfn + offset;
}

// ---- hierarchy ------------------------------------------------------------
int PAMR_compose_hierarchy(int min_lev, int max_lev, int num, int *lev, double *bbox, double t) {
  if (!need_ctx("PAMR_compose_hierarchy")) return 0;
  if (C->mgh) {
    std::printf("PAMR_compose_hierarchy: error ... cannot re-compose with an existing MG hierarhcy\n");
    return 0;
  }
  if (C->A) {
    std::printf("PAMR_compose_hierarchy: error ... re-composing an existing hierarchy is not supported\n");
    return 0;
  }
  if (min_lev != 1 || max_lev < 1 || max_lev > kMaxLevs || num < 1 || !lev || !bbox) {
    std::printf("PAMR_compose_hierarchy: error ... levels %i..%i: the hierarchy must start at level 1\n", min_lev,
                max_lev);
    return 0;
  }
  for (int l = 0; l + 1 < max_lev; l++)
    if (C->rho_sp[l] != 2) {
      std::printf("PAMR_compose_hierarchy: error ... only rho_sp = 2 is supported\n");
      return 0;
    }
  int nbase = 0;
  std::vector<int> nbox(std::max(0, max_lev - 1), 0);
  std::vector<double> boxes;
  for (int l = 2; l <= max_lev; l++)
    for (int i = 0; i < num; i++)
      if (lev[i] == l) {
        nbox[l - 2]++;
        boxes.insert(boxes.end(), bbox + 6 * i, bbox + 6 * i + 6);
      }
  for (int i = 0; i < num; i++) {
    if (lev[i] < min_lev || lev[i] > max_lev) {
      std::printf("PAMR_compose_hierarchy: error ... box %i has level %i outside %i..%i\n", i, lev[i], min_lev,
                  max_lev);
      return 0;
    }
    if (lev[i] == 1) {
      nbase++;
      for (int d = 0; d < 3; d++) {
        const double fz = C->dx[0][d] / 2;
        if (std::fabs(bbox[6 * i + 2 * d] - C->bbox[2 * d]) >= fz ||
            std::fabs(bbox[6 * i + 2 * d + 1] - C->bbox[2 * d + 1]) >= fz) {
          std::printf("PAMR_compose_hierarchy: error ... level 1 must be one grid over the whole domain\n");
          return 0;
        }
      }
    }
  }
  if (nbase != 1) {
    std::printf("PAMR_compose_hierarchy: error ... level 1 must be one grid over the whole domain\n");
    return 0;
  }
  for (int l = 0; l < max_lev - 1; l++)
    if (nbox[l] == 0) {
      std::printf("PAMR_compose_hierarchy: error ... level %i has no grids\n", l + 2);
      return 0;
    }
  if (C->mcw[0] != C->mcw[1] || C->mcw[0] != C->mcw[2]) {
    std::printf("PAMR_compose_hierarchy: error ... the MG coarse width must be the same in every dimension\n");
    return 0;
  }
  pmg_amr *A = pmg_amr_create(C->shape, C->bbox, max_lev - 1, nbox.data(), boxes.data(), C->mcw[0], 0);
  if (!A) {
    std::printf("PAMR_compose_hierarchy: error ... building the device hierarchy failed\n");
    return 0;
  }
  C->A = A;
  C->namr = max_lev - 1;
  C->nbase = pmg_amr_base_level(A);
  C->nlev = pmg_amr_nlev(A);
  C->stream = (hipStream_t)pmg_amr_stream(A);
  for (Var &v : C->vars) {
    v.amr_field.clear();
    v.mg_field = -1;
    if (v.a.in_amrh)
      for (int k = 0; k < v.a.num_tl; k++) v.amr_field.push_back(pmg_amr_add_field(A));
    if (v.a.in_mgh && !v.a.in_amrh) v.mg_field = pmg_amr_add_field(A);
    for (int f : v.amr_field)
      if (f < 0) goto oom;
    if (v.a.in_mgh && !v.a.in_amrh && v.mg_field < 0) goto oom;
  }
  C->tf.assign(C->gfns, 0);
  for (int l = 0; l < max_lev; l++) C->t[l] = t;
  C->it = Iter();
  build_views();
  // src/pamr.c:309
  if (C->ex_on && !init_ex_mask(C->A, C->nbase, C->nlev)) {
    std::printf("PAMR_compose_hierarchy: error ... filling the excision mask failed\n");
    return 0;
  }
  return 1;
oom:
  std::printf("PAMR_compose_hierarchy: error ... out of device memory (at most %d grid functions)\n",
              PMG_MAX_EXTRA);
  pmg_amr_destroy(A);
  C->A = nullptr;
  return 0;
}

int PAMR_build_mgh(int min_lev, int max_lev, int tl) {
  if (!need_ctx("PAMR_build_mgh")) return 0;
  if (C->mgh) {
    std::printf("PAMR_build_mgh: error ... MG hierarhcy already exists\n");
    return 0;
  }
  if (!C->A) {
    std::printf("PAMR_build_mgh: error ... AMR hierarchy does not exist\n");
    return 0;
  }
  if (min_lev < 1 || min_lev > C->namr + 1 || max_lev != C->namr + 1) {
    std::printf("PAMR_build_mgh: error ... only min_lev in 1..%i and max_lev = %i (the finest AMR level) are "
                "supported\n", C->namr + 1, C->namr + 1);
    return 0;
  }
  if (tl < 1) {
    std::printf("PAMR_build_mgh: error ... invalid time level %i\n", tl);
    return 0;
  }
  // from AMR level 1 the MG levels are A's own; from a refined level an
  // island hierarchy over A's levels, built once and kept (src/gh.c:1222-1365)
  pmg_amr *M = C->A;
  if (min_lev > 1) {
    M = nullptr;
    for (auto &sub : C->subs)
      if (sub.first == min_lev) M = sub.second;
    if (!M) {
      M = pmg_amr_create_mgh(C->A, min_lev);
      if (!M) {
        std::printf("PAMR_build_mgh: error ... building the MG hierarchy from level %i failed\n", min_lev);
        return 0;
      }
      C->subs.emplace_back(min_lev, M);
    }
  }
  C->M = M;
  C->mg_min = min_lev;
  C->mgh = true;
  C->mgh_tl = tl;
  const int mb = mg_nbase(), ml = mg_nlev();
  // freshly allocated MG grid functions are zero (src/gh.c:895-898,1150-1157)
  for (const Var &v : C->vars) {
    if (!v.a.in_mgh) continue;
    const int f = v.a.in_amrh ? v.amr_field[std::min<size_t>(tl, v.amr_field.size()) - 1] : v.mg_field;
    const int Ltop = v.a.in_amrh ? mb - 1 : ml;
    for (int L = 1; L <= Ltop; L++) {
      double *p = pmg_amr_field(M, L, 0, f);
      if (hipMemsetAsync(p, 0, pmg_amr_pool_points(M, L) * sizeof(double), C->stream) != hipSuccess) return 0;
    }
  }
  build_views();
  // the excision mask of the new MG levels (src/gh.c:1322)
  if (C->ex_on && mb > 1 && !init_ex_mask(M, 1, mb - 1)) {
    std::printf("PAMR_build_mgh: error ... filling the excision mask failed\n");
    C->mgh = false;
    C->M = nullptr;
    build_views();
    return 0;
  }
  // initial injection into each new MG level (src/gh.c:1237-1239,1334)
  PAMR_thaw_tf_bits();
  PAMR_set_tf_bits(1, PAMR_MGH, PAMR_TF_MGH_INIT);
  PAMR_freeze_tf_bits();
  for (int L = mb; L > 1; L--)
    if (!PAMR_inject(L, 1, PAMR_MGH)) {
      PAMR_thaw_tf_bits();
      return 0;
    }
  PAMR_thaw_tf_bits();
  C->it = Iter();
  return ml;
}

int PAMR_destroy_mgh(void) {
  if (!need_ctx("PAMR_destroy_mgh")) return 0;
  if (!C->mgh) return 1;
  if (C->stream) (void)hipStreamSynchronize(C->stream);
  C->mgh = false;
  C->M = nullptr;
  C->comm_coarsest = 0;
  C->it = Iter();
  build_views();
  return 1;
}

int PAMR_get_max_lev(int hier) {
  if (!need_ctx("PAMR_get_max_lev")) return 0;
  if (hier == PAMR_AMRH) {
    if (!C->A) { std::printf("PAMR_get_max_lev: error ... no AMR hierarchy\n"); return 0; }
    return C->namr + 1;
  }
  if (hier == PAMR_MGH) {
    if (!C->mgh) { std::printf("PAMR_get_max_lev: error ... no MG hierarchy\n"); return 0; }
    return mg_nlev();
  }
  std::printf("PAMR_get_max_lev: error ... invalid hierarchy\n");
  return 0;
}

int PAMR_get_min_lev(int hier) {
  if (!need_ctx("PAMR_get_min_lev")) return 0;
  if ((hier == PAMR_AMRH && !C->A) || (hier == PAMR_MGH && !C->mgh) || (hier != PAMR_AMRH && hier != PAMR_MGH)) {
    std::printf("PAMR_get_min_lev: error ... invalid hierarchy\n");
    return 0;
  }
  return 1;
}

// ---- excision (src/pamr.c:1983-2037) ----------------------------------------------
int PAMR_excision_on(char *ex_mask_var, void (*app_fill_ex_mask_fnc)(double *, int, int *, double *, double),
                     double excised, int initialize_now) {
  if (!C) {
    std::printf("PAMR_excision_on: error ... no current context\n");
    return 0;
  }
  Var *f = find_var(ex_mask_var);
  if (!f) {
    std::printf("PAMR_excision_on: error ... no variable %s\n", ex_mask_var ? ex_mask_var : "(null)");
    return 0;
  }
  if (!f->a.in_amrh || !f->a.in_mgh) {
    std::printf("PAMR_excision_on: error ... variable %s must be defined in both the MG and AMR hierarchies\n",
                ex_mask_var);
    return 0;
  }
  if (f->a.num_tl != 1) {
    std::printf("PAMR_excision_on: error ... variable %s can only have 1 time level\n", ex_mask_var);
    return 0;
  }
  C->ex_on = true;
  C->excised = excised;
  C->ex_var = f->name;
  C->ex_fill = app_fill_ex_mask_fnc;
  if (initialize_now && C->A) {
    if (!init_ex_mask(C->A, C->nbase, C->nlev) || (C->mgh && mg_nbase() > 1 && !init_ex_mask(hobj(PAMR_MGH), 1, mg_nbase() - 1))) {
      std::printf("PAMR_excision_on: error ... filling the excision mask failed\n");
      return 0;
    }
  }
  return 1;
}

void PAMR_excision_off(void) {
  if (!C) {
    std::printf("PAMR_excision_off: error ... no current context\n");
    return;
  }
  C->ex_on = false;
}

void PAMR_set_time(int lev, double t) {
  if (need_ctx("PAMR_set_time") && lev >= 1 && lev <= kMaxLevs) C->t[lev - 1] = t;
}
double PAMR_get_time(int lev) {
  if (!need_ctx("PAMR_get_time") || lev < 1 || lev > kMaxLevs) return 0;
  return C->t[lev - 1];
}

// ---- transfer bits ----------------------------------------------------------
void PAMR_freeze_tf_bits(void) {
  if (need_ctx("PAMR_freeze_tf_bits")) C->frozen = true;
}
void PAMR_thaw_tf_bits(void) {
  if (need_ctx("PAMR_thaw_tf_bits")) C->frozen = false;
}
void PAMR_clear_tf_bits(void) {
  if (!need_ctx("PAMR_clear_tf_bits")) return;
  if (!C->A) { std::printf("PAMR_clear_tf_bits: error ... no hierarchy\n"); return; }
  std::fill(C->tf.begin(), C->tf.end(), 0);
}
void PAMR_set_tf_bit(int gf, int val) {
  if (!need_ctx("PAMR_set_tf_bit")) return;
  if (!C->A) { std::printf("PAMR_set_tf_bits: error ... no hierarchy\n"); return; }
  if (gf < 1 || gf > C->gfns) { std::printf("PAMR_set_tf_bits: error ... gf=%i out of range\n", gf); return; }
  PAMR_freeze_tf_bits();
  C->tf[gf - 1] = val;
}

// src/pamr.c:1843-1903
void PAMR_set_tf_bits(int tl, int hierarchy, int stage) {
  if (!need_ctx("PAMR_set_tf_bits")) return;
  if (!C->A) { std::printf("PAMR_set_tf_bits: error ... no hierarchy\n"); return; }
  if (C->frozen) return;
  PAMR_clear_tf_bits();
  for (const Var &v : C->vars) {
    int op = 0, gs = 1, ge = 0;
    const Attr &a = v.a;
    if (hierarchy == PAMR_AMRH && a.in_amrh) {
      if ((stage == PAMR_TF_SYNC && (op = a.amr_sync)) || (stage == PAMR_TF_COMPOSE && (op = a.regrid_transfer)) ||
          (stage == PAMR_TF_INJECT && (op = a.amr_inject)) || (stage == PAMR_TF_INTERP && (op = a.amr_interp)) ||
          (stage == PAMR_TF_BDY_INTERP && (op = a.amr_bdy_interp))) {
        if (tl > 0 && tl <= a.num_tl) {
          gs = v.sgfn + tl - 1;
          ge = gs + 1;
        } else if (tl < 0) {
          gs = v.sgfn;
          ge = gs + a.num_tl;
        }
      }
    } else if (hierarchy == PAMR_MGH && a.in_mgh) {
      if ((stage == PAMR_TF_SYNC && (op = a.mg_sync)) ||
          (stage == PAMR_TF_INJECT && (op = a.mg_inject) && !a.mg_noinj_to_amr) ||
          (stage == PAMR_TF_INJECT_TO_MG_LEV && (op = a.mg_inject)) ||
          (stage == PAMR_TF_INTERP && (op = a.mg_interp)) ||
          (stage == PAMR_TF_MGH_INIT && a.in_amrh && (op = a.mg_inject))) {
        gs = v.sgfn + (a.in_amrh ? a.num_tl : 0);
        ge = gs + 1;
      }
    }
    for (int i = gs; i < ge; i++) C->tf[i - 1] = op;
  }
}

// ---- transfers ----------------------------------------------------------------
int PAMR_sync(int l, int tl, int hierarchy, int AMR_bdy_width) {
  if (!need_ctx("PAMR_sync")) return 0;
  if (AMR_bdy_width < 0) { std::printf("PAMR_sync: error ... AMR_bdy_width must be >=0\n"); return 0; }
  const int Lp = plevel(l, hierarchy);
  if (!Lp) { std::printf("PAMR_sync: error ... level %i out of range\n", l); return 0; }
  if (AMR_bdy_width > 1) { std::printf("PAMR_sync: error ... AMR_bdy_width > 1 is not supported\n"); return 0; }
  PAMR_set_tf_bits(tl, hierarchy, PAMR_TF_SYNC);
  std::vector<int> f, ops;
  selected(f, ops);
  int rc;
  if (hierarchy == PAMR_MGH && C->comm_coarsest == Lp) {  // only the coarsest grids communicate
    if (AMR_bdy_width != 1) { std::printf("PAMR_sync: error ... coarsest-only sync needs AMR_bdy_width 1\n"); return 0; }
    rc = pmg_amr_sync_fields_coarsest(hobj(hierarchy), Lp, (int)f.size(), f.data());
  } else {
    rc = pmg_amr_sync_fields(hobj(hierarchy), Lp, AMR_bdy_width, (int)f.size(), f.data());
  }
  if (rc != 0) {
    std::printf("PAMR_sync: error ... device transfer failed\n");
    return 0;
  }
  return 1;
}

int PAMR_inject(int lf, int tl, int hierarchy) {
  if (!need_ctx("PAMR_inject")) return 0;
  const int Lp = plevel(lf, hierarchy), Lc = plevel(lf - 1, hierarchy);
  if (lf < 2 || !Lp || !Lc) { std::printf("PAMR_inject: error ... level %i out of range\n", lf); return 0; }
  if (hierarchy == PAMR_MGH && Lc < mg_nbase())  // the coarse level is not shared with the AMR hierarchy
    PAMR_set_tf_bits(tl, hierarchy, PAMR_TF_INJECT_TO_MG_LEV);
  else
    PAMR_set_tf_bits(tl, hierarchy, PAMR_TF_INJECT);
  std::vector<int> f, ops;
  selected(f, ops);
  for (int op : ops)
    if (op < 1 || op > 3) {
      std::printf("PAMR_inject: error ... injection operator %i is not supported\n", op);
      return 0;
    }
  const int mf = mask_field();
  pmg_amr *P = hobj(hierarchy);
  const int rc = mf >= 0 ? pmg_amr_inject_fields_ex(P, Lp, (int)f.size(), f.data(), ops.data(), mf, C->excised)
                         : pmg_amr_inject_fields(P, Lp, (int)f.size(), f.data(), ops.data());
  if (rc != 0) {
    std::printf("PAMR_inject: error ... device transfer failed\n");
    return 0;
  }
  return 1;
}

int PAMR_interp(int lc, int tl, int hierarchy) {
  if (!need_ctx("PAMR_interp")) return 0;
  const int Lp = plevel(lc, hierarchy), Lf = plevel(lc + 1, hierarchy);
  if (!Lp || !Lf) { std::printf("PAMR_interp: error ... level %i out of range\n", lc); return 0; }
  PAMR_set_tf_bits(tl, hierarchy, PAMR_TF_INTERP);
  std::vector<int> f, ops;
  selected(f, ops);
  for (int op : ops)
    if (op != PAMR_SECOND_ORDER && op != PAMR_FOURTH_ORDER) {
      std::printf("PAMR_interp: error ... interpolation order %i is not supported\n", op);
      return 0;
    }
  for (int ord : {PAMR_SECOND_ORDER, PAMR_FOURTH_ORDER}) {
    std::vector<int> sel;
    for (size_t q = 0; q < f.size(); q++)
      if (ops[q] == ord) sel.push_back(f[q]);
    const int mf = mask_field();
    pmg_amr *P = hobj(hierarchy);
    const int rc = mf >= 0 ? pmg_amr_interp_fields_ex(P, Lp, ord, (int)sel.size(), sel.data(), mf, C->excised)
                           : pmg_amr_interp_fields_ord(P, Lp, ord, (int)sel.size(), sel.data());
    if (rc != 0) {
      std::printf("PAMR_interp: error ... device transfer failed\n");
      return 0;
    }
  }
  return 1;
}

// ---- iterator and getters ----------------------------------------------------------
int PAMR_init_s_iter(int l, int hier, int all) {
  (void)all;  // one rank: every grid is local
  if (!need_ctx("PAMR_init_s_iter")) return 0;
  if (l < 1 || l > kMaxLevs) { std::printf("PAMR_init_s_iter: error ... invalid level %i\n", l); return 0; }
  C->it = Iter();
  if (hier == PAMR_MGH) {
    if (!C->mgh) { std::printf("PAMR_init_s_iter: error ... no MG hierarchy\n"); return 0; }
  } else if (hier == PAMR_AMRH) {
    if (!C->A) { std::printf("PAMR_init_s_iter: error ... no AMR hierarchy\n"); return 0; }
  } else {
    std::printf("PAMR_init_s_iter: error ... unknown hierarchy %i\n", hier);
    return 0;
  }
  const int Lp = plevel(l, hier);
  if (!Lp) {
    std::printf("PAMR_init_s_iter: error ... no level %i in %s hierarchy\n", l, hier == PAMR_MGH ? "MG" : "AMR");
    return 0;
  }
  C->it.hier = hier;
  C->it.L = l;
  C->it.g = views_of(hier)[Lp - 1].empty() ? -1 : 0;
  return C->it.g >= 0 ? 1 : 0;
}

int PAMR_next_g(void) {
  if (!need_ctx("PAMR_next_g")) return 0;
  if (C->it.g < 0) return 0;
  const int Lp = plevel(C->it.L, C->it.hier);
  C->it.g++;
  if (!Lp || C->it.g >= (int)views_of(C->it.hier)[Lp - 1].size()) C->it.g = -1;
  return C->it.g >= 0 ? 1 : 0;
}

int PAMR_push_iter(void) {
  if (!need_ctx("PAMR_push_iter")) return 0;
  if (C->sp == kIterStack) { std::printf("PAMR_push_iter: error ... stack is full\n"); return 0; }
  C->stack[C->sp++] = C->it;
  return 1;
}

int PAMR_pop_iter(void) {
  if (!need_ctx("PAMR_pop_iter")) return 0;
  if (C->sp == 0) { std::printf("PAMR_pop_iter: error ... stack is empty\n"); return 0; }
  C->it = C->stack[--C->sp];
  return C->it.g >= 0 ? 1 : 0;
}

int PAMR_get_g_rank(int *rank) {
  if (!cur_view("PAMR_get_g_rank")) return 0;
  *rank = 0;
  return 1;
}
int PAMR_get_g_dim(int *dim) {
  if (!cur_view("PAMR_get_g_dim")) return 0;
  *dim = 3;
  return 1;
}
int PAMR_get_g_shape(int *shape) {
  if (!cur_view("PAMR_get_g_shape")) return 0;
  double bb[6];
  return pmg_amr_grid_info(hobj(C->it.hier), plevel(C->it.L, C->it.hier), C->it.g, shape, bb) == 0;
}
int PAMR_get_g_bbox(double *bbox) {
  if (!cur_view("PAMR_get_g_bbox")) return 0;
  int shp[3];
  return pmg_amr_grid_info(hobj(C->it.hier), plevel(C->it.L, C->it.hier), C->it.g, shp, bbox) == 0;
}
int PAMR_get_g_ghost_width(int *ghost_width) {
  if (!cur_view("PAMR_get_g_ghost_width")) return 0;
  for (int i = 0; i < 6; i++) ghost_width[i] = 0;
  return 1;
}
int PAMR_get_g_t(double *t) {
  if (!cur_view("PAMR_get_g_t")) return 0;
  *t = level_time(C->it.L, C->it.hier);
  return 1;
}
int PAMR_get_g_ngfs(int *ngfs) {
  if (!cur_view("PAMR_get_g_ngfs")) return 0;
  *ngfs = C->gfns;
  return 1;
}
int PAMR_get_g_x(double **x) {
  View *w = cur_view("PAMR_get_g_x");
  if (!w) return 0;
  for (int d = 0; d < 3; d++) x[d] = w->x[d].data();
  return 1;
}
int PAMR_get_g_gfs(double **gfs) {
  View *w = cur_view("PAMR_get_g_gfs");
  if (!w) return 0;
  for (int i = 0; i < C->gfns; i++) gfs[i] = w->gfs[i];
  return 1;
}
int PAMR_get_g_coarsest(int *coarsest) {
  if (!cur_view("PAMR_get_g_coarsest")) return 0;
  *coarsest = C->it.hier == PAMR_MGH ? pmg_amr_grid_coarsest(hobj(PAMR_MGH), C->it.L, C->it.g) == 1 : 0;
  return 1;
}
int PAMR_get_g_level(int *L) {
  if (!cur_view("PAMR_get_g_level")) return 0;
  *L = C->it.L;
  return 1;
}
int PAMR_get_g_comm(int *comm) {
  View *w = cur_view("PAMR_get_g_comm");
  if (!w) return 0;
  *comm = w->comm;
  return 1;
}
int PAMR_set_g_comm(int comm) {
  View *w = cur_view("PAMR_set_g_comm");
  if (!w) return 0;
  w->comm = comm;
  return 1;
}

int PAMR_get_g_attribs(int *rank, int *dim, int *shape, double *bbox, int *ghost_width, double *t, int *ngfs,
                       double **x, double **gfs) {
  if (!cur_view("PAMR_get_g_attribs")) return 0;
  int ret = 1;
  ret *= PAMR_get_g_rank(rank);
  ret *= PAMR_get_g_dim(dim);
  ret *= PAMR_get_g_shape(shape);
  ret *= PAMR_get_g_bbox(bbox);
  ret *= PAMR_get_g_ghost_width(ghost_width);
  ret *= PAMR_get_g_t(t);
  ret *= PAMR_get_g_ngfs(ngfs);
  ret *= PAMR_get_g_x(x);
  ret *= PAMR_get_g_gfs(gfs);
  return ret;
}

// ---- grid-function output (src/pamr.c:1908-1978) ---------------------------
static int do_save_gfn(const char *who, char *name, int hier, int tl, int L, double t, char *pre_tag,
                       char *post_tag, int close) {
  if (!need_ctx(who)) return 0;
  const int gfn = PAMR_get_gfn(name, hier, tl);
  if (!gfn) return 0;
  char out_name[512];
  if (hier == PAMR_AMRH) {
    if (!C->A) { std::printf("%s: AMR hierarchy does not exist\n", who); return 0; }
    std::snprintf(out_name, sizeof out_name, "%s%s_tl%i%s_%i", pre_tag ? pre_tag : "", name, tl,
                  post_tag ? post_tag : "", 0);
  } else {
    if (!C->mgh) { std::printf("%s: MG hierarchy does not exist\n", who); return 0; }
    std::snprintf(out_name, sizeof out_name, "%s%s_MG%s_%i", pre_tag ? pre_tag : "", name,
                  post_tag ? post_tag : "", 0);
  }
  const int lmin = PAMR_get_min_lev(hier), lmax = PAMR_get_max_lev(hier);
  const int ls = L < 0 ? lmin : L, le = L < 0 ? lmax : L;
  if (ls < lmin || le > lmax) {
    std::printf("%s: error ... no level %i in the %s hierarchy\n", who, L, hier == PAMR_MGH ? "MG" : "AMR");
    return 0;
  }
  if (hipStreamSynchronize(C->stream) != hipSuccess) {
    std::printf("%s: error ... device stream failed\n", who);
    return 0;
  }
  PAMR_gft_out_bbox_fn out = gft_out_app ? gft_out_app : pgf_out_bbox;
  std::vector<double> host;
  for (int l = ls; l <= le; l++) {
    const int Lp = plevel(l, hier);
    const std::vector<View> &vs = views_of(hier)[Lp - 1];
    for (int g = 0; g < (int)vs.size(); g++) {
      const double *dev = vs[g].gfs[gfn - 1];
      if (!dev) continue;  // the grid function does not exist on this grid
      int shp[3];
      double bb[6];
      pmg_amr_grid_info(hobj(hier), Lp, g, shp, bb);
      host.resize((size_t)shp[0] * shp[1] * shp[2]);
      if (hipMemcpy(host.data(), dev, host.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        std::printf("%s: error ... device copy failed\n", who);
        return 0;
      }
      out(out_name, t == -1.0 ? level_time(l, hier) : t, shp, 3, bb, host.data());
    }
  }
  if (close) {
    if (gft_out_app) {
      if (gft_close_app) gft_close_app(out_name);
    } else {
      pgf_close(out_name);
    }
  }
  return 1;
}

int PAMR_save_gfn(char *name, int hier, int tl, int L, double t, char *pre_tag, char *post_tag) {
  return do_save_gfn("PAMR_save_gfn", name, hier, tl, L, t, pre_tag, post_tag, 0);
}

int PAMR_save_gfn_close(char *name, int hier, int tl, int L, double t, char *pre_tag, char *post_tag) {
  return do_save_gfn("PAMR_save_gfn_close", name, hier, tl, L, t, pre_tag, post_tag, 1);
}

void PAMR_set_gft_writer(PAMR_gft_out_bbox_fn out_bbox, PAMR_gft_close_fn close) {
  gft_out_app = out_bbox;
  gft_close_app = out_bbox ? close : nullptr;
}

// ===========================================================================
// AMRD multigrid driver
// ===========================================================================
void amrd_set_app_MG_hooks(amrd_real_hook app_MG_residual, amrd_real_hook app_MG_relax, amrd_void_hook app_L_op) {
  D.residual = app_MG_residual;
  D.relax = app_MG_relax;
  D.lop = app_L_op;
}

void amrd_set_app_MG_level_hooks(amrd_level_relax_hook relax, amrd_level_real_hook residual,
                                 amrd_level_void_hook lop) {
  D.lrelax = relax;
  D.lresidual = residual;
  D.llop = lop;
}


// amrd/io.c:843-937 (stage 4), AMRD's elliptic_vars_t0
int amrd_mg_def_vars_t0(const char *elliptic_vars_t0, const char *mg_fw_restr, const char *mg_interp2) {
  if (!C) {
    std::printf("amrd_mg_def_vars_t0: error ... no current context\n");
    return 0;
  }
  const std::vector<std::string> t0 = split_list(elliptic_vars_t0);
  if (!def_mg_work_vars("amrd_mg_def_vars_t0", t0, split_list(mg_fw_restr), split_list(mg_interp2))) return 0;
  D.ell_t0 = t0;
  return 1;
}

// AMRD's MG_cnst_data_vars (amrd/io.c:354-358, 1102): MG variables of the application
int amrd_mg_def_cnst_data_vars(const char *vars) {
  if (!C) {
    std::printf("amrd_mg_def_cnst_data_vars: error ... no current context\n");
    return 0;
  }
  const std::vector<std::string> v = split_list(vars);
  for (const std::string &n : v) {
    Var *f = find_var(n.c_str());
    if (!f || !f->a.in_mgh) {
      std::printf("amrd_mg_def_cnst_data_vars: error ... <%s> is not defined in the MG hierarchy\n", n.c_str());
      return 0;
    }
  }
  D.cnst = v;
  return 1;
}

void amrd_set_is_t0(int is_t0) { D.is_t0 = is_t0 ? 1 : 0; }

// amrd/amrd.c:109-112
void amrd_set_elliptic_vars_t0_init(amrd_void_hook app_elliptic_vars_t0_init) {
  D.t0_init = app_elliptic_vars_t0_init;
}

// the app_t0_cnst_data argument of amrd() (amrd/amrd.c:21,51)
void amrd_set_app_t0_cnst_data(amrd_void_hook app_t0_cnst_data) { D.t0_cnst = app_t0_cnst_data; }

// amrd/io.c:843-937 (stage 1) and 988-1005
int amrd_mg_def_vars(const char *elliptic_vars, const char *mg_fw_restr, const char *mg_interp2, int do_ex) {
  if (!C) {
    std::printf("amrd_mg_def_vars: error ... no current context\n");
    return 0;
  }
  const std::vector<std::string> ell = split_list(elliptic_vars), fw = split_list(mg_fw_restr),
                                 i2 = split_list(mg_interp2);
  int pbt[6] = {PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN};
  if (!def_mg_work_vars("amrd_mg_def_vars", ell, fw, i2)) return 0;
  if (!PAMR_def_var_full((char *)"cmask", 1, 1, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC,
                         PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_NO_TRANSFER, pbt))
    return 0;
  if (do_ex &&
      !PAMR_def_var_full((char *)"chr", 1, 1, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP, PAMR_NO_SYNC,
                         PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_NO_TRANSFER, pbt))
    return 0;
  D.ell = ell;
  D.do_ex = do_ex;
  return 1;
}

void amrd_mg_set_params(int pre_swp, int pst_swp, int min_iter, int max_iter, double tol, double crtol,
                        int max_citer, double w0_r, double w0_i) {
  D.pre = pre_swp;
  D.pst = pst_swp;
  D.min_iter = min_iter;
  D.max_iter = max_iter;
  D.tol = tol;
  D.crtol = crtol;
  D.max_citer = max_citer;
  D.w0_r = w0_r;
  D.w0_i = w0_i;
}

void amrd_mg_set_trace(int MG_trace, int MG_DV_trace, double MG_DV_trace_t_on, double MG_DV_trace_t_off) {
  D.trace = MG_trace;
  D.dv_trace = MG_DV_trace;
  D.dv_t_on = MG_DV_trace_t_on;
  D.dv_t_off = MG_DV_trace_t_off;
}

int amrd_mg_def_dump_vars(const char *hyperbolic_vars, const char *MGH_work_vars, const char *AMRH_work_in_MGH_vars) {
  D.hyp = split_list(hyperbolic_vars);
  D.mgh_work = split_list(MGH_work_vars);
  D.amrh_work_mg = split_list(AMRH_work_in_MGH_vars);
  return 1;
}

// amrd/mg.c:707-766
void mg_dump(int L, char *tag, int giter, int which) {
  if (!D.dv_trace || !C) return;
  const double t_lf = PAMR_get_time(PAMR_get_max_lev(PAMR_AMRH));
  if (D.dv_t_on > t_lf || D.dv_t_off < t_lf) return;
  if (giter == 0 && D.dv_pgiter != giter) D.dv_lcount++;
  D.dv_pgiter = giter;
  const double t = D.dv_lcount + giter / (D.max_iter + 2.0);
  std::vector<std::string> ell = D.ell;
  if (D.is_t0) ell.insert(ell.end(), D.ell_t0.begin(), D.ell_t0.end());
  auto save = [&](const std::string &v) {
    PAMR_save_gfn((char *)v.c_str(), PAMR_MGH, 0, L, t, tag, (char *)"");
  };
  for (const std::string &v : ell) save(v);
  if (which < 2) {
    for (const std::string &v : ell) {
      save(v + "_lop");
      save(v + "_rhs");
      save(v + "_res");
    }
    if (which < 1) {
      for (const std::string &v : D.hyp) save(v);
      save("cmask");
      for (const std::string &v : D.mgh_work) save(v);
      for (const std::string &v : D.amrh_work_mg) save(v);
    }
  }
}

// amrd/mg.c:84-201
int amrd_def_tre_vars(const char *tre_vars) {
  if (!C) {
    std::printf("amrd_def_tre_vars: error ... no current context\n");
    return 0;
  }
  int pbt[6] = {PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN};
  for (const std::string &v : split_list(tre_vars)) {
    const std::string name = v + "_tre";
    const int in_mgh = D.ntre == 0 ? 1 : 0;  // the first one carries the MG estimate (amrd/io.c:823-827)
    if (!PAMR_def_var_full((char *)name.c_str(), 1, in_mgh, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP,
                           PAMR_NO_SYNC, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 1, PAMR_SECOND_ORDER, pbt))
      return 0;
    if (in_mgh) D.tre = name;
    D.ntre++;
  }
  return 1;
}

int amrd_mg_def_extrap_vars(const char *amr_interp2, const char *amr_transfer2) {
  if (!C) {
    std::printf("amrd_mg_def_extrap_vars: error ... no current context\n");
    return 0;
  }
  if (D.ell.empty()) {
    std::printf("amrd_mg_def_extrap_vars: error ... no elliptic variables (amrd_mg_def_vars)\n");
    return 0;
  }
  const std::vector<std::string> i2 = split_list(amr_interp2), t2 = split_list(amr_transfer2);
  int pbt[6] = {PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN, PAMR_UNKNOWN};
  for (const std::string &v : D.ell) {
    // amrd/io.c:898-934: the attributes left by the preceding <v>_rhs, then
    // AMR-only with one time level
    if (!PAMR_def_var_full((char *)(v + "_brs").c_str(), 1, 0, 1, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_INTERP,
                           PAMR_NO_SYNC, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 0, PAMR_NO_TRANSFER, pbt))
      return 0;
    for (const char *suf : {"_extrap_tm1", "_extrap_tm2"}) {
      const std::string name = v + suf;
      int amr_interp = PAMR_FOURTH_ORDER, amr_bdy_interp = PAMR_NO_INTERP, regrid = PAMR_FOURTH_ORDER;
      if (in_list(name, i2)) amr_interp = amr_bdy_interp = PAMR_SECOND_ORDER;
      if (in_list(name, t2)) regrid = PAMR_SECOND_ORDER;
      if (!PAMR_def_var_full((char *)name.c_str(), 1, 0, 1, PAMR_STRAIGHT_INJECT, amr_interp, amr_bdy_interp,
                             PAMR_NO_SYNC, PAMR_NO_INJECT, PAMR_NO_INTERP, PAMR_NO_SYNC, 1, regrid, pbt))
        return 0;
    }
  }
  return 1;
}


int amrd_mg_extrap(int lev, int method, double tnp1, double dt, double tnp1_lm1, double dt_lm1) {
  if (!C || !C->A) {
    std::printf("amrd_mg_extrap: error ... no AMR hierarchy\n");
    return 0;
  }
  const int Lp = plevel(lev, PAMR_AMRH);
  if (!Lp) {
    std::printf("amrd_mg_extrap: error ... level %i out of range\n", lev);
    return 0;
  }
  if (method != 0 && method != 1) {
    std::printf("evolve: invalid MG_extrap_method\n");
    return 0;
  }
  std::vector<ExtrapGfns> E;
  if (!extrap_gfns(E, "amrd_mg_extrap")) return 0;
  if (lev <= 1) return 1;  // amrd/evolve.c:349
  const double f1 = (tnp1 - (tnp1_lm1 - 2 * dt_lm1)) / dt_lm1;
  const double f2 = -(tnp1 - (tnp1_lm1 - dt_lm1)) / dt_lm1;
  for (const ExtrapGfns &e : E) {
    long long n;
    double *v1 = pool(Lp, e.v1, n, PAMR_AMRH), *v2 = pool(Lp, e.v2, n, PAMR_AMRH), *tm1 = pool(Lp, e.tm1, n, PAMR_AMRH), *tm2 = pool(Lp, e.tm2, n, PAMR_AMRH);
    if (!v1 || !v2 || !tm1 || !tm2) return 0;
    if (method == 0)
      hipLaunchKernelGGL(k_extrap_etm1, dim3(nblocks(n)), dim3(kNT), 0, C->stream, v1, tm1, tm2, f1, f2, n);
    else
      hipLaunchKernelGGL(k_extrap_tn, dim3(nblocks(n)), dim3(kNT), 0, C->stream, v1, v2, tm1, tm2, dt / dt_lm1, n);
    if (hipGetLastError() != hipSuccess) return 0;
  }
  return 1;
}

int amrd_set_mg_brs_vars(int L1, int L2) {
  if (!C || !C->A) {
    std::printf("amrd_set_mg_brs_vars: error ... no AMR hierarchy\n");
    return 0;
  }
  std::vector<ExtrapGfns> E;
  if (!extrap_gfns(E, "amrd_set_mg_brs_vars")) return 0;
  for (int lev = L1; lev <= L2; lev++) {
    const int Lp = plevel(lev, PAMR_AMRH);
    if (!Lp) continue;
    for (const ExtrapGfns &e : E) {
      long long n;
      double *brs = pool(Lp, e.brs, n, PAMR_AMRH), *v2 = pool(Lp, e.v2, n, PAMR_AMRH);
      if (!brs || !v2 || hipMemcpyAsync(brs, v2, n * sizeof(double), hipMemcpyDeviceToDevice, C->stream) != hipSuccess)
        return 0;
    }
  }
  return 1;
}

int amrd_update_mg_extrap_vars(int L1, int Lf, double eps_c) {
  if (!C || !C->A) {
    std::printf("amrd_update_mg_extrap_vars: error ... no AMR hierarchy\n");
    return 0;
  }
  std::vector<ExtrapGfns> E;
  if (!extrap_gfns(E, "amrd_update_mg_extrap_vars")) return 0;
  int net_rho = 1;
  if (L1 > 1) {  // amrd/evolve.c:1671-1676
    net_rho = C->rho_tm[L1 - 1];
    L1++;
  }
  for (int lev = L1; lev <= Lf; lev++) {
    const int Lp = plevel(lev, PAMR_AMRH);
    if (Lp) {
      const int eps = lev > 1 && eps_c != 0;
      const double f = eps ? 1.0 - eps_c / ((double)net_rho) : 0.0;
      for (const ExtrapGfns &e : E) {
        long long n;
        double *tm1 = pool(Lp, e.tm1, n, PAMR_AMRH), *tm2 = pool(Lp, e.tm2, n, PAMR_AMRH), *v2 = pool(Lp, e.v2, n, PAMR_AMRH),
               *brs = pool(Lp, e.brs, n, PAMR_AMRH);
        if (!tm1 || !tm2 || !v2 || !brs) return 0;
        hipLaunchKernelGGL(k_update_extrap, dim3(nblocks(n)), dim3(kNT), 0, C->stream, tm1, tm2, v2, brs, eps, f, n);
        if (hipGetLastError() != hipSuccess) return 0;
      }
    }
    net_rho *= C->rho_tm[L1 - 1];  // amrd/evolve.c:1700: the factor of level L1, every level
  }
  return 1;
}

double amrd_vcycle(int giter) { return amrd_vcycle_tre(giter, 0); }

double amrd_vcycle_tre(int giter, int save_tre) {
  if (!C || !C->mgh) {
    std::printf("amrd_vcycle: error ... no MG hierarchy\n");
    return -1.0;
  }
  if ((!D.relax && !D.lrelax) || (!D.residual && !D.lresidual) || (!D.lop && !D.llop)) {
    std::printf("amrd_vcycle: error ... MG hooks not set\n");
    return -1.0;
  }
  std::vector<EllGfns> E;
  int cmask = 0;
  if (!ell_gfns(E, cmask)) {
    std::printf("amrd_vcycle: error ... MG variables not defined (amrd_mg_def_vars)\n");
    return -1.0;
  }
  const int Lf = PAMR_get_max_lev(PAMR_MGH);
  double res[kMaxLevs + 2];
  for (double &r : res) r = 0.0;
  int iter = 0;
  if (!clear_mg_vars(Lf, E)) return -1.0;
  for (int L = Lf; L > 0; L--) {
    if (L < Lf) {
      res[L + 1] = residual_level(L + 1);
      set_res_sync(E);
      if (!PAMR_sync(L + 1, 0, PAMR_MGH, 0)) return -1.0;
      PAMR_thaw_tf_bits();
      if (!clear_mg_vars(L, E)) return -1.0;
      if (!PAMR_inject(L + 1, 0, PAMR_MGH)) return -1.0;
      if (!PAMR_sync(L, 0, PAMR_MGH, 0)) return -1.0;
      if (!set_cmask(L, 1, cmask)) return -1.0;
      lop_level(L);
      if (!compute_mg_rhs(L, E, cmask, save_tre)) return -1.0;
    }
    if (!set_cmask(L, 0, cmask)) return -1.0;
    mg_dump(L, (char *)"before_preswp_", giter, 0);
    for (iter = 0; iter < D.pre; iter++) {
      res[L] = relax_level(L, 0);
      if (D.trace > 3) { trace_indent(L, Lf); std::printf("(L%i(%i):%12.8e)\n", L, iter, res[L]); }
      if (!PAMR_sync(L, 0, PAMR_MGH, iter == 0 ? 0 : 1)) return -1.0;
    }
    if (D.trace > 2) { trace_indent(L, Lf); std::printf("L%i(%i):%12.8e\n", L, iter, res[L]); }
    mg_dump(L, (char *)"after_preswp_", giter, 0);
  }
  for (int L = 1; L <= Lf; L++) {
    const double ctol = res[L] * D.crtol;
    if (L > 1 && !compute_cgc(L, E)) return -1.0;
    mg_dump(L, (char *)"after_cgc_", giter, 1);
    for (iter = 0; iter < D.pst; iter++) {
      res[L] = relax_level(L, 0);
      if (D.trace > 3) { trace_indent(L, Lf); std::printf("(L%i(%i):%12.8e)\n", L, iter, res[L]); }
      if (!PAMR_sync(L, 0, PAMR_MGH, iter == 0 ? 0 : 1)) return -1.0;
    }
    if (is_coarsest(L)) {  // amrd/mg.c:172-190
      res[L] = ctol + 1;
      set_comm(L, true);
      while (iter < D.max_citer && res[L] > ctol) {
        res[L] = relax_level(L, 1);
        if (!PAMR_sync(L, 0, PAMR_MGH, 1)) return -1.0;
        iter++;
      }
      if (iter == D.max_citer && D.cwarn) {
        std::printf("WARNING ... failed to solve elliptics on coarsest grid to within %12.8e in %i iterations; "
                    "res=%12.8e\n", ctol, D.max_citer, res[L]);
        D.cwarn--;
      }
      set_comm(L, false);
      D.last_citer = iter;
    }
    if (L == Lf) res[Lf] = residual_level(L);
    if (D.trace > 2) { trace_indent(L, Lf); std::printf("L%i(%i):%12.8e\n", L, iter, res[L]); }
    mg_dump(L, (char *)"after_pstswp_", giter, 1);
  }
  for (int L = 1; L <= Lf; L++) D.last[L - 1] = res[L];
  D.last_nlev = Lf;
  return res[Lf];
}

// tstep's iteration of the evolution and the elliptic equations on level lev
// (amrd/evolve.c:449-500), without the dissipation and i-1 boundary steps
int amrd_evo_iterate(int lev, amrd_evolve_hook app_evolve, amrd_real_hook app_evo_residual, int evo_min_iter,
                     int evo_max_iter, double evo_tol, int evo_ssc, double *evo_res_out, double *mg_res_out) {
  if (!C || !C->A || !app_evolve || !app_evo_residual) {
    std::printf("amrd_evo_iterate: error ... no AMR hierarchy or hooks\n");
    return -1;
  }
  const int Lf = PAMR_get_max_lev(PAMR_AMRH);
  const bool mg = lev == Lf && !D.ell.empty();
  double mg_res = 0.0;
  if (mg) {
    mg_res = D.tol + 1;
    if (!PAMR_build_mgh(Lf, Lf, 1)) {
      std::printf("tstep: PAMR_build_mgh() failed\n");
      return -1;
    }
  }
  int iter = 0;
  double evo_res = evo_tol + 1;
  while ((iter < evo_min_iter || evo_res > evo_tol || mg_res > D.tol) && iter < evo_max_iter) {
    if (evo_res > evo_tol || !evo_ssc || iter < evo_min_iter) {
      int valid = PAMR_init_s_iter(lev, PAMR_AMRH, 0);  // call_app_1int (amrd/util.c)
      while (valid) {
        app_evolve(iter + 1);
        valid = PAMR_next_g();
      }
      if (!PAMR_sync(lev, 1, PAMR_AMRH, iter == 0 ? 0 : 1)) return -1;
      double r = 0.0;
      bool first = true;
      valid = PAMR_init_s_iter(lev, PAMR_AMRH, 0);  // call_rr_app, max over the grids
      while (valid) {
        const double g = app_evo_residual();
        r = first ? g : std::max(g, r);
        first = false;
        valid = PAMR_next_g();
      }
      evo_res = r;
    }
    if (mg && (!evo_ssc || mg_res > D.tol || iter < evo_min_iter)) {
      mg_res = amrd_vcycle_tre(iter, 0);
      if (mg_res < 0) {
        PAMR_destroy_mgh();
        return -1;
      }
    }
    iter++;
  }
  if (iter == evo_max_iter && (evo_res > evo_tol || mg_res > D.tol))
    std::printf("\nWARNING ... failed to solve evolution (mg) equations to within %lf (%lf) in %i iterations "
                "(lev=%i): res=%12.8e (%12.8e)\n", evo_tol, D.tol, evo_max_iter, lev, evo_res, mg_res);
  if (mg) PAMR_destroy_mgh();
  if (evo_res_out) *evo_res_out = evo_res;
  if (mg_res_out) *mg_res_out = mg_res;
  return iter;
}

int amrd_solve_elliptics(int L, int tl, double *res_out) { return amrd_solve_elliptics_tre(L, tl, 0, res_out); }

// amrd/mg.c:20-79 (no t0 initial-data or constraint hooks)
int amrd_solve_elliptics_tre(int L, int tl, int save_tre, double *res_out) {
  if (!C || !C->A) {
    std::printf("amrd_solve_elliptics: error ... no AMR hierarchy\n");
    return -1;
  }
  const int Lf_amr = PAMR_get_max_lev(PAMR_AMRH);
  if ((D.ell.empty() && !D.is_t0) || (D.ell.empty() && D.ell_t0.empty() && D.is_t0)) return 0;  // mg.c:28-29
  if (!PAMR_build_mgh(L, Lf_amr, tl)) {
    std::printf("solve_elliptics: PAMR_build_mgh() failed\n");
    return -1;
  }
  const int Lf = PAMR_get_max_lev(PAMR_MGH);
  if (D.trace) {  // mg.c:35-41
    std::printf("\n\n");
    for (int l = 0; l < L - 1; l++) std::printf(" ");
    std::printf("MG L = %i - %i, MG Lf=%i, tl=%i ...", L, Lf_amr, Lf, tl);
  }
  if (D.trace > 1) std::printf("\n");
  if (D.t0_init && D.is_t0 && !t0_pass(D.t0_init, D.ell_t0, Lf)) {  // mg.c:45-52
    PAMR_destroy_mgh();
    return -1;
  }
  double res = D.tol + 1;
  int it = 0;
  while ((it < D.min_iter || res > D.tol) && it < D.max_iter) {
    res = amrd_vcycle_tre(it, save_tre);
    if (res < 0) {
      PAMR_destroy_mgh();
      return -1;
    }
    if (!D.cnst.empty() && D.is_t0 && (!D.t0_cnst || !t0_pass(D.t0_cnst, D.cnst, Lf))) {  // mg.c:60-68
      PAMR_destroy_mgh();
      return -1;
    }
    it++;
    if (D.trace > 1) std::printf("iter %i\t\tres(Lf)=%12.8e\n", it, res);
  }
  if (it == D.max_iter && res > D.tol)
    std::printf("\nWARNING ... failed to solve elliptics to within %12.8e in %i iterations; res=%12.8e\n", D.tol,
                D.max_iter, res);
  else if (D.trace)
    std::printf("  iters=%i\n", it);
  std::fflush(stdout);
  PAMR_destroy_mgh();
  if (res_out) *res_out = res;
  return it;
}

// The reference's own names and signatures (amrd/mg.h:32,37), so that AMRD
// code calling them (amrd/evolve.c:285,483,980,1156,1192,1257,1294) links
// unchanged.  A failed solve stops the program as AMRD_stop does
// (amrd/util.c:27-34, amrd/mg.c:31): message, context freed, exit(1).
void solve_elliptics(int L, int tl, int save_tre) {
  if (amrd_solve_elliptics_tre(L, tl, save_tre, nullptr) >= 0) return;
  std::printf("\nsolve_elliptics: PAMR_build_mgh() failed\n\nStopping.\n");
  std::fflush(stdout);
  if (C) PAMR_free_context(1);
  std::exit(1);
}

double vcycle(int giter, int save_tre) { return amrd_vcycle_tre(giter, save_tre); }

int amrd_mg_last_norms(double *lev_norms, int *citer) {
  if (!D.last_nlev) return 0;
  if (lev_norms)
    for (int L = 0; L < D.last_nlev; L++) lev_norms[L] = D.last[L];
  if (citer) *citer = D.last_citer;
  return 1;
}

void *amrd_gpu_stream(void) { return (C && C->A) ? (void *)C->stream : nullptr; }

}  // extern "C"
