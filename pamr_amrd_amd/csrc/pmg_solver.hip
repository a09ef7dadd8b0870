// pmg_solver.hip -- the uniform-hierarchy FAS multigrid solver on one GPU.
//
// A pmg_solver is the device-resident counterpart of an AMRD MG hierarchy
// built over a single base grid (PAMR_build_mgh, src/gh.c:1222-1365): every
// MG level is one grid, coarsened by 2 while calc_island_no allows it
// (src/gh.c:1427-1467).  pmg_vcycle executes amrd/mg.c:84-201 with the
// plain hooks on the device; the whole cycle is enqueued on one stream with
// no host synchronisation (the coarsest-grid loop and its stopping test run
// inside one workgroup), so it can be captured once in a hipGraph and
// replayed.  Mixed precision (FP32 inner V-cycles inside FP64 defect
// correction) reuses the same kernels instantiated for float.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pmg_kernels.h"
#include "pamr_mg_gpu.h"

extern "C" int pmg_net_reserve(pmg_net *N, long long points);
extern "C" int pmg_net_correct(pmg_net *N, const float *x, float scale, double *u, double coef, int nx,
                               int ny, int nz, void *stream);
extern "C" int pmg_net_recorrect(pmg_net *N, double *u, double coef, int nx, int ny, int nz, void *stream);

using namespace pmg;

namespace {

// ---------------------------------------------------------------------------
// set_cmask_bdy / set_cmask_child for one grid (amrd/mg.c:362-488): the ON
// index box (0-based, inclusive).  ok = false when the box is empty.
// ---------------------------------------------------------------------------
struct Box {
  int lo[3], hi[3];
  bool ok;
};

Box cmask_bdy_box(const int *shape, const double *gbbox, const double *base, const double *gdx) {
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = 0;
    b.hi[d] = shape[d] - 1;
    if ((gbbox[2 * d] - base[2 * d]) > gdx[d] / 2) b.lo[d]++;
    if ((base[2 * d + 1] - gbbox[2 * d + 1]) > gdx[d] / 2) b.hi[d]--;
  }
  b.ok = true;
  for (int d = 0; d < 3; d++)
    if (b.lo[d] < 0 || b.lo[d] > b.hi[d] || b.hi[d] >= shape[d]) b.ok = false;
  return b;
}

Box cmask_child_box(const int *shape, const double *gbbox, const double *base, const double *gdx,
                    const double *cb) {
  Box b;
  for (int d = 0; d < 3; d++) {
    int s = (int)std::max((cb[2 * d] - gbbox[2 * d]) / gdx[d] + 0.5, 0.0);
    int e = (int)std::min((cb[2 * d + 1] - gbbox[2 * d]) / gdx[d] + 0.5, (double)(shape[d] - 1));
    if (s > 0 || (gbbox[2 * d] - base[2 * d]) > gdx[d] / 2) s++;
    if (e < (shape[d] - 1) || (base[2 * d + 1] - gbbox[2 * d + 1]) > gdx[d] / 2) e--;
    b.lo[d] = s;
    b.hi[d] = e;
  }
  b.ok = true;
  for (int d = 0; d < 3; d++)
    if (b.lo[d] < 0 || b.lo[d] > b.hi[d] || b.hi[d] >= shape[d]) b.ok = false;
  return b;
}

bool is_pow2(double v) {
  int e;
  double m = std::frexp(v, &e);
  return m == 0.5 && v > 1e-300 && v < 1e300;
}

template <typename T>
Spacing<T> make_spacing(const double *hdx) {
  Spacing<T> sp;
  T dx = (T)hdx[0], dy = (T)hdx[1], dz = (T)hdx[2];
  sp.dx2 = dx * dx;
  sp.dy2 = dy * dy;
  sp.dz2 = dz * dz;
  sp.idx2 = T(1) / sp.dx2;
  sp.idy2 = T(1) / sp.dy2;
  sp.idz2 = T(1) / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  sp.pow2 = is_pow2((double)sp.dx2) && is_pow2((double)sp.dy2) && is_pow2((double)sp.dz2);
  return sp;
}

inline unsigned nblk256(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------------------
// solver state
// ---------------------------------------------------------------------------
template <typename T>
struct LevelArrays {
  T *u = nullptr, *ua = nullptr, *rho = nullptr, *rhs = nullptr, *res = nullptr, *lop = nullptr;
  T *a = nullptr, *c = nullptr;
};

struct Level {
  int n[3];
  size_t size;
  double hdx[3], gdx[3];
  Spacing<double> sp64;
  Spacing<float> sp32;
  Box rbox;  // relax/residual box: cmask_bdy box clipped to the interior
  Box cbox;  // child mask box (levels below the finest)
  LevelArrays<double> d;
  LevelArrays<float> f;
  int cur[2] = {0, 0};      // ping-pong buffer holding u (0: u, 1: ua), FP64 / FP32,
                            // tracked while a cycle is enqueued
  uint8_t *chr = nullptr;   // excision mask, 1 = excised (sweeps, hooks)
  double *chrd = nullptr;   // the chr grid function itself (m_util_r8 _ex transfers)
};

struct pmg_solver {
  int nlev = 0, op = 0, do_ex = 0;
  double ex = 1.0;
  double bbox[6];
  Level lev[kMaxLev];
  int pre = 2, pst = 2, max_citer = 5000, rop_u = 3, rop_res = 3, rop_src = 3, iord = 2;
  double crtol = 1e-3, w0_r = 1.0, w0_i = 1.0;
  hipStream_t stream = nullptr;
  NormWs ws{nullptr, nullptr, 0, 0};
  double *d_norms = nullptr;  // folded norms: FP64 slots, FP32 slots, coarse-solve outputs
  int *d_citer = nullptr;     // 2 ints (FP64 cycle, FP32 cycle)
  double *d_snap = nullptr;   // saved state of level 1 for the coarsest loop's cycle watch
  double *h_pinned = nullptr; // pinned readback
  int sweep_variant = 1, zchunk = 32;
  int lv_variant[2][kMaxLev], lv_zchunk[2][kMaxLev];  // per-level sweep configuration, [0] FP64, [1] FP32
  int use_graph = 1;
  int use_bottom = 1;  // fuse the small bottom levels into one workgroup
  int bottom_max_n = 17;  // ... those with at most this many points per side
  int bottom_lds = 1;     // ... with their arrays in LDS when they fit
  int fp32_exact_bottom = 1;  // FP32 inner cycles: exact sine-transform solve of a 17^3 bottom level
  int use_fused_restrict = 2;  // finest residual + injection in one pass (1: res ring; 2, 3: restriction streamed in z)
  int rr_czchunk[2] = {8, 8};   // its coarse planes per block, [0] FP64, [1] FP32
  int use_sweep2 = 0;          // pairs of sweeps temporally blocked on large levels
                               // (measured slower than single sweeps; off by default)
  int sweep2_zchunk = 32;
  hipGraphExec_t g64 = nullptr;   // captured FP64 V-cycle
  hipGraphExec_t g32 = nullptr;   // captured mixed-precision outer iteration
  bool fp32_ready = false;
  int last_citer = 0;
  // learned mode
  pmg_net *net = nullptr;
  int net_enabled = 1;
  int net_skip = 0;          // solves left before the network is tried again
  int net_retry_period = 20; // doubles after every rejection (up to 1280), back to 20 on acceptance
  double t_net_ms = -1, t_vc_ms = -1, rho_vc = 0.1;
  int net_accepted = 0, net_rejected = 0;
  // history warm start: per-block Gram partials and their fold
  double *d_hist = nullptr;
  double *h_hist = nullptr;
  // flexible GMRES: V (m+1 vectors), Z (m), w, a zero source, reductions
  int kry_m = 0;
  double *kry = nullptr;
  double *kry_red = nullptr;
  double *kry_host = nullptr;
  hipGraphExec_t gprec = nullptr;  // captured FP32 V-cycle (the preconditioner)
  int prec_cycles = 1;             // FP32 V-cycles per preconditioner application
  // kernel timing inside V-cycles (pmg_time_kernel which 3): HIP events
  // around the finest level's FP64 sweep launches of ungraphed cycles
  hipEvent_t *tk_ev = nullptr;
  int tk_cap = 0, tk_n = 0, tk_sweeps = 0;
  // learned correction in the band next to the excised region (pmg_band.hip)
  struct Band {
    int depth = 0, core = 0, side = 0, margin = 3, ncore = 0;
    long long npts = 0;
    int *d_org = nullptr;          // crop origins, 3 ints per core
    uint8_t *d_band = nullptr;     // band mask of the finest level
    float *d_x = nullptr;          // network input crops [ncore][side^3]
    std::vector<long long> h_pts;  // linear indices of the band points (training sampler)
    pmg_net *net = nullptr;
    double ts = 1.0;               // output scale: correction = ts * ||r|| * N(r / ||r||)
    int use_dc = 0, use_fgmres = 0;
    double rho_plain = -1, t_plain = -1;   // defect correction: plain iteration (EMA)
    double rho_net = -1, t_net = -1;       // ... iterations with the correction (EMA)
    double fg_rate_plain = 0, fg_rate_net = 0;  // FGMRES: log10 reduction per ms of a whole solve
    int fg_skip = 0;
    int accepted = 0, rejected = 0;
    hipGraphExec_t gdc = nullptr;   // captured correction of the defect-correction iteration
    hipGraphExec_t gpc = nullptr;   // captured correction of the preconditioner output
    // the correction inside the FP32 V-cycle, after the finest level's
    // coarse-grid correction (instead of after the cycle): on when inner = 1
    int inner = 0;
    pmg_net *net_in = nullptr;      // its network and output scale (trained on in-cycle crops)
    double ts_in = 1.0;
    long long n_upd = 0;            // updatable finest points (the RMS norms' count)
    bool enq_inner = false;         // set while an FP32 cycle with it is enqueued
    const double *cur_norm = nullptr;  // RMS of that cycle's finest source (device)
    hipGraphExec_t g32i = nullptr;     // mixed iteration with the in-cycle correction
    hipGraphExec_t gpreci = nullptr;   // preconditioner with the in-cycle correction
    // training crops taken inside the cycle instead of correcting (pmg_band_sample_inner)
    const double *samp_ustar = nullptr;
    int *samp_org = nullptr;
    int samp_n = 0;
    float *samp_x = nullptr, *samp_t = nullptr, *samp_w = nullptr;
  } band;
};

// norm slot layout in d_norms
enum {
  NS_RELAX = 0,                 // + L-1: last relax norm per level (FP64 cycle)
  NS_RESID = kMaxLev,           // + L-1: residual norm per level
  NS_FINAL = 2 * kMaxLev,       // final residual of the cycle
  NS_CPRE = 2 * kMaxLev + 1,    // coarsest pre-sweep norm
  NS_R0 = 2 * kMaxLev + 2,      // standalone finest residual
  NS_OUTER = 2 * kMaxLev + 3,   // mixed: outer FP64 residual
  kNormSlots = 2 * kMaxLev + 8,
  NS_F32 = kNormSlots,          // FP32 inner-cycle slots start here
  NS_COARSE = 2 * kNormSlots,   // coarse solve outputs: pre64, post64, pre32, post32
  NS_PSRC = 2 * kNormSlots + 4, // RMS of the preconditioner's source (unit 2-norm Krylov vector)
  kNormTotal = 2 * kNormSlots + 8
};

namespace {

int alloc_level_arrays(pmg_solver *S, Level &Lv, bool f32) {
  const size_t n = Lv.size;
  if (!f32) {
    auto &d = Lv.d;
    for (double **p : {&d.u, &d.ua, &d.rho, &d.rhs, &d.res, &d.lop}) {
      PMG_HOST_CHECK(hipMalloc(p, n * sizeof(double)));
      PMG_HOST_CHECK(hipMemsetAsync(*p, 0, n * sizeof(double), S->stream));
    }
    if (S->op == 1) {
      for (double **p : {&d.a, &d.c}) {
        PMG_HOST_CHECK(hipMalloc(p, n * sizeof(double)));
        PMG_HOST_CHECK(hipMemsetAsync(*p, 0, n * sizeof(double), S->stream));
      }
    }
    if (S->do_ex) {
      PMG_HOST_CHECK(hipMalloc(&Lv.chr, n));
      PMG_HOST_CHECK(hipMemsetAsync(Lv.chr, 0, n, S->stream));
      PMG_HOST_CHECK(hipMalloc(&Lv.chrd, n * sizeof(double)));
    }
  } else {
    auto &f = Lv.f;
    for (float **p : {&f.u, &f.ua, &f.rho, &f.rhs, &f.res, &f.lop}) {
      PMG_HOST_CHECK(hipMalloc(p, n * sizeof(float)));
      PMG_HOST_CHECK(hipMemsetAsync(*p, 0, n * sizeof(float), S->stream));
    }
    if (S->op == 1) {
      for (float **p : {&f.a, &f.c}) {
        PMG_HOST_CHECK(hipMalloc(p, n * sizeof(float)));
      }
    }
  }
  return 0;
}

template <typename T>
void free_arrays(LevelArrays<T> &A) {
  for (T **p : {&A.u, &A.ua, &A.rho, &A.rhs, &A.res, &A.lop, &A.a, &A.c}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

// ---------------------------------------------------------------------------
// kernel launchers (templated on precision)
// ---------------------------------------------------------------------------
template <typename T>
LevelArrays<T> &arrays(Level &Lv);
template <>
LevelArrays<double> &arrays<double>(Level &Lv) { return Lv.d; }
template <>
LevelArrays<float> &arrays<float>(Level &Lv) { return Lv.f; }
template <typename T>
const Spacing<T> &spacing(const Level &Lv);
template <>
const Spacing<double> &spacing<double>(const Level &Lv) { return Lv.sp64; }
template <>
const Spacing<float> &spacing<float>(const Level &Lv) { return Lv.sp32; }

template <typename T>
int &curbuf(Level &Lv) { return Lv.cur[std::is_same<T, float>::value ? 1 : 0]; }
// the buffer holding u and its ping-pong twin
template <typename T>
T *ucur(Level &Lv) { LevelArrays<T> &A = arrays<T>(Lv); return curbuf<T>(Lv) ? A.ua : A.u; }
template <typename T>
T *uoth(Level &Lv) { LevelArrays<T> &A = arrays<T>(Lv); return curbuf<T>(Lv) ? A.u : A.ua; }

// sweep tile variants (TX, TY, NT, planes in flight, LDS ring planes)
#define PMG_SWEEP_VARIANTS(X)                                                                                   \
  X(0, 64, 8, 256, 1, 4) X(1, 64, 16, 512, 1, 4) X(2, 32, 16, 512, 1, 4) X(3, 128, 8, 512, 1, 4)             \
  X(4, 64, 32, 512, 1, 4) X(5, 128, 16, 1024, 1, 4) X(6, 128, 16, 1024, 2, 4) X(7, 64, 16, 512, 2, 4)        \
  X(8, 128, 8, 512, 2, 4) X(9, 128, 16, 1024, 2, 5) X(10, 128, 16, 1024, 1, 5) X(11, 128, 8, 512, 1, 5)
constexpr int kNumSweepVariants = 12;

template <typename T, int OP, bool EX, bool P2, bool RHS0>
int launch_sweep_v(pmg_solver *S, int variant, SweepP<T> P) {
  P.zchunk = std::max(1, std::min(P.zchunk, P.nz - 2));
  int tx = 64, ty = 8;
  switch (variant) {
#define PMG_TXY(V, TXV, TYV, NTV, PFV, RGV) case V: tx = TXV; ty = TYV; break;
    PMG_SWEEP_VARIANTS(PMG_TXY)
#undef PMG_TXY
    default: return -1;
  }
  dim3 grid(((P.nx - 2 + tx - 1) / tx) * ((P.ny - 2 + ty - 1) / ty) * ((P.nz - 2 + P.zchunk - 1) / P.zchunk));
  P.nb = (int)grid.x;
  if (P.nb > S->ws.cap) return -3;
  switch (variant) {
#define PMG_LSW(V, TXV, TYV, NTV, PFV, RGV)                                                         \
  case V:                                                                                           \
    hipLaunchKernelGGL((k_sweep<T, OP, EX, P2, RHS0, TXV, TYV, NTV, PFV, RGV>), grid, dim3(NTV), 0, \
                       S->stream, P);                                                               \
    break;
    PMG_SWEEP_VARIANTS(PMG_LSW)
#undef PMG_LSW
  }
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int OP, bool EX>
int launch_sweep(pmg_solver *S, int v, SweepP<T> P, bool p2, bool rhs0) {
  if (p2) return rhs0 ? launch_sweep_v<T, OP, EX, true, true>(S, v, P) : launch_sweep_v<T, OP, EX, true, false>(S, v, P);
  return rhs0 ? launch_sweep_v<T, OP, EX, false, true>(S, v, P) : launch_sweep_v<T, OP, EX, false, false>(S, v, P);
}

// two sweeps in one pass (temporal blocking), 64 x 16 tiles, 1024 threads,
// on levels with at least kSweep2MinN points per side
constexpr int kSweep2MinN = 65;
constexpr int kS2TX = 64, kS2TY = 16, kS2NT = 1024;

template <typename T, int OP, bool EX, bool P2, bool RHS0>
int launch_sweep2_v(pmg_solver *S, SweepP<T> P) {
  P.zchunk = std::max(1, std::min(P.zchunk, P.nz - 2));
  dim3 grid(((P.nx - 2 + kS2TX - 1) / kS2TX) * ((P.ny - 2 + kS2TY - 1) / kS2TY) *
            ((P.nz - 2 + P.zchunk - 1) / P.zchunk));
  P.nb = (int)grid.x;
  if (P.nb > S->ws.cap) return -3;
  hipLaunchKernelGGL((k_sweep2<T, OP, EX, P2, RHS0, kS2TX, kS2TY, kS2NT>), grid, dim3(kS2NT), 0, S->stream, P);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int OP, bool EX>
int launch_sweep2(pmg_solver *S, SweepP<T> P, bool p2, bool rhs0) {
  if (p2) return rhs0 ? launch_sweep2_v<T, OP, EX, true, true>(S, P) : launch_sweep2_v<T, OP, EX, true, false>(S, P);
  return rhs0 ? launch_sweep2_v<T, OP, EX, false, true>(S, P) : launch_sweep2_v<T, OP, EX, false, false>(S, P);
}

// levels where pairs of sweeps run temporally blocked
bool use_sweep2(const pmg_solver *S, const Level &Lv) {
  return S->use_sweep2 && Lv.n[0] >= kSweep2MinN && Lv.n[1] >= kSweep2MinN && Lv.n[2] >= kSweep2MinN;
}

// n red-black sweeps on level L (ping-pong u <-> ua; odd counts copy back)
template <typename T>
int sweeps(pmg_solver *S, int L, int n, int slot, bool rhs0) {
  Level &Lv = S->lev[L - 1];
  LevelArrays<T> &A = arrays<T>(Lv);
  if (n <= 0) return 0;
  if (!Lv.rbox.ok) return 0;
  SweepP<T> P;
  P.rho = A.rho; P.rhs = A.rhs; P.a = A.a; P.c = A.c; P.chr = Lv.chr;
  P.nx = Lv.n[0]; P.ny = Lv.n[1]; P.nz = Lv.n[2];
  P.bx0 = Lv.rbox.lo[0]; P.bx1 = Lv.rbox.hi[0];
  P.by0 = Lv.rbox.lo[1]; P.by1 = Lv.rbox.hi[1];
  P.bz0 = Lv.rbox.lo[2]; P.bz1 = Lv.rbox.hi[2];
  constexpr int pi = std::is_same<T, float>::value ? 1 : 0;
  P.zchunk = S->lv_zchunk[pi][L - 1];
  const int variant = S->lv_variant[pi][L - 1];
  P.sp = spacing<T>(Lv);
  P.ws = S->ws;
  P.slot = slot;
  const bool p2 = P.sp.pow2 != 0;
  const bool two = use_sweep2(S, Lv);
  int done = 0, launches = 0;
  T *const b0 = ucur<T>(Lv), *const b1 = uoth<T>(Lv);
  while (done < n) {
    P.uin = (launches % 2 == 0) ? b0 : b1;
    P.uout = (launches % 2 == 0) ? b1 : b0;
    int rc;
    if (two && n - done >= 2) {
      SweepP<T> Q = P;
      Q.zchunk = S->sweep2_zchunk;
      if (S->op == 0) rc = S->do_ex ? launch_sweep2<T, 0, true>(S, Q, p2, rhs0) : launch_sweep2<T, 0, false>(S, Q, p2, rhs0);
      else rc = S->do_ex ? launch_sweep2<T, 1, true>(S, Q, p2, rhs0) : launch_sweep2<T, 1, false>(S, Q, p2, rhs0);
      done += 2;
    } else {
      if (S->op == 0) rc = S->do_ex ? launch_sweep<T, 0, true>(S, variant, P, p2, rhs0) : launch_sweep<T, 0, false>(S, variant, P, p2, rhs0);
      else rc = S->do_ex ? launch_sweep<T, 1, true>(S, variant, P, p2, rhs0) : launch_sweep<T, 1, false>(S, variant, P, p2, rhs0);
      done += 1;
    }
    if (rc) return rc;
    launches++;
  }
  // an odd number of launches leaves u in the twin buffer: later kernels of
  // this cycle read it there, settle() moves it back at the end of the level
  curbuf<T>(Lv) ^= (launches & 1);
  return 0;
}

// u back in its primary buffer (both buffers then agree)
template <typename T>
int settle(pmg_solver *S, int L) {
  Level &Lv = S->lev[L - 1];
  LevelArrays<T> &A = arrays<T>(Lv);
  if (curbuf<T>(Lv) == 0) return 0;
  hipLaunchKernelGGL(k_copy<T>, dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, A.u, A.ua, Lv.size);
  PMG_HOST_CHECK(hipGetLastError());
  curbuf<T>(Lv) = 0;
  return 0;
}

template <typename T, int OP, bool EX, bool P2>
int launch_residual_t(pmg_solver *S, ResidP<T> P, bool rhs0, bool zero) {
  // z depth per block: long marches on big levels (HBM reuse), short ones on
  // small levels (latency: a block's planes are processed one after another)
  P.zchunk = std::max(1, std::min(std::min(RZC, std::max(2, (P.nz - 2) / 16)), P.nz - 2));
  const int nzc = (P.nz - 2 + P.zchunk - 1) / P.zchunk;
  dim3 grid(((P.nx - 2 + RTX - 1) / RTX) * ((P.ny - 2 + RTY - 1) / RTY) * nzc);
  const int nb = (int)grid.x;
  P.nb = nb;
  if (nb > S->ws.cap) return -3;
#define PMG_LR(R0, Z0, O32) hipLaunchKernelGGL((k_residual<T, OP, EX, P2, R0, Z0, O32>), grid, dim3(RTX * RTY), 0, S->stream, P)
  if (P.neg32) { PMG_LR(true, true, true); }
  else if (rhs0) { if (zero) PMG_LR(true, true, false); else PMG_LR(true, false, false); }
  else { if (zero) PMG_LR(false, true, false); else PMG_LR(false, false, false); }
#undef PMG_LR
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int residual(pmg_solver *S, int L, int slot, bool rhs0, bool zero, float *neg32 = nullptr) {
  Level &Lv = S->lev[L - 1];
  LevelArrays<T> &A = arrays<T>(Lv);
  ResidP<T> P;
  P.u = ucur<T>(Lv); P.rho = A.rho; P.rhs = A.rhs; P.a = A.a; P.c = A.c; P.chr = Lv.chr;
  P.res = A.res;
  P.neg32 = neg32;
  P.nx = Lv.n[0]; P.ny = Lv.n[1]; P.nz = Lv.n[2];
  if (Lv.rbox.ok) {
    P.bx0 = Lv.rbox.lo[0]; P.bx1 = Lv.rbox.hi[0];
    P.by0 = Lv.rbox.lo[1]; P.by1 = Lv.rbox.hi[1];
    P.bz0 = Lv.rbox.lo[2]; P.bz1 = Lv.rbox.hi[2];
  } else {
    P.bx0 = P.by0 = P.bz0 = 1;
    P.bx1 = P.by1 = P.bz1 = 0;
  }
  P.sp = spacing<T>(Lv);
  P.ws = S->ws;
  P.slot = slot;
  const bool p2 = P.sp.pow2 != 0;
#define PMG_RS(OPV, EXV) \
  (p2 ? launch_residual_t<T, OPV, EXV, true>(S, P, rhs0, zero) : launch_residual_t<T, OPV, EXV, false>(S, P, rhs0, zero))
  if (S->op == 0) return S->do_ex ? PMG_RS(0, true) : PMG_RS(0, false);
  return S->do_ex ? PMG_RS(1, true) : PMG_RS(1, false);
#undef PMG_RS
}

// restriction of one field from level Lf to Lf-1 (whole grid)
template <typename T>
int restrict_field(pmg_solver *S, int Lfine, const T *src, T *dst, T *alt, int kind, bool use_ex) {
  Level &F = S->lev[Lfine - 1], &Cl = S->lev[Lfine - 2];
  if (kind == 0) return 0;
  const size_t n = Cl.size;
  (void)use_ex;
#define PMG_RW(K) hipLaunchKernelGGL((k_restrict_whole<T, K, false>), grid3(Cl.n[0], Cl.n[1], Cl.n[2]), dim3(256), 0, S->stream, src, dst, alt, (const T *)nullptr, T(0), F.n[0], F.n[1], F.n[2], Cl.n[0], Cl.n[1], Cl.n[2])
  if (kind == 1) PMG_RW(1);
  else if (kind == 2) PMG_RW(2);
  else PMG_RW(3);
#undef PMG_RW
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

namespace {
__global__ void k_chr_bytes(uint8_t *dst, const double *chr, double ex, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = chr[q] == ex ? 1 : 0;
}

int restrict_field_ex64(pmg_solver *S, int Lfine, const double *src, double *dst, double *alt,
                        int kind) {
  Level &F = S->lev[Lfine - 1], &Cl = S->lev[Lfine - 2];
  const size_t n = Cl.size;
  const double *chrv = F.chrd;
#define PMG_RWE(K) hipLaunchKernelGGL((k_restrict_whole<double, K, true>), grid3(Cl.n[0], Cl.n[1], Cl.n[2]), dim3(256), 0, S->stream, src, dst, alt, chrv, S->ex, F.n[0], F.n[1], F.n[2], Cl.n[0], Cl.n[1], Cl.n[2])
  if (kind == 1) PMG_RWE(1);
  else if (kind == 2) PMG_RWE(2);
  else PMG_RWE(3);
#undef PMG_RWE
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int restrict_any(pmg_solver *S, int Lfine, const T *src, T *dst, T *alt, int kind) {
  if (kind == 0) return 0;
  if constexpr (std::is_same<T, double>::value) {
    if (S->do_ex) return restrict_field_ex64(S, Lfine, src, dst, alt, kind);
  }
  return restrict_field<T>(S, Lfine, src, dst, alt, kind, false);
}

template <typename T>
int lop_rhs(pmg_solver *S, int L) {
  Level &Lv = S->lev[L - 1];
  LevelArrays<T> &A = arrays<T>(Lv);
  LopRhsP<T> P;
  P.u = A.u; P.rho = A.rho; P.res = A.res; P.a = A.a; P.c = A.c; P.chr = Lv.chr;
  P.rhs = A.rhs; P.lop = A.lop;
  P.nx = Lv.n[0]; P.ny = Lv.n[1]; P.nz = Lv.n[2];
  if (Lv.cbox.ok) {
    P.cx0 = Lv.cbox.lo[0]; P.cx1 = Lv.cbox.hi[0];
    P.cy0 = Lv.cbox.lo[1]; P.cy1 = Lv.cbox.hi[1];
    P.cz0 = Lv.cbox.lo[2]; P.cz1 = Lv.cbox.hi[2];
  } else {
    P.cx0 = P.cy0 = P.cz0 = 1;
    P.cx1 = P.cy1 = P.cz1 = 0;
  }
  P.w0_r = (T)S->w0_r;
  P.sp = spacing<T>(Lv);
  const bool p2 = P.sp.pow2 != 0;
  const dim3 nb = grid3(Lv.n[0], Lv.n[1], Lv.n[2]);
#define PMG_LO(OPV, EXV, P2V) hipLaunchKernelGGL((k_lop_rhs<T, OPV, EXV, P2V>), nb, dim3(256), 0, S->stream, P)
  if (S->op == 0) {
    if (S->do_ex) { if (p2) PMG_LO(0, true, true); else PMG_LO(0, true, false); }
    else { if (p2) PMG_LO(0, false, true); else PMG_LO(0, false, false); }
  } else {
    if (S->do_ex) { if (p2) PMG_LO(1, true, true); else PMG_LO(1, true, false); }
    else { if (p2) PMG_LO(1, false, true); else PMG_LO(1, false, false); }
  }
#undef PMG_LO
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pmg_dminterp3d(double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                              int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni,
                              int nj, int nk, int si1, int sj1, int sk1, int irho, int jrho,
                              int krho, int ord, const double *chr1, double ex, int do_ex,
                              hipStream_t stream);

template <typename T>
int cgc(pmg_solver *S, int L) {
  Level &F = S->lev[L - 1], &Cl = S->lev[L - 2];
  LevelArrays<T> &AF = arrays<T>(F);
  LevelArrays<T> &AC = arrays<T>(Cl);
  if (S->iord != 2) return -4;
  bool ex = false;
  if constexpr (std::is_same<T, double>::value) ex = S->do_ex != 0;
  if (!ex) {
    hipLaunchKernelGGL(k_cgc_diff<T>, dim3(nblk256(Cl.size)), dim3(256), 0, S->stream, AC.res, AC.u, AC.lop,
                       (T)S->w0_i, Cl.size);
    hipLaunchKernelGGL(k_cgc<T>, grid3(Cl.n[0], Cl.n[1], Cl.n[2]), dim3(256), 0, S->stream, ucur<T>(F), uoth<T>(F), AC.res,
                       F.n[0], F.n[1], F.n[2], Cl.n[0], Cl.n[1], Cl.n[2], (const uint8_t *)F.chr);
    PMG_HOST_CHECK(hipGetLastError());
    return 0;
  }
  if constexpr (std::is_same<T, double>::value) {
    // excision: coarse res, dminterp3d_ex into the fine res, clear_ex, add
    hipLaunchKernelGGL(k_cgc_diff<double>, dim3(nblk256(Cl.size)), dim3(256), 0, S->stream, AC.res,
                       AC.u, AC.lop, S->w0_i, Cl.size);
    PMG_HOST_CHECK(hipGetLastError());
    int rc = pmg_dminterp3d(AC.res, AF.res, Cl.n[0], Cl.n[1], Cl.n[2], F.n[0], F.n[1], F.n[2], 1, 1, 1,
                            1, 1, 1, Cl.n[0], Cl.n[1], Cl.n[2], 1, 1, 1, 2, 2, 2, 2,
                            Cl.chrd, S->ex, 1, S->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear_ex<double>, dim3(nblk256(Cl.size)), dim3(256), 0, S->stream, AC.res,
                       Cl.chr, Cl.size);
    hipLaunchKernelGGL(k_cgc_add_ex<double>, grid3(F.n[0], F.n[1], F.n[2]), dim3(256), 0, S->stream, ucur<double>(F),
                       uoth<double>(F), AF.res, F.chr, F.n[0], F.n[1], F.n[2]);
    PMG_HOST_CHECK(hipGetLastError());
  }
  return 0;
}

template <typename T>
int coarse_solve(pmg_solver *S, int L, double *norm_pre, double *norm_post, int *citer,
                 int max_citer) {
  Level &Lv = S->lev[L - 1];
  LevelArrays<T> &A = arrays<T>(Lv);
  CoarseP<T> P;
  P.u = A.u; P.rho = A.rho; P.rhs = A.rhs; P.a = A.a; P.c = A.c; P.chr = Lv.chr;
  P.nx = Lv.n[0]; P.ny = Lv.n[1]; P.nz = Lv.n[2];
  if (Lv.rbox.ok) {
    P.bx0 = Lv.rbox.lo[0]; P.bx1 = Lv.rbox.hi[0];
    P.by0 = Lv.rbox.lo[1]; P.by1 = Lv.rbox.hi[1];
    P.bz0 = Lv.rbox.lo[2]; P.bz1 = Lv.rbox.hi[2];
  } else {
    P.bx0 = P.by0 = P.bz0 = 1;
    P.bx1 = P.by1 = P.bz1 = 0;
  }
  P.sp = spacing<T>(Lv);
  P.npre = S->pre; P.npst = S->pst; P.max_citer = max_citer; P.crtol = S->crtol;
  P.norm_pre = norm_pre; P.norm_post = norm_post; P.citer = citer;
  P.snap = L == 1 ? reinterpret_cast<T *>(S->d_snap) : nullptr;
  const int narr = (S->op == 1) ? 5 : 3;
  const size_t lds = Lv.size * narr * sizeof(T);
  const bool use_lds = lds <= 60 * 1024;
  const bool p2 = P.sp.pow2 != 0;
  constexpr int NT = 256;
#define PMG_CS(OPV, EXV, P2V, LDSV) hipLaunchKernelGGL((k_coarse_solve<T, OPV, EXV, P2V, LDSV, NT>), dim3(1), dim3(NT), LDSV ? lds : 0, S->stream, P)
#define PMG_CS2(OPV, EXV) \
  if (p2) { if (use_lds) PMG_CS(OPV, EXV, true, true); else PMG_CS(OPV, EXV, true, false); } \
  else { if (use_lds) PMG_CS(OPV, EXV, false, true); else PMG_CS(OPV, EXV, false, false); }
  if (S->op == 0) { if (S->do_ex) { PMG_CS2(0, true) } else { PMG_CS2(0, false) } }
  else { if (S->do_ex) { PMG_CS2(1, true) } else { PMG_CS2(1, false) } }
#undef PMG_CS2
#undef PMG_CS
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// one FAS V-cycle, enqueued (amrd/mg.c:84-201).  T = double is the
// reference method.  T = float is the inner cycle of the mixed-precision
// solver: its finest rho is the (negated) FP64 defect, rhs on the finest
// level is zero and coarse-level rho are zero (for a linear operator the
// coarse rho cancels between the FAS rhs and the coarse relaxation).
// ---------------------------------------------------------------------------
int fold_norms(pmg_solver *S, int slot0, int nslots) {
  hipLaunchKernelGGL(k_norm_fold, dim3(nslots), dim3(kFoldThreads), 0, S->stream, S->ws, slot0, 1,
                     S->d_norms + slot0);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// The small levels at the bottom of the V (every dimension at most
// kBottomMaxN points, below the finest level, no excision) run as one
// single-workgroup kernel; returns the top level of that range or 0.
constexpr int kBottomMaxN = 17;

int bottom_top(const pmg_solver *S) {
  if (S->do_ex || S->nlev < 3) return 0;
  int Lb = 0;
  for (int L = 1; L < S->nlev && L <= kBotMax; L++) {
    const Level &Lv = S->lev[L - 1];
    const int mx = std::min(kBottomMaxN, S->bottom_max_n);
    if (Lv.n[0] <= mx && Lv.n[1] <= mx && Lv.n[2] <= mx && Lv.rbox.ok) Lb = L;
    else break;
  }
  return Lb >= 2 ? Lb : 0;
}

// whether the LDS bottom kernel can take lds bytes of dynamic LDS: the
// device's per-workgroup LDS less the kernel's static LDS (level table,
// reduction slots), with the dynamic limit raised to match once per variant
template <typename T, int OP, bool P2>
bool bottom_lds_fits_v(size_t lds) {
  constexpr int kDev = 64;
  static std::mutex mu;
  static int dev_limit[kDev];
  static size_t static_lds[kDev];
  static bool known[kDev], raised[kDev];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDev) return false;
  std::lock_guard<std::mutex> g(mu);
  if (!known[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return false;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_bottom<T, OP, P2, true>)) != hipSuccess)
      return false;
    static_lds[dev] = fa.sharedSizeBytes;
    dev_limit[dev] = v;
    known[dev] = true;
  }
  if (lds + static_lds[dev] > (size_t)dev_limit[dev]) return false;
  if (!raised[dev]) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bottom<T, OP, P2, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            dev_limit[dev] - (int)static_lds[dev]) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    raised[dev] = true;
  }
  return true;
}

template <typename T>
bool bottom_lds_fits(int op, bool p2, size_t lds, T *) {
  if (op == 0) return p2 ? bottom_lds_fits_v<T, 0, true>(lds) : bottom_lds_fits_v<T, 0, false>(lds);
  return p2 ? bottom_lds_fits_v<T, 1, true>(lds) : bottom_lds_fits_v<T, 1, false>(lds);
}

// ---------------------------------------------------------------------------
// Exact solve of a 17^3 level for the FP32 inner cycles (plain Poisson, no
// excision; no parity claim: the reference relaxes there).  The FAS
// equation of the level, L[u] = lap_h u - rho = rhs, is solved for the 15^3
// interior with u's boundary values as Dirichlet data, by the discrete sine
// transform that diagonalises the 7-point Laplacian: with Q the orthonormal
// DST-I basis of each axis (Q = S / sqrt(8), S[i][m] = sin((i+1)(m+1) pi / 16)),
// lap_h = Q Lambda Q^T, lambda = -(4/h^2) (sin^2(i' pi/32) + sin^2(j' pi/32) +
// sin^2(k' pi/32)), so u = (1/512) S3 (Lambda^-1 (S3 b)), b = rhs + rho - (the
// boundary neighbours' u / h^2).  One workgroup, 15^3 values in LDS, six
// 15-point transforms.
// ---------------------------------------------------------------------------
constexpr int kDstM = 15;  // interior points per axis (a level of 17)
constexpr int kDstThreads = 1024;

__global__ __launch_bounds__(kDstThreads) void k_dst_solve17(float *__restrict__ u, float *__restrict__ ua,
                                                     const float *__restrict__ rho, const float *__restrict__ rhs,
                                                     float h2inv, double *cnorm, int *citer) {
  constexpr int M = kDstM, N = M + 2, V = M * M * M;
  __shared__ float X[V], Y[V];
  __shared__ float sn[M][M];
  __shared__ float lam[M];
  const int t = threadIdx.x;
  for (int q = t; q < M * M; q += kDstThreads) {
    const int i = q / M, m = q % M;
    sn[i][m] = (float)std::sin((double)((i + 1) * (m + 1)) * M_PI / 16.0);
  }
  if (t < M) {
    const double sv = std::sin((double)(t + 1) * M_PI / 32.0);
    lam[t] = (float)(4.0 * sv * sv);
  }
  for (int q = t; q < V; q += kDstThreads) {
    const int i = q % M, j = (q / M) % M, k = q / (M * M);
    const int I = i + 1, J = j + 1, K = k + 1;
    const size_t p = ix3(I, J, K, N, N);
    float b = rhs[p] + rho[p];
    if (I == 1) b -= u[p - 1] * h2inv;
    if (I == M) b -= u[p + 1] * h2inv;
    if (J == 1) b -= u[p - N] * h2inv;
    if (J == M) b -= u[p + N] * h2inv;
    if (K == 1) b -= u[p - N * N] * h2inv;
    if (K == M) b -= u[p + N * N] * h2inv;
    X[q] = b;
  }
  __syncthreads();
  // Y = S along x, X = S along y, Y = S along z (and the same back)
  auto tx = [&](const float *in, float *out) {
    for (int q = t; q < V; q += kDstThreads) {
      const int i = q % M, row = q - i;
      float acc = 0.f;
      for (int m = 0; m < M; m++) acc += sn[i][m] * in[row + m];
      out[q] = acc;
    }
  };
  auto ty = [&](const float *in, float *out) {
    for (int q = t; q < V; q += kDstThreads) {
      const int i = q % M, j = (q / M) % M, k = q / (M * M);
      float acc = 0.f;
      for (int m = 0; m < M; m++) acc += sn[j][m] * in[(k * M + m) * M + i];
      out[q] = acc;
    }
  };
  auto tz = [&](const float *in, float *out) {
    for (int q = t; q < V; q += kDstThreads) {
      const int ij = q % (M * M), k = q / (M * M);
      float acc = 0.f;
      for (int m = 0; m < M; m++) acc += sn[k][m] * in[m * M * M + ij];
      out[q] = acc;
    }
  };
  tx(X, Y);
  __syncthreads();
  ty(Y, X);
  __syncthreads();
  tz(X, Y);
  __syncthreads();
  for (int q = t; q < V; q += kDstThreads) {
    const int i = q % M, j = (q / M) % M, k = q / (M * M);
    // lap_h eigenvalue -(lam_i + lam_j + lam_k) / h^2; 1/512 normalises S3 S3
    Y[q] = Y[q] / (-(lam[i] + lam[j] + lam[k]) * h2inv * 512.0f);
  }
  __syncthreads();
  tx(Y, X);
  __syncthreads();
  ty(X, Y);
  __syncthreads();
  tz(Y, X);
  __syncthreads();
  for (int q = t; q < V; q += kDstThreads) {
    const int i = q % M, j = (q / M) % M, k = q / (M * M);
    const size_t p = ix3(i + 1, j + 1, k + 1, N, N);
    u[p] = X[q];
    ua[p] = X[q];
  }
  if (t == 0) {
    cnorm[0] = 0.0;
    cnorm[1] = 0.0;
    *citer = 0;
  }
}

template <typename T>
int bottom(pmg_solver *S, int Lb, int norm_base, double *cnorm, int *citer) {
  if constexpr (std::is_same<T, float>::value) {
    // FP32 inner cycles: an exact solve of a 17^3 top bottom level instead
    // of the single-workgroup V below it (plain Poisson, no excision)
    Level &Lv = S->lev[Lb - 1];
    const bool cube17 = Lv.n[0] == kDstM + 2 && Lv.n[1] == kDstM + 2 && Lv.n[2] == kDstM + 2;
    const bool full = Lv.rbox.ok && Lv.rbox.lo[0] == 1 && Lv.rbox.lo[1] == 1 && Lv.rbox.lo[2] == 1 &&
                      Lv.rbox.hi[0] == kDstM && Lv.rbox.hi[1] == kDstM && Lv.rbox.hi[2] == kDstM;
    const bool iso = Lv.hdx[0] == Lv.hdx[1] && Lv.hdx[1] == Lv.hdx[2];
    if (S->fp32_exact_bottom && S->op == 0 && !S->do_ex && cube17 && full && iso) {
      int rc;
      if ((rc = lop_rhs<float>(S, Lb))) return rc;
      LevelArrays<float> &A = arrays<float>(Lv);
      hipLaunchKernelGGL(k_dst_solve17, dim3(1), dim3(kDstThreads), 0, S->stream, A.u, A.ua, A.rho, A.rhs,
                         (float)(1.0 / (Lv.hdx[0] * Lv.hdx[0])), cnorm, citer);
      PMG_HOST_CHECK(hipGetLastError());
      return 0;
    }
  }
  BottomP<T> P;
  P.nb = Lb;
  for (int L = 1; L <= Lb; L++) {
    Level &Lv = S->lev[L - 1];
    LevelArrays<T> &A = arrays<T>(Lv);
    BotLev<T> &B = P.lev[L - 1];
    B.u = A.u; B.rho = A.rho; B.a = A.a; B.c = A.c; B.rhs = A.rhs; B.res = A.res; B.lop = A.lop;
    B.nx = Lv.n[0]; B.ny = Lv.n[1]; B.nz = Lv.n[2];
    B.bx0 = Lv.rbox.lo[0]; B.bx1 = Lv.rbox.hi[0];
    B.by0 = Lv.rbox.lo[1]; B.by1 = Lv.rbox.hi[1];
    B.bz0 = Lv.rbox.lo[2]; B.bz1 = Lv.rbox.hi[2];
    if (Lv.cbox.ok) {
      B.cx0 = Lv.cbox.lo[0]; B.cx1 = Lv.cbox.hi[0];
      B.cy0 = Lv.cbox.lo[1]; B.cy1 = Lv.cbox.hi[1];
      B.cz0 = Lv.cbox.lo[2]; B.cz1 = Lv.cbox.hi[2];
    } else {
      B.cx0 = B.cy0 = B.cz0 = 1;
      B.cx1 = B.cy1 = B.cz1 = 0;
    }
    B.sp = spacing<T>(Lv);
  }
  P.npre = S->pre; P.npst = S->pst; P.max_citer = S->max_citer; P.crtol = S->crtol;
  P.w0_r = (T)S->w0_r; P.w0_i = (T)S->w0_i;
  P.rop_u = S->rop_u; P.rop_res = S->rop_res;
  P.ws = S->ws;
  P.relax_slot0 = norm_base + NS_RELAX;
  P.resid_slot0 = norm_base + NS_RESID;
  P.norm_pre = cnorm; P.norm_post = cnorm + 1; P.citer = citer;
  P.snap = reinterpret_cast<T *>(S->d_snap);
  const bool p2 = P.lev[0].sp.pow2 != 0;
  for (int L = 2; L <= Lb; L++)
    if ((P.lev[L - 1].sp.pow2 != 0) != p2) return -5;
  // the LDS variant when every array of every bottom level fits beside the
  // kernel's static LDS
  size_t lds = 0;
  for (int l = 0; l < Lb; l++)
    lds += (size_t)(S->op == 1 ? 7 : 5) * P.lev[l].nx * P.lev[l].ny * P.lev[l].nz * sizeof(T);
  const bool use_lds = S->bottom_lds && bottom_lds_fits(S->op, p2, lds, (T *)nullptr);
#define PMG_BOT(OPV, P2V)                                                                                           \
  do {                                                                                                              \
    if (use_lds)                                                                                                    \
      hipLaunchKernelGGL((k_bottom<T, OPV, P2V, true>), dim3(1), dim3(kBotThreads), lds, S->stream, P);            \
    else                                                                                                            \
      hipLaunchKernelGGL((k_bottom<T, OPV, P2V, false>), dim3(1), dim3(kBotThreads), 0, S->stream, P);             \
  } while (0)
  if (S->op == 0) { if (p2) PMG_BOT(0, true); else PMG_BOT(0, false); }
  else { if (p2) PMG_BOT(1, true); else PMG_BOT(1, false); }
#undef PMG_BOT
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// finest-level residual + full-weight injection of u and f_res in one pass
// (no excision, full weighting for both, fine = 2 coarse - 1), in the FP64
// cycle and in the FP32 inner cycle; the finest f_res is then never stored
template <typename T>
bool can_fuse_resid_restrict(const pmg_solver *S) {
  if (!S->use_fused_restrict || S->do_ex || S->rop_u != 3 || S->rop_res != 3 || S->nlev < 2) return false;
  const Level &F = S->lev[S->nlev - 1], &C = S->lev[S->nlev - 2];
  for (int d = 0; d < 3; d++)
    if (F.n[d] != 2 * C.n[d] - 1) return false;
  // the residual box must be the whole interior (ZERO semantics off it)
  return F.rbox.ok && F.rbox.lo[0] == 1 && F.rbox.lo[1] == 1 && F.rbox.lo[2] == 1 && F.rbox.hi[0] == F.n[0] - 2 &&
         F.rbox.hi[1] == F.n[1] - 2 && F.rbox.hi[2] == F.n[2] - 2;
}

// workgroups of the fused kernel for a coarse level of cn points and a
// given number of coarse planes per workgroup: the ring kernel tiles every
// coarse column, the streamed one columns 0 .. cn-2 (never more workgroups)
template <typename T>
int resid_restrict_blocks(const int *cn, int czchunk, bool streamed = false) {
  constexpr int CX = std::is_same<T, double>::value ? 16 : 32, CY = 8;
  const int tx = streamed ? rr_tiles(cn[0], CX) : (cn[0] + CX - 1) / CX;
  const int ty = streamed ? rr_tiles(cn[1], CY) : (cn[1] + CY - 1) / CY;
  return tx * ty * ((cn[2] + czchunk - 1) / czchunk);
}

template <typename T, int OP, bool P2>
int launch_resid_restrict_t(pmg_solver *S, ResRestrP<T> P) {
  constexpr int CX = std::is_same<T, double>::value ? 16 : 32, CY = 8, NT = 256;
  const int cn[3] = {P.cnx, P.cny, P.cnz};
  const dim3 grid(resid_restrict_blocks<T>(cn, P.czchunk, S->use_fused_restrict >= 2));
  P.nb = (int)grid.x;
  if (P.nb > S->ws.cap) return -3;
  if (S->use_fused_restrict == 2)
    hipLaunchKernelGGL((k_resid_restrict_s<T, OP, P2, CX, CY, NT, 1>), grid, dim3(NT), 0, S->stream, P);
  else if (S->use_fused_restrict == 3)
    hipLaunchKernelGGL((k_resid_restrict_s<T, OP, P2, CX, CY, NT, 2>), grid, dim3(NT), 0, S->stream, P);
  else
    hipLaunchKernelGGL((k_resid_restrict<T, OP, P2, CX, CY, NT>), grid, dim3(NT), 0, S->stream, P);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
int resid_restrict(pmg_solver *S, int slot) {
  Level &F = S->lev[S->nlev - 1], &C = S->lev[S->nlev - 2];
  LevelArrays<T> &AF = arrays<T>(F);
  LevelArrays<T> &AC = arrays<T>(C);
  ResRestrP<T> P;
  P.u = ucur<T>(F); P.rho = AF.rho; P.a = AF.a; P.c = AF.c;
  P.cu = AC.u; P.cua = AC.ua; P.cres = AC.res;
  P.nx = F.n[0]; P.ny = F.n[1]; P.nz = F.n[2];
  P.cnx = C.n[0]; P.cny = C.n[1]; P.cnz = C.n[2];
  P.czchunk = std::max(1, std::min(S->rr_czchunk[std::is_same<T, float>::value ? 1 : 0], C.n[2]));
  P.sp = spacing<T>(F);
  P.ws = S->ws;
  P.slot = slot;
  const bool p2 = P.sp.pow2 != 0;
  if (S->op == 0) return p2 ? launch_resid_restrict_t<T, 0, true>(S, P) : launch_resid_restrict_t<T, 0, false>(S, P);
  return p2 ? launch_resid_restrict_t<T, 1, true>(S, P) : launch_resid_restrict_t<T, 1, false>(S, P);
}

int band_inner(pmg_solver *S);

// sweeps(), with HIP events around the launches of the finest FP64 level
// while pmg_time_kernel (which 3) collects in-cycle durations
template <typename T>
int timed_sweeps(pmg_solver *S, int L, int n, int slot, bool rhs0) {
  const bool tk = std::is_same<T, double>::value && L == S->nlev && S->tk_ev && n > 0 &&
                  S->tk_n + 2 <= S->tk_cap;
  if (tk) PMG_HOST_CHECK(hipEventRecord(S->tk_ev[S->tk_n], S->stream));
  const int rc = sweeps<T>(S, L, n, slot, rhs0);
  if (tk) {
    PMG_HOST_CHECK(hipEventRecord(S->tk_ev[S->tk_n + 1], S->stream));
    S->tk_n += 2;
    S->tk_sweeps += n;
  }
  return rc;
}

template <typename T>
int enqueue_vcycle(pmg_solver *S, bool final_residual, int norm_base, double *cnorm, int *citer) {
  const int Lf = S->nlev;
  const int Lb = S->use_bottom ? bottom_top(S) : 0;
  int rc;
  for (int L = Lf; L >= 1; L--) {
    if (L < Lf) {
      LevelArrays<T> &AF = arrays<T>(S->lev[L]);
      LevelArrays<T> &AC = arrays<T>(S->lev[L - 1]);
      if (L + 1 == Lf && can_fuse_resid_restrict<T>(S)) {
        if ((rc = resid_restrict<T>(S, norm_base + NS_RESID + L))) return rc;
      } else {
        // residual on L+1; the finest level was cleared at the cycle start
        if ((rc = residual<T>(S, L + 1, norm_base + NS_RESID + L, L + 1 == Lf, L + 1 == Lf))) return rc;
        // PAMR_inject(L+1): u and u_res (sources are constant, restricted once)
        if ((rc = restrict_any<T>(S, L + 1, ucur<T>(S->lev[L]), AC.u, AC.ua, S->rop_u))) return rc;
        if ((rc = restrict_any<T>(S, L + 1, AF.res, AC.res, (T *)nullptr, S->rop_res))) return rc;
      }
      if (L == Lb) {
        if ((rc = bottom<T>(S, Lb, norm_base, cnorm, citer))) return rc;
        break;
      }
      if ((rc = lop_rhs<T>(S, L))) return rc;
    }
    if (L == 1) {
      if ((rc = coarse_solve<T>(S, 1, cnorm, cnorm + 1, citer, S->max_citer))) return rc;
    } else {
      if ((rc = timed_sweeps<T>(S, L, S->pre, norm_base + NS_RELAX + L - 1, L == Lf))) return rc;
    }
  }
  for (int L = Lb ? Lb + 1 : 2; L <= Lf; L++) {
    if ((rc = cgc<T>(S, L))) return rc;
    if constexpr (std::is_same<T, float>::value) {
      // the learned band correction of the finest level's coarse-grid correction
      if (L == Lf && S->band.enq_inner && (rc = band_inner(S))) return rc;
    }
    if ((rc = timed_sweeps<T>(S, L, S->pst, norm_base + NS_RELAX + L - 1, L == Lf))) return rc;
    if ((rc = settle<T>(S, L))) return rc;
  }
  if (final_residual) {
    if ((rc = residual<T>(S, Lf, norm_base + NS_FINAL, true, true))) return rc;
  }
  return 0;
}

int capture_or_run(pmg_solver *S, hipGraphExec_t *cache, int (*enq)(pmg_solver *)) {
  if (!S->use_graph) return enq(S);
  if (!*cache) {
    hipGraph_t g;
    PMG_HOST_CHECK(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    int rc = enq(S);
    hipGraph_t g2 = nullptr;
    hipError_t e = hipStreamEndCapture(S->stream, &g2);
    if (rc) { if (g2) (void)hipGraphDestroy(g2); return rc; }
    if (e != hipSuccess) { std::fprintf(stderr, "pamr_mg: graph capture failed: %s\n", hipGetErrorString(e)); return -1; }
    g = g2;
    PMG_HOST_CHECK(hipGraphInstantiate(cache, g, nullptr, nullptr, 0));
    PMG_HOST_CHECK(hipGraphDestroy(g));
  }
  PMG_HOST_CHECK(hipGraphLaunch(*cache, S->stream));
  return 0;
}

int enq_v64(pmg_solver *S) {
  int rc = enqueue_vcycle<double>(S, true, 0, S->d_norms + NS_COARSE, S->d_citer);
  if (rc) return rc;
  return fold_norms(S, 0, kNormSlots);
}

// Mixed precision outer iteration (defect correction).  On entry the FP32
// source of the finest level holds -r (r = L64[u], the FP64 defect) and the
// FP32 correction e is zero; the graph runs the FP32 V-cycle on A e = -r,
// adds e to u in FP64 (zeroing e again), and recomputes the FP64 defect of
// the updated u -- its norm is the convergence test, its negation in FP32
// the next source.
__global__ __launch_bounds__(256) void k_add_f32(double *u, float *e, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) {
    u[q] = u[q] + (double)e[q];
    e[q] = 0.0f;
  }
}

int enq_mixed(pmg_solver *S) {
  const int Lf = S->nlev;
  Level &F = S->lev[Lf - 1];
  int rc;
  if ((rc = enqueue_vcycle<float>(S, false, NS_F32, S->d_norms + NS_COARSE + 2, S->d_citer + 1))) return rc;
  hipLaunchKernelGGL(k_add_f32, dim3(nblk256(F.size)), dim3(256), 0, S->stream, F.d.u, F.f.u, F.size);
  PMG_HOST_CHECK(hipGetLastError());
  if ((rc = residual<double>(S, Lf, NS_OUTER, true, true, F.f.rho))) return rc;
  return fold_norms(S, 0, 2 * kNormSlots);
}

int ensure_fp32(pmg_solver *S) {
  if (S->fp32_ready) return 0;
  for (int L = 1; L <= S->nlev; L++)
    if (alloc_level_arrays(S, S->lev[L - 1], true)) return -1;
  // coefficients in FP32 on every level (restricted FP64 values rounded)
  for (int L = 1; L <= S->nlev; L++) {
    Level &Lv = S->lev[L - 1];
    if (S->op == 1) {
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.f.a, Lv.d.a, Lv.size);
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.f.c, Lv.d.c, Lv.size);
    }
  }
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  S->fp32_ready = true;
  return 0;
}

// The learned band correction of the defect-correction iteration: on entry
// the FP32 source of the finest level holds -r and d_norms[NS_OUTER] its
// RMS norm; u += ts * ||r|| * N(r / ||r||) at the band points, then -r is
// recomputed where that changed it (each core grown by one point).
int enq_band_dc(pmg_solver *S) {
  Level &F = S->lev[S->nlev - 1];
  auto &B = S->band;
  const double *nrm = S->d_norms + NS_OUTER;
  if (pmg_band_gather_f32(F.f.rho, F.n[0], F.n[1], F.n[2], B.d_org, B.ncore, B.side, -1.0f, nrm, B.d_x, S->stream))
    return -1;
  const float *y = pmg_net_infer_async(B.net, B.d_x, 1.0f, B.side, B.side, B.side, B.ncore, S->stream);
  if (!y) return -1;
  if (pmg_band_scatter(F.d.u, nullptr, nullptr, y, B.d_band, F.n[0], F.n[1], F.n[2], B.d_org, B.ncore, B.side,
                       B.margin, B.ts, nrm, S->stream))
    return -1;
  return pmg_band_resid(S->op, S->do_ex, F.d.u, F.d.rho, F.d.a, F.d.c, F.chr, F.f.rho, F.n[0], F.n[1], F.n[2],
                        B.d_org, B.ncore, B.side, B.margin, F.hdx[0], F.hdx[1], F.hdx[2], S->stream);
}

// The same correction of the preconditioner's output (flexible GMRES): the
// FP32 residual of the V-cycle's z for its source, its norm folded on the
// device, z += ts * ||r|| * N(r / ||r||) at the band points.
int enq_band_pc(pmg_solver *S) {
  Level &F = S->lev[S->nlev - 1];
  auto &B = S->band;
  const int slot = NS_F32 + NS_FINAL;
  int rc;
  if ((rc = residual<float>(S, S->nlev, slot, true, true))) return rc;
  if ((rc = fold_norms(S, slot, 1))) return rc;
  const double *nrm = S->d_norms + slot;
  if (pmg_band_gather_f32(F.f.res, F.n[0], F.n[1], F.n[2], B.d_org, B.ncore, B.side, 1.0f, nrm, B.d_x, S->stream))
    return -1;
  const float *y = pmg_net_infer_async(B.net, B.d_x, 1.0f, B.side, B.side, B.side, B.ncore, S->stream);
  if (!y) return -1;
  return pmg_band_scatter(nullptr, nullptr, ucur<float>(F), y, B.d_band, F.n[0], F.n[1], F.n[2], B.d_org, B.ncore,
                          B.side, B.margin, B.ts, nrm, S->stream);
}

// Inside the FP32 V-cycle, after the finest coarse-grid correction: u32 +=
// ts * |src| * N(r32 / |src|) at the band points, r32 the FP32 residual at
// that moment and |src| the RMS of the cycle's finest source (the cycle is a
// linear map of its source, so this scaling is exact); or, while training
// crops are requested, gather them instead.
int band_inner(pmg_solver *S) {
  Level &F = S->lev[S->nlev - 1];
  auto &B = S->band;
  if (!B.cur_norm || B.ncore < 1) return -1;
  float *u32 = ucur<float>(F);
  if (B.samp_x)
    return pmg_band_pairs_res32(S->op, S->do_ex, u32, F.f.rho, F.f.a, F.f.c, F.chr, B.samp_ustar, F.d.u, B.d_band,
                                F.n[0], F.n[1], F.n[2], B.samp_org, B.samp_n, B.side, B.cur_norm, F.hdx[0], F.hdx[1],
                                F.hdx[2], B.samp_x, B.samp_t, B.samp_w, S->stream);
  if (!B.net_in) return -1;
  if (pmg_band_gather_res32(S->op, S->do_ex, u32, F.f.rho, F.f.a, F.f.c, F.chr, F.n[0], F.n[1], F.n[2], B.d_org,
                            B.ncore, B.side, 1.0f, B.cur_norm, F.hdx[0], F.hdx[1], F.hdx[2], B.d_x, S->stream))
    return -1;
  const float *y = pmg_net_infer_async(B.net_in, B.d_x, 1.0f, B.side, B.side, B.side, B.ncore, S->stream);
  if (!y) return -1;
  return pmg_band_scatter(nullptr, nullptr, u32, y, B.d_band, F.n[0], F.n[1], F.n[2], B.d_org, B.ncore, B.side,
                          B.margin, B.ts_in, B.cur_norm, S->stream);
}

int enq_mixed_inner(pmg_solver *S) {
  S->band.enq_inner = true;
  S->band.cur_norm = S->d_norms + NS_OUTER;
  const int rc = enq_mixed(S);
  S->band.enq_inner = false;
  return rc;
}

int enq_prec_inner(pmg_solver *S) {
  S->band.enq_inner = true;
  S->band.cur_norm = S->d_norms + NS_PSRC;
  const int rc = enqueue_vcycle<float>(S, false, NS_F32, S->d_norms + NS_COARSE + 2, S->d_citer + 1);
  S->band.enq_inner = false;
  return rc;
}

void band_graphs_reset(pmg_solver *S) {
  auto &B = S->band;
  for (hipGraphExec_t *g : {&B.gdc, &B.gpc, &B.g32i, &B.gpreci})
    if (*g) { (void)hipGraphExecDestroy(*g); *g = nullptr; }
}

void invalidate_graphs(pmg_solver *S) {
  if (S->g64) { (void)hipGraphExecDestroy(S->g64); S->g64 = nullptr; }
  if (S->g32) { (void)hipGraphExecDestroy(S->g32); S->g32 = nullptr; }
  if (S->gprec) { (void)hipGraphExecDestroy(S->gprec); S->gprec = nullptr; }
  band_graphs_reset(S);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

pmg_solver *pmg_create(int nx, int ny, int nz, const double *bbox, int min_cwidth, int op,
                       int do_ex, double ex) {
  if (nx < 3 || ny < 3 || nz < 3 || !bbox || (op != 0 && op != 1)) return nullptr;
  pmg_solver *S = new pmg_solver();
  S->op = op;
  S->do_ex = do_ex ? 1 : 0;
  S->ex = ex;
  std::memcpy(S->bbox, bbox, 6 * sizeof(double));
  if (hipStreamCreate(&S->stream) != hipSuccess) { delete S; return nullptr; }
  // build_mgh coarsening (src/gh.c:1275-1340) via calc_island_no rules
  int shapes[kMaxLev][3];
  int nl = 1;
  shapes[0][0] = nx; shapes[0][1] = ny; shapes[0][2] = nz;
  double dx0[3];
  for (int d = 0; d < 3; d++) dx0[d] = (bbox[2 * d + 1] - bbox[2 * d]) / (shapes[0][d] - 1);
  for (int cf = 2; nl < kMaxLev; cf *= 2) {
    int island = 1;
    for (int d = 0; d < 3; d++) {
      int sh = (int)((bbox[2 * d + 1] - bbox[2 * d]) / dx0[d] + 1.5);
      if ((((sh - 1) % cf) && ((sh - 1) / cf + 1) >= min_cwidth) || (((sh - 1) / cf + 1) < min_cwidth))
        island = 0;
      else if (((sh - 1) / cf / 2 + 1) < min_cwidth && island != 0)
        island = -1;
    }
    if (island == 0) break;
    for (int d = 0; d < 3; d++) shapes[nl][d] = (int)((bbox[2 * d + 1] - bbox[2 * d]) / (dx0[d] * cf) + 1.5);
    nl++;
  }
  S->nlev = nl;
  for (int L = 1; L <= nl; L++) {
    Level &Lv = S->lev[L - 1];
    for (int d = 0; d < 3; d++) Lv.n[d] = shapes[nl - L][d];
    Lv.size = (size_t)Lv.n[0] * Lv.n[1] * Lv.n[2];
    for (int d = 0; d < 3; d++) {
      // coordinates by repeated addition (src/gh.c:905-906)
      double dx = (bbox[2 * d + 1] - bbox[2 * d]) / (Lv.n[d] - 1);
      double x0 = bbox[2 * d], x1 = x0 + dx, x2 = x1 + dx;
      Lv.hdx[d] = x1 - x0;
      Lv.gdx[d] = x2 - x1;
    }
    Lv.sp64 = make_spacing<double>(Lv.hdx);
    Lv.sp32 = make_spacing<float>(Lv.hdx);
    Box b = cmask_bdy_box(Lv.n, bbox, bbox, Lv.gdx);
    // relax/residual act on points 2..N-1 (1-based) where cmask is ON
    for (int d = 0; d < 3; d++) {
      b.lo[d] = std::max(b.lo[d], 1);
      b.hi[d] = std::min(b.hi[d], Lv.n[d] - 2);
      if (b.lo[d] > b.hi[d]) b.ok = false;
    }
    Lv.rbox = b;
    Lv.cbox = cmask_child_box(Lv.n, bbox, bbox, Lv.gdx, bbox);
    if (alloc_level_arrays(S, Lv, false)) { pmg_destroy(S); return nullptr; }
  }
  for (int L = 1; L <= nl; L++) {
    const int n = S->lev[L - 1].n[2];
    for (int pi = 0; pi < 2; pi++) {
      S->lv_variant[pi][L - 1] = 1;
      S->lv_zchunk[pi][L - 1] = n >= 257 ? 32 : (n >= 129 ? 16 : (n >= 65 ? 8 : 4));
    }
  }
  // coarse planes per block of the fused finest residual + injection
  // (streamed kernel), the fastest measured: 6 (FP64) and 4 (FP32) on a
  // coarse level of 129 planes, 8 and 8 from 257 planes; smaller grids 8
  if (nl >= 2 && S->lev[nl - 2].n[2] >= 129 && S->lev[nl - 2].n[2] < 257) {
    S->rr_czchunk[0] = 6;
    S->rr_czchunk[1] = 4;
  }
  // norm workspace sized for the largest launch
  // partial-norm capacity: the largest launch over the finest level
  const Level &F = S->lev[nl - 1];
  const int rblk = ((F.n[0] + RTX - 1) / RTX) * ((F.n[1] + RTY - 1) / RTY) * ((F.n[2] + RZC - 1) / RZC);
  const int sblk = ((F.n[0] + 61) / 64) * ((F.n[1] + 5) / 8) * ((F.n[2] + 5) / 8);
  int cap = std::max(1 << 14, std::max(rblk, sblk) + 64);
  S->ws.cap = cap;
  S->ws.nslots = 2 * kNormSlots;
  if (hipMalloc(&S->ws.partials, sizeof(double) * 2 * (size_t)cap * S->ws.nslots) != hipSuccess ||
      hipMalloc(&S->ws.slot_nb, sizeof(int) * S->ws.nslots) != hipSuccess ||
      hipMalloc(&S->d_norms, sizeof(double) * kNormTotal) != hipSuccess ||
      hipMalloc(&S->d_citer, sizeof(int) * 2) != hipSuccess ||
      hipMalloc(&S->d_snap, sizeof(double) * kSnap * S->lev[0].size) != hipSuccess ||
      hipHostMalloc(&S->h_pinned, sizeof(double) * (kNormTotal + 4)) != hipSuccess) {
    pmg_destroy(S);
    return nullptr;
  }
  (void)hipMemsetAsync(S->ws.slot_nb, 0, sizeof(int) * S->ws.nslots, S->stream);
  (void)hipMemsetAsync(S->d_norms, 0, sizeof(double) * kNormTotal, S->stream);
  (void)hipMemsetAsync(S->d_citer, 0, sizeof(int) * 2, S->stream);
  if (S->do_ex) {
    // default chr: nothing excised (every value ex + 1)
    for (int L = 1; L <= nl; L++) {
      std::vector<double> h(S->lev[L - 1].size, ex + 1.0);
      if (hipMemcpy(S->lev[L - 1].chrd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        pmg_destroy(S);
        return nullptr;
      }
    }
  }
  if (hipStreamSynchronize(S->stream) != hipSuccess) { pmg_destroy(S); return nullptr; }
  return S;
}

void pmg_destroy(pmg_solver *S) {
  if (!S) return;
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  invalidate_graphs(S);
  for (int L = 0; L < S->nlev; L++) {
    free_arrays(S->lev[L].d);
    free_arrays(S->lev[L].f);
    if (S->lev[L].chr) (void)hipFree(S->lev[L].chr);
    if (S->lev[L].chrd) (void)hipFree(S->lev[L].chrd);
  }
  if (S->ws.partials) (void)hipFree(S->ws.partials);
  if (S->ws.slot_nb) (void)hipFree(S->ws.slot_nb);
  if (S->d_norms) (void)hipFree(S->d_norms);
  if (S->d_citer) (void)hipFree(S->d_citer);
  if (S->d_snap) (void)hipFree(S->d_snap);
  if (S->h_pinned) (void)hipHostFree(S->h_pinned);
  if (S->d_hist) (void)hipFree(S->d_hist);
  if (S->h_hist) (void)hipHostFree(S->h_hist);
  if (S->kry) (void)hipFree(S->kry);
  if (S->kry_red) (void)hipFree(S->kry_red);
  if (S->kry_host) (void)hipHostFree(S->kry_host);
  if (S->band.d_org) (void)hipFree(S->band.d_org);
  if (S->band.d_band) (void)hipFree(S->band.d_band);
  if (S->band.d_x) (void)hipFree(S->band.d_x);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
}

int pmg_nlev(const pmg_solver *S) { return S ? S->nlev : 0; }

int pmg_level_shape(const pmg_solver *S, int L, int *shape) {
  if (!S || L < 1 || L > S->nlev) return -1;
  for (int d = 0; d < 3; d++) shape[d] = S->lev[L - 1].n[d];
  return 0;
}

int pmg_set_params(pmg_solver *S, int pre, int pst, int max_citer, double crtol, double w0_r,
                   double w0_i, int rop_u, int rop_res, int rop_src, int interp_ord) {
  if (!S || pre < 0 || pst < 0 || max_citer < 0 || interp_ord != 2) return -1;
  if (rop_u < 0 || rop_u > 3 || rop_res < 0 || rop_res > 3 || rop_src < 0 || rop_src > 3) return -1;
  S->pre = pre; S->pst = pst; S->max_citer = max_citer; S->crtol = crtol;
  S->w0_r = w0_r; S->w0_i = w0_i; S->rop_u = rop_u; S->rop_res = rop_res; S->rop_src = rop_src;
  S->iord = interp_ord;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_sweep_config(pmg_solver *S, int variant, int zchunk) {
  // the same tile variant and z chunk for every level
  if (!S || variant < 0 || variant >= kNumSweepVariants || zchunk < 1) return -1;
  for (int L = 0; L < S->nlev; L++)
    for (int pi = 0; pi < 2; pi++) {
      S->lv_variant[pi][L] = variant;
      S->lv_zchunk[pi][L] = zchunk;
    }
  invalidate_graphs(S);
  return 0;
}

int pmg_set_level_sweep_config(pmg_solver *S, int L, int variant, int zchunk) {
  int rc = pmg_set_level_sweep_config_prec(S, L, 0, variant, zchunk);
  return rc ? rc : pmg_set_level_sweep_config_prec(S, L, 1, variant, zchunk);
}

int pmg_set_level_sweep_config_prec(pmg_solver *S, int L, int prec, int variant, int zchunk) {
  if (!S || L < 1 || L > S->nlev || prec < 0 || prec > 1 || variant < 0 || variant >= kNumSweepVariants ||
      zchunk < 1)
    return -1;
  S->lv_variant[prec][L - 1] = variant;
  S->lv_zchunk[prec][L - 1] = zchunk;
  invalidate_graphs(S);
  return 0;
}

int pmg_num_sweep_variants(void) { return kNumSweepVariants; }

int pmg_set_graph(pmg_solver *S, int on) {
  if (!S) return -1;
  S->use_graph = on ? 1 : 0;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_resid_restrict_zchunk(pmg_solver *S, int prec, int czchunk) {
  if (!S || (prec != 0 && prec != 1) || czchunk < 1 || czchunk > 64) return -1;
  if (S->nlev >= 2) {
    // the launch must fit the partial-norm workspace
    const int *cn = S->lev[S->nlev - 2].n;
    const int cz = std::min(czchunk, cn[2]);
    const int nb = prec ? resid_restrict_blocks<float>(cn, cz) : resid_restrict_blocks<double>(cn, cz);
    if (nb > S->ws.cap) {
      std::fprintf(stderr, "pamr_mg: %d coarse planes per workgroup need %d workgroups, more than the %d the norm "
                   "workspace holds\n", czchunk, nb, S->ws.cap);
      return -3;
    }
  }
  S->rr_czchunk[prec] = czchunk;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_bottom_fusion(pmg_solver *S, int on) {
  if (!S || on < 0 || on > kBottomMaxN) return -1;
  S->use_bottom = on ? 1 : 0;
  S->bottom_max_n = on > 1 ? on : kBottomMaxN;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_fp32_exact_bottom(pmg_solver *S, int on) {
  if (!S) return -1;
  S->fp32_exact_bottom = on ? 1 : 0;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_bottom_lds(pmg_solver *S, int on) {
  if (!S || on < 0 || on > 1) return -1;
  S->bottom_lds = on;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_sweep_blocking(pmg_solver *S, int on, int zchunk) {
  if (!S || zchunk < 0) return -1;
  S->use_sweep2 = on ? 1 : 0;
  if (zchunk > 0) S->sweep2_zchunk = zchunk;
  invalidate_graphs(S);
  return 0;
}

int pmg_set_restrict_fusion(pmg_solver *S, int on) {
  if (!S || on < 0 || on > 3) return -1;
  S->use_fused_restrict = on;
  invalidate_graphs(S);
  return 0;
}

void *pmg_stream(pmg_solver *S) { return S ? (void *)S->stream : nullptr; }

double *pmg_field(pmg_solver *S, int L, int which) {
  if (!S || L < 1 || L > S->nlev) return nullptr;
  LevelArrays<double> &d = S->lev[L - 1].d;
  switch (which) {
    case PMG_F_U: return d.u;
    case PMG_F_RHO: return d.rho;
    case PMG_F_RHS: return d.rhs;
    case PMG_F_RES: return d.res;
    case PMG_F_LOP: return d.lop;
    case PMG_F_A: return d.a;
    case PMG_F_C: return d.c;
    default: return nullptr;
  }
}

int pmg_upload(pmg_solver *S, int L, int which, const double *host) {
  double *d = pmg_field(S, L, which);
  if (!d || !host) return -1;
  const size_t n = S->lev[L - 1].size;
  PMG_HOST_CHECK(hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, S->stream));
  if (which == PMG_F_U)
    PMG_HOST_CHECK(hipMemcpyAsync(S->lev[L - 1].d.ua, host, n * sizeof(double), hipMemcpyHostToDevice, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int pmg_download(pmg_solver *S, int L, int which, double *host) {
  double *d = pmg_field(S, L, which);
  if (!d || !host) return -1;
  PMG_HOST_CHECK(hipMemcpyAsync(host, d, S->lev[L - 1].size * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int pmg_set_chr(pmg_solver *S, int L, const double *chr_host) {
  if (!S || !S->do_ex || L < 1 || L > S->nlev || !chr_host) return -1;
  Level &Lv = S->lev[L - 1];
  PMG_HOST_CHECK(hipMemcpyAsync(Lv.chrd, chr_host, Lv.size * sizeof(double), hipMemcpyHostToDevice, S->stream));
  hipLaunchKernelGGL(k_chr_bytes, dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.chr, Lv.chrd,
                     S->ex, Lv.size);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int pmg_download_cmask(pmg_solver *S, int L, int kind, double *host) {
  // materialise the cmask the reference would hold: kind 0 = set_cmask_bdy
  // (during sweeps), 1 = set_cmask_child (during L_op / compute_MG_rhs)
  if (!S || L < 1 || L > S->nlev || !host) return -1;
  Level &Lv = S->lev[L - 1];
  Box b = kind == 0 ? cmask_bdy_box(Lv.n, S->bbox, S->bbox, Lv.gdx)
                    : cmask_child_box(Lv.n, S->bbox, S->bbox, Lv.gdx, S->bbox);
  for (size_t q = 0; q < Lv.size; q++) host[q] = 0.0;
  if (!b.ok) return 0;
  for (int k = b.lo[2]; k <= b.hi[2]; k++)
    for (int j = b.lo[1]; j <= b.hi[1]; j++)
      for (int i = b.lo[0]; i <= b.hi[0]; i++) host[ix3(i, j, k, Lv.n[0], Lv.n[1])] = 1.0;
  return 0;
}

int pmg_init(pmg_solver *S) {
  // PAMR_build_mgh's TF_MGH_INIT injection (src/gh.c:1237-1239,1334): u and
  // the in-AMRH sources/coefficients restricted level by level; the MG
  // variables res/rhs/lop start at zero
  if (!S) return -1;
  int rc;
  for (int L = S->nlev; L > 1; L--) {
    Level &F = S->lev[L - 1], &Cl = S->lev[L - 2];
    if ((rc = restrict_any<double>(S, L, F.d.u, Cl.d.u, nullptr, S->rop_u))) return rc;
    if ((rc = restrict_any<double>(S, L, F.d.rho, Cl.d.rho, nullptr, S->rop_src))) return rc;
    if (S->op == 1) {
      // the coefficients of the (project-defined) variable-coefficient
      // operator have no excised value: restricted without the excision
      // renormalisation, also when excision is on
      if ((rc = restrict_field<double>(S, L, F.d.a, Cl.d.a, nullptr, S->rop_src, false))) return rc;
      if ((rc = restrict_field<double>(S, L, F.d.c, Cl.d.c, nullptr, S->rop_src, false))) return rc;
    }
  }
  for (int L = 1; L <= S->nlev; L++) {
    Level &Lv = S->lev[L - 1];
    hipLaunchKernelGGL(k_copy<double>, dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.d.ua, Lv.d.u, Lv.size);
    PMG_HOST_CHECK(hipMemsetAsync(Lv.d.rhs, 0, Lv.size * sizeof(double), S->stream));
    PMG_HOST_CHECK(hipMemsetAsync(Lv.d.res, 0, Lv.size * sizeof(double), S->stream));
    PMG_HOST_CHECK(hipMemsetAsync(Lv.d.lop, 0, Lv.size * sizeof(double), S->stream));
  }
  if (S->op == 1 && S->fp32_ready) {
    for (int L = 1; L <= S->nlev; L++) {
      Level &Lv = S->lev[L - 1];
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.f.a, Lv.d.a, Lv.size);
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk256(Lv.size)), dim3(256), 0, S->stream, Lv.f.c, Lv.d.c, Lv.size);
    }
  }
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int pmg_vcycle_async(pmg_solver *S) {
  if (!S) return -1;
  return capture_or_run(S, &S->g64, enq_v64);
}

int pmg_read_norms(pmg_solver *S, double *res_out, double *lev_norms, int *citer) {
  if (!S) return -1;
  PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms, sizeof(double) * kNormTotal, hipMemcpyDeviceToHost, S->stream));
  int ci[2];
  PMG_HOST_CHECK(hipMemcpyAsync(ci, S->d_citer, sizeof(int) * 2, hipMemcpyDeviceToHost, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  if (res_out) *res_out = S->h_pinned[NS_FINAL];
  if (lev_norms) {
    for (int L = 1; L <= S->nlev; L++) lev_norms[L - 1] = S->h_pinned[NS_RELAX + L - 1];
    lev_norms[0] = S->h_pinned[NS_COARSE + 1];
    lev_norms[S->nlev - 1] = S->h_pinned[NS_FINAL];
  }
  if (citer) *citer = ci[0];
  return 0;
}

int pmg_vcycle(pmg_solver *S, double *res_out, double *lev_norms, int *citer) {
  int rc = pmg_vcycle_async(S);
  if (rc) return rc;
  return pmg_read_norms(S, res_out, lev_norms, citer);
}

int pmg_residual(pmg_solver *S, double *norm_out) {
  if (!S) return -1;
  int rc = residual<double>(S, S->nlev, NS_R0, true, true);
  if (rc) return rc;
  if ((rc = fold_norms(S, NS_R0, 1))) return rc;
  PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_R0, sizeof(double), hipMemcpyDeviceToHost, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  if (norm_out) *norm_out = S->h_pinned[0];
  return 0;
}

int pmg_solve(pmg_solver *S, int min_iter, int max_iter, double tol, double *norms, int *iters) {
  // solve_elliptics loop (amrd/mg.c:56-72): absolute RMS tolerance
  if (!S) return -1;
  int it = 0;
  double res = tol + 1;
  while ((it < min_iter || res > tol) && it < max_iter) {
    int rc = pmg_vcycle(S, &res, nullptr, nullptr);
    if (rc) return rc;
    if (norms) norms[it] = res;
    it++;
  }
  if (iters) *iters = it;
  return 0;
}

int pmg_solve_rel(pmg_solver *S, int mode, double rtol, int max_iter, double *hist, int *iters,
                  double *r0_out) {
  // Relative-residual solve: ||r_k|| / ||r_0|| <= rtol with ||.|| the RMS
  // norm of the residual hook and r_0 the residual of the initial u.
  // mode 0: plain FP64 V-cycles; mode 1: FP32 inner V-cycles in FP64 defect
  // correction (the FP64 residual, the update and the test stay FP64).
  // *r0_out > 0 on entry: that norm is the reference (e.g. the residual of
  // u = 0 when the solve starts from a warm start); the test is then
  // ||r_k|| <= rtol * reference.
  if (!S || max_iter < 0) return -1;
  double r0;
  int rc;
  if (mode == 1 || mode == 2 || mode == 3) {
    // the FP64 defect of the initial u: its norm is r0 and, in FP32, it is
    // the source of the first inner cycle
    if ((rc = ensure_fp32(S))) return rc;
    Level &F = S->lev[S->nlev - 1];
    if ((rc = residual<double>(S, S->nlev, NS_OUTER, true, true, F.f.rho))) return rc;
    if ((rc = fold_norms(S, NS_OUTER, 1))) return rc;
    PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_OUTER, sizeof(double), hipMemcpyDeviceToHost, S->stream));
    PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
    r0 = S->h_pinned[0];
  } else if ((rc = pmg_residual(S, &r0))) {
    return rc;
  }
  const double rref = (r0_out && *r0_out > 0) ? *r0_out : r0;
  if (r0_out) *r0_out = rref;
  int it = 0;
  double r = r0;
  r0 = rref;
  if (mode == 0) {
    while (it < max_iter && r > rtol * r0) {
      if ((rc = pmg_vcycle(S, &r, nullptr, nullptr))) return rc;
      if (hist) hist[it] = r;
      it++;
    }
  } else if (mode == 1) {
    while (it < max_iter && r > rtol * r0) {
      if ((rc = capture_or_run(S, &S->g32, enq_mixed))) return rc;
      PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_OUTER, sizeof(double), hipMemcpyDeviceToHost, S->stream));
      PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
      r = S->h_pinned[0];
      if (hist) hist[it] = r;
      it++;
    }
  } else if (mode == 2) {
    // mixed precision with the learned correction before each FP32
    // V-cycle, kept only when its contraction per unit time beats the
    // plain cycle's (measured); otherwise undone and retried later
    if (!S->net) return -1;
    if (S->do_ex) {
      // the correction kernel writes every interior point, excised ones too
      std::fprintf(stderr, "pamr_mg: pmg_solve_rel mode 2 does not support excision\n");
      return -1;
    }
    Level &F = S->lev[S->nlev - 1];
    const double h2 = F.hdx[0] * F.hdx[0];
    EventPair ev;
    PMG_HOST_CHECK(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    auto read_outer = [&](double *out) -> int {
      PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_OUTER, sizeof(double), hipMemcpyDeviceToHost, S->stream));
      PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
      *out = S->h_pinned[0];
      return 0;
    };
    const bool try_net = S->net_enabled && S->net_skip == 0;
    if (S->net_skip > 0) S->net_skip--;
    bool net_on = try_net;
    while (it < max_iter && r > rtol * r0) {
      if (net_on && r > 0) {
        // buffers and weight fragments outside the timed region: the
        // acceptance test compares steady-state costs
        if ((rc = pmg_net_reserve(S->net, (long long)F.n[0] * F.n[1] * F.n[2]))) return rc;
        // u before the step, in the twin buffer (no FP64 sweep runs in this
        // mode): a rejected step restores it exactly
        PMG_HOST_CHECK(hipMemcpyAsync(F.d.ua, F.d.u, F.size * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
        PMG_HOST_CHECK(hipEventRecord(e0, S->stream));
        if ((rc = pmg_net_correct(S->net, F.f.rho, (float)(-1.0 / r), F.d.u, r * h2, F.n[0], F.n[1], F.n[2],
                                  S->stream))) return rc;
        if ((rc = residual<double>(S, S->nlev, NS_OUTER, true, true, F.f.rho))) return rc;
        if ((rc = fold_norms(S, NS_OUTER, 1))) return rc;
        PMG_HOST_CHECK(hipEventRecord(e1, S->stream));
        double rn;
        if ((rc = read_outer(&rn))) return rc;
        float ms = 0.f;
        PMG_HOST_CHECK(hipEventElapsedTime(&ms, e0, e1));
        S->t_net_ms = S->t_net_ms < 0 ? ms : 0.5 * (S->t_net_ms + ms);
        const double tv = S->t_vc_ms > 0 ? S->t_vc_ms : ms;
        const bool better = rn < r && (std::log(rn / r) / S->t_net_ms) < (std::log(S->rho_vc) / tv);
        if (better) {
          S->net_accepted++;
          S->net_retry_period = 20;
          r = rn;
          if (hist) hist[it] = r;
          it++;
          continue;
        }
        // undo: u restored from its copy; the defect is recomputed for the V-cycle
        S->net_rejected++;
        PMG_HOST_CHECK(hipMemcpyAsync(F.d.u, F.d.ua, F.size * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
        if ((rc = residual<double>(S, S->nlev, NS_OUTER, true, true, F.f.rho))) return rc;
        net_on = false;
        S->net_skip = S->net_retry_period;
        S->net_retry_period = std::min(2 * S->net_retry_period, 1280);
      }
      PMG_HOST_CHECK(hipEventRecord(e0, S->stream));
      if ((rc = capture_or_run(S, &S->g32, enq_mixed))) return rc;
      PMG_HOST_CHECK(hipEventRecord(e1, S->stream));
      double rn;
      if ((rc = read_outer(&rn))) return rc;
      float ms = 0.f;
      PMG_HOST_CHECK(hipEventElapsedTime(&ms, e0, e1));
      S->t_vc_ms = S->t_vc_ms < 0 ? ms : 0.5 * (S->t_vc_ms + ms);
      if (r > 0 && rn > 0) S->rho_vc = 0.5 * S->rho_vc + 0.5 * (rn / r);
      r = rn;
      if (hist) hist[it] = r;
      it++;
    }
  } else if (mode == 3) {
    // mixed precision with the learned band correction (excision) before
    // each FP32 V-cycle.  Every iteration is judged by the FP64 residual:
    // its contraction per unit time must beat the plain iteration's (both
    // measured, running means); the first iteration that does not ends the
    // correction for the rest of the solve (plain iterations from there on).
    auto &B = S->band;
    pmg_net *act = B.inner ? B.net_in : B.net;
    if (!act || B.ncore < 1) return -1;
    if ((rc = pmg_net_reserve(act, (long long)B.ncore * B.side * B.side * B.side))) return rc;
    EventPair ev;
    PMG_HOST_CHECK(ev.create());
    bool net_on = B.use_dc != 0;
    while (it < max_iter && r > rtol * r0) {
      // the first iteration of a solve is plain and not measured: its
      // defect (of the initial guess) is unlike the post-cycle defects the
      // network was trained on, and its contraction unlike the later ones
      const bool first = it == 0;
      const bool with_net = !first && net_on && B.rho_plain > 0 && B.t_plain > 0;
      bool fresh;  // graph capture in this iteration
      if (with_net && B.inner) fresh = !B.g32i;
      else if (with_net) fresh = !B.gdc || !S->g32;
      else fresh = !S->g32;
      PMG_HOST_CHECK(hipEventRecord(ev.e0, S->stream));
      if (with_net && B.inner) {
        if ((rc = capture_or_run(S, &B.g32i, enq_mixed_inner))) return rc;
      } else {
        if (with_net && (rc = capture_or_run(S, &B.gdc, enq_band_dc))) return rc;
        if ((rc = capture_or_run(S, &S->g32, enq_mixed))) return rc;
      }
      PMG_HOST_CHECK(hipEventRecord(ev.e1, S->stream));
      PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_OUTER, sizeof(double), hipMemcpyDeviceToHost,
                                    S->stream));
      PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
      const double rn = S->h_pinned[0];
      float ms = 0.f;
      PMG_HOST_CHECK(hipEventElapsedTime(&ms, ev.e0, ev.e1));
      const double rho = r > 0 ? rn / r : 1.0;
      if (with_net) {
        if (!fresh) B.t_net = B.t_net < 0 ? ms : 0.5 * (B.t_net + ms);
        B.rho_net = B.rho_net < 0 ? rho : 0.5 * (B.rho_net + rho);
        const double t = fresh ? (B.t_net > 0 ? B.t_net : B.t_plain) : (double)ms;
        if (rho < 1.0 && std::log(rho) / t < std::log(B.rho_plain) / B.t_plain) {
          B.accepted++;
        } else {
          B.rejected++;
          net_on = false;
        }
      } else if (!first && rho > 0 && rho < 1.0) {
        if (!fresh) B.t_plain = B.t_plain < 0 ? ms : 0.5 * (B.t_plain + ms);
        B.rho_plain = B.rho_plain < 0 ? rho : 0.5 * (B.rho_plain + rho);
      }
      r = rn;
      if (hist) hist[it] = r;
      it++;
    }
  } else {
    return -1;
  }
  if (iters) *iters = it;
  return 0;
}

int pmg_set_net(pmg_solver *S, pmg_net *N) {
  if (!S) return -1;
  S->net = N;
  S->net_enabled = 1;
  S->net_skip = 0;
  S->net_retry_period = 20;
  S->t_net_ms = -1;
  return 0;
}

int pmg_net_stats(pmg_solver *S, int *accepted, int *rejected, double *t_net_ms, double *t_vc_ms) {
  if (!S) return -1;
  if (accepted) *accepted = S->net_accepted;
  if (rejected) *rejected = S->net_rejected;
  if (t_net_ms) *t_net_ms = S->t_net_ms;
  if (t_vc_ms) *t_vc_ms = S->t_vc_ms;
  return 0;
}

// u (finest, both sweep buffers) = ca * a + cb * b (device FP64, b may be NULL)
__global__ __launch_bounds__(256) void k_lincomb(double *u, double *ua, const double *a, const double *b,
                                                 double ca, double cb, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const double v = b ? ca * a[q] + cb * b[q] : ca * a[q];
  u[q] = v;
  ua[q] = v;
}

int pmg_set_solution_lincomb(pmg_solver *S, const double *a, const double *b, double ca, double cb) {
  if (!S || !a) return -1;
  Level &F = S->lev[S->nlev - 1];
  hipLaunchKernelGGL(k_lincomb, dim3(nblk256(F.size)), dim3(256), 0, S->stream, F.d.u, F.d.ua, a, b, ca, cb,
                     F.size);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// History warm start (a minimal-residual projection onto earlier solves, in
// the spirit of P. F. Fischer, "Projection techniques for iterative solution
// of Ax = b with successive right-hand sides", CMAME 163 (1998) 193-204).
// Earlier sources F_0 (newest) .. F_{K-1} and their solutions U_k are given;
// since L is linear, u0 = sum c_k U_k has the residual
// (rho - sum c_k F_k) + sum c_k r_k, so c is chosen to minimise
// ||rho - sum c_k F_k||_2.  The F_k of a smooth sequence are nearly
// collinear, so the normal equations are formed in the backward-difference
// basis B_j = nabla^j F_0 = sum_i (-1)^i C(j,i) F_i, whose vectors shrink
// geometrically but stay far from collinear: each Gram entry keeps its
// relative accuracy, and a diagonal scaling leaves a well-conditioned
// system.  The K x K solve runs on the host in long double.
// ---------------------------------------------------------------------------
constexpr int kHistMax = 8;
constexpr int kHistBlocks = 1024;
constexpr int kHistM = kHistMax * (kHistMax + 1) / 2 + kHistMax + 1;

struct HistPtrs {
  const double *p[kHistMax];
};

__host__ __device__ constexpr int binom_c(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}

// per block: G (upper triangle, row-major), g = B^T rho, rho^T rho
template <int K>
__global__ __launch_bounds__(256) void k_hist_gram(HistPtrs F, const double *rho, size_t n, double *partials) {
  constexpr int M = K * (K + 1) / 2 + K + 1;
  double acc[M];
#pragma unroll
  for (int m = 0; m < M; m++) acc[m] = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    double f[K], b[K];
#pragma unroll
    for (int k = 0; k < K; k++) f[k] = F.p[k][q];
#pragma unroll
    for (int j = 0; j < K; j++) {
      double v = 0.0;
#pragma unroll
      for (int i = 0; i <= j; i++) v += ((i & 1) ? -1.0 : 1.0) * (double)binom_c(j, i) * f[i];
      b[j] = v;
    }
    const double r = rho[q];
    int m = 0;
#pragma unroll
    for (int j = 0; j < K; j++)
#pragma unroll
      for (int l = j; l < K; l++) acc[m++] += b[j] * b[l];
#pragma unroll
    for (int j = 0; j < K; j++) acc[m++] += b[j] * r;
    acc[m] += r * r;
  }
  __shared__ double red[256 / kWave][M];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int m = 0; m < M; m++) {
    double v = acc[m];
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    if (lane == 0) red[wid][m] = v;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < M; m += 256) {
    double v = 0.0;
    for (int w = 0; w < 256 / kWave; w++) v += red[w][m];
    partials[(size_t)blockIdx.x * M + m] = v;
  }
}

// fold of per-block partials: out[m] = sum_b partials[b][m], one workgroup
// per m; thread t sums b = t, t + 256, ... and the workgroup adds the 256
// sums in a fixed tree (deterministic for fixed nb)
__global__ __launch_bounds__(256) void k_hist_fold(const double *partials, int nb, int M, double *out) {
  const int m = blockIdx.x;
  double v = 0.0, unused = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += partials[(size_t)b * M + m];
  block_reduce2<256>(v, unused);
  if (threadIdx.x == 0) out[m] = v;
}

// per block: sum (rho - sum_k c_k F_k)^2 and sum rho^2
template <int K>
__global__ __launch_bounds__(256) void k_hist_misfit(HistPtrs F, const double *rho, const double *c, size_t n,
                                                     double *partials) {
  double e2 = 0.0, r2 = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    const double r = rho[q];
    double v = r;
#pragma unroll
    for (int k = 0; k < K; k++) v -= c[k] * F.p[k][q];
    e2 += v * v;
    r2 += r * r;
  }
  block_reduce2<256>(e2, r2);
  if (threadIdx.x == 0) {
    partials[(size_t)blockIdx.x * 2] = e2;
    partials[(size_t)blockIdx.x * 2 + 1] = r2;
  }
}

// u (finest, both sweep buffers) = sum_k c_k U_k
template <int K>
__global__ __launch_bounds__(256) void k_hist_lincomb(double *u, double *ua, HistPtrs U, const double *c,
                                                      size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < K; k++) v += c[k] * U.p[k][q];
  u[q] = v;
  ua[q] = v;
}

// Solves G a = g (G symmetric positive semi-definite, K x K) after scaling by
// the diagonal, with a relative ridge for exactly dependent histories, by
// Gaussian elimination with partial pivoting in long double.
void hist_solve(int K, const double *Gup, const double *g, long double *a) {
  long double A[kHistMax][kHistMax + 1], d[kHistMax];
  int m = 0;
  for (int j = 0; j < K; j++)
    for (int l = j; l < K; l++) {
      A[j][l] = A[l][j] = Gup[m];
      m++;
    }
  for (int j = 0; j < K; j++) d[j] = A[j][j] > 0 ? 1.0L / std::sqrt((long double)A[j][j]) : 0.0L;
  for (int j = 0; j < K; j++) {
    for (int l = 0; l < K; l++) A[j][l] *= d[j] * d[l];
    A[j][j] += 1e-15L;  // ridge on the unit diagonal: exactly dependent histories
    A[j][K] = g[j] * d[j];
  }
  for (int col = 0; col < K; col++) {
    int piv = col;
    for (int r = col + 1; r < K; r++)
      if (std::fabs((double)A[r][col]) > std::fabs((double)A[piv][col])) piv = r;
    if (piv != col)
      for (int l = 0; l <= K; l++) std::swap(A[col][l], A[piv][l]);
    if (A[col][col] == 0) continue;
    for (int r = col + 1; r < K; r++) {
      const long double f = A[r][col] / A[col][col];
      for (int l = col; l <= K; l++) A[r][l] -= f * A[col][l];
    }
  }
  for (int r = K - 1; r >= 0; r--) {
    long double v = A[r][K];
    for (int l = r + 1; l < K; l++) v -= A[r][l] * a[l];
    a[r] = A[r][r] != 0 ? v / A[r][r] : 0.0L;
  }
  for (int j = 0; j < K; j++) a[j] *= d[j];
}

enum { HIST_GRAM = 0, HIST_LINCOMB = 1, HIST_MISFIT = 2 };

template <int K>
int hist_launch(pmg_solver *S, const HistPtrs &F, const HistPtrs &U, double *dc, int what) {
  Level &Lf = S->lev[S->nlev - 1];
  constexpr int M = K * (K + 1) / 2 + K + 1;
  if (what == HIST_MISFIT) {
    hipLaunchKernelGGL(k_hist_misfit<K>, dim3(kHistBlocks), dim3(256), 0, S->stream, F, Lf.d.rho, dc, Lf.size,
                       S->d_hist);
    PMG_HOST_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_hist_fold, dim3(2), dim3(256), 0, S->stream, S->d_hist, kHistBlocks, 2,
                       S->d_hist + (size_t)kHistBlocks * kHistM);
  } else if (what == HIST_GRAM) {
    hipLaunchKernelGGL(k_hist_gram<K>, dim3(kHistBlocks), dim3(256), 0, S->stream, F, Lf.d.rho, Lf.size, S->d_hist);
    PMG_HOST_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_hist_fold, dim3(M), dim3(256), 0, S->stream, S->d_hist, kHistBlocks, M,
                       S->d_hist + (size_t)kHistBlocks * kHistM);
  } else {
    hipLaunchKernelGGL(k_hist_lincomb<K>, dim3(nblk256(Lf.size)), dim3(256), 0, S->stream, Lf.d.u, Lf.d.ua, U, dc,
                       Lf.size);
  }
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int hist_dispatch(pmg_solver *S, int K, const HistPtrs &F, const HistPtrs &U, double *dc, int what) {
  switch (K) {
    case 1: return hist_launch<1>(S, F, U, dc, what);
    case 2: return hist_launch<2>(S, F, U, dc, what);
    case 3: return hist_launch<3>(S, F, U, dc, what);
    case 4: return hist_launch<4>(S, F, U, dc, what);
    case 5: return hist_launch<5>(S, F, U, dc, what);
    case 6: return hist_launch<6>(S, F, U, dc, what);
    case 7: return hist_launch<7>(S, F, U, dc, what);
    case 8: return hist_launch<8>(S, F, U, dc, what);
  }
  return -1;
}

}  // namespace

extern "C" {

int pmg_history_warm_start(pmg_solver *S, int K, const double *const *F, const double *const *U, double *c_out,
                           double *rel_out) {
  if (!S || K < 1 || K > kHistMax || !F || !U) return -1;
  for (int k = 0; k < K; k++)
    if (!F[k] || !U[k]) return -1;
  if (!S->d_hist) {
    PMG_HOST_CHECK(hipMalloc(&S->d_hist, sizeof(double) * ((size_t)kHistBlocks * kHistM + kHistM + kHistMax)));
    PMG_HOST_CHECK(hipHostMalloc(&S->h_hist, sizeof(double) * (kHistM + kHistMax), hipHostMallocDefault));
  }
  HistPtrs hf{}, hu{};
  for (int k = 0; k < K; k++) {
    hf.p[k] = F[k];
    hu.p[k] = U[k];
  }
  const int M = K * (K + 1) / 2 + K + 1;
  double *d_fold = S->d_hist + (size_t)kHistBlocks * kHistM;
  if (hist_dispatch(S, K, hf, hu, nullptr, HIST_GRAM)) return -1;
  PMG_HOST_CHECK(hipMemcpyAsync(S->h_hist, d_fold, sizeof(double) * M, hipMemcpyDeviceToHost, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  const double *Gup = S->h_hist, *g = S->h_hist + K * (K + 1) / 2;
  long double a[kHistMax];
  hist_solve(K, Gup, g, a);
  // back to the F basis: B_j = sum_i (-1)^i C(j,i) F_i
  long double c[kHistMax];
  for (int i = 0; i < K; i++) {
    c[i] = 0.0L;
    for (int j = i; j < K; j++) c[i] += a[j] * (long double)(((i & 1) ? -1 : 1) * binom_c(j, i));
  }
  double *hc = S->h_hist + kHistM;
  for (int k = 0; k < K; k++) hc[k] = (double)c[k];
  double *dc = d_fold + kHistM;
  PMG_HOST_CHECK(hipMemcpyAsync(dc, hc, sizeof(double) * K, hipMemcpyHostToDevice, S->stream));
  if (hist_dispatch(S, K, hf, hu, dc, HIST_LINCOMB)) return -1;
  if (rel_out) {  // the misfit measured directly (one more pass over the sources)
    if (hist_dispatch(S, K, hf, hu, dc, HIST_MISFIT)) return -1;
    PMG_HOST_CHECK(hipMemcpyAsync(S->h_hist, d_fold, sizeof(double) * 2, hipMemcpyDeviceToHost, S->stream));
  }
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  if (rel_out) *rel_out = S->h_hist[1] > 0 ? std::sqrt(S->h_hist[0] / S->h_hist[1]) : 0.0;
  if (c_out)
    for (int k = 0; k < K; k++) c_out[k] = hc[k];
  return 0;
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Flexible GMRES(m) on the finest level, right-preconditioned by one FP32
// V-cycle from zero (Y. Saad, "A flexible inner-outer preconditioned GMRES
// algorithm", SIAM J. Sci. Comput. 14 (1993) 461-469).  The operator is the
// linear part of the FP64 residual hook: A v = R(v) with a zero source (the
// residual kernel with rho = 0; excised and boundary points give 0, so every
// Krylov vector stays on the unknowns).  The preconditioner is the mixed
// mode's inner cycle: writing v as the FP32 source and cycling from zero
// gives z ~ A^{-1} v.  Orthogonalisation: classical Gram-Schmidt applied
// twice, each pass one fused kernel over all basis vectors (dots, then the
// update with the new norm).  All reductions are per block in a fixed grid
// and folded in block order: deterministic.
// ---------------------------------------------------------------------------
constexpr int kKryMax = 16;
constexpr int kKryBlocks = 2048;

struct VecPtrs {
  const double *p[kKryMax + 1];
};

// partials[b][i] = sum_q w[q] V_i[q], i < nv
__global__ __launch_bounds__(256) void k_kry_dots(const double *w, VecPtrs V, int nv, size_t n, double *partials) {
  double acc[kKryMax + 1];
#pragma unroll
  for (int i = 0; i <= kKryMax; i++) acc[i] = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    const double wv = w[q];
#pragma unroll
    for (int i = 0; i <= kKryMax; i++)
      if (i < nv) acc[i] += wv * V.p[i][q];
  }
  __shared__ double red[256 / kWave][kKryMax + 1];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i <= kKryMax; i++) {
    double v = acc[i];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    if (lane == 0) red[wid][i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nv; i += 256) {
    double v = 0.0;
    for (int w2 = 0; w2 < 256 / kWave; w2++) v += red[w2][i];
    partials[(size_t)blockIdx.x * nv + i] = v;
  }
}

// w -= sum_i c_i V_i (c on the device); partials[b] = sum of the new w^2
__global__ __launch_bounds__(256) void k_kry_update(double *w, VecPtrs V, int nv, const double *c, size_t n,
                                                    double *partials) {
  double cc[kKryMax + 1];
#pragma unroll
  for (int i = 0; i <= kKryMax; i++) cc[i] = i < nv ? c[i] : 0.0;
  double s = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    double v = w[q];
#pragma unroll
    for (int i = 0; i <= kKryMax; i++)
      if (i < nv) v -= cc[i] * V.p[i][q];
    w[q] = v;
    s += v * v;
  }
  double cnt = 0.0;
  block_reduce2<256>(s, cnt);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// w -= sum_i c_i V_i, and partials[b][i] = sum_q V_i[q] w_new[q] (i < nv),
// partials[b][nv] = sum w_new^2: the first Gram-Schmidt update fused with the
// dot products of the second pass
__global__ __launch_bounds__(256) void k_kry_update_dots(double *w, VecPtrs V, int nv, const double *c, size_t n,
                                                         double *partials) {
  double cc[kKryMax + 1], acc[kKryMax + 2];
#pragma unroll
  for (int i = 0; i <= kKryMax; i++) cc[i] = i < nv ? c[i] : 0.0;
#pragma unroll
  for (int i = 0; i <= kKryMax + 1; i++) acc[i] = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    double vi[kKryMax + 1];
    double v = w[q];
#pragma unroll
    for (int i = 0; i <= kKryMax; i++) {
      vi[i] = i < nv ? V.p[i][q] : 0.0;
      v -= cc[i] * vi[i];
    }
    w[q] = v;
#pragma unroll
    for (int i = 0; i <= kKryMax; i++) acc[i] += vi[i] * v;
    acc[kKryMax + 1] += v * v;
  }
  __shared__ double red[256 / kWave][kKryMax + 2];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i <= kKryMax + 1; i++) {
    double v = acc[i];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    if (lane == 0) red[wid][i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= nv; i += 256) {
    const int src = i < nv ? i : kKryMax + 1;
    double v = 0.0;
    for (int w2 = 0; w2 < 256 / kWave; w2++) v += red[w2][src];
    partials[(size_t)blockIdx.x * (nv + 1) + i] = v;
  }
}

__global__ __launch_bounds__(256) void k_kry_scale(double *dst, const double *src, double f, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = f * src[q];
}

// preconditioner input: FP32 source = v, FP32 solution = 0
__global__ __launch_bounds__(256) void k_kry_prep(float *rho32, float *u32, const double *v, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) {
    rho32[q] = (float)v[q];
    u32[q] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_kry_to64(double *dst, const float *src, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = (double)src[q];
}

// u (both sweep buffers) += sum_j y_j Z_j
__global__ __launch_bounds__(256) void k_kry_xupdate(double *u, double *ua, VecPtrs Z, int nz, const double *y,
                                                     size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  double v = u[q];
  for (int j = 0; j < nz; j++) v += y[j] * Z.p[j][q];
  u[q] = v;
  ua[q] = v;
}

int enq_prec(pmg_solver *S) {
  return enqueue_vcycle<float>(S, false, NS_F32, S->d_norms + NS_COARSE + 2, S->d_citer + 1);
}

// FP64 residual hook of an arbitrary u with source rho into res (finest level)
int resid_vec(pmg_solver *S, const double *u, const double *rho, double *res, int slot) {
  Level &Lv = S->lev[S->nlev - 1];
  LevelArrays<double> &A = arrays<double>(Lv);
  ResidP<double> P;
  P.u = u; P.rho = rho; P.rhs = A.rhs; P.a = A.a; P.c = A.c; P.chr = Lv.chr;
  P.res = res;
  P.neg32 = nullptr;
  P.nx = Lv.n[0]; P.ny = Lv.n[1]; P.nz = Lv.n[2];
  if (Lv.rbox.ok) {
    P.bx0 = Lv.rbox.lo[0]; P.bx1 = Lv.rbox.hi[0];
    P.by0 = Lv.rbox.lo[1]; P.by1 = Lv.rbox.hi[1];
    P.bz0 = Lv.rbox.lo[2]; P.bz1 = Lv.rbox.hi[2];
  } else {
    P.bx0 = P.by0 = P.bz0 = 1;
    P.bx1 = P.by1 = P.bz1 = 0;
  }
  P.sp = spacing<double>(Lv);
  P.ws = S->ws;
  P.slot = slot;
  const bool p2 = P.sp.pow2 != 0;
#define PMG_RV(OPV, EXV) \
  (p2 ? launch_residual_t<double, OPV, EXV, true>(S, P, true, true) : launch_residual_t<double, OPV, EXV, false>(S, P, true, true))
  if (S->op == 0) return S->do_ex ? PMG_RV(0, true) : PMG_RV(0, false);
  return S->do_ex ? PMG_RV(1, true) : PMG_RV(1, false);
#undef PMG_RV
}

int kry_fold(pmg_solver *S, int M, double *host) {
  double *fold = S->kry_red + (size_t)kKryBlocks * (kKryMax + 1);
  hipLaunchKernelGGL(k_hist_fold, dim3(M), dim3(256), 0, S->stream, S->kry_red, kKryBlocks, M, fold);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipMemcpyAsync(host, fold, sizeof(double) * M, hipMemcpyDeviceToHost, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int kry_dots(pmg_solver *S, const double *w, const VecPtrs &V, int nv, double *host) {
  Level &F = S->lev[S->nlev - 1];
  hipLaunchKernelGGL(k_kry_dots, dim3(kKryBlocks), dim3(256), 0, S->stream, w, V, nv, F.size, S->kry_red);
  PMG_HOST_CHECK(hipGetLastError());
  return kry_fold(S, nv, host);
}

int kry_update(pmg_solver *S, double *w, const VecPtrs &V, int nv, const double *c_host, double *norm2) {
  Level &F = S->lev[S->nlev - 1];
  double *dc = S->kry_red + (size_t)kKryBlocks * (kKryMax + 1) + (kKryMax + 1);
  double *hc = S->kry_host + 2 * (kKryMax + 1);
  for (int i = 0; i < nv; i++) hc[i] = c_host[i];
  PMG_HOST_CHECK(hipMemcpyAsync(dc, hc, sizeof(double) * nv, hipMemcpyHostToDevice, S->stream));
  hipLaunchKernelGGL(k_kry_update, dim3(kKryBlocks), dim3(256), 0, S->stream, w, V, nv, dc, F.size, S->kry_red);
  PMG_HOST_CHECK(hipGetLastError());
  return kry_fold(S, 1, norm2);
}

// w -= V c_host; h_out[0..nv-1] = V^T w_new, *norm2 = w_new . w_new
int kry_update_dots(pmg_solver *S, double *w, const VecPtrs &V, int nv, const double *c_host, double *h_out,
                    double *norm2) {
  Level &F = S->lev[S->nlev - 1];
  double *dc = S->kry_red + (size_t)kKryBlocks * (kKryMax + 1) + (kKryMax + 1);
  double *hc = S->kry_host + 2 * (kKryMax + 1);
  for (int i = 0; i < nv; i++) hc[i] = c_host[i];
  PMG_HOST_CHECK(hipMemcpyAsync(dc, hc, sizeof(double) * nv, hipMemcpyHostToDevice, S->stream));
  hipLaunchKernelGGL(k_kry_update_dots, dim3(kKryBlocks), dim3(256), 0, S->stream, w, V, nv, dc, F.size,
                     S->kry_red);
  PMG_HOST_CHECK(hipGetLastError());
  double tmp[kKryMax + 2];
  int rc = kry_fold(S, nv + 1, tmp);
  if (rc) return rc;
  for (int i = 0; i < nv; i++) h_out[i] = tmp[i];
  *norm2 = tmp[nv];
  return 0;
}

int ensure_krylov(pmg_solver *S, int m) {
  if (S->kry && S->kry_m >= m) return 0;
  if (S->kry) { PMG_HOST_CHECK(hipStreamSynchronize(S->stream)); (void)hipFree(S->kry); S->kry = nullptr; }
  const size_t n = S->lev[S->nlev - 1].size;
  const size_t nvec = (size_t)(2 * m + 2);
  PMG_HOST_CHECK(hipMalloc(&S->kry, nvec * n * sizeof(double)));
  PMG_HOST_CHECK(hipMemsetAsync(S->kry, 0, nvec * n * sizeof(double), S->stream));
  if (!S->kry_red) {
    PMG_HOST_CHECK(hipMalloc(&S->kry_red, sizeof(double) * ((size_t)kKryBlocks * (kKryMax + 1) + 4 * (kKryMax + 1))));
    PMG_HOST_CHECK(hipHostMalloc(&S->kry_host, sizeof(double) * 4 * (kKryMax + 1), hipHostMallocDefault));
  }
  S->kry_m = m;
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

}  // namespace

extern "C" {

int pmg_solve_fgmres(pmg_solver *S, int m, double rtol, int max_iter, double *hist, int *iters, double *r0_out) {
  if (!S || m < 1 || m > kKryMax - 1 || max_iter < 0) return -1;
  int rc;
  if ((rc = ensure_fp32(S))) return rc;
  if ((rc = ensure_krylov(S, m))) return rc;
  Level &F = S->lev[S->nlev - 1];
  const size_t n = F.size;
  const unsigned nb = nblk256(n);
  double *V[kKryMax + 1], *Z[kKryMax];
  for (int i = 0; i <= m; i++) V[i] = S->kry + (size_t)i * n;
  for (int i = 0; i < m; i++) Z[i] = S->kry + (size_t)(m + 1 + i) * n;
  double *w = S->kry + (size_t)(2 * m + 1) * n;
  double *hh = S->kry_host;  // [0, kKryMax]: dots / norms
  auto true_residual = [&](double *rms) -> int {
    int r;
    if ((r = resid_vec(S, ucur<double>(F), F.d.rho, w, NS_R0))) return r;
    if ((r = fold_norms(S, NS_R0, 1))) return r;
    PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_R0, sizeof(double), hipMemcpyDeviceToHost, S->stream));
    PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
    *rms = S->h_pinned[0];
    return 0;
  };
  double r_rms;
  if ((rc = true_residual(&r_rms))) return rc;
  const double rref = (r0_out && *r0_out > 0) ? *r0_out : r_rms;
  if (r0_out) *r0_out = rref;
  int it = 0;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  // learned band correction of every preconditioner output (excision): on
  // once a plain solve has measured the plain rate, kept while a solve with
  // it reduces the residual faster per unit time than the plain solves do
  auto &B = S->band;
  pmg_net *act = B.inner ? B.net_in : B.net;
  // use_fgmres 2: always on (a measurement of the correction's effect, no policy)
  bool band_on = B.use_fgmres && act && B.ncore > 0 && (B.fg_rate_plain > 0 || B.use_fgmres == 2);
  if (band_on && B.use_fgmres != 2 && B.fg_skip > 0) {
    B.fg_skip--;
    band_on = false;
  }
  if (band_on && (rc = pmg_net_reserve(act, (long long)B.ncore * B.side * B.side * B.side))) return rc;
  if (band_on && B.inner) {
    // the preconditioner's source is a Krylov vector of unit 2-norm: its RMS
    // over the updatable points scales the in-cycle correction
    S->h_pinned[1] = B.n_upd > 0 ? 1.0 / std::sqrt((double)B.n_upd) : 1.0;
    PMG_HOST_CHECK(hipMemcpyAsync(S->d_norms + NS_PSRC, S->h_pinned + 1, sizeof(double), hipMemcpyHostToDevice,
                                  S->stream));
    PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  }
  const double r_start = r_rms;
  const auto t_start = std::chrono::steady_clock::now();
  while (it < max_iter && r_rms > rtol * rref) {
    // b = -R(u) (in w): 2-norm, and the RMS-to-2-norm factor of this grid
    VecPtrs Pw{};
    Pw.p[0] = w;
    if ((rc = kry_dots(S, w, Pw, 1, hh))) return rc;
    const double beta = std::sqrt(hh[0]);
    if (!(beta > 0)) break;
    const double scale = beta / r_rms;
    hipLaunchKernelGGL(k_kry_scale, dim3(nb), dim3(256), 0, S->stream, V[0], w, -1.0 / beta, n);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    const double target = rtol * rref * scale;
    int k = 0;
    for (int j = 0; j < m && it < max_iter; j++) {
      // Z_j = M V_j: the FP32 V-cycle from zero with V_j as its source
      hipLaunchKernelGGL(k_kry_prep, dim3(nb), dim3(256), 0, S->stream, F.f.rho, F.f.u, V[j], n);
      PMG_HOST_CHECK(hipGetLastError());
      // prec_cycles FP32 V-cycles on the source V_j, the first from zero;
      // the band correction (when on) after a single cycle, or between
      // cycles, where the next cycle smooths what it adds
      for (int cyc = 0; cyc < S->prec_cycles; cyc++) {
        if (cyc > 0 && band_on && !B.inner && (rc = capture_or_run(S, &B.gpc, enq_band_pc))) return rc;
        if (band_on && B.inner) {
          if ((rc = capture_or_run(S, &B.gpreci, enq_prec_inner))) return rc;
        } else if ((rc = capture_or_run(S, &S->gprec, enq_prec))) {
          return rc;
        }
      }
      if (band_on && !B.inner && S->prec_cycles == 1 && (rc = capture_or_run(S, &B.gpc, enq_band_pc))) return rc;
      hipLaunchKernelGGL(k_kry_to64, dim3(nb), dim3(256), 0, S->stream, Z[j], F.f.u, n);
      PMG_HOST_CHECK(hipGetLastError());
      // V_{j+1} = A Z_j, orthogonalised against V_0..V_j by classical
      // Gram-Schmidt, repeated when the first pass removed more than 1 - 1/2
      // of w's squared norm (the criterion of Daniel, Gragg, Kaufman and
      // Stewart, Math. Comp. 30 (1976) 772-795)
      if ((rc = resid_vec(S, Z[j], nullptr, V[j + 1], NS_OUTER + 1))) return rc;
      VecPtrs PV{};
      for (int i = 0; i <= j; i++) PV.p[i] = V[i];
      PV.p[j + 1] = V[j + 1];
      double h1[kKryMax + 2], h2[kKryMax + 1], nrm2 = 0.0;
      if ((rc = kry_dots(S, V[j + 1], PV, j + 2, h1))) return rc;
      const double before2 = h1[j + 1];
      PV.p[j + 1] = nullptr;
      if ((rc = kry_update_dots(S, V[j + 1], PV, j + 1, h1, h2, &nrm2))) return rc;
      const bool again = nrm2 < 0.5 * before2;
      if (again && (rc = kry_update(S, V[j + 1], PV, j + 1, h2, &nrm2))) return rc;
      double col[kKryMax + 1];
      for (int i = 0; i <= j; i++) col[i] = h1[i] + (again ? h2[i] : 0.0);
      col[j + 1] = std::sqrt(nrm2);
      if (col[j + 1] > 0)
        hipLaunchKernelGGL(k_kry_scale, dim3(nb), dim3(256), 0, S->stream, V[j + 1], V[j + 1], 1.0 / col[j + 1], n);
      // Givens rotations
      for (int i = 0; i < j; i++) {
        const double t = cs[i] * col[i] + sn[i] * col[i + 1];
        col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
        col[i] = t;
      }
      const double den = std::hypot(col[j], col[j + 1]);
      cs[j] = den > 0 ? col[j] / den : 1.0;
      sn[j] = den > 0 ? col[j + 1] / den : 0.0;
      col[j] = den;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = col[i];
      k = j + 1;
      if (hist) hist[it] = std::fabs(g[j + 1]) / scale;
      it++;
      if (std::fabs(g[j + 1]) <= target || !(col[j] > 0) || nrm2 == 0.0) break;
    }
    // y = H^{-1} g, u += Z y
    for (int i = k - 1; i >= 0; i--) {
      double v = g[i];
      for (int l = i + 1; l < k; l++) v -= H[(size_t)i * m + l] * y[l];
      y[i] = H[(size_t)i * m + i] != 0 ? v / H[(size_t)i * m + i] : 0.0;
    }
    VecPtrs PZ{};
    for (int i = 0; i < k; i++) {
      PZ.p[i] = Z[i];
      hh[kKryMax + 1 + i] = y[i];
    }
    double *dy = S->kry_red + (size_t)kKryBlocks * (kKryMax + 1) + 2 * (kKryMax + 1);
    PMG_HOST_CHECK(hipMemcpyAsync(dy, hh + kKryMax + 1, sizeof(double) * k, hipMemcpyHostToDevice, S->stream));
    hipLaunchKernelGGL(k_kry_xupdate, dim3(nb), dim3(256), 0, S->stream, F.d.u, F.d.ua, PZ, k, dy, n);
    PMG_HOST_CHECK(hipGetLastError());
    if ((rc = true_residual(&r_rms))) return rc;
    if (hist && it > 0) hist[it - 1] = r_rms;
  }
  if ((B.net || B.net_in) && B.ncore > 0 && it > 0 && r_rms > 0 && r_start > r_rms) {
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    const double rate = std::log10(r_start / r_rms) / ms;
    if (!band_on) {
      B.fg_rate_plain = B.fg_rate_plain > 0 ? 0.5 * (B.fg_rate_plain + rate) : rate;
    } else {
      B.fg_rate_net = B.fg_rate_net > 0 ? 0.5 * (B.fg_rate_net + rate) : rate;
      if (rate > B.fg_rate_plain) {
        B.accepted++;
      } else {
        B.rejected++;
        B.fg_skip = 4;  // plain solves re-measure the plain rate before the next trial
      }
    }
  }
  if (iters) *iters = it;
  return 0;
}

int pmg_set_fgmres_prec_cycles(pmg_solver *S, int cycles) {
  if (!S || cycles < 1 || cycles > 4) return -1;
  S->prec_cycles = cycles;
  return 0;
}

int pmg_zero_solution(pmg_solver *S) {
  // initial guess u = 0 on the finest level (both sweep buffers); coarser u
  // are overwritten by injection in every cycle
  if (!S) return -1;
  Level &F = S->lev[S->nlev - 1];
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk256(F.size)), dim3(256), 0, S->stream, F.d.u, 0.0, F.size);
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk256(F.size)), dim3(256), 0, S->stream, F.d.ua, 0.0, F.size);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_time_kernel(pmg_solver *S, int which, int L, int prec, int nrep, double *ms_avg) {
  // Average duration of one launch of a solver kernel, timed with HIP
  // events on the solver's stream over nrep back-to-back launches on the
  // current data.  which: 0 = red-black sweep (ping-pong pairs, so u returns
  // to its buffer), 1 = residual.  prec: 0 FP64, 1 FP32.
  if (!S || L < 1 || L > S->nlev || nrep < 2 || !ms_avg) return -1;
  if (prec == 1 && ensure_fp32(S)) return -1;
  if (which == 3) {
    // the finest FP64 sweep inside nrep ungraphed V-cycles: events around
    // each pre- and post-smoothing group, the average per launch
    if (prec != 0 || L != S->nlev) return -1;
    std::vector<hipEvent_t> evs((size_t)4 * nrep, nullptr);
    int rc = 0;
    for (auto &e : evs)
      if (hipEventCreate(&e) != hipSuccess) { rc = -1; break; }
    const int keep = S->use_graph;
    if (!rc) {
      S->tk_ev = evs.data();
      S->tk_cap = (int)evs.size();
      S->tk_n = S->tk_sweeps = 0;
      S->use_graph = 0;
      for (int r = 0; r < nrep && !rc; r++) rc = enq_v64(S);
      S->use_graph = keep;
      if (!rc && hipStreamSynchronize(S->stream) != hipSuccess) rc = -1;
      double tot = 0.0;
      for (int q = 0; !rc && q + 1 < S->tk_n; q += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, evs[q], evs[q + 1]) != hipSuccess) rc = -1;
        tot += ms;
      }
      if (!rc) *ms_avg = S->tk_sweeps > 0 ? tot / S->tk_sweeps : 0.0;
    }
    S->tk_ev = nullptr;
    S->tk_cap = S->tk_n = S->tk_sweeps = 0;
    for (auto &e : evs)
      if (e) (void)hipEventDestroy(e);
    return rc;
  }
  EventPair ev;
  PMG_HOST_CHECK(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  nrep -= nrep % 2;
  const bool fine = (L == S->nlev);
  int rc = 0;
  PMG_HOST_CHECK(hipEventRecord(e0, S->stream));
  if (which == 2) {
    // the finest residual fused with the full-weight injection into L-1
    if (!fine || !(prec ? can_fuse_resid_restrict<float>(S) : can_fuse_resid_restrict<double>(S))) return -1;
  }
  if (which == 0) {
    rc = prec ? sweeps<float>(S, L, nrep, NS_F32 + NS_RELAX + L - 1, fine)
              : sweeps<double>(S, L, nrep, NS_RELAX + L - 1, fine);
  } else if (which == 2) {
    for (int r = 0; r < nrep && !rc; r++)
      rc = prec ? resid_restrict<float>(S, NS_F32 + NS_RESID + L - 1) : resid_restrict<double>(S, NS_RESID + L - 1);
  } else {
    for (int r = 0; r < nrep && !rc; r++)
      rc = prec ? residual<float>(S, L, NS_F32 + NS_RESID + L - 1, fine, fine)
                : residual<double>(S, L, NS_RESID + L - 1, fine, fine);
  }
  PMG_HOST_CHECK(hipEventRecord(e1, S->stream));
  if (!rc && which == 0) rc = prec ? settle<float>(S, L) : settle<double>(S, L);
  PMG_HOST_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  PMG_HOST_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *ms_avg = (double)ms / nrep;
  return rc;
}

// ---------------------------------------------------------------------------
// Learned band correction (excision; no reference counterpart)
// ---------------------------------------------------------------------------
int pmg_band_setup(pmg_solver *S, int depth, int core) {
  if (!S || !S->do_ex || depth < 1 || depth > 16 || core < 4 || core > 64) return -1;
  Level &F = S->lev[S->nlev - 1];
  auto &B = S->band;
  const int nx = F.n[0], ny = F.n[1], nz = F.n[2];
  std::vector<double> chr(F.size);
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  PMG_HOST_CHECK(hipMemcpy(chr.data(), F.chrd, F.size * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<unsigned char> band(F.size);
  const long long np = pmg_band_mask(chr.data(), nx, ny, nz, S->ex, depth, band.data());
  if (np < 0) return -1;
  const int margin = 3;
  const int maxc = ((nx + core - 1) / core) * ((ny + core - 1) / core) * ((nz + core - 1) / core);
  std::vector<int> org(3 * (size_t)maxc);
  const int nc = pmg_band_cores(band.data(), nx, ny, nz, core, margin, org.data(), maxc);
  if (nc < 0) return -1;
  if (B.d_org) PMG_HOST_CHECK(hipFree(B.d_org));
  if (B.d_band) PMG_HOST_CHECK(hipFree(B.d_band));
  if (B.d_x) PMG_HOST_CHECK(hipFree(B.d_x));
  B.d_org = nullptr;
  B.d_band = nullptr;
  B.d_x = nullptr;
  const int side = core + 2 * margin;
  PMG_HOST_CHECK(hipMalloc(&B.d_band, F.size));
  PMG_HOST_CHECK(hipMemcpy(B.d_band, band.data(), F.size, hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMalloc(&B.d_org, sizeof(int) * 3 * (size_t)std::max(nc, 1)));
  if (nc > 0) {
    PMG_HOST_CHECK(hipMemcpy(B.d_org, org.data(), sizeof(int) * 3 * (size_t)nc, hipMemcpyHostToDevice));
    PMG_HOST_CHECK(hipMalloc(&B.d_x, sizeof(float) * (size_t)nc * side * side * side));
  }
  long long nupd = 0;
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++)
        if (chr[ix3(i, j, k, nx, ny)] != S->ex) nupd++;
  B.n_upd = nupd;
  B.h_pts.clear();
  B.h_pts.reserve((size_t)np);
  for (size_t p = 0; p < F.size; p++)
    if (band[p]) B.h_pts.push_back((long long)p);
  B.depth = depth;
  B.core = core;
  B.side = side;
  B.margin = margin;
  B.ncore = nc;
  B.npts = np;
  band_graphs_reset(S);
  return nc;
}

int pmg_band_info(pmg_solver *S, long long *npts, int *ncore, int *side) {
  if (!S) return -1;
  if (npts) *npts = S->band.npts;
  if (ncore) *ncore = S->band.ncore;
  if (side) *side = S->band.side;
  return 0;
}

int pmg_band_download(pmg_solver *S, unsigned char *host) {
  if (!S || !host || !S->band.d_band) return -1;
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  PMG_HOST_CHECK(hipMemcpy(host, S->band.d_band, S->lev[S->nlev - 1].size, hipMemcpyDeviceToHost));
  return 0;
}

int pmg_set_band_net(pmg_solver *S, pmg_net *N, double ts, int use_dc, int use_fgmres) {
  if (!S || (N && !(ts > 0))) return -1;
  auto &B = S->band;
  B.net = N;
  B.ts = ts;
  B.use_dc = use_dc;
  B.use_fgmres = use_fgmres;
  B.rho_net = B.t_net = -1;
  B.fg_rate_net = 0;
  B.fg_skip = 0;
  B.accepted = B.rejected = 0;
  band_graphs_reset(S);
  return 0;
}

int pmg_set_band_net_inner(pmg_solver *S, pmg_net *N, double ts) {
  if (!S || (N && !(ts > 0))) return -1;
  S->band.net_in = N;
  S->band.ts_in = ts;
  band_graphs_reset(S);
  return 0;
}

int pmg_band_set_inner(pmg_solver *S, int on) {
  if (!S) return -1;
  S->band.inner = on ? 1 : 0;
  return 0;
}

int pmg_band_sample_inner(pmg_solver *S, const double *ustar, int ncrop, unsigned long long seed, float *x, float *t,
                          float *w, double *rnorm) {
  // one mixed iteration (no graph) whose FP32 cycle, after the finest
  // coarse-grid correction, writes training crops centred on seeded random
  // band points: x = r32 / |r|, t = (ustar - u - u32) / |r| (the error of the
  // cycle's correction at that moment), w = band; |r| = the outer defect's
  // RMS.  Needs the state a mode-1 pmg_solve_rel leaves (FP32 source -r).
  // *rnorm: the outer defect's RMS after the iteration.
  if (!S || !ustar || !x || !t || !w || ncrop < 1 || !S->fp32_ready) return -1;
  auto &B = S->band;
  if (B.h_pts.empty()) return -1;
  Level &F = S->lev[S->nlev - 1];
  const int nx = F.n[0], ny = F.n[1];
  std::vector<int> org(3 * (size_t)ncrop);
  unsigned long long z = seed;
  for (int c = 0; c < ncrop; c++) {
    z += 0x9E3779B97F4A7C15ull;  // splitmix64
    unsigned long long v = z;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    const long long p = B.h_pts[(size_t)(v % B.h_pts.size())];
    org[3 * c] = (int)(p % nx) - B.side / 2;
    org[3 * c + 1] = (int)((p / nx) % ny) - B.side / 2;
    org[3 * c + 2] = (int)(p / ((long long)nx * ny)) - B.side / 2;
  }
  int *dorg = nullptr;
  PMG_HOST_CHECK(hipMalloc(&dorg, sizeof(int) * org.size()));
  PMG_HOST_CHECK(hipMemcpyAsync(dorg, org.data(), sizeof(int) * org.size(), hipMemcpyHostToDevice, S->stream));
  B.samp_ustar = ustar;
  B.samp_org = dorg;
  B.samp_n = ncrop;
  B.samp_x = x;
  B.samp_t = t;
  B.samp_w = w;
  B.enq_inner = true;
  B.cur_norm = S->d_norms + NS_OUTER;
  int rc = enq_mixed(S);
  B.enq_inner = false;
  B.samp_x = B.samp_t = B.samp_w = nullptr;
  B.samp_ustar = nullptr;
  B.samp_org = nullptr;
  PMG_HOST_CHECK(hipMemcpyAsync(S->h_pinned, S->d_norms + NS_OUTER, sizeof(double), hipMemcpyDeviceToHost,
                                S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  PMG_HOST_CHECK(hipFree(dorg));
  if (rnorm) *rnorm = S->h_pinned[0];
  return rc;
}

int pmg_band_enable(pmg_solver *S, int use_dc, int use_fgmres) {
  if (!S) return -1;
  S->band.use_dc = use_dc;
  S->band.use_fgmres = use_fgmres;
  return 0;
}

int pmg_band_stats(pmg_solver *S, int *accepted, int *rejected, double *out6) {
  if (!S) return -1;
  const auto &B = S->band;
  if (accepted) *accepted = B.accepted;
  if (rejected) *rejected = B.rejected;
  if (out6) {
    out6[0] = B.rho_plain;
    out6[1] = B.t_plain;
    out6[2] = B.rho_net;
    out6[3] = B.t_net;
    out6[4] = B.fg_rate_plain;
    out6[5] = B.fg_rate_net;
  }
  return 0;
}

int pmg_band_sample(pmg_solver *S, const double *ustar, double xs, double ts, int ncrop, unsigned long long seed,
                    float *x, float *t, float *w) {
  // training crops centred on random band points of the current state:
  // x = xs * (FP32 finest source), t = ts * (ustar - u), w = band indicator
  if (!S || !ustar || !x || !t || !w || ncrop < 1) return -1;
  auto &B = S->band;
  if (B.h_pts.empty() || !S->fp32_ready) return -1;
  Level &F = S->lev[S->nlev - 1];
  const int nx = F.n[0], ny = F.n[1];
  std::vector<int> org(3 * (size_t)ncrop);
  unsigned long long z = seed;
  for (int c = 0; c < ncrop; c++) {
    // splitmix64 (S. Vigna): one draw per crop
    z += 0x9E3779B97F4A7C15ull;
    unsigned long long v = z;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    const long long p = B.h_pts[(size_t)(v % B.h_pts.size())];
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long long)nx * ny));
    org[3 * c] = i - B.side / 2;
    org[3 * c + 1] = j - B.side / 2;
    org[3 * c + 2] = k - B.side / 2;
  }
  int *dorg = nullptr;
  PMG_HOST_CHECK(hipMalloc(&dorg, sizeof(int) * org.size()));
  PMG_HOST_CHECK(hipMemcpyAsync(dorg, org.data(), sizeof(int) * org.size(), hipMemcpyHostToDevice, S->stream));
  int rc = pmg_band_pairs(F.f.rho, (float)xs, ustar, F.d.u, ts, B.d_band, F.n[0], F.n[1], F.n[2], dorg, ncrop,
                          B.side, x, t, w, S->stream);
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  PMG_HOST_CHECK(hipFree(dorg));
  return rc;
}

int pmg_band_correct(pmg_solver *S, int which) {
  // one band correction now: which 0 = the defect-correction form (FP64 u,
  // from the FP32 source -r and the outer norm, then -r updated), 1 = the
  // preconditioner form (FP32 u from its FP32 residual)
  if (!S || !S->band.net || S->band.ncore < 1 || !S->fp32_ready) return -1;
  auto &B = S->band;
  int rc;
  if ((rc = pmg_net_reserve(B.net, (long long)B.ncore * B.side * B.side * B.side))) return rc;
  rc = which == 0 ? enq_band_dc(S) : enq_band_pc(S);
  if (rc) return rc;
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

// FP32 arrays of a level (mixed precision): which as pmg_field (u, rho, rhs, res, lop)
int pmg_download32(pmg_solver *S, int L, int which, float *host) {
  if (!S || L < 1 || L > S->nlev || !host || !S->fp32_ready) return -1;
  Level &Lv = S->lev[L - 1];
  const float *src = nullptr;
  switch (which) {
    case 0: src = ucur<float>(Lv); break;
    case 1: src = Lv.f.rho; break;
    case 2: src = Lv.f.rhs; break;
    case 3: src = Lv.f.res; break;
    case 4: src = Lv.f.lop; break;
    default: return -1;
  }
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  PMG_HOST_CHECK(hipMemcpy(host, src, Lv.size * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pmg_upload32(pmg_solver *S, int L, int which, const float *host) {
  if (!S || L < 1 || L > S->nlev || !host) return -1;
  if (ensure_fp32(S)) return -1;
  Level &Lv = S->lev[L - 1];
  float *dst = nullptr;
  switch (which) {
    case 0: dst = ucur<float>(Lv); break;
    case 1: dst = Lv.f.rho; break;
    default: return -1;
  }
  PMG_HOST_CHECK(hipMemcpyAsync(dst, host, Lv.size * sizeof(float), hipMemcpyHostToDevice, S->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(S->stream));
  return 0;
}

int pmg_last_citer(pmg_solver *S, int which) {
  if (!S) return -1;
  int ci[2] = {0, 0};
  if (hipMemcpy(ci, S->d_citer, sizeof(int) * 2, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return ci[which ? 1 : 0];
}

}  // extern "C"
