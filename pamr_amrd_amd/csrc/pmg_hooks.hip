// pmg_hooks.hip -- device implementations of the application MG hooks.
//
// These are the grid-function operators the AMRD driver calls per grid
// (include/amrd.h:137-159): relax (one red-black Gauss-Seidel sweep),
// residual and L_op.  They act where cmask == AMRD_CMASK_ON (1.0) on points
// 2..N-1 (1-based) and, when a chr array is given, skip excised points
// (chr == ex).  op 0 is the plain Poisson operator of the nps example
// (w/examples/nps/num_w.f:18-157), op 1 the variable-coefficient operator.
// Arrays are device pointers; cmask and chr are double grid functions as in
// the reference.  The norm (RMS of the residuals the hook evaluated) is
// written to a device double, so the caller decides when to synchronise.
#include <mutex>

#include "pmg_ops.h"
#include "../../include/pamr_mg_gpu.h"

using namespace pmg;

namespace {

constexpr int NT = 256;

// Partial-norm scratch of pmg_hook_relax_noalloc: a fixed device array of
// the library (one copy per GPU, held until the process ends) holding the
// two colour slots of (sum, count) partials, kScratchBlocks blocks each,
// then the two per-slot block counts.  Nothing in it needs clearing: every
// launch writes all partials it folds and block 0 writes its slot's count.
constexpr int kScratchBlocks = 1 << 19;
constexpr size_t kScratchPartials = 2 * 2 * (size_t)kScratchBlocks;
static __device__ __align__(256) double scratch[kScratchPartials + 2];

constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the device address of this GPU's scratch (looked up once per device, with
// that device current)
int scratch_address(double **out) {
  int dev = -1;
  PMG_HOST_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    std::fprintf(stderr, "pamr_mg: device %d beyond the %d the no-alloc hook supports\n", dev, kMaxDevices);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    PMG_HOST_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

struct HookGeom {
  int nx, ny, nz;
  size_t sy, sz;
};

__device__ __forceinline__ bool hook_on(const double *cmask, const double *chr, double ex,
                                        size_t p) {
  if (cmask[p] != 1.0) return false;
  if (chr && chr[p] == ex) return false;
  return true;
}

// one colour pass of the red-black sweep (num_w.f:130-152): pass 0 updates
// points with mod(i+j+k+pass,2)==0 in 1-based indices.
template <int OP, bool P2>
__global__ __launch_bounds__(NT) void k_hook_relax_pass(double *V, const double *rhs,
                                                        const double *rho, const double *a,
                                                        const double *c, const double *cmask,
                                                        const double *chr, double ex, HookGeom g,
                                                        Spacing<double> sp, int pass, NormWs ws,
                                                        int nb) {
  const int mi = g.nx - 2, mj = g.ny - 2, mk = g.nz - 2;
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  double s = 0.0, cnt = 0.0;
  if (q < (size_t)mi * mj * mk) {
    int i = 1 + (int)(q % mi), j = 1 + (int)((q / mi) % mj), k = 1 + (int)(q / ((size_t)mi * mj));
    size_t p = ix3(i, j, k, g.nx, g.ny);
    if (((i + j + k + 3 + pass) & 1) == 0 && hook_on(cmask, chr, ex, p)) {
      double jac;
      double res = lop_at<double, OP, P2>(V, rho, a, c, p, g.sy, g.sz, sp, &jac) - rhs[p];
      V[p] = V[p] - res / jac;
      s = res * res;
      cnt = 1.0;
    }
  }
  norm_partial<NT>(s, cnt, ws, pass, blockIdx.x, nb);
}

template <int OP, bool P2>
__global__ __launch_bounds__(NT) void k_hook_lop(double *LV, const double *V, const double *rho,
                                                 const double *a, const double *c,
                                                 const double *cmask, const double *chr,
                                                 double ex, HookGeom g, Spacing<double> sp) {
  const int mi = g.nx - 2, mj = g.ny - 2, mk = g.nz - 2;
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  if (q >= (size_t)mi * mj * mk) return;
  int i = 1 + (int)(q % mi), j = 1 + (int)((q / mi) % mj), k = 1 + (int)(q / ((size_t)mi * mj));
  size_t p = ix3(i, j, k, g.nx, g.ny);
  if (hook_on(cmask, chr, ex, p)) LV[p] = lop_at<double, OP, P2>(V, rho, a, c, p, g.sy, g.sz, sp, nullptr);
}

// residual (num_w.f:59-90): lop into res, then res -= rhs, norm
template <int OP, bool P2>
__global__ __launch_bounds__(NT) void k_hook_residual(double *res, const double *rhs,
                                                      const double *V, const double *rho,
                                                      const double *a, const double *c,
                                                      const double *cmask, const double *chr,
                                                      double ex, HookGeom g, Spacing<double> sp,
                                                      NormWs ws, int nb) {
  const int mi = g.nx - 2, mj = g.ny - 2, mk = g.nz - 2;
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  double s = 0.0, cnt = 0.0;
  if (q < (size_t)mi * mj * mk) {
    int i = 1 + (int)(q % mi), j = 1 + (int)((q / mi) % mj), k = 1 + (int)(q / ((size_t)mi * mj));
    size_t p = ix3(i, j, k, g.nx, g.ny);
    if (hook_on(cmask, chr, ex, p)) {
      double r = lop_at<double, OP, P2>(V, rho, a, c, p, g.sy, g.sz, sp, nullptr) - rhs[p];
      res[p] = r;
      s = r * r;
      cnt = 1.0;
    }
  }
  norm_partial<NT>(s, cnt, ws, 0, blockIdx.x, nb);
}

// Host helpers ---------------------------------------------------------------
bool is_pow2(double v) {
  int e;
  double m = std::frexp(v, &e);
  return m == 0.5 && v > 1e-300 && v < 1e300;
}

Spacing<double> make_spacing(double dx, double dy, double dz) {
  Spacing<double> sp;
  sp.dx2 = dx * dx;
  sp.dy2 = dy * dy;
  sp.dz2 = dz * dz;
  sp.idx2 = 1.0 / sp.dx2;
  sp.idy2 = 1.0 / sp.dy2;
  sp.idz2 = 1.0 / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  sp.pow2 = is_pow2(sp.dx2) && is_pow2(sp.dy2) && is_pow2(sp.dz2);
  return sp;
}

// Hook-layer norm workspace: two slots (the two colour passes of a relax
// sweep are folded together).  The hook layer is single-threaded like the
// reference driver.
NormWs g_ws = {nullptr, nullptr, 0, 2};

int ensure_ws(int nb) {
  if (nb <= g_ws.cap) return 0;
  if (g_ws.partials) {
    PMG_HOST_CHECK(hipDeviceSynchronize());
    PMG_HOST_CHECK(hipFree(g_ws.partials));
  }
  int cap = nb < 4096 ? 4096 : nb;
  PMG_HOST_CHECK(hipMalloc(&g_ws.partials, sizeof(double) * 2 * 2 * cap));
  if (!g_ws.slot_nb) {
    PMG_HOST_CHECK(hipMalloc(&g_ws.slot_nb, sizeof(int) * 2));
    PMG_HOST_CHECK(hipMemset(g_ws.slot_nb, 0, sizeof(int) * 2));
  }
  g_ws.cap = cap;
  return 0;
}

// the two colour passes and the norm fold of a relax sweep over nb blocks,
// with the partial norms in ws
int relax_launch(int op, double *V, const double *rhs, const double *rho, const double *a, const double *c,
                 const double *cmask, const double *chr, double ex, const HookGeom &g, const Spacing<double> &sp,
                 int nb, NormWs ws, double *norm_dev, hipStream_t stream) {
  for (int pass = 0; pass <= 1; pass++) {
#define PMG_HR(OPV, P2V)                                                                         \
  hipLaunchKernelGGL((k_hook_relax_pass<OPV, P2V>), dim3(nb), dim3(NT), 0, stream, V, rhs, rho, \
                     a, c, cmask, chr, ex, g, sp, pass, ws, nb)
    if (op == 0) { if (sp.pow2) PMG_HR(0, true); else PMG_HR(0, false); }
    else { if (sp.pow2) PMG_HR(1, true); else PMG_HR(1, false); }
#undef PMG_HR
    PMG_HOST_CHECK(hipGetLastError());
  }
  if (norm_dev) {
    hipLaunchKernelGGL(k_norm_fold, dim3(1), dim3(kFoldThreads), 0, stream, ws, 0, 2, norm_dev);
    PMG_HOST_CHECK(hipGetLastError());
  }
  return 0;
}

}  // namespace

extern "C" {

int pmg_hook_relax(int op, double *V, const double *rhs, const double *rho, const double *a,
                   const double *c, const double *cmask, const double *chr, double ex, int nx,
                   int ny, int nz, double dx, double dy, double dz, double *norm_dev,
                   hipStream_t stream) {
  if (nx < 3 || ny < 3 || nz < 3 || !V || !rhs || !rho || !cmask) return -1;
  if (op == 1 && (!a || !c)) return -1;
  HookGeom g{nx, ny, nz, (size_t)nx, (size_t)nx * ny};
  Spacing<double> sp = make_spacing(dx, dy, dz);
  size_t n = (size_t)(nx - 2) * (ny - 2) * (nz - 2);
  int nb = (int)((n + NT - 1) / NT);
  if (ensure_ws(nb)) return -1;
  return relax_launch(op, V, rhs, rho, a, c, cmask, chr, ex, g, sp, nb, g_ws, norm_dev, stream);
}

int pmg_hook_relax_noalloc(int op, double *V, const double *rhs, const double *rho, const double *a,
                           const double *c, const double *cmask, const double *chr, double ex, int nx,
                           int ny, int nz, double dx, double dy, double dz, double *norm_dev,
                           hipStream_t stream) {
  if (nx < 3 || ny < 3 || nz < 3 || !V || !rhs || !rho || !cmask) return -1;
  if (op == 1 && (!a || !c)) return -1;
  const size_t n = (size_t)(nx - 2) * (ny - 2) * (nz - 2);
  if ((n + NT - 1) / NT > (size_t)kScratchBlocks) {
    std::fprintf(stderr, "pmg_hook_relax_noalloc: %zu interior points exceed the %zu its scratch holds\n", n,
                 (size_t)kScratchBlocks * NT);
    return -2;
  }
  double *base = nullptr;
  if (scratch_address(&base)) return -1;
  NormWs ws{base, reinterpret_cast<int *>(base + kScratchPartials), kScratchBlocks, 2};
  HookGeom g{nx, ny, nz, (size_t)nx, (size_t)nx * ny};
  Spacing<double> sp = make_spacing(dx, dy, dz);
  const int nb = (int)((n + NT - 1) / NT);
  return relax_launch(op, V, rhs, rho, a, c, cmask, chr, ex, g, sp, nb, ws, norm_dev, stream);
}

int pmg_hook_residual(int op, double *res, const double *rhs, const double *V, const double *rho,
                      const double *a, const double *c, const double *cmask, const double *chr,
                      double ex, int nx, int ny, int nz, double dx, double dy, double dz,
                      double *norm_dev, hipStream_t stream) {
  if (nx < 3 || ny < 3 || nz < 3 || !V || !rhs || !rho || !cmask || !res) return -1;
  if (op == 1 && (!a || !c)) return -1;
  HookGeom g{nx, ny, nz, (size_t)nx, (size_t)nx * ny};
  Spacing<double> sp = make_spacing(dx, dy, dz);
  size_t n = (size_t)(nx - 2) * (ny - 2) * (nz - 2);
  int nb = (int)((n + NT - 1) / NT);
  if (ensure_ws(nb)) return -1;
#define PMG_HS(OPV, P2V)                                                                        \
  hipLaunchKernelGGL((k_hook_residual<OPV, P2V>), dim3(nb), dim3(NT), 0, stream, res, rhs, V, \
                     rho, a, c, cmask, chr, ex, g, sp, g_ws, nb)
  if (op == 0) { if (sp.pow2) PMG_HS(0, true); else PMG_HS(0, false); }
  else { if (sp.pow2) PMG_HS(1, true); else PMG_HS(1, false); }
#undef PMG_HS
  PMG_HOST_CHECK(hipGetLastError());
  if (norm_dev) {
    hipLaunchKernelGGL(k_norm_fold, dim3(1), dim3(kFoldThreads), 0, stream, g_ws, 0, 1, norm_dev);
    PMG_HOST_CHECK(hipGetLastError());
  }
  return 0;
}

int pmg_hook_lop(int op, double *LV, const double *V, const double *rho, const double *a,
                 const double *c, const double *cmask, const double *chr, double ex, int nx,
                 int ny, int nz, double dx, double dy, double dz, hipStream_t stream) {
  if (nx < 3 || ny < 3 || nz < 3 || !V || !LV || !rho || !cmask) return -1;
  if (op == 1 && (!a || !c)) return -1;
  HookGeom g{nx, ny, nz, (size_t)nx, (size_t)nx * ny};
  Spacing<double> sp = make_spacing(dx, dy, dz);
  size_t n = (size_t)(nx - 2) * (ny - 2) * (nz - 2);
  int nb = (int)((n + NT - 1) / NT);
#define PMG_HL(OPV, P2V)                                                                    \
  hipLaunchKernelGGL((k_hook_lop<OPV, P2V>), dim3(nb), dim3(NT), 0, stream, LV, V, rho, a, \
                     c, cmask, chr, ex, g, sp)
  if (op == 0) { if (sp.pow2) PMG_HL(0, true); else PMG_HL(0, false); }
  else { if (sp.pow2) PMG_HL(1, true); else PMG_HL(1, false); }
#undef PMG_HL
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
