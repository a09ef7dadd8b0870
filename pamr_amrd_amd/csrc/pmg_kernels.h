// pmg_kernels.h -- templated device kernels of the uniform-hierarchy solver.
//
// The kernels implement, level by level, the operations of one FAS V-cycle
// (amrd/mg.c:84-201) over a single grid per level.  They are templated on
// the arithmetic type (FP64 for the reference method, FP32 for the inner
// cycles of mixed-precision defect correction), on the operator (0 plain
// Poisson, 1 variable coefficient) and on excision.
#pragma once
#include "pmg_ops.h"
#include <type_traits>

namespace pmg {

// ---------------------------------------------------------------------------
// Fused red-black Gauss-Seidel sweep (both colours in one pass).
//
// A block owns a TX x TY tile of interior (x,y) columns and a chunk of z
// planes.  It marches in z with a 4-plane LDS ring holding u over the tile
// plus a 2-point halo: pass 0 ("red", mod(i+j+k+pass,2)==0 in the 1-based
// indices of num_w.f:134) is evaluated one plane ahead over tile+1, pass 1
// on the tile, so each plane is read from HBM once and written once.  The
// sweep reads u_in and writes u_out (ping-pong buffers): blocks never see a
// neighbour's updated values, so the redundant halo updates are identical
// to the owner's and results equal the sequential colour-by-colour sweep
// bit for bit.  Boundary points are not written: both buffers keep the same
// boundary shell (maintained by the kernels that write boundaries).
// ---------------------------------------------------------------------------
template <typename T>
struct SweepP {
  const T *uin;
  T *uout;
  const T *rho, *rhs, *a, *c;
  const uint8_t *chr;
  int nx, ny, nz;
  int bx0, bx1, by0, by1, bz0, bz1;  // updatable box, inclusive, within the interior
  int zchunk;
  Spacing<T> sp;
  NormWs ws;
  int slot;  // norm workspace slot
  int nb;
};

// Position of colour-slot q of a sweep phase on plane k.  HALO (pass 0):
// tile + 1 ring; otherwise (pass 1) the tile.  Returns false when the slot
// falls outside the region.  (lx, ly) index the LDS plane, whose origin is
// the global point (x0-2, y0-2).
template <int TX, int TY, bool HALO>
__device__ __forceinline__ bool sweep_pos(int q, int x0, int y0, int k, int pass, int &lx,
                                          int &ly) {
  constexpr int W = HALO ? TX + 2 : TX, H = HALO ? TY + 2 : TY, OFF = HALO ? 1 : 2;
  constexpr int W2 = (W + 1) / 2;
  if (q >= W2 * H) return false;
  ly = OFF + q / W2;
  const int j = y0 - 2 + ly;
  const int want = (j + k + 3 + pass) & 1;  // parity of i: (i+j+k+3+pass) even
  lx = OFF + ((want - (x0 - 2 + OFF)) & 1) + 2 * (q % W2);
  return lx <= OFF + W - 1;
}

template <typename T>
__device__ __forceinline__ bool sweep_updatable(const SweepP<T> &P, int i, int j, int k) {
  return i >= P.bx0 && i <= P.bx1 && j >= P.by0 && j <= P.by1 && k >= P.bz0 && k <= P.bz1;
}

// The variable-coefficient sweep keeps the coefficient a in an LDS ring of
// the same shape as u's when both rings fit (its seven reads per point then
// come from LDS, like u's); otherwise a is read from global memory.
template <typename T, int OP, int TX, int TY, int RING = 4>
constexpr bool sweep_a_ring() {
  return OP == 1 && 2 * RING * (TY + 4) * (TX + 4) * sizeof(T) <= 128 * 1024;
}

// LDS column of local x lx in a row of LX (even) values stored colour-split:
// even lx in the first half of the row, odd lx in the second.  The points of
// one colour in a row all have the same lx parity, so a wave updating them
// reads and writes consecutive 8-byte words (no bank conflicts), and so do
// their x, y and z neighbours.
template <int LX>
__device__ __forceinline__ int csx(int lx) {
  static_assert(LX % 2 == 0, "colour-split rows need an even length");
  return (lx & 1) * (LX / 2) + (lx >> 1);
}

// Slot of plane k (k >= -2) in an LDS ring of RING planes.
template <int RING>
__device__ __forceinline__ int ring_ix(int k) {
  if constexpr (RING == 4) return k & 3; else return (k + 2 * RING) % RING;
}

// Flat LDS offset (within a plane) of local position (lx, ly), and the
// offset from it to its east neighbour in the colour-split row: +LX/2 for an
// even lx, 1 - LX/2 for an odd one (the west neighbour is one word before
// the east one).
template <int LX>
__device__ __forceinline__ int cs_off(int lx, int ly) { return ly * LX + csx<LX>(lx); }
template <int LX>
__device__ __forceinline__ int cs_east(int lx) { return (lx & 1) ? 1 - LX / 2 : LX / 2; }

// One colour update at flat LDS offset `lo` of plane k (`de`: cs_east).
// rho/rhs are the values at that point (prefetched by the caller); sA is a's
// ring (AR).
template <typename T, int OP, bool P2, bool RHS0, int TX, int TY, bool AR = false, int RING = 4>
__device__ __forceinline__ T sweep_point(const SweepP<T> &P, T (*sU)[TY + 4][TX + 4], int lo,
                                         int de, int k, size_t p, T rho, T rhs,
                                         T (*sA)[TY + 4][TX + 4] = nullptr) {
  constexpr int LX = TX + 4;
  T *S0 = &sU[ring_ix<RING>(k)][0][0];
  const T *Sm = &sU[ring_ix<RING>(k - 1)][0][0];
  const T *Sp = &sU[ring_ix<RING>(k + 1)][0][0];
  const int ce = lo + de, cw = ce - 1;
  const T uc = S0[lo];
  T jac, L;
  if constexpr (OP == 0) {
    L = lop_poisson<T, P2>(uc, S0[ce], S0[cw], S0[lo + LX], S0[lo - LX], Sp[lo], Sm[lo], rho, P.sp);
    jac = P.sp.jac;
  } else if constexpr (AR) {
    const T *A0 = &sA[ring_ix<RING>(k)][0][0];
    const T *Am = &sA[ring_ix<RING>(k - 1)][0][0];
    const T *Ap = &sA[ring_ix<RING>(k + 1)][0][0];
    L = lop_varcoef<T, P2>(uc, S0[ce], S0[cw], S0[lo + LX], S0[lo - LX], Sp[lo], Sm[lo], A0[lo], A0[ce],
                           A0[cw], A0[lo + LX], A0[lo - LX], Ap[lo], Am[lo], P.c[p], rho, P.sp, &jac);
  } else {
    const T *__restrict__ a = P.a;
    const size_t sy = (size_t)P.nx, sz = (size_t)P.nx * P.ny;
    L = lop_varcoef<T, P2>(uc, S0[ce], S0[cw], S0[lo + LX], S0[lo - LX], Sp[lo], Sm[lo], a[p], a[p + 1],
                           a[p - 1], a[p + sy], a[p - sy], a[p + sz], a[p - sz], P.c[p], rho, P.sp, &jac);
  }
  T res;
  if constexpr (RHS0) res = L - T(0); else res = L - rhs;
  if constexpr (std::is_same<T, float>::value && OP == 0)
    S0[lo] = uc - res * P.sp.ijac;  // FP32 inner cycles: no IEEE division
  else
    S0[lo] = uc - div_jac<T, OP>(res, jac, P.sp);  // FP64: the reference's V - res/Jac, bit for bit
  return res;
}

// Prefetch state of one phase: per slot the point index (or -1) and its
// rho/rhs values, loaded one step before use.
template <typename T, int N>
struct PhaseRegs {
  long long p[N];
  int lo[N], fl[N];  // as in SlotTab
  T rho[N], rhs[N];
};

// Per-thread slot table of one phase for both plane parities, computed once
// per block: LDS position (flat colour-split offset `lo`; `fl` = 2 * cs_east
// + 1 when the slot lies in the tile itself, whose residuals the norm
// counts), xy offset of the point, and whether the slot is inside the tile
// region and the updatable x/y range.
template <int N>
struct SlotTab {
  int lo[2][N], fl[2][N];
  int off[2][N];  // i + nx*j, or -1
};

template <typename T, int TX, int TY, int NT, bool HALO, int N>
__device__ __forceinline__ void slot_table(const SweepP<T> &P, int x0, int y0, int pass, SlotTab<N> &S) {
#pragma unroll
  for (int par = 0; par < 2; par++)
#pragma unroll
    for (int m = 0; m < N; m++) {
      int lx = 0, ly = 0;
      S.off[par][m] = -1;
      // plane parity par: the colour test uses (j + k + 3 + pass), k = par mod 2
      if (!sweep_pos<TX, TY, HALO>(threadIdx.x + m * NT, x0, y0, par, pass, lx, ly)) continue;
      const int i = x0 - 2 + lx, j = y0 - 2 + ly;
      S.lo[par][m] = cs_off<TX + 4>(lx, ly);
      S.fl[par][m] = 2 * cs_east<TX + 4>(lx) + ((lx >= 2 && lx < TX + 2 && ly >= 2 && ly < TY + 2) ? 1 : 0);
      if (i >= P.bx0 && i <= P.bx1 && j >= P.by0 && j <= P.by1) S.off[par][m] = i + P.nx * j;
    }
}

// kpar: the plane whose parity selects the table (k itself, except for the
// prologue's pass 0 over the tile, which takes the pass-1 table of the plane
// above: the colours of pass 0 on plane k are those of pass 1 on k + 1).
template <typename T, bool EX, bool RHS0, int N>
__device__ __forceinline__ void phase_prefetch_t(const SweepP<T> &P, int k, const SlotTab<N> &S,
                                                 PhaseRegs<T, N> &R, int kpar) {
  // the plane parity selects between the two register tables with a
  // select, not an index (a run-time index would move them to scratch)
  //
  // The loads are unconditional (slots outside the region read element 0 and
  // are marked p = -1): a load behind a branch makes the compiler's wait
  // counting fall back to waiting for every load in flight, which would
  // stall each march step on the plane prefetched for the steps after it.
  const bool odd = (kpar & 1) != 0;
  const bool zok = k >= P.bz0 && k <= P.bz1;
  const size_t base = (size_t)k * P.nx * P.ny;
#pragma unroll
  for (int m = 0; m < N; m++) {
    const int off = odd ? S.off[1][m] : S.off[0][m];
    const bool ok = zok && off >= 0;
    const size_t p = ok ? base + (size_t)off : 0;
    bool act = ok;
    if constexpr (EX) act = act && !P.chr[p];
    R.p[m] = act ? (long long)p : -1;
    R.lo[m] = odd ? S.lo[1][m] : S.lo[0][m];
    R.fl[m] = odd ? S.fl[1][m] : S.fl[0][m];
    R.rho[m] = P.rho[p];
    if constexpr (!RHS0) R.rhs[m] = P.rhs[p];
  }
}

template <typename T, int OP, bool P2, bool RHS0, int TX, int TY, int N, bool AR, int RING>
__device__ __forceinline__ void phase_run(const SweepP<T> &P, T (*sU)[TY + 4][TX + 4],
                                          T (*sA)[TY + 4][TX + 4], int k, const PhaseRegs<T, N> &R, bool count,
                                          double &s, double &cnt) {
#pragma unroll
  for (int m = 0; m < N; m++) {
    if (R.p[m] < 0) continue;
    const T res = sweep_point<T, OP, P2, RHS0, TX, TY, AR, RING>(P, sU, R.lo[m], R.fl[m] >> 1, k, (size_t)R.p[m],
                                                                 R.rho[m], RHS0 ? T(0) : R.rhs[m], sA);
    if (count && (R.fl[m] & 1)) {
      s += (double)res * (double)res;
      cnt += 1.0;
    }
  }
}

// Fused sweep.  Step k finalises plane k: pass 0 on plane k+1 (tile+1),
// pass 1 on plane k (tile), store plane k.  While a step computes, the u
// values of plane k+3 and the rho/rhs values of the next step are already
// in flight in registers, so HBM latency overlaps the stencil work.
// One tile of the fused sweep: columns [x0, x0+TX) x [y0, y0+TY) of the
// interior, planes [z0, z1).  `bid` is the tile's slot in the norm
// workspace.  Shared by the single-grid and the batched (all grids of a
// level) kernels.
// PF: u planes in flight ahead of the march (1: one plane issued per step and
// committed the next; 2: two, for latency-bound tiles such as FP32).  Planes
// past z1+1, which no pass of the chunk reads, are not fetched.
// RING: LDS planes.  With 4, the slot plane k+3 is committed into is that of
// plane k-1, which pass 1 on plane k reads, so a step needs three barriers
// (after pass 0, after pass 1, after the commit).  With 5 the commit touches
// no slot the step reads: it follows pass 1 directly and one barrier covers
// both, two barriers per step.
template <typename T, int OP, bool EX, bool P2, bool RHS0, int TX, int TY, int NT, int PF = 1, int RING = 4>
__device__ __forceinline__ void sweep_tile(const SweepP<T> &P, int x0, int y0, int z0, int z1, int bid) {
  static_assert(PF == 1 || PF == 2, "prefetch depth");
  static_assert(RING == 4 || RING == 5, "LDS ring");
  constexpr int LX = TX + 4, LY = TY + 4;
  constexpr int NLD = (LX * LY + NT - 1) / NT;
  constexpr int N0 = (((TX + 3) / 2) * (TY + 2) + NT - 1) / NT;
  constexpr int N1 = (((TX + 1) / 2) * TY + NT - 1) / NT;
  constexpr bool AR = sweep_a_ring<T, OP, TX, TY, RING>();
  __shared__ T sU[RING][LY][LX];
  __shared__ T sAbuf[AR ? RING * LY * LX : 1];
  T(*sA)[LY][LX] = AR ? reinterpret_cast<T(*)[LY][LX]>(sAbuf) : nullptr;
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const T *__restrict__ uin = P.uin;
  const T *__restrict__ ain = P.a;
  T *__restrict__ uout = P.uout;
  double s = 0.0, cnt = 0.0;

  // u (and a) planes are prefetched into rotating register sets
  struct Ld {
    T u[NLD];
    T a[AR ? NLD : 1];
  };
  Ld ldA, ldB, ldC;
  // the last plane the chunk reads: pass 0 also runs on plane z1 (pass 1 on
  // z1-1 needs its updated points) and reads z1+1
  const int kmax = min(z1 + 1, nz - 1);
  // unconditional loads (positions outside the grid or the chunk read
  // element 0, never committed): see phase_prefetch_t
  auto load_issue = [&](Ld &ld, int k) {
#pragma unroll
    for (int m = 0; m < NLD; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % LX, ly = q / LX;
      const int i = x0 - 2 + lx, j = y0 - 2 + ly;
      const bool ok = k >= 0 && k <= kmax && q < LX * LY && i >= 0 && i < nx && j >= 0 && j < ny;
      const size_t p = ok ? ix3(i, j, k, nx, ny) : 0;
      ld.u[m] = uin[p];
      if constexpr (AR) ld.a[m] = ain[p];
    }
  };
  auto load_commit = [&](const Ld &ld, int k) {
    if (k < 0 || k > kmax) return;
    T(*S)[LX] = sU[ring_ix<RING>(k)];
#pragma unroll
    for (int m = 0; m < NLD; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % LX, ly = q / LX;
      const int i = x0 - 2 + lx, j = y0 - 2 + ly;
      if (q < LX * LY && i >= 0 && i < nx && j >= 0 && j < ny) {
        S[ly][csx<LX>(lx)] = ld.u[m];
        if constexpr (AR) sA[ring_ix<RING>(k)][ly][csx<LX>(lx)] = ld.a[m];
      }
    }
  };
  // fixed trip count (unrolled): no loop of global stores inside the march
  constexpr int NST = (TX * TY + NT - 1) / NT;
  auto store = [&](int k) {
    T(*S)[LX] = sU[ring_ix<RING>(k)];
#pragma unroll
    for (int m = 0; m < NST; m++) {
      const int q = threadIdx.x + m * NT;
      const int tx = q % TX, ty = q / TX;
      const int i = x0 + tx, j = y0 + ty;
      if (q < TX * TY && i <= nx - 2 && j <= ny - 2) uout[ix3(i, j, k, nx, ny)] = S[ty + 2][csx<LX>(tx + 2)];
    }
  };

  // prologue: planes z0-2 .. z0+2, pass 0 on z0-1 (tile) and z0 (tile+1).
  // Every load of the prologue and of the first march step is requested
  // before any is used (one memory round trip instead of one per plane).
  SlotTab<N0> tab0;
  SlotTab<N1> tab1;
  slot_table<T, TX, TY, NT, true, N0>(P, x0, y0, 0, tab0);
  slot_table<T, TX, TY, NT, false, N1>(P, x0, y0, 1, tab1);
  // the rho/rhs prefetch sets alternate with the steps like the u sets (no
  // register copies of values still in flight)
  PhaseRegs<T, N0> pa0, pb0, pc0;
  PhaseRegs<T, N1> pa1, pb1, pc1;
  {
    Ld q0, q1, q2, q3, q4;
    load_issue(q0, z0 - 2);
    load_issue(q1, z0 - 1);
    load_issue(q2, z0);
    load_issue(q3, z0 + 1);
    load_issue(q4, z0 + 2);
    PhaseRegs<T, N1> pro1;  // pass 0 on plane z0-1 over the tile
    PhaseRegs<T, N0> pro0;  // pass 0 on plane z0 over tile+1
    phase_prefetch_t<T, EX, RHS0, N1>(P, z0 - 1, tab1, pro1, z0);
    phase_prefetch_t<T, EX, RHS0, N0>(P, z0, tab0, pro0, z0);
    load_issue(ldA, z0 + 3);
    if constexpr (PF == 2) load_issue(ldB, z0 + 4);
    phase_prefetch_t<T, EX, RHS0, N0>(P, z0 + 1, tab0, pa0, z0 + 1);
    phase_prefetch_t<T, EX, RHS0, N1>(P, z0, tab1, pa1, z0);
    load_commit(q0, z0 - 2);
    load_commit(q1, z0 - 1);
    load_commit(q2, z0);
    load_commit(q3, z0 + 1);
    __syncthreads();
    phase_run<T, OP, P2, RHS0, TX, TY, N1, AR, RING>(P, sU, sA, z0 - 1, pro1, false, s, cnt);
    phase_run<T, OP, P2, RHS0, TX, TY, N0, AR, RING>(P, sU, sA, z0, pro0, true, s, cnt);
    __syncthreads();
    load_commit(q4, z0 + 2);
    __syncthreads();
  }
  // step k: issue plane k+3+PF and the next step's rho/rhs; pass 0 on k+1;
  // pass 1 on k; store k; commit plane k+3 (issued PF steps earlier)
  auto step = [&](int k, const Ld &ldc, Ld &ldi, const PhaseRegs<T, N0> &cur0, const PhaseRegs<T, N1> &cur1,
                  PhaseRegs<T, N0> &nxt0, PhaseRegs<T, N1> &nxt1) {
    load_issue(ldi, k + 3 + PF);
    phase_prefetch_t<T, EX, RHS0, N0>(P, k + 2, tab0, nxt0, k + 2);
    phase_prefetch_t<T, EX, RHS0, N1>(P, k + 1, tab1, nxt1, k + 1);
    phase_run<T, OP, P2, RHS0, TX, TY, N0, AR, RING>(P, sU, sA, k + 1, cur0, k + 1 < z1, s, cnt);
    __syncthreads();
    phase_run<T, OP, P2, RHS0, TX, TY, N1, AR, RING>(P, sU, sA, k, cur1, true, s, cnt);
    if constexpr (RING == 5) {
      load_commit(ldc, k + 3);
      __syncthreads();
      store(k);
    } else {
      __syncthreads();
      store(k);
      load_commit(ldc, k + 3);
      __syncthreads();
    }
  };
  if constexpr (PF == 1) {
    for (int k = z0; k < z1; k += 2) {
      step(k, ldA, ldB, pa0, pa1, pb0, pb1);
      if (k + 1 < z1) step(k + 1, ldB, ldA, pb0, pb1, pa0, pa1);
    }
  } else {
    // set (k - z0) % 3 holds plane k+3; the set committed one step earlier
    // takes plane k+5
    for (int k = z0; k < z1; k += 3) {
      step(k, ldA, ldC, pa0, pa1, pb0, pb1);
      if (k + 1 < z1) step(k + 1, ldB, ldA, pb0, pb1, pc0, pc1);
      if (k + 2 < z1) step(k + 2, ldC, ldB, pc0, pc1, pa0, pa1);
    }
  }
  norm_partial<NT>(s, cnt, P.ws, P.slot, bid, P.nb);
}

template <typename T, int OP, bool EX, bool P2, bool RHS0, int TX, int TY, int NT, int PF = 1, int RING = 4>
__global__ __launch_bounds__(NT) void k_sweep(SweepP<T> P) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int nty = (P.ny - 2 + TY - 1) / TY, ntx = (P.nx - 2 + TX - 1) / TX;
  const int ty = tile % nty, tx = (tile / nty) % ntx, tz = tile / (nty * ntx);
  const int z0 = 1 + tz * P.zchunk;
  sweep_tile<T, OP, EX, P2, RHS0, TX, TY, NT, PF, RING>(P, 1 + tx * TX, 1 + ty * TY, z0,
                                                        min(z0 + P.zchunk, P.nz - 1), blockIdx.x);
}

// ---------------------------------------------------------------------------
// Two red-black sweeps in one pass over HBM (temporal blocking).  The four
// colour passes -- sweep A pass 0 / pass 1, sweep B pass 0 / pass 1 -- run
// as a pipeline in one z-march: at step k, phase p works on plane
// k + 3 - p over the tile plus a (3 - p)-point ring, so every value a phase
// reads was produced by the previous phase in this block.  The LDS holds a
// 6-plane ring of the tile plus a 4-point margin.  Each u value crosses HBM
// once per two sweeps (plus the margins); results equal two sequential
// sweeps bit for bit, and the norm is that of sweep B (the last one).
// ---------------------------------------------------------------------------
constexpr int kS2Margin = 4, kS2Planes = 6;

// colour slot q of phase region (tile + R ring) on plane k; LDS margin M
template <int TX, int TY, int R, int M>
__device__ __forceinline__ bool sweep2_pos(int q, int x0, int y0, int k, int pass, int &lx, int &ly) {
  constexpr int W = TX + 2 * R, H = TY + 2 * R, OFF = M - R;
  constexpr int W2 = (W + 1) / 2;
  if (q >= W2 * H) return false;
  ly = OFF + q / W2;
  const int j = y0 - M + ly;
  const int want = (j + k + 3 + pass) & 1;  // parity of i: (i+j+k+3+pass) even
  lx = OFF + ((want - (x0 - M + OFF)) & 1) + 2 * (q % W2);
  return lx <= OFF + W - 1;
}

template <typename T, int OP, bool EX, bool P2, bool RHS0, int TX, int TY, int NT>
__device__ __forceinline__ void sweep2_tile(const SweepP<T> &P, int x0, int y0, int z0, int z1, int bid) {
  constexpr int M = kS2Margin, NPL = kS2Planes;
  constexpr int LX = TX + 2 * M, LY = TY + 2 * M;
  constexpr int NLD = (LX * LY + NT - 1) / NT;
  constexpr int NS = ((((TX + 7) / 2) * (TY + 6)) + NT - 1) / NT;  // slots of the widest phase
  __shared__ T sU[NPL][LY][LX];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const T *__restrict__ uin = P.uin;
  T *__restrict__ uout = P.uout;
  auto ring = [](int k) { return (k + 2 * NPL) % NPL; };
  double s = 0.0, cnt = 0.0;

  // per-thread slot tables of the four phases for both plane parities
  int tlx[4][2][NS], tly[4][2][NS], toff[4][2][NS];
  auto table = [&](auto PHc) {
    constexpr int ph = decltype(PHc)::value, R = 3 - ph;
#pragma unroll
    for (int par = 0; par < 2; par++)
#pragma unroll
      for (int m = 0; m < NS; m++) {
        int lx = 0, ly = 0;
        toff[ph][par][m] = -1;
        tlx[ph][par][m] = 0;
        tly[ph][par][m] = 0;
        if (!sweep2_pos<TX, TY, R, M>(threadIdx.x + m * NT, x0, y0, par, ph & 1, lx, ly)) continue;
        const int i = x0 - M + lx, j = y0 - M + ly;
        tlx[ph][par][m] = lx;
        tly[ph][par][m] = ly;
        if (i >= P.bx0 && i <= P.bx1 && j >= P.by0 && j <= P.by1) toff[ph][par][m] = i + nx * j;
      }
  };
  table(std::integral_constant<int, 0>());
  table(std::integral_constant<int, 1>());
  table(std::integral_constant<int, 2>());
  table(std::integral_constant<int, 3>());

  auto load_direct = [&](int k) {
    if (k < 0 || k >= nz) return;
    T(*S)[LX] = sU[ring(k)];
    for (int q = threadIdx.x; q < LX * LY; q += NT) {
      const int lx = q % LX, ly = q / LX;
      const int i = x0 - M + lx, j = y0 - M + ly;
      if (i >= 0 && i < nx && j >= 0 && j < ny) S[ly][lx] = uin[ix3(i, j, k, nx, ny)];
    }
  };
  T ldA[NLD], ldB[NLD];
  auto load_issue = [&](T(&ld)[NLD], int k) {
#pragma unroll
    for (int m = 0; m < NLD; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % LX, ly = q / LX;
      const int i = x0 - M + lx, j = y0 - M + ly;
      if (k >= 0 && k < nz && q < LX * LY && i >= 0 && i < nx && j >= 0 && j < ny) ld[m] = uin[ix3(i, j, k, nx, ny)];
    }
  };
  auto load_commit = [&](const T(&ld)[NLD], int k) {
    if (k < 0 || k >= nz) return;
    T(*S)[LX] = sU[ring(k)];
#pragma unroll
    for (int m = 0; m < NLD; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % LX, ly = q / LX;
      const int i = x0 - M + lx, j = y0 - M + ly;
      if (q < LX * LY && i >= 0 && i < nx && j >= 0 && j < ny) S[ly][lx] = ld[m];
    }
  };
  // one colour phase on plane k (in place in the LDS ring); the phase index
  // is a compile-time constant so the slot tables stay in registers
  auto phase = [&](auto PHc, int k, bool count) {
    constexpr int ph = decltype(PHc)::value;
    if (k < P.bz0 || k > P.bz1) return;
    const bool odd = (k & 1) != 0;
    const size_t base = (size_t)k * nx * ny;
    T(*S0)[LX] = sU[ring(k)];
    T(*Sm)[LX] = sU[ring(k - 1)];
    T(*Sp)[LX] = sU[ring(k + 1)];
#pragma unroll
    for (int m = 0; m < NS; m++) {
      const int off = odd ? toff[ph][1][m] : toff[ph][0][m];
      if (off < 0) continue;
      const int lx = odd ? tlx[ph][1][m] : tlx[ph][0][m];
      const int ly = odd ? tly[ph][1][m] : tly[ph][0][m];
      const size_t p = base + (size_t)off;
      if constexpr (EX) { if (P.chr[p]) continue; }
      const T uc = S0[ly][lx];
      T jac, L;
      if constexpr (OP == 0) {
        L = lop_poisson<T, P2>(uc, S0[ly][lx + 1], S0[ly][lx - 1], S0[ly + 1][lx], S0[ly - 1][lx], Sp[ly][lx],
                               Sm[ly][lx], P.rho[p], P.sp);
        jac = P.sp.jac;
      } else {
        const T *__restrict__ a = P.a;
        const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
        L = lop_varcoef<T, P2>(uc, S0[ly][lx + 1], S0[ly][lx - 1], S0[ly + 1][lx], S0[ly - 1][lx], Sp[ly][lx],
                               Sm[ly][lx], a[p], a[p + 1], a[p - 1], a[p + sy], a[p - sy], a[p + sz], a[p - sz],
                               P.c[p], P.rho[p], P.sp, &jac);
      }
      T res;
      if constexpr (RHS0) res = L - T(0); else res = L - P.rhs[p];
      if constexpr (std::is_same<T, float>::value && OP == 0) S0[ly][lx] = uc - res * P.sp.ijac;
      else S0[ly][lx] = uc - div_jac<T, OP>(res, jac, P.sp);
      if (count && lx >= M && lx < M + TX && ly >= M && ly < M + TY) {
        s += (double)res * (double)res;
        cnt += 1.0;
      }
    }
  };
  auto store = [&](int k) {
    T(*S)[LX] = sU[ring(k)];
    for (int q = threadIdx.x; q < TX * TY; q += NT) {
      const int tx = q % TX, ty = q / TX;
      const int i = x0 + tx, j = y0 + ty;
      if (i <= nx - 2 && j <= ny - 2) uout[ix3(i, j, k, nx, ny)] = S[ty + M][tx + M];
    }
  };

  // steps k = z0-6 .. z1-1: phase p on plane k+3-p when that plane is in
  // [z0-(3-p), z1-1+(3-p)]; planes z0-4 .. z1+3 pass through the ring
  const int kb = z0 - 6;
  load_direct(kb + 2);
  load_direct(kb + 3);
  load_direct(kb + 4);
  load_issue(ldB, kb + 5);
  __syncthreads();
  auto run = [&](auto PHc, int k) {
    constexpr int ph = decltype(PHc)::value, d = 3 - ph;
    const int kp = k + d;
    if (kp >= z0 - d && kp <= z1 - 1 + d) phase(PHc, kp, ph >= 2 && kp >= z0 && kp < z1);
    __syncthreads();
  };
  auto step = [&](int k, T(&ldi)[NLD], const T(&ldc)[NLD]) {
    load_issue(ldi, k + 6);
    run(std::integral_constant<int, 0>(), k);
    run(std::integral_constant<int, 1>(), k);
    run(std::integral_constant<int, 2>(), k);
    run(std::integral_constant<int, 3>(), k);
    if (k >= z0) store(k);
    load_commit(ldc, k + 5);
    __syncthreads();
  };
  for (int k = kb; k < z1; k += 2) {
    step(k, ldA, ldB);
    if (k + 1 < z1) step(k + 1, ldB, ldA);
  }
  norm_partial<NT>(s, cnt, P.ws, P.slot, bid, P.nb);
}

template <typename T, int OP, bool EX, bool P2, bool RHS0, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_sweep2(SweepP<T> P) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int nty = (P.ny - 2 + TY - 1) / TY, ntx = (P.nx - 2 + TX - 1) / TX;
  const int ty = tile % nty, tx = (tile / nty) % ntx, tz = tile / (nty * ntx);
  const int z0 = 1 + tz * P.zchunk;
  sweep2_tile<T, OP, EX, P2, RHS0, TX, TY, NT>(P, 1 + tx * TX, 1 + ty * TY, z0, min(z0 + P.zchunk, P.nz - 1),
                                               blockIdx.x);
}

// ---------------------------------------------------------------------------
// Residual res = L[u] - rhs on updatable points (num_w.f:59-90).  Other
// interior points: ZERO mode writes 0 (clear_MG_vars then residual,
// amrd/mg.c:91,104 on the finest level); KEEP mode leaves them (coarser
// levels, whose other points already hold the injected values).  Boundary
// points are never written: on the levels where ZERO mode is used their
// res (and the FP32 defect) are zero from allocation and no kernel writes
// them, which is what clear_MG_vars leaves there.
//
// A block owns an RTX x RTY tile of interior (x, y) columns and marches in
// z: the u plane k (tile + 1-point ring) sits in LDS for the x/y
// neighbours, u(k-1), u(k), u(k+1) of the own column in registers, so each
// u value is read from HBM once (plus the ring).
// ---------------------------------------------------------------------------
template <typename T>
struct ResidP {
  const T *u, *rho, *rhs, *a, *c;
  const uint8_t *chr;
  T *res;
  float *neg32;  // OUT32 mode: writes (float)(-res) here instead of res
  int nx, ny, nz;
  int bx0, bx1, by0, by1, bz0, bz1;
  int zchunk;
  Spacing<T> sp;
  NormWs ws;
  int slot;
  int nb;
};

constexpr int RTX = 64, RTY = 8, RZC = 16;
constexpr int kResRing = 2 * RTX + 2 * RTY + 4;

template <typename T, int OP, bool EX, bool P2, bool RHS0, bool ZERO, bool OUT32>
__global__ __launch_bounds__(RTX * RTY) void k_residual(ResidP<T> P) {
  __shared__ T sP[RTY + 2][RTX + 2];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const size_t sy = nx, sz = (size_t)nx * ny;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int nty = (ny - 2 + RTY - 1) / RTY, ntx = (nx - 2 + RTX - 1) / RTX;
  const int ty = tile % nty, tx = (tile / nty) % ntx, tz = tile / (nty * ntx);
  const int x0 = 1 + tx * RTX, y0 = 1 + ty * RTY;
  const int z0 = 1 + tz * P.zchunk, z1 = min(z0 + P.zchunk, nz - 1);
  const int lx = threadIdx.x % RTX, ly = threadIdx.x / RTX;
  const int i = x0 + lx, j = y0 + ly;
  const bool inside = i <= nx - 1 && j <= ny - 1;  // a grid point (interior or boundary)
  const bool interior = i <= nx - 2 && j <= ny - 2;
  const bool inxy = interior && i >= P.bx0 && i <= P.bx1 && j >= P.by0 && j <= P.by1;
  // ring point of this thread (threads 0 .. kResRing-1)
  int hx = -1, hy = -1;
  {
    const int t = threadIdx.x;
    if (t < RTX + 2) { hx = t; hy = 0; }
    else if (t < 2 * (RTX + 2)) { hx = t - (RTX + 2); hy = RTY + 1; }
    else if (t < 2 * (RTX + 2) + RTY) { hx = 0; hy = 1 + t - 2 * (RTX + 2); }
    else if (t < kResRing) { hx = RTX + 1; hy = 1 + t - 2 * (RTX + 2) - RTY; }
  }
  const int hi = x0 - 1 + hx, hj = y0 - 1 + hy;
  const bool hvalid = hx >= 0 && hi <= nx - 1 && hj <= ny - 1;
  const T *__restrict__ u = P.u;
  double s = 0.0, cnt = 0.0;
  T um = T(0), uc = T(0), up = T(0);
  if (inside) {
    um = u[ix3(i, j, z0 - 1, nx, ny)];
    uc = u[ix3(i, j, z0, nx, ny)];
  }
  T hv = hvalid ? u[ix3(hi, hj, z0, nx, ny)] : T(0);
  for (int k = z0; k < z1; k++) {
    const size_t p = ix3(i, j, k, nx, ny);
    if (inside) up = u[p + sz];
    T hn = T(0);
    if (hvalid && k + 1 < nz) hn = u[ix3(hi, hj, k + 1, nx, ny)];
    T rhov = T(0), rhsv = T(0);
    bool on = inxy && k >= P.bz0 && k <= P.bz1;
    if constexpr (EX) { if (interior) on = on && !P.chr[p]; }
    if (on) {
      rhov = P.rho ? P.rho[p] : T(0);  // no source array: a zero source (FGMRES's operator)
      if constexpr (!RHS0) rhsv = P.rhs[p];
    }
    __syncthreads();  // the previous plane's readers are done
    sP[ly + 1][lx + 1] = uc;
    if (hvalid) sP[hy][hx] = hv;
    __syncthreads();
    if (on) {
      T L;
      if constexpr (OP == 0) {
        L = lop_poisson<T, P2>(uc, sP[ly + 1][lx + 2], sP[ly + 1][lx], sP[ly + 2][lx + 1], sP[ly][lx + 1], up, um,
                               rhov, P.sp);
      } else {
        const T *__restrict__ a = P.a;
        L = lop_varcoef<T, P2>(uc, sP[ly + 1][lx + 2], sP[ly + 1][lx], sP[ly + 2][lx + 1], sP[ly][lx + 1], up, um,
                               a[p], a[p + 1], a[p - 1], a[p + sy], a[p - sy], a[p + sz], a[p - sz], P.c[p], rhov,
                               P.sp, nullptr);
      }
      T r;
      if constexpr (RHS0) r = L - T(0); else r = L - rhsv;
      if constexpr (OUT32) P.neg32[p] = (float)(-r); else P.res[p] = r;
      s += (double)r * (double)r;
      cnt += 1.0;
    } else if constexpr (ZERO) {
      if (interior) {
        if constexpr (OUT32) P.neg32[p] = 0.0f; else P.res[p] = T(0);
      }
    }
    um = uc;
    uc = up;
    hv = hn;
  }
  norm_partial<RTX * RTY>(s, cnt, P.ws, P.slot, blockIdx.x, P.nb);
}

// ---------------------------------------------------------------------------
// Whole-grid restriction fine (nf) -> coarse (nc = (nf-1)/2+1), the
// local_transfer coarsen call of a single-grid level (src/transfer.c:1126-
// 1149 with i1=i2=1, stride 2).  kind: 1 straight, 2 half weight, 3 full
// weight.  Boundary points are also written to `alt` (the second ping-pong
// buffer of u) when given.
// ---------------------------------------------------------------------------
// Elementwise kernels over a grid use 64 x 4 thread blocks over (x, y) and
// one z plane per blockIdx.z: no 64-bit index division per point.
constexpr int EX_ = 64, EY_ = 4;
inline dim3 grid3(int nx, int ny, int nz) {
  return dim3((nx + EX_ - 1) / EX_, (ny + EY_ - 1) / EY_, nz);
}
#define PMG_IJK(nx, ny)                                          \
  const int i = blockIdx.x * EX_ + (int)(threadIdx.x % EX_);     \
  const int j = blockIdx.y * EY_ + (int)(threadIdx.x / EX_);     \
  const int k = blockIdx.z;                                      \
  if (i >= (nx) || j >= (ny)) return;

template <typename T, int KIND, bool EXC>
__global__ __launch_bounds__(256) void k_restrict_whole(const T *__restrict__ u1,
                                                        T *__restrict__ u2, T *__restrict__ alt,
                                                        const T *__restrict__ chr1, T ex,
                                                        int nx1, int ny1, int nz1, int nx2,
                                                        int ny2, int nz2) {
  PMG_IJK(nx2, ny2)
  const size_t q = ix3(i, j, k, nx2, ny2);
  const int i0 = 1 + 2 * i, j0 = 1 + 2 * j, k0 = 1 + 2 * k;
  T v;
  if constexpr (KIND == 1) {
    v = u1[ix3(i0 - 1, j0 - 1, k0 - 1, nx1, ny1)];
  } else if constexpr (KIND == 2) {
    v = hw_point<T, EXC>(u1, chr1, ex, nx1, ny1, nz1, i0, j0, k0);
  } else {
    if constexpr (EXC) v = fw_point_ex<T>(u1, chr1, ex, nx1, ny1, nz1, i0, j0, k0, i + 1);
    else v = fw_point<T>(u1, nx1, ny1, nz1, i0, j0, k0);
  }
  u2[q] = v;
  if (alt && (i == 0 || j == 0 || k == 0 || i == nx2 - 1 || j == ny2 - 1 || k == nz2 - 1))
    alt[q] = v;
}

// ---------------------------------------------------------------------------
// L_op + compute_MG_rhs fused (amrd/mg.c:124-126, 271-304, save_tre = 0):
//   lop = L[u] on interior points of the child-mask box (not excised), 0
//   elsewhere (cleared by clear_MG_vars); rhs = -res*w0_r + lop inside the
//   child-mask box, 0 outside; then lop = u.
// ---------------------------------------------------------------------------
template <typename T>
struct LopRhsP {
  const T *u, *rho, *res, *a, *c;
  const uint8_t *chr;
  T *rhs, *lop;
  int nx, ny, nz;
  int cx0, cx1, cy0, cy1, cz0, cz1;  // child-mask ON box (may include boundary)
  T w0_r;
  Spacing<T> sp;
};

template <typename T, int OP, bool EX, bool P2>
__global__ __launch_bounds__(256) void k_lop_rhs(LopRhsP<T> P) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  PMG_IJK(nx, ny)
  const size_t q = ix3(i, j, k, nx, ny);
  const bool on = i >= P.cx0 && i <= P.cx1 && j >= P.cy0 && j <= P.cy1 && k >= P.cz0 && k <= P.cz1;
  const bool interior = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
  T lopv = T(0);
  bool act = on && interior;
  if constexpr (EX) act = act && !P.chr[q];
  if (act) lopv = lop_at<T, OP, P2>(P.u, P.rho, P.a, P.c, q, (size_t)nx, (size_t)nx * ny, P.sp, nullptr);
  P.rhs[q] = on ? (-P.res[q]) * P.w0_r + lopv : T(0);
  P.lop[q] = P.u[q];
}

// ---------------------------------------------------------------------------
// Coarse-grid correction, no excision (amrd/mg.c:323-355): the coarse
// correction R = (u_c - lop_c)/w0_i is first stored in the coarse f_res
// (k_cgc_diff, as the reference does), then u_f += P(R) where P is the
// dminterp3d trilinear composition (x, then y, then z phases; c0=c1=0.5 for
// rho=2) evaluated per fine point -- the interpolated fine f_res is consumed
// directly instead of being stored.  Boundary points also go to `alt`.
// ---------------------------------------------------------------------------
//
// One thread per coarse point (ci, cj, ck): it reads the (up to) 8 corners
// of its coarse cell once and updates the fine points 2c + r, r in {0,1}^3
// inside the grid, with the per-point expression of the x/y/z phases.
template <typename T>
__global__ __launch_bounds__(256) void k_cgc(T *__restrict__ u, T *__restrict__ alt,
                                             const T *__restrict__ Rc, int nx, int ny, int nz, int cnx,
                                             int cny, int cnz, const uint8_t *__restrict__ chr) {
  const int ci = blockIdx.x * EX_ + (int)(threadIdx.x % EX_);
  const int cj = blockIdx.y * EY_ + (int)(threadIdx.x / EX_);
  const int ck = blockIdx.z;
  if (ci >= cnx || cj >= cny) return;
  const T c0 = T(1 - (1 * 1.0) / 2), c1 = T((1 * 1.0) / 2);
  const bool xi = 2 * ci + 1 <= nx - 1, yi = 2 * cj + 1 <= ny - 1, zi = 2 * ck + 1 <= nz - 1;
  T r[2][2][2];  // [c][b][a] corners (a along x)
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const bool need = (a == 0 || xi) && (b == 0 || yi) && (c == 0 || zi);
        r[c][b][a] = need ? Rc[ix3(ci + a, cj + b, ck + c, cnx, cny)] : T(0);
      }
#pragma unroll
  for (int rk = 0; rk < 2; rk++) {
    if (rk == 1 && !zi) continue;
#pragma unroll
    for (int rj = 0; rj < 2; rj++) {
      if (rj == 1 && !yi) continue;
#pragma unroll
      for (int ri = 0; ri < 2; ri++) {
        if (ri == 1 && !xi) continue;
        const int i = 2 * ci + ri, j = 2 * cj + rj, k = 2 * ck + rk;
        auto X = [&](int b, int c) -> T { return ri == 0 ? r[c][b][0] : c0 * r[c][b][0] + c1 * r[c][b][1]; };
        auto Y = [&](int c) -> T { return rj == 0 ? X(0, c) : c0 * X(0, c) + c1 * X(1, c); };
        T v = rk == 0 ? Y(0) : c0 * Y(0) + c1 * Y(1);
        const size_t q = ix3(i, j, k, nx, ny);
        if (chr && chr[q]) v = T(0);
        const T nu = u[q] + v;
        u[q] = nu;
        if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) alt[q] = nu;
      }
    }
  }
}

// coarse res = (u - lop)/w0_i (first loop of compute_cgc)
template <typename T>
__global__ __launch_bounds__(256) void k_cgc_diff(T *res, const T *u, const T *lop, T w0_i,
                                                  size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) res[q] = (u[q] - lop[q]) / w0_i;
}

// clear_ex (src/transfer.c:930-935) on res, then u += res (whole grid)
template <typename T>
__global__ __launch_bounds__(256) void k_cgc_add_ex(T *u, T *alt, T *res, const uint8_t *chr,
                                                    int nx, int ny, int nz) {
  PMG_IJK(nx, ny)
  const size_t q = ix3(i, j, k, nx, ny);
  T r = res[q];
  if (chr[q]) { r = T(0); res[q] = r; }
  const T nu = u[q] + r;
  u[q] = nu;
  if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) alt[q] = nu;
}

template <typename T>
__global__ __launch_bounds__(256) void k_clear_ex(T *f, const uint8_t *chr, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n && chr[q]) f[q] = T(0);
}

// ---------------------------------------------------------------------------
// Coarsest level in one workgroup (amrd/mg.c:135-141,161-190 at L=1): the
// pre-sweeps, then ctol = res_pre*MG_crtol, the post-sweeps, then relax
// until res <= ctol or iter == MG_max_citer (iter continues from pst).
// u, rho, rhs (and a, c, chr) live in LDS when they fit; sweeps update in
// place colour by colour, as the hook does.
// ---------------------------------------------------------------------------
template <typename T>
struct CoarseP {
  T *u;
  const T *rho, *rhs, *a, *c;
  const uint8_t *chr;
  int nx, ny, nz;
  int bx0, bx1, by0, by1, bz0, bz1;
  Spacing<T> sp;
  int npre, npst, max_citer;
  double crtol;
  double *norm_pre, *norm_post;  // device outputs (res[1] after pre / final)
  int *citer;                    // device output: iteration count at exit
  T *snap;                       // a saved state of the grid for cycle_watch (or null)
};

// q / m and q % m for 0 <= q < 2^22: a float reciprocal estimate, off by
// at most one, corrected by one integer step (exact).  A few instructions
// instead of the integer-division sequence, which dominates the index
// arithmetic of the single-workgroup kernels below.
struct FastDiv {
  int m;
  float inv;
};
__device__ __forceinline__ FastDiv make_fdiv(int m) { return FastDiv{m, 1.0f / (float)(m > 0 ? m : 1)}; }
__device__ __forceinline__ int fdivmod(int q, const FastDiv &d, int &rem) {
  int t = (int)((float)q * d.inv);
  int r = q - t * d.m;
  if (r < 0) {
    t--;
    r += d.m;
  } else if (r >= d.m) {
    t++;
    r -= d.m;
  }
  rem = r;
  return t;
}
// (i, j, k) of the linear index q of an mi x mj x * box, i fastest
__device__ __forceinline__ void box_ijk(int q, const FastDiv &di, const FastDiv &dj, int &i, int &j, int &k) {
  const int t = fdivmod(q, di, i);
  k = fdivmod(t, dj, j);
}

// The coarsest loop `while (iter < max_citer && res > ctol) relax` of
// amrd/mg.c:179-186 can reach a state that the relaxation maps back onto
// itself after lam sweeps (a fixed point or cycle at rounding level, res
// still above ctol); the reference then relaxes until max_citer.  A sweep is
// a deterministic function of u (fixed point order, fixed reduction tree), so
// from there states and residuals repeat with period lam, and the state and
// residual at max_citer are those after (max_citer - iter) mod lam further
// sweeps.  The cycle is found with Brent's method (one saved state, replaced
// whenever the distance to it reaches a power of two; R. P. Brent, BIT 20,
// 1980), one bitwise comparison of the grid per sweep, watched from the
// kWatchStart-th sweep of the loop on.  The outcome is the same bits and the
// same iteration count as every sweep run.
constexpr int kSnap = 1;        // saved states (of the coarsest grid)
constexpr int kWatchStart = 16;  // loop sweeps before the watch starts
__device__ __forceinline__ bool same_bits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_int(a) == __float_as_int(b); }

struct CycleWatch {
  int power = 1, lam = 0;
  bool saved = false;
};

// called after sweep `iter` (uniform across the block); returns -1 to go on,
// else the number of further sweeps after which u is the state of sweep
// max_citer
template <typename T, int NT>
__device__ int cycle_watch(const T *u, size_t n, T *snap, CycleWatch &cw, int iter, int max_citer) {
  auto save = [&]() {
    for (size_t q = threadIdx.x; q < n; q += NT) snap[q] = u[q];
    __syncthreads();  // before the next sweep writes u
  };
  if (!cw.saved) {
    save();
    cw.saved = true;
    return -1;
  }
  cw.lam++;
  int diff = 0;
  for (size_t q = threadIdx.x; q < n; q += NT) diff |= !same_bits(u[q], snap[q]);
  diff = __syncthreads_or(diff);
  if (!diff) return (max_citer - iter) % cw.lam;
  if (cw.lam == cw.power) {
    save();
    cw.power <<= 1;
    cw.lam = 0;
  }
  return -1;
}

// The coarsest loop by one wavefront.  When the relax box has at most 64
// points, the block kernels give each point to lane q of wave 0 and every
// other wave contributes zero to the norm reduction, whose first stage is the
// wave's shuffle tree; so one wave running the same sweeps and the same tree
// produces the same bits, without a block barrier per half sweep.  Colour
// passes are ordered by waiting for every outstanding memory operation of
// the wave (the arrays are in LDS or in the wave's own global writes).
constexpr int kWaveCoarseMax = 512;  // grid points staged for the wave path
template <typename T>
struct WaveCoarse {
  T *u;
  const T *rho, *rhs, *a, *c;
  const uint8_t *chr;
  int nx, ny, bx0, by0, bz0, mi, mj, npts;
  Spacing<T> sp;
  int npre, npst, max_citer;
  double crtol;
  double *slot;  // norm partial (sum, count) written after every sweep, or null
  int *slot_nb;
  T *snap;       // a saved state for the cycle watch, or null
  size_t n;      // grid points (for the watch)
};

// a wavefront-scope fence orders the wave's memory operations for the
// compiler (the hardware keeps a wave's accesses in order); the wait makes
// every earlier store of the wave complete before any later load issues
__device__ __forceinline__ void wave_wait_all() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// one sweep by lanes 0..63 of wave 0; returns the RMS of the residuals
// (lane-uniform) and writes the norm partial
template <typename T, int OP, bool EX, bool P2, bool IJAC>
__device__ double wave_sweep(const WaveCoarse<T> &W) {
  const int lane = (int)threadIdx.x;
  double s = 0.0, cnt = 0.0;
  const size_t sy = W.nx, sz = (size_t)W.nx * W.ny;
  const FastDiv dmi = make_fdiv(W.mi), dmj = make_fdiv(W.mj);
  for (int pass = 0; pass <= 1; pass++) {
    if (lane < W.npts) {
      int i, j, k;
      box_ijk(lane, dmi, dmj, i, j, k);
      i += W.bx0; j += W.by0; k += W.bz0;
      const size_t p = ix3(i, j, k, W.nx, W.ny);
      if (((i + j + k + 3 + pass) & 1) == 0 && !(EX && W.chr[p])) {
        T jac;
        const T r = lop_at<T, OP, P2>(W.u, W.rho, W.a, W.c, p, sy, sz, W.sp, &jac) - W.rhs[p];
        if constexpr (IJAC) W.u[p] = W.u[p] - r * W.sp.ijac;
        else W.u[p] = W.u[p] - div_jac<T, OP>(r, jac, W.sp);
        s += (double)r * (double)r;
        cnt += 1.0;
      }
    }
    wave_wait_all();
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    s += __shfl_down(s, off, kWave);
    cnt += __shfl_down(cnt, off, kWave);
  }
  double res = 0.0;
  if (lane == 0) {
    res = sqrt(s / cnt);
    if (W.slot) {
      W.slot[0] = s;
      W.slot[1] = cnt;
      *W.slot_nb = 1;
    }
  }
  return __shfl(res, 0, kWave);
}

// cycle_watch for one wave (the saved state in global memory; each lane
// reads back only what it wrote)
template <typename T>
__device__ int cycle_watch_wave(const WaveCoarse<T> &W, CycleWatch &cw, int iter) {
  const int lane = (int)threadIdx.x;
  auto save = [&]() {
    for (size_t q = lane; q < W.n; q += kWave) W.snap[q] = W.u[q];
    wave_wait_all();
  };
  if (!cw.saved) {
    save();
    cw.saved = true;
    return -1;
  }
  cw.lam++;
  int diff = 0;
  for (size_t q = lane; q < W.n; q += kWave) diff |= !same_bits(W.u[q], W.snap[q]);
  if (!__any(diff)) return (W.max_citer - iter) % cw.lam;
  if (cw.lam == cw.power) {
    save();
    cw.power <<= 1;
    cw.lam = 0;
  }
  return -1;
}

// pre sweeps, post sweeps and the loop of amrd/mg.c:135-141, 161-190 at the
// coarsest level, by wave 0; returns (lane-uniform) the final residual and
// sets iter and the pre-loop residual
template <typename T, int OP, bool EX, bool P2, bool IJAC>
__device__ double wave_coarsest(const WaveCoarse<T> &W, int &iter_out, double &res_pre) {
  double res = 0.0;
  for (int it = 0; it < W.npre; it++) res = wave_sweep<T, OP, EX, P2, IJAC>(W);
  const double ctol = res * W.crtol;
  res_pre = res;
  int iter;
  for (iter = 0; iter < W.npst; iter++) res = wave_sweep<T, OP, EX, P2, IJAC>(W);
  res = ctol + 1;
  CycleWatch cw;
  int wit = 0;
  while (iter < W.max_citer && res > ctol) {
    res = wave_sweep<T, OP, EX, P2, IJAC>(W);
    iter++;
    if (W.snap && ++wit >= kWatchStart && res > ctol && iter < W.max_citer) {
      const int r = cycle_watch_wave<T>(W, cw, iter);
      if (r >= 0) {
        for (int k = 0; k < r; k++) res = wave_sweep<T, OP, EX, P2, IJAC>(W);
        iter = W.max_citer;
        break;
      }
    }
  }
  iter_out = iter;
  return res;
}

template <typename T, int OP, bool EX, bool P2, bool LDS, int NT>
__global__ __launch_bounds__(NT) void k_coarse_solve(CoarseP<T> P) {
  extern __shared__ unsigned char smem[];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  T *u = P.u;
  const T *rho = P.rho, *rhs = P.rhs, *a = P.a, *c = P.c;
  if constexpr (LDS) {
    T *su = reinterpret_cast<T *>(smem);
    T *srho = su + n, *srhs = srho + n;
    T *sa = srhs + n, *sc = sa + (OP == 1 ? n : 0);
    for (size_t q = threadIdx.x; q < n; q += NT) {
      su[q] = P.u[q];
      srho[q] = P.rho[q];
      srhs[q] = P.rhs[q];
      if constexpr (OP == 1) { sa[q] = P.a[q]; sc[q] = P.c[q]; }
    }
    __syncthreads();
    u = su; rho = srho; rhs = srhs;
    if constexpr (OP == 1) { a = sa; c = sc; }
  }
  const int mi = P.bx1 - P.bx0 + 1, mj = P.by1 - P.by0 + 1, mk = P.bz1 - P.bz0 + 1;
  const int npts = (mi > 0 && mj > 0 && mk > 0) ? mi * mj * mk : 0;
  const FastDiv dmi = make_fdiv(mi), dmj = make_fdiv(mj);
  __shared__ double bres;
  auto sweep = [&]() -> double {
    double s = 0.0, cnt = 0.0;
    for (int pass = 0; pass <= 1; pass++) {
      for (int q = threadIdx.x; q < npts; q += NT) {
        int i, j, k;
        box_ijk(q, dmi, dmj, i, j, k);
        i += P.bx0; j += P.by0; k += P.bz0;
        if (((i + j + k + 3 + pass) & 1) != 0) continue;
        const size_t p = ix3(i, j, k, nx, ny);
        if constexpr (EX) { if (P.chr[p]) continue; }
        T jac;
        const T r = lop_at<T, OP, P2>(u, rho, a, c, p, sy, sz, P.sp, &jac) - rhs[p];
        u[p] = u[p] - div_jac<T, OP>(r, jac, P.sp);
        s += (double)r * (double)r;
        cnt += 1.0;
      }
      __syncthreads();
    }
    block_reduce2<NT>(s, cnt);
    if (threadIdx.x == 0) bres = sqrt(s / cnt);
    __syncthreads();
    return bres;
  };
  if constexpr (LDS) {
    if (npts <= kWave) {  // one wave, same bits (see wave_coarsest)
      if (threadIdx.x < kWave) {
        WaveCoarse<T> W{u, rho, rhs, a, c, P.chr, nx, ny, P.bx0, P.by0, P.bz0, mi, mj, npts, P.sp,
                        P.npre, P.npst, P.max_citer, P.crtol, nullptr, nullptr, P.snap, n};
        int iter;
        double rpre;
        const double res = wave_coarsest<T, OP, EX, P2, false>(W, iter, rpre);
        if (threadIdx.x == 0) {
          if (P.norm_pre) P.norm_pre[0] = rpre;
          if (P.norm_post) P.norm_post[0] = res;
          if (P.citer) P.citer[0] = iter;
        }
      }
      __syncthreads();
      for (size_t q = threadIdx.x; q < n; q += NT) P.u[q] = u[q];
      return;
    }
  }
  double res = 0.0;
  for (int it = 0; it < P.npre; it++) res = sweep();
  const double ctol = res * P.crtol;
  if (threadIdx.x == 0 && P.norm_pre) P.norm_pre[0] = res;
  int iter = 0;
  for (iter = 0; iter < P.npst; iter++) res = sweep();
  res = ctol + 1;
  CycleWatch cw;
  int wit = 0;
  while (iter < P.max_citer && res > ctol) {
    res = sweep();
    iter++;
    if (P.snap && ++wit >= kWatchStart && res > ctol && iter < P.max_citer) {
      const int r = cycle_watch<T, NT>(u, n, P.snap, cw, iter, P.max_citer);
      if (r >= 0) {
        for (int k = 0; k < r; k++) res = sweep();
        iter = P.max_citer;
        break;
      }
    }
  }
  if constexpr (LDS) {
    for (size_t q = threadIdx.x; q < n; q += NT) P.u[q] = u[q];
  }
  if (threadIdx.x == 0) {
    if (P.norm_post) P.norm_post[0] = res;
    if (P.citer) P.citer[0] = iter;
  }
}

// ---------------------------------------------------------------------------
// Finest-level residual fused with PAMR_inject of u and f_res (full
// weighting) into the next coarser level (amrd/mg.c:103-113 for L+1 = Lf).
// On the finest level rhs = 0 and res is zero off the updatable points, so
// res is computed where the restriction needs it and never stored: a block
// owns CX x CY coarse columns and marches over its coarse z chunk, keeping
// 4 fine u planes and 3 fine res planes (tile plus the restriction ring) in
// LDS.  Per fine point u and rho are read once (plus the halo ring); the
// coarse u (and its ping-pong twin on the boundary) and res are written.
// The residual norm of the fine level is accumulated over the fine points
// each block owns.  Values are identical to residual + dmfwr3d.
// ---------------------------------------------------------------------------
template <typename T>
struct ResRestrP {
  const T *u, *rho, *a, *c;
  T *cu, *cua, *cres;
  int nx, ny, nz, cnx, cny, cnz;
  int czchunk;
  Spacing<T> sp;
  NormWs ws;
  int slot;
  int nb;
};

template <typename T, int OP, bool P2, int CX, int CY, int NT>
__global__ __launch_bounds__(NT) void k_resid_restrict(ResRestrP<T> P) {
  constexpr int UX = 2 * CX + 3, UY = 2 * CY + 3;  // u tile: fine x in [2cx0-2, 2cx0+2CX]
  constexpr int RX = 2 * CX + 1, RY = 2 * CY + 1;  // res tile: fine x in [2cx0-1, 2cx0+2CX-1]
  constexpr int NLU = (UX * UY + NT - 1) / NT, NLR = (RX * RY + NT - 1) / NT;
  __shared__ T sU[4][UY][UX];
  __shared__ T sR[3][RY][RX];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ntx = (P.cnx + CX - 1) / CX, nty = (P.cny + CY - 1) / CY;
  const int ty = tile % nty, tx = (tile / nty) % ntx, tz = tile / (nty * ntx);
  const int cx0 = tx * CX, cy0 = ty * CY;
  const int ck0 = tz * P.czchunk, ck1 = min(ck0 + P.czchunk, P.cnz);
  const int ux0 = 2 * cx0 - 2, uy0 = 2 * cy0 - 2;  // global fine index of u tile origin
  const int rx0 = 2 * cx0 - 1, ry0 = 2 * cy0 - 1;  // global fine index of res tile origin
  double s = 0.0, cnt = 0.0;
  // u planes: loaded into registers one step ahead, then committed to LDS
  T ldU[NLU];
  auto issue_u = [&](int k) {
#pragma unroll
    for (int m = 0; m < NLU; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % UX, ly = q / UX;
      const int i = ux0 + lx, j = uy0 + ly;
      ldU[m] = (q < UX * UY && k >= 0 && k < nz && i >= 0 && i < nx && j >= 0 && j < ny)
                   ? P.u[ix3(i, j, k, nx, ny)] : T(0);
    }
  };
  auto commit_u = [&](int k) {
    T(*S)[UX] = sU[(k + 4) & 3];
#pragma unroll
    for (int m = 0; m < NLU; m++) {
      const int q = threadIdx.x + m * NT;
      if (q < UX * UY) S[q / UX][q % UX] = ldU[m];
    }
  };
  // rho of the res points of a plane, one step ahead
  T ldR[NLR];
  auto issue_rho = [&](int k) {
    const bool kin = k >= 1 && k <= nz - 2;
#pragma unroll
    for (int m = 0; m < NLR; m++) {
      const int q = threadIdx.x + m * NT;
      const int i = rx0 + q % RX, j = ry0 + q / RX;
      ldR[m] = (q < RX * RY && kin && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2)
                   ? P.rho[ix3(i, j, k, nx, ny)] : T(0);
    }
  };
  auto res_plane = [&](int k, const T (&rho)[NLR]) {
    T(*R)[RX] = sR[(k + 3) % 3];
    T(*U0)[UX] = sU[(k + 4) & 3];
    T(*Um)[UX] = sU[(k + 3) & 3];
    T(*Up)[UX] = sU[(k + 5) & 3];
    const bool kin = k >= 1 && k <= nz - 2;
    const bool kown = k >= 2 * ck0 && k < 2 * ck1;
#pragma unroll
    for (int m = 0; m < NLR; m++) {
      const int q = threadIdx.x + m * NT;
      if (q >= RX * RY) continue;
      const int lx = q % RX, ly = q / RX;
      const int i = rx0 + lx, j = ry0 + ly;
      T r = T(0);
      if (kin && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
        const int ax = lx + 1, ay = ly + 1;  // position in the u tile
        if constexpr (OP == 0) {
          r = lop_poisson<T, P2>(U0[ay][ax], U0[ay][ax + 1], U0[ay][ax - 1], U0[ay + 1][ax], U0[ay - 1][ax],
                                 Up[ay][ax], Um[ay][ax], rho[m], P.sp) - T(0);
        } else {
          const size_t p = ix3(i, j, k, nx, ny);
          const size_t sy = nx, sz = (size_t)nx * ny;
          const T *__restrict__ a = P.a;
          r = lop_varcoef<T, P2>(U0[ay][ax], U0[ay][ax + 1], U0[ay][ax - 1], U0[ay + 1][ax], U0[ay - 1][ax],
                                 Up[ay][ax], Um[ay][ax], a[p], a[p + 1], a[p - 1], a[p + sy], a[p - sy],
                                 a[p + sz], a[p - sz], P.c[p], rho[m], P.sp, nullptr) - T(0);
        }
        if (kown && i >= 2 * cx0 && i < 2 * cx0 + 2 * CX && j >= 2 * cy0 && j < 2 * cy0 + 2 * CY) {
          s += (double)r * (double)r;
          cnt += 1.0;
        }
      }
      R[ly][lx] = r;
    }
  };
  auto coarse_plane = [&](int kc) {
    const int t = threadIdx.x;
    if (t >= CX * CY) return;
    const int ci = cx0 + t % CX, cj = cy0 + t / CX;
    if (ci >= P.cnx || cj >= P.cny) return;
    const int i0 = 2 * ci + 1, j0 = 2 * cj + 1, k0 = 2 * kc + 1;  // 1-based fine
    auto Racc = [&](int a, int b, int c) -> T { return sR[(c - 1 + 3) % 3][b - 1 - ry0][a - 1 - rx0]; };
    auto Uacc = [&](int a, int b, int c) -> T { return sU[(c - 1 + 4) & 3][b - 1 - uy0][a - 1 - ux0]; };
    const T vr = fw_point_acc<T>(Racc, nx, ny, nz, i0, j0, k0);
    const T vu = fw_point_acc<T>(Uacc, nx, ny, nz, i0, j0, k0);
    const size_t q = ix3(ci, cj, kc, P.cnx, P.cny);
    P.cres[q] = vr;
    P.cu[q] = vu;
    if (ci == 0 || cj == 0 || kc == 0 || ci == P.cnx - 1 || cj == P.cny - 1 || kc == P.cnz - 1) P.cua[q] = vu;
  };
  // fine planes needed: res at 2ck0-1 .. min(2ck1-1, nz-1); u one more each side
  const int kf0 = 2 * ck0 - 1, kf1 = min(2 * ck1 - 1, nz - 1);
  issue_u(kf0 - 1);
  commit_u(kf0 - 1);
  issue_u(kf0);
  commit_u(kf0);
  issue_u(kf0 + 1);
  issue_rho(kf0);
  for (int k = kf0; k <= kf1; k++) {
    commit_u(k + 1);  // plane k+1 (its slot held plane k-3, no longer read)
    T rho[NLR];
#pragma unroll
    for (int m = 0; m < NLR; m++) rho[m] = ldR[m];
    if (k + 2 <= kf1 + 1) issue_u(k + 2);
    if (k + 1 <= kf1) issue_rho(k + 1);
    __syncthreads();
    res_plane(k, rho);
    __syncthreads();
    // coarse plane kc is complete once fine plane min(2kc+1, nz-1) is done
    if (k & 1) {
      const int kc = (k - 1) / 2;  // k = 2kc+1
      if (kc >= ck0 && kc < ck1) coarse_plane(kc);
    } else if (k == nz - 1 && k / 2 >= ck0 && k / 2 < ck1) {
      coarse_plane(k / 2);
    }
    __syncthreads();
  }
  norm_partial<NT>(s, cnt, P.ws, P.slot, blockIdx.x, P.nb);
}

// The same fused residual + full-weight injection with the restriction
// streamed through the z march: dmfwr3d's 27-term sum for an interior coarse
// point adds its terms in the order plane km, plane k0, plane kp (within a
// plane row by row), left to right, so it can be carried in a register per
// coarse column and extended as each fine plane is formed -- the same
// additions in the same order, hence the same bits.  Only the current fine
// residual plane is kept (two LDS slots) and u needs a ring of three planes,
// and each fine plane costs two barriers instead of three.  Coarse points on
// the x or y faces (lower-dimensional stencils in another order) take
// dmfwr3d's point formula on the fine u in global memory, their res is zero
// (the fine res vanishes on the boundary); the z-boundary planes use the
// single-plane stencil as the plane passes.
// ---------------------------------------------------------------------------
template <typename T, typename V>
__device__ __forceinline__ T fw_plane_chain(T acc, bool first, const V &v, T c) {
  // terms (di, dj) in dmfwr3d's order: dj = -1, 0, 1 outer, di = -1, 0, 1
  // inner, weights c * (1 2 1; 2 4 2; 1 2 1)
  const T w1 = c, w2 = T(2) * c, w4 = T(4) * c;
  acc = first ? w1 * v(-1, -1) : acc + w1 * v(-1, -1);
  acc = acc + w2 * v(0, -1);
  acc = acc + w1 * v(1, -1);
  acc = acc + w2 * v(-1, 0);
  acc = acc + w4 * v(0, 0);
  acc = acc + w2 * v(1, 0);
  acc = acc + w1 * v(-1, 1);
  acc = acc + w2 * v(0, 1);
  acc = acc + w1 * v(1, 1);
  return acc;
}

// Tiles cover the coarse columns 0 .. cnx-2 (and rows 0 .. cny-2): on the
// usual 2^m + 1 grids that is a whole number of tiles, with no tail tile
// marching for a single column.  The last column and row are x/y faces
// (point formula, no march) and go to the tiles at the high edges.
inline __host__ __device__ int rr_tiles(int cn, int c) { return cn > 1 ? (cn - 1 + c - 1) / c : 1; }

template <typename T, int OP, bool P2, int CX, int CY, int NT, int PF>
__global__ __launch_bounds__(NT) void k_resid_restrict_s(ResRestrP<T> P) {
  static_assert(NT >= CX * CY && NT >= CX + CY + 1, "one coarse column per thread");
  static_assert(PF == 1 || PF == 2, "planes in flight");
  constexpr int UX = 2 * CX + 3, UY = 2 * CY + 3;  // u tile: fine x in [2cx0-2, 2cx0+2CX]
  constexpr int RX = 2 * CX + 1, RY = 2 * CY + 1;  // res tile: fine x in [2cx0-1, 2cx0+2CX-1]
  constexpr int NLU = (UX * UY + NT - 1) / NT, NLR = (RX * RY + NT - 1) / NT;
  __shared__ T sU[3][UY][UX];
  __shared__ T sR[2][RY][RX];
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ntx = rr_tiles(P.cnx, CX), nty = rr_tiles(P.cny, CY);
  const int ty = tile % nty, tx = (tile / nty) % ntx, tz = tile / (nty * ntx);
  const int cx0 = tx * CX, cy0 = ty * CY;
  const int ck0 = tz * P.czchunk, ck1 = min(ck0 + P.czchunk, P.cnz);
  const int ux0 = 2 * cx0 - 2, uy0 = 2 * cy0 - 2;
  const int rx0 = 2 * cx0 - 1, ry0 = 2 * cy0 - 1;
  double s = 0.0, cnt = 0.0;
  // u and rho of the next PF fine planes in flight, in registers (buffer
  // plane % PF)
  T ldU[PF][NLU], ldR[PF][NLR];
  auto issue_u = [&](int k, T (&buf)[NLU]) {
#pragma unroll
    for (int m = 0; m < NLU; m++) {
      const int q = threadIdx.x + m * NT;
      const int lx = q % UX, ly = q / UX;
      const int i = ux0 + lx, j = uy0 + ly;
      buf[m] = (q < UX * UY && k >= 0 && k < nz && i >= 0 && i < nx && j >= 0 && j < ny)
                   ? P.u[ix3(i, j, k, nx, ny)] : T(0);
    }
  };
  auto commit_u = [&](int k, const T (&buf)[NLU]) {
    T(*S)[UX] = sU[(k + 3) % 3];
#pragma unroll
    for (int m = 0; m < NLU; m++) {
      const int q = threadIdx.x + m * NT;
      if (q < UX * UY) S[q / UX][q % UX] = buf[m];
    }
  };
  auto issue_rho = [&](int k, T (&buf)[NLR]) {
    const bool kin = k >= 1 && k <= nz - 2;
#pragma unroll
    for (int m = 0; m < NLR; m++) {
      const int q = threadIdx.x + m * NT;
      const int i = rx0 + q % RX, j = ry0 + q / RX;
      buf[m] = (q < RX * RY && kin && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2)
                   ? P.rho[ix3(i, j, k, nx, ny)] : T(0);
    }
  };
  auto res_plane = [&](int k, const T (&rho)[NLR]) {
    T(*R)[RX] = sR[k & 1];
    T(*U0)[UX] = sU[(k + 3) % 3];
    T(*Um)[UX] = sU[(k + 2) % 3];
    T(*Up)[UX] = sU[(k + 4) % 3];
    const bool kin = k >= 1 && k <= nz - 2;
    const bool kown = k >= 2 * ck0 && k < 2 * ck1;
#pragma unroll
    for (int m = 0; m < NLR; m++) {
      const int q = threadIdx.x + m * NT;
      if (q >= RX * RY) continue;
      const int lx = q % RX, ly = q / RX;
      const int i = rx0 + lx, j = ry0 + ly;
      T r = T(0);
      if (kin && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
        const int ax = lx + 1, ay = ly + 1;
        if constexpr (OP == 0) {
          r = lop_poisson<T, P2>(U0[ay][ax], U0[ay][ax + 1], U0[ay][ax - 1], U0[ay + 1][ax], U0[ay - 1][ax],
                                 Up[ay][ax], Um[ay][ax], rho[m], P.sp) - T(0);
        } else {
          const size_t p = ix3(i, j, k, nx, ny);
          const size_t sy = nx, sz = (size_t)nx * ny;
          const T *__restrict__ a = P.a;
          r = lop_varcoef<T, P2>(U0[ay][ax], U0[ay][ax + 1], U0[ay][ax - 1], U0[ay + 1][ax], U0[ay - 1][ax],
                                 Up[ay][ax], Um[ay][ax], a[p], a[p + 1], a[p - 1], a[p + sy], a[p - sy],
                                 a[p + sz], a[p - sz], P.c[p], rho[m], P.sp, nullptr) - T(0);
        }
        if (kown && i >= 2 * cx0 && i < 2 * cx0 + 2 * CX && j >= 2 * cy0 && j < 2 * cy0 + 2 * CY) {
          s += (double)r * (double)r;
          cnt += 1.0;
        }
      }
      R[ly][lx] = r;
    }
  };
  auto put = [&](int ci, int cj, int kc, T vu, T vr) {
    const size_t q = ix3(ci, cj, kc, P.cnx, P.cny);
    P.cres[q] = vr;
    P.cu[q] = vu;
    if (ci == 0 || cj == 0 || kc == 0 || ci == P.cnx - 1 || cj == P.cny - 1 || kc == P.cnz - 1) P.cua[q] = vu;
  };
  // this thread's coarse column (columns 0 .. cnx-2, rows 0 .. cny-2)
  const int t = threadIdx.x;
  const int ci = cx0 + t % CX, cj = cy0 + t / CX;
  const bool mine = t < CX * CY && ci < P.cnx - 1 && cj < P.cny - 1;
  const bool col_in = ci >= 1 && cj >= 1;
  const int lx = 2 * (ci - cx0) + 1, ly = 2 * (cj - cy0) + 1;  // fine centre in the res tile
  T accU = T(0), accR = T(0);
  auto restrict_plane = [&](int f) {
    if (!mine || !col_in) return;
    T(*Uf)[UX] = sU[(f + 3) % 3];
    T(*Rf)[RX] = sR[f & 1];
    auto vu = [&](int di, int dj) -> T { return Uf[ly + 1 + dj][lx + 1 + di]; };
    auto vr = [&](int di, int dj) -> T { return Rf[ly + dj][lx + di]; };
    if ((f & 1) == 0) {
      const int kc = f / 2;
      if (kc < ck0 || kc >= ck1) return;
      if (kc == 0 || kc == P.cnz - 1) {  // single-plane stencil of the z faces
        put(ci, cj, kc, fw_plane_chain<T>(T(0), true, vu, T(1)) / T(16.0),
            fw_plane_chain<T>(T(0), true, vr, T(1)) / T(16.0));
      } else {
        accU = fw_plane_chain<T>(accU, false, vu, T(2));
        accR = fw_plane_chain<T>(accR, false, vr, T(2));
      }
    } else {
      const int kf = (f - 1) / 2, ks = (f + 1) / 2;  // f = kp of kf, km of ks
      if (kf >= ck0 && kf < ck1 && kf >= 1 && kf <= P.cnz - 2) {
        accU = fw_plane_chain<T>(accU, false, vu, T(1));
        accR = fw_plane_chain<T>(accR, false, vr, T(1));
        put(ci, cj, kf, accU / T(64.0), accR / T(64.0));
      }
      if (ks >= ck0 && ks < ck1 && ks >= 1 && ks <= P.cnz - 2) {
        accU = fw_plane_chain<T>(T(0), true, vu, T(1));
        accR = fw_plane_chain<T>(T(0), true, vr, T(1));
      }
    }
  };
  // fine planes: the km plane of the first coarse plane to the kp plane of
  // the last (the z faces: the plane itself)
  const int kf0 = max(2 * ck0 - 1, 0), kf1 = min(2 * ck1 - 1, nz - 1);
  {
    T tmp[NLU];
    issue_u(kf0 - 1, tmp);
    commit_u(kf0 - 1, tmp);
    issue_u(kf0, tmp);
    commit_u(kf0, tmp);
  }
  for (int d = 1; d <= PF; d++) {
    if (PF == 1 || ((kf0 + d) & 1) == 0) issue_u(kf0 + d, ldU[0]);
    else issue_u(kf0 + d, ldU[PF - 1]);
  }
  for (int d = 0; d < PF; d++) {
    if (PF == 1 || ((kf0 + d) & 1) == 0) issue_rho(kf0 + d, ldR[0]);
    else issue_rho(kf0 + d, ldR[PF - 1]);
  }
  // one step of the march; B = f % PF (a compile-time constant, so the
  // register buffers are indexed statically)
  auto step = [&](int f, auto parity) {
    constexpr int B = decltype(parity)::value, B1 = (B + 1) % PF;
    // slot of plane f+1 held f-2: last read by the residual of f-1 and the
    // restriction of f-2, both before the previous barrier
    commit_u(f + 1, ldU[B1]);
    T rho[NLR];
#pragma unroll
    for (int m = 0; m < NLR; m++) rho[m] = ldR[B][m];
    if (f + 1 + PF <= kf1 + 1) issue_u(f + 1 + PF, ldU[B1]);
    if (f + PF <= kf1) issue_rho(f + PF, ldR[B]);
    __syncthreads();
    res_plane(f, rho);  // res slot f&1 was last read by the restriction of f-2
    __syncthreads();
    restrict_plane(f);
  };
  for (int f = kf0; f <= kf1; f++) {
    if (PF == 2 && (f & 1)) step(f, std::integral_constant<int, (PF == 2 ? 1 : 0)>{});
    else step(f, std::integral_constant<int, 0>{});
  }
  // x/y-face columns: dmfwr3d's point formula on the fine u, after the march
  // (it reads global memory and would otherwise hold registers through it):
  // columns 0 and rows 0 of the tile, the last column (cnx-1) on the tiles of
  // the high x edge, the last row (cny-1) on those of the high y edge
  auto put_face = [&](int fi, int fj) {
    for (int kc = ck0; kc < ck1; kc++)
      put(fi, fj, kc, fw_point<T>(P.u, nx, ny, nz, 2 * fi + 1, 2 * fj + 1, 2 * kc + 1), T(0));
  };
  if (mine && !col_in) put_face(ci, cj);
  if (tx == ntx - 1 && t <= CY) {
    const int fj = t < CY ? cy0 + t : P.cny - 1;
    if ((t < CY && fj < P.cny - 1) || (t == CY && ty == nty - 1)) put_face(P.cnx - 1, fj);
  }
  if (ty == nty - 1 && t > CY && t <= CY + CX) {
    const int fi = cx0 + t - CY - 1;
    if (fi < P.cnx - 1) put_face(fi, P.cny - 1);
  }
  norm_partial<NT>(s, cnt, P.ws, P.slot, blockIdx.x, P.nb);
}

// ---------------------------------------------------------------------------
// Bottom of the V-cycle in one workgroup.  Levels 1..nb (all small) are
// processed from L_op + compute_MG_rhs on the top bottom level, through its
// pre-sweeps, residual and injection, down to the coarsest solve and back up
// through the coarse-grid corrections and post-sweeps (amrd/mg.c:84-201
// restricted to those levels).  The caller has already computed the residual
// of the level above and injected u and f_res into the top bottom level.
// Sweeps update in place colour by colour, as the hook does (the fused
// ping-pong sweep used on large levels gives the same values), so a whole
// bottom V costs one launch instead of about ten per level.
// ---------------------------------------------------------------------------
template <typename T>
struct BotLev {
  T *u;
  const T *rho, *a, *c;
  T *rhs, *res, *lop;
  int nx, ny, nz;
  int bx0, bx1, by0, by1, bz0, bz1;  // relax / residual box
  int cx0, cx1, cy0, cy1, cz0, cz1;  // child-mask ON box
  Spacing<T> sp;
};

constexpr int kBotMax = 6;
constexpr int kBotThreads = 1024;

template <typename T>
struct BottomP {
  BotLev<T> lev[kBotMax];  // lev[0] = MG level 1 (coarsest)
  int nb;
  int npre, npst, max_citer;
  double crtol;
  T w0_r, w0_i;
  int rop_u, rop_res;
  NormWs ws;
  int relax_slot0;  // level L's last relax norm -> slot relax_slot0 + L - 1
  int resid_slot0;  // level L's residual norm -> slot resid_slot0 + L - 1
  double *norm_pre, *norm_post;
  int *citer;
  T *snap;  // a saved state of the coarsest level for cycle_watch (or null)
};

// LDS: every array of the bottom levels (u, rho, rhs, res, lop, and a, c for
// op 1) lives in dynamic LDS for the whole V, so the many barrier-separated
// phases of the small levels read and write LDS instead of L2.  u, rho, res
// (a, c) come in, u, rhs, res and lop go back out; the arithmetic is the
// same, so the results are the same bits as without LDS.
template <typename T, int OP, bool P2, bool LDS = false>
__global__ __launch_bounds__(kBotThreads) void k_bottom(BottomP<T> P) {
  constexpr int NT = kBotThreads;
  __shared__ double bres;
  extern __shared__ unsigned char bot_smem[];
  __shared__ BotLev<T> slv[LDS ? kBotMax : 1];
  const BotLev<T> *lv = P.lev;
  if constexpr (LDS) {
    T *m = reinterpret_cast<T *>(bot_smem);
    for (int l = 0; l < P.nb; l++) {
      const BotLev<T> &G = P.lev[l];
      const int n = G.nx * G.ny * G.nz;
      T *u = m, *rho = m + n, *rhs = m + 2 * n, *res = m + 3 * n, *lop = m + 4 * n;
      m += 5 * n;
      T *a = nullptr, *c = nullptr;
      if constexpr (OP == 1) {
        a = m;
        c = m + n;
        m += 2 * n;
      }
      for (int q = threadIdx.x; q < n; q += NT) {
        u[q] = G.u[q];
        rho[q] = G.rho[q];
        res[q] = G.res[q];
        if constexpr (OP == 1) {
          a[q] = G.a[q];
          c[q] = G.c[q];
        }
      }
      if (threadIdx.x == 0) {
        slv[l] = G;
        slv[l].u = u; slv[l].rho = rho; slv[l].rhs = rhs; slv[l].res = res; slv[l].lop = lop;
        slv[l].a = a; slv[l].c = c;
      }
    }
    __syncthreads();
    lv = slv;
  }
  auto put_norm = [&](int slot, double s, double c) {
    block_reduce2<NT>(s, c);
    if (threadIdx.x == 0) {
      double *p = P.ws.partials + (size_t)slot * P.ws.cap * 2;
      p[0] = s;
      p[1] = c;
      P.ws.slot_nb[slot] = 1;
      bres = sqrt(s / c);
    }
    __syncthreads();
    return bres;
  };
  auto box_pts = [&](const BotLev<T> &L, int &mi, int &mj, int &mk) {
    mi = L.bx1 - L.bx0 + 1; mj = L.by1 - L.by0 + 1; mk = L.bz1 - L.bz0 + 1;
    return (mi > 0 && mj > 0 && mk > 0) ? mi * mj * mk : 0;
  };
  // one red-black sweep in place; returns (sum res^2, count)
  auto sweep = [&](const BotLev<T> &L, double &s, double &cnt) {
    int mi, mj, mk;
    const int npts = box_pts(L, mi, mj, mk);
    const size_t sy = L.nx, sz = (size_t)L.nx * L.ny;
    s = 0.0; cnt = 0.0;
    const FastDiv dmi = make_fdiv(mi), dmj = make_fdiv(mj);
    for (int pass = 0; pass <= 1; pass++) {
      for (int q = threadIdx.x; q < npts; q += NT) {
        int i, j, k;
        box_ijk(q, dmi, dmj, i, j, k);
        i += L.bx0; j += L.by0; k += L.bz0;
        if (((i + j + k + 3 + pass) & 1) != 0) continue;
        const size_t p = ix3(i, j, k, L.nx, L.ny);
        T jac;
        const T r = lop_at<T, OP, P2>(L.u, L.rho, L.a, L.c, p, sy, sz, L.sp, &jac) - L.rhs[p];
        if constexpr (std::is_same<T, float>::value && OP == 0) L.u[p] = L.u[p] - r * L.sp.ijac;
        else L.u[p] = L.u[p] - div_jac<T, OP>(r, jac, L.sp);
        s += (double)r * (double)r;
        cnt += 1.0;
      }
      __syncthreads();
    }
  };
  auto lop_rhs = [&](const BotLev<T> &L) {
    const int n = L.nx * L.ny * L.nz;
    const size_t sy = L.nx, sz = (size_t)L.nx * L.ny;
    const FastDiv dnx = make_fdiv(L.nx), dny = make_fdiv(L.ny);
    for (int q = threadIdx.x; q < n; q += NT) {
      int i, j, k;
      box_ijk(q, dnx, dny, i, j, k);
      const bool on = i >= L.cx0 && i <= L.cx1 && j >= L.cy0 && j <= L.cy1 && k >= L.cz0 && k <= L.cz1;
      const bool interior = i >= 1 && i <= L.nx - 2 && j >= 1 && j <= L.ny - 2 && k >= 1 && k <= L.nz - 2;
      T lopv = T(0);
      if (on && interior) lopv = lop_at<T, OP, P2>(L.u, L.rho, L.a, L.c, (size_t)q, sy, sz, L.sp, nullptr);
      L.rhs[q] = on ? (-L.res[q]) * P.w0_r + lopv : T(0);
      L.lop[q] = L.u[q];
    }
    __syncthreads();
  };
  auto residual = [&](const BotLev<T> &L, double &s, double &cnt) {
    int mi, mj, mk;
    const int npts = box_pts(L, mi, mj, mk);
    const size_t sy = L.nx, sz = (size_t)L.nx * L.ny;
    s = 0.0; cnt = 0.0;
    const FastDiv dmi = make_fdiv(mi), dmj = make_fdiv(mj);
    for (int q = threadIdx.x; q < npts; q += NT) {
      int i, j, k;
      box_ijk(q, dmi, dmj, i, j, k);
      i += L.bx0; j += L.by0; k += L.bz0;
      const size_t p = ix3(i, j, k, L.nx, L.ny);
      const T r = lop_at<T, OP, P2>(L.u, L.rho, L.a, L.c, p, sy, sz, L.sp, nullptr) - L.rhs[p];
      L.res[p] = r;
      s += (double)r * (double)r;
      cnt += 1.0;
    }
    __syncthreads();
  };
  auto restrict_to = [&](const T *f, T *cdst, const BotLev<T> &F, const BotLev<T> &C, int kind) {
    const int n = C.nx * C.ny * C.nz;
    const FastDiv dnx = make_fdiv(C.nx), dny = make_fdiv(C.ny);
    for (int q = threadIdx.x; q < n; q += NT) {
      int ci, cj, ck;
      box_ijk(q, dnx, dny, ci, cj, ck);
      const int i0 = 1 + 2 * ci, j0 = 1 + 2 * cj, k0 = 1 + 2 * ck;
      T v;
      if (kind == 1) v = f[ix3(i0 - 1, j0 - 1, k0 - 1, F.nx, F.ny)];
      else if (kind == 2) v = hw_point<T, false>(f, nullptr, T(0), F.nx, F.ny, F.nz, i0, j0, k0);
      else v = fw_point<T>(f, F.nx, F.ny, F.nz, i0, j0, k0);
      cdst[q] = v;
    }
  };
  auto cgc = [&](const BotLev<T> &F, const BotLev<T> &C) {
    const int nc = C.nx * C.ny * C.nz;
    for (int q = threadIdx.x; q < nc; q += NT) C.res[q] = (C.u[q] - C.lop[q]) / P.w0_i;
    __syncthreads();
    const int n = F.nx * F.ny * F.nz;
    const T c0 = T(1 - (1 * 1.0) / 2), c1 = T((1 * 1.0) / 2);
    const FastDiv dnx = make_fdiv(F.nx), dny = make_fdiv(F.ny);
    for (int q = threadIdx.x; q < n; q += NT) {
      int i, j, k;
      box_ijk(q, dnx, dny, i, j, k);
      const int ai = i >> 1, ri = i & 1, aj = j >> 1, rj = j & 1, ak = k >> 1, rk = k & 1;
      auto R = [&](int a, int b, int c) -> T { return C.res[ix3(a, b, c, C.nx, C.ny)]; };
      auto X = [&](int b, int c) -> T { return ri == 0 ? R(ai, b, c) : c0 * R(ai, b, c) + c1 * R(ai + 1, b, c); };
      auto Y = [&](int c) -> T { return rj == 0 ? X(aj, c) : c0 * X(aj, c) + c1 * X(aj + 1, c); };
      const T v = rk == 0 ? Y(ak) : c0 * Y(ak) + c1 * Y(ak + 1);
      F.u[q] = F.u[q] + v;
    }
    __syncthreads();
  };

  const int top = P.nb - 1;
  double s, cnt;
  for (int l = top; l >= 1; l--) {
    const BotLev<T> &L = lv[l];
    lop_rhs(L);
    for (int it = 0; it < P.npre; it++) sweep(L, s, cnt);
    residual(L, s, cnt);
    put_norm(P.resid_slot0 + l, s, cnt);
    if (P.rop_u) restrict_to(L.u, lv[l - 1].u, L, lv[l - 1], P.rop_u);
    if (P.rop_res) restrict_to(L.res, lv[l - 1].res, L, lv[l - 1], P.rop_res);
    __syncthreads();
  }
  // coarsest level (amrd/mg.c:135-141, 161-190 at L = 1)
  constexpr bool IJAC = std::is_same<T, float>::value && OP == 0;
  __shared__ T cb[LDS ? 1 : (OP == 1 ? 5 : 3) * kWaveCoarseMax];
  int cmi, cmj, cmk;
  const int cnpts = box_pts(lv[0], cmi, cmj, cmk);
  const size_t cn0 = (size_t)lv[0].nx * lv[0].ny * lv[0].nz;
  if (cnpts <= kWave && cn0 <= (size_t)kWaveCoarseMax) {
    // one wave, same bits (see wave_coarsest); staged into LDS unless there
    const BotLev<T> &L = lv[0];
    lop_rhs(L);
    T *wu = L.u;
    const T *wrho = L.rho, *wrhs = L.rhs, *wa = L.a, *wc = L.c;
    if constexpr (!LDS) {
      T *cu = cb, *crho = cb + cn0, *crhs = cb + 2 * cn0, *ca = nullptr, *cc = nullptr;
      if constexpr (OP == 1) {
        ca = cb + 3 * cn0;
        cc = cb + 4 * cn0;
      }
      for (size_t q = threadIdx.x; q < cn0; q += NT) {
        cu[q] = L.u[q];
        crho[q] = L.rho[q];
        crhs[q] = L.rhs[q];
        if constexpr (OP == 1) {
          ca[q] = L.a[q];
          cc[q] = L.c[q];
        }
      }
      __syncthreads();
      wu = cu; wrho = crho; wrhs = crhs; wa = ca; wc = cc;
    }
    if (threadIdx.x < kWave) {
      WaveCoarse<T> W{wu, wrho, wrhs, wa, wc, nullptr, L.nx, L.ny, L.bx0, L.by0, L.bz0, cmi, cmj, cnpts, L.sp,
                      P.npre, P.npst, P.max_citer, P.crtol,
                      P.ws.partials + (size_t)P.relax_slot0 * P.ws.cap * 2, P.ws.slot_nb + P.relax_slot0,
                      P.snap, cn0};
      int iter;
      double rpre;
      const double res = wave_coarsest<T, OP, false, P2, IJAC>(W, iter, rpre);
      if (threadIdx.x == 0) {
        if (P.norm_pre) P.norm_pre[0] = rpre;
        if (P.norm_post) P.norm_post[0] = res;
        if (P.citer) P.citer[0] = iter;
      }
      if constexpr (!LDS)
        for (size_t q = threadIdx.x; q < cn0; q += kWave) L.u[q] = wu[q];
    }
    __syncthreads();
  } else {
    const BotLev<T> &L = lv[0];
    lop_rhs(L);
    double res = 0.0;
    for (int it = 0; it < P.npre; it++) { sweep(L, s, cnt); res = put_norm(P.relax_slot0, s, cnt); }
    const double ctol = res * P.crtol;
    if (threadIdx.x == 0 && P.norm_pre) P.norm_pre[0] = res;
    int iter;
    for (iter = 0; iter < P.npst; iter++) { sweep(L, s, cnt); res = put_norm(P.relax_slot0, s, cnt); }
    res = ctol + 1;
    CycleWatch cw;
    int wit = 0;
    const size_t n0 = (size_t)L.nx * L.ny * L.nz;
    while (iter < P.max_citer && res > ctol) {
      sweep(L, s, cnt);
      res = put_norm(P.relax_slot0, s, cnt);
      iter++;
      if (P.snap && ++wit >= kWatchStart && res > ctol && iter < P.max_citer) {
        const int r = cycle_watch<T, NT>(L.u, n0, P.snap, cw, iter, P.max_citer);
        if (r >= 0) {
          for (int k = 0; k < r; k++) {
            sweep(L, s, cnt);
            res = put_norm(P.relax_slot0, s, cnt);
          }
          iter = P.max_citer;
          break;
        }
      }
    }
    if (threadIdx.x == 0) {
      if (P.norm_post) P.norm_post[0] = res;
      if (P.citer) P.citer[0] = iter;
    }
  }
  for (int l = 1; l <= top; l++) {
    const BotLev<T> &L = lv[l];
    cgc(L, lv[l - 1]);
    for (int it = 0; it < P.npst; it++) sweep(L, s, cnt);
    if (P.npst > 0) put_norm(P.relax_slot0 + l, s, cnt);
  }
  if constexpr (LDS) {
    __syncthreads();
    for (int l = 0; l < P.nb; l++) {
      const BotLev<T> &G = P.lev[l];
      const int n = G.nx * G.ny * G.nz;
      for (int q = threadIdx.x; q < n; q += NT) {
        G.u[q] = lv[l].u[q];
        G.rhs[q] = lv[l].rhs[q];
        G.res[q] = lv[l].res[q];
        G.lop[q] = lv[l].lop[q];
      }
    }
  }
}

// elementwise helpers ---------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_copy(T *dst, const T *src, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *dst, T v, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = v;
}

template <typename TD, typename TS>
__global__ __launch_bounds__(256) void k_convert(TD *dst, const TS *src, size_t n) {
  size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = (TD)src[q];
}

}  // namespace pmg
