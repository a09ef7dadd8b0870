// pmg_learned.hip -- the learned smoother / warm-start network on gfx950.
//
// Network: 3 layers of 3x3x3 convolutions, 1 -> H -> H -> 1 channels
// (H = 16), GELU (tanh form) after layers 1 and 2, zero padding.  It maps a
// normalised residual r/s to a normalised error estimate; the solver applies
// u += s * h^2 * N(r/s) and keeps the step only when the FP64 residual says
// it contracts (pmg_solver's learned mode).
//
// Kernels are implicit GEMMs on the MFMA units: every layer is computed as a
// 16 -> 16 channel convolution in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16): a wavefront owns 16 consecutive points of an
// x row and all 16 output channels, and accumulates the 27 taps as 27 MFMAs
// with K = 16 input channels.  Layer 1's single input channel and layer 3's
// single output channel are zero-padded to 16 (their padding rows/columns
// stay exactly zero under training).  Activations are channel-last BF16
// [point][16]; pre-activations FP32.
//
// Training kernels: backward-data is the same convolution with flipped taps
// and transposed weights, fused with the GELU derivative; backward-weight is
// a GEMM over points (K = points) per tap with per-block partials folded
// deterministically; Adam on FP32 master weights.
#include <cmath>
#include <cstring>
#include <vector>

#include "pmg_common.h"
#include "../../include/pamr_mg_gpu.h"

using namespace pmg;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int NC = 16;      // channels (padded)
constexpr int NTAP = 27;

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ inline float gelu_f(float x) {
  const float k = 0.7978845608028654f;
  return 0.5f * x * (1.0f + tanhf(k * (x + 0.044715f * x * x * x)));
}
__device__ inline float gelu_grad_f(float x) {
  const float k = 0.7978845608028654f;
  const float x2 = x * x;
  const float t = tanhf(k * (x + 0.044715f * x2 * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * x2);
}

struct Geom {
  int nx, ny, nz, nb;  // box shape and number of boxes (batch)
};

// ---------------------------------------------------------------------------
// MFMA lane layout self-test: D = A * B for 16x16 bf16 operands.  Lane l
// holds A[l%16][4*(l/16)+j], B[4*(l/16)+j][l%16] and D[4*(l/16)+i][l%16].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mfma_probe(const uint16_t *A, const uint16_t *B, float *D) {
  const int l = threadIdx.x, m = l & 15, g = l >> 4;
  bf16x4 a, b;
  for (int j = 0; j < 4; j++) {
    a[j] = (short)A[m * 16 + 4 * g + j];
    b[j] = (short)B[(4 * g + j) * 16 + m];
  }
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int i = 0; i < 4; i++) D[(4 * g + i) * 16 + m] = acc[i];
}

// The K = 32 form (v_mfma_f32_16x16x32_bf16): lane l holds A[l%16][8*(l/16)+j]
// and B[8*(l/16)+j][l%16], j = 0..7; D as above.
__global__ __launch_bounds__(64) void k_mfma32_probe(const uint16_t *A, const uint16_t *B, float *D) {
  const int l = threadIdx.x, m = l & 15, g = l >> 4;
  uint16_t a[8], b[8];
  for (int j = 0; j < 8; j++) {
    a[j] = A[m * 32 + 8 * g + j];
    b[j] = B[(8 * g + j) * 16 + m];
  }
  bf16x8 av, bv;
  __builtin_memcpy(&av, a, 16);
  __builtin_memcpy(&bv, b, 16);
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int i = 0; i < 4; i++) D[(4 * g + i) * 16 + m] = acc[i];
}

// ---------------------------------------------------------------------------
// 16 -> 16 convolution.  Grid (ceil(nx/16), ceil(ny/4), nz*nb), 4 waves per
// block, one x row segment of 16 points per wave.
//   EPI 0: z = conv + bias -> Z (fp32), A = bf16(gelu(z))
//   EPI 1: z = conv + bias -> Z (fp32) only (last layer)
//   EPI 2: backward data: v = conv * gelu'(aux) -> Z (fp32), A = bf16(v)
// Wf: weight fragments [27][64 lanes][4] bf16 (B operand order).
// ---------------------------------------------------------------------------
template <int EPI>
__global__ __launch_bounds__(256) void k_conv16(const uint16_t *__restrict__ X,
                                                const uint16_t *__restrict__ Wf,
                                                const float *__restrict__ bias,
                                                const float *__restrict__ aux,
                                                float *__restrict__ Zout,
                                                uint16_t *__restrict__ Aout, Geom g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 16, y = blockIdx.y * 4 + w;
  const int z = blockIdx.z % g.nz, b = blockIdx.z / g.nz;
  if (y >= g.ny) return;  // uniform per wave; no block barrier below
  const int m = lane & 15, grp = lane >> 4;
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const uint16_t *Xb = X + (size_t)b * box * NC;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
  for (int tap = 0; tap < NTAP; tap++) {
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const int xx = x0 + m + dx, yy = y + dy, zz = z + dz;
    bf16x4 a = {0, 0, 0, 0};
    if (xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz)
      a = *(const bf16x4 *)(Xb + (((size_t)zz * g.ny + yy) * g.nx + xx) * NC + 4 * grp);
    const bf16x4 bw = *(const bf16x4 *)(Wf + ((size_t)tap * 64 + lane) * 4);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int xx = x0 + 4 * grp + i;
    if (xx >= g.nx) continue;
    const size_t pp = (size_t)b * box + ((size_t)z * g.ny + y) * g.nx + xx;
    const size_t p = pp * NC + m;
    float v = acc[i];
    if constexpr (EPI == 0) {
      v += bias[m];
      Zout[p] = v;
      Aout[p] = f2bf(gelu_f(v));
    } else if constexpr (EPI == 1) {
      Zout[p] = v + bias[m];
    } else {
      v *= gelu_grad_f(aux[p]);
      Zout[p] = v;
      Aout[p] = f2bf(v);
    }
  }
}

// GELU (tanh form) through the logistic function: 0.5 x (1 + tanh u) =
// x / (1 + e^(-2u)), u = k (x + 0.044715 x^3); the inference epilogues use it
// (the training kernels keep tanhf; the two agree to FP32 rounding)
__device__ inline float gelu_fast(float x) {
  const float k = 0.7978845608028654f;
  const float u = k * (x + 0.044715f * x * x * x);
  return x / (1.0f + __expf(-2.0f * u));
}

// Stores one MFMA result tile of BF16 activations channel-last: lane (m, grp)
// holds channel m of points 4grp + i (v[i], i = 0..3); after a shuffle lane L
// holds channels 4(L%4) .. +3 of point L/4 and writes them as one 8-byte store,
// so the wave writes 16 points x 32 B contiguously.  `row` is the element
// offset of the group's point 0, `npts` how many of its 16 points exist.
__device__ inline void store_act16(uint16_t *__restrict__ Aout, size_t row, int npts, const uint16_t (&v)[4]) {
  const int L = threadIdx.x & 63;
  const unsigned w01 = (unsigned)v[0] | ((unsigned)v[1] << 16), w23 = (unsigned)v[2] | ((unsigned)v[3] << 16);
  const int p = L >> 2, c0 = (L & 3) * 4, i = p & 3, sg = p >> 2;
  unsigned out[2];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int src = (c0 + k) + 16 * sg;
    const unsigned a = (unsigned)__builtin_amdgcn_ds_bpermute(src * 4, (int)w01);
    const unsigned b = (unsigned)__builtin_amdgcn_ds_bpermute(src * 4, (int)w23);
    const unsigned word = i < 2 ? a : b;
    const unsigned h = (i & 1) ? (word >> 16) : (word & 0xffffu);
    if (k & 1) out[k >> 1] |= h << 16;
    else out[k >> 1] = h;
  }
  if (p < npts) *(uint2 *)(Aout + (row + (size_t)p) * NC + c0) = make_uint2(out[0], out[1]);
}

// Hidden and last layer of the inference path with the input staged in LDS:
// a block owns a CTX x CTY tile of (x, y) columns and marches ZC planes in
// z, keeping a ring of 4 input planes (tile plus a 1-point halo, 16 BF16
// channels = 32 B per point, zero outside the box) in LDS.  Wave w computes
// rows 2w, 2w+1 of the tile, two 16-point segments each.  The 27 taps go in
// pairs through the K = 32 MFMA (14 per segment and plane, the last pair
// padded with zeros): lanes 0-31 read 16 channel bytes of the pair's first
// tap, lanes 32-63 of its second, as one 16-byte LDS read each; the 14 weight
// fragments stay in registers.  One barrier per plane: plane z+2 is stored
// into the slot plane z-2 left at the start of step z, plane z+3 is fetched
// into registers, and step z reads z-1..z+1.
//   EPI 3: A = bf16(gelu(conv + bias));  EPI 4: Z[p] = conv + bias, channel 0
constexpr int CTX = 32, CTY = 8, CLX = CTX + 2, CLY = CTY + 2;
constexpr int CCHUNK = CLY * CLX * 2;                  // 16-byte chunks per staged plane
constexpr int CLOAD = (CCHUNK + 255) / 256;            // per thread
constexpr int NPAIR = (NTAP + 1) / 2;                  // tap pairs of the K = 32 form
template <int EPI, int ZC>
__global__ __launch_bounds__(256) void k_conv16_tile(const uint16_t *__restrict__ X,
                                                     const uint16_t *__restrict__ Wp,
                                                     const float *__restrict__ bias, float *__restrict__ Zout,
                                                     uint16_t *__restrict__ Aout, Geom g) {
  __shared__ uint4 sX[4][CLY][CLX][2];
  const int ntx = (g.nx + CTX - 1) / CTX, nty = (g.ny + CTY - 1) / CTY, nzc = (g.nz + ZC - 1) / ZC;
  int t = blockIdx.x;
  const int tx = t % ntx;
  t /= ntx;
  const int ty = t % nty;
  t /= nty;
  const int zc = t % nzc, b = t / nzc;
  const int x0 = tx * CTX, y0 = ty * CTY, z0 = zc * ZC, z1 = min(g.nz, z0 + ZC);
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const uint16_t *Xb = X + (size_t)b * box * NC;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, m = lane & 15, grp = lane >> 4;
  uint4 reg[CLOAD];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < CLOAD; q++) {
      const int idx = (int)threadIdx.x + q * 256;
      reg[q] = make_uint4(0u, 0u, 0u, 0u);
      if (idx >= CCHUNK || k < 0 || k >= g.nz) continue;
      const int ly = idx / (CLX * 2), rem = idx % (CLX * 2), lx = rem >> 1, h = rem & 1;
      const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
      if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny)
        reg[q] = ((const uint4 *)(Xb + (((size_t)k * g.ny + gy) * g.nx + gx) * NC))[h];
    }
  };
  auto store = [&](int k) {
    uint4 *S = &sX[k & 3][0][0][0];
#pragma unroll
    for (int q = 0; q < CLOAD; q++) {
      const int idx = (int)threadIdx.x + q * 256;
      if (idx < CCHUNK) S[idx] = reg[q];
    }
  };
  // the 14 weight fragments stay in registers for the whole march
  bf16x8 wk[NPAIR];
#pragma unroll
  for (int pr = 0; pr < NPAIR; pr++) wk[pr] = *(const bf16x8 *)(Wp + ((size_t)pr * 64 + lane) * 8);
  const int half = grp & 1, second = grp >> 1;  // channel half; first or second tap of the pair
  const float bm = bias[EPI == 3 ? m : 0];
  fetch(z0 - 1);
  store(z0 - 1);
  fetch(z0);
  store(z0);
  fetch(z0 + 1);
  store(z0 + 1);
  fetch(z0 + 2);
  __syncthreads();
  for (int z = z0; z < z1; z++) {
    store(z + 2);  // the slot of plane z-2, free since the barrier ending step z-1
    fetch(z + 3);
    floatx4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pr = 0; pr < NPAIR; pr++) {
      const int tap = 2 * pr + second;
      const bool live = tap < NTAP;
      const int tp = live ? tap : 0;
      const int dx = tp % 3 - 1, dy = (tp / 3) % 3 - 1, dz = tp / 9 - 1;
      const int slot = (z + dz) & 3;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = 2 * w + (q >> 1), sg = q & 1;
        bf16x8 a = *(const bf16x8 *)&sX[slot][r + 1 + dy][16 * sg + m + 1 + dx][half];
        if (!live) a = bf16x8{};
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wk[pr], acc[q], 0, 0, 0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int y = y0 + 2 * w + (q >> 1);
      if (y >= g.ny) continue;  // uniform per wave
      const int xs = x0 + 16 * (q & 1);
      const size_t row = (size_t)b * box + ((size_t)z * g.ny + y) * g.nx + xs;
      if constexpr (EPI == 3) {
        uint16_t v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = f2bf(gelu_fast(acc[q][i] + bm));
        store_act16(Aout, row, min(16, g.nx - xs), v);
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int xx = xs + 4 * grp + i;
          if (m == 0 && xx < g.nx) Zout[row + 4 * grp + i] = acc[q][i] + bm;
        }
      }
    }
    __syncthreads();
  }
}

// Layer 1 of the inference path on the same tiles: the plane tile of
// bf16(scale * x) (one channel, 2 B per point, zero outside the box) is staged
// in a 4-plane LDS ring; lane (m, grp) gathers its taps 8grp + j from LDS for
// one K = 32 MFMA whose K is the 27 taps (5 zero).
template <int ZC>
__global__ __launch_bounds__(256) void k_conv_in_tile(const float *__restrict__ x, float scale,
                                                      const uint16_t *__restrict__ Wfi,
                                                      const float *__restrict__ bias, uint16_t *__restrict__ Aout,
                                                      Geom g) {
  __shared__ uint16_t sx[4][CLY][CLX];
  const int ntx = (g.nx + CTX - 1) / CTX, nty = (g.ny + CTY - 1) / CTY, nzc = (g.nz + ZC - 1) / ZC;
  int t = blockIdx.x;
  const int tx = t % ntx;
  t /= ntx;
  const int ty = t % nty;
  t /= nty;
  const int zc = t % nzc, b = t / nzc;
  const int x0 = tx * CTX, y0 = ty * CTY, z0 = zc * ZC, z1 = min(g.nz, z0 + ZC);
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const float *xb = x + (size_t)b * box;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, m = lane & 15, grp = lane >> 4;
  constexpr int NPT = CLY * CLX, NLD = (NPT + 255) / 256;
  uint16_t reg[NLD];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < NLD; q++) {
      const int idx = (int)threadIdx.x + q * 256;
      reg[q] = 0;
      if (idx >= NPT || k < 0 || k >= g.nz) continue;
      const int ly = idx / CLX, lx = idx % CLX, gx = x0 - 1 + lx, gy = y0 - 1 + ly;
      if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) reg[q] = f2bf(xb[((size_t)k * g.ny + gy) * g.nx + gx] * scale);
    }
  };
  auto store = [&](int k) {
    uint16_t *S = &sx[k & 3][0][0];
#pragma unroll
    for (int q = 0; q < NLD; q++) {
      const int idx = (int)threadIdx.x + q * 256;
      if (idx < NPT) S[idx] = reg[q];
    }
  };
  const bf16x8 bw = *(const bf16x8 *)(Wfi + (size_t)lane * 8);
  const float bm = bias[m];
  fetch(z0 - 1);
  store(z0 - 1);
  fetch(z0);
  store(z0);
  fetch(z0 + 1);
  store(z0 + 1);
  fetch(z0 + 2);
  __syncthreads();
  for (int z = z0; z < z1; z++) {
    store(z + 2);
    fetch(z + 3);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = 2 * w + (q >> 1), sg = q & 1;
      uint16_t av[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int tap = 8 * grp + j;
        av[j] = 0;
        if (tap < NTAP) {
          const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
          av[j] = sx[(z + dz) & 3][r + 1 + dy][16 * sg + m + 1 + dx];
        }
      }
      bf16x8 a;
      __builtin_memcpy(&a, av, 16);
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc, 0, 0, 0);
      const int y = y0 + r;
      if (y < g.ny) {  // uniform per wave
        const int xs = x0 + 16 * sg;
        uint16_t v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = f2bf(gelu_fast(acc[i] + bm));
        store_act16(Aout, (size_t)b * box + ((size_t)z * g.ny + y) * g.nx + xs, min(16, g.nx - xs), v);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][ci][co] = sum_p X[p+off(tap)][ci] * dZ[p][co];
// tap 27 is the bias (input channel 0 == 1).  Grid (nchunk, 28): block
// (chunk, tap), 4 waves each summing every 4th group of 16 points of the
// chunk with MFMA (M = ci, N = co, K = points), then a fixed-order fold of
// the 4 waves into partial[chunk][tap][16][16].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv16_wgrad(const uint16_t *__restrict__ X,
                                                      const uint16_t *__restrict__ dZ,
                                                      float *__restrict__ partial, Geom g,
                                                      int chunk) {
  __shared__ float red[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tap = blockIdx.y;
  const int m = lane & 15, grp = lane >> 4;
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const size_t total = box * g.nb;
  const size_t p0 = (size_t)blockIdx.x * chunk;
  const size_t p1 = p0 + chunk < total ? p0 + chunk : total;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (size_t base = p0 + 16 * w; base < p1; base += 64) {
    bf16x4 a = {0, 0, 0, 0}, bb = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const size_t p = base + 4 * grp + j;
      if (p >= p1) continue;
      bb[j] = (short)dZ[p * NC + m];
      if (tap == NTAP) {
        a[j] = (short)(m == 0 ? 0x3f80 : 0);  // bf16(1.0) on channel 0
      } else {
        const size_t q = p % box;
        const size_t bidx = p / box;
        const int x = (int)(q % g.nx), y = (int)((q / g.nx) % g.ny), z = (int)(q / ((size_t)g.nx * g.ny));
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz)
          a[j] = (short)X[(bidx * box + ((size_t)zz * g.ny + yy) * g.nx + xx) * NC + m];
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
  // lane holds D[ci = 4*grp + i][co = m]
#pragma unroll
  for (int i = 0; i < 4; i++) red[w][(4 * grp + i) * 16 + m] = acc[i];
  __syncthreads();
  const int e = threadIdx.x;  // 256 entries
  const float v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
  partial[((size_t)blockIdx.x * (NTAP + 1) + tap) * 256 + e] = v;
}

// fold partials over chunks (fixed order) -> grad[28][256]
__global__ __launch_bounds__(256) void k_fold_wgrad(const float *__restrict__ partial,
                                                    float *__restrict__ grad, int nchunk) {
  const int tap = blockIdx.x, e = threadIdx.x;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += partial[((size_t)c * (NTAP + 1) + tap) * 256 + e];
  grad[tap * 256 + e] = s;
}

// loss = 0.5 * mean over counted points of (Z[p][0] - T[p])^2 ; dZ[p][0] =
// (Z - T)/count (bf16), other channels 0.  Points within `margin` of a box
// face are not counted.  Per-block partial sums of the squared error.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ Z, const float *__restrict__ T,
                                              const float *__restrict__ W, const double *__restrict__ wcount,
                                              uint16_t *__restrict__ dZ, double *__restrict__ part,
                                              Geom g, int margin, double inv_count) {
  // W (optional): per-point weights 0/1 on top of the margin rule; the
  // mean then runs over the weighted points, whose count *wcount holds
  __shared__ double red[256];
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (wcount) inv_count = *wcount > 0 ? 1.0 / *wcount : 0.0;
  double e2 = 0.0;
  if (p < box * g.nb) {
    const size_t q = p % box;
    const int x = (int)(q % g.nx), y = (int)((q / g.nx) % g.ny), z = (int)(q / ((size_t)g.nx * g.ny));
    bool on = x >= margin && x < g.nx - margin && y >= margin && y < g.ny - margin &&
              z >= margin && z < g.nz - margin;
    if (W) on = on && W[p] != 0.f;
    float d = 0.f;
    if (on) {
      d = Z[p * NC] - T[p];
      e2 = (double)d * (double)d;
    }
    dZ[p * NC] = f2bf((float)(d * inv_count));
#pragma unroll
    for (int c = 1; c < NC; c++) dZ[p * NC + c] = 0;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// partial counts of the weighted points (margin rule and W != 0), one per block
__global__ __launch_bounds__(256) void k_wcount(const float *__restrict__ W, Geom g, int margin,
                                                double *__restrict__ part) {
  __shared__ double red[256];
  const size_t box = (size_t)g.nx * g.ny * g.nz;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (p < box * g.nb) {
    const size_t q = p % box;
    const int x = (int)(q % g.nx), y = (int)((q / g.nx) % g.ny), z = (int)(q / ((size_t)g.nx * g.ny));
    if (x >= margin && x < g.nx - margin && y >= margin && y < g.ny - margin && z >= margin &&
        z < g.nz - margin && W[p] != 0.f)
      v = 1.0;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_fold_d(const double *__restrict__ part, int n, double *out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// input packing: X[p][0] = bf16(src[p] * scale), channels 1..15 = 0
template <typename TS>
__global__ __launch_bounds__(256) void k_pack_input(const TS *__restrict__ src, uint16_t *__restrict__ X,
                                                    size_t n, float scale) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  bf16x4 v = {0, 0, 0, 0};
  v[0] = (short)f2bf((float)src[p] * scale);
  const bf16x4 zero = {0, 0, 0, 0};
  bf16x4 *d = (bf16x4 *)(X + p * NC);
  d[0] = v;
  d[1] = zero;
  d[2] = zero;
  d[3] = zero;
}

// Adam on FP32 master weights (only the entries `mask` marks trainable)
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ m1,
                                              float *__restrict__ m2, const float *__restrict__ gr,
                                              const uint8_t *__restrict__ mask, int n, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !mask[i]) return;
  const float g = gr[i];
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float b = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// master layout per layer: W[tap][ci][co] (27*256 floats) then bias[co]
// (16 floats), i.e. 28 rows of 256 with the bias in row 27, entries 0..15
constexpr int LAYER_FLOATS = (NTAP + 1) * 256;

// fragments: forward Wf[tap][lane][j] = W[tap][4*(lane/16)+j][lane%16];
// backward-data (flipped taps, transposed): Wb[tap][lane][j] =
// W[26-tap][lane%16][4*(lane/16)+j]
__global__ __launch_bounds__(256) void k_pack_frags(const float *__restrict__ W, uint16_t *__restrict__ Wf,
                                                    uint16_t *__restrict__ Wb) {
  const int idx = blockIdx.x * 256 + threadIdx.x;  // over 27*64*4
  if (idx >= NTAP * 64 * 4) return;
  const int tap = idx / 256, lane = (idx / 4) % 64, j = idx % 4;
  const int m = lane & 15, grp = lane >> 4;
  Wf[idx] = f2bf(W[tap * 256 + (4 * grp + j) * 16 + m]);
  Wb[idx] = f2bf(W[(26 - tap) * 256 + m * 16 + (4 * grp + j)]);
}

// inference fragments for the K = 32 MFMA.  Layer 1 (K = taps of input
// channel 0): Wfi[lane][j] = W[tap = 8(lane/16) + j][ci 0][co lane%16], 0 past
// tap 26.  Hidden / last layer (K = a tap pair x 16 channels):
// Wp[pr][lane][j] = W[tap = 2 pr + (lane/16)/2][ci = 8((lane/16)%2) + j][co lane%16]
__global__ __launch_bounds__(256) void k_pack_frags_in(const float *__restrict__ W, uint16_t *__restrict__ Wfi) {
  const int idx = blockIdx.x * 256 + threadIdx.x;  // over 64*8
  if (idx >= 64 * 8) return;
  const int lane = idx / 8, j = idx % 8;
  const int tap = 8 * (lane >> 4) + j;
  Wfi[idx] = tap < NTAP ? f2bf(W[tap * 256 + (lane & 15)]) : (uint16_t)0;
}
__global__ __launch_bounds__(256) void k_pack_frags_pair(const float *__restrict__ W, uint16_t *__restrict__ Wp) {
  const int idx = blockIdx.x * 256 + threadIdx.x;  // over NPAIR*64*8
  if (idx >= NPAIR * 64 * 8) return;
  const int pr = idx / 512, lane = (idx / 8) % 64, j = idx % 8, grp = lane >> 4;
  const int tap = 2 * pr + (grp >> 1), ci = 8 * (grp & 1) + j;
  Wp[idx] = tap < NTAP ? f2bf(W[tap * 256 + ci * 16 + (lane & 15)]) : (uint16_t)0;
}

// u += coef * Z[p * STRIDE] at interior points of the grid (box = the grid)
template <typename TU, int STRIDE>
__global__ __launch_bounds__(256) void k_apply_correction(TU *__restrict__ u, const float *__restrict__ Z,
                                                          int nx, int ny, int nz, double coef) {
  const int i = blockIdx.x * 64 + (int)(threadIdx.x % 64);
  const int j = blockIdx.y * 4 + (int)(threadIdx.x / 64);
  const int k = blockIdx.z;
  if (i < 1 || j < 1 || k < 1 || i >= nx - 1 || j >= ny - 1 || k >= nz - 1) return;
  const size_t p = ix3(i, j, k, nx, ny);
  u[p] = u[p] + (TU)(coef * (double)Z[p * STRIDE]);
}

}  // namespace

// ===========================================================================
// network object
// ===========================================================================
struct pmg_net {
  int hidden = 16;
  float *w = nullptr;       // 3 layers * LAYER_FLOATS master weights (device)
  float *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  uint8_t *mask = nullptr;  // trainable entries
  uint16_t *wf = nullptr, *wb = nullptr;  // 3 layers * 27*64*4 fragments each
  uint16_t *wfi = nullptr;                // layer 1 with K = taps (inference path), 64*8
  uint16_t *wp = nullptr;                 // hidden and last layer, tap pairs: 2 * NPAIR*64*8
  int adam_t = 0;
  // activation buffers (grown on demand): X0, A1, A2, dA (bf16 [P][16]); Z1, Z2, Z3, dZ (fp32)
  size_t cap = 0;
  uint16_t *x0 = nullptr, *a1 = nullptr, *a2 = nullptr, *d16 = nullptr, *d16b = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *dz = nullptr;
  // inference buffers (their own capacity: inference allocates no training buffers)
  size_t icap = 0;
  uint16_t *ia1 = nullptr, *ia2 = nullptr;  // bf16 [P][16]
  float *y = nullptr;                       // output, one FP32 per point
  float *partial = nullptr;
  size_t partial_cap = 0;
  double *lpart = nullptr, *lsum = nullptr;
  size_t lpart_cap = 0;
  hipStream_t stream = nullptr;
  bool frags_valid = false;
};

namespace {

int net_ensure(pmg_net *N, size_t P) {
  if (P <= N->cap) return 0;
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  for (void *p : {(void *)N->x0, (void *)N->a1, (void *)N->a2, (void *)N->d16, (void *)N->d16b,
                  (void *)N->z1, (void *)N->z2, (void *)N->z3, (void *)N->dz})
    if (p) PMG_HOST_CHECK(hipFree(p));
  const size_t b16 = P * NC * sizeof(uint16_t), b32 = P * NC * sizeof(float);
  PMG_HOST_CHECK(hipMalloc(&N->x0, b16));
  PMG_HOST_CHECK(hipMalloc(&N->a1, b16));
  PMG_HOST_CHECK(hipMalloc(&N->a2, b16));
  PMG_HOST_CHECK(hipMalloc(&N->d16, b16));
  PMG_HOST_CHECK(hipMalloc(&N->d16b, b16));
  PMG_HOST_CHECK(hipMalloc(&N->z1, b32));
  PMG_HOST_CHECK(hipMalloc(&N->z2, b32));
  PMG_HOST_CHECK(hipMalloc(&N->z3, b32));
  PMG_HOST_CHECK(hipMalloc(&N->dz, b32));
  N->cap = P;
  return 0;
}

int net_ensure_infer(pmg_net *N, size_t P) {
  if (P <= N->icap) return 0;
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  for (void *p : {(void *)N->ia1, (void *)N->ia2, (void *)N->y})
    if (p) PMG_HOST_CHECK(hipFree(p));
  PMG_HOST_CHECK(hipMalloc(&N->ia1, P * NC * sizeof(uint16_t)));
  PMG_HOST_CHECK(hipMalloc(&N->ia2, P * NC * sizeof(uint16_t)));
  PMG_HOST_CHECK(hipMalloc(&N->y, P * sizeof(float)));
  N->icap = P;
  return 0;
}

int net_frags(pmg_net *N) {
  if (N->frags_valid) return 0;
  for (int l = 0; l < 3; l++) {
    hipLaunchKernelGGL(k_pack_frags, dim3((NTAP * 256 + 255) / 256), dim3(256), 0, N->stream,
                       N->w + (size_t)l * LAYER_FLOATS, N->wf + (size_t)l * NTAP * 256,
                       N->wb + (size_t)l * NTAP * 256);
  }
  hipLaunchKernelGGL(k_pack_frags_in, dim3(2), dim3(256), 0, N->stream, N->w, N->wfi);
  for (int l = 1; l < 3; l++)
    hipLaunchKernelGGL(k_pack_frags_pair, dim3((NPAIR * 512 + 255) / 256), dim3(256), 0, N->stream,
                       N->w + (size_t)l * LAYER_FLOATS, N->wp + (size_t)(l - 1) * NPAIR * 512);
  PMG_HOST_CHECK(hipGetLastError());
  N->frags_valid = true;
  return 0;
}

dim3 conv_grid(const Geom &g) { return dim3((g.nx + 15) / 16, (g.ny + 3) / 4, g.nz * g.nb); }

// z planes one block of the LDS-tiled inference kernels marches
constexpr int kTilePlanes = 16;

// inference: N(scale * x) into N->y (one FP32 per point).  Layer 1 reads x
// directly (K = taps), the hidden layer keeps only the BF16 activations, the
// last layer writes channel 0 only -- no pre-activations are stored
int net_infer(pmg_net *N, const float *x, float scale, const Geom &g) {
  int rc;
  if ((rc = net_frags(N))) return rc;
  const float *b1 = N->w + 0 * LAYER_FLOATS + NTAP * 256;
  const float *b2 = N->w + 1 * LAYER_FLOATS + NTAP * 256;
  const float *b3 = N->w + 2 * LAYER_FLOATS + NTAP * 256;
  constexpr int TZ = kTilePlanes;
  const dim3 gt((unsigned)(((g.nx + CTX - 1) / CTX) * ((g.ny + CTY - 1) / CTY) * ((g.nz + TZ - 1) / TZ) * g.nb));
  hipLaunchKernelGGL(k_conv_in_tile<TZ>, gt, dim3(256), 0, N->stream, x, scale, N->wfi, b1, N->ia1, g);
  hipLaunchKernelGGL((k_conv16_tile<3, TZ>), gt, dim3(256), 0, N->stream, N->ia1, N->wp, b2, (float *)nullptr,
                     N->ia2, g);
  hipLaunchKernelGGL((k_conv16_tile<4, TZ>), gt, dim3(256), 0, N->stream, N->ia2, N->wp + NPAIR * 512, b3, N->y,
                     (uint16_t *)nullptr, g);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// forward of the 3 layers on X0 (already packed); Z3[p][0] is the output
int net_forward(pmg_net *N, const Geom &g) {
  int rc;
  if ((rc = net_frags(N))) return rc;
  const dim3 gr = conv_grid(g);
  const float *b1 = N->w + 0 * LAYER_FLOATS + NTAP * 256;
  const float *b2 = N->w + 1 * LAYER_FLOATS + NTAP * 256;
  const float *b3 = N->w + 2 * LAYER_FLOATS + NTAP * 256;
  hipLaunchKernelGGL(k_conv16<0>, gr, dim3(256), 0, N->stream, N->x0, N->wf, b1, (const float *)nullptr,
                     N->z1, N->a1, g);
  hipLaunchKernelGGL(k_conv16<0>, gr, dim3(256), 0, N->stream, N->a1, N->wf + NTAP * 256, b2,
                     (const float *)nullptr, N->z2, N->a2, g);
  hipLaunchKernelGGL(k_conv16<1>, gr, dim3(256), 0, N->stream, N->a2, N->wf + 2 * NTAP * 256, b3,
                     (const float *)nullptr, N->z3, (uint16_t *)nullptr, g);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int net_wgrad(pmg_net *N, const Geom &g, const uint16_t *X, const uint16_t *dZ, float *grad_out) {
  const size_t total = (size_t)g.nx * g.ny * g.nz * g.nb;
  const int chunk = 4096;
  const int nchunk = (int)((total + chunk - 1) / chunk);
  const size_t need = (size_t)nchunk * (NTAP + 1) * 256;
  if (need > N->partial_cap) {
    PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
    if (N->partial) PMG_HOST_CHECK(hipFree(N->partial));
    PMG_HOST_CHECK(hipMalloc(&N->partial, need * sizeof(float)));
    N->partial_cap = need;
  }
  hipLaunchKernelGGL(k_conv16_wgrad, dim3(nchunk, NTAP + 1), dim3(256), 0, N->stream, X, dZ, N->partial, g,
                     chunk);
  hipLaunchKernelGGL(k_fold_wgrad, dim3(NTAP + 1), dim3(256), 0, N->stream, N->partial, grad_out, nchunk);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// forward + loss + backward; gradients into N->grad (3 * LAYER_FLOATS)
int net_loss_backward(pmg_net *N, const Geom &g, const float *T, const float *W, int margin, double *loss_out) {
  int rc;
  if ((rc = net_forward(N, g))) return rc;
  const size_t total = (size_t)g.nx * g.ny * g.nz * g.nb;
  const int nblk = (int)((total + 255) / 256);
  if ((size_t)nblk > N->lpart_cap) {
    PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
    if (N->lpart) PMG_HOST_CHECK(hipFree(N->lpart));
    PMG_HOST_CHECK(hipMalloc(&N->lpart, (size_t)nblk * sizeof(double)));
    N->lpart_cap = nblk;
  }
  const long long cx = g.nx - 2 * margin, cy = g.ny - 2 * margin, cz = g.nz - 2 * margin;
  double count = (double)(cx > 0 ? cx : 0) * (cy > 0 ? cy : 0) * (cz > 0 ? cz : 0) * g.nb;
  if (W) {
    // the weighted count, folded on the device into lsum[1] before the loss uses it
    hipLaunchKernelGGL(k_wcount, dim3(nblk), dim3(256), 0, N->stream, W, g, margin, N->lpart);
    hipLaunchKernelGGL(k_fold_d, dim3(1), dim3(256), 0, N->stream, N->lpart, nblk, N->lsum + 1);
  }
  hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, N->stream, N->z3, T, W, W ? N->lsum + 1 : (const double *)nullptr,
                     N->d16, N->lpart, g, margin, 1.0 / count);
  hipLaunchKernelGGL(k_fold_d, dim3(1), dim3(256), 0, N->stream, N->lpart, nblk, N->lsum);
  PMG_HOST_CHECK(hipGetLastError());
  const dim3 gr = conv_grid(g);
  // layer 3: dW3 from A2 and dZ3; dZ2 = conv^T(dZ3) * gelu'(Z2)
  if ((rc = net_wgrad(N, g, N->a2, N->d16, N->grad + 2 * LAYER_FLOATS))) return rc;
  hipLaunchKernelGGL(k_conv16<2>, gr, dim3(256), 0, N->stream, N->d16, N->wb + 2 * NTAP * 256,
                     (const float *)nullptr, N->z2, N->dz, N->d16b, g);
  // layer 2
  if ((rc = net_wgrad(N, g, N->a1, N->d16b, N->grad + 1 * LAYER_FLOATS))) return rc;
  hipLaunchKernelGGL(k_conv16<2>, gr, dim3(256), 0, N->stream, N->d16b, N->wb + 1 * NTAP * 256,
                     (const float *)nullptr, N->z1, N->dz, N->d16, g);
  // layer 1
  if ((rc = net_wgrad(N, g, N->x0, N->d16, N->grad))) return rc;
  PMG_HOST_CHECK(hipGetLastError());
  if (loss_out) {
    double l[2];
    PMG_HOST_CHECK(hipMemcpyAsync(l, N->lsum, 2 * sizeof(double), hipMemcpyDeviceToHost, N->stream));
    PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
    if (W) count = l[1];
    *loss_out = count > 0 ? 0.5 * l[0] / count : 0.0;
  }
  return 0;
}

}  // namespace

extern "C" {

int pmg_net_mfma32_selftest(const float *A, const float *B, float *D_out) {
  // host A (16x32), B (32x16) rounded to bf16; D_out = A*B from one K = 32 MFMA
  std::vector<uint16_t> ha(512), hb(512);
  for (int i = 0; i < 512; i++) { ha[i] = f2bf(A[i]); hb[i] = f2bf(B[i]); }
  uint16_t *dA, *dB;
  float *dD;
  PMG_HOST_CHECK(hipMalloc(&dA, 1024));
  PMG_HOST_CHECK(hipMalloc(&dB, 1024));
  PMG_HOST_CHECK(hipMalloc(&dD, 1024));
  PMG_HOST_CHECK(hipMemcpy(dA, ha.data(), 1024, hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMemcpy(dB, hb.data(), 1024, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma32_probe, dim3(1), dim3(64), 0, 0, dA, dB, dD);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipMemcpy(D_out, dD, 1024, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return 0;
}

int pmg_net_mfma_selftest(const float *A, const float *B, float *D_out) {
  // host A (16x16), B (16x16) rounded to bf16; D_out = A*B from one MFMA
  std::vector<uint16_t> ha(256), hb(256);
  for (int i = 0; i < 256; i++) { ha[i] = f2bf(A[i]); hb[i] = f2bf(B[i]); }
  uint16_t *dA, *dB;
  float *dD;
  PMG_HOST_CHECK(hipMalloc(&dA, 512));
  PMG_HOST_CHECK(hipMalloc(&dB, 512));
  PMG_HOST_CHECK(hipMalloc(&dD, 1024));
  PMG_HOST_CHECK(hipMemcpy(dA, ha.data(), 512, hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMemcpy(dB, hb.data(), 512, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dA, dB, dD);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipMemcpy(D_out, dD, 1024, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return 0;
}

pmg_net *pmg_net_create(int hidden, unsigned long long seed) {
  if (hidden != 16) return nullptr;
  pmg_net *N = new pmg_net();
  if (hipStreamCreate(&N->stream) != hipSuccess) { delete N; return nullptr; }
  const size_t nf = 3 * (size_t)LAYER_FLOATS;
  std::vector<float> w(nf, 0.f);
  std::vector<uint8_t> mask(nf, 0);
  // splitmix64-driven uniform init, He-like scale; padding stays zero
  unsigned long long st = seed;
  auto rnd = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const int cin[3] = {1, 16, 16}, cout[3] = {16, 16, 1};
  for (int l = 0; l < 3; l++) {
    const double sc = std::sqrt(3.0 / (27.0 * cin[l]));
    for (int t = 0; t < NTAP; t++)
      for (int ci = 0; ci < cin[l]; ci++)
        for (int co = 0; co < cout[l]; co++) {
          const size_t idx = (size_t)l * LAYER_FLOATS + t * 256 + ci * 16 + co;
          w[idx] = (float)((2.0 * rnd() - 1.0) * sc);
          mask[idx] = 1;
        }
    for (int co = 0; co < cout[l]; co++) mask[(size_t)l * LAYER_FLOATS + NTAP * 256 + co] = 1;
  }
  bool ok = hipMalloc(&N->w, nf * 4) == hipSuccess && hipMalloc(&N->m1, nf * 4) == hipSuccess &&
            hipMalloc(&N->m2, nf * 4) == hipSuccess && hipMalloc(&N->grad, nf * 4) == hipSuccess &&
            hipMalloc(&N->mask, nf) == hipSuccess &&
            hipMalloc(&N->wf, 3 * NTAP * 256 * sizeof(uint16_t)) == hipSuccess &&
            hipMalloc(&N->wb, 3 * NTAP * 256 * sizeof(uint16_t)) == hipSuccess &&
            hipMalloc(&N->wfi, 64 * 8 * sizeof(uint16_t)) == hipSuccess &&
            hipMalloc(&N->wp, 2 * NPAIR * 512 * sizeof(uint16_t)) == hipSuccess &&
            hipMalloc(&N->lsum, 2 * sizeof(double)) == hipSuccess;
  if (!ok) { pmg_net_destroy(N); return nullptr; }
  (void)hipMemcpy(N->w, w.data(), nf * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(N->mask, mask.data(), nf, hipMemcpyHostToDevice);
  (void)hipMemset(N->m1, 0, nf * 4);
  (void)hipMemset(N->m2, 0, nf * 4);
  (void)hipMemset(N->grad, 0, nf * 4);
  return N;
}

void pmg_net_destroy(pmg_net *N) {
  if (!N) return;
  if (N->stream) (void)hipStreamSynchronize(N->stream);
  for (void *p : {(void *)N->w, (void *)N->m1, (void *)N->m2, (void *)N->grad, (void *)N->mask,
                  (void *)N->wf, (void *)N->wb, (void *)N->x0, (void *)N->a1, (void *)N->a2,
                  (void *)N->d16, (void *)N->d16b, (void *)N->z1, (void *)N->z2, (void *)N->z3,
                  (void *)N->dz, (void *)N->ia1, (void *)N->ia2, (void *)N->y, (void *)N->wfi, (void *)N->wp,
                  (void *)N->partial, (void *)N->lpart, (void *)N->lsum})
    if (p) (void)hipFree(p);
  if (N->stream) (void)hipStreamDestroy(N->stream);
  delete N;
}

// Weights in the oracle's layout (oracle/mg_oracle.h): w1[16][1][27], b1[16],
// w2[16][16][27], b2[16], w3[1][16][27], b3[1]; FP64 on the host.
int pmg_net_set_weights(pmg_net *N, const double *w1, const double *b1, const double *w2,
                        const double *b2, const double *w3, const double *b3) {
  if (!N) return -1;
  std::vector<float> w(3 * (size_t)LAYER_FLOATS, 0.f);
  for (int co = 0; co < 16; co++) {
    for (int t = 0; t < NTAP; t++) {
      w[0 * LAYER_FLOATS + t * 256 + 0 * 16 + co] = (float)w1[co * 27 + t];
      for (int ci = 0; ci < 16; ci++)
        w[1 * LAYER_FLOATS + t * 256 + ci * 16 + co] = (float)w2[(co * 16 + ci) * 27 + t];
    }
    w[0 * LAYER_FLOATS + NTAP * 256 + co] = (float)b1[co];
    w[1 * LAYER_FLOATS + NTAP * 256 + co] = (float)b2[co];
  }
  for (int ci = 0; ci < 16; ci++)
    for (int t = 0; t < NTAP; t++) w[2 * LAYER_FLOATS + t * 256 + ci * 16 + 0] = (float)w3[ci * 27 + t];
  w[2 * LAYER_FLOATS + NTAP * 256 + 0] = (float)b3[0];
  PMG_HOST_CHECK(hipMemcpy(N->w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
  N->frags_valid = false;
  return 0;
}

static void from_master(const std::vector<float> &w, double *w1, double *b1, double *w2, double *b2,
                        double *w3, double *b3) {
  for (int co = 0; co < 16; co++) {
    for (int t = 0; t < NTAP; t++) {
      w1[co * 27 + t] = w[0 * LAYER_FLOATS + t * 256 + co];
      for (int ci = 0; ci < 16; ci++) w2[(co * 16 + ci) * 27 + t] = w[1 * LAYER_FLOATS + t * 256 + ci * 16 + co];
    }
    b1[co] = w[0 * LAYER_FLOATS + NTAP * 256 + co];
    b2[co] = w[1 * LAYER_FLOATS + NTAP * 256 + co];
  }
  for (int ci = 0; ci < 16; ci++)
    for (int t = 0; t < NTAP; t++) w3[ci * 27 + t] = w[2 * LAYER_FLOATS + t * 256 + ci * 16];
  b3[0] = w[2 * LAYER_FLOATS + NTAP * 256];
}

int pmg_net_get_weights(pmg_net *N, double *w1, double *b1, double *w2, double *b2, double *w3,
                        double *b3) {
  if (!N) return -1;
  std::vector<float> w(3 * (size_t)LAYER_FLOATS);
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  PMG_HOST_CHECK(hipMemcpy(w.data(), N->w, w.size() * 4, hipMemcpyDeviceToHost));
  from_master(w, w1, b1, w2, b2, w3, b3);
  return 0;
}

// Forward on nb boxes of nx*ny*nz points: x (device fp32 [nb][P]) -> y
// (device fp32 [nb][P]).
int pmg_net_forward(pmg_net *N, const float *x, float *y, int nx, int ny, int nz, int nb) {
  if (!N || !x || !y || nx < 1 || ny < 1 || nz < 1 || nb < 1) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure(N, P)) return -1;
  Geom g{nx, ny, nz, nb};
  hipLaunchKernelGGL(k_pack_input<float>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, N->stream, x, N->x0,
                     P, 1.0f);
  int rc = net_forward(N, g);
  if (rc) return rc;
  // y = channel 0 of Z3
  PMG_HOST_CHECK(hipMemcpy2DAsync(y, sizeof(float), N->z3, NC * sizeof(float), sizeof(float), P,
                                  hipMemcpyDeviceToDevice, N->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  return 0;
}

// Loss and gradients (oracle layout, host FP64) for inputs x and targets t
// (device fp32 [nb][P]); margin = points near box faces excluded from the loss.
int pmg_net_loss_grad(pmg_net *N, const float *x, const float *t, int nx, int ny, int nz, int nb,
                      int margin, double *loss, double *gw1, double *gb1, double *gw2, double *gb2,
                      double *gw3, double *gb3) {
  if (!N || !x || !t) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure(N, P)) return -1;
  Geom g{nx, ny, nz, nb};
  hipLaunchKernelGGL(k_pack_input<float>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, N->stream, x, N->x0,
                     P, 1.0f);
  int rc = net_loss_backward(N, g, t, nullptr, margin, loss);
  if (rc) return rc;
  std::vector<float> gr(3 * (size_t)LAYER_FLOATS);
  PMG_HOST_CHECK(hipMemcpy(gr.data(), N->grad, gr.size() * 4, hipMemcpyDeviceToHost));
  from_master(gr, gw1, gb1, gw2, gb2, gw3, gb3);
  return 0;
}

// Loss and gradients as pmg_net_loss_grad, the mean taken over the points
// with w != 0 (w: device fp32 [nb][P]) that also satisfy the margin rule.
int pmg_net_loss_grad_w(pmg_net *N, const float *x, const float *t, const float *w, int nx, int ny, int nz,
                        int nb, int margin, double *loss, double *gw1, double *gb1, double *gw2, double *gb2,
                        double *gw3, double *gb3) {
  if (!N || !x || !t || !w) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure(N, P)) return -1;
  Geom g{nx, ny, nz, nb};
  hipLaunchKernelGGL(k_pack_input<float>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, N->stream, x, N->x0,
                     P, 1.0f);
  int rc = net_loss_backward(N, g, t, w, margin, loss);
  if (rc) return rc;
  std::vector<float> gr(3 * (size_t)LAYER_FLOATS);
  PMG_HOST_CHECK(hipMemcpy(gr.data(), N->grad, gr.size() * 4, hipMemcpyDeviceToHost));
  from_master(gr, gw1, gb1, gw2, gb2, gw3, gb3);
  return 0;
}

// One Adam step on a batch (device fp32 inputs/targets; w optional weights
// as in pmg_net_loss_grad_w, NULL for every point past the margin).
int pmg_net_train_step_w(pmg_net *N, const float *x, const float *t, const float *w, int nx, int ny, int nz,
                         int nb, int margin, float lr, double *loss) {
  if (!N || !x || !t) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure(N, P)) return -1;
  Geom g{nx, ny, nz, nb};
  hipLaunchKernelGGL(k_pack_input<float>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, N->stream, x, N->x0,
                     P, 1.0f);
  int rc = net_loss_backward(N, g, t, w, margin, loss);
  if (rc) return rc;
  N->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)N->adam_t), c2 = 1.f - std::pow(b2, (float)N->adam_t);
  const int nf = 3 * LAYER_FLOATS;
  hipLaunchKernelGGL(k_adam, dim3((nf + 255) / 256), dim3(256), 0, N->stream, N->w, N->m1, N->m2, N->grad,
                     N->mask, nf, lr, b1, b2, c1, c2, 1e-8f);
  PMG_HOST_CHECK(hipGetLastError());
  N->frags_valid = false;
  return 0;
}

// One Adam step on a batch (device fp32 inputs/targets).
int pmg_net_train_step(pmg_net *N, const float *x, const float *t, int nx, int ny, int nz, int nb,
                       int margin, float lr, double *loss) {
  if (!N || !x || !t) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure(N, P)) return -1;
  Geom g{nx, ny, nz, nb};
  hipLaunchKernelGGL(k_pack_input<float>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, N->stream, x, N->x0,
                     P, 1.0f);
  int rc = net_loss_backward(N, g, t, nullptr, margin, loss);
  if (rc) return rc;
  N->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)N->adam_t), c2 = 1.f - std::pow(b2, (float)N->adam_t);
  const int nf = 3 * LAYER_FLOATS;
  hipLaunchKernelGGL(k_adam, dim3((nf + 255) / 256), dim3(256), 0, N->stream, N->w, N->m1, N->m2, N->grad,
                     N->mask, nf, lr, b1, b2, c1, c2, 1e-8f);
  PMG_HOST_CHECK(hipGetLastError());
  N->frags_valid = false;
  return 0;
}

void *pmg_net_stream(pmg_net *N) { return N ? (void *)N->stream : nullptr; }

// Learned correction on one grid, on the caller's stream: y = N(scale * x)
// (x device FP32, e.g. the negated defect the mixed solver keeps), then
// u += coef * y at interior points.  The output stays in the network's
// buffer so pmg_net_recorrect can undo or repeat the step.
int pmg_net_correct(pmg_net *N, const float *x, float scale, double *u, double coef, int nx,
                    int ny, int nz, void *stream) {
  if (!N || !x || !u) return -1;
  hipStream_t st = (hipStream_t)stream;
  const size_t P = (size_t)nx * ny * nz;
  if (net_ensure_infer(N, P)) return -1;
  hipStream_t keep = N->stream;
  N->stream = st;
  int rc = net_infer(N, x, scale, Geom{nx, ny, nz, 1});
  N->stream = keep;
  if (rc) return rc;
  hipLaunchKernelGGL((k_apply_correction<double, 1>), dim3((nx + 63) / 64, (ny + 3) / 4, nz), dim3(256), 0, st, u,
                     N->y, nx, ny, nz, coef);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_net_recorrect(pmg_net *N, double *u, double coef, int nx, int ny, int nz, void *stream) {
  if (!N || !u || (size_t)nx * ny * nz > N->icap) return -1;
  hipLaunchKernelGGL((k_apply_correction<double, 1>), dim3((nx + 63) / 64, (ny + 3) / 4, nz), dim3(256), 0,
                     (hipStream_t)stream, u, N->y, nx, ny, nz, coef);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_net_reserve(pmg_net *N, long long points) {
  if (!N || points < 1) return -1;
  if (net_ensure_infer(N, (size_t)points)) return -1;
  if (net_frags(N)) return -1;
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  return 0;
}

// y = N(scale * x) on nb boxes, enqueued on `stream` with no host
// synchronisation (capturable in a graph once pmg_net_reserve has sized the
// buffers and packed the weights: then nothing is allocated or packed
// here).  Returns the network-owned output buffer (fp32 [nb][P], valid until
// the next inference), or NULL on error.
float *pmg_net_infer_async(pmg_net *N, const float *x, float scale, int nx, int ny, int nz, int nb, void *stream) {
  if (!N || !x || nx < 1 || ny < 1 || nz < 1 || nb < 1) return nullptr;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure_infer(N, P)) return nullptr;
  hipStream_t keep = N->stream;
  N->stream = (hipStream_t)stream;
  const int rc = net_infer(N, x, scale, Geom{nx, ny, nz, nb});
  N->stream = keep;
  return rc ? nullptr : N->y;
}

int pmg_net_infer(pmg_net *N, const float *x, float scale, float *y, int nx, int ny, int nz, int nb) {
  if (!N || !x || !y || nx < 1 || ny < 1 || nz < 1 || nb < 1) return -1;
  const size_t P = (size_t)nx * ny * nz * nb;
  if (net_ensure_infer(N, P)) return -1;
  int rc = net_infer(N, x, scale, Geom{nx, ny, nz, nb});
  if (rc) return rc;
  PMG_HOST_CHECK(hipMemcpyAsync(y, N->y, P * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(N->stream));
  return 0;
}

// Crops for training: out[c][z][y][x] = scale * (a[p] - b[p]) (b may be
// NULL) for cubes of side `side` at host-given origins (ox, oy, oz).
__global__ void k_gather_crops(const double *__restrict__ a, const double *__restrict__ b, int nx, int ny,
                               const int *__restrict__ org, int side, double scale, float *__restrict__ out) {
  const int c = blockIdx.y;
  const size_t n = (size_t)side * side * side;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const int x = (int)(q % side), y = (int)((q / side) % side), z = (int)(q / ((size_t)side * side));
  const size_t p = ix3(org[3 * c] + x, org[3 * c + 1] + y, org[3 * c + 2] + z, nx, ny);
  const double v = b ? a[p] - b[p] : a[p];
  out[(size_t)c * n + q] = (float)(scale * v);
}

int pmg_gather_crops(const double *a, const double *b, int nx, int ny, int nz, const int *origins_host,
                     int ncrop, int side, double scale, float *out, void *stream) {
  if (!a || !out || ncrop < 1 || side < 1) return -1;
  for (int c = 0; c < ncrop; c++)
    for (int d = 0; d < 3; d++) {
      const int o = origins_host[3 * c + d], n = d == 0 ? nx : (d == 1 ? ny : nz);
      if (o < 0 || o + side > n) return -1;
    }
  int *dorg;
  PMG_HOST_CHECK(hipMalloc(&dorg, sizeof(int) * 3 * ncrop));
  PMG_HOST_CHECK(hipMemcpyAsync(dorg, origins_host, sizeof(int) * 3 * ncrop, hipMemcpyHostToDevice,
                                (hipStream_t)stream));
  const size_t n = (size_t)side * side * side;
  hipLaunchKernelGGL(k_gather_crops, dim3((unsigned)((n + 255) / 256), ncrop), dim3(256), 0, (hipStream_t)stream,
                     a, b, nx, ny, dorg, side, scale, out);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipStreamSynchronize((hipStream_t)stream));
  PMG_HOST_CHECK(hipFree(dorg));
  return 0;
}

}  // extern "C"
