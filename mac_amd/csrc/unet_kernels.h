// unet_kernels.h -- kernels of the 3-D U-Net that proposes the initial
// pressure guess.  Activations are BF16, channels-last ([z][y][x][c]),
// accumulation FP32; master weights FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct Vol {
  int D, H, W;
};

__device__ inline float bf2f(bf16 v) { return (float)v; }
__device__ inline bf16 f2bf(float v) { return (bf16)v; }

// 3x3x3 convolution, zero padding 1, stride 1, as an implicit GEMM on the
// BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate):
//   M = output voxels (one wave: 16 consecutive x in each of ROWS consecutive
//   y rows), N = Cout (NB blocks of 16), K = 27 * Cin ordered (tap, cin),
//   tap = 9 (dz+1) + 3 (dy+1) + (dx+1).
// Lane l holds A row l % 16 and the 8 consecutive K of group l / 16 (one tap,
// 8 channels: one 16-byte load of the channels-last input), and B column
// l % 16 with the same 8 K (one 16-byte load of the [Cout][Kpad] weights).
// Each B fragment feeds ROWS matrix-core operations and each A fragment NB.
// Epilogue: + bias, optional ReLU, optional accumulation into the output.
// Cin = 1 << lcin with Cin >= 8; channel strides and offsets multiples of 8;
// W a multiple of 16; H a multiple of ROWS.
template <int NB, int ROWS>
__global__ void __launch_bounds__(256) k_conv3(const bf16* __restrict__ in, int in_cs, int in_co, int lcin, Vol v,
                                               const bf16* __restrict__ wt, int Kpad, const float* __restrict__ bias,
                                               bf16* __restrict__ out, int out_cs, int out_co, int relu, int accum) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nxw = v.W >> 4, nyw = v.H / ROWS;
  if (wave >= v.D * nyw * nxw) return;
  const int xw = wave % nxw, y0 = ((wave / nxw) % nyw) * ROWS, z = wave / (nxw * nyw);
  const int x0 = xw << 4;
  const int i = lane & 15, g = lane >> 4;
  const int Cin = 1 << lcin;
  floatx4 acc[ROWS][NB];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[r][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int K = 27 * Cin;
  const bf16* wrow = wt + (size_t)i * Kpad;
  for (int kc = 0; kc < Kpad; kc += 32) {
    const int k = kc + 8 * g;
    bf16x8 bfr[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bfr[nb] = *(const bf16x8*)(wrow + (size_t)16 * nb * Kpad + k);
    const int tap = k >> lcin, ci = k & (Cin - 1);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int xx = x0 + i + dx, zz = z + dz;
    const bool okxz = (k < K) && xx >= 0 && xx < v.W && zz >= 0 && zz < v.D;
    const bf16* base = in + ((size_t)zz * v.H * v.W + xx) * in_cs + in_co + ci;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int yy = y0 + r + dy;
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = (bf16)0.0f;
      if (okxz && yy >= 0 && yy < v.H) a = *(const bf16x8*)(base + (size_t)yy * v.W * in_cs);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bfr[nb], acc[r][nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = 16 * nb + i;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 4 * g + j;
        bf16* o = out + ((size_t)(z * v.H + y0 + r) * v.W + x0 + row) * out_cs + out_co + co;
        float val = acc[r][nb][j] + bv;
        if (relu) val = val > 0.f ? val : 0.f;
        if (accum) val += bf2f(*o);
        *o = f2bf(val);
      }
    }
}

// The same convolution with the input tile staged in LDS: a block of four
// waves computes a 16 x 16 (x, y) tile of one z plane (wave w: rows 4w..4w+3),
// and first loads the 18 x 18 x 3 input halo of its CIN channels once
// (zero outside the volume), so each of the 27 taps reads the LDS copy
// instead of re-reading global memory.  CIN = 1 << lcin in {8, 16, 32};
// H and W multiples of 16.  Same arithmetic and K order as k_conv3.
template <int NB, int CIN>
__global__ void __launch_bounds__(256) k_conv3_lds(const bf16* __restrict__ in, int in_cs, int in_co, int lcin, Vol v,
                                                   const bf16* __restrict__ wt, int Kpad, const float* __restrict__ bias,
                                                   bf16* __restrict__ out, int out_cs, int out_co, int relu, int accum) {
  constexpr int TX = 18, TY = 18, TZ = 3, C8 = CIN / 8;
  __shared__ bf16x8 tile[TZ * TY * TX * C8];
  const int nxt = v.W >> 4, nyt = v.H >> 4;
  const int b = blockIdx.x;
  const int x0 = (b % nxt) << 4, Y0 = ((b / nxt) % nyt) << 4, z = b / (nxt * nyt);
  for (int e = threadIdx.x; e < TZ * TY * TX * C8; e += blockDim.x) {
    const int c8 = e % C8, sidx = e / C8;
    const int lx = sidx % TX, ly = (sidx / TX) % TY, lz = sidx / (TX * TY);
    const int gx = x0 + lx - 1, gy = Y0 + ly - 1, gz = z + lz - 1;
    bf16x8 val;
#pragma unroll
    for (int j = 0; j < 8; ++j) val[j] = (bf16)0.0f;
    if (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H && gz >= 0 && gz < v.D)
      val = *(const bf16x8*)(in + ((size_t)(gz * v.H + gy) * v.W + gx) * in_cs + in_co + 8 * c8);
    tile[e] = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  floatx4 acc[4][NB];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[r][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int K = 27 * CIN;
  const bf16* wrow = wt + (size_t)i * Kpad;
  // weight fragments one chunk ahead, so their load overlaps the chunk's work
  bf16x8 bnext[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) bnext[nb] = *(const bf16x8*)(wrow + (size_t)16 * nb * Kpad + 8 * g);
  for (int kc = 0; kc < Kpad; kc += 32) {
    const int k = kc + 8 * g;
    bf16x8 bfr[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bfr[nb] = bnext[nb];
    if (kc + 32 < Kpad) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) bnext[nb] = *(const bf16x8*)(wrow + (size_t)16 * nb * Kpad + k + 32);
    }
    const int tap = k >> lcin, c8 = (k & (CIN - 1)) >> 3;
    const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // offsets + 1 into the tile
    const bool ok = k < K;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = (bf16)0.0f;
      if (ok) a = tile[((dz * TY + 4 * w + r + dy) * TX + i + dx) * C8 + c8];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bfr[nb], acc[r][nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = 16 * nb + i;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 4 * g + j;
        bf16* o = out + ((size_t)(z * v.H + Y0 + 4 * w + r) * v.W + x0 + row) * out_cs + out_co + co;
        float val = acc[r][nb][j] + bv;
        if (relu) val = val > 0.f ? val : 0.f;
        if (accum) val += bf2f(*o);
        *o = f2bf(val);
      }
    }
}

// Weight gradient of a 3x3x3 convolution, FP32:
//   dW[co][tap*Cin + ci] += sum_v G[v][co] * X[v + off(tap)][ci]
// Block = (tap, chunk of x-rows); thread = one 4x4 (co, ci) tile and a
// residue class of x; rows are staged in LDS.  The R threads sharing a tile
// are reduced with a fixed butterfly, so the result is reproducible; each
// block writes its partial, summed in chunk order by k_wgrad_reduce.
__global__ void __launch_bounds__(256) k_conv3_wgrad(const bf16* __restrict__ G, int g_cs, int g_co, int Cout,
                                                     const bf16* __restrict__ X, int x_cs, int x_co, int Cin, Vol v,
                                                     int rows_per_chunk, float* __restrict__ partial) {
  __shared__ float gs[64 * 64];  // [x][co]
  __shared__ float xs[64 * 64];  // [x][ci]
  const int tap = blockIdx.x % 27, chunk = blockIdx.x / 27;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int ntc = Cin >> 2, ntiles = (Cout >> 2) * ntc;
  const int R = 256 / ntiles;  // threads per tile (power of 2, <= 32)
  const int tile = threadIdx.x / R, r = threadIdx.x % R;
  const int co0 = (tile / ntc) * 4, ci0 = (tile % ntc) * 4;
  float acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
  const int nrows = v.D * v.H;
  const int r0 = chunk * rows_per_chunk, r1 = min(nrows, r0 + rows_per_chunk);
  for (int row = r0; row < r1; ++row) {
    const int y = row % v.H, z = row / v.H;
    const int yy = y + dy, zz = z + dz;
    __syncthreads();
    for (int e = threadIdx.x; e < v.W * Cout; e += blockDim.x) {
      const int x = e / Cout, c = e % Cout;
      gs[x * Cout + c] = bf2f(G[((size_t)(z * v.H + y) * v.W + x) * g_cs + g_co + c]);
    }
    const bool rowok = (yy >= 0 && yy < v.H && zz >= 0 && zz < v.D);
    for (int e = threadIdx.x; e < v.W * Cin; e += blockDim.x) {
      const int x = e / Cin, c = e % Cin;
      const int xx = x + dx;
      float val = 0.f;
      if (rowok && xx >= 0 && xx < v.W) val = bf2f(X[((size_t)(zz * v.H + yy) * v.W + xx) * x_cs + x_co + c]);
      xs[x * Cin + c] = val;
    }
    __syncthreads();
    if (tile < ntiles) {
      for (int x = r; x < v.W; x += R) {
        float gv[4], xv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) gv[a] = gs[x * Cout + co0 + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) xv[b] = xs[x * Cin + ci0 + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] += gv[a] * xv[b];
      }
    }
  }
  // butterfly over the R lanes of a tile (they are consecutive lanes of one wave)
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      for (int o = 1; o < R; o <<= 1) acc[a][b] += __shfl_xor(acc[a][b], o);
  if (tile < ntiles && r == 0) {
    float* P = partial + ((size_t)chunk * 27 + tap) * Cout * Cin;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) P[(co0 + a) * Cin + ci0 + b] = acc[a][b];
  }
}

// Weight gradient of a 3x3x3 convolution on the BF16 matrix cores:
//   dW_tap[co][ci] = sum_v G[v][co] X[v + off(tap)][ci]
// one GEMM per tap with M = Cout, N = Cin, K = voxels, 32 x-voxels of a row
// per v_mfma_f32_16x16x32_bf16.  Block = (tap, chunk of (z, y) rows).  Per
// row the G row and the tap's shifted X row are staged in LDS channel-major
// ([c][x], row stride S = Wp + 8 so that the 16 rows of a fragment start in
// different banks; zero beyond the volume and in padded channels), so an
// operand fragment -- one channel, 8 consecutive voxels -- is one 16-byte LDS
// read: lane l supplies A row co0 + l % 16 and B column ci0 + l % 16 for the
// voxels kc + 8 (l / 16) .. + 7, and holds C[co0 + 4 (l / 16) + j][ci0 + l % 16].
// Wave w takes the 16 x 16 (co, ci) tiles w, w + 4, ...  Products of BF16
// values are exact in FP32 and the sums run in a fixed order, so the result
// is reproducible.  Each block writes its partial, summed in chunk order by
// k_wgrad_reduce.  W <= WG_MAXW, Cout a multiple of 16, Cin a multiple of 8
// and at most 64.
constexpr int WG_MAXW = 128;
constexpr int WG_S = WG_MAXW + 8;
__global__ void __launch_bounds__(256) k_conv3_wgrad_mfma(const bf16* __restrict__ G, int g_cs, int g_co, int Cout,
                                                          const bf16* __restrict__ X, int x_cs, int x_co, int Cin, Vol v,
                                                          int rows_per_chunk, float* __restrict__ partial) {
  __shared__ __align__(16) bf16 lds[128 * WG_S];  // [Cout][S] then [CinP][S]
  const int tap = blockIdx.x % 27, chunk = blockIdx.x / 27;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int Wp = (v.W + 31) & ~31, S = Wp + 8;
  const int nci = (Cin + 15) >> 4, CinP = nci << 4, nco = Cout >> 4;
  const int ntiles = nco * nci;
  bf16* gs = lds;
  bf16* xs = lds + Cout * S;
  for (int e = threadIdx.x; e < (Cout + CinP) * S; e += blockDim.x) lds[e] = (bf16)0.0f;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  floatx4 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nrows = v.D * v.H;
  const int r0 = chunk * rows_per_chunk, r1 = min(nrows, r0 + rows_per_chunk);
  const int c8o = Cout >> 3, c8i = Cin >> 3;
  for (int row = r0; row < r1; ++row) {
    const int y = row % v.H, z = row / v.H;
    const int yy = y + dy, zz = z + dz;
    const bool rowok = (yy >= 0 && yy < v.H && zz >= 0 && zz < v.D);
    __syncthreads();
    for (int e = threadIdx.x; e < v.W * c8o; e += blockDim.x) {
      const int x = e / c8o, c8 = e % c8o;
      const bf16x8 val = *(const bf16x8*)(G + ((size_t)(z * v.H + y) * v.W + x) * g_cs + g_co + 8 * c8);
#pragma unroll
      for (int j = 0; j < 8; ++j) gs[(8 * c8 + j) * S + x] = val[j];
    }
    for (int e = threadIdx.x; e < v.W * c8i; e += blockDim.x) {
      const int x = e / c8i, c8 = e % c8i;
      const int xx = x + dx;
      bf16x8 val;
#pragma unroll
      for (int j = 0; j < 8; ++j) val[j] = (bf16)0.0f;
      if (rowok && xx >= 0 && xx < v.W) val = *(const bf16x8*)(X + ((size_t)(zz * v.H + yy) * v.W + xx) * x_cs + x_co + 8 * c8);
#pragma unroll
      for (int j = 0; j < 8; ++j) xs[(8 * c8 + j) * S + x] = val[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = w + 4 * k;
      if (t >= ntiles) break;
      const int co0 = (t / nci) << 4, ci0 = (t % nci) << 4;
      const bf16* ap = gs + (co0 + i) * S + 8 * g;
      const bf16* bp = xs + (ci0 + i) * S + 8 * g;
      for (int kc = 0; kc < Wp; kc += 32)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*(const bf16x8*)(ap + kc), *(const bf16x8*)(bp + kc), acc[k], 0, 0, 0);
    }
  }
  float* P = partial + ((size_t)chunk * 27 + tap) * Cout * Cin;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = w + 4 * k;
    if (t >= ntiles) break;
    const int co0 = (t / nci) << 4, ci = ((t % nci) << 4) + i;
    if (ci < Cin) {
#pragma unroll
      for (int j = 0; j < 4; ++j) P[(size_t)(co0 + 4 * g + j) * Cin + ci] = acc[k][j];
    }
  }
}

// Weight gradient of the 2x2x2 transposed convolution on the matrix cores:
//   dW[o][co][ci] = sum over coarse voxels v of G[2v + o][co] X[v][ci]
// one GEMM per octant o (M = Cout, N = Cin, K = coarse voxels, 32 x-voxels of
// a coarse row per v_mfma_f32_16x16x32_bf16), staged and tiled as in
// k_conv3_wgrad_mfma.  Block = (octant, chunk of coarse (z, y) rows); X is
// channels-last with Cin channels.  Partials [chunk][o][co][ci], summed in
// chunk order by k_up_wgrad_reduce.  vc.W <= WG_MAXW; Cout and Cin multiples
// of 16 with Cout + Cin <= 128.
__global__ void __launch_bounds__(256) k_up_wgrad_mfma(const bf16* __restrict__ G, int g_cs, int g_co, int Cout, Vol vc,
                                                       const bf16* __restrict__ X, int Cin, int rows_per_chunk,
                                                       float* __restrict__ partial) {
  __shared__ __align__(16) bf16 lds[128 * WG_S];
  const int o = blockIdx.x & 7, chunk = blockIdx.x >> 3;
  const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
  const int Wp = (vc.W + 31) & ~31, S = Wp + 8;
  const int nci = Cin >> 4, nco = Cout >> 4, ntiles = nco * nci;
  const int Hf = 2 * vc.H, Wf = 2 * vc.W;
  bf16* gs = lds;
  bf16* xs = lds + Cout * S;
  for (int e = threadIdx.x; e < (Cout + Cin) * S; e += blockDim.x) lds[e] = (bf16)0.0f;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  floatx4 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nrows = vc.D * vc.H;
  const int r0 = chunk * rows_per_chunk, r1 = min(nrows, r0 + rows_per_chunk);
  const int c8o = Cout >> 3, c8i = Cin >> 3;
  for (int row = r0; row < r1; ++row) {
    const int y = row % vc.H, z = row / vc.H;
    __syncthreads();
    for (int e = threadIdx.x; e < vc.W * c8o; e += blockDim.x) {
      const int x = e / c8o, c8 = e % c8o;
      const bf16x8 val =
          *(const bf16x8*)(G + ((size_t)((2 * z + oz) * Hf + 2 * y + oy) * Wf + 2 * x + ox) * g_cs + g_co + 8 * c8);
#pragma unroll
      for (int j = 0; j < 8; ++j) gs[(8 * c8 + j) * S + x] = val[j];
    }
    for (int e = threadIdx.x; e < vc.W * c8i; e += blockDim.x) {
      const int x = e / c8i, c8 = e % c8i;
      const bf16x8 val = *(const bf16x8*)(X + ((size_t)(z * vc.H + y) * vc.W + x) * Cin + 8 * c8);
#pragma unroll
      for (int j = 0; j < 8; ++j) xs[(8 * c8 + j) * S + x] = val[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = w + 4 * k;
      if (t >= ntiles) break;
      const int co0 = (t / nci) << 4, ci0 = (t % nci) << 4;
      const bf16* ap = gs + (co0 + i) * S + 8 * g;
      const bf16* bp = xs + (ci0 + i) * S + 8 * g;
      for (int kc = 0; kc < Wp; kc += 32)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*(const bf16x8*)(ap + kc), *(const bf16x8*)(bp + kc), acc[k], 0, 0, 0);
    }
  }
  float* P = partial + ((size_t)chunk * 8 + o) * Cout * Cin;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = w + 4 * k;
    if (t >= ntiles) break;
    const int co0 = (t / nci) << 4, ci = ((t % nci) << 4) + i;
#pragma unroll
    for (int j = 0; j < 4; ++j) P[(size_t)(co0 + 4 * g + j) * Cin + ci] = acc[k][j];
  }
}

// dW[o][co][ci] = sum over chunks (in order) of partial[chunk][o][co][ci].
__global__ void k_up_wgrad_reduce(int nchunk, int Cout, int Cin, const float* __restrict__ partial, float* __restrict__ dw) {
  const int n = 8 * Cout * Cin;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += partial[(size_t)c * n + e];
    dw[e] = s;
  }
}

// dW[co][tap*Cin+ci] = sum over chunks (in order) of partial[chunk][tap][co][ci].
__global__ void k_wgrad_reduce(int nchunk, int Cout, int Cin, const float* __restrict__ partial, float* __restrict__ dw) {
  const int n = 27 * Cout * Cin;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int tap = e / (Cout * Cin), co = (e / Cin) % Cout, ci = e % Cin;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += partial[(size_t)c * n + e];
    dw[(size_t)co * 27 * Cin + tap * Cin + ci] = s;
  }
}

// Bias gradient: db[c] = sum_v G[v][c]; one block per channel, fixed order.
__global__ void k_bias_grad(const bf16* __restrict__ G, int cs, int co, size_t nvox, float* __restrict__ db) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (size_t v = threadIdx.x; v < nvox; v += blockDim.x) s += bf2f(G[v * cs + co + c]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = sh[0];
}

// Forward/backward weight layouts for the matrix-core convolution:
//   fwd[co][tap*Cin + ci]            (K padded to Kpad with zeros)
//   bwd[ci][tap'*Cout + co] = w[co][tap*Cin+ci], tap' = 26 - tap (mirror),
// so that the data gradient is the same convolution kernel on the bwd weights.
__global__ void k_pack_weights(int Cout, int Cin, const float* __restrict__ w, bf16* __restrict__ fwd, int Kpad,
                               bf16* __restrict__ bwd, int Kpad2) {
  const int nf = Cout * Kpad, nbw = Cin * Kpad2;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nf + nbw; e += gridDim.x * blockDim.x) {
    if (e < nf) {
      const int co = e / Kpad, k = e % Kpad;
      fwd[e] = f2bf(k < 27 * Cin ? w[(size_t)co * 27 * Cin + k] : 0.f);
    } else if (bwd) {
      const int e2 = e - nf;
      const int ci = e2 / Kpad2, k = e2 % Kpad2;
      float val = 0.f;
      if (k < 27 * Cout) {
        const int tp = k / Cout, co = k % Cout, tap = 26 - tp;
        val = w[(size_t)co * 27 * Cin + tap * Cin + ci];
      }
      bwd[e2] = f2bf(val);
    }
  }
}

// G = dY * (Y > 0), in place on dY (ReLU backward).
__global__ void k_relu_mask(size_t nvox, int C, bf16* __restrict__ dy, int dy_cs, int dy_co, const bf16* __restrict__ y, int y_cs,
                            int y_co) {
  const size_t n = nvox * C;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const size_t v = e / C;
    const int c = (int)(e % C);
    if (!(bf2f(y[v * y_cs + y_co + c]) > 0.f)) dy[v * dy_cs + dy_co + c] = f2bf(0.f);
  }
}

// 2x2x2 average pooling (fine volume v -> coarse v/2).
__global__ void k_pool_fwd(const bf16* __restrict__ in, int in_cs, int in_co, int C, Vol vf, bf16* __restrict__ out) {
  const int Dc = vf.D / 2, Hc = vf.H / 2, Wc = vf.W / 2;
  const size_t n = (size_t)Dc * Hc * Wc * C;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const size_t vc = e / C;
    const int x = (int)(vc % Wc), y = (int)((vc / Wc) % Hc), z = (int)(vc / ((size_t)Wc * Hc));
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
      const int xx = 2 * x + (o & 1), yy = 2 * y + ((o >> 1) & 1), zz = 2 * z + (o >> 2);
      s += bf2f(in[((size_t)(zz * vf.H + yy) * vf.W + xx) * in_cs + in_co + c]);
    }
    out[e] = f2bf(s * 0.125f);
  }
}

// Pool backward, accumulated into the fine gradient: dIn[child] += dOut[parent]/8.
__global__ void k_pool_bwd(const bf16* __restrict__ dout, int C, Vol vf, bf16* __restrict__ din, int din_cs, int din_co) {
  const size_t n = (size_t)vf.D * vf.H * vf.W * C;
  const int Hc = vf.H / 2, Wc = vf.W / 2;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const size_t vv = e / C;
    const int x = (int)(vv % vf.W), y = (int)((vv / vf.W) % vf.H), z = (int)(vv / ((size_t)vf.W * vf.H));
    const float g = bf2f(dout[((size_t)((z / 2) * Hc + y / 2) * Wc + x / 2) * C + c]) * 0.125f;
    bf16* p = din + vv * din_cs + din_co + c;
    *p = f2bf(bf2f(*p) + g);
  }
}

// Transposed 2x2x2 convolution, stride 2 (coarse vc -> fine vf), + bias, ReLU:
//   out[2v + o][co] = relu(b[co] + sum_ci W[o][co][ci] in[v][ci]).
__global__ void k_up_fwd(const bf16* __restrict__ in, int Cin, Vol vc, const float* __restrict__ w, const float* __restrict__ b,
                         int Cout, bf16* __restrict__ out, int out_cs, int out_co) {
  const int Df = 2 * vc.D, Hf = 2 * vc.H, Wf = 2 * vc.W;
  const size_t n = (size_t)Df * Hf * Wf * Cout;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int co = (int)(e % Cout);
    const size_t vv = e / Cout;
    const int x = (int)(vv % Wf), y = (int)((vv / Wf) % Hf), z = (int)(vv / ((size_t)Wf * Hf));
    const int o = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
    const bf16* ip = in + ((size_t)((z / 2) * vc.H + y / 2) * vc.W + x / 2) * Cin;
    const float* wp = w + ((size_t)o * Cout + co) * Cin;
    float s = b[co];
    for (int ci = 0; ci < Cin; ++ci) s += wp[ci] * bf2f(ip[ci]);
    out[vv * out_cs + out_co + co] = f2bf(s > 0.f ? s : 0.f);
  }
}

// FP32 -> BF16 copy (the transposed-convolution weights for the matrix cores).
__global__ void k_to_bf16(size_t n, const float* __restrict__ w, bf16* __restrict__ out) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) out[e] = f2bf(w[e]);
}

// Transposed convolution forward on the matrix cores: for each octant o,
// out[2v+o][co] = relu(b[co] + sum_ci W[o][co][ci] in[v][ci]), a GEMM of the
// coarse voxels (rows) by Cin (K) by Cout (columns).  One wave: 16 voxels of
// a coarse x row, all 8 octants, Cout = 16 NB; Cin a multiple of 32.  Weights
// wb [8][Cout][Cin] in BF16.  MFMA 16x16x32 operand layout: lane l holds A
// row l%16 and B column l%16 at k = 8(l/16) + j; D rows 4(l/16) + j.
template <int NB>
__global__ void __launch_bounds__(256) k_up_fwd_mfma(const bf16* __restrict__ in, int Cin, Vol vc, const bf16* __restrict__ wb,
                                                     const float* __restrict__ b, bf16* __restrict__ out, int out_cs, int out_co) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nxw = vc.W >> 4;
  if (wave >= vc.D * vc.H * nxw) return;
  const int xw = wave % nxw, y = (wave / nxw) % vc.H, z = wave / (nxw * vc.H);
  const int i = lane & 15, g = lane >> 4;
  const int Cout = 16 * NB;
  const size_t vrow = (size_t)(z * vc.H + y) * vc.W + (xw << 4);
  const int Hf = 2 * vc.H, Wf = 2 * vc.W;
  for (int o = 0; o < 8; ++o) {
    floatx4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Cin; kc += 32) {
      const int k = kc + 8 * g;
      const bf16x8 a = *(const bf16x8*)(in + (vrow + i) * Cin + k);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const bf16x8 bb = *(const bf16x8*)(wb + ((size_t)o * Cout + 16 * nb + i) * Cin + k);
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[nb], 0, 0, 0);
      }
    }
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = 16 * nb + i;
      const float bv = b[co];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int xc = (xw << 4) + 4 * g + j;
        const size_t fv = ((size_t)(2 * z + oz) * Hf + 2 * y + oy) * Wf + 2 * xc + ox;
        const float val = acc[nb][j] + bv;
        out[fv * out_cs + out_co + co] = f2bf(val > 0.f ? val : 0.f);
      }
    }
  }
}

// Transposed conv backward-data: dIn[v][ci] = sum_{o,co} W[o][co][ci] G[2v+o][co].
__global__ void k_up_bwd_data(const bf16* __restrict__ g, int g_cs, int g_co, int Cout, Vol vc, const float* __restrict__ w,
                              int Cin, bf16* __restrict__ din) {
  const size_t n = (size_t)vc.D * vc.H * vc.W * Cin;
  const int Hf = 2 * vc.H, Wf = 2 * vc.W;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int ci = (int)(e % Cin);
    const size_t v = e / Cin;
    const int x = (int)(v % vc.W), y = (int)((v / vc.W) % vc.H), z = (int)(v / ((size_t)vc.W * vc.H));
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
      const int xx = 2 * x + (o & 1), yy = 2 * y + ((o >> 1) & 1), zz = 2 * z + (o >> 2);
      const bf16* gp = g + ((size_t)(zz * Hf + yy) * Wf + xx) * g_cs + g_co;
      for (int co = 0; co < Cout; ++co) s += w[((size_t)o * Cout + co) * Cin + ci] * bf2f(gp[co]);
    }
    din[e] = f2bf(s);
  }
}

// Transposed conv weight/bias gradient: one thread per (o, co, ci) (+ bias),
// summing over coarse voxels in order.
__global__ void k_up_wgrad(const bf16* __restrict__ g, int g_cs, int g_co, int Cout, Vol vc, const bf16* __restrict__ in, int Cin,
                           float* __restrict__ dw, float* __restrict__ db) {
  const int n = 8 * Cout * Cin + Cout;
  const int Hf = 2 * vc.H, Wf = 2 * vc.W;
  const size_t nv = (size_t)vc.D * vc.H * vc.W;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    float s = 0.f;
    if (e < 8 * Cout * Cin) {
      const int o = e / (Cout * Cin), co = (e / Cin) % Cout, ci = e % Cin;
      for (size_t v = 0; v < nv; ++v) {
        const int x = (int)(v % vc.W), y = (int)((v / vc.W) % vc.H), z = (int)(v / ((size_t)vc.W * vc.H));
        const int xx = 2 * x + (o & 1), yy = 2 * y + ((o >> 1) & 1), zz = 2 * z + (o >> 2);
        s += bf2f(g[((size_t)(zz * Hf + yy) * Wf + xx) * g_cs + g_co + co]) * bf2f(in[v * Cin + ci]);
      }
      dw[e] = s;
    } else {
      const int co = e - 8 * Cout * Cin;
      for (size_t v = 0; v < nv * 8; ++v) s += bf2f(g[v * g_cs + g_co + co]);
      db[co] = s;
    }
  }
}

// Head: 1x1x1 convolution C -> 1, FP32 output.
__global__ void k_head_fwd(const bf16* __restrict__ in, int C, size_t nvox, const float* __restrict__ w, const float* __restrict__ b,
                           float* __restrict__ out) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    float s = b[0];
    for (int c = 0; c < C; ++c) s += w[c] * bf2f(in[v * C + c]);
    out[v] = s;
  }
}

__global__ void k_head_bwd_data(const float* __restrict__ gout, int C, size_t nvox, const float* __restrict__ w,
                                bf16* __restrict__ din) {
  const size_t n = nvox * C;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    din[e] = f2bf(w[e % C] * gout[e / C]);
}

// Head weight/bias gradient with per-block partials (reduced in order on the host side of the launch).
__global__ void k_head_wgrad(const float* __restrict__ gout, const bf16* __restrict__ in, int C, size_t nvox,
                             float* __restrict__ partial) {
  __shared__ float sh[256];
  for (int c = 0; c <= C; ++c) {
    float s = 0.f;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x)
      s += gout[v] * (c < C ? bf2f(in[v * C + c]) : 1.f);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * (C + 1) + c] = sh[0];
    __syncthreads();
  }
}

// Masked MSE: loss partials and d loss / d out = 2 m (p - t) / sum(m).
__global__ void k_mse(size_t n, const float* __restrict__ p, const float* __restrict__ t, const float* __restrict__ m, float inv_msum,
                      float* __restrict__ gout, float* __restrict__ partial) {
  __shared__ float sh[256];
  float s = 0.f;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) {
    const float d = (p[v] - t[v]) * m[v];
    s += d * d;
    gout[v] = 2.f * d * inv_msum;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Adam (bias-corrected step sizes passed in).
__global__ void k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v2,
                       float lr_t, float b1, float b2, float eps) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v2[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v2[e] = vv;
    w[e] -= lr_t * mm / (sqrtf(vv) + eps);
  }
}

// Host FP32 input [V][5] -> BF16 [V][8] (channels 5..7 zero).
__global__ void k_pack_input(size_t nvox, const float* __restrict__ in5, bf16* __restrict__ out8) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvox * 8; e += (size_t)gridDim.x * blockDim.x) {
    const size_t v = e >> 3;
    const int c = (int)(e & 7);
    out8[e] = f2bf(c < 5 ? in5[v * 5 + c] : 0.f);
  }
}

// ------------------------------------------------ brick grid <-> volume ----
// The network sees the grid's bounding box as a dense D x H x W volume (x
// fastest); a cell of brick b with chunk key (kx, ky, kz) (shifted so that the
// minimum of each axis is 0) and in-brick index t = x + 8y + 64z sits at voxel
// (8kx + x) + W (8ky + y + H (8kz + z)).  Scales: s = rms of the right-hand side
// over the unknowns; inputs are f / s and p ih2 / s, the output is p ih2 /
// (s out_scale).  A zero right-hand side gives a zero guess.

struct GuessGrid {
  int nb;
  const int* keys;        // 3 per brick
  const uint8_t* flags;   // bit 7 exists, bits 0-2 material (2 = Solid)
  const uint8_t* code;    // bit 6: unknown of the pressure system
  const double* ff;       // ||f||^2 over the unknowns
  double ndom, ih2m;
  int D, H, W;
};

__device__ inline double guess_inv_scale(const GuessGrid& g) {
  const double ff = *g.ff;
  return ff > 0.0 ? sqrt(g.ndom / ff) : 0.0;
}

__device__ inline size_t guess_voxel(const GuessGrid& g, int b, int t) {
  const int X = 8 * g.keys[3 * b] + (t & 7), Y = 8 * g.keys[3 * b + 1] + ((t >> 3) & 7), Z = 8 * g.keys[3 * b + 2] + (t >> 6);
  return (size_t)X + (size_t)g.W * ((size_t)Y + (size_t)g.H * Z);
}

// every voxel: no right-hand side, no pressure, solid (cells outside the grid)
__global__ void k_guess_fill(size_t nvox, float* __restrict__ in5) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nvox * 5; e += (size_t)gridDim.x * blockDim.x)
    in5[e] = (e % 5 == 3) ? 1.f : 0.f;
}

// in5 from the grid: f / s, fluid (unknown), air (existing, not solid, not an
// unknown: p = 0), solid (the rest), prev ih2 / s.
__global__ void k_guess_pack(GuessGrid g, const double* __restrict__ f, const double* __restrict__ prev, float* __restrict__ in5) {
  const double is = guess_inv_scale(g);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)g.nb * 512; e += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(e >> 9), t = (int)(e & 511);
    const size_t v = guess_voxel(g, b, t);
    const bool dom = (g.code[e] & 0x40) != 0;
    const uint8_t fl = g.flags[e];
    const bool air = !dom && (fl & 0x80) && (fl & 7) != 2;
    float* o = in5 + 5 * v;
    o[0] = dom ? (float)(f[e] * is) : 0.f;
    o[1] = dom ? 1.f : 0.f;
    o[2] = air ? 1.f : 0.f;
    o[3] = (dom || air) ? 0.f : 1.f;
    o[4] = (dom && prev) ? (float)(prev[e] * g.ih2m * is) : 0.f;
  }
}

// guess x from the network output (zero off the unknowns)
__global__ void k_guess_unpack(GuessGrid g, const float* __restrict__ out, float out_scale, double* __restrict__ x) {
  const double is = guess_inv_scale(g);
  const double sc = is > 0.0 ? (double)out_scale / (is * g.ih2m) : 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)g.nb * 512; e += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(e >> 9), t = (int)(e & 511);
    const bool dom = (g.code[e] & 0x40) != 0;
    x[e] = dom ? (double)out[guess_voxel(g, b, t)] * sc : 0.0;
  }
}

// training target p ih2 / (s out_scale) and mask on the unknowns (target and
// mask zero-filled beforehand)
__global__ void k_guess_target(GuessGrid g, const double* __restrict__ p, float out_scale, float* __restrict__ tgt,
                               float* __restrict__ mask) {
  const double is = guess_inv_scale(g);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)g.nb * 512; e += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(e >> 9), t = (int)(e & 511);
    if (!(g.code[e] & 0x40)) continue;
    const size_t v = guess_voxel(g, b, t);
    tgt[v] = (float)(p[e] * g.ih2m * is / out_scale);
    mask[v] = 1.f;
  }
}
