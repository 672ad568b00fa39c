// rf_lu.hip -- blocked right-looking LU of the assembled system and the
// solves of the direct solvers.
//
// Reference: magma_sgesv_gpu (src/simulation.cu:403-418) and LAPACK sgesv
// (fortran/solve_system.incl:3): LU with row pivoting, then L y = P b, U x = y.
//
// Factorisation of F, the FP32 system written into the factor storage (the
// residuals come from the system's own form), in outer panels of
// kLuPanel = W columns.  Panel [K0, KE):
//   1. its W x W diagonal block D is factored in place, 64-column blocks at a
//      time: diag (one wavefront, partial pivoting among the block's rows,
//      the row order kept as a permutation perm[k0 + c] = source row of row
//      c), l21 and u12 (the block's triangular solves for D's rows and
//      columns), dupdate (D's Schur complement, FP32);
//   2. the inverses of D's factors, Uinv = U_D^-1 and Tfwd = the forward
//      operator of D (the block permutations and L_D^-1 as the forward solve
//      applies them), by running the panel solves on the unit vectors;
//   3. the panel's rows below D, L21 = A21 Uinv, and its rows of U for the
//      trailing columns, U12 = Tfwd A12, each one MFMA GEMM with K = W from
//      packed copies of A21 / A12 (written to F).
// The panels are taken in groups of kLuGroup: within a group each panel's
// columns are updated by the group's earlier panels (left-looking), and the
// trailing matrix right of the group once by all of them, A22 -= L21 U12
// with K = kLuGroup W, on MFMA from packs in the mode's element type, each
// 32-row block of the L21 pack and 32-column block of the U12 pack scaled by
// a power of two that puts its largest magnitude in [2^14, 2^15) (the update
// divides the scales out exactly).  A one-group look-ahead overlaps the next
// group's panels (aux stream) with the trailing update of the columns right
// of that group.
//
// Pivoting is restricted to the 64-row diagonal blocks (the system's identity-
// dominated diagonal blocks make that pivot choice safe; the refinement below
// verifies every solve and the solver falls back to GMRES when it does not
// converge).  The columns left of a block are not permuted: the forward solve
// subtracts their contribution in the block's original row order and then
// applies the block's permutation, which is the same L y = P b.
//
// Every solve is followed by FP64 iterative refinement with residuals of the
// FP32 A (r = b - A x accumulated in FP64), so the returned x carries a
// verified ||b - A x|| / ||b||.
#include <cmath>
#include <vector>

#include "rf_lu.h"

namespace rf {

// ---------------------------------------------------------------------------
// diag: GEPP of the kb x kb block at (k0, k0) by one wavefront, lane = row.
struct DiagLane {
  float row[kLuNb];  // the lane's row of the block
  bool used;         // chosen as a pivot row already
  int pos;           // its row in the factored block
  int zero_col;      // first column with a zero pivot (1-based), 0 if none
};

// Maximum of an unsigned over the wavefront by DPP (xor 1, xor 2, half-row and
// row mirrors, then row broadcasts 15 and 31): lane 63 ends with the maximum.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  auto step = [](unsigned x, unsigned y) { return x > y ? x : y; };
  v = step(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xf, 0xf, false)));
  v = step(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xf, 0xf, false)));
  v = step(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xf, 0xf, false)));
  v = step(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xf, 0xf, false)));
  v = step(v, static_cast<unsigned>(
                  __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x142, 0xa, 0xf, false)));
  v = step(v, static_cast<unsigned>(
                  __builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x143, 0xc, 0xf, false)));
  return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Elimination step for column C (compile-time, so the row stays in registers).
// The pivot row is broadcast from the pivot lane's registers (v_readlane: the
// pivot lane is uniform over the wavefront); rows already chosen get the
// multiplier 0, so their update fma(-0, p, x) leaves them exactly as they are
// (finite p).
template <int C>
__device__ __forceinline__ void diag_step(DiagLane& d, int kb, int lane) {
  if constexpr (C < kLuNb) {
    if (C >= kb) return;
    // pivot: largest |row[C]| among the rows not yet chosen.  The key orders
    // by |value| (IEEE bits of a non-negative float order as integers, the
    // lowest 6 mantissa bits dropped) and then by the lower row; NaN ranks
    // highest; chosen rows have key 0, every candidate a key >= 64.
    const float a = d.row[C];
    const unsigned bits = __builtin_bit_cast(unsigned, fabsf(a)) >> 6;
    const unsigned key = d.used ? 0u : (((bits + 1u) << 6) | static_cast<unsigned>(63 - lane));
    const int pr = 63 - static_cast<int>(wave_max_u32(key) & 63u);
    if (lane == pr) {
      d.used = true;
      d.pos = C;
    }
    float p[kLuNb - C];
#pragma unroll
    for (int c = C; c < kLuNb; ++c)
      p[c - C] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, d.row[c]), pr));
    const float pv = p[0];
    if (pv == 0.0f && d.zero_col == 0) d.zero_col = C + 1;
    const float inv = (pv != 0.0f) ? 1.0f / pv : 0.0f;
    const float l = d.used ? 0.0f : d.row[C] * inv;
    if (!d.used) d.row[C] = l;
#pragma unroll
    for (int cc = C + 1; cc < kLuNb; ++cc) d.row[cc] = fmaf(-l, p[cc - C], d.row[cc]);
    diag_step<C + 1>(d, kb, lane);
  }
}

__global__ __launch_bounds__(64) void lu_diag_kernel(int64_t lda, float* __restrict__ F, int64_t k0, int kb,
                                                     int* __restrict__ perm, int* __restrict__ info) {
  const int lane = threadIdx.x;
  DiagLane d;
#pragma unroll
  for (int c = 0; c < kLuNb; ++c) d.row[c] = (c < kb && lane < kb) ? F[(k0 + c) * lda + k0 + lane] : 0.0f;
  d.used = lane >= kb;  // rows past the block never pivot
  d.pos = lane;
  d.zero_col = 0;
  diag_step<0>(d, kb, lane);
  const int pos = d.pos, zero_col = d.zero_col;
  const float* row = d.row;
  if (lane < kb) {
#pragma unroll
    for (int c = 0; c < kLuNb; ++c)
      if (c < kb) F[(k0 + c) * lda + k0 + pos] = row[c];
    perm[k0 + pos] = static_cast<int>(k0) + lane;
  }
  if (lane == 0 && zero_col != 0 && info[0] == 0) info[0] = static_cast<int>(k0) + zero_col;
}

// l21 and u12 of the 64-column block at k0, one launch: workgroups
// [0, nl) solve x U11 = a for 256 rows of [k1, e) each (lane = row; the
// block's columns of F), the rest L11 y = P a for 256 columns of [k1, e)
// each (lane = column; the block's rows of F gathered by its permutation,
// L11 unit lower).  The two parts write disjoint parts of F and read only the
// factored block besides their own entries.  Column-oriented substitution:
// once an entry is final, its multiple of U's row (L's column) leaves every
// later entry (independent FMAs rather than one dependent chain per entry).
__global__ __launch_bounds__(256) void lu_l21_u12_kernel(int64_t k1, int64_t e, int64_t lda, float* __restrict__ F,
                                                         int64_t k0, const int* __restrict__ perm, int nl) {
  // l21: Mt[j][c] = U[j][c];  u12: Mt[j][r] = L[r][j]  (16-byte rows)
  __shared__ __attribute__((aligned(16))) float Mt[kLuNb][kLuNb + 4];
  __shared__ float rd[kLuNb];
  __shared__ int src[kLuNb];
  const int tid = threadIdx.x;
  const bool lpart = static_cast<int>(blockIdx.x) < nl;
  {  // the factored block, every load of the thread in flight together
    constexpr int kPer = kLuNb * kLuNb / 256;
    float bv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid + 256 * u;
      bv[u] = F[(k0 + (q >> 6)) * lda + k0 + (q & 63)];
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid + 256 * u, hi = q >> 6, lo = q & 63;
      if (lpart)
        Mt[lo][hi] = bv[u];  // U[lo][hi]
      else
        Mt[hi][lo] = bv[u];  // L[lo][hi]
    }
  }
  if (tid < kLuNb) {
    if (lpart) {
      const float d = F[(k0 + tid) * lda + k0 + tid];
      rd[tid] = (d != 0.0f) ? 1.0f / d : 0.0f;
    } else {
      src[tid] = min(max(perm[k0 + tid] - static_cast<int>(k0), 0), kLuNb - 1);
    }
  }
  __syncthreads();
  float x[kLuNb];
  if (lpart) {
    const int64_t r = k1 + static_cast<int64_t>(blockIdx.x) * 256 + tid;
    if (r >= e) return;  // no barrier follows
#pragma unroll
    for (int c = 0; c < kLuNb; ++c) x[c] = F[(k0 + c) * lda + r];
#pragma unroll
    for (int j = 0; j < kLuNb; ++j) {
      const float xj = x[j] * rd[j];
      x[j] = xj;
#pragma unroll
      for (int c = j + 1; c < kLuNb; ++c) x[c] = fmaf(-xj, Mt[j][c], x[c]);
    }
#pragma unroll
    for (int c = 0; c < kLuNb; ++c) F[(k0 + c) * lda + r] = x[c];
  } else {
    const int64_t c = k1 + static_cast<int64_t>(blockIdx.x - nl) * 256 + tid;
    if (c >= e) return;  // no barrier follows
    float* col = F + c * lda + k0;
#pragma unroll
    for (int r = 0; r < kLuNb; ++r) x[r] = col[src[r]];
#pragma unroll
    for (int j = 0; j < kLuNb - 1; ++j) {
      const float xj = x[j];
#pragma unroll
      for (int r = j + 1; r < kLuNb; ++r) x[r] = fmaf(-Mt[j][r], xj, x[r]);
    }
#pragma unroll
    for (int r = 0; r < kLuNb; ++r) col[r] = x[r];
  }
}

// dupdate: F[k1:e, k1:e] -= L[k1:e, k0:k1] U[k0:k1, k1:e] (the diagonal
// block's Schur complement after block k0, K = 64, FP32).  64 x 64 tile per
// workgroup, 4 x 4 outputs per thread.
__global__ __launch_bounds__(256) void lu_dupdate_kernel(int64_t lda, float* __restrict__ F, int64_t k0, int64_t k1,
                                                         int64_t e) {
  __shared__ float Ls[kLuNb][kLuNb + 1];  // [k][row]
  __shared__ float Us[kLuNb][kLuNb + 1];  // [k][col]
  const int tid = threadIdx.x;
  const int64_t r0 = k1 + static_cast<int64_t>(blockIdx.x) * kLuNb;
  const int64_t cc0 = k1 + static_cast<int64_t>(blockIdx.y) * kLuNb;
  // staging: every load of the thread in flight together (addresses clamped
  // to the last row / column, the values past e replaced by zeros)
  constexpr int kPer = kLuNb * kLuNb / 256;
  float lv[kPer], uv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int q = tid + 256 * u, k = q >> 6, i = q & 63;
    lv[u] = F[(k0 + k) * lda + min(r0 + i, e - 1)];
    uv[u] = F[min(cc0 + k, e - 1) * lda + k0 + i];  // column k, row i of U
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int q = tid + 256 * u, k = q >> 6, i = q & 63;
    Ls[k][i] = (r0 + i < e) ? lv[u] : 0.0f;
    Us[i][k] = (cc0 + k < e) ? uv[u] : 0.0f;
  }
  __syncthreads();
  const int tr = (tid & 15) * 4, tc = (tid >> 4) * 4;
  float acc[4][4] = {};
#pragma unroll 8
  for (int k = 0; k < kLuNb; ++k) {
    float a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = Ls[k][tr + i];
      b[i] = Us[k][tc + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
  }
  if (r0 + kLuNb <= e && cc0 + kLuNb <= e) {  // a whole tile (uniform): its 16 loads in flight together
    float cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) cv[j][i] = F[(cc0 + tc + j) * lda + r0 + tr + i];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) F[(cc0 + tc + j) * lda + r0 + tr + i] = cv[j][i] - acc[i][j];
    return;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t col = cc0 + tc + j;
    if (col >= e) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = r0 + tr + i;
      if (row < e) F[col * lda + row] -= acc[i][j];
    }
  }
}

// Power-of-two scale that puts amax in [2^14, 2^15) (1 for amax = 0 or not
// finite), exponent clamped to [-60, 60] so that products of two scales and
// their reciprocals stay normal FP32 numbers.
__device__ __forceinline__ float block_scale(float amax) {
  if (!(amax > 0.0f) || !(amax < INFINITY)) return 1.0f;
  int e = 0;
  (void)frexpf(amax, &e);  // amax in [2^(e-1), 2^e)
  const int k = min(max(15 - e, -60), 60);
  return ldexpf(1.0f, k);
}

template <typename T>
__device__ __forceinline__ void store8(T* dst, const float* v) {
  if constexpr (sizeof(T) == 2) {
    union {
      T h[8];
      uint4 u;
    } pk;
#pragma unroll
    for (int i = 0; i < 8; ++i) pk.h[i] = static_cast<T>(v[i]);
    *reinterpret_cast<uint4*>(dst) = pk.u;
  } else {
    *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(dst + 4) = make_float4(v[4], v[5], v[6], v[7]);
  }
}

// Packs of kw values per row (kw a multiple of kPackChunk: whole panels),
// K-contiguous, in the GEMM's element type; one workgroup (256 threads) per
// 32 rows of the pack.  With scale != nullptr each 32-row block is multiplied
// by block_scale(its largest magnitude over all kw values), which is stored in
// scale[block].
//   rows of F:    pack row q = F[r0 + q, k0 .. k0 + kw)   (strided in F: LDS transpose)
//   columns of F: pack row q = F[k0 .. k0 + kw, c0 + q]   (contiguous in F)
// Both read with many loads of a thread in flight (unrolled loops, clamped
// addresses instead of guarded loads).
constexpr int kPackChunk = 256;  // columns of F per pass of the row pack
template <typename T>
__global__ __launch_bounds__(256) void lu_pack_rows_kernel(int64_t lda, const float* __restrict__ F, int64_t r0,
                                                           int64_t r1, int64_t k0, int kw, T* __restrict__ dst,
                                                           float* __restrict__ scale) {
  constexpr int KC = kPackChunk, PER = KC / 8;  // per thread and pass: PER columns of its row
  __shared__ float tile[32][KC + 1];
  __shared__ float red[8];
  const int tid = threadIdx.x, lane32 = tid & 31, grp = tid >> 5;
  const int64_t rb = r0 + static_cast<int64_t>(blockIdx.x) * 32;
  const bool live = rb + lane32 < r1;
  const float* const frow = F + min(rb + lane32, r1 - 1);  // rows past r1 read the last row, store zeros
  float s = 1.0f;
  if (scale) {
    float m[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = grp; k < kw; k += 64) {
#pragma unroll
      for (int u = 0; u < 8; ++u) m[u] = fmaxf(m[u], fabsf(frow[(k0 + k + 8 * u) * lda]));
    }
    float mm = live ? fmaxf(fmaxf(fmaxf(m[0], m[1]), fmaxf(m[2], m[3])), fmaxf(fmaxf(m[4], m[5]), fmaxf(m[6], m[7])))
                    : 0.0f;
    for (int o = 16; o > 0; o >>= 1) mm = fmaxf(mm, __shfl_xor(mm, o, 32));
    if (lane32 == 0) red[grp] = mm;
    __syncthreads();
    mm = red[0];
    for (int g = 1; g < 8; ++g) mm = fmaxf(mm, red[g]);
    s = block_scale(mm);
    if (tid == 0) scale[blockIdx.x] = s;
  }
  const int q = tid >> 3, part = tid & 7;  // writer: pack row q, 8 values at 64 c + 8 part
  for (int kc = 0; kc < kw; kc += KC) {
    float v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = frow[(k0 + kc + grp * PER + i) * lda];
    __syncthreads();  // the previous pass's reads of the tile are done
#pragma unroll
    for (int i = 0; i < PER; ++i) tile[lane32][grp * PER + i] = live ? v[i] : 0.0f;
    __syncthreads();
    if (rb + q < r1) {
#pragma unroll
      for (int c = 0; c < KC / 64; ++c) {
        float w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = tile[q][64 * c + part * 8 + i] * s;
        store8<T>(dst + (rb - r0 + q) * kw + kc + 64 * c + part * 8, w);
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lu_pack_cols_kernel(int64_t lda, const float* __restrict__ F, int64_t c0,
                                                           int64_t c1, int64_t k0, int kw, T* __restrict__ dst,
                                                           float* __restrict__ scale) {
  __shared__ float red[8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t cb = c0 + static_cast<int64_t>(blockIdx.x) * 32;
  float s = 1.0f;
  if (scale) {
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int q = wave; q < 32; q += 4) {
      const int64_t c = cb + q;
      if (c >= c1) break;
      const float4* col = reinterpret_cast<const float4*>(F + c * lda + k0);
      for (int k = lane; k < kw / 4; k += 256) {  // kw / 4 a multiple of 256
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float4 v = col[k + 64 * u];
          m[u] = fmaxf(m[u], fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
      }
    }
    float mm = fmaxf(fmaxf(m[0], m[1]), fmaxf(m[2], m[3]));
    for (int o = 32; o > 0; o >>= 1) mm = fmaxf(mm, __shfl_xor(mm, o, 64));
    if (lane == 0) red[wave] = mm;
    __syncthreads();
    mm = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    s = block_scale(mm);
    if (tid == 0) scale[blockIdx.x] = s;
  }
  for (int q = wave; q < 32; q += 4) {  // a wave per column: 64 lanes x 8 values per 512
    const int64_t c = cb + q;
    if (c >= c1) break;
    const float* col = F + c * lda + k0;
    for (int k0l = lane * 8; k0l < kw; k0l += 4 * 512) {  // kw a multiple of 2048 or of 512
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0l + 512 * u;
        if (k < kw) {
          const float4 a = *reinterpret_cast<const float4*>(col + k);
          const float4 b = *reinterpret_cast<const float4*>(col + k + 4);
          const float v[8] = {a.x * s, a.y * s, a.z * s, a.w * s, b.x * s, b.y * s, b.z * s, b.w * s};
          store8<T>(dst + (c - c0) * kw + k, v);
        }
      }
    }
  }
}
static_assert(kLuPanel % kPackChunk == 0 && kLuPanel % 1024 == 0, "packs take whole chunks");

// ---------------------------------------------------------------------------
// Solves with the factors (FP64 vectors, FP32 factors), by panels of
// kLuPanel rows: one workgroup solves the panel's diagonal part block by
// block (its values in LDS), a row-parallel GEMV applies the panel's columns
// to the rows outside it.

// Inverses of the diagonal blocks, once per factorisation (every solve and
// refinement sweep reuses them): block b at k0 = 64 b, lane = column c of the
// inverse, FP64 substitution on the FP32 factors.  Stored transposed,
// Dinv[b][0][c][r] = (L_bb^-1)[r][c] and Dinv[b][1][c][r] = (U_bb^-1)[r][c],
// so that lane r of a solve reads row r of the inverse coalesced.
__global__ __launch_bounds__(128) void lu_block_inverse_kernel(int64_t n, int64_t lda, const float* __restrict__ F,
                                                               int64_t first_block, double* __restrict__ Dinv) {
  __shared__ float blk[kLuNb][kLuNb + 1];  // [col][row]
  // X[w][r][c]: wavefront w's solution (w = 0: L^-1, w = 1: U^-1), lane c's
  // column (one-off kernel: LDS, not registers)
  __shared__ double X[2][kLuNb][kLuNb];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t bidx = first_block + blockIdx.x;
  const int64_t k0 = bidx * kLuNb;
  const int kb = static_cast<int>(min(static_cast<int64_t>(kLuNb), n - k0));
#pragma unroll 8
  for (int c = w; c < kLuNb; c += 2)
    blk[c][lane] = (c < kb && lane < kb) ? F[(k0 + c) * lda + k0 + lane] : (c == lane ? 1.0f : 0.0f);
  __syncthreads();
  double (*Xw)[kLuNb] = X[w];
  double* out = Dinv + (bidx * 2 + w) * kLuNb * kLuNb;
  if (w == 0) {
    // L X = I (unit lower): x_r = delta_rc - sum_{j<r} L[r][j] x_j
    for (int r = 0; r < kLuNb; ++r) {
      double v = (r == lane) ? 1.0 : 0.0;
#pragma unroll 8
      for (int j = 0; j < r; ++j) v = fma(-static_cast<double>(blk[j][r]), Xw[j][lane], v);
      Xw[r][lane] = v;
    }
  } else {
    // U X = I: x_r = (delta_rc - sum_{j>r} U[r][j] x_j) / U[r][r]
    for (int r = kLuNb - 1; r >= 0; --r) {
      double v = (r == lane) ? 1.0 : 0.0;
#pragma unroll 8
      for (int j = r + 1; j < kLuNb; ++j) v = fma(-static_cast<double>(blk[j][r]), Xw[j][lane], v);
      Xw[r][lane] = v / static_cast<double>(blk[r][r]);
    }
  }
  for (int r = 0; r < kLuNb; ++r) out[lane * kLuNb + r] = Xw[r][lane];
}

// sum_c F[(k0 + c) lda + row] ys[c] over one block's columns (p = F + k0 lda +
// row); a full block issues its 64 loads together.
__device__ __forceinline__ double row_block_dot(const float* __restrict__ p, int64_t lda, const double* ys, int kb) {
  double s0 = 0.0, s1 = 0.0;
  if (kb == kLuNb) {
    float v[kLuNb];
#pragma unroll
    for (int c = 0; c < kLuNb; ++c) v[c] = p[c * lda];
#pragma unroll
    for (int c = 0; c < kLuNb; c += 2) {
      s0 = fma(static_cast<double>(v[c]), ys[c], s0);
      s1 = fma(static_cast<double>(v[c + 1]), ys[c + 1], s1);
    }
  } else {
    for (int c = 0; c < kb; ++c) s0 = fma(static_cast<double>(p[c * lda]), ys[c], s0);
  }
  return s0 + s1;
}

// out_r = sum_c Minv[c][r] t[c] for the lane's row r (Minv transposed, 64 x 64).
__device__ __forceinline__ double block_inverse_row(const double* __restrict__ Minv, const double* t, int r, int kb) {
  double s0 = 0.0, s1 = 0.0;
  if (kb == kLuNb) {
    double v[kLuNb];
#pragma unroll
    for (int c = 0; c < kLuNb; ++c) v[c] = Minv[c * kLuNb + r];
#pragma unroll
    for (int c = 0; c < kLuNb; c += 2) {
      s0 = fma(v[c], t[c], s0);
      s1 = fma(v[c + 1], t[c + 1], s1);
    }
  } else {
    for (int c = 0; c < kb; ++c) s0 = fma(Minv[c * kLuNb + r], t[c], s0);
  }
  return s0 + s1;
}

constexpr int kSolveThreads = 1024;  // >= the rows of a panel outside one block

// Forward: y[K0:KE] <- L_panel^-1 P y[K0:KE] (block permutations applied
// after the contributions of the columns left of each block).
__global__ __launch_bounds__(kSolveThreads) void lu_fwd_panel_kernel(int64_t lda, const float* __restrict__ F,
                                                                     const int* __restrict__ perm,
                                                                     const double* __restrict__ Dinv, int64_t K0,
                                                                     int64_t KE, double* __restrict__ y) {
  __shared__ double ys[kLuPanel];
  __shared__ double tb[kLuNb];
  const int tid = threadIdx.x;
  const int len = static_cast<int>(KE - K0);
  for (int r = tid; r < len; r += kSolveThreads) ys[r] = y[K0 + r];
  __syncthreads();
  for (int b0 = 0; b0 < len; b0 += kLuNb) {
    const int kb = min(kLuNb, len - b0);
    const int64_t k0 = K0 + b0;
    if (tid < kb) {  // the block's values in factored order
      const int src = min(max(perm[k0 + tid] - static_cast<int>(k0), 0), kb - 1);
      tb[tid] = ys[b0 + src];
    }
    __syncthreads();
    if (tid < kb) ys[b0 + tid] = block_inverse_row(Dinv + (k0 / kLuNb) * 2 * kLuNb * kLuNb, tb, tid, kb);
    __syncthreads();
    const int r = b0 + kb + tid;
    if (r < len) ys[r] -= row_block_dot(F + k0 * lda + K0 + r, lda, ys + b0, kb);
    __syncthreads();
  }
  for (int r = tid; r < len; r += kSolveThreads) y[K0 + r] = ys[r];
}

// Backward: y[K0:KE] <- U_panel^-1 y[K0:KE].
__global__ __launch_bounds__(kSolveThreads) void lu_bwd_panel_kernel(int64_t lda, const float* __restrict__ F,
                                                                     const double* __restrict__ Dinv, int64_t K0,
                                                                     int64_t KE, double* __restrict__ y) {
  __shared__ double ys[kLuPanel];
  __shared__ double tb[kLuNb];
  const int tid = threadIdx.x;
  const int len = static_cast<int>(KE - K0);
  for (int r = tid; r < len; r += kSolveThreads) ys[r] = y[K0 + r];
  __syncthreads();
  for (int b0 = ((len - 1) / kLuNb) * kLuNb; b0 >= 0; b0 -= kLuNb) {
    const int kb = min(kLuNb, len - b0);
    const int64_t k0 = K0 + b0;
    if (tid < kb) tb[tid] = ys[b0 + tid];
    __syncthreads();
    if (tid < kb)
      ys[b0 + tid] = block_inverse_row(Dinv + (k0 / kLuNb) * 2 * kLuNb * kLuNb + kLuNb * kLuNb, tb, tid, kb);
    __syncthreads();
    if (tid < b0) ys[tid] -= row_block_dot(F + k0 * lda + K0 + tid, lda, ys + b0, kb);
    __syncthreads();
  }
  for (int r = tid; r < len; r += kSolveThreads) y[K0 + r] = ys[r];
}

// t[r] = sum_k M[r W + k] x[k], r < W, for one of a panel's FP32 solve
// inverses M (row-major W x W): one wavefront per row (16-byte loads along the
// row, FP64 sums, reduced across the wavefront), kBgRows rows per workgroup.
constexpr int kBgRows = 8;
__global__ __launch_bounds__(256) void lu_block_gemv_kernel(const float* __restrict__ M,
                                                            const double* __restrict__ x, double* __restrict__ t) {
  constexpr int W = kLuPanel, Q = W / 256;  // float4 loads per lane and row
  constexpr int R = kBgRows / 4;            // rows per wavefront
  __shared__ double xs[W];
  for (int k = threadIdx.x; k < W; k += 256) xs[k] = x[k];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rw = blockIdx.x * kBgRows + wave * R;
  float4 v[R][Q];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int q = 0; q < Q; ++q)
      v[i][q] = reinterpret_cast<const float4*>(M + static_cast<int64_t>(rw + i) * W)[lane + 64 * q];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double* xk = xs + 4 * (lane + 64 * q);
      s0 = fma(static_cast<double>(v[i][q].x), xk[0], s0);
      s1 = fma(static_cast<double>(v[i][q].y), xk[1], s1);
      s0 = fma(static_cast<double>(v[i][q].z), xk[2], s0);
      s1 = fma(static_cast<double>(v[i][q].w), xk[3], s1);
    }
    double sum = s0 + s1;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (lane == 0) t[rw + i] = sum;
  }
}
static_assert(kLuPanel % kBgRows == 0 && kBgRows % 4 == 0 && kLuPanel % 256 == 0, "block GEMV tiling");

// y[r0:r1] -= F[r0:r1, c0:c0+w] x[0:w] (w <= W; the rows disjoint from the
// columns' block; r0 and lda multiples of 4).  Workgroup = kSgRows rows x all
// w columns: wavefront v takes columns [v W / kSgWaves, (v + 1) W / kSgWaves),
// lane 2 consecutive rows (one 8-byte load per column), FP64 sums; the
// wavefronts' partial sums are added in LDS in a fixed order.  128 rows per
// workgroup (256 before, lane 4 rows: the same sums in the same order) give a
// panel with r rows below it r / 128 workgroups, more of the chip for the
// panels past the first third.  With copy != nullptr, workgroup 0 also
// writes x to copy[0:w] (the solved block back into the vector).
constexpr int kSgWaves = 16, kSgRows = 128;
__global__ __launch_bounds__(64 * kSgWaves) void lu_solve_gemv_kernel(int64_t lda, const float* __restrict__ F,
                                                                       int64_t c0, int w, int64_t r0, int64_t r1,
                                                                       const double* __restrict__ x,
                                                                       double* __restrict__ y,
                                                                       double* __restrict__ copy) {
  constexpr int CW = kLuPanel / kSgWaves;  // columns per wavefront
  constexpr int U = 16;                    // loads in flight per lane
  __shared__ double xs[kLuPanel];
  __shared__ double part[kSgWaves][kSgRows];
  const int tid = threadIdx.x;
  for (int c = tid; c < w; c += 64 * kSgWaves) xs[c] = x[c];
  if (copy != nullptr && blockIdx.x == 0)
    for (int c = tid; c < w; c += 64 * kSgWaves) copy[c] = x[c];
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int64_t rt = r0 + static_cast<int64_t>(blockIdx.x) * kSgRows;  // the workgroup's first row
  const int64_t rb = rt + 2 * lane;                                     // the lane's first row
  const int cb = wave * CW, ce = min(w, cb + CW);
  double s[2] = {0.0, 0.0};
  if (rb + 2 <= r1) {
    const float* p = F + c0 * lda + rb;
    int c = cb;
    for (; c + U <= ce; c += U) {
      float2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const float2*>(p + (c + u) * lda);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double xc = xs[c + u];
        s[0] = fma(static_cast<double>(v[u].x), xc, s[0]);
        s[1] = fma(static_cast<double>(v[u].y), xc, s[1]);
      }
    }
    for (; c < ce; ++c) {
      const float2 v = *reinterpret_cast<const float2*>(p + c * lda);
      s[0] = fma(static_cast<double>(v.x), xs[c], s[0]);
      s[1] = fma(static_cast<double>(v.y), xs[c], s[1]);
    }
  } else if (rb < r1) {
    for (int c = cb; c < ce; ++c) s[0] = fma(static_cast<double>(F[(c0 + c) * lda + rb]), xs[c], s[0]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) part[wave][2 * lane + i] = s[i];
  __syncthreads();
  if (tid < kSgRows && rt + tid < r1) {
    double acc = part[0][tid];
#pragma unroll
    for (int v = 1; v < kSgWaves; ++v) acc += part[v][tid];
    y[rt + tid] -= acc;
  }
}
static_assert(kSgRows == 2 * 64 && kLuPanel % kSgWaves == 0 && 64 * kSgWaves >= kSgRows, "solve GEMV tiling");

// The inverses of a factored W x W diagonal block D = F[K0:K0+W, K0:K0+W]:
// the panel solves above applied to unit vectors, kInvCols of them per
// workgroup (FP32 vectors in LDS, FP64 block-inverse products).
//   FWD: Tfwd = the forward operator (block permutations, L_D^-1), written
//        row-major: out[r W + k] = Tfwd[r][k] (the A operand of U12 = Tfwd A12);
//   BWD: Uinv = U_D^-1, written by columns: out[c W + r] = Uinv[r][c] (the
//        B operand of L21 = A21 Uinv).
// Both come from one launch: workgroups [0, W / kInvCols) form Tfwd, the
// rest Uinv (the two sweeps are independent).  Each is also written in FP32,
// row-major (sol[r W + k]), for the solves (lu_block_gemv_kernel).
constexpr int kInvCols = 8;
template <typename T, bool FWD>
__device__ __forceinline__ void panel_inverse_block(int64_t lda, const float* __restrict__ F,
                                                    const int* __restrict__ perm, const double* __restrict__ Dinv,
                                                    int64_t K0, int blk, T* __restrict__ out,
                                                    float* __restrict__ sol, float (*__restrict__ ys)[kLuPanel],
                                                    float (*__restrict__ tb)[kLuNb], double* __restrict__ Ms) {
  constexpr int W = kLuPanel;
  const int tid = threadIdx.x;
  const int j0 = blk * kInvCols;
  for (int q = tid; q < kInvCols * W; q += 256) {
    const int c = q / W, r = q % W;
    ys[c][r] = (r == j0 + c) ? 1.0f : 0.0f;
  }
  __syncthreads();
  constexpr int nb = W / kLuNb;
  for (int bi = 0; bi < nb; ++bi) {
    const int b0 = (FWD ? bi : nb - 1 - bi) * kLuNb;
    const int64_t k0 = K0 + b0;
    const double* Minv = Dinv + (k0 / kLuNb) * 2 * kLuNb * kLuNb + (FWD ? 0 : kLuNb * kLuNb);
    // the block inverse into LDS, all its loads in flight together
    constexpr int kPer = kLuNb * kLuNb / 256;
    double mv[kPer];
#pragma unroll
    for (int t = 0; t < kPer; ++t) mv[t] = Minv[t * 256 + tid];
#pragma unroll
    for (int t = 0; t < kPer; ++t) Ms[t * 256 + tid] = mv[t];
    for (int q = tid; q < kInvCols * kLuNb; q += 256) {
      const int c = q >> 6, i = q & 63;
      const int src = FWD ? min(max(perm[k0 + i] - static_cast<int>(k0), 0), kLuNb - 1) : i;
      tb[c][i] = ys[c][b0 + src];
    }
    __syncthreads();
    for (int q = tid; q < kInvCols * kLuNb; q += 256) {
      const int c = q >> 6, i = q & 63;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < kLuNb; ++j) acc[j & 3] = fma(Ms[j * kLuNb + i], static_cast<double>(tb[c][j]), acc[j & 3]);
      ys[c][b0 + i] = static_cast<float>((acc[0] + acc[1]) + (acc[2] + acc[3]));
    }
    __syncthreads();
    const int ra = FWD ? b0 + kLuNb : 0, rz = FWD ? W : b0;
    for (int r = ra + tid; r < rz; r += 256) {
      const float* p = F + k0 * lda + K0 + r;
      float acc[kInvCols] = {};
      float f[kLuNb];  // the row's 64 values, all loads in flight together
#pragma unroll
      for (int j = 0; j < kLuNb; ++j) f[j] = p[j * lda];
#pragma unroll
      for (int j = 0; j < kLuNb; ++j) {
#pragma unroll
        for (int c = 0; c < kInvCols; ++c) acc[c] = fmaf(f[j], ys[c][b0 + j], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < kInvCols; ++c) ys[c][r] -= acc[c];
    }
    __syncthreads();
  }
  for (int r = tid; r < W; r += 256) {
    float v[kInvCols];
#pragma unroll
    for (int c = 0; c < kInvCols; ++c) v[c] = ys[c][r];
    if (FWD) store8<T>(out + static_cast<int64_t>(r) * W + j0, v);
    store8<float>(sol + static_cast<int64_t>(r) * W + j0, v);
  }
  if (!FWD) {
    for (int q = tid; q < kInvCols * W / 8; q += 256) {
      const int c = q / (W / 8), r = (q % (W / 8)) * 8;
      store8<T>(out + static_cast<int64_t>(j0 + c) * W + r, &ys[c][r]);
    }
  }
}
static_assert(kInvCols == 8, "the inverse kernel stores 8 values per thread");

template <typename T>
__global__ __launch_bounds__(256) void lu_panel_inverses_kernel(int64_t lda, const float* __restrict__ F,
                                                                const int* __restrict__ perm,
                                                                const double* __restrict__ Dinv, int64_t K0,
                                                                T* __restrict__ tfwd, T* __restrict__ uinv,
                                                                float* __restrict__ tsol, float* __restrict__ usol) {
  constexpr int G = kLuPanel / kInvCols;  // workgroups per inverse
  __shared__ __attribute__((aligned(16))) float ys[kInvCols][kLuPanel];  // the workgroup's vectors
  __shared__ float tb[kInvCols][kLuNb];
  __shared__ double Ms[kLuNb * kLuNb];  // the current diagonal block's inverse
  if (static_cast<int>(blockIdx.x) < G)
    panel_inverse_block<T, true>(lda, F, perm, Dinv, K0, blockIdx.x, tfwd, tsol, ys, tb, Ms);
  else
    panel_inverse_block<T, false>(lda, F, perm, Dinv, K0, blockIdx.x - G, uinv, usol, ys, tb, Ms);
}

static unsigned grid_of(int64_t len, int per) { return static_cast<unsigned>((len + per - 1) / per); }

// Blocked right-looking LU with pivoting inside the 64-wide diagonal blocks,
// panels of W = kLuPanel columns taken in groups of G = kLuGroup, so that the
// trailing matrix is updated once per group, with K = G W (the C traffic per
// flop of one update per panel, divided by G).  Group [K0, KG), panels
// i = 0 .. G-1 at K_i = K0 + i W, each brought up to date left-looking:
//   for each earlier panel j: its rows of panel i's columns -= the panels
//   before j (K = j W), then U12_j over panel i's columns; panel i's columns
//   below -= the panels before i (K = i W); factor panel i (diagonal blocks,
//   inverses, L21 into F).
// When columns remain right of the group (main stream): for each panel j, its
// rows of the trailing columns -= the panels before j, then U12_j over them;
// then the trailing update A22 -= L21 U12 with K = G W, from packs of F's
// columns [K0, KG) and rows [K0, KG).
// Look-ahead: the update first covers the next group's columns; the aux
// stream then brings that group's panels up to date, and packs their L21 for
// the group's trailing update, while the main stream updates the rest.  The
// two streams touch disjoint columns of F and disjoint buffers (the group's
// panels use Y, X, Ti[], Ui, the update packs of slot 0 and the L21 pack of
// the group's parity; the main stream's preparation of the same group, which
// uses X, Ti[] and slot 1, finishes before the next group's panels start, and
// its update reads the other parity's L21 pack).
template <typename T>
static hipError_t lu_factor_t(LuWorkspace& ws, LuMode mode, hipStream_t s, LuResult* res) {
  constexpr int G = kLuGroup;
  const int64_t n = ws.n, lda = ws.lda, W = kLuPanel, GW = G * kLuPanel;
  const bool scaled = mode != LuMode::kFP32;
  float* F = ws.F;
  auto Lj = [&](int64_t K0) { return static_cast<T*>(ws.At[(K0 / GW) & 1]); };  // by the group's parity
  auto sAj = [&](int64_t K0) { return ws.sA[(K0 / GW) & 1]; };
  T* Uj = static_cast<T*>(ws.Bt);
  T* X = static_cast<T*>(ws.Xm);
  T* Y = static_cast<T*>(ws.Ya);
  T* Ui = static_cast<T*>(ws.Uinv);
  auto Ti = [&](int slot) { return static_cast<T*>(ws.Tinv) + static_cast<int64_t>(slot) * W * W; };
  auto at = [&](int64_t r, int64_t c) { return F + c * lda + r; };
  auto sc = [&](float* p) { return scaled ? p : nullptr; };
  hipError_t e = hipSuccess;
  int nev = 0;
  // C -= A B on stream st; the updates of the columns right of a group (main
  // stream, `count`) are timed and their flops counted
  auto sub = [&](int64_t m, int64_t nc, int64_t K, const T* A, const T* B, float* C, const float* sa,
                 const float* sb, hipStream_t st, bool count, int cap = 0) -> hipError_t {
    const bool timed = count && nev < LuWorkspace::kEvents;
    if (timed) (void)hipEventRecord(ws.ev[2 * nev], s);
    hipError_t ee = gemm_packed(mode, false, m, nc, K, A, K, B, K, C, lda, st, sa, sb, cap);
    if (ee != hipSuccess) return ee;
    if (timed) {
      (void)hipEventRecord(ws.ev[2 * nev + 1], s);
      ++nev;
      res->gemm_flop += 2.0 * static_cast<double>(m) * static_cast<double>(nc) * static_cast<double>(K);
    }
    return hipSuccess;
  };
  // F[r0:r1, c0:c1] -= F[r0:r1, K0:K0+kw] F[K0:K0+kw, c0:c1] from the packs of slot ps
  auto update = [&](int64_t r0, int64_t r1, int64_t c0, int64_t c1, int64_t K0, int64_t kw, int ps,
                    hipStream_t st, bool count) -> hipError_t {
    T* Pr = static_cast<T*>(ws.Pr[ps]);
    T* Pc = static_cast<T*>(ws.Pc[ps]);
    hipLaunchKernelGGL((lu_pack_rows_kernel<T>), dim3(grid_of(r1 - r0, 32)), dim3(256), 0, st, lda, F, r0, r1, K0,
                       static_cast<int>(kw), Pr, sc(ws.sPr[ps]));
    hipLaunchKernelGGL((lu_pack_cols_kernel<T>), dim3(grid_of(c1 - c0, 32)), dim3(256), 0, st, lda, F, c0, c1, K0,
                       static_cast<int>(kw), Pc, sc(ws.sPc[ps]));
    return sub(r1 - r0, c1 - c0, kw, Pr, Pc, at(r0, c0), sc(ws.sPr[ps]), sc(ws.sPc[ps]), st, count);
  };
  // panel [K0, KE) on stream st: diagonal blocks, block inverses and, when
  // rows remain below, the forward operator of D into Ti(slot) and L21 into F
  auto factor_panel = [&](int64_t K0, int64_t KE, int slot, hipStream_t st) -> hipError_t {
    for (int64_t k0 = K0; k0 < KE; k0 += kLuNb) {
      const int kb = static_cast<int>(std::min<int64_t>(kLuNb, KE - k0));
      const int64_t k1 = k0 + kb;
      hipLaunchKernelGGL(lu_diag_kernel, dim3(1), dim3(64), 0, st, lda, F, k0, kb, ws.perm, ws.info);
      if (k1 < KE) {  // a full block (only the last block of the matrix is short)
        const int nl = static_cast<int>(grid_of(KE - k1, 256));
        hipLaunchKernelGGL(lu_l21_u12_kernel, dim3(2 * nl), dim3(256), 0, st, k1, KE, lda, F, k0, ws.perm, nl);
        hipLaunchKernelGGL(lu_dupdate_kernel, dim3(grid_of(KE - k1, kLuNb), grid_of(KE - k1, kLuNb)), dim3(256), 0,
                           st, lda, F, k0, k1, KE);
      }
    }
    hipLaunchKernelGGL(lu_block_inverse_kernel, dim3(grid_of(KE - K0, kLuNb)), dim3(128), 0, st, n, lda, F,
                       K0 / kLuNb, ws.Dinv);
    if (KE >= n) return hipGetLastError();
    hipLaunchKernelGGL((lu_panel_inverses_kernel<T>), dim3(2 * W / kInvCols), dim3(256), 0, st, lda, F, ws.perm,
                       ws.Dinv, K0, Ti(slot), Ui, ws.Tsol + (K0 / W) * W * W, ws.Usol + (K0 / W) * W * W);
    // L21 = A21 U_D^-1
    hipLaunchKernelGGL((lu_pack_rows_kernel<T>), dim3(grid_of(n - KE, 32)), dim3(256), 0, st, lda, F, KE, n, K0,
                       static_cast<int>(W), Y, nullptr);
    hipError_t ee = gemm_packed(mode, true, n - KE, W, W, Y, W, Ui, W, at(KE, K0), lda, st);
    if (ee != hipSuccess) return ee;
    return hipGetLastError();
  };
  // U12 of the full panel [K0, K0 + W) over columns [c0, c1): Ti(slot) A12
  auto form_u12 = [&](int64_t K0, int64_t c0, int64_t c1, int slot, hipStream_t st) -> hipError_t {
    hipLaunchKernelGGL((lu_pack_cols_kernel<T>), dim3(grid_of(c1 - c0, 32)), dim3(256), 0, st, lda, F, c0, c1, K0,
                       static_cast<int>(W), X, nullptr);
    hipError_t ee = gemm_packed(mode, true, W, c1 - c0, W, Ti(slot), W, X, W, at(K0, c0), lda, st);
    if (ee != hipSuccess) return ee;
    return hipGetLastError();
  };
  // the panels of the group at K0, brought up to date and factored, on st
  auto group_panels = [&](int64_t K0, hipStream_t st) -> hipError_t {
    hipError_t ee = hipSuccess;
    for (int i = 0; i < G; ++i) {
      const int64_t Ki = K0 + i * W;
      if (Ki >= n) break;
      const int64_t KEi = std::min<int64_t>(n, Ki + W);
      for (int j = 0; j < i; ++j) {  // rows of panel j over panel i's columns, then U12_j there
        const int64_t Kj = K0 + j * W;
        if (j > 0 && (ee = update(Kj, Kj + W, Ki, KEi, K0, j * W, 0, st, false)) != hipSuccess) return ee;
        if ((ee = form_u12(Kj, Ki, KEi, j, st)) != hipSuccess) return ee;
      }
      if (i > 0 && (ee = update(Ki, n, Ki, KEi, K0, i * W, 0, st, false)) != hipSuccess) return ee;
      if ((ee = factor_panel(Ki, KEi, i, st)) != hipSuccess) return ee;
    }
    // the L21 pack of the trailing update, here (off the main stream for the
    // look-ahead groups): the group's L21 rows below it are final
    if (K0 + GW < n)
      hipLaunchKernelGGL((lu_pack_rows_kernel<T>), dim3(grid_of(n - K0 - GW, 32)), dim3(256), 0, st, lda, F, K0 + GW,
                         n, K0, static_cast<int>(GW), Lj(K0), sc(sAj(K0)));
    return hipGetLastError();
  };
  // the group's rows of the columns right of it, and the U12 pack of the
  // update (main stream); the group's panels are all full here
  auto group_prepare = [&](int64_t K0) -> hipError_t {
    const int64_t KG = K0 + GW;
    hipError_t ee = hipSuccess;
    for (int j = 0; j < G; ++j) {
      const int64_t Kj = K0 + j * W;
      if (j > 0 && (ee = update(Kj, Kj + W, KG, n, K0, j * W, 1, s, true)) != hipSuccess) return ee;
      if ((ee = form_u12(Kj, KG, n, j, s)) != hipSuccess) return ee;
    }
    hipLaunchKernelGGL((lu_pack_cols_kernel<T>), dim3(grid_of(n - KG, 32)), dim3(256), 0, s, lda, F, KG, n, K0,
                       static_cast<int>(GW), Uj, sc(ws.sB));
    return hipGetLastError();
  };
  // A22[KG:n, c0:c1] -= L21 U12[c0:c1], K = G W
  auto trailing = [&](int64_t KG, int64_t c0, int64_t c1, int cap) -> hipError_t {
    return sub(n - KG, c1 - c0, GW, Lj(KG - GW), Uj + (c0 - KG) * GW, at(KG, c0), sc(sAj(KG - GW)),
               scaled ? ws.sB + (c0 - KG) / 32 : nullptr, s, true, cap);
  };
  if ((e = group_panels(0, s)) != hipSuccess) return e;
  for (int64_t K0 = 0; K0 + GW < n; K0 += GW) {
    const int64_t KG = K0 + GW;
    if ((e = group_prepare(K0)) != hipSuccess) return e;
    // the next group's columns, then (look-ahead) that group's panels on the
    // aux stream while the main stream updates the rest
    const int64_t NE = std::min<int64_t>(n, KG + GW);
    if ((e = trailing(KG, KG, NE, 0)) != hipSuccess) return e;
    if ((e = hipEventRecord(ws.next_cols, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(ws.aux, ws.next_cols, 0)) != hipSuccess) return e;
    if ((e = group_panels(KG, ws.aux)) != hipSuccess) return e;
    if ((e = hipEventRecord(ws.next_done, ws.aux)) != hipSuccess) return e;
    // in the last groups the panels' chain outlasts this update: it leaves
    // CUs free for the chain's small kernels (the mixed modes; the FP32
    // update is several times longer than the chain)
    const int cap = (mode != LuMode::kFP32 && n - KG <= kLuTailRows) ? kLuTailWorkgroups : 0;
    if (NE < n && (e = trailing(KG, NE, n, cap)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(s, ws.next_done, 0)) != hipSuccess) return e;
  }
  res->gemm_launches = nev;
  return hipGetLastError();
}

static hipError_t lu_factor(LuWorkspace& ws, LuMode mode, hipStream_t s, LuResult* res) {
  switch (mode) {
    case LuMode::kFP32:
      return lu_factor_t<float>(ws, mode, s, res);
    case LuMode::kBF16:
      return lu_factor_t<__bf16>(ws, mode, s, res);
    case LuMode::kFP16:
      return lu_factor_t<_Float16>(ws, mode, s, res);
  }
  return hipErrorInvalidValue;
}

// y <- (LU)^-1 P y.  Panels with rows below them (their FP32 solve inverses
// stored): the block's forward operator (or U_D^-1) times y's block into the
// scratch t, then the rows outside the block -= the panel times t, the same
// launch writing t back into y.  The last panel (no inverses): the
// block-by-block panel kernels, in place.
static hipError_t lu_apply_inverse(LuWorkspace& ws, double* y, hipStream_t s) {
  const int64_t n = ws.n, lda = ws.lda, W = kLuPanel;
  const int64_t KL = ((n - 1) / W) * W;  // the last panel
  double* t = ws.vec;
  auto gemv = [&](int64_t c0, int w, int64_t r0, int64_t r1, const double* xv, double* copy) {
    const unsigned grid = std::max(1u, grid_of(r1 - r0, kSgRows));
    hipLaunchKernelGGL(lu_solve_gemv_kernel, dim3(grid), dim3(64 * kSgWaves), 0, s, lda, ws.F, c0, w, r0, r1, xv,
                       y, copy);
  };
  for (int64_t K0 = 0; K0 < KL; K0 += W) {
    hipLaunchKernelGGL(lu_block_gemv_kernel, dim3(W / kBgRows), dim3(256), 0, s, ws.Tsol + (K0 / W) * W * W, y + K0,
                       t);
    gemv(K0, static_cast<int>(W), K0 + W, n, t, y + K0);
  }
  hipLaunchKernelGGL(lu_fwd_panel_kernel, dim3(1), dim3(kSolveThreads), 0, s, lda, ws.F, ws.perm, ws.Dinv, KL, n, y);
  hipLaunchKernelGGL(lu_bwd_panel_kernel, dim3(1), dim3(kSolveThreads), 0, s, lda, ws.F, ws.Dinv, KL, n, y);
  if (KL > 0) gemv(KL, static_cast<int>(n - KL), 0, KL, y + KL, nullptr);
  for (int64_t K0 = KL - W; K0 >= 0; K0 -= W) {
    hipLaunchKernelGGL(lu_block_gemv_kernel, dim3(W / kBgRows), dim3(256), 0, s, ws.Usol + (K0 / W) * W * W, y + K0,
                       t);
    gemv(K0, static_cast<int>(W), 0, K0, t, y + K0);
  }
  return hipGetLastError();
}

hipError_t lu_workspace_create(LuMode mode, int64_t n, int64_t lda, LuWorkspace* ws) {
  if (n < 1 || lda < n || lda % 4 != 0) return hipErrorInvalidValue;  // the solves load 4 rows at a time
  ws->mode = mode;
  ws->n = n;
  ws->lda = lda;
  const size_t esz = mode == LuMode::kFP32 ? 4 : 2;
  const size_t pack = esz * static_cast<size_t>(n) * kLuPanel;
  const size_t inv = esz * static_cast<size_t>(kLuPanel) * kLuPanel;
  const size_t scales = sizeof(float) * static_cast<size_t>((n + 31) / 32);
  const size_t dinv = sizeof(double) * 2 * kLuNb * kLuNb * static_cast<size_t>((n + kLuNb - 1) / kLuNb);
  hipError_t e;
  if ((e = hipMalloc(&ws->F, sizeof(float) * static_cast<size_t>(lda) * n)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->perm, sizeof(int) * static_cast<size_t>(n))) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->info, sizeof(int))) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->vec, sizeof(double) * static_cast<size_t>(n))) != hipSuccess) return e;
  constexpr int G = kLuGroup;
  const size_t wide = esz * static_cast<size_t>(n) * (G - 1) * kLuPanel;     // n rows, (G - 1) panels deep
  const size_t narrow = esz * static_cast<size_t>(kLuPanel) * (G - 1) * kLuPanel;
  for (void** p : {&ws->At[0], &ws->At[1], &ws->Bt})  // the group's panels
    if ((e = hipMalloc(p, G * pack)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Pr[0], wide)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Pc[0], narrow)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Pr[1], narrow)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Pc[1], wide)) != hipSuccess) return e;
  for (void** p : {&ws->Xm, &ws->Ya})
    if ((e = hipMalloc(p, pack)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Tinv, G * inv)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Uinv, inv)) != hipSuccess) return e;
  const size_t sol = sizeof(float) * static_cast<size_t>((n - 1) / kLuPanel) * kLuPanel * kLuPanel;
  if (sol > 0)
    for (float** p : {&ws->Tsol, &ws->Usol})
      if ((e = hipMalloc(p, sol)) != hipSuccess) return e;
  for (float** p : {&ws->sA[0], &ws->sA[1], &ws->sB, &ws->sPr[0], &ws->sPr[1], &ws->sPc[0], &ws->sPc[1]})
    if ((e = hipMalloc(p, scales)) != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&ws->aux, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&ws->next_cols, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&ws->next_done, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->Dinv, dinv)) != hipSuccess) return e;
  for (auto& ev : ws->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
  ws->bytes = static_cast<int64_t>(sizeof(float) * static_cast<size_t>(lda) * n + sizeof(int) * (n + 1) +
                                   sizeof(double) * static_cast<size_t>(n) + (3 * G + 2) * pack + 2 * wide +
                                   2 * narrow + (G + 1) * inv + 2 * sol + 7 * scales + dinv);
  return hipSuccess;
}

int64_t lu_workspace_bytes(const LuWorkspace& ws) { return ws.bytes; }

void lu_workspace_destroy(LuWorkspace* ws) {
  if (ws->aux) (void)hipStreamSynchronize(ws->aux);
  for (void* p : {static_cast<void*>(ws->F), static_cast<void*>(ws->perm), static_cast<void*>(ws->info),
                  static_cast<void*>(ws->vec), ws->At[0], ws->At[1], ws->Bt, ws->Pr[0], ws->Pr[1], ws->Pc[0],
                  ws->Pc[1], ws->Xm, ws->Ya, ws->Tinv, ws->Uinv, static_cast<void*>(ws->sA[0]),
                  static_cast<void*>(ws->sA[1]), static_cast<void*>(ws->sB),
                  static_cast<void*>(ws->sPr[0]), static_cast<void*>(ws->sPr[1]), static_cast<void*>(ws->sPc[0]),
                  static_cast<void*>(ws->sPc[1]), static_cast<void*>(ws->Dinv), static_cast<void*>(ws->Tsol),
                  static_cast<void*>(ws->Usol)})
    if (p) (void)hipFree(p);
  if (ws->next_cols) (void)hipEventDestroy(ws->next_cols);
  if (ws->next_done) (void)hipEventDestroy(ws->next_done);
  if (ws->aux) (void)hipStreamDestroy(ws->aux);
  for (auto& ev : ws->ev)
    if (ev) (void)hipEventDestroy(ev);
  *ws = LuWorkspace{};
}

hipError_t lu_solve(LuWorkspace& ws, LuMode mode, const LuSystem& sys, int64_t n, const double* b, double* x,
                    double* work, double* dpart, double* hdev, double tol, int maxit, hipStream_t s,
                    LuResult* result) {
  hipError_t e;
  *result = LuResult{};
  if (mode != ws.mode || n != ws.n || !sys.fill || !sys.residual) return hipErrorInvalidValue;
  if ((e = sys.fill(ws.F, ws.lda, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(ws.info, 0, sizeof(int), s)) != hipSuccess) return e;
  if ((e = lu_factor(ws, mode, s, result)) != hipSuccess) return e;
  int info = 0;
  if ((e = hipMemcpyAsync(&info, ws.info, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  for (int k = 0; k < result->gemm_launches; ++k) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ws.ev[2 * k], ws.ev[2 * k + 1]) == hipSuccess) result->gemm_ms += ms;
  }
  result->info = info;
  if (info != 0) return hipSuccess;
  double* r = work;
  // ||b||
  if ((e = mdot(n, 1, b, n, b, dpart, hdev, s)) != hipSuccess) return e;
  double bb = 0.0;
  if ((e = hipMemcpyAsync(&bb, hdev, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const double bnorm = std::sqrt(bb);
  // x = (LU)^-1 P b
  if ((e = axpby(n, 1.0, b, 0.0, x, s)) != hipSuccess) return e;
  if ((e = lu_apply_inverse(ws, x, s)) != hipSuccess) return e;
  double res = 0.0;
  for (int it = 0;; ++it) {
    if ((e = sys.residual(x, r, s)) != hipSuccess) return e;
    if ((e = mdot(n, 1, r, n, r, dpart, hdev, s)) != hipSuccess) return e;
    double rr = 0.0;
    if ((e = hipMemcpyAsync(&rr, hdev, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    res = (bnorm > 0.0) ? std::sqrt(rr) / bnorm : 0.0;
    result->refinement_iterations = it;
    if (!(res > tol) || it >= maxit || !std::isfinite(res)) break;
    if ((e = lu_apply_inverse(ws, r, s)) != hipSuccess) return e;
    if ((e = axpby(n, 1.0, r, 1.0, x, s)) != hipSuccess) return e;
  }
  result->relative_residual = res;
  result->converged = res <= tol;
  return hipSuccess;
}

}  // namespace rf
