// rf_gemm.hip -- C -= A B (or C = A B) on MFMA for the LU factorisation.
//
// The operands are packed K-contiguous: row r of At holds A[r][0..K) and row c
// of Bt holds B[0..K)[c] (the LU writes these packs while it solves its
// triangular blocks), in the element type of the mode:
//   float    v_mfma_f32_32x32x2_f32    (the plain FP32 --magma path)
//   __bf16   v_mfma_f32_32x32x16_bf16
//   _Float16 v_mfma_f32_32x32x16_f16
// always with FP32 accumulation into the FP32 C (column-major, ldc).
//
// Per-block scaling: the packs of the mixed modes may hold each 32-row block
// of A (32-column block of B) multiplied by a power of two, sa[r / 32]
// (sb[c / 32]); the epilogue multiplies each 32 x 32 accumulator block by
// 1 / (sa sb), exactly, before it updates C.
//
// Tiling: a workgroup of WM x WN waves owns a 256x256 tile of C; each wave
// computes a (256/WM) x (256/WN) sub-tile as 32x32 MFMA blocks.  K advances
// in stages of 128 bytes per row (64 BF16/FP16 or 32 FP32 values),
// double-buffered in LDS (2 x 2 x 256 rows of 144 bytes; the 16-byte pad
// keeps the 16-byte fragment reads of 8 lanes on distinct banks).  Global
// loads (16 bytes per lane, 8 lanes per 128-byte row segment) run PF stages
// ahead of the MFMAs through registers.
//
// MFMA operand layouts (per lane l):
//   32x32x16 (2-byte types): A[i = l % 32][k = 8 (l / 32) .. +8], likewise B;
//   32x32x2  (FP32): A[i = l % 32][k = l / 32]; a 16-byte read of k = 4 (l / 32)
//            .. +4 feeds four MFMAs, the j-th pairing k = j and k = 4 + j.
//   Output accumulator i: column l % 32, row 8 (i / 4) + 4 (l / 32) + i % 4.
#include <type_traits>

#include "rf_lu.h"

namespace rf {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int kTile = 256;
// Tile order: with kGemmGroupRows > 0 the grid is one-dimensional and each of
// the 8 XCDs (workgroup w runs on XCD w % 8) takes a contiguous run of the
// tiles ordered in groups of kGemmGroupRows row tiles (rows fastest inside a
// group, then columns), so that the workgroups resident on an XCD share a few
// A and B panels (with K = 3072, each tile's A and B panels are 1.5 MB each:
// row tiles fastest, the 32 workgroups on an XCD read 32 different A panels,
// ~400 MB over the chip, more than the 256 MB MALL; groups of 4 rows keep
// them in it); 0: a 2-D grid, row tiles fastest.  In the LU at n = 75 000
// (BF16 / FP16 ms per step, trailing GEMM TFLOP/s): 0 -> 488 / 438 (743 /
// 860), 4 -> 466 / 422 (790 / 897), 8 -> 468 / 425, 16 -> 480 / 439.
constexpr int kGemmGroupRows = 4;
constexpr int kStageBytes = 128;
constexpr int kLdsRowBytes = kStageBytes + 16;

// PERSIST: a grid smaller than the tile count; workgroup b takes the tiles of
// the virtual workgroups b, b + gridDim.x, ... (gridDim.x a multiple of 8, so
// that every one of them runs on b's XCD), with a barrier between tiles.
template <typename T, int WM, int WN, int PF, bool STORE, bool PERSIST = false>
__global__ __launch_bounds__(64 * WM * WN) void gemm_packed_sub_kernel(int m, int n, int K,
                                                                       const T* __restrict__ At, int64_t ldak,
                                                                       const T* __restrict__ Bt, int64_t ldbk,
                                                                       float* __restrict__ C, int64_t ldc,
                                                                       int vec4, const float* __restrict__ sa,
                                                                       const float* __restrict__ sb) {
  constexpr int NT = 64 * WM * WN;              // threads
  constexpr int E = 16 / sizeof(T);             // elements per 16-byte chunk
  constexpr int BK = kStageBytes / sizeof(T);   // K per stage
  constexpr int RS = kLdsRowBytes / sizeof(T);  // padded LDS row
  constexpr int TM = kTile / WM / 32;           // MFMA blocks per wave along m
  constexpr int TN = kTile / WN / 32;           // ... along n
  constexpr int CH = kTile * 8 / NT;            // 16-byte chunks per thread and operand
  constexpr int RSTEP = NT / 8;                 // rows covered by one chunk round
  static_assert(PF == 1 || PF == 2, "prefetch distance");
  __shared__ __attribute__((aligned(16))) T As[2][kTile][RS];
  __shared__ __attribute__((aligned(16))) T Bs[2][kTile][RS];

  static_assert(!PERSIST || kGemmGroupRows > 0, "the persistent grid walks the grouped order");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave % WM) * (kTile / WM), wn = (wave / WM) * (kTile / WN);
  for (int vw = static_cast<int>(blockIdx.x);; vw += static_cast<int>(gridDim.x)) {  // virtual workgroup
    int64_t m0, n0;
    if constexpr (kGemmGroupRows > 0) {
      const int mt = (m + kTile - 1) / kTile, nt = (n + kTile - 1) / kTile, tiles = mt * nt;
      const int per = (tiles + 7) / 8;
      if constexpr (PERSIST) {
        if (vw >= 8 * per) return;           // every virtual workgroup done
        if (vw != static_cast<int>(blockIdx.x)) __syncthreads();  // the previous tile's LDS reads are done
      }
      const int v = (vw & 7) * per + (vw >> 3);  // this XCD's run of the grouped order
      if (v >= tiles) {                        // padding of the grid (the whole workgroup)
        if constexpr (PERSIST) continue;
        return;
      }
      const int first = v / (kGemmGroupRows * nt) * kGemmGroupRows;
      const int rows = min(kGemmGroupRows, mt - first);
      const int within = v - first * nt;
      m0 = static_cast<int64_t>(first + within % rows) * kTile;
      n0 = static_cast<int64_t>(within / rows) * kTile;
    } else {
      m0 = static_cast<int64_t>(blockIdx.x) * kTile;
      n0 = static_cast<int64_t>(blockIdx.y) * kTile;
    }
    const int srow = tid >> 3, spart = tid & 7;

    v16f acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.0f;

    // Rows past m (columns past n) of a ragged edge tile load the last valid
    // row (column) instead: they only feed accumulator rows (columns) that the
    // epilogue does not store, and the loads stay branch-free.
    // two register slots, named apart: a slot index that is a run-time value
    // anywhere would put the slots in scratch memory
    uint4 ra0[CH], rb0[CH], ra1[CH], rb1[CH];
    const T* arow[CH];
    const T* brow[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int64_t r = min(m0 + srow + RSTEP * i, static_cast<int64_t>(m) - 1);
      const int64_t c = min(n0 + srow + RSTEP * i, static_cast<int64_t>(n) - 1);
      arow[i] = At + r * ldak + spart * E;
      brow[i] = Bt + c * ldbk + spart * E;
    }
    auto load = [&](auto slot_c, int k0) {
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const uint4 a = *reinterpret_cast<const uint4*>(arow[i] + k0);
        const uint4 b = *reinterpret_cast<const uint4*>(brow[i] + k0);
        if constexpr (decltype(slot_c)::value == 0) {
          ra0[i] = a;
          rb0[i] = b;
        } else {
          ra1[i] = a;
          rb1[i] = b;
        }
      }
    };
    auto store = [&](auto slot_c, int buf) {
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        constexpr bool first = decltype(slot_c)::value == 0;
        *reinterpret_cast<uint4*>(&As[buf][srow + RSTEP * i][spart * E]) = first ? ra0[i] : ra1[i];
        *reinterpret_cast<uint4*>(&Bs[buf][srow + RSTEP * i][spart * E]) = first ? rb0[i] : rb1[i];
      }
    };
    auto compute = [&](int cur) {
      if constexpr (std::is_same<T, float>::value) {
#pragma unroll
        for (int kg = 0; kg < BK; kg += 8) {
          v4f a[TM], b[TN];
#pragma unroll
          for (int t = 0; t < TM; ++t)
            a[t] = *reinterpret_cast<const v4f*>(&As[cur][wm + 32 * t + (lane & 31)][kg + 4 * (lane >> 5)]);
#pragma unroll
          for (int t = 0; t < TN; ++t)
            b[t] = *reinterpret_cast<const v4f*>(&Bs[cur][wn + 32 * t + (lane & 31)][kg + 4 * (lane >> 5)]);
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int ta = 0; ta < TM; ++ta)
#pragma unroll
              for (int tb = 0; tb < TN; ++tb)
                acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ta][j], b[tb][j], acc[ta][tb], 0, 0, 0);
        }
      } else {
        using V8 = typename std::conditional<std::is_same<T, __bf16>::value, v8bf, v8h>::type;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
          V8 a[TM], b[TN];
#pragma unroll
          for (int t = 0; t < TM; ++t)
            a[t] = *reinterpret_cast<const V8*>(&As[cur][wm + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
          for (int t = 0; t < TN; ++t)
            b[t] = *reinterpret_cast<const V8*>(&Bs[cur][wn + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
          for (int ta = 0; ta < TM; ++ta)
#pragma unroll
            for (int tb = 0; tb < TN; ++tb) {
              if constexpr (std::is_same<T, __bf16>::value)
                acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
              else
                acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
            }
        }
      }
    };

    using Slot0 = std::integral_constant<int, 0>;
    using Slot1 = std::integral_constant<int, 1>;
    const int stages = K / BK;
    load(Slot0{}, 0);
    store(Slot0{}, 0);
    if constexpr (PF == 1) {
      __syncthreads();
      for (int s = 0; s < stages; ++s) {
        const int cur = s & 1;
        if (s + 1 < stages) load(Slot0{}, (s + 1) * BK);
        compute(cur);
        if (s + 1 < stages) store(Slot0{}, cur ^ 1);
        __syncthreads();
      }
    } else {
      // register slot (s & 1) holds stage s until it is stored, then stage s + 2;
      // the loop is unrolled by two so that every slot index is a constant
      if (stages > 1) load(Slot1{}, BK);
      __syncthreads();
      auto step = [&](auto cur_c, int s) {
        constexpr int cur = decltype(cur_c)::value;
        if (s + 1 < stages) store(std::integral_constant<int, cur ^ 1>{}, cur ^ 1);  // stage s + 1
        if (s + 2 < stages) load(cur_c, (s + 2) * BK);
        compute(cur);
        __syncthreads();
      };
      for (int s = 0; s < stages; s += 2) {
        step(Slot0{}, s);
        if (s + 1 < stages) step(Slot1{}, s + 1);
      }
    }

    // C -= acc / (sa sb), or C = acc (STORE).  A whole tile with 16-byte C
    // access (the same for every thread of the workgroup) reads half its
    // wave's C blocks at a time, all loads in flight together, into the
    // registers the K loop no longer needs; other tiles take the general path.
    if (!STORE && vec4 && m0 + kTile <= m && n0 + kTile <= n) {
      constexpr int TH = TN >= 4 ? 1 : 4 / TN;  // row blocks per round: 16 loads in flight
      static_assert(TM % TH == 0, "whole rounds of the wave's row blocks");
      // the wave's block scales first (uniform: scalar loads), so that no wait
      // for them falls between the C loads
      const int uw = __builtin_amdgcn_readfirstlane(wave);
      const int64_t um = m0 + (uw % WM) * (kTile / WM), un = n0 + (uw / WM) * (kTile / WN);
      float w[TM][TN];
#pragma unroll
      for (int ta = 0; ta < TM; ++ta)
#pragma unroll
        for (int tb = 0; tb < TN; ++tb) w[ta][tb] = 1.0f;
      if (sa) {
        float ra[TM], rb[TN];
#pragma unroll
        for (int ta = 0; ta < TM; ++ta) ra[ta] = sa[(um + 32 * ta) >> 5];
#pragma unroll
        for (int tb = 0; tb < TN; ++tb) rb[tb] = sb[(un + 32 * tb) >> 5];
#pragma unroll
        for (int ta = 0; ta < TM; ++ta)
#pragma unroll
          for (int tb = 0; tb < TN; ++tb) w[ta][tb] = 1.0f / (ra[ta] * rb[tb]);  // exact: powers of two
      }
      const float* cbase = C + (un + (lane & 31)) * ldc + um + 4 * (lane >> 5);
#pragma unroll
      for (int h = 0; h < TM / TH; ++h) {
        v4f cv[TH][TN][4];
#pragma unroll
        for (int a = 0; a < TH; ++a)
#pragma unroll
          for (int tb = 0; tb < TN; ++tb)
#pragma unroll
            for (int g = 0; g < 4; ++g)
              cv[a][tb][g] = *reinterpret_cast<const v4f*>(cbase + 32 * tb * ldc + 32 * (h * TH + a) + 8 * g);
#pragma unroll
        for (int a = 0; a < TH; ++a)
#pragma unroll
          for (int tb = 0; tb < TN; ++tb) {
            const int ta = h * TH + a;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
              v4f c = cv[a][tb][g];
#pragma unroll
              for (int i = 0; i < 4; ++i) c[i] = fmaf(-w[ta][tb], acc[ta][tb][4 * g + i], c[i]);
              *reinterpret_cast<v4f*>(const_cast<float*>(cbase) + 32 * tb * ldc + 32 * ta + 8 * g) = c;
            }
          }
      }
      if constexpr (PERSIST) continue;
      return;
    }
#pragma unroll
    for (int ta = 0; ta < TM; ++ta)
#pragma unroll
      for (int tb = 0; tb < TN; ++tb) {
        const int64_t col = n0 + wn + 32 * tb + (lane & 31);
        if (col >= n) continue;
        float* cc = C + col * ldc;
        float w = 1.0f;
        if (!STORE && sa) {  // exact: powers of two
          const int64_t rb = min(m0 + wm + 32 * ta, static_cast<int64_t>(m) - 1) >> 5;  // rows past m: unused
          w = 1.0f / (sa[rb] * sb[col >> 5]);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int64_t row = m0 + wm + 32 * ta + 8 * g + 4 * (lane >> 5);
          if (vec4 && row + 3 < m) {
            v4f c;
            if constexpr (STORE) {
              c[0] = acc[ta][tb][4 * g + 0];
              c[1] = acc[ta][tb][4 * g + 1];
              c[2] = acc[ta][tb][4 * g + 2];
              c[3] = acc[ta][tb][4 * g + 3];
            } else {
              c = *reinterpret_cast<const v4f*>(cc + row);
              c[0] = fmaf(-w, acc[ta][tb][4 * g + 0], c[0]);
              c[1] = fmaf(-w, acc[ta][tb][4 * g + 1], c[1]);
              c[2] = fmaf(-w, acc[ta][tb][4 * g + 2], c[2]);
              c[3] = fmaf(-w, acc[ta][tb][4 * g + 3], c[3]);
            }
            *reinterpret_cast<v4f*>(cc + row) = c;
          } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (row + i < m) {
                if constexpr (STORE)
                  cc[row + i] = acc[ta][tb][4 * g + i];
                else
                  cc[row + i] = fmaf(-w, acc[ta][tb][4 * g + i], cc[row + i]);
              }
          }
        }
      }
    if constexpr (!PERSIST) return;
  }
}

// Launch shape per element type: 8 waves of 128x64 (two per SIMD, 128
// accumulators each), loads two stages ahead.  Measured at m = n = 37888,
// K = 1024 (BF16): 565-583 TFLOP/s, against 550 for 4 waves of 128x128
// loading one stage ahead and 475-530 for 256x128 tiles with two workgroups
// per CU (with or without a persistent, staggered schedule).  The C update
// is what holds it there: its read-modify-write moves 8 bytes per element of
// C (11.5 GB at this size, the same order as the MFMA time at peak), and the
// same loop without it runs at about 1.5 PFLOP/s.  The loop itself reads 24
// 16-byte LDS fragments per 32 MFMAs and stores each stage once: per CU and
// stage about as many LDS cycles as MFMA cycles.  In the LU at n = 75 000
// (K = 2048, the C loads of the epilogue batched): BF16 710 TFLOP/s, against
// 603 for 4 waves of 128x128 (one per SIMD, one stage ahead; two stages
// ahead spills).
template <typename T>
struct GemmShape {
  static constexpr int WM = 2, WN = 4, PF = 2;
};

int gemm_stage_k(LuMode mode) { return mode == LuMode::kFP32 ? kStageBytes / 4 : kStageBytes / 2; }

template <typename T>
static void launch_gemm(dim3 grid, bool store, int m, int n, int K, const void* At, int64_t ldak, const void* Bt,
                        int64_t ldbk, float* C, int64_t ldc, int vec4, const float* sa, const float* sb,
                        hipStream_t s, bool persist) {
  using S = GemmShape<T>;
  if (persist)
    hipLaunchKernelGGL((gemm_packed_sub_kernel<T, S::WM, S::WN, S::PF, false, true>), grid, dim3(64 * S::WM * S::WN),
                       0, s, m, n, K, static_cast<const T*>(At), ldak, static_cast<const T*>(Bt), ldbk, C, ldc, vec4,
                       sa, sb);
  else if (store)
    hipLaunchKernelGGL((gemm_packed_sub_kernel<T, S::WM, S::WN, S::PF, true>), grid, dim3(64 * S::WM * S::WN), 0, s,
                       m, n, K, static_cast<const T*>(At), ldak, static_cast<const T*>(Bt), ldbk, C, ldc, vec4, sa,
                       sb);
  else
    hipLaunchKernelGGL((gemm_packed_sub_kernel<T, S::WM, S::WN, S::PF, false>), grid, dim3(64 * S::WM * S::WN), 0,
                       s, m, n, K, static_cast<const T*>(At), ldak, static_cast<const T*>(Bt), ldbk, C, ldc, vec4, sa,
                       sb);
}

hipError_t gemm_packed_sub(LuMode mode, int64_t m, int64_t n, int64_t K, const void* At, int64_t ldak,
                           const void* Bt, int64_t ldbk, float* C, int64_t ldc, hipStream_t s, const float* sa,
                           const float* sb, int max_workgroups) {
  return gemm_packed(mode, false, m, n, K, At, ldak, Bt, ldbk, C, ldc, s, sa, sb, max_workgroups);
}

hipError_t gemm_packed(LuMode mode, bool store, int64_t m, int64_t n, int64_t K, const void* At, int64_t ldak,
                       const void* Bt, int64_t ldbk, float* C, int64_t ldc, hipStream_t s, const float* sa,
                       const float* sb, int max_workgroups) {
  if (m <= 0 || n <= 0 || K <= 0) return hipSuccess;
  const int bk = gemm_stage_k(mode);
  const int esz = mode == LuMode::kFP32 ? 4 : 2;
  // K a whole number of stages; 16-byte aligned operand rows
  if (K % bk != 0 || (ldak * esz) % 16 != 0 || (ldbk * esz) % 16 != 0 ||
      (reinterpret_cast<uintptr_t>(At) & 15) != 0 || (reinterpret_cast<uintptr_t>(Bt) & 15) != 0 ||
      m > INT32_MAX || n > INT32_MAX)
    return hipErrorInvalidValue;
  const int vec4 = ((reinterpret_cast<uintptr_t>(C) & 15) == 0 && (ldc & 3) == 0) ? 1 : 0;
  // tile order: see kGemmGroupRows
  const unsigned mt = static_cast<unsigned>((m + kTile - 1) / kTile), nt = static_cast<unsigned>((n + kTile - 1) / kTile);
  dim3 grid = kGemmGroupRows > 0 ? dim3((mt * nt + 7) / 8 * 8) : dim3(mt, nt);
  // a cap below the tile count (subtracting form): the persistent form on a
  // whole number of workgroups per XCD
  const int cap = max_workgroups / 8 * 8;
  const bool persist = !store && kGemmGroupRows > 0 && cap > 0 && grid.x > static_cast<unsigned>(cap);
  if (persist) grid = dim3(static_cast<unsigned>(cap));
  const int mi = static_cast<int>(m), ni = static_cast<int>(n), ki = static_cast<int>(K);
  if ((sa == nullptr) != (sb == nullptr)) return hipErrorInvalidValue;
  switch (mode) {
    case LuMode::kFP32:
      launch_gemm<float>(grid, store, mi, ni, ki, At, ldak, Bt, ldbk, C, ldc, vec4, sa, sb, s, persist);
      break;
    case LuMode::kBF16:
      launch_gemm<__bf16>(grid, store, mi, ni, ki, At, ldak, Bt, ldbk, C, ldc, vec4, sa, sb, s, persist);
      break;
    case LuMode::kFP16:
      launch_gemm<_Float16>(grid, store, mi, ni, ki, At, ldak, Bt, ldbk, C, ldc, vec4, sa, sb, s, persist);
      break;
  }
  return hipGetLastError();
}

}  // namespace rf
