// dmsim.hip -- CDNA4 (gfx950) kernels and the C ABI of the Pauli-basis
// density-matrix engine.  See include/dmsim.h for the contract.
//
// Layout: a flat vector of N = 4^n coefficients; the 2-bit digit at
// "position" p = n-1-q of the flat index is the Pauli component of qubit q
// (stride 4^p).  All kernels update the vector in place; every element is
// read and written by exactly one thread (or one workgroup for tiles), so
// in-place updates are race free.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "dmsim.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxQubits = 20;
constexpr int kFamilies = 10;  // profiling slots, see dm_profile_read

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(DM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
  } while (0)

// ---------------------------------------------------------------------------
// Small vector helpers
// ---------------------------------------------------------------------------

template <typename T, int W>
struct alignas(W * sizeof(T)) Pack {
  T v[W];
};

template <typename P, typename T>
__device__ __forceinline__ P ldp(const T *p) {
  return *reinterpret_cast<const P *>(p);
}
template <typename P, typename T>
__device__ __forceinline__ void stp(T *p, const P &x) {
  *reinterpret_cast<P *>(p) = x;
}

// 16-byte loads / stores, with the nontemporal hint when NT (streamed data
// not re-read before eviction; measured faster on gfx950 for strides of
// >= 16 coefficients, slower for stride 4).
template <typename T, bool NT>
__device__ __forceinline__ Pack<T, 16 / sizeof(T)> ld16(const T *p) {
  using P = Pack<T, 16 / sizeof(T)>;
  if (!NT) return ldp<P>(p);
  using V = T __attribute__((ext_vector_type(16 / sizeof(T))));
  const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
  P r;
#pragma unroll
  for (int e = 0; e < (int)(16 / sizeof(T)); ++e) r.v[e] = v[e];
  return r;
}
template <typename T, bool NT>
__device__ __forceinline__ void st16(T *p, const Pack<T, 16 / sizeof(T)> &x) {
  if (!NT) {
    stp(p, x);
    return;
  }
  using V = T __attribute__((ext_vector_type(16 / sizeof(T))));
  V v;
#pragma unroll
  for (int e = 0; e < (int)(16 / sizeof(T)); ++e) v[e] = x.v[e];
  __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
}
// Insert a zero 2-bit digit at each of the ascending positions pos[0..m-1].
__device__ __forceinline__ uint64_t insert_digits(uint64_t r, const int *pos, int m) {
  for (int j = 0; j < m; ++j) {
    const int s = 2 * pos[j];
    const uint64_t low = r & ((1ull << s) - 1ull);
    r = ((r >> s) << (s + 2)) | low;
  }
  return r;
}

__host__ __device__ __forceinline__ uint64_t pow4(int p) { return 1ull << (2 * p); }

// ---------------------------------------------------------------------------
// K1: single-qubit PTM (rows X,Y,Z), any stride.
// ---------------------------------------------------------------------------

template <typename T>
struct Rows3 {
  T r[3][4];
};

template <typename T>
Rows3<T> make_rows3(const double *R) {
  Rows3<T> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) m.r[i][j] = (T)R[4 * i + j];
  return m;
}

// Stride S = 4^p >= W: each unit covers W consecutive low indices of one
// (hi) slab; the four components are four coalesced W-vectors.
template <typename T, bool USE_I, bool NT>
__global__ __launch_bounds__(kBlock) void k_1q_wide(T *__restrict__ a, uint64_t units, int sh,
                                                    uint64_t S, Rows3<T> m) {
  constexpr int W = 16 / sizeof(T);
  using P = Pack<T, W>;
  const uint64_t mask = (1ull << sh) - 1ull;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < units; u += step) {
    T *b = a + ((u >> sh) * 4ull) * S + (u & mask) * W;
    P x0;
    if (USE_I) x0 = ld16<T, NT>(b);
    const P x1 = ld16<T, NT>(b + S), x2 = ld16<T, NT>(b + 2 * S), x3 = ld16<T, NT>(b + 3 * S);
    P y1, y2, y3;
#pragma unroll
    for (int e = 0; e < W; ++e) {
      T s1 = m.r[0][1] * x1.v[e] + m.r[0][2] * x2.v[e] + m.r[0][3] * x3.v[e];
      T s2 = m.r[1][1] * x1.v[e] + m.r[1][2] * x2.v[e] + m.r[1][3] * x3.v[e];
      T s3 = m.r[2][1] * x1.v[e] + m.r[2][2] * x2.v[e] + m.r[2][3] * x3.v[e];
      if (USE_I) {
        s1 += m.r[0][0] * x0.v[e];
        s2 += m.r[1][0] * x0.v[e];
        s3 += m.r[2][0] * x0.v[e];
      }
      y1.v[e] = s1;
      y2.v[e] = s2;
      y3.v[e] = s3;
    }
    st16<T, NT>(b + S, y1);
    st16<T, NT>(b + 2 * S, y2);
    st16<T, NT>(b + 3 * S, y3);
  }
}

// Stride 1 (the last qubit): the four components of a group are adjacent.
template <typename T, bool USE_I>
__global__ __launch_bounds__(kBlock) void k_1q_p0(T *__restrict__ a, uint64_t groups, Rows3<T> m) {
  using P = Pack<T, 4>;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += step) {
    P x = ldp<P>(a + 4 * g);
    P y;
    y.v[0] = x.v[0];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T s = m.r[r][1] * x.v[1] + m.r[r][2] * x.v[2] + m.r[r][3] * x.v[3];
      if (USE_I) s += m.r[r][0] * x.v[0];
      y.v[r + 1] = s;
    }
    stp(a + 4 * g, y);
  }
}

// ---------------------------------------------------------------------------
// K2: CX as a signed permutation of the 16 (control, target) pairs, with an
// optional per-pair scale.  Pairs are addressed in position order (hi, lo);
// CTRL_HI says whether the control qubit sits at the higher position.
// ---------------------------------------------------------------------------

template <typename T>
struct Scale16 {
  T s[16];
};

// Output pair (c,t) <- sign * input pair src(c,t), for the 12 moving pairs.
// (0,2)<-(3,2) (3,2)<-(0,2) (0,3)<-(3,3) (3,3)<-(0,3) (1,0)<-(1,1) (1,1)<-(1,0)
// (1,2)<-(2,3) (1,3)<--(2,2) (2,0)<-(2,1) (2,1)<-(2,0) (2,2)<--(1,3) (2,3)<-(1,2)
template <typename V, typename T, bool CTRL_HI, bool SCALED>
__device__ __forceinline__ void cx_map(V (&x)[4][4], const Scale16<T> &sc) {
  auto X = [&](int c, int t) -> V & { return CTRL_HI ? x[c][t] : x[t][c]; };
  auto S = [&](int c, int t) -> T { return SCALED ? sc.s[4 * c + t] : T(1); };
  V t02 = X(0, 2), t32 = X(3, 2), t03 = X(0, 3), t33 = X(3, 3);
  V t10 = X(1, 0), t11 = X(1, 1), t12 = X(1, 2), t13 = X(1, 3);
  V t20 = X(2, 0), t21 = X(2, 1), t22 = X(2, 2), t23 = X(2, 3);
  X(0, 2) = t32 * S(0, 2);
  X(3, 2) = t02 * S(3, 2);
  X(0, 3) = t33 * S(0, 3);
  X(3, 3) = t03 * S(3, 3);
  X(1, 0) = t11 * S(1, 0);
  X(1, 1) = t10 * S(1, 1);
  X(1, 2) = t23 * S(1, 2);
  X(1, 3) = -t22 * S(1, 3);
  X(2, 0) = t21 * S(2, 0);
  X(2, 1) = t20 * S(2, 1);
  X(2, 2) = -t13 * S(2, 2);
  X(2, 3) = t12 * S(2, 3);
  if (SCALED) {
    X(0, 1) = X(0, 1) * S(0, 1);
    X(3, 0) = X(3, 0) * S(3, 0);
    X(3, 1) = X(3, 1) * S(3, 1);
  }
}

// Whether pair (i_hi, i_lo) in position order is written by the CX map.
__host__ __device__ constexpr bool cx_writes(bool ctrl_hi, bool scaled, int ih, int il) {
  return (ih == 0 && il == 0) ? false
         : scaled             ? true
         : ctrl_hi            ? !((ih == 0 && il == 1) || (ih == 3 && il == 0) || (ih == 3 && il == 1))
                              : !((il == 0 && ih == 1) || (il == 3 && ih == 0) || (il == 3 && ih == 1));
}

// Low position >= 1 (stride >= W): units of W consecutive low indices.
template <typename T, bool CTRL_HI, bool SCALED, bool NT>
__global__ __launch_bounds__(kBlock) void k_cx_wide(T *__restrict__ a, uint64_t units, int sh_lo,
                                                    int sh_mid, uint64_t Sh, uint64_t Sl,
                                                    Scale16<T> sc) {
  constexpr int W = 16 / sizeof(T);
  using P = Pack<T, W>;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < units; u += step) {
    const uint64_t lo = u & ((1ull << sh_lo) - 1ull);
    const uint64_t rest = u >> sh_lo;
    const uint64_t mid = rest & ((1ull << sh_mid) - 1ull);
    const uint64_t hi = rest >> sh_mid;
    T *b = a + hi * 4ull * Sh + mid * 4ull * Sl + lo * W;
    P x[4][4];
#pragma unroll
    for (int ih = 0; ih < 4; ++ih)
#pragma unroll
      for (int il = 0; il < 4; ++il)
        if (cx_writes(CTRL_HI, SCALED, ih, il)) x[ih][il] = ld16<T, NT>(b + ih * Sh + il * Sl);
    // per-lane permutation
    P y[4][4];
#pragma unroll
    for (int e = 0; e < W; ++e) {
      T v[4][4];
#pragma unroll
      for (int ih = 0; ih < 4; ++ih)
#pragma unroll
        for (int il = 0; il < 4; ++il) v[ih][il] = cx_writes(CTRL_HI, SCALED, ih, il) ? x[ih][il].v[e] : T(0);
      cx_map<T, T, CTRL_HI, SCALED>(v, sc);
#pragma unroll
      for (int ih = 0; ih < 4; ++ih)
#pragma unroll
        for (int il = 0; il < 4; ++il) y[ih][il].v[e] = v[ih][il];
    }
#pragma unroll
    for (int ih = 0; ih < 4; ++ih)
#pragma unroll
      for (int il = 0; il < 4; ++il)
        if (cx_writes(CTRL_HI, SCALED, ih, il)) st16<T, NT>(b + ih * Sh + il * Sl, y[ih][il]);
  }
}

// Low position 0: each of the four high components is a contiguous quad.
template <typename T, bool CTRL_HI, bool SCALED>
__global__ __launch_bounds__(kBlock) void k_cx_p0(T *__restrict__ a, uint64_t groups, int sh_mid,
                                                  uint64_t Sh, Scale16<T> sc) {
  using P = Pack<T, 4>;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += step) {
    const uint64_t mid = g & ((1ull << sh_mid) - 1ull);
    const uint64_t hi = g >> sh_mid;
    T *b = a + hi * 4ull * Sh + mid * 4ull;
    T v[4][4];
#pragma unroll
    for (int ih = 0; ih < 4; ++ih) {
      P q = ldp<P>(b + ih * Sh);
#pragma unroll
      for (int il = 0; il < 4; ++il) v[ih][il] = q.v[il];
    }
    cx_map<T, T, CTRL_HI, SCALED>(v, sc);
#pragma unroll
    for (int ih = 0; ih < 4; ++ih) {
      bool any = false;
#pragma unroll
      for (int il = 0; il < 4; ++il) any = any || cx_writes(CTRL_HI, SCALED, ih, il);
      if (any) {
        P q;
#pragma unroll
        for (int il = 0; il < 4; ++il) q.v[il] = v[ih][il];
        stp(b + ih * Sh, q);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K2': dense two-qubit PTM (16x16), pairs in position order (hi, lo).
// ---------------------------------------------------------------------------

template <typename T>
struct Mat16 {
  T m[16][16];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_2q_dense(T *__restrict__ a, uint64_t groups, int sh_lo,
                                                     int sh_mid, uint64_t Sh, uint64_t Sl,
                                                     Mat16<T> R, uint32_t rowmask) {
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < groups; u += step) {
    const uint64_t lo = u & ((1ull << sh_lo) - 1ull);
    const uint64_t rest = u >> sh_lo;
    const uint64_t mid = rest & ((1ull << sh_mid) - 1ull);
    const uint64_t hi = rest >> sh_mid;
    T *b = a + hi * 4ull * Sh + mid * 4ull * Sl + lo;
    T x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = b[(j >> 2) * Sh + (j & 3) * Sl];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (rowmask & (1u << r)) {
        T s = T(0);
#pragma unroll
        for (int c = 0; c < 16; ++c) s += R.m[r][c] * x[c];
        b[(r >> 2) * Sh + (r & 3) * Sl] = s;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K3: register tile of M qubits (positions pos[0..M-1], ascending): each
// thread owns the 4^M coefficients that differ only on those qubits and
// applies one 4x4 map per tile qubit.  Used for memory noise with a thermal
// feed (x,y <- h x,h y; z <- g z + d a_I) where the update is not diagonal.
// ---------------------------------------------------------------------------

template <typename T, int M>
struct TileMaps {
  int pos[M];
  T A[M][4][4];
};

template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_tile_maps(T *__restrict__ a, uint64_t units,
                                                      TileMaps<T, M> tm) {
  constexpr int E = 1 << (2 * M);
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  uint64_t S[M];
#pragma unroll
  for (int j = 0; j < M; ++j) S[j] = pow4(tm.pos[j]);
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < units; u += step) {
    const uint64_t base = insert_digits(u, tm.pos, M);
    T x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint64_t off = 0;
#pragma unroll
      for (int j = 0; j < M; ++j) off += (uint64_t)((e >> (2 * j)) & 3) * S[j];
      x[e] = a[base + off];
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (((e >> (2 * j)) & 3) != 0) continue;
        const int st = 1 << (2 * j);
        T v0 = x[e], v1 = x[e + st], v2 = x[e + 2 * st], v3 = x[e + 3 * st];
        x[e + st] = tm.A[j][1][0] * v0 + tm.A[j][1][1] * v1 + tm.A[j][1][2] * v2 + tm.A[j][1][3] * v3;
        x[e + 2 * st] = tm.A[j][2][0] * v0 + tm.A[j][2][1] * v1 + tm.A[j][2][2] * v2 + tm.A[j][2][3] * v3;
        x[e + 3 * st] = tm.A[j][3][0] * v0 + tm.A[j][3][1] * v1 + tm.A[j][3][2] * v2 + tm.A[j][3][3] * v3;
      }
    }
#pragma unroll
    for (int e = 1; e < E; ++e) {
      uint64_t off = 0;
#pragma unroll
      for (int j = 0; j < M; ++j) off += (uint64_t)((e >> (2 * j)) & 3) * S[j];
      a[base + off] = x[e];
    }
  }
}

// ---------------------------------------------------------------------------
// K4: diagonal scalings.
//   k_diag_tab:  a_i *= prod_p tab[p][digit_p(i)]       (per-qubit tables)
//   k_diag_pow:  a_i *= hpow[#X/Y digits] * gpow[#Z digits] (same map on all)
//   k_product:   a_i  = scale * prod_p vec[p][digit_p(i)]  (product states)
// ---------------------------------------------------------------------------

struct Tab4 {
  double t[kMaxQubits * 4];
};

template <typename T, bool WRITE_ONLY>
__global__ __launch_bounds__(kBlock) void k_diag_tab(T *__restrict__ a, uint64_t units, int n, Tab4 tab,
                                                     double scale) {
  constexpr int W = 16 / sizeof(T);
  using P = Pack<T, W>;
  __shared__ double s_tab[kMaxQubits * 4];
  for (int i = threadIdx.x; i < 4 * n; i += kBlock) s_tab[i] = tab.t[i];
  __syncthreads();
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < units; u += step) {
    P x;
    if (!WRITE_ONLY) x = ldp<P>(a + u * W);
#pragma unroll
    for (int e = 0; e < W; ++e) {
      const uint64_t i = u * W + e;
      double f = scale;
      for (int p = 0; p < n; ++p) f *= s_tab[4 * p + ((i >> (2 * p)) & 3)];
      x.v[e] = WRITE_ONLY ? (T)f : (T)(x.v[e] * f);
    }
    stp(a + u * W, x);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_pow(T *__restrict__ a, uint64_t units, int n,
                                                     double h, double g) {
  constexpr int W = 16 / sizeof(T);
  using P = Pack<T, W>;
  __shared__ T s_h[kMaxQubits + 1], s_g[kMaxQubits + 1];
  if (threadIdx.x == 0) {
    double ph = 1.0, pg = 1.0;
    for (int k = 0; k <= n; ++k) {
      s_h[k] = (T)ph;
      s_g[k] = (T)pg;
      ph *= h;
      pg *= g;
    }
  }
  __syncthreads();
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  const uint64_t m55 = 0x5555555555555555ull;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < units; u += step) {
    P x = ldp<P>(a + u * W);
#pragma unroll
    for (int e = 0; e < W; ++e) {
      const uint64_t i = u * W + e;
      const uint64_t lo = i & m55, hi = (i >> 1) & m55;
      const int nz = __popcll(lo & hi);
      const int nxy = __popcll(lo ^ hi);
      x.v[e] = x.v[e] * (s_h[nxy] * s_g[nz]);
    }
    stp(a + u * W, x);
  }
}

// ---------------------------------------------------------------------------
// K5: fused k-qubit block on the matrix cores.
//
// The block's 4^K components (rows) times W contiguous "column" indices form
// a tile; C = S * B with S the 4^K x 4^K block PTM.  The tile rows are the
// digits at positions P[0..K-1]; the columns are the lowest L positions not
// in P (4^L = W); the remaining positions enumerate tiles.  B is staged in
// LDS; for K = 3 S^T is staged in LDS too, for K = 4 (256x256) it is read
// through the L2.  v_mfma_*_16x16x4 operand layout (per lane l of a wave):
//   A[16x4]: A[l%16][l/16]   B[4x16]: B[l/16][l%16]
//   D[16x16] in 4 registers: D[MfmaT<T>::row(l/16, i)][l%16]
// ---------------------------------------------------------------------------

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaT;
// Accumulator row of register v in lane group lk = lane/16 (column lane%16),
// as measured on gfx950: f32 holds rows 4*lk+v, f64 holds rows lk+4*v.
template <>
struct MfmaT<double> {
  typedef dx4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct MfmaT<float> {
  typedef fx4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

struct BlockGeom {
  int n;
  int pos[4];       // block positions, ascending (digit j of the row index)
  int lpos[4];      // column positions, ascending
  int rpos[kMaxQubits];  // tile-enumerating positions, ascending
  int nr;
  uint64_t tiles;
};

template <typename T, int K, int L>
__global__ __launch_bounds__(kBlock) void k_block_mfma(T *__restrict__ a, const T *__restrict__ ST,
                                                       BlockGeom geo) {
  constexpr int ROWS = 1 << (2 * K);
  constexpr int W = 1 << (2 * L);
  constexpr int PAD = (sizeof(T) == 4 && (W % 32) == 0) ? 16 : 0;
  constexpr int LDW = W + PAD;
  constexpr bool ST_IN_LDS = (K <= 3);
  constexpr int LDR = ROWS + ((sizeof(T) == 4 && (ROWS % 32) == 0) ? 16 : 0);
  using MT = MfmaT<T>;
  using Acc = typename MT::acc;

  __shared__ T s_b[ROWS * LDW];
  __shared__ T s_st[ST_IN_LDS ? ROWS * LDR : 1];
  __shared__ uint64_t s_offr[ROWS];
  __shared__ uint64_t s_offc[W];

  const int tid = threadIdx.x;
  for (int r = tid; r < ROWS; r += kBlock) {
    uint64_t o = 0;
    for (int j = 0; j < K; ++j) o += (uint64_t)((r >> (2 * j)) & 3) << (2 * geo.pos[j]);
    s_offr[r] = o;
  }
  for (int c = tid; c < W; c += kBlock) {
    uint64_t o = 0;
    for (int j = 0; j < L; ++j) o += (uint64_t)((c >> (2 * j)) & 3) << (2 * geo.lpos[j]);
    s_offc[c] = o;
  }
  if (ST_IN_LDS) {
    for (int i = tid; i < ROWS * ROWS; i += kBlock) s_st[(i / ROWS) * LDR + (i % ROWS)] = ST[i];
  }
  __syncthreads();

  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int lr = lane & 15;   // row (A) / column (B, D) within a 16-block
  const int lk = lane >> 4;   // k index (A, B) / row quad (D)

  // Work split per wave: K=3 -> 16 row-blocks? no: 4 row-blocks x (W/16) col-blocks.
  constexpr int RB = ROWS / 16;          // row blocks
  constexpr int CB = W / 16;             // column blocks
  constexpr int NB = RB * CB;            // 16x16 output blocks per tile
  constexpr int PER_WAVE = NB / 4;       // blocks per wave (4 waves)
  static_assert(NB % 4 == 0, "tile must split over 4 waves");

  for (uint64_t t = blockIdx.x; t < geo.tiles; t += gridDim.x) {
    // tile base: deposit t's digits on the enumerating positions
    uint64_t base = 0;
    {
      uint64_t r = t;
      for (int j = 0; j < geo.nr; ++j) {
        base += (r & 3ull) << (2 * geo.rpos[j]);
        r >>= 2;
      }
    }
    // stage B
    for (int e = tid; e < ROWS * W; e += kBlock) {
      const int r = e / W, c = e % W;
      s_b[r * LDW + c] = a[base + s_offr[r] + s_offc[c]];
    }
    __syncthreads();

    Acc acc[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) acc[i] = Acc{0, 0, 0, 0};

    // output block index ob = wave * PER_WAVE + i -> (rb, cb) row-major over CB
#pragma unroll 4
    for (int k0 = 0; k0 < ROWS; k0 += 4) {
      const int kk = k0 + lk;
#pragma unroll
      for (int i = 0; i < PER_WAVE; ++i) {
        const int ob = wave * PER_WAVE + i;
        const int rb = ob / CB, cb = ob % CB;
        const T av = ST_IN_LDS ? s_st[kk * LDR + rb * 16 + lr] : ST[(uint64_t)kk * ROWS + rb * 16 + lr];
        const T bv = s_b[kk * LDW + cb * 16 + lr];
        acc[i] = MT::mma(av, bv, acc[i]);
      }
    }
    // write D straight to HBM
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
      const int ob = wave * PER_WAVE + i;
      const int rb = ob / CB, cb = ob % CB;
      const uint64_t oc = s_offc[cb * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = rb * 16 + MT::row(lk, j);
        a[base + s_offr[row] + oc] = acc[i][j];
      }
    }
    __syncthreads();  // s_b is restaged by the next tile
  }
}

// ---------------------------------------------------------------------------
// K5': tile program.  One workgroup stages a tile of 4^M coefficients -- the
// digits at positions tpos[0..M-1] (ascending; tpos[j] == j for j < w, so
// runs of 4^w coefficients are contiguous in HBM) with the other positions
// fixed -- in LDS, applies a short list of operations to it and writes it
// back: one HBM round trip for a whole group of gates.  Operations address
// tile digits:
//   TOP_1Q     digit d0, 12 values (rows X,Y,Z of a 4x4 PTM)
//   TOP_CX     control digit d0, target digit d1, k=1: 16 per-pair scales
//   TOP_2Q     digits d0, d1, 256 values (16x16 PTM, pair index 4 i_d0 + i_d1)
//   TOP_DENSE  k = 3 or 4 digits d0..d{k-1}: 4^k x 4^k PTM on the matrix
//              cores, values stored transposed (ST[col][row]); the other
//              tile digits index the 4^(M-k) columns.
//   TOP_STAGE  digits d0 < d1 and k members, the k records that follow it
//              (TOP_1Q on d0 or d1, TOP_CX or TOP_2Q on {d0, d1} in either
//              order), applied in order to each group of 16 coefficients in
//              registers: one LDS round trip for all of them.
// The kernels run TOP_STAGE and TOP_DENSE records; the host turns every
// other operation of a pass into a stage member (group_stages), marks the
// single-qubit members whose PTM is unital (first column (0, 0, 0): no term
// in the identity component) as TOP_1QU, and gives every CX member its 16
// scales (ones for a plain CX).
// ---------------------------------------------------------------------------

enum { TOP_1Q = 1, TOP_CX = 2, TOP_2Q = 3, TOP_DENSE = 4, TOP_STAGE = 6, TOP_1QU = 7 };

struct TileOp {
  int32_t kind, k, d[4], param;
  // Swizzled passes: the order in which a thread's group index g enumerates
  // the tile digits the operation leaves free (3 bits per digit, first digit
  // in bits 0-2); the first ones cover distinct bank fields (lds_swz).
  int32_t lanes;
  // FP32 stages on digits >= 1 with 8-byte LDS accesses along digit 0: the
  // order of the free digits other than digit 0, the first two covering the
  // bank fields of LDS index bits 2-3 and 4-5.
  int32_t qlanes;
};

struct TileGeom {
  int m;                  // tile digits (== template M)
  int w;                  // leading digits with tpos[j] == j
  int tpos[8];            // HBM position of each tile digit (ascending)
  int nr;
  int rpos[kMaxQubits];
  uint64_t tiles;
  // ident: tile index i sits at LDS index i.  Otherwise the LDS index is
  // lds_swz(i), and the operations enumerate their elements in the order of
  // TileOp::lanes.
  int ident;
};

// Swizzled LDS index.  The 64 banks of 4 bytes are LDS index bits 0-5 (tile
// digits 0-2); bits 2-3 and 4-5 take the XOR of tile digits 3 and 4, so the
// bank fields are [d0, d1 ^ d3, d2 ^ d4].  An operation on digit 1 or 2
// still has a free digit (3 or 4) that sweeps the field it occupies, and
// the quads of tile digit 0 stay contiguous.  A bijection on [0, 4^M) for
// M >= 5 that only permutes bits 2-5.
template <bool SWZ>
__device__ __forceinline__ int lds_swz(int i) {
  return SWZ ? (i ^ (((i >> 6) & 15) << 2)) : i;
}

// Tile index with the base-4 digits of g placed at the tile digits listed in
// ord (3 bits each), the operation's own digits left zero.
__device__ __forceinline__ int lane_deposit(int g, int ord, int nf) {
  int b = 0;
  for (int k = 0; k < nf; ++k) b |= ((g >> (2 * k)) & 3) << (2 * ((ord >> (3 * k)) & 7));
  return b;
}

// A quad (4 consecutive coefficients: 16 bytes in FP32, 32 in FP64) from or
// to HBM, with the nontemporal hint when NT.
template <typename T, bool NT>
__device__ __forceinline__ Pack<T, 4> ldq(const T *p) {
  constexpr int H = 16 / sizeof(T);
  Pack<T, 4> q;
#pragma unroll
  for (int h = 0; h < 4 / H; ++h) {
    const Pack<T, H> v = ld16<T, NT>(p + h * H);
#pragma unroll
    for (int e = 0; e < H; ++e) q.v[h * H + e] = v.v[e];
  }
  return q;
}
template <typename T, bool NT>
__device__ __forceinline__ void stq(T *p, const Pack<T, 4> &q) {
  constexpr int H = 16 / sizeof(T);
#pragma unroll
  for (int h = 0; h < 4 / H; ++h) {
    Pack<T, H> v;
#pragma unroll
    for (int e = 0; e < H; ++e) v.v[e] = q.v[h * H + e];
    st16<T, NT>(p + h * H, v);
  }
}

template <typename T>
__device__ __forceinline__ void lds_put_quad(T *s, int lb, const Pack<T, 4> &q) {
  *reinterpret_cast<Pack<T, 4> *>(&s[lb]) = q;
}
template <typename T>
__device__ __forceinline__ Pack<T, 4> lds_get_quad(const T *s, int lb) {
  return *reinterpret_cast<const Pack<T, 4> *>(&s[lb]);
}

// LDS accesses at a byte offset.  A stage addresses a coefficient as the
// XOR of a per-thread offset and a uniform one (scalar registers); in bytes
// that is one vector instruction per access instead of two.
template <typename V, typename T>
__device__ __forceinline__ V lds_ld(const T *s, uint32_t off) {
  return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(s) + off);
}
template <typename V, typename T>
__device__ __forceinline__ void lds_st(T *s, uint32_t off, const V &x) {
  *reinterpret_cast<V *>(reinterpret_cast<char *>(s) + off) = x;
}

__device__ __forceinline__ int ins_zero(int g, int j) {
  const int lo = g & ((1 << (2 * j)) - 1);
  return ((g >> (2 * j)) << (2 * j + 2)) | lo;
}

template <typename T, int M>
__device__ __forceinline__ void tile_dense(T *s, const TileOp &op, const T *__restrict__ ST) {
  // rows: digits op.d[0..k-1] (row digit j -> tile digit d[j]); cols: the
  // remaining tile digits in ascending order.
  using MT = MfmaT<T>;
  using Acc = typename MT::acc;
  const int k = op.k;
  const int rows = 1 << (2 * k);
  const int cols = 1 << (2 * (M - k));
  int cdig[8];
  {
    int c = 0;
    for (int j = 0; j < M; ++j) {
      bool inrow = false;
      for (int r = 0; r < k; ++r) inrow = inrow || (op.d[r] == j);
      if (!inrow) cdig[c++] = j;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int CB = cols / 16;
  const int NB = (rows / 16) * CB;
  const int per_wave = NB / 4;  // 4 for (M,k) = (6,3), (6,4)
  Acc acc[4];
  int rb_[4], cb_[4], coff[4];
  for (int i = 0; i < 4; ++i) {
    acc[i] = Acc{0, 0, 0, 0};
    const int ob = wave * per_wave + i;
    rb_[i] = ob / CB;
    cb_[i] = ob % CB;
    const int col = cb_[i] * 16 + lr;
    int o = 0;
    for (int j = 0; j < M - k; ++j) o |= ((col >> (2 * j)) & 3) << (2 * cdig[j]);
    coff[i] = o;
  }
  for (int k0 = 0; k0 < rows; k0 += 4) {
    const int kk = k0 + lk;
    int roff = 0;
    for (int j = 0; j < k; ++j) roff |= ((kk >> (2 * j)) & 3) << (2 * op.d[j]);
    for (int i = 0; i < per_wave; ++i) {
      const T av = ST[(size_t)kk * rows + rb_[i] * 16 + lr];
      const T bv = s[roff + coff[i]];
      acc[i] = MT::mma(av, bv, acc[i]);
    }
  }
  __syncthreads();  // every wave has read its B fragments
  for (int i = 0; i < per_wave; ++i) {
    for (int v = 0; v < 4; ++v) {
      const int row = rb_[i] * 16 + MT::row(lk, v);
      int roff = 0;
      for (int j = 0; j < k; ++j) roff |= ((row >> (2 * j)) & 3) << (2 * op.d[j]);
      s[roff + coff[i]] = acc[i][v];
    }
  }
}

template <typename T, int M>
__device__ __forceinline__ uint64_t tile_offset(int t, int w, int lomask, const TileGeom &geo) {
  uint64_t off = (uint64_t)(t & lomask);
  for (int j = w; j < M; ++j) off |= (uint64_t)((t >> (2 * j)) & 3) << (2 * geo.tpos[j]);
  return off;
}

// a * b + c with one rounding (v_fma_f32 / v_fma_f64).  The library is
// compiled without floating-point contraction (its host planner must match
// the Python planner bit for bit), so the register stages ask for it.
template <typename T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// Single-qubit member on stage digit a (LO) or b: rows X, Y, Z of the PTM;
// AFF adds the identity-component column (non-unital maps).
template <typename T, int NG, bool LO, bool AFF>
__device__ __forceinline__ void member_1q(T (&v)[NG][4][4], const T *__restrict__ pv) {
  T r[3][4];
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) r[x][y] = pv[4 * x + y];
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const T x0 = LO ? v[g][0][y] : v[g][y][0], x1 = LO ? v[g][1][y] : v[g][y][1];
      const T x2 = LO ? v[g][2][y] : v[g][y][2], x3 = LO ? v[g][3][y] : v[g][y][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T o = fmad(r[c][3], x3, fmad(r[c][2], x2, AFF ? fmad(r[c][1], x1, r[c][0] * x0) : r[c][1] * x1));
        if (LO) v[g][c + 1][y] = o;
        else v[g][y][c + 1] = o;
      }
    }
}

// CX member, control digit a (LO: pair (c, t) is v[c][t]) or b (v[t][c]),
// with its 16 per-pair scales.
template <typename T, int NG, bool LO>
__device__ __forceinline__ void member_cx(T (&v)[NG][4][4], const T *__restrict__ pv) {
  Scale16<T> sv;
#pragma unroll
  for (int u = 0; u < 16; ++u) sv.s[u] = pv[u];
#pragma unroll
  for (int g = 0; g < NG; ++g) cx_map<T, T, LO, true>(v[g], sv);
}

// The member operations of a stage on NG groups of 16 coefficients in
// registers: v[g][x][y], x the component index along stage digit a (the
// lower one), y along the other digit.
template <typename T, bool GEN2Q, int NG>
__device__ __forceinline__ void stage_apply(T (&v)[NG][4][4], int a, const TileOp *__restrict__ mem, int nm,
                                            const T *__restrict__ prm) {
  for (int q = 0; q < nm; ++q) {
    const TileOp m = mem[q];
    const T *pv = prm + m.param;
    const bool lo = m.d[0] == a;
    if (m.kind == TOP_1QU) {
      if (lo) member_1q<T, NG, true, false>(v, pv);
      else member_1q<T, NG, false, false>(v, pv);
    } else if (m.kind == TOP_1Q) {
      if (lo) member_1q<T, NG, true, true>(v, pv);
      else member_1q<T, NG, false, true>(v, pv);
    } else if (!GEN2Q || m.kind == TOP_CX) {
      if (lo) member_cx<T, NG, true>(v, pv);
      else member_cx<T, NG, false>(v, pv);
    } else {
      // 16x16 PTM, pair index 4 i_d0 + i_d1 (first row the identity's)
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        T w[16];
#pragma unroll
        for (int r = 1; r < 16; ++r) {
          T acc = T(0);
#pragma unroll
          for (int c = 0; c < 16; ++c) acc = fmad(pv[16 * r + c], lo ? v[g][c >> 2][c & 3] : v[g][c & 3][c >> 2], acc);
          w[r] = acc;
        }
#pragma unroll
        for (int r = 1; r < 16; ++r) {
          if (lo) v[g][r >> 2][r & 3] = w[r];
          else v[g][r & 3][r >> 2] = w[r];
        }
      }
    }
  }
}

// The operations of a tile program on the LDS tile s (all threads of the
// workgroup take part; a barrier follows each operation).  Every operation
// is a stage (TOP_STAGE, digits ja < jb) or, in DENSE variants, a matrix-core
// PTM.  A stage reads each group of 16 coefficients (the 4 x 4 components
// along ja and jb) once, applies its members in registers and writes the 15
// that can change (the (I, I) component never does).  Access forms:
//  - ja = 0: one quad (16 bytes along digit 0) per component of jb;
//  - VEC, FP32, ja >= 1, at least 32 coefficients per thread (7-digit tiles
//    of 512 threads): the two groups that differ in the low bit of digit 0,
//    one 8-byte access per component pair;
//  - otherwise one 4- or 8-byte access per coefficient.
// SWZ: the tile is stored at lds_swz of each tile index, and a stage
// enumerates its groups in the digit order of TileOp::lanes (quad and
// element forms) or TileOp::qlanes (8-byte form), so that the lanes of a
// wave fall on distinct banks.  VEC is off in the 64-register one-tile
// kernel of 6-digit tiles (measured on MI355X: the 8-byte form's registers
// spill there).
template <typename T, int M, bool DENSE, bool GEN2Q, int NT, bool SWZ, bool VEC = true>
__device__ __forceinline__ void tile_ops(T *s, const TileOp *__restrict__ ops, int nops, const T *__restrict__ prm) {
  constexpr int E = 1 << (2 * M);
  constexpr int GPT = (E / 16 + NT - 1) / NT;  // groups of 16 per thread
  using Q = Pack<T, 4>;
  const int tid = threadIdx.x;
  for (int i = 0; i < nops; ++i) {
    const TileOp op = ops[i];
    if (DENSE && op.kind == TOP_DENSE) {
      if (M == 6) tile_dense<T, M>(s, op, prm + op.param);
      __syncthreads();
      continue;
    }
    // TOP_STAGE: the members follow the header in the list
    const int ja = op.d[0], jb = op.d[1];
    const TileOp *mem = ops + i + 1;
    const int nm = op.k;
    i += nm;
    uint32_t ka[4], kb[4];  // byte offsets of the components along ja, jb
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ka[u] = lds_swz<SWZ>(u << (2 * ja)) * (uint32_t)sizeof(T);
      kb[u] = lds_swz<SWZ>(u << (2 * jb)) * (uint32_t)sizeof(T);
    }
    if (ja == 0) {
      // VEC kernels with 32+ FP32 coefficients per thread (7-digit tiles of
      // 512 threads: 2 groups) load all of a thread's groups before the
      // first operation; the others take one group at a time (their
      // register budget)
      constexpr int GB = (VEC && sizeof(T) == 4 && E / NT >= 32) ? GPT : 1;
      for (int g0 = tid; g0 < E / 16; g0 += GB * NT) {
        uint32_t bq[GB];
        T v[GB][4][4];
#pragma unroll
        for (int q = 0; q < GB; ++q) {
          const int g = g0 + q * NT;
          bq[q] = (SWZ ? lds_swz<true>(lane_deposit(g, op.lanes, M - 2)) : ins_zero(ins_zero(g, 0), jb)) * (uint32_t)sizeof(T);
          if (g < E / 16) {
#pragma unroll
            for (int y = 0; y < 4; ++y) {
              const Q qd = lds_ld<Q>(s, bq[q] ^ kb[y]);
#pragma unroll
              for (int x = 0; x < 4; ++x) v[q][x][y] = qd.v[x];
            }
          }
        }
        stage_apply<T, GEN2Q, GB>(v, 0, mem, nm, prm);
#pragma unroll
        for (int q = 0; q < GB; ++q) {
          if (g0 + q * NT >= E / 16) continue;
          // the (I, I) coefficient never changes; left visible, the compiler
          // drops it from the first quad's store and splits that store into
          // 4-byte writes that hit each bank 4 times
          asm volatile("" : "+v"(v[q][0][0]));
#pragma unroll
          for (int y = 0; y < 4; ++y) {
            Q qd;
#pragma unroll
            for (int x = 0; x < 4; ++x) qd.v[x] = v[q][x][y];
            lds_st(s, bq[q] ^ kb[y], qd);
          }
        }
      }
    } else if (VEC && sizeof(T) == 4 && E / NT >= 32) {
      // (with fewer than 32 coefficients per thread the 8-byte form would
      // leave threads idle)
      using P2 = Pack<T, 2>;
      constexpr int UN = E / 32;
      for (int un = tid; un < UN; un += NT) {
        const uint32_t b = lds_swz<SWZ>(((un & 1) << 1) | lane_deposit(un >> 1, op.qlanes, M - 3)) * (uint32_t)sizeof(T);
        T v[2][4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) {
            const P2 p = lds_ld<P2>(s, b ^ ka[x] ^ kb[y]);
            v[0][x][y] = p.v[0];
            v[1][x][y] = p.v[1];
          }
        stage_apply<T, GEN2Q, 2>(v, ja, mem, nm, prm);
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y)
            if (x | y) {
              P2 p;
              p.v[0] = v[0][x][y];
              p.v[1] = v[1][x][y];
              lds_st(s, b ^ ka[x] ^ kb[y], p);
            }
      }
    } else {
      for (int g = tid; g < E / 16; g += NT) {
        const uint32_t b = (SWZ ? lds_swz<true>(lane_deposit(g, op.lanes, M - 2)) : ins_zero(ins_zero(g, ja), jb)) * (uint32_t)sizeof(T);
        T v[1][4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) v[0][x][y] = lds_ld<T>(s, b ^ ka[x] ^ kb[y]);
        stage_apply<T, GEN2Q, 1>(v, ja, mem, nm, prm);
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y)
            if (x | y) lds_st(s, b ^ ka[x] ^ kb[y], v[0][x][y]);
      }
    }
    __syncthreads();
  }
}

// A fused pass: each workgroup loads a tile (the M tile digits of geo) into
// LDS, applies the pass's operations and writes it back, looping over tiles
// with stride gridDim.x and prefetching its next tile into registers.  The
// host launches grid = tiles or a persistent grid; mode bit 0 (persistent
// grids) keeps consecutive tiles on one XCD, bit 1 (FP32 runs of >= 4^3)
// makes the tile loads and stores nontemporal.  SWZ: the tile sits in LDS
// at lds_swz of each index.
template <typename T, int M, bool DENSE, bool GEN2Q, int NT, bool SWZ>
__global__ __launch_bounds__(NT) void k_tile_prog(T *__restrict__ a, TileGeom geo, const TileOp *__restrict__ ops,
                                                      int nops, const T *__restrict__ prm, int mode) {
  constexpr int E = 1 << (2 * M);
  constexpr int QPT = (E / 4 + NT - 1) / NT;  // quads per thread (4 for M=6, 1 for M=5)
  using Q = Pack<T, 4>;
  __shared__ __align__(32) T s[E];
  const int tid = threadIdx.x;
  const int w = geo.w;
  const int lomask = (1 << (2 * w)) - 1;
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
  // L2).  With remap set (persistent grids over short runs, where
  // neighbouring tiles share HBM lines) consecutive tiles go to workgroups of
  // the same XCD: logical id = (b % 8) * (grid / 8) + b / 8.
  const uint64_t g8 = gridDim.x / 8;
  const bool remap = mode & 1;
  const uint64_t lid = (remap && gridDim.x % 8 == 0) ? (blockIdx.x % 8) * g8 + blockIdx.x / 8 : blockIdx.x;
  auto tile_base = [&](uint64_t tile) {
    uint64_t base = 0, r = tile;
    for (int j = 0; j < geo.nr; ++j) {
      base |= (r & 3ull) << (2 * geo.rpos[j]);
      r >>= 2;
    }
    return base;
  };
  // quad c = tid + i * NT is tile indices 4c..4c+3, 4 consecutive
  // coefficients in HBM and (swizzled or not) in LDS
  // nontemporal loads and stores (FP32 runs of >= 4^3 coefficients); runs
  // of 4^2 share HBM lines between neighbouring tiles through L2
  const bool nt = mode & 2;
  Q buf[QPT];
  auto load_tile = [&](uint64_t b) {
    if (nt) {
#pragma unroll
      for (int i = 0; i < QPT; ++i) {
        const int c = tid + i * NT;
        if (c < E / 4) buf[i] = ldq<T, true>(a + b + tile_offset<T, M>(4 * c, w, lomask, geo));
      }
    } else {
#pragma unroll
      for (int i = 0; i < QPT; ++i) {
        const int c = tid + i * NT;
        if (c < E / 4) buf[i] = ldq<T, false>(a + b + tile_offset<T, M>(4 * c, w, lomask, geo));
      }
    }
  };
  uint64_t tile = lid;
  if (tile < geo.tiles) load_tile(tile_base(tile));
  for (; tile < geo.tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(tile);
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) lds_put_quad(s, lds_swz<SWZ>(4 * c), buf[i]);
    }
    // prefetch the next tile while this one is computed
    const uint64_t next = tile + gridDim.x;
    if (next < geo.tiles) load_tile(tile_base(next));
    __syncthreads();
    tile_ops<T, M, DENSE, GEN2Q, NT, SWZ>(s, ops, nops, prm);
    if (nt) {
#pragma unroll
      for (int i = 0; i < QPT; ++i) {
        const int c = tid + i * NT;
        if (c < E / 4) stq<T, true>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo), lds_get_quad(s, lds_swz<SWZ>(4 * c)));
      }
    } else {
#pragma unroll
      for (int i = 0; i < QPT; ++i) {
        const int c = tid + i * NT;
        if (c < E / 4) stq<T, false>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo), lds_get_quad(s, lds_swz<SWZ>(4 * c)));
      }
    }
    __syncthreads();
  }
}

// The first or last stage of a one-tile pass, fused with the HBM copy: DIR
// 0 reads the stage's groups of 16 coefficients from HBM and writes its
// result to the LDS tile, DIR 1 reads the groups from LDS and writes the
// result to HBM, DIR 2 (a pass of one stage) goes from HBM to HBM without
// LDS.  Saves one LDS round trip on either side.  Only for stages whose
// group order is contiguous in HBM and conflict-free in LDS (stage_io_ok):
// digit 0 with a digit >= 3 (one quad per component; lanes over digits 1, 2
// first), or two digits >= 3 (lanes over digits 0-2 first; FP32 7-digit
// tiles pair the groups that differ in the low bit of digit 0 into 8-byte
// accesses).  A thread loads all its groups before the first operation.
// NTM: nontemporal HBM accesses (one-tile passes over whole states; the
// batch's small states keep them cached).
__host__ __device__ __forceinline__ bool stage_io_ok(const TileOp &op) {
  return (op.d[0] == 0 && op.d[1] >= 3) || op.d[0] >= 3;
}

template <typename T, int M, bool GEN2Q, int NT, bool SWZ, bool VEC, int DIR, bool NTM = true>
__device__ __forceinline__ void stage_io(T *__restrict__ a, uint64_t base, int w, int lomask, const TileGeom &geo,
                                         T *s, const TileOp *__restrict__ ops, int at, const T *__restrict__ prm) {
  constexpr int E = 1 << (2 * M);
  constexpr int GPT = (E / 16 + NT - 1) / NT;  // groups per thread
  using Q = Pack<T, 4>;
  const TileOp op = ops[at];
  const int ja = op.d[0], jb = op.d[1];
  const TileOp *mem = ops + at + 1;
  const int nm = op.k;
  const int tid = threadIdx.x;
  uint32_t ka[4], kb[4];  // LDS byte offsets of the components along ja, jb
  uint64_t ha[4], hb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    ka[u] = lds_swz<SWZ>(u << (2 * ja)) * (uint32_t)sizeof(T);
    kb[u] = lds_swz<SWZ>(u << (2 * jb)) * (uint32_t)sizeof(T);
    ha[u] = (uint64_t)u << (2 * geo.tpos[ja]);
    hb[u] = (uint64_t)u << (2 * geo.tpos[jb]);
  }
  if (VEC && sizeof(T) == 4 && ja != 0) {
    // one unit per thread: the two groups that differ in the low bit of
    // digit 0, one 8-byte access per component pair
    using P2 = Pack<T, 2>;
    const int un = tid;
    if (un < E / 32) {
      const int bt = ((un & 1) << 1) | lane_deposit(un >> 1, op.qlanes, M - 3);
      const uint32_t lb = lds_swz<SWZ>(bt) * (uint32_t)sizeof(T);
      T *hp = a + base + tile_offset<T, M>(bt, w, lomask, geo);
      T v[2][4][4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          P2 p;
          if (DIR == 1) {
            p = lds_ld<P2>(s, lb ^ ka[x] ^ kb[y]);
          } else {
            using V2 = T __attribute__((ext_vector_type(2)));
            const V2 *qp = reinterpret_cast<const V2 *>(hp + ha[x] + hb[y]);
            const V2 q = NTM ? __builtin_nontemporal_load(qp) : *qp;
            p.v[0] = q[0];
            p.v[1] = q[1];
          }
          v[0][x][y] = p.v[0];
          v[1][x][y] = p.v[1];
        }
      stage_apply<T, GEN2Q, 2>(v, ja, mem, nm, prm);
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          if (DIR == 0) {
            P2 p;
            p.v[0] = v[0][x][y];
            p.v[1] = v[1][x][y];
            lds_st(s, lb ^ ka[x] ^ kb[y], p);
          } else {
            using V2 = T __attribute__((ext_vector_type(2)));
            V2 q;
            q[0] = v[0][x][y];
            q[1] = v[1][x][y];
            if (NTM) __builtin_nontemporal_store(q, reinterpret_cast<V2 *>(hp + ha[x] + hb[y]));
            else *reinterpret_cast<V2 *>(hp + ha[x] + hb[y]) = q;
          }
        }
    }
    return;
  }
  uint32_t lb[GPT];
  T *hp[GPT];
  T v[GPT][4][4];
#pragma unroll
  for (int q = 0; q < GPT; ++q) {
    const int g = tid + q * NT;
    const int bt = SWZ ? lane_deposit(g, op.lanes, M - 2) : ins_zero(ins_zero(g, ja), jb);
    lb[q] = lds_swz<SWZ>(bt) * (uint32_t)sizeof(T);
    hp[q] = a + base + tile_offset<T, M>(bt, w, lomask, geo);
    if (ja == 0) {
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const Q qd = DIR == 1 ? lds_ld<Q>(s, lb[q] ^ kb[y]) : ldq<T, NTM>(hp[q] + hb[y]);
#pragma unroll
        for (int x = 0; x < 4; ++x) v[q][x][y] = qd.v[x];
      }
    } else {
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          v[q][x][y] = DIR == 1 ? lds_ld<T>(s, lb[q] ^ ka[x] ^ kb[y])
                                : NTM ? __builtin_nontemporal_load(hp[q] + ha[x] + hb[y]) : hp[q][ha[x] + hb[y]];
    }
  }
  stage_apply<T, GEN2Q, GPT>(v, ja, mem, nm, prm);
#pragma unroll
  for (int q = 0; q < GPT; ++q) {
    if (ja == 0) {
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        Q qd;
#pragma unroll
        for (int x = 0; x < 4; ++x) qd.v[x] = v[q][x][y];
        if (DIR == 0) lds_st(s, lb[q] ^ kb[y], qd);
        else stq<T, NTM>(hp[q] + hb[y], qd);
      }
    } else {
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          if (DIR == 0) lds_st(s, lb[q] ^ ka[x] ^ kb[y], v[q][x][y]);
          else if (NTM) __builtin_nontemporal_store(v[q][x][y], hp[q] + ha[x] + hb[y]);
          else hp[q][ha[x] + hb[y]] = v[q][x][y];
        }
    }
  }
}

// One tile per workgroup (grid = tiles, tile = block index): no prefetch of
// a next tile, so fewer registers during the operations and more resident
// workgroups per CU, which overlap each other's HBM and LDS phases.  Runs
// are >= 4^3 coefficients: nontemporal loads and stores (NTM; measured
// faster on MI355X for FP32 runs, 6 % on 7-digit tiles; FP64 keeps plain
// accesses).  FP32: 8 waves per SIMD (64
// VGPRs), 7 for swizzled tiles (72 VGPRs: their XOR addresses cannot use
// instruction offsets; 64 spilled); 7-digit tiles 4 waves (128 VGPRs), the
// two 512-thread workgroups their 64 KB of LDS allow per CU.
template <typename T, int M, bool GEN2Q, int NT, bool SWZ, bool NTM = true>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? ((1 << (2 * M)) / NT >= 32 ? 4 : SWZ && NT <= 256 ? 7 : 8) : 1))) void k_tile_stream(T *__restrict__ a, TileGeom geo, const TileOp *__restrict__ ops,
                                                        int nops, const T *__restrict__ prm, int io, int last_at) {
  constexpr int E = 1 << (2 * M);
  constexpr int QPT = (E / 4 + NT - 1) / NT;
  using Q = Pack<T, 4>;
  __shared__ __align__(32) T s[E];
  const uint64_t tile = blockIdx.x;
  if (tile >= geo.tiles) return;  // uniform over the workgroup
  const int tid = threadIdx.x;
  const int w = geo.w;
  const int lomask = (1 << (2 * w)) - 1;
  uint64_t base = 0, r = tile;
  for (int j = 0; j < geo.nr; ++j) {
    base |= (r & 3ull) << (2 * geo.rpos[j]);
    r >>= 2;
  }
  // io bit 0: the first operation is a stage that reads HBM itself, bit 1:
  // the stage with header last_at (the last operation) writes HBM itself
  constexpr bool VEC = E / NT >= 32;
  if ((io & 3) == 3 && last_at == 0) {
    stage_io<T, M, GEN2Q, NT, SWZ, VEC, 2, NTM>(a, base, w, lomask, geo, s, ops, 0, prm);
    return;
  }
  int i0 = 0;
  if (io & 1) {
    stage_io<T, M, GEN2Q, NT, SWZ, VEC, 0, NTM>(a, base, w, lomask, geo, s, ops, 0, prm);
    i0 = 1 + ops[0].k;
  } else if constexpr (VEC) {
    // 2 workgroups per CU: every load of the tile is issued before the first
    // LDS write (and every LDS read before the first store), so a thread
    // waits for HBM once.  6-digit tiles (8 workgroups per CU, 64 VGPRs)
    // keep one quad in flight per thread, which needs no buffer registers.
    Q buf[QPT];
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) buf[i] = ldq<T, NTM>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo));
    }
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) lds_put_quad(s, lds_swz<SWZ>(4 * c), buf[i]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) lds_put_quad(s, lds_swz<SWZ>(4 * c), ldq<T, NTM>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo)));
    }
  }
  __syncthreads();
  const int i1 = (io & 2) ? last_at : nops;
  tile_ops<T, M, false, GEN2Q, NT, SWZ, VEC>(s, ops + i0, i1 - i0, prm);
  if (io & 2) {
    stage_io<T, M, GEN2Q, NT, SWZ, VEC, 1, NTM>(a, base, w, lomask, geo, s, ops, last_at, prm);
  } else if constexpr (VEC) {
    Q buf[QPT];
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) buf[i] = lds_get_quad(s, lds_swz<SWZ>(4 * c));
    }
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) stq<T, NTM>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo), buf[i]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int c = tid + i * NT;
      if (c < E / 4) stq<T, NTM>(a + base + tile_offset<T, M>(4 * c, w, lomask, geo), lds_get_quad(s, lds_swz<SWZ>(4 * c)));
    }
  }
}

// ---------------------------------------------------------------------------
// K6: reductions (FP64 accumulation).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double s_part[kBlock / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_part[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += s_part[w];
  return t;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq(const T *__restrict__ a, uint64_t N, double *part) {
  double s = 0.0;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += step) {
    const double v = (double)a[i];
    s += v * v;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename T, typename TB>
__global__ __launch_bounds__(kBlock) void k_dot(const T *__restrict__ a, const TB *__restrict__ b,
                                                uint64_t N, double *part) {
  double s = 0.0;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += step)
    s += (double)a[i] * b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// per block: max |a-b|, sum a*b, sum a*a, sum b*b
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_diff(const TA *__restrict__ a, const TB *__restrict__ b, uint64_t N,
                                                 double *part) {
  double mx = 0.0, ab = 0.0, aa = 0.0, bb = 0.0;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += step) {
    const double x = (double)a[i], y = (double)b[i];
    mx = fmax(mx, fabs(x - y));
    ab += x * y;
    aa += x * x;
    bb += y * y;
  }
  __shared__ double s_m[kBlock];
  s_m[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + off]);
    __syncthreads();
  }
  const double m = s_m[0];
  __syncthreads();
  ab = block_sum(ab);
  __syncthreads();
  aa = block_sum(aa);
  __syncthreads();
  bb = block_sum(bb);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = m;
    part[4 * blockIdx.x + 1] = ab;
    part[4 * blockIdx.x + 2] = aa;
    part[4 * blockIdx.x + 3] = bb;
  }
}

__global__ __launch_bounds__(kBlock) void k_sum_parts(const double *part, int nparts, double *out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) s += part[i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[0] = s;
}

// Gather the coefficients with measured qubits' components in the basis
// support and I on every other qubit. j's digit t (base 2 or 4, most
// significant first) is measured qubit t.
struct GatherSpec {
  int m;
  int base_log;           // 1 (components {0,k}) or 2 (all four)
  int comp;               // k for base 2
  uint64_t stride[kMaxQubits];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const T *__restrict__ a, uint64_t count, GatherSpec gs,
                                                   double *out) {
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < count; j += step) {
    uint64_t idx = 0;
    for (int t = 0; t < gs.m; ++t) {
      const int sh = gs.base_log * (gs.m - 1 - t);
      const uint64_t d = (j >> sh) & ((1ull << gs.base_log) - 1ull);
      const uint64_t c = gs.base_log == 1 ? (d ? (uint64_t)gs.comp : 0ull) : d;
      idx += c * gs.stride[t];
    }
    out[j] = (double)a[idx];
  }
}

// One contraction step over digit t: (pre, din, post) -> (pre, 2, post),
// out[p,b,q] = sum_x w[b][x] in[p,x,q].
__global__ __launch_bounds__(kBlock) void k_contract(const double *__restrict__ in, double *__restrict__ out,
                                                     uint64_t pre, int din, uint64_t post, double w00,
                                                     double w01, double w02, double w03, double w10,
                                                     double w11, double w12, double w13) {
  const uint64_t total = pre * 2 * post;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t o = (uint64_t)blockIdx.x * kBlock + threadIdx.x; o < total; o += step) {
    const uint64_t q = o % post;
    const uint64_t b = (o / post) & 1ull;
    const uint64_t p = o / (2 * post);
    const double *src = in + p * din * post + q;
    const double w0 = b ? w10 : w00, w1 = b ? w11 : w01, w2 = b ? w12 : w02, w3 = b ? w13 : w03;
    double s = w0 * src[0] + w1 * src[post];
    if (din == 4) s += w2 * src[2 * post] + w3 * src[3 * post];
    out[o] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_chop(T *__restrict__ a, uint64_t N, T thr) {
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += step)
    if (fabs((double)a[i]) < (double)thr) a[i] = T(0);
}

template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_convert(const TS *__restrict__ s, TD *__restrict__ d, uint64_t N) {
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += step) d[i] = (TD)s[i];
}

// ---------------------------------------------------------------------------
// K7: Pauli -> matrix conversion.  Per qubit, (I,X,Y,Z) -> (I+Z, X-iY, X+iY,
// I-Z) on complex values; then mu_q = 2 r_q + c_q is scattered to
// (row, col) with qubit 0 the most significant bit.
// ---------------------------------------------------------------------------

// First qubit: real coefficients in, complex values out (out of place).
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_dense_first(const TS *__restrict__ in, double2 *__restrict__ out,
                                                        uint64_t groups, int p) {
  const uint64_t S = pow4(p);
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += step) {
    const uint64_t hi = g >> (2 * p), lo = g & (S - 1);
    const uint64_t b = hi * 4 * S + lo;
    const double x0 = (double)in[b], x1 = (double)in[b + S], x2 = (double)in[b + 2 * S], x3 = (double)in[b + 3 * S];
    // mu0 = I + Z; mu1 = X - iY; mu2 = X + iY; mu3 = I - Z
    out[b] = make_double2(x0 + x3, 0.0);
    out[b + S] = make_double2(x1, -x2);
    out[b + 2 * S] = make_double2(x1, x2);
    out[b + 3 * S] = make_double2(x0 - x3, 0.0);
  }
}

// Every other qubit, in place on the complex buffer: each thread owns the
// four entries of its group.
__global__ __launch_bounds__(kBlock) void k_dense_pass(double2 *__restrict__ buf, uint64_t groups, int p) {
  const uint64_t S = pow4(p);
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += step) {
    const uint64_t hi = g >> (2 * p), lo = g & (S - 1);
    const uint64_t b = hi * 4 * S + lo;
    const double2 x0 = buf[b], x1 = buf[b + S], x2 = buf[b + 2 * S], x3 = buf[b + 3 * S];
    buf[b] = make_double2(x0.x + x3.x, x0.y + x3.y);
    buf[b + S] = make_double2(x1.x + x2.y, x1.y - x2.x);
    buf[b + 2 * S] = make_double2(x1.x - x2.y, x1.y + x2.x);
    buf[b + 3 * S] = make_double2(x0.x - x3.x, x0.y - x3.y);
  }
}

__device__ __forceinline__ uint64_t spread_bits(uint64_t x) {
  // bit i -> bit 2i (x < 2^32)
  x &= 0xffffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// Output entries [o0, o0 + count) of the row-major 2^n x 2^n matrix: entry
// (row, col) is the buffer value at the bit-interleaved index
// sum_q (2 r_q + c_q) 4^p(q).
__global__ __launch_bounds__(kBlock) void k_dense_scatter(const double2 *__restrict__ b, double2 *__restrict__ out,
                                                          int n, uint64_t o0, uint64_t count) {
  const uint64_t dim = 1ull << n;
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += step) {
    const uint64_t o = o0 + i;
    const uint64_t row = o >> n, col = o & (dim - 1);
    out[i] = b[(spread_bits(row) << 1) | spread_bits(col)];
  }
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// K8: whole-circuit kernel for small registers (n <= 7).  One workgroup holds
// one circuit's entire state (4^n coefficients, at most 128 KB in FP64) in
// LDS, applies its operations (tile ops over all n digits; TOP_ALLQ applies
// one 1q map to every qubit, the per-level memory noise), then writes the
// Z-basis ensemble probabilities 2^n values and its start / end GPU clock.
// ---------------------------------------------------------------------------

enum { TOP_ALLQ = 5 };

struct BatchCirc {
  int32_t n, nops;
  int64_t op_off;   // first TileOp
  int64_t out_off;  // first probability
  double err;       // measurement polarisation retention
};

__device__ __forceinline__ void stamp_now(uint64_t *where) { *where = wall_clock64(); }

__global__ void k_stamp(uint64_t *where) { stamp_now(where); }

template <typename T>
__global__ __launch_bounds__(1024) void k_circuit_lds(const BatchCirc *__restrict__ circs, const TileOp *__restrict__ allops,
                                                      const T *__restrict__ prm, double *__restrict__ probs,
                                                      uint64_t *__restrict__ stamps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *s = reinterpret_cast<T *>(smem_raw);
  const BatchCirc c = circs[blockIdx.x];
  const int n = c.n;
  const int E = 1 << (2 * n);
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) stamp_now(&stamps[2 * blockIdx.x]);
  const T amp = (T)ldexp(1.0, -n);
  for (int i = tid; i < E; i += nt) {
    // |0...0>: 2^-n on indices whose digits are all I or Z
    const uint32_t lo = (uint32_t)i & 0x55555555u, hi = ((uint32_t)i >> 1) & 0x55555555u;
    s[i] = (lo ^ hi) ? T(0) : amp;
  }
  __syncthreads();
  const TileOp *ops = allops + c.op_off;
  for (int oi = 0; oi < c.nops; ++oi) {
    const TileOp op = ops[oi];
    const T *pv = prm + op.param;
    if (op.kind == TOP_1Q || op.kind == TOP_ALLQ) {
      const int j0 = op.kind == TOP_1Q ? op.d[0] : 0;
      const int j1 = op.kind == TOP_1Q ? op.d[0] + 1 : n;
      for (int j = j0; j < j1; ++j) {
        const int st = 1 << (2 * j);
        for (int g = tid; g < E / 4; g += nt) {
          const int b = ins_zero(g, j);
          const T x0 = s[b], x1 = s[b + st], x2 = s[b + 2 * st], x3 = s[b + 3 * st];
          s[b + st] = pv[0] * x0 + pv[1] * x1 + pv[2] * x2 + pv[3] * x3;
          s[b + 2 * st] = pv[4] * x0 + pv[5] * x1 + pv[6] * x2 + pv[7] * x3;
          s[b + 3 * st] = pv[8] * x0 + pv[9] * x1 + pv[10] * x2 + pv[11] * x3;
        }
        __syncthreads();
      }
      continue;
    }
    if (op.kind == TOP_CX || op.kind == TOP_2Q) {
      const int jc = op.d[0], jt = op.d[1];
      const int jl = min(jc, jt), jh = max(jc, jt);
      const int sc = 1 << (2 * jc), stt = 1 << (2 * jt);
      for (int g = tid; g < E / 16; g += nt) {
        const int b = ins_zero(ins_zero(g, jl), jh);
        T v[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int t = 0; t < 4; ++t) v[a][t] = s[b + a * sc + t * stt];
        if (op.kind == TOP_CX) {
          Scale16<T> sv;
#pragma unroll
          for (int q = 0; q < 16; ++q) sv.s[q] = op.k ? pv[q] : T(1);
          cx_map<T, T, true, true>(v, sv);
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int t = 0; t < 4; ++t)
              if (a | t) s[b + a * sc + t * stt] = v[a][t];
        } else {
#pragma unroll
          for (int r = 1; r < 16; ++r) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < 16; ++q) acc += pv[16 * r + q] * v[q >> 2][q & 3];
            s[b + (r >> 2) * sc + (r & 3) * stt] = acc;
          }
        }
      }
    }
    __syncthreads();
  }
  // Z ensemble: gather the {I,Z}^n sub-lattice into the free LDS after the
  // state (2^n values), contract one qubit at a time in place (FP64).
  double *v = reinterpret_cast<double *>(smem_raw + (size_t)E * sizeof(T));
  const int M = 1 << n;
  for (int x = tid; x < M; x += nt) {
    int idx = 0;
    for (int q = 0; q < n; ++q)  // bit q of x -> qubit n-1-q -> position q
      if ((x >> q) & 1) idx |= 3 << (2 * q);
    v[x] = (double)s[idx];
  }
  __syncthreads();
  for (int q = 0; q < n; ++q) {  // butterfly on bit q: (u, w) -> (u + e w, u - e w)
    for (int x = tid; x < M / 2; x += nt) {
      const int lo = x & ((1 << q) - 1);
      const int b = ((x >> q) << (q + 1)) | lo;
      const double u = v[b], w = v[b | (1 << q)];
      v[b] = u + c.err * w;
      v[b | (1 << q)] = u - c.err * w;
    }
    __syncthreads();
  }
  // outcome b: bit of qubit k is bit (n-1-k) of b (qubit 0 most significant);
  // in v, qubit k is bit n-1-k as well, so the order already matches
  for (int x = tid; x < M; x += nt) probs[c.out_off + x] = v[x];
  __syncthreads();
  if (tid == 0) stamp_now(&stamps[2 * blockIdx.x + 1]);
}

// ---------------------------------------------------------------------------
// K8 (HBM tier): batched circuits whose state lives in HBM.
//
// Each circuit is a chain of passes over its own state in one arena:
//   BP_INIT   write |0...0> (2^-n on every index whose digits are I or Z),
//   BP_TILE   one 6-digit tile pass (the tile operations of a fused block,
//             or a single gate / memory-noise group as a one-pass block),
//   BP_PROBS  gather the {I,Z}^n sub-lattice (2^n values), Walsh-Hadamard in
//             LDS (FP64), write the Z-ensemble probabilities.
// A pass is split into items of consecutive tiles.  Round r holds pass r of
// every circuit that has one; the items of a round are independent (distinct
// circuits, or distinct tiles of one pass), and one launch per round works
// through them with a grid-stride loop.  Rounds follow each other on one
// stream, so a pass sees the previous pass's writes through the kernel
// boundary: no workgroup ever waits for another.
// ---------------------------------------------------------------------------

enum { BP_INIT = 0, BP_TILE = 1, BP_PROBS = 2 };

struct BatchPass {
  TileGeom geo;       // BP_TILE: the pass's tile (6 digits, or 7 in FP32)
  int32_t kind;
  int32_t circ;       // HBM-tier circuit index (clock stamps)
  int32_t n;          // register size
  int32_t nops;       // BP_TILE: tile operations
  int32_t io;         // BP_TILE: stage_io bits (first / last stage reads / writes HBM)
  int32_t last_at;    // BP_TILE: header index of the last operation
  int64_t top_off;    // BP_TILE: first tile operation
  int64_t state_off;  // first coefficient of the circuit's state in the arena
  int64_t out_off;    // BP_PROBS: first probability
  double err;         // BP_PROBS: measurement polarisation retention
};

struct BatchItem {
  uint32_t pass;    // index into the pass table
  uint32_t tile0;   // first tile (BP_INIT: first 4096-coefficient chunk)
  uint32_t ntiles;
  uint32_t pad;
};

constexpr int kQueueThreads = 256;
constexpr int kQueueTile = 6;          // tile digits of HBM-tier passes
constexpr int kQueueMaxQubits = 12;    // 2^12 FP64 probabilities fill the LDS buffer
constexpr int kQueueLdsBytes = 32768;  // one FP64 6-digit tile, or 2^12 FP64 values
// FP32 passes whose 6-digit tile would leave runs of 4^2 coefficients take
// 7-digit tiles (64 KB of LDS, 512 threads, as the one-tile kernel), run by
// a second launch per round over those items only.
constexpr int kQueueTile7 = 7;
constexpr int kQueueThreads7 = 512;
template <int M>
struct QueueShape {
  static constexpr int threads = M == kQueueTile7 ? kQueueThreads7 : kQueueThreads;
  static constexpr int lds = M == kQueueTile7 ? 65536 : kQueueLdsBytes;
};

template <typename T, bool SWZ, int M>
__device__ __forceinline__ void queue_tiles(T *s, T *a, const BatchPass &bp, uint32_t t0, uint32_t nt,
                                            const TileOp *__restrict__ tops, const T *__restrict__ prm) {
  constexpr int NT = QueueShape<M>::threads;
  constexpr int E = 1 << (2 * M);
  constexpr int QPT = E / 4 / NT;
  constexpr bool VEC = E / NT >= 32;
  // 7-digit passes hold single-qubit and CX members only (the host refuses
  // others): the 16x16 member's registers would spill at 128 per lane
  constexpr bool GEN2Q = M != kQueueTile7;
  static_assert(E * sizeof(T) <= QueueShape<M>::lds, "the tile fits the round kernel's LDS");
  using Q = Pack<T, 4>;
  const int tid = threadIdx.x;
  const int w = bp.geo.w;
  const int lomask = (1 << (2 * w)) - 1;
  const TileOp *ops = tops + bp.top_off;
  for (uint32_t tile = t0; tile < t0 + nt; ++tile) {
    uint64_t base = 0, r = tile;
    for (int j = 0; j < bp.geo.nr; ++j) {
      base |= (r & 3ull) << (2 * bp.geo.rpos[j]);
      r >>= 2;
    }
    // the first and last stage read and write HBM themselves where
    // stage_io allows (bp.io), as in the one-tile kernel
    if ((bp.io & 3) == 3 && bp.last_at == 0) {
      stage_io<T, M, GEN2Q, NT, SWZ, VEC, 2, false>(a, base, w, lomask, bp.geo, s, ops, 0, prm);
      continue;
    }
    int i0 = 0;
    if (bp.io & 1) {
      stage_io<T, M, GEN2Q, NT, SWZ, VEC, 0, false>(a, base, w, lomask, bp.geo, s, ops, 0, prm);
      i0 = 1 + ops[0].k;
    } else {
      Q buf[QPT];
#pragma unroll
      for (int i = 0; i < QPT; ++i)
        buf[i] = ldp<Q>(a + base + tile_offset<T, M>(4 * (tid + i * NT), w, lomask, bp.geo));
#pragma unroll
      for (int i = 0; i < QPT; ++i) lds_put_quad(s, lds_swz<SWZ>(4 * (tid + i * NT)), buf[i]);
    }
    __syncthreads();
    const int i1 = (bp.io & 2) ? bp.last_at : bp.nops;
    tile_ops<T, M, false, GEN2Q, NT, SWZ>(s, ops + i0, i1 - i0, prm);
    if (bp.io & 2) {
      stage_io<T, M, GEN2Q, NT, SWZ, VEC, 1, false>(a, base, w, lomask, bp.geo, s, ops, bp.last_at, prm);
    } else {
#pragma unroll
      for (int i = 0; i < QPT; ++i) {
        const int c = tid + i * NT;
        stp(a + base + tile_offset<T, M>(4 * c, w, lomask, bp.geo), lds_get_quad(s, lds_swz<SWZ>(4 * c)));
      }
    }
    __syncthreads();  // LDS reads of this tile before the next tile's writes
  }
}

// The M-digit tile items [i0, i0 + gridDim.x) of one round: one workgroup
// per item of one tile, without the round kernel's loops, so the registers
// are those of the one-tile kernel (in the grid-stride round kernel the tile
// code sat at the 106-SGPR limit and the 7-digit form spilled to scratch).
// Waves per SIMD as the one-tile kernel: 7-digit FP32 tiles 4 (the two
// 64 KB workgroups per CU), 6-digit FP32 tiles 6.
template <typename T, int M>
__global__ __launch_bounds__(QueueShape<M>::threads) __attribute__((amdgpu_waves_per_eu(
    sizeof(T) == 4 ? (M == kQueueTile7 ? 4 : 6) : 1))) void k_batch_tile(const BatchPass *__restrict__ passes,
                                                                          const BatchItem *__restrict__ items,
                                                                          uint32_t i0,
                                                                          const TileOp *__restrict__ tops,
                                                                          const T *__restrict__ prm,
                                                                          T *__restrict__ arena) {
  __shared__ __align__(32) T s[1 << (2 * M)];
  const BatchItem it = items[i0 + blockIdx.x];
  const BatchPass &bp = passes[it.pass];
  T *a = arena + bp.state_off;
  if (bp.geo.ident) queue_tiles<T, false, M>(s, a, bp, it.tile0, 1, tops, prm);
  else queue_tiles<T, true, M>(s, a, bp, it.tile0, 1, tops, prm);
}

// The INIT and PROBS items [i0, i1) of one round, grid-stride over
// workgroups.
template <typename T>
__global__ __launch_bounds__(kQueueThreads) void k_batch_round(const BatchPass *__restrict__ passes,
                                                               const BatchItem *__restrict__ items, uint32_t i0,
                                                               uint32_t i1, const TileOp *__restrict__ tops,
                                                               const T *__restrict__ prm, T *__restrict__ arena,
                                                               double *__restrict__ probs,
                                                               uint64_t *__restrict__ stamps) {
  __shared__ __align__(32) unsigned char raw[kQueueLdsBytes];
  const int tid = threadIdx.x;
  for (uint32_t t = i0 + blockIdx.x; t < i1; t += gridDim.x) {
    const BatchItem it = items[t];
    const BatchPass &bp = passes[it.pass];
    T *a = arena + bp.state_off;
    if (bp.kind == BP_INIT) {
      if (tid == 0 && it.tile0 == 0) stamp_now(&stamps[2 * bp.circ]);
      const T amp = (T)ldexp(1.0, -bp.n);
      const uint64_t c0 = (uint64_t)it.tile0 << (2 * kQueueTile);
      const uint64_t c1 = c0 + ((uint64_t)it.ntiles << (2 * kQueueTile));
      for (uint64_t i = c0 + 4 * tid; i < c1; i += 4 * kQueueThreads) {
        Pack<T, 4> q;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint64_t x = i + c;
          const uint64_t lo = x & 0x5555555555555555ull, hi = (x >> 1) & 0x5555555555555555ull;
          q.v[c] = (lo ^ hi) ? T(0) : amp;
        }
        stp(a + i, q);
      }
    } else {
      // Z ensemble: bit q of x is the Z/I choice of qubit n-1-q (position q);
      // the butterfly on bit q applies (u, w) -> (u + e w, u - e w), and the
      // outcome index has qubit 0 as its most significant bit, as v does
      double *v = reinterpret_cast<double *>(raw);
      const int n = bp.n, nv = 1 << n;
      for (int x = tid; x < nv; x += kQueueThreads) {
        uint64_t idx = 0;
        for (int q = 0; q < n; ++q)
          if ((x >> q) & 1) idx |= 3ull << (2 * q);
        v[x] = (double)a[idx];
      }
      __syncthreads();
      for (int q = 0; q < n; ++q) {
        for (int x = tid; x < nv / 2; x += kQueueThreads) {
          const int lo = x & ((1 << q) - 1);
          const int b = ((x >> q) << (q + 1)) | lo;
          const double u = v[b], wv = v[b | (1 << q)];
          v[b] = u + bp.err * wv;
          v[b | (1 << q)] = u - bp.err * wv;
        }
        __syncthreads();
      }
      for (int x = tid; x < nv; x += kQueueThreads) probs[bp.out_off + x] = v[x];
      if (tid == 0) stamp_now(&stamps[2 * bp.circ + 1]);
    }
    __syncthreads();  // the LDS buffer is free before the next item
  }
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================

struct ProfSlot {
  uint64_t launches = 0;
  double ms = 0.0;
};

struct ProfPending {
  hipEvent_t a, b;
  int fam;
};

// A fused pass parsed from its parameter record (see DM_OP_FUSED in dmsim.h).
struct FusedPass {
  int m = 0;
  TileGeom geo;
  std::vector<TileOp> ops;     // param offsets relative to this pass's values
  std::vector<double> vals;    // values in kernel layout (FP64; converted at upload)
};

struct dm_state {
  int n = 0;
  int dtype = DM_FP64;
  uint64_t N = 0;
  int device = 0;
  int cu_count = 256;
  void *d_a = nullptr;
  hipStream_t stream = nullptr;
  bool owns_stream = true;
  // scratch
  double *d_red = nullptr;      // reduction partials + results
  size_t red_cap = 0;           // doubles
  void *d_mat = nullptr;        // program image: tile-operation records, block matrices (storage dtype)
  size_t mat_cap = 0;           // bytes
  double *d_aux = nullptr;      // gathered vectors / overlap operand
  size_t aux_cap = 0;           // bytes
  // memory noise with a thermal feed as tile passes: the passes, and their
  // records and values on the device, for the last (f, p, g)
  std::vector<FusedPass> mn_passes;
  std::vector<size_t> mn_top_off, mn_val_off;
  void *d_mn = nullptr;
  size_t mn_cap = 0;
  double mn_key[3] = {0.0, 0.0, -1.0};
  // profiling
  bool prof = false;
  int prof_depth = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ProfSlot slots[kFamilies];
  std::vector<ProfPending> pending;      // recorded, not yet resolved
  std::vector<hipEvent_t> spare_events;  // reusable
  hipEvent_t cur_a = nullptr;
};

namespace {

// One thread per work unit: for streaming passes a full grid measured faster
// on gfx950 than a capped grid-stride launch.  The kernels keep 64-bit
// grid-stride loops; the cap (2^22 workgroups = 2^30 work-items) keeps a
// launch below the 2^32 work-item limit of one grid dimension.
int grid_for(const dm_state *h, uint64_t units) {
  (void)h;
  const uint64_t want = (units + kBlock - 1) / kBlock;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(want, 1ull << 22));
}

int ensure(void **p, size_t *cap, size_t bytes) {
  if (*cap >= bytes) return DM_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_TRY(hipMalloc(p, bytes));
  *cap = bytes;
  return DM_OK;
}

int check_handle(const dm_state *h) {
  if (!h) return fail(DM_ERR_ARG, "null handle");
  if (!h->d_a) return fail(DM_ERR_STATE, "handle has no device state");
  return DM_OK;
}

int check_qubit(const dm_state *h, int q) {
  if (q < 0 || q >= h->n) return fail(DM_ERR_ARG, "qubit %d out of range for %d qubits", q, h->n);
  return DM_OK;
}

hipEvent_t take_event(dm_state *h) {
  if (!h->spare_events.empty()) {
    hipEvent_t e = h->spare_events.back();
    h->spare_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets the outermost library operation with a pair of HIP events on the
// handle's stream (nested scopes of composite operations fold into the outer
// one).  Nothing synchronises here; dm_profile_read resolves the pairs.
struct ProfScope {
  dm_state *h;
  int fam;
  bool outer;
  hipEvent_t a = nullptr;
  ProfScope(dm_state *h_, int fam_) : h(h_), fam(fam_), outer(h_->prof_depth++ == 0) {
    if (h->prof && outer) {
      a = take_event(h);
      if (a) hipEventRecord(a, h->stream);
    }
  }
  ~ProfScope() {
    h->prof_depth--;
    if (!h->prof || !outer || !a) return;
    hipEvent_t b = take_event(h);
    if (!b) return;
    hipEventRecord(b, h->stream);
    h->pending.push_back({a, b, fam});
  }
};

int resolve_profile(dm_state *h) {
  if (h->pending.empty()) return DM_OK;
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (auto &pp : h->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, pp.a, pp.b);
    h->slots[pp.fam].launches += 1;
    h->slots[pp.fam].ms += ms;
    h->spare_events.push_back(pp.a);
    h->spare_events.push_back(pp.b);
  }
  h->pending.clear();
  return DM_OK;
}

int last_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(DM_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return DM_OK;
}

template <typename T>
int launch_1q(dm_state *h, int q, const double *R) {
  const int p = h->n - 1 - q;
  bool use_i = R[0] != 0.0 || R[4] != 0.0 || R[8] != 0.0;
  Rows3<T> m = make_rows3<T>(R);
  T *a = static_cast<T *>(h->d_a);
  constexpr int W = 16 / sizeof(T);
  ProfScope ps(h, 0);
  if (p == 0) {
    const uint64_t groups = h->N / 4;
    if (use_i)
      hipLaunchKernelGGL((k_1q_p0<T, true>), dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream, a, groups, m);
    else
      hipLaunchKernelGGL((k_1q_p0<T, false>), dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream, a, groups, m);
  } else {
    const uint64_t S = pow4(p);
    const uint64_t units = h->N / 4 / W;
    int sh = 2 * p - (W == 2 ? 1 : 2);
    const dim3 grid(grid_for(h, units));
    if (p >= 2) {  // runs of >= 16 coefficients: nontemporal
      if (use_i) hipLaunchKernelGGL((k_1q_wide<T, true, true>), grid, dim3(kBlock), 0, h->stream, a, units, sh, S, m);
      else hipLaunchKernelGGL((k_1q_wide<T, false, true>), grid, dim3(kBlock), 0, h->stream, a, units, sh, S, m);
    } else {
      if (use_i) hipLaunchKernelGGL((k_1q_wide<T, true, false>), grid, dim3(kBlock), 0, h->stream, a, units, sh, S, m);
      else hipLaunchKernelGGL((k_1q_wide<T, false, false>), grid, dim3(kBlock), 0, h->stream, a, units, sh, S, m);
    }
  }
  return last_launch("k_1q");
}

template <typename T, bool CTRL_HI, bool SCALED>
int launch_cx_t(dm_state *h, int ph, int pl, const Scale16<T> &sc) {
  T *a = static_cast<T *>(h->d_a);
  constexpr int W = 16 / sizeof(T);
  const uint64_t Sh = pow4(ph), Sl = pow4(pl);
  const int sh_mid = 2 * (ph - pl - 1);
  if (pl == 0) {
    const uint64_t groups = h->N / 16;
    hipLaunchKernelGGL((k_cx_p0<T, CTRL_HI, SCALED>), dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream, a,
                       groups, sh_mid, Sh, sc);
  } else {
    const uint64_t units = h->N / 16 / W;
    const int sh_lo = 2 * pl - (W == 2 ? 1 : 2);
    if (pl >= 2)  // runs of >= 16 coefficients: nontemporal
      hipLaunchKernelGGL((k_cx_wide<T, CTRL_HI, SCALED, true>), dim3(grid_for(h, units)), dim3(kBlock), 0, h->stream,
                         a, units, sh_lo, sh_mid, Sh, Sl, sc);
    else
      hipLaunchKernelGGL((k_cx_wide<T, CTRL_HI, SCALED, false>), dim3(grid_for(h, units)), dim3(kBlock), 0,
                         h->stream, a, units, sh_lo, sh_mid, Sh, Sl, sc);
  }
  return last_launch("k_cx");
}

template <typename T>
int launch_cx(dm_state *h, int c, int t, const double *scale16) {
  const int pc = h->n - 1 - c, pt = h->n - 1 - t;
  const bool ctrl_hi = pc > pt;
  const int ph = std::max(pc, pt), pl = std::min(pc, pt);
  Scale16<T> sc;
  bool scaled = false;
  for (int i = 0; i < 16; ++i) {
    sc.s[i] = scale16 ? (T)scale16[i] : T(1);
    if (scale16 && scale16[i] != 1.0 && i != 0) scaled = true;
  }
  ProfScope ps(h, 1);
  if (ctrl_hi) {
    return scaled ? launch_cx_t<T, true, true>(h, ph, pl, sc) : launch_cx_t<T, true, false>(h, ph, pl, sc);
  }
  return scaled ? launch_cx_t<T, false, true>(h, ph, pl, sc) : launch_cx_t<T, false, false>(h, ph, pl, sc);
}

template <typename T>
int launch_2q(dm_state *h, int qa, int qb, const double *R) {
  const int pa = h->n - 1 - qa, pb = h->n - 1 - qb;
  const int ph = std::max(pa, pb), pl = std::min(pa, pb);
  // pair index in position order (i_hi, i_lo); R uses (i_qa, i_qb)
  Mat16<T> M;
  uint32_t rowmask = 0;
  for (int r = 0; r < 16; ++r) {
    const int rh = r >> 2, rl = r & 3;
    const int ra = pa > pb ? rh : rl, rb = pa > pb ? rl : rh;
    const int rsrc = 4 * ra + rb;
    bool ident = true;
    for (int c = 0; c < 16; ++c) {
      const int ch = c >> 2, cl = c & 3;
      const int ca = pa > pb ? ch : cl, cb = pa > pb ? cl : ch;
      const double v = R[16 * rsrc + 4 * ca + cb];
      M.m[r][c] = (T)v;
      if (v != (r == c ? 1.0 : 0.0)) ident = false;
    }
    if (!ident) rowmask |= 1u << r;
  }
  const uint64_t groups = h->N / 16;
  const int sh_lo = 2 * pl, sh_mid = 2 * (ph - pl - 1);
  ProfScope ps(h, 2);
  hipLaunchKernelGGL((k_2q_dense<T>), dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream,
                     static_cast<T *>(h->d_a), groups, sh_lo, sh_mid, pow4(ph), pow4(pl), M, rowmask);
  return last_launch("k_2q_dense");
}

template <typename T, int M>
int launch_tile_maps(dm_state *h, const int *pos, const double (*A)[4][4]) {
  TileMaps<T, M> tm;
  for (int j = 0; j < M; ++j) {
    tm.pos[j] = pos[j];
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) tm.A[j][r][c] = (T)A[j][r][c];
  }
  const uint64_t units = h->N >> (2 * M);
  hipLaunchKernelGGL((k_tile_maps<T, M>), dim3(grid_for(h, units)), dim3(kBlock), 0, h->stream,
                     static_cast<T *>(h->d_a), units, tm);
  return last_launch("k_tile_maps");
}

template <typename T>
int launch_memory_noise_tiles(dm_state *h, double f, double p, double g);

template <typename T>
int launch_memory_noise(dm_state *h, double f, double p, double g) {
  const double od = sqrt(g) * f;
  const double dd = (1 - g) * (2 * p - 1);
  ProfScope ps(h, 3);
  if (dd != 0.0 && h->n >= 6) return launch_memory_noise_tiles<T>(h, f, p, g);
  if (dd == 0.0) {
    constexpr int W = 16 / sizeof(T);
    const uint64_t units = h->N / W;
    hipLaunchKernelGGL((k_diag_pow<T>), dim3(grid_for(h, units)), dim3(kBlock), 0, h->stream,
                       static_cast<T *>(h->d_a), units, h->n, od, g);
    return last_launch("k_diag_pow");
  }
  double A[3][4][4];
  for (int j = 0; j < 3; ++j) {
    memset(A[j], 0, sizeof(A[j]));
    A[j][0][0] = 1.0;
    A[j][1][1] = od;
    A[j][2][2] = od;
    A[j][3][3] = g;
    A[j][3][0] = dd;
  }
  // position 0 (the last qubit): contiguous quads, the vectorised 1q kernel
  {
    const double R[12] = {0, od, 0, 0, 0, 0, od, 0, dd, 0, 0, g};
    int rc = launch_1q<T>(h, h->n - 1, R);
    if (rc) return rc;
  }
  int p0 = 1;
  while (p0 < h->n) {
    const int m = std::min(3, h->n - p0);
    int pos[3] = {p0, p0 + 1, p0 + 2};
    int rc = m == 3 ? launch_tile_maps<T, 3>(h, pos, A)
                    : m == 2 ? launch_tile_maps<T, 2>(h, pos, A) : launch_tile_maps<T, 1>(h, pos, A);
    if (rc) return rc;
    p0 += m;
  }
  return DM_OK;
}

template <typename T>
int launch_diag_tab(dm_state *h, const double *tab_by_pos, bool write_only, double scale) {
  Tab4 tab;
  memset(&tab, 0, sizeof(tab));
  memcpy(tab.t, tab_by_pos, sizeof(double) * 4 * h->n);
  constexpr int W = 16 / sizeof(T);
  const uint64_t units = h->N / W;
  ProfScope ps(h, write_only ? 7 : 4);
  if (write_only)
    hipLaunchKernelGGL((k_diag_tab<T, true>), dim3(grid_for(h, units)), dim3(kBlock), 0, h->stream,
                       static_cast<T *>(h->d_a), units, h->n, tab, scale);
  else
    hipLaunchKernelGGL((k_diag_tab<T, false>), dim3(grid_for(h, units)), dim3(kBlock), 0, h->stream,
                       static_cast<T *>(h->d_a), units, h->n, tab, scale);
  return last_launch("k_diag_tab");
}

template <typename T, int K, int L>
int launch_block_t(dm_state *h, const int *pos_sorted, const T *dST) {
  BlockGeom geo;
  memset(&geo, 0, sizeof(geo));
  geo.n = h->n;
  bool inblk[kMaxQubits] = {false};
  for (int j = 0; j < K; ++j) {
    geo.pos[j] = pos_sorted[j];
    inblk[pos_sorted[j]] = true;
  }
  int nl = 0;
  for (int p = 0; p < h->n && nl < L; ++p)
    if (!inblk[p]) {
      geo.lpos[nl++] = p;
      inblk[p] = true;
    }
  if (nl < L) return fail(DM_ERR_ARG, "block needs at least %d qubits", K + L);
  geo.nr = 0;
  for (int p = 0; p < h->n; ++p)
    if (!inblk[p]) geo.rpos[geo.nr++] = p;
  geo.tiles = 1ull << (2 * geo.nr);
  const int grid = (int)std::min<uint64_t>(geo.tiles, (uint64_t)h->cu_count * 8);
  hipLaunchKernelGGL((k_block_mfma<T, K, L>), dim3(grid), dim3(kBlock), 0, h->stream, static_cast<T *>(h->d_a),
                     dST, geo);
  return last_launch("k_block_mfma");
}

// Block matrix in kernel order: S is indexed by sum_j i_{qubits[j]} 4^j; the
// kernel's row digit j is the j-th ascending position.  Writes ST[col][row]
// (rows*rows values of T) and the ascending positions.
template <typename T>
void prep_block_matrix(int n, const int *qubits, int k, const double *S, T *st, int *pos_sorted) {
  const int rows = 1 << (2 * k);
  int pos[4], order[4];
  for (int j = 0; j < k; ++j) {
    pos[j] = n - 1 - qubits[j];
    order[j] = j;
  }
  std::sort(order, order + k, [&](int x, int y) { return pos[x] < pos[y]; });
  for (int j = 0; j < k; ++j) pos_sorted[j] = pos[order[j]];
  std::vector<int> remap(rows);
  for (int r = 0; r < rows; ++r) {
    int src = 0;
    for (int j = 0; j < k; ++j) src |= ((r >> (2 * j)) & 3) << (2 * order[j]);
    remap[r] = src;
  }
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < rows; ++c) st[(size_t)c * rows + r] = (T)S[(size_t)remap[r] * rows + remap[c]];
}

template <typename T>
int launch_block_dev(dm_state *h, int k, const int *pos_sorted, const T *dST) {
  ProfScope ps(h, k == 4 ? 8 : 5);
  if (k == 3) return launch_block_t<T, 3, 3>(h, pos_sorted, dST);
  if (k == 4) return launch_block_t<T, 4, 2>(h, pos_sorted, dST);
  return fail(DM_ERR_ARG, "block size k=%d not supported (3 or 4)", k);
}

template <typename T>
int launch_block(dm_state *h, const int *qubits, int k, const double *S) {
  const int rows = 1 << (2 * k);
  std::vector<T> st((size_t)rows * rows);
  int pos_sorted[4];
  prep_block_matrix<T>(h->n, qubits, k, S, st.data(), pos_sorted);
  int rc = ensure(&h->d_mat, &h->mat_cap, st.size() * sizeof(T));
  if (rc) return rc;
  // synchronous copy: the staging vector is released on return, and the
  // previous block kernel reading d_mat has to finish first
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(h->d_mat, st.data(), st.size() * sizeof(T), hipMemcpyHostToDevice));
  return launch_block_dev<T>(h, k, pos_sorted, static_cast<const T *>(h->d_mat));
}

int reserve_red(dm_state *h, size_t doubles) {
  return ensure(reinterpret_cast<void **>(&h->d_red), &h->red_cap, doubles * sizeof(double));
}

template <typename T>
int do_sumsq(dm_state *h, double *out) {
  const int grid = std::min(grid_for(h, h->N), 4096);
  int rc = reserve_red(h, (size_t)grid + 8);
  if (rc) return rc;
  ProfScope ps(h, 6);
  hipLaunchKernelGGL((k_sumsq<T>), dim3(grid), dim3(kBlock), 0, h->stream, static_cast<const T *>(h->d_a), h->N,
                     h->d_red);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, h->stream, h->d_red, grid, h->d_red + grid);
  rc = last_launch("k_sumsq");
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, h->d_red + grid, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DM_OK;
}

template <typename T>
int do_dot(dm_state *h, const double *d_b, double *out) {
  const int grid = std::min(grid_for(h, h->N), 4096);
  int rc = reserve_red(h, (size_t)grid + 8);
  if (rc) return rc;
  ProfScope ps(h, 6);
  hipLaunchKernelGGL((k_dot<T, double>), dim3(grid), dim3(kBlock), 0, h->stream, static_cast<const T *>(h->d_a),
                     d_b, h->N, h->d_red);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, h->stream, h->d_red, grid, h->d_red + grid);
  rc = last_launch("k_dot");
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, h->d_red + grid, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DM_OK;
}

// Host <-> device copies and dot products over the coefficients [off, off +
// count), the host side in type H (double or float).  Equal types copy
// directly; otherwise each chunk of at most kStageBytes passes through the
// handle's staging buffer and is converted on the device.  Nothing here
// allocates more than one chunk, so a 17-qubit state streams through a few
// hundred megabytes of device memory.
constexpr uint64_t kStageBytes = 256ull << 20;

template <typename T, typename H>
int copy_range(dm_state *h, uint64_t off, uint64_t count, H *host, bool to_host) {
  T *a = static_cast<T *>(h->d_a) + off;
  if (std::is_same<T, H>::value) {
    HIP_TRY(hipMemcpyAsync(to_host ? (void *)host : (void *)a, to_host ? (const void *)a : (const void *)host,
                           count * sizeof(T), to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return last_launch("copy_range");
  }
  const uint64_t chunk = kStageBytes / sizeof(H);
  HIP_TRY(hipStreamSynchronize(h->stream));  // the staging buffer may still be read by earlier work
  if (int rc = ensure(reinterpret_cast<void **>(&h->d_aux), &h->aux_cap, std::min(count, chunk) * sizeof(H)))
    return rc;
  H *stage = reinterpret_cast<H *>(h->d_aux);
  for (uint64_t i = 0; i < count; i += chunk) {
    const uint64_t c = std::min(chunk, count - i);
    if (to_host) {
      hipLaunchKernelGGL((k_convert<T, H>), dim3(grid_for(h, c)), dim3(kBlock), 0, h->stream, a + i, stage, c);
      HIP_TRY(hipMemcpyAsync(host + i, stage, c * sizeof(H), hipMemcpyDeviceToHost, h->stream));
    } else {
      HIP_TRY(hipMemcpyAsync(stage, host + i, c * sizeof(H), hipMemcpyHostToDevice, h->stream));
      hipLaunchKernelGGL((k_convert<H, T>), dim3(grid_for(h, c)), dim3(kBlock), 0, h->stream, stage, a + i, c);
    }
    if (int rc = last_launch("k_convert")) return rc;
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DM_OK;
}

template <typename T, typename H>
int dot_range(dm_state *h, uint64_t off, uint64_t count, const H *host, double *out) {
  const T *a = static_cast<const T *>(h->d_a) + off;
  const uint64_t chunk = kStageBytes / sizeof(H);
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (int rc = ensure(reinterpret_cast<void **>(&h->d_aux), &h->aux_cap, std::min(count, chunk) * sizeof(H)))
    return rc;
  H *stage = reinterpret_cast<H *>(h->d_aux);
  const int grid = 2048;
  if (int rc = reserve_red(h, (size_t)grid + 8)) return rc;
  double total = 0.0;
  for (uint64_t i = 0; i < count; i += chunk) {
    const uint64_t c = std::min(chunk, count - i);
    HIP_TRY(hipMemcpyAsync(stage, host + i, c * sizeof(H), hipMemcpyHostToDevice, h->stream));
    const int g = std::min(grid_for(h, c), grid);
    hipLaunchKernelGGL((k_dot<T, H>), dim3(g), dim3(kBlock), 0, h->stream, a + i, stage, c, h->d_red);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, h->stream, h->d_red, g, h->d_red + grid);
    if (int rc = last_launch("k_dot")) return rc;
    double part = 0.0;
    HIP_TRY(hipMemcpyAsync(&part, h->d_red + grid, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    total += part;
  }
  *out = total;
  return DM_OK;
}

template <typename T>
int do_probabilities(dm_state *h, const int *qubits, int m, char basis, const double *nvec, double err,
                     double *out) {
  GatherSpec gs;
  memset(&gs, 0, sizeof(gs));
  gs.m = m;
  double w[2][4] = {{1, 0, 0, 0}, {1, 0, 0, 0}};
  if (basis == 'N') {
    gs.base_log = 2;
    for (int k = 0; k < 3; ++k) {
      w[0][k + 1] = nvec[k] * err;
      w[1][k + 1] = -nvec[k] * err;
    }
  } else {
    gs.base_log = 1;
    gs.comp = basis == 'X' ? 1 : basis == 'Y' ? 2 : 3;
    w[0][1] = err;
    w[1][1] = -err;
  }
  for (int t = 0; t < m; ++t) gs.stride[t] = pow4(h->n - 1 - qubits[t]);
  const uint64_t count = 1ull << (gs.base_log * m);
  const uint64_t outn = 1ull << m;
  // aux holds two ping-pong buffers of `count` doubles
  int rc = ensure(reinterpret_cast<void **>(&h->d_aux), &h->aux_cap, 2 * count * sizeof(double));
  if (rc) return rc;
  double *buf0 = h->d_aux, *buf1 = h->d_aux + count;
  ProfScope ps(h, 6);
  hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(h, count)), dim3(kBlock), 0, h->stream,
                     static_cast<const T *>(h->d_a), count, gs, buf0);
  double *src = buf0, *dst = buf1;
  const uint64_t din = 1ull << gs.base_log;
  for (int t = 0; t < m; ++t) {
    const uint64_t pre = 1ull << t;
    uint64_t post = 1;
    for (int r = t + 1; r < m; ++r) post *= din;
    hipLaunchKernelGGL(k_contract, dim3(grid_for(h, pre * 2 * post)), dim3(kBlock), 0, h->stream, src, dst, pre,
                       (int)din, post, w[0][0], w[0][1], w[0][2], w[0][3], w[1][0], w[1][1], w[1][2], w[1][3]);
    std::swap(src, dst);
  }
  rc = last_launch("k_gather/k_contract");
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, src, outn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  // marginal over the unmeasured qubits: each contributes a factor 2 (sum of
  // its two outcome rows on the identity component)
  const double fac = ldexp(1.0, h->n - m);
  for (uint64_t i = 0; i < outn; ++i) out[i] *= fac;
  return DM_OK;
}

// Largest staging chunk of the dense output (bytes): the n per-qubit passes
// run in place on one complex buffer of 4^n entries, then the scatter writes
// the row-major matrix through this chunk to the host.
constexpr uint64_t kDenseChunkBytes = 1ull << 31;

template <typename T>
int do_to_dense(dm_state *h, double *out_c128) {
  const uint64_t N = h->N;
  const size_t bytes = N * sizeof(double2);
  const uint64_t chunk = std::min<uint64_t>(N, kDenseChunkBytes / sizeof(double2));
  double2 *buf = nullptr, *stage = nullptr;
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&buf), bytes, h->stream));
  hipError_t e1 = hipMallocAsync(reinterpret_cast<void **>(&stage), chunk * sizeof(double2), h->stream);
  if (e1 != hipSuccess) {
    hipFreeAsync(buf, h->stream);
    hipStreamSynchronize(h->stream);
    return fail(DM_ERR_HIP, "dense output of %d qubits: device buffer of %.1f GB: %s", h->n, bytes / 1e9,
                hipGetErrorString(e1));
  }
  const uint64_t groups = N / 4;
  ProfScope ps(h, 6);
  // qubit order 0..n-1, i.e. position n-1 down to 0 (the per-qubit maps commute)
  for (int q = 0; q < h->n; ++q) {
    const int p = h->n - 1 - q;
    if (q == 0)
      hipLaunchKernelGGL((k_dense_first<T>), dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream,
                         static_cast<const T *>(h->d_a), buf, groups, p);
    else
      hipLaunchKernelGGL(k_dense_pass, dim3(grid_for(h, groups)), dim3(kBlock), 0, h->stream, buf, groups, p);
  }
  int rc = last_launch("k_dense");
  for (uint64_t o0 = 0; rc == DM_OK && o0 < N; o0 += chunk) {
    const uint64_t cnt = std::min<uint64_t>(chunk, N - o0);
    hipLaunchKernelGGL(k_dense_scatter, dim3(grid_for(h, cnt)), dim3(kBlock), 0, h->stream, buf, stage, h->n, o0,
                       cnt);
    rc = last_launch("k_dense_scatter");
    if (rc) break;
    hipError_t e = hipMemcpyAsync(reinterpret_cast<double2 *>(out_c128) + o0, stage, cnt * sizeof(double2),
                                  hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) rc = fail(DM_ERR_HIP, "dense download failed: %s", hipGetErrorString(e));
  }
  hipFreeAsync(buf, h->stream);
  hipFreeAsync(stage, h->stream);
  HIP_TRY(hipStreamSynchronize(h->stream));
  return rc;
}

int param_count(const dm_state *h, const dm_op &op) {
  switch (op.kind) {
    case DM_OP_1Q: return 12;
    case DM_OP_CX: return 0;
    case DM_OP_CX_SCALED: return 16;
    case DM_OP_2Q: return 256;
    case DM_OP_BLOCK: return 1 << (4 * op.k);
    case DM_OP_MEMORY_NOISE: return 3;
    case DM_OP_DIAG: return 4 * h->n;
    case DM_OP_FUSED: return 0;  // variable length, checked by the parser
    default: return -1;
  }
}

int validate_op(const dm_state *h, const dm_op &op, int i, int64_t nparams) {
  const int need = param_count(h, op);
  if (need < 0) return fail(DM_ERR_ARG, "op %d: unknown kind %d", i, op.kind);
  int nq = 0;
  switch (op.kind) {
    case DM_OP_1Q: nq = 1; break;
    case DM_OP_CX: case DM_OP_CX_SCALED: case DM_OP_2Q: nq = 2; break;
    case DM_OP_BLOCK:
      if (op.k != 3 && op.k != 4) return fail(DM_ERR_ARG, "op %d: block size %d not supported", i, op.k);
      if (h->n < op.k + (op.k == 3 ? 3 : 2)) return fail(DM_ERR_ARG, "op %d: block too large for %d qubits", i, h->n);
      nq = op.k;
      break;
    default: nq = 0;
  }
  for (int j = 0; j < nq; ++j) {
    if (op.qubits[j] < 0 || op.qubits[j] >= h->n)
      return fail(DM_ERR_ARG, "op %d: qubit %d out of range for %d qubits", i, op.qubits[j], h->n);
    for (int m = 0; m < j; ++m)
      if (op.qubits[m] == op.qubits[j]) return fail(DM_ERR_ARG, "op %d: repeated qubit %d", i, op.qubits[j]);
  }
  if (need > 0 && (op.param < 0 || op.param + need > nparams))
    return fail(DM_ERR_ARG, "op %d: parameters [%lld, %lld) outside the %lld given", i, (long long)op.param,
                (long long)(op.param + need), (long long)nparams);
  return DM_OK;
}

// Stages.  An operation applied on its own costs one LDS round trip of the
// whole tile; operations whose digits fit in two tile digits are gathered
// into stages that apply them in registers, one round trip per stage.  An
// operation joins the first stage, at or after the last stage that touches
// its digits, whose digit set stays within two digits (it commutes with
// every later stage, which acts on other digits); otherwise it opens a new
// stage at the end.  Single-qubit operations that end up alone on their
// digit become one product PTM, run as a stage with digit 0 (or digit 3 when
// the digit is 0).  A dense operation is a stage of its own and stays a
// TOP_DENSE record.  Every other record of the result is a TOP_STAGE header
// followed by its members.
void group_stages(FusedPass *fp) {
  struct Stage {
    unsigned mask;
    std::vector<int> mem;
    bool dense;
  };
  std::vector<Stage> st;
  const int nops = (int)fp->ops.size();
  for (int i = 0; i < nops; ++i) {
    const TileOp &op = fp->ops[i];
    const int nd = op.kind == TOP_1Q ? 1 : op.kind == TOP_DENSE ? op.k : 2;
    unsigned D = 0;
    for (int x = 0; x < nd; ++x) D |= 1u << op.d[x];
    int j = -1;
    for (int k = 0; k < (int)st.size(); ++k)
      if (st[k].mask & D) j = k;
    int best = -1;
    if (op.kind != TOP_DENSE)
      for (int k = std::max(j, 0); k < (int)st.size() && best < 0; ++k)
        if (!st[k].dense && __builtin_popcount(st[k].mask | D) <= 2) best = k;
    if (best < 0) {
      st.push_back(Stage{D, {i}, op.kind == TOP_DENSE});
    } else {
      st[best].mask |= D;
      st[best].mem.push_back(i);
    }
  }
  // The one-tile kernel fuses the first and the last stage with the HBM copy
  // when their digits allow it (stage_io_ok).  A stage that shares no digit
  // with any stage before it may move to the front, one that shares none
  // with any stage after it to the back (it commutes with the others).
  auto io_ok = [&](const Stage &x) {
    if (x.dense) return false;
    TileOp hd;
    memset(&hd, 0, sizeof(hd));
    if (__builtin_popcount(x.mask) == 1) {
      const int d = __builtin_ctz(x.mask);
      hd.d[1] = d == 0 ? 3 : d;
    } else {
      hd.d[0] = __builtin_ctz(x.mask);
      hd.d[1] = 31 - __builtin_clz(x.mask);
    }
    return stage_io_ok(hd);
  };
  const int ns = (int)st.size();
  if (ns > 1 && !io_ok(st[0])) {
    unsigned before = 0;
    for (int k = 0; k < ns; ++k) {
      if (!(st[k].mask & before) && io_ok(st[k])) {
        std::rotate(st.begin(), st.begin() + k, st.begin() + k + 1);
        break;
      }
      before |= st[k].mask;
    }
  }
  if (ns > 2 && !io_ok(st[ns - 1])) {
    unsigned after = 0;
    for (int k = ns - 1; k >= 1; --k) {
      if (!(st[k].mask & after) && io_ok(st[k])) {
        std::rotate(st.begin() + k, st.begin() + k + 1, st.end());
        break;
      }
      after |= st[k].mask;
    }
  }
  std::vector<TileOp> out;
  out.reserve(nops + st.size());
  for (const Stage &s : st) {
    if (s.dense) {
      out.push_back(fp->ops[s.mem[0]]);
      continue;
    }
    TileOp hd;
    memset(&hd, 0, sizeof(hd));
    hd.kind = TOP_STAGE;
    if (__builtin_popcount(s.mask) == 1) {
      // P = R_last ... R_first over the components (I, X, Y, Z); each R has
      // the identity's first row and rows X, Y, Z from the operation
      TileOp op = fp->ops[s.mem[0]];
      if (s.mem.size() > 1) {
        double P[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
        for (int i : s.mem) {
          double R[4][4] = {{1, 0, 0, 0}};
          const double *r = &fp->vals[fp->ops[i].param];
          for (int c = 0; c < 3; ++c)
            for (int y = 0; y < 4; ++y) R[c + 1][y] = r[4 * c + y];
          double Q[4][4];
          for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
              double acc = 0.0;
              for (int z = 0; z < 4; ++z) acc += R[x][z] * P[z][y];
              Q[x][y] = acc;
            }
          memcpy(P, Q, sizeof(P));
        }
        op.param = (int32_t)fp->vals.size();
        for (int c = 1; c < 4; ++c)
          for (int y = 0; y < 4; ++y) fp->vals.push_back(P[c][y]);
      }
      const int d = op.d[0];
      hd.k = 1;
      hd.d[0] = 0;
      hd.d[1] = d == 0 ? 3 : d;
      out.push_back(hd);
      out.push_back(op);
      continue;
    }
    hd.k = (int32_t)s.mem.size();
    hd.d[0] = __builtin_ctz(s.mask);
    hd.d[1] = 31 - __builtin_clz(s.mask);
    out.push_back(hd);
    for (int i : s.mem) out.push_back(fp->ops[i]);
  }
  for (TileOp &op : out) {
    if (op.kind == TOP_1Q) {
      const double *r = &fp->vals[op.param];
      if (r[0] == 0.0 && r[4] == 0.0 && r[8] == 0.0) op.kind = TOP_1QU;
    } else if (op.kind == TOP_CX && op.k == 0) {
      op.k = 1;
      op.param = (int32_t)fp->vals.size();
      fp->vals.insert(fp->vals.end(), 16, 1.0);
    }
  }
  fp->ops.swap(out);
}

int parse_fused(const dm_state *h, const double *p, int64_t avail, int opi, FusedPass *fp) {
  int64_t at = 0;
  auto next = [&](double *out) -> bool {
    if (at >= avail) return false;
    *out = p[at++];
    return true;
  };
  auto next_int = [&](int *out) -> bool {
    double v;
    if (!next(&v) || v != (double)(int)v) return false;
    *out = (int)v;
    return true;
  };
  int m = 0, nq = 0;
  if (!next_int(&m) || m < 5 || m > 7) return fail(DM_ERR_ARG, "op %d: fused tile size must be 5, 6 or 7", opi);
  if (h->n < m) return fail(DM_ERR_ARG, "op %d: fused tile of %d qubits on %d qubits", opi, m, h->n);
  if (!next_int(&nq) || nq < 0 || nq > m) return fail(DM_ERR_ARG, "op %d: bad fused qubit count", opi);
  bool intile[kMaxQubits] = {false};
  int count = 0;
  for (int i = 0; i < nq; ++i) {
    int q;
    if (!next_int(&q) || q < 0 || q >= h->n) return fail(DM_ERR_ARG, "op %d: bad fused qubit", opi);
    const int pos = h->n - 1 - q;
    if (!intile[pos]) {
      intile[pos] = true;
      ++count;
    }
  }
  for (int pos = 0; pos < h->n && count < m; ++pos)
    if (!intile[pos]) {
      intile[pos] = true;
      ++count;
    }
  if (!intile[0]) return fail(DM_ERR_ARG, "op %d: fused tile must hold the last qubit's position", opi);
  memset(&fp->geo, 0, sizeof(fp->geo));
  fp->m = m;
  fp->geo.m = m;
  int digit_of_pos[kMaxQubits];
  int j = 0;
  for (int pos = 0; pos < h->n; ++pos)
    if (intile[pos]) {
      digit_of_pos[pos] = j;
      fp->geo.tpos[j++] = pos;
    } else {
      fp->geo.rpos[fp->geo.nr++] = pos;
    }
  int w = 0;
  while (w < m && fp->geo.tpos[w] == w) ++w;
  fp->geo.w = w;
  fp->geo.tiles = 1ull << (2 * fp->geo.nr);
  int nsub = 0;
  if (!next_int(&nsub) || nsub < 0) return fail(DM_ERR_ARG, "op %d: bad fused op count", opi);
  for (int i = 0; i < nsub; ++i) {
    int kind, kq, nv, qs[4];
    if (!next_int(&kind) || !next_int(&kq) || kq < 0 || kq > 4)
      return fail(DM_ERR_ARG, "op %d.%d: bad fused sub-op header", opi, i);
    for (int x = 0; x < 4; ++x)
      if (!next_int(&qs[x])) return fail(DM_ERR_ARG, "op %d.%d: bad fused sub-op qubits", opi, i);
    if (!next_int(&nv) || nv < 0 || at + nv > avail) return fail(DM_ERR_ARG, "op %d.%d: bad value count", opi, i);
    TileOp op;
    memset(&op, 0, sizeof(op));
    op.kind = kind;
    op.k = kq;
    for (int x = 0; x < kq; ++x) {
      if (qs[x] < 0 || qs[x] >= h->n || !intile[h->n - 1 - qs[x]])
        return fail(DM_ERR_ARG, "op %d.%d: qubit %d is not in the fused tile", opi, i, qs[x]);
      op.d[x] = digit_of_pos[h->n - 1 - qs[x]];
      for (int y = 0; y < x; ++y)
        if (qs[y] == qs[x]) return fail(DM_ERR_ARG, "op %d.%d: repeated qubit", opi, i);
    }
    int want = -1;
    if (kind == TOP_1Q && kq == 1) want = 12;
    else if (kind == TOP_CX && kq == 2) want = (nv == 16) ? 16 : 0;
    else if (kind == TOP_2Q && kq == 2) want = 256;
    else if (kind == TOP_DENSE && (kq == 3 || kq == 4) && m == 6) want = 1 << (4 * kq);
    if (want < 0 || nv != want) return fail(DM_ERR_ARG, "op %d.%d: bad fused sub-op (kind %d, %d qubits, %d values)",
                                           opi, i, kind, kq, nv);
    if (kind == TOP_CX) op.k = (nv == 16) ? 1 : 0;
    op.param = (int32_t)fp->vals.size();
    const double *v = p + at;
    if (kind == TOP_2Q) {
      for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 16; ++c)
          if (r == 0 && v[c] != (c == 0 ? 1.0 : 0.0))
            return fail(DM_ERR_ARG, "op %d.%d: two-qubit PTM is not trace preserving", opi, i);
    }
    if (kind == TOP_DENSE) {
      const int rows = 1 << (2 * kq);
      for (int c = 0; c < rows; ++c)
        for (int r = 0; r < rows; ++r) fp->vals.push_back(v[(size_t)r * rows + c]);  // transposed
    } else {
      fp->vals.insert(fp->vals.end(), v, v + nv);
    }
    at += nv;
    fp->ops.push_back(op);
  }
  // LDS layout.  The 64 banks (4 bytes each) are LDS index bits 0-5, tile
  // digits 0-2.  Stages on digits 3 and up, and stages on digit 0 and a
  // digit >= 3 (one quad access per component), fall on distinct banks in
  // the identity layout.  A pass with any other stage (digit 1 or 2 without
  // digit 0, or digit 0 with digit 1 or 2), whose lanes would hit each bank
  // 4 or 16 times (measured on MI355X: a CX on digits 0 and 1 costs 3.5x a
  // CX on digits 3 and 4), stores the tile swizzled (lds_swz, bank fields
  // [d0, d1^d3, d2^d4]) and gives each stage an order of its free digits
  // whose first ones cover distinct fields: d0, then d1 (else d3), then d2
  // (else d4), then the rest ascending.  Matrix-core passes keep the
  // identity layout.
  group_stages(fp);
  bool dense = false, swz = false;
  for (size_t i = 0; i < fp->ops.size(); ++i) {
    const TileOp &op = fp->ops[i];
    dense = dense || op.kind == TOP_DENSE;
    if (op.kind != TOP_STAGE) continue;
    swz = swz || (op.d[0] == 0 ? op.d[1] < 3 : op.d[0] < 3);
    i += op.k;  // members run in registers
  }
  swz = swz && !dense;
  fp->geo.ident = swz ? 0 : 1;
  // group orders of the stage headers (members run in registers)
  if (swz)
    for (TileOp &op : fp->ops) {
      if (op.kind != TOP_STAGE) continue;
      bool busy[8] = {false};
      busy[op.d[0]] = busy[op.d[1]] = true;
      int ord[8], nf = 0;
      auto take = [&](int d) {
        if (d < m && !busy[d]) {
          busy[d] = true;
          ord[nf++] = d;
          return true;
        }
        return false;
      };
      take(0);
      if (!take(1)) take(3);
      if (!take(2)) take(4);
      for (int d = 0; d < m; ++d) take(d);
      int packed = 0;
      for (int k = 0; k < nf; ++k) packed |= ord[k] << (3 * k);
      op.lanes = packed;
    }
  for (TileOp &op : fp->ops) {
    if (op.kind != TOP_STAGE) continue;
    bool busy[8] = {false};
    busy[0] = busy[op.d[0]] = busy[op.d[1]] = true;
    int ord[8], nf = 0;
    auto take = [&](int d) {
      if (d < m && !busy[d]) {
        busy[d] = true;
        ord[nf++] = d;
        return true;
      }
      return false;
    };
    if (!take(1)) take(3);
    if (!take(2)) take(4);
    for (int d = 1; d < m; ++d) take(d);
    int packed = 0;
    for (int k = 0; k < nf; ++k) packed |= ord[k] << (3 * k);
    op.qlanes = packed;
  }
  return DM_OK;
}

// Workgroups of one tile-program variant that fit on a CU at once (registers,
// LDS, waves), from the runtime's occupancy calculator; cached per variant.
template <typename T, int MM, bool D, bool G, int NTH, bool L>
int tile_resident_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tile_prog<T, MM, D, G, NTH, L>, NTH, 0) != hipSuccess ||
        nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}


template <typename T, int MM, bool G, int NTH, bool L>
int stream_resident_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tile_stream<T, MM, G, NTH, L>, NTH, 0) != hipSuccess ||
        nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}

// stage_io bits of a pass (bit 0: its first operation is a stage that can
// read HBM itself, bit 1: its last one can write HBM; stage_io_ok) and the
// header index of its last operation.
int stage_io_flags(const FusedPass &fp, int *last) {
  const int nops = (int)fp.ops.size();
  int last_stage = 0, io = 0;
  for (int i = 0; i < nops; i += fp.ops[i].kind == TOP_STAGE ? 1 + fp.ops[i].k : 1) last_stage = i;
  if (nops > 0 && fp.ops[0].kind == TOP_STAGE && stage_io_ok(fp.ops[0])) io |= 1;
  if (nops > 0 && fp.ops[last_stage].kind == TOP_STAGE && stage_io_ok(fp.ops[last_stage])) io |= 2;
  *last = last_stage;
  return io;
}

template <typename T>
int launch_fused(dm_state *h, const FusedPass &fp, const TileOp *d_ops, const T *d_vals) {
  ProfScope ps(h, 9);
  const int nops = (int)fp.ops.size();
  bool dense = false, gen = false;
  for (const TileOp &op : fp.ops) {
    dense = dense || op.kind == TOP_DENSE;
    gen = gen || op.kind == TOP_2Q;
  }
  T *a = static_cast<T *>(h->d_a);
  // the one-tile kernel fuses the first and the last stage with the HBM
  // copy where their group order suits both HBM and LDS
  int last_stage = 0;
  const int io = stage_io_flags(fp, &last_stage);
  // FP64 7-digit one-tile passes (1024 threads): the HBM-side stages of the
  // 16-byte form (stages on digit 0) ran slower than the staging copy
  // (16 qubits: 17.7 against 15.0 ms), those of the other forms faster
  int io_no_quad = io;
  if ((io & 1) && fp.ops[0].d[0] == 0) io_no_quad &= ~1;
  if ((io & 2) && fp.ops[last_stage].d[0] == 0) io_no_quad &= ~2;
  // one variant per feature set keeps the register footprint of the common
  // 1q + CX passes small; 7-qubit tiles (64 / 128 KB of LDS) use 512 threads
  // Grid shape, from 16-qubit sweeps on MI355X:
  //  - FP32, runs of >= 4^3 coefficients, at least 28 waves of the
  //    one-tile kernel resident per CU (6-digit tiles: 64 VGPRs, 8
  //    workgroups of 256 threads; 7-digit tiles stay below, their 64 KB of
  //    LDS allows 2): k_tile_stream, one tile per workgroup in tile order
  //    (the hardware overlaps the load, LDS and store phases of different
  //    workgroups);
  //  - FP64 6-digit tiles over runs of >= 4^3: k_tile_stream with plain
  //    accesses (110 VGPRs, 4 waves per SIMD; k_tile_prog's 146-162 leave
  //    3); FP64 7-digit tiles over such runs (without 16x16 members): the
  //    same with 1024 threads (113 VGPRs, 4 waves per SIMD in the one
  //    128 KB workgroup per CU, against 2 for the persistent kernel of 512,
  //    which spills at 1024);
  //  - other runs of >= 4^3 with 2 or more resident workgroups per CU:
  //    k_tile_prog with one tile per workgroup, whose loads all issue before
  //    the LDS writes; in FP32, where k_tile_prog's registers leave one
  //    resident workgroup and k_tile_stream's leave two, k_tile_stream;
  //  - otherwise (runs of 4^2, whose neighbouring tiles share HBM lines, or
  //    one resident workgroup per CU): persistent workgroups that prefetch
  //    their next tile, 8 (resp. 4) per resident slot, with consecutive
  //    tiles kept on one XCD.
  // FP32 tiles over runs of >= 4^3 load and store nontemporally (FP64
  // passes measured 1.3x slower with the hint).
  // Measured and rejected for FP32 7-digit tiles over runs of 4^3: the
  // persistent prefetching kernel (one resident workgroup per CU: 1.56x the
  // headline pass time) and one-tile workgroups of 1024 threads (64 VGPRs,
  // 8 waves per SIMD: 1.36x).
#define TILE_LAUNCH_L(MM, D, G, NTH, L)                                                                          \
  do {                                                                                                           \
    const int per_cu = tile_resident_per_cu<T, MM, D, G, NTH, L>();                                             \
    /* one workgroup per tile only where tiles x NTH work items fit a grid */                                 \
    const bool long_runs = fp.geo.w >= 3 && fp.geo.tiles * (uint64_t)NTH <= 0xffffffffull;                    \
    const int s_cu = D ? 0 : stream_resident_per_cu<T, MM, G, NTH, L>();                                        \
    if (long_runs && !D && sizeof(T) == 4 && (s_cu * NTH / 64 >= 28 || (per_cu < 2 && s_cu >= 2))) {            \
      hipLaunchKernelGGL((k_tile_stream<T, MM, G, NTH, L>), dim3((unsigned)fp.geo.tiles), dim3(NTH), 0,          \
                         h->stream, a, fp.geo, d_ops, nops, d_vals, io, last_stage);                             \
      break;                                                                                                     \
    }                                                                                                            \
    if constexpr (sizeof(T) == 8 && (MM == 6 || NTH == 1024) && !D) {                                            \
      if (long_runs) {                                                                                           \
        hipLaunchKernelGGL((k_tile_stream<T, MM, G, NTH, L, false>), dim3((unsigned)fp.geo.tiles), dim3(NTH), 0, \
                           h->stream, a, fp.geo, d_ops, nops, d_vals, MM == 7 ? io_no_quad : io, last_stage);    \
        break;                                                                                                   \
      }                                                                                                          \
    }                                                                                                            \
    const bool full = long_runs && per_cu >= 2;                                                                  \
    const uint64_t slots = (uint64_t)h->cu_count * per_cu * (per_cu >= 2 ? 8 : 4);                              \
    uint64_t g = full ? fp.geo.tiles : std::min<uint64_t>(fp.geo.tiles, slots);                                 \
    if (!full && g >= 8) g -= g % 8;                                                                             \
    hipLaunchKernelGGL((k_tile_prog<T, MM, D, G, NTH, L>), dim3((unsigned)g), dim3(NTH), 0, h->stream, a,       \
                       fp.geo, d_ops, nops, d_vals, (full ? 0 : 1) | (sizeof(T) == 4 && fp.geo.w >= 3 ? 2 : 0)); \
  } while (0)
#define TILE_LAUNCH(MM, D, G, NTH)               \
  do {                                           \
    if (fp.geo.ident) TILE_LAUNCH_L(MM, D, G, NTH, false); \
    else TILE_LAUNCH_L(MM, D, G, NTH, true);     \
  } while (0)
  if (fp.m == 7 && gen) TILE_LAUNCH(7, false, true, 512);
  else if (fp.m == 7 && sizeof(T) == 8 && fp.geo.w >= 3) TILE_LAUNCH(7, false, false, 1024);
  else if (fp.m == 7) TILE_LAUNCH(7, false, false, 512);
  else if (fp.m == 6 && dense) TILE_LAUNCH_L(6, true, true, 256, false);
  else if (fp.m == 6 && gen) TILE_LAUNCH(6, false, true, 256);
  else if (fp.m == 6) TILE_LAUNCH(6, false, false, 256);
  else if (gen) TILE_LAUNCH(5, false, true, 256);
  else TILE_LAUNCH(5, false, false, 256);
#undef TILE_LAUNCH_L
#undef TILE_LAUNCH
  return last_launch("k_tile_prog");
}

// Memory noise with a thermal feed, d = (1-g)(2p-1) != 0 (the z component
// of every qubit takes d a_I, so the channel is not diagonal), as tile
// passes of the 1-qubit channel R = [X: od x; Y: od y; Z: g z + d a_I] on
// every position: the first pass holds positions 0 .. m-1 (runs of 4^m
// coefficients), each further pass the three lowest positions (runs of 4^3)
// and m - 3 new ones; m = 6 in FP64 (32 KB of LDS), 7 in FP32 (64 KB).
// 1 + ceil((n - m) / (m - 3)) passes, against 1 + ceil((n - 1) / 3) for
// 3-qubit register tiles.  The passes and their device records are kept per
// handle for the last (f, p, g).
template <typename T>
int launch_memory_noise_tiles(dm_state *h, double f, double p, double g) {
  const int n = h->n;
  const int m = sizeof(T) == 4 && n >= 7 ? 7 : 6;
  if (!(h->mn_key[0] == f && h->mn_key[1] == p && h->mn_key[2] == g)) {
    const double od = sqrt(g) * f, dd = (1 - g) * (2 * p - 1);
    const double R[12] = {0, od, 0, 0, 0, 0, od, 0, dd, 0, 0, g};
    std::vector<FusedPass> passes;
    auto add_pass = [&](int lo, int hi) -> int {  // channel on positions [lo, hi); tile filled from position 0
      std::vector<double> rec = {(double)m, (double)(hi - lo)};
      for (int pos = lo; pos < hi; ++pos) rec.push_back(n - 1 - pos);
      rec.push_back(hi - lo);
      for (int pos = lo; pos < hi; ++pos) {
        const double head[7] = {(double)TOP_1Q, 1.0, (double)(n - 1 - pos), 0, 0, 0, 12.0};
        rec.insert(rec.end(), head, head + 7);
        rec.insert(rec.end(), R, R + 12);
      }
      FusedPass fp;
      if (int rc = parse_fused(h, rec.data(), (int64_t)rec.size(), 0, &fp)) return rc;
      passes.push_back(std::move(fp));
      return DM_OK;
    };
    if (int rc = add_pass(0, std::min(n, m))) return rc;
    for (int lo = m; lo < n; lo += m - 3)
      if (int rc = add_pass(lo, std::min(n, lo + m - 3))) return rc;
    // device image: [records of every pass | values (T)], values 256-byte aligned
    std::vector<TileOp> tops;
    std::vector<T> vals;
    std::vector<size_t> top_off, val_off;
    for (const FusedPass &fp : passes) {
      top_off.push_back(tops.size());
      val_off.push_back(vals.size());
      tops.insert(tops.end(), fp.ops.begin(), fp.ops.end());
      for (double v : fp.vals) vals.push_back((T)v);
    }
    const size_t vals_at = (tops.size() * sizeof(TileOp) + 255) & ~(size_t)255;
    const size_t bytes = vals_at + vals.size() * sizeof(T);
    HIP_TRY(hipStreamSynchronize(h->stream));  // earlier passes may still read the old image
    if (int rc = ensure(&h->d_mn, &h->mn_cap, bytes)) return rc;
    HIP_TRY(hipMemcpy(h->d_mn, tops.data(), tops.size() * sizeof(TileOp), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(static_cast<char *>(h->d_mn) + vals_at, vals.data(), vals.size() * sizeof(T),
                      hipMemcpyHostToDevice));
    for (size_t &o : val_off) o = vals_at + o * sizeof(T);
    h->mn_passes = std::move(passes);
    h->mn_top_off = std::move(top_off);
    h->mn_val_off = std::move(val_off);
    h->mn_key[0] = f;
    h->mn_key[1] = p;
    h->mn_key[2] = g;
  }
  for (size_t i = 0; i < h->mn_passes.size(); ++i) {
    const TileOp *d_ops = static_cast<const TileOp *>(h->d_mn) + h->mn_top_off[i];
    const T *d_vals = reinterpret_cast<const T *>(static_cast<const char *>(h->d_mn) + h->mn_val_off[i]);
    if (int rc = launch_fused<T>(h, h->mn_passes[i], d_ops, d_vals)) return rc;
  }
  return DM_OK;
}

// A program split into a host phase (stage: validate, parse the fused
// passes, lay out every tile-operation record and block matrix in one host
// image), an upload phase (the image into device memory) and a launch phase
// that only enqueues kernels on the handle's stream (so it can be captured in
// a graph).  The device image is [tile-operation records | values (T)], the
// values starting at a 256-byte boundary.
struct Prepared {
  std::vector<dm_op> ops;
  std::vector<double> params;
  std::vector<FusedPass> fused;
  std::vector<int> fused_of;
  std::vector<size_t> offs, top_off;
  std::vector<int> pos;
  std::vector<TileOp> tops;            // host image: records
  std::vector<unsigned char> values;   // host image: values, storage type
  size_t vals_at = 0;                  // byte offset of the values in the image
  const void *d_mat = nullptr;         // device values (after upload)
  const void *d_ops = nullptr;         // device records (after upload)
  size_t image_bytes() const { return vals_at + values.size(); }
};

template <typename T>
int stage_program(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams, Prepared *pp) {
  pp->ops.assign(ops, ops + nops);
  pp->params.assign(params, params + (nparams > 0 ? nparams : 0));
  pp->fused.clear();
  pp->fused_of.assign(nops, -1);
  size_t pool = 0;
  for (int i = 0; i < nops; ++i) {
    if (int rc = validate_op(h, ops[i], i, nparams)) return rc;
    if (ops[i].kind == DM_OP_BLOCK) pool += (size_t)1 << (4 * ops[i].k);
    if (ops[i].kind == DM_OP_FUSED) {
      if (ops[i].param < 0 || ops[i].param >= nparams) return fail(DM_ERR_ARG, "op %d: bad parameter offset", i);
      FusedPass fp;
      if (int rc = parse_fused(h, params + ops[i].param, nparams - ops[i].param, i, &fp)) return rc;
      pp->fused_of[i] = (int)pp->fused.size();
      pool += fp.vals.size();
      pp->fused.push_back(std::move(fp));
    }
  }
  pp->offs.assign(nops, 0);
  pp->pos.assign(4 * (size_t)nops, 0);
  pp->top_off.assign(pp->fused.size(), 0);
  // Tile operations are gathered for every fused pass, also when no pass
  // has parameter values (noiseless CX only): the kernels read them from
  // the records.
  pp->tops.clear();
  std::vector<T> staging(pool);
  size_t off = 0;
  for (int i = 0; i < nops; ++i) {
    if (ops[i].kind == DM_OP_BLOCK) {
      prep_block_matrix<T>(h->n, ops[i].qubits, ops[i].k, params + ops[i].param, staging.data() + off,
                           &pp->pos[4 * (size_t)i]);
      pp->offs[i] = off;
      off += (size_t)1 << (4 * ops[i].k);
    } else if (ops[i].kind == DM_OP_FUSED) {
      FusedPass &fp = pp->fused[pp->fused_of[i]];
      pp->offs[i] = off;
      for (double v : fp.vals) staging[off++] = (T)v;
      pp->top_off[pp->fused_of[i]] = pp->tops.size();
      pp->tops.insert(pp->tops.end(), fp.ops.begin(), fp.ops.end());
    }
  }
  pp->vals_at = (pp->tops.size() * sizeof(TileOp) + 255) & ~(size_t)255;
  pp->values.resize(staging.size() * sizeof(T));
  if (!staging.empty()) memcpy(pp->values.data(), staging.data(), pp->values.size());
  return DM_OK;
}

// Copies the host image to dev (the caller owns dev, at least
// image_bytes() long) after the handle's earlier work, which may still read
// the previous image there, has finished.
int upload_image(dm_state *h, Prepared *pp, void *dev) {
  if (pp->tops.empty() && pp->values.empty()) return DM_OK;
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (!pp->tops.empty())
    HIP_TRY(hipMemcpy(dev, pp->tops.data(), pp->tops.size() * sizeof(TileOp), hipMemcpyHostToDevice));
  if (!pp->values.empty())
    HIP_TRY(hipMemcpy(static_cast<char *>(dev) + pp->vals_at, pp->values.data(), pp->values.size(),
                      hipMemcpyHostToDevice));
  pp->d_ops = dev;
  pp->d_mat = static_cast<const char *>(dev) + pp->vals_at;
  return DM_OK;
}

// Stage, then upload into the handle's own buffer (grown with hipMalloc).
template <typename T>
int prepare_program(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams, Prepared *pp) {
  if (int rc = stage_program<T>(h, ops, nops, params, nparams, pp)) return rc;
  if (pp->tops.empty() && pp->values.empty()) return DM_OK;
  if (h->mat_cap < pp->image_bytes()) HIP_TRY(hipStreamSynchronize(h->stream));  // the old buffer may be in use
  if (int rc = ensure(&h->d_mat, &h->mat_cap, pp->image_bytes())) return rc;
  return upload_image(h, pp, h->d_mat);
}

template <typename T>
int launch_program(dm_state *h, const Prepared &pp) {
  const T *dmat = static_cast<const T *>(pp.d_mat);
  const double *params = pp.params.data();
  for (size_t i = 0; i < pp.ops.size(); ++i) {
    const dm_op &op = pp.ops[i];
    const double *pv = pp.params.empty() ? nullptr : params + op.param;
    int rc = DM_OK;
    switch (op.kind) {
      case DM_OP_1Q: rc = launch_1q<T>(h, op.qubits[0], pv); break;
      case DM_OP_CX: rc = launch_cx<T>(h, op.qubits[0], op.qubits[1], nullptr); break;
      case DM_OP_CX_SCALED: rc = launch_cx<T>(h, op.qubits[0], op.qubits[1], pv); break;
      case DM_OP_2Q: rc = launch_2q<T>(h, op.qubits[0], op.qubits[1], pv); break;
      case DM_OP_BLOCK: rc = launch_block_dev<T>(h, op.k, &pp.pos[4 * i], dmat + pp.offs[i]); break;
      case DM_OP_FUSED: {
        const int f = pp.fused_of[i];
        rc = launch_fused<T>(h, pp.fused[f], static_cast<const TileOp *>(pp.d_ops) + pp.top_off[f],
                             dmat + pp.offs[i]);
        break;
      }
      case DM_OP_MEMORY_NOISE:
        if (!(pv[0] == 1.0 && pv[2] == 1.0)) rc = launch_memory_noise<T>(h, pv[0], pv[1], pv[2]);
        break;
      case DM_OP_DIAG: {
        double tab[4 * kMaxQubits];
        for (int p = 0; p < h->n; ++p)
          for (int c = 0; c < 4; ++c) tab[4 * p + c] = pv[4 * (h->n - 1 - p) + c];
        rc = launch_diag_tab<T>(h, tab, false, 1.0);
        break;
      }
      default: rc = fail(DM_ERR_ARG, "op %zu: unknown kind", i);
    }
    if (rc) return rc;
  }
  return DM_OK;
}

template <typename T>
int run_program(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams) {
  Prepared pp;
  if (int rc = prepare_program<T>(h, ops, nops, params, nparams, &pp)) return rc;
  return launch_program<T>(h, pp);
}

}  // namespace

// Program image of dm_run_program_noalloc: one fixed array per device, held
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory; its device address comes from hipGetSymbolAddress
// (scratch_address), once per device.
constexpr size_t kNoAllocScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on `device`, looked up on the first call for
// that device with that device current (the caller's device is restored).
int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(DM_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(DM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return DM_OK;
}

// dm_run_program with the program image in the device's fixed scratch array
// instead of the handle's growable buffer: stage on the host, refuse an
// image larger than the array, copy it there after the handle's earlier work
// has finished, enqueue the same kernels.
template <typename T>
int run_program_noalloc(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams) {
  Prepared pp;
  if (int rc = stage_program<T>(h, ops, nops, params, nparams, &pp)) return rc;
  if (pp.image_bytes() > kNoAllocScratchBytes)
    return fail(DM_ERR_ARG, "program image of %zu bytes exceeds the %zu-byte scratch of dm_run_program_noalloc",
                pp.image_bytes(), kNoAllocScratchBytes);
  if (!pp.tops.empty() || !pp.values.empty()) {
    void *dev = nullptr;
    if (int rc = scratch_address(h->device, &dev)) return rc;
    if (int rc = upload_image(h, &pp, dev)) return rc;
  }
  return launch_program<T>(h, pp);
}

}  // namespace

// A batch of independent circuits run together (see dm_batch_create).
struct dm_batch {
  int dtype = DM_FP64;
  int ncirc = 0;
  std::vector<int> nq;
  std::vector<int64_t> prob_off;
  int64_t nprobs = 0;
  std::vector<double> err;
  // LDS tier: circuits held entirely in one workgroup's LDS, grouped by width
  std::vector<int> small;
  std::vector<int> small_width_start;  // per width 1..7: first index into small-ordered arrays
  void *d_circs = nullptr, *d_tops = nullptr, *d_vals = nullptr;
  int lds_threads = 256;
  // HBM tier: the states in one arena, the pass table, the items in round
  // order (round r: pass r of every circuit), and the tile operations and
  // values of every pass
  std::vector<int> large;
  void *d_arena = nullptr, *d_passes = nullptr, *d_items = nullptr, *d_qtops = nullptr, *d_qvals = nullptr;
  std::vector<uint32_t> round_start;  // items of round r: [round_start[r], round_start[r + 1])
  // round r: INIT and PROBS items [round_start[r], round_split6[r]), 6-digit
  // tile items [round_split6[r], round_split[r]) and 7-digit tile items
  // [round_split[r], round_start[r + 1]), one tile per tile item
  std::vector<uint32_t> round_split6, round_split;
  uint32_t nitems = 0;
  int64_t npasses = 0;
  int queue_grid = 0;  // resident workgroups of k_batch_round on the device
  hipStream_t master = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, ev0 = nullptr, ev1 = nullptr;
  // a round's tile launches run beside its INIT / PROBS launch
  hipStream_t side6 = nullptr, side7 = nullptr;
  hipEvent_t forkr = nullptr, join6 = nullptr, join7 = nullptr;
  double *d_probs = nullptr;
  uint64_t *d_stamps = nullptr;  // start / end clock per circuit: LDS tier first, then the HBM tier
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double clock_khz = 1.0e5;
  int64_t launches = 0;  // kernel launches one run enqueues
};

namespace {

// Tile ops of a small circuit: every position is a tile digit (digit j = position j).
int small_circuit_ops(int n, const dm_op *ops, int nops, const double *params, int64_t nparams, int ci,
                      std::vector<TileOp> *tops, std::vector<double> *vals) {
  auto digit = [&](int q) { return n - 1 - q; };
  auto push = [&](int kind, int k, const int *qs, const double *v, int nv) {
    TileOp op;
    memset(&op, 0, sizeof(op));
    op.kind = kind;
    op.k = k;
    for (int x = 0; x < 4; ++x) op.d[x] = qs ? digit(qs[x] < 0 ? 0 : qs[x]) : 0;
    op.param = (int32_t)vals->size();
    vals->insert(vals->end(), v, v + nv);
    tops->push_back(op);
  };
  dm_state probe;
  probe.n = n;
  for (int i = 0; i < nops; ++i) {
    const dm_op &op = ops[i];
    if (int rc = validate_op(&probe, op, i, nparams)) return rc;
    const double *pv = params + op.param;
    int qs[4] = {op.qubits[0], op.qubits[1], op.qubits[2], op.qubits[3]};
    switch (op.kind) {
      case DM_OP_1Q: push(TOP_1Q, 1, qs, pv, 12); break;
      case DM_OP_CX: push(TOP_CX, 0, qs, pv, 0); break;
      case DM_OP_CX_SCALED: push(TOP_CX, 1, qs, pv, 16); break;
      case DM_OP_2Q: {
        for (int c = 0; c < 16; ++c)
          if (pv[c] != (c == 0 ? 1.0 : 0.0))
            return fail(DM_ERR_ARG, "circuit %d op %d: two-qubit PTM is not trace preserving", ci, i);
        push(TOP_2Q, 2, qs, pv, 256);
        break;
      }
      case DM_OP_MEMORY_NOISE: {
        const double f = pv[0], p = pv[1], g = pv[2];
        if (f == 1.0 && g == 1.0) break;
        const double od = sqrt(g) * f, dd = (1 - g) * (2 * p - 1);
        const double R[12] = {0, od, 0, 0, 0, 0, od, 0, dd, 0, 0, g};
        push(TOP_ALLQ, n, nullptr, R, 12);
        break;
      }
      case DM_OP_FUSED: {
        // the whole register is the tile: inline the sub-operations
        // (record: m, nq, q..., nsub, then (kind, kq, q0..q3, nv, v...) each)
        int64_t at = op.param;
        const double *p = params;
        if (at + 2 > nparams) return fail(DM_ERR_ARG, "circuit %d op %d: truncated fused record", ci, i);
        const int nq = (int)p[at + 1];
        at += 2 + nq;
        if (nq < 0 || at + 1 > nparams) return fail(DM_ERR_ARG, "circuit %d op %d: truncated fused record", ci, i);
        const int nsub = (int)p[at++];
        for (int j = 0; j < nsub; ++j) {
          if (at + 7 > nparams) return fail(DM_ERR_ARG, "circuit %d op %d: truncated fused record", ci, i);
          const int kind = (int)p[at], kq = (int)p[at + 1];
          int sq[4] = {(int)p[at + 2], (int)p[at + 3], (int)p[at + 4], (int)p[at + 5]};
          const int nv = (int)p[at + 6];
          at += 7;
          if (nv < 0 || at + nv > nparams) return fail(DM_ERR_ARG, "circuit %d op %d: bad fused values", ci, i);
          if (kq < 0 || kq > 4) return fail(DM_ERR_ARG, "circuit %d op %d: bad fused sub-op size %d", ci, i, kq);
          for (int x = 0; x < kq; ++x)
            if (sq[x] < 0 || sq[x] >= n) return fail(DM_ERR_ARG, "circuit %d op %d: bad qubit", ci, i);
          if (kind == TOP_1Q && kq == 1 && nv == 12) push(TOP_1Q, 1, sq, p + at, 12);
          else if (kind == TOP_CX && kq == 2 && (nv == 0 || nv == 16)) push(TOP_CX, nv ? 1 : 0, sq, p + at, nv);
          else if (kind == TOP_2Q && kq == 2 && nv == 256) push(TOP_2Q, 2, sq, p + at, 256);
          else return fail(DM_ERR_ARG, "circuit %d op %d: sub-op kind %d not supported in LDS circuits", ci, i, kind);
          at += nv;
        }
        break;
      }
      default:
        return fail(DM_ERR_ARG, "circuit %d op %d: kind %d not supported in LDS circuits", ci, i, op.kind);
    }
  }
  return DM_OK;
}

size_t lds_bytes(int n, int dtype) {
  return ((size_t)1 << (2 * n)) * (dtype == DM_FP64 ? 8 : 4) + ((size_t)1 << n) * 8;
}

// The tile pass of one plain operation (a 1q / CX / 2q PTM): a one-operation
// fused record over its qubits, parsed like any fused pass.
int plain_op_pass(const dm_state *h, int kind, int kq, const int *qs, const double *v, int nv, int opi,
                  FusedPass *fp) {
  std::vector<double> rec = {(double)kQueueTile, (double)kq};
  for (int x = 0; x < kq; ++x) rec.push_back(qs[x]);
  rec.push_back(1.0);
  rec.push_back(kind);
  rec.push_back(kq);
  for (int x = 0; x < 4; ++x) rec.push_back(x < kq ? qs[x] : 0);
  rec.push_back(nv);
  rec.insert(rec.end(), v, v + nv);
  return parse_fused(h, rec.data(), (int64_t)rec.size(), opi, fp);
}

// The 6-digit tile passes of one program operation of an HBM-tier circuit.
// Memory noise becomes one 1-qubit channel per qubit, positions 0-5 in the
// first pass and 5 more per further pass.
int queue_op_passes(const dm_state *h, const dm_op &op, const double *params, int64_t nparams, int ci, int i,
                    std::vector<FusedPass> *out) {
  if (int rc = validate_op(h, op, i, nparams)) return rc;
  const double *pv = params + op.param;
  FusedPass fp;
  int rc = DM_OK;
  switch (op.kind) {
    case DM_OP_1Q: rc = plain_op_pass(h, TOP_1Q, 1, op.qubits, pv, 12, i, &fp); break;
    case DM_OP_CX: rc = plain_op_pass(h, TOP_CX, 2, op.qubits, pv, 0, i, &fp); break;
    case DM_OP_CX_SCALED: rc = plain_op_pass(h, TOP_CX, 2, op.qubits, pv, 16, i, &fp); break;
    case DM_OP_2Q: rc = plain_op_pass(h, TOP_2Q, 2, op.qubits, pv, 256, i, &fp); break;
    case DM_OP_FUSED: {
      if (op.param < 0 || op.param >= nparams) return fail(DM_ERR_ARG, "circuit %d op %d: bad parameter offset", ci, i);
      rc = parse_fused(h, params + op.param, nparams - op.param, i, &fp);
      if (rc) return rc;
      if (fp.m != kQueueTile && !(fp.m == kQueueTile7 && h->dtype == DM_FP32))
        return fail(DM_ERR_ARG, "circuit %d op %d: batched circuits take %d-digit fused tiles (or %d in FP32), not %d",
                    ci, i, kQueueTile, kQueueTile7, fp.m);
      for (const TileOp &t : fp.ops) {
        if (t.kind == TOP_DENSE) return fail(DM_ERR_ARG, "circuit %d op %d: dense tile operations are not batched", ci, i);
        if (t.kind == TOP_2Q && fp.m == kQueueTile7)
          return fail(DM_ERR_ARG, "circuit %d op %d: batched %d-digit passes take single-qubit and CX operations only",
                      ci, i, kQueueTile7);
      }
      break;
    }
    case DM_OP_MEMORY_NOISE: {
      const double f = pv[0], p = pv[1], g = pv[2];
      if (f == 1.0 && g == 1.0) return DM_OK;
      const double od = sqrt(g) * f, dd = (1 - g) * (2 * p - 1);
      const double R[12] = {0, od, 0, 0, 0, 0, od, 0, dd, 0, 0, g};
      for (int lo = 0; lo < h->n; lo = (lo == 0 ? kQueueTile : lo + kQueueTile - 1)) {
        const int hi = std::min(h->n, lo == 0 ? kQueueTile : lo + kQueueTile - 1);
        std::vector<double> rec = {(double)kQueueTile, (double)(hi - lo)};
        for (int pos = lo; pos < hi; ++pos) rec.push_back(h->n - 1 - pos);
        rec.push_back(hi - lo);
        for (int pos = lo; pos < hi; ++pos) {
          const double head[7] = {(double)TOP_1Q, 1.0, (double)(h->n - 1 - pos), 0, 0, 0, 12.0};
          rec.insert(rec.end(), head, head + 7);
          rec.insert(rec.end(), R, R + 12);
        }
        FusedPass mp;
        if (int r2 = parse_fused(h, rec.data(), (int64_t)rec.size(), i, &mp)) return r2;
        out->push_back(std::move(mp));
      }
      return DM_OK;
    }
    default:
      return fail(DM_ERR_ARG, "circuit %d op %d: kind %d is not batched", ci, i, op.kind);
  }
  if (rc) return rc;
  out->push_back(std::move(fp));
  return DM_OK;
}

template <typename T>
int enqueue_batch(dm_batch *b) {
  HIP_TRY(hipEventRecord(b->fork, b->master));
  HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
  // LDS tier on the side stream: one launch per width
  for (int w = 1; w <= 7; ++w) {
    const int lo = b->small_width_start[w], hi = b->small_width_start[w + 1];
    if (hi <= lo) continue;
    hipLaunchKernelGGL((k_circuit_lds<T>), dim3(hi - lo), dim3(b->lds_threads), lds_bytes(w, b->dtype), b->side,
                       static_cast<const BatchCirc *>(b->d_circs) + lo, static_cast<const TileOp *>(b->d_tops),
                       static_cast<const T *>(b->d_vals), b->d_probs, b->d_stamps + 2 * (size_t)lo);
    if (int rc = last_launch("k_circuit_lds")) return rc;
  }
  // HBM tier: the rounds in order on the master stream.  A round's INIT /
  // PROBS items (master), 6-digit tile items (side6) and 7-digit tile items
  // (side7, FP32) belong to distinct circuits: the three launches run side
  // by side, forked from and joined back into master.
  for (size_t r = 0; r + 1 < b->round_start.size(); ++r) {
    const uint32_t i0 = b->round_start[r], i6 = b->round_split6[r], i7 = b->round_split[r],
                   i1 = b->round_start[r + 1];
    const bool t6 = i7 > i6, t7 = sizeof(T) == 4 && i1 > i7;
    if (t6 || t7) {
      HIP_TRY(hipEventRecord(b->forkr, b->master));
      if (t6) HIP_TRY(hipStreamWaitEvent(b->side6, b->forkr, 0));
      if (t7) HIP_TRY(hipStreamWaitEvent(b->side7, b->forkr, 0));
    }
    if (i6 > i0) {
      const int grid = (int)std::min<uint32_t>(i6 - i0, (uint32_t)b->queue_grid);
      hipLaunchKernelGGL((k_batch_round<T>), dim3(grid), dim3(kQueueThreads), 0, b->master,
                         static_cast<const BatchPass *>(b->d_passes), static_cast<const BatchItem *>(b->d_items), i0,
                         i6, static_cast<const TileOp *>(b->d_qtops), static_cast<const T *>(b->d_qvals),
                         static_cast<T *>(b->d_arena), b->d_probs, b->d_stamps + 2 * b->small.size());
      if (int rc = last_launch("k_batch_round")) return rc;
    }
    if (t6) {
      hipLaunchKernelGGL((k_batch_tile<T, kQueueTile>), dim3(i7 - i6), dim3(kQueueThreads), 0, b->side6,
                         static_cast<const BatchPass *>(b->d_passes), static_cast<const BatchItem *>(b->d_items), i6,
                         static_cast<const TileOp *>(b->d_qtops), static_cast<const T *>(b->d_qvals),
                         static_cast<T *>(b->d_arena));
      if (int rc = last_launch("k_batch_tile")) return rc;
      HIP_TRY(hipEventRecord(b->join6, b->side6));
      HIP_TRY(hipStreamWaitEvent(b->master, b->join6, 0));
    }
    if constexpr (sizeof(T) == 4) {
      if (t7) {
        hipLaunchKernelGGL((k_batch_tile<float, kQueueTile7>), dim3(i1 - i7), dim3(kQueueThreads7), 0, b->side7,
                           static_cast<const BatchPass *>(b->d_passes), static_cast<const BatchItem *>(b->d_items),
                           i7, static_cast<const TileOp *>(b->d_qtops), static_cast<const float *>(b->d_qvals),
                           static_cast<float *>(b->d_arena));
        if (int rc = last_launch("k_batch_tile")) return rc;
        HIP_TRY(hipEventRecord(b->join7, b->side7));
        HIP_TRY(hipStreamWaitEvent(b->master, b->join7, 0));
      }
    }
  }
  HIP_TRY(hipEventRecord(b->join, b->side));
  HIP_TRY(hipStreamWaitEvent(b->master, b->join, 0));
  return DM_OK;
}

}  // namespace

namespace dmsim_internal {
int set_error(int code, const char *msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}
}  // namespace dmsim_internal

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" int dm_batch_destroy(dm_batch *b);

template <typename T>
static int batch_create_t(dm_batch *b, const int64_t *op_start, const dm_op *ops, const double *params,
                          int64_t nparams, int lds_max_qubits, int items_per_pass) {
  int cu = 256, dev = 0;
  {
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cu = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      b->clock_khz = khz;
  }
  // LDS circuits grouped by width (probabilities stay in circuit order)
  std::vector<BatchCirc> circs;
  std::vector<TileOp> tops;
  std::vector<double> vals;
  b->small_width_start.assign(9, 0);
  for (int w = 1; w <= 7; ++w) {
    b->small_width_start[w] = (int)b->small.size();
    if (w > lds_max_qubits) continue;
    for (int c = 0; c < b->ncirc; ++c) {
      if (b->nq[c] != w) continue;
      BatchCirc bc;
      bc.n = w;
      bc.op_off = (int64_t)tops.size();
      if (int rc = small_circuit_ops(w, ops + op_start[c], (int)(op_start[c + 1] - op_start[c]), params, nparams, c,
                                     &tops, &vals))
        return rc;
      bc.nops = (int32_t)(tops.size() - bc.op_off);
      bc.out_off = b->prob_off[c];
      bc.err = b->err[c];
      circs.push_back(bc);
      b->small.push_back(c);
    }
  }
  b->small_width_start[8] = (int)b->small.size();
  for (int w = 1; w <= 7; ++w) b->launches += b->small_width_start[w + 1] > b->small_width_start[w] ? 1 : 0;

  // HBM tier: every circuit as INIT, its tile passes, PROBS; states in one arena
  std::vector<BatchPass> passes;
  std::vector<TileOp> qtops;
  std::vector<double> qvals;
  std::vector<std::vector<int>> chain;  // pass-table indices of each HBM circuit, in order
  std::vector<double> work;             // per HBM circuit: bytes moved (ticket order)
  uint64_t arena = 0;
  for (int c = 0; c < b->ncirc; ++c) {
    if (b->nq[c] <= lds_max_qubits) continue;
    const int j = (int)b->large.size();
    b->large.push_back(c);
    dm_state probe;
    probe.n = b->nq[c];
    probe.dtype = b->dtype;
    probe.N = pow4(probe.n);
    BatchPass bp;
    memset(&bp, 0, sizeof(bp));
    bp.circ = j;
    bp.n = probe.n;
    bp.state_off = (int64_t)arena;
    arena += (probe.N + 63) & ~63ull;  // 256-byte aligned states
    std::vector<int> ids;
    bp.kind = BP_INIT;
    ids.push_back((int)passes.size());
    passes.push_back(bp);
    for (int64_t i = op_start[c]; i < op_start[c + 1]; ++i) {
      std::vector<FusedPass> fps;
      if (int rc = queue_op_passes(&probe, ops[i], params, nparams, c, (int)(i - op_start[c]), &fps)) return rc;
      for (FusedPass &fp : fps) {
        BatchPass tp = bp;
        tp.kind = BP_TILE;
        tp.geo = fp.geo;
        tp.nops = (int32_t)fp.ops.size();
        int last_at = 0;
        tp.io = stage_io_flags(fp, &last_at);
        tp.last_at = last_at;
        tp.top_off = (int64_t)qtops.size();
        const size_t vbase = qvals.size();
        for (TileOp t : fp.ops) {
          t.param += (int32_t)vbase;
          qtops.push_back(t);
        }
        qvals.insert(qvals.end(), fp.vals.begin(), fp.vals.end());
        ids.push_back((int)passes.size());
        passes.push_back(tp);
      }
    }
    bp.kind = BP_PROBS;
    bp.out_off = b->prob_off[c];
    bp.err = b->err[c];
    ids.push_back((int)passes.size());
    passes.push_back(bp);
    work.push_back((double)ids.size() * (double)probe.N);
    chain.push_back(std::move(ids));
  }
  if (qvals.size() > (size_t)INT32_MAX) return fail(DM_ERR_ARG, "batch: too many tile-operation values");
  b->npasses = (int64_t)passes.size();

  // Items in round order: round r holds pass r of every circuit that has
  // one, circuits with the most work first.  A pass of t tiles splits into
  // items of ceil(t / items_per_pass) tiles.
  std::vector<BatchItem> items;
  {
    std::vector<int> order(chain.size());
    for (size_t j = 0; j < order.size(); ++j) order[j] = (int)j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    size_t rounds = 0;
    for (auto &ids : chain) rounds = std::max(rounds, ids.size());
    for (size_t r = 0; r < rounds; ++r) {
      b->round_start.push_back((uint32_t)items.size());
      // the round's INIT and PROBS items, then its 6-digit tile items, then
      // its 7-digit ones (tile items: one tile each)
      for (int group = 0; group < 3; ++group) {
        if (group == 1) b->round_split6.push_back((uint32_t)items.size());
        if (group == 2) b->round_split.push_back((uint32_t)items.size());
        for (int j : order) {
          if (r >= chain[j].size()) continue;
          const int pi = chain[j][r];
          const BatchPass &bp = passes[pi];
          const int g = bp.kind != BP_TILE ? 0 : bp.geo.m == kQueueTile7 ? 2 : 1;
          if (g != group) continue;
          const uint64_t units =
              bp.kind == BP_PROBS ? 1 : pow4(bp.n - (bp.kind == BP_TILE ? bp.geo.m : kQueueTile));
          const uint64_t per =
              g ? 1 : std::max<uint64_t>(1, (units + items_per_pass - 1) / items_per_pass);
          for (uint64_t t0 = 0; t0 < units; t0 += per) {
            BatchItem it;
            it.pass = (uint32_t)pi;
            it.tile0 = (uint32_t)t0;
            it.ntiles = (uint32_t)std::min<uint64_t>(per, units - t0);
            it.pad = 0;
            items.push_back(it);
          }
          if (items.size() > (size_t)UINT32_MAX / 2) return fail(DM_ERR_ARG, "batch: too many work items");
        }
      }
    }
    if (rounds) b->round_start.push_back((uint32_t)items.size());
  }
  b->nitems = (uint32_t)items.size();

  HIP_TRY(hipStreamCreateWithFlags(&b->master, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
  HIP_TRY(hipStreamCreateWithFlags(&b->side6, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&b->side7, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&b->forkr, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&b->join6, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&b->join7, hipEventDisableTiming));
  HIP_TRY(hipEventCreate(&b->ev0));
  HIP_TRY(hipEventCreate(&b->ev1));
  auto upload = [](void **dst, const void *src, size_t bytes) -> int {
    HIP_TRY(hipMalloc(dst, std::max<size_t>(bytes, 16)));
    if (bytes) HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return DM_OK;
  };
  if (!circs.empty()) {
    std::vector<T> tv(vals.begin(), vals.end());
    if (int rc = upload(&b->d_circs, circs.data(), circs.size() * sizeof(BatchCirc))) return rc;
    if (int rc = upload(&b->d_tops, tops.data(), tops.size() * sizeof(TileOp))) return rc;
    if (int rc = upload(&b->d_vals, tv.data(), tv.size() * sizeof(T))) return rc;
    size_t need = 0;
    for (auto &bc : circs) need = std::max(need, lds_bytes(bc.n, b->dtype));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_circuit_lds<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  }
  if (b->nitems) {
    std::vector<T> tv(qvals.begin(), qvals.end());
    HIP_TRY(hipMalloc(&b->d_arena, arena * sizeof(T)));
    if (int rc = upload(&b->d_passes, passes.data(), passes.size() * sizeof(BatchPass))) return rc;
    if (int rc = upload(&b->d_items, items.data(), items.size() * sizeof(BatchItem))) return rc;
    if (int rc = upload(&b->d_qtops, qtops.data(), qtops.size() * sizeof(TileOp))) return rc;
    if (int rc = upload(&b->d_qvals, tv.data(), tv.size() * sizeof(T))) return rc;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_round<T>, kQueueThreads, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    b->queue_grid = cu * per_cu;
    for (size_t r = 0; r + 1 < b->round_start.size(); ++r)
      b->launches += (b->round_split6[r] > b->round_start[r] ? 1 : 0) +
                     (b->round_split[r] > b->round_split6[r] ? 1 : 0) +
                     (b->round_start[r + 1] > b->round_split[r] ? 1 : 0);
  }
  HIP_TRY(hipMalloc(&b->d_probs, std::max<int64_t>(1, b->nprobs) * sizeof(double)));
  HIP_TRY(hipMalloc(&b->d_stamps, 2 * (size_t)b->ncirc * sizeof(uint64_t)));
  HIP_TRY(hipDeviceSynchronize());
  return DM_OK;
}

template <typename TA, typename TB>
static int diff_t(dm_state *a, dm_state *b, double *out4) {
  const int grid = std::min(grid_for(a, a->N), 2048);
  if (int rc = reserve_red(a, 4 * (size_t)grid)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  hipLaunchKernelGGL((k_diff<TA, TB>), dim3(grid), dim3(kBlock), 0, a->stream, static_cast<const TA *>(a->d_a),
                     static_cast<const TB *>(b->d_a), a->N, a->d_red);
  if (int rc = last_launch("k_diff")) return rc;
  std::vector<double> part(4 * (size_t)grid);
  HIP_TRY(hipMemcpyAsync(part.data(), a->d_red, part.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  double m = 0, ab = 0, aa = 0, bb = 0;
  for (int i = 0; i < grid; ++i) {
    m = std::max(m, part[4 * i]);
    ab += part[4 * i + 1];
    aa += part[4 * i + 2];
    bb += part[4 * i + 3];
  }
  out4[0] = m;
  out4[1] = ab;
  out4[2] = aa;
  out4[3] = bb;
  return DM_OK;
}

#define DISPATCH(h, fn, ...) ((h)->dtype == DM_FP64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

int dm_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return DM_ERR_ARG;
  snprintf(buf, len, "%s", g_err);
  return DM_OK;
}

int dm_device_info(int device, int *n_devices, int *cu_count, uint64_t *hbm_bytes) {
  int nd = 0;
  HIP_TRY(hipGetDeviceCount(&nd));
  if (n_devices) *n_devices = nd;
  if (device < 0 || device >= nd) return fail(DM_ERR_ARG, "device %d not present (%d visible)", device, nd);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
  return DM_OK;
}

int dm_create(int n_qubits, int dtype, dm_state **out) {
  if (!out) return fail(DM_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > kMaxQubits)
    return fail(DM_ERR_ARG, "number of qubits %d outside [1, %d]", n_qubits, kMaxQubits);
  if (dtype != DM_FP64 && dtype != DM_FP32) return fail(DM_ERR_ARG, "unknown dtype %d", dtype);
  dm_state *h = new dm_state();
  h->n = n_qubits;
  h->dtype = dtype;
  h->N = pow4(n_qubits);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, h->device);
    if (e == hipSuccess) h->cu_count = prop.multiProcessorCount;
  }
  if (e == hipSuccess) e = hipStreamCreate(&h->stream);
  const size_t es = dtype == DM_FP64 ? sizeof(double) : sizeof(float);
  if (e == hipSuccess) e = hipMalloc(&h->d_a, h->N * es);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e != hipSuccess) {
    const int rc = fail(DM_ERR_HIP, "dm_create(%d qubits, %s): %s", n_qubits, dtype ? "fp32" : "fp64",
                        hipGetErrorString(e));
    dm_destroy(h);
    return rc;
  }
  *out = h;
  return DM_OK;
}

int dm_destroy(dm_state *h) {
  if (!h) return DM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->d_a) hipFree(h->d_a);
  if (h->d_red) hipFree(h->d_red);
  if (h->d_mat) hipFree(h->d_mat);
  if (h->d_aux) hipFree(h->d_aux);
  if (h->d_mn) hipFree(h->d_mn);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  for (auto &pp : h->pending) {
    hipEventDestroy(pp.a);
    hipEventDestroy(pp.b);
  }
  for (auto e : h->spare_events) hipEventDestroy(e);
  if (h->stream && h->owns_stream) hipStreamDestroy(h->stream);
  delete h;
  return DM_OK;
}

int dm_num_qubits(const dm_state *h, int *n) {
  if (!h || !n) return fail(DM_ERR_ARG, "null argument");
  *n = h->n;
  return DM_OK;
}

int dm_dtype(const dm_state *h, int *dtype) {
  if (!h || !dtype) return fail(DM_ERR_ARG, "null argument");
  *dtype = h->dtype;
  return DM_OK;
}

int dm_device_ptr(const dm_state *h, void **ptr) {
  if (int rc = check_handle(h)) return rc;
  if (!ptr) return fail(DM_ERR_ARG, "null argument");
  *ptr = h->d_a;
  return DM_OK;
}

int dm_init_product(dm_state *h, const double *per_qubit4, double scale) {
  if (int rc = check_handle(h)) return rc;
  if (!per_qubit4) return fail(DM_ERR_ARG, "null per-qubit vectors");
  // table row p = position p = qubit n-1-p
  double tab[4 * kMaxQubits];
  for (int p = 0; p < h->n; ++p)
    for (int c = 0; c < 4; ++c) tab[4 * p + c] = per_qubit4[4 * (h->n - 1 - p) + c];
  return DISPATCH(h, launch_diag_tab, h, tab, true, scale);
}

int dm_upload(dm_state *h, const double *host) {
  if (int rc = check_handle(h)) return rc;
  if (!host) return fail(DM_ERR_ARG, "null host buffer");
  ProfScope ps(h, 7);
  if (h->dtype == DM_FP64) {
    HIP_TRY(hipMemcpyAsync(h->d_a, host, h->N * sizeof(double), hipMemcpyHostToDevice, h->stream));
  } else {
    double *tmp = nullptr;
    HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&tmp), h->N * sizeof(double), h->stream));
    HIP_TRY(hipMemcpyAsync(tmp, host, h->N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(h, h->N)), dim3(kBlock), 0, h->stream, tmp,
                       static_cast<float *>(h->d_a), h->N);
    HIP_TRY(hipFreeAsync(tmp, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return last_launch("upload");
}

static int download_impl(dm_state *h, double *host) {
  if (h->dtype == DM_FP64) {
    HIP_TRY(hipMemcpyAsync(host, h->d_a, h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  } else {
    double *tmp = nullptr;
    HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&tmp), h->N * sizeof(double), h->stream));
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(h, h->N)), dim3(kBlock), 0, h->stream,
                       static_cast<const float *>(h->d_a), tmp, h->N);
    HIP_TRY(hipMemcpyAsync(host, tmp, h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipFreeAsync(tmp, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return last_launch("download");
}

int dm_download(dm_state *h, double *host) {
  if (int rc = check_handle(h)) return rc;
  if (!host) return fail(DM_ERR_ARG, "null host buffer");
  return download_impl(h, host);
}

int dm_download_chop(dm_state *h, double thr, double *host) {
  if (int rc = check_handle(h)) return rc;
  if (!host) return fail(DM_ERR_ARG, "null host buffer");
  if (thr > 0) {
    if (h->dtype == DM_FP64)
      hipLaunchKernelGGL((k_chop<double>), dim3(grid_for(h, h->N)), dim3(kBlock), 0, h->stream,
                         static_cast<double *>(h->d_a), h->N, thr);
    else
      hipLaunchKernelGGL((k_chop<float>), dim3(grid_for(h, h->N)), dim3(kBlock), 0, h->stream,
                         static_cast<float *>(h->d_a), h->N, (float)thr);
  }
  return download_impl(h, host);
}

int dm_read_coeffs(dm_state *h, const uint64_t *index, int k, double *out) {
  if (int rc = check_handle(h)) return rc;
  if (k < 0 || (k > 0 && (!index || !out))) return fail(DM_ERR_ARG, "bad index list");
  const size_t es = h->dtype == DM_FP64 ? 8 : 4;
  for (int i = 0; i < k; ++i) {
    if (index[i] >= h->N) return fail(DM_ERR_ARG, "coefficient index %llu out of range", (unsigned long long)index[i]);
    if (h->dtype == DM_FP64) {
      HIP_TRY(hipMemcpyAsync(out + i, static_cast<char *>(h->d_a) + index[i] * es, es, hipMemcpyDeviceToHost,
                             h->stream));
    } else {
      float f = 0.f;
      HIP_TRY(hipMemcpyAsync(&f, static_cast<char *>(h->d_a) + index[i] * es, es, hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      out[i] = f;
    }
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DM_OK;
}

int dm_apply_1q(dm_state *h, int q, const double R[12]) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = check_qubit(h, q)) return rc;
  if (!R) return fail(DM_ERR_ARG, "null matrix");
  return DISPATCH(h, launch_1q, h, q, R);
}

int dm_apply_cx(dm_state *h, int control, int target, const double *scale16) {
  if (int rc = check_handle(h)) return rc;
  if (control == target || control < 0 || target < 0 || control >= h->n || target >= h->n)
    return fail(DM_ERR_ARG, "Qubit Labels out of bound in CX Gate (%d, %d; n=%d)", control, target, h->n);
  return DISPATCH(h, launch_cx, h, control, target, scale16);
}

int dm_apply_2q(dm_state *h, int qa, int qb, const double R[256]) {
  if (int rc = check_handle(h)) return rc;
  if (qa == qb || qa < 0 || qb < 0 || qa >= h->n || qb >= h->n)
    return fail(DM_ERR_ARG, "bad qubit pair (%d, %d; n=%d)", qa, qb, h->n);
  if (!R) return fail(DM_ERR_ARG, "null matrix");
  return DISPATCH(h, launch_2q, h, qa, qb, R);
}

int dm_memory_noise(dm_state *h, double f, double p, double g) {
  if (int rc = check_handle(h)) return rc;
  if (f == 1.0 && g == 1.0) return DM_OK;  // identity (up to the sign of zeros)
  return DISPATCH(h, launch_memory_noise, h, f, p, g);
}

int dm_apply_diag(dm_state *h, const double *tab4n) {
  if (int rc = check_handle(h)) return rc;
  if (!tab4n) return fail(DM_ERR_ARG, "null table");
  double tab[4 * kMaxQubits];
  for (int p = 0; p < h->n; ++p)
    for (int c = 0; c < 4; ++c) tab[4 * p + c] = tab4n[4 * (h->n - 1 - p) + c];
  return DISPATCH(h, launch_diag_tab, h, tab, false, 1.0);
}

int dm_apply_block(dm_state *h, const int *qubits, int k, const double *S) {
  if (int rc = check_handle(h)) return rc;
  if (!qubits || !S) return fail(DM_ERR_ARG, "null argument");
  if (k != 3 && k != 4) return fail(DM_ERR_ARG, "block size k=%d not supported (3 or 4)", k);
  for (int j = 0; j < k; ++j) {
    if (int rc = check_qubit(h, qubits[j])) return rc;
    for (int i = 0; i < j; ++i)
      if (qubits[i] == qubits[j]) return fail(DM_ERR_ARG, "repeated qubit %d in block", qubits[j]);
  }
  const int L = (k == 3) ? 3 : 2;
  if (h->n < k + L) return fail(DM_ERR_ARG, "block of %d qubits needs n >= %d (n=%d)", k, k + L, h->n);
  return DISPATCH(h, launch_block, h, qubits, k, S);
}

int dm_probabilities(dm_state *h, const int *qubits, int m, char basis, const double *nvec, double err,
                     double *out) {
  if (int rc = check_handle(h)) return rc;
  if (m < 1 || m > h->n || !qubits || !out) return fail(DM_ERR_ARG, "bad measured-qubit list");
  if (basis != 'X' && basis != 'Y' && basis != 'Z' && basis != 'N') return fail(DM_ERR_ARG, "bad basis '%c'", basis);
  if (basis == 'N' && !nvec) return fail(DM_ERR_ARG, "basis N needs a direction");
  if (basis == 'N' && m > 13) return fail(DM_ERR_ARG, "basis N gathers 4^m values; m=%d too large", m);
  for (int t = 0; t < m; ++t) {
    if (int rc = check_qubit(h, qubits[t])) return rc;
    if (t && qubits[t] <= qubits[t - 1]) return fail(DM_ERR_ARG, "measured qubits must be ascending");
  }
  return DISPATCH(h, do_probabilities, h, qubits, m, basis, nvec, err, out);
}

int dm_trace_purity(dm_state *h, double *tr, double *purity) {
  if (int rc = check_handle(h)) return rc;
  if (tr) {
    uint64_t zero = 0;
    double a0 = 0;
    if (int rc = dm_read_coeffs(h, &zero, 1, &a0)) return rc;
    *tr = ldexp(a0, h->n);
  }
  if (purity) {
    double s = 0;
    if (int rc = DISPATCH(h, do_sumsq, h, &s)) return rc;
    *purity = ldexp(s, h->n);
  }
  return DM_OK;
}

int dm_overlap(dm_state *h, const double *other_host, double *out) {
  if (int rc = check_handle(h)) return rc;
  if (!other_host || !out) return fail(DM_ERR_ARG, "null argument");
  double *d_b = nullptr;
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&d_b), h->N * sizeof(double), h->stream));
  hipError_t e = hipMemcpyAsync(d_b, other_host, h->N * sizeof(double), hipMemcpyHostToDevice, h->stream);
  int rc = DM_OK;
  if (e != hipSuccess) rc = fail(DM_ERR_HIP, "overlap upload failed: %s", hipGetErrorString(e));
  double s = 0;
  if (rc == DM_OK) rc = DISPATCH(h, do_dot, h, d_b, &s);
  hipFreeAsync(d_b, h->stream);
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (rc == DM_OK) *out = ldexp(s, h->n);
  return rc;
}

static int check_range(const dm_state *h, uint64_t offset, uint64_t count, int host_dtype, const void *host) {
  if (int rc = check_handle(h)) return rc;
  if (host_dtype != DM_FP64 && host_dtype != DM_FP32) return fail(DM_ERR_ARG, "unknown host dtype %d", host_dtype);
  if (offset > h->N || count > h->N - offset)
    return fail(DM_ERR_ARG, "range [%llu, +%llu) outside the %llu coefficients", (unsigned long long)offset,
                (unsigned long long)count, (unsigned long long)h->N);
  if (count && !host) return fail(DM_ERR_ARG, "null host buffer");
  return DM_OK;
}

#define STATE_DISPATCH(fn, H, ...) ((h)->dtype == DM_FP64 ? fn<double, H>(__VA_ARGS__) : fn<float, H>(__VA_ARGS__))

int dm_download_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, void *host) {
  if (int rc = check_range(h, offset, count, host_dtype, host)) return rc;
  if (!count) return DM_OK;
  ProfScope ps(h, 7);
  if (host_dtype == DM_FP64)
    return STATE_DISPATCH(copy_range, double, h, offset, count, static_cast<double *>(host), true);
  return STATE_DISPATCH(copy_range, float, h, offset, count, static_cast<float *>(host), true);
}

int dm_upload_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, const void *host) {
  if (int rc = check_range(h, offset, count, host_dtype, host)) return rc;
  if (!count) return DM_OK;
  ProfScope ps(h, 7);
  if (host_dtype == DM_FP64)
    return STATE_DISPATCH(copy_range, double, h, offset, count,
                          const_cast<double *>(static_cast<const double *>(host)), false);
  return STATE_DISPATCH(copy_range, float, h, offset, count, const_cast<float *>(static_cast<const float *>(host)),
                        false);
}

int dm_dot_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, const void *host, double *sum) {
  if (int rc = check_range(h, offset, count, host_dtype, host)) return rc;
  if (!sum) return fail(DM_ERR_ARG, "null output");
  *sum = 0.0;
  if (!count) return DM_OK;
  if (host_dtype == DM_FP64)
    return STATE_DISPATCH(dot_range, double, h, offset, count, static_cast<const double *>(host), sum);
  return STATE_DISPATCH(dot_range, float, h, offset, count, static_cast<const float *>(host), sum);
}
#undef STATE_DISPATCH

int dm_to_dense(dm_state *h, double *out_c128) {
  if (int rc = check_handle(h)) return rc;
  if (!out_c128) return fail(DM_ERR_ARG, "null output");
  return DISPATCH(h, do_to_dense, h, out_c128);
}

int dm_run_program(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams) {
  if (int rc = check_handle(h)) return rc;
  if (nops < 0 || (nops > 0 && !ops)) return fail(DM_ERR_ARG, "bad op list");
  if (nparams < 0 || (nparams > 0 && !params)) return fail(DM_ERR_ARG, "bad parameter array");
  return DISPATCH(h, run_program, h, ops, nops, params, nparams);
}

int dm_run_program_noalloc(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams) {
  if (int rc = check_handle(h)) return rc;
  if (nops < 0 || (nops > 0 && !ops)) return fail(DM_ERR_ARG, "bad op list");
  if (nparams < 0 || (nparams > 0 && !params)) return fail(DM_ERR_ARG, "bad parameter array");
  return DISPATCH(h, run_program_noalloc, h, ops, nops, params, nparams);
}

int dm_batch_destroy(dm_batch *b) {
  if (!b) return DM_OK;
  if (b->master) hipStreamSynchronize(b->master);
  if (b->side) hipStreamSynchronize(b->side);
  if (b->side6) hipStreamSynchronize(b->side6);
  if (b->side7) hipStreamSynchronize(b->side7);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (void *p : {b->d_circs, b->d_tops, b->d_vals, b->d_arena, b->d_passes, b->d_items, b->d_qtops, b->d_qvals,
                  (void *)b->d_probs, (void *)b->d_stamps})
    if (p) hipFree(p);
  for (hipEvent_t e : {b->fork, b->join, b->ev0, b->ev1, b->forkr, b->join6, b->join7})
    if (e) hipEventDestroy(e);
  for (hipStream_t st : {b->side, b->side6, b->side7, b->master})
    if (st) hipStreamDestroy(st);
  delete b;
  return DM_OK;
}

int dm_batch_create(int dtype, int ncirc, const int *n_qubits, const int64_t *op_start, const dm_op *ops,
                    const double *params, int64_t nparams, const double *meas_err, int lds_max_qubits,
                    int lds_threads, int items_per_pass, dm_batch **out) {
  if (!out) return fail(DM_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (dtype != DM_FP64 && dtype != DM_FP32) return fail(DM_ERR_ARG, "unknown dtype %d", dtype);
  if (ncirc < 1 || !n_qubits || !op_start || !meas_err) return fail(DM_ERR_ARG, "bad circuit list");
  if (lds_threads != 256 && lds_threads != 512 && lds_threads != 1024)
    return fail(DM_ERR_ARG, "lds_threads must be 256, 512 or 1024");
  if (lds_max_qubits < kQueueTile - 1 || lds_max_qubits > 7)
    return fail(DM_ERR_ARG, "lds_max_qubits must lie in [%d, 7]", kQueueTile - 1);
  if (items_per_pass < 1 || items_per_pass > 65536) return fail(DM_ERR_ARG, "items_per_pass must lie in [1, 65536]");
  if (nparams < 0 || (nparams > 0 && !params)) return fail(DM_ERR_ARG, "bad parameter array");
  for (int c = 0; c < ncirc; ++c) {
    if (n_qubits[c] < 1 || n_qubits[c] > kQueueMaxQubits)
      return fail(DM_ERR_ARG, "circuit %d: %d qubits outside [1, %d]", c, n_qubits[c], kQueueMaxQubits);
    if (op_start[c + 1] < op_start[c] || (op_start[c + 1] > op_start[c] && !ops))
      return fail(DM_ERR_ARG, "circuit %d: bad op range", c);
  }
  dm_batch *b = new dm_batch();
  b->dtype = dtype;
  b->ncirc = ncirc;
  b->lds_threads = lds_threads;
  for (int c = 0; c < ncirc; ++c) {
    b->nq.push_back(n_qubits[c]);
    b->prob_off.push_back(b->nprobs);
    b->nprobs += (int64_t)1 << n_qubits[c];
    b->err.push_back(meas_err[c]);
  }
  const int rc = dtype == DM_FP64
                     ? batch_create_t<double>(b, op_start, ops, params, nparams, lds_max_qubits, items_per_pass)
                     : batch_create_t<float>(b, op_start, ops, params, nparams, lds_max_qubits, items_per_pass);
  if (rc) {
    dm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return DM_OK;
}

int dm_batch_info(const dm_batch *b, int64_t *hbm_circuits, int64_t *passes, int64_t *items, int64_t *launches) {
  if (!b) return fail(DM_ERR_ARG, "null batch");
  if (hbm_circuits) *hbm_circuits = (int64_t)b->large.size();
  if (passes) *passes = b->npasses;
  if (items) *items = b->nitems;
  if (launches) *launches = b->launches;
  return DM_OK;
}

int dm_batch_run(dm_batch *b, int use_graph, double *probs_out, double *circuit_us, double *total_ms) {
  if (!b) return fail(DM_ERR_ARG, "null batch");
  if (use_graph && !b->exec) {
    HIP_TRY(hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal));
    const int rc = b->dtype == DM_FP64 ? enqueue_batch<double>(b) : enqueue_batch<float>(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->master, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(DM_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(hipEventRecord(b->ev0, b->master));
  if (use_graph) {
    HIP_TRY(hipGraphLaunch(b->exec, b->master));
  } else {
    const int rc = b->dtype == DM_FP64 ? enqueue_batch<double>(b) : enqueue_batch<float>(b);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b->ev1, b->master));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (total_ms) *total_ms = ms;
  if (probs_out) HIP_TRY(hipMemcpy(probs_out, b->d_probs, b->nprobs * sizeof(double), hipMemcpyDeviceToHost));
  if (circuit_us) {
    std::vector<uint64_t> st(2 * (size_t)b->ncirc);
    HIP_TRY(hipMemcpy(st.data(), b->d_stamps, st.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    // stamps are stored LDS-tier circuits first, then HBM-tier circuits; report in circuit order
    std::vector<int> order(b->small);
    order.insert(order.end(), b->large.begin(), b->large.end());
    for (size_t j = 0; j < order.size(); ++j)
      circuit_us[order[j]] = (double)(st[2 * j + 1] - st[2 * j]) / b->clock_khz * 1e3;
  }
  return DM_OK;
}

int dm_compare(dm_state *a, dm_state *b, double *max_abs_diff, double *dot, double *norm2_a, double *norm2_b) {
  if (int rc = check_handle(a)) return rc;
  if (int rc = check_handle(b)) return rc;
  if (a->n != b->n) return fail(DM_ERR_ARG, "states of %d and %d qubits", a->n, b->n);
  double o[4];
  int rc;
  if (a->dtype == DM_FP64 && b->dtype == DM_FP64) rc = diff_t<double, double>(a, b, o);
  else if (a->dtype == DM_FP64) rc = diff_t<double, float>(a, b, o);
  else if (b->dtype == DM_FP64) rc = diff_t<float, double>(a, b, o);
  else rc = diff_t<float, float>(a, b, o);
  if (rc) return rc;
  if (max_abs_diff) *max_abs_diff = o[0];
  if (dot) *dot = o[1];
  if (norm2_a) *norm2_a = o[2];
  if (norm2_b) *norm2_b = o[3];
  return DM_OK;
}

int dm_synchronize(dm_state *h) {
  if (int rc = check_handle(h)) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return last_launch("stream");
}

int dm_profile(dm_state *h, int enable) {
  if (int rc = check_handle(h)) return rc;
  h->prof = enable != 0;
  return DM_OK;
}

int dm_profile_read(dm_state *h, int family, uint64_t *launches, double *ms) {
  if (int rc = check_handle(h)) return rc;
  if (family < 0 || family >= kFamilies) return fail(DM_ERR_ARG, "bad kernel family %d", family);
  if (int rc = resolve_profile(h)) return rc;
  if (launches) *launches = h->slots[family].launches;
  if (ms) *ms = h->slots[family].ms;
  return DM_OK;
}

int dm_profile_reset(dm_state *h) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = resolve_profile(h)) return rc;
  for (auto &s : h->slots) s = ProfSlot();
  return DM_OK;
}

}  // extern "C"
