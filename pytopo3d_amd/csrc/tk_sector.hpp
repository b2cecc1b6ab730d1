// Walsh-Hadamard sector form of the H8 element operator, shared by the
// streaming element kernel (tk_operator2.hip), the single-workgroup coarse
// V-cycle tail (tk_tail.hip) and the persistent coarse cycle (tk_pcycle.hip).  The 8 corner values of each displacement
// component are transformed per 2x2 face (index f = ox + 2 oy) and then
// between the bottom (z = 0) and top (z = 1) faces; in that basis KE splits
// into the 8 parity sectors described below, scaled by 1/64 on the host.
#pragma once

#include "tk_common.hpp"

namespace tk {

// Non-zero sector entries, 45 per KE, in this order (row-major within a block,
// rows and columns = the member components listed):
//   chi 0: comps {0,1,2} (9)  chi 1: {1,2} (4)  chi 2: {0,2} (4)
//   chi 3: {0,1} (4) + comp 2 alone (1)  chi 4: {0,1} (4)
//   chi 5: {0,2} (4) + comp 1 alone (1)  chi 6: {1,2} (4) + comp 0 alone (1)
//   chi 7: {0,1,2} (9)
constexpr int kSecN = 45;
constexpr int kSecOff[8] = {0, 9, 13, 17, 22, 26, 31, 36};

template <typename T>
struct SparseKE {
  T v[kSecN];
};

// In-plane transform of the 4 values of a face, index f = ox + 2 oy:
// butterflies (lo, hi) -> (lo + hi, hi - lo) along x, then along y.
template <typename T>
__device__ __forceinline__ void face_fwd(T v[4]) {
  T a = v[0] + v[1], b = v[1] - v[0], c = v[2] + v[3], d = v[3] - v[2];
  v[0] = a + c;
  v[2] = c - a;
  v[1] = b + d;
  v[3] = d - b;
}

// Inverse butterflies (sm, df) -> (sm - df, sm + df) along x, then along y.
template <typename T>
__device__ __forceinline__ void face_inv(T v[4]) {
  T a = v[0] - v[1], b = v[0] + v[1], c = v[2] - v[3], d = v[2] + v[3];
  v[0] = a - c;
  v[2] = a + c;
  v[1] = b - d;
  v[3] = b + d;
}

// Sector blocks: a[c][m] (spectrum, mode 0 unused) -> o[c][m] = s (B a)[c][m],
// with o[c][0] = 0 implied (not written).
template <typename T>
__device__ __forceinline__ void sector_apply(const T* __restrict__ k, T s, const T a[3][8],
                                             T o[3][8]) {
  // chi 0: members (0, 1), (1, 2), (2, 4)
  {
    const T* b = k + kSecOff[0];
    const T x0 = a[0][1], x1 = a[1][2], x2 = a[2][4];
    o[0][1] = s * (b[0] * x0 + b[1] * x1 + b[2] * x2);
    o[1][2] = s * (b[3] * x0 + b[4] * x1 + b[5] * x2);
    o[2][4] = s * (b[6] * x0 + b[7] * x1 + b[8] * x2);
  }
  // chi 1: (1, 3), (2, 5)   [(0, 0) is the x translation]
  {
    const T* b = k + kSecOff[1];
    const T x1 = a[1][3], x2 = a[2][5];
    o[1][3] = s * (b[0] * x1 + b[1] * x2);
    o[2][5] = s * (b[2] * x1 + b[3] * x2);
  }
  // chi 2: (0, 3), (2, 6)   [(1, 0) is the y translation]
  {
    const T* b = k + kSecOff[2];
    const T x0 = a[0][3], x2 = a[2][6];
    o[0][3] = s * (b[0] * x0 + b[1] * x2);
    o[2][6] = s * (b[2] * x0 + b[3] * x2);
  }
  // chi 3: (0, 2), (1, 1) coupled; (2, 7) alone
  {
    const T* b = k + kSecOff[3];
    const T x0 = a[0][2], x1 = a[1][1];
    o[0][2] = s * (b[0] * x0 + b[1] * x1);
    o[1][1] = s * (b[2] * x0 + b[3] * x1);
    o[2][7] = s * (b[4] * a[2][7]);
  }
  // chi 4: (0, 5), (1, 6)   [(2, 0) is the z translation]
  {
    const T* b = k + kSecOff[4];
    const T x0 = a[0][5], x1 = a[1][6];
    o[0][5] = s * (b[0] * x0 + b[1] * x1);
    o[1][6] = s * (b[2] * x0 + b[3] * x1);
  }
  // chi 5: (0, 4), (2, 1) coupled; (1, 7) alone
  {
    const T* b = k + kSecOff[5];
    const T x0 = a[0][4], x2 = a[2][1];
    o[0][4] = s * (b[0] * x0 + b[1] * x2);
    o[2][1] = s * (b[2] * x0 + b[3] * x2);
    o[1][7] = s * (b[4] * a[1][7]);
  }
  // chi 6: (1, 4), (2, 2) coupled; (0, 7) alone
  {
    const T* b = k + kSecOff[6];
    const T x1 = a[1][4], x2 = a[2][2];
    o[1][4] = s * (b[0] * x1 + b[1] * x2);
    o[2][2] = s * (b[2] * x1 + b[3] * x2);
    o[0][7] = s * (b[4] * a[0][7]);
  }
  // chi 7: (0, 6), (1, 5), (2, 3)
  {
    const T* b = k + kSecOff[7];
    const T x0 = a[0][6], x1 = a[1][5], x2 = a[2][3];
    o[0][6] = s * (b[0] * x0 + b[1] * x1 + b[2] * x2);
    o[1][5] = s * (b[3] * x0 + b[4] * x1 + b[5] * x2);
    o[2][3] = s * (b[6] * x0 + b[7] * x1 + b[8] * x2);
  }
}

// Index of sector entry (chi, row component c, column component c') in the
// 45 entries of tk_sector.hpp, -1 where the block is zero.
static __constant__ signed char kSecIdx[8][3][3] = {
    {{0, 1, 2}, {3, 4, 5}, {6, 7, 8}},
    {{-1, -1, -1}, {-1, 9, 10}, {-1, 11, 12}},
    {{13, -1, 14}, {-1, -1, -1}, {15, -1, 16}},
    {{17, 18, -1}, {19, 20, -1}, {-1, -1, 21}},
    {{22, 23, -1}, {24, 25, -1}, {-1, -1, -1}},
    {{26, -1, 27}, {-1, 30, -1}, {28, -1, 29}},
    {{35, -1, -1}, {-1, 31, 32}, {-1, 33, 34}},
    {{36, 37, 38}, {39, 40, 41}, {42, 43, 44}}};

// Value of lane (l ^ M) within each group of 8 lanes, M in 1..7: a DPP quad
// permutation for M < 4; the half-row mirror (l ^ 7) followed by the quad
// permutation of M ^ 7 for M >= 4.  Every lane of the wave must be active.
template <int M>
__device__ __forceinline__ int lane_xor_i(int v) {
  static_assert(M > 0 && M < 8, "lanes of one 8-lane group");
  constexpr int kQuad[4] = {0xE4, 0xB1, 0x4E, 0x1B};   // quad_perm: l, l ^ 1, l ^ 2, l ^ 3
  if constexpr (M >= 4) {
    v = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    if constexpr ((M ^ 7) != 0) v = __builtin_amdgcn_mov_dpp(v, kQuad[M ^ 7], 0xF, 0xF, false);
    return v;
  } else {
    return __builtin_amdgcn_mov_dpp(v, kQuad[M], 0xF, 0xF, false);
  }
}
template <int M>
__device__ __forceinline__ float lane_xor(float v) {
  return __builtin_bit_cast(float, lane_xor_i<M>(__builtin_bit_cast(int, v)));
}
template <int M>
__device__ __forceinline__ double lane_xor(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)lane_xor_i<M>((int)(unsigned)u);
  const unsigned hi = (unsigned)lane_xor_i<M>((int)(unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// One butterfly of the corner transform across the lanes of an element, on
// the corner bit M: (lo, hi) -> (lo + hi, hi - lo); the inverse
// (sm, df) -> (sm - df, sm + df).  The same butterflies as face_fwd / face_inv
// and the z butterfly of the element kernel.
template <int M, typename T>
__device__ __forceinline__ T wht_fwd(T v, bool hi) {
  const T p = lane_xor<M>(v);
  return hi ? v - p : v + p;
}
template <int M, typename T>
__device__ __forceinline__ T wht_inv(T v, bool hi) {
  const T p = lane_xor<M>(v);
  return hi ? p + v : v - p;
}

// Lane j's sector rows for the 8-lane element product: output component c,
// input component c' of the block that couples mode j of c with mode
// j ^ e_c ^ e_c' of c' (0 where the block is zero).
template <typename T>
__device__ __forceinline__ void sec8_rows(const T* sec, int j, T kc[3][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int c2 = 0; c2 < 3; ++c2) {
      const int k = kSecIdx[j ^ (1 << c)][c][c2];
      kc[c][c2] = k >= 0 ? sec[k] : T(0);
    }
}

// The element product s KE u over 8 lanes, lane j holding corner j (j = ox +
// 2 oy + 4 oz) of the element: u, its (masked) corner values, become o, that
// corner's outputs.  Three butterflies across the 8 lanes give the
// Walsh-Hadamard modes (lane j holds mode j of each component), the sector
// blocks couple mode j of component c with modes j ^ e_c ^ e_c' of the others
// (three lane exchanges), and the inverse butterflies return to corners.
// Every lane of the wave must take part.
template <typename T>
__device__ __forceinline__ void sec8_apply(const T kc[3][3], T s, int j, const T u_in[3], T o[3]) {
  const bool h1 = j & 1, h2 = j & 2, h4 = j & 4;
  T u[3] = {u_in[0], u_in[1], u_in[2]};
#pragma unroll
  for (int c = 0; c < 3; ++c) u[c] = wht_fwd<1>(u[c], h1);
#pragma unroll
  for (int c = 0; c < 3; ++c) u[c] = wht_fwd<2>(u[c], h2);
#pragma unroll
  for (int c = 0; c < 3; ++c) u[c] = wht_fwd<4>(u[c], h4);
  const T g30 = lane_xor<3>(u[0]), g31 = lane_xor<3>(u[1]);
  const T g50 = lane_xor<5>(u[0]), g52 = lane_xor<5>(u[2]);
  const T g61 = lane_xor<6>(u[1]), g62 = lane_xor<6>(u[2]);
  o[0] = s * (kc[0][0] * u[0] + kc[0][1] * g31 + kc[0][2] * g52);
  o[1] = s * (kc[1][0] * g30 + kc[1][1] * u[1] + kc[1][2] * g62);
  o[2] = s * (kc[2][0] * g50 + kc[2][1] * g61 + kc[2][2] * u[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = wht_inv<1>(o[c], h1);
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = wht_inv<2>(o[c], h2);
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = wht_inv<4>(o[c], h4);
}

// Reciprocal of the on-the-fly diagonal: the hardware reciprocal in FP32
// (1 ulp; it only scales the smoother's correction), a rounded division in FP64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

}  // namespace tk
