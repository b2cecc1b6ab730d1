// Walsh-Hadamard sector form of the H8 element operator, shared by the
// streaming element kernel (tk_operator2.hip) and the single-workgroup coarse
// V-cycle tail (tk_tail.hip).  The 8 corner values of each displacement
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

// Reciprocal of the on-the-fly diagonal: the hardware reciprocal in FP32
// (1 ulp; it only scales the smoother's correction), a rounded division in FP64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

}  // namespace tk
