// Element-centric matrix-free H8 operator (the fast path of K u).
//
// KE of an axis-aligned orthotropic cube commutes with the three mirror
// reflections of the element (the reference checks this invariance,
// tests/test_stiffness.py:69-78).  In the Walsh-Hadamard basis of the 8 corner
// values of each displacement component the 24x24 KE therefore splits into 8
// independent 3x3 blocks, one per parity sector chi: the members of sector chi
// are (component c, mode m = chi XOR e_c).  Per element:
//     a = W^T u_e   (3 x 8-point butterflies, 72 add/sub)
//     a'_chi = s_e B_chi a_chi   (8 x 3x3 blocks, 72 FMA + 24 mul)
//     y_e = W a'    (72 add/sub)
// with B = W^T KE W / 64 built and checked on the host.  This is about 240
// operations per element instead of 1152 for the dense 24x24 product.
//
// Scatter without atomics: a workgroup owns a 15 x 15 column of nodes in (y, x)
// and streams along z.  Its 16 x 16 threads evaluate the 16 x 16 elements
// touching those nodes (one halo row and column is re-evaluated by the
// neighbouring tile), exchange the six off-diagonal corner contributions of
// each element through LDS, and every node thread sums its 8 contributions in
// a fixed order, so results are deterministic.  The top-plane partial sums of
// one element layer stay in registers as the bottom-plane start of the next.
#include "tk_plan.hpp"

namespace tk {

template <typename T>
struct SectorKE {
  T b[8][9];   // b[chi][3*c + c'], sector member c has mode chi ^ (1 << c)
};

template <typename T>
struct Epilogue {
  T* y;            // MODE 0/1 output
  const T* b;      // MODE 1/2/3 right-hand side
  T* xo;           // MODE 2/3 updated iterate
  T* d;            // MODE 2/3 Chebyshev direction
  const T* dinv;   // MODE 2/3 inverse diagonal
  T c0, c1, c2;
};

enum { kModeApply = 0, kModeResidual = 1, kModeChebFirst = 2, kModeChebStep = 3 };

constexpr int kTile = 16;          // threads per tile edge
constexpr int kOut = kTile - 1;    // owned nodes per tile edge
constexpr int kPl = kTile + 1;     // node-plane edge held in LDS

// forward butterflies on 8 values in canonical corner order t = ox + 2 oy + 4 oz
template <typename T>
__device__ __forceinline__ void wht_fwd(T v[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (!(t & b)) {
        const T lo = v[t], hi = v[t | b];
        v[t] = lo + hi;
        v[t | b] = hi - lo;
      }
}

template <typename T>
__device__ __forceinline__ void wht_inv(T v[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (!(t & b)) {
        const T sm = v[t], df = v[t | b];
        v[t] = sm - df;
        v[t | b] = sm + df;
      }
}

// canonical index t of corner a, and back
__device__ __forceinline__ constexpr int t_of(int a) {
  return corner_x(a) + 2 * corner_y(a) + 4 * corner_z(a);
}

// Element product: u[c][t] -> y[c][t] = s * (KE u)[c][t]
template <typename T>
__device__ __forceinline__ void element_apply(const T* __restrict__ kb, T s, T u[3][8]) {
  wht_fwd(u[0]);
  wht_fwd(u[1]);
  wht_fwd(u[2]);
  T out[3][8];
#pragma unroll
  for (int chi = 0; chi < 8; ++chi) {
    const T a0 = u[0][chi ^ 1], a1 = u[1][chi ^ 2], a2 = u[2][chi ^ 4];
    const T* b = kb + 9 * chi;
    out[0][chi ^ 1] = s * (b[0] * a0 + b[1] * a1 + b[2] * a2);
    out[1][chi ^ 2] = s * (b[3] * a0 + b[4] * a1 + b[5] * a2);
    out[2][chi ^ 4] = s * (b[6] * a0 + b[7] * a1 + b[8] * a2);
  }
  wht_inv(out[0]);
  wht_inv(out[1]);
  wht_inv(out[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < 8; ++t) u[c][t] = out[c][t];
}

// Corner slots exchanged through LDS: t = 2 (x=0,y=1), 3 (x=1,y=1), 1 (x=1,y=0)
// for the bottom plane, and the same +4 for the top plane.
__device__ __forceinline__ constexpr int slot_of_t(int t) {
  // t: 1 -> 0, 2 -> 1, 3 -> 2, 5 -> 3, 6 -> 4, 7 -> 5
  return (t & 3) - 1 + 3 * (t >> 2);
}

template <typename T>
struct PlaneRegs {
  T own[3];      // own node, unmasked
  uint8_t fix;   // own node fixed bits (7 outside the domain)
  T ext[3];      // extra halo node, masked
};

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_apply_el(Grid g, SectorKE<float> Kv,
                                                  const T* __restrict__ Kg,
                                                  const T* __restrict__ s,
                                                  const uint8_t* __restrict__ fixed,
                                                  const T* __restrict__ u, Epilogue<T> ep,
                                                  int nty, int zc, int with_dot,
                                                  const int* __restrict__ done, double* partials,
                                                  unsigned* counter, Scalars* sc) {
  if (done && *done) return;
  __shared__ T un[2][3][kPl * kPl];       // two node planes of the tile (masked u)
  __shared__ T yo[6][3][kTile * kTile];   // exchanged corner contributions
  __shared__ T kb[72];                    // sector blocks of KE

  const int tid = threadIdx.x;
  // FP64 blocks are staged in LDS (144 SGPRs would spill); FP32 blocks stay in
  // scalar registers (72 SGPRs), which measured faster than LDS broadcasts.
  if (sizeof(T) == 8 && tid < 72) kb[tid] = Kg[tid];
  const int ly = tid & (kTile - 1), lx = tid >> 4;
  // XCD-aware order: z-chunks of a tile, then y, then x neighbours share an L2
  const int lg = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int chunk = lg % gridDim.y, tile = lg / gridDim.y;
  const int Y0 = (tile % nty) * kOut, X0 = (tile / nty) * kOut;
  const int Z0 = chunk * zc;
  const int Z1 = min(Z0 + zc, g.nnz);
  const int ey = Y0 - 1 + ly, ex = X0 - 1 + lx;   // element == owned node (tile coords)
  const bool col_el = (unsigned)ey < (unsigned)g.nely && (unsigned)ex < (unsigned)g.nelx;
  const bool col_node = (unsigned)ey < (unsigned)g.nny && (unsigned)ex < (unsigned)g.nnx;
  const bool owner = ly >= 1 && lx >= 1 && col_node;
  const int nxy = g.nnx * g.nny;
  const int exy = g.nelx * g.nely;

  // extra plane nodes loaded by threads 0..32: row i = 16 (j = 0..16), column j = 16 (i = 0..15)
  int xi = -1, xj = -1;
  if (tid < kPl) { xi = kTile; xj = tid; }
  else if (tid < kPl + kTile) { xi = tid - kPl; xj = kTile; }
  const int xny = Y0 - 1 + xi, xnx = X0 - 1 + xj;
  const bool ext_node = xi >= 0 && (unsigned)xny < (unsigned)g.nny && (unsigned)xnx < (unsigned)g.nnx;

  // global loads of node plane pz (registers only)
  auto fetch = [&](int pz, PlaneRegs<T>& R) {
    R.own[0] = R.own[1] = R.own[2] = T(0);
    R.fix = 7;
    if (col_node && (unsigned)pz < (unsigned)g.nnz) {
      const int n = ey + ex * g.nny + pz * nxy;
      R.fix = fixed[n];
      R.own[0] = u[3 * n];
      R.own[1] = u[3 * n + 1];
      R.own[2] = u[3 * n + 2];
    }
    R.ext[0] = R.ext[1] = R.ext[2] = T(0);
    if (ext_node && (unsigned)pz < (unsigned)g.nnz) {
      const int n = xny + xnx * g.nny + pz * nxy;
      const uint8_t f = fixed[n];
      R.ext[0] = (f & 1) ? T(0) : u[3 * n];
      R.ext[1] = (f & 2) ? T(0) : u[3 * n + 1];
      R.ext[2] = (f & 4) ? T(0) : u[3 * n + 2];
    }
  };
  auto store = [&](int buf, const PlaneRegs<T>& R) {
#pragma unroll
    for (int c = 0; c < 3; ++c) un[buf][c][ly + kPl * lx] = ((R.fix >> c) & 1) ? T(0) : R.own[c];
    if (xi >= 0)
#pragma unroll
      for (int c = 0; c < 3; ++c) un[buf][c][xi + kPl * xj] = R.ext[c];
  };

  T ue[3][8];
  T part[3] = {T(0), T(0), T(0)};   // partial sum of the node at the current top plane
  double dot = 0.0;

  PlaneRegs<T> cur, nxt;
  fetch(Z0 - 1, cur);
  fetch(Z0, nxt);
  store(0, cur);
  store(1, nxt);
  T raw_prev[3] = {cur.own[0], cur.own[1], cur.own[2]};
  uint8_t fix_prev = cur.fix;
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int oy = (t >> 1) & 1, ox = t & 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[c][t] = un[0][c][(ly + oy) + kPl * (lx + ox)];
  }
  int buf = 1;   // buffer holding plane ez + 1

  for (int ez = Z0 - 1; ez < Z1; ++ez) {
    // issue the global loads of the next iteration and of this epilogue early
    const bool more = ez + 1 < Z1;
    PlaneRegs<T> pre;
    if (more) fetch(ez + 2, pre);
    const bool out_now = owner && ez >= Z0;
    const int n_out = ey + ex * g.nny + ez * nxy;
    T eb[3] = {T(0), T(0), T(0)}, ed[3] = {T(0), T(0), T(0)}, edi[3] = {T(0), T(0), T(0)};
    if (out_now && MODE != kModeApply) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        eb[c] = ep.b[3 * n_out + c];
        if (MODE == kModeChebFirst || MODE == kModeChebStep) edi[c] = ep.dinv[3 * n_out + c];
        if (MODE == kModeChebStep) ed[c] = ep.d[3 * n_out + c];
      }
    }
    const bool el = col_el && (unsigned)ez < (unsigned)g.nelz;
    const T se = el ? s[ey + ex * g.nely + ez * exy] : T(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int oy = (t >> 1) & 1, ox = t & 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) ue[c][t + 4] = un[buf][c][(ly + oy) + kPl * (lx + ox)];
    }
    T w[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int t = 0; t < 8; ++t) w[c][t] = ue[c][t];
    if constexpr (sizeof(T) == 8) element_apply<T>(kb, se, w);
    else element_apply<T>(&Kv.b[0][0], se, w);   // se = 0 outside the domain gives zero contributions
#pragma unroll
    for (int t = 1; t < 8; ++t) {
      if (t == 4) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) yo[slot_of_t(t)][c][tid] = w[c][t];
    }
    __syncthreads();
    // node (ly, lx) at the bottom plane ez: own corner t=0, (ly-1,lx) t=2,
    // (ly,lx-1) t=1, (ly-1,lx-1) t=3; top plane: the same + 4.
    if (owner) {
      const int tym = tid - 1, txm = tid - kTile, tbm = tid - kTile - 1;
      T bot[3], top[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bot[c] = part[c] + w[c][0] + yo[slot_of_t(2)][c][tym] + yo[slot_of_t(1)][c][txm] +
                 yo[slot_of_t(3)][c][tbm];
        top[c] = w[c][4] + yo[slot_of_t(6)][c][tym] + yo[slot_of_t(5)][c][txm] +
                 yo[slot_of_t(7)][c][tbm];
      }
      if (out_now) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int i = 3 * n_out + c;
          const T val = ((fix_prev >> c) & 1) ? raw_prev[c] : bot[c];
          if (MODE == kModeApply) {
            ep.y[i] = val;
            dot += (double)raw_prev[c] * (double)val;
          } else if (MODE == kModeResidual) {
            ep.y[i] = eb[c] - val;
          } else if (MODE == kModeChebFirst) {
            const T dd = ep.c0 * edi[c] * (eb[c] - val);
            ep.d[i] = dd;
            ep.xo[i] = raw_prev[c] + dd;
          } else {
            const T dd = ep.c1 * ed[c] + ep.c2 * edi[c] * (eb[c] - val);
            ep.d[i] = dd;
            ep.xo[i] = raw_prev[c] + dd;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) part[c] = top[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int t = 0; t < 4; ++t) ue[c][t] = ue[c][t + 4];
      raw_prev[c] = nxt.own[c];
    }
    fix_prev = nxt.fix;
    if (more) {
      store(buf ^ 1, pre);
      nxt = pre;
    }
    __syncthreads();
    buf ^= 1;
  }
  if (MODE == kModeApply && with_dot) {
    double v[1] = {dot};
    block_partials<1>(v, partials);
  }
}

// ---------------------------------------------------------------------------
// Host: sector decomposition of KE.  Returns false when KE is not
// reflection-symmetric (then the dense node kernel is used).
static bool build_sectors(const double* KE, double out[8][9]) {
  // W[(a,c),(m,c)] = prod_d sigma_{a,d}^{m_d}, sigma = -1 (offset 0) / +1 (offset 1)
  auto psi = [](int a, int m) {
    const int o[3] = {corner_x(a), corner_y(a), corner_z(a)};
    double v = 1.0;
    for (int d = 0; d < 3; ++d)
      if ((m >> d) & 1) v *= o[d] ? 1.0 : -1.0;
    return v;
  };
  // B[(m,c),(m',c')] = sum_{a,b} psi(a,m) KE[(a,c),(b,c')] psi(b,m') / 64
  static double B[24][24];
  double bmax = 0.0;
  for (int m = 0; m < 8; ++m)
    for (int c = 0; c < 3; ++c)
      for (int m2 = 0; m2 < 8; ++m2)
        for (int c2 = 0; c2 < 3; ++c2) {
          double acc = 0.0;
          for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
              acc += psi(a, m) * KE[(3 * a + c) * 24 + 3 * b + c2] * psi(b, m2);
          B[3 * m + c][3 * m2 + c2] = acc / 64.0;
          bmax = fmax(bmax, fabs(acc / 64.0));
        }
  // sector membership: (c, m) belongs to chi = m ^ (1 << c)
  double off = 0.0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      const int ci = i % 3, mi = i / 3, cj = j % 3, mj = j / 3;
      if ((mi ^ (1 << ci)) != (mj ^ (1 << cj))) off = fmax(off, fabs(B[i][j]));
    }
  if (off > 1e-12 * bmax) return false;
  for (int chi = 0; chi < 8; ++chi)
    for (int c = 0; c < 3; ++c)
      for (int c2 = 0; c2 < 3; ++c2) {
        const int m = chi ^ (1 << c), m2 = chi ^ (1 << c2);
        out[chi][3 * c + c2] = B[3 * m + c][3 * m2 + c2];
      }
  return true;
}

bool operator_setup(Plan& P) {
  double sec[8][9];
  P.sector_ok = build_sectors(P.KE, sec);
  for (int chi = 0; chi < 8; ++chi)
    for (int k = 0; k < 9; ++k) {
      P.sec64[chi][k] = sec[chi][k];
      P.sec32[chi][k] = (float)sec[chi][k];
    }
  if (hipMalloc((void**)&P.sec64_dev, sizeof(double) * 72) != hipSuccess ||
      hipMalloc((void**)&P.sec32_dev, sizeof(float) * 72) != hipSuccess ||
      hipMemcpy(P.sec64_dev, P.sec64, sizeof(double) * 72, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P.sec32_dev, P.sec32, sizeof(float) * 72, hipMemcpyHostToDevice) != hipSuccess)
    P.sector_ok = false;
  return P.sector_ok;
}

template <typename T> static const T* sectors_of(const Plan& P);
template <> const double* sectors_of<double>(const Plan& P) { return P.sec64_dev; }
template <> const float* sectors_of<float>(const Plan& P) { return P.sec32_dev; }

static void el_grid(const Grid& g, int& nty, int& zc, dim3& grid) {
  nty = (g.nny + kOut - 1) / kOut;
  const int ntx = (g.nnx + kOut - 1) / kOut;
  // z chunks: enough workgroups to fill the chip, few re-evaluated halo layers
  const int tiles = nty * ntx;
  int nch = (2048 + tiles - 1) / tiles;
  if (nch > g.nnz) nch = g.nnz;
  if (nch < 1) nch = 1;
  int zc0 = (g.nnz + nch - 1) / nch;
  if (zc0 < 8 && g.nnz >= 8) zc0 = 8;
  nch = (g.nnz + zc0 - 1) / zc0;
  zc = (g.nnz + nch - 1) / nch;
  grid = dim3(tiles, nch);
}

template <typename T, int MODE>
static void launch_el(Plan& P, const Level& L, const T* sl, const T* u, const Epilogue<T>& ep,
                      int with_dot, const int* done, hipStream_t s) {
  int nty, zc;
  dim3 grid;
  el_grid(L.g, nty, zc, grid);
  // the reduction scratch holds 4 values per block of the largest 1-D launch
  while ((long long)grid.x * grid.y > P.partials_len / 4 && grid.y > 1) {
    const int nch = (int)grid.y - 1;
    zc = (L.g.nnz + nch - 1) / nch;
    grid.y = (L.g.nnz + zc - 1) / zc;
  }
  SectorKE<float> kv;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 9; ++j) kv.b[i][j] = P.sec32[i][j];
  hipLaunchKernelGGL((k_apply_el<T, MODE>), grid, dim3(256), 0, s, L.g, kv, sectors_of<T>(P), sl,
                     L.fixed, u, ep, nty, zc, with_dot, done, P.partials, P.counter, P.sc);
  if (MODE == kModeApply && with_dot) finalize(P, kFinAlpha, (int)(grid.x * grid.y), s);
}

template <typename T>
void el_apply(Plan& P, const Level& L, const T* sl, const T* u, T* y, int with_dot,
              const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.y = y;
  launch_el<T, kModeApply>(P, L, sl, u, ep, with_dot, done, s);
}

template <typename T>
void el_residual(Plan& P, const Level& L, const T* sl, const T* b, const T* x, T* r,
                 const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.y = r;
  ep.b = b;
  launch_el<T, kModeResidual>(P, L, sl, x, ep, 0, done, s);
}

template <typename T>
void el_cheb(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, const T* x, T* xo,
             T* d, double c0, double c1, double c2, int first, const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.b = b;
  ep.xo = xo;
  ep.d = d;
  ep.dinv = dinv;
  ep.c0 = (T)c0;
  ep.c1 = (T)c1;
  ep.c2 = (T)c2;
  if (first) launch_el<T, kModeChebFirst>(P, L, sl, x, ep, 0, done, s);
  else launch_el<T, kModeChebStep>(P, L, sl, x, ep, 0, done, s);
}

#define TK_INST2(T)                                                                              \
  template void el_apply<T>(Plan&, const Level&, const T*, const T*, T*, int, const int*,       \
                            hipStream_t);                                                       \
  template void el_residual<T>(Plan&, const Level&, const T*, const T*, const T*, T*,           \
                               const int*, hipStream_t);                                        \
  template void el_cheb<T>(Plan&, const Level&, const T*, const T*, const T*, const T*, T*, T*, \
                           double, double, double, int, const int*, hipStream_t);
TK_INST2(double)
TK_INST2(float)
#undef TK_INST2

}  // namespace tk
