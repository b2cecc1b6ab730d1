// Single-workgroup tail of the V-cycle: on the second-coarsest level the
// pre-smoothing, the residual and its restriction; the dense coarsest solve;
// then the prolongation and the post-smoothing back on the second-coarsest
// level.  Those two levels hold a few hundred nodes (9x5x5 and 5x3x3 at
// 256x128x128): level by level they cost a dozen launches per V-cycle of ~5 us
// each, nearly all launch and memory latency.  Here one workgroup keeps the
// level state in LDS and runs all of it in one launch, separated by barriers.
//
// The arithmetic is that of the level-by-level kernels: K u in the
// Walsh-Hadamard sector form of the element kernel (one thread per element
// writes its 24 corner outputs to LDS, one thread per node sums its <= 8 in a
// fixed order; inputs masked at fixed DOFs, identity rows there), the
// Chebyshev steps d = c1 d + c2 D^-1 (b - K x), x' = x + d, the trilinear
// restriction and prolongation of tk_transfer.hpp, and the stored dense
// inverse of the coarsest operator.  (A node-gather operator with the dense
// KE, one node per thread, took ~10 us per step on one compute unit.)
#include "tk_plan.hpp"
#include "tk_sector.hpp"
#include "tk_transfer.hpp"

#include <type_traits>

namespace tk {

constexpr int kTailThreads = 256;

// LDS layout of the tail for level g and coarsest grid gc (element type T):
// sector entries, x (two buffers), d, b, D^-1, element scales, the element
// outputs (24 per element), coarse right-hand side and solution, optionally a
// copy of the coarsest inverse (16-byte aligned), then the fixed bits of both
// grids.  Offsets in elements of T.
template <typename T>
__host__ __device__ inline size_t tail_inv_offset(const Grid& g, const Grid& gc) {
  const size_t vals = 48 + 5 * (size_t)g.ndof + g.nele + 24 * (size_t)g.nele + 2 * (size_t)gc.ndof;
  constexpr size_t a = 16 / sizeof(T);
  return (vals + a - 1) / a * a;
}

template <typename T>
static size_t tail_lds(const Grid& g, const Grid& gc, bool inv_lds) {
  const size_t vals = tail_inv_offset<T>(g, gc) + (inv_lds ? (size_t)gc.ndof * gc.ndof : 0);
  return vals * sizeof(T) + (size_t)g.nnode + (size_t)gc.nnode;
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_tail(
    Grid g, Grid gc, const T* __restrict__ sec, const T* __restrict__ sl,
    const uint8_t* __restrict__ fixed, const uint8_t* __restrict__ fixed_c,
    const T* __restrict__ dinv, const T* __restrict__ b, const T* __restrict__ inv,
    T* __restrict__ xout, TailCoef cf, int inv_lds, const int* __restrict__ done) {
  if (*done) return;
  extern __shared__ __align__(16) unsigned char tail_mem[];
  T* kb = reinterpret_cast<T*>(tail_mem);   // 45 sector entries (48 slots)
  // x, two buffers; chosen by offset (a runtime-indexed pointer array loses
  // the LDS address space and turns every access into a flat one)
  T* X0 = kb + 48;
  auto X = [&](int w) -> T* { return X0 + (w ? g.ndof : 0); };
  T* D = X0 + 2 * g.ndof;
  T* B = D + g.ndof;
  T* DI = B + g.ndof;
  T* S = DI + g.ndof;
  T* EY = S + g.nele;          // element outputs: [element][corner f + 4 oz][component]
  T* BC = EY + 24 * g.nele;
  T* XC = BC + gc.ndof;
  T* IV = kb + tail_inv_offset<T>(g, gc);   // the coarsest inverse, when inv_lds
  uint8_t* F = reinterpret_cast<uint8_t*>(IV + (inv_lds ? gc.ndof * gc.ndof : 0));
  uint8_t* FC = F + g.nnode;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T c0 = (T)cf.c0;
  const int n2 = gc.ndof * gc.ndof;
  if (inv_lds) {
    // 16-byte copies, several in flight per thread
    const int n16 = n2 * (int)sizeof(T) / 16;
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(inv);
    uint4* dst = reinterpret_cast<uint4*>(IV);
#pragma unroll 8
    for (int i = tid; i < n16; i += nt) dst[i] = src[i];
    for (int i = n16 * (16 / (int)sizeof(T)) + tid; i < n2; i += nt) IV[i] = inv[i];
  }
  for (int i = tid; i < kSecN; i += nt) kb[i] = sec[i];
#pragma unroll 4
  for (int i = tid; i < g.ndof; i += nt) {
    B[i] = b[i];
    DI[i] = dinv[i];
  }
  for (int i = tid; i < g.nele; i += nt) S[i] = sl[i];
  for (int i = tid; i < g.nnode; i += nt) F[i] = fixed[i];
  for (int i = tid; i < gc.nnode; i += nt) FC[i] = fixed_c[i];
  __syncthreads();

  const int nxy = g.nnx * g.nny, exy = g.nelx * g.nely;
  // EY = the element products s_e KE u_e (inputs masked at fixed DOFs), in the
  // Walsh-Hadamard sector form of tk_sector.hpp: per component the corner
  // values of the bottom (oz = 0) and top face, index f = ox + 2 oy, are
  // face-transformed and combined, the sector blocks applied, and the outputs
  // transformed back; corner (ox, oy, oz) is EY[e][f + 4 oz]
  auto element_products = [&](const T* u) {
    for (int e = tid; e < g.nele; e += nt) {
      const int ey = e % g.nely, t = e / g.nely, ex = t % g.nelx, ez = t / g.nelx;
      T Fb[3][4], Ft[3][4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const int m0 = (ey + (f >> 1)) + (ex + (f & 1)) * g.nny + ez * nxy;
        const int m1 = m0 + nxy;
        const uint8_t f0 = F[m0], f1 = F[m1];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Fb[c][f] = ((f0 >> c) & 1) ? T(0) : u[3 * m0 + c];
          Ft[c][f] = ((f1 >> c) & 1) ? T(0) : u[3 * m1 + c];
        }
      }
      T a[3][8], o[3][8];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        face_fwd(Fb[c]);
        face_fwd(Ft[c]);
        a[c][0] = T(0);
#pragma unroll
        for (int f = 1; f < 4; ++f) a[c][f] = Fb[c][f] + Ft[c][f];
#pragma unroll
        for (int f = 0; f < 4; ++f) a[c][f + 4] = Ft[c][f] - Fb[c][f];
      }
      sector_apply<T>(kb, S[e], a, o);
      T* y = EY + 24 * e;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T nb[4], nt4[4];
        nb[0] = -o[c][4];
        nt4[0] = o[c][4];
#pragma unroll
        for (int f = 1; f < 4; ++f) {
          nb[f] = o[c][f] - o[c][f + 4];
          nt4[f] = o[c][f] + o[c][f + 4];
        }
        face_inv(nb);
        face_inv(nt4);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          y[3 * f + c] = nb[f];
          y[3 * (f + 4) + c] = nt4[f];
        }
      }
    }
    __syncthreads();
  };
  // (K u)[node n]: the <= 8 element outputs at n, summed in a fixed order
  auto node_sum = [&](int n, T acc[3]) {
    int iy, ix, iz;
    grid_node_coords(g, n, iy, ix, iz);
    acc[0] = acc[1] = acc[2] = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int oy = (k >> 1) & 1, ox = k & 1, oz = k >> 2;
      const int ey = iy - oy, ex = ix - ox, ez = iz - oz;
      if ((unsigned)ey >= (unsigned)g.nely || (unsigned)ex >= (unsigned)g.nelx ||
          (unsigned)ez >= (unsigned)g.nelz)
        continue;
      const T* y = EY + 24 * (ey + ex * g.nely + ez * exy) + 3 * ((ox + 2 * oy) + 4 * oz);
      acc[0] += y[0];
      acc[1] += y[1];
      acc[2] += y[2];
    }
  };

  int cur = 0;
  // one Chebyshev step from X(cur) into X(cur ^ 1): d = c1 d + c2 D^-1 (b - K x)
  // (first: d = c2 D^-1 (b - K x))
  auto step = [&](bool first, T c1, T c2) {
    const T* x = X(cur);
    T* xo = X(cur ^ 1);
    element_products(x);
    for (int n = tid; n < g.nnode; n += nt) {
      T acc[3];
      node_sum(n, acc);
      const uint8_t f = F[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int i = 3 * n + c;
        const T yc = ((f >> c) & 1) ? x[i] : acc[c];
        const T dd = first ? c2 * DI[i] * (B[i] - yc) : c1 * D[i] + c2 * DI[i] * (B[i] - yc);
        D[i] = dd;
        xo[i] = x[i] + dd;
      }
    }
    __syncthreads();
    cur ^= 1;
  };

  // pre-smoothing from x = 0: the first step is d = c0 D^-1 b
  for (int i = tid; i < g.ndof; i += nt) {
    const T dd = c0 * DI[i] * B[i];
    D[i] = dd;
    X(cur)[i] = dd;
  }
  __syncthreads();
  for (int k = 1; k < cf.deg; ++k) step(false, (T)cf.c1[k], (T)cf.c2[k]);

  // residual (0 on fixed DOFs) into the free buffer, restricted to the coarsest grid
  {
    const T* x = X(cur);
    T* r = X(cur ^ 1);
    element_products(x);
    for (int n = tid; n < g.nnode; n += nt) {
      T acc[3];
      node_sum(n, acc);
      const uint8_t f = F[n];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        r[3 * n + c] = ((f >> c) & 1) ? T(0) : B[3 * n + c] - acc[c];
    }
    __syncthreads();
    for (int I = tid; I < gc.nnode; I += nt) {
      int Iy, Ix, Iz;
      grid_node_coords(gc, I, Iy, Ix, Iz);
      T o[3];
      restrict_at(g, Iy, Ix, Iz, FC[I], r, o);
      BC[3 * I] = o[0];
      BC[3 * I + 1] = o[1];
      BC[3 * I + 2] = o[2];
    }
    __syncthreads();
  }

  // coarsest solve with the stored inverse (symmetric: row i read as column i,
  // coalesced over the threads), accumulated in double; from the LDS copy when
  // it fits, else from global memory with many loads in flight
  if (inv_lds) {
    for (int i = tid; i < gc.ndof; i += nt) {
      double acc = 0.0;
#pragma unroll 8
      for (int j = 0; j < gc.ndof; ++j) acc += (double)IV[j * gc.ndof + i] * (double)BC[j];
      XC[i] = (T)acc;
    }
  } else {
    for (int i = tid; i < gc.ndof; i += nt) {
      double acc = 0.0;
#pragma unroll 16
      for (int j = 0; j < gc.ndof; ++j) acc += (double)inv[j * gc.ndof + i] * (double)BC[j];
      XC[i] = (T)acc;
    }
  }
  __syncthreads();

  // prolongation onto the free DOFs, then the post-smoothing
  for (int n = tid; n < g.nnode; n += nt) prolong_node(gc, g, F, XC, X(cur), n);
  __syncthreads();
  step(true, T(0), c0);
  for (int k = 1; k < cf.deg; ++k) step(false, (T)cf.c1[k], (T)cf.c2[k]);

  for (int i = tid; i < g.ndof; i += nt) xout[i] = X(cur)[i];
}

void tail_setup(Plan& P) {
  P.tail_level = -1;
  // the sector form needs KE's mirror symmetry (as the element kernel does)
  if (P.nlev < 3 || !P.sector_ok) return;
  const int lt = P.nlev - 2;
  const Grid& g = P.lv[lt].g;
  const Grid& gc = P.lv[lt + 1].g;
  const int deg = lt >= 2 ? P.opt.coarse_degree : P.opt.smooth_degree;
  if (g.nnode > kTailMaxNodes || deg > kTailMaxDeg || gc.ndof > kCoarseMaxDof) return;
  const size_t lds64 = tail_lds<double>(g, gc, false), lds32 = tail_lds<float>(g, gc, false);
  if ((int)lds64 > P.lds_limit) return;
  // the coarsest inverse joins the LDS state where it fits
  P.tail_inv_lds64 = (int)tail_lds<double>(g, gc, true) <= P.lds_limit;
  P.tail_inv_lds32 = (int)tail_lds<float>(g, gc, true) <= P.lds_limit;
  if (hipFuncSetAttribute((const void*)k_tail<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)tail_lds<double>(g, gc, P.tail_inv_lds64)) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)tail_lds<float>(g, gc, P.tail_inv_lds32)) != hipSuccess)
    return;   // the level-by-level cycle stays in use
  (void)lds32;
  P.tail_level = lt;
}

template <typename T>
void launch_tail(Plan& P, const T* b, T* x, const TailCoef& cf, hipStream_t s) {
  const int lt = P.tail_level;
  const Level& L = P.lv[lt];
  const Level& C = P.lv[lt + 1];
  const LevelBuf<T>* B;
  const T* sec;
  const T* inv;
  bool inv_lds;
  if constexpr (std::is_same_v<T, double>) {
    B = &L.f64;
    sec = P.sec64_dev;
    inv = P.coarse_inv64;
    inv_lds = P.tail_inv_lds64;
  } else {
    B = &L.f32;
    sec = P.sec32_dev;
    inv = P.coarse_inv32;
    inv_lds = P.tail_inv_lds32;
  }
  hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(kTailThreads), tail_lds<T>(L.g, C.g, inv_lds), s,
                     L.g, C.g, sec, B->s, L.fixed, C.fixed, B->dinv, b, inv, x, cf, (int)inv_lds,
                     &P.sc->done);
}

template void launch_tail<float>(Plan&, const float*, float*, const TailCoef&, hipStream_t);
template void launch_tail<double>(Plan&, const double*, double*, const TailCoef&, hipStream_t);

}  // namespace tk
