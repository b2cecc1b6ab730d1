// Single-workgroup tail of the V-cycle: on the second-coarsest level the
// pre-smoothing, the residual and its restriction; the dense coarsest solve;
// then the prolongation and the post-smoothing back on the second-coarsest
// level.  Those two levels hold a few hundred nodes (9x5x5 and 5x3x3 at
// 256x128x128): level by level they cost a dozen launches per V-cycle of ~5 us
// each, nearly all launch and memory latency.  Here one workgroup of 1024
// threads keeps the level state on chip and runs all of it in one launch,
// separated by barriers.
//
// The work is spread so that each phase is a short chain of instructions per
// lane (the level is far too small to be limited by throughput):
//   - element products: 8 lanes per element, one per corner.  The
//     Walsh-Hadamard transform of the sector form (tk_sector.hpp) runs as
//     three butterflies across the 8 lanes (DPP lane exchanges), the sector
//     blocks couple each lane's modes with two partner lanes' modes, and the
//     inverse butterflies leave corner j's outputs on lane j, which stores
//     them to LDS;
//   - node phase: one thread per DOF sums its <= 8 element outputs in a fixed
//     corner order (deterministic) and applies the Chebyshev step
//     d = c1 d + c2 D^-1 (b - K x), x' = x + d, with b, D^-1, d and x held in
//     its registers (x is also written to LDS for the next element phase);
//   - the trilinear restriction and prolongation of tk_transfer.hpp, and the
//     stored dense inverse of the coarsest operator, its product split into
//     partial sums over the threads.
// Inputs are masked at fixed DOFs and those rows are the identity, as in the
// level-by-level kernels.
#include "tk_plan.hpp"
#include "tk_sector.hpp"
#include "tk_transfer.hpp"

#include <type_traits>

namespace tk {

constexpr int kTailThreads = 1024;
constexpr int kTailElPass = 2;    // element passes: 128 elements of 8 lanes each per pass
constexpr int kTailDofPass = 2;   // DOFs owned per thread

// LDS layout of the tail for level g and coarsest grid gc (element type T),
// offsets in elements of T: x, the residual, the element outputs (24 rows of
// ne_pad, then one zero), coarse right-hand side and solution, the coarse
// solve's partial sums (kTailThreads doubles), optionally the coarsest inverse
// (16-byte aligned); then the fixed bits of both grids.
__host__ __device__ inline int tail_ne_pad(const Grid& g) {
  // a multiple of 64 plus 8: the 64 lanes of a wave (8 elements x 8 corners)
  // write 64 different banks
  return (g.nele + 63) / 64 * 64 + 8;
}
template <typename T>
__host__ __device__ inline size_t tail_inv_offset(const Grid& g, const Grid& gc) {
  const size_t vals = 2 * (size_t)g.ndof + 24 * (size_t)tail_ne_pad(g) + 1 + 2 * (size_t)gc.ndof +
                      kTailThreads * sizeof(double) / sizeof(T);
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
  if (done && *done) return;
  extern __shared__ __align__(16) unsigned char tail_mem[];
  const int NE = tail_ne_pad(g);
  T* X = reinterpret_cast<T*>(tail_mem);   // iterate
  T* R = X + g.ndof;                        // residual
  T* EY = R + g.ndof;                       // element outputs [corner j][component c][element]
  const int zslot = 24 * NE;                // EY[zslot] = 0: the output of a missing element
  T* BC = EY + zslot + 1;
  T* XC = BC + gc.ndof;
  double* PS = reinterpret_cast<double*>(XC + gc.ndof);
  T* IV = X + tail_inv_offset<T>(g, gc);    // the coarsest inverse, when inv_lds
  uint8_t* F = reinterpret_cast<uint8_t*>(IV + (inv_lds ? gc.ndof * gc.ndof : 0));
  uint8_t* FC = F + g.nnode;
  const int tid = threadIdx.x;
  const int n2 = gc.ndof * gc.ndof;
  const int nxy = g.nnx * g.nny, exy = g.nelx * g.nely;

  // ---- element lanes: element e = pass * 128 + tid / 8, corner j = tid % 8
  // (j = ox + 2 oy + 4 oz, the bit order of the corner transform)
  const int j = tid & 7;
  const int npass = (g.nele + 127) >> 7;
  T kc[3][3];   // this lane's sector rows: output component c, input component c'
  sec8_rows(sec, j, kc);
  int nd[kTailElPass];       // corner node of the lane's element (-1: no element)
  T se[kTailElPass];         // element scale
#pragma unroll
  for (int p = 0; p < kTailElPass; ++p) {
    nd[p] = -1;
    se[p] = T(0);
    const int e = p * 128 + (tid >> 3);
    if (p < npass && e < g.nele) {
      const int ey = e % g.nely, t = e / g.nely, ex = t % g.nelx, ez = t / g.nelx;
      nd[p] = (ey + ((j >> 1) & 1)) + (ex + (j & 1)) * g.nny + (ez + (j >> 2)) * nxy;
      se[p] = sl[e];
    }
  }
  // ---- owned DOFs: i = q * kTailThreads + tid; b, D^-1, d and x in registers
  T rb[kTailDofPass], rdi[kTailDofPass], rd[kTailDofPass], rx[kTailDofPass];
  int ebase[kTailDofPass];   // EY offset of corner 0's element for the DOF's component
  int emask[kTailDofPass];   // bit k: the node's corner-k element exists; bit 8: DOF fixed
#pragma unroll
  for (int q = 0; q < kTailDofPass; ++q) {
    const int i = q * kTailThreads + tid;
    rb[q] = rdi[q] = rd[q] = rx[q] = T(0);
    ebase[q] = 0;
    emask[q] = 0;
    if (i < g.ndof) {
      rb[q] = b[i];
      rdi[q] = dinv[i];
      const int n = i / 3, c = i - 3 * n;
      int iy, ix, iz;
      grid_node_coords(g, n, iy, ix, iz);
      ebase[q] = c * NE + iy + ix * g.nely + iz * exy;
      int m = ((fixed[n] >> c) & 1) << 8;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
        if ((unsigned)(iy - oy) < (unsigned)g.nely && (unsigned)(ix - ox) < (unsigned)g.nelx &&
            (unsigned)(iz - oz) < (unsigned)g.nelz)
          m |= 1 << k;
      }
      emask[q] = m;
    }
  }
  if (inv_lds) {
    // 16-byte copies, several in flight per thread
    const int n16 = n2 * (int)sizeof(T) / 16;
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(inv);
    uint4* dst = reinterpret_cast<uint4*>(IV);
#pragma unroll 4
    for (int i = tid; i < n16; i += kTailThreads) dst[i] = src[i];
    for (int i = n16 * (16 / (int)sizeof(T)) + tid; i < n2; i += kTailThreads) IV[i] = inv[i];
  }
  for (int i = tid; i < g.nnode; i += kTailThreads) F[i] = fixed[i];
  for (int i = tid; i < gc.nnode; i += kTailThreads) FC[i] = fixed_c[i];
  if (tid == 0) EY[zslot] = T(0);
  __syncthreads();
  // fixed bits of the element lanes' corner nodes
  int cfix[kTailElPass];
#pragma unroll
  for (int p = 0; p < kTailElPass; ++p) cfix[p] = nd[p] >= 0 ? F[nd[p]] : 7;

  // EY = the element products s_e KE u_e (inputs masked at fixed DOFs) in the
  // sector form, 8 lanes per element (sec8_apply)
  auto element_products = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int p = 0; p < kTailElPass; ++p) {
      if (p >= npass) break;   // uniform: every lane of the wave takes part
      T u[3], o[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = (nd[p] >= 0 && !((cfix[p] >> c) & 1)) ? X[3 * nd[p] + c] : T(0);
      sec8_apply(kc, se[p], j, u, o);
      const int e = p * 128 + (tid >> 3);
      if (nd[p] >= 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) EY[(3 * j + c) * NE + e] = o[c];
    }
    __syncthreads();
  };
  // (K u) at owned DOF q: the <= 8 element outputs, summed in corner order
  auto node_sum = [&](int q) __attribute__((always_inline)) -> T {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
      const int a = ebase[q] + 3 * k * NE - (oy + ox * g.nely + oz * exy);
      acc += EY[((emask[q] >> k) & 1) ? a : zslot];
    }
    return acc;
  };
  // one Chebyshev step d = c1 d + c2 D^-1 (b - K x), x += d (first: c1 = 0)
  auto step = [&](bool first, T c1, T c2) __attribute__((always_inline)) {
    element_products();
#pragma unroll
    for (int q = 0; q < kTailDofPass; ++q) {
      const int i = q * kTailThreads + tid;
      if (i < g.ndof) {
        const T ku = node_sum(q);
        const T yc = (emask[q] >> 8) ? rx[q] : ku;
        const T dd = first ? c2 * rdi[q] * (rb[q] - yc) : c1 * rd[q] + c2 * rdi[q] * (rb[q] - yc);
        rd[q] = dd;
        rx[q] += dd;
        X[i] = rx[q];
      }
    }
    __syncthreads();
  };

  // pre-smoothing from x = 0: the first step is d = c0 D^-1 b
  const T c0 = (T)cf.c0;
#pragma unroll
  for (int q = 0; q < kTailDofPass; ++q) {
    const int i = q * kTailThreads + tid;
    if (i < g.ndof) {
      rd[q] = c0 * rdi[q] * rb[q];
      rx[q] = rd[q];
      X[i] = rx[q];
    }
  }
  __syncthreads();
  for (int k = 1; k < cf.deg; ++k) step(false, (T)cf.c1[k], (T)cf.c2[k]);

  // residual (0 on fixed DOFs), restricted to the coarsest grid
  element_products();
#pragma unroll
  for (int q = 0; q < kTailDofPass; ++q) {
    const int i = q * kTailThreads + tid;
    if (i < g.ndof) R[i] = (emask[q] >> 8) ? T(0) : rb[q] - node_sum(q);
  }
  __syncthreads();
  if (tid < gc.ndof) {
    const int I = tid / 3, c = tid - 3 * I;
    int Iy, Ix, Iz;
    grid_node_coords(gc, I, Iy, Ix, Iz);
    T o[3];
    restrict_at(g, Iy, Ix, Iz, FC[I], R, o);
    BC[tid] = o[c];
  }
  __syncthreads();

  // coarsest solve with the stored inverse (symmetric: column i read as row i),
  // split over the threads: part p of output i sums j = p, p + np, ... in
  // double, then the np parts are added in order
  {
    const int n = gc.ndof, np = kTailThreads / n;
    if (tid < np * n) {
      const int p = tid / n, i = tid - p * n;
      double acc = 0.0;
      if (inv_lds) {
#pragma unroll 4
        for (int jj = p; jj < n; jj += np) acc += (double)IV[jj * n + i] * (double)BC[jj];
      } else {
#pragma unroll 8
        for (int jj = p; jj < n; jj += np) acc += (double)inv[jj * n + i] * (double)BC[jj];
      }
      PS[tid] = acc;
    }
    __syncthreads();
    if (tid < n) {
      double acc = 0.0;
      for (int p = 0; p < np; ++p) acc += PS[p * n + tid];
      XC[tid] = (T)acc;
    }
    __syncthreads();
  }

  // prolongation onto the free DOFs, then the post-smoothing
#pragma unroll
  for (int q = 0; q < kTailDofPass; ++q) {
    const int i = q * kTailThreads + tid;
    if (i < g.ndof && !(emask[q] >> 8)) {
      const int n = i / 3, c = i - 3 * n;
      int iy, ix, iz;
      grid_node_coords(g, n, iy, ix, iz);
      const int y0 = iy >> 1, y1 = (iy + 1) >> 1, x0 = ix >> 1, x1 = (ix + 1) >> 1;
      const int z0 = iz >> 1, z1 = (iz + 1) >> 1;
      const T wy[2] = {(iy & 1) ? T(0.5) : T(1), (iy & 1) ? T(0.5) : T(0)};
      const T wx[2] = {(ix & 1) ? T(0.5) : T(1), (ix & 1) ? T(0.5) : T(0)};
      const T wz[2] = {(iz & 1) ? T(0.5) : T(1), (iz & 1) ? T(0.5) : T(0)};
      const int cxy = gc.nnx * gc.nny;
      T a = T(0);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ky = k & 1, kx = (k >> 1) & 1, kz = k >> 2;
        const int I = (ky ? y1 : y0) + (kx ? x1 : x0) * gc.nny + (kz ? z1 : z0) * cxy;
        a += wy[ky] * wx[kx] * wz[kz] * XC[3 * I + c];
      }
      rx[q] += a;
      X[i] = rx[q];
    }
  }
  __syncthreads();
  step(true, T(0), c0);
  for (int k = 1; k < cf.deg; ++k) step(false, (T)cf.c1[k], (T)cf.c2[k]);

#pragma unroll
  for (int q = 0; q < kTailDofPass; ++q) {
    const int i = q * kTailThreads + tid;
    if (i < g.ndof) xout[i] = rx[q];
  }
}

void tail_setup(Plan& P) {
  P.tail_level = -1;
  // the sector form needs KE's mirror symmetry (as the element kernel does)
  if (P.nlev < 3 || !P.sector_ok) return;
  const int lt = P.nlev - 2;
  const Grid& g = P.lv[lt].g;
  const Grid& gc = P.lv[lt + 1].g;
  const int deg = lt >= 2 ? P.opt.coarse_degree : P.opt.smooth_degree;
  if (g.nnode > kTailMaxNodes || deg > kTailMaxDeg || gc.ndof > kCoarseMaxDof ||
      g.nele > 128 * kTailElPass || g.ndof > kTailThreads * kTailDofPass || gc.ndof > kTailThreads)
    return;
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

// done == nullptr: run whatever the solver's flag says (tk_tail_apply)
template <typename T>
void launch_tail(Plan& P, const T* b, T* x, const TailCoef& cf, hipStream_t s, const int* done) {
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
                     done);
}

template void launch_tail<float>(Plan&, const float*, float*, const TailCoef&, hipStream_t,
                                 const int*);
template void launch_tail<double>(Plan&, const double*, double*, const TailCoef&, hipStream_t,
                                  const int*);

}  // namespace tk
