// Shared definitions of the MI355X (gfx950) SIMP hot-path library.
//
// Layout conventions (identical to PyTopo3D's, pytopo3d/utils/filter.py:169-174
// and pytopo3d/utils/assembly.py:121-129):
//   element e = iy + ix*nely + iz*nelx*nely        (Fortran order of (nely,nelx,nelz))
//   node    n = iy + ix*(nely+1) + iz*(nelx+1)*(nely+1)
//   DOF       = 3*n + c, c in {x, y, z}            (array-of-structs, 24 B per node)
// H8 local corner a has (x, y, z) offset kCornerX/Y/Z[a]
// (pytopo3d/utils/assembly.py:17-26).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace tk {

constexpr int kBlock = 256;           // default workgroup size (4 waves of 64)
constexpr int kMaxLevels = 12;        // multigrid levels
constexpr int kCoarseMaxDof = 192;    // coarsest level is factorised densely in LDS
constexpr int kMaxStencil = 1024;     // density-filter stencil entries (rmin <= ~6)

// element flags
constexpr uint8_t kSolid = 1;         // protected solid: rho = 1
constexpr uint8_t kVoid = 2;          // protected void (obstacle): rho = 0, wins over solid

__host__ __device__ constexpr int corner_x(int a) { return (0xCC >> a) & 1; }  // 0,0,1,1,0,0,1,1
__host__ __device__ constexpr int corner_y(int a) { return (0x66 >> a) & 1; }  // 0,1,1,0,0,1,1,0
__host__ __device__ constexpr int corner_z(int a) { return (0xF0 >> a) & 1; }  // 0,0,0,0,1,1,1,1

// Geometry of one grid (fine level or a multigrid level).
struct Grid {
  int nelx, nely, nelz;   // elements
  int nnx, nny, nnz;      // nodes = elements + 1
  int nele, nnode, ndof;
};

inline Grid make_grid(int nelx, int nely, int nelz) {
  Grid g;
  g.nelx = nelx; g.nely = nely; g.nelz = nelz;
  g.nnx = nelx + 1; g.nny = nely + 1; g.nnz = nelz + 1;
  g.nele = nelx * nely * nelz;
  g.nnode = g.nnx * g.nny * g.nnz;
  g.ndof = 3 * g.nnode;
  return g;
}

// Device-resident scalars of one plan.  Kernels read and write these, so the
// host only synchronises when a caller asks for a scalar result.
struct Scalars {
  // conjugate gradients
  double rz, pq, alpha, beta, rr, bb, tol2, rr0;
  int iter, done, maxit, stalled;
  int pc_err;                // a barrier wait of the persistent coarse cycle expired
  // optimality-criteria bisection
  double lo, hi, lam, best_lam, best_err, target, V, ratio_scale, tol;
  double ka, kva, kb, kvb;   // evaluated multipliers nearest the crossing: V(ka) > target >= V(kb)
  double lam_prev;           // the previous bisection's kept multiplier (first probe)
  double slope_prev;         // its -dV/dlog(lambda) near the crossing (0: unknown)
  double gfac;               // step factor of the one-sided search
  double slot_lam[4], slot_v[4];   // last two midpoints on each side of the target
  int slot_state[4], slot_step[4];
  int steps, oc_done, max_steps, evals;
  int spec, nspec, redo, phase;  // probes on, probes made, repeat without probes, checks
  int check, pad1;               // slot being measured after the last step
  // general results
  double c;            // compliance
  double change;       // max |x_new - x| over free elements
  double vol_sum;      // sum of rho over free elements
  double gray_cnt;     // count of 0.05 < rho < 0.95 over free elements
  double sum0, sum1;   // scratch sums
  double norm0, norm1; // scratch norms
  double zs;           // scale of the FP32 preconditioner input (flexible CG)
  double qz;           // q.z of the last flexible-CG update (next beta)
};

// ---------------------------------------------------------------------------
// Deterministic reductions.  Partial results are combined in a fixed order
// (wave shuffle tree, then waves, then blocks strided by thread), so equal
// inputs and launch shapes give bitwise equal results on every run.
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += __shfl_down(v[i], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[i][wid] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = sh[i][0];
      for (int w = 1; w < nw; ++w) s += sh[i][w];
      v[i] = s;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double block_max(double v) {
  __shared__ double sh[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    double s = sh[0];
    for (int w = 1; w < nw; ++w) s = fmax(s, sh[w]);
    v = s;
  }
  __syncthreads();
  return v;
}

// Linear block id and block count of a 1-, 2- or 3-D grid.
__device__ __forceinline__ int block_id() {
  return (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}
__device__ __forceinline__ int block_count() { return (int)(gridDim.x * gridDim.y * gridDim.z); }

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs of the chip, each with its own L2.  Remapping the linear block id so
// that every XCD receives a contiguous range of logical tiles keeps the
// overlapping halos of neighbouring tiles in the same L2 instead of fetching
// them again from HBM.  The map is a bijection of [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_logical(int bid, int nb) {
  const int x = bid % kXcds, k = bid / kXcds;
  const int q = nb / kXcds, r = nb % kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Two-phase grid reductions: every block writes its partial value(s) to
// partials[i * G + block] with no atomics, and a one-block finalize kernel
// (tk_reduce.hip) combines them in a fixed order and applies the scalar
// update.  (A single device-scope counter touched by every block, to let the
// last block finish the sum, serialises at the memory side at tens of
// nanoseconds per block.)  All threads of the block must call (no early
// return).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
    const int G = block_count(), b = block_id();
#pragma unroll
    for (int i = 0; i < NV; ++i) partials[i * G + b] = v[i];
  }
}

__device__ __forceinline__ void block_partial_max(double v, double* partials, int slot) {
  v = block_max(v);
  if (threadIdx.x == 0) partials[slot * block_count() + block_id()] = v;
}

inline int nblocks(long long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

}  // namespace tk
