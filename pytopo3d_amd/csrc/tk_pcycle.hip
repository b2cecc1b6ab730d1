// Persistent coarse cycle: the multigrid levels between the launch-per-step
// levels and the single-workgroup tail (32x16x16 and 16x8x8 elements at
// 256x128x128) in one launch per direction of the V-cycle, on the 32 compute
// units of one XCD.
//
// A launch per smoothing step costs ~5 us on these levels whatever their size
// (dispatch, a round of loads, the drain of the stores); with ~11 steps per
// level and direction that is most of their time.  Here the launch has 8 x 36
// workgroups (the dispatcher deals them to the 8 XCDs in turn); the first 32
// that the hardware's XCC_ID register places on XCD 0 take part, ranked by an
// atomic ticket in that XCD's L2 (the others return at once), each owning a
// brick of the level's nodes and keeping its b, D^-1, d and x in registers.
// The down and up launches draw tickets from, and raise flags in, words of
// their own; each resets the other's once its own participants are ranked
// (the other launch runs only after this one has finished).  One smoothing
// step:
//   - the node values of every element touching the brick (the brick and a
//     one-node halo) are loaded into LDS with agent-scope loads: they miss the
//     compute unit's L1 and read the XCD's L2, where the other workgroups'
//     stores of the previous step are;
//   - one thread per element forms its sector-form product (tk_sector.hpp)
//     and writes the 24 corner outputs to LDS; one thread per owned DOF sums
//     its <= 8 in a fixed order and applies d = c1 d + c2 D^-1 (b - K x),
//     x += d (fixed DOFs: identity rows, inputs masked);
//   - the new x is stored (the level's two iterate buffers alternate, so no
//     workgroup overwrites what another may still read), the stores are
//     waited for, and the 32 workgroups meet at a barrier: each stores the
//     number of barriers it has reached into a flag word of its own (a cache
//     line each) and one wave polls all 32 flags with agent-scope loads until
//     every one has reached it (~1 us with the stores, measured; an arrival
//     count that all 32 add to took ~2.3 us, its atomics serialising in L2).
// The restriction to the next level is gathered by that level's owners after
// a barrier, the prolongation by the finer level's owners.  Every wait is
// bounded (~50 ms of the constant-rate clock): on expiry the launch records an
// error, raises the solver's stop flag and returns, so it cannot hang; the
// solve then fails with a message and the plan falls back to a launch per step.
#include "tk_plan.hpp"
#include "tk_sector.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace tk {

constexpr int kPcWG = 32;            // participating workgroups (one XCD)
constexpr int kPcBlocks = 8 * 36;    // blocks per launch (8 XCDs): >= 32 land on XCD 0
static_assert(kPcSyncWords >= 64 + 2 * 32 * kPcWG, "ticket and flag words");
constexpr int kPcMaxBrickNodes = 512;
constexpr int kPcMaxEl = 1024;       // elements touching one brick
constexpr int kPcMaxBox = 1024;      // their nodes

template <typename T>
constexpr int pc_threads() {
  return sizeof(T) == 4 ? 1024 : 512;   // FP64: 256 registers per lane for the sector form
}

template <typename T>
struct PcLevel {
  Grid g;
  const uint8_t* fixed;
  const T* s;       // element scales
  const T* dinv;    // inverse diagonal (1 on fixed DOFs)
  T* b;             // right-hand side
  T* x0;            // iterates (alternating)
  T* x1;
  T* r;             // residual
  int nb[3];        // bricks along y, x, z (product kPcWG)
};

template <typename T>
struct PcArgs {
  int nlev;                      // persistent levels, finest first
  PcLevel<T> lv[kPcMaxLev];
  Grid gt;                       // the tail level below them
  const uint8_t* fixed_t;
  T* bt;                         // its right-hand side (written on the way down)
  const T* xt;                   // its solution (read on the way up)
  const T* sec;                  // the 45 sector entries
  TailCoef cf;
  const int* done;
  unsigned* sync;                // tickets: [16] down, [48] up launch; flags of
  Scalars* sc;                   // workgroup w: [64 + 32 w] down, [1088 + 32 w] up
  int box_max;                   // LDS layout: node-box values (3 box_max), element
  int epad_max;                  // outputs (24 epad_max + 1), the 45 sector entries
  unsigned long long spin;       // barrier wait limit, ticks of the 100 MHz realtime clock
};

// Device-memory accesses of data other workgroups write during the launch: the
// pointers arrive inside the argument struct, so they are marked global here
// (generic ones compile to flat instructions).
template <typename T>
using pc_gptr = __attribute__((address_space(1))) T*;

// s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4): (size - 1) << 11 | offset << 6 | id 20
__device__ __forceinline__ int pc_xcc() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xF; }
template <typename T>
__device__ __forceinline__ T pc_load(const T* p) {
  return __hip_atomic_load((pc_gptr<const T>)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void pc_store(T* p, T v) {
  *(pc_gptr<T>)p = v;
}

// Barrier number j (from 0) of a launch among its kPcWG workgroups, rank w;
// flags: the launch's flag words (32 apart).  Returns false once any wait has
// expired (then every later barrier returns at once).
__device__ bool pc_barrier(unsigned* flags, int w, unsigned j, Scalars* sc,
                           unsigned long long spin) {
  __builtin_amdgcn_s_waitcnt(0);   // this thread's stores are in L2
  __syncthreads();
  __shared__ int ok;
  if (threadIdx.x < 64) {   // one wave
    int good = pc_load(&sc->pc_err) == 0;
    if (good) {
      if (threadIdx.x == 0) pc_store(flags + 32 * w, j + 1);
      const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
      while (true) {
        const unsigned f =
            threadIdx.x < kPcWG ? pc_load(flags + 32 * threadIdx.x) : 0xffffffffu;
        if (__all(f >= j + 1)) break;
        if (__builtin_amdgcn_s_memrealtime() - t0 > spin) {
          if (threadIdx.x == 0) {
            atomicExch(&sc->pc_err, 1);
            atomicExch(&sc->done, 1);
          }
          good = 0;
          break;
        }
      }
    }
    if (threadIdx.x == 0) ok = good;
  }
  __syncthreads();
  return ok;
}

// Clears the other launch's flags and ticket (rank 0, after its first barrier).
__device__ __forceinline__ void pc_reset_other(unsigned* flags, unsigned* ticket) {
  if (threadIdx.x < kPcWG) pc_store(flags + 32 * threadIdx.x, 0u);
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
}

// Rank of this workgroup among the participants (tickets from *ticket), or
// -1: not on XCD 0, stopped solver, or a 33rd block on XCD 0.
__device__ int pc_rank(unsigned* ticket, const int* done, Scalars* sc) {
  if (pc_xcc() != 0) return -1;   // the same for every thread of the workgroup
  // one thread reads the flags (another workgroup of the launch may be
  // setting them) and decides for the whole workgroup
  __shared__ int rank;
  if (threadIdx.x == 0) rank = (*done || sc->pc_err) ? kPcWG : (int)atomicAdd(ticket, 1u);
  __syncthreads();
  return rank < kPcWG ? rank : -1;
}

// Element-output row length for nel elements: a multiple of 64 plus 8
__host__ __device__ inline int pc_epad(int nel) { return (nel + 63) / 64 * 64 + 8; }

// Geometry of workgroup w's brick on a level: owned nodes [n0, n1) per axis
// (y, x, z), the elements touching them (origin e0, extent ne) and their nodes
// (origin e0, extent ne + 1).
struct PcBrick {
  int n0[3], n1[3], e0[3], ne[3];
  int nbn[3];   // node-box extent
  int nown, nel, nbox;
};

__host__ __device__ inline void pc_brick(const Grid& g, const int nb[3], int w, PcBrick& B) {
  const int nn[3] = {g.nny, g.nnx, g.nnz}, nel[3] = {g.nely, g.nelx, g.nelz};
  const int wi[3] = {w % nb[0], (w / nb[0]) % nb[1], w / (nb[0] * nb[1])};
  B.nown = B.nel = B.nbox = 1;
  for (int a = 0; a < 3; ++a) {
    B.n0[a] = (int)((long long)nn[a] * wi[a] / nb[a]);
    B.n1[a] = (int)((long long)nn[a] * (wi[a] + 1) / nb[a]);
    B.e0[a] = B.n0[a] > 0 ? B.n0[a] - 1 : 0;
    const int e1 = B.n1[a] - 1 < nel[a] - 1 ? B.n1[a] - 1 : nel[a] - 1;
    B.ne[a] = e1 - B.e0[a] + 1;
    B.nbn[a] = B.ne[a] + 1;
    B.nown *= B.n1[a] - B.n0[a];
    B.nel *= B.ne[a];
    B.nbox *= B.nbn[a];
  }
}

// One level's per-thread state in the persistent kernel: the owned DOFs
// (registers), the thread's elements and its node-box load slots.
template <typename T, int NT>
struct PcState {
  static constexpr int DP = (3 * kPcMaxBrickNodes + NT - 1) / NT;   // owned DOFs
  static constexpr int EP = (kPcMaxEl + NT - 1) / NT;               // elements
  static constexpr int LS = (3 * kPcMaxBox + NT - 1) / NT;          // node-box values
  PcBrick B;
  int epad;          // element-output row length
  T b[DP], di[DP], d[DP], x[DP];
  int gi[DP];        // global DOF (-1: none)
  int eb[DP];        // element-output offset of the DOF's corner-0 element
  int em[DP];        // bits 0-7: corner-k element exists; bit 8: fixed DOF
  int ec[EP];        // node-box index of the element's corner 0 (-1: none)
  T es[EP];          // element scale
  int lo[LS];        // global DOF of node-box value s * NT + tid (-1: zero, -2: none)

  __device__ __forceinline__ void init(const PcLevel<T>& L, int w, int tid) {
    const Grid& g = L.g;
    pc_brick(g, L.nb, w, B);
    epad = pc_epad(B.nel);
    const int oy = B.n1[0] - B.n0[0], ox = B.n1[1] - B.n0[1];
    const int nxy = g.nnx * g.nny, exy = g.nelx * g.nely;
#pragma unroll
    for (int q = 0; q < DP; ++q) {
      const int k = q * NT + tid;
      gi[q] = -1;
      eb[q] = em[q] = 0;
      b[q] = di[q] = d[q] = x[q] = T(0);
      if (k < 3 * B.nown) {
        const int n = k / 3, c = k - 3 * n;
        const int iy = B.n0[0] + n % oy, ix = B.n0[1] + (n / oy) % ox, iz = B.n0[2] + n / (oy * ox);
        const int gn = iy + ix * g.nny + iz * nxy;
        gi[q] = 3 * gn + c;
        di[q] = L.dinv[gi[q]];
        eb[q] = c * epad + (iy - B.e0[0]) + (ix - B.e0[1]) * B.ne[0] + (iz - B.e0[2]) * B.ne[0] * B.ne[1];
        int m = ((L.fixed[gn] >> c) & 1) << 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int cx = j & 1, cy = (j >> 1) & 1, cz = j >> 2;
          if ((unsigned)(iy - cy) < (unsigned)g.nely && (unsigned)(ix - cx) < (unsigned)g.nelx &&
              (unsigned)(iz - cz) < (unsigned)g.nelz)
            m |= 1 << j;
        }
        em[q] = m;
      }
    }
#pragma unroll
    for (int p = 0; p < EP; ++p) {
      const int e = p * NT + tid;
      ec[p] = -1;
      es[p] = T(0);
      if (e < B.nel) {
        const int ly = e % B.ne[0], lx = (e / B.ne[0]) % B.ne[1], lz = e / (B.ne[0] * B.ne[1]);
        ec[p] = ly + lx * B.nbn[0] + lz * B.nbn[0] * B.nbn[1];
        es[p] = L.s[(B.e0[0] + ly) + (B.e0[1] + lx) * g.nely + (B.e0[2] + lz) * exy];
      }
    }
#pragma unroll
    for (int s = 0; s < LS; ++s) {
      const int v = s * NT + tid;
      lo[s] = -2;
      if (v < 3 * B.nbox) {
        const int n = v / 3, c = v - 3 * n;
        const int gy = B.e0[0] + n % B.nbn[0], gx = B.e0[1] + (n / B.nbn[0]) % B.nbn[1];
        const int gz = B.e0[2] + n / (B.nbn[0] * B.nbn[1]);
        const int gn = gy + gx * g.nny + gz * nxy;
        lo[s] = ((L.fixed[gn] >> c) & 1) ? -1 : 3 * gn + c;
      }
    }
  }

  // EY = the element products of x (global) for the brick's elements
  __device__ __forceinline__ void products(const T* xg, T* XN, T* EY, const T* kb, int tid) const {
    // all loads issued before any is used (unconditional, clamped index)
    T v[LS];
#pragma unroll
    for (int s = 0; s < LS; ++s) v[s] = pc_load(xg + (lo[s] > 0 ? lo[s] : 0));
#pragma unroll
    for (int s = 0; s < LS; ++s)
      if (lo[s] > -2) XN[s * NT + tid] = lo[s] >= 0 ? v[s] : T(0);
    if (tid == 0) EY[24 * epad] = T(0);
    __syncthreads();
    const int sy = B.nbn[0], sxy = B.nbn[0] * B.nbn[1];
#pragma unroll
    for (int p = 0; p < EP; ++p) {
      if (ec[p] < 0) continue;
      T Fb[3][4], Ft[3][4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const int m0 = ec[p] + (f >> 1) + (f & 1) * sy, m1 = m0 + sxy;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Fb[c][f] = XN[3 * m0 + c];
          Ft[c][f] = XN[3 * m1 + c];
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
      sector_apply<T>(kb, es[p], a, o);
      const int e = p * NT + tid;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T nb[4], nt[4];
        nb[0] = -o[c][4];
        nt[0] = o[c][4];
#pragma unroll
        for (int f = 1; f < 4; ++f) {
          nb[f] = o[c][f] - o[c][f + 4];
          nt[f] = o[c][f] + o[c][f + 4];
        }
        face_inv(nb);
        face_inv(nt);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          EY[(3 * f + c) * epad + e] = nb[f];
          EY[(3 * (f + 4) + c) * epad + e] = nt[f];
        }
      }
      __builtin_amdgcn_sched_barrier(0);   // one element's temporaries at a time
    }
    __syncthreads();
  }

  // (K x) at owned DOF q from the element outputs, corner order
  __device__ __forceinline__ T node_sum(int q, const T* EY) const {
    const int zs = 24 * epad;
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int cx = j & 1, cy = (j >> 1) & 1, cz = j >> 2;
      const int a = eb[q] + 3 * j * epad - (cy + cx * B.ne[0] + cz * B.ne[0] * B.ne[1]);
      acc += EY[((em[q] >> j) & 1) ? a : zs];
    }
    return acc;
  }
  __device__ __forceinline__ bool fixed_dof(int q) const { return em[q] >> 8; }
};

// Component c of R rf (trilinear restriction, R = P^T) at coarse DOF i, the
// fine field read with agent-scope loads; zero on the coarse node's fixed DOFs.
template <typename T>
__device__ __forceinline__ T pc_restrict(const Grid& f, const Grid& cg, const uint8_t* fixed_c,
                                         const T* rf, int i) {
  const int I = i / 3, c = i - 3 * I;
  const int Iy = I % cg.nny, Ix = (I / cg.nny) % cg.nnx, Iz = I / (cg.nny * cg.nnx);
  if ((fixed_c[I] >> c) & 1) return T(0);
  T acc = T(0);
  const int nxy = f.nnx * f.nny;
#pragma unroll
  for (int kz = 0; kz < 3; ++kz) {
    const int jz = 2 * Iz + kz - 1;
    if ((unsigned)jz >= (unsigned)f.nnz) continue;
    const T wz = kz == 1 ? T(1) : T(0.5);
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int jx = 2 * Ix + kx - 1;
      if ((unsigned)jx >= (unsigned)f.nnx) continue;
      const T wzx = wz * (kx == 1 ? T(1) : T(0.5));
#pragma unroll
      for (int ky = 0; ky < 3; ++ky) {
        const int jy = 2 * Iy + ky - 1;
        if ((unsigned)jy >= (unsigned)f.nny) continue;
        acc += wzx * (ky == 1 ? T(1) : T(0.5)) * pc_load(rf + 3 * (jy + jx * f.nny + jz * nxy) + c);
      }
    }
  }
  return acc;
}

// (P xc) at fine DOF i (component c of fine node (iy, ix, iz)), agent-scope loads.
template <typename T>
__device__ __forceinline__ T pc_prolong(const Grid& cg, const T* xc, int iy, int ix, int iz, int c) {
  const int y0 = iy >> 1, y1 = (iy + 1) >> 1, x0 = ix >> 1, x1 = (ix + 1) >> 1;
  const int z0 = iz >> 1, z1 = (iz + 1) >> 1;
  const T wy[2] = {(iy & 1) ? T(0.5) : T(1), (iy & 1) ? T(0.5) : T(0)};
  const T wx[2] = {(ix & 1) ? T(0.5) : T(1), (ix & 1) ? T(0.5) : T(0)};
  const T wz[2] = {(iz & 1) ? T(0.5) : T(1), (iz & 1) ? T(0.5) : T(0)};
  const int cxy = cg.nnx * cg.nny;
  T a = T(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ky = k & 1, kx = (k >> 1) & 1, kz = k >> 2;
    const int I = (ky ? y1 : y0) + (kx ? x1 : x0) * cg.nny + (kz ? z1 : z0) * cxy;
    a += wy[ky] * wx[kx] * wz[kz] * pc_load(xc + 3 * I + c);
  }
  return a;
}

// node-box values, then the element outputs (24 rows of epad_max, one zero),
// then the 45 sector entries (48 slots)
template <typename T>
__host__ __device__ inline size_t pc_lds_bytes(int box_max, int epad_max) {
  return (3 * (size_t)box_max + 24 * (size_t)epad_max + 1 + 48) * sizeof(T);
}

// One Chebyshev step on a level: d = c1 d + c2 D^-1 (b - K x), x += d, x read
// from xin, written to xout, then the barrier.
template <typename T, int NT>
__device__ __forceinline__ bool pc_step(PcState<T, NT>& S, const T* xin, T* xout, T c1, T c2,
                                        T* XN, T* EY, const T* kb, unsigned* flags, int w,
                                        unsigned& nbar, Scalars* sc, unsigned long long spin,
                                        int tid) {
  S.products(xin, XN, EY, kb, tid);
#pragma unroll
  for (int q = 0; q < PcState<T, NT>::DP; ++q) {
    if (S.gi[q] < 0) continue;
    const T ku = S.node_sum(q, EY);
    const T yc = S.fixed_dof(q) ? S.x[q] : ku;
    const T dd = c1 * S.d[q] + c2 * S.di[q] * (S.b[q] - yc);
    S.d[q] = dd;
    S.x[q] += dd;
    pc_store(xout + S.gi[q], S.x[q]);
  }
  return pc_barrier(flags, w, nbar++, sc, spin);
}

template <typename T>
__global__ __launch_bounds__(pc_threads<T>()) void k_pcycle_down(PcArgs<T> A) {
  constexpr int NT = pc_threads<T>();
  const int w = pc_rank(A.sync + 16, A.done, A.sc), tid = threadIdx.x;
  if (w < 0) return;
  unsigned* flags = A.sync + 64;
  unsigned nbar = 0;
  extern __shared__ __align__(16) unsigned char pc_mem[];
  T* XN = reinterpret_cast<T*>(pc_mem);
  T* EY = XN + 3 * A.box_max;
  T* KB = EY + 24 * A.epad_max + 1;
  if (tid < kSecN) KB[tid] = A.sec[tid];
  const TailCoef& cf = A.cf;
  const T c0 = (T)cf.c0;
  PcState<T, NT> S;
  bool ok = true;
  for (int l = 0; l < A.nlev && ok; ++l) {
    const PcLevel<T>& L = A.lv[l];
    S.init(L, w, tid);
    // right-hand side: the finer level's restricted residual (the first level's
    // was written by the launch before), kept for the way up
#pragma unroll
    for (int q = 0; q < PcState<T, NT>::DP; ++q) {
      if (S.gi[q] < 0) continue;
      if (l == 0) {
        S.b[q] = L.b[S.gi[q]];
      } else {
        S.b[q] = pc_restrict(A.lv[l - 1].g, L.g, L.fixed, A.lv[l - 1].r, S.gi[q]);
        pc_store(L.b + S.gi[q], S.b[q]);
      }
      S.d[q] = c0 * S.di[q] * S.b[q];   // first step from x = 0
      S.x[q] = S.d[q];
      pc_store(L.x0 + S.gi[q], S.x[q]);
    }
    ok = pc_barrier(flags, w, nbar++, A.sc, A.spin);
    if (l == 0 && w == 0) pc_reset_other(A.sync + 64 + 32 * kPcWG, A.sync + 48);
    for (int k = 1; k < cf.deg && ok; ++k)
      ok = pc_step(S, (k & 1) ? L.x0 : L.x1, (k & 1) ? L.x1 : L.x0, (T)cf.c1[k], (T)cf.c2[k], XN,
                   EY, KB, flags, w, nbar, A.sc, A.spin, tid);
    if (!ok) break;
    // residual of the pre-smoothed x (in x[(deg - 1) & 1]), 0 on fixed DOFs
    S.products(((cf.deg - 1) & 1) ? L.x1 : L.x0, XN, EY, KB, tid);
#pragma unroll
    for (int q = 0; q < PcState<T, NT>::DP; ++q)
      if (S.gi[q] >= 0) pc_store(L.r + S.gi[q], S.fixed_dof(q) ? T(0) : S.b[q] - S.node_sum(q, EY));
    ok = pc_barrier(flags, w, nbar++, A.sc, A.spin);
  }
  if (!ok) return;
  // the tail level's right-hand side, its DOFs spread over all threads
  const PcLevel<T>& L = A.lv[A.nlev - 1];
  for (int i = w * NT + tid; i < A.gt.ndof; i += kPcWG * NT)
    pc_store(A.bt + i, pc_restrict(L.g, A.gt, A.fixed_t, L.r, i));
}

template <typename T>
__global__ __launch_bounds__(pc_threads<T>()) void k_pcycle_up(PcArgs<T> A) {
  constexpr int NT = pc_threads<T>();
  const int w = pc_rank(A.sync + 48, A.done, A.sc), tid = threadIdx.x;
  if (w < 0) return;
  unsigned* flags = A.sync + 64 + 32 * kPcWG;
  unsigned nbar = 0;
  extern __shared__ __align__(16) unsigned char pc_mem[];
  T* XN = reinterpret_cast<T*>(pc_mem);
  T* EY = XN + 3 * A.box_max;
  T* KB = EY + 24 * A.epad_max + 1;
  if (tid < kSecN) KB[tid] = A.sec[tid];
  const TailCoef& cf = A.cf;
  const T c0 = (T)cf.c0;
  // the pre-smoothed x of every level is in x[pd]; the prolongated one goes to
  // x[pd ^ 1] and the deg steps end in x0
  const int pd = (cf.deg - 1) & 1;
  PcState<T, NT> S;
  bool ok = true;
  for (int l = A.nlev - 1; l >= 0 && ok; --l) {
    const PcLevel<T>& L = A.lv[l];
    T* xb[2] = {L.x0, L.x1};
    S.init(L, w, tid);
    const Grid& cg = l + 1 < A.nlev ? A.lv[l + 1].g : A.gt;
    const T* xc = l + 1 < A.nlev ? A.lv[l + 1].x0 : A.xt;
    const int nxy = L.g.nnx * L.g.nny;
#pragma unroll
    for (int q = 0; q < PcState<T, NT>::DP; ++q) {
      if (S.gi[q] < 0) continue;
      S.b[q] = L.b[S.gi[q]];
      S.x[q] = pc_load(xb[pd] + S.gi[q]);
      if (!S.fixed_dof(q)) {
        const int n = S.gi[q] / 3, c = S.gi[q] - 3 * n;
        S.x[q] += pc_prolong(cg, xc, n % L.g.nny, (n / L.g.nny) % L.g.nnx, n / nxy, c);
      }
      pc_store(xb[pd ^ 1] + S.gi[q], S.x[q]);
    }
    ok = pc_barrier(flags, w, nbar++, A.sc, A.spin);
    if (l == A.nlev - 1 && w == 0) pc_reset_other(A.sync + 64, A.sync + 16);
    int cur = pd ^ 1;
    for (int k = 0; k < cf.deg && ok; ++k) {
      const T c1 = k ? (T)cf.c1[k] : T(0), c2 = k ? (T)cf.c2[k] : c0;
      ok = pc_step(S, xb[cur], xb[cur ^ 1], c1, c2, XN, EY, KB, flags, w, nbar, A.sc, A.spin,
                   tid);
      cur ^= 1;
    }
  }
}

// ---------------------------------------------------------------- host

__global__ void k_xcc_ids(int* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = pc_xcc();
}

// Brick counts (y, x, z) with product kPcWG for grid g: the fewest elements
// per brick; false when no split keeps a brick within the limits.  Raises
// box_max and epad_max to the split's largest node box and element rows.
static bool pc_split(const Grid& g, int nb[3], int& box_max, int& epad_max) {
  int best = -1, bbox = 0;
  for (int by = 1; by <= kPcWG; by *= 2)
    for (int bx = 1; by * bx <= kPcWG; bx *= 2) {
      const int bz = kPcWG / (by * bx);
      if (by > g.nny || bx > g.nnx || bz > g.nnz) continue;
      const int t[3] = {by, bx, bz};
      int mel = 0, mown = 0, mbox = 0;
      for (int w = 0; w < kPcWG; ++w) {
        PcBrick B;
        pc_brick(g, t, w, B);
        mel = std::max(mel, B.nel);
        mown = std::max(mown, B.nown);
        mbox = std::max(mbox, B.nbox);
      }
      if (mel > kPcMaxEl || mown > kPcMaxBrickNodes || mbox > kPcMaxBox) continue;
      if (best < 0 || mel < best) {
        best = mel;
        nb[0] = by;
        nb[1] = bx;
        nb[2] = bz;
        bbox = mbox;
      }
    }
  if (best < 0) return false;
  box_max = std::max(box_max, bbox);
  epad_max = std::max(epad_max, pc_epad(best));
  return true;
}

void pcycle_setup(Plan& P) {
  P.pc_first = -1;
  P.pc_nlev = 0;
  const int lt = P.tail_level;
  if (!P.opt.persistent_coarse || lt < 3 || !P.sector_ok || !P.use_el) return;
  // levels above the tail, coarse ones of at most kMidMinEl elements (the same
  // smoothing coefficients as the tail), not the Galerkin level, and small
  // enough for 32 bricks
  int first = lt, box_max = 0, epad_max = 0;
  for (int l = lt - 1; l >= 2 && lt - l <= kPcMaxLev; --l) {
    int bm = box_max, em = epad_max;
    if (l == P.gal_level || P.lv[l].g.nele > kMidMinEl ||
        !pc_split(P.lv[l].g, P.pc_nb[l], bm, em) ||
        pc_lds_bytes<double>(bm, em) > (size_t)P.lds_limit)
      break;
    box_max = bm;
    epad_max = em;
    first = l;
  }
  if (first == lt) return;
  // a launch of kPcBlocks must place at least kPcWG blocks on XCD 0 (the
  // XCC_ID register reads as the dispatcher deals the blocks)
  int* ids = nullptr;
  std::vector<int> h(kPcBlocks, -1);
  bool rr = hipMalloc(&ids, sizeof(int) * h.size()) == hipSuccess;
  if (rr) {
    hipLaunchKernelGGL(k_xcc_ids, dim3((unsigned)h.size()), dim3(64), 0, 0, ids);
    rr = hipMemcpy(h.data(), ids, sizeof(int) * h.size(), hipMemcpyDeviceToHost) == hipSuccess;
  }
  if (ids) (void)hipFree(ids);
  rr = rr && std::count(h.begin(), h.end(), 0) >= kPcWG;
  if (!rr ||
      hipFuncSetAttribute((const void*)k_pcycle_down<float>,
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)pc_lds_bytes<float>(box_max, epad_max)) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_pcycle_up<float>,
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)pc_lds_bytes<float>(box_max, epad_max)) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_pcycle_down<double>,
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)pc_lds_bytes<double>(box_max, epad_max)) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_pcycle_up<double>,
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)pc_lds_bytes<double>(box_max, epad_max)) != hipSuccess)
    return;
  P.pc_first = first;
  P.pc_nlev = lt - first;
  P.pc_box_max = box_max;
  P.pc_epad_max = epad_max;
}

template <typename T>
static PcArgs<T> pc_args(Plan& P, const TailCoef& cf) {
  PcArgs<T> A{};
  A.nlev = P.pc_nlev;
  for (int i = 0; i < P.pc_nlev; ++i) {
    Level& L = P.lv[P.pc_first + i];
    LevelBuf<T>& B = std::is_same_v<T, double> ? (LevelBuf<T>&)L.f64 : (LevelBuf<T>&)L.f32;
    PcLevel<T>& D = A.lv[i];
    D.g = L.g;
    D.fixed = L.fixed;
    D.s = B.s;
    D.dinv = B.dinv;
    D.b = B.b;
    D.x0 = B.x;
    D.x1 = B.xa;
    D.r = B.r;
    for (int a = 0; a < 3; ++a) D.nb[a] = P.pc_nb[P.pc_first + i][a];
  }
  Level& Lt = P.lv[P.tail_level];
  LevelBuf<T>& Bt = std::is_same_v<T, double> ? (LevelBuf<T>&)Lt.f64 : (LevelBuf<T>&)Lt.f32;
  A.gt = Lt.g;
  A.fixed_t = Lt.fixed;
  A.bt = Bt.b;
  A.xt = Bt.x;
  if constexpr (std::is_same_v<T, double>)
    A.sec = P.sec64_dev;
  else
    A.sec = P.sec32_dev;
  A.cf = cf;
  A.done = &P.sc->done;
  A.sync = P.pc_sync;
  A.sc = P.sc;
  A.box_max = P.pc_box_max;
  A.epad_max = P.pc_epad_max;
  A.spin = P.pc_spin;
  return A;
}

// The way down from level P.pc_first (right-hand side in its b): smoothing,
// residuals and restrictions, ending with the tail level's right-hand side in
// its b.  The way up from the tail level's x: the first level's result is in
// its x (LevelBuf::x).
template <typename T>
void launch_pcycle_down(Plan& P, const TailCoef& cf, hipStream_t s) {
  hipLaunchKernelGGL(k_pcycle_down<T>, dim3(kPcBlocks), dim3(pc_threads<T>()),
                     pc_lds_bytes<T>(P.pc_box_max, P.pc_epad_max), s, pc_args<T>(P, cf));
}
template <typename T>
void launch_pcycle_up(Plan& P, const TailCoef& cf, hipStream_t s) {
  hipLaunchKernelGGL(k_pcycle_up<T>, dim3(kPcBlocks), dim3(pc_threads<T>()),
                     pc_lds_bytes<T>(P.pc_box_max, P.pc_epad_max), s, pc_args<T>(P, cf));
}

template void launch_pcycle_down<float>(Plan&, const TailCoef&, hipStream_t);
template void launch_pcycle_down<double>(Plan&, const TailCoef&, hipStream_t);
template void launch_pcycle_up<float>(Plan&, const TailCoef&, hipStream_t);
template void launch_pcycle_up<double>(Plan&, const TailCoef&, hipStream_t);

}  // namespace tk
