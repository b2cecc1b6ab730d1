// Allocation-free variant of tk_energy: element energies, dc/drho and the
// compliance of one displacement field without a plan.  The only device memory
// beyond the caller's inputs and outputs is `scratch` below, a file-scope
// array every call on a device shares: the finalize scalars at its start, the
// per-block partials of c after them.  Calls allocate nothing; calls that
// overlap on one device race on that array.
#include "pytopo3d_hip.h"
#include "tk_plan.hpp"
#include "tk_sector.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

namespace tk {

bool sector_form(const double* KE, double out[kSecN]);
void launch_energy_sector_raw(const Grid& g, const double* sec, const double* U,
                              const double* rho, double penal, double* ce, double* dc_drho,
                              double* partials, Scalars* sc, hipStream_t s);
void finalize_raw(int op, int G, const double* partials, Scalars* sc, hipStream_t s);

// 64 Ki partials of 256 elements each after 1 KiB of scalars: 513 KiB per device
constexpr int kScalarDoubles = 128;
constexpr int kNoAllocPartials = 1 << 16;
constexpr long long kNoAllocMaxElements = (long long)kNoAllocPartials * kBlock;
static_assert(sizeof(Scalars) <= kScalarDoubles * sizeof(double), "scalars exceed their slot");

static __device__ __align__(256) double scratch[kScalarDoubles + kNoAllocPartials];

// Device address of `scratch` on each device, looked up once per device with
// that device current.  In host code the array's name is only a stand-in for
// the symbol: kernels and copies get these addresses.
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(std::string("hipGetDevice failed: ") + hipGetErrorString(e));
    return 2;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index out of range for the allocation-free entry points");
    return 2;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return 2;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Sector entries of the last KE seen on this thread (the form is built once
// per KE, not per call).
struct SectorCache {
  double ke[576];
  double sec[kSecN];
  bool valid = false, ok = false;
};
static thread_local SectorCache g_sec;

__global__ void k_store_c(const Scalars* __restrict__ sc, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = sc->c;
}

}  // namespace tk

extern "C" int tk_energy_noalloc(int nelx, int nely, int nelz, const double* KE,
                                 const double* U, const double* rho, double penal, double* ce,
                                 double* dc_drho, double* c_dev, double* c_host, void* stream) {
  using namespace tk;
  if (nelx < 1 || nely < 1 || nelz < 1) {
    set_error("nelx, nely, and nelz must all be positive");
    return 1;
  }
  const long long nele = (long long)nelx * nely * nelz;
  if (nele > kNoAllocMaxElements) {
    set_error("tk_energy_noalloc handles at most " + std::to_string(kNoAllocMaxElements) +
              " elements (its static scratch holds " + std::to_string(kNoAllocPartials) +
              " partial sums); got " + std::to_string(nele));
    return 1;
  }
  const long long nn = (long long)(nelx + 1) * (nely + 1) * (nelz + 1);
  if (3 * nn >= (1LL << 31) - 1024) {
    set_error("grid too large: the DOF count must fit in 31 bits");
    return 1;
  }
  if (!KE || !U || !rho) {
    set_error("KE (host), U and rho are required");
    return 1;
  }
  if (!(std::isfinite(penal) && penal > 0)) {
    set_error("penal must be positive, got " + std::to_string(penal));
    return 1;
  }
  if (!g_sec.valid || std::memcmp(g_sec.ke, KE, sizeof(g_sec.ke)) != 0) {
    std::memcpy(g_sec.ke, KE, sizeof(g_sec.ke));
    g_sec.ok = sector_form(KE, g_sec.sec);
    g_sec.valid = true;
  }
  if (!g_sec.ok) {
    set_error("tk_energy_noalloc needs a KE with the element's mirror symmetry (every "
              "axis-aligned orthotropic lk_H8 KE has it); use tk_energy on a plan otherwise");
    return 1;
  }
  double* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  Scalars* sc = reinterpret_cast<Scalars*>(base);
  double* partials = base + kScalarDoubles;
  hipStream_t s = (hipStream_t)stream;
  const Grid g = make_grid(nelx, nely, nelz);
  // every partial the finalize reads is written by this launch (one per block)
  launch_energy_sector_raw(g, g_sec.sec, U, rho, penal, ce, dc_drho, partials, sc, s);
  finalize_raw(kFinC, nblocks(g.nele), partials, sc, s);
  if (c_dev) hipLaunchKernelGGL(k_store_c, dim3(1), dim3(64), 0, s, sc, c_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return 2;
  }
  if (c_host) {
    e = hipMemcpyAsync(c_host, &sc->c, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("reading c failed: ") + hipGetErrorString(e));
      return 2;
    }
  }
  return 0;
}
