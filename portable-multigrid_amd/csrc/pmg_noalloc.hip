// pmg_noalloc.hip -- preconditioned CG on a fixed, file-scope device work
// array instead of vectors the library allocates (pmg_cg_solve_noalloc).
//
// `scratch` is one static device array per GPU the library runs on, held
// until the process ends.  Host code never passes the array's name to a
// kernel or a copy: the name is a host-side stand-in, not a device address.
// Its address on the current device is looked up once per device with
// hipGetSymbolAddress and kept; every use goes through that pointer.
//
// Layout of one call's work space inside scratch (offsets in bytes):
//   [0, 64)                 8 CG scalars (cleared by a fill kernel per call)
//   [256, 256 + 8 kDotBlocks)   partial sums of the dot products
//   kHeader + i * vb        r, z, p, q (i = 0..3), vb = n s rounded up to 256
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two host threads, or two streams without ordering) race.
#include <mutex>

#include "pmg_library.hpp"

namespace pmg {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kAlign = 256;
constexpr size_t kHeader = ((kAlign + sizeof(double) * kDotBlocks) + kAlign - 1) / kAlign * kAlign;
static_assert(kHeader < kScratchBytes, "CG scratch header exceeds the array");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
void *scratch_on_current_device() {
  int dev = -1;
  PMG_HIP_CHECK(hipGetDevice(&dev));
  PMG_REQUIRE(dev >= 0 && dev < kMaxDevices, PMG_ERR_UNSUPPORTED,
              "device index beyond the CG scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    PMG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PMG_REQUIRE(p != nullptr, PMG_ERR_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

}  // namespace

int64_t cg_noalloc_max_dofs(int precision) {
  const size_t vb = (kScratchBytes - kHeader) / 4 / kAlign * kAlign;
  return static_cast<int64_t>(vb / element_size(precision));
}

CGResult solve_cg_noalloc(const Operator &op, int precond, const VCycleMultigrid *mg,
                          const void *b, void *x, double rel_tol, int max_it, hipStream_t s) {
  const int64_t n = op.m();
  const size_t es = element_size(op.precision());
  const size_t vb = round_up(static_cast<size_t>(n) * es);
  PMG_REQUIRE(n <= cg_noalloc_max_dofs(op.precision()) && kHeader + 4 * vb <= kScratchBytes,
              PMG_ERR_SIZE,
              "pmg_cg_solve_noalloc: " + std::to_string(n) + " DoFs exceed the fixed work array (" +
                  std::to_string(kScratchBytes >> 20) + " MiB, at most " +
                  std::to_string(cg_noalloc_max_dofs(op.precision())) +
                  " DoFs at this precision); use pmg_cg_solve");
  auto *base = static_cast<unsigned char *>(scratch_on_current_device());
  CGSpace w;
  w.sc = reinterpret_cast<double *>(base);
  w.partials = reinterpret_cast<double *>(base + kAlign);
  w.r = base + kHeader;
  w.z = base + kHeader + vb;
  w.p = base + kHeader + 2 * vb;
  w.q = base + kHeader + 3 * vb;
  w.host_rr = nullptr;  // r.r is copied back at each convergence check
  // the previous call's scalars are still there: start from zero
  launch_fill<double>(8, w.sc, 0.0, s);
  return solve_cg(op, precond, mg, b, x, rel_tol, max_it, w, s);
}

}  // namespace pmg
