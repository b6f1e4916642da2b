// Allocation-free entry point (no reference counterpart), for callers that
// cannot allocate device memory: code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc.
//
// rfhip_op_energy_noalloc is rfhip_op_state_energy (rhs.rs:1509-1544) on
// caller-owned device arrays. Its only scratch, the per-block partial sums
// of the energy reduction, lives in the file-scope device array `scratch`
// below (kMaxEnergyPartials doubles, 32 KiB, held by every GPU the library
// runs on until the process ends). Host code never passes the array's name
// to a kernel: rfhip_ctx_create looks up its device address once per device
// (noalloc_init) and the entry point passes that address.
//
// Every call on one GPU shares that GPU's array. The reduction's first kernel
// writes every partial the second reads, so nothing has to start at zero;
// two calls that overlap on one GPU (different streams, or different host
// threads) race on the partials and must not be issued.
#include <mutex>

#include "rfhip_impl.hpp"

namespace {
// one partial per block of energy_elems_per_block() (256) elements:
// K <= 4096 * 256 = 1,048,576 elements
constexpr int kMaxEnergyPartials = 4096;
}  // namespace

static __device__ __align__(256) double scratch[kMaxEnergyPartials];

namespace rfhip {
namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device ordinal

double *scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

bool on_device(const void *p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error of an unregistered pointer
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}
}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  RF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double *>(p);
}

}  // namespace rfhip

using namespace rfhip;

extern "C" int rfhip_op_energy_noalloc(rfhip_op *op, const double *y, const double *mass_ref, const double *det,
                                       const double *inv_eps, const double *inv_mu, double *energy) {
  try {
    RF_REQUIRE(op && y && mass_ref && det && inv_eps && inv_mu && energy, "null argument");
    const int64_t K = op->n_elem;
    const int per = energy_elems_per_block();
    RF_REQUIRE(K >= 1 && K <= (int64_t)kMaxEnergyPartials * per,
               "rfhip_op_energy_noalloc: " + std::to_string(K) + " elements exceed the scratch of " +
                   std::to_string(kMaxEnergyPartials) + " partial sums (at most " +
                   std::to_string((int64_t)kMaxEnergyPartials * per) + " elements)");
    RF_HIP(hipSetDevice(op->ctx->device));
    double *part = scratch_on(op->ctx->device);
    RF_REQUIRE(part != nullptr, "the scratch array of this device was not set up (no context on it)");
    for (const void *q : {(const void *)y, (const void *)mass_ref, (const void *)det, (const void *)inv_eps,
                          (const void *)inv_mu, (const void *)energy})
      RF_REQUIRE(on_device(q), "rfhip_op_energy_noalloc takes device pointers only");
    const int nblk = (int)((K + per - 1) / per);
    launch_energy_scaled(y, mass_ref, det, inv_eps, inv_mu, op->np, K, part, nblk, 0.5, energy, op->stream);
    RF_HIP(hipStreamSynchronize(op->stream));  // *energy is ready on return, as in rfhip_op_state_energy
    return RFHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return RFHIP_EINVAL;
  }
}
