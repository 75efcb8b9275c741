// Green's function / overlap entry point that allocates nothing.
//
// afqmc_greens_function_noalloc is afqmc_greens_function without a context:
// the trial orbitals come in as arguments and the work buffers of the overlap
// LU and the inverse (the O matrices, their inverses and the per-walker
// log|det| / phase partials) live in `scratch`, a static device array of
// AFQMC_NOALLOC_SCRATCH_BYTES that the code object carries.  Each GPU the
// library runs on holds its own copy until the process ends; its device
// address is looked up once per device (hipGetSymbolAddress with that device
// current) and kept.  The kernels are the ones afqmc_greens_function launches.
// Every call on one GPU shares that GPU's array, so calls that overlap in time
// on one GPU (two host threads, or two streams whose work interleaves) race.
// The kernels write every scratch entry they read, so nothing is cleared.
#include <cstdio>
#include <mutex>

#include "afqmc_internal.hpp"
#include "../../include/afqmc_hip.h"

namespace {

constexpr size_t kScratchBytes = AFQMC_NOALLOC_SCRATCH_BYTES;
constexpr size_t kHalf = kScratchBytes / 2;  // O and its partials | O^{-1}
static_assert(kHalf % 256 == 0, "both halves stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];
thread_local char g_err[256] = "";

int nerr(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

// Device address of `scratch` on the current device, looked up on first use.
int scratch_ptr(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return nerr(AFQMC_ERR_HIP, "no current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      snprintf(g_err, sizeof(g_err), "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
               hipGetErrorString(e));
      return AFQMC_ERR_HIP;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return AFQMC_OK;
}

}  // namespace

extern "C" {

size_t afqmc_noalloc_scratch_bytes(int32_t nwalkers, int32_t nalpha, int32_t nbeta) {
  if (nwalkers < 1 || nalpha < 1 || nbeta < 0) return 0;
  return 2 * afqmc::gf_work_bytes(nwalkers, nalpha, nbeta);
}

const char* afqmc_noalloc_last_error(void) { return g_err; }

int afqmc_greens_function_noalloc(int32_t nwalkers, int32_t nbasis, int32_t nalpha, int32_t nbeta, int32_t rhf,
                                  const double* psia, const double* psib, const double* phia, const double* phib,
                                  double* ghalfa, double* ghalfb, double* ovlp, const double* log_shift,
                                  int32_t* walker_flags, void* stream) {
  g_err[0] = '\0';
  if (nwalkers < 1 || nbasis < 1 || nalpha < 1 || nbeta < 0)
    return nerr(AFQMC_ERR_SHAPE, "nwalkers, nbasis, nalpha must be positive and nbeta non-negative");
  if (nalpha > 64 || nbeta > 64 || nalpha > nbasis || nbeta > nbasis)
    return nerr(AFQMC_ERR_SHAPE, "nalpha and nbeta must be at most 64 and at most nbasis");
  const bool two = nbeta > 0 && !rhf;
  if (!psia || !phia || !ovlp || (two && (!psib || !phib)))
    return nerr(AFQMC_ERR_INVALID_ARG, "null trial, walker or overlap array");
  const bool want_ghalf = ghalfa != nullptr;
  if (want_ghalf && two && ghalfb == nullptr)
    return nerr(AFQMC_ERR_INVALID_ARG, "Ghalf of both spins, or of neither (overlap only)");
  const size_t half = afqmc::gf_work_bytes(nwalkers, nalpha, nbeta);
  if (half > kHalf) {
    snprintf(g_err, sizeof(g_err),
             "scratch of %zu bytes needed for %d walkers (nalpha %d, nbeta %d) exceeds the static %zu bytes",
             2 * half, nwalkers, nalpha, nbeta, kScratchBytes);
    return AFQMC_ERR_SHAPE;
  }
  void* base = nullptr;
  int r = scratch_ptr(&base);
  if (r) return r;
  afqmc::GfArgs a{};
  a.nwalkers = nwalkers;
  a.nbasis = nbasis;
  a.nalpha = nalpha;
  a.nbeta = nbeta;
  a.rhf = rhf ? 1 : 0;
  a.phia = phia;
  a.phib = two ? phib : nullptr;
  a.psia = psia;
  a.psib = two ? psib : nullptr;
  a.ghalfa = ghalfa;
  a.ghalfb = two ? ghalfb : nullptr;
  a.ovlp = ovlp;
  a.log_shift = log_shift;
  a.flags = nullptr;
  a.wflags = walker_flags;
  a.psi_real = 0;
  double* obuf = static_cast<double*>(base);
  double* oinv = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kHalf);
  const hipError_t e = afqmc::launch_gf(a, want_ghalf, obuf, oinv, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "Green's function launch failed: %s", hipGetErrorString(e));
    return e == hipErrorInvalidValue ? AFQMC_ERR_SHAPE : AFQMC_ERR_HIP;
  }
  return AFQMC_OK;
}

}  // extern "C"
