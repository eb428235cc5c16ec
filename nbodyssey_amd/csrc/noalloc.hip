// nb_compute_forces_noalloc: the force evaluation of nb_compute_forces for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  The caller passes
// device arrays for the positions and the accelerations; every scratch
// buffer (sort keys and values, tree nodes, COM frontier, look-back words,
// all-pairs partial sums) is carved out of one file-scope device array,
// `scratch`, which each GPU holds from its first use until the process ends.
//
// hipGetSymbolAddress gives each device's own address of that array; it is
// looked up once per device, with the device current, and only that address
// is handed to kernels and copies (the array's name is a host-side stand-in).
//
// Calls share the array of their GPU and keep the last call's contents, so
// calls that overlap on one GPU (two streams, two threads) race.  What must
// start at zero is cleared by kernels on the caller's stream (no memset
// nodes, so the call can be captured into a graph and replayed).
#include "nb_internal.hpp"
#include "nbody.h"

#include <mutex>

namespace nb {
namespace {

// Sized for NB_NOALLOC_MAX_N particles; checked against the carve below at
// each device's set-up.
constexpr size_t kScratchBytes = size_t(12) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
bool g_ready[kMaxDevices] = {};
// Per device: the Barnes-Hut buffers carved for NB_NOALLOC_MAX_N particles
// out of that device's copy of `scratch`.  The layout is the same for every
// call, whatever its n, so the sort state and the epoch-tagged look-back
// words stay where the previous call left them; the all-pairs partial sums
// use the per-particle part only (from keys_a on), never those words.
BhScratch g_bufs[kMaxDevices];

// This device's buffers; on first use, looks up the device's address of
// `scratch`, checks that both layouts fit and clears the sort state, the COM
// list counters and the look-back words.
const BhScratch& device_buffers(int dev, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ready[dev]) return g_bufs[dev];
  void* ptr = nullptr;
  NB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  size_t need = 0;
  const BhScratch b = carve_bh_scratch(ptr, NB_NOALLOC_MAX_N, &need);
  const size_t fixed = (size_t)(reinterpret_cast<char*>(b.keys_a) - static_cast<char*>(ptr));
  const size_t naive = (size_t)allpairs_split(NB_NOALLOC_MAX_N) * NB_NOALLOC_MAX_N *
                       sizeof(float4);
  if (need > kScratchBytes || fixed + naive > kScratchBytes)
    throw Error(kErrState, "static scratch too small for NB_NOALLOC_MAX_N particles");
  launch_fill_u32(s, reinterpret_cast<unsigned*>(b.state), 0u, sizeof(SortState) / 4);
  launch_fill_u32(s, reinterpret_cast<unsigned*>(b.frontier), 0u, kComCounterInts);
  launch_fill_u32(s, reinterpret_cast<unsigned*>(b.lookback), 0u,
                  lookback_words(b.max_tiles) * 2);
  g_bufs[dev] = b;
  g_ready[dev] = true;
  return g_bufs[dev];
}

void need_device_array(const void* p, int dev, const char* what) {
  if (!p) throw Error(kErrInvalid, std::string(what) + " is NULL");
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) (void)hipGetLastError();  // unknown to HIP: plain host memory
  const bool on_device = a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
  if (e != hipSuccess || !on_device || a.device != dev)
    throw Error(kErrInvalid, std::string(what) + " must be device memory of the current device");
}

}  // namespace
}  // namespace nb

extern "C" int nb_compute_forces_noalloc(const float* d_pos4, int n, const nb_params* params,
                                         float* d_acc4, void* stream) {
  using namespace nb;
  try {
    if (!params) throw Error(kErrInvalid, "params is NULL");
    const Params p = from_c(params);
    validate(p);
    if (n < 1 || n > NB_NOALLOC_MAX_N)
      throw Error(kErrInvalid, "n = " + std::to_string(n) + " is outside [1, " +
                                   std::to_string(NB_NOALLOC_MAX_N) +
                                   "], the size of the static scratch");
    int dev = 0;
    NB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(kErrDevice, "device index out of range");
    need_device_array(d_pos4, dev, "d_pos4");
    need_device_array(d_acc4, dev, "d_acc4");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const BhScratch& b = device_buffers(dev, s);
    float4* pos = reinterpret_cast<float4*>(const_cast<float*>(d_pos4));
    float4* acc = reinterpret_cast<float4*>(d_acc4);
    if (p.force == 1 && n >= 2) {
      // histograms and tile tickets start at zero, whatever an earlier call
      // on this device left behind
      launch_sort_init(s, b.state);
      enqueue_bh_force(s, b, pos, nullptr, acc, n, p, false, nullptr);
    } else if (p.force == 1) {
      // n < 2: zero force (forces_bh.cu:466-469)
      launch_fill_u32(s, reinterpret_cast<unsigned*>(acc), 0u, (size_t)n * 4);
    } else {
      launch_allpairs(s, pos, n, p.eps * p.eps, p.G, acc, reinterpret_cast<float4*>(b.keys_a),
                      allpairs_split(n), p.precision == 1, nullptr, 0.f);
    }
    return NB_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return NB_ERR_DEVICE;
  }
}
