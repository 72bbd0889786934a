// Scratch of the allocation-free entry point (evoc_mi355x_knn_graph_noalloc).
//
// One file-scope device array, `scratch`, of EVOC_MI355X_NOALLOC_SCRATCH_BYTES:
// every GPU the library runs on holds its own copy, at its own address, from
// the time this code object is loaded on it until the process ends.  In host
// code the array's name is a host-side stand-in, not a device address, so the
// host never passes it anywhere: noalloc_scratch_init looks up each device's
// address once (hipGetSymbolAddress, with that device current) and every user
// takes that pointer.  Calls on one device share the array; the lock returned
// with it serialises them within this process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "evoc_host.h"
#include "evoc_mi355x.h"

static __device__ __align__(256) uint8_t scratch[EVOC_MI355X_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_init_mu;
bool g_looked_up[MAX_DEVICES] = {};
void* g_ptr[MAX_DEVICES] = {};
std::mutex g_call_mu[MAX_DEVICES];

}  // namespace

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return EVOC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_looked_up[device]) return g_ptr[device] ? EVOC_OK : EVOC_ERR_HIP;
  g_looked_up[device] = true;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) {
    (void)hipGetLastError();
    g_looked_up[device] = false;  // not current: look it up on a later call
    return EVOC_ERR_HIP;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return EVOC_ERR_HIP;
  }
  g_ptr[device] = p;
  return EVOC_OK;
}

void* noalloc_scratch(int device, std::mutex** lock) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (lock) *lock = &g_call_mu[device];
  return g_ptr[device];
}
