// scratch.hip -- device scratch of the allocation-free RHS entry point
// (hype_rhs_noalloc).
//
// A file-scope device array holds what hype_rhs keeps in context-owned
// device memory: the instance descriptor, the time state whose wave-speed
// and error slots the stage kernel reduces into, the tile table and the
// left-inlet table.  Every GPU the library runs on holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel or
// a copy (on the host it is only a stand-in, not a device address): the
// first call on a device looks up that device's address of the array with
// hipGetSymbolAddress, with the device current, and every later call uses
// the stored address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../../include/hype2d.h"
#include "../hype_internal.h"

static __device__ __align__(256) unsigned char scratch[HYPE_SCRATCH_BYTES];

static_assert(sizeof(hype_inst) <= HYPE_SCRATCH_STATE_OFF, "instance descriptor overlaps the time state");
static_assert(HYPE_SCRATCH_STATE_OFF + sizeof(hype_dstate) <= HYPE_SCRATCH_TILES_OFF,
              "time state overlaps the tile table");
static_assert(HYPE_SCRATCH_INLET_OFF <= HYPE_SCRATCH_BYTES, "tile table exceeds the scratch");

#define HYPE_SCRATCH_DEVICES 64

static void* g_scratch_dev[HYPE_SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

extern "C" int hyped_scratch(void** ptr, size_t* bytes) {
  int dev = -1;
  *ptr = nullptr;
  if (hipGetDevice(&dev) != hipSuccess) return HYPE_E_DEVICE;
  if (dev < 0 || dev >= HYPE_SCRATCH_DEVICES) return HYPE_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HYPE_E_DEVICE;
    g_scratch_dev[dev] = p;
  }
  *ptr = g_scratch_dev[dev];
  if (bytes) *bytes = HYPE_SCRATCH_BYTES;
  return HYPE_OK;
}
