/*
 * The allocation-free entry point, libpt_amd_energy_noalloc: the call of
 * libpt_amd_energy on device-resident inputs for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference libpt.
 *
 * Every device buffer the call needs beyond the caller's inputs comes
 * from `scratch`, one static array per GPU, which each GPU the library
 * runs on holds until the process ends.  Its device address is looked up
 * once per device, with that device current, at the first call on it;
 * host code only ever passes that address (never the array's name) to
 * the engine.  Calls on one GPU share the array and would race on it, so
 * the entry point holds a per-device lock for the whole call: calls from
 * several threads run one after another.  The array keeps the last
 * call's contents; the engine clears what must start at zero (the
 * per-item energies and the scale maximum) on every call.
 */
#include <hip/hip_runtime.h>

#include <math.h>

#include <mutex>

#include "../../include/libpt_amd.h"
#include "pt_engine.h"

static __device__ __align__(256) unsigned char
    scratch[LIBPT_AMD_NOALLOC_SCRATCH];

namespace {

constexpr int MAXDEV = 64;

/* device address of `scratch` on each device (null until looked up) and
 * the lock that serialises the calls on it */
void *g_scratch[MAXDEV];
std::mutex g_lock[MAXDEV];

} /* namespace */

extern "C" double
libpt_amd_energy_noalloc(int call, int form, size_t oa, size_t va,
    size_t ob, size_t vb, const void *const *dev_arrays, void *stream)
{
	pt_set_error(nullptr);
	const int dev = pt_device_init();
	if (dev < 0)
		return NAN;
	if (dev >= MAXDEV) {
		pt_set_error("device index beyond the allocation-free entry "
		    "point's table");
		return NAN;
	}
	std::lock_guard<std::mutex> lk(g_lock[dev]);
	if (!g_scratch[dev]) {
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr) {
			pt_set_error("hipGetSymbolAddress of the static scratch "
			    "failed");
			return NAN;
		}
		g_scratch[dev] = ptr;
	}
	return pt_noalloc_energy(call, form, oa, va, ob, vb, dev_arrays,
	    g_scratch[dev], LIBPT_AMD_NOALLOC_SCRATCH, stream);
}
