// Allocation-free convolutional degridding (croc_convdegrid_noalloc).
//
// croc_convdegrid bins its coordinates into scratch buffers the context
// allocates and grows on demand (hipMalloc).  Callers that must not
// allocate -- code under HIP graph capture, tight serving loops, harnesses
// that forbid it -- call this variant instead: the same binning and tile
// kernels, with every scratch buffer (located records, tile-sorted records,
// bin counters and offsets, scan space, work items, status word and queue
// heads) carved from one file-scope static device array.  Each GPU holds its
// own copy of the array from module load until the process ends.
//
// Host code never uses the array's name as a device address: its address on
// each device is looked up once, with that device current, when the first
// context of the device is created (noalloc_init, called by croc_ctx_create)
// and kept per device; the kernels receive pointers derived from it.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; the binning clears what must start at zero (bin counters, status,
// queue heads) with fill kernels, so no state carries over from one call to
// the next.  Calls that overlap on one GPU (two streams, two host threads)
// race on the array and must be serialised by the caller.
#include <mutex>

#include "internal.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)CROC_NOALLOC_SCRATCH_MIB << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mutex;
char *g_scratch[MAX_DEVICES] = {};  // device address of `scratch` per device

}  // namespace

int noalloc_init(int device) {
  CROC_REQUIRE(device >= 0 && device < MAX_DEVICES, CROC_EINVAL,
               "device %d beyond the %d devices the static scratch table holds", device,
               MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return CROC_OK;
  int cur = -1;
  CROC_HIP(hipGetDevice(&cur));
  CROC_REQUIRE(cur == device, CROC_EINVAL, "noalloc_init: device %d is not current", device);
  void *ptr = nullptr;
  CROC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  CROC_REQUIRE(ptr != nullptr, CROC_EHIP, "hipGetSymbolAddress returned no address");
  g_scratch[device] = (char *)ptr;
  return CROC_OK;
}

extern "C" int croc_convdegrid_noalloc(croc_ctx *ctx, const void *grid, int N,
                                       const double *uvw, int64_t nv, int64_t stride,
                                       double lam, void *vis_out, const void *kern, int Q,
                                       int separable, int nplanes, const int64_t *plane_off,
                                       const int32_t *plane_s, int max_support,
                                       int plane_mode, const int32_t *vplane, double wstep,
                                       int wbin0) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(grid && kern && plane_off && plane_s, CROC_EINVAL, "null array argument");
  CROC_REQUIRE(nv == 0 || (uvw && vis_out), CROC_EINVAL, "null uvw / vis_out");
  CROC_REQUIRE(stride >= 2, CROC_EINVAL, "uvw stride must be >= 2");
  CROC_REQUIRE(plane_mode != CROC_PLANE_INDEX || vplane, CROC_EINVAL,
               "CROC_PLANE_INDEX needs vplane");
  CROC_REQUIRE(plane_mode != CROC_PLANE_WBIN || (stride >= 3 && wstep > 0), CROC_EINVAL,
               "CROC_PLANE_WBIN needs a w column and wstep > 0");
  CROC_REQUIRE(lam > 0, CROC_EINVAL, "lam must be positive");
  CROC_REQUIRE(nv >= 0 && nv <= CROC_NOALLOC_MAX_NV, CROC_EINVAL,
               "croc_convdegrid_noalloc takes at most %d visibilities (%lld given)",
               CROC_NOALLOC_MAX_NV, (long long)nv);
  if (nv == 0) return CROC_OK;
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    base = g_scratch[ctx->device];
  }
  if (!base) {  // a context made before the table existed: look it up now
    if ((rc = noalloc_init(ctx->device))) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    base = g_scratch[ctx->device];
  }
  WsSet set;  // every slot carved from the static array, nothing allocated
  set.arena = base;
  set.arena_size = NOALLOC_SCRATCH_BYTES;
  BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
            plane_mode, vplane, wstep, wbin0, max_support};
  BinPlan plan;
  rc = bin_visibilities(ctx, set, a, max_support, ITEMS_BY_GROUP, &plan);
  if (rc == CROC_ENOMEM) return CROC_EINVAL;  // beyond the static maximum (message set)
  if (rc) return rc;
  rc = degrid_binned(ctx, plan, grid, N, vis_out, kern, separable);
  if (rc) return rc;
  return read_status(ctx, plan);
}
