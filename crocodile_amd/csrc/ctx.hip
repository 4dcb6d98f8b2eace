// Context, workspace and error plumbing of libcrocgpu.
#include <stdarg.h>
#include <stdio.h>

#include "internal.h"

static thread_local char g_err[512] = "";

void croc_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char *croc_last_error(void) { return g_err; }

extern "C" int croc_version(void) { return 1; }

int croc_status_to_code(uint32_t st) {
  if (st & ST_DOMAIN) {
    croc_set_error("grid coordinate outside [-0.5, 0.5)");
    return CROC_EDOMAIN;
  }
  if (st & ST_PLANE) {
    croc_set_error("w-bin / plane index outside the kernel bank");
    return CROC_EPLANE;
  }
  if (st & ST_RANGE) {
    croc_set_error("time / frequency window or array range outside a baseline's data");
    return CROC_EINVAL;
  }
  if (st & ST_BANK) {
    croc_set_error("kernel bank larger than 2^32 complex elements");
    return CROC_EINVAL;
  }
  if (st & ST_SUPPORT) {
    croc_set_error("a kernel plane's support exceeds max_support");
    return CROC_EINVAL;
  }
  if (st & ST_BOUNDS) {
    croc_set_error("kernel window extends beyond the grid");
    return CROC_EBOUNDS;
  }
  return CROC_OK;
}

int ctx_enter(croc_ctx *ctx) {
  CROC_REQUIRE(ctx != nullptr, CROC_EINVAL, "null context");
  CROC_HIP(hipSetDevice(ctx->device));
  return CROC_OK;
}

void *ws_get_in(croc_ctx *ctx, WsSet &set, WsSlot slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (set.ws_size[slot] >= bytes) return set.ws[slot];
  if (set.arena) {  // fixed arena: carve, never allocate
    const size_t off = (set.arena_used + 255) & ~(size_t)255;
    if (off > set.arena_size || bytes > set.arena_size - off) {
      croc_set_error("scratch needs more than the %zu-byte static scratch array "
                     "(%zu bytes requested at offset %zu)", set.arena_size, bytes, off);
      return nullptr;
    }
    set.ws[slot] = set.arena + off;
    set.ws_size[slot] = bytes;
    set.arena_used = off + bytes;
    return set.ws[slot];
  }
  if (set.ws[slot]) {
    // The stream may still use the old buffer.
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      croc_set_error("stream synchronisation failed while growing workspace");
      return nullptr;
    }
    (void)hipFree(set.ws[slot]);
    set.ws[slot] = nullptr;
    set.ws_size[slot] = 0;
  }
  size_t want = bytes + bytes / 4;  // headroom against repeated growth
  void *p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    croc_set_error("hipMalloc of %zu bytes failed", want);
    return nullptr;
  }
  set.ws[slot] = p;
  set.ws_size[slot] = want;
  return p;
}

void *ws_get(croc_ctx *ctx, WsSlot slot, size_t bytes) {
  return ws_get_in(ctx, ctx->ws, slot, bytes);
}

void ws_free(WsSet &set) {
  if (set.arena) return;  // not ours
  for (int i = 0; i < WS_COUNT; i++) {
    if (set.ws[i]) (void)hipFree(set.ws[i]);
    set.ws[i] = nullptr;
    set.ws_size[i] = 0;
  }
}

void *host_stage(croc_ctx *ctx, size_t bytes) {
  if (ctx->stage_size >= bytes) return ctx->stage;
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  ctx->stage = nullptr;
  ctx->stage_size = 0;
  const size_t want = bytes + bytes / 4;
  if (hipHostMalloc(&ctx->stage, want, hipHostMallocDefault) != hipSuccess) {
    ctx->stage = nullptr;
    croc_set_error("hipHostMalloc of %zu bytes failed", want);
    return nullptr;
  }
  ctx->stage_size = want;
  return ctx->stage;
}

extern "C" int croc_ctx_create(int device, croc_ctx **out) {
  CROC_REQUIRE(out != nullptr, CROC_EINVAL, "null output pointer");
  int ndev = 0;
  CROC_HIP(hipGetDeviceCount(&ndev));
  CROC_REQUIRE(device >= 0 && device < ndev, CROC_EINVAL,
               "device %d not present (%d devices)", device, ndev);
  CROC_HIP(hipSetDevice(device));
  int rc = noalloc_init(device);
  if (rc) return rc;
  croc_ctx *ctx = new croc_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    ctx->num_cu = prop.multiProcessorCount;
  if (hipHostMalloc((void **)&ctx->host_status, 64, hipHostMallocDefault) !=
      hipSuccess) {
    delete ctx;
    croc_set_error("hipHostMalloc failed");
    return CROC_ENOMEM;
  }
  for (int i = 0; i < 4; i++) {
    if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
      croc_set_error("hipEventCreate failed");
      return CROC_EHIP;
    }
  }
  *out = ctx;
  return CROC_OK;
}

extern "C" int croc_ctx_destroy(croc_ctx *ctx) {
  if (!ctx) return CROC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ws_free(ctx->ws);
  for (int i = 0; i < 4; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  for (int d = 0; d < 2; d++)
    for (hipEvent_t e : ctx->tev[d]) (void)hipEventDestroy(e);
  if (ctx->host_status) (void)hipHostFree(ctx->host_status);
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  delete ctx;
  return CROC_OK;
}

extern "C" int croc_ctx_set_stream(croc_ctx *ctx, void *stream) {
  CROC_REQUIRE(ctx != nullptr, CROC_EINVAL, "null context");
  ctx->stream = (hipStream_t)stream;
  return CROC_OK;
}

extern "C" int croc_ctx_set_tuning(croc_ctx *ctx, int tile, int chunk,
                                   int group_kib) {
  CROC_REQUIRE(ctx != nullptr, CROC_EINVAL, "null context");
  CROC_REQUIRE(tile == 0 || (tile >= 16 && tile <= 80 && tile % 16 == 0),
               CROC_EINVAL, "tile must be 0 or a multiple of 16 in [16, 80]");
  CROC_REQUIRE(chunk >= 64, CROC_EINVAL, "chunk must be >= 64");
  CROC_REQUIRE(group_kib >= 0, CROC_EINVAL, "group_kib must be >= 0");
  ctx->tile = tile;
  ctx->chunk = chunk;
  ctx->group_kib = group_kib;
  return CROC_OK;
}

extern "C" int croc_last_kernel_ms(croc_ctx *ctx, float *grid_ms,
                                   float *degrid_ms) {
  CROC_REQUIRE(ctx != nullptr, CROC_EINVAL, "null context");
  CROC_HIP(hipSetDevice(ctx->device));
  float g = 0.f, d = 0.f;
  if (ctx->ev_grid) {
    CROC_HIP(hipEventSynchronize(ctx->ev[1]));
    CROC_HIP(hipEventElapsedTime(&g, ctx->ev[0], ctx->ev[1]));
  }
  if (ctx->ev_degrid) {
    CROC_HIP(hipEventSynchronize(ctx->ev[3]));
    CROC_HIP(hipEventElapsedTime(&d, ctx->ev[2], ctx->ev[3]));
  }
  if (grid_ms) *grid_ms = g;
  if (degrid_ms) *degrid_ms = d;
  return CROC_OK;
}

void tap_timer_reset(croc_ctx *ctx, int dir) { ctx->tev_n[dir] = 0; }

int tap_timer_begin(croc_ctx *ctx, int dir) {
  std::vector<hipEvent_t> &v = ctx->tev[dir];
  const size_t k = (size_t)ctx->tev_n[dir];
  while (v.size() < 2 * k + 2) {  // events are made once and reused
    hipEvent_t e = nullptr;
    CROC_HIP(hipEventCreate(&e));
    v.push_back(e);
  }
  CROC_HIP(hipEventRecord(v[2 * k], ctx->stream));
  return CROC_OK;
}

int tap_timer_end(croc_ctx *ctx, int dir) {
  CROC_HIP(hipEventRecord(ctx->tev[dir][2 * (size_t)ctx->tev_n[dir] + 1], ctx->stream));
  ctx->tev_n[dir]++;
  return CROC_OK;
}

extern "C" int croc_last_tap_kernel_ms(croc_ctx *ctx, float *grid_ms, int *grid_launches,
                                       float *degrid_ms, int *degrid_launches) {
  CROC_REQUIRE(ctx != nullptr, CROC_EINVAL, "null context");
  CROC_HIP(hipSetDevice(ctx->device));
  float tot[2] = {0.f, 0.f};
  for (int d = 0; d < 2; d++) {
    for (int k = 0; k < ctx->tev_n[d]; k++) {
      float ms = 0.f;
      CROC_HIP(hipEventSynchronize(ctx->tev[d][2 * (size_t)k + 1]));
      CROC_HIP(hipEventElapsedTime(&ms, ctx->tev[d][2 * (size_t)k], ctx->tev[d][2 * (size_t)k + 1]));
      tot[d] += ms;
    }
  }
  if (grid_ms) *grid_ms = tot[0];
  if (grid_launches) *grid_launches = ctx->tev_n[0];
  if (degrid_ms) *degrid_ms = tot[1];
  if (degrid_launches) *degrid_launches = ctx->tev_n[1];
  return CROC_OK;
}

int persistent_blocks(croc_ctx *ctx, int per_cu) {
  // A multiple of 8 so that every XCD receives the same number of blocks.
  int g = ctx->num_cu * per_cu;
  return (g + 7) / 8 * 8;
}
