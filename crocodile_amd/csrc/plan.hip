// Plans: the binning of one set of coordinates, reused by any number of
// gridding and degridding calls.  An iterative reconstruction grids and
// degrids the same visibilities many times (the major cycle of
// scripts/image_dataset.py-style imaging followed by prediction); the
// one-shot croc_convgrid / croc_convdegrid bin on every call, a plan bins
// once.  croc_plan_bin validates every coordinate (the reference's
// exceptions are raised there); the plan calls are then fully asynchronous.
#include "internal.h"

struct croc_plan {
  int device = 0;
  WsSet ws;           // the plan's own buffers (records, work items, scan space)
  BinPlan bp;
  bool ready = false;
  int N = 0, Q = 0, separable = 0, nplanes = 0;
  int64_t nv = 0;
  const int64_t *plane_off = nullptr;
  const int32_t *plane_s = nullptr;
};

namespace {

int plan_check(croc_ctx *ctx, const croc_plan *plan) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(plan != nullptr, CROC_EINVAL, "null plan");
  CROC_REQUIRE(plan->ready, CROC_EINVAL, "plan holds no successful binning");
  CROC_REQUIRE(plan->device == ctx->device, CROC_EINVAL, "plan belongs to another device");
  return CROC_OK;
}

}  // namespace

extern "C" int croc_plan_create(croc_ctx *ctx, croc_plan **out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(out != nullptr, CROC_EINVAL, "null output pointer");
  croc_plan *p = new croc_plan();
  p->device = ctx->device;
  *out = p;
  return CROC_OK;
}

extern "C" int croc_plan_destroy(croc_ctx *ctx, croc_plan *plan) {
  if (!plan) return CROC_OK;
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  (void)hipStreamSynchronize(ctx->stream);
  ws_free(plan->ws);
  delete plan;
  return CROC_OK;
}

extern "C" int croc_plan_bin(croc_ctx *ctx, croc_plan *plan, int N, const double *uvw,
                             int64_t nv, int64_t stride, double lam, int Q, int separable,
                             int nplanes, const int64_t *plane_off, const int32_t *plane_s,
                             int max_support, int plane_mode, const int32_t *vplane,
                             double wstep, int wbin0) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  CROC_REQUIRE(plan != nullptr, CROC_EINVAL, "null plan");
  CROC_REQUIRE(plan->device == ctx->device, CROC_EINVAL, "plan belongs to another device");
  plan->ready = false;
  CROC_REQUIRE(plane_off && plane_s, CROC_EINVAL, "null array argument");
  CROC_REQUIRE(nv == 0 || uvw, CROC_EINVAL, "null uvw");
  CROC_REQUIRE(stride >= 2, CROC_EINVAL, "uvw stride must be >= 2");
  CROC_REQUIRE(plane_mode != CROC_PLANE_INDEX || vplane, CROC_EINVAL,
               "CROC_PLANE_INDEX needs vplane");
  CROC_REQUIRE(plane_mode != CROC_PLANE_WBIN || (stride >= 3 && wstep > 0),
               CROC_EINVAL, "CROC_PLANE_WBIN needs a w column and wstep > 0");
  CROC_REQUIRE(lam > 0, CROC_EINVAL, "lam must be positive");
  plan->N = N;
  plan->Q = Q;
  plan->separable = separable;
  plan->nplanes = nplanes;
  plan->nv = nv;
  plan->plane_off = plane_off;
  plan->plane_s = plane_s;
  if (nv > 0) {
    BinArgs a{N, uvw, nv, stride, lam, Q, nplanes, plane_s, plane_off, separable,
              plane_mode, vplane, wstep, wbin0, max_support};
    rc = bin_visibilities(ctx, plan->ws, a, max_support, ITEMS_BY_TILE | ITEMS_BY_GROUP,
                         &plan->bp);
    if (rc) return rc;
    rc = read_status(ctx, plan->bp);
    if (rc) return rc;
  }
  plan->ready = true;
  return CROC_OK;
}

extern "C" int croc_convgrid_plan(croc_ctx *ctx, croc_plan *plan, void *grid,
                                  const void *vis, const void *kern) {
  int rc = plan_check(ctx, plan);
  if (rc) return rc;
  CROC_REQUIRE(grid && kern, CROC_EINVAL, "null array argument");
  if (plan->nv == 0) return CROC_OK;
  CROC_REQUIRE(vis != nullptr, CROC_EINVAL, "null vis");
  rc = zero_u32_async(ctx, plan->bp.qhead, XCD_QUEUES + 1);  // work queues restart
  if (rc) return rc;
  return grid_binned(ctx, plan->bp, grid, plan->N, plan->nv, vis, kern, plan->Q,
                     plan->separable, plan->nplanes, plan->plane_off, plan->plane_s);
}

extern "C" int croc_convdegrid_plan(croc_ctx *ctx, croc_plan *plan, const void *grid,
                                    void *vis_out, const void *kern) {
  int rc = plan_check(ctx, plan);
  if (rc) return rc;
  CROC_REQUIRE(grid && kern, CROC_EINVAL, "null array argument");
  if (plan->nv == 0) return CROC_OK;
  CROC_REQUIRE(vis_out != nullptr, CROC_EINVAL, "null vis_out");
  rc = zero_u32_async(ctx, plan->bp.qhead, XCD_QUEUES + 1);  // work queues restart
  if (rc) return rc;
  return degrid_binned(ctx, plan->bp, grid, plan->N, vis_out, kern, plan->separable);
}
