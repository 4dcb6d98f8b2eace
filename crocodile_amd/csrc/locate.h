// Location of one visibility for the gridding kernels: FP64 frac_coord
// (synthesis.py:419-449), kernel plane (make_gcf :451-461, w-bin of
// w_cache_imaging :793), window and table-offset validation, and its bin
// (tile, plane-group) for the binning of bin.hip.
#pragma once

#include "internal.h"

constexpr uint32_t LOC_INVALID = 0xffffffffu;

struct BinGeom {
  int T, ntx, ngroups;
  int64_t group_elems;  // plane-group width in bank elements
};

__device__ __forceinline__ uint4 locate(const BinArgs &a, const BinGeom &g,
                                        int64_t i, uint32_t *st) {
  const double *pu = a.uvw + i * a.stride;
  const double px = grid_fraction(pu[0], a.lam);
  const double py = grid_fraction(pu[1], a.lam);
  int ix, xf, iy, yf;
  const bool okx = frac_coord_dev(a.N, a.Q, px, &ix, &xf);
  const bool oky = frac_coord_dev(a.N, a.Q, py, &iy, &yf);
  if (!(okx && oky)) {
    *st |= ST_DOMAIN;
    return make_uint4(0, 0, 0, LOC_INVALID);
  }
  int64_t plane = 0;
  if (a.plane_mode == CROC_PLANE_INDEX) {
    plane = a.vplane[i];
  } else if (a.plane_mode == CROC_PLANE_WBIN) {
    plane = wbin_dev(pu[2], a.wstep) - a.wbin0;
  }
  if (plane < 0 || plane >= a.nplanes) {
    *st |= ST_PLANE;
    return make_uint4(0, 0, 0, LOC_INVALID);
  }
  const int S = a.plane_s[plane];
  if (S < 1 || S > a.max_support) {  // LDS regions are sized for max_support
    *st |= ST_SUPPORT;
    return make_uint4(0, 0, 0, LOC_INVALID);
  }
  const int x0 = ix - S / 2, y0 = iy - S / 2;
  if (x0 < 0 || y0 < 0 || x0 + S > a.N || y0 + S > a.N) {
    *st |= ST_BOUNDS;
    return make_uint4(0, 0, 0, LOC_INVALID);
  }
  const int64_t poff = a.plane_off[plane];
  const int64_t toff =
      a.separable ? poff : poff + ((int64_t)yf * a.Q + xf) * S * S;
  const int64_t tend = toff + (a.separable ? (int64_t)a.Q * S : (int64_t)S * S);
  if (poff < 0 || tend > (int64_t)UINT32_MAX) {
    *st |= ST_BANK;
    return make_uint4(0, 0, 0, LOC_INVALID);
  }
  const int grp = (int)min((int64_t)g.ngroups - 1, toff / g.group_elems);
  const uint32_t tile = (uint32_t)((y0 / g.T) * g.ntx + x0 / g.T);
  return make_uint4((uint32_t)x0 | ((uint32_t)y0 << 16),
                    (uint32_t)S | ((uint32_t)xf << 8) | ((uint32_t)yf << 20),
                    (uint32_t)toff, tile * (uint32_t)g.ngroups + (uint32_t)grp);
}

