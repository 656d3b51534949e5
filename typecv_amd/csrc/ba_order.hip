// ba_order.hip — the solve entries' point ordering on the device, included by ba_engine.hip.
//
// The one-pass Schur product needs each workgroup's OP_TILES consecutive point tiles to observe at
// most OP_SPAN cameras, which holds when points come ordered by the cameras that see them. The
// points are put in the lowest, then the highest camera of each track (stable, directional points
// after the others) when that order fits where the given one does not or writes fewer partial
// rows: these kernels measure each track's cameras, check the index words the reordering relies
// on, gather the reordered scene and index, and put the solved points back in the given order.
// The choice and the sorts of the keys run on the host (counting sorts); each camera's list of the
// reordered observations is a stable radix sort of their cameras on the device.
#pragma once

#include <hipcub/hipcub.hpp>

namespace tcvba {

constexpr int ORDER_THREADS = 256;

// One thread per point: the extremes of its track's cameras, plain and wrapped round the camera
// loop (a camera k < C/2 counted as k + C) as four words (lo, hi, wrapped lo, wrapped hi; lo > hi
// for an empty track), and whether any of its observations names another point or no camera of C
// (flag[0] = 1).
__global__ __launch_bounds__(ORDER_THREADS) void k_order_spans(const uint32_t* ocam,
                                                               const uint32_t* opt,
                                                               const uint32_t* ps, uint32_t P,
                                                               uint32_t C, uint32_t* span,
                                                               uint32_t* flag) {
  const uint32_t p = blockIdx.x * ORDER_THREADS + threadIdx.x;
  if (p >= P) return;
  const uint32_t half = C / 2;
  uint32_t lo = UINT32_MAX, hi = 0, wlo = UINT32_MAX, whi = 0;
  bool bad = false;
  for (uint32_t i = ps[p]; i < ps[p + 1]; i++) {
    const uint32_t cam = ocam[i], v = cam < half ? cam + C : cam;
    bad |= opt[i] != p || cam >= C;
    lo = min(lo, cam);
    hi = max(hi, cam);
    wlo = min(wlo, v);
    whi = max(whi, v);
  }
  span[4 * (size_t)p] = lo;
  span[4 * (size_t)p + 1] = hi;
  span[4 * (size_t)p + 2] = wlo;
  span[4 * (size_t)p + 3] = whi;
  if (bad) flag[0] = 1;
}

// One workgroup per camera: whether its list names observations of the index (< M) that name it,
// ascending (flag[0] = 1 otherwise), as the engine's own checks require.
__global__ __launch_bounds__(ORDER_THREADS) void k_order_lists(const uint32_t* ocam,
                                                               const uint32_t* cs,
                                                               const uint32_t* co, uint32_t M,
                                                               uint32_t* flag) {
  const uint32_t k = blockIdx.x, n0 = cs[k], n1 = cs[k + 1];
  bool bad = false;
  for (uint32_t n = n0 + threadIdx.x; n < n1; n += ORDER_THREADS) {
    const uint32_t i = co[n];
    bad |= i >= M || ocam[min(i, M - 1)] != k || (n > n0 && co[n - 1] >= i);
  }
  if (bad) flag[0] = 1;
}

// One thread per new point j (the given point order[j]): its coordinates, and its observations'
// cameras, points and pixels in their given order from nps[j] on.
__global__ __launch_bounds__(ORDER_THREADS) void k_order_gather(
    const uint32_t* order, const uint32_t* ps, const uint32_t* nps, const uint32_t* ocam,
    const double* points, const double* pixels, uint32_t P, uint32_t* ncam, uint32_t* npt,
    double* npoints, double* npixels) {
  const uint32_t j = blockIdx.x * ORDER_THREADS + threadIdx.x;
  if (j >= P) return;
  const uint32_t g = order[j];
  for (int k = 0; k < 3; k++) npoints[3 * (size_t)j + k] = points[3 * (size_t)g + k];
  for (uint32_t i = ps[g], n = nps[j]; i < ps[g + 1]; i++, n++) {
    ncam[n] = ocam[i];
    npt[n] = j;
    npixels[2 * (size_t)n] = pixels[2 * (size_t)i];
    npixels[2 * (size_t)n + 1] = pixels[2 * (size_t)i + 1];
  }
}

__global__ __launch_bounds__(ORDER_THREADS) void k_order_iota(uint32_t* v, uint32_t n) {
  const uint32_t i = blockIdx.x * ORDER_THREADS + threadIdx.x;
  if (i < n) v[i] = i;
}

// The solved points back in the given order: point order[j] of `out` takes solved point j.
__global__ __launch_bounds__(ORDER_THREADS) void k_order_scatter(const uint32_t* order,
                                                                 const double* solved, uint32_t P,
                                                                 double* out_points) {
  const uint32_t j = blockIdx.x * ORDER_THREADS + threadIdx.x;
  if (j >= P) return;
  const size_t g = order[j];
  for (int k = 0; k < 3; k++) out_points[3 * g + k] = solved[3 * (size_t)j + k];
}

}  // namespace tcvba
