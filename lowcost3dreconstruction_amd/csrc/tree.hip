// tree.hip -- the tree order of a batch of clouds (see Tree in lcr_common.hpp).
//
// The bounding-volume tree answers the nearest-neighbour queries the ring search
// hands over: points far from the cloud, chiefly the unbounded fitness pass
// (Registration::getFitnessScore, fine_registration.cpp:126).  Built over the
// cell-sorted order its upper nodes would be thin slabs of whole grid planes;
// built over the Morton order of the points every node bounds a compact region,
// so a far query prunes most of the tree at once.  The order is a radix sort of
// (cloud, 24-bit Morton code) keys carrying each point's cell-sorted position;
// the points are then gathered into that order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"
#include "lcr_search.hpp"

namespace lcr {

namespace {

// key = (cloud, Morton code): 32-bit keys while the cloud index fits in the
// bits above the code, 64-bit otherwise.
template <class Key>
__global__ void k_tree_keys(const GridDesc* __restrict__ grids, const float4* __restrict__ pts,
                            const int* __restrict__ cloud_of_sorted, int64_t n,
                            Key* __restrict__ keys, int* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int c = cloud_of_sorted[j];
  const float4 p = pts[j];
  keys[j] = ((Key)c << kTreeKeyBits) | (Key)tree_key(grids[c], p.x, p.y, p.z);
  vals[j] = (int)j;
}

__global__ void k_tree_gather(const float4* __restrict__ pts, const int* __restrict__ tpos, int64_t n,
                              float4* __restrict__ tpts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tpts[i] = pts[tpos[i]];
}

int key_bits(int nclouds) {
  int b = 0;
  while ((1 << b) < nclouds) ++b;
  return kTreeKeyBits + b;
}

bool narrow_keys(int nclouds) { return key_bits(nclouds) <= 32; }

}  // namespace

size_t tree_order_tmp_bytes(int64_t npts, int nclouds) {
  size_t bytes = 0;
  if (npts <= 0) return 0;
  const hipError_t e =
      narrow_keys(nclouds)
          ? hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr,
                                               (uint32_t*)nullptr, (const int*)nullptr, (int*)nullptr,
                                               (int)npts, 0, key_bits(nclouds))
          : hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr,
                                               (uint64_t*)nullptr, (const int*)nullptr, (int*)nullptr,
                                               (int)npts, 0, key_bits(nclouds));
  return e == hipSuccess ? bytes : 0;
}

hipError_t launch_tree_order(const GridDesc* grids, int nclouds, const float4* pts,
                             const int* cloud_of_sorted, int64_t npts, uint64_t* keys,
                             uint64_t* keys_sorted, int* vals, int* tpos, void* tmp, size_t tmp_bytes,
                             float4* tpts, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((npts + 255) / 256);
  size_t bytes = tmp_bytes;
  hipError_t e;
  if (narrow_keys(nclouds)) {
    uint32_t* k = reinterpret_cast<uint32_t*>(keys);
    uint32_t* ks = reinterpret_cast<uint32_t*>(keys_sorted);
    hipLaunchKernelGGL(k_tree_keys<uint32_t>, dim3(blocks), dim3(256), 0, s, grids, pts,
                       cloud_of_sorted, npts, k, vals);
    e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k, ks, vals, tpos, (int)npts, 0,
                                           key_bits(nclouds), s);
  } else {
    hipLaunchKernelGGL(k_tree_keys<uint64_t>, dim3(blocks), dim3(256), 0, s, grids, pts,
                       cloud_of_sorted, npts, keys, vals);
    e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, keys_sorted, vals, tpos, (int)npts, 0,
                                           key_bits(nclouds), s);
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_tree_gather, dim3(blocks), dim3(256), 0, s, pts, tpos, npts, tpts);
  return hipSuccess;
}

}  // namespace lcr
