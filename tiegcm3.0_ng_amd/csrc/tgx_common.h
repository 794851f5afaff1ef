// Shared device/host helpers of libtgx (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define TGX_WAVE 64
// Column tile of the tile kernels and stencil preludes: a block of 256
// threads takes TGX_TILE_X x TGX_TILE_Y columns.
#define TGX_TILE_X 32
#define TGX_TILE_Y 8

// Record a HIP failure in the library's error slot and return the status code.
int tgx_set_error(int code, const char *fmt, ...);
// True on the process's main thread.  Per-thread device resources are
// released when a worker thread exits; the main thread's are kept until the
// process ends (its thread-local destructors run during exit, when the HIP
// runtime may already be shutting down).
bool tgx_on_main_thread();
#define TGX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return tgx_set_error(2, "%s:%d %s: %s", __FILE__, __LINE__, #call,           \
                           hipGetErrorString(e_));                                 \
  } while (0)

// Launches are 1-D over columns.  The hardware hands consecutive block ids
// to the 8 XCDs round-robin; this remap gives every XCD one contiguous range
// of block ids (neighbouring latitude rows then share one L2).  Bijective on
// [0, nblocks) for any nblocks.
__device__ __forceinline__ int tgx_xcd_block(int nblocks) {
  const int nxcd = 8;
  int b = blockIdx.x;
  int x = b % nxcd, slot = b / nxcd;
  int q = nblocks / nxcd, r = nblocks % nxcd;
  return x * q + (x < r ? x : r) + slot;
}
