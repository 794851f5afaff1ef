// Shared device/host helpers of libtgx (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define TGX_WAVE 64

// Record a HIP failure in the library's error slot and return the status code.
int tgx_set_error(int code, const char *fmt, ...);
#define TGX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return tgx_set_error(2, "%s:%d %s: %s", __FILE__, __LINE__, #call,           \
                           hipGetErrorString(e_));                                 \
  } while (0)

// Grid-stride helpers.  Launches are 1-D over columns; blocks are remapped so
// that each of the 8 XCDs (which receive blocks round-robin) works on one
// contiguous range of columns, keeping neighbouring latitude rows in one L2.
__device__ __forceinline__ int tgx_xcd_block(int nblocks) {
  const int nxcd = 8;
  int b = blockIdx.x;
  int per = (nblocks + nxcd - 1) / nxcd;
  int xcd = b % nxcd, slot = b / nxcd;
  int nb = xcd * per + slot;
  // the last XCD ranges may be short: fold overflow back (keeps a bijection
  // because per*nxcd >= nblocks and we only drop ids >= nblocks)
  return nb;
}
