// Layout conversion between the model's arrays and the library's level
// planes.  The Fortran model holds a 3-D nest field as f(nk, lond0:lond1,
// latd0:latd1), level fastest (fields_ng.F90:226-261); the fused step reads
// longitude-fastest level planes (tgx_step.h).  A block converts a run of
// PACK_TC neighbouring columns of one latitude row: their nk * PACK_TC values
// are one contiguous span of the model array, read with coalesced loads into
// LDS, and written back one level at a time as PACK_TC consecutive doubles of
// a plane (and the reverse for unpack).  Nothing is transposed through HBM.
#include <hip/hip_runtime.h>

#include "../../include/tgx.h"
#include "../../include/tgx_step.h"
#include "tgx_common.h"

namespace tgx {
namespace pack {

constexpr int PACK_TC = 32;       // columns per block
constexpr int PACK_MAX_NK = 240;  // LDS: PACK_TC * (nk + 1) doubles <= 64 KiB

template <bool TO_PLANES>
__global__ void __launch_bounds__(256) levels_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                     int nk, int nlond, int nlatd, int ld) {
  extern __shared__ double t[];  // [PACK_TC][nk + 1]
  const int ntx = (nlond + PACK_TC - 1) / PACK_TC;
  const int j = blockIdx.x / ntx, i0 = (blockIdx.x % ntx) * PACK_TC;
  const int nc = nlond - i0 < PACK_TC ? nlond - i0 : PACK_TC;  // columns of this block
  const int span = nc * nk;
  const size_t kf = ((size_t)j * nlond + i0) * (size_t)nk;       // first value in the model array
  const size_t plane = (size_t)ld * nlatd, row = (size_t)j * ld + i0;
  if (TO_PLANES) {
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
      const int c = e / nk, k = e - c * nk;
      t[c * (nk + 1) + k] = src[kf + e];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
      const int k = e / nc, c = e - k * nc;
      dst[(size_t)k * plane + row + c] = t[c * (nk + 1) + k];
    }
  } else {
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
      const int k = e / nc, c = e - k * nc;
      t[c * (nk + 1) + k] = src[(size_t)k * plane + row + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
      const int c = e / nk, k = e - c * nk;
      dst[kf + e] = t[c * (nk + 1) + k];
    }
  }
}

}  // namespace pack
}  // namespace tgx

namespace {
int launch_levels(bool to_planes, const double *src, double *dst, int nk, int nlond, int nlatd, int ld,
                  void *stream) {
  using namespace tgx::pack;
  if (!src || !dst || nk < 1 || nk > PACK_MAX_NK || nlond < 1 || nlatd < 1 || ld < nlond)
    return tgx_set_error(TGX_ERR_ARG, "tgx_ng_%s_levels: bad shape nk=%d nlond=%d nlatd=%d ld=%d (nk <= %d)",
                         to_planes ? "pack" : "unpack", nk, nlond, nlatd, ld, PACK_MAX_NK);
  const int ntx = (nlond + PACK_TC - 1) / PACK_TC;
  const size_t lds = sizeof(double) * PACK_TC * (size_t)(nk + 1);
  const dim3 grid((unsigned)(ntx * nlatd)), block(256);
  if (to_planes)
    hipLaunchKernelGGL(levels_kernel<true>, grid, block, lds, (hipStream_t)stream, src, dst, nk, nlond, nlatd, ld);
  else
    hipLaunchKernelGGL(levels_kernel<false>, grid, block, lds, (hipStream_t)stream, src, dst, nk, nlond, nlatd, ld);
  TGX_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int tgx_ng_pack_levels(const double *src, double *dst, int nk, int nlond, int nlatd, int ld,
                                  void *stream) {
  return launch_levels(true, src, dst, nk, nlond, nlatd, ld, stream);
}

extern "C" int tgx_ng_unpack_levels(const double *src, double *dst, int nk, int nlond, int nlatd, int ld,
                                    void *stream) {
  return launch_levels(false, src, dst, nk, nlond, nlatd, ld, stream);
}
