// PSF blur of an image around the projector (the reference's PSFReconStep,
// model/recon_step.py:56-115): the image seen as a matrix X of nxy = nx * ny
// rows (voxel columns) by nz columns (slices), blurred across columns by a
// sparse nxy x nxy matrix S (CSR) and along z by a dense nz x nz matrix:
//   out[r, z] = sum_z' (sum_k S[r, c_k] X[c_k, z']) M[z', z]
// Forward (recon_step.py:79-90, S psf_xy and X psf_z^T): M = psf_z^T.
// Backward (recon_step.py:103-110, S^T psf_xy and X psf_z): M = psf_z, with
// S^T passed as its own CSR.  The reference forms X M first and then S (X M);
// the product is associative, so the sums here differ from its order by FP32
// rounding only.
#include "common.h"

namespace srfx {
namespace {

constexpr int kPsfRowsPerBlock = 4; // one wave of 64 lanes per row
constexpr int kPsfMaxNz = 2048;     // 4 rows x nz floats of LDS

// One wave per output row.  Lanes stride over z: first the sparse sum of the
// neighbour rows (each a contiguous run of nz floats, so the lanes' reads are
// coalesced) into the wave's LDS row t, then out[r, z] = sum_z' t[z'] M[z', z]
// with t broadcast from LDS and M read along z (coalesced, cached).
__global__ __launch_bounds__(256) void psf_apply_kernel(const int *__restrict__ row_ptr,
                                                        const int *__restrict__ col,
                                                        const float *__restrict__ val, int nrows,
                                                        const float *__restrict__ m, int nz,
                                                        const float *__restrict__ img,
                                                        float *__restrict__ out)
{
    extern __shared__ float t_lds[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * kPsfRowsPerBlock + w;
    float *t = t_lds + w * nz;
    if (r < nrows) {
        const int k0 = row_ptr[r], k1 = row_ptr[r + 1];
        for (int z = lane; z < nz; z += 64) {
            float acc = 0.0f;
            for (int k = k0; k < k1; k++)
                acc = fmaf(val[k], img[(int64_t)col[k] * nz + z], acc);
            t[z] = acc;
        }
    }
    __syncthreads();
    if (r < nrows) {
        for (int z = lane; z < nz; z += 64) {
            float acc = 0.0f;
            for (int zp = 0; zp < nz; zp++)
                acc = fmaf(t[zp], m[(int64_t)zp * nz + z], acc);
            out[(int64_t)r * nz + z] = acc;
        }
    }
}

} // namespace
} // namespace srfx

extern "C" int srfx_psf_apply(const int32_t *row_ptr, const int32_t *col, const float *val,
                              int32_t nrows, const float *m, int32_t nz, const float *img,
                              float *out, void *stream)
{
    using namespace srfx;
    if (nrows < 0 || nz < 1 || (nrows > 0 && (!row_ptr || !col || !val || !m || !img || !out)))
        return SRFX_EINVAL;
    if (img == out)
        return SRFX_EINVAL; // rows read neighbours: no in-place blur
    if (nz > kPsfMaxNz)
        return SRFX_EUNSUPPORTED;
    if (nrows == 0)
        return SRFX_OK;
    const int blocks = (nrows + kPsfRowsPerBlock - 1) / kPsfRowsPerBlock;
    hipLaunchKernelGGL(psf_apply_kernel, dim3(blocks), dim3(64 * kPsfRowsPerBlock),
                       sizeof(float) * kPsfRowsPerBlock * nz, (hipStream_t)stream, row_ptr, col,
                       val, nrows, m, nz, img, out);
    return launch_status();
}
