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
// with t from LDS and M read along z (coalesced, cached), z' over the band
// z - below .. z + above outside which M is zero (an axial blur is banded;
// the skipped terms are exact zeros).
__global__ __launch_bounds__(256) void psf_apply_kernel(const int *__restrict__ row_ptr,
                                                        const int *__restrict__ col,
                                                        const float *__restrict__ val, int nrows,
                                                        const float *__restrict__ m, int nz,
                                                        int below, int above,
                                                        const float *__restrict__ img,
                                                        float *__restrict__ out)
{
    extern __shared__ float t_lds[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // wave-uniform row: the CSR entries come through scalar loads
    const int r = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kPsfRowsPerBlock + w);
    float *t = t_lds + w * nz;
    if (r < nrows) {
        const int k0 = row_ptr[r], k1 = row_ptr[r + 1];
        // 256 columns per pass, 4 per lane, so that each CSR entry issues up
        // to 4 independent row loads
        for (int zb = 0; zb < nz; zb += 256) {
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
            for (int k = k0; k < k1; k++) {
                const float v = val[k];
                const float *src = img + (int64_t)col[k] * nz + zb + lane;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (zb + lane + 64 * j < nz)
                        acc[j] = fmaf(v, src[64 * j], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (zb + lane + 64 * j < nz)
                    t[zb + lane + 64 * j] = acc[j];
        }
    }
    __syncthreads();
    if (r < nrows) {
        for (int z = lane; z < nz; z += 64) {
            float acc = 0.0f;
            const int z0 = max(0, z - below), z1 = min(nz - 1, z + above);
            for (int zp = z0; zp <= z1; zp++)
                acc = fmaf(t[zp], m[(int64_t)zp * nz + z], acc);
            out[(int64_t)r * nz + z] = acc;
        }
    }
}

} // namespace
} // namespace srfx

extern "C" int srfx_psf_apply(const int32_t *row_ptr, const int32_t *col, const float *val,
                              int32_t nrows, const float *m, int32_t nz, int32_t below,
                              int32_t above, const float *img, float *out, void *stream)
{
    using namespace srfx;
    if (nrows < 0 || nz < 1 || below < 0 || above < 0 ||
        (nrows > 0 && (!row_ptr || !col || !val || !m || !img || !out)))
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
                       val, nrows, m, nz, below, above, img, out);
    return launch_status();
}
