// BF16 MFMA batched GEMM for the Fourier-neural-operator surrogate: every
// matrix product of its forward and backward passes (lifting, pointwise
// channel mixing, per-mode spectral mixing, projection, and their weight and
// input gradients) runs through this kernel.
//
//   C[z] (+)= A[z] * B[z],  z = 0 .. batch-1
//   A: M x K bf16, element (m, k) at A + z*sAz + m*sAm + k*sAk
//   B: K x N bf16, element (k, n) at B + z*sBz + k*sBk + n*sBn
//   C: M x N fp32, element (m, n) at C + z*sCz + m*sCm + n
//
// General strides let the backward pass read transposed operands in place.
// Products are exact in FP32 and accumulate in FP32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16, gfx950).
#include <hip/hip_runtime.h>

#include "cgfd_fno.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, PAD = 8;

struct GemmArgs {
    const unsigned short *A, *B;
    float *C;
    long sAz, sAm, sAk, sBz, sBk, sBn, sCz, sCm;
    int M, N, K, accumulate;
};

// 256 lanes = 4 wavefronts; wavefront w owns the 32x32 block
// (32*(w/2), 32*(w%2)) of the 64x64 output tile as 2x2 MFMA 16x16 blocks.
// A and B tiles are staged in LDS with K contiguous, so each lane reads its
// 8 consecutive k values of a row (A) or column (B) as one 16-byte vector.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ unsigned short As[BM][BK + PAD];
    __shared__ unsigned short Bs[BN][BK + PAD];
    const int z = blockIdx.z, m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const unsigned short *A = g.A + z * g.sAz;
    const unsigned short *B = g.B + z * g.sBz;
    float *C = g.C + z * g.sCz;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int lr = tid >> 2, lk = (tid & 3) * 8; // staging slot: one row (or column), 8 k values
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        {
            const int m = m0 + lr;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + lk + j;
                As[lr][lk + j] = (m < g.M && k < g.K) ? A[m * g.sAm + k * g.sAk] : (unsigned short)0;
            }
            const int n = n0 + lr;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + lk + j;
                Bs[lr][lk + j] = (n < g.N && k < g.K) ? B[k * g.sBk + n * g.sBn] : (unsigned short)0;
            }
        }
        __syncthreads();
        const int kq = (lane >> 4) * 8;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm + bi * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn + bj * 16 + (lane & 15)][kq]);
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[bi][bj], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // lane l holds rows 4*(l/16)+i, column l%16 of each 16x16 block
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + wm + bi * 16 + (lane >> 4) * 4 + i;
                const int n = n0 + wn + bj * 16 + (lane & 15);
                if (m < g.M && n < g.N) {
                    float *p = C + m * g.sCm + n;
                    *p = g.accumulate ? *p + acc[bi][bj][i] : acc[bi][bj][i];
                }
            }
}

} // namespace

extern "C" int cgfd_fno_gemm_bf16(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk,
                                  long sBn, float *C, long sCz, long sCm, int M, int N, int K, int batch,
                                  int accumulate, void *stream)
{
    if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || batch <= 0) return CGFD_FNO_EINVAL;
    const long tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    if (tm > 65535 || batch > 65535 || tn > 0x7fffffffL) return CGFD_FNO_EINVAL;
    GemmArgs g{(const unsigned short *)A, (const unsigned short *)B, C, sAz, sAm, sAk, sBz, sBk, sBn, sCz, sCm,
               M, N, K, accumulate ? 1 : 0};
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)tn, (unsigned)tm, (unsigned)batch), dim3(256), 0,
                       (hipStream_t)stream, g);
    return hipGetLastError() == hipSuccess ? 0 : CGFD_FNO_EHIP;
}
