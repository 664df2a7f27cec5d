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

// How a block stages an operand tile: 0 element by element (any strides),
// 1 eight k-consecutive elements per 16-byte load (unit k stride), 2 eight
// m- (or n-) consecutive elements per 16-byte load, transposed into LDS.
// Modes 1 and 2 need a 16-byte aligned base and the other strides in
// multiples of 8 elements; edge groups fall back to single loads.
enum { STAGE_ANY = 0, STAGE_K = 1, STAGE_MN = 2 };

struct GemmArgs {
    const unsigned short *A, *B;
    float *C;
    long sAz, sAm, sAk, sBz, sBk, sBn, sCz, sCm;
    int M, N, K, accumulate, amode, bmode;
};

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// Stage the 64 x 32 tile (rows r0.., k0..) of an operand whose element (r, k)
// sits at P + r*sR + k*sK into T[r][k] (k contiguous in LDS).
template <int MODE>
__device__ __forceinline__ void stage(unsigned short (*T)[BK + PAD], const unsigned short *P, long sR, long sK,
                                      int r0, int k0, int R, int K, int tid)
{
    if (MODE == STAGE_MN) {
        // thread: 8 consecutive rows at one k
        const int rg = (tid & 7) * 8, kk = tid >> 3;
        const int r = r0 + rg, k = k0 + kk;
        u16x8 v;
        if (k < K && r + 7 < R) {
            v = *reinterpret_cast<const u16x8 *>(P + r * sR + k * sK);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (k < K && r + i < R) ? P[(r + i) * sR + k * sK] : (unsigned short)0;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) T[rg + i][kk] = v[i];
    } else {
        // thread: 8 consecutive k of one row
        const int lr = tid >> 2, lk = (tid & 3) * 8;
        const int r = r0 + lr, k = k0 + lk;
        if (MODE == STAGE_K && r < R && k + 7 < K) {
            *reinterpret_cast<u16x8 *>(&T[lr][lk]) = *reinterpret_cast<const u16x8 *>(P + r * sR + k);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                T[lr][lk + j] = (r < R && k + j < K) ? P[r * sR + (k + j) * sK] : (unsigned short)0;
        }
    }
}

// 256 lanes = 4 wavefronts; wavefront w owns the 32x32 block
// (32*(w/2), 32*(w%2)) of the 64x64 output tile as 2x2 MFMA 16x16 blocks.
// A and B tiles are staged in LDS with K contiguous, so each lane reads its
// 8 consecutive k values of a row (A) or column (B) as one 16-byte vector.
template <int AM, int BMODE>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) unsigned short As[BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[BN][BK + PAD];
    const int z = blockIdx.z, m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const unsigned short *A = g.A + z * g.sAz;
    const unsigned short *B = g.B + z * g.sBz;
    float *C = g.C + z * g.sCz;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        // A rows are m (stride sAm), B "rows" are n (stride sBn)
        stage<AM>(As, A, g.sAm, g.sAk, m0, k0, g.M, g.K, tid);
        stage<BMODE>(Bs, B, g.sBn, g.sBk, n0, k0, g.N, g.K, tid);
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

template <int AM>
void launch_b(const GemmArgs &g, dim3 grid, hipStream_t s)
{
    if (g.bmode == STAGE_K) hipLaunchKernelGGL((gemm_bf16_kernel<AM, STAGE_K>), grid, dim3(256), 0, s, g);
    else if (g.bmode == STAGE_MN) hipLaunchKernelGGL((gemm_bf16_kernel<AM, STAGE_MN>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_bf16_kernel<AM, STAGE_ANY>), grid, dim3(256), 0, s, g);
}

// staging mode of an operand: unit stride along k (STAGE_K) or along its
// rows (STAGE_MN), with a 16-byte aligned base and the other strides in
// multiples of 8 elements; otherwise element by element
int stage_mode(const void *p, long sZ, long sR, long sK)
{
    if ((reinterpret_cast<unsigned long>(p) & 15) != 0 || sZ % 8 != 0) return STAGE_ANY;
    if (sK == 1 && sR % 8 == 0) return STAGE_K;
    if (sR == 1 && sK % 8 == 0) return STAGE_MN;
    return STAGE_ANY;
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
               M, N, K, accumulate ? 1 : 0, stage_mode(A, sAz, sAm, sAk), stage_mode(B, sBz, sBn, sBk)};
    const dim3 grid((unsigned)tn, (unsigned)tm, (unsigned)batch);
    const hipStream_t s = (hipStream_t)stream;
    if (g.amode == STAGE_K) launch_b<STAGE_K>(g, grid, s);
    else if (g.amode == STAGE_MN) launch_b<STAGE_MN>(g, grid, s);
    else launch_b<STAGE_ANY>(g, grid, s);
    return hipGetLastError() == hipSuccess ? 0 : CGFD_FNO_EHIP;
}
