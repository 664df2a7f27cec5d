// Kernels of the Fourier-neural-operator surrogate: every matrix product of
// its forward and backward passes (lifting, pointwise channel mixing,
// per-mode spectral mixing, the truncated transforms, projection, and their
// weight and input gradients) with the bias, residual and GELU of a layer
// fused into the product's epilogue; the GELU backward; and the Adam update
// of all parameters in one launch.
//
//   C[z] = act((C[z] +) A[z] * B[z] + bias + D[z]),  z = 0 .. batch-1
//   (and, on request, C[z] again in bf16: the operand of the next product)
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
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

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
    // epilogue
    const float *bias; // [N] or null
    const float *D;    // addend, (m, n) at D + z*sDz + m*sDm + n, or null
    long sDz, sDm;
    float *pre;        // the value before the activation, laid out as C, or null
    int act;           // CGFD_FNO_ACT_*
    unsigned short *Cb; // bf16 copy of the result, (m, n) at Cb + z*sEz + m*sEm + n, or null
    long sEz, sEm;
};

// One lane's share of a 64 x 32 operand tile (rows r0.., k0..) whose element
// (r, k) sits at P + r*sR + k*sK: eight bf16 values, loaded into registers
// so that the next tile's loads overlap the current tile's products.
template <int MODE>
__device__ __forceinline__ u16x8 fetch(const unsigned short *P, long sR, long sK, int r0, int k0, int R, int K,
                                       int tid)
{
    u16x8 v;
    if (MODE == STAGE_MN) {
        // 8 consecutive rows at one k
        const int r = r0 + (tid & 7) * 8, k = k0 + (tid >> 3);
        if (k < K && r + 7 < R) {
            v = *reinterpret_cast<const u16x8 *>(P + r * sR + k * sK);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (k < K && r + i < R) ? P[(r + i) * sR + k * sK] : (unsigned short)0;
        }
    } else {
        // 8 consecutive k of one row
        const int r = r0 + (tid >> 2), k = k0 + (tid & 3) * 8;
        if (MODE == STAGE_K && r < R && k + 7 < K) {
            v = *reinterpret_cast<const u16x8 *>(P + r * sR + k);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (r < R && k + j < K) ? P[r * sR + (k + j) * sK] : (unsigned short)0;
        }
    }
    return v;
}

// Store a fetched share into T[r][k] (k contiguous in LDS).
template <int MODE>
__device__ __forceinline__ void park(unsigned short (*T)[BK + PAD], u16x8 v, int tid)
{
    if (MODE == STAGE_MN) {
        const int rg = (tid & 7) * 8, kk = tid >> 3;
#pragma unroll
        for (int i = 0; i < 8; ++i) T[rg + i][kk] = v[i];
    } else {
        *reinterpret_cast<u16x8 *>(&T[tid >> 2][(tid & 3) * 8]) = v;
    }
}

// bf16 of an fp32 value, rounded to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// GELU, erf form: x/2 (1 + erf(x/sqrt 2)), and its derivative
// Phi(x) + x phi(x).
__device__ __forceinline__ float gelu(float x) { return x * 0.5f * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// 256 lanes = 4 wavefronts; wavefront w owns the 32x32 block
// (32*(w/2), 32*(w%2)) of the 64x64 output tile as 2x2 MFMA 16x16 blocks.
// A and B tiles are staged in LDS with K contiguous, so each lane reads its
// 8 consecutive k values of a row (A) or column (B) as one 16-byte vector.
// Two LDS buffers: the next tile is fetched into registers before the
// current tile's products and parked in the other buffer after them, so one
// barrier per K step suffices.
template <int AM, int BMODE>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) unsigned short As[2][BM][BK + PAD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[2][BN][BK + PAD];
    const int z = blockIdx.z, m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const unsigned short *A = g.A + z * g.sAz;
    const unsigned short *B = g.B + z * g.sBz;
    float *C = g.C + z * g.sCz;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // A rows are m (stride sAm), B "rows" are n (stride sBn)
    park<AM>(As[0], fetch<AM>(A, g.sAm, g.sAk, m0, 0, g.M, g.K, tid), tid);
    park<BMODE>(Bs[0], fetch<BMODE>(B, g.sBn, g.sBk, n0, 0, g.N, g.K, tid), tid);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        u16x8 na, nb;
        if (more) {
            na = fetch<AM>(A, g.sAm, g.sAk, m0, k0 + BK, g.M, g.K, tid);
            nb = fetch<BMODE>(B, g.sBn, g.sBk, n0, k0 + BK, g.N, g.K, tid);
        }
        const int kq = (lane >> 4) * 8;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[cur][wm + bi * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn + bj * 16 + (lane & 15)][kq]);
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[bi][bj], 0, 0, 0);
            }
        }
        if (more) {
            park<AM>(As[cur ^ 1], na, tid);
            park<BMODE>(Bs[cur ^ 1], nb, tid);
        }
        __syncthreads();
        cur ^= 1;
    }
    // lane l holds rows 4*(l/16)+i, column l%16 of each 16x16 block
    const float *D = g.D ? g.D + z * g.sDz : nullptr;
    float *pre = g.pre ? g.pre + z * g.sCz : nullptr;
    unsigned short *Cb = g.Cb ? g.Cb + z * g.sEz : nullptr;
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
                    float v = acc[bi][bj][i];
                    if (g.accumulate) v = *p + v;
                    if (g.bias) v = v + g.bias[n];
                    if (D) v = v + D[m * g.sDm + n];
                    if (pre) pre[m * g.sCm + n] = v;
                    if (g.act == CGFD_FNO_ACT_GELU) v = gelu(v);
                    *p = v;
                    if (Cb) Cb[m * g.sEm + n] = to_bf16(v);
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


// gx may alias gy: each element is read before it is written, by one lane.
__global__ void __launch_bounds__(256) gelu_bwd_kernel(const float *gy, const float *__restrict__ pre, float *gx,
                                                       unsigned short *__restrict__ gxb, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = gy[i] * gelu_grad(pre[i]);
        gx[i] = v;
        if (gxb) gxb[i] = to_bf16(v);
    }
}

// One Adam step (no weight decay, no amsgrad), as torch.optim.Adam computes
// it: m += (1-b1) (g - m); v = b2 v + (1-b2) g^2;
// p -= (lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps), with the scalars formed
// in double and rounded once to float; with clear set the gradient is zeroed
// as it is read.  Block b updates chunk b: elements
// [off, off + CGFD_FNO_ADAM_CHUNK) of tensor t, with (t, off) = chunks[2b],
// chunks[2b+1] and the tensor's (p, g, m, v, n) = tensors[5t .. 5t+4].
__global__ void __launch_bounds__(256) adam_kernel(const long long *__restrict__ tensors,
                                                   const long long *__restrict__ chunks, float w1, float b2,
                                                   float w2, float lr_bc1, float bc2_sqrt, float eps, int clear)
{
    const long long t = chunks[2 * blockIdx.x], off = chunks[2 * blockIdx.x + 1];
    float *p = reinterpret_cast<float *>(tensors[5 * t]);
    float *g = reinterpret_cast<float *>(tensors[5 * t + 1]);
    float *m = reinterpret_cast<float *>(tensors[5 * t + 2]);
    float *v = reinterpret_cast<float *>(tensors[5 * t + 3]);
    const long long n = tensors[5 * t + 4];
    const long long end = off + CGFD_FNO_ADAM_CHUNK < n ? off + CGFD_FNO_ADAM_CHUNK : n;
    for (long long i = off + threadIdx.x; i < end; i += blockDim.x) {
        const float gi = g[i];
        if (clear) g[i] = 0.0f; // the gradient is read once: the next backward accumulates into zeros
        const float mi = m[i] + w1 * (gi - m[i]);
        const float vi = v[i] * b2 + w2 * (gi * gi);
        m[i] = mi;
        v[i] = vi;
        p[i] = p[i] + -lr_bc1 * (mi / (sqrtf(vi) / bc2_sqrt + eps));
    }
}

} // namespace

extern "C" int cgfd_fno_gemm_bf16_ex(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk,
                                     long sBn, float *C, long sCz, long sCm, int M, int N, int K, int batch,
                                     int accumulate, const float *bias, const float *D, long sDz, long sDm,
                                     float *pre, int act, void *C_bf16, long sEz, long sEm, void *stream)
{
    if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || batch <= 0) return CGFD_FNO_EINVAL;
    if (act != CGFD_FNO_ACT_NONE && act != CGFD_FNO_ACT_GELU) return CGFD_FNO_EINVAL;
    const long tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    if (tm > 65535 || batch > 65535 || tn > 0x7fffffffL) return CGFD_FNO_EINVAL;
    GemmArgs g{(const unsigned short *)A, (const unsigned short *)B, C, sAz, sAm, sAk, sBz, sBk, sBn, sCz, sCm,
               M, N, K, accumulate ? 1 : 0, stage_mode(A, sAz, sAm, sAk), stage_mode(B, sBz, sBn, sBk),
               bias, D, sDz, sDm, pre, act, (unsigned short *)C_bf16, sEz, sEm};
    const dim3 grid((unsigned)tn, (unsigned)tm, (unsigned)batch);
    const hipStream_t s = (hipStream_t)stream;
    if (g.amode == STAGE_K) launch_b<STAGE_K>(g, grid, s);
    else if (g.amode == STAGE_MN) launch_b<STAGE_MN>(g, grid, s);
    else launch_b<STAGE_ANY>(g, grid, s);
    return hipGetLastError() == hipSuccess ? 0 : CGFD_FNO_EHIP;
}

extern "C" int cgfd_fno_gemm_bf16(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk,
                                  long sBn, float *C, long sCz, long sCm, int M, int N, int K, int batch,
                                  int accumulate, void *stream)
{
    return cgfd_fno_gemm_bf16_ex(A, sAz, sAm, sAk, B, sBz, sBk, sBn, C, sCz, sCm, M, N, K, batch, accumulate,
                                 nullptr, nullptr, 0, 0, nullptr, CGFD_FNO_ACT_NONE, nullptr, 0, 0, stream);
}

extern "C" int cgfd_fno_gelu_bwd(const float *gy, const float *pre, float *gx, void *gx_bf16, long n, void *stream)
{
    if (!gy || !pre || !gx || n < 0) return CGFD_FNO_EINVAL;
    if (n == 0) return 0;
    const long blocks = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, gy, pre, gx,
                       (unsigned short *)gx_bf16, n);
    return hipGetLastError() == hipSuccess ? 0 : CGFD_FNO_EHIP;
}

extern "C" int cgfd_fno_adam(const long long *tensors, const long long *chunks, int nchunks, double lr,
                             double beta1, double beta2, double eps, long long step, int clear_grad, void *stream)
{
    if (!tensors || !chunks || nchunks < 0 || step < 1 || !(lr >= 0.0) || !(eps > 0.0) || !(beta1 >= 0.0) ||
        !(beta1 < 1.0) || !(beta2 >= 0.0) || !(beta2 < 1.0))
        return CGFD_FNO_EINVAL;
    if (nchunks == 0) return 0;
    const double bc1 = 1.0 - __builtin_pow(beta1, (double)step);
    const double bc2 = 1.0 - __builtin_pow(beta2, (double)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)nchunks), dim3(256), 0, (hipStream_t)stream, tensors, chunks,
                       (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2), (float)(lr / bc1),
                       (float)__builtin_sqrt(bc2), (float)eps, clear_grad ? 1 : 0);
    return hipGetLastError() == hipSuccess ? 0 : CGFD_FNO_EHIP;
}
