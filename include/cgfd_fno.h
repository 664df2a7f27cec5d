/*
 * cgfd_fno.h -- C-ABI of the kernels of the Fourier-neural-operator
 * surrogate: BF16 MFMA products with fused epilogues, the GELU backward and
 * the Adam update (no reference counterpart: the reference has no learned
 * model; the surrogate maps source location, mechanism and moment to surface
 * peak-ground-velocity maps and station seismograms for a fixed medium,
 * trained on runs of this engine).
 */
#ifndef CGFD_FNO_H
#define CGFD_FNO_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_FNO_EINVAL (-1)
#define CGFD_FNO_EHIP (-3)

/* epilogue activations of cgfd_fno_gemm_bf16_ex */
#define CGFD_FNO_ACT_NONE 0
#define CGFD_FNO_ACT_GELU 1 /* x/2 (1 + erf(x / sqrt 2)) */

/* elements of one tensor that one block of cgfd_fno_adam updates */
#define CGFD_FNO_ADAM_CHUNK 8192

/* Batched GEMM on the matrix cores, bf16 inputs, fp32 accumulation:
 *   C[z] = A[z] * B[z]   (accumulate = 0)   or   C[z] += A[z] * B[z]
 * for z < batch.  A is M x K with element (m, k) at A + z*sAz + m*sAm + k*sAk,
 * B is K x N with (k, n) at B + z*sBz + k*sBk + n*sBn (bf16 elements, strides
 * in elements), C is M x N fp32 with (m, n) at C + z*sCz + m*sCm + n.  The
 * caller guarantees that every addressed element lies in its buffer.
 * Asynchronous on `stream` (a hipStream_t; null = the default stream).
 * M/64 and batch must not exceed 65535. */
int cgfd_fno_gemm_bf16(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk, long sBn,
                       float *C, long sCz, long sCm, int M, int N, int K, int batch, int accumulate, void *stream);

/* The same product with a layer's epilogue fused into it:
 *   pre[z] = (C[z] +) A[z] * B[z] + bias + D[z],   C[z] = act(pre[z])
 * bias: N fp32 values added to every row, or null.  D: an M x N fp32 addend
 * with (m, n) at D + z*sDz + m*sDm + n, or null.  pre: where the value before
 * the activation is stored (same layout as C; what the activation's backward
 * needs), or null.  act: CGFD_FNO_ACT_NONE or CGFD_FNO_ACT_GELU.  C_bf16:
 * where C is stored again rounded to bf16 (nearest even), (m, n) at C_bf16 +
 * z*sEz + m*sEm + n (the next product's operand), or null. */
int cgfd_fno_gemm_bf16_ex(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk, long sBn,
                          float *C, long sCz, long sCm, int M, int N, int K, int batch, int accumulate,
                          const float *bias, const float *D, long sDz, long sDm, float *pre, int act, void *C_bf16,
                          long sEz, long sEm, void *stream);

/* Backward of GELU over n fp32 elements: gx = gy * GELU'(pre), and, when
 * gx_bf16 is not null, the same values rounded to bf16 (nearest even) there
 * (the operand of the next backward products).  gx may alias gy. */
int cgfd_fno_gelu_bwd(const float *gy, const float *pre, float *gx, void *gx_bf16, long n, void *stream);

/* One Adam step (torch.optim.Adam without weight decay or amsgrad) over many
 * fp32 tensors in one launch.  tensors: device array of 5 int64 per tensor
 * (parameter, gradient, first moment, second moment addresses, element
 * count).  chunks: device array of 2 int64 per block (tensor index, first
 * element, a multiple of CGFD_FNO_ADAM_CHUNK below the count), nchunks
 * blocks.  step counts from 1; the bias corrections 1 - beta^step, lr /
 * (1 - beta1^step) and sqrt(1 - beta2^step) are formed in double.  With
 * clear_grad, each gradient element is set to zero once read (the next
 * backward pass then accumulates into zeros, with no separate clearing). */
int cgfd_fno_adam(const long long *tensors, const long long *chunks, int nchunks, double lr, double beta1,
                  double beta2, double eps, long long step, int clear_grad, void *stream);

#ifdef __cplusplus
}
#endif

#endif
