// cgfd_device.h -- device-side data layout and arithmetic helpers shared by
// the gfx950 kernels.
//
// Arithmetic contract: every expression below reproduces the rounding of the
// reference's host build (fes478/CoPara-CGFD3D src/calculate.cu compiled
// without FP contraction): one-sided stencils are accumulated in FP64 (the
// reference's coefficients are double literals, typenew.h:52-88) and rounded
// to the storage type once; everything else is storage-type arithmetic in
// source order.  The library is compiled with -ffp-contract=off; the only
// fused multiply-adds are the explicit ones in div_fixed().
#pragma once
#include <hip/hip_runtime.h>

namespace cgfd {

// one-sided DRP/opt MacCormack coefficients (typenew.h:52-56)
constexpr double DA_1 = -0.30874;
constexpr double DA0 = -0.6326;
constexpr double DA1 = 1.23300;
constexpr double DA2 = -0.3334;
constexpr double DA3 = 0.04168;
// Gottlieb 2-4 MacCormack, used two rows below the free surface (typenew.h:76-79)
constexpr double M24A0 = -1.1666666666666667;
constexpr double M24A1 = 1.3333333333333333;
constexpr double M24A2 = -0.16666666666666666;

enum { VX, VY, VZ, TXX, TYY, TZZ, TXY, TXZ, TYZ, NF };
enum { XIX, XIY, XIZ, ETX, ETY, ETZ, ZTX, ZTY, ZTZ, JAC, NDRV };

// Padded x-fastest layout of one 3-D field:
//   index(x, y, z) = base + ((long)z * nyp + y) * nxp + x,   0 <= x < ni ...
// nxp >= ni + 3 (rounded to 64 floats), nyp = nj + 3, and base = 4 planes, so
// every stencil neighbour up to 3 cells outside the valid box lands on a zero
// guard cell: the x overflow of one row lands in the zero tail of the previous
// or current row, the y overflow in the zero rows [nj, nyp) of the previous or
// current plane, and the z overflow in the 4 zero planes at either end.
struct Geom {
    int ni, nj, nk;
    int nxp, nyp;
    long plane;  // nxp * nyp
    long fs;     // element stride between fields of one set
    long base;   // index of (0,0,0)
};

__device__ __forceinline__ long gidx(const Geom &g, int x, int y, int z)
{
    return g.base + ((long)z * g.nyp + y) * g.nxp + x;
}

// Correctly rounded s/h for a fixed divisor h, given rh = RN(1/h): one
// product and a fused remainder correction (Markstein's theorem: with rh the
// correctly rounded reciprocal and q within one ulp of s/h, q + (s - h q) rh
// rounds to RN(s/h)).  Used for the FP64 stencil sums, whose magnitude keeps
// them far from the FP64 subnormal range.
__device__ __forceinline__ double div_fixed(double s, double h, double rh)
{
    double q = s * rh;
    double r = __builtin_fma(-q, h, s);
    return __builtin_fma(r, rh, q);
}

// One-sided DRP derivative from five taps in ascending position.
//   S = +1 (forward):  taps at offsets -1..+3, summed as strF (typenew.h:58)
//   S = -1 (backward): taps at offsets -3..+1, summed as strB (typenew.h:59),
//                      i.e. starting from the far (-3) end; the divisor is -h.
// The taps are storage-type values; T may be R itself or R already widened
// to double (the widening is exact, so the sum is the same).
template <typename R, int S, typename T>
__device__ __forceinline__ R dfd(T t0, T t1, T t2, T t3, T t4, double h, double rh)
{
    double s;
    if (S > 0)
        s = DA_1 * (double)t0 + DA0 * (double)t1 + DA1 * (double)t2 + DA2 * (double)t3 + DA3 * (double)t4;
    else
        s = DA3 * (double)t0 + DA2 * (double)t1 + DA1 * (double)t2 + DA0 * (double)t3 + DA_1 * (double)t4;
    double q = div_fixed(s, h, rh);
    return (R)(S > 0 ? q : -q);
}

// DRP sum in strF order for an arbitrary increment, as CalTIMG applies it to
// its 7-point traction vectors (calculate.cu:5502-5504): taps v[3-inc],
// v[3], v[3+inc], v[3+2inc], v[3+3inc] with inc = +-1.
template <typename R, int S>
__device__ __forceinline__ double drp_sum_strf(const R *v)
{
    return DA_1 * (double)v[3 - S] + DA0 * (double)v[3] + DA1 * (double)v[3 + S] + DA2 * (double)v[3 + 2 * S] +
           DA3 * (double)v[3 + 3 * S];
}

// typenew.h:44-45 (the reference's ABS/MAX, including their handling of
// signed zeros)
template <typename R> __device__ __forceinline__ R rabs(R a) { return a > 0 ? a : -a; }
template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return a >= b ? a : b; }

} // namespace cgfd
